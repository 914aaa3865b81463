/*
 * remd_oracle.h -- CPU restatement of the reference REMD per-replica path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libremd_gpu.so, the remd_md driver) never links it.
 *
 * Every function restates the reference's CPU semantics (the makefile.gnu
 * build: SWITCHING + FOUR_SITE tuned kernel, nlane = 1, one OpenMP thread)
 * and cites the reference file:line it follows.  Where the reference has no
 * usable semantics (3/5-site site roles, slit wall, 2D/3D barostat factors)
 * the oracle implements the same rule as the GPU kernels; those places say
 * so, with the reason, in remd_oracle.c.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - bit-for-bit / 1e-13 agreement with oracle/_ref (the reference's own CPU
 *     translation units, compiled by oracle/Makefile) on the TIP4P tube;
 *   - closed forms: switching identities, two-molecule analytic site sums,
 *     tube wall table at r = 0, 10-4 slit wall, exchange pair lists,
 *     glibc rand() stream, NP-Hamiltonian drift O(dt^2), |q| = 1;
 *   - the reference's own fixture test/input/gpu0000.chk.
 */
#ifndef REMD_ORACLE_H
#define REMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Units (src/unit.h:3-18): out = {unit_mass, unit_length, unit_energy,
 * unit_time, unit_coulomb, unit_temp, unit_press, Na, ele}. */
void orc_units(double out[9]);

#define ORC_MAXSITE 5

typedef struct {
  int nsite;
  double r[ORC_MAXSITE][3]; /* body-frame site coordinates (reduced = A) */
  double m[ORC_MAXSITE];    /* reduced mass (oxygen = 1) */
  double sigma[ORC_MAXSITE];/* A */
  double eps[ORC_MAXSITE];  /* kJ/mol */
  double q[ORC_MAXSITE];    /* reduced charge, q^2/r in kJ/mol */
  double mass;              /* molecule mass */
  double inertia[3];        /* principal moments (src/variables.h:189-206) */
} orc_moltype;

/* Parse the k-th molecule block of a moltype.txt text
 * (src/molecules.cxx:69-131).  Returns 0 on success. */
int orc_moltype_parse(const char *text, int k, orc_moltype *out);

typedef struct {
  int nmol;
  int tmode, pmode, cmode;  /* tconstant, pconstant, confined */
  double dt, dthalf;        /* reduced time */
  double rcut, rswitch;     /* A */
  double wall_length;       /* A (tube radius / slit width) */
  int nfwall;
  const double *fwall;      /* nfwall x {f(r)/r form, energy} */
  int sort_interval;        /* 10 in the reference */
  double margin;            /* 0.5 A in the reference */
  int fp32;                 /* 1: pair math in float (shipped USE_FP32 build) */
  orc_moltype mt;
  /* state, AoS per molecule: r,v,f [3*nmol]; q,p,n [4*nmol] */
  double *r, *v, *f, *q, *p, *n;
  int *id;                  /* molecule identity, permuted with the sort */
  double *site_f;           /* nsite*nmol*3 site forces (switched + wall) */
  double s, Ps, Q;
  double Pv[3], W;
  double L[3];
  double T, P;              /* reduced */
  double gkT, H0;
  /* properties (src/variables.h:6-49) */
  double lj, clmb, wall, pot, vir[3];
  double tra, trav[3], rot, kin, tsm, tsp, tst, bst, tot, ham, Tk, prs[4];
  /* neighbour window (src/molecules.h:77-78) */
  int *jstart, *jend;
  int sort_axis;            /* 2 (z) for bulk/tube, 0 (x) for the slit */
  int *wfull;               /* contract: 1 where the window is the whole ring */
} orc_system;

int  orc_system_alloc(orc_system *s, int nmol);
void orc_system_free(orc_system *s);

double orc_volume(const orc_system *s);       /* src/molecules.h:357-366 */
double orc_bottom_area(const orc_system *s);  /* src/molecules.h:367-376 */

/* Stable ascending argsort of keys + cyclic window of half-width w
 * (src/molecules.cxx:541-585).  order[k] = index of k-th smallest key;
 * js/je are computed on the sorted keys. */
void orc_sort_window(const double *key, int n, double w,
                     int *order, int *js, int *je);
/* Window only, on already sorted keys. */
void orc_window(const double *sorted_key, int n, double w, int *js, int *je);
/* Contract (no reference counterpart): full[i] = 1 when molecule i's
 * window is the whole ring, i.e. the scan wrapped onto i itself, or the
 * first key outside on one side lies inside the other side's arc (every
 * key is within w of key i in one direction or the other).  The
 * reference then visits only the js..je range, which misses pairs. */
void orc_window_full(const double *sorted_key, int n, double w, const int *js,
                     const int *je, int *full);

/* Force pass (src/molecules.cxx:531-629): optional sort + window, M2A,
 * switched LJ + Coulomb, wall, A2M.  Sets f, n, site_f, lj, clmb, wall,
 * pot, vir. */
void orc_calc_force(orc_system *s, int do_sort, double margin);

/* One step ExecuteStep<MODE> (src/integrator.cxx:646-706). */
void orc_step(orc_system *s, int do_sort, double margin);
/* ExecuteSteps (src/integrator.cxx:708-803) incl. CalcProperties. */
void orc_execute_steps(orc_system *s, int nsteps);

void orc_calc_properties(orc_system *s);       /* src/molecules.cxx:702-734 */
void orc_initialize_property(orc_system *s);   /* src/molecules.h:260-270 */
void orc_set_temperature(orc_system *s, double T); /* src/molecules.h:332 */
void orc_change_temperature(orc_system *s, double T); /* molecules.h:333-353 */
void orc_kill_momentum(orc_system *s);         /* src/molecules.cxx:168-187 */
void orc_kill_angular_momentum_tube(orc_system *s); /* molecules.cxx:189-210 */
void orc_velocity_scaling(orc_system *s);      /* src/molecules.cxx:261-273 */
void orc_ang_velocity_scaling(orc_system *s);  /* src/molecules.cxx:296-311 */
void orc_update_hamiltonian(orc_system *s);    /* src/molecules.h:316-319 */
/* Average vector in the reference's enum order (src/remdinfo.h:27-41,
 * src/molecules.h:380-402); out[19]. */
void orc_get_average(const orc_system *s, double out[19]);

/* Wall tables (reduced units; R = radius / slit width). */
void orc_wall_table_tube(int nfwall, double R, double sgm, double eps,
                         double rho, double *fwall);   /* force_calculator.cxx:117-161 */
void orc_wall_table_slit(int nfwall, double H, double sgm, double eps,
                         double rho, double *fwall);   /* contract: 10-4 graphene wall */

/* Replica exchange bookkeeping (src/remdinfo.cxx:54-108, 348-359;
 * src/remd.cxx:464-538; src/remdinfo.h:276-359). */
int  orc_exchange_list(int ntemp, int npress, int type, int *a, int *b);
void orc_ladder_geometric(int ntemp, int npress, double tmin, double tmax,
                          double pmin, double pmax, double *T, double *P);
/* One ReplicaExchange call.  uniforms == NULL draws glibc rand()/RAND_MAX in
 * pair-list order, otherwise uses uniforms[k] for the k-th pair.  Returns the
 * number of accepted swaps. */
int  orc_replica_exchange(int ntemp, int npress, long step, int pmode,
                          const double *pot, const double *vol, int *index,
                          double *T, double *P, unsigned char *is_exchanged,
                          const double *uniforms, int *accepted /* per pair, may be NULL */);

/* glibc PRNG plumbing so tests can reproduce the reference stream. */
void orc_srand(unsigned seed);
int  orc_rand(void);

#ifdef __cplusplus
}
#endif
#endif
