// ref_harness.cpp -- drives the reference's own CPU translation units
// (force_calculator.cxx, integrator.cxx, molecules.cxx, remdinfo.cxx, built
// from /root/reference/src by oracle/Makefile into oracle/_ref/) through a
// plain C interface, so tests can pin the oracle restatement against the
// reference itself.  TEST INFRASTRUCTURE ONLY; this file is the project's
// harness, the reference sources are compiled where they lie and never
// copied.  The reference's io_manager/remd/main units need the un-vendored
// VMD molfile plugin and are not built; nothing here stands in for them.
#include "molecules.h"
#include "remdinfo.h"

#include <cstring>
#include <sstream>
#include <string>

namespace {
struct Handle {
  Molecules *md;
};

Parameter make_param(const char *mtl_path, int nmol, int nsite, double dt_fs,
                     double rcut, double rswitch, int tconstant,
                     double tstat_mass, int pconstant, double bstat_mass,
                     int interval, int confined, int nfwall,
                     double wall_length, double sgm_wall, double eps_wall,
                     double rho_wall, double temperature_K,
                     double pressure_Pa) {
  Parameter p;
  p.input_prefix = "";
  p.mtl_in = mtl_path;
  p.nmol = nmol;
  p.natom = nmol * nsite;
  p.nsite = nsite;
  p.dt = dt_fs;
  p.rcut = rcut;
  p.rswitch = rswitch;
  p.tconstant = tconstant;
  p.tstat_mass = tstat_mass;
  p.pconstant = pconstant;
  p.bstat_mass = bstat_mass;
  p.interval = interval;
  p.confined = confined;
  p.nfwall = nfwall;
  p.wall_length = wall_length;
  p.sgm_wall = sgm_wall;
  p.eps_wall = eps_wall;
  p.rho_wall = rho_wall;
  p.temperature = temperature_K;
  p.pressure = pressure_Pa;
  p.nthreads = 1;
  return p;
}
}  // namespace

extern "C" {

void *ref_create(const char *mtl_path, int nmol, int nsite, double dt_fs,
                 double rcut, double rswitch, int tconstant, double tstat_mass,
                 int pconstant, double bstat_mass, int interval, int confined,
                 int nfwall, double wall_length, double sgm_wall,
                 double eps_wall, double rho_wall, double temperature_K,
                 double pressure_Pa) {
  Parameter p = make_param(mtl_path, nmol, nsite, dt_fs, rcut, rswitch,
                           tconstant, tstat_mass, pconstant, bstat_mass,
                           interval, confined, nfwall, wall_length, sgm_wall,
                           eps_wall, rho_wall, temperature_K, pressure_Pa);
  Handle *h = new Handle;
  h->md = new Molecules(p);
  return h;
}

void ref_destroy(void *hp) {
  Handle *h = (Handle *)hp;
  delete h->md;
  delete h;
}

// State in the reference's internal representation: r fractional, v scaled
// by L*s, p scaled by s (what ReadCheckPoint leaves, io_manager.cxx:577-581).
void ref_set_state(void *hp, const double *r, const double *v,
                   const double *q, const double *p, const int *id,
                   const double *scal /* s Ps Q Pv0 Pv1 Pv2 W L0 L1 L2 */) {
  Molecules *md = ((Handle *)hp)->md;
  const double mass = [&] {
    double m = 0;
    for (auto &a : md->mlist[0].a) m += a.m;
    return m;
  }();
  for (int i = 0; i < md->nmol; i++) {
    Molecule &m = md->mlcl[i];
    for (int d = 0; d < 3; d++) m.r[d] = r[3 * i + d];
    for (int d = 0; d < 3; d++) m.v[d] = v[3 * i + d];
    for (int d = 0; d < 4; d++) m.q[d] = q[4 * i + d];
    for (int d = 0; d < 4; d++) m.p[d] = p[4 * i + d];
    m.m = mass;
    m.i = md->mlist[0].i;
    m.type = 0;
    m.id = id[i];
  }
  md->tst->s = scal[0];
  md->tst->Ps = scal[1];
  md->tst->Q = scal[2];
  md->bst->Pv[0] = scal[3];
  md->bst->Pv[1] = scal[4];
  md->bst->Pv[2] = scal[5];
  md->bst->W = scal[6];
  md->L[0] = scal[7];
  md->L[1] = scal[8];
  md->L[2] = scal[9];
}

void ref_get_state(void *hp, double *r, double *v, double *f, double *q,
                   double *p, double *n, int *id, double *scal) {
  Molecules *md = ((Handle *)hp)->md;
  for (int i = 0; i < md->nmol; i++) {
    const Molecule &m = md->mlcl[i];
    for (int d = 0; d < 3; d++) r[3 * i + d] = m.r[d];
    for (int d = 0; d < 3; d++) v[3 * i + d] = m.v[d];
    for (int d = 0; d < 3; d++) f[3 * i + d] = m.f[d];
    for (int d = 0; d < 4; d++) q[4 * i + d] = m.q[d];
    for (int d = 0; d < 4; d++) p[4 * i + d] = m.p[d];
    for (int d = 0; d < 4; d++) n[4 * i + d] = m.n[d];
    id[i] = m.id;
  }
  scal[0] = md->tst->s;
  scal[1] = md->tst->Ps;
  scal[2] = md->tst->Q;
  scal[3] = md->bst->Pv[0];
  scal[4] = md->bst->Pv[1];
  scal[5] = md->bst->Pv[2];
  scal[6] = md->bst->W;
  scal[7] = md->L[0];
  scal[8] = md->L[1];
  scal[9] = md->L[2];
}

// Site forces after the last CalcForcePot (Atom::f, variables.h:80-87).
void ref_get_site_forces(void *hp, double *af) {
  Molecules *md = ((Handle *)hp)->md;
  for (int i = 0; i < md->natom; i++)
    for (int d = 0; d < 3; d++) af[3 * i + d] = md->atom[i].f[d];
}

void ref_set_temperature(void *hp, double T) { ((Handle *)hp)->md->SetTemperature(T); }
void ref_set_pressure(void *hp, double P) { ((Handle *)hp)->md->SetPressure(P); }
void ref_initialize_property(void *hp) { ((Handle *)hp)->md->InitializeProperty(); }
void ref_calc_force(void *hp, int do_sort, double mergin) {
  ((Handle *)hp)->md->CalcForcePot(do_sort != 0, mergin);
}
void ref_execute_steps(void *hp) { ((Handle *)hp)->md->ExecuteSteps(); }
void ref_calc_properties(void *hp) { ((Handle *)hp)->md->CalcProperties(); }
void ref_update_hamiltonian(void *hp) { ((Handle *)hp)->md->UpdateHamiltonian(); }
void ref_velocity_scaling(void *hp) { ((Handle *)hp)->md->VelocityScaling(); }
void ref_ang_velocity_scaling(void *hp) { ((Handle *)hp)->md->AngVelocityScaling(); }
void ref_kill_momentum(void *hp) { ((Handle *)hp)->md->KillMomentum(); }
void ref_change_temperature(void *hp, double T) { ((Handle *)hp)->md->ChangeTemperature(T); }

// lj clmb wall pot vir0 vir1 vir2 tra rot kin tot ham T prs0 prs1 prs2 prs3 H0 gkT
void ref_get_props(void *hp, double *out) {
  const Property &pr = ((Handle *)hp)->md->prop;
  out[0] = pr.lj;
  out[1] = pr.clmb;
  out[2] = pr.wall;
  out[3] = pr.pot;
  out[4] = pr.vir[0];
  out[5] = pr.vir[1];
  out[6] = pr.vir[2];
  out[7] = pr.tra;
  out[8] = pr.rot;
  out[9] = pr.kin;
  out[10] = pr.tot;
  out[11] = pr.ham;
  out[12] = pr.T;
  out[13] = pr.prs[0];
  out[14] = pr.prs[1];
  out[15] = pr.prs[2];
  out[16] = pr.prs[3];
  out[17] = pr.H0;
  out[18] = pr.gkT;
}

void ref_get_average(void *hp, double *out) {
  Average a = ((Handle *)hp)->md->GetAverage();
  for (int i = 0; i < Average::NUM_ELEM; i++) out[i] = a.sum[i];
}

void ref_get_window(void *hp, int *js, int *je) {
  Molecules *md = ((Handle *)hp)->md;
  for (int i = 0; i < md->nmol; i++) {
    js[i] = md->jstart[i];
    je[i] = md->jend[i];
  }
}

int ref_nfwall(void *hp) { return ((Handle *)hp)->md->fclcltr->nfwall; }
void ref_wall_table(void *hp, double *out) {
  ForceCalculator *fc = ((Handle *)hp)->md->fclcltr;
  for (int i = 0; i < fc->nfwall; i++) {
    out[2 * i] = fc->fwall[i][0];
    out[2 * i + 1] = fc->fwall[i][1];
  }
}

// Molecule type after MakeMolTypeList: per site r[3] m s e q; inertia[3].
int ref_moltype(void *hp, double *sites, double *inertia) {
  Molecules *md = ((Handle *)hp)->md;
  const MolType &mt = md->mlist[0];
  for (size_t j = 0; j < mt.a.size(); j++) {
    sites[7 * j + 0] = mt.a[j].r[0];
    sites[7 * j + 1] = mt.a[j].r[1];
    sites[7 * j + 2] = mt.a[j].r[2];
    sites[7 * j + 3] = mt.a[j].m;
    sites[7 * j + 4] = mt.a[j].s;
    sites[7 * j + 5] = mt.a[j].e;
    sites[7 * j + 6] = mt.a[j].q;
  }
  inertia[0] = mt.i[0];
  inertia[1] = mt.i[1];
  inertia[2] = mt.i[2];
  return (int)mt.a.size();
}

double ref_dt(void *hp) { return ((Handle *)hp)->md->dt; }

// REMDInfo ladder and pair lists (remdinfo.cxx:54-108, 140-266, 348-359).
// T/P in K and Pa; outputs reduced T/P (as REMDInfo stores them).
int ref_remdinfo(int ntemp, int npress, double tmin, double tmax, double pmin,
                 double pmax, double *T, double *P, int *lists /* 4 x cap x 2 */,
                 int cap, int *len) {
  Parameter p;
  p.nreplica = ntemp * npress;
  p.ntemp = ntemp;
  p.npress = npress;
  p.temp_min = tmin;
  p.temp_max = tmax;
  p.press_min = pmin;
  p.press_max = pmax;
  p.cond_mode = 0;
  // ~REMDInfo deletes `new`-ed histograms with delete[]
  // (remdinfo.cxx:303), so the object is intentionally never destroyed.
  REMDInfo *info = new REMDInfo(p);
  for (int i = 0; i < p.nreplica; i++) {
    T[i] = info->GetTemperature(i);
    P[i] = info->GetPressure(i);
  }
  for (int t = 0; t < 4; t++) {
    len[t] = info->GetListLength(t);
    for (int k = 0; k < len[t] && k < cap; k++) {
      Pair pr = info->GetPair(t, k);
      lists[(t * cap + k) * 2 + 0] = pr.a;
      lists[(t * cap + k) * 2 + 1] = pr.b;
    }
  }
  return p.nreplica;
}

}  // extern "C"
