/*
 * remd_gpu.h -- C ABI of the MI355X REMD per-replica GPU path.
 *
 * Drop-in boundary for the reference's `class REMDGPU` (src/remdgpu.h:92-238):
 * plain pointers, sizes and status codes; no C++ or framework types.  Every
 * entry point returns 0 on success and a negative REMD_E* code on failure;
 * remd_gpu_last_error() gives the message (the reference prints and exit()s,
 * src/remdgpu.cu:208-212, 646-649; here exit() is left to the driver).
 *
 * Units are the reference's reduced units (src/unit.h:3-18): A, kJ/mol,
 * oxygen mass; temperature T/unit_temp, pressure P/unit_press, time
 * dt_fs*1e-15/unit_time.  Molecule state uses the reference's internal
 * representation (what IOManager::ReadCheckPoint leaves,
 * src/io_manager.cxx:577-581): r fractional in [0,1), v = velocity * L * s,
 * p = quaternion momentum * s.  Arrays are replica-major AoS:
 * r[(rep*nmol + i)*3 + d], q[(rep*nmol + i)*4 + d].
 *
 * Threading: one host thread per handle; calls on one handle are blocking
 * and ordered (as REMDGPU's).  Handles may share a batch (remd_batch_*).
 */
#ifndef REMD_GPU_H
#define REMD_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REMD_GPU_ABI_VERSION 1

enum {
  REMD_OK = 0,
  REMD_EINVAL = -1,   /* invalid argument / unsupported configuration */
  REMD_EHIP = -2,     /* HIP runtime error */
  REMD_ENOMEM = -3,   /* device or host allocation failed / budget exceeded */
  REMD_ESTATE = -4,   /* call not valid in the current state */
};

enum { REMD_PREC_MIXED = 0, REMD_PREC_FP64 = 1 };

/* One REMD job: the fields REMDGPU::REMDGPU reads from Molecules/Parameter
 * (src/remdgpu.cu:178-212; src/parameter.h:24-88; src/variables.h:159-207). */
typedef struct {
  int nmol;              /* molecules per replica, 2..4096 (the global-state path keeps
                            the site records of replicas beyond one workgroup's LDS
                            in global memory) */
  int nsite;             /* 3, 4 or 5 */
  int nreplica;
  int tconstant;         /* 0 NVE, 1 Nose-Poincare, 2 velocity scaling */
  int pconstant;         /* 0 none, 1 isotropic (bulk), 2 z (tube), 3 area (slit) */
  int confined;          /* 0 bulk, 1 nanotube, 2 nano-slit */
  int interval;          /* MD steps per exchange interval */
  int sort_interval;     /* neighbour rebuild period, 10 in the reference */
  double dt;             /* reduced time step */
  double rcut, rswitch;  /* A */
  double margin;         /* A, neighbour window margin (0.5 in the reference) */
  double wall_length;    /* A: tube radius or slit width */
  int nfwall;            /* wall table nodes */
  double sgm_wall, eps_wall, rho_wall; /* A, kJ/mol, A^-2 */
  /* molecule type, reduced units (src/molecules.cxx:98-110) */
  double site_r[5][3];
  double site_m[5], site_sigma[5], site_eps[5], site_q[5];
  double mass, inertia[3];
  int precision;         /* REMD_PREC_MIXED: FP32 pair math, FP64 accumulation */
  int block_threads;     /* 0: default; 128, 256 or 512 (global-state path) */
  int chunks;            /* j-chunks per molecule in the pair pass; 0: heuristic */
  int path;              /* REMD_PATH_AUTO, _RESIDENT (state in LDS, replicas may span
                            several workgroups), _SPLIT (as RESIDENT with the state in a
                            private global copy per workgroup, for replicas too large for
                            the LDS) or _GLOBAL (state in HBM, one workgroup per replica);
                            AUTO takes RESIDENT when the replica fits, else SPLIT when
                            that fits, else GLOBAL */
} remd_gpu_desc;

enum { REMD_PATH_AUTO = 0, REMD_PATH_RESIDENT = 1, REMD_PATH_GLOBAL = 2, REMD_PATH_SPLIT = 3 };

/* Per-replica extended variables (src/variables.h:137-157) and box. */
typedef struct {
  double s, Ps, Q;
  double Pv[3], W;
  double L[3];
  double T, P;           /* reduced conditions */
} remd_gpu_scalars;

typedef struct remd_gpu remd_gpu;

/* REMDGPU::REMDGPU (src/remdgpu.cu:178-503): uploads the replicas, builds the
 * wall table on the device, sorts, computes forces and sets H0 = total
 * energy (the reference's dt = 0 interval, :461-478).  ids may be NULL. */
int remd_gpu_create(remd_gpu **out, const remd_gpu_desc *desc, const remd_gpu_scalars *sc,
                    const double *r, const double *v, const double *q, const double *p,
                    const int *ids);

/* REMDGPU::ExecuteSteps (src/remdgpu.cu:656-681): one exchange interval for
 * every replica in one launch.  T[rep], P[rep] are the replica's conditions
 * (reduced); a replica whose T or P differs from its current conditions is
 * first rescaled as REMD::ReplicaExchange does on the CPU path
 * (src/remd.cxx:526-536).  Blocking. */
int remd_gpu_execute_steps(remd_gpu *g, const double *T, const double *P);

/* As remd_gpu_execute_steps but with one launch per replica, serially (the
 * plain one-replica-per-launch form used as the batching baseline). */
int remd_gpu_execute_steps_plain(remd_gpu *g, const double *T, const double *P);

/* As remd_gpu_execute_steps, allocating nothing (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops).  The launch's scratch -- group
 * and abort words, the replica table, the split-global block table and the
 * exchange rows of replicas split over workgroups -- lives in one static
 * device array of 16 MiB per GPU (held until the process ends); a job that
 * needs more gets REMD_EINVAL.  The array keeps the last call's contents
 * and every call on one GPU shares it: calls that overlap on one GPU race.
 * Same results, bit for bit, as remd_gpu_execute_steps. */
int remd_gpu_execute_steps_noalloc(remd_gpu *g, const double *T, const double *P);

/* REMDGPU::GetAverages (src/remdgpu.h:208-233) in the Average enum order of
 * src/remdinfo.h:27-41: LJ CLMB WALL VOL TRA ROT T P Px Py Pz TOT DRF TSTs
 * TSTv BSTx BSTy BSTz DUMMY (T in K, pressures in Pa, as the CPU path). */
int remd_gpu_get_averages(remd_gpu *g, int rep, double out[19]);

/* pot = LJ + CLMB + WALL and volume per replica (src/remd.cxx:247-251). */
int remd_gpu_get_pot_vol(remd_gpu *g, double *pot, double *vol);

/* REMDGPU::ConvertVariablesToHost (src/remdgpu.cu:806-854) for one replica:
 * arrays of nmol rows in the current (sorted) order; any pointer may be NULL. */
int remd_gpu_download_state(remd_gpu *g, int rep, double *r, double *v, double *f, double *q,
                            double *p, double *n, int *ids, remd_gpu_scalars *sc);

/* Checkpoint of the job's exact device state at an interval boundary --
 * what REMDGPU::ConvertVariablesToHost (src/remdgpu.cu:806-854) hands to
 * IOManager::WriteCheckPoint (src/io_manager.cxx:593-634), plus the forces,
 * molecule ids, H0 and energies the text checkpoint drops -- as one opaque
 * blob of remd_gpu_checkpoint_bytes(g) bytes.  Loading it into a job
 * created with the same descriptor continues the run bit for bit (the
 * reference re-derives forces and H0 from its text checkpoint instead,
 * src/io_manager.cxx:528-591).  A blob of another shape, pair precision or
 * chunk count is refused (REMD_EINVAL). */
size_t remd_gpu_checkpoint_bytes(remd_gpu *g);
int remd_gpu_checkpoint_save(remd_gpu *g, void *buf, size_t bytes);
int remd_gpu_checkpoint_load(remd_gpu *g, const void *buf, size_t bytes);

/* REMDGPU::OutputCDV (src/remdgpu.cu:700-804): one .cdv text file per
 * replica, <prefix>R<rep:04>s<step:07>.cdv. */
int remd_gpu_write_cdv(remd_gpu *g, const char *prefix, unsigned step);

void remd_gpu_destroy(remd_gpu *g);
const char *remd_gpu_last_error(void);
int remd_gpu_abi_version(void);

/* ---- test and measurement hooks (no reference counterpart) ---- */
/* Force pass only on the resident state (CalcForcePot, src/molecules.cxx:
 * 531-629); do_sort rebuilds the window with the given margin. */
int remd_gpu_compute_forces(remd_gpu *g, int do_sort, double margin);
/* Keep per-site forces and the window of every force pass (debug buffers). */
int remd_gpu_set_debug(remd_gpu *g, int on);
int remd_gpu_get_site_forces(remd_gpu *g, int rep, double *af /* nmol*nsite*3 */);
/* Window of the last sort: the reference's js/je (src/molecules.h:77-78)
 * and full[i] = 1 where the window is the whole ring (may be NULL). */
int remd_gpu_get_window(remd_gpu *g, int rep, int *js, int *je, int *full);
/* Average duration (ms) of the interval kernel over the last
 * remd_gpu_execute_steps call, from HIP events on the job's stream. */
int remd_gpu_last_kernel_ms(remd_gpu *g, double *ms);
/* Phase timers of the last interval (REMD_PHASE_TIMERS builds; zeros
 * otherwise): out[16] in 100 MHz ticks accumulated over the job so far. */
int remd_gpu_phase_timers(remd_gpu *g, int rep, double out[16]);
/* Chosen launch shape. */
int remd_gpu_launch_shape(remd_gpu *g, int *block_threads, int *chunks, int *lds_bytes);
/* Path of the job and the workgroups per replica its last launch used. */
int remd_gpu_launch_path(remd_gpu *g, int *path, int *parts);
/* Measure each (block_threads, chunks) candidate with timed force passes and
 * keep the fastest (project autotuner). */
int remd_gpu_autotune(remd_gpu *g, int reps_per_candidate);

/* ForceCalculator::CalcWallForce (src/force_calculator.cxx:117-200) on the
 * device: confined 1 tube of radius R, 2 slit of width R; out[nfwall*2]. */
int remd_gpu_wall_table(int confined, int nfwall, double R, double sgm, double eps, double rho,
                        double *out);

/* Standalone neighbour build (src/molecules.cxx:541-585 semantics): for each
 * of nrow rows of n keys in [0,1): stable ascending order and the cyclic
 * window [js, je] of half-width w on the sorted keys. */
int remd_gpu_sort_window(const double *keys, int nrow, int n, double w, int *order, int *js,
                         int *je);

/* Device Metropolis exchange (REMD::ReplicaExchange, src/remd.cxx:486-538)
 * for one job, using the energies and volumes of the last interval on the
 * device and host-drawn uniforms in pair-list order; updates index/T/P and
 * flags the exchanged replicas for the next remd_gpu_execute_pending().
 * Returns the number of pairs in accepted (may be NULL). */
int remd_gpu_exchange_device(remd_gpu *g, int ntemp, int npress, long step, int *index, double *T,
                             double *P, const double *uniforms, int *accepted, int *npairs);
/* Runs an interval with the conditions already on the device (after
 * remd_gpu_exchange_device). */
int remd_gpu_execute_pending(remd_gpu *g);

/* ---- batching: several jobs, one launch per interval ---- */
typedef struct remd_batch remd_batch;
int remd_batch_create(remd_batch **out);
/* Adds a job (any model, path and pair precision). */
int remd_batch_add(remd_batch *b, remd_gpu *g);
int remd_batch_remove(remd_batch *b, remd_gpu *g);
/* Stage conditions for a job's next batched interval. */
int remd_gpu_set_conditions(remd_gpu *g, const double *T, const double *P);
/* One interval for every job of the batch: every resident-path replica in
 * one kernel launch (split over 1..8 workgroups each by the batch's packing
 * plan), then one launch per global-state kernel class.  The first
 * intervals after the batch changes time the candidate plans; results do
 * not depend on the plan. */
int remd_batch_execute(remd_batch *b);
/* remd_batch_execute in two halves, so that several batches (each on its
 * own stream) run concurrently: launch enqueues the interval and the result
 * copies and returns; poll returns 1 once they are done (0 before); wait
 * blocks until then and finishes the bookkeeping (abort check, timing,
 * autotuner).  Between launch and wait the batch's jobs must not be staged,
 * read or changed, and the batch must not be changed (REMD_ESTATE). */
int remd_batch_launch(remd_batch *b);
int remd_batch_poll(remd_batch *b);
int remd_batch_wait(remd_batch *b);
/* Time candidate packing plans on the first intervals (default on); off:
 * keep the cost model's best plan. */
int remd_batch_set_autotune(remd_batch *b, int on);
/* Fix the packing plan: workgroups per replica for each job, in the order
 * the jobs were added (1..min(8, nmol / 8) for a resident-path job, 1 for
 * a global-state job); parts == NULL restores the automatic plan.  Adding
 * or removing a job also restores it. */
int remd_batch_set_parts(remd_batch *b, const int *parts, int n);
int remd_batch_last_kernel_ms(remd_batch *b, double *ms);
void remd_batch_destroy(remd_batch *b);
/* Device bytes a job of this shape keeps resident (admission control). */
size_t remd_gpu_footprint(const remd_gpu_desc *desc);

/* glibc-compatible rand() stream per job (the reference's exchange RNG,
 * src/remd.cxx:465), reproducible per job regardless of other jobs. */
typedef struct { int r[34]; int k; } remd_rng;
void remd_rng_seed(remd_rng *s, unsigned seed);
int remd_rng_next(remd_rng *s);

/* Exchange pair list `type` 0..3 (REMDInfo::ExchangeList,
 * src/remdinfo.cxx:54-108) of an ntemp x npress grid: the condition slots
 * a[k], b[k] of pair k in list order (either may be NULL); returns the
 * number of pairs. */
int remd_exchange_pairs(int ntemp, int npress, int type, int *a, int *b);

/* nint exchange intervals back to back on the device, with no host round
 * trip between them: the interval loop of REMD::Run
 * (src/remd.cxx:241-301: ExecuteSteps, GetAverages, ReplicaExchange,
 * :486-538) for one job.  Interval k runs at the conditions the exchange
 * after interval k-1 left (the first at T, P as given, as
 * remd_gpu_execute_steps does); then the Metropolis exchange of step
 * step0 + k (pair list (step0 + k) % 4) runs on the device with the
 * uniforms remd_exchange_host would draw from rng (rand()/RAND_MAX, one per
 * pair in list order), drawn on the host before the first launch; rng
 * advances by as many draws.  The same decisions as nint rounds of
 * remd_gpu_execute_steps + remd_gpu_get_pot_vol + remd_exchange_host.
 * index (by condition slot), T, P (by replica): in/out.  accepted (may be
 * NULL): per-pair flags of every interval in order; averages (may be
 * NULL): nint x nreplica x 19, each interval's GetAverages; kernel_ms (may
 * be NULL): nint interval kernel times.  Returns the number of accepted
 * swaps, or a negative error. */
int remd_gpu_execute_intervals(remd_gpu *g, int ntemp, int npress, long step0, int nint, int *index,
                               double *T, double *P, remd_rng *rng, int *accepted, double *averages,
                               double *kernel_ms);

/* REMD::ReplicaExchange (src/remd.cxx:486-538) on the host for one job:
 * pair list step%4 (src/remdinfo.cxx:54-108), Metropolis with one
 * rng draw per pair in list order, REMDInfo::ExchangeAccepted/Rejected
 * (src/remdinfo.h:276-359) on index/T/P (reduced; T, P indexed by
 * replica, index by condition slot).  is_exchanged[rep] (may be NULL) and
 * accepted[pair] (may be NULL) report the outcome; returns the number of
 * accepted swaps or a negative error. */
int remd_exchange_host(int ntemp, int npress, long step, int pconstant, const double *pot,
                       const double *vol, int *index, double *T, double *P,
                       unsigned char *is_exchanged, remd_rng *rng, int *accepted, int *npairs);

#ifdef __cplusplus
}
#endif
#endif
