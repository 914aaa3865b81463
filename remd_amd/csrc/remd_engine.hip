// remd_engine.hip -- host side of the C ABI in include/remd_gpu.h.
//
// A remd_gpu handle owns one REMD job's device-resident state (the
// reference's REMDGPU, src/remdgpu.h:92-238): molecule blocks, scalar blocks,
// the wall table and the per-replica descriptors the persistent kernel walks.
// A remd_batch packs the replicas of several handles into one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/remd_gpu.h"
#include "remd_common.h"

namespace remd {
hipError_t launch_interval(const LaunchArgs &a, int nt, int nsite, int prec, int lds, hipStream_t st);
hipError_t set_kernel_lds_limit(int bytes);
hipError_t launch_resident(const LaunchArgs &a, int nsite, int prec, int ens, int lds, hipStream_t st);
hipError_t resident_occupancy(int nsite, int prec, int ens, int lds, int *per_cu);
hipError_t launch_resident_batch(const LaunchArgs &a, int lds, hipStream_t st);
hipError_t resident_batch_occupancy(int lds, int *per_cu);
int resident_lds(int nmol, int nsite, int prec, int site_dbg, int gm);
hipError_t launch_wall_tube(int nfwall, double R, double sgm06, double coefE, double *fw, hipStream_t st);
hipError_t launch_wall_slit(int nfwall, double H, double sgm, double eps, double rho, double *fw,
                            hipStream_t st);
hipError_t launch_sort_window(const double *keys, int nrow, int n, double w, int *order, int *js, int *je,
                              hipStream_t st);
hipError_t launch_exchange(const ExchangeJob *jobs, int njobs, hipStream_t st);
hipError_t launch_scratch_setup(const ReplicaDesc *src, ReplicaDesc *dst, int nrep, double *xbase,
                                size_t xstride, unsigned long long *sync, int nsync, unsigned *blk, int nb,
                                int parts, unsigned variant, hipStream_t st);
int resident_grid(int nrep, int parts);
int lds_bytes_host(int nmol, int nsite, int prec, int nt, bool grec = false);
}  // namespace remd

using namespace remd;

// ------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(REMD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

extern "C" const char *remd_gpu_last_error(void) { return g_err.c_str(); }
extern "C" int remd_gpu_abi_version(void) { return REMD_GPU_ABI_VERSION; }

// ------------------------------------------------------------- units
// src/unit.h:3-18
static const double kNa = 6.0221367e+23, kKB = 1.3806488e-23;
static double unit_energy() { return 1e3 / kNa; }
static double unit_temp() { return unit_energy() / kKB; }
static double unit_press() { return unit_energy() / (1e-10 * 1e-10 * 1e-10); }

namespace remd {
static inline int align16h(int x) { return (x + 15) & ~15; }
int lds_bytes_host(int nmol, int nsite, int prec, int nt, bool grec) {
  const int stride = (nsite == 5) ? 18 : (nsite == 4 ? 16 : 12);
  const int rec = grec ? 0 : nmol * stride * (prec ? 8 : 4);
  const int sortb = 20 * nmol;
  const int region0 = align16h(rec > sortb ? rec : sortb);
  const int win = align16h(4 * nmol);
  const int misc = (nt / 64) * 32 * 8 + (16 * 64 + 16) * 4;
  return region0 + win + align16h(misc);
}
}  // namespace remd

// Bytes of one replica's site records (FP32 or FP64 site positions and
// body-frame offsets, src/remdgpu.cu:2607-2651 layout stride).
static size_t rec_bytes(int nmol, int nsite, int prec) {
  const int stride = (nsite == 5) ? 18 : (nsite == 4 ? 16 : 12);
  return (size_t)nmol * stride * (prec ? 8 : 4);
}

static int device_lds_limit() {
  static int lim = -1;
  if (lim < 0) {
    int dev = 0;
    hipGetDevice(&dev);
    int v = 0, o = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 65536;
    if (hipDeviceGetAttribute(&o, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && o > v) v = o;
    lim = v;
    set_kernel_lds_limit(lim);
  }
  return lim;
}

// -------------------------------------------------------------- job
// Largest number of workgroups one replica is split over (resident path).
constexpr int GMAX = 8;

struct remd_gpu {
  remd_gpu_desc desc;
  JobParams prm;
  int device = 0;            // HIP device the job was created on
  int nrep = 0, N = 0, NS = 0;
  int nt = 256, lds = 0;
  bool resident = false;     // state in LDS for the launch (resident_kernel)
  bool gm = false;           // resident kernel, split-global variant: state in per-part global copies
  double *d_gbuf = nullptr;  // split-global: nrep x GMAX x gm_part_doubles(N)
  char *d_grec = nullptr;    // global-state path: site records in global memory (rec_bytes each)
  bool grec = false;         // the global-state path's site records do not fit the LDS
  unsigned *d_blk1 = nullptr;  // split-global: workgroup table of the last (nrep, parts) launch shape
  int blk1_nrep = 0, blk1_parts = 0, blk1_n = 0;
  int capacity = 0;          // resident workgroups the device holds at once
  double wfrac = 1.0;        // window fraction of the sort axis, ~2 (rcut + margin) / L
  int last_parts = 1;
  unsigned long long *d_sync = nullptr;  // nrep group words + the abort word
  double *d_xbuf = nullptr;    // nrep x 2 x xbuf_parity_doubles(N) exchange rows
  unsigned *h_abort = nullptr; // pinned copy of the abort word
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // device
  JobParams *d_job = nullptr;
  double *d_mol = nullptr;
  int *d_mid = nullptr;
  double *d_sc = nullptr;
  double *d_fwall = nullptr;
  double *d_site = nullptr;
  int *d_win = nullptr;
  ReplicaDesc *d_reps = nullptr;
  // host
  std::vector<ReplicaDesc> h_reps;
  double *h_sc = nullptr;  // pinned mirror of the scalar blocks
  bool debug = false;
  // device exchange buffers
  int *d_index = nullptr;
  double *d_T = nullptr, *d_P = nullptr, *d_u = nullptr;
  double **d_scptr = nullptr;
  int *d_acc = nullptr;
  ExchangeJob *d_xjob = nullptr;
  // remd_gpu_execute_intervals: grow-only buffers and events, kept per job
  double *xi_u = nullptr, *xi_hist = nullptr;
  int *xi_acc = nullptr;
  ExchangeJob *xi_jobs = nullptr;
  size_t xi_ucap = 0, xi_acap = 0, xi_hcap = 0, xi_jcap = 0;
  std::vector<hipEvent_t> xi_ev;
};

// Every resident replica of a batch runs in ONE launch of
// resident_batch_kernel (any water model and precision).  Global-state
// replicas are launched per class = (workgroup size, sites, pair precision),
// each class on its own stream, after the resident launch, so a long
// global-state launch never holds the CUs a waiting part needs.
constexpr int NCLASS = 12;
static int kclass(int nsite, int prec, int nt) { return (nt == 512 ? 6 : 0) + (nsite - 3) * 2 + prec; }
static int class_nt(int c) { return c >= 6 ? 512 : 256; }

struct remd_batch {
  std::vector<remd_gpu *> jobs;
  ReplicaDesc *d_table = nullptr;  // resident rows (heaviest first), then the global-state classes
  int table_cap = 0;
  int ntab = 0;
  bool dirty = true;
  // resident rows
  int res_n = 0, res_lds = 0, res_capacity = 0;
  std::vector<int> row_job;  // resident row -> index into jobs
  // global-state classes
  int cls_off[NCLASS] = {0}, cls_n[NCLASS] = {0}, cls_lds[NCLASS] = {0};
  // packing plan: workgroups per replica of each job, and its block table
  std::vector<int> plan;
  unsigned *d_blk = nullptr;
  int blk_cap = 0, nblk = 0;
  bool split = false;  // some replica of the plan has parts > 1
  // autotuner: candidate plans (the cost model's best estimates first, then
  // hill climbing around the fastest timed plan), each timed on one interval
  // after an untimed one; the fastest is kept until the batch changes.
  // Results do not depend on the plan.
  std::vector<std::vector<int>> cands;
  std::vector<double> cand_ms;
  int cand_cur = 0;
  int warm = 0;
  bool tuned = true;
  int autotune = 1;
  std::vector<int> forced;  // plan fixed by remd_batch_set_parts (empty: automatic)
  unsigned long long *d_sync = nullptr;  // res_n group words + the abort word
  unsigned *h_abort = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t cstream[NCLASS] = {nullptr};
  hipEvent_t cdone[NCLASS] = {nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, efork = nullptr;
  hipEvent_t edone = nullptr;  // end of a launch's result copies (remd_batch_poll)
  bool in_flight = false;      // launched, not yet waited for
  double last_ms = 0.0;
};

// Pair-pass chunk count: balance passes over the workgroup against the
// per-lane window length (estimated from density along the sort axis).
static int pick_chunks(int N, int nt, double window_frac) {
  const double wcount = std::max(1.0, window_frac * N);
  int best = 1;
  double best_cost = 1e300;
  for (int C = 1; C <= 16; C *= 2) {
    const int passes = (N * C + nt - 1) / nt;
    const double cost = passes * std::ceil(wcount / C) + 6.0 * std::log2((double)C) * passes;
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      best = C;
    }
  }
  return best;
}

static int validate(const remd_gpu_desc *d, const remd_gpu_scalars *sc) {
  if (!d || !sc) return fail(REMD_EINVAL, "null descriptor");
  if (d->nsite < 3 || d->nsite > 5) return fail(REMD_EINVAL, "nsite must be 3, 4 or 5");
  if (d->nmol < 2 || d->nmol > 4096) return fail(REMD_EINVAL, "nmol must be in [2, 4096]");
  if (d->nreplica < 1) return fail(REMD_EINVAL, "nreplica must be >= 1");
  if (d->tconstant < 0 || d->tconstant > 2) return fail(REMD_EINVAL, "tconstant must be 0, 1 or 2");
  if (d->confined < 0 || d->confined > 2) return fail(REMD_EINVAL, "confined must be 0, 1 or 2");
  // valid ensemble/confinement pairs (src/remd.cxx:35-36 and GetBottomArea)
  const int P = d->pconstant, Cm = d->confined;
  const bool okp = (P == 0) || (Cm == 0 && P == 1) || (Cm == 1 && P == 2) || (Cm == 2 && P == 3);
  if (!okp) return fail(REMD_EINVAL, "pconstant not valid for this confinement (bulk:1, tube:2, slit:3)");
  if (d->interval < 0) return fail(REMD_EINVAL, "interval must be >= 0");
  if (d->sort_interval < 1) return fail(REMD_EINVAL, "sort_interval must be >= 1");
  if (!(d->rcut > d->rswitch) || d->rswitch <= 0) return fail(REMD_EINVAL, "need 0 < rswitch < rcut");
  if (Cm > 0 && d->nfwall < 3) return fail(REMD_EINVAL, "confined systems need nfwall >= 3");
  if (Cm > 0 && !(d->wall_length > 0)) return fail(REMD_EINVAL, "wall_length must be > 0");
  if (!(d->site_eps[0] > 0)) return fail(REMD_EINVAL, "site 0 must carry the LJ interaction");
  for (int a = 1; a < d->nsite; a++)
    if (d->site_eps[a] != 0.0) return fail(REMD_EINVAL, "only site 0 may carry LJ parameters");
  if (d->nsite >= 4 && d->site_q[0] != 0.0)
    return fail(REMD_EINVAL, "4/5-site models must have an uncharged site 0");
  if (d->block_threads != 0 && d->block_threads != 128 && d->block_threads != 256 && d->block_threads != 512)
    return fail(REMD_EINVAL, "block_threads must be 0, 128, 256 or 512");
  if (d->chunks < 0 || d->chunks > 64) return fail(REMD_EINVAL, "chunks must be in [0, 64]");
  if (d->path < 0 || d->path > 3)
    return fail(REMD_EINVAL, "path must be 0 (auto), 1 (resident), 2 (global) or 3 (split)");
  if (d->precision != 0 && d->precision != 1) return fail(REMD_EINVAL, "precision must be 0 or 1");
  const int axis = (Cm == 2) ? 0 : 2;
  for (int r = 0; r < d->nreplica; r++) {
    const double *L = sc[r].L;
    if (!(sc[r].s > 0) || !(sc[r].Q > 0) || !(sc[r].W > 0)) return fail(REMD_EINVAL, "s, Q and W must be > 0");
    if (!(sc[r].T > 0)) return fail(REMD_EINVAL, "T must be > 0");
    if (!(L[axis] > 2.0 * (d->rcut + d->margin)))
      return fail(REMD_EINVAL, "box along the sort axis must exceed 2*(rcut+margin)");
    const int per[3] = {Cm == 0 || Cm == 2, Cm == 0 || Cm == 2, Cm == 0 || Cm == 1};
    for (int k = 0; k < 3; k++)
      if (per[k] && !(L[k] > 2.0 * d->rcut)) return fail(REMD_EINVAL, "periodic box length must exceed 2*rcut");
  }
  return 0;
}

static void fill_params(const remd_gpu_desc *d, JobParams &J) {
  memset(&J, 0, sizeof(J));
  J.nmol = d->nmol;
  J.nsite = d->nsite;
  J.qbeg = (d->nsite == 3) ? 0 : 1;
  J.tmode = d->tconstant;
  J.pmode = d->pconstant;
  J.cmode = d->confined;
  J.sort_axis = (d->confined == 2) ? 0 : 2;
  J.per[0] = J.per[1] = (d->confined == 0 || d->confined == 2);
  J.per[2] = (d->confined == 0 || d->confined == 1);
  J.interval = d->interval;
  J.sort_interval = d->sort_interval;
  J.nfwall = d->nfwall;
  J.precision = d->precision;
  // bulk and slit windows are slabs of the box, mostly beyond the cutoff:
  // screen the partners first; tube windows are mostly inside it
  J.compact = d->confined == 1 ? 0 : 1;
  // the rank sort's O(N^2 / threads) comparisons lose to the radix sort's
  // 16 passes from about this size (measured); REMD_SORT_RADIX_MIN overrides
  {
    const char *e = std::getenv("REMD_SORT_RADIX_MIN");
    J.radix_min = e ? std::atoi(e) : RADIX_MIN_DEFAULT;
  }
  // REMD_ENSEMBLE_FIXED=0 runs every job on the instantiation that reads
  // the modes from the job (the tests compare the two)
  {
    const char *e = std::getenv("REMD_ENSEMBLE_FIXED");
    J.ens_fixed = (e && std::atoi(e) == 0) ? 0 : 1;
  }
  J.dt = d->dt;
  J.dthalf = d->dt * 0.5;
  J.rcut = d->rcut;
  J.rswitch = d->rswitch;
  J.margin = d->margin;
  J.wall_length = d->wall_length;
  {
    const double sigma_c = 3.4;  // src/molecules.h:369-371
    const double wl = (d->wall_length - 0.5 * sigma_c);
    J.bottom_area = M_PI * wl * wl;
  }
  double m = 0;
  for (int a = 0; a < d->nsite; a++) m += d->site_m[a];
  J.mass = d->mass > 0 ? d->mass : m;
  for (int k = 0; k < 3; k++) J.inertia[k] = d->inertia[k];
  J.inv_mass = 1.0 / J.mass;
  for (int k = 0; k < 3; k++) J.inv_inertia[k] = J.inertia[k] != 0.0 ? 1.0 / J.inertia[k] : 0.0;
  for (int a = 0; a < d->nsite; a++) {
    for (int k = 0; k < 3; k++) J.site_r[a][k] = d->site_r[a][k];
    J.q[a] = d->site_q[a];
  }
  // MakeParamList diagonal (src/force_calculator.cxx:226-228): sigma of the
  // pair = (s_i + s_j)/2, epsilon = sqrt(e_i e_j); 4 eps as the tuned kernel
  const double sg = 0.5 * (d->site_sigma[0] + d->site_sigma[0]);
  J.lj_sig2 = sg * sg;
  J.lj_eps4 = 4.0 * std::sqrt(d->site_eps[0] * d->site_eps[0]);
  J.unit_temp = unit_temp();
  J.unit_press = unit_press();
  {
    const double rlc = d->rswitch - d->rcut;
    const double coef = 1.0 / (rlc * rlc * rlc * rlc * rlc);
    J.f_rc = (float)d->rcut;
    J.f_rl = (float)d->rswitch;
    J.f_rc2 = (float)(d->rcut * d->rcut);
    J.f_rl2 = (float)(d->rswitch * d->rswitch);
    J.f_coef = (float)coef;
    J.f_coef30 = (float)(30.0 * coef);
    J.f_sig2 = (float)J.lj_sig2;
    J.f_eps4 = (float)J.lj_eps4;
    J.f_eps48 = (float)(12.0 * J.lj_eps4);
    for (int a = 0; a < MAXSITE; a++)
      for (int b = 0; b < MAXSITE; b++) J.f_qq[a][b] = (float)(J.q[a] * J.q[b]);
  }
}

extern "C" size_t remd_gpu_footprint(const remd_gpu_desc *d) {
  if (!d) return 0;
  const size_t N = (size_t)d->nmol, R = (size_t)d->nreplica;
  size_t b = R * NFIELD * N * sizeof(double) + R * N * sizeof(int) + R * NSC * sizeof(double);
  b += (size_t)std::max(d->nfwall, 0) * 2 * sizeof(double) + R * sizeof(ReplicaDesc) + sizeof(JobParams);
  b += R * (sizeof(double) * 3 + sizeof(int) * 2 + sizeof(double *)) + 4096;
  // resident path (an upper bound when the job takes the global-state path):
  // exchange rows, group words, abort word
  b += R * 2 * xbuf_parity_doubles((int)N) * sizeof(double) + (R + 1) * sizeof(unsigned long long);
  // split-global private copies (replicas of 512 or more molecules may take that path)
  if (N >= 512) b += R * GMAX * gm_part_doubles((int)N) * sizeof(double);
  // global-state path site records in global memory (replicas beyond the LDS)
  if (N >= 1024) b += R * rec_bytes((int)N, d->nsite, d->precision);
  return b;
}

static void free_job(remd_gpu *g) {
  if (!g) return;
  hipFree(g->d_job);
  hipFree(g->d_mol);
  hipFree(g->d_mid);
  hipFree(g->d_sc);
  hipFree(g->d_fwall);
  hipFree(g->d_site);
  hipFree(g->d_win);
  hipFree(g->d_reps);
  hipFree(g->d_index);
  hipFree(g->d_T);
  hipFree(g->d_P);
  hipFree(g->d_u);
  hipFree(g->d_scptr);
  hipFree(g->d_acc);
  hipFree(g->d_xjob);
  hipFree(g->xi_u);
  hipFree(g->xi_hist);
  hipFree(g->xi_acc);
  hipFree(g->xi_jobs);
  for (hipEvent_t e : g->xi_ev) hipEventDestroy(e);
  g->xi_ev.clear();
  hipFree(g->d_sync);
  hipFree(g->d_xbuf);
  hipFree(g->d_gbuf);
  hipFree(g->d_grec);
  hipFree(g->d_blk1);
  if (g->h_abort) hipHostFree(g->h_abort);
  if (g->h_sc) hipHostFree(g->h_sc);
  if (g->ev0) hipEventDestroy(g->ev0);
  if (g->ev1) hipEventDestroy(g->ev1);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
}

static int build_wall(remd_gpu *g) {
  const remd_gpu_desc &d = g->desc;
  if (d.confined == 0) return 0;
  HIPCHK(hipMalloc(&g->d_fwall, sizeof(double) * 2 * d.nfwall));
  if (d.confined == 1) {
    // float powf of sigma as the reference (force_calculator.cxx:123)
    const double sgm06 = (double)powf((float)d.sgm_wall, 6.0f);
    const double coefE = M_PI * d.rho_wall * d.eps_wall * sgm06;
    HIPCHK(launch_wall_tube(d.nfwall, d.wall_length, sgm06, coefE, g->d_fwall, g->stream));
  } else {
    HIPCHK(launch_wall_slit(d.nfwall, d.wall_length, d.sgm_wall, d.eps_wall, d.rho_wall, g->d_fwall, g->stream));
  }
  return 0;
}

static void refresh_reps(remd_gpu *g) {
  g->h_reps.resize(g->nrep);
  for (int r = 0; r < g->nrep; r++) {
    ReplicaDesc &R = g->h_reps[r];
    R.job = g->d_job;
    R.mol = g->d_mol + (size_t)r * NFIELD * g->N;
    R.mid = g->d_mid + (size_t)r * g->N;
    R.sc = g->d_sc + (size_t)r * NSC;
    R.site_f = g->debug ? g->d_site + (size_t)r * g->N * g->NS * 3 : nullptr;
    R.win = g->debug ? g->d_win + (size_t)r * g->N * 3 : nullptr;
    R.xbuf = g->d_xbuf ? g->d_xbuf + (size_t)r * 2 * xbuf_parity_doubles(g->N) : nullptr;
    R.gbuf = g->d_gbuf ? g->d_gbuf + (size_t)r * GMAX * gm_part_doubles(g->N) : nullptr;
    R.grec = g->d_grec ? (void *)(g->d_grec + (size_t)r * rec_bytes(g->N, g->NS, g->prm.precision)) : nullptr;
  }
}

static int upload_reps(remd_gpu *g) {
  refresh_reps(g);
  HIPCHK(hipMemcpyAsync(g->d_reps, g->h_reps.data(), sizeof(ReplicaDesc) * g->nrep, hipMemcpyHostToDevice,
                        g->stream));
  return 0;
}

// The same-XCD group barrier (L1-only invalidation) may be switched off
// with REMD_SAME_XCD_BARRIER=0; the parts then always use agent-scope fences.
static int same_xcd_barrier_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = std::getenv("REMD_SAME_XCD_BARRIER");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v;
}

static int same_xcd_barrier_mode() {
  static int v = -1;
  if (v < 0) {
    const char *e = std::getenv("REMD_BARRIER_MODE");
    v = e ? (std::atoi(e) & 3) : 0;
  }
  return v;
}

// Workgroups per replica for a resident launch of nrep replicas that has
// `capacity` co-resident workgroup slots to itself: every part of every
// replica must be resident at once (the parts wait for each other), and a
// part keeps at least 8 molecules.
static int pick_parts(int capacity, int nrep, int N) {
  if (nrep <= 0) return 1;
  int p = capacity / nrep;
  p = std::min(p, GMAX);
  p = std::min(p, std::max(1, N / 8));
  return p >= 2 ? p : 1;
}

// Ensemble code of the job (selects the resident-kernel instantiation).
static int job_ens(const remd_gpu *g) {
  return g->prm.ens_fixed ? ens_code(g->prm.tmode, g->prm.pmode, g->prm.cmode) : ENS_RUNTIME;
}

// A resident-path launch of nrep replicas with a.parts workgroups each: the
// LDS kernel, or for the split-global variant the table-driven kernel with
// the same workgroup mapping (parts of a replica on block indices equal
// mod 8).
static int launch_res(remd_gpu *g, LaunchArgs &a, hipStream_t st) {
  if (!g->gm) {
    HIPCHK(launch_resident(a, g->NS, g->prm.precision, job_ens(g), g->lds, st));
    return 0;
  }
  const int parts = a.parts < 1 ? 1 : a.parts;
  if (g->blk1_nrep != a.nrep || g->blk1_parts != parts || !g->d_blk1) {
    const int nb = parts <= 1 ? a.nrep : 8 * parts * ((a.nrep + 7) / 8);
    std::vector<unsigned> blk(nb, ~0u);
    const unsigned v = (unsigned)(6 + (g->NS - 3) * 2 + g->prm.precision);
    for (int b = 0; b < nb; b++) {
      int rep, part;
      if (parts <= 1) {
        rep = b;
        part = 0;
      } else {
        const int slot = b >> 3;
        rep = (slot / parts) * 8 + (b & 7);
        part = slot % parts;
      }
      if (rep < a.nrep) blk[b] = (unsigned)rep << 12 | v << 8 | (unsigned)parts << 4 | (unsigned)part;
    }
    HIPCHK(hipStreamSynchronize(st));
    hipFree(g->d_blk1);
    g->d_blk1 = nullptr;
    HIPCHK(hipMalloc(&g->d_blk1, sizeof(unsigned) * nb));
    HIPCHK(hipMemcpy(g->d_blk1, blk.data(), sizeof(unsigned) * nb, hipMemcpyHostToDevice));
    g->blk1_nrep = a.nrep;
    g->blk1_parts = parts;
    g->blk1_n = nb;
  }
  a.blk = g->d_blk1;
  a.nblk = g->blk1_n;
  HIPCHK(launch_resident_batch(a, g->lds, st));
  return 0;
}

static int launch(remd_gpu *g, int mode, int do_sort, double margin, const ReplicaDesc *reps, int nrep,
                  bool timed) {
  LaunchArgs a;
  a.reps = reps;
  a.nrep = nrep;
  a.mode = mode;
  a.do_sort = do_sort;
  a.margin = margin;
  a.parts = 1;
  a.same_xcd_ok = same_xcd_barrier_enabled();
  a.same_xcd_mode = same_xcd_barrier_mode();
  a.sync = g->d_sync;
  a.abort = g->d_sync ? (unsigned *)(g->d_sync + g->nrep) : nullptr;
  if (g->resident) {
    a.parts = pick_parts(g->capacity, nrep, g->N);
    if (a.parts > 1) HIPCHK(hipMemsetAsync(g->d_sync, 0, sizeof(unsigned long long) * (g->nrep + 1), g->stream));
  }
  g->last_parts = a.parts;
  if (timed) HIPCHK(hipEventRecord(g->ev0, g->stream));
  if (g->resident) {
    int rc = launch_res(g, a, g->stream);
    if (rc) return rc;
  } else {
    HIPCHK(launch_interval(a, g->nt, g->NS, g->prm.precision, g->lds, g->stream));
  }
  if (timed) HIPCHK(hipEventRecord(g->ev1, g->stream));
  if (a.parts > 1)
    HIPCHK(hipMemcpyAsync(g->h_abort, a.abort, sizeof(unsigned), hipMemcpyDeviceToHost, g->stream));
  return 0;
}

// Stream sync + the launch's abort word: a resident launch whose parts could
// not all be resident gives up (no state is written back) instead of waiting.
static int finish(remd_gpu *g) {
  HIPCHK(hipStreamSynchronize(g->stream));
  if (g->last_parts > 1 && *g->h_abort != 0u) {
    *g->h_abort = 0u;
    return fail(REMD_EHIP, "resident launch aborted: the parts of a replica were not co-resident "
                           "(another kernel held the device)");
  }
  return 0;
}

static int sync_scalars(remd_gpu *g) {
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_sc, sizeof(double) * NSC * g->nrep, hipMemcpyDeviceToHost, g->stream));
  return finish(g);
}

extern "C" int remd_gpu_create(remd_gpu **out, const remd_gpu_desc *d, const remd_gpu_scalars *sc,
                               const double *r, const double *v, const double *q, const double *p,
                               const int *ids) {
  if (!out) return fail(REMD_EINVAL, "null output handle");
  *out = nullptr;
  int rc = validate(d, sc);
  if (rc) return rc;
  if (!r || !v || !q || !p) return fail(REMD_EINVAL, "null state array");
  remd_gpu *g = new remd_gpu;
  g->desc = *d;
  g->nrep = d->nreplica;
  g->N = d->nmol;
  g->NS = d->nsite;
  fill_params(d, g->prm);
  // global-state path: 512-thread workgroups (two waves per SIMD) for
  // replicas of 512 molecules or more, 256 below
  g->nt = d->block_threads ? d->block_threads : (d->nmol >= 512 ? 512 : 256);
  const int lim = device_lds_limit();
  // window fraction along the sort axis ~ 2 (rcut + margin) / L
  const double wf = std::min(1.0, 2.0 * (d->rcut + d->margin) / sc[0].L[g->prm.sort_axis]);
  const bool bt_ok = d->block_threads == 0 || d->block_threads == RNT;
  const int rlds = resident_lds(g->N, g->NS, g->prm.precision, 0, 0);
  const int glds = resident_lds(g->N, g->NS, g->prm.precision, 0, 1);
  const bool fits = bt_ok && rlds <= lim;
  const bool fits_gm = bt_ok && glds <= lim;
  if (d->path == REMD_PATH_RESIDENT && !fits) {
    free_job(g);
    return fail(REMD_EINVAL, "replica does not fit the resident path (" + std::to_string(rlds) + " > " +
                                 std::to_string(lim) + " bytes of LDS, or block_threads != " +
                                 std::to_string(RNT) + ")");
  }
  if (d->path == REMD_PATH_SPLIT && !fits_gm) {
    free_job(g);
    return fail(REMD_EINVAL, "replica does not fit the split-global path (" + std::to_string(glds) + " > " +
                                 std::to_string(lim) + " bytes of LDS, or block_threads != " +
                                 std::to_string(RNT) + ")");
  }
  // AUTO takes the split-global variant for a replica too large for the
  // LDS: in a batch it joins the one resident launch (the global-state
  // classes only start after it), and alone it splits over up to GMAX
  // workgroups
  g->gm = d->path == REMD_PATH_SPLIT || (d->path == REMD_PATH_AUTO && !fits && fits_gm);
  g->resident = d->path == REMD_PATH_RESIDENT || (d->path == REMD_PATH_AUTO && fits) || g->gm;
  if (g->resident) {
    g->nt = RNT;
    g->lds = g->gm ? glds : rlds;
    int per_cu = 0, ncu = 0, dev = 0;
    hipGetDevice(&dev);
    const hipError_t oe = g->gm ? resident_batch_occupancy(g->lds, &per_cu)
                                : resident_occupancy(g->NS, g->prm.precision, job_ens(g), g->lds, &per_cu);
    if (oe != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1 ||
        ncu < 1) {
      free_job(g);
      return fail(REMD_EHIP, "occupancy query for the resident kernel failed");
    }
    g->capacity = per_cu * ncu;
    g->wfrac = wf;
    // chunks: the (molecule, chunk) items of one part fill its 256 lanes at
    // the part count this job gets alone; fixed for the job's lifetime, so
    // results do not depend on the part count of a later launch
    const int gn = pick_parts(g->capacity, g->nrep, g->N);
    const int nb = (g->N + EB - 1) / EB;
    const int M = ((nb + gn - 1) / gn) * EB;  // molecules of the largest part
    const int wcount = std::max(1, (int)std::ceil(wf * g->N));
    g->prm.chunks = d->chunks ? d->chunks : std::max(1, std::min(std::min(64, RNT / M), wcount));
  } else {
    if (d->chunks && (d->chunks & (d->chunks - 1))) {
      free_job(g);
      return fail(REMD_EINVAL, "chunks must be a power of two on the global-state path");
    }
    // the default 512 threads need more reduction scratch than 256: a
    // replica at the LDS limit takes the smaller workgroup; beyond it the
    // site records move to global memory (read through the caches)
    if (!d->block_threads && g->nt == 512 && lds_bytes_host(g->N, g->NS, g->prm.precision, 512) > lim)
      g->nt = 256;
    if (lds_bytes_host(g->N, g->NS, g->prm.precision, g->nt) > lim) {
      g->grec = true;
      if (!d->block_threads && g->N >= 512) g->nt = 512;
    }
    g->prm.chunks = d->chunks ? d->chunks : pick_chunks(g->N, g->nt, wf);
    g->lds = lds_bytes_host(g->N, g->NS, g->prm.precision, g->nt, g->grec);
    if (g->lds > lim) {
      free_job(g);
      return fail(REMD_EINVAL, "replica too large for workgroup LDS (" + std::to_string(g->lds) + " > " +
                                   std::to_string(lim) + " bytes)");
    }
  }
#define CK(...)                        \
  do {                                 \
    int rc_ = (__VA_ARGS__);           \
    if (rc_) {                         \
      std::string m_ = g_err;          \
      free_job(g);                     \
      return fail(rc_, m_);            \
    }                                  \
  } while (0)
  CK([&]() -> int {
    HIPCHK(hipGetDevice(&g->device));
    HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&g->ev0));
    HIPCHK(hipEventCreate(&g->ev1));
    const size_t N = g->N, R = g->nrep;
    HIPCHK(hipMalloc(&g->d_job, sizeof(JobParams)));
    HIPCHK(hipMalloc(&g->d_mol, sizeof(double) * NFIELD * N * R));
    HIPCHK(hipMalloc(&g->d_mid, sizeof(int) * N * R));
    HIPCHK(hipMalloc(&g->d_sc, sizeof(double) * NSC * R));
    HIPCHK(hipMalloc(&g->d_reps, sizeof(ReplicaDesc) * R));
    HIPCHK(hipHostMalloc(&g->h_sc, sizeof(double) * NSC * R, hipHostMallocDefault));
    if (g->resident) {
      HIPCHK(hipMalloc(&g->d_sync, sizeof(unsigned long long) * (R + 1)));
      HIPCHK(hipMemsetAsync(g->d_sync, 0, sizeof(unsigned long long) * (R + 1), g->stream));
      HIPCHK(hipMalloc(&g->d_xbuf, sizeof(double) * R * 2 * xbuf_parity_doubles((int)N)));
      if (g->gm) HIPCHK(hipMalloc(&g->d_gbuf, sizeof(double) * R * GMAX * gm_part_doubles((int)N)));
      HIPCHK(hipHostMalloc(&g->h_abort, sizeof(unsigned), hipHostMallocDefault));
      *g->h_abort = 0u;
    }
    if (g->grec) HIPCHK(hipMalloc(&g->d_grec, R * rec_bytes((int)N, g->NS, g->prm.precision)));
    return 0;
  }());
  CK(build_wall(g));
  g->prm.fwall = g->d_fwall;
  CK([&]() -> int {
    HIPCHK(hipMemcpyAsync(g->d_job, &g->prm, sizeof(JobParams), hipMemcpyHostToDevice, g->stream));
    const size_t N = g->N, R = g->nrep;
    std::vector<double> mol(NFIELD * N * R, 0.0);
    std::vector<int> mid(N * R);
    for (size_t rr = 0; rr < R; rr++) {
      double *m = mol.data() + rr * NFIELD * N;
      for (size_t i = 0; i < N; i++) {
        const size_t k = rr * N + i;
        for (int dd = 0; dd < 3; dd++) {
          m[(F_RX + dd) * N + i] = r[3 * k + dd];
          m[(F_VX + dd) * N + i] = v[3 * k + dd];
        }
        for (int dd = 0; dd < 4; dd++) {
          m[(F_Q0 + dd) * N + i] = q[4 * k + dd];
          m[(F_P0 + dd) * N + i] = p[4 * k + dd];
        }
        mid[k] = ids ? ids[k] : (int)i;
      }
    }
    HIPCHK(hipMemcpyAsync(g->d_mol, mol.data(), sizeof(double) * mol.size(), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(g->d_mid, mid.data(), sizeof(int) * mid.size(), hipMemcpyHostToDevice, g->stream));
    for (size_t rr = 0; rr < R; rr++) {
      double *s = g->h_sc + rr * NSC;
      memset(s, 0, sizeof(double) * NSC);
      s[S_S] = sc[rr].s;
      s[S_PS] = sc[rr].Ps;
      s[S_Q] = sc[rr].Q;
      for (int dd = 0; dd < 3; dd++) s[S_PV0 + dd] = sc[rr].Pv[dd];
      s[S_W] = sc[rr].W;
      for (int dd = 0; dd < 3; dd++) s[S_L0 + dd] = sc[rr].L[dd];
      s[S_T] = s[S_TNEW] = sc[rr].T;
      s[S_P] = s[S_PNEW] = sc[rr].P;
    }
    HIPCHK(hipMemcpyAsync(g->d_sc, g->h_sc, sizeof(double) * NSC * R, hipMemcpyHostToDevice, g->stream));
    int rc2 = upload_reps(g);
    if (rc2) return rc2;
    // the staging vectors above are pageable and go out of scope here
    HIPCHK(hipStreamSynchronize(g->stream));
    return 0;
  }());
  CK(launch(g, MODE_INIT, 1, 0.0, g->d_reps, g->nrep, false));
  CK(sync_scalars(g));
#undef CK
  *out = g;
  return 0;
}

extern "C" void remd_gpu_destroy(remd_gpu *g) {
  if (g && g->stream) hipStreamSynchronize(g->stream);
  free_job(g);
}

extern "C" int remd_gpu_set_conditions(remd_gpu *g, const double *T, const double *P) {
  if (!g || !T || !P) return fail(REMD_EINVAL, "null argument");
  for (int r = 0; r < g->nrep; r++) {
    double *s = g->h_sc + (size_t)r * NSC;
    if (!(T[r] > 0)) return fail(REMD_EINVAL, "T must be > 0");
    const bool changed = (T[r] != s[S_T]) || (P[r] != s[S_P]);
    s[S_TNEW] = T[r];
    s[S_PNEW] = P[r];
    s[S_XFLAG] = changed ? 1.0 : 0.0;
  }
  HIPCHK(hipMemcpyAsync(g->d_sc, g->h_sc, sizeof(double) * NSC * g->nrep, hipMemcpyHostToDevice, g->stream));
  return 0;
}

extern "C" int remd_gpu_execute_steps(remd_gpu *g, const double *T, const double *P) {
  int rc = remd_gpu_set_conditions(g, T, P);
  if (rc) return rc;
  rc = launch(g, MODE_INTERVAL, 0, 0.0, g->d_reps, g->nrep, true);
  if (rc) return rc;
  rc = sync_scalars(g);
  if (rc) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  return 0;
}

// ---------------------------------------------- allocation-free interval
// remd_gpu_execute_steps_noalloc keeps the scratch of its launch -- group
// words and abort word, the replica table, the split-global block table and
// the per-step exchange rows of replicas split over workgroups -- in this
// file-scope device array (one copy per GPU, held until the process ends),
// so the call allocates nothing.  Host code reaches the array only through
// the address hipGetSymbolAddress returns for the device (kept per device
// below); the array's name is not a device address on the host.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAXDEV = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[MAXDEV];

// Device address of this device's scratch array, looked up once per device
// with that device current.
static int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= MAXDEV) return fail(REMD_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(REMD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" int remd_gpu_execute_steps_noalloc(remd_gpu *g, const double *T, const double *P) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  const int parts = g->resident ? pick_parts(g->capacity, g->nrep, g->N) : 1;
  const int nsync = g->nrep + 1;
  const int nb = g->gm ? resident_grid(g->nrep, parts) : 0;
  const size_t xstride = 2 * xbuf_parity_doubles(g->N);
  const size_t off_tab = align256(sizeof(unsigned long long) * nsync);
  const size_t off_blk = off_tab + align256(sizeof(ReplicaDesc) * g->nrep);
  const size_t off_x = off_blk + align256(sizeof(unsigned) * nb);
  const size_t need = off_x + (parts > 1 ? sizeof(double) * xstride * g->nrep : 0);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(REMD_EINVAL, "the job needs " + std::to_string(need) + " bytes of launch scratch; the static "
                             "scratch of remd_gpu_execute_steps_noalloc holds " +
                             std::to_string(NOALLOC_SCRATCH_BYTES) + " (use remd_gpu_execute_steps)");
  void *base = nullptr;
  int rc = scratch_address(g->device, &base);
  if (rc) return rc;
  rc = remd_gpu_set_conditions(g, T, P);
  if (rc) return rc;
  char *sb = (char *)base;
  unsigned long long *sync = (unsigned long long *)sb;
  ReplicaDesc *tab = (ReplicaDesc *)(sb + off_tab);
  unsigned *blk = (unsigned *)(sb + off_blk);
  double *xb = parts > 1 ? (double *)(sb + off_x) : nullptr;
  const unsigned variant = (unsigned)(6 + (g->NS - 3) * 2 + g->prm.precision);
  HIPCHK(launch_scratch_setup(g->d_reps, tab, g->nrep, xb, xstride, sync, nsync, nb ? blk : nullptr, nb, parts,
                              variant, g->stream));
  LaunchArgs a;
  a.reps = tab;
  a.nrep = g->nrep;
  a.mode = MODE_INTERVAL;
  a.do_sort = 0;
  a.margin = 0.0;
  a.parts = parts;
  a.same_xcd_ok = same_xcd_barrier_enabled();
  a.same_xcd_mode = same_xcd_barrier_mode();
  a.sync = sync;
  a.abort = (unsigned *)(sync + g->nrep);
  g->last_parts = parts;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if (g->gm) {
    a.blk = blk;
    a.nblk = nb;
    HIPCHK(launch_resident_batch(a, g->lds, g->stream));
  } else if (g->resident) {
    HIPCHK(launch_resident(a, g->NS, g->prm.precision, job_ens(g), g->lds, g->stream));
  } else {
    HIPCHK(launch_interval(a, g->nt, g->NS, g->prm.precision, g->lds, g->stream));
  }
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  if (parts > 1)
    HIPCHK(hipMemcpyAsync(g->h_abort, a.abort, sizeof(unsigned), hipMemcpyDeviceToHost, g->stream));
  rc = sync_scalars(g);
  if (rc) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  return 0;
}

extern "C" int remd_gpu_execute_pending(remd_gpu *g) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  int rc = launch(g, MODE_INTERVAL, 0, 0.0, g->d_reps, g->nrep, true);
  if (rc) return rc;
  rc = sync_scalars(g);
  if (rc) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  return 0;
}

extern "C" int remd_gpu_execute_steps_plain(remd_gpu *g, const double *T, const double *P) {
  int rc = remd_gpu_set_conditions(g, T, P);
  if (rc) return rc;
  // one replica per launch, one launch after another; a resident replica
  // takes as many workgroups as it can use
  const int parts = g->resident ? pick_parts(g->capacity, 1, g->N) : 1;
  g->last_parts = parts;
  if (parts > 1) HIPCHK(hipMemsetAsync(g->d_sync, 0, sizeof(unsigned long long) * (g->nrep + 1), g->stream));
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  for (int r = 0; r < g->nrep; r++) {
    LaunchArgs a;
    a.reps = g->d_reps + r;
    a.nrep = 1;
    a.mode = MODE_INTERVAL;
    a.do_sort = 0;
    a.margin = 0.0;
    a.parts = parts;
    a.same_xcd_ok = same_xcd_barrier_enabled();
    a.same_xcd_mode = same_xcd_barrier_mode();
    a.sync = g->d_sync ? g->d_sync + r : nullptr;
    a.abort = g->d_sync ? (unsigned *)(g->d_sync + g->nrep) : nullptr;
    if (g->resident) {
      rc = launch_res(g, a, g->stream);
      if (rc) return rc;
    } else {
      HIPCHK(launch_interval(a, g->nt, g->NS, g->prm.precision, g->lds, g->stream));
    }
  }
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  if (parts > 1)
    HIPCHK(hipMemcpyAsync(g->h_abort, g->d_sync + g->nrep, sizeof(unsigned), hipMemcpyDeviceToHost, g->stream));
  rc = sync_scalars(g);
  if (rc) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  return 0;
}

extern "C" int remd_gpu_compute_forces(remd_gpu *g, int do_sort, double margin) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  if (!do_sort && !g->debug) return fail(REMD_ESTATE, "reusing the window needs remd_gpu_set_debug(g, 1)");
  int rc = launch(g, MODE_FORCE, do_sort, margin, g->d_reps, g->nrep, true);
  if (rc) return rc;
  rc = sync_scalars(g);
  if (rc) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  g->last_ms = ms;
  return 0;
}

extern "C" int remd_gpu_set_debug(remd_gpu *g, int on) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  if (on && !g->d_site) {
    HIPCHK(hipMalloc(&g->d_site, sizeof(double) * g->nrep * g->N * g->NS * 3));
    HIPCHK(hipMalloc(&g->d_win, sizeof(int) * g->nrep * g->N * 3));
    HIPCHK(hipMemsetAsync(g->d_site, 0, sizeof(double) * g->nrep * g->N * g->NS * 3, g->stream));
    HIPCHK(hipMemsetAsync(g->d_win, 0, sizeof(int) * g->nrep * g->N * 3, g->stream));
  }
  if (g->resident) {
    // the site-force partials take extra LDS on the resident path
    const int lds = resident_lds(g->N, g->NS, g->prm.precision, on != 0, g->gm ? 1 : 0);
    if (lds > device_lds_limit()) return fail(REMD_EINVAL, "site-force debug buffers do not fit this replica's LDS");
    g->lds = lds;
  }
  g->debug = on != 0;
  int rc = upload_reps(g);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int remd_gpu_get_site_forces(remd_gpu *g, int rep, double *af) {
  if (!g || !af || rep < 0 || rep >= g->nrep) return fail(REMD_EINVAL, "bad argument");
  if (!g->d_site) return fail(REMD_ESTATE, "debug buffers are off");
  HIPCHK(hipMemcpy(af, g->d_site + (size_t)rep * g->N * g->NS * 3, sizeof(double) * g->N * g->NS * 3,
                   hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int remd_gpu_get_window(remd_gpu *g, int rep, int *js, int *je, int *full) {
  if (!g || !js || !je || rep < 0 || rep >= g->nrep) return fail(REMD_EINVAL, "bad argument");
  if (!g->d_win) return fail(REMD_ESTATE, "debug buffers are off");
  std::vector<int> w(3 * g->N);
  HIPCHK(hipMemcpy(w.data(), g->d_win + (size_t)rep * g->N * 3, sizeof(int) * 3 * g->N, hipMemcpyDeviceToHost));
  for (int i = 0; i < g->N; i++) {
    js[i] = w[3 * i];
    je[i] = w[3 * i + 1];
    if (full) full[i] = w[3 * i + 2];
  }
  return 0;
}

extern "C" int remd_gpu_last_kernel_ms(remd_gpu *g, double *ms) {
  if (!g || !ms) return fail(REMD_EINVAL, "null argument");
  *ms = g->last_ms;
  return 0;
}

extern "C" int remd_gpu_phase_timers(remd_gpu *g, int rep, double out[16]) {
  if (!g || !out || rep < 0 || rep >= g->nrep) return fail(REMD_EINVAL, "bad argument");
  for (int k = 0; k < 16; k++) out[k] = g->h_sc[(size_t)rep * NSC + S_DIAG0 + k];
  return 0;
}

extern "C" int remd_gpu_launch_shape(remd_gpu *g, int *bt, int *ch, int *lds) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  if (bt) *bt = g->nt;
  if (ch) *ch = g->prm.chunks;
  if (lds) *lds = g->lds;
  return 0;
}

extern "C" int remd_gpu_launch_path(remd_gpu *g, int *path, int *parts) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  if (path) *path = g->gm ? REMD_PATH_SPLIT : (g->resident ? REMD_PATH_RESIDENT : REMD_PATH_GLOBAL);
  if (parts) *parts = g->last_parts;
  return 0;
}

// Time `reps` force passes (after one warm-up) with the current shape.
static int time_force_passes(remd_gpu *g, int reps, double *ms_sum) {
  double tsum = 0;
  for (int k = 0; k <= reps; k++) {
    int rc = launch(g, MODE_FORCE, 1, g->prm.margin, g->d_reps, g->nrep, true);
    if (rc) return rc;
    rc = finish(g);
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    if (k > 0) tsum += ms;
  }
  *ms_sum = tsum;
  return 0;
}

static int autotune_resident(remd_gpu *g, int reps) {
  const int c0 = g->prm.chunks;
  const int cands[6] = {c0 / 2, c0 - 1, c0, c0 + 1, c0 + 2, 2 * c0};
  double best = 1e300;
  int bc = c0;
  for (int k = 0; k < 6; k++) {
    const int C = cands[k];
    if (C < 1 || C > 64 || (k > 0 && C == cands[k - 1])) continue;
    g->prm.chunks = C;
    HIPCHK(hipMemcpyAsync(g->d_job, &g->prm, sizeof(JobParams), hipMemcpyHostToDevice, g->stream));
    double t = 0;
    int rc = time_force_passes(g, reps, &t);
    if (rc) return rc;
    if (t < best) {
      best = t;
      bc = C;
    }
  }
  g->prm.chunks = bc;
  HIPCHK(hipMemcpyAsync(g->d_job, &g->prm, sizeof(JobParams), hipMemcpyHostToDevice, g->stream));
  int rc = launch(g, MODE_FORCE, 1, g->prm.margin, g->d_reps, g->nrep, false);
  if (rc) return rc;
  return sync_scalars(g);
}

extern "C" int remd_gpu_autotune(remd_gpu *g, int reps) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  if (reps < 1) reps = 3;
  if (g->resident) return autotune_resident(g, reps);
  const int nts[3] = {128, 256, 512};
  double best = 1e300;
  int bnt = g->nt, bc = g->prm.chunks;
  const int lim = device_lds_limit();
  for (int a = 0; a < 3; a++) {
    for (int C = 1; C <= 16; C *= 2) {
      const int lds = lds_bytes_host(g->N, g->NS, g->prm.precision, nts[a], g->grec);
      if (lds > lim) continue;
      g->nt = nts[a];
      g->lds = lds;
      g->prm.chunks = C;
      HIPCHK(hipMemcpyAsync(g->d_job, &g->prm, sizeof(JobParams), hipMemcpyHostToDevice, g->stream));
      double tsum = 0;
      for (int k = 0; k <= reps; k++) {
        int rc = launch(g, MODE_FORCE, 1, g->prm.margin, g->d_reps, g->nrep, true);
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(g->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        if (k > 0) tsum += ms;  // first launch is a warm-up
      }
      if (tsum < best) {
        best = tsum;
        bnt = nts[a];
        bc = C;
      }
    }
  }
  g->nt = bnt;
  g->prm.chunks = bc;
  g->lds = lds_bytes_host(g->N, g->NS, g->prm.precision, bnt, g->grec);
  HIPCHK(hipMemcpyAsync(g->d_job, &g->prm, sizeof(JobParams), hipMemcpyHostToDevice, g->stream));
  // restore the force/energies of the resident state for the chosen shape
  int rc = launch(g, MODE_FORCE, 1, g->prm.margin, g->d_reps, g->nrep, false);
  if (rc) return rc;
  return sync_scalars(g);
}

extern "C" int remd_gpu_get_averages(remd_gpu *g, int rep, double out[19]) {
  if (!g || !out || rep < 0 || rep >= g->nrep) return fail(REMD_EINVAL, "bad argument");
  const double *s = g->h_sc + (size_t)rep * NSC;
  out[0] = s[S_LJ];
  out[1] = s[S_CLMB];
  out[2] = s[S_WALL];
  out[3] = s[S_VOL];
  out[4] = s[S_TRA];
  out[5] = s[S_ROT];
  out[6] = s[S_TK];
  out[7] = s[S_PRS3];
  out[8] = s[S_PRS0];
  out[9] = s[S_PRS1];
  out[10] = s[S_PRS2];
  out[11] = s[S_TOT];
  out[12] = s[S_HAM];
  out[13] = s[S_S];
  out[14] = s[S_PS];
  out[15] = s[S_PV0];
  out[16] = s[S_PV1];
  out[17] = s[S_PV2];
  out[18] = 0.0;
  return 0;
}

extern "C" int remd_gpu_get_pot_vol(remd_gpu *g, double *pot, double *vol) {
  if (!g) return fail(REMD_EINVAL, "null handle");
  for (int r = 0; r < g->nrep; r++) {
    const double *s = g->h_sc + (size_t)r * NSC;
    if (pot) pot[r] = s[S_LJ] + s[S_CLMB] + s[S_WALL];
    if (vol) vol[r] = s[S_VOL];
  }
  return 0;
}

extern "C" int remd_gpu_download_state(remd_gpu *g, int rep, double *r, double *v, double *f, double *q,
                                       double *p, double *n, int *ids, remd_gpu_scalars *sc) {
  if (!g || rep < 0 || rep >= g->nrep) return fail(REMD_EINVAL, "bad argument");
  const size_t N = g->N;
  std::vector<double> m(NFIELD * N);
  HIPCHK(hipMemcpyAsync(m.data(), g->d_mol + (size_t)rep * NFIELD * N, sizeof(double) * NFIELD * N,
                        hipMemcpyDeviceToHost, g->stream));
  std::vector<int> mid(N);
  HIPCHK(hipMemcpyAsync(mid.data(), g->d_mid + (size_t)rep * N, sizeof(int) * N, hipMemcpyDeviceToHost,
                        g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  for (size_t i = 0; i < N; i++) {
    for (int d = 0; d < 3; d++) {
      if (r) r[3 * i + d] = m[(F_RX + d) * N + i];
      if (v) v[3 * i + d] = m[(F_VX + d) * N + i];
      if (f) f[3 * i + d] = m[(F_FX + d) * N + i];
    }
    for (int d = 0; d < 4; d++) {
      if (q) q[4 * i + d] = m[(F_Q0 + d) * N + i];
      if (p) p[4 * i + d] = m[(F_P0 + d) * N + i];
      if (n) n[4 * i + d] = m[(F_N0 + d) * N + i];
    }
    if (ids) ids[i] = mid[i];
  }
  if (sc) {
    const double *s = g->h_sc + (size_t)rep * NSC;
    sc->s = s[S_S];
    sc->Ps = s[S_PS];
    sc->Q = s[S_Q];
    for (int d = 0; d < 3; d++) sc->Pv[d] = s[S_PV0 + d];
    sc->W = s[S_W];
    for (int d = 0; d < 3; d++) sc->L[d] = s[S_L0 + d];
    sc->T = s[S_T];
    sc->P = s[S_P];
  }
  return 0;
}

// ------------------------------------------------------------ checkpoint
// Exact device state of a job at an interval boundary: every replica's
// molecule block (r, v, f, q, p, n in the internal representation of
// REMDGPU::ConvertVariablesToHost, src/remdgpu.cu:806-854), molecule ids
// and scalar block (thermostat, barostat, box, conditions, H0, energies,
// virial).  Loading it into a job of the same shape continues the run bit
// for bit: the next interval starts with a sort and uses the stored forces,
// so nothing is recomputed (the reference's text checkpoint,
// src/io_manager.cxx:528-634, re-derives forces and H0 at start-up).
struct CkptHeader {
  char magic[8];          // "REMDCKPT"
  int version;            // 1
  int nmol, nsite, nrep;
  int nfield, nsc;
  int precision, chunks;  // the launch shape that fixes the summation order
  int pad;
};

extern "C" size_t remd_gpu_checkpoint_bytes(remd_gpu *g) {
  if (!g) return 0;
  const size_t N = g->N, R = g->nrep;
  return sizeof(CkptHeader) + R * (NFIELD * N * sizeof(double) + NSC * sizeof(double) + N * sizeof(int));
}

extern "C" int remd_gpu_checkpoint_save(remd_gpu *g, void *buf, size_t bytes) {
  if (!g || !buf) return fail(REMD_EINVAL, "null argument");
  if (bytes < remd_gpu_checkpoint_bytes(g)) return fail(REMD_EINVAL, "checkpoint buffer too small");
  const size_t N = g->N, R = g->nrep;
  CkptHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, "REMDCKPT", 8);
  h.version = 1;
  h.nmol = g->N;
  h.nsite = g->NS;
  h.nrep = g->nrep;
  h.nfield = NFIELD;
  h.nsc = NSC;
  h.precision = g->prm.precision;
  h.chunks = g->prm.chunks;
  char *o = (char *)buf;
  std::memcpy(o, &h, sizeof(h));
  o += sizeof(h);
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(o, g->d_mol, sizeof(double) * NFIELD * N * R, hipMemcpyDeviceToHost));
  o += sizeof(double) * NFIELD * N * R;
  HIPCHK(hipMemcpy(o, g->d_sc, sizeof(double) * NSC * R, hipMemcpyDeviceToHost));
  o += sizeof(double) * NSC * R;
  HIPCHK(hipMemcpy(o, g->d_mid, sizeof(int) * N * R, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int remd_gpu_checkpoint_load(remd_gpu *g, const void *buf, size_t bytes) {
  if (!g || !buf) return fail(REMD_EINVAL, "null argument");
  if (bytes < sizeof(CkptHeader)) return fail(REMD_EINVAL, "checkpoint too short");
  CkptHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, "REMDCKPT", 8) != 0 || h.version != 1)
    return fail(REMD_EINVAL, "not a remd_gpu checkpoint (or another version)");
  if (h.nmol != g->N || h.nsite != g->NS || h.nrep != g->nrep || h.nfield != NFIELD || h.nsc != NSC)
    return fail(REMD_EINVAL, "checkpoint shape (nmol, nsite, nreplica) differs from the job's");
  if (h.precision != g->prm.precision || h.chunks != g->prm.chunks)
    return fail(REMD_EINVAL, "checkpoint of a job with another pair precision or chunk count");
  if (bytes < remd_gpu_checkpoint_bytes(g)) return fail(REMD_EINVAL, "checkpoint truncated");
  const size_t N = g->N, R = g->nrep;
  const char *o = (const char *)buf + sizeof(h);
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(g->d_mol, o, sizeof(double) * NFIELD * N * R, hipMemcpyHostToDevice));
  o += sizeof(double) * NFIELD * N * R;
  std::memcpy(g->h_sc, o, sizeof(double) * NSC * R);
  HIPCHK(hipMemcpy(g->d_sc, g->h_sc, sizeof(double) * NSC * R, hipMemcpyHostToDevice));
  o += sizeof(double) * NSC * R;
  HIPCHK(hipMemcpy(g->d_mid, o, sizeof(int) * N * R, hipMemcpyHostToDevice));
  return 0;
}

// body_to_space (src/rotation.h:29-48)
static void b2s(const double *a, const double *v, double *t) {
  const double a0sq = a[0] * a[0], a1sq = a[1] * a[1], a2sq = a[2] * a[2], a3sq = a[3] * a[3];
  const double a01 = a[0] * a[1], a02 = a[0] * a[2], a03 = a[0] * a[3];
  const double a12 = a[1] * a[2], a13 = a[1] * a[3], a23 = a[2] * a[3];
  t[0] = (a0sq + a1sq - a2sq - a3sq) * v[0] + 2. * (a12 - a03) * v[1] + 2. * (a13 + a02) * v[2];
  t[1] = 2. * (a12 + a03) * v[0] + (a0sq - a1sq + a2sq - a3sq) * v[1] + 2. * (a23 - a01) * v[2];
  t[2] = 2. * (a13 - a02) * v[0] + 2. * (a23 + a01) * v[1] + (a0sq - a1sq - a2sq + a3sq) * v[2];
}

extern "C" int remd_gpu_write_cdv(remd_gpu *g, const char *prefix, unsigned step) {
  if (!g || !prefix) return fail(REMD_EINVAL, "null argument");
  const size_t N = g->N;
  std::vector<double> r(3 * N), q(4 * N);
  for (int rep = 0; rep < g->nrep; rep++) {
    remd_gpu_scalars sc;
    int rc = remd_gpu_download_state(g, rep, r.data(), nullptr, nullptr, q.data(), nullptr, nullptr, nullptr, &sc);
    if (rc) return rc;
    char name[4096];
    snprintf(name, sizeof(name), "%sR%04ds%07u.cdv", prefix, rep, step);
    FILE *fp = fopen(name, "w");
    if (!fp) return fail(REMD_EINVAL, std::string("cannot open ") + name);
    const double *L = sc.L;
    const double R = g->desc.wall_length;
    if (g->desc.confined == 1)
      fprintf(fp, "' box_sx=%g box_sy=%g box_sz=%g box_ex=%g box_ey=%g box_ez=%g", -R, -R, -0.5 * L[2], R, R,
              0.5 * L[2]);
    else if (g->desc.confined == 2)
      fprintf(fp, "' box_sx=%g box_sy=%g box_sz=%g box_ex=%g box_ey=%g box_ez=%g", -0.5 * L[0], -0.5 * L[1],
              -0.5 * R, 0.5 * L[0], 0.5 * L[1], 0.5 * R);
    else
      fprintf(fp, "' box_sx=%g box_sy=%g box_sz=%g box_ex=%g box_ey=%g box_ez=%g", -0.5 * L[0], -0.5 * L[1],
              -0.5 * L[2], 0.5 * L[0], 0.5 * L[1], 0.5 * L[2]);
    fprintf(fp, " r0=1.0 r1=0.8 r2=0.8 c0=(1.0,0.0,0.0) c1=(1.0,1.0,1.0) c2=(1.0,1.0,1.0)"
                " bond0_c=(1.0,1.0,1.0) st0=\"T=%g K, P=%g MPa\" st0_pos=(-2.0,-2.0)\n",
            sc.T * unit_temp(), sc.P * unit_press() / 1e6);
    for (size_t i = 0; i < N; i++) {
      for (int a = 0; a < g->NS; a++) {
        double t[3];
        b2s(&q[4 * i], g->prm.site_r[a], t);
        fprintf(fp, "%zu %d %g %g %g\n", i, a, (r[3 * i] - 0.5) * L[0] + t[0], (r[3 * i + 1] - 0.5) * L[1] + t[1],
                (r[3 * i + 2] - 0.5) * L[2] + t[2]);
      }
    }
    fclose(fp);
  }
  return 0;
}

// ------------------------------------------------------------- wall table
extern "C" int remd_gpu_wall_table(int confined, int nfwall, double R, double sgm, double eps, double rho,
                                   double *out) {
  if (!out || nfwall < 3 || !(R > 0) || (confined != 1 && confined != 2))
    return fail(REMD_EINVAL, "bad wall table arguments");
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * 2 * nfwall));
  hipError_t e;
  if (confined == 1) {
    const double sgm06 = (double)powf((float)sgm, 6.0f);
    e = launch_wall_tube(nfwall, R, sgm06, M_PI * rho * eps * sgm06, d, nullptr);
  } else {
    e = launch_wall_slit(nfwall, R, sgm, eps, rho, d, nullptr);
  }
  if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(double) * 2 * nfwall, hipMemcpyDeviceToHost);
  hipFree(d);
  if (e != hipSuccess) return fail(REMD_EHIP, std::string("wall table: ") + hipGetErrorString(e));
  return 0;
}

// --------------------------------------------------------- sort / window
extern "C" int remd_gpu_sort_window(const double *keys, int nrow, int n, double w, int *order, int *js,
                                    int *je) {
  if (!keys || !order || !js || !je || nrow < 1 || n < 2 || n > 4096)
    return fail(REMD_EINVAL, "bad sort arguments");
  const int lds = 20 * n + 8 + (16 * 64 + 16) * 4;
  if (lds > device_lds_limit()) return fail(REMD_EINVAL, "row too long for LDS");
  double *dk = nullptr;
  int *dout = nullptr;
  HIPCHK(hipMalloc(&dk, sizeof(double) * (size_t)nrow * n));
  hipError_t e = hipMalloc(&dout, sizeof(int) * 3 * (size_t)nrow * n);
  if (e == hipSuccess) e = hipMemcpy(dk, keys, sizeof(double) * (size_t)nrow * n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_sort_window(dk, nrow, n, w, dout, dout + (size_t)nrow * n, dout + 2 * (size_t)nrow * n, nullptr);
  if (e == hipSuccess) e = hipMemcpy(order, dout, sizeof(int) * (size_t)nrow * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(js, dout + (size_t)nrow * n, sizeof(int) * (size_t)nrow * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(je, dout + 2 * (size_t)nrow * n, sizeof(int) * (size_t)nrow * n, hipMemcpyDeviceToHost);
  hipFree(dk);
  hipFree(dout);
  if (e != hipSuccess) return fail(REMD_EHIP, std::string("sort_window: ") + hipGetErrorString(e));
  return 0;
}

// -------------------------------------------------------------- exchange
static int pair_count(int dt, int dp, int type) {
  if (type == 0) return dp * (dt / 2);
  if (type == 1) return (dp / 2) * dt;
  if (type == 2) return dp * std::max(0, dt / 2 - (dt + 1) % 2);
  return std::max(0, dp / 2 - (dp + 1) % 2) * dt;
}

extern "C" int remd_gpu_exchange_device(remd_gpu *g, int ntemp, int npress, long step, int *index, double *T,
                                        double *P, const double *u, int *accepted, int *npairs) {
  if (!g || !index || !T || !P || !u) return fail(REMD_EINVAL, "null argument");
  if (ntemp * npress != g->nrep) return fail(REMD_EINVAL, "ntemp*npress must equal nreplica");
  const int R = g->nrep;
  {
    // index must be a permutation of the replicas: the kernel dereferences it
    std::vector<char> seen(R, 0);
    for (int k = 0; k < R; k++) {
      if (index[k] < 0 || index[k] >= R || seen[index[k]]) return fail(REMD_EINVAL, "index is not a permutation");
      seen[index[k]] = 1;
      if (!(T[k] > 0)) return fail(REMD_EINVAL, "T must be > 0");
    }
  }
  const int ec = (int)(step % 4);
  const int np = pair_count(ntemp, npress, ec);
  if (!g->d_index) {
    HIPCHK(hipMalloc(&g->d_index, sizeof(int) * R));
    HIPCHK(hipMalloc(&g->d_T, sizeof(double) * R));
    HIPCHK(hipMalloc(&g->d_P, sizeof(double) * R));
    HIPCHK(hipMalloc(&g->d_u, sizeof(double) * (R + 1)));
    HIPCHK(hipMalloc(&g->d_acc, sizeof(int) * (R + 1)));
    HIPCHK(hipMalloc(&g->d_scptr, sizeof(double *) * R));
    HIPCHK(hipMalloc(&g->d_xjob, sizeof(ExchangeJob)));
    std::vector<double *> ptrs(R);
    for (int r = 0; r < R; r++) ptrs[r] = g->d_sc + (size_t)r * NSC;
    HIPCHK(hipMemcpy(g->d_scptr, ptrs.data(), sizeof(double *) * R, hipMemcpyHostToDevice));
  }
  ExchangeJob X;
  X.ntemp = ntemp;
  X.npress = npress;
  X.pmode = g->desc.pconstant;
  X.ec_type = ec;
  X.index = g->d_index;
  X.T = g->d_T;
  X.P = g->d_P;
  X.u = g->d_u;
  X.sc = g->d_scptr;
  X.accepted = g->d_acc;
  X.hist = nullptr;
  X.sync = nullptr;
  HIPCHK(hipMemcpyAsync(g->d_xjob, &X, sizeof(X), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_index, index, sizeof(int) * R, hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_T, T, sizeof(double) * R, hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_P, P, sizeof(double) * R, hipMemcpyHostToDevice, g->stream));
  if (np > 0) HIPCHK(hipMemcpyAsync(g->d_u, u, sizeof(double) * np, hipMemcpyHostToDevice, g->stream));
  HIPCHK(launch_exchange(g->d_xjob, 1, g->stream));
  HIPCHK(hipMemcpyAsync(index, g->d_index, sizeof(int) * R, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(T, g->d_T, sizeof(double) * R, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(P, g->d_P, sizeof(double) * R, hipMemcpyDeviceToHost, g->stream));
  std::vector<int> acc(np > 0 ? np : 1);
  if (np > 0) HIPCHK(hipMemcpyAsync(acc.data(), g->d_acc, sizeof(int) * np, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (accepted)
    for (int k = 0; k < np; k++) accepted[k] = acc[k];
  if (npairs) *npairs = np;
  // keep the host mirror's conditions in step with the device
  return sync_scalars(g);
}

// GetAverages of one replica's scalar block (Average enum order,
// src/remdinfo.h:27-41).
static void averages_of(const double *s, double out[19]) {
  out[0] = s[S_LJ];
  out[1] = s[S_CLMB];
  out[2] = s[S_WALL];
  out[3] = s[S_VOL];
  out[4] = s[S_TRA];
  out[5] = s[S_ROT];
  out[6] = s[S_TK];
  out[7] = s[S_PRS3];
  out[8] = s[S_PRS0];
  out[9] = s[S_PRS1];
  out[10] = s[S_PRS2];
  out[11] = s[S_TOT];
  out[12] = s[S_HAM];
  out[13] = s[S_S];
  out[14] = s[S_PS];
  out[15] = s[S_PV0];
  out[16] = s[S_PV1];
  out[17] = s[S_PV2];
  out[18] = 0.0;
}

// Grows a per-job device buffer of remd_gpu_execute_intervals to n elements.
template <class T>
static int grow(T **p, size_t *cap, size_t n) {
  if (n <= *cap) return 0;
  hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIPCHK(hipMalloc(p, sizeof(T) * n));
  *cap = n;
  return 0;
}

extern "C" int remd_gpu_execute_intervals(remd_gpu *g, int ntemp, int npress, long step0, int nint, int *index,
                                          double *T, double *P, remd_rng *rng, int *accepted, double *averages,
                                          double *kernel_ms) {
  if (!g || !index || !T || !P || !rng) return fail(REMD_EINVAL, "null argument");
  if (nint < 1 || step0 < 0) return fail(REMD_EINVAL, "nint must be >= 1 and step0 >= 0");
  if (ntemp < 1 || npress < 1 || ntemp * npress != g->nrep) return fail(REMD_EINVAL, "ntemp*npress must equal nreplica");
  const int R = g->nrep;
  {
    // index must be a permutation of the replicas: the exchange kernel dereferences it
    std::vector<char> seen(R, 0);
    for (int k = 0; k < R; k++) {
      if (index[k] < 0 || index[k] >= R || seen[index[k]]) return fail(REMD_EINVAL, "index is not a permutation");
      seen[index[k]] = 1;
    }
  }
  // the uniforms remd_exchange_host would draw, interval after interval
  std::vector<int> off(nint + 1, 0);
  for (int k = 0; k < nint; k++) off[k + 1] = off[k] + pair_count(ntemp, npress, (int)((step0 + k) % 4));
  const int U = off[nint];
  std::vector<double> u(U > 0 ? U : 1);
  for (int k = 0; k < U; k++) u[k] = (double)remd_rng_next(rng) / (double)2147483647;  // rand()/RAND_MAX
  int rc = remd_gpu_set_conditions(g, T, P);
  if (rc) return rc;
  if (!g->d_index) {
    HIPCHK(hipMalloc(&g->d_index, sizeof(int) * R));
    HIPCHK(hipMalloc(&g->d_T, sizeof(double) * R));
    HIPCHK(hipMalloc(&g->d_P, sizeof(double) * R));
    HIPCHK(hipMalloc(&g->d_u, sizeof(double) * (R + 1)));
    HIPCHK(hipMalloc(&g->d_acc, sizeof(int) * (R + 1)));
    HIPCHK(hipMalloc(&g->d_scptr, sizeof(double *) * R));
    HIPCHK(hipMalloc(&g->d_xjob, sizeof(ExchangeJob)));
    std::vector<double *> ptrs(R);
    for (int r = 0; r < R; r++) ptrs[r] = g->d_sc + (size_t)r * NSC;
    HIPCHK(hipMemcpy(g->d_scptr, ptrs.data(), sizeof(double *) * R, hipMemcpyHostToDevice));
  }
  if ((rc = grow(&g->xi_u, &g->xi_ucap, (size_t)(U > 0 ? U : 1))) ||
      (rc = grow(&g->xi_acc, &g->xi_acap, (size_t)(U > 0 ? U : 1))) ||
      (rc = grow(&g->xi_jobs, &g->xi_jcap, (size_t)nint)) ||
      (averages && (rc = grow(&g->xi_hist, &g->xi_hcap, (size_t)NSC * R * nint))))
    return rc;
  while (g->xi_ev.size() < 2 * (size_t)nint) {
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    g->xi_ev.push_back(e);
  }
  std::vector<ExchangeJob> X(nint);
  for (int k = 0; k < nint; k++) {
    X[k].ntemp = ntemp;
    X[k].npress = npress;
    X[k].pmode = g->desc.pconstant;
    X[k].ec_type = (int)((step0 + k) % 4);
    X[k].index = g->d_index;
    X[k].T = g->d_T;
    X[k].P = g->d_P;
    X[k].u = g->xi_u + off[k];
    X[k].sc = g->d_scptr;
    X[k].accepted = g->xi_acc + off[k];
    X[k].hist = averages ? g->xi_hist + (size_t)k * NSC * R : nullptr;
    // the exchange kernel clears the group words for the next launch
    X[k].sync = (g->resident && k + 1 < nint) ? g->d_sync : nullptr;
  }
  HIPCHK(hipMemcpyAsync(g->xi_jobs, X.data(), sizeof(ExchangeJob) * nint, hipMemcpyHostToDevice, g->stream));
  if (U > 0) HIPCHK(hipMemcpyAsync(g->xi_u, u.data(), sizeof(double) * U, hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_index, index, sizeof(int) * R, hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_T, T, sizeof(double) * R, hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_P, P, sizeof(double) * R, hipMemcpyHostToDevice, g->stream));
  LaunchArgs a;
  a.reps = g->d_reps;
  a.nrep = R;
  a.mode = MODE_INTERVAL;
  a.do_sort = 0;
  a.margin = 0.0;
  a.parts = g->resident ? pick_parts(g->capacity, R, g->N) : 1;
  a.same_xcd_ok = same_xcd_barrier_enabled();
  a.same_xcd_mode = same_xcd_barrier_mode();
  a.sync = g->d_sync;
  a.abort = g->d_sync ? (unsigned *)(g->d_sync + R) : nullptr;
  g->last_parts = a.parts;
  // the abort word is cleared once: a launch that gives up leaves it set,
  // every later launch of the call then gives up at once, and the call fails
  if (a.parts > 1) HIPCHK(hipMemsetAsync(g->d_sync, 0, sizeof(unsigned long long) * (R + 1), g->stream));
  for (int k = 0; k < nint; k++) {
    HIPCHK(hipEventRecord(g->xi_ev[2 * k], g->stream));
    if (g->resident) {
      rc = launch_res(g, a, g->stream);
      if (rc) return rc;
    } else {
      HIPCHK(launch_interval(a, g->nt, g->NS, g->prm.precision, g->lds, g->stream));
    }
    HIPCHK(hipEventRecord(g->xi_ev[2 * k + 1], g->stream));
    HIPCHK(launch_exchange(g->xi_jobs + k, 1, g->stream));
  }
  if (a.parts > 1)
    HIPCHK(hipMemcpyAsync(g->h_abort, a.abort, sizeof(unsigned), hipMemcpyDeviceToHost, g->stream));
  std::vector<int> acc(U > 0 ? U : 1);
  std::vector<double> hist(averages ? (size_t)NSC * R * nint : 0);
  HIPCHK(hipMemcpyAsync(index, g->d_index, sizeof(int) * R, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(T, g->d_T, sizeof(double) * R, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(P, g->d_P, sizeof(double) * R, hipMemcpyDeviceToHost, g->stream));
  if (U > 0) HIPCHK(hipMemcpyAsync(acc.data(), g->xi_acc, sizeof(int) * U, hipMemcpyDeviceToHost, g->stream));
  if (averages)
    HIPCHK(hipMemcpyAsync(hist.data(), g->xi_hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost, g->stream));
  rc = sync_scalars(g);
  if (rc) return rc;
  int nacc = 0;
  for (int k = 0; k < U; k++) {
    nacc += acc[k];
    if (accepted) accepted[k] = acc[k];
  }
  if (averages)
    for (int k = 0; k < nint; k++)
      for (int r = 0; r < R; r++) averages_of(&hist[((size_t)k * R + r) * NSC], averages + ((size_t)k * R + r) * 19);
  for (int k = 0; k < nint; k++) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, g->xi_ev[2 * k], g->xi_ev[2 * k + 1]));
    if (kernel_ms) kernel_ms[k] = ms;
    g->last_ms = ms;
  }
  return nacc;
}

// ----------------------------------------------------------------- batch
extern "C" int remd_batch_create(remd_batch **out) {
  if (!out) return fail(REMD_EINVAL, "null output");
  remd_batch *b = new remd_batch;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->efork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->edone, hipEventDisableTiming);
  for (int c = 0; c < NCLASS && e == hipSuccess; c++) {
    e = hipStreamCreateWithFlags(&b->cstream[c], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->cdone[c], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipHostMalloc(&b->h_abort, sizeof(unsigned), hipHostMallocDefault);
  if (e != hipSuccess) {
    remd_batch_destroy(b);
    return fail(REMD_EHIP, hipGetErrorString(e));
  }
  *b->h_abort = 0u;
  *out = b;
  return 0;
}

extern "C" int remd_batch_add(remd_batch *b, remd_gpu *g) {
  if (!b || !g) return fail(REMD_EINVAL, "null argument");
  if (b->in_flight) return fail(REMD_ESTATE, "the batch has an interval in flight (remd_batch_wait first)");
  if (std::find(b->jobs.begin(), b->jobs.end(), g) != b->jobs.end()) return 0;
  b->jobs.push_back(g);
  b->dirty = true;
  return 0;
}

extern "C" int remd_batch_remove(remd_batch *b, remd_gpu *g) {
  if (!b || !g) return fail(REMD_EINVAL, "null argument");
  if (b->in_flight) return fail(REMD_ESTATE, "the batch has an interval in flight (remd_batch_wait first)");
  auto it = std::find(b->jobs.begin(), b->jobs.end(), g);
  if (it == b->jobs.end()) return fail(REMD_EINVAL, "job not in batch");
  b->jobs.erase(it);
  b->dirty = true;
  return 0;
}

extern "C" int remd_batch_set_autotune(remd_batch *b, int on) {
  if (!b) return fail(REMD_EINVAL, "null batch");
  b->autotune = on != 0;
  b->dirty = true;
  return 0;
}

extern "C" int remd_batch_set_parts(remd_batch *b, const int *parts, int n) {
  if (!b) return fail(REMD_EINVAL, "null batch");
  if (b->in_flight) return fail(REMD_ESTATE, "the batch has an interval in flight (remd_batch_wait first)");
  if (!parts) {
    b->forced.clear();
  } else {
    if (n != (int)b->jobs.size()) return fail(REMD_EINVAL, "one part count per job of the batch");
    for (int k = 0; k < n; k++) {
      const remd_gpu *g = b->jobs[k];
      const int pmax = g->resident ? std::min(GMAX, std::max(1, g->N / 8)) : 1;
      if (parts[k] < 1 || parts[k] > pmax)
        return fail(REMD_EINVAL, "job " + std::to_string(k) + ": parts must be 1.." + std::to_string(pmax));
    }
    b->forced.assign(parts, parts + n);
  }
  b->dirty = true;
  return 0;
}

// Relative time of one MD step of one replica of g split over p
// workgroups: the integration every part repeats (O(N)), the rank sort
// (O(N^2) compares over 256 lanes), the pair work over the sort-axis
// window shared out among the parts (twice as dear in FP64), and for p > 1
// the exchange rows and the group barrier.  Coefficients are per-phase
// timings of the resident kernel on 144 TIP4P in a nanotube, in
// microseconds; the autotuner corrects what this model gets wrong.
static double step_cost(const remd_gpu *g, int p) {
  const double N = g->N, ns = g->NS;
  const double pair = 2.2e-4 * N * N * g->wfrac * ns * ns * (g->prm.precision ? 2.0 : 1.0);
  return 1.3e-2 * N * ns + 6.6e-5 * N * N + pair / p + (p > 1 ? 2.6e-2 * N : 0.0);
}

static double interval_cost(const remd_gpu *g, int p) { return step_cost(g, p) * std::max(1, g->desc.interval); }

static unsigned variant_of(const remd_gpu *g) {
  return (unsigned)((g->gm ? 6 : 0) + (g->NS - 3) * 2 + g->prm.precision);
}

// Estimated time of one batched interval under plan P: the split replicas
// hold their workgroups for their (longest) interval; the single-part
// replicas fill the remaining slots meanwhile and every slot afterwards.
static double plan_time(const remd_batch *b, const std::vector<int> &P) {
  double tg = 0.0, w1 = 0.0, m1 = 0.0;
  double used = 0.0;
  for (size_t k = 0; k < b->jobs.size(); k++) {
    const remd_gpu *g = b->jobs[k];
    if (!g->resident) continue;
    const double t = interval_cost(g, P[k]);
    if (P[k] > 1) {
      tg = std::max(tg, t);
      used += (double)g->nrep * P[k];
    } else {
      w1 += g->nrep * t;
      m1 = std::max(m1, t);
    }
  }
  const double cap = std::max(1, b->res_capacity), fr = cap - used;
  const double T = fr * tg >= w1 ? tg : tg + (w1 - fr * tg) / cap;
  return std::max(T, m1);
}

// Block table of plan P: first the parts of every split replica, packed
// into the 8 block columns (block index mod 8, one XCD each) with all parts
// of a replica in one column, more parts first; then the single-part
// replicas, heaviest first.  False when the split parts need more than
// capacity / 8 rows: they must all be resident at once.
static bool layout_plan(const remd_batch *b, const std::vector<int> &P, std::vector<unsigned> *out) {
  std::vector<unsigned> col[8];
  int h[8] = {0};
  std::vector<int> split, single;
  for (int r = 0; r < b->res_n; r++) (P[b->row_job[r]] > 1 ? split : single).push_back(r);
  std::stable_sort(split.begin(), split.end(),
                   [&](int x, int y) { return P[b->row_job[x]] > P[b->row_job[y]]; });
  for (int r : split) {
    const int p = P[b->row_job[r]];
    int x = 0;
    for (int k = 1; k < 8; k++)
      if (h[k] < h[x]) x = k;
    const unsigned v = variant_of(b->jobs[b->row_job[r]]);
    for (int q = 0; q < p; q++) col[x].push_back((unsigned)r << 12 | v << 8 | (unsigned)p << 4 | (unsigned)q);
    h[x] += p;
  }
  int H = 0;
  for (int x = 0; x < 8; x++) H = std::max(H, h[x]);
  if (H > b->res_capacity / 8) return false;
  if (out) {
    out->clear();
    for (int s = 0; s < H; s++)
      for (int x = 0; x < 8; x++) out->push_back(s < h[x] ? col[x][s] : ~0u);
    for (int r : single) out->push_back((unsigned)r << 12 | variant_of(b->jobs[b->row_job[r]]) << 8 | 1u << 4);
  }
  return true;
}

// Candidate plans: from one workgroup per replica, repeatedly give the job
// whose replicas take longest one more part, as long as the split parts
// still fit; keep the (up to 4) plans estimated within 25% of the best.
static void make_candidates(remd_batch *b) {
  const size_t nj = b->jobs.size();
  std::vector<int> P(nj, 1), capped(nj, 0);
  std::vector<std::pair<double, std::vector<int>>> seen{{plan_time(b, P), P}};
  for (;;) {
    int k = -1;
    double worst = -1.0;
    for (size_t j = 0; j < nj; j++) {
      const remd_gpu *g = b->jobs[j];
      if (!g->resident || capped[j] || P[j] >= std::min(GMAX, std::max(1, g->N / 8))) continue;
      const double t = interval_cost(g, P[j]);
      if (t > worst) {
        worst = t;
        k = (int)j;
      }
    }
    if (k < 0) break;
    std::vector<int> Q = P;
    Q[k]++;
    if (!layout_plan(b, Q, nullptr)) {
      capped[k] = 1;
      continue;
    }
    P = Q;
    seen.push_back({plan_time(b, P), P});
  }
  std::stable_sort(seen.begin(), seen.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
  b->cands.clear();
  for (const auto &c : seen) {
    if (b->cands.size() >= 4 || c.first > 1.25 * seen[0].first) break;
    b->cands.push_back(c.second);
  }
}

static int apply_plan(remd_batch *b, const std::vector<int> &P) {
  std::vector<unsigned> blk;
  if (!layout_plan(b, P, &blk)) return fail(REMD_EINVAL, "batch plan does not fit the device");
  const int n = (int)blk.size();
  if (n > b->blk_cap) {
    hipFree(b->d_blk);
    b->d_blk = nullptr;
    HIPCHK(hipMalloc(&b->d_blk, sizeof(unsigned) * std::max(n, 1)));
    b->blk_cap = n;
  }
  if (n) HIPCHK(hipMemcpy(b->d_blk, blk.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice));
  b->nblk = n;
  b->plan = P;
  b->split = false;
  for (size_t k = 0; k < b->jobs.size(); k++)
    if (b->jobs[k]->resident && P[k] > 1) b->split = true;
  return 0;
}

// Table of every replica of the batch: the resident rows heaviest first,
// then the global-state rows by class, heaviest first within a class
// (workgroups are dispatched in block order, so long replicas start early).
static int rebuild_table(remd_batch *b) {
  struct Row {
    int cls;  // -1: resident
    double w;
    int job;
    ReplicaDesc d;
  };
  std::vector<Row> rows;
  for (int c = 0; c < NCLASS; c++) b->cls_lds[c] = b->cls_n[c] = b->cls_off[c] = 0;
  b->res_n = b->res_lds = 0;
  for (size_t k = 0; k < b->jobs.size(); k++) {
    remd_gpu *g = b->jobs[k];
    refresh_reps(g);
    const int c = g->resident ? -1 : kclass(g->NS, g->prm.precision, g->nt);
    const double w = interval_cost(g, 1);
    for (int r = 0; r < g->nrep; r++) rows.push_back({c, -w, (int)k, g->h_reps[r]});
    if (g->resident) {
      b->res_lds = std::max(b->res_lds, g->lds);
      b->res_n += g->nrep;
    } else {
      b->cls_lds[c] = std::max(b->cls_lds[c], g->lds);
      b->cls_n[c] += g->nrep;
    }
  }
  if (b->res_n >= (1 << 20)) return fail(REMD_EINVAL, "too many resident replicas in one batch");
  if (b->res_lds > device_lds_limit()) return fail(REMD_EINVAL, "batch LDS exceeds the device limit");
  b->res_capacity = 0;
  if (b->res_n) {
    int per_cu = 0, ncu = 0, dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(resident_batch_occupancy(b->res_lds, &per_cu));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu < 1) return fail(REMD_EINVAL, "the resident batch kernel does not fit a CU");
    b->res_capacity = per_cu * ncu;
  }
  std::stable_sort(rows.begin(), rows.end(),
                   [](const Row &x, const Row &y) { return x.cls != y.cls ? x.cls < y.cls : x.w < y.w; });
  int off = b->res_n;
  for (int c = 0; c < NCLASS; c++) {
    b->cls_off[c] = off;
    off += b->cls_n[c];
  }
  const int n = (int)rows.size();
  b->row_job.resize(b->res_n);
  for (int r = 0; r < b->res_n; r++) b->row_job[r] = rows[r].job;
  if (n > b->table_cap) {
    hipFree(b->d_table);
    b->d_table = nullptr;
    HIPCHK(hipMalloc(&b->d_table, sizeof(ReplicaDesc) * std::max(n, 1)));
    b->table_cap = n;
  }
  hipFree(b->d_sync);
  b->d_sync = nullptr;
  HIPCHK(hipMalloc(&b->d_sync, sizeof(unsigned long long) * (b->res_n + 1)));
  *b->h_abort = 0u;
  std::vector<ReplicaDesc> flat(n);
  for (int k = 0; k < n; k++) flat[k] = rows[k].d;
  if (n) HIPCHK(hipMemcpyAsync(b->d_table, flat.data(), sizeof(ReplicaDesc) * n, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->ntab = n;
  bool fixed = false;
  if (!b->forced.empty() && b->forced.size() == b->jobs.size()) {
    if (!layout_plan(b, b->forced, nullptr))
      return fail(REMD_EINVAL, "the fixed parts need more workgroups than the device holds at once");
    b->cands.assign(1, b->forced);
    fixed = true;
  } else {
    b->forced.clear();
    make_candidates(b);
  }
  b->cand_ms.assign(b->cands.size(), 0.0);
  b->cand_cur = 0;
  b->warm = 1;
  b->tuned = !b->autotune || fixed;
  int rc = apply_plan(b, b->cands[0]);
  if (rc) return rc;
  b->dirty = false;
  return 0;
}

// Hill climbing after the model's candidates: the plans one part away (for
// one job, up or down) from the fastest plan timed so far, not yet timed
// and fitting the device, best estimate first; at most 16 plans are timed
// per batch composition.
static void add_neighbours(remd_batch *b) {
  constexpr size_t MAXTIMED = 16;
  if (b->cands.size() >= MAXTIMED) return;
  const size_t best = (size_t)(std::min_element(b->cand_ms.begin(), b->cand_ms.end()) - b->cand_ms.begin());
  const std::vector<int> base = b->cands[best];
  std::vector<std::pair<double, std::vector<int>>> nb;
  for (size_t k = 0; k < b->jobs.size(); k++) {
    const remd_gpu *g = b->jobs[k];
    if (!g->resident) continue;
    const int pmax = std::min(GMAX, std::max(1, g->N / 8));
    for (int dp = -1; dp <= 1; dp += 2) {
      std::vector<int> P = base;
      P[k] += dp;
      if (P[k] < 1 || P[k] > pmax) continue;
      if (std::find(b->cands.begin(), b->cands.end(), P) != b->cands.end()) continue;
      if (!layout_plan(b, P, nullptr)) continue;
      nb.push_back({plan_time(b, P), P});
    }
  }
  std::stable_sort(nb.begin(), nb.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
  for (const auto &c : nb) {
    if (b->cands.size() >= MAXTIMED) break;
    b->cands.push_back(c.second);
    b->cand_ms.push_back(0.0);
  }
}

extern "C" int remd_batch_launch(remd_batch *b) {
  if (!b) return fail(REMD_EINVAL, "null batch");
  if (b->in_flight) return fail(REMD_ESTATE, "the batch's previous interval is still in flight");
  if (b->jobs.empty()) return 0;
  for (remd_gpu *g : b->jobs)
    if (!g->resident && g->nt != 256 && g->nt != 512)
      return fail(REMD_EINVAL, "batched global-state jobs use 256- or 512-thread workgroups");
  if (b->dirty) {
    int rc = rebuild_table(b);
    if (rc) return rc;
  }
  // conditions were staged on each job's stream; order them before the launch
  for (remd_gpu *g : b->jobs) HIPCHK(hipStreamSynchronize(g->stream));
  if (b->split) HIPCHK(hipMemsetAsync(b->d_sync, 0, sizeof(unsigned long long) * (b->res_n + 1), b->stream));
  for (size_t k = 0; k < b->jobs.size(); k++) b->jobs[k]->last_parts = b->jobs[k]->resident ? b->plan[k] : 1;
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  // the global-state classes start with the resident launch when no replica
  // of it is split over workgroups; otherwise after it, so that no
  // global-state workgroup holds a CU a waiting part needs
  bool fork = false;
  if (!b->split) {
    HIPCHK(hipEventRecord(b->efork, b->stream));
    fork = true;
  }
  if (b->nblk) {
    LaunchArgs a;
    a.reps = b->d_table;
    a.nrep = b->res_n;
    a.mode = MODE_INTERVAL;
    a.do_sort = 0;
    a.margin = 0.0;
    a.parts = 1;
    a.blk = b->d_blk;
    a.nblk = b->nblk;
    a.same_xcd_ok = same_xcd_barrier_enabled();
    a.same_xcd_mode = same_xcd_barrier_mode();
    a.sync = b->d_sync;
    a.abort = (unsigned *)(b->d_sync + b->res_n);
    HIPCHK(launch_resident_batch(a, b->res_lds, b->stream));
  }
  for (int c = 0; c < NCLASS; c++) {
    if (!b->cls_n[c]) continue;
    if (b->cls_lds[c] > device_lds_limit()) return fail(REMD_EINVAL, "batch LDS exceeds the device limit");
    if (!fork) {
      HIPCHK(hipEventRecord(b->efork, b->stream));
      fork = true;
    }
    HIPCHK(hipStreamWaitEvent(b->cstream[c], b->efork, 0));
    LaunchArgs a;
    a.reps = b->d_table + b->cls_off[c];
    a.nrep = b->cls_n[c];
    a.mode = MODE_INTERVAL;
    a.do_sort = 0;
    a.margin = 0.0;
    a.parts = 1;
    a.same_xcd_ok = 0;
    a.same_xcd_mode = 0;
    a.sync = nullptr;
    a.abort = nullptr;
    const int nsite = 3 + (c % 6) / 2, prec = c % 2;
    HIPCHK(launch_interval(a, class_nt(c), nsite, prec, b->cls_lds[c], b->cstream[c]));
  }
  // join
  for (int c = 0; c < NCLASS; c++) {
    if (!b->cls_n[c]) continue;
    HIPCHK(hipEventRecord(b->cdone[c], b->cstream[c]));
    HIPCHK(hipStreamWaitEvent(b->stream, b->cdone[c], 0));
  }
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  if (b->split)
    HIPCHK(hipMemcpyAsync(b->h_abort, b->d_sync + b->res_n, sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
  for (remd_gpu *g : b->jobs)
    HIPCHK(hipMemcpyAsync(g->h_sc, g->d_sc, sizeof(double) * NSC * g->nrep, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipEventRecord(b->edone, b->stream));
  b->in_flight = true;
  return 0;
}

extern "C" int remd_batch_poll(remd_batch *b) {
  if (!b) return fail(REMD_EINVAL, "null batch");
  if (!b->in_flight) return 1;
  const hipError_t e = hipEventQuery(b->edone);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(REMD_EHIP, hipGetErrorString(e));
}

extern "C" int remd_batch_wait(remd_batch *b) {
  if (!b) return fail(REMD_EINVAL, "null batch");
  if (!b->in_flight) return 0;
  b->in_flight = false;
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->split && *b->h_abort != 0u) {
    *b->h_abort = 0u;
    return fail(REMD_EHIP, "resident batch launch aborted: the parts of a replica were not co-resident");
  }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  if (!b->tuned) {
    if (b->warm > 0) {
      b->warm--;
    } else {
      b->cand_ms[b->cand_cur] = ms;
      if (b->cand_cur + 1 >= (int)b->cands.size()) add_neighbours(b);
      if (b->cand_cur + 1 < (int)b->cands.size()) {
        b->cand_cur++;
      } else {
        b->tuned = true;
        b->cand_cur = (int)(std::min_element(b->cand_ms.begin(), b->cand_ms.end()) - b->cand_ms.begin());
      }
      int rc = apply_plan(b, b->cands[b->cand_cur]);
      if (rc) return rc;
    }
  }
  return 0;
}

extern "C" int remd_batch_execute(remd_batch *b) {
  int rc = remd_batch_launch(b);
  if (rc) return rc;
  return remd_batch_wait(b);
}

extern "C" int remd_batch_last_kernel_ms(remd_batch *b, double *ms) {
  if (!b || !ms) return fail(REMD_EINVAL, "null argument");
  *ms = b->last_ms;
  return 0;
}

extern "C" void remd_batch_destroy(remd_batch *b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  hipFree(b->d_table);
  hipFree(b->d_sync);
  hipFree(b->d_blk);
  if (b->h_abort) hipHostFree(b->h_abort);
  for (int c = 0; c < NCLASS; c++) {
    if (b->cdone[c]) hipEventDestroy(b->cdone[c]);
    if (b->cstream[c]) hipStreamDestroy(b->cstream[c]);
  }
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->efork) hipEventDestroy(b->efork);
  if (b->edone) hipEventDestroy(b->edone);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

// ---------------------------------------------------------------- rng
// glibc random() TYPE_3 (the generator behind rand()): additive lagged
// Fibonacci r[i] = r[i-3] + r[i-31] (mod 2^32), seeded with the
// 16807 mod (2^31-1) Lehmer sequence, first 310 outputs discarded, each
// output r[i] >> 1 (as described by P. Selinger, "The GLIBC pseudo-random
// number generator").  State: the last 34 words, k = next index mod 34.
extern "C" void remd_rng_seed(remd_rng *s, unsigned seed) {
  if (seed == 0) seed = 1;
  int32_t r[344];
  r[0] = (int32_t)seed;
  for (int i = 1; i < 31; i++) {
    long long w = (16807LL * r[i - 1]) % 2147483647LL;
    if (w < 0) w += 2147483647LL;
    r[i] = (int32_t)w;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 34; i < 344; i++) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
  for (int i = 0; i < 34; i++) s->r[i] = r[310 + i];
  s->k = 0;  // next index 344 == 310 + 34  ->  slot 0 of the ring
}

extern "C" int remd_rng_next(remd_rng *s) {
  const int k = s->k;
  const uint32_t v = (uint32_t)s->r[(k + 34 - 31) % 34] + (uint32_t)s->r[(k + 34 - 3) % 34];
  s->r[k] = (int32_t)v;
  s->k = (k + 1) % 34;
  return (int)(v >> 1);
}

// ------------------------------------------------------- host exchange
// Exchange pair list `type` (REMDInfo::ExchangeList, src/remdinfo.cxx:54-108):
// condition slots (a, b) in list order; 0: T-even, 1: P-even, 2: T-odd,
// 3: P-odd neighbours of the ntemp x npress grid (slot = t + p * ntemp).
template <class F>
static void for_each_pair(int dt, int dp, int type, F &&visit) {
  if (type == 0) {
    for (int j = 0; j < dp; ++j)
      for (int i = 0; i < dt / 2; ++i) visit(2 * i + j * dt, 2 * i + 1 + j * dt);
  } else if (type == 1) {
    for (int j = 0; j < dp / 2; ++j)
      for (int i = 0; i < dt; ++i) visit(i + 2 * j * dt, i + (2 * j + 1) * dt);
  } else if (type == 2) {
    for (int j = 0; j < dp; ++j)
      for (int i = 0; i < dt / 2 - (dt + 1) % 2; ++i) visit(2 * i + j * dt + 1, 2 * i + 1 + j * dt + 1);
  } else {
    for (int j = 0; j < dp / 2 - (dp + 1) % 2; ++j)
      for (int i = 0; i < dt; ++i) visit(i + (2 * j + 1) * dt, i + (2 * j + 2) * dt);
  }
}

extern "C" int remd_exchange_pairs(int dt, int dp, int type, int *a, int *b) {
  if (dt < 1 || dp < 1 || type < 0 || type > 3) return fail(REMD_EINVAL, "bad argument");
  int k = 0;
  for_each_pair(dt, dp, type, [&](int x, int y) {
    if (a) a[k] = x;
    if (b) b[k] = y;
    k++;
  });
  return k;
}

extern "C" int remd_exchange_host(int dt, int dp, long step, int pconstant, const double *pot, const double *vol,
                                  int *index, double *T, double *P, unsigned char *isx, remd_rng *rng,
                                  int *accepted, int *npairs) {
  if (dt < 1 || dp < 1 || !pot || !vol || !index || !T || !P || !rng) return fail(REMD_EINVAL, "bad argument");
  const int n = dt * dp;
  for (int k = 0; k < n; k++)
    if (index[k] < 0 || index[k] >= n) return fail(REMD_EINVAL, "index out of range");
  if (isx)
    for (int r = 0; r < n; r++) isx[r] = 0;
  int k = 0, nacc = 0;
  auto visit = [&](int a, int b) {
    const int ind1 = index[a], ind2 = index[b];
    const double beta1 = 1. / T[ind1], beta2 = 1. / T[ind2];
    double delta = (beta2 - beta1) * (pot[ind1] - pot[ind2]);
    if (pconstant > 0) delta += (beta2 * P[ind2] - beta1 * P[ind1]) * (vol[ind1] - vol[ind2]);
    const double u = (double)remd_rng_next(rng) / (double)2147483647;  // rand()/RAND_MAX
    const int acc = exp(-delta) > u;
    if (accepted) accepted[k] = acc;
    k++;
    if (acc) {
      if (isx) isx[ind1] = isx[ind2] = 1;
      std::swap(T[ind1], T[ind2]);
      std::swap(P[ind1], P[ind2]);
      std::swap(index[a], index[b]);
      nacc++;
    } else if (isx) {
      isx[ind1] = isx[ind2] = 0;
    }
  };
  for_each_pair(dt, dp, (int)(step % 4), visit);
  if (npairs) *npairs = k;
  return nacc;
}
