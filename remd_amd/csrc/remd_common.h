// remd_common.h -- definitions shared by the HIP kernels and the host engine.
//
// Device memory layout (one REMD job = one set of replicas of one system):
//   molecule block, per replica: NFIELD arrays of nmol doubles (SoA, sorted
//     molecule order): r[3] (fractional), v[3] (velocity*L*s), f[3],
//     q[4] (quaternion), p[4] (quaternion momentum * s), n[4] (torque in the
//     quaternion-4 form of src/molecules.cxx:409-418);
//   mid: nmol ints, molecule identity permuted with the sort;
//   scalar block, per replica: NSC doubles (thermostat, barostat, box,
//     conditions and the properties of src/variables.h:6-49).
#pragma once

#include <stdint.h>

namespace remd {

constexpr int MAXSITE = 5;

// molecule block field offsets (in units of nmol doubles)
enum Field {
  F_RX = 0, F_RY, F_RZ,
  F_VX, F_VY, F_VZ,
  F_FX, F_FY, F_FZ,
  F_Q0, F_Q1, F_Q2, F_Q3,
  F_P0, F_P1, F_P2, F_P3,
  F_N0, F_N1, F_N2, F_N3,
  NFIELD
};

// replica scalar block (doubles)
enum Scalar {
  S_S = 0, S_PS, S_Q,
  S_PV0, S_PV1, S_PV2, S_W,
  S_L0, S_L1, S_L2,
  S_T, S_P,            // current conditions (reduced)
  S_TNEW, S_PNEW,      // conditions requested for the next interval
  S_XFLAG,             // 1.0: the replica was exchanged; rescale first
  S_GKT, S_H0,
  S_LJ, S_CLMB, S_WALL, S_POT,
  S_VIR0, S_VIR1, S_VIR2,
  S_TRA0, S_TRA1, S_TRA2, S_TRA, S_ROT, S_KIN,
  S_TSM, S_TSP, S_TST, S_BST, S_TOT, S_HAM, S_TK,
  S_PRS0, S_PRS1, S_PRS2, S_PRS3,
  S_VOL,
  S_STEPS,             // steps integrated so far (diagnostic)
  S_DIAG0,             // 16 phase timers (REMD_PHASE_TIMERS builds only)
  NSC = S_DIAG0 + 16
};

// kernel modes
enum Mode {
  MODE_INIT = 0,       // sort (margin 0) + force + properties, H0 = tot
  MODE_INTERVAL = 1,   // [exchange rescale] + interval steps + properties
  MODE_FORCE = 2,      // force pass only (test hook), optional sort
};

struct JobParams {
  int nmol, nsite, qbeg;        // qbeg: first charged site (0: 3-site, 1: 4/5-site)
  int tmode, pmode, cmode;      // tconstant, pconstant, confined
  int sort_axis;                // 2 (z) bulk/tube, 0 (x) slit
  int per[3];                   // periodic axes
  int interval, sort_interval;
  int nfwall;
  int chunks;                   // j-chunks per i-molecule in the pair pass (power of two on the
                                // global-state path, any 1..64 on the resident path)
  int precision;                // 0 mixed (FP32 pair, FP64 acc), 1 FP64
  int compact;                  // 1: screen partners by COM distance before the pair terms
  int radix_min;                // resident path: molecules from which the neighbour sort is the
                                // wave-ballot radix sort (the rank sort below; same order)
  int ens_fixed;                // resident path: 1 runs jobs of the ENS_*_NP ensembles on the
                                // instantiations with the modes fixed (same bits as 0)
  double dt, dthalf;
  double rcut, rswitch, margin;
  double wall_length;           // A
  double bottom_area;           // pi (R - 1.7)^2 for the tube
  double mass, inertia[3];
  double inv_mass, inv_inertia[3]; // 1 / mass, 1 / inertia (host-computed)
  double site_r[MAXSITE][3];    // body-frame site coordinates
  double q[MAXSITE];            // reduced charges
  double lj_sig2, lj_eps4;      // site 0 LJ: sigma^2, 4 epsilon
  double unit_temp, unit_press;
  const double *fwall;          // nfwall x {f, e}
  // FP32 pair constants, precomputed on the host so the pair pass reads
  // them with scalar loads: rc, rl, rc^2, rl^2, 1/(rl-rc)^5, 30/(rl-rc)^5,
  // sigma^2, 4 eps, 48 eps, q_a q_b
  float f_rc, f_rl, f_rc2, f_rl2, f_coef, f_coef30, f_sig2, f_eps4, f_eps48;
  float f_qq[MAXSITE][MAXSITE];
};

struct ReplicaDesc {
  const JobParams *job;         // device pointer
  double *mol;                  // NFIELD * nmol doubles
  int *mid;                     // nmol
  double *sc;                   // NSC doubles
  double *site_f;               // optional: nsite*nmol*3 (last force pass)
  int *win;                     // optional: 3*nmol window (js, je, whole-ring flag)
  double *xbuf;                 // resident path with parts > 1: exchange rows (xbuf_parity_doubles)
  double *gbuf;                 // split-global path: one private state copy per part (gm_part_doubles)
  void *grec;                   // global-state path, replicas whose site records do not fit the
                                // LDS: the records in global memory (nmol x stride x FP32/FP64)
};

// Resident path: threads per workgroup (4 waves: one per SIMD of a CU).
constexpr int RNT = 256;

// Resident path: default molecule count from which the neighbour sort is
// the radix sort rather than the rank sort.
constexpr int RADIX_MIN_DEFAULT = 768;

// Ensemble code of a job: tconstant 0..2, pconstant 0..3, confined 0..2;
// ENS_RUNTIME: the modes are read from the job.
constexpr int ENS_RUNTIME = -1;
constexpr int ens_code(int tconstant, int pconstant, int confined) {
  return (tconstant * 4 + pconstant) * 3 + confined;
}
constexpr int ens_tconstant(int code) { return code / 12; }
constexpr int ens_pconstant(int code) { return (code / 3) % 4; }
constexpr int ens_confined(int code) { return code % 3; }
// Ensembles the resident kernels (single-job and batch) are also built for
// with the modes fixed (FP32 pair math): Nose-Poincare thermostat with the
// 1D barostat in a tube, the area barostat in a slit, the 3D barostat in
// bulk.
constexpr int ENS_TUBE_NP = ens_code(1, 2, 1);
constexpr int ENS_SLIT_NP = ens_code(1, 3, 2);
constexpr int ENS_BULK_NP = ens_code(1, 1, 0);

// Resident path: molecules per energy block.  A replica's parts own whole
// blocks, and the energy/virial totals are sums of block sums.
constexpr int EB = 4;

// Resident path exchange buffer, per replica and step parity: one 64-byte
// row per molecule (F[3], n[4]) and one per energy block (lj, clmb, wall,
// vir[3]).  A part owns whole blocks, so the rows of one cache line have one
// writer (block rows at part boundaries excepted).
constexpr int XR = 8;  // doubles per molecule row
constexpr int XB = 8;  // doubles per energy-block row
__host__ __device__ inline size_t xbuf_parity_doubles(int nmol) {
  return (size_t)nmol * XR + (size_t)((nmol + EB - 1) / EB) * XB;
}

// Split-global path: a part's private copy of the replica state in global
// memory (NFIELD x nmol doubles, the energy-block sums 6 x nblk, then nmol
// ints), in doubles.
__host__ __device__ inline size_t gm_part_doubles(int nmol) {
  return (size_t)NFIELD * nmol + (size_t)6 * ((nmol + EB - 1) / EB) + (size_t)((nmol + 1) / 2);
}

struct LaunchArgs {
  const ReplicaDesc *reps;
  int nrep;
  int mode;
  int do_sort;                  // MODE_FORCE only
  double margin;                // MODE_FORCE only
  int parts;                    // resident path: workgroups per replica (blk == nullptr)
  const unsigned *blk = nullptr; // resident batch launch: per-workgroup entry (rep << 12 |
                                 // variant << 8 | parts << 4 | part), ~0u for an idle workgroup
  int nblk = 0;                  // workgroups of a resident batch launch
  unsigned long long *sync;     // resident path, parts > 1: nrep group words, zeroed (low 32
                                // bits: arrivals; high 32: XCD histogram of the parts)
  unsigned *abort;              // resident path, parts > 1: abort word, zeroed
  int same_xcd_ok;              // resident path: may use the same-XCD group barrier
  int same_xcd_mode;            // same-XCD barrier: bit 0 agent-scope release, bit 1 agent-scope
                                // acquire (else: stores drained to L2 / device-scope loads)
};

// One job's device exchange (exchange_kernel): grid, pair list type, the
// condition arrays and the uniforms in pair-list order.
struct ExchangeJob {
  int ntemp, npress, pmode, ec_type;
  int *index;         // condition slot -> replica
  double *T, *P;      // per replica conditions (reduced)
  const double *u;    // uniforms, pair-list order
  double *const *sc;  // per replica scalar blocks
  int *accepted;      // per pair, may be null
  // remd_gpu_execute_intervals: the interval's scalar blocks (nrep x NSC,
  // contiguous from sc[0]) copied to hist before the exchange, and the
  // group words of the next interval launch cleared (both may be null)
  double *hist;
  unsigned long long *sync;
};

}  // namespace remd
