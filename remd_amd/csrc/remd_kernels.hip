// remd_kernels.hip -- the REMD per-replica hot path for gfx950 (MI355X).
//
// One persistent workgroup per replica runs a whole exchange interval:
//   Nose-Poincare thermostat + Andersen barostat rigid-body splitting
//   D6 D5 D4 D3 D2 D1 D2 D3 (PsPv) D4 | force | D5 D6
//   (src/integrator.cxx:646-706, the CPU canonical order), with
//   - the z-sorted neighbour window rebuilt every sort_interval steps by a
//     wave-ballot LSD radix sort in LDS (replaces src/sort.cu +
//     src/quicksort.cu; window semantics of src/molecules.cxx:541-585),
//   - M2A into LDS site records (FP32 centred COM + FP32 site offsets,
//     rotated in FP64; src/molecules.cxx:319-360, src/rotation.h:29-48),
//   - the switched LJ + Coulomb pair pass (src/force_calculator.cxx:803-989
//     semantics) with FP32 pair math and FP64 accumulation, work items
//     (i, chunk) mapped to adjacent lanes and reduced by a fixed xor
//     butterfly (deterministic, independent of what else is resident),
//   - the wall table lookup (src/force_calculator.cxx:1494-1531) and A2M
//     torque (src/molecules.cxx:401-439) in FP64,
//   - block reductions in FP64 with a fixed tree.
// All O(N) integrator state is FP64; only the O(N * neighbours) pair math is
// FP32 (precision == 1 switches it to FP64 for the precision study).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "remd_common.h"

namespace remd {

constexpr int WAVE = 64;

// Pointers read out of ReplicaDesc would be generic to the compiler; the
// device view below carries their address spaces in the loaded types, so
// job parameters load through the scalar unit (constant) and state arrays
// with global rather than flat instructions.
#define REMD_AS_CONST __attribute__((address_space(4)))
#define REMD_AS_GLOBAL __attribute__((address_space(1)))
struct DescD {
  const REMD_AS_CONST JobParams *job;
  REMD_AS_GLOBAL double *mol;
  REMD_AS_GLOBAL int *mid;
  REMD_AS_GLOBAL double *sc;
  REMD_AS_GLOBAL double *site_f;
  REMD_AS_GLOBAL int *win;
  REMD_AS_GLOBAL double *xbuf;
  REMD_AS_GLOBAL double *gbuf;
  REMD_AS_GLOBAL char *grec;
};
static_assert(sizeof(DescD) == sizeof(ReplicaDesc), "device view of ReplicaDesc");
template <class T>
__device__ __forceinline__ T *gen(REMD_AS_GLOBAL T *p) {
  return (T *)p;
}
typedef const REMD_AS_CONST JobParams &JRef;

// Ensemble of a resident-path instantiation.  ENS_RUNTIME reads the
// thermostat, barostat and confinement modes from the job; a code from
// ens_code() fixes them at compile time.  That removes the branches on the
// modes (and the registers that hold them) and the minimum image on axes
// that are not periodic (a no-op there), so both forms give the same bits.
template <int EM>
__device__ __forceinline__ int tmode_of(JRef J) {
  if constexpr (EM < 0) return J.tmode;
  else return ens_tconstant(EM);
}
template <int EM>
__device__ __forceinline__ int pmode_of(JRef J) {
  if constexpr (EM < 0) return J.pmode;
  else return ens_pconstant(EM);
}
template <int EM>
__device__ __forceinline__ int cmode_of(JRef J) {
  if constexpr (EM < 0) return J.cmode;
  else return ens_confined(EM);
}
// periodic axes follow the confinement (bit d: axis d; x, y not in the
// tube, z not in the slit); -1: read from the job
template <int EM>
__device__ __host__ constexpr int per_mask_of() {
  return EM < 0 ? -1 : (ens_confined(EM) == 1 ? 4 : (ens_confined(EM) == 2 ? 3 : 7));
}
// the sort axis follows the confinement: x for the slit, z otherwise
template <int EM>
__device__ __forceinline__ int sort_axis_of(JRef J) {
  if constexpr (EM < 0) return J.sort_axis;
  else return ens_confined(EM) == 2 ? 0 : 2;
}



// Optional phase timers (constant 100 MHz clock), accumulated by thread 0
// into 16 accumulators (the S_DIAG0.. words of the replica's scalar block,
// or LDS on the resident path); compiled out by default.
#ifdef REMD_PHASE_TIMERS
#define PT_DECL uint64_t pt_t0 = wall_clock64();
#define PT_MARK(k, sc)                                        \
  do {                                                        \
    const uint64_t pt_t1 = wall_clock64();                    \
    if (threadIdx.x == 0 && (sc)) (sc)[(k)] += (double)(pt_t1 - pt_t0); \
    pt_t0 = pt_t1;                                            \
  } while (0)
#else
#define PT_DECL
#define PT_MARK(k, sc) \
  do {                 \
  } while (0)
#endif

// ------------------------------------------------------------ LDS layout
template <int NS>
struct Rec {
  // COM (3) + NS sites x 3, padded for 8/16-byte LDS reads
  static constexpr int STRIDE = (NS == 5) ? 18 : (NS == 4 ? 16 : 12);
};

__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

// Bytes of dynamic LDS a replica of nmol molecules needs.
__host__ __device__ inline int lds_bytes(int nmol, int nsite, int prec, int nt, bool grec = false) {
  const int stride = (nsite == 5) ? 18 : (nsite == 4 ? 16 : 12);
  const int rec = grec ? 0 : nmol * stride * (prec ? 8 : 4);
  const int sortb = 20 * nmol;
  int region0 = align16(rec > sortb ? rec : sortb);
  const int win = align16(4 * nmol);
  const int misc = (nt / WAVE) * 32 * 8 + (16 * 64 + 16) * 4;
  return region0 + win + align16(misc);
}

struct Lds {
  char *region0;
  char *rec;     // site records: region 0, or global memory (grec)
  uint16_t *js, *je;
  double *red;   // NW x 32 doubles
  int *cnt;      // radix counts
};

template <int NT>
__device__ inline Lds lds_carve(char *smem, int nmol, int nsite, int prec, char *grec = nullptr) {
  const int stride = (nsite == 5) ? 18 : (nsite == 4 ? 16 : 12);
  const int rec = grec ? 0 : nmol * stride * (prec ? 8 : 4);
  const int sortb = 20 * nmol;
  const int region0 = align16(rec > sortb ? rec : sortb);
  Lds L;
  L.region0 = smem;
  L.rec = grec ? grec : smem;
  L.js = (uint16_t *)(smem + region0);
  L.je = L.js + nmol;
  char *misc = smem + region0 + align16(4 * nmol);
  L.red = (double *)misc;
  L.cnt = (int *)(misc + (NT / WAVE) * 32 * 8);
  return L;
}

// ------------------------------------------------------------ reductions
// Lane exchange inside rows of 16 lanes (DPP control CTRL) of a double.
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a full wave; every lane receives the same bits.
// Inside each row of 16 lanes: quad_perm swaps (lane ^ 1, lane ^ 2), then
// the half-row mirror (lane 7 - l) and the row mirror (lane 15 - l); each
// step adds the same two operands as the partner lane (a + b == b + a
// exactly), so a row's lanes agree bit for bit.  The four row sums are read
// from lanes 0, 16, 32 and 48 and added in a fixed order.
__device__ inline double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const double r0 = __hiloint2double(__builtin_amdgcn_readlane(hi, 0), __builtin_amdgcn_readlane(lo, 0));
  const double r1 = __hiloint2double(__builtin_amdgcn_readlane(hi, 16), __builtin_amdgcn_readlane(lo, 16));
  const double r2 = __hiloint2double(__builtin_amdgcn_readlane(hi, 32), __builtin_amdgcn_readlane(lo, 32));
  const double r3 = __hiloint2double(__builtin_amdgcn_readlane(hi, 48), __builtin_amdgcn_readlane(lo, 48));
  return (r0 + r1) + (r2 + r3);
}

// Sum K doubles over the workgroup; every thread receives the same result.
// With SYNC_AFTER = false the caller guarantees a workgroup barrier before
// red is written again.
template <int NT, int K, bool SYNC_AFTER = true>
__device__ inline void block_sum(double (&v)[K], double *red) {
  constexpr int NW = NT / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) red[w * K + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    double s = 0.0;
    for (int ww = 0; ww < NW; ww++) s += red[ww * K + k];
    v[k] = s;
  }
  if (SYNC_AFTER) __syncthreads();
}

// ------------------------------------------------------- replica scalars
struct Sc {
  double s, Ps, Q, Pv[3], W, L[3], T, P, gkT, H0;
  double lj, clmb, wall, pot, vir[3];
  double iQ, iW;  // 1 / Q, 1 / W
};

__device__ inline void load_sc(Sc &c, const double *sc) {
  c.s = sc[S_S]; c.Ps = sc[S_PS]; c.Q = sc[S_Q];
  c.Pv[0] = sc[S_PV0]; c.Pv[1] = sc[S_PV1]; c.Pv[2] = sc[S_PV2]; c.W = sc[S_W];
  c.iQ = 1.0 / c.Q;
  c.iW = 1.0 / c.W;
  c.L[0] = sc[S_L0]; c.L[1] = sc[S_L1]; c.L[2] = sc[S_L2];
  c.T = sc[S_T]; c.P = sc[S_P]; c.gkT = sc[S_GKT]; c.H0 = sc[S_H0];
  c.lj = sc[S_LJ]; c.clmb = sc[S_CLMB]; c.wall = sc[S_WALL]; c.pot = sc[S_POT];
  c.vir[0] = sc[S_VIR0]; c.vir[1] = sc[S_VIR1]; c.vir[2] = sc[S_VIR2];
}

__device__ inline void store_sc(const Sc &c, double *sc) {
  sc[S_S] = c.s; sc[S_PS] = c.Ps; sc[S_Q] = c.Q;
  sc[S_PV0] = c.Pv[0]; sc[S_PV1] = c.Pv[1]; sc[S_PV2] = c.Pv[2]; sc[S_W] = c.W;
  sc[S_L0] = c.L[0]; sc[S_L1] = c.L[1]; sc[S_L2] = c.L[2];
  sc[S_T] = c.T; sc[S_P] = c.P; sc[S_GKT] = c.gkT; sc[S_H0] = c.H0;
  sc[S_LJ] = c.lj; sc[S_CLMB] = c.clmb; sc[S_WALL] = c.wall; sc[S_POT] = c.pot;
  sc[S_VIR0] = c.vir[0]; sc[S_VIR1] = c.vir[1]; sc[S_VIR2] = c.vir[2];
}

// Molecules::GetVolume / GetBottomArea (src/molecules.h:357-376)
template <int EM = ENS_RUNTIME>
__device__ inline double volume(JRef J, double L0, double L1, double L2) {
  if (cmode_of<EM>(J) == 1) return J.bottom_area * L2;
  if (cmode_of<EM>(J) == 2) return L0 * L1 * J.wall_length;
  return L0 * L1 * L2;
}

// --------------------------------------------------------------- rotations
// body_to_space (src/rotation.h:29-48)
__device__ inline void body_to_space(const double *a, const double *v, double *t) {
  const double a0sq = a[0] * a[0], a1sq = a[1] * a[1], a2sq = a[2] * a[2], a3sq = a[3] * a[3];
  const double a01 = a[0] * a[1], a02 = a[0] * a[2], a03 = a[0] * a[3];
  const double a12 = a[1] * a[2], a13 = a[1] * a[3], a23 = a[2] * a[3];
  t[0] = (a0sq + a1sq - a2sq - a3sq) * v[0] + 2. * (a12 - a03) * v[1] + 2. * (a13 + a02) * v[2];
  t[1] = 2. * (a12 + a03) * v[0] + (a0sq - a1sq + a2sq - a3sq) * v[1] + 2. * (a23 - a01) * v[2];
  t[2] = 2. * (a13 - a02) * v[0] + 2. * (a23 + a01) * v[1] + (a0sq - a1sq - a2sq + a3sq) * v[2];
}

// space_to_body (src/rotation.h:7-27)
__device__ inline void space_to_body(const double *a, const double *v, double *t) {
  const double a0sq = a[0] * a[0], a1sq = a[1] * a[1], a2sq = a[2] * a[2], a3sq = a[3] * a[3];
  const double a01 = a[0] * a[1], a02 = a[0] * a[2], a03 = a[0] * a[3];
  const double a12 = a[1] * a[2], a13 = a[1] * a[3], a23 = a[2] * a[3];
  t[0] = (a0sq + a1sq - a2sq - a3sq) * v[0] + 2. * (a12 + a03) * v[1] + 2. * (a13 - a02) * v[2];
  t[1] = 2. * (a12 - a03) * v[0] + (a0sq - a1sq + a2sq - a3sq) * v[1] + 2. * (a23 + a01) * v[2];
  t[2] = 2. * (a13 + a02) * v[0] + 2. * (a23 - a01) * v[1] + (a0sq - a1sq - a2sq + a3sq) * v[2];
}

// sin and cos of a rotation sub-step angle, from their Taylor series (no
// library call, so the five rotations of a step stay short, branch-light
// code).  |x| < 0.5 rad (every free rotation of an MD step at these
// temperatures): sin through x^15, cos through x^16; truncation below
// 1e-19 relative, under FP64 rounding of the result.  Larger angles:
// Cody-Waite reduction by pi/2 with a three-part constant (exact products
// for |k| < 2^20) to |r| <= pi/4, then sin through r^19 and cos through
// r^20 (truncation below 1e-21) and the quadrant's sign and swap.
__device__ inline void sincos_poly(double x, double *sn, double *cs, bool wide) {
  const double x2 = x * x;
  if (!wide) {
    *sn = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
          x2 * (-1.0 / 39916800.0 + x2 * (1.0 / 6227020800.0 + x2 * (-1.0 / 1307674368000.0))))))));
    *cs = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0 +
          x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0 + x2 * (1.0 / 20922789888000.0))))))));
  } else {
    *sn = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
          x2 * (-1.0 / 39916800.0 + x2 * (1.0 / 6227020800.0 + x2 * (-1.0 / 1307674368000.0 +
          x2 * (1.0 / 355687428096000.0 + x2 * (-1.0 / 121645100408832000.0))))))))));
    *cs = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0 +
          x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0 + x2 * (1.0 / 20922789888000.0 +
          x2 * (-1.0 / 6402373705728000.0 + x2 * (1.0 / 2432902008176640000.0))))))))));
  }
}

__device__ inline void step_sincos(double x, double *sn, double *cs) {
  if (fabs(x) < 0.5) {
    sincos_poly(x, sn, cs, false);
    return;
  }
  // pi/2 = P1 + P2 + P3: P1 and P2 carry 33 significant bits each, so k * P1
  // and k * P2 are exact for |k| < 2^20
  const double P1 = 1.5707963267341256;
  const double P2 = 6.077100506303966e-11;
  const double P3 = 2.0222662487959506e-21;
  const double k = rint(x * 0.63661977236758134308);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  double s, c;
  sincos_poly(r, &s, &c, true);
  const int q = (int)k & 3;
  const double ss = (q & 1) ? c : s;
  const double cc = (q & 1) ? s : c;
  *sn = (q & 2) ? -ss : ss;
  *cs = ((q + 1) & 2) ? -cc : cc;
}

// Free rotation about body axis k (src/integrator.cxx:467-563, P1..P3 of
// src/integrator.h:321-346).  Returns xi.
__device__ inline double rotate(double *q, double *p, int axis, double coef_xi, double invI, double h) {
  double Pq[4], Pp[4];
  if (axis == 0) {
    Pq[0] = -q[1]; Pq[1] = q[0]; Pq[2] = q[3]; Pq[3] = -q[2];
    Pp[0] = -p[1]; Pp[1] = p[0]; Pp[2] = p[3]; Pp[3] = -p[2];
  } else if (axis == 1) {
    Pq[0] = -q[2]; Pq[1] = -q[3]; Pq[2] = q[0]; Pq[3] = q[1];
    Pp[0] = -p[2]; Pp[1] = -p[3]; Pp[2] = p[0]; Pp[3] = p[1];
  } else {
    Pq[0] = -q[3]; Pq[1] = q[2]; Pq[2] = -q[1]; Pq[3] = q[0];
    Pp[0] = -p[3]; Pp[1] = p[2]; Pp[2] = -p[1]; Pp[3] = p[0];
  }
  const double xi = (p[0] * Pq[0] + p[1] * Pq[1] + p[2] * Pq[2] + p[3] * Pq[3]) * coef_xi * invI;
  double sn, c;
  step_sincos(xi * h, &sn, &c);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const double qn = q[d] * c + Pq[d] * sn;
    const double pn = p[d] * c + Pp[d] * sn;
    q[d] = qn;
    p[d] = pn;
  }
  return xi;
}

// ------------------------------------------------------------- integrator
template <int EM = ENS_RUNTIME>
__device__ inline int teff(JRef J) { return tmode_of<EM>(J) == 1 ? 1 : 0; }

// D6 (src/integrator.cxx:634-644)
template <int EM = ENS_RUNTIME>
__device__ inline void D6(JRef J, Sc &c) {
  if (teff<EM>(J)) {
    const double tmp = 1.0 + (c.Ps * (0.5 * c.iQ) * J.dthalf);
    c.s *= tmp * tmp;
    c.Ps /= tmp;
  }
}

// scalar part of D5 (src/integrator.cxx:581-600); pconstant 3 is the area
// barostat (Lx = Ly = sqrt(A), conjugate momentum in Pv[0]).
template <int EM = ENS_RUNTIME>
__device__ inline void D5_scalars(JRef J, Sc &c) {
  if (teff<EM>(J)) {
    c.Ps -= c.pot * J.dthalf;
    if (pmode_of<EM>(J) > 0) c.Ps -= (c.P * volume<EM>(J, c.L[0], c.L[1], c.L[2])) * J.dthalf;
  }
  if (pmode_of<EM>(J) == 1)
    c.Pv[0] += ((c.vir[0] + c.vir[1] + c.vir[2]) / (3.0 * c.L[0] * c.L[1] * c.L[2]) - c.P) * c.s * J.dthalf;
  if (pmode_of<EM>(J) == 2) c.Pv[2] += c.s * (c.vir[2] / c.L[2] - c.P * J.bottom_area) * J.dthalf;
  if (pmode_of<EM>(J) == 3) {
    const double A = c.L[0] * c.L[1];
    c.Pv[0] += c.s * ((c.vir[0] + c.vir[1]) / (2.0 * A) - c.P * J.wall_length) * J.dthalf;
  }
}

// D4 (src/integrator.cxx:604-632; pconstant 1 with the Pv^2/(2W) kinetic
// term of remdgpu.cu:2090-2095)
template <int EM = ENS_RUNTIME>
__device__ inline void D4(JRef J, Sc &c) {
  if (pmode_of<EM>(J) == 1) {
    if (teff<EM>(J)) c.Ps -= 0.5 * c.Pv[0] * c.Pv[0] * c.iW * J.dthalf;
    const double V = volume<EM>(J, c.L[0], c.L[1], c.L[2]) + c.Pv[0] * c.s * c.iW * J.dthalf;
    const double l = cbrt(V);
    c.L[0] = c.L[1] = c.L[2] = l;
  }
  if (pmode_of<EM>(J) == 2) {
    if (teff<EM>(J)) c.Ps -= 0.5 * c.Pv[2] * c.Pv[2] * c.iW * J.dthalf;
    c.L[2] += c.s * c.Pv[2] * c.iW * J.dthalf;
  }
  if (pmode_of<EM>(J) == 3) {
    if (teff<EM>(J)) c.Ps -= 0.5 * c.Pv[0] * c.Pv[0] * c.iW * J.dthalf;
    const double A = c.L[0] * c.L[1] + c.s * c.Pv[0] * c.iW * J.dthalf;
    const double l = sqrt(A);
    c.L[0] = c.L[1] = l;
  }
}

// D5 on one molecule (src/integrator.cxx:564-580): v += f L s dt/2 / m,
// p += S(q) n 2 s dt/2.
__device__ inline void D5_mol(JRef J, const Sc &c, double *v, const double *f,
                              double *p, const double *q, const double *n) {
  const double cm = c.s * J.dthalf * J.inv_mass;
  v[0] += f[0] * (c.L[0] * cm);
  v[1] += f[1] * (c.L[1] * cm);
  v[2] += f[2] * (c.L[2] * cm);
  const double cp = 2.0 * c.s * J.dthalf;
  const double d0 = q[0] * n[0] - q[1] * n[1] - q[2] * n[2] - q[3] * n[3];
  const double d1 = q[1] * n[0] + q[0] * n[1] - q[3] * n[2] + q[2] * n[3];
  const double d2 = q[2] * n[0] + q[3] * n[1] + q[0] * n[2] - q[1] * n[3];
  const double d3 = q[3] * n[0] - q[2] * n[1] + q[1] * n[2] + q[0] * n[3];
  p[0] += d0 * cp;
  p[1] += d1 * cp;
  p[2] += d2 * cp;
  p[3] += d3 * cp;
}

// ------------------------------------------------------------ properties
// Molecules::CalcProperties (src/molecules.cxx:702-734) over the replica.
template <int NT, int EM = ENS_RUNTIME>
__device__ void calc_properties(JRef J, const double *mol, int N, Sc &c, double *sc,
                                double *red) {
  double acc[4] = {0, 0, 0, 0};
  const double *I = (const double *)J.inertia;
  for (int i = threadIdx.x; i < N; i += NT) {
    const double vx = mol[F_VX * N + i], vy = mol[F_VY * N + i], vz = mol[F_VZ * N + i];
    acc[0] += (vx * vx) * J.mass;
    acc[1] += (vy * vy) * J.mass;
    acc[2] += (vz * vz) * J.mass;
    const double q0 = mol[F_Q0 * N + i], q1 = mol[F_Q1 * N + i], q2 = mol[F_Q2 * N + i], q3 = mol[F_Q3 * N + i];
    const double p0 = mol[F_P0 * N + i], p1 = mol[F_P1 * N + i], p2 = mol[F_P2 * N + i], p3 = mol[F_P3 * N + i];
    const double coef = 0.25 / c.s;
    double x0 = (-p0 * q1 + p1 * q0 + p2 * q3 - p3 * q2) * ((1.0 / I[0]) * coef);
    double x1 = (-p0 * q2 - p1 * q3 + p2 * q0 + p3 * q1) * ((1.0 / I[1]) * coef);
    double x2 = (-p0 * q3 + p1 * q2 - p2 * q1 + p3 * q0) * ((1.0 / I[2]) * coef);
    acc[3] += I[0] * x0 * x0 * 2.0 + I[1] * x1 * x1 * 2.0 + I[2] * x2 * x2 * 2.0;
  }
  block_sum<NT, 4>(acc, red);
  double tra[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double cf = 1.0 / (c.L[d] * c.s);
    tra[d] = acc[d] * (cf * cf * 0.5);
  }
  const double tr = tra[0] + tra[1] + tra[2];
  const double rot = acc[3];
  const double kin = tr + rot;
  const double tsm = 0.5 * c.Ps * c.Ps / c.Q;
  const double tsp = c.gkT * log(c.s);
  double bst = 0.0;
  if (pmode_of<EM>(J) > 0) {
    if (cmode_of<EM>(J) == 0)
      bst = 0.5 * (c.Pv[0] * c.Pv[0] + c.Pv[1] * c.Pv[1] + c.Pv[2] * c.Pv[2]) / c.W + c.P * c.L[0] * c.L[1] * c.L[2];
    else if (cmode_of<EM>(J) == 1)
      bst = 0.5 * c.Pv[2] * c.Pv[2] / c.W + c.P * volume<EM>(J, c.L[0], c.L[1], c.L[2]);
    else
      bst = 0.5 * c.Pv[0] * c.Pv[0] / c.W + c.P * volume<EM>(J, c.L[0], c.L[1], c.L[2]);
  }
  double tot = c.pot + kin;
  if (tmode_of<EM>(J) == 1) tot += tsm + tsp;
  if (pmode_of<EM>(J) > 0) tot += bst;
  const double V = volume<EM>(J, c.L[0], c.L[1], c.L[2]);
  if (threadIdx.x == 0 && sc) {
    sc[S_TRA0] = tra[0]; sc[S_TRA1] = tra[1]; sc[S_TRA2] = tra[2];
    sc[S_TRA] = tr; sc[S_ROT] = rot; sc[S_KIN] = kin;
    sc[S_TSM] = tsm; sc[S_TSP] = tsp; sc[S_TST] = tsm + tsp; sc[S_BST] = bst;
    sc[S_TOT] = tot; sc[S_HAM] = tot - c.H0;
    sc[S_TK] = kin * J.unit_temp / (3. * N);
    for (int d = 0; d < 3; d++) sc[S_PRS0 + d] = (tra[d] * 2.0 + c.vir[d]) / V * J.unit_press;
    sc[S_PRS3] = ((tra[0] * 2.0 + c.vir[0]) + (tra[1] * 2.0 + c.vir[1]) + (tra[2] * 2.0 + c.vir[2])) /
                 (3. * V) * J.unit_press;
    sc[S_VOL] = V;
  }
  __syncthreads();
}

// Returns tot (same value on every thread) without writing sc.
template <int NT, int EM = ENS_RUNTIME>
__device__ double total_energy(JRef J, const double *mol, int N, const Sc &c, double *red) {
  double acc[4] = {0, 0, 0, 0};
  const double *I = (const double *)J.inertia;
  for (int i = threadIdx.x; i < N; i += NT) {
    const double vx = mol[F_VX * N + i], vy = mol[F_VY * N + i], vz = mol[F_VZ * N + i];
    acc[0] += (vx * vx) * J.mass;
    acc[1] += (vy * vy) * J.mass;
    acc[2] += (vz * vz) * J.mass;
    const double q0 = mol[F_Q0 * N + i], q1 = mol[F_Q1 * N + i], q2 = mol[F_Q2 * N + i], q3 = mol[F_Q3 * N + i];
    const double p0 = mol[F_P0 * N + i], p1 = mol[F_P1 * N + i], p2 = mol[F_P2 * N + i], p3 = mol[F_P3 * N + i];
    const double coef = 0.25 / c.s;
    double x0 = (-p0 * q1 + p1 * q0 + p2 * q3 - p3 * q2) * ((1.0 / I[0]) * coef);
    double x1 = (-p0 * q2 - p1 * q3 + p2 * q0 + p3 * q1) * ((1.0 / I[1]) * coef);
    double x2 = (-p0 * q3 + p1 * q2 - p2 * q1 + p3 * q0) * ((1.0 / I[2]) * coef);
    acc[3] += I[0] * x0 * x0 * 2.0 + I[1] * x1 * x1 * 2.0 + I[2] * x2 * x2 * 2.0;
  }
  block_sum<NT, 4>(acc, red);
  double tra = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double cf = 1.0 / (c.L[d] * c.s);
    tra += acc[d] * (cf * cf * 0.5);
  }
  const double kin = tra + acc[3];
  double tot = c.pot + kin;
  if (tmode_of<EM>(J) == 1) tot += 0.5 * c.Ps * c.Ps / c.Q + c.gkT * log(c.s);
  if (pmode_of<EM>(J) > 0) {
    if (cmode_of<EM>(J) == 0)
      tot += 0.5 * (c.Pv[0] * c.Pv[0] + c.Pv[1] * c.Pv[1] + c.Pv[2] * c.Pv[2]) / c.W + c.P * c.L[0] * c.L[1] * c.L[2];
    else if (cmode_of<EM>(J) == 1)
      tot += 0.5 * c.Pv[2] * c.Pv[2] / c.W + c.P * volume<EM>(J, c.L[0], c.L[1], c.L[2]);
    else
      tot += 0.5 * c.Pv[0] * c.Pv[0] / c.W + c.P * volume<EM>(J, c.L[0], c.L[1], c.L[2]);
  }
  return tot;
}

// KillMomentum + KillAngularMomentumInTube (src/molecules.cxx:168-210)
template <int NT, int EM = ENS_RUNTIME>
__device__ void kill_momentum(JRef J, double *mol, int N, const Sc &c, double *red,
                              bool tube_twice) {
  double acc[4] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < N; i += NT) {
    acc[0] += mol[F_VX * N + i] * J.mass;
    acc[1] += mol[F_VY * N + i] * J.mass;
    acc[2] += mol[F_VZ * N + i] * J.mass;
    acc[3] += J.mass;
  }
  block_sum<NT, 4>(acc, red);
  const double sum[3] = {acc[0] / acc[3], acc[1] / acc[3], acc[2] / acc[3]};
  for (int i = threadIdx.x; i < N; i += NT) {
    if (cmode_of<EM>(J) == 1) {
      mol[F_VZ * N + i] -= sum[2];
    } else if (cmode_of<EM>(J) == 2) {
      mol[F_VX * N + i] -= sum[0];
      mol[F_VY * N + i] -= sum[1];
    } else {
      mol[F_VX * N + i] -= sum[0];
      mol[F_VY * N + i] -= sum[1];
      mol[F_VZ * N + i] -= sum[2];
    }
  }
  __syncthreads();
  const int reps = (cmode_of<EM>(J) == 1) ? (tube_twice ? 2 : 1) : 0;
  for (int rpt = 0; rpt < reps; rpt++) {
    double a2[2] = {0, 0};
    for (int i = threadIdx.x; i < N; i += NT) {
      const double rx = (mol[F_RX * N + i] - 0.5) * c.L[0];
      const double ry = (mol[F_RY * N + i] - 0.5) * c.L[1];
      const double vx = mol[F_VX * N + i] / (c.L[0] * c.s);
      const double vy = mol[F_VY * N + i] / (c.L[1] * c.s);
      a2[0] += J.mass * (rx * vy - ry * vx);
      a2[1] += J.mass;
    }
    block_sum<NT, 2>(a2, red);
    const double mom = a2[0] / a2[1];
    for (int i = threadIdx.x; i < N; i += NT) {
      const double rx = (mol[F_RX * N + i] - 0.5) * c.L[0];
      const double ry = (mol[F_RY * N + i] - 0.5) * c.L[1];
      const double r02 = rx * rx + ry * ry;
      if (r02 > 1.0) {
        mol[F_VX * N + i] += mom * ry / r02 * c.L[0] * c.s;
        mol[F_VY * N + i] -= mom * rx / r02 * c.L[1] * c.s;
      }
    }
    __syncthreads();
  }
}

// VelocityScaling + AngVelocityScaling (src/molecules.cxx:261-311)
template <int NT>
__device__ void velocity_scaling(JRef J, double *mol, int N, Sc &c, double *red) {
  double acc[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < N; i += NT) {
    acc[0] += mol[F_VX * N + i] * mol[F_VX * N + i] * J.mass;
    acc[1] += mol[F_VY * N + i] * mol[F_VY * N + i] * J.mass;
    acc[2] += mol[F_VZ * N + i] * mol[F_VZ * N + i] * J.mass;
  }
  block_sum<NT, 3>(acc, red);
  double tra = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double cf = 1.0 / (c.L[d] * c.s);
    tra += acc[d] * (cf * cf * 0.5);
  }
  if (tra != 0.0) {
    const double coef = sqrt(1.5 * N * c.T / tra);
    for (int i = threadIdx.x; i < N; i += NT) {
      mol[F_VX * N + i] *= coef;
      mol[F_VY * N + i] *= coef;
      mol[F_VZ * N + i] *= coef;
    }
    __syncthreads();
    kill_momentum<NT>(J, mol, N, c, red, true);
    c.s = 1.0;
    c.Ps = 0.0;
    c.Pv[0] = c.Pv[1] = c.Pv[2] = 0.0;
  }
  // AngVelocityScaling
  double rr[1] = {0.0};
  const double *I = (const double *)J.inertia;
  for (int i = threadIdx.x; i < N; i += NT) {
    const double q0 = mol[F_Q0 * N + i], q1 = mol[F_Q1 * N + i], q2 = mol[F_Q2 * N + i], q3 = mol[F_Q3 * N + i];
    const double p0 = mol[F_P0 * N + i], p1 = mol[F_P1 * N + i], p2 = mol[F_P2 * N + i], p3 = mol[F_P3 * N + i];
    const double coef = 0.25 / c.s;
    double x0 = (-p0 * q1 + p1 * q0 + p2 * q3 - p3 * q2) * ((1.0 / I[0]) * coef);
    double x1 = (-p0 * q2 - p1 * q3 + p2 * q0 + p3 * q1) * ((1.0 / I[1]) * coef);
    double x2 = (-p0 * q3 + p1 * q2 - p2 * q1 + p3 * q0) * ((1.0 / I[2]) * coef);
    rr[0] += I[0] * x0 * x0 * 2.0 + I[1] * x1 * x1 * 2.0 + I[2] * x2 * x2 * 2.0;
  }
  block_sum<NT, 1>(rr, red);
  if (rr[0] != 0.0) {
    const double coef = sqrt(1.5 * N * c.T / rr[0]);
    for (int i = threadIdx.x; i < N; i += NT)
      for (int d = 0; d < 4; d++) mol[(F_P0 + d) * N + i] *= coef;
    __syncthreads();
  }
}

// ------------------------------------------------------ neighbour build
// Wave-ballot LSD radix sort of 64-bit keys with 16-bit payloads, 4-bit
// digits, stable: within a 64-lane group the rank is the popcount of the
// lanes below with the same digit; groups are scanned in element order.
template <int NT>
__device__ void radix_sort(uint64_t *keyA, uint16_t *idxA, uint64_t *keyB, uint16_t *idxB, int *cnt,
                           int n) {
  constexpr int NW = NT / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const int ntiles = (n + NT - 1) / NT;
  const int ng = ntiles * NW;
  const uint64_t below = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
#pragma unroll 1
  for (int pass = 0; pass < 16; ++pass) {
    const int shift = 4 * pass;
    for (int t = 0; t < ntiles; ++t) {
      const int e = t * NT + threadIdx.x;
      const bool valid = e < n;
      const int dg = valid ? (int)((keyA[e] >> shift) & 15ull) : 16;
      const int g = t * NW + w;
      for (int v = 0; v < 16; ++v) {
        const uint64_t b = __ballot(dg == v);
        if (lane == v) cnt[v * ng + g] = __popcll(b);
      }
    }
    __syncthreads();
    if (threadIdx.x < 16) {
      const int v = threadIdx.x;
      int s = 0;
      for (int g = 0; g < ng; ++g) {
        const int c = cnt[v * ng + g];
        cnt[v * ng + g] = s;
        s += c;
      }
      cnt[16 * ng + v] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int v = 0; v < 16; ++v) {
        const int c = cnt[16 * ng + v];
        cnt[16 * ng + v] = s;
        s += c;
      }
    }
    __syncthreads();
    for (int t = 0; t < ntiles; ++t) {
      const int e = t * NT + threadIdx.x;
      const bool valid = e < n;
      const uint64_t key = valid ? keyA[e] : 0ull;
      const uint16_t id = valid ? idxA[e] : (uint16_t)0;
      const int dg = (int)((key >> shift) & 15ull);
      uint64_t m = __ballot(valid);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int bit = (dg >> k) & 1;
        const uint64_t b = __ballot(valid && bit);
        m &= bit ? b : ~b;
      }
      const int rank = __popcll(m & below);
      const int g = t * NW + w;
      if (valid) {
        const int pos = cnt[16 * ng + dg] + cnt[dg * ng + g] + rank;
        keyB[pos] = key;
        idxB[pos] = id;
      }
    }
    __syncthreads();
    uint64_t *tk = keyA; keyA = keyB; keyB = tk;
    uint16_t *ti = idxA; idxA = idxB; idxB = ti;
  }
}

// Window bounds for nlane = 1 (src/molecules.cxx:556-583) on sorted keys z:
// js = first index scanning down from i (cyclic) with z < z_i - w,
// je = first index scanning up (cyclic) with z > z_i + w.
__device__ inline int count_less(const double *z, int n, double x) {
  int lo = 0, hi = n;  // first index with z >= x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (z[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ inline int count_leq(const double *z, int n, double x) {
  int lo = 0, hi = n;  // first index with z > x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (z[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
// The scan wraps at most twice for w < 2: a second wrap stops at once, on
// index n - 1 (downward) or 0 (upward).
__device__ inline void window_of(const double *z, int n, int i, double w, int &js, int &je) {
  double lb = z[i] - w;
  int k = count_less(z, n, lb);
  if (k >= 1) js = k - 1;
  else {
    lb += 1.0;
    k = count_less(z, n, lb);
    js = (k >= 1) ? k - 1 : n - 1;
  }
  double ub = z[i] + w;
  int m = count_leq(z, n, ub);
  if (m < n) je = m;
  else {
    ub -= 1.0;
    m = count_leq(z, n, ub);
    je = (m < n) ? m : 0;
  }
}

// Contract (no reference counterpart): the window is the whole ring when
// the scan wrapped onto i itself, or when the first key outside on one
// side lies inside the other side's arc; the reference's js..je range then
// misses pairs.  Same comparisons as the scan's bounds.
__device__ inline bool window_full(const double *z, int i, double w, int js, int je) {
  const bool a_up = (js > i) ? (z[js] <= z[i] + w) : (z[js] <= (z[i] + w) - 1.0);
  const bool b_dn = (je < i) ? (z[je] >= z[i] - w) : (z[je] >= (z[i] - w) + 1.0);
  return js == i || je == i || a_up || b_dn;
}

// Sort the replica's molecules by the sort-axis coordinate, permute every
// state array, rebuild js/je (CalcForcePot's doSort branch,
// src/molecules.cxx:543-585).
template <int NT>
__device__ void sort_and_window(JRef J, double *mol, int *mid, int N, double Laxis,
                                double margin, const Lds &S, int *win_dbg) {
  uint64_t *keyA = (uint64_t *)S.region0;
  uint64_t *keyB = keyA + N;
  uint16_t *idxA = (uint16_t *)(keyB + N);
  uint16_t *idxB = idxA + N;
  const double *ax = mol + (F_RX + J.sort_axis) * N;
  for (int i = threadIdx.x; i < N; i += NT) {
    const double z = ax[i] + 0.0;  // -0.0 -> +0.0
    keyA[i] = (uint64_t)__double_as_longlong(z);
    idxA[i] = (uint16_t)i;
  }
  __syncthreads();
  radix_sort<NT>(keyA, idxA, keyB, idxB, S.cnt, N);
  // permute the molecule block through LDS (keyB area as staging)
  double *tmp = (double *)keyB;
  for (int f = 0; f < NFIELD; ++f) {
    double *col = mol + f * N;
    for (int i = threadIdx.x; i < N; i += NT) tmp[i] = col[i];
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += NT) col[k] = tmp[idxA[k]];
    __syncthreads();
  }
  {
    int *tmpi = (int *)keyB;
    for (int i = threadIdx.x; i < N; i += NT) tmpi[i] = mid[i];
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += NT) mid[k] = tmpi[idxA[k]];
    __syncthreads();
  }
  // sorted coordinate into LDS, then the windows
  double *zs = (double *)keyA;
  for (int i = threadIdx.x; i < N; i += NT) zs[i] = ax[i];
  __syncthreads();
  const double w = (J.rcut + margin) / Laxis;
  for (int i = threadIdx.x; i < N; i += NT) {
    int js, je;
    window_of(zs, N, i, w, js, je);
    const bool full = window_full(zs, i, w, js, je);
    if (win_dbg) {
      win_dbg[3 * i] = js;
      win_dbg[3 * i + 1] = je;
      win_dbg[3 * i + 2] = full ? 1 : 0;
    }
    if (js >= je) je += N;
    // a whole-ring window, or a scanned range that misses i: every other
    // molecule is a candidate
    if (full || !((js <= i && i <= je) || (js <= i + N && i + N <= je))) {
      js = i + 1;
      je = i + N - 1;
    }
    S.js[i] = (uint16_t)js;
    S.je[i] = (uint16_t)je;
  }
  __syncthreads();
}

// ---------------------------------------------------------------- forces
template <int NS, typename PT>
struct PairK {
  PT rc, rl, rc2, rl2, coef, coef30;
  PT rs2;  // screening radius^2, slightly above rc^2 (pair_acc decides)
  PT sig2, eps4, eps48;
  PT qq[NS][NS];
  PT L[3], Li[3];  // L: the box length on periodic axes, 0 on the others (no wrap there)
};

template <int NS, typename AT = double>
struct Acc {
  AT site[NS][3];
  AT corr[3];
  AT vir[3];
  AT lj, clmb;
};

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double rcp(double x) { return 1.0 / x; }
__device__ inline float rnd(float x) { return rintf(x); }
__device__ inline double rnd(double x) { return rint(x); }

// Minimum image of a COM difference on axis d.  PER < 0: every axis, with
// L = 0 on the non-periodic ones; PER >= 0: the mask of periodic axes fixed
// by the ensemble at compile time, and the non-periodic axes skip the
// no-op.  x - 0 * round(x / L) is x bit for bit for every finite x (a zero
// difference of two centred coordinates is +0, and +0 - 0 * 0 == +0), so
// both give the same bits.
template <int PER, int NS, typename PT>
__device__ __forceinline__ PT min_image(PT x, const PairK<NS, PT> &k, int d) {
  if constexpr (PER < 0) return x - k.L[d] * rnd(x * k.Li[d]);
  else return ((PER >> d) & 1) ? x - k.L[d] * rnd(x * k.Li[d]) : x;
}

// One ordered molecule pair (i, j): switched LJ (site 0 - site 0) + Coulomb
// over charged site pairs, CPU tuned-kernel semantics
// (src/force_calculator.cxx:874-989).  Pair math in PT, accumulation in AT.
// Branch-free: a pair at or beyond the cutoff gets switching factors 0 (as
// the CPU kernel's r02 = min(rc^2, r02) does, :883), so every sum receives
// a signed zero and keeps its bits; `self` (j == i) moves the partner far
// outside the cutoff first, so that every term stays finite.
template <int NS, int QB, typename PT, int PER, typename AT>
__device__ __forceinline__ void pair_acc(const PT *gi, const PT (&ai)[NS][3], const PT *rj,
                                         const PairK<NS, PT> &k, bool self, Acc<NS, AT> &acc) {
  PT dg[3];
#pragma unroll
  for (int d = 0; d < 3; d++) dg[d] = min_image<PER>(gi[d] - rj[d], k, d);
  dg[2] = self ? PT(1.0e4) : dg[2];
  const PT r2 = dg[0] * dg[0] + dg[1] * dg[1] + dg[2] * dg[2];
  const bool inside = r2 < k.rc2;
  PT sw, dswr;
  {
    const PT ri = rsq(r2);
    const PT r = r2 * ri;
    const PT c_ = r - k.rc;
    const PT l_ = r - k.rl;
    const PT s5 = k.coef * c_ * c_ * c_ * (PT(10) * l_ * l_ - PT(5) * l_ * c_ + c_ * c_);
    const PT d5 = k.coef30 * c_ * c_ * l_ * l_ * ri;
    const bool switched = r2 > k.rl2;
    sw = inside ? (switched ? s5 : PT(1)) : PT(0);
    dswr = (inside && switched) ? d5 : PT(0);
  }
  PT di[NS][3];
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int d = 0; d < 3; d++) di[a][d] = dg[d] + ai[a][d];
  PT fs[NS][3];
  PT elj, ecl = PT(0);
  {  // LJ between the two site-0 (oxygen) sites
    const PT x = di[0][0] - rj[3], y = di[0][1] - rj[4], z = di[0][2] - rj[5];
    const PT rr = x * x + y * y + z * z;
    const PT ri2 = rcp(rr);
    const PT s2 = k.sig2 * ri2;
    const PT s6 = s2 * s2 * s2;
    elj = k.eps4 * s6 * (s6 - PT(1));
    const PT f = k.eps48 * s6 * (s6 - PT(0.5)) * ri2;
    fs[0][0] = f * x;
    fs[0][1] = f * y;
    fs[0][2] = f * z;
  }
#pragma unroll
  for (int a = 1; a < NS; a++) fs[a][0] = fs[a][1] = fs[a][2] = PT(0);
#pragma unroll
  for (int a = QB; a < NS; a++) {
#pragma unroll
    for (int b = QB; b < NS; b++) {
      const PT x = di[a][0] - rj[3 + 3 * b], y = di[a][1] - rj[4 + 3 * b], z = di[a][2] - rj[5 + 3 * b];
      const PT rr = x * x + y * y + z * z;
      const PT ri = rsq(rr);
      const PT e = k.qq[a][b] * ri;
      const PT f = e * ri * ri;
      ecl += e;
      fs[a][0] += f * x;
      fs[a][1] += f * y;
      fs[a][2] += f * z;
    }
  }
  // switched site forces into the item's sums (one multiply-add each) and
  // the pair's COM force -dS/dr E g + S sum_a f_a for the virial
  const PT cr = -dswr * (elj + ecl);
  PT fsum[3] = {fs[0][0], fs[0][1], fs[0][2]};
#pragma unroll
  for (int a = 1; a < NS; a++) {
    fsum[0] += fs[a][0];
    fsum[1] += fs[a][1];
    fsum[2] += fs[a][2];
  }
#pragma unroll
  for (int a = 0; a < NS; a++) {
    acc.site[a][0] += (AT)(sw * fs[a][0]);
    acc.site[a][1] += (AT)(sw * fs[a][1]);
    acc.site[a][2] += (AT)(sw * fs[a][2]);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    acc.corr[d] += (AT)(cr * dg[d]);
    const PT F = cr * dg[d] + sw * fsum[d];
    acc.vir[d] += (AT)(F * dg[d]);
  }
  acc.lj += (AT)(sw * elj);
  acc.clmb += (AT)(sw * ecl);
}

template <int NS, typename PT>
__device__ __forceinline__ void load_rec(PT (&r)[3 + 3 * NS], const PT *src) {
#pragma unroll
  for (int u = 0; u < 3 + 3 * NS; u++) r[u] = src[u];
}

// The partners of work item (i, c): jj = js + c, js + c + C, ... <= je
// (indices mod N), each through pair_acc, in this order.  The record of the
// next partner is loaded while the current pair is computed.  With
// `compact` the COM distances of the next 32 partners are screened first
// (radius slightly above the cutoff) and only the ones inside go through
// pair_acc, still in order: the same pairs and the same sums, with less
// work where most partners lie beyond the cutoff (bulk and slit windows).
template <int NS, int QB, typename PT, int PER, typename AT>
__device__ __forceinline__ void pair_window_run(const PT *rec, int N, int i, int js, int je, int c0, int C,
                                                bool compact, const PT *gi, const PT (&ai)[NS][3],
                                                const PairK<NS, PT> &k, Acc<NS, AT> &acc) {
  constexpr int ST = Rec<NS>::STRIDE;
  constexpr int RL = 3 + 3 * NS;
  if (!compact) {
    int jj = js + c0;
    if (jj > je) return;
    int j = (jj >= N) ? jj - N : jj;
    PT rj[RL];
    load_rec<NS, PT>(rj, rec + j * ST);
    for (;;) {
      const int jn = jj + C;
      const bool more = jn <= je;
      const int jx = more ? ((jn >= N) ? jn - N : jn) : j;
      PT rn[RL];
      load_rec<NS, PT>(rn, rec + jx * ST);
      pair_acc<NS, QB, PT, PER>(gi, ai, rj, k, j == i, acc);
      if (!more) break;
      jj = jn;
      j = jx;
#pragma unroll
      for (int u = 0; u < RL; u++) rj[u] = rn[u];
    }
    return;
  }
  for (int base = js + c0; base <= je; base += 32 * C) {
    unsigned mask = 0u;
    {
      int jj = base;
      int j = (jj >= N) ? jj - N : jj;
      PT c[3] = {rec[j * ST], rec[j * ST + 1], rec[j * ST + 2]};
      for (int u = 0; u < 32; u++) {
        const int jn = jj + C;
        const bool more = u < 31 && jn <= je;
        const int jx = more ? ((jn >= N) ? jn - N : jn) : j;
        const PT cn[3] = {rec[jx * ST], rec[jx * ST + 1], rec[jx * ST + 2]};
        PT dg[3];
#pragma unroll
        for (int d = 0; d < 3; d++) dg[d] = min_image<PER>(gi[d] - c[d], k, d);
        const PT r2 = dg[0] * dg[0] + dg[1] * dg[1] + dg[2] * dg[2];
        if (j != i && r2 < k.rs2) mask |= 1u << u;
        if (!more) break;
        jj = jn;
        j = jx;
        c[0] = cn[0];
        c[1] = cn[1];
        c[2] = cn[2];
      }
    }
    if (!mask) continue;
    int u = __builtin_ctz(mask);
    mask &= mask - 1u;
    int j = base + u * C;
    j = (j >= N) ? j - N : j;
    PT rj[RL];
    load_rec<NS, PT>(rj, rec + j * ST);
    for (;;) {
      const bool more = mask != 0u;
      int jx = j;
      if (more) {
        jx = base + __builtin_ctz(mask) * C;
        jx = (jx >= N) ? jx - N : jx;
        mask &= mask - 1u;
      }
      PT rn[RL];
      load_rec<NS, PT>(rn, rec + jx * ST);
      pair_acc<NS, QB, PT, PER>(gi, ai, rj, k, false, acc);
      if (!more) break;
      j = jx;
#pragma unroll
      for (int v = 0; v < RL; v++) rj[v] = rn[v];
    }
  }
}

// One work item's partner sums into `acc` (zero on entry).  FP32 pair math
// sums its chunk (about window/C partners) in FP32 and converts once; FP64
// pair math sums in FP64.  The chunks are then reduced in FP64.  A chunk's
// partners do not depend on how a replica is split over workgroups, so the
// results stay the same bits for any split.
template <int NS, int QB, typename PT, int PER = -1>
__device__ __forceinline__ void pair_window(const PT *rec, int N, int i, int js, int je, int c0, int C,
                                            bool compact, const PT *gi, const PT (&ai)[NS][3],
                                            const PairK<NS, PT> &k, Acc<NS> &acc) {
  if constexpr (sizeof(PT) == 4) {
    Acc<NS, float> part;
#pragma unroll
    for (int a = 0; a < NS; a++) part.site[a][0] = part.site[a][1] = part.site[a][2] = 0.0f;
    part.corr[0] = part.corr[1] = part.corr[2] = 0.0f;
    part.vir[0] = part.vir[1] = part.vir[2] = 0.0f;
    part.lj = part.clmb = 0.0f;
    pair_window_run<NS, QB, PT, PER>(rec, N, i, js, je, c0, C, compact, gi, ai, k, part);
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int d = 0; d < 3; d++) acc.site[a][d] = (double)part.site[a][d];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      acc.corr[d] = (double)part.corr[d];
      acc.vir[d] = (double)part.vir[d];
    }
    acc.lj = (double)part.lj;
    acc.clmb = (double)part.clmb;
  } else {
    pair_window_run<NS, QB, PT, PER>(rec, N, i, js, je, c0, C, compact, gi, ai, k, acc);
  }
}

// M2A: LDS records (COM centred on the box, site offsets R(q) b).
template <int NT, int NS, typename PT>
__device__ void m2a(JRef J, const double *mol, int N, double L0, double L1, double L2, PT *rec) {
  constexpr int ST = Rec<NS>::STRIDE;
  for (int i = threadIdx.x; i < N; i += NT) {
    PT *o = rec + i * ST;
    o[0] = (PT)((mol[F_RX * N + i] - 0.5) * L0);
    o[1] = (PT)((mol[F_RY * N + i] - 0.5) * L1);
    o[2] = (PT)((mol[F_RZ * N + i] - 0.5) * L2);
    const double q[4] = {mol[F_Q0 * N + i], mol[F_Q1 * N + i], mol[F_Q2 * N + i], mol[F_Q3 * N + i]};
#pragma unroll
    for (int a = 0; a < NS; a++) {
      double t[3];
      body_to_space(q, (const double *)J.site_r[a], t);
      o[3 + 3 * a] = (PT)t[0];
      o[4 + 3 * a] = (PT)t[1];
      o[5 + 3 * a] = (PT)t[2];
    }
  }
}

// Pair pass + wall + A2M for every molecule of the replica.  Writes f, n
// into the molecule block and the energy/virial totals into c.
template <int NT, int NS, typename PT>
__device__ void force_pass(JRef J, double *mol, int N, Sc &c, const Lds &S,
                           double *site_dbg, double *pt_sc) {
  PT_DECL
  constexpr int QB = (NS == 3) ? 0 : 1;
  constexpr int ST = Rec<NS>::STRIDE;
  PT *rec = (PT *)S.rec;
  m2a<NT, NS, PT>(J, mol, N, c.L[0], c.L[1], c.L[2], rec);
  __syncthreads();
  PT_MARK(4, pt_sc);

  PairK<NS, PT> k;
  if (sizeof(PT) == 4) {
    k.rc = J.f_rc;
    k.rl = J.f_rl;
    k.rc2 = J.f_rc2;
    k.rl2 = J.f_rl2;
    k.coef = J.f_coef;
    k.coef30 = J.f_coef30;
    k.sig2 = J.f_sig2;
    k.eps4 = J.f_eps4;
    k.eps48 = J.f_eps48;
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int b = 0; b < NS; b++) k.qq[a][b] = J.f_qq[a][b];
  } else {
    k.rc = (PT)J.rcut;
    k.rl = (PT)J.rswitch;
    k.rc2 = (PT)(J.rcut * J.rcut);
    k.rl2 = (PT)(J.rswitch * J.rswitch);
    const double rlc = J.rswitch - J.rcut;
    const double coef = 1.0 / (rlc * rlc * rlc * rlc * rlc);
    k.coef = (PT)coef;
    k.coef30 = (PT)(30.0 * coef);
    k.sig2 = (PT)J.lj_sig2;
    k.eps4 = (PT)J.lj_eps4;
    k.eps48 = (PT)(12.0 * J.lj_eps4);
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int b = 0; b < NS; b++) k.qq[a][b] = (PT)(J.q[a] * J.q[b]);
  }
  k.rs2 = k.rc2 * (PT)1.0001;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    k.L[d] = J.per[d] ? (PT)c.L[d] : PT(0);
    k.Li[d] = (PT)(1.0 / c.L[d]);
  }

  const int C = J.chunks;
  const int lc = __ffs(C) - 1;  // log2(C)
  double e_lj = 0.0, e_cl = 0.0, e_wall = 0.0, vir[3] = {0.0, 0.0, 0.0};
  const double dr = (J.cmode == 1) ? J.wall_length / (double)J.nfwall
                                   : 0.5 * J.wall_length / (double)J.nfwall;
  const double dri = 1.0 / dr;
  const int total = N << lc;
  for (int base = 0; base < total; base += NT) {
    const int t = base + threadIdx.x;
    const bool active = t < total;
    const int i = active ? (t >> lc) : 0;
    const int cc = t & (C - 1);
    Acc<NS> acc;
#pragma unroll
    for (int a = 0; a < NS; a++) acc.site[a][0] = acc.site[a][1] = acc.site[a][2] = 0.0;
    acc.corr[0] = acc.corr[1] = acc.corr[2] = 0.0;
    acc.vir[0] = acc.vir[1] = acc.vir[2] = 0.0;
    acc.lj = acc.clmb = 0.0;
    if (active) {
      const PT *ri = rec + i * ST;
      PT gi[3] = {ri[0], ri[1], ri[2]};
      PT ai[NS][3];
#pragma unroll
      for (int a = 0; a < NS; a++) {
        ai[a][0] = ri[3 + 3 * a];
        ai[a][1] = ri[4 + 3 * a];
        ai[a][2] = ri[5 + 3 * a];
      }
      pair_window<NS, QB, PT>(rec, N, i, S.js[i], S.je[i], cc, C, J.compact != 0, gi, ai, k, acc);
    }
    // fixed xor butterfly over the C lanes of molecule i
    for (int m = 1; m < C; m <<= 1) {
#pragma unroll
      for (int a = 0; a < NS; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) acc.site[a][d] += __shfl_xor(acc.site[a][d], m, WAVE);
#pragma unroll
  for (int d = 0; d < 3; d++) {
        acc.corr[d] += __shfl_xor(acc.corr[d], m, WAVE);
        acc.vir[d] += __shfl_xor(acc.vir[d], m, WAVE);
      }
      acc.lj += __shfl_xor(acc.lj, m, WAVE);
      acc.clmb += __shfl_xor(acc.clmb, m, WAVE);
    }
    if (active && cc == 0) {
      // wall on the site-0 (oxygen) position, FP64
      const double q[4] = {mol[F_Q0 * N + i], mol[F_Q1 * N + i], mol[F_Q2 * N + i], mol[F_Q3 * N + i]};
      double F[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
        double s = acc.corr[d];
#pragma unroll
        for (int a = 0; a < NS; a++) s += acc.site[a][d];
        F[d] = s;
      }
      if (J.cmode == 1 || J.cmode == 2) {
        double b0[3];
        body_to_space(q, (const double *)J.site_r[0], b0);
        if (J.cmode == 1) {
          const double x = (mol[F_RX * N + i] - 0.5) * c.L[0] + b0[0];
          const double y = (mol[F_RY * N + i] - 0.5) * c.L[1] + b0[1];
          const double r01 = sqrt(x * x + y * y);
          int idx = (int)(r01 / dr);
          if (idx > J.nfwall - 2) idx = J.nfwall - 2;
          const double f = (J.fwall[2 * (idx + 1)] - J.fwall[2 * idx]) * dri * (r01 - dr * idx) + J.fwall[2 * idx];
          const double e = (J.fwall[2 * (idx + 1) + 1] - J.fwall[2 * idx + 1]) * dri * (r01 - dr * idx) + J.fwall[2 * idx + 1];
          acc.site[0][0] += f * x;
          acc.site[0][1] += f * y;
          F[0] += f * x;
          F[1] += f * y;
          e_wall += e;
        } else {
          const double z = (mol[F_RZ * N + i] - 0.5) * c.L[2] + b0[2];
          const double az = fabs(z);
          int idx = (int)(az / dr);
          if (idx > J.nfwall - 2) idx = J.nfwall - 2;
          const double f = (J.fwall[2 * (idx + 1)] - J.fwall[2 * idx]) * dri * (az - dr * idx) + J.fwall[2 * idx];
          const double e = (J.fwall[2 * (idx + 1) + 1] - J.fwall[2 * idx + 1]) * dri * (az - dr * idx) + J.fwall[2 * idx + 1];
          acc.site[0][2] += f * z;
          F[2] += f * z;
          e_wall += e;
        }
      }
      // A2M torque (src/molecules.cxx:409-418)
      double n4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < NS; a++) {
        double fb[3];
        space_to_body(q, acc.site[a], fb);
        const double *br = (const double *)J.site_r[a];
        n4[0] += br[0] * fb[0] + br[1] * fb[1] + br[2] * fb[2];
        n4[1] += br[1] * fb[2] - br[2] * fb[1];
        n4[2] += br[2] * fb[0] - br[0] * fb[2];
        n4[3] += br[0] * fb[1] - br[1] * fb[0];
      }
      mol[F_FX * N + i] = F[0];
      mol[F_FY * N + i] = F[1];
      mol[F_FZ * N + i] = F[2];
      mol[F_N0 * N + i] = n4[0];
      mol[F_N1 * N + i] = n4[1];
      mol[F_N2 * N + i] = n4[2];
      mol[F_N3 * N + i] = n4[3];
      if (site_dbg) {
#pragma unroll
        for (int a = 0; a < NS; a++)
#pragma unroll
          for (int d = 0; d < 3; d++) site_dbg[(i * NS + a) * 3 + d] = acc.site[a][d];
      }
      e_lj += acc.lj;
      e_cl += acc.clmb;
      vir[0] += acc.vir[0];
      vir[1] += acc.vir[1];
      vir[2] += acc.vir[2];
    }
  }
  PT_MARK(5, pt_sc);
  double tot[6] = {e_lj, e_cl, e_wall, vir[0], vir[1], vir[2]};
  block_sum<NT, 6>(tot, S.red);
  PT_MARK(6, pt_sc);
  c.lj = 0.5 * tot[0];
  c.clmb = 0.5 * tot[1];
  c.wall = tot[2];
  c.vir[0] = 0.5 * tot[3];
  c.vir[1] = 0.5 * tot[4];
  c.vir[2] = 0.5 * tot[5];
  c.pot = c.lj + c.clmb + c.wall;
}

// ----------------------------------------------------------- one step
// First half of ExecuteStep (src/integrator.cxx:651-678): D6, D5 scalars,
// D4, then per molecule D5 (old box), D3 D2 D1 D2 D3 (new box), PsPv, D4.
// First half of a step (src/integrator.cxx:646-685).  With FUSE the second
// half of the previous step (src/integrator.cxx:686-703) runs first, in the
// same molecule loop: its scalar part (D5 scalars, D6) before the loop and
// its D5 on each molecule, with the box and s it saw, right before this
// step's D5 on that molecule -- the same operations in the same order as
// second_half() followed by first_half(), one pass over the molecules.
template <int NT, bool FUSE = false, int EM = ENS_RUNTIME>
__device__ void first_half(JRef J, double *mol, int N, Sc &c, double *red) {
  Sc c_prev;  // D5 of the previous step's second half: box and s before its scalars
  if (FUSE) {
    c_prev.s = c.s;
    c_prev.L[0] = c.L[0];
    c_prev.L[1] = c.L[1];
    c_prev.L[2] = c.L[2];
    D5_scalars<EM>(J, c);
    D6<EM>(J, c);
  }
  D6<EM>(J, c);
  Sc c_old;  // D5 on the molecules uses the box and s from before D4
  c_old.s = c.s;
  c_old.L[0] = c.L[0];
  c_old.L[1] = c.L[1];
  c_old.L[2] = c.L[2];
  D5_scalars<EM>(J, c);
  D4<EM>(J, c);
  const bool T1 = teff<EM>(J) == 1;
  const bool needPv = (pmode_of<EM>(J) > 0) || T1;
  double sums[4] = {0, 0, 0, 0};  // sum_Pv[3], sum_Ps
  // PsPv factors (box and s are fixed through the molecule loop)
  // reciprocals of s and of the box once; the factors below are products
  const double si = 1.0 / c.s;
  double Li[3], ipv[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    Li[d] = 1.0 / c.L[d];
    ipv[d] = (Li[d] * Li[d]) * (si * si);
  }
  const double logs = T1 ? log(c.s) : 0.0;
  double fpv = 0.0;
  if (pmode_of<EM>(J) == 1) fpv = c.s / (volume<EM>(J, c.L[0], c.L[1], c.L[2]) * 3.0) * J.dthalf;
  if (pmode_of<EM>(J) == 2) fpv = c.s * Li[2] * J.dthalf;
  if (pmode_of<EM>(J) == 3) fpv = c.s / (2.0 * c.L[0] * c.L[1]) * J.dthalf;
  {
    const double coef_xi = 0.25 * si;
    double coef_r[3];
#pragma unroll
    for (int d = 0; d < 3; d++) coef_r[d] = Li[d] * Li[d] * si * J.dt;
    const double *I = (const double *)J.inertia;
    const double *iI = (const double *)J.inv_inertia;
    for (int i = threadIdx.x; i < N; i += NT) {
      double r[3] = {mol[F_RX * N + i], mol[F_RY * N + i], mol[F_RZ * N + i]};
      double v[3] = {mol[F_VX * N + i], mol[F_VY * N + i], mol[F_VZ * N + i]};
      const double f[3] = {mol[F_FX * N + i], mol[F_FY * N + i], mol[F_FZ * N + i]};
      double q[4] = {mol[F_Q0 * N + i], mol[F_Q1 * N + i], mol[F_Q2 * N + i], mol[F_Q3 * N + i]};
      double p[4] = {mol[F_P0 * N + i], mol[F_P1 * N + i], mol[F_P2 * N + i], mol[F_P3 * N + i]};
      const double n[4] = {mol[F_N0 * N + i], mol[F_N1 * N + i], mol[F_N2 * N + i], mol[F_N3 * N + i]};
      if (FUSE) D5_mol(J, c_prev, v, f, p, q, n);
      D5_mol(J, c_old, v, f, p, q, n);
      double xi;
      xi = rotate(q, p, 2, coef_xi, iI[2], J.dthalf);
      if (T1) sums[3] += 0.5 * xi * xi * I[2];
      xi = rotate(q, p, 1, coef_xi, iI[1], J.dthalf);
      if (T1) sums[3] += 0.5 * xi * xi * I[1];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        r[d] += v[d] * coef_r[d];
        if (r[d] < 0.0) r[d] += 1.0;
        if (r[d] >= 1.0) r[d] -= 1.0;
      }
      xi = rotate(q, p, 0, 0.25 * si, iI[0], J.dt);
      if (needPv) {
        sums[0] += 2.0 * J.mass * v[0] * v[0];
        sums[1] += 2.0 * J.mass * v[1] * v[1];
        sums[2] += 2.0 * J.mass * v[2] * v[2];
      }
      if (T1) sums[3] += xi * xi * I[0];
      xi = rotate(q, p, 1, coef_xi, iI[1], J.dthalf);
      if (T1) sums[3] += 0.5 * xi * xi * I[1];
      xi = rotate(q, p, 2, coef_xi, iI[2], J.dthalf);
      if (T1) sums[3] += 0.5 * xi * xi * I[2];
      mol[F_RX * N + i] = r[0]; mol[F_RY * N + i] = r[1]; mol[F_RZ * N + i] = r[2];
      mol[F_VX * N + i] = v[0]; mol[F_VY * N + i] = v[1]; mol[F_VZ * N + i] = v[2];
      mol[F_Q0 * N + i] = q[0]; mol[F_Q1 * N + i] = q[1]; mol[F_Q2 * N + i] = q[2]; mol[F_Q3 * N + i] = q[3];
      mol[F_P0 * N + i] = p[0]; mol[F_P1 * N + i] = p[1]; mol[F_P2 * N + i] = p[2]; mol[F_P3 * N + i] = p[3];
    }
  }

  // includes __syncthreads (molecule writes visible); the next write of red
  // follows later workgroup barriers
  block_sum<NT, 4, false>(sums, red);
  // PsPv (src/integrator.cxx:145-171)
  if (needPv) {
    for (int d = 0; d < 3; d++) sums[d] *= ipv[d];
  }
  if (pmode_of<EM>(J) == 1) c.Pv[0] += (sums[0] + sums[1] + sums[2]) * fpv;
  if (pmode_of<EM>(J) == 2) c.Pv[2] += sums[2] * fpv;
  if (pmode_of<EM>(J) == 3) c.Pv[0] += (sums[0] + sums[1]) * fpv;
  if (T1) c.Ps += (0.25 * (sums[0] + sums[1] + sums[2]) + 2.0 * sums[3] - (1.0 + logs) * c.gkT + c.H0) * J.dt;
  D4<EM>(J, c);
}

// Second half (src/integrator.cxx:686-703): D5 on every molecule with the
// new forces, D5 scalars, D6.  Ends with a workgroup barrier.
template <int NT, int EM = ENS_RUNTIME>
__device__ void second_half(JRef J, double *mol, int N, Sc &c) {
  for (int i = threadIdx.x; i < N; i += NT) {
    double v[3] = {mol[F_VX * N + i], mol[F_VY * N + i], mol[F_VZ * N + i]};
    const double f[3] = {mol[F_FX * N + i], mol[F_FY * N + i], mol[F_FZ * N + i]};
    const double q[4] = {mol[F_Q0 * N + i], mol[F_Q1 * N + i], mol[F_Q2 * N + i], mol[F_Q3 * N + i]};
    double p[4] = {mol[F_P0 * N + i], mol[F_P1 * N + i], mol[F_P2 * N + i], mol[F_P3 * N + i]};
    const double n[4] = {mol[F_N0 * N + i], mol[F_N1 * N + i], mol[F_N2 * N + i], mol[F_N3 * N + i]};
    D5_mol(J, c, v, f, p, q, n);
    mol[F_VX * N + i] = v[0]; mol[F_VY * N + i] = v[1]; mol[F_VZ * N + i] = v[2];
    mol[F_P0 * N + i] = p[0]; mol[F_P1 * N + i] = p[1]; mol[F_P2 * N + i] = p[2]; mol[F_P3 * N + i] = p[3];
  }
  D5_scalars<EM>(J, c);
  D6<EM>(J, c);
  __syncthreads();
}

template <int NT, int NS, typename PT>
__device__ void md_step(JRef J, double *mol, int *mid, int N, Sc &c, const Lds &S,
                        bool do_sort, const DescD &R) {
  PT_DECL
  first_half<NT>(J, mol, N, c, S.red);
  PT_MARK(0, gen(R.sc) + S_DIAG0);
  // force
  if (do_sort) sort_and_window<NT>(J, mol, mid, N, J.sort_axis == 0 ? c.L[0] : c.L[2], J.margin, S, gen(R.win));
  PT_MARK(1, gen(R.sc) + S_DIAG0);
  force_pass<NT, NS, PT>(J, mol, N, c, S, gen(R.site_f), gen(R.sc) + S_DIAG0);  // ends with block_sum (synced)
  PT_MARK(2, gen(R.sc) + S_DIAG0);
  second_half<NT>(J, mol, N, c);
  PT_MARK(3, gen(R.sc) + S_DIAG0);
}

// --------------------------------------------------------- replica driver
template <int NT, int NS, typename PT>
__device__ void run_replica(const LaunchArgs &a, const DescD &R, JRef J, char *smem) {
  const int N = J.nmol;
  double *mol = gen(R.mol);
  int *mid = gen(R.mid);
  Lds S = lds_carve<NT>(smem, N, NS, J.precision, gen(R.grec));
  Sc c;
  load_sc(c, gen(R.sc));

  if (a.mode == MODE_INIT) {
    // InitializeProperty (src/molecules.h:260-270): sort with margin 0,
    // force, gkT, CalcProperties, H0 = tot.
    sort_and_window<NT>(J, mol, mid, N, J.sort_axis == 0 ? c.L[0] : c.L[2], 0.0, S, gen(R.win));
    force_pass<NT, NS, PT>(J, mol, N, c, S, gen(R.site_f), gen(R.sc) + S_DIAG0);
    c.gkT = 6.0 * (double)N * c.T;
    const double tot = total_energy<NT>(J, mol, N, c, S.red);
    c.H0 = tot;
    calc_properties<NT>(J, mol, N, c, gen(R.sc), S.red);
    if (threadIdx.x == 0) store_sc(c, gen(R.sc));
    return;
  }
  if (a.mode == MODE_FORCE) {
    if (a.do_sort) sort_and_window<NT>(J, mol, mid, N, J.sort_axis == 0 ? c.L[0] : c.L[2], a.margin, S, gen(R.win));
    else {
      // reuse the stored window
      for (int i = threadIdx.x; i < N; i += NT) {
        int js = gen(R.win) ? gen(R.win)[3 * i] : 0;
        int je = gen(R.win) ? gen(R.win)[3 * i + 1] : N - 1;
        const bool full = gen(R.win) ? gen(R.win)[3 * i + 2] != 0 : true;
        if (js >= je) je += N;
        if (full || !((js <= i && i <= je) || (js <= i + N && i + N <= je))) {
          js = i + 1;
          je = i + N - 1;
        }
        S.js[i] = (uint16_t)js;
        S.je[i] = (uint16_t)je;
      }
      __syncthreads();
    }
    force_pass<NT, NS, PT>(J, mol, N, c, S, gen(R.site_f), gen(R.sc) + S_DIAG0);
    calc_properties<NT>(J, mol, N, c, gen(R.sc), S.red);
    if (threadIdx.x == 0) store_sc(c, gen(R.sc));
    return;
  }

  // MODE_INTERVAL
  const double xflag = gen(R.sc)[S_XFLAG];
  if (xflag != 0.0) {
    // REMD::ReplicaExchange for an exchanged replica
    // (src/remd.cxx:526-536): ChangeTemperature, KillMomentum, SetPressure,
    // CalcProperties, UpdateHamiltonian.
    const double Tn = gen(R.sc)[S_TNEW];
    const double coef = sqrt(Tn / c.T);
    for (int i = threadIdx.x; i < N; i += NT) {
      for (int d = 0; d < 3; d++) mol[(F_VX + d) * N + i] *= coef / c.s;
      for (int d = 0; d < 4; d++) mol[(F_P0 + d) * N + i] *= coef / c.s;
    }
    c.Ps *= coef;
    if (J.pmode == 1) c.Pv[0] *= coef * c.s;
    if (J.pmode == 2) c.Pv[2] *= coef * c.s;
    if (J.pmode == 3) c.Pv[0] *= coef * c.s;
    c.T = Tn;
    c.gkT = 6.0 * (double)N * Tn;
    c.s = 1.0;
    __syncthreads();
    kill_momentum<NT>(J, mol, N, c, S.red, false);
    c.P = gen(R.sc)[S_PNEW];
    c.H0 = total_energy<NT>(J, mol, N, c, S.red);
  }
  // the driver scales once per interval before ExecuteSteps
  // (src/remd.cxx:259-262), which scales again at every step
  if (J.tmode == 2) velocity_scaling<NT>(J, mol, N, c, S.red);
  for (int st = 0; st < J.interval; ++st) {
    if (J.tmode == 2) velocity_scaling<NT>(J, mol, N, c, S.red);
    const bool do_sort = (st % J.sort_interval) == 0;
    md_step<NT, NS, PT>(J, mol, mid, N, c, S, do_sort, R);
  }
  calc_properties<NT>(J, mol, N, c, gen(R.sc), S.red);
  if (threadIdx.x == 0) {
    store_sc(c, gen(R.sc));
    gen(R.sc)[S_XFLAG] = 0.0;
    gen(R.sc)[S_STEPS] += (double)J.interval;
  }
}

template <int NT, int NS, typename PT>
__global__ __launch_bounds__(NT) void interval_kernel(LaunchArgs a) {
  extern __shared__ __align__(16) char smem[];
  if ((int)blockIdx.x >= a.nrep) return;
  const DescD R = ((const DescD *)a.reps)[blockIdx.x];
  JRef J = *R.job;
  run_replica<NT, NS, PT>(a, R, J, smem);
}

// ====================================================== replica-resident path
// For replicas whose state fits in LDS.  The replica's molecule arrays are
// loaded into LDS at the start of a launch and written back at its end.  A
// replica may be split over G workgroups ("parts"): every part integrates
// all molecules (the same instructions on the same data, so the copies stay
// bit-identical) and computes the forces of a contiguous molecule range; the
// parts exchange per-molecule forces, torques, energies and virials through
// global memory once per step behind a group barrier.  Each molecule's force
// is computed whole by one part with the job's fixed chunk count, and the
// energies are reduced from per-molecule values in a fixed order, so results
// do not depend on G, on batching or on placement.

constexpr unsigned MAXSPIN = 1u << 20; // group-barrier polls before the launch gives up

// Per-item partials of the pair pass: COM force (3), quaternion torque (4),
// virial (3), lj, clmb, wall energy -- plus the site forces (3 NS) when the
// site-force debug buffer is on.
constexpr int NPART = 13;
__host__ __device__ inline int npart(int nsite, bool site_dbg) { return NPART + (site_dbg ? 3 * nsite : 0); }

struct LdsR {
  double *mol;      // NFIELD x N replica state
  int *mid;         // N molecule identities
  char *region0;    // site records, or the sort scratch
  uint16_t *js, *je;
  double *part;     // NV x NT pair partials, value-major
  int part_n;       // doubles in the partial area
  double *E;        // energy-block sums lj, clmb, wall, vir[3] (6 x nblk)
  double *red;      // block_sum scratch
  int *flag;        // barrier status broadcast (16 ints)
  double *ptimer;   // 16 phase-timer accumulators (REMD_PHASE_TIMERS builds)
};

__host__ __device__ inline int resident_region0(int nmol, int nsite, int prec) {
  const int stride = (nsite == 5) ? 18 : (nsite == 4 ? 16 : 12);
  const int rec = nmol * stride * (prec ? 8 : 4);
  const int sortb = align16(10 * nmol) + 8 * nmol;  // keys, ranks, one staged field
  return align16(rec > sortb ? rec : sortb);
}

// Bytes of dynamic LDS the resident path needs for one replica (gm: the
// split-global variant, whose state lives in global memory).
__host__ __device__ inline int resident_lds_bytes(int nmol, int nsite, int prec, int nt, bool site_dbg,
                                                  bool gm = false) {
  return (gm ? 0 : align16(NFIELD * 8 * nmol) + align16(4 * nmol)) + resident_region0(nmol, nsite, prec) +
         align16(4 * nmol) + npart(nsite, site_dbg) * 8 * nt +
         (gm ? 0 : align16(6 * 8 * ((nmol + EB - 1) / EB))) + (nt / WAVE) * 32 * 8 + 64 + 128;
}

template <int NT, int NS>
__device__ inline LdsR carve_resident(char *smem, int N, int prec, bool site_dbg, double *gmol = nullptr,
                                      int *gmid = nullptr) {
  LdsR L;
  char *p = smem;
  if (gmol) {
    L.mol = gmol;
    L.mid = gmid;
  } else {
    L.mol = (double *)p;
    p += align16(NFIELD * 8 * N);
    L.mid = (int *)p;
    p += align16(4 * N);
  }
  L.region0 = p;
  p += resident_region0(N, NS, prec);
  L.js = (uint16_t *)p;
  L.je = L.js + N;
  p += align16(4 * N);
  L.part = (double *)p;
  L.part_n = npart(NS, site_dbg) * NT;
  p += npart(NS, site_dbg) * 8 * NT;
  if (gmol) {
    L.E = gmol + (size_t)NFIELD * N;  // block sums in the private global copy
  } else {
    L.E = (double *)p;
    p += align16(6 * 8 * ((N + EB - 1) / EB));
  }
  L.red = (double *)p;
  p += (NT / WAVE) * 32 * 8;
  L.flag = (int *)p;
  p += 64;
  L.ptimer = (double *)p;
  return L;
}

struct Group {
  int parts, part;
  unsigned long long *ctr;  // low 32 bits: arrivals of this replica's parts; high 32 bits: 4-bit
                            // count of parts per XCD (added at the first barrier); zero at start
  unsigned *abort;  // nonzero once any wait of the launch timed out
  double *xbuf;     // 2 x xbuf_parity_doubles(N) exchange rows (double-buffered by step parity)
  unsigned epoch;   // group barriers completed in this launch
  bool same_xcd;    // every part of the replica shares one L2
  int mode;         // same-XCD barrier variant (LaunchArgs::same_xcd_mode)
};

// XCD (XCC) this workgroup runs on.
__device__ inline unsigned xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u; }

// Barrier across the parts of one replica.  Every thread first makes its
// own global stores visible to the other parts; one thread counts in and
// polls the counter (bounded; a timeout raises the launch-wide abort flag so
// every part leaves); then the other parts' stores are made visible to this
// part's loads.  When all parts share one XCD (and so one L2, the coherence
// point of its CUs) the stores are only drained to L2 and the exchange rows
// are read with device-scope loads that bypass the CU's L1 (an L1-only
// invalidate was measured not to suffice); otherwise agent-scope release and
// acquire fences (L2 write-back / invalidate) and ordinary loads.  Returns
// false when the launch is aborting.
// `census` (the launch's first barrier): each part also adds 1 to the
// 4-bit field of its XCD in the high half of the group word, and flag[1]
// receives whether all parts share one XCD -- read from the same word that
// completes the barrier, so every part sees every part's field and all
// parts decide alike.
template <int NT>
__device__ bool group_wait(Group &g, int *flag, double *pt_sc = nullptr, bool census = false) {
  PT_DECL
  (void)pt_sc;
  if (g.same_xcd && !(g.mode & 1)) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  PT_MARK(8, pt_sc);
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned long long target = (unsigned long long)g.parts * (g.epoch + 1u);
    const unsigned long long inc = 1ull + (census ? (1ull << (32 + 4 * xcc_id())) : 0ull);
    __hip_atomic_fetch_add(g.ctr, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    unsigned long long v;
    while (((v = __hip_atomic_load(g.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & 0xffffffffull) <
           target) {
      ++spins;
      if ((spins & 63u) == 0u) {
        if (__hip_atomic_load(g.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
          ok = 0;
          break;
        }
        if (spins > MAXSPIN) {
          __hip_atomic_store(g.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
      }
      __builtin_amdgcn_s_sleep(2);
    }
    *flag = ok;
    if (census) {
      const unsigned hist = (unsigned)(v >> 32);
      int nz = 0;
      for (int x = 0; x < 8; x++) nz += ((hist >> (4 * x)) & 15u) != 0u;
      flag[1] = (nz == 1) ? 1 : 0;
    }
  }
  __syncthreads();
  PT_MARK(9, pt_sc);
  const bool ok = *flag != 0;
  if (!g.same_xcd || (g.mode & 2)) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  g.epoch++;
  // no barrier needed before flag is rewritten: thread 0 writes it again
  // only after the workgroup barrier that opens the next call, which every
  // thread reaches after its read above
  PT_MARK(10, pt_sc);
  return ok;
}

// Sorted position rk[i] of each key ax[i] (stable: equal keys keep their
// order) by the wave-ballot radix sort of the keys' order-preserving bit
// patterns, in `region` (20 N bytes + the digit counts): keys A/B at 0 and
// 8N, indices A/B at 16N and 18N, counts at 20N; 16 passes leave the result
// in A.  rk (N 16-bit entries) must lie outside [0, 8N) and [16N, ...).
// Out of line: inlined into the resident kernel it perturbs the register
// allocation of the whole step loop.
template <int NT>
__device__ __attribute__((noinline)) void radix_ranks(const double *ax, int N, char *region, uint16_t *rk) {
  uint64_t *keyA = (uint64_t *)region, *keyB = keyA + N;
  uint16_t *idxA = (uint16_t *)(region + 16 * N), *idxB = idxA + N;
  int *cnt = (int *)(region + 20 * N);
  for (int i = threadIdx.x; i < N; i += NT) {
    const uint64_t u = (uint64_t)__double_as_longlong(ax[i] + 0.0);  // -0.0 -> +0.0
    keyA[i] = (u >> 63) ? ~u : (u | (1ull << 63));
    idxA[i] = (uint16_t)i;
  }
  __syncthreads();
  radix_sort<NT>(keyA, idxA, keyB, idxB, cnt, N);
  for (int k = threadIdx.x; k < N; k += NT) rk[idxA[k]] = (uint16_t)k;
}

// Stable rank sort of the sort-axis coordinate in LDS (the same order as a
// stable sort on the double keys: rank = #smaller + #equal-and-earlier),
// permutation of the state, then the windows by binary search on the sorted
// keys (src/molecules.cxx:541-585 semantics; raw js/je and the whole-ring
// flag go to win_dbg when given).
template <int NT, int NS, int EM = ENS_RUNTIME>
__device__ void rank_sort_resident(JRef J, double *mol, int *mid, int N, double Laxis,
                                   double margin, const LdsR &S, int *win_dbg) {
  double *z = (double *)S.region0;
  uint16_t *rk = (uint16_t *)(S.region0 + 8 * N);
  double *ax = mol + (F_RX + sort_axis_of<EM>(J)) * N;
  // large replicas: the wave-ballot radix sort of the keys' order-preserving
  // bit patterns (stable, so the same order as the rank sort), in region 0:
  // keys A/B at 0 and 8N bytes, indices A/B at 16N and 18N, digit counts
  // after; 16 passes leave the result in A.  rk then goes to [8N, 10N).
  const int ng = ((N + NT - 1) / NT) * (NT / WAVE);
  const bool radix = N >= J.radix_min && 20 * N + 4 * 17 * ng <= resident_region0(N, NS, J.precision);
  if (radix) {
    radix_ranks<NT>(ax, N, S.region0, rk);
  } else {
    // rank = #smaller + #equal-and-earlier, the N x N comparisons split into
    // (molecule, segment of partners) items so that every thread takes part:
    // the segment count minimises the longest thread's comparisons; the
    // segments' counts meet in an integer per molecule (order-free)
    int *cnt = (int *)(S.region0 + align16(10 * N));
    int nseg = 1, best = 1 << 30;
    for (int g = 1; g <= 8; g++) {
      const int len = (N + g - 1) / g;
      const int cost = ((N * g + NT - 1) / NT) * len + 8 * g;
      if (cost < best) {
        best = cost;
        nseg = g;
      }
    }
    const int len = (N + nseg - 1) / nseg;
    for (int i = threadIdx.x; i < N; i += NT) {
      z[i] = ax[i] + 0.0;  // -0.0 -> +0.0
      cnt[i] = 0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < N * nseg; t += NT) {
      const int i = t % N, seg = t / N;
      const double zi = z[i];
      const int j0 = seg * len, j1 = min(N, j0 + len);
      int r = 0;
#pragma unroll 8
      for (int j = j0; j < j1; ++j) {
        const double zj = z[j];
        r += (j < i) ? (zj <= zi ? 1 : 0) : (zj < zi ? 1 : 0);
      }
      if (nseg == 1) cnt[i] = r;
      else atomicAdd(&cnt[i], r);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += NT) rk[i] = (uint16_t)cnt[i];
  }
  __syncthreads();
  // permute the state through the larger of the pair-partial area (idle
  // during the sort) and the rest of region 0, as many fields per round as
  // fit.  F and n are not permuted: the force pass that follows every sort
  // rewrites them for every molecule before anything reads them.  Field k of
  // the NPF permuted ones: r, v (k < 6), then q, p.
  constexpr int NPF = NFIELD - 7;
  const int cap_part = S.part_n / N;
  const int cap_r0 = (resident_region0(N, NS, J.precision) - align16(10 * N)) / (8 * N);
  double *tmp = cap_part >= cap_r0 ? S.part : (double *)(S.region0 + align16(10 * N));
  const int fit = cap_part >= cap_r0 ? cap_part : cap_r0;
  const int per = fit < 1 ? 1 : (fit > NPF ? NPF : fit);
  for (int f0 = 0; f0 < NPF; f0 += per) {
    const int nf = (f0 + per <= NPF) ? per : NPF - f0;
    for (int i = threadIdx.x; i < N; i += NT) {
      const int ri = rk[i];
      for (int f = 0; f < nf; ++f) {
        const int fld = (f0 + f < 6) ? f0 + f : f0 + f + (F_Q0 - 6);
        tmp[f * N + ri] = mol[fld * N + i];
      }
    }
    __syncthreads();
    for (int f = 0; f < nf; ++f) {
      const int fld = (f0 + f < 6) ? f0 + f : f0 + f + (F_Q0 - 6);
      for (int k = threadIdx.x; k < N; k += NT) mol[fld * N + k] = tmp[f * N + k];
    }
    __syncthreads();
  }
  {
    int *ti = (int *)tmp;
    for (int i = threadIdx.x; i < N; i += NT) ti[rk[i]] = mid[i];
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += NT) mid[i] = ti[i];
    __syncthreads();
  }
  const double w = (J.rcut + margin) / Laxis;
  for (int i = threadIdx.x; i < N; i += NT) {
    int js, je;
    window_of(ax, N, i, w, js, je);
    const bool full = window_full(ax, i, w, js, je);
    if (win_dbg) {
      win_dbg[3 * i] = js;
      win_dbg[3 * i + 1] = je;
      win_dbg[3 * i + 2] = full ? 1 : 0;
    }
    if (js >= je) je += N;
    if (full || !((js <= i && i <= je) || (js <= i + N && i + N <= je))) {
      js = i + 1;
      je = i + N - 1;
    }
    S.js[i] = (uint16_t)js;
    S.je[i] = (uint16_t)je;
  }
  __syncthreads();
}

// Window of the last sort, from the debug copy (MODE_FORCE without a sort).
template <int NT>
__device__ void window_from_debug(const int *win, int N, const LdsR &S) {
  for (int i = threadIdx.x; i < N; i += NT) {
    int js = win ? win[3 * i] : 0;
    int je = win ? win[3 * i + 1] : N - 1;
    const bool full = win ? win[3 * i + 2] != 0 : true;
    if (js >= je) je += N;
    if (full || !((js <= i && i <= je) || (js <= i + N && i + N <= je))) {
      js = i + 1;
      je = i + N - 1;
    }
    S.js[i] = (uint16_t)js;
    S.je[i] = (uint16_t)je;
  }
  __syncthreads();
}

// Pair pass for molecules [i0, i1) of the replica (switched LJ + Coulomb,
// src/force_calculator.cxx:803-989 semantics; wall src/force_calculator.cxx:
// 1494-1531; A2M torque src/molecules.cxx:401-439).  Work item (i, c): chunk
// c of C runs over i's window with stride C; the C partials of i are summed
// in chunk order by one thread.  Outputs per molecule: F[3], n[4] into the
// LDS state (and into the global exchange row xrow when given); per energy
// block of EB molecules: the lj, clmb, wall, vir[3] sums into Bk (6 x nblk,
// value-major) and into the global block row xblk when given.  i0 is a
// block boundary.
template <int NT, int NS, typename PT, int EM = ENS_RUNTIME>
__device__ void force_range(JRef J, double *mol, int N, const Sc &c, const LdsR &S, int i0,
                            int i1, double *xrow, double *Bk, int nblk, double *xblk, double *site_dbg,
                            double *pt_sc = nullptr) {
  PT_DECL
  (void)pt_sc;
  constexpr int QB = (NS == 3) ? 0 : 1;
  constexpr int ST = Rec<NS>::STRIDE;
  const PT *rec = (const PT *)S.region0;
  PairK<NS, PT> k;
  if (sizeof(PT) == 4) {
    k.rc = J.f_rc;
    k.rl = J.f_rl;
    k.rc2 = J.f_rc2;
    k.rl2 = J.f_rl2;
    k.coef = J.f_coef;
    k.coef30 = J.f_coef30;
    k.sig2 = J.f_sig2;
    k.eps4 = J.f_eps4;
    k.eps48 = J.f_eps48;
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int b = 0; b < NS; b++) k.qq[a][b] = J.f_qq[a][b];
  } else {
    k.rc = (PT)J.rcut;
    k.rl = (PT)J.rswitch;
    k.rc2 = (PT)(J.rcut * J.rcut);
    k.rl2 = (PT)(J.rswitch * J.rswitch);
    const double rlc = J.rswitch - J.rcut;
    const double coef = 1.0 / (rlc * rlc * rlc * rlc * rlc);
    k.coef = (PT)coef;
    k.coef30 = (PT)(30.0 * coef);
    k.sig2 = (PT)J.lj_sig2;
    k.eps4 = (PT)J.lj_eps4;
    k.eps48 = (PT)(12.0 * J.lj_eps4);
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int b = 0; b < NS; b++) k.qq[a][b] = (PT)(J.q[a] * J.q[b]);
  }
  k.rs2 = k.rc2 * (PT)1.0001;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    k.L[d] = J.per[d] ? (PT)c.L[d] : PT(0);
    k.Li[d] = (PT)(1.0 / c.L[d]);
  }
  const int C = J.chunks;
  const int B = (NT / C) / EB * EB;  // molecules per round, whole energy blocks
  const int NV = npart(NS, site_dbg != nullptr);
  const double dr = (cmode_of<EM>(J) == 1) ? J.wall_length / (double)J.nfwall : 0.5 * J.wall_length / (double)J.nfwall;
  const double dri = 1.0 / dr;
  const int t = threadIdx.x;
  for (int ib = i0; ib < i1; ib += B) {
    const int ie = (ib + B < i1) ? ib + B : i1;
    const int items = (ie - ib) * C;
    if (t < items) {
      const int li = t / C;
      const int i = ib + li;
      const int cc = t - li * C;
      Acc<NS> acc;
#pragma unroll
      for (int a = 0; a < NS; a++) acc.site[a][0] = acc.site[a][1] = acc.site[a][2] = 0.0;
      acc.corr[0] = acc.corr[1] = acc.corr[2] = 0.0;
      acc.vir[0] = acc.vir[1] = acc.vir[2] = 0.0;
      acc.lj = acc.clmb = 0.0;
      const double q[4] = {mol[F_Q0 * N + i], mol[F_Q1 * N + i], mol[F_Q2 * N + i], mol[F_Q3 * N + i]};
      // wall on the site-0 position (chunk 0 of each molecule), FP64: the
      // table rows are fetched before the pair loop so their latency hides
      // behind it (src/force_calculator.cxx:1494-1531)
      const bool wall = cc == 0 && (cmode_of<EM>(J) == 1 || cmode_of<EM>(J) == 2);
      double wx = 0.0, wy = 0.0, wz = 0.0, wu = 0.0, f0 = 0.0, f1 = 0.0, e0 = 0.0, e1 = 0.0;
      int widx = 0;
      if (wall) {
        double b0[3];
        body_to_space(q, (const double *)J.site_r[0], b0);
        if (cmode_of<EM>(J) == 1) {
          wx = (mol[F_RX * N + i] - 0.5) * c.L[0] + b0[0];
          wy = (mol[F_RY * N + i] - 0.5) * c.L[1] + b0[1];
          wu = sqrt(wx * wx + wy * wy);
        } else {
          wz = (mol[F_RZ * N + i] - 0.5) * c.L[2] + b0[2];
          wu = fabs(wz);
        }
        widx = (int)(wu / dr);
        if (widx > J.nfwall - 2) widx = J.nfwall - 2;
        f0 = J.fwall[2 * widx];
        e0 = J.fwall[2 * widx + 1];
        f1 = J.fwall[2 * (widx + 1)];
        e1 = J.fwall[2 * (widx + 1) + 1];
      }
      const PT *ri = rec + i * ST;
      PT gi[3] = {ri[0], ri[1], ri[2]};
      PT ai[NS][3];
#pragma unroll
      for (int a = 0; a < NS; a++) {
        ai[a][0] = ri[3 + 3 * a];
        ai[a][1] = ri[4 + 3 * a];
        ai[a][2] = ri[5 + 3 * a];
      }
      pair_window<NS, QB, PT, per_mask_of<EM>()>(rec, N, i, S.js[i], S.je[i], cc, C, J.compact != 0, gi, ai, k,
                                              acc);
      double e_wall = 0.0;
      if (wall) {
        const double f = (f1 - f0) * dri * (wu - dr * widx) + f0;
        e_wall = (e1 - e0) * dri * (wu - dr * widx) + e0;
        if (cmode_of<EM>(J) == 1) {
          acc.site[0][0] += f * wx;
          acc.site[0][1] += f * wy;
        } else {
          acc.site[0][2] += f * wz;
        }
      }
      // this item's COM force and A2M torque (src/molecules.cxx:401-439;
      // linear in the site forces, so chunks can be summed afterwards)
      double F[3], n4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double sum = acc.corr[d];
#pragma unroll
        for (int a = 0; a < NS; a++) sum += acc.site[a][d];
        F[d] = sum;
      }
#pragma unroll
      for (int a = 0; a < NS; a++) {
        double fb[3];
        space_to_body(q, acc.site[a], fb);
        const double *br = (const double *)J.site_r[a];
        n4[0] += br[0] * fb[0] + br[1] * fb[1] + br[2] * fb[2];
        n4[1] += br[1] * fb[2] - br[2] * fb[1];
        n4[2] += br[2] * fb[0] - br[0] * fb[2];
        n4[3] += br[0] * fb[1] - br[1] * fb[0];
      }
      double *P = S.part;
      P[0 * NT + t] = F[0];
      P[1 * NT + t] = F[1];
      P[2 * NT + t] = F[2];
      P[3 * NT + t] = n4[0];
      P[4 * NT + t] = n4[1];
      P[5 * NT + t] = n4[2];
      P[6 * NT + t] = n4[3];
      P[7 * NT + t] = acc.vir[0];
      P[8 * NT + t] = acc.vir[1];
      P[9 * NT + t] = acc.vir[2];
      P[10 * NT + t] = acc.lj;
      P[11 * NT + t] = acc.clmb;
      P[12 * NT + t] = e_wall;
      if (site_dbg) {
#pragma unroll
        for (int a = 0; a < NS; a++)
#pragma unroll
          for (int d = 0; d < 3; d++) P[(NPART + 3 * a + d) * NT + t] = acc.site[a][d];
      }
    }
    __syncthreads();
    PT_MARK(12, pt_sc);
    // Chunk partials -> per-molecule values, one (value group, molecule)
    // per thread: group 0 the COM force (partials 0-2), 1 the quaternion
    // torque (3-6), 2 the virial (7-9), 3 lj, clmb, wall (10-12).  Each
    // value is its chunks added in chunk order.  A group spans M4 lanes
    // (whole lane quads), so the EB = 4 molecules of an energy block sit on
    // one lane quad (rounds start on block boundaries).
    const int M = ie - ib;
    const int M4 = (M + EB - 1) / EB * EB;
    for (int w = t; w < 4 * M4; w += NT) {
      const int grp = w / M4;
      const int li = w - grp * M4;
      const bool mine = li < M;
      const int u0 = grp == 0 ? 0 : (grp == 1 ? 3 : (grp == 2 ? 7 : 10));
      const int nu = grp == 1 ? 4 : 3;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      if (mine) {
        const double *P = S.part + (size_t)u0 * NT + li * C;
        // chunk-outer so each chunk's loads issue together
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (u < nu) v[u] = P[u * NT];
        for (int cc = 1; cc < C; cc++) {
          double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (u < nu) x[u] = P[u * NT + cc];
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (u < nu) v[u] += x[u];
        }
        // F, n into the state (and the exchange row)
        const int i = ib + li;
        if (grp == 0) {
          mol[F_FX * N + i] = v[0];
          mol[F_FY * N + i] = v[1];
          mol[F_FZ * N + i] = v[2];
          if (xrow) {
            double *o = xrow + (size_t)i * XR;
            o[0] = v[0];
            o[1] = v[1];
            o[2] = v[2];
          }
        } else if (grp == 1) {
          mol[F_N0 * N + i] = v[0];
          mol[F_N1 * N + i] = v[1];
          mol[F_N2 * N + i] = v[2];
          mol[F_N3 * N + i] = v[3];
          if (xrow) {
            double *o = xrow + (size_t)i * XR;
            o[3] = v[0];
            o[4] = v[1];
            o[5] = v[2];
            o[6] = v[3];
          }
        }
      }
      // energy-block sums (groups 2 and 3): lane 4k adds its quad's values
      // in molecule order (absent molecules are zeros); quads are whole, so
      // every lane a quad_perm reads is active
      double bs[3];
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const double x1 = dpp_mov<0x55>(v[u]);  // quad_perm [1, 1, 1, 1]
        const double x2 = dpp_mov<0xAA>(v[u]);  // quad_perm [2, 2, 2, 2]
        const double x3 = dpp_mov<0xFF>(v[u]);  // quad_perm [3, 3, 3, 3]
        bs[u] = ((v[u] + x1) + x2) + x3;
      }
      if (grp >= 2 && mine && (li & 3) == 0) {
        const int b = (ib + li) / EB;
        const int r0 = grp == 3 ? 0 : 3;  // block rows: lj, clmb, wall, vir0, vir1, vir2
#pragma unroll
        for (int u = 0; u < 3; u++) Bk[(r0 + u) * nblk + b] = bs[u];
        if (xblk) {
#pragma unroll
          for (int u = 0; u < 3; u++) xblk[(size_t)b * XB + r0 + u] = bs[u];
        }
      }
    }
    if (site_dbg) {
      for (int li = t; li < M; li += NT) {
        const double *P = S.part + li * C;
        for (int u = NPART; u < NV; u++) {
          double sum = P[u * NT];
          for (int cc = 1; cc < C; cc++) sum += P[u * NT + cc];
          site_dbg[(ib + li) * NS * 3 + (u - NPART)] = sum;
        }
      }
    }
    __syncthreads();
    PT_MARK(13, pt_sc);
  }
}

// Exchange row of molecule k (COM force, quaternion torque) written by
// another part.  `direct`: device-scope loads that bypass the CU's L1 (the
// same-XCD barrier drains stores only to L2).
__device__ inline void xrow_load(const double *o, bool direct, double (&v)[7]) {
  if (direct) {
#pragma unroll
    for (int u = 0; u < 7; u++) v[u] = __hip_atomic_load(o + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
#pragma unroll
    for (int u = 0; u < 7; u++) v[u] = o[u];
  }
}

__device__ inline void xrow_store(double *mol, int N, int k, const double (&v)[7]) {
  mol[F_FX * N + k] = v[0];
  mol[F_FY * N + k] = v[1];
  mol[F_FZ * N + k] = v[2];
  mol[F_N0 * N + k] = v[3];
  mol[F_N1 * N + k] = v[4];
  mol[F_N2 * N + k] = v[5];
  mol[F_N3 * N + k] = v[6];
}

// Totals of the six block sums (lj, clmb, wall, virial diagonal) over the
// replica's nblk energy blocks, computed by every wave for itself (no LDS
// round trip, no workgroup barrier).  The order is a fixed tree of NT
// slots: slot s adds blocks s, s + NT, ... in order, the slots of each
// 64-slot group are added by wave_sum, and the groups are added in group
// order after a 0.0 -- the order of block_sum<NT> over slots, whatever the
// parts.  Block b comes from this part's LDS rows (Bk, blocks [b0, b1),
// or all of them when xblk is null) or from the exchange area.
template <int NT>
__device__ inline void block_totals(const double *Bk, const double *xblk, int nblk, int b0, int b1,
                                    bool direct, double (&tot)[6]) {
  constexpr int NW = NT / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int u = 0; u < 6; u++) tot[u] = 0.0;
  for (int ww = 0; ww < NW; ww++) {
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (ww * WAVE < nblk) {  // wave-uniform
      for (int b = ww * WAVE + lane; b < nblk; b += NT) {
        double e[6];
        if (xblk == nullptr || (b >= b0 && b < b1)) {
#pragma unroll
          for (int u = 0; u < 6; u++) e[u] = Bk[u * nblk + b];
        } else if (direct) {
#pragma unroll
          for (int u = 0; u < 6; u++)
            e[u] = __hip_atomic_load(xblk + (size_t)b * XB + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
#pragma unroll
          for (int u = 0; u < 6; u++) e[u] = xblk[(size_t)b * XB + u];
        }
#pragma unroll
        for (int u = 0; u < 6; u++) p[u] += e[u];
      }
#pragma unroll
      for (int u = 0; u < 6; u++) p[u] = wave_sum(p[u]);
    }
#pragma unroll
    for (int u = 0; u < 6; u++) tot[u] += p[u];
  }
}

// Forces of the whole replica into the LDS state and the energy/virial
// totals into c: this part's molecules [i0, i1) (whole energy blocks), the
// exchange with the other parts -- F and n per molecule, the six energy and
// virial sums per block of EB molecules -- then the fixed-order reduction of
// the block sums.  A block is summed in molecule order by one thread and
// the blocks by a fixed tree, so the totals do not depend on the parts.
template <int NT, int NS, typename PT, int EM = ENS_RUNTIME>
__device__ bool forces_resident(JRef J, double *mol, int N, Sc &c, const LdsR &S, Group &g,
                                int i0, int i1, double *site_dbg, double *pt_sc = nullptr) {
  PT_DECL
  (void)pt_sc;
  m2a<NT, NS, PT>(J, mol, N, c.L[0], c.L[1], c.L[2], (PT *)S.region0);
  __syncthreads();
  PT_MARK(4, pt_sc);
  const int nblk = (N + EB - 1) / EB;
  const int b0 = i0 / EB, b1 = (i1 + EB - 1) / EB;
  double *xrow = g.parts > 1 ? g.xbuf + (size_t)(g.epoch & 1u) * xbuf_parity_doubles(N) : nullptr;
  double *xblk = xrow ? xrow + (size_t)N * XR : nullptr;
  double *Bk = S.E;  // 6 x nblk block sums
  force_range<NT, NS, PT, EM>(J, mol, N, c, S, i0, i1, xrow, Bk, nblk, xblk, site_dbg, pt_sc);
  PT_MARK(5, pt_sc);
  const bool split = g.parts > 1;
  if (split) {
    if (!group_wait<NT>(g, S.flag, pt_sc)) return false;
    PT_MARK(6, pt_sc);
  }
  const bool direct = split && g.same_xcd && !(g.mode & 2);
  // the other parts' rows: the first row of each thread is loaded before
  // the block totals' loads so the two latencies overlap, and written to
  // the state after them
  int k = threadIdx.x;
  const bool row0 = split && k < N && !(k >= i0 && k < i1);
  double rv[7];
  if (row0) xrow_load(xrow + (size_t)k * XR, direct, rv);
  double tot[6];
  block_totals<NT>(Bk, split ? xblk : nullptr, nblk, b0, b1, direct, tot);
  if (row0) xrow_store(mol, N, k, rv);
  if (split) {
    for (k += NT; k < N; k += NT) {
      if (k >= i0 && k < i1) continue;
      xrow_load(xrow + (size_t)k * XR, direct, rv);
      xrow_store(mol, N, k, rv);
    }
  }
  PT_MARK(11, pt_sc);
  __syncthreads();
  c.lj = 0.5 * tot[0];
  c.clmb = 0.5 * tot[1];
  c.wall = tot[2];
  c.vir[0] = 0.5 * tot[3];
  c.vir[1] = 0.5 * tot[4];
  c.vir[2] = 0.5 * tot[5];
  c.pot = c.lj + c.clmb + c.wall;
  PT_MARK(7, pt_sc);
  return true;
}

// GM (split-global variant): the replica's state is a private copy per
// part in global memory (R.gbuf) instead of LDS; everything else as above.
template <int NT, int NS, typename PT, bool GM = false, int EM = ENS_RUNTIME>
__device__ __attribute__((always_inline)) void run_resident(const LaunchArgs &a, const DescD &R, JRef J, char *smem,
                             Group &g) {
  const int N = J.nmol;
  double *gmol = nullptr;
  int *gmid = nullptr;
  if (GM) {
    gmol = gen(R.gbuf) + (size_t)g.part * gm_part_doubles(N);
    gmid = (int *)(gmol + (size_t)NFIELD * N + (size_t)6 * ((N + EB - 1) / EB));
  }
  LdsR S = carve_resident<NT, NS>(smem, N, J.precision, R.site_f != nullptr, gmol, gmid);
  double *mol = S.mol;
  int *mid = S.mid;
  for (int k = threadIdx.x; k < NFIELD * N; k += NT) mol[k] = gen(R.mol)[k];
  for (int i = threadIdx.x; i < N; i += NT) mid[i] = gen(R.mid)[i];
  Sc c;
  load_sc(c, gen(R.sc));
  const double xflag = gen(R.sc)[S_XFLAG];
  const double Tn = gen(R.sc)[S_TNEW];
  const double Pn = gen(R.sc)[S_PNEW];
  if (threadIdx.x < 16) S.ptimer[threadIdx.x] = 0.0;
  __syncthreads();
  const bool lead = g.part == 0;
  if (g.parts > 1) {
    // one fully fenced barrier first; it also tells whether every part of
    // the replica runs on one XCD
    g.same_xcd = false;
    if (!group_wait<NT>(g, S.flag, nullptr, true)) return;
    g.same_xcd = a.same_xcd_ok && S.flag[1] != 0;
    __syncthreads();
  }
  // this part's molecules: whole energy blocks
  const int nblk = (N + EB - 1) / EB;
  const int i0 = (int)(((long long)g.part * nblk) / g.parts) * EB;
  const int i1e = (int)(((long long)(g.part + 1) * nblk) / g.parts) * EB;
  const int i1 = i1e < N ? i1e : N;
  double *site_dbg = gen(R.site_f);
  int *win_dbg = lead ? gen(R.win) : nullptr;
  double *sc_out = lead ? gen(R.sc) : nullptr;
  const double Laxis0 = sort_axis_of<EM>(J) == 0 ? c.L[0] : c.L[2];

  if (a.mode == MODE_INIT) {
    // InitializeProperty (src/molecules.h:260-270)
    rank_sort_resident<NT, NS, EM>(J, mol, mid, N, Laxis0, 0.0, S, win_dbg);
    if (!forces_resident<NT, NS, PT, EM>(J, mol, N, c, S, g, i0, i1, site_dbg)) return;
    __syncthreads();  // S.red is read by forces_resident's last reduction
    c.gkT = 6.0 * (double)N * c.T;
    c.H0 = total_energy<NT, EM>(J, mol, N, c, S.red);
    calc_properties<NT, EM>(J, mol, N, c, sc_out, S.red);
  } else if (a.mode == MODE_FORCE) {
    if (a.do_sort) rank_sort_resident<NT, NS, EM>(J, mol, mid, N, Laxis0, a.margin, S, win_dbg);
    else window_from_debug<NT>(gen(R.win), N, S);
    if (!forces_resident<NT, NS, PT, EM>(J, mol, N, c, S, g, i0, i1, site_dbg)) return;
    __syncthreads();  // S.red is read by forces_resident's last reduction
    calc_properties<NT, EM>(J, mol, N, c, sc_out, S.red);
  } else {
    if (xflag != 0.0) {
      // REMD::ReplicaExchange for an exchanged replica (src/remd.cxx:526-536)
      const double coef = sqrt(Tn / c.T);
      for (int i = threadIdx.x; i < N; i += NT) {
        for (int d = 0; d < 3; d++) mol[(F_VX + d) * N + i] *= coef / c.s;
        for (int d = 0; d < 4; d++) mol[(F_P0 + d) * N + i] *= coef / c.s;
      }
      c.Ps *= coef;
      if (pmode_of<EM>(J) == 1) c.Pv[0] *= coef * c.s;
      if (pmode_of<EM>(J) == 2) c.Pv[2] *= coef * c.s;
      if (pmode_of<EM>(J) == 3) c.Pv[0] *= coef * c.s;
      c.T = Tn;
      c.gkT = 6.0 * (double)N * Tn;
      c.s = 1.0;
      __syncthreads();
      kill_momentum<NT, EM>(J, mol, N, c, S.red, false);
      c.P = Pn;
      c.H0 = total_energy<NT, EM>(J, mol, N, c, S.red);
    }
    // the driver scales once per interval before ExecuteSteps
    // (src/remd.cxx:259-262), which scales again at every step
    if (tmode_of<EM>(J) == 2) velocity_scaling<NT>(J, mol, N, c, S.red);
    double *pt = lead ? S.ptimer : nullptr;
    (void)pt;
    // the second half of step st - 1 runs fused into the first half of step
    // st (not with per-step velocity scaling, which comes between them)
    const bool fuse = tmode_of<EM>(J) != 2;
    for (int st = 0; st < J.interval; ++st) {
      PT_DECL
      if (tmode_of<EM>(J) == 2) velocity_scaling<NT>(J, mol, N, c, S.red);
      if (fuse && st > 0) first_half<NT, true, EM>(J, mol, N, c, S.red);
      else first_half<NT, false, EM>(J, mol, N, c, S.red);
      PT_MARK(0, pt);
      if ((st % J.sort_interval) == 0)
        rank_sort_resident<NT, NS, EM>(J, mol, mid, N, sort_axis_of<EM>(J) == 0 ? c.L[0] : c.L[2], J.margin, S, win_dbg);
      PT_MARK(1, pt);
      if (!forces_resident<NT, NS, PT, EM>(J, mol, N, c, S, g, i0, i1, site_dbg, pt)) return;
      PT_MARK(2, pt);
      if (!fuse || st == J.interval - 1) second_half<NT, EM>(J, mol, N, c);
      PT_MARK(3, pt);
    }
    calc_properties<NT, EM>(J, mol, N, c, sc_out, S.red);
  }
  // every part has read the replica's global state before the lead rewrites it
  if (g.parts > 1 && !group_wait<NT>(g, S.flag)) return;
  if (lead) {
    for (int k = threadIdx.x; k < NFIELD * N; k += NT) gen(R.mol)[k] = mol[k];
    for (int i = threadIdx.x; i < N; i += NT) gen(R.mid)[i] = mid[i];
    if (threadIdx.x == 0) {
#ifdef REMD_PHASE_TIMERS
      for (int k = 0; k < 14; k++) gen(R.sc)[S_DIAG0 + k] += S.ptimer[k];
      gen(R.sc)[S_DIAG0 + 14] = (double)xcc_id();
      gen(R.sc)[S_DIAG0 + 15] = g.same_xcd ? 1.0 : 0.0;
#endif
      store_sc(c, gen(R.sc));
      if (a.mode == MODE_INTERVAL) {
        gen(R.sc)[S_XFLAG] = 0.0;
        gen(R.sc)[S_STEPS] += (double)J.interval;
      }
    }
  }
}

// grid: parts == 1 -> one workgroup per replica; parts > 1 -> 8 * parts *
// ceil(nrep / 8) workgroups, the parts of a replica on block indices that
// are equal mod 8 (workgroups are dispatched to the 8 XCDs round-robin, so a
// replica's parts share one L2).
template <int NT, int NS, typename PT, int EM = ENS_RUNTIME>
__global__ __launch_bounds__(NT) void resident_kernel(LaunchArgs a) {
  extern __shared__ __align__(16) char smem[];
  int rep, part;
  if (a.parts <= 1) {
    rep = blockIdx.x;
    part = 0;
  } else {
    const int b = blockIdx.x;
    const int slot = b >> 3;
    rep = (slot / a.parts) * 8 + (b & 7);
    part = slot % a.parts;
  }
  if (rep >= a.nrep) return;
  const DescD R = ((const DescD *)a.reps)[rep];
  JRef J = *R.job;
  Group g;
  g.parts = a.parts < 1 ? 1 : a.parts;
  g.part = part;
  g.ctr = a.sync + rep;
  g.abort = a.abort;
  g.same_xcd = false;
  g.mode = a.same_xcd_mode;
  g.xbuf = gen(R.xbuf);
  g.epoch = 0;
  run_resident<NT, NS, PT, false, EM>(a, R, J, smem, g);
}

// One launch for every resident replica of a batch, whatever its water
// model and pair precision.  Workgroup b runs the entry blk[b] = rep << 12 |
// variant << 8 | parts << 4 | part (variant = 2 (nsite - 3) + prec, plus 6
// for the split-global variant whose state is in global memory), or
// nothing for ~0u.  The host lists the parts of every split replica first,
// the parts of one replica on block indices equal mod 8 (one XCD), all of
// them within the workgroups the device holds at once; the single-part
// replicas follow, heaviest first, and queue for workgroup slots as they
// free up.  Workgroups are dispatched in index order, so every split
// replica has all its parts resident before a queued replica starts, and
// no part ever waits on a workgroup that is not yet running.
// The batch kernel calls the twelve variants as functions (one register
// allocation each), and for FP32 jobs of the ENS_*_NP ensembles the same
// variants with the modes fixed; the single-job kernel inlines its one
// variant.
template <int NS, typename PT, bool GM, int EM = ENS_RUNTIME>
__device__ __attribute__((noinline)) void run_resident_call(const LaunchArgs &a, const DescD &R, JRef J, char *smem,
                                                            Group &g) {
  run_resident<RNT, NS, PT, GM, EM>(a, R, J, smem, g);
}

// The FP32 variants of a batch entry with the ensemble modes fixed at
// compile time (EM, one of the ENS_*_NP codes).  False for FP64 entries.
template <int EM>
__device__ inline bool run_resident_fixed(unsigned variant, const LaunchArgs &a, const DescD &R, JRef J, char *smem,
                                          Group &g) {
  switch (variant) {
    case 0: run_resident_call<3, float, false, EM>(a, R, J, smem, g); return true;
    case 6: run_resident_call<3, float, true, EM>(a, R, J, smem, g); return true;
    case 2: run_resident_call<4, float, false, EM>(a, R, J, smem, g); return true;
    case 8: run_resident_call<4, float, true, EM>(a, R, J, smem, g); return true;
    case 4: run_resident_call<5, float, false, EM>(a, R, J, smem, g); return true;
    case 10: run_resident_call<5, float, true, EM>(a, R, J, smem, g); return true;
    default: return false;
  }
}

__global__ __launch_bounds__(RNT) void resident_batch_kernel(LaunchArgs a) {
  extern __shared__ __align__(16) char smem[];
  const unsigned e = a.blk[blockIdx.x];
  if (e == ~0u) return;
  const int rep = (int)(e >> 12);
  if (rep >= a.nrep) return;
  const DescD R = ((const DescD *)a.reps)[rep];
  JRef J = *R.job;
  Group g;
  const int parts = (int)((e >> 4) & 15u);
  g.parts = parts < 1 ? 1 : parts;
  g.part = (int)(e & 15u);
  g.ctr = a.sync + rep;
  g.abort = a.abort;
  g.same_xcd = false;
  g.mode = a.same_xcd_mode;
  g.xbuf = gen(R.xbuf);
  g.epoch = 0;
  const unsigned variant = (e >> 8) & 15u;
  const int ens = J.ens_fixed ? ens_code(J.tmode, J.pmode, J.cmode) : ENS_RUNTIME;
  if (ens == ENS_TUBE_NP && run_resident_fixed<ENS_TUBE_NP>(variant, a, R, J, smem, g)) return;
  if (ens == ENS_SLIT_NP && run_resident_fixed<ENS_SLIT_NP>(variant, a, R, J, smem, g)) return;
  if (ens == ENS_BULK_NP && run_resident_fixed<ENS_BULK_NP>(variant, a, R, J, smem, g)) return;
  switch (variant) {
    case 0: run_resident_call<3, float, false>(a, R, J, smem, g); break;
    case 6: run_resident_call<3, float, true>(a, R, J, smem, g); break;
    case 1: run_resident_call<3, double, false>(a, R, J, smem, g); break;
    case 7: run_resident_call<3, double, true>(a, R, J, smem, g); break;
    case 2: run_resident_call<4, float, false>(a, R, J, smem, g); break;
    case 8: run_resident_call<4, float, true>(a, R, J, smem, g); break;
    case 3: run_resident_call<4, double, false>(a, R, J, smem, g); break;
    case 9: run_resident_call<4, double, true>(a, R, J, smem, g); break;
    case 4: run_resident_call<5, float, false>(a, R, J, smem, g); break;
    case 10: run_resident_call<5, float, true>(a, R, J, smem, g); break;
    case 5: run_resident_call<5, double, false>(a, R, J, smem, g); break;
    case 11: run_resident_call<5, double, true>(a, R, J, smem, g); break;
    default: break;
  }
}

// ---------------------------------------------------------------- b1
// ForceCalculator::CalcWallForce, tube (src/force_calculator.cxx:117-161):
// one thread per table node, 10000-point phi sum in FP64 with float sqrt as
// the reference, in the reference's summation order.
__global__ void wall_tube_energy_kernel(int nfwall, double R, double sgm06, double coefE, double *fw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfwall) return;
  const double dr = R / (double)nfwall;
  const int ninteg = 10000;
  const double dphi = M_PI / (double)ninteg;
  const double R02 = R * R;
  const double r = (double)i * dr;
  const double r02 = r * r;
  double sum = 0.0;
  for (int j = 0; j < ninteg; j++) {
    const double phi = M_PI - (double)j * dphi;
    const double rho_min = (double)sqrtf((float)(R02 + r02 - 2.0 * R * r * cos(phi)));
    const double rho_mini = 1.0 / rho_min;
    const double rho_min05i = rho_mini * rho_mini * rho_mini * rho_mini * rho_mini;
    double tmp = 63.0 / 32.0 * sgm06 * rho_min05i * rho_min05i * rho_mini;
    tmp -= 3.0 * rho_min05i;
    sum += tmp;
  }
  fw[2 * i + 1] = coefE * sum * R * dphi;
}

__global__ void wall_tube_force_kernel(int nfwall, double R, double *fw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double dr = R / (double)nfwall;
  if (i >= 1 && i < nfwall - 1) {
    const double r = (double)i * dr;
    fw[2 * i] = -(fw[2 * (i + 1) + 1] - fw[2 * (i - 1) + 1]) / (2.0 * dr * r);
  }
}

__global__ void wall_tube_edges_kernel(int nfwall, double *fw) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fw[0] = fw[2];
  fw[2 * (nfwall - 1)] = (fw[2 * (nfwall - 2)] - fw[2 * (nfwall - 3)]) * 2.0 + fw[2 * (nfwall - 3)];
}

// Slit wall (10-4 graphene sheets at z = +-H/2; see the oracle's contract).
__device__ inline double w104_e(double h, double sgm, double eps, double rho) {
  const double x = sgm / h, x2 = x * x, x4 = x2 * x2, x10 = x4 * x4 * x2;
  return 2.0 * M_PI * rho * eps * sgm * sgm * (0.4 * x10 - x4);
}
__device__ inline double w104_dedh(double h, double sgm, double eps, double rho) {
  const double x = sgm / h, x2 = x * x, x4 = x2 * x2, x10 = x4 * x4 * x2;
  return 2.0 * M_PI * rho * eps * sgm * sgm * (-4.0 * x10 + 4.0 * x4) / h;
}
__global__ void wall_slit_kernel(int nfwall, double H, double sgm, double eps, double rho, double *fw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfwall) return;
  const double dr = 0.5 * H / (double)nfwall;
  const double z = (double)i * dr;
  const double hd = 0.5 * H + z, hu = 0.5 * H - z;
  fw[2 * i + 1] = w104_e(hd, sgm, eps, rho) + w104_e(hu, sgm, eps, rho);
  const double dedz = w104_dedh(hd, sgm, eps, rho) - w104_dedh(hu, sgm, eps, rho);
  fw[2 * i] = (i == 0) ? 0.0 : -dedz / z;
}
__global__ void wall_slit_edge_kernel(double *fw) {
  if (threadIdx.x == 0 && blockIdx.x == 0) fw[0] = fw[2];
}

// ----------------------------------------------------------------- c1
// Standalone sort + window for the parity hook: one workgroup per row of
// n keys (fractional coordinates in [0,1)).
template <int NT>
__global__ __launch_bounds__(NT) void sort_window_kernel(const double *keys, int n, double w, int *order,
                                                         int *js_out, int *je_out) {
  extern __shared__ __align__(16) char smem[];
  const double *z = keys + (size_t)blockIdx.x * n;
  uint64_t *keyA = (uint64_t *)smem;
  uint64_t *keyB = keyA + n;
  uint16_t *idxA = (uint16_t *)(keyB + n);
  uint16_t *idxB = idxA + n;
  int *cnt = (int *)(idxB + n + (n & 1));
  for (int i = threadIdx.x; i < n; i += NT) {
    keyA[i] = (uint64_t)__double_as_longlong(z[i] + 0.0);
    idxA[i] = (uint16_t)i;
  }
  __syncthreads();
  radix_sort<NT>(keyA, idxA, keyB, idxB, cnt, n);
  double *zs = (double *)keyB;
  for (int i = threadIdx.x; i < n; i += NT) zs[i] = z[idxA[i]];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NT) {
    int js, je;
    window_of(zs, n, i, w, js, je);
    order[(size_t)blockIdx.x * n + i] = idxA[i];
    js_out[(size_t)blockIdx.x * n + i] = js;
    je_out[(size_t)blockIdx.x * n + i] = je;
  }
}

// ----------------------------------------------------------------- e1
// Metropolis over one pair list (REMD::ReplicaExchange,
// src/remd.cxx:486-538 with REMDInfo::ExchangeAccepted/Rejected,
// src/remdinfo.h:276-359) on device, fed uniforms drawn by the host in
// pair-list order.  One thread per job; writes the new conditions and the
// exchange flag into each replica's scalar block.

// Slots (a, b) of pair k of exchange list `type` on a dt x dp grid, in the
// list order of REMDInfo::ExchangeList (src/remdinfo.cxx:54-108); returns
// the number of pairs (k beyond it: a = b = -1).
__device__ inline int exchange_pair(int dt, int dp, int type, int k, int &a, int &b) {
  int n, i, j;
  a = b = -1;
  if (type == 0) {
    const int w = dt / 2;
    n = dp * w;
    if (k >= n || w == 0) return n;
    j = k / w, i = k - j * w;
    a = 2 * i + j * dt;
    b = a + 1;
  } else if (type == 1) {
    n = (dp / 2) * dt;
    if (k >= n) return n;
    j = k / dt, i = k - j * dt;
    a = i + 2 * j * dt;
    b = a + dt;
  } else if (type == 2) {
    const int w = dt / 2 - (dt + 1) % 2;
    n = w > 0 ? dp * w : 0;
    if (k >= n) return n;
    j = k / w, i = k - j * w;
    a = 2 * i + j * dt + 1;
    b = a + 1;
  } else {
    const int h = dp / 2 - (dp + 1) % 2;
    n = h > 0 ? h * dt : 0;
    if (k >= n) return n;
    j = k / dt, i = k - j * dt;
    a = i + (2 * j + 1) * dt;
    b = a + dt;
  }
  return n;
}

// REMD::ReplicaExchange (src/remd.cxx:486-538) on the device, one workgroup
// per job.  The pairs of one list share no slot (so no replica), so they are
// decided side by side, one thread per pair, each with its own uniform
// u[k] (pair-list order): the decisions and swaps of the serial loop.
__global__ void exchange_kernel(const ExchangeJob *jobs, int njobs) {
  const int jid = blockIdx.x;
  if (jid >= njobs) return;
  const ExchangeJob &X = jobs[jid];
  const int dt = X.ntemp, dp = X.npress;
  const int nrep = dt * dp;
  if (X.hist)
    for (int i = threadIdx.x; i < nrep * NSC; i += blockDim.x) X.hist[i] = X.sc[0][i];
  if (X.sync)
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) X.sync[r] = 0ull;
  __syncthreads();
  for (int r = threadIdx.x; r < nrep; r += blockDim.x) X.sc[r][S_XFLAG] = 0.0;
  __syncthreads();
  int a, b;
  const int np = exchange_pair(dt, dp, X.ec_type, 0, a, b);
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    exchange_pair(dt, dp, X.ec_type, k, a, b);
    const int ind1 = X.index[a], ind2 = X.index[b];
    const double beta1 = 1. / X.T[ind1], beta2 = 1. / X.T[ind2];
    const double pot1 = X.sc[ind1][S_POT], pot2 = X.sc[ind2][S_POT];
    double delta = (beta2 - beta1) * (pot1 - pot2);
    if (X.pmode > 0) {
      const double v1 = X.sc[ind1][S_VOL], v2 = X.sc[ind2][S_VOL];
      delta += (beta2 * X.P[ind2] - beta1 * X.P[ind1]) * (v1 - v2);
    }
    const int acc = exp(-delta) > X.u[k];
    if (X.accepted) X.accepted[k] = acc;
    if (acc) {
      X.sc[ind1][S_XFLAG] = 1.0;
      X.sc[ind2][S_XFLAG] = 1.0;
      double t = X.T[ind1]; X.T[ind1] = X.T[ind2]; X.T[ind2] = t;
      t = X.P[ind1]; X.P[ind1] = X.P[ind2]; X.P[ind2] = t;
      X.index[a] = ind2;
      X.index[b] = ind1;
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
    X.sc[r][S_TNEW] = X.T[r];
    X.sc[r][S_PNEW] = X.P[r];
  }
}

// ------------------------------------------------ scratch-table set-up
// Set-up of an interval launch whose scratch sits in one caller-free buffer
// (the allocation-free entry point): zeroes the nsync group words (the
// abort word is the last), copies the job's replica table with replica k's
// exchange rows moved to xbase + k * xstride (none when xbase is null), and
// writes the block table of a split-global launch with the single-job
// mapping (the parts of a replica on block indices equal mod 8).
__global__ void scratch_setup_kernel(const ReplicaDesc *src, ReplicaDesc *dst, int nrep, double *xbase,
                                     size_t xstride, unsigned long long *sync, int nsync, unsigned *blk,
                                     int nb, int parts, unsigned variant) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nsync) sync[k] = 0ull;
  if (k < nrep) {
    ReplicaDesc d = src[k];
    d.xbuf = xbase ? xbase + (size_t)k * xstride : nullptr;
    dst[k] = d;
  }
  if (k < nb) {
    int rep, part;
    if (parts <= 1) {
      rep = k;
      part = 0;
    } else {
      const int slot = k >> 3;
      rep = (slot / parts) * 8 + (k & 7);
      part = slot % parts;
    }
    blk[k] = rep < nrep ? ((unsigned)rep << 12 | variant << 8 | (unsigned)parts << 4 | (unsigned)part) : ~0u;
  }
}

// ------------------------------------------------------------ launchers
// Kernel class of a replica: (block threads, sites, pair precision).
template <int NT>
static hipError_t launch_nt(const LaunchArgs &a, int nsite, int prec, int lds, hipStream_t st) {
  const dim3 g(a.nrep), b(NT);
  const int key = nsite * 2 + prec;
  switch (key) {
    case 6: hipLaunchKernelGGL((interval_kernel<NT, 3, float>), g, b, lds, st, a); break;
    case 8: hipLaunchKernelGGL((interval_kernel<NT, 4, float>), g, b, lds, st, a); break;
    case 10: hipLaunchKernelGGL((interval_kernel<NT, 5, float>), g, b, lds, st, a); break;
    case 7: hipLaunchKernelGGL((interval_kernel<NT, 3, double>), g, b, lds, st, a); break;
    case 9: hipLaunchKernelGGL((interval_kernel<NT, 4, double>), g, b, lds, st, a); break;
    case 11: hipLaunchKernelGGL((interval_kernel<NT, 5, double>), g, b, lds, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Resident path: block size 256 only.  The FP32 jobs of the ensembles
// ENS_*_NP run an instantiation with the modes fixed, the rest the one that
// reads them from the job (same bits either way).
static const void *resident_fn(int nsite, int prec, int ens) {
  if (prec == 0) {
#define REMD_RES_ENS(E)                                                  \
  if (ens == E) {                                                        \
    if (nsite == 3) return (const void *)resident_kernel<RNT, 3, float, E>; \
    if (nsite == 4) return (const void *)resident_kernel<RNT, 4, float, E>; \
    if (nsite == 5) return (const void *)resident_kernel<RNT, 5, float, E>; \
  }
    REMD_RES_ENS(ENS_TUBE_NP)
    REMD_RES_ENS(ENS_SLIT_NP)
    REMD_RES_ENS(ENS_BULK_NP)
#undef REMD_RES_ENS
  }
  switch (nsite * 2 + prec) {
    case 6: return (const void *)resident_kernel<RNT, 3, float>;
    case 8: return (const void *)resident_kernel<RNT, 4, float>;
    case 10: return (const void *)resident_kernel<RNT, 5, float>;
    case 7: return (const void *)resident_kernel<RNT, 3, double>;
    case 9: return (const void *)resident_kernel<RNT, 4, double>;
    case 11: return (const void *)resident_kernel<RNT, 5, double>;
    default: return nullptr;
  }
}

int resident_grid(int nrep, int parts) { return parts <= 1 ? nrep : 8 * parts * ((nrep + 7) / 8); }

int resident_lds(int nmol, int nsite, int prec, int site_dbg, int gm) {
  return resident_lds_bytes(nmol, nsite, prec, RNT, site_dbg != 0, gm != 0);
}

hipError_t launch_resident(const LaunchArgs &a, int nsite, int prec, int ens, int lds, hipStream_t st) {
  if (a.nrep <= 0) return hipSuccess;
  const void *f = resident_fn(nsite, prec, ens);
  if (!f) return hipErrorInvalidValue;
  LaunchArgs arg = a;
  void *args[] = {&arg};
  const hipError_t e = hipLaunchKernel(f, dim3(resident_grid(a.nrep, a.parts)), dim3(RNT), args, (size_t)lds, st);
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_resident_batch(const LaunchArgs &a, int lds, hipStream_t st) {
  if (a.nblk <= 0 || !a.blk) return hipSuccess;
  hipLaunchKernelGGL(resident_batch_kernel, dim3(a.nblk), dim3(RNT), lds, st, a);
  return hipGetLastError();
}

hipError_t resident_batch_occupancy(int lds, int *per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, (const void *)resident_batch_kernel, RNT,
                                                      (size_t)lds);
}

// Workgroups of the resident kernel one CU holds at once with lds bytes.
hipError_t resident_occupancy(int nsite, int prec, int ens, int lds, int *per_cu) {
  const void *f = resident_fn(nsite, prec, ens);
  if (!f) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, RNT, (size_t)lds);
}

hipError_t launch_interval(const LaunchArgs &a, int nt, int nsite, int prec, int lds, hipStream_t st) {
  if (a.nrep <= 0) return hipSuccess;
  if (nt == 256) return launch_nt<256>(a, nsite, prec, lds, st);
  if (nt == 512) return launch_nt<512>(a, nsite, prec, lds, st);
  if (nt == 128) return launch_nt<128>(a, nsite, prec, lds, st);
  return hipErrorInvalidValue;
}

template <int NT>
static hipError_t set_lds_nt(int bytes) {
  hipError_t e = hipSuccess;
#define REMD_SETLDS(NS_, PT_)                                                                              \
  if (e == hipSuccess)                                                                                      \
    e = hipFuncSetAttribute((const void *)interval_kernel<NT, NS_, PT_>,                                    \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  REMD_SETLDS(3, float)
  REMD_SETLDS(4, float)
  REMD_SETLDS(5, float)
  REMD_SETLDS(3, double)
  REMD_SETLDS(4, double)
  REMD_SETLDS(5, double)
#undef REMD_SETLDS
  return e;
}

hipError_t set_kernel_lds_limit(int bytes) {
  hipError_t e = hipSuccess;
  const int ens[4] = {ENS_RUNTIME, ENS_TUBE_NP, ENS_SLIT_NP, ENS_BULK_NP};
  for (int k = 0; k < 4 && e == hipSuccess; k++)
    for (int ns = 3; ns <= 5 && e == hipSuccess; ns++)
      for (int pr = 0; pr < 2 && e == hipSuccess; pr++)
        e = hipFuncSetAttribute(resident_fn(ns, pr, ens[k]), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)resident_batch_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            bytes);
  if (e == hipSuccess) e = set_lds_nt<256>(bytes);
  if (e == hipSuccess) e = set_lds_nt<512>(bytes);
  if (e == hipSuccess) e = set_lds_nt<128>(bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)sort_window_kernel<256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            bytes);
  return e;
}

hipError_t launch_wall_tube(int nfwall, double R, double sgm06, double coefE, double *fw, hipStream_t st) {
  const int nb = (nfwall + 255) / 256;
  hipLaunchKernelGGL(wall_tube_energy_kernel, dim3(nb), dim3(256), 0, st, nfwall, R, sgm06, coefE, fw);
  hipLaunchKernelGGL(wall_tube_force_kernel, dim3(nb), dim3(256), 0, st, nfwall, R, fw);
  hipLaunchKernelGGL(wall_tube_edges_kernel, dim3(1), dim3(64), 0, st, nfwall, fw);
  return hipGetLastError();
}

hipError_t launch_wall_slit(int nfwall, double H, double sgm, double eps, double rho, double *fw,
                            hipStream_t st) {
  const int nb = (nfwall + 255) / 256;
  hipLaunchKernelGGL(wall_slit_kernel, dim3(nb), dim3(256), 0, st, nfwall, H, sgm, eps, rho, fw);
  hipLaunchKernelGGL(wall_slit_edge_kernel, dim3(1), dim3(64), 0, st, fw);
  return hipGetLastError();
}

hipError_t launch_sort_window(const double *keys, int nrow, int n, double w, int *order, int *js, int *je,
                              hipStream_t st) {
  const int lds = 20 * n + 8 + (16 * 64 + 16) * 4;
  hipLaunchKernelGGL(sort_window_kernel<256>, dim3(nrow), dim3(256), lds, st, keys, n, w, order, js, je);
  return hipGetLastError();
}

hipError_t launch_scratch_setup(const ReplicaDesc *src, ReplicaDesc *dst, int nrep, double *xbase,
                                size_t xstride, unsigned long long *sync, int nsync, unsigned *blk, int nb,
                                int parts, unsigned variant, hipStream_t st) {
  int n = nsync > nrep ? nsync : nrep;
  if (nb > n) n = nb;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(scratch_setup_kernel, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, nrep, xbase, xstride,
                     sync, nsync, blk, nb, parts, variant);
  return hipGetLastError();
}

hipError_t launch_exchange(const ExchangeJob *jobs, int njobs, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(exchange_kernel, dim3(njobs), dim3(256), 0, st, jobs, njobs);
  return hipGetLastError();
}

}  // namespace remd
