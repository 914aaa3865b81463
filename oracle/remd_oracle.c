/*
 * remd_oracle.c -- CPU restatement of the reference REMD per-replica path.
 * TEST INFRASTRUCTURE ONLY (see remd_oracle.h).  Never linked by the product.
 *
 * Operation order in the FP64 force path follows the reference's tuned
 * kernel expression by expression (src/force_calculator.cxx:803-989,
 * 1060-1472) so that, on the TIP4P tube, the oracle agrees with the
 * reference's own FP64 build (oracle/_ref) to the last bit or close to it.
 * Build with -ffp-contract=off (oracle/Makefile).
 */
#include "remd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

/* ---------------------------------------------------------------- units */
/* src/unit.h:3-18 */
static const double kNa = 6.0221367e+23;
static const double kE0 = 8.8542e-12;
static const double kEle = 1.6021761e-19;
static const double kKB = 1.3806488e-23;

static double u_mass(void) { return 15.9994e-3 / kNa; }
static double u_length(void) { return 1.0e-10; }
static double u_energy(void) { return 1e3 / kNa; }
static double u_time(void) {
  return sqrt(u_mass() * u_length() * u_length() / u_energy());
}
static double u_coulomb(void) {
  return sqrt(4. * M_PI * kE0 * u_energy() * u_length());
}
static double u_temp(void) { return u_energy() / kKB; }
static double u_press(void) {
  return u_energy() / (u_length() * u_length() * u_length());
}

void orc_units(double out[9]) {
  out[0] = u_mass();
  out[1] = u_length();
  out[2] = u_energy();
  out[3] = u_time();
  out[4] = u_coulomb();
  out[5] = u_temp();
  out[6] = u_press();
  out[7] = kNa;
  out[8] = kEle;
}

/* glibc rand() is the reference's exchange RNG (src/remd.cxx:465). */
void orc_srand(unsigned seed) { srand(seed); }
int orc_rand(void) { return rand(); }

/* ------------------------------------------------------------- moltype */
/* MolType::CalcIner (src/variables.h:189-206) */
static void calc_inertia(orc_moltype *mt) {
  double xx = 0, yy = 0, zz = 0;
  for (int j = 0; j < mt->nsite; j++) {
    const double m = mt->m[j];
    const double x = mt->r[j][0], y = mt->r[j][1], z = mt->r[j][2];
    xx += m * (y * y + z * z);
    yy += m * (z * z + x * x);
    zz += m * (x * x + y * y);
  }
  mt->inertia[0] = xx;
  mt->inertia[1] = yy;
  mt->inertia[2] = zz;
}

/* Molecules::MakeMolTypeList(std::string) (src/molecules.cxx:69-131):
 * line-oriented; a block starts with a name whose first character is a
 * letter, followed by natom lines "name x y z m sigma eps q". */
int orc_moltype_parse(const char *text, int k, orc_moltype *out) {
  const char *p = text;
  int found = -1;
  char line[512];
  while (*p) {
    const char *e = strchr(p, '\n');
    size_t len = e ? (size_t)(e - p) : strlen(p);
    if (len >= sizeof(line)) len = sizeof(line) - 1;
    memcpy(line, p, len);
    line[len] = 0;
    p = e ? e + 1 : p + len;
    char name[128];
    int natom = 0;
    if (sscanf(line, "%127s", name) != 1) continue;
    if (!isalpha((unsigned char)name[0])) continue;
    if (sscanf(line, "%127s %d", name, &natom) != 2) continue;
    found++;
    orc_moltype mt;
    memset(&mt, 0, sizeof(mt));
    if (natom < 1 || natom > ORC_MAXSITE) return -2;
    mt.nsite = natom;
    for (int i = 0; i < natom; i++) {
      const char *e2 = strchr(p, '\n');
      size_t l2 = e2 ? (size_t)(e2 - p) : strlen(p);
      if (l2 >= sizeof(line)) l2 = sizeof(line) - 1;
      memcpy(line, p, l2);
      line[l2] = 0;
      p = e2 ? e2 + 1 : p + l2;
      char an[64];
      double r0, r1, r2, m, s, ep, q;
      if (sscanf(line, "%63s %lf %lf %lf %lf %lf %lf %lf", an, &r0, &r1, &r2,
                 &m, &s, &ep, &q) != 8)
        return -3;
      /* normalisation, src/molecules.cxx:98-103 */
      mt.r[i][0] = r0 * (1e-10 / u_length());
      mt.r[i][1] = r1 * (1e-10 / u_length());
      mt.r[i][2] = r2 * (1e-10 / u_length());
      mt.m[i] = m * (1e-03 / kNa / u_mass());
      mt.sigma[i] = s * (1e-10 / u_length());
      mt.eps[i] = ep * (1e+03 / kNa / u_energy());
      mt.q[i] = q * (kEle / u_coulomb());
    }
    if (found == k) {
      /* the reference normalises r *= 1e-10/unit_length as a vector op;
       * 1e-10/1e-10 == 1 exactly, so the values are the parsed ones. */
      double mass = 0;
      for (int i = 0; i < natom; i++) mass += mt.m[i];
      mt.mass = mass;
      calc_inertia(&mt);
      *out = mt;
      return 0;
    }
  }
  return -1;
}

/* -------------------------------------------------------------- system */
int orc_system_alloc(orc_system *s, int nmol) {
  memset(s, 0, sizeof(*s));
  s->nmol = nmol;
  s->r = calloc(3 * (size_t)nmol, sizeof(double));
  s->v = calloc(3 * (size_t)nmol, sizeof(double));
  s->f = calloc(3 * (size_t)nmol, sizeof(double));
  s->q = calloc(4 * (size_t)nmol, sizeof(double));
  s->p = calloc(4 * (size_t)nmol, sizeof(double));
  s->n = calloc(4 * (size_t)nmol, sizeof(double));
  s->site_f = calloc(3 * ORC_MAXSITE * (size_t)nmol, sizeof(double));
  s->id = calloc((size_t)nmol, sizeof(int));
  s->jstart = calloc((size_t)nmol, sizeof(int));
  s->jend = calloc((size_t)nmol, sizeof(int));
  s->wfull = calloc((size_t)nmol, sizeof(int));
  s->sort_interval = 10;
  s->margin = 0.5;
  s->sort_axis = 2;
  s->s = 1.0;
  for (int i = 0; i < nmol; i++) {
    s->q[4 * i] = 1.0;
    s->id[i] = i;
  }
  return (s->r && s->v && s->f && s->q && s->p && s->n && s->site_f && s->id &&
          s->jstart && s->jend && s->wfull)
             ? 0
             : -1;
}

void orc_system_free(orc_system *s) {
  free(s->wfull);
  free(s->r);
  free(s->v);
  free(s->f);
  free(s->q);
  free(s->p);
  free(s->n);
  free(s->site_f);
  free(s->id);
  free(s->jstart);
  free(s->jend);
  memset(s, 0, sizeof(*s));
}

/* Molecules::GetBottomArea (src/molecules.h:367-376): pi (R - sigma_c/2)^2,
 * sigma_c = 3.4 A. */
double orc_bottom_area(const orc_system *s) {
  const double sigma_c = 3.4;
  const double wl = (s->wall_length - 0.5 * sigma_c) * 1e-10 / u_length();
  return M_PI * wl * wl;
}

/* Molecules::GetVolume (src/molecules.h:357-366). */
double orc_volume(const orc_system *s) {
  if (s->cmode == 1) return orc_bottom_area(s) * s->L[2];
  if (s->cmode == 2) {
    const double wl = s->wall_length * 1e-10 / u_length();
    return s->L[0] * s->L[1] * wl;
  }
  return s->L[0] * s->L[1] * s->L[2];
}

/* -------------------------------------------------------- sort/window */
/* Stable merge sort of indices by key: equals std::sort with the
 * reference's comparator (src/molecules.cxx:478, 546) for distinct keys,
 * and breaks ties by the current index. */
static void merge_sort_idx(const double *key, int *idx, int *tmp, int n) {
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n;
      int hi = lo + 2 * w < n ? lo + 2 * w : n;
      int a = lo, b = mid, o = lo;
      while (a < mid && b < hi) tmp[o++] = (key[idx[b]] < key[idx[a]]) ? idx[b++] : idx[a++];
      while (a < mid) tmp[o++] = idx[a++];
      while (b < hi) tmp[o++] = idx[b++];
    }
    memcpy(idx, tmp, sizeof(int) * (size_t)n);
  }
}

/* jstart/jend for nlane = 1 (src/molecules.cxx:556-583): js is the first
 * index scanning downward from i (cyclically) whose key is below z_i - w,
 * je the first index scanning upward whose key exceeds z_i + w. */
void orc_window(const double *z, int n, double w, int *js, int *je) {
  for (int i = 0; i < n; i++) {
    int s = i;
    double lb = z[s] - w;
    while (z[s] >= lb) {
      s--;
      if (s < 0) {
        s += n;
        lb += 1.0;
      }
    }
    js[i] = s;
    int e = i;
    double ub = z[e] + w;
    while (z[e] <= ub) {
      e++;
      if (e >= n) {
        e -= n;
        ub -= 1.0;
      }
    }
    je[i] = e;
  }
}

void orc_window_full(const double *z, int n, double w, const int *js,
                     const int *je, int *full) {
  for (int i = 0; i < n; i++) {
    const int a = js[i], b = je[i];
    /* a inside the upward arc, b inside the downward arc, with the scan's
     * own bounds (z_i + w, minus 1 after a wrap; z_i - w, plus 1) */
    const int a_up = (a > i) ? (z[a] <= z[i] + w) : (z[a] <= (z[i] + w) - 1.0);
    const int b_dn = (b < i) ? (z[b] >= z[i] - w) : (z[b] >= (z[i] - w) + 1.0);
    full[i] = (a == i || b == i || a_up || b_dn) ? 1 : 0;
  }
}

void orc_sort_window(const double *key, int n, double w, int *order, int *js,
                     int *je) {
  int *tmp = malloc(sizeof(int) * (size_t)n);
  double *sk = malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) order[i] = i;
  merge_sort_idx(key, order, tmp, n);
  for (int i = 0; i < n; i++) sk[i] = key[order[i]];
  orc_window(sk, n, w, js, je);
  free(tmp);
  free(sk);
}

static void permute(double *a, int width, const int *order, int n, double *buf) {
  for (int i = 0; i < n; i++)
    for (int d = 0; d < width; d++) buf[width * i + d] = a[width * order[i] + d];
  memcpy(a, buf, sizeof(double) * (size_t)width * (size_t)n);
}

static void sort_molecules(orc_system *s, double margin) {
  const int n = s->nmol;
  const int ax = s->sort_axis;
  double *key = malloc(sizeof(double) * (size_t)n);
  int *order = malloc(sizeof(int) * (size_t)n);
  int *tmp = malloc(sizeof(int) * (size_t)n);
  double *buf = malloc(sizeof(double) * 4 * (size_t)n);
  for (int i = 0; i < n; i++) key[i] = s->r[3 * i + ax];
  for (int i = 0; i < n; i++) order[i] = i;
  merge_sort_idx(key, order, tmp, n);
  permute(s->r, 3, order, n, buf);
  permute(s->v, 3, order, n, buf);
  permute(s->f, 3, order, n, buf);
  permute(s->q, 4, order, n, buf);
  permute(s->p, 4, order, n, buf);
  permute(s->n, 4, order, n, buf);
  for (int i = 0; i < n; i++) tmp[i] = s->id[order[i]];
  memcpy(s->id, tmp, sizeof(int) * (size_t)n);
  /* window on the sorted coordinate (src/molecules.cxx:556-583) */
  for (int i = 0; i < n; i++) key[i] = s->r[3 * i + ax];
  orc_window(key, n, (s->rcut + margin) / s->L[ax], s->jstart, s->jend);
  orc_window_full(key, n, (s->rcut + margin) / s->L[ax], s->jstart, s->jend, s->wfull);
  free(key);
  free(order);
  free(tmp);
  free(buf);
}

/* ----------------------------------------------------------- rotation */
/* body_to_space / space_to_body (src/rotation.h:7-48) */
static void body_to_space(const double *a, const double *v, double *t) {
  const double a0sq = a[0] * a[0], a1sq = a[1] * a[1], a2sq = a[2] * a[2],
               a3sq = a[3] * a[3];
  const double a01 = a[0] * a[1], a02 = a[0] * a[2], a03 = a[0] * a[3];
  const double a12 = a[1] * a[2], a13 = a[1] * a[3];
  const double a23 = a[2] * a[3];
  t[0] = (a0sq + a1sq - a2sq - a3sq) * v[0];
  t[0] += 2. * (a12 - a03) * v[1];
  t[0] += 2. * (a13 + a02) * v[2];
  t[1] = 2. * (a12 + a03) * v[0];
  t[1] += (a0sq - a1sq + a2sq - a3sq) * v[1];
  t[1] += 2. * (a23 - a01) * v[2];
  t[2] = 2. * (a13 - a02) * v[0];
  t[2] += 2. * (a23 + a01) * v[1];
  t[2] += (a0sq - a1sq - a2sq + a3sq) * v[2];
}

static void space_to_body(const double *a, const double *v, double *t) {
  const double a0sq = a[0] * a[0], a1sq = a[1] * a[1], a2sq = a[2] * a[2],
               a3sq = a[3] * a[3];
  const double a01 = a[0] * a[1], a02 = a[0] * a[2], a03 = a[0] * a[3];
  const double a12 = a[1] * a[2], a13 = a[1] * a[3];
  const double a23 = a[2] * a[3];
  t[0] = (a0sq + a1sq - a2sq - a3sq) * v[0];
  t[0] += 2. * (a12 + a03) * v[1];
  t[0] += 2. * (a13 - a02) * v[2];
  t[1] = 2. * (a12 - a03) * v[0];
  t[1] += (a0sq - a1sq + a2sq - a3sq) * v[1];
  t[1] += 2. * (a23 + a01) * v[2];
  t[2] = 2. * (a13 + a02) * v[0];
  t[2] += 2. * (a23 - a01) * v[1];
  t[2] += (a0sq - a1sq - a2sq + a3sq) * v[2];
}

/* ------------------------------------------------------- pair kernels */
/* The reference's literal constants are double (force_calculator.cxx:781-789)
 * and stay double in its FP32 build, so the float variant keeps them. */
#define ZERO 0.0
#define HALF 0.5
#define ONE 1.0
#define FIVE 5.0
#define TEN 10.0
#define TWELVE 12.0
#define THIRTY 30.0

#define FP double
#define FPSQRT sqrt
#define FPROUND round
#define PAIR_FN pair_forces_f64
#include "remd_oracle_pair.inc"
#undef FP
#undef FPSQRT
#undef FPROUND
#undef PAIR_FN

#define FP float
#define FPSQRT sqrtf
#define FPROUND roundf
#define PAIR_FN pair_forces_f32
#include "remd_oracle_pair.inc"
#undef FP
#undef FPSQRT
#undef FPROUND
#undef PAIR_FN

/* --------------------------------------------------------------- wall */
/* ForceCalculator::CalcWallForce, confined_dim == 1
 * (src/force_calculator.cxx:117-161).  sqrtf/powf are float as in the
 * reference; everything else double. */
void orc_wall_table_tube(int nfwall, double R, double sgm, double eps,
                         double rho, double *fwall) {
  const double wall_length = R;
  const double dr = wall_length / (double)nfwall;
  const int ninteg = 10000;
  const double dphi = M_PI / (double)ninteg;
  const double sgm06 = powf((float)sgm, 6);
  const double R02 = wall_length * wall_length;
  for (int i = 0; i < nfwall; i++) {
    const double r = (double)i * dr;
    const double r02 = r * r;
    double tmp, sum = 0.0;
    for (int j = 0; j < ninteg; j++) {
      const double phi = M_PI - (double)j * dphi;
      const double rho_min =
          sqrtf((float)(R02 + r02 - 2.0 * wall_length * r * cos(phi)));
      const double rho_mini = 1.0 / rho_min;
      const double rho_min05i =
          rho_mini * rho_mini * rho_mini * rho_mini * rho_mini;
      tmp = 63.0 / 32.0 * sgm06 * rho_min05i * rho_min05i * rho_mini;
      tmp -= 3.0 * rho_min05i;
      sum += tmp;
    }
    fwall[2 * i + 1] = M_PI * rho * eps * sgm06 * sum * wall_length * dphi;
  }
  for (int i = 1; i < nfwall - 1; i++) {
    const double r = (double)i * dr;
    fwall[2 * i + 0] =
        -(fwall[2 * (i + 1) + 1] - fwall[2 * (i - 1) + 1]) / (2.0 * dr * r);
  }
  fwall[0] = fwall[2];
  fwall[2 * (nfwall - 1)] =
      (fwall[2 * (nfwall - 2)] - fwall[2 * (nfwall - 3)]) * 2.0 +
      fwall[2 * (nfwall - 3)];
}

/* Slit wall (contract; the reference's slit table at
 * src/force_calculator.cxx:163-199 uses sigma/epsilon as raw 9-3
 * coefficients and its lookup asserts false at :1525).  Two graphene-like
 * sheets of areal density rho at z = -H/2 and z = +H/2, each integrated LJ
 * (10-4 wall): E(h) = 2 pi rho eps sigma^2 [(2/5)(sigma/h)^10 - (sigma/h)^4].
 * Table index i <-> |z| = i * dr, dr = (H/2)/nfwall, the same spacing as the
 * reference slit table (0.5 * i * wall_length / nfwall).  Entry 0 is the
 * force in the f(z)/z form (F_z = f * z) and entry 1 the energy, as for the
 * tube; the derivative is analytic. */
static double wall104_e(double h, double sgm, double eps, double rho) {
  const double x = sgm / h;
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double x10 = x4 * x4 * x2;
  return 2.0 * M_PI * rho * eps * sgm * sgm * (0.4 * x10 - x4);
}
static double wall104_dedh(double h, double sgm, double eps, double rho) {
  const double x = sgm / h;
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double x10 = x4 * x4 * x2;
  /* d/dh [0.4 x^10 - x^4] = (-4 x^10 + 4 x^4) / h */
  return 2.0 * M_PI * rho * eps * sgm * sgm * (-4.0 * x10 + 4.0 * x4) / h;
}
void orc_wall_table_slit(int nfwall, double H, double sgm, double eps,
                         double rho, double *fwall) {
  const double dr = 0.5 * H / (double)nfwall;
  for (int i = 0; i < nfwall; i++) {
    const double z = (double)i * dr;
    const double hd = 0.5 * H + z; /* distance to the lower sheet */
    const double hu = 0.5 * H - z; /* distance to the upper sheet */
    const double e = wall104_e(hd, sgm, eps, rho) + wall104_e(hu, sgm, eps, rho);
    /* dE/dz = dE/dhd - dE/dhu */
    const double dedz =
        wall104_dedh(hd, sgm, eps, rho) - wall104_dedh(hu, sgm, eps, rho);
    fwall[2 * i + 1] = e;
    fwall[2 * i + 0] = (i == 0) ? 0.0 : -dedz / z;
  }
  if (nfwall > 1) fwall[0] = fwall[2];
}

/* ForceCalculator::Confined, FOUR_SITE branch
 * (src/force_calculator.cxx:1494-1531): O site (site 0) only; linear
 * interpolation between table nodes (int)(r/dr) and +1.  The slit branch
 * is the contract lookup on |z - Lz/2|. */
static void wall_forces(orc_system *s, const double *atom_r) {
  const int nsite = s->mt.nsite;
  double wall = 0.0;
  if (s->cmode == 1) {
    const double dr = s->wall_length / (double)s->nfwall;
    const double dri = 1.0 / dr;
    for (int i = 0; i < s->nmol; i++) {
      const int ai = nsite * i;
      const double rx = atom_r[3 * ai + 0] - s->L[0] * 0.5;
      const double ry = atom_r[3 * ai + 1] - s->L[1] * 0.5;
      const double r01 = sqrt(rx * rx + ry * ry);
      int index = (int)(r01 / dr);
      if (index > s->nfwall - 2) index = s->nfwall - 2;
      const double *fw = s->fwall;
      const double f = (fw[2 * (index + 1)] - fw[2 * index]) * dri * (r01 - dr * index) + fw[2 * index];
      const double e = (fw[2 * (index + 1) + 1] - fw[2 * index + 1]) * dri * (r01 - dr * index) + fw[2 * index + 1];
      s->site_f[3 * ai + 0] += f * rx;
      s->site_f[3 * ai + 1] += f * ry;
      s->f[3 * i + 0] += f * rx;
      s->f[3 * i + 1] += f * ry;
      /* wall_omp is FP (float in the USE_FP32 build, force_calculator.h:64) */
      if (s->fp32) wall = (double)(float)(wall + e); /* float += double */
      else wall += e;
    }
  } else if (s->cmode == 2) {
    const double dr = 0.5 * s->wall_length / (double)s->nfwall;
    const double dri = 1.0 / dr;
    for (int i = 0; i < s->nmol; i++) {
      const int ai = nsite * i;
      const double rz = atom_r[3 * ai + 2] - s->L[2] * 0.5;
      const double az = fabs(rz);
      int index = (int)(az / dr);
      if (index > s->nfwall - 2) index = s->nfwall - 2;
      const double *fw = s->fwall;
      const double f = (fw[2 * (index + 1)] - fw[2 * index]) * dri * (az - dr * index) + fw[2 * index];
      const double e = (fw[2 * (index + 1) + 1] - fw[2 * index + 1]) * dri * (az - dr * index) + fw[2 * index + 1];
      s->site_f[3 * ai + 2] += f * rz;
      s->f[3 * i + 2] += f * rz;
      wall += e;
    }
  }
  s->wall = wall;
}

/* ------------------------------------------------------------- force */
void orc_calc_force(orc_system *s, int do_sort, double margin) {
  const int n = s->nmol;
  const int nsite = s->mt.nsite;
  if (do_sort) sort_molecules(s, margin);
  /* ConvertToAtoms (src/molecules.cxx:319-360): atom position = r*L + R(q) b */
  double *atom_r = malloc(sizeof(double) * 3 * (size_t)nsite * (size_t)n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < nsite; d++) {
      double b2s[3];
      body_to_space(&s->q[4 * i], s->mt.r[d], b2s);
      double *a = &atom_r[3 * (nsite * i + d)];
      a[0] = s->r[3 * i + 0] * s->L[0] + b2s[0];
      a[1] = s->r[3 * i + 1] * s->L[1] + b2s[1];
      a[2] = s->r[3 * i + 2] * s->L[2] + b2s[2];
    }
  }
  if (s->fp32)
    pair_forces_f32(s, atom_r);
  else
    pair_forces_f64(s, atom_r);
  if (s->cmode > 0) wall_forces(s, atom_r);
  else s->wall = 0.0;
  /* ConvertFromAtoms (src/molecules.cxx:401-439) */
  for (int i = 0; i < n; i++) {
    double nn[4] = {0, 0, 0, 0};
    for (int d = 0; d < nsite; d++) {
      double fb[3];
      space_to_body(&s->q[4 * i], &s->site_f[3 * (nsite * i + d)], fb);
      const double *br = s->mt.r[d];
      nn[0] += 0.0 + br[0] * fb[0] + br[1] * fb[1] + br[2] * fb[2];
      nn[1] += br[1] * fb[2] - br[2] * fb[1];
      nn[2] += br[2] * fb[0] - br[0] * fb[2];
      nn[3] += br[0] * fb[1] - br[1] * fb[0];
    }
    for (int d = 0; d < 4; d++) s->n[4 * i + d] = nn[d];
  }
  s->pot = s->lj + s->clmb + s->wall;
  free(atom_r);
}

/* --------------------------------------------------------- integrator */
/* P1/P2/P3 (src/integrator.h:321-346) */
static void P1(const double *q, double *t) { t[0] = -q[1]; t[1] = q[0]; t[2] = q[3]; t[3] = -q[2]; }
static void P2(const double *q, double *t) { t[0] = -q[2]; t[1] = -q[3]; t[2] = q[0]; t[3] = q[1]; }
static void P3(const double *q, double *t) { t[0] = -q[3]; t[1] = q[2]; t[2] = -q[1]; t[3] = q[0]; }
static double dot4(const double *a, const double *b) {
  double t = 0.0;
  for (int d = 0; d < 4; d++) t += a[d] * b[d];
  return t;
}

/* Free rotation about body axis k by angle xi*h (src/integrator.cxx:467-563). */
static void rotate(double *q, double *p, int axis, double coef_xi, double inertia,
                   double h, double *xi_out) {
  double Pq[4], Pp[4];
  if (axis == 0) { P1(q, Pq); P1(p, Pp); }
  else if (axis == 1) { P2(q, Pq); P2(p, Pp); }
  else { P3(q, Pq); P3(p, Pp); }
  const double xi = dot4(p, Pq) * coef_xi / inertia;
  const double xidt = xi * h;
  const double c = cos(xidt), sn = sin(xidt);
  for (int d = 0; d < 4; d++) {
    const double qn = q[d] * c + Pq[d] * sn;
    const double pn = p[d] * c + Pp[d] * sn;
    q[d] = qn;
    p[d] = pn;
  }
  *xi_out = xi;
}

static int tmode_eff(const orc_system *s) {
  /* T_SCALE (2) integrates with the thermostat off (src/integrator.cxx:735-766) */
  return s->tmode == 1 ? 1 : 0;
}

/* D6 (src/integrator.cxx:634-644) */
static void D6(orc_system *s) {
  if (tmode_eff(s) == 1) {
    const double tmp = 1.0 + (s->Ps / (2.0 * s->Q) * s->dthalf);
    s->s *= tmp * tmp;
    s->Ps /= tmp;
  }
}

/* D5 (src/integrator.cxx:564-601); the 2D branch is the area barostat
 * (one area momentum, Lx = Ly = sqrt(A)) of the reference's SQUARE CUDA
 * branch (src/remdgpu.cu:2103-2107), not the CPU path's L[0]-only update. */
static void D5(orc_system *s) {
  const double cv0 = s->L[0] * s->s * s->dthalf;
  const double cv1 = s->L[1] * s->s * s->dthalf;
  const double cv2 = s->L[2] * s->s * s->dthalf;
  const double coef_p = 2.0 * s->s * s->dthalf;
  const double m = s->mt.mass;
  for (int i = 0; i < s->nmol; i++) {
    double *v = &s->v[3 * i];
    const double *f = &s->f[3 * i];
    v[0] += f[0] * cv0 / m;
    v[1] += f[1] * cv1 / m;
    v[2] += f[2] * cv2 / m;
    const double *q = &s->q[4 * i];
    const double *n = &s->n[4 * i];
    double *p = &s->p[4 * i];
    /* S(q) (src/integrator.h:303-310) times n */
    const double S0[4] = {q[0], -q[1], -q[2], -q[3]};
    const double S1[4] = {q[1], q[0], -q[3], q[2]};
    const double S2[4] = {q[2], q[3], q[0], -q[1]};
    const double S3[4] = {q[3], -q[2], q[1], q[0]};
    const double d0 = dot4(S0, n), d1 = dot4(S1, n), d2 = dot4(S2, n), d3 = dot4(S3, n);
    p[0] += d0 * coef_p;
    p[1] += d1 * coef_p;
    p[2] += d2 * coef_p;
    p[3] += d3 * coef_p;
  }
  if (tmode_eff(s) == 1) {
    s->Ps -= s->pot * s->dthalf;
    if (s->pmode > 0) s->Ps -= (s->P * orc_volume(s)) * s->dthalf;
  }
  if (s->pmode == 1)
    s->Pv[0] += ((s->vir[0] + s->vir[1] + s->vir[2]) / (3.0 * s->L[0] * s->L[1] * s->L[2]) - s->P) * s->s * s->dthalf;
  if (s->pmode == 2)
    s->Pv[2] += s->s * (s->vir[2] / s->L[2] - s->P * orc_bottom_area(s)) * s->dthalf;
  if (s->pmode == 3) {
    /* area barostat: -dU/dA - P * H with Lx = Ly = sqrt(A) */
    const double A = s->L[0] * s->L[1];
    s->Pv[0] += s->s * ((s->vir[0] + s->vir[1]) / (2.0 * A) - s->P * s->wall_length) * s->dthalf;
  }
}

/* D4 (src/integrator.cxx:604-632).  pconstant 1 keeps the barostat kinetic
 * energy Pv^2/(2W) of BarostatEnergy (molecules.cxx:666) consistent, as the
 * reference GPU kernel does (remdgpu.cu:2090-2095). */
static void D4(orc_system *s) {
  if (s->pmode == 1) {
    if (tmode_eff(s) == 1) s->Ps -= 0.5 * s->Pv[0] * s->Pv[0] / s->W * s->dthalf;
    const double V = orc_volume(s) + s->Pv[0] * s->s / s->W * s->dthalf;
    const double l = cbrt(V);
    s->L[0] = s->L[1] = s->L[2] = l;
  }
  if (s->pmode == 2) {
    if (tmode_eff(s) == 1) s->Ps -= 0.5 * s->Pv[2] * s->Pv[2] / s->W * s->dthalf;
    s->L[2] += s->s * s->Pv[2] / s->W * s->dthalf;
  }
  if (s->pmode == 3) {
    if (tmode_eff(s) == 1) s->Ps -= 0.5 * s->Pv[0] * s->Pv[0] / s->W * s->dthalf;
    const double A = s->L[0] * s->L[1] + s->s * s->Pv[0] / s->W * s->dthalf;
    const double l = sqrt(A);
    s->L[0] = s->L[1] = l;
  }
}

/* D3 D2 D1 D2 D3 + PsPv (src/integrator.cxx:120-172, 467-563) */
static void D3D2D1D2D3(orc_system *s) {
  const int T1 = tmode_eff(s) == 1;
  const int needPv = (s->pmode > 0) || T1;
  double sum_Pv[3] = {0, 0, 0};
  double sum_Ps = 0.0;
  const double coef_xi = 0.25 / s->s;
  const double *I = s->mt.inertia;
  const double m = s->mt.mass;
  double xi;
  /* D3 */
  for (int i = 0; i < s->nmol; i++) {
    rotate(&s->q[4 * i], &s->p[4 * i], 2, coef_xi, I[2], s->dthalf, &xi);
    if (T1) sum_Ps += 0.5 * xi * xi * I[2];
  }
  /* D2 */
  for (int i = 0; i < s->nmol; i++) {
    rotate(&s->q[4 * i], &s->p[4 * i], 1, coef_xi, I[1], s->dthalf, &xi);
    if (T1) sum_Ps += 0.5 * xi * xi * I[1];
  }
  /* D1 */
  {
    const double si = 1.0 / s->s;
    double coef_r[3];
    for (int d = 0; d < 3; d++) {
      const double Li = 1.0 / s->L[d];
      coef_r[d] = Li * Li * si * s->dt;
    }
    const double coef_xi1 = 0.25 * si;
    for (int i = 0; i < s->nmol; i++) {
      double *r = &s->r[3 * i];
      const double *v = &s->v[3 * i];
      for (int d = 0; d < 3; d++) {
        r[d] += v[d] * coef_r[d];
        if (r[d] < 0.0) r[d] += 1.0;
        if (r[d] >= 1.0) r[d] -= 1.0;
      }
      rotate(&s->q[4 * i], &s->p[4 * i], 0, coef_xi1, I[0], s->dt, &xi);
      if (needPv) {
        sum_Pv[0] += 2.0 * m * v[0] * v[0];
        sum_Pv[1] += 2.0 * m * v[1] * v[1];
        sum_Pv[2] += 2.0 * m * v[2] * v[2];
      }
      if (T1) sum_Ps += xi * xi * I[0];
    }
  }
  /* D2 */
  for (int i = 0; i < s->nmol; i++) {
    rotate(&s->q[4 * i], &s->p[4 * i], 1, coef_xi, I[1], s->dthalf, &xi);
    if (T1) sum_Ps += 0.5 * xi * xi * I[1];
  }
  /* D3 */
  for (int i = 0; i < s->nmol; i++) {
    rotate(&s->q[4 * i], &s->p[4 * i], 2, coef_xi, I[2], s->dthalf, &xi);
    if (T1) sum_Ps += 0.5 * xi * xi * I[2];
  }
  /* PsPv (src/integrator.cxx:145-171) */
  if (needPv) {
    for (int d = 0; d < 3; d++) sum_Pv[d] /= s->L[d] * s->L[d] * s->s * s->s;
  }
  if (s->pmode == 1)
    s->Pv[0] += s->s * (sum_Pv[0] + sum_Pv[1] + sum_Pv[2]) / (orc_volume(s) * 3.0) * s->dthalf;
  if (s->pmode == 2) s->Pv[2] += s->s * sum_Pv[2] / s->L[2] * s->dthalf;
  if (s->pmode == 3) {
    /* -d(KE_x+KE_y)/dA = (KE_x+KE_y)/A, KE_d = sum_Pv[d]/4 */
    const double A = s->L[0] * s->L[1];
    s->Pv[0] += s->s * (sum_Pv[0] + sum_Pv[1]) / (2.0 * A) * s->dthalf;
  }
  if (T1)
    s->Ps += (0.25 * (sum_Pv[0] + sum_Pv[1] + sum_Pv[2]) + 2.0 * sum_Ps -
              (1.0 + log(s->s)) * s->gkT + s->H0) * s->dt;
}

void orc_step(orc_system *s, int do_sort, double margin) {
  D6(s);
  D5(s);
  D4(s);
  D3D2D1D2D3(s);
  D4(s);
  orc_calc_force(s, do_sort, margin);
  D5(s);
  D6(s);
}

/* -------------------------------------------------------- properties */
static void translational_energy(orc_system *s, double *tmp) {
  double coef[3], kin[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) coef[d] = 1.0 / (s->L[d] * s->s);
  for (int i = 0; i < s->nmol; i++)
    for (int d = 0; d < 3; d++) kin[d] += (s->v[3 * i + d] * s->v[3 * i + d]) * s->mt.mass;
  for (int d = 0; d < 3; d++) tmp[d] = kin[d] * (coef[d] * coef[d] * 0.5);
}

static double rotational_energy(orc_system *s) {
  const double coef = 0.25 / s->s;
  double rot[3] = {0, 0, 0};
  const double *I = s->mt.inertia;
  for (int i = 0; i < s->nmol; i++) {
    const double *q = &s->q[4 * i], *p = &s->p[4 * i];
    double t[4], xi[3];
    P1(q, t); xi[0] = dot4(p, t);
    P2(q, t); xi[1] = dot4(p, t);
    P3(q, t); xi[2] = dot4(p, t);
    for (int d = 0; d < 3; d++) {
      xi[d] *= (1.0 / I[d]) * coef;
      rot[d] += I[d] * xi[d] * xi[d] * 2.0;
    }
  }
  return rot[0] + rot[1] + rot[2];
}

static double barostat_energy(orc_system *s) {
  if (s->pmode == 0) return 0.0;
  if (s->cmode == 0)
    return 0.5 * (s->Pv[0] * s->Pv[0] + s->Pv[1] * s->Pv[1] + s->Pv[2] * s->Pv[2]) / s->W +
           s->P * s->L[0] * s->L[1] * s->L[2];
  if (s->cmode == 1) return 0.5 * s->Pv[2] * s->Pv[2] / s->W + s->P * orc_volume(s);
  /* slit, contract (area barostat) */
  return 0.5 * s->Pv[0] * s->Pv[0] / s->W + s->P * orc_volume(s);
}

void orc_calc_properties(orc_system *s) {
  double tmp[3];
  translational_energy(s, tmp);
  s->trav[0] = tmp[0];
  s->trav[1] = tmp[1];
  s->trav[2] = tmp[2];
  s->tra = tmp[0] + tmp[1] + tmp[2];
  s->rot = rotational_energy(s);
  s->kin = s->tra + s->rot;
  s->Tk = s->kin * u_temp() / (3. * s->nmol);
  s->tsm = 0.5 * s->Ps * s->Ps / s->Q;
  s->tsp = s->gkT * log(s->s);
  s->tst = s->tsm + s->tsp;
  s->bst = barostat_energy(s);
  s->tot = s->pot + s->kin;
  if (s->tmode == 1) s->tot += s->tst;
  if (s->pmode > 0) s->tot += s->bst;
  s->ham = s->tot - s->H0;
  const double volume = orc_volume(s);
  for (int d = 0; d < 3; d++) s->prs[d] = (tmp[d] * 2.0 + s->vir[d]) / volume * u_press();
  s->prs[3] = ((tmp[0] * 2.0 + s->vir[0]) + (tmp[1] * 2.0 + s->vir[1]) + (tmp[2] * 2.0 + s->vir[2])) /
              (3. * volume) * u_press();
}

void orc_set_temperature(orc_system *s, double T) {
  s->T = T;
  s->gkT = 6.0 * (double)s->nmol * T;
}

void orc_initialize_property(orc_system *s) {
  orc_calc_force(s, 1, 0.0);
  s->gkT = 6.0 * (double)s->nmol * s->T;
  orc_calc_properties(s);
  s->H0 = s->tot;
  s->ham = s->tot - s->H0;
}

void orc_update_hamiltonian(orc_system *s) {
  s->H0 = s->tot;
  s->ham = 0.0;
}

void orc_kill_angular_momentum_tube(orc_system *s) {
  double mom = 0., msum = 0.;
  for (int i = 0; i < s->nmol; i++) {
    double r[3], v[3];
    for (int d = 0; d < 3; d++) {
      r[d] = (s->r[3 * i + d] - 0.5) * s->L[d];
      v[d] = s->v[3 * i + d] / (s->L[d] * s->s);
    }
    const double m = s->mt.mass;
    mom += m * (r[0] * v[1] - r[1] * v[0]);
    msum += m;
  }
  mom /= msum;
  for (int i = 0; i < s->nmol; i++) {
    double r[3];
    for (int d = 0; d < 3; d++) r[d] = (s->r[3 * i + d] - 0.5) * s->L[d];
    const double r02 = r[0] * r[0] + r[1] * r[1];
    if (r02 > 1.0) {
      s->v[3 * i + 0] += mom * r[1] / r02 * s->L[0] * s->s;
      s->v[3 * i + 1] -= mom * r[0] / r02 * s->L[1] * s->s;
    }
  }
}

void orc_kill_momentum(orc_system *s) {
  double sum[3] = {0, 0, 0}, m = 0.;
  for (int i = 0; i < s->nmol; i++) {
    for (int d = 0; d < 3; d++) sum[d] += s->v[3 * i + d] * s->mt.mass;
    m += s->mt.mass;
  }
  for (int d = 0; d < 3; d++) sum[d] /= m;
  if (s->cmode == 1) {
    for (int i = 0; i < s->nmol; i++) s->v[3 * i + 2] -= sum[2];
    orc_kill_angular_momentum_tube(s);
  } else if (s->cmode == 2) {
    for (int i = 0; i < s->nmol; i++) {
      s->v[3 * i + 0] -= sum[0];
      s->v[3 * i + 1] -= sum[1];
    }
  } else {
    for (int i = 0; i < s->nmol; i++)
      for (int d = 0; d < 3; d++) s->v[3 * i + d] -= sum[d];
  }
}

void orc_velocity_scaling(orc_system *s) {
  double tmp[3];
  translational_energy(s, tmp);
  const double tra = tmp[0] + tmp[1] + tmp[2];
  if (tra == 0.) return;
  const double coef = sqrt(1.5 * s->nmol * s->T / tra);
  for (int i = 0; i < 3 * s->nmol; i++) s->v[i] *= coef;
  orc_kill_momentum(s);
  if (s->cmode == 1) orc_kill_angular_momentum_tube(s);
  s->s = 1.0;
  s->Ps = 0.0;
  s->Pv[0] = s->Pv[1] = s->Pv[2] = 0.0;
}

void orc_ang_velocity_scaling(orc_system *s) {
  const double rot = rotational_energy(s);
  if (rot == 0.) return;
  const double coef = sqrt(1.5 * s->nmol * s->T / rot);
  for (int i = 0; i < 4 * s->nmol; i++) s->p[i] *= coef;
}

void orc_change_temperature(orc_system *s, double T) {
  const double coef = sqrt(T / s->T);
  for (int i = 0; i < s->nmol; i++) {
    for (int d = 0; d < 3; d++) s->v[3 * i + d] *= coef / s->s;
    for (int d = 0; d < 4; d++) s->p[4 * i + d] *= coef / s->s;
  }
  s->Ps *= coef;
  if (s->pmode == 1) s->Pv[0] *= coef * s->s;
  if (s->pmode == 2) s->Pv[2] *= coef * s->s;
  if (s->pmode == 3) s->Pv[0] *= coef * s->s; /* area momentum */
  orc_set_temperature(s, T);
  s->s = 1.0;
}

void orc_execute_steps(orc_system *s, int nsteps) {
  for (int st = 0; st < nsteps; st++) {
    if (s->tmode == 2) {
      orc_velocity_scaling(s);
      orc_ang_velocity_scaling(s);
    }
    const int do_sort = (st % s->sort_interval) == 0;
    orc_step(s, do_sort, s->margin);
  }
  orc_calc_properties(s);
}

/* Average::sum in enum order (src/remdinfo.h:27-41) as filled by
 * Molecules::GetAverage (src/molecules.h:380-402). */
void orc_get_average(const orc_system *s, double out[19]) {
  out[0] = s->lj;
  out[1] = s->clmb;
  out[2] = s->wall;
  out[3] = orc_volume(s);
  out[4] = s->tra;
  out[5] = s->rot;
  out[6] = s->Tk;
  out[7] = s->prs[3];
  out[8] = s->prs[0];
  out[9] = s->prs[1];
  out[10] = s->prs[2];
  out[11] = s->tot;
  out[12] = s->ham;
  out[13] = s->s;
  out[14] = s->Ps;
  out[15] = s->Pv[0];
  out[16] = s->Pv[1];
  out[17] = s->Pv[2];
  out[18] = 0.0;
}

/* ----------------------------------------------------------- exchange */
/* ExchangeList::ExchangeList (src/remdinfo.cxx:54-108) */
int orc_exchange_list(int dim_temp, int dim_press, int type, int *a, int *b) {
  int c = 0;
  if (type == 0) {
    for (int j = 0; j < dim_press; ++j)
      for (int i = 0; i < dim_temp / 2; ++i) {
        a[c] = 2 * i + j * dim_temp;
        b[c] = 2 * i + 1 + j * dim_temp;
        c++;
      }
  } else if (type == 1) {
    for (int j = 0; j < dim_press / 2; ++j)
      for (int i = 0; i < dim_temp; ++i) {
        a[c] = i + 2 * j * dim_temp;
        b[c] = i + (2 * j + 1) * dim_temp;
        c++;
      }
  } else if (type == 2) {
    for (int j = 0; j < dim_press; ++j)
      for (int i = 0; i < dim_temp / 2 - (dim_temp + 1) % 2; ++i) {
        a[c] = 2 * i + j * dim_temp + 1;
        b[c] = 2 * i + 1 + j * dim_temp + 1;
        c++;
      }
  } else {
    for (int j = 0; j < dim_press / 2 - (dim_press + 1) % 2; ++j)
      for (int i = 0; i < dim_temp; ++i) {
        a[c] = i + (2 * j + 1) * dim_temp;
        b[c] = i + (2 * j + 2) * dim_temp;
        c++;
      }
  }
  return c;
}

/* REMDInfo::SetConditionGeometrical (src/remdinfo.cxx:348-359): powf. */
void orc_ladder_geometric(int dim_temp, int dim_press, double tmin, double tmax,
                          double pmin, double pmax, double *T, double *P) {
  const double t = (dim_temp == 1) ? 1.0 : powf((float)(tmax / tmin), (float)(1. / (double)(dim_temp - 1)));
  const double p = (dim_press == 1) ? 0.0 : ((pmax - pmin) / (double)(dim_press - 1));
  for (int j = 0; j < dim_press; ++j)
    for (int i = 0; i < dim_temp; ++i) {
      const int ind = i + j * dim_temp;
      T[ind] = tmin * powf((float)t, (float)i);
      P[ind] = pmin + p * j;
    }
}

/* REMD::ReplicaExchange + CalcExchangeProb (src/remd.cxx:464-538) with
 * REMDInfo::ExchangeAccepted/Rejected (src/remdinfo.h:276-359). */
int orc_replica_exchange(int ntemp, int npress, long step, int pmode,
                         const double *pot, const double *vol, int *index,
                         double *T, double *P, unsigned char *is_exchanged,
                         const double *uniforms, int *accepted) {
  const int nrep = ntemp * npress;
  int *la = malloc(sizeof(int) * (size_t)(nrep + 1));
  int *lb = malloc(sizeof(int) * (size_t)(nrep + 1));
  for (int i = 0; i < nrep; i++) is_exchanged[i] = 0;
  const int ec_type = (int)(step % 4);
  const int len = orc_exchange_list(ntemp, npress, ec_type, la, lb);
  int nacc = 0;
  for (int k = 0; k < len; k++) {
    const int ind1 = index[la[k]];
    const int ind2 = index[lb[k]];
    const double beta1 = 1. / T[ind1];
    const double beta2 = 1. / T[ind2];
    double delta = (beta2 - beta1) * (pot[ind1] - pot[ind2]);
    if (pmode > 0) delta += (beta2 * P[ind2] - beta1 * P[ind1]) * (vol[ind1] - vol[ind2]);
    const double prob = exp(-delta);
    const double u = uniforms ? uniforms[k] : ((double)rand() / (double)RAND_MAX);
    const int acc = prob > u;
    if (accepted) accepted[k] = acc;
    if (acc) {
      is_exchanged[index[la[k]]] = 1;
      is_exchanged[index[lb[k]]] = 1;
      double t = T[index[la[k]]];
      T[index[la[k]]] = T[index[lb[k]]];
      T[index[lb[k]]] = t;
      t = P[index[la[k]]];
      P[index[la[k]]] = P[index[lb[k]]];
      P[index[lb[k]]] = t;
      const int it = index[la[k]];
      index[la[k]] = index[lb[k]];
      index[lb[k]] = it;
      nacc++;
    } else {
      is_exchanged[index[la[k]]] = 0;
      is_exchanged[index[lb[k]]] = 0;
    }
  }
  free(la);
  free(lb);
  return nacc;
}
