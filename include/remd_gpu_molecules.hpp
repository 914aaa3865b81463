/*
 * remd_gpu_molecules.hpp -- the reference's host classes bound to the C ABI
 * of remd_gpu.h.  Header-only C++ for the reference's own build: include it
 * after the reference's "molecules.h" (class Molecules, src/molecules.h:33-70;
 * Molecule, Thermostat, Barostat, MolType in src/variables.h:89-200; Average
 * in src/remdinfo.h:27-60).  It holds the data mapping the REMDGPU adapter
 * and the Molecules-side hooks need, so each adapter method is one call:
 *
 *   REMDGPU::REMDGPU                 -> remd_amd::create(m, nreplica)
 *   REMDGPU::ConvertVariablesToHost  -> remd_amd::to_host(h, m, nreplica)
 *   REMDGPU::GetAverages             -> remd_amd::averages(h, rep)
 *   Molecules::CalcForcePot          -> remd_amd::calc_force_pot(h, mol, ...)
 *   Molecules::ExecuteSteps          -> remd_amd::execute_steps(h, mol)
 *
 * State crosses in the reference's internal representation (what
 * Molecules keeps between steps): r fractional, v = velocity * L * s,
 * p = quaternion momentum * s.  Rows come back in the device's sorted order
 * with the original molecule index beside them.
 */
#ifndef REMD_GPU_MOLECULES_HPP
#define REMD_GPU_MOLECULES_HPP

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "remd_gpu.h"

namespace remd_amd {

// The reference prints and exits on a GPU error (src/remdgpu.cu:208-212).
inline void check(int rc, const char *what) {
  if (rc != REMD_OK) {
    std::fprintf(stderr, "error: %s: %s\n", what, remd_gpu_last_error());
    std::exit(EXIT_FAILURE);
  }
}

// The job fields REMDGPU::REMDGPU reads from replica 0 (src/remdgpu.cu:178-212).
inline remd_gpu_desc desc_of(const Molecules &m0, unsigned nreplica) {
  const Parameter &p = m0.param;
  remd_gpu_desc d = {};
  d.nmol = m0.nmol;
  d.nsite = p.nsite;
  d.nreplica = (int)nreplica;
  d.tconstant = p.tconstant;
  d.pconstant = p.pconstant;
  d.confined = p.confined;
  d.interval = p.interval;
  d.sort_interval = 10;  // Molecules::ExecuteSteps default (src/molecules.h:305)
  d.dt = m0.dt;
  d.rcut = p.rcut;
  d.rswitch = p.rswitch;
  d.margin = 0.5;        // src/molecules.h:305
  d.wall_length = p.wall_length;
  d.nfwall = p.nfwall;
  d.sgm_wall = p.sgm_wall;
  d.eps_wall = p.eps_wall;
  d.rho_wall = p.rho_wall;
  const MolType &mt = m0.mlist[0];
  for (int a = 0; a < p.nsite; a++) {
    for (int k = 0; k < 3; k++) d.site_r[a][k] = mt.a[a].r[k];
    d.site_m[a] = mt.a[a].m;
    d.site_sigma[a] = mt.a[a].s;
    d.site_eps[a] = mt.a[a].e;
    d.site_q[a] = mt.a[a].q;
  }
  d.mass = m0.mlcl[0].m;
  for (int k = 0; k < 3; k++) d.inertia[k] = mt.i[k];
  d.precision = REMD_PREC_MIXED;  // FP32 pair math, FP64 accumulation
  return d;
}

// Extended variables, box and conditions of one replica (src/variables.h:137-157).
inline remd_gpu_scalars scalars_of(const Molecules &x) {
  remd_gpu_scalars s;
  s.s = x.tst->s;
  s.Ps = x.tst->Ps;
  s.Q = x.tst->Q;
  for (int d = 0; d < 3; d++) {
    s.Pv[d] = x.bst->Pv[d];
    s.L[d] = x.L[d];
  }
  s.W = x.bst->W;
  s.T = x.T;
  s.P = x.P;
  return s;
}

// REMDGPU::REMDGPU: uploads nrep replicas; the device sorts, computes the
// forces and sets H0 as the reference's dt = 0 interval does.
inline remd_gpu *create(Molecules **m, unsigned nrep, int precision = REMD_PREC_MIXED) {
  remd_gpu_desc d = desc_of(*m[0], nrep);
  d.precision = precision;
  const size_t nmol = (size_t)m[0]->nmol;
  std::vector<remd_gpu_scalars> sc(nrep);
  std::vector<double> r(3 * nrep * nmol), v(r.size()), q(4 * nrep * nmol), p(q.size());
  std::vector<int> ids(nrep * nmol);
  for (unsigned rep = 0; rep < nrep; rep++) {
    const Molecules &x = *m[rep];
    sc[rep] = scalars_of(x);
    for (size_t i = 0; i < nmol; i++) {
      const Molecule &a = x.mlcl[i];
      const size_t o = rep * nmol + i;
      for (int k = 0; k < 3; k++) {
        r[3 * o + k] = a.r[k];
        v[3 * o + k] = a.v[k];
      }
      for (int k = 0; k < 4; k++) {
        q[4 * o + k] = a.q[k];
        p[4 * o + k] = a.p[k];
      }
      ids[o] = (int)i;
    }
  }
  remd_gpu *h = nullptr;
  check(remd_gpu_create(&h, &d, sc.data(), r.data(), v.data(), q.data(), p.data(), ids.data()), "REMDGPU");
  return h;
}

// REMDGPU::ConvertVariablesToHost (src/remdgpu.cu:806-854): every replica's
// state, forces and torques into its Molecules, by original molecule index.
inline void to_host(remd_gpu *h, Molecules **m, unsigned nrep) {
  const size_t nmol = (size_t)m[0]->nmol;
  std::vector<double> r(3 * nmol), v(3 * nmol), f(3 * nmol), q(4 * nmol), p(4 * nmol), n(4 * nmol);
  std::vector<int> id(nmol);
  for (unsigned rep = 0; rep < nrep; rep++) {
    remd_gpu_scalars s;
    check(remd_gpu_download_state(h, (int)rep, r.data(), v.data(), f.data(), q.data(), p.data(), n.data(),
                                  id.data(), &s),
          "ConvertVariablesToHost");
    Molecules &x = *m[rep];
    for (size_t k = 0; k < nmol; k++) {
      Molecule &a = x.mlcl[id[k]];
      for (int d = 0; d < 3; d++) {
        a.r[d] = r[3 * k + d];
        a.v[d] = v[3 * k + d];
        a.f[d] = f[3 * k + d];
      }
      for (int d = 0; d < 4; d++) {
        a.q[d] = q[4 * k + d];
        a.p[d] = p[4 * k + d];
        a.n[d] = n[4 * k + d];
      }
    }
    x.tst->s = s.s;
    x.tst->Ps = s.Ps;
    for (int d = 0; d < 3; d++) {
      x.bst->Pv[d] = s.Pv[d];
      x.L[d] = s.L[d];
    }
    x.T = s.T;
    x.P = s.P;
  }
}

// REMDGPU::GetAverages (src/remdgpu.h:208-233), Average enum order.
inline Average averages(remd_gpu *h, unsigned rep) {
  Average a;
  double out[19];
  check(remd_gpu_get_averages(h, (int)rep, out), "GetAverages");
  for (int k = 0; k < Average::NUM_ELEM; k++) a.sum[k] = out[k];
  return a;
}

// Molecules::CalcForcePot (src/molecules.h:290) for a one-replica job that
// holds this replica: sort (with the margin), forces on the device, then
// COM forces, quaternion torques and the potential terms into the host copy.
inline void calc_force_pot(remd_gpu *h, Molecules &x, bool do_sort = true, double margin = 0.0) {
  const size_t nmol = (size_t)x.nmol;
  std::vector<double> f(3 * nmol), n(4 * nmol);
  std::vector<int> id(nmol);
  check(remd_gpu_compute_forces(h, do_sort ? 1 : 0, margin), "CalcForcePot");
  check(remd_gpu_download_state(h, 0, nullptr, nullptr, f.data(), nullptr, nullptr, n.data(), id.data(), nullptr),
        "CalcForcePot");
  for (size_t k = 0; k < nmol; k++) {
    Molecule &a = x.mlcl[id[k]];
    for (int d = 0; d < 3; d++) a.f[d] = f[3 * k + d];
    for (int d = 0; d < 4; d++) a.n[d] = n[4 * k + d];
  }
  const Average av = averages(h, 0);
  x.prop.lj = av.sum[Average::LJ];
  x.prop.clmb = av.sum[Average::CLMB];
  x.prop.wall = av.sum[Average::WALL];
  x.prop.pot = x.prop.lj + x.prop.clmb + x.prop.wall;
}

// Molecules::ExecuteSteps (src/molecules.h:305, src/integrator.cxx:604-803)
// for a one-replica job: `interval` steps at the replica's (T, P), then the
// state back into the host copy.
inline void execute_steps(remd_gpu *h, Molecules &x) {
  double T = x.T, P = x.P;
  check(remd_gpu_execute_steps(h, &T, &P), "ExecuteSteps");
  Molecules *one[1] = {&x};
  to_host(h, one, 1);
}

}  // namespace remd_amd

#endif
