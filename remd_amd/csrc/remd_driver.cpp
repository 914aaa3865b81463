// remd_driver.cpp -- `remd_md x.inp`: the REMD driver of the reference's
// md.out (src/main.cxx, src/remd.cxx) over the C ABI of libremd_gpu.
//
// Reads the same input files -- the Parameter key/value file
// (src/parameter.h:170-247), the molecule types (src/molecules.cxx:69-131)
// and one checkpoint per replica (src/io_manager.cxx:528-591) -- builds the
// (T, P) ladder (src/remdinfo.cxx:140-266, 348-398), applies the start-up
// velocity scaling (src/remd.cxx:102-110), and runs the interval loop of
// REMD::ExecuteREMD (src/remd.cxx:274-466): one remd_gpu_execute_steps per
// interval, the " rep" lines of the GPU path, the per-condition binary logs
// (src/remd.cxx:150-193, 891-931), the replica exchange with the glibc
// rand() stream (src/remd.cxx:464-538), CDV snapshots and checkpoints.
// Not supported (the program stops with a message): gen_vel, adjust_center,
// cond_mode 1/4 (heat-capacity ladders), rem_type 1 (designed walk), the
// condition adjustor, MPI.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "remd_gpu.h"

namespace {

// src/unit.h:3-18
const double kNa = 6.0221367e+23;
const double kE0 = 8.8542e-12;
const double kEle = 1.6021761e-19;
const double kKB = 1.3806488e-23;
const double kUnitMass = 15.9994e-3 / kNa;
const double kUnitLength = 1.0e-10;
const double kUnitEnergy = 1e3 / kNa;
const double kUnitTime = std::sqrt(kUnitMass * kUnitLength * kUnitLength / kUnitEnergy);
const double kUnitCoulomb = std::sqrt(4. * M_PI * kE0 * kUnitEnergy * kUnitLength);
const double kUnitTemp = kUnitEnergy / kKB;
const double kUnitPress = kUnitEnergy / (kUnitLength * kUnitLength * kUnitLength);

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "error: %s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}

void check(int rc, const char *what) {
  if (rc < 0) die(std::string(what) + ": " + remd_gpu_last_error());
}

// Parameter (src/parameter.h:24-88) -- the keys this path reads.
struct Param {
  std::string input_prefix, output_prefix, chk_in = "null", chk_out = "null", cdv_out = "null",
      mtl_in = "moltype.txt";
  int nmol = 0, natom = 0, nsite = 4;
  double dt = 2.0, rcut = 10.0, rswitch = 8.0;
  int tconstant = 0, pconstant = 0;
  double tstat_mass = -1.0, bstat_mass = -1.0;
  long ninterval = 0, interval = 0, energy_interval = 0, snap_interval = 0, chk_interval = 0;
  int ngpus = 0, nreplica = 1, ntemp = 1, npress = 1;
  double temp_max = 300.0, temp_min = 300.0, press_max = 1.0, press_min = 1.0;
  int gen_vel = 0, cond_mode = 0, restart = 0, adjust_center = 0, init_scaling = 0, rem_type = 0;
  int confined = 0, nfwall = 0;
  double wall_length = 0.0, sgm_wall = 3.19, eps_wall = 0.3135, rho_wall = 0.3816;
  int adjustor = 0;
};

Param read_param(const std::string &path) {
  std::ifstream is(path);
  if (!is) die("open " + path + " failed");
  Param p;
  for (std::string line; std::getline(is, line);) {
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag.empty() || !std::isalpha((unsigned char)tag[0])) continue;
#define KEY(name)          \
  if (tag == #name) {      \
    ss >> p.name;          \
    continue;              \
  }
    KEY(input_prefix) KEY(output_prefix) KEY(chk_in) KEY(chk_out) KEY(cdv_out) KEY(mtl_in)
    KEY(nmol) KEY(natom) KEY(nsite) KEY(dt) KEY(rcut) KEY(rswitch)
    KEY(tconstant) KEY(tstat_mass) KEY(pconstant) KEY(bstat_mass)
    KEY(ninterval) KEY(interval) KEY(energy_interval) KEY(snap_interval) KEY(chk_interval)
    KEY(ngpus) KEY(nreplica) KEY(ntemp) KEY(npress)
    KEY(temp_max) KEY(temp_min) KEY(press_max) KEY(press_min)
    KEY(gen_vel) KEY(cond_mode) KEY(restart) KEY(adjust_center) KEY(init_scaling) KEY(rem_type)
    KEY(confined) KEY(nfwall) KEY(wall_length) KEY(sgm_wall) KEY(eps_wall) KEY(rho_wall) KEY(adjustor)
#undef KEY
  }
  return p;
}

// First block of a moltype file (src/molecules.cxx:69-131), reduced units.
struct MolType {
  int nsite = 0;
  double r[5][3], m[5], sigma[5], eps[5], q[5];
  double mass = 0, inertia[3] = {0, 0, 0};
};

MolType read_moltype(const std::string &path) {
  std::ifstream is(path);
  if (!is) die("open " + path + " failed");
  std::vector<std::string> lines;
  for (std::string l; std::getline(is, l);) lines.push_back(l);
  MolType mt;
  for (size_t i = 0; i < lines.size(); i++) {
    std::istringstream ss(lines[i]);
    std::string name;
    int nat = 0;
    if (!(ss >> name) || name.empty() || !std::isalpha((unsigned char)name[0]) || !(ss >> nat)) continue;
    if (nat < 3 || nat > 5 || i + nat >= lines.size() + 0) die("bad molecule block in " + path);
    mt.nsite = nat;
    for (int a = 0; a < nat; a++) {
      std::istringstream as(lines[i + 1 + a]);
      std::string an;
      double x, y, z, m, s, e, q;
      if (!(as >> an >> x >> y >> z >> m >> s >> e >> q)) die("bad site line in " + path);
      mt.r[a][0] = x * (1e-10 / kUnitLength);
      mt.r[a][1] = y * (1e-10 / kUnitLength);
      mt.r[a][2] = z * (1e-10 / kUnitLength);
      mt.m[a] = m * (1e-3 / kNa / kUnitMass);
      mt.sigma[a] = s * (1e-10 / kUnitLength);
      mt.eps[a] = e * (1e3 / kNa / kUnitEnergy);
      mt.q[a] = q * (kEle / kUnitCoulomb);
    }
    // MolType::CalcIner (src/variables.h:189-206)
    for (int a = 0; a < nat; a++) {
      const double mm = mt.m[a], x = mt.r[a][0], y = mt.r[a][1], z = mt.r[a][2];
      mt.mass += mm;
      mt.inertia[0] += mm * (y * y + z * z);
      mt.inertia[1] += mm * (z * z + x * x);
      mt.inertia[2] += mm * (x * x + y * y);
    }
    return mt;
  }
  die("no molecule block in " + path);
}

// IOManager::ReadCheckPoint (src/io_manager.cxx:528-591): internal
// representation r / L, v * L * s, p * s.
struct Replica {
  std::vector<double> r, v, q, p;
  remd_gpu_scalars sc;
};

Replica read_chk(const std::string &path, int nmol) {
  std::ifstream is(path);
  if (!is) die("open " + path + " failed");
  int n = 0, natom = 0;
  is >> n >> natom;
  if (n != nmol) die(path + ": nmol " + std::to_string(n) + " != " + std::to_string(nmol));
  Replica R;
  R.r.resize(3 * n);
  R.v.resize(3 * n);
  R.q.resize(4 * n);
  R.p.resize(4 * n);
  for (int i = 0; i < n; i++) {
    double m, I[3];
    int type, id;
    for (int d = 0; d < 3; d++) is >> R.r[3 * i + d];
    for (int d = 0; d < 3; d++) is >> R.v[3 * i + d];
    for (int d = 0; d < 4; d++) is >> R.q[4 * i + d];
    for (int d = 0; d < 4; d++) is >> R.p[4 * i + d];
    is >> m >> I[0] >> I[1] >> I[2] >> type >> id;
  }
  remd_gpu_scalars &s = R.sc;
  is >> s.s >> s.Ps >> s.Q >> s.Pv[0] >> s.Pv[1] >> s.Pv[2] >> s.W >> s.L[0] >> s.L[1] >> s.L[2];
  if (!is) die(path + ": truncated checkpoint");
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      R.r[3 * i + d] /= s.L[d];
      R.v[3 * i + d] *= s.L[d] * s.s;
    }
    for (int d = 0; d < 4; d++) R.p[4 * i + d] *= s.s;
  }
  return R;
}

// IOManager::WriteCheckPoint (src/io_manager.cxx:593-634).
void write_chk(const std::string &path, const MolType &mt, const std::vector<double> &r,
               const std::vector<double> &v, const std::vector<double> &q, const std::vector<double> &p,
               const std::vector<int> &ids, const remd_gpu_scalars &s) {
  const int n = (int)ids.size();
  std::vector<int> row(n);
  for (int k = 0; k < n; k++) row[ids[k]] = k;  // original molecule order
  std::ofstream os(path);
  if (!os) die("open " + path + " failed");
  os << std::setprecision(16);
  os << n << " " << n * mt.nsite << "\n";
  for (int i = 0; i < n; i++) {
    const int k = row[i];
    for (int d = 0; d < 3; d++) os << " " << r[3 * k + d] * s.L[d];
    for (int d = 0; d < 3; d++) os << " " << v[3 * k + d] / (s.L[d] * s.s);
    for (int d = 0; d < 4; d++) os << " " << q[4 * k + d];
    for (int d = 0; d < 4; d++) os << " " << p[4 * k + d] / s.s;
    os << " " << mt.mass << " " << mt.inertia[0] << " " << mt.inertia[1] << " " << mt.inertia[2] << " 0 "
       << i * mt.nsite << "\n";
  }
  os << s.s << " " << s.Ps << " " << s.Q << "\n";
  os << s.Pv[0] << " " << s.Pv[1] << " " << s.Pv[2] << " " << s.W << "\n";
  os << s.L[0] << " " << s.L[1] << " " << s.L[2] << "\n";
  os << "0\n";
}

// KillMomentum + KillAngularMomentumInTube (src/molecules.cxx:168-210).
void kill_momentum(Replica &R, const MolType &mt, int confined, int n) {
  double sum[3] = {0, 0, 0}, msum = 0;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) sum[d] += R.v[3 * i + d] * mt.mass;
    msum += mt.mass;
  }
  for (int d = 0; d < 3; d++) sum[d] /= msum;
  for (int i = 0; i < n; i++) {
    if (confined == 1) {
      R.v[3 * i + 2] -= sum[2];
    } else if (confined == 2) {
      R.v[3 * i + 0] -= sum[0];
      R.v[3 * i + 1] -= sum[1];
    } else {
      for (int d = 0; d < 3; d++) R.v[3 * i + d] -= sum[d];
    }
  }
  if (confined != 1) return;
  const double *L = R.sc.L;
  const double s = R.sc.s;
  double mom = 0., ms = 0.;
  for (int i = 0; i < n; i++) {
    const double rx = (R.r[3 * i] - 0.5) * L[0], ry = (R.r[3 * i + 1] - 0.5) * L[1];
    const double vx = R.v[3 * i] / (L[0] * s), vy = R.v[3 * i + 1] / (L[1] * s);
    mom += mt.mass * (rx * vy - ry * vx);
    ms += mt.mass;
  }
  mom /= ms;
  for (int i = 0; i < n; i++) {
    const double rx = (R.r[3 * i] - 0.5) * L[0], ry = (R.r[3 * i + 1] - 0.5) * L[1];
    const double r02 = rx * rx + ry * ry;
    if (r02 > 1.0) {
      R.v[3 * i] += mom * ry / r02 * L[0] * s;
      R.v[3 * i + 1] -= mom * rx / r02 * L[1] * s;
    }
  }
}

// VelocityScaling + AngVelocityScaling (src/molecules.cxx:261-311).
void velocity_scaling(Replica &R, const MolType &mt, int confined, int n, double T) {
  double tra = 0;
  for (int d = 0; d < 3; d++) {
    double acc = 0;
    for (int i = 0; i < n; i++) acc += R.v[3 * i + d] * R.v[3 * i + d] * mt.mass;
    const double cf = 1.0 / (R.sc.L[d] * R.sc.s);
    tra += acc * (cf * cf * 0.5);
  }
  if (tra != 0.0) {
    const double coef = std::sqrt(1.5 * n * T / tra);
    for (double &x : R.v) x *= coef;
    kill_momentum(R, mt, confined, n);
    if (confined == 1) kill_momentum(R, mt, confined, n);
    R.sc.s = 1.0;
    R.sc.Ps = 0.0;
    R.sc.Pv[0] = R.sc.Pv[1] = R.sc.Pv[2] = 0.0;
  }
  double rot = 0;
  for (int i = 0; i < n; i++) {
    const double *q = &R.q[4 * i], *p = &R.p[4 * i];
    const double cf = 0.25 / R.sc.s;
    const double x0 = (-p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2]) * ((1.0 / mt.inertia[0]) * cf);
    const double x1 = (-p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1]) * ((1.0 / mt.inertia[1]) * cf);
    const double x2 = (-p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0]) * ((1.0 / mt.inertia[2]) * cf);
    rot += mt.inertia[0] * x0 * x0 * 2.0 + mt.inertia[1] * x1 * x1 * 2.0 + mt.inertia[2] * x2 * x2 * 2.0;
  }
  if (rot != 0.0) {
    const double coef = std::sqrt(1.5 * n * T / rot);
    for (double &x : R.p) x *= coef;
  }
}

// Run state kept across a restart (the reference keeps none of it: its
// step counter, exchange index, rand() stream and acceptance/tunnel counts
// start afresh, src/remd.cxx:25-198): the next interval number, the
// condition-slot -> replica index, the replica conditions, the exchange
// stream, acceptance per pair of each exchange list
// (REMDInfo::PairList, src/remdinfo.h:141-189), the tunnel counters
// (src/remdinfo.h:191-218) and the job's exact device state
// (remd_gpu_checkpoint_save).
struct RunState {
  long step = 0;
  std::vector<int> index;
  std::vector<double> T, P;
  remd_rng rng;
  std::vector<long> acc[4], rej[4];
  std::vector<int> tc_count;
  std::vector<unsigned char> tc_max, tc_min;
};

const char kRunMagic[8] = {'R', 'E', 'M', 'D', 'R', 'U', 'N', '1'};

template <class T>
void put(std::ofstream &os, const T *p, size_t n) {
  os.write((const char *)p, sizeof(T) * n);
}
template <class T>
void get(std::ifstream &is, T *p, size_t n, const std::string &path) {
  is.read((char *)p, sizeof(T) * n);
  if (!is) die(path + ": truncated run state");
}

void write_run_state(const std::string &path, const RunState &st, remd_gpu *g) {
  const size_t nb = remd_gpu_checkpoint_bytes(g);
  std::vector<char> blob(nb);
  check(remd_gpu_checkpoint_save(g, blob.data(), nb), "checkpoint");
  const std::string tmp = path + ".tmp";
  {
    std::ofstream os(tmp, std::ios::binary);
    if (!os) die("open " + tmp + " failed");
    os.write(kRunMagic, 8);
    const int nrep = (int)st.index.size();
    put(os, &st.step, 1);
    put(os, &nrep, 1);
    put(os, st.index.data(), nrep);
    put(os, st.T.data(), nrep);
    put(os, st.P.data(), nrep);
    put(os, &st.rng, 1);
    for (int t = 0; t < 4; t++) {
      const int n = (int)st.acc[t].size();
      put(os, &n, 1);
      put(os, st.acc[t].data(), n);
      put(os, st.rej[t].data(), n);
    }
    put(os, st.tc_count.data(), nrep);
    put(os, st.tc_max.data(), nrep);
    put(os, st.tc_min.data(), nrep);
    put(os, &nb, 1);
    os.write(blob.data(), (std::streamsize)nb);
    if (!os) die("write " + tmp + " failed");
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) die("rename " + tmp + " failed");
}

void read_run_state(const std::string &path, RunState &st, remd_gpu *g) {
  std::ifstream is(path, std::ios::binary);
  if (!is) die("open " + path + " failed");
  char magic[8];
  get(is, magic, 8, path);
  if (std::memcmp(magic, kRunMagic, 8) != 0) die(path + ": not a remd_md run state");
  int nrep = 0;
  get(is, &st.step, 1, path);
  get(is, &nrep, 1, path);
  if (nrep != (int)st.index.size()) die(path + ": nreplica differs from the input");
  get(is, st.index.data(), nrep, path);
  get(is, st.T.data(), nrep, path);
  get(is, st.P.data(), nrep, path);
  get(is, &st.rng, 1, path);
  for (int t = 0; t < 4; t++) {
    int n = 0;
    get(is, &n, 1, path);
    if (n != (int)st.acc[t].size()) die(path + ": exchange lists differ from the input");
    get(is, st.acc[t].data(), n, path);
    get(is, st.rej[t].data(), n, path);
  }
  get(is, st.tc_count.data(), nrep, path);
  get(is, st.tc_max.data(), nrep, path);
  get(is, st.tc_min.data(), nrep, path);
  size_t nb = 0;
  get(is, &nb, 1, path);
  std::vector<char> blob(nb);
  get(is, blob.data(), nb, path);
  check(remd_gpu_checkpoint_load(g, blob.data(), nb), "checkpoint load");
}

// REMDInfo::OutputAcceptRate (src/remdinfo.cxx:762-820): per pressure row
// the T-even and T-odd pairs interleaved, then per temperature column the
// P-even and P-odd pairs, "Ti Pi Tj Pj rate" with the conditions of the
// pair's slots.
void write_accept_rate(const std::string &path, const RunState &st, int dt, int dp,
                       const std::vector<int> (&pa)[4], const std::vector<int> (&pb)[4]) {
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) die("open " + path + " failed");
  std::fprintf(fp, "#Ti,Pi,Tj,Pj,AR\n");
  auto line = [&](int type, int i) {
    const int ind0 = st.index[pa[type][i]], ind1 = st.index[pb[type][i]];
    const double rate = (double)st.acc[type][i] / (double)(st.acc[type][i] + st.rej[type][i]);
    std::fprintf(fp, "%lf %lf %lf %lf %lf\n", st.T[ind0], st.P[ind0], st.T[ind1], st.P[ind1], rate);
  };
  const int ll[4] = {(int)pa[0].size(), (int)pa[1].size(), (int)pa[2].size(), (int)pa[3].size()};
  for (int p = 0; p < dp; p++) {
    const int i0 = p * ll[0] / dp, i2 = p * ll[2] / dp;
    for (int i = 0; i < ll[0] / dp; i++) {
      line(0, i + i0);
      if (i >= ll[2] / dp) break;
      line(2, i + i2);
    }
    std::fprintf(fp, "\n");
  }
  for (int t = 0; t < dt; t++) {
    const int i1 = t * ll[1] / dt, i3 = t * ll[3] / dt;
    for (int i = 0; i < ll[1] / dt; i++) {
      line(1, i + i1);
      if (i >= ll[3] / dt) break;
      line(3, i + i3);
    }
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

std::string fmt_default(double x) {  // std::ostream default formatting, as in file names
  std::ostringstream s;
  s << x;
  return s.str();
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) die("usage: remd_md input.inp");
  const Param prm = read_param(argv[1]);
  if (prm.gen_vel == 1) die("gen_vel 1 is not supported: start from checkpoints (chk_in)");
  if (prm.adjust_center == 1) die("adjust_center 1 is not supported");
  if (prm.adjustor == 1) die("the condition adjustor is not supported");
  if (prm.rem_type == 1) die("rem_type 1 (designed walk) is not supported");
  if (prm.chk_in == "null") die("chk_in is required (one checkpoint per replica)");
  const int nrep = prm.ntemp * prm.npress;
  if (nrep != prm.nreplica) die("nreplica != ntemp * npress");
  if (prm.confined == 1 && (prm.pconstant == 1 || prm.pconstant == 3)) die("tube needs pconstant 0 or 2");
  if (prm.confined == 2 && (prm.pconstant == 1 || prm.pconstant == 2)) die("slit needs pconstant 0 or 3");

  const MolType mt = read_moltype(prm.input_prefix + prm.mtl_in);
  if (mt.nsite != prm.nsite) die("nsite of the input != sites of the molecule type");

  // ladder on reduced values (src/remdinfo.cxx:168-171, 348-398)
  const double tmax = prm.temp_max / kUnitTemp, tmin = prm.temp_min / kUnitTemp;
  const double pmax = prm.press_max / kUnitPress, pmin = prm.press_min / kUnitPress;
  std::vector<double> T(nrep), P(nrep);
  if (prm.cond_mode == 0 || prm.cond_mode == 2) {
    const double t = (prm.ntemp == 1) ? 1.0 : powf((float)(tmax / tmin), (float)(1. / (double)(prm.ntemp - 1)));
    const double pp = (prm.npress == 1) ? 0.0 : ((pmax - pmin) / (double)(prm.npress - 1));
    for (int j = 0; j < prm.npress; ++j)
      for (int i = 0; i < prm.ntemp; ++i) {
        T[i + j * prm.ntemp] = tmin * powf((float)t, (float)i);
        P[i + j * prm.ntemp] = pmin + pp * j;
      }
    if (prm.cond_mode == 2) {
      const std::string f = prm.input_prefix + "/condition.dat";
      std::ifstream is(f);
      if (!is) die("open " + f + " failed");
      int count = 0;
      for (std::string line; std::getline(is, line) && count < nrep;) {
        std::istringstream ss(line);
        int idx;
        if (ss >> idx >> P[count] >> T[count]) count++;
      }
      if (count != nrep) die("# of conditions in condition.dat != nreplica");
    }
  } else if (prm.cond_mode == 3) {
    const double t = (prm.ntemp == 1) ? 0.0 : ((tmax - tmin) / (double)(prm.ntemp - 1));
    const double pp = (prm.npress == 1) ? 0.0 : ((pmax - pmin) / (double)(prm.npress - 1));
    for (int j = 0; j < prm.npress; ++j)
      for (int i = 0; i < prm.ntemp; ++i) {
        T[i + j * prm.ntemp] = tmin + t * i;
        P[i + j * prm.ntemp] = pmin + pp * j;
      }
  } else {
    die("cond_mode must be 0, 2 or 3");
  }

  // replicas (src/remd.cxx:68-113)
  std::vector<Replica> reps(nrep);
  for (int i = 0; i < nrep; i++) {
    std::ostringstream name;
    name << prm.input_prefix << prm.chk_in << std::setw(4) << std::setfill('0') << i << ".chk";
    reps[i] = read_chk(name.str(), prm.nmol);
    reps[i].sc.T = T[i];
    reps[i].sc.P = P[i];
    if (prm.init_scaling == 1) velocity_scaling(reps[i], mt, prm.confined, prm.nmol, T[i]);
    if (prm.tstat_mass >= 0.0) {
      if (!(prm.tstat_mass > 0.0)) die("tstat_mass must be > 0");
      reps[i].sc.Q = prm.tstat_mass;
      reps[i].sc.s = 1.0;
      reps[i].sc.Ps = 0.0;
    }
  }

  remd_gpu_desc d;
  std::memset(&d, 0, sizeof(d));
  d.nmol = prm.nmol;
  d.nsite = prm.nsite;
  d.nreplica = nrep;
  d.tconstant = prm.tconstant;
  d.pconstant = prm.pconstant;
  d.confined = prm.confined;
  d.interval = (int)prm.interval;
  d.sort_interval = 10;                                  // src/integrator.cxx:729
  d.dt = prm.dt * 1e-15 / kUnitTime;                     // src/molecules.h:92
  d.rcut = prm.rcut;
  d.rswitch = prm.rswitch;
  d.margin = 0.5;                                        // src/molecules.h:305
  d.wall_length = prm.wall_length;
  d.nfwall = prm.nfwall;
  d.sgm_wall = prm.sgm_wall;
  d.eps_wall = prm.eps_wall;
  d.rho_wall = prm.rho_wall;
  for (int a = 0; a < mt.nsite; a++) {
    for (int k = 0; k < 3; k++) d.site_r[a][k] = mt.r[a][k];
    d.site_m[a] = mt.m[a];
    d.site_sigma[a] = mt.sigma[a];
    d.site_eps[a] = mt.eps[a];
    d.site_q[a] = mt.q[a];
  }
  d.mass = mt.mass;
  for (int k = 0; k < 3; k++) d.inertia[k] = mt.inertia[k];
  d.precision = REMD_PREC_MIXED;

  const size_t N = prm.nmol;
  std::vector<double> r(3 * N * nrep), v(3 * N * nrep), q(4 * N * nrep), p(4 * N * nrep);
  std::vector<remd_gpu_scalars> sc(nrep);
  for (int i = 0; i < nrep; i++) {
    std::copy(reps[i].r.begin(), reps[i].r.end(), r.begin() + 3 * N * i);
    std::copy(reps[i].v.begin(), reps[i].v.end(), v.begin() + 3 * N * i);
    std::copy(reps[i].q.begin(), reps[i].q.end(), q.begin() + 4 * N * i);
    std::copy(reps[i].p.begin(), reps[i].p.end(), p.begin() + 4 * N * i);
    sc[i] = reps[i].sc;
  }
  remd_gpu *g = nullptr;
  check(remd_gpu_create(&g, &d, sc.data(), r.data(), v.data(), q.data(), p.data(), nullptr), "create");

  // per-condition binary logs (src/remd.cxx:150-193)
  const std::string out = prm.output_prefix;
  std::vector<FILE *> logs(nrep, nullptr);
  for (int c = 0; c < nrep; c++) {
    const std::string f = out + "/P" + fmt_default(P[c]) + "T" + fmt_default(T[c]) + ".log";
    logs[c] = std::fopen(f.c_str(), prm.restart == 1 ? "ab" : "wb");
    if (!logs[c]) die("open " + f + " failed");
    if (prm.restart != 1) {
      const double cond[3] = {T[c], P[c], 1.0};
      const int nelem = 19;
      std::fwrite(cond, sizeof(double), 3, logs[c]);
      std::fwrite(&nelem, sizeof(int), 1, logs[c]);
    }
  }

  RunState st;
  st.index.resize(nrep);
  for (int i = 0; i < nrep; i++) st.index[i] = i;
  st.T = T;  // conditions indexed by replica
  st.P = P;
  remd_rng_seed(&st.rng, 1);  // rand() without srand()
  std::vector<int> pa[4], pb[4];
  for (int t = 0; t < 4; t++) {
    const int n = remd_exchange_pairs(prm.ntemp, prm.npress, t, nullptr, nullptr);
    check(n, "exchange list");
    pa[t].resize(n);
    pb[t].resize(n);
    remd_exchange_pairs(prm.ntemp, prm.npress, t, pa[t].data(), pb[t].data());
    st.acc[t].assign(n, 0);
    st.rej[t].assign(n, 0);
  }
  st.tc_count.assign(nrep, 0);
  st.tc_max.assign(nrep, 0);
  st.tc_min.assign(nrep, 0);
  // restart: the exact state of the previous run when it left one (written
  // beside its checkpoints), else the text checkpoints alone, as the
  // reference does
  if (prm.restart == 1) {
    const std::string f = prm.input_prefix + prm.chk_in + ".state";
    if (std::ifstream(f).good()) {
      read_run_state(f, st, g);
      std::fprintf(stderr, "restart: continuing from interval %ld (%s)\n", st.step, f.c_str());
    }
  }
  std::vector<int> &index = st.index;
  std::vector<double> &Trep = st.T, &Prep = st.P;
  std::vector<double> pot(nrep), vol(nrep), ave(19 * (size_t)nrep);
  std::vector<int> accepted(nrep + 1);
  const long nint = prm.interval > 0 ? prm.ninterval : 0;
  const long step0 = st.step;
  std::cout << std::setprecision(6);
  for (long s = 0; s < nint; ++s) {
    const long step = st.step;
    check(remd_gpu_execute_steps(g, Trep.data(), Prep.data()), "ExecuteSteps");
    check(remd_gpu_get_pot_vol(g, pot.data(), vol.data()), "GetAverages");
    for (int rep = 0; rep < nrep; rep++) {
      double *a = &ave[19 * (size_t)rep];
      check(remd_gpu_get_averages(g, rep, a), "GetAverages");
      std::cout << " rep " << rep << " " << step << " " << pot[rep] << " " << vol[rep] << " " << a[0] << " "
                << a[1] << " " << a[2] << " " << a[6];
      if (prm.confined == 1) std::cout << " " << a[10];
      else if (prm.confined == 2) std::cout << " " << a[8] << " " << a[9];
      std::cout << " " << a[4] << " " << a[5] << " " << a[13] << " " << a[14];
      if (prm.confined == 1) std::cout << " " << a[17];
      else if (prm.confined == 2) std::cout << " " << a[15] << " " << a[16];
      std::cout << " " << a[12] << " " << a[11] << "\n";
    }
    for (int c = 0; c < nrep; c++) std::fwrite(&ave[19 * (size_t)index[c]], sizeof(double), 19, logs[c]);
    if (prm.rem_type == 0) {
      int np = 0;
      check(remd_exchange_host(prm.ntemp, prm.npress, step, prm.pconstant, pot.data(), vol.data(), index.data(),
                               Trep.data(), Prep.data(), nullptr, &st.rng, accepted.data(), &np),
            "ReplicaExchange");
      const int type = (int)(step % 4);
      for (int k = 0; k < np; k++) (accepted[k] ? st.acc[type][k] : st.rej[type][k])++;
      // REMDInfo::CheckTunneling (src/remdinfo.h:389-392): the replica at the
      // lowest slot reaches the minimum, the one at the highest the maximum
      auto reach = [&](int rep, bool is_max) {
        (is_max ? st.tc_max : st.tc_min)[rep] = 1;
        if (st.tc_max[rep] && st.tc_min[rep]) {
          st.tc_count[rep]++;
          (is_max ? st.tc_min : st.tc_max)[rep] = 0;
        }
      };
      reach(index[0], false);
      reach(index[nrep - 1], true);
    }
    if (prm.snap_interval > 0 && step % prm.snap_interval == 0 && prm.cdv_out != "null")
      check(remd_gpu_write_cdv(g, (out + prm.cdv_out).c_str(), (unsigned)(step / prm.snap_interval)), "OutputCDV");
    st.step = step + 1;
    const bool last = s + 1 == nint;
    if (prm.chk_out != "null" && ((prm.chk_interval > 0 && step % prm.chk_interval == 0) || last)) {
      std::vector<double> rr(3 * N), vv(3 * N), qq(4 * N), pq(4 * N);
      std::vector<int> ids(N);
      for (int rep = 0; rep < nrep; rep++) {
        remd_gpu_scalars ss;
        check(remd_gpu_download_state(g, rep, rr.data(), vv.data(), nullptr, qq.data(), pq.data(), nullptr,
                                      ids.data(), &ss),
              "ConvertVariablesToHost");
        std::ostringstream name;
        name << out << prm.chk_out << std::setw(4) << std::setfill('0') << index[rep] << ".chk";
        write_chk(name.str(), mt, rr, vv, qq, pq, ids, ss);
      }
      write_run_state(out + prm.chk_out + ".state", st, g);
    }
  }
  for (FILE *f : logs) std::fclose(f);
  if (nint > 0) {
    const long step = st.step;
    // index.dat (REMD::OutputBackUp, src/remd.cxx:816-836)
    FILE *fi = std::fopen((out + "/index.dat").c_str(), "w");
    if (fi) {
      for (int i = 0; i < nrep; i++) std::fprintf(fi, "%4d %4d\n", i, index[i]);
      std::fclose(fi);
    }
    // end-of-run outputs of src/remd.cxx:452-458
    write_accept_rate(out + "/accept_rate.dat", st, prm.ntemp, prm.npress, pa, pb);
    // REMDInfo::OutputTunnelCount (src/remdinfo.cxx:822-832)
    FILE *ft = std::fopen((out + "/tunnel_count.dat").c_str(), "w");
    if (!ft) die("open tunnel_count.dat failed");
    for (int rep = 0; rep < nrep; rep++) std::fprintf(ft, " %d", st.tc_count[rep]);
    std::fprintf(ft, "\n");
    std::fclose(ft);
    // REMD::OutputEnergy (src/remd.cxx:838-849): pot (+ P V) per replica
    {
      std::ofstream os(out + "/energy.dat");
      if (!os) die("open energy.dat failed");
      os << std::setw(8) << step;
      for (int i = 0; i < nrep; i++) {
        double e = pot[i];
        if (prm.pconstant > 0) e += vol[i] * Prep[i];
        os << " " << std::setw(16) << std::setprecision(16) << e;
      }
      os << std::endl;
    }
    // REMD::OutputIndex (src/remd.cxx:851-879): appended unless step is 0
    {
      const char *mode = step == 0 ? "w" : "a";
      FILE *fpt = std::fopen((out + "/index_temp.dat").c_str(), mode);
      FILE *fpp = std::fopen((out + "/index_press.dat").c_str(), mode);
      if (!fpt || !fpp) die("open index_temp.dat / index_press.dat failed");
      std::fprintf(fpt, " %7ld", step);
      std::fprintf(fpp, " %7ld", step);
      for (int i = 0; i < nrep; i++) {
        std::fprintf(fpt, " %lf", Trep[i]);
        std::fprintf(fpp, " %lf", Prep[i]);
      }
      std::fprintf(fpt, "\n");
      std::fprintf(fpp, "\n");
      std::fclose(fpt);
      std::fclose(fpp);
    }
  }
  (void)step0;
  remd_gpu_destroy(g);
  return 0;
}
