// tlgk_lbm_gpu -- cloning-sampler driver on one MI355X.
//
// Drop-in for the reference's MPI drivers TLGK_LBM_POPS_PERTURB
// (src/mpi_LBM_GK_pops_perturb.cpp) and, with `order = forcing`,
// TLGK_LBM_FORCING (src/mpi_LBM_GK_forcing.cpp): same positional input
// file, same initial-state files, same outputs (output_TLGK.out,
// rvalues.dat and labels.dat, clone_<j>/data_force.datout; in forcing order
// weights_evolution_<t>, copies_evolution_<t> and
// clone_<j>/evolution_<t>_clone_<j> instead of data_force.datout), with
// every clone on the GPU behind the C ABI of include/lbc.h instead of one
// MPI rank per block of clones.
//
//   tlgk_lbm_gpu <inputfile>
//
// Input: the 11 positional fields of the reference driver
// (src/mpi_LBM_GK_pops_perturb.cpp:70-80)
//     Nc T dT eps L tau U0 alpha path_to_folder master_folder facquVTK
// optionally followed by `key = value` lines:
//     order       pops (default) | forcing   positional order; `forcing` reads
//                 input_TEMPLATE's order Nc T dT dT0 eps L tau U0 alpha path
//                 master (src/mpi_LBM_GK_forcing.cpp:62-80) and runs the
//                 forcing driver: each interval's first floor(dT0*T0) steps
//                 with the body force (1 + eps*map)*beta0, map a Gaussian
//                 random field drawn per clone and interval on the device,
//                 then floor((dT-dT0)*T0) uniform steps; drag scaled by 1/F0
//                 with no shift, no snapshot perturbation (:137-144,333-400)
//     field_sigma forcing: covariance length of the fields (0 = 32)
//     dump_fields forcing: 1 writes each interval's fields to
//                 fields_<t>.datout in the master folder (Nc maps of Dx*Dy
//                 doubles, map[y + Dy*x])
//     mode        ref (default) | channel
//     storage     f64 (default) | f32 | f16 | bf16
//     rng         glibc (default in ref mode: the reference's srand/rand and
//                 drand48 streams) | philox (device draws)
//     ranks       p: reproduce a p-rank run's draw streams and total_R order
//     init        files (default in ref mode) | equilibrium
//     init_amp    velocity noise amplitude of the equilibrium initial state
//     force_io    1 (default) | 0: per-clone force files and VTK frames
//     vtk         1 (default) | 0: with force_io, write each clone's VTK
//                 frames every facquVTK steps of every interval, numbered
//                 t*floor(M/facquVTK) + k as the reference does (:450-455)
//     zero_mean   1 (default) | 0: subtract AVG = 1.775 from F/F0
//     seed        Philox key
//     threads     step-kernel workgroup size (0 = tuned)
//     passes      LBM passes per HBM sweep, 1-3 (0 = tuned)
//     nt          non-temporal population stores: 1 off, 2 on (0 = tuned)
//     flags       the whole lbc_config.flags word (instead of passes / nt)
//   channel mode (no reference counterpart; lattice given explicitly):
//     Lx Ly side cx cy u_in k sigma steps intervals
//   extra instances, batched into the same launches (channel mode):
//     instance = Lx Ly side cx cy tau u_in n_clones steps k sigma
//
// In ref mode with rng = glibc and ranks = 1 the run consumes rand() and
// drand48() exactly as the reference's rank 0 does, and the weights are
// exp(alpha * s_) evaluated on the host with glibc, so the selections and
// copy tables are the reference's.  In forcing order the reference also
// draws its fields from rand() (randNormal.cpp), which this driver does not
// (the fields come from Philox on the device), so only drand48's floor-step
// stream is shared with the reference there.
#include <sys/stat.h>

#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "lbc.h"
#include "vtk_writer.h"

namespace {

constexpr int kNN = 10;         // snapshots mixed by the perturbation (NN)
constexpr double kAVG = 1.775;  // ZERO_MEAN shift (src/mpi_LBM_GK_pops_perturb.cpp:27)

struct Input {
  int Nc = 0;
  double T = 1, dT = 1, dT0 = 0, eps = 1;
  int L = 0;
  double tau = 1, U0 = 1, alpha = 1;
  std::string path, master;
  int facquVTK = 1;
  std::map<std::string, std::string> kv;
  std::vector<std::vector<double>> extra_instances;
};

bool make_dir(const std::string& d) {
  if (mkdir(d.c_str(), 0755) == 0 || errno == EEXIST) return true;
  return false;
}

bool read_input(const char* file, Input& in, std::string& err) {
  std::ifstream f(file);
  if (!f.is_open()) {
    err = "COULD NOT OPEN INPUT FILE";
    return false;
  }
  std::vector<std::string> pos;
  std::string line;
  std::vector<std::string> tail;
  while (std::getline(f, line)) {
    auto eq = line.find('=');
    if (eq != std::string::npos) {
      tail.push_back(line);
      continue;
    }
    std::istringstream ls(line);
    std::string tok;
    while (ls >> tok) pos.push_back(tok);
  }
  for (const auto& t : tail) {
    auto eq = t.find('=');
    std::string k = t.substr(0, eq), v = t.substr(eq + 1);
    auto trim = [](std::string s) {
      size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    k = trim(k);
    v = trim(v);
    if (k == "instance") {
      std::istringstream vs(v);
      std::vector<double> vals;
      double x;
      while (vs >> x) vals.push_back(x);
      if (vals.size() != 11) {
        err = "instance lines need 11 values";
        return false;
      }
      in.extra_instances.push_back(vals);
    } else {
      in.kv[k] = v;
    }
  }
  const bool forcing = in.kv.count("order") && in.kv["order"] == "forcing";
  if (pos.size() < (forcing ? 11u : 10u)) {
    err = "input file has too few positional fields";
    return false;
  }
  size_t i = 0;
  auto D = [&](double& d) { d = std::atof(pos[i++].c_str()); };
  auto I = [&](int& v) { v = std::atoi(pos[i++].c_str()); };
  I(in.Nc);
  D(in.T);
  D(in.dT);
  if (forcing) D(in.dT0);
  D(in.eps);
  I(in.L);
  D(in.tau);
  D(in.U0);
  D(in.alpha);
  in.path = pos[i++];
  in.master = pos[i++];
  if (!forcing && i < pos.size()) I(in.facquVTK);
  return true;
}

std::string kv(const Input& in, const char* k, const char* d) {
  auto it = in.kv.find(k);
  return it == in.kv.end() ? std::string(d) : it->second;
}
double kvd(const Input& in, const char* k, double d) {
  auto it = in.kv.find(k);
  return it == in.kv.end() ? d : std::atof(it->second.c_str());
}
long kvi(const Input& in, const char* k, long d) {
  auto it = in.kv.find(k);
  return it == in.kv.end() ? d : std::atol(it->second.c_str());
}

// Tee of stdout and output_TLGK.out, as the reference writes both.
struct Log {
  std::ofstream f;
  template <class T>
  Log& operator<<(const T& v) {
    std::cout << v;
    f << v;
    return *this;
  }
  Log& operator<<(std::ostream& (*m)(std::ostream&)) {
    std::cout << m;
    f << m;
    return *this;
  }
};

bool read_pops(const std::string& file, size_t n, double* dst) {
  std::ifstream p(file, std::ios::binary);
  if (!p.is_open()) return false;
  p.read(reinterpret_cast<char*>(dst), n * sizeof(double));
  return (size_t)p.gcount() == n * sizeof(double);
}

int check(int rc, lbc_ctx* c, const char* what) {
  if (rc != LBC_OK) {
    std::cerr << what << " failed (" << rc << "): " << lbc_last_error(c)
              << std::endl;
    std::exit(2);
  }
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <inputfile>" << std::endl;
    return 1;
  }
  Input in;
  std::string err;
  Log log;
  log.f.open("output_TLGK.out");
  if (!read_input(argv[1], in, err)) {
    log << err << " " << argv[1] << std::endl;
    return 1;
  }
  std::cout << "All processes sucessfully read input file " << argv[1] << std::endl;

  const bool channel = kv(in, "mode", "ref") == "channel";
  const bool forcing = !channel && kv(in, "order", "pops") == "forcing";
  const bool dump_fields = forcing && kvi(in, "dump_fields", 0) != 0;
  // VTK frames of every clone (src/mpi_LBM_GK_pops_perturb.cpp:404-455);
  // the forcing driver writes none
  const bool vtk = kvi(in, "force_io", 1) != 0 && !forcing &&
                   kvi(in, "vtk", 1) != 0 && in.facquVTK > 0;
  const std::string st = kv(in, "storage", "f64");
  const int storage = st == "f32" ? LBC_STORE_F32
                      : st == "f16" ? LBC_STORE_F16
                      : st == "bf16" ? LBC_STORE_BF16
                                     : LBC_STORE_F64;
  const bool glibc_rng = kv(in, "rng", channel ? "philox" : "glibc") == "glibc";
  const int p = (int)kvi(in, "ranks", 1);
  const bool force_io = kvi(in, "force_io", 1) != 0;
  const bool zero_mean = kvi(in, "zero_mean", 1) != 0;
  const int Nc = in.Nc;
  if (Nc < 1 || p < 1 || Nc % p) {
    log << "Nc must be positive and divisible by ranks" << std::endl;
    return 1;
  }
  const int local_Nc = Nc / p;

  // derived parameters (src/mpi_LBM_GK_pops_perturb.cpp:116-151)
  lbc_instance inst{};
  int nbrTimeSteps = (int)std::floor(in.T / in.dT);
  if (!channel) {
    const int Lx = in.L, Ly = in.L;
    const int Dy = 4 * Ly + 1, Dx = Dy;
    inst.Lx = Dx;
    inst.Ly = Dy;
    inst.xmin = (Dx - 1) / 2;
    inst.xmax = inst.xmin + Lx;
    inst.ymin = (Dy - 1) / 2 - Ly / 2;
    inst.ymax = inst.ymin + Ly;
    const double beta0 = (1. / (Dx - 1)) * ((double)Lx / (Dy - 1)) * in.U0 * in.U0;
    const double T0 = Lx / in.U0;
    const double F0 = (in.U0 * in.U0) * Lx * 0.5;
    inst.beta = beta0;
    inst.force_scale = 1. / F0;
    inst.force_shift = zero_mean ? kAVG : 0.0;
    inst.dt = 1.0 / T0;
    inst.steps = (int)std::floor(in.dT * T0);
    inst.eps = in.eps;
    inst.n_snapshots = kNN;
    if (forcing) {
      // two phases of lbmTimeSteps1 and lbmTimeSteps2 steps, F*oneOvF0
      // (src/mpi_LBM_GK_forcing.cpp:137-144,358,389)
      const int M1 = (int)std::floor(in.dT0 * T0);
      const int M2 = (int)std::floor((in.dT - in.dT0) * T0);
      inst.steps = M1 + M2;
      inst.forced_steps = M1;
      inst.field_eps = in.eps;
      inst.field_sigma = kvd(in, "field_sigma", 0.0);
      inst.force_shift = 0.0;
      inst.eps = 0.0;
      inst.n_snapshots = 0;
    }
    log << "READ input file" << std::endl << "---------------" << std::endl;
    log << "  Nc = " << Nc << std::endl;
    log << "  T = " << in.T << " dT = " << in.dT << std::endl;
    log << "  L = " << Lx << " Dx = " << Dx << " Dy = " << Dy << std::endl;
    log << "  beta0 = " << beta0 << std::endl;
    log << "---------------------" << std::endl << "---------------------" << std::endl;
  } else {
    inst.Lx = (int)kvi(in, "Lx", 200);
    inst.Ly = (int)kvi(in, "Ly", 50);
    const int side = (int)kvi(in, "side", inst.Ly / 5);
    const int cx = (int)kvi(in, "cx", inst.Lx / 4);
    const int cy = (int)kvi(in, "cy", inst.Ly / 2);
    inst.xmin = cx - side / 2;
    inst.xmax = inst.xmin + side;
    inst.ymin = cy - side / 2;
    inst.ymax = inst.ymin + side;
    inst.u_in = kvd(in, "u_in", in.U0);
    inst.force_scale = 1.0;
    inst.force_shift = 0.0;
    inst.dt = 1.0;
    inst.steps = (int)kvi(in, "steps", 100);
    inst.noise_sigma = kvd(in, "sigma", 1e-4);
    nbrTimeSteps = (int)kvi(in, "intervals", nbrTimeSteps);
    log << "READ input file (channel)" << std::endl;
    log << "  Nc = " << Nc << " Lx = " << inst.Lx << " Ly = " << inst.Ly
        << " steps/interval = " << inst.steps << " intervals = " << nbrTimeSteps
        << std::endl;
  }
  inst.tau = in.tau;
  inst.n_clones = Nc;
  if (vtk) inst.macro_every = in.facquVTK;
  inst.bias = channel ? kvd(in, "k", in.alpha) : in.alpha;
  inst.rank_blocks = p;

  std::vector<lbc_instance> insts{inst};
  for (const auto& v : in.extra_instances) {
    lbc_instance e{};
    e.Lx = (int)v[0];
    e.Ly = (int)v[1];
    const int side = (int)v[2];
    e.xmin = (int)v[3] - side / 2;
    e.xmax = e.xmin + side;
    e.ymin = (int)v[4] - side / 2;
    e.ymax = e.ymin + side;
    e.tau = v[5];
    e.u_in = v[6];
    e.n_clones = (int)v[7];
    e.steps = (int)v[8];
    e.bias = v[9];
    e.noise_sigma = v[10];
    e.force_scale = 1.0;
    e.dt = 1.0;
    e.rank_blocks = 1;
    insts.push_back(e);
  }
  if (!channel && insts.size() > 1) {
    log << "extra instances need mode = channel" << std::endl;
    return 1;
  }

  lbc_config cfg{};
  cfg.mode = channel ? LBC_MODE_CHANNEL : LBC_MODE_REF;
  cfg.storage = storage;
  cfg.n_instances = (int)insts.size();
  cfg.device = 0;
  cfg.seed = (uint64_t)kvi(in, "seed", 0);
  cfg.threads = (int)kvi(in, "threads", 0);
  cfg.flags = (int)kvi(in, "flags", kvi(in, "vec", 0));
  if (in.kv.count("passes"))
    cfg.flags = (cfg.flags & ~(3 << 5)) | (int)(kvi(in, "passes", 0) & 3) << 5;
  if (in.kv.count("nt"))
    cfg.flags = (cfg.flags & ~(3 << 3)) | (int)(kvi(in, "nt", 0) & 3) << 3;
  if ((cfg.flags & 7) == 0 && (cfg.flags >> 3)) cfg.flags |= 1;
  cfg.instances = insts.data();
  lbc_ctx* ctx = nullptr;
  check(lbc_create(&cfg, &ctx), nullptr, "lbc_create");

  // output folders (src/mpi_LBM_GK_pops_perturb.cpp:224-249)
  make_dir(in.master);
  std::ofstream labelsFile(in.master + "labels.dat", std::ios::binary);
  std::cout << "creating label file " << in.master << "labels.dat" << std::endl;
  log << "Created parent folder " << in.master << std::endl;
  std::vector<std::string> folder(Nc);
  for (int j = 0; j < Nc; j++) {
    folder[j] = in.master + "clone_" + std::to_string(j) + "/";
    if (force_io) make_dir(folder[j]);
  }

  const size_t Nsite = (size_t)inst.Lx * inst.Ly;
  const size_t Npop = 9 * Nsite;
  const std::string init = kv(in, "init", channel ? "equilibrium" : "files");
  if (init == "files") {
    // initial states and snapshots named in popfiles_list.dat
    // (src/mpi_LBM_GK_pops_perturb.cpp:251-384)
    std::ifstream list(in.path + "popfiles_list.dat");
    if (!list.is_open()) {
      log << "ERROR : COULD NOT OPEN INDEX FILE " << in.path
          << "popfiles_list.dat" << std::endl;
      return 1;
    }
    log << "Index file " << in.path << "popfiles_list.dat successfully opened"
        << std::endl << "-------------" << std::endl;
    std::vector<std::string> names;
    std::string nm;
    while (list >> nm) names.push_back(nm);
    if ((int)names.size() < Nc) {
      log << "ERROR : index file lists fewer files than clones" << std::endl;
      return 1;
    }
    std::vector<double> buf(Npop);
    // with p ranks the master hands the first Nc-local_Nc names to clones
    // local_Nc..Nc-1 and keeps the next local_Nc for clones 0..local_Nc-1
    for (int j = 0; j < Nc; j++) {
      const int idx = (j >= local_Nc) ? j - local_Nc : (Nc - local_Nc) + j;
      log << "    Copy " << j << " init. on " << names[idx] << std::endl;
      if (!read_pops(in.path + names[idx], Npop, buf.data())) {
        log << "ERROR : could not read " << in.path + names[idx] << std::endl;
        return 1;
      }
      check(lbc_load_states(ctx, 0, j, 1, buf.data()), ctx, "lbc_load_states");
    }
    if (!channel && !forcing) {
      if ((int)names.size() < kNN) {
        log << "ERROR WITH READING OF POPULATIONS FOR PERTURB IN MASTER" << std::endl;
        return 1;
      }
      std::vector<double> snaps(kNN * Npop);
      for (int n = 0; n < kNN; n++)
        if (!read_pops(in.path + names[n], Npop, snaps.data() + n * Npop)) {
          log << "ERROR WITH READING OF POPULATIONS FOR PERTURB IN MASTER" << std::endl;
          return 1;
        }
      check(lbc_set_snapshots(ctx, 0, snaps.data()), ctx, "lbc_set_snapshots");
    }
  } else {
    for (int i = 0; i < (int)insts.size(); i++)
      check(lbc_init_equilibrium(ctx, i, kvd(in, "init_amp", 1e-3)), ctx,
            "lbc_init_equilibrium");
  }
  log << "--------------------- " << std::endl << "--------------------- " << std::endl;

  // Draw streams of the reference: every rank calls srand(rank+1) and
  // discards 1000 rand() (:176-178); drand48 is never seeded, so every rank
  // starts from glibc's never-seeded state: X = 0 (the library's static
  // drand48 state is zero-filled and its first use only installs the
  // multiplier 0x5DEECE66D and addend 0xB), whose first draw is 11 / 2^48.
  // Rank 0's streams drive selection.
  srand(1);
  for (int i = 0; i < 1000; i++) (void)rand();
  std::vector<std::array<unsigned short, 3>> d48(p);
  for (auto& s : d48) s = {0, 0, 0};
  std::deque<int32_t> rand_queue;

  const int M = inst.steps;
  std::vector<double> R_record;
  std::vector<double> A((size_t)Nc * M), obs(Nc), logw(Nc), s(Nc), u(Nc);
  std::vector<int32_t> rint(Nc), nb0(Nc), nb1(Nc), pairs(2 * Nc), labels(Nc);
  std::vector<uint8_t> marks(Nc);
  std::vector<double> pref((size_t)Nc * kNN);
  std::vector<double> fields(dump_fields ? (size_t)Nc * Nsite : 0);
  const bool fast = channel && !glibc_rng && !force_io;

  if (fast) {
    // everything on the device: one graph replay per interval
    check(lbc_advance(ctx, nbrTimeSteps), ctx, "lbc_advance");
    std::vector<double> R(nbrTimeSteps);
    int nR = 0;
    check(lbc_get_R_history(ctx, 0, R.data(), nbrTimeSteps, &nR), ctx,
          "lbc_get_R_history");
    R_record.assign(R.begin(), R.begin() + nR);
  }
  for (int t = 0; !fast && t < nbrTimeSteps; t++) {
    log << "This is cloning step " << t << std::endl;
    // perturbation prefactors: each rank's marked clones in order
    // (generatePerturbedState.cpp:15 via :412-416)
    const double* pf = nullptr;
    if (!channel && !forcing && glibc_rng) {
      check(lbc_get_marks(ctx, 0, marks.data()), ctx, "lbc_get_marks");
      for (int j = 0; j < Nc; j++) {
        if (!marks[j]) continue;
        auto& st48 = d48[j / local_Nc];
        for (int n = 0; n < kNN; n++) pref[(size_t)j * kNN + n] = erand48(st48.data());
      }
      pf = pref.data();
    }
    check(lbc_run_interval(ctx, pf), ctx, "lbc_run_interval");
    if (force_io) {
      check(lbc_get_drag(ctx, 0, A.data()), ctx, "lbc_get_drag");
      for (int j = 0; j < Nc; j++) {
        // forcing: one file per clone and interval (src/mpi_LBM_GK_forcing.cpp:325-330)
        std::ofstream o = forcing
            ? std::ofstream(folder[j] + "evolution_" + std::to_string(t) +
                                "_clone_" + std::to_string(j),
                            std::ios::binary)
            : std::ofstream(folder[j] + "data_force.datout",
                            std::ios::binary | std::ios::app);
        o.write(reinterpret_cast<const char*>(&A[(size_t)j * M]), M * sizeof(double));
      }
    }
    if (vtk) {
      // frames of step k*facquVTK of interval t are fluid_t<t*nbrVTK + k>.vtk
      // (nbrVTK = floor(M/facquVTK): with facquVTK > M every interval
      // rewrites frame 0, as the reference does)
      const int E = in.facquVTK;
      const int frames = (M + E - 1) / E;
      const int nbrVTK = M / E;
      const size_t per = (size_t)frames * 3 * Nsite;
      const int chunk = (int)std::max<size_t>(1, ((size_t)1 << 27) / per);
      std::vector<double> mac;
      for (int j0 = 0; j0 < Nc; j0 += chunk) {
        const int n = std::min(chunk, Nc - j0);
        mac.resize(per * n);
        check(lbc_get_macro(ctx, 0, j0, n, mac.data()), ctx, "lbc_get_macro");
        for (int j = j0; j < j0 + n; j++)
          for (int k = 0; k < frames; k++) {
            const double* m = mac.data() + per * (j - j0) + (size_t)k * 3 * Nsite;
            const int Lx = inst.Lx;
            auto plane = [m, Lx, Nsite](int q) {
              return [m, Lx, Nsite, q](int X, int Y) {
                return m[q * Nsite + (size_t)Y * Lx + X];
              };
            };
            if (!write_fluid_vtk(t * nbrVTK + k, inst.Lx, inst.Ly, folder[j],
                                 plane(0), plane(1), plane(2))) {
              log << "ERROR : could not write a VTK frame in " << folder[j]
                  << std::endl;
              return 1;
            }
          }
      }
    }
    if (dump_fields) {
      check(lbc_get_field_maps(ctx, 0, fields.data()), ctx, "lbc_get_field_maps");
      std::ofstream o(in.master + "fields_" + std::to_string(t) + ".datout",
                      std::ios::binary);
      o.write(reinterpret_cast<const char*>(fields.data()),
              fields.size() * sizeof(double));
    }
    check(lbc_get_logweights(ctx, 0, obs.data(), logw.data()), ctx,
          "lbc_get_logweights");
    lbc_selection sel{nb0.data(), nb1.data(), pairs.data(), labels.data(), 0, 0, 0, 0};
    if (glibc_rng) {
      for (int j = 0; j < Nc; j++) s[j] = std::exp(inst.bias * obs[j]);  // :467
      for (int j = 0; j < Nc; j++) u[j] = erand48(d48[0].data());        // :517
      while ((int)rand_queue.size() < Nc) rand_queue.push_back(rand());
      for (int j = 0; j < Nc; j++) rint[j] = rand_queue[j];
      check(lbc_select(ctx, 0, s.data(), u.data(), rint.data(), &sel), ctx,
            "lbc_select");
      for (int i = 0; i < sel.n_rint_used; i++) rand_queue.pop_front();
    } else {
      check(lbc_select(ctx, 0, nullptr, nullptr, nullptr, &sel), ctx, "lbc_select");
      for (int j = 0; j < Nc; j++) s[j] = std::exp(logw[j]);
    }
    for (int i = 1; i < (int)insts.size(); i++)
      check(lbc_select(ctx, i, nullptr, nullptr, nullptr, nullptr), ctx, "lbc_select");
    if (forcing) {
      // weights_evolution_t / copies_evolution_t (src/mpi_LBM_GK_forcing.cpp:302-316,448,461,510)
      std::ofstream wf(in.master + "weights_evolution_" + std::to_string(t), std::ios::binary);
      wf.write(reinterpret_cast<const char*>(s.data()), Nc * sizeof(double));
      std::ofstream cf(in.master + "copies_evolution_" + std::to_string(t), std::ios::binary);
      cf.write(reinterpret_cast<const char*>(nb0.data()), Nc * sizeof(int32_t));
      cf.write(reinterpret_cast<const char*>(nb1.data()), Nc * sizeof(int32_t));
    }
    // report as the master does (:576-631)
    log << "  * Evolution and cloning/pruning - DONE " << std::endl;
    log << "    nbCreatedCopies : ";
    int ss = 0;
    for (int j = 0; j < Nc; j++) {
      log << j << "|" << nb1[j] << " ";
      ss += nb1[j];
    }
    log << std::endl;
    log << "    Total # copies after cloning/pruning : " << ss << std::endl;
    R_record.push_back(sel.total_R);
    log << "  * Creation of communication table - DONE " << std::endl;
    labelsFile.write(reinterpret_cast<const char*>(labels.data()), Nc * sizeof(int32_t));
    labelsFile.flush();
    for (int i = 0; i < sel.n_pairs; i++)
      log << "    " << pairs[2 * i] << "->" << pairs[2 * i + 1] << " | ";
    log << std::endl;
    log << "  * Now doing the " << sel.n_pairs << " communications " << std::endl;
    check(lbc_gather(ctx), ctx, "lbc_gather");
    // R values so far (the reference writes them once at the end, :705-714)
    std::ofstream rv(in.master + "rvalues.dat", std::ios::binary);
    rv.write(reinterpret_cast<const char*>(R_record.data()),
             R_record.size() * sizeof(double));
  }
  {
    std::ofstream rv(in.master + "rvalues.dat", std::ios::binary);
    rv.write(reinterpret_cast<const char*>(R_record.data()),
             R_record.size() * sizeof(double));
  }
  if (kvi(in, "dump_states", 0)) {
    std::vector<double> buf(Npop);
    for (int j = 0; j < Nc; j++) {
      make_dir(folder[j]);
      check(lbc_store_states(ctx, 0, j, 1, buf.data()), ctx, "lbc_store_states");
      std::ofstream o(folder[j] + "pops_final.datout", std::ios::binary);
      o.write(reinterpret_cast<const char*>(buf.data()), Npop * sizeof(double));
    }
  }
  lbc_destroy(ctx);
  return 0;
}
