// lbm_warmup -- single-lattice body-force run on one MI355X.
//
// Drop-in for the reference's warm-up program run_bodyforce
// (src/lbm_src/main_bodyforce.cpp): same input file, same outputs, with the
// lattice stepped on the GPU behind the C ABI of include/lbc.h (REF physics,
// FP64: the same bits as the reference).
//
//   lbm_warmup [input.datin]
//
// Input (main_bodyforce.cpp:36-46), whitespace separated:
//     nbOfChunks nbOfTimeSteps L tau Ma folderName inputPopsFileName
//     facquVtk facquForce facquU
// optionally followed by `key = value` lines:
//     snap_every, snap_count  also write the last snap_count states, snap_every
//                 steps apart (the final state last), as folderName/pops_<i>.
//                 datout plus folderName/popfiles_list.dat: the initial-state
//                 and snapshot set of the cloning driver
//     storage     f64 (default) | f32 | f16 | bf16
//     threads, passes  step-kernel shape (0 = tuned)
//
// Outputs, as the reference writes them (main_bodyforce.cpp:66-182):
//     folderName/parameters.datout
//     folderName/data_force.datout  raw stress-integral force on the state
//                                   after step n, for n % facquForce == 0
//     folderName/ux_t.datout        ux of the state entering step n (incl.
//                                   the beta/2rho shift) at the reference's
//                                   probe site, for n % facquU == 0
//     folderName + "fluid_t<i>.vtk" density - 1 and velocity at step
//                                   n = i*facquVtk: the fields the previous
//                                   step computed (the initial state's fields
//                                   at n = 0)
//     folderName/pops.datout        final populations, AoS FP64
// The reference's first frame holds uninitialised memory when it starts from
// a pops file; this program writes that state's fields instead.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "lbc.h"
#include "vtk_writer.h"

namespace {

constexpr int kC[9][2] = {{0, 0}, {1, 0},  {0, 1},   {-1, 0}, {0, -1},
                          {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
constexpr double kW[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                          1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

struct Fields {
  std::vector<double> rho, ux, uy;  // idx(x,y) = y + Dy*x
};

// Macroscopic fields of an AoS state, IDX(x,y,k) = k + 9*(y + Dy*x), in the
// reference's operation order: as the step computes them
// (streamCollCompute.cpp:17-26, with the beta/2rho shift) or as
// initializeFields does (initialize_lattice_arrays.cpp:25-42, without).
void fields_of(const std::vector<double>& f, int Dx, int Dy, double beta,
               bool step, Fields& out) {
  out.rho.assign((size_t)Dx * Dy, 0.0);
  out.ux.assign((size_t)Dx * Dy, 0.0);
  out.uy.assign((size_t)Dx * Dy, 0.0);
  for (int x = 0; x < Dx; x++)
    for (int y = 0; y < Dy; y++) {
      double r = 0.0, u = 0.0, v = 0.0;
      for (int k = 0; k < 9; k++) {
        const double ft = f[k + 9 * ((size_t)y + (size_t)Dy * x)];
        r += ft;
        u += ft * kC[k][0];
        v += ft * kC[k][1];
      }
      const size_t i = (size_t)y + (size_t)Dy * x;
      if (step) {
        u = u / r + 0.5 * beta / r;
        v /= r;
      } else {
        u = u / r;
        v = v / r;
      }
      out.rho[i] = r;
      out.ux[i] = u;
      out.uy[i] = v;
    }
}

// One VTK frame of fields in the reference's idx(x,y) layout.
bool write_vtk(int t, int Dx, int Dy, const Fields& F, const std::string& folder) {
  auto at = [Dy](const std::vector<double>& v) {
    return [&v, Dy](int X, int Y) { return v[(size_t)Y + (size_t)Dy * X]; };
  };
  return write_fluid_vtk(t, Dx, Dy, folder, at(F.rho), at(F.ux), at(F.uy));
}

bool make_dir(const std::string& d) {
  return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST;
}

void check(int rc, lbc_ctx* ctx, const char* what) {
  if (rc != LBC_OK) {
    std::cerr << "lbm_warmup: " << what << " failed (" << rc
              << "): " << lbc_last_error(ctx) << std::endl;
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char** argv) {
  const std::string input_name = argc > 1 ? argv[1] : "input.datin";
  std::ifstream input_file(input_name);
  if (!input_file) {
    std::cerr << "lbm_warmup: cannot open " << input_name << std::endl;
    return 1;
  }
  int nbOfChunks, nbOfTimeSteps, L, facquVtk, facquForce, facquU;
  double tau, Ma;
  std::string folderName, inputPopsFileName;
  input_file >> nbOfChunks >> nbOfTimeSteps >> L >> tau >> Ma >> folderName >>
      inputPopsFileName >> facquVtk >> facquForce >> facquU;
  if (!input_file || nbOfTimeSteps < 1 || L < 2 || facquVtk < 1 ||
      facquForce < 1 || facquU < 1) {
    std::cerr << "lbm_warmup: bad input file " << input_name
              << " (nbOfChunks nbOfTimeSteps L tau Ma folderName "
                 "inputPopsFileName facquVtk facquForce facquU)"
              << std::endl;
    return 1;
  }
  std::map<std::string, std::string> kv;
  for (std::string line; std::getline(input_file, line);) {
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      const size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  auto kvi = [&](const char* k, long d) {
    return kv.count(k) ? std::strtol(kv[k].c_str(), nullptr, 10) : d;
  };

  // derived parameters (main_bodyforce.cpp:49-58)
  const int Ly = L, Lx = L;
  const int Dy = 4 * Ly + 1, Dx = Dy;
  const int xmin = (Dx - 1) / 2, xmax = xmin + Lx;
  const int ymin = (Dy - 1) / 2 - Ly / 2, ymax = ymin + Ly;
  const double ot = 1. / 3;
  const double cs = 1. / std::sqrt(3);
  const double u0 = cs * cs * Ma;
  const double nu = ot * (tau - 0.5);
  const double beta = 8 * nu * u0 / ((Dy - 1) / 2) / ((Dy - 1) / 2);

  if (!make_dir(folderName) || !make_dir(folderName + "/vtk_fluid/")) {
    std::cerr << "lbm_warmup: cannot create " << folderName << std::endl;
    return 1;
  }
  {
    std::ofstream param(folderName + "/parameters.datout");
    param << "Number of timesteps : " << nbOfChunks << "X" << nbOfTimeSteps
          << std::endl;
    param << "L : " << Lx << std::endl;
    param << "Dx : " << Dx << std::endl;
    param << "Dy : " << Dy << std::endl;
    param << "tau : " << tau << std::endl;
    param << "beta : " << beta << std::endl;
  }

  const size_t Npop = (size_t)Dx * Dy * 9;
  std::vector<double> state(Npop);
  if (inputPopsFileName != "0") {
    std::ifstream popFile(inputPopsFileName, std::ios::binary);
    popFile.read(reinterpret_cast<char*>(state.data()), Npop * sizeof(double));
    if (!popFile) {
      std::cerr << "lbm_warmup: cannot read " << Dx << "x" << Dy
                << " populations from " << inputPopsFileName << std::endl;
      return 1;
    }
    std::cout << "Initialized populations taken from " << inputPopsFileName
              << std::endl;
  } else {
    std::cout << "Initializing pops to equilibrium value" << std::endl;
    for (size_t i = 0; i < Npop; i++) state[i] = kW[i % 9];
  }

  lbc_instance inst{};
  inst.Lx = Dx;
  inst.Ly = Dy;
  inst.xmin = xmin;
  inst.xmax = xmax;
  inst.ymin = ymin;
  inst.ymax = ymax;
  inst.n_clones = 1;
  inst.steps = 1;
  inst.tau = tau;
  inst.beta = beta;
  inst.force_scale = 1.0;  // raw computeForceOnSquare
  inst.force_shift = 0.0;
  inst.dt = 1.0;
  inst.rank_blocks = 1;
  const std::string st = kv.count("storage") ? kv["storage"] : "f64";
  const int storage = st == "f32"    ? LBC_STORE_F32
                      : st == "f16"  ? LBC_STORE_F16
                      : st == "bf16" ? LBC_STORE_BF16
                                     : LBC_STORE_F64;
  lbc_config cfg{};
  cfg.mode = LBC_MODE_REF;
  cfg.storage = storage;
  cfg.n_instances = 1;
  cfg.threads = (int)kvi("threads", 0);
  cfg.flags = (int)(kvi("passes", 0) & 3) << 5;
  if (cfg.flags) cfg.flags |= 1;
  cfg.instances = &inst;
  lbc_ctx* ctx = nullptr;
  check(lbc_create(&cfg, &ctx), nullptr, "lbc_create");
  check(lbc_load_states(ctx, 0, 0, 1, state.data()), ctx, "lbc_load_states");

  const long snap_every = kvi("snap_every", 0), snap_count = kvi("snap_count", 0);
  std::vector<std::string> snap_names;
  auto maybe_snapshot = [&](int n, const std::vector<double>& s) {
    if (snap_every < 1 || snap_count < 1) return;
    const long back = nbOfTimeSteps - n;
    if (back % snap_every || back / snap_every >= snap_count) return;
    const std::string name =
        "pops_" + std::to_string(snap_names.size()) + ".datout";
    std::ofstream o(folderName + "/" + name, std::ios::binary);
    o.write(reinterpret_cast<const char*>(s.data()), Npop * sizeof(double));
    snap_names.push_back(name);
  };

  Fields prev, cur;
  fields_of(state, Dx, Dy, beta, false, prev);  // initializeFields
  std::ofstream uxFile(folderName + "/ux_t.datout", std::ios::binary);
  std::ofstream forceFile(folderName + "/data_force.datout", std::ios::binary);
  // the reference samples ux[idx(Dx/4,Dy/4)] through an unparenthesised
  // macro, idx(i,j) = (j+Dy*i), so the index is Dy/4 + Dy*Dx/4 in integer
  // arithmetic: site (x, y) = (index / Dy, index % Dy), not (Dx/4, Dy/4)
  const size_t probe = (size_t)(Dy / 4 + Dy * Dx / 4);
  int tt = 0;
  for (int n = 0; n < nbOfTimeSteps; n++) {
    if (n % facquVtk == 0) {
      if (!write_vtk(tt, Dx, Dy, prev, folderName)) {
        std::cerr << "lbm_warmup: cannot write VTK frame " << tt << std::endl;
        return 1;
      }
      tt++;
    }
    // the state entering step n is needed for its fields (ux probe now, the
    // next frame) and for the snapshot set
    const bool want_ux = n % facquU == 0;
    const bool want_frame = (n + 1) % facquVtk == 0 && n + 1 < nbOfTimeSteps;
    const long back = nbOfTimeSteps - n;
    const bool want_snap = snap_every > 0 && back % snap_every == 0 &&
                           back / snap_every < snap_count;
    if (want_ux || want_frame || want_snap) {
      check(lbc_store_states(ctx, 0, 0, 1, state.data()), ctx, "lbc_store_states");
      fields_of(state, Dx, Dy, beta, true, cur);
      maybe_snapshot(n, state);
    }
    check(lbc_run_interval(ctx, nullptr), ctx, "lbc_run_interval");
    if (n % facquForce == 0) {
      double F = 0.0;
      check(lbc_get_drag(ctx, 0, &F), ctx, "lbc_get_drag");
      forceFile.write(reinterpret_cast<const char*>(&F), sizeof(double));
    }
    if (want_ux)
      uxFile.write(reinterpret_cast<const char*>(&cur.ux[probe]), sizeof(double));
    if (want_frame) prev.rho.swap(cur.rho), prev.ux.swap(cur.ux), prev.uy.swap(cur.uy);
  }
  forceFile.close();
  uxFile.close();
  check(lbc_store_states(ctx, 0, 0, 1, state.data()), ctx, "lbc_store_states");
  maybe_snapshot(nbOfTimeSteps, state);
  {
    std::ofstream pops(folderName + "/pops.datout", std::ios::binary);
    pops.write(reinterpret_cast<const char*>(state.data()), Npop * sizeof(double));
    if (!pops) {
      std::cerr << "lbm_warmup: cannot write " << folderName << "/pops.datout"
                << std::endl;
      return 1;
    }
  }
  if (!snap_names.empty()) {
    std::ofstream list(folderName + "/popfiles_list.dat");
    for (const auto& nm : snap_names) list << nm << "\n";
  }
  lbc_destroy(ctx);
  return 0;
}
