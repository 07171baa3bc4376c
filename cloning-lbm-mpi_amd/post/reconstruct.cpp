// reconstruct -- ancestral drag trajectories of a cloning run.
//
// Drop-in for post_processing/reconstruction.cpp, with its hard-coded
// parameters (run folder, clone count, run and cloning times) taken from the
// command line:
//
//   reconstruct <run_dir/> <Nc> <n_intervals> <steps_per_interval>
//               [out_dir/] [slot ... | all]
//
// Inputs are the files the cloning driver writes:
//   run_dir/labels.dat                  Nc int32 per cloning step: the parent
//                                       slot of every slot after the step
//   run_dir/clone_<k>/data_force.datout steps_per_interval doubles per
//                                       interval, appended, for slot k
// For every requested slot j (default: slot 0, as the reference program
// does; "all" for every slot) the program walks the genealogy backwards
// from the last interval and writes out_dir/clone_<j>/data_force.datout:
// n_intervals * steps_per_interval doubles, the per-step drag along slot
// j's ancestral line (reconstruction.cpp:59-104).  out_dir defaults to
// "outputFolder/".  Exit status 0 on success, 1 on bad arguments or I/O
// errors (the reference has no error path).
#include <sys/stat.h>

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

bool make_dir(const std::string& d) {
  return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST;
}

int usage() {
  std::cerr << "usage: reconstruct <run_dir/> <Nc> <n_intervals> "
               "<steps_per_interval> [out_dir/] [slot ... | all]\n";
  return 1;
}

bool parse_int(const char* s, long& v) {
  char* end = nullptr;
  errno = 0;
  v = std::strtol(s, &end, 10);
  return errno == 0 && end != s && *end == '\0';
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) return usage();
  std::string dir = argv[1];
  if (!dir.empty() && dir.back() != '/') dir += '/';
  long Nc, nt, ndt;
  if (!parse_int(argv[2], Nc) || !parse_int(argv[3], nt) ||
      !parse_int(argv[4], ndt) || Nc < 1 || nt < 1 || ndt < 1)
    return usage();
  std::string out = "outputFolder/";
  int a = 5;
  if (a < argc) {
    long tmp;
    if (std::strcmp(argv[a], "all") != 0 && !parse_int(argv[a], tmp)) {
      out = argv[a++];
      if (!out.empty() && out.back() != '/') out += '/';
    }
  }
  std::vector<long> slots;
  for (; a < argc; a++) {
    if (std::strcmp(argv[a], "all") == 0) {
      for (long j = 0; j < Nc; j++) slots.push_back(j);
      continue;
    }
    long j;
    if (!parse_int(argv[a], j) || j < 0 || j >= Nc) return usage();
    slots.push_back(j);
  }
  if (slots.empty()) slots.push_back(0);

  // the whole genealogy: record t holds the parents after cloning step t
  std::ifstream lf(dir + "labels.dat", std::ios::binary);
  if (!lf) {
    std::cerr << "reconstruct: cannot open " << dir << "labels.dat\n";
    return 1;
  }
  std::vector<int32_t> labels((size_t)Nc * (nt > 1 ? nt - 1 : 1));
  if (nt > 1) {
    lf.read(reinterpret_cast<char*>(labels.data()),
            (std::streamsize)(sizeof(int32_t) * labels.size()));
    if (!lf) {
      std::cerr << "reconstruct: labels.dat holds fewer than " << nt - 1
                << " cloning steps of " << Nc << " slots\n";
      return 1;
    }
  }

  if (!make_dir(out)) {
    std::cerr << "reconstruct: cannot create " << out << "\n";
    return 1;
  }
  std::vector<double> f((size_t)nt * ndt);
  for (long j : slots) {
    long k = j;  // slot holding the ancestor during interval t
    for (long t = nt - 1; t >= 0; t--) {
      const std::string name =
          dir + "clone_" + std::to_string(k) + "/data_force.datout";
      std::ifstream ff(name, std::ios::binary);
      ff.seekg((std::streamoff)t * ndt * (std::streamoff)sizeof(double),
               std::ios::beg);
      ff.read(reinterpret_cast<char*>(&f[(size_t)t * ndt]),
              (std::streamsize)(ndt * sizeof(double)));
      if (!ff) {
        std::cerr << "reconstruct: " << name << " holds no interval " << t
                  << "\n";
        return 1;
      }
      if (t > 0) {
        k = labels[(size_t)(t - 1) * Nc + k];
        if (k < 0 || k >= Nc) {
          std::cerr << "reconstruct: labels.dat step " << t - 1
                    << " names slot " << k << " outside [0, " << Nc << ")\n";
          return 1;
        }
      }
    }
    const std::string folder = out + "clone_" + std::to_string(j) + "/";
    if (!make_dir(folder)) {
      std::cerr << "reconstruct: cannot create " << folder << "\n";
      return 1;
    }
    std::ofstream o(folder + "data_force.datout", std::ios::binary);
    o.write(reinterpret_cast<const char*>(f.data()),
            (std::streamsize)(sizeof(double) * f.size()));
    if (!o) {
      std::cerr << "reconstruct: cannot write " << folder
                << "data_force.datout\n";
      return 1;
    }
  }
  return 0;
}
