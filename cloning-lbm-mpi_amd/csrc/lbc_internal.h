// Internal definitions shared by the HIP kernels and the host side of the
// C ABI.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbc {

// Instance descriptor as the kernels see it (device memory, one per
// instance).  Populations of clone c, direction q, site (x,y) live at
// buf[par][c*clone_stride + q*plane + y*Lx + x]   (SoA [clone][q][y][x]).
struct InstDev {
  void* buf[2];            // ping-pong population buffers
  long long clone_stride;  // elements per clone = 9*plane
  int plane;               // elements per direction plane = Lx*Ly
  int Lx, Ly, xmin, xmax, ymin, ymax;
  int n_clones, steps;
  int ring_n;              // drag slots: 2*(sx+1) + 2*(sy+1)
  int n_snapshots;
  int rank_blocks;
  int clone_base;          // first global clone index of this instance
  double tau, beta, u_in;
  double inlet_k;          // 1 / (1 - u_in)
  double scale, shift, dt, bias;
  double eps, sigma;
  double* drag;            // [n_clones][steps]
  double* obs;             // [n_clones]  dt*sum A
  double* logw;            // [n_clones]
  double* weights;         // [n_clones]  selection input
  double* u;               // [n_clones]  floor-step uniforms
  int32_t* rint;           // [n_clones]  prune/add draws
  int32_t* nb_before;      // [n_clones]
  int32_t* nb_after;       // [n_clones]
  int32_t* pairs;          // [2*n_clones]
  int32_t* labels;         // [n_clones]  parent of each slot
  int32_t* sel_info;       // [4]: n_pairs, n_rint_used, status, spare
  double* sel_R;           // [2]: total_R, log total_R
  double* R_hist;          // [max_intervals]
  uint8_t* marks;          // [n_clones]
  double* prefactors;      // [n_clones][n_snapshots]
  double* snapshots;       // [n_snapshots][9*plane] FP64 SoA
  double* norms;           // [n_clones][2] REF perturbation norms
  const uint32_t* wave_info;  // two words per block of 64 sites (wave_word,
                              // wave_lanes)
  // random-field body force (forcing mode); forced_steps = 0: off
  int forced_steps;        // leading steps of an interval that use it
  int fmx, fmy;            // circulant embedding (2Lx-1) x (2Ly-1)
  double field_eps;        // beta(x,y) = (1 + field_eps*map(x,y))*beta
  double* field_map;       // [n_clones][plane] fields, site order y*Lx + x
  float* field_map32;      // the same rounded to FP32 (shifted storage only)
  double* field_sqrt_ev;   // [fmx*fmy] square roots of the eigenvalues
  double* field_work;      // [chunk][fmx*fmy] complex (re, im), FFT in place;
                           // FP32 pairs for the shifted storages
  // macroscopic output: rho, ux, uy at steps 0, E, 2E, ... (E = 0: off)
  int macro_every, macro_frames;
  double* macro;           // [n_clones][macro_frames][3][plane]
  int max_intervals;
  int pad_;
};

// Device-resident counters shared by all instances.
struct Counters {
  int parity[64];          // which buffer holds the post-stream state
  int interval[64];        // cloning steps completed per instance
};

constexpr int kMaxInstances = 64;
// Selection without allocation (lbc_select_static): largest population, and
// devices whose scratch address is kept.
constexpr int kStaticMaxClones = 8192;
constexpr int kMaxDevices = 64;

// Work item of the step kernel: one clone of one instance.
struct WorkItem {
  int inst;
  int clone;
};

enum Storage { kF64 = 0, kF32 = 1, kF16 = 2, kBF16 = 3 };

struct Launch {
  const InstDev* insts;
  const WorkItem* items;
  Counters* counters;
  int n_items;
  int threads;
  int ring_max;
  int mode;
  int storage;
  int vec;           // sites per thread per iteration of the step kernel (1)
  int nt;            // non-temporal population stores
  int passes;        // LBM passes per HBM sweep (1, 2 or 3)
  int rch;           // LDS ring sites per direction (fused)
  int forced;        // some instance has forced steps (random-field force)
  // clones per workgroup of the one-pass step kernel (1 or 2); with 2 it
  // runs over items2, pairs of clones of one instance (a lone clone's
  // partner is {inst, -1}: it runs the barriers only)
  int cpw;
  const WorkItem* items2;
  int n_items2;
  hipStream_t stream;
};

// Kernel launchers (lbc_kernels.hip).
hipError_t launch_interval(const Launch& L);
// The step kernel with the random-field body force (lbc_forced.hip; the
// reference mode only).
hipError_t launch_interval_forced(const Launch& L);
// LDS ring sites per direction the fused step kernel needs with `threads`
// threads for lattices of row length up to max_lx in physics mode `mode`,
// and the LDS bytes of the whole launch (0 if it does not fit in one CU's
// 160 KiB).
int fused_ring_sites(int threads, int max_lx, int mode);

// Wave table of an instance: for the block of 64 consecutive sites starting
// at site0 = 64*b, the first site's x (bits 0-14) and y (bits 16-30), and
// whether the block holds an x = 0 or x = Lx-1 site (bit 15) or a wall-row
// or square-box site (bit 31).  Lattices narrower than 64 sites mark every
// block with both.  Coordinates must stay below 2^15.
constexpr uint32_t kWaveCoordMask = 0x7FFFu;
constexpr int kWaveEdgeBit = 15;
constexpr int kWaveSpecialBit = 31;

inline uint32_t wave_word(int Lx, int Ly, int x0, int x1, int y0, int y1,
                          int site0) {
  const int x = site0 % Lx, y = site0 / Lx;
  bool edge = true, special = true;
  if (Lx >= 64) {
    const int xe = x + 63;  // last site's x before wrapping to row y + 1
    const bool wraps = xe >= Lx;
    edge = (x == 0) || (xe >= Lx - 1);
    const int ylast = wraps ? y + 1 : y;
    const bool wall = (y == 0) || (ylast >= Ly - 1);
    const bool box0 = y >= y0 && y <= y1 && x <= x1 && (wraps ? Lx - 1 : xe) >= x0;
    const bool box1 = wraps && ylast >= y0 && ylast <= y1 && xe - Lx >= x0;
    special = wall || box0 || box1;
  }
  return (uint32_t)x | ((uint32_t)edge << kWaveEdgeBit) | ((uint32_t)y << 16) |
         ((uint32_t)special << kWaveSpecialBit);
}

// Second word of a block's wave-table entry: how the step kernel runs the
// block (bits 30-31) and, for lattices at least 64 wide, the lane holding
// the block's x = 0 site (bits 0-6) and the lane holding its x = Lx-1 site
// (bits 8-14), kWaveNoLane where there is none (such a block holds at most
// one of each).
//   kWavePlain:   64 lattice sites, none of them on a wall row, in the
//                 square's box or at an x boundary;
//   kWaveXEdge:   64 lattice sites, x-boundary lanes but no wall-row or box
//                 site;
//   kWaveGeneric: wall-row or box sites, or lanes past the lattice's last
//                 site (Lx >= 64);
//   kWaveNarrow:  Lx < 64: several x-boundary lanes per block, found from
//                 the lanes' coordinates.
constexpr uint32_t kWaveNoLane = 0x7Fu;
constexpr int kWaveKindShift = 30;
constexpr int kWavePlain = 0, kWaveXEdge = 1, kWaveGeneric = 2, kWaveNarrow = 3;

inline uint32_t wave_lanes(int Lx, int Ly, int x0, int x1, int y0, int y1,
                           int plane, int site0) {
  if (Lx < 64) return (uint32_t)kWaveNarrow << kWaveKindShift;
  uint32_t l0 = kWaveNoLane, l1 = kWaveNoLane;
  for (int l = 0; l < 64 && site0 + l < plane; l++) {
    const int x = (site0 + l) % Lx;
    if (x == 0) l0 = (uint32_t)l;
    if (x == Lx - 1) l1 = (uint32_t)l;
  }
  int kind = kWavePlain;
  if (site0 + 64 > plane) {
    kind = kWaveGeneric;
  } else {
    const uint32_t w = wave_word(Lx, Ly, x0, x1, y0, y1, site0);
    if ((w >> kWaveSpecialBit) & 1u)
      kind = kWaveGeneric;
    else if ((w >> kWaveEdgeBit) & 1u)
      kind = kWaveXEdge;
  }
  return l0 | (l1 << 8) | ((uint32_t)kind << kWaveKindShift);
}

// Bounce-back directions of the square ring at (x,y) (the writes of
// computeSquareBounceBack_TEST, src/lbm_src/square.cpp:24-50).
__host__ __device__ inline int square_mask(int x, int y, int x0, int x1,
                                           int y0, int y1) {
  if (x < x0 || x > x1 || y < y0 || y > y1) return 0;
  const bool inx = (x > x0 && x < x1), iny = (y > y0 && y < y1);
  if (inx && iny) return 0;
  if (y == y0 && inx) return (1 << 4) | (1 << 7) | (1 << 8);
  if (y == y1 && inx) return (1 << 2) | (1 << 5) | (1 << 6);
  if (x == x0 && iny) return (1 << 3) | (1 << 6) | (1 << 7);
  if (x == x1 && iny) return (1 << 1) | (1 << 5) | (1 << 8);
  if (x == x0) return (y == y0) ? (1 << 7) : (1 << 6);
  return (y == y1) ? (1 << 5) : (1 << 8);
}

// Wall bounce-back directions (src/lbm_src/domain_noSlipWalls.cpp:10-21).
__host__ __device__ inline int wall_mask(int y, int Ly) {
  int m = 0;
  if (y == 0) m |= (1 << 2) | (1 << 5) | (1 << 6);
  if (y == Ly - 1) m |= (1 << 4) | (1 << 7) | (1 << 8);
  return m;
}

// Grouping of the M+1 passes of an interval into HBM sweeps with P passes
// per sweep (P >= 2): pass n starts a group of P when it is a
// stream-and-collide pass (1 <= n, n+P-1 <= M-1), the group does not
// straddle the last forced step M1 (all its passes collide steps < M1 or
// all >= M1), and no pass of it is a macroscopic-output step (a multiple of
// E > 0); every other pass runs alone.  The step kernel walks the same
// rule.
__host__ __device__ constexpr bool fused_group_at(int n, int M, int P, int M1,
                                                  int E) {
  return n >= 1 && n + P - 1 <= M - 1 && (n + P - 1 < M1 || n >= M1) &&
         (E <= 0 || (n + E - 1) / E * E > n + P - 1);
}

// HBM sweeps of that grouping; each sweep swaps the ping-pong buffers.
constexpr int fused_sweeps(int M, int P, int M1 = 0, int E = 0) {
  int s = 0;
  for (int n = 0; n <= M; s++) n += fused_group_at(n, M, P, M1, E) ? P : 1;
  return s;
}
// LDS bytes of one ring of the multi-pass step kernel: nine direction
// planes of rch slots plus two mirror slots each, elements of es bytes,
// rounded up to 16 bytes.
__host__ __device__ constexpr size_t ring_bytes(int rch, int es) {
  return ((size_t)9 * (size_t)(rch + 2) * (size_t)es + 15) / 16 * 16;
}
size_t fused_lds_bytes(int storage, int ring_max, int rch, int passes);
hipError_t launch_flip(const Launch& L, int n_inst, const int* flip_mask_dev);
// The device address of the selection scratch (kStaticMaxClones doubles) on
// the current device, looked up on the first call per device.
hipError_t static_scratch(void** out);
// One selection of I (a descriptor held by the caller, passed to the kernel
// by value; I.weights must be the current device's scratch) with the
// caller's draws, on `stream`.
hipError_t launch_select_static(const InstDev& I, hipStream_t stream);
hipError_t launch_scgf(const InstDev* insts, int inst, int n, double dT,
                       double* lam, hipStream_t stream);
// Random fields of the forcing mode (lbc_forced.hip), for clones
// [clone0, clone0 + n) of one instance (clone0 even) through its FFT work
// slots 0..(n+1)/2-1, slot k serving clones clone0 + 2k and clone0 + 2k + 1:
// draw fills each slot with complex standard normals (Philox4x32-10 counter
// (pair, even clone, interval, kFieldStream + inst)) times the square-rooted
// eigenvalues; after the caller's in-place 2-D forward FFT, extract writes
// the real part of the lattice block to the even clone's map and the
// imaginary part to the odd one's.  fmn = fmx*fmy.  f32: the work slots
// hold single-precision complex values (a C2C transform; the shifted
// storages), else double (Z2Z; FP64 storage).
constexpr uint32_t kFieldStream = 0xF1E1D000u;
hipError_t launch_field_draw(const InstDev* insts, const Counters* counters,
                             int inst, int clone0, int n, int fmn,
                             uint32_t seed_lo, uint32_t seed_hi, bool f32,
                             hipStream_t stream);
hipError_t launch_field_extract(const InstDev* insts, int inst, int clone0,
                                int n, int plane, bool f32,
                                hipStream_t stream);
// field_map32 from field_map for all clones of an instance (maps the caller
// set); nothing when the instance has no FP32 copy (has_f32 false).
hipError_t launch_field_narrow(const InstDev* insts, int inst, int n_clones,
                               int plane, bool has_f32, hipStream_t stream);
hipError_t launch_perturb(const Launch& L, int philox_prefactors,
                          uint32_t seed_lo, uint32_t seed_hi);
hipError_t launch_select(const InstDev* insts, Counters* counters, int inst0,
                         int n_inst, int n_max, int weights_from_host, int draws_from_host,
                         uint32_t seed_lo, uint32_t seed_hi,
                         hipStream_t stream);
// fuse0: channel copies with noise are collided into the next interval's
// pass 0 (k_gather).
hipError_t launch_gather(const Launch& L, uint32_t seed_lo, uint32_t seed_hi,
                         int fuse0 = 0);
hipError_t launch_init_eq(const Launch& L, int inst, int n_clones, double amp,
                          uint32_t seed_lo, uint32_t seed_hi);
hipError_t launch_aos_to_soa(int storage, const InstDev* inst_dev,
                             const Counters* counters, int inst, int which,
                             int clone0, int n, const double* aos_dev,
                             hipStream_t stream);
hipError_t launch_soa_to_aos(int storage, const InstDev* inst_dev,
                             const Counters* counters, int inst, int which,
                             int clone0, int n, double* aos_dev,
                             hipStream_t stream);
hipError_t launch_snap_to_soa(const InstDev* inst_dev, int inst, int n,
                              const double* aos_dev, int Lx, int Ly,
                              double* snaps_dev, hipStream_t stream);

}  // namespace lbc
