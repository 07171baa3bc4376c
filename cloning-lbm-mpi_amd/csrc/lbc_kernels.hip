// HIP kernels of the cloning-sampler inner loop for gfx950 (MI355X).
//
//  k_interval  one persistent workgroup per clone runs a whole cloning
//              interval: pull-scheme stream + BGK collide in SoA
//              [clone][q][y][x] layout, bounce-back walls and square, the
//              mode's x boundaries, fused per-clone drag reduced in LDS in a
//              fixed order, FP64 time integration of the drag.
//  k_select    one workgroup per instance: weights, mean weight, floor
//              step, prune/add, copy table by two prefix sums.
//  k_gather    copies selected parents into their destination slots, with
//              Philox Gaussian population noise (channel mode).
//  k_perturb   snapshot-mix perturbation of marked clones (reference mode).
//  k_init_eq, k_aos_to_soa, k_soa_to_aos, k_snap_to_soa: state set-up and
//              exchange with the reference's AoS FP64 layout.
//
// The step kernel's device code lives in lbc_step.h; this unit instantiates
// it for the uniform body force, lbc_forced.hip for the random-field force.
//
// FP64 storage reproduces the reference's arithmetic order exactly
// (floating-point contraction is off in those paths), so its lattice fields
// and drag match the reference bit for bit.  Shifted storages hold f - w_q
// and collide in FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "lbc_internal.h"
#include "lbc_step.h"

namespace lbc {
namespace {

// Running scaled-cumulant estimate from the log mean weights of the first n
// intervals, in the form of post_processing/read_rvalue.py:14-21: lam[0] = 0;
// for i >= 1, T accumulates dT by repeated addition and
// lam[i] = (1/T) * sum_{j=1}^{i-1} log R_j, summed in order (the first
// interval's R never enters).  One lane: the sum is sequential.
__global__ void k_scgf(const InstDev* __restrict__ insts, int inst, int n,
                       double dT, double* __restrict__ lam) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double* logR = insts[inst].R_hist;  // (R, log R) per interval
  double s = 0.0, T = 0.0;
  lam[0] = 0.0;
  for (int i = 1; i < n; i++) {
    if (i >= 2) s = s + logR[2 * (i - 1) + 1];
    T = T + dT;
    const double oneOvT = 1. / T;
    lam[i] = oneOvT * s;
  }
}

__global__ void k_flip(const InstDev* __restrict__ insts, Counters* counters,
                       int n_inst, const int* __restrict__ flip_mask) {
  const int i = threadIdx.x;
  if (i < n_inst && flip_mask[i]) counters->parity[i] ^= 1;
}

// ---------------------------------------------------------------- selection
// Exclusive scan of v[0..n) in LDS by the whole block; returns the total.
// part must hold blockDim.x ints.
__device__ int block_scan_excl(int* v, int n, int* part) {
  const int nt = blockDim.x, t = threadIdx.x;
  const int chunk = (n + nt - 1) / nt;
  const int b = t * chunk, e = min(n, b + chunk);
  int s = 0;
  for (int i = b; i < e; i++) s += v[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    int add = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  const int total = part[nt - 1];
  int run = part[t] - s;
  for (int i = b; i < e; i++) {
    const int x = v[i];
    v[i] = run;
    run += x;
  }
  __syncthreads();
  return total;
}

// Selection of the cloning step (src/mpi_LBM_GK_pops_perturb.cpp:489-632)
// for instance `inst` described by I, in one workgroup.  `interval` keys the
// device draws and the R history slot; interval_out (if not null) receives
// interval + 1.  LDS: nb[N], temp[2N], work[N], part[blockDim].
__device__ __forceinline__ void select_body(const InstDev& I, int inst,
                                            int interval, int* interval_out,
                                            int weights_from_host,
                                            int draws_from_host,
                                            uint32_t seed_lo,
                                            uint32_t seed_hi) {
  extern __shared__ int sm[];
  const int N = I.n_clones;
  const int t = threadIdx.x, nt = blockDim.x;
  int* nb = sm;
  int* temp = nb + N;
  int* work = temp + 2 * N;
  int* part = work + N;
  __shared__ double sh_total, sh_shift;
  __shared__ int sh_bad, sh_used;

  // weights and draws
  if (t == 0) sh_bad = 0;
  __syncthreads();
  double lmax = -1e308, lmin = 1e308;
  if (!weights_from_host) {
    for (int j = t; j < N; j += nt) {
      const double l = I.logw[j];
      lmax = fmax(lmax, l);
      lmin = fmin(lmin, l);
    }
  }
  // block max/min of log-weights through LDS (reuse part as doubles)
  double* pd = (double*)(part);
  __syncthreads();
  if (!weights_from_host) {
    if (t == 0) sh_shift = -1e308;
    __syncthreads();
    // wave-level reduction then one thread folds the waves
    for (int off = 32; off > 0; off >>= 1) {
      lmax = fmax(lmax, __shfl_down(lmax, off, 64));
      lmin = fmin(lmin, __shfl_down(lmin, off, 64));
    }
    if ((t & 63) == 0) {
      pd[2 * (t >> 6)] = lmax;
      pd[2 * (t >> 6) + 1] = lmin;
    }
    __syncthreads();
    if (t == 0) {
      double mx = -1e308, mn = 1e308;
      for (int w = 0; w < (nt + 63) / 64; w++) {
        mx = fmax(mx, pd[2 * w]);
        mn = fmin(mn, pd[2 * w + 1]);
      }
      // keep the reference's arithmetic (shift 0) unless exp() would
      // overflow or underflow
      sh_shift = (mx < 700.0 && mn > -700.0) ? 0.0 : mx;
      if (!(mx == mx) || !(mn == mn) || mx > 1e300 || mn < -1e300) sh_bad = 1;
    }
    __syncthreads();
    const double sh = sh_shift;
    for (int j = t; j < N; j += nt) I.weights[j] = exp(I.logw[j] - sh);
  } else if (t == 0) {
    sh_shift = 0.0;
  }
  if (!draws_from_host) {
    for (int j = t; j < N; j += nt) {
      const U4 r = philox((uint32_t)j, (uint32_t)interval, (uint32_t)inst,
                          0x5E1EC7u, seed_lo, seed_hi);
      I.u[j] = u53(r.x, r.y);
      I.rint[j] = (int32_t)(r.z >> 1);
    }
  }
  __syncthreads();

  // total_R: sequential, in rank blocks (src/mpi_LBM_GK_pops_perturb.cpp:489-497,512)
  if (t == 0) {
#pragma clang fp contract(off)
    const int p = I.rank_blocks > 0 ? I.rank_blocks : 1;
    const int local = N / p;
    double total = 0.0;
    for (int r = 0; r < p; r++) {
      double R = 0.0;
      for (int j = r * local; j < (r + 1) * local; j++) R += I.weights[j];
      if (r == 0) total = R;
      else total += R;
    }
    total /= N;
    sh_total = total;
    if (!(total > 0.0) || !(total < 1e308)) sh_bad = 1;
  }
  __syncthreads();
  const double total = sh_total;

  // floor step (:514-520)
  for (int j = t; j < N; j += nt) {
    const double v = floor(I.weights[j] / total + I.u[j]);
    int c = 0;
    if (v >= 0.0 && v <= 2.0 * N) c = (int)v;
    else sh_bad = 1;
    nb[j] = c;
    I.nb_before[j] = c;
    work[j] = c;
  }
  __syncthreads();
  // multiset of parents (:524-532)
  const int ncp = sh_bad ? 0 : block_scan_excl(work, N, part);
  if (sh_bad || ncp < 1 || ncp > 2 * N) {
    // non-finite or inconsistent weights: keep every clone, report the error
    for (int j = t; j < N; j += nt) {
      I.nb_after[j] = 1;
      I.labels[j] = j;
      I.marks[j] = 0;
    }
    if (t == 0) {
      I.sel_info[0] = 0;
      I.sel_info[1] = 0;
      I.sel_info[2] = -5;
      I.sel_R[0] = total;
      I.sel_R[1] = log(total) + sh_shift;
    }
    return;
  }
  for (int j = t; j < N; j += nt) {
    const int o = work[j];
    for (int i = 0; i < nb[j]; i++) temp[o + i] = j;
  }
  __syncthreads();
  // prune / add (:540-571), sequential by construction
  if (t == 0) {
    const int dN = ncp - N;
    int used = 0;
    if (dN > 0) {
      for (int i = 0; i < dN; i++) {
        const int idx = I.rint[used++] % (ncp - i);
        const int c = temp[idx];
        temp[idx] = temp[ncp - i - 1];
        nb[c]--;
      }
    } else if (dN < 0) {
      for (int i = 0; i < -dN; i++) {
        const int idx = I.rint[used++] % ncp;
        const int c = temp[idx];
        temp[ncp + i] = c;
        nb[c]++;
      }
    }
    sh_used = used;
  }
  __syncthreads();
  // copy table (:590-619): the j-th extra copy (parents ascending) goes to
  // the j-th empty slot (ascending)
  for (int j = t; j < N; j += nt) {
    I.nb_after[j] = nb[j];
    work[j] = nb[j] > 1 ? nb[j] - 1 : 0;
    temp[j] = nb[j] == 0 ? 1 : 0;
  }
  __syncthreads();
  const int n_pairs = block_scan_excl(work, N, part);
  block_scan_excl(temp, N, part);
  // empty-slot list into temp[N..2N)
  for (int j = t; j < N; j += nt)
    if (nb[j] == 0) temp[N + temp[j]] = j;
  __syncthreads();
  for (int j = t; j < N; j += nt) I.labels[j] = j;
  __syncthreads();
  for (int j = t; j < N; j += nt) {
    const int e = nb[j] > 1 ? nb[j] - 1 : 0;
    for (int i = 0; i < e; i++) {
      const int pidx = work[j] + i;
      const int dst = temp[N + pidx];
      I.pairs[2 * pidx] = j;
      I.pairs[2 * pidx + 1] = dst;
      I.labels[dst] = j;
    }
  }
  __syncthreads();
  for (int j = t; j < N; j += nt) I.marks[j] = (I.labels[j] != j) ? 1 : 0;
  if (t == 0) {
    const double lr = log(total) + sh_shift;
    I.sel_info[0] = n_pairs;
    I.sel_info[1] = sh_used;
    I.sel_info[2] = 0;
    I.sel_R[0] = sh_shift == 0.0 ? total : exp(lr);
    I.sel_R[1] = lr;
    if (interval < I.max_intervals) {
      I.R_hist[2 * interval] = I.sel_R[0];
      I.R_hist[2 * interval + 1] = lr;
    }
    if (interval_out) *interval_out = interval + 1;
  }
}

__global__ void __launch_bounds__(1024)
    k_select(const InstDev* __restrict__ insts, Counters* counters, int inst0,
             int weights_from_host, int draws_from_host, uint32_t seed_lo,
             uint32_t seed_hi) {
  const int inst = inst0 + blockIdx.x;
  select_body(insts[inst], inst, counters->interval[inst],
              &counters->interval[inst], weights_from_host, draws_from_host,
              seed_lo, seed_hi);
}

// The same selection for a descriptor passed by value (no device-side
// descriptor or counters), with the caller's draws: lbc_select_static.
__global__ void __launch_bounds__(1024) k_select_static(InstDev I) {
  select_body(I, 0, 0, nullptr, 0, 1, 0u, 0u);
}

// ---------------------------------------------------------------- gather
// Destination slot k <- parent labels[k]: in channel mode with noise site
// by site, adding Gaussian noise to every population outside the solid
// interior; otherwise as a plain block copy.  The interval counter has
// already been advanced by the selection.
//
// fuse0 (channel mode with noise on shifted storage, no macroscopic
// output): the copy goes no further than the next interval's collide-only
// pass 0.  The slot's noisy state, rounded to storage as the plain copy
// stores it, is collided here and written where pass 0 writes
// (buf[parity ^ 1]); the slot's label
// becomes -1 - parent, which tells the step kernel that pass 0 is done (and
// which it turns back into the parent).  The parents' buf[parity] stays
// unwritten until the step kernel, which runs after this launch, so no
// copy reads a state being overwritten.  Results are those of the plain
// copy followed by pass 0, bit for bit; the copy's own read and write of
// the slot's populations are gone.
template <int MODE, int ST>
__global__ void __launch_bounds__(256)
    k_gather(const InstDev* __restrict__ insts,
             const WorkItem* __restrict__ items,
             const Counters* __restrict__ counters, uint32_t seed_lo,
             uint32_t seed_hi, int fuse0) {
  using T = Traits<ST>;
  using S = typename T::S;
  const WorkItem wi = items[blockIdx.x];
  const InstDev* __restrict__ I = insts + wi.inst;
  const int parent = I->labels[wi.clone];
  if (parent == wi.clone) return;
  const int par = counters->parity[wi.inst];
  const int interval = counters->interval[wi.inst];
  const S* src = (const S*)I->buf[par] + (long long)parent * I->clone_stride;
  S* dst = (S*)I->buf[par] + (long long)wi.clone * I->clone_stride;
  const int Lx = I->Lx, plane = I->plane;
  const double sigma = I->sigma;
  const bool noisy = (MODE == kModeChannel) && sigma > 0.0;
  // (FP64 stores its copies: its noise, on libm, makes the fused copy longer
  // than the pass 0 it saves)
  const bool fused = ST != kF64 && noisy && fuse0 && I->macro_every <= 0;
  // pass 0 of the next interval: the slot's post-collision populations
  Geo g{};
  g.Lx = Lx;
  g.plane = plane;
  const PassConst<ST> pc(g, MODE);
  const HbmIO<ST> out0{
      clone_rsrc((S*)I->buf[par ^ 1] + (long long)wi.clone * I->clone_stride,
                 (int)(I->clone_stride * (long long)sizeof(S))),
      plane};
  const DriveOf<false, ST> drive{I->beta};
  if (!noisy) {
    // a plain copy of the whole clone block, padding included: clone_stride
    // is a multiple of 128 elements, so the block is a whole number of
    // 16-byte vectors at 256-byte aligned addresses
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u* __restrict__ vs = (const v4u*)src;
    v4u* __restrict__ vd = (v4u*)dst;
    const long long nv = I->clone_stride * (long long)sizeof(S) / 16;
    const int bd = blockDim.x;
    long long i = threadIdx.x;
    // four loads in flight per thread before their stores
    for (; i + 3LL * bd < nv; i += 4LL * bd) {
      const v4u a = vs[i], b = vs[i + bd], c = vs[i + 2 * bd],
                d = vs[i + 3 * bd];
      vd[i] = a;
      vd[i + bd] = b;
      vd[i + 2 * bd] = c;
      vd[i + 3 * bd] = d;
    }
    for (; i < nv; i += bd) vd[i] = vs[i];
    return;
  }
  // site coordinates advance by the stride without a division per site
  const int bd = blockDim.x;
  const int dy = bd / Lx, dx = bd - dy * Lx;
  int x = threadIdx.x % Lx, y = threadIdx.x / Lx;
  for (int s = threadIdx.x; s < plane; s += bd) {
    typename T::C f[9];
#pragma unroll
    for (int q = 0; q < 9; q++) f[q] = T::ld(src + q * plane + s);
    const bool interior = x > I->xmin && x < I->xmax && y > I->ymin && y < I->ymax;
    if (!interior) {
      const U4 r = philox((uint32_t)s, (uint32_t)wi.clone, (uint32_t)interval,
                          (uint32_t)wi.inst, seed_lo, seed_hi);
      if constexpr (!T::shifted) {
        double nrm[10];
        site_normals_d(r, nrm);
#pragma unroll
        for (int q = 0; q < 9; q++) f[q] += sigma * nrm[q];
      } else {
        float nrm[10];
        site_normals_f(r, nrm);
#pragma unroll
        for (int q = 0; q < 9; q++) f[q] += (float)sigma * nrm[q];
      }
    }
    if (fused) {
      // the stored copy's elements, in the register form pass 0 loads
      pop_t<ST> p[9];
#pragma unroll
      for (int q = 0; q < 9; q++)
        p[q] = elem_pop<ST>(T::pack(f[q]));
      collide_store<MODE, ST, false, HbmIO<ST>>(p, false, out0, pc, drive,
                                                I->beta, I->tau, s, plane,
                                                kCollide);
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) T::st(dst + q * plane + s, f[q]);
    }
    x += dx;
    y += dy;
    if (x >= Lx) {
      x -= Lx;
      y += 1;
    }
  }
  if (fused) {
    __syncthreads();  // every thread has read the label
    if (threadIdx.x == 0) I->labels[wi.clone] = -1 - parent;
  }
}

// ---------------------------------------------------------------- perturb
// Snapshot-mix perturbation of marked clones
// (src/perturbation_routines/generatePerturbedState.cpp:7-36):
// f' = sum_n a_n P_n, r = |f|/|f'|, f <- (f + eps*r*f') / (1 + eps*r*sum a).
template <int ST>
__global__ void __launch_bounds__(256)
    k_perturb(const InstDev* __restrict__ insts,
              const WorkItem* __restrict__ items,
              const Counters* __restrict__ counters, int philox_prefactors,
              uint32_t seed_lo, uint32_t seed_hi) {
  using T = Traits<ST>;
  using S = typename T::S;
  const WorkItem wi = items[blockIdx.x];
  const InstDev* __restrict__ I = insts + wi.inst;
  const int NN = I->n_snapshots;
  // an instance without snapshots (the forcing mode) is not perturbed: with
  // no snapshot the mix f' would be zero and r = |f|/|f'| infinite; its
  // marks only record the last selection's copies
  if (NN <= 0 || !I->marks[wi.clone]) return;
  const int plane = I->plane;
  const int par = counters->parity[wi.inst];
  S* f = (S*)I->buf[par] + (long long)wi.clone * I->clone_stride;
  __shared__ double a[64];
  __shared__ double red[2][256];
  __shared__ double sh_w, sh_inv, sh_eps;
  const int t = threadIdx.x;
  if (t < NN) {
    if (philox_prefactors) {
      const int interval = counters->interval[wi.inst];
      const U4 r = philox((uint32_t)t, (uint32_t)wi.clone, (uint32_t)interval,
                          (uint32_t)wi.inst | 0xB0000000u, seed_lo, seed_hi);
      a[t] = u53(r.x, r.y);
    } else {
      a[t] = I->prefactors[(long long)wi.clone * NN + t];
    }
  }
  __syncthreads();
  double n0 = 0.0, np = 0.0;
  const int total = 9 * plane;
  for (int i = t; i < total; i += blockDim.x) {
    const int q = i / plane;
    double v = (double)T::ld(f + i);
    if (T::shifted) v += WD[q];
    double fp = 0.0;
    for (int nn = 0; nn < NN; nn++)
      fp += a[nn] * I->snapshots[(long long)nn * total + i];
    n0 += v * v;
    np += fp * fp;
  }
  red[0][t] = n0;
  red[1][t] = np;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma clang fp contract(off)
    double sum = 0.0;
    for (int nn = 0; nn < NN; nn++) sum += a[nn];
    const double eps = I->eps;
    const double weight = sqrt(red[0][0]) / sqrt(red[1][0]);
    const double A = 1.0 + sum * eps * weight;
    sh_w = weight;
    sh_inv = 1. / A;
    sh_eps = eps;
    I->norms[2 * wi.clone] = red[0][0];
    I->norms[2 * wi.clone + 1] = red[1][0];
  }
  __syncthreads();
  const double weight = sh_w, inv = sh_inv, eps = sh_eps;
  for (int i = t; i < total; i += blockDim.x) {
#pragma clang fp contract(off)
    const int q = i / plane;
    double v = (double)T::ld(f + i);
    if (T::shifted) v += WD[q];
    double fp = 0.0;
    for (int nn = 0; nn < NN; nn++)
      fp += a[nn] * I->snapshots[(long long)nn * total + i];
    v += eps * fp * weight;
    v = v * inv;
    if (T::shifted) v -= WD[q];
    T::st(f + i, (typename T::C)v);
  }
  __syncthreads();
  if (t == 0) I->marks[wi.clone] = 0;
}

// ---------------------------------------------------------------- set-up
template <int ST>
__global__ void k_init_eq(const InstDev* __restrict__ insts,
                          const Counters* __restrict__ counters, int inst,
                          double amp, uint32_t seed_lo, uint32_t seed_hi) {
  using T = Traits<ST>;
  using S = typename T::S;
  const InstDev* __restrict__ I = insts + inst;
  const int clone = blockIdx.x;
  const int par = counters->parity[inst];
  S* f = (S*)I->buf[par] + (long long)clone * I->clone_stride;
  const int Lx = I->Lx, plane = I->plane;
  for (int s = threadIdx.x; s < plane; s += blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    double feq[9];
    if (x > I->xmin && x < I->xmax && y > I->ymin && y < I->ymax) {
#pragma unroll
      for (int k = 0; k < 9; k++) feq[k] = WD[k];
    } else {
#pragma clang fp contract(off)
      const U4 r = philox((uint32_t)s, (uint32_t)clone, 0xFFFFFFFFu,
                          (uint32_t)inst, seed_lo, seed_hi);
      const double two32inv = 2.3283064365386963e-10;
      const double ux = I->u_in + amp * (2.0 * ((double)r.x * two32inv) - 1.0);
      const double uy = amp * (2.0 * ((double)r.y * two32inv) - 1.0);
      const double u2 = -1.5 * (ux * ux + uy * uy);
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const double eu = (double)CX[k] * ux + (double)CY[k] * uy;
        const double eueu = 4.5 * eu * eu;
        feq[k] = WD[k] * 1.0 * (1.0 + 3.0 * eu + eueu + u2);
      }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double v = T::shifted ? feq[k] - WD[k] : feq[k];
      T::st(f + k * plane + s, (typename T::C)v);
    }
  }
}

// AoS FP64 [clone][x][y][k] <-> SoA storage of clones clone0..clone0+n-1 in
// buf[which ^ parity] (which = 0: current state, 1: the other buffer).
template <int ST>
__global__ void k_aos_to_soa(const InstDev* __restrict__ insts,
                             const Counters* __restrict__ counters, int inst,
                             int which, int clone0, const double* __restrict__ aos) {
  using T = Traits<ST>;
  using S = typename T::S;
  const InstDev* __restrict__ I = insts + inst;
  const int c = blockIdx.y;
  const int Lx = I->Lx, Ly = I->Ly, plane = I->plane;
  const int par = counters->parity[inst] ^ which;
  S* f = (S*)I->buf[par] + (long long)(clone0 + c) * I->clone_stride;
  const double* src = aos + (long long)c * 9 * plane;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < plane;
       s += gridDim.x * blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    const double* p = src + 9LL * (y + (long long)Ly * x);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double v = T::shifted ? p[k] - WD[k] : p[k];
      T::st(f + k * plane + s, (typename T::C)v);
    }
  }
}

template <int ST>
__global__ void k_soa_to_aos(const InstDev* __restrict__ insts,
                             const Counters* __restrict__ counters, int inst,
                             int which, int clone0, double* __restrict__ aos) {
  using T = Traits<ST>;
  using S = typename T::S;
  const InstDev* __restrict__ I = insts + inst;
  const int c = blockIdx.y;
  const int Lx = I->Lx, Ly = I->Ly, plane = I->plane;
  const int par = counters->parity[inst] ^ which;
  const S* f = (const S*)I->buf[par] + (long long)(clone0 + c) * I->clone_stride;
  double* dst = aos + (long long)c * 9 * plane;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < plane;
       s += gridDim.x * blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    double* p = dst + 9LL * (y + (long long)Ly * x);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      double v = (double)T::ld(f + k * plane + s);
      if (T::shifted) v += WD[k];
      p[k] = v;
    }
  }
}

__global__ void k_snap_to_soa(int n, const double* __restrict__ aos, int Lx,
                              int Ly, double* __restrict__ snaps) {
  const int c = blockIdx.y;
  const int plane = Lx * Ly;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < plane;
       s += gridDim.x * blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    const double* p = aos + (long long)c * 9 * plane + 9LL * (y + (long long)Ly * x);
#pragma unroll
    for (int k = 0; k < 9; k++)
      snaps[(long long)c * 9 * plane + (long long)k * plane + s] = p[k];
  }
}


template <int MODE, int ST>
hipError_t launch_gather_t(const Launch& L, uint32_t lo, uint32_t hi,
                           int fuse0) {
  hipLaunchKernelGGL((k_gather<MODE, ST>), dim3(L.n_items), dim3(256), 0,
                     L.stream, L.insts, L.items, L.counters, lo, hi, fuse0);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_gather_m(const Launch& L, uint32_t lo, uint32_t hi,
                           int fuse0) {
  switch (L.storage) {
    case kF64: return launch_gather_t<MODE, kF64>(L, lo, hi, fuse0);
    case kF32: return launch_gather_t<MODE, kF32>(L, lo, hi, fuse0);
    case kF16: return launch_gather_t<MODE, kF16>(L, lo, hi, fuse0);
    case kBF16: return launch_gather_t<MODE, kBF16>(L, lo, hi, fuse0);
  }
  return hipErrorInvalidValue;
}

}  // namespace

int fused_ring_sites(int threads, int max_lx, int mode) {
  const int L = (pull_reach(mode, max_lx) + threads - 1) / threads;
  int R = 1;
  while (R < 2 * L + 2) R <<= 1;
  const int rch = R * threads;
  return rch < 1024 ? 1024 : rch;  // compiled ring sizes: 1024 .. 8192
}

size_t fused_lds_bytes(int storage, int ring_max, int rch, int passes) {
  if (rch > 8192 || passes < 2) return 0;
  const size_t es = storage == kF64 ? 8 : storage == kF32 ? 4 : 2;
  const size_t b = sizeof(double) * 2 * passes * (size_t)ring_max +
                   (passes - 1) * ring_bytes(rch, (int)es);
  return b <= 160 * 1024 ? b : 0;
}

hipError_t launch_interval(const Launch& L) {
  if (L.n_items <= 0) return hipSuccess;
  if (L.forced) return launch_interval_forced(L);
  return L.mode == kModeRef ? launch_interval_m<kModeRef, false>(L)
                            : launch_interval_m<kModeChannel, false>(L);
}

hipError_t launch_scgf(const InstDev* insts, int inst, int n, double dT,
                       double* lam, hipStream_t stream) {
  hipLaunchKernelGGL(k_scgf, dim3(1), dim3(64), 0, stream, insts, inst, n, dT,
                     lam);
  return hipGetLastError();
}

hipError_t launch_flip(const Launch& L, int n_inst, const int* flip_mask_dev) {
  hipLaunchKernelGGL(k_flip, dim3(1), dim3(64), 0, L.stream, L.insts,
                     L.counters, n_inst, flip_mask_dev);
  return hipGetLastError();
}

hipError_t launch_perturb(const Launch& L, int philox_prefactors,
                          uint32_t lo, uint32_t hi) {
  if (L.n_items <= 0) return hipSuccess;
  switch (L.storage) {
    case kF64:
      hipLaunchKernelGGL(k_perturb<kF64>, dim3(L.n_items), dim3(256), 0,
                         L.stream, L.insts, L.items, L.counters,
                         philox_prefactors, lo, hi);
      break;
    case kF32:
      hipLaunchKernelGGL(k_perturb<kF32>, dim3(L.n_items), dim3(256), 0,
                         L.stream, L.insts, L.items, L.counters,
                         philox_prefactors, lo, hi);
      break;
    case kF16:
      hipLaunchKernelGGL(k_perturb<kF16>, dim3(L.n_items), dim3(256), 0,
                         L.stream, L.insts, L.items, L.counters,
                         philox_prefactors, lo, hi);
      break;
    case kBF16:
      hipLaunchKernelGGL(k_perturb<kBF16>, dim3(L.n_items), dim3(256), 0,
                         L.stream, L.insts, L.items, L.counters,
                         philox_prefactors, lo, hi);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_select(const InstDev* insts, Counters* counters, int inst0,
                         int n_inst, int n_max, int weights_from_host,
                         int draws_from_host, uint32_t lo, uint32_t hi,
                         hipStream_t stream) {
  const int threads = 1024;
  const size_t lds = sizeof(int) * (4 * (size_t)n_max + threads);
  // above 64 KiB the kernel needs the attribute on the current device; it
  // is set on every such launch (a cheap call) rather than cached, so that
  // contexts on other devices or threads never skip it
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_select,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_select, dim3(n_inst), dim3(threads), lds, stream, insts,
                     counters, inst0, weights_from_host, draws_from_host, lo,
                     hi);
  return hipGetLastError();
}

// ------------------------------------------------ selection without allocation
// The one scratch buffer of lbc_select_static: the FP64 weights of at most
// kStaticMaxClones clones.  Every device the library runs on holds its own
// copy until the process ends; calls on one device share it.
static __device__ __align__(256) double scratch[kStaticMaxClones];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t static_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    // the array's device address on the current device (the host-side name
    // of the array is not a device address)
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

hipError_t launch_select_static(const InstDev& I, hipStream_t stream) {
  const int threads = 1024;
  const size_t lds = sizeof(int) * (4 * (size_t)I.n_clones + threads);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_select_static,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_select_static, dim3(1), dim3(threads), lds, stream, I);
  return hipGetLastError();
}

hipError_t launch_gather(const Launch& L, uint32_t lo, uint32_t hi,
                         int fuse0) {
  if (L.n_items <= 0) return hipSuccess;
  return L.mode == kModeRef ? launch_gather_m<kModeRef>(L, lo, hi, 0)
                            : launch_gather_m<kModeChannel>(L, lo, hi, fuse0);
}

hipError_t launch_init_eq(const Launch& L, int inst, int n_clones, double amp,
                          uint32_t lo, uint32_t hi) {
  switch (L.storage) {
    case kF64:
      hipLaunchKernelGGL(k_init_eq<kF64>, dim3(n_clones), dim3(256), 0,
                         L.stream, L.insts, L.counters, inst, amp, lo, hi);
      break;
    case kF32:
      hipLaunchKernelGGL(k_init_eq<kF32>, dim3(n_clones), dim3(256), 0,
                         L.stream, L.insts, L.counters, inst, amp, lo, hi);
      break;
    case kF16:
      hipLaunchKernelGGL(k_init_eq<kF16>, dim3(n_clones), dim3(256), 0,
                         L.stream, L.insts, L.counters, inst, amp, lo, hi);
      break;
    case kBF16:
      hipLaunchKernelGGL(k_init_eq<kBF16>, dim3(n_clones), dim3(256), 0,
                         L.stream, L.insts, L.counters, inst, amp, lo, hi);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_aos_to_soa(int storage, const InstDev* inst_dev,
                             const Counters* counters, int inst, int which,
                             int clone0, int n, const double* aos_dev,
                             hipStream_t stream) {
  dim3 grid(64, n);
  switch (storage) {
    case kF64:
      hipLaunchKernelGGL(k_aos_to_soa<kF64>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kF32:
      hipLaunchKernelGGL(k_aos_to_soa<kF32>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kF16:
      hipLaunchKernelGGL(k_aos_to_soa<kF16>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kBF16:
      hipLaunchKernelGGL(k_aos_to_soa<kBF16>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(int storage, const InstDev* inst_dev,
                             const Counters* counters, int inst, int which,
                             int clone0, int n, double* aos_dev,
                             hipStream_t stream) {
  dim3 grid(64, n);
  switch (storage) {
    case kF64:
      hipLaunchKernelGGL(k_soa_to_aos<kF64>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kF32:
      hipLaunchKernelGGL(k_soa_to_aos<kF32>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kF16:
      hipLaunchKernelGGL(k_soa_to_aos<kF16>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
    case kBF16:
      hipLaunchKernelGGL(k_soa_to_aos<kBF16>, grid, dim3(256), 0, stream,
                         inst_dev, counters, inst, which, clone0, aos_dev);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_snap_to_soa(const InstDev*, int, int n, const double* aos_dev,
                              int Lx, int Ly, double* snaps_dev,
                              hipStream_t stream) {
  hipLaunchKernelGGL(k_snap_to_soa, dim3(64, n), dim3(256), 0, stream, n,
                     aos_dev, Lx, Ly, snaps_dev);
  return hipGetLastError();
}

}  // namespace lbc
