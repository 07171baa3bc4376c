// Forcing mode of the cloning sampler for gfx950 (MI355X): the step kernel
// instantiated with the random-field body force, and the two kernels around
// the FFT that turn Philox normals into the clones' Gaussian random fields.
//
// The fields follow src/perturbation_routines/generate_random_field.cpp: a
// stationary Gaussian field with covariance exp(-r^2 / (2 sigma^2)) on the
// lattice, drawn by circulant embedding in an fmx x fmy = (2Lx-1) x (2Ly-1)
// torus.  With the embedding's eigenvalues lambda (the forward 2-D DFT of the
// wrapped covariance divided by fmx*fmy, computed once by the host), a field
// is Re F on the lattice block, F the forward 2-D DFT of sqrt(lambda) * Z and
// Z complex with independent standard normal parts (:127-178).  Im F is a
// second field of the same law, independent of Re F (the real and imaginary
// parts of a circulant embedding draw; Dietrich & Newsam 1997, Wood & Chan
// 1994), which the reference discards: here clones 2k and 2k + 1 of an
// instance take Re F and Im F of one transform, so an interval runs half the
// transforms.  The host runs the DFT (hipFFT, batched over clone pairs;
// Z2Z for FP64 storage, C2C for the shifted storages) between k_field_draw
// and k_field_extract.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbc_internal.h"
#include "lbc_step.h"

namespace lbc {
namespace {

// One Philox block per pair of embedding entries m = 2p, 2p + 1 of a clone
// pair: Box-Muller gives (re, im) of Z_2p from its first two words and of
// Z_2p+1 from its last two; the counter is the even clone's.  The pair index
// p runs over blockIdx.x * blockDim.x + threadIdx.x, the transform slot
// (clones clone0 + 2*slot and clone0 + 2*slot + 1) over blockIdx.y.
// W: the transform's precision (double for FP64 storage, float for the
// shifted storages, whose collision reads the maps in FP32).
template <typename W>
__global__ void __launch_bounds__(256)
    k_field_draw(const InstDev* __restrict__ insts,
                 const Counters* __restrict__ counters, int inst, int clone0,
                 int fmn, uint32_t seed_lo, uint32_t seed_hi) {
  const InstDev* __restrict__ I = insts + inst;
  const int slot = blockIdx.y;
  const int clone = clone0 + 2 * slot;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * p;
  if (m >= fmn) return;
  const uint32_t interval = (uint32_t)counters->interval[inst];
  const U4 r = philox((uint32_t)p, (uint32_t)clone, interval,
                      kFieldStream + (uint32_t)inst, seed_lo, seed_hi);
  double z[4];
  box_muller_d(r, z);
  W* __restrict__ w =
      reinterpret_cast<W*>(I->field_work) + 2 * ((long long)slot * fmn + m);
  const double a0 = I->field_sqrt_ev[m];
  w[0] = (W)(a0 * z[0]);
  w[1] = (W)(a0 * z[1]);
  if (m + 1 < fmn) {
    const double a1 = I->field_sqrt_ev[m + 1];
    w[2] = (W)(a1 * z[2]);
    w[3] = (W)(a1 * z[3]);
  }
}

// map[clone][y*Lx + x] = Re F[y + fmy*x] for the even clone of a slot and
// Im F[y + fmy*x] for the odd one (generate_random_field.cpp:170-178 with the
// lattice's x as the embedding's row index); clone0 + n ends the chunk.  The
// FP32 copy of the maps, where the instance has one, is written alongside.
template <typename W>
__global__ void __launch_bounds__(256)
    k_field_extract(const InstDev* __restrict__ insts, int inst, int clone0,
                    int n) {
  const InstDev* __restrict__ I = insts + inst;
  const int slot = blockIdx.y;
  const int Lx = I->Lx, plane = I->plane, fmy = I->fmy;
  const long long fmn = (long long)I->fmx * fmy;
  const W* __restrict__ F =
      reinterpret_cast<const W*>(I->field_work) + 2 * (long long)slot * fmn;
  const int c = clone0 + 2 * slot;
  const bool odd = 2 * slot + 1 < n;
  double* __restrict__ re = I->field_map + (long long)c * plane;
  double* __restrict__ im = re + plane;
  float* __restrict__ re32 =
      I->field_map32 ? I->field_map32 + (long long)c * plane : nullptr;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < plane;
       s += gridDim.x * blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    const long long k = 2 * (y + (long long)fmy * x);
    const double a = (double)F[k], b = (double)F[k + 1];
    re[s] = a;
    if (odd) im[s] = b;
    if (re32) {
      re32[s] = (float)a;
      if (odd) re32[s + plane] = (float)b;
    }
  }
}

// field_map32 = (float)field_map over all n*plane entries of an instance
// (maps the caller supplied).
__global__ void __launch_bounds__(256)
    k_field_narrow(const InstDev* __restrict__ insts, int inst, long long n) {
  const InstDev* __restrict__ I = insts + inst;
  for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
       s += (long long)gridDim.x * blockDim.x)
    I->field_map32[s] = (float)I->field_map[s];
}

}  // namespace

hipError_t launch_interval_forced(const Launch& L) {
  if (L.n_items <= 0) return hipSuccess;
  if (L.mode != kModeRef) return hipErrorInvalidValue;
  return launch_interval_m<kModeRef, true>(L);
}

hipError_t launch_field_draw(const InstDev* insts, const Counters* counters,
                             int inst, int clone0, int n, int fmn,
                             uint32_t seed_lo, uint32_t seed_hi, bool f32,
                             hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int pairs = (fmn + 1) / 2;
  const dim3 grid((pairs + 255) / 256, (n + 1) / 2);
  if (f32)
    hipLaunchKernelGGL(k_field_draw<float>, grid, dim3(256), 0, stream, insts,
                       counters, inst, clone0, fmn, seed_lo, seed_hi);
  else
    hipLaunchKernelGGL(k_field_draw<double>, grid, dim3(256), 0, stream, insts,
                       counters, inst, clone0, fmn, seed_lo, seed_hi);
  return hipGetLastError();
}

hipError_t launch_field_extract(const InstDev* insts, int inst, int clone0,
                                int n, int plane, bool f32,
                                hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int bx = (plane + 255) / 256 < 64 ? (plane + 255) / 256 : 64;
  const dim3 grid(bx, (n + 1) / 2);
  if (f32)
    hipLaunchKernelGGL(k_field_extract<float>, grid, dim3(256), 0, stream,
                       insts, inst, clone0, n);
  else
    hipLaunchKernelGGL(k_field_extract<double>, grid, dim3(256), 0, stream,
                       insts, inst, clone0, n);
  return hipGetLastError();
}

hipError_t launch_field_narrow(const InstDev* insts, int inst, int n_clones,
                               int plane, bool has_f32, hipStream_t stream) {
  if (!has_f32) return hipSuccess;
  const long long n = (long long)n_clones * plane;
  const long long want = (n + 255) / 256;
  const int blocks = (int)(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(k_field_narrow, dim3(blocks), dim3(256), 0, stream, insts,
                     inst, n);
  return hipGetLastError();
}

}  // namespace lbc
