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
// Z complex with independent standard normal parts (:127-178).  The host runs
// the DFT (hipFFT, batched over clones) between k_field_draw and
// k_field_extract.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbc_internal.h"
#include "lbc_step.h"

namespace lbc {
namespace {

// One Philox block per pair of embedding entries m = 2p, 2p + 1 of a clone:
// Box-Muller gives (re, im) of Z_2p from its first two words and of Z_2p+1
// from its last two.  The pair index p runs over blockIdx.x * blockDim.x +
// threadIdx.x, the clone over blockIdx.y.
__global__ void __launch_bounds__(256)
    k_field_draw(const InstDev* __restrict__ insts,
                 const Counters* __restrict__ counters, int inst, int clone0,
                 int fmn, uint32_t seed_lo, uint32_t seed_hi) {
  const InstDev* __restrict__ I = insts + inst;
  const int slot = blockIdx.y;
  const int clone = clone0 + slot;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * p;
  if (m >= fmn) return;
  const uint32_t interval = (uint32_t)counters->interval[inst];
  const U4 r = philox((uint32_t)p, (uint32_t)clone, interval,
                      kFieldStream + (uint32_t)inst, seed_lo, seed_hi);
  double z[4];
  box_muller_d(r, z);
  double* __restrict__ w = I->field_work + 2 * ((long long)slot * fmn + m);
  const double a0 = I->field_sqrt_ev[m];
  w[0] = a0 * z[0];
  w[1] = a0 * z[1];
  if (m + 1 < fmn) {
    const double a1 = I->field_sqrt_ev[m + 1];
    w[2] = a1 * z[2];
    w[3] = a1 * z[3];
  }
}

// map[clone][y*Lx + x] = Re F[y + fmy*x] (generate_random_field.cpp:170-178
// with the lattice's x as the embedding's row index).
__global__ void __launch_bounds__(256)
    k_field_extract(const InstDev* __restrict__ insts, int inst, int clone0) {
  const InstDev* __restrict__ I = insts + inst;
  const int slot = blockIdx.y;
  const int Lx = I->Lx, plane = I->plane, fmy = I->fmy;
  const long long fmn = (long long)I->fmx * fmy;
  const double* __restrict__ F = I->field_work + 2 * (long long)slot * fmn;
  double* __restrict__ map =
      I->field_map + (long long)(clone0 + slot) * plane;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < plane;
       s += gridDim.x * blockDim.x) {
    const int x = s % Lx, y = s / Lx;
    map[s] = F[2 * (y + (long long)fmy * x)];
  }
}

}  // namespace

hipError_t launch_interval_forced(const Launch& L) {
  if (L.n_items <= 0) return hipSuccess;
  if (L.mode != kModeRef) return hipErrorInvalidValue;
  return launch_interval_m<kModeRef, true>(L);
}

hipError_t launch_field_draw(const InstDev* insts, const Counters* counters,
                             int inst, int clone0, int n, int fmn,
                             uint32_t seed_lo, uint32_t seed_hi,
                             hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int pairs = (fmn + 1) / 2;
  const dim3 grid((pairs + 255) / 256, n);
  hipLaunchKernelGGL(k_field_draw, grid, dim3(256), 0, stream, insts, counters,
                     inst, clone0, fmn, seed_lo, seed_hi);
  return hipGetLastError();
}

hipError_t launch_field_extract(const InstDev* insts, int inst, int clone0,
                                int n, int plane, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int bx = (plane + 255) / 256 < 64 ? (plane + 255) / 256 : 64;
  hipLaunchKernelGGL(k_field_extract, dim3(bx, n), dim3(256), 0, stream, insts,
                     inst, clone0);
  return hipGetLastError();
}

}  // namespace lbc
