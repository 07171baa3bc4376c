// stream_probe -- practical HBM streaming ceilings of one MI355X, for the
// step kernel's roofline.
//
//   stream_probe
//
// Prints one JSON line of best-of-five rates in GB/s (bytes read + written
// per second):
//   copy        float4 non-temporal copy of 2 GiB (2 GiB read, 2 GiB written)
//   read, write float4 non-temporal read-only / write-only passes over 2 GiB
//   pattern_f16 the step kernel's HBM pattern without its arithmetic: 4096
//               clones of 400x100 sites in [clone][q][y][x] FP16 planes, one
//               workgroup of 512 per clone, every pass reading each site's
//               nine pulled populations (the source shifted by the
//               direction, clamped at the lattice edge) and writing nine,
//               20 passes ping-ponged between two buffers
// All accesses are in bounds by construction (indices clamped to the
// lattice, grid-stride loops bounded by the element count).
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                               \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "stream_probe: %s: %s\n", #x,                    \
                   hipGetErrorString(e_));                                  \
      return 1;                                                             \
    }                                                                       \
  } while (0)

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
    k_copy(const f4* __restrict__ a, f4* __restrict__ b, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ void __launch_bounds__(256)
    k_read(const f4* __restrict__ a, float* __restrict__ out, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  float s = 0.0f;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const f4 v = __builtin_nontemporal_load(a + i);
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1234.5f) out[0] = s;  // keeps the loads; the data are zeros
}

__global__ void __launch_bounds__(256) k_write(f4* __restrict__ b, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  const f4 v = {1.0f, 2.0f, 3.0f, 4.0f};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(v, b + i);
}

constexpr int kCX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int kCY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

__global__ void __launch_bounds__(512)
    k_pattern(unsigned short* __restrict__ a, unsigned short* __restrict__ b,
              int Lx, int Ly, long cstride, int passes) {
  const int plane = Lx * Ly;
  unsigned short* A = a + (long)blockIdx.x * cstride;
  unsigned short* B = b + (long)blockIdx.x * cstride;
  for (int p = 0; p < passes; p++) {
    const unsigned short* src = (p & 1) ? B : A;
    unsigned short* dst = (p & 1) ? A : B;
    for (int s = threadIdx.x; s < plane; s += blockDim.x) {
      const int x = s % Lx, y = s / Lx;
      unsigned short v[9];
#pragma unroll
      for (int q = 0; q < 9; q++) {
        int xs = x - kCX[q], ys = y - kCY[q];
        xs = xs < 0 ? 0 : (xs >= Lx ? Lx - 1 : xs);
        ys = ys < 0 ? 0 : (ys >= Ly ? Ly - 1 : ys);
        v[q] = src[(long)q * plane + (long)ys * Lx + xs];
      }
#pragma unroll
      for (int q = 0; q < 9; q++)
        __builtin_nontemporal_store((unsigned short)(v[q] + 1),
                                    dst + (long)q * plane + s);
    }
    __syncthreads();
  }
}

}  // namespace

int main() {
  const long bytes = 2L << 30;
  const long n4 = bytes / 16;
  f4 *a = nullptr, *b = nullptr;
  float* out = nullptr;
  CK(hipMalloc((void**)&a, bytes));
  CK(hipMalloc((void**)&b, bytes));
  CK(hipMalloc((void**)&out, 64));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  double rate[4] = {0, 0, 0, 0};
  for (int kind = 0; kind < 3; kind++) {
    for (int r = 0; r < 5; r++) {
      CK(hipEventRecord(e0));
      if (kind == 0)
        hipLaunchKernelGGL(k_copy, dim3(1024), dim3(256), 0, 0, a, b, n4);
      else if (kind == 1)
        hipLaunchKernelGGL(k_read, dim3(2048), dim3(256), 0, 0, a, out, n4);
      else
        hipLaunchKernelGGL(k_write, dim3(65536), dim3(256), 0, 0, b, n4);
      CK(hipGetLastError());
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, e0, e1));
      const double moved = (kind == 0 ? 2.0 : 1.0) * bytes;
      if (moved / ms / 1e6 > rate[kind]) rate[kind] = moved / ms / 1e6;
    }
  }
  CK(hipFree(a));
  CK(hipFree(b));

  const int Lx = 400, Ly = 100, N = 4096, passes = 20;
  const long cstride = ((long)9 * Lx * Ly + 127) / 128 * 128;
  unsigned short *pa = nullptr, *pb = nullptr;
  CK(hipMalloc((void**)&pa, 2 * cstride * N));
  CK(hipMalloc((void**)&pb, 2 * cstride * N));
  CK(hipMemset(pa, 0, 2 * cstride * N));
  CK(hipMemset(pb, 0, 2 * cstride * N));
  for (int r = 0; r < 5; r++) {
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_pattern, dim3(N), dim3(512), 0, 0, pa, pb, Lx, Ly,
                       cstride, passes);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double moved = 2.0 * 18 * Lx * Ly * (double)N * passes;
    if (moved / ms / 1e6 > rate[3]) rate[3] = moved / ms / 1e6;
  }
  std::printf("{\"copy\": %.0f, \"read\": %.0f, \"write\": %.0f, "
              "\"pattern_f16\": %.0f, \"unit\": \"GB/s\"}\n",
              rate[0], rate[1], rate[2], rate[3]);
  return 0;
}
