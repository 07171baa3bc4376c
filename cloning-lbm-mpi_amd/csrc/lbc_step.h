// Step-kernel device code of the cloning-sampler inner loop for gfx950
// (MI355X), shared by the translation units that instantiate it:
// storage traits, Philox, the collision, HBM and LDS access, the site
// passes, the multi-pass sweeps and the interval kernels with their
// launchers.  Internal; included by lbc_kernels.hip and lbc_forced.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lbc_internal.h"

namespace lbc {
namespace {

// D2Q9 velocity set and weights in the reference's ordering
// (src/mpi_LBM_GK_pops_perturb.cpp:31-32).
constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double WD[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                          1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr float WF[9] = {4.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f,
                         1.0f / 9.0f,  1.0f / 9.0f,  1.0f / 36.0f,
                         1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};

constexpr int kModeRef = 0;
constexpr int kModeChannel = 1;

enum Phase { kCollide = 0, kBoth = 1, kStream = 2 };

// ---------------------------------------------------------------- storage
template <int ST>
struct Traits;

template <>
struct Traits<kF64> {
  using S = double;
  using C = double;
  static constexpr bool shifted = false;
  __device__ static C ld(const S* p) { return *p; }
  __device__ static void st(S* p, C v) { *p = v; }
  __device__ static S pack(C v) { return v; }
  __device__ static C rest(int q) { return WD[q]; }
};

template <>
struct Traits<kF32> {
  using S = float;
  using C = float;
  static constexpr bool shifted = true;
  __device__ static C ld(const S* p) { return *p; }
  __device__ static void st(S* p, C v) { *p = v; }
  __device__ static S pack(C v) { return v; }
  __device__ static C rest(int) { return 0.0f; }
};

// FP32 value to half, rounded to FP32 first: the empty asm keeps the
// compiler from folding the operation that produced v into one
// mixed-precision instruction that rounds to half once, which it does in
// some kernels and not in others (the collision's own half results are
// spelled out in fma_store).
__device__ inline _Float16 to_half(float v) {
  asm("" : "+v"(v));
  return (_Float16)v;
}

template <>
struct Traits<kF16> {
  using S = _Float16;
  using C = float;
  static constexpr bool shifted = true;
  __device__ static C ld(const S* p) { return (float)*p; }
  __device__ static void st(S* p, C v) { *p = to_half(v); }
  __device__ static S pack(C v) { return to_half(v); }
  __device__ static C rest(int) { return 0.0f; }
};

// bfloat16 bits, rounded to nearest even by the hardware conversion
// (v_cvt_pk_bf16_f32 on gfx950: one instruction where a software rounding
// takes five).
__device__ inline uint16_t f32_to_bf16(float v) {
  return __builtin_bit_cast(uint16_t, (__bf16)v);
}

template <>
struct Traits<kBF16> {
  using S = uint16_t;
  using C = float;
  static constexpr bool shifted = true;
  __device__ static C ld(const S* p) { return __uint_as_float((uint32_t)*p << 16); }
  __device__ static void st(S* p, C v) { *p = f32_to_bf16(v); }
  __device__ static S pack(C v) { return f32_to_bf16(v); }
  __device__ static C rest(int) { return 0.0f; }
};

// ---------------------------------------------------------------- f16 in place
// An FP16 population as loaded: its half in the low 16 bits of a VGPR.  The
// collision consumes it through gfx950's mixed-precision FMA v_fma_mix_f32,
// which reads f16 operands exactly, so no separate conversion is issued.
// Each form rounds once, as the same operation on the converted value does
// (checked on the device against v_cvt_f32_f16 for every f16 bit pattern,
// subnormals included); the product takes a -0.0 addend so that a zero
// product keeps its sign.
// The register is held as the 32-bit value the zero-extending loads
// (buffer_load_ushort, ds_read_u16) leave, upper half zero: typed as a
// 16-bit value it would be masked with 0xffff before every use as an
// asm operand, nine VALU instructions per site.
struct H16 {
  uint32_t v;  // the half in the low 16 bits, zeros above
};

__device__ inline float to_f(float a) { return a; }
__device__ inline float to_f(H16 a) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)a.v);
}
// a + b, a - b, x + a, s * a on populations a, b (float or f16 in place)
__device__ inline float padd(float a, float b) { return a + b; }
__device__ inline float psub(float a, float b) { return a - b; }
__device__ inline float paddf(float a, float x) { return x + a; }
__device__ inline float pmul(float s, float a) { return s * a; }
__device__ inline float padd(H16 a, H16 b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a.v), "v"(b.v));
  return r;
}
__device__ inline float psub(H16 a, H16 b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a.v), "v"(b.v));
  return r;
}
__device__ inline float paddf(H16 a, float x) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(a.v), "v"(x));
  return r;
}
__device__ inline float pmul(float s, H16 a) {
  float r;
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,1,0]"
      : "=v"(r) : "v"(s), "v"(a.v), "s"(-0.0f));
  return r;
}
// s * a + c on a population a (one rounding)
__device__ inline float pfma(float s, float a, float c) { return fmaf(s, a, c); }
__device__ inline float pfma(float s, H16 a, float c) {
  float r;
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,1,0]"
      : "=v"(r) : "v"(s), "v"(a.v), "v"(c));
  return r;
}

// Register form of a pulled population: f16 in place for FP16 storage, the
// compute type of the storage otherwise.
template <int ST>
struct PopT {
  using P = typename Traits<ST>::C;
};
template <>
struct PopT<kF16> {
  using P = H16;
};
template <int ST>
using pop_t = typename PopT<ST>::P;

// The compute-type value of a population in register form, and its storage
// element (exact: an f16 in place is its own element).
template <int ST>
__device__ inline typename Traits<ST>::C pop_c(typename Traits<ST>::C v) {
  return v;
}
template <int ST>
__device__ inline float pop_c(H16 v) {
  return to_f(v);
}
template <int ST>
__device__ inline typename Traits<ST>::S pop_pack(typename Traits<ST>::C v) {
  return Traits<ST>::pack(v);
}
template <int ST>
__device__ inline _Float16 pop_pack(H16 v) {
  return __builtin_bit_cast(_Float16, (unsigned short)v.v);
}

// Register form of a storage element (FP16: its bits zero-extended).
template <int ST>
__device__ inline pop_t<ST> elem_pop(typename Traits<ST>::S e) {
  if constexpr (ST == kF16)
    return H16{(uint32_t)__builtin_bit_cast(unsigned short, e)};
  else
    return Traits<ST>::ld(&e);
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., SC'11).  Each round's two 32x32 -> 64-bit
// products are one v_mad_u64_u32 each (low and high word together) and
// its two three-way XORs one v_bitop3_b32 each (truth table 0x96).
struct U4 {
  uint32_t x, y, z, w;
};

__device__ inline U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 =
        __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    const uint32_t n2 =
        __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  return U4{c0, c1, c2, c3};
}

// 53-bit uniform in [0,1) from two words.
__device__ inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
         (1.0 / 9007199254740992.0);
}

__device__ inline void box_muller_d(U4 r, double n[4]) {
  const double two32inv = 2.3283064365386963e-10;
  const double twopi = 6.283185307179586;
  const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
  for (int h = 0; h < 2; h++) {
    double u1 = ((double)rr[2 * h] + 1.0) * two32inv;
    double u2 = (double)rr[2 * h + 1] * two32inv;
    double rad = sqrt(-2.0 * log(u1));
    double ang = twopi * u2;
    n[2 * h] = rad * cos(ang);
    n[2 * h + 1] = rad * sin(ang);
  }
}

// Copy noise of one site (channel mode; no reference counterpart): nine
// normals from one Philox4x32-10 draw, as five Box-Muller pairs.  Pair p
// takes bits [25p, 25p + 25) of the 128-bit draw (words x, y, z, w from the
// least significant bit): a 13-bit a, u1 = (a + 1) / 2^13 in (0, 1], then a
// 12-bit b, u2 = b / 2^12 in [0, 1); n[2p] = sqrt(-2 ln u1) cos(2 pi u2) and
// n[2p + 1] = sqrt(-2 ln u1) sin(2 pi u2).  n[9] is not used.
__device__ inline uint32_t draw_bits(const U4& r, int off, int len) {
  const uint32_t w[5] = {r.x, r.y, r.z, r.w, 0u};
  const int i = off >> 5, sh = off & 31;
  const uint64_t two = ((uint64_t)w[i + 1] << 32) | w[i];
  return (uint32_t)(two >> sh) & ((1u << len) - 1u);
}

__device__ inline void site_normals_d(const U4& r, double n[10]) {
  const double twopi = 6.283185307179586;
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const double u1 = ((double)draw_bits(r, 25 * p, 13) + 1.0) * 0x1p-13;
    const double u2 = (double)draw_bits(r, 25 * p + 13, 12) * 0x1p-12;
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = twopi * u2;
    n[2 * p] = rad * cos(ang);
    n[2 * p + 1] = rad * sin(ang);
  }
}

// The same normals in FP32 on gfx950's transcendental instructions
// (v_log_f32 in base 2, v_sqrt_f32, v_cos_f32 / v_sin_f32 taking the angle
// in revolutions), each good to about 1e-6 relative; u1 and u2 are exact.
__device__ inline void site_normals_f(const U4& r, float n[10]) {
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const float u1 = ((float)draw_bits(r, 25 * p, 13) + 1.0f) * 0x1p-13f;
    const float u2 = (float)draw_bits(r, 25 * p + 13, 12) * 0x1p-12f;
    // -2 ln(u1) = -2 ln(2) log2(u1)
    const float rad =
        __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    n[2 * p] = rad * __builtin_amdgcn_cosf(u2);
    n[2 * p + 1] = rad * __builtin_amdgcn_sinf(u2);
  }
}

// ---------------------------------------------------------------- physics
// BGK collision with Guo forcing in the reference's operation order
// (src/lbm_src/streamCollCompute.cpp:34-52), contraction off.
__device__ inline void collide_f64(double (&f)[9], double beta, double tau) {
#pragma clang fp contract(off)
  const double omega = 1. / tau;
  const double omega1 = 1.0 - omega;
  const double cforce = 1.0 - 0.5 * omega;
  double rho = 0.0, ux = 0.0, uy = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double ft = f[k];
    rho += ft;
    ux += ft * (double)CX[k];
    uy += ft * (double)CY[k];
  }
  ux = ux / rho + 0.5 * beta / rho;
  uy /= rho;
  const double u2 = -1.5 * (ux * ux + uy * uy);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double eu = (double)CX[k] * ux + (double)CY[k] * uy;
    const double eueu = 4.5 * eu * eu;
    const double feq = WD[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
    const double fd = WD[k] * cforce * beta *
                      (3. * ((double)CX[k] - ux) + 9. * (double)CX[k] * eu);
    f[k] = f[k] * omega1 + feq * omega + fd;
  }
}

// Same collision on shifted populations d = f - w, FP32, written pairwise:
// for each pair of opposite directions (+e, -e) the equilibrium splits into
// a common part t = drho + rho(4.5 eu^2 - 1.5 u^2) and an odd part
// s = 3 rho eu, so feq_q - w_q = w_q (t +/- s).
// a*b + c as a storage element of precision ST.  FP16: the exact a*b + c
// rounded once to half (v_fma_mixlo_f16); FP32: fmaf; BF16: fmaf, then
// round-to-nearest-even to bfloat16.  Spelling out the half case keeps the
// compiler from choosing between one rounding (fma_mix) and two (fma, then
// convert) differently in different kernels.
template <int ST>
__device__ inline typename Traits<ST>::S fma_store(float a, float b, float c) {
  if constexpr (ST == kF16) {
    unsigned r;
    asm("v_fma_mixlo_f16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return __builtin_bit_cast(_Float16, (unsigned short)r);
  } else {
    return Traits<ST>::pack(fmaf(a, b, c));
  }
}

// s*d + c as a storage element of precision ST for a population d in
// register form (f16 in place: read by the FMA itself).
template <int ST>
__device__ inline typename Traits<ST>::S fma_store_pop(float s, float d, float c) {
  return fma_store<ST>(s, d, c);
}
template <int ST>
__device__ inline _Float16 fma_store_pop(float s, H16 d, float c) {
  unsigned r;
  asm("v_fma_mixlo_f16 %0, %1, %2, %3 op_sel_hi:[0,1,0]"
      : "=v"(r) : "v"(s), "v"(d.v), "v"(c));
  return __builtin_bit_cast(_Float16, (unsigned short)r);
}

// Channel x boundary of one lane, folded into the shifted storages'
// collision (channel mode, no reference counterpart): the inlet (kind 1,
// x = 0) and the outlet (kind 2, x = Lx-1) set the site to the equilibrium
// of their boundary moments (channel_x_boundary states the rules), and the
// collision of an equilibrium is that equilibrium, so the lane takes those
// moments and omega = 1: d_q' = w_q (t +/- s) of (rho_b, u_b), rounded once.
struct LaneBC {
  int kind;        // 0: interior, 1: inlet, 2: outlet
  float u_in, kin;  // inlet velocity and 1 / (1 - u_in)
};

// The reference mode's collision on FP32 storage: BGK with the Guo forcing
// of src/lbm_src/streamCollCompute.cpp:7-48 in FP32, the equilibrium from
// u = (j + beta/2) / rho through rho-scaled coefficients, and (1 - omega)
// d_q taken as its own product.  The reference mode's flow is driven by a
// body force some 1e-6 of the populations, and its steady state carries
// each site's rounding pattern from step to step; measured against FP64 at
// 129x129 with 4096 clones, this form keeps the FP32 drag within 2.1e-6
// (the form of collide_shifted: 1.5e-5).  In the 2-byte storages the
// storage rounding dominates, and collide_shifted is as accurate.
template <int ST, class P>
__device__ inline void collide_shifted_ref(const P (&d)[9],
                                           typename Traits<ST>::S (&o)[9],
                                           float omega, float beta) {
  float drho = paddf(d[8], (padd(d[0], d[1]) + padd(d[2], d[3])) +
                               (padd(d[4], d[5]) + padd(d[6], d[7])));
  const float jx = (psub(d[1], d[3]) + psub(d[5], d[6])) + psub(d[8], d[7]);
  const float jy = (psub(d[2], d[4]) + padd(d[5], d[6])) - padd(d[7], d[8]);
  const float rho = 1.0f + drho;
  const float inv = __builtin_amdgcn_rcpf(rho);
  const float ux = fmaf(0.5f * beta, inv, jx * inv);
  const float uy = jy * inv;
  const float usq = fmaf(ux, ux, uy * uy);
  const float base = fmaf(-1.5f * rho, usq, drho);
  const float r45 = 4.5f * rho, r3 = 3.0f * rho;
  const float om1 = 1.0f - omega;
  const float o0 = omega * (4.0f / 9.0f), o1 = omega * (1.0f / 9.0f),
              o2 = omega * (1.0f / 36.0f);
  // eu of directions 1, 2, 5, 6 (their opposites 3, 4, 7, 8 have -eu)
  const float e1 = ux, e2 = uy, e5 = ux + uy, e6 = uy - ux;
  const float t1 = fmaf(r45 * e1, e1, base), s1 = r3 * e1;
  const float t2 = fmaf(r45 * e2, e2, base), s2 = r3 * e2;
  const float t5 = fmaf(r45 * e5, e5, base), s5 = r3 * e5;
  const float t6 = fmaf(r45 * e6, e6, base), s6 = r3 * e6;
  // d_q' = omega_q (feq_q - w_q) + (1 - omega) d_q, omega_q = omega w_q
  const float oq[9] = {o0, o1, o1, o1, o1, o2, o2, o2, o2};
  const float tq[9] = {base,    t1 + s1, t2 + s2, t1 - s1, t2 - s2,
                       t5 + s5, t6 + s6, t5 - s5, t6 - s6};
  float n[9];
#pragma unroll
  for (int q = 0; q < 9; q++) n[q] = fmaf(oq[q], tq[q], pmul(om1, d[q]));
  // Guo forcing w_q (1 - omega/2) beta (3 (cx - ux) + 9 cx eu)
  const float G = (1.0f - 0.5f * omega) * beta;
  const float g0 = (4.0f / 9.0f) * G, g1 = (1.0f / 9.0f) * G,
              g2 = (1.0f / 36.0f) * G;
  const float a = 6.0f * ux, c = 9.0f * uy, m3ux = -3.0f * ux;
  o[0] = fma_store<ST>(g0, m3ux, n[0]);
  o[1] = fma_store<ST>(g1, 3.0f + a, n[1]);
  o[3] = fma_store<ST>(g1, a - 3.0f, n[3]);
  o[2] = fma_store<ST>(g1, m3ux, n[2]);
  o[4] = fma_store<ST>(g1, m3ux, n[4]);
  o[5] = fma_store<ST>(g2, (3.0f + a) + c, n[5]);
  o[7] = fma_store<ST>(g2, (a - 3.0f) + c, n[7]);
  o[6] = fma_store<ST>(g2, (a - 3.0f) - c, n[6]);
  o[8] = fma_store<ST>(g2, (3.0f + a) - c, n[8]);
}

template <int MODE, int ST, bool BC = false, class P>
__device__ inline void collide_shifted(const P (&d)[9],
                                       typename Traits<ST>::S (&o)[9],
                                       float omega, float beta,
                                       LaneBC bc = LaneBC{0, 0.0f, 0.0f}) {
  // Every fused multiply-add is written out (the file compiles with
  // contraction off), so the rounding is fixed by this source and every
  // kernel that inlines it produces the same bits, with populations given
  // as floats or as f16 in place.
  if constexpr (MODE == kModeRef && ST == kF32) {
    collide_shifted_ref<ST>(d, o, omega, beta);
    return;
  }
  // moments, the density sharing the x momentum's partial sums: A and B
  // collect the directions with cx = +1 (1, 5, 8) and cx = -1 (3, 6, 7)
  const float A = paddf(d[8], padd(d[1], d[5]));
  const float B = paddf(d[7], padd(d[3], d[6]));
  const float Z = paddf(d[0], padd(d[2], d[4]));  // cx = 0: 0, 2, 4
  float drho = (A + B) + Z;
  const float jx = A - B;
  const float jy = paddf(d[6], padd(d[2], d[5])) - paddf(d[8], padd(d[4], d[7]));
  float rho = 1.0f + drho;
  const float inv = __builtin_amdgcn_rcpf(rho);
  float ux = jx * inv;
  if (MODE == kModeRef) ux = fmaf(0.5f * beta, inv, ux);
  float uy = jy * inv;
  // the momentum rho u: (jx, jy), plus the half-step force in REF mode
  float Jx = MODE == kModeRef ? jx + 0.5f * beta : jx, Jy = jy;
  float om1 = 1.0f - omega;
  float base = fmaf(-1.5f, fmaf(Jx, ux, Jy * uy), drho);  // drho - 1.5 rho u^2
  if constexpr (BC) {
    // A boundary lane collides its boundary moments with omega = 1: its
    // result is their equilibrium, w_q t_q.  The pass's coefficients below
    // carry omega w_q, so the lane scales its momentum and base by
    // 1 / omega and drops the (1 - omega) d_q term; selects, no branches.
    const bool in = bc.kind == 1, edge = bc.kind != 0;
    // inlet: rho_in - 1 = (D + u_in) / (1 - u_in) from the populations that
    // streamed in from the domain, D = Z + 2 B (the weights of the known
    // directions sum to 1 - (w1 + w5 + w8) and the w-shifts cancel); the
    // outlet holds the reference density with its own velocity
    const float dr_in = (fmaf(2.0f, B, Z) + bc.u_in) * bc.kin;
    const float dr_b = in ? dr_in : 0.0f;
    const float ux_b = in ? bc.u_in : ux, uy_b = in ? 0.0f : uy;
    const float rho_b = 1.0f + dr_b;
    const float sc = __builtin_amdgcn_rcpf(omega);
    const float Jx_b = (rho_b * ux_b) * sc, Jy_b = (rho_b * uy_b) * sc;
    const float base_b =
        fmaf(-1.5f, fmaf(rho_b * ux_b, ux_b, (rho_b * uy_b) * uy_b), dr_b) * sc;
    ux = edge ? ux_b : ux;
    uy = edge ? uy_b : uy;
    Jx = edge ? Jx_b : Jx;
    Jy = edge ? Jy_b : Jy;
    base = edge ? base_b : base;
    om1 = edge ? 0.0f : om1;
  }
  const float o0 = omega * (4.0f / 9.0f), o1 = omega * (1.0f / 9.0f),
              o2 = omega * (1.0f / 36.0f);
  // feq_q - w_q = w_q (base + 4.5 rho eu^2 +/- 3 rho eu) for eu = e.u of
  // directions 1, 2, 5, 6 (their opposites 3, 4, 7, 8 have -eu); with
  // omega_q = omega w_q folded in and J = rho eu,
  // omega_q (feq_q - w_q) = J (4.5 omega_q eu +/- 3 omega_q) + omega_q base,
  // two FMAs per direction on coefficients that are constants of the pass
  const float c1a = 4.5f * o1, c1b = 3.0f * o1, b1 = o1 * base;
  const float c2a = 4.5f * o2, c2b = 3.0f * o2, b2 = o2 * base;
  const float e5 = ux + uy, e6 = uy - ux, J5 = Jx + Jy, J6 = Jy - Jx;
  const float tq[9] = {o0 * base,
                       fmaf(Jx, fmaf(c1a, ux, c1b), b1),
                       fmaf(Jy, fmaf(c1a, uy, c1b), b1),
                       fmaf(Jx, fmaf(c1a, ux, -c1b), b1),
                       fmaf(Jy, fmaf(c1a, uy, -c1b), b1),
                       fmaf(J5, fmaf(c2a, e5, c2b), b2),
                       fmaf(J6, fmaf(c2a, e6, c2b), b2),
                       fmaf(J5, fmaf(c2a, e5, -c2b), b2),
                       fmaf(J6, fmaf(c2a, e6, -c2b), b2)};
  // d_q' = (1 - omega) d_q + omega_q (feq_q - w_q)
  if (MODE != kModeRef) {
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = fma_store_pop<ST>(om1, d[q], tq[q]);
  } else {
    float n[9];
#pragma unroll
    for (int q = 0; q < 9; q++) n[q] = pfma(om1, d[q], tq[q]);
    // Guo forcing w_q (1 - omega/2) beta (3 (cx - ux) + 9 cx eu)
    const float G = (1.0f - 0.5f * omega) * beta;
    const float g0 = (4.0f / 9.0f) * G, g1 = (1.0f / 9.0f) * G,
                g2 = (1.0f / 36.0f) * G;
    const float a = 6.0f * ux, c = 9.0f * uy, m3ux = -3.0f * ux;
    o[0] = fma_store<ST>(g0, m3ux, n[0]);
    o[1] = fma_store<ST>(g1, 3.0f + a, n[1]);
    o[3] = fma_store<ST>(g1, a - 3.0f, n[3]);
    o[2] = fma_store<ST>(g1, m3ux, n[2]);
    o[4] = fma_store<ST>(g1, m3ux, n[4]);
    o[5] = fma_store<ST>(g2, (3.0f + a) + c, n[5]);
    o[7] = fma_store<ST>(g2, (a - 3.0f) + c, n[7]);
    o[6] = fma_store<ST>(g2, (a - 3.0f) - c, n[6]);
    o[8] = fma_store<ST>(g2, (3.0f + a) - c, n[8]);
  }
}

// Moments and non-equilibrium stress of one node in the order of
// src/lbm_src/force.cpp:26-46 (FP64, contraction off).
__device__ inline void node_stress_f64(const double (&f)[9], double& rho_o,
                                       double& pxx_o, double& pxy_o) {
#pragma clang fp contract(off)
  double rho = 0.0, ux = 0.0, uy = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    rho += f[k];
    ux += f[k] * (double)CX[k];
    uy += f[k] * (double)CY[k];
  }
  uy /= rho;
  ux /= rho;
  double pxx = 0.0, pxy = 0.0;
  const double u2 = -1.5 * (ux * ux + uy * uy);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double eu = (double)CX[k] * ux + (double)CY[k] * uy;
    const double eueu = 4.5 * eu * eu;
    const double feq = WD[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
    const double fneq = f[k] - feq;
    pxx += fneq * (double)CX[k] * (double)CX[k];
    pxy += fneq * (double)CX[k] * (double)CY[k];
  }
  rho_o = rho;
  pxx_o = pxx;
  pxy_o = pxy;
}

// Shifted FP32 version: returns drho and the non-equilibrium stresses.
__device__ inline void node_stress_shifted(const float (&d)[9], float& drho_o,
                                           float& pxx_o, float& pxy_o) {
  const float drho = ((d[0] + d[1]) + (d[2] + d[3])) +
                     ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
  const float jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
  const float jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
  const float rho = 1.0f + drho;
  const float inv = 1.0f / rho;
  const float ux = jx * inv, uy = jy * inv;
  const float usq = 1.5f * (ux * ux + uy * uy);
  float pxx = 0.0f, pxy = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const float eu = (float)CX[k] * ux + (float)CY[k] * uy;
    const float deq = WF[k] * (drho + rho * (3.0f * eu + 4.5f * eu * eu - usq));
    const float ne = d[k] - deq;
    pxx += ne * (float)(CX[k] * CX[k]);
    pxy += ne * (float)(CX[k] * CY[k]);
  }
  drho_o = drho;
  pxx_o = pxx;
  pxy_o = pxy;
}

// Channel x boundaries on the post-stream populations of one site
// (no reference counterpart; the oracle states the same rules):
//  inlet  x = 0:    feq(rho_in, (u_in, 0)), rho_in = (f0 + f2 + f4 +
//                   2 (f3 + f6 + f7)) / (1 - u_in) from the populations that
//                   streamed in from the domain;
//  outlet x = Lx-1: after the zero-gradient ghost copy, feq(1, u) with u the
//                   site's own velocity (the outlet holds the reference
//                   density).
template <int ST>
__device__ inline void channel_x_boundary(typename Traits<ST>::C (&f)[9],
                                          bool inlet, double u_in,
                                          double inlet_k) {
  using T = Traits<ST>;
  if constexpr (!T::shifted) {
#pragma clang fp contract(off)
    double rho, ux, uy;
    if (inlet) {
      rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u_in);
      ux = u_in;
      uy = 0.0;
    } else {
      double r = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        r += f[k];
        jx += f[k] * (double)CX[k];
        jy += f[k] * (double)CY[k];
      }
      ux = jx / r;
      uy = jy / r;
      rho = 1.0;
    }
    const double u2 = -1.5 * (ux * ux + uy * uy);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double eu = (double)CX[k] * ux + (double)CY[k] * uy;
      const double eueu = 4.5 * eu * eu;
      f[k] = WD[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
    }
  } else {
    // Shifted storage (f = w + d): both rules in one branch-free evaluation,
    // so a wave holding inlet and outlet lanes runs it once.
    const float ui = (float)u_in;
    const float kin = (float)inlet_k;  // 1 / (1 - u_in)
    const float D = (f[0] + f[2] + f[4]) + 2.0f * (f[3] + f[6] + f[7]);
    const float dr = ((f[0] + f[1]) + (f[2] + f[3])) +
                     ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
    const float jx = (f[1] - f[3]) + (f[5] - f[6]) + (f[8] - f[7]);
    const float jy = (f[2] - f[4]) + (f[5] + f[6]) - (f[7] + f[8]);
    const float inv = __builtin_amdgcn_rcpf(1.0f + dr);
    // inlet: rho_in - 1 = (D + u_in) / (1 - u_in), since the weights of the
    // known directions sum to 1 - (w1 + w5 + w8) and w-shifts cancel
    const float drho = inlet ? (D + ui) * kin : 0.0f;
    const float ux = inlet ? ui : jx * inv;
    const float uy = inlet ? 0.0f : jy * inv;
    const float rho = 1.0f + drho;
    const float usq = 1.5f * fmaf(ux, ux, uy * uy);
    const float A = fmaf(-rho, usq, drho);
    const float r45 = 4.5f * rho, r3 = 3.0f * rho;
    const float e1 = ux, e2 = uy, e5 = ux + uy, e6 = uy - ux;
    const float t1 = fmaf(r45 * e1, e1, A), s1 = r3 * e1;
    const float t2 = fmaf(r45 * e2, e2, A), s2 = r3 * e2;
    const float t5 = fmaf(r45 * e5, e5, A), s5 = r3 * e5;
    const float t6 = fmaf(r45 * e6, e6, A), s6 = r3 * e6;
    f[0] = WF[0] * A;
    f[1] = WF[1] * (t1 + s1);
    f[3] = WF[3] * (t1 - s1);
    f[2] = WF[2] * (t2 + s2);
    f[4] = WF[4] * (t2 - s2);
    f[5] = WF[5] * (t5 + s5);
    f[7] = WF[7] * (t5 - s5);
    f[6] = WF[6] * (t6 + s6);
    f[8] = WF[8] * (t6 - s6);
  }
}

// Per-instance scalars held in registers by the step kernel.
struct Geo {
  int Lx, Ly, x0, x1, y0, y1, plane;
  int sx, sy;
};

// Buffer-resource access to one clone's populations: the per-direction
// plane offsets are wave-uniform (SGPR soffset), so a load or store costs no
// vector address arithmetic.  dword3 0x00020000 is the raw-buffer format of
// gfx9-family targets.
__device__ inline __amdgpu_buffer_rsrc_t clone_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes,
                                           0x00020000);
}

template <int ST>
struct BufIO;

template <>
struct BufIO<kF64> {
  using R = double;  // raw element as loaded
  __device__ static R raw(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
  }
  __device__ static double cvt(R v) { return v; }
  __device__ static double pop(R v) { return v; }
  __device__ static double ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return raw(r, vo, so);
  }
  template <bool NT>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
        vo, so, NT ? 2 : 0);
  }
  template <bool NT>
  __device__ static void st_raw(__amdgpu_buffer_rsrc_t r, int vo, int so,
                                double v) {
    st<NT>(r, vo, so, v);
  }
};

template <>
struct BufIO<kF32> {
  using R = float;
  __device__ static R raw(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
  }
  __device__ static float cvt(R v) { return v; }
  __device__ static float pop(R v) { return v; }
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return raw(r, vo, so);
  }
  template <bool NT>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo,
                                          so, NT ? 2 : 0);
  }
  template <bool NT>
  __device__ static void st_raw(__amdgpu_buffer_rsrc_t r, int vo, int so,
                                float v) {
    st<NT>(r, vo, so, v);
  }
};

template <>
struct BufIO<kF16> {
  using R = uint32_t;  // zero-extended (buffer_load_ushort)
  __device__ static R raw(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
  }
  __device__ static float cvt(R v) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)v);
  }
  __device__ static H16 pop(R v) { return H16{v}; }
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return cvt(raw(r, vo, so));
  }
  template <bool NT>
  __device__ static void st_raw(__amdgpu_buffer_rsrc_t r, int vo, int so,
                                _Float16 v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v),
                                          r, vo, so, NT ? 2 : 0);
  }
  template <bool NT>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, float v) {
    st_raw<NT>(r, vo, so, to_half(v));
  }
};

template <>
struct BufIO<kBF16> {
  using R = unsigned short;
  __device__ static R raw(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
  }
  __device__ static float cvt(R v) { return __uint_as_float((uint32_t)v << 16); }
  __device__ static float pop(R v) { return cvt(v); }
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return cvt(raw(r, vo, so));
  }
  template <bool NT>
  __device__ static void st_raw(__amdgpu_buffer_rsrc_t r, int vo, int so,
                                uint16_t v) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, vo, so, NT ? 2 : 0);
  }
  template <bool NT>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, int vo, int so, float v) {
    st_raw<NT>(r, vo, so, f32_to_bf16(v));
  }
};

// Population storage of one clone as seen by one pass: HBM through a
// buffer resource (soffset carries the wave-uniform plane offset), or the
// LDS ring of the two-level sweep, which holds `rch` consecutive sites
// (a power of two) per direction.
template <int ST>
struct HbmIO {
  __amdgpu_buffer_rsrc_t r;
  int plane;
  static constexpr int ES = (int)sizeof(typename Traits<ST>::S);
  __device__ typename Traits<ST>::C ld(int q, int site) const {
    return BufIO<ST>::ld(r, (q * plane + site) * ES, 0);
  }
  template <bool NT>
  __device__ void st(int q, int site, typename Traits<ST>::S v) const {
    BufIO<ST>::template st_raw<NT>(r, (q * plane + site) * ES, 0, v);
  }
};

//
// The rings sit from LDS address 0 (the step kernels declare no static LDS,
// so their dynamic LDS starts there), ring k at k * BYTES, and are addressed
// absolutely: ring and direction plane go into the instruction's offset
// field and no base is added.  Each direction plane holds its RCH slots
// followed by copies of slots 0 and 1 (written with them), so that the
// three source columns of a pulled row are one masked base plus an
// immediate offset of 0, 1 or 2 sites.
template <int ST, int RCH, int RING = 0>
struct LdsIO {
  typedef __attribute__((address_space(3))) typename Traits<ST>::S LS;
  static constexpr int ES = (int)sizeof(typename Traits<ST>::S);
  static constexpr int BMASK = RCH * ES - 1;  // ring bytes per direction - 1
  static constexpr int PLANEB = (RCH + 2) * ES;  // one direction plane
  static constexpr int BYTES = (int)ring_bytes(RCH, ES);  // one ring
  // A DS instruction's offset field holds 16 bits.  A ring whose planes
  // reach past it (ring 1 of the 2-byte storages at 2048-site chunks) keeps
  // a base OB in the address register instead: (b & BMASK) | OB, one
  // v_and_or_b32 where the other rings take one v_and_b32, with OB a
  // multiple of the ring bytes per direction (so | adds) chosen so that
  // every plane's remaining offset fits the field.
  static constexpr int OB = [] {
    constexpr int last = RING * BYTES + 8 * PLANEB + 2 * ES;  // past plane 8
    if (RING == 0 || last <= 65536) return 0;
    const int ob = (RING * BYTES) & ~BMASK;
    return last - ob <= 65536 ? ob : 0;
  }();
  __device__ static LS* at_abs(int q, int b) {
    return (LS*)(size_t)(unsigned)(RING * BYTES + q * PLANEB + b);
  }
  __device__ static LS* at(int q, int boff) {
    if constexpr (OB != 0) {
      unsigned v;
      // (gfx9 VOP3 reads one scalar operand at most: OB stays in a VGPR)
      asm("v_and_or_b32 %0, %1, %2, %3"
          : "=v"(v) : "v"(boff), "s"(BMASK), "v"(OB));
      return (LS*)(size_t)(v + (unsigned)(RING * BYTES - OB + q * PLANEB));
    } else {
      return at_abs(q, boff & BMASK);
    }
  }
  __device__ typename Traits<ST>::C ld_byte(int q, int boff) const {
    typename Traits<ST>::S e = *at(q, boff);
    return Traits<ST>::ld(&e);
  }
  // register form of the element at byte b of plane q (b in [0, PLANEB))
  __device__ pop_t<ST> pop_abs(int q, int b) const { return elem_pop(at_abs(q, b)); }
  // register form of the element at p (FP16: its bits zero-extended by
  // the load itself)
  __device__ static pop_t<ST> elem_pop(const LS* p) {
    if constexpr (ST == kF16) {
      typedef __attribute__((address_space(3))) const unsigned short LU;
      return H16{(uint32_t)*(LU*)p};
    } else {
      typename Traits<ST>::S e = *p;
      return Traits<ST>::ld(&e);
    }
  }
  __device__ pop_t<ST> pop_byte(int q, int boff) const {
    return pop_abs(q, boff & BMASK);
  }
  __device__ typename Traits<ST>::C ld(int q, int site) const {
    return ld_byte(q, site * ES);
  }
  template <bool NT>
  __device__ void st(int q, int site, typename Traits<ST>::S v) const {
    *at(q, site * ES) = v;
  }
};

// Collision of one site's populations (register form: the compute type,
// or f16 in place); the result comes out as storage elements.  BC: the
// lane's channel x boundary folded in (shifted storages).
template <int MODE, int ST, bool BC = false, class P>
__device__ inline void collide_any(const P (&f)[9],
                                   typename Traits<ST>::S (&o)[9], double beta,
                                   double tau, LaneBC bc = LaneBC{0, 0.0f, 0.0f}) {
  using T = Traits<ST>;
  if constexpr (!T::shifted) {
    double ff[9];
#pragma unroll
    for (int q = 0; q < 9; q++) ff[q] = (double)f[q];
    collide_f64(ff, MODE == kModeRef ? beta : 0.0, tau);
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = ff[q];
  } else {
    collide_shifted<MODE, ST, BC>(f, o, (float)(1.0 / tau), (float)beta, bc);
  }
}

// Body force of a pass.  Drive<false>: the instance's uniform beta.
// Drive<true, MT>: in a forced pass (map set) the per-site force
// beta(x, y) = (1 + eps*map[site])*beta of
// src/lbm_src/streamCollCompute.cpp:63, in FP64 with the reference's
// operation order; map holds the clone's random field in site order, FP64
// for FP64 storage, its FP32 rounding for the shifted storages (which
// collide in FP32).
template <bool FRC, class MT = double>
struct Drive {
  double beta;
  static constexpr bool kFrame = false;
  static constexpr bool kMap = false;  // the force is one number
  __device__ double at(int) const { return beta; }
};

template <class MT>
struct Drive<true, MT> {
  double beta, eps;
  const MT* __restrict__ map;  // null outside the forced passes
  static constexpr bool kFrame = false;
  static constexpr bool kMap = true;
  __device__ double at(int site) const {
#pragma clang fp contract(off)
    return map ? (1.0 + eps * (double)map[site]) * beta : beta;
  }
};

// Element type of the field maps the step kernel reads in storage ST.
template <int ST>
using map_t = typename std::conditional<ST == kF64, double, float>::type;

template <bool FRC, int ST>
using DriveOf = Drive<FRC, map_t<ST>>;

template <bool FRC, int ST>
__device__ inline DriveOf<FRC, ST> drive_of(const InstDev* __restrict__ I,
                                            int clone, bool forced) {
  if constexpr (FRC) {
    const map_t<ST>* base;
    if constexpr (ST == kF64)
      base = I->field_map;
    else
      base = I->field_map32;
    return DriveOf<FRC, ST>{
        I->beta, I->field_eps,
        forced ? base + (long long)clone * I->plane : nullptr};
  } else {
    return DriveOf<FRC, ST>{I->beta};
  }
}

// Drive of a pass that runs alone (a sweep of one pass): the body force,
// plus where the step's macroscopic fields go when the pass collides a
// macroscopic-output step (null otherwise).
template <class DV>
struct LonePass {
  DV dv;
  double* __restrict__ out;  // [3][plane]: rho, ux, uy
  static constexpr bool kFrame = true;
  __device__ double at(int site) const { return dv.at(site); }
};

// Output planes of pass n of a clone when step n is an output step
// (n < M, n a multiple of macro_every), else null.
__device__ inline double* macro_out(const InstDev* __restrict__ I, int clone,
                                    int n, int M) {
  const int E = I->macro_every;
  if (E <= 0 || n >= M || n % E) return nullptr;
  return I->macro + ((long long)clone * I->macro_frames + n / E) * 3LL * I->plane;
}

// rho, ux, uy of a site as the collision of its step computes them
// (src/lbm_src/streamCollCompute.cpp:17-26: FP64 sums in direction order,
// ux/rho + beta/(2 rho), uy/rho; shifted storages the FP32 moments of
// collide_shifted).  A solid interior site holds the rest state, as the
// reference's reset leaves it (src/mpi_LBM_GK_pops_perturb.cpp:430-439).
template <int MODE, int ST>
__device__ inline void store_macro(const typename Traits<ST>::C (&f)[9],
                                   bool solid, double beta,
                                   double* __restrict__ out, int site,
                                   int plane) {
  double rho, ux, uy;
  if constexpr (!Traits<ST>::shifted) {
#pragma clang fp contract(off)
    rho = 0.0;
    ux = 0.0;
    uy = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double ft = solid ? WD[k] : (double)f[k];
      rho += ft;
      ux += ft * (double)CX[k];
      uy += ft * (double)CY[k];
    }
    ux = ux / rho + 0.5 * beta / rho;
    uy /= rho;
  } else {
    float d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = solid ? 0.0f : (float)f[k];
    const float drho = ((d[0] + d[1]) + (d[2] + d[3])) +
                       ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
    const float jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
    const float jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
    const float r = 1.0f + drho;
    const float inv = __builtin_amdgcn_rcpf(r);
    float u = jx * inv;
    if (MODE == kModeRef) u = fmaf(0.5f * (float)beta, inv, u);
    rho = r;
    ux = u;
    uy = jy * inv;
  }
  out[site] = rho;
  out[plane + site] = ux;
  out[2 * plane + site] = uy;
}

// Storage elements of a site: collided (any pass but stream-only) or packed
// as they are.
template <int MODE, int ST, bool BC = false, class P>
__device__ inline void finish_site(const P (&f)[9],
                                   typename Traits<ST>::S (&o)[9], int PH,
                                   double beta, double tau,
                                   LaneBC bc = LaneBC{0, 0.0f, 0.0f}) {
  if (PH != kStream) {
    collide_any<MODE, ST, BC>(f, o, beta, tau, bc);
  } else if constexpr (BC) {
    // no collision to fold the x boundary into: a boundary lane stores its
    // boundary equilibrium
    typename Traits<ST>::C fc[9];
#pragma unroll
    for (int q = 0; q < 9; q++) fc[q] = pop_c<ST>(f[q]);
    if (bc.kind != 0)
      channel_x_boundary<ST>(fc, bc.kind == 1, (double)bc.u_in, (double)bc.kin);
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = Traits<ST>::pack(fc[q]);
  } else {
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = pop_pack<ST>(f[q]);
  }
}

// Drag term of a ring node from its post-stream populations, written to
// the slot(s) of its face(s) in the reference's face order.
template <int MODE, int ST>
__device__ inline void ring_drag(const Geo& g, int x, int y,
                                 const typename Traits<ST>::C (&f)[9],
                                 double* __restrict__ ring, double tau) {
  using T = Traits<ST>;
  if (MODE == kModeRef) {
    double tW, tE, tN, tS;
    if (!T::shifted) {
#pragma clang fp contract(off)
      double rho, pxx, pxy;
      double ff[9];
#pragma unroll
      for (int q = 0; q < 9; q++) ff[q] = (double)f[q];
      node_stress_f64(ff, rho, pxx, pxy);
      const double ot = 1. / 3.;
      const double cf = 1.0 - .5 * (1. / tau);
      tW = rho * ot + cf * pxx;
      tE = -rho * ot - cf * pxx;
      tN = -cf * pxy;
      tS = +cf * pxy;
    } else {
      float dr, pxx, pxy;
      float dd[9];
#pragma unroll
      for (int q = 0; q < 9; q++) dd[q] = (float)f[q];
      node_stress_shifted(dd, dr, pxx, pxy);
      const float cf = 1.0f - 0.5f / (float)tau;
      // the constant 1/3 of each node cancels between the west and east
      // faces, which have the same number of nodes
      tW = (double)(dr * (1.0f / 3.0f) + cf * pxx);
      tE = -tW;
      tN = (double)(-cf * pxy);
      tS = -tN;
    }
    if (x == g.x0) ring[y - g.y0] = tW;
    if (x == g.x1) ring[(g.sy + 1) + (y - g.y0)] = tE;
    if (y == g.y1) ring[2 * (g.sy + 1) + (x - g.x0)] = tN;
    if (y == g.y0) ring[2 * (g.sy + 1) + (g.sx + 1) + (x - g.x0)] = tS;
  } else {
    const int sm = square_mask(x, y, g.x0, g.x1, g.y0, g.y1);
    double t = 0.0;
    if (!T::shifted) {
#pragma clang fp contract(off)
#pragma unroll
      for (int q = 0; q < 9; q++)
        if ((sm >> q) & 1) t += (double)(-2 * CX[q]) * (double)f[q];
    } else {
      float tf = 0.0f;
#pragma unroll
      for (int q = 0; q < 9; q++)
        if ((sm >> q) & 1) tf += (float)(-2 * CX[q]) * (float)f[q];
      t = (double)tf;
    }
    if (x == g.x0) ring[y - g.y0] = t;
    else if (x == g.x1) ring[(g.sy + 1) + (y - g.y0)] = t;
    else if (y == g.y1) ring[2 * (g.sy + 1) + (x - g.x0)] = t;
    else ring[2 * (g.sy + 1) + (g.sx + 1) + (x - g.x0)] = t;
  }
}

// Uses all nine loaded registers at one point, so the compiler issues the
// nine loads back to back and waits once, instead of interleaving a
// conversion (and a wait) after each load.  Emits no instruction.
template <class R>
__device__ inline void join9(R (&v)[9]) {
  asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]),
               "+v"(v[5]), "+v"(v[6]), "+v"(v[7]), "+v"(v[8]));
}

// Pulls of a site.  From HBM: one buffer load per direction at the site's
// byte offset with the wave-uniform soffset q*plane - cy*Lx - cx; from the
// LDS ring: the ring slot of site - cy*Lx - cx.  voL / voR (adjL / adjR)
// fold in the x boundary for directions with cx = +1 / -1 (waves holding
// an x = 0 or x = Lx-1 site); pull9_rows serves every other wave.
template <int ST, int RCH, int RING>
__device__ inline void pull9(const LdsIO<ST, RCH, RING>& io, pop_t<ST> (&f)[9],
                             int site, int adjL, int adjR, int rowb) {
  constexpr int ES = LdsIO<ST, RCH, RING>::ES;
  // source byte offsets: column x - cx (x boundary folded in), then the row
  // y - cy one row-length (rowb bytes) away
  const int b0 = site * ES;
  const int bP = b0 + (adjL - 1) * ES, bM = b0 + (adjR + 1) * ES;
  const int col[9] = {b0, bP, b0, bM, b0, bP, bM, bM, bP};
#pragma unroll
  for (int q = 0; q < 9; q++)
    f[q] = io.pop_byte(q, CY[q] > 0 ? col[q] - rowb
                                    : (CY[q] < 0 ? col[q] + rowb : col[q]));
  asm volatile("" ::: "memory");  // see pull9_rows
  if constexpr (ST == kF16) {
    // FP16: the values leave this branch through an (empty) asm, not as
    // zero-extended loads, so the compiler does not narrow their merge with
    // pull9_rows' values to 16 bits, which would mask them with 0xffff
    // before every use on the common path
    asm volatile("" : "+v"(f[0].v), "+v"(f[1].v), "+v"(f[2].v), "+v"(f[3].v),
                 "+v"(f[4].v), "+v"(f[5].v), "+v"(f[6].v), "+v"(f[7].v),
                 "+v"(f[8].v));
  }
}

// The same pulls away from the x boundary: one masked ring offset per
// source row (y - 1, y, y + 1), one site left of the site's column, and the
// columns x - 1, x, x + 1 at 0, 1, 2 sites from it (the plane's two mirror
// slots cover a row that wraps the ring).
template <int ST, int RCH, int RING>
__device__ inline void pull9_rows(const LdsIO<ST, RCH, RING>&,
                                  pop_t<ST> (&f)[9], int site, int rowb) {
  using IO = LdsIO<ST, RCH, RING>;
  constexpr int ES = IO::ES;
  const int b = site * ES - ES;
  // the three rows as element pointers into plane 0 of the ring; plane and
  // column are constant element indices (the instruction's offset field)
  typename IO::LS* row[3] = {IO::at(0, b - rowb), IO::at(0, b), IO::at(0, b + rowb)};
#pragma unroll
  for (int q = 0; q < 9; q++)
    f[q] = IO::elem_pop(&row[1 - CY[q]][q * (RCH + 2) + (1 - CX[q])]);
  // keeps these loads in this branch: merged with the edge waves' pulls
  // they would take a per-direction address register each
  asm volatile("" ::: "memory");
}

template <int ST>
struct PassConst;

template <int ST>
__device__ inline void own9(const HbmIO<ST>& io, typename Traits<ST>::C (&f)[9],
                            int site, const PassConst<ST>& pc) {
  constexpr int ES = HbmIO<ST>::ES;
  typename BufIO<ST>::R v[9];
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = BufIO<ST>::raw(io.r, site * ES, pc.sown(q));
  join9(v);
#pragma unroll
  for (int q = 0; q < 9; q++) f[q] = BufIO<ST>::cvt(v[q]);
}

template <int ST, int RCH, int RING>
__device__ inline void own9(const LdsIO<ST, RCH, RING>& io,
                            typename Traits<ST>::C (&f)[9], int site,
                            const PassConst<ST>&) {
#pragma unroll
  for (int q = 0; q < 9; q++) f[q] = io.ld(q, site);
}

template <int ST, bool NT>
__device__ inline void put9(const HbmIO<ST>& io, const typename Traits<ST>::S (&o)[9],
                            int site, const PassConst<ST>& pc, int = -1) {
  constexpr int ES = HbmIO<ST>::ES;
#pragma unroll
  for (int q = 0; q < 9; q++)
    BufIO<ST>::template st_raw<NT>(io.r, site * ES, pc.sown(q), o[q]);
}

// wblock: the wave's block index (sites 64 * wblock ..), when the wave's
// 64 lanes hold consecutive sites; then only the first wave of every RCH
// sites tests its lanes for the mirrored slots 0 and 1 (a scalar branch);
// otherwise (-1) every lane tests its slot.
template <int ST, bool NT, int RCH, int RING>
__device__ inline void put9(const LdsIO<ST, RCH, RING>& io,
                            const typename Traits<ST>::S (&o)[9], int site,
                            const PassConst<ST>&, int wblock = -1) {
  using IO = LdsIO<ST, RCH, RING>;
#pragma unroll
  for (int q = 0; q < 9; q++) io.template st<NT>(q, site, o[q]);
  // slots 0 and 1 are also kept after the plane's last slot
  if (wblock >= 0 && (wblock & (RCH / 64 - 1)) != 0) return;
  const int slot = site & (RCH - 1);
  if (slot < 2) {
#pragma unroll
    for (int q = 0; q < 9; q++) *IO::at_abs(q, (RCH + slot) * IO::ES) = o[q];
  }
}

// Per-pass constants of the pull and store addressing: two wave-uniform
// byte counts from which every direction's soffset follows with a scalar
// multiply-add.  renew() hides their values from the compiler once per
// chunk, so it recomputes the nine soffsets with scalar instructions where
// they are used instead of keeping eighteen of them live in SGPRs across
// the sweep (which, at eight waves per SIMD, would spill them to VGPR
// lanes and cost a vector instruction per reload).
template <int ST>
struct PassConst {
  static constexpr int ES = (int)sizeof(typename Traits<ST>::S);
  int planeb;   // one direction plane in bytes
  int rowb;     // one lattice row in bytes
  int adjL, adjR;
  __device__ PassConst(const Geo& g, int mode) {
    planeb = g.plane * ES;
    rowb = g.Lx * ES;
    adjL = (mode == kModeRef) ? g.Lx : 1;
    adjR = (mode == kModeRef) ? -g.Lx : -1;
  }
  // own-site soffset of direction q
  __device__ int sown(int q) const { return q * planeb; }
  // pull soffset of direction q (HBM source): the plane, minus the source's
  // row (cy) and column (cx) offsets, one row less for cx = -1, whose
  // voffset is one row ahead
  __device__ int soff(int q) const {
    return q * planeb - CY[q] * rowb - CX[q] * ES - (CX[q] < 0 ? rowb : 0);
  }
  __device__ void renew() { asm volatile("" : "+s"(planeb), "+s"(rowb)); }
};

// What one wave's 64 consecutive sites touch (decoded from the instance's
// wave table, one word per block of 64 sites, read with a scalar load):
// wave-uniform, so the per-site branches on it are scalar and most waves
// compute no lane coordinates at all.
struct WaveGeo {
  int kind;      // kWavePlain, kWaveXEdge, kWaveGeneric or kWaveNarrow
  int x, y;      // first site
  bool edge;     // holds an x = 0 or x = Lx-1 site
  bool special;  // holds a wall-row site or a site of the square's box
  int l0, l1;    // Lx >= 64: the lanes of the x = 0 and x = Lx-1 sites
                 // (kWaveNoLane: none)
  int block;     // the block index: sites 64 * block .. 64 * block + 63
};

__device__ inline WaveGeo wave_of(const InstDev* __restrict__ I, int block) {
  const __attribute__((address_space(4))) uint32_t* t =
      (const __attribute__((address_space(4))) uint32_t*)I->wave_info;
  const unsigned i = 2u * (unsigned)block;
  const uint32_t w = t[i], wl = t[i + 1];
  WaveGeo g;
  g.kind = (int)(wl >> kWaveKindShift);
  g.x = (int)(w & kWaveCoordMask);
  g.edge = (w >> kWaveEdgeBit) & 1u;
  g.y = (int)((w >> 16) & kWaveCoordMask);
  g.special = (w >> kWaveSpecialBit) & 1u;
  g.l0 = (int)(wl & kWaveNoLane);
  g.l1 = (int)((wl >> 8) & kWaveNoLane);
  g.block = block;
  return g;
}


// Lane coordinates inside a wave (edge / special waves only).
__device__ inline void lane_xy(const Geo& g, const WaveGeo& w, int lane, int& x,
                               int& y) {
  x = w.x + lane;
  y = w.y;
  if (g.Lx >= 64) {
    if (x >= g.Lx) {
      x -= g.Lx;
      y += 1;
    }
  } else {
    const int r = x / g.Lx;
    y += r;
    x -= r * g.Lx;
  }
}

// x-boundary offsets of a lane for the pulls of directions with cx = +1 / -1
// (periodic wrap in REF mode, inlet/outlet ghost columns in channel mode);
// zero away from the x boundaries.
template <int ST>
__device__ inline void x_adjust(const Geo& g, const WaveGeo& w, int lane,
                                const PassConst<ST>& pc, int& aL, int& aR) {
  aL = aR = 0;
  if (w.kind == kWaveNarrow) {
    int x, y;
    lane_xy(g, w, lane, x, y);
    aL = (x == 0) ? pc.adjL : 0;
    aR = (x == g.Lx - 1) ? pc.adjR : 0;
  } else if (w.kind != kWavePlain) {
    aL = (lane == w.l0) ? pc.adjL : 0;
    aR = (lane == w.l1) ? pc.adjR : 0;
  }
}

// Issues the nine HBM pulls of a site (raw elements; their wait comes at
// the first use, so a caller can issue them a chunk ahead).
template <int ST>
__device__ inline void pull_issue(const Geo& g, const WaveGeo& w, int lane,
                                  int site, const HbmIO<ST>& src,
                                  const PassConst<ST>& pc,
                                  typename BufIO<ST>::R (&v)[9]) {
  constexpr int ES = HbmIO<ST>::ES;
  int aL, aR;
  x_adjust(g, w, lane, pc, aL, aR);
  // A raw buffer access is range-checked on voffset alone, so voffset must
  // stay non-negative: directions with cx = -1 carry +Lx sites in voffset
  // and -Lx in their soffset (adjR >= -Lx).
  const int vo = site * ES, voL = vo + aL * ES, voR = vo + (aR + g.Lx) * ES;
#pragma unroll
  for (int q = 0; q < 9; q++)
    v[q] = BufIO<ST>::raw(src.r, CX[q] > 0 ? voL : (CX[q] < 0 ? voR : vo),
                          pc.soff(q));
}

// The nine pulled populations of a site, in register form.
template <int ST>
__device__ inline void pull_site(const Geo& g, const WaveGeo& w, int lane,
                                 int site, const HbmIO<ST>& src,
                                 const PassConst<ST>& pc, pop_t<ST> (&f)[9]) {
  typename BufIO<ST>::R v[9];
  pull_issue(g, w, lane, site, src, pc, v);
  join9(v);
#pragma unroll
  for (int q = 0; q < 9; q++) f[q] = BufIO<ST>::pop(v[q]);
}

// Makes a reloaded FP16 population opaque to the compiler (no instruction):
// merged as a zero-extended load with the row pulls' values it would be
// narrowed to 16 bits and masked before every use.
template <class P>
__device__ inline void opaque_pop(P&) {}
__device__ inline void opaque_pop(H16& v) { asm volatile("" : "+v"(v.v)); }

// The x-boundary lanes of a wave pulled by rows (Lx >= 64, at most one lane
// of each kind): the x = 0 lane reloads its cx = +1 directions from column
// adjL - 1, the x = Lx-1 lane its cx = -1 directions from column
// Lx + adjR (the periodic wrap in REF mode, the lane's own column, i.e. the
// zero-gradient ghost, in channel mode).
template <int ST, int RCH, int RING>
__device__ inline void pull_edge_lanes(const LdsIO<ST, RCH, RING>& io,
                                       pop_t<ST> (&f)[9], int site,
                                       const WaveGeo& w, int lane,
                                       const PassConst<ST>& pc) {
  constexpr int ES = LdsIO<ST, RCH, RING>::ES;
  if (lane == w.l0) {
    const int b = (site - 1 + pc.adjL) * ES;
    pop_t<ST> v1 = io.pop_byte(1, b), v5 = io.pop_byte(5, b - pc.rowb),
              v8 = io.pop_byte(8, b + pc.rowb);
    opaque_pop(v1);
    opaque_pop(v5);
    opaque_pop(v8);
    f[1] = v1;
    f[5] = v5;
    f[8] = v8;
  }
  if (lane == w.l1) {
    const int b = (site + 1 + pc.adjR) * ES;
    pop_t<ST> v3 = io.pop_byte(3, b), v6 = io.pop_byte(6, b - pc.rowb),
              v7 = io.pop_byte(7, b + pc.rowb);
    opaque_pop(v3);
    opaque_pop(v6);
    opaque_pop(v7);
    f[3] = v3;
    f[6] = v6;
    f[7] = v7;
  }
}

template <int ST, int RCH, int RING>
__device__ inline void pull_site(const Geo& g, const WaveGeo& w, int lane,
                                 int site, const LdsIO<ST, RCH, RING>& src,
                                 const PassConst<ST>& pc, pop_t<ST> (&f)[9]) {
  if (w.kind == kWaveNarrow) {
    int aL, aR;
    x_adjust(g, w, lane, pc, aL, aR);
    pull9(src, f, site, aL, aR, pc.rowb);
  } else {
    pull9_rows(src, f, site, pc.rowb);
    if (w.kind != kWavePlain) pull_edge_lanes(src, f, site, w, lane, pc);
  }
}

// A site's own populations in register form (the collide-only pass).
template <int ST>
__device__ inline void own_pop(const HbmIO<ST>& io, pop_t<ST> (&f)[9], int site,
                               const PassConst<ST>& pc) {
  constexpr int ES = HbmIO<ST>::ES;
  typename BufIO<ST>::R v[9];
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = BufIO<ST>::raw(io.r, site * ES, pc.sown(q));
  join9(v);
#pragma unroll
  for (int q = 0; q < 9; q++) f[q] = BufIO<ST>::pop(v[q]);
}

template <int ST, int RCH, int RING>
__device__ inline void own_pop(const LdsIO<ST, RCH, RING>& io, pop_t<ST> (&f)[9],
                               int site, const PassConst<ST>&) {
#pragma unroll
  for (int q = 0; q < 9; q++) f[q] = io.pop_byte(q, site * LdsIO<ST, RCH, RING>::ES);
}

// The rest of one site of one pass, given its nine pulled populations f
// (own populations in the collide-only pass).  Waves holding wall or square
// lanes load each lane's own populations and take the bounce-back ones
// from there (ring lanes add their drag term, solid lanes store the rest
// state); channel x-boundary lanes apply the boundary rule (folded into the
// collision in the shifted storages).  Collision (except in the
// stream-only pass) and nine stores follow.  Boundary waves and the rest
// run separate copies of the collide-and-store tail, so the common path
// carries neither the solid-lane select nor register copies where the two
// would merge.
template <int MODE, int ST, bool NT, class SRC, bool BC = false, class DST,
          class DV, class P>
__device__ inline void collide_store(P (&f)[9], bool solid, const DST& dst,
                                     const PassConst<ST>& pc, const DV& dv,
                                     double beta, double tau, int site,
                                     int plane, int PH,
                                     LaneBC bc = LaneBC{0, 0.0f, 0.0f},
                                     int wblock = -1) {
  using T = Traits<ST>;
  if constexpr (DV::kFrame) {
    if (dv.out) {
      typename T::C fc[9];
#pragma unroll
      for (int q = 0; q < 9; q++) fc[q] = pop_c<ST>(f[q]);
      if constexpr (BC) {
        // the boundary lane's fields are those of its boundary equilibrium
        if (bc.kind != 0)
          channel_x_boundary<ST>(fc, bc.kind == 1, (double)bc.u_in,
                                 (double)bc.kin);
      }
      store_macro<MODE, ST>(fc, solid, beta, dv.out, site, plane);
    }
  }
  typename T::S o[9];
  finish_site<MODE, ST, BC>(f, o, PH, beta, tau, bc);
  if (solid) {
    // the reference resets the solid interior to w every step; its
    // post-collision values never reach a fluid node
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = T::pack(T::rest(q));
  }
  put9<ST, NT>(dst, o, site, pc, wblock);
}

// x or y at bit-select precision: mk all ones picks a, zero picks b (one
// v_bfi_b32 on the 32-bit register, for floats and f16 in place alike).
__device__ inline float pick(uint32_t mk, float a, float b) {
  return __uint_as_float((__float_as_uint(a) & mk) | (__float_as_uint(b) & ~mk));
}
__device__ inline H16 pick(uint32_t mk, H16 a, H16 b) {
  return H16{(a.v & mk) | (b.v & ~mk)};
}

// Bounce-back directions of the square's ring for a site at (dx, dy) from
// the box's lower-left corner (the writes of computeSquareBounceBack_TEST,
// src/lbm_src/square.cpp:24-50, as square_mask), without branches: a
// 16-bit field of a 64-bit table per column class (left edge, right edge,
// strictly inside), selected by the row class (bottom, top, strictly
// inside) of the site.
__device__ inline int square_mask_sel(int dx, int dy, int sx, int sy) {
  constexpr uint64_t kLeft = (0x80ull << 16) | (0x40ull << 32) | (0xC8ull << 48);
  constexpr uint64_t kRight = (0x100ull << 16) | (0x20ull << 32) | (0x122ull << 48);
  constexpr uint64_t kIn = (0x190ull << 16) | (0x64ull << 32);
  const bool ix = (unsigned)(dx - 1) < (unsigned)(sx - 1);
  const bool iy = (unsigned)(dy - 1) < (unsigned)(sy - 1);
  const uint64_t t = dx == 0 ? kLeft : (dx == sx ? kRight : (ix ? kIn : 0ull));
  const int sh = dy == 0 ? 16 : (dy == sy ? 32 : (iy ? 48 : 0));
  return (int)((t >> sh) & 0x1FFu);
}

// The boundary lanes of a wave on shifted storage (wall rows, the square's
// box, and in channel mode x boundaries; Lx >= 64), computed for every lane
// with selects instead of per-lane branches: bounce-back from the site's
// own populations `own` (src/lbm_src/domain_noSlipWalls.cpp:10-21,
// src/lbm_src/square.cpp:24-50), the ring nodes' drag terms as ring_drag
// computes them (REF: the stress integral's face terms,
// src/lbm_src/force.cpp:26-46; channel: momentum exchange, FP32 in
// direction order), the inlet and outlet folded into the collision
// (channel), the solid interior stored at rest.  The same results as the
// general path, bit for bit.
template <int MODE, int ST, bool NT, class SRC, class DST, class DV, class P>
__device__ inline void boundary_lanes(
    const Geo& g, const WaveGeo& w, int lane, int site, const P (&f)[9],
    const P (&own)[9], const DST& dst, const PassConst<ST>& pc,
    double* __restrict__ ring, const DV& dv, double beta, double tau,
    const InstDev* __restrict__ I, int PH) {
  int x = w.x + lane, y = w.y;
  if (x >= g.Lx) {  // the block's second row
    x -= g.Lx;
    y += 1;
  }
  const int dx = x - g.x0, dy = y - g.y0;
  const int sm = square_mask_sel(dx, dy, g.sx, g.sy);
  const int m = sm | wall_mask(y, g.Ly);
  const bool solid = (unsigned)(dx - 1) < (unsigned)(g.sx - 1) &&
                     (unsigned)(dy - 1) < (unsigned)(g.sy - 1);
  P fb[9];
  fb[0] = f[0];
#pragma unroll
  for (int q = 1; q < 9; q++)
    fb[q] = pick((uint32_t)__builtin_amdgcn_sbfe(m, q, 1), own[OPP[q]], f[q]);
  if (MODE == kModeRef && sm != 0) {
    // ring node: its faces' stress terms (one face, two at a corner)
    float dd[9];
#pragma unroll
    for (int q = 0; q < 9; q++) dd[q] = pop_c<ST>(fb[q]);
    float dr, pxx, pxy;
    node_stress_shifted(dd, dr, pxx, pxy);
    const float cf = 1.0f - 0.5f / (float)tau;
    const double tW = (double)(dr * (1.0f / 3.0f) + cf * pxx);
    const double tN = (double)(-cf * pxy);
    if (dx == 0) ring[dy] = tW;
    if (dx == g.sx) ring[(g.sy + 1) + dy] = -tW;
    if (dy == g.sy) ring[2 * (g.sy + 1) + dx] = tN;
    if (dy == 0) ring[2 * (g.sy + 1) + (g.sx + 1) + dx] = -tN;
  } else if (sm != 0) {
    // ring node: F = sum over its bounced directions of -2 cx f_q
    float tf = 0.0f;
#pragma unroll
    for (int q = 1; q < 9; q++) {
      if (CX[q] == 0) continue;  // their terms are zeros
      const uint32_t mk = (uint32_t)__builtin_amdgcn_sbfe(sm, q, 1);
      tf += __uint_as_float(__float_as_uint(pmul((float)(-2 * CX[q]), fb[q])) & mk);
    }
    const int slot =
        dx == 0 ? dy
                : (dx == g.sx ? (g.sy + 1) + dy
                              : (dy == g.sy ? 2 * (g.sy + 1) + dx
                                            : 2 * (g.sy + 1) + (g.sx + 1) + dx));
    ring[slot] = (double)tf;
  }
  if (MODE == kModeRef ||
      (w.l0 == (int)kWaveNoLane && w.l1 == (int)kWaveNoLane)) {
    collide_store<MODE, ST, NT, SRC>(fb, solid, dst, pc, dv, beta, tau, site,
                                     g.plane, PH, LaneBC{0, 0.0f, 0.0f},
                                     w.block);
  } else {
    const LaneBC bc{lane == w.l0 ? 1 : (lane == w.l1 ? 2 : 0),
                    (float)I->u_in, (float)I->inlet_k};
    collide_store<MODE, ST, NT, SRC, true>(fb, solid, dst, pc, dv, beta, tau,
                                           site, g.plane, PH, bc, w.block);
  }
}

// OP: the site's own populations were loaded ahead into own_pre (the HBM
// pass of a sweep); otherwise the boundary lanes load them from src.
template <int MODE, int ST, bool NT, bool OP = false, class SRC, class DST,
          class DV>
__device__ inline void site_finish(const Geo& g, const WaveGeo& w, int lane,
                                   int site, pop_t<ST> (&f)[9],
                                   const SRC& src, const DST& dst,
                                   const PassConst<ST>& pc,
                                   double* __restrict__ ring, const DV& dv,
                                   double beta, double tau,
                                   const InstDev* __restrict__ I, int PH,
                                   const pop_t<ST>* own_pre = nullptr) {
  using T = Traits<ST>;
  using C = typename T::C;
  if (w.kind == kWavePlain) {
    // 64 lattice sites with no boundary lane: the common path
    collide_store<MODE, ST, NT, SRC>(f, false, dst, pc, dv, beta, tau, site,
                                     g.plane, PH, LaneBC{0, 0.0f, 0.0f},
                                     w.block);
    return;
  }
  if (site >= g.plane) return;  // a lane past the lattice's last site
  // channel x boundary folded into the collision (shifted storages)
  constexpr bool kFold = MODE == kModeChannel && T::shifted;
  bool solid = false;
  if constexpr (kFold) {
    // waves whose only boundary lanes are their (at most two) x-boundary
    // lanes: the inlet or outlet rule folded into the collision of those
    // lanes, no lane coordinates
    if (PH != kCollide && w.kind == kWaveXEdge) {
      const LaneBC bc{lane == w.l0 ? 1 : (lane == w.l1 ? 2 : 0),
                      (float)I->u_in, (float)I->inlet_k};
      collide_store<MODE, ST, NT, SRC, true>(f, false, dst, pc, dv, beta, tau,
                                             site, g.plane, PH, bc, w.block);
      return;
    }
  }
  if constexpr (T::shifted) {
    if (PH != kCollide && w.kind == kWaveGeneric) {
      pop_t<ST> own[9];
      if constexpr (OP) {
#pragma unroll
        for (int q = 0; q < 9; q++) own[q] = own_pre[q];
      } else {
        own_pop(src, own, site, pc);
      }
      boundary_lanes<MODE, ST, NT, SRC>(g, w, lane, site, f, own, dst, pc,
                                        ring, dv, beta, tau, I, PH);
      return;
    }
  }
  // REF-mode x-boundary lanes need nothing beyond their pulls (the periodic
  // wrap), so wide waves with only those take the common path
  const bool edge_rule =
      w.edge && (MODE == kModeChannel || w.kind == kWaveNarrow);
  if (PH != kCollide && (edge_rule || w.special)) {
    C fc[9];
#pragma unroll
    for (int q = 0; q < 9; q++) fc[q] = pop_c<ST>(f[q]);
    int x, y;
    lane_xy(g, w, lane, x, y);
    LaneBC bc{0, (float)I->u_in, (float)I->inlet_k};
    const bool xb = MODE == kModeChannel && (x == 0 || x == g.Lx - 1);
    bool in_box = false;
    if (w.special) {
      C own[9];
      own9(src, own, site, pc);
      const int m = wall_mask(y, g.Ly) |
                    square_mask(x, y, g.x0, g.x1, g.y0, g.y1);
#pragma unroll
      for (int q = 0; q < 9; q++)
        if ((m >> q) & 1) fc[q] = own[OPP[q]];
      in_box = x >= g.x0 && x <= g.x1 && y >= g.y0 && y <= g.y1;
      solid = in_box && x > g.x0 && x < g.x1 && y > g.y0 && y < g.y1;
    }
    if (xb) {
      if constexpr (kFold)
        bc.kind = (x == 0) ? 1 : 2;
      else
        channel_x_boundary<ST>(fc, x == 0, I->u_in, I->inlet_k);
    }
    if (in_box && !solid) ring_drag<MODE, ST>(g, x, y, fc, ring, tau);
    collide_store<MODE, ST, NT, SRC, kFold>(fc, solid, dst, pc, dv, beta, tau,
                                            site, g.plane, PH, bc, w.block);
    return;
  }
  collide_store<MODE, ST, NT, SRC>(f, false, dst, pc, dv, beta, tau, site,
                                   g.plane, PH, LaneBC{0, 0.0f, 0.0f}, w.block);
}

// One site of one pass: the collide-only pass loads the site's own
// populations, every other pass pulls its nine populations (x boundary
// folded in), then site_finish.
template <int MODE, int ST, bool NT, class SRC, class DST, class DV>
__device__ inline void site_step(const Geo& g, const WaveGeo& w, int lane,
                                 int site, const SRC& src, const DST& dst,
                                 const PassConst<ST>& pc,
                                 double* __restrict__ ring, const DV& dv,
                                 double beta, double tau,
                                 const InstDev* __restrict__ I, int PH) {
  pop_t<ST> f[9];
  if (PH == kCollide)
    own_pop(src, f, site, pc);
  else
    pull_site(g, w, lane, site, src, pc, f);
  site_finish<MODE, ST, NT>(g, w, lane, site, f, src, dst, pc, ring, dv, beta,
                            tau, I, PH);
}

// One HBM-to-HBM pass of one thread over its sites tid, tid + T, ...
template <int MODE, int ST, bool NT, class DV>
__device__ inline void site_pass(const Geo& g, int tid, int nthr,
                                 const HbmIO<ST>& src, const HbmIO<ST>& dst,
                                 double* __restrict__ ring, const DV& dv,
                                 double tau, const InstDev* __restrict__ I,
                                 int PH) {
  PassConst<ST> pc(g, MODE);
  const int lane = tid & 63;
  const int wblk = __builtin_amdgcn_readfirstlane(tid) >> 6;  // wave in chunk
  const int bpc = nthr >> 6;                                  // waves per chunk
  const int nch = (g.plane + nthr - 1) / nthr;
  int site = tid;
  for (int c = 0; c < nch; c++) {
    pc.renew();
    // (site_finish skips lanes past the lattice; the guard here keeps the
    // body force's map read inside the clone's map)
    if (site < g.plane)
      site_step<MODE, ST, NT>(g, wave_of(I, c * bpc + wblk), lane, site, src,
                              dst, pc, ring, dv, dv.at(site), tau, I, PH);
    site += nthr;
  }
}

// Reach of a pull in sites: Lx + 1 with clamped x boundaries; with the
// periodic x wrap a site at x = 0 pulls from x = Lx-1 one row away, up to
// 2*Lx - 1 sites ahead (and symmetrically behind).
__host__ __device__ inline int pull_reach(int mode, int Lx) {
  return mode == kModeRef ? 2 * Lx : Lx + 1;
}

// Workgroup barrier that orders LDS traffic only: LDS writes before it are
// visible to LDS reads after it, while HBM loads and stores in flight keep
// going (__syncthreads would wait for them too).  Inside one sweep nothing
// reads HBM that the sweep writes, so this is all the chunk steps need.  The
// memory clobber keeps the compiler from moving memory accesses across.
__device__ inline void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// NP (2 or 3) stream-and-collide passes in one sweep over the lattice.  The
// first pass reads HBM and writes its post-collision populations into LDS
// ring 0; each later pass runs D chunks behind the one before (D = L for
// two passes, L + 1 for three), reads the previous ring, and writes the
// next ring or (the last pass) HBM.  A chunk is T consecutive sites (one per
// thread); L*T >= the pull reach, so every pull finds its source already in
// the ring, and a ring (>= 2L + 2 chunks) is never overwritten while a later
// pass still needs it.  One LDS barrier per chunk step (between the two
// passes, or after the three); the sweep's caller ends it with a full
// barrier.  drag[k*ring_max] holds pass k's drag slots.
template <int MODE, int ST, bool NT, int RCH, int NP, class DV>
__device__ inline void fused_sweep(const Geo& g, int tid, int nthr,
                                   const HbmIO<ST>& src, const HbmIO<ST>& dst,
                                   double* drag, int ring_max, const DV& dv,
                                   double tau, const InstDev* __restrict__ I) {
  static_assert(NP == 2 || NP == 3, "two or three passes per sweep");
  // HBM prefetch distance in chunks: each wave keeps PF chunks' first-pass
  // loads in flight.  Measured at 400x100 in FP16: two chunks are 1-2%
  // faster than one at two passes per sweep (three gain nothing more); at
  // three passes, where the rings leave half the waves, three chunks are
  // 3.5% faster than two (four gain nothing more)
  constexpr int PF = NP == 3 ? 3 : 2;
  using R = typename BufIO<ST>::R;
  PassConst<ST> pc(g, MODE);
  const int CH = nthr;
  const int L = (pull_reach(MODE, g.Lx) + CH - 1) / CH;
  const int nch = (g.plane + CH - 1) / CH;
  // lag in chunks of each pass behind the one before.  Two passes: L, with
  // the chunk step's barrier between the passes.  Three passes: L + 1, so
  // every pull reads a ring chunk written in an earlier chunk step and one
  // barrier at the end of the step orders them all (the rings hold the same
  // 2L + 2 chunks: a pass's pulls in step c reach chunks c-2L-1 .. c-1 of
  // the ring the pass before writes chunk c of)
  const int D = NP == 3 ? L + 1 : L;
  const int total = nch + (NP - 1) * D;  // chunk steps of the sweep
  const int lane = tid & 63;
  const int wblk = __builtin_amdgcn_readfirstlane(tid) >> 6;  // wave in chunk
  const int bpc = CH >> 6;                                    // waves per chunk
  const LdsIO<ST, RCH, 0> r0{};
  const LdsIO<ST, RCH, 1> r1{};
  // The first pass's HBM pulls are issued PF chunks ahead, so they travel
  // while the passes of the chunks before compute.  They are issued for
  // every lane and every step; past the last chunk they repeat the last
  // chunk's loads (served from L2, unused), and lanes past the lattice read
  // padding or, past the clone, zeros (raw buffer range check).  A
  // conditional load would make the compiler carry the old values along
  // and wait.
  // With PF > 1 the steps rotate through PF register sets (the loop body is
  // unrolled PF times) and the loads are issued on every step, so the wait
  // before a set's use counts the other sets' loads as still in flight;
  // their wave-table lookups are clamped to the table.
  R pre[PF][9];
#pragma unroll
  for (int u = 0; u < PF; u++)
    pull_issue(g, wave_of(I, min(u, nch - 1) * bpc + wblk), lane,
               min(u, nch - 1) * CH + tid, src, pc, pre[u]);
  // Boundary blocks on shifted storage take bounce-back values from the
  // sites' own populations; for the HBM pass they are loaded one chunk step
  // ahead (issued before that step's pulls, so waiting for them leaves the
  // newest chunk's pulls in flight).
  constexpr bool kOwnAhead = Traits<ST>::shifted;
  R ownr[9];
  auto own_issue = [&](int cc) {
    if constexpr (kOwnAhead) {
      if (wave_of(I, cc * bpc + wblk).kind == kWaveGeneric) {
        constexpr int ES = HbmIO<ST>::ES;
#pragma unroll
        for (int q = 0; q < 9; q++)
          ownr[q] = BufIO<ST>::raw(src.r, (cc * CH + tid) * ES, pc.sown(q));
      }
    }
  };
  own_issue(0);
  // the body force of the site a pass handles in chunk cc.  With field maps
  // in FP64 or FP32 storage the map load is issued at the top of a step,
  // ahead of the pulls and LDS barriers its collision waits behind; the
  // 2-byte storages (bound by VALU issue, and measured slower with the
  // guarded early load) load it at the collision, and a uniform force needs
  // no load at all.
  constexpr bool kEarly = DV::kMap && (ST == kF64 || ST == kF32);
  auto beta_at = [&](int cc) {
    if constexpr (!kEarly) {
      return 0.0;
    } else {
      const int s = cc * CH + tid;
      return (cc >= 0 && cc < nch && s < g.plane) ? dv.at(s) : 0.0;
    }
  };
  // (lanes past the lattice, which site_finish skips, read the map at the
  // last site instead of past the clone's map)
  auto beta_of = [&](double early, int s) {
    if constexpr (kEarly)
      return early;
    else if constexpr (DV::kMap)
      return dv.at(min(s, g.plane - 1));
    else
      return dv.at(s);
  };
  auto step = [&](int c, R (&pre)[9]) {
    pc.renew();
    const double b0 = beta_at(c);
    const double b1 = NP == 3 ? beta_at(c - D) : 0.0;
    const double b2 = beta_at(c - (NP - 1) * D);
    if (PF > 1 || c < nch) {
      const int s = c * CH + tid;
      pop_t<ST> f[9];
#pragma unroll
      for (int q = 0; q < 9; q++) f[q] = BufIO<ST>::pop(pre[q]);
      if ((unsigned)c < (unsigned)nch) {
        pop_t<ST> own[9];
#pragma unroll
        for (int q = 0; q < 9; q++) own[q] = BufIO<ST>::pop(ownr[q]);
        site_finish<MODE, ST, false, kOwnAhead>(
            g, wave_of(I, c * bpc + wblk), lane, s, f, src, r0, pc, drag, dv,
            beta_of(b0, s), tau, I, kBoth, own);
      }
      // then chunk c + PF's loads into the same registers, once their values
      // are consumed (loads cannot move above a memory clobber), so no
      // register copies are needed; the next chunk's own populations first
      asm volatile("" ::: "memory");
      if ((unsigned)(c + 1) < (unsigned)nch) own_issue(c + 1);
      const int cp = min(c + PF, nch - 1);
      pull_issue(g, wave_of(I, cp * bpc + wblk), lane, cp * CH + tid, src, pc,
                 pre);
    }
    if constexpr (NP == 2) lds_barrier();
    if constexpr (NP == 3) {
      const int cm = c - D;
      if ((unsigned)cm < (unsigned)nch) {
        const int s = cm * CH + tid;
        site_step<MODE, ST, false>(g, wave_of(I, cm * bpc + wblk), lane, s, r0,
                                   r1, pc, drag + ring_max, dv, beta_of(b1, s),
                                   tau, I, kBoth);
      }
    }
    const int cz = c - (NP - 1) * D;
    if ((unsigned)cz < (unsigned)nch) {
      const int s = cz * CH + tid;
      double* dz = drag + (NP - 1) * ring_max;
      if constexpr (NP == 3)
        site_step<MODE, ST, NT>(g, wave_of(I, cz * bpc + wblk), lane, s, r1,
                                dst, pc, dz, dv, beta_of(b2, s), tau, I, kBoth);
      else
        site_step<MODE, ST, NT>(g, wave_of(I, cz * bpc + wblk), lane, s, r0,
                                dst, pc, dz, dv, beta_of(b2, s), tau, I, kBoth);
    }
    if constexpr (NP == 3) lds_barrier();
  };
  // a multiple of PF steps; steps past the last do nothing.  The PF steps
  // are written out: a loop over them, which the compiler may leave rolled
  // once the body is large, would index the register sets indirectly.
  static_assert(PF == 2 || PF == 3, "two or three register sets");
  for (int c = 0; c < total; c += PF) {
    step(c, pre[0]);
    step(c + 1, pre[1]);
    if constexpr (PF == 3) step(c + 2, pre[2]);
  }
}

// Sum of the drag slots in the reference's order: west face (y ascending),
// east face, north face (x ascending), south face
// (src/lbm_src/force.cpp:23,54,84,114,141).
template <int MODE>
__device__ inline double ring_sum(const double* ring, int sx, int sy) {
#pragma clang fp contract(off)
  double fW = 0.0, fE = 0.0, fN = 0.0, fS = 0.0;
  for (int i = 0; i <= sy; i++) fW += ring[i];
  for (int i = 0; i <= sy; i++) fE += ring[(sy + 1) + i];
  const int lo = (MODE == kModeRef) ? 0 : 1;
  const int hi = (MODE == kModeRef) ? sx : sx - 1;
  for (int i = lo; i <= hi; i++) fN += ring[2 * (sy + 1) + i];
  for (int i = lo; i <= hi; i++) fS += ring[2 * (sy + 1) + (sx + 1) + i];
  return fW + fE + fN + fS;
}

struct KernelSetup {
  Geo g;
  int M;
  double beta, tau, scale, shift;
  void* A;
  void* B;
  double* drag;
  int clone_bytes;
};

template <int ST>
__device__ inline KernelSetup setup_clone(const InstDev* __restrict__ I,
                                          const WorkItem& wi, int par) {
  using S = typename Traits<ST>::S;
  KernelSetup k;
  k.g.Lx = I->Lx;
  k.g.Ly = I->Ly;
  k.g.x0 = I->xmin;
  k.g.x1 = I->xmax;
  k.g.y0 = I->ymin;
  k.g.y1 = I->ymax;
  k.g.plane = I->plane;
  k.g.sx = k.g.x1 - k.g.x0;
  k.g.sy = k.g.y1 - k.g.y0;
  k.M = I->steps;
  k.beta = I->beta;
  k.tau = I->tau;
  k.scale = I->scale;
  k.shift = I->shift;
  k.A = (S*)I->buf[par] + (long long)wi.clone * I->clone_stride;
  k.B = (S*)I->buf[par ^ 1] + (long long)wi.clone * I->clone_stride;
  k.drag = I->drag + (long long)wi.clone * k.M;
  k.clone_bytes = (int)(I->clone_stride * (long long)sizeof(S));
  return k;
}

// Drag of pass n (n >= 1) from its slots: F*scale - shift, accumulated in
// FP64 in pass order (src/mpi_LBM_GK_pops_perturb.cpp:445-459).
template <int MODE>
__device__ inline void take_drag(const KernelSetup& k, const double* ring, int n,
                                 double& acc) {
#pragma clang fp contract(off)
  double F = ring_sum<MODE>(ring, k.g.sx, k.g.sy);
  F = F * k.scale;
  F = F - k.shift;
  k.drag[n - 1] = F;
  acc += F;
}

template <int MODE>
__device__ inline void finish_clone(const InstDev* __restrict__ I,
                                    const WorkItem& wi, double acc) {
#pragma clang fp contract(off)
  const double o = acc * I->dt;
  I->obs[wi.clone] = o;
  I->logw[wi.clone] = I->bias * o;
}

// One cloning interval of one clone per workgroup, one pass per sweep.  The
// post-stream state S_0 is in buf[parity]; pass 0 collides it, passes
// 1..M-1 stream and collide, pass M streams only, so the interval ends with
// S_M in buf[parity ^ ((M+1)&1)].  Drag F_n is taken on S_n in pass n.
// Pass n collides step n; with FRC the passes n < forced_steps collide with
// the clone's random-field force (src/mpi_LBM_GK_forcing.cpp:335-364).
// CPW clones per workgroup: with 2, items holds pairs of clones of one
// instance (so both halves make the same passes and meet the same
// barriers), each half of the workgroup runs one of them with its own drag
// slots, and a padding item {inst, -1} only meets the barriers.
template <int MODE, int ST, bool NT, bool FRC, int CPW>
__global__ void __launch_bounds__(1024)
    k_interval(const InstDev* __restrict__ insts,
               const WorkItem* __restrict__ items,
               const Counters* __restrict__ counters, int ring_max) {
  extern __shared__ double ring_lds[];
  const int nthr = (int)blockDim.x / CPW;
  // the half of the workgroup, uniform over each wave (nthr is a multiple
  // of 64)
  const int sub =
      CPW == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x / nthr);
  const int tid = (int)threadIdx.x - sub * nthr;
  WorkItem wi = items[blockIdx.x * CPW + sub];
  const bool live = wi.clone >= 0;
  if (!live) wi.clone = 0;  // set up on clone 0 of the instance, never written
  const InstDev* __restrict__ I = insts + wi.inst;
  const KernelSetup k = setup_clone<ST>(I, wi, counters->parity[wi.inst]);
  const HbmIO<ST> ioA{clone_rsrc(k.A, k.clone_bytes), k.g.plane};
  const HbmIO<ST> ioB{clone_rsrc(k.B, k.clone_bytes), k.g.plane};
  const int M1 = FRC ? I->forced_steps : 0;
  // a copy collided by k_gather (label -1 - parent) has had its pass 0
  const bool done0 = live && I->labels[wi.clone] < 0;
  double acc = 0.0;
  for (int n = 0; n <= k.M; ++n) {
    const HbmIO<ST>& src = (n & 1) ? ioB : ioA;
    const HbmIO<ST>& dst = (n & 1) ? ioA : ioB;
    double* ring = ring_lds + (2 * sub + (n & 1)) * ring_max;
    const int ph = (n == 0) ? kCollide : (n == k.M ? kStream : kBoth);
    const LonePass<DriveOf<FRC, ST>> lp{drive_of<FRC, ST>(I, wi.clone, n < M1),
                                  macro_out(I, wi.clone, n, k.M)};
    if (live && (n > 0 || !done0))
      site_pass<MODE, ST, NT>(k.g, tid, nthr, src, dst, ring, lp, k.tau, I, ph);
    __syncthreads();
    if (live && n >= 1 && tid == 0) take_drag<MODE>(k, ring, n, acc);
  }
  if (live && tid == 0) {
    finish_clone<MODE>(I, wi, acc);
    if (done0) I->labels[wi.clone] = -1 - I->labels[wi.clone];
  }
}

// The same interval with NP passes per HBM sweep where it can: the
// collide-only pass 0 runs alone, passes 1..M-1 (stream and collide) run NP
// at a time through the LDS rings (fused_sweep; the remainder alone), and
// the stream-only pass M runs alone, fused_sweeps(M, NP, M1, E) sweeps in
// all, so the interval ends with S_M in
// buf[parity ^ (fused_sweeps(M, NP, M1, E) & 1)].  With FRC no sweep mixes
// forced passes (n < M1 = forced_steps) with the rest, and macroscopic
// output steps (multiples of E = macro_every) run alone (fused_group_at),
// so the force is one choice per sweep and output comes from lone passes.
// LDS: NP-1 population rings of 9*rch sites, then 2*NP drag slot arrays
// (NP per sweep, double-buffered over sweeps).
template <int MODE, int ST, bool NT, int RCH, int NP, bool FRC>
__device__ inline void interval_fused(const InstDev* __restrict__ insts,
                                      const WorkItem* __restrict__ items,
                                      const Counters* __restrict__ counters,
                                      int ring_max) {
  extern __shared__ double lds[];
  const WorkItem wi = items[blockIdx.x];
  const InstDev* __restrict__ I = insts + wi.inst;
  const KernelSetup k = setup_clone<ST>(I, wi, counters->parity[wi.inst]);
  const int nthr = blockDim.x, tid = threadIdx.x;
  const HbmIO<ST> ioA{clone_rsrc(k.A, k.clone_bytes), k.g.plane};
  const HbmIO<ST> ioB{clone_rsrc(k.B, k.clone_bytes), k.g.plane};
  double* const slots = lds + (NP - 1) * (LdsIO<ST, RCH>::BYTES / 8);
  const int M1 = FRC ? I->forced_steps : 0;
  const int E = I->macro_every;
  // a copy collided by k_gather (label -1 - parent) has had its pass 0,
  // which is a sweep of its own
  const bool done0 = I->labels[wi.clone] < 0;
  double acc = 0.0;
  int sweep = 0;
  for (int n = 0; n <= k.M; sweep++) {
    const HbmIO<ST>& src = (sweep & 1) ? ioB : ioA;
    const HbmIO<ST>& dst = (sweep & 1) ? ioA : ioB;
    double* drag = slots + (sweep & 1) * NP * ring_max;
    const DriveOf<FRC, ST> dv = drive_of<FRC, ST>(I, wi.clone, n < M1);
    if (fused_group_at(n, k.M, NP, M1, E)) {
      fused_sweep<MODE, ST, NT, RCH, NP>(k.g, tid, nthr, src, dst, drag,
                                         ring_max, dv, k.tau, I);
      __syncthreads();
      if (tid == 0)
        for (int p = 0; p < NP; p++)
          take_drag<MODE>(k, drag + p * ring_max, n + p, acc);
      n += NP;
    } else {
      const int ph = (n == 0) ? kCollide : (n == k.M ? kStream : kBoth);
      const LonePass<DriveOf<FRC, ST>> lp{dv, macro_out(I, wi.clone, n, k.M)};
      if (n > 0 || !done0)
        site_pass<MODE, ST, NT>(k.g, tid, nthr, src, dst, drag, lp, k.tau, I,
                                ph);
      __syncthreads();
      if (n >= 1 && tid == 0) take_drag<MODE>(k, drag, n, acc);
      n += 1;
    }
  }
  if (tid == 0) {
    finish_clone<MODE>(I, wi, acc);
    if (done0) I->labels[wi.clone] = -1 - I->labels[wi.clone];
  }
}

// Largest workgroup of the multi-pass kernel in storage ST.  An FP64 ring
// of 1024-site chunks (at least four of them, 295 KB) never fits the CU's
// LDS, so FP64 runs at most 512 threads, and the bound lets its waves hold
// up to 256 VGPRs instead of spilling at 128.
template <int ST>
constexpr int fused_max_threads() {
  return ST == kF64 ? 512 : 1024;
}

template <int MODE, int ST, bool NT, int RCH, int NP, bool FRC>
__global__ void __launch_bounds__(fused_max_threads<ST>())
    k_interval_fused(const InstDev* __restrict__ insts,
                     const WorkItem* __restrict__ items,
                     const Counters* __restrict__ counters, int ring_max) {
  interval_fused<MODE, ST, NT, RCH, NP, FRC>(insts, items, counters, ring_max);
}

// 2-byte storages, two passes per sweep: a 512-thread workgroup's ring
// takes a quarter of the CU's LDS, so registers are capped for eight waves
// per SIMD and four workgroups share a CU (uncapped, the SGPR count allows
// seven waves).
template <int MODE, int ST, bool NT, int RCH, bool FRC>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
    k_interval_fused_w8(const InstDev* __restrict__ insts,
                        const WorkItem* __restrict__ items,
                        const Counters* __restrict__ counters, int ring_max) {
  interval_fused<MODE, ST, NT, RCH, 2, FRC>(insts, items, counters, ring_max);
}

template <class K>
hipError_t raise_lds_limit(K kernel, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute((const void*)kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds);
}

// Launchers of the step kernel for one body-force kind (FRC), dispatched
// on the context's shape: passes per sweep, ring size, non-temporal stores
// and storage.
template <int MODE, int ST, bool NT, int RCH, int NP, bool FRC>
hipError_t launch_fused(const Launch& L) {
  const size_t lds = fused_lds_bytes(ST, L.ring_max, RCH, NP);
  if (!lds || L.threads > fused_max_threads<ST>()) return hipErrorInvalidValue;
  if constexpr (sizeof(typename Traits<ST>::S) == 2 && NP == 2) {
    hipError_t e =
        raise_lds_limit(k_interval_fused_w8<MODE, ST, NT, RCH, FRC>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_interval_fused_w8<MODE, ST, NT, RCH, FRC>),
                       dim3(L.n_items), dim3(L.threads), lds, L.stream, L.insts,
                       L.items, L.counters, L.ring_max);
  } else {
    hipError_t e =
        raise_lds_limit(k_interval_fused<MODE, ST, NT, RCH, NP, FRC>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_interval_fused<MODE, ST, NT, RCH, NP, FRC>),
                       dim3(L.n_items), dim3(L.threads), lds, L.stream, L.insts,
                       L.items, L.counters, L.ring_max);
  }
  return hipGetLastError();
}

template <int MODE, int ST, bool NT, int NP, bool FRC>
hipError_t launch_fused_r(const Launch& L) {
  switch (L.rch) {
    case 1024: return launch_fused<MODE, ST, NT, 1024, NP, FRC>(L);
    case 2048: return launch_fused<MODE, ST, NT, 2048, NP, FRC>(L);
    case 4096: return launch_fused<MODE, ST, NT, 4096, NP, FRC>(L);
    case 8192: return launch_fused<MODE, ST, NT, 8192, NP, FRC>(L);
  }
  return hipErrorInvalidValue;
}

template <int MODE, int ST, bool NT, bool FRC>
hipError_t launch_interval_k(const Launch& L) {
  if (L.passes == 2) return launch_fused_r<MODE, ST, NT, 2, FRC>(L);
  if (L.passes == 3) return launch_fused_r<MODE, ST, NT, 3, FRC>(L);
  if (L.cpw == 2) {
    if (2 * L.threads > 1024) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * 4 * (size_t)L.ring_max;
    hipLaunchKernelGGL((k_interval<MODE, ST, NT, FRC, 2>), dim3(L.n_items2 / 2),
                       dim3(2 * L.threads), lds, L.stream, L.insts, L.items2,
                       L.counters, L.ring_max);
    return hipGetLastError();
  }
  const size_t lds = sizeof(double) * 2 * (size_t)L.ring_max;
  hipLaunchKernelGGL((k_interval<MODE, ST, NT, FRC, 1>), dim3(L.n_items),
                     dim3(L.threads), lds, L.stream, L.insts, L.items,
                     L.counters, L.ring_max);
  return hipGetLastError();
}

template <int MODE, int ST, bool FRC>
hipError_t launch_interval_t(const Launch& L) {
  return L.nt ? launch_interval_k<MODE, ST, true, FRC>(L)
              : launch_interval_k<MODE, ST, false, FRC>(L);
}

template <int MODE, bool FRC>
hipError_t launch_interval_m(const Launch& L) {
  switch (L.storage) {
    case kF64: return launch_interval_t<MODE, kF64, FRC>(L);
    case kF32: return launch_interval_t<MODE, kF32, FRC>(L);
    case kF16: return launch_interval_t<MODE, kF16, FRC>(L);
    case kBF16: return launch_interval_t<MODE, kBF16, FRC>(L);
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace lbc
