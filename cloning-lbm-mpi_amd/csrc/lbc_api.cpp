// Host side of the C ABI declared in include/lbc.h: context and device
// buffer management, launch sequencing on the context's stream, HIP-graph
// capture of a whole cloning step, event timing of the step kernel, and the
// random fields of the forcing mode (embedding spectrum, per-interval draws
// through hipFFT).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lbc.h"
#include "lbc_internal.h"

using lbc::Counters;
using lbc::InstDev;
using lbc::WorkItem;

namespace {

thread_local std::string g_err;
// error text of lbc_select_static, which must not allocate (string literals)
thread_local const char* g_err_static = nullptr;

constexpr int kMaxClones = 8192;
constexpr int kMaxIntervals = 1 << 16;
// forcing mode: largest embedding torus, and the FFT work buffer's budget
constexpr long long kMaxFieldEntries = 1LL << 26;
constexpr size_t kFieldWorkBytes = size_t(1) << 30;
// macroscopic output buffer limit (64 GiB of FP64)
constexpr long long kMaxMacroDoubles = 1LL << 33;

size_t elem_bytes(int storage) {
  switch (storage) {
    case LBC_STORE_F64: return 8;
    case LBC_STORE_F32: return 4;
    default: return 2;
  }
}

}  // namespace

struct lbc_ctx {
  lbc_config cfg{};
  std::vector<lbc_instance> inst;
  hipStream_t stream = nullptr;
  std::vector<InstDev> idh;           // host copies of the descriptors
  InstDev* idd = nullptr;             // device descriptors
  Counters* counters = nullptr;
  WorkItem* items = nullptr;
  int n_items = 0;
  WorkItem* items2 = nullptr;         // the same clones in same-instance pairs
  int n_items2 = 0;
  int* flip_mask = nullptr;
  double* scgf = nullptr;             // running estimate of one instance
  std::vector<int> parity;            // host mirror of counters->parity
  std::vector<int> selected;          // selection done since last gather
  std::vector<void*> allocs;
  double* staging = nullptr;          // AoS FP64 transfer buffer (device)
  size_t staging_bytes = 0;
  int threads = 256;
  int vec = 1;                        // sites per thread per iteration (step kernel)
  int nt = 0;                         // non-temporal population stores
  int passes = 1;                     // LBM passes per HBM sweep (LDS rings)
  int cpw = 1;                        // clones per workgroup (one-pass kernel)
  int rch = 0;                        // LDS ring sites per direction
  int max_lx = 0;
  int ring_max = 0;
  int n_max = 0;
  std::string err;
  std::vector<hipEvent_t> ev;         // pairs around step-kernel launches
  int ev_used = 0;
  // step kernel | the rest | the rest with the copies collided into the
  // next interval's pass 0
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graphs[3] = {nullptr, nullptr, nullptr};
  bool timing_from_graph = false;
  std::vector<float> graph_ms;
  // forcing mode, per instance
  bool forced = false;                // some instance has forced steps
  std::vector<hipfftHandle> fft;      // batched in-place 2-D Z2Z plans
  std::vector<int> fchunk;            // transforms per FFT batch (two clones each)
  std::vector<char> maps_set;         // next interval uses injected maps
  // copy kernel of the last lbc_gather: events around it, bytes it moved
  std::vector<int> last_copies;       // per instance, from the last lbc_select
  hipEvent_t gev[2] = {nullptr, nullptr};
  bool gather_timed = false;
  double gather_bytes = 0.0;
  // lbc_gather_fused left copies collided into the next interval's pass 0:
  // the states are not readable until an interval has run
  bool pass0_pending = false;
};

namespace {

// Makes `dev` the calling thread's current HIP device for one API call and
// gives the caller's device back when the call returns: a context's buffers,
// stream, events, graphs and FFT plans all belong to the context's device,
// and allocations, synchronous copies and event creation follow the current
// device, not the stream.
class DeviceScope {
 public:
  explicit DeviceScope(int dev) {
    int cur = -1;
    if (dev < 0 || hipGetDevice(&cur) != hipSuccess) return;
    if (cur == dev) {
      ok_ = true;
      return;
    }
    ok_ = hipSetDevice(dev) == hipSuccess;
    if (ok_) prev_ = cur;
  }
  ~DeviceScope() {
    if (prev_ >= 0) (void)hipSetDevice(prev_);
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
  bool ok() const { return ok_; }

 private:
  int prev_ = -1;
  bool ok_ = false;
};

#define ON_CTX_DEVICE(c) DeviceScope dev_scope_((c) ? (c)->cfg.device : -1)

int fail(lbc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_err = msg;
  g_err_static = nullptr;
  return code;
}

int fail_static(int code, const char* msg) {
  g_err_static = msg;
  return code;
}

#define HIPC(ctx, expr)                                                      \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail((ctx), LBC_E_HIP,                                          \
                  std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

int validate(const lbc_config* cfg, std::string& why) {
  if (!cfg || !cfg->instances) return why = "null config", -1;
  if (cfg->mode != LBC_MODE_REF && cfg->mode != LBC_MODE_CHANNEL)
    return why = "mode must be LBC_MODE_REF or LBC_MODE_CHANNEL", -1;
  if (cfg->storage < LBC_STORE_F64 || cfg->storage > LBC_STORE_BF16)
    return why = "unknown storage", -1;
  if (cfg->n_instances < 1 || cfg->n_instances > lbc::kMaxInstances)
    return why = "n_instances must be in [1, 64]", -1;
  if (cfg->threads != 0 &&
      (cfg->threads < 64 || cfg->threads > 1024 || cfg->threads % 64))
    return why = "threads must be 0 or a multiple of 64 in [64, 1024]", -1;
  {
    const int v = cfg->flags & 7;
    const int ntm = (cfg->flags >> 3) & 3;
    const int cpm = (cfg->flags >> 7) & 3;
    const int pm = (cfg->flags >> 5) & 3;
    if ((cfg->flags & ~511) || (v != 0 && v != 1) || ntm == 3 || cpm == 3 ||
        (cpm == 2 && pm > 1))
      return why = "flags: bits 0-2 sites per thread (0 or 1), bits 3-4 "
                   "non-temporal stores (0 tuned, 1 off, 2 on), bits 5-6 "
                   "LBM passes per HBM sweep (0 tuned, 1, 2 or 3), bits 7-8 "
                   "clones per workgroup (0 tuned, 1, or 2 with one pass per "
                   "sweep)", -1;
  }
  for (int i = 0; i < cfg->n_instances; i++) {
    const lbc_instance& s = cfg->instances[i];
    char buf[160];
    auto bad = [&](const char* m) {
      snprintf(buf, sizeof buf, "instance %d: %s", i, m);
      why = buf;
      return -1;
    };
    if (s.Lx < 8 || s.Ly < 8) return bad("lattice must be at least 8x8");
    if ((long long)s.Lx * s.Ly * 9 > (1LL << 30) || s.Lx > 32767 ||
        s.Ly > 32767)
      return bad("lattice too large");
    {
      // the step kernels address one clone through a buffer resource with
      // 32-bit byte offsets
      const long long cstride = ((long long)9 * s.Lx * s.Ly + 127) / 128 * 128;
      if (cstride * (long long)elem_bytes(cfg->storage) >= (1LL << 31))
        return bad("lattice too large for this storage: one clone must stay "
                   "below 2 GiB");
    }
    if (s.xmin < 2 || s.xmax > s.Lx - 3 || s.ymin < 2 || s.ymax > s.Ly - 3)
      return bad("square must keep two fluid nodes from every boundary");
    if (s.xmax - s.xmin < 2 || s.ymax - s.ymin < 2)
      return bad("square side must be at least 2");
    if (s.n_clones < 1 || s.n_clones > kMaxClones)
      return bad("n_clones must be in [1, 8192]");
    if (s.steps < 1) return bad("steps must be >= 1");
    if (!(s.tau > 0.5)) return bad("tau must exceed 0.5");
    if (s.rank_blocks < 1 || s.n_clones % s.rank_blocks)
      return bad("rank_blocks must divide n_clones");
    if (s.n_snapshots < 0 || s.n_snapshots > 64)
      return bad("n_snapshots must be in [0, 64]");
    if (!std::isfinite(s.force_scale) || !std::isfinite(s.force_shift) ||
        !std::isfinite(s.dt) || !std::isfinite(s.bias) ||
        !std::isfinite(s.beta) || !std::isfinite(s.u_in))
      return bad("non-finite parameter");
    if (s.forced_steps < 0 || s.forced_steps > s.steps)
      return bad("forced_steps must be in [0, steps]");
    if (s.macro_every < 0) return bad("macro_every must be >= 0");
    if (s.macro_every > 0 &&
        (long long)s.n_clones * ((s.steps + s.macro_every - 1) / s.macro_every) *
                3 * s.Lx * s.Ly > kMaxMacroDoubles)
      return bad("macroscopic output too large: raise macro_every");
    if (s.forced_steps > 0) {
      if (cfg->mode != LBC_MODE_REF)
        return bad("forced steps need LBC_MODE_REF (the body-force mode)");
      if (!std::isfinite(s.field_eps) || !std::isfinite(s.field_sigma) ||
          s.field_sigma < 0)
        return bad("field_eps must be finite and field_sigma finite and >= 0");
      if ((long long)(2 * s.Lx - 1) * (2 * s.Ly - 1) > kMaxFieldEntries)
        return bad("lattice too large for the random-field embedding");
    }
  }
  return 0;
}

template <class T>
int dalloc(lbc_ctx* c, T** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, LBC_E_NOMEM,
                std::string("hipMalloc(") + std::to_string(bytes) +
                    "): " + hipGetErrorString(e));
  }
  c->allocs.push_back((void*)*p);
  return LBC_OK;
}

lbc::Launch make_launch(lbc_ctx* c) {
  lbc::Launch L{};
  L.insts = c->idd;
  L.items = c->items;
  L.counters = c->counters;
  L.n_items = c->n_items;
  L.threads = c->threads;
  L.ring_max = c->ring_max;
  L.mode = c->cfg.mode;
  L.storage = c->cfg.storage;
  L.vec = c->vec;
  L.nt = c->nt;
  L.passes = c->passes;
  L.rch = c->rch;
  L.forced = c->forced ? 1 : 0;
  L.cpw = c->cpw;
  L.items2 = c->items2;
  L.n_items2 = c->n_items2;
  L.stream = c->stream;
  return L;
}

uint32_t seed_lo(const lbc_ctx* c) { return (uint32_t)(c->cfg.seed & 0xFFFFFFFFu); }
uint32_t seed_hi(const lbc_ctx* c) { return (uint32_t)(c->cfg.seed >> 32); }

int ensure_staging(lbc_ctx* c, size_t bytes) {
  if (bytes <= c->staging_bytes) return LBC_OK;
  if (c->staging) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
  }
  hipError_t e = hipMalloc((void**)&c->staging, bytes);
  if (e != hipSuccess) return fail(c, LBC_E_NOMEM, "staging allocation failed");
  c->staging_bytes = bytes;
  return LBC_OK;
}

#define FFTC(ctx, expr)                                                      \
  do {                                                                       \
    hipfftResult r_ = (expr);                                                \
    if (r_ != HIPFFT_SUCCESS)                                                \
      return fail((ctx), LBC_E_HIP,                                          \
                  std::string(#expr) + ": hipFFT error " +                   \
                      std::to_string((int)r_));                              \
  } while (0)

// Random-field set-up of forced instance i: map and FFT buffers, the square
// roots of the circulant embedding's eigenvalues, and the batched plan.
// generate_random_field.cpp:44-106 builds the wrapped covariance
// G[j + fmy*r] = f(min(r, fmx-r), min(j, fmy-j)) / (fmx*fmy) with
// f(x, y) = exp(-(x*x + y*y) / (2 sigma^2)) (covariance_fctn.cpp:3-10);
// :112-158 take Re of its forward 2-D DFT as the eigenvalues and zero the
// negative ones.  The test there, abs(a) > 1e-15, calls the int abs of
// <cstdlib> on the double, so every negative eigenvalue above -1 is zeroed
// and only a <= -1 is flagged (and turns the field into NaN); such an
// embedding is rejected here.
int field_setup(lbc_ctx* c, int i) {
  InstDev& d = c->idh[i];
  const lbc_instance& s = c->inst[i];
  const double sigma = s.field_sigma > 0 ? s.field_sigma : 32.0;
  const int fmx = 2 * s.Lx - 1, fmy = 2 * s.Ly - 1;
  const size_t fmn = (size_t)fmx * fmy;
  d.forced_steps = s.forced_steps;
  d.field_eps = s.field_eps;
  d.fmx = fmx;
  d.fmy = fmy;
  // The shifted storages collide with FP32 maps, so their transforms run in
  // single precision (C2C) at half the bytes; FP64 storage runs Z2Z.
  const bool f32 = c->cfg.storage != LBC_STORE_F64;
  const size_t per = 2 * (f32 ? sizeof(float) : sizeof(double)) * fmn;
  // One transform serves two clones (the real and imaginary parts).  The
  // plan runs its whole batch every time, so the transforms are spread
  // evenly over the fewest batches that fit the work buffer.
  const size_t transforms = ((size_t)s.n_clones + 1) / 2;
  const size_t fit = std::max<size_t>(1, kFieldWorkBytes / per);
  const size_t batches = (transforms + fit - 1) / fit;
  const int chunk = (int)((transforms + batches - 1) / batches);
  // the embedding's spectrum below is one FP64 transform in the same buffer
  const size_t work_bytes =
      std::max(per * chunk, 2 * sizeof(double) * fmn);
  c->fchunk[i] = chunk;
  const size_t map_bytes = sizeof(double) * (size_t)s.n_clones * d.plane;
  if (dalloc(c, &d.field_map, map_bytes) ||
      dalloc(c, &d.field_sqrt_ev, sizeof(double) * fmn) ||
      dalloc(c, &d.field_work, work_bytes))
    return LBC_E_NOMEM;
  HIPC(c, hipMemsetAsync(d.field_map, 0, map_bytes, c->stream));
  // the shifted storages collide in FP32 and read an FP32 copy of the maps:
  // half the bytes of each forced step's map read
  if (c->cfg.storage != LBC_STORE_F64) {
    if (dalloc(c, &d.field_map32, map_bytes / 2)) return LBC_E_NOMEM;
    HIPC(c, hipMemsetAsync(d.field_map32, 0, map_bytes / 2, c->stream));
  }
  HIPC(c, hipMemsetAsync(d.field_work, 0, work_bytes, c->stream));

  std::vector<double> G(2 * fmn, 0.0);
  const double N = (double)fmn;
  for (int r = 0; r < fmx; r++) {
    const int dx = std::min(r, fmx - r);
    for (int j = 0; j < fmy; j++) {
      const int dy = std::min(j, fmy - j);
      const double dd = (double)(dx * dx + dy * dy);
      G[2 * (j + (size_t)fmy * r)] = std::exp(-dd / (2 * sigma * sigma)) / N;
    }
  }
  auto* work = reinterpret_cast<hipfftDoubleComplex*>(d.field_work);
  HIPC(c, hipMemcpyAsync(work, G.data(), sizeof(double) * G.size(),
                         hipMemcpyHostToDevice, c->stream));
  hipfftHandle one = nullptr;
  FFTC(c, hipfftPlan2d(&one, fmx, fmy, HIPFFT_Z2Z));
  hipfftResult fr = hipfftSetStream(one, c->stream);
  if (fr == HIPFFT_SUCCESS) fr = hipfftExecZ2Z(one, work, work, HIPFFT_FORWARD);
  const hipError_t he = hipMemcpyAsync(G.data(), work, sizeof(double) * G.size(),
                                       hipMemcpyDeviceToHost, c->stream);
  const hipError_t hs = hipStreamSynchronize(c->stream);
  (void)hipfftDestroy(one);
  FFTC(c, fr);
  HIPC(c, he);
  HIPC(c, hs);
  std::vector<double> sq(fmn);
  for (size_t m = 0; m < fmn; m++) {
    double a = G[2 * m];
    if (a < 0) {
      if (std::trunc(-a) >= 1.0)
        return fail(c, LBC_E_ARG,
                    "random field: the circulant embedding has an eigenvalue "
                    "<= -1");
      a = 0.0;
    }
    sq[m] = std::sqrt(a);
  }
  HIPC(c, hipMemcpyAsync(d.field_sqrt_ev, sq.data(), sizeof(double) * fmn,
                         hipMemcpyHostToDevice, c->stream));
  int dims[2] = {fmx, fmy};
  FFTC(c, hipfftPlanMany(&c->fft[i], 2, dims, nullptr, 1, (int)fmn, nullptr, 1,
                         (int)fmn, f32 ? HIPFFT_C2C : HIPFFT_Z2Z, chunk));
  FFTC(c, hipfftSetStream(c->fft[i], c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

// Random fields of the coming interval for every forced instance: Philox
// normals, the batched FFT and the lattice block, chunk by chunk, each
// transform giving two clones' fields; an instance whose maps the caller set
// (lbc_set_field_maps) keeps them once.
int enqueue_fields(lbc_ctx* c) {
  for (int i = 0; i < c->cfg.n_instances; i++) {
    const InstDev& d = c->idh[i];
    if (d.forced_steps <= 0) continue;
    if (c->maps_set[i]) {
      c->maps_set[i] = 0;
      continue;
    }
    const bool f32 = c->cfg.storage != LBC_STORE_F64;
    // fchunk transforms per batch, two clones per transform
    const int per_batch = 2 * c->fchunk[i];
    for (int c0 = 0; c0 < d.n_clones; c0 += per_batch) {
      const int n = std::min(per_batch, d.n_clones - c0);
      HIPC(c, lbc::launch_field_draw(c->idd, c->counters, i, c0, n,
                                     d.fmx * d.fmy, seed_lo(c), seed_hi(c),
                                     f32, c->stream));
      if (f32) {
        auto* w = reinterpret_cast<hipfftComplex*>(d.field_work);
        FFTC(c, hipfftExecC2C(c->fft[i], w, w, HIPFFT_FORWARD));
      } else {
        auto* w = reinterpret_cast<hipfftDoubleComplex*>(d.field_work);
        FFTC(c, hipfftExecZ2Z(c->fft[i], w, w, HIPFFT_FORWARD));
      }
      HIPC(c, lbc::launch_field_extract(c->idd, i, c0, n, d.plane, f32,
                                        c->stream));
    }
  }
  return LBC_OK;
}

int check_inst(lbc_ctx* c, int inst) {
  if (!c) return fail(nullptr, LBC_E_ARG, "null context");
  if (inst < 0 || inst >= c->cfg.n_instances)
    return fail(c, LBC_E_ARG, "instance index out of range");
  return LBC_OK;
}

// Calls that read, replace or reselect the states refuse while copies wait
// in the next interval's pass 0 (lbc_gather_fused).
int check_no_pending(lbc_ctx* c) {
  if (c && c->pass0_pending)
    return fail(c, LBC_E_STATE,
                "the copies of lbc_gather_fused are pending in the next "
                "interval's pass 0: run an interval first");
  return LBC_OK;
}

// Buffer flips of one interval: one per HBM sweep (one per pass with one
// pass per sweep).
int interval_flip(const lbc_ctx* c, int i) {
  const int passes = c->inst[i].steps + 1;
  return c->passes > 1
             ? lbc::fused_sweeps(c->inst[i].steps, c->passes,
                                 c->inst[i].forced_steps,
                                 c->inst[i].macro_every) & 1
             : passes & 1;
}

int upload_flips(lbc_ctx* c) {
  std::vector<int> flips(lbc::kMaxInstances, 0);
  for (int i = 0; i < c->cfg.n_instances; i++) flips[i] = interval_flip(c, i);
  HIPC(c, hipMemcpy(c->flip_mask, flips.data(), sizeof(int) * flips.size(),
                    hipMemcpyHostToDevice));
  return LBC_OK;
}

// Step kernel + parity flip for every instance, bracketed by events.
int enqueue_interval(lbc_ctx* c, bool timed) {
  lbc::Launch L = make_launch(c);
  if (timed) {
    if (c->ev_used + 2 > (int)c->ev.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        HIPC(c, hipEventCreate(&e));
        c->ev.push_back(e);
      }
    }
    HIPC(c, hipEventRecord(c->ev[c->ev_used], c->stream));
  }
  HIPC(c, lbc::launch_interval(L));
  if (timed) {
    HIPC(c, hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    c->ev_used += 2;
  }
  HIPC(c, lbc::launch_flip(L, c->cfg.n_instances, c->flip_mask));
  for (int i = 0; i < c->cfg.n_instances; i++) c->parity[i] ^= interval_flip(c, i);
  return LBC_OK;
}

int enqueue_perturb(lbc_ctx* c, bool philox) {
  if (c->cfg.mode != LBC_MODE_REF) return LBC_OK;
  bool any = false;
  for (auto& s : c->inst) any |= s.n_snapshots > 0;
  if (!any) return LBC_OK;
  lbc::Launch L = make_launch(c);
  HIPC(c, lbc::launch_perturb(L, philox ? 1 : 0, seed_lo(c), seed_hi(c)));
  return LBC_OK;
}

int autotune(lbc_ctx* c) {
  // Time short intervals (at most 48 steps: long enough that the share of
  // the lone first and last passes is close to a full interval's, so the
  // ranking of passes per sweep holds) at each shape on the freshly
  // initialised state; the caller re-initialises the state after.
  std::vector<InstDev> tmp = c->idh;
  for (auto& d : tmp) {
    d.steps = std::min(d.steps, 48);
    d.forced_steps = std::min(d.forced_steps, d.steps);
    d.macro_every = 0;
  }
  HIPC(c, hipMemcpyAsync(c->idd, tmp.data(), sizeof(InstDev) * tmp.size(),
                         hipMemcpyHostToDevice, c->stream));
  const int tcands[3] = {256, 512, 1024};
  const int vfixed = c->cfg.flags & 7;
  const int ntmode = (c->cfg.flags >> 3) & 3;  // 0 tuned, 1 off, 2 on
  const int pmode = (c->cfg.flags >> 5) & 3;   // 0 tuned, else passes per sweep
  const int cpmode = (c->cfg.flags >> 7) & 3;  // 0 tuned, else clones per group
  float best = 1e30f;
  int best_t = 256, best_v = 1, best_nt = 0, best_p = 1, best_cpw = 1;
  hipEvent_t e0, e1;
  HIPC(c, hipEventCreate(&e0));
  HIPC(c, hipEventCreate(&e1));
  for (int passes = 1; passes <= 3; passes++)
  for (int nt = 0; nt < 2; nt++)
  for (int v : {1})
  for (int cpw = 1; cpw <= 2; cpw++) {
    if (vfixed && v != vfixed) continue;
    if ((ntmode == 1 && nt) || (ntmode == 2 && !nt)) continue;
    if (pmode && passes != pmode) continue;
    // two clones per workgroup: the one-pass kernel only (the multi-pass
    // kernels' LDS rings are per clone, so two clones need twice the LDS
    // of one and no more waves fit per CU)
    if (cpw == 2 && passes != 1) continue;
    if (cpmode && cpw != cpmode) continue;
    for (int t : tcands) {
      if (c->cfg.threads && t != c->cfg.threads) continue;
      if (cpw * t > 1024) continue;
      const int rch = lbc::fused_ring_sites(t, c->max_lx, c->cfg.mode);
      if (passes > 1 &&
          !lbc::fused_lds_bytes(c->cfg.storage, c->ring_max, rch, passes))
        continue;
      lbc::Launch L = make_launch(c);
      L.threads = t;
      L.vec = v;
      L.nt = nt;
      L.passes = passes;
      L.rch = rch;
      L.cpw = cpw;
      // a shape the device refuses to launch (a configuration error,
      // reported at launch) is skipped; the others are still timed
      if (lbc::launch_interval(L) != hipSuccess) {  // warm-up
        (void)hipGetLastError();
        continue;
      }
      HIPC(c, hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; r++) HIPC(c, lbc::launch_interval(L));
      HIPC(c, hipEventRecord(e1, c->stream));
      HIPC(c, hipEventSynchronize(e1));
      float ms = 0;
      HIPC(c, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_t = t;
        best_v = v;
        best_nt = nt;
        best_p = passes;
        best_cpw = cpw;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best >= 1e30f)
    return fail(c, LBC_E_HIP, "autotune: no step-kernel shape could be launched");
  c->threads = best_t;
  c->vec = best_v;
  c->nt = best_nt;
  c->passes = best_p;
  c->cpw = best_cpw;
  c->rch = lbc::fused_ring_sites(best_t, c->max_lx, c->cfg.mode);
  HIPC(c, hipMemcpyAsync(c->idd, c->idh.data(), sizeof(InstDev) * c->idh.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

}  // namespace

extern "C" {

const char* lbc_last_error(const lbc_ctx* ctx) {
  if (ctx && !ctx->err.empty()) return ctx->err.c_str();
  if (g_err_static) return g_err_static;
  return g_err.c_str();
}

int lbc_threads(const lbc_ctx* ctx) { return ctx ? ctx->threads : 0; }

int lbc_sites_per_thread(const lbc_ctx* ctx) { return ctx ? ctx->vec : 0; }

int lbc_step_shape(const lbc_ctx* ctx) {
  if (!ctx) return 0;
  return ctx->vec | ((ctx->nt ? 2 : 1) << 3) | (ctx->passes << 5) |
         (ctx->cpw << 7);
}

void lbc_destroy(lbc_ctx* c) {
  if (!c) return;
  ON_CTX_DEVICE(c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 3; i++) {
    if (c->gexec[i]) (void)hipGraphExecDestroy(c->gexec[i]);
    if (c->graphs[i]) (void)hipGraphDestroy(c->graphs[i]);
  }
  for (auto e : c->ev) (void)hipEventDestroy(e);
  for (auto e : c->gev)
    if (e) (void)hipEventDestroy(e);
  for (auto p : c->fft)
    if (p) (void)hipfftDestroy(p);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->staging) (void)hipFree(c->staging);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int lbc_create(const lbc_config* cfg, lbc_ctx** out) {
  if (!out) return fail(nullptr, LBC_E_ARG, "null output pointer");
  *out = nullptr;
  std::string why;
  if (validate(cfg, why)) return fail(nullptr, LBC_E_ARG, why);
  lbc_ctx* c = new lbc_ctx();
  c->cfg = *cfg;
  c->inst.assign(cfg->instances, cfg->instances + cfg->n_instances);
  c->cfg.instances = c->inst.data();
  auto bail = [&](int code) {
    lbc_destroy(c);
    return code;
  };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return bail(fail(nullptr, LBC_E_HIP, "no HIP device available"));
  if (cfg->device < 0 || cfg->device >= ndev)
    return bail(fail(nullptr, LBC_E_ARG, "device ordinal out of range"));
  DeviceScope scope(cfg->device);
  if (!scope.ok()) return bail(fail(nullptr, LBC_E_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(nullptr, LBC_E_HIP, "hipStreamCreate failed"));

  const int ni = cfg->n_instances;
  const size_t eb = elem_bytes(cfg->storage);
  c->idh.resize(ni);
  c->parity.assign(ni, 0);
  c->selected.assign(ni, 0);
  c->fft.assign(ni, nullptr);
  c->fchunk.assign(ni, 0);
  c->maps_set.assign(ni, 0);
  c->last_copies.assign(ni, 0);
  int clone_base = 0;
  for (int i = 0; i < ni; i++) {
    const lbc_instance& s = c->inst[i];
    InstDev& d = c->idh[i];
    std::memset(&d, 0, sizeof d);
    d.Lx = s.Lx;
    d.Ly = s.Ly;
    d.xmin = s.xmin;
    d.xmax = s.xmax;
    d.ymin = s.ymin;
    d.ymax = s.ymax;
    d.plane = s.Lx * s.Ly;
    // clones start on 128-element boundaries
    d.clone_stride = ((long long)9 * d.plane + 127) / 128 * 128;
    d.n_clones = s.n_clones;
    d.steps = s.steps;
    d.ring_n = 2 * (s.xmax - s.xmin + 1) + 2 * (s.ymax - s.ymin + 1);
    d.n_snapshots = cfg->mode == LBC_MODE_REF ? s.n_snapshots : 0;
    d.rank_blocks = s.rank_blocks;
    d.clone_base = clone_base;
    clone_base += s.n_clones;
    d.tau = s.tau;
    d.beta = cfg->mode == LBC_MODE_REF ? s.beta : 0.0;
    d.u_in = cfg->mode == LBC_MODE_CHANNEL ? s.u_in : 0.0;
    d.scale = s.force_scale;
    d.shift = s.force_shift;
    d.dt = s.dt;
    d.bias = s.bias;
    d.eps = s.eps;
    d.sigma = s.noise_sigma;
    d.inlet_k = 1.0 / (1.0 - d.u_in);
    d.max_intervals = kMaxIntervals;
    c->ring_max = std::max(c->ring_max, d.ring_n);
    c->max_lx = std::max(c->max_lx, d.Lx);
    c->n_max = std::max(c->n_max, s.n_clones);
    const size_t n = (size_t)s.n_clones;
    int rc = LBC_OK;
    for (int b = 0; b < 2 && rc == LBC_OK; b++)
      rc = dalloc(c, &d.buf[b], (size_t)d.clone_stride * n * eb);
    if (rc == LBC_OK) rc = dalloc(c, &d.drag, sizeof(double) * n * s.steps);
    if (rc == LBC_OK) rc = dalloc(c, &d.obs, sizeof(double) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.logw, sizeof(double) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.weights, sizeof(double) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.u, sizeof(double) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.rint, sizeof(int32_t) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.nb_before, sizeof(int32_t) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.nb_after, sizeof(int32_t) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.pairs, sizeof(int32_t) * 2 * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.labels, sizeof(int32_t) * n);
    if (rc == LBC_OK) rc = dalloc(c, &d.sel_info, sizeof(int32_t) * 4);
    if (rc == LBC_OK) rc = dalloc(c, &d.sel_R, sizeof(double) * 2);
    if (rc == LBC_OK)
      rc = dalloc(c, &d.R_hist, sizeof(double) * 2 * (size_t)kMaxIntervals);
    if (rc == LBC_OK) rc = dalloc(c, &d.marks, n);
    if (rc == LBC_OK)
      rc = dalloc(c, &d.prefactors,
                  sizeof(double) * n * std::max(1, d.n_snapshots));
    if (rc == LBC_OK)
      rc = dalloc(c, &d.snapshots,
                  sizeof(double) * 9 * (size_t)d.plane * d.n_snapshots);
    if (rc == LBC_OK) rc = dalloc(c, &d.norms, sizeof(double) * 2 * n);
    if (rc != LBC_OK) return bail(rc);
    if (hipMemsetAsync(d.marks, 0, n, c->stream) != hipSuccess ||
        hipMemsetAsync(d.drag, 0, sizeof(double) * n * s.steps, c->stream) !=
            hipSuccess)
      return bail(fail(c, LBC_E_HIP, "hipMemsetAsync failed"));
    std::vector<int32_t> ident(n);
    for (size_t j = 0; j < n; j++) ident[j] = (int32_t)j;
    if (hipMemcpy(d.labels, ident.data(), sizeof(int32_t) * n,
                  hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(c, LBC_E_HIP, "label init failed"));
    // wave table: two words per block of 64 sites, plus two chunks' worth
    // of blocks past the lattice (the last, partial chunk and the step
    // kernel's one-chunk lookahead)
    std::vector<uint32_t> waves(2 * ((d.plane + 63) / 64 + 2 * 1024 / 64), 0u);
    for (size_t b = 0; 2 * b < waves.size(); b++) {
      if ((long long)b * 64 < d.plane)
        waves[2 * b] = lbc::wave_word(d.Lx, d.Ly, d.xmin, d.xmax, d.ymin,
                                      d.ymax, 64 * (int)b);
      waves[2 * b + 1] = lbc::wave_lanes(d.Lx, d.Ly, d.xmin, d.xmax, d.ymin,
                                         d.ymax, d.plane, 64 * (int)b);
    }
    uint32_t* wt = nullptr;
    if (dalloc(c, &wt, sizeof(uint32_t) * waves.size()))
      return bail(LBC_E_NOMEM);
    if (hipMemcpy(wt, waves.data(), sizeof(uint32_t) * waves.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(c, LBC_E_HIP, "wave table upload failed"));
    d.wave_info = wt;
    if (s.forced_steps > 0) {
      c->forced = true;
      if (int rc2 = field_setup(c, i)) return bail(rc2);
    }
    if (s.macro_every > 0) {
      d.macro_every = s.macro_every;
      d.macro_frames = (s.steps + s.macro_every - 1) / s.macro_every;
      const size_t mb = sizeof(double) * 3 * (size_t)d.plane * d.macro_frames * n;
      if (dalloc(c, &d.macro, mb)) return bail(LBC_E_NOMEM);
      if (hipMemsetAsync(d.macro, 0, mb, c->stream) != hipSuccess)
        return bail(fail(c, LBC_E_HIP, "hipMemsetAsync failed"));
    }
  }
  if (dalloc(c, &c->idd, sizeof(InstDev) * ni) ||
      dalloc(c, &c->counters, sizeof(Counters)) ||
      dalloc(c, &c->flip_mask, sizeof(int) * lbc::kMaxInstances) ||
      dalloc(c, &c->scgf, sizeof(double) * (size_t)kMaxIntervals))
    return bail(LBC_E_NOMEM);
  // work items: biggest lattices first so the tail of the launch is short
  std::vector<int> order(ni);
  for (int i = 0; i < ni; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)c->idh[a].plane * c->idh[a].steps >
           (long long)c->idh[b].plane * c->idh[b].steps;
  });
  std::vector<WorkItem> items;
  for (int i : order)
    for (int j = 0; j < c->inst[i].n_clones; j++) items.push_back({i, j});
  c->n_items = (int)items.size();
  if (dalloc(c, &c->items, sizeof(WorkItem) * items.size())) return bail(LBC_E_NOMEM);
  // the same clones in pairs of one instance for two clones per workgroup
  // (both halves then make the same passes); an odd clone count leaves the
  // last clone with a padding partner {inst, -1}
  std::vector<WorkItem> items2;
  for (int i : order) {
    const int n = c->inst[i].n_clones;
    for (int j = 0; j < n; j++) items2.push_back({i, j});
    if (n & 1) items2.push_back({i, -1});
  }
  c->n_items2 = (int)items2.size();
  if (dalloc(c, &c->items2, sizeof(WorkItem) * items2.size()))
    return bail(LBC_E_NOMEM);
  Counters zero{};
  if (hipMemcpy(c->idd, c->idh.data(), sizeof(InstDev) * ni,
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->items, items.data(), sizeof(WorkItem) * items.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->items2, items2.data(), sizeof(WorkItem) * items2.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->counters, &zero, sizeof zero, hipMemcpyHostToDevice) !=
          hipSuccess)
    return bail(fail(c, LBC_E_HIP, "descriptor upload failed"));

  // valid initial state everywhere: equilibrium of the mode's inflow
  lbc::Launch L = make_launch(c);
  for (int i = 0; i < ni; i++) {
    if (lbc::launch_init_eq(L, i, c->inst[i].n_clones, 0.0, 0, 0) != hipSuccess)
      return bail(fail(c, LBC_E_HIP, "init kernel launch failed"));
  }
  const int vfixed = cfg->flags & 7;
  const int ntmode = (cfg->flags >> 3) & 3;
  const int pmode = (cfg->flags >> 5) & 3;
  if (cfg->threads == 0 || vfixed == 0 || ntmode == 0 || pmode == 0) {
    int rc = autotune(c);
    if (rc) return bail(rc);
    L = make_launch(c);
    for (int i = 0; i < ni; i++)
      if (lbc::launch_init_eq(L, i, c->inst[i].n_clones, 0.0, 0, 0) != hipSuccess)
        return bail(fail(c, LBC_E_HIP, "init kernel launch failed"));
  } else {
    c->threads = cfg->threads;
    c->vec = vfixed;
    c->nt = ntmode == 2;
    c->passes = pmode;
    c->cpw = ((cfg->flags >> 7) & 3) == 2 ? 2 : 1;
    if (c->cpw == 2 && 2 * c->threads > 1024)
      return bail(fail(c, LBC_E_ARG,
                       "step kernel: two clones per workgroup take at most "
                       "512 threads each"));
    c->rch = lbc::fused_ring_sites(c->threads, c->max_lx, c->cfg.mode);
    if (c->passes > 1 &&
        !lbc::fused_lds_bytes(cfg->storage, c->ring_max, c->rch, c->passes))
      return bail(fail(c, LBC_E_ARG,
                       "step kernel: the LDS rings of this many passes per "
                       "sweep do not fit with these threads and storage"));
  }
  if (int rc = upload_flips(c)) return bail(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(c, LBC_E_HIP, "initialisation failed"));
  *out = c;
  return LBC_OK;
}

int lbc_load_states(lbc_ctx* c, int inst, int clone0, int n, const double* aos) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (int rc = check_no_pending(c)) return rc;
  const lbc_instance& s = c->inst[inst];
  if (!aos || n < 1 || clone0 < 0 || clone0 + n > s.n_clones)
    return fail(c, LBC_E_ARG, "clone range out of bounds");
  const size_t bytes = sizeof(double) * 9 * (size_t)s.Lx * s.Ly * n;
  if (int rc = ensure_staging(c, bytes)) return rc;
  HIPC(c, hipMemcpyAsync(c->staging, aos, bytes, hipMemcpyHostToDevice, c->stream));
  HIPC(c, lbc::launch_aos_to_soa(c->cfg.storage, c->idd, c->counters, inst, 0,
                                 clone0, n, c->staging, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

static int store_which(lbc_ctx* c, int inst, int which, int clone0, int n,
                       double* aos) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (int rc = check_no_pending(c)) return rc;
  const lbc_instance& s = c->inst[inst];
  if (!aos || n < 1 || clone0 < 0 || clone0 + n > s.n_clones)
    return fail(c, LBC_E_ARG, "clone range out of bounds");
  const size_t bytes = sizeof(double) * 9 * (size_t)s.Lx * s.Ly * n;
  if (int rc = ensure_staging(c, bytes)) return rc;
  HIPC(c, lbc::launch_soa_to_aos(c->cfg.storage, c->idd, c->counters, inst,
                                 which, clone0, n, c->staging, c->stream));
  HIPC(c, hipMemcpyAsync(aos, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_store_states(lbc_ctx* c, int inst, int clone0, int n, double* aos) {
  return store_which(c, inst, 0, clone0, n, aos);
}

int lbc_init_equilibrium(lbc_ctx* c, int inst, double amp) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (int rc = check_no_pending(c)) return rc;
  if (!std::isfinite(amp) || amp < 0) return fail(c, LBC_E_ARG, "bad amplitude");
  lbc::Launch L = make_launch(c);
  HIPC(c, lbc::launch_init_eq(L, inst, c->inst[inst].n_clones, amp, seed_lo(c),
                              seed_hi(c)));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_set_snapshots(lbc_ctx* c, int inst, const double* aos) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  const InstDev& d = c->idh[inst];
  if (d.n_snapshots <= 0)
    return fail(c, LBC_E_STATE, "instance has no snapshot slots (REF mode, n_snapshots > 0)");
  if (!aos) return fail(c, LBC_E_ARG, "null snapshots");
  const size_t bytes = sizeof(double) * 9 * (size_t)d.plane * d.n_snapshots;
  if (int rc = ensure_staging(c, bytes)) return rc;
  HIPC(c, hipMemcpyAsync(c->staging, aos, bytes, hipMemcpyHostToDevice, c->stream));
  HIPC(c, lbc::launch_snap_to_soa(c->idd, inst, d.n_snapshots, c->staging, d.Lx,
                                  d.Ly, d.snapshots, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_set_marks(lbc_ctx* c, int inst, const uint8_t* marks) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (!marks) return fail(c, LBC_E_ARG, "null marks");
  HIPC(c, hipMemcpyAsync(c->idh[inst].marks, marks, c->inst[inst].n_clones,
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_get_marks(lbc_ctx* c, int inst, uint8_t* marks) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (!marks) return fail(c, LBC_E_ARG, "null marks");
  HIPC(c, hipMemcpyAsync(marks, c->idh[inst].marks, c->inst[inst].n_clones,
                         hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_run_interval(lbc_ctx* c, const double* prefactors) {
  ON_CTX_DEVICE(c);
  if (!c) return fail(nullptr, LBC_E_ARG, "null context");
  bool philox = true;
  if (prefactors && c->cfg.mode == LBC_MODE_REF) {
    philox = false;
    size_t off = 0;
    for (int i = 0; i < c->cfg.n_instances; i++) {
      const InstDev& d = c->idh[i];
      const size_t cnt = (size_t)d.n_clones * d.n_snapshots;
      if (cnt)
        HIPC(c, hipMemcpyAsync(d.prefactors, prefactors + off,
                               sizeof(double) * cnt, hipMemcpyHostToDevice,
                               c->stream));
      off += cnt;
    }
  }
  c->ev_used = 0;
  c->timing_from_graph = false;
  if (int rc = enqueue_perturb(c, philox)) return rc;
  if (int rc = enqueue_fields(c)) return rc;
  if (int rc = enqueue_interval(c, true)) return rc;
  c->pass0_pending = false;
  return LBC_OK;
}

int lbc_get_drag(lbc_ctx* c, int inst, double* A) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (!A) return fail(c, LBC_E_ARG, "null output");
  const InstDev& d = c->idh[inst];
  HIPC(c, hipMemcpyAsync(A, d.drag, sizeof(double) * d.n_clones * d.steps,
                         hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_get_logweights(lbc_ctx* c, int inst, double* obs, double* logw) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  const InstDev& d = c->idh[inst];
  if (obs)
    HIPC(c, hipMemcpyAsync(obs, d.obs, sizeof(double) * d.n_clones,
                           hipMemcpyDeviceToHost, c->stream));
  if (logw)
    HIPC(c, hipMemcpyAsync(logw, d.logw, sizeof(double) * d.n_clones,
                           hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_select(lbc_ctx* c, int inst, const double* weights, const double* u,
               const int32_t* rint, lbc_selection* out) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (int rc = check_no_pending(c)) return rc;
  const InstDev& d = c->idh[inst];
  const size_t n = (size_t)d.n_clones;
  if ((u == nullptr) != (rint == nullptr))
    return fail(c, LBC_E_ARG, "u and rint must both be given or both be NULL");
  if (weights)
    HIPC(c, hipMemcpyAsync(d.weights, weights, sizeof(double) * n,
                           hipMemcpyHostToDevice, c->stream));
  if (u) {
    for (size_t j = 0; j < n; j++)
      if (rint[j] < 0) return fail(c, LBC_E_ARG, "rint must be non-negative");
    HIPC(c, hipMemcpyAsync(d.u, u, sizeof(double) * n, hipMemcpyHostToDevice,
                           c->stream));
    HIPC(c, hipMemcpyAsync(d.rint, rint, sizeof(int32_t) * n,
                           hipMemcpyHostToDevice, c->stream));
  }
  // one-instance launch through a descriptor pointer offset
  HIPC(c, lbc::launch_select(c->idd, c->counters, inst, 1, d.n_clones,
                             weights ? 1 : 0, u ? 1 : 0, seed_lo(c), seed_hi(c),
                             c->stream));
  c->selected[inst] = 1;
  int32_t info[4];
  double R[2];
  HIPC(c, hipMemcpyAsync(info, d.sel_info, sizeof info, hipMemcpyDeviceToHost,
                         c->stream));
  HIPC(c, hipMemcpyAsync(R, d.sel_R, sizeof R, hipMemcpyDeviceToHost, c->stream));
  if (out) {
    if (out->nb_before)
      HIPC(c, hipMemcpyAsync(out->nb_before, d.nb_before, 4 * n,
                             hipMemcpyDeviceToHost, c->stream));
    if (out->nb_after)
      HIPC(c, hipMemcpyAsync(out->nb_after, d.nb_after, 4 * n,
                             hipMemcpyDeviceToHost, c->stream));
    if (out->labels)
      HIPC(c, hipMemcpyAsync(out->labels, d.labels, 4 * n,
                             hipMemcpyDeviceToHost, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->last_copies[inst] = info[0];
  if (out) {
    out->n_pairs = info[0];
    out->n_rint_used = info[1];
    out->total_R = R[0];
    out->log_total_R = R[1];
    if (out->pairs && info[0] > 0) {
      HIPC(c, hipMemcpy(out->pairs, d.pairs, sizeof(int32_t) * 2 * info[0],
                        hipMemcpyDeviceToHost));
    }
  }
  if (info[2] != 0)
    return fail(c, LBC_E_NUMERIC,
                "selection: non-finite or out-of-range weights");
  return LBC_OK;
}

static int gather_impl(lbc_ctx* c, bool fuse0) {
  ON_CTX_DEVICE(c);
  if (!c) return fail(nullptr, LBC_E_ARG, "null context");
  if (int rc = check_no_pending(c)) return rc;
  for (int i = 0; i < c->cfg.n_instances; i++)
    if (!c->selected[i])
      return fail(c, LBC_E_STATE, "lbc_gather: every instance needs lbc_select first");
  if (!c->gev[0]) {
    hipEvent_t e[2] = {nullptr, nullptr};
    for (int k = 0; k < 2; k++) {
      const hipError_t he = hipEventCreate(&e[k]);
      if (he != hipSuccess) {
        if (e[0]) (void)hipEventDestroy(e[0]);
        return fail(c, LBC_E_HIP,
                    std::string("hipEventCreate: ") + hipGetErrorString(he));
      }
    }
    c->gev[0] = e[0];
    c->gev[1] = e[1];
  }
  lbc::Launch L = make_launch(c);
  HIPC(c, hipEventRecord(c->gev[0], c->stream));
  HIPC(c, lbc::launch_gather(L, seed_lo(c), seed_hi(c), fuse0 ? 1 : 0));
  HIPC(c, hipEventRecord(c->gev[1], c->stream));
  // every copy reads the parent's populations and writes the destination's
  double bytes = 0.0;
  for (int i = 0; i < c->cfg.n_instances; i++)
    bytes += 2.0 * 9.0 * c->idh[i].plane * (double)elem_bytes(c->cfg.storage) *
             c->last_copies[i];
  c->gather_bytes = bytes;
  c->gather_timed = true;
  for (int i = 0; i < c->cfg.n_instances; i++) c->selected[i] = 0;
  if (fuse0 && c->cfg.mode == LBC_MODE_CHANNEL) c->pass0_pending = true;
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_gather(lbc_ctx* c) { return gather_impl(c, false); }

int lbc_gather_fused(lbc_ctx* c) { return gather_impl(c, true); }

namespace {

// Capture one cloning step of every instance as graphs: [step kernel] and
// [parity flip, selection, gather, perturbation of the new copies (REF)],
// the latter twice: with the copies stored (graph 1) and, for the channel
// mode's noisy copies, with them collided into the next interval's pass 0
// (graph 2; k_gather's fuse0), which only a following step kernel may
// replay after.  Ordinary events recorded between the replays time the
// step kernel alone.
int capture_step(lbc_ctx* c) {
  lbc::Launch L = make_launch(c);
  for (int part = 0; part < 3; part++) {
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    bool ok = true;
    if (part == 0) {
      ok = lbc::launch_interval(L) == hipSuccess;
    } else {
      ok = lbc::launch_flip(L, c->cfg.n_instances, c->flip_mask) == hipSuccess;
      if (ok)
        ok = lbc::launch_select(c->idd, c->counters, 0, c->cfg.n_instances,
                                c->n_max, 0, 0, seed_lo(c), seed_hi(c),
                                c->stream) == hipSuccess;
      if (ok)
        ok = lbc::launch_gather(L, seed_lo(c), seed_hi(c), part == 2) ==
             hipSuccess;
      bool snapshots = false;
      for (const auto& s : c->inst) snapshots |= s.n_snapshots > 0;
      if (ok && c->cfg.mode == LBC_MODE_REF && snapshots)
        ok = lbc::launch_perturb(L, 1, seed_lo(c), seed_hi(c)) == hipSuccess;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (!ok || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return fail(c, LBC_E_HIP, "graph capture failed");
    }
    hipError_t ei = hipGraphInstantiate(&c->gexec[part], g, nullptr, nullptr, 0);
    c->graphs[part] = g;
    if (ei != hipSuccess) {
      c->gexec[part] = nullptr;
      return fail(c, LBC_E_HIP, "graph instantiation failed");
    }
  }
  return LBC_OK;
}

}  // namespace

int lbc_advance(lbc_ctx* c, int n_intervals) {
  ON_CTX_DEVICE(c);
  if (!c) return fail(nullptr, LBC_E_ARG, "null context");
  if (n_intervals < 0) return fail(c, LBC_E_ARG, "negative interval count");
  if (c->pass0_pending && n_intervals == 0)
    return check_no_pending(c);
  if (!c->gexec[0] || !c->gexec[1] || !c->gexec[2])
    if (int rc = capture_step(c)) return rc;
  c->timing_from_graph = false;
  c->ev_used = 0;
  if (n_intervals > 0)
    if (int rc = enqueue_perturb(c, true)) return rc;
  for (int k = 0; k < n_intervals; k++) {
    if (c->ev_used + 2 > (int)c->ev.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        HIPC(c, hipEventCreate(&e));
        c->ev.push_back(e);
      }
    }
    if (c->forced)
      if (int rc = enqueue_fields(c)) return rc;
    HIPC(c, hipEventRecord(c->ev[c->ev_used], c->stream));
    HIPC(c, hipGraphLaunch(c->gexec[0], c->stream));
    HIPC(c, hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    // the last interval's copies are stored, so the states are whole when
    // the call returns; the others go into the next interval's pass 0
    HIPC(c, hipGraphLaunch(c->gexec[k + 1 < n_intervals ? 2 : 1], c->stream));
    c->ev_used += 2;
    for (int i = 0; i < c->cfg.n_instances; i++)
      c->parity[i] ^= interval_flip(c, i);
    c->pass0_pending = false;
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_get_R_history(lbc_ctx* c, int inst, double* R, int max_n, int* n_out) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  Counters h;
  HIPC(c, hipMemcpyAsync(&h, c->counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  const int n = std::min(h.interval[inst], kMaxIntervals);
  if (n_out) *n_out = n;
  if (R && max_n > 0) {
    std::vector<double> tmp(2 * (size_t)std::min(n, max_n));
    if (!tmp.empty()) {
      HIPC(c, hipMemcpy(tmp.data(), c->idh[inst].R_hist,
                        sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
      for (size_t t = 0; t < tmp.size() / 2; t++) R[t] = tmp[2 * t];
    }
  }
  return LBC_OK;
}

int lbc_get_scgf(lbc_ctx* c, int inst, double dT, double* lam, int max_n,
                 int* n_out) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  if (!(dT > 0.0)) return fail(c, LBC_E_ARG, "dT must be positive");
  Counters h;
  HIPC(c, hipMemcpyAsync(&h, c->counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  const int n = std::min(h.interval[inst], kMaxIntervals);
  if (n_out) *n_out = n;
  if (lam && max_n > 0 && n > 0) {
    HIPC(c, lbc::launch_scgf(c->idd, inst, n, dT, c->scgf, c->stream));
    HIPC(c, hipMemcpyAsync(lam, c->scgf, sizeof(double) * std::min(n, max_n),
                           hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
  }
  return LBC_OK;
}

int lbc_set_field_maps(lbc_ctx* c, int inst, const double* maps) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  const InstDev& d = c->idh[inst];
  if (d.forced_steps <= 0)
    return fail(c, LBC_E_STATE, "instance has no forced steps");
  if (!maps) return fail(c, LBC_E_ARG, "null maps");
  const size_t plane = (size_t)d.plane;
  std::vector<double> tmp(plane * d.n_clones);
  for (int j = 0; j < d.n_clones; j++)
    for (int x = 0; x < d.Lx; x++)
      for (int y = 0; y < d.Ly; y++)
        tmp[j * plane + (size_t)y * d.Lx + x] =
            maps[j * plane + y + (size_t)d.Ly * x];
  // the shifted storages step with the FP32 copy: keep its widening, so
  // that lbc_get_field_maps reports the values the interval runs with
  if (d.field_map32)
    for (double& v : tmp) v = (double)(float)v;
  HIPC(c, hipMemcpyAsync(d.field_map, tmp.data(), sizeof(double) * tmp.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, lbc::launch_field_narrow(c->idd, inst, d.n_clones, d.plane,
                                   d.field_map32 != nullptr, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->maps_set[inst] = 1;
  return LBC_OK;
}

int lbc_get_field_maps(lbc_ctx* c, int inst, double* maps) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  const InstDev& d = c->idh[inst];
  if (d.forced_steps <= 0)
    return fail(c, LBC_E_STATE, "instance has no forced steps");
  if (!maps) return fail(c, LBC_E_ARG, "null maps");
  const size_t plane = (size_t)d.plane;
  std::vector<double> tmp(plane * d.n_clones);
  HIPC(c, hipMemcpyAsync(tmp.data(), d.field_map, sizeof(double) * tmp.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < d.n_clones; j++)
    for (int x = 0; x < d.Lx; x++)
      for (int y = 0; y < d.Ly; y++)
        maps[j * plane + y + (size_t)d.Ly * x] =
            tmp[j * plane + (size_t)y * d.Lx + x];
  return LBC_OK;
}

int lbc_get_macro(lbc_ctx* c, int inst, int clone0, int n, double* out) {
  ON_CTX_DEVICE(c);
  if (int rc = check_inst(c, inst)) return rc;
  const InstDev& d = c->idh[inst];
  if (d.macro_every <= 0)
    return fail(c, LBC_E_STATE, "instance has no macroscopic output (macro_every = 0)");
  if (!out || n < 1 || clone0 < 0 || clone0 + n > d.n_clones)
    return fail(c, LBC_E_ARG, "clone range out of bounds");
  const size_t per = 3 * (size_t)d.plane * d.macro_frames;
  HIPC(c, hipMemcpyAsync(out, d.macro + per * clone0, sizeof(double) * per * n,
                         hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_sync(lbc_ctx* c) {
  ON_CTX_DEVICE(c);
  if (!c) return fail(nullptr, LBC_E_ARG, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  return LBC_OK;
}

int lbc_step_kernel_ms(lbc_ctx* c, double* avg_ms, int* n_launches) {
  ON_CTX_DEVICE(c);
  if (!c || !avg_ms) return fail(c, LBC_E_ARG, "null argument");
  HIPC(c, hipStreamSynchronize(c->stream));
  double tot = 0;
  int n = 0;
  if (c->timing_from_graph) {
    for (float ms : c->graph_ms) {
      tot += ms;
      n++;
    }
  } else {
    for (int i = 0; i + 1 < c->ev_used; i += 2) {
      float ms = 0;
      HIPC(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
      tot += ms;
      n++;
    }
  }
  *avg_ms = n ? tot / n : 0.0;
  if (n_launches) *n_launches = n;
  return LBC_OK;
}

int lbc_gather_kernel_ms(lbc_ctx* c, double* ms, double* bytes) {
  ON_CTX_DEVICE(c);
  if (!c || !ms) return fail(c, LBC_E_ARG, "null argument");
  if (!c->gather_timed) return fail(c, LBC_E_STATE, "no lbc_gather call to time");
  HIPC(c, hipEventSynchronize(c->gev[1]));
  float t = 0.0f;
  HIPC(c, hipEventElapsedTime(&t, c->gev[0], c->gev[1]));
  *ms = t;
  if (bytes) *bytes = c->gather_bytes;
  return LBC_OK;
}

int lbc_select_static(int n, int rank_blocks, const double* logw,
                      const double* u, const int32_t* rint, int32_t* nb_before,
                      int32_t* nb_after, int32_t* pairs, int32_t* labels,
                      uint8_t* marks, int32_t* info, double* R, void* stream) {
  // host-side checks first: nothing is launched for a rejected call
  if (n < 1 || n > LBC_STATIC_MAX_CLONES)
    return fail_static(LBC_E_ARG, "lbc_select_static: n must be in "
                                  "[1, LBC_STATIC_MAX_CLONES]");
  if (rank_blocks < 1 || n % rank_blocks != 0)
    return fail_static(LBC_E_ARG, "lbc_select_static: rank_blocks must be "
                                  "positive and divide n");
  if (!logw || !u || !rint || !nb_before || !nb_after || !pairs || !labels ||
      !marks || !info || !R)
    return fail_static(LBC_E_ARG, "lbc_select_static: null device array");
  void* w = nullptr;
  if (lbc::static_scratch(&w) != hipSuccess)
    return fail_static(LBC_E_HIP, "lbc_select_static: scratch address lookup "
                                  "failed");
  lbc::InstDev I{};
  I.n_clones = n;
  I.rank_blocks = rank_blocks;
  I.logw = const_cast<double*>(logw);
  I.weights = static_cast<double*>(w);
  // read only: the kernel writes the draws only when it draws them itself
  I.u = const_cast<double*>(u);
  I.rint = const_cast<int32_t*>(rint);
  I.nb_before = nb_before;
  I.nb_after = nb_after;
  I.pairs = pairs;
  I.labels = labels;
  I.marks = marks;
  I.sel_info = info;
  I.sel_R = R;
  I.max_intervals = 0;  // no R history
  if (lbc::launch_select_static(I, static_cast<hipStream_t>(stream)) !=
      hipSuccess)
    return fail_static(LBC_E_HIP, "lbc_select_static: launch failed");
  return LBC_OK;
}

int lbc_step_host(int mode, const lbc_instance* geom, double* fin, double* fout,
                  double* force) {
  if (!geom || !fin || !fout) return fail(nullptr, LBC_E_ARG, "null argument");
  lbc_instance s = *geom;
  s.n_clones = 1;
  s.steps = 1;
  s.force_scale = 1.0;
  s.force_shift = 0.0;
  s.dt = 1.0;
  s.rank_blocks = 1;
  s.n_snapshots = 0;
  s.forced_steps = 0;
  s.macro_every = 0;
  lbc_config cfg{};
  cfg.mode = mode;
  cfg.storage = LBC_STORE_F64;
  cfg.n_instances = 1;
  cfg.threads = 256;
  cfg.flags = 1 | (1 << 3) | (1 << 5);
  cfg.instances = &s;
  lbc_ctx* c = nullptr;
  int rc = lbc_create(&cfg, &c);
  if (rc) return rc;
  rc = lbc_load_states(c, 0, 0, 1, fin);
  if (!rc) rc = lbc_run_interval(c, nullptr);
  if (!rc) rc = store_which(c, 0, 0, 0, 1, fout);
  if (!rc) rc = store_which(c, 0, 1, 0, 1, fin);
  if (!rc && force) rc = lbc_get_drag(c, 0, force);
  if (rc) g_err = c->err;
  lbc_destroy(c);
  return rc;
}

}  // extern "C"
