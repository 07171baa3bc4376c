"""ctypes binding of the C ABI in include/lbc.h (liblbc.so, built in this
directory).  Python mirror of the cloning-sampler inner loop for tests,
bench.py and scripting; the production host driver is driver/tlgk_lbm_gpu.cpp.

Every call goes to the HIP library: there is no CPU fallback, and loading
fails loudly when liblbc.so is missing.

Population arrays are FP64 in the reference's AoS layout
IDX(x,y,k) = k + 9*(y + Ly*x) (src/lbm_src/global.h:4): numpy shape
(n, Lx, Ly, 9).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbc.so")

MODE_REF = 0
MODE_CHANNEL = 1
STORE_F64 = 0
STORE_F32 = 1
STORE_F16 = 2
STORE_BF16 = 3
STORAGE_NAMES = {STORE_F64: "f64", STORE_F32: "f32", STORE_F16: "f16",
                 STORE_BF16: "bf16"}
STORAGE_BYTES = {STORE_F64: 8, STORE_F32: 4, STORE_F16: 2, STORE_BF16: 2}

# Entry points declared in include/lbc.h.
EXPORTS = (
    "lbc_create", "lbc_destroy", "lbc_last_error", "lbc_load_states",
    "lbc_store_states", "lbc_init_equilibrium", "lbc_set_snapshots",
    "lbc_set_marks", "lbc_get_marks", "lbc_run_interval", "lbc_get_drag",
    "lbc_get_logweights", "lbc_select", "lbc_gather", "lbc_gather_fused",
    "lbc_advance",
    "lbc_get_R_history", "lbc_get_scgf", "lbc_set_field_maps",
    "lbc_get_field_maps", "lbc_get_macro", "lbc_sync", "lbc_step_kernel_ms",
    "lbc_gather_kernel_ms", "lbc_threads",
    "lbc_sites_per_thread", "lbc_step_shape", "lbc_step_host",
    "lbc_select_static",
)

STATIC_MAX_CLONES = 8192  # LBC_STATIC_MAX_CLONES


class Instance(C.Structure):
    _fields_ = [
        ("Lx", C.c_int), ("Ly", C.c_int),
        ("xmin", C.c_int), ("xmax", C.c_int), ("ymin", C.c_int),
        ("ymax", C.c_int),
        ("n_clones", C.c_int), ("steps", C.c_int),
        ("tau", C.c_double), ("beta", C.c_double), ("u_in", C.c_double),
        ("force_scale", C.c_double), ("force_shift", C.c_double),
        ("dt", C.c_double), ("bias", C.c_double), ("eps", C.c_double),
        ("noise_sigma", C.c_double),
        ("rank_blocks", C.c_int), ("n_snapshots", C.c_int),
        ("forced_steps", C.c_int), ("field_eps", C.c_double),
        ("field_sigma", C.c_double), ("macro_every", C.c_int),
    ]


class Config(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("storage", C.c_int), ("n_instances", C.c_int),
        ("device", C.c_int), ("seed", C.c_uint64), ("threads", C.c_int),
        ("flags", C.c_int), ("instances", C.POINTER(Instance)),
    ]


class Selection(C.Structure):
    _fields_ = [
        ("nb_before", C.POINTER(C.c_int32)), ("nb_after", C.POINTER(C.c_int32)),
        ("pairs", C.POINTER(C.c_int32)), ("labels", C.POINTER(C.c_int32)),
        ("n_pairs", C.c_int32), ("n_rint_used", C.c_int32),
        ("total_R", C.c_double), ("log_total_R", C.c_double),
    ]


class LbcError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load liblbc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbcError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                       "or make -C cloning-lbm-mpi_amd")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.lbc_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.lbc_destroy.argtypes = [vp]
    L.lbc_destroy.restype = None
    L.lbc_last_error.argtypes = [vp]
    L.lbc_last_error.restype = C.c_char_p
    L.lbc_load_states.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.lbc_store_states.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.lbc_init_equilibrium.argtypes = [vp, C.c_int, C.c_double]
    L.lbc_set_snapshots.argtypes = [vp, C.c_int, dp]
    L.lbc_set_marks.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8)]
    L.lbc_get_marks.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8)]
    L.lbc_run_interval.argtypes = [vp, dp]
    L.lbc_get_drag.argtypes = [vp, C.c_int, dp]
    L.lbc_get_logweights.argtypes = [vp, C.c_int, dp, dp]
    L.lbc_select.argtypes = [vp, C.c_int, dp, dp, ip, C.POINTER(Selection)]
    L.lbc_gather.argtypes = [vp]
    L.lbc_gather_fused.argtypes = [vp]
    L.lbc_advance.argtypes = [vp, C.c_int]
    L.lbc_get_R_history.argtypes = [vp, C.c_int, dp, C.c_int,
                                    C.POINTER(C.c_int)]
    L.lbc_get_scgf.argtypes = [vp, C.c_int, C.c_double, dp, C.c_int,
                               C.POINTER(C.c_int)]
    L.lbc_set_field_maps.argtypes = [vp, C.c_int, dp]
    L.lbc_get_field_maps.argtypes = [vp, C.c_int, dp]
    L.lbc_get_macro.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.lbc_sync.argtypes = [vp]
    L.lbc_step_kernel_ms.argtypes = [vp, dp, C.POINTER(C.c_int)]
    L.lbc_gather_kernel_ms.argtypes = [vp, dp, dp]
    L.lbc_threads.argtypes = [vp]
    L.lbc_sites_per_thread.argtypes = [vp]
    L.lbc_step_shape.argtypes = [vp]
    L.lbc_step_host.argtypes = [C.c_int, C.POINTER(Instance), dp, dp, dp]
    L.lbc_select_static.argtypes = [C.c_int, C.c_int] + [vp] * 11
    _lib = L
    return L


def _dp(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _sized(a: np.ndarray, n: int, what: str) -> np.ndarray:
    """The library reads exactly n elements from a host array: check it."""
    if a.size != n:
        raise ValueError(f"{what}: expected {n} values, got {a.size}")
    return a


def _ip(a):
    if a is None:
        return None
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@dataclass
class InstanceSpec:
    """One sampler instance (fields of lbc_instance)."""
    Lx: int
    Ly: int
    xmin: int
    xmax: int
    ymin: int
    ymax: int
    n_clones: int
    steps: int
    tau: float
    beta: float = 0.0
    u_in: float = 0.0
    force_scale: float = 1.0
    force_shift: float = 0.0
    dt: float = 1.0
    bias: float = 0.0
    eps: float = 0.0
    noise_sigma: float = 0.0
    rank_blocks: int = 1
    n_snapshots: int = 0
    forced_steps: int = 0
    field_eps: float = 0.0
    field_sigma: float = 0.0
    macro_every: int = 0

    def c(self) -> Instance:
        return Instance(self.Lx, self.Ly, self.xmin, self.xmax, self.ymin,
                        self.ymax, self.n_clones, self.steps, self.tau,
                        self.beta, self.u_in, self.force_scale,
                        self.force_shift, self.dt, self.bias, self.eps,
                        self.noise_sigma, self.rank_blocks, self.n_snapshots,
                        self.forced_steps, self.field_eps, self.field_sigma,
                        self.macro_every)

    @property
    def sites(self) -> int:
        return self.Lx * self.Ly


def ref_box(L: int, tau: float, U0: float, n_clones: int, steps: int,
            alpha: float, eps: float = 0.0, n_snapshots: int = 0,
            zero_mean: bool = True, rank_blocks: int = 1) -> InstanceSpec:
    """The reference driver's geometry and scalings for square side L
    (src/mpi_LBM_GK_pops_perturb.cpp:131-150, 27)."""
    Dy = 4 * L + 1
    Dx = Dy
    xmin = (Dx - 1) // 2
    ymin = (Dy - 1) // 2 - L // 2
    beta0 = (1.0 / (Dx - 1)) * (float(L) / (Dy - 1)) * U0 * U0
    T0 = L / U0
    F0 = (U0 * U0) * L * 0.5
    return InstanceSpec(Dx, Dy, xmin, xmin + L, ymin, ymin + L, n_clones,
                        steps, tau, beta=beta0, force_scale=1.0 / F0,
                        force_shift=1.775 if zero_mean else 0.0,
                        dt=1.0 / T0, bias=alpha, eps=eps,
                        rank_blocks=rank_blocks, n_snapshots=n_snapshots)


def forcing_box(L: int, tau: float, U0: float, n_clones: int, dT: float,
                dT0: float, eps: float, alpha: float,
                field_sigma: float = 0.0) -> InstanceSpec:
    """The forcing driver's geometry, scalings and two-phase interval for
    square side L (src/mpi_LBM_GK_forcing.cpp:131-148): floor(dT0*T0) forced
    steps then floor((dT - dT0)*T0) uniform ones, T0 = L/U0, drag scaled by
    1/F0 with no shift, random-field amplitude eps."""
    T0 = L / U0
    m1 = math.floor(dT0 * T0)
    m2 = math.floor((dT - dT0) * T0)
    spec = ref_box(L, tau, U0, n_clones, m1 + m2, alpha, zero_mean=False)
    spec.forced_steps = m1
    spec.field_eps = eps
    spec.field_sigma = field_sigma
    return spec


def channel(Lx: int, Ly: int, side: int, cx: int, cy: int, tau: float,
            u_in: float, n_clones: int, steps: int, k: float,
            sigma: float = 0.0) -> InstanceSpec:
    """Channel flow past a square of `side` centred at (cx, cy); the ring
    runs from cx - side/2 to cx - side/2 + side."""
    xmin = cx - side // 2
    ymin = cy - side // 2
    return InstanceSpec(Lx, Ly, xmin, xmin + side, ymin, ymin + side,
                        n_clones, steps, tau, u_in=u_in, bias=k,
                        noise_sigma=sigma)


@dataclass
class SelectionResult:
    nb_before: np.ndarray
    nb_after: np.ndarray
    pairs: np.ndarray
    labels: np.ndarray
    n_rint_used: int
    total_R: float
    log_total_R: float


class Sampler:
    """A context of the HIP library: one or more sampler instances sharing a
    physics mode and a storage precision."""

    def __init__(self, specs, mode: int = MODE_CHANNEL,
                 storage: int = STORE_F64, seed: int = 0, threads: int = 0,
                 device: int = 0, vec: int = 0, nt: int | None = None,
                 fused: bool | None = None, passes: int | None = None,
                 cpw: int = 0):
        """threads = 0 tunes the step kernel's shape at creation; with
        threads given, unspecified nt / passes default to off / one pass.
        passes: LBM passes per HBM sweep (1, 2 or 3); fused=True is
        passes=2, fused=False passes=1.  cpw: clones per workgroup (1, or 2
        with one pass per sweep; 0 = tuned where the tuner runs, else 1)."""
        if isinstance(specs, InstanceSpec):
            specs = [specs]
        self.specs = list(specs)
        self.mode = mode
        self.storage = storage
        L = lib()
        arr = (Instance * len(self.specs))(*[s.c() for s in self.specs])
        self._arr = arr
        if threads and not vec:
            vec = 1
        if nt is None:
            nt = False if threads else None
        if passes is None and fused is not None:
            passes = 2 if fused else 1
        if passes is None:
            passes = 1 if threads else 0
        flags = (vec | ((0 if nt is None else (2 if nt else 1)) << 3)
                 | (passes << 5) | (cpw << 7))
        cfg = Config(mode, storage, len(self.specs), device, seed, threads,
                     flags, C.cast(arr, C.POINTER(Instance)))
        h = C.c_void_p()
        rc = L.lbc_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LbcError(f"lbc_create failed ({rc}): "
                           f"{L.lbc_last_error(None).decode()}")
        self._h = h

    # -- helpers
    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = lib().lbc_last_error(self._h).decode()
            raise LbcError(f"{what} failed ({rc}): {msg}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().lbc_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def threads(self) -> int:
        return lib().lbc_threads(self._h)

    @property
    def vec(self) -> int:
        return lib().lbc_sites_per_thread(self._h)

    @property
    def nt(self) -> bool:
        """Non-temporal population stores in the step kernel."""
        return ((lib().lbc_step_shape(self._h) >> 3) & 3) == 2

    @property
    def passes(self) -> int:
        """LBM passes per HBM sweep of the step kernel (LDS rings between)."""
        return (lib().lbc_step_shape(self._h) >> 5) & 3

    @property
    def cpw(self) -> int:
        """Clones per workgroup of the step kernel (1 or 2)."""
        return (lib().lbc_step_shape(self._h) >> 7) & 3

    @property
    def fused(self) -> bool:
        return self.passes > 1

    # -- state
    def load_states(self, inst: int, clone0: int, aos: np.ndarray) -> None:
        s = self.specs[inst]
        a = np.ascontiguousarray(aos, np.float64).reshape(-1, s.Lx, s.Ly, 9)
        self._check(lib().lbc_load_states(self._h, inst, clone0, a.shape[0],
                                          _dp(a)), "lbc_load_states")

    def store_states(self, inst: int, clone0: int = 0, n: int | None = None):
        s = self.specs[inst]
        n = s.n_clones - clone0 if n is None else n
        out = np.zeros((n, s.Lx, s.Ly, 9))
        self._check(lib().lbc_store_states(self._h, inst, clone0, n, _dp(out)),
                    "lbc_store_states")
        return out

    def init_equilibrium(self, inst: int, amp: float) -> None:
        self._check(lib().lbc_init_equilibrium(self._h, inst, amp),
                    "lbc_init_equilibrium")

    def set_snapshots(self, inst: int, snaps: np.ndarray) -> None:
        s = self.specs[inst]
        a = _sized(np.ascontiguousarray(snaps, np.float64),
                   s.n_snapshots * s.Lx * s.Ly * 9, "snapshots")
        self._check(lib().lbc_set_snapshots(self._h, inst, _dp(a)),
                    "lbc_set_snapshots")

    def set_marks(self, inst: int, marks) -> None:
        m = _sized(np.ascontiguousarray(marks, np.uint8),
                   self.specs[inst].n_clones, "marks")
        self._check(lib().lbc_set_marks(
            self._h, inst, m.ctypes.data_as(C.POINTER(C.c_uint8))),
            "lbc_set_marks")

    def get_marks(self, inst: int) -> np.ndarray:
        m = np.zeros(self.specs[inst].n_clones, np.uint8)
        self._check(lib().lbc_get_marks(
            self._h, inst, m.ctypes.data_as(C.POINTER(C.c_uint8))),
            "lbc_get_marks")
        return m

    # -- the loop
    def run_interval(self, prefactors: np.ndarray | None = None) -> None:
        p = None
        if prefactors is not None:
            p = np.ascontiguousarray(prefactors, np.float64)
            if self.mode == MODE_REF:
                _sized(p, sum(sp.n_clones * sp.n_snapshots
                              for sp in self.specs), "prefactors")
        self._check(lib().lbc_run_interval(self._h, _dp(p)),
                    "lbc_run_interval")

    def drag(self, inst: int) -> np.ndarray:
        s = self.specs[inst]
        out = np.zeros((s.n_clones, s.steps))
        self._check(lib().lbc_get_drag(self._h, inst, _dp(out)),
                    "lbc_get_drag")
        return out

    def logweights(self, inst: int):
        n = self.specs[inst].n_clones
        obs = np.zeros(n)
        lw = np.zeros(n)
        self._check(lib().lbc_get_logweights(self._h, inst, _dp(obs), _dp(lw)),
                    "lbc_get_logweights")
        return obs, lw

    def select(self, inst: int, weights=None, u=None, rint=None
               ) -> SelectionResult:
        n = self.specs[inst].n_clones
        w = (None if weights is None else
             _sized(np.ascontiguousarray(weights, np.float64), n, "weights"))
        uu = (None if u is None else
              _sized(np.ascontiguousarray(u, np.float64), n, "u"))
        rr = (None if rint is None else
              _sized(np.ascontiguousarray(rint, np.int32), n, "rint"))
        nb0 = np.zeros(n, np.int32)
        nb1 = np.zeros(n, np.int32)
        pairs = np.zeros(2 * n, np.int32)
        labels = np.zeros(n, np.int32)
        out = Selection(_ip(nb0), _ip(nb1), _ip(pairs), _ip(labels), 0, 0,
                        0.0, 0.0)
        self._check(lib().lbc_select(self._h, inst, _dp(w), _dp(uu), _ip(rr),
                                     C.byref(out)), "lbc_select")
        return SelectionResult(nb0, nb1, pairs[: 2 * out.n_pairs].reshape(-1, 2),
                               labels, out.n_rint_used, out.total_R,
                               out.log_total_R)

    def gather(self) -> None:
        self._check(lib().lbc_gather(self._h), "lbc_gather")

    def gather_fused(self) -> None:
        """The copies collided into the next interval's pass 0 (CHANNEL with
        noise): no state may be read, replaced or reselected before the next
        run_interval or advance."""
        self._check(lib().lbc_gather_fused(self._h), "lbc_gather_fused")

    def advance(self, n_intervals: int) -> None:
        self._check(lib().lbc_advance(self._h, n_intervals), "lbc_advance")

    def R_history(self, inst: int) -> np.ndarray:
        n = C.c_int()
        self._check(lib().lbc_get_R_history(self._h, inst, None, 0,
                                            C.byref(n)), "lbc_get_R_history")
        out = np.zeros(max(n.value, 1))
        self._check(lib().lbc_get_R_history(self._h, inst, _dp(out), n.value,
                                            C.byref(n)), "lbc_get_R_history")
        return out[: n.value]

    def scgf(self, inst: int, dT: float) -> np.ndarray:
        """Running scaled-cumulant estimate after every interval so far
        (FP64, on the device), read_rvalue.py's form with interval length
        dT."""
        n = C.c_int()
        self._check(lib().lbc_get_scgf(self._h, inst, dT, None, 0, C.byref(n)),
                    "lbc_get_scgf")
        out = np.zeros(max(n.value, 1))
        self._check(lib().lbc_get_scgf(self._h, inst, dT, _dp(out), n.value,
                                       C.byref(n)), "lbc_get_scgf")
        return out[: n.value]

    def set_field_maps(self, inst: int, maps: np.ndarray) -> None:
        """Forcing mode: the next interval's random fields, shape
        (n_clones, Lx, Ly) (the reference's map[y + Ly*x] layout)."""
        s = self.specs[inst]
        a = np.ascontiguousarray(maps, np.float64)
        if a.shape != (s.n_clones, s.Lx, s.Ly):
            raise ValueError(f"field maps: expected shape "
                             f"{(s.n_clones, s.Lx, s.Ly)}, got {a.shape}")
        self._check(lib().lbc_set_field_maps(self._h, inst, _dp(a)),
                    "lbc_set_field_maps")

    def field_maps(self, inst: int) -> np.ndarray:
        s = self.specs[inst]
        out = np.zeros((s.n_clones, s.Lx, s.Ly))
        self._check(lib().lbc_get_field_maps(self._h, inst, _dp(out)),
                    "lbc_get_field_maps")
        return out

    def macro(self, inst: int, clone0: int = 0, n: int | None = None):
        """Macroscopic fields of the last interval (macro_every > 0):
        shape (n, frames, 3, Ly, Lx) -- rho, ux, uy at steps 0, E, 2E, ..."""
        s = self.specs[inst]
        n = s.n_clones - clone0 if n is None else n
        frames = -(-s.steps // s.macro_every) if s.macro_every > 0 else 1
        out = np.zeros((n, frames, 3, s.Ly, s.Lx))
        self._check(lib().lbc_get_macro(self._h, inst, clone0, n, _dp(out)),
                    "lbc_get_macro")
        return out

    def sync(self) -> None:
        self._check(lib().lbc_sync(self._h), "lbc_sync")

    def step_kernel_ms(self):
        ms = C.c_double()
        n = C.c_int()
        self._check(lib().lbc_step_kernel_ms(self._h, C.byref(ms), C.byref(n)),
                    "lbc_step_kernel_ms")
        return ms.value, n.value

    def gather_kernel_ms(self):
        """(device ms, population bytes moved) of the last gather's copy
        kernel."""
        ms = C.c_double()
        b = C.c_double()
        self._check(lib().lbc_gather_kernel_ms(self._h, C.byref(ms), C.byref(b)),
                    "lbc_gather_kernel_ms")
        return ms.value, b.value


def step_host(mode: int, spec: InstanceSpec, fin: np.ndarray):
    """One reference step on one host lattice through the HIP library:
    returns (post-collision fin, next state fout, raw force)."""
    a = np.ascontiguousarray(fin, np.float64).copy()
    out = np.zeros_like(a)
    F = np.zeros(1)
    inst = spec.c()
    rc = lib().lbc_step_host(mode, C.byref(inst), _dp(a), _dp(out), _dp(F))
    if rc != 0:
        raise LbcError(f"lbc_step_host failed ({rc}): "
                       f"{lib().lbc_last_error(None).decode()}")
    return a, out, float(F[0])


def select_static(n: int, rank_blocks: int, logw: int, u: int, rint: int,
                  nb_before: int, nb_after: int, pairs: int, labels: int,
                  marks: int, info: int, R: int, stream: int = 0) -> None:
    """lbc_select_static on device arrays given by address (the caller owns
    and sizes them; see include/lbc.h); enqueued on `stream`."""
    rc = lib().lbc_select_static(n, rank_blocks, logw, u, rint, nb_before,
                                 nb_after, pairs, labels, marks, info, R,
                                 stream or None)
    if rc != 0:
        raise LbcError(f"lbc_select_static failed ({rc}): "
                       f"{lib().lbc_last_error(None).decode()}")


def bytes_per_lup(storage: int) -> int:
    """Algorithmic HBM bytes per lattice update: 9 reads + 9 writes."""
    return 18 * STORAGE_BYTES[storage]


def clone_mlups(specs, intervals: int, seconds: float) -> float:
    lups = sum(s.n_clones * s.sites * s.steps for s in specs) * intervals
    return lups / seconds / 1e6 if seconds > 0 else math.nan
