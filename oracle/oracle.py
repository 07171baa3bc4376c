"""ctypes bindings to the CPU FP64 oracle (oracle/liboracle.so) and, where it
was built, to the reference kernels linked by oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.

Arrays are numpy FP64 in the reference's AoS layout [x][y][k]
(src/lbm_src/global.h:4): shape (Lx, Ly, 9), C order.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

MODE_REF = 0
MODE_CHANNEL = 1

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])


class Geom(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("Lx", C.c_int),
        ("Ly", C.c_int),
        ("xmin", C.c_int),
        ("xmax", C.c_int),
        ("ymin", C.c_int),
        ("ymax", C.c_int),
        ("tau", C.c_double),
        ("beta", C.c_double),
        ("u_in", C.c_double),
    ]


def ref_box_geom(L: int, tau: float, U0: float) -> Geom:
    """REF geometry of the reference driver for square side L
    (src/mpi_LBM_GK_pops_perturb.cpp:131-133, 144)."""
    Dy = 4 * L + 1
    Dx = Dy
    xmin = (Dx - 1) // 2
    xmax = xmin + L
    ymin = (Dy - 1) // 2 - L // 2
    ymax = ymin + L
    beta0 = (1.0 / (Dx - 1)) * (float(L) / (Dy - 1)) * U0 * U0
    return Geom(MODE_REF, Dx, Dy, xmin, xmax, ymin, ymax, tau, beta0, 0.0)


def channel_geom(Lx: int, Ly: int, side: int, cx: int, cy: int, tau: float,
                 u_in: float) -> Geom:
    """CHANNEL geometry: square ring xmin = cx - side/2 .. xmin + side."""
    xmin = cx - side // 2
    ymin = cy - side // 2
    return Geom(MODE_CHANNEL, Lx, Ly, xmin, xmin + side, ymin, ymin + side,
                tau, 0.0, u_in)


_lib = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        gp = C.POINTER(Geom)
        L.orc_step.argtypes = [gp, dp, dp]
        L.orc_step.restype = C.c_double
        L.orc_force_stress.argtypes = [gp, dp]
        L.orc_force_stress.restype = C.c_double
        L.orc_force_mea.argtypes = [gp, dp]
        L.orc_force_mea.restype = C.c_double
        L.orc_run_steps.argtypes = [gp, dp, dp, C.c_int, C.c_double, C.c_double,
                                    C.c_double, dp]
        L.orc_run_steps.restype = C.c_double
        L.orc_run_clones.argtypes = [gp, dp, dp, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, dp]
        L.orc_run_clones.restype = C.c_double
        L.orc_perturb.argtypes = [dp, dp, C.c_int, C.c_long, C.c_double, dp]
        L.orc_perturb.restype = None
        L.orc_select.argtypes = [C.c_int, dp, C.c_int, dp, ip, ip, ip, ip, ip,
                                 dp, C.POINTER(C.c_int)]
        L.orc_select.restype = C.c_int
        u32p = C.POINTER(C.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_philox4x32_10.restype = None
        L.orc_box_muller4.argtypes = [u32p, dp]
        L.orc_box_muller4.restype = None
        L.orc_site_normals.argtypes = [u32p, dp]
        L.orc_site_normals.restype = None
        L.orc_noise.argtypes = [gp, dp, C.c_double, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_noise.restype = None
        L.orc_init_channel.argtypes = [gp, dp, C.c_double, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_init_channel.restype = None
        L.orc_feq.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_bb_mask.argtypes = [gp, C.c_int, C.c_int]
        L.orc_bb_mask.restype = C.c_int
        L.orc_scgf.argtypes = [dp, C.c_int, C.c_double, dp]
        L.orc_scgf.restype = None
        L.orc_step_map.argtypes = [gp, dp, dp, dp, C.c_double]
        L.orc_step_map.restype = C.c_double
        L.orc_run_steps_forced.argtypes = [gp, dp, dp, C.c_int, C.c_int, dp,
                                           C.c_double, C.c_double, C.c_double,
                                           C.c_double, dp]
        L.orc_run_steps_forced.restype = C.c_double
        L.orc_field_normals.argtypes = [C.c_uint32] * 5 + [C.c_int, dp]
        L.orc_macro.argtypes = [gp, dp, dp, C.c_double, dp, dp, dp]
        L.orc_macro.restype = None
        L.orc_field_normals.restype = None
        L.orc_glibc_srand.argtypes = [C.c_uint]
        L.orc_glibc_rand_n.argtypes = [C.c_int, ip]
        L.orc_glibc_seed48_default.argtypes = []
        L.orc_glibc_drand48_n.argtypes = [C.c_int, dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        dp = C.POINTER(C.c_double)
        R.ref_set_geometry.argtypes = [C.c_int] * 6
        R.ref_step.argtypes = [dp, dp, dp, dp, dp, C.c_double, C.c_double]
        R.ref_step_spatial.argtypes = [dp, dp, dp, dp, dp, C.c_double, dp,
                                       C.c_double, C.c_double]
        R.ref_force.argtypes = [dp, C.c_double]
        R.ref_force.restype = C.c_double
        R.ref_run_clones.argtypes = [dp, dp, C.c_int, C.c_int, C.c_double,
                                     C.c_double, dp]
        R.ref_run_clones.restype = None
        R.ref_perturb.argtypes = [dp, dp, C.c_double, C.c_int, C.c_int]
        R.ref_seed48.argtypes = [C.c_ushort] * 3
        R.ref_drand48.restype = C.c_double
        _ref = R
    return _ref


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _u32p(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# --------------------------------------------------------------------------
# thin numpy wrappers

def step(g: Geom, fin: np.ndarray, fout: np.ndarray) -> float:
    return lib().orc_step(C.byref(g), _dp(fin), _dp(fout))


def force(g: Geom, f: np.ndarray) -> float:
    if g.mode == MODE_REF:
        return lib().orc_force_stress(C.byref(g), _dp(f))
    return lib().orc_force_mea(C.byref(g), _dp(f))


def run_steps(g: Geom, state: np.ndarray, M: int, scale=1.0, shift=0.0,
              dt=1.0):
    """M steps in place; returns (acc, A[M])."""
    scratch = np.empty_like(state)
    A = np.zeros(M)
    acc = lib().orc_run_steps(C.byref(g), _dp(state), _dp(scratch), M, scale,
                              shift, dt, _dp(A))
    return acc, A


def run_clones(g: Geom, states: np.ndarray, M: int, scale=1.0, shift=0.0,
               dt=1.0):
    n = states.shape[0]
    scratch = np.empty_like(states[0])
    acc = np.zeros(n)
    lib().orc_run_clones(C.byref(g), _dp(states), _dp(scratch), n, M, scale,
                         shift, dt, _dp(acc))
    return acc


def step_map(g: Geom, fin: np.ndarray, fout: np.ndarray, fmap: np.ndarray,
             eps: float) -> float:
    """One REF step with the per-site force (1 + eps*map)*beta; fmap has
    shape (Lx, Ly) (the reference's idx(x,y) = y + Ly*x layout)."""
    m = np.ascontiguousarray(fmap, np.float64)
    return lib().orc_step_map(C.byref(g), _dp(fin), _dp(fout), _dp(m), eps)


def macro(g: Geom, f: np.ndarray, fmap: np.ndarray | None = None,
          eps: float = 0.0) -> np.ndarray:
    """rho, ux, uy of state f as the step computes them; shape (3, Lx, Ly)."""
    out = np.zeros((3, g.Lx, g.Ly))
    m = None if fmap is None else np.ascontiguousarray(fmap, np.float64)
    lib().orc_macro(C.byref(g), _dp(np.ascontiguousarray(f)),
                    None if m is None else _dp(m), eps, _dp(out[0]),
                    _dp(out[1]), _dp(out[2]))
    return out


def run_steps_forced(g: Geom, state: np.ndarray, M1: int, M: int,
                     fmap: np.ndarray, eps: float, scale=1.0, shift=0.0,
                     dt=1.0):
    """Forcing-mode interval in place: M1 forced steps, then M - M1 uniform
    ones; returns (acc, A[M])."""
    scratch = np.empty_like(state)
    A = np.zeros(M)
    m = np.ascontiguousarray(fmap, np.float64)
    acc = lib().orc_run_steps_forced(C.byref(g), _dp(state), _dp(scratch), M1,
                                     M, _dp(m), eps, scale, shift, dt, _dp(A))
    return acc, A


# --------------------------------------------------------------------------
# random fields of the forcing mode
# (src/perturbation_routines/generate_random_field.cpp:9-181)

def field_embedding(Lx: int, Ly: int, sigma: float = 32.0) -> np.ndarray:
    """Wrapped covariance of the circulant embedding, shape (2Lx-1, 2Ly-1):
    G[r, j] = f(min(r, fmx-r), min(j, fmy-j)) / (fmx*fmy) with
    f(x, y) = exp(-(x*x + y*y) / (2 sigma^2)) (covariance_fctn.cpp:3-10;
    generate_random_field.cpp:44-106 for Lx = Ly).  The exponentials come
    from math.exp (the C library's exp), one per distinct squared lag."""
    fmx, fmy = 2 * Lx - 1, 2 * Ly - 1
    r = np.arange(fmx)
    j = np.arange(fmy)
    dx = np.minimum(r, fmx - r)
    dy = np.minimum(j, fmy - j)
    d = dx[:, None] ** 2 + dy[None, :] ** 2
    uniq, inv = np.unique(d, return_inverse=True)
    vals = np.array([math.exp(-float(v) / (2 * sigma * sigma)) for v in uniq])
    return vals[inv].reshape(d.shape) / float(fmx * fmy)


def field_eigenvalues(Lx: int, Ly: int, sigma: float = 32.0) -> np.ndarray:
    """Eigenvalues of the embedding, negative ones set to zero
    (generate_random_field.cpp:127-158; its abs() is the int one, so every
    negative eigenvalue above -1 is zeroed)."""
    lam = np.fft.fft2(field_embedding(Lx, Ly, sigma)).real
    if (lam <= -1.0).any():
        raise ValueError("embedding eigenvalue <= -1")
    return np.where(lam < 0, 0.0, lam)


def field_normals(seed: int, clone: int, interval: int, instance: int,
                  shape) -> np.ndarray:
    """The complex normals Z of one clone's draw, shape (fmx, fmy)."""
    n = int(shape[0] * shape[1])
    z = np.zeros(2 * n)
    lib().orc_field_normals(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF,
                            clone, interval, instance, n, _dp(z))
    return (z[0::2] + 1j * z[1::2]).reshape(shape)


def field_from_normals(lam: np.ndarray, Z: np.ndarray, Lx: int,
                       Ly: int) -> np.ndarray:
    """Field on the lattice, shape (Lx, Ly): Re of the forward 2-D DFT of
    sqrt(lam)*Z on the block x < Lx, y < Ly
    (generate_random_field.cpp:158-178)."""
    return np.fft.fft2(np.sqrt(lam) * Z).real[:Lx, :Ly].copy()


def field_of_clone(lam: np.ndarray, seed: int, clone: int, interval: int,
                   instance: int, Lx: int, Ly: int) -> np.ndarray:
    """The field the device gives `clone`: clones 2k and 2k + 1 share the
    transform of clone 2k's normals and take its real and imaginary parts
    (two independent fields of the reference's law; the reference keeps
    only the real part of a fresh draw per clone)."""
    Z = field_normals(seed, clone - clone % 2, interval, instance, lam.shape)
    F = np.fft.fft2(np.sqrt(lam) * Z)[:Lx, :Ly]
    return (F.real if clone % 2 == 0 else F.imag).copy()


def field_covariance(lam: np.ndarray, Lx: int, Ly: int) -> np.ndarray:
    """Covariance of the field at lag (dx, dy), dx < Lx, dy < Ly, implied by
    the (clipped) eigenvalues: sum_k lam_k cos(2 pi k.d / M)."""
    return np.fft.fft2(lam).real[:Lx, :Ly].copy()


def scgf(logR: np.ndarray, dT: float) -> np.ndarray:
    """Running scaled-cumulant estimate lam[0..n) from log mean weights."""
    logR = np.ascontiguousarray(logR, np.float64)
    lam = np.zeros(len(logR))
    lib().orc_scgf(_dp(logR), len(logR), dT, _dp(lam))
    return lam


def perturb(fin: np.ndarray, snaps: np.ndarray, eps: float, a: np.ndarray):
    NN = snaps.shape[0]
    lib().orc_perturb(_dp(fin), _dp(np.ascontiguousarray(snaps)), NN,
                      fin.size, eps, _dp(np.ascontiguousarray(a, np.float64)))


def select(s: np.ndarray, u: np.ndarray, rint: np.ndarray, p_blocks: int = 1):
    Nc = s.size
    s = np.ascontiguousarray(s, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    rint = np.ascontiguousarray(rint, np.int32)
    nb0 = np.zeros(Nc, np.int32)
    nb1 = np.zeros(Nc, np.int32)
    pairs = np.zeros(2 * Nc, np.int32)
    labels = np.zeros(Nc, np.int32)
    tot = C.c_double()
    used = C.c_int()
    n = lib().orc_select(Nc, _dp(s), p_blocks, _dp(u), _ip(rint), _ip(nb0),
                         _ip(nb1), _ip(pairs), _ip(labels), C.byref(tot),
                         C.byref(used))
    if n < 0:
        raise ValueError("orc_select rejected its input")
    return dict(nb_before=nb0, nb_after=nb1, pairs=pairs[: 2 * n].reshape(-1, 2),
                labels=labels, total_R=tot.value, n_rint=used.value)


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_u32p(c), _u32p(k), _u32p(out))
    return out


def box_muller4(r) -> np.ndarray:
    rr = np.ascontiguousarray(r, np.uint32)
    n = np.zeros(4)
    lib().orc_box_muller4(_u32p(rr), _dp(n))
    return n


def site_normals(r) -> np.ndarray:
    """The ten normals orc_site_normals makes of one Philox block (nine used
    as one site's copy noise)."""
    rr = np.ascontiguousarray(r, np.uint32)
    n = np.zeros(10)
    lib().orc_site_normals(_u32p(rr), _dp(n))
    return n


def noise(g: Geom, f: np.ndarray, sigma: float, seed: int, clone: int,
          interval: int, instance: int = 0) -> None:
    lib().orc_noise(C.byref(g), _dp(f), sigma, seed & 0xFFFFFFFF,
                    (seed >> 32) & 0xFFFFFFFF, clone, interval, instance)


def init_channel(g: Geom, amp: float, seed: int, clone: int,
                 instance: int = 0) -> np.ndarray:
    f = np.zeros((g.Lx, g.Ly, 9))
    lib().orc_init_channel(C.byref(g), _dp(f), amp, seed & 0xFFFFFFFF,
                           (seed >> 32) & 0xFFFFFFFF, clone, instance)
    return f


def feq(rho: float, ux: float, uy: float) -> np.ndarray:
    out = np.zeros(9)
    lib().orc_feq(rho, ux, uy, _dp(out))
    return out


def bb_mask(g: Geom, x: int, y: int) -> int:
    return lib().orc_bb_mask(C.byref(g), x, y)


def glibc_rand(seed: int, n: int, discard: int = 0) -> np.ndarray:
    """srand(seed), discard `discard` draws, return the next n rand()."""
    L = lib()
    L.orc_glibc_srand(seed)
    if discard:
        tmp = np.zeros(discard, np.int32)
        L.orc_glibc_rand_n(discard, _ip(tmp))
    out = np.zeros(n, np.int32)
    L.orc_glibc_rand_n(n, _ip(out))
    return out


def glibc_drand48(n: int, skip: int = 0) -> np.ndarray:
    """drand48() from the default (never-seeded) state."""
    L = lib()
    L.orc_glibc_seed48_default()
    if skip:
        tmp = np.zeros(skip)
        L.orc_glibc_drand48_n(skip, _dp(tmp))
    out = np.zeros(n)
    L.orc_glibc_drand48_n(n, _dp(out))
    return out


# --------------------------------------------------------------------------
# reference kernels (oracle/_ref)

def ref_step(g: Geom, fin: np.ndarray, fout: np.ndarray) -> float:
    """One reference step (fin <- post-collision, fout <- new state); returns
    the raw stress-integral force on fout, as the driver computes it."""
    R = ref()
    R.ref_set_geometry(g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax)
    S = g.Lx * g.Ly
    rho = np.zeros(S)
    ux = np.zeros(S)
    uy = np.zeros(S)
    R.ref_step(_dp(fin), _dp(fout), _dp(rho), _dp(ux), _dp(uy), g.beta, g.tau)
    return R.ref_force(_dp(fout), 1.0 / g.tau)


def ref_set_geometry(g: Geom) -> None:
    """Point the reference kernels' geometry globals at g (before threads
    start running ref_run_clones on it)."""
    ref().ref_set_geometry(g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax)


def ref_run_clones(g: Geom, states: np.ndarray, M: int) -> np.ndarray:
    """The reference driver's step loop on n clones (states (n, Lx, Ly, 9),
    advanced in place by M production steps of the reference's own kernels);
    returns each clone's sum of raw forces.  ref_set_geometry(g) first."""
    n = states.shape[0]
    scratch = np.empty_like(states[0])
    acc = np.zeros(n)
    ref().ref_run_clones(_dp(states), _dp(scratch), n, M, g.beta, g.tau,
                         _dp(acc))
    return acc


def ref_step_fields(g: Geom, fin: np.ndarray, fout: np.ndarray,
                    fmap: np.ndarray | None = None, eps: float = 0.0):
    """One reference step (uniform, or spatial with fmap); returns the
    fields rho, ux, uy the reference kernel writes, shape (3, Lx, Ly)."""
    R = ref()
    R.ref_set_geometry(g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax)
    out = np.zeros((3, g.Lx, g.Ly))
    if fmap is None:
        R.ref_step(_dp(fin), _dp(fout), _dp(out[0]), _dp(out[1]), _dp(out[2]),
                   g.beta, g.tau)
    else:
        m = np.ascontiguousarray(fmap, np.float64).copy()
        R.ref_step_spatial(_dp(fin), _dp(fout), _dp(out[0]), _dp(out[1]),
                           _dp(out[2]), g.beta, _dp(m), g.tau, eps)
    return out


def ref_step_spatial(g: Geom, fin: np.ndarray, fout: np.ndarray,
                     fmap: np.ndarray, eps: float) -> float:
    """One reference forced step (src/mpi_LBM_GK_forcing.cpp:337-351) with
    map fmap of shape (Lx, Ly); returns the raw force on fout."""
    R = ref()
    R.ref_set_geometry(g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax)
    S = g.Lx * g.Ly
    rho = np.zeros(S)
    ux = np.zeros(S)
    uy = np.zeros(S)
    m = np.ascontiguousarray(fmap, np.float64).copy()
    R.ref_step_spatial(_dp(fin), _dp(fout), _dp(rho), _dp(ux), _dp(uy), g.beta,
                       _dp(m), g.tau, eps)
    return R.ref_force(_dp(fout), 1.0 / g.tau)


def ref_perturb_default_stream(fin: np.ndarray, snaps: np.ndarray, eps: float,
                               skip: int = 0) -> np.ndarray:
    """Run the reference generatePerturbedState from drand48's default state
    (after `skip` draws); returns the NN prefactors it consumed."""
    R = ref()
    R.ref_seed48(0, 0, 0)
    for _ in range(skip):
        R.ref_drand48()
    NN = snaps.shape[0]
    R.ref_perturb(_dp(fin), _dp(np.ascontiguousarray(snaps)), eps, NN, fin.size)
    R.ref_seed48(0, 0, 0)
    for _ in range(skip):
        R.ref_drand48()
    return np.array([R.ref_drand48() for _ in range(NN)])
