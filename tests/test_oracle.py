"""Pins the CPU oracle: against golden fixtures written by the reference's own
compiled kernels, against the live reference build where it exists, and
against closed forms and invariants that a plausible mistake would break."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _geom_from(O, fx, mode=0):
    Lx, Ly, x0, x1, y0, y1 = [int(v) for v in fx["geom"]]
    return O.Geom(mode, Lx, Ly, x0, x1, y0, y1, float(fx["tau"]),
                  float(fx["beta"]), 0.0)


@pytest.mark.parametrize("L", [4, 8])
def test_oracle_matches_reference_fixture(O, L):
    fx = np.load(os.path.join(GOLDEN, f"ref_steps_L{L}.npz"))
    g = _geom_from(O, fx)
    a = fx["S0"].copy()
    b = np.zeros_like(a)
    keep = list(fx["keep"])
    for n in range(len(fx["forces"])):
        F = O.step(g, a, b)
        if n in keep:
            i = keep.index(n)
            assert np.array_equal(a, fx["post"][i]), f"post-collision step {n}"
            assert np.array_equal(b, fx["states"][i]), f"state step {n}"
        assert F == fx["forces"][n], f"force step {n}"
        a, b = b, a


def test_oracle_perturbation_matches_reference_fixture(O):
    fx = np.load(os.path.join(GOLDEN, "ref_perturb_L4.npz"))
    f = fx["fin"].copy()
    O.perturb(f, fx["snaps"], float(fx["eps"]), fx["prefactors"])
    assert np.array_equal(f, fx["out"])


def test_perturbation_prefactors_are_drand48_default_stream(O):
    fx = np.load(os.path.join(GOLDEN, "ref_perturb_L4.npz"))
    d = np.load(os.path.join(GOLDEN, "glibc_streams.npz"))["drand48_default"]
    assert np.array_equal(fx["prefactors"], d[:3])


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
    "_ref", "libref.so")), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("L,tau,U0,seed", [(4, 0.6, 0.05, 1), (5, 0.53, 0.08, 2),
                                           (6, 0.515, 0.03, 3), (9, 0.7, 0.1, 4)])
def test_oracle_matches_live_reference(O, L, tau, U0, seed):
    g = O.ref_box_geom(L, tau, U0)
    rng = np.random.default_rng(seed)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + 1e-2 * rng.standard_normal(),
                            U0 + 1e-2 * rng.standard_normal(),
                            1e-2 * rng.standard_normal())
    a, b, c, d = f.copy(), np.zeros_like(f), f.copy(), np.zeros_like(f)
    for _ in range(5):
        Fo = O.step(g, a, b)
        Fr = O.ref_step(g, c, d)
        assert Fo == Fr
        assert np.array_equal(a, c) and np.array_equal(b, d)
        a, b, c, d = b, a, d, c
    # perturbation from the same drand48 position
    snaps = np.stack([a * (1 + 1e-3 * rng.standard_normal(a.shape))
                      for _ in range(4)])
    x1 = a.copy()
    pref = O.ref_perturb_default_stream(x1, snaps, 0.1, skip=7)
    x2 = a.copy()
    O.perturb(x2, snaps, 0.1, pref)
    assert np.array_equal(x1, x2)


# ---------------------------------------------------------------- RNG pins
def _drand48_closed_form(n):
    a, c, m = 0x5DEECE66D, 0xB, 1 << 48
    x = 0  # glibc's never-seeded state
    out = []
    for _ in range(n):
        x = (a * x + c) % m
        out.append(x / float(m))
    return np.array(out)


def _glibc_random_r_closed_form(seed, n, discard):
    """glibc random() TYPE_3: r[0]=seed, r[i]=16807*r[i-1] mod (2^31-1),
    r[31..33] copies, then additive lagged Fibonacci r[i]=r[i-31]+r[i-3]
    (mod 2^32); the first 310 values after r[33] are discarded and each
    output is r >> 1."""
    r = [seed]
    for i in range(1, 31):
        r.append((16807 * r[i - 1]) % 2147483647)
    for i in range(31, 34):
        r.append(r[i - 31])
    out = []
    total = 310 + discard + n
    for k in range(total):
        i = 34 + k
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
        if k >= 310 + discard:
            out.append(r[i] >> 1)
    return np.array(out, dtype=np.int64)


def test_never_seeded_drand48_is_the_zero_state(O):
    """A fresh process that never seeds drand48, as each reference rank,
    draws 11/2^48 first: glibc's static state starts at X = 0."""
    import subprocess
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import NEVER_SEEDED
    out = subprocess.run([sys.executable, "-c", NEVER_SEEDED], check=True,
                         capture_output=True, text=True).stdout.split()
    d = np.array([float(v) for v in out])
    assert d[0] == 11 / 2 ** 48
    assert np.array_equal(d, _drand48_closed_form(64))
    assert np.array_equal(d, O.glibc_drand48(64))


def test_glibc_streams(O):
    fx = np.load(os.path.join(GOLDEN, "glibc_streams.npz"))
    assert np.array_equal(O.glibc_drand48(64), fx["drand48_default"])
    assert np.array_equal(fx["drand48_default"], _drand48_closed_form(64))
    r = O.glibc_rand(1, 64, discard=1000)
    assert np.array_equal(r, fx["rand_seed1_after1000"])
    assert np.array_equal(r.astype(np.int64),
                          _glibc_random_r_closed_form(1, 64, 1000))


def test_philox_known_answers(O):
    # Random123 known-answer vectors for philox4x32-10
    kat = [
        ([0, 0, 0, 0], [0, 0],
         [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2,
         [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
         [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, exp in kat:
        assert O.philox(ctr, key).tolist() == exp


def test_box_muller_moments(O):
    n = []
    for i in range(20000):
        n.extend(O.box_muller4(O.philox([i, 7, 0, 0], [1, 2])))
    n = np.array(n)
    assert abs(n.mean()) < 4 / np.sqrt(n.size)
    assert abs(n.var() - 1) < 6 * np.sqrt(2 / n.size)
    # fourth moment of a standard normal is 3
    assert abs(np.mean(n ** 4) - 3) < 0.1


# ---------------------------------------------------------------- physics
def test_rest_state_is_fixed_point_with_zero_drag(O):
    g = O.ref_box_geom(6, 0.6, 0.05)
    g.beta = 0.0
    f = np.broadcast_to(O.W, (g.Lx, g.Ly, 9)).copy()
    b = np.zeros_like(f)
    F = O.step(g, f, b)
    np.testing.assert_allclose(b, f, rtol=0, atol=1e-15)
    assert abs(F) < 1e-14


def test_ref_mass_conserved(O):
    g = O.ref_box_geom(6, 0.55, 0.05)
    rng = np.random.default_rng(0)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + 1e-2 * rng.standard_normal(),
                            0.05 + 1e-2 * rng.standard_normal(),
                            1e-2 * rng.standard_normal())
    fluid = np.ones((g.Lx, g.Ly), bool)
    fluid[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = False
    m0 = f[fluid].sum()
    acc, _ = O.run_steps(g, f, 20)
    assert abs(f[fluid].sum() - m0) / m0 < 1e-13


def test_ref_y_mirror_symmetry(O):
    """Even L centres the square in y; a y-mirrored state stays mirrored."""
    g = O.ref_box_geom(6, 0.55, 0.05)
    mir = np.array([0, 1, 4, 3, 2, 8, 7, 6, 5])
    rng = np.random.default_rng(1)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range((g.Ly + 1) // 2):
            ux = 0.05 + 1e-2 * rng.standard_normal()
            # the middle row is its own mirror image: no transverse velocity
            uy = 1e-2 * rng.standard_normal() if 2 * y != g.Ly - 1 else 0.0
            rho = 1 + 1e-2 * rng.standard_normal()
            f[x, y] = O.feq(rho, ux, uy)
            f[x, g.Ly - 1 - y] = O.feq(rho, ux, -uy)
    f[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = O.W
    assert g.ymin + g.ymax == g.Ly - 1
    O.run_steps(g, f, 10)
    np.testing.assert_allclose(f, f[:, ::-1, :][:, :, mir], rtol=0, atol=1e-13)


def test_channel_uniform_flow_stays_equilibrium_away_from_obstacle(O):
    g = O.channel_geom(60, 30, 6, 40, 15, 0.6, 0.05)
    feq = O.feq(1.0, 0.05, 0.0)
    f = np.broadcast_to(feq, (g.Lx, g.Ly, 9)).copy()
    f[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = O.W
    b = np.zeros_like(f)
    O.step(g, f, b)
    # columns 0..20 (inlet side, two nodes clear of the square's influence),
    # rows 1..Ly-2 (away from the walls)
    np.testing.assert_allclose(b[0:20, 1:-1], np.broadcast_to(feq, b[0:20, 1:-1].shape),
                               rtol=0, atol=2e-16)


def test_channel_momentum_exchange_agrees_with_stress_integral(O):
    """On a developed flow the two drag functionals agree to discretisation
    error (a sign, index or factor-2 mistake in either would not)."""
    g = O.channel_geom(120, 40, 8, 30, 20, 0.56, 0.05)
    f = O.init_channel(g, 1e-3, 0, 0)
    O.run_steps(g, f, 3000)
    Fm = O.lib().orc_force_mea(__import__("ctypes").byref(g), O._dp(f))
    Fs = O.lib().orc_force_stress(__import__("ctypes").byref(g), O._dp(f))
    assert Fm > 0 and Fs > 0
    assert abs(Fm - Fs) / Fm < 0.1, (Fm, Fs)


def test_channel_mass_flux_balances_at_steady_state(O):
    g = O.channel_geom(80, 24, 4, 20, 12, 0.7, 0.04)
    f = O.init_channel(g, 0.0, 0, 0)
    O.run_steps(g, f, 6000)
    jx = f[:, :, [1, 5, 8]].sum(-1) - f[:, :, [3, 6, 7]].sum(-1)
    q_in = jx[2].sum()
    q_out = jx[-3].sum()
    assert abs(q_in - q_out) / q_in < 1e-3


def test_site_normals_bit_layout(O):
    """Pair p of a site's noise reads bits [25p, 25p + 25) of the Philox
    block as one 128-bit number (word 0 least significant): a 13-bit a, then
    a 12-bit b; u1 = (a + 1) / 2^13, u2 = b / 2^12.  The block is built here
    from chosen fields with integer arithmetic, so a wrong offset, word
    order, field width or the +1 changes the normals."""
    rng = np.random.default_rng(11)
    for trial in range(200):
        a = rng.integers(0, 1 << 13, 5)
        b = rng.integers(0, 1 << 12, 5)
        if trial == 0:
            a[:] = 0          # u1 = 2^-13: the largest radius
            b[:] = 0          # angle 0: n = (radius, 0)
        if trial == 1:
            a[:] = (1 << 13) - 1  # u1 = 1: radius 0
        big = 0
        for p in range(5):
            big |= int(a[p]) << (25 * p)
            big |= int(b[p]) << (25 * p + 13)
        big |= int(rng.integers(0, 8)) << 125  # the unused top bits
        r = [(big >> (32 * k)) & 0xFFFFFFFF for k in range(4)]
        n = O.site_normals(r)
        for p in range(5):
            rad = math.sqrt(-2.0 * math.log((a[p] + 1) / 8192.0))
            ang = 2.0 * math.pi * b[p] / 4096.0
            assert abs(n[2 * p] - rad * math.cos(ang)) < 1e-14
            assert abs(n[2 * p + 1] - rad * math.sin(ang)) < 1e-14
    n0 = O.site_normals([0, 0, 0, 0])
    assert n0[0] == pytest.approx(math.sqrt(26.0 * math.log(2.0)), rel=1e-15)
    assert n0[1] == 0.0


def test_site_normals_distribution(O):
    """Each of the nine normals of a site, over 20000 Philox blocks of
    consecutive sites: mean 0, variance 1, no correlation between
    directions, and the normal distribution function to within the
    Kolmogorov-Smirnov bound for that many draws (the 13-bit u1 cuts the
    tail beyond 4.2 sigma, which moves nothing measurable here)."""
    n = np.array([O.site_normals(O.philox([s, 3, 5, 0], [9, 1]))[:9]
                  for s in range(20000)])
    assert np.all(np.abs(n.mean(0)) < 0.03)
    assert np.all(np.abs(n.std(0) - 1.0) < 0.02)
    c = np.corrcoef(n.T)
    assert np.abs(c - np.eye(9)).max() < 0.04
    from scipy import stats
    for q in range(9):
        assert stats.kstest(n[:, q], "norm").statistic < 0.015


def test_noise_is_added_to_fluid_only(O):
    g = O.channel_geom(30, 20, 4, 10, 10, 0.6, 0.05)
    f = O.init_channel(g, 1e-3, 3, 1)
    f0 = f.copy()
    O.noise(g, f, 1e-4, 3, 1, 1, 0)
    d = f - f0
    assert np.all(d[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] == 0)
    fluid = np.ones((g.Lx, g.Ly), bool)
    fluid[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = False
    assert abs(d[fluid].std() / 1e-4 - 1) < 0.05


# ---------------------------------------------------------------- selection
def test_selection_invariants(O):
    rng = np.random.default_rng(5)
    for N in (1, 2, 3, 16, 100, 1000):
        for _ in range(5):
            s = np.exp(rng.normal(0, 1.0, N))
            r = O.select(s, rng.random(N), rng.integers(0, 2**31 - 1, N))
            nb = r["nb_after"]
            assert nb.sum() == N and nb.min() >= 0
            srcs, dsts = r["pairs"][:, 0], r["pairs"][:, 1]
            assert len(set(dsts)) == len(dsts)
            assert np.all(nb[dsts] == 0)
            assert np.all(np.diff(srcs) >= 0)
            for i in np.unique(srcs):
                assert np.sum(srcs == i) == nb[i] - 1
            lab = r["labels"]
            assert np.array_equal(np.bincount(lab, minlength=N), nb)


def test_selection_equal_weights_is_identity(O):
    N = 64
    r = O.select(np.full(N, 4.0), np.random.default_rng(0).random(N) * 0.999,
                 np.zeros(N, np.int32))
    assert len(r["pairs"]) == 0
    assert np.array_equal(r["labels"], np.arange(N))
    assert r["total_R"] == 4.0


def test_selection_floor_step_is_unbiased(O):
    """E[floor(s/R + U)] = s/R."""
    rng = np.random.default_rng(2)
    N = 8
    s = np.array([0.1, 0.5, 1.0, 1.3, 2.2, 0.7, 1.9, 0.3])
    R = s.mean()
    acc = np.zeros(N)
    T = 20000
    for _ in range(T):
        acc += O.select(s, rng.random(N), rng.integers(0, 2**31 - 1, N))["nb_before"]
    mean = acc / T
    sd = np.sqrt((s / R % 1) * (1 - s / R % 1) / T) + 1e-12
    assert np.all(np.abs(mean - s / R) < 5 * sd + 1e-3)


def test_selection_prune_is_uniform_without_replacement(O):
    """With nb_before fixed, the kills of parent j average dN*nb_j/NcPrime
    (hypergeometric mean)."""
    rng = np.random.default_rng(3)
    s = np.array([3.0, 3.0, 0.5, 0.5, 1.0, 1.0])  # R = 1.5 -> s/R = 2,2,1/3,1/3,2/3,2/3
    u = np.array([0.5, 0.5, 0.9, 0.9, 0.5, 0.5])   # nb = 2,2,1,1,1,1 -> NcPrime 8
    kills = np.zeros(6)
    T = 20000
    for _ in range(T):
        r = O.select(s, u, rng.integers(0, 2**31 - 1, 6))
        assert np.array_equal(r["nb_before"], [2, 2, 1, 1, 1, 1])
        kills += r["nb_before"] - r["nb_after"]
    exp = 2 * np.array([2, 2, 1, 1, 1, 1]) / 8
    assert np.all(np.abs(kills / T - exp) < 0.02)


def test_selection_table_matches_reference_printout(O):
    """The reference's copy-table loop (src/mpi_LBM_GK_pops_perturb.cpp:
    595-619) printed `9->0 | 12->1 | 12->5` for nb = {0:0, 1:0, 5:0, 9:2,
    12:3, all others 1} on a 16-clone run; reproduce that table by feeding
    weights and draws that yield exactly those counts."""
    N = 16
    nb = np.ones(N, int)
    nb[[0, 1, 5]] = 0
    nb[9] = 2
    nb[12] = 3
    s = nb.astype(float) + 0.25  # R = 1.25 -> s/R in (0, 3)
    R = s.mean()
    u = (nb - s / R) + 0.5 / R   # floor(s/R + u) = nb exactly
    assert np.all((u >= 0) & (u < 1))
    r = O.select(s, u, np.zeros(N, np.int32))
    assert np.array_equal(r["nb_before"], nb)
    assert r["pairs"].tolist() == [[9, 0], [12, 1], [12, 5]]


def test_scgf_zero_bias_gives_unit_mean_weight(O):
    N = 32
    s = np.exp(0.0 * np.random.default_rng(0).normal(size=N))
    r = O.select(s, np.random.default_rng(1).random(N), np.zeros(N, np.int32))
    assert r["total_R"] == 1.0


def test_brute_force_selection_tiny(O):
    """N=3: enumerate floor draws on a grid and every prune/add draw residue;
    the multiset after prune equals the floor multiset minus a uniformly
    chosen sub-multiset, and the table is the greedy first-empty-slot one."""
    s = np.array([2.0, 0.5, 0.5])  # s/R = 2, .5, .5
    for ua in (0.1, 0.6):
        for ub in (0.2, 0.7):
            u = np.array([0.3, ua, ub])
            nb0 = np.floor(s / s.mean() + u).astype(int)
            for r0 in range(6):
                r = O.select(s, u, np.array([r0, r0 + 1, r0 + 2], np.int32))
                assert np.array_equal(r["nb_before"], nb0)
                assert r["nb_after"].sum() == 3
                empties = [k for k in range(3) if r["nb_after"][k] == 0]
                extras = [i for i in range(3) for _ in range(max(r["nb_after"][i] - 1, 0))]
                assert r["pairs"].tolist() == [[i, k] for i, k in zip(extras, empties)]


def test_scgf_estimator_closed_forms(O):
    """read_rvalue.py's running estimate: constant log R = c gives
    lam_i = (i - 1) c / (i dT); the first interval's R never enters; zero bias
    (R = 1 every interval) gives zero; a sign or off-by-one in either the sum
    or the time fails one of these."""
    n, dT, c = 50, 0.3, 0.0173
    lam = O.scgf(np.full(n, c), dT)
    i = np.arange(1, n)
    assert lam[0] == 0.0
    np.testing.assert_allclose(lam[1:], (i - 1) * c / (i * dT), rtol=1e-13)
    logR = np.random.default_rng(3).normal(0, 0.1, n)
    a, b = logR.copy(), logR.copy()
    b[0] = 40.0
    assert np.array_equal(O.scgf(a, dT), O.scgf(b, dT))
    assert not np.any(O.scgf(np.zeros(n), dT))
    # the last estimate is the mean log R over intervals 1..n-2 scaled by
    # (n-2)/((n-1) dT)
    np.testing.assert_allclose(O.scgf(logR, dT)[-1],
                               logR[1:n - 1].sum() / ((n - 1) * dT), rtol=1e-12)


# ---------------------------------------------------------------- forcing mode
def test_oracle_forced_interval_matches_reference_fixture(O):
    fx = np.load(os.path.join(GOLDEN, "ref_forced_L4.npz"))
    g = _geom_from(O, fx)
    M1, M, eps = int(fx["M1"]), int(fx["M"]), float(fx["eps"])
    a = fx["S0"].copy()
    b = np.zeros_like(a)
    for n in range(M):
        F = O.step_map(g, a, b, fx["map"], eps) if n < M1 else O.step(g, a, b)
        if n == 0:
            assert np.array_equal(a, fx["post0"])
        assert F == fx["forces"][n], f"force step {n}"
        a, b = b, a
    assert np.array_equal(a, fx["final"])
    s = fx["S0"].copy()
    acc, A = O.run_steps_forced(g, s, M1, M, fx["map"], eps, 1.0, 0.0, 1.0)
    assert np.array_equal(s, fx["final"])
    assert np.array_equal(A, fx["forces"])
    tot = 0.0
    for v in fx["forces"]:
        tot += v
    assert acc == tot


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
    "_ref", "libref.so")), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("L,eps,seed", [(4, 0.9, 5), (6, 0.05, 6)])
def test_oracle_spatial_step_matches_live_reference(O, L, eps, seed):
    g = O.ref_box_geom(L, 0.56, 0.07)
    rng = np.random.default_rng(seed)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + 1e-2 * rng.standard_normal(),
                            0.07 + 1e-2 * rng.standard_normal(),
                            1e-2 * rng.standard_normal())
    fmap = rng.standard_normal((g.Lx, g.Ly))
    a, b = f.copy(), np.zeros_like(f)
    ra, rb = f.copy(), np.zeros_like(f)
    for n in range(4):
        F = O.step_map(g, a, b, fmap, eps)
        Fr = O.ref_step_spatial(g, ra, rb, fmap, eps)
        assert np.array_equal(a, ra) and np.array_equal(b, rb), f"step {n}"
        assert F == Fr
        a, b = b, a
        ra, rb = rb, ra


def test_spatial_force_reduces_to_uniform(O):
    g = O.ref_box_geom(4, 0.6, 0.06)
    rng = np.random.default_rng(3)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + 1e-2 * rng.standard_normal(),
                            0.06 + 1e-2 * rng.standard_normal(),
                            1e-2 * rng.standard_normal())
    # a zero map, or eps = 0, is the uniform step bit for bit
    for fmap, eps in [(np.zeros((g.Lx, g.Ly)), 0.8),
                      (rng.standard_normal((g.Lx, g.Ly)), 0.0)]:
        a, b = f.copy(), np.zeros_like(f)
        c, d = f.copy(), np.zeros_like(f)
        assert O.step_map(g, a, b, fmap, eps) == O.step(g, c, d)
        assert np.array_equal(a, c) and np.array_equal(b, d)
    # a constant map c is the uniform step with beta' = (1 + eps*c)*beta
    cval, eps = 0.37, 0.5
    a, b = f.copy(), np.zeros_like(f)
    O.step_map(g, a, b, np.full((g.Lx, g.Ly), cval), eps)
    g2 = O.ref_box_geom(4, 0.6, 0.06)
    g2.beta = (1.0 + eps * cval) * g.beta
    c, d = f.copy(), np.zeros_like(f)
    O.step(g2, c, d)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    # the map is read as map[y + Ly*x]: a force on one site changes only the
    # post-collision populations of that site
    fmap = np.zeros((g.Lx, g.Ly))
    fmap[3, 11] = 1.0
    a, b = f.copy(), np.zeros_like(f)
    c, d = f.copy(), np.zeros_like(f)
    O.step_map(g, a, b, fmap, 0.5)
    O.step(g, c, d)
    diff = np.argwhere((a != c).any(axis=2))
    assert diff.tolist() == [[3, 11]]


def _reference_embedding_loops(n, sigma):
    """Literal restatement of generate_random_field.cpp:44-111 (Rows/Cols,
    row/col padding, G = row/N) with covariance_fctn.cpp:3-10, for small n."""
    M = 2 * n - 1
    N = M * M
    f = lambda x, y: math.exp(-float(x * x + y * y) / (2 * sigma * sigma))
    Rows = [[f(R, y) for y in range(n)] for R in range(n)]
    Cols = [[f(-R, y) for y in range(n)] for R in range(n)]
    row = [[0.0] * M for _ in range(M)]
    col = [[0.0] * M for _ in range(n)]
    for R in range(n):
        for j in range(n):
            row[R][j] = Rows[R][j]
            col[R][j] = Cols[R][j]
        ii = 0
        for j in range(n, M):
            ii += 1
            row[R][j] = Cols[R][n - ii]
            col[R][j] = Rows[R][n - ii]
    ii = 0
    for R in range(n, M):
        ii += 1
        for j in range(M):
            row[R][j] = col[n - ii][j]
    return np.array([[row[R][j] / N for j in range(M)] for R in range(M)])


@pytest.mark.parametrize("n,sigma", [(3, 32.0), (5, 32.0), (8, 1.7), (6, 3.0)])
def test_field_embedding_matches_reference_construction(O, n, sigma):
    assert np.array_equal(O.field_embedding(n, n, sigma),
                          _reference_embedding_loops(n, sigma))


@pytest.mark.parametrize("Lx,Ly,sigma", [(241, 241, 32.0), (60, 40, 3.0),
                                         (33, 70, 2.0)])
def test_field_covariance_is_the_target_without_clipping(O, Lx, Ly, sigma):
    G = O.field_embedding(Lx, Ly, sigma)
    raw = np.fft.fft2(G).real
    assert raw.min() > -1e-15  # nothing clipped beyond rounding
    lam = O.field_eigenvalues(Lx, Ly, sigma)
    cov = O.field_covariance(lam, Lx, Ly)
    dx = np.arange(Lx)[:, None]
    dy = np.arange(Ly)[None, :]
    target = np.exp(-(dx * dx + dy * dy) / (2 * sigma * sigma))
    assert np.abs(cov - target).max() < 1e-12


def test_field_eigenvalue_clipping_is_the_int_abs_of_the_reference(O):
    # generate_random_field.cpp:147 tests abs(a) > 1e-15 on a double with
    # <cmath>, <cstdlib> and "complex.h" included: that abs is the int one,
    # so a negative eigenvalue above -1 is zeroed, not flagged
    import shutil
    import subprocess
    import tempfile

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = ('#include <cmath>\n#include <cstdlib>\n#include "complex.h"\n'
           '#include <cstdio>\nint main(){double a=-0.5;'
           'std::printf("%d %d\\n",(int)(abs(a)>1e-15),(int)(abs(-1.5)>1e-15));}\n')
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "p.cpp"), "w") as fh:
            fh.write(src)
        subprocess.run(["g++", "-w", "-o", os.path.join(d, "p"),
                        os.path.join(d, "p.cpp")], check=True)
        out = subprocess.run([os.path.join(d, "p")], capture_output=True,
                             text=True, check=True).stdout.split()
    assert out == ["0", "1"]
    # small lattices at sigma 32 have negative eigenvalues; the oracle zeroes
    # them and the field's variance then exceeds one
    lam_raw = np.fft.fft2(O.field_embedding(17, 17, 32.0)).real
    assert lam_raw.min() < -1e-3
    lam = O.field_eigenvalues(17, 17, 32.0)
    assert np.array_equal(lam, np.where(lam_raw < 0, 0.0, lam_raw))
    assert O.field_covariance(lam, 17, 17)[0, 0] > 1.0


def test_field_normals_follow_the_philox_layout(O):
    Z = O.field_normals(0x1234_5678_9ABC, clone=5, interval=3, instance=2,
                        shape=(3, 5))
    flat = Z.ravel()
    for p in range(8):
        r = O.philox([p, 5, 3, 0xF1E1D000 + 2], [0x56789ABC, 0x1234])
        v = O.box_muller4(r)
        assert flat[2 * p] == complex(v[0], v[1])
        if 2 * p + 1 < flat.size:
            assert flat[2 * p + 1] == complex(v[2], v[3])


def test_field_sample_statistics(O):
    Lx, Ly, sigma = 24, 16, 2.5
    lam = O.field_eigenvalues(Lx, Ly, sigma)
    cov = O.field_covariance(lam, Lx, Ly)
    rng = np.random.default_rng(11)
    fields = np.stack([
        O.field_from_normals(lam, rng.standard_normal(lam.shape)
                             + 1j * rng.standard_normal(lam.shape), Lx, Ly)
        for _ in range(600)])
    assert abs(fields.mean()) < 0.05
    for dx, dy in [(0, 0), (1, 0), (0, 2), (3, 3), (6, 0)]:
        a = fields[:, : Lx - dx, : Ly - dy]
        b = fields[:, dx:, dy:]
        est = (a * b).mean()
        assert abs(est - cov[dx, dy]) < 0.06, (dx, dy, est, cov[dx, dy])


# ---------------------------------------------------------------- macroscopic output
def test_oracle_macro_matches_reference_fixture(O):
    fx = np.load(os.path.join(GOLDEN, "ref_macro_L4.npz"))
    g = _geom_from(O, fx)
    assert np.array_equal(O.macro(g, fx["S0"]), fx["uniform"])
    assert np.array_equal(O.macro(g, fx["S0"], fx["map"], float(fx["eps"])),
                          fx["spatial"])


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
    "_ref", "libref.so")), reason="reference build (oracle/_ref) not present")
def test_oracle_macro_matches_live_reference(O):
    g = O.ref_box_geom(6, 0.6, 0.08)
    rng = np.random.default_rng(12)
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + 1e-2 * rng.standard_normal(),
                            0.08 + 1e-2 * rng.standard_normal(),
                            1e-2 * rng.standard_normal())
    fmap = rng.standard_normal((g.Lx, g.Ly))
    for m, eps in [(None, 0.0), (fmap, 0.3)]:
        want = O.ref_step_fields(g, f.copy(), np.zeros_like(f), m, eps)
        assert np.array_equal(O.macro(g, f, m, eps), want)


def test_macro_of_equilibrium_recovers_its_moments(O):
    # feq(rho, u) has density rho and momentum rho*u, so the fields are rho,
    # u_x + beta/(2 rho) and u_y (to rounding)
    g = O.ref_box_geom(4, 0.6, 0.05)
    rng = np.random.default_rng(2)
    f = np.empty((g.Lx, g.Ly, 9))
    want = np.empty((3, g.Lx, g.Ly))
    for x in range(g.Lx):
        for y in range(g.Ly):
            r, u, v = 1 + 0.1 * rng.random(), 0.05 * rng.random(), 0.03 * rng.random()
            f[x, y] = O.feq(r, u, v)
            want[:, x, y] = (r, u + 0.5 * g.beta / r, v)
    np.testing.assert_allclose(O.macro(g, f), want, rtol=1e-13, atol=1e-15)


def test_oracle_is_clean_under_sanitizers():
    """Every oracle function, driven by oracle/san_check.c, built with
    AddressSanitizer and UndefinedBehaviorSanitizer: no report, exit 0."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc or make not available")
    d = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "oracle")
    b = subprocess.run(["make", "-s", "-C", d, "san"], capture_output=True,
                       text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([os.path.join(d, "_san", "san_check")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("san_check: ok")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.skipif(not __import__("oracle").ref_available(),
                    reason="reference kernels not built")
def test_reference_clone_loop_equals_its_steps(O):
    """ref_run_clones (the reference's kernels in the driver's loop, used to
    time the reference on the CPU) is the same sequence of ref_step calls."""
    g = O.ref_box_geom(4, 0.53, 0.05)
    rng = np.random.default_rng(8)
    states = np.stack([np.broadcast_to(O.W, (g.Lx, g.Ly, 9)) *
                       (1 + 1e-3 * rng.standard_normal((g.Lx, g.Ly, 1)))
                       for _ in range(3)])
    states[:, g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = O.W
    for M in (4, 5):
        want = states.copy()
        acc_want = np.zeros(3)
        for j in range(3):
            a, b = want[j].copy(), np.zeros_like(want[j])
            for _ in range(M):
                acc_want[j] += O.ref_step(g, a, b)
                a, b = b, a
            want[j] = a
        got = states.copy()
        O.ref_set_geometry(g)
        acc = O.ref_run_clones(g, got, M)
        assert np.array_equal(got, want) and np.array_equal(acc, acc_want)
