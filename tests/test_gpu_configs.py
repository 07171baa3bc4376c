"""The channel workloads at their real shapes (lattice, square, tau, inlet
velocity, bias, noise, clones, steps per interval), checked through
properties that do not depend on size: FP64 clones bit-exact against the
oracle over a whole interval, every step-kernel shape giving the same bits,
low-precision storage within the drag bar of FP64, finite positive mean
weights over a run."""
import numpy as np
import pytest

from test_gpu_parity import geom_of

pytestmark = pytest.mark.gpu


def _oracle_interval(O, g, state, M):
    s = state.copy()
    acc, A = O.run_steps(g, s, M)
    return acc, A, s


def test_small_channel_run_200x50(O, lbc):
    """200x50, square 10 at (50, 25), tau 0.515, u_in 0.05, k 0.01,
    sigma 1e-4, 16 clones, 100 steps: the first interval of every clone
    bit-exact against the oracle, then 19 more cloning steps on the device."""
    spec = lbc.channel(200, 50, 10, 50, 25, 0.515, 0.05, 16, 100, k=0.01,
                       sigma=1e-4)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print("case: channel 200x50, 16 clones, 100 steps, 20 intervals", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=0) as s:
        s.init_equilibrium(0, 1e-3)
        init = s.store_states(0)
        for j in range(16):
            assert np.array_equal(init[j], O.init_channel(g, 1e-3, 0, j))
        s.run_interval()
        obs, _ = s.logweights(0)
        got = s.store_states(0)
        drag = s.drag(0)
        for j in range(16):
            acc, A, fin = _oracle_interval(O, g, init[j], 100)
            assert obs[j] == acc and np.array_equal(drag[j], A), j
            assert np.array_equal(got[j], fin), j
        s.select(0)
        s.gather()
        s.advance(19)
        R = s.R_history(0)
    assert len(R) == 20 and np.all(np.isfinite(R)) and np.all(R > 0)
    assert np.all(np.abs(np.log(R)) < 1.0)


def test_fp64_channel_400x100_512_clones(O, lbc):
    """400x100, square 20 at (100, 50), tau 0.53, k 0.005, sigma 1e-4, 512
    clones, 200 steps, FP64: first and last clone bit-exact against the
    oracle over the interval."""
    spec = lbc.channel(400, 100, 20, 100, 50, 0.53, 0.05, 512, 200, k=0.005,
                       sigma=1e-4)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print("case: channel 400x100, 512 clones FP64, 200 steps", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=4) as s:
        s.init_equilibrium(0, 1e-3)
        init = s.store_states(0, 0, 1), s.store_states(0, 511, 1)
        s.run_interval()
        obs, _ = s.logweights(0)
        last = s.store_states(0, 511, 1)
    for (j, st) in ((0, init[0][0]), (511, init[1][0])):
        acc, _, fin = _oracle_interval(O, g, st, 200)
        assert obs[j] == acc, j
        if j == 511:
            assert np.array_equal(last[0], fin)


BATCH = [(200, 50, 1024), (300, 75, 512), (400, 100, 384), (600, 150, 128)]


def _batch_specs(lbc):
    return [lbc.channel(Lx, Ly, Ly // 5, Lx // 4, Ly // 2, 0.53, 0.05, n, 200,
                        k=k, sigma=1e-4)
            for k in (0.002, 0.01) for (Lx, Ly, n) in BATCH]


def test_batched_eight_samplers_4096_clones(O, lbc):
    """Eight samplers ({200x50, 300x75, 400x100, 600x150} x k {0.002, 0.01},
    square Ly/5 at (Lx/4, Ly/2), tau 0.53, 4096 clones in all, 200 steps) in
    one launch: one clone of every sampler bit-exact in FP64, and FP16
    within the drag bar of FP64 on every clone."""
    specs = _batch_specs(lbc)
    print("case: batched 8 samplers, 4096 clones, 200 steps", flush=True)
    with lbc.Sampler(specs, mode=lbc.MODE_CHANNEL, seed=6) as s64:
        for i in range(8):
            s64.init_equilibrium(i, 1e-3)
        init = [s64.store_states(i, sp.n_clones // 2, 1)[0]
                for i, sp in enumerate(specs)]
        s64.run_interval()
        obs64 = [s64.logweights(i)[0] for i in range(8)]
    for i, sp in enumerate(specs):
        g = geom_of(O, sp, O.MODE_CHANNEL)
        j = sp.n_clones // 2
        acc, _, _ = _oracle_interval(O, g, init[i], 200)
        assert obs64[i][j] == acc, (i, j)
    # the same seeded initial states, rounded to the storage
    with lbc.Sampler(specs, mode=lbc.MODE_CHANNEL, storage=lbc.STORE_F16,
                     seed=6) as s16:
        for i in range(8):
            s16.init_equilibrium(i, 1e-3)
        s16.run_interval()
        for i in range(8):
            obs16 = s16.logweights(i)[0]
            rel = np.abs(obs16 - obs64[i]) / np.abs(obs64[i])
            print(f"sampler {i}: FP16 drag rel err max {rel.max():.2e}")
            assert rel.max() < 1e-3, i


def test_large_channel_800x200(O, lbc):
    """800x200, square 40 at (200, 100), tau 0.52, u_in 0.04, k 0.002,
    sigma 5e-5, 400 steps: all 2048 clones run one FP16 interval; on 48 of
    them one, two and three passes per sweep give the same bits and stay
    within the drag bar of FP64, and one clone is bit-exact in FP64."""
    def spec_of(n):
        return lbc.channel(800, 200, 40, 200, 100, 0.52, 0.04, n, 400, k=0.002,
                           sigma=5e-5)
    print("case: channel 800x200, 2048 clones FP16, 400 steps", flush=True)
    with lbc.Sampler(spec_of(2048), mode=lbc.MODE_CHANNEL,
                     storage=lbc.STORE_F16, seed=8) as s:
        s.init_equilibrium(0, 1e-3)
        s.run_interval()
        obs, _ = s.logweights(0)
        s.select(0)
        s.gather()
        assert np.all(np.isfinite(obs))
        R = s.R_history(0)
        assert np.isfinite(R[0]) and R[0] > 0
    spec = spec_of(48)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=8) as s64:
        s64.init_equilibrium(0, 1e-3)
        init = s64.store_states(0)
        s64.run_interval()
        obs64 = s64.logweights(0)[0]
    acc, _, _ = _oracle_interval(O, g, init[0], 400)
    assert obs64[0] == acc
    out = []
    for passes in (1, 2, 3):
        with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, storage=lbc.STORE_F16,
                         threads=512 if passes < 3 else 256,
                         passes=passes) as s16:
            s16.load_states(0, 0, init)
            s16.run_interval()
            out.append((s16.logweights(0)[0], s16.store_states(0, 0, 4)))
    for p in (1, 2):
        assert np.array_equal(out[0][0], out[p][0]), p + 1
        assert np.array_equal(out[0][1], out[p][1]), p + 1
    rel = np.abs(out[0][0] - obs64) / np.abs(obs64)
    print(f"800x200 FP16 drag rel err max {rel.max():.2e}")
    assert rel.max() < 1e-3
