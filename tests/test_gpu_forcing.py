"""Forcing mode on the GPU (src/mpi_LBM_GK_forcing.cpp:333-395): the first
forced_steps steps of every interval run with the per-site body force
(1 + eps*map)*beta, map a Gaussian random field per clone and interval.

The step kernel with injected maps must match the FP64 oracle bit for bit in
every kernel shape, including sweeps split at the last forced step.  The
device's random fields must equal the oracle's restatement of
src/perturbation_routines/generate_random_field.cpp given the same Philox
normals: to 1e-12 where the embedding's eigenvalues are well separated from
zero, and to 1e-5 (absolute, fields of unit variance) where most of them sit
at rounding level, because rocFFT and numpy round those tiny eigenvalues
differently and their square roots then differ by ~1e-9 each.
"""
import numpy as np
import pytest

from test_gpu_parity import rand_state

pytestmark = pytest.mark.gpu


def _states(O, g, n, seed, u0=0.05):
    rng = np.random.default_rng(seed)
    return np.stack([rand_state(O, g, rng, amp=1e-2, u0=u0) for _ in range(n)])


# (FP64 rings of three passes fit in LDS with 256 threads only)
@pytest.mark.parametrize("threads,passes,M1", [(256, 1, 7), (512, 1, 0),
                                               (256, 2, 6), (512, 2, 7),
                                               (256, 3, 5), (256, 3, 6),
                                               (256, 2, 20)])
def test_forced_interval_f64_bitexact(O, lbc, threads, passes, M1):
    L, M, N, eps = 8, 20, 4, 0.6
    spec = lbc.ref_box(L, 0.53, 0.05, N, M, alpha=2.0, zero_mean=False)
    spec.forced_steps = M1
    spec.field_eps = eps
    g = O.ref_box_geom(L, 0.53, 0.05)
    states = _states(O, g, N, 3)
    maps = np.random.default_rng(4).standard_normal((N, g.Lx, g.Ly))
    print(f"case: forced interval F64 L={L} M={M} M1={M1} threads={threads} "
          f"passes={passes}", flush=True)
    if M1 == 0:
        with pytest.raises(lbc.LbcError):
            with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=threads,
                             passes=passes) as s:
                s.set_field_maps(0, maps)  # no forced steps: refused
        return
    with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=threads,
                     passes=passes) as s:
        s.load_states(0, 0, states)
        s.set_field_maps(0, maps)
        assert np.array_equal(s.field_maps(0), maps)
        s.run_interval()
        got = s.store_states(0)
        drag = s.drag(0)
        obs, _ = s.logweights(0)
        assert np.array_equal(s.field_maps(0), maps)
    for j in range(N):
        ref = states[j].copy()
        acc, A = O.run_steps_forced(g, ref, M1, M, maps[j], eps,
                                    spec.force_scale, spec.force_shift,
                                    spec.dt)
        assert np.array_equal(got[j], ref), f"clone {j} state"
        assert np.array_equal(drag[j], A), f"clone {j} drag"
        assert obs[j] == acc, f"clone {j} observable"


@pytest.mark.parametrize("sigma,atol", [(1.5, 1e-12), (0.0, 1e-5)])
def test_device_fields_match_oracle(O, lbc, sigma, atol):
    L, N, seed = 4, 7, 0x5EED_0001_2345  # odd: the last transform has one clone
    spec = lbc.ref_box(L, 0.55, 0.05, N, 5, alpha=1.0, zero_mean=False)
    spec.forced_steps = 3
    spec.field_eps = 0.4
    spec.field_sigma = sigma
    g = O.ref_box_geom(L, 0.55, 0.05)
    lam = O.field_eigenvalues(g.Lx, g.Ly, sigma if sigma > 0 else 32.0)
    states = _states(O, g, N, 8)
    print(f"case: device fields vs oracle L={L} sigma={sigma}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, seed=seed, threads=256,
                     passes=2) as s:
        s.load_states(0, 0, states)
        for t in range(2):
            start = s.store_states(0)
            s.run_interval()
            maps = s.field_maps(0)
            got = s.store_states(0)
            for j in range(N):
                want = O.field_of_clone(lam, seed, j, t, 0, g.Lx, g.Ly)
                if j % 2 == 0:  # the even clone's is the reference's form
                    Z = O.field_normals(seed, j, t, 0, lam.shape)
                    assert np.array_equal(
                        want, O.field_from_normals(lam, Z, g.Lx, g.Ly))
                err = np.abs(maps[j] - want).max()
                assert err < atol, f"interval {t} clone {j}: {err:.3g}"
                # the interval ran with the maps it reports
                ref = start[j].copy()
                O.run_steps_forced(g, ref, 3, 5, maps[j], 0.4,
                                   spec.force_scale, spec.force_shift, spec.dt)
                assert np.array_equal(got[j], ref), f"interval {t} clone {j}"
            assert not np.array_equal(maps[0], maps[1])
            s.select(0)
            s.gather()


def test_device_fields_large_lattice_batches_and_statistics(O, lbc):
    """Side-241 lattice (the smallest whose sigma-32 embedding needs no
    clipping): 600 clones need 300 transforms, two clones each; at most 290
    fit the work buffer, so they run as two batches of 150 (clones 0-299 and
    300-599); clones on both sides of the batch boundary match the
    oracle, and the sample covariance over all fields, real and imaginary
    parts alike, matches the target exp(-r^2 / 2048)."""
    L, N, seed = 60, 600, 77
    spec = lbc.ref_box(L, 0.55, 0.05, N, 1, alpha=1.0, zero_mean=False)
    spec.forced_steps = 1
    spec.field_eps = 0.1
    g = O.ref_box_geom(L, 0.55, 0.05)
    assert g.Lx == 241
    lam = O.field_eigenvalues(g.Lx, g.Ly, 32.0)
    print(f"case: device fields 241x241, {N} clones", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, seed=seed, threads=256,
                     passes=1) as s:
        s.run_interval()
        maps = s.field_maps(0)
    assert np.all(np.isfinite(maps))
    for j in (0, 1, 298, 299, 300, 301, 599):
        want = O.field_of_clone(lam, seed, j, 0, 0, g.Lx, g.Ly)
        err = np.abs(maps[j] - want).max()
        assert err < 1e-5, f"clone {j}: {err:.3g}"
    assert abs(maps.mean()) < 0.1
    for part in (maps[0::2], maps[1::2], maps):  # Re, Im, both
        for dx, dy in [(0, 0), (16, 0), (0, 32), (24, 24), (64, 0)]:
            a = part[:, : g.Lx - dx, : g.Ly - dy]
            b = part[:, dx:, dy:]
            est = float((a * b).mean())
            target = np.exp(-(dx * dx + dy * dy) / 2048.0)
            assert abs(est - target) < 0.12, (dx, dy, est, target)
    # the two parts of one transform are independent: no cross-correlation
    cross = float((maps[0::2] * maps[1::2]).mean())
    assert abs(cross) < 0.1, cross


@pytest.mark.parametrize("storage", [2, 1])
@pytest.mark.parametrize("M1", [4, 5])
def test_forced_multi_pass_sweeps_equal_single_passes(O, lbc, storage, M1):
    """Shifted storages: two or three passes per sweep give the same bits as
    one pass per sweep when the forced steps end inside a would-be sweep."""
    spec = lbc.ref_box(40, 0.55, 0.05, 3, 9, alpha=2.0, zero_mean=False)
    spec.forced_steps = M1
    spec.field_eps = 0.5
    g = O.ref_box_geom(40, 0.55, 0.05)
    st = _states(O, g, 3, 1)
    maps = np.random.default_rng(2).standard_normal((3, g.Lx, g.Ly))
    print(f"case: forced multi-pass storage={storage} M1={M1}", flush=True)
    out = []
    for passes in (1, 2, 3):
        with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage, threads=256,
                         passes=passes) as s:
            s.load_states(0, 0, st)
            s.set_field_maps(0, maps)
            s.run_interval()
            out.append((s.store_states(0), s.drag(0), s.logweights(0)[1]))
    for p, o in zip((2, 3), out[1:]):
        for name, u, v in zip(("states", "drag", "logw"), out[0], o):
            assert np.array_equal(u, v), f"{p} passes: {name} differ"
    # and close to the FP64 oracle
    for j in range(3):
        ref = st[j].copy()
        acc, A = O.run_steps_forced(g, ref, M1, 9, maps[j], 0.5,
                                    spec.force_scale, spec.force_shift, spec.dt)
        rel = np.abs(out[0][1][j] - A).max() / np.abs(A).max()
        print(f"storage {storage} clone {j}: drag rel err {rel:.2e}")
        assert rel < (1e-2 if storage == 2 else 1e-4), (j, rel)


@pytest.mark.parametrize("storage", [0, 2])
def test_forced_two_clones_per_workgroup_equal_one(O, lbc, storage):
    """The one-pass step kernel with the random-field force gives the same
    bits with two clones per workgroup (three clones: one pair and a padded
    one) as with one."""
    spec = lbc.ref_box(40, 0.55, 0.05, 3, 9, alpha=2.0, zero_mean=False)
    spec.forced_steps = 5
    spec.field_eps = 0.5
    g = O.ref_box_geom(40, 0.55, 0.05)
    st = _states(O, g, 3, 1)
    maps = np.random.default_rng(2).standard_normal((3, g.Lx, g.Ly))
    print(f"case: forced two clones per workgroup storage={storage}", flush=True)
    out = []
    for cpw in (1, 2):
        with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage, threads=256,
                         passes=1, cpw=cpw) as s:
            assert s.cpw == cpw
            s.load_states(0, 0, st)
            s.set_field_maps(0, maps)
            s.run_interval()
            out.append((s.store_states(0), s.drag(0), s.logweights(0)[1]))
    for name, u, v in zip(("states", "drag", "logw"), out[0], out[1]):
        assert np.array_equal(u, v), f"two clones per workgroup: {name} differ"


@pytest.mark.parametrize("N", [7, 1])
@pytest.mark.parametrize("storage", [0, 1, 2, 3])
def test_fields_of_every_storage_match_oracle(O, lbc, storage, N):
    """The shifted storages draw their fields with a single-precision
    transform of the same Philox normals: within 1e-4 of the oracle's FP64
    fields (a wrong eigenvalue weight, part or clone pairing is O(1) off),
    and the reported FP64 maps are exact widenings of FP32 values.  FP64
    storage matches to 1e-12.  One clone: a single transform slot, whose
    single-precision work buffer is smaller than the FP64 spectrum
    transform run in it at set-up."""
    L, seed = 4, 0x5EED_0001_2345
    spec = lbc.ref_box(L, 0.55, 0.05, N, 5, alpha=1.0, zero_mean=False)
    spec.forced_steps = 3
    spec.field_eps = 0.4
    spec.field_sigma = 1.5
    g = O.ref_box_geom(L, 0.55, 0.05)
    lam = O.field_eigenvalues(g.Lx, g.Ly, 1.5)
    print(f"case: fields storage={storage} clones={N}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage, seed=seed,
                     threads=256, passes=2) as s:
        s.load_states(0, 0, _states(O, g, N, 8))
        for t in range(2):
            s.run_interval()
            maps = s.field_maps(0)
            if storage != 0:
                assert np.array_equal(maps, maps.astype(np.float32))
            atol = 1e-12 if storage == 0 else 1e-4
            for j in range(N):
                want = O.field_of_clone(lam, seed, j, t, 0, g.Lx, g.Ly)
                err = np.abs(maps[j] - want).max()
                assert err < atol, f"interval {t} clone {j}: {err:.3g}"
            s.select(0)  # advances the interval the next draws use
            s.gather()


@pytest.mark.parametrize("storage", [1, 2])
def test_shifted_storage_reads_fp32_maps(O, lbc, storage):
    """Shifted storages read the FP32 rounding of the field maps: injected
    maps and their FP32 roundings give the same bits, and an interval run
    with device-drawn fields equals the same interval re-run with those
    fields injected (the draw's FP32 copy is the rounding of its FP64 maps)."""
    spec = lbc.ref_box(12, 0.55, 0.05, 5, 9, alpha=2.0, zero_mean=False)
    spec.forced_steps = 5
    spec.field_eps = 0.5
    spec.field_sigma = 4.0
    g = O.ref_box_geom(12, 0.55, 0.05)
    st = _states(O, g, 5, 6)
    maps = np.random.default_rng(9).standard_normal((5, g.Lx, g.Ly))
    print(f"case: fp32 maps storage={storage}", flush=True)

    def run(m, seed=3):
        with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage, seed=seed,
                         threads=256, passes=2) as s:
            s.load_states(0, 0, st)
            if m is not None:
                s.set_field_maps(0, m)
            s.run_interval()
            return s.store_states(0), s.drag(0), s.field_maps(0)

    a = run(maps)
    b = run(maps.astype(np.float32).astype(np.float64))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the maps reported are the FP32 values the interval ran with
    assert np.array_equal(a[2], maps.astype(np.float32).astype(np.float64))
    c = run(maps + 1e-3)  # the maps are used
    assert not np.array_equal(a[1], c[1])
    drawn = run(None)
    again = run(drawn[2])
    assert np.array_equal(drawn[0], again[0])
    assert np.array_equal(drawn[1], again[1])


def test_forced_advance_graph_matches_stepwise(lbc):
    """lbc_advance draws each interval's fields outside the captured graph;
    the result equals the same sequence issued call by call."""
    spec = lbc.forcing_box(8, 0.55, 0.05, 8, dT=0.1, dT0=0.04, eps=0.5,
                           alpha=400.0, field_sigma=3.0)
    # T0 = 160: floor(0.04*160) = 6 forced, floor(0.06*160) = 9 uniform
    assert spec.forced_steps == 6 and spec.steps == 15
    print("case: forced advance graph vs stepwise, REF F64", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, seed=5) as a, \
            lbc.Sampler(spec, mode=lbc.MODE_REF, seed=5) as b:
        a.init_equilibrium(0, 1e-3)
        b.init_equilibrium(0, 1e-3)
        a.advance(3)
        copies = 0
        for _ in range(3):
            b.run_interval()
            copies += len(b.select(0).pairs)
            b.gather()
        # the bias is large enough that clones are copied: the forcing mode
        # has no snapshots, so a copy is not perturbed in either path
        assert copies > 0
        assert np.array_equal(a.field_maps(0), b.field_maps(0))
        sa = a.store_states(0)
        assert np.all(np.isfinite(sa)) and np.array_equal(sa, b.store_states(0))
        Ra, Rb = a.R_history(0), b.R_history(0)
        assert len(Ra) == 3 and np.array_equal(Ra, Rb)
        assert np.all(np.isfinite(Ra)) and np.all(Ra > 0)


def test_forced_instance_beside_a_perturbed_one(lbc):
    """A forced instance (no snapshots) and a snapshot-perturbed instance in
    one context, with copies in both: the perturbation kernel runs for the
    second instance's copies and leaves the first's alone, in the
    call-by-call sequence and in the captured graphs (a copy of a clone with
    no snapshots to mix was once turned into NaN)."""
    a = lbc.forcing_box(8, 0.55, 0.05, 8, dT=0.1, dT0=0.04, eps=0.5,
                        alpha=400.0, field_sigma=3.0)
    b = lbc.ref_box(6, 0.6, 0.05, 8, 12, alpha=400.0, eps=0.1, n_snapshots=3)
    print("case: forced + perturbed instances with copies", flush=True)
    for graph in (False, True):
        with lbc.Sampler([a, b], mode=lbc.MODE_REF, seed=9) as s:
            s.init_equilibrium(0, 1e-3)
            s.init_equilibrium(1, 1e-3)
            s.set_snapshots(1, s.store_states(1, 0, 3))
            copies = [0, 0]
            if graph:
                s.advance(3)
            else:
                for _ in range(3):
                    s.run_interval()
                    for i in range(2):
                        copies[i] += len(s.select(i).pairs)
                    s.gather()
                assert copies[0] > 0 and copies[1] > 0
            for i in range(2):
                assert np.all(np.isfinite(s.store_states(i))), (graph, i)
                R = s.R_history(i)
                assert len(R) == 3 and np.all(np.isfinite(R)), (graph, i)


def test_forced_and_uniform_instances_batched(O, lbc):
    """A forced and an unforced instance in one context (one launch per
    interval) give the same bits as each in a context of its own."""
    a = lbc.ref_box(4, 0.55, 0.05, 3, 10, alpha=1.0, zero_mean=False)
    a.forced_steps = 4
    a.field_eps = 0.5
    b = lbc.ref_box(6, 0.6, 0.05, 2, 12, alpha=2.0, zero_mean=False)
    ga = O.ref_box_geom(4, 0.55, 0.05)
    gb = O.ref_box_geom(6, 0.6, 0.05)
    sa = _states(O, ga, 3, 31)
    sb = _states(O, gb, 2, 32)
    maps = np.random.default_rng(33).standard_normal((3, ga.Lx, ga.Ly))
    print("case: forced + uniform instances batched, REF F64", flush=True)

    def run(specs, loads):
        with lbc.Sampler(specs, mode=lbc.MODE_REF, threads=256, passes=2) as s:
            for i, (st, fmap) in enumerate(loads):
                s.load_states(i, 0, st)
                if fmap is not None:
                    s.set_field_maps(i, fmap)
            s.run_interval()
            return [(s.store_states(i), s.drag(i)) for i in range(len(specs))]

    both = run([a, b], [(sa, maps), (sb, None)])
    alone_a = run([a], [(sa, maps)])[0]
    alone_b = run([b], [(sb, None)])[0]
    for got, want in ((both[0], alone_a), (both[1], alone_b)):
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # and the forced instance is the oracle's forced interval
    for j in range(3):
        ref = sa[j].copy()
        _, A = O.run_steps_forced(ga, ref, 4, 10, maps[j], 0.5, a.force_scale,
                                  a.force_shift, a.dt)
        assert np.array_equal(both[0][0][j], ref) and np.array_equal(both[0][1][j], A)
