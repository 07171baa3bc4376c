"""GPU parity of the HIP path against the CPU FP64 oracle, through the C ABI.

FP64 storage must match the oracle bit for bit (lattice fields, per-step
drag, interval observable, selection, copies).  Quantities that go through
libm transcendental functions on the device (noise, perturbation norms,
exp of log-weights) are compared at 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_state(O, g, rng, amp=1e-3, u0=0.05):
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            f[x, y] = O.feq(1 + amp * rng.standard_normal(),
                            u0 + amp * rng.standard_normal(),
                            amp * rng.standard_normal())
    f += 1e-4 * amp * rng.standard_normal(f.shape)
    # solid interior at rest, as the reference's reset leaves it
    f[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax, :] = O.W
    return f


def geom_of(O, spec, mode):
    return O.Geom(mode, spec.Lx, spec.Ly, spec.xmin, spec.xmax, spec.ymin,
                  spec.ymax, spec.tau, spec.beta if mode == O.MODE_REF else 0.0,
                  spec.u_in if mode == O.MODE_CHANNEL else 0.0)


@pytest.mark.parametrize("L", [4, 8])
def test_ref_step_host_bitexact(O, lbc, L):
    spec = lbc.ref_box(L, 0.53, 0.05, 1, 1, 1.0)
    g = O.ref_box_geom(L, 0.53, 0.05)
    print(f"case: lbc_step_host REF L={L}", flush=True)
    fin = rand_state(O, g, np.random.default_rng(L))
    a, b = fin.copy(), np.zeros_like(fin)
    F = O.step(g, a, b)
    pc, nxt, Fg = lbc.step_host(lbc.MODE_REF, spec, fin)
    assert np.array_equal(nxt, b)
    assert np.array_equal(pc, a)
    assert Fg == F


@pytest.mark.parametrize("storage_threads,passes", [(256, 1), (1024, 1),
                                                  (256, 2), (512, 2), (256, 3)])
def test_ref_interval_f64_bitexact(O, lbc, storage_threads, passes):
    L, M, N = 8, 37, 5
    spec = lbc.ref_box(L, 0.515, 0.05, N, M, alpha=3.0)
    g = O.ref_box_geom(L, 0.515, 0.05)
    rng = np.random.default_rng(11)
    states = np.stack([rand_state(O, g, rng) for _ in range(N)])
    print(f"case: REF interval F64 L={L} M={M} N={N} threads={storage_threads} "
          f"passes={passes}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=storage_threads,
                     passes=passes) as s:
        s.load_states(0, 0, states)
        s.run_interval()
        got = s.store_states(0)
        A = s.drag(0)
        obs, lw = s.logweights(0)
    for j in range(N):
        ref = states[j].copy()
        acc, Aref = O.run_steps(g, ref, M, spec.force_scale, spec.force_shift,
                                spec.dt)
        assert np.array_equal(got[j], ref), f"clone {j} fields"
        assert np.array_equal(A[j], Aref), f"clone {j} drag"
        assert obs[j] == acc
        assert lw[j] == spec.bias * acc


@pytest.mark.parametrize("M,passes", [(1, 1), (2, 1), (24, 1),
                                      (1, 2), (2, 2), (3, 2), (24, 2),
                                      (1, 3), (2, 3), (3, 3), (4, 3), (5, 3),
                                      (24, 3)])
def test_channel_interval_f64_bitexact(O, lbc, M, passes):
    """Bit-exact against the oracle for every interval length that changes
    how the passes group into sweeps (lone collide pass, groups of 2 or 3
    stream-collide passes, remainders, lone stream pass)."""
    spec = lbc.channel(80, 40, 8, 20, 20, 0.53, 0.05, 3, M, k=0.01)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print(f"case: CHANNEL interval F64 80x40 M={M} passes={passes}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=5, threads=256,
                     passes=passes) as s:
        s.init_equilibrium(0, 1e-3)
        init = s.store_states(0)
        for j in range(3):
            ref0 = O.init_channel(g, 1e-3, 5, j)
            assert np.array_equal(init[j], ref0), f"init clone {j}"
        s.run_interval()
        got = s.store_states(0)
        A = s.drag(0)
        obs, lw = s.logweights(0)
    for j in range(3):
        ref = init[j].copy()
        acc, Aref = O.run_steps(g, ref, M)
        assert np.array_equal(got[j], ref), f"clone {j} fields"
        assert np.array_equal(A[j], Aref), f"clone {j} drag"
        assert obs[j] == acc
        assert lw[j] == 0.01 * acc


@pytest.mark.parametrize("Lx,Ly,side,cx,cy",
                         [(8, 8, 2, 4, 4), (63, 12, 4, 20, 6),
                          (64, 12, 4, 20, 6), (65, 13, 4, 20, 6),
                          (130, 9, 2, 60, 4)])
@pytest.mark.parametrize("mode", ["channel", "ref"])
def test_lattice_shapes_f64_bitexact(O, lbc, Lx, Ly, side, cx, cy, mode):
    """Lattice widths where a 64-site wave covers several rows (8, 63),
    exactly one (64), or straddles a row end (65, 130), the smallest lattice
    accepted (8x8), a 9-row lattice whose walls are close to the square,
    both x-boundary rules (inlet/outlet, periodic), one, two and three
    passes per sweep: FP64 fields, drag and observables bit-exact against
    the oracle over 7 steps."""
    M = 7
    spec = lbc.channel(Lx, Ly, side, cx, cy, 0.56, 0.05, 2, M, k=0.01)
    if mode == "ref":
        spec.u_in = 0.0
        spec.beta = 1e-5
    m = O.MODE_CHANNEL if mode == "channel" else O.MODE_REF
    g = geom_of(O, spec, m)
    rng = np.random.default_rng(Lx * 100 + Ly)
    states = np.stack([rand_state(O, g, rng, amp=1e-2) for _ in range(2)])
    for passes in (1, 2, 3):
        print(f"case: {mode} {Lx}x{Ly} passes={passes}", flush=True)
        try:
            s = lbc.Sampler(spec, mode=lbc.MODE_CHANNEL if mode == "channel"
                            else lbc.MODE_REF, threads=256, passes=passes)
        except lbc.LbcError as e:
            # the periodic wrap reaches 2*Lx sites: at 130 wide the two FP64
            # rings of three passes need 2048 sites each and exceed LDS
            assert "do not fit" in str(e) and mode == "ref" and Lx == 130
            assert passes == 3
            continue
        with s:
            s.load_states(0, 0, states)
            s.run_interval()
            got = s.store_states(0)
            A = s.drag(0)
            obs, _ = s.logweights(0)
        for j in range(2):
            ref = states[j].copy()
            acc, Aref = O.run_steps(g, ref, M, spec.force_scale,
                                    spec.force_shift, spec.dt)
            assert np.array_equal(got[j], ref), (passes, j)
            assert np.array_equal(A[j], Aref), (passes, j)
            assert obs[j] == acc


def test_sixty_four_instances_in_one_context(O, lbc):
    """The most instances a context takes (64), lattices of 16..31 x 12,
    one to three clones each, biases and interval lengths varying: every
    instance's clones bit-exact against the oracle after one interval, then
    per-instance selection and copies, then two more cloning steps from the
    captured graphs with finite mean weights everywhere."""
    specs = []
    for i in range(64):
        Lx = 16 + (i % 16)
        specs.append(lbc.channel(Lx, 12, 2, Lx // 3, 6, 0.55 + 0.001 * i, 0.05,
                                 1 + i % 3, 2 + i % 4, k=0.001 * (i + 1),
                                 sigma=1e-4))
    print("case: 64 instances in one context", flush=True)
    with lbc.Sampler(specs, mode=lbc.MODE_CHANNEL, seed=11, threads=256,
                     passes=2) as s:
        for i in range(64):
            s.init_equilibrium(i, 1e-3)
        init = [s.store_states(i) for i in range(64)]
        s.run_interval()
        for i, sp in enumerate(specs):
            g = geom_of(O, sp, O.MODE_CHANNEL)
            got = s.store_states(i)
            obs, _ = s.logweights(i)
            for j in range(sp.n_clones):
                ref = init[i][j].copy()
                acc, _ = O.run_steps(g, ref, sp.steps)
                assert np.array_equal(got[j], ref), (i, j)
                assert obs[j] == acc, (i, j)
        for i in range(64):
            s.select(i)
        s.gather()
        s.advance(2)
        for i in range(64):
            R = s.R_history(i)
            assert len(R) == 3 and np.all(np.isfinite(R)) and np.all(R > 0), i


@pytest.mark.parametrize("passes", [1, 2, 3])
def test_batched_uneven_instances_f64_bitexact(O, lbc, passes):
    """Eight samplers of four lattice sizes and two biases in one launch per
    interval, with different clone counts and interval lengths: every
    instance's fields, per-step drag and log-weights equal the oracle's after
    two intervals, and each instance selects on its own weights."""
    dims = [(40, 10), (60, 15), (80, 20), (120, 30)]
    specs = []
    for i, (Lx, Ly) in enumerate(dims):
        for j, k in enumerate((0.002, 0.01)):
            specs.append(lbc.channel(Lx, Ly, Ly // 5, Lx // 4, Ly // 2, 0.53,
                                     0.05, 2 + i + j, 5 + j, k=k))
    print(f"case: batched uneven instances passes={passes}", flush=True)
    rng = np.random.default_rng(11)
    with lbc.Sampler(specs, mode=lbc.MODE_CHANNEL, seed=9, threads=256,
                     passes=passes) as s:
        init = []
        for i in range(len(specs)):
            s.init_equilibrium(i, 1e-3)
            init.append(s.store_states(i))
        s.run_interval()
        s.run_interval()
        for i, spec in enumerate(specs):
            g = geom_of(O, spec, O.MODE_CHANNEL)
            got = s.store_states(i)
            A = s.drag(i)
            obs, lw = s.logweights(i)
            for c in range(spec.n_clones):
                assert np.array_equal(init[i][c],
                                      O.init_channel(g, 1e-3, 9, c, i)), \
                    f"instance {i} init clone {c}"
                ref = init[i][c].copy()
                O.run_steps(g, ref, spec.steps)
                acc, Aref = O.run_steps(g, ref, spec.steps)
                assert np.array_equal(got[c], ref), f"instance {i} clone {c}"
                assert np.array_equal(A[c], Aref), f"instance {i} drag {c}"
                assert obs[c] == acc
                assert lw[c] == spec.bias * acc
            N = spec.n_clones
            w = np.exp(rng.normal(0, 1, N))
            u = rng.random(N)
            r = rng.integers(0, 2**31 - 1, N, dtype=np.int64).astype(np.int32)
            sel = s.select(i, weights=w, u=u, rint=r)
            ref = O.select(w, u, r)
            assert np.array_equal(sel.pairs, ref["pairs"])
            assert np.array_equal(sel.labels, ref["labels"])


@pytest.mark.parametrize("N,p", [(1, 1), (2, 1), (7, 1), (64, 4), (1000, 8),
                                 (4096, 1), (8192, 2)])
def test_selection_matches_oracle(O, lbc, N, p):
    rng = np.random.default_rng(N)
    spec = lbc.channel(32, 16, 4, 8, 8, 0.6, 0.05, N, 1, k=0.0)
    print(f"case: selection N={N} p={p}", flush=True)
    spec.rank_blocks = p
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, threads=256) as s:
        for trial in range(3):
            w = np.exp(rng.normal(0, 0.5 + trial, N))
            u = rng.random(N)
            r = rng.integers(0, 2**31 - 1, N, dtype=np.int64).astype(np.int32)
            ref = O.select(w, u, r, p_blocks=p)
            got = s.select(0, weights=w, u=u, rint=r)
            assert np.array_equal(got.nb_before, ref["nb_before"])
            assert np.array_equal(got.nb_after, ref["nb_after"])
            assert np.array_equal(got.pairs, ref["pairs"])
            assert np.array_equal(got.labels, ref["labels"])
            assert got.total_R == ref["total_R"]
            assert got.n_rint_used == ref["n_rint"]


def test_gather_copies_parents_exactly(O, lbc):
    L, N = 4, 8
    spec = lbc.ref_box(L, 0.53, 0.05, N, 3, alpha=1.0)
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(3)
    states = np.stack([rand_state(O, g, rng) for _ in range(N)])
    print("case: gather REF F64", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=256) as s:
        s.load_states(0, 0, states)
        with pytest.raises(lbc.LbcError):
            s.gather_kernel_ms()  # nothing gathered yet
        w = np.array([5.0, 0.1, 0.1, 3.0, 0.1, 1.0, 0.1, 0.1])
        u = rng.random(N)
        r = rng.integers(0, 2**31 - 1, N).astype(np.int32)
        sel = s.select(0, weights=w, u=u, rint=r)
        ref = O.select(w, u, r)
        assert np.array_equal(sel.labels, ref["labels"])
        assert len(sel.pairs) > 0
        s.gather()
        gms, gbytes = s.gather_kernel_ms()
        got = s.store_states(0)
        marks = s.get_marks(0)
    # the copy kernel's bytes: parent read + destination written, FP64
    copies = int(np.sum(sel.labels != np.arange(N)))
    assert copies == len(sel.pairs)
    assert gbytes == 2 * 9 * g.Lx * g.Ly * 8 * copies and gms > 0
    for k in range(N):
        assert np.array_equal(got[k], states[sel.labels[k]])
        assert marks[k] == (sel.labels[k] != k)


def test_channel_noise_on_copies(O, lbc):
    spec = lbc.channel(48, 24, 6, 12, 12, 0.55, 0.05, 6, 2, k=0.0, sigma=1e-4)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print("case: gather CHANNEL F64 with noise", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=123, threads=256) as s:
        s.init_equilibrium(0, 1e-3)
        before = s.store_states(0)
        w = np.array([4.0, 0.01, 0.01, 1.0, 0.5, 0.5])
        u = np.full(6, 0.5)
        r = np.arange(6, dtype=np.int32)
        sel = s.select(0, weights=w, u=u, rint=r)
        s.gather()
        after = s.store_states(0)
    assert len(sel.pairs) > 0
    for k in range(6):
        exp = before[sel.labels[k]].copy()
        if sel.labels[k] != k:
            O.noise(g, exp, 1e-4, 123, k, 1, 0)
            np.testing.assert_allclose(after[k], exp, rtol=0, atol=1e-17)
            assert not np.array_equal(after[k], before[sel.labels[k]])
        else:
            assert np.array_equal(after[k], exp)


def test_channel_noise_on_copies_f32(O, lbc):
    """FP32-shifted storage draws the same Philox words as FP64, keeps 24
    bits of each and turns them into normals with the hardware
    log2/sqrt/sin/cos.  Each copied population's noise is the FP64 oracle's
    to within FP32 rounding of the shifted value (< 4e-9 here) plus 1e-3 of
    the noise (the 24-bit uniforms; ~1e-7 typically, more only where the
    radius draw is tiny); a wrong or missing draw is off by ~100%."""
    spec = lbc.channel(48, 24, 6, 12, 12, 0.55, 0.05, 6, 2, k=0.0, sigma=1e-4)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print("case: gather CHANNEL F32 with noise", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, storage=lbc.STORE_F32,
                     seed=123, threads=256) as s:
        s.init_equilibrium(0, 1e-3)
        before = s.store_states(0)
        w = np.array([4.0, 0.01, 0.01, 1.0, 0.5, 0.5])
        u = np.full(6, 0.5)
        r = np.arange(6, dtype=np.int32)
        sel = s.select(0, weights=w, u=u, rint=r)
        s.gather()
        after = s.store_states(0)
    assert len(sel.pairs) > 0
    interior = np.zeros((g.Lx, g.Ly), bool)
    interior[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = True
    for k in range(6):
        exp = before[sel.labels[k]].copy()
        if sel.labels[k] != k:
            O.noise(g, exp, 1e-4, 123, k, 1, 0)
            got_n = after[k] - before[sel.labels[k]]
            ref_n = exp - before[sel.labels[k]]
            err = np.abs(got_n - ref_n) - (8e-9 + 1e-3 * np.abs(ref_n))
            assert err.max() <= 0, (k, err.max())
            moved = np.abs(got_n)
            assert np.all(moved[~interior].max(axis=-1) > 0)
            assert np.array_equal(after[k][interior], before[sel.labels[k]][interior])
        else:
            assert np.array_equal(after[k], exp)


def test_ref_perturbation(O, lbc):
    L, N, NN, M = 4, 4, 3, 5
    spec = lbc.ref_box(L, 0.53, 0.05, N, M, alpha=1.0, eps=0.3, n_snapshots=NN)
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(17)
    states = np.stack([rand_state(O, g, rng) for _ in range(N)])
    snaps = np.stack([rand_state(O, g, rng, amp=1e-2) for _ in range(NN)])
    pref = rng.random((N, NN))
    marks = np.array([0, 1, 1, 0], np.uint8)
    print("case: REF snapshot perturbation F64", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=256) as s:
        s.load_states(0, 0, states)
        s.set_snapshots(0, snaps)
        s.set_marks(0, marks)
        s.run_interval(prefactors=pref)
        got = s.store_states(0)
        A = s.drag(0)
        assert not s.get_marks(0).any()
    for j in range(N):
        ref = states[j].copy()
        if marks[j]:
            O.perturb(ref, snaps, 0.3, pref[j])
        acc, Aref = O.run_steps(g, ref, M, spec.force_scale, spec.force_shift,
                                spec.dt)
        np.testing.assert_allclose(got[j], ref, rtol=1e-12, atol=0)
        np.testing.assert_allclose(A[j], Aref, rtol=1e-9, atol=1e-12)
        if not marks[j]:
            assert np.array_equal(got[j], ref)


@pytest.mark.parametrize("storage,tol", [(1, 2e-6), (2, 6e-5)])
def test_ref_perturbation_shifted_storage(O, lbc, storage, tol):
    """The snapshot perturbation on FP32- and FP16-shifted states: the mix
    is computed in FP64 from the stored values plus the rest weights, then
    stored shifted again.  After the perturbation and one step, every clone
    is within the storage's rounding of the FP64 oracle's, and the marked
    clones moved by far more than that."""
    L, N, NN = 4, 4, 3
    spec = lbc.ref_box(L, 0.53, 0.05, N, 1, alpha=1.0, eps=0.3, n_snapshots=NN)
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(29)
    states = np.stack([rand_state(O, g, rng) for _ in range(N)])
    snaps = np.stack([rand_state(O, g, rng, amp=1e-2) for _ in range(NN)])
    pref = rng.random((N, NN))
    marks = np.array([0, 1, 1, 0], np.uint8)
    print(f"case: REF snapshot perturbation storage={storage}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage,
                     threads=256) as s:
        s.load_states(0, 0, states)
        start = s.store_states(0)  # the states as stored
        s.set_snapshots(0, snaps)
        s.set_marks(0, marks)
        s.run_interval(prefactors=pref)
        got = s.store_states(0)
    for j in range(N):
        ref = start[j].copy()
        if marks[j]:
            O.perturb(ref, snaps, 0.3, pref[j])
        plain = start[j].copy()
        O.run_steps(g, ref, 1, spec.force_scale, spec.force_shift, spec.dt)
        O.run_steps(g, plain, 1, spec.force_scale, spec.force_shift, spec.dt)
        err = np.abs(got[j] - ref).max()
        assert err < tol, (j, err)
        if marks[j]:
            assert np.abs(ref - plain).max() > 20 * tol


def test_advance_graph_matches_stepwise(lbc):
    """lbc_advance (captured graph, device draws) equals the same sequence
    issued call by call."""
    spec = lbc.channel(64, 32, 6, 16, 16, 0.55, 0.05, 16, 10, k=2.0,
                       sigma=1e-4)
    print("case: advance graph vs stepwise, CHANNEL F64", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=9, threads=256) as a, \
            lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=9, threads=256) as b:
        a.init_equilibrium(0, 1e-3)
        b.init_equilibrium(0, 1e-3)
        a.advance(3)
        for _ in range(3):
            b.run_interval()
            b.select(0)
            b.gather()
        assert np.array_equal(a.store_states(0), b.store_states(0))
        Ra, Rb = a.R_history(0), b.R_history(0)
        assert len(Ra) == 3 and np.array_equal(Ra, Rb)
        assert np.all(np.isfinite(Ra)) and np.all(Ra > 0)


@pytest.mark.parametrize("storage,passes", [(0, 1), (0, 2), (2, 3), (1, 3),
                                            (3, 2)])
def test_advance_fused_copies_match_stepwise(lbc, storage, passes):
    """Inside lbc_advance the noisy copies of every interval but the last
    are collided straight into the next interval's pass 0 (k_gather with
    fuse0), and that pass is skipped for them.  The states, drags and R
    history equal the call-by-call sequence (stored copies, full pass 0) bit
    for bit in every storage and sweep shape, copies happen in the fused
    intervals, and a plain interval afterwards still matches (the labels
    are restored)."""
    spec = lbc.channel(96, 40, 8, 24, 20, 0.55, 0.05, 24, 12, k=4.0,
                       sigma=1e-4)
    print(f"case: advance with fused copies vs stepwise, CHANNEL storage "
          f"{storage} passes {passes}", flush=True)
    kw = dict(mode=lbc.MODE_CHANNEL, storage=storage, seed=21, threads=256,
              nt=True, passes=passes)
    with lbc.Sampler(spec, **kw) as a, lbc.Sampler(spec, **kw) as b:
        a.init_equilibrium(0, 1e-3)
        b.init_equilibrium(0, 1e-3)
        a.advance(4)
        copies = []
        for _ in range(4):
            b.run_interval()
            copies.append(len(b.select(0).pairs))
            b.gather()
        assert sum(copies[:3]) > 0, copies
        assert np.array_equal(a.store_states(0), b.store_states(0))
        assert np.array_equal(a.drag(0), b.drag(0))
        assert np.array_equal(a.R_history(0), b.R_history(0))
        a.run_interval()
        b.run_interval()
        assert np.array_equal(a.store_states(0), b.store_states(0))
        assert np.array_equal(a.logweights(0)[1], b.logweights(0)[1])


@pytest.mark.parametrize("storage,passes", [(0, 1), (2, 3), (3, 2)])
def test_gather_fused_matches_gather(lbc, storage, passes):
    """lbc_gather_fused followed by an interval equals lbc_gather followed
    by the same interval, bit for bit (states, drags, log-weights); in
    between, the calls that read or reselect states are refused."""
    spec = lbc.channel(96, 40, 8, 24, 20, 0.55, 0.05, 12, 10, k=0.0,
                       sigma=1e-4)
    print(f"case: gather_fused vs gather, CHANNEL storage {storage} passes "
          f"{passes}", flush=True)
    kw = dict(mode=lbc.MODE_CHANNEL, storage=storage, seed=5, threads=256,
              nt=True, passes=passes)
    w = np.where(np.arange(12) % 3 == 0, 1.0, 1e-9)
    u = np.full(12, 0.5)
    r = np.arange(12, dtype=np.int32)
    with lbc.Sampler(spec, **kw) as a, lbc.Sampler(spec, **kw) as b:
        for s in (a, b):
            s.init_equilibrium(0, 1e-3)
            s.run_interval()
        sa = a.select(0, weights=w, u=u, rint=r)
        sb = b.select(0, weights=w, u=u, rint=r)
        assert np.array_equal(sa.labels, sb.labels) and len(sa.pairs) > 0
        a.gather_fused()
        b.gather()
        with pytest.raises(lbc.LbcError, match=r"\(-4\)"):
            a.store_states(0)
        with pytest.raises(lbc.LbcError, match=r"\(-4\)"):
            a.select(0, weights=w, u=u, rint=r)
        a.run_interval()
        b.run_interval()
        assert np.array_equal(a.store_states(0), b.store_states(0))
        assert np.array_equal(a.drag(0), b.drag(0))
        assert np.array_equal(a.logweights(0)[1], b.logweights(0)[1])
        # labels are restored: a plain cycle afterwards still agrees
        for s in (a, b):
            s.select(0, weights=w, u=u, rint=r)
            s.gather()
            s.run_interval()
        assert np.array_equal(a.store_states(0), b.store_states(0))


@pytest.mark.parametrize("storage,tol", [(1, 2e-6), (2, 1e-3), (3, 1e-2)])
def test_low_precision_channel_interval(O, lbc, storage, tol):
    """FP32-collide on shifted storage vs the FP64 oracle from identical
    states over one interval: per-clone interval drag within `tol` relative."""
    spec = lbc.channel(200, 50, 10, 50, 25, 0.515, 0.05, 4, 100, k=0.01)
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print(f"case: low precision storage={storage} 200x50 M=100", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=0, threads=512) as s64:
        s64.init_equilibrium(0, 1e-3)
        s64.run_interval()  # develop a little flow in FP64
        init = s64.store_states(0)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, storage=storage,
                     threads=512) as s:
        s.load_states(0, 0, init)
        s.run_interval()
        obs, _ = s.logweights(0)
        got = s.store_states(0)
    for j in range(4):
        ref = init[j].copy()
        acc, _ = O.run_steps(g, ref, 100)
        rel = abs(obs[j] - acc) / abs(acc)
        print(f"storage {storage} clone {j}: drag sum {obs[j]:.9g} vs "
              f"{acc:.9g} rel {rel:.2e}")
        assert rel < tol
        assert np.max(np.abs(got[j] - ref)) < 50 * tol


@pytest.mark.parametrize("storage,tol", [(1, 2e-6), (2, 2e-4), (3, 5e-3)])
def test_low_precision_ref_interval(O, lbc, storage, tol):
    """The reference mode (periodic x, Guo body force, stress-integral drag)
    on shifted storage vs the FP64 oracle over one 80-step interval of the
    reference driver's L = 8 box, from developed-flow states: interval drag
    within `tol` relative, fields within 10 * tol."""
    spec = lbc.ref_box(8, 0.6, 0.1, 3, 80, alpha=1.0, zero_mean=False)
    g = O.ref_box_geom(8, 0.6, 0.1)
    rng = np.random.default_rng(41)
    init = np.stack([rand_state(O, g, rng, amp=1e-2, u0=0.1)
                     for _ in range(3)])
    print(f"case: REF low precision storage={storage} 33x33 M=80", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage,
                     threads=256) as s:
        s.load_states(0, 0, init)
        start = s.store_states(0)
        s.run_interval()
        obs, _ = s.logweights(0)
        got = s.store_states(0)
    for j in range(3):
        ref = start[j].copy()
        acc, _ = O.run_steps(g, ref, 80, spec.force_scale, spec.force_shift,
                             spec.dt)
        rel = abs(obs[j] - acc) / abs(acc)
        print(f"storage {storage} clone {j}: drag {obs[j]:.9g} vs {acc:.9g} "
              f"rel {rel:.2e}")
        assert rel < tol
        assert np.max(np.abs(got[j] - ref)) < 10 * tol


@pytest.mark.parametrize("storage", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_multi_pass_sweeps_equal_single_passes(O, lbc, storage, mode):
    """Two or three passes per sweep through the LDS rings give bit-identical
    states, drag and weights to one pass per sweep (the rings hold the
    storage precision), in both modes, with a lattice row longer than one
    chunk."""
    if mode == 0:
        spec = lbc.ref_box(40, 0.55, 0.05, 3, 7, alpha=2.0)   # 161x161
    else:
        spec = lbc.channel(300, 40, 8, 75, 20, 0.55, 0.05, 3, 8, k=0.5)
    print(f"case: multi-pass vs single-pass sweeps storage={storage} "
          f"mode={mode}", flush=True)
    out = []
    for passes in (1, 2, 3):
        with lbc.Sampler(spec, mode=mode, storage=storage, seed=3, threads=256,
                         passes=passes) as s:
            if mode == 0:
                g = O.ref_box_geom(40, 0.55, 0.05)
                rng = np.random.default_rng(1)
                st = np.stack([rand_state(O, g, rng) for _ in range(3)])
                s.load_states(0, 0, st)
            else:
                s.init_equilibrium(0, 1e-3)
            s.run_interval()
            s.run_interval()
            out.append((s.store_states(0), s.drag(0), s.logweights(0)[1]))
    (a, da, la) = out[0]
    for p, (b, db, lb) in zip((2, 3), out[1:]):
        for name, u, v in (("states", a, b), ("drag", da, db),
                           ("logw", la, lb)):
            bad = np.argwhere(u != v)
            assert bad.size == 0, (
                f"{p} passes, {name}: {len(bad)} of {u.size} differ, first at "
                f"{bad[0].tolist()} ({u[tuple(bad[0])]!r} vs "
                f"{v[tuple(bad[0])]!r})")


@pytest.mark.parametrize("storage", [0, 2])
@pytest.mark.parametrize("mode", [0, 1])
def test_two_clones_per_workgroup_equal_one(lbc, storage, mode):
    """Two clones per workgroup of the one-pass step kernel (pairs of clones
    of one instance, a padding partner for an odd count) give bit-identical
    states, drag and weights to one clone per workgroup, for two instances
    of different lattices and odd clone counts."""
    if mode == 0:
        specs = [lbc.ref_box(8, 0.55, 0.05, 5, 9, alpha=2.0),
                 lbc.ref_box(4, 0.55, 0.05, 3, 9, alpha=2.0)]
    else:
        specs = [lbc.channel(300, 40, 8, 75, 20, 0.55, 0.05, 5, 9, k=0.5),
                 lbc.channel(120, 30, 6, 30, 15, 0.55, 0.05, 3, 9, k=0.5)]
    print(f"case: two clones per workgroup vs one, storage={storage} "
          f"mode={mode}", flush=True)
    out = []
    for cpw in (1, 2):
        with lbc.Sampler(specs, mode=mode, storage=storage, seed=5,
                         threads=256, passes=1, cpw=cpw) as s:
            assert s.cpw == cpw
            for i in range(len(specs)):
                s.init_equilibrium(i, 1e-3)
            s.run_interval()
            s.run_interval()
            out.append([(s.store_states(i), s.drag(i), s.logweights(i)[1])
                        for i in range(len(specs))])
    for i in range(len(specs)):
        for name, u, v in zip(("states", "drag", "logw"), out[0][i],
                              out[1][i]):
            bad = np.argwhere(u != v)
            assert bad.size == 0, (
                f"instance {i}, {name}: {len(bad)} of {u.size} differ, first "
                f"at {bad[0].tolist()}")


def test_scgf_estimate_matches_oracle(O, lbc):
    """The device's running large-deviation estimate equals the oracle's
    restatement of read_rvalue.py on the sampler's own mean weights; with
    zero bias every mean weight is exactly 1 and the estimate exactly 0."""
    print("case: SCGF estimate 80x40, 64 clones, 12 cloning steps", flush=True)
    dT = 20.0
    for k in (0.05, 0.0):
        spec = lbc.channel(80, 40, 8, 20, 20, 0.53, 0.05, 64, 20, k=k,
                           sigma=1e-4)
        with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=2, threads=256) as s:
            s.init_equilibrium(0, 1e-3)
            s.advance(12)
            R = s.R_history(0)
            lam = s.scgf(0, dT)
        assert len(R) == 12 and len(lam) == 12
        if k == 0.0:
            assert np.all(R == 1.0) and not np.any(lam)
        else:
            np.testing.assert_allclose(lam, O.scgf(np.log(R), dT), rtol=1e-12,
                                       atol=1e-16)
            assert np.any(lam != 0.0)
