"""Macroscopic output of the step kernel (macro_every > 0): rho, ux, uy of
every clone at steps 0, E, 2E, ... of an interval, the fields the
reference's step computes and its VTK writer prints
(src/lbm_src/streamCollCompute.cpp:17-26,44-45,
src/mpi_LBM_GK_pops_perturb.cpp:450-455).  FP64 frames are bit-exact against
the oracle in every kernel shape; output steps run as lone passes, so the
sweep grouping changes and the states and drag must stay bit-exact too."""
import numpy as np
import pytest

from test_gpu_parity import geom_of, rand_state

pytestmark = pytest.mark.gpu


def _oracle_frames(O, g, state, M, E, M1=0, fmap=None, eps=0.0, scale=1.0,
                   shift=0.0):
    """Step the oracle through one interval; frames (n_frames, 3, Lx, Ly) and
    the per-step drag."""
    a = state.copy()
    b = np.zeros_like(a)
    frames, drag = [], []
    for n in range(M):
        m = fmap if n < M1 else None
        if n % E == 0:
            frames.append(O.macro(g, a, m, eps))
        F = O.step_map(g, a, b, m, eps) if m is not None else O.step(g, a, b)
        drag.append(F * scale - shift)
        a, b = b, a
    return np.array(frames), np.array(drag), a


@pytest.mark.parametrize("threads,passes,M1", [(256, 1, 0), (256, 2, 0),
                                               (512, 2, 7), (256, 3, 0),
                                               (256, 3, 7)])
def test_ref_macro_frames_f64_bitexact(O, lbc, threads, passes, M1):
    L, M, N, E, eps = 8, 20, 3, 5, 0.6
    spec = lbc.ref_box(L, 0.53, 0.05, N, M, alpha=2.0, zero_mean=False)
    spec.macro_every = E
    spec.forced_steps = M1
    spec.field_eps = eps
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(21)
    states = np.stack([rand_state(O, g, rng, amp=1e-2) for _ in range(N)])
    maps = rng.standard_normal((N, g.Lx, g.Ly))
    print(f"case: REF macro frames F64 M={M} E={E} M1={M1} threads={threads} "
          f"passes={passes}", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_REF, threads=threads,
                     passes=passes) as s:
        s.load_states(0, 0, states)
        if M1:
            s.set_field_maps(0, maps)
        s.run_interval()
        mac = s.macro(0)
        got = s.store_states(0)
        drag = s.drag(0)
    assert mac.shape == (N, 4, 3, g.Ly, g.Lx)
    for j in range(N):
        fr, dr, fin = _oracle_frames(O, g, states[j], M, E, M1, maps[j], eps,
                                     spec.force_scale, spec.force_shift)
        assert np.array_equal(mac[j].transpose(0, 1, 3, 2), fr), f"clone {j}"
        assert np.array_equal(got[j], fin), f"clone {j} state"
        assert np.array_equal(drag[j], dr), f"clone {j} drag"


def test_channel_macro_frames_f64_bitexact(O, lbc):
    spec = lbc.channel(80, 40, 6, 20, 20, 0.55, 0.05, 2, 12, k=1.0)
    spec.macro_every = 4
    g = geom_of(O, spec, O.MODE_CHANNEL)
    print("case: CHANNEL macro frames F64 80x40 M=12 E=4", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=2, threads=256,
                     passes=2) as s:
        s.init_equilibrium(0, 1e-3)
        init = s.store_states(0)
        s.run_interval()
        mac = s.macro(0)
    for j in range(2):
        fr, _, _ = _oracle_frames(O, g, init[j], 12, 4)
        assert np.array_equal(mac[j].transpose(0, 1, 3, 2), fr), f"clone {j}"


@pytest.mark.parametrize("storage", [2, 1])
def test_shifted_macro_frames_agree_across_shapes(O, lbc, storage):
    """FP16/FP32 frames are identical in one, two and three passes per sweep
    and close to the FP64 oracle's."""
    spec = lbc.ref_box(40, 0.55, 0.05, 2, 9, alpha=1.0, zero_mean=False)
    spec.macro_every = 3
    g = O.ref_box_geom(40, 0.55, 0.05)
    rng = np.random.default_rng(5)
    st = np.stack([rand_state(O, g, rng, amp=1e-2) for _ in range(2)])
    print(f"case: shifted macro frames storage={storage}", flush=True)
    out = []
    for passes in (1, 2, 3):
        with lbc.Sampler(spec, mode=lbc.MODE_REF, storage=storage, threads=256,
                         passes=passes) as s:
            s.load_states(0, 0, st)
            s.run_interval()
            out.append(s.macro(0))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    tol = 2e-3 if storage == 2 else 1e-6
    for j in range(2):
        fr, _, _ = _oracle_frames(O, g, st[j], 9, 3)
        err = np.abs(out[0][j].transpose(0, 1, 3, 2) - fr).max()
        print(f"storage {storage} clone {j}: max field error {err:.2e}")
        assert err < tol
