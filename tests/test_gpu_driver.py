"""End-to-end parity of the driver (tlgk_lbm_gpu) against the reference's
driver loop (src/mpi_LBM_GK_pops_perturb.cpp:387-714) restated with the
CPU oracle and glibc's generators: same input file, same pops files, same
labels.dat / rvalues.dat / data_force.datout."""
import math
import os
import subprocess
from collections import deque

import numpy as np
import pytest

import refloop
from conftest import PKG
from refloop import make_pops, parse_output

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG, "driver", "tlgk_lbm_gpu")
RECON = os.path.join(PKG, "post", "reconstruct")


def write_case(tmp, O, L, Nc, T, dT, eps, tau, U0, alpha, extra=""):
    g = O.ref_box_geom(L, tau, U0)
    pops = make_pops(O, L, max(Nc, 10), L * 100 + Nc, tau, U0)
    data, _, inp = refloop.write_case(tmp, pops, Nc, T, dT, eps, L, tau, U0,
                                      alpha, extra)
    return g, pops, inp, data


def reference_loop(O, g, pops, Nc, T, dT, eps, U0, alpha, L, p=1):
    r = refloop.reference_loop(O, g, pops, Nc, T, dT, eps, U0, alpha, L, p)
    return r["labels"], r["R"], r["forces"], r["states"]


@pytest.mark.parametrize("Nc,alpha", [(6, 40.0), (8, 0.0)])
def test_driver_matches_reference_loop(tmp_path, O, Nc, alpha):
    L, T, dT, eps, tau, U0 = 4, 1.0, 0.25, 0.02, 0.53, 0.05
    g, pops, inp, _ = write_case(tmp_path, O, L, Nc, T, dT, eps, tau, U0,
                                 alpha, extra="dump_states = 1\n")
    print(f"case: driver REF L={L} Nc={Nc} alpha={alpha}", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    labels_ref, R_ref, forces_ref, states_ref = reference_loop(
        O, g, pops, Nc, T, dT, eps, U0, alpha, L)
    run = tmp_path / "run"
    labels = np.fromfile(run / "labels.dat", dtype=np.int32).reshape(-1, Nc)
    assert np.array_equal(labels, np.array(labels_ref))
    R = np.fromfile(run / "rvalues.dat")
    np.testing.assert_allclose(R, R_ref, rtol=1e-12, atol=0)
    # the first interval precedes any perturbation: bit-exact
    for j in range(Nc):
        Fj = np.fromfile(run / f"clone_{j}" / "data_force.datout").reshape(-1, 20)
        assert np.array_equal(Fj[0], forces_ref[j][0])
        np.testing.assert_allclose(Fj, np.array(forces_ref[j]), rtol=1e-10,
                                   atol=1e-12)
        fin = np.fromfile(run / f"clone_{j}" / "pops_final.datout").reshape(
            g.Lx, g.Ly, 9)
        np.testing.assert_allclose(fin, states_ref[j], rtol=1e-11, atol=0)
    out = (tmp_path / "output_TLGK.out").read_text()
    assert "nbCreatedCopies" in out and "Total # copies after cloning/pruning" in out
    if alpha == 0.0:
        assert np.all(R == 1.0)
    # the ancestral drag trajectory of every final slot, reconstructed from
    # the driver's files, follows the reference loop's genealogy
    nt = labels.shape[0]
    r = subprocess.run([RECON, str(run) + "/", str(Nc), str(nt), "20",
                        str(tmp_path / "traj") + "/", "all"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for j in range(Nc):
        want, k = [None] * nt, j
        for t in range(nt - 1, -1, -1):
            want[t] = forces_ref[k][t]
            if t > 0:
                k = labels_ref[t - 1][k]
        got = np.fromfile(tmp_path / "traj" / f"clone_{j}" /
                          "data_force.datout").reshape(nt, 20)
        np.testing.assert_allclose(got, np.array(want), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("Nc,p", [(8, 2), (9, 3)])
def test_driver_reproduces_a_multi_rank_run(tmp_path, O, Nc, p):
    """ranks = p: the initial-state assignment, the per-rank prefactor
    streams, the master's draws and the rank-blocked total_R of a p-rank
    run of the reference driver."""
    L, T, dT, eps, tau, U0, alpha = 4, 1.0, 0.25, 0.02, 0.53, 0.05, 40.0
    g, pops, inp, _ = write_case(tmp_path, O, L, Nc, T, dT, eps, tau, U0,
                                 alpha, extra=f"ranks = {p}\ndump_states = 1\n")
    print(f"case: driver REF L={L} Nc={Nc} ranks={p}", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    labels_ref, R_ref, forces_ref, states_ref = reference_loop(
        O, g, pops, Nc, T, dT, eps, U0, alpha, L, p=p)
    run = tmp_path / "run"
    labels = np.fromfile(run / "labels.dat", dtype=np.int32).reshape(-1, Nc)
    assert np.array_equal(labels, np.array(labels_ref))
    # a slot of rank >= 1 was overwritten before the last interval, so its
    # perturbation drew from that rank's own stream
    assert any(lab[j] != j for lab in labels_ref[:-1]
               for j in range(Nc // p, Nc))
    np.testing.assert_allclose(np.fromfile(run / "rvalues.dat"), R_ref,
                               rtol=1e-12, atol=0)
    for j in range(Nc):
        Fj = np.fromfile(run / f"clone_{j}" / "data_force.datout").reshape(-1, 20)
        assert np.array_equal(Fj[0], forces_ref[j][0])
        fin = np.fromfile(run / f"clone_{j}" / "pops_final.datout").reshape(
            g.Lx, g.Ly, 9)
        np.testing.assert_allclose(fin, states_ref[j], rtol=1e-11, atol=0)


REF_RUN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "ref_driver_L4.npz")


@pytest.mark.parametrize("i", range(5))
def test_driver_reproduces_the_reference_driver_run(tmp_path, i):
    """The reference's own TLGK_LBM_POPS_PERTURB, run under mpiexec with
    p = 1, 2, 3 ranks (fixture ref_driver_L4.npz): the driver with
    `ranks = p` on the same input and pops files writes the same labels.dat
    and the same copy tables, rvalues.dat within 1e-12, and every slot's
    first-interval forces bit for bit (later intervals start from perturbed
    copies, whose norms the device sums in another order: 1e-10)."""
    z = np.load(REF_RUN)
    Nc, p, T, dT, eps, L, tau, U0, alpha = z[f"c{i}_params"]
    Nc, p, L = int(Nc), int(p), int(L)
    _, _, inp = refloop.write_case(tmp_path, z["pops"], Nc, T, dT, eps, L,
                                   tau, U0, alpha, extra=f"ranks = {p}\n")
    print(f"case: reference run {i}: Nc={Nc} ranks={p} alpha={alpha}",
          flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    run = tmp_path / "run"
    labels = np.fromfile(run / "labels.dat", dtype=np.int32).reshape(-1, Nc)
    assert np.array_equal(labels, z[f"c{i}_labels"])
    np.testing.assert_allclose(np.fromfile(run / "rvalues.dat"),
                               z[f"c{i}_rvalues"], rtol=1e-12, atol=0)
    nb, pairs = parse_output((tmp_path / "output_TLGK.out").read_text(), Nc)
    assert np.array_equal(np.array(nb), z[f"c{i}_nb_after"])
    want = np.split(z[f"c{i}_pairs"], np.cumsum(z[f"c{i}_npairs"])[:-1])
    assert len(pairs) == len(want)
    for got, ref in zip(pairs, want):
        assert np.array_equal(got, ref)
    F = z[f"c{i}_forces"]
    for j in range(Nc):
        Fj = np.fromfile(run / f"clone_{j}" / "data_force.datout").reshape(
            F.shape[1], F.shape[2])
        assert np.array_equal(Fj[0], F[j][0]), f"slot {j}"
        np.testing.assert_allclose(Fj, F[j], rtol=1e-10, atol=1e-12)


def _channel_input(tmp, Nc, extra):
    inp = tmp / "input"
    # positional fields: Nc T dT eps L tau U0 alpha path master facquVTK
    inp.write_text(f"{Nc}\n1\n1\n0\n4\n0.55\n0.05\n1.0\n{tmp}/\n"
                   f"{tmp}/run/\n5\nmode = channel\nLx = 80\nLy = 40\n"
                   "side = 6\ncx = 20\ncy = 20\nu_in = 0.05\nk = 1.0\n"
                   "sigma = 1e-4\nsteps = 12\nintervals = 3\nseed = 3\n"
                   + extra)
    return inp


@pytest.mark.parametrize("force_io,storage,extra",
                         [(1, "f64", ""), (0, "f64", ""),
                          (1, "f16", "passes = 3\nthreads = 256\n"),
                          (0, "bf16", "")],
                         ids=["loop-f64", "device-f64", "loop-f16-pinned",
                              "device-bf16"])
def test_driver_channel_mode_matches_the_library(tmp_path, lbc, force_io,
                                                 storage, extra):
    """mode = channel: the driver's interval loop (force_io = 1) and its
    all-device path (force_io = 0: one lbc_advance of every interval), in
    FP64 and in shifted storage, with a tuned or a pinned kernel shape, give
    the same labels, mean weights and per-step drag as the same calls made
    through the library directly, bit for bit."""
    Nc = 12
    inp = _channel_input(tmp_path, Nc, f"force_io = {force_io}\nvtk = 0\n"
                                       f"dump_states = 1\nstorage = {storage}\n"
                                       + extra)
    code = {"f64": lbc.STORE_F64, "f16": lbc.STORE_F16,
            "bf16": lbc.STORE_BF16}[storage]
    print(f"case: channel driver force_io={force_io} storage={storage} "
          f"{extra!r}", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    run = tmp_path / "run"
    spec = lbc.channel(80, 40, 6, 20, 20, 0.55, 0.05, Nc, 12, k=1.0,
                       sigma=1e-4)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, storage=code, seed=3) as s:
        s.init_equilibrium(0, 1e-3)
        if force_io:
            labels, drag = [], []
            for _ in range(3):
                s.run_interval()
                drag.append(s.drag(0))
                labels.append(s.select(0).labels)
                s.gather()
        else:
            s.advance(3)
        R = s.R_history(0)
        final = s.store_states(0)
    assert np.array_equal(np.fromfile(run / "rvalues.dat"), R)
    if force_io:
        got = np.fromfile(run / "labels.dat", dtype=np.int32).reshape(3, Nc)
        assert np.array_equal(got, np.array(labels))
        for j in range(Nc):
            Fj = np.fromfile(run / f"clone_{j}" / "data_force.datout")
            assert np.array_equal(Fj.reshape(3, 12),
                                  np.array([d[j] for d in drag]))
    for j in range(Nc):
        fin = np.fromfile(run / f"clone_{j}" / "pops_final.datout").reshape(
            80, 40, 9)
        assert np.array_equal(fin, final[j])


def test_driver_batches_extra_channel_instances(tmp_path, lbc):
    """`instance = ...` lines add samplers to the driver's launches; the
    first sampler's results are those of the same batched context built
    through the library."""
    Nc = 8
    inp = _channel_input(tmp_path, Nc, "force_io = 0\ndump_states = 1\n"
                                       "instance = 60 30 6 15 15 0.55 0.05 "
                                       "16 12 0.5 1e-4\n")
    print("case: channel driver with an extra instance", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    specs = [lbc.channel(80, 40, 6, 20, 20, 0.55, 0.05, Nc, 12, k=1.0,
                         sigma=1e-4),
             lbc.channel(60, 30, 6, 15, 15, 0.55, 0.05, 16, 12, k=0.5,
                         sigma=1e-4)]
    with lbc.Sampler(specs, mode=lbc.MODE_CHANNEL, seed=3) as s:
        s.init_equilibrium(0, 1e-3)
        s.init_equilibrium(1, 1e-3)
        s.advance(3)
        R = s.R_history(0)
        final = s.store_states(0)
    assert np.array_equal(np.fromfile(tmp_path / "run" / "rvalues.dat"), R)
    for j in range(Nc):
        fin = np.fromfile(tmp_path / "run" / f"clone_{j}" /
                          "pops_final.datout").reshape(80, 40, 9)
        assert np.array_equal(fin, final[j])


WARMUP = os.path.join(PKG, "driver", "lbm_warmup")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_warmup_matches_reference_program(tmp_path):
    """lbm_warmup against every output of the reference's run_bodyforce
    (src/lbm_src/main_bodyforce.cpp) on the same input file: force and
    velocity-probe series and final populations bit for bit, parameters
    and VTK frames byte for byte."""
    ref = np.load(os.path.join(GOLDEN, "bodyforce_L4.npz"))
    (tmp_path / "input.datin").write_text(str(ref["input"]))
    print("case: warm-up L=4, 100 steps from rest", flush=True)
    r = subprocess.run([WARMUP], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    out = tmp_path / "out"
    assert (out / "parameters.datout").read_text() == str(ref["parameters"])
    assert np.array_equal(np.fromfile(out / "data_force.datout"), ref["force"])
    assert np.array_equal(np.fromfile(out / "ux_t.datout"), ref["ux"])
    assert np.array_equal(np.fromfile(out / "pops.datout"), ref["pops"])
    for name, text in zip(ref["vtk_names"], ref["vtk"]):
        assert (out / str(name)).read_text() == str(text), name


def test_warmup_snapshot_set_feeds_the_cloning_driver(tmp_path):
    """The warm-up's snapshot set (pops_<i>.datout + popfiles_list.dat) is
    the cloning driver's initial-state input: the last snapshot is the final
    state, and a driver run starting from the set completes."""
    (tmp_path / "input.datin").write_text(
        "1\n120\n4\n0.53\n0.1\nwarm/\n0\n1000\n1000\n1000\n"
        "snap_every = 5\nsnap_count = 10\n")
    print("case: warm-up snapshot set L=4", flush=True)
    r = subprocess.run([WARMUP], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    warm = tmp_path / "warm"
    names = (warm / "popfiles_list.dat").read_text().split()
    assert names == [f"pops_{i}.datout" for i in range(10)]
    assert np.array_equal(np.fromfile(warm / names[-1]),
                          np.fromfile(warm / "pops.datout"))
    assert not np.array_equal(np.fromfile(warm / names[0]),
                              np.fromfile(warm / names[-1]))
    inp = tmp_path / "input_tlgk"
    inp.write_text(f"4\n0.5\n0.25\n0.01\n4\n0.53\n0.05\n5.0\n{warm}/\n"
                   f"{tmp_path}/clones/\n1000\n")
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert np.fromfile(tmp_path / "clones" / "labels.dat",
                       dtype=np.int32).size == 4 * 2


def test_forcing_driver_matches_reference_forcing_loop(tmp_path, O):
    """`order = forcing`: the forcing driver's loop
    (src/mpi_LBM_GK_forcing.cpp:297-601) restated with the oracle on the
    fields the driver dumps: per-interval force files, weights and copies
    files, labels and R values.  The fields themselves match the oracle's
    generator on the same Philox normals."""
    L, Nc, T, dT, dT0, eps, tau, U0, alpha = 4, 6, 1.0, 0.25, 0.1, 0.5, 0.53, \
        0.05, 20.0
    g, pops, _, data = write_case(tmp_path, O, L, Nc, T, dT, eps, tau, U0,
                                  alpha)
    inp = tmp_path / "input_forcing"
    inp.write_text(f"{Nc}\n{T}\n{dT}\n{dT0}\n{eps}\n{L}\n{tau}\n{U0}\n{alpha}\n"
                   f"{data}/\n{tmp_path}/frun/\n"
                   "order = forcing\ndump_fields = 1\n")
    print(f"case: forcing driver L={L} Nc={Nc}", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    run = tmp_path / "frun"
    T0 = L / U0
    M1 = math.floor(dT0 * T0)
    M = M1 + math.floor((dT - dT0) * T0)
    nT = math.floor(T / dT)
    F0 = U0 * U0 * L * 0.5
    fields = [np.fromfile(run / f"fields_{t}.datout").reshape(Nc, g.Lx, g.Ly)
              for t in range(nT)]
    lam = O.field_eigenvalues(g.Lx, g.Ly, 32.0)
    for t in (0, nT - 1):
        for j in range(Nc):
            want = O.field_of_clone(lam, 0, j, t, 0, g.Lx, g.Ly)
            assert np.abs(fields[t][j] - want).max() < 1e-5, (t, j)
    # the reference loop on the oracle with those fields: rank 0's rand()
    # after srand(1) and 1000 discards, drand48 from its never-seeded state
    states = [pops[j].copy() for j in range(Nc)]
    lib = O.lib()
    lib.orc_glibc_srand(1)
    tmpi = np.zeros(1000, np.int32)
    lib.orc_glibc_rand_n(1000, O._ip(tmpi))
    lib.orc_glibc_seed48_default()
    rq = deque()
    labels_ref, R_ref = [], []
    for t in range(nT):
        acc = np.zeros(Nc)
        for j in range(Nc):
            acc[j], A = O.run_steps_forced(g, states[j], M1, M, fields[t][j],
                                           eps, 1.0 / F0, 0.0, 1.0 / T0)
            got = np.fromfile(run / f"clone_{j}" / f"evolution_{t}_clone_{j}")
            assert np.array_equal(got, A), (t, j)
        s = np.array([math.exp(alpha * a) for a in acc])
        assert np.array_equal(np.fromfile(run / f"weights_evolution_{t}"), s)
        u = np.zeros(Nc)
        lib.orc_glibc_drand48_n(Nc, O._dp(u))
        while len(rq) < Nc:
            one = np.zeros(1, np.int32)
            lib.orc_glibc_rand_n(1, O._ip(one))
            rq.append(int(one[0]))
        sel = O.select(s, u, np.array(list(rq)[:Nc], np.int32))
        for _ in range(sel["n_rint"]):
            rq.popleft()
        copies = np.fromfile(run / f"copies_evolution_{t}", dtype=np.int32)
        assert np.array_equal(copies, np.concatenate([sel["nb_before"],
                                                      sel["nb_after"]]))
        labels_ref.append(sel["labels"])
        R_ref.append(sel["total_R"])
        for src, dst in sel["pairs"]:
            states[dst] = states[src].copy()
    labels = np.fromfile(run / "labels.dat", dtype=np.int32).reshape(-1, Nc)
    assert np.array_equal(labels, np.array(labels_ref))
    np.testing.assert_allclose(np.fromfile(run / "rvalues.dat"), R_ref,
                               rtol=1e-12, atol=0)
    # the run did clone and prune
    assert any(not np.array_equal(lb, np.arange(Nc)) for lb in labels_ref)


def _vtk_text(rho_ux_uy, Dx, Dy):
    """The reference's VTK frame (src/lbm_src/write_vtk.cpp:8-53) as text;
    C++ stream defaults print a double like printf's %g."""
    rho, ux, uy = rho_ux_uy
    lines = ["# vtk DataFile Version 3.0", "fluid_state", "ASCII",
             "DATASET RECTILINEAR_GRID", f"DIMENSIONS {Dx} {Dy} 1",
             f"X_COORDINATES {Dx} float",
             "".join(f"{i} " for i in range(Dx)),
             f"Y_COORDINATES {Dy} float",
             "".join(f"{j} " for j in range(Dy)),
             "Z_COORDINATES 1 float", "0", f"POINT_DATA {Dx * Dy}",
             "SCALARS density float 1", "LOOKUP_TABLE default"]
    lines += [f"{rho[X, Y] - 1.0:g}" for Y in range(Dy) for X in range(Dx)]
    lines.append("VECTORS velocity_vector float")
    lines += [f"{ux[X, Y]:g} {uy[X, Y]:g} 0" for Y in range(Dy)
              for X in range(Dx)]
    return "\n".join(lines) + "\n"


def test_driver_vtk_frames_match_reference_loop(tmp_path, O):
    """facquVTK = 5 with 20 steps per interval: every clone writes frames
    fluid_t<4t + k>.vtk of steps 0, 5, 10, 15 (src/mpi_LBM_GK_pops_perturb.cpp:
    404-455).  The first interval's frames equal the reference's text for the
    oracle's states byte for byte."""
    L, Nc, T, dT, eps, tau, U0, alpha = 4, 3, 0.5, 0.25, 0.02, 0.53, 0.05, 5.0
    g, pops, _, data = write_case(tmp_path, O, L, Nc, T, dT, eps, tau, U0,
                                  alpha)
    inp = tmp_path / "input_vtk"
    inp.write_text(f"{Nc}\n{T}\n{dT}\n{eps}\n{L}\n{tau}\n{U0}\n{alpha}\n"
                   f"{data}/\n{tmp_path}/vrun/\n5\n")
    print(f"case: driver VTK frames L={L} Nc={Nc}", flush=True)
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    run = tmp_path / "vrun"
    for j in range(Nc):
        a = pops[j].copy()
        b = np.zeros_like(a)
        for n in range(20):
            if n % 5 == 0:
                text = (run / f"clone_{j}" / f"fluid_t{n // 5}.vtk").read_text()
                assert text == _vtk_text(O.macro(g, a), g.Lx, g.Ly), (j, n)
            O.step(g, a, b)
            a, b = b, a
        # the second interval's frames follow as 4..7
        for k in range(4, 8):
            assert (run / f"clone_{j}" / f"fluid_t{k}.vtk").exists()
        assert not (run / f"clone_{j}" / "fluid_t8.vtk").exists()
