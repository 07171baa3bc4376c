#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ by running the reference's
own kernel code, compiled from /root/reference by `make -C oracle ref` into
oracle/_ref/libref.so, and glibc's generators that the reference calls.

Fixtures (numpy .npz, FP64 unless noted):
  ref_steps_L4.npz, ref_steps_L8.npz
      reference-geometry lattice (square side L, Dx = Dy = 4L+1), seeded
      non-equilibrium initial state S0, the state and post-collision array
      after steps 1, 2 and the last (indices in `keep`), and the raw
      stress-integral force after every step
      (src/lbm_src/streamCollCompute.cpp, domain_noSlipWalls.cpp, square.cpp,
      force.cpp and the interior reset of src/mpi_LBM_GK_pops_perturb.cpp).
  ref_perturb_L4.npz
      generatePerturbedState (src/perturbation_routines/
      generatePerturbedState.cpp) on a seeded state with 3 snapshots and
      drand48 in its never-seeded state (X = 0); the prefactors it drew are
      recorded.
  glibc_streams.npz
      rand() after srand(1) and 1000 discards (the master rank's stream,
      src/mpi_LBM_GK_pops_perturb.cpp:176-178), and drand48() in a fresh
      process that never seeds it (the reference never does).
  bodyforce_L4.npz
      every output of the reference's warm-up program (src/lbm_src/
      main_bodyforce.cpp, built by `make -C oracle ref` into
      oracle/_ref/run_bodyforce) for square side 4 from rest, 100 steps:
      the input file, parameters.datout, data_force.datout, ux_t.datout,
      pops.datout and the VTK frames fluid_t<i>.vtk (text).
  ref_forced_L4.npz
      a forcing-mode interval (src/mpi_LBM_GK_forcing.cpp:335-395) on the
      square-side-4 lattice: seeded state S0 and random map (the reference's
      map[idx(x,y)] as an (Lx, Ly) array), M1 steps of
      streamingAndCollisionComputeMacroBodyForceSpatial
      (src/lbm_src/streamCollCompute.cpp:50-93) then M - M1 uniform steps,
      each with walls, square and interior reset; the post-collision array of
      the first step, the final state and the raw force after every step.

  ref_macro_L4.npz
      the macroscopic fields rho, ux, uy (each (Lx, Ly), idx(x,y) layout)
      that the reference's step kernels write for a seeded state: uniform
      (streamingAndCollisionComputeMacroBodyForce) and spatial with a map
      (...Spatial), src/lbm_src/streamCollCompute.cpp:17-26,44-45.

  ref_driver_L4.npz
      the reference's production cloning driver (src/mpi_LBM_GK_pops_perturb
      .cpp, built by `make -C oracle ref` into oracle/_ref/tlgk_pops_perturb
      with the macros FORCE_IO, ZERO_MEAN, _SMART_PERTURB) run under
      `mpiexec -n p` on square-side-4 lattices: five cases (Nc, p, alpha) =
      (6, 1, 10), (8, 2, 20), (9, 3, 40), (9, 1, 20), (6, 3, 10), each 8
      cloning intervals of 20 steps from the same ten seeded pops files
      (`pops`, also the ten perturbation snapshots).  Per case `c<i>_`:
      params [Nc, p, T, dT, eps, L, tau, U0, alpha], labels.dat, rvalues.dat,
      every slot's data_force.datout, and from output_TLGK.out the
      nbCreatedCopies after the prune/add and the copy pairs (flat, with
      npairs per interval).  The alphas were chosen so that every case has
      an interval that prunes and one that adds two or more clones.

Run from the repository root: python tests/golden/make_golden.py [name ...]
(names: steps perturb glibc bodyforce forced macro driver; default all).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def seeded_state(g, rng, amp=2e-3):
    f = np.empty((g.Lx, g.Ly, 9))
    for x in range(g.Lx):
        for y in range(g.Ly):
            rho = 1 + amp * rng.standard_normal()
            ux = 0.05 + amp * rng.standard_normal()
            uy = amp * rng.standard_normal()
            u2 = -1.5 * (ux * ux + uy * uy)
            for k in range(9):
                eu = O.CX[k] * ux + O.CY[k] * uy
                f[x, y, k] = O.W[k] * rho * (1 + 3 * eu + 4.5 * eu * eu + u2)
    f += 1e-5 * rng.standard_normal(f.shape)
    f[g.xmin + 1:g.xmax, g.ymin + 1:g.ymax, :] = O.W
    return f


def steps_fixture(L, nsteps, seed, tau, U0):
    g = O.ref_box_geom(L, tau, U0)
    rng = np.random.default_rng(seed)
    s0 = seeded_state(g, rng)
    a, b = s0.copy(), np.zeros_like(s0)
    states, posts, forces = [], [], []
    keep = [0, 1, nsteps - 1]  # steps whose arrays are stored
    for _ in range(nsteps):
        F = O.ref_step(g, a, b)
        posts.append(a.copy())
        a, b = b, a
        states.append(a.copy())
        forces.append(F)
    np.savez_compressed(
        os.path.join(OUT, f"ref_steps_L{L}.npz"), L=L, tau=tau, U0=U0,
        geom=np.array([g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax]),
        beta=g.beta, S0=s0, keep=np.array(keep),
        states=np.array([states[i] for i in keep]),
        post=np.array([posts[i] for i in keep]), forces=np.array(forces))


def perturb_fixture():
    L = 4
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(99)
    fin = seeded_state(g, rng)
    snaps = np.stack([seeded_state(g, rng, amp=1e-2) for _ in range(3)])
    out = fin.copy()
    a = O.ref_perturb_default_stream(out, snaps, 0.25)
    np.savez_compressed(os.path.join(OUT, "ref_perturb_L4.npz"), fin=fin,
                        snaps=snaps, eps=0.25, prefactors=a, out=out)


NEVER_SEEDED = ("import ctypes; c = ctypes.CDLL('libc.so.6'); "
                "c.drand48.restype = ctypes.c_double; "
                "print(' '.join(repr(c.drand48()) for _ in range(64)))")


def glibc_fixture():
    r = O.glibc_rand(1, 64, discard=1000)
    # drand48 in a fresh process that never seeds it, as the reference's
    # ranks never do
    d = np.array([float(v) for v in subprocess.run(
        [sys.executable, "-c", NEVER_SEEDED], check=True, capture_output=True,
        text=True).stdout.split()])
    np.savez_compressed(os.path.join(OUT, "glibc_streams.npz"),
                        rand_seed1_after1000=r, drand48_default=d)


BODYFORCE_INPUT = "1\n100\n4\n0.53\n0.1\nout/\n0\n50\n1\n1\n"


def bodyforce_fixture():
    exe = os.path.join(ROOT, "oracle", "_ref", "run_bodyforce")
    run = os.path.join(ROOT, "oracle", "_ref", "bodyforce_run")
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(run)
    with open(os.path.join(run, "input.datin"), "w") as f:
        f.write(BODYFORCE_INPUT)
    subprocess.run([exe], cwd=run, check=True, stdout=subprocess.DEVNULL)
    out = os.path.join(run, "out")
    vtk = sorted(n for n in os.listdir(out) if n.startswith("fluid_t"))
    text = lambda n: open(os.path.join(out, n)).read()
    np.savez_compressed(
        os.path.join(OUT, "bodyforce_L4.npz"), input=BODYFORCE_INPUT,
        parameters=text("parameters.datout"),
        force=np.fromfile(os.path.join(out, "data_force.datout")),
        ux=np.fromfile(os.path.join(out, "ux_t.datout")),
        pops=np.fromfile(os.path.join(out, "pops.datout")),
        vtk_names=np.array(vtk), vtk=np.array([text(n) for n in vtk]))


def forced_fixture(L=4, M1=5, M=9, eps=0.4, seed=21):
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(seed)
    s0 = seeded_state(g, rng)
    fmap = rng.standard_normal((g.Lx, g.Ly))
    a, b = s0.copy(), np.zeros_like(s0)
    forces, post0 = [], None
    for n in range(M):
        F = (O.ref_step_spatial(g, a, b, fmap, eps) if n < M1
             else O.ref_step(g, a, b))
        if n == 0:
            post0 = a.copy()
        a, b = b, a
        forces.append(F)
    np.savez_compressed(
        os.path.join(OUT, f"ref_forced_L{L}.npz"), L=L, tau=g.tau,
        geom=np.array([g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax]),
        beta=g.beta, S0=s0, map=fmap, eps=eps, M1=M1, M=M, post0=post0,
        final=a, forces=np.array(forces))


def macro_fixture(L=4, eps=0.7, seed=31):
    g = O.ref_box_geom(L, 0.53, 0.05)
    rng = np.random.default_rng(seed)
    s0 = seeded_state(g, rng)
    fmap = rng.standard_normal((g.Lx, g.Ly))
    uni = O.ref_step_fields(g, s0.copy(), np.zeros_like(s0))
    spa = O.ref_step_fields(g, s0.copy(), np.zeros_like(s0), fmap, eps)
    np.savez_compressed(
        os.path.join(OUT, f"ref_macro_L{L}.npz"), L=L, tau=g.tau,
        geom=np.array([g.Lx, g.Ly, g.xmin, g.xmax, g.ymin, g.ymax]),
        beta=g.beta, S0=s0, map=fmap, eps=eps, uniform=uni, spatial=spa)


DRIVER_CASES = [(6, 1, 10.0), (8, 2, 20.0), (9, 3, 40.0), (9, 1, 20.0),
                (6, 3, 10.0)]
DRIVER_RUN = dict(T=2.0, dT=0.25, eps=0.02, L=4, tau=0.53, U0=0.05)


def driver_fixture():
    import pathlib

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refloop import make_pops, parse_output, write_case

    exe = os.path.join(ROOT, "oracle", "_ref", "tlgk_pops_perturb")
    mpiexec = os.path.join(os.environ.get("MPI_HOME", "/opt/conda"), "bin",
                           "mpiexec")
    if not os.path.exists(exe):
        sys.exit("oracle/_ref/tlgk_pops_perturb missing: run `make -C oracle "
                 "ref` where /root/reference and MPI exist")
    c = DRIVER_RUN
    pops = make_pops(O, c["L"], 10, 2024, c["tau"], c["U0"])
    M = int(c["dT"] * c["L"] / c["U0"])
    nT = int(c["T"] / c["dT"])
    out = dict(pops=pops)
    for i, (Nc, p, alpha) in enumerate(DRIVER_CASES):
        tmpdir = tempfile.TemporaryDirectory()
        run = pathlib.Path(tmpdir.name)
        _, _, inp = write_case(run, pops, Nc, c["T"], c["dT"], c["eps"],
                               c["L"], c["tau"], c["U0"], alpha)
        subprocess.run([mpiexec, "-n", str(p), exe, str(inp)], cwd=run,
                       check=True, stdout=subprocess.DEVNULL, timeout=600)
        res = run / "run"
        nb, pairs = parse_output((run / "output_TLGK.out").read_text(), Nc)
        assert len(nb) == nT and len(pairs) == nT
        out[f"c{i}_params"] = np.array([Nc, p, c["T"], c["dT"], c["eps"],
                                        c["L"], c["tau"], c["U0"], alpha])
        out[f"c{i}_labels"] = np.fromfile(res / "labels.dat",
                                          np.int32).reshape(nT, Nc)
        out[f"c{i}_rvalues"] = np.fromfile(res / "rvalues.dat")
        out[f"c{i}_forces"] = np.stack([
            np.fromfile(res / f"clone_{j}" / "data_force.datout")
            for j in range(Nc)]).reshape(Nc, nT, M)
        out[f"c{i}_nb_after"] = np.array(nb)
        out[f"c{i}_npairs"] = np.array([len(q) for q in pairs], np.int32)
        out[f"c{i}_pairs"] = np.concatenate(pairs).astype(np.int32)
        tmpdir.cleanup()
    np.savez_compressed(os.path.join(OUT, "ref_driver_L4.npz"), **out)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` "
                 "where /root/reference exists")
    jobs = {
        "steps": lambda: (steps_fixture(4, 12, 4, 0.53, 0.05),
                          steps_fixture(8, 6, 8, 0.515, 0.05)),
        "perturb": perturb_fixture,
        "glibc": glibc_fixture,
        "bodyforce": bodyforce_fixture,
        "forced": forced_fixture,
        "macro": macro_fixture,
        "driver": driver_fixture,
    }
    for name in sys.argv[1:] or list(jobs):
        jobs[name]()
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
