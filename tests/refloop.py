"""The reference driver's interval loop restated on the CPU oracle, and the
input files it reads; shared by the driver tests.

Test infrastructure only: it calls the oracle (oracle/oracle.py) and glibc's
generators as restated there.  Reference: src/mpi_LBM_GK_pops_perturb.cpp
(input :66-84, derived parameters :116-151, RNG seeding :176-178, file
hand-out :251-319, snapshots :321-384, interval loop :387-714)."""
import math
from collections import deque

import numpy as np

AVG = 1.775  # src/mpi_LBM_GK_pops_perturb.cpp:27 (ZERO_MEAN shift)
NN = 10      # snapshots mixed by generatePerturbedState (:43)


def make_pops(O, L, n, seed, tau=0.53, U0=0.05, amp=2e-3):
    """n reference-geometry lattices (square side L), each a seeded
    near-equilibrium state in the pops-file layout [x][y][k], with the
    square's interior at rest."""
    g = O.ref_box_geom(L, tau, U0)
    rng = np.random.default_rng(seed)
    pops = np.empty((n, g.Lx, g.Ly, 9))
    for i in range(n):
        for x in range(g.Lx):
            for y in range(g.Ly):
                pops[i, x, y] = O.feq(1 + amp * rng.standard_normal(),
                                      U0 * (1 + 0.2 * rng.standard_normal()),
                                      amp * rng.standard_normal())
        pops[i, g.xmin + 1:g.xmax, g.ymin + 1:g.ymax] = O.W
    return pops


def write_case(tmp, pops, Nc, T, dT, eps, L, tau, U0, alpha, extra="",
               facq=1000):
    """Pops files, popfiles_list.dat and the positional input file
    (input_TEMPLATE order: Nc T dT eps L tau U0 alpha path master facquVTK)
    under tmp; returns (data folder, file names, input path)."""
    data = tmp / "data"
    data.mkdir()
    names = []
    for i, f in enumerate(pops):
        name = f"pops_{i}.datout"
        f.tofile(data / name)
        names.append(name)
    (data / "popfiles_list.dat").write_text("\n".join(names) + "\n")
    inp = tmp / "input"
    inp.write_text(f"{Nc}\n{T}\n{dT}\n{eps}\n{L}\n{tau}\n{U0}\n{alpha}\n"
                   f"{data}/\n{tmp}/run/\n{facq}\n{extra}")
    return data, names, inp


def reference_loop(O, g, pops, Nc, T, dT, eps, U0, alpha, L, p=1):
    """The reference driver's interval loop with p MPI ranks of Nc/p clones,
    on the oracle: the master hands the first Nc - Nc/p file names to ranks
    1.. and keeps the next Nc/p (:273-292); every rank reads the first NN
    names as snapshots (:321-384) and draws its marked clones' prefactors
    from its own drand48 (never seeded, so each rank's stream starts from
    the default state); the master draws the floor-step uniforms from its
    drand48 after its own prefactors (:517), the prune/add ints from its
    rand() (srand(1), 1000 discarded, :176-178, :545, :563), and sums
    total_R in rank blocks (:489-497).

    Returns per interval the labels, total_R, nbCreatedCopies before and
    after the prune/add, the copy pairs and deltaN; per clone slot its force
    series, and the final states."""
    T0 = L / U0
    F0 = U0 * U0 * L * 0.5
    dt = 1.0 / T0
    M = int(math.floor(dT * T0))
    nT = int(math.floor(T / dT))
    local = Nc // p
    first = [j - local if j >= local else (Nc - local) + j for j in range(Nc)]
    states = [pops[first[j]].copy() for j in range(Nc)]
    snaps = np.ascontiguousarray(pops[:NN])
    lib = O.lib()
    lib.orc_glibc_srand(1)
    tmpi = np.zeros(1000, np.int32)
    lib.orc_glibc_rand_n(1000, O._ip(tmpi))
    lib.orc_glibc_seed48_default()
    default48 = np.zeros(100000)  # drand48's sequence from its default state
    lib.orc_glibc_drand48_n(default48.size, O._dp(default48))
    pos = [0] * p

    def drand(n, rank=0):
        out = default48[pos[rank]:pos[rank] + n]
        pos[rank] += n
        return out

    rq = deque()
    marks = np.zeros(Nc, bool)
    forces = [[] for _ in range(Nc)]
    out = dict(labels=[], R=[], nb_before=[], nb_after=[], pairs=[],
               deltaN=[])
    for t in range(nT):
        for j in range(Nc):
            if marks[j]:
                O.perturb(states[j], snaps, eps, drand(NN, j // local))
        acc = np.zeros(Nc)
        for j in range(Nc):
            acc[j], A = O.run_steps(g, states[j], M, 1.0 / F0, AVG, dt)
            forces[j].append(A)
        s = np.array([math.exp(alpha * a) for a in acc])
        u = drand(Nc)
        while len(rq) < Nc:
            tmp1 = np.zeros(1, np.int32)
            lib.orc_glibc_rand_n(1, O._ip(tmp1))
            rq.append(int(tmp1[0]))
        sel = O.select(s, u, np.array(list(rq)[:Nc], np.int32), p_blocks=p)
        for _ in range(sel["n_rint"]):
            rq.popleft()
        out["labels"].append(sel["labels"])
        out["R"].append(sel["total_R"])
        out["nb_before"].append(sel["nb_before"])
        out["nb_after"].append(sel["nb_after"])
        out["pairs"].append(sel["pairs"])
        out["deltaN"].append(int(sel["nb_before"].sum()) - Nc)
        marks[:] = False
        for src, dst in sel["pairs"]:
            states[dst] = states[src].copy()
            marks[dst] = True
    out["forces"] = [np.array(f) for f in forces]
    out["states"] = states
    return out


def parse_output(text, Nc):
    """nbCreatedCopies after the prune/add and the copy pairs of every
    cloning step, from the master's output_TLGK.out
    (src/mpi_LBM_GK_pops_perturb.cpp:576-631)."""
    nb, pairs = [], []
    lines = text.splitlines()
    for i, line in enumerate(lines):
        if line.strip().startswith("nbCreatedCopies :"):
            fields = line.split(":", 1)[1].split()
            row = np.zeros(Nc, np.int32)
            for f in fields:
                j, n = f.split("|")
                row[int(j)] = int(n)
            nb.append(row)
        if line.strip() == "* Creation of communication table - DONE":
            table = lines[i + 1]
            prs = [tuple(int(v) for v in item.split("->"))
                   for item in table.split("|") if "->" in item]
            pairs.append(np.array(prs, np.int32).reshape(-1, 2))
    return nb, pairs
