"""The selection and the driver loop against the reference's own cloning
driver (src/mpi_LBM_GK_pops_perturb.cpp), compiled from its sources and run
under mpiexec with 1, 2 and 3 ranks: fixture tests/golden/ref_driver_L4.npz
(written by tests/golden/make_golden.py).  CPU only: the oracle's restated
loop (tests/refloop.py) must reproduce every output file of the reference
run, and orc_select its prune and add steps on hand-worked draws."""
import os

import numpy as np
import pytest

from refloop import reference_loop

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "ref_driver_L4.npz")


def _cases():
    z = np.load(GOLDEN)
    n = len([k for k in z.files if k.endswith("_params")])
    return z, n


@pytest.fixture(scope="module")
def fixture():
    return _cases()[0]


@pytest.mark.parametrize("i", range(_cases()[1]))
def test_oracle_loop_reproduces_the_reference_driver(O, fixture, i):
    """Labels, mean weights, every slot's per-step force over all eight
    intervals (perturbed ones included), nbCreatedCopies after the
    prune/add and the copy pairs: all identical to the reference run."""
    z = fixture
    Nc, p, T, dT, eps, L, tau, U0, alpha = z[f"c{i}_params"]
    Nc, p, L = int(Nc), int(p), int(L)
    g = O.ref_box_geom(L, tau, U0)
    r = reference_loop(O, g, z["pops"], Nc, T, dT, eps, U0, alpha, L, p=p)
    print(f"case {i}: Nc={Nc} ranks={p} alpha={alpha} deltaN={r['deltaN']}")
    # the case exercises the prune branch and an add of two or more clones
    assert max(r["deltaN"]) > 0 and min(r["deltaN"]) <= -2
    assert np.array_equal(np.array(r["labels"]), z[f"c{i}_labels"])
    assert np.array_equal(np.array(r["R"]), z[f"c{i}_rvalues"])
    assert np.array_equal(np.array(r["nb_after"]), z[f"c{i}_nb_after"])
    pairs = np.split(z[f"c{i}_pairs"], np.cumsum(z[f"c{i}_npairs"])[:-1])
    for t, q in enumerate(pairs):
        assert np.array_equal(r["pairs"][t], q), f"interval {t}"
    for j in range(Nc):
        assert np.array_equal(r["forces"][j], z[f"c{i}_forces"][j]), f"slot {j}"


def test_select_add_draws_from_the_original_multiset(O):
    """Hand-worked add step (src/mpi_LBM_GK_pops_perturb.cpp:524-532,
    558-571, 590-619).  Weights [2, 1, 0.6, 0.7, 0.7] sum to exactly 5, so
    total_R = 1 and with zero uniforms nb = [2, 1, 0, 0, 0]: NcPrime = 3,
    temp = [0, 0, 1], deltaN = -2.  Add 0 draws 2 % 3 = 2: clone 1, which
    is also appended as temp[3].  Add 1 draws 3 % 3 = 0: clone 0 (an index
    taken modulo the grown count, 3 % 4 = 3, would pick clone 1 again).
    nb ends [3, 2, 0, 0, 0]; the table fills slots 2, 3 from clone 0 and
    slot 4 from clone 1."""
    s = np.array([2.0, 1.0, 0.6, 0.7, 0.7])
    sel = O.select(s, np.zeros(5), np.array([2, 3, 99, 99, 99], np.int32))
    assert sel["total_R"] == 1.0
    assert list(sel["nb_before"]) == [2, 1, 0, 0, 0]
    assert sel["n_rint"] == 2
    assert list(sel["nb_after"]) == [3, 2, 0, 0, 0]
    assert sel["pairs"].tolist() == [[0, 2], [0, 3], [1, 4]]
    assert list(sel["labels"]) == [0, 1, 0, 0, 1]


def test_select_prune_swaps_the_last_entry_in(O):
    """Hand-worked prune step (src/mpi_LBM_GK_pops_perturb.cpp:540-556).
    Weights [3, 0.5, 0.5, 0.5, 0.5] (total_R = 1) with uniforms
    [0, 0.6, 0.6, 0.6, 0.6] give nb = [3, 1, 1, 1, 1]: NcPrime = 7,
    temp = [0, 0, 0, 1, 2, 3, 4], deltaN = 2.  Prune 0 draws 9 % 7 = 2:
    clone 0 dies once and temp[2] takes temp[6] = 4.  Prune 1 draws
    8 % 6 = 2: clone 4 (without the swap it would be clone 0 again; modulo
    7 it would be index 1, clone 0).  nb ends [2, 1, 1, 1, 0] and clone 0
    refills slot 4."""
    s = np.array([3.0, 0.5, 0.5, 0.5, 0.5])
    u = np.array([0.0, 0.6, 0.6, 0.6, 0.6])
    sel = O.select(s, u, np.array([9, 8, 99, 99, 99], np.int32))
    assert sel["total_R"] == 1.0
    assert list(sel["nb_before"]) == [3, 1, 1, 1, 1]
    assert sel["n_rint"] == 2
    assert list(sel["nb_after"]) == [2, 1, 1, 1, 0]
    assert sel["pairs"].tolist() == [[0, 4]]
    assert list(sel["labels"]) == [0, 1, 2, 3, 0]
