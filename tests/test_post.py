"""Post-processing of a cloning run's output files (CPU): the ancestral
trajectory reconstruction of post_processing/reconstruction.cpp."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RECON = os.path.join(ROOT, "cloning-lbm-mpi_amd", "post", "reconstruct")


@pytest.fixture(params=["reconstruct", "reconstruct_san"])
def tool(request):
    """The program as shipped, and the same source built with
    AddressSanitizer + UndefinedBehaviorSanitizer (any report aborts it)."""
    path = os.path.join(os.path.dirname(RECON), request.param)
    if not os.path.exists(path):
        pytest.skip(f"post/{request.param} is not built "
                    "(make -C cloning-lbm-mpi_amd)")
    return path


def fake_run(path, Nc, nt, ndt, seed):
    """A run with a genealogy built forward in time: after every cloning
    step each slot receives a random parent.  Slot k's force record of
    interval t encodes (k, t, step) so a trajectory shows which slot it
    read at every interval.  Returns the parents, parents[t][k] = slot that
    slot k descends from across cloning step t."""
    rng = np.random.default_rng(seed)
    parents = rng.integers(0, Nc, size=(nt, Nc)).astype(np.int32)
    os.makedirs(path, exist_ok=True)
    parents.tofile(os.path.join(path, "labels.dat"))
    for k in range(Nc):
        d = os.path.join(path, f"clone_{k}")
        os.makedirs(d, exist_ok=True)
        rec = np.array([[1000.0 * k + 10.0 * t + s for s in range(ndt)]
                        for t in range(nt)])
        rec.tofile(os.path.join(d, "data_force.datout"))
    return parents


def lineage(parents, j, nt):
    """Slot holding slot j's ancestor during each interval, walking the
    forward-built parents from the last interval back."""
    line = [j]
    for t in range(nt - 1, 0, -1):
        line.append(int(parents[t - 1][line[-1]]))
    return line[::-1]


@pytest.mark.parametrize("Nc,nt,ndt", [(1, 1, 4), (5, 4, 3), (16, 30, 7)])
def test_reconstruct_follows_the_genealogy(tmp_path, tool, Nc, nt, ndt):
    run = str(tmp_path / "run") + "/"
    out = str(tmp_path / "out") + "/"
    parents = fake_run(run, Nc, nt, ndt, seed=Nc * 100 + nt)
    r = subprocess.run([tool, run, str(Nc), str(nt), str(ndt), out, "all"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for j in range(Nc):
        got = np.fromfile(os.path.join(out, f"clone_{j}", "data_force.datout"))
        assert got.size == nt * ndt
        got = got.reshape(nt, ndt)
        line = lineage(parents, j, nt)
        for t in range(nt):
            np.testing.assert_array_equal(
                got[t], 1000.0 * line[t] + 10.0 * t + np.arange(ndt))


def test_reconstruct_identity_genealogy_returns_own_records(tmp_path, tool):
    """With every slot its own parent the trajectory is the slot's file."""
    Nc, nt, ndt = 6, 5, 4
    run = str(tmp_path / "run") + "/"
    fake_run(run, Nc, nt, ndt, seed=1)
    np.tile(np.arange(Nc, dtype=np.int32), nt).tofile(
        os.path.join(run, "labels.dat"))
    out = str(tmp_path / "out") + "/"
    r = subprocess.run([tool, run, str(Nc), str(nt), str(ndt), out, "2", "5"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for j in (2, 5):
        got = np.fromfile(os.path.join(out, f"clone_{j}", "data_force.datout"))
        own = np.fromfile(os.path.join(run, f"clone_{j}", "data_force.datout"))
        np.testing.assert_array_equal(got, own)
    assert not os.path.exists(os.path.join(out, "clone_0"))


def test_reconstruct_default_is_slot_zero_and_rejects_bad_input(tmp_path, tool):
    Nc, nt, ndt = 3, 3, 2
    run = str(tmp_path / "run") + "/"
    fake_run(run, Nc, nt, ndt, seed=2)
    r = subprocess.run([tool, run, str(Nc), str(nt), str(ndt)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(tmp_path / "outputFolder" / "clone_0" /
                          "data_force.datout")
    assert not os.path.exists(tmp_path / "outputFolder" / "clone_1")
    # more intervals than the files hold, a slot out of range, a bad label
    for args in ([run, str(Nc), str(nt + 1), str(ndt)],
                 [run, str(Nc), str(nt), str(ndt), "7"]):
        r = subprocess.run([tool] + args, capture_output=True, text=True,
                           cwd=tmp_path)
        assert r.returncode == 1
    np.full(Nc * nt, 9, np.int32).tofile(os.path.join(run, "labels.dat"))
    r = subprocess.run([tool, run, str(Nc), str(nt), str(ndt)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "outside" in r.stderr
