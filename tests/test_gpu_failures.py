"""Failure handling on the device, forced from the host so that every kernel
still runs to completion: non-finite populations, weights that would
overflow exp(), degenerate selections, and driver inputs that are missing
or short (the reference only prints such errors and carries on,
src/mpi_LBM_GK_pops_perturb.cpp:66-111,321-384)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG
from refloop import make_pops, write_case

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG, "driver", "tlgk_lbm_gpu")


def _channel(lbc, k=0.01, n=4, steps=3):
    return lbc.channel(48, 24, 6, 12, 12, 0.55, 0.05, n, steps, k=k,
                       sigma=1e-4)


def test_nan_population_fails_selection_and_keeps_every_clone(lbc):
    spec = _channel(lbc)
    print("case: NaN population in clone 2, CHANNEL F64 48x24", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=5, threads=256) as s:
        s.init_equilibrium(0, 1e-3)
        st = s.store_states(0)
        st[2, 15, 12, :] = np.nan  # a node of the square's east face
        s.load_states(0, 0, st)
        s.run_interval()
        obs, logw = s.logweights(0)
        assert np.isnan(logw[2]) and np.all(np.isfinite(np.delete(logw, 2)))
        before = s.store_states(0)
        with pytest.raises(lbc.LbcError, match="non-finite"):
            s.select(0)
        # the failed selection keeps every clone: the copy moves nothing
        s.gather()
        after = s.store_states(0)
        marks = s.get_marks(0)
    ok = np.ones(4, bool)
    ok[2] = False
    assert np.array_equal(after[ok], before[ok])
    assert np.isnan(after[2]).any()
    assert not marks.any()


def test_weights_beyond_exp_range_select_in_the_log_domain(lbc):
    """A bias that puts log-weights beyond 800 (exp overflows at 709): the
    selection shifts by the largest log-weight, so the mean weight is
    reported through its logarithm and the copies are still made."""
    spec = _channel(lbc, k=2e6, n=64, steps=4)
    print("case: bias 2e6, 64 clones", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=7, threads=256) as s:
        s.init_equilibrium(0, 1e-3)
        s.run_interval()
        _, logw = s.logweights(0)
        assert np.abs(logw).max() > 800 and np.all(np.isfinite(logw))
        sel = s.select(0)
        s.gather()
    m = logw.max()
    log_mean = m + np.log(np.mean(np.exp(logw - m)))
    assert np.isfinite(sel.log_total_R)
    assert abs(sel.log_total_R - log_mean) < 1e-9 * abs(log_mean)
    assert sel.nb_after.sum() == 64 and len(sel.pairs) > 0
    # the heaviest clone is never pruned
    assert sel.nb_after[np.argmax(logw)] >= 1


def test_equal_weights_give_identity_and_no_copies(lbc):
    spec = _channel(lbc, n=16)
    print("case: equal weights, 16 clones", flush=True)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=9, threads=256) as s:
        s.init_equilibrium(0, 1e-3)
        s.run_interval()
        before = s.store_states(0)
        sel = s.select(0, weights=np.ones(16), u=np.full(16, 0.5),
                       rint=np.arange(16, dtype=np.int32))
        s.gather()
        after = s.store_states(0)
        marks = s.get_marks(0)
    assert np.array_equal(sel.labels, np.arange(16))
    assert len(sel.pairs) == 0 and sel.total_R == 1.0
    assert np.array_equal(after, before) and not marks.any()


def test_driver_reports_missing_and_short_population_files(tmp_path, O):
    print("case: driver with missing index file, then a short pops file",
          flush=True)
    pops = make_pops(O, 4, 10, 404)
    data, names, inp = write_case(tmp_path, pops, 4, 0.6, 0.3, 0.1, 4, 0.53,
                                  0.05, 1.0)
    os.rename(data / "popfiles_list.dat", data / "moved.dat")
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    log = (tmp_path / "output_TLGK.out").read_text()
    assert r.returncode == 1 and "COULD NOT OPEN INDEX FILE" in log
    os.rename(data / "moved.dat", data / "popfiles_list.dat")
    raw = (data / names[1]).read_bytes()
    (data / names[1]).write_bytes(raw[: len(raw) // 2])
    r = subprocess.run([DRIVER, str(inp)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    log = (tmp_path / "output_TLGK.out").read_text()
    assert r.returncode == 1 and "could not read" in log
