"""The C-ABI library loads and exports every entry point include/lbc.h
declares; configuration errors are reported without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbc_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree(lbc):
    assert declared_functions() == sorted(lbc.EXPORTS)


def test_library_exports_every_declared_function(lbc):
    L = lbc.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_driver_built_and_linked():
    exe = os.path.join(ROOT, "cloning-lbm-mpi_amd", "driver", "tlgk_lbm_gpu")
    assert os.path.exists(exe) and os.access(exe, os.X_OK)


def _bad(lbc, **kw):
    spec = lbc.channel(80, 40, 8, 20, 20, 0.53, 0.05, 4, 10, k=0.01)
    mode = kw.pop("mode", lbc.MODE_CHANNEL)
    storage = kw.pop("storage", lbc.STORE_F64)
    threads = kw.pop("threads", 0)
    vec = kw.pop("vec", 0)
    shape = {k: kw.pop(k) for k in ("passes", "cpw") if k in kw}
    for k, v in kw.items():
        setattr(spec, k, v)
    with pytest.raises(lbc.LbcError) as e:
        lbc.Sampler(spec, mode=mode, storage=storage, threads=threads, vec=vec,
                    **shape)
    assert "(-1)" in str(e.value)
    return str(e.value)


def test_create_rejects_invalid_configurations(lbc):
    assert "tau" in _bad(lbc, tau=0.5)
    assert "square" in _bad(lbc, xmin=1)
    assert "square" in _bad(lbc, ymax=38)
    assert "n_clones" in _bad(lbc, n_clones=0)
    assert "n_clones" in _bad(lbc, n_clones=8193)
    assert "steps" in _bad(lbc, steps=0)
    assert "rank_blocks" in _bad(lbc, rank_blocks=3)
    assert "lattice" in _bad(lbc, Lx=6)
    # one FP64 clone of 6000x6000 is 2.6 GB: beyond 32-bit byte offsets;
    # the same lattice fits in FP16 storage, so only the bytes are refused
    assert "2 GiB" in _bad(lbc, Lx=6000, Ly=6000)
    assert "mode" in _bad(lbc, mode=7)
    assert "storage" in _bad(lbc, storage=9)
    assert "threads" in _bad(lbc, threads=100)
    assert "sites per thread" in _bad(lbc, threads=256, vec=2)
    # two clones per workgroup: the one-pass kernel only; 3 is no count
    assert "clones per workgroup" in _bad(lbc, threads=256, passes=2, cpw=2)
    assert "clones per workgroup" in _bad(lbc, threads=256, passes=1, cpw=3)
    assert "non-finite" in _bad(lbc, bias=float("nan"))
    # forcing mode and macroscopic output
    assert "forced_steps" in _bad(lbc, forced_steps=11)
    assert "forced_steps" in _bad(lbc, forced_steps=-1)
    assert "LBC_MODE_REF" in _bad(lbc, forced_steps=3)  # channel mode
    assert "field_sigma" in _bad(lbc, mode=lbc.MODE_REF, forced_steps=3,
                                 field_sigma=-1.0)
    assert "field_eps" in _bad(lbc, mode=lbc.MODE_REF, forced_steps=3,
                               field_eps=float("inf"))
    assert "macro_every" in _bad(lbc, macro_every=-2)


def test_null_context_calls_fail_cleanly(lbc):
    L = lbc.lib()
    assert L.lbc_run_interval(None, None) == -1
    assert L.lbc_gather(None) == -1
    assert L.lbc_sync(None) == -1
    assert b"null" in L.lbc_last_error(None)


def test_host_arrays_are_checked_before_the_library_reads_them(lbc):
    """The C ABI reads a fixed number of elements from caller arrays; the
    Python mirror checks type, layout and length first."""
    import numpy as np
    with pytest.raises(ValueError):
        lbc._sized(np.zeros(3), 4, "weights")
    assert lbc._sized(np.zeros(4), 4, "weights").size == 4
    with pytest.raises(TypeError):
        lbc._dp(np.zeros(4, np.float32))
    with pytest.raises(TypeError):
        lbc._dp(np.zeros((4, 4))[:, 0])
    with pytest.raises(TypeError):
        lbc._ip(np.zeros(4, np.int64))


def test_select_static_rejects_without_launching(lbc):
    """lbc_select_static checks its arguments on the host before it looks up
    its scratch or launches anything, so these run without a GPU."""
    fake = 1 << 20  # never dereferenced: every call below is rejected
    args = [fake] * 10
    for n, p, what in ((0, 1, "n must"), (lbc.STATIC_MAX_CLONES + 1, 1,
                                          "n must"), (10, 3, "rank_blocks")):
        with pytest.raises(lbc.LbcError) as e:
            lbc.select_static(n, p, *args)
        assert "(-1)" in str(e.value) and what in str(e.value)
    with pytest.raises(lbc.LbcError) as e:
        lbc.select_static(8, 1, fake, 0, *args[2:])
    assert "null device array" in str(e.value)
