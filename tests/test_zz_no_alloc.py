"""lbc_select_static, the selection entry point that allocates nothing,
against lbc_select on the same device log-weights and draws: bit-exact
counts, pairs, labels, marks, status and mean weight, on two calls in a row
(the second with fewer clones, so it runs over what the first left in the
shared scratch).  The test allocates the device arrays itself through the
HIP runtime the library is linked against (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """Device buffers from the library's own HIP runtime (the same loaded
    copy of libamdhip64), copied to and from numpy arrays."""

    def __init__(self):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.c_int]
        self.ptrs = []

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0
        return p.value

    def get(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _context_selection(lbc, n, p, k, seed):
    """A channel interval of n clones, then lbc_select with host draws."""
    spec = lbc.channel(40, 20, 4, 10, 10, 0.55, 0.05, n, 6, k=k)
    spec.rank_blocks = p
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    rint = rng.integers(0, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    with lbc.Sampler(spec, mode=lbc.MODE_CHANNEL, seed=seed) as s:
        s.init_equilibrium(0, 2e-2)
        s.run_interval()
        _, logw = s.logweights(0)
        sel = s.select(0, u=u, rint=rint)
        marks = s.get_marks(0)
    return logw, u, rint, sel, marks


@pytest.mark.parametrize("calls", [((4096, 4, 300.0), (700, 1, 300.0))])
def test_select_static_matches_select(lbc, calls):
    dev = DeviceArrays()
    adjusted = 0
    try:
        for n, p, k in calls:
            print(f"case: lbc_select_static n={n} rank_blocks={p} k={k}",
                  flush=True)
            logw, u, rint, sel, marks = _context_selection(lbc, n, p, k, n)
            # every call copies; adjusted counts the calls that prune or add
            assert len(sel.pairs) > 0
            adjusted += int(sel.nb_before.sum()) != n
            i32 = np.full(n, -7, np.int32)
            host = dict(nb0=i32, nb1=i32, labels=i32,
                        pairs=np.full(2 * n, -7, np.int32),
                        marks=np.full(n, 7, np.uint8),
                        info=np.full(4, -7, np.int32), R=np.zeros(2))
            d = {key: dev.put(v) for key, v in host.items()}
            lbc.select_static(n, p, dev.put(logw), dev.put(u), dev.put(rint),
                              d["nb0"], d["nb1"], d["pairs"], d["labels"],
                              d["marks"], d["info"], d["R"])
            got = {key: dev.get(d[key], v) for key, v in host.items()}
            info = got["info"]
            npairs = len(sel.pairs)
            assert info[0] == npairs and info[1] == sel.n_rint_used
            assert info[2] == 0
            assert np.array_equal(got["nb0"], sel.nb_before)
            assert np.array_equal(got["nb1"], sel.nb_after)
            assert np.array_equal(got["pairs"][:2 * npairs], sel.pairs.ravel())
            assert np.array_equal(got["labels"], sel.labels)
            assert np.array_equal(got["marks"], marks)
            assert got["R"][0] == sel.total_R
            assert got["R"][1] == sel.log_total_R
    finally:
        dev.free()
    assert adjusted > 0
