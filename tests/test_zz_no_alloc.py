"""The allocation-free diffusive step (trmc_diffusive_tridiag_f32_noalloc)
against the allocating entry point trmc_diffusive_tridiag_f32, bit for bit, on
two calls in a row (the second smaller, on the same static scratch), through
the C-ABI with device-resident arrays.  Device memory comes from the HIP
runtime the library itself is linked against (ctypes), in the test only.
Also: a size above the limit and an invalid reach record are argument errors."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_trmc
from test_diffusive import NODES, _data

trmc = load_trmc()
H2D, D2H = 1, 2


def _hip():
    trmc.lib()                                  # loads the runtime the library uses
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DevArray:
    """A device copy of a numpy array (test plumbing)."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.a.nbytes, 1)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _run_noalloc(hip, rp, dx, df, ce, qp, out0):
    arrs = [DevArray(hip, a) for a in (rp, dx, df, ce, qp, out0)]
    try:
        trmc.diffusive_tridiag_noalloc(*[d.p.value for d in arrs], len(rp))
        return arrs[5].get()
    finally:
        for d in arrs:
            d.free()


@pytest.mark.gpu
def test_noalloc_step_equals_the_allocating_one(gpu):
    hip = _hip()
    for nreach, seed in ((3000, 5), (700, 6)):       # the second call is smaller
        rp, dx, df, ce, qp = _data(nreach, seed=seed)
        out0 = np.random.default_rng(seed + 1).random(nreach * NODES).astype(np.float32)  # beyond ncomp: kept
        want = trmc.diffusive_tridiag(rp, dx, df, ce, qp, qp_out=out0)
        got = _run_noalloc(hip, rp, dx, df, ce, qp, out0)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), nreach


@pytest.mark.gpu
def test_noalloc_rejects_sizes_and_bad_records(gpu):
    hip = _hip()
    with pytest.raises(trmc.TrmcError) as e:
        trmc.diffusive_tridiag_noalloc(1, 1, 1, 1, 1, 1, trmc.DIFF_NOALLOC_MAX_REACHES + 1)
    assert e.value.code == trmc.ERR_ARG
    rp, dx, df, ce, qp = _data(200, seed=9)
    want = trmc.diffusive_tridiag(rp, dx, df, ce, qp)
    bad = rp.copy()
    bad["ncomp"][17] = NODES + 1
    arrs = [DevArray(hip, a) for a in (bad, dx, df, ce, qp, np.zeros(200 * NODES, np.float32))]
    try:
        with pytest.raises(trmc.TrmcError) as e:
            trmc.diffusive_tridiag_noalloc(*[d.p.value for d in arrs], 200)
        assert e.value.code == trmc.ERR_ARG
        got = arrs[5].get().reshape(200, NODES)
    finally:
        for d in arrs:
            d.free()
    w = want.reshape(200, NODES)
    keep = np.arange(200) != 17
    assert np.array_equal(got[keep].view(np.uint32), w[keep].view(np.uint32))
    assert not got[17].any()                       # the invalid reach is untouched
