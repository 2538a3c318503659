"""jpxlg_score_pair_noalloc through the C ABI: the no-allocation score (its
intermediate buffers in the library's static device scratch array) equals
jpxlg_score_pair on the same pair, on two calls in a row (the second with a
smaller image, so stale contents of the shared array would show), and a pair
too large for the array is refused before any work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(w, h, seed):
    rng = np.random.default_rng(seed)
    ref = rng.random((3, h, w)).astype(np.float32)
    cand = np.clip(ref + rng.standard_normal(ref.shape).astype(np.float32) * np.float32(0.03), 0, 1)
    return ref, cand.astype(np.float32)


class _Hip:
    """Device buffers for the test's inputs, from the HIP runtime the library
    itself loaded (the same shared object, found by its soname)."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        self.h = h

    def upload(self, a: np.ndarray) -> ctypes.c_void_p:
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        self.h.hipFree(p)


def _noalloc(ctx, dref, dcand, w, h):
    score = ctypes.c_double()
    terms = np.zeros(108, np.float64)
    nterms = ctypes.c_uint32()
    rc = ctx.lib.jpxlg_score_pair_noalloc(ctx.handle, w, h, dref, dcand, ctypes.byref(score),
                                          terms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nterms))
    return rc, score.value, terms[: nterms.value]


def test_noalloc_score_equals_score_pair():
    import jpxl_amd
    from jpxl_amd import _lib

    ctx = jpxl_amd.Context(0)
    hip = _Hip()
    try:
        for w, h, seed in [(1024, 1000, 1), (333, 170, 2)]:
            ref, cand = _pair(w, h, seed)
            want, want_terms = ctx.score_pair(ref, cand)
            dref, dcand = hip.upload(ref), hip.upload(cand)
            rc, got, terms = _noalloc(ctx, dref, dcand, w, h)
            assert rc == 0, ctx.lib.jpxlg_last_error(ctx.handle)
            assert got == want
            assert np.array_equal(terms, want_terms)
            # the inputs are read only
            assert np.array_equal(hip.download(dref, ref), ref) and np.array_equal(hip.download(dcand, cand), cand)
            hip.free(dref)
            hip.free(dcand)
        # too large for the scratch array: refused with a clear error, before any work
        big = np.zeros((3, 1400, 1400), np.float32)
        dbig = hip.upload(big)
        rc, _, _ = _noalloc(ctx, dbig, dbig, 1400, 1400)
        hip.free(dbig)
        assert rc == _lib.E_UNSUPPORTED
        assert b"scratch array" in ctx.lib.jpxlg_last_error(ctx.handle)
    finally:
        ctx.close()
