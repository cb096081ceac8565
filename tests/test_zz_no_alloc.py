"""The allocation-free normalize_cells (sev_normalize_cells_noalloc) against
the host entry point sev_normalize_cells on the same counts: bit for bit, on
two calls in a row (the second with fewer cells, so totals left in the shared
scratch by the first call would show), in both output types; sizes beyond
the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib(sev):
    L = sev.load()
    return L, L.sev_normalize_cells_noalloc


def _counts(sev, orc, cells, genes, density, seed):
    t = sev.nb_tables(genes, seed=seed, density=density)
    return orc.synth_csc(cells, t)


def _host(sev, m, n, cp, rv, cnt, method, dtype):
    out = np.zeros(len(rv), dtype)
    L = sev.load()
    code = 0 if dtype == np.float32 else 1
    cp1 = np.ascontiguousarray(cp + 1, np.int64)  # Julia's 1-based Int64 layout
    rv1 = np.ascontiguousarray(rv.astype(np.int64) + 1)
    cnt32 = np.ascontiguousarray(cnt, np.int32)
    rc = L.sev_normalize_cells(m, n, cp1.ctypes.data, rv1.ctypes.data, cnt32.ctypes.data, 0,
                               method, 1e4, code, out.ctypes.data)
    assert rc == 0, L.sev_last_error()
    return out


class _Hip:
    """The HIP runtime that libsevero_hip.so links (same soname, so the same
    loaded instance): device buffers and a stream for the caller's side."""

    def __init__(self, sev):
        sev.load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h = h
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _device(sev, f, m, n, cp, rv, cnt, method, dtype):
    hip = _Hip(sev)
    try:
        d_cp = hip.put(cp.astype(np.int64))
        d_rv = hip.put(rv.astype(np.int32))
        d_cnt = hip.put(cnt.astype(np.int32))
        out = np.full(len(rv), np.nan, dtype)
        d_out = hip.put(out)
        st = C.c_void_p()
        assert hip.h.hipStreamCreate(C.byref(st)) == 0
        rc = f(m, n, len(rv), d_cp, d_rv, d_cnt, method, 1e4, 0 if dtype == np.float32 else 1,
               d_out, st)
        assert rc == 0, sev.load().sev_last_error()
        assert hip.h.hipStreamSynchronize(st) == 0
        hip.h.hipStreamDestroy(st)
        return hip.get(d_out, out)
    finally:
        hip.free()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", [0, 1])  # lognormalize, relativecounts
def test_noalloc_normalize_matches_host_entry(sev, orc, dtype, method):
    _, f = _lib(sev)
    # two calls in a row on one GPU: the second has fewer cells and genes
    for cells, genes, seed in ((30000, 3000, 1), (7000, 1200, 2)):
        cp, rv, cnt = _counts(sev, orc, cells, genes, 0.05, seed)
        got = _device(sev, f, cells, genes, cp, rv, cnt, method, dtype)
        ref = _host(sev, cells, genes, cp, rv, cnt, method, dtype)
        assert np.array_equal(got, ref), (cells, genes)


def test_noalloc_rejects_too_many_cells(sev):
    L, f = _lib(sev)
    rc = f((1 << 23) + 1, 1, 0, None, None, None, 0, 1e4, 0, None, None)
    assert rc == -1
    assert b"scratch" in L.sev_last_error()
