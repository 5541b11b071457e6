"""pgw_fstatnet_noalloc through the C ABI: caller-owned device buffers (hipMalloc'd by the test
through the HIP runtime the library itself uses), partial sums in the library's fixed device
scratch.  Two calls in a row, the second with fewer points, each against pgw_fstatnet on the same
points (same kernels: bit-equal); sizes beyond the scratch are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from cases import segment

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from polgraw_allsky_amd import pgw
    pgw.lib()  # loads the HIP runtime the library links
    H = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        a = np.zeros(shape)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _points(seg, n, seed):
    inj = seg.injection
    truth = np.array([inj["f"], inj["fdot"], inj["delta"], inj["alpha"]])
    sc = np.array([2e-5, 2e-10, 2e-3, 2e-3])
    return truth + np.random.default_rng(seed).uniform(-1, 1, (n, 4)) * sc


@pytest.mark.parametrize("name", ["net-2det-86k", "gaps-3det-30k"])
def test_noalloc_equals_fstatnet(name, hip, abi):
    from polgraw_allsky_amd import pgw
    seg = segment(name)
    h = pgw.Handle(seg, abi.PREC_FP64)
    L = pgw.lib()
    nmax = 1024
    bp, bo, bg = DevBuf(hip, 32 * nmax), DevBuf(hip, 88 * nmax), DevBuf(hip, 32 * nmax)
    try:
        for call, n in enumerate((1024, 300)):
            pts = _points(seg, n, seed=call)
            ref_out, ref_g = h.fstatnet(pts)
            bp.put(pts)
            st = L.pgw_fstatnet_noalloc(h.h, 0, bp.p, n, bo.p, bg.p)
            print(name, "call", call, "points", n, "status", st, flush=True)
            assert st == 0
            out, g = bo.get((n, 11)), bg.get((n, 4))
            # out[4] = sqrt(2 (F - 2)) is NaN where F < 2, in both
            assert np.array_equal(out, ref_out, equal_nan=True)
            assert np.array_equal(g, ref_g)
            # values only (no gradient buffer)
            st = L.pgw_fstatnet_noalloc(h.h, 0, bp.p, n, bo.p, None)
            assert st == 0
            assert np.array_equal(bo.get((n, 11)), h.fstatnet(pts, grad=False)[0], equal_nan=True)
        # beyond the scratch (npts * chunks * 26 doubles > 8 MiB): refused on the host
        nch = -(-seg.sett["N"] // 4096)
        too_many = (8 << 20) // (26 * 8) // nch + 1
        assert L.pgw_fstatnet_noalloc(h.h, 0, bp.p, too_many, bo.p, bg.p) == abi.PGW_ERR_SCRATCH
        assert L.pgw_fstatnet_noalloc(h.h, 0, None, 10, bo.p, None) == abi.PGW_ERR_ARG
        assert L.pgw_fstatnet_noalloc(h.h, 5, bp.p, 10, bo.p, None) == abi.PGW_ERR_ARG
    finally:
        for b in (bp, bo, bg):
            b.free()
        h.close()
