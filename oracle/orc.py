"""ctypes access to the CPU oracle (liborc.so) and to the reference's compiled auxi.c (_ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
import pgwload  # noqa: E402

pgwload.load()
from polgraw_allsky_amd import abi  # noqa: E402
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        # ORC_LIB: another build of the oracle (the sanitizer builds of oracle/Makefile)
        path = Path(os.environ.get("ORC_LIB", HERE / "liborc.so"))
        if not path.exists():
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(str(path))
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.Settings), C.POINTER(abi.Detector), C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_job_core.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Range),
                                   C.POINTER(abi.Opts), _dp, _ip, _dp, C.c_int, _ip]
        L.orc_fstat_one.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(abi.Opts), _dp]
        L.orc_resampled.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_spindown_range.argtypes = [C.POINTER(abi.Settings), C.c_int, C.c_int,
                                         C.POINTER(abi.Range), C.POINTER(abi.Opts), _ip, _ip]
        L.orc_settings.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(abi.Settings)]
        L.orc_fft_plan_init.argtypes = [C.c_void_p, C.c_int]
        L.orc_fft_plan_free.argtypes = [C.c_void_p]
        L.orc_fft_exec.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_lin2ast.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                  _dp, _dp, _dp, _dp]
        L.orc_rogcvir.argtypes = [C.c_double, C.c_double, _dp]
        L.orc_modvir.argtypes = [C.c_double] * 6 + [_dp, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_spline.argtypes = [_dp, C.c_int, _dp, _dp]
        L.orc_splintpad.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_var.restype = C.c_double
        L.orc_var.argtypes = [_dp, C.c_int]
        L.orc_fstat_norm.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_fstatnet_batch.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp]
        L.orc_amoeba.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_mads.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_mads_stencil.argtypes = [C.POINTER(C.c_byte)]
        _lib = L
    return _lib


def ref():
    """The reference's own auxi.c compiled by oracle/Makefile, or None if absent."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libref_auxi.so"
        if not path.exists():
            return None
        R = C.CDLL(str(path))
        R.lin2ast.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        R.spline.argtypes = [_dp, C.c_int, _dp]
        R.splintpad.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        R.var.restype = C.c_double
        R.var.argtypes = [_dp, C.c_int]
        R.FStat.restype = C.c_double
        R.FStat.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        R.ast2lin.restype = C.c_int
        R.ast2lin.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        # spline() keeps a static scratch sized by its first call: size it once, large.
        big = 1 << 21
        y = np.zeros(big, np.complex128)
        y2 = np.zeros(big, np.complex128)
        R.spline(y.ctypes.data_as(_dp), big, y2.ctypes.data_as(_dp))
        R._spline_cap = big
        _ref = R
    return _ref


def cptr(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class FFTPlan(C.Structure):
    _fields_ = [("n", C.c_int), ("tw", C.c_void_p), ("work", C.c_void_p)]


def fft(x: np.ndarray, direction: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    p = FFTPlan()
    if lib().orc_fft_plan_init(C.byref(p), len(x)) != 0:
        raise ValueError("bad FFT length")
    try:
        lib().orc_fft_exec(C.byref(p), cptr(x), direction, None)
    finally:
        lib().orc_fft_plan_free(C.byref(p))
    return x


class Search:
    """Oracle search state for one segment."""

    def __init__(self, seg, nthreads: int = 1, lines=None):
        self.seg = seg
        self.sett = abi.make_settings(seg, lines)
        self.dets = abi.make_detectors(seg)
        self.h = lib().orc_create(C.byref(self.sett), self.dets, len(seg.dets), nthreads)
        if not self.h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def job_core(self, pm, mm, nn, rng, opts, cap=1 << 16):
        sgnlv = np.zeros(abi.NPAR * cap)
        sgnlc = C.c_int(0)
        fnum = C.c_int(0)
        st = lib().orc_job_core(self.h, pm, mm, nn, C.byref(rng), C.byref(opts), None,
                                C.byref(sgnlc), sgnlv.ctypes.data_as(_dp), cap, C.byref(fnum))
        if st != 0:
            raise RuntimeError(f"orc_job_core status {st}")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR), fnum.value

    def fstat(self, pm, mm, nn, ss, opts):
        F = np.zeros(self.sett.nmax - self.sett.nmin)
        st = lib().orc_fstat_one(self.h, pm, mm, nn, ss, C.byref(opts), F.ctypes.data_as(_dp))
        return F if st == 0 else None

    def resampled(self, pm, mm, nn, det):
        N = self.sett.N
        a = np.zeros(N, np.complex128)
        b = np.zeros(N, np.complex128)
        st = lib().orc_resampled(self.h, pm, mm, nn, det, cptr(a), cptr(b))
        return (a, b) if st == 0 else None

    def fstatnet(self, pts, grad=True):
        """Fstatnet at points (n, 4) = (omega, omegadot, delta, alpha): (out (n, 11), dF/dpt (n, 4))."""
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 4))
        out = np.zeros((len(pts), 11))
        g = np.zeros((len(pts), 4)) if grad else None
        st = lib().orc_fstatnet_batch(self.h, pts.ctypes.data_as(_dp), len(pts), out.ctypes.data_as(_dp),
                                      g.ctypes.data_as(_dp) if grad else None)
        if st != 0:
            raise RuntimeError("orc_fstatnet failed")
        return out, g

    def amoeba(self, point, pc2, tol=1e-7, maxeval=20000):
        """followup's Nelder-Mead from point; returns (best point, its 11 values, evaluations)."""
        x = np.ascontiguousarray(np.asarray(point, dtype=np.float64).copy())
        step = np.ascontiguousarray(np.asarray(pc2, dtype=np.float64))
        out = np.zeros(11)
        nev = lib().orc_amoeba(self.h, x.ctypes.data_as(_dp), step.ctypes.data_as(_dp), tol, maxeval,
                               out.ctypes.data_as(_dp))
        return x, out, nev

    def mads(self, point, p2, nod, limit=700):
        """followup's invertedMADS from point; returns (final centre, its 11 values, rounds)."""
        x = np.ascontiguousarray(np.asarray(point, dtype=np.float64).copy())
        p = np.ascontiguousarray(np.asarray(p2, dtype=np.float64))
        out = np.zeros(11)
        r = lib().orc_mads(self.h, x.ctypes.data_as(_dp), p.ctypes.data_as(_dp), nod, limit,
                           out.ctypes.data_as(_dp))
        return x, out, r

    def spindown_range(self, mm, nn, rng, opts):
        lo, hi = C.c_int(), C.c_int()
        lib().orc_spindown_range(C.byref(self.sett), mm, nn, C.byref(rng), C.byref(opts),
                                 C.byref(lo), C.byref(hi))
        return lo.value, hi.value


def mads_stencil() -> np.ndarray:
    """The oracle's 80 mesh offsets (signs) in the reference's order."""
    st = np.zeros((80, 4), np.int8)
    lib().orc_mads_stencil(st.ctypes.data_as(C.POINTER(C.c_byte)))
    return st


def default_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
