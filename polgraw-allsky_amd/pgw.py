"""The HIP search core (libpgw_hip.so) through its C ABI (include/pgw.h).

Raises at load time if the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import abi

LIB_PATH = Path(__file__).resolve().parent / "libpgw_hip.so"
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


class PGWError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = lib().pgw_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} ({status})")


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        L.pgw_search_settings.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(abi.Settings)]
        L.pgw_rogcvir.argtypes = [C.c_double, C.c_double, _dp]
        L.pgw_sig2.restype = C.c_double
        L.pgw_sig2.argtypes = [_dp, C.c_int]
        L.pgw_init.argtypes = [C.POINTER(C.c_void_p), C.POINTER(abi.Settings), C.POINTER(abi.Detector),
                               C.c_int, C.c_int]
        L.pgw_job_core.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Range),
                                   C.POINTER(abi.Opts), _ip, _dp, C.c_int, _ip]
        L.pgw_job_batch.argtypes = [C.c_void_p, _ip, C.c_int, C.POINTER(abi.Range), C.POINTER(abi.Opts),
                                    _ip, _dp, C.c_int, _ip]
        _fp = C.POINTER(C.c_float)
        L.pgw_job_core_f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Range),
                                     C.POINTER(abi.Opts), _ip, _fp, C.c_int, _ip]
        L.pgw_job_batch_f.argtypes = [C.c_void_p, _ip, C.c_int, C.POINTER(abi.Range), C.POINTER(abi.Opts),
                                      _ip, _fp, C.c_int, _ip]
        L.pgw_debug_fstat.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Opts), _dp]
        L.pgw_debug_resampled.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.pgw_init_multi.argtypes = [C.POINTER(C.c_void_p), C.POINTER(abi.Settings), C.POINTER(abi.Detector),
                                     C.c_int, C.c_int, C.c_int]
        L.pgw_job_batch_multi.argtypes = [C.c_void_p, _ip, C.c_int, C.POINTER(abi.Range), C.POINTER(abi.Opts),
                                          _ip, _dp, C.c_int, _ip, _ip]
        L.pgw_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Opts),
                                   _ip, _ip, _ip]
        L.pgw_autotune_range.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.Opts),
                                         C.POINTER(abi.Range), C.c_int, _ip, _ip, _ip]
        L.pgw_tune_string.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.pgw_fstatnet.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _dp]
        L.pgw_fstatnet_multi.argtypes = [C.c_void_p, _ip, _dp, C.c_int, _dp, _dp]
        # device addresses as integers (caller-owned device buffers)
        L.pgw_fstatnet_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.pgw_job_spotlight.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, C.c_int, C.POINTER(abi.Opts),
                                        _ip, _dp, C.c_int, _ip]
        L.pgw_get_settings.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.Settings)]
        L.pgw_search.argtypes = [C.c_void_p, C.POINTER(abi.Range), C.POINTER(abi.Opts), C.POINTER(abi.Output),
                                 _ip]
        L.pgw_read_checkpoint.argtypes = [C.c_char_p, C.POINTER(abi.Range), _ip]
        L.pgw_time_stages.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(abi.Opts), C.c_int, _dp]
        L.pgw_copy_bench.argtypes = [C.c_size_t, C.c_int, _dp]
        L.pgw_hbm_probe.argtypes = [C.c_int, C.c_size_t, C.c_int, _dp]
        L.pgw_last_gpu_ms.restype = C.c_double
        L.pgw_last_gpu_ms.argtypes = [C.c_void_p]
        L.pgw_kernel_times.argtypes = [C.c_void_p, C.c_int, _dp]
        L.pgw_destroy.argtypes = [C.c_void_p]
        L.pgw_strerror.restype = C.c_char_p
        L.pgw_strerror.argtypes = [C.c_int]
        _lib = L
    return _lib


EXPORTED = ["pgw_search_settings", "pgw_rogcvir", "pgw_sig2", "pgw_init", "pgw_job_core", "pgw_job_batch",
            "pgw_job_core_f", "pgw_job_batch_f",
            "pgw_init_multi", "pgw_job_batch_multi", "pgw_autotune", "pgw_autotune_range", "pgw_tune_string", "pgw_fstatnet", "pgw_fstatnet_multi", "pgw_fstatnet_noalloc",
            "pgw_get_settings",
            "pgw_search", "pgw_read_checkpoint", "pgw_job_spotlight",
            "pgw_debug_fstat", "pgw_debug_resampled", "pgw_time_stages", "pgw_copy_bench", "pgw_hbm_probe",
            "pgw_last_gpu_ms",
            "pgw_kernel_times", "pgw_destroy", "pgw_strerror"]


def search_settings(fpo: float, dt: float, N: int, fftpad: int) -> abi.Settings:
    """search_settings() sizes for a segment of N samples (pgw_search_settings)."""
    s = abi.Settings()
    st = lib().pgw_search_settings(fpo, dt, N, fftpad, C.byref(s))
    if st != 0:
        raise PGWError(st, "pgw_search_settings")
    return s


def rogcvir(ephi: float, egam: float) -> np.ndarray:
    """Amplitude-modulation coefficients c1..c9 of a detector (pgw_rogcvir)."""
    c = np.zeros(9)
    st = lib().pgw_rogcvir(ephi, egam, c.ctypes.data_as(_dp))
    if st != 0:
        raise PGWError(st, "pgw_rogcvir")
    return c


def sig2(x) -> float:
    """crf0 * var(xDat) (pgw_sig2)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return float(lib().pgw_sig2(x.ctypes.data_as(_dp), len(x)))


def read_checkpoint(qname, rng):
    """read_checkpoints(): (status, FNum); status 1 = the state file is empty (run finished)."""
    fn = C.c_int(0)
    st = lib().pgw_read_checkpoint(str(qname).encode(), C.byref(rng), C.byref(fn))
    if st < 0:
        raise PGWError(st, "pgw_read_checkpoint")
    return st, fn.value


def hbm_probe(nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """GB/s this device reaches with streaming kernels over nbytes: copy (read + write bytes), read
    only and write only (pgw_hbm_probe)."""
    out = {}
    for kind, name in ((0, "copy"), (1, "read"), (2, "write")):
        ms = C.c_double()
        st = lib().pgw_hbm_probe(kind, nbytes, reps, C.byref(ms))
        if st != 0:
            raise PGWError(st, "pgw_hbm_probe")
        moved = 2 * nbytes if kind == 0 else nbytes
        out[name] = round(moved / ms.value / 1e6, 1)
    return out


def copy_bench(nbytes: int, reps: int = 20) -> float:
    """Average ms of a plain device copy of nbytes (read + write)."""
    ms = C.c_double()
    st = lib().pgw_copy_bench(nbytes, reps, C.byref(ms))
    if st != 0:
        raise PGWError(st, "pgw_copy_bench")
    return ms.value


class Handle:
    """One segment resident on the GPU (pgw_init ... pgw_destroy)."""

    def __init__(self, seg, precision: int = abi.PREC_FP64, lines=None):
        segs = seg if isinstance(seg, (list, tuple)) else [seg]
        self.seg = segs[0]
        self.segs = segs
        self.precision = precision
        sett = (abi.Settings * len(segs))()
        for k, s in enumerate(segs):
            sett[k] = abi.make_settings(s, lines)
        self._setts = sett
        self.sett = sett[0]
        self.dets = abi.make_detectors_multi(segs)
        self.h = C.c_void_p()
        st = lib().pgw_init_multi(C.byref(self.h), sett, self.dets, len(segs), len(segs[0].dets), precision)
        if st != 0:
            raise PGWError(st, "pgw_init")

    def close(self):
        if self.h:
            lib().pgw_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # record buffers are np.empty: the library writes the first sgnlc records and only those are
    # read back, so no page of the rest is ever touched
    def job_core(self, pm, mm, nn, rng, opts, cap=1 << 16):
        sgnlv = np.empty(abi.NPAR * cap)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_core(self.h, pm, mm, nn, C.byref(rng), C.byref(opts), C.byref(sgnlc),
                                sgnlv.ctypes.data_as(_dp), cap, C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_core")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR).copy(), fnum.value

    def job_batch(self, sky, rng, opts, cap=1 << 18):
        sky = np.ascontiguousarray(np.asarray(sky, dtype=np.int32).reshape(-1, 3))
        sgnlv = np.empty(abi.NPAR * cap)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_batch(self.h, sky.ctypes.data_as(_ip), len(sky), C.byref(rng), C.byref(opts),
                                 C.byref(sgnlc), sgnlv.ctypes.data_as(_dp), cap, C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_batch")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR).copy(), fnum.value

    def job_core_f(self, pm, mm, nn, rng, opts, cap=1 << 16):
        """job_core with float records (COMP_FLOAT drivers): (records (n, 5) float32, FNum)."""
        sgnlv = np.empty(abi.NPAR * cap, np.float32)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_core_f(self.h, pm, mm, nn, C.byref(rng), C.byref(opts), C.byref(sgnlc),
                                  sgnlv.ctypes.data_as(C.POINTER(C.c_float)), cap, C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_core_f")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR).copy(), fnum.value

    def job_batch_f(self, sky, rng, opts, cap=1 << 18):
        """job_batch with float records (COMP_FLOAT drivers): (records (n, 5) float32, FNum)."""
        sky = np.ascontiguousarray(np.asarray(sky, dtype=np.int32).reshape(-1, 3))
        sgnlv = np.empty(abi.NPAR * cap, np.float32)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_batch_f(self.h, sky.ctypes.data_as(_ip), len(sky), C.byref(rng), C.byref(opts),
                                   C.byref(sgnlc), sgnlv.ctypes.data_as(C.POINTER(C.c_float)), cap, C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_batch_f")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR).copy(), fnum.value

    def job_spotlight(self, pm, sky, opts, cap=1 << 18):
        """sky: [(mm, nn, [ss, ...]), ...] as a spotlight file lists them; returns (records, FNum)."""
        mm = np.ascontiguousarray([x[0] for x in sky], dtype=np.int32)
        nn = np.ascontiguousarray([x[1] for x in sky], dtype=np.int32)
        noss = np.ascontiguousarray([len(x[2]) for x in sky], dtype=np.int32)
        ss = np.ascontiguousarray([v for x in sky for v in x[2]] or [0], dtype=np.int32)
        sgnlv = np.empty(abi.NPAR * cap)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_spotlight(self.h, pm, mm.ctypes.data_as(_ip), nn.ctypes.data_as(_ip),
                                     noss.ctypes.data_as(_ip), ss.ctypes.data_as(_ip), len(sky), C.byref(opts),
                                     C.byref(sgnlc), sgnlv.ctypes.data_as(_dp), cap, C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_spotlight")
        return sgnlv[:abi.NPAR * sgnlc.value].reshape(-1, abi.NPAR).copy(), fnum.value

    def job_batch_multi(self, work, rng, opts, cap=1 << 18):
        """work rows (segment, pm, mm, nn); returns (records, work index per record, templates)."""
        work = np.ascontiguousarray(np.asarray(work, dtype=np.int32).reshape(-1, 4))
        sgnlv = np.empty(abi.NPAR * cap)
        owner = np.empty(cap, np.int32)
        sgnlc, fnum = C.c_int(0), C.c_int(0)
        st = lib().pgw_job_batch_multi(self.h, work.ctypes.data_as(_ip), len(work), C.byref(rng), C.byref(opts),
                                       C.byref(sgnlc), sgnlv.ctypes.data_as(_dp), cap,
                                       owner.ctypes.data_as(_ip), C.byref(fnum))
        if st != 0:
            raise PGWError(st, "pgw_job_batch_multi")
        n = sgnlc.value
        return sgnlv[:abi.NPAR * n].reshape(-1, abi.NPAR).copy(), owner[:n].copy(), fnum.value

    def fstatnet(self, pts, seg=0, grad=True):
        """Fstatnet at points (n, 4): (out (n, 11) with out[:, 5] = -F, dF/dpts (n, 4) or None)."""
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 4))
        out = np.zeros((len(pts), 11))
        g = np.zeros((len(pts), 4)) if grad else None
        st = lib().pgw_fstatnet(self.h, seg, pts.ctypes.data_as(_dp), len(pts), out.ctypes.data_as(_dp),
                                g.ctypes.data_as(_dp) if grad else None)
        if st != 0:
            raise PGWError(st, "pgw_fstatnet")
        return out, g

    def fstatnet_multi(self, seg, pts, grad=True):
        """Fstatnet at points (n, 4) of segments seg (n,) in one call: (out (n, 11), dF/dpts or None)."""
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 4))
        seg = np.ascontiguousarray(np.asarray(seg, dtype=np.int32).reshape(-1))
        if len(seg) != len(pts):
            raise ValueError("one segment per point")
        out = np.zeros((len(pts), 11))
        g = np.zeros((len(pts), 4)) if grad else None
        st = lib().pgw_fstatnet_multi(self.h, seg.ctypes.data_as(_ip), pts.ctypes.data_as(_dp), len(pts),
                                      out.ctypes.data_as(_dp), g.ctypes.data_as(_dp) if grad else None)
        if st != 0:
            raise PGWError(st, "pgw_fstatnet_multi")
        return out, g

    def search(self, rng, opts, prefix, ident=1, band=666, label="", qname=None, every=64, max_sky=0, FNum=0,
               float_records=False):
        """search(): trigger files under prefix, optional checkpoint state; returns FNum."""
        out = abi.Output(prefix=str(prefix).encode(), ident=ident, band=band, label=label.encode(),
                         qname=str(qname).encode() if qname else None, checkpoint_every=every, max_sky=max_sky,
                         float_records=int(float_records))
        fn = C.c_int(FNum)
        st = lib().pgw_search(self.h, C.byref(rng), C.byref(opts), C.byref(out), C.byref(fn))
        if st != 0:
            raise PGWError(st, "pgw_search")
        return fn.value

    def autotune(self, seg, pm, mm, nn, opts, rng=None, nsky=4):
        """pgw_autotune, or with a spindown range rng pgw_autotune_range (calls of nsky copies of
        the sky point with the templates a search with that range forms there)."""
        tb, na, seq = C.c_int(), C.c_int(), C.c_int()
        if rng is None:
            st = lib().pgw_autotune(self.h, seg, pm, mm, nn, C.byref(opts), C.byref(tb), C.byref(na), C.byref(seq))
        else:
            st = lib().pgw_autotune_range(self.h, seg, pm, mm, nn, C.byref(opts), C.byref(rng), nsky, C.byref(tb),
                                          C.byref(na), C.byref(seq))
        if st != 0:
            raise PGWError(st, "pgw_autotune")
        return dict(tb=tb.value, row_na=na.value, col_seq=seq.value)

    def tune_string(self):
        """The handle's kernel variants in PGW_TUNE syntax (pgw_tune_string)."""
        buf = C.create_string_buffer(256)
        st = lib().pgw_tune_string(self.h, buf, len(buf))
        if st != 0:
            raise PGWError(st, "pgw_tune_string")
        return buf.value.decode()

    def fstat(self, pm, mm, nn, ss, opts):
        F = np.zeros(self.sett.nmax - self.sett.nmin)
        st = lib().pgw_debug_fstat(self.h, pm, mm, nn, ss, C.byref(opts), F.ctypes.data_as(_dp))
        if st < 0:
            raise PGWError(st, "pgw_debug_fstat")
        return F if st == 0 else None

    def resampled(self, pm, mm, nn, det):
        N = self.sett.N
        a = np.zeros(2 * N)
        b = np.zeros(2 * N)
        st = lib().pgw_debug_resampled(self.h, pm, mm, nn, det, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp))
        if st < 0:
            raise PGWError(st, "pgw_debug_resampled")
        if st != 0:
            return None
        return a.view(np.complex128), b.view(np.complex128)

    def time_stages(self, pm, mm, nn, s0, nt, opts, reps=5):
        ms = np.zeros(5)
        st = lib().pgw_time_stages(self.h, pm, mm, nn, s0, nt, C.byref(opts), reps, ms.ctypes.data_as(_dp))
        if st != 0:
            raise PGWError(st, "pgw_time_stages")
        return dict(sky=ms[0], col=ms[1], row=ms[2], thr=ms[3], nt=int(ms[4]))

    def kernel_times(self, every: int) -> dict:
        """Totals of the kernel timing since the previous call (then reset); later searches record
        every `every`-th template launch (0: off) (pgw_kernel_times)."""
        out = np.zeros(5)
        st = lib().pgw_kernel_times(self.h, int(every), out.ctypes.data_as(_dp))
        if st != 0:
            raise PGWError(st, "pgw_kernel_times")
        return dict(col_ms=out[0], row_ms=out[1], thr_ms=out[2], launches=int(out[3]), templates=int(out[4]))

    @property
    def last_gpu_ms(self) -> float:
        return lib().pgw_last_gpu_ms(self.h)
