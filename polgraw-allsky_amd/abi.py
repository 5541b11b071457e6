"""ctypes mirror of include/pgw.h (structs only) and builders from a synthetic Segment."""
from __future__ import annotations

import ctypes as C

import numpy as np

NPAR = 5
MAXL = 2048
MAX_DETECTORS = 8

PGW_OK = 0
PGW_ERR_ARG = -1
PGW_ERR_CAPACITY = -2
PGW_ERR_DEVICE = -3
PGW_ERR_NOMEM = -4
PGW_ERR_UNSUPPORTED = -5
PGW_ERR_NODEVICE = -6
PGW_ERR_IO = -7
PGW_ERR_SCRATCH = -8

PREC_FP64 = 0
PREC_FP32 = 1


class Settings(C.Structure):
    _fields_ = [
        ("fpo", C.c_double), ("dt", C.c_double), ("B", C.c_double), ("oms", C.c_double),
        ("omr", C.c_double), ("Smin", C.c_double), ("Smax", C.c_double),
        ("sepsm", C.c_double), ("cepsm", C.c_double),
        ("nfft", C.c_int), ("N", C.c_int), ("nfftf", C.c_int), ("nmax", C.c_int),
        ("nmin", C.c_int), ("nd", C.c_int), ("interpftpad", C.c_int), ("fftpad", C.c_int),
        ("Ninterp", C.c_int), ("nifo", C.c_int),
        ("M", C.c_double * 16),
        ("numlines_band", C.c_int),
        ("lines", (C.c_double * 2) * MAXL),
    ]


class Detector(C.Structure):
    _fields_ = [
        ("xDat", C.POINTER(C.c_double)), ("DetSSB", C.POINTER(C.c_double)),
        ("sig2", C.c_double), ("sphir", C.c_double), ("cphir", C.c_double),
        ("amod", C.c_double * 9),
    ]


class Range(C.Structure):
    _fields_ = [
        ("pmr", C.c_int * 2), ("mr", C.c_int * 2), ("nr", C.c_int * 2), ("spndr", C.c_int * 2),
        ("pst", C.c_int), ("mst", C.c_int), ("nst", C.c_int), ("sst", C.c_int),
    ]


class Output(C.Structure):
    _fields_ = [("prefix", C.c_char_p), ("ident", C.c_int), ("band", C.c_int), ("label", C.c_char_p),
                ("qname", C.c_char_p), ("checkpoint_every", C.c_int), ("max_sky", C.c_int),
                ("float_records", C.c_int)]


class Opts(C.Structure):
    _fields_ = [("trl", C.c_double), ("white_flag", C.c_int), ("s0_flag", C.c_int),
                ("use_range", C.c_int)]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_settings(seg, lines=None) -> Settings:
    s = Settings()
    d = seg.sett
    for k in ("fpo", "dt", "B", "oms", "omr", "Smin", "Smax"):
        setattr(s, k, float(d[k]))
    s.sepsm = float(np.sin(seg.epsm))
    s.cepsm = float(np.cos(seg.epsm))
    for k in ("nfft", "N", "nfftf", "nmax", "nmin", "nd", "interpftpad", "fftpad", "Ninterp"):
        setattr(s, k, int(d[k]))
    s.nifo = len(seg.dets)
    for i in range(16):
        s.M[i] = float(seg.M[i])
    if lines is not None:
        s.numlines_band = len(lines)
        for k, (lo, hi) in enumerate(lines):
            s.lines[k][0] = lo
            s.lines[k][1] = hi
    return s


def make_detectors(seg):
    """Detector array; keeps references to the numpy buffers alive on the returned object."""
    return make_detectors_multi([seg])


def make_detectors_multi(segs):
    """Detectors of several segments, segment-major (ifo[k*nifo + d])."""
    nifo = len(segs[0].dets)
    arr = (Detector * (nifo * len(segs)))()
    keep = []
    for s, seg in enumerate(segs):
        assert len(seg.dets) == nifo
        for k in range(nifo):
            x = np.ascontiguousarray(seg.xDat[k], dtype=np.float64)
            e = np.ascontiguousarray(seg.DetSSB[k], dtype=np.float64)
            keep += [x, e]
            a = arr[s * nifo + k]
            a.xDat = dptr(x)
            a.DetSSB = dptr(e)
            a.sig2 = float(seg.sig2[k])
            a.sphir = float(np.sin(seg.phir[k]))
            a.cphir = float(np.cos(seg.phir[k]))
            for i in range(9):
                a.amod[i] = float(seg.amod[k][i])
    arr._keep = keep
    return arr


def make_range(seg, spndr=None) -> Range:
    r = Range()
    rr = seg.range
    sp = spndr if spndr is not None else rr["spndr"]
    r.spndr[0], r.spndr[1] = sp
    r.nr[0], r.nr[1] = rr["nr"]
    r.mr[0], r.mr[1] = rr["mr"]
    r.pmr[0], r.pmr[1] = rr["pmr"]
    r.pst, r.mst, r.nst, r.sst = rr["pmr"][0], rr["mr"][0], rr["nr"][0], sp[0]
    return r


def make_opts(trl=20.0, white=False, s0=False, use_range=True) -> Opts:
    return Opts(trl=trl, white_flag=int(white), s0_flag=int(s0), use_range=int(use_range))
