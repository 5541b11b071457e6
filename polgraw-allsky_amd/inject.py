"""Software injections: the reference's ``-addsig`` path.

With ``-addsig <file>`` the reference adds a simulated signal to the frame's data and searches
a box of +-gsize grid points around it instead of the range file
(search/network/src-cpu/main.c:115-121, add_signal() in init.c:462-651).  The signal file is
what ``sigen`` prints (sigen.c:360-367):

    amp <h0> <gsize> <reffr>        or      snr <rho> <gsize> <reffr>
    f fdot delta alpha a1 a2 a3 a4  (8 doubles: frequency and spindown in radians per sample at
                                     the reference frame reffr, sky position, the four amplitudes
                                     of Phys. Rev. D 82, 022005, Eqs. 2.13a-d)

Host-side input preparation, like the file readers: the search itself runs on the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .synth import C_EPSMA, amplitudes, ast2lin, lin2ast, modvir


@dataclass
class Signal:
    kind: str          # "amp" (value = h0) or "snr" (value = network signal-to-noise ratio)
    value: float
    gsize: int
    reffr: int
    sgnlo: np.ndarray  # 8 doubles


def read_signal(path) -> Signal:
    """Parses a signal file the way add_signal() scans it (init.c:479-506)."""
    tok = open(path).read().split()
    if not tok or tok[0] not in ("amp", "snr"):
        raise ValueError(f"{path}: the first word must be 'amp' or 'snr'")
    if len(tok) < 12:
        raise ValueError(f"{path}: {len(tok)} words, a signal file holds 12")
    return Signal(kind=tok[0], value=float(tok[1]), gsize=int(tok[2]), reffr=int(tok[3]),
                  sgnlo=np.array([float(t) for t in tok[4:12]]))


def write_signal(path, sig: Signal) -> None:
    """Writes a signal file in sigen's layout (sigen.c:360-367)."""
    with open(path, "w") as f:
        f.write("%s %e\n%d\n%d\n" % (sig.kind, sig.value, sig.gsize, sig.reffr))
        f.write("".join("%.16e\n" % v for v in sig.sgnlo))


def random_signal(rng, sett: dict, kind: str = "snr", value: float = 10.0, gsize: int = 2,
                  reffr: int = 0) -> Signal:
    """A random signal as sigen draws it (sigen.c:270-367), without its known-line and near-pole
    options: spindown uniform in [-Smax, 0]; hemisphere 1 or 2; sky position uniform over the
    disk of linear sky coordinates (be1^2 + be2^2 <= 1) mapped by lin2ast; initial phase and
    polarisation angle uniform in [0, 2 pi), cos(inclination) uniform in [-1, 1].  The frequency
    is uniform over the searched band [2 pi nmin, 2 pi nmax) / (fftpad nfft) at frame reffr (the
    reference's draw, pi U (nmax - nmin) / nfft + nmin / nfft, is near that band for fftpad 2).
    rng: a numpy Generator."""
    L = sett["fftpad"] * sett["nfft"]
    fdot = -sett["Smax"] * rng.uniform()
    f = 2.0 * math.pi * (sett["nmin"] + rng.uniform() * (sett["nmax"] - sett["nmin"])) / L
    pm = 1 + int(rng.integers(0, 2))
    while True:
        be1, be2 = 2.0 * rng.uniform() - 1.0, 2.0 * rng.uniform() - 1.0
        if be1 * be1 + be2 * be2 <= 1.0:
            break
    delta, alpha = lin2ast(be1, be2, pm)
    ph_o, psik = 2.0 * math.pi * rng.uniform(), 2.0 * math.pi * rng.uniform()
    amp = amplitudes(ph_o, psik, 2.0 * rng.uniform() - 1.0)
    return Signal(kind=kind, value=value, gsize=gsize, reffr=reffr,
                  sgnlo=np.array([f, fdot, delta, alpha, *amp]))


def _cround(x: float) -> int:
    """C's round(): halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def signal_range(sett: dict, M, sig: Signal, ident: int):
    """The search box around the injection (init.c:516-572): the grid coordinates solving
    M^T x = (f, fdot, be1 cof, be2 cof) with f moved from frame reffr to frame ident, rounded,
    +-gsize.  Returns (range dict, sgnlo with f moved, cof).  A frequency outside [0, pi] after the
    move is an error (the reference exits with status 171 below the band, 187 above it)."""
    N = int(sett["N"])
    s = np.array(sig.sgnlo, dtype=np.float64)
    s[0] += -2.0 * s[1] * N * (sig.reffr - ident)
    if s[0] < 0:
        raise ValueError("injected frequency below the band (the reference exits with 171)")
    if s[0] > math.pi:
        raise ValueError("injected frequency above the band (the reference exits with 187)")
    cof = sett["oms"] + s[0]
    pm, be = ast2lin(s[3], s[2], C_EPSMA)
    rhs = np.array([s[0], s[1], be[0] * cof, be[1] * cof])
    A = np.asarray(M, dtype=np.float64).reshape(4, 4).T   # element (r, c) = M[4c + r]
    x = np.linalg.solve(A, rhs)
    ss, nn, mm = _cround(x[1]), _cround(x[2]), _cround(x[3])
    g = int(sig.gsize)
    rng = dict(spndr=[ss - g, ss + g], nr=[nn - g, nn + g], mr=[mm - g, mm + g], pmr=[pm, pm])
    return rng, s, cof


def signal_series(frame, s: np.ndarray, cof: float, shift_frames: int) -> list:
    """Per detector, the unit-amplitude signal of add_signal() (init.c:574-619): amplitude
    modulation of the detector (modvir, settings.c:281-325), phase f i + fdot i^2 + (cof + 2 fdot i)
    (n . DetSSB_i) minus the phase accumulated over shift_frames = reffr - ident frames."""
    N = int(frame.sett["N"])
    i = np.arange(N, dtype=np.float64)
    omr = frame.sett["omr"]
    cosmodf, sinmodf = np.cos(omr * i), np.sin(omr * i)
    sd, cd, sa, ca = math.sin(s[2]), math.cos(s[2]), math.sin(s[3]), math.cos(s[3])
    D = float(N * shift_frames)
    phaseshift = s[0] * N * shift_frames + s[1] * D ** 2
    out = []
    for k in range(len(frame.dets)):
        aa, bb = modvir(sa, ca, sd, cd, math.sin(frame.phir[k]), math.cos(frame.phir[k]),
                        frame.amod[k], cosmodf, sinmodf)
        e = np.asarray(frame.DetSSB[k], dtype=np.float64).reshape(-1, 3)
        shift = (ca * cd) * e[:, 0] + (sa * cd) * e[:, 1] + sd * e[:, 2]
        ph = s[0] * i + s[1] * (i * i) + (cof + 2.0 * s[1] * i) * shift - phaseshift
        c, sn = np.cos(ph), np.sin(ph)
        out.append(s[4] * aa * c + s[6] * aa * sn + s[5] * bb * c + s[7] * bb * sn)
    return out


def add_signal(frame, sig: Signal, ident: int) -> dict:
    """add_signal() (init.c:462-651) on a loaded frame: adds h0 times the signal to every
    non-zero sample of each detector (gaps stay zero), sets frame.range to the box around the
    injection and frame.injection to what was added, and returns the range.  With an 'snr' file
    h0 = snr / sqrt(sum of the squared signal over all detectors and samples).  frame.sig2 is
    left as it was: the reference computes the noise variances before the injection
    (init_arrays() precedes add_signal() in main.c:81-121)."""
    rng, s, cof = signal_range(frame.sett, frame.M, sig, ident)
    series = signal_series(frame, s, cof, sig.reffr - ident)
    h0 = sig.value if sig.kind == "amp" else 0.0
    if sig.kind == "snr" and sig.value:
        total = math.fsum(float(np.dot(v, v)) for v in series)
        h0 = sig.value / math.sqrt(total)
    for k, v in enumerate(series):
        x = np.array(frame.xDat[k], dtype=np.float64, copy=True)
        nz = x != 0.0
        x[nz] += h0 * v[nz]
        frame.xDat[k] = x
    frame.range = rng
    frame.injection = dict(h0=h0, sgnlo=s, pm=rng["pmr"][0],
                           grid=np.array([0, (rng["spndr"][0] + rng["spndr"][1]) // 2,
                                          (rng["nr"][0] + rng["nr"][1]) // 2,
                                          (rng["mr"][0] + rng["mr"][1]) // 2]))
    return rng
