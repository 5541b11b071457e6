"""The reference's file formats on either side of the search core.

Inputs of a frame (one time segment) as init_arrays()/read_grid() read them
(search/network/src-cpu/init.c:287-457, settings.c:134-232):

  <dtaprefix>/<ident %03d>/grid.bin                         int fftpad, 16 doubles M (column-major)
  <dtaprefix>/<ident %03d>/<IFO>/xdatc_<ident>_<band><label>.bin   N doubles, zeros mark gaps
  <dtaprefix>/<ident %03d>/<IFO>/DetSSB.bin                  3N doubles (detector position w.r.t. the
                                                              solar-system barycentre, units of c dt),
                                                              then phir, then epsm
  range file                                                 8 ints: spndr0 spndr1 nr0 nr1 mr0 mr1 pmr0 pmr1
  spotlight file                                             pm nsky, then per sky: mm nn noss ss_1..ss_noss
  <dtaprefix>/O1LinesToBeCleaned_<IFO>_v1.txt               known lines (7 columns, '%' comments)

Output: triggers_<ident %03d>_<band %04d><label>_<pm>.bin, records of 5 doubles
(jobcore.c:92-93,141), written by pgw_search.

`load_frame` builds the same segment object the tests and the benchmark use (settings, grid
matrix, detector arrays), with the settings, noise variances and amplitude-modulation
coefficients from the library's own C entry points.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

RAD_TO_DEG = 180.0 / math.pi
C_OMEGA_R = 7.2921151467064e-5           # search/network/src-cpu/settings.h:31
C_SIDDAY = 2.0 * math.pi / C_OMEGA_R     # settings.h:32

# Detector sites: geographical latitude, longitude and orientation gamma [rad]
# (search/network/src-cpu/settings.c:174-232)
DETECTORS = {
    "H1": ((46 + (27 + 18.528 / 60.0) / 60.0) / RAD_TO_DEG,
           -(119 + (24 + 27.5657 / 60.0) / 60.0) / RAD_TO_DEG,
           170.9994 / RAD_TO_DEG),
    "L1": ((30 + (33 + 46.4196 / 60.0) / 60.0) / RAD_TO_DEG,
           -(90 + (46 + 27.2654 / 60.0) / 60.0) / RAD_TO_DEG,
           242.7165 / RAD_TO_DEG),
    "V1": ((43.0 + 37.0 / 60.0 + 53.0880 / 3600.0) / RAD_TO_DEG,
           (10.0 + 30.0 / 60.0 + 16.1885 / 3600.0) / RAD_TO_DEG,
           (135.0 - (19.0 + 25.0 / 60.0 + 57.96 / 3600.0)) / RAD_TO_DEG),
}


def _read_exact(path, dtype, count: int) -> np.ndarray:
    a = np.fromfile(path, dtype=dtype, count=count)
    if len(a) != count:
        raise ValueError(f"{path}: {len(a)} values, expected {count}")
    return a


def read_grid(path):
    """grid.bin (init.c:306-315): (fftpad, M[16])."""
    with open(path, "rb") as f:
        head = f.read(4 + 16 * 8)
    if len(head) != 4 + 16 * 8:
        raise ValueError(f"{path}: short grid file")
    fftpad = int(np.frombuffer(head[:4], dtype=np.int32)[0])
    M = np.frombuffer(head[4:], dtype=np.float64).copy()
    return fftpad, M


def read_xdat(path, N: int) -> np.ndarray:
    """xdatc_*.bin (init.c:349-352): the first N doubles."""
    return _read_exact(path, np.float64, N)


def read_detssb(path, N: int):
    """DetSSB.bin (init.c:381-397): (DetSSB[3N], phir, epsm)."""
    a = _read_exact(path, np.float64, 3 * N + 2)
    return a[:3 * N].copy(), float(a[3 * N]), float(a[3 * N + 1])


def read_range(path) -> dict:
    """Range file (init.c:679-683): 8 integers spndr0 spndr1 nr0 nr1 mr0 mr1 pmr0 pmr1."""
    v = [int(x) for x in Path(path).read_text().split()[:8]]
    if len(v) != 8:
        raise ValueError(f"{path}: expected 8 integers")
    return dict(spndr=v[0:2], nr=v[2:4], mr=v[4:6], pmr=v[6:8])


def read_spotlight(path):
    """Spotlight range file (search/spotlight/src/init.c:392-442): "pm nskypos", then per sky
    position "mm nn noss ss_1 .. ss_noss".  Returns (pm, [(mm, nn, [ss, ...]), ...])."""
    v = [int(x) for x in Path(path).read_text().split()]
    if len(v) < 2:
        raise ValueError(f"{path}: missing header")
    pm, nsky, k, sky = v[0], v[1], 2, []
    while k < len(v):
        if k + 3 > len(v):
            raise ValueError(f"{path}: truncated sky position {len(sky)}")
        mm, nn, noss = v[k:k + 3]
        if noss < 0 or k + 3 + noss > len(v):
            raise ValueError(f"{path}: bad spindown count at sky position {len(sky)}")
        sky.append((mm, nn, v[k + 3:k + 3 + noss]))
        k += 3 + noss
    if len(sky) != nsky:
        raise ValueError(f"{path}: {len(sky)} sky positions, header says {nsky}")
    return pm, sky


FDOT_MAX = 0.5e-8  # line-broadening spindown limit of read_lines (settings.c:434)


def read_lines(path, N: int, dt: float, DetSSB) -> list:
    """Known lines of one detector (read_lines, settings.c:326-533) as [left, right] in Hz.

    File columns: spacing or frequency (Hz), type (0 singlet, 1 comb with fixed width, 2 comb with
    scaling width), offset of the first harmonic (Hz), first and last harmonic index, left and
    right width (Hz); lines starting with '%' are comments.  Every width is replaced by the
    broadening of the barycentring, 2 fdotmax (N dt + |max d(t E)/dt|) + f |max dE|, from the
    detector's ephemeris (scaled by the harmonic index for type 2).  A missing file gives no lines;
    blank lines are skipped."""
    try:
        text = Path(path).read_text()
    except FileNotFoundError:
        return []
    rows = [[float(t) for t in ln.replace("\t", " ").split()] for ln in text.splitlines()
            if ln.strip() and not ln.startswith("%")]
    E = np.asarray(DetSSB, dtype=np.float64)[:3 * N].reshape(N, 3)
    dE = np.abs(np.diff(E, axis=0)).max(axis=0)
    i = np.arange(N, dtype=np.float64)[:, None]
    dtE = (np.abs(np.diff(E * i, axis=0)) * dt).max(axis=0)
    norm_dtE, norm_dE = math.sqrt(float(np.sum(dtE ** 2))), math.sqrt(float(np.sum(dE ** 2)))

    def width(f):
        return 2.0 * FDOT_MAX * (N * dt + norm_dtE) + f * norm_dE

    out = []
    for r in rows:
        kind = int(r[1])
        if kind == 0:
            out.append([r[0] - width(r[0]), r[0] + width(r[0])])
        elif kind in (1, 2):
            for k in range(int(r[3]), int(r[4]) + 1):
                f = r[2] + k * r[0]
                w = width(f) * (k if kind == 2 else 1)
                out.append([f - w, f + w])
    return out


def narrow_down_lines(narrowdown: float) -> list:
    """Band edges excluded by -narrowdown (narrow_down_band, settings.c:535-550), in radians."""
    return [[0.0, math.pi / 2 - narrowdown], [math.pi / 2 + narrowdown, math.pi]]


def lines_in_band(fpo: float, B: float, det_lines, start=()) -> list:
    """Known lines overlapping [fpo, fpo + B], clipped and mapped to [0, pi] (lines_in_band,
    settings.c:552-597), after the entries of `start` (the narrow-down ranges)."""
    out = [list(x) for x in start]
    bs, be = fpo, fpo + B
    for lines in det_lines:
        for l, r in lines:
            if not (r < bs or l > be):
                ll, lr = max(l, bs), min(r, be)
                out.append([(ll - bs) / B * math.pi, (lr - bs) / B * math.pi])
    return out


def veto_fraction(lines) -> float:
    """Fraction of the band [0, pi] covered by the lines (fraction_of_band_vetoed,
    settings.c:625-650: sorted, overlaps with the previous range counted once)."""
    ls = sorted((float(a), float(b)) for a, b in lines)
    tot = 0.0
    for i, (a, b) in enumerate(ls):
        tot += (b - ls[i - 1][1]) if i > 0 and a < ls[i - 1][1] else (b - a)
    return tot / math.pi


def band_fully_vetoed(lines) -> bool:
    """True when the sorted ranges leave no gap in [0, pi] (check_if_band_is_fully_vetoed,
    settings.c:599-623; the reference then ends the run successfully)."""
    ll = 0.0
    for a, b in sorted((float(a), float(b)) for a, b in lines):
        if a > ll:
            return False
        ll = b
    return not ll < math.pi


MAXL = 2048  # search/network/src-cpu/struct.h:13


def band_lines(frame: "Frame", dtaprefix, narrowdown: float = 0.5 * math.pi, veto: bool = True) -> list:
    """The vetoed ranges of a frame's band, in band radians, the way main() assembles them
    (search/network/src-cpu/main.c:83-111): the narrow-down edges when narrowdown < pi/2, then,
    with veto, every detector's known lines (<dtaprefix>/O1LinesToBeCleaned_<IFO>_v1.txt) inside
    the band.
    The result goes to Handle(..., lines=...).  More than MAXL ranges raise ValueError."""
    s = frame.sett
    start = narrow_down_lines(narrowdown) if narrowdown < 0.5 * math.pi else []
    det_lines = [] if not veto else [read_lines(Path(dtaprefix) / f"O1LinesToBeCleaned_{d}_v1.txt", s["N"], s["dt"], e)
                 for d, e in zip(frame.dets, frame.DetSSB)]
    out = lines_in_band(s["fpo"], s["B"], det_lines, start) if veto else start
    if len(out) > MAXL:
        raise ValueError(f"{len(out)} vetoed ranges, at most {MAXL}")
    return out


def read_triggers(path, dtype=np.float64) -> np.ndarray:
    """triggers_*.bin: (n, 5) records (frequency, spindown, delta, alpha, SNR); dtype float32 for
    the files of a COMP_FLOAT build (FLOAT_TYPE = float, search/network/src-cpu/auxi.h:16-23)."""
    a = np.fromfile(path, dtype=dtype)
    if len(a) % 5:
        raise ValueError(f"{path}: not a whole number of 5-value records")
    return a.reshape(-1, 5)


def fpo_from_band(band: int, dt: float) -> float:
    """Reference band frequency when -fpo is not given (init.c:232-240)."""
    return 10.0 + 0.96875 * band * (0.5 / dt)


def samples(nod: int, dt: float) -> int:
    """N = round(nod C_SIDDAY / dt) with the reference's round (settings.c:32, auxi.h:14)."""
    return int(math.floor(nod * C_SIDDAY / dt + 0.5))


@dataclass
class Frame:
    """One frame's inputs; the fields of the synthetic Segment, so the same ABI builders apply."""
    sett: dict
    M: np.ndarray
    dets: list
    xDat: list
    DetSSB: list
    phir: list
    epsm: float
    amod: list
    sig2: list
    injection: dict = field(default_factory=dict)
    range: dict = field(default_factory=dict)


def _settings_dict(s) -> dict:
    keys = ("fpo", "dt", "B", "oms", "omr", "Smin", "Smax", "nfft", "N", "nfftf", "nmax", "nmin", "nd",
            "interpftpad", "fftpad", "Ninterp")
    return {k: getattr(s, k) for k in keys}


def load_frame(dtaprefix, ident: int, band: int, dt: float = 2.0, nod: int = 2, label: str = "",
               fpo: float | None = None, dets=None, range_file=None, addsig=None) -> Frame:
    """Reads a frame the way main() does (read_grid, search_settings, detectors_settings,
    init_arrays, rogcvir, then add_signal or set_search_range; search/network/src-cpu/main.c:72-121).
    dets: detector names in order (default: the subdirectories holding this band's data, sorted).
    addsig: a signal file (the -addsig option): the signal is added to the data after the noise
    variances are taken and the range is the box around it; range_file is then not read."""
    from . import pgw
    base = Path(dtaprefix) / f"{ident:03d}"
    xname = f"xdatc_{ident:03d}_{band:04d}{label}.bin"
    if dets is None:
        dets = sorted(d for d in os.listdir(base) if (base / d / xname).is_file())
    if not dets:
        raise FileNotFoundError(f"no detector data {xname} under {base}")
    for d in dets:
        if d not in DETECTORS:
            raise ValueError(f"unknown detector {d}")
    fftpad, M = read_grid(base / "grid.bin")
    N = samples(nod, dt)
    s = pgw.search_settings(fpo if fpo is not None else fpo_from_band(band, dt), dt, N, fftpad)
    xs, ssb, phir, amod, sig2 = [], [], [], [], []
    epsm = None
    for d in dets:
        x = read_xdat(base / d / xname, N)
        e, ph, ep = read_detssb(base / d / "DetSSB.bin", N)
        if epsm is not None and math.sin(ep) != math.sin(epsm):  # init.c:430-436
            raise ValueError(f"epsm of {d} differs from the other detectors")
        epsm = ep
        xs.append(x)
        ssb.append(e)
        phir.append(ph)
        amod.append(pgw.rogcvir(DETECTORS[d][0], DETECTORS[d][2]))
        sig2.append(pgw.sig2(x))
    rng = read_range(range_file) if range_file and not addsig else {}
    fr = Frame(sett=_settings_dict(s), M=M, dets=list(dets), xDat=xs, DetSSB=ssb, phir=phir, epsm=epsm,
               amod=amod, sig2=sig2, range=rng)
    if addsig:
        from . import inject
        inject.add_signal(fr, inject.read_signal(addsig), ident)
    return fr
