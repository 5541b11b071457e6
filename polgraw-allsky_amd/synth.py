"""Seeded synthetic inputs for the coherent F-statistic search core.

Everything a segment needs, generated in memory instead of read from the
reference's files (``xdatc_*.bin``, ``DetSSB.bin``, ``grid.bin``):

* white Gaussian noise per detector (numpy PCG64 generator, seeded);
* a circular-orbit + sidereal-rotation ephemeris in the ``DetSSB`` convention
  (detector position w.r.t. the solar-system barycentre in units of c*dt,
  equatorial frame; gen2day/get_barycenter.c:51-84, gen2day/ExtractBand.c:336-342);
* a grid-generating matrix ``M`` with the zero pattern of the reference's
  grid files (M1=M2=M3=M6=M7=M11=0), M0 = 2 pi / (fftpad*nfft), the other
  entries scaled from testdata/test-gpu-cpu/001/grid.bin to the segment length;
* an injected signal following add_signal() (search/network/src-cpu/init.c:562-642)
  with amplitudes from sigen.c:348-357.

This is input generation, not part of the search path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .io import DETECTORS

RAD_TO_DEG = 180.0 / math.pi
C_OMEGA_R = 7.2921151467064e-5          # settings.h:31
C_EPSMA = 84381.448 / 3600.0 / RAD_TO_DEG  # settings.h:20
C_SPEED_OF_LIGHT = 299792.458           # km/s
C_AU = 1.49597870691e8                  # km
C_EARTH_R = 6378.140                    # km
C_YEARSEC = 365.25 * 86400.0

# Detector sites: latitude, longitude, orientation gamma [rad] (settings.c:174-232)
SITES = DETECTORS

# grid.bin of testdata/test-gpu-cpu/001 (N=86164, nfft=2^17, fftpad=1), column-major
_FIXTURE_N = 86164
_FIXTURE_M = {
    0: 4.79368996e-05, 4: 1.06749473e-04, 8: 8.71359861e-03, 12: -1.03495650e-02,
    5: -1.07995468e-09, 9: 6.07898667e-10, 13: 1.30440156e-09,
    10: -8.96932636e+01, 14: 8.72851497e+01, 15: 1.11152781e+02,
}


def settings_dict(N: int, dt: float = 2.0, fftpad: int = 1, fpo: float | None = None,
                  band: int = 666) -> dict:
    """Sizes from search_settings() (settings.c:16-81) for a segment of N samples."""
    if fpo is None:
        fpo = 10.0 + 0.96875 * band * (0.5 / dt)  # init.c:240
    B = 0.5 / dt
    nfft = 1 << int(math.ceil(math.log(N) / math.log(2.0)))
    return dict(
        N=N, dt=dt, fpo=fpo, B=B, oms=2.0 * math.pi * fpo * dt, omr=C_OMEGA_R * dt,
        nfft=nfft, fftpad=fftpad, nfftf=fftpad * nfft, interpftpad=2, Ninterp=2 * nfft,
        nmin=int(fftpad * 4096 * B), nmax=int((nfft // 2 - 4096 * B) * fftpad), nd=2,
        Smax=2.0 * math.pi * 0.5e-8 * dt * dt, Smin=2.0 * math.pi * 0.5e-9 * dt * dt,
    )


def grid_matrix(N: int, nfft: int, fftpad: int) -> np.ndarray:
    """Synthetic grid-generating matrix (16 doubles, column-major like grid.bin)."""
    r = _FIXTURE_N / float(N)
    M = np.zeros(16)
    M[0] = 2.0 * math.pi / (fftpad * nfft)
    for k in (4, 8, 12):
        M[k] = _FIXTURE_M[k] / _FIXTURE_M[0] * M[0]
    for k in (5, 9, 13):
        M[k] = _FIXTURE_M[k] * r * r
    for k in (10, 14, 15):
        M[k] = _FIXTURE_M[k] * r
    return M


def rogcvir(ephi: float, egam: float) -> np.ndarray:
    """Amplitude-modulation coefficients c1..c9 (settings.c:253-275, JKS98 Eqs. 12-13)."""
    s2g, c2g = math.sin(2 * egam), math.cos(2 * egam)
    sp, cp = math.sin(ephi), math.cos(ephi)
    return np.array([
        0.25 * s2g * (1 + sp * sp), -0.5 * c2g * sp, 0.5 * s2g * math.sin(2 * ephi),
        -c2g * cp, 0.75 * s2g * cp * cp, c2g * sp,
        0.5 * s2g * (1.0 + sp * sp), c2g * cp, 0.5 * s2g * math.sin(2 * ephi)])


def modvir(sinal, cosal, sindel, cosdel, sphir, cphir, c, cosmodf, sinmodf):
    """a(t), b(t) of JKS98 (settings.c:281-325), vectorised."""
    cosalfr = cosal * cphir + sinal * sphir
    sinalfr = sinal * cphir - cosal * sphir
    c2d = cosdel * cosdel
    c2sd = sindel * cosdel
    cc = cosalfr * cosmodf + sinalfr * sinmodf
    ss = sinalfr * cosmodf - cosalfr * sinmodf
    c2s = 2.0 * cc * cc
    cs = cc * ss
    aa = (c[0] * (2.0 - c2d) * c2s + c[1] * (2.0 - c2d) * 2.0 * cs + c[2] * c2sd * cc
          + c[3] * c2sd * ss - c[0] * (2.0 - c2d) + c[4] * c2d)
    bb = (c[5] * sindel * c2s + c[6] * sindel * 2.0 * cs + c[7] * cosdel * cc
          + c[8] * cosdel * ss - c[5] * sindel)
    return aa, bb


def ephemeris(N: int, dt: float, site: str, orbit_phase: float, phir: float) -> np.ndarray:
    """DetSSB (3N doubles, samples): circular orbit in the ecliptic rotated by the
    obliquity into equatorial coordinates, plus the detector's diurnal motion."""
    lat = SITES[site][0]
    t = np.arange(N, dtype=np.float64) * dt
    r_orb = C_AU / (C_SPEED_OF_LIGHT * dt)
    r_e = C_EARTH_R / (C_SPEED_OF_LIGHT * dt)
    lam = orbit_phase + 2.0 * math.pi * t / C_YEARSEC
    x_ecl, y_ecl = r_orb * np.cos(lam), r_orb * np.sin(lam)
    ce, se = math.cos(C_EPSMA), math.sin(C_EPSMA)
    lst = phir + C_OMEGA_R * t
    out = np.empty((N, 3))
    out[:, 0] = x_ecl + r_e * math.cos(lat) * np.cos(lst)
    out[:, 1] = y_ecl * ce + r_e * math.cos(lat) * np.sin(lst)
    out[:, 2] = y_ecl * se + r_e * math.sin(lat)
    return out.reshape(-1)


def ast2lin(alfa: float, delta: float, epsm: float):
    """Hemisphere and linear sky coordinates (auxi.c:26-53)."""
    be0 = math.cos(epsm) * math.sin(alfa) * math.cos(delta) + math.sin(epsm) * math.sin(delta)
    be1 = math.cos(alfa) * math.cos(delta)
    d1 = math.asin(be0 * math.sin(epsm) + math.sqrt(max(0.0, 1.0 - be0 * be0 - be1 * be1)) * math.cos(epsm)) - delta
    pm = 1 if abs(d1) < 10.0 * np.finfo(float).eps else 2
    return pm, (be0, be1)


def amplitudes(ph_o: float, psik: float, hoc: float) -> np.ndarray:
    """The four amplitudes of a signal of initial phase ph_o, polarisation angle psik and
    cos(inclination) hoc (sigen.c:348-357; Phys. Rev. D 82, 022005, Eqs. 2.13a-d)."""
    hop = (1.0 + hoc * hoc) / 2.0
    c2p, s2p, cph, sph = math.cos(2 * psik), math.sin(2 * psik), math.cos(ph_o), math.sin(ph_o)
    return np.array([c2p * hop * cph - s2p * hoc * sph, s2p * hop * cph + c2p * hoc * sph,
                     -c2p * hop * sph - s2p * hoc * cph, -s2p * hop * sph + c2p * hoc * cph])


def lin2ast(be1: float, be2: float, pm: int, epsm: float = C_EPSMA):
    """(delta, alpha) of linear sky coordinates in hemisphere pm (auxi.c:15-24, with the
    declination and right ascension of jobcore.c:261-262)."""
    sepsm, cepsm = math.sin(epsm), math.cos(epsm)
    sindel = be1 * sepsm - (2 * pm - 3) * math.sqrt(1.0 - be1 * be1 - be2 * be2) * cepsm
    cosdel = math.sqrt(1.0 - sindel * sindel)
    sinal = (be1 - sepsm * sindel) / (cepsm * cosdel)
    cosal = be2 / cosdel
    return math.asin(sindel), math.fmod(math.atan2(sinal, cosal) + 2 * math.pi, 2 * math.pi)


@dataclass
class Segment:
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
    injections: list = field(default_factory=list)


def _sig2(x: np.ndarray) -> float:
    """crf0 * var(x) (init.c:359-370, auxi.c:124-139)."""
    N = len(x)
    nz = int(np.count_nonzero(x == 0.0))
    crf0 = N / (N - nz)
    mean = np.sum(x) / N
    return crf0 * float(np.sum((x - mean) ** 2) / (N - 1))


def make_segment(N: int = 86400, dets=("H1",), sigmas=(1.0,), dt: float = 2.0, fftpad: int = 1,
                 seed: int = 1, h0: float = 0.1, inject: bool = True, gap_frac: float = 0.0,
                 nsky: tuple = (16, 16), nspin: int = 512, band: int = 666) -> Segment:
    """One seeded segment for a network of detectors, with one injection and a
    search range of nsky sky points x nspin spindowns centred on it."""
    rng = np.random.default_rng(seed)
    sett = settings_dict(N, dt=dt, fftpad=fftpad, band=band)
    M = grid_matrix(N, sett["nfft"], fftpad)
    orbit_phase = rng.uniform(0, 2 * math.pi)
    phir0 = rng.uniform(0, 2 * math.pi)
    t = np.arange(N, dtype=np.float64)
    cosmodf, sinmodf = np.cos(sett["omr"] * t), np.sin(sett["omr"] * t)

    xs, ssbs, phirs, amods = [], [], [], []
    for k, d in enumerate(dets):
        lat, lon, gam = SITES[d]
        phir = (phir0 + lon) % (2 * math.pi)   # local sidereal angle at t=0
        x = rng.standard_normal(N) * sigmas[k]
        if gap_frac > 0:
            # zeroed gaps at seeded random positions covering gap_frac of the samples
            ngap = max(1, int(rng.integers(3, 9)))
            total = int(gap_frac * N)
            lens = rng.multinomial(total, np.ones(ngap) / ngap)
            starts = np.sort(rng.choice(N - total, size=ngap, replace=False))
            off = 0
            for s0, ln in zip(starts, lens):
                x[s0 + off:s0 + off + ln] = 0.0
                off += ln
        xs.append(x)
        ssbs.append(ephemeris(N, dt, d, orbit_phase, phir))
        phirs.append(phir)
        amods.append(rogcvir(lat, gam))

    inj = {}
    rng_s = {}
    if inject:
        f = rng.uniform(0.1 * math.pi, 0.9 * math.pi)
        fdot = -1e-8 * rng.uniform(0.0, 1.0)
        Mmat = M.reshape(4, 4).T
        oms = sett["oms"]
        for _ in range(1000):
            # uniform sky, redrawn until the whole search patch lies inside the grid disk
            sindel = rng.uniform(-1.0, 1.0)
            delta = math.asin(sindel)
            alpha = rng.uniform(0.0, 2 * math.pi)
            _, be = ast2lin(alpha, delta, C_EPSMA)
            g = np.linalg.solve(Mmat, np.array([f, fdot, be[0] * (oms + f), be[1] * (oms + f)]))
            nn0, mm0 = int(round(g[2])), int(round(g[3]))
            ok = True
            for dn in (-(nsky[0] // 2), nsky[0] - nsky[0] // 2 - 1):
                for dm in (-(nsky[1] // 2), nsky[1] - nsky[1] // 2 - 1):
                    a1 = (nn0 + dn) * M[10] + (mm0 + dm) * M[14]
                    a2 = (nn0 + dn) * M[11] + (mm0 + dm) * M[15]
                    ok &= (a1 * a1 + a2 * a2) / (oms * oms) <= 0.95
            if ok:
                break
        ph_o, psik = rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)
        amp = amplitudes(ph_o, psik, rng.uniform(-1.0, 1.0))
        cof = sett["oms"] + f
        sa, ca, sd, cd = math.sin(alpha), math.cos(alpha), math.sin(delta), math.cos(delta)
        nS = np.array([ca * cd, sa * cd, sd])
        for k in range(len(dets)):
            aa, bb = modvir(sa, ca, sd, cd, math.sin(phirs[k]), math.cos(phirs[k]), amods[k],
                            cosmodf, sinmodf)
            shift = ssbs[k].reshape(-1, 3) @ nS
            phase = f * t + fdot * t * t + (cof + 2.0 * fdot * t) * shift
            sig = (amp[0] * aa * np.cos(phase) + amp[2] * aa * np.sin(phase)
                   + amp[1] * bb * np.cos(phase) + amp[3] * bb * np.sin(phase))
            nzm = xs[k] != 0.0
            xs[k][nzm] += h0 * sig[nzm]
        pm, be = ast2lin(alpha, delta, C_EPSMA)
        g = np.linalg.solve(Mmat, np.array([f, fdot, be[0] * cof, be[1] * cof]))
        inj = dict(f=f, fdot=fdot, delta=delta, alpha=alpha, amp=amp, h0=h0, pm=pm,
                   grid=np.round(g).astype(int))
        ss0, nn0, mm0 = inj["grid"][1], inj["grid"][2], inj["grid"][3]
    else:
        pm, ss0, nn0, mm0 = 1, 0, 0, 0
    rng_s = dict(spndr=[int(ss0 - nspin // 2), int(ss0 - nspin // 2 + nspin - 1)],
                 nr=[int(nn0 - nsky[0] // 2), int(nn0 - nsky[0] // 2 + nsky[0] - 1)],
                 mr=[int(mm0 - nsky[1] // 2), int(mm0 - nsky[1] // 2 + nsky[1] - 1)],
                 pmr=[int(pm), int(pm)])
    sig2 = [_sig2(x) for x in xs]
    return Segment(sett=sett, M=M, dets=list(dets), xDat=xs, DetSSB=ssbs, phir=phirs,
                   epsm=C_EPSMA, amod=amods, sig2=sig2, injection=inj, range=rng_s)


def make_injections(N: int = 86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), dt: float = 2.0, fftpad: int = 2,
                    seed: int = 1, nsig: int = 64, h0_range=(0.05, 0.3), band: int = 666) -> Segment:
    """One seeded noise segment carrying nsig injected signals (add_signal(),
    search/network/src-cpu/init.c:562-642, summed).  Signal k's frequency is drawn from the middle
    half of the k-th of nsig equal strata of [0.1 pi, 0.9 pi], so neighbours are at least half a
    stratum apart (~270 Fourier bins for 64 signals at N=86400), far beyond a signal's Doppler and
    spindown spread (< 20 bins): each candidate's F sees its own signal.  Spindown
    U[-1e-8, 0] rad/sample^2, uniform sky, psi and phi0 U[0, 2 pi), cos(iota) U[-1, 1],
    h0 U[h0_range] (units of the first detector's sigma).  seg.injections lists the signals
    (f, fdot, delta, alpha, amp, h0, pm) in frequency order."""
    rng = np.random.default_rng(seed)
    base = make_segment(N=N, dets=dets, sigmas=sigmas, dt=dt, fftpad=fftpad, seed=seed, inject=False,
                        nsky=(1, 1), nspin=1, band=band)
    sett = base.sett
    t = np.arange(N, dtype=np.float64)
    cosmodf, sinmodf = np.cos(sett["omr"] * t), np.sin(sett["omr"] * t)
    width = 0.8 * math.pi / nsig
    cof0 = sett["oms"]
    injs = []
    for k in range(nsig):
        f = 0.1 * math.pi + width * (k + rng.uniform(0.25, 0.75))
        fdot = -1e-8 * rng.uniform(0.0, 1.0)
        delta = math.asin(rng.uniform(-1.0, 1.0))
        alpha = rng.uniform(0.0, 2 * math.pi)
        ph_o, psik = rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)
        amp = amplitudes(ph_o, psik, rng.uniform(-1.0, 1.0))
        h0 = rng.uniform(*h0_range) * sigmas[0]
        pm, _ = ast2lin(alpha, delta, C_EPSMA)
        injs.append(dict(f=f, fdot=fdot, delta=delta, alpha=alpha, amp=amp, h0=h0, pm=pm))
    for d in range(len(dets)):
        sig = np.zeros(N)
        for j in injs:
            sa, ca, sd, cd = math.sin(j["alpha"]), math.cos(j["alpha"]), math.sin(j["delta"]), math.cos(j["delta"])
            aa, bb = modvir(sa, ca, sd, cd, math.sin(base.phir[d]), math.cos(base.phir[d]), base.amod[d],
                            cosmodf, sinmodf)
            shift = base.DetSSB[d].reshape(-1, 3) @ np.array([ca * cd, sa * cd, sd])
            phase = j["f"] * t + j["fdot"] * t * t + (cof0 + j["f"] + 2.0 * j["fdot"] * t) * shift
            cph, sph = np.cos(phase), np.sin(phase)
            a = j["amp"]
            sig += j["h0"] * (a[0] * aa * cph + a[2] * aa * sph + a[1] * bb * cph + a[3] * bb * sph)
        nzm = base.xDat[d] != 0.0
        base.xDat[d][nzm] += sig[nzm]
    base.sig2 = [_sig2(x) for x in base.xDat]
    base.injection = {}
    base.injections = injs
    return base


def sky_list(seg: Segment, max_points: int | None = None):
    """(pm, mm, nn) in the reference's loop order (jobcore.c:90-98) over seg.range."""
    out = []
    r = seg.range
    for pm in range(r["pmr"][0], r["pmr"][1] + 1):
        for mm in range(r["mr"][0], r["mr"][1] + 1):
            for nn in range(r["nr"][0], r["nr"][1] + 1):
                out.append((pm, mm, nn))
    return out if max_points is None else out[:max_points]


# The measured configurations: make_segment arguments (seed added by workload_segments) and text.
WORKLOADS = {
    "network": (dict(N=86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, h0=0.1, nsky=(16, 16), nspin=512),
                "network: 2 detectors (H1, L1), N=86400, nfft=2^17, fftpad=2, L=2^18, 256 sky x 512 spindowns, trl=20"),
    "plain": (dict(N=86400, dets=("H1",), sigmas=(1.0,), fftpad=1, h0=0.3, nsky=(2, 2), nspin=32),
              "plain: 1 detector (H1), N=86400, nfft=L=2^17, fftpad=1, 4 sky x 32 spindowns, trl=20"),
    "precision": (dict(N=259200, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, h0=0.06, nsky=(32, 32), nspin=256),
                  "precision: 2 detectors (H1, L1), N=259200, nfft=2^18, fftpad=2, L=2^19, 1024 sky x 256 spindowns, "
                  "trl=20"),
}
BATCHING_TEXT = ("batching: 3 detectors (H1, L1, V1), 64 segments of N in [60000, 86400] with 5-30 % gaps, "
                 "nfft 2^16 or 2^17, fftpad=1, 32 sky x 512 spindowns per segment, trl=20")


def workload_segment(name: str, seed: int = 2024) -> Segment:
    """The segment of a single-segment configuration (plain, network, precision)."""
    return make_segment(seed=seed, **WORKLOADS[name][0])


def batching_segments(seed: int = 2024, nseg: int = 64) -> list:
    """The batching configuration: nseg segments of 3 detectors, N uniform in [60000, 86400]
    (nfft 2^16 or 2^17), zeroed gaps over a fraction uniform in [0.05, 0.3], one injection
    (h0 = 0.12) each, 4 x 8 sky points x 512 spindowns around it."""
    rs = np.random.default_rng(seed)
    segs = []
    for k in range(nseg):
        N = int(rs.integers(60000, 86401))
        segs.append(make_segment(N=N, dets=("H1", "L1", "V1"), sigmas=(1.0, 1.3, 1.1), fftpad=1,
                                 seed=seed + 1 + k, h0=0.12, gap_frac=float(rs.uniform(0.05, 0.3)),
                                 nsky=(4, 8), nspin=512))
    return segs


def differentiation_workload(seed: int = 2024, nseg: int = 16, per_seg: int = 64, mismatch: float = 0.25):
    """The differentiation configuration: nseg * per_seg injections (h0 in [0.05, 0.3]) in the
    2-detector N=86400 setup, as nseg noise realisations of H1 and L1 (sigma 1 and 1.3, fftpad 2)
    with per_seg signals each at separated frequencies (make_injections).

    Cells come from the phase metric G at the injection (followup.phase_metric): the cell along
    parameter k is 1 / sqrt(G_kk), the step that moves the phase by one radian rms, the spacing of
    a covering grid.  (The grid matrix of make_segment is scaled from the reference's test grid;
    its sky vectors do not follow this ephemeris's Doppler direction, so its cells are far wider
    than an F peak.)  The candidate point of each injection is drawn uniformly inside the ellipsoid
    d^T G d <= mismatch around the truth, where a grid search of minimal match 1 - mismatch leaves
    its best template.  Returns (segments, truth (n, 4), candidates (n, 4), cells (n, 4), segment
    of each (n,), h0 (n,))."""
    from .followup import phase_metric
    segs = [make_injections(N=86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=seed + 101 + k,
                            nsig=per_seg) for k in range(nseg)]
    rs = np.random.default_rng(seed)
    truth, starts, cells, seg_of, h0 = [], [], [], [], []
    for k, sg in enumerate(segs):
        oms, N = sg.sett["oms"], sg.sett["N"]
        for j in sg.injections:
            t = np.array([j["f"], j["fdot"], j["delta"], j["alpha"]])
            G = phase_metric(sg.DetSSB, N, oms, t)
            G = G + 1e-9 * np.trace(G) * np.eye(4)
            z = rs.standard_normal(4)
            z *= math.sqrt(mismatch) * rs.uniform() ** 0.25 / np.linalg.norm(z)
            d = np.linalg.solve(np.linalg.cholesky(G).T, z)          # d^T G d = |z|^2
            truth.append(t)
            starts.append(t + d)
            cells.append(1.0 / np.sqrt(np.diag(G)))
            seg_of.append(k)
            h0.append(j["h0"])
    return segs, np.array(truth), np.array(starts), np.array(cells), np.array(seg_of), np.array(h0)

