"""Seeded segments shared by the parity tests (sizes the oracle finishes in seconds)."""
import functools

import numpy as np

import pgwload

pgwload.load()
from polgraw_allsky_amd import synth  # noqa: E402

# name -> make_segment kwargs
CASES = {
    # one detector, fftpad 1: 3 FStat blocks + a raw tail of 2048 bins
    "plain-1det-32k": dict(N=32768, dets=("H1",), sigmas=(1.0,), fftpad=1, seed=5, h0=0.6, nsky=(3, 3), nspin=12),
    # two detectors, fftpad 2
    "net-2det-32k": dict(N=32768, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=7, h0=0.5, nsky=(3, 3),
                         nspin=12),
    # three detectors with zeroed gaps, N not a power of two
    "gaps-3det-30k": dict(N=30000, dets=("H1", "L1", "V1"), sigmas=(1.0, 1.3, 1.1), fftpad=1, seed=9, h0=0.5,
                          gap_frac=0.2, nsky=(2, 2), nspin=8),
    # the network configuration's shape (N=86400, nfft=2^17, fftpad 2, L=2^18)
    "net-2det-86k": dict(N=86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=3, h0=0.1, nsky=(2, 2),
                         nspin=16),
    # the plain configuration's shape: one detector, nfft = L = 2^17, FStat raw tail of 2048 bins
    "plain-1det-86k": dict(N=86400, dets=("H1",), sigmas=(1.0,), fftpad=1, seed=13, h0=0.15, nsky=(2, 2),
                           nspin=8),
    # the precision configuration's shape: N=259200, nfft=2^18, fftpad 2, L=2^19
    "prec-2det-259k": dict(N=259200, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=17, h0=0.06,
                           nsky=(1, 1), nspin=4),
    # the most detectors a handle takes (8; sites repeat with independent noise)
    "net-8det-16k": dict(N=16384, dets=("H1", "L1", "V1", "H1", "L1", "V1", "H1", "L1"),
                         sigmas=(1.0, 1.3, 1.1, 0.9, 1.2, 1.0, 1.4, 1.1), fftpad=2, seed=32, h0=0.4,
                         nsky=(2, 2), nspin=6),
    # smallest supported F-statistic FFT: N=12000 -> nfft = L = 2^14 (1024 x 16 split)
    "tiny-1det-12k": dict(N=12000, dets=("H1",), sigmas=(1.0,), fftpad=1, seed=23, h0=0.8, nsky=(2, 2),
                          nspin=6),
    # largest: N=500000 -> nfft = 2^19, fftpad 2, L = 2^20 (4096 x 256 split), Ninterp = 2^20
    "huge-2det-500k": dict(N=500000, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=29, h0=0.1,
                           nsky=(1, 1), nspin=3),
    # the batching configuration's shorter segments: N=60000 -> nfft = L = 2^16, three detectors, gaps
    "batch-3det-60k": dict(N=60000, dets=("H1", "L1", "V1"), sigmas=(1.0, 1.3, 1.1), fftpad=1, seed=19, h0=0.12,
                           gap_frac=0.15, nsky=(2, 2), nspin=8),
}


@functools.lru_cache(maxsize=None)
def segment(name):
    return synth.make_segment(**CASES[name])


def injected_sky(seg):
    """(pm, mm, nn) of the grid point nearest the injection."""
    g = seg.injection["grid"]
    return seg.injection["pm"], int(g[3]), int(g[2])


def phase_scales(seg, point):
    """Steps of about one radian of phase for each of (omega, omegadot, delta, alpha) at point."""
    from polgraw_allsky_amd import followup
    return followup.phase_scales(seg.DetSSB, seg.sett["N"], seg.sett["oms"], point)


def noise_free_signal(N=32768, dets=("H1", "L1"), seed=5, h0=0.05, fftpad=2, amp=None):
    """Segment whose data are a noise-free signal at off-grid parameters (sig2 = 1): the
    injection formula of init.c:600-615 with amplitudes amp.  Returns (seg, truth, amp, rho2)."""
    import math
    seg = synth.make_segment(N=N, dets=dets, sigmas=(1.0,) * len(dets), fftpad=fftpad, seed=seed,
                             inject=False, nsky=(1, 1), nspin=3)
    rng = np.random.default_rng(seed + 1000)
    s = seg.sett
    f = rng.uniform(0.2 * math.pi, 0.8 * math.pi)
    fdot = -rng.uniform(0.1, 0.9) * 1e-8
    delta = math.asin(rng.uniform(-0.9, 0.9))
    alpha = rng.uniform(0.0, 2 * math.pi)
    amp = rng.uniform(-1.0, 1.0, 4) if amp is None else np.asarray(amp, dtype=float)
    t = np.arange(N, dtype=float)
    cm, sm = np.cos(s["omr"] * t), np.sin(s["omr"] * t)
    sa, ca, sd, cd = math.sin(alpha), math.cos(alpha), math.sin(delta), math.cos(delta)
    rho2 = 0.0
    for k in range(len(dets)):
        aa, bb = synth.modvir(sa, ca, sd, cd, math.sin(seg.phir[k]), math.cos(seg.phir[k]), seg.amod[k], cm, sm)
        shift = seg.DetSSB[k].reshape(-1, 3) @ np.array([ca * cd, sa * cd, sd])
        ph = f * t + fdot * t * t + (s["oms"] + f + 2 * fdot * t) * shift
        sig = h0 * (amp[0] * aa * np.cos(ph) + amp[2] * aa * np.sin(ph) + amp[1] * bb * np.cos(ph)
                    + amp[3] * bb * np.sin(ph))
        seg.xDat[k] = sig
        seg.sig2[k] = 1.0
        rho2 += float(np.sum(sig ** 2))
    return seg, np.array([f, fdot, delta, alpha]), amp * h0, rho2


FP32_MARGIN = 1e-4  # the FP32 form's bound on relative F error against the FP64 path


def fp32_record_diff(got, ref, trl, margin=FP32_MARGIN):
    """FP32-form records `got` against FP64 records `ref` (rows of 5 doubles; F = snr^2/2 + 2).

    The rule: records are keyed by their template (fields 1-3) and frequency (field 0), which the
    host computes identically in both forms.  A record in both sets must agree in F within
    `margin` relative.  A record in one set only is allowed only where FP32 rounding of F (at most
    `margin` relative) can change the outcome of the run-max threshold:
      (a) threshold: its F lies within margin * trl of trl, so its run can appear or vanish;
      (b) run maximum: the other set has an unmatched record of the same template whose F agrees
          within `margin`, i.e. the maximum of a run moved between two bins of nearly equal F.
    Returns (max relative dF over matched records, matched count, [(record, reason)] of allowed
    exceptions, [record] violations)."""
    F = lambda snr: snr * snr / 2 + 2  # noqa: E731
    kg = {tuple(r[:4]): F(r[4]) for r in got}
    kr = {tuple(r[:4]): F(r[4]) for r in ref}
    common = set(kg) & set(kr)
    rel = max((abs(kg[k] - kr[k]) / kr[k] for k in common), default=0.0)
    viol = [k for k in common if abs(kg[k] - kr[k]) > margin * kr[k]]
    only_g = {k: kg[k] for k in set(kg) - common}
    only_r = {k: kr[k] for k in set(kr) - common}
    allowed = []
    for mine, other in ((only_g, only_r), (only_r, only_g)):
        for k, f in mine.items():
            if abs(f - trl) <= margin * trl:
                allowed.append((k, "threshold"))
            elif any(o[1:4] == k[1:4] and abs(fo - f) <= margin * max(f, fo) for o, fo in other.items()):
                allowed.append((k, "run maximum"))
            else:
                viol.append(k)
    return rel, len(common), allowed, viol
