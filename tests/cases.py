"""Seeded segments shared by the parity tests (sizes the oracle finishes in seconds)."""
import functools

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
