"""The full-size parity cases: the measured configurations' inputs at their real shapes.

  plain-full      the plain configuration in full: 1 detector, N=86400, L=2^17, 4 sky x 32 spindowns
  network-full    the network configuration in full: 2 detectors, N=86400, L=2^18, 256 sky x 512
  precision-8sky  the precision configuration's first 8 sky points x 256 spindowns (N=259200, L=2^19)
  batching-64seg  every one of the batching configuration's 64 segments (3 detectors, N in
                  [60000, 86400], gaps, L=2^16 or 2^17) at its injection's sky point x 512 spindowns
"""
import functools

import pgwload

pgwload.load()
from polgraw_allsky_amd import abi, synth  # noqa: E402

FULL_CASES = ["plain-full", "network-full", "precision-8sky", "batching-64seg"]
SEED = 2024


@functools.lru_cache(maxsize=None)
def case_segments(name):
    if name == "batching-64seg":
        return tuple(synth.batching_segments(seed=SEED))
    return (synth.workload_segment(name.split("-")[0], seed=SEED),)


def case_items(name):
    """(segments, [(segment, pm, mm, nn, range)]) in the order the records are appended."""
    segs = case_segments(name)
    items = []
    if name == "batching-64seg":
        for k, sg in enumerate(segs):
            g = sg.injection["grid"]
            items.append((k, sg.injection["pm"], int(g[3]), int(g[2]), abi.make_range(sg)))
    else:
        sg = segs[0]
        rng = abi.make_range(sg)
        sky = synth.sky_list(sg)
        if name == "precision-8sky":
            sky = sky[:8]
        items = [(0, pm, mm, nn, rng) for pm, mm, nn in sky]
    return list(segs), items
