"""Batching form: segments of uneven length in one handle, templates of many sky points per
launch, and the autotuner; every record checked against the CPU oracle run per segment."""
import os

import numpy as np
import pytest

from polgraw_allsky_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


def _segments():
    # three detectors, gaps 5-30 %, N spanning both nfft = 2^16 and 2^17
    out = []
    for k, (N, gap) in enumerate([(60000, 0.05), (86400, 0.3), (64000, 0.15), (70000, 0.1)]):
        out.append(synth.make_segment(N=N, dets=("H1", "L1", "V1"), sigmas=(1.0, 1.3, 1.1), fftpad=1,
                                      seed=100 + k, h0=0.15, gap_frac=gap, nsky=(2, 2), nspin=6))
    return out


def test_multi_segment_batch_equals_oracle(pgw, abi, orc):
    segs = _segments()
    assert {s.sett["nfft"] for s in segs} == {65536, 131072}
    opts = abi.make_opts(trl=15.0)
    h = pgw.Handle(segs, abi.PREC_FP64)
    work, expect, fn_r = [], [], 0
    for k, seg in enumerate(segs):
        for pm, mm, nn in synth.sky_list(seg):
            work.append((k, pm, mm, nn))
    # interleave segments so groups have to split on nfft changes
    work = work[::2] + work[1::2]
    # one shared range (the spindown window of segment 0) for all items
    rng = abi.make_range(segs[0])
    oracles = [orc.Search(s, nthreads=8) for s in segs]
    for i, (k, pm, mm, nn) in enumerate(work):
        r, f = oracles[k].job_core(pm, mm, nn, rng, opts)
        expect.append((i, r))
        fn_r += f
    got, owner, fn_g = h.job_batch_multi(work, rng, opts)
    assert fn_g == fn_r
    ref = np.concatenate([r for _, r in expect])
    ref_owner = np.concatenate([np.full(len(r), i) for i, r in expect])
    print("work items", len(work), "templates", fn_r, "records", len(ref), flush=True)
    assert got.shape == ref.shape
    assert np.array_equal(owner, ref_owner)
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_keeps_results(pgw, abi, prec):
    from cases import injected_sky, segment
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=12.0)
    h = pgw.Handle(seg, abi.PREC_FP32 if prec == "fp32" else abi.PREC_FP64)
    sky = synth.sky_list(seg)
    before, fa = h.job_batch(sky, rng, opts)
    choice = h.autotune(0, *injected_sky(seg), opts)
    print("autotune", choice, flush=True)
    assert choice["tb"] in (64, 128, 256) and choice["row_na"] in (1, 2, 8) and choice["col_seq"] in (0, 1)
    after, fb = h.job_batch(sky, rng, opts)
    # the variants compute the same sums; the compiler's FMA contraction and the FStat partial sums
    # order differ between them, so values move by rounding only: same templates and bins, F within
    # 1e-6 relative
    assert fa == fb and before.shape == after.shape
    assert np.array_equal(before[:, :4], after[:, :4])
    Fa, Fb = before[:, 4] ** 2 / 2 + 2, after[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fa - Fb) / Fa) <= 1e-6


def test_kernel_times_cover_every_launch(pgw, abi):
    """pgw_kernel_times: events around every template launch (or every k-th) of a search; totals
    reset per call."""
    from cases import segment
    seg = segment("net-2det-32k")
    os.environ["PGW_TUNE"] = "tb=16"
    try:
        h = pgw.Handle(seg, abi.PREC_FP32)
    finally:
        del os.environ["PGW_TUNE"]
    sky = synth.sky_list(seg)
    rng, opts = abi.make_range(seg), abi.make_opts(trl=12.0)
    assert h.kernel_times(1)["launches"] == 0
    _, fn = h.job_batch(sky, rng, opts)
    kt = h.kernel_times(3)
    print(kt, flush=True)
    assert kt["templates"] == fn and kt["launches"] >= 3
    assert kt["col_ms"] > 0 and kt["row_ms"] > 0 and kt["thr_ms"] > 0
    h.job_batch(sky, rng, opts)  # every third launch
    k3 = h.kernel_times(0)
    assert k3["launches"] == (kt["launches"] + 2) // 3
    h.job_batch(sky, rng, opts)  # recording off
    assert h.kernel_times(0)["launches"] == 0
