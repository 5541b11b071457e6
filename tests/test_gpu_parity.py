"""HIP search core vs the CPU oracle on the same seeded inputs (FP64 path).

Tolerances: candidate identities (ss, bin, sky) identical, so the records'
frequency/spindown/sky fields are bit-equal; F within 1e-9 relative, measured as
|dF| / max(|F|, 1) so that the handful of bins with F near 0 (absolute FFT round-off) do
not dominate.
"""
import numpy as np
import pytest

from cases import CASES, injected_sky, segment

pytestmark = pytest.mark.gpu

FTOL = 1e-9


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))


@pytest.mark.parametrize("name", list(CASES))
def test_resampled_series(name, pgw, abi, orc):
    seg = segment(name)
    pm, mm, nn = injected_sky(seg)
    h = pgw.Handle(seg, abi.PREC_FP64)
    o = orc.Search(seg, nthreads=8)
    for det in range(len(seg.dets)):
        ga, gb = h.resampled(pm, mm, nn, det)
        ra, rb = o.resampled(pm, mm, nn, det)
        scale = max(np.abs(ra).max(), 1e-300)
        print(name, det, np.abs(ga - ra).max() / scale, np.abs(gb - rb).max() / scale, flush=True)
        assert np.abs(ga - ra).max() <= 1e-11 * scale
        assert np.abs(gb - rb).max() <= 1e-11 * np.abs(rb).max()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("white", [False, True])
def test_fstat_one_template(name, white, pgw, abi, orc):
    seg = segment(name)
    pm, mm, nn = injected_sky(seg)
    ss = int(seg.injection["grid"][1])
    opts = abi.make_opts(white=white)
    h = pgw.Handle(seg, abi.PREC_FP64)
    o = orc.Search(seg, nthreads=8)
    Fg = h.fstat(pm, mm, nn, ss, opts)
    Fr = o.fstat(pm, mm, nn, ss, opts)
    err = _rel(Fg, Fr)
    print(name, white, "max F", Fr.max(), "rel err", err, flush=True)
    assert err <= FTOL


def _compare_candidates(got, ref, fn_g, fn_r):
    assert fn_g == fn_r
    assert got.shape == ref.shape, (got.shape, ref.shape)
    if len(ref):
        assert np.array_equal(got[:, :4], ref[:, :4])
        Fg = got[:, 4] ** 2 / 2 + 2
        Fr = ref[:, 4] ** 2 / 2 + 2
        assert np.max(np.abs(Fg - Fr) / Fr) <= FTOL


@pytest.mark.parametrize("name", list(CASES))
def test_job_core_candidates(name, pgw, abi, orc):
    seg = segment(name)
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0)
    h = pgw.Handle(seg, abi.PREC_FP64)
    o = orc.Search(seg, nthreads=8)
    pm, mm, nn = injected_sky(seg)
    got, fn_g = h.job_core(pm, mm, nn, rng, opts)
    ref, fn_r = o.job_core(pm, mm, nn, rng, opts)
    print(name, "templates", fn_r, "candidates", len(ref), flush=True)
    _compare_candidates(got, ref, fn_g, fn_r)
    assert len(ref) > 0  # the injection is above threshold at its own grid point


def test_job_batch_equals_job_core_loop(pgw, abi, orc):
    from polgraw_allsky_amd import synth
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=12.0)
    h = pgw.Handle(seg, abi.PREC_FP64)
    sky = synth.sky_list(seg)
    got, fn_g = h.job_batch(sky, rng, opts)
    o = orc.Search(seg, nthreads=8)
    recs, fn_r = [], 0
    for pm, mm, nn in sky:
        r, f = o.job_core(pm, mm, nn, rng, opts)
        recs.append(r)
        fn_r += f
    ref = np.concatenate(recs) if recs else np.zeros((0, 5))
    print("batch templates", fn_r, "candidates", len(ref), flush=True)
    _compare_candidates(got, ref, fn_g, fn_r)
