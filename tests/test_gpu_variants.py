"""Kernel variants (chosen per process with PGW_TUNE, or by the autotuner) give the oracle's
records: row pass with 16 or 8 elements per thread (the latter also with the transforms' later
exchanges inside each wave, row8=3/4), with and without the half-spectrum pruning,
both spectra together, column pass with shared LDS, spline chunk lengths, the spline with its
untranspose and evaluation fused or as three kernels (spl_fuse), launch and group
sizes, thresholds folded into the next column launch or launched on their own, a device
candidate buffer smaller than a group's candidates (the group runs again into a buffer of the
needed size), FFT rows capped at 2048 or 1024 points (longer columns), groups launched as captured HIP
graphs (graph=1); for FP32 also column passes of 256 or 512 threads (col_wide), the template stage as one persistent work-queue kernel
(fused=1, lag flag), the
demodulation phase in radians (rev=0) and the loads-first column pass (col_lf=1, built for 16
register rows: with l1max=1024 the columns of 128 points have them).  FP64 to 1e-9 relative in
F; FP32 against the FP64 oracle by the rule of cases.fp32_record_diff."""
import os

import numpy as np
import pytest

from cases import fp32_record_diff, injected_sky, segment

pytestmark = pytest.mark.gpu

VARIANTS = ["", "row8=0", "row8=1", "row8=2", "half=0", "row8=0,half=0", "row8=2,half=0", "row_na=2,row8=0",
            "col_seq=1", "spl_k=8", "spl_k=16", "tb=32,group=3", "thr_fuse=0", "tb=8", "candcap=16",
            "candcap=16,group=1", "l1max=2048", "l1max=2048,row8=0", "l1max=1024", "row8=3", "row8=4",
            "row8=3,half=0", "row8=4,half=0", "row8=3,l1max=2048", "row8=3,l1max=1024", "row8=4,tb=8",
            "col_nt=1", "col_nt=1,tb=8", "col_nt=1,col_seq=0", "col_nt=1,l1max=1024", "row_nt=0",
            "row_nt=0,row8=0", "row_nt=0,row8=1", "f_nt=0", "f_nt=0,row8=0", "graph=1", "graph=1,group=1,tb=8",
            "spl_fuse=0", "spl_fuse=0,spl_k=16"]
CASES = ["plain-1det-32k", "net-2det-32k", "net-2det-86k"]


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


@pytest.fixture(scope="module")
def oracle_runs(orc, abi):
    out = {}
    for name in CASES:
        seg = segment(name)
        pm, mm, nn = injected_sky(seg)
        sky = [(pm, mm, nn), (pm, mm, nn + 1)]
        rng, opts = abi.make_range(seg), abi.make_opts(trl=12.0)
        S = orc.Search(seg, nthreads=8)
        recs, fn = [], 0
        for s in sky:
            r, f = S.job_core(*s, rng, opts)
            recs.append(r)
            fn += f
        out[name] = (sky, np.concatenate(recs), fn)
    return out


def _run(pgw, abi, seg, prec, tune, sky):
    old = os.environ.get("PGW_TUNE")
    os.environ["PGW_TUNE"] = tune
    try:
        h = pgw.Handle(seg, prec)
    finally:
        if old is None:
            del os.environ["PGW_TUNE"]
        else:
            os.environ["PGW_TUNE"] = old
    return h.job_batch(sky, abi.make_range(seg), abi.make_opts(trl=12.0))


@pytest.mark.parametrize("tune", VARIANTS)
@pytest.mark.parametrize("name", CASES)
def test_variant_fp64(pgw, abi, oracle_runs, name, tune):
    sky, ref, fr = oracle_runs[name]
    got, fg = _run(pgw, abi, segment(name), abi.PREC_FP64, tune, sky)
    print(name, repr(tune), "records", len(got), len(ref), flush=True)
    assert fg == fr and got.shape == ref.shape and len(ref) > 0
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.parametrize("tune", ["", "row8=1", "row8=2", "half=0", "row_na=2", "spl_k=8", "thr_fuse=0",
                                  "l1max=2048", "l1max=1024", "rev=0", "rev=0,tb=8",
                                  "col_lf=1", "col_lf=1,rev=0", "col_lf=1,tb=8",
                                  "col_lf=1,l1max=1024", "col_lf=1,l1max=1024,rev=0", "row8=3", "row8=4",
                                  "row8=3,half=0", "row8=3,l1max=1024", "col_nt=1", "col_nt=1,tb=8",
                                  "col_nt=1,col_lf=1,l1max=1024", "row_nt=0", "row_nt=0,row_na=2", "f_nt=0",
                                  "graph=1", "graph=1,group=1,tb=8", "fused=1", "fused=1,flag=0",
                                  "fused=1,tb=8,flag=2", "fused=1,l1max=1024", "fused=1,col_nt=1,tb=256",
                                  "fused=1,graph=1", "col_wide=0", "col_wide=0,col_nt=1,tb=8", "col_wide=1,col_seq=0",
                                  "col_wide=1,thr_fuse=0,tb=8"])
@pytest.mark.parametrize("name", CASES)
def test_variant_fp32(pgw, abi, oracle_runs, name, tune):
    sky, ref, fr = oracle_runs[name]
    got, fg = _run(pgw, abi, segment(name), abi.PREC_FP32, tune, sky)
    print(name, repr(tune), "records", len(got), len(ref), flush=True)
    assert fg == fr
    rel, nmatch, allowed, viol = fp32_record_diff(got, ref, 12.0)
    print("max rel dF", rel, "allowed", allowed, flush=True)
    assert not viol, viol[:5]
    assert nmatch > 0


@pytest.mark.parametrize("tune", ["tb=256", "tb=200", "tb=128"])
def test_long_spindown_range(pgw, abi, orc, tune):
    """300 spindowns of one sky point: launches of more templates than the kernel-argument table
    holds take het1, sgnlt1 from the descriptors; records equal the oracle's."""
    seg = segment("net-2det-32k")
    pm, mm, nn = injected_sky(seg)
    s0 = seg.range["spndr"][0] - 150
    rng, opts = abi.make_range(seg, spndr=[s0, s0 + 299]), abi.make_opts(trl=12.0)
    ref, fr = orc.Search(seg, nthreads=8).job_core(pm, mm, nn, rng, opts)
    old = os.environ.get("PGW_TUNE")
    os.environ["PGW_TUNE"] = tune
    try:
        h = pgw.Handle(seg, abi.PREC_FP64)
    finally:
        if old is None:
            del os.environ["PGW_TUNE"]
        else:
            os.environ["PGW_TUNE"] = old
    got, fg = h.job_core(pm, mm, nn, rng, opts)
    print(tune, "records", len(got), flush=True)
    assert fg == fr == 300 and got.shape == ref.shape and len(ref) > 0
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.parametrize("name", ["plain-1det-32k", "net-2det-32k"])
def test_low_threshold_runs(pgw, abi, orc, name):
    """trl = 5 on noise: several per cent of the bins exceed it, so runs of many bins cross lanes,
    register rows and FStat blocks; every run-maximum record equals the oracle's."""
    seg = segment(name)
    pm, mm, nn = injected_sky(seg)
    rng, opts = abi.make_range(seg), abi.make_opts(trl=5.0)
    ref, fr = orc.Search(seg, nthreads=8).job_core(pm, mm, nn + 1, rng, opts)
    got, fg = pgw.Handle(seg, abi.PREC_FP64).job_core(pm, mm, nn + 1, rng, opts, cap=1 << 21)
    print(name, "records", len(got), flush=True)
    assert fg == fr and got.shape == ref.shape and len(ref) > 1000
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.parametrize("name", ["tiny-1det-12k", "huge-2det-500k"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_extreme_fft_lengths(pgw, abi, orc, name, prec):
    """The smallest and largest supported F-statistic FFT (L = 2^14 with a 16-point column pass,
    L = 2^20 with a 256-point column pass and Ninterp = 2^20) against the oracle, with the row
    pass variants of each precision."""
    seg = segment(name)
    pm, mm, nn = injected_sky(seg)
    sky = [(pm, mm, nn), (pm, mm + 1, nn)]
    rng, opts = abi.make_range(seg), abi.make_opts(trl=12.0)
    S = orc.Search(seg, nthreads=8)
    ref = np.concatenate([S.job_core(*s, rng, opts)[0] for s in sky])
    for tune in (["row8=0", "row8=2"] if prec == "fp64" else ["", "row_na=2"]):
        old = os.environ.get("PGW_TUNE")
        os.environ["PGW_TUNE"] = tune
        try:
            h = pgw.Handle(seg, abi.PREC_FP64 if prec == "fp64" else abi.PREC_FP32)
        finally:
            if old is None:
                del os.environ["PGW_TUNE"]
            else:
                os.environ["PGW_TUNE"] = old
        got, fg = h.job_batch(sky, rng, opts)
        h.close()
        print(name, prec, repr(tune), "records", len(got), len(ref), flush=True)
        F = lambda snr: snr * snr / 2 + 2  # noqa: E731
        if prec == "fp64":
            assert got.shape == ref.shape and len(ref) > 0
            assert np.array_equal(got[:, :4], ref[:, :4])
            assert np.max(np.abs(F(got[:, 4]) - F(ref[:, 4])) / F(ref[:, 4])) <= 1e-9
        else:
            kg = {tuple(r[:4]): F(r[4]) for r in got}
            strong = [r for r in ref if F(r[4]) > 12.05]
            assert strong and all(tuple(r[:4]) in kg for r in strong)
            assert max(abs(kg[tuple(r[:4])] - F(r[4])) / F(r[4]) for r in strong) <= 1e-4


@pytest.mark.parametrize("white", [False, True])
def test_eight_detectors_and_whitenoise(pgw, abi, orc, white):
    """Eight detectors (the handle's maximum), and --whitenoise (no FStat normalisation,
    jobcore.c:650): FP64 records equal the oracle's."""
    seg = segment("net-8det-16k")
    pm, mm, nn = injected_sky(seg)
    rng, opts = abi.make_range(seg), abi.make_opts(trl=12.0, white=white)
    ref, fr = orc.Search(seg, nthreads=8).job_core(pm, mm, nn, rng, opts)
    got, fg = pgw.Handle(seg, abi.PREC_FP64).job_core(pm, mm, nn, rng, opts)
    print("white", white, "records", len(got), flush=True)
    assert fg == fr and got.shape == ref.shape and len(ref) > 0
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replays(pgw, abi, oracle_runs, prec):
    """graph=1: each group's launch sequence is captured into a HIP graph and launched as one;
    the handle keeps one executable graph per buffer set and updates it in place on later groups.
    Five calls in a row on one handle (two groups each, so both graphs are replayed and updated),
    with another handle's eager launches and synchronisations between them: every call's records
    equal the oracle's."""
    name = "net-2det-32k"
    sky, ref, fr = oracle_runs[name]
    seg = segment(name)
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_FP32
    old = os.environ.get("PGW_TUNE")
    os.environ["PGW_TUNE"] = "graph=1,group=1"
    try:
        h = pgw.Handle(seg, p)
    finally:
        if old is None:
            del os.environ["PGW_TUNE"]
        else:
            os.environ["PGW_TUNE"] = old
    eager = pgw.Handle(seg, p)
    rng, opts = abi.make_range(seg), abi.make_opts(trl=12.0)
    for call in range(5):
        got, fg = h.job_batch(sky, rng, opts)
        eager.job_batch(sky[::-1], rng, opts)
        print(prec, "call", call, "records", len(got), len(ref), flush=True)
        assert fg == fr
        if prec == "fp64":
            assert got.shape == ref.shape and len(ref) > 0
            assert np.array_equal(got[:, :4], ref[:, :4])
            Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
            assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9
        else:
            rel, nmatch, allowed, viol = fp32_record_diff(got, ref, 12.0)
            assert not viol and nmatch > 0, viol[:5]
    assert "graph=1" in h.tune_string()
    h.close()
    eager.close()


def test_tune_string(pgw, abi):
    """pgw_tune_string reports the variants PGW_TUNE pinned, and a handle created with that string
    reports the same string (defaults included); the autotuner's choice is reported in the same
    form."""
    seg = segment("net-2det-32k")
    old = os.environ.get("PGW_TUNE")
    try:
        os.environ["PGW_TUNE"] = "tb=64,col_seq=0,row8=4,col_nt=1,row_nt=0,rev=0"
        s = pgw.Handle(seg, abi.PREC_FP32).tune_string()
        kv = dict(x.split("=") for x in s.split(","))
        assert (kv["tb"], kv["col_seq"], kv["row8"], kv["col_nt"], kv["row_nt"], kv["rev"]) == \
            ("64", "0", "4", "1", "0", "0"), s
        os.environ["PGW_TUNE"] = s
        assert pgw.Handle(seg, abi.PREC_FP32).tune_string() == s
    finally:
        if old is None:
            os.environ.pop("PGW_TUNE", None)
        else:
            os.environ["PGW_TUNE"] = old
    h = pgw.Handle(seg, abi.PREC_FP64)
    pm, mm, nn = injected_sky(seg)
    choice = h.autotune(0, pm, mm, nn, abi.make_opts(trl=12.0))
    kv = dict(x.split("=") for x in h.tune_string().split(","))
    print(choice, kv, flush=True)
    assert int(kv["tb"]) == choice["tb"] and int(kv["col_seq"]) == choice["col_seq"]
    # tuned over the templates of the workload's spindown range (pgw_autotune_range)
    h32 = pgw.Handle(seg, abi.PREC_FP32)
    choice = h32.autotune(0, pm, mm, nn, abi.make_opts(trl=12.0), abi.make_range(seg), 8)
    kv = dict(x.split("=") for x in h32.tune_string().split(","))
    print(choice, kv, flush=True)
    assert int(kv["tb"]) == choice["tb"] and int(kv["col_seq"]) == choice["col_seq"]
