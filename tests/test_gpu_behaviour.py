"""Reference behaviours of job_core reproduced by the HIP path (all against the CPU oracle).

jobcore.c:424-444 spindown range without a range file; --nospindown (jobcore.c:444,456); the
known-line veto (jobcore.c:676-681); the sky disk check (jobcore.c:248); the record buffer
capacity (an error code instead of the reference's silent overflow); job_batch == job_core loop;
an ephemeris outside the fused spline kernel's assumption (an error, and the three-kernel path);
and the FP32 form against the FP64 oracle.
"""
import numpy as np
import pytest

from cases import fp32_record_diff, injected_sky, segment

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


def _same(got, ref, fn_g, fn_r, ftol=1e-9):
    assert fn_g == fn_r
    assert got.shape == ref.shape, (got.shape, ref.shape)
    if len(ref):
        assert np.array_equal(got[:, :4], ref[:, :4])
        Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
        assert np.max(np.abs(Fg - Fr) / Fr) <= ftol


def test_spindown_range_from_smin_smax(pgw, abi, orc):
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0, use_range=False)
    pm, mm, nn = injected_sky(seg)
    o = orc.Search(seg, nthreads=8)
    lo, hi = o.spindown_range(mm, nn, rng, opts)
    assert hi - lo + 1 > 12  # wider than the range-file case
    got, fg = pgw.Handle(seg).job_core(pm, mm, nn, rng, opts)
    ref, fr = o.job_core(pm, mm, nn, rng, opts)
    print("templates", fr, "records", len(ref), flush=True)
    _same(got, ref, fg, fr)


def test_nospindown(pgw, abi, orc):
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=10.0, s0=True)
    pm, mm, nn = injected_sky(seg)
    got, fg = pgw.Handle(seg).job_core(pm, mm, nn, rng, opts)
    ref, fr = orc.Search(seg, nthreads=4).job_core(pm, mm, nn, rng, opts)
    assert fr == 1 and np.all(ref[:, 1] == 0.0)
    _same(got, ref, fg, fr)


def test_line_veto(pgw, abi, orc):
    seg = segment("plain-1det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0)
    pm, mm, nn = injected_sky(seg)
    full, _ = orc.Search(seg, nthreads=4).job_core(pm, mm, nn, rng, opts)
    f0 = float(np.median(full[:, 0]))
    lines = [(f0 - 2e-4, f0 + 2e-4), (0.01, 0.02)]
    got, fg = pgw.Handle(seg, lines=lines).job_core(pm, mm, nn, rng, opts)
    ref, fr = orc.Search(seg, nthreads=4, lines=lines).job_core(pm, mm, nn, rng, opts)
    assert 0 < len(ref) < len(full)
    _same(got, ref, fg, fr)


def test_outside_disk_and_batch_mix(pgw, abi, orc):
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=12.0)
    pm, mm, nn = injected_sky(seg)
    far = (pm, mm + 400, nn + 400)  # (al1^2 + al2^2)/oms^2 > 1: skipped, no templates
    h = pgw.Handle(seg)
    got, fg = h.job_core(*far, rng, opts)
    assert fg == 0 and len(got) == 0
    sky = [far, (pm, mm, nn), far, (pm, mm, nn - 1)]
    got, fg = h.job_batch(sky, rng, opts)
    o = orc.Search(seg, nthreads=8)
    recs, fr = [], 0
    for s in sky:
        r, f = o.job_core(*s, rng, opts)
        recs.append(r)
        fr += f
    _same(got, np.concatenate(recs), fg, fr)


def test_capacity_error(pgw, abi):
    seg = segment("net-2det-32k")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=10.0)
    pm, mm, nn = injected_sky(seg)
    h = pgw.Handle(seg)
    full, _ = h.job_core(pm, mm, nn, rng, opts)
    assert len(full) > 2
    with pytest.raises(pgw.PGWError) as e:
        h.job_core(pm, mm, nn, rng, opts, cap=2)
    assert e.value.status == abi.PGW_ERR_CAPACITY


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_nonmonotonic_shift_is_an_error(pgw, abi, prec):
    """The fused spline kernel evaluates each output in the workgroup that holds its knots, which
    needs the abscissae interpftpad (i - shftf_i) to increase with i (physical ephemerides change
    the shift by far less than a sample per sample).  An ephemeris with a jump of thousands of
    samples breaks that: the search must fail with PGW_ERR_DEVICE, and the three-kernel path
    (spl_fuse=0, which evaluates every output on its own) must run.  The jump is 3000 samples of
    shift along the sky point's direction (from a record of the unmodified segment), so the
    abscissae fall back by 6000 knots, more than one workgroup's knots."""
    import dataclasses
    import os
    base = segment("plain-1det-32k")
    N = base.sett["N"]
    rng = abi.make_range(base)
    opts = abi.make_opts(trl=20.0)
    pm, mm, nn = injected_sky(base)
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_FP32
    recs, fn = pgw.Handle(base, p).job_core(pm, mm, nn, rng, opts)
    assert len(recs) > 0
    delta, alpha = float(recs[0][2]), float(recs[0][3])
    nsrc = np.array([np.cos(alpha) * np.cos(delta), np.sin(alpha) * np.cos(delta), np.sin(delta)])
    ssb = base.DetSSB[0].copy().reshape(N, 3)
    ssb[N // 2:] += 3000.0 * nsrc
    seg = dataclasses.replace(base, DetSSB=[ssb.reshape(-1)])
    rng = abi.make_range(seg)
    with pytest.raises(pgw.PGWError) as e:
        pgw.Handle(seg, p).job_core(pm, mm, nn, rng, opts)
    assert e.value.status == abi.PGW_ERR_DEVICE
    old = os.environ.get("PGW_TUNE")
    os.environ["PGW_TUNE"] = "spl_fuse=0"
    try:
        h = pgw.Handle(seg, p)
    finally:
        if old is None:
            os.environ.pop("PGW_TUNE", None)
        else:
            os.environ["PGW_TUNE"] = old
    recs, fn = h.job_core(pm, mm, nn, rng, opts)
    assert fn > 0


def test_small_shift_jump_is_exact_or_an_error(pgw, abi):
    """A jump of 10 samples of shift drops the abscissae by 18 knots: the outputs after it belong
    to the workgroup of the knots before it unless the drop crosses a workgroup's first knot.  The
    fused spline kernel must then give the three-kernel path's records (fields 0-3 equal, F within
    1e-9) or fail with PGW_ERR_DEVICE, never other records."""
    import dataclasses
    import os
    base = segment("plain-1det-32k")
    N = base.sett["N"]
    rng = abi.make_range(base)
    opts = abi.make_opts(trl=20.0)
    pm, mm, nn = injected_sky(base)
    recs, _ = pgw.Handle(base, abi.PREC_FP64).job_core(pm, mm, nn, rng, opts)
    delta, alpha = float(recs[0][2]), float(recs[0][3])
    nsrc = np.array([np.cos(alpha) * np.cos(delta), np.sin(alpha) * np.cos(delta), np.sin(delta)])
    ssb = base.DetSSB[0].copy().reshape(N, 3)
    ssb[N // 2 + 777:] += 10.0 * nsrc
    seg = dataclasses.replace(base, DetSSB=[ssb.reshape(-1)])
    rng = abi.make_range(seg)
    old = os.environ.get("PGW_TUNE")
    os.environ["PGW_TUNE"] = "spl_fuse=0"
    try:
        h = pgw.Handle(seg, abi.PREC_FP64)
    finally:
        if old is None:
            os.environ.pop("PGW_TUNE", None)
        else:
            os.environ["PGW_TUNE"] = old
    ref, fr = h.job_core(pm, mm, nn, rng, opts)
    assert fr > 0 and len(ref) > 0
    try:
        got, fg = pgw.Handle(seg, abi.PREC_FP64).job_core(pm, mm, nn, rng, opts)
    except pgw.PGWError as e:
        assert e.status == abi.PGW_ERR_DEVICE
        print("fused spline: PGW_ERR_DEVICE (the drop crosses a workgroup's first knot)")
        return
    assert fg == fr and got.shape == ref.shape
    assert np.array_equal(got[:, :4], ref[:, :4])
    assert np.max(np.abs(got[:, 4] - ref[:, 4]) / ref[:, 4]) <= 1e-9
    print("fused spline: records equal to the three-kernel path's")


@pytest.mark.parametrize("name", ["net-2det-86k", "plain-1det-86k", "batch-3det-60k", "prec-2det-259k"])
def test_fp32_form_vs_fp64_oracle(name, pgw, abi, orc):
    """FP32 series/FFTs (FP64 phases and spline coefficients) vs the FP64 oracle: max relative dF
    <= 1e-4 on bins with F >= 1, and the records by the rule of cases.fp32_record_diff."""
    seg = segment(name)
    pm, mm, nn = injected_sky(seg)
    ss = int(seg.injection["grid"][1])
    opts = abi.make_opts(trl=20.0)
    h = pgw.Handle(seg, abi.PREC_FP32)
    o = orc.Search(seg, nthreads=8)
    Fg = h.fstat(pm, mm, nn, ss, opts)
    Fr = o.fstat(pm, mm, nn, ss, opts)
    sel = Fr >= 1.0
    rel = np.max(np.abs(Fg - Fr)[sel] / Fr[sel])
    print(name, "max rel dF (F>=1)", rel, "max |dF|", np.abs(Fg - Fr).max(), flush=True)
    assert rel <= 1e-4
    rng = abi.make_range(seg)
    got, fg = h.job_core(pm, mm, nn, rng, opts)
    ref, fr = o.job_core(pm, mm, nn, rng, opts)
    assert fg == fr
    rel, nmatch, allowed, viol = fp32_record_diff(got, ref, 20.0)
    print(name, "records", len(got), len(ref), "max rel dF", rel, "allowed", allowed, flush=True)
    assert not viol, viol[:5]
    assert nmatch > 0
