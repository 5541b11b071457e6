"""Software injections (the reference's -addsig path, search/network/src-cpu/init.c:462-651).

Pinned by what the mathematics fixes, not by a second copy of add_signal's formulas:
* the search box: a signal placed exactly at a grid point through the search's own forward map
  (jobcore.c:243-262,456: al1/al2 -> lin2ast sky position, spindown ss M5 + nn M9 + mm M13) comes
  back at that grid point (a transposed M, a wrong hemisphere or a wrong frequency move fails);
* the moves between frames: a signal given at frame reffr and added to frame ident equals, sample
  for sample, the same signal added at frame reffr to a longer frame that runs on through frame
  ident (time translation; a wrong sign of the frequency move or of the phase shift fails);
* the snr normalisation: the squared injected series sums to snr^2;
* the device search on the reference's own test frame finds the injection, with the oracle's records.
"""
import math
import types
from pathlib import Path

import numpy as np
import pytest

from polgraw_allsky_amd import inject, io, synth

FIX = Path(__file__).resolve().parent / "golden" / "ref-001"


def lin2ast(be1, be2, pm, sepsm, cepsm):
    """Sky position of linear coordinates (auxi.c:15-24), as (delta, alpha)."""
    sindel = be1 * sepsm - (2 * pm - 3) * math.sqrt(1.0 - be1 * be1 - be2 * be2) * cepsm
    cosdel = math.sqrt(1.0 - sindel * sindel)
    sinal = (be1 - sepsm * sindel) / (cepsm * cosdel)
    cosal = be2 / cosdel
    return math.asin(sindel), math.fmod(math.atan2(sinal, cosal) + 2 * math.pi, 2 * math.pi)


def signal_at(sett, M, ss, nn, mm, pm, f, d, kind="amp", value=1.0, gsize=1, amps=(0.6, -0.3, 0.2, 0.7)):
    """A signal file that puts the injection at grid point (ss, nn, mm, pm) of frame ident = reffr - d,
    with frequency f (radians per sample) at that frame's first sample."""
    cof = sett["oms"] + f
    al1 = nn * M[10] + mm * M[14]
    al2 = nn * M[11] + mm * M[15]
    e = synth.C_EPSMA
    delta, alpha = lin2ast(al1 / cof, al2 / cof, pm, math.sin(e), math.cos(e))
    fdot = ss * M[5] + nn * M[9] + mm * M[13]
    f_ref = f + 2.0 * fdot * sett["N"] * d          # the frequency at frame reffr
    return inject.Signal(kind=kind, value=value, gsize=gsize, reffr=7 + d,
                         sgnlo=np.array([f_ref, fdot, delta, alpha, *amps]))


@pytest.fixture(scope="module")
def frame():
    return io.load_frame(FIX, 1, 666, range_file=FIX / "range.dat")


def test_signal_file_round_trip(tmp_path):
    s = inject.Signal("snr", 12.5, 2, 10, np.array([1.25, -3e-9, 0.3, 4.1, 0.5, -0.25, 0.75, 0.125]))
    p = tmp_path / "sig.txt"
    inject.write_signal(p, s)
    lines = p.read_text().splitlines()
    assert lines[:3] == ["snr 1.250000e+01", "2", "10"] and len(lines) == 11
    r = inject.read_signal(p)
    assert (r.kind, r.value, r.gsize, r.reffr) == ("snr", 12.5, 2, 10)
    assert np.array_equal(r.sgnlo, s.sgnlo)
    p.write_text("hrss 1 2 3\n" + "0\n" * 8)
    with pytest.raises(ValueError):
        inject.read_signal(p)
    p.write_text("amp 1 2 3\n0 0 0\n")
    with pytest.raises(ValueError):
        inject.read_signal(p)


@pytest.mark.parametrize("pm", [1, 2])
@pytest.mark.parametrize("d", [0, 3, -2])
@pytest.mark.parametrize("point", [(10, 0, 1), (-37, 1, -2), (4, -1, 2)])
def test_box_centres_on_grid_point(frame, point, pm, d):
    ss, nn, mm = point
    sig = signal_at(frame.sett, frame.M, ss, nn, mm, pm, 1.5, d, gsize=2)
    rng, s, cof = inject.signal_range(frame.sett, frame.M, sig, ident=7)
    assert rng == dict(spndr=[ss - 2, ss + 2], nr=[nn - 2, nn + 2], mr=[mm - 2, mm + 2], pmr=[pm, pm])
    assert s[0] == pytest.approx(1.5, abs=1e-9) and cof == pytest.approx(frame.sett["oms"] + 1.5, abs=1e-9)


def test_frequency_outside_band_is_an_error(frame):
    for f in (-0.01, math.pi + 0.01):
        sig = signal_at(frame.sett, frame.M, 0, 0, 0, 1, f, 0)
        with pytest.raises(ValueError):
            inject.signal_range(frame.sett, frame.M, sig, ident=7)


def _line_frame(N, phir, dt=2.0):
    """A frame for the series alone: detector at the barycentre (no Roemer delay), H1's
    amplitude modulation with sidereal angle phir at the first sample."""
    sett = synth.settings_dict(N, dt=dt)
    lat, _, gam = synth.SITES["H1"]
    return types.SimpleNamespace(sett=sett, dets=["H1"], phir=[phir], amod=[synth.rogcvir(lat, gam)],
                                 DetSSB=[np.zeros(3 * N)])


@pytest.mark.parametrize("frames_later", [1, 2])
def test_series_is_coherent_across_frames(frames_later):
    """Frame ident = reffr + k of N samples follows k N samples after frame reffr starts: the signal
    added to it equals the tail of the same signal added at frame reffr to a frame of (k+1) N samples."""
    N, k = 12000, frames_later
    fdot = -2.5e-9
    s_ref = np.array([0.8, fdot, 0.4, 2.0, 0.6, -0.3, 0.2, 0.7])
    sig = inject.Signal("amp", 1.0, 1, 5, s_ref)
    long_fr = _line_frame((k + 1) * N, 0.3)
    omr = long_fr.sett["omr"]
    later = _line_frame(N, math.fmod(0.3 + omr * k * N, 2 * math.pi))
    s_long = s_ref.copy()
    s_late = s_ref.copy()
    s_late[0] += -2.0 * fdot * N * (sig.reffr - (sig.reffr + k))
    a = inject.signal_series(long_fr, s_long, long_fr.sett["oms"] + s_long[0], 0)[0]
    b = inject.signal_series(later, s_late, later.sett["oms"] + s_late[0], -k)[0]
    assert np.max(np.abs(a[k * N:] - b)) <= 1e-7


def test_snr_normalisation_and_gaps(synth):
    seg = synth.make_segment(N=20000, dets=("H1", "L1"), sigmas=(1.0, 1.2), seed=4, inject=False,
                             gap_frac=0.0, nsky=(1, 1), nspin=1)
    old = [x.copy() for x in seg.xDat]
    sig2 = list(seg.sig2)
    sig = signal_at(seg.sett, seg.M, 3, 0, 0, 1, 1.0, 0, kind="snr", value=9.0)
    sig.reffr = 11
    inject.add_signal(seg, sig, ident=11)
    tot = sum(float(np.sum((x - o) ** 2)) for x, o in zip(seg.xDat, old))
    assert tot == pytest.approx(81.0, rel=1e-9)
    assert seg.sig2 == sig2
    # zeroed samples stay zero; the others get h0 times the series
    gapped = synth.make_segment(N=20000, dets=("H1",), seed=6, inject=False, gap_frac=0.3, nsky=(1, 1), nspin=1)
    zeros = gapped.xDat[0] == 0.0
    amp = inject.Signal("amp", 0.5, 1, 11, sig.sgnlo)
    inject.add_signal(gapped, amp, ident=11)
    assert zeros.sum() > 0.25 * 20000 and np.all(gapped.xDat[0][zeros] == 0.0)
    assert gapped.injection["h0"] == 0.5


def test_load_frame_with_signal_file(frame, tmp_path):
    sig = signal_at(frame.sett, frame.M, 10, 0, 1, 1, 1.5, 0, kind="snr", value=15.0)
    sig.reffr = 1
    p = tmp_path / "sig.txt"
    inject.write_signal(p, sig)
    fr = io.load_frame(FIX, 1, 666, range_file=FIX / "range.dat", addsig=p)
    assert fr.range == dict(spndr=[9, 11], nr=[-1, 1], mr=[0, 2], pmr=[1, 1])
    assert fr.sig2 == frame.sig2
    for x, x0 in zip(fr.xDat, frame.xDat):
        assert np.array_equal(x == 0, x0 == 0)
        assert not np.array_equal(x, x0)


@pytest.fixture(scope="module")
def injected(tmp_path_factory):
    """The reference's test frame with a signal at grid point (ss, nn, mm) = (10, 0, 1).  The
    reference's snr assumes unit noise; this frame's noise has variances 3.4 and 7.4 and gaps, so
    snr 60 gives a signal-to-noise ratio of about 21 in the data."""
    base = io.load_frame(FIX, 1, 666)
    sig = signal_at(base.sett, base.M, 10, 0, 1, 1, 1.5, 0, kind="snr", value=60.0)
    sig.reffr = 1
    p = tmp_path_factory.mktemp("sig") / "sig.txt"
    inject.write_signal(p, sig)
    fr = io.load_frame(FIX, 1, 666, addsig=p)
    rho = math.sqrt(sum(float(np.sum((x - y) ** 2)) / s for x, y, s in zip(fr.xDat, base.xDat, base.sig2)))
    return fr, sig, rho


def _oracle_records(fr, orc, abi, trl):
    S = orc.Search(fr, nthreads=8)
    r = abi.make_range(fr)
    opts = abi.make_opts(trl=trl)
    recs = []
    for mm in range(r.mr[0], r.mr[1] + 1):
        for nn in range(r.nr[0], r.nr[1] + 1):
            rec, _ = S.job_core(1, mm, nn, r, opts)
            recs.append(rec)
    return np.concatenate(recs)


def test_oracle_recovers_injection(injected, orc, abi):
    """The oracle's search of the box finds the injection: the template at the injected grid point
    has a record within a frequency bin of the injected frequency with about the signal-to-noise
    ratio in the data, and no record of the box is far from the injected frequency (threshold
    F = 40: the noise of this frame alone reaches F = 25 in the box)."""
    fr, sig, rho = injected
    rec = _oracle_records(fr, orc, abi, 40.0)
    binw = 2 * math.pi / (fr.sett["fftpad"] * fr.sett["nfft"])
    at = rec[(rec[:, 1] == sig.sgnlo[1]) & (np.abs(rec[:, 0] - sig.sgnlo[0]) <= binw)]
    print("rho", rho, "records", len(rec), "at the grid point", at, flush=True)
    assert len(at) == 1
    assert at[0, 2] == pytest.approx(sig.sgnlo[2], abs=1e-3)
    assert at[0, 3] == pytest.approx(sig.sgnlo[3], abs=1e-3)
    assert 0.8 * rho <= at[0, 4] <= 1.2 * rho
    assert np.all(np.abs(rec[:, 0] - sig.sgnlo[0]) <= 0.02)
    assert rec[:, 4].max() <= 1.2 * rho


@pytest.mark.gpu
def test_device_search_recovers_injection(injected, orc, abi, tmp_path):
    """pgw_search over the box around the injection writes the oracle's records."""
    from polgraw_allsky_amd import pgw
    fr, sig, _ = injected
    ref = _oracle_records(fr, orc, abi, 40.0)
    h = pgw.Handle(fr, abi.PREC_FP64)
    fn = h.search(abi.make_range(fr), abi.make_opts(trl=40.0), tmp_path, ident=1, band=666)
    got = io.read_triggers(tmp_path / "triggers_001_0666_1.bin")
    print("templates", fn, "records", len(got), "best", got[np.argmax(got[:, 4])], flush=True)
    assert fn == 27
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


def test_random_signal_draws(frame):
    """sigen's draws (sigen.c:270-357): spindown in [-Smax, 0], frequency in the searched band, sky
    position on the grid disk in the drawn hemisphere, and amplitudes with the invariants of their
    parametrisation: a1^2 + a2^2 + a3^2 + a4^2 = hop^2 + hoc^2 and a1 a4 - a2 a3 = hop hoc
    (hop = (1 + hoc^2) / 2) for one cos(inclination) hoc in [-1, 1]."""
    s = frame.sett
    L = s["fftpad"] * s["nfft"]
    rng = np.random.default_rng(99)
    pms = set()
    for _ in range(200):
        sig = inject.random_signal(rng, s, kind="amp", value=0.5, gsize=2, reffr=4)
        f, fdot, delta, alpha = sig.sgnlo[:4]
        assert (sig.kind, sig.value, sig.gsize, sig.reffr) == ("amp", 0.5, 2, 4)
        assert -s["Smax"] <= fdot <= 0
        assert 2 * math.pi * s["nmin"] / L <= f < 2 * math.pi * s["nmax"] / L
        pm, be = synth.ast2lin(alpha, delta, synth.C_EPSMA)
        assert be[0] ** 2 + be[1] ** 2 <= 1 + 1e-12
        pms.add(pm)
        a = sig.sgnlo[4:]
        q = float(np.sum(a * a))
        d = a[0] * a[3] - a[1] * a[2]
        # d = hop hoc fixes hoc (h^3 + h = 2 d has one real root); q must then be hop^2 + hoc^2
        r = np.roots([1.0, 0.0, 1.0, -2.0 * d])
        h = float(r[np.argmin(np.abs(r.imag))].real)
        assert abs(h) <= 1 + 1e-9
        assert q == pytest.approx(((1 + h * h) / 2) ** 2 + h * h, abs=1e-9)
    assert pms == {1, 2}


def test_amplitude_invariants():
    rng = np.random.default_rng(5)
    for _ in range(50):
        ph_o, psik, hoc = rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(-1, 1)
        a = synth.amplitudes(ph_o, psik, hoc)
        hop = (1 + hoc * hoc) / 2
        assert float(np.sum(a * a)) == pytest.approx(hop * hop + hoc * hoc, abs=1e-14)
        assert a[0] * a[3] - a[1] * a[2] == pytest.approx(hop * hoc, abs=1e-14)


@pytest.mark.gpu
def test_injection_recovery_rate(abi):
    """The reference's sensitivity pipeline in miniature (sigen's random signals, the -addsig search
    of the +-gsize box, detection): 12 random signals of snr 60 (about 20 in this frame's data,
    whose noise variances are 3.4 and 7.4, with gaps) added one at a time to the reference's test
    frame and searched on the device in the FP32 form.  The templates of the box lie within
    gsize + 1/2 grid steps of the signal in spindown and sky, which moves the frequency of the
    best template by at most (|M4| + |M8| + |M12|) (gsize + 1/2) (the grid's frequency row).
    Detection: a record in that window with at least 0.7 of the signal-to-noise ratio in the data,
    where the same box of the frame without the signal has no record (this frame has strong
    non-Gaussian features elsewhere in the band).  All 12 are detected."""
    import copy
    from polgraw_allsky_amd import pgw
    base = io.load_frame(FIX, 1, 666)
    M = base.M
    gsize = 1
    win = (abs(M[4]) + abs(M[8]) + abs(M[12])) * (gsize + 0.5)
    opts = abi.make_opts(trl=20.0)
    h_noise = pgw.Handle(base, abi.PREC_FP32)
    rng = np.random.default_rng(2024)
    found = 0
    for k in range(12):
        sig = inject.random_signal(rng, base.sett, kind="snr", value=60.0, gsize=gsize, reffr=1)
        fr = copy.deepcopy(base)
        box = inject.add_signal(fr, sig, ident=1)
        rho = math.sqrt(sum(float(np.sum((x - y) ** 2)) / s for x, y, s in zip(fr.xDat, base.xDat, base.sig2)))
        sky = [(box["pmr"][0], mm, nn) for mm in range(box["mr"][0], box["mr"][1] + 1)
               for nn in range(box["nr"][0], box["nr"][1] + 1)]
        rng_box = abi.make_range(fr)
        h = pgw.Handle(fr, abi.PREC_FP32)
        try:
            rec, fn = h.job_batch(sky, rng_box, opts)
        finally:
            h.close()
        ref, _ = h_noise.job_batch(sky, rng_box, opts)
        near = lambda r: r[np.abs(r[:, 0] - sig.sgnlo[0]) <= win] if len(r) else r  # noqa: E731
        w, wn = near(rec), near(ref)
        best = float(w[:, 4].max()) if len(w) else 0.0
        print(k, "f %.3f rho %.1f templates %d records %d in window %d best %.1f noise in window %d"
              % (sig.sgnlo[0], rho, fn, len(rec), len(w), best, len(wn)), flush=True)
        assert len(wn) == 0
        found += best >= 0.7 * rho
    h_noise.close()
    assert found == 12
