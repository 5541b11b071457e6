"""The reference's own test inputs (testdata/test-gpu-cpu/001: Gaussian-noise H1/L1 frame, band
666, N=86164, fftpad 1; range of search/network/src-gpu/range.dat) read through the reference's
file formats, searched by the oracle and by the device, and written as trigger files."""
import math
from pathlib import Path

import numpy as np
import pytest

from polgraw_allsky_amd import io

FIX = Path(__file__).resolve().parent / "golden" / "ref-001"


@pytest.fixture(scope="module")
def frame():
    return io.load_frame(FIX, 1, 666, range_file=FIX / "range.dat")


def test_readers(frame, abi):
    s = frame.sett
    assert frame.dets == ["H1", "L1"]
    assert (s["N"], s["nfft"], s["fftpad"], s["nfftf"], s["nmin"], s["nmax"]) == (86164, 1 << 17, 1, 1 << 17, 1024, 64512)
    assert frame.range == dict(spndr=[-40, 40], nr=[-1, 1], mr=[-2, 2], pmr=[1, 1])
    # grid.bin zero pattern and M0 = 2 pi / (fftpad nfft)
    assert all(frame.M[k] == 0 for k in (1, 2, 3, 6, 7, 11))
    assert frame.M[0] == pytest.approx(2 * math.pi / (1 << 17), rel=1e-8)
    # DetSSB.bin tail: epsm is the obliquity of the ecliptic (84381.448 arcsec); positions ~1 AU / (c dt)
    assert frame.epsm == pytest.approx(84381.448 / 3600 * math.pi / 180, abs=1e-12)
    for e in frame.DetSSB:
        r = np.linalg.norm(e.reshape(-1, 3), axis=1)
        assert 240 < r.min() and r.max() < 255
    # sig2 = crf0 * var (n - 1, zeros included)
    for x, v in zip(frame.xDat, frame.sig2):
        nz = np.count_nonzero(x == 0)
        assert v == pytest.approx(len(x) / (len(x) - nz) * np.var(x, ddof=1), rel=1e-10)


def test_triggers_round_trip(tmp_path):
    rec = np.random.default_rng(1).standard_normal((7, 5))
    p = tmp_path / "triggers_001_0666_1.bin"
    rec.tofile(p)
    assert np.array_equal(io.read_triggers(p), rec)
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(ValueError):
        io.read_triggers(p)


def test_fixture_gaps(frame):
    """The frame has long zeroed gaps (42 % of H1, 53 % of L1 samples), N not a power of two."""
    z = [np.count_nonzero(x == 0) / len(x) for x in frame.xDat]
    assert 0.4 < z[0] < 0.45 and 0.5 < z[1] < 0.55


@pytest.fixture(scope="module")
def oracle_records(frame, orc, abi):
    S = orc.Search(frame, nthreads=8)
    r = abi.make_range(frame)
    opts = abi.make_opts(trl=12.0)
    recs, fn = [], 0
    for mm in range(r.mr[0], r.mr[1] + 1):
        for nn in range(r.nr[0], r.nr[1] + 1):
            rec, f = S.job_core(1, mm, nn, r, opts)
            recs.append(rec)
            fn += f
    return np.concatenate(recs), fn


@pytest.mark.gpu
def test_fixture_search_fp64(frame, abi, oracle_records, tmp_path):
    from polgraw_allsky_amd import pgw
    ref, fn_ref = oracle_records
    h = pgw.Handle(frame, abi.PREC_FP64)
    fn = h.search(abi.make_range(frame), abi.make_opts(trl=12.0), tmp_path, ident=1, band=666)
    got = io.read_triggers(tmp_path / "triggers_001_0666_1.bin")
    print("templates", fn, "records", len(got), flush=True)
    assert fn == fn_ref == 81 * 15
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.gpu
def test_fixture_search_fp32(frame, abi, oracle_records):
    """FP32 form on the same inputs against the FP64 oracle's records by the rule of
    cases.fp32_record_diff: matched F within 1e-4; a record in one set only where its F is within
    1e-4 trl of the threshold or a run's maximum moved between bins whose F agree within 1e-4."""
    from polgraw_allsky_amd import pgw
    from cases import fp32_record_diff
    ref, fn_ref = oracle_records
    h = pgw.Handle(frame, abi.PREC_FP32)
    got, fn = h.job_batch([(1, mm, nn) for mm in range(-2, 3) for nn in range(-1, 2)], abi.make_range(frame),
                          abi.make_opts(trl=12.0))
    rel, nmatch, allowed, viol = fp32_record_diff(got, ref, 12.0)
    print("records fp32", len(got), "fp64", len(ref), "matched", nmatch, "max rel dF", rel, "allowed", allowed,
          flush=True)
    assert fn == fn_ref
    assert not viol, viol[:5]
    assert nmatch > 0


def test_known_lines(frame, tmp_path):
    """Known-line veto inputs: widths from the ephemeris, band mapping and narrow-down, on the
    reference frame's DetSSB."""
    s = frame.sett
    f0 = s["fpo"] + 0.3 * s["B"]
    p = tmp_path / "O1LinesToBeCleaned_H1_v1.txt"
    p.write_text("% comment\n"
                 f"{f0}\t0\t0\t0\t0\t0.001\t0.001\n"          # singlet inside the band
                 f"0.05\t1\t{s['fpo'] - 0.02}\t0\t3\t0\t0\n"    # comb: harmonics at fpo-0.02 .. fpo+0.13
                 f"0.05\t2\t{s['fpo'] + 0.4}\t1\t2\t0\t0\n")   # scaling comb, beyond the band
    lines = io.read_lines(p, s["N"], s["dt"], frame.DetSSB[0])
    assert len(lines) == 1 + 4 + 2
    w0 = (lines[0][1] - lines[0][0]) / 2
    # the broadening is the orbital Doppler spread ~1e-4 f plus the spindown term
    assert 0.5e-4 * f0 < w0 < 2e-4 * f0
    assert lines[0][0] + w0 == pytest.approx(f0)
    # scaling comb: harmonic k has k times the width of its frequency
    for k, (a, b) in zip((1, 2), lines[5:]):
        f = s["fpo"] + 0.4 + k * 0.05
        assert (a + b) / 2 == pytest.approx(f)
    band = io.lines_in_band(s["fpo"], s["B"], [lines])
    # the singlet maps to 0.3 pi; combs below fpo are clipped at 0; the scaling comb is outside
    assert any(a < 0.3 * math.pi < b for a, b in band)
    assert all(0 <= a <= b <= math.pi for a, b in band)
    assert not any(a > 0.9 * math.pi for a, b in band)
    nd = io.narrow_down_lines(0.4)
    assert nd == [[0.0, math.pi / 2 - 0.4], [math.pi / 2 + 0.4, math.pi]]
    both = io.lines_in_band(s["fpo"], s["B"], [lines], start=nd)
    assert both[:2] == nd
    assert io.veto_fraction(nd) == pytest.approx(1 - 0.8 / math.pi)
    assert io.veto_fraction([[0, 1], [0.5, 2], [3, 3.1]]) == pytest.approx(2.1 / math.pi)
    assert io.read_lines(tmp_path / "missing.txt", s["N"], s["dt"], frame.DetSSB[0]) == []


def test_band_veto_assembly(frame, tmp_path):
    """band_lines: narrow-down first, then the detectors' lines in the band; fully vetoed bands."""
    s = frame.sett
    assert io.band_lines(frame, tmp_path) == []                      # no lines files, no narrow-down
    assert io.band_lines(frame, tmp_path, narrowdown=0.2) == io.narrow_down_lines(0.2)
    f0 = s["fpo"] + 0.7 * s["B"]
    (tmp_path / "O1LinesToBeCleaned_L1_v1.txt").write_text(f"{f0} 0 0 0 0 0 0\n")
    got = io.band_lines(frame, tmp_path, narrowdown=0.2)
    assert len(got) == 3 and got[:2] == io.narrow_down_lines(0.2)
    assert got[2][0] < 0.7 * math.pi < got[2][1]
    assert io.band_lines(frame, tmp_path, narrowdown=0.2, veto=False) == io.narrow_down_lines(0.2)
    assert not io.band_fully_vetoed(got)
    assert io.band_fully_vetoed(io.narrow_down_lines(0.0))           # [0, pi/2] and [pi/2, pi]
    assert io.band_fully_vetoed([[1.0, math.pi], [0.0, 1.5]])         # out of order, overlapping
    assert not io.band_fully_vetoed([[0.0, 1.0], [1.1, math.pi]])     # a gap at 1..1.1
    assert not io.band_fully_vetoed([[0.0, 3.0]])                     # open at the top


@pytest.mark.gpu
def test_fixture_veto_fp64(frame, abi, orc, tmp_path):
    """Known-line veto and narrow-down from the reference's inputs: the device's records equal the
    oracle's with the same ranges, none of them inside a vetoed range, and fewer than without."""
    from polgraw_allsky_amd import pgw
    s = frame.sett
    (tmp_path / "O1LinesToBeCleaned_H1_v1.txt").write_text(
        f"% singlet and a comb\n{s['fpo'] + 0.6 * s['B']} 0 0 0 0 0 0\n0.01 1 {s['fpo']} 10 16 0 0\n")
    lines = io.band_lines(frame, tmp_path, narrowdown=1.2)
    assert len(lines) == 2 + 1 + 7
    r, opts = abi.make_range(frame), abi.make_opts(trl=12.0)
    got, fg = pgw.Handle(frame, abi.PREC_FP64, lines=lines).job_core(1, 0, 0, r, opts)
    ref, fr = orc.Search(frame, nthreads=8, lines=lines).job_core(1, 0, 0, r, opts)
    full, _ = pgw.Handle(frame, abi.PREC_FP64).job_core(1, 0, 0, r, opts)
    print("records", len(got), "unvetoed", len(full), flush=True)
    assert fg == fr and got.shape == ref.shape and 0 < len(got) < len(full)
    assert np.array_equal(got[:, :4], ref[:, :4])
    assert np.max(np.abs(got[:, 4] - ref[:, 4]) / ref[:, 4]) <= 1e-9
    for lo, hi in lines:
        assert not np.any((got[:, 0] >= lo) & (got[:, 0] <= hi))
