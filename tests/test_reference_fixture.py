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
    """FP32 form on the same inputs: every oracle record away from the threshold (F > 12.01) is
    found at the same template and bin, F within 1e-4."""
    from polgraw_allsky_amd import pgw
    ref, _ = oracle_records
    h = pgw.Handle(frame, abi.PREC_FP32)
    got, fn = h.job_batch([(1, mm, nn) for mm in range(-2, 3) for nn in range(-1, 2)], abi.make_range(frame),
                          abi.make_opts(trl=12.0))
    kg = {tuple(r[:4]): r[4] for r in got}
    F = lambda snr: snr * snr / 2 + 2  # noqa: E731
    strong = [r for r in ref if F(r[4]) > 12.01]
    hit = [k for k in (tuple(r[:4]) for r in strong) if k in kg]
    assert len(hit) >= 0.999 * len(strong)
    rel = max(abs(F(kg[tuple(r[:4])]) - F(r[4])) / F(r[4]) for r in strong if tuple(r[:4]) in kg)
    assert rel <= 1e-4
