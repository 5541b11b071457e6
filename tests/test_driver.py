"""search() driver (jobcore.c:51-184): trigger files in the reference's format and order,
checkpoint state files (init.c:868-917) and resume."""
import numpy as np
import pytest

from cases import segment


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


def _range(abi, seg, pmr=(1, 2)):
    r = abi.make_range(seg)
    r.pmr[0], r.pmr[1] = pmr
    r.pst = pmr[0]
    return r


def test_read_checkpoint(pgw, abi, tmp_path):
    seg = segment("net-2det-32k")
    q = tmp_path / "state_001_0666.dat"
    r = _range(abi, seg)
    r.pst, r.mst, r.nst, r.sst = 9, 9, 9, 9
    assert pgw.read_checkpoint(q, r) == (0, 0)            # no file: from the beginning
    assert (r.pst, r.mst, r.nst, r.sst) == (r.pmr[0], r.mr[0], r.nr[0], r.spndr[0])
    q.write_text("2 -3 4 17 1234\n")
    assert pgw.read_checkpoint(q, r) == (0, 1234)
    assert (r.pst, r.mst, r.nst, r.sst) == (2, -3, 4, 17)
    q.write_text("")
    assert pgw.read_checkpoint(q, r)[0] == 1              # empty: the run has finished
    q.write_text("2 x\n")
    with pytest.raises(pgw.PGWError):
        pgw.read_checkpoint(q, r)


def _oracle_triggers(orc, abi, seg, r, opts):
    S = orc.Search(seg, nthreads=8)
    out, fnum = {}, 0
    for pm in range(r.pmr[0], r.pmr[1] + 1):
        recs = []
        for mm in range(r.mr[0], r.mr[1] + 1):
            for nn in range(r.nr[0], r.nr[1] + 1):
                rec, fn = S.job_core(pm, mm, nn, r, opts)
                recs.append(rec)
                fnum += fn
        out[pm] = np.concatenate(recs) if recs else np.zeros((0, 5))
    return out, fnum


def _read(path):
    return np.fromfile(path, dtype=np.float64).reshape(-1, 5)


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a[:, :4], b[:, :4])
    Fa, Fb = a[:, 4] ** 2 / 2 + 2, b[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fa - Fb) / Fb, initial=0) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("every", [2, 64])
def test_search_trigger_files(pgw, abi, orc, tmp_path, every):
    """Both hemispheres of a 3x3 sky patch: one file per hemisphere whose records are the
    oracle's job_core records in the reference's (pm, mm, nn, ss, bin) order."""
    seg = segment("net-2det-32k")
    opts = abi.make_opts(trl=12.0)
    ref, fn_ref = _oracle_triggers(orc, abi, seg, _range(abi, seg), opts)
    h = pgw.Handle(seg, abi.PREC_FP64)
    fn = h.search(_range(abi, seg), opts, tmp_path, ident=1, band=666, every=every)
    assert fn == fn_ref
    for pm in (1, 2):
        f = tmp_path / f"triggers_001_0666_{pm}.bin"
        assert f.exists()
        _same(_read(f), ref[pm])
    print("records", {pm: len(ref[pm]) for pm in ref}, flush=True)


@pytest.mark.gpu
def test_search_resume(pgw, abi, tmp_path):
    """A run stopped after 5 sky points and resumed from its state file writes the same trigger
    files and FNum as an uninterrupted run; the state file ends empty."""
    seg = segment("net-2det-32k")
    opts = abi.make_opts(trl=12.0)
    h = pgw.Handle(seg, abi.PREC_FP64)
    full = tmp_path / "full"
    full.mkdir()
    fn_full = h.search(_range(abi, seg), opts, full, every=3)
    part = tmp_path / "part"
    part.mkdir()
    q = part / "state_001_0666.dat"
    fn = h.search(_range(abi, seg), opts, part, qname=q, every=3, max_sky=5)
    st, fn_state = pgw.read_checkpoint(q, r := _range(abi, seg))
    assert st == 0 and fn_state == fn
    # resumed twice more, the second time across the hemisphere boundary
    for _ in range(2):
        fn = h.search(r, opts, part, qname=q, every=3, max_sky=7, FNum=fn_state)
        st, fn_state = pgw.read_checkpoint(q, r := _range(abi, seg))
        if st == 1:
            break
    if st == 0:
        fn = h.search(r, opts, part, qname=q, every=3, FNum=fn_state)
    assert pgw.read_checkpoint(q, _range(abi, seg))[0] == 1
    assert fn == fn_full
    for pm in (1, 2):
        _same(_read(part / f"triggers_001_0666_{pm}.bin"), _read(full / f"triggers_001_0666_{pm}.bin"))


@pytest.mark.gpu
def test_float_records(pgw, abi, tmp_path):
    """COMP_FLOAT drivers (FLOAT_TYPE = float, cpu/auxi.h:16-23): pgw_job_core_f, pgw_job_batch_f and
    float trigger files carry each field of the double records rounded once to float, with the
    same record count, order and FNum."""
    from polgraw_allsky_amd import io
    seg = segment("net-2det-32k")
    opts = abi.make_opts(trl=12.0)
    r = _range(abi, seg)
    h = pgw.Handle(seg, abi.PREC_FP64)
    sky = [(1, mm, nn) for mm in range(r.mr[0], r.mr[1] + 1) for nn in range(r.nr[0], r.nr[1] + 1)]
    d, fd = h.job_batch(sky, r, opts)
    f, ff = h.job_batch_f(sky, r, opts)
    assert ff == fd and f.dtype == np.float32 and len(d) > 0
    assert np.array_equal(f, d.astype(np.float32))
    d1, fd1 = h.job_core(*sky[4], r, opts)
    f1, ff1 = h.job_core_f(*sky[4], r, opts)
    assert ff1 == fd1 and np.array_equal(f1, d1.astype(np.float32))
    # a record buffer one short: the count needed, as with double records
    with pytest.raises(pgw.PGWError) as e:
        h.job_batch_f(sky, r, opts, cap=len(d) - 1)
    assert e.value.status == abi.PGW_ERR_CAPACITY
    for fl, sub in ((False, "d"), (True, "f")):
        (tmp_path / sub).mkdir()
        h.search(_range(abi, seg), opts, tmp_path / sub, every=4, float_records=fl)
    for pm in (1, 2):
        dd = io.read_triggers(tmp_path / "d" / f"triggers_001_0666_{pm}.bin")
        ff_ = io.read_triggers(tmp_path / "f" / f"triggers_001_0666_{pm}.bin", dtype=np.float32)
        assert np.array_equal(ff_, dd.astype(np.float32))
    print("records", len(d), flush=True)
