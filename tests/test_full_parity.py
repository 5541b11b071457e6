"""The configurations at their full shapes against the oracle's committed records.

CPU: each fixture (tests/golden/oracle/*.npz, written by tests/golden/make_oracle_fixtures.py)
is tied to the current oracle by re-running the oracle on its richest work item, bit-exact.

GPU: every work item of each case through the C ABI, FP64: records in the reference's order with
fields 0-3 bit-equal and F within 1e-9 relative, and the same template count; FP32 form against
the same FP64 records by the rule of cases.fp32_record_diff (matched F within 1e-4; a record in
one set only where F is within 1e-4 trl of the threshold or where a run's maximum moved between
bins whose F agree within 1e-4).  Cases (tests/full_cases.py): the plain configuration in full
(4 x 32), the network configuration in full (256 x 512 = 131072 templates), the precision
configuration's first 8 sky points x 256, and all 64 batching segments at one sky point x 512.
"""
from pathlib import Path

import numpy as np
import pytest

from cases import fp32_record_diff
from full_cases import FULL_CASES, case_items

FIX = Path(__file__).resolve().parent / "golden" / "oracle"


def load(name):
    d = np.load(FIX / f"{name}.npz")
    return d["work"], d["records"], d["counts"], d["fnum"]


def per_item(records, counts):
    off = np.concatenate([[0], np.cumsum(counts)])
    return [records[off[i]:off[i + 1]] for i in range(len(counts))]


@pytest.mark.parametrize("name", FULL_CASES)
def test_fixture_is_the_oracle(name, abi, orc):
    work, records, counts, fnum = load(name)
    segs, items = case_items(name)
    assert len(items) == len(work) == len(counts) == len(fnum)
    for (k, pm, mm, nn, rng), w in zip(items, work):
        assert (k, pm, mm, nn, rng.spndr[0], rng.spndr[1]) == tuple(int(v) for v in w)
    i = int(np.argmax(counts))
    k, pm, mm, nn, rng = items[i]
    r, fn = orc.Search(segs[k], nthreads=8).job_core(pm, mm, nn, rng, abi.make_opts(trl=20.0))
    print(name, "items", len(items), "records", int(counts.sum()), "checked item", i, "with", len(r), flush=True)
    assert fn == fnum[i]
    assert np.array_equal(r, per_item(records, counts)[i])
    assert counts.sum() > 0


def _run_device(name, prec):
    from polgraw_allsky_amd import abi, pgw
    segs, items = case_items(name)
    h = pgw.Handle(segs, prec)
    opts = abi.make_opts(trl=20.0)
    out, fns = [], []
    if name == "batching-64seg":
        # one call per segment: each segment's own spindown range
        for k, pm, mm, nn, rng in items:
            r, _, f = h.job_batch_multi([(k, pm, mm, nn)], rng, opts)
            out.append(r)
            fns.append(f)
    else:
        rng = items[0][4]
        sky = [(pm, mm, nn) for _, pm, mm, nn, _ in items]
        for i in range(0, len(sky), 64):
            r, f = h.job_batch(sky[i:i + 64], rng, opts)
            out.append(r)
            fns.append(f)
    h.close()
    return np.concatenate(out), int(np.sum(fns))


@pytest.mark.gpu
@pytest.mark.parametrize("name", FULL_CASES)
def test_full_fp64(name, abi):
    work, ref, counts, fnum = load(name)
    got, fn = _run_device(name, abi.PREC_FP64)
    print(name, "templates", fn, "records", len(got), "oracle", len(ref), flush=True)
    assert fn == int(fnum.sum())
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :4], ref[:, :4])
    Fg, Fr = got[:, 4] ** 2 / 2 + 2, ref[:, 4] ** 2 / 2 + 2
    assert np.max(np.abs(Fg - Fr) / Fr) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("name", FULL_CASES)
def test_full_fp32(name, abi):
    work, ref, counts, fnum = load(name)
    got, fn = _run_device(name, abi.PREC_FP32)
    rel, nmatch, allowed, viol = fp32_record_diff(got, ref, 20.0)
    print(name, "records fp32", len(got), "fp64 oracle", len(ref), "matched", nmatch, "max rel dF", rel,
          "allowed", allowed, "violations", viol[:5], flush=True)
    assert fn == int(fnum.sum())
    assert not viol
    assert nmatch >= len(ref) - len(allowed)
