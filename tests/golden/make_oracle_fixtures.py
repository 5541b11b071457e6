"""Writes the oracle's candidate records for the full-size parity cases to tests/golden/oracle/.

    make -C oracle && python tests/golden/make_oracle_fixtures.py [case ...]

The oracle is the project's FP64 C restatement of the reference CPU job_core (oracle/), the same
code the CPU tests run; these fixtures spare the GPU tests its minutes of CPU time at the full
sizes (the network case alone is ~130k templates).  Each file holds, per work item (segment, pm,
mm, nn, first and last spindown), the records the oracle appends and its template count;
tests/test_full_parity.py re-runs the oracle on one item of each file to tie the file to the
current oracle, and compares the device's records for every item with it.
"""
from __future__ import annotations

import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
import pgwload  # noqa: E402

pgwload.load()
import orc  # noqa: E402
from full_cases import FULL_CASES, case_items  # noqa: E402
from polgraw_allsky_amd import abi  # noqa: E402

OUT = HERE / "oracle"


def make(name: str, threads: int) -> Path:
    segs, items = case_items(name)
    opts = abi.make_opts(trl=20.0)
    searches = {}
    recs, counts, fnums = [], [], []
    t0 = time.time()
    for i, (k, pm, mm, nn, rng) in enumerate(items):
        if k not in searches:
            searches[k] = orc.Search(segs[k], nthreads=threads)
        r, fn = searches[k].job_core(pm, mm, nn, rng, opts)
        recs.append(np.array(r, dtype=np.float64).reshape(-1, 5))
        counts.append(len(r))
        fnums.append(fn)
        if i % 16 == 15:
            print(f"{name}: {i + 1}/{len(items)} items, {sum(counts)} records, {time.time() - t0:.0f} s", flush=True)
    work = np.array([(k, pm, mm, nn, rng.spndr[0], rng.spndr[1]) for k, pm, mm, nn, rng in items], dtype=np.int64)
    path = OUT / f"{name}.npz"
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(path, work=work, records=np.concatenate(recs) if recs else np.zeros((0, 5)),
                        counts=np.array(counts, dtype=np.int64), fnum=np.array(fnums, dtype=np.int64))
    print(f"{name}: {len(items)} items, {sum(fnums)} templates, {sum(counts)} records in {time.time() - t0:.0f} s -> "
          f"{path}", flush=True)
    return path


if __name__ == "__main__":
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    for name in sys.argv[1:] or list(FULL_CASES):
        make(name, threads)
