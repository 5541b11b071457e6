"""The persistent fused kernel's job order (fused_job in pgw_fused.hpp), compiled for the host with
hipcc and run on the CPU: over many launch shapes (templates per launch, column tiles, row jobs,
threshold jobs, lag), every job of every template is handed out exactly once, and each template's
column jobs all come before its row jobs, which all come before its threshold jobs, so every
dependency wait of the kernel is on jobs with lower tickets (which is what makes the waits end)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = "/opt/rocm/bin/hipcc"

SRC = r'''
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <vector>
#include "pgw_fused.hpp"
using namespace pgw;
int main() {
  long bad = 0, cases = 0;
  for (int nt : {1, 2, 3, 7, 16, 64, 128, 255, 256})
    for (int ct : {1, 16, 32, 64})
      for (int rj : {1, 16, 32, 64})
        for (int tj : {1, 4, 16})
          for (int lag : {0, 1, 8, 64}) {
            FusedArgs a{};
            a.nt = nt; a.ct = ct; a.rj = rj; a.tj = tj; a.D = std::min(lag, nt / 2);
            a.o1 = a.D * a.ct;
            a.o2 = a.o1 + a.D * (a.ct + a.rj);
            a.o3 = a.o2 + (a.nt - 2 * a.D) * (a.ct + a.rj + a.tj);
            a.o4 = a.o3 + a.D * (a.rj + a.tj);
            a.total = a.o4 + a.D * a.tj;
            const int per = ct + rj + tj;
            std::vector<int> seen((size_t)nt * per, -1);
            std::vector<int> lastcol(nt, -1), lastrow(nt, -1), firstrow(nt, 1 << 30), firstthr(nt, 1 << 30);
            if (a.total != nt * per) bad++;
            for (int j = 0; j < a.total; j++) {
              int kind, t, i;
              fused_job(a, j, kind, t, i);
              const int lim = kind == 0 ? ct : kind == 1 ? rj : tj;
              if (kind < 0 || kind > 2 || t < 0 || t >= nt || i < 0 || i >= lim) { bad++; continue; }
              const size_t id = (size_t)t * per + (kind == 0 ? 0 : kind == 1 ? ct : ct + rj) + i;
              if (seen[id] >= 0) bad++;
              seen[id] = j;
              if (kind == 0) lastcol[t] = std::max(lastcol[t], j);
              if (kind == 1) { lastrow[t] = std::max(lastrow[t], j); firstrow[t] = std::min(firstrow[t], j); }
              if (kind == 2) firstthr[t] = std::min(firstthr[t], j);
            }
            for (int v : seen) if (v < 0) bad++;
            for (int t = 0; t < nt; t++) if (!(lastcol[t] < firstrow[t] && lastrow[t] < firstthr[t])) bad++;
            cases++;
          }
  printf("%ld %ld\n", cases, bad);
  return bad != 0;
}
'''


@pytest.mark.skipif(not Path(HIPCC).exists() or shutil.which("g++") is None, reason="hipcc not installed")
def test_fused_job_order(tmp_path):
    src = tmp_path / "jobs.cpp"
    src.write_text(SRC)
    exe = tmp_path / "jobs"
    r = subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", str(ROOT / "polgraw-allsky_amd" / "csrc"), "-I",
                        str(ROOT / "include"), "-o", str(exe), str(src)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    cases, bad = map(int, out.stdout.split())
    assert out.returncode == 0 and bad == 0 and cases == 1728, out.stdout
