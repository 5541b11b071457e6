"""Copies the reference's own test inputs into tests/golden/ref-001 (data only).

Source: testdata/test-gpu-cpu/001 (Gaussian-noise H1/L1 frame 001, band 666: grid.bin, and per
detector xdatc_001_0666.bin and DetSSB.bin) and search/network/src-gpu/range.dat (the search range
the reference's GPU test uses: spindowns -40..40, nn -1..1, mm -2..2, hemisphere 1).
Run in a tree where the reference is mounted at /root/reference:

    python3 tests/golden/copy_reference_fixture.py
"""
import shutil
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "ref-001"


def main():
    src = REF / "testdata" / "test-gpu-cpu" / "001"
    dst = OUT / "001"
    for rel in ("grid.bin", "H1/xdatc_001_0666.bin", "H1/DetSSB.bin", "L1/xdatc_001_0666.bin", "L1/DetSSB.bin"):
        (dst / rel).parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(src / rel, dst / rel)
    shutil.copyfile(REF / "search" / "network" / "src-gpu" / "range.dat", OUT / "range.dat")
    print("copied into", OUT)


if __name__ == "__main__":
    main()
