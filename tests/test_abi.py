"""The C ABI: every symbol include/pgw.h declares is exported by the built library (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    txt = (ROOT / "include" / "pgw.h").read_text()
    return sorted(set(re.findall(r"\b(pgw_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("pgw_init", "pgw_job_core", "pgw_job_batch", "pgw_destroy"):
        assert n in names


def test_library_exports_all_declared_symbols():
    from polgraw_allsky_amd import pgw
    if not pgw.LIB_PATH.exists():
        pytest.fail("libpgw_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(pgw.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(pgw.EXPORTED) == declared()


def test_host_helpers_match_oracle(abi, orc):
    """Host-only helpers (no device work): settings, rogcvir, sig2 equal the oracle's bit for bit."""
    import numpy as np
    from polgraw_allsky_amd import pgw
    L = pgw.lib()
    for N, fftpad in ((86400, 2), (259200, 2), (60000, 1), (86164, 1)):
        a, b = abi.Settings(), abi.Settings()
        assert L.pgw_search_settings(171.3, 2.0, N, fftpad, ctypes.byref(a)) == 0
        assert orc.lib().orc_settings(171.3, 2.0, N, fftpad, ctypes.byref(b)) == 0
        assert bytes(a) == bytes(b)
    c1 = np.zeros(9)
    c2 = np.zeros(9)
    L.pgw_rogcvir(0.81, 2.98, c1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    orc.lib().orc_rogcvir(0.81, 2.98, c2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.array_equal(c1, c2)
    x = np.random.default_rng(0).standard_normal(1000)
    x[100:200] = 0
    dp = ctypes.POINTER(ctypes.c_double)
    crf0 = 1000 / 900
    assert L.pgw_sig2(x.ctypes.data_as(dp), 1000) == pytest.approx(crf0 * orc.lib().orc_var(x.ctypes.data_as(dp), 1000), rel=0, abs=0)
