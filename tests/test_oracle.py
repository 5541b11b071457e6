"""Pins the CPU oracle (oracle/pgw_oracle.c) before it is trusted as the checker.

Each function is checked against something independent of its own text:
  * the reference's own auxi.c compiled by oracle/Makefile into oracle/_ref (bit-exact);
  * library routines: numpy.fft (pocketfft), scipy.interpolate.CubicSpline, numpy var;
  * closed forms and brute force: DFT of deltas/cisoids, band-limited resampling at
    half-sample shifts, detector-tensor antenna patterns, FStat block means, run scans.
The reference's job_core needs FFTW3 and GSL, which this image lacks, so it is not built; its
composition is pinned by the closed-form resampling test and by grid-exact injection recovery.
"""
import ctypes as C
import math

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

_dp = C.POINTER(C.c_double)


def dp(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("reference tree not mounted: oracle/_ref not built")
    return r


# ------------------------------------------------------------------ FFT --
@pytest.mark.parametrize("logn", [1, 2, 3, 5, 8, 11, 13, 16, 17])
def test_fft_matches_numpy(orc, logn):
    n = 1 << logn
    rng = np.random.default_rng(logn)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = orc.fft(x, -1)
    b = orc.fft(x, +1)
    scale = np.abs(np.fft.fft(x)).max()
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-14 * logn * scale
    assert np.abs(b - n * np.fft.ifft(x)).max() <= 1e-14 * logn * scale


def test_fft_closed_forms(orc):
    n = 4096
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.abs(orc.fft(d) - 1).max() < 1e-15
    j = 123
    cis = np.exp(2j * np.pi * j * np.arange(n) / n)
    f = orc.fft(cis, -1)
    expect = np.zeros(n, complex)
    expect[j] = n
    assert np.abs(f - expect).max() < 1e-9
    # forward sign convention: a delay by one sample multiplies bin k by exp(-2 pi i k/n)
    e = np.zeros(n, complex)
    e[1] = 1
    assert np.abs(orc.fft(e, -1) - np.exp(-2j * np.pi * np.arange(n) / n)).max() < 1e-14


# --------------------------------------------------------------- auxi.c --
def test_lin2ast_matches_reference(orc, ref):
    rng = np.random.default_rng(3)
    sepsm, cepsm = math.sin(0.40909280422232897), math.cos(0.40909280422232897)
    for _ in range(200):
        r, th = math.sqrt(rng.uniform()) * 0.999, rng.uniform(0, 2 * math.pi)
        be1, be2, pm = r * math.cos(th), r * math.sin(th), int(rng.integers(1, 3))
        a = [C.c_double() for _ in range(4)]
        b = [C.c_double() for _ in range(4)]
        orc.lib().orc_lin2ast(be1, be2, pm, sepsm, cepsm, *[C.byref(v) for v in a])
        ref.lin2ast(be1, be2, pm, sepsm, cepsm, *[C.byref(v) for v in b])
        assert [v.value for v in a] == [v.value for v in b]


def test_lin2ast_ast2lin_round_trip(orc, ref):
    """ast2lin (reference) then lin2ast (oracle) returns the sky position (closed form).
    The reference's ast2lin decides the hemisphere with a 10*DBL_EPSILON test (auxi.c:47) that
    rounding often fails, so the hemisphere here is the one that reproduces delta."""
    epsm = 0.40909280422232897
    rng = np.random.default_rng(4)
    for _ in range(200):
        alpha, delta = rng.uniform(0, 2 * math.pi), math.asin(rng.uniform(-0.99, 0.99))
        be = np.zeros(2)
        ref.ast2lin(alpha, delta, epsm, dp(be))
        best = None
        for pm in (1, 2):
            s = [C.c_double() for _ in range(4)]
            orc.lib().orc_lin2ast(be[0], be[1], pm, math.sin(epsm), math.cos(epsm), *[C.byref(v) for v in s])
            vals = [v.value for v in s]
            if best is None or abs(vals[2] - math.sin(delta)) < abs(best[2] - math.sin(delta)):
                best = vals
        sinal, cosal, sindel, cosdel = best
        assert abs(sindel - math.sin(delta)) < 1e-12
        assert abs(math.atan2(sinal, cosal) % (2 * math.pi) - alpha) < 1e-9


@pytest.mark.parametrize("n", [3, 4, 17, 1000, 65536])
def test_spline_matches_reference(orc, ref, n):
    rng = np.random.default_rng(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y2o = np.zeros(n, complex)
    u = np.zeros(n, complex)
    y2r = np.zeros(n, complex)
    orc.lib().orc_spline(dp(y.view(np.float64)), n, dp(y2o.view(np.float64)), dp(u.view(np.float64)))
    ref.spline(dp(y.view(np.float64)), n, dp(y2r.view(np.float64)))
    assert np.array_equal(y2o, y2r)


def test_spline_is_natural_cubic_spline(orc):
    """Library routine: scipy's natural cubic spline on unit knots has the same second derivatives."""
    from scipy.interpolate import CubicSpline
    n = 5000
    rng = np.random.default_rng(7)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y2 = np.zeros(n, complex)
    u = np.zeros(n, complex)
    orc.lib().orc_spline(dp(y.view(np.float64)), n, dp(y2.view(np.float64)), dp(u.view(np.float64)))
    cs_r = CubicSpline(np.arange(n), y.real, bc_type="natural")
    cs_i = CubicSpline(np.arange(n), y.imag, bc_type="natural")
    d2 = cs_r(np.arange(n), 2) + 1j * cs_i(np.arange(n), 2)
    assert np.abs(y2 - d2).max() < 1e-10 * np.abs(d2).max()


@pytest.mark.parametrize("N", [100, 4096, 30001])
def test_splintpad_matches_reference(orc, ref, N):
    rng = np.random.default_rng(N)
    ya = rng.standard_normal(2 * N + 8) + 1j * rng.standard_normal(2 * N + 8)
    shftf = rng.uniform(-12, 12, N)
    shftf[:3] = [0.0, 0.5, -0.25]
    shftf[-2:] = [0.4, -0.6]         # samples that fall outside [0, 2N-1] -> exactly 0
    out_o = np.zeros(N, complex)
    out_r = np.zeros(N, complex)
    orc.lib().orc_splintpad(dp(ya.view(np.float64)), dp(shftf), N, 2, dp(out_o.view(np.float64)))
    ref.splintpad(dp(ya.view(np.float64)), dp(shftf), N, 2, dp(out_r.view(np.float64)))
    assert np.array_equal(out_o, out_r)
    x = 2 * (np.arange(N) - shftf)
    assert np.all(out_o[(x < 0) | (x > 2 * N - 1)] == 0)


def test_splint_reproduces_linear_functions(orc):
    """Natural cubic splines are exact on linear data (closed form)."""
    N = 2000
    t = np.arange(2 * N + 4, dtype=float)
    ya = (0.3 - 0.01j) + (0.002 + 0.004j) * t
    shftf = np.random.default_rng(1).uniform(-3, 3, N)
    out = np.zeros(N, complex)
    orc.lib().orc_splintpad(dp(ya.view(np.float64)), dp(shftf), N, 2, dp(out.view(np.float64)))
    x = 2 * (np.arange(N) - shftf)
    ok = (x >= 0) & (x <= 2 * N - 1)
    assert np.abs(out[ok] - ((0.3 - 0.01j) + (0.002 + 0.004j) * x[ok])).max() < 1e-11


def test_var_matches_reference_and_numpy(orc, ref):
    x = np.random.default_rng(2).standard_normal(10001) * 1.7 + 0.3
    x[::7] = 0.0
    assert orc.lib().orc_var(dp(x), len(x)) == ref.var(dp(x), len(x))
    assert orc.lib().orc_var(dp(x), len(x)) == pytest.approx(np.var(x, ddof=1), rel=1e-12)


@pytest.mark.parametrize("M", [4096, 63488, 126976, 4095, 9000])
def test_fstat_norm_matches_reference(orc, ref, M):
    F = np.random.default_rng(M).exponential(2.0, M) + 0.01
    a, b = F.copy(), F.copy()
    orc.lib().orc_fstat_norm(dp(a), M, 4096)
    ref.FStat(dp(b), M, 4096, 0)
    assert np.array_equal(a, b)
    nblk = M // 4096
    for j in range(nblk):   # every full block is normalised to mean 2 (closed form)
        assert a[4096 * j:4096 * (j + 1)].mean() == pytest.approx(2.0, rel=1e-12)
    assert np.array_equal(a[4096 * nblk:], F[4096 * nblk:])  # the tail stays raw


# -------------------------------------------------------------- modvir --
def test_modvir_matches_detector_tensor(orc, synth):
    """a^2 + b^2 equals F+^2 + Fx^2 of the detector tensor (JKS98, arms at 90 degrees), computed
    here from the site geometry: arms at gamma +- 45 deg counter-clockwise from local East,
    rotated by the local sidereal angle phir + omr t."""
    omr = 7.2921151467064e-5 * 2.0
    t = np.arange(0, 43200, 37.0)
    cm, sm = np.cos(omr * t), np.sin(omr * t)
    rng = np.random.default_rng(5)
    for site in ("H1", "L1", "V1"):
        lat, _, gam = synth.SITES[site]
        c = np.zeros(9)
        orc.lib().orc_rogcvir(lat, gam, dp(c))
        for _ in range(5):
            alpha, delta, phir = rng.uniform(0, 2 * np.pi), math.asin(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
            aa, bb = np.zeros(len(t)), np.zeros(len(t))
            orc.lib().orc_modvir(math.sin(alpha), math.cos(alpha), math.sin(delta), math.cos(delta), math.sin(phir),
                                 math.cos(phir), dp(c), len(t), dp(cm), dp(sm), dp(aa), dp(bb))
            lst = phir + omr * t
            E = np.stack([-np.sin(lst), np.cos(lst), 0 * lst], -1)
            Nn = np.stack([-math.sin(lat) * np.cos(lst), -math.sin(lat) * np.sin(lst), math.cos(lat) + 0 * lst], -1)
            u = math.cos(gam + np.pi / 4) * E + math.sin(gam + np.pi / 4) * Nn
            v = math.cos(gam - np.pi / 4) * E + math.sin(gam - np.pi / 4) * Nn
            k = np.array([math.cos(delta) * math.cos(alpha), math.cos(delta) * math.sin(alpha), math.sin(delta)])
            p = np.array([math.sin(alpha), -math.cos(alpha), 0.0])
            q = np.cross(k, p)
            up, uq, vp, vq = u @ p, u @ q, v @ p, v @ q
            Fp = 0.5 * (up ** 2 - uq ** 2 - vp ** 2 + vq ** 2)
            Fx = up * uq - vp * vq
            assert np.abs(aa ** 2 + bb ** 2 - (Fp ** 2 + Fx ** 2)).max() < 1e-13
            # same polarisation frame up to handedness and a constant rotation 2 psi0:
            # (a + ib) / (F+ - iFx) is one constant unit number for all t
            z = Fp - 1j * Fx
            ratio = (aa + 1j * bb)[np.abs(z) > 0.1] / z[np.abs(z) > 0.1]
            assert np.abs(np.abs(ratio) - 1).max() < 1e-9
            assert np.abs(ratio / ratio[0] - 1).max() < 1e-9


# ------------------------------------------------------------ resampling --
class FFTPlan(C.Structure):
    _fields_ = [("n", C.c_int), ("tw", C.c_void_p), ("work", C.c_void_p)]


def resample(orc, xa, shftf, nfft):
    N = len(xa)
    L = orc.lib()
    L.orc_resample.argtypes = [C.c_void_p, C.c_void_p] + [_dp] * 3 + [C.c_int] * 3 + [_dp] * 5
    p1, p2 = FFTPlan(), FFTPlan()
    assert L.orc_fft_plan_init(C.byref(p1), nfft) == 0 and L.orc_fft_plan_init(C.byref(p2), 2 * nfft) == 0
    wa, wb, wk = (np.zeros(2 * nfft, complex) for _ in range(3))
    oa, ob = np.zeros(N, complex), np.zeros(N, complex)
    xa = np.ascontiguousarray(xa, complex)
    xb = (1j * xa).copy()
    L.orc_resample(C.byref(p1), C.byref(p2), dp(xa.view(float)), dp(xb.view(float)), dp(shftf), N, nfft, 2,
                   dp(wa.view(float)), dp(wb.view(float)), dp(wk.view(float)), dp(oa.view(float)),
                   dp(ob.view(float)))
    L.orc_fft_plan_free(C.byref(p1))
    L.orc_fft_plan_free(C.byref(p2))
    assert np.abs(ob - 1j * oa).max() < 1e-12 * max(1, np.abs(oa).max())
    return oa


@pytest.mark.parametrize("k0", [0, 1, 37, 2047, 2048, -1, -2047])
def test_resampling_band_limited_closed_form(orc, k0):
    """For a cisoid on the bin grid (no zero padding), FFT -> zero-insert -> inverse FFT -> scale is
    exact band-limited interpolation, so a half-sample shift reproduces exp(2 pi i k0 (i-1/2)/nfft)
    exactly at every spline knot (jobcore.c:319-370; the Nyquist bin nfft/2 counts as positive)."""
    nfft = 4096
    N = nfft
    i = np.arange(N)
    x = np.exp(2j * np.pi * k0 * i / nfft)
    shftf = np.full(N, 0.5)
    out = resample(orc, x, shftf, nfft)
    k = k0 if k0 <= nfft // 2 else k0 - nfft
    expect = np.exp(2j * np.pi * k * (i - 0.5) / nfft)
    expect[0] = 0.0   # abscissa 2*(0 - 0.5) < 0 -> exactly 0
    assert np.abs(out - expect).max() < 1e-11
    # zero shift returns the input samples themselves
    out0 = resample(orc, x, np.zeros(N), nfft)
    assert np.abs(out0 - x).max() < 1e-11


def test_resampling_general_shift_is_spline_of_band_limited_interpolant(orc):
    """Library routine: for zero-padded data the output equals scipy's natural cubic spline through
    the band-limited (trigonometric) interpolant at half-sample knots."""
    from scipy.interpolate import CubicSpline
    N, nfft = 3000, 4096
    rng = np.random.default_rng(11)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    shftf = rng.uniform(-5, 5, N)
    out = resample(orc, x, shftf, nfft)
    X = np.fft.fft(np.concatenate([x, np.zeros(nfft - N)]))
    k = np.fft.fftfreq(nfft, 1.0 / nfft)
    k[nfft // 2] = nfft // 2
    knots = np.arange(2 * N) / 2.0
    yk = (np.exp(2j * np.pi * np.outer(knots, k) / nfft) @ X) / nfft
    cr = CubicSpline(np.arange(2 * N), yk.real, bc_type="natural")
    ci = CubicSpline(np.arange(2 * N), yk.imag, bc_type="natural")
    xs = 2 * (np.arange(N) - shftf)
    ok = (xs >= 0) & (xs <= 2 * N - 1)
    expect = np.where(ok, cr(np.clip(xs, 0, 2 * N - 1)) + 1j * ci(np.clip(xs, 0, 2 * N - 1)), 0)
    assert np.abs(out - expect).max() < 1e-9 * np.abs(expect).max()


# ------------------------------------------------------------- threshold --
def runmax_bruteforce(F, nmin, nmax, trl):
    """Independent formulation: maximal runs of F > trl via edges, ii = last index of the run maximum."""
    above = np.zeros(len(F) + 2, bool)
    above[nmin + 1:nmax + 1] = F[nmin:nmax] > trl
    edges = np.flatnonzero(np.diff(above.astype(int)))
    out = []
    for s, e in zip(edges[::2], edges[1::2]):
        seg = F[s:e]
        ii = s + len(seg) - 1 - int(np.argmax(seg[::-1]))
        out.append((ii, F[ii]))
    return out


def runmax_oracle(orc, F, nmin, nmax, trl):
    cap = len(F)
    ii = np.zeros(cap, np.int32)
    Fc = np.zeros(cap)
    L = orc.lib()
    L.orc_runmax.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _dp, C.c_int]
    n = L.orc_runmax(dp(F), nmin, nmax, trl, ii.ctypes.data_as(C.POINTER(C.c_int)), dp(Fc), cap)
    return list(zip(ii[:n].tolist(), Fc[:n].tolist()))


@settings(max_examples=300, deadline=None)
@given(st.lists(st.sampled_from([0.0, 5.0, 20.0, 20.5, 21.0, 30.0, 30.0, 44.0]), min_size=1, max_size=60),
       st.integers(0, 5), st.integers(0, 5))
def test_runmax_matches_bruteforce(orc, vals, lo_cut, hi_cut):
    F = np.array(vals, float)
    nmin, nmax = min(lo_cut, len(F)), max(min(lo_cut, len(F)), len(F) - hi_cut)
    assert runmax_oracle(orc, F, nmin, nmax, 20.0) == runmax_bruteforce(F, nmin, nmax, 20.0)


# -------------------------------------------------------- spindown range --
def test_spindown_range_property(orc, abi, synth):
    """Without a range file every spindown in [-Smax, Smin] is covered up to one grid step
    (jobcore.c:424-439)."""
    seg = synth.make_segment(N=8192, dets=("H1",), inject=False, nsky=(1, 1), nspin=1)
    sett = abi.make_settings(seg)
    rngs = abi.make_range(seg)
    opts = abi.make_opts(use_range=False)
    M = seg.M
    lo, hi = C.c_int(), C.c_int()
    for nn in range(-3, 4):
        for mm in range(-3, 4):
            orc.lib().orc_spindown_range(C.byref(sett), mm, nn, C.byref(rngs), C.byref(opts), C.byref(lo), C.byref(hi))
            ss = np.arange(lo.value, hi.value + 1)
            sd = ss * M[5] + nn * M[9] + mm * M[13]
            step = abs(M[5])
            assert sd.min() <= -seg.sett["Smax"] + step and sd.max() >= seg.sett["Smin"] - step
            assert sd.min() >= -seg.sett["Smax"] - step and sd.max() <= seg.sett["Smin"] + step


@pytest.mark.parametrize("sign", [-1.0, 1.0])
@pytest.mark.parametrize("s0", [False, True])
def test_spindown_range_endpoints(orc, abi, synth, sign, s0):
    """Endpoints of jobcore.c:424-439 at hand-derived values: with M5 = sign * 2^-30 and the sky
    terms nn M9 + mm M13 = 1.25 * 2^-30 (nn = 3, mm = 1), Smin = 3.7 * 2^-30 and Smax = 12.3 * 2^-30
    (every quotient exact up to the rounding of 3.7 and 12.3), the raw bounds are
    (Smin - 1.25) / M5 = -2.45 sign and -(1.25 + Smax) / M5 = -13.55 sign; truncation toward zero
    gives -2 and 13 for M5 < 0, and 2 and -13 swapped into [-13, 2] for M5 > 0 (floor would give
    -3 or -14); --nospindown keeps only the upper one."""
    seg = synth.make_segment(N=8192, dets=("H1",), inject=False, nsky=(1, 1), nspin=1)
    u = 2.0 ** -30
    seg.M[5], seg.M[9], seg.M[13] = sign * u, 0.25 * u, 0.5 * u
    seg.sett["Smin"], seg.sett["Smax"] = 3.7 * u, 12.3 * u
    sett = abi.make_settings(seg)
    rngs = abi.make_range(seg)
    opts = abi.make_opts(use_range=False, s0=s0)
    lo, hi = C.c_int(), C.c_int()
    orc.lib().orc_spindown_range(C.byref(sett), 1, 3, C.byref(rngs), C.byref(opts), C.byref(lo), C.byref(hi))
    want = (-2, 13) if sign < 0 else (-13, 2)
    if s0:
        want = (want[1], want[1])
    assert (lo.value, hi.value) == want


def test_het0_sign_at_band_edge(orc, abi, synth):
    """het0 = fmod(nn M8 + mm M12, M0) keeps the sign of its argument (jobcore.c:264).  A noise-free
    signal on the template of bin nmin of a sky point whose argument is negative (het0 in (-M0, 0))
    is found at bin nmin with the template's frequency 2 pi nmin / L + het0 + het1; with het0
    wrapped into [0, M0) its peak would fall at nmin - 1, outside [nmin, nmax), and the run's
    record at nmin would carry a frequency one bin (M0) higher."""
    seg = synth.make_segment(N=32768, dets=("H1", "L1"), sigmas=(1.0, 1.0), fftpad=2, seed=21, inject=False,
                             nsky=(1, 1), nspin=3)
    s, M = seg.sett, seg.M
    pm, mm, nn, ss = 1, 0, -1, 4
    arg = nn * M[8] + mm * M[12]
    het0 = math.fmod(arg, M[0])
    assert arg < 0 and -M[0] < het0 < 0
    oms = s["oms"]
    al1, al2 = nn * M[10] + mm * M[14], nn * M[11] + mm * M[15]
    sv = [C.c_double() for _ in range(4)]
    orc.lib().orc_lin2ast(al1 / oms, al2 / oms, pm, math.sin(seg.epsm), math.cos(seg.epsm), *[C.byref(v) for v in sv])
    sinal, cosal, sindel, cosdel = (v.value for v in sv)
    het1 = math.fmod(ss * M[4], M[0]) % M[0]
    ii = s["nmin"]
    f = 2 * math.pi * ii / s["nfftf"] + het0 + het1
    fdot = ss * M[5] + nn * M[9] + mm * M[13]
    t = np.arange(s["N"], dtype=float)
    cm, sm = np.cos(s["omr"] * t), np.sin(s["omr"] * t)
    for k in range(2):
        aa, bb = synth.modvir(sinal, cosal, sindel, cosdel, math.sin(seg.phir[k]), math.cos(seg.phir[k]),
                              seg.amod[k], cm, sm)
        shift = seg.DetSSB[k].reshape(-1, 3) @ np.array([cosal * cosdel, sinal * cosdel, sindel])
        ph = f * t + fdot * t * t + (oms + f + 2 * fdot * t) * shift
        seg.xDat[k] = 0.08 * (aa * np.cos(ph) + bb * np.sin(ph))
        seg.sig2[k] = 1.0
    seg.range = dict(spndr=[ss, ss], nr=[nn, nn], mr=[mm, mm], pmr=[pm, pm])
    recs, fn = orc.Search(seg, nthreads=4).job_core(pm, mm, nn, abi.make_range(seg), abi.make_opts(trl=5.0, white=True))
    assert fn == 1 and len(recs) > 0
    best = recs[np.argmax(recs[:, 4])]
    assert abs(best[0] - f) < 1e-12
    # the record's bin is nmin: its frequency minus the template's offsets is 2 pi nmin / L
    assert round((best[0] - het0 - het1) * s["nfftf"] / (2 * math.pi)) == ii


# ----------------------------------------------------------- composition --
def test_openmp_equals_serial(orc, abi):
    from cases import injected_sky, segment
    seg = segment("net-2det-32k")
    rng, opts = abi.make_range(seg), abi.make_opts(trl=10.0)
    a, fa = orc.Search(seg, nthreads=1).job_core(*injected_sky(seg), rng, opts)
    b, fb = orc.Search(seg, nthreads=6).job_core(*injected_sky(seg), rng, opts)
    assert fa == fb and np.array_equal(a, b)


def test_grid_exact_injection_recovery(orc, abi, synth):
    """A noise-free signal placed exactly on a grid template (sky from lin2ast of the grid point,
    frequency on a bin, spindown on the spindown lattice) is found at that template: frequency to
    1e-12 rad/sample, the grid spindown and sky, and F = rho^2/2 within 1% (sigma^2 = 1, no FStat
    normalisation), where rho^2 = sum of the squared signal, computed directly from the injection."""
    seg = synth.make_segment(N=32768, dets=("H1", "L1"), sigmas=(1.0, 1.0), fftpad=2, seed=21, inject=False,
                             nsky=(1, 1), nspin=3)
    s, M = seg.sett, seg.M
    pm, mm, nn, ss, ii = 1, 2, -3, 5, 9000
    oms = s["oms"]
    al1, al2 = nn * M[10] + mm * M[14], nn * M[11] + mm * M[15]
    sv = [C.c_double() for _ in range(4)]
    orc.lib().orc_lin2ast(al1 / oms, al2 / oms, pm, math.sin(seg.epsm), math.cos(seg.epsm), *[C.byref(v) for v in sv])
    sinal, cosal, sindel, cosdel = (v.value for v in sv)
    het0 = math.fmod(nn * M[8] + mm * M[12], M[0])
    het1 = math.fmod(ss * M[4], M[0]) % M[0]
    f = 2 * math.pi * ii / s["nfftf"] + het0 + het1
    fdot = ss * M[5] + nn * M[9] + mm * M[13]
    t = np.arange(s["N"], dtype=float)
    cm, sm = np.cos(s["omr"] * t), np.sin(s["omr"] * t)
    h0, rho2 = 0.08, 0.0
    for k in range(2):
        aa, bb = synth.modvir(sinal, cosal, sindel, cosdel, math.sin(seg.phir[k]), math.cos(seg.phir[k]),
                              seg.amod[k], cm, sm)
        shift = seg.DetSSB[k].reshape(-1, 3) @ np.array([cosal * cosdel, sinal * cosdel, sindel])
        ph = f * t + fdot * t * t + (oms + f + 2 * fdot * t) * shift   # add_signal, init.c:607-615
        sig = h0 * (aa * np.cos(ph) + bb * np.sin(ph))
        seg.xDat[k] = sig
        seg.sig2[k] = 1.0
        rho2 += np.sum(sig ** 2)
    seg.range = dict(spndr=[ss - 1, ss + 1], nr=[nn, nn], mr=[mm, mm], pmr=[pm, pm])
    rngs = abi.make_range(seg)
    opts = abi.make_opts(trl=5.0, white=True)
    recs, fn = orc.Search(seg, nthreads=8).job_core(pm, mm, nn, rngs, opts)
    assert fn == 3
    best = recs[np.argmax(recs[:, 4])]
    assert abs(best[0] - f) < 1e-12
    assert best[1] == fdot
    assert best[2] == pytest.approx(math.asin(sindel), abs=1e-15)
    Fbest = best[4] ** 2 / 2 + 2
    assert Fbest == pytest.approx(rho2 / 2, rel=0.01)
