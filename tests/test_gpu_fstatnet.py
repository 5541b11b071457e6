"""Differentiation form: the device Fstatnet (followup.c:413-534) and its analytic gradient against
the oracle, torch gradcheck through the autograd wrapper, and a 20-step gradient ascent against
the followup's Nelder-Mead (followup.c:1126-1173, restated in the oracle)."""
import numpy as np
import pytest

from cases import injected_sky, noise_free_signal, phase_scales, segment

pytestmark = pytest.mark.gpu

# FP64 on both sides; the sums run in different orders (tree vs sequential), 172800 terms
F_RTOL = 1e-10
G_RTOL = 1e-8


@pytest.fixture(scope="module")
def pgw():
    from polgraw_allsky_amd import pgw as p
    p.lib()
    return p


def _points(seg, n, seed, spread=2.0):
    inj = seg.injection
    truth = np.array([inj["f"], inj["fdot"], inj["delta"], inj["alpha"]])
    sc = phase_scales(seg, truth)
    rng = np.random.default_rng(seed)
    return truth + rng.uniform(-spread, spread, (n, 4)) * sc, sc


def _compare(seg, pts, sc, out_g, g_g, out_o, g_o):
    F_g, F_o = -out_g[:, 5], -out_o[:, 5]
    assert np.max(np.abs(F_g - F_o) / F_o) < F_RTOL
    # the other outputs (amplitudes, snr, point) to the same relative accuracy of their scale;
    # out[4] = sqrt(2 (F - 2)) is NaN where F < 2, as in the reference, on both sides
    assert np.array_equal(np.isnan(out_g), np.isnan(out_o))
    og, oo = np.nan_to_num(out_g), np.nan_to_num(out_o)
    scale = np.max(np.abs(oo), axis=0) + 1e-300
    assert np.max(np.abs(og - oo) / scale) < 1e-8
    # gradient per unit scaled step, relative to F
    err = np.abs(g_g - g_o) * sc / F_o[:, None]
    assert np.max(err) < G_RTOL, np.max(err)


@pytest.mark.parametrize("name", ["net-2det-32k", "gaps-3det-30k"])
def test_fstatnet_matches_oracle(pgw, abi, orc, name):
    seg = segment(name)
    pts, sc = _points(seg, 64, seed=1)
    h = pgw.Handle(seg, abi.PREC_FP64)
    out_g, g_g = h.fstatnet(pts)
    out_o, g_o = orc.Search(seg, nthreads=8).fstatnet(pts)
    print(name, "F range", (-out_o[:, 5]).min(), (-out_o[:, 5]).max(), flush=True)
    _compare(seg, pts, sc, out_g, g_g, out_o, g_o)


def test_fstatnet_1024_points_network(pgw, abi, orc):
    """The differentiation configuration: 2 detectors, N = 86400, 1024 candidate points."""
    seg = segment("net-2det-86k")
    pts, sc = _points(seg, 1024, seed=2, spread=4.0)
    h = pgw.Handle(seg, abi.PREC_FP64)
    out_g, g_g = h.fstatnet(pts)
    print("1024 points: gpu ms", h.last_gpu_ms, flush=True)
    out_o, g_o = orc.Search(seg, nthreads=8).fstatnet(pts)
    _compare(seg, pts, sc, out_g, g_g, out_o, g_o)
    # without the gradient the values are the same to the last bits (the values-only kernel takes its
    # sin and cos from another FP64 implementation: 1.1e-16 absolute error against 1 ulp)
    out_n, g_n = h.fstatnet(pts, grad=False)
    assert g_n is None
    assert np.allclose(out_n, out_g, rtol=1e-12, atol=1e-12, equal_nan=True)


def test_fstatnet_empty_and_errors(pgw, abi):
    seg = segment("net-2det-32k")
    h = pgw.Handle(seg, abi.PREC_FP32)   # Fstatnet is FP64 in either handle precision
    out, g = h.fstatnet(np.zeros((0, 4)))
    assert out.shape == (0, 11)
    with pytest.raises(pgw.PGWError):
        h.fstatnet(np.zeros((1, 4)), seg=3)


def test_gradcheck_torch(pgw, abi):
    import torch
    from polgraw_allsky_amd import followup
    seg = segment("net-2det-32k")
    pts, sc = _points(seg, 3, seed=4, spread=1.0)
    h = pgw.Handle(seg, abi.PREC_FP64)
    fs = followup.FStatNet(h)
    base = torch.from_numpy(pts)
    scale = torch.from_numpy(np.broadcast_to(sc, pts.shape).copy())

    def f(u):
        return followup.fstat(base + u * scale, fs)

    u = torch.zeros_like(base, requires_grad=True)
    # one unit of u moves the phase by about a radian.  The FP64 phase (1e5..1e6 rad) is rounded
    # at ~1e-10 rad per sample, so steps of 1e-4 rad keep that rounding 1e-6 below the difference
    # while the O(eps^2) truncation stays below 1e-7 relative.
    assert torch.autograd.gradcheck(f, (u,), eps=1e-4, atol=1e-3, rtol=1e-5)
    # backward equals the device gradient
    Fv = f(u)
    Fv.sum().backward()
    _, g = h.fstatnet(pts)
    assert np.allclose(u.grad.numpy(), g * sc, rtol=1e-12, atol=0)


def test_gradient_ascent_vs_simplex(pgw, abi, orc):
    """From 4 points around an injected signal (2 detectors, N = 86400), 20 batched quasi-Newton
    steps with the device gradient reach the F of the followup's Nelder-Mead (run to tol 1e-10
    from the same starts) and F at the injection within 1e-5, with 21 device evaluations per point
    instead of the simplex's 100-250.  Near the maximum F is flat to 1e-6 along a combination of
    frequency and sky position, so the end points are compared through F.  (Over a segment much
    shorter than a day that direction is nearly degenerate and the simplex can drift along it.)"""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(N=86400, seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    sc = phase_scales(seg, truth)
    rng = np.random.default_rng(9)
    starts = truth + rng.uniform(-0.5, 0.5, (4, 4)) * sc
    h = pgw.Handle(seg, abi.PREC_FP64)
    G = np.stack([followup.phase_metric(seg.DetSSB, seg.sett["N"], seg.sett["oms"], p) for p in starts])
    x, F, nev = followup.ascend(followup.FStatNet(h), starts, G, steps=20)
    S = orc.Search(seg, nthreads=8)
    F0 = -S.fstatnet(truth[None, :], grad=False)[0][0, 5]
    for k in range(len(starts)):
        xn, out_n, nev_n = S.amoeba(starts[k], 0.5 * sc, tol=1e-10)
        print(f"start {k}: ascent F {F[k]:.9f} ({nev} evals), simplex F {-out_n[5]:.9f} ({nev_n} evals)",
              flush=True)
        assert F[k] >= -out_n[5] * (1 - 1e-5)
        assert F[k] >= F0 * (1 - 1e-5)


def test_batched_amoeba_on_device(pgw, abi, orc):
    """followup.amoeba fed by the device's Fstatnet, 8 simplices in each batched call, against the
    oracle's Nelder-Mead from the same starts: both converge, to F within 1e-8 relative (their F
    differ by ~1e-12 relative, which can change a near-tie decision and so the path)."""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(N=86400, seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    sc = phase_scales(seg, truth)
    starts = truth + np.random.default_rng(5).uniform(-0.8, 0.8, (8, 4)) * sc
    h = pgw.Handle(seg, abi.PREC_FP64)
    x, out, nev, conv = followup.amoeba(followup.FStatNet(h), starts, 0.5 * sc, tol=1e-9)
    assert conv.all()
    S = orc.Search(seg, nthreads=8)
    same = 0
    for k in range(len(starts)):
        xr, outr, nevr = S.amoeba(starts[k], 0.5 * sc, tol=1e-9)
        print(f"start {k}: device F {-out[k, 5]:.10f} ({nev[k]} evals), oracle F {-outr[5]:.10f} ({nevr})",
              flush=True)
        assert abs(out[k, 5] - outr[5]) <= 1e-8 * abs(outr[5])
        same += int(nev[k] == nevr)
    assert same >= len(starts) // 2


def test_batched_mads_on_device(pgw, abi, orc):
    """followup.mads fed by the device's Fstatnet (80 mesh points of every unfinished candidate
    per call) against the same search fed by the oracle: the same end points' F within 1e-8
    relative and the same number of rounds for most candidates (F differ by ~1e-12 relative, which
    can change a near-tie)."""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(N=86400, seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    sc = phase_scales(seg, truth)
    nod = seg.sett["N"] * seg.sett["dt"] / 86400.0
    starts = truth + np.random.default_rng(7).uniform(-0.8, 0.8, (6, 4)) * sc
    h = pgw.Handle(seg, abi.PREC_FP64)
    x, out, rounds = followup.mads(followup.FStatNet(h), starts, 2.0 * sc, nod)
    S = orc.Search(seg, nthreads=8)

    def ofs(pts, grad=False, owner=None):
        return S.fstatnet(np.asarray(pts).reshape(-1, 4), grad=False)[0], None

    xo, outo, ro = followup.mads(ofs, starts, 2.0 * sc, nod)
    print("rounds device", rounds, "oracle", ro, flush=True)
    assert np.all(np.abs(out[:, 5] - outo[:, 5]) <= 1e-8 * np.abs(outo[:, 5]))
    assert np.sum(rounds == ro) >= len(starts) // 2
    F0 = -h.fstatnet(starts, grad=False)[0][:, 5]
    assert np.all(-out[:, 5] > F0)
