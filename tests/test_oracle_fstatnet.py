"""Pins the oracle's Fstatnet (followup.c:413-534), its gradient and the Nelder-Mead followup.

Independent checks: the noise-free closed form (F = rho^2/2 and amplitude estimates = injected
amplitudes at the true point), central differences of F for the gradient, the grid search's F
at its grid candidate, and the simplex converging onto the injected parameters."""
import math

import numpy as np
import pytest

from cases import noise_free_signal, phase_scales
from polgraw_allsky_amd import synth


@pytest.fixture(scope="module")
def clean(orc):
    seg, truth, amp, rho2 = noise_free_signal()
    return seg, truth, amp, rho2, orc.Search(seg, nthreads=8)


def closed_form(seg, truth, amp):
    """Noise-free limit of Fstatnet at the true point: with A = sum a^2, B = sum b^2, C = sum a b
    over detectors (modvir at the true sky, sig2 = 1), Xa = (amp0 A + amp1 C - i (amp2 A + amp3 C))/2
    and Xb = (amp0 C + amp1 B - i (amp2 C + amp3 B))/2 up to the sums of cos 2phi, sin 2phi terms."""
    s = seg.sett
    t = np.arange(s["N"], dtype=float)
    cm, sm = np.cos(s["omr"] * t), np.sin(s["omr"] * t)
    d, al = truth[2], truth[3]
    A = B = Cab = 0.0
    for k in range(len(seg.dets)):
        a, b = synth.modvir(math.sin(al), math.cos(al), math.sin(d), math.cos(d), math.sin(seg.phir[k]),
                            math.cos(seg.phir[k]), seg.amod[k], cm, sm)
        A += np.sum(a * a)
        B += np.sum(b * b)
        Cab += np.sum(a * b)
    Xa = 0.5 * complex(amp[0] * A + amp[1] * Cab, -(amp[2] * A + amp[3] * Cab))
    Xb = 0.5 * complex(amp[0] * Cab + amp[1] * B, -(amp[2] * Cab + amp[3] * B))
    F = abs(Xa) ** 2 / A + abs(Xb) ** 2 / B
    est = np.array([2 * Xa.real / A, 2 * Xb.real / B, -2 * Xa.imag / A, -2 * Xb.imag / B])
    return F, est, Xa, Xb


def test_noise_free_closed_form(clean):
    seg, truth, amp, rho2, S = clean
    out, _ = S.fstatnet(truth[None, :], grad=False)
    out = out[0]
    F = -out[5]
    Fc, est, Xa, Xb = closed_form(seg, truth, amp)
    assert F == pytest.approx(Fc, rel=1e-3)
    assert np.allclose(out[:4], est, rtol=0, atol=1e-3 * np.max(np.abs(est)))
    assert np.array_equal(out[6:10], truth)
    assert out[4] == pytest.approx(math.sqrt(2 * (F - 2)))
    # out[10] = sqrt((2 Re Xa)^2 + (2 Re Xb)^2 + (2 Im Xa)^2 + (2 Im Xb)^2)
    assert out[10] == pytest.approx(2 * math.sqrt(abs(Xa) ** 2 + abs(Xb) ** 2), rel=1e-3)


def test_gradient_central_differences(clean):
    seg, truth, amp, rho2, S = clean
    sc = phase_scales(seg, truth)
    rng = np.random.default_rng(3)
    pts = truth[None, :] + (rng.uniform(-0.3, 0.3, (4, 4)) * sc)
    out, g = S.fstatnet(pts)
    h = 1e-4
    for k in range(4):
        e = np.zeros(4)
        e[k] = h * sc[k]
        fp, _ = S.fstatnet(pts + e, grad=False)
        fm, _ = S.fstatnet(pts - e, grad=False)
        fd = (-fp[:, 5] + fm[:, 5]) / (2 * e[k])
        # compared per unit scaled step, relative to F
        err = np.abs(g[:, k] - fd) * sc[k] / (-out[:, 5])
        assert np.max(err) < 1e-6, (k, err)


def test_maximum_at_truth(orc):
    """With the signal in the a-channel only (amp1 = amp3 = 0) the statistic peaks at the true
    parameters: the scaled gradient there is small next to its size a quarter step away.  (With
    both channels and a nonzero a-b overlap the peak of this statistic moves off the truth.)"""
    seg, truth, amp, rho2 = noise_free_signal(seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    S = orc.Search(seg, nthreads=8)
    sc = phase_scales(seg, truth)
    _, g0 = S.fstatnet(truth[None, :])
    for k in range(4):
        p = truth.copy()
        p[k] += 0.25 * sc[k]
        _, g1 = S.fstatnet(p[None, :])
        assert abs(g0[0, k] * sc[k]) < 0.1 * abs(g1[0, k] * sc[k]), k


def test_agrees_with_grid_search(orc, abi, synth):
    """Fstatnet at the grid search's best candidate reproduces that candidate's F (no FStat
    normalisation) to the grid search's own approximation (barycentring by resampling and spline,
    about 1 %); a sign or index slip in the phase model would leave F at the noise level."""
    from cases import injected_sky, segment
    seg = segment("net-2det-32k")
    S = orc.Search(seg, nthreads=8)
    pm, mm, nn = injected_sky(seg)
    recs, _ = S.job_core(pm, mm, nn, abi.make_range(seg), abi.make_opts(trl=10.0, white=True))
    best = recs[np.argmax(recs[:, 4])]
    Fgrid = best[4] ** 2 / 2 + 2
    out, _ = S.fstatnet(best[None, :4], grad=False)
    assert -out[0, 5] == pytest.approx(Fgrid, rel=0.03)


def test_amoeba_reaches_a_maximum(orc):
    """The simplex started half a scaled step from the injection ends at a stationary point of F
    (scaled gradient below 1e-3 F) with F at least F(truth)."""
    seg, truth, amp, rho2 = noise_free_signal(seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    S = orc.Search(seg, nthreads=8)
    sc = phase_scales(seg, truth)
    start = truth + np.array([0.6, -0.5, 0.4, -0.4]) * sc
    x, out, nev = S.amoeba(start, 0.5 * sc, tol=1e-12)
    assert nev > 0
    F0, _ = S.fstatnet(truth[None, :], grad=False)
    assert -out[5] >= -F0[0, 5] * (1 - 1e-9)
    _, g = S.fstatnet(x[None, :])
    assert np.max(np.abs(g[0] * sc)) < 1e-3 * -out[5]


class _OracleFS:
    """The followup's batched driver fed by the oracle's Fstatnet (CPU): the same arithmetic as the
    oracle's own Nelder-Mead, so the two must walk identical simplices."""

    def __init__(self, S):
        self.S = S

    def __call__(self, pts, grad=False, owner=None):
        out, _ = self.S.fstatnet(np.asarray(pts, dtype=np.float64).reshape(-1, 4), grad=False)
        return out, None


def test_batched_amoeba_equals_reference_steps(orc):
    """followup.amoeba (batched over simplices) against the oracle's restatement of followup.c's
    amoeba, one start at a time: identical best points, outputs and evaluation counts."""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(seed=8, amp=[0.7, 0.0, -0.4, 0.0])
    S = orc.Search(seg, nthreads=8)
    sc = phase_scales(seg, truth)
    starts = truth + np.random.default_rng(3).uniform(-0.8, 0.8, (5, 4)) * sc
    x, out, nev, conv = followup.amoeba(_OracleFS(S), starts, 0.5 * sc, tol=1e-9)
    assert conv.all()
    for k in range(len(starts)):
        xr, outr, nevr = S.amoeba(starts[k], 0.5 * sc, tol=1e-9)
        assert nev[k] == nevr
        assert np.array_equal(x[k], xr)
        assert np.array_equal(out[k], outr)
    # an evaluation budget ends a simplex unconverged
    _, _, nev2, conv2 = followup.amoeba(_OracleFS(S), starts[:2], 0.5 * sc, tol=1e-14, maxeval=12)
    assert not conv2.any() and np.all(nev2 >= 12)


def test_mads_mesh_order_is_the_references(orc):
    """The oracle's 80 mesh offsets against entries read off the reference's array assignments
    (followup.c:573-672; index j of array[j], 1-based), each of the 80 nonzero sign vectors once,
    and the product's stencil in the same order."""
    from polgraw_allsky_amd import followup
    st = orc.mads_stencil()
    hand = {1: (1, 1, 1, 1), 2: (1, 1, 1, -1), 9: (-1, 1, 1, 1), 16: (-1, -1, -1, -1),
            17: (0, 1, 1, 1), 24: (0, -1, -1, -1), 25: (1, 0, 1, 1), 33: (1, 1, 0, 1), 41: (1, 1, 1, 0),
            48: (-1, -1, -1, 0), 49: (0, 0, 1, 1), 53: (0, 1, 0, 1), 55: (0, 1, 1, 0), 60: (0, -1, -1, 0),
            61: (1, 0, 0, 1), 63: (1, 0, 1, 0), 65: (1, 1, 0, 0), 66: (1, -1, 0, 0), 70: (-1, 0, -1, 0),
            72: (-1, -1, 0, 0), 73: (0, 1, 0, 0), 75: (0, 0, 0, 1), 77: (0, 0, 1, 0), 79: (1, 0, 0, 0),
            80: (-1, 0, 0, 0)}
    for j, v in hand.items():
        assert tuple(st[j - 1]) == v, (j, st[j - 1], v)
    assert len({tuple(r) for r in st}) == 80 and not np.any(np.all(st == 0, axis=1))
    assert np.array_equal(st.astype(np.float64), followup.MADS_STENCIL)


def test_batched_mads_equals_oracle_mads(orc):
    """followup.mads (batched over candidates and mesh points, fed by the oracle's Fstatnet) against
    the oracle's invertedMADS (orc_mads): identical final points, outputs and rounds; the end point
    improves on the start."""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(N=8192, seed=8, amp=[2.8, 0.0, -1.6, 0.0])
    S = orc.Search(seg, nthreads=2)
    sc = phase_scales(seg, truth)
    nod = seg.sett["N"] * seg.sett["dt"] / 86400.0
    starts = truth + np.random.default_rng(6).uniform(-0.6, 0.6, (3, 4)) * sc
    p2 = 0.5 * sc  # a coarse mesh limit relative to the short segment keeps the search short
    x, out, rounds = followup.mads(_OracleFS(S), starts, p2, nod)
    assert np.all(rounds < 700)
    for k in range(len(starts)):
        xr, outr, cr = S.mads(starts[k], p2, nod)
        print("start", k, "rounds", rounds[k], cr, "F", -out[k, 5], -outr[5], flush=True)
        assert rounds[k] == cr
        assert np.array_equal(x[k], xr) and np.array_equal(out[k], outr, equal_nan=True)
        F0 = -S.fstatnet(starts[k][None, :], grad=False)[0][0, 5]
        assert -out[k, 5] > F0


def test_oracle_mads_round_limit(orc):
    """The round limit ends the search (followup.c:740-742): with mesh steps of 6e-7 of the mesh
    limit (nod = 1e5 days) the oracle stops after exactly `limit` rounds, as the batched driver does."""
    from polgraw_allsky_amd import followup
    seg, truth, amp, rho2 = noise_free_signal(N=8192, seed=8, amp=[2.8, 0.0, -1.6, 0.0])
    S = orc.Search(seg, nthreads=2)
    sc = phase_scales(seg, truth)
    x, out, r = S.mads(truth + 0.1 * sc, 0.5 * sc, 1e5, limit=5)
    xb, outb, rb = followup.mads(_OracleFS(S), (truth + 0.1 * sc)[None, :], 0.5 * sc, 1e5, limit=5)
    assert r == 5 and rb[0] == 5
    assert np.array_equal(x, xb[0]) and np.array_equal(out, outb[0])
