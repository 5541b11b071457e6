"""Followup of candidates with the exact network F-statistic (followup/src/followup.c).

`FStatNet` is the reference's Fstatnet (followup.c:413-534) for batches of points on the device,
with its analytic gradient; `fstat` makes it differentiable in torch; `ascend` is a batched
quasi-Newton ascent started from the phase metric, an alternative to the simplex search of
followup.c:1126-1173 (`amoeba`) when the gradient is available.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import pgw


def _phase_derivatives(DetSSB, N: int, oms: float, point, stride: int = 1):
    """d phi / d(omega, omegadot, delta, alpha) at every stride-th sample of every detector, for
    phi = omega (i + shft) + omegadot i^2 + (oms + 2 omegadot i) shft (followup.c:455-457)."""
    om, omd, de, al = (float(v) for v in point)
    i = np.arange(0, N, stride, dtype=np.float64)
    nS = np.array([np.cos(al) * np.cos(de), np.sin(al) * np.cos(de), np.sin(de)])
    nSd = np.array([-np.cos(al) * np.sin(de), -np.sin(al) * np.sin(de), np.cos(de)])
    nSa = np.array([-np.sin(al) * np.cos(de), np.cos(al) * np.cos(de), 0.0])
    d = [[], [], [], []]
    for R in DetSSB:
        R = np.asarray(R, dtype=np.float64)[:3 * N].reshape(N, 3)[::stride]
        shft = R @ nS
        K = om + oms + 2 * omd * i
        d[0].append(i + shft)
        d[1].append(i * i + 2 * i * shft)
        d[2].append(K * (R @ nSd))
        d[3].append(K * (R @ nSa))
    return np.stack([np.concatenate(v) for v in d])


def phase_scales(DetSSB, N: int, oms: float, point) -> np.ndarray:
    """Parameter steps that change the phase by about one radian rms over the segment: for
    parameter k, 1 / std_i(d phi / d theta_k) (the mean is removed: a constant phase does not
    change F).  DetSSB: the detectors' 3N barycentric positions."""
    return 1.0 / np.std(_phase_derivatives(DetSSB, N, oms, point), axis=1)


def phase_metric(DetSSB, N: int, oms: float, point, stride: int = 8) -> np.ndarray:
    """Phase metric G_jk = Cov_i(d phi/d theta_j, d phi/d theta_k) over the samples of all detectors.
    Near a maximum F(theta + d) ~ F(theta) (1 - d^T G d), so G^-1 grad F / (2 F) is a Newton step
    and sqrt(d^T G d) is the rms phase change of a step in radians.  The derivatives are smooth in
    time, so every stride-th sample gives the covariance to high accuracy."""
    D = _phase_derivatives(DetSSB, N, oms, point, stride)
    D = D - D.mean(axis=1, keepdims=True)
    return D @ D.T / D.shape[1]


class FStatNet:
    """Fstatnet at batches of points of one segment of a pgw.Handle (FP64 on the device)."""

    def __init__(self, handle: pgw.Handle, seg: int = 0):
        self.h = handle
        self.seg = seg

    def __call__(self, pts, grad: bool = True, owner=None):
        """pts (n, 4) = (omega, omegadot, delta, alpha) -> (out (n, 11), dF/dpts (n, 4) or None).
        out[:, 5] = -F as the reference; the gradient is of F.  owner (the candidate each point
        belongs to) is not needed with one segment."""
        return self.h.fstatnet(pts, seg=self.seg, grad=grad)


class MultiFStatNet:
    """Fstatnet for candidates spread over the segments of one pgw.Handle: candidate k lives in
    segment seg_of[k].  The searches below pass each point's candidate (`owner`); all points go to
    the device in one call (pgw_fstatnet_multi)."""

    def __init__(self, handle: pgw.Handle, seg_of):
        self.h = handle
        self.seg_of = np.asarray(seg_of, dtype=np.int64)

    def __call__(self, pts, grad: bool = True, owner=None):
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 4)
        if owner is None:
            if len(pts) != len(self.seg_of):
                raise ValueError("points of several segments need their owners")
            owner = np.arange(len(pts))
        seg = self.seg_of[np.asarray(owner, dtype=np.int64)]
        return self.h.fstatnet_multi(seg, pts, grad=grad)


def _torch_fn():
    import torch

    class FstatFn(torch.autograd.Function):
        """F(pts) with the device's analytic gradient: pts (n, 4) float64 tensor -> F (n,)."""

        @staticmethod
        def forward(ctx, pts, fs):
            out, g = fs(pts.detach().cpu().numpy(), grad=True)
            ctx.save_for_backward(torch.from_numpy(g).to(pts.device))
            return torch.from_numpy(-out[:, 5]).to(pts.device)

        @staticmethod
        def backward(ctx, gF):
            (g,) = ctx.saved_tensors
            return gF[:, None] * g, None

    return FstatFn


def fstat(pts, fs: FStatNet):
    """Differentiable F at pts (torch float64 (n, 4))."""
    return _torch_fn().apply(pts, fs)


def ascend(fs: FStatNet, start, metric, steps: int = 20, trust0: float = 1.0, trust_max: float = 4.0):
    """Batched quasi-Newton ascent of F from start (n, 4), one device evaluation per step.

    metric: (4, 4) or (n, 4, 4) phase metrics (phase_metric at each start).  The inverse Hessian
    starts from (2 F G)^-1, the Newton step of the quadratic model F (1 - d^T G d), and BFGS updates
    it from the exact gradients, which brings in the curvature of the amplitude modulation that the
    phase metric leaves out.  Each step is cut to a trust radius in rms radians of phase
    (sqrt(d^T G d)); it is kept if F increases (radius doubles, up to trust_max), else rejected
    (radius / 4).  All candidates are evaluated in one device batch per step.
    Returns (points, F, evaluations per candidate)."""
    x = np.array(start, dtype=np.float64).reshape(-1, 4)
    n = len(x)
    G = np.broadcast_to(np.asarray(metric, dtype=np.float64), (n, 4, 4)).copy()
    # regularise the near-degenerate directions of short segments
    G += 1e-9 * np.trace(G, axis1=1, axis2=2)[:, None, None] * np.eye(4)
    own = np.arange(n)
    out, g = fs(x, owner=own)
    F = -out[:, 5]
    H = np.linalg.inv(2 * F[:, None, None] * G)
    trust = np.full(n, float(trust0))
    eye = np.eye(4)
    nev = 1
    for _ in range(steps):
        d = np.einsum("nij,nj->ni", H, g)
        rms = np.sqrt(np.maximum(np.einsum("ni,nij,nj->n", d, G, d), 1e-300))
        d *= np.minimum(1.0, trust / rms)[:, None]
        out_t, g_t = fs(x + d, owner=own)
        nev += 1
        Ft = -out_t[:, 5]
        ok = Ft > F
        for k in np.nonzero(ok)[0]:
            y = g[k] - g_t[k]                  # change of the gradient of -F
            ys = y @ d[k]
            if ys > 0:                         # curvature condition; skip the update otherwise
                r = 1.0 / ys
                A = eye - r * np.outer(d[k], y)
                H[k] = A @ H[k] @ A.T + r * np.outer(d[k], d[k])
        x[ok] += d[ok]
        F[ok], g[ok] = Ft[ok], g_t[ok]
        trust = np.where(ok, np.minimum(2 * trust, trust_max), trust / 4)
    return x, F, nev


def amoeba(fs: FStatNet, starts, steps, tol: float = 1e-7, maxeval: int = 20000):
    """The followup's Nelder-Mead on -F (followup/src/followup.c:1126-1173) for many starting
    points at once: simplex k starts at starts[k] with vertex i (< 4) moved by steps[k][i] along
    parameter i; every round's trial vertices of all unfinished simplices go to the device in one
    batched Fstatnet call.  Per simplex the steps are the reference's: reflection; expansion when
    the reflected vertex is the new best; contraction when it is still the worst, and a shrink
    toward the best vertex when that fails; stop when |fmax - fmin| < (|fmax| + |fmin|) tol + 1e-14
    (or after maxeval evaluations, which leaves that simplex's best vertex unconverged).
    Returns (best points (n, 4), their 11 outputs (n, 11), evaluations (n,), converged (n,))."""
    D = 4
    x0 = np.asarray(starts, dtype=np.float64).reshape(-1, D)
    st = np.broadcast_to(np.asarray(steps, dtype=np.float64), x0.shape)
    n = len(x0)
    sx = np.repeat(x0[:, None, :], D + 1, axis=1)          # (n, 5, 4)
    for i in range(D):
        sx[:, i, i] += st[:, i]
    out = fs(sx.reshape(-1, D), grad=False, owner=np.repeat(np.arange(n), D + 1))[0].reshape(n, D + 1, 11)
    nev = np.full(n, D + 1)
    done = np.zeros(n, bool)
    conv = np.zeros(n, bool)
    ar = np.arange(n)

    def evaluate(owner, pts):
        return fs(pts, grad=False, owner=owner)[0] if len(owner) else np.zeros((0, 11))

    while True:
        act = np.nonzero(~done)[0]
        if not len(act):
            break
        f = out[act, :, 5]
        # ordering with the reference's tie rules
        first = f[:, 0] > f[:, 1]
        ihi = np.where(first, 0, 1)
        ilo = np.where(first, 1, 0)
        inhi = ilo.copy()
        for i in range(2, D + 1):
            fi = f[:, i]
            lo_ = fi <= f[np.arange(len(act)), ilo]
            hi_ = ~lo_ & (fi > f[np.arange(len(act)), ihi])
            nh_ = ~lo_ & ~hi_ & (fi > f[np.arange(len(act)), inhi])
            inhi = np.where(hi_, ihi, np.where(nh_, i, inhi))
            ihi = np.where(hi_, i, ihi)
            ilo = np.where(lo_, i, ilo)
        a = np.arange(len(act))
        mid = np.zeros((len(act), D))
        for i in range(D + 1):
            mid += np.where((ihi != i)[:, None], sx[act, i, :], 0.0)
        mid /= D
        line = sx[act, ihi, :] - mid
        fmax, fmin = f[a, ihi], f[a, ilo]
        fin = np.abs(fmax - fmin) < (np.abs(fmax) + np.abs(fmin)) * tol + 1e-14
        conv[act[fin]] = True
        stop = fin | (nev[act] >= maxeval)
        done[act[stop]] = True
        keep = ~stop
        act, ihi, ilo, inhi, mid, line = act[keep], ihi[keep], ilo[keep], inhi[keep], mid[keep], line[keep]
        if not len(act):
            continue

        def trial(sel, scale):
            """Try mid + scale line for simplices act[sel]; accept where it beats their worst."""
            idx = act[sel]
            pts = mid[sel] + scale * line[sel]
            o = evaluate(idx, pts)
            nev[idx] += 1
            acc = o[:, 5] < out[idx, ihi[sel], 5]
            out[idx[acc], ihi[sel][acc]] = o[acc]
            sx[idx[acc], ihi[sel][acc]] = pts[acc]
            return acc

        allsel = np.arange(len(act))
        trial(allsel, -1.0)                                     # reflection
        fw = out[act, ihi, 5]
        best = fw < out[act, ilo, 5]
        if best.any():
            trial(np.nonzero(best)[0], -2.0)                     # expansion
        worst = ~best & (fw > out[act, inhi, 5])
        if worst.any():
            sel = np.nonzero(worst)[0]
            acc = trial(sel, 0.5)                                # contraction
            shr = sel[~acc]
            if len(shr):                                         # shrink toward the best vertex
                idx = act[shr]
                pts, where = [], []
                for k, s in zip(idx, shr):
                    for i in range(D + 1):
                        if i == ilo[s]:
                            continue
                        sx[k, i] = (sx[k, ilo[s]] + sx[k, i]) * 0.5
                        pts.append(sx[k, i])
                        where.append((k, i))
                o = evaluate(np.array([k for k, _ in where]), np.array(pts))
                for (k, i), oo in zip(where, o):
                    out[k, i] = oo
                nev[idx] += D
    f = out[:, :, 5]
    # the reference returns the best vertex of its last ordering (ties resolved as there)
    best = np.array([_last_ilo(f[k]) for k in range(n)])
    return sx[ar, best], out[ar, best], nev, conv


def _last_ilo(f):
    """The best vertex as the reference's ordering loop names it (ties resolved as there)."""
    if f[0] > f[1]:
        ilo = 1
    else:
        ilo = 0
    for i in range(2, len(f)):
        if f[i] <= f[ilo]:
            ilo = i
    return ilo


# The 80 points of the 3^4 mesh around the centre as multiples (+1, 0, -1) of the mesh sizes of
# (omega, omegadot, delta, alpha), in the order the reference fills its mesh array
# (followup/src/followup.c:573-672): 16 vertices (all four moved; omega slowest, + before -), 32
# edges (one kept, in parameter order; the rest in the vertices' order), 24 faces (two kept) and 8
# cells (one moved: omegadot, alpha, delta, omega).  The order decides only exact ties: the first
# strict minimum wins (the reference's OpenMP loop leaves ties to its schedule).
def _mads_stencil() -> np.ndarray:
    vert = list(itertools.product((1.0, -1.0), repeat=4))
    edges = []
    for z in range(4):
        for v in itertools.product((1.0, -1.0), repeat=3):
            p = list(v)
            p.insert(z, 0.0)
            edges.append(tuple(p))
    faces = [(0, 0, d, a) for d in (1, -1) for a in (1, -1)]
    for f in (0, 1, -1):
        for s_ in ((1, -1) if f == 0 else (0,)):
            faces += [(f, s_, 0, 1), (f, s_, 0, -1), (f, s_, 1, 0), (f, s_, -1, 0)]
        if f != 0:
            faces += [(f, 1, 0, 0), (f, -1, 0, 0)]
    cells = [(0, 1, 0, 0), (0, -1, 0, 0), (0, 0, 0, 1), (0, 0, 0, -1), (0, 0, 1, 0), (0, 0, -1, 0),
             (1, 0, 0, 0), (-1, 0, 0, 0)]
    return np.array(vert + edges + faces + cells, dtype=np.float64)


MADS_STENCIL = _mads_stencil()


def mads(fs: FStatNet, starts, p2, nod: float, limit: int = 700):
    """The followup's mesh search (invertedMADS, followup/src/followup.c:537-754) for many
    candidates at once: each round evaluates, for every unfinished candidate, the 80 points of the
    3^4 mesh around its current best point (mesh sizes `param`) in one batched Fstatnet call.  A
    better point becomes the centre and the mesh restarts at (1 - scale) p2; otherwise the mesh
    grows by (1 - scale) p2; scale = 1 - 0.06 / nod.  A candidate finishes when every mesh size
    exceeds p2, or after `limit` rounds.  Returns (points (n, 4), their 11 outputs (n, 11), rounds
    (n,))."""
    x = np.asarray(starts, dtype=np.float64).reshape(-1, 4).copy()
    n = len(x)
    p2 = np.broadcast_to(np.asarray(p2, dtype=np.float64), x.shape)
    step = (1.0 - (1.0 - 0.01 * 6.0 / nod)) * p2
    param = step.copy()
    best = fs(x, grad=False, owner=np.arange(n))[0]
    rounds = np.zeros(n, int)
    act = np.ones(n, bool)
    while act.any():
        idx = np.nonzero(act)[0]
        rounds[idx] += 1
        pts = x[idx, None, :] + MADS_STENCIL[None, :, :] * param[idx, None, :]
        out = fs(pts.reshape(-1, 4), grad=False, owner=np.repeat(idx, 80))[0].reshape(len(idx), 80, 11)
        j = np.argmin(out[:, :, 5], axis=1)              # first minimum in stencil order
        cand = out[np.arange(len(idx)), j]
        better = cand[:, 5] < best[idx, 5]
        bi = idx[better]
        x[bi] = pts[np.arange(len(idx))[better], j[better]]
        best[bi] = cand[better]
        param[bi] = step[bi]
        gi = idx[~better]
        param[gi] = param[gi] + step[gi]
        act[idx] = np.any(param[idx] <= p2[idx], axis=1) & (rounds[idx] < limit)
    return x, best, rounds
