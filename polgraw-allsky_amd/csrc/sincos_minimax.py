"""Coefficients of the FP64 sin/cos polynomials of sincos_cw (pgw_kernels.hpp).

On |r| <= pi/4, with z = r^2:  sin r = r + r^3 S(z),  cos r = 1 - z C(z), where S and C are the
polynomials (degree 5 each) that minimise the largest absolute error of sin and of cos over the
interval (Remez exchange, weights r^3 and r^2), computed in 200-bit arithmetic and printed highest
power first as the doubles the kernel uses.  Run: python3 sincos_minimax.py
"""
import mpmath as mp

mp.mp.prec = 200
B = (mp.pi / 4) ** 2
A = B * mp.mpf(10) ** -8  # z = 0 itself has weight 0; start just above it


def s_of_z(z):
    r = mp.sqrt(z)
    return (mp.sin(r) - r) / (r * z)


def c_of_z(z):
    return (1 - mp.cos(mp.sqrt(z))) / z


def remez(f, w, deg, iters=40, grid=4000):
    """Minimax polynomial p (coefficients, lowest first) of max |w (p - f)| on [A, B] and that max."""
    n = deg + 2
    xs = [(A + B) / 2 - (B - A) / 2 * mp.cos(mp.pi * i / (n - 1)) for i in range(n)]
    pts = [A + (B - A) * i / grid for i in range(grid + 1)]
    for _ in range(iters):
        M = mp.matrix(n, n)
        y = mp.matrix(n, 1)
        for i, x in enumerate(xs):
            for j in range(deg + 1):
                M[i, j] = x ** j
            M[i, deg + 1] = (-1) ** i / w(x)
            y[i] = f(x)
        sol = mp.lu_solve(M, y)
        c = [sol[j] for j in range(deg + 1)]
        ev = [w(z) * (sum(c[j] * z ** j for j in range(deg + 1)) - f(z)) for z in pts]
        ext = [0] + [i for i in range(1, grid) if (ev[i] - ev[i - 1]) * (ev[i + 1] - ev[i]) <= 0] + [grid]
        sel = []
        for i in ext:  # alternating extrema, the largest of each run of one sign
            if sel and mp.sign(ev[i]) == mp.sign(ev[sel[-1]]):
                if abs(ev[i]) > abs(ev[sel[-1]]):
                    sel[-1] = i
            else:
                sel.append(i)
        while len(sel) > n:
            sel.pop(0 if abs(ev[sel[0]]) < abs(ev[sel[-1]]) else -1)
        if len(sel) < n:
            break
        xs = [pts[i] for i in sel]
    return c, max(abs(e) for e in ev)


if __name__ == "__main__":
    cs, es = remez(s_of_z, lambda z: mp.sqrt(z) ** 3, 5)
    cc, ec = remez(c_of_z, lambda z: z, 5)
    print("sin: max abs error", mp.nstr(es, 3), "S (highest first):", [repr(float(x)) for x in cs[::-1]])
    print("cos: max abs error", mp.nstr(ec, 3), "C (highest first):", [repr(float(x)) for x in cc[::-1]])
