"""Coefficients of the near-field sin/cos polynomials in p2p_math.h (provenance script).

The near field evaluates e^{i pi u} for u = k r / pi, reduced to y = u - rint(u) in
[-1/2, 1/2].  On that interval

    sin(pi y) ~= y * S(y^2),   S of degree 4 in y^2 (odd polynomial of degree 9 in y),
    cos(pi y) ~= C(y^2),       C of degree 5 in y^2 (even polynomial of degree 10 in y),

chosen by the Remez exchange algorithm (E. Ya. Remez, 1934; the second algorithm as in
Cheney, "Introduction to Approximation Theory", ch. 3) for the smallest maximum ABSOLUTE
error of sin(pi y) and cos(pi y) over y in [0, 1/2] (the polynomials are odd / even, so
the bound holds on [-1/2, 1/2]).  Arithmetic in mpmath at 50 digits; the coefficients are
then rounded to binary32.

Run:  python hfmm0_amd/csrc/sincos_remez.py   (prints the C initialisers and the errors)
"""
import mpmath as mp

mp.mp.dps = 50


def remez(f, w, nterms, a, b, iters=30):
    """Minimise max_{t in [a,b]} |w(t) (P(t) - f(t))| over polynomials P of nterms terms."""
    n = nterms
    # initial reference: Chebyshev extrema
    xs = [(a + b) / 2 + (b - a) / 2 * mp.cos(mp.pi * i / (n)) for i in range(n + 1)][::-1]
    coef = None
    for _ in range(iters):
        # solve sum_j c_j x_i^j + (-1)^i E / w(x_i) = f(x_i)
        A = mp.matrix(n + 1, n + 1)
        rhs = mp.matrix(n + 1, 1)
        for i, x in enumerate(xs):
            for j in range(n):
                A[i, j] = x ** j
            A[i, n] = (-1) ** i / w(x)
            rhs[i] = f(x)
        sol = mp.lu_solve(A, rhs)
        coef = [sol[j] for j in range(n)]

        def err(t):
            return w(t) * (mp.polyval(coef[::-1], t) - f(t))
        # new reference: the extrema of err between sign changes, located on a fine grid
        # and refined by golden-section search
        grid = [a + (b - a) * i / 4000 for i in range(4001)]
        vals = [err(t) for t in grid]
        ext = []
        i = 0
        while i < len(grid):
            j = i
            s = mp.sign(vals[i])
            while j + 1 < len(grid) and mp.sign(vals[j + 1]) == s:
                j += 1
            k = max(range(i, j + 1), key=lambda m: abs(vals[m]))
            lo, hi = grid[max(k - 1, 0)], grid[min(k + 1, len(grid) - 1)]
            x = grid[k]
            if 0 < k < len(grid) - 1:
                g = (mp.sqrt(5) - 1) / 2
                for _ in range(80):
                    c1, c2 = hi - g * (hi - lo), lo + g * (hi - lo)
                    if abs(err(c1)) > abs(err(c2)):
                        hi = c2
                    else:
                        lo = c1
                x = (lo + hi) / 2
            ext.append(x)
            i = j + 1
        if len(ext) < n + 1:
            break
        # keep n+1 alternating extrema with the largest errors
        while len(ext) > n + 1:
            if abs(err(ext[0])) < abs(err(ext[-1])):
                ext.pop(0)
            else:
                ext.pop()
        xs = ext
    return coef


def main():
    a, b = mp.mpf(0), mp.mpf(1) / 4       # t = y^2, y in [0, 1/2]
    # sin: P(t) ~ sin(pi sqrt t)/sqrt t, weight sqrt t (absolute error of y P(y^2))
    fs = lambda t: mp.sin(mp.pi * mp.sqrt(t)) / mp.sqrt(t) if t > 0 else mp.pi
    ws = lambda t: mp.sqrt(t) if t > 0 else mp.mpf(10) ** -30
    fc = lambda t: mp.cos(mp.pi * mp.sqrt(t))
    wc = lambda t: mp.mpf(1)
    cs = remez(fs, ws, 5, a, b)
    cc = remez(fc, wc, 6, a, b)
    import numpy as np
    f32 = lambda v: float(np.float32(float(v)))
    print("sin coefficients (y, y^3, ..., y^9):", ", ".join("%.9ef" % f32(c) for c in cs))
    print("cos coefficients (1, y^2, ..., y^10):", ", ".join("%.9ef" % f32(c) for c in cc))
    ys = [mp.mpf(i) / 20000 for i in range(10001)]
    es = max(abs(y * mp.polyval([mp.mpf(f32(c)) for c in cs][::-1], y * y) - mp.sin(mp.pi * y)) for y in ys)
    ec = max(abs(mp.polyval([mp.mpf(f32(c)) for c in cc][::-1], y * y) - mp.cos(mp.pi * y)) for y in ys)
    print("max |error| with binary32 coefficients, exact arithmetic: sin %.3e, cos %.3e" % (es, ec))


if __name__ == "__main__":
    main()
