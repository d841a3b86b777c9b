"""Coefficients of the kernels' FP64 exp and log polynomials (ice3_physics.hpp: exp_f64,
log_f64).  Not run by the build: the coefficients are literals in the header; this script
shows how they were obtained and measures the result.

Each polynomial interpolates its target at Chebyshev nodes (a near-minimax fit), in 60-digit
decimal arithmetic, and the coefficients are rounded to double:

* exp: e^r = 1 + r (1 + r q(r)) on |r| <= 0.3466 (just over ln2 / 2), q(r) = 1/2 + r q2(r)
  with q2 of degree 8 fitted to ((e^r - 1 - r)/r^2 - 1/2)/r, so that q(0) = 1/2 is the exact
  inline constant;
* log: log m = 2s + s z p(z), s = (m - 1)/(m + 1), z = s^2 in [0, ((sqrt2 - 1)/(sqrt2 + 1))^2],
  p of degree 6 fitted to (log m - 2s)/(s z).

The error of the double Horner chain, each step a correctly rounded FMA, is then measured
exactly against 60-digit references at random arguments.

Usage: python3 fit_exp_log.py
"""
import random
from decimal import Decimal as D, getcontext

getcontext().prec = 60
PI = D("3.14159265358979323846264338327950288419716939937510582097494459")


def cos(t):
    c, term, i = D(0), D(1), 0
    while abs(term) > D("1e-58"):
        c += term
        i += 2
        term = -term * t * t / (i * (i - 1))
    return c


def solve(a, b):
    n = len(b)
    m = [row[:] + [b[i]] for i, row in enumerate(a)]
    for c in range(n):
        p = max(range(c, n), key=lambda r: abs(m[r][c]))
        m[c], m[p] = m[p], m[c]
        for r in range(n):
            if r != c:
                f = m[r][c] / m[c][c]
                m[r] = [m[r][k] - f * m[c][k] for k in range(n + 1)]
    return [m[i][n] / m[i][i] for i in range(n)]


def fit(f, lo, hi, deg):
    """Coefficients (lowest degree first) of the polynomial interpolating f at deg+1
    Chebyshev nodes of [lo, hi]."""
    lo, hi = D(lo), D(hi)
    nodes = [(lo + hi) / 2 + (hi - lo) / 2 * cos((2 * k + 1) * PI / (2 * (deg + 1)))
             for k in range(deg + 1)]
    return solve([[x ** j if j else D(1) for j in range(deg + 1)] for x in nodes], [f(x) for x in nodes])


def fma(a, b, c):
    return float(D(a) * D(b) + D(c))


def q2(x):
    """((e^x - 1 - x)/x^2 - 1/2)/x, by its series near 0 (cancellation)."""
    if abs(x) < D("1e-6"):
        return sum(x ** k / D(_fact(k + 3)) for k in range(12))
    return ((x.exp() - 1 - x) / (x * x) - D("0.5")) / x


def _fact(n):
    r = 1
    for i in range(2, n + 1):
        r *= i
    return r


def p_log(z):
    if z == 0:
        return D(2) / 3
    s = z.sqrt()
    return (((1 + s) / (1 - s)).ln() - 2 * s) / (s * z)


def main():
    rng = random.Random(1)
    L = 0.3466
    c = [0.5] + [float(v) for v in fit(q2, -L, L, 8)]
    worst = 0.0
    for _ in range(20000):
        r = rng.uniform(-L, L)
        p = c[-1]
        for k in range(len(c) - 2, -1, -1):
            p = fma(p, r, c[k])
        p = fma(fma(p, r, 1.0), r, 1.0)
        ex = D(r).exp()
        worst = max(worst, float(abs((D(p) - ex) / ex) / D(2.0 ** -52)))
    print("exp: q coefficients (r^0 first)", [v.hex() for v in c], "max error %.3f ulp" % worst)

    zmax = ((2 ** 0.5 - 1) / (2 ** 0.5 + 1)) ** 2
    c = [float(v) for v in fit(p_log, 0, zmax * 1.0001, 6)]
    worst = 0.0
    for _ in range(20000):
        m = rng.uniform(2 ** -0.5, 2 ** 0.5)
        f = m - 1.0
        s = float(D(f) / D(2.0 + f))
        z = s * s
        p = c[-1]
        for k in range(len(c) - 2, -1, -1):
            p = fma(p, z, c[k])
        r = fma(s * z, p, s + s)
        ex = D(m).ln()
        if ex != 0:
            worst = max(worst, float(abs((D(r) - ex) / ex) / D(2.0 ** -52)))
    print("log: p coefficients (z^0 first)", [v.hex() for v in c], "max error of log m %.3f ulp" % worst)


if __name__ == "__main__":
    main()
