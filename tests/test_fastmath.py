"""The kernels' FP64 exp and log (ice3_physics.hpp: exp_f64, log_f64) against the C library.

They are built from textbook reductions (log: frexp and 2 atanh((m-1)/(m+1)); exp: n ln2)
with near-minimax polynomials (csrc/fit_exp_log.py), so they are pinned here by what the
mathematics fixes: the C library's values (numpy: glibc, within 1 ulp), the exact
special cases, and identities exp(log x) = x and log(2^k) = k ln 2.  The CPU build runs the
same header as the kernels; every step is a correctly rounded IEEE operation, except that the
kernels' division (log's (m - 1)/(m + 1)) is a reciprocal refined by Newton steps, within 1 ulp.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_stages"))
import host_lib  # noqa: E402


def ulps(y, ref):
    return np.abs(y - ref) / np.spacing(np.abs(ref))


def test_log_accuracy():
    rng = np.random.default_rng(7)
    x = np.concatenate([
        10.0 ** rng.uniform(-307, 308, 200000),           # every binade
        1.0 + rng.uniform(-0.3, 0.45, 100000),             # near 1, both sides of sqrt(1/2)..sqrt(2)
        1.0 + 10.0 ** rng.uniform(-15, -3, 20000) * rng.choice([-1, 1], 20000),
        rng.uniform(1e-323, 2.2e-308, 2000),                # subnormal
        np.array([np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 0), np.sqrt(2.0), 0.5, 1.0, 2.0,
                  np.finfo(float).max, np.finfo(float).tiny, 5e-324]),
    ])
    y = host_lib.elementwise("log", x)
    ref = np.log(x)
    err = ulps(y, ref)
    nz = ref != 0
    assert err[nz].max() <= 2.0, (err.max(), x[np.argmax(err)])
    assert y[x == 1.0].tolist() == [0.0] * int((x == 1.0).sum())
    k = np.arange(-1074, 1024)
    yk = host_lib.elementwise("log", np.ldexp(1.0, k))
    assert np.abs(yk - k * np.log(2.0)).max() <= 4 * np.spacing(1024 * np.log(2.0))


def test_exp_accuracy():
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-745.0, 709.7, 200000),
        rng.uniform(-1.0, 1.0, 100000),
        10.0 ** rng.uniform(-300, -1, 20000) * rng.choice([-1, 1], 20000),
        np.array([0.0, -0.0, 1.0, -1.0, 709.78, -708.39, -744.0, 0.5 * np.log(2.0)]),
    ])
    y = host_lib.elementwise("exp", x)
    ref = np.exp(x)
    normal = ref >= np.finfo(float).tiny
    assert ulps(y[normal], ref[normal]).max() <= 2.0
    sub = ~normal & (ref > 0)
    assert np.abs(y[sub] - ref[sub]).max() <= 2 * 5e-324     # subnormal results: absolute
    assert y[x == 0.0].tolist() == [1.0] * int((x == 0.0).sum())


def test_special_values():
    nan, inf = np.nan, np.inf
    yl = host_lib.elementwise("log", np.array([0.0, -0.0, inf, -1.0, -inf, nan]))
    assert yl[0] == -inf and yl[1] == -inf and yl[2] == inf
    assert np.isnan(yl[3:]).all()
    ye = host_lib.elementwise("exp", np.array([inf, -inf, nan, 710.0, 1e300, -800.0, -1e300]))
    assert ye[0] == inf and ye[1] == 0.0 and np.isnan(ye[2])
    assert ye[3] == inf and ye[4] == inf and ye[5] == 0.0 and ye[6] == 0.0


def test_round_trip():
    rng = np.random.default_rng(3)
    x = 10.0 ** rng.uniform(-300, 300, 50000)
    y = host_lib.elementwise("exp", host_lib.elementwise("log", x))
    # exp amplifies the log's absolute error by |log x| <= 691
    assert (np.abs(y / x - 1.0) <= 700 * 2.3e-16).all()
