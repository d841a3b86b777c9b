"""Comparison helpers for the parity tests."""
import numpy as np

def rel_err(a, b, floor=0.0):
    """Elementwise |a-b| / max(|b|, floor); NaN where exactly one side is NaN, 0 where both are."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both = np.isnan(a) & np.isnan(b)
    one = np.isnan(a) ^ np.isnan(b)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(s > 0, d / s, np.where(d > 0, np.inf, 0.0))
    r = np.where(both, 0.0, r)
    r = np.where(one, np.inf, r)
    return r


MIXING_RATIOS = ("rv_out", "rc_out", "ri_out", "hlc_hrc", "hli_hri", "rvt", "rct", "rrt", "rit",
                 "rst", "rgt")
SOURCES = ("rvs", "rcs", "ris", "rrs", "rss", "rgs", "evap3d")
PRECIP = ("inprc", "inprr", "inpri", "inprs", "inprg")


def floor_for(name, tol, dt=50.0):
    """Scale below which rel_err measures absolute error, chosen so that
    rel_err <= tol  <=>  |a-b| <= max(tol*|b|, atol), with atol = 1e-18 kg/kg on
    mixing ratios, 1e-18/dt kg/kg/s on their sources and 1e-21 m/s on surface
    precipitation (1e-18 kg m-2 s-1 of water); other fields have no floor."""
    if name in MIXING_RATIOS:
        atol = 1e-18
    elif name in SOURCES:
        atol = 1e-18 / dt
    elif name in PRECIP:
        atol = 1e-21
    else:
        return 0.0
    return atol / tol
