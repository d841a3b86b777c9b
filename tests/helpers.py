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
                 "rst", "rgt", "rht")
SOURCES = ("rvs", "rcs", "ris", "rrs", "rss", "rgs", "rhs", "evap3d")
PRECIP = ("inprc", "inprr", "inpri", "inprs", "inprg", "inprh")


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


def check_with_flips(out, ref, names, flips, tol, max_bad=2e-3, max_flips=1e-2):
    """Parity with branch-flip accounting.

    flips: points whose ICE4 time splitting took a different number of tendency calls
    than the reference (a threshold decided the other way by a last-bit difference).
    A point is *divergent* when any of its outputs misses max(tol*|ref|, atol).  Every
    divergent point must be a flip, divergent points must stay below max_bad of the
    points and flips below max_flips.  Per-column outputs are checked on the columns
    without a divergent point.  Returns (worst error over non-divergent points,
    number of flips, number of divergent points)."""
    errs = {k: rel_err(out[k], ref[k], floor_for(k, tol)) for k in names}
    bad = np.zeros(flips.shape, dtype=bool)
    for r in errs.values():
        if r.shape == flips.shape:
            bad |= ~(r <= tol)
    worst = {}
    for k, r in errs.items():
        if r.shape == flips.shape:
            r = np.where(bad, 0.0, r)
        elif r.ndim == 1:
            r = np.where(bad.any(axis=0), 0.0, r)
        worst[k] = float(r.max())
    assert not (bad & ~flips).any(), f"{int((bad & ~flips).sum())} points beyond {tol} without a branch flip"
    assert bad.mean() <= max_bad, f"{int(bad.sum())} divergent points of {bad.size}"
    assert flips.mean() <= max_flips, f"{int(flips.sum())} flips of {flips.size}"
    over = {k: v for k, v in worst.items() if not v <= tol}
    assert not over, over
    return worst, int(flips.sum()), int(bad.sum())


def flips_of(out, ref, calls="calls", iters="iters"):
    """Branch flips of the ICE4 time splitting between two runs of the same points: a
    different number of tendency calls, or the pristine-ice concentration reset to 0 on one
    side only (RAIN_ICE sets cit = 0 when ri ends a sub-step at or below 0, and where a
    process such as OCND2's RILARS empties the ice, the sign of ri's rounding residual
    decides it; micro/mode_ice4_stepping.F90:471-474)."""
    return (out[calls] != ref[iters]) | ((out["ci"] == 0.0) != (ref["ci"] == 0.0))
