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


def ill_conditioned(run, s, names, tol, fields, delta=4.5e-16, draws=2, seed=20260090):
    """Points whose oracle outputs are not determined to tol by FP64 inputs: run(s) against
    run(s') where every field of `fields` is multiplied elementwise by 1 +- delta (a random
    sign per value; delta is two units in the last place), for `draws` sign patterns.  A
    point where any output of `names` moves by more than max(tol*|ref|, atol) sits on a
    discontinuity of the reference's algorithm (a threshold or a time-step limiter that a
    last-bit change decides the other way, without changing the number of tendency calls),
    so two correct FP64 evaluations may disagree there: it is counted as a branch flip."""
    base = run(s)
    rng = np.random.Generator(np.random.Philox(seed))
    mask = None
    for _ in range(draws):
        sp = dict(s)
        for k in fields:
            v = np.asarray(s[k], dtype=np.float64)
            sp[k] = v * (1.0 + delta * rng.choice([-1.0, 1.0], size=v.shape))
        o = run(sp)
        for k in names:
            r = rel_err(o[k], base[k], floor_for(k, tol))
            if r.ndim != 2:
                continue
            m = ~(r <= tol)
            mask = m if mask is None else (mask | m)
    return mask


def reproduced_by_oracle(out, run, s, names, tol, fields, points, scales=(1e-15, 1e-14, 1e-13)):
    """Of the given (level, column) points, those where the oracle, run on the point's column
    with one input value of the point changed by a relative 1e-15..1e-13 (one field at a
    time, both signs), gives out's values at the point within tol for every output of
    `names`.  Such a point sits on a discontinuity of the reference's algorithm (a species
    emptied to exactly 0 or to a rounding residue, a threshold met to the last bit) and out
    took the branch that a neighbouring input takes: a branch flip, shown by the oracle
    itself.  A wrong result is not reproduced this way."""
    found = np.zeros(len(points), dtype=bool)
    for n, (k, i) in enumerate(points):
        col = {f: (np.ascontiguousarray(v[:, i:i + 1]) if getattr(v, "ndim", 0) == 2 else
                   (np.ascontiguousarray(v[i:i + 1]) if getattr(v, "ndim", 0) == 1 else v))
               for f, v in s.items()}
        for f in fields:
            for e in scales:
                for sg in (1.0, -1.0):
                    sp = dict(col)
                    v = col[f].copy()
                    v[k, 0] = v[k, 0] * (1.0 + sg * e)
                    sp[f] = v
                    o = run(sp)
                    ok = True
                    for q in names:
                        if np.ndim(o[q]) != 2:
                            continue
                        r = rel_err(o[q][k, 0], out[q][k, i], floor_for(q, tol))
                        if not r <= tol:
                            ok = False
                            break
                    if ok:
                        found[n] = True
                        break
                if found[n]:
                    break
            if found[n]:
                break
    return found


# The input source and state at t of each output source (rain_ice state names).
SOURCE_INPUTS = {"ths": ("ths", "tht"), "rvs": ("rvs", "rvt"), "rcs": ("rcs", "rct"),
                 "rrs": ("rrs", "rrt"), "ris": ("ris", "rit"), "rss": ("rss", "rst"),
                 "rgs": ("rgs", "rgt"), "rhs": ("rhs", "rht")}


def cancellation_residuals(out, ref, s, names, tol, dt, ulps=8):
    """Points where every output of `names` beyond the bar differs from ref by at most `ulps`
    units in the last place of the input it was computed from (the input source, or the state
    at t over dt): an output that cancels its input to a rounding residue (RRHONG freezing all
    the rain, a species emptied by its sink) is fixed only to that, and where that residue is
    above the bar's absolute floor two correct FP64 evaluations may disagree there."""
    eps = np.finfo(np.float64).eps
    beyond = np.zeros(np.shape(ref[names[0]]), dtype=bool)
    within = np.ones_like(beyond)
    for k in names:
        if np.ndim(ref[k]) != 2:
            continue
        b = ~(rel_err(out[k], ref[k], floor_for(k, tol, dt)) <= tol)
        beyond |= b
        if k in SOURCE_INPUTS:
            src, t = SOURCE_INPUTS[k]
            scale = np.maximum(np.abs(s[src]), np.abs(s[t]) / dt)
            within &= ~b | (np.abs(out[k] - ref[k]) <= ulps * eps * scale)
        else:
            within &= ~b
    return beyond & within


def explained_flips(out, ref, run, s, names, tol, fields, calls="calls", iters="iters", ties=None):
    """The branch flips of out against the oracle's ref: flips_of, then ill_conditioned, then
    the remaining divergent points where the oracle's wet/dry growth choice rests on rounding
    residuals (ties: the oracle's "tie" branch record, oracle_lib.rounding_ties), then those
    that reproduced_by_oracle shows on a discontinuity."""
    flips = flips_of(out, ref, calls, iters) | ill_conditioned(run, s, names, tol, fields)
    bad = np.zeros(flips.shape, dtype=bool)
    for k in names:
        r = rel_err(out[k], ref[k], floor_for(k, tol))
        if r.shape == flips.shape:
            bad |= ~(r <= tol)
    if ties is not None:
        flips |= bad & ties
    pts = [tuple(p) for p in np.argwhere(bad & ~flips)]
    if pts and len(pts) <= 64:
        rep = reproduced_by_oracle(out, run, s, names, tol, fields, pts)
        for p, r in zip(pts, rep):
            flips[p] = flips[p] or r
    return flips


def every_point(res, f, make_state, run, names, state_fields, tol, dt=50.0, slab=16384, label=""):
    """Compares a GPU step's results res (device tensors, (nlev, ncol) or (ncol,)) with the
    oracle at every point.  The oracle runs on slabs of `slab` columns in parallel threads
    (its C entry points release the interpreter lock; its branch record is per thread):
    make_state(g) builds the rain_ice state of the fields g of one slab, run(s, **kw) runs the
    oracle's rain_ice variant on it.  A point beyond the bar must be a flip (flips_of); the
    divergent points that keep the oracle's call count are checked on their column alone:
    explained_flips with the ties of the oracle's branch record (wet/dry choices and graupel
    shares that rest on rounding residuals), and cancellation_residuals.  Returns (worst error
    per output over the other points, flips, divergent points, columns checked alone)."""
    import concurrent.futures as cf
    import os

    from oracle import oracle_lib
    ncol = f["th"].shape[1]
    B = oracle_lib.BRANCHES

    def oracle_slab(a):
        g = {k: np.ascontiguousarray(v[:, a:a + slab] if np.ndim(v) == 2 else v[a:a + slab])
             for k, v in f.items()}
        st = make_state(g)
        return a, st, run(st, with_iters=True)

    def cols(d, c):
        return {k: (np.ascontiguousarray(v[:, c]) if np.ndim(v) == 2 else
                    (np.ascontiguousarray(v[c]) if np.ndim(v) == 1 else v)) for k, v in d.items()}

    worst, nf, nb, checked = {}, 0, 0, 0
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        for a, st, ref in pool.map(oracle_slab, range(0, ncol, slab)):
            got = {k: v[..., a:a + slab].detach().cpu().numpy() for k, v in res.items()}
            flips = flips_of(got, ref)
            bad = np.zeros(flips.shape, dtype=bool)
            for k in names:
                r = rel_err(got[k], ref[k], floor_for(k, tol, dt))
                if r.shape == flips.shape:
                    bad |= ~(r <= tol)
            c = np.unique(np.nonzero(bad & ~flips)[1])
            if c.size:
                sc, gc, rc = cols(st, c), cols(got, c), cols(ref, c)
                _, rec = oracle_lib.record_branches(run, sc)
                ties = (((rec >> B.index("tie")) | (rec >> B.index("share_tie"))) & 1).astype(bool)
                fc = explained_flips(gc, rc, run, sc, names, tol, state_fields, ties=ties)
                fc |= cancellation_residuals(gc, rc, sc, names, tol, dt)
                for k, j in np.argwhere(bad[:, c] & ~flips[:, c] & ~fc)[:20]:
                    errs = {q: float(rel_err(gc[q][k, j], rc[q][k, j], floor_for(q, tol, dt)))
                            for q in names if np.ndim(gc[q]) == 2}
                    print(f"{label} level {k} column {a + c[j]}: calls {gc['calls'][k, j]} / "
                          f"{rc['iters'][k, j]}, oracle record {[n for i, n in enumerate(B) if rec[k, j] >> i & 1]}, "
                          f"not explained, beyond the bar "
                          f"{ {q: e for q, e in errs.items() if not e <= tol} }", flush=True)
                print(f"{label} columns {a}-{a + slab - 1}: {int((bad & ~flips).sum())} points beyond the "
                      f"bar with the oracle's call count, in {c.size} columns checked alone; "
                      f"{int((bad[:, c] & ~flips[:, c] & fc).sum())} explained", flush=True)
                flips[:, c] |= fc
                checked += c.size
            w, n1, n2 = check_with_flips(got, ref, names, flips, tol)
            worst = {k: max(worst.get(k, 0.0), v) for k, v in w.items()}
            nf, nb = nf + n1, nb + n2
    return worst, nf, nb, checked
