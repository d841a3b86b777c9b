"""Parity of the HIP kernels (through the C ABI) with the compiled reference's golden
vectors and with the oracle, plus size-independent properties at benchmark sizes.

FP64 bar: relative error <= 1e-11 with an absolute floor of 1e-18 kg/kg on mixing
ratios (1e-18/dt kg/kg/s on their sources).  Points whose ICE4 time-splitting
trajectory differs from the reference's (a threshold decided the other way by a
last-bit difference: different number of tendency calls) are counted as branch
flips and reported.  A point that misses the bar must be such a flip, and fewer
than 0.2 % of the points may miss it (helpers.check_with_flips)."""
import os

import numpy as np
import pytest

from helpers import check_with_flips, floor_for, rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-11
ADJ_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
           "cph", "lv", "ls", "rvs", "rcs", "ris", "ths")
RAIN_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr",
            "inprc", "inprr", "inpri", "inprs", "inprg")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _np(t):
    return t.detach().cpu().numpy()


def _adjust_gpu(f, dt=50.0, precision="fp64"):
    from ice3_amd import IceAdjust
    ia = IceAdjust()
    args = [f[k] for k in ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri",
                           "rr", "rs", "rg", "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")]
    res = ia(f["sigqsat"], *args, timestep=dt, precision=precision, return_srcs=True)
    out = {k: _np(v) for k, v in zip(ADJ_OUT + ("srcs",), res)}
    return out


def _rain_state(s):
    """RainIceJAX state-dict names from the oracle/reference field names."""
    m = dict(exn="exn", dzz="dzz", rhodj="rhodj", rhodref="rhodref", exnref="exn_ref", pres="pabs",
             cf="cldfr", hlc_hrc="hlc_hrc", hlc_hcf="hlc_hcf", hli_hri="hli_hri", hli_hcf="hli_hcf",
             sigs="sigs", th_t="tht", rv_t="rvt", rc_t="rct", rr_t="rrt", ri_t="rit", rs_t="rst",
             rg_t="rgt", ci_t="ci", ths="ths", rvs="rvs", rcs="rcs", rrs="rrs", ris="ris",
             rss="rss", rgs="rgs")
    return {k: s[v] for k, v in m.items()}


def _rain_gpu(s, dt=50.0, precision="fp64"):
    from ice3_amd import RainIce
    new, diags = RainIce()(_rain_state(s), dt, precision=precision, return_calls=True)
    out = {k: _np(new[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "evap3d", "rainfr")}
    out["ci"] = _np(new["ci_t"])
    out.update({k: _np(v) for k, v in diags.items()})
    return out


def _check(out, ref, names, flips=None, tol=TOL):
    worst = {}
    for k in names:
        r = rel_err(out[k], ref[k], floor_for(k, tol))
        if flips is not None and r.shape == flips.shape:
            r = np.where(flips, 0.0, r)
        elif flips is not None and r.ndim == 1:  # per column: skip columns with a flip
            r = np.where(flips.any(axis=0), 0.0, r)
        worst[k] = float(r.max())
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, bad
    return worst


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_ice_adjust_golden(dev, golden_dir, tag):
    g = np.load(os.path.join(golden_dir, "ice_adjust_fp64.npz"))
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in ADJ_OUT + ("srcs",)}
    out = _adjust_gpu(f, float(g["dt"]))
    for k in ADJ_OUT:
        assert np.array_equal(np.isnan(out[k]), np.isnan(ref[k])), k
    w = _check(out, ref, ADJ_OUT + ("srcs",))
    print("ice_adjust golden max rel err", tag, max(w.values()))


ADJ_FIELDS = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
              "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")


def _every_adjust_point(out, ref, f, run, names, label):
    """Every point of an ICE_ADJUST result against the oracle's: within the FP64 bar, except
    points where the oracle itself is not determined to the bar by FP64 inputs
    (helpers.ill_conditioned: its outputs move beyond the bar when the inputs change by 2 ulp,
    as a cloud fraction 0.5 + 0.36 atan(1.55 Q1) that cancels to near 0), which must stay
    below 1e-5 of the points."""
    from helpers import ill_conditioned
    bad = np.zeros(f["th"].shape, dtype=bool)
    for k in names:
        bad |= ~(rel_err(out[k], ref[k], floor_for(k, TOL)) <= TOL)
    ill = ill_conditioned(run, f, names, TOL, ADJ_FIELDS) if bad.any() else bad
    for k, i in np.argwhere(bad)[:10]:
        print(f"level {k} column {i}: ill-conditioned {bool(ill[k, i])}, "
              f"{ {q: (float(out[q][k, i]), float(ref[q][k, i])) for q in names if not rel_err(out[q][k, i], ref[q][k, i], floor_for(q, TOL)) <= TOL} }")
    assert not (bad & ~ill).any(), np.argwhere(bad & ~ill)[:10]
    assert bad.mean() <= 1e-5, int(bad.sum())
    w = _check(out, ref, names, flips=bad)
    print(f"{label}, every point: {int(bad.sum())} ill-conditioned points beyond the bar of "
          f"{bad.size}, worst {max(w.values()):.2e} on the others")


def test_ice_adjust_vs_oracle_config2_full(dev):
    """ICE_ADJUST alone on the whole 256x256x90 workload it is measured on (the "adjust"
    generator, seed 20260002, dt 50 s): every one of its 5,898,240 points against the oracle
    (_every_adjust_point)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib
    f = synthetic.generate(256 * 256, 90, seed=20260002, variant="adjust", dt=50.0)
    _every_adjust_point(_adjust_gpu(f), oracle_lib.ice_adjust(f), f, oracle_lib.ice_adjust,
                        ADJ_OUT + ("srcs",), "ice_adjust 256x256x90")


def test_ice_adjust_ocnd2_vs_oracle_config2_full(dev):
    """ICE_ADJUST with OCND2 on the 256x256x90 workload it is measured on (the same inputs,
    the generator's level heights, ice cloud weight 1): every point against the oracle
    (_every_adjust_point)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib
    f = synthetic.generate(256 * 256, 90, seed=20260002, variant="adjust", dt=50.0)
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.ones(256 * 256)
    _every_adjust_point(_adjust_ocnd2_gpu(f), oracle_lib.ice_adjust_ocnd2(f), f, oracle_lib.ice_adjust_ocnd2,
                        OCND2_CHECKED, "ice_adjust OCND2 256x256x90")


OCND2_X = ("icldfr", "wcldfr", "ssio", "ssiu", "ifr")
OCND2_CHECKED = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
                 "rvs", "rcs", "ris", "ths", "srcs") + OCND2_X


def _adjust_ocnd2_gpu(f, dt=50.0):
    from ice3_amd import IceAdjustOCND2
    args = [f[k] for k in ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri",
                           "rr", "rs", "rg", "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")]
    res = IceAdjustOCND2()(f["sigqsat"], *args, timestep=dt, zz=f["zz"], ice_cld_wgt=f["ice_cld_wgt"],
                           return_srcs=True)
    return {k: _np(v) for k, v in zip(ADJ_OUT + OCND2_X + ("srcs",), res)}


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_ice_adjust_ocnd2_golden(dev, golden_dir, tag):
    """ICE_ADJUST with OCND2 (ICECLOUD over the TIWMX tables, liquid-only condensate) against
    the compiled reference's vectors."""
    g = np.load(os.path.join(golden_dir, "ice_adjust_ocnd2_fp64.npz"))
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in OCND2_CHECKED}
    out = _adjust_ocnd2_gpu(f, float(g["dt"]))
    for k in OCND2_CHECKED:
        assert np.array_equal(np.isnan(out[k]), np.isnan(ref[k])), k
    w = _check(out, ref, OCND2_CHECKED)
    print("ice_adjust OCND2 golden max rel err", tag, max(w.values()))


def test_ice_adjust_ocnd2_vs_oracle_slab(dev):
    from ice3_amd import synthetic
    from oracle import oracle_lib
    f = synthetic.generate(2048, 90, seed=20260036)
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.random.default_rng(7).uniform(0.5, 1.5, 2048)
    out = _adjust_ocnd2_gpu(f)
    ref = oracle_lib.ice_adjust_ocnd2(f)
    w = _check(out, ref, OCND2_CHECKED)
    print("ice_adjust OCND2 vs oracle, 2048 columns: max rel err", max(w.values()))


def _rain_case(g, tag):
    p = "" if tag == "base" else tag + "_"
    s = {k[len(p) + 3:]: g[k] for k in g.files if k.startswith(p + "in_")}
    return s, {k: g[p + "out_" + k] for k in RAIN_OUT}


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_golden(dev, golden_dir, tag):
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    s, ref = _rain_case(g, tag)
    out = _rain_gpu(s, float(g["dt"]))
    iters = oracle_lib.rain_ice(s, float(g["dt"]), with_iters=True)["iters"]
    flips = out["calls"] != iters
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"rain_ice golden {tag}: {nf} branch flips, {nb} divergent points of {flips.size}, "
          f"worst {max(w.values()):.2e}")


def test_pipeline_vs_oracle(dev):
    """Fused ice_adjust + rain_ice on the benchmark generator, 256 columns."""
    from ice3_amd import AdjustRainIce, synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(256, 90, seed=20260003, variant="pipeline")
    res = AdjustRainIce()({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    a = oracle_lib.ice_adjust(f)
    s = ref_lib.pipeline_state(f, a)
    ref = oracle_lib.rain_ice(s, with_iters=True)
    flips = out["calls"] != ref["iters"]
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"pipeline: {nf} branch flips, {nb} divergent points of {flips.size}, worst {max(w.values()):.2e}")
    _check(out, a, ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf"))


# Columns of the 512x512x90 workload compared with the oracle: a stride coprime to 64, so
# that the sampled columns fall on every lane of the microphysics kernel's 64-point chunks
# (and of the column walk's 64-thread blocks), parked-and-resumed points included.
FULL_STRIDE = 127


@pytest.fixture(scope="module")
def full512():
    """The fused benchmark workload (512x512 columns x 90 levels, seed 20260003) and the
    oracle's results on every FULL_STRIDE-th column (2,065 columns, 185,850 points)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(512 * 512, 90, seed=20260003, variant="pipeline")
    cols = np.arange(0, 512 * 512, FULL_STRIDE)
    g = {k: np.ascontiguousarray(v[:, cols] if v.ndim == 2 else v[cols]) for k, v in f.items()}
    ref = oracle_lib.rain_ice(ref_lib.pipeline_state(g, oracle_lib.ice_adjust(g)), with_iters=True)
    lanes = np.unique(cols % 64)
    assert lanes.size == 64, lanes.size
    return f, cols, ref


def _column_budget(dev, f, res, dt=50.0):
    """Max over columns of |change of the column's total water - surface precipitation| /
    column water (the sources carry the water; precipitation leaves at the ground)."""
    rho = torch.as_tensor(f["rho_dry_ref"], device=dev)
    dz = torch.as_tensor(f["dzz"], device=dev)
    w0 = sum(torch.as_tensor(f[k], device=dev) for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs"))
    w1 = sum(res[k] for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs"))
    col0 = (rho * dz * w0).sum(0) * dt
    col1 = (rho * dz * w1).sum(0) * dt
    out = -1000.0 * (res["inprr"] + res["inprs"] + res["inprg"]) * dt
    return ((col1 - col0 - out).abs() / col0).max().item()


def _sampled(res, cols, dev):
    idx = torch.as_tensor(cols, device=dev)
    return {k: _np(v.index_select(v.dim() - 1, idx)) for k, v in res.items()}


def test_pipeline_full_size_water_budget(dev, full512):
    """512x512 columns x 90 levels (the fused benchmark workload): every column's total
    water changes only by its surface precipitation; sources stay finite and
    non-negative for the condensed species; and every 127th column (2,065 columns over all
    64 lanes, 185,850 points) matches the oracle at the FP64 bar, branch flips counted."""
    from ice3_amd import AdjustRainIce
    f, cols, ref = full512
    res = AdjustRainIce()({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True)
    err = _column_budget(dev, f, res)
    print("full-size column water budget max rel err", err)
    assert err < 1e-12
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        assert torch.isfinite(res[k]).all() and (res[k] >= 0).all(), k
    got = _sampled(res, cols, dev)
    del res
    flips = got["calls"] != ref["iters"]
    w, nf, nb = check_with_flips(got, ref, RAIN_OUT, flips, TOL)
    lane_flips = np.bincount((cols[None, :] % 64 * np.ones((90, 1), dtype=int))[flips], minlength=64)
    print(f"512x512 step, sampled columns on all 64 lanes: {nf} branch flips, {nb} divergent points "
          f"of {flips.size}, worst {max(w.values()):.2e}; lanes with a flip: {int((lane_flips > 0).sum())}")


RAIN_STATE = ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "ths", "rvs", "rcs", "rrs", "ris",
              "rss", "rgs")


def test_pipeline_full_size_every_point(dev, full512):
    """512x512 columns x 90 levels (the fused benchmark workload), every one of its
    23,592,960 points against the oracle at the FP64 bar, branch flips counted
    (helpers.every_point: the oracle on 16 slabs of 16,384 columns in parallel threads).  A
    point beyond the bar must be a flip: a different number of tendency calls, or (checked on
    its column alone, as the 2,048-column tests do) an output that last-bit changes of the
    inputs move beyond the bar, a wet/dry choice or a graupel share of riming or accretion
    that the oracle records as resting on rounding residuals, a value the oracle reproduces
    from an input changed in its last digits, or a source that cancels its input to a
    rounding residue (within 8 units in the last place of the input source or the state at t
    over dt, as where RRHONG freezes all the rain)."""
    from helpers import every_point
    from ice3_amd import AdjustRainIce
    from oracle import oracle_lib, ref_lib
    f = full512[0]
    res = AdjustRainIce()({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True)
    worst, nf, nb, checked = every_point(
        res, f, lambda g: ref_lib.pipeline_state(g, oracle_lib.ice_adjust(g)),
        lambda x, **kw: oracle_lib.rain_ice(x, 50.0, **kw), RAIN_OUT, RAIN_STATE, TOL, label="fused step")
    npts = 90 * 512 * 512
    print(f"512x512 step, every point: {nf} branch flips ({nf / npts:.1e}), {nb} divergent points "
          f"(all at flips) of {npts}, worst {max(worst.values()):.2e}; {checked} columns checked "
          f"alone for flips other than a call-count change")
    assert nb <= nf and nf <= 1e-2 * npts


# Mixed path (config 4) against the FP64 oracle: largest error relative to the field's
# largest magnitude, and RMS error relative to the field's RMS, per output field.
MIXED_MAX_REL = 1e-4
MIXED_RMS_REL = 2e-5


def test_mixed_precision_full_size_vs_oracle(dev, full512):
    """The mixed path (float process rates, slopes and lookups; double saturation adjustment,
    state update, time splitting and sedimentation fluxes) on the 512x512x90 benchmark
    workload, at every point: each output within MIXED_MAX_REL of the oracle field's largest
    magnitude and MIXED_RMS_REL of its RMS (the oracle on 16 slabs of 16,384 columns in
    parallel threads), the points where the ice concentration is reset to 0 on one side only
    (a branch flip) counted apart and below 1e-4 of the points; the condensed species' sources non-negative and each column's total
    water changed only by its surface precipitation (the budget is carried in double)."""
    import concurrent.futures as cf
    from ice3_amd import AdjustRainIce
    from oracle import oracle_lib, ref_lib
    f = full512[0]
    res = AdjustRainIce()({**f, "rhodref": f["rho_dry_ref"]}, 50.0, precision="mixed")
    err = _column_budget(dev, f, res)
    print("mixed, full-size column water budget max rel err", err)
    assert err < 1e-12
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        assert bool((res[k] >= 0).all()), k
    names = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "inprr", "inprs", "inprg")
    ncol, slab = 512 * 512, 16384

    def oracle_slab(a):
        g = {k: np.ascontiguousarray(v[:, a:a + slab] if v.ndim == 2 else v[a:a + slab]) for k, v in f.items()}
        return a, oracle_lib.rain_ice(ref_lib.pipeline_state(g, oracle_lib.ice_adjust(g)))

    # per field: max |err|, max |ref|, sum err^2, sum ref^2, count
    acc = {k: [0.0, 0.0, 0.0, 0.0, 0] for k in names}
    resets = 0
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        for a, ref in pool.map(oracle_slab, range(0, ncol, slab)):
            for k in names:
                g = _np(res[k][..., a:a + slab])
                d = np.abs(g - ref[k])
                if k == "ci":
                    # the ice concentration reset to 0 on one side only: a branch decided by
                    # the sign of ri's rounding residue (micro/mode_ice4_stepping.F90:471-474)
                    one = (g == 0.0) != (ref[k] == 0.0)
                    resets += int(one.sum())
                    d = np.where(one, 0.0, d)
                m = acc[k]
                m[0], m[1] = max(m[0], float(d.max())), max(m[1], float(np.abs(ref[k]).max()))
                m[2], m[3], m[4] = m[2] + float((d ** 2).sum()), m[3] + float((ref[k] ** 2).sum()), m[4] + d.size
    stats = {k: (m[0] / m[1] if m[1] > 0 else 0.0, float(np.sqrt(m[2] / m[3])) if m[3] > 0 else 0.0)
             for k, m in acc.items()}
    print("mixed vs oracle (512x512x90, every point): max err / field max, RMS err / field RMS",
          {k: f"{a:.1e}, {b:.1e}" for k, (a, b) in stats.items()},
          f"; ice concentration reset on one side only at {resets} points")
    bad = {k: v for k, v in stats.items() if not (v[0] <= MIXED_MAX_REL and v[1] <= MIXED_RMS_REL)}
    assert not bad, bad
    assert resets <= 1e-4 * 90 * ncol, resets


def _fused_raw(lib, ctx, dev, f, ncol, nlev, ld, stages=(7,), calls=False):
    """The fused step through the C ABI on fields padded to a row stride ld >= ncol;
    stages: the ICE3_STAGE_* masks of successive ice3_adjust_rain_ice_stages calls."""
    import ctypes
    from ice3_amd import _lib
    from ice3_amd._fields import ptr
    g = {}
    for k in _lib.PIPE_IN_FIELDS:
        v = torch.as_tensor(f[k], dtype=torch.float64, device=dev)
        if v.dim() == 2:
            pad = torch.full((nlev, ld), -1.0, dtype=torch.float64, device=dev)
            pad[:, :ncol] = v
            v = pad
        g[k] = v.contiguous()
    o = {k: torch.full((nlev, ld), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
    o.update({k: torch.full((ncol,), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
    c = torch.zeros(nlev * ld, dtype=torch.int32, device=dev)
    cnt = torch.zeros(3, dtype=torch.int64, device=dev)
    st = _lib.Stats(ptr(c), ptr(cnt))
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    cout = _lib.PipeOut(*[ptr(o[k]) for k in _lib.PIPE_OUT_FIELDS])
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    for m in stages:
        rc = lib.ice3_adjust_rain_ice_stages(ctx, 0, ncol, nlev, ld, 50.0, ctypes.byref(cin),
                                             ctypes.byref(cout), ctypes.byref(st), m, sp)
        assert rc == 0, rc
    torch.cuda.synchronize()
    res = {k: (v[:, :ncol] if v.dim() == 2 else v).cpu().numpy() for k, v in o.items()}
    res["calls"] = c.view(nlev, ld)[:, :ncol].cpu().numpy()
    res["counters"] = cnt.cpu().numpy()
    return res


def test_fused_layout_and_stages_bitwise(dev):
    """The microphysics kernel runs 64-column chunks with the points still being
    time-stepped parked and resumed across chunks: a row stride ld > ncol, a partial last
    chunk of each level, and the step split into separate stage calls must all give the
    dense single-call results bit for bit (the per-point arithmetic does not depend on the
    order in which points are advanced)."""
    from ice3_amd import _lib, synthetic
    lib, ctx = _lib.load(), _lib.context(0)
    ncol, nlev = 1000, 12                      # 15 full chunks + 40 columns per level
    f = synthetic.generate(ncol, nlev, seed=20260011, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    dense = _fused_raw(lib, ctx, dev, f, ncol, nlev, ncol)
    assert dense["counters"][1] == 0 and (dense["calls"] > 1).any()
    for ld, stages in ((1003, (7,)), (1000, (1 | 2, 4)), (1064, (1, 2, 4))):
        other = _fused_raw(lib, ctx, dev, f, ncol, nlev, ld, stages)
        for k in _lib.PIPE_OUT_FIELDS + ("calls", "counters"):
            assert np.array_equal(np.nan_to_num(other[k], nan=1.0), np.nan_to_num(dense[k], nan=1.0)), (ld, stages, k)


def _fused_hail_raw(lib, ctx, dev, f, ncol, nlev, ld):
    """The fused KRR=7 step (ice3_adjust_rain_ice_hail) through the C ABI on fields padded to a
    row stride ld >= ncol."""
    import ctypes
    from ice3_amd import _lib
    from ice3_amd._fields import ptr

    def padded(v, fill):
        v = torch.as_tensor(v, dtype=torch.float64, device=dev)
        if v.dim() == 2:
            pad = torch.full((nlev, ld), fill, dtype=torch.float64, device=dev)
            pad[:, :ncol] = v
            v = pad
        return v.contiguous()
    g = {k: padded(f[k], -1.0) for k in _lib.PIPE_IN_FIELDS}
    rh, rhs = padded(f["rh"], -1.0), padded(f["rhs"], -1.0)
    o = {k: torch.full((nlev, ld), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
    o.update({k: torch.full((ncol,), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
    o["rhs"] = torch.full((nlev, ld), -7.0, dtype=torch.float64, device=dev)
    o["inprh"] = torch.full((ncol,), -7.0, dtype=torch.float64, device=dev)
    c = torch.zeros(nlev * ld, dtype=torch.int32, device=dev)
    cnt = torch.zeros(3, dtype=torch.int64, device=dev)
    st = _lib.Stats(ptr(c), ptr(cnt))
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    cout = _lib.PipeOut(*[ptr(o[k]) for k in _lib.PIPE_OUT_FIELDS])
    hail = _lib.PipeHail(ptr(rh), ptr(rhs), ptr(o["rhs"]), ptr(o["inprh"]))
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib.ice3_adjust_rain_ice_hail(ctx, ncol, nlev, ld, 50.0, ctypes.byref(cin), ctypes.byref(cout),
                                       ctypes.byref(hail), ctypes.byref(st), sp)
    assert rc == 0, rc
    torch.cuda.synchronize()
    res = {k: (v[:, :ncol] if v.dim() == 2 else v).cpu().numpy() for k, v in o.items()}
    res["calls"] = c.view(nlev, ld)[:, :ncol].cpu().numpy()
    res["counters"] = cnt.cpu().numpy()
    return res


@pytest.mark.parametrize("ncol,level", [(1000, None), (333, 1)])
def test_fused_hail_layout_bitwise(dev, ncol, level):
    """The persistent hail microphysics (64-column chunks, parks) with a row stride ld > ncol,
    a partial last chunk of each level and a single level (one level of the 12-level
    generator) gives the dense call's results bit for bit (KRR=7: rh = rg U(0.2, 2), its
    source rh / dt)."""
    from ice3_amd import _lib, synthetic
    lib, ctx = _lib.load(), _lib.context(0)
    f = synthetic.generate(ncol, 12, seed=20260061, variant="pipeline")
    if level is not None:
        f = {k: np.ascontiguousarray(v[level:level + 1]) if v.ndim == 2 else v for k, v in f.items()}
    nlev = f["th"].shape[0]
    f["rhodref"] = f["rho_dry_ref"]
    rng = np.random.Generator(np.random.Philox(20260062))
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
    f["rhs"] = f["rh"] / 50.0
    dense = _fused_hail_raw(lib, ctx, dev, f, ncol, nlev, ncol)
    assert dense["counters"][1] == 0 and (dense["calls"] > 1).any()
    for ld in (ncol + 3, ncol + 64):
        other = _fused_hail_raw(lib, ctx, dev, f, ncol, nlev, ld)
        for k in dense:
            assert np.array_equal(np.nan_to_num(other[k], nan=1.0), np.nan_to_num(dense[k], nan=1.0)), (ld, k)


def test_nonfinite_inputs(dev):
    """ice3_count_nonfinite counts NaN and infinities per field (and nothing in the row
    padding beyond ncol); the drop-ins' check_inputs=True names the fields that hold any
    and leaves clean results unchanged; without the check, a NaN or an infinity in one
    column changes no other column, and every point still ends its time step."""
    import ctypes
    from ice3_amd import AdjustRainIce, IceAdjust, _lib, synthetic
    from ice3_amd._fields import ptr
    lib = _lib.load()
    gen = torch.Generator().manual_seed(5)
    ncol, ld, nlev = 300, 320, 90
    fields = [torch.rand((nlev, ld), generator=gen, dtype=torch.float64) for _ in range(3)]
    fields.append(torch.rand((ld,), generator=gen, dtype=torch.float64))
    fields[0][5, 7] = float("nan"); fields[0][89, 299] = float("inf")
    fields[2][0, 0] = float("-inf"); fields[3][17] = float("nan")
    for f in fields[:3]:
        f[:, ncol:] = float("nan")  # padding beyond ncol is not read
    fields[3][ncol:] = float("nan")
    want = [int((~torch.isfinite(f[..., :ncol])).sum()) for f in fields]
    assert want == [2, 0, 1, 1]
    dfields = [f.to(dev) for f in fields]
    counts = torch.zeros(4, dtype=torch.int64, device=dev)
    fp = (ctypes.c_void_p * 4)(*[ptr(f).value for f in dfields])
    rows = (ctypes.c_int * 4)(nlev, nlev, nlev, 1)
    for _ in range(2):  # accumulates
        assert lib.ice3_count_nonfinite(4, fp, rows, ncol, ld, ptr(counts), None) == _lib.ICE3_OK
    assert counts.tolist() == [2 * w for w in want]

    f = synthetic.generate(64, 90, seed=20260015, variant="pipeline")
    step = AdjustRainIce()
    clean = step({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True)
    checked = step({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True, check_inputs=True)
    for k in clean:
        assert torch.equal(clean[k], checked[k]), k
    names = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
             "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
    for fld, val in (("th", np.nan), ("rrs", np.inf), ("sigs", np.nan), ("rhodref", np.inf)):
        g = {k: np.array(v, copy=True) for k, v in f.items()}
        g["rhodref"] = g["rho_dry_ref"]
        g[fld][[0, 30, 60, 89], 3] = val
        with pytest.raises(ValueError, match=fld):
            step(g, 50.0, check_inputs=True)
        if fld in names:
            with pytest.raises(ValueError, match=fld):
                IceAdjust()(g["sigqsat"], *[g[k] for k in names], timestep=50.0, check_inputs=True)
        res = step(g, 50.0, return_calls=True)
        assert step.last_counters["capped_points"] == 0
        for k in clean:
            a, b = _np(res[k]), _np(clean[k])
            keep = np.arange(a.shape[-1]) != 3
            assert np.array_equal(a[..., keep], b[..., keep], equal_nan=True), (fld, k)
    print("non-finite inputs: counted per field, named by check_inputs, confined to their column")


def test_empty_and_tiny_domains(dev):
    """An empty domain (no column or no level) is a call that does nothing and succeeds,
    through the C ABI (NULL field pointers) and the drop-ins (empty arrays in, empty arrays
    out); negative sizes are still rejected.  One column of two levels (a 64-point chunk
    with one live lane, a two-level sedimentation walk) matches the oracle."""
    import ctypes
    from ice3_amd import AdjustRainIce, IceAdjust, RainIce, _lib, synthetic
    from oracle import oracle_lib, ref_lib
    lib, ctx = _lib.load(), _lib.context(0)
    pin, pout = _lib.PipeIn(), _lib.PipeOut()
    ain, aout = _lib.AdjustIn(), _lib.AdjustOut()
    rin, rio, rout = _lib.RainIn(), _lib.RainIO(), _lib.RainOut()
    for ncol, nlev in ((0, 90), (16, 0), (0, 0)):
        for prec in (0, 1):
            assert lib.ice3_ice_adjust(ctx, prec, ncol, nlev, ncol, 50.0, ctypes.byref(ain),
                                       ctypes.byref(aout), None) == _lib.ICE3_OK
            assert lib.ice3_rain_ice(ctx, prec, ncol, nlev, ncol, 50.0, ctypes.byref(rin), ctypes.byref(rio),
                                     ctypes.byref(rout), None, None) == _lib.ICE3_OK
            assert lib.ice3_adjust_rain_ice(ctx, prec, ncol, nlev, ncol, 50.0, ctypes.byref(pin),
                                            ctypes.byref(pout), None, None) == _lib.ICE3_OK
    assert lib.ice3_adjust_rain_ice(ctx, 0, -1, 90, 0, 50.0, ctypes.byref(pin), ctypes.byref(pout),
                                    None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_adjust_rain_ice(ctx, 0, 0, 90, 0, 0.0, ctypes.byref(pin), ctypes.byref(pout),
                                    None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_adjust_rain_ice(ctx, 7, 0, 90, 0, 50.0, ctypes.byref(pin), ctypes.byref(pout),
                                    None, None) == _lib.ICE3_EINVAL

    f = synthetic.generate(4, 90, seed=20260013, variant="pipeline")
    empty = {k: (v[:, :0] if v.ndim == 2 else v[:0]) for k, v in f.items()}
    names = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
             "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
    res = IceAdjust()(empty["sigqsat"], *[empty[k] for k in names], timestep=50.0)
    assert len(res) == 16 and all(tuple(r.shape) == (90, 0) for r in res)
    state = {k: v[:, :0] for k, v in _rain_state(ref_lib.pipeline_state(f, oracle_lib.ice_adjust(f))).items()}
    new, diags = RainIce()(state, 50.0)
    assert tuple(new["rrs"].shape) == (90, 0) and tuple(diags["inprr"].shape) == (0,)
    out = AdjustRainIce()({**empty, "rhodref": empty["rho_dry_ref"]}, 50.0)
    assert tuple(out["ths"].shape) == (90, 0)

    g = synthetic.generate(1, 2, seed=20260014, variant="pipeline")
    res = AdjustRainIce()({**g, "rhodref": g["rho_dry_ref"]}, 50.0, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    ref = oracle_lib.rain_ice(ref_lib.pipeline_state(g, oracle_lib.ice_adjust(g)), with_iters=True)
    check_with_flips(out, ref, RAIN_OUT, out["calls"] != ref["iters"], TOL, max_bad=1.0, max_flips=1.0)


def test_substep_cap_status(dev):
    """The sub-step cap through the C ABI: with the default cap no point stops and
    ice3_ctx_status returns ICE3_OK; with a cap of one tendency call, the status query (no
    ice3_stats passed to the step) returns ICE3_ESUBSTEP and the number of points that needed
    more calls, then resets; the capped points still conserve each column's water (their
    external tendencies count once); the drop-in raises."""
    import ctypes
    from ice3_amd import AdjustRainIce, _lib, synthetic
    lib, ctx = _lib.load(), _lib.context(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    n = ctypes.c_ulonglong(7)
    assert lib.ice3_ctx_status(ctx, sp, ctypes.byref(n)) == _lib.ICE3_OK  # clears earlier calls
    f = synthetic.generate(512, 90, seed=20260016, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    ncol, nlev = 512, 90
    free = _fused_raw(lib, ctx, dev, f, ncol, nlev, ncol)
    assert lib.ice3_ctx_status(ctx, sp, ctypes.byref(n)) == _lib.ICE3_OK and n.value == 0
    more = np.abs(free["calls"]) > 1
    assert lib.ice3_ctx_set_substep_cap(ctx, 0) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_set_substep_cap(ctx, 1) == _lib.ICE3_OK
    try:
        g = {k: torch.as_tensor(f[k], device=dev) for k in _lib.PIPE_IN_FIELDS}
        o = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
        o.update({k: torch.empty(ncol, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
        from ice3_amd._fields import ptr
        cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
        cout = _lib.PipeOut(*[ptr(o[k]) for k in _lib.PIPE_OUT_FIELDS])
        assert lib.ice3_adjust_rain_ice(ctx, 0, ncol, nlev, ncol, 50.0, ctypes.byref(cin), ctypes.byref(cout),
                                        None, sp) == _lib.ICE3_OK
        assert lib.ice3_ctx_status(ctx, sp, ctypes.byref(n)) == _lib.ICE3_ESUBSTEP
        assert n.value == int(more.sum()) > 0, (n.value, int(more.sum()))
        assert lib.ice3_ctx_status(ctx, sp, None) == _lib.ICE3_OK  # reset by the query
        err = _column_budget(dev, f, o)
        print(f"cap of one call: {n.value} capped points reported, column water budget max rel err {err:.1e}")
        assert err < 1e-12
        for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
            a = _np(o[k])
            assert np.array_equal(a[~more], free[k][~more]), k
        with pytest.raises(_lib.Ice3Error, match="sub-step cap"):
            AdjustRainIce()(f, 50.0)
    finally:
        assert lib.ice3_ctx_set_substep_cap(ctx, 4096) == _lib.ICE3_OK
        lib.ice3_ctx_status(ctx, sp, None)


def _to_jax(a, nx):
    """(nlev, ncol) -> (nx, ncol // nx, nlev), the JAX components' layout (vertical last)."""
    nlev, ncol = a.shape
    return np.ascontiguousarray(a.T.reshape(nx, ncol // nx, nlev))


def _from_jax(a):
    a = _np(a) if hasattr(a, "detach") else np.asarray(a)
    return a.reshape(-1, a.shape[-1]).T


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_ice_adjust_golden_jax_layout(dev, golden_dir, tag):
    """IceAdjustJAX's call on (nx, ny, nz) arrays, the default layout of IceAdjust: the golden
    vectors reshaped to JAX arrays (sigqsat as a 3-D field too, as IceAdjustJAX takes it)
    give the 16-tuple in (nx, ny, nz) that matches the compiled reference at the FP64 bar."""
    from ice3_amd import IceAdjust
    g = np.load(os.path.join(golden_dir, "ice_adjust_fp64.npz"))
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    nlev, ncol = f["th"].shape
    nx = 4
    names = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
             "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
    sq3 = _to_jax(np.broadcast_to(f["sigqsat"][None, :], (nlev, ncol)), nx)
    res = IceAdjust()(sq3, *[_to_jax(f[k], nx) for k in names], timestep=float(g["dt"]))
    assert len(res) == 16 and all(tuple(r.shape) == (nx, ncol // nx, nlev) for r in res)
    out = {k: _from_jax(v) for k, v in zip(ADJ_OUT, res)}
    ref = {k: g[f"{tag}_out_{k}"] for k in ADJ_OUT}
    for k in ADJ_OUT:
        assert np.array_equal(np.isnan(out[k]), np.isnan(ref[k])), k
    w = _check(out, ref, ADJ_OUT)
    print("ice_adjust golden, JAX (nx, ny, nz) arrays: max rel err", tag, max(w.values()))


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_golden_jax_layout(dev, golden_dir, tag):
    """RainIceJAX's state dict of (ngpblks, nproma, nflevg) arrays, the default layout of
    RainIce: the golden vectors in that form give sources in that form and surface
    precipitation as (ngpblks, nproma), matching the compiled reference at the FP64 bar."""
    from ice3_amd import RainIce
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    s, ref = _rain_case(g, tag)
    nlev, ncol = s["tht"].shape
    nblk = 4
    state = {k: _to_jax(v, nblk) for k, v in _rain_state(s).items()}
    new, diags = RainIce()(state, float(g["dt"]), return_calls=True)
    shp = (nblk, ncol // nblk, nlev)
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci_t", "evap3d", "rainfr"):
        assert tuple(new[k].shape) == shp, k
    for k in ("inprc", "inprr", "inpri", "inprs", "inprg"):
        assert tuple(diags[k].shape) == shp[:2], k
    assert new["th_t"] is state["th_t"]  # the t-state is returned as given (Fortran semantics)
    out = {k: _from_jax(new[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "evap3d", "rainfr")}
    out["ci"] = _from_jax(new["ci_t"])
    out.update({k: _np(diags[k]).reshape(-1) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")})
    out["calls"] = _from_jax(diags["calls"])
    iters = oracle_lib.rain_ice(s, float(g["dt"]), with_iters=True)["iters"]
    flips = out["calls"] != iters
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"rain_ice golden {tag}, JAX (ngpblks, nproma, nflevg) state: {nf} flips, worst {max(w.values()):.2e}")


def test_fused_step_jax_layout_bitwise(dev):
    """AdjustRainIce on (nx, ny, nz) arrays gives, bit for bit, its (nlev, ncol) results
    reshaped, per-column outputs as (nx, ny)."""
    from ice3_amd import AdjustRainIce, synthetic
    f = synthetic.generate(256, 90, seed=20260017, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    step = AdjustRainIce()
    col = step(f, 50.0, return_calls=True)
    j = {k: (_to_jax(v, 16) if v.ndim == 2 else v.reshape(16, 16)) for k, v in f.items()}
    jx = step(j, 50.0, return_calls=True)
    for k, v in col.items():
        a = _np(jx[k])
        if v.dim() == 2:
            assert a.shape == (16, 16, 90), k
            a = _from_jax(a)
        else:
            assert a.shape == (16, 16), k
            a = a.reshape(-1)
        assert np.array_equal(a, _np(v), equal_nan=True), k


# ------------------------------------------------------------- RAIN_ICE with OCND2

def _with_heights(f, seed):
    from ice3_amd import synthetic
    ncol = f["th"].shape[1]
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.random.Generator(np.random.Philox(seed)).uniform(0.5, 1.5, ncol)
    return f


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_ocnd2_golden(dev, golden_dir, tag):
    """RAIN_ICE with LOCND2 through RainIceOCND2 against the compiled reference's vectors,
    flips counted (call counts and the ri <= 0 ice-concentration reset)."""
    from helpers import flips_of
    from ice3_amd import RainIceOCND2
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_ocnd2_fp64.npz"))
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in RAIN_OUT}
    state = _rain_state(s)
    state.update({k: s[k] for k in ("icldfr", "ssio", "ssiu", "ifr")})
    new, diags = RainIceOCND2()(state, float(g["dt"]), return_calls=True)
    out = {k: _np(new[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "evap3d", "rainfr")}
    out["ci"] = _np(new["ci_t"])
    out.update({k: _np(v) for k, v in diags.items()})
    ref["iters"] = oracle_lib.rain_ice_ocnd2(s, float(g["dt"]), with_iters=True)["iters"]
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips_of(out, ref), TOL)
    print(f"rain_ice OCND2 golden {tag}: {nf} flips, {nb} divergent of {out['calls'].size}, worst {max(w.values()):.2e}")


def test_pipeline_ocnd2_vs_oracle(dev):
    """The fused step with OCND2 in both halves on the benchmark generator (2,048 columns)
    against the oracle, flips counted; the column water budget holds; the two OCND2 entry
    points (fused, and ice_adjust then rain_ice) agree bit for bit."""
    from helpers import flips_of
    from ice3_amd import AdjustRainIceOCND2, IceAdjustOCND2, RainIceOCND2, synthetic
    from oracle import oracle_lib, ref_lib
    f = _with_heights(synthetic.generate(2048, 90, seed=20260048, variant="pipeline"), 20260049)
    f["rhodref"] = f["rho_dry_ref"]
    res = AdjustRainIceOCND2()(f, 50.0, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    a = oracle_lib.ice_adjust_ocnd2(f)
    ref = oracle_lib.rain_ice_ocnd2(ref_lib.pipeline_state_ocnd2(f, a), with_iters=True)
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips_of(out, ref), TOL)
    _check(out, a, ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf", "icldfr", "wcldfr", "ssio", "ssiu", "ifr"))
    print(f"fused OCND2 step, 2048 columns: {nf} flips, {nb} divergent of {out['calls'].size}, "
          f"worst {max(w.values()):.2e}")
    assert _column_budget(dev, f, res) < 1e-12
    # the same step through the two separate OCND2 entry points
    names = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
             "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
    adj = IceAdjustOCND2()(f["sigqsat"], *[f[k] for k in names], timestep=50.0, zz=f["zz"],
                            ice_cld_wgt=f["ice_cld_wgt"])
    ad = dict(zip(ADJ_OUT + OCND2_X, adj))
    t = lambda k: torch.as_tensor(f[k], device=dev)  # noqa: E731
    state = dict(exn=t("exn"), dzz=t("dzz"), rhodj=t("rhodj"), rhodref=t("rho_dry_ref"), exnref=t("exn_ref"),
                 pres=t("pabs"), cf=ad["cldfr"], hlc_hrc=ad["hlc_hrc"], hlc_hcf=ad["hlc_hcf"],
                 hli_hri=ad["hli_hri"], hli_hcf=ad["hli_hcf"], sigs=t("sigs"),
                 th_t=ad["ths"] * 50.0, rv_t=ad["rvs"] * 50.0, rc_t=ad["rcs"] * 50.0, rr_t=t("rrs") * 50.0,
                 ri_t=ad["ris"] * 50.0, rs_t=t("rss") * 50.0, rg_t=t("rgs") * 50.0, ci_t=t("ci"),
                 ths=ad["ths"], rvs=ad["rvs"], rcs=ad["rcs"], rrs=t("rrs"), ris=ad["ris"], rss=t("rss"),
                 rgs=t("rgs"), icldfr=ad["icldfr"], ssio=ad["ssio"], ssiu=ad["ssiu"], ifr=ad["ifr"])
    new, diags = RainIceOCND2()(state, 50.0)
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "evap3d", "rainfr"):
        assert torch.equal(new[k], res[k]), k
    assert torch.equal(new["ci_t"], res["ci"])
    for k in ("inprr", "inpri", "inprs", "inprg"):
        assert torch.equal(diags[k], res[k]), k


def test_pipeline_ocnd2_full_size(dev, full512):
    """The fused OCND2 step at 512x512x90: column water budget, non-negative condensed
    sources, and every point against the oracle, flips counted (helpers.every_point)."""
    from helpers import every_point
    from ice3_amd import AdjustRainIceOCND2
    from oracle import oracle_lib, ref_lib
    f = _with_heights(dict(full512[0]), 20260050)
    res = AdjustRainIceOCND2()({**f, "rhodref": f["rho_dry_ref"]}, 50.0, return_calls=True)
    err = _column_budget(dev, f, res)
    assert err < 1e-12, err
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        assert bool((res[k] >= 0).all()), k
    w, nf, nb, checked = every_point(
        res, f, lambda g: ref_lib.pipeline_state_ocnd2(g, oracle_lib.ice_adjust_ocnd2(g)),
        lambda x, **kw: oracle_lib.rain_ice_ocnd2(x, 50.0, **kw), RAIN_OUT, RAIN_STATE, TOL,
        label="fused OCND2 step")
    print(f"fused OCND2 step 512x512x90, every point: budget {err:.1e}, {nf} flips, {nb} divergent "
          f"of {90 * 512 * 512}, worst {max(w.values()):.2e}; {checked} columns checked alone")


@pytest.mark.parametrize("kind", ["plain", "ocnd2", "hail"])
def test_persistent_and_point_kernels_bitwise(dev, kind):
    """The persistent microphysics kernel (64-point chunks, LDS parking) and the one-thread-
    per-point kernel (taken when a context sets a sub-step cap other than the default; 4095
    is never reached here) give the same results bit for bit, plain, OCND2 and with hail
    (KRR=7: rh = rg U(0.2, 2) and its source rh / dt)."""
    from ice3_amd import AdjustRainIce, AdjustRainIceHail, AdjustRainIceOCND2, _lib, synthetic
    lib, ctx = _lib.load(), _lib.context(0)
    f = _with_heights(synthetic.generate(1500, 90, seed=20260051, variant="pipeline"), 20260052)
    f["rhodref"] = f["rho_dry_ref"]
    if kind == "hail":
        rng = np.random.Generator(np.random.Philox(20260053))
        f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
        f["rhs"] = f["rh"] / 50.0
    step = {"plain": AdjustRainIce, "ocnd2": AdjustRainIceOCND2, "hail": AdjustRainIceHail}[kind]()
    a = step(f, 50.0, return_calls=True)
    assert lib.ice3_ctx_set_substep_cap(ctx, 4095) == _lib.ICE3_OK
    try:
        b = step(f, 50.0, return_calls=True)
    finally:
        assert lib.ice3_ctx_set_substep_cap(ctx, 4096) == _lib.ICE3_OK
    assert int((a["calls"] > 1).sum()) > 1000
    for k in a:
        assert torch.equal(torch.nan_to_num(a[k], nan=1.0), torch.nan_to_num(b[k], nan=1.0)), k
