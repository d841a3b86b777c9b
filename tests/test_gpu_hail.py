"""Hail (KRR=7, HCLOUD='ICE4') on the GPU, through the C ABI's hail entry points and the
IceAdjustHail / RainIceHail / AdjustRainIceHail drop-ins: against the compiled reference's
hail golden vectors, against the oracle on the benchmark generator with hail added (2,048
columns, and sampled columns of the 512x512x90 workload), the column water budget with
hail and surface hail, and the reduction to KRR=6 without hail and graupel.

FP64 bar as in test_gpu_parity: relative error <= 1e-11 with absolute floors; branch flips
are the points whose number of tendency calls or ice-concentration reset differs from the
oracle's, the points where the oracle is not determined to the bar by FP64 inputs, the
points where the oracle's graupel or hail wet/dry growth choice rests on rounding residuals,
and the points where the oracle reproduces the GPU's value from an input changed in its last
digits (helpers.explained_flips)."""
import os

import numpy as np
import pytest

from helpers import check_with_flips, explained_flips, floor_for, rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-11
DT = 50.0
ADJ_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
           "cph", "lv", "ls", "rvs", "rcs", "ris", "ths")
ADJ_IN = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
          "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
HAIL_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr", "inprc",
            "inprr", "inpri", "inprs", "inprg", "rhs", "inprh")
HAIL_STATE = ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "rht", "ths", "rvs", "rcs", "rrs",
              "ris", "rss", "rgs", "rhs")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_run(dt=DT):
    from oracle import oracle_lib
    return lambda x, **kw: oracle_lib.rain_ice_hail(x, dt, **kw)


def _ties(s, dt=DT):
    from oracle import oracle_lib
    return oracle_lib.rounding_ties(oracle_lib.rain_ice_hail, s, dt)


def _hail_state(s):
    """RainIceHail state-dict names from the oracle/reference field names."""
    m = dict(exn="exn", dzz="dzz", rhodj="rhodj", rhodref="rhodref", exnref="exn_ref", pres="pabs",
             cf="cldfr", hlc_hrc="hlc_hrc", hlc_hcf="hlc_hcf", hli_hri="hli_hri", hli_hcf="hli_hcf",
             sigs="sigs", th_t="tht", rv_t="rvt", rc_t="rct", rr_t="rrt", ri_t="rit", rs_t="rst",
             rg_t="rgt", rh_t="rht", ci_t="ci", ths="ths", rvs="rvs", rcs="rcs", rrs="rrs", ris="ris",
             rss="rss", rgs="rgs", rhs="rhs")
    return {k: s[v] for k, v in m.items()}


def _hail_pipeline(ncol, seed, hail=True, graupel=True):
    """The benchmark generator with a hail mixing ratio where it holds graupel (rh = rg times
    U(0.2, 2)), rhs = rh / dt as for the other species."""
    from ice3_amd import synthetic
    f = synthetic.generate(ncol, 90, seed=seed, variant="pipeline")
    rng = np.random.Generator(np.random.Philox(seed + 7))
    if not graupel:
        f["rg"] = np.zeros_like(f["rg"])
        f["rgs"] = np.zeros_like(f["rgs"])
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape)) if hail \
        else np.zeros_like(f["rg"])
    f["rhs"] = f["rh"] / DT
    f["rhodref"] = f["rho_dry_ref"]
    return f


def _budget(dev, f, res, dt=DT):
    """Max over columns of |change of the column's total water (hail included) - surface
    precipitation (hail included)| / column water."""
    t = lambda k: torch.as_tensor(f[k], device=dev)  # noqa: E731
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")
    mass = t("rho_dry_ref") * t("dzz")
    col0 = (mass * sum(t(k) for k in ws)).sum(0) * dt
    col1 = (mass * sum(res[k] for k in ws)).sum(0) * dt
    out = -1000.0 * (res["inprr"] + res["inprs"] + res["inprg"] + res["inprh"]) * dt
    return ((col1 - col0 - out).abs() / col0).max().item()


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_ice_adjust_hail_golden(dev, golden_dir, tag):
    """ICE_ADJUST with KRR=7 (IceAdjustHail) against the compiled reference's vectors."""
    from ice3_amd import IceAdjustHail
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    p = f"{tag}_adj_in_"
    f = {k[len(p):]: g[k] for k in g.files if k.startswith(p)}
    res = IceAdjustHail()(f["sigqsat"], *[f[k] for k in ADJ_IN], timestep=float(g["dt"]), rh=f["rh"])
    out = {k: _np(v) for k, v in zip(ADJ_OUT, res)}
    worst = {}
    for k in ADJ_OUT:
        ref = g[f"{tag}_adj_out_{k}"]
        assert np.array_equal(np.isnan(out[k]), np.isnan(ref)), k
        worst[k] = float(np.nanmax(rel_err(out[k], ref, floor_for(k, TOL))))
    assert max(worst.values()) <= TOL, worst
    print(f"ice_adjust hail golden {tag}: worst {max(worst.values()):.2e}")


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_hail_golden(dev, golden_dir, tag):
    """RAIN_ICE with KRR=7 (RainIceHail: hail sedimentation, ICE4_FAST_RH, graupel-to-hail
    conversion, hail totals and negativities) against the compiled reference's vectors."""
    from ice3_amd import RainIceHail
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    dt = float(g["dt"])
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in HAIL_OUT}
    new, diags = RainIceHail()(_hail_state(s), dt, return_calls=True)
    out = {k: _np(new[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs", "evap3d", "rainfr")}
    out["ci"] = _np(new["ci_t"])
    out.update({k: _np(v) for k, v in diags.items()})
    run = _oracle_run(dt)
    o = run(s, with_iters=True)
    flips = explained_flips(out, o, run, s, HAIL_OUT, TOL, HAIL_STATE, ties=_ties(s, dt))
    w, nf, nb = check_with_flips(out, ref, HAIL_OUT, flips, TOL)
    print(f"rain_ice hail golden {tag}: {nf} flips, {nb} divergent of {flips.size}, worst {max(w.values()):.2e}")


def test_pipeline_hail_vs_oracle(dev):
    """The fused step with hail on the benchmark generator with hail added (2,048 columns)
    against the oracle, flips counted; the water budget with hail holds; the fused entry point
    and the two separate ones (IceAdjustHail then RainIceHail) agree bit for bit."""
    from ice3_amd import AdjustRainIceHail, IceAdjustHail, RainIceHail
    from oracle import oracle_lib, ref_lib
    f = _hail_pipeline(2048, 20260081)
    res = AdjustRainIceHail()(f, DT, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    a = oracle_lib.ice_adjust_hail(f)
    s = ref_lib.pipeline_state_hail(f, a)
    run = _oracle_run()
    ref = run(s, with_iters=True)
    flips = explained_flips(out, ref, run, s, HAIL_OUT, TOL, HAIL_STATE, ties=_ties(s))
    w, nf, nb = check_with_flips(out, ref, HAIL_OUT, flips, TOL)
    for k in ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf"):
        assert rel_err(out[k], a[k], floor_for(k, TOL)).max() <= TOL, k
    err = _budget(dev, f, res)
    print(f"fused hail step, 2048 columns: {nf} flips, {nb} divergent of {flips.size}, "
          f"worst {max(w.values()):.2e}, water budget {err:.1e}")
    assert err < 1e-12
    assert int((out["inprh"] > 0).sum()) > 100
    # the same step through the two separate hail entry points
    adj = IceAdjustHail()(f["sigqsat"], *[f[k] for k in ADJ_IN], timestep=DT, rh=f["rh"])
    ad = dict(zip(ADJ_OUT, adj))
    t = lambda k: torch.as_tensor(f[k], device=dev)  # noqa: E731
    state = dict(exn=t("exn"), dzz=t("dzz"), rhodj=t("rhodj"), rhodref=t("rho_dry_ref"), exnref=t("exn_ref"),
                 pres=t("pabs"), cf=ad["cldfr"], hlc_hrc=ad["hlc_hrc"], hlc_hcf=ad["hlc_hcf"],
                 hli_hri=ad["hli_hri"], hli_hcf=ad["hli_hcf"], sigs=t("sigs"),
                 th_t=ad["ths"] * DT, rv_t=ad["rvs"] * DT, rc_t=ad["rcs"] * DT, rr_t=t("rrs") * DT,
                 ri_t=ad["ris"] * DT, rs_t=t("rss") * DT, rg_t=t("rgs") * DT, rh_t=t("rhs") * DT,
                 ci_t=t("ci"), ths=ad["ths"], rvs=ad["rvs"], rcs=ad["rcs"], rrs=t("rrs"), ris=ad["ris"],
                 rss=t("rss"), rgs=t("rgs"), rhs=t("rhs"))
    new, diags = RainIceHail()(state, DT)
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs", "evap3d", "rainfr"):
        assert torch.equal(new[k], res[k]), k
    assert torch.equal(new["ci_t"], res["ci"])
    for k in ("inprr", "inpri", "inprs", "inprg", "inprh"):
        assert torch.equal(diags[k], res[k]), k


# A stride coprime to 64 over the 512x512 columns: 516 sampled columns, all lanes of a
# wavefront.
def test_pipeline_hail_full_size(dev):
    """The fused hail step at 512x512x90 (the benchmark workload with hail added): column
    water budget with hail, finite non-negative condensed sources, and every point against
    the oracle at the FP64 bar, flips counted (helpers.every_point)."""
    from helpers import every_point
    from ice3_amd import AdjustRainIceHail
    from oracle import oracle_lib, ref_lib
    f = _hail_pipeline(512 * 512, 20260003)
    res = AdjustRainIceHail()(f, DT, return_calls=True)
    err = _budget(dev, f, res)
    assert err < 1e-12, err
    for k in ("rcs", "rrs", "ris", "rss", "rgs", "rhs"):
        assert bool(torch.isfinite(res[k]).all()) and bool((res[k] >= 0).all()), k
    w, nf, nb, checked = every_point(
        res, f, lambda g: ref_lib.pipeline_state_hail(g, oracle_lib.ice_adjust_hail(g)),
        _oracle_run(), HAIL_OUT, HAIL_STATE, TOL, dt=DT, label="fused hail step")
    print(f"fused hail step 512x512x90, every point: budget {err:.1e}, {nf} flips, {nb} divergent "
          f"of {90 * 512 * 512}, worst {max(w.values()):.2e}; {checked} columns checked alone")


def test_hail_without_hail_or_graupel_is_krr6(dev):
    """Without hail and graupel nothing feeds the hail processes: the KRR=7 fused step gives
    the KRR=6 step's results bit for bit (its point kernel against the persistent one), with
    the hail source and surface rate at 0."""
    from ice3_amd import AdjustRainIce, AdjustRainIceHail
    f = _hail_pipeline(1024, 20260083, hail=False, graupel=False)
    a = AdjustRainIce()(f, DT, return_calls=True)
    b = AdjustRainIceHail()(f, DT, return_calls=True)
    for k in a:
        assert torch.equal(torch.nan_to_num(a[k], nan=1.0), torch.nan_to_num(b[k], nan=1.0)), k
    assert bool((b["rhs"] == 0).all()) and bool((b["inprh"] == 0).all())
