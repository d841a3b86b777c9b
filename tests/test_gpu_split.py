"""The split sedimentation (PARAM_ICE CSEDIM='SPLI', k_sedim_split) on the GPU, selected on
the context (ice3_ctx_set_sedimentation) or on the drop-ins (sedimentation="SPLI"), against
the oracle's restatement (tests/test_oracle_split.py pins it):

* the sedimentation stage alone (the split scheme and its negativity correction) of the
  fused step;
* rain_ice, the fused step, hail and OCND2 with SPLI, the microphysics included, flips
  counted as in test_gpu_parity and test_gpu_hail;
* one to four slots per lane (nlev 40, 90, 150, 256) and partial workgroups of columns;
* the column water budget at 512x512x90, and the argument checks.

FP64 bar: relative error <= 1e-11 with the absolute floors of helpers.floor_for."""
import contextlib
import ctypes

import numpy as np
import pytest

from helpers import check_with_flips, explained_flips, floor_for, rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-11
DT = 50.0
RAIN_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr",
            "inprc", "inprr", "inpri", "inprs", "inprg")
RAIN_STATE = ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "ths", "rvs", "rcs", "rrs", "ris",
              "rss", "rgs")
HAIL_OUT = RAIN_OUT + ("rhs", "inprh")
HAIL_STATE = RAIN_STATE + ("rht", "rhs")
SEDIM_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "inprc", "inprr", "inpri", "inprs",
             "inprg")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _np(t):
    return t.detach().cpu().numpy()


@contextlib.contextmanager
def _oracle_spli():
    from oracle import oracle_lib
    oracle_lib.set_sedimentation("SPLI", 0.8)
    try:
        yield
    finally:
        oracle_lib.set_sedimentation("STAT")


def _spli(fn):
    """fn run with the oracle's sedimentation switched to SPLI."""
    def run(*a, **kw):
        with _oracle_spli():
            return fn(*a, **kw)
    return run


def _flips(out, s, names, fields, fn):
    from oracle import oracle_lib
    run = _spli(lambda x, **kw: fn(x, DT, **kw))
    ref = run(s, with_iters=True)
    ties = _spli(oracle_lib.rounding_ties)(fn, s, DT)
    return ref, explained_flips(out, ref, run, s, names, TOL, fields, ties=ties)


def _rain_state(s, hail=False):
    m = dict(exn="exn", dzz="dzz", rhodj="rhodj", rhodref="rhodref", exnref="exn_ref", pres="pabs",
             cf="cldfr", hlc_hrc="hlc_hrc", hlc_hcf="hlc_hcf", hli_hri="hli_hri", hli_hcf="hli_hcf",
             sigs="sigs", th_t="tht", rv_t="rvt", rc_t="rct", rr_t="rrt", ri_t="rit", rs_t="rst",
             rg_t="rgt", ci_t="ci", ths="ths", rvs="rvs", rcs="rcs", rrs="rrs", ris="ris",
             rss="rss", rgs="rgs")
    if hail:
        m.update(rh_t="rht", rhs="rhs")
    return {k: s[v] for k, v in m.items()}


def _rain_out(new, diags, hail=False):
    keys = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "evap3d", "rainfr") + (("rhs",) if hail else ())
    out = {k: _np(new[k]) for k in keys}
    out["ci"] = _np(new["ci_t"])
    out.update({k: _np(v) for k, v in diags.items()})
    return out


def _pipeline(ncol, seed, nlev=90):
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(ncol, nlev, seed=seed, variant="pipeline")
    s = ref_lib.pipeline_state(f, oracle_lib.ice_adjust(f))
    s["ci"] = f["ci"]
    return f, s


def test_sedimentation_stage_vs_oracle(dev):
    """The fused step's ADJUST then SEDIMENTATION stages with SPLI on the context: the
    sources after the split sedimentation and its negativity correction, and the surface
    rates, against the oracle's ICE4_SEDIMENTATION with SPLI on the oracle's adjusted state."""
    from ice3_amd import _lib
    from ice3_amd._fields import ptr, stream_ptr
    from oracle import oracle_lib
    from ice3_amd import synthetic
    from oracle import ref_lib
    ncol, nlev = 1000, 90
    # layers ten times thinner than the generator's (the stages read dzz only here) and more
    # pristine ice: rain, snow and graupel then take 2 to 26 sub-steps
    f = synthetic.generate(ncol, nlev, seed=20260311, variant="pipeline")
    f["dzz"] = f["dzz"] * 0.1
    f["ri"] = f["ri"] + 5e-4
    f["ris"] = f["ri"] / DT
    s = ref_lib.pipeline_state(f, oracle_lib.ice_adjust(f))
    ref = oracle_lib.sedimentation_spli(s, DT, maxcfl=0.8)
    for q in (0, 2, 3):
        assert (ref["nsub"][q] > 4).sum() > 100, q
    lib, ctx = _lib.load(), _lib.context(dev.index)
    g = {k: torch.as_tensor(v, device=dev) for k, v in f.items()}
    g["rhodref"] = g["rho_dry_ref"]
    out = {k: torch.zeros((nlev, ncol), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
    out.update({k: torch.zeros(ncol, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
    _lib.check(lib.ice3_ctx_set_sedimentation(ctx, 1, 0.8), "set SPLI")
    try:
        _lib.check(lib.ice3_adjust_rain_ice_stages(ctx, _lib.PREC_FP64, ncol, nlev, ncol, DT,
                                                   ctypes.byref(cin), ctypes.byref(cout), None,
                                                   _lib.STAGE_ADJUST | _lib.STAGE_SEDIMENTATION,
                                                   stream_ptr(dev)), "stages")
        torch.cuda.synchronize()
    finally:
        lib.ice3_ctx_set_sedimentation(ctx, 0, 0.8)
    worst = {}
    for k in SEDIM_OUT:
        r = rel_err(_np(out[k]), ref[k], floor_for(k, TOL))
        worst[k] = float(r.max())
    print("SPLI stage, 1000 columns:", {k: f"{v:.1e}" for k, v in worst.items()})
    assert max(worst.values()) <= TOL, worst


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_split_golden_inputs_vs_oracle(dev, golden_dir, tag):
    """RainIce(sedimentation="SPLI") on the compiled reference's rain_ice golden inputs against
    the oracle's rain_ice with SPLI, flips counted."""
    import os
    from ice3_amd import RainIce
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    p = "" if tag == "base" else tag + "_"
    s = {k[len(p) + 3:]: g[k] for k in g.files if k.startswith(p + "in_")}
    new, diags = RainIce(sedimentation="SPLI")(_rain_state(s), DT, return_calls=True)
    out = _rain_out(new, diags)
    ref, flips = _flips(out, s, RAIN_OUT, RAIN_STATE, oracle_lib.rain_ice)
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"rain_ice SPLI {tag}: {nf} flips, {nb} divergent of {flips.size}, worst {max(w.values()):.2e}")
    # the context is back on the statistical scheme afterwards
    new2, diags2 = RainIce()(_rain_state(s), DT)
    stat = oracle_lib.rain_ice(s, DT)
    assert rel_err(_np(diags2["inprr"]), stat["inprr"], floor_for("inprr", TOL)).max() <= TOL


@pytest.mark.parametrize("nlev,ncol", [(40, 37), (90, 100), (150, 33), (256, 21)])
def test_pipeline_split_levels_vs_oracle(dev, nlev, ncol):
    """AdjustRainIce(sedimentation="SPLI") with one to four level slots per lane and columns
    that leave the last workgroup partly empty, against the oracle; water budget."""
    from ice3_amd import AdjustRainIce
    from oracle import oracle_lib
    f, s = _pipeline(ncol, 20260320 + nlev, nlev)
    res = AdjustRainIce(sedimentation="SPLI")({**f, "rhodref": f["rho_dry_ref"]}, DT, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    ref, flips = _flips(out, s, RAIN_OUT, RAIN_STATE, oracle_lib.rain_ice)
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    mass = f["rho_dry_ref"] * f["dzz"]
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    col0 = (mass * sum(s[k] for k in ws)).sum(0) * DT
    col1 = (mass * sum(out[k] for k in ws)).sum(0) * DT
    pr = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"]) * DT
    budget = float(np.max(np.abs(col1 - col0 - pr) / col0))
    print(f"fused SPLI nlev={nlev} ncol={ncol}: {nf} flips, {nb} divergent, worst {max(w.values()):.2e}, "
          f"budget {budget:.1e}")
    assert budget < 1e-12


def test_pipeline_split_vs_oracle_and_mixed(dev):
    """The fused step with SPLI on 1,024 benchmark columns against the oracle (FP64), and the
    mixed path with SPLI: its FP64 split sedimentation and float microphysics stay within
    1e-4 of each field's maximum of the FP64 step."""
    from ice3_amd import AdjustRainIce
    from oracle import oracle_lib
    f, s = _pipeline(1024, 20260312)
    ari = AdjustRainIce(sedimentation="SPLI")
    res = ari({**f, "rhodref": f["rho_dry_ref"]}, DT, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    ref, flips = _flips(out, s, RAIN_OUT, RAIN_STATE, oracle_lib.rain_ice)
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"fused SPLI 1024 columns: {nf} flips, {nb} divergent of {flips.size}, worst {max(w.values()):.2e}")
    mixed = ari({**f, "rhodref": f["rho_dry_ref"]}, DT, precision="mixed")
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
        d = np.abs(_np(mixed[k]) - out[k]).max() / np.abs(out[k]).max()
        assert d < 1e-4, (k, d)


def test_hail_split_vs_oracle(dev):
    """RainIceHail and AdjustRainIceHail with SPLI (hail sediments in sub-steps too) against the
    oracle, flips counted; budget with hail; the fused and separate entry points agree."""
    from ice3_amd import AdjustRainIceHail
    from oracle import oracle_lib, ref_lib
    from ice3_amd import synthetic
    f = synthetic.generate(1024, 90, seed=20260313, variant="pipeline")
    rng = np.random.Generator(np.random.Philox(20260313))
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
    f["rhs"] = f["rh"] / DT
    f["rhodref"] = f["rho_dry_ref"]
    res = AdjustRainIceHail(sedimentation="SPLI")(f, DT, return_calls=True)
    out = {k: _np(v) for k, v in res.items()}
    s = ref_lib.pipeline_state_hail(f, oracle_lib.ice_adjust_hail(f))
    ref, flips = _flips(out, s, HAIL_OUT, HAIL_STATE, oracle_lib.rain_ice_hail)
    w, nf, nb = check_with_flips(out, ref, HAIL_OUT, flips, TOL)
    mass = f["rho_dry_ref"] * f["dzz"]
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")
    col0 = (mass * sum(s[k] for k in ws)).sum(0) * DT
    col1 = (mass * sum(out[k] for k in ws)).sum(0) * DT
    pr = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"] + out["inprh"]) * DT
    budget = float(np.max(np.abs(col1 - col0 - pr) / col0))
    print(f"fused hail SPLI: {nf} flips, {nb} divergent of {flips.size}, worst {max(w.values()):.2e}, "
          f"budget {budget:.1e}")
    assert budget < 1e-12
    assert int((out["inprh"] > 0).sum()) > 50


def test_ocnd2_split_vs_oracle(dev):
    """RainIceOCND2 with SPLI: the split kernel also writes the model level height the OCND2
    microphysics reads; against the oracle's rain_ice_ocnd2 with SPLI."""
    from ice3_amd import RainIceOCND2, synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(512, 90, seed=20260314, variant="pipeline")
    z1 = synthetic.level_profile(90)[2]
    f["zz"] = np.ascontiguousarray(np.broadcast_to(z1[:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.ones(512)
    a = oracle_lib.ice_adjust_ocnd2(f)
    s = ref_lib.pipeline_state_ocnd2(f, a)
    state = _rain_state(s)
    state.update({k: s[k] for k in ("icldfr", "ssio", "ssiu", "ifr")})
    new, diags = RainIceOCND2(sedimentation="SPLI")(state, DT, return_calls=True)
    out = _rain_out(new, diags)
    names = RAIN_OUT
    ref, flips = _flips(out, s, names, RAIN_STATE, oracle_lib.rain_ice_ocnd2)
    w, nf, nb = check_with_flips(out, ref, names, flips, TOL)
    print(f"rain_ice OCND2 SPLI: {nf} flips, {nb} divergent of {flips.size}, worst {max(w.values()):.2e}")


def test_split_full_size_budget(dev):
    """The fused step with SPLI at 512x512x90: column water budget, finite non-negative
    condensed sources, every point against the oracle with SPLI at the FP64 bar, flips
    counted (helpers.every_point; the oracle's scheme is set once for all its threads), and
    the statistical step is unchanged after it (the drop-in sets the scheme back)."""
    from helpers import every_point
    from ice3_amd import AdjustRainIce, synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(512 * 512, 90, seed=20260003, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    stat0 = AdjustRainIce()(f, DT)
    res = AdjustRainIce(sedimentation="SPLI")(f, DT, return_calls=True)
    with _oracle_spli():
        w, nf, nb, checked = every_point(
            res, f, lambda g: ref_lib.pipeline_state(g, oracle_lib.ice_adjust(g)),
            lambda x, **kw: oracle_lib.rain_ice(x, DT, **kw), RAIN_OUT, RAIN_STATE, TOL, dt=DT,
            label="fused SPLI step")
    print(f"fused SPLI 512x512x90, every point: {nf} flips, {nb} divergent of {90 * 512 * 512}, "
          f"worst {max(w.values()):.2e}; {checked} columns checked alone")
    res.pop("calls")
    t = lambda k: torch.as_tensor(f[k], device=dev)  # noqa: E731
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    mass = t("rho_dry_ref") * t("dzz")
    col0 = (mass * sum(t(k) for k in ws)).sum(0) * DT
    col1 = (mass * sum(res[k] for k in ws)).sum(0) * DT
    pr = -1000.0 * (res["inprr"] + res["inprs"] + res["inprg"]) * DT
    err = ((col1 - col0 - pr).abs() / col0).max().item()
    print(f"fused SPLI 512x512x90: water budget {err:.1e}")
    assert err < 1e-12
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        assert bool(torch.isfinite(res[k]).all()) and bool((res[k] >= 0).all()), k
    stat1 = AdjustRainIce()(f, DT)
    for k in stat0:
        assert torch.equal(torch.nan_to_num(stat0[k]), torch.nan_to_num(stat1[k])), k


def test_split_argument_checks(dev):
    """Bad schemes and CFL bounds are refused; with SPLI, more than 256 levels and the
    surrogate entry points are refused, and nothing runs."""
    from ice3_amd import _lib
    lib, ctx = _lib.load(), _lib.context(dev.index)
    assert lib.ice3_ctx_set_sedimentation(ctx, 2, 0.8) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_set_sedimentation(ctx, 1, 0.0) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_set_sedimentation(ctx, 1, float("nan")) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_set_sedimentation(None, 1, 0.8) == _lib.ICE3_EINVAL
    from ice3_amd import AdjustRainIce, synthetic
    f = synthetic.generate(8, 257, seed=1, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    with pytest.raises(_lib.Ice3Error):
        AdjustRainIce(sedimentation="SPLI")(f, DT)
    res = AdjustRainIce()(f, DT)  # the statistical scheme has no level limit
    assert bool(torch.isfinite(res["rrs"]).all())
    with pytest.raises(ValueError):
        AdjustRainIce(sedimentation="NONE")
    # with SPLI the microphysics stage runs in the call of the sedimentation stage (which keeps
    # its state at t); ADJUST in one call then SEDIMENTATION + MICROPHYSICS in the next gives
    # the one-call step bit for bit
    from ice3_amd._fields import ptr, stream_ptr
    f = synthetic.generate(100, 90, seed=2, variant="pipeline")
    g = {k: torch.as_tensor(v, device=dev) for k, v in f.items()}
    g["rhodref"] = g["rho_dry_ref"]

    def outs():
        o = {k: torch.zeros((90, 100), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
        o.update({k: torch.zeros(100, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
        return o, _lib.PipeOut(*[ptr(o[k]) for k in _lib.PIPE_OUT_FIELDS])
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    o1, c1 = outs()
    o2, c2 = outs()
    _lib.check(lib.ice3_ctx_set_sedimentation(ctx, 1, 0.8), "set SPLI")
    try:
        call = lambda c, st: lib.ice3_adjust_rain_ice_stages(  # noqa: E731
            ctx, _lib.PREC_FP64, 100, 90, 100, DT, ctypes.byref(cin), ctypes.byref(c), None, st, stream_ptr(dev))
        assert call(c1, _lib.STAGE_MICROPHYSICS) == _lib.ICE3_EINVAL
        assert call(c1, _lib.STAGE_ADJUST | _lib.STAGE_MICROPHYSICS) == _lib.ICE3_EINVAL
        _lib.check(call(c1, 7), "one call")
        _lib.check(call(c2, _lib.STAGE_ADJUST), "adjust")
        _lib.check(call(c2, _lib.STAGE_SEDIMENTATION | _lib.STAGE_MICROPHYSICS), "sedimentation + microphysics")
        torch.cuda.synchronize()
    finally:
        lib.ice3_ctx_set_sedimentation(ctx, 0, 0.8)
    for k in _lib.PIPE_OUT_FIELDS:
        assert torch.equal(o1[k], o2[k]), k
