"""The oracle (C restatement) against the compiled reference's golden vectors, the
mathematics the path must satisfy, and (in the build container) the reference itself."""
import os

import numpy as np
import pytest

from helpers import floor_for, rel_err
from ice3_amd import synthetic
from oracle import oracle_lib, ref_lib

ADJ_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
           "cph", "lv", "ls", "rvs", "rcs", "ris", "ths", "srcs")
RAIN_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr",
            "inprc", "inprr", "inpri", "inprs", "inprg")


def _adjust_case(g, tag):
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    return f, {k: g[f"{tag}_out_{k}"] for k in ADJ_OUT}


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_ice_adjust_matches_reference_golden(golden_dir, tag):
    g = np.load(os.path.join(golden_dir, "ice_adjust_fp64.npz"))
    f, ref = _adjust_case(g, tag)
    o = oracle_lib.ice_adjust(f, float(g["dt"]))
    for k in ADJ_OUT:
        r = rel_err(o[k], ref[k], floor_for(k, 1e-14))
        assert np.all(np.isnan(o[k]) == np.isnan(ref[k])), k
        assert r.max() <= 1e-14, (k, r.max())


def _rain_case(g, tag):
    p = "" if tag == "base" else tag + "_"
    s = {k[len(p) + 3:]: g[k] for k in g.files if k.startswith(p + "in_")}
    return s, {k: g[p + "out_" + k] for k in RAIN_OUT}


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_matches_reference_golden(golden_dir, tag):
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    s, ref = _rain_case(g, tag)
    o = oracle_lib.rain_ice(s, float(g["dt"]))
    assert o["status"] == 0
    for k in RAIN_OUT:
        r = rel_err(o[k], ref[k], floor_for(k, 1e-12))
        assert r.max() <= 1e-12, (k, r.max())


def test_golden_fixtures_exercise_the_processes(golden_dir):
    """The rain_ice fixtures must cover precipitating, mixed-phase, melting and
    freezing points, and multi-call time splitting (so that parity is meaningful)."""
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    s = {}
    for tag in ("base", "cov"):
        si, _ = _rain_case(g, tag)
        for k, v in si.items():
            s[k] = np.concatenate([s[k], v], axis=1) if k in s else v
    o = oracle_lib.rain_ice(s, float(g["dt"]), with_iters=True)
    t = s["tht"] * s["exn"]
    assert (o["iters"] >= 2).sum() > 100
    assert ((s["rrt"] > 1e-4) & (t < 273.16)).sum() > 20       # supercooled rain
    assert ((s["rgt"] > 1e-4) & (t > 273.16)).sum() > 8        # melting graupel
    assert ((s["rst"] > 1e-4) & (s["rct"] > 1e-6)).sum() > 5   # riming
    assert (t < 273.16 - 35).sum() > 100                       # homogeneous freezing range
    assert (o["inprr"] > 0).sum() > 3 and (o["inprg"] > 0).sum() > 3


def test_ice_adjust_conserves_water_and_energy():
    """Sources: d(rvs+rcs+ris) = 0 and the theta increment is the latent heat release
    (micro/ice_adjust.F90:313-345) -- a closed form independent of the restatement."""
    f = synthetic.generate(64, 90, seed=7)
    o = oracle_lib.ice_adjust(f)
    tot0 = f["rvs"] + f["rcs"] + f["ris"]
    tot1 = o["rvs"] + o["rcs"] + o["ris"]
    assert np.max(np.abs(tot1 - tot0) / np.abs(tot0)) < 1e-14
    dth = (o["rcs"] - f["rcs"]) * o["lv"] / (o["cph"] * f["exn_ref"]) + \
          (o["ris"] - f["ris"]) * o["ls"] / (o["cph"] * f["exn_ref"])
    assert np.allclose(o["ths"] - f["ths"], dth, rtol=1e-9, atol=1e-12)
    assert np.all((o["cldfr"] >= 0) & (o["cldfr"] <= 1))


def test_ice_adjust_closed_forms():
    """CB02 cases and esat: one point in each Q1 regime, recomputed from the formulas."""
    f = synthetic.generate(16, 90, seed=11)
    o = oracle_lib.ice_adjust(f)
    c = {}
    for line in open(oracle_lib.CONSTANTS):
        p = line.split()
        c[p[0]] = float(p[2])
    t = f["th"] * f["exn"]
    esw = np.exp(c["XALPW"] - c["XBETAW"] / t - c["XGAMW"] * np.log(t))
    assert np.all(esw > 0)
    # warm, dry point: no condensate -> cldfr 0 implies rc_out == 0 and ri_out == 0
    dry = o["cldfr"] == 0
    assert dry.sum() > 10
    assert np.all(o["rc_out"][dry] == 0) and np.all(o["ri_out"][dry] == 0)
    # lv/ls closed forms
    assert np.allclose(o["lv"], c["XLVTT"] + (c["XCPV"] - c["XCL"]) * (t - c["XTT"]), rtol=1e-15)
    assert np.allclose(o["ls"], c["XLSTT"] + (c["XCPV"] - c["XCI"]) * (t - c["XTT"]), rtol=1e-15)


def _pipeline(ncol, seed):
    f = synthetic.generate(ncol, 90, seed=seed, variant="pipeline")
    a = oracle_lib.ice_adjust(f)
    s = ref_lib.pipeline_state(f, {**a, "ths": a["ths"], "rvs": a["rvs"], "rcs": a["rcs"], "ris": a["ris"]})
    s["ci"] = f["ci"]
    return f, s


def test_ice_adjust_branch_record(golden_dir):
    """ICE_ADJUST's branch record (oracle_record_adjust_branches) against what the compiled
    reference's golden outputs show of each branch: no cloud where the cloud fraction is that
    of the mass flux alone; the adjusted-state limiter where it divided 0/0 (rv_out NaN with
    no mass-flux condensate); a droplet or ice pdf case 1 exactly where its cloud fraction is
    zero, case 0 where it is at least half the mass-flux fraction, case 2 below; the sign of
    ZW1 / ZW2 from the change of rcs / ris without mass-flux condensate; the Q1 interval
    against INQ1.  Recording leaves the outputs bitwise unchanged."""
    g = np.load(os.path.join(golden_dir, "ice_adjust_fp64.npz"))
    seen = {}
    for tag in ("base", "mf"):
        f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
        ref = {k[len(tag) + 5:]: g[k] for k in g.files if k.startswith(tag + "_out_")}
        dt = float(g["dt"])
        o, flags = oracle_lib.record_adjust_branches(oracle_lib.ice_adjust, f, dt)
        plain = oracle_lib.ice_adjust(f, dt)
        for k in plain:
            assert np.array_equal(o[k], plain[k], equal_nan=True), k
        b = oracle_lib.adjust_branch_fields(flags)
        for n, v in b.items():
            seen.setdefault(n, set()).update(np.unique(v).tolist())
        cf = f["cf_mf"]
        assert np.array_equal(b["nocloud"] == 1, ref["cldfr"] == np.minimum(1.0, cf))
        nomf = (f["rc_mf"] == 0) & (f["ri_mf"] == 0)
        assert np.array_equal((b["mfout"] == 1) & nomf, np.isnan(ref["rv_out"]) & nomf)
        for pdf, hcf in (("pdfc", ref["hlc_hcf"]), ("pdfi", ref["hli_hcf"])):
            on = cf > 0
            assert np.array_equal((b[pdf] == 1) & on, (hcf == 0) & on), pdf
            assert np.all(hcf[(b[pdf] == 0) & on] >= 0.5 * cf[(b[pdf] == 0) & on] * (1 - 1e-12)), pdf
            assert np.all(hcf[(b[pdf] == 2) & on] <= 0.5 * cf[(b[pdf] == 2) & on] * (1 + 1e-12)), pdf
        # without mass-flux condensate the sources change by the limited ZW1 / ZW2 alone
        dc, di = ref["rcs"] - f["rcs"], ref["ris"] - f["ris"]
        assert np.all(dc[(b["zw1neg"] == 1) & nomf] <= 0) and np.all(dc[(b["zw1neg"] == 0) & nomf] >= 0)
        assert np.all(di[(b["zw2neg"] == 1) & nomf] <= 0) and np.all(di[(b["zw2neg"] == 0) & nomf] >= 0)
        assert np.any(dc[(b["zw1neg"] == 1) & nomf] < 0) and np.any(di[(b["zw2neg"] == 1) & nomf] < 0)
        # Q1 <= 0: INQ1 <= 0; 0 < Q1 <= 2: 0 <= INQ1 <= 4; Q1 > 2: INQ1 >= 4 (INQ1 + 22 stored)
        q, i = b["q1"], b["inq1"]
        assert np.all(i[q == 0] <= 22) and np.all((i[q == 1] >= 22) & (i[q == 1] <= 26))
        assert np.all(i[q == 2] >= 26) and np.all(i <= 32) and not np.any(q == 3)
    assert seen["q1"] == {0, 1, 2} and seen["pdfc"] == {0, 1, 2} and seen["pdfi"] == {0, 1, 2}
    for n in ("nocloud", "zw1neg", "zw2neg", "mflimit", "mfout"):
        assert seen[n] == {0, 1}, n


def test_rain_ice_column_water_budget():
    """Total water of each column changes only by what falls out of its lowest level:
    sum_k rho dz (sum of water sources)(after-before) dt = -rho_lw (inprr+inprs+inprg) dt."""
    f, s = _pipeline(48, 21)
    o = oracle_lib.rain_ice(s)
    dt = 50.0
    w = lambda d: sum(d[k] for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs"))  # noqa: E731
    col = lambda x: np.sum(s["rhodref"] * s["dzz"] * x, axis=0) * dt  # noqa: E731
    dw = col(w(o)) - col(w(s))
    out = -1000.0 * (o["inprr"] + o["inprs"] + o["inprg"]) * dt
    scale = col(w(s))
    assert np.max(np.abs(dw - out) / scale) < 1e-12


def test_rain_ice_identity_without_hydrometeors_above_freezing():
    """No condensate and T > TT: no process is active, so rain_ice leaves the sources alone."""
    f, s = _pipeline(4, 5)
    for k in ("rct", "rrt", "rit", "rst", "rgt"):
        s[k] = np.zeros_like(s[k])
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        s[k] = np.zeros_like(s[k])
    s["tht"] = np.maximum(s["tht"], 290.0 / s["exn"])
    s["ths"] = s["tht"] / 50.0
    o = oracle_lib.rain_ice(s)
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
        assert np.array_equal(o[k], s[k]), k
    assert np.all(o["inprr"] == 0)


def test_rain_ice_full_freezing_below_minus_35():
    """T < TT-35 with rain: RRHONG freezes all of it within the step (ICE4_RRHONG)."""
    f, s = _pipeline(2, 9)
    s["tht"] = np.full_like(s["tht"], 230.0) / s["exn"]
    s["ths"] = s["tht"] / 50.0
    s["rrt"] = np.full_like(s["rrt"], 1e-3)
    s["rrs"] = s["rrt"] / 50.0
    s["dzz"] = s["dzz"] * 1e6  # no sedimentation in or out
    o = oracle_lib.rain_ice(s)
    assert np.all(np.abs(o["rrs"]) < 1e-12)
    assert np.all(o["rgs"] > s["rgs"])


@pytest.mark.skipif(not ref_lib.available(), reason="compiled reference not built here")
@pytest.mark.parametrize("seed", [101, 202])
def test_oracle_matches_compiled_reference_fresh_seeds(seed):
    f = synthetic.generate(24, 90, seed=seed, variant="pipeline")
    a_ref = ref_lib.ice_adjust(f)
    a = oracle_lib.ice_adjust(f)
    for k in ("rvs", "rcs", "ris", "ths", "cldfr", "srcs"):
        assert rel_err(a[k], a_ref[k], floor_for(k, 1e-14)).max() <= 1e-14, k
    s = ref_lib.pipeline_state(f, a_ref)
    r_ref = ref_lib.rain_ice(s, nproma=1)
    r = oracle_lib.rain_ice(s)
    for k in RAIN_OUT:
        assert rel_err(r[k], r_ref[k], floor_for(k, 1e-12)).max() <= 1e-12, k


# ------------------------------------------------------------------------- OCND2 branch

OCND2_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
             "rvs", "rcs", "ris", "ths", "srcs", "icldfr", "wcldfr", "ssio", "ssiu", "ifr")


def _ocnd2_case(g, tag):
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    return f, {k: g[f"{tag}_out_{k}"] for k in OCND2_OUT}


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_ice_adjust_ocnd2_matches_reference_golden(golden_dir, tag):
    """ICE_ADJUST with LOCND2 (ICECLOUD, the TIWMX tables, the liquid-only condensate)
    against the compiled reference's vectors."""
    g = np.load(os.path.join(golden_dir, "ice_adjust_ocnd2_fp64.npz"))
    f, ref = _ocnd2_case(g, tag)
    o = oracle_lib.ice_adjust_ocnd2(f, float(g["dt"]))
    for k in OCND2_OUT:
        assert np.all(np.isnan(o[k]) == np.isnan(ref[k])), k
        r = rel_err(o[k], ref[k], floor_for(k, 1e-14))
        assert np.nanmax(np.where(np.isnan(ref[k]), 0.0, r)) <= 1e-14, (k, r.max())


def test_ice_adjust_ocnd2_invariants():
    """What the OCND2 branch fixes whatever its tables: the condensation leaves the ice
    alone (ri_out = ri and, without mass flux, the ice source unchanged), total water is
    conserved point by point, PIFR is its initial 10, PWCLDFR is the statistical cloud
    fraction, and the fractions stay in [0, 1]."""
    f = synthetic.generate(48, 90, seed=20260035, variant="adjust")
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.full(48, 1.0)
    o = oracle_lib.ice_adjust_ocnd2(f)
    assert np.array_equal(o["ris"], f["ris"])
    # the adjusted state POUT_* is NaN where its mass-flux limiter divides 0/0
    # (ice_adjust.F90:449-466 with rc_mf = ri_mf = 0 and a negative adjusted vapour)
    ok = np.isfinite(o["ri_out"])
    assert ok.mean() > 0.8
    assert np.array_equal(o["ri_out"][ok], f["ri"][ok])
    water = f["rv"] + f["rc"]
    assert np.all(np.abs(o["rv_out"] + o["rc_out"] - water)[ok] <= 4e-16 * water[ok])
    assert np.all(o["ifr"] == 10.0)
    for k in ("cldfr", "icldfr", "wcldfr"):
        assert np.all((o[k] >= 0.0) & (o[k] <= 1.0)), k
    assert np.all(o["ssiu"] >= -1.0) and np.all(o["ssiu"] <= 0.0)


@pytest.mark.skipif(not ref_lib.available(), reason="compiled reference not built here")
def test_ocnd2_oracle_matches_compiled_reference_fresh_seed():
    f = synthetic.generate(40, 90, seed=303, variant="adjust")
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.linspace(0.2, 2.0, 40)
    r = ref_lib.ice_adjust_ocnd2(f)
    o = oracle_lib.ice_adjust_ocnd2(f)
    for k in ("rvs", "rcs", "ris", "ths", "cldfr", "srcs", "icldfr", "wcldfr", "ssio", "ssiu", "ifr"):
        assert rel_err(o[k], r[k], floor_for(k, 1e-14)).max() <= 1e-14, k


# ------------------------------------------------------------- RAIN_ICE with OCND2

def _rain_ocnd2_case(g, tag):
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    return s, {k: g[f"{tag}_out_{k}"] for k in RAIN_OUT}


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_ocnd2_matches_reference_golden(golden_dir, tag):
    """RAIN_ICE with LOCND2 (mask on PSSIO, tabulated nucleation, ZESI/ZAI from the TIWMX
    tables, OCND2 collision factors, ICE4_FAST_RI_RS) against the compiled reference."""
    g = np.load(os.path.join(golden_dir, "rain_ice_ocnd2_fp64.npz"))
    s, ref = _rain_ocnd2_case(g, tag)
    o = oracle_lib.rain_ice_ocnd2(s, float(g["dt"]))
    assert o["status"] == 0
    for k in RAIN_OUT:
        r = rel_err(o[k], ref[k], floor_for(k, 1e-12))
        assert r.max() <= 1e-12, (k, r.max())


def test_rain_ice_ocnd2_fixtures_exercise_the_branch(golden_dir):
    """The OCND2 fixtures reach what the branch adds: ice-supersaturated points inside the ice
    cloud and subsaturated ones outside it with pristine ice (ICE4_FAST_RI_RS), nucleation
    below 0 C outside the microphysics mask, mask points kept only by PSSIO, multi-call time
    splitting; and the branch changes the results against the plain RAIN_ICE."""
    g = np.load(os.path.join(golden_dir, "rain_ice_ocnd2_fp64.npz"))
    for tag in ("base", "cov"):
        s, ref = _rain_ocnd2_case(g, tag)
        o = oracle_lib.rain_ice_ocnd2(s, float(g["dt"]), with_iters=True)
        t = s["tht"] * s["exn"]
        ri = s["rit"] > 1e-10
        assert ((s["ssio"] > 0) & (s["icldfr"] > 0.02) & ri & (s["ci"] > 0)).sum() > 50, tag
        assert ((s["ssiu"] < 0) & (s["icldfr"] < 0.98) & ri & (s["ci"] > 0)).sum() > 20, tag
        hyd = np.stack([s[k] for k in ("rct", "rrt", "rit", "rst", "rgt")])
        only_ssio = (s["ssio"] > 0) & ~(hyd > 1e-15).any(0)
        assert only_ssio.sum() > 0 or tag == "cov", tag
        assert ((o["iters"] == 0) & (t < 273.16)).sum() > 0 or tag == "cov", tag
        assert (o["iters"] >= 2).sum() > 100, tag
        plain = oracle_lib.rain_ice(s, float(g["dt"]))
        assert np.abs(plain["ris"] - ref["ris"]).max() > 1e-3 * np.abs(ref["ris"]).max(), tag


def test_rain_ice_ocnd2_column_water_budget():
    """The OCND2 processes move water between species (vapour -> ice in RVDEPI, ice -> snow in
    RILARS, nucleation takes no vapour): each column's total water still changes only by its
    surface precipitation, and every condensed source ends non-negative."""
    f = synthetic.generate(64, 90, seed=20260046, variant="pipeline")
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.full(64, 1.0)
    a = oracle_lib.ice_adjust_ocnd2(f)
    s = ref_lib.pipeline_state_ocnd2(f, a)
    o = oracle_lib.rain_ice_ocnd2(s)
    dt = 50.0
    mass = f["rho_dry_ref"] * f["dzz"]
    w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    col0 = (mass * sum(s[k] for k in w)).sum(0) * dt
    col1 = (mass * sum(o[k] for k in w)).sum(0) * dt
    pr = -1000.0 * (o["inprr"] + o["inprs"] + o["inprg"]) * dt
    assert np.max(np.abs(col1 - col0 - pr) / col0) < 1e-13
    for k in ("rcs", "rrs", "ris", "rss", "rgs"):
        assert np.all(o[k] >= 0.0), k


@pytest.mark.skipif(not ref_lib.available(), reason="compiled reference not built here")
def test_rain_ice_ocnd2_oracle_matches_compiled_reference_fresh_seed():
    f = synthetic.generate(40, 90, seed=404, variant="pipeline")
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.linspace(0.2, 2.0, 40)
    s = ref_lib.pipeline_state_ocnd2(f, ref_lib.ice_adjust_ocnd2(f))
    r = ref_lib.rain_ice_ocnd2(s, nproma=1)
    o = oracle_lib.rain_ice_ocnd2(s)
    for k in RAIN_OUT:
        assert rel_err(o[k], r[k], floor_for(k, 1e-12)).max() <= 1e-12, k


# ------------------------------------------------------------- hail (KRR=7)

RAIN_HAIL_OUT = RAIN_OUT + ("rhs", "inprh")


def _hail_case(g, tag):
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    return s, {k: g[f"{tag}_out_{k}"] for k in RAIN_HAIL_OUT}


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_ice_adjust_hail_matches_reference_golden(golden_dir, tag):
    """ICE_ADJUST with KRR=7: the hail mixing ratio in the heat capacity."""
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    p = f"{tag}_adj_in_"
    f = {k[len(p):]: g[k] for k in g.files if k.startswith(p)}
    o = oracle_lib.ice_adjust_hail(f, float(g["dt"]))
    for k in ADJ_OUT:
        ref = g[f"{tag}_adj_out_{k}"]
        assert np.all(np.isnan(o[k]) == np.isnan(ref)), k
        assert rel_err(o[k], ref, floor_for(k, 1e-14)).max() <= 1e-14, k


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_rain_ice_hail_matches_reference_golden(golden_dir, tag):
    """RAIN_ICE with KRR=7 (ICE4_FAST_RH, the SWETH/GWETH/RWETH kernels, graupel-to-hail
    conversion, hail sedimentation and negativity correction) against the compiled reference."""
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    s, ref = _hail_case(g, tag)
    o = oracle_lib.rain_ice_hail(s, float(g["dt"]))
    assert o["status"] == 0
    for k in RAIN_HAIL_OUT:
        r = rel_err(o[k], ref[k], floor_for(k, 1e-12))
        assert r.max() <= 1e-12, (k, r.max())


def test_rain_ice_hail_fixtures_exercise_the_branch(golden_dir):
    """The hail fixtures take every branch of ICE4_FAST_RH (wet growth, dry growth, conversion
    of hail to graupel, melting) and the graupel-to-hail conversion of ICE4_FAST_RG, and hail
    changes the results against KRR=6."""
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    total, change = {}, 0.0
    for tag in ("base", "cov"):
        s, ref = _hail_case(g, tag)
        o, fl = oracle_lib.record_branches(oracle_lib.rain_ice_hail, s, float(g["dt"]))
        for b, n in oracle_lib.branch_counts(fl).items():
            total[b] = total.get(b, 0) + n
        plain = oracle_lib.rain_ice(s, float(g["dt"]))
        change = max(change, np.abs(plain["rgs"] - ref["rgs"]).max() / np.abs(ref["rgs"]).max())
    assert change > 1e-2, change
    for b in ("weth", "dryh", "dryhg", "hmltr", "wetgh", "wetg", "dryg"):
        assert total[b] >= 5, (b, total)


def _hail_pipeline(ncol, seed, hail=True, graupel=True):
    f = synthetic.generate(ncol, 90, seed=seed, variant="pipeline")
    rng = np.random.Generator(np.random.Philox(seed + 7))
    if not graupel:
        f["rg"] = np.zeros_like(f["rg"])
        f["rgs"] = np.zeros_like(f["rgs"])
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape)) if hail else np.zeros_like(f["rg"])
    f["rhs"] = f["rh"] / 50.0
    return f


def test_rain_ice_hail_column_water_budget():
    """With hail, each column's total water (vapour, the five condensed species and hail)
    changes only by its surface precipitation, hail included, and every condensed source ends
    non-negative."""
    f = _hail_pipeline(64, 20260075)
    a = oracle_lib.ice_adjust_hail(f)
    s = ref_lib.pipeline_state_hail(f, a)
    o = oracle_lib.rain_ice_hail(s)
    dt = 50.0
    mass = f["rho_dry_ref"] * f["dzz"]
    w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")
    col0 = (mass * sum(s[k] for k in w)).sum(0) * dt
    col1 = (mass * sum(o[k] for k in w)).sum(0) * dt
    pr = -1000.0 * (o["inprr"] + o["inprs"] + o["inprg"] + o["inprh"]) * dt
    assert np.max(np.abs(col1 - col0 - pr) / col0) < 1e-13
    assert (o["inprh"] > 0).sum() > 10
    for k in ("rcs", "rrs", "ris", "rss", "rgs", "rhs"):
        assert np.all(o[k] >= 0.0), k


def test_rain_ice_hail_without_hail_or_graupel_is_krr6():
    """Without hail and graupel nothing feeds the hail processes (no graupel wet growth to
    convert, no hail to grow, melt or sediment): KRR=7 gives KRR=6's results bit for bit and
    leaves the hail source and surface rate at 0."""
    f = _hail_pipeline(48, 20260077, hail=False, graupel=False)
    a6 = oracle_lib.ice_adjust(f)
    a7 = oracle_lib.ice_adjust_hail(f)
    for k in ADJ_OUT:
        assert np.array_equal(a6[k], a7[k], equal_nan=True), k
    s = ref_lib.pipeline_state_hail(f, a7)
    o6 = oracle_lib.rain_ice(s)
    o7 = oracle_lib.rain_ice_hail(s)
    for k in RAIN_OUT:
        assert np.array_equal(o6[k], o7[k]), k
    assert np.all(o7["rhs"] == 0.0) and np.all(o7["inprh"] == 0.0)


@pytest.mark.skipif(not ref_lib.available(), reason="compiled reference not built here")
def test_rain_ice_hail_oracle_matches_compiled_reference_fresh_seed():
    f = _hail_pipeline(40, 505)
    s = ref_lib.pipeline_state_hail(f, ref_lib.ice_adjust_hail(f))
    r = ref_lib.rain_ice_hail(s, nproma=1)
    o = oracle_lib.rain_ice_hail(s)
    for k in RAIN_HAIL_OUT:
        assert rel_err(o[k], r[k], floor_for(k, 1e-12)).max() <= 1e-12, k
