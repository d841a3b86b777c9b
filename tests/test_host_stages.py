"""The kernels' own stage source (dwarf-p-ice3_amd/csrc/ice3_stages.hpp), compiled for
the CPU by g++ (tests/host_stages), against the oracle, the reference's golden
vectors and under AddressSanitizer / UndefinedBehaviorSanitizer.

This checks the arithmetic and indexing of the device code without a GPU; the GPU
parity tests then check the hipcc build of the same source."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import check_with_flips, floor_for, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "host_stages"))
import host_lib  # noqa: E402

TOL = 1e-11
RAIN_OUT = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr",
            "inprc", "inprr", "inpri", "inprs", "inprg")
ADJ_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
           "cph", "lv", "ls", "rvs", "rcs", "ris", "ths")


def _worst(out, ref, names, flips=None, tol=TOL):
    worst = {}
    for k in names:
        r = rel_err(out[k], ref[k], floor_for(k, tol))
        if flips is not None and r.shape == flips.shape:
            r = np.where(flips, 0.0, r)
        elif flips is not None and r.ndim == 1:
            r = np.where(flips.any(axis=0), 0.0, r)
        worst[k] = float(np.nanmax(r))
    return worst


def _pipeline_input(ncol, seed=20260003):
    from ice3_amd import synthetic
    f = synthetic.generate(ncol, 90, seed=seed, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    return f


OCND2_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
             "rvs", "rcs", "ris", "ths", "srcs", "icldfr", "wcldfr", "ssio", "ssiu", "ifr")


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_host_ice_adjust_ocnd2_golden(golden_dir, tag):
    """The kernels' OCND2 ICE_ADJUST source on the CPU against the compiled reference's vectors
    (its TIWMX saturation pressures come from the kernels' FP64 exp/log, not the library's)."""
    g = np.load(os.path.join(golden_dir, "ice_adjust_ocnd2_fp64.npz"))
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in OCND2_OUT}
    out = host_lib.ice_adjust_ocnd2(f, float(g["dt"]))
    for k in OCND2_OUT:
        assert np.array_equal(np.isnan(out[k]), np.isnan(ref[k])), k
    w = _worst(out, ref, OCND2_OUT)
    assert max(w.values()) <= TOL, w


def test_host_pipeline_matches_oracle():
    from oracle import oracle_lib, ref_lib
    f = _pipeline_input(128)
    out = host_lib.adjust_rain_ice(f, 50.0, 0)
    a = oracle_lib.ice_adjust(f)
    ref = oracle_lib.rain_ice(ref_lib.pipeline_state(f, a), with_iters=True)
    flips = out["calls"] != ref["iters"]
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    w.update(_worst(out, a, ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf")))
    print(f"host stages vs oracle, fused step: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")
    assert max(w.values()) <= TOL, w
    # counters: total calls and microphysics points agree with the per-point record
    calls = out["calls"]
    assert int(out["counters"][0]) == int(np.abs(calls).sum())
    assert int(out["counters"][2]) == int((calls != 0).sum())
    assert int(out["counters"][1]) == 0


def test_host_substep_cap_counts_external_tendencies_once():
    """A point stopped by the sub-step cap keeps the processes integrated so far and counts
    its external tendencies (here the sedimentation of rr, ri, rs, rg) once, through the
    sources: with a cap of one call, every column's water still changes only by its surface
    precipitation, the capped points are those that needed more calls, and the other points
    are unchanged."""
    lib = host_lib.load()
    f = _pipeline_input(128, seed=20260021)
    dt = 50.0
    free = host_lib.adjust_rain_ice(f, dt, 0)
    assert lib.hst_set_substep_cap(1) == 0
    try:
        capped = host_lib.adjust_rain_ice(f, dt, 0)
    finally:
        assert lib.hst_set_substep_cap(4096) == 0
    more = np.abs(free["calls"]) > 1
    assert more.sum() > 100
    assert int(capped["counters"][1]) == int(more.sum())
    assert np.array_equal(capped["calls"] < 0, more)
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
        assert np.array_equal(capped[k][~more], free[k][~more]), k
    rho, dz = f["rhodref"], f["dzz"]
    w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    col0 = (rho * dz * sum(f[k] for k in w)).sum(0) * dt
    col1 = (rho * dz * sum(capped[k] for k in w)).sum(0) * dt
    prec = -1000.0 * (capped["inprr"] + capped["inprs"] + capped["inprg"]) * dt
    err = np.max(np.abs(col1 - col0 - prec) / col0)
    print("cap of one call: capped points", int(more.sum()), "column water budget max rel err", err)
    assert err < 1e-12


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_host_rain_ice_golden(golden_dir, tag):
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    p = "" if tag == "base" else tag + "_"
    s = {k[len(p) + 3:]: g[k] for k in g.files if k.startswith(p + "in_")}
    ref = {k: g[p + "out_" + k] for k in RAIN_OUT}
    out = host_lib.rain_ice(s, float(g["dt"]), 0)
    iters = oracle_lib.rain_ice(s, float(g["dt"]), with_iters=True)["iters"]
    flips = out["calls"] != iters
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    print(f"host stages rain_ice golden {tag}: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")


def test_host_mixed_precision_water_budget():
    """Mixed path (float rates and lookups): every column's total water changes only by
    its surface precipitation, and the sources stay close to FP64."""
    f = _pipeline_input(256, seed=20260009)
    dt = 50.0
    o64 = host_lib.adjust_rain_ice(f, dt, 0)
    o32 = host_lib.adjust_rain_ice(f, dt, 1)
    rho, dz = f["rhodref"], f["dzz"]
    w0 = sum(f[k] for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs"))
    for o in (o64, o32):
        w1 = sum(o[k] for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs"))
        col0 = (rho * dz * w0).sum(0) * dt
        col1 = (rho * dz * w1).sum(0) * dt
        prec = -1000.0 * (o["inprr"] + o["inprs"] + o["inprg"]) * dt
        assert np.max(np.abs(col1 - col0 - prec) / col0) < 1e-12
    for k in ("rrs", "rss", "rgs"):
        scale = np.abs(o64[k]).max()
        assert np.abs(o32[k] - o64[k]).max() <= 1e-2 * scale, k


def test_host_stages_sanitized(tmp_path):
    """ASan + UBSan (float-to-int casts included) over the fused step, FP64 and mixed."""
    from ice3_amd import _lib as L
    exe = host_lib.build_sanitized()
    f = _pipeline_input(64, seed=20260010)
    for k in L.PIPE_IN_FIELDS:
        np.ascontiguousarray(f[k], dtype=np.float64).tofile(str(tmp_path / f"{k}.bin"))
    for prec in ("0", "1"):
        r = subprocess.run([exe, str(tmp_path), "64", "90", "50", prec], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        assert "calls" in r.stdout


def test_host_stages_nonfinite_inputs_stay_in_their_column(tmp_path):
    """NaN and infinities in the inputs of the fused step (kernels' stage source on the CPU,
    FP64 and mixed): every other column's outputs are bitwise those of the clean run, no
    point reaches the sub-step cap, and the sanitizer build (ASan, UBSan with float-to-int
    casts) stays clean, so no table index or loop bound is taken from a non-finite value."""
    from ice3_amd import _lib as L
    f = _pipeline_input(16, seed=20260010)
    bad = {k: np.array(v, dtype=np.float64, copy=True) for k, v in f.items()}
    plan = (("th", 1, np.nan), ("pabs", 2, np.inf), ("rhodref", 4, np.nan), ("sigs", 5, np.nan),
            ("rrs", 6, np.inf), ("rgs", 7, -np.inf), ("ci", 8, np.nan), ("dzz", 9, np.nan),
            ("rc_mf", 10, np.inf), ("sigqsat", 11, np.nan), ("exn", 12, np.nan), ("rv", 13, np.nan))
    for k, col, val in plan:
        if bad[k].ndim == 2:
            bad[k][[0, 30, 60, 89], col] = val
        else:
            bad[k][col] = val
    hit = sorted({col for _, col, _ in plan})
    keep = np.setdiff1d(np.arange(16), hit)
    for prec in (0, 1):
        o0 = host_lib.adjust_rain_ice(f, 50.0, prec)
        o1 = host_lib.adjust_rain_ice(bad, 50.0, prec)
        assert o1["counters"][1] == 0
        for k, v in o0.items():
            if k == "counters":
                continue
            assert np.array_equal(o1[k][..., keep], v[..., keep], equal_nan=True), (prec, k)
    exe = host_lib.build_sanitized()
    for k in L.PIPE_IN_FIELDS:
        np.ascontiguousarray(bad[k], dtype=np.float64).tofile(str(tmp_path / f"{k}.bin"))
    for prec in ("0", "1"):
        r = subprocess.run([exe, str(tmp_path), "16", "90", "50", prec], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        assert "capped 0" in r.stdout


def test_tiwmx_tables_closed_forms():
    """The TIWMX tables the kernels read (built on the host as INI_TIWMX does) against the
    functions of micro/mode_tiwmx_fun.F90 evaluated here in numpy at every tabulated
    temperature: ESATW, ESATI (water above 273.15 K), AA2, BB3, AM3 (0 above 273.15 K) and
    REDIN, each from its own closed form and the reference's constants."""
    import ctypes
    from helpers import rel_err
    lib = host_lib.load()
    n = 37316 - 13200 + 1
    out = np.zeros(6 * n)
    assert lib.hst_tiwmx(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 6
    tab = out.reshape(6, n)
    c = {}
    for line in open(os.path.join(os.path.dirname(__file__), "golden", "constants.txt")):
        p = line.split()
        c[p[0]] = float(p[2])
    t = np.arange(13200, 37317) / 100.0
    esw = np.exp(c["XALPW"] - c["XBETAW"] / t - c["XGAMW"] * np.log(t))
    esi = np.exp(c["XALPI"] - c["XBETAI"] / t - c["XGAMI"] * np.log(t))
    want = [esw, np.where(t > 273.15, esw, np.where(t >= c["XTT"], esw, esi))]
    for q, (name, w) in enumerate(zip(("ESW", "ESI"), want)):
        err = rel_err(tab[q], w, 1e-300).max()
        print(f"TIWMX {name}: max rel err vs closed form {err:.1e}")
        assert err < 2e-14, name  # numpy vs libm exp/log last-ulp noise on arguments near 60
    # the functions built on them, from the tables' own pressures (no tabulated T lies strictly
    # between 273.15 K and XTT, so ESATI is ESTABI below XTT and ESTABW from XTT up)
    esw, esati = tab[0], np.where(t >= c["XTT"], tab[0], tab[1])
    want = [esw, tab[1],
            (c["XLSTT"] + (c["XCPV"] - c["XCI"]) * (t - c["XTT"])) ** 2 /
            (2.38e-2 + 0.0071e-2 * (t - c["XTT"])) / (t ** 2 * c["XRV"]),
            c["XRV"] / (0.211e-4 * (t / c["XTT"]) ** 1.94 * c["XP00"]) * t / esati,
            np.where(t > 273.15, 0.0, 1000.0 * np.exp(12.96 * (esw / esati - 1.0) - 0.639)),
            1.0 - (1.0 - np.clip((c["XTT"] - t) / 20.0, 0, 1)) /
            (np.clip((c["XTT"] - t) / 20.0, 0, 1) ** 3 + (1.0 - np.clip((c["XTT"] - t) / 20.0, 0, 1)) ** 3) ** 0.333]
    for q, (name, w) in enumerate(zip(("ESW", "ESI", "A2", "BB3", "AM3", "REDIN"), want)):
        if q < 2:
            continue
        # REDIN is 1 minus a ratio close to 1 just below 0 C: its error is absolute (floor 1)
        err = rel_err(tab[q], w, 1.0 if name == "REDIN" else 1e-300).max()
        print(f"TIWMX {name}: max rel err vs closed form {err:.1e}")
        assert err < 4e-15, name


def _ocnd2_input(ncol, seed):
    from ice3_amd import synthetic
    f = _pipeline_input(ncol, seed=seed)
    f["zz"] = np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(90)[2][:, None], f["th"].shape))
    f["ice_cld_wgt"] = np.random.Generator(np.random.Philox(seed + 1)).uniform(0.5, 1.5, ncol)
    return f


def test_host_pipeline_ocnd2_matches_oracle():
    """The kernels' stage source with OCND2 in both halves (ICE_ADJUST's ice cloud, then
    RAIN_ICE with the tabulated nucleation and fields, collision factors and ICE4_FAST_RI_RS)
    against the oracle at the FP64 bar, flips counted; the water budget holds."""
    from helpers import flips_of
    from oracle import oracle_lib, ref_lib
    f = _ocnd2_input(128, 20260047)
    out = host_lib.adjust_rain_ice_ocnd2(f, 50.0)
    a = oracle_lib.ice_adjust_ocnd2(f)
    ref = oracle_lib.rain_ice_ocnd2(ref_lib.pipeline_state_ocnd2(f, a), with_iters=True)
    flips = flips_of(out, ref)
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips, TOL)
    w.update(_worst(out, a, ("cldfr", "icldfr", "wcldfr", "ssio", "ssiu", "ifr")))
    print(f"host stages vs oracle, fused OCND2 step: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")
    assert max(w.values()) <= TOL, w
    rho, dz = f["rhodref"], f["dzz"]
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    col0 = (rho * dz * sum(f[k] for k in ws)).sum(0) * 50.0
    col1 = (rho * dz * sum(out[k] for k in ws)).sum(0) * 50.0
    pr = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"]) * 50.0
    assert np.max(np.abs(col1 - col0 - pr) / col0) < 1e-12


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_host_rain_ice_ocnd2_golden(golden_dir, tag):
    from helpers import flips_of
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_ocnd2_fp64.npz"))
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in RAIN_OUT}
    out = host_lib.rain_ice_ocnd2(s, float(g["dt"]))
    ref["iters"] = oracle_lib.rain_ice_ocnd2(s, float(g["dt"]), with_iters=True)["iters"]
    w, nf, nb = check_with_flips(out, ref, RAIN_OUT, flips_of(out, ref), TOL)
    print(f"host stages rain_ice OCND2 golden {tag}: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")


# ------------------------------------------------------------- hail (KRR=7)

HAIL_OUT = RAIN_OUT + ("rhs", "inprh")
HAIL_STATE = ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "rht", "ths", "rvs", "rcs", "rrs",
              "ris", "rss", "rgs", "rhs")


def _hail_flips(out, s, run):
    """Branch flips of a hail case against the oracle (helpers.explained_flips)."""
    from helpers import explained_flips
    from oracle import oracle_lib
    ref = run(s, with_iters=True)
    ties = oracle_lib.rounding_ties(lambda x: run(x), s)
    return explained_flips(out, ref, run, s, HAIL_OUT, TOL, HAIL_STATE, ties=ties), ref


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_host_rain_ice_hail_golden(golden_dir, tag):
    """The kernels' hail stage source (KRR=7) against the compiled reference's hail golden
    vectors at the FP64 bar, flips counted."""
    from oracle import oracle_lib
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    ref = {k: g[f"{tag}_out_{k}"] for k in HAIL_OUT}
    out = host_lib.rain_ice_hail(s, float(g["dt"]))
    flips, _ = _hail_flips(out, s, lambda x, **kw: oracle_lib.rain_ice_hail(x, float(g["dt"]), **kw))
    w, nf, nb = check_with_flips(out, ref, HAIL_OUT, flips, TOL)
    print(f"host stages rain_ice hail golden {tag}: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")
    p = f"{tag}_adj_in_"
    a = host_lib.ice_adjust_hail({k[len(p):]: g[k] for k in g.files if k.startswith(p)}, float(g["dt"]))
    aref = {k: g[f"{tag}_adj_out_{k}"] for k in ADJ_OUT}
    wa = _worst(a, aref, ADJ_OUT)
    assert max(wa.values()) <= TOL, wa


def test_host_pipeline_hail_matches_oracle():
    """The fused step with hail (ICE_ADJUST with the hail heat capacity, hail sedimentation in
    the same column walk, the KRR=7 microphysics) against the oracle, flips counted; column
    water (hail and surface hail included) is conserved."""
    from oracle import oracle_lib, ref_lib
    f = _pipeline_input(96, 20260079)
    rng = np.random.Generator(np.random.Philox(20260080))
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
    f["rhs"] = f["rh"] / 50.0
    out = host_lib.adjust_rain_ice_hail(f, 50.0)
    a = oracle_lib.ice_adjust_hail(f)
    s = ref_lib.pipeline_state_hail(f, a)
    flips, _ = _hail_flips(out, s, lambda x, **kw: oracle_lib.rain_ice_hail(x, **kw))
    ref = oracle_lib.rain_ice_hail(s)
    w, nf, nb = check_with_flips(out, ref, HAIL_OUT, flips, TOL)
    print(f"host stages vs oracle, fused hail step: flips {nf}, divergent {nb}, worst {max(w.values()):.2e}")
    rho, dz = f["rhodref"], f["dzz"]
    ws = ("rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")
    col0 = (rho * dz * sum(f[k] for k in ws)).sum(0) * 50.0
    col1 = (rho * dz * sum(out[k] for k in ws)).sum(0) * 50.0
    pr = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"] + out["inprh"]) * 50.0
    assert np.max(np.abs(col1 - col0 - pr) / col0) < 1e-12
