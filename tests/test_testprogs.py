"""The PHYEX testprogs dataset reader (ice3_amd.testprogs).

The reference's reproduction files (data/ice_adjust.nc, data/rain_ice.nc) are not in its
checkout, so the reader is checked on files written here in the layout its tests read
(tests/components/test_ice_adjust_fortran.py:227-262, tests/components/test_rain_ice_fortran.py:46-128):
synthetic inputs, stored top-down like PHYEX's AROME arrays, with the oracle's results as the
file's outputs.  CPU tests check that every field lands at its level, column and species; the
GPU tests run the drop-ins on what the reader returns and compare with the file's outputs, at the
FP64 bar (the outputs are the oracle's) and at the reference tests' tolerances.  When
$ICE3_TESTPROGS_DATA names a directory holding PHYEX's own files, they are run too."""
import os

import numpy as np
import pytest

from helpers import check_with_flips, floor_for, rel_err

NB, NPROMA, NLEV = 3, 8, 15   # ngpblks, nproma, nflevg
NCOL = NB * NPROMA
DT = 50.0


def _file_layout(a, top_down=True):
    """(nlev, ncol) bottom-up -> (ngpblks, nflevg, nproma), top-down if asked."""
    a = a[::-1] if top_down else a
    return np.ascontiguousarray(a.reshape(NLEV, NB, NPROMA).transpose(1, 0, 2))


def _write(path, variables, attrs=None):
    from scipy.io import netcdf_file
    with netcdf_file(path, "w", version=2) as nc:
        for k, v in dict(ngpblks=NB, nproma=NPROMA, nflevg=NLEV, krr=6, krr1=7).items():
            nc.createDimension(k, v)
        for name, (dims, arr) in variables.items():
            nc.createVariable(name, "d", dims)[:] = arr
        for k, v in (attrs or {}).items():
            setattr(nc, k, v)


P3 = ("ngpblks", "nflevg", "nproma")


def _species(arrays, dim, top_down):
    return (("ngpblks", dim, "nflevg", "nproma"),
            np.stack([_file_layout(a, top_down) for a in arrays], axis=1))


def _adjust_inputs():
    from ice3_amd import synthetic
    return synthetic.generate(NCOL, NLEV, seed=20260005)


def _write_ice_adjust(path, f, out=None, top_down=True):
    L = lambda a: (P3, _file_layout(a, top_down))  # noqa: E731
    v = {"PPABSM": L(f["pabs"]), "PEXNREF": L(f["exn"]), "PRHODREF": L(f["rho_dry_ref"]),
         "PRHODJ": L(f["rho_dry_ref"]), "PSIGS": L(f["sigs"]), "PCF_MF": L(f["cf_mf"]),
         "PRC_MF": L(f["rc_mf"]), "PRI_MF": L(f["ri_mf"]), "PTHS": L(f["ths"]),
         "ZRS": _species([f[k] for k in ("th", "rv", "rc", "rr", "ri", "rs", "rg")], "krr1", top_down),
         "PRS": _species([f["rvs"], f["rcs"], f["rvs"] * 0, f["ris"], f["rvs"] * 0, f["rvs"] * 0],
                         "krr", top_down),
         "ZSIGQSAT": (("ngpblks", "nproma"), f["sigqsat"].reshape(NB, NPROMA))}
    if out is not None:
        z = np.zeros_like(out["rvs"])
        v["PRS_OUT"] = _species([out["rvs"], out["rcs"], z, out["ris"], z, z], "krr", top_down)
        v["PTHS_OUT"] = L(out["ths"])
        v["PCLDFR_OUT"] = L(out["cldfr"])
        v["PSRCS_OUT"] = L(out["srcs"])
        for k in ("hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf"):
            v[f"P{k.upper()}_OUT"] = L(out[k])
    _write(path, v, {"PTSTEP": DT})


def _rain_state():
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(NCOL, NLEV, seed=20260006, variant="pipeline")
    return ref_lib.pipeline_state(f, oracle_lib.ice_adjust(f))


def _write_rain_ice(path, s, out=None, top_down=True):
    L = lambda a: (P3, _file_layout(a, top_down))  # noqa: E731
    v = {"PEXNREF": L(s["exn"]), "PDZZ": L(s["dzz"]), "PRHODJ": L(s["rhodj"]),
         "PRHODREF": L(s["rhodref"]), "PPABSM": L(s["pabs"]), "PCLDFR": L(s["cldfr"]),
         "PTHT": L(s["tht"]), "PSIGS": L(s["sigs"]), "PCIT": L(s["ci"]), "PTHS": L(s["ths"]),
         "PRT": _species([s[k] for k in ("rvt", "rct", "rrt", "rit", "rst", "rgt")], "krr", top_down),
         "PRS": _species([s[k] for k in ("rvs", "rcs", "rrs", "ris", "rss", "rgs")], "krr", top_down)}
    for k in ("hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf"):
        v[f"P{k.upper()}"] = L(s[k])
    if out is not None:
        v["PRS_OUT"] = _species([out[k] for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")],
                                "krr1", top_down)
        for k in ("inprc", "inprr", "inprs", "inprg"):
            v[f"P{k.upper()}_OUT"] = (("ngpblks", "nproma"), out[k].reshape(NB, NPROMA))
    _write(path, v, {"PTSTEP": DT})


ADJ_ARGS = {"pabs": "pabs", "sigs": "sigs", "th": "th", "exn": "exn", "exn_ref": "exn",
            "rho_dry_ref": "rho_dry_ref", "rv": "rv", "rc": "rc", "ri": "ri", "rr": "rr", "rs": "rs",
            "rg": "rg", "cf_mf": "cf_mf", "rc_mf": "rc_mf", "ri_mf": "ri_mf", "rvs": "rvs",
            "rcs": "rcs", "ris": "ris", "ths": "ths", "sigqsat": "sigqsat"}
RAIN_STATE = dict(exn="exn", exnref="exn", dzz="dzz", rhodj="rhodj", rhodref="rhodref", pres="pabs",
                  cf="cldfr", sigs="sigs", hlc_hrc="hlc_hrc", hlc_hcf="hlc_hcf", hli_hri="hli_hri",
                  hli_hcf="hli_hcf", ci_t="ci", th_t="tht", ths="ths", rv_t="rvt", rc_t="rct",
                  rr_t="rrt", ri_t="rit", rs_t="rst", rg_t="rgt", rvs="rvs", rcs="rcs", rrs="rrs",
                  ris="ris", rss="rss", rgs="rgs")


@pytest.mark.parametrize("top_down", [True, False])
def test_ice_adjust_layout(tmp_path, top_down):
    from ice3_amd import testprogs
    from oracle import oracle_lib
    f = _adjust_inputs()
    out = oracle_lib.ice_adjust(f)
    p = str(tmp_path / "ice_adjust.nc")
    _write_ice_adjust(p, f, out, top_down)
    ds = testprogs.read(p)
    assert testprogs.top_down(ds) == top_down
    args, dt, expected = testprogs.ice_adjust_case(ds)
    assert dt == DT
    for k, src in ADJ_ARGS.items():
        assert np.array_equal(args[k], f[src]), k
    for k in ("rvs", "rcs", "ris", "ths", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf", "srcs"):
        assert np.array_equal(expected[k], out[k]), k


@pytest.mark.parametrize("top_down", [True, False])
def test_rain_ice_layout(tmp_path, top_down):
    from ice3_amd import testprogs
    s = _rain_state()
    out = {k: s[k] * 2.0 for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")}
    out.update({k: np.arange(NCOL, dtype=np.float64) + j for j, k in enumerate(("inprc", "inprr", "inprs", "inprg"))})
    p = str(tmp_path / "rain_ice.nc")
    _write_rain_ice(p, s, out, top_down)
    state, dt, expected = testprogs.rain_ice_case(testprogs.read(p))
    for k, src in RAIN_STATE.items():
        assert np.array_equal(state[k], s[src]), k
    for k, v in out.items():
        assert np.array_equal(expected[k], v), k


def test_missing_fields_and_formats(tmp_path):
    from ice3_amd import testprogs
    p = str(tmp_path / "partial.nc")
    _write(p, {"PPABSM": (P3, np.ones((NB, NLEV, NPROMA)))})
    with pytest.raises(KeyError, match="PEXNREF"):
        testprogs.ice_adjust_case(testprogs.read(p))
    h = tmp_path / "hdf5.nc"
    h.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(RuntimeError, match="nccopy"):
        testprogs.read(str(h))
    assert testprogs.find("ice_adjust.nc", str(tmp_path)) is None


# ------------------------------------------------------------------------------------ GPU


def _need_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run_ice_adjust(args, dt):
    from ice3_amd import IceAdjust
    names = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
             "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
    res = IceAdjust()(args["sigqsat"], *[args[k] for k in names], timestep=dt, return_srcs=True)
    keys = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
            "cph", "lv", "ls", "rvs", "rcs", "ris", "ths", "srcs")
    return {k: v.detach().cpu().numpy() for k, v in zip(keys, res)}


def _run_rain_ice(state, dt):
    from ice3_amd import RainIce
    new, diags = RainIce()(state, dt, return_calls=True)
    out = {k: new[k].detach().cpu().numpy() for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")}
    out.update({k: v.detach().cpu().numpy() for k, v in diags.items()})
    return out


def _reference_tolerances(out, expected):
    from ice3_amd import testprogs
    for k, ref in expected.items():
        atol, rtol = testprogs.TOLERANCES.get(k, (1e-5, 1e-4))
        np.testing.assert_allclose(out[k], ref, atol=atol, rtol=rtol, err_msg=k)


@pytest.mark.gpu
def test_ice_adjust_file_through_dropin(tmp_path):
    _need_gpu()
    from ice3_amd import testprogs
    from oracle import oracle_lib
    f = _adjust_inputs()
    p = str(tmp_path / "ice_adjust.nc")
    _write_ice_adjust(p, f, oracle_lib.ice_adjust(f))
    args, dt, expected = testprogs.ice_adjust_case(testprogs.read(p))
    out = _run_ice_adjust(args, dt)
    for k, ref in expected.items():
        r = rel_err(out[k], ref, floor_for(k, 1e-11))
        assert r.max() <= 1e-11, (k, float(r.max()))
    _reference_tolerances(out, expected)


@pytest.mark.gpu
def test_rain_ice_file_through_dropin(tmp_path):
    """Top-down file: the reader must flip it for the sedimentation to run downwards."""
    _need_gpu()
    from ice3_amd import testprogs
    from oracle import oracle_lib
    s = _rain_state()
    ref = oracle_lib.rain_ice(s, DT, with_iters=True)
    p = str(tmp_path / "rain_ice.nc")
    _write_rain_ice(p, s, ref)
    state, dt, expected = testprogs.rain_ice_case(testprogs.read(p))
    out = _run_rain_ice(state, dt)
    flips = out["calls"] != ref["iters"]
    w, nf, nb = check_with_flips(out, expected, tuple(expected), flips, 1e-11)
    print(f"rain_ice through the reader: {nf} flips, {nb} divergent points, worst {max(w.values()):.2e}")
    _reference_tolerances(out, expected)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ice_adjust.nc", "rain_ice.nc"])
def test_phyex_testprogs_files(name):
    """PHYEX's own reproduction data, when $ICE3_TESTPROGS_DATA provides it."""
    from ice3_amd import testprogs
    path = testprogs.find(name)
    if path is None:
        pytest.skip(f"{name}: PHYEX testprogs data not provided (absent from the reference checkout)")
    _need_gpu()
    ds = testprogs.read(path)
    if name == "ice_adjust.nc":
        args, dt, expected = testprogs.ice_adjust_case(ds)
        out = _run_ice_adjust(args, dt)
    else:
        state, dt, expected = testprogs.rain_ice_case(ds)
        out = _run_rain_ice(state, dt)
    _reference_tolerances(out, expected)
