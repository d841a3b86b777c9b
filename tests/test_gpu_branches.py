"""Per-point branch accounting of the microphysics: the GPU's record of the process branches
that each point's tendency calls took (ice3_rain_ice_branches, one thread per point) against
the oracle's record (oracle_record_branches), on the compiled reference's golden inputs (KRR=6
and hail) and on 2,048 columns of the benchmark generator.

A point whose tendency-call count differs between the GPU and the oracle is a branch flip.
Every other point must record the same branches, except where the oracle marks a choice that
rests on rounding residuals: a graupel or hail wet/dry choice decided by the signs of rounding
residuals (its "tie" bit), or a graupel share of the snow freezing budget (riming, rain
accretion) that is a rounding residual (its "share_tie" bit).  Such a branch is decided by the
last bits of the rates.  The bits are those of include/ice3.h's ICE3_BR_* and
oracle/ice3_oracle.h's ORACLE_BR_* (the same order).

ICE_ADJUST's record (ice3_ice_adjust_branches against oracle_record_adjust_branches: the Q1
interval, no cloud, the ZW1/ZW2 signs, the mass-flux limiters, the TRIANGLE pdf cases, INQ1;
ICE3_ABR_* / ORACLE_ABR_*) must be the oracle's at every point."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DT = 50.0
TIE_BIT = 13  # then SHARE_TIE: the oracle's marks of branches decided by rounding residuals
PROCESS_BITS = (1 << TIE_BIT) - 1


def _bits(*names):
    from oracle import oracle_lib
    return sum(1 << oracle_lib.BRANCHES.index(n) for n in names)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


def _gpu_branches(dev, s, dt, hail):
    """ice3_rain_ice_branches on the oracle-format rain_ice state s."""
    import ctypes
    from ice3_amd import _lib
    from ice3_amd._fields import ptr
    lib, ctx = _lib.load(), _lib.context(0)
    nlev, ncol = s["tht"].shape

    def t(v):
        return torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device=dev).clone()
    g = {k: t(s[k]) for k in _lib.RAIN_IN_FIELDS}
    io = {k: t(s[k]) for k in _lib.RAIN_IO_FIELDS}
    out = {k: torch.empty((nlev, ncol) if k in ("evap3d", "rainfr") else (ncol,), dtype=torch.float64,
                          device=dev) for k in _lib.RAIN_OUT_FIELDS}
    br = torch.full((nlev, ncol), -1, dtype=torch.int32, device=dev)
    rin = _lib.RainIn(*[ptr(g[k]) for k in _lib.RAIN_IN_FIELDS])
    rio = _lib.RainIO(*[ptr(io[k]) for k in _lib.RAIN_IO_FIELDS])
    rout = _lib.RainOut(*[ptr(out[k]) for k in _lib.RAIN_OUT_FIELDS])
    h = None
    if hail:
        hv = {"rht": t(s["rht"]), "rhs": t(s["rhs"]), "inprh": torch.empty(ncol, dtype=torch.float64, device=dev)}
        h = ctypes.byref(_lib.Hail(*[ptr(hv[k]) for k in _lib.HAIL_FIELDS]))
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib.ice3_rain_ice_branches(ctx, ncol, nlev, ncol, dt, ctypes.byref(rin), ctypes.byref(rio),
                                    ctypes.byref(rout), h, ptr(br), sp)
    assert rc == _lib.ICE3_OK, rc
    torch.cuda.synchronize()
    return _np(br).astype(np.uint32)


def _gpu_calls(s, dt, hail):
    """Tendency calls per point of the production rain_ice drop-in."""
    from ice3_amd import RainIce, RainIceHail
    m = dict(exn="exn", dzz="dzz", rhodj="rhodj", rhodref="rhodref", exnref="exn_ref", pres="pabs",
             cf="cldfr", hlc_hrc="hlc_hrc", hlc_hcf="hlc_hcf", hli_hri="hli_hri", hli_hcf="hli_hcf",
             sigs="sigs", th_t="tht", rv_t="rvt", rc_t="rct", rr_t="rrt", ri_t="rit", rs_t="rst",
             rg_t="rgt", ci_t="ci", ths="ths", rvs="rvs", rcs="rcs", rrs="rrs", ris="ris", rss="rss",
             rgs="rgs")
    if hail:
        m.update(rh_t="rht", rhs="rhs")
    state = {k: s[v] for k, v in m.items()}
    _, diags = (RainIceHail() if hail else RainIce())(state, dt, return_calls=True)
    return _np(diags["calls"])


def _compare(dev, s, dt, hail, label):
    from oracle import oracle_lib
    fn = oracle_lib.rain_ice_hail if hail else oracle_lib.rain_ice
    o, ref = oracle_lib.record_branches(fn, s, dt, with_iters=True)
    got = _gpu_branches(dev, s, dt, hail)
    calls = _gpu_calls(s, dt, hail)
    assert got.shape == ref.shape and bool((got >> TIE_BIT == 0).all())
    flips = calls != o["iters"]
    # a wet/dry tie decides the wet/dry-dependent branches, a share tie riming and accretion
    tie = ((ref >> TIE_BIT) & 1).astype(bool)
    share = ((ref >> (TIE_BIT + 1)) & 1).astype(bool)
    waived = np.where(tie, _bits("wetg", "dryg", "wetgh", "weth", "dryh", "dryhg"), 0) | \
        np.where(share, _bits("riming", "raccsg"), 0)
    differ = ((got ^ ref) & PROCESS_BITS) != 0
    unexplained = (((got ^ ref) & PROCESS_BITS & ~waived.astype(np.uint32)) != 0) & ~flips
    tie = tie | share
    n_gpu = oracle_lib.branch_counts(got)
    n_ref = oracle_lib.branch_counts(ref & PROCESS_BITS)
    taken = {b: (n_gpu[b], n_ref[b]) for b in oracle_lib.BRANCHES[:TIE_BIT] if n_ref[b] or n_gpu[b]}
    print(f"{label}: {flips.sum()} call-count flips, {(differ & ~flips).sum()} other points with a "
          f"different record ({(differ & ~flips & tie).sum()} at rounding ties), {unexplained.sum()} "
          f"unexplained, of {got.size}; points per branch (GPU, oracle): {taken}")
    assert unexplained.sum() == 0, np.argwhere(unexplained)[:10]
    return taken


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_branch_record_golden(dev, golden_dir, tag):
    """KRR=6, on the compiled reference's rain_ice golden inputs."""
    g = np.load(os.path.join(golden_dir, "rain_ice_fp64.npz"))
    p = "" if tag == "base" else tag + "_"
    s = {k[len(p) + 3:]: g[k] for k in g.files if k.startswith(p + "in_")}
    taken = _compare(dev, s, float(g["dt"]), False, f"rain_ice golden {tag}")
    assert sum(n for n, _ in taken.values()) > 0


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_branch_record_hail_golden(dev, golden_dir, tag):
    """KRR=7, on the compiled reference's hail golden inputs (every hail branch is taken there)."""
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    s = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    taken = _compare(dev, s, float(g["dt"]), True, f"rain_ice hail golden {tag}")
    assert sum(n for n, _ in taken.values()) > 0


def test_branch_record_pipeline(dev):
    """KRR=6 on 2,048 columns of the benchmark generator (the oracle's ICE_ADJUST first)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(2048, 90, seed=20260071, variant="pipeline")
    s = ref_lib.pipeline_state(f, oracle_lib.ice_adjust(f))
    taken = _compare(dev, s, DT, False, "pipeline 2048 columns")
    for b in ("riming", "raccsg", "wetg", "dryg"):
        assert taken.get(b, (0, 0))[1] > 0, (b, taken)


def test_branch_record_hail_pipeline(dev):
    """KRR=7 on 2,048 columns of the benchmark generator with hail where there is graupel
    (rh = rg U(0.2, 2), its source rh / dt; the oracle's ICE_ADJUST with hail first)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(2048, 90, seed=20260072, variant="pipeline")
    rng = np.random.Generator(np.random.Philox(20260073))
    f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
    f["rhs"] = f["rh"] / DT
    s = ref_lib.pipeline_state_hail(f, oracle_lib.ice_adjust_hail(f))
    taken = _compare(dev, s, DT, True, "hail pipeline 2048 columns")
    for b in ("wetgh", "weth", "dryh", "hmltr"):
        assert taken.get(b, (0, 0))[1] > 0, (b, taken)


def _gpu_adjust_branches(dev, f, dt, hail):
    """ice3_ice_adjust_branches on the oracle-format ICE_ADJUST fields f."""
    import ctypes
    from ice3_amd import _lib
    from ice3_amd._fields import ptr
    lib, ctx = _lib.load(), _lib.context(0)
    nlev, ncol = f["th"].shape

    def t(v):
        return torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device=dev).clone()
    g = {k: t(f[k]) for k in _lib.ADJ_IN_FIELDS}
    out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=dev) for k in _lib.ADJ_OUT_FIELDS}
    rh = t(f["rh"]) if hail else None
    br = torch.full((nlev, ncol), -1, dtype=torch.int32, device=dev)
    ain = _lib.AdjustIn(*[ptr(g[k]) for k in _lib.ADJ_IN_FIELDS])
    aout = _lib.AdjustOut(*[ptr(out[k]) for k in _lib.ADJ_OUT_FIELDS])
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib.ice3_ice_adjust_branches(ctx, ncol, nlev, ncol, dt, ctypes.byref(ain),
                                      ptr(rh) if hail else None, ctypes.byref(aout), ptr(br), sp)
    assert rc == _lib.ICE3_OK, rc
    torch.cuda.synchronize()
    return _np(br).astype(np.uint32), {k: _np(v) for k, v in out.items()}


def _compare_adjust(dev, f, dt, hail, label):
    from oracle import oracle_lib
    fn = oracle_lib.ice_adjust_hail if hail else oracle_lib.ice_adjust
    o, ref = oracle_lib.record_adjust_branches(fn, f, dt)
    got, out = _gpu_adjust_branches(dev, f, dt, hail)
    # the hook's outputs are those of the production ICE_ADJUST: the oracle's to the parity
    # tolerance (1e-11 relative, floors of 1e-18 kg/kg on mixing ratios and 1e-18/dt on sources)
    from helpers import floor_for, rel_err
    for k in ("t", "rv_out", "rc_out", "ri_out", "cldfr", "rvs", "rcs", "ris", "ths"):
        assert rel_err(out[k], o[k], floor_for(k, 1e-11, dt)).max() <= 1e-11, k
    fg, fr = oracle_lib.adjust_branch_fields(got), oracle_lib.adjust_branch_fields(ref)
    taken = {n: np.unique(fr[n]).tolist() for n in fr if n != "inq1"}
    differ = {n: int((fg[n] != fr[n]).sum()) for n in fg}
    print(f"{label}: points with a different record per field {differ} of {got.size}; "
          f"values taken (oracle): {taken}, INQ1 + 22 in [{fr['inq1'].min()}, {fr['inq1'].max()}]")
    assert got.shape == ref.shape and bool((got >> 17 == 0).all())
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:10]
    return taken


@pytest.mark.parametrize("tag", ["base", "mf"])
def test_adjust_branch_record_golden(dev, golden_dir, tag):
    """ICE_ADJUST (KRR=6) on the compiled reference's golden inputs: the record is the
    oracle's at every point."""
    g = np.load(os.path.join(golden_dir, "ice_adjust_fp64.npz"))
    f = {k[len(tag) + 4:]: g[k] for k in g.files if k.startswith(tag + "_in_")}
    taken = _compare_adjust(dev, f, float(g["dt"]), False, f"ice_adjust golden {tag}")
    assert len(taken["q1"]) >= 2 and len(taken["pdfc"]) >= 1


@pytest.mark.parametrize("tag", ["base", "cov"])
def test_adjust_branch_record_hail_golden(dev, golden_dir, tag):
    """ICE_ADJUST with hail (KRR=7) on the compiled reference's hail golden inputs."""
    g = np.load(os.path.join(golden_dir, "rain_ice_hail_fp64.npz"))
    p = tag + "_adj_in_"
    f = {k[len(p):]: g[k] for k in g.files if k.startswith(p)}
    _compare_adjust(dev, f, float(g["dt"]), True, f"ice_adjust hail golden {tag}")


@pytest.mark.parametrize("hail", [False, True])
def test_adjust_branch_record_pipeline(dev, hail):
    """ICE_ADJUST on 2,048 columns of the benchmark generator (with hail where there is
    graupel for KRR=7): every branch value is taken and the record is the oracle's."""
    from ice3_amd import synthetic
    f = synthetic.generate(2048, 90, seed=20260076, variant="pipeline")
    if hail:
        rng = np.random.Generator(np.random.Philox(20260077))
        f["rh"] = np.ascontiguousarray(f["rg"] * rng.uniform(0.2, 2.0, f["rg"].shape))
    taken = _compare_adjust(dev, f, DT, hail, f"ice_adjust pipeline 2048 columns{' hail' if hail else ''}")
    assert taken["q1"] == [0, 1, 2] and taken["nocloud"] == [0, 1], taken
