"""ctypes binding to the C restatement of the oracle (oracle/_build/libice3_oracle.so).

Test infrastructure: tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
bench.py use it as the checker.  The constants it runs with are the compiled
reference's own values, read from tests/golden/constants.txt.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_HERE, "_build", "libice3_oracle.so")
CONSTANTS = os.path.join(_ROOT, "tests", "golden", "constants.txt")

ADJ_IN = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
          "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths")
ADJ_OUT = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
           "cph", "lv", "ls", "srcs", "th_out", "rvs", "rcs", "ris", "ths")
RAIN_IN = ("exn", "dzz", "rhodj", "rhodref", "exn_ref", "pabs", "cldfr", "hlc_hrc", "hlc_hcf",
           "hli_hri", "hli_hcf", "tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "sigs")
RAIN_IO = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci", "evap3d", "rainfr")
RAIN_COL = ("inprc", "inprr", "inpri", "inprs", "inprg")

_lib = None


def build():
    """Compile the C restatement (gcc, a second or two)."""
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=_ROOT, check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_load_constants.argtypes = [ctypes.c_char_p]
        rc = lib.oracle_load_constants(CONSTANTS.encode())
        if rc != 0:
            raise RuntimeError(f"oracle_load_constants({CONSTANTS}) failed: {rc}")
        _lib = lib
    return _lib


def _arr(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*[p.ctypes.data for p in ptrs])


def _chk(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def ice_adjust(f: dict, dt: float = 50.0) -> dict:
    lib = _load()
    nlev, ncol = f["th"].shape
    ins = [_chk(f[k]) for k in ADJ_IN]
    outs = {k: np.empty((nlev, ncol)) for k in ADJ_OUT}
    sq = _chk(f["sigqsat"])
    rc = lib.oracle_ice_adjust(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
                               ctypes.c_void_p(sq.ctypes.data), _arr(ins),
                               _arr([outs[k] for k in ADJ_OUT]))
    if rc != 0:
        raise RuntimeError(f"oracle_ice_adjust failed: {rc}")
    return outs


OCND2_X = ("icldfr", "wcldfr", "ssio", "ssiu", "ifr")


def ice_adjust_ocnd2(f: dict, dt: float = 50.0) -> dict:
    """ICE_ADJUST with LOCND2; ``f`` adds ``zz`` (nlev, ncol) and ``ice_cld_wgt`` (ncol)."""
    lib = _load()
    nlev, ncol = f["th"].shape
    ins = [_chk(f[k]) for k in ADJ_IN]
    outs = {k: np.empty((nlev, ncol)) for k in ADJ_OUT}
    xs = {k: np.empty((nlev, ncol)) for k in OCND2_X}
    sq, w, zz = _chk(f["sigqsat"]), _chk(f["ice_cld_wgt"]), _chk(f["zz"])
    rc = lib.oracle_ice_adjust_ocnd2(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
                                     ctypes.c_void_p(sq.ctypes.data), ctypes.c_void_p(w.ctypes.data),
                                     ctypes.c_void_p(zz.ctypes.data), _arr(ins),
                                     _arr([outs[k] for k in ADJ_OUT]), _arr([xs[k] for k in OCND2_X]))
    if rc != 0:
        raise RuntimeError(f"oracle_ice_adjust_ocnd2 failed: {rc}")
    outs.update(xs)
    return outs


def rain_ice(s: dict, dt: float = 50.0, with_iters: bool = False) -> dict:
    """``s`` uses the keys of RAIN_IN plus the sources ths..rgs and ci."""
    lib = _load()
    nlev, ncol = s["tht"].shape
    ins = [_chk(s[k]) for k in RAIN_IN]
    io = {k: (np.array(s[k], dtype=np.float64, order="C") if k in s else np.zeros((nlev, ncol)))
          for k in RAIN_IO}
    col = {k: np.zeros(ncol) for k in RAIN_COL}
    iters = np.zeros((nlev, ncol), dtype=np.int32)
    rc = lib.oracle_rain_ice(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _arr(ins),
                             _arr([io[k] for k in RAIN_IO]), _arr([col[k] for k in RAIN_COL]),
                             ctypes.c_void_p(iters.ctypes.data))
    if rc == -1:
        raise RuntimeError("oracle_rain_ice: constants not loaded")
    out = dict(io)
    out.update(col)
    out["status"] = rc
    if with_iters:
        out["iters"] = iters
    return out


RAIN_OCND2 = ("icldfr", "ssio", "ssiu", "ifr")


def rain_ice_ocnd2(s: dict, dt: float = 50.0, with_iters: bool = False) -> dict:
    """RAIN_ICE with LOCND2; ``s`` as for rain_ice plus ICE_ADJUST's OCND2 outputs
    ``icldfr, ssio, ssiu, ifr``."""
    lib = _load()
    nlev, ncol = s["tht"].shape
    ins = [_chk(s[k]) for k in RAIN_IN]
    x4 = [_chk(s[k]) for k in RAIN_OCND2]
    io = {k: (np.array(s[k], dtype=np.float64, order="C") if k in s else np.zeros((nlev, ncol)))
          for k in RAIN_IO}
    col = {k: np.zeros(ncol) for k in RAIN_COL}
    iters = np.zeros((nlev, ncol), dtype=np.int32)
    rc = lib.oracle_rain_ice_ocnd2(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _arr(ins),
                                   _arr(x4), _arr([io[k] for k in RAIN_IO]), _arr([col[k] for k in RAIN_COL]),
                                   ctypes.c_void_p(iters.ctypes.data))
    if rc in (-1, -3):
        raise RuntimeError(f"oracle_rain_ice_ocnd2 failed: {rc}")
    out = dict(io)
    out.update(col)
    out["status"] = rc
    if with_iters:
        out["iters"] = iters
    return out


def ice_adjust_hail(f: dict, dt: float = 50.0) -> dict:
    """ICE_ADJUST with KRR = 7; ``f`` adds the hail mixing ratio ``rh``."""
    lib = _load()
    nlev, ncol = f["th"].shape
    ins = [_chk(f[k]) for k in ADJ_IN]
    outs = {k: np.empty((nlev, ncol)) for k in ADJ_OUT}
    sq, rh = _chk(f["sigqsat"]), _chk(f["rh"])
    rc = lib.oracle_ice_adjust_hail(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
                                    ctypes.c_void_p(sq.ctypes.data), _arr(ins),
                                    ctypes.c_void_p(rh.ctypes.data), _arr([outs[k] for k in ADJ_OUT]))
    if rc != 0:
        raise RuntimeError(f"oracle_ice_adjust_hail failed: {rc}")
    return outs


def rain_ice_hail(s: dict, dt: float = 50.0, with_iters: bool = False) -> dict:
    """RAIN_ICE with KRR = 7; ``s`` as for rain_ice plus the hail state ``rht`` and source
    ``rhs``.  Returns rain_ice's outputs plus the updated ``rhs`` and ``inprh``."""
    lib = _load()
    nlev, ncol = s["tht"].shape
    ins = [_chk(s[k]) for k in RAIN_IN]
    io = {k: (np.array(s[k], dtype=np.float64, order="C") if k in s else np.zeros((nlev, ncol)))
          for k in RAIN_IO}
    col = {k: np.zeros(ncol) for k in RAIN_COL}
    rht = _chk(s["rht"])
    rhs = np.array(s["rhs"], dtype=np.float64, order="C")
    inprh = np.zeros(ncol)
    iters = np.zeros((nlev, ncol), dtype=np.int32)
    rc = lib.oracle_rain_ice_hail(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _arr(ins),
                                  ctypes.c_void_p(rht.ctypes.data), ctypes.c_void_p(rhs.ctypes.data),
                                  ctypes.c_void_p(inprh.ctypes.data), _arr([io[k] for k in RAIN_IO]),
                                  _arr([col[k] for k in RAIN_COL]), ctypes.c_void_p(iters.ctypes.data))
    if rc == -1:
        raise RuntimeError("oracle_rain_ice_hail: constants (with the hail tables) not loaded")
    out = dict(io)
    out.update(col)
    out.update(rhs=rhs, inprh=inprh, status=rc)
    if with_iters:
        out["iters"] = iters
    return out


SEDIM = {"STAT": 0, "SPLI": 1}


def set_sedimentation(scheme: str = "STAT", maxcfl: float = 0.8):
    """Sedimentation scheme of the rain_ice calls that follow (PARAM_ICE CSEDIM, and
    XSPLIT_MAXCFL for "SPLI"; the AROME default is "STAT")."""
    lib = _load()
    rc = lib.oracle_set_sedimentation(ctypes.c_int(SEDIM[scheme]), ctypes.c_double(maxcfl))
    if rc != 0:
        raise ValueError(f"oracle_set_sedimentation({scheme}, {maxcfl}) failed: {rc}")


SPLIT_SPECIES = ("r", "i", "s", "g", "h")


def sedim_split(f: dict, dt: float = 50.0, krr: int = 6, maxcfl: float = 0.8) -> dict:
    """ICE4_SEDIMENTATION_SPLIT alone.  ``f`` holds ``dzz``, ``rhodref`` and, for each species
    x of rain, pristine ice, snow, graupel (and hail with krr = 7), the mixing ratio at t
    ``rxt`` and the source ``rxs`` (nlev, ncol).  Returns the updated sources ``rxs``, the
    surface rates ``inprx`` (ncol) and the sub-step counts ``nsub`` (5, ncol)."""
    lib = _load()
    nlev, ncol = f["dzz"].shape
    sp = SPLIT_SPECIES[:5 if krr == 7 else 4]
    dummy = np.zeros((nlev, ncol))
    rt = [_chk(f[f"r{x}t"]) if x in sp else dummy for x in SPLIT_SPECIES]
    rs = [np.array(f[f"r{x}s"], dtype=np.float64, order="C") if x in sp else dummy.copy()
          for x in SPLIT_SPECIES]
    inpr = [np.zeros(ncol) for _ in SPLIT_SPECIES]
    nsub = np.zeros((5, ncol), dtype=np.int32)
    dzz, rho = _chk(f["dzz"]), _chk(f["rhodref"])
    rc = lib.oracle_sedim_split(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
                                ctypes.c_int(krr), ctypes.c_double(maxcfl),
                                ctypes.c_void_p(dzz.ctypes.data), ctypes.c_void_p(rho.ctypes.data),
                                _arr(rt), _arr(rs), _arr(inpr), ctypes.c_void_p(nsub.ctypes.data))
    if rc != 0:
        raise RuntimeError(f"oracle_sedim_split failed: {rc}")
    out = {f"r{x}s": rs[q] for q, x in enumerate(SPLIT_SPECIES) if x in sp}
    out.update({f"inpr{x}": inpr[q] for q, x in enumerate(SPLIT_SPECIES) if x in sp})
    out["nsub"] = nsub
    return out


SPLI_T = ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "rht")
SPLI_S = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")
SPLI_PR = ("inprc", "inprr", "inpri", "inprs", "inprg", "inprh")


def sedimentation_spli(s: dict, dt: float = 50.0, krr: int = 6, maxcfl: float = 0.8) -> dict:
    """ICE4_SEDIMENTATION with CSEDIM='SPLI' as RAIN_ICE runs it before the microphysics: the
    split sedimentation, then the negativity correction of every source.  ``s`` holds
    ``dzz``, ``rhodref``, ``exn``, the state at t ``tht .. rgt`` (``rht`` with krr = 7) and
    the sources ``ths .. rgs`` (``rhs``).  Returns the updated sources, the surface rates
    ``inprc .. inprg`` (``inprh``) and the sub-step counts ``nsub`` (5, ncol)."""
    lib = _load()
    nlev, ncol = s["dzz"].shape
    n = 8 if krr == 7 else 7
    dummy = np.zeros((nlev, ncol))
    t = [_chk(s[k]) for k in SPLI_T[:n]] + [dummy] * (8 - n)
    src = [np.array(s[k], dtype=np.float64, order="C") for k in SPLI_S[:n]] + [dummy.copy()] * (8 - n)
    pr = [np.zeros(ncol) for _ in SPLI_PR]
    nsub = np.zeros((5, ncol), dtype=np.int32)
    dzz, rho, exn = _chk(s["dzz"]), _chk(s["rhodref"]), _chk(s["exn"])
    rc = lib.oracle_sedimentation_spli(ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
                                       ctypes.c_int(krr), ctypes.c_double(maxcfl),
                                       ctypes.c_void_p(dzz.ctypes.data), ctypes.c_void_p(rho.ctypes.data),
                                       ctypes.c_void_p(exn.ctypes.data), _arr(t), _arr(src), _arr(pr),
                                       ctypes.c_void_p(nsub.ctypes.data))
    if rc != 0:
        raise RuntimeError(f"oracle_sedimentation_spli failed: {rc}")
    out = {k: src[q] for q, k in enumerate(SPLI_S[:n])}
    out.update({k: pr[q] for q, k in enumerate(SPLI_PR[:5 if krr == 6 else 6])})
    out["nsub"] = nsub
    return out


BRANCHES = ("rrhong", "rimltc", "riming", "raccsg", "smltg", "wetg", "dryg", "gmltr", "wetgh",
            "weth", "dryh", "dryhg", "hmltr", "tie", "share_tie")


def record_branches(fn, *a, **kw):
    """Runs one rain_ice call of this module (fn) with the branch record on; returns its
    outputs and, per point, the mask of the ORACLE_BR_* branches taken (BRANCHES order)."""
    lib = _load()
    s = a[0]
    flags = np.zeros(s["tht"].shape, dtype=np.uint32)
    lib.oracle_record_branches(ctypes.c_void_p(flags.ctypes.data))
    try:
        out = fn(*a, **kw)
    finally:
        lib.oracle_record_branches(ctypes.c_void_p(None))
    return out, flags


# ICE_ADJUST's branch record (oracle_record_adjust_branches): (name, first bit, width).
ADJ_BRANCHES = (("q1", 0, 2), ("nocloud", 2, 1), ("zw1neg", 3, 1), ("zw2neg", 4, 1),
                ("mflimit", 5, 1), ("mfout", 6, 1), ("pdfc", 7, 2), ("pdfi", 9, 2), ("inq1", 11, 6))


def record_adjust_branches(fn, f, *a, **kw):
    """Runs one ice_adjust call of this module (fn: ice_adjust or ice_adjust_hail on the
    fields f) with ICE_ADJUST's branch record on; returns its outputs and the per-point mask
    (ADJ_BRANCHES layout)."""
    lib = _load()
    flags = np.zeros(np.shape(f["th"]), dtype=np.uint32)
    lib.oracle_record_adjust_branches(ctypes.c_void_p(flags.ctypes.data))
    try:
        out = fn(f, *a, **kw)
    finally:
        lib.oracle_record_adjust_branches(ctypes.c_void_p(None))
    return out, flags


def adjust_branch_fields(flags) -> dict:
    """The ADJ_BRANCHES fields of an ICE_ADJUST branch record."""
    return {n: (flags >> b) & ((1 << w) - 1) for n, b, w in ADJ_BRANCHES}


def branch_counts(flags) -> dict:
    """Number of points that took each branch."""
    return {b: int(((flags >> i) & 1).sum()) for i, b in enumerate(BRANCHES)}


def rounding_ties(fn, *a, **kw):
    """Points where a graupel or hail wet/dry growth choice of the oracle's run rests on the
    signs of rounding residuals (the "tie" branch record): their branch is decided by the
    last bits of the rates."""
    _, flags = record_branches(fn, *a, **kw)
    return ((flags >> BRANCHES.index("tie")) & 1).astype(bool)
