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


BRANCHES = ("rrhong", "rimltc", "riming", "raccsg", "smltg", "wetg", "dryg", "gmltr", "wetgh",
            "weth", "dryh", "dryhg", "hmltr", "tie")


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


def branch_counts(flags) -> dict:
    """Number of points that took each branch."""
    return {b: int(((flags >> i) & 1).sum()) for i, b in enumerate(BRANCHES)}


def rounding_ties(fn, *a, **kw):
    """Points where a graupel or hail wet/dry growth choice of the oracle's run rests on the
    signs of rounding residuals (the "tie" branch record): their branch is decided by the
    last bits of the rates."""
    _, flags = record_branches(fn, *a, **kw)
    return ((flags >> BRANCHES.index("tie")) & 1).astype(bool)
