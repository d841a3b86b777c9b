"""ctypes binding to the compiled FP64 PHYEX reference (oracle/_ref/libphyex_ref.so).

Test infrastructure: used to write golden vectors (tests/golden/make_golden.py),
to pin the C restatement in oracle/, and as the CPU baseline of bench.py when
the library has been built.  The library is produced in this container by
``make -f oracle/ref_build.mk`` from the reference sources plus
oracle/ref_driver.F90.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libphyex_ref.so")
# The optimised build of the same sources (-O3, x86-64-v3 with FMA contraction): timing only,
# never a checker (its arithmetic is not the plain IEEE sequence of the parity build).
FAST_LIB_PATH = os.path.join(_HERE, "_ref", "fast", "libphyex_ref.so")
_lib = None
_init_state = None


def available(fast: bool = False) -> bool:
    return os.path.exists(FAST_LIB_PATH if fast else LIB_PATH)


def use_fast_build():
    """Load the optimised timing build instead of the parity build (before the first call;
    for the CPU baseline of bench.py)."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the reference library is already loaded")
    LIB_PATH = FAST_LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{LIB_PATH} not built (make -f oracle/ref_build.mk)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ref_init.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        _lib.ref_set_nproma.argtypes = [ctypes.c_int]
        _lib.ref_dump_constants.argtypes = [ctypes.c_char_p, ctypes.c_int]
    return _lib


def init(dt: float = 50.0, nproma: int = 1, krr: int = 6):
    """INI_CST + INI_RAIN_ICE for the pinned AROME configuration: HCLOUD='ICE3' (krr=6) or
    'ICE4' (krr=7, hail)."""
    global _init_state
    lib = _load()
    if _init_state is None or _init_state[0] != dt or _init_state[2] != krr:
        lib.ref_init(dt, nproma, krr)
        _init_state = (dt, nproma, krr)
    elif _init_state[1] != nproma:
        lib.ref_set_nproma(nproma)
        _init_state = (dt, nproma, krr)
    return lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def dump_constants(path: str, dt: float = 50.0):
    """Constants and tables after the HCLOUD='ICE4' initialisation: the 'ICE3' one gives the
    same values with XRTMIN(1:6) only."""
    lib = init(dt, krr=7)
    b = path.encode()
    lib.ref_dump_constants(b, len(b))


# The compiled reference keeps its work arrays on the stack (Fortran automatic
# arrays sized NIJT x NKT); calls are made on column chunks of this width.
CHUNK = 32


def _chunked(fn, d: dict, *a, **kw) -> dict:
    ncol = next(v for v in d.values() if getattr(v, "ndim", 0) == 2).shape[1]
    if ncol <= CHUNK:
        return fn(d, *a, **kw)
    parts = []
    for c0 in range(0, ncol, CHUNK):
        sub = {k: np.ascontiguousarray(v[..., c0:c0 + CHUNK]) for k, v in d.items()}
        parts.append(fn(sub, *a, **kw))
    return {k: np.concatenate([p[k] for p in parts], axis=-1) for k in parts[0]}


def ice_adjust(f: dict, dt: float = 50.0) -> dict:
    """ICE_ADJUST on a synthetic field dict; returns the 16 JAX-order outputs + srcs."""
    return _chunked(_ice_adjust, f, dt)


def _ice_adjust(f: dict, dt: float) -> dict:
    lib = init(dt)
    nlev, ncol = f["th"].shape
    rvs, rcs, ris, ths = (f[k].copy() for k in ("rvs", "rcs", "ris", "ths"))
    names = ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf", "rv_out", "rc_out", "ri_out",
             "th_out", "srcs", "cph", "lv", "ls")
    out = {k: np.zeros((nlev, ncol)) for k in names}
    lib.ref_ice_adjust(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _p(f["sigqsat"]),
        *[_p(f[k]) for k in ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri",
                             "rr", "rs", "rg", "cf_mf", "rc_mf", "ri_mf")],
        _p(rvs), _p(rcs), _p(ris), _p(ths),
        *[_p(out[k]) for k in names])
    out["t"] = out["th_out"] * f["exn"]
    out.update(rvs=rvs, rcs=rcs, ris=ris, ths=ths)
    return out


def ice_adjust_hail(f: dict, dt: float = 50.0) -> dict:
    """ICE_ADJUST with KRR = 7: ``f`` adds the hail mixing ratio ``rh`` to the ice_adjust
    fields (it enters the heat capacity only)."""
    return _chunked(_ice_adjust_hail, f, dt)


def _ice_adjust_hail(f: dict, dt: float) -> dict:
    lib = init(dt, krr=7)
    nlev, ncol = f["th"].shape
    rvs, rcs, ris, ths = (f[k].copy() for k in ("rvs", "rcs", "ris", "ths"))
    names = ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf", "rv_out", "rc_out", "ri_out",
             "th_out", "srcs", "cph", "lv", "ls")
    out = {k: np.zeros((nlev, ncol)) for k in names}
    lib.ref_ice_adjust_hail(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _p(f["sigqsat"]),
        *[_p(f[k]) for k in ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri",
                             "rr", "rs", "rg", "rh", "cf_mf", "rc_mf", "ri_mf")],
        _p(rvs), _p(rcs), _p(ris), _p(ths),
        *[_p(out[k]) for k in names])
    out["t"] = out["th_out"] * f["exn"]
    out.update(rvs=rvs, rcs=rcs, ris=ris, ths=ths)
    return out


OCND2_OUT = ("cldfr", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf", "rv_out", "rc_out", "ri_out",
             "th_out", "srcs", "icldfr", "wcldfr", "ssio", "ssiu", "ifr")


def ice_adjust_ocnd2(f: dict, dt: float = 50.0) -> dict:
    """ICE_ADJUST with LOCND2 = .TRUE.; ``f`` adds ``zz`` (level heights, (nlev, ncol)) and
    ``ice_cld_wgt`` (per column) to the ice_adjust fields.  Returns the updated sources and
    the outputs named in OCND2_OUT."""
    return _chunked(_ice_adjust_ocnd2, f, dt)


def _ice_adjust_ocnd2(f: dict, dt: float) -> dict:
    lib = init(dt)
    nlev, ncol = f["th"].shape
    rvs, rcs, ris, ths = (f[k].copy() for k in ("rvs", "rcs", "ris", "ths"))
    out = {k: np.zeros((nlev, ncol)) for k in OCND2_OUT}
    lib.ref_ice_adjust_ocnd2(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt), _p(f["sigqsat"]),
        _p(f["ice_cld_wgt"]), _p(f["zz"]),
        *[_p(f[k]) for k in ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri",
                             "rr", "rs", "rg", "cf_mf", "rc_mf", "ri_mf")],
        _p(rvs), _p(rcs), _p(ris), _p(ths),
        *[_p(out[k]) for k in OCND2_OUT])
    out.update(rvs=rvs, rcs=rcs, ris=ris, ths=ths)
    return out


def rain_ice(s: dict, dt: float = 50.0, nproma: int = 1) -> dict:
    """Reinstated RAIN_ICE; ``s`` holds exn, dzz, rhodj, rhodref, exn_ref, pabs, ci, cldfr,
    hlc_hrc, hlc_hcf, hli_hri, hli_hcf, tht, rvt, rct, rrt, rit, rst, rgt, sigs and the
    sources ths..rgs.  Returns updated sources, ci, diagnostics."""
    return _chunked(_rain_ice, s, dt, nproma)


def _rain_ice(s: dict, dt: float, nproma: int) -> dict:
    lib = init(dt, nproma)
    nlev, ncol = s["tht"].shape
    io = {k: s[k].copy() for k in ("ci", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
                                   "ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")}
    col = {k: np.zeros(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")}
    o2 = {k: np.zeros((nlev, ncol)) for k in ("evap3d", "rainfr")}
    lib.ref_rain_ice(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
        *[_p(s[k]) for k in ("exn", "dzz", "rhodj", "rhodref", "exn_ref", "pabs")],
        _p(io["ci"]), _p(s["cldfr"]),
        *[_p(io[k]) for k in ("hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf")],
        *[_p(s[k]) for k in ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "sigs")],
        *[_p(io[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")],
        *[_p(col[k]) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")],
        _p(o2["evap3d"]), _p(o2["rainfr"]))
    io.update(col)
    io.update(o2)
    return io


def rain_ice_ocnd2(s: dict, dt: float = 50.0, nproma: int = 1) -> dict:
    """RAIN_ICE with LOCND2 = .TRUE.; ``s`` holds rain_ice's fields plus ICE_ADJUST's OCND2
    outputs ``icldfr, ssio, ssiu, ifr``."""
    return _chunked(_rain_ice_ocnd2, s, dt, nproma)


def _rain_ice_ocnd2(s: dict, dt: float, nproma: int) -> dict:
    lib = init(dt, nproma)
    nlev, ncol = s["tht"].shape
    io = {k: s[k].copy() for k in ("ci", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
                                   "ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")}
    col = {k: np.zeros(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")}
    o2 = {k: np.zeros((nlev, ncol)) for k in ("evap3d", "rainfr")}
    lib.ref_rain_ice_ocnd2(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
        *[_p(s[k]) for k in ("exn", "dzz", "rhodj", "rhodref", "exn_ref", "pabs")],
        _p(io["ci"]), _p(s["cldfr"]),
        *[_p(io[k]) for k in ("hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf")],
        *[_p(np.ascontiguousarray(s[k], dtype=np.float64)) for k in ("icldfr", "ssio", "ssiu", "ifr")],
        *[_p(s[k]) for k in ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "sigs")],
        *[_p(io[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")],
        *[_p(col[k]) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")],
        _p(o2["evap3d"]), _p(o2["rainfr"]))
    io.update(col)
    io.update(o2)
    return io


def rain_ice_hail(s: dict, dt: float = 50.0, nproma: int = 1) -> dict:
    """RAIN_ICE with KRR = 7; ``s`` holds rain_ice's fields plus the hail state ``rht`` and
    source ``rhs``.  Returns rain_ice's outputs plus the updated ``rhs`` and ``inprh``."""
    return _chunked(_rain_ice_hail, s, dt, nproma)


def _rain_ice_hail(s: dict, dt: float, nproma: int) -> dict:
    lib = init(dt, nproma, krr=7)
    nlev, ncol = s["tht"].shape
    io = {k: s[k].copy() for k in ("ci", "hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf",
                                   "ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")}
    col = {k: np.zeros(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg", "inprh")}
    o2 = {k: np.zeros((nlev, ncol)) for k in ("evap3d", "rainfr")}
    lib.ref_rain_ice_hail(
        ctypes.c_int(ncol), ctypes.c_int(nlev), ctypes.c_double(dt),
        *[_p(s[k]) for k in ("exn", "dzz", "rhodj", "rhodref", "exn_ref", "pabs")],
        _p(io["ci"]), _p(s["cldfr"]),
        *[_p(io[k]) for k in ("hlc_hrc", "hlc_hcf", "hli_hri", "hli_hcf")],
        *[_p(s[k]) for k in ("tht", "rvt", "rct", "rrt", "rit", "rst", "rgt", "rht", "sigs")],
        *[_p(io[k]) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "rhs")],
        *[_p(col[k]) for k in ("inprc", "inprr", "inpri", "inprs", "inprg", "inprh")],
        _p(o2["evap3d"]), _p(o2["rainfr"]))
    io.update(col)
    io.update(o2)
    return io


def pipeline_state_hail(f: dict, adj: dict, dt: float = 50.0) -> dict:
    """rain_ice inputs of the fused step with hail: pipeline_state plus the hail state and
    source (rht = rhs * dt, as for the other species)."""
    s = pipeline_state(f, adj, dt)
    s.update(rht=f["rhs"] * dt, rhs=f["rhs"])
    return s


def pipeline_state_ocnd2(f: dict, adj: dict, dt: float = 50.0) -> dict:
    """rain_ice inputs of the fused step with OCND2: pipeline_state plus ICE_ADJUST's OCND2
    outputs (adj from ice_adjust_ocnd2)."""
    s = pipeline_state(f, adj, dt)
    s.update({k: adj[k] for k in ("icldfr", "ssio", "ssiu", "ifr")})
    return s


def pipeline_state(f: dict, adj: dict, dt: float = 50.0) -> dict:
    """rain_ice inputs for the fused ice_adjust + rain_ice workload.

    The t-state handed to rain_ice is the post-adjustment S-variables times dt
    (the synthetic sources carry no external forcing, rxs = rx/dt).  The POUT_*
    adjusted state is not used: its mass-flux limiter (micro/ice_adjust.F90:449-466)
    divides 0/0 when CB02 condenses more than the total water and rc_mf = ri_mf = 0,
    whereas the S-variables are limited (:313-345) and stay finite.
    """
    return {
        "exn": f["exn"], "dzz": f["dzz"], "rhodj": f["rhodj"], "rhodref": f["rho_dry_ref"],
        "exn_ref": f["exn_ref"], "pabs": f["pabs"], "ci": f["ci"], "cldfr": adj["cldfr"],
        "hlc_hrc": adj["hlc_hrc"], "hlc_hcf": adj["hlc_hcf"], "hli_hri": adj["hli_hri"],
        "hli_hcf": adj["hli_hcf"], "tht": adj["ths"] * dt, "rvt": adj["rvs"] * dt,
        "rct": adj["rcs"] * dt, "rrt": f["rrs"] * dt, "rit": adj["ris"] * dt, "rst": f["rss"] * dt,
        "rgt": f["rgs"] * dt, "sigs": f["sigs"],
        "ths": adj["ths"], "rvs": adj["rvs"], "rcs": adj["rcs"], "rrs": f["rrs"], "ris": adj["ris"],
        "rss": f["rss"], "rgs": f["rgs"],
    }
