"""Seeded synthetic columns for the ice_adjust / rain_ice workloads.

Every field is a C-contiguous float64 array of shape ``(nlev, ncol)``: the
column index is contiguous, which is byte-identical to the reference's
Fortran ``(NIJT, NKT)`` arrays (``bridge/phyex_bridge.F90:96-118``).  Level 0
is the lowest model level (the reference bridge's ``NKL=1`` ordering,
``bridge/phyex_bridge.F90:128``).

Profile (the generator the benchmarks are quoted on):

* pressure ``p_k = 1000 hPa * (10/1000)**(k/(nlev-1))``;
* heights from the hypsometric equation on the undisturbed temperature
  profile ``T(z) = max(288.15 - 6.5e-3 z, 215)``, integrated upward from
  ``z_0 = 0`` with the temperature of the level below;
* ``T = T(z) + N(0, 1 K)`` per point; ``RH ~ U(0.3, 1.05)`` and
  ``rv = RH * qsat_w(T, p)`` with the reference's ``esat_w``
  (``micro/condensation.F90:407``);
* ``rc, ri, rr, rs, rg ~ LogNormal(ln 1e-5, 1)``, each zeroed independently
  with probability 0.6;
* ``sigma_s ~ U(0, 2e-4)``, ``sigqsat = 0.02`` per column, mass-flux inputs 0;
* ``exner = (p/p00)**(Rd/cpd)``, ``theta = T/exner``, reference exner equal
  to exner, dry density ``p * eps/(eps+rv) / (Rd T)``, ``rhodj = rhodref``;
* ``dzz_k = z_{k+1} - z_k`` (top level repeats the one below);
* ``ci = LogNormal(ln 1e5 m-3, 1)`` where ``ri > 0``, else 0;
* sources with no external forcing: ``rxs = rx/dt``, ``ths = theta/dt``.

The ``"pipeline"`` variant (the ice_adjust + rain_ice workload) adds, in 20 %
of the columns, one precipitating band of 10-30 contiguous levels where
``rr, rs, rg ~ U(1e-4, 5e-3)``; mixed-phase layers (253-273 K) are present in
every column because the lapse-rate profile crosses that range near 2.3-5.4 km.

The random stream is numpy's counter-based Philox bit generator, so a
``(seed, ncol, nlev)`` triple always gives the same columns.
"""
from __future__ import annotations

import numpy as np

# INI_CST closed forms (aux/mode_ini_cst.F90:135-157)
_AVOGADRO = 6.0221367e23
_BOLTZ = 1.380658e-23
_MD = 28.9644e-3
_MV = 18.0153e-3
RD = _AVOGADRO * _BOLTZ / _MD
RV = _AVOGADRO * _BOLTZ / _MV
EPSILO = _MV / _MD
CPD = 3.5 * RD
CPV = 4.0 * RV
CL = 4.218e3
TT = 273.16
LVTT = 2.5008e6
ESTT = 611.14
GAMW = (CL - CPV) / RV
BETAW = LVTT / RV + GAMW * TT
ALPW = np.log(ESTT) + BETAW / TT + GAMW * np.log(TT)
G = 9.80665
P00 = 1.0e5

FIELDS_ADJUST = (
    "pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr", "rs", "rg",
    "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths",
)


def level_profile(nlev: int):
    """Pressure (Pa), undisturbed temperature (K) and height (m) per level."""
    k = np.arange(nlev, dtype=np.float64)
    p = 1.0e5 * (10.0 / 1000.0) ** (k / (nlev - 1))
    t = np.empty(nlev)
    z = np.empty(nlev)
    z[0] = 0.0
    t[0] = 288.15
    for i in range(1, nlev):
        z[i] = z[i - 1] + RD * t[i - 1] / G * np.log(p[i - 1] / p[i])
        t[i] = max(288.15 - 6.5e-3 * z[i], 215.0)
    return p, t, z


def generate(ncol: int, nlev: int = 90, seed: int = 20260002, variant: str = "adjust",
             dt: float = 50.0) -> dict:
    """Return a dict of float64 ``(nlev, ncol)`` fields (plus ``sigqsat`` of shape ``(ncol,)``)."""
    if variant not in ("adjust", "pipeline"):
        raise ValueError(f"unknown variant {variant!r}")
    rng = np.random.Generator(np.random.Philox(seed))
    shape = (nlev, ncol)
    p1, t1, z1 = level_profile(nlev)
    p = np.broadcast_to(p1[:, None], shape).copy()
    t = t1[:, None] + rng.standard_normal(shape)
    rh = rng.uniform(0.3, 1.05, shape)
    esw = np.exp(ALPW - BETAW / t - GAMW * np.log(t))
    qsat = EPSILO * esw / (p - esw)
    rv = rh * qsat
    hyd = {}
    for name in ("rc", "ri", "rr", "rs", "rg"):
        val = rng.lognormal(np.log(1.0e-5), 1.0, shape)
        keep = rng.uniform(0.0, 1.0, shape) >= 0.6
        hyd[name] = np.where(keep, val, 0.0)
    if variant == "pipeline":
        precip = rng.uniform(0.0, 1.0, ncol) < 0.2
        width = rng.integers(10, 31, ncol)
        start = rng.integers(0, nlev, ncol)
        lev = np.arange(nlev)[:, None]
        band = precip[None, :] & (lev >= start[None, :]) & (lev < (start + width)[None, :])
        for name in ("rr", "rs", "rg"):
            val = rng.uniform(1.0e-4, 5.0e-3, shape)
            hyd[name] = np.where(band, val, hyd[name])
    sigs = rng.uniform(0.0, 2.0e-4, shape)
    ci_val = rng.lognormal(np.log(1.0e5), 1.0, shape)
    exn = (p / P00) ** (RD / CPD)
    th = t / exn
    rhodref = p * EPSILO / (EPSILO + rv) / (RD * t)
    dz1 = np.empty(nlev)
    dz1[:-1] = np.diff(z1)
    dz1[-1] = dz1[-2]
    f = {
        "pabs": p, "sigs": sigs, "th": th, "exn": exn, "exn_ref": exn.copy(),
        "rho_dry_ref": rhodref, "rv": rv, "rc": hyd["rc"], "ri": hyd["ri"], "rr": hyd["rr"],
        "rs": hyd["rs"], "rg": hyd["rg"],
        "cf_mf": np.zeros(shape), "rc_mf": np.zeros(shape), "ri_mf": np.zeros(shape),
        "rvs": rv / dt, "rcs": hyd["rc"] / dt, "ris": hyd["ri"] / dt, "ths": th / dt,
        "rrs": hyd["rr"] / dt, "rss": hyd["rs"] / dt, "rgs": hyd["rg"] / dt,
        "rhodj": rhodref.copy(),
        "dzz": np.broadcast_to(dz1[:, None], shape).copy(),
        "ci": np.where(hyd["ri"] > 0.0, ci_val, 0.0),
        "sigqsat": np.full(ncol, 0.02),
    }
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}
