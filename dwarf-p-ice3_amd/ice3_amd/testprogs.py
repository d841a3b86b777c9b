"""Reader for the PHYEX testprogs reproduction datasets (``data/ice_adjust.nc`` and
``data/rain_ice.nc`` of the reference checkout), mapped onto the drop-ins' "columns" layout.

Layout, as the reference's own tests read these files
(tests/components/test_ice_adjust_fortran.py:227-262, tests/components/test_ice_adjust_jax.py:449-520,
tests/components/test_rain_ice_fortran.py:46-128, docs/agent/data_ice_adjust.md):

* point fields are ``(ngpblks, nflevg, nproma)``, per-column fields ``(ngpblks, nproma)``;
* species fields are ``(ngpblks, nspecies, nflevg, nproma)``.  ice_adjust: ``ZRS`` = (th, rv, rc,
  rr, ri, rs, rg) at t, ``PRS`` = sources (rv, rc, rr, ri, rs, rg), ``PRS_OUT`` the same sources
  after the adjustment, ``PTHS`` the theta source (updated in place by the Fortran).
  rain_ice: ``PRT`` = t-state (rv .. rg), ``PRS`` = sources, ``PRS_OUT`` = (ths, rvs, .., rgs);
* column ``c = block * nproma + jproma`` (the tests' reshape of ``(ngpblks, nproma)``);
* the Exner function at t is ``PEXNREF`` (the files carry no other).

Vertical order.  The kernels number levels from the ground up.  A dataset's order is read from
its pressure: when the first level has the lower mean pressure it runs top-down (PHYEX's AROME
arrays do) and every point field is flipped.  ice_adjust is pointwise, so the order only matters
for rain_ice, whose sedimentation runs down the column.

Format.  scipy's netCDF reader covers the classic formats (CDF-1, CDF-2).  A netCDF-4 file is
HDF5 underneath and needs an HDF5 reader, which this environment lacks; loading one raises with
the conversion to use (``nccopy -k classic``).
"""
from __future__ import annotations

import os

import numpy as np

# Tolerances of the reference's reproduction tests: (atol, rtol)
# (tests/components/test_ice_adjust_jax.py:567-630, tests/components/test_rain_ice_fortran.py:215-224).
TOLERANCES = {"ths": (1e-5, 1e-4), "rvs": (1e-5, 1e-4), "rcs": (1e-5, 1e-4), "rrs": (1e-5, 1e-4),
              "ris": (1e-5, 1e-4), "rss": (1e-5, 1e-4), "rgs": (1e-5, 1e-4), "cldfr": (1e-3, 1e-2),
              "hlc_hrc": (1e-5, 1e-4), "hlc_hcf": (1e-3, 1e-2), "hli_hri": (1e-5, 1e-4),
              "hli_hcf": (1e-3, 1e-2), "inprr": (1e-5, 1e-4), "inprs": (1e-5, 1e-4),
              "inprg": (1e-5, 1e-4), "inprc": (1e-5, 1e-4)}

_ICE_ADJUST_INPUTS = ("PPABSM", "PEXNREF", "PRHODREF", "PSIGS", "PCF_MF", "PRC_MF", "PRI_MF",
                      "PTHS", "ZRS", "PRS", "ZSIGQSAT")
_RAIN_ICE_INPUTS = ("PEXNREF", "PDZZ", "PRHODJ", "PRHODREF", "PPABSM", "PCLDFR", "PTHT", "PRT",
                    "PHLC_HRC", "PHLC_HCF", "PHLI_HRI", "PHLI_HCF", "PCIT", "PSIGS", "PTHS", "PRS")


class Dataset:
    """Variables (float64 numpy arrays), dimension sizes and global attributes of one file."""

    def __init__(self, variables: dict, dims: dict, attrs: dict):
        self.variables, self.dims, self.attrs = variables, dims, attrs

    def __contains__(self, name):
        return name in self.variables

    def __getitem__(self, name):
        return self.variables[name]


def read(path: str) -> Dataset:
    """Load a classic-format netCDF file completely into memory."""
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] != b"CDF":
        if magic == b"\x89HDF":
            raise RuntimeError(f"{path}: netCDF-4 (HDF5) file; no HDF5 reader is installed here. "
                               "Convert it with `nccopy -k classic` first.")
        raise RuntimeError(f"{path}: not a netCDF file")
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        variables = {k: np.array(v.data, dtype=np.float64) for k, v in nc.variables.items()}
        dims = {k: int(v) for k, v in nc.dimensions.items() if v is not None}
        attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in nc._attributes.items()}
    return Dataset(variables, dims, attrs)


def _require(ds: Dataset, names):
    missing = [n for n in names if n not in ds]
    if missing:
        raise KeyError(f"dataset lacks {', '.join(missing)}")


def top_down(ds: Dataset) -> bool:
    """True when level 0 of the file is the top of the column (lowest mean pressure)."""
    p = ds["PPABSM"]
    return float(p[:, 0, :].mean()) < float(p[:, -1, :].mean())


def _columns(a: np.ndarray, flip: bool) -> np.ndarray:
    """(ngpblks, nflevg, nproma) -> (nlev, ncol), level 0 the lowest."""
    nb, nl, npr = a.shape
    c = np.ascontiguousarray(np.transpose(a, (1, 0, 2)).reshape(nl, nb * npr))
    return np.ascontiguousarray(c[::-1]) if flip else c


def _species(a: np.ndarray, j: int, flip: bool) -> np.ndarray:
    return _columns(a[:, j], flip)


def _column_field(a: np.ndarray) -> np.ndarray:
    """(ngpblks, nproma) -> (ncol,)"""
    return np.ascontiguousarray(a.reshape(-1))


def timestep(ds: Dataset, default: float = 50.0) -> float:
    return float(ds.attrs.get("PTSTEP", default))


def ice_adjust_case(ds: Dataset):
    """Return ``(args, timestep, expected)``: ``args`` are IceAdjust's keyword arguments as
    (nlev, ncol) arrays (``sigqsat`` per column), ``expected`` the file's results that exist
    among rvs, rcs, ris, ths, cldfr, hlc_hrc, hlc_hcf, hli_hri, hli_hcf, srcs."""
    _require(ds, _ICE_ADJUST_INPUTS)
    flip = top_down(ds)
    col = lambda n: _columns(ds[n], flip)  # noqa: E731
    zrs, prs = ds["ZRS"], ds["PRS"]
    args = dict(sigqsat=_column_field(ds["ZSIGQSAT"]), pabs=col("PPABSM"), sigs=col("PSIGS"),
                th=_species(zrs, 0, flip), exn=col("PEXNREF"), exn_ref=col("PEXNREF"),
                rho_dry_ref=col("PRHODREF"), rv=_species(zrs, 1, flip), rc=_species(zrs, 2, flip),
                ri=_species(zrs, 4, flip), rr=_species(zrs, 3, flip), rs=_species(zrs, 5, flip),
                rg=_species(zrs, 6, flip), cf_mf=col("PCF_MF"), rc_mf=col("PRC_MF"),
                ri_mf=col("PRI_MF"), rvs=_species(prs, 0, flip), rcs=_species(prs, 1, flip),
                ris=_species(prs, 3, flip), ths=col("PTHS"))
    expected = {}
    if "PRS_OUT" in ds:
        out = ds["PRS_OUT"]
        expected.update(rvs=_species(out, 0, flip), rcs=_species(out, 1, flip),
                        ris=_species(out, 3, flip))
    for key, name in (("ths", "PTHS_OUT"), ("cldfr", "PCLDFR_OUT"), ("hlc_hrc", "PHLC_HRC_OUT"),
                      ("hlc_hcf", "PHLC_HCF_OUT"), ("hli_hri", "PHLI_HRI_OUT"),
                      ("hli_hcf", "PHLI_HCF_OUT"), ("srcs", "PSRCS_OUT")):
        if name in ds:
            expected[key] = col(name)
    return args, timestep(ds), expected


def rain_ice_case(ds: Dataset):
    """Return ``(state, dt, expected)``: ``state`` is RainIce's state dict (RainIceJAX names)
    as (nlev, ncol) arrays, ``expected`` the file's results that exist among ths .. rgs
    (``PRS_OUT``) and the surface precipitation inprc, inprr, inprs, inprg (``PINPR*_OUT``)."""
    _require(ds, _RAIN_ICE_INPUTS)
    flip = top_down(ds)
    col = lambda n: _columns(ds[n], flip)  # noqa: E731
    prt, prs = ds["PRT"], ds["PRS"]
    state = dict(exn=col("PEXNREF"), exnref=col("PEXNREF"), dzz=col("PDZZ"), rhodj=col("PRHODJ"),
                 rhodref=col("PRHODREF"), pres=col("PPABSM"), cf=col("PCLDFR"), sigs=col("PSIGS"),
                 hlc_hrc=col("PHLC_HRC"), hlc_hcf=col("PHLC_HCF"), hli_hri=col("PHLI_HRI"),
                 hli_hcf=col("PHLI_HCF"), ci_t=col("PCIT"), th_t=col("PTHT"), ths=col("PTHS"))
    for j, s in enumerate(("rv", "rc", "rr", "ri", "rs", "rg")):
        state[f"{s}_t"] = _species(prt, j, flip)
        state[f"{s}s"] = _species(prs, j, flip)
    expected = {}
    if "PRS_OUT" in ds:
        out = ds["PRS_OUT"]
        for j, k in enumerate(("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")):
            expected[k] = _species(out, j, flip)
    for k in ("inprc", "inprr", "inprs", "inprg"):
        name = f"P{k.upper()}_OUT"
        if name in ds:
            expected[k] = _column_field(ds[name])
    return state, timestep(ds), expected


def find(name: str, directory: str | None = None) -> str | None:
    """Path of a testprogs file (``ice_adjust.nc``, ``rain_ice.nc``) in ``directory`` or
    ``$ICE3_TESTPROGS_DATA``, or None when it is not there."""
    d = directory or os.environ.get("ICE3_TESTPROGS_DATA")
    if not d:
        return None
    p = os.path.join(d, name)
    return p if os.path.isfile(p) else None
