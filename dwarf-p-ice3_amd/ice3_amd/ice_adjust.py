"""IceAdjust: drop-in for the reference's IceAdjustJAX (src/ice3/jax/ice_adjust.py:40-230).

Same call signature, argument names and 16-tuple result
``(t, rv_out, rc_out, ri_out, cldfr, hlc_hrc, hlc_hcf, hli_hri, hli_hcf, cph, lv, ls,
rvs, rcs, ris, ths)``; the arithmetic is PHYEX ICE_ADJUST (micro/ice_adjust.F90) with
the sigma_s subgrid scheme, evaluated by the HIP kernel ``k_ice_adjust``.

Semantics follow the Fortran, where the JAX stencil diverges from it:
* ``t, rv_out, rc_out, ri_out`` are ICE_ADJUST's POUT_* adjusted state including the
  mass-flux condensate (micro/ice_adjust.F90:449-472).  Like the Fortran, they are NaN
  where CB02 condenses more than the total water while rc_mf = ri_mf = 0 (the
  limiter divides 0/0 there); the sources rvs..ths are limited and stay finite.
* ``sigqsat`` is per column, as PSIGQSAT(NIJT); a (..., nz) array is accepted and its
  lowest level is used.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._fields import as_columns, check_layout, from_columns, ptr, resolve_layout, stream_ptr

_PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}
_OUT_ORDER = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri",
              "hli_hcf", "cph", "lv", "ls", "rvs", "rcs", "ris", "ths")


# The switches of a Phyex object that the kernels implement in one setting only, with the
# value the pinned configuration needs (the reference's enumerations:
# src/ice3/phyex_common/nebn.py:6-46 and ice_parameters.py:10-68).  A Phyex("AROME") object
# carries these values (nebn.py:261-266, ice_parameters.py:408-481); a changed one raises.
# LSEDIC and LCRIAUTI are not checked: the Python AROME defaults set them (ice_parameters.py:412,
# 454) where PHYEX's Fortran AROME configuration, which the kernels follow, does not.
_PINNED = {
    "nebn": {"LSUBG_COND": 1, "LSIGMAS": 1, "LSTATNW": 0, "LHGT_QS": 0, "CONDENS": 0,
             "LAMBDA3": 0, "FRAC_ICE_ADJUST": 0},
    "param_icen": {"LCOND2": 0, "SUBG_MF_PDF": 1, "LRED": 0, "MRSTEP": 0.0,
                   "TSTEP_TS": 0.0, "LWARM": 1, "LFEEDBACKT": 1, "LEVLIMIT": 1, "LNULLWETG": 1,
                   "LWETGPOST": 1, "LCRFLIMIT": 1, "LSNOW_T": 0, "LSEDIM_AFTER": 0, "LDEPOSC": 0,
                   "SNOW_RIMING": 0, "SUBG_RC_RR_ACCR": 0, "SUBG_RR_EVAP": 0, "SUBG_PR_PDF": 0,
                   "SUBG_AUCV_RC": 1, "SUBG_AUCV_RI": 0, "PRISTINE_ICE": "PLAT"},
}


def sedimentation_of(cfg) -> tuple[str, float]:
    """The sedimentation scheme a configuration object selects: its ``param_icen.SEDIM``
    (the reference's Sedim enum, src/ice3/phyex_common/ice_parameters.py:62-64: 0 = SPLI,
    1 = STAT) and ``param_icen.SPLIT_MAXCFL``; ("STAT", 0.8) without them (AROME)."""
    p = getattr(cfg, "param_icen", None)
    sedim = getattr(p, "SEDIM", 1)
    sedim = getattr(sedim, "value", sedim)
    if sedim not in (0, 1):
        raise ValueError(f"param_icen.SEDIM = {sedim!r}: 0 (SPLI) or 1 (STAT)")
    return ("SPLI" if sedim == 0 else "STAT"), float(getattr(p, "SPLIT_MAXCFL", 0.8))


def check_configuration(cfg, what: str = "phyex", ocnd2: bool = False):
    """Accept the configurations that mean the pinned AROME ICE3 one: None, "AROME", or a
    configuration object with ``program == "AROME"`` (the reference's ``Phyex("AROME")``,
    src/ice3/phyex_common/phyex.py:20-47).  When the object carries ``nebn`` and
    ``param_icen``, every switch the kernels implement in one setting only must hold that
    setting (``LCOND2`` = 1 for the OCND2 entry point, 0 otherwise).  Anything else raises
    ValueError."""
    prog = getattr(cfg, "program", cfg)
    if not (prog is None or (isinstance(prog, str) and prog.upper() == "AROME")):
        raise ValueError(f"{what}: only the pinned AROME ICE3 configuration is implemented, got {prog!r}")
    sedimentation_of(cfg)  # SEDIM: either scheme (rain_ice reads it)
    for group, fields in _PINNED.items():
        obj = getattr(cfg, group, None)
        if obj is None:
            continue
        for name, want in fields.items():
            if name == "LCOND2":
                want = 1 if ocnd2 else 0
            have = getattr(obj, name, want)
            if have != want:
                raise ValueError(f"{what}: {group}.{name} = {have!r}; the kernels implement "
                                 f"{want!r} only (pinned AROME ICE3 configuration)")


class IceAdjust:
    """ICE_ADJUST on MI355X.

    Parameters
    ----------
    phyex : the configuration, kept for API compatibility: None, "AROME" or an AROME
        ``Phyex`` object; the kernels implement the pinned AROME ICE3 configuration (KRR=6,
        sigma_s subgrid condensation, CB02) and anything else raises.
    jit : accepted for API compatibility (kernels are compiled ahead of time).
    layout : "auto" (default): each call reads the layout from its arrays' rank, so that
        the JAX components' (nx, ny, nz) arrays (vertical index last) and the Fortran
        (nlev, ncol) memory (columns contiguous, no copy) are both taken as they are;
        "jax" or "columns" forces one.  Results come back in the layout of the inputs.
    device : CUDA device (default: current).

    The call takes IceAdjustJAX's arguments; ``check_inputs=True`` first counts the NaN and
    infinite input values on the device (``ice3_count_nonfinite``) and raises ValueError
    naming the fields that hold any (the reference does not check; off by default).
    """

    _OCND2 = False

    def __init__(self, phyex=None, jit: bool = True, layout: str = "auto", device=None):
        check_configuration(phyex, ocnd2=self._OCND2)
        self.layout = check_layout(layout)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self._lib = _lib.load()
        self._ctx = _lib.context(self.device.index)

    def _inputs(self, sigqsat, args, layout):
        """Columns-layout device tensors of the fields and the per-column sigqsat."""
        f = {k: as_columns(v, layout, device=self.device, name=k) for k, v in args.items()}
        nlev, ncol = f["th"].shape
        for k, v in f.items():
            if tuple(v.shape) != (nlev, ncol):
                raise ValueError(f"{k}: shape {tuple(v.shape)} != {(nlev, ncol)}")
        sq = self._per_column(sigqsat, "sigqsat", ncol, layout)
        return f, sq, nlev, ncol

    def _per_column(self, x, name, ncol, layout):
        t = torch.as_tensor(x, dtype=torch.float64, device=self.device)
        if t.dim() > 1 and t.numel() != ncol:  # a full field: per column, take the lowest level
            t = as_columns(t, layout, device=self.device, name=name)[0]
        t = t.reshape(-1).contiguous()
        if t.numel() != ncol:
            raise ValueError(f"{name}: {t.numel()} values for {ncol} columns")
        return t

    def __call__(self, sigqsat, pabs, sigs, th, exn, exn_ref, rho_dry_ref, rv, rc, ri, rr, rs, rg,
                 cf_mf, rc_mf, ri_mf, rvs, rcs, ris, ths, timestep: float,
                 precision: str = "fp64", return_srcs: bool = False, check_inputs: bool = False):
        args = dict(pabs=pabs, sigs=sigs, th=th, exn=exn, exn_ref=exn_ref, rho_dry_ref=rho_dry_ref,
                    rv=rv, rc=rc, ri=ri, rr=rr, rs=rs, rg=rg, cf_mf=cf_mf, rc_mf=rc_mf, ri_mf=ri_mf,
                    rvs=rvs, rcs=rcs, ris=ris, ths=ths)
        shape = tuple(th.shape) if hasattr(th, "shape") else tuple(torch.as_tensor(th).shape)
        layout = resolve_layout(self.layout, th)
        f, sq, nlev, ncol = self._inputs(sigqsat, args, layout)
        if check_inputs:
            _lib.require_finite({**f, "sigqsat": sq}, self.device)
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) for k in _OUT_ORDER}
        srcs = torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) if return_srcs else None
        cin = _lib.AdjustIn(*[ptr(f[k]) for k in _lib.ADJ_IN_FIELDS[:-1]], ptr(sq))
        cout = _lib.AdjustOut(*[ptr(out[k]) for k in _lib.ADJ_OUT_FIELDS[:-1]], ptr(srcs))
        with torch.cuda.device(self.device):
            rc_ = self._lib.ice3_ice_adjust(self._ctx, _PREC[precision], ncol, nlev, ncol, float(timestep),
                                            ctypes.byref(cin), ctypes.byref(cout), stream_ptr(self.device))
        _lib.check(rc_, "ice3_ice_adjust")
        res = [out[k] for k in _OUT_ORDER]
        res = [from_columns(r, layout, shape) for r in res]
        if srcs is not None:
            srcs = from_columns(srcs, layout, shape)
        return (*res, srcs) if return_srcs else tuple(res)


class IceAdjustOCND2(IceAdjust):
    """ICE_ADJUST with PARAM_ICE LOCND2 = .TRUE. on MI355X (``ice3_ice_adjust_ocnd2``): the
    liquid-only condensation and the separate ice cloud of ICECLOUD over the TIWMX tables
    (micro/condensation.F90:263-265,373-394,582-637, micro/mode_icecloud.F90).

    The reference's JAX component pins OCND2 = False (src/ice3/jax/ice_adjust.py:116); this
    class follows the Fortran ICE_ADJUST called with LOCND2.  The call takes IceAdjust's
    arguments plus ``zz`` (level heights in m, increasing upward, PZZ) and ``ice_cld_wgt``
    (PICE_CLD_WGT, per column), in FP64, and returns IceAdjust's 16-tuple followed by
    ``icldfr, wcldfr, ssio, ssiu, ifr`` (PICLDFR, PWCLDFR, PSSIO, PSSIU, PIFR).  A ``phyex``
    object must have ``param_icen.LCOND2 = 1``.
    """

    _OCND2 = True

    def __call__(self, sigqsat, pabs, sigs, th, exn, exn_ref, rho_dry_ref, rv, rc, ri, rr, rs, rg,
                 cf_mf, rc_mf, ri_mf, rvs, rcs, ris, ths, timestep: float, zz=None,
                 ice_cld_wgt=None, return_srcs: bool = False, check_inputs: bool = False):
        if zz is None or ice_cld_wgt is None:
            raise ValueError("the OCND2 branch needs zz (level heights) and ice_cld_wgt")
        args = dict(pabs=pabs, sigs=sigs, th=th, exn=exn, exn_ref=exn_ref, rho_dry_ref=rho_dry_ref,
                    rv=rv, rc=rc, ri=ri, rr=rr, rs=rs, rg=rg, cf_mf=cf_mf, rc_mf=rc_mf, ri_mf=ri_mf,
                    rvs=rvs, rcs=rcs, ris=ris, ths=ths, zz=zz)
        shape = tuple(th.shape) if hasattr(th, "shape") else tuple(torch.as_tensor(th).shape)
        layout = resolve_layout(self.layout, th)
        f, sq, nlev, ncol = self._inputs(sigqsat, args, layout)
        wgt = self._per_column(ice_cld_wgt, "ice_cld_wgt", ncol, layout)
        if nlev == 1:
            raise ValueError("the OCND2 branch needs at least two levels")
        if check_inputs:
            _lib.require_finite({**f, "sigqsat": sq, "ice_cld_wgt": wgt}, self.device)
        names = _OUT_ORDER + _lib.OCND2_OUT_FIELDS
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) for k in names}
        srcs = torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) if return_srcs else None
        cin = _lib.AdjustIn(*[ptr(f[k]) for k in _lib.ADJ_IN_FIELDS[:-1]], ptr(sq))
        cin2 = _lib.Ocnd2In(ptr(f["zz"]), ptr(wgt))
        cout = _lib.AdjustOut(*[ptr(out[k]) for k in _lib.ADJ_OUT_FIELDS[:-1]], ptr(srcs))
        cout2 = _lib.Ocnd2Out(*[ptr(out[k]) for k in _lib.OCND2_OUT_FIELDS])
        with torch.cuda.device(self.device):
            rc_ = self._lib.ice3_ice_adjust_ocnd2(self._ctx, ncol, nlev, ncol, float(timestep),
                                                  ctypes.byref(cin), ctypes.byref(cin2), ctypes.byref(cout),
                                                  ctypes.byref(cout2), stream_ptr(self.device))
        _lib.check(rc_, "ice3_ice_adjust_ocnd2")
        res = [from_columns(out[k], layout, shape) for k in names]
        if srcs is not None:
            srcs = from_columns(srcs, layout, shape)
        return (*res, srcs) if return_srcs else tuple(res)


class IceAdjustHail(IceAdjust):
    """ICE_ADJUST with KRR = 7 (hail, HCLOUD='ICE4') on MI355X (``ice3_ice_adjust_hail``, FP64):
    the hail mixing ratio enters the heat capacity (micro/ice_adjust.F90:503-507).  The
    reference reaches it through the Fortran ICE_ADJUST with KRR=7 and PRH.  The call takes
    IceAdjust's arguments plus ``rh`` (hail mixing ratio at t) and returns IceAdjust's 16-tuple.
    """

    def __call__(self, sigqsat, pabs, sigs, th, exn, exn_ref, rho_dry_ref, rv, rc, ri, rr, rs, rg,
                 cf_mf, rc_mf, ri_mf, rvs, rcs, ris, ths, timestep: float, rh=None,
                 return_srcs: bool = False, check_inputs: bool = False):
        if rh is None:
            raise ValueError("KRR=7 needs the hail mixing ratio rh")
        args = dict(pabs=pabs, sigs=sigs, th=th, exn=exn, exn_ref=exn_ref, rho_dry_ref=rho_dry_ref,
                    rv=rv, rc=rc, ri=ri, rr=rr, rs=rs, rg=rg, cf_mf=cf_mf, rc_mf=rc_mf, ri_mf=ri_mf,
                    rvs=rvs, rcs=rcs, ris=ris, ths=ths, rh=rh)
        shape = tuple(th.shape) if hasattr(th, "shape") else tuple(torch.as_tensor(th).shape)
        layout = resolve_layout(self.layout, th)
        f, sq, nlev, ncol = self._inputs(sigqsat, args, layout)
        if check_inputs:
            _lib.require_finite({**f, "sigqsat": sq}, self.device)
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) for k in _OUT_ORDER}
        srcs = torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) if return_srcs else None
        cin = _lib.AdjustIn(*[ptr(f[k]) for k in _lib.ADJ_IN_FIELDS[:-1]], ptr(sq))
        cout = _lib.AdjustOut(*[ptr(out[k]) for k in _lib.ADJ_OUT_FIELDS[:-1]], ptr(srcs))
        with torch.cuda.device(self.device):
            rc_ = self._lib.ice3_ice_adjust_hail(self._ctx, ncol, nlev, ncol, float(timestep),
                                                 ctypes.byref(cin), ptr(f["rh"]), ctypes.byref(cout),
                                                 stream_ptr(self.device))
        _lib.check(rc_, "ice3_ice_adjust_hail")
        res = [from_columns(out[k], layout, shape) for k in _OUT_ORDER]
        if srcs is not None:
            srcs = from_columns(srcs, layout, shape)
        return (*res, srcs) if return_srcs else tuple(res)
