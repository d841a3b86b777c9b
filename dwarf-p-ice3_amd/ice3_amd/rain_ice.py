"""RainIce / AdjustRainIce: drop-ins for the reference's RainIceJAX (src/ice3/jax/rain_ice.py:20-420).

``RainIce(constants)(state, dt)`` returns ``(state, diagnostics)`` like RainIceJAX: the
state dict's sources ``ths, rvs, rcs, rrs, ris, rss, rgs`` and ``ci_t`` are replaced by
their updated values, ``evap3d``/``rainfr`` are added, and the diagnostics hold the
surface precipitation ``inprc, inprr, inpri, inprs, inprg``.  Arithmetic is PHYEX
RAIN_ICE (micro/rain_ice.F90) with the sedimentation before the microphysics, the
per-point ICE4 time splitting, the total tendencies and the negativity correction,
evaluated by the HIP kernels ``k_sedim`` (the statistical sedimentation, one column per
thread) or ``k_sedim_split`` (``sedimentation="SPLI"``: the split scheme, one column per
wavefront) and ``k_micro`` (the persistent per-point microphysics).

State keys read (RainIceJAX names): exn, rhodref, pres, dzz, th_t, rv_t, rc_t, rr_t, ri_t,
rs_t, rg_t, ci_t, ths, rvs, rcs, rrs, ris, rss, rgs; optional: exnref (default exn),
rhodj (default rhodref), cf (cloud fraction, default 0), sigs (default 0), hlc_hrc,
hlc_hcf, hli_hri, hli_hcf (default 0).

Layouts: with ``layout="auto"`` (the default) each call takes RainIceJAX's
``(ngpblks, nproma, nflevg)`` arrays (vertical index last; any rank of 3 or more) or the
Fortran ``(nlev, ncol)`` memory (columns contiguous, no copy), read from the rank of
``th_t``; results come back in the same layout, and the per-column diagnostics
``inpr*`` with the shape of the inputs without their vertical axis, as RainIceJAX
returns them.

Where the returned state differs from RainIceJAX's: RainIceJAX overwrites the sources
with the microphysics tendency alone and returns the time-split, negativity-corrected
t-state in ``th_t..rg_t`` (src/ice3/jax/rain_ice.py:347-373).  Here the semantics are the
Fortran RAIN_ICE's: the sources are the input sources plus the sedimentation and the
microphysics tendencies, corrected for negativities (micro/rain_ice.F90:754-823), and
``th_t..rg_t``, ``exn``, ``rhodref`` and the other inputs are returned as they were given.
The state at t + dt is ``x_t + dt * xs``.

``check_inputs=True`` (RainIce, AdjustRainIce and IceAdjust) first counts the NaN and
infinite input values on the device (``ice3_count_nonfinite``) and raises ValueError
naming the fields that hold any; the reference has no such check, so it is off by
default.
"""
from __future__ import annotations

import contextlib
import ctypes

import torch

from . import _lib
from ._fields import as_columns, check_layout, from_columns, ptr, resolve_layout, stream_ptr
from .ice_adjust import check_configuration, sedimentation_of

_PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}


class _Base:
    """sedimentation: "STAT" (ICE4_SEDIMENTATION_STAT, the AROME configuration) or "SPLI"
    (ICE4_SEDIMENTATION_SPLIT with CFL bound split_max_cfl, then the negativity correction);
    by default what the configuration object selects (param_icen.SEDIM and SPLIT_MAXCFL),
    else "STAT"."""

    def __init__(self, constants=None, layout: str = "auto", device=None,
                 sedimentation: str | None = None, split_max_cfl: float | None = None):
        sedim, cfl = "STAT", 0.8
        if not isinstance(constants, dict):  # a constants dict is accepted and not read
            check_configuration(constants, "constants")
            sedim, cfl = sedimentation_of(constants)
        self.sedimentation = (sedimentation or sedim).upper()
        self.split_max_cfl = float(cfl if split_max_cfl is None else split_max_cfl)
        if self.sedimentation not in _lib.SEDIM or not self.split_max_cfl > 0.0:
            raise ValueError(f"sedimentation {sedimentation!r} / split_max_cfl {split_max_cfl!r}: "
                             "\"STAT\" or \"SPLI\" with a CFL bound > 0")
        self.layout = check_layout(layout)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self._lib = _lib.load()
        self._ctx = _lib.context(self.device.index)
        self.last_counters = None

    @contextlib.contextmanager
    def _scheme(self):
        """The context's sedimentation scheme set to this component's for its launches (the
        context is shared by the components of a device), then back to STAT."""
        _lib.check(self._lib.ice3_ctx_set_sedimentation(self._ctx, _lib.SEDIM[self.sedimentation],
                                                        self.split_max_cfl), "ice3_ctx_set_sedimentation")
        try:
            yield
        finally:
            self._lib.ice3_ctx_set_sedimentation(self._ctx, _lib.SEDIM["STAT"], 0.8)

    def _col(self, x, name, layout):
        return as_columns(x, layout, device=self.device, name=name)

    @staticmethod
    def _shape(x):
        return tuple(x.shape) if hasattr(x, "shape") else tuple(torch.as_tensor(x).shape)

    def _stats(self, n, with_calls):
        counters = torch.zeros(3, dtype=torch.int64, device=self.device)
        calls = torch.zeros(n, dtype=torch.int32, device=self.device) if with_calls else None
        return counters, calls, _lib.Stats(ptr(calls), ptr(counters))

    def _finish_counters(self, counters):
        c = counters.cpu().tolist()
        self.last_counters = {"tendency_calls": c[0], "capped_points": c[1], "micro_points": c[2]}
        if c[1]:
            raise _lib.Ice3Error(f"{c[1]} points hit the ICE4 sub-step cap ({_lib.load().ice3_strerror(-3).decode()})")


class RainIce(_Base):
    """RAIN_ICE on MI355X (see module docstring)."""

    _ENTRY = "ice3_rain_ice"

    def _extra_inputs(self, state, layout, nlev, ncol):
        return {}

    def _extra_results(self, extra, new, diags, layout, shape):
        pass

    def _launch(self, precision, ncol, nlev, dt, cin, extra, cio, cout, st):
        return self._lib.ice3_rain_ice(self._ctx, _PREC[precision], ncol, nlev, ncol, float(dt),
                                       ctypes.byref(cin), ctypes.byref(cio), ctypes.byref(cout),
                                       ctypes.byref(st), stream_ptr(self.device))

    def __call__(self, state: dict, dt: float, lsedim_after: bool = False, ldeposc: bool = False,
                 ldsoft: bool = False, lfeedbackt: bool = True, max_iterations: int | None = None,
                 precision: str = "fp64", return_calls: bool = False, check_inputs: bool = False):
        if lsedim_after or ldeposc or ldsoft or not lfeedbackt:
            raise ValueError("only LSEDIM_AFTER=F, LDEPOSC=F, LFEEDBACKT=T are implemented")
        shape = self._shape(state["th_t"])
        layout = resolve_layout(self.layout, state["th_t"])
        g = lambda k: self._col(state[k], k, layout)  # noqa: E731
        th = g("th_t")
        nlev, ncol = th.shape
        z = torch.zeros((nlev, ncol), dtype=torch.float64, device=self.device)
        opt = lambda k, d: self._col(state[k], k, layout) if k in state else d  # noqa: E731
        exn = g("exn")
        rhodref = g("rhodref")
        inp = dict(exn=exn, dzz=g("dzz"), rhodj=opt("rhodj", rhodref), rhodref=rhodref,
                   exn_ref=opt("exnref", exn), pabs=g("pres"), cldfr=opt("cf", z),
                   hlc_hrc=opt("hlc_hrc", z), hlc_hcf=opt("hlc_hcf", z), hli_hri=opt("hli_hri", z),
                   hli_hcf=opt("hli_hcf", z), sigs=opt("sigs", z), tht=th, rvt=g("rv_t"), rct=g("rc_t"),
                   rrt=g("rr_t"), rit=g("ri_t"), rst=g("rs_t"), rgt=g("rg_t"))
        # sources and ci are updated in place on private copies
        io = {k: g(k).clone() for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")}
        io["ci"] = g("ci_t").clone()
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device) for k in ("evap3d", "rainfr")}
        out.update({k: torch.empty(ncol, dtype=torch.float64, device=self.device)
                    for k in ("inprc", "inprr", "inpri", "inprs", "inprg")})
        for k, v in list(inp.items()) + list(io.items()):
            if tuple(v.shape) != (nlev, ncol):
                raise ValueError(f"{k}: shape {tuple(v.shape)} != {(nlev, ncol)}")
        if check_inputs:
            _lib.require_finite({**inp, **io}, self.device)
        extra = self._extra_inputs(state, layout, nlev, ncol)
        if check_inputs:
            _lib.require_finite(extra, self.device)
        counters, calls, st = self._stats(nlev * ncol, return_calls)
        cin = _lib.RainIn(*[ptr(inp[k]) for k in _lib.RAIN_IN_FIELDS])
        cio = _lib.RainIO(*[ptr(io[k]) for k in _lib.RAIN_IO_FIELDS])
        cout = _lib.RainOut(*[ptr(out[k]) for k in _lib.RAIN_OUT_FIELDS])
        with torch.cuda.device(self.device), self._scheme():
            rc = self._launch(precision, ncol, nlev, dt, cin, extra, cio, cout, st)
        _lib.check(rc, self._ENTRY)
        self._finish_counters(counters)
        new = dict(state)
        for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
            new[k] = from_columns(io[k], layout, shape)
        new["ci_t"] = from_columns(io["ci"], layout, shape)
        new["evap3d"] = from_columns(out["evap3d"], layout, shape)
        new["rainfr"] = from_columns(out["rainfr"], layout, shape)
        diags = {k: from_columns(out[k], layout, shape) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")}
        self._extra_results(extra, new, diags, layout, shape)
        if return_calls:
            diags["calls"] = from_columns(calls.view(nlev, ncol), layout, shape)
        return new, diags


class AdjustRainIce(_Base):
    """ICE_ADJUST followed by RAIN_ICE in one device pass (the fused time step).

    Inputs are the IceAdjust arguments plus ``rhodj, dzz, ci, rrs, rss, rgs``; the
    rain_ice t-state is the adjusted sources times dt.  Returns a dict with the final
    sources, ``cldfr, hlc_*, hli_*, ci, evap3d, rainfr`` and the surface precipitation.
    """

    def __call__(self, f: dict, dt: float, precision: str = "fp64", return_calls: bool = False,
                 check_inputs: bool = False):
        return self._fused(f, dt, precision, return_calls, check_inputs)

    def _fused(self, f, dt, precision, return_calls, check_inputs, ocnd2=False, hail=False):
        shape = self._shape(f["th"])
        layout = resolve_layout(self.layout, f["th"])
        g = {k: self._col(f[k], k, layout) for k in _lib.PIPE_IN_FIELDS if k != "sigqsat"}
        nlev, ncol = g["th"].shape
        sq = self._per_column(f["sigqsat"], "sigqsat", ncol, layout)
        extra = {}
        if ocnd2:
            extra["zz"] = self._col(f["zz"], "zz", layout)
            wgt = self._per_column(f["ice_cld_wgt"], "ice_cld_wgt", ncol, layout)
            if nlev < 2:
                raise ValueError("the OCND2 branch needs at least two levels")
        if hail:
            extra["rh"] = self._col(f["rh"], "rh", layout)
            extra["rhs"] = self._col(f["rhs"], "rhs", layout)
        for k, v in list(g.items()) + list(extra.items()):
            if tuple(v.shape) != (nlev, ncol):
                raise ValueError(f"{k}: shape {tuple(v.shape)} != {(nlev, ncol)}")
        if check_inputs:
            _lib.require_finite({**g, **extra, "sigqsat": sq, **({"ice_cld_wgt": wgt} if ocnd2 else {})},
                                self.device)
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device)
               for k in _lib.PIPE_OUT_FIELDS[:15]}
        out.update({k: torch.empty(ncol, dtype=torch.float64, device=self.device)
                    for k in _lib.PIPE_OUT_FIELDS[15:]})
        counters, calls, st = self._stats(nlev * ncol, return_calls)
        cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS[:-1]], ptr(sq))
        cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
        with torch.cuda.device(self.device), self._scheme():
            if ocnd2:
                x = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.device)
                     for k in _lib.OCND2_OUT_FIELDS}
                cin2 = _lib.Ocnd2In(ptr(extra["zz"]), ptr(wgt))
                cout2 = _lib.Ocnd2Out(*[ptr(x[k]) for k in _lib.OCND2_OUT_FIELDS])
                rc = self._lib.ice3_adjust_rain_ice_ocnd2(self._ctx, ncol, nlev, ncol, float(dt),
                                                          ctypes.byref(cin), ctypes.byref(cin2),
                                                          ctypes.byref(cout), ctypes.byref(cout2),
                                                          ctypes.byref(st), stream_ptr(self.device))
                out.update(x)
            elif hail:
                out["rhs"] = torch.empty((nlev, ncol), dtype=torch.float64, device=self.device)
                out["inprh"] = torch.empty(ncol, dtype=torch.float64, device=self.device)
                h = _lib.PipeHail(ptr(extra["rh"]), ptr(extra["rhs"]), ptr(out["rhs"]), ptr(out["inprh"]))
                rc = self._lib.ice3_adjust_rain_ice_hail(self._ctx, ncol, nlev, ncol, float(dt),
                                                         ctypes.byref(cin), ctypes.byref(cout), ctypes.byref(h),
                                                         ctypes.byref(st), stream_ptr(self.device))
            else:
                rc = self._lib.ice3_adjust_rain_ice(self._ctx, _PREC[precision], ncol, nlev, ncol, float(dt),
                                                    ctypes.byref(cin), ctypes.byref(cout), ctypes.byref(st),
                                                    stream_ptr(self.device))
        _lib.check(rc, "ice3_adjust_rain_ice_ocnd2" if ocnd2 else
                   ("ice3_adjust_rain_ice_hail" if hail else "ice3_adjust_rain_ice"))
        self._finish_counters(counters)
        res = {k: from_columns(v, layout, shape) for k, v in out.items()}
        if return_calls:
            res["calls"] = from_columns(calls.view(nlev, ncol), layout, shape)
        return res

    def _per_column(self, x, name, ncol, layout):
        t = torch.as_tensor(x, dtype=torch.float64, device=self.device)
        if t.dim() > 1 and t.numel() != ncol:  # a full field: per column, the lowest level
            t = self._col(t, name, layout)[0]
        t = t.reshape(-1).contiguous()
        if t.numel() != ncol:
            raise ValueError(f"{name}: {t.numel()} values for {ncol} columns")
        return t


class RainIceOCND2(RainIce):
    """RAIN_ICE with PARAM_ICE LOCND2 = .TRUE. on MI355X (``ice3_rain_ice_ocnd2``, FP64).

    The reference reaches this branch through the Fortran RAIN_ICE only (its JAX component
    has no OCND2).  The state takes RainIce's keys plus ICE_ADJUST's OCND2 outputs
    ``icldfr, ssio, ssiu, ifr`` (PICLDFR, PSSIO, PSSIU, PIFR: ``IceAdjustOCND2`` returns them);
    results as RainIce.  A ``constants`` object must have ``param_icen.LCOND2 = 1``.
    """

    _ENTRY = "ice3_rain_ice_ocnd2"

    def __init__(self, constants=None, layout: str = "auto", device=None,
                 sedimentation: str | None = None, split_max_cfl: float | None = None):
        sedim, cfl = None, None
        if not isinstance(constants, dict):
            check_configuration(constants, "constants", ocnd2=True)
            sedim, cfl = sedimentation_of(constants)
        super().__init__({}, layout, device, sedimentation or sedim,
                         cfl if split_max_cfl is None else split_max_cfl)

    def _extra_inputs(self, state, layout, nlev, ncol):
        x = {k: self._col(state[k], k, layout) for k in _lib.RAIN_OCND2_IN_FIELDS}
        for k, v in x.items():
            if tuple(v.shape) != (nlev, ncol):
                raise ValueError(f"{k}: shape {tuple(v.shape)} != {(nlev, ncol)}")
        return x

    def _launch(self, precision, ncol, nlev, dt, cin, extra, cio, cout, st):
        if precision != "fp64":
            raise ValueError("the OCND2 branch is implemented in FP64 only")
        cin2 = _lib.RainOcnd2In(*[ptr(extra[k]) for k in _lib.RAIN_OCND2_IN_FIELDS])
        return self._lib.ice3_rain_ice_ocnd2(self._ctx, ncol, nlev, ncol, float(dt), ctypes.byref(cin),
                                             ctypes.byref(cin2), ctypes.byref(cio), ctypes.byref(cout),
                                             ctypes.byref(st), stream_ptr(self.device))


class AdjustRainIceOCND2(AdjustRainIce):
    """The fused step with OCND2 in both halves (``ice3_adjust_rain_ice_ocnd2``, FP64):
    IceAdjustOCND2's ICE_ADJUST, then RainIceOCND2's sedimentation and microphysics on the
    adjusted state and ice cloud.  Inputs are AdjustRainIce's plus ``zz`` (level heights, m)
    and ``ice_cld_wgt`` (per column); the result adds ``icldfr, wcldfr, ssio, ssiu, ifr``."""

    def __init__(self, constants=None, layout: str = "auto", device=None,
                 sedimentation: str | None = None, split_max_cfl: float | None = None):
        sedim, cfl = None, None
        if not isinstance(constants, dict):
            check_configuration(constants, "constants", ocnd2=True)
            sedim, cfl = sedimentation_of(constants)
        super().__init__({}, layout, device, sedimentation or sedim,
                         cfl if split_max_cfl is None else split_max_cfl)

    def __call__(self, f: dict, dt: float, return_calls: bool = False, check_inputs: bool = False):
        return self._fused(f, dt, "fp64", return_calls, check_inputs, ocnd2=True)


class RainIceHail(RainIce):
    """RAIN_ICE with KRR = 7 (hail, HCLOUD='ICE4') on MI355X (``ice3_rain_ice_hail``, FP64).

    The reference reaches hail through the Fortran RAIN_ICE with KRR=7 and PRHT / PRHS /
    PINPRH (micro/rain_ice.F90:18,247-249); its JAX component has no hail.  The state takes
    RainIce's keys plus ``rh_t`` (hail mixing ratio at t) and ``rhs`` (its source); the
    returned state has the updated ``rhs`` and the diagnostics add ``inprh`` (surface hail).
    """

    _ENTRY = "ice3_rain_ice_hail"

    def _extra_inputs(self, state, layout, nlev, ncol):
        x = {"rht": self._col(state["rh_t"], "rh_t", layout),
             "rhs": self._col(state["rhs"], "rhs", layout).clone()}
        for k in ("rht", "rhs"):
            if tuple(x[k].shape) != (nlev, ncol):
                raise ValueError(f"{k}: shape {tuple(x[k].shape)} != {(nlev, ncol)}")
        return x

    def _launch(self, precision, ncol, nlev, dt, cin, extra, cio, cout, st):
        if precision != "fp64":
            raise ValueError("the hail path is implemented in FP64 only")
        extra["inprh"] = torch.empty(ncol, dtype=torch.float64, device=self.device)
        h = _lib.Hail(ptr(extra["rht"]), ptr(extra["rhs"]), ptr(extra["inprh"]))
        return self._lib.ice3_rain_ice_hail(self._ctx, ncol, nlev, ncol, float(dt), ctypes.byref(cin),
                                            ctypes.byref(cio), ctypes.byref(cout), ctypes.byref(h),
                                            ctypes.byref(st), stream_ptr(self.device))

    def _extra_results(self, extra, new, diags, layout, shape):
        new["rhs"] = from_columns(extra["rhs"], layout, shape)
        diags["inprh"] = from_columns(extra["inprh"], layout, shape)


class AdjustRainIceHail(AdjustRainIce):
    """The fused step with hail (``ice3_adjust_rain_ice_hail``, FP64): IceAdjustHail's
    ICE_ADJUST, then RainIceHail's sedimentation and microphysics on the adjusted state.  Inputs
    are AdjustRainIce's plus ``rh`` (hail mixing ratio at t) and ``rhs`` (hail source; times dt,
    the hail t-state of the microphysics, as for the other species); the result adds ``rhs``
    (final hail source) and ``inprh``."""

    def __call__(self, f: dict, dt: float, return_calls: bool = False, check_inputs: bool = False):
        return self._fused(f, dt, "fp64", return_calls, check_inputs, hail=True)
