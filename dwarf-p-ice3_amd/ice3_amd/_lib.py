"""ctypes binding of libice3_hip.so (C ABI in include/ice3.h).

The library is built in-tree by ``ice3_amd.build.build()`` (hipcc, gfx950) and
loaded from ``ice3_amd/_lib/libice3_hip.so``.  There is no fallback: when the
library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libice3_hip.so")

ICE3_OK, ICE3_EINVAL, ICE3_EDEVICE, ICE3_ESUBSTEP, ICE3_ECONFIG, ICE3_ETOOLARGE = 0, -1, -2, -3, -4, -5
SURROGATE_SCRATCH_BYTES = 64 << 20
ICE3_MAX_CHECK_FIELDS = 32
PREC_FP64, PREC_MIXED = 0, 1

_P = ctypes.c_void_p


def _struct(name, fields):
    return type(name, (ctypes.Structure,), {"_fields_": [(f, _P) for f in fields]})


ADJ_IN_FIELDS = ("pabs", "sigs", "th", "exn", "exn_ref", "rho_dry_ref", "rv", "rc", "ri", "rr",
                 "rs", "rg", "cf_mf", "rc_mf", "ri_mf", "rvs", "rcs", "ris", "ths", "sigqsat")
ADJ_OUT_FIELDS = ("t", "rv_out", "rc_out", "ri_out", "cldfr", "hlc_hrc", "hlc_hcf", "hli_hri",
                  "hli_hcf", "cph", "lv", "ls", "rvs", "rcs", "ris", "ths", "srcs")
OCND2_IN_FIELDS = ("zz", "ice_cld_wgt")
OCND2_OUT_FIELDS = ("icldfr", "wcldfr", "ssio", "ssiu", "ifr")
RAIN_IN_FIELDS = ("exn", "dzz", "rhodj", "rhodref", "exn_ref", "pabs", "cldfr", "hlc_hrc",
                  "hlc_hcf", "hli_hri", "hli_hcf", "sigs", "tht", "rvt", "rct", "rrt", "rit",
                  "rst", "rgt")
RAIN_IO_FIELDS = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci")
RAIN_OUT_FIELDS = ("evap3d", "rainfr", "inprc", "inprr", "inpri", "inprs", "inprg")
RAIN_OCND2_IN_FIELDS = ("icldfr", "ssio", "ssiu", "ifr")
PIPE_IN_FIELDS = ("pabs", "sigs", "th", "exn", "exn_ref", "rhodref", "rhodj", "dzz", "rv", "rc",
                  "ri", "rr", "rs", "rg", "ci", "cf_mf", "rc_mf", "ri_mf", "ths", "rvs", "rcs",
                  "rrs", "ris", "rss", "rgs", "sigqsat")
PIPE_OUT_FIELDS = ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "cldfr", "hlc_hrc",
                   "hlc_hcf", "hli_hri", "hli_hcf", "ci", "evap3d", "rainfr", "inprc", "inprr",
                   "inpri", "inprs", "inprg")

AdjustIn = _struct("ice3_ice_adjust_in", ADJ_IN_FIELDS)
AdjustOut = _struct("ice3_ice_adjust_out", ADJ_OUT_FIELDS)
Ocnd2In = _struct("ice3_ocnd2_in", OCND2_IN_FIELDS)
Ocnd2Out = _struct("ice3_ocnd2_out", OCND2_OUT_FIELDS)
RainIn = _struct("ice3_rain_ice_in", RAIN_IN_FIELDS)
RainIO = _struct("ice3_rain_ice_inout", RAIN_IO_FIELDS)
RainOut = _struct("ice3_rain_ice_out", RAIN_OUT_FIELDS)
RainOcnd2In = _struct("ice3_rain_ocnd2_in", RAIN_OCND2_IN_FIELDS)
HAIL_FIELDS = ("rht", "rhs", "inprh")
PIPE_HAIL_FIELDS = ("rh", "rhs", "rhs_out", "inprh")
PipeIn = _struct("ice3_pipeline_in", PIPE_IN_FIELDS)
PipeOut = _struct("ice3_pipeline_out", PIPE_OUT_FIELDS)
Stats = _struct("ice3_stats", ("calls", "counters"))
Hail = _struct("ice3_hail", HAIL_FIELDS)
PipeHail = _struct("ice3_pipeline_hail", PIPE_HAIL_FIELDS)

STAGE_ADJUST, STAGE_SEDIMENTATION, STAGE_MICROPHYSICS = 1, 2, 4
SEDIM = {"STAT": 0, "SPLI": 1}   # ICE3_SEDIM_STAT, ICE3_SEDIM_SPLI
NFEAT, NTARG = 18, 8


class SurrogateNorm(ctypes.Structure):
    """ice3_surrogate_norm (include/ice3.h)."""
    _fields_ = [("feat_mean", ctypes.c_float * NFEAT), ("feat_scale", ctypes.c_float * NFEAT),
                ("feat_lo", ctypes.c_float * NFEAT), ("feat_hi", ctypes.c_float * NFEAT),
                ("targ_unit", ctypes.c_float * NTARG), ("targ_mean", ctypes.c_float * NTARG),
                ("targ_std", ctypes.c_float * NTARG)]


class Surrogate(ctypes.Structure):
    """ice3_surrogate (include/ice3.h)."""
    _fields_ = [("weights_packed", _P), ("params", _P), ("norm", _P), ("budget_tol", ctypes.c_float),
                ("neg_tol", ctypes.c_float)]


EXPORTS = ("ice3_ctx_create", "ice3_ctx_destroy", "ice3_strerror", "ice3_dump_constants",
           "ice3_ctx_set_substep_cap", "ice3_ctx_status", "ice3_ctx_set_sedimentation",
           "ice3_ice_adjust", "ice3_ice_adjust_ocnd2", "ice3_rain_ice", "ice3_adjust_rain_ice",
           "ice3_rain_ice_ocnd2", "ice3_adjust_rain_ice_ocnd2",
           "ice3_ice_adjust_hail", "ice3_rain_ice_hail", "ice3_adjust_rain_ice_hail",
           "ice3_rain_ice_branches", "ice3_ice_adjust_branches",
           "ice3_stage_timing",
           "ice3_stage_times", "ice3_adjust_rain_ice_stages", "ice3_adjust_rain_ice_surrogate",
           "ice3_adjust_rain_ice_surrogate_noalloc",
           "ice3_surrogate_features", "ice3_surrogate_work_bytes", "ice3_mlp_sizes", "ice3_mlp_forward", "ice3_mlp_backward",
           "ice3_mlp_adam", "ice3_count_nonfinite")

_lib = None


class Ice3Error(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Ice3Error(f"{LIB_PATH} is missing: build it with ice3_amd.build.build()")
        lib = ctypes.CDLL(LIB_PATH)
        lib.ice3_ctx_create.argtypes = [ctypes.POINTER(_P)]
        lib.ice3_ctx_destroy.argtypes = [_P]
        lib.ice3_ctx_destroy.restype = None
        lib.ice3_strerror.argtypes = [ctypes.c_int]
        lib.ice3_strerror.restype = ctypes.c_char_p
        lib.ice3_ctx_set_substep_cap.argtypes = [_P, ctypes.c_int]
        lib.ice3_ctx_status.argtypes = [_P, _P, ctypes.POINTER(ctypes.c_ulonglong)]
        lib.ice3_ctx_set_sedimentation.argtypes = [_P, ctypes.c_int, ctypes.c_double]
        lib.ice3_dump_constants.argtypes = [_P, ctypes.c_char_p, ctypes.c_long]
        lib.ice3_dump_constants.restype = ctypes.c_long
        common = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.ice3_ice_adjust.argtypes = common + [ctypes.POINTER(AdjustIn), ctypes.POINTER(AdjustOut), _P]
        lib.ice3_ice_adjust_ocnd2.argtypes = [
            _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(AdjustIn),
            ctypes.POINTER(Ocnd2In), ctypes.POINTER(AdjustOut), ctypes.POINTER(Ocnd2Out), _P]
        lib.ice3_rain_ice.argtypes = common + [ctypes.POINTER(RainIn), ctypes.POINTER(RainIO),
                                               ctypes.POINTER(RainOut), ctypes.POINTER(Stats), _P]
        lib.ice3_adjust_rain_ice.argtypes = common + [ctypes.POINTER(PipeIn), ctypes.POINTER(PipeOut),
                                                      ctypes.POINTER(Stats), _P]
        c5 = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.ice3_rain_ice_ocnd2.argtypes = c5 + [ctypes.POINTER(RainIn), ctypes.POINTER(RainOcnd2In),
                                                 ctypes.POINTER(RainIO), ctypes.POINTER(RainOut),
                                                 ctypes.POINTER(Stats), _P]
        lib.ice3_adjust_rain_ice_ocnd2.argtypes = c5 + [ctypes.POINTER(PipeIn), ctypes.POINTER(Ocnd2In),
                                                        ctypes.POINTER(PipeOut), ctypes.POINTER(Ocnd2Out),
                                                        ctypes.POINTER(Stats), _P]
        lib.ice3_ice_adjust_hail.argtypes = c5 + [ctypes.POINTER(AdjustIn), _P, ctypes.POINTER(AdjustOut), _P]
        lib.ice3_rain_ice_hail.argtypes = c5 + [ctypes.POINTER(RainIn), ctypes.POINTER(RainIO),
                                                ctypes.POINTER(RainOut), ctypes.POINTER(Hail),
                                                ctypes.POINTER(Stats), _P]
        lib.ice3_rain_ice_branches.argtypes = c5 + [ctypes.POINTER(RainIn), ctypes.POINTER(RainIO),
                                                    ctypes.POINTER(RainOut), ctypes.POINTER(Hail),
                                                    _P, _P]
        lib.ice3_ice_adjust_branches.argtypes = c5 + [ctypes.POINTER(AdjustIn), _P,
                                                      ctypes.POINTER(AdjustOut), _P, _P]
        lib.ice3_adjust_rain_ice_hail.argtypes = c5 + [ctypes.POINTER(PipeIn), ctypes.POINTER(PipeOut),
                                                       ctypes.POINTER(PipeHail), ctypes.POINTER(Stats), _P]
        lib.ice3_adjust_rain_ice_stages.argtypes = common + [
            ctypes.POINTER(PipeIn), ctypes.POINTER(PipeOut), ctypes.POINTER(Stats), ctypes.c_int, _P]
        lib.ice3_adjust_rain_ice_surrogate.argtypes = [
            _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(PipeIn),
            ctypes.POINTER(PipeOut), ctypes.POINTER(Surrogate), _P, _P, ctypes.POINTER(Stats), _P]
        lib.ice3_adjust_rain_ice_surrogate_noalloc.argtypes = [
            _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(PipeIn),
            ctypes.POINTER(PipeOut), ctypes.POINTER(Surrogate), _P, ctypes.POINTER(Stats), _P]
        lib.ice3_surrogate_work_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ice3_surrogate_work_bytes.restype = ctypes.c_size_t
        lib.ice3_surrogate_features.argtypes = [
            _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(PipeIn),
            ctypes.POINTER(PipeOut), _P, _P, _P, _P, _P]
        lib.ice3_mlp_sizes.argtypes = [ctypes.POINTER(ctypes.c_int)] * 5
        lib.ice3_mlp_forward.argtypes = [ctypes.c_int] + [_P] * 10
        lib.ice3_mlp_backward.argtypes = [ctypes.c_int] + [_P] * 8
        lib.ice3_mlp_adam.argtypes = [ctypes.c_int] + [ctypes.c_float] * 4 + [_P] * 7
        lib.ice3_count_nonfinite.argtypes = [ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_int, ctypes.c_int, _P, _P]
        lib.ice3_stage_timing.argtypes = [_P, ctypes.c_int]
        lib.ice3_stage_times.argtypes = [_P, ctypes.POINTER(ctypes.c_float)]
        _lib = lib
    return _lib


def require_finite(fields: dict, device):
    """Raise ValueError naming every field of ``fields`` (name -> float64 CUDA tensor, a
    (nlev, ncol) field or an (ncol,) per-column one, contiguous) that holds a NaN or an
    infinity, with the number found.  One ice3_count_nonfinite launch on the device's
    current stream, then a synchronising read of the counts."""
    import torch
    from ._fields import ptr, stream_ptr
    items = [(k, v) for k, v in fields.items() if v is not None and v.numel()]
    if not items:
        return
    if len(items) > ICE3_MAX_CHECK_FIELDS:
        raise ValueError(f"at most {ICE3_MAX_CHECK_FIELDS} fields per check")
    ncol = items[0][1].shape[-1]
    for k, v in items:
        if v.dtype != torch.float64 or not v.is_contiguous() or v.dim() not in (1, 2) or v.shape[-1] != ncol:
            raise ValueError(f"{k}: expected a contiguous float64 field of {ncol} columns")
    counts = torch.zeros(len(items), dtype=torch.int64, device=device)
    fptr = (_P * len(items))(*[ptr(v).value for _, v in items])
    rows = (ctypes.c_int * len(items))(*[v.shape[0] if v.dim() == 2 else 1 for _, v in items])
    lib = load()
    with torch.cuda.device(device):
        check(lib.ice3_count_nonfinite(len(items), fptr, rows, ncol, ncol, ptr(counts), stream_ptr(device)),
              "ice3_count_nonfinite")
    n = counts.tolist()
    bad = [f"{k} ({c})" for (k, _), c in zip(items, n) if c]
    if bad:
        raise ValueError("non-finite input values (NaN or infinity): " + ", ".join(bad))


def check(rc: int, what: str):
    if rc != ICE3_OK:
        msg = load().ice3_strerror(rc).decode()
        raise Ice3Error(f"{what} failed ({rc}): {msg}")


def host_constants() -> dict:
    """The library's derived constants and tables, built on the host (no GPU needed)."""
    lib = load()
    n = lib.ice3_dump_constants(None, None, 0)
    if n < 0:
        raise Ice3Error(f"ice3_dump_constants failed: {n}")
    buf = ctypes.create_string_buffer(n + 1)
    lib.ice3_dump_constants(None, buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        p = line.split()
        out[p[0]] = [float(x) for x in p[2:]]
    return out


_ctx = {}


def context(device: int | None = None):
    """Per-device context (constants uploaded once per device, kept until exit)."""
    import torch

    dev = torch.cuda.current_device() if device is None else device
    if dev not in _ctx:
        lib = load()
        with torch.cuda.device(dev):
            h = _P()
            check(lib.ice3_ctx_create(ctypes.byref(h)), "ice3_ctx_create")
        _ctx[dev] = h
    return _ctx[dev]
