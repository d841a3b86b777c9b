"""Build and load the CPU build of the device stages (tests only).

``load()`` compiles host_stages.cpp with g++ (plain, -O2, no FMA contraction) into
tests/host_stages/_build/libhost_stages.so and returns a ctypes handle whose
``hst_*`` functions take the same C-ABI structs as include/ice3.h, with host
(numpy) pointers.  ``build_sanitized()`` builds the stand-alone driver with
AddressSanitizer and UndefinedBehaviorSanitizer (including float-to-int casts).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "dwarf-p-ice3_amd", "csrc")
BUILD = os.path.join(HERE, "_build")
SOURCES = [os.path.join(HERE, "host_stages.cpp"), os.path.join(CSRC, "ice3_params.cpp")]
INCLUDES = ["-I" + os.path.join(HERE, "shim"), "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
CXX = os.environ.get("CXX_HOST", "g++")

_lib = None


def _stale(target):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    deps = SOURCES + [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "shim", "hip", "hip_runtime.h")]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(args, target):
    os.makedirs(BUILD, exist_ok=True)
    tmp = f"{target}.{os.getpid()}.tmp"  # parallel test workers may build at once
    cmd = [CXX, "-std=c++17", "-ffp-contract=off", *INCLUDES, *args, *SOURCES, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host build failed:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    os.replace(tmp, target)
    return target


def load():
    """Build (when sources changed) and load libhost_stages.so."""
    global _lib
    if _lib is None:
        from ice3_amd import _lib as L
        so = os.path.join(BUILD, "libhost_stages.so")
        if _stale(so):
            _compile(["-O2", "-fPIC", "-shared"], so)
        lib = ctypes.CDLL(so)
        common = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.hst_ice_adjust.argtypes = common + [ctypes.POINTER(L.AdjustIn), ctypes.POINTER(L.AdjustOut)]
        lib.hst_ice_adjust_ocnd2.argtypes = common + [
            ctypes.POINTER(L.AdjustIn), ctypes.POINTER(L.Ocnd2In), ctypes.POINTER(L.AdjustOut),
            ctypes.POINTER(L.Ocnd2Out)]
        lib.hst_rain_ice.argtypes = [ctypes.c_int] + common + [
            ctypes.POINTER(L.RainIn), ctypes.POINTER(L.RainIO), ctypes.POINTER(L.RainOut), ctypes.POINTER(L.Stats)]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.hst_exp_f64.argtypes = [dp, dp, ctypes.c_long]
        lib.hst_log_f64.argtypes = [dp, dp, ctypes.c_long]
        lib.hst_adjust_rain_ice.argtypes = [ctypes.c_int] + common + [
            ctypes.POINTER(L.PipeIn), ctypes.POINTER(L.PipeOut), ctypes.POINTER(L.Stats)]
        lib.hst_set_substep_cap.argtypes = [ctypes.c_int]
        lib.hst_ice_adjust_hail.argtypes = common + [ctypes.POINTER(L.AdjustIn), ctypes.c_void_p,
                                                     ctypes.POINTER(L.AdjustOut)]
        lib.hst_rain_ice_hail.argtypes = common + [
            ctypes.POINTER(L.RainIn), ctypes.POINTER(L.RainIO), ctypes.POINTER(L.RainOut),
            ctypes.POINTER(L.Hail), ctypes.POINTER(L.Stats)]
        lib.hst_adjust_rain_ice_hail.argtypes = common + [
            ctypes.POINTER(L.PipeIn), ctypes.POINTER(L.PipeOut), ctypes.POINTER(L.PipeHail),
            ctypes.POINTER(L.Stats)]
        lib.hst_tiwmx.argtypes = [dp]
        lib.hst_rain_ice_ocnd2.argtypes = common + [
            ctypes.POINTER(L.RainIn), ctypes.POINTER(L.RainOcnd2In), ctypes.POINTER(L.RainIO),
            ctypes.POINTER(L.RainOut), ctypes.POINTER(L.Stats)]
        lib.hst_adjust_rain_ice_ocnd2.argtypes = common + [
            ctypes.POINTER(L.PipeIn), ctypes.POINTER(L.Ocnd2In), ctypes.POINTER(L.PipeOut),
            ctypes.POINTER(L.Ocnd2Out), ctypes.POINTER(L.Stats)]
        _lib = lib
    return _lib


def build_sanitized():
    """Build the sanitizer driver executable (see host_stages.cpp, HST_MAIN)."""
    exe = os.path.join(BUILD, "host_stages_san")
    if _stale(exe):
        _compile(["-O1", "-g", "-DHST_MAIN", "-fno-omit-frame-pointer",
                  "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all"], exe)
    return exe


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(None)


def adjust_rain_ice(f: dict, dt: float = 50.0, prec: int = 0):
    """Fused step on host arrays; f holds the ice3_pipeline_in fields ((nlev, ncol) float64,
    sigqsat (ncol,)).  Returns the ice3_pipeline_out fields plus calls and counters."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = f["th"].shape
    inp = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.PIPE_IN_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in L.PIPE_OUT_FIELDS[:15]}
    out.update({k: np.empty(ncol) for k in L.PIPE_OUT_FIELDS[15:]})
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    cin = L.PipeIn(*[_p(inp[k]) for k in L.PIPE_IN_FIELDS])
    cout = L.PipeOut(*[_p(out[k]) for k in L.PIPE_OUT_FIELDS])
    st = L.Stats(_p(calls), _p(counters))
    rc = lib.hst_adjust_rain_ice(prec, ncol, nlev, ncol, dt, ctypes.byref(cin), ctypes.byref(cout), ctypes.byref(st))
    assert rc == 0
    out["calls"] = calls
    out["counters"] = counters
    return out


def ice_adjust_ocnd2(f: dict, dt: float = 50.0):
    """ICE_ADJUST with OCND2 on host arrays: the ice_adjust fields plus zz and ice_cld_wgt."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = f["th"].shape
    inp = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.ADJ_IN_FIELDS}
    x_in = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.OCND2_IN_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in L.ADJ_OUT_FIELDS}
    x_out = {k: np.empty((nlev, ncol)) for k in L.OCND2_OUT_FIELDS}
    rc = lib.hst_ice_adjust_ocnd2(
        ncol, nlev, ncol, dt, ctypes.byref(L.AdjustIn(*[_p(inp[k]) for k in L.ADJ_IN_FIELDS])),
        ctypes.byref(L.Ocnd2In(*[_p(x_in[k]) for k in L.OCND2_IN_FIELDS])),
        ctypes.byref(L.AdjustOut(*[_p(out[k]) for k in L.ADJ_OUT_FIELDS])),
        ctypes.byref(L.Ocnd2Out(*[_p(x_out[k]) for k in L.OCND2_OUT_FIELDS])))
    assert rc == 0
    return {**out, **x_out}


def rain_ice(s: dict, dt: float = 50.0, prec: int = 0):
    """rain_ice on host arrays with the oracle's field names (tht.., ths.., ci)."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = s["tht"].shape
    inp = {k: np.ascontiguousarray(s[k], dtype=np.float64) for k in L.RAIN_IN_FIELDS}
    io = {k: np.array(s[k], dtype=np.float64, order="C", copy=True) for k in L.RAIN_IO_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in ("evap3d", "rainfr")}
    out.update({k: np.empty(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")})
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    cin = L.RainIn(*[_p(inp[k]) for k in L.RAIN_IN_FIELDS])
    cio = L.RainIO(*[_p(io[k]) for k in L.RAIN_IO_FIELDS])
    cout = L.RainOut(*[_p(out[k]) for k in L.RAIN_OUT_FIELDS])
    st = L.Stats(_p(calls), _p(counters))
    rc = lib.hst_rain_ice(prec, ncol, nlev, ncol, dt, ctypes.byref(cin), ctypes.byref(cio),
                          ctypes.byref(cout), ctypes.byref(st))
    assert rc == 0
    res = {**io, **out, "calls": calls, "counters": counters}
    return res


def elementwise(fn: str, x):
    """y = fn(x) for the kernels' FP64 exp ("exp") or log ("log") on a float64 array."""
    import numpy as np
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    getattr(lib, f"hst_{fn}_f64")(x.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size)
    return y


def adjust_rain_ice_ocnd2(f: dict, dt: float = 50.0):
    """The fused step with OCND2 on host arrays: the ice3_pipeline_in fields plus zz and
    ice_cld_wgt.  Returns the pipeline outputs, the OCND2 outputs, calls and counters."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = f["th"].shape
    inp = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.PIPE_IN_FIELDS}
    zz, wgt = np.ascontiguousarray(f["zz"], dtype=np.float64), np.ascontiguousarray(f["ice_cld_wgt"], dtype=np.float64)
    out = {k: np.empty((nlev, ncol)) for k in L.PIPE_OUT_FIELDS[:15]}
    out.update({k: np.empty(ncol) for k in L.PIPE_OUT_FIELDS[15:]})
    x = {k: np.empty((nlev, ncol)) for k in L.OCND2_OUT_FIELDS}
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    cin = L.PipeIn(*[_p(inp[k]) for k in L.PIPE_IN_FIELDS])
    cout = L.PipeOut(*[_p(out[k]) for k in L.PIPE_OUT_FIELDS])
    st = L.Stats(_p(calls), _p(counters))
    rc = lib.hst_adjust_rain_ice_ocnd2(ncol, nlev, ncol, dt, ctypes.byref(cin), ctypes.byref(L.Ocnd2In(_p(zz), _p(wgt))),
                                       ctypes.byref(cout), ctypes.byref(L.Ocnd2Out(*[_p(x[k]) for k in L.OCND2_OUT_FIELDS])),
                                       ctypes.byref(st))
    assert rc == 0
    out.update(x)
    out["calls"] = calls
    out["counters"] = counters
    return out


def rain_ice_ocnd2(s: dict, dt: float = 50.0):
    """RAIN_ICE with OCND2 on host arrays (rain_ice's fields plus icldfr, ssio, ssiu, ifr)."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = s["tht"].shape
    names = {"exn_ref": "exn_ref"}
    inp = {k: np.ascontiguousarray(s[names.get(k, k)], dtype=np.float64) for k in L.RAIN_IN_FIELDS}
    x4 = [np.ascontiguousarray(s[k], dtype=np.float64) for k in L.RAIN_OCND2_IN_FIELDS]
    io = {k: np.array(s[k if k != "ci" else "ci"], dtype=np.float64, order="C") for k in L.RAIN_IO_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in ("evap3d", "rainfr")}
    out.update({k: np.empty(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg")})
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    rc = lib.hst_rain_ice_ocnd2(ncol, nlev, ncol, dt, ctypes.byref(L.RainIn(*[_p(inp[k]) for k in L.RAIN_IN_FIELDS])),
                                ctypes.byref(L.RainOcnd2In(*[_p(a) for a in x4])),
                                ctypes.byref(L.RainIO(*[_p(io[k]) for k in L.RAIN_IO_FIELDS])),
                                ctypes.byref(L.RainOut(*[_p(out[k]) for k in L.RAIN_OUT_FIELDS])),
                                ctypes.byref(L.Stats(_p(calls), _p(counters))))
    assert rc == 0
    io.update(out)
    io["calls"] = calls
    io["counters"] = counters
    return io


def ice_adjust_hail(f: dict, dt: float = 50.0):
    """ICE_ADJUST with KRR=7 on host arrays: the ice_adjust fields plus the hail mixing ratio rh."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = f["th"].shape
    inp = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.ADJ_IN_FIELDS}
    rh = np.ascontiguousarray(f["rh"], dtype=np.float64)
    out = {k: np.empty((nlev, ncol)) for k in L.ADJ_OUT_FIELDS}
    rc = lib.hst_ice_adjust_hail(ncol, nlev, ncol, dt, ctypes.byref(L.AdjustIn(*[_p(inp[k]) for k in L.ADJ_IN_FIELDS])),
                                 _p(rh), ctypes.byref(L.AdjustOut(*[_p(out[k]) for k in L.ADJ_OUT_FIELDS])))
    assert rc == 0
    return out


def rain_ice_hail(s: dict, dt: float = 50.0):
    """RAIN_ICE with KRR=7 on host arrays (rain_ice's fields plus rht and rhs)."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = s["tht"].shape
    inp = {k: np.ascontiguousarray(s[k], dtype=np.float64) for k in L.RAIN_IN_FIELDS}
    io = {k: np.array(s[k], dtype=np.float64, order="C") for k in L.RAIN_IO_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in ("evap3d", "rainfr")}
    out.update({k: np.empty(ncol) for k in ("inprc", "inprr", "inpri", "inprs", "inprg", "inprh")})
    rht = np.ascontiguousarray(s["rht"], dtype=np.float64)
    rhs = np.array(s["rhs"], dtype=np.float64, order="C")
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    rc = lib.hst_rain_ice_hail(ncol, nlev, ncol, dt, ctypes.byref(L.RainIn(*[_p(inp[k]) for k in L.RAIN_IN_FIELDS])),
                               ctypes.byref(L.RainIO(*[_p(io[k]) for k in L.RAIN_IO_FIELDS])),
                               ctypes.byref(L.RainOut(*[_p(out[k]) for k in L.RAIN_OUT_FIELDS])),
                               ctypes.byref(L.Hail(_p(rht), _p(rhs), _p(out["inprh"]))),
                               ctypes.byref(L.Stats(_p(calls), _p(counters))))
    assert rc == 0
    io.update(out)
    io.update(rhs=rhs, calls=calls, counters=counters)
    return io


def adjust_rain_ice_hail(f: dict, dt: float = 50.0):
    """The fused step with hail on host arrays: the ice3_pipeline_in fields plus rh and rhs."""
    import numpy as np
    from ice3_amd import _lib as L
    lib = load()
    nlev, ncol = f["th"].shape
    inp = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in L.PIPE_IN_FIELDS}
    out = {k: np.empty((nlev, ncol)) for k in L.PIPE_OUT_FIELDS[:15]}
    out.update({k: np.empty(ncol) for k in L.PIPE_OUT_FIELDS[15:]})
    rh, rhs = (np.ascontiguousarray(f[k], dtype=np.float64) for k in ("rh", "rhs"))
    rhs_out, inprh = np.empty((nlev, ncol)), np.empty(ncol)
    calls = np.zeros((nlev, ncol), dtype=np.int32)
    counters = np.zeros(3, dtype=np.uint64)
    rc = lib.hst_adjust_rain_ice_hail(ncol, nlev, ncol, dt,
                                      ctypes.byref(L.PipeIn(*[_p(inp[k]) for k in L.PIPE_IN_FIELDS])),
                                      ctypes.byref(L.PipeOut(*[_p(out[k]) for k in L.PIPE_OUT_FIELDS])),
                                      ctypes.byref(L.PipeHail(_p(rh), _p(rhs), _p(rhs_out), _p(inprh))),
                                      ctypes.byref(L.Stats(_p(calls), _p(counters))))
    assert rc == 0
    out.update(rhs=rhs_out, inprh=inprh, calls=calls, counters=counters)
    return out
