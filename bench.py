"""Benchmark: grid columns/s of the fused ice_adjust + rain_ice step (90 levels, FP64).

Workload (the headline configuration): 512 x 512 = 262,144 synthetic columns x 90
levels, dt = 50 s, generator variant "pipeline" (precipitating bands in 20 % of the
columns), seed 20260003.  A step is one pass of ICE_ADJUST -> statistical
sedimentation -> ICE4 microphysics -> total tendencies -> negativity correction over
all columns: one ice3_adjust_rain_ice call, i.e. the kernels k_adjust_sedim (ICE_ADJUST
and the sedimentation in one column walk) and k_micro on one stream.  Inputs are
resident in HBM before the timed region;
the K steps are bracketed by synchronisations, each step is timed by its own pair of
HIP events on the launch stream, and the reported time is the median step (the mean
is reported beside it); every call's return code is checked after the timed loop.

Also reported: the duration of each stage (HIP events recorded by the library
around each kernel, in a second pass of K steps), the roofline of the dominant
kernel from its algorithmic bytes per point (k_micro: 18 float64 reads + 10
writes = 224 B/point), the whole step's algorithmic rate (25 float64 inputs + 15
outputs = 320 B/point plus 48 B/column), the mixed-precision variant with its
per-field error against FP64, and a CPU baseline (the compiled FP64 PHYEX
reference when its library is present, else the C restatement of the oracle)
timed on a bounded sample on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--ncol N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "dwarf-p-ice3_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

NLEV = 90
DT = 50.0
SEED = 20260003
BYTES_PER_POINT = 320.0   # whole step: 25 inputs + 15 outputs, float64
BYTES_PER_COLUMN = 48.0   # sigqsat + 5 surface precipitation fields
# algorithmic bytes of each stage (per point, per column) as launched by the fused step
STAGE_BYTES = {"ice_adjust": (232.0, 8.0),      # 19 reads + 10 writes; sigqsat
               "sedimentation": (80.0, 40.0),   # 6 reads + 4 writes; 5 precipitation fields
               # one column walk: 23 reads + 13 writes (ri once); sigqsat + precipitation
               "ice_adjust+sedimentation": (288.0, 48.0),
               "microphysics": (224.0, 0.0)}    # 18 reads + 10 writes
STAGE_KERNELS = {"ice_adjust": "k_ice_adjust", "sedimentation": "k_sedim",
                 "ice_adjust+sedimentation": "k_adjust_sedim", "microphysics": "k_micro"}
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP64_VALU_PEAK_TFLOPS = 78.6  # MI355X FP64 vector
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def _profiled(table: str, kernel: str, ncol: int):
    """Entry `kernel` of `table` in traffic.json (profiled on the headline workload), else None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("workload", {}).get("columns") != ncol or t["workload"].get("levels") != NLEV:
        return None
    return t.get(table, {}).get(kernel)


def measured_traffic(kernel: str, ncol: int):
    """Profiled HBM bytes per dispatch of `kernel` for the headline workload, else None."""
    return _profiled("bytes_per_dispatch", kernel, ncol)


def measured_fp64_flops(kernel: str, ncol: int):
    """FP64 VALU FLOPs per dispatch of `kernel` for the headline workload (rocprofv3
    SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes), else None."""
    return _profiled("fp64_flops_per_dispatch", kernel, ncol)


def cpu_baseline(budget_s: float = 20.0, workers: int | None = None) -> dict:
    """Time the CPU path on a bounded sample of the same workload, before the GPU is touched.

    Kind "reference": the FP64 flang build of the PHYEX sources (oracle/_ref), run by
    independent worker processes (its Fortran module state is not thread-safe), each on
    its own slab of columns from the same generator, with NPROMICRO=128 cache blocking.
    Kind "port": the C restatement of the oracle.  Each worker generates its slab, then
    times only the physics; the rate is the total column count over the slowest worker.
    """
    import multiprocessing as mp

    from oracle import ref_lib

    kind = "reference" if ref_lib.available() else "port"
    build = None
    if kind == "reference":
        if ref_lib.available(fast=True):
            ref_lib.use_fast_build()  # inherited by the forked workers
            build = "flang -O3 -march=x86-64-v3 -fdefault-real-8 (FMA contraction)"
        else:
            build = "flang -O2 -fdefault-real-8 (x86-64 baseline, no FMA)"
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    workers = workers or max(1, min(16, ncpu))
    per_col = _cpu_slab((kind, 256, SEED))[1] / 256          # calibration on one core
    ncol_w = int(max(64, min(16384, budget_s / per_col / 1.5)))
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        res = pool.map_async(_cpu_slab, [(kind, ncol_w, SEED + 1 + w) for w in range(workers)]).get(
            40 * budget_s + 300)
    wall = max(t for _, t in res)
    ncol = ncol_w * workers
    return {"value": ncol / wall, "unit": "columns/s", "cores": workers, "kind": kind, "build": build,
            "sample": f"{workers} processes x {ncol_w} columns x {NLEV} levels of the same generator "
                      f"(seeds {SEED + 1}..{SEED + workers}), fused ice_adjust+rain_ice FP64, "
                      f"slowest worker {wall:.1f} s",
            "ice_adjust_1024x90": _cpu_adjust_1024(kind),
            "jax_cpu": "not timed: jax is not installed on this image and there is no package index"}


def _cpu_adjust_1024(kind: str, runs: int = 7) -> dict:
    """The reference's own CPU case: ice_adjust alone, FP64, 1024 columns x 90 levels of the
    "adjust" generator, one process, median of `runs` calls (the call includes the
    binding's column chunking and copies)."""
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(1024, NLEV, seed=20260001, variant="adjust", dt=DT)
    if kind == "reference":
        ref_lib.init(DT, 128)
        fn = lambda: ref_lib.ice_adjust(f, DT)  # noqa: E731
    else:
        fn = lambda: oracle_lib.ice_adjust(f, DT)  # noqa: E731
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ms = sorted(ts)[runs // 2] * 1e3
    return {"ms": ms, "columns_per_s": 1024 / (ms * 1e-3), "cores": 1, "kind": kind}


def _cpu_slab(arg):
    kind, ncol, seed = arg
    from ice3_amd import synthetic
    from oracle import oracle_lib, ref_lib
    f = synthetic.generate(ncol, NLEV, seed=seed, variant="pipeline", dt=DT)
    if kind == "reference":
        ref_lib.init(DT, 128)
        t0 = time.perf_counter()
        a = ref_lib.ice_adjust(f, DT)
        ref_lib.rain_ice(ref_lib.pipeline_state(f, a, DT), DT, nproma=128)
    else:
        t0 = time.perf_counter()
        a = oracle_lib.ice_adjust(f, DT)
        oracle_lib.rain_ice(ref_lib.pipeline_state(f, a, DT), DT)
    return ncol, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncol", type=int, default=512 * 512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjust", action="store_true", help="skip the ice_adjust-alone line")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate line")
    ap.add_argument("--no-split", action="store_true", help="skip the split-sedimentation line")
    ap.add_argument("--no-configs", action="store_true", help="skip the OCND2 and hail lines")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device twice (a plain copy of known size, "
                         "to scale HBM byte counters read by a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(columns are independent: a multi-GPU run would give each GPU ncol/N columns)",
              file=sys.stderr)
        sys.exit(2)

    t_start = time.time()

    def progress(what):
        print(f"bench: {what} ({time.time() - t_start:.0f} s)", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_baseline()
    progress("CPU baseline done" if cpu else "CPU baseline skipped")

    import numpy as np
    import torch

    from ice3_amd import _lib, synthetic
    from ice3_amd._fields import ptr

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    ncol = args.ncol
    f = synthetic.generate(ncol, NLEV, seed=SEED, variant="pipeline", dt=DT)
    f["rhodref"] = f.pop("rho_dry_ref")
    g = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    del f
    lib = _lib.load()
    ctx = _lib.context(0)

    def outputs():
        o = {k: torch.empty((NLEV, ncol), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
        o.update({k: torch.empty(ncol, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
        return o

    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)

    def call(prec, cout):
        return lib.ice3_adjust_rain_ice(ctx, prec, ncol, NLEV, ncol, DT, ctypes.byref(cin),
                                        ctypes.byref(cout), None, sp)

    def run(prec, out, steps, warmup):
        """(median, mean) ms per step over `steps` steps, each bracketed by its own HIP events
        on the launch stream; every call's return code is checked after the timed loop."""
        cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
        return timed_steps(lambda: call(prec, cout), stream, steps, warmup)

    def stage_times_of(fn, steps, merged):
        """Average duration of each stage's kernel over `steps` calls of fn, from the library's
        HIP events (merged: ICE_ADJUST and the sedimentation reported together, where one
        kernel runs both)."""
        _lib.check(lib.ice3_stage_timing(ctx, 1), "ice3_stage_timing")
        acc = [0.0, 0.0, 0.0]
        ms3 = (ctypes.c_float * 6)()
        for _ in range(steps):
            _lib.check(fn(), "timed step")
            _lib.check(lib.ice3_stage_times(ctx, ms3), "ice3_stage_times")
            acc = [x + y for x, y in zip(acc, ms3)]
        _lib.check(lib.ice3_stage_timing(ctx, 0), "ice3_stage_timing")
        if merged:
            return {"ice_adjust+sedimentation": (acc[0] + acc[1]) / steps, "microphysics": acc[2] / steps}
        return dict(zip(("ice_adjust", "sedimentation", "microphysics"), [x / steps for x in acc]))

    def stage_times(prec, out, steps, merged=None):
        """Stage times of the fused step in precision prec (merged by default for the FP64
        statistical step, whose column walk runs ICE_ADJUST and the sedimentation)."""
        cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
        return stage_times_of(lambda: call(prec, cout), steps,
                              merged if merged is not None else prec == _lib.PREC_FP64)

    out64 = outputs()
    progress("inputs resident on the GPU")
    ms, ms_mean = run(_lib.PREC_FP64, out64, args.steps, args.warmup)
    stages = stage_times(_lib.PREC_FP64, out64, args.steps)
    progress(f"FP64 step timed: {ms:.3f} ms median")
    npts = ncol * NLEV
    alg_bytes = BYTES_PER_POINT * npts + BYTES_PER_COLUMN * ncol
    dom = max(stages, key=stages.get)
    bp, bc = STAGE_BYTES[dom]
    dom_achieved = (bp * npts + bc * ncol) / (stages[dom] * 1e-3) / 1e9
    res = {
        "metric": "grid columns/s (90 lev) ice_adjust+rain_ice, FP64",
        "value": ncol / (ms * 1e-3),
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "ms_per_step_mean": ms_mean,
        "timing": f"median of {args.steps} steps, each between its own HIP events on the launch stream",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fused ice_adjust+rain_ice, statistical sedimentation, FP64, "
                               "512x512 columns x 90 levels, dt=50 s",
                   "columns": ncol, "levels": NLEV, "dt_s": DT, "generator": "pipeline", "seed": SEED},
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNELS[dom], "achieved": dom_achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom_achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(STAGE_KERNELS[dom], ncol),
                     "bytes_per_point": bp, "bytes_per_column": bc, "ms_per_launch": stages[dom]},
        "valu_roofline": valu_roofline(STAGE_KERNELS[dom], stages[dom], ncol),
        "stages_ms": stages,
        "step_algorithmic_GBs": alg_bytes / (ms * 1e-3) / 1e9,
        "cpu_baseline": cpu,
        "input_check": input_check(lib, g, ncol, stream, sp, args.steps),
    }
    def budget_err(out):
        """Max over columns of |column water change - surface precipitation| / column water."""
        w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
        mass = g["rhodref"] * g["dzz"]
        col0 = (mass * sum(g[k] for k in w)).sum(0) * DT
        col1 = (mass * sum(out[k] for k in w)).sum(0) * DT
        prec = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"]) * DT
        return ((col1 - col0 - prec).abs() / col0).max().item()

    def substeps(prec, out):
        """ICE4 time-splitting statistics of one extra (untimed) step: tendency calls per
        point, their histogram, and the mean over 64-point groups (one wavefront of the
        point kernel) of the largest count, which sets how many passes a wavefront makes."""
        cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
        calls = torch.zeros(npts, dtype=torch.int32, device=dev)
        cnt = torch.zeros(3, dtype=torch.int64, device=dev)
        st = _lib.Stats(ptr(calls), ptr(cnt))
        _lib.check(lib.ice3_adjust_rain_ice(ctx, prec, ncol, NLEV, ncol, DT, ctypes.byref(cin),
                                            ctypes.byref(cout), ctypes.byref(st), sp), "substep step")
        torch.cuda.synchronize()
        c = calls.abs().to(torch.int64)
        hist = torch.bincount(c).tolist()
        n_calls, capped, n_micro = cnt.tolist()
        wmax = c.view(NLEV, ncol // 64, 64).amax(-1).double() if ncol % 64 == 0 else None
        return {"tendency_calls": n_calls, "points": npts, "points_with_microphysics": n_micro,
                "calls_per_microphysics_point": n_calls / max(n_micro, 1),
                "capped_points": capped, "histogram_calls": hist,
                "wavefront_max_calls_mean": wmax.mean().item() if wmax is not None else None}

    res["water_budget_max_rel_err"] = budget_err(out64)
    res["ice4_substeps"] = substeps(_lib.PREC_FP64, out64)
    if not args.no_split:
        # the same step with the split sedimentation (PARAM_ICE CSEDIM='SPLI', XSPLIT_MAXCFL 0.8)
        outs = outputs()
        _lib.check(lib.ice3_ctx_set_sedimentation(ctx, _lib.SEDIM["SPLI"], 0.8), "SPLI")
        try:
            ms_s, _ = run(_lib.PREC_FP64, outs, args.steps, args.warmup)
            st_s = stage_times(_lib.PREC_FP64, outs, args.steps, merged=False)
        finally:
            _lib.check(lib.ice3_ctx_set_sedimentation(ctx, _lib.SEDIM["STAT"], 0.8), "STAT")
        res.setdefault("variants", {})["split_sedimentation"] = {
            "columns_per_s": ncol / (ms_s * 1e-3), "ms_per_step": ms_s, "speedup_vs_stat": ms / ms_s,
            "stages_ms": st_s, "water_budget_max_rel_err": budget_err(outs),
            "note": "fused step with CSEDIM='SPLI': k_ice_adjust, k_sedim_split (FP64; it also keeps "
                    "the state at t of th/rv/rc for the microphysics), k_micro"}
        progress(f"SPLI step timed: {ms_s:.3f} ms median")
        del outs
    if not args.no_configs:
        # the other configurations of the same step on the same inputs: OCND2 in ICE_ADJUST
        # and RAIN_ICE (PARAMI%LOCND2), and KRR=7 hail
        V = res.setdefault("variants", {})
        V["ocnd2"] = ocnd2_step(lib, ctx, dev, g, cin, sp, stream, args, outputs, budget_err,
                                stage_times_of, ms)
        progress(f"OCND2 step timed: {V['ocnd2']['ms_per_step']:.3f} ms median")
        V["hail"] = hail_step(lib, ctx, dev, g, cin, sp, stream, args, outputs, stage_times_of, ms)
        progress(f"hail step timed: {V['hail']['ms_per_step']:.3f} ms median")
    if not args.no_mixed:
        out32 = outputs()
        ms32, _ = run(_lib.PREC_MIXED, out32, args.steps, args.warmup)
        stages32 = stage_times(_lib.PREC_MIXED, out32, args.steps)
        err = {}
        for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs"):
            d = (out32[k] - out64[k]).abs()
            scale = out64[k].abs().max().item()
            err[k] = {"max_abs": d.max().item(), "rms_abs": d.pow(2).mean().sqrt().item(),
                      "max_rel_to_field_max": d.max().item() / scale if scale > 0 else 0.0}
        progress(f"mixed step timed: {ms32:.3f} ms median")
        res.setdefault("variants", {})["mixed"] = {"columns_per_s": ncol / (ms32 * 1e-3), "ms_per_step": ms32,
                                     "speedup_vs_fp64": ms / ms32, "stages_ms": stages32,
                                     "water_budget_max_rel_err": budget_err(out32),
                                     "error_vs_fp64": err}
    if not args.no_surrogate:
        del out64
        if not args.no_mixed:
            del out32
        res.setdefault("variants", {})["surrogate_1024x1024"] = surrogate_variant(lib, ctx, dev, g, args)
        progress("surrogate line done")
    del g
    if not args.no_adjust:
        res.setdefault("variants", {})["ice_adjust_256x256"] = ice_adjust_alone(lib, ctx, dev, args)
        res["variants"]["ice_adjust_ocnd2_256x256"] = ice_adjust_alone(lib, ctx, dev, args, ocnd2=True)
        res["variants"]["ice_adjust_1024x90"] = ice_adjust_alone(lib, ctx, dev, args, ncol=1024)
    print(json.dumps(res))


def valu_roofline(kernel: str, ms: float, ncol: int) -> dict:
    """FP64 VALU roofline of `kernel`: its profiled FP64 FLOPs per dispatch over the
    event-timed duration `ms` of this run, against the FP64 vector peak."""
    flops = measured_fp64_flops(kernel, ncol)
    tf = flops / (ms * 1e-3) / 1e12 if flops else None
    return {"bound": "valu_fp64", "kernel": kernel, "achieved": tf, "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tf / FP64_VALU_PEAK_TFLOPS if tf else None,
            "flops_per_launch": flops, "ms_per_launch": ms,
            "flops_source": "rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 on this workload (profiles/traffic.json)"}


def ocnd2_step(lib, ctx, dev, g, cin, sp, stream, args, outputs, budget_err, stage_times_of, ms_stat):
    """The fused step with OCND2 in both halves (ice3_adjust_rain_ice_ocnd2) on the headline
    inputs, with the generator's level heights in every column and PICE_CLD_WGT drawn from
    U(0.5, 1.5)."""
    import numpy as np
    import torch
    from ice3_amd import _lib, synthetic
    from ice3_amd._fields import ptr
    ncol = g["th"].shape[1]
    zz = torch.as_tensor(np.ascontiguousarray(np.broadcast_to(synthetic.level_profile(NLEV)[2][:, None],
                                                              (NLEV, ncol))), device=dev)
    wgt = torch.as_tensor(np.random.Generator(np.random.Philox(SEED + 1)).uniform(0.5, 1.5, ncol), device=dev)
    out = outputs()
    x = {k: torch.empty((NLEV, ncol), dtype=torch.float64, device=dev) for k in _lib.OCND2_OUT_FIELDS}
    cin2 = _lib.Ocnd2In(ptr(zz), ptr(wgt))
    cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
    cout2 = _lib.Ocnd2Out(*[ptr(x[k]) for k in _lib.OCND2_OUT_FIELDS])

    def fn():
        return lib.ice3_adjust_rain_ice_ocnd2(ctx, ncol, NLEV, ncol, DT, ctypes.byref(cin), ctypes.byref(cin2),
                                              ctypes.byref(cout), ctypes.byref(cout2), None, sp)
    ms, _ = timed_steps(fn, stream, args.steps, args.warmup)
    return {"columns_per_s": ncol / (ms * 1e-3), "ms_per_step": ms, "speedup_vs_stat_krr6": ms_stat / ms,
            "stages_ms": stage_times_of(fn, args.steps, False),
            "water_budget_max_rel_err": budget_err(out),
            "note": "fused step with PARAMI%LOCND2: k_ice_adjust_ocnd2, k_sedim<double> (which also "
                    "writes the model level height), k_micro<double, OCND2> (persistent)"}


def hail_step(lib, ctx, dev, g, cin, sp, stream, args, outputs, stage_times_of, ms_stat):
    """The fused step with KRR=7 (ice3_adjust_rain_ice_hail) on the headline inputs with hail
    added where there is graupel: rh = rg x U(0.2, 2), its source rh / dt."""
    import numpy as np
    import torch
    from ice3_amd import _lib
    from ice3_amd._fields import ptr
    ncol = g["th"].shape[1]
    rng = np.random.Generator(np.random.Philox(SEED + 7))
    rh = g["rg"] * torch.as_tensor(rng.uniform(0.2, 2.0, (NLEV, ncol)), device=dev)
    rhs = rh / DT
    rhs_out = torch.empty_like(rh)
    inprh = torch.empty(ncol, dtype=torch.float64, device=dev)
    out = outputs()
    cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
    hail = _lib.PipeHail(ptr(rh), ptr(rhs), ptr(rhs_out), ptr(inprh))

    def fn():
        return lib.ice3_adjust_rain_ice_hail(ctx, ncol, NLEV, ncol, DT, ctypes.byref(cin), ctypes.byref(cout),
                                             ctypes.byref(hail), None, sp)
    ms, _ = timed_steps(fn, stream, args.steps, args.warmup)
    stages = stage_times_of(fn, args.steps, True)
    # column water budget with hail: change of the column's water = surface precipitation
    w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    mass = g["rhodref"] * g["dzz"]
    col0 = (mass * (sum(g[k] for k in w) + rhs)).sum(0) * DT
    col1 = (mass * (sum(out[k] for k in w) + rhs_out)).sum(0) * DT
    prec = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"] + inprh) * DT
    budget = ((col1 - col0 - prec).abs() / col0).max().item()
    return {"columns_per_s": ncol / (ms * 1e-3), "ms_per_step": ms, "speedup_vs_stat_krr6": ms_stat / ms,
            "stages_ms": stages, "water_budget_max_rel_err": budget,
            "note": "fused step with KRR=7: k_adjust_sedim_hail (column walk), then the hail "
                    "microphysics k_micro<double, false, HAIL> (persistent, one wavefront per SIMD)"}


def timed_steps(fn, stream, steps, warmup):
    """Run fn (one step, returning a C return code) warmup times, synchronise, then `steps`
    times with a HIP event pair around each call on `stream`; returns (median, mean) ms.
    Return codes are collected and checked after the loop so the checks do not perturb it."""
    import statistics

    import torch
    from ice3_amd import _lib
    for _ in range(warmup):
        _lib.check(fn(), "warmup")
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    rcs = []
    for e0, e1 in ev:
        e0.record(stream)
        rcs.append(fn())
        e1.record(stream)
    torch.cuda.synchronize()
    for rc in rcs:
        _lib.check(rc, "timed step")
    ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    return statistics.median(ms), sum(ms) / len(ms)


MLP_FLOP_PER_POINT = 2.0 * (14 * 128 + 3 * 128 * 128 + 128 * 8)   # 103,936
MFMA_BF16_PEAK_TFLOPS = 2500.0
# Floor of the network stage.  The matrix cores run the padded widths (32 inputs, 32 outputs):
MLP_FLOP_EXECUTED = 2.0 * (32 * 128 + 3 * 128 * 128 + 128 * 32)   # 114,688 per point
MLP_ACTIVATIONS = 4 * 128                                         # SiLU evaluations per point
# Full-rate FP32 VALU lane operations per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9
# VALU issue slots per activation in k_mlp_infer's epilogue as compiled: v_exp_f32 and
# v_rcp_f32 (quarter rate, 4 slots each), 2 v_add_f32, 2 v_mul_f32, a bf16 convert and an
# accumulator read; with a single quarter-rate operation and the bias add, multiply and
# convert, 7.
EPILOGUE_SLOTS_AS_BUILT, EPILOGUE_SLOTS_MIN = 14, 7


def input_check(lib, g, ncol, stream, sp, steps):
    """Cost of the opt-in input check (ice3_count_nonfinite over the fused step's 26 input
    fields, one HBM pass), timed with HIP events on the stream it is launched on."""
    import torch
    from ice3_amd import _lib
    names = list(_lib.PIPE_IN_FIELDS)
    fields = (ctypes.c_void_p * len(names))(*[g[k].data_ptr() for k in names])
    rows = (ctypes.c_int * len(names))(*[g[k].shape[0] if g[k].dim() == 2 else 1 for k in names])
    counts = torch.zeros(len(names), dtype=torch.int64, device=g["th"].device)
    cp = ctypes.c_void_p(counts.data_ptr())

    def call():
        return lib.ice3_count_nonfinite(len(names), fields, rows, ncol, ncol, cp, sp)
    ms, _ = timed_steps(call, stream, steps, 1)
    nbytes = sum(g[k].numel() for k in names) * 8
    return {"kernel": "k_count_nonfinite", "fields": len(names), "ms": ms,
            "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
            "nonfinite_per_call": int(counts.sum().item()) // (steps + 1)}


def surrogate_variant(lib, ctx, dev, g, args, reps: int = 4):
    """The learned surrogate at 1024x1024 columns x 90 levels: the 512x512 slab of the
    headline workload followed by three more slabs of fresh generator columns (seeds
    SEED + 1001..1003).  FP64 step vs surrogate step (network + FP64 fallback), per-field
    error of the surrogate against FP64, fallback statistics, water budget, matrix-core
    roofline of the network kernel (103,936 FLOP per point)."""
    import torch

    from ice3_amd import _lib, synthetic
    from ice3_amd import surrogate as S
    from ice3_amd._fields import ptr
    if not os.path.exists(S.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights (" + os.path.relpath(S.DEFAULT_WEIGHTS, ROOT) + ")"}
    sg = S.TrainedSurrogate.load(S.DEFAULT_WEIGHTS, device=dev.index)
    per = g["th"].shape[1]
    G = {k: torch.empty((v.shape[0], per * reps) if v.dim() == 2 else (per * reps,), dtype=v.dtype, device=dev)
         for k, v in g.items()}
    for k, v in g.items():
        G[k][..., :per].copy_(v)
    for r in range(1, reps):
        f = synthetic.generate(per, NLEV, seed=SEED + 1000 + r, variant="pipeline", dt=DT)
        f["rhodref"] = f.pop("rho_dry_ref")
        for k, v in f.items():
            G[k][..., r * per:(r + 1) * per].copy_(torch.from_numpy(v))
        del f
    nlev, ncol = G["th"].shape
    npts = nlev * ncol

    def outputs():
        o = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
        o.update({k: torch.empty(ncol, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
        return o
    cin = _lib.PipeIn(*[ptr(G[k]) for k in _lib.PIPE_IN_FIELDS])
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    ref, out = outputs(), outputs()
    cref = _lib.PipeOut(*[ptr(ref[k]) for k in _lib.PIPE_OUT_FIELDS])
    cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
    work = torch.empty(lib.ice3_surrogate_work_bytes(ncol, nlev), dtype=torch.uint8, device=dev)
    sgc = torch.zeros(5, dtype=torch.int64, device=dev)
    h = sg.handle()

    def plain():
        return lib.ice3_adjust_rain_ice(ctx, _lib.PREC_FP64, ncol, nlev, ncol, DT, ctypes.byref(cin),
                                        ctypes.byref(cref), None, sp)

    def surrogate():
        return lib.ice3_adjust_rain_ice_surrogate(ctx, ncol, nlev, ncol, DT, ctypes.byref(cin), ctypes.byref(cout),
                                                  ctypes.byref(h), ptr(work), ptr(sgc), None, sp)

    ms_plain, _ = timed_steps(plain, stream, args.steps, args.warmup)
    ms_sg, _ = timed_steps(surrogate, stream, args.steps, args.warmup)
    _lib.check(lib.ice3_stage_timing(ctx, 1), "ice3_stage_timing")
    acc = [0.0] * 6
    ms4 = (ctypes.c_float * 6)()
    for _ in range(args.steps):
        _lib.check(surrogate(), "timed surrogate step")
        _lib.check(lib.ice3_stage_times(ctx, ms4), "ice3_stage_times")
        acc = [x + y for x, y in zip(acc, ms4)]
    _lib.check(lib.ice3_stage_timing(ctx, 0), "ice3_stage_timing")
    acc = [acc[0] + acc[1]] + acc[2:]  # ICE_ADJUST and the sedimentation are one kernel
    stages = dict(zip(("ice_adjust+sedimentation", "features", "network", "decode", "fallback"),
                      [x / args.steps for x in acc]))
    sgc.zero_()
    _lib.check(surrogate(), "surrogate step")
    torch.cuda.synchronize()
    c = sgc.cpu().tolist()
    err = {}
    for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci"):
        d = (out[k] - ref[k]).abs()
        scale = ref[k].abs().max().item()
        err[k] = {"max_abs": d.max().item(), "rms_abs": d.pow(2).mean().sqrt().item(),
                  "max_rel_to_field_max": d.max().item() / scale if scale > 0 else 0.0,
                  "rms_rel_to_field_rms": (d.pow(2).mean().sqrt() / ref[k].pow(2).mean().sqrt()).item()}
    w = ("rvs", "rcs", "rrs", "ris", "rss", "rgs")
    mass = G["rhodref"] * G["dzz"]
    col0 = (mass * sum(G[k] for k in w)).sum(0) * DT
    col1 = (mass * sum(out[k] for k in w)).sum(0) * DT
    pr = -1000.0 * (out["inprr"] + out["inprs"] + out["inprg"]) * DT
    budget = ((col1 - col0 - pr).abs() / col0).max().item()
    tflops = MLP_FLOP_PER_POINT * npts / (stages["network"] * 1e-3) / 1e12
    # what the network stage cannot go below on this hardware, and the step built on it
    mfma_ms = MLP_FLOP_EXECUTED * npts / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e3
    acts = MLP_ACTIVATIONS * npts
    ep_ms, ep_min_ms = (acts * n / VALU_LANE_OPS * 1e3 for n in (EPILOGUE_SLOTS_AS_BUILT, EPILOGUE_SLOTS_MIN))
    floor = {"network_mfma_ms": mfma_ms, "network_epilogue_valu_ms_as_built": ep_ms,
             "network_epilogue_valu_ms_one_transcendental": ep_min_ms,
             "network_measured_ms": stages["network"],
             "surrogate_step_floor_ms": stages["ice_adjust+sedimentation"] + max(mfma_ms, ep_min_ms)
             + stages["fallback"],
             "fp64_step_ms": ms_plain,
             "basis": "matrix-core time at the dense bf16 peak; SiLU epilogue at the full-rate FP32 VALU "
                      "issue rate (quarter-rate exp/rcp = 4 slots); step floor = the column walk and the "
                      "FP64 fallback as measured, plus the network at its floor, with features and decode "
                      "fused away"}
    return {"columns": ncol, "levels": nlev, "weights": os.path.relpath(S.DEFAULT_WEIGHTS, ROOT),
            "train_pairs": int(sg.meta.get("pairs", 0)), "budget_tol": sg.budget_tol, "neg_tol": sg.neg_tol,
            "fp64_ms_per_step": ms_plain, "fp64_columns_per_s": ncol / (ms_plain * 1e-3),
            "surrogate_ms_per_step": ms_sg, "surrogate_columns_per_s": ncol / (ms_sg * 1e-3),
            "speedup_vs_fp64": ms_plain / ms_sg, "stages_ms": stages,
            "points": npts, "network_points": c[0], "fallback": {"out_of_range": c[1], "negative_source": c[2],
                                                                 "water_residual": c[3], "no_hydrometeor": c[4]},
            "network_fraction": c[0] / npts, "water_budget_max_rel_err": budget, "error_vs_fp64": err,
            "floor": floor,
            "roofline": {"bound": "mfma", "kernel": "k_mlp_infer", "achieved": tflops,
                         "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                         "flop_per_point": MLP_FLOP_PER_POINT, "ms_per_launch": stages["network"]}}


def ice_adjust_alone(lib, ctx, dev, args, ncol: int = 256 * 256, ocnd2: bool = False):
    """ice3_ice_adjust alone, FP64, 256x256 columns x 90 levels of the "adjust" generator:
    19 inputs + 16 outputs (the IceAdjustJAX tuple) = 280 B/point + 8 B/column.  With ocnd2,
    ice3_ice_adjust_ocnd2 (the LOCND2 branch): the level heights are one more input and
    PICLDFR, PWCLDFR, PSSIO, PSSIU, PIFR five more outputs (328 B/point), and PICE_CLD_WGT
    one more per-column input (16 B/column)."""
    import torch

    from ice3_amd import _lib, synthetic
    from ice3_amd._fields import ptr
    f = synthetic.generate(ncol, NLEV, seed=20260002, variant="adjust", dt=DT)
    g = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    del f
    out = {k: torch.empty((NLEV, ncol), dtype=torch.float64, device=dev) for k in _lib.ADJ_OUT_FIELDS[:-1]}
    cin = _lib.AdjustIn(*[ptr(g[k]) for k in _lib.ADJ_IN_FIELDS])
    cout = _lib.AdjustOut(*[ptr(out[k]) for k in _lib.ADJ_OUT_FIELDS[:-1]], ctypes.c_void_p(None))
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    if ocnd2:
        z1 = synthetic.level_profile(NLEV)[2]
        zz = torch.from_numpy(z1).to(dev)[:, None].expand(NLEV, ncol).contiguous()
        wgt = torch.ones(ncol, dtype=torch.float64, device=dev)
        x_out = {k: torch.empty((NLEV, ncol), dtype=torch.float64, device=dev) for k in _lib.OCND2_OUT_FIELDS}
        cin2 = _lib.Ocnd2In(ptr(zz), ptr(wgt))
        cout2 = _lib.Ocnd2Out(*[ptr(x_out[k]) for k in _lib.OCND2_OUT_FIELDS])
        bpp, bpc, kernel = 328.0, 16.0, "k_ice_adjust_ocnd2"

        def call():
            return lib.ice3_ice_adjust_ocnd2(ctx, ncol, NLEV, ncol, DT, ctypes.byref(cin), ctypes.byref(cin2),
                                             ctypes.byref(cout), ctypes.byref(cout2), sp)
    else:
        bpp, bpc, kernel = 280.0, 8.0, "k_ice_adjust"

        def call():
            return lib.ice3_ice_adjust(ctx, _lib.PREC_FP64, ncol, NLEV, ncol, DT, ctypes.byref(cin),
                                       ctypes.byref(cout), sp)
    ms, _ = timed_steps(call, stream, args.steps, args.warmup)
    gbs = (bpp * ncol * NLEV + bpc * ncol) / (ms * 1e-3) / 1e9
    return {"columns_per_s": ncol / (ms * 1e-3), "ms_per_step": ms, "columns": ncol,
            "kernel": kernel, "bytes_per_point": bpp, "achieved_GBs": gbs,
            "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}


if __name__ == "__main__":
    main()
