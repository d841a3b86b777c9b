"""The allocation-free surrogate step (ice3_adjust_rain_ice_surrogate_noalloc, static device
scratch) against the same step with a caller-allocated work area
(ice3_adjust_rain_ice_surrogate), through the C ABI: bit-for-bit equal outputs, counters
and per-point call counts on two calls in a row, the second smaller and with a row stride
ld > ncol (the static scratch then holds the first call's contents); a problem larger than
the scratch is rejected with ICE3_ETOOLARGE before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch.device("cuda", 0)


def _surrogate(dev):
    import os
    from ice3_amd import surrogate as S
    if not os.path.exists(S.DEFAULT_WEIGHTS):
        pytest.skip("no committed surrogate weights")
    return S.TrainedSurrogate.load(S.DEFAULT_WEIGHTS, device=dev.index)


def _fields(dev, ncol, nlev, ld, seed):
    from ice3_amd import _lib, synthetic
    f = synthetic.generate(ncol, nlev, seed=seed, variant="pipeline")
    f["rhodref"] = f["rho_dry_ref"]
    g = {}
    for k in _lib.PIPE_IN_FIELDS:
        v = torch.as_tensor(f[k], device=dev)
        if v.dim() == 2 and ld != ncol:
            pad = torch.full((nlev, ld), -1.0, dtype=torch.float64, device=dev)
            pad[:, :ncol] = v
            v = pad
        g[k] = v.contiguous()
    return g


def _run(lib, ctx, sg, g, ncol, nlev, ld, dev, work):
    from ice3_amd import _lib
    from ice3_amd._fields import ptr
    o = {k: torch.full((nlev, ld), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
    o.update({k: torch.full((ncol,), -7.0, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
    calls = torch.full((nlev * ld,), -9, dtype=torch.int32, device=dev)
    cnt = torch.zeros(3, dtype=torch.int64, device=dev)
    sgc = torch.zeros(5, dtype=torch.int64, device=dev)
    st = _lib.Stats(ptr(calls), ptr(cnt))
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    cout = _lib.PipeOut(*[ptr(o[k]) for k in _lib.PIPE_OUT_FIELDS])
    h = sg.handle()
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if work is None:
        rc = lib.ice3_adjust_rain_ice_surrogate_noalloc(ctx, ncol, nlev, ld, 50.0, ctypes.byref(cin),
                                                        ctypes.byref(cout), ctypes.byref(h), ptr(sgc),
                                                        ctypes.byref(st), sp)
    else:
        rc = lib.ice3_adjust_rain_ice_surrogate(ctx, ncol, nlev, ld, 50.0, ctypes.byref(cin), ctypes.byref(cout),
                                                ctypes.byref(h), ptr(work), ptr(sgc), ctypes.byref(st), sp)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in o.items()}
    res.update(calls=calls.cpu().numpy(), counters=cnt.cpu().numpy(), sg_counters=sgc.cpu().numpy())
    return rc, res


def test_noalloc_surrogate_matches_work_area_entry_point(dev):
    from ice3_amd import _lib
    lib, ctx = _lib.load(), _lib.context(dev.index)
    sg = _surrogate(dev)
    cases = ((6000, 90, 6000, 20260601), (1000, 90, 1003, 20260602))  # second: smaller, ld > ncol
    for ncol, nlev, ld, seed in cases:
        need = lib.ice3_surrogate_work_bytes(ncol, nlev)
        assert need <= _lib.SURROGATE_SCRATCH_BYTES, need
        g = _fields(dev, ncol, nlev, ld, seed)
        work = torch.empty(need, dtype=torch.uint8, device=dev)
        rc_w, ref = _run(lib, ctx, sg, g, ncol, nlev, ld, dev, work)
        rc_s, got = _run(lib, ctx, sg, g, ncol, nlev, ld, dev, None)
        assert rc_w == _lib.ICE3_OK and rc_s == _lib.ICE3_OK, (rc_w, rc_s)
        for k in ref:
            assert np.array_equal(got[k], ref[k], equal_nan=True), (ncol, k)
        assert ref["sg_counters"].sum() == ncol * nlev and ref["sg_counters"][0] > 0
        print(f"noalloc == work-area entry point at {ncol}x{nlev} (ld {ld}): network {ref['sg_counters'][0]}, "
              f"fallback {ref['sg_counters'][1:].sum()}, scratch {need / 2**20:.1f} MiB")

    # a problem larger than the static scratch: rejected before any launch, outputs untouched
    ncol, nlev = 8000, 90
    assert lib.ice3_surrogate_work_bytes(ncol, nlev) > _lib.SURROGATE_SCRATCH_BYTES
    g = _fields(dev, ncol, nlev, ncol, 20260603)
    rc, got = _run(lib, ctx, sg, g, ncol, nlev, ncol, dev, None)
    assert rc == _lib.ICE3_ETOOLARGE
    assert b"scratch" in lib.ice3_strerror(rc)
    assert all(np.all(got[k] == -7.0) for k in _lib.PIPE_OUT_FIELDS)
