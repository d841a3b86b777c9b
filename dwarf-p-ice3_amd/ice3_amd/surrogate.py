"""Learned surrogate of the ICE4 microphysics: training on the GPU and the fused step with it.

The surrogate is a per-point MLP (18 features -> 4 x 128 SiLU -> 8 outputs) whose
weights run in bf16 on the matrix cores with fp32 accumulation.  It is trained here,
on the GPU, by the library's own kernels (``ice3_mlp_forward`` / ``ice3_mlp_backward``
/ ``ice3_mlp_adam``), on state -> tendency pairs that the FP64 kernels produce
(``make_pairs``): the input of the microphysics stage of the fused step
(``ice3_surrogate_features``) and the change that stage makes to each source.

``AdjustRainIceSurrogate`` runs the fused step with the network in place of the
microphysics (``ice3_adjust_rain_ice_surrogate``); every point the network cannot
vouch for (features outside the training range, a water residual above ``budget_tol``,
a source negative beyond ``neg_tol``, or no hydrometeor) is recomputed by the plain FP64
kernel; accepted points pass the same negativity correction as the plain kernel's.

The surrogate has no counterpart in the reference; it is a build addition beside the
drop-in components.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from . import _lib, synthetic
from ._fields import ptr

NFEAT, NTARG = _lib.NFEAT, _lib.NTARG
FEATURES = ("T", "ln_p", "rhodref", "ln_rv", "ln_rc", "ln_rr", "ln_ri", "ln_rs", "ln_rg", "ln_1p_ci",
            "sigs", "exn", "dt", "cldfr", "sed_rr", "sed_ri", "sed_rs", "sed_rg")
TARGETS = ("d_ths", "d_rvs", "d_rcs", "d_rrs", "d_ris", "d_rss", "d_rgs", "d_ci")
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate.npz")


def sizes():
    lib = _lib.load()
    v = [ctypes.c_int() for _ in range(5)]
    _lib.check(lib.ice3_mlp_sizes(*[ctypes.byref(x) for x in v]), "ice3_mlp_sizes")
    return dict(zip(("n_params", "n_weights", "tile", "in_width", "out_width"), [x.value for x in v]))


def _sp(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


# ------------------------------------------------------------------ training pairs

def _pipe_buffers(f: dict, dev):
    g = {k: torch.as_tensor(np.ascontiguousarray(f[k]), dtype=torch.float64, device=dev)
         for k in _lib.PIPE_IN_FIELDS}
    nlev, ncol = g["th"].shape
    out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
    out.update({k: torch.empty(ncol, dtype=torch.float64, device=dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
    return g, out


def pairs_from_state(f: dict, dt: float = 50.0, device=None):
    """Features and targets of every microphysics point of one pipeline state.

    Runs ICE_ADJUST and the sedimentation (FP64 kernels), reads the surrogate features
    of the state the microphysics stage sees, runs the FP64 microphysics and returns
    (raw features (n, 14) float32, targets (n, 8) float64) as device tensors, where
    target v is the change of source v (th, rv, rc, rr, ri, rs, rg) and of ci."""
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
    lib = _lib.load()
    ctx = _lib.context(dev.index)
    g, out = _pipe_buffers(f, dev)
    nlev, ncol = g["th"].shape
    cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS])
    cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
    sp = _sp(dev)
    _lib.check(lib.ice3_adjust_rain_ice_stages(ctx, _lib.PREC_FP64, ncol, nlev, ncol, dt, ctypes.byref(cin),
                                               ctypes.byref(cout), None,
                                               _lib.STAGE_ADJUST | _lib.STAGE_SEDIMENTATION, sp),
               "adjust + sedimentation")
    npts = nlev * ncol
    raw = torch.empty((npts, NFEAT), dtype=torch.float32, device=dev)
    micro = torch.empty(npts, dtype=torch.uint8, device=dev)
    _lib.check(lib.ice3_surrogate_features(ctx, ncol, nlev, ncol, dt, ctypes.byref(cin), ctypes.byref(cout),
                                           None, ptr(raw), None, ptr(micro), sp), "features")
    src = torch.stack([out[k].reshape(-1).clone() for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs")]
                      + [g["ci"].reshape(-1).clone()], dim=1)
    _lib.check(lib.ice3_adjust_rain_ice_stages(ctx, _lib.PREC_FP64, ncol, nlev, ncol, dt, ctypes.byref(cin),
                                               ctypes.byref(cout), None, _lib.STAGE_MICROPHYSICS, sp),
               "microphysics")
    fin = torch.stack([out[k].reshape(-1) for k in ("ths", "rvs", "rcs", "rrs", "ris", "rss", "rgs", "ci")],
                      dim=1)
    keep = micro.bool()
    return raw[keep], (fin - src)[keep]


def make_pairs(n_pairs: int, ncol: int = 65536, seed0: int = 20260100, dt: float = 50.0, device=None,
               log=None):
    """At least n_pairs (features, targets) from successive seeds of the pipeline generator."""
    feats, targs, n, s = [], [], 0, seed0
    while n < n_pairs:
        f = synthetic.generate(ncol, 90, seed=s, variant="pipeline", dt=dt)
        f["rhodref"] = f.pop("rho_dry_ref")
        x, y = pairs_from_state(f, dt, device)
        feats.append(x)
        targs.append(y)
        n += x.shape[0]
        s += 1
        if log:
            log(f"pairs: seed {s - 1}: {x.shape[0]} points, total {n}")
    return torch.cat(feats)[:n_pairs], torch.cat(targs)[:n_pairs]


# ------------------------------------------------------------------ normalisation

def fit_norm(raw: torch.Tensor, targ: torch.Tensor) -> dict:
    """Feature standardisation and range; target transform u = asinh(y / unit),
    standardised.  unit = median |y| over the points where y != 0."""
    r = raw.double()
    mean = r.mean(0)
    std = r.std(0)
    scale = torch.where(std > 1e-12 * (1 + mean.abs()), 1.0 / std, torch.zeros_like(std))
    lo, hi = r.min(0).values, r.max(0).values
    unit = torch.empty(NTARG, dtype=torch.float64, device=targ.device)
    for v in range(NTARG):
        a = targ[:, v].abs()
        nz = a[a > 0]
        unit[v] = nz.median() if nz.numel() else 1.0
    u = torch.asinh(targ / unit)
    return {"feat_mean": mean.float().cpu().numpy(), "feat_scale": scale.float().cpu().numpy(),
            "feat_lo": lo.float().cpu().numpy(), "feat_hi": hi.float().cpu().numpy(),
            "targ_unit": unit.float().cpu().numpy(), "targ_mean": u.mean(0).float().cpu().numpy(),
            "targ_std": u.std(0).clamp_min(1e-6).float().cpu().numpy()}


def encode_features(raw: torch.Tensor, norm: dict) -> torch.Tensor:
    """(n, 14) raw features -> (n, 32) bf16 network inputs (same arithmetic as the kernels)."""
    dev = raw.device
    m = torch.as_tensor(norm["feat_mean"], device=dev)
    s = torch.as_tensor(norm["feat_scale"], device=dev)
    x = torch.zeros((raw.shape[0], 32), dtype=torch.float32, device=dev)
    x[:, :NFEAT] = (raw.float() - m) * s
    return x.to(torch.bfloat16)


def encode_targets(targ: torch.Tensor, norm: dict) -> torch.Tensor:
    dev = targ.device
    unit = torch.as_tensor(norm["targ_unit"], dtype=torch.float64, device=dev)
    mean = torch.as_tensor(norm["targ_mean"], dtype=torch.float64, device=dev)
    std = torch.as_tensor(norm["targ_std"], dtype=torch.float64, device=dev)
    return ((torch.asinh(targ / unit) - mean) / std).float()


def decode_targets(y: torch.Tensor, norm: dict) -> torch.Tensor:
    dev = y.device
    unit = torch.as_tensor(norm["targ_unit"], device=dev)
    mean = torch.as_tensor(norm["targ_mean"], device=dev)
    std = torch.as_tensor(norm["targ_std"], device=dev)
    return unit * torch.sinh(mean + std * y.float())


# ------------------------------------------------------------------ the network

class MLP:
    """Parameters (fp32 master + bf16 copies) and the training kernels' driver."""

    def __init__(self, device=None, seed: int = 0, params: np.ndarray | None = None):
        self.dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.lib = _lib.load()
        sz = sizes()
        self.n_params, self.n_weights = sz["n_params"], sz["n_weights"]
        if params is None:
            params = init_params(seed)
        assert params.shape == (self.n_params,)
        self.params = torch.as_tensor(params, dtype=torch.float32, device=self.dev).clone()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.wrm = torch.empty(self.n_weights, dtype=torch.bfloat16, device=self.dev)
        self.wpk = torch.empty(self.n_weights, dtype=torch.bfloat16, device=self.dev)
        self.step = 0
        self.sync()

    def sync(self):
        """Rewrite the bf16 copies from the fp32 parameters."""
        _lib.check(self.lib.ice3_mlp_adam(1, 0.0, 0.0, 0.0, 0.0, ptr(self.params), None, None, None,
                                          ptr(self.wrm), ptr(self.wpk), _sp(self.dev)), "ice3_mlp_adam")

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """Network outputs (n, 8) for bf16 inputs (n, 32)."""
        n = x.shape[0]
        y = torch.empty((n, NTARG), dtype=torch.float32, device=self.dev)
        _lib.check(self.lib.ice3_mlp_forward(n, ptr(self.wpk), ptr(self.params), ptr(x), None, ptr(y),
                                             None, None, None, None, _sp(self.dev)), "ice3_mlp_forward")
        return y

    def _buffers(self, n):
        if getattr(self, "_bn", 0) != n:
            tiles = (n + 127) // 128
            self.zsave = torch.empty((4, n, 128), dtype=torch.float32, device=self.dev)
            self.hsave = torch.empty((4, n, 128), dtype=torch.bfloat16, device=self.dev)
            self.dy = torch.empty((n, 32), dtype=torch.float32, device=self.dev)
            self.loss_part = torch.empty(tiles, dtype=torch.float32, device=self.dev)
            self.gpart = torch.empty((tiles, self.n_params), dtype=torch.float32, device=self.dev)
            self._bn = n

    def gradient(self, x: torch.Tensor, t: torch.Tensor):
        """Sum of squared errors and the gradient of the mean over the batch (into self.grad)."""
        n = x.shape[0]
        self._buffers(n)
        sp = _sp(self.dev)
        _lib.check(self.lib.ice3_mlp_forward(n, ptr(self.wpk), ptr(self.params), ptr(x), ptr(t), None,
                                             ptr(self.zsave), ptr(self.hsave), ptr(self.dy), ptr(self.loss_part),
                                             sp), "ice3_mlp_forward")
        _lib.check(self.lib.ice3_mlp_backward(n, ptr(self.wrm), ptr(x), ptr(self.zsave), ptr(self.hsave),
                                              ptr(self.dy), ptr(self.gpart), ptr(self.grad), sp),
                   "ice3_mlp_backward")
        return self.loss_part.sum()

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        _lib.check(self.lib.ice3_mlp_adam(self.step, lr, b1, b2, eps, ptr(self.params), ptr(self.m),
                                          ptr(self.v), ptr(self.grad), ptr(self.wrm), ptr(self.wpk),
                                          _sp(self.dev)), "ice3_mlp_adam")


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the 14 used inputs and the hidden layers, small output layer,
    zero biases; padded input rows and output columns are zero (and stay zero: their
    gradients vanish)."""
    rng = np.random.Generator(np.random.Philox(seed))
    dims = [(32, 128), (128, 128), (128, 128), (128, 128), (128, 32)]
    ws = []
    for l, (k, n) in enumerate(dims):
        fan_in = NFEAT if l == 0 else k
        lim = math.sqrt(6.0 / fan_in) * (0.5 if l == 4 else 1.0)
        w = rng.uniform(-lim, lim, (k, n))
        if l == 0:
            w[NFEAT:, :] = 0.0
        if l == 4:
            w[:, NTARG:] = 0.0
        ws.append(w.ravel())
    return np.concatenate(ws + [np.zeros(4 * 128 + 32)]).astype(np.float32)


def train(mlp: MLP, x: torch.Tensor, t: torch.Tensor, epochs: int = 4, batch: int = 32768,
          lr: float = 2e-3, lr_min: float = 5e-5, seed: int = 0, log=None, val=None):
    """Adam with cosine learning-rate decay over shuffled minibatches; returns the per-epoch
    mean squared error (normalised targets) and, with val=(x, t), the validation error."""
    n = x.shape[0]
    steps = epochs * (n // batch)
    gen = torch.Generator(device=x.device)
    gen.manual_seed(seed)
    hist = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device=x.device, generator=gen)
        tot = torch.zeros((), dtype=torch.float64, device=x.device)
        for b in range(n // batch):
            idx = perm[b * batch:(b + 1) * batch]
            tot += mlp.gradient(x[idx], t[idx]).double()
            lr_k = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, steps)))
            mlp.adam(lr_k)
            k += 1
        mse = tot.item() / ((n // batch) * batch * NTARG)
        rec = {"epoch": ep, "train_mse": mse}
        if val is not None:
            rec["val_mse"] = evaluate(mlp, *val)
        hist.append(rec)
        if log:
            log(rec)
    return hist


def evaluate(mlp: MLP, x: torch.Tensor, t: torch.Tensor, batch: int = 1 << 20) -> float:
    tot = 0.0
    for b in range(0, x.shape[0], batch):
        y = mlp.forward(x[b:b + batch])
        tot += ((y - t[b:b + batch]) ** 2).sum().item()
    return tot / (x.shape[0] * NTARG)


# ------------------------------------------------------------------ trained model

class TrainedSurrogate:
    """Trained parameters + normalisation, resident on one device."""

    def __init__(self, params: np.ndarray, norm: dict, budget_tol: float = 0.1, neg_tol: float = 0.25,
                 device=None, meta=None):
        self.mlp = MLP(device=device, params=params)
        self.dev = self.mlp.dev
        self.norm = {k: np.asarray(v, dtype=np.float32) for k, v in norm.items()}
        self.budget_tol = float(budget_tol)
        self.neg_tol = float(neg_tol)
        self.meta = dict(meta or {})
        cn = _lib.SurrogateNorm()
        for k in ("feat_mean", "feat_scale", "feat_lo", "feat_hi", "targ_unit", "targ_mean", "targ_std"):
            getattr(cn, k)[:] = [float(x) for x in self.norm[k]]
        raw = bytes(cn)
        self.norm_dev = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.dev)

    def handle(self) -> "_lib.Surrogate":
        return _lib.Surrogate(ptr(self.mlp.wpk), ptr(self.mlp.params), ptr(self.norm_dev), self.budget_tol,
                              self.neg_tol)

    def save(self, path: str):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.mlp.params.cpu().numpy(), budget_tol=self.budget_tol, neg_tol=self.neg_tol,
                 **{f"norm_{k}": v for k, v in self.norm.items()},
                 **{f"meta_{k}": np.asarray(v) for k, v in self.meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device=None):
        z = np.load(path, allow_pickle=False)
        norm = {k[5:]: z[k] for k in z.files if k.startswith("norm_")}
        meta = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}
        return cls(z["params"], norm, float(z["budget_tol"]), float(z["neg_tol"]), device=device, meta=meta)


class AdjustRainIceSurrogate:
    """The fused step with the surrogate microphysics (see the module docstring).

    Same inputs and result dict as ``AdjustRainIce``; ``last_counters`` holds the
    network / fallback statistics of the last call."""

    def __init__(self, surrogate: TrainedSurrogate):
        self.sg = surrogate
        self.dev = surrogate.dev
        self.lib = _lib.load()
        self.ctx = _lib.context(self.dev.index)
        self.last_counters = None

    def __call__(self, f: dict, dt: float, return_calls: bool = False):
        g = {k: base_col(f[k], self.dev) for k in _lib.PIPE_IN_FIELDS if k != "sigqsat"}
        nlev, ncol = g["th"].shape
        sq = torch.as_tensor(f["sigqsat"], dtype=torch.float64, device=self.dev).reshape(-1)[:ncol].contiguous()
        out = {k: torch.empty((nlev, ncol), dtype=torch.float64, device=self.dev) for k in _lib.PIPE_OUT_FIELDS[:15]}
        out.update({k: torch.empty(ncol, dtype=torch.float64, device=self.dev) for k in _lib.PIPE_OUT_FIELDS[15:]})
        work = torch.empty(self.lib.ice3_surrogate_work_bytes(ncol, nlev), dtype=torch.uint8, device=self.dev)
        sgc = torch.zeros(5, dtype=torch.int64, device=self.dev)
        counters = torch.zeros(3, dtype=torch.int64, device=self.dev)
        calls = torch.zeros(nlev * ncol, dtype=torch.int32, device=self.dev) if return_calls else None
        st = _lib.Stats(ptr(calls), ptr(counters))
        cin = _lib.PipeIn(*[ptr(g[k]) for k in _lib.PIPE_IN_FIELDS[:-1]], ptr(sq))
        cout = _lib.PipeOut(*[ptr(out[k]) for k in _lib.PIPE_OUT_FIELDS])
        h = self.sg.handle()
        with torch.cuda.device(self.dev):
            rc = self.lib.ice3_adjust_rain_ice_surrogate(self.ctx, ncol, nlev, ncol, float(dt), ctypes.byref(cin),
                                                         ctypes.byref(cout), ctypes.byref(h), ptr(work),
                                                         ptr(sgc), ctypes.byref(st), _sp(self.dev))
        _lib.check(rc, "ice3_adjust_rain_ice_surrogate")
        c = sgc.cpu().tolist()
        cc = counters.cpu().tolist()
        self.last_counters = {"network": c[0], "out_of_range": c[1], "negative": c[2], "water_residual": c[3],
                              "no_hydrometeor": c[4], "fallback_tendency_calls": cc[0], "capped_points": cc[1]}
        if cc[1]:
            raise _lib.Ice3Error(f"{cc[1]} fallback points hit the ICE4 sub-step cap")
        res = dict(out)
        if return_calls:
            res["calls"] = calls.view(nlev, ncol)
        return res


def base_col(x, dev):
    t = torch.as_tensor(x, dtype=torch.float64, device=dev)
    if t.dim() != 2:
        raise ValueError(f"expected (nlev, ncol) fields, got {tuple(t.shape)}")
    return t.contiguous()


def train_surrogate(n_pairs: int = 50_000_000, epochs: int = 6, batch: int = 32768, lr: float = 2e-3,
                    ncol: int = 262144, seed: int = 0, budget_tol: float = 0.1, neg_tol: float = 0.25,
                    device=None, log=print):
    """Generate n_pairs FP64 pairs, fit the normalisation, train; returns (TrainedSurrogate, history)."""
    import time
    t0 = time.time()
    raw, targ = make_pairs(n_pairs, ncol=ncol, device=device, log=log)
    log(f"pairs: {raw.shape[0]} in {time.time() - t0:.1f} s")
    norm = fit_norm(raw, targ)
    x = encode_features(raw, norm)
    t = encode_targets(targ, norm)
    del raw, targ
    nval = min(1 << 20, x.shape[0] // 20)
    xv, tv = x[:nval], t[:nval]
    mlp = MLP(device=device, seed=seed)
    t1 = time.time()
    hist = train(mlp, x[nval:], t[nval:], epochs=epochs, batch=batch, lr=lr, seed=seed, log=log, val=(xv, tv))
    torch.cuda.synchronize()
    log(f"training: {epochs} epochs over {x.shape[0] - nval} pairs in {time.time() - t1:.1f} s")
    sg = TrainedSurrogate(mlp.params.cpu().numpy(), norm, budget_tol=budget_tol, neg_tol=neg_tol, device=device,
                          meta={"pairs": x.shape[0], "epochs": epochs, "batch": batch, "lr": lr,
                                "val_mse": hist[-1]["val_mse"], "train_mse": hist[-1]["train_mse"]})
    return sg, hist


def main(argv=None):
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="train the rain_ice surrogate on the GPU")
    ap.add_argument("--pairs", type=int, default=50_000_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--ncol", type=int, default=262144)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)
    torch.cuda.set_device(0)
    sg, hist = train_surrogate(a.pairs, a.epochs, a.batch, a.lr, a.ncol, log=log)
    sg.save(a.out)
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
