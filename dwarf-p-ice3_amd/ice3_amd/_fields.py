"""Array plumbing shared by the drop-in components: layout checks and pointers."""
from __future__ import annotations

import ctypes

import torch


def as_columns(x, layout: str, *, device, name: str) -> torch.Tensor:
    """Return a contiguous float64 CUDA tensor of shape (nlev, ncol).

    layout "columns": x is (nlev, ncol) with columns contiguous, the reference's
    Fortran (NIJT, NKT) memory (no copy when already float64 on the device).
    layout "jax": x has the JAX components' (..., nz) shape, vertical index last;
    it is transposed to (nz, ncol).
    """
    t = torch.as_tensor(x, dtype=torch.float64, device=device)
    if layout == "columns":
        if t.dim() != 2:
            raise ValueError(f"{name}: expected a (nlev, ncol) array, got {tuple(t.shape)}")
        return t.contiguous()
    if t.dim() < 2:
        raise ValueError(f"{name}: expected a (..., nz) array, got {tuple(t.shape)}")
    return t.reshape(-1, t.shape[-1]).t().contiguous()


def from_columns(t: torch.Tensor, layout: str, shape) -> torch.Tensor:
    """Inverse of as_columns for results: a (nlev, ncol) field back to ``shape``, or a
    per-column (ncol,) result to ``shape[:-1]`` (the JAX layout without its vertical axis)."""
    if layout == "columns":
        return t
    if t.dim() == 1:
        return t.reshape(tuple(shape[:-1]))
    return t.t().reshape(shape)


LAYOUTS = ("auto", "columns", "jax")


def check_layout(layout: str) -> str:
    if layout not in LAYOUTS:
        raise ValueError(f"layout must be one of {LAYOUTS}, got {layout!r}")
    return layout


def resolve_layout(layout: str, x) -> str:
    """The layout of one call: "auto" reads it from the array rank of ``x`` (a field of the
    call): three or more dimensions are the JAX components' (nx, ny, nz) /
    (ngpblks, nproma, nflevg) arrays, vertical index last; two are the Fortran
    (nlev, ncol) memory, columns contiguous."""
    if layout != "auto":
        return layout
    nd = getattr(x, "ndim", None)
    if nd is None:
        nd = torch.as_tensor(x).dim()
    return "jax" if nd >= 3 else "columns"


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(None)


def stream_ptr(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
