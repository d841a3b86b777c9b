"""ice3_amd -- MI355X (gfx950) implementation of the ICE3 microphysics column kernels.

Drop-in components for the reference's JAX API (src/ice3/jax/ice_adjust.py,
src/ice3/jax/rain_ice.py): ``IceAdjust`` (same arguments and 16-tuple result as
``IceAdjustJAX``), ``RainIce`` (same ``(state, dt) -> (state, diagnostics)`` call as
``RainIceJAX``) and ``AdjustRainIce`` (both, fused on the device); the ``*OCND2``
variants run PHYEX's PARAM_ICE LOCND2 branch of the Fortran, the ``*Hail`` variants its
KRR=7 (HCLOUD='ICE4') hail configuration.  Arithmetic is
the PHYEX CY50T1 Fortran's, in FP64 (``precision="fp64"``) or with FP32 process
rates (``precision="mixed"``); everything runs in hand-written HIP kernels from
libice3_hip.so, called through its C ABI (include/ice3.h).
"""
from .ice_adjust import IceAdjust, IceAdjustHail, IceAdjustOCND2
from .rain_ice import (AdjustRainIce, AdjustRainIceHail, AdjustRainIceOCND2, RainIce, RainIceHail,
                       RainIceOCND2)

__all__ = ["IceAdjust", "IceAdjustOCND2", "IceAdjustHail", "RainIce", "RainIceOCND2",
           "RainIceHail", "AdjustRainIce", "AdjustRainIceOCND2", "AdjustRainIceHail"]
