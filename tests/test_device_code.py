"""Properties of the gfx950 device code itself, read from the assembly hipcc emits for each HIP
source with the library's own build flags (CPU only: hipcc cross-compiles).

Every kernel keeps its private data in its own registers or a small scratch frame of register
spills: no device-side calls (every helper is inlined), no flat load or store (private memory
is reached only by scratch instructions at fixed frame offsets, global memory only by global
instructions), and no pointer formed from the private aperture base.  A flat access is the one
kind of memory instruction that the hardware translates through the private and LDS
apertures, so code without any cannot raise an aperture violation."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dwarf-p-ice3_amd"))

from ice3_amd import build as B  # noqa: E402

HIP_SOURCES = [s for s in B.SOURCES if s.endswith(".hip")]


def _assembly(src: str, tmp_path) -> str:
    if not os.path.exists(B.HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    out = tmp_path / (os.path.splitext(src)[0] + ".s")
    cmd = [B.HIPCC, *B.FLAGS, *B.SOURCE_FLAGS.get(src, []), "--cuda-device-only", "-S",
           os.path.join(B.CSRC, src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def _kernels(asm: str) -> dict:
    """Kernel name -> its .amdhsa_kernel descriptor block."""
    return {m.group(1): m.group(2) for m in
            re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S)}


@pytest.mark.parametrize("src", HIP_SOURCES)
def test_no_calls_no_flat_access(src, tmp_path):
    asm = _assembly(src, tmp_path)
    ks = _kernels(asm)
    assert ks, "no kernels found in the assembly"
    code = [ln.strip() for ln in asm.splitlines() if ln.startswith("\t") and not ln.startswith("\t.")]
    flat = [ln for ln in code if ln.startswith("flat_")]
    calls = [ln for ln in code if ln.startswith(("s_swappc", "s_setpc", "s_call"))]
    aperture = [ln for ln in code if "src_private_base" in ln or "src_shared_base" in ln]
    assert not flat, flat[:5]
    assert not calls, calls[:5]
    assert not aperture, aperture[:5]
    for name, desc in ks.items():
        assert ".amdhsa_uses_dynamic_stack 0" in desc, name
        frame = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", desc).group(1))
        # register spills only: 24 bytes per lane in k_micro, at most 12 in the split
        # sedimentation's instances (the one for 193 to 256 levels with hail)
        assert frame <= 128, (name, frame)
