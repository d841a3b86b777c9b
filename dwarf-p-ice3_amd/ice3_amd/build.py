"""In-tree build of libice3_hip.so for gfx950 (hipcc, no JIT cache)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(HERE, "_lib", "libice3_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ("ice3_kernels.hip", "ice3_mlp.hip", "ice3_params.cpp")
# -ffp-contract=off keeps FP64 a*b+c as two roundings, the sequence of the CPU reference.
# Packed FP32 instructions (v_pk_*_f32) are disabled for the device code: some of their
# forms have given wrong results on this hardware without any error.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]
# Per-source flags.
# -disable-machine-licm for the physics kernels: machine-level loop-invariant code motion
# hoists the FP64 constants of every inlined exp/log out of k_micro's persistent loop into
# VGPRs, which then spill to scratch under the 256-register budget of two wavefronts per
# SIMD (each polynomial step of a log became a dependent scratch reload); rebuilding the
# constants at their uses is cheaper (k_micro<double>: 256 VGPRs + 104 B/lane of spills ->
# 216 VGPRs, none; -10 %).  The matrix-core kernels (ice3_mlp.hip) keep it: their hoisted
# LDS addresses pay.
# The iterative ILP scheduler, for both: the kernels run at a fixed two (k_micro), four
# (k_adjust_sedim) or one (k_mlp_infer) wavefronts per SIMD and hide latency mostly within
# a wavefront (k_micro -7 %, k_mlp_infer -5 %; results bit for bit the same).
# -fapprox-func for the physics kernels' device code: divisions become a reciprocal and
# Newton steps (FP64: two steps and a residual correction, within 1 ulp and in practice
# correctly rounded; FP32: the hardware reciprocal, 1 ulp) instead of the scaled IEEE
# sequence.  The FP64 outputs of the fused 512x512x90 step are bitwise those of the IEEE
# build, 6 % faster; the mixed path is 8 % faster.  Every divisor on the path is a positive
# finite quantity or sits in a branch that excludes zero.
ILP = ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]
SOURCE_FLAGS = {"ice3_kernels.hip": ["-mllvm", "-disable-machine-licm", "-Xarch_device", "-fapprox-func",
                                     *ILP],
                "ice3_mlp.hip": ILP}


def build(verbose: bool = False) -> str:
    """Regenerate the constants header from the host builder, then compile every HIP
    source into ice3_amd/_lib/libice3_hip.so; returns its path."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_params_inc", os.path.join(CSRC, "gen_params_inc.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    gen.generate(os.path.join(CSRC, "ice3_params_gen.inc"))
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    objs = []

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stderr}")
        if verbose and r.stderr:
            print(r.stderr)

    for src in SOURCES:
        obj = os.path.join(os.path.dirname(OUT), os.path.splitext(src)[0] + ".o")
        run([HIPCC, *FLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    for obj in objs:
        os.remove(obj)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
