"""Build libwmles_amd.so (HIP, gfx950) in-tree.

    python wmles_jax_amd/build.py          # or build() from Python (no package import)

Every translation unit is compiled with packed FP32 disabled.  The ghost-fill /
geometry and reduction units are compiled without floating-point contraction so
that their results round exactly like the reference's float64 host arithmetic; the
RHS unit contracts multiply-adds into FMAs.  The FP64 and Hybrid FP32-bulk flux kernels
(flux_ilp.hip) are scheduled by LLVM's max-ILP machine scheduler.  The persistent stage
kernel (stage.hip) and its bodies (stage_off.hip, stage_fast.hip) are relocatable device
code, device-linked into the same library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwmles_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WMX_OFFLOAD_ARCH", "gfx950")

COMMON = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wno-unused-result"]

UNITS = {
    "ghost.hip": ["-ffp-contract=off"],
    "reduce.hip": ["-ffp-contract=off"],
    "rhs.hip": ["-ffp-contract=fast"],
    "flux.hip": ["-ffp-contract=fast"],
    "flux_ilp.hip": ["-ffp-contract=fast", "-mllvm", "-amdgpu-sched-strategy=max-ilp"],
    "flux3.hip": ["-ffp-contract=fast"],
    "mlp.hip": ["-ffp-contract=fast"],
    "ops.hip": ["-ffp-contract=fast"],
    "capi.hip": [],
    # the persistent stage kernel and its bodies: relocatable device code, each body unit
    # with the FP contraction of the unit whose kernels it shares code with
    "stage_off.hip": ["-ffp-contract=off", "-fgpu-rdc"],
    "stage_fast.hip": ["-ffp-contract=fast", "-fgpu-rdc"],
    "stage.hip": ["-fgpu-rdc"],
}


def _compile(unit, flags):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, *flags, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, len(UNITS))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB + ".tmp"
    # -fgpu-rdc --hip-link: the relocatable units are device-linked (with LTO; each
    # instruction keeps its unit's contraction) beside the self-contained ones
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fgpu-rdc", "--hip-link", "-shared", "-fPIC", "-o", tmp,
           *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
