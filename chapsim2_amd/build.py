"""Build libchapsim_hip.so (HIP, gfx950) in-tree.

    python -m chapsim2_amd.build        # or chapsim2_amd.build.build()

Objects go to chapsim2_amd/_build/, the library to chapsim2_amd/lib/.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libchapsim_hip.so")
ARCH = os.environ.get("CHS_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "chs_plan.cpp",
    "chs_geom.cpp",
    "chs_exact.cpp",
    "chs_tune.hip",
    "chs_lineop.hip",
    "chs_lineop_x.hip",
    "chs_poisson.hip",
    "chs_flow.hip",
    "chs_api.hip",
    "chs_props.hip",
    "chs_mlp.hip",
]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 instructions (v_pk_*_f32) are disabled for all device code: the
# compiler forms them from plain float code and some of their forms have given
# silently wrong results on MI355X.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(SRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + DEVICE_FLAGS
    else:
        cmd = [HIPCC, "-c", "-x", "c++", path, "-o", obj] + COMMON + ["-D__HIP_PLATFORM_AMD__"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
