"""Build recipe of the native library hornsculpt_amd/libhsbem.so (gfx950).

Host C++20 sources and HIP kernels are compiled by hipcc in parallel and linked
into one shared library with the C ABI of include/hsbem.h. The library is built
in-tree so it travels with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libhsbem.so"
OBJECTS = PACKAGE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOSTCXX = os.environ.get("HSBEM_HOSTCXX", "/opt/rocm/llvm/bin/clang++")
COMMON = ["-std=c++20", "-O3", "-fPIC", "-fvisibility=hidden", "-I", str(ROOT / "include")]
# gfx950 kernels: no packed FP32 instructions, hardware FP atomics on coarse-grained memory.
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HOST_SOURCES = [
    "host/space.cpp", "host/rules.cpp", "host/adaptive.cpp", "host/plan.cpp", "host/mesh_check.cpp",
    "host/support_tree.cpp", "host/support_near_plan.cpp", "host/fmm_levels.cpp", "host/fmm_rotation.cpp", "host/fmm_transfer.cpp",
    "host/fmm_scaled.cpp", "host/acoustic_analysis.cpp", "host/translation_plan_cache.cpp", "host/frequency_schedule.cpp",
    "autotune.cpp", "solver.cpp", "near_operator.cpp", "fmm_operator.cpp", "server.cpp", "capi.cpp",
]
DEVICE_SOURCES = ["hip/assembly.hip", "hip/linalg.hip", "hip/field.hip", "hip/fmm.hip", "hip/fmm_hf.hip", "hip/noalloc.hip", "hip/gmres_graph.hip", "hip/estimator.hip"]


def _compile(source: str, device: bool) -> Path:
    src = CSRC / source
    obj = OBJECTS / (source.replace("/", "_") + ".o")
    if device:
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HOSTCXX, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    OBJECTS.mkdir(exist_ok=True)
    jobs = [(s, False) for s in HOST_SOURCES] + [(s, True) for s in DEVICE_SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as pool:
        objects = list(pool.map(lambda j: _compile(*j), jobs))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects), "-lpthread"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
