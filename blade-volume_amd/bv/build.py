"""Builds libbvhip.so in-tree for gfx950 with hipcc.

Each source compiles to its own object in parallel (under _build/, outside
version control), then one hipcc call links the shared library.

Flags:
  -ffp-contract=off                          no a*b+c fusion: plane tests round as the reference's FP32 code
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division / sqrt (the reference's f32 `/` and sqrt)
  -munsafe-fp-atomics                        hardware global float atomics for gradient scatter
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
OBJ_DIR = os.path.join(os.path.dirname(PKG_DIR), "_build", "bvhip")
SOURCES = ("bv_scene.hip", "bv_trace.hip", "bv_walk.hip", "bv_adjoint.hip", "bv_server.hip", "bv_bvh.hip",
           "bv_splat.hip", "bv_gauss.hip", "bv_cech.hip", "bv_compact.hip", "bv_graph.hip", "bv_delaunay.hip")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


# Host-side AddressSanitizer build (the device code is compiled as usual):
# every -fsanitize= applies to the host pass only.
ASAN_HOST_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
ASAN_LIB = os.path.join(os.path.dirname(PKG_DIR), "_build", "asan", "libbvhip_asan.so")


def _compile(src: str, asan: bool = False) -> str:
    obj_dir = os.path.join(os.path.dirname(OBJ_DIR), "bvhip_asan") if asan else OBJ_DIR
    os.makedirs(obj_dir, exist_ok=True)
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIP_FLAGS, *(ASAN_HOST_FLAGS if asan else []), "-c", "-o", obj, os.path.join(CSRC, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} ({r.returncode}):\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int = 0, asan: bool = False) -> str:
    """Compile every source into lib/libbvhip.so (always recompiles).  With
    asan, a host-AddressSanitizer build goes to _build/asan/libbvhip_asan.so
    instead (load it through BV_LIBRARY with the sanitizer runtime preloaded)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, asan), sources))
    out = ASAN_LIB if asan else os.path.join(LIB_DIR, "libbvhip.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    # the sanitizer build links no runtime: its __asan symbols bind to the
    # preloaded libclang_rt.asan when the library is loaded
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{r.stderr}")
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import sys

    print(build(verbose=True, asan="--asan" in sys.argv[1:]))
