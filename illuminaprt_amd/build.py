"""Builds libilm_gpu.so in-tree with hipcc for gfx950, then the C++ plug-in.

Every source is compiled as HIP (device + host) into illuminaprt_amd/_build/
and linked into illuminaprt_amd/libilm_gpu.so. The packed-FP32 feature is
disabled for all device code. The plug-in layer (plugin/ilm_plugin.cpp, host
C++ over the C ABI) is built with g++ into libilm_plugin.so, next to the
sandbox program that loads it (ilm_plugin_sandbox).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libilm_gpu.so")
PLUGIN = os.path.join(HERE, "libilm_plugin.so")
SANDBOX = os.path.join(HERE, "ilm_plugin_sandbox")
PLUGIN_SRC = os.path.join(HERE, "plugin")
ARCH = "gfx950"

SOURCES = ["api.cpp", "bvh_build.cpp", "scenegen.cpp", "objfile.cpp", "persistent.cpp", "stats.cpp", "trace.hip", "lbvh.hip", "space.hip", "inst.hip", "post.hip", "pt.hip", "nrc.hip", "icache.hip"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


# Per-source flags. The cache MLP's kernels keep their MFMA accumulators in
# VGPRs instead of AGPRs: their occupancy is set by LDS, so the registers are
# there, and results in AGPRs must be copied out before every ReLU and store.
EXTRA_FLAGS = {"nrc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, *EXTRA_FLAGS.get(src, []), "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if noise:
            sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_plugin(verbose)
    return LIB


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_plugin(verbose: bool = False) -> None:
    """The plug-in links libilm_gpu.so and finds it through $ORIGIN. The
    sandbox links the plug-in's classes (as the reference's sandbox links its
    core library) and still resolves the plug-in's entry symbols with dlopen."""
    cxx = shutil.which("g++") or "g++"
    flags = ["-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include")]
    tmp = PLUGIN + ".tmp"
    _run([cxx, *flags, "-fPIC", "-shared", os.path.join(PLUGIN_SRC, "ilm_plugin.cpp"),
          os.path.join(PLUGIN_SRC, "ilm_environment.cpp"), "-o", tmp,
          "-L", HERE, "-lilm_gpu", "-Wl,-rpath,$ORIGIN", "-lpthread"], "plug-in build")
    os.replace(tmp, PLUGIN)
    tmp = SANDBOX + ".tmp"
    _run([cxx, *flags, os.path.join(PLUGIN_SRC, "plugin_sandbox.cpp"), "-o", tmp,
          "-L", HERE, "-lilm_plugin", "-lilm_gpu", "-Wl,-rpath,$ORIGIN", "-ldl"], "sandbox build")
    os.replace(tmp, SANDBOX)
    if verbose:
        print(f"built {PLUGIN} and {SANDBOX}")


if __name__ == "__main__":
    build(verbose=True)
