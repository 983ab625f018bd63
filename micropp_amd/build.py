"""Build micropp_amd/libmicropp_amd.so for gfx950 (MI355X), in-tree.

HIP sources (csrc/*.hip) go through hipcc for gfx950; host-only C++
(csrc/*.cpp) through g++.  Contraction is off by default (the FP64 path
reproduces the reference arithmetic bit for bit); packed FP32 instructions
are disabled for every HIP source.

Sanitizer builds (host code only; the GPU code is built as usual):
libmicropp_amd_asan.so (AddressSanitizer) and libmicropp_amd_tsan.so
(ThreadSanitizer), whose host code of every source is instrumented; the
.cpp sources then go through clang++, so that one sanitizer runtime (the
one the executable links, e.g. tests/drivers/sanitize_driver.cpp) serves
the whole process.

    python -m micropp_amd.build                     # or: python micropp_amd/build.py
    python -m micropp_amd.build --sanitize address  # or thread
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libmicropp_amd.so")
ARCH = os.environ.get("MICROPP_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
             f"-I{INC}", f"-I{CSRC}"]
# per-file contraction override (kernels whose arithmetic is not pinned bit-for-bit)
FAST_CONTRACT = {"mixed.hip"}
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INC}", f"-I{CSRC}",
             "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


CLANGXX = "/opt/rocm/llvm/bin/clang++"
SANITIZERS = {"address": "asan", "thread": "tsan"}


def lib_path(sanitize: str | None = None) -> str:
    return LIB if not sanitize else os.path.join(HERE, f"libmicropp_amd_{SANITIZERS[sanitize]}.so")


def _compile(src: str, sanitize: str | None = None) -> str:
    bdir = BUILD if not sanitize else os.path.join(BUILD, SANITIZERS[sanitize])
    obj = os.path.join(bdir, os.path.basename(src) + ".o")
    san = [f"-fsanitize={sanitize}", "-fno-omit-frame-pointer", "-g"] if sanitize else []
    if src.endswith(".hip"):
        flags = list(HIP_FLAGS)
        if os.path.basename(src) in FAST_CONTRACT:
            flags[flags.index("-ffp-contract=off")] = "-ffp-contract=fast"
        host = [x for f in san for x in ("-Xarch_host", f)]  # host code only
        cmd = [HIPCC] + flags + host + ["-c", src, "-o", obj]
    elif sanitize:
        cmd = [CLANGXX] + CXX_FLAGS + san + ["-c", src, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, sanitize: str | None = None) -> str:
    if sanitize and sanitize not in SANITIZERS:
        raise ValueError(f"sanitize must be one of {sorted(SANITIZERS)}")
    os.makedirs(BUILD if not sanitize else os.path.join(BUILD, SANITIZERS[sanitize]), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda f: _compile(f, sanitize), srcs))
    out = lib_path(sanitize)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs + [
        "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="build micropp_amd for gfx950")
    ap.add_argument("--sanitize", choices=sorted(SANITIZERS), default=None)
    build(verbose=True, sanitize=ap.parse_args().sanitize)
    sys.exit(0)
