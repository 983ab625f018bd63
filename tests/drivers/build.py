"""Build the C and Fortran drivers of the micropp3_* interface into tests/drivers/bin/.

c_driver.c (gcc) and f_driver.f90 (ROCm flang, when present, with the
drop-in modules micropp_amd/fortran/libmaterial.f95 and libmicropp.f95) link
micropp_amd/libmicropp_amd.so with an rpath relative to the binary, so the
built tree can move.  sanitize_driver.cpp links a sanitizer build of the
library (clang++, the same -fsanitize).  Run after the library is built:

    python tests/drivers/build.py            # C and Fortran drivers
    python tests/drivers/build.py address    # sanitize_driver_asan (or: thread)
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
LIBDIR = os.path.join(ROOT, "micropp_amd")
FORTRAN = os.path.join(LIBDIR, "fortran")
RPATH = "-Wl,-rpath,$ORIGIN/../../../micropp_amd"


def flang():
    for c in ("/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang", shutil.which("flang") or ""):
        if c and os.path.exists(c):
            return c
    return None


CLANGXX = "/opt/rocm/llvm/bin/clang++"


def build_sanitize(sanitize: str) -> str:
    """sanitize_driver.cpp against libmicropp_amd_<asan|tsan>.so (python -m micropp_amd.build --sanitize ...),
    with the same sanitizer, whose runtime clang links into the executable."""
    tag = {"address": "asan", "thread": "tsan"}[sanitize]
    os.makedirs(BIN, exist_ok=True)
    exe = os.path.join(BIN, f"sanitize_driver_{tag}")
    cmd = [CLANGXX, "-std=c++17", "-O1", "-g", f"-fsanitize={sanitize}", "-fno-omit-frame-pointer", "-Wno-non-c-typedef-for-linkage",
           f"-I{ROOT}/include", "-o", exe, os.path.join(HERE, "sanitize_driver.cpp"), f"-L{LIBDIR}",
           f"-lmicropp_amd_{tag}", RPATH]
    subprocess.run(cmd, check=True)
    return exe


def build() -> dict:
    os.makedirs(BIN, exist_ok=True)
    out = {}
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Werror", f"-I{ROOT}/include", "-o", os.path.join(BIN, "c_driver"),
           os.path.join(HERE, "c_driver.c"), f"-L{LIBDIR}", "-lmicropp_amd", RPATH]
    subprocess.run(cmd, check=True)
    out["c"] = os.path.join(BIN, "c_driver")
    fc = flang()
    if fc:
        # the drop-in modules first (their .mod files into bin/), then the driver that uses them
        mods = [os.path.join(FORTRAN, f) for f in ("libmaterial.f95", "libmicropp.f95")]
        cmd = [fc, "-O2", "-o", os.path.join(BIN, "f_driver")] + mods + [os.path.join(HERE, "f_driver.f90"),
               "-module-dir", BIN, f"-L{LIBDIR}", "-lmicropp_amd", RPATH]
        subprocess.run(cmd, check=True, cwd=BIN)
        out["fortran"] = os.path.join(BIN, "f_driver")
    return out


if __name__ == "__main__":
    import sys

    print(build_sanitize(sys.argv[1]) if len(sys.argv) > 1 else build())
