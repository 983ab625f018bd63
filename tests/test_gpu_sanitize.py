"""GPU: the library's host code under AddressSanitizer.

tests/drivers/sanitize_driver.cpp, linked against libmicropp_amd_asan.so
(python -m micropp_amd.build --sanitize address: the host code of every
source instrumented, the GPU code as usual) and built by
__graft_entry__.build(), runs the host paths that allocate, copy between host
and device or run host threads: the batch of GP slots, concurrent solvers on
host threads, the multi-RHS solver with its coarse level, the mixed Newton,
restart files and VTU output.  Any AddressSanitizer report ends it with a
non-zero status (halt_on_error).  Leak detection is off (the HIP runtime's
own allocations at exit are not this library's).
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "drivers", "bin", "sanitize_driver_asan")


def test_host_code_under_address_sanitizer():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    if not os.path.exists(EXE):
        pytest.skip(f"{EXE} not built")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:protect_shadow_gap=0")
    env.pop("MPPD_GP_BATCH", None)
    print("running sanitize_driver_asan", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    lines = r.stdout.splitlines()
    assert lines[-1] == "done"
    got = dict(ln.split(": ", 1) for ln in lines if ": " in ln)
    # the batch of GP slots and the concurrent solvers give the same results bit for bit
    assert got["batch j2"] == got["concurrent j2"]
