"""GPU: C and Fortran callers of the micropp3_* interface give the same results as the Python binding.

tests/drivers/c_driver.c and f_driver.f90 (built by __graft_entry__.build()
into tests/drivers/bin/) run a two-GP plastic sphere RVE for three load
steps through micropp3_new (subiterations on, 4 of them, use_A0 off) and print
flags, stress and ctan; the same calls through micropp3_new_ext from Python
must reproduce them bit for bit (one library, deterministic kernels).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "drivers", "bin")


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def parse(text):
    recs = []
    for line in text.splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "flags":
            recs.append({"flags": tuple(int(v) for v in tok[1:])})
        elif tok[0] in ("sig", "ctan"):
            recs[-1][tok[0]] = np.array([float(v) for v in tok[1:]])
    return recs


def python_records(mp):
    mats = [mp.material(1, 1e6, 0.3, 1e5, 1e3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    m = mp.Micropp3(6, mp.MIC_SPHERE, (0.25, 0, 0, 0), mats, ngp=2, coupling=[mp.FE_FULL, mp.FE_ONE_WAY],
                    subiterations=1, nsubiterations=4, mpi_rank=0, use_A0=0, its_with_A0=1, write_log=0)
    recs = []
    for t in range(3):
        for g in range(2):
            eps = np.zeros(6)
            eps[0] = 1e-3 * (t + 1) * (g + 1)
            eps[3] = -4e-4 * (t + 1)
            m.set_strain(g, eps)
        m.homogenize()
        for g in range(2):
            recs.append({"flags": (t, g, m.is_non_linear(g), m.get_cost(g), int(m.has_converged(g)),
                                   int(m.has_subiterated(g))),
                         "sig": m.get_stress(g), "ctan": m.get_ctan(g).ravel()})
        m.update_vars()
    return recs


@pytest.mark.parametrize("name", ["c_driver", "f_driver"])
def test_driver_matches_python(mp, name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    print(f"running {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, ref = parse(r.stdout), python_records(mp)
    assert len(got) == len(ref) == 6
    for a, b in zip(got, ref):
        assert a["flags"] == tuple(int(v) for v in b["flags"])
        assert np.array_equal(a["sig"], b["sig"])
        assert np.array_equal(a["ctan"], b["ctan"])
    assert any(rec["flags"][2] for rec in got)  # the plastic matrix yields at some point
