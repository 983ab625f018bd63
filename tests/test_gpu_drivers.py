"""GPU: C and Fortran callers of the micropp3_* interface against the oracle and the Python binding.

tests/drivers/c_driver.c and f_driver.f90 (built by __graft_entry__.build()
into tests/drivers/bin/; the Fortran one through the drop-in modules
libmicropp / libmaterial of micropp_amd/fortran/) run a two-GP plastic sphere
RVE for three load steps through micropp3_new (subiterations on, 4 of them,
use_A0 off) and print flags, stress and ctan.
- Against the oracle (the CPU restatement of the reference, same calls):
  sigma and ctan within 1e-6 relative, the same non-linear / converged /
  subiterated flags, the cost within the +-1 CG iteration per linear solve a
  different reduction order allows.
- Against the same calls through micropp3_new_ext from Python: bit for bit
  (one library, deterministic kernels).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import bindings as ob

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


def oracle_records(oracle):
    mats = [ob.material(1, 1e6, 0.3, 1e5, 1e3), ob.material(0, 1e7, 0.25), ob.material(0, 1e7, 0.25)]
    p, k = ob.make_params(6, ob.MIC_SPHERE, (0.25, 0, 0, 0), mats, ngp=2, coupling=[ob.FE_FULL, ob.FE_ONE_WAY],
                          subiterations=1, nsubiterations=4)
    m = oracle.micro(p, k)
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
    m.close()
    return recs


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", ["c_driver", "f_driver"])
def test_driver_matches_oracle(mp, oracle, name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    print(f"running {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, ref = parse(r.stdout), oracle_records(oracle)
    assert len(got) == len(ref) == 6
    for a, b in zip(got, ref):
        es, ec = rel(a["sig"], b["sig"]), rel(a["ctan"], b["ctan"])
        print(f"{name} step/gp {a['flags'][:2]}: flags {a['flags'][2:]} oracle {b['flags'][2:]}; sigma {es:.2e} "
              f"ctan {ec:.2e}", flush=True)
        assert es < 1e-6 and ec < 1e-6
        fa, fb = a["flags"], tuple(int(v) for v in b["flags"])
        assert fa[:3] == fb[:3] and fa[4:] == fb[4:]
        assert abs(fa[3] - fb[3]) <= 4 * 7 * 5  # +-1 per CG solve: 7 Newton solves of <= 4, <= 5 subiterations


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
