"""GPU: micropp<3>::homogenize running FE GPs concurrently on several solvers.

When the FE GPs of an instance do not run as one batch (MPPD_GP_BATCH=0), the
reference's OpenMP loop over GPs (src/homogenize.cpp:72) becomes up to MPPD_GP_WORKERS
solvers, each with its own HIP stream, taking GPs in turn from host threads.
A GP's kernels and their order do not depend on which solver runs it, so the
bar against one solver (MPPD_GP_WORKERS=1) is bit-for-bit equality of sigma,
ctan, cost and flags, and of the device state u_n, u_k, vars_n, vars_k, step
after step -- with history allocation and subiterations on the way.
"""
import numpy as np
import pytest

from test_gpu_gp_batch import J2, KW, couplings, ramps, record, same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def make(mp, **kw):
    cp = couplings(mp)
    return mp.Micropp3(6, mp.MIC_SPHERE, (0.3, 0, 0, 0), [mp.material(*s) for s in J2], ngp=len(cp), coupling=cp,
                       **kw)


def step(m, rp, t, monkeypatch, workers, batch=None):
    """One load step; MPPD_GP_WORKERS / MPPD_GP_BATCH are read at an instance's first homogenize."""
    monkeypatch.setenv("MPPD_GP_WORKERS", str(workers))
    if batch is None:
        monkeypatch.delenv("MPPD_GP_BATCH", raising=False)
    else:
        monkeypatch.setenv("MPPD_GP_BATCH", str(batch))
    for g in range(m.ngp):
        m.set_strain(g, rp[g][t])
    m.homogenize()
    r = record(m)
    m.update_vars()
    return r


CASES = {
    "mixed": dict(mixed_precision=1, batch=0),
    "use_A0": dict(use_A0=1, its_with_A0=1, batch=0),
    "mixed_A0": dict(mixed_precision=1, use_A0=1, its_with_A0=1, batch=0),
    "fp64_no_batch": dict(batch=0),
}


@pytest.mark.parametrize("case", list(CASES))
def test_concurrent_gps_equal_one_solver(mp, monkeypatch, case):
    kw = dict(CASES[case])
    batch = kw.pop("batch", None)
    steps = 3
    one, many = make(mp, lin_stress=0, **KW, **kw), make(mp, lin_stress=0, **KW, **kw)
    rp = ramps(one.ngp, steps, seed=11)
    nl = sub = 0
    for t in range(steps):
        a = step(one, rp, t, monkeypatch, 1, batch)
        b = step(many, rp, t, monkeypatch, 3, batch)
        print(f"{case} step {t}: flags " + " ".join(str(r[2]) for r in a), flush=True)
        same(a, b, f"{case} step {t}")
        nl += sum(r[2][0] for r in a)
        sub += sum(r[2][3] for r in a)
    assert nl > 0, "the case must allocate history"
    if not case.startswith("mixed"):
        assert sub > 0, "the case must subiterate"
