"""GPU: calc_fields and VTU output (reference src/average.cpp:84-112, src/output.cpp:29-214).

The element fields come from a gfx950 kernel with the reference's summation
order, so they equal the oracle's calc_fields bit for bit on the same
displacements and history, and the .vtu file micropp<3>::output writes equals
the oracle's write_vtu of the GP's own u_k and vars_n byte for byte.  The
oracle's writer is itself pinned byte for byte to the compiled reference
(tests/test_oracle.py::test_vtu_bytewise).
"""
import os

import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu

LAWS = {
    "elastic": [(0, 1e6, 0.3), (0, 1e7, 0.25), (0, 1e7, 0.25)],
    "plastic": [(1, 1e6, 0.3, 1e5, 2e3), (0, 1e7, 0.25), (0, 1e7, 0.25)],
    "damage": [(2, 1e6, 0.3, 0, 0, 1e5), (0, 1e7, 0.25), (0, 1e7, 0.25)],
}
SIZE = (5, 4, 6)


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def rand_hist(rng, n, kind):
    h = np.zeros(n)
    if kind == "plastic":
        h[:] = rng.normal(0, 1e-4, n)
        h[6::7] = np.abs(rng.normal(0, 1e-3, n // 7))
    elif kind == "damage":
        h[0::7] = 1e5 / np.sqrt(1e6) * (1 + np.abs(rng.normal(0, 0.2, n // 7)))
        h[1::7] = rng.uniform(0, 0.5, n // 7)
    return h


@pytest.mark.parametrize("kind", list(LAWS))
def test_elem_fields_bitwise(mp, oracle, kind):
    rng = np.random.default_rng(11)
    rv = mp.Rve(SIZE, mp.MIC_SPHERE, (0.3, 0, 0, 0), [mp.material(*s) for s in LAWS[kind]])
    sol = mp.Solver(rv)
    orve = oracle.rve(SIZE, mp.MIC_SPHERE, (0.3, 0, 0, 0), [ob.material(*s) for s in LAWS[kind]])
    u = rng.normal(0, 1e-3, rv.nrow)
    hist = rand_hist(rng, rv.nvars, kind) if kind != "elastic" else None
    ud = mp.DeviceArray.from_numpy(u)
    hd = mp.DeviceArray.from_numpy(hist) if hist is not None else None
    for h, hdev in ((None, None), (hist, hd)):
        es, ss = sol.elem_fields(ud, hdev)
        eo, so = orve.calc_fields(u, h)
        assert np.array_equal(es, eo) and np.array_equal(ss, so)


@pytest.mark.parametrize("kind", ["plastic", "damage"])
def test_output_vtu_bytewise(mp, oracle, kind, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    cp = [mp.FE_ONE_WAY, mp.FE_FULL, mp.FE_LINEAR]
    kw = dict(lin_stress=0, nr_max_its=4, nr_rel_tol=1e-6, cg_rel_tol=1e-10, cg_max_its=3000)
    m = mp.Micropp3(SIZE, mp.MIC_SPHERE, (0.3, 0, 0, 0), [mp.material(*s) for s in LAWS[kind]], ngp=3, coupling=cp,
                    **kw)
    orve = oracle.rve(SIZE, mp.MIC_SPHERE, (0.3, 0, 0, 0), [ob.material(*s) for s in LAWS[kind]])
    amp = {"plastic": 6e-3, "damage": 0.05}[kind]
    d = np.random.default_rng(3).uniform(-1, 1, 6) * amp
    nl = 0
    for t in range(2):
        for g in range(3):
            m.set_strain(g, d * (t + 1) / 2 * (1 + 0.3 * g))
        m.homogenize()
        for g in range(2):
            m.output(g, f"gpu_{g}_{t}")
            u_k, vars_n = m.state(g, 1), m.state(g, 2)
            nl += vars_n is not None
            ref = orve.write_vtu(u_k, vars_n, tmp_path / f"orc_{g}_{t}")
            a, b = open(f"gpu_{g}_{t}.vtu", "rb").read(), open(ref, "rb").read()
            assert a == b, f"gp {g} step {t}: first difference at byte " \
                           f"{next(j for j in range(min(len(a), len(b))) if a[j] != b[j])}"
        m.update_vars()
    assert nl > 0, "the case must write a GP with history"
    m.output2(1, 42, 7)
    assert open("micropp-42-7.vtu", "rb").read() == open(orve.write_vtu(m.state(1, 1), m.state(1, 2),
                                                                         tmp_path / "orc_o2"), "rb").read()
    # a GP without a displacement field (FE_LINEAR) reports instead of writing
    m.output(2, "none")
    assert not os.path.exists("none.vtu")
