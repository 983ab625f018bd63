"""GPU: the restart file and the profiling log against files the compiled reference wrote.

The fixtures (tests/golden/make_restart_log.py) come from the reference micropp<3>: a rank-3 instance of a
4^3-node J2 RVE with four GPs (FE_FULL, FE_ONE_WAY, FE_LINEAR, FE_FULL) over three load steps, its restart
file after the second (micropp-restart-3-7.bin, src/output.cpp:216-262 and include/gp.hpp:107-124), its log
(micropp-profiling-3.log, src/output.cpp:264-282), and the third step of a fresh reference instance that read
that restart file.

- read_restart of the reference's file: the allocated flags and the device vars_n / u_n equal the file bit for
  bit, and write_restart of that state gives the same bytes; the next step then matches the reference's own
  restart continuation (sigma and ctan within 1e-6, flags equal, cost within the +-1 CG iteration per linear
  solve a different reduction order allows).
- write_restart after the same two steps on the GPU: the reference's layout and flags, values within 1e-6.
- the log over the three steps: the reference's header, line structure, GP ids, non-linear and converged
  flags, with each cost the instance's get_cost and within the tolerance above of the reference's (which the
  oracle reproduces exactly, FE_FULL stale addition included: tests/test_oracle.py).
"""
import json
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


@pytest.fixture(scope="module")
def fixture():
    with open(os.path.join(GOLDEN, "restart_log_ref.json")) as f:
        doc = json.load(f)
    with open(os.path.join(GOLDEN, doc["restart_file"]), "rb") as f:
        raw = f.read()
    return doc, raw


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def parse_restart(raw, ngp, nvars, nndim):
    """per GP: (allocated, vars_n, u_n) of the reference layout (1-byte bool, then nvars + nndim doubles)"""
    out, o = [], 0
    for _ in range(ngp):
        a = raw[o] != 0
        o += 1
        if a:
            v = np.frombuffer(raw, np.float64, nvars, o)
            u = np.frombuffer(raw, np.float64, nndim, o + 8 * nvars)
            o += 8 * (nvars + nndim)
            out.append((True, v, u))
        else:
            out.append((False, None, None))
    assert o == len(raw)
    return out


def instance(mp, doc, **kw):
    mats = [mp.material(*m) for m in doc["materials"]]
    return mp.Micropp3(doc["size"], doc["micro_type"], doc["geo"], mats, ngp=len(doc["coupling"]),
                       coupling=doc["coupling"], mpi_rank=doc["mpi_rank"], **kw)


def cost_tol(doc):
    """+-1 CG iteration per linear solve: a GP's step runs at most nr_max_its Newton solves for sigma and, under
    FE_FULL, as many for each of the six perturbations (the reference defaults: 4)."""
    return 4 * 7


def test_read_reference_restart(mp, fixture, tmp_path, monkeypatch):
    doc, raw = fixture
    monkeypatch.chdir(tmp_path)
    rid = doc["restart_id"]
    shutil.copy(os.path.join(GOLDEN, doc["restart_file"]), tmp_path / doc["restart_file"])
    m = instance(mp, doc)
    ngp = len(doc["coupling"])
    m.read_restart(rid)
    parsed = parse_restart(raw, ngp, m.nvars, m.nndim)
    for g, (a, v, u) in enumerate(parsed):
        assert m.is_non_linear(g) == int(a), g
        if a:
            assert np.array_equal(m.state(g, 2), v) and np.array_equal(m.state(g, 0), u), g
    m.write_restart(rid + 1)
    with open(tmp_path / f"micropp-restart-{doc['mpi_rank']}-{rid + 1}.bin", "rb") as f:
        assert f.read() == raw
    # the next load step from the restored state, against the reference's continuation
    t = doc["restart_after_step"] + 1
    for g in range(ngp):
        m.set_strain(g, doc["strains"][t][g])
    m.homogenize()
    for g, r in enumerate(doc["continuation"]):
        es, ec = rel(m.get_stress(g), r["sigma"]), rel(m.get_ctan(g).ravel(), r["ctan"])
        print(f"continuation gp {g}: sigma {es:.2e} ctan {ec:.2e} cost {m.get_cost(g)} vs {r['cost']}, flags "
              f"{(m.is_non_linear(g), int(m.has_converged(g)))} vs {(r['non_linear'], r['converged'])}", flush=True)
        assert es < 1e-6 and ec < 1e-6
        assert (m.is_non_linear(g), int(m.has_converged(g))) == (r["non_linear"], r["converged"])
        assert abs(m.get_cost(g) - r["cost"]) <= cost_tol(doc)
    m.close()


def test_write_restart_and_log_like_reference(mp, fixture, tmp_path, monkeypatch):
    doc, raw = fixture
    monkeypatch.chdir(tmp_path)
    ngp = len(doc["coupling"])
    m = instance(mp, doc, write_log=1)
    costs = []
    for t, rec in enumerate(doc["steps"]):
        for g in range(ngp):
            m.set_strain(g, doc["strains"][t][g])
        m.homogenize()
        costs.append([m.get_cost(g) for g in range(ngp)])
        for g, r in enumerate(rec):
            es = rel(m.get_stress(g), r["sigma"])
            print(f"step {t} gp {g}: sigma {es:.2e}, cost {m.get_cost(g)} vs reference {r['cost']}", flush=True)
            assert es < 1e-6
            assert (m.is_non_linear(g), int(m.has_converged(g))) == (r["non_linear"], r["converged"])
            assert abs(m.get_cost(g) - r["cost"]) <= cost_tol(doc)
        m.update_vars()
        if t == doc["restart_after_step"]:
            m.write_restart(doc["restart_id"])
            with open(tmp_path / doc["restart_file"], "rb") as f:
                mine = f.read()
            assert len(mine) == len(raw)
            for g, (a, b) in enumerate(zip(parse_restart(mine, ngp, m.nvars, m.nndim),
                                           parse_restart(raw, ngp, m.nvars, m.nndim))):
                assert a[0] == b[0], g
                if a[0]:
                    assert rel(a[1], b[1]) < 1e-6 and rel(a[2], b[2]) < 1e-6, g
    m.close()  # closes the log
    with open(tmp_path / f"micropp-profiling-{doc['mpi_rank']}.log") as f:
        mine = f.read().splitlines()
    ref = doc["log"].splitlines()
    assert len(mine) == len(ref) and mine[0] == ref[0]
    for t in range(len(doc["steps"])):
        o = 1 + t * (1 + ngp)
        assert mine[o] == ref[o] == f"log_id : {t}"
        for g in range(ngp):
            a, b = mine[o + 1 + g].split("\t"), ref[o + 1 + g].split("\t")
            assert a[0] == b[0] == "" and a[1] == b[1] == str(g)
            assert a[2] == b[2] and a[4] == b[4]  # non-linear, converged
            assert int(a[3]) == costs[t][g]
            assert abs(int(a[3]) - int(b[3])) <= cost_tol(doc)
