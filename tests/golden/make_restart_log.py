"""Writes the restart and profiling-log fixtures from the compiled reference (oracle/_ref).

Run from the repository root after `make -C oracle ref` (the Micropp sources must be present):

    python tests/golden/make_restart_log.py

Outputs (committed):
- tests/golden/micropp-restart-3-7.bin: micropp<3>::write_restart(7) of a rank-3 instance after two load steps
  (src/output.cpp:216-262, include/gp.hpp:107-124): per GP a bool, then vars_n and u_n when it has history;
- tests/golden/restart_log_ref.json: the workload, the per-step sigma / ctan / cost / flags of that instance
  (three steps; the third after the restart was written), its profiling log micropp-profiling-3.log
  (write_log, src/output.cpp:264-282, header from src/micropp.cpp:173-182), and the third step of a second,
  fresh instance that read the restart file first (the reference's own restart continuation).

The workload: 4^3 nodes, MIC_SPHERE (r 0.3), J2 matrix (E 1e6, nu 0.3, Ka 1e5, Sy 1e4), elastic sphere (E 1e7,
nu 0.25); GPs FE_FULL / FE_ONE_WAY / FE_LINEAR / FE_FULL, the first two and the last strained past yield, the
third small; the reference's defaults otherwise (lin_stress on, CG 1e-5 / 1000, Newton 4 / 1e-3).
"""
import json
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bindings as ob  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")
RANK, RESTART_ID = 3, 7
SIZE = 4
GEO = (0.3, 0, 0, 0)
MATS = [(1, 1e6, 0.3, 1e5, 1e4, 0.0), (0, 1e7, 0.25, 0, 0, 0.0), (0, 1e7, 0.25, 0, 0, 0.0)]
COUPLING = [ob.FE_FULL, ob.FE_ONE_WAY, ob.FE_LINEAR, ob.FE_FULL]
DIRS = [[1.0, -0.3, 0.2, 0.5, 0.0, -0.2], [-0.4, 1.0, 0.1, 0.0, 0.6, 0.0], [1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
        [0.2, 0.3, -1.0, 0.0, 0.0, 0.4]]
AMPS = [2e-2, 1.5e-2, 1e-4, 1.2e-2]
STEPS = 3


def strains(t):
    """Strain of every GP at load step t (t = 0, 1, 2)."""
    return [list(np.array(DIRS[g]) * AMPS[g] * (t + 1) / STEPS) for g in range(len(COUPLING))]


def params():
    return ob.make_params(SIZE, ob.MIC_SPHERE, GEO, [ob.material(*m) for m in MATS], ngp=len(COUPLING),
                          coupling=COUPLING)


def record(m):
    return [{"sigma": m.get_stress(g).tolist(), "ctan": m.get_ctan(g).ravel().tolist(), "cost": int(m.get_cost(g)),
             "non_linear": int(m.is_non_linear(g)), "converged": int(m.has_converged(g))}
            for g in range(len(COUPLING))]


def main():
    ref = ob.Ref()
    cwd = os.getcwd()
    tmp = tempfile.mkdtemp()
    try:
        os.chdir(tmp)
        p, keep = params()
        a = ref.micro_rank(p, keep, rank=RANK, write_log=1)
        steps = []
        for t in range(STEPS):
            for g, e in enumerate(strains(t)):
                a.set_strain(g, e)
            a.homogenize()
            steps.append(record(a))
            a.update_vars()
            if t == STEPS - 2:
                a.write_restart(RESTART_ID)
        a.close()  # closes the log
        with open(f"micropp-profiling-{RANK}.log") as f:
            log = f.read()
        fn = f"micropp-restart-{RANK}-{RESTART_ID}.bin"
        p2, keep2 = params()
        b = ref.micro_rank(p2, keep2, rank=RANK, write_log=0)
        b.read_restart(RESTART_ID)
        for g, e in enumerate(strains(STEPS - 1)):
            b.set_strain(g, e)
        b.homogenize()
        cont = record(b)
        b.close()
        shutil.copy(fn, os.path.join(OUT, fn))
        raw = open(fn, "rb").read()
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp)
    doc = {"size": SIZE, "micro_type": ob.MIC_SPHERE, "geo": list(GEO), "materials": MATS, "coupling": COUPLING,
           "mpi_rank": RANK, "restart_id": RESTART_ID, "restart_after_step": STEPS - 2,
           "strains": [strains(t) for t in range(STEPS)], "steps": steps, "log": log,
           "restart_file": fn, "restart_bytes": len(raw), "continuation": cont}
    with open(os.path.join(OUT, "restart_log_ref.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {fn} ({len(raw)} bytes) and restart_log_ref.json; log:\n{log}")


if __name__ == "__main__":
    main()
