"""CPU: the benchmark's contract pieces that need no GPU.

- `--gpus N` with N > 1 exits with status 2 and says why, before any GPU work (one GPU is measured;
  RVEs would shard across GPUs by estimated cost).
- The headline workload is seeded: 8 random spheres with radii U[0.05, 0.15] (the geometry
  parameters of the random-sphere micro type) and 6 macro strains with components U[-5e-3, 5e-3].
- The HBM traffic figures that bench.py reports are read from the committed rocprofv3 summaries,
  and the algorithmic bytes per unit match what those summaries state.
"""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and not p.stdout


def test_headline_workload_is_seeded():
    geo, strains = bench.workload()
    geo2, strains2 = bench.workload()
    assert geo == geo2 and np.array_equal(strains, strains2)
    assert geo[1:] == (8.0, 0.05, 0.15)
    assert strains.shape == (6, 6) and np.abs(strains).max() <= 5e-3
    assert not np.array_equal(strains, bench.workload(seed=bench.SEED + 1)[1])


def test_traffic_summaries_parse():
    t32, src32 = bench.pmc_traffic("k_m_spmv12_50cube_traffic.md")
    t64, src64 = bench.pmc_traffic("k_cg_step1_50cube_traffic.md")
    assert src32 and src64
    assert 1.0 < t32 / (bench.BYTES_PER_NODE_SPMV32 * 51 ** 3) < 1.2  # measured over algorithmic bytes
    assert 1.0 < t64 / (bench.BYTES_PER_ROW_SPMV64 * 3 * 51 ** 3) < 1.2
    assert bench.pmc_traffic("no_such_summary.md") == (None, None)


def test_algorithmic_bytes_match_the_summaries():
    assert bench.BYTES_PER_NODE_SPMV32 == 252 * 4 + 2 * 3 * 12 * 4 == 1296
    assert bench.BYTES_PER_ROW_SPMV64 == 81 * 8 + 16 == 664
    with open(os.path.join(ROOT, "profiles", "k_m_spmv12_50cube_kernel_trace.md")) as f:
        assert "1296 B/node" in f.read()
    with open(os.path.join(ROOT, "profiles", "k_cg_step1_50cube_kernel_trace.md")) as f:
        assert "664 B/row" in f.read()
