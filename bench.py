#!/usr/bin/env python3
"""RVE homogenization throughput on one MI355X: RVE solves/s (sigma + ctan).

Workload (the 50^3 mixed-vs-FP64 case): one RVE of 50x50x50 hexes (51^3
nodes, 397 953 DOF) with 8 seeded random spheres (centres U[0,1)^3, radii
U[0.05, 0.15]; matrix E=1e6 nu=0.3, spheres E=1e7 nu=0.25, linear elastic)
and 6 seeded macro strains with components U[-5e-3, 5e-3].  One step does
what micropp<3> does for 6 FE Gauss points sharing the RVE: ctan from the six
1e-8 unit-strain solves, then sigma of each strain from its own solve and the
volume average, all to the reference CG's stopping metric ||D^-1 r|| <= 1e-8
of its initial value.  A step yields 6 RVE solves (sigma + ctan).

Paths
  mixed (value): the 12 solves of a step run together -- FP32 multi-RHS
        Jacobi-PCG inner solves, FP64 residual / update / averages.
  plain: the reference algorithm on the GPU, one FP64 Newton/CG solve after
        another (ELL SpMV with explicit columns).
roofline: the dominant kernel of the mixed path (the FP32 multi-RHS SpMV),
  timed with HIP events on its own stream; algorithmic bytes per node =
  244 x 4 (coefficients incl. the pad slot) + 2 x 3K x 4 (p in, Ap out).
cpu_baseline: the oracle (C restatement of the reference; OpenMP over the
  solves as the reference's homogenize) on this host, one step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SIZE = 51
SEED = 2024
N_STRAINS = 6
K = 12
CG_TOL, CG_MAX = 1e-8, 2000
D_EPS = 1e-8
HBM_PEAK_GBS = 8000.0
BYTES_PER_ROW_SPMV64 = 81 * 12 + 16
BYTES_PER_NODE_SPMV32 = 244 * 4 + 2 * 3 * K * 4


def workload(seed=SEED):
    rng = np.random.default_rng(seed)
    strains = rng.uniform(-5e-3, 5e-3, (N_STRAINS, 6))
    geo = (float(seed), 8.0, 0.05, 0.15)
    return geo, strains


def step_plain(mp, rve, solver, u, strains):
    """ctan (6 unit solves from u = 0) + sigma of every strain, one FP64 solve at a time."""
    ctan = np.zeros((6, 6))
    kw = dict(cg_rel_tol=CG_TOL, cg_max_its=CG_MAX)
    cg_its = 0
    for i in range(6):
        mp.load().mppd_memset(u.ptr, 0, u.nbytes)
        e = np.zeros(6)
        e[i] = D_EPS
        cg_its += solver.newton(e, u, None, **kw)[1]
        ctan[:, i] = solver.ave_stress(u) / D_EPS
    sig = np.zeros((len(strains), 6))
    for k, eps in enumerate(strains):
        mp.load().mppd_memset(u.ptr, 0, u.nbytes)
        cg_its += solver.newton(eps, u, None, **kw)[1]
        sig[k] = solver.ave_stress(u)
    return sig, ctan, cg_its


def step_mixed(mp, msolver, strains):
    sig, ctan, stats, resid, ok = mp.homogenize_mixed(msolver, strains, D_EPS, rel_tol=CG_TOL,
                                                      max_inner_its=CG_MAX)
    if not ok:
        raise RuntimeError(f"mixed solve did not converge (residual ratios {resid})")
    return sig, ctan, stats


def timed(fn, steps, warmup, mp, label):
    for w in range(warmup):
        fn()
        print(f"{label} warmup {w} done", file=sys.stderr, flush=True)
    mp.load().mppd_synchronize()
    t0 = time.perf_counter()
    out = None
    for s in range(steps):
        out = fn()
        print(f"{label} step {s} done", file=sys.stderr, flush=True)
    mp.load().mppd_synchronize()
    return (time.perf_counter() - t0) / steps, out


def cpu_baseline(geo, strains):
    from oracle import bindings as ob

    if not os.path.exists(ob.ORACLE_SO):
        ob.ensure_built(ref=False)
    o = ob.Oracle()
    mats = [ob.material(0, 1e6, 0.3), ob.material(0, 1e7, 0.25), ob.material(0, 1e7, 0.25)]
    ngp = len(strains)
    t0 = time.perf_counter()
    p, keep = ob.make_params(SIZE, ob.MIC3D_RANDOM_SPHERES, geo, mats, ngp=ngp, coupling=[ob.FE_FULL] * ngp,
                             lin_stress=0, cg_rel_tol=CG_TOL, cg_max_its=CG_MAX)
    m = o.micro(p, keep)  # constructor: the six ctan solves (OpenMP over them)
    for g in range(ngp):
        m.set_strain(g, strains[g])
    m.homogenize()  # the six sigma solves (OpenMP over the GPs)
    dt = time.perf_counter() - t0
    sig = np.array([m.get_stress(g) for g in range(ngp)])
    threads = min(o.lib.orc_nthreads(), ngp)
    return ngp / dt, threads, sig, m.ctan_lin(), dt


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def pmc_traffic(summary):
    """HBM bytes per launch measured by the FETCH_SIZE / WRITE_SIZE runs whose summary (commands, copy-probe
    scaling) is profiles/<summary>; None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("| total traffic"):
                mb = float(line.split("**")[1].split()[0])
                return mb * 1e6, "profiles/" + summary
    return None, None


COPY_PROBE_BYTES = 1 << 30


def copy_probe():
    """Three copies of a 1 GiB buffer: each reads and writes COPY_PROBE_BYTES, the yardstick for the
    FETCH_SIZE / WRITE_SIZE counters of the solver kernels."""
    import torch

    a = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--profile-iters", type=int, default=100)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of COPY_PROBE_BYTES (scales HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); RVEs are independent, so a "
              "batch would shard across GPUs with no exchange (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import micropp_amd as mp

    if not mp.gpu_available():
        raise SystemExit("bench.py: no HIP device")
    geo, strains = workload()
    mats = [mp.material(0, 1e6, 0.3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    rve = mp.Rve(SIZE, mp.MIC3D_RANDOM_SPHERES, geo, mats)

    # ---- headline: mixed-precision multi-RHS path ----
    ms = mp.MSolver(rve, K)
    dt_mixed, (sig_m, ctan_m, stats) = timed(lambda: step_mixed(mp, ms, strains), args.steps, args.warmup, mp,
                                             "mixed")
    value = N_STRAINS / dt_mixed
    spmv_ms, iter_ms = ms.profile(args.profile_iters)
    nodes = SIZE ** 3
    spmv_bytes = BYTES_PER_NODE_SPMV32 * nodes
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    traffic, traffic_src = pmc_traffic(f"k_m_spmv{K}_50cube_traffic.md")

    line = {
        "metric": "RVE solves/s (sigma+ctan) at 50^3, 1 GPU",
        "value": value,
        "unit": "RVE solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner CG / f64 residual, update and averages",
        "data": "synthetic (seeded random-sphere RVE and macro strains)",
        "config": {"workload": "50^3 hex RVE, 8 random spheres, 6 macro strains; sigma + ctan per strain "
                               "(12 solves per step)",
                   "path": "mixed precision, 12 right-hand sides per pass", "rows": 3 * nodes,
                   "rve_solves_per_step": N_STRAINS, "rel_tol": CG_TOL,
                   "outer_steps": stats[0], "inner_iterations": stats[1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": f"k_m_spmv<{K}> (FP32 node-block SpMV, {K} RHS, + p.Ap)",
                     "bytes_per_launch": spmv_bytes, "avg_ms": spmv_ms, "inner_iteration_ms": iter_ms},
    }

    # ---- plain FP64 path (the reference algorithm, one solve at a time) ----
    if not args.no_plain:
        solver = mp.Solver(rve)
        u = mp.DeviceArray(rve.nrow)
        psteps = max(1, min(args.steps, 3))
        dt_plain, (sig_p, ctan_p, cg_its) = timed(lambda: step_plain(mp, rve, solver, u, strains), psteps, 1, mp,
                                                  "plain")
        b = mp.DeviceArray(rve.nrow)
        solver.set_displ_bc(strains[0], u)
        solver.assembly_mat(u)
        solver.assembly_rhs(u, None, b)
        p_spmv_ms, p_iter_ms = mp.cg_profile(solver, b, args.profile_iters)
        p_bytes = BYTES_PER_ROW_SPMV64 * 3 * nodes
        p_ach = p_bytes / (p_spmv_ms * 1e-3) / 1e9
        p_traffic, p_src = pmc_traffic("k_cg_step1_50cube_traffic.md")
        line["plain_fp64"] = {"value": N_STRAINS / dt_plain, "ms_per_step": 1e3 * dt_plain,
                              "cg_iterations_per_step": cg_its,
                              "roofline": {"bound": "hbm", "achieved": p_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": p_ach / HBM_PEAK_GBS, "traffic": p_traffic,
                                           "traffic_source": p_src, "kernel": "k_cg_step1 (FP64 ELL SpMV + p.Ap)",
                                           "bytes_per_launch": p_bytes, "avg_ms": p_spmv_ms,
                                           "cg_iteration_ms": p_iter_ms}}
        line["speedup_vs_plain_fp64"] = value / (N_STRAINS / dt_plain)
        line["mixed_vs_fp64_rel_err"] = {"sigma": rel(sig_m, sig_p), "ctan": rel(ctan_m, ctan_p)}

    if not args.no_cpu:
        cv, threads, csig, cctan, cdt = cpu_baseline(geo, strains)
        line["cpu_baseline"] = {"value": cv, "unit": "RVE solves/s", "cores": threads, "kind": "port",
                                "sample": f"one step (6 ctan + 6 sigma solves at 50^3, FP64, CG 1e-8) on the "
                                          f"oracle, {cdt:.1f} s"}
        line["mixed_vs_cpu_rel_err"] = {"sigma": rel(sig_m, csig), "ctan": rel(ctan_m, cctan)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
