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
        another (the reference ELL values; the SpMV computes the stencil
        columns instead of loading them, with bit-identical results).
roofline: the dominant kernel of the mixed path (the FP32 multi-RHS SpMV),
  timed with HIP events on its own stream; algorithmic bytes per node =
  252 x 4 (coefficients incl. the 9 row pads) + 2 x 3K x 4 (p in, Ap out).
Secondary lines (not value; --no-batch skips them):
  drop_in: the same workload through the micropp3_* interface (one mixed-
        precision micropp<3> instance with 6 Gauss points; its constructor
        and one homogenize per step).
  batched: the batched J2 (64 RVEs of 30^3, 2 cyclic steps) and damage (16
        RVEs of 20..50^3, 3 damaging ramp steps) workloads, mixed precision
        against the FP64 batch, and the adjoint workload (4 RVEs of 40^3).
cpu_baseline: the oracle (C restatement of the reference; OpenMP over the
  Gauss points as the reference's homogenize) on this host: one sigma solve
  per thread at once, and one solve on one thread; an RVE solve counts as 2
  linear solves, as in the workload (12 solves = 6 RVE solves).
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
BYTES_PER_ROW_SPMV64 = 81 * 8 + 16  # values + p in + Ap out (columns are computed)
BYTES_PER_NODE_SPMV32 = 252 * 4 + 2 * 3 * K * 4


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


def cpu_threads():
    """(OpenMP threads the oracle runs with, logical CPUs this process may use, os.cpu_count())."""
    from oracle import bindings as ob

    if not os.path.exists(ob.ORACLE_SO):
        ob.ensure_built(ref=False)
    o = ob.Oracle()
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return o.lib.orc_nthreads(), usable, os.cpu_count()


def _cpu_solves(geo, strains):
    """One sigma solve per strain (an FE_ONE_WAY Gauss point each, no ctan) on the oracle, OpenMP over the
    Gauss points as the reference's homogenize (src/homogenize.cpp:72): (seconds, sigma)."""
    from oracle import bindings as ob

    o = ob.Oracle()
    mats = [ob.material(0, 1e6, 0.3), ob.material(0, 1e7, 0.25), ob.material(0, 1e7, 0.25)]
    ngp = len(strains)
    p, keep = ob.make_params(SIZE, ob.MIC3D_RANDOM_SPHERES, geo, mats, ngp=ngp, coupling=[ob.FE_ONE_WAY] * ngp,
                             lin_stress=0, calc_ctan_lin=0, cg_rel_tol=CG_TOL, cg_max_its=CG_MAX)
    m = o.micro(p, keep)
    for g in range(ngp):
        m.set_strain(g, strains[g])
    t0 = time.perf_counter()
    m.homogenize()
    dt = time.perf_counter() - t0
    sig = np.array([m.get_stress(g) for g in range(ngp)])
    m.close()
    return dt, sig


def cpu_baseline(geo, strains):
    """The oracle on this host, in the unit of `value`: one step of the workload is 12 linear solves (6 ctan
    unit-strain solves shared by the 6 strains + 6 sigma solves) for 6 RVE solves, so an RVE solve counts as 2
    linear solves of the 50^3 RVE to CG 1e-8.

    all cores: one solve per OpenMP thread at once (as many Gauss points as threads, the reference's
    parallel loop), RVE solves/s = solves / seconds / 2;
    one core: one solve on one thread, RVE solves/s = 1 / seconds / 2.
    Also returns the workload's sigma from the first 6 solves (for the error line)."""
    threads, usable, ncpu = cpu_threads()
    rng = np.random.default_rng(SEED + 1)
    batch = np.vstack([strains, rng.uniform(-5e-3, 5e-3, (max(0, threads - len(strains)), 6))])[:max(threads, 1)]
    dt_all, sig_all = _cpu_solves(geo, batch)
    dt_one, _ = _cpu_solves(geo, strains[:1])
    return {"value": len(batch) / dt_all / 2, "unit": "RVE solves/s", "cores": threads, "kind": "port",
            "sample": f"{len(batch)} sigma solves at 50^3 (FP64 Jacobi-PCG to 1e-8, one per OpenMP thread) on the "
                      f"oracle, {dt_all:.1f} s; an RVE solve counted as 2 linear solves, as in the GPU workload "
                      f"(12 solves = 6 RVE solves)",
            "value_1core": 1.0 / dt_one / 2, "unit_1core": "RVE solves/s (same unit as value)",
            "sample_1core": f"one sigma solve at 50^3 on one thread, {dt_one:.1f} s",
            "omp_threads": threads, "usable_cpus": usable, "os_cpu_count": ncpu}, sig_all[:len(strains)]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def drop_in(mp, geo, strains, steps):
    """The headline workload through the drop-in interface (micropp3_new_ext / micropp3_set_strain /
    micropp3_homogenize / micropp3_get_stress / micropp3_get_ctan): one micropp<3> instance with mixed
    precision and 6 FE_ONE_WAY Gauss points sharing the 50^3 RVE.  The constructor sets up the RVE and the
    multi-RHS solver and defers the six ctan unit-strain solves (calc_ctan_lin_fe_models); the first homogenize
    solves them in one 12-field pass with the 6 Gauss points' Newton solves (sigma by calc_ave_stress).  One
    step = a fresh instance + one homogenize: 6 RVE solves (sigma + ctan), the same 12 linear solves as a step
    of `value`.  Returns (dict, sigma, ctan of the last step)."""
    mats = [mp.material(0, 1e6, 0.3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    kw = dict(lin_stress=0, cg_rel_tol=CG_TOL, cg_max_its=CG_MAX, mixed_precision=1)
    ngp = len(strains)
    t_new, t_hom = [], []
    sig = ctan = None
    for s in range(steps + 1):  # the first is untimed
        mp.load().mppd_synchronize()
        t0 = time.perf_counter()
        m = mp.Micropp3(SIZE, mp.MIC3D_RANDOM_SPHERES, geo, mats, ngp=ngp, coupling=[mp.FE_ONE_WAY] * ngp, **kw)
        t1 = time.perf_counter()
        for g in range(ngp):
            m.set_strain(g, strains[g])
        m.homogenize()
        sig = np.array([m.get_stress(g) for g in range(ngp)])
        ctan = m.get_ctan(0)
        t2 = time.perf_counter()
        conv = all(m.has_converged(g) for g in range(ngp))
        m.close()
        if not conv:
            raise RuntimeError("drop-in route: a Gauss point did not converge")
        if s:
            t_new.append(t1 - t0)
            t_hom.append(t2 - t1)
        print(f"drop-in step {s} done", file=sys.stderr, flush=True)
    ctor, hom = float(np.mean(t_new)), float(np.mean(t_hom))
    return {"value": ngp / (ctor + hom), "unit": "RVE solves/s",
            "path": "micropp3_new_ext (mixed_precision) + micropp3_homogenize, 6 FE_ONE_WAY GPs of one 50^3 RVE",
            "steps": steps, "constructor_ms": 1e3 * ctor, "homogenize_ms": 1e3 * hom,
            "homogenize_linear_solves_per_s": 2 * ngp / hom}, sig, ctan


def batched_lines():
    """Secondary lines (not `value`): the batched nonlinear workloads at their real RVE sizes with reduced RVE
    and step counts, mixed precision against the FP64 batch, and the adjoint workload."""
    out = {}
    out["batch_j2"] = batch_j2(64, 2, mixed=True)
    out["batch_damage"] = batch_damage(16, 3, mixed=True)
    out["adjoint"] = adjoint_workload(16, 0)
    return out


# ---------------------------------------------------------------------------
# secondary workloads (not the headline): batched nonlinear RVEs and adjoint derivatives
# ---------------------------------------------------------------------------
def _batch_timed(mp, descs, ramps, params, label):
    """Batched homogenize + update_vars over the load steps: (seconds, info, stats, tuning, sigma, mixed stats)."""
    b = mp.Batch(descs, **params)
    mp.load().mppd_synchronize()
    t0 = time.perf_counter()
    for t in range(len(ramps[0])):
        b.set_strains(np.array([r[t] for r in ramps]))
        b.homogenize()
        b.update_vars()
        print(f"{label} batched step {t} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    prof = b.spmv_profile(bool(params.get("mixed_precision")))
    out = (dt, b.info(), b.stats(), b.tuning(), b.stress(), b.mixed_stats() if params.get("mixed_precision") else None,
           prof)
    del b
    return out


def _batch_run(mp, descs, ramps, params, label):
    """Batched homogenize over the load steps, then the same RVEs one micropp<3> instance after another (a
    subset, for the per-RVE rate); returns the measurement dict.  With mixed_precision, the FP64 batch of the
    same RVEs is timed instead of the one-by-one runs, and sigma is compared with it."""
    nrve, steps = len(descs), len(ramps[0])
    dt_b, info, stats, (_, spmv_slots, spmv_xcd), sig_b, mstats, (spmv_ms, spmv_bytes) = _batch_timed(
        mp, descs, ramps, params, label)
    common = {"rves": nrve, "load_steps": steps, "batched_s": dt_b,
              "nonlinear_rves": int(info[:, 0].sum()), "converged_rves": int(info[:, 2].sum()),
              "graph_chunk": int(stats[1]), "spmv_slots_per_round": spmv_slots, "spmv_blocks_per_xcd": spmv_xcd,
              "cg_iterations_queued": int(stats[0]), "host_syncs_last_step": int(stats[4]),
              "batched_newton_iterations_last_step": int(stats[5]),
              "spmv": {"precision": "fp32" if mstats is not None else "fp64", "ms": spmv_ms,
                       "bytes_per_launch": spmv_bytes, "achieved_GBs": spmv_bytes / spmv_ms / 1e6,
                       "frac_of_hbm_peak": spmv_bytes / spmv_ms / 1e6 / HBM_PEAK_GBS}}
    if mstats is not None:
        dt64, info64, stats64, _, sig64, _, (ms64, bytes64) = _batch_timed(
            mp, descs, ramps, dict(params, mixed_precision=0), label + " fp64")
        err = np.abs(sig_b - sig64).max(axis=1) / np.abs(sig64).max(axis=1)
        both = (info[:, 2] != 0) & (info64[:, 2] != 0)  # the last step's Newton converged in both
        mism = int(np.any(info[:, [0, 2]] != info64[:, [0, 2]], axis=1).sum())
        return dict(common, value=nrve * steps / dt_b, unit="RVE solves/s (sigma+ctan, FE_FULL)",
                    precision="FP32 inner CG, FP64 defect correction", fp64_batched_rve_solves_per_s=nrve * steps / dt64,
                    speedup_vs_fp64_batch=dt64 / dt_b, sigma_max_rel_err_vs_fp64=float(err.max()),
                    sigma_max_rel_err_vs_fp64_converged=float(err[both].max()) if both.any() else None,
                    converged_rves_both=int(both.sum()),
                    same_flags_as_fp64=mism == 0, rves_with_other_flags_than_fp64=mism,
                    fp64_cg_iterations_queued=int(stats64[0]), fp64_host_syncs_last_step=int(stats64[4]),
                    fp64_spmv={"ms": ms64, "bytes_per_launch": bytes64, "achieved_GBs": bytes64 / ms64 / 1e6,
                               "frac_of_hbm_peak": bytes64 / ms64 / 1e6 / HBM_PEAK_GBS},
                    mixed_outer_steps=mstats[0], mixed_fp64_fallbacks=mstats[1], mixed_short_solves=mstats[2])
    nseq = min(nrve, 4)
    kw = {k: v for k, v in params.items() if k != "graph_chunk"}
    t0 = time.perf_counter()
    sig_s = []
    for i in range(nseq):
        d = descs[i]
        m = mp.Micropp3(d["size"], d["micro_type"], d["geo"], d["materials"], ngp=1, coupling=[d["coupling"]], **kw)
        for t in range(steps):
            m.set_strain(0, ramps[i][t])
            m.homogenize()
            m.update_vars()
        sig_s.append(m.get_stress(0))
        m.close()
        print(f"{label} single RVE {i} done", file=sys.stderr, flush=True)
    dt_s = (time.perf_counter() - t0) / nseq
    return dict(common, value=nrve * steps / dt_b, unit="RVE solves/s (sigma+ctan, FE_FULL)",
                one_by_one_rve_solves_per_s=steps / dt_s, speedup_vs_one_by_one=(nrve * steps / dt_b) / (steps / dt_s),
                batched_equals_one_by_one=bool(np.array_equal(np.array(sig_s), sig_b[:nseq])))


J2_PARAMS = dict(lin_stress=0, nr_max_its=4, nr_rel_tol=1e-3, cg_rel_tol=1e-8, cg_max_its=2000)
DAMAGE_PARAMS = dict(lin_stress=0, nr_max_its=10, nr_rel_tol=1e-6, cg_rel_tol=1e-8, cg_max_its=2000)
# The reference damage law (src/material.cpp:206-269) leaves the linear range when r = sqrt(sigma_lin : eps)
# passes Xt / sqrt(E), and its hardening law holds q = 0 (D = 1) below r0 = Ey / sqrt(E) with Ey = 1e5 fixed
# in the law.  Xt = 1e5 = Ey puts the threshold at r0 (the reference's own benchmark-damage choice).  With
# E = 1e6, r0 = 100 and r ~ 1.2e3 |eps|: a ramp to |eps| = 1e-2 stays linear; one to 0.15 damages after its
# fourth step.
DAMAGE_XT = 1e5
DAMAGE_PEAK = 0.15
DAMAGE_RAMP_PEAK = 0.12  # the measured ramps: |eps| 0.04, 0.08, 0.12 in 3 steps (damage from the second)


def j2_fibre_workload(mp, nrve, steps, seed=3003):
    """RVEs of 30^3 hexes, a z-fibre of radius U[0.1, 0.3], J2 matrix (E 1e6, nu 0.3, Sy 1e4, Ka 1e5), elastic
    fibre (E 1e7, nu 0.25), FE_FULL; cyclic uniaxial + shear strain of amplitude 2e-3 (20 steps a cycle).
    The workload names no Newton tolerance: the reference defaults (4 iterations, 1e-3), CG 1e-8.
    Returns (RVE descriptors, per-RVE strain lists, batch parameters)."""
    rng = np.random.default_rng(seed)
    mats = [mp.material(1, 1e6, 0.3, 1e5, 1e4), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    descs = [dict(size=31, micro_type=mp.MIC_CILI_FIB_Z, geo=(rng.uniform(0.1, 0.3), 0, 0, 0), materials=mats,
                  coupling=mp.FE_FULL) for _ in range(nrve)]
    ramps = []
    for _ in range(nrve):
        d = np.zeros(6)
        d[0], d[3] = 2e-3, rng.uniform(-2e-3, 2e-3)
        ramps.append([d * np.sin(2 * np.pi * (t + 1) / 20) / np.sin(2 * np.pi / 20) for t in range(steps)])
    return descs, ramps, dict(J2_PARAMS)


def damage_layer_workload(mp, nrve, steps, seed=2002, peak=DAMAGE_PEAK, ramp_steps=10):
    """RVEs of n in {20, 30, 40, 50} hexes (seeded mix), a y-layer of width U[0.1, 0.3], damage matrix (E 1e6,
    nu 0.3, Xt 1e5), elastic fibre (E 1e7, nu 0.25), FE_FULL; each RVE a seeded strain direction ramped
    linearly to |eps| = peak over ramp_steps steps (the first `steps` of them); Newton tolerance 1e-6, CG 1e-8.
    Returns (RVE descriptors, per-RVE strain lists, batch parameters)."""
    rng = np.random.default_rng(seed)
    mats = [mp.material(2, 1e6, 0.3, 0, 0, DAMAGE_XT), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    descs = [dict(size=int(rng.choice([20, 30, 40, 50])) + 1, micro_type=mp.MIC_LAYER_Y,
                  geo=(rng.uniform(0.1, 0.3), 0, 0, 0), materials=mats, coupling=mp.FE_FULL) for _ in range(nrve)]
    ramps = []
    for _ in range(nrve):
        d = rng.normal(size=6)
        d *= peak / np.linalg.norm(d)
        ramps.append([d * (t + 1) / ramp_steps for t in range(steps)])
    return descs, ramps, dict(DAMAGE_PARAMS)


def sphere_rves(nrve, seed=4004):
    """Geometry of the adjoint workload: 40^3 hexes, 1-10 seeded spheres with radii U[0.05, 0.2]
    (MIC3D_RANDOM_SPHERES geo_params), and two macro strains U[-5e-3, 5e-3] per RVE."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nrve):
        geo = (float(rng.integers(1, 1 << 30)), float(rng.integers(1, 11)), 0.05, 0.2)
        out.append((geo, rng.uniform(-5e-3, 5e-3, (2, 6))))
    return out


def batch_j2(nrve, steps, mixed=False):
    """The batched J2 workload (j2_fibre_workload): batched vs one-by-one, or mixed vs FP64 batch."""
    import micropp_amd as mp

    nrve, steps = nrve or 32, steps or 2
    descs, ramps, params = j2_fibre_workload(mp, nrve, steps)
    params["mixed_precision"] = int(mixed)
    out = {"metric": "batched J2 RVEs (cylindrical fibre), RVE solves/s", "workload":
           f"{nrve} RVEs of 30^3 hexes, z-fibre radius U[0.1,0.3], J2 matrix, {steps} cyclic load steps"}
    out.update(_batch_run(mp, descs, ramps, params, "batch_j2"))
    return out


def batch_damage(nrve, steps, mixed=False):
    """The batched damage workload (damage_layer_workload) on a ramp of `steps` steps to |eps| = 0.12, which
    damages the matrix from its second step: batched vs one-by-one, or mixed vs FP64 batch."""
    import micropp_amd as mp

    nrve, steps = nrve or 16, steps or 3
    descs, ramps, params = damage_layer_workload(mp, nrve, steps, peak=DAMAGE_RAMP_PEAK, ramp_steps=steps)
    params["mixed_precision"] = int(mixed)
    out = {"metric": "batched damage RVEs (fibre layer, uneven sizes), RVE solves/s", "workload":
           f"{nrve} RVEs of 20..50^3 hexes, y-layer width U[0.1,0.3], damage matrix, {steps} ramp steps "
           f"to |eps| = {DAMAGE_RAMP_PEAK}"}
    out.update(_batch_run(mp, descs, ramps, params, "batch_damage"))
    return out


def adjoint_workload(nrve, _steps, fd_rves=4):
    """Elastic RVEs of 40^3 hexes with 1-10 seeded spheres (radii U[0.05, 0.2]): d sigma / d eps and
    d sigma / d (E, nu) per phase from the adjoint (one multi-RHS solve per RVE: 6 adjoint + 2 forward fields),
    the whole batch at once (mppd_msolver_adjoint_many, 4 host threads, each solver on its own stream) and one RVE
    after another, vs the forward approach (6 perturbed solves for ctan + central differences for the 4 material
    parameters) on the first fd_rves RVEs."""
    import micropp_amd as mp

    nrve = nrve or 4
    spec = [(0, 1e6, 0.3), (0, 1e7, 0.25), (0, 1e7, 0.25)]
    rv = sphere_rves(nrve)
    rves = [mp.Rve(41, mp.MIC3D_RANDOM_SPHERES, geo, [mp.material(*s) for s in spec]) for geo, _ in rv]
    ms = [mp.MSolver(r, 8) for r in rves]
    strains = np.array([s for _, s in rv])
    mp.load().mppd_synchronize()
    t0 = time.perf_counter()
    ctan_b, sig_b, dmat_b, _, _, ok_b = mp.adjoint_many(ms, strains, workers=4)
    t_many = time.perf_counter() - t0
    print("adjoint batch done", file=sys.stderr, flush=True)
    t_adj = t_fwd = 0.0
    err_ctan = err_fd = 0.0
    same = ok_b
    for i, ((geo, st), m) in enumerate(zip(rv, ms)):
        t0 = time.perf_counter()
        ctan, sig, dmat, _, _, ok = m.adjoint(st)
        t_adj += time.perf_counter() - t0
        same = same and ok and np.array_equal(ctan, ctan_b[i]) and np.array_equal(dmat, dmat_b[i])
        if i >= fd_rves:
            continue
        # forward: ctan from 6 unit-strain solves + sigma, then central differences (step 1e-6 relative) for
        # (E, nu) of phases 0 and 1, each a fresh RVE and solve
        t0 = time.perf_counter()
        sg, ctan_f, _, _, _ = mp.homogenize_mixed(m, st)
        fd = np.zeros((2, 6, 2, 2))
        for p in range(2):
            for w in (1, 2):
                h = spec[p][w] * 1e-6
                res = []
                for sgn in (1, -1):
                    sp = [list(s) for s in spec]
                    sp[p][w] += sgn * h
                    rr = mp.Rve(41, mp.MIC3D_RANDOM_SPHERES, geo, [mp.material(*s) for s in sp])
                    res.append(mp.MSolver(rr, 8).solve(st, rel_tol=1e-12, max_outer=30)[0])
                fd[:, :, p, w - 1] = (res[0] - res[1]) / (2 * h)
        t_fwd += time.perf_counter() - t0
        err_ctan = max(err_ctan, rel(ctan, ctan_f))
        err_fd = max(err_fd, rel(dmat[:, :, :2, :], fd))
        print(f"adjoint RVE {i} done", file=sys.stderr, flush=True)
    nfd = min(fd_rves, nrve)
    return {"metric": "adjoint derivatives (40^3 elastic RVEs), RVEs/s", "value": nrve / t_many,
            "unit": "RVEs/s (ctan + d sigma/d(E,nu) for 2 strains)", "rves": nrve,
            "path": "mppd_msolver_adjoint_many, 4 host threads, one stream per solver",
            "batch_ms_per_rve": 1e3 * t_many / nrve, "one_by_one_ms_per_rve": 1e3 * t_adj / nrve,
            "batch_vs_one_by_one": t_adj / t_many, "batch_equals_one_by_one": bool(same),
            "forward_fd_ms_per_rve": 1e3 * t_fwd / nfd, "speedup_vs_forward": (t_fwd / nfd) / (t_many / nrve),
            "forward_compared_rves": nfd, "ctan_rel_err_vs_forward": err_ctan, "dmat_rel_err_vs_central_fd": err_fd}


def newton_mixed_workload(nrve, steps):
    """The 50^3 random-sphere RVE with a J2 matrix (E=1e6, nu=0.3, Sy=1e4, Ka=1e5) and elastic spheres, one
    micropp<3> Gauss point (FE_ONE_WAY, sigma from calc_ave_stress, CG 1e-8): load steps of a yielding strain
    ramp, FP64 Newton (the reference algorithm) vs Newton with mixed-precision linear solves."""
    import micropp_amd as mp

    steps = steps or 2
    geo, strains = workload()
    mats = [mp.material(1, 1e6, 0.3, 1e5, 1e4), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    kw = dict(lin_stress=0, calc_ctan_lin=0, nr_max_its=4, nr_rel_tol=1e-6, cg_rel_tol=1e-8, cg_max_its=4000)
    d = strains[0] / np.abs(strains[0]).max() * 2e-2  # peak component 2e-2: the matrix yields
    res = {}
    for mixed in (0, 1):
        m = mp.Micropp3(51, mp.MIC3D_RANDOM_SPHERES, geo, mats, ngp=1, coupling=[mp.FE_ONE_WAY],
                        mixed_precision=mixed, **kw)
        sig, cost, t = [], 0, 0.0
        for k in range(steps):
            m.set_strain(0, d * (k + 1) / steps)
            mp.load().mppd_synchronize()
            t0 = time.perf_counter()
            m.homogenize()
            t += time.perf_counter() - t0
            sig.append(m.get_stress(0))
            cost += m.get_cost(0)
            m.update_vars()
            print(f"newton_mixed: mixed={mixed} step {k} done", file=sys.stderr, flush=True)
        res[mixed] = (np.array(sig), cost, t, m.is_non_linear(0))
    return {"metric": "J2 Newton solves/s at 50^3 (FP64 vs mixed-precision linear solves)",
            "value": steps / res[1][2], "unit": "Newton solves/s (one Gauss point, mixed precision)",
            "workload": f"50^3 random-sphere RVE, J2 matrix, {steps} load steps to peak strain 2e-2",
            "fp64_solves_per_s": steps / res[0][2], "speedup_vs_fp64": res[0][2] / res[1][2],
            "iterations_fp64": res[0][1], "iterations_mixed": res[1][1], "nonlinear": int(res[1][3]),
            "sigma_rel_err_vs_fp64": rel(res[1][0], res[0][0])}


def gp_concurrent_workload(nrve, steps):
    """One micropp<3> instance with mixed-precision Newton: nrve FE_ONE_WAY Gauss points of a 30^3 z-fibre RVE
    with a J2 matrix (E=1e6, nu=0.3, Sy=1e4, Ka=1e5), each with its own seeded strain ramp (peak components
    U[-3e-3, 3e-3]).  With the batch off (MPPD_GP_BATCH=0): MPPD_GP_WORKERS=1 (one solver, GPs one after
    another) and =4 (four solvers on their own streams, GPs concurrently), whose results must be equal bit for
    bit; then the default route, the GPs as one batch with mixed-precision solves."""
    import micropp_amd as mp

    nrve = nrve or 16
    steps = steps or 2
    mats = [mp.material(1, 1e6, 0.3, 1e5, 1e4), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    kw = dict(lin_stress=0, calc_ctan_lin=0, nr_max_its=4, nr_rel_tol=1e-6, cg_rel_tol=1e-8, cg_max_its=4000,
              mixed_precision=1)
    rng = np.random.default_rng(SEED)
    d = rng.uniform(-3e-3, 3e-3, (nrve, 6))
    res = {}
    for workers in (1, 4, "batch"):
        if workers == "batch":
            os.environ.pop("MPPD_GP_BATCH")
        else:
            os.environ["MPPD_GP_BATCH"] = "0"
            os.environ["MPPD_GP_WORKERS"] = str(workers)
        m = mp.Micropp3(31, mp.MIC_CILI_FIB_Z, (0.2, 0, 0, 0), mats, ngp=nrve, coupling=[mp.FE_ONE_WAY] * nrve,
                        **kw)
        sig, t = [], 0.0
        for k in range(steps):
            for g in range(nrve):
                m.set_strain(g, d[g] * (k + 1) / steps)
            mp.load().mppd_synchronize()
            t0 = time.perf_counter()
            m.homogenize()
            t += time.perf_counter() - t0
            sig.append([m.get_stress(g) for g in range(nrve)])
            m.update_vars()
            print(f"gp_concurrent: workers={workers} step {k} done", file=sys.stderr, flush=True)
        res[workers] = (np.array(sig), t, sum(m.is_non_linear(g) for g in range(nrve)))
    os.environ.pop("MPPD_GP_WORKERS")
    s4, sb = res[4][0], res["batch"][0]
    return {"metric": "Gauss-point solves/s, mixed-precision Newton at 30^3 (one vs four concurrent solvers)",
            "value": nrve * steps / res[4][1], "unit": "GP solves/s (4 solvers)",
            "workload": f"{nrve} GPs of a 30^3 z-fibre RVE, J2 matrix, {steps} load steps",
            "one_solver_gp_solves_per_s": nrve * steps / res[1][1], "speedup": res[1][1] / res[4][1],
            "nonlinear_gps": int(res[4][2]), "bit_equal": bool(np.array_equal(res[1][0], res[4][0])),
            "batched_gp_solves_per_s": nrve * steps / res["batch"][1],
            "batched_vs_4_solvers": res[4][1] / res["batch"][1],
            "batched_sigma_max_rel_err_vs_4_solvers": float(np.abs(sb - s4).max() / np.abs(s4).max())}


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
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the secondary lines (drop-in route, batched J2 / damage, adjoint)")
    ap.add_argument("--profile-iters", type=int, default=100)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of COPY_PROBE_BYTES (scales HBM counter readings)")
    ap.add_argument("--workload", default="mixed50",
                    choices=["mixed50", "batch_j2", "batch_damage", "adjoint", "newton_mixed", "gp_concurrent"],
                    help="mixed50 (default, the headline); the others measure the batched and adjoint paths")
    ap.add_argument("--rves", type=int, default=0, help="RVEs of a batch / adjoint workload (0: its default)")
    ap.add_argument("--load-steps", type=int, default=0, help="load steps of a batch workload (0: its default)")
    ap.add_argument("--mixed", action="store_true",
                    help="batch_j2 / batch_damage: mixed-precision linear solves (compared with the FP64 batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); RVEs are independent, so a batch "
              "would shard across GPUs by estimated cost, with only sigma and ctan gathered back", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.workload != "mixed50":
        fn = {"batch_j2": batch_j2, "batch_damage": batch_damage, "adjoint": adjoint_workload,
              "newton_mixed": newton_mixed_workload, "gp_concurrent": gp_concurrent_workload}[args.workload]
        extra = {"mixed": True} if args.mixed and args.workload in ("batch_j2", "batch_damage") else {}
        print(json.dumps(fn(args.rves, args.load_steps, **extra)), flush=True)
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
                     "kernel": f"k_m_spmv<{K}> (FP32 node-block SpMV, {K} RHS in 4-field groups, + p.Ap)",
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
                                           "traffic_source": p_src, "kernel": "k_cg_step1 (FP64 ELL SpMV, stencil columns computed, + p.Ap)",
                                           "bytes_per_launch": p_bytes, "avg_ms": p_spmv_ms,
                                           "cg_iteration_ms": p_iter_ms}}
        line["speedup_vs_plain_fp64"] = value / (N_STRAINS / dt_plain)
        line["mixed_vs_fp64_rel_err"] = {"sigma": rel(sig_m, sig_p), "ctan": rel(ctan_m, ctan_p)}

    if not args.no_batch:
        d, sig_d, ctan_d = drop_in(mp, geo, strains, 2)
        d["vs_value"] = d["value"] / value
        d["sigma_rel_err_vs_mixed"] = rel(sig_d, sig_m)
        d["ctan_rel_err_vs_mixed"] = rel(ctan_d, ctan_m)
        if not args.no_plain:
            d["sigma_rel_err_vs_fp64"] = rel(sig_d, sig_p)
            d["ctan_rel_err_vs_fp64"] = rel(ctan_d, ctan_p)
        line["drop_in"] = d
        line["batched"] = batched_lines()

    if not args.no_cpu:
        line["cpu_baseline"], csig = cpu_baseline(geo, strains)
        line["mixed_vs_cpu_rel_err"] = {"sigma": rel(sig_m, csig)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
