"""GPU: batches of RVEs with mixed-precision linear solves (mppd_batch_params::mixed_precision).

Every Newton linear solve of every RVE runs as FP64 defect correction around an
FP32 Jacobi-PCG on an FP32 copy of its matrix, to the reference CG's metric in
FP64 (the rules of mppd_newton_mixed).  The bar is the FP64 batch, itself equal
bit for bit to micropp<3> instances and within 1e-6 of the oracle:
- sigma within 1e-6 relative and ctan within 1e-4 (ctan is a forward difference
  over a 1e-8 strain step, so it multiplies solve-level differences by ~1e8 /
  |sigma|) for every RVE, elastic, J2 and damaged, whose FP64 Newton converged
  at this and every earlier step (a linear solve whose FP32 inner CG breaks
  down or stalls short of the tolerance within the outer-step cap is redone by
  the FP64 CG, so every solve meets the same tolerance); 1e-4 / 1e-2 where a
  Newton stopped at its cap unconverged, whose path may cross a singular
  tangent (tests/test_oracle.py::test_capped_damage_newton_crosses_a_singular_tangent);
- the same nonlinear (history) and convergence flags.
"""
import numpy as np
import pytest

from test_gpu_batch import KW, SPECS, mixed_batch, rel, run_batch
import test_gpu_gp_batch as gpb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


@pytest.mark.parametrize("its_with_A0", [0, 1])
def test_mixed_batch_matches_fp64_batch(mp, its_with_A0):
    descs = mixed_batch(mp)
    kw = dict(KW, use_A0=1, its_with_A0=its_with_A0) if its_with_A0 else dict(KW)
    steps = 2
    r64, s64 = run_batch(mp, descs, steps, **kw)
    r32, s32 = run_batch(mp, descs, steps, mixed_precision=1, **kw)
    print(f"CG iterations queued: fp64 {s64[0]:.0f}, mixed {s32[0]:.0f}", flush=True)
    conv_so_far = [True] * len(descs)
    for t in range(steps):
        sig64, ctan64, info64 = r64[t]
        sig32, ctan32, info32 = r32[t]
        for i, d in enumerate(descs):
            conv_so_far[i] = conv_so_far[i] and bool(info64[i][2])
            ts, tc = (1e-6, 1e-4) if conv_so_far[i] else (1e-4, 1e-2)
            es, ec = rel(sig32[i], sig64[i]), rel(ctan32[i], ctan64[i])
            print(f"step {t} rve {i} ({d['kind']}): info mixed {tuple(info32[i])} fp64 {tuple(info64[i])}; "
                  f"sigma {es:.2e} ctan {ec:.2e}", flush=True)
            assert es < ts and ec < tc
            assert info32[i][0] == info64[i][0] and info32[i][2] == info64[i][2]


def test_mixed_batch_spmv_shape_does_not_change_results(mp):
    """The inner FP32 SpMV's tuned shapes (slots in flight, block order) change only load scheduling: the
    mixed batch's results are the same bit for bit in the extreme shapes."""
    descs = mixed_batch(mp)[:4]
    ra, _ = run_batch(mp, descs, 1, mixed_precision=1, spmv_chunk=9, block_order=0, **KW)
    rb, _ = run_batch(mp, descs, 1, mixed_precision=1, spmv_chunk=81, block_order=1, **KW)
    for x, y in zip(ra[0], rb[0]):
        assert np.array_equal(x, y)


def test_mixed_batch_fp64_fallback_equals_fp64_batch(mp, monkeypatch):
    """The fallback path, forced from the host (MPPD_FORCE_FP64_FALLBACK=1: every FP32 inner solve counts as
    broken down): each linear solve is redone by the FP64 batch CG from du = 0, so sigma, ctan, cost and flags
    equal the FP64 batch's bit for bit."""
    descs = mixed_batch(mp)
    r64, _ = run_batch(mp, descs, 2, **KW)
    monkeypatch.setenv("MPPD_FORCE_FP64_FALLBACK", "1")
    rfb, _ = run_batch(mp, descs, 2, mixed_precision=1, **KW)
    for t in range(2):
        for x, y in zip(r64[t], rfb[t]):
            assert np.array_equal(x, y), f"step {t}"


def test_mixed_batch_stats(mp):
    """An elastic batch: every solve converges in FP32 steps (no FP64 fallback, no short solve)."""
    descs = [d for d in mixed_batch(mp) if d["kind"] == "elastic"]
    b = mp.Batch([dict(size=d["size"], micro_type=d["micro_type"], geo=d["geo"],
                       materials=[mp.material(*s) for s in SPECS[d["kind"]]], coupling=d["coupling"])
                  for d in descs], mixed_precision=1, **KW)
    b.set_strains(np.tile([1e-3, -2e-4, 3e-4, 5e-4, 0.0, -1e-4], (len(descs), 1)))
    b.homogenize()
    outer, fallbacks, short = b.mixed_stats()
    print(f"outer steps {outer}, fallbacks {fallbacks}, short {short}", flush=True)
    assert outer > 0 and fallbacks == 0 and short == 0


@pytest.mark.parametrize("slots,lin_stress", [(-1, 0), (2, 0), (2, 1)])
def test_micropp_mixed_instance_runs_as_mixed_batch(mp, monkeypatch, slots, lin_stress):
    """A micropp<3> instance with mixed_precision runs its FE GPs as a batch with mixed-precision solves (all
    5 FE GPs in one batch, or groups of 2 with an idle slot in the last); against the same instance on
    concurrent solvers (MPPD_GP_BATCH=0, the node-block FP32 CG of mppd_newton_mixed): sigma within 1e-6,
    ctan within 1e-4, the same nonlinear and convergence flags, over three J2 load steps with history
    allocation."""
    steps = 3
    bat, seq = (gpb.make(mp, lin_stress=lin_stress, mixed_precision=1, **gpb.KW) for _ in range(2))
    rp = gpb.ramps(bat.ngp, steps)
    nl = 0
    for t in range(steps):
        a = gpb.step(bat, rp, t, monkeypatch, slots)
        b = gpb.step(seq, rp, t, monkeypatch, 0)
        for g, (ra, rb) in enumerate(zip(a, b)):
            es, ec = rel(ra[0], rb[0]), rel(ra[1], rb[1])
            print(f"step {t} gp {g}: flags batch {ra[2]} solvers {rb[2]}; sigma {es:.2e} ctan {ec:.2e}", flush=True)
            assert es < 1e-6 and ec < 1e-4
            assert ra[2][0] == rb[2][0] and ra[2][2] == rb[2][2]
            nl += ra[2][0]
    assert nl > 0, "the case must allocate history"


def test_mixed_batch_coarse_level_against_jacobi(mp, monkeypatch):
    """The two-level preconditioner of the batch's inner FP32 CG (a coarse grid per RVE from its initial
    tangent; 4 elements per coarse element at 20^3, 2 at 12^3, 4 / 2 / 4 at 20 x 12 x 16) against Jacobi alone (MPPD_COARSE=0 at
    mppd_batch_create), both against the FP64 batch: J2 and damage RVEs of two sizes over two load steps.  The
    coarse level cuts the inner iterations, and sigma / ctan keep the mixed-precision bars."""
    descs = [dict(size=21, micro_type=mp.MIC_CILI_FIB_Z, geo=(0.25, 0, 0, 0), kind="j2", coupling=mp.FE_FULL),
             dict(size=13, micro_type=mp.MIC_LAYER_Y, geo=(0.3, 0, 0, 0), kind="damage", coupling=mp.FE_ONE_WAY),
             dict(size=21, micro_type=mp.MIC_SPHERE, geo=(0.3, 0, 0, 0), kind="plastic", coupling=mp.FE_FULL),
             dict(size=(21, 13, 17), micro_type=mp.MIC_SPHERE, geo=(0.25, 0, 0, 0), kind="plastic",
                  coupling=mp.FE_ONE_WAY)]
    kw = dict(KW, cg_rel_tol=1e-8)
    steps = 2
    r64, _ = run_batch(mp, descs, steps, **kw)
    r2l, s2l = run_batch(mp, descs, steps, mixed_precision=1, **kw)
    monkeypatch.setenv("MPPD_COARSE", "0")
    rj, sj = run_batch(mp, descs, steps, mixed_precision=1, **kw)
    monkeypatch.delenv("MPPD_COARSE")
    print(f"inner iterations queued: two-level {s2l[0]:.0f}, Jacobi {sj[0]:.0f}", flush=True)
    assert s2l[0] < 0.8 * sj[0]
    conv_so_far = [True] * len(descs)
    for t in range(steps):
        for i, d in enumerate(descs):
            conv_so_far[i] = conv_so_far[i] and bool(r64[t][2][i][2])
            ts, tc = (1e-6, 1e-4) if conv_so_far[i] else (1e-4, 1e-2)
            for lab, r in (("two-level", r2l), ("jacobi", rj)):
                es, ec = rel(r[t][0][i], r64[t][0][i]), rel(r[t][1][i], r64[t][1][i])
                print(f"step {t} rve {i} ({d['kind']}) {lab}: sigma {es:.2e} ctan {ec:.2e}", flush=True)
                assert es < ts and ec < tc
                assert r[t][2][i][0] == r64[t][2][i][0] and r[t][2][i][2] == r64[t][2][i][2]
