"""GPU: the Newton loop of several elastic Gauss points in one multi-RHS pass
(mppd_msolver_newton), and micropp<3> with mixed precision on an all-elastic
RVE, whose constructor ctan and FE Gauss points run through it.

Against the FP64 path (mppd_newton, itself bit-exact with the oracle's
assembly and within 1e-6 of it), per Gauss point:
- the same Newton iteration count and convergence flag (the reference's
  exit quirk included: at nr_max_its the Newton stops unconverged);
- sigma and the displacement within 1e-8 relative at CG 1e-10 (the FP32
  inner solves are corrected to the FP64 residual metric);
- a warm start from the previous step's displacement, as homogenize gives
  it (u_n), reproduced.
micropp<3>: sigma, ctan and u_k within 1e-6 of the FP64 instance and of the
oracle, the same flags, over two load steps, with 18 FE GPs (passes of 16
and 2) beside FE_LINEAR and MIX_RULE_CHAMIS GPs.
"""
import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu

SPEC = [(0, 1e6, 0.3), (0, 1e7, 0.25), (0, 1e7, 0.25)]
GEO = (17.0, 6.0, 0.05, 0.2)


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def strains(n, seed, amp=5e-3):
    return np.random.default_rng(seed).uniform(-amp, amp, (n, 6))


def test_msolver_newton_matches_fp64_newton(mp):
    r = mp.Rve(13, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in SPEC])
    n = 6
    kw = dict(cg_rel_tol=1e-10, cg_max_its=4000)
    ms, sol = mp.MSolver(r, 8), mp.Solver(r)
    u_m = [mp.DeviceArray(r.nrow) for _ in range(n)]
    u_f = [mp.DeviceArray(r.nrow) for _ in range(n)]
    for step, e in enumerate((strains(n, 1), strains(n, 2))):
        # step 0 from zero; step 1 warm-started from step 0's displacements
        sig, out, stats = ms.newton(e, u_in=u_m if step else None, u_out=u_m, **kw)
        print(f"step {step}: out {out.tolist()}, stats {stats}", flush=True)
        for k in range(n):
            fo = sol.newton(e[k], u_f[k], **kw)
            sf = sol.ave_stress(u_f[k])
            es, eu = rel(sig[k], sf), rel(u_m[k].numpy(), u_f[k].numpy())
            print(f"  gp {k}: fp64 {fo}, sigma {es:.2e}, u {eu:.2e}", flush=True)
            assert (out[k][0], out[k][2]) == (fo[0], fo[2])
            assert es < 1e-8 and eu < 1e-8


def test_msolver_newton_exit_quirk_and_zero_strain(mp):
    """nr_max_its = 1: one Newton iteration, then the loop ends unconverged (src/solve.cpp:43-78), as the FP64
    path reports it; a zero strain from zero displacement converges at the first test with no iteration."""
    r = mp.Rve(9, mp.MIC_SPHERE, (0.3, 0, 0, 0), [mp.material(*s) for s in SPEC])
    ms, sol = mp.MSolver(r, 4), mp.Solver(r)
    e = np.vstack([strains(3, 5), np.zeros((1, 6))])
    sig, out, _ = ms.newton(e, nr_max_its=1, cg_rel_tol=1e-10, cg_max_its=4000)
    for k in range(4):
        u = mp.DeviceArray(r.nrow)
        fo = sol.newton(e[k], u, nr_max_its=1, cg_rel_tol=1e-10, cg_max_its=4000)
        print(f"gp {k}: multi {out[k].tolist()} fp64 {fo}", flush=True)
        assert (out[k][0], out[k][2]) == (fo[0], fo[2])
    assert out[:3, 2].tolist() == [0, 0, 0] and out[:3, 0].tolist() == [1, 1, 1]
    assert out[3].tolist() == [0, 0, 1] and not sig[3].any()


def make(mp, mixed, ngp, coupling, **kw):
    return mp.Micropp3(11, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in SPEC], ngp=ngp,
                       coupling=coupling, mixed_precision=mixed, **kw)


@pytest.mark.parametrize("lin_stress", [0, 1])
def test_micropp_elastic_mixed_runs_multi_rhs(mp, oracle, lin_stress):
    kw = dict(lin_stress=lin_stress, cg_rel_tol=1e-10, cg_max_its=4000)
    cp = [mp.FE_ONE_WAY] * 12 + [mp.FE_FULL] * 6 + [mp.FE_LINEAR, mp.MIX_RULE_CHAMIS]
    ngp = len(cp)
    a, b = make(mp, 1, ngp, cp, **kw), make(mp, 0, ngp, cp, **kw)
    p, k = ob.make_params(11, ob.MIC3D_RANDOM_SPHERES, GEO, [ob.material(*s) for s in SPEC], ngp=ngp, coupling=cp,
                          **kw)
    o = oracle.micro(p, k)
    print(f"ctan_lin: multi vs FP64 {rel(a.ctan_lin(), b.ctan_lin()):.2e}, vs oracle "
          f"{rel(a.ctan_lin(), o.ctan_lin()):.2e}", flush=True)
    assert rel(a.ctan_lin(), b.ctan_lin()) < 1e-6 and rel(a.ctan_lin(), o.ctan_lin()) < 1e-6
    for t in range(2):
        e = strains(ngp, 10 + t)
        for m in (a, b, o):
            for g in range(ngp):
                m.set_strain(g, e[g])
            m.homogenize()
        costs = []
        for g in range(ngp):
            sa, sb, so = a.get_stress(g), b.get_stress(g), o.get_stress(g)
            ca = a.get_ctan(g)
            es, eo = rel(sa, sb), rel(sa, so)
            print(f"step {t} gp {g} ({cp[g]}): sigma vs FP64 {es:.2e} vs oracle {eo:.2e}; cost {a.get_cost(g)}, "
                  f"converged {a.has_converged(g)}", flush=True)
            assert es < 1e-6 and eo < 1e-6
            assert rel(ca, b.get_ctan(g)) < 1e-6
            assert a.has_converged(g) == b.has_converged(g) == o.has_converged(g)
            assert a.is_non_linear(g) == 0
            if cp[g] in (mp.FE_ONE_WAY, mp.FE_FULL):
                assert rel(a.state(g, 1), b.state(g, 1)) < 1e-6  # u_k
                costs.append(a.get_cost(g))
        # the 18 FE GPs ran as passes of 16 and 2: the GPs of a pass share its inner iteration count
        assert len(set(costs[:16])) == 1 and len(set(costs[16:])) == 1, costs
        for m in (a, b, o):
            m.update_vars()


def test_micropp_elastic_mixed_subiterates_on_the_one_gp_path(mp):
    """nr_max_its = 1 leaves every Newton unconverged; with subiterations on, those GPs are solved again on the
    one-GP path, which subiterates: flags as the FP64 instance's, sigma within 1e-6."""
    kw = dict(lin_stress=0, nr_max_its=1, subiterations=1, nsubiterations=2, cg_rel_tol=1e-10, cg_max_its=4000)
    cp = [mp.FE_ONE_WAY] * 4
    a, b = make(mp, 1, 4, cp, **kw), make(mp, 0, 4, cp, **kw)
    e = strains(4, 21)
    for m in (a, b):
        for g in range(4):
            m.set_strain(g, e[g])
        m.homogenize()
    for g in range(4):
        print(f"gp {g}: multi ({a.has_converged(g)}, {a.has_subiterated(g)}) fp64 ({b.has_converged(g)}, "
              f"{b.has_subiterated(g)})", flush=True)
        assert (a.has_converged(g), a.has_subiterated(g)) == (b.has_converged(g), b.has_subiterated(g)) == (False, True)
        assert rel(a.get_stress(g), b.get_stress(g)) < 1e-6


@pytest.mark.parametrize("ngp", [6, 12])
def test_micropp_ctan_solves_ride_the_first_pass(mp, ngp):
    """The constructor's six ctan solves are deferred: the first homogenize solves them in its first pass,
    beside the GPs (6 GPs: one 12-field pass; 12 GPs: passes of 10 + 6 ctan fields and 2 GPs).  Against an
    instance that reads its ctan before homogenize (the ctan solved on its own): the same ctan, sigma and u_k bit
    for bit (each field's solve does not depend on the other fields of its pass), and within 1e-6 of the FP64
    instance."""
    kw = dict(lin_stress=0, cg_rel_tol=1e-10, cg_max_its=4000)
    cp = [mp.FE_ONE_WAY] * ngp
    a, b, c = make(mp, 1, ngp, cp, **kw), make(mp, 1, ngp, cp, **kw), make(mp, 0, ngp, cp, **kw)
    ctan_b = b.ctan_lin()  # solved on its own, now
    e = strains(ngp, 33)
    for m in (a, b, c):
        for g in range(ngp):
            m.set_strain(g, e[g])
        m.homogenize()
    assert np.array_equal(a.ctan_lin(), ctan_b)
    assert rel(a.ctan_lin(), c.ctan_lin()) < 1e-6
    for g in range(ngp):
        assert np.array_equal(a.get_stress(g), b.get_stress(g))
        assert np.array_equal(a.get_ctan(g), b.get_ctan(g))
        assert np.array_equal(a.state(g, 1), b.state(g, 1))
        assert rel(a.get_stress(g), c.get_stress(g)) < 1e-6
        assert a.has_converged(g) == c.has_converged(g)
