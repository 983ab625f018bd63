"""GPU: the mixed-precision multi-RHS path against the FP64 path and the oracle.

Bar: averaged stress and ctan within 1e-6 relative of the FP64 results (the
FP32 inner solves are refined by FP64 defect correction to the same 1e-8
preconditioned-residual tolerance the FP64 CG uses).
"""
import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu

GEO = (4242.0, 8.0, 0.05, 0.15)


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def elastic(mod):
    return [mod.material(0, 1e6, 0.3), mod.material(0, 1e7, 0.25), mod.material(0, 1e7, 0.25)]


@pytest.mark.parametrize("size", [11, 21])
def test_mixed_matches_fp64_and_oracle(mp, oracle, size):
    rng = np.random.default_rng(size)
    strains = rng.uniform(-5e-3, 5e-3, (6, 6))
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp))
    ms = mp.MSolver(r, 12)
    sig, ctan, stats, resid, ok = mp.homogenize_mixed(ms, strains)
    print(f"size {size}: outer {stats[0]} inner {stats[1]} max resid ratio {resid.max():.2e}", flush=True)
    assert ok and resid.max() < 1e-8
    # FP64 GPU path: micropp<3> with six GPs
    kw = dict(lin_stress=0, cg_rel_tol=1e-8, cg_max_its=3000)
    m = mp.Micropp3(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp), ngp=6, coupling=[mp.FE_FULL] * 6, **kw)
    for g in range(6):
        m.set_strain(g, strains[g])
    m.homogenize()
    sig64 = np.array([m.get_stress(g) for g in range(6)])
    print("sigma rel", rel(sig, sig64), "ctan rel", rel(ctan, m.ctan_lin()), flush=True)
    assert rel(sig, sig64) < 1e-6
    assert rel(ctan, m.ctan_lin()) < 1e-6
    # the oracle (CPU restatement of the reference)
    p, k = ob.make_params(size, ob.MIC3D_RANDOM_SPHERES, GEO, elastic(ob), ngp=6, coupling=[ob.FE_FULL] * 6, **kw)
    o = oracle.micro(p, k)
    for g in range(6):
        o.set_strain(g, strains[g])
    o.homogenize()
    assert rel(sig, np.array([o.get_stress(g) for g in range(6)])) < 1e-6
    assert rel(ctan, o.ctan_lin()) < 1e-6
    # displacement fields
    u64 = m.state(0, 1)
    assert rel(ms.u(6), u64) < 1e-6


@pytest.mark.parametrize("size", [(3, 3, 3), (4, 5, 7), (9, 6, 13), (17, 5, 8)])
def test_mixed_odd_shapes(mp, size):
    """Non-cubic and tiny grids (one interior node at 3^3; x-rows shorter than a 64-node slice, so the
    DPP neighbour lanes and the clamped addresses cross rows and slices): sigma and ctan against the
    FP64 micropp<3> path."""
    rng = np.random.default_rng(sum(size))
    strains = rng.uniform(-5e-3, 5e-3, (6, 6))
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp))
    sig, ctan, stats, resid, ok = mp.homogenize_mixed(mp.MSolver(r, 12), strains)
    assert ok and resid.max() < 1e-8
    kw = dict(lin_stress=0, cg_rel_tol=1e-8, cg_max_its=3000)
    m = mp.Micropp3(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp), ngp=6, coupling=[mp.FE_FULL] * 6, **kw)
    for g in range(6):
        m.set_strain(g, strains[g])
    m.homogenize()
    sig64 = np.array([m.get_stress(g) for g in range(6)])
    print(f"{size}: outer {stats[0]} inner {stats[1]}; sigma rel {rel(sig, sig64):.1e} ctan rel "
          f"{rel(ctan, m.ctan_lin()):.1e}", flush=True)
    assert rel(sig, sig64) < 1e-6 and rel(ctan, m.ctan_lin()) < 1e-6


@pytest.mark.parametrize("K", [4, 8, 16])
def test_mixed_rhs_counts_agree(mp, K):
    r = mp.Rve(11, mp.MIC_SPHERE, (0.2, 0, 0, 0), elastic(mp))
    eps = np.random.default_rng(3).uniform(-5e-3, 5e-3, (K, 6))
    ref = mp.MSolver(r, 16).solve(eps)[0]
    sig, stats, resid, ok = mp.MSolver(r, K).solve(eps)
    assert ok
    assert rel(sig, ref) < 1e-7
    # fewer active strains than K: the idle fields stay zero and do not hold the solve back
    sig2, _, _, ok2 = mp.MSolver(r, K).solve(eps[:1])
    assert ok2 and rel(sig2, ref[:1]) < 1e-7


def test_mixed_patch_test(mp):
    """Homogeneous RVE: sigma = C eps for every field."""
    m3 = [mp.material(0, 1e6, 0.3)] * 3
    r = mp.Rve(21, mp.MIC_HOMOGENEOUS, (0.1,) * 4, m3)
    eps = np.random.default_rng(5).uniform(-5e-3, 5e-3, (4, 6))
    sig, stats, resid, ok = mp.MSolver(r, 4).solve(eps)
    lam, mu = 0.3 * 1e6 / (1.3 * 0.4), 1e6 / 2.6
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[range(3, 6), range(3, 6)] = mu
    assert ok and rel(sig, eps @ C.T) < 1e-8


def test_mixed_rejects_nonlinear_and_bad_k(mp):
    plastic = [mp.material(1, 1e6, 0.3, 1e5, 1e3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    with pytest.raises(mp.MppdError):
        mp.MSolver(mp.Rve(5, mp.MIC_SPHERE, (0.2,) * 4, plastic), 4)
    with pytest.raises(mp.MppdError):
        mp.MSolver(mp.Rve(5, mp.MIC_SPHERE, (0.2,) * 4, elastic(mp)), 6)


def test_mixed_full_size(mp):
    """50^3 elements, the benchmark geometry: mixed vs FP64 on two strains."""
    r = mp.Rve(51, mp.MIC3D_RANDOM_SPHERES, (2024.0, 8.0, 0.05, 0.15), elastic(mp))
    strains = np.random.default_rng(7).uniform(-5e-3, 5e-3, (2, 6))
    sig, ctan, stats, resid, ok = mp.homogenize_mixed(mp.MSolver(r, 8), strains)
    assert ok
    s = mp.Solver(r)
    u = mp.DeviceArray(r.nrow)
    for k in range(2):
        mp.load().mppd_memset(u.ptr, 0, u.nbytes)
        s.newton(strains[k], u, cg_rel_tol=1e-8, cg_max_its=3000)
        s64 = s.ave_stress(u)
        print(f"50^3 strain {k}: rel {rel(sig[k], s64):.2e}; stats {stats}", flush=True)
        assert rel(sig[k], s64) < 1e-6


@pytest.mark.parametrize("K", [4, 12, 16])
def test_mixed_solve_is_deterministic(mp, K):
    """Repeated solves agree bit for bit: the SpMV's LDS staging (direct global-to-LDS loads shared by the
    field-group waves) and every reduction run in a fixed order, so any lost synchronisation would show as a
    difference between runs.  Several fields stop early, so waves with stopped fields are exercised too."""
    r = mp.Rve(31, mp.MIC3D_RANDOM_SPHERES, (77.0, 6.0, 0.05, 0.2), elastic(mp))
    rng = np.random.default_rng(5)
    eps = rng.uniform(-5e-3, 5e-3, (K, 6))
    eps[::3] *= 1e-3  # small fields reach the absolute floor at other iterations
    ms = mp.MSolver(r, K)
    first, st0, _, ok0 = ms.solve(eps)
    assert ok0
    for rep in range(3):
        sig, st, _, ok = mp.MSolver(r, K).solve(eps) if rep == 1 else ms.solve(eps)
        assert ok and st == st0
        assert np.array_equal(sig, first), f"repeat {rep}: max diff {np.abs(sig - first).max():.3e}"


@pytest.mark.parametrize("size", [(51, 11, 41), (13, 9, 17), (31, 31, 31)])
def test_coarse_level_against_jacobi(mp, monkeypatch, size):
    """The two-level preconditioner of the inner CG (coarse grids with h = 5, 2, 2 / 2, 2, 2 / 3, 3, 3 fine
    elements per coarse element) changes the inner iterations only: sigma and ctan agree with the
    Jacobi-only solve (MPPD_COARSE=0) and the FP64 path, and the inner iterations drop."""
    rng = np.random.default_rng(sum(size))
    strains = rng.uniform(-5e-3, 5e-3, (6, 6))
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp))
    sig, ctan, stats, resid, ok = mp.homogenize_mixed(mp.MSolver(r, 12), strains)
    monkeypatch.setenv("MPPD_COARSE", "0")
    sig_j, ctan_j, stats_j, resid_j, ok_j = mp.homogenize_mixed(mp.MSolver(r, 12), strains)
    monkeypatch.delenv("MPPD_COARSE")
    print(f"{size}: inner iterations {stats[1]} (Jacobi {stats_j[1]}); sigma rel {rel(sig, sig_j):.1e} ctan rel "
          f"{rel(ctan, ctan_j):.1e}", flush=True)
    assert ok and ok_j and resid.max() < 1e-8 and resid_j.max() < 1e-8
    assert rel(sig, sig_j) < 1e-7 and rel(ctan, ctan_j) < 1e-7
    assert stats[1] < 0.7 * stats_j[1]
    kw = dict(lin_stress=0, cg_rel_tol=1e-8, cg_max_its=3000)
    m = mp.Micropp3(size, mp.MIC3D_RANDOM_SPHERES, GEO, elastic(mp), ngp=6, coupling=[mp.FE_FULL] * 6, **kw)
    for g in range(6):
        m.set_strain(g, strains[g])
    m.homogenize()
    assert rel(sig, np.array([m.get_stress(g) for g in range(6)])) < 1e-6
    assert rel(ctan, m.ctan_lin()) < 1e-6
