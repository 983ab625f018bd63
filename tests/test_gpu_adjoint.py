"""GPU: adjoint derivatives of the volume-averaged stress (elastic RVEs).

d sigma / d eps from the six adjoint fields is checked against the ctan of
the forward solves (the reference's calc_ctan_lin_fe_models algorithm, on the
GPU and in the oracle); d sigma / d (E, nu) of each phase against central
finite differences with a relative step of 1e-6.  The derivatives are also
wrapped in a torch.autograd.Function (test harness only) and checked with
torch.autograd.gradcheck.
"""
import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu

SPEC = [(0, 1e6, 0.3), (0, 1e7, 0.25), (0, 1e7, 0.25)]
GEO = (31.0, 5.0, 0.08, 0.2)


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def sigma_of(mp, size, spec, strains, rel_tol=1e-13):
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in spec])
    sig, _, _, ok = mp.MSolver(r, 8).solve(strains, rel_tol=rel_tol, max_outer=30)
    assert ok
    return sig


@pytest.mark.parametrize("size", [9, 17])
def test_adjoint_ctan_matches_forward_and_oracle(mp, oracle, size):
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in SPEC])
    ctan_adj, _, _, stats, resid, ok = mp.MSolver(r, 8).adjoint(rel_tol=1e-10)
    assert ok
    _, ctan_fwd, _, _, ok2 = mp.homogenize_mixed(mp.MSolver(r, 12), np.zeros((1, 6)), rel_tol=1e-10)
    assert ok2
    print(f"size {size}: adjoint vs forward ctan rel {rel(ctan_adj, ctan_fwd):.2e}; stats {stats}", flush=True)
    assert rel(ctan_adj, ctan_fwd) < 1e-6
    # the oracle's calc_ctan_lin_fe_models (FP64 CG, six 1e-8 perturbations)
    p, k = ob.make_params(size, ob.MIC3D_RANDOM_SPHERES, GEO, [ob.material(*s) for s in SPEC], ngp=1,
                          coupling=[ob.FE_FULL], cg_rel_tol=1e-10, cg_max_its=4000)
    o = oracle.micro(p, k)
    print(f"  vs oracle ctan rel {rel(ctan_adj, o.ctan_lin()):.2e}", flush=True)
    assert rel(ctan_adj, o.ctan_lin()) < 1e-6
    # symmetry of the elastic tangent
    assert rel(ctan_adj, ctan_adj.T) < 1e-7


def test_adjoint_material_derivatives_vs_central_fd(mp):
    size = 7
    strains = np.random.default_rng(4).uniform(-5e-3, 5e-3, (2, 6))
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in SPEC])
    ctan, sig, dmat, stats, resid, ok = mp.MSolver(r, 8).adjoint(strains, rel_tol=1e-12, max_outer=30)
    assert ok
    assert rel(sig, sigma_of(mp, size, SPEC, strains)) < 1e-9
    assert rel(sig, strains @ ctan.T) < 1e-8  # linear elasticity: sigma = ctan eps
    et = r.elem_types()
    for p in range(3):
        for w, name in ((1, "E"), (2, "nu")):
            if not (et == p).any():
                assert np.all(dmat[:, :, p, w - 1] == 0.0)
                continue
            h = SPEC[p][w] * 1e-6
            sp, sm = [list(s) for s in SPEC], [list(s) for s in SPEC]
            sp[p][w] += h
            sm[p][w] -= h
            fd = (sigma_of(mp, size, sp, strains) - sigma_of(mp, size, sm, strains)) / (2 * h)
            got = dmat[:, :, p, w - 1]
            print(f"phase {p} d/d{name}: rel err vs central FD {rel(got, fd):.2e}", flush=True)
            assert rel(got, fd) < 1e-5


def test_adjoint_autograd_gradcheck(mp):
    torch = pytest.importorskip("torch")
    size = 6
    base = np.array([[s[1], s[2]] for s in SPEC[:2]])  # (E, nu) of phases 0, 1

    class AdjointSigma(torch.autograd.Function):
        """sigma(eps, theta) with theta = (E, nu) of phases 0 and 1 relative to base; backward from the adjoint."""

        @staticmethod
        def forward(ctx, eps, theta):
            spec = [list(s) for s in SPEC]
            th = theta.detach().numpy().reshape(2, 2) * base
            for p in range(2):
                spec[p][1], spec[p][2] = th[p]
            r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in spec])
            ctan, sig, dmat, _, _, ok = mp.MSolver(r, 8).adjoint(eps.detach().numpy()[None] * 1e-3, rel_tol=1e-13,
                                                                 max_outer=30)
            assert ok
            jac_t = torch.from_numpy(dmat[0, :, :2, :] * base[None]).reshape(6, 4)  # d sigma / d theta
            ctx.save_for_backward(torch.from_numpy(ctan * 1e-3), jac_t)
            return torch.from_numpy(sig[0])

        @staticmethod
        def backward(ctx, g):
            ctan, jac_t = ctx.saved_tensors
            return ctan.T @ g, jac_t.T @ g

    eps = torch.tensor([1.0, -0.5, 0.3, 0.2, -0.1, 0.4], dtype=torch.float64, requires_grad=True)
    theta = torch.ones(4, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(AdjointSigma.apply, (eps, theta), eps=1e-6, atol=1e-4, rtol=1e-5)


def test_adjoint_many_equals_one_by_one(mp):
    """mppd_msolver_adjoint_many (the batched adjoint: solvers of 6 RVEs of three sizes, on 3 host threads,
    each solver on its own stream) gives each RVE's ctan, sigma, d sigma / d (E, nu) and residual ratios bit for
    bit as mppd_msolver_adjoint one RVE after another; a solver given twice is refused."""
    rng = np.random.default_rng(21)
    sizes = [9, 13, 17, 9, 13, 17]
    rves = [mp.Rve(n, mp.MIC3D_RANDOM_SPHERES, (float(100 + i), float(1 + i % 5), 0.08, 0.2),
                   [mp.material(*s) for s in SPEC]) for i, n in enumerate(sizes)]
    ms = [mp.MSolver(r, 8) for r in rves]
    strains = rng.uniform(-5e-3, 5e-3, (len(ms), 2, 6))
    one = [m.adjoint(strains[i]) for i, m in enumerate(ms)]
    ctan, sig, dmat, stats, relr, ok = mp.adjoint_many(ms, strains, workers=3)
    assert ok
    for i in range(len(ms)):
        assert np.array_equal(ctan[i], one[i][0]) and np.array_equal(sig[i], one[i][1])
        assert np.array_equal(dmat[i], one[i][2]) and tuple(stats[i]) == one[i][3]
        assert np.array_equal(relr[i], one[i][4])
    # ctan alone (n = 0), all solvers on one thread
    c0 = mp.adjoint_many(ms, None, workers=1)[0]
    for i in range(len(ms)):
        assert np.array_equal(c0[i], ms[i].adjoint()[0])
    with pytest.raises(mp.MppdError):
        mp.adjoint_many([ms[0], ms[0]], strains[:2])


def test_adjoint_argument_checks(mp):
    r = mp.Rve(5, mp.MIC_SPHERE, (0.2, 0, 0, 0), [mp.material(*s) for s in SPEC])
    with pytest.raises(mp.MppdError):
        mp.MSolver(r, 8).adjoint(np.zeros((3, 6)))  # 6 + 3 > K = 8


def test_adjoint_failed_first_allocation_is_not_kept(mp, monkeypatch):
    """The adjoint buffers of an msolver are made on the first mppd_msolver_adjoint call.  When that allocation
    fails (forced from the host with MPPD_FAIL_ALLOC=1), none of them may stay behind: the next call allocates
    afresh and equals, bit for bit, the adjoint of an msolver that never failed."""
    r = mp.Rve(9, mp.MIC3D_RANDOM_SPHERES, GEO, [mp.material(*s) for s in SPEC])
    strains = np.array([[2e-3, -1e-3, 5e-4, 1e-3, 0.0, -5e-4]])
    good, ms = mp.MSolver(r, 8), mp.MSolver(r, 8)
    ref = good.adjoint(strains, rel_tol=1e-10)
    monkeypatch.setenv("MPPD_FAIL_ALLOC", "1")
    with pytest.raises(mp.MppdError, match="allocation"):
        ms.adjoint(strains, rel_tol=1e-10)
    monkeypatch.delenv("MPPD_FAIL_ALLOC")
    got = ms.adjoint(strains, rel_tol=1e-10)
    assert got[5] and ref[5]
    for a, b in zip(got[:3], ref[:3]):
        assert np.array_equal(a, b)
