"""GPU: mppd_cg_noalloc, the Jacobi-PCG whose working set lives in the library's fixed device scratch array,
against mppd_cg (the same kernels on the solver's own buffers) through the C ABI.

Two calls in a row, the second on a smaller RVE (the scratch keeps the first call's contents), each bit-equal
to mppd_cg in x, iterations and r.z, and within the CG tolerance of the oracle's ell_solve_cgpd; then a solve
with A0 (the initial matrix of a stiffer RVE of the same grid), and the size limit, refused on the host
before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def elastic(mp, e_matrix=1e6, e_incl=1e7):
    return [mp.material(0, e_matrix, 0.3), mp.material(0, e_incl, 0.25), mp.material(0, e_incl, 0.25)]


def system(mp, n, eps, mats=None):
    """An assembled RVE of n^3 nodes (centred sphere) and the right-hand side of the strain eps."""
    r = mp.Rve((n, n, n), mp.MIC_SPHERE, (0.2, 0, 0, 0), mats or elastic(mp))
    s = mp.Solver(r)
    u = mp.DeviceArray(r.nrow)
    s.set_displ_bc(eps, u)
    s.assembly_mat(u)
    b = mp.DeviceArray(r.nrow)
    s.assembly_rhs(u, None, b)
    return r, s, b


def noalloc(mp, s, b, x, min_err, rel_err, max_its, A0=None):
    its, err = C.c_int(), C.c_double()
    rc = mp.load().mppd_cg_noalloc(s.h, None if A0 is None else A0.h, b.ptr, x.ptr, min_err, rel_err, max_its,
                                   C.byref(its), C.byref(err))
    return rc, its.value, err.value


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n, tol", [(31, 1e-8), (11, 1e-10)])
def test_noalloc_cg_equals_cg(mp, oracle, n, tol):
    """The larger RVE first, then the smaller one (parametrize order): the second call runs on a scratch array
    that still holds the first call's vectors."""
    print(f"mppd_cg_noalloc vs mppd_cg: {n}^3 nodes, tol {tol}", flush=True)
    r, s, b = system(mp, n, [5e-3, -1e-3, 2e-3, 1e-3, 0.0, -2e-3])
    x_ref, x_na = mp.DeviceArray(r.nrow), mp.DeviceArray(r.nrow)
    its_ref, err_ref = s.cg(b, x_ref, 1e-50, tol, 4000)
    rc, its, err = noalloc(mp, s, b, x_na, 1e-50, tol, 4000)
    assert rc == 0, mp.load().mppd_last_error()
    a, ref = x_na.numpy(), x_ref.numpy()
    print(f"  its {its} vs {its_ref}, err {err:.6e} vs {err_ref:.6e}, max |dx| {np.abs(a - ref).max():.1e}",
          flush=True)
    assert its == its_ref and err == err_ref
    assert np.array_equal(a, ref)
    cols, vals = s.ell()
    xo, io, _ = oracle.ell_cg(r.nrow // 3, 3, 81, 13, cols, vals, b.numpy(), 1e-50, tol, 4000)
    assert abs(its - io) <= 1 and rel(a, xo) < 10 * tol


def test_noalloc_cg_with_A0(mp):
    """A0: the matrix of a 3x stiffer RVE of the same grid, solved for this RVE's right-hand side."""
    eps = [2e-3, 0, 0, 0, 1e-3, 0]
    r, s, b = system(mp, 15, eps)
    _, s0, _ = system(mp, 15, eps, elastic(mp, 3e6, 3e7))
    x_ref, x_na = mp.DeviceArray(r.nrow), mp.DeviceArray(r.nrow)
    its_ref, err_ref = s.cg(b, x_ref, 1e-50, 1e-9, 2000, A0=s0)
    rc, its, err = noalloc(mp, s, b, x_na, 1e-50, 1e-9, 2000, A0=s0)
    assert rc == 0, mp.load().mppd_last_error()
    assert its == its_ref and err == err_ref
    assert np.array_equal(x_na.numpy(), x_ref.numpy())
    # the stiffer matrix gives a third of the displacement of the RVE's own matrix
    x_own = mp.DeviceArray(r.nrow)
    s.cg(b, x_own, 1e-50, 1e-9, 2000)
    assert rel(3 * x_na.numpy(), x_own.numpy()) < 1e-6


def test_noalloc_cg_refuses_rves_beyond_the_scratch(mp):
    """52^3 nodes = 421 824 rows > MPPD_NOALLOC_MAX_ROWS: refused with MPPD_ERR_ARG before any kernel runs."""
    r = mp.Rve((52, 52, 52), mp.MIC_SPHERE, (0.2, 0, 0, 0), elastic(mp))
    s = mp.Solver(r)
    b, x = mp.DeviceArray(r.nrow), mp.DeviceArray(r.nrow)
    rc, _, _ = noalloc(mp, s, b, x, 1e-50, 1e-8, 100)
    assert rc == -1
    assert b"MPPD_NOALLOC_MAX_ROWS" in mp.load().mppd_last_error()
