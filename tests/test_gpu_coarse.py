"""GPU: the coarse level of the multi-RHS solver's two-level preconditioner.

The solver's FP32 coarse inverse (mppd_msolver_coarse) against the inverse of
the Galerkin matrix P^T A P formed here independently: A is the FP64 matrix
the FP64 path assembles (mppd_assembly_mat, downloaded as ELL), P the
trilinear hats of the interior coarse nodes (numpy), the inverse numpy's LU.
This pins the device assembly of A_c (per-phase tables, coarse elements, band
tiles) and its banded tile Cholesky and inverse together.

Bar: max |ainv - inv(P^T A P)| <= 1e-6 max |inv(P^T A P)| (FP32 rounding of
the entries is 6e-8 of each entry; the FP64 factorisation adds about the
condition number times 1e-16), and the inverse exactly symmetric.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

GEO = (777.0, 6.0, 0.08, 0.2)


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def hats(n_nodes, h, m):
    """n_nodes x m: the hat of interior coarse node c (1 .. m) at fine index i, max(0, 1 - |i - c h| / h)."""
    i = np.arange(n_nodes)[:, None]
    c = np.arange(1, m + 1)[None, :]
    return np.maximum(0.0, 1.0 - np.abs(i - c * h) / h)


def galerkin(mp, r, dims):
    nc, hx, hy, hz, mx, my, mz = (int(v) for v in dims)
    nx, ny, nz = r.size
    s = mp.Solver(r)
    s.assembly_mat(None)
    cols, vals = s.ell()
    rows = np.repeat(np.arange(r.nrow), cols.shape[1])
    A = sp.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(r.nrow, r.nrow))
    # node order x fastest (n = (k ny + j) nx + i), coarse node order likewise, DOF = node * 3 + d
    Pn = sp.kron(sp.csr_matrix(hats(nz, hz, mz)),
                 sp.kron(sp.csr_matrix(hats(ny, hy, my)), sp.csr_matrix(hats(nx, hx, mx))))
    P = sp.kron(Pn, sp.identity(3)).tocsr()
    assert P.shape == (r.nrow, nc)
    return (P.T @ A @ P).toarray()


@pytest.mark.parametrize("size", [(21, 21, 21), (21, 13, 17), (41, 41, 41), (51, 51, 51)])
def test_coarse_inverse_matches_galerkin_inverse(mp, size):
    mats = [mp.material(0, 1e6, 0.3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    r = mp.Rve(size, mp.MIC3D_RANDOM_SPHERES, GEO, mats)
    ms = mp.MSolver(r, 4)
    dims, ainv = ms.coarse()
    print(f"size {size}: coarse dims {dims.tolist()}", flush=True)
    assert dims[0] > 0 and ainv.shape == (dims[0], dims[0])
    assert np.array_equal(ainv, ainv.T)
    ref = np.linalg.inv(galerkin(mp, r, dims))
    err = np.abs(ainv.astype(np.float64) - ref).max() / np.abs(ref).max()
    print(f"size {size}: max |ainv - inv(P^T A P)| / max |inv| = {err:.2e}", flush=True)
    assert err < 1e-6


def test_no_coarse_level_on_a_small_grid(mp):
    """A side of 3 elements has no aligned coarse grid with 2 coarse elements of >= 2 fine ones: Jacobi alone."""
    mats = [mp.material(0, 1e6, 0.3), mp.material(0, 1e7, 0.25), mp.material(0, 1e7, 0.25)]
    r = mp.Rve((4, 9, 9), mp.MIC3D_RANDOM_SPHERES, GEO, mats)
    dims, ainv = mp.MSolver(r, 4).coarse()
    assert dims[0] == 0 and ainv is None
