"""Pins the CPU oracle (oracle/micropp_oracle.c) before anything is checked against it.

Three independent kinds of evidence:
  1. the golden values the Micropp reference's own tests hold
     (tests/golden/reference_tables.json);
  2. bit-for-bit agreement with the compiled reference (oracle/_ref), for
     every function the GPU path restates;
  3. closed forms and independent numpy restatements (patch test, a dense
     global stiffness built from an independent B-matrix, Hill bounds,
     splitmix64 and the sphere centroid rule).
"""
import numpy as np
import pytest

from oracle.bindings import (FE_FULL, FE_LINEAR, FE_ONE_WAY, MIC3D_8, MIC3D_RANDOM_SPHERES, MIC_CILI_FIB_Z,
                             MIC_HOMOGENEOUS, MIC_LAYER_Y, MIC_SPHERE, MIX_RULE_CHAMIS, NVAR, make_params, material)

# geometry parameters that give both phases for every reference micro type
GEO = {0: 0.1, 1: 0.2, 2: 0.3, 3: 0.2, 4: 0.2, 5: 0.1, 6: 0.1, 7: 0.1, 8: 0.1, 9: 0.1, 10: 0.1, 11: 0.1, 12: 0.1}
ELAST2 = [material(0, 1e6, 0.3), material(0, 1e7, 0.25), material(0, 1e7, 0.25)]


def run_steps(lib, tab, n=None):
    m = material(*tab["material"])
    mats = [m, material(0, 3.0e7, 0.25), material(0, 3.0e7, 0.25)]
    p, k = make_params(n or tab["n"], MIC_HOMOGENEOUS, mats=mats, lin_stress=0)
    micro = lib.micro(p, k)
    t, dt = 0.0, tab["t_final"] / tab["steps"]
    eps = np.zeros(6)
    out = []
    for _ in range(tab["steps"]):
        eps[tab["dir"]] = tab["eps_max"] * t
        micro.set_strain(0, eps)
        micro.homogenize()
        out.append(micro.get_stress(0))
        micro.update_vars()
        t += dt
    return np.array(out)


# --------------------------------------------------------------------------
# 1. reference golden tables
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["benchmark_elastic", "benchmark_plastic", "benchmark_damage"])
def test_reference_sigma_tables(oracle, tables, name):
    tab = tables[name]
    sig = run_steps(oracle, tab)
    assert np.all(np.abs(sig - np.array(tab["sig"])) < tab["tol"]), name


def test_reference_ell_1(oracle, tables):
    t = tables["test_ell_1"]
    cols, vals = oracle.ell2d_build(t["nx"], t["ny"], t["Ae"])
    assert np.array_equal(cols.ravel(), np.array(t["cols"]))
    for i in list(range(4)) + list(range(5, 9)):
        assert np.array_equal(vals[i], np.array(t["dir_row"], float))
    assert np.all(np.abs(vals[4] - np.array(t["int_row"])) < t["int_row_tol"])
    b = np.zeros(9)
    b[t["b_index"]] = 1.0
    x, its, err = oracle.ell_cg(9, 1, 9, 4, cols, vals, b, t["cg_min_err"], t["cg_rel_err"], t["cg_max_its"])
    assert np.all(np.abs(x - np.array(t["x"])) < t["x_tol"])
    assert err < t["cg_err_max"] and its == t["cg_its"]


def test_reference_ell_2_sizes(oracle, tables):
    n = tables["test_ell_2"]["n"]
    cols = oracle.ell_cols3d(n, n, n)
    assert cols.shape == (3 * n ** 3, 3 * tables["test_ell_2"]["nnz_per_field"])


def test_reference_elem_nodes(oracle, tables):
    t = tables["test_get_elem_nodes"]
    for (ex, ey, ez), want in t["cases"]:
        assert list(oracle.elem_nodes(t["nx"], t["ny"], ex, ey, ez)) == want


def _mats(rows):
    return [material(*r) for r in rows]


def test_reference_batch_invariant_3d_4(oracle, tables):
    t = tables["test3d_4"]
    p, k = make_params(t["n"], MIC_SPHERE, (0.1, 0.1, 0.1, 0.1), _mats(t["materials"]), ngp=t["ngp"])
    m = oracle.micro(p, k)
    eps = np.zeros(6)
    for _ in range(t["steps"]):
        eps[t["dir"]] += t["d_eps"]
        for g in range(t["ngp"]):
            m.set_strain(g, eps)
        m.homogenize()
        s0 = m.get_stress(0)
        for g in range(1, t["ngp"]):
            assert np.all(np.abs(m.get_stress(g)[:3] - s0[:3]) < t["tol"])
        m.update_vars()


def test_reference_batch_invariant_3d_5(oracle, tables):
    t = tables["test3d_5"]
    p, k = make_params(t["n"], MIC_SPHERE, (0.1, 0.1, 0.1, 0.1), _mats(t["materials"]), ngp=t["ngp"])
    m = oracle.micro(p, k)
    e1, e2 = np.zeros(6), np.zeros(6)
    for _ in range(t["steps"]):
        e1[t["dir"]] += t["d_eps_1"]
        e2[t["dir"]] += t["d_eps_2"]
        for g in range(t["ngp"]):
            m.set_strain(g, e1 if g % 2 else e2)
        m.homogenize()
        s = [m.get_stress(g) for g in range(t["ngp"])]
        for g in range(2, t["ngp"]):
            assert np.all(np.abs(s[g][:3] - s[g % 2][:3]) < t["tol"])
        m.update_vars()


# --------------------------------------------------------------------------
# 2. bit-for-bit against the compiled reference
# --------------------------------------------------------------------------
@pytest.mark.parametrize("size", [(2, 2, 2), (3, 4, 5), (6, 6, 6)])
def test_cols_bitwise(oracle, ref, size):
    assert np.array_equal(oracle.ell_cols3d(*size), ref.ell_cols3d(*size))


@pytest.mark.parametrize("mtype", list(range(13)))
def test_phase_maps_bitwise(oracle, ref, mtype):
    for size in [(3, 3, 3), (4, 4, 4), (11, 11, 11), (12, 12, 12), (5, 8, 11), (22, 22, 22)]:
        geo = (GEO[mtype],) * 4
        p, k = make_params(size, mtype, geo, ELAST2, calc_ctan_lin=0)
        rm = ref.micro(p, k)
        et, rc = oracle.elem_types(mtype, geo, *size)
        assert rc == 0
        assert np.array_equal(et, rm.elem_types()), (mtype, size)
        assert np.array_equal(oracle.bmat(*size), rm.bmat())


def _rand_state(rng, nvars, kind):
    v = np.zeros(nvars)
    h = v.reshape(-1, NVAR)
    if kind == "plastic":
        h[:, :6] = rng.normal(0, 2e-3, (h.shape[0], 6))
        h[:, 6] = rng.uniform(0, 1e-2, h.shape[0])
    elif kind == "damage":
        h[:, 0] = rng.uniform(0, 3e2, h.shape[0])
        h[:, 1] = rng.uniform(-0.5, 0.5, h.shape[0])
    return v


MATSETS = {
    "elastic": [material(0, 1e6, 0.3), material(0, 1e7, 0.25), material(0, 1e7, 0.25)],
    "plastic": [material(1, 1e6, 0.3, 1e5, 1e3), material(0, 1e7, 0.25), material(1, 2e6, 0.3, 1e4, 5e2)],
    "damage": [material(2, 1e6, 0.3, 0, 0, 1e3), material(0, 1e7, 0.25), material(2, 1e6, 0.3, 0, 0, 1e3)],
}


@pytest.mark.parametrize("kind", ["elastic", "plastic", "damage"])
def test_materials_bitwise(oracle, ref, kind):
    rng = np.random.default_rng(1)
    m = MATSETS[kind][0]
    for trial in range(200):
        eps = rng.normal(0, 10 ** rng.uniform(-5, -1), 6)
        hist = _rand_state(rng, NVAR, kind) if trial % 2 else None
        assert np.array_equal(oracle.mat_stress(m, eps, hist), ref.mat_stress(m, eps, hist))
        assert np.array_equal(oracle.mat_ctan(m, eps, hist), ref.mat_ctan(m, eps, hist))
        ho, hr = np.zeros(NVAR), np.zeros(NVAR)
        assert oracle.mat_evolute(m, eps, hist, ho) == ref.mat_evolute(m, eps, hist, hr)
        assert np.array_equal(ho, hr)


@pytest.mark.parametrize("kind", ["elastic", "plastic", "damage"])
def test_assembly_bitwise(oracle, ref, kind):
    rng = np.random.default_rng(2)
    size = (5, 4, 6)
    p, k = make_params(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind], calc_ctan_lin=0)
    rm = ref.micro(p, k)
    orve = oracle.rve(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind])
    u = rng.normal(0, 1e-3, rm.nndim)
    vars_ = _rand_state(rng, rm.nvars, kind) if kind != "elastic" else None
    for v in (None, vars_):
        assert np.array_equal(orve.assembly_mat(u, v), rm.assembly_mat(u, v))
        bo, no = orve.assembly_rhs(u, v)
        br, nr = rm.assembly_rhs(u, v)
        assert np.array_equal(bo, br) and no == nr
        assert np.array_equal(orve.ave_stress(u, v), rm.ave_stress(u, v))
        vo, nlo = orve.vars_new(u, v)
        vr, nlr = rm.vars_new(u, v)
        assert nlo == nlr and np.array_equal(vo, vr)
    eps = rng.normal(0, 1e-3, 6)
    u1, u2 = u.copy(), u.copy()
    orve.set_displ_bc(eps, u1)
    rm.set_displ_bc(eps, u2)
    assert np.array_equal(u1, u2)


def test_spmv_cg_bitwise(oracle, ref):
    size = (6, 6, 6)
    p, k = make_params(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS["elastic"], calc_ctan_lin=0)
    rm = ref.micro(p, k)
    u = np.zeros(rm.nndim)
    rm.set_displ_bc([1e-3, -2e-3, 5e-4, 1e-3, 0, 2e-3], u)
    vals = rm.assembly_mat(u)
    b, _ = rm.assembly_rhs(u)
    cols = oracle.ell_cols3d(*size)
    x = np.random.default_rng(3).normal(size=rm.nndim)
    assert np.array_equal(oracle.ell_mvp(cols, vals, x), ref.ell_mvp(cols, vals, x))
    nn = rm.nndim // 3
    for tol in (1e-5, 1e-8, 1e-12):
        xo, io, eo = oracle.ell_cg(nn, 3, 81, 13, cols, vals, b, 1e-50, tol, 2000)
        xr, ir, er = ref.ell_cg(nn, 3, 81, 13, cols, vals, b, 1e-50, tol, 2000)
        assert io == ir and eo == er and np.array_equal(xo, xr)


@pytest.mark.parametrize("kind", ["elastic", "plastic", "damage"])
def test_newton_bitwise(oracle, ref, kind):
    size = (6, 6, 6)
    p, k = make_params(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind], calc_ctan_lin=0)
    rm = ref.micro(p, k)
    orve = oracle.rve(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind])
    rng = np.random.default_rng(4)
    vars_ = _rand_state(rng, rm.nvars, kind) if kind != "elastic" else None
    eps = np.array([4e-3, -1e-3, 1e-3, 2e-3, -3e-3, 1e-3])
    for tol in (1e-5, 1e-8):
        uo, ur = np.zeros(rm.nndim), np.zeros(rm.nndim)
        ro = orve.newton(eps, uo, vars_, cg_rel_tol=tol)
        rr = rm.newton(eps, ur, vars_, cg_rel_tol=tol)
        assert ro == rr and np.array_equal(uo, ur)


def _drive(lib, p, k, strains, steps):
    m = lib.micro(p, k)
    rec = []
    for s in range(steps):
        for g in range(p.ngp):
            m.set_strain(g, strains[g] * (s + 1))
        m.homogenize()
        rec.append([(m.get_stress(g), m.get_ctan(g), m.get_cost(g), m.has_converged(g), m.has_subiterated(g),
                     m.is_non_linear(g)) for g in range(p.ngp)])
        m.update_vars()
    states = [[m.state(g, w) for w in range(4)] for g in range(p.ngp)]
    return rec, states, m.ctan_lin(), m.volume_fractions()


@pytest.mark.parametrize("kind,lin_stress,subits", [("plastic", 0, 0), ("damage", 0, 1), ("plastic", 1, 1),
                                                    ("elastic", 0, 0)])
def test_instance_bitwise(oracle, ref, kind, lin_stress, subits):
    coupling = [FE_FULL, FE_ONE_WAY, FE_LINEAR, MIX_RULE_CHAMIS]
    p, k = make_params(5, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind], ngp=4, coupling=coupling,
                       lin_stress=lin_stress, subiterations=subits, nsubiterations=3, nr_max_its=2)
    rng = np.random.default_rng(5)
    amp = {"elastic": 2e-3, "plastic": 2e-3, "damage": 2e-3}[kind]
    strains = [rng.uniform(-amp, amp, 6) for _ in range(4)]
    a = _drive(oracle, p, k, strains, 4)
    b = _drive(ref, p, k, strains, 4)
    for sa, sb in zip(a[0], b[0]):
        for ga, gb in zip(sa, sb):
            assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1])
            assert ga[2:] == gb[2:]
    for ga, gb in zip(a[1], b[1]):
        for xa, xb in zip(ga, gb):
            assert (xa is None) == (xb is None)
            if xa is not None:
                assert np.array_equal(xa, xb)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_instance_use_A0_bitwise(oracle, ref):
    p, k = make_params(5, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS["plastic"], ngp=2, coupling=[FE_FULL, FE_ONE_WAY],
                       use_A0=1, its_with_A0=1, lin_stress=0)
    strains = [np.array([3e-3, 0, 0, 1e-3, 0, 0]), np.array([0, -2e-3, 1e-3, 0, 0, 1e-3])]
    a = _drive(oracle, p, k, strains, 3)
    b = _drive(ref, p, k, strains, 3)
    for sa, sb in zip(a[0], b[0]):
        for ga, gb in zip(sa, sb):
            assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1]) and ga[2:] == gb[2:]


# --------------------------------------------------------------------------
# 3. closed forms and independent restatements
# --------------------------------------------------------------------------
def c_iso(E, nu):
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def independent_B(dx, dy, dz):
    """Trilinear B matrices at the 2x2x2 Gauss points, written from the shape functions."""
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    g = 1 / np.sqrt(3)
    Bs = []
    for q in s * g:
        dN = np.stack([s[:, 0] * (1 + s[:, 1] * q[1]) * (1 + s[:, 2] * q[2]) / 8 * 2 / dx,
                       s[:, 1] * (1 + s[:, 0] * q[0]) * (1 + s[:, 2] * q[2]) / 8 * 2 / dy,
                       s[:, 2] * (1 + s[:, 0] * q[0]) * (1 + s[:, 1] * q[1]) / 8 * 2 / dz], 1)
        B = np.zeros((6, 24))
        for a in range(8):
            B[0, 3 * a], B[1, 3 * a + 1], B[2, 3 * a + 2] = dN[a]
            B[3, 3 * a], B[3, 3 * a + 1] = dN[a, 1], dN[a, 0]
            B[4, 3 * a], B[4, 3 * a + 2] = dN[a, 2], dN[a, 0]
            B[5, 3 * a + 1], B[5, 3 * a + 2] = dN[a, 2], dN[a, 1]
        Bs.append(B)
    return Bs


def test_bmat_independent(oracle):
    # CONSTXG is a truncated 1/sqrt(3); the difference is far below 1e-12 relative
    Bs = independent_B(1 / 3, 1 / 4, 1 / 5)
    b = oracle.bmat(4, 5, 6)
    for gp in range(8):
        assert np.allclose(b[gp], Bs[gp], rtol=1e-12, atol=1e-12)


def test_dense_stiffness_independent(oracle):
    """Global stiffness from an independent numpy assembly vs the oracle's ELL."""
    size = (4, 5, 4)
    nx, ny, nz = size
    orve = oracle.rve(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS["elastic"])
    et, _ = oracle.elem_types(MIC_SPHERE, (0.3, 0, 0, 0), *size)
    vals = orve.assembly_mat(np.zeros(nx * ny * nz * 3))
    cols = oracle.ell_cols3d(*size)
    N = nx * ny * nz * 3
    A = np.zeros((N, N))
    for r in range(N):
        for s in range(81):
            A[r, cols[r, s]] += vals[r, s]
    dx, dy, dz = 1 / (nx - 1), 1 / (ny - 1), 1 / (nz - 1)
    Bs = independent_B(dx, dy, dz)
    wg = dx * dy * dz / 8
    K = np.zeros((N, N))
    Cs = [c_iso(1e6, 0.3), c_iso(1e7, 0.25)]
    pos = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    for ez in range(nz - 1):
        for ey in range(ny - 1):
            for ex in range(nx - 1):
                e = (ez * (ny - 1) + ey) * (nx - 1) + ex
                Ke = sum(B.T @ Cs[et[e]] @ B for B in Bs) * wg
                dofs = np.array([((ez + c) * ny * nx + (ey + b) * nx + ex + a) * 3 + d for a, b, c in pos
                                 for d in range(3)])
                K[np.ix_(dofs, dofs)] += Ke
    bc = np.zeros(N, bool)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
                    bc[((k * ny + j) * nx + i) * 3:((k * ny + j) * nx + i) * 3 + 3] = True
    K[bc, :] = 0
    K[bc, bc] = 1
    assert np.allclose(A, K, rtol=1e-11, atol=1e-11 * np.abs(K).max())
    # interior block symmetric; BC rows identity, BC columns kept (reference quirk)
    Ii = ~bc
    assert np.allclose(A[np.ix_(Ii, Ii)], A[np.ix_(Ii, Ii)].T, rtol=0, atol=1e-9 * np.abs(A).max())
    # linear rhs is -K u on the interior
    u = np.random.default_rng(6).normal(0, 1e-3, N)
    b, _ = orve.assembly_rhs(u)
    ref_b = np.zeros(N)
    Ku = np.zeros(N)  # unconstrained K u
    for ez in range(nz - 1):
        for ey in range(ny - 1):
            for ex in range(nx - 1):
                e = (ez * (ny - 1) + ey) * (nx - 1) + ex
                dofs = np.array([((ez + c) * ny * nx + (ey + b_) * nx + ex + a) * 3 + d for a, b_, c in pos
                                 for d in range(3)])
                Ke = sum(B.T @ Cs[et[e]] @ B for B in Bs) * wg
                Ku[dofs] += Ke @ u[dofs]
    ref_b[~bc] = -Ku[~bc]
    assert np.allclose(b, ref_b, rtol=1e-10, atol=1e-10 * np.abs(ref_b).max())


def test_cg_matches_dense_solve(oracle):
    size = (4, 4, 5)
    orve = oracle.rve(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS["elastic"])
    N = 4 * 4 * 5 * 3
    u = np.zeros(N)
    orve.set_displ_bc([2e-3, 0, -1e-3, 0, 1e-3, 0], u)
    vals = orve.assembly_mat(u)
    b, _ = orve.assembly_rhs(u)
    cols = oracle.ell_cols3d(*size)
    A = np.zeros((N, N))
    for r in range(N):
        for s in range(81):
            A[r, cols[r, s]] += vals[r, s]
    x_dense = np.linalg.solve(A, b)
    x, its, _ = oracle.ell_cg(N // 3, 3, 81, 13, cols, vals, b, 1e-50, 1e-13, 2000)
    assert np.allclose(x, x_dense, rtol=1e-9, atol=1e-9 * np.abs(x_dense).max())


def test_patch_homogeneous(oracle):
    """Affine u is in the trilinear FE space: ctan = C and sigma = C eps."""
    mats = [material(0, 1e6, 0.3)] * 3
    p, k = make_params(5, MIC_HOMOGENEOUS, mats=mats, cg_rel_tol=1e-12, lin_stress=0)
    m = oracle.micro(p, k)
    C = c_iso(1e6, 0.3)
    assert np.allclose(m.ctan_lin(), C, rtol=1e-7, atol=1e-7 * C.max())
    eps = np.array([1e-3, -2e-3, 5e-4, 1e-3, -1e-3, 2e-3])
    m.set_strain(0, eps)
    m.homogenize()
    assert np.allclose(m.get_stress(0), C @ eps, rtol=1e-7, atol=1e-7 * np.abs(C @ eps).max())


def test_hill_bounds_and_symmetry(oracle):
    p, k = make_params(9, MIC_SPHERE, (0.3, 0, 0, 0), ELAST2, cg_rel_tol=1e-10)
    m = oracle.micro(p, k)
    ct = m.ctan_lin()
    Vm, Vf = m.volume_fractions()
    assert abs(Vm + Vf - 1) < 1e-12
    Cm, Cf = c_iso(1e6, 0.3), c_iso(1e7, 0.25)
    CV = Vm * Cm + Vf * Cf
    CR = np.linalg.inv(Vm * np.linalg.inv(Cm) + Vf * np.linalg.inv(Cf))
    assert np.allclose(ct, ct.T, rtol=1e-6, atol=1e-6 * ct.max())
    sym = 0.5 * (ct + ct.T)
    assert np.linalg.eigvalsh(CV - sym).min() > -1e-6 * ct.max()
    assert np.linalg.eigvalsh(sym - CR).min() > -1e-6 * ct.max()


def _splitmix(seed):
    s = seed

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return (z ^ (z >> 31)) >> 11

    return lambda: nxt() * 2.0 ** -53


def test_random_spheres_independent(oracle):
    geo = (12345.0, 8.0, 0.05, 0.15)
    sph = oracle.random_spheres(geo)
    u = _splitmix(12345)
    want = []
    for _ in range(8):
        c = [u(), u(), u()]
        want.append(c + [0.05 + (0.15 - 0.05) * u()])
    assert np.array_equal(sph, np.array(want))
    for n in (7, 12):
        et, rc = oracle.elem_types(MIC3D_RANDOM_SPHERES, geo, n)
        h = 1.0 / (n - 1)
        idx = np.arange(n - 1)
        cz, cy, cx = np.meshgrid(idx * h + h / 2, idx * h + h / 2, idx * h + h / 2, indexing="ij")
        inside = np.zeros(cx.shape, bool)
        for x0, y0, z0, r in want:
            d = np.sqrt(((cx - x0) * (cx - x0) + (cy - y0) * (cy - y0)) + (cz - z0) * (cz - z0))
            inside |= d < r
        assert rc == 0 and np.array_equal(et, inside.ravel().astype(np.int32))


def test_invalid_micro_type(oracle):
    et, rc = oracle.elem_types(99, (0.1,) * 4, 4)
    assert rc == -1


def test_cylinder_axis_quirk(oracle):
    """A centroid exactly on the cylinder axis point gives 0/0 = NaN, classified as matrix."""
    et, _ = oracle.elem_types(MIC_CILI_FIB_Z, (0.2,) * 4, 4)
    et = et.reshape(3, 3, 3)
    assert et[1, 1, 1] == 0 and et[0, 1, 1] == 1


def test_layer_and_mic8_volume(oracle):
    et, _ = oracle.elem_types(MIC_LAYER_Y, (0.3,) * 4, 11)
    assert abs(et.mean() - 0.3) < 1e-12
    et, _ = oracle.elem_types(MIC3D_8, (0.1,) * 4, 21)
    assert set(np.unique(et)) == {0, 1, 2}


@pytest.mark.parametrize("kind", ["elastic", "plastic", "damage"])
def test_vtu_bytewise(oracle, ref, kind, tmp_path):
    """calc_fields + write_vtu (src/average.cpp:84-112, src/output.cpp:67-214): the oracle's file equals the
    compiled reference's byte for byte, with and without history, on a non-cubic grid."""
    rng = np.random.default_rng(5)
    size = (5, 4, 6)
    p, k = make_params(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind], calc_ctan_lin=0)
    rm = ref.micro(p, k)
    orve = oracle.rve(size, MIC_SPHERE, (0.3, 0, 0, 0), MATSETS[kind])
    u = rng.normal(0, 1e-3, rm.nndim)
    vars_ = _rand_state(rng, rm.nvars, kind) if kind != "elastic" else None
    for i, v in enumerate((None, vars_)):
        fo = orve.write_vtu(u, v, tmp_path / f"orc{i}")
        fr = rm.write_vtu(u, v, str(tmp_path / f"ref{i}"))
        a, b = open(fo, "rb").read(), open(fr, "rb").read()
        assert a == b, f"first difference at byte {next(j for j in range(min(len(a), len(b))) if a[j] != b[j])}"
        # the full-precision fields: element volumes are equal, so their mean stress is calc_ave_stress
        es, ss = orve.calc_fields(u, v)
        sig = orve.ave_stress(u, v)
        assert np.allclose(ss.mean(axis=0), sig, rtol=1e-12, atol=1e-12 * np.abs(sig).max())


# ---------------------------------------------------------------------------
# restart file and profiling log: fixtures written by the compiled reference
# (tests/golden/make_restart_log.py)
# ---------------------------------------------------------------------------
def restart_fixture():
    import json
    import os

    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(d, "restart_log_ref.json")) as f:
        doc = json.load(f)
    with open(os.path.join(d, doc["restart_file"]), "rb") as f:
        raw = f.read()
    return doc, raw


def parse_restart(raw, ngp, nvars, nndim):
    """The reference restart layout (include/gp.hpp:107-124): per GP a 1-byte bool, then vars_n and u_n
    (doubles, native order) when it is set.  Returns [(allocated, vars_n, u_n)]."""
    out, o = [], 0
    for _ in range(ngp):
        a = raw[o] != 0
        o += 1
        if a:
            v = np.frombuffer(raw, np.float64, nvars, o)
            o += 8 * nvars
            u = np.frombuffer(raw, np.float64, nndim, o)
            o += 8 * nndim
            out.append((True, v, u))
        else:
            out.append((False, None, None))
    assert o == len(raw)
    return out


def restart_params(doc):
    from oracle import bindings as ob

    return ob.make_params(doc["size"], doc["micro_type"], doc["geo"], [ob.material(*m) for m in doc["materials"]],
                          ngp=len(doc["coupling"]), coupling=doc["coupling"])


def test_oracle_reproduces_reference_log_and_restart(oracle):
    """The oracle over the fixture's three steps: sigma, ctan, cost (the FE_FULL stale addition included,
    src/homogenize.cpp:267-269) and flags equal the compiled reference's bit for bit, and its state after the
    second step equals the restart file the reference wrote then, byte for byte."""
    doc, raw = restart_fixture()
    p, k = restart_params(doc)
    m = oracle.micro(p, k)
    ngp = len(doc["coupling"])
    for t, rec in enumerate(doc["steps"]):
        for g in range(ngp):
            m.set_strain(g, doc["strains"][t][g])
        m.homogenize()
        for g in range(ngp):
            r = rec[g]
            assert m.get_stress(g).tolist() == r["sigma"], (t, g)
            assert m.get_ctan(g).ravel().tolist() == r["ctan"], (t, g)
            assert (m.get_cost(g), m.is_non_linear(g), int(m.has_converged(g))) == (r["cost"], r["non_linear"],
                                                                                     r["converged"]), (t, g)
        m.update_vars()
        if t == doc["restart_after_step"]:
            for g, (a, v, u) in enumerate(parse_restart(raw, ngp, m.nvars, m.nndim)):
                assert a == bool(m.is_non_linear(g))
                if a:
                    assert np.array_equal(m.state(g, 2), v) and np.array_equal(m.state(g, 0), u)
    # the log the reference wrote: a header, then per homogenize "log_id : i" and one line per GP
    lines = doc["log"].splitlines()
    assert lines[0] == "#<gp_id>  <non-linear>  <cost>  <converged>"
    assert len(lines) == 1 + len(doc["steps"]) * (1 + ngp)
    for t, rec in enumerate(doc["steps"]):
        blk = lines[1 + t * (1 + ngp):1 + (t + 1) * (1 + ngp)]
        assert blk[0] == f"log_id : {t}"
        for g in range(ngp):
            assert blk[1 + g] == f"\t{g}\t{rec[g]['non_linear']}\t{rec[g]['cost']}\t{rec[g]['converged']}"


# ---------------------------------------------------------------------------
# conditioning of a damaged RVE's Newton path (the FP64 control for the
# mixed-precision damage bars of the GPU tests)
# ---------------------------------------------------------------------------
def test_capped_damage_newton_crosses_a_singular_tangent(oracle):
    """The damage point of the mixed-precision Newton GPU test (9x10x11 nodes, MIC_SPHERE 0.3, damage matrix
    Xt = 1e4, FE_FULL GP 0 at its second load step, CG 1e-12, Newton 6 / 1e-8), restated on the oracle: the FP64
    Newton oscillates and stops at its cap unconverged, and one of its tangents is numerically singular on the
    interior DOFs (smallest / largest singular value of D^-1 A below 1e-12).  There a linear solution meeting
    the CG tolerance is not unique: adding the tangent's near-null vector, scaled so that the residual
    ||D^-1 (b - A du)|| stays within 20x the tolerance, changes the returned sigma by more than 1e-3.  Any two
    correct linear solvers (the FP64 CG, a mixed-precision defect correction) may therefore return sigma that
    differ far beyond 1e-6 at such a point; the bar 1e-6 applies only where the Newton converges."""
    import scipy.sparse as sp

    law = [(2, 1e6, 0.3, 0, 0, 1e4), (0, 1e7, 0.25), (0, 1e7, 0.25)]
    size = (9, 10, 11)
    mats = [material(*s) for s in law]
    kw = dict(lin_stress=0, nr_max_its=6, nr_rel_tol=1e-8, cg_rel_tol=1e-12, cg_max_its=4000)
    p, k = make_params(size, MIC_SPHERE, (0.3, 0, 0, 0), mats, ngp=2, coupling=[FE_FULL, FE_ONE_WAY], **kw)
    m = oracle.micro(p, k)
    rng = np.random.default_rng(3)
    d = [rng.uniform(-1, 1, 6) * 0.015, rng.uniform(-1, 1, 6) * 0.015 * 0.7]
    for g in range(2):
        m.set_strain(g, d[g] / 3)
    m.homogenize()
    m.update_vars()
    u_n, vars_n = m.state(0, 0), m.state(0, 2)
    m.close()
    from oracle.bindings import OracleRVE

    r = OracleRVE(oracle, size, MIC_SPHERE, (0.3, 0, 0, 0), mats)
    nx, ny, nz = size
    cols = np.zeros(r.nndim * 81, dtype=np.int32)
    oracle.lib.orc_ell_cols3d(nx, ny, nz, cols)
    cols = cols.reshape(-1, 81)
    tol = kw["cg_rel_tol"]

    def newton(perturb_at=None):
        eps = d[0] * 2 / 3
        u = u_n.copy()
        r.set_displ_bc(eps, u)
        b, norm = r.assembly_rhs(u, vars_n)
        norm0, conds, resid = norm, [], None
        for it in range(kw["nr_max_its"]):
            if norm < 1e-10 or norm < norm0 * kw["nr_rel_tol"]:
                return r.ave_stress(u, vars_n), conds, True, resid
            vals = r.assembly_mat(u, vars_n)
            A = sp.csr_matrix((vals.ravel().copy(), cols.ravel().astype(np.int64),
                               np.arange(0, r.nndim * 81 + 1, 81)), shape=(r.nndim, r.nndim))
            A.sum_duplicates()
            x, _, _ = oracle.ell_cg(r.nndim // 3, 3, 81, 13, cols, vals, b, 1e-50, tol, kw["cg_max_its"])
            kinv = 1.0 / A.diagonal()
            diag = A.diagonal()
            off = np.abs(A).sum(axis=1).A1 - np.abs(diag)
            inner = np.where(~((off == 0) & (diag == 1.0)))[0]  # Dirichlet rows are unit rows
            M = (sp.diags(kinv) @ A).tocsr()[inner][:, inner].toarray()
            _, S, Vt = np.linalg.svd(M)
            conds.append(S[-1] / S[0])
            if perturb_at == it:
                v = np.zeros(r.nndim)
                v[inner] = Vt[-1]
                bn = np.linalg.norm(kinv * b)
                # the largest multiple of v whose residual stays within 20x the tolerance
                c = 1.0
                while np.linalg.norm(kinv * (b - A @ (x + c * v))) > 20 * tol * bn:
                    c *= 0.5
                x = x + c * v
                resid = np.linalg.norm(kinv * (b - A @ x)) / bn
            u = u + x
            b, norm = r.assembly_rhs(u, vars_n)
        return r.ave_stress(u, vars_n), conds, False, resid

    s0, conds, converged, _ = newton()
    worst = int(np.argmin(conds))
    s1, _, _, resid = newton(worst)
    spread = np.abs(s1 - s0).max() / np.abs(s0).max()
    print(f"singular value ratios per Newton iteration {['%.1e' % c for c in conds]}; perturbed at {worst} "
          f"(residual {resid:.1e}): sigma spread {spread:.2e}", flush=True)
    assert not converged
    assert conds[worst] < 1e-12
    assert resid < 20 * tol and spread > 1e-3
