"""ctypes bindings for the CPU oracle and the compiled reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

``Oracle`` wraps oracle/liboracle.so (the C restatement, micropp_oracle.c).
``Ref`` wraps oracle/_ref/libmicropp_ref.so (Micropp's own sources compiled
by oracle/Makefile with ref_shim.cpp); it exists only where the Micropp tree
was present at build time.  Both expose the same methods, so a test can run
one case through each and compare.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmicropp_ref.so")
REF_SRC = os.environ.get("MICROPP_REF", "/root/reference")

MIC_HOMOGENEOUS, MIC_SPHERE, MIC_LAYER_Y, MIC_CILI_FIB_X, MIC_CILI_FIB_Z, MIC_CILI_FIB_XZ = range(6)
MIC_QUAD_FIB_XYZ, MIC_QUAD_FIB_XZ, MIC_QUAD_FIB_XZ_BROKEN_X, MIC3D_SPHERES, MIC3D_8 = range(6, 11)
MIC3D_FIBS_20_ORDER, MIC3D_FIBS_20_DISORDER, MIC3D_RANDOM_SPHERES = 11, 12, 13
FE_LINEAR, FE_ONE_WAY, FE_FULL, MIX_RULE_CHAMIS = range(4)
NVAR = 7


class Material(C.Structure):
    """Layout of the reference ``struct material_base`` (include/material_base.h:40-45)."""

    _fields_ = [(n, C.c_double) for n in ("E", "nu", "Ka", "Sy", "k", "mu", "lambda_", "Xt")] + [
        ("type", C.c_int)
    ]


class Params(C.Structure):
    _fields_ = [
        ("ngp", C.c_int),
        ("size", C.c_int * 3),
        ("type", C.c_int),
        ("geo", C.c_double * 4),
        ("mat", Material * 3),
        ("coupling", C.POINTER(C.c_int)),
        ("subiterations", C.c_int),
        ("nsubiterations", C.c_int),
        ("nr_max_its", C.c_int),
        ("nr_max_tol", C.c_double),
        ("nr_rel_tol", C.c_double),
        ("cg_max_its", C.c_int),
        ("cg_abs_tol", C.c_double),
        ("cg_rel_tol", C.c_double),
        ("calc_ctan_lin", C.c_int),
        ("use_A0", C.c_int),
        ("its_with_A0", C.c_int),
        ("lin_stress", C.c_int),
    ]


def material(type_, E, nu, Ka=0.0, Sy=0.0, Xt=0.0):
    """``material_set`` (src/material.c:26-38) with the reference expressions."""
    m = Material()
    m.E, m.nu, m.Ka, m.Sy, m.Xt, m.type = E, nu, Ka, Sy, Xt, type_
    m.k = E / (3.0 * (1.0 - 2.0 * nu))
    m.mu = E / (2.0 * (1.0 + nu))
    m.lambda_ = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return m


def ensure_built(ref: bool = True, force: bool = False) -> None:
    """Build the oracle (and the reference, when its source tree is here); force: recompile everything."""
    targets = ["oracle"]
    if ref and os.path.isdir(os.path.join(REF_SRC, "src")):
        targets.append("ref")
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", HERE, f"REF={REF_SRC}"] + targets,
                   check=True)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_VP = C.c_void_p


def _dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class _Lib:
    prefix = ""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        p = self.prefix
        L = self.lib
        PD = C.POINTER(C.c_double)
        sig = {
            "mat_stress": (None, [C.POINTER(Material), _D, PD, _D]),
            "mat_ctan": (None, [C.POINTER(Material), _D, PD, _D]),
            "mat_evolute": (C.c_int, [C.POINTER(Material), _D, PD, PD]),
            "ell_cols3d": (None, [C.c_int, C.c_int, C.c_int, _I]),
            "ell_mvp": (None, [C.c_int, C.c_int, _I, _D, _D, _D]),
            "ell_cg": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _I, _D, _D, _D, C.c_double, C.c_double,
                                 C.c_int, PD]),
            "micro_new": (_VP, [C.POINTER(Params)]),
            "micro_free": (None, [_VP]),
            "set_strain": (None, [_VP, C.c_int, _D]),
            "get_stress": (None, [_VP, C.c_int, _D]),
            "get_ctan": (None, [_VP, C.c_int, _D]),
            "homogenize": (None, [_VP]),
            "homogenize_linear": (None, [_VP]),
            "update_vars": (None, [_VP]),
            "is_non_linear": (C.c_int, [_VP, C.c_int]),
            "get_non_linear_gps": (C.c_int, [_VP]),
            "get_cost": (C.c_long, [_VP, C.c_int]),
            "has_converged": (C.c_int, [_VP, C.c_int]),
            "has_subiterated": (C.c_int, [_VP, C.c_int]),
            "get_ctan_lin": (None, [_VP, _D]),
            "get_volume_fractions": (None, [_VP, _D]),
            "get_state": (C.c_int, [_VP, C.c_int, C.c_int, _D]),
            "nthreads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = args

    def f(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- materials ----
    def mat_stress(self, m, eps, hist=None):
        s = np.zeros(6)
        self.f("mat_stress")(C.byref(m), np.ascontiguousarray(eps, float), _dptr(hist), s)
        return s

    def mat_ctan(self, m, eps, hist=None):
        c = np.zeros(36)
        self.f("mat_ctan")(C.byref(m), np.ascontiguousarray(eps, float), _dptr(hist), c)
        return c.reshape(6, 6)

    def mat_evolute(self, m, eps, hist_old=None, hist_new=None):
        return self.f("mat_evolute")(C.byref(m), np.ascontiguousarray(eps, float), _dptr(hist_old),
                                     _dptr(hist_new))

    # ---- ELL ----
    def ell_cols3d(self, nx, ny, nz):
        c = np.zeros(nx * ny * nz * 3 * 81, dtype=np.int32)
        self.f("ell_cols3d")(nx, ny, nz, c)
        return c.reshape(-1, 81)

    def ell_mvp(self, cols, vals, x):
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, float)
        nrow, nnz = vals.shape
        y = np.zeros(nrow)
        self.f("ell_mvp")(nrow, nnz, cols, vals, np.ascontiguousarray(x, float), y)
        return y

    def ell_cg(self, nn, nfield, nnz, shift, cols, vals, b, min_err=1e-50, rel_err=1e-5, max_its=1000):
        x = np.zeros(nn * nfield)
        err = C.c_double(0)
        its = self.f("ell_cg")(nn, nfield, nnz, shift, np.ascontiguousarray(cols, np.int32),
                               np.ascontiguousarray(vals, float), np.ascontiguousarray(b, float), x, min_err,
                               rel_err, max_its, C.byref(err))
        return x, its, err.value


class Micro:
    """One micropp<3> instance of either library (same call surface)."""

    def __init__(self, lib: _Lib, params: Params, keep=None):
        self.lib = lib
        self._keep = keep
        self.params = params
        self.h = lib.f("micro_new")(C.byref(params))
        if not self.h:
            raise ValueError("micro_new failed")
        self.ngp = params.ngp
        nx, ny, nz = params.size
        self.nndim = nx * ny * nz * 3
        self.nvars = (nx - 1) * (ny - 1) * (nz - 1) * 8 * NVAR

    def close(self):
        if self.h:
            self.lib.f("micro_free")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_strain(self, gp, eps):
        self.lib.f("set_strain")(self.h, gp, np.ascontiguousarray(eps, float))

    def get_stress(self, gp):
        s = np.zeros(6)
        self.lib.f("get_stress")(self.h, gp, s)
        return s

    def get_ctan(self, gp):
        c = np.zeros(36)
        self.lib.f("get_ctan")(self.h, gp, c)
        return c.reshape(6, 6)

    def homogenize(self):
        self.lib.f("homogenize")(self.h)

    def homogenize_linear(self):
        self.lib.f("homogenize_linear")(self.h)

    def update_vars(self):
        self.lib.f("update_vars")(self.h)

    def is_non_linear(self, gp):
        return self.lib.f("is_non_linear")(self.h, gp)

    def get_non_linear_gps(self):
        return self.lib.f("get_non_linear_gps")(self.h)

    def get_cost(self, gp):
        return self.lib.f("get_cost")(self.h, gp)

    def has_converged(self, gp):
        return bool(self.lib.f("has_converged")(self.h, gp))

    def has_subiterated(self, gp):
        return bool(self.lib.f("has_subiterated")(self.h, gp))

    def ctan_lin(self):
        c = np.zeros(36)
        self.lib.f("get_ctan_lin")(self.h, c)
        return c.reshape(6, 6)

    def volume_fractions(self):
        v = np.zeros(2)
        self.lib.f("get_volume_fractions")(self.h, v)
        return v

    def state(self, gp, which):
        n = self.nndim if which < 2 else self.nvars
        out = np.zeros(n)
        ok = self.lib.f("get_state")(self.h, gp, which, out)
        return out if ok else None


def make_params(size, micro_type=MIC_HOMOGENEOUS, geo=(0.1, 0.1, 0.1, 0.1), mats=None, ngp=1, coupling=None,
                **kw):
    p = Params()
    p.ngp = ngp
    s = (size, size, size) if np.isscalar(size) else size
    p.size[:] = list(s)
    p.type = micro_type
    p.geo[:] = list(geo)
    mats = mats or [material(0, 1e6, 0.3)] * 3
    for i in range(3):
        p.mat[i] = mats[i]
    keep = None
    if coupling is not None:
        keep = (C.c_int * ngp)(*coupling)
        p.coupling = C.cast(keep, C.POINTER(C.c_int))
    defaults = dict(subiterations=0, nsubiterations=10, nr_max_its=4, nr_max_tol=1e-10, nr_rel_tol=1e-3,
                    cg_max_its=1000, cg_abs_tol=1e-50, cg_rel_tol=1e-5, calc_ctan_lin=1, use_A0=0,
                    its_with_A0=1, lin_stress=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p, keep


class Oracle(_Lib):
    prefix = "orc_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        PD = C.POINTER(C.c_double)
        L.orc_bmat.argtypes = [C.c_int, C.c_int, C.c_int, _D]
        L.orc_elem_types.argtypes = [C.c_int, _D, C.c_int, C.c_int, C.c_int, _I]
        L.orc_elem_types.restype = C.c_int
        L.orc_random_spheres.argtypes = [_D, _D, C.c_int]
        L.orc_random_spheres.restype = C.c_int
        L.orc_ell_cols2d.argtypes = [C.c_int, C.c_int, _I]
        L.orc_elem_nodes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _I]
        L.orc_ell_add2d_all.argtypes = [C.c_int, C.c_int, _D, _D]
        L.orc_ell_bc2d.argtypes = [C.c_int, C.c_int, _D]
        L.orc_rve_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _D, C.POINTER(Material)]
        L.orc_rve_new.restype = _VP
        L.orc_rve_free.argtypes = [_VP]
        L.orc_set_displ_bc.argtypes = [_VP, _D, _D]
        L.orc_assembly_rhs.argtypes = [_VP, _D, PD, _D]
        L.orc_assembly_rhs.restype = C.c_double
        L.orc_assembly_mat.argtypes = [_VP, _D, PD, _D]
        L.orc_ave_stress.argtypes = [_VP, _D, PD, _D]
        L.orc_vars_new.argtypes = [_VP, _D, PD, _D]
        L.orc_vars_new.restype = C.c_int
        L.orc_calc_fields.argtypes = [_VP, _D, PD, _D, _D]
        L.orc_write_vtu.argtypes = [_VP, _D, PD, C.c_char_p]
        L.orc_write_vtu.restype = C.c_int
        L.orc_newton.argtypes = [_VP, _D, PD, _D, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                 C.c_double, PD, C.c_int, _I]

    def bmat(self, nx, ny=None, nz=None):
        ny = ny or nx
        nz = nz or nx
        b = np.zeros(8 * 6 * 24)
        self.lib.orc_bmat(nx, ny, nz, b)
        return b.reshape(8, 6, 24)

    def elem_types(self, micro_type, geo, nx, ny=None, nz=None):
        ny = ny or nx
        nz = nz or nx
        et = np.zeros((nx - 1) * (ny - 1) * (nz - 1), dtype=np.int32)
        rc = self.lib.orc_elem_types(micro_type, np.asarray(geo, float), nx, ny, nz, et)
        return et, rc

    def elem_nodes(self, nx, ny, ex, ey, ez=0):
        n = np.zeros(8, dtype=np.int32)
        self.lib.orc_elem_nodes(nx, ny, ex, ey, ez, n)
        return n

    def random_spheres(self, geo, max_spheres=64):
        out = np.zeros(max_spheres * 4)
        n = self.lib.orc_random_spheres(np.asarray(geo, float), out, max_spheres)
        return out[: n * 4].reshape(n, 4)

    def ell2d_build(self, nx, ny, Ae):
        cols = np.zeros(nx * ny * 9, dtype=np.int32)
        vals = np.zeros(nx * ny * 9)
        self.lib.orc_ell_cols2d(nx, ny, cols)
        self.lib.orc_ell_add2d_all(nx, ny, np.asarray(Ae, float), vals)
        self.lib.orc_ell_bc2d(nx, ny, vals)
        return cols.reshape(-1, 9), vals.reshape(-1, 9)

    def rve(self, size, micro_type, geo, mats):
        return OracleRVE(self, size, micro_type, geo, mats)

    def micro(self, params, keep=None):
        return Micro(self, params, keep)


class OracleRVE:
    def __init__(self, o: Oracle, size, micro_type, geo, mats):
        self.o = o
        s = (size, size, size) if np.isscalar(size) else tuple(size)
        self.size = s
        arr = (Material * 3)(*mats)
        self.h = o.lib.orc_rve_new(s[0], s[1], s[2], micro_type, np.asarray(geo, float), arr)
        if not self.h:
            raise ValueError("invalid micro type")
        self.nndim = s[0] * s[1] * s[2] * 3
        self.nelem = (s[0] - 1) * (s[1] - 1) * (s[2] - 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.orc_rve_free(self.h)
            self.h = None

    def set_displ_bc(self, eps, u):
        self.o.lib.orc_set_displ_bc(self.h, np.asarray(eps, float), u)

    def assembly_rhs(self, u, vars_=None):
        b = np.zeros(self.nndim)
        n = self.o.lib.orc_assembly_rhs(self.h, np.ascontiguousarray(u), _dptr(vars_), b)
        return b, n

    def assembly_mat(self, u, vars_=None):
        v = np.zeros(self.nndim * 81)
        self.o.lib.orc_assembly_mat(self.h, np.ascontiguousarray(u), _dptr(vars_), v)
        return v.reshape(-1, 81)

    def ave_stress(self, u, vars_=None):
        s = np.zeros(6)
        self.o.lib.orc_ave_stress(self.h, np.ascontiguousarray(u), _dptr(vars_), s)
        return s

    def vars_new(self, u, vars_old=None):
        vn = np.zeros(self.nelem * 8 * NVAR)
        nl = self.o.lib.orc_vars_new(self.h, np.ascontiguousarray(u), _dptr(vars_old), vn)
        return vn, nl

    def calc_fields(self, u, vars_=None):
        es, ss = np.zeros(self.nelem * 6), np.zeros(self.nelem * 6)
        self.o.lib.orc_calc_fields(self.h, np.ascontiguousarray(u), _dptr(vars_), es, ss)
        return es.reshape(-1, 6), ss.reshape(-1, 6)

    def write_vtu(self, u, vars_, filename):
        """Writes <filename>.vtu; returns its path."""
        rc = self.o.lib.orc_write_vtu(self.h, np.ascontiguousarray(u), _dptr(vars_), str(filename).encode())
        if rc:
            raise OSError(f"cannot write {filename}.vtu")
        return f"{filename}.vtu"

    def newton(self, eps, u, vars_old=None, nr_max_its=4, nr_max_tol=1e-10, nr_rel_tol=1e-3, cg_max_its=1000,
               cg_abs_tol=1e-50, cg_rel_tol=1e-5, A0=None, its_with_A0=1):
        out = np.zeros(3, dtype=np.int32)
        self.o.lib.orc_newton(self.h, np.asarray(eps, float), _dptr(vars_old), u, nr_max_its, nr_max_tol,
                              nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol, _dptr(A0), its_with_A0, out)
        return tuple(int(v) for v in out)


class Ref(_Lib):
    prefix = "ref_"

    def __init__(self, path=REF_SO):
        super().__init__(path)
        L = self.lib
        PD = C.POINTER(C.c_double)
        L.ref_ell2d_build.argtypes = [C.c_int, C.c_int, _D, _I, _D]
        for n, res, args in [
            ("nrow", C.c_int, [_VP]),
            ("nelem", C.c_int, [_VP]),
            ("bmat", None, [_VP, _D]),
            ("elem_types", None, [_VP, _I]),
            ("set_displ_bc", None, [_VP, _D, _D]),
            ("assembly_rhs", C.c_double, [_VP, _D, PD, _D]),
            ("assembly_mat", None, [_VP, _D, PD, _D]),
            ("ave_stress", None, [_VP, _D, PD, _D]),
            ("vars_new", C.c_int, [_VP, _D, PD, _D]),
            ("newton", None, [_VP, _D, PD, _D, C.c_double, C.c_double, C.c_int, _I]),
            ("write_vtu", C.c_int, [_VP, _D, PD, C.c_char_p]),
            ("micro_new_rank", _VP, [C.POINTER(Params), C.c_int, C.c_int]),
            ("write_restart", None, [_VP, C.c_int]),
            ("read_restart", None, [_VP, C.c_int]),
        ]:
            f = getattr(L, "ref_" + n)
            f.restype = res
            f.argtypes = args

    def ell2d_build(self, nx, ny, Ae):
        cols = np.zeros(nx * ny * 9, dtype=np.int32)
        vals = np.zeros(nx * ny * 9)
        self.lib.ref_ell2d_build(nx, ny, np.asarray(Ae, float), cols, vals)
        return cols.reshape(-1, 9), vals.reshape(-1, 9)

    def micro(self, params, keep=None):
        return RefMicro(self, params, keep)

    def micro_rank(self, params, keep=None, rank=0, write_log=0):
        """An instance of MPI rank `rank` (the restart / log file names), writing the profiling log
        micropp-profiling-<rank>.log into the current directory when write_log is set."""
        return RefMicro(self, params, keep, rank=rank, write_log=write_log)


class RefMicro(Micro):
    """A reference micropp<3> instance, plus the protected members the shim exposes."""

    def __init__(self, lib, params, keep=None, rank=None, write_log=0):
        if rank is None:
            super().__init__(lib, params, keep)
            return
        self.lib, self._keep, self.params = lib, keep, params
        self.h = lib.lib.ref_micro_new_rank(C.byref(params), int(rank), int(write_log))
        if not self.h:
            raise ValueError("micro_new_rank failed")
        self.ngp = params.ngp
        nx, ny, nz = params.size
        self.nndim = nx * ny * nz * 3
        self.nvars = (nx - 1) * (ny - 1) * (nz - 1) * 8 * NVAR

    def write_restart(self, restart_id):
        """micropp-restart-<rank>-<id>.bin in the current directory."""
        self.lib.lib.ref_write_restart(self.h, int(restart_id))

    def read_restart(self, restart_id):
        self.lib.lib.ref_read_restart(self.h, int(restart_id))

    def bmat(self):
        b = np.zeros(8 * 6 * 24)
        self.lib.lib.ref_bmat(self.h, b)
        return b.reshape(8, 6, 24)

    def elem_types(self):
        et = np.zeros(self.lib.lib.ref_nelem(self.h), dtype=np.int32)
        self.lib.lib.ref_elem_types(self.h, et)
        return et

    def set_displ_bc(self, eps, u):
        self.lib.lib.ref_set_displ_bc(self.h, np.asarray(eps, float), u)

    def assembly_rhs(self, u, vars_=None):
        b = np.zeros(self.nndim)
        n = self.lib.lib.ref_assembly_rhs(self.h, np.ascontiguousarray(u), _dptr(vars_), b)
        return b, n

    def assembly_mat(self, u, vars_=None):
        v = np.zeros(self.nndim * 81)
        self.lib.lib.ref_assembly_mat(self.h, np.ascontiguousarray(u), _dptr(vars_), v)
        return v.reshape(-1, 81)

    def ave_stress(self, u, vars_=None):
        s = np.zeros(6)
        self.lib.lib.ref_ave_stress(self.h, np.ascontiguousarray(u), _dptr(vars_), s)
        return s

    def vars_new(self, u, vars_old=None):
        vn = np.zeros(self.nvars)
        nl = self.lib.lib.ref_vars_new(self.h, np.ascontiguousarray(u), _dptr(vars_old), vn)
        return vn, nl

    def write_vtu(self, u, vars_, filename):
        """calc_fields + write_vtu of the reference; returns the file's path."""
        self.lib.lib.ref_write_vtu(self.h, np.ascontiguousarray(u), _dptr(vars_), str(filename).encode())
        return f"{filename}.vtu"

    def newton(self, eps, u, vars_old=None, cg_abs_tol=1e-50, cg_rel_tol=1e-5, cg_max_its=1000):
        out = np.zeros(3, dtype=np.int32)
        self.lib.lib.ref_newton(self.h, np.asarray(eps, float), _dptr(vars_old), u, cg_abs_tol, cg_rel_tol,
                                cg_max_its, out)
        return tuple(int(v) for v in out)
