"""micropp_amd -- MI355X (gfx950) implementation of Micropp's RVE homogenization hot path.

Python mirror of the C interface in include/micropp_c.h (the reference's
``micropp3_*`` API, src/micropp_c.cpp) and of the kernel-level ABI in
include/mppd.h, loaded with ctypes from the in-tree ``libmicropp_amd.so``.

The library is the product; this module only binds it.  If the shared
library is missing, importing the bindings raises: there is no fallback.
Build it with ``python -m micropp_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmicropp_amd.so")

# reference include/types.hpp:85-99 (+ the seeded sphere list), :115
(MIC_HOMOGENEOUS, MIC_SPHERE, MIC_LAYER_Y, MIC_CILI_FIB_X, MIC_CILI_FIB_Z, MIC_CILI_FIB_XZ, MIC_QUAD_FIB_XYZ,
 MIC_QUAD_FIB_XZ, MIC_QUAD_FIB_XZ_BROKEN_X, MIC3D_SPHERES, MIC3D_8, MIC3D_FIBS_20_ORDER, MIC3D_FIBS_20_DISORDER,
 MIC3D_RANDOM_SPHERES) = range(14)
FE_LINEAR, FE_ONE_WAY, FE_FULL, MIX_RULE_CHAMIS = range(4)
MATERIAL_ELASTIC, MATERIAL_PLASTIC, MATERIAL_DAMAGE = range(3)
NUM_VAR_GP = 7


class MaterialBase(C.Structure):
    """``struct material_base`` (include/material_base.h)."""

    _fields_ = [(n, C.c_double) for n in ("E", "nu", "Ka", "Sy", "k", "mu", "lambda_", "Xt")] + [
        ("type", C.c_int)
    ]


class Micropp3Params(C.Structure):
    """``struct micropp3_params`` (include/micropp_c.h)."""

    _fields_ = [
        ("ngp", C.c_int),
        ("size", C.c_int * 3),
        ("type", C.c_int),
        ("geo_params", C.c_double * 4),
        ("materials", MaterialBase * 3),
        ("coupling", C.POINTER(C.c_int)),
        ("subiterations", C.c_int),
        ("nsubiterations", C.c_int),
        ("mpi_rank", C.c_int),
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
        ("write_log", C.c_int),
        ("mixed_precision", C.c_int),
    ]


class Micropp3Handle(C.Structure):
    _fields_ = [("ptr", C.c_void_p)]


class NewtonParams(C.Structure):
    """``mppd_newton_params`` (include/mppd.h)."""

    _fields_ = [("nr_max_its", C.c_int), ("nr_max_tol", C.c_double), ("nr_rel_tol", C.c_double),
                ("cg_max_its", C.c_int), ("cg_abs_tol", C.c_double), ("cg_rel_tol", C.c_double),
                ("its_with_A0", C.c_int)]


# every symbol include/*.h declares, with (restype, argtypes)
_VP, _I, _D, _PD, _PI = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)
_H = C.POINTER(Micropp3Handle)
SIGNATURES = {
    # material_base.h
    "material_set": (None, [C.POINTER(MaterialBase), _I, _D, _D, _D, _D, _D]),
    "material_print": (None, [C.POINTER(MaterialBase)]),
    # micropp_c.h
    "micropp3_new": (None, [_H, _I, _PI, _I, _PD, C.POINTER(MaterialBase), _PI, _I, _I]),
    "micropp3_free": (None, [_H]),
    "micropp3_set_strain": (None, [_H, _I, _PD]),
    "micropp3_get_stress": (None, [_H, _I, _PD]),
    "micropp3_get_ctan": (None, [_H, _I, _PD]),
    "micropp3_homogenize": (None, [_H]),
    "micropp3_homogenize_linear": (None, [_H]),
    "micropp3_update_vars": (None, [_H]),
    "micropp3_is_non_linear": (C.c_bool, [_H, _I]),
    "micropp3_get_cost": (_I, [_H, _I]),
    "micropp3_has_converged": (C.c_bool, [_H, _I]),
    "micropp3_has_subiterated": (C.c_bool, [_H, _I]),
    "micropp3_output": (None, [_H, _I, C.c_char_p]),
    "micropp3_print_info": (None, [_H]),
    "micropp3_get_non_linear_gps": (_I, [_H]),
    "micropp3_write_restart": (None, [_H, _I]),
    "micropp3_read_restart": (None, [_H, _I]),
    "micropp3_output2": (None, [_H, _I, _I, _I]),
    "micropp3_params_default": (None, [C.POINTER(Micropp3Params)]),
    "micropp3_new_ext": (None, [_H, C.POINTER(Micropp3Params)]),
    "micropp3_get_state": (_I, [_H, _I, _I, _PD]),
    "micropp3_get_ctan_lin": (None, [_H, _PD]),
    "micropp3_get_volume_fractions": (None, [_H, _PD]),
    # mppd.h
    "mppd_last_error": (C.c_char_p, []),
    "mppd_device_count": (_I, [_PI]),
    "mppd_malloc": (_I, [C.POINTER(C.c_void_p), C.c_size_t]),
    "mppd_free": (_I, [_VP]),
    "mppd_memcpy_h2d": (_I, [_VP, _VP, C.c_size_t]),
    "mppd_memcpy_d2h": (_I, [_VP, _VP, C.c_size_t]),
    "mppd_memcpy_d2d": (_I, [_VP, _VP, C.c_size_t]),
    "mppd_memset": (_I, [_VP, _I, C.c_size_t]),
    "mppd_synchronize": (_I, []),
    "mppd_solver_copy": (_I, [_VP, _VP, _VP, C.c_size_t]),
    "mppd_solver_fill": (_I, [_VP, _VP, _I, C.c_size_t]),
    "mppd_mem_info": (_I, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mppd_rve_create": (_I, [C.POINTER(C.c_void_p), _PI, _I, _PD, C.POINTER(MaterialBase)]),
    "mppd_rve_destroy": (_I, [_VP]),
    "mppd_rve_info": (_I, [_VP, _PI, _PI, _PI]),
    "mppd_rve_bmat": (_I, [_VP, _PD]),
    "mppd_rve_elem_types": (_I, [_VP, _PI]),
    "mppd_rve_volume_fractions": (_I, [_VP, _PD]),
    "mppd_material_eval": (_I, [C.POINTER(MaterialBase), _I, _VP, _VP, _VP, _VP, _VP, _VP]),
    "mppd_solver_create": (_I, [C.POINTER(C.c_void_p), _VP]),
    "mppd_solver_destroy": (_I, [_VP]),
    "mppd_ell_download": (_I, [_VP, _PI, _PD]),
    "mppd_ell_upload_vals": (_I, [_VP, _PD]),
    "mppd_set_displ_bc": (_I, [_VP, _PD, _VP]),
    "mppd_assembly_rhs": (_I, [_VP, _VP, _VP, _VP, _PD]),
    "mppd_assembly_mat": (_I, [_VP, _VP, _VP]),
    "mppd_ell_mvp": (_I, [_VP, _VP, _VP]),
    "mppd_cg": (_I, [_VP, _VP, _VP, _VP, _D, _D, _I, _PI, _PD]),
    "mppd_cg_noalloc": (_I, [_VP, _VP, _VP, _VP, _D, _D, _I, _PI, _PD]),
    "mppd_ave_stress": (_I, [_VP, _VP, _VP, _PD]),
    "mppd_vars_new": (_I, [_VP, _VP, _VP, _VP, _PI]),
    "mppd_elem_fields": (_I, [_VP, _VP, _VP, _PD, _PD]),
    "mppd_newton": (_I, [_VP, _VP, _PD, _VP, _VP, C.POINTER(NewtonParams), _PI]),
    "mppd_cg_profile": (_I, [_VP, _VP, _I, _PD, _PD]),
    "mppd_msolver_create": (_I, [C.POINTER(C.c_void_p), _VP, _I]),
    "mppd_msolver_destroy": (_I, [_VP]),
    "mppd_msolver_solve": (_I, [_VP, _I, _PD, C.c_void_p, _PD, _PI, _PD]),
    "mppd_msolver_get_u": (_I, [_VP, _I, _PD]),
    "mppd_msolver_adjoint": (_I, [_VP, _I, _PD, C.c_void_p, _PD, _PD, _PD, _PI, _PD]),
    "mppd_msolver_adjoint_many": (_I, [_I, C.c_void_p, _I, _PD, C.c_void_p, _I, _PD, _PD, _PD, _PI, _PD]),
    "mppd_msolver_profile": (_I, [_VP, _I, _PD, _PD]),
    "mppd_msolver_coarse": (_I, [_VP, _PI, C.POINTER(C.c_float)]),
    "mppd_msolver_newton": (_I, [_VP, _I, _PD, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _PD, _PI, _PI]),
    "mppd_newton_mixed": (_I, [_VP, _VP, _PD, _VP, _VP, C.POINTER(NewtonParams), _VP, _PI, _PI]),
    "mppd_batch_params_default": (None, [_VP]),
    "mppd_batch_create": (_I, [C.POINTER(C.c_void_p), _I, _VP, _VP]),
    "mppd_batch_destroy": (_I, [_VP]),
    "mppd_batch_set_strains": (_I, [_VP, _PD]),
    "mppd_batch_homogenize": (_I, [_VP]),
    "mppd_batch_get_stress": (_I, [_VP, _PD]),
    "mppd_batch_get_ctan": (_I, [_VP, _PD]),
    "mppd_batch_get_info": (_I, [_VP, _PI]),
    "mppd_batch_update_vars": (_I, [_VP]),
    "mppd_batch_stats": (_I, [_VP, _PD]),
    "mppd_batch_mixed_stats": (_I, [_VP, C.POINTER(C.c_long)]),
    "mppd_batch_slot_bytes": (_I, [_VP, C.POINTER(C.c_size_t)]),
    "mppd_batch_tuning": (_I, [_VP, _PI]),
    "mppd_batch_spmv_profile": (_I, [_VP, _I, _I, _PD, _PD]),
    "mppd_batch_load": (_I, [_VP, _I, _VP, _VP, _VP]),
    "mppd_batch_store": (_I, [_VP, _I, _VP, _VP, _VP]),
}


class BatchRve(C.Structure):
    """``mppd_batch_rve`` (include/mppd.h)."""

    _fields_ = [("size", C.c_int * 3), ("micro_type", C.c_int), ("geo_params", C.c_double * 4),
                ("materials", MaterialBase * 3), ("coupling", C.c_int)]


class BatchParams(C.Structure):
    """``mppd_batch_params`` (include/mppd.h)."""

    _fields_ = [("newton", NewtonParams), ("lin_stress", C.c_int), ("calc_ctan_lin", C.c_int),
                ("subiterations", C.c_int), ("nsubiterations", C.c_int), ("graph_chunk", C.c_int),
                ("spmv_chunk", C.c_int), ("block_order", C.c_int), ("use_A0", C.c_int),
                ("mixed_precision", C.c_int)]


class MixedParams(C.Structure):
    """``mppd_mixed_params`` (include/mppd.h)."""

    _fields_ = [("rel_tol", C.c_double), ("abs_tol", C.c_double), ("inner_tol", C.c_double),
                ("max_inner_its", C.c_int), ("max_outer", C.c_int)]

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libmicropp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m micropp_amd.build`")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class MppdError(RuntimeError):
    pass


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().mppd_last_error().decode(errors="replace")
        raise MppdError(f"{what} failed ({rc}): {msg}")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_PD)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_PI)


def gpu_available() -> bool:
    n = C.c_int(0)
    try:
        return load().mppd_device_count(C.byref(n)) == 0 and n.value > 0
    except (RuntimeError, OSError):
        return False


def material(type_, E, nu, Ka=0.0, Sy=0.0, Xt=0.0) -> MaterialBase:
    """material_set through the library (reference src/material.c:26-38)."""
    m = MaterialBase()
    load().material_set(C.byref(m), int(type_), float(E), float(nu), float(Ka), float(Sy), float(Xt))
    return m


# ---------------------------------------------------------------------------
# micropp3 (include/micropp_c.h)
# ---------------------------------------------------------------------------
class Micropp3:
    """One ``micropp<3>`` instance through the C interface.

    Keyword arguments are the fields of ``micropp_params_t``; the defaults are
    the reference's (include/types.hpp:43-83).
    """

    def __init__(self, size, micro_type=MIC_HOMOGENEOUS, geo_params=(0.1, 0.1, 0.1, 0.1), materials=None, ngp=1,
                 coupling=None, **kw):
        lib = load()
        p = Micropp3Params()
        lib.micropp3_params_default(C.byref(p))
        p.ngp = ngp
        p.size[:] = list((size,) * 3 if np.isscalar(size) else size)
        p.type = micro_type
        p.geo_params[:] = list(geo_params)
        mats = materials or [material(0, 1e6, 0.3)] * 3
        for i in range(3):
            p.materials[i] = mats[i]
        self._coupling = None
        if coupling is not None:
            self._coupling = (C.c_int * ngp)(*coupling)
            p.coupling = C.cast(self._coupling, _PI)
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError(f"unknown parameter {k}")
            setattr(p, k, v)
        self.params = p
        self.h = Micropp3Handle()
        lib.micropp3_new_ext(C.byref(self.h), C.byref(p))
        if not self.h.ptr:
            raise MppdError("micropp3_new_ext failed (see stderr)")
        self.ngp = ngp
        nx, ny, nz = p.size
        self.nndim = 3 * nx * ny * nz
        self.nvars = (nx - 1) * (ny - 1) * (nz - 1) * 8 * NUM_VAR_GP

    def close(self):
        if getattr(self, "h", None) is not None and self.h.ptr:
            load().micropp3_free(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_strain(self, gp, eps):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        load().micropp3_set_strain(C.byref(self.h), gp, _dp(e))

    def get_stress(self, gp):
        s = np.zeros(6)
        load().micropp3_get_stress(C.byref(self.h), gp, _dp(s))
        return s

    def get_ctan(self, gp):
        c = np.zeros(36)
        load().micropp3_get_ctan(C.byref(self.h), gp, _dp(c))
        return c.reshape(6, 6)

    def homogenize(self):
        load().micropp3_homogenize(C.byref(self.h))

    def homogenize_linear(self):
        load().micropp3_homogenize_linear(C.byref(self.h))

    def update_vars(self):
        load().micropp3_update_vars(C.byref(self.h))

    def is_non_linear(self, gp):
        return int(load().micropp3_is_non_linear(C.byref(self.h), gp))

    def get_non_linear_gps(self):
        return load().micropp3_get_non_linear_gps(C.byref(self.h))

    def get_cost(self, gp):
        return load().micropp3_get_cost(C.byref(self.h), gp)

    def has_converged(self, gp):
        return bool(load().micropp3_has_converged(C.byref(self.h), gp))

    def has_subiterated(self, gp):
        return bool(load().micropp3_has_subiterated(C.byref(self.h), gp))

    def output(self, gp, filename):
        """write_vtu of GP gp into <filename>.vtu."""
        load().micropp3_output(C.byref(self.h), gp, str(filename).encode())

    def output2(self, gp, elem_global, time_step):
        """micropp-<elem_global>-<time_step>.vtu in the working directory."""
        load().micropp3_output2(C.byref(self.h), gp, elem_global, time_step)

    def write_restart(self, restart_id):
        load().micropp3_write_restart(C.byref(self.h), restart_id)

    def read_restart(self, restart_id):
        load().micropp3_read_restart(C.byref(self.h), restart_id)

    def print_info(self):
        load().micropp3_print_info(C.byref(self.h))

    def state(self, gp, which):
        n = self.nndim if which < 2 else self.nvars
        out = np.zeros(n)
        ok = load().micropp3_get_state(C.byref(self.h), gp, which, _dp(out))
        return out if ok else None

    def ctan_lin(self):
        c = np.zeros(36)
        load().micropp3_get_ctan_lin(C.byref(self.h), _dp(c))
        return c.reshape(6, 6)

    def volume_fractions(self):
        v = np.zeros(2)
        load().micropp3_get_volume_fractions(C.byref(self.h), _dp(v))
        return v


# ---------------------------------------------------------------------------
# kernel-level ABI (include/mppd.h)
# ---------------------------------------------------------------------------
class DeviceArray:
    """A device buffer (mppd_malloc) with numpy round trips."""

    def __init__(self, n, dtype=np.float64, host=None):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _chk(load().mppd_malloc(C.byref(p), max(self.nbytes, 1)), "mppd_malloc")
        self.ptr = p
        if host is not None:
            self.upload(host)
        else:
            _chk(load().mppd_memset(self.ptr, 0, max(self.nbytes, 1)), "mppd_memset")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        return cls(a.size, a.dtype, a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _chk(load().mppd_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes), "h2d")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _chk(load().mppd_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            load().mppd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    return None if a is None else a.ptr


class Rve:
    """mppd_rve: geometry, phase map and shape functions of one RVE on the GPU."""

    def __init__(self, size, micro_type, geo, materials):
        s = np.array(list((size,) * 3 if np.isscalar(size) else size), dtype=np.int32)
        g = np.ascontiguousarray(geo, dtype=np.float64)
        mats = (MaterialBase * 3)(*materials)
        h = C.c_void_p()
        _chk(load().mppd_rve_create(C.byref(h), _ip(s), int(micro_type), _dp(g), mats), "mppd_rve_create")
        self.h = h
        nrow, nelem, nvars = C.c_int(), C.c_int(), C.c_int()
        _chk(load().mppd_rve_info(h, C.byref(nrow), C.byref(nelem), C.byref(nvars)), "mppd_rve_info")
        self.nrow, self.nelem, self.nvars = nrow.value, nelem.value, nvars.value
        self.size = tuple(int(v) for v in s)

    def bmat(self):
        b = np.zeros(8 * 6 * 24)
        _chk(load().mppd_rve_bmat(self.h, _dp(b)), "mppd_rve_bmat")
        return b.reshape(8, 6, 24)

    def elem_types(self):
        et = np.zeros(self.nelem, dtype=np.int32)
        _chk(load().mppd_rve_elem_types(self.h, _ip(et)), "mppd_rve_elem_types")
        return et

    def volume_fractions(self):
        v = np.zeros(2)
        _chk(load().mppd_rve_volume_fractions(self.h, _dp(v)), "mppd_rve_volume_fractions")
        return v

    def __del__(self):
        try:
            if self.h:
                load().mppd_rve_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Solver:
    """mppd_solver: the ELL matrix and CG workspace of one RVE."""

    def __init__(self, rve: Rve):
        h = C.c_void_p()
        _chk(load().mppd_solver_create(C.byref(h), rve.h), "mppd_solver_create")
        self.h = h
        self.rve = rve
        self.nrow = rve.nrow

    def __del__(self):
        try:
            if self.h:
                load().mppd_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def ell(self):
        cols = np.zeros(self.nrow * 81, dtype=np.int32)
        vals = np.zeros(self.nrow * 81)
        _chk(load().mppd_ell_download(self.h, _ip(cols), _dp(vals)), "mppd_ell_download")
        return cols.reshape(-1, 81), vals.reshape(-1, 81)

    def upload_vals(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.float64)
        _chk(load().mppd_ell_upload_vals(self.h, _dp(v)), "mppd_ell_upload_vals")

    def set_displ_bc(self, eps, u: DeviceArray):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        _chk(load().mppd_set_displ_bc(self.h, _dp(e), u.ptr), "mppd_set_displ_bc")

    def assembly_rhs(self, u: DeviceArray, vars_: DeviceArray | None, b: DeviceArray):
        n = C.c_double()
        _chk(load().mppd_assembly_rhs(self.h, u.ptr, _ptr(vars_), b.ptr, C.byref(n)), "mppd_assembly_rhs")
        return n.value

    def assembly_mat(self, u: DeviceArray | None, vars_: DeviceArray | None = None):
        _chk(load().mppd_assembly_mat(self.h, _ptr(u), _ptr(vars_)), "mppd_assembly_mat")

    def mvp(self, x: DeviceArray, y: DeviceArray):
        _chk(load().mppd_ell_mvp(self.h, x.ptr, y.ptr), "mppd_ell_mvp")

    def cg(self, b: DeviceArray, x: DeviceArray, min_err=1e-50, rel_err=1e-5, max_its=1000, A0=None):
        its, err = C.c_int(), C.c_double()
        _chk(load().mppd_cg(self.h, None if A0 is None else A0.h, b.ptr, x.ptr, min_err, rel_err, max_its,
                            C.byref(its), C.byref(err)), "mppd_cg")
        return its.value, err.value

    def ave_stress(self, u: DeviceArray, vars_: DeviceArray | None = None):
        s = np.zeros(6)
        _chk(load().mppd_ave_stress(self.h, u.ptr, _ptr(vars_), _dp(s)), "mppd_ave_stress")
        return s

    def vars_new(self, u: DeviceArray, vars_old: DeviceArray | None, vars_new: DeviceArray):
        nl = C.c_int()
        _chk(load().mppd_vars_new(self.h, u.ptr, _ptr(vars_old), vars_new.ptr, C.byref(nl)), "mppd_vars_new")
        return nl.value

    def elem_fields(self, u: DeviceArray, vars_: DeviceArray | None = None):
        """calc_fields: (nelem x 6 strain, nelem x 6 stress)."""
        ne = self.rve.nelem
        es, ss = np.zeros(ne * 6), np.zeros(ne * 6)
        _chk(load().mppd_elem_fields(self.h, u.ptr, _ptr(vars_), _dp(es), _dp(ss)), "mppd_elem_fields")
        return es.reshape(ne, 6), ss.reshape(ne, 6)

    def newton(self, eps, u: DeviceArray, vars_old: DeviceArray | None = None, A0=None, nr_max_its=4,
               nr_max_tol=1e-10, nr_rel_tol=1e-3, cg_max_its=1000, cg_abs_tol=1e-50, cg_rel_tol=1e-5,
               its_with_A0=1):
        p = NewtonParams(nr_max_its, nr_max_tol, nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol, its_with_A0)
        e = np.ascontiguousarray(eps, dtype=np.float64)
        out = np.zeros(3, dtype=np.int32)
        _chk(load().mppd_newton(self.h, None if A0 is None else A0.h, _dp(e), _ptr(vars_old), u.ptr, C.byref(p),
                                _ip(out)), "mppd_newton")
        return tuple(int(v) for v in out)

    def newton_mixed(self, eps, u: DeviceArray, vars_old: DeviceArray | None = None, nr_max_its=4, nr_max_tol=1e-10,
                     nr_rel_tol=1e-3, cg_max_its=1000, cg_abs_tol=1e-50, cg_rel_tol=1e-5, inner_tol=1e-5,
                     max_inner_its=None, max_outer=10, A0=None, its_with_A0=1):
        """newton_raphson with FP32-inner / FP64-refined linear solves: ((its, solver_its, converged),
        (outer steps, iterations, FP64 fallbacks, solves stopped short by max_outer)).  A0 as in newton."""
        p = NewtonParams(nr_max_its, nr_max_tol, nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol, its_with_A0)
        mpar = MixedParams(cg_rel_tol, cg_abs_tol, inner_tol, max_inner_its or cg_max_its, max_outer)
        e = np.ascontiguousarray(eps, dtype=np.float64)
        out = np.zeros(3, dtype=np.int32)
        st = np.zeros(4, dtype=np.int32)
        _chk(load().mppd_newton_mixed(self.h, None if A0 is None else A0.h, _dp(e), _ptr(vars_old), u.ptr, C.byref(p),
                                      C.cast(C.byref(mpar), C.c_void_p), _ip(out), _ip(st)), "mppd_newton_mixed")
        return tuple(int(v) for v in out), tuple(int(v) for v in st)


class MSolver:
    """mppd_msolver: FP32-inner / FP64-refined solves of up to K strains on one elastic RVE."""

    NOT_CONVERGED = -5

    def __init__(self, rve: Rve, nrhs: int = 12):
        h = C.c_void_p()
        _chk(load().mppd_msolver_create(C.byref(h), rve.h, nrhs), "mppd_msolver_create")
        self.h = h
        self.rve = rve
        self.K = nrhs

    def __del__(self):
        try:
            if self.h:
                load().mppd_msolver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def solve(self, strains, rel_tol=1e-8, abs_tol=1e-50, inner_tol=1e-5, max_inner_its=2000, max_outer=10):
        """Averaged stress of each strain (n x 6), stats {outer, inner its}, final residual ratios."""
        e = np.ascontiguousarray(strains, dtype=np.float64).reshape(-1, 6)
        n = e.shape[0]
        sig = np.zeros((n, 6))
        stats = np.zeros(2, dtype=np.int32)
        rel = np.zeros(n)
        p = MixedParams(rel_tol, abs_tol, inner_tol, max_inner_its, max_outer)
        rc = load().mppd_msolver_solve(self.h, n, _dp(e), C.cast(C.byref(p), C.c_void_p), _dp(sig), _ip(stats),
                                       _dp(rel))
        if rc not in (0, self.NOT_CONVERGED):
            _chk(rc, "mppd_msolver_solve")
        return sig, (int(stats[0]), int(stats[1])), rel, rc == 0

    def adjoint(self, strains=None, rel_tol=1e-8, abs_tol=1e-50, inner_tol=1e-5, max_inner_its=2000, max_outer=10):
        """Adjoint derivatives: (ctan 6x6 = dsigma/deps, sigma n x 6, dsigma/d(E, nu) n x 6 x 3 x 2 [s][v][phase][E|nu],
        stats, residual ratios, converged)."""
        e = np.zeros((0, 6)) if strains is None else np.ascontiguousarray(strains, dtype=np.float64).reshape(-1, 6)
        n = e.shape[0]
        ctan = np.zeros(36)
        sig = np.zeros((max(n, 1), 6))
        dmat = np.zeros((max(n, 1), 36))
        stats = np.zeros(2, dtype=np.int32)
        rel = np.zeros(6 + n)
        p = MixedParams(rel_tol, abs_tol, inner_tol, max_inner_its, max_outer)
        e_arg = e if n else np.zeros((1, 6))
        rc = load().mppd_msolver_adjoint(self.h, n, _dp(np.ascontiguousarray(e_arg)), C.cast(C.byref(p), C.c_void_p),
                                         _dp(sig), _dp(ctan), _dp(dmat), _ip(stats), _dp(rel))
        if rc not in (0, self.NOT_CONVERGED):
            _chk(rc, "mppd_msolver_adjoint")
        return (ctan.reshape(6, 6), sig[:n], dmat[:n].reshape(n, 6, 3, 2), (int(stats[0]), int(stats[1])), rel,
                rc == 0)

    def u(self, k):
        out = np.zeros(self.rve.nrow)
        _chk(load().mppd_msolver_get_u(self.h, k, _dp(out)), "mppd_msolver_get_u")
        return out

    def newton(self, strains, u_in=None, u_out=None, nr_max_its=4, nr_max_tol=1e-10, nr_rel_tol=1e-3,
               cg_max_its=1000, cg_abs_tol=1e-50, cg_rel_tol=1e-5, mixed=None):
        """newton_raphson of n <= K Gauss points at once (mppd_msolver_newton): u_in / u_out are lists of
        DeviceArray or None (zero start / no output).  mixed: (inner_tol, max_inner_its, max_outer) or None.
        Returns (sigma n x 6, out n x 3 = its, cost, converged, stats (outer steps, inner iterations))."""
        e = np.ascontiguousarray(strains, dtype=np.float64).reshape(-1, 6)
        n = e.shape[0]
        p = NewtonParams(nr_max_its, nr_max_tol, nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol, 1)
        ins = (C.c_void_p * n)(*[None if a is None else a.ptr for a in (u_in or [None] * n)])
        outs = (C.c_void_p * n)(*[None if a is None else a.ptr for a in (u_out or [None] * n)])
        mpar = None if mixed is None else MixedParams(cg_rel_tol, cg_abs_tol, *mixed)
        sig = np.zeros((n, 6))
        out = np.zeros((n, 3), dtype=np.int32)
        stats = np.zeros(2, dtype=np.int32)
        _chk(load().mppd_msolver_newton(self.h, n, _dp(e), C.cast(ins, C.c_void_p), C.cast(outs, C.c_void_p),
                                        C.cast(C.byref(p), C.c_void_p),
                                        None if mpar is None else C.cast(C.byref(mpar), C.c_void_p), _dp(sig),
                                        _ip(out), _ip(stats)), "mppd_msolver_newton")
        return sig, out, (int(stats[0]), int(stats[1]))

    def profile(self, iters):
        a, t = C.c_double(), C.c_double()
        _chk(load().mppd_msolver_profile(self.h, iters, C.byref(a), C.byref(t)), "mppd_msolver_profile")
        return a.value, t.value

    def coarse(self, inverse=True):
        """The coarse level of the two-level inner preconditioner: (dims {nc, hx, hy, hz, mx, my, mz}, the FP32
        inverse of the Galerkin coarse matrix, nc x nc, or None).  nc = 0 without a coarse level."""
        dims = np.zeros(7, dtype=np.int32)
        _chk(load().mppd_msolver_coarse(self.h, _ip(dims), None), "mppd_msolver_coarse")
        nc = int(dims[0])
        if not (inverse and nc):
            return dims, None
        ainv = np.zeros((nc, nc), dtype=np.float32)
        _chk(load().mppd_msolver_coarse(self.h, _ip(dims), ainv.ctypes.data_as(C.POINTER(C.c_float))),
             "mppd_msolver_coarse")
        return dims, ainv


def adjoint_many(msolvers, strains=None, workers=0, rel_tol=1e-8, abs_tol=1e-50, inner_tol=1e-5, max_inner_its=2000,
                 max_outer=10):
    """MSolver.adjoint of many RVEs at once (mppd_msolver_adjoint_many: up to `workers` host threads, 0: 4, each
    solver on its own stream).  strains: None or (nm, n, 6).  Returns (ctan nm x 6 x 6, sigma nm x n x 6,
    dsigma/d(E, nu) nm x n x 6 x 3 x 2, stats nm x 2, residual ratios nm x (6 + n), converged)."""
    nm = len(msolvers)
    e = np.zeros((nm, 0, 6)) if strains is None else np.ascontiguousarray(strains, dtype=np.float64).reshape(nm, -1, 6)
    n = e.shape[1]
    ctan = np.zeros((nm, 36))
    sig = np.zeros((nm, max(n, 1), 6))
    dmat = np.zeros((nm, max(n, 1), 36))
    stats = np.zeros((nm, 2), dtype=np.int32)
    rel = np.zeros((nm, 6 + n))
    p = MixedParams(rel_tol, abs_tol, inner_tol, max_inner_its, max_outer)
    hs = (C.c_void_p * nm)(*[m.h for m in msolvers])
    e_arg = np.ascontiguousarray(e if n else np.zeros((nm, 1, 6)))
    rc = load().mppd_msolver_adjoint_many(nm, C.cast(hs, C.c_void_p), n, _dp(e_arg), C.cast(C.byref(p), C.c_void_p),
                                          int(workers), _dp(sig), _dp(ctan), _dp(dmat), _ip(stats), _dp(rel))
    if rc not in (0, MSolver.NOT_CONVERGED):
        _chk(rc, "mppd_msolver_adjoint_many")
    return (ctan.reshape(nm, 6, 6), sig[:, :n], dmat[:, :n].reshape(nm, n, 6, 3, 2), stats, rel, rc == 0)


def homogenize_mixed(msolver: MSolver, strains, d_eps=1e-8, **kw):
    """sigma of each strain and the linear-FE ctan from one multi-RHS solve.

    The six unit-strain fields of calc_ctan_lin_fe_models (reference
    src/micropp.cpp:255-284; eps = d_eps e_i from u = 0, ctan[:, i] =
    sigma / d_eps) and the strains' own fields run together.
    """
    strains = np.asarray(strains, dtype=np.float64).reshape(-1, 6)
    eps = np.vstack([np.eye(6) * d_eps, strains])
    sig, stats, rel, ok = msolver.solve(eps, **kw)
    ctan = sig[:6].T / d_eps
    return sig[6:], ctan, stats, rel, ok


class Batch:
    """mppd_batch: many RVEs (own size, geometry, materials, coupling, strain and history), one launch per
    finite-element operation for all of them; each RVE follows micropp<3>'s per-GP algorithm."""

    def __init__(self, rves, **params):
        """rves: list of dicts with size, micro_type, geo, materials (3), coupling (FE_* id).
        params: newton fields (nr_max_its, nr_max_tol, nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol,
        its_with_A0), lin_stress, calc_ctan_lin, subiterations, nsubiterations, graph_chunk, spmv_chunk,
        block_order, use_A0, mixed_precision."""
        p = BatchParams()
        load().mppd_batch_params_default(C.byref(p))
        for k, v in params.items():
            if hasattr(p.newton, k):
                setattr(p.newton, k, v)
            elif k in dict(BatchParams._fields_):
                setattr(p, k, int(v))
            else:
                raise TypeError(f"unknown batch parameter {k}")
        arr = (BatchRve * len(rves))()
        for i, d in enumerate(rves):
            s = d["size"]
            s = (s,) * 3 if np.isscalar(s) else tuple(s)
            arr[i].size[:] = [int(x) for x in s]
            arr[i].micro_type = int(d["micro_type"])
            arr[i].geo_params[:] = [float(x) for x in d["geo"]]
            for k in range(3):
                arr[i].materials[k] = d["materials"][k]
            arr[i].coupling = int(d.get("coupling", FE_ONE_WAY))
        h = C.c_void_p()
        _chk(load().mppd_batch_create(C.byref(h), len(rves), C.cast(arr, C.c_void_p), C.cast(C.byref(p),
                                                                                             C.c_void_p)),
             "mppd_batch_create")
        self.h = h
        self.n = len(rves)

    def __del__(self):
        try:
            if self.h:
                load().mppd_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_strains(self, eps):
        e = np.ascontiguousarray(eps, dtype=np.float64).reshape(self.n, 6)
        _chk(load().mppd_batch_set_strains(self.h, _dp(e)), "mppd_batch_set_strains")

    def homogenize(self):
        _chk(load().mppd_batch_homogenize(self.h), "mppd_batch_homogenize")

    def stress(self):
        out = np.zeros((self.n, 6))
        _chk(load().mppd_batch_get_stress(self.h, _dp(out)), "mppd_batch_get_stress")
        return out

    def ctan(self):
        out = np.zeros((self.n, 6, 6))
        _chk(load().mppd_batch_get_ctan(self.h, _dp(out)), "mppd_batch_get_ctan")
        return out

    def info(self):
        """n x 4: is_non_linear, cost, converged, subiterated."""
        out = np.zeros((self.n, 4), dtype=np.int32)
        _chk(load().mppd_batch_get_info(self.h, _ip(out)), "mppd_batch_get_info")
        return out

    def update_vars(self):
        _chk(load().mppd_batch_update_vars(self.h), "mppd_batch_update_vars")

    def stats(self):
        """CG iterations queued, graph chunk length, graph launches, ms of the last homogenize, host
        synchronisations and batched Newton iterations of the last homogenize."""
        out = np.zeros(6)
        _chk(load().mppd_batch_stats(self.h, _dp(out)), "mppd_batch_stats")
        return out

    def mixed_stats(self):
        """mixed_precision totals: (outer steps, FP64 fallbacks, solves stopped short by the outer cap)."""
        out = np.zeros(3, dtype=np.int64)
        _chk(load().mppd_batch_mixed_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_long))), "mppd_batch_mixed_stats")
        return tuple(int(v) for v in out)

    def tuning(self):
        """(graph chunk length, SpMV slots per load round, SpMV blocks grouped per XCD)."""
        out = np.zeros(3, dtype=np.int32)
        _chk(load().mppd_batch_tuning(self.h, _ip(out)), "mppd_batch_tuning")
        return tuple(int(v) for v in out)

    def spmv_profile(self, fp32=False, iters=50):
        """(ms per launch, algorithmic bytes per launch) of the batch's CG SpMV over all RVEs."""
        ms, nb = C.c_double(), C.c_double()
        _chk(load().mppd_batch_spmv_profile(self.h, int(fp32), int(iters), C.byref(ms), C.byref(nb)),
             "mppd_batch_spmv_profile")
        return ms.value, nb.value


def cg_profile(solver: "Solver", b: DeviceArray, iters: int):
    """Average (SpMV kernel ms, whole CG iteration ms) over ``iters`` iterations."""
    a, t = C.c_double(), C.c_double()
    _chk(load().mppd_cg_profile(solver.h, b.ptr, iters, C.byref(a), C.byref(t)), "mppd_cg_profile")
    return a.value, t.value


def material_eval(m: MaterialBase, eps: np.ndarray, hist: np.ndarray | None = None):
    """Stress, tangent and history update of ``m`` at n points, on the GPU."""
    eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1, 6)
    n = eps.shape[0]
    de = DeviceArray.from_numpy(eps)
    dh = DeviceArray.from_numpy(np.ascontiguousarray(hist, dtype=np.float64)) if hist is not None else None
    ds, dc, dhn = DeviceArray(n * 6), DeviceArray(n * 36), DeviceArray(n * NUM_VAR_GP)
    dnl = DeviceArray(n, np.int32)
    _chk(load().mppd_material_eval(C.byref(m), n, de.ptr, _ptr(dh), ds.ptr, dc.ptr, dhn.ptr, dnl.ptr),
         "mppd_material_eval")
    return (ds.numpy().reshape(n, 6), dc.numpy().reshape(n, 6, 6), dhn.numpy().reshape(n, NUM_VAR_GP),
            dnl.numpy())
