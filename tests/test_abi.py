"""The C-ABI library loads and exports every entry point include/*.h declares.

CPU-only: no compute call is made.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_c_functions():
    names = set()
    for h in ("material_base.h", "micropp_c.h", "mppd.h"):
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:struct\s+)?\w+[\s\*]+(\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib():
    import micropp_amd

    if not os.path.exists(micropp_amd.LIB_PATH):
        from micropp_amd import build

        build.build()
    return micropp_amd.load()


def test_headers_parse():
    names = declared_c_functions()
    assert {"micropp3_new", "micropp3_homogenize", "mppd_cg", "mppd_ell_mvp", "material_set"} <= names
    assert len(names) > 40


def test_c_symbols_exported(lib):
    import micropp_amd

    missing = [n for n in declared_c_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the same set
    assert declared_c_functions() == set(micropp_amd.SIGNATURES)


def test_cpp_api_exported(lib):
    import micropp_amd

    out = subprocess.run(["nm", "-DC", "--defined-only", micropp_amd.LIB_PATH], capture_output=True,
                         text=True).stdout
    for meth in ["micropp<3>::micropp(micropp_params_t const&)", "micropp<3>::~micropp()",
                 "micropp<3>::set_strain(int, double const*)", "micropp<3>::get_stress(int, double*) const",
                 "micropp<3>::get_ctan(int, double*) const", "micropp<3>::homogenize()",
                 "micropp<3>::homogenize_linear()", "micropp<3>::update_vars()",
                 "micropp<3>::is_non_linear(int) const", "micropp<3>::get_non_linear_gps() const",
                 "micropp<3>::get_cost(int) const", "micropp<3>::has_converged(int) const",
                 "micropp<3>::has_subiterated(int) const", "micropp<3>::output(int, char const*)",
                 "micropp<3>::output2(int, int, int)", "micropp<3>::write_restart(int) const",
                 "micropp<3>::read_restart(int) const", "micropp<3>::print_info() const"]:
        assert meth in out, meth


def test_material_set_matches_reference_formula(lib):
    """material_set is pure host arithmetic (no device needed)."""
    import micropp_amd as mp

    m = mp.material(1, 3.0e7, 0.25, 1e7, 1e5, 0.0)
    assert m.k == 3.0e7 / (3.0 * (1.0 - 2.0 * 0.25))
    assert m.mu == 3.0e7 / (2.0 * (1.0 + 0.25))
    assert m.lambda_ == 0.25 * 3.0e7 / ((1.0 + 0.25) * (1.0 - 2.0 * 0.25))
    assert m.type == 1 and m.Ka == 1e7 and m.Sy == 1e5


def test_no_oracle_in_product():
    """The product package never imports or links the oracle."""
    pkg = os.path.join(ROOT, "micropp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f


def test_c_and_fortran_callers_build(lib, tmp_path):
    """The C header compiles as C99 and links; the Fortran drop-in modules libmaterial / libmicropp
    (micropp_amd/fortran/) compile and a Fortran caller that uses them links."""
    import shutil

    libdir = os.path.join(ROOT, "micropp_amd")
    drv = os.path.join(ROOT, "tests", "drivers")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{INC}", "-o", str(tmp_path / "c_driver"),
                    os.path.join(drv, "c_driver.c"), f"-L{libdir}", "-lmicropp_amd"], check=True)
    fc = next((c for c in ("/opt/rocm/bin/amdflang", shutil.which("flang") or "") if c and os.path.exists(c)), None)
    if fc is None:
        pytest.skip("no Fortran compiler")
    mods = [os.path.join(libdir, "fortran", f) for f in ("libmaterial.f95", "libmicropp.f95")]
    subprocess.run([fc, "-o", str(tmp_path / "f_driver")] + mods + [os.path.join(drv, "f_driver.f90"), "-module-dir",
                    str(tmp_path), f"-L{libdir}", "-lmicropp_amd"], check=True, cwd=str(tmp_path))
