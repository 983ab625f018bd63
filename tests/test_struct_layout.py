"""CPU: the ctypes mirrors of the C ABI structs have the C layout.

A small C program, compiled with gcc against include/, prints sizeof and the
offset of every field the Python mirror declares; a field the C struct lacks
fails the compile.  A mismatch would pass wrong parameters silently (the batch
and Newton parameter structs grow as options are added), so sizes and offsets
must agree exactly.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
C_NAME = {"lambda_": "lambda"}  # Python keywords renamed in the mirrors


def mirrors():
    import micropp_amd as mp

    return {
        "struct material_base": mp.MaterialBase,
        "struct micropp3_params": mp.Micropp3Params,
        "mppd_newton_params": mp.NewtonParams,
        "mppd_mixed_params": mp.MixedParams,
        "mppd_batch_rve": mp.BatchRve,
        "mppd_batch_params": mp.BatchParams,
    }


def test_ctypes_mirrors_match_c_layout(tmp_path):
    import ctypes as C

    ms = mirrors()
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "mppd.h"', '#include "micropp_c.h"',
           "int main(void) {"]
    for cname, cls in ms.items():
        src.append(f'  printf("{cname}|<sizeof>|%zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            src.append(f'  printf("{cname}|{fname}|%zu\\n", offsetof({cname}, {C_NAME.get(fname, fname)}));')
    src += ["  return 0;", "}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", INC, str(c), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    seen = 0
    for line in out:
        cname, what, val = line.split("|")
        cls = ms[cname]
        want = C.sizeof(cls) if what == "<sizeof>" else getattr(cls, what).offset
        assert int(val) == want, f"{cname}.{what}: C {val}, ctypes {want}"
        seen += 1
    assert seen == sum(1 + len(cls._fields_) for cls in ms.values())
