"""The ctypes mirrors of the C-ABI's structs (siconos_amd/capi.py) against the
header: a small C program built with gcc from include/fc3d_hip.h prints every
field's offset and each struct's size, which must equal what ctypes lays out.
A field added to one side only would otherwise shift every later field."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from siconos_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = [("fc3d_hip_params", capi.HipParams), ("fc3d_hip_stats", capi.HipStats),
           ("fc3d_hip_batch_stats", capi.HipBatchStats), ("fc2d_hip_params", capi.Fc2dParams),
           ("fc2d_hip_stats", capi.Fc2dStats), ("SolverOptions", capi.SolverOptions),
           ("Callback", capi.Callback), ("FrictionContactProblem", capi.FrictionContactProblem),
           ("NumericsMatrix", capi.NumericsMatrix),
           ("NumericsSparseMatrix", capi.NumericsSparseMatrix),
           ("CSparseMatrix", capi.CSparseMatrix),
           ("SparseBlockStructuredMatrix", capi.SparseBlockStructuredMatrix)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to read the header")
def test_struct_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fc3d_hip.h"',
             '#include "fc2d_hip.h"', 'int main(void) {']
    for cname, cls in STRUCTS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=gnu11", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in STRUCTS:
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
