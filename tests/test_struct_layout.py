"""The ctypes mirrors of the C-ABI structs match include/cismgpu.h: same
size, and every field at the offset the C compiler gives the member of the
same name (gcc on the header; no library or device needed)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from cism_amd import _lib, basal, ensemble, grid, velo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = [
    ("cismgpu_grid", grid.CGrid),
    ("cismgpu_solve_opts", _lib.SolveOpts),
    ("cismgpu_solve_info", _lib.SolveInfo),
    ("cismgpu_assembly_in", velo.AssemblyIn),
    ("cismgpu_assembly_info", velo.AssemblyInfo),
    ("cismgpu_diva_in", velo.DivaIn),
    ("cismgpu_ensemble_opts", ensemble.EnsOpts),
    ("cismgpu_ensemble_info", ensemble.EnsInfo),
    ("cismgpu_calcbeta_in", basal.CalcbetaIn),
]


def test_ctypes_mirrors_match_header(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cismgpu.h"', "int main(void) {"]
    for cname, cls in MIRRORS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, *_ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in MIRRORS:
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, *_ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
