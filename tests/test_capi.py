"""The C-ABI library builds, loads without a GPU and exports every entry
point that include/cismgpu.h declares; without a device it fails loudly."""
import ctypes
import os
import re
import subprocess

import pytest

from cism_amd import _lib
from cism_amd.grid import Grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cismgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cismgpu_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cismgpu_pcg_solver_standard_3d", "cismgpu_pcg_solver_chrongear_3d",
              "cismgpu_matvec_multiply_structured_3d", "cismgpu_system_assemble_3d",
              "cismgpu_pcg_solver_standard_2d", "cismgpu_pcg_solver_chrongear_2d",
              "cismgpu_matvec_multiply_structured_2d", "cismgpu_system_create_2d", "cismgpu_system_set_matrix_2d",
              "cismgpu_system_assemble_2d", "cismgpu_system_diva_setup", "cismgpu_system_diva_velocity",
              "cismgpu_system_l1l2_velocity"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s(cismgpu_\w+)", out.stdout))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(L, n)
    assert set(_lib.EXPORTS) <= exported


def test_no_device_fails_loudly():
    L = _lib.load()
    if L.cismgpu_device_count() > 0:
        pytest.skip("a GPU is visible")
    g = Grid.from_owned(8, 8, 5).cstruct()
    h = ctypes.c_void_p()
    st = L.cismgpu_system_create(ctypes.byref(g), ctypes.byref(h))
    assert st == 10 and not h.value
    assert "device" in _lib.last_error().lower()


def test_bad_grid_is_rejected():
    L = _lib.load()
    g = Grid.from_owned(8, 8, 5)
    g.ilo = 1   # owned bounds inside the halo
    h = ctypes.c_void_p()
    assert L.cismgpu_system_create(ctypes.byref(g.cstruct()), ctypes.byref(h)) == 3


def test_2d_wrappers_check_shapes_before_any_device_call():
    """The 2-D host functions take (nx-1, ny-1, 9) stencils and (nx-1, ny-1)
    vectors; 3-D-shaped arrays are rejected on the host."""
    import numpy as np
    from cism_amd import pcg
    g = Grid.from_owned(8, 8, 5)
    act = np.ones((g.nvx, g.nvy), np.int32, order="F")
    A3 = [g.zeros_mat() for _ in range(4)]
    v3 = g.zeros_vec()
    with pytest.raises(ValueError):
        pcg.pcg_solver_chrongear_2d(g, act, *A3, v3, v3, v3.copy(order="F"), v3.copy(order="F"), 1, 5, 1e-8, 10)
    A2 = [np.zeros((g.nvx, g.nvy, 9), order="F") for _ in range(4)]
    with pytest.raises(ValueError):
        pcg.matvec_multiply_structured_2d(g, act, *A2, v3, v3)


def test_half_storage_size_guard():
    """The half (symmetric) storage SpMVs index one level of slots with
    32-bit offsets, so they are chosen only while nyo * nxp * 56 < 2^31;
    larger grids keep the full-row kernels (host-only query, no device)."""
    L = _lib.load()

    def ok(n, m, nz=11):
        return L.cismgpu_half_storage_supported(ctypes.byref(Grid.from_owned(n, m, nz).cstruct()))

    assert ok(1024, 1024) == 1 and ok(6000, 6000) == 1
    # 6208 x 6208: nxp = 6208, 6208 * 6208 * 56 = 2.158e9 >= 2^31
    assert ok(6208, 6208) == 0
    # the owned row width is padded to 64: 6145 owned columns count as 6208
    assert ok(6145, 6208) == 0 and ok(6144, 6208) == 1
    g = Grid.from_owned(8, 8, 5)
    g.ilo = 1
    assert L.cismgpu_half_storage_supported(ctypes.byref(g.cstruct())) == 0
