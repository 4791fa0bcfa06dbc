"""Loader of the in-tree HIP library cism_amd/libcismgpu.so (no fallback).

The product path has no CPU implementation: if the library is missing or no
GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes
import os

from .grid import CGrid

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcismgpu.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_G = ctypes.POINTER(CGrid)
_lib = None

STATUS = {
    0: "ok", 1: "NaN in PCG scalars", 2: "zero pivot in tridiagonal preconditioner", 3: "bad argument",
    4: "singular element Jacobian", 5: "asymmetric element matrix", 6: "negative or NaN basal traction",
    10: "HIP runtime failure",
    11: "unsupported configuration", 12: "iterative refinement did not converge",
}

# symbols declared in include/cismgpu.h (checked by the CPU test suite)
EXPORTS = (
    "cismgpu_last_error", "cismgpu_device_count",
    "cismgpu_pcg_solver_standard_3d", "cismgpu_pcg_solver_chrongear_3d", "cismgpu_matvec_multiply_structured_3d",
    "cismgpu_system_create", "cismgpu_system_destroy", "cismgpu_system_set_matrix", "cismgpu_system_get_matrix",
    "cismgpu_system_set_x", "cismgpu_system_get_x", "cismgpu_system_solve", "cismgpu_system_matvec",
    "cismgpu_system_bench_spmv", "cismgpu_system_halo_x", "cismgpu_system_assemble_3d", "cismgpu_system_get_efvs",
    "cismgpu_system_residual", "cismgpu_system_zero_x",
    "cismgpu_ensemble_create", "cismgpu_ensemble_destroy", "cismgpu_ensemble_autotune", "cismgpu_ensemble_solve",
    "cismgpu_ensemble_get_x", "cismgpu_ensemble_member_info", "cismgpu_bench_stream",
    "cismgpu_system_save_assembly", "cismgpu_system_restore_assembly",
    "cismgpu_system_work_vectors", "cismgpu_system_vector_op", "cismgpu_system_residual_velocity",
    "cismgpu_pcg_solver_standard_2d", "cismgpu_pcg_solver_chrongear_2d", "cismgpu_matvec_multiply_structured_2d",
    "cismgpu_system_create_2d", "cismgpu_system_set_matrix_2d", "cismgpu_system_assemble_2d", "cismgpu_system_diva_setup",
    "cismgpu_system_diva_velocity", "cismgpu_system_l1l2_velocity", "cismgpu_half_storage_supported",
    "cismgpu_residual_vector_3d_noalloc", "cismgpu_reprosum_calc",
    "cismgpu_system_calcbeta_setup", "cismgpu_system_calcbeta", "cismgpu_dropin_release",
    "cismgpu_system_internal_stress", "cismgpu_system_internal_stress_2d",
    "cismgpu_system_basal_friction",
)


class SolveOpts(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int32), ("precond", ctypes.c_int32), ("ncheck", ctypes.c_int32),
                ("maxiters", ctypes.c_int32), ("tolerance", ctypes.c_double), ("precision", ctypes.c_int32),
                ("inner_tol", ctypes.c_double), ("max_outer", ctypes.c_int32), ("time_spmv", ctypes.c_int32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("err", ctypes.c_double), ("niters", ctypes.c_int32), ("outer_iters", ctypes.c_int32),
                ("solve_ms", ctypes.c_double), ("spmv_ms", ctypes.c_double), ("spmv_count", ctypes.c_int32),
                ("spmv_kind", ctypes.c_int32), ("setup_ms", ctypes.c_double)]


class CismGpuError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: status {status} ({STATUS.get(status, '?')}): {last_error()}")


def load(path: str | None = None):
    """Load libcismgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(p)
    _declare(L)
    _lib = L
    return L


def _declare(L):
    L.cismgpu_last_error.restype = ctypes.c_char_p
    L.cismgpu_device_count.restype = ctypes.c_int32
    solver = [_G, _I, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
              ctypes.c_int32, _D, _I]
    for n in ("cismgpu_pcg_solver_standard_3d", "cismgpu_pcg_solver_chrongear_3d", "cismgpu_pcg_solver_standard_2d",
              "cismgpu_pcg_solver_chrongear_2d"):
        getattr(L, n).argtypes = solver
        getattr(L, n).restype = ctypes.c_int32
    for n in ("cismgpu_matvec_multiply_structured_3d", "cismgpu_matvec_multiply_structured_2d"):
        getattr(L, n).argtypes = [_G, _I, _D, _D, _D, _D, _D, _D, _D, _D]
        getattr(L, n).restype = ctypes.c_int32
    vp = ctypes.c_void_p
    L.cismgpu_system_create.argtypes = [_G, ctypes.POINTER(vp)]
    L.cismgpu_system_create_2d.argtypes = [_G, ctypes.POINTER(vp)]
    L.cismgpu_system_set_matrix_2d.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int32]
    L.cismgpu_system_destroy.argtypes = [vp]
    L.cismgpu_system_set_matrix.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int32]
    L.cismgpu_system_get_matrix.argtypes = [vp, _D, _D, _D, _D, _D, _D]
    L.cismgpu_system_set_x.argtypes = [vp, _D, _D]
    L.cismgpu_system_get_x.argtypes = [vp, _D, _D]
    L.cismgpu_system_solve.argtypes = [vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cismgpu_system_matvec.argtypes = [vp, _D, _D]
    L.cismgpu_system_bench_spmv.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _D, _D]
    L.cismgpu_system_zero_x.argtypes = [vp]
    L.cismgpu_residual_vector_3d_noalloc.argtypes = [_G] + [ctypes.c_void_p] * 13
    L.cismgpu_residual_vector_3d_noalloc.restype = ctypes.c_int32
    L.cismgpu_reprosum_calc.argtypes = [_D, ctypes.c_int64, ctypes.c_int32, _D]
    L.cismgpu_reprosum_calc.restype = ctypes.c_int32
    L.cismgpu_half_storage_supported.argtypes = [_G]
    L.cismgpu_half_storage_supported.restype = ctypes.c_int32
    L.cismgpu_bench_stream.argtypes = [ctypes.c_int64, ctypes.c_int32, _D, _D, _D]
    L.cismgpu_bench_stream.restype = ctypes.c_int32
    for n in ("cismgpu_system_zero_x", "cismgpu_system_create", "cismgpu_system_destroy", "cismgpu_system_set_matrix",
              "cismgpu_system_create_2d", "cismgpu_system_set_matrix_2d",
              "cismgpu_system_get_matrix", "cismgpu_system_set_x", "cismgpu_system_get_x", "cismgpu_system_solve",
              "cismgpu_system_matvec", "cismgpu_system_bench_spmv"):
        getattr(L, n).restype = ctypes.c_int32


def bench_stream(nbytes: int = 2 << 30, nrep: int = 10):
    """Measured HBM bandwidth (GB/s) of STREAM-style copy, triad and read kernels."""
    L = load()
    c, t, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(L.cismgpu_bench_stream(int(nbytes), int(nrep), ctypes.byref(c), ctypes.byref(t), ctypes.byref(r)),
          "bench_stream")
    return dict(copy=c.value, triad=t.value, read=r.value)


def reprosum_calc(arr) -> "list[float]":
    """cism_reprosum_calc on the device: exact, order-independent sums of the
    columns of arr (nsummands, nflds) (a 1-D array is one field)."""
    import numpy as np
    a = np.asfortranarray(np.asarray(arr, dtype=np.float64))
    if a.ndim == 1:
        a = a.reshape(-1, 1, order="F")
    out = np.zeros(a.shape[1])
    check(load().cismgpu_reprosum_calc(a.ctypes.data_as(_D), a.shape[0], a.shape[1], out.ctypes.data_as(_D)),
          "reprosum_calc")
    return out


def last_error() -> str:
    if _lib is None:
        return ""
    return (_lib.cismgpu_last_error() or b"").decode(errors="replace")


def check(status: int, what: str):
    if status != 0:
        raise CismGpuError(status, what)
