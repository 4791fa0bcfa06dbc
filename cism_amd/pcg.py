"""Host mirror of ``module glissade_velo_higher_pcg`` for the 3-D solvers.

Same names, argument meaning and error behaviour as the reference
(glissade_velo_higher_pcg.F90:56-58, :827-838, :2632-2643, :5528-5540),
implemented by the HIP library through its C ABI.  The reference aborts via
``write_log(GM_FATAL)`` when alpha/eta1 is NaN; here that raises
:class:`cism_amd._lib.CismGpuError` with status 1.  Non-convergence is not an
error: ``niters == maxiters`` and ``err`` holds the last checked residual.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .grid import Grid

HO_PRECOND_NONE = 0
HO_PRECOND_DIAG = 1
HO_PRECOND_SIA = 2
HO_PRECOND_TRIDIAG_LOCAL = 3
HO_PRECOND_TRIDIAG_GLOBAL = 4

FP64, MIXED_FP32, MIXED_BF16 = 0, 1, 2
ALGO_STANDARD, ALGO_CHRONGEAR = 2, 3


def _d(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.f_contiguous:
        raise TypeError("arrays must be float64 in Fortran order")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _mask(m) -> np.ndarray:
    return np.asfortranarray(np.asarray(m, dtype=np.int32))


def _check_shapes(grid: Grid, A, vecs):
    for a in A:
        if a.shape != grid.mat_shape:
            raise ValueError(f"matrix shape {a.shape} != {grid.mat_shape}")
    for v in vecs:
        if v.shape != grid.vec_shape:
            raise ValueError(f"vector shape {v.shape} != {grid.vec_shape}")


def _solver(name, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, linear_solve_ncheck,
            tolerance, maxiters):
    L = _lib.load()
    _check_shapes(grid, (Auu, Auv, Avu, Avv), (bu, bv, xu, xv))
    act = _mask(active_vertex)
    g = grid.cstruct()
    err = ctypes.c_double(float("nan"))
    nit = ctypes.c_int32(0)
    st = getattr(L, name)(ctypes.byref(g), act.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          _d(Auu), _d(Auv), _d(Avu), _d(Avv), _d(bu), _d(bv), _d(xu), _d(xv),
                          int(precond), int(linear_solve_ncheck), float(tolerance), int(maxiters),
                          ctypes.byref(err), ctypes.byref(nit))
    _lib.check(st, name)
    return err.value, nit.value


def pcg_solver_standard_3d(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                           linear_solve_ncheck, tolerance, maxiters):
    """Standard preconditioned CG; xu, xv updated in place.  Returns (err, niters)."""
    return _solver("cismgpu_pcg_solver_standard_3d", grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv,
                   precond, linear_solve_ncheck, tolerance, maxiters)


def pcg_solver_chrongear_3d(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                            linear_solve_ncheck, tolerance, maxiters):
    """Chronopoulos-Gear PCG; xu, xv updated in place.  Returns (err, niters)."""
    return _solver("cismgpu_pcg_solver_chrongear_3d", grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv,
                   precond, linear_solve_ncheck, tolerance, maxiters)


def matvec_multiply_structured_3d(grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    """y = A x on owned active vertices (zero elsewhere).  Returns (yu, yv)."""
    L = _lib.load()
    _check_shapes(grid, (Auu, Auv, Avu, Avv), (xu, xv))
    yu = grid.zeros_vec()
    yv = grid.zeros_vec()
    act = _mask(active_vertex)
    g = grid.cstruct()
    st = L.cismgpu_matvec_multiply_structured_3d(ctypes.byref(g), act.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                 _d(Auu), _d(Auv), _d(Avu), _d(Avv), _d(xu), _d(xv), _d(yu), _d(yv))
    _lib.check(st, "matvec_multiply_structured_3d")
    return yu, yv


class System:
    """Device-resident linear system (matrix, rhs and iterate kept in HBM)."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self._L = _lib.load()
        self._h = ctypes.c_void_p()
        self._g = grid.cstruct()
        _lib.check(self._L.cismgpu_system_create(ctypes.byref(self._g), ctypes.byref(self._h)), "system_create")

    def close(self):
        if self._h:
            self._L.cismgpu_system_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_matrix(self, active_vertex, Auu, Auv, Avu, Avv, bu, bv):
        _check_shapes(self.grid, (Auu, Auv, Avu, Avv), (bu, bv))
        act = _mask(active_vertex)
        P = lambda a: a.ctypes.data
        _lib.check(self._L.cismgpu_system_set_matrix(self._h, P(act), P(Auu), P(Auv), P(Avu), P(Avv), P(bu), P(bv), 0),
                   "system_set_matrix")

    def set_matrix_device(self, active_ptr, Auu, Auv, Avu, Avv, bu, bv):
        """Device pointers (ints) to reference-layout arrays."""
        _lib.check(self._L.cismgpu_system_set_matrix(self._h, active_ptr, Auu, Auv, Avu, Avv, bu, bv, 1),
                   "system_set_matrix")

    def get_matrix(self):
        g = self.grid
        A = [g.zeros_mat() for _ in range(4)]
        b = [g.zeros_vec(), g.zeros_vec()]
        _lib.check(self._L.cismgpu_system_get_matrix(self._h, *[_d(a) for a in A], _d(b[0]), _d(b[1])),
                   "system_get_matrix")
        return A, b

    def set_x(self, xu, xv):
        _check_shapes(self.grid, (), (xu, xv))
        _lib.check(self._L.cismgpu_system_set_x(self._h, _d(xu), _d(xv)), "system_set_x")

    def get_x(self):
        xu = self.grid.zeros_vec()
        xv = self.grid.zeros_vec()
        _lib.check(self._L.cismgpu_system_get_x(self._h, _d(xu), _d(xv)), "system_get_x")
        return xu, xv

    def zero_x(self):
        _lib.check(self._L.cismgpu_system_zero_x(self._h), "system_zero_x")

    def matvec(self):
        yu = self.grid.zeros_vec()
        yv = self.grid.zeros_vec()
        _lib.check(self._L.cismgpu_system_matvec(self._h, _d(yu), _d(yv)), "system_matvec")
        return yu, yv

    def solve(self, *, algo=ALGO_CHRONGEAR, precond=HO_PRECOND_SIA, ncheck=5, maxiters=200, tolerance=1e-8,
              precision=FP64, inner_tol=1e-4, max_outer=10, time_spmv=False, raise_on_error=True):
        o = _lib.SolveOpts(algo, precond, ncheck, maxiters, tolerance, precision, inner_tol, max_outer,
                           1 if time_spmv else 0)
        info = _lib.SolveInfo()
        st = self._L.cismgpu_system_solve(self._h, ctypes.byref(o), ctypes.byref(info))
        res = dict(status=st, err=info.err, niters=info.niters, outer_iters=info.outer_iters,
                   solve_ms=info.solve_ms, spmv_ms=info.spmv_ms, spmv_count=info.spmv_count,
                   spmv_kind=info.spmv_kind)
        if raise_on_error and st != 0:
            raise _lib.CismGpuError(st, "system_solve")
        return res

    def bench_spmv(self, kind=0, nrep=20):
        ms = ctypes.c_double()
        by = ctypes.c_double()
        _lib.check(self._L.cismgpu_system_bench_spmv(self._h, kind, nrep, ctypes.byref(ms), ctypes.byref(by)),
                   "system_bench_spmv")
        return ms.value, by.value
