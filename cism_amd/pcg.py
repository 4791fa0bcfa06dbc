"""Host mirror of ``module glissade_velo_higher_pcg``: the 3-D and 2-D solvers.

Same names, argument meaning and error behaviour as the reference
(glissade_velo_higher_pcg.F90:56-58, :827-838, :2632-2643, :5528-5540 for
3-D; :67-78, :1541-1552, :5656-5663 for 2-D),
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


def _shapes(grid: Grid, two_d: bool):
    if two_d:
        return (grid.nvx, grid.nvy, 9), (grid.nvx, grid.nvy)
    return grid.mat_shape, grid.vec_shape


def _check_shapes(grid: Grid, A, vecs, two_d: bool = False):
    ms, vs = _shapes(grid, two_d)
    for a in A:
        if a.shape != ms:
            raise ValueError(f"matrix shape {a.shape} != {ms}")
    for v in vecs:
        if v.shape != vs:
            raise ValueError(f"vector shape {v.shape} != {vs}")


def _solver(name, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, linear_solve_ncheck,
            tolerance, maxiters):
    L = _lib.load()
    _check_shapes(grid, (Auu, Auv, Avu, Avv), (bu, bv, xu, xv), name.endswith("_2d"))
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


def _matvec(name, grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    L = _lib.load()
    two_d = name.endswith("_2d")
    _check_shapes(grid, (Auu, Auv, Avu, Avv), (xu, xv), two_d)
    vs = _shapes(grid, two_d)[1]
    yu = np.zeros(vs, order="F")
    yv = np.zeros(vs, order="F")
    act = _mask(active_vertex)
    g = grid.cstruct()
    st = getattr(L, "cismgpu_" + name)(ctypes.byref(g), act.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       _d(Auu), _d(Auv), _d(Avu), _d(Avv), _d(xu), _d(xv), _d(yu), _d(yv))
    _lib.check(st, name)
    return yu, yv


def matvec_multiply_structured_3d(grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    """y = A x on owned active vertices (zero elsewhere).  Returns (yu, yv)."""
    return _matvec("matvec_multiply_structured_3d", grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv)


def pcg_solver_standard_2d(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                           linear_solve_ncheck, tolerance, maxiters):
    """Standard PCG on the 9-point stencils A(nx-1, ny-1, 9); x(nx-1, ny-1) updated
    in place (grid.nz is ignored).  precond 2 is taken as 1 (glide_setup.F90:1363-1366).
    Returns (err, niters)."""
    return _solver("cismgpu_pcg_solver_standard_2d", grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv,
                   precond, linear_solve_ncheck, tolerance, maxiters)


def pcg_solver_chrongear_2d(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                            linear_solve_ncheck, tolerance, maxiters):
    """Chronopoulos-Gear PCG on the 9-point stencils; same contract as the standard 2-D solver."""
    return _solver("cismgpu_pcg_solver_chrongear_2d", grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv,
                   precond, linear_solve_ncheck, tolerance, maxiters)


def matvec_multiply_structured_2d(grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    """y = A x for the 9-point stencils on owned active vertices.  Returns (yu, yv)."""
    return _matvec("matvec_multiply_structured_2d", grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv)


class System:
    """Device-resident linear system (matrix, rhs and iterate kept in HBM).

    ``two_d=True`` makes a one-level system for the 2-D solves: matrices are
    the (nx-1, ny-1, 9) stencils, vectors (nx-1, ny-1), and ``get_matrix``
    returns one-level 27-point rows with the stencil in m = 10..18.
    """

    def __init__(self, grid: Grid, two_d: bool = False):
        self.grid = grid
        self.two_d = bool(two_d)
        self._L = _lib.load()
        self._h = ctypes.c_void_p()
        self._g = grid.cstruct()
        create = self._L.cismgpu_system_create_2d if self.two_d else self._L.cismgpu_system_create
        _lib.check(create(ctypes.byref(self._g), ctypes.byref(self._h)), "system_create")
        self._vs = (grid.nvx, grid.nvy) if self.two_d else grid.vec_shape

    def _zeros_vec(self):
        return np.zeros(self._vs, order="F")

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

    def _set(self):
        return self._L.cismgpu_system_set_matrix_2d if self.two_d else self._L.cismgpu_system_set_matrix

    def set_matrix(self, active_vertex, Auu, Auv, Avu, Avv, bu, bv):
        for a in (Auu, Auv, Avu, Avv, bu, bv):
            _d(a)
        _check_shapes(self.grid, (Auu, Auv, Avu, Avv), (bu, bv), self.two_d)
        act = _mask(active_vertex)
        P = lambda a: a.ctypes.data
        _lib.check(self._set()(self._h, P(act), P(Auu), P(Auv), P(Avu), P(Avv), P(bu), P(bv), 0),
                   "system_set_matrix")

    def set_matrix_device(self, active_ptr, Auu, Auv, Avu, Avv, bu, bv):
        """Device pointers (ints) to reference-layout arrays."""
        _lib.check(self._set()(self._h, active_ptr, Auu, Auv, Avu, Avv, bu, bv, 1), "system_set_matrix")

    def get_matrix(self):
        g = self.grid
        ms = (27, 1, g.nvx, g.nvy) if self.two_d else g.mat_shape
        A = [np.zeros(ms, order="F") for _ in range(4)]
        b = [self._zeros_vec(), self._zeros_vec()]
        _lib.check(self._L.cismgpu_system_get_matrix(self._h, *[_d(a) for a in A], _d(b[0]), _d(b[1])),
                   "system_get_matrix")
        return A, b

    def set_x(self, xu, xv):
        _check_shapes(self.grid, (), (xu, xv), self.two_d)
        _lib.check(self._L.cismgpu_system_set_x(self._h, _d(xu), _d(xv)), "system_set_x")

    def get_x(self):
        xu = self._zeros_vec()
        xv = self._zeros_vec()
        _lib.check(self._L.cismgpu_system_get_x(self._h, _d(xu), _d(xv)), "system_get_x")
        return xu, xv

    def zero_x(self):
        _lib.check(self._L.cismgpu_system_zero_x(self._h), "system_zero_x")

    def matvec(self):
        yu = self._zeros_vec()
        yv = self._zeros_vec()
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
                   spmv_kind=info.spmv_kind, setup_ms=info.setup_ms)
        if raise_on_error and st != 0:
            raise _lib.CismGpuError(st, "system_solve")
        return res

    def bench_spmv(self, kind=0, nrep=20):
        ms = ctypes.c_double()
        by = ctypes.c_double()
        _lib.check(self._L.cismgpu_system_bench_spmv(self._h, kind, nrep, ctypes.byref(ms), ctypes.byref(by)),
                   "system_bench_spmv")
        return ms.value, by.value
