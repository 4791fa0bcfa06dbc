"""Batched ensemble of independent velocity solves (one persistent launch).

Each member is assembled on the GPU by :class:`~cism_amd.velo.BPSolver`;
the ensemble copies the assembled systems into batched storage and solves
all of them with FP64 iterative refinement around an FP32 (or BF16-matrix)
Chronopoulos-Gear inner solve in a single persistent kernel captured in a
hipGraph.  A per-size-bucket autotuner picks how many lanes share a column.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .velo import BPSolver


class EnsOpts(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("inner_tol", ctypes.c_double), ("inner_maxiters", ctypes.c_int32),
                ("max_outer", ctypes.c_int32), ("ncheck", ctypes.c_int32), ("bf16", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32)]


class EnsInfo(ctypes.Structure):
    _fields_ = [("solve_ms", ctypes.c_double), ("steps", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("max_outer", ctypes.c_int32), ("max_inner", ctypes.c_int32), ("total_inner", ctypes.c_int64),
                ("max_err", ctypes.c_double), ("barrier_error", ctypes.c_int32), ("grid_blocks", ctypes.c_int32),
                ("matrix_storage", ctypes.c_int32)]


def _declare(L):
    if getattr(L, "_ens_declared", False):
        return
    vp = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int32)
    L.cismgpu_ensemble_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.cismgpu_ensemble_destroy.argtypes = [vp]
    L.cismgpu_ensemble_autotune.argtypes = [vp, I, D]
    L.cismgpu_ensemble_solve.argtypes = [vp, ctypes.POINTER(EnsOpts), ctypes.POINTER(EnsInfo)]
    L.cismgpu_ensemble_get_x.argtypes = [vp, ctypes.c_int32, D, D]
    L.cismgpu_ensemble_member_info.argtypes = [vp, ctypes.c_int32, D, I, I, I]
    for n in ("cismgpu_ensemble_create", "cismgpu_ensemble_destroy", "cismgpu_ensemble_autotune",
              "cismgpu_ensemble_solve", "cismgpu_ensemble_get_x", "cismgpu_ensemble_member_info"):
        getattr(L, n).restype = ctypes.c_int32
    L._ens_declared = True


class Ensemble:
    def __init__(self, problems, keep_members=False):
        self._L = _lib.load()
        _declare(self._L)
        self.problems = list(problems)
        self.solvers = []
        self.assembly_ms = 0.0
        for p in self.problems:
            s = BPSolver(p)
            info = s.assemble(p.uvel, p.vvel)
            self.assembly_ms += info["total_ms"]
            self.solvers.append(s)
        handles = (ctypes.c_void_p * len(self.solvers))(*[s.sys.handle.value for s in self.solvers])
        self._h = ctypes.c_void_p()
        _lib.check(self._L.cismgpu_ensemble_create(len(self.solvers), handles, ctypes.byref(self._h)),
                   "ensemble_create")
        if not keep_members:
            for s in self.solvers:
                s.sys.close()
            self.solvers = []

    def close(self):
        if self._h:
            self._L.cismgpu_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        split = (ctypes.c_int32 * 3)()
        ms = (ctypes.c_double * 9)()
        _lib.check(self._L.cismgpu_ensemble_autotune(self._h, split, ms), "ensemble_autotune")
        return list(split), np.array(ms[:]).reshape(3, 3)

    def solve(self, *, tolerance=1e-10, inner_tol=1e-3, inner_maxiters=400, max_outer=20, ncheck=5, bf16=False,
              blocks_per_cu=0, raise_on_error=True):
        o = EnsOpts(tolerance, inner_tol, inner_maxiters, max_outer, ncheck, 1 if bf16 else 0, blocks_per_cu)
        info = EnsInfo()
        st = self._L.cismgpu_ensemble_solve(self._h, ctypes.byref(o), ctypes.byref(info))
        res = {k: getattr(info, k) for k, _ in EnsInfo._fields_}
        res["status"] = st
        if raise_on_error and st != 0:
            raise _lib.CismGpuError(st, "ensemble_solve")
        return res

    def get_x(self, m):
        g = self.problems[m].grid
        xu, xv = g.zeros_vec(), g.zeros_vec()
        D = ctypes.POINTER(ctypes.c_double)
        _lib.check(self._L.cismgpu_ensemble_get_x(self._h, m, xu.ctypes.data_as(D), xv.ctypes.data_as(D)),
                   "ensemble_get_x")
        return xu, xv

    def member_info(self, m):
        e = ctypes.c_double()
        o, t, f = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self._L.cismgpu_ensemble_member_info(self._h, m, ctypes.byref(e), ctypes.byref(o), ctypes.byref(t),
                                                        ctypes.byref(f)), "ensemble_member_info")
        return dict(err=e.value, outer=o.value, total_inner=t.value, flags=f.value)
