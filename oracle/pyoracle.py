"""ctypes bindings of the CPU restatement (oracle/libcism_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import this.
Arrays use the reference layouts (Fortran order), see cism_amd/grid.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcism_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _declare(L):
    from cism_amd.grid import CGrid
    G = ctypes.POINTER(CGrid)
    solver_args = [G, _I, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int32, ctypes.c_int32,
                   ctypes.c_double, ctypes.c_int32, _D, _I]
    for name in ("orc_pcg_solver_standard_3d", "orc_pcg_solver_chrongear_3d"):
        f = getattr(L, name)
        f.argtypes = solver_args
        f.restype = ctypes.c_int
    L.orc_pcg_solver_chrongear_3d_omp.argtypes = solver_args[:-2] + [ctypes.c_int32, _D, _I]
    L.orc_pcg_solver_chrongear_3d_omp.restype = ctypes.c_int
    L.orc_matvec_3d.argtypes = [G, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_matvec_2d.argtypes = [G, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    for name in ("orc_pcg_solver_standard_2d", "orc_pcg_solver_chrongear_2d"):
        getattr(L, name).argtypes = solver_args
        getattr(L, name).restype = ctypes.c_int
    L.orc_halo3d.argtypes = [G, ctypes.c_int, _D]
    L.orc_global_sum.argtypes = [G, _D, _D]
    L.orc_global_sum.restype = ctypes.c_double
    L.orc_ddpdd_sum.argtypes = [ctypes.c_int64, _D, _D, _D]
    L.orc_reprosum_calc.argtypes = [ctypes.c_int64, _D]
    L.orc_reprosum_calc.restype = ctypes.c_double
    L.orc_tridiag_solver.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D]
    L.orc_set_trace.argtypes = [_D, ctypes.c_int32]
    L.orc_trace_len.restype = ctypes.c_int32
    L.orc_fe_tables.argtypes = [_D, _D, _D, _D]
    i = ctypes.c_int
    d = ctypes.c_double
    L.orc_calclsrf.argtypes = [i, i, _D, _D, d, _D]
    L.orc_get_masks.argtypes = [i, i, _D, _D, d, d, _I, _I, _I, _I]
    L.orc_stagger.argtypes = [i, i, _D, _D]
    L.orc_surface_elevation_gradient.argtypes = [i, i, d, d, _I, _I, _D, _D, d, d, i, d, _D, _D]
    L.orc_slope_angle.argtypes = [i, i, d, d, _D, _D]
    L.orc_vertex_masks.argtypes = [i, i, i, _I, _I, _I]
    L.orc_flwafact.argtypes = [i, i, i, i, d, _D, _D]
    L.orc_load_vector_gravity_local.argtypes = [i, i, i, _D, d, d, _I, _D, _D, _D, _D, _D]
    L.orc_element_matrix.argtypes = [_D, _D, _D, _D, _D, d, i, d, d, d, _D, _D, _D, _D, _D]
    L.orc_element_matrix.restype = ctypes.c_int
    L.orc_assemble_stiffness_matrix_3d.argtypes = [i, i, i, _D, i, _I, _D, _D, _D, _D, _D, _D, _D,
                                                   i, d, d, d, _D, _D, _D, _D, _D]
    L.orc_assemble_stiffness_matrix_3d.restype = ctypes.c_int
    L.orc_basal_sliding_bc_local.argtypes = [i, i, i, d, d, _I, _D, _D, _D, _D]
    L.orc_dirichlet_bc_3d.argtypes = [i, i, i, i, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_check_symmetry_assembled_3d.argtypes = [i, i, i, i, i, i, i, _I, _D, _D, _D, _D]
    L.orc_check_symmetry_assembled_3d.restype = ctypes.c_int64
    L.orc_count_nonzeros_3d.argtypes = [i, i, i, i, i, i, i, _I, _D, _D, _D, _D]
    L.orc_count_nonzeros_3d.restype = ctypes.c_int64
    L.orc_residual_velocity_3d.argtypes = [i, i, i, i, i, i, i, i, _D, _D, _D, _D]
    L.orc_residual_velocity_3d.restype = ctypes.c_double
    L.orc_stagger_masked.argtypes = [i, i, _D, _I, _D]
    L.orc_load_vector_lateral_bc.argtypes = [i, i, i, _D, i, i, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_grounded_fraction_no_glp.argtypes = [i, i, _I, _I, _I, _D]
    L.orc_vertical_average.argtypes = [i, i, i, _D, _D, _D]
    L.orc_element_matrix_2d_ssa.argtypes = [_D, _D, _D, _D, _D, d, i, d, d, d, _D, _D, _D, _D, _D]
    L.orc_element_matrix_2d_ssa.restype = ctypes.c_int
    L.orc_assemble_stiffness_matrix_2d_ssa.argtypes = [i, i, i, _D, i, _I, _D, _D, _D, _D, _D, _D, i, d, d, d,
                                                       _D, _D, _D, _D, _D]
    L.orc_assemble_stiffness_matrix_2d_ssa.restype = ctypes.c_int
    L.orc_basal_sliding_bc_2d_local.argtypes = [i, i, d, d, _I, _D, _D, _D]
    L.orc_dirichlet_bc_2d.argtypes = [i, i, i, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_check_symmetry_assembled_2d.argtypes = [i, i, i, i, i, i, _I, _D, _D, _D, _D]
    L.orc_check_symmetry_assembled_2d.restype = ctypes.c_int64
    L.orc_count_nonzeros_2d.argtypes = [i, i, i, i, i, i, _I, _D, _D, _D, _D]
    L.orc_count_nonzeros_2d.restype = ctypes.c_int64
    L.orc_residual_velocity_2d.argtypes = [i, i, i, i, i, i, i, _D, _D, _D, _D]
    L.orc_residual_velocity_2d.restype = ctypes.c_double
    L.orc_slope_angle_xy.argtypes = [i, i, d, d, _D, _D, _D, _D]
    L.orc_slope_angle_staggered.argtypes = [i, i, d, d, _D, _D, _D, _D]
    L.orc_cell_halo.argtypes = [i, i, i, _D]
    L.orc_calcbeta.argtypes = [i, i, i, i, i, _D] + [_D] * 7 + [_I] + [_D] * 5
    L.orc_compute_internal_stress.argtypes = [i, i, i, _D, i, _I, _D, _D, _D, _D, _D, i, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, _D, _D, _D]
    L.orc_compute_internal_stress.restype = ctypes.c_int
    L.orc_compute_internal_stress_approx.argtypes = [i, i, i, _D, i, _I, _D, _D, _D, _D, i, _D, i, _D,
                                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, _D]
    L.orc_compute_internal_stress_approx.restype = ctypes.c_int
    L.orc_basal_friction.argtypes = [i, i, i, i, i, _I, _I, _D, _D, _D, _D, _D, _D]
    L.orc_compute_grounded_fraction.argtypes = [_D, _D]
    L.orc_compute_grounded_fraction.restype = ctypes.c_int
    L.orc_grounded_fraction.argtypes = [i, i, i, i, i, i, _D, _D, ctypes.c_double, _I, _I, _I, _D, _D]
    L.orc_grounded_fraction.restype = ctypes.c_int
    L.orc_calcbeta.restype = ctypes.c_int
    L.orc_assemble_stiffness_matrix_2d_diva.argtypes = [i, i, i, _D, _D, i, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D,
                                                        _D, _D, i, d, d, d, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_assemble_stiffness_matrix_2d_diva.restype = ctypes.c_int
    L.orc_diva_beta_eff.argtypes = [i, i, i, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_basal_sliding_bc_2d_diva_local.argtypes = [i, i, d, d, _I, _D, _D, _D, _D, _D, _D]
    L.orc_compute_3d_velocity_diva.argtypes = [i, i, i, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D,
                                               _D, _D]
    L.orc_load_vector_gravity_standard.argtypes = [i, i, i, i, _D, _I, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_load_vector_gravity_standard.restype = ctypes.c_int
    L.orc_basal_sliding_bc_standard.argtypes = [i, i, i, i, i, i, _I, _D, _D, _D, _D, _D, _D, _D]
    L.orc_basal_sliding_bc_standard.restype = ctypes.c_int
    L.orc_l1l2_tau_parallel.argtypes = [d, d, d]
    L.orc_l1l2_tau_parallel.restype = ctypes.c_double
    L.orc_assemble_stiffness_matrix_2d_l1l2.argtypes = [i, i, i, _D, i, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, i, d,
                                                        d, d, _D, _D, _D, _D, _D]
    L.orc_assemble_stiffness_matrix_2d_l1l2.restype = ctypes.c_int
    L.orc_compute_3d_velocity_l1l2.argtypes = [i, i, i, i, i, i, i, i, i, _D, d, d, _I, _I, _I, _I, _I, _D, _D, _D,
                                               _D, _D, _D, _D, _D, i, d, _D, _D]
    L.orc_compute_3d_velocity_l1l2.restype = ctypes.c_int


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous), a.dtype
    return a.ctypes.data_as(_D)


def ip(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_I)


def as_mask(m) -> np.ndarray:
    return np.asfortranarray(np.asarray(m, dtype=np.int32))


# ---------------------------------------------------------------- solver path
def pcg_solve(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, *, algo="chrongear",
              precond=2, ncheck=5, tolerance=1e-8, maxiters=200, trace=False, nthreads=None):
    """Run the restated solver in place on xu, xv.  Returns (status, err, niters[, trace])."""
    L = lib()
    g = grid.cstruct()
    err = ctypes.c_double(float("nan"))
    nit = ctypes.c_int32(0)
    act = as_mask(active_vertex)
    tbuf = None
    if trace:
        tbuf = np.zeros(3 * (maxiters + 1))
        L.orc_set_trace(dp(tbuf), maxiters + 1)
    else:
        L.orc_set_trace(None, 0)
    args = [ctypes.byref(g), ip(act), dp(Auu), dp(Auv), dp(Avu), dp(Avv), dp(bu), dp(bv), dp(xu), dp(xv),
            int(precond), int(ncheck), float(tolerance), int(maxiters)]
    if nthreads is not None:
        st = L.orc_pcg_solver_chrongear_3d_omp(*args, int(nthreads), ctypes.byref(err), ctypes.byref(nit))
    elif algo == "chrongear":
        st = L.orc_pcg_solver_chrongear_3d(*args, ctypes.byref(err), ctypes.byref(nit))
    else:
        st = L.orc_pcg_solver_standard_3d(*args, ctypes.byref(err), ctypes.byref(nit))
    out = (st, err.value, nit.value)
    if trace:
        n = L.orc_trace_len()
        out = out + (tbuf[: 3 * n].reshape(n, 3).copy(),)
        L.orc_set_trace(None, 0)
    return out


def matvec(grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    L = lib()
    g = grid.cstruct()
    yu = grid.zeros_vec()
    yv = grid.zeros_vec()
    L.orc_matvec_3d(ctypes.byref(g), ip(as_mask(active_vertex)), dp(Auu), dp(Auv), dp(Avu), dp(Avv),
                    dp(xu), dp(xv), dp(yu), dp(yv))
    return yu, yv


def pcg_solve_2d(grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, *, algo="chrongear", precond=1,
                 ncheck=5, tolerance=1e-8, maxiters=200, trace=False):
    """Restated pcg_solver_{standard,chrongear}_2d in place on xu, xv (nx-1, ny-1);
    A(nx-1, ny-1, 9).  Returns (status, err, niters[, trace])."""
    L = lib()
    g = grid.cstruct()
    err = ctypes.c_double(float("nan"))
    nit = ctypes.c_int32(0)
    act = as_mask(active_vertex)
    tbuf = None
    if trace:
        tbuf = np.zeros(3 * (maxiters + 1))
        L.orc_set_trace(dp(tbuf), maxiters + 1)
    else:
        L.orc_set_trace(None, 0)
    args = [ctypes.byref(g), ip(act), dp(Auu), dp(Auv), dp(Avu), dp(Avv), dp(bu), dp(bv), dp(xu), dp(xv),
            int(precond), int(ncheck), float(tolerance), int(maxiters), ctypes.byref(err), ctypes.byref(nit)]
    f = L.orc_pcg_solver_chrongear_2d if algo == "chrongear" else L.orc_pcg_solver_standard_2d
    st = f(*args)
    out = (st, err.value, nit.value)
    if trace:
        n = L.orc_trace_len()
        out = out + (tbuf[: 3 * n].reshape(n, 3).copy(),)
        L.orc_set_trace(None, 0)
    return out


def matvec_2d(grid, active_vertex, Auu, Auv, Avu, Avv, xu, xv):
    L = lib()
    g = grid.cstruct()
    yu = np.zeros((grid.nvx, grid.nvy), order="F")
    yv = np.zeros((grid.nvx, grid.nvy), order="F")
    L.orc_matvec_2d(ctypes.byref(g), ip(as_mask(active_vertex)), dp(Auu), dp(Auv), dp(Avu), dp(Avv),
                    dp(xu), dp(xv), dp(yu), dp(yv))
    return yu, yv


def halo(grid, a: np.ndarray):
    L = lib()
    g = grid.cstruct()
    nlev = a.shape[0] if a.ndim == 3 else 1
    L.orc_halo3d(ctypes.byref(g), nlev, dp(a))


def global_sum(grid, a1, a2=None):
    L = lib()
    g = grid.cstruct()
    return L.orc_global_sum(ctypes.byref(g), dp(a1), dp(a2) if a2 is not None else None)


def ddpdd_sum(arr: np.ndarray):
    L = lib()
    hi = ctypes.c_double()
    lo = ctypes.c_double()
    a = np.ascontiguousarray(arr, dtype=np.float64)
    L.orc_ddpdd_sum(a.size, dp(a), ctypes.byref(hi), ctypes.byref(lo))
    return hi.value, lo.value


def reprosum(arr: np.ndarray) -> float:
    """cism_reprosum_calc (integer vectors) of a 1-D array of summands."""
    a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
    return lib().orc_reprosum_calc(a.size, dp(a))


def tridiag(sb, d, sp, rhs):
    L = lib()
    n = len(d)
    out = np.zeros(n)
    L.orc_tridiag_solver(n, dp(np.ascontiguousarray(sb, float)), dp(np.ascontiguousarray(d, float)),
                         dp(np.ascontiguousarray(sp, float)), dp(np.ascontiguousarray(rhs, float)), dp(out))
    return out


# ------------------------------------------------------------ assembly path
def fe_tables():
    L = lib()
    t = [np.zeros((8, 8), order="F") for _ in range(4)]
    L.orc_fe_tables(*[dp(a) for a in t])
    return t   # phi, dphi_dxr, dphi_dyr, dphi_dzr as (n, p)


def element_matrix(x, y, z, u, v, flwafact, whichefvs=2, efvs_constant=0.0, effstrain_min=1e-8, n_glen=3.0):
    L = lib()
    K = [np.zeros((8, 8)) for _ in range(4)]   # C order [nr][nc]
    em = ctypes.c_double()
    f = lambda a: dp(np.ascontiguousarray(a, dtype=np.float64))
    st = L.orc_element_matrix(f(x), f(y), f(z), f(u), f(v), flwafact, whichefvs, efvs_constant,
                              effstrain_min, n_glen, *[dp(k) for k in K], ctypes.byref(em))
    return st, K, em.value


def geometry(problem):
    """Setup part of glissade_velo_higher_solve for the 3-D BP path
    (glissade_velo_higher.F90:1247-1847): lsrf/usrf, masks, staggered fields,
    surface gradient with its halo, active masks, flwafact."""
    L = lib()
    p = problem
    g = p.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    lsrf = np.zeros((nx, ny), order="F")
    L.orc_calclsrf(nx, ny, dp(p.thck), dp(p.topg), p.eus, dp(lsrf))
    usrf = np.asfortranarray(np.maximum(0.0, p.thck + lsrf))
    ice = np.zeros((nx, ny), np.int32, order="F")
    flt = np.zeros_like(ice); ocn = np.zeros_like(ice); land = np.zeros_like(ice)
    L.orc_get_masks(nx, ny, dp(p.thck), dp(p.topg), p.eus, p.thklim, ip(ice), ip(flt), ip(ocn), ip(land))
    stagthck = np.zeros((nx - 1, ny - 1), order="F")
    stagusrf = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_stagger(nx, ny, dp(p.thck), dp(stagthck))
    L.orc_stagger(nx, ny, dp(usrf), dp(stagusrf))
    dsdx = np.zeros((nx - 1, ny - 1), order="F")
    dsdy = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_surface_elevation_gradient(nx, ny, p.dx, p.dy, ip(ice), ip(land), dp(usrf), dp(p.thck),
                                     p.thklim, p.thck_gradient_ramp, p.gradient_margin, p.max_slope,
                                     dp(dsdx), dp(dsdy))
    halo(g, dsdx)
    halo(g, dsdy)
    acell = np.zeros((nx, ny), np.int32, order="F")
    avert = np.zeros((nx - 1, ny - 1), np.int32, order="F")
    L.orc_vertex_masks(nx, ny, 2, ip(ice), ip(acell), ip(avert))
    flwafact = np.zeros((nz - 1, nx, ny), order="F")
    L.orc_flwafact(nx, ny, nz, 2, p.n_glen, dp(np.asfortranarray(p.flwa)), dp(flwafact))
    staglsrf = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_stagger(nx, ny, dp(lsrf), dp(staglsrf))
    theta = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_slope_angle(nx - 1, ny - 1, p.dx, p.dy, dp(staglsrf), dp(theta))
    fground = grounded_fraction(p, ice, flt, land)
    # thickness and surface at vertices for the standard lateral BC: mean over
    # ice-covered or land cells (glissade_velo_higher.F90:1510-1515, 2006-2016)
    ipl = np.asfortranarray(((ice == 1) | (land == 1)).astype(np.int32))
    stagthck_lat = np.zeros((nx - 1, ny - 1), order="F")
    stagusrf_lat = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_stagger_masked(nx, ny, dp(p.thck), ip(ipl), dp(stagthck_lat))
    L.orc_stagger_masked(nx, ny, dp(usrf), ip(ipl), dp(stagusrf_lat))
    return dict(lsrf=lsrf, usrf=usrf, ice_mask=ice, floating_mask=flt, ocean_mask=ocn, land_mask=land,
                stagthck=stagthck, stagusrf=stagusrf, dusrf_dx=dsdx, dusrf_dy=dsdy,
                active_cell=acell, active_vertex=avert, flwafact=flwafact, theta_basal=theta,
                f_ground=fground, stagthck_lateral=stagthck_lat, stagusrf_lateral=stagusrf_lat)


def grounded_fraction(problem, ice, flt, land):
    """f_ground at the vertices for the problem's which_ho_ground (0, 1),
    which_ho_flotation_function (0-3) and which_ho_fground_no_glp (0, 1)
    (glissade_grounded_fraction, glissade_grounding_line.F90:59-674); the
    defaults (0, 2, 0) are orc_grounded_fraction_no_glp."""
    p = problem
    g = p.grid
    wg = getattr(p, "which_ho_ground", 0)
    wf = getattr(p, "which_ho_flotation_function", 2)
    wn = getattr(p, "which_ho_fground_no_glp", 0)
    if wf == 3 and getattr(p, "topg_raised", None) is None:
        # the reference computes the flotation function for every option first
        raise ValueError("which_ho_flotation_function 3 needs topg_raised")
    fground = np.zeros((g.nx - 1, g.ny - 1), order="F")
    if wg == 0 and wn == 0:
        lib().orc_grounded_fraction_no_glp(g.nx, g.ny, ip(ice), ip(flt), ip(land), dp(fground))
        return fground
    ff = np.zeros((g.nx, g.ny), order="F")
    f64 = lambda a: np.asfortranarray(np.array(a, dtype=np.float64))
    bed = p.topg_raised if wf == 3 else p.topg   # option 3: option 2 on the raised bed (:299-326)
    st = lib().orc_grounded_fraction(g.nx, g.ny, g.global_bc, wg, wf, wn, dp(f64(p.thck)), dp(f64(bed)), p.eus,
                                     ip(ice), ip(flt), ip(land), dp(ff), dp(fground))
    if st:
        raise ValueError(f"grounded fraction: status {st}")
    return fground


def internal_stress(problem, uvel, vvel, geo=None, approx=2, efvs=None):
    """compute_internal_stress (glissade_velo_higher.F90:7046-7281) of a
    3-D velocity (nz, nx-1, ny-1) for whichapprox ``approx`` (0 SIA, 1 SSA,
    2 BP, 3 L1L2 with the layer viscosity ``efvs`` (nz-1, nx, ny), 4 DIVA):
    dict of tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff, each
    (nz-1, nx, ny)."""
    p = problem
    g = p.grid
    geo = geo or geometry(p)
    xV, yV = p.vertex_coords()
    nfl = (g.nz - 1) * g.nx * g.ny
    tau = np.zeros(6 * nfl)
    if approx == 3 and efvs is None:
        raise ValueError("internal_stress: L1L2 needs the layer viscosity efvs")
    ef = None if efvs is None else np.asfortranarray(efvs, dtype=np.float64)
    st = lib().orc_compute_internal_stress_approx(
        g.nx, g.ny, g.nz, dp(np.ascontiguousarray(p.sigma, dtype=np.float64)), 2, ip(geo["active_cell"]), dp(xV),
        dp(yV), dp(geo["stagusrf"]), dp(geo["stagthck"]), int(approx), dp(geo["flwafact"]), p.whichefvs,
        None if ef is None else dp(ef), p.efvs_constant, p.effstrain_min, p.n_glen,
        dp(np.asfortranarray(uvel, dtype=np.float64)), dp(np.asfortranarray(vvel, dtype=np.float64)), dp(tau))
    if st:
        raise RuntimeError(f"oracle internal stress failed with status {st}")
    names = ("tau_xz", "tau_yz", "tau_xx", "tau_yy", "tau_xy", "tau_eff")
    return {k: np.asfortranarray(tau[q * nfl:(q + 1) * nfl].reshape((g.nz - 1, g.nx, g.ny), order="F"))
            for q, k in enumerate(names)}


def basal_friction(grid, which, active_cell, active_vertex, ubas, vbas, beta):
    """compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042)
    with its cell halo (:3807), and btractx/btracty = beta u_b (:3817-3818):
    (bfricflx (nx, ny), btractx, btracty (nx-1, ny-1))."""
    g = grid
    f64 = lambda a: np.asfortranarray(np.array(a, dtype=np.float64))
    i32 = lambda a: np.asfortranarray(np.array(a, dtype=np.int32))
    bf = np.zeros((g.nx, g.ny), order="F")
    bx = np.zeros((g.nx - 1, g.ny - 1), order="F")
    by = np.zeros((g.nx - 1, g.ny - 1), order="F")
    ub, vb, b = f64(ubas), f64(vbas), f64(beta)
    ac, av = i32(active_cell), i32(active_vertex)
    lib().orc_basal_friction(g.nx, g.ny, 2, g.global_bc, int(which), ip(ac), ip(av), dp(ub), dp(vb), dp(b), dp(bf),
                             dp(bx), dp(by))
    return bf, bx, by


def compute_grounded_fraction(f4):
    """compute_grounded_fraction (glissade_grounding_line.F90:678-1080) of the
    corner values (SW, SE, NE, NW) of one staggered cell."""
    f = np.ascontiguousarray(f4, dtype=np.float64)
    out = np.zeros(1)
    st = lib().orc_compute_grounded_fraction(dp(f), dp(out))
    if st:
        raise ValueError("compute_grounded_fraction: a pattern the reference stops on")
    return float(out[0])


def no_penetration_masks(grid):
    """staggered_no_penetration_mask (parallel_mpi.F90:10187-10231) on one
    task: u mask 1 on vertex columns local_ewn-uhalo.. and ..lhalo, v mask
    on the matching rows, then the staggered halo of both; (u, v) int32
    (nx-1, ny-1) masks, zero unless global_bc is 2."""
    g = grid
    mu = np.zeros((g.nx - 1, g.ny - 1), order="F")
    mv = np.zeros((g.nx - 1, g.ny - 1), order="F")
    if g.global_bc == 2:
        mu[g.nx - 2 - 1:, :] = 1.0        # umask(local_ewn-uhalo:, :) with uhalo = 2
        mu[:2, :] = 1.0                   # umask(:lhalo, :)
        mv[:, g.ny - 2 - 1:] = 1.0
        mv[:, :2] = 1.0
        halo(g, mu)
        halo(g, mv)
    return mu.astype(np.int32), mv.astype(np.int32)


def calcbeta(problem, geo, ubas, vbas):
    """glissade_calcbeta (glissade_basal_traction.F90:69-719) for the
    problem's friction law (cism_amd.basal.Friction) from the basal velocity
    (nx-1, ny-1): orc_calcbeta, then the staggered halo (:703).  A negative
    or NaN beta raises ValueError naming the vertex (the reference's
    GM_FATAL, :687-699)."""
    fr = problem.friction
    g = problem.grid
    f64 = lambda a: None if a is None else np.asfortranarray(np.array(a, dtype=np.float64))
    prm = np.ascontiguousarray(fr.params(problem.eus, getattr(problem, "dx", 0.0), getattr(problem, "dy", 0.0)))
    flwa_basal = f64(fr.flwa_basal if fr.flwa_basal is not None else np.asarray(problem.flwa)[g.nz - 2])
    bpmp = None if fr.bpmp_mask is None else np.asfortranarray(np.array(fr.bpmp_mask, dtype=np.int32))
    flds = [f64(getattr(fr, k)) for k in ("effecpress_stag", "coulomb_c", "powerlaw_c", "mintauf",
                                           "c_space_factor_stag")]
    bext, fg, topg, thck = f64(fr.beta_external), f64(geo["f_ground"]), f64(problem.topg), f64(problem.thck)
    u, v = f64(ubas), f64(vbas)
    b = np.zeros((g.nx - 1, g.ny - 1), order="F")
    opt = lambda a: None if a is None else dp(a)
    st = lib().orc_calcbeta(g.nx, g.ny, g.global_bc, fr.whichbabc, fr.which_beta_limit, dp(prm), dp(u), dp(v),
                            *[opt(a) for a in flds], ip(bpmp) if bpmp is not None else None, opt(bext), dp(fg),
                            opt(topg), opt(thck), opt(flwa_basal), dp(b))
    if st < 0:
        raise ValueError(f"which_ho_babc {fr.whichbabc} is not covered")
    if st > 0:
        t = st - 1
        raise ValueError(f"invalid beta {b.ravel(order='F')[t]} at vertex i={t % (g.nx - 1) + 1} "
                         f"j={t // (g.nx - 1) + 1}")
    halo(g, b)
    return b


def beta_internal(problem, geo, ubas=None, vbas=None):
    """glissade_calcbeta for the constant/large/external cases
    (glissade_basal_traction.F90:202, 304, 399), times f_ground with
    beta_grounded_min = 0 (:650-672), then the staggered halo (:697); a
    problem with a friction law: :func:`calcbeta` of the basal velocity."""
    p = problem
    g = p.grid
    if getattr(p, "friction", None) is not None:
        return calcbeta(p, geo, ubas, vbas)
    if p.whichbabc == 0:
        b = np.full((g.nx - 1, g.ny - 1), p.beta_const, order="F")
    elif p.whichbabc == 4:
        b = np.full((g.nx - 1, g.ny - 1), 1.0e10, order="F")
    else:
        b = np.asfortranarray(p.beta.copy())
    fg = geo["f_ground"]
    for j in range(g.ny - 1):
        for i in range(g.nx - 1):
            b[i, j] = b[i, j] * fg[i, j]
    halo(g, b)
    return b


def stagger(var):
    """glissade_stagger, margin 0, of a cell field (orc_stagger)."""
    nx, ny = var.shape
    out = np.zeros((nx - 1, ny - 1), order="F")
    lib().orc_stagger(nx, ny, dp(np.asfortranarray(var, dtype=np.float64)), dp(out))
    return out


def gravity_load(p, geo, xV, yV, loadu, loadv):
    """load_vector_gravity with which_ho_assemble_taud (p.assemble_taud: 1 local,
    the default, :4311-4342; 0 standard, :4350-4446)."""
    L = lib()
    g = p.grid
    av = geo["active_vertex"]
    if getattr(p, "assemble_taud", 1) == 0:
        st = L.orc_load_vector_gravity_standard(g.nx, g.ny, g.nz, 2, dp(np.ascontiguousarray(p.sigma, dtype=np.float64)),
                                                ip(geo["active_cell"]), dp(xV), dp(yV), dp(geo["stagusrf"]),
                                                dp(geo["stagthck"]), dp(geo["dusrf_dx"]), dp(geo["dusrf_dy"]),
                                                dp(loadu), dp(loadv))
        if st:
            raise RuntimeError(f"oracle standard gravity load failed with status {st}")
        return
    from cism_amd.grid import stagwbndsigma
    swb = stagwbndsigma(p.sigma)
    L.orc_load_vector_gravity_local(g.nx, g.ny, g.nz, dp(swb), p.dx, p.dy, ip(av), dp(geo["stagthck"]),
                                    dp(geo["dusrf_dx"]), dp(geo["dusrf_dy"]), dp(loadu), dp(loadv))


def basal_2d(p, geo, xV, yV, av, binternal, A):
    """basal_sliding_bc_2d with which_ho_assemble_beta (p.assemble_beta: 1 local,
    :8373-8384; 0 standard, flat Jacobian, :8386-8527)."""
    L = lib()
    g = p.grid
    if getattr(p, "assemble_beta", 1) == 0:
        st = L.orc_basal_sliding_bc_standard(g.nx, g.ny, g.nz, 2, 1, 0, ip(geo["active_cell"]), dp(xV), dp(yV),
                                             None, dp(binternal), None, dp(A[0]), dp(A[3]))
        if st:
            raise RuntimeError(f"oracle standard 2-D basal assembly failed with status {st}")
    else:
        L.orc_basal_sliding_bc_2d_local(g.nx, g.ny, p.dx, p.dy, ip(av), dp(binternal), dp(A[0]), dp(A[3]))


def assemble(problem, uvel=None, vvel=None, geo=None):
    """One assembly of the linear BP system in the reference's order
    (glissade_velo_higher.F90:2815-2931): stiffness, basal BC, b = load,
    Dirichlet rows, halo of A and b, symmetry fix."""
    L = lib()
    p = problem
    g = p.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    geo = geo or geometry(p)
    uvel = np.asfortranarray(p.uvel.copy() if uvel is None else uvel)
    vvel = np.asfortranarray(p.vvel.copy() if vvel is None else vvel)
    av = geo["active_vertex"]
    own = g.owned_slice()
    inact = (av[own[1], own[2]] == 0)
    uo = uvel[own]; vo = vvel[own]
    uo[:, inact] = 0.0; vo[:, inact] = 0.0
    uvel[own] = uo; vvel[own] = vo
    xV, yV = p.vertex_coords()
    loadu = g.zeros_vec(); loadv = g.zeros_vec()
    from cism_amd.grid import stagwbndsigma
    swb = stagwbndsigma(p.sigma)
    gravity_load(p, geo, xV, yV, loadu, loadv)
    sigma_ = np.ascontiguousarray(p.sigma, dtype=np.float64)
    L.orc_load_vector_lateral_bc(nx, ny, nz, dp(sigma_), 2, int(p.assemble_lateral), ip(geo["ocean_mask"]),
                                 ip(geo["active_cell"]), dp(xV), dp(yV), dp(geo["usrf"]), dp(p.thck),
                                 dp(geo["stagusrf_lateral"]), dp(geo["stagthck_lateral"]), dp(loadu), dp(loadv))
    A = [g.zeros_mat() for _ in range(4)]
    efvs = np.zeros((nz - 1, nx, ny), order="F")
    sigma = np.ascontiguousarray(p.sigma, dtype=np.float64)
    st = L.orc_assemble_stiffness_matrix_3d(nx, ny, nz, dp(sigma), 2, ip(geo["active_cell"]), dp(xV), dp(yV),
                                            dp(uvel), dp(vvel), dp(geo["stagusrf"]), dp(geo["stagthck"]),
                                            dp(geo["flwafact"]), p.whichefvs, p.efvs_constant,
                                            p.effstrain_min, p.n_glen, dp(efvs), *[dp(a) for a in A])
    if st:
        raise RuntimeError(f"oracle assembly failed with status {st}")
    binternal = beta_internal(p, geo, uvel[nz - 1], vvel[nz - 1])
    if p.whichbabc != 6:
        if getattr(p, "assemble_beta", 1) == 0:
            st = L.orc_basal_sliding_bc_standard(nx, ny, nz, 2, 0, 1, ip(geo["active_cell"]), dp(xV), dp(yV),
                                                 dp(stagger(geo["lsrf"])), dp(binternal), None, dp(A[0]), dp(A[3]))
            if st:
                raise RuntimeError(f"oracle standard basal assembly failed with status {st}")
        else:
            L.orc_basal_sliding_bc_local(nx, ny, nz, p.dx, p.dy, ip(av), dp(binternal), dp(geo["theta_basal"]),
                                         dp(A[0]), dp(A[3]))
    bu = loadu.copy(order="F"); bv = loadv.copy(order="F")
    umask = np.zeros(g.vec_shape, np.int32, order="F")
    if p.whichbabc == 6:
        umask[nz - 1, :, :] = 1
    kin = getattr(p, "kinbcmask", None)       # kinematic-BC columns (:1444-1455)
    if kin is not None:
        umask[:, np.asarray(kin) == 1] = 1
    vmask = umask.copy(order="F")
    npu, npv = no_penetration_masks(g)        # (:1461-1471)
    umask[:, npu == 1] = 1
    vmask[:, npv == 1] = 1
    L.orc_dirichlet_bc_3d(nx, ny, nz, 2, ip(av), ip(umask), ip(vmask), dp(uvel), dp(vvel),
                          *[dp(a) for a in A], dp(bu), dp(bv))
    for a in A:
        L.orc_halo3d(ctypes.byref(g.cstruct()), 27 * nz, dp(a))
    halo(g, bu)
    halo(g, bv)
    nwarn = L.orc_check_symmetry_assembled_3d(nx, ny, nz, g.ilo, g.ihi, g.jlo, g.jhi, ip(av), *[dp(a) for a in A])
    nnz = L.orc_count_nonzeros_3d(nx, ny, nz, g.ilo, g.ihi, g.jlo, g.jhi, ip(av), *[dp(a) for a in A])
    return dict(Auu=A[0], Auv=A[1], Avu=A[2], Avv=A[3], bu=bu, bv=bv, efvs=efvs, uvel=uvel, vvel=vvel,
                active_vertex=av, loadu=loadu, loadv=loadv, umask=umask, vmask=vmask,
                beta_internal=binternal, nwarn_asym=nwarn, nnz=nnz, geo=geo)


def element_matrix_2d_ssa(x, y, u, v, h, flwafact_2d, whichefvs=2, efvs_constant=0.0, effstrain_min=1e-8,
                          n_glen=3.0):
    L = lib()
    K = [np.zeros((4, 4)) for _ in range(4)]   # C order [nr][nc]
    em = ctypes.c_double()
    f = lambda a: dp(np.ascontiguousarray(a, dtype=np.float64))
    st = L.orc_element_matrix_2d_ssa(f(x), f(y), f(u), f(v), f(h), flwafact_2d, whichefvs, efvs_constant,
                                     effstrain_min, n_glen, *[dp(k) for k in K], ctypes.byref(em))
    return st, K, em.value


def vertical_average(sigma, var):
    """glissade_vertical_average of a layer field (nz-1, nx, ny)."""
    L = lib()
    nz = len(sigma)
    _, nx, ny = var.shape
    out = np.zeros((nx, ny), order="F")
    L.orc_vertical_average(nx, ny, nz, dp(np.ascontiguousarray(sigma, dtype=np.float64)),
                           dp(np.asfortranarray(var)), dp(out))
    return out


def geometry_diva(problem, geo):
    """DIVA inputs of glissade_velo_higher_solve with diva_slope_correction
    (glissade_velo_higher.F90:2078-2140): slope angles of the mid-column
    surface z_mean = (usrf + lsrf)/2 at cells (glissade_slope_angle, with the
    cell halo) and vertices (glissade_slope_angle_staggered), the slope
    factors cos^2 + 4 sin^2, the basal slope angles of the staggered lower
    surface used by basal_sliding_bc_2d_diva (:8591-8605), and the
    ice-or-land cell mask (:1510-1515)."""
    L = lib()
    p = problem
    g = p.grid
    nx, ny = g.nx, g.ny
    zmean = np.asfortranarray(0.5 * (geo["usrf"] + geo["lsrf"]))
    th = np.zeros((nx, ny), order="F"); thx = np.zeros_like(th); thy = np.zeros_like(th)
    L.orc_slope_angle_xy(nx, ny, p.dx, p.dy, dp(zmean), dp(th), dp(thx), dp(thy))
    L.orc_cell_halo(nx, ny, g.global_bc, dp(thx))
    L.orc_cell_halo(nx, ny, g.global_bc, dp(thy))
    sth = np.zeros((nx - 1, ny - 1), order="F"); sthx = np.zeros_like(sth); sthy = np.zeros_like(sth)
    L.orc_slope_angle_staggered(nx, ny, p.dx, p.dy, dp(zmean), dp(sth), dp(sthx), dp(sthy))
    sf = lambda t: np.asfortranarray(np.cos(t) ** 2 + 4.0 * np.sin(t) ** 2)
    staglsrf = np.zeros((nx - 1, ny - 1), order="F")
    L.orc_stagger(nx, ny, dp(geo["lsrf"]), dp(staglsrf))
    tb = np.zeros((nx - 1, ny - 1), order="F"); tbx = np.zeros_like(tb); tby = np.zeros_like(tb)
    L.orc_slope_angle_xy(nx - 1, ny - 1, p.dx, p.dy, dp(staglsrf), dp(tb), dp(tbx), dp(tby))
    ipl = np.asfortranarray(((geo["ice_mask"] == 1) | (geo["land_mask"] == 1)).astype(np.int32))
    return dict(theta_slope_x=thx, theta_slope_y=thy, slope_factor_x=sf(thx), slope_factor_y=sf(thy),
                stag_theta_slope_x=sthx, stag_theta_slope_y=sthy, stag_slope_factor_x=sf(sthx),
                stag_slope_factor_y=sf(sthy), theta_basal_x=tbx, theta_basal_y=tby, ice_plus_land_mask=ipl)


def diva_state(problem):
    """Cold-start DIVA state: efvs_qp_3d (nz-1, 4, nx, ny) and the basal
    traction (nx-1, ny-1) start at zero."""
    g = problem.grid
    return dict(efvs_qp_3d=np.zeros((g.nz - 1, 4, g.nx, g.ny), order="F"),
                btractx=np.zeros((g.nx - 1, g.ny - 1), order="F"),
                btracty=np.zeros((g.nx - 1, g.ny - 1), order="F"))


def assemble_2d(problem, uvel_2d=None, vvel_2d=None, geo=None, approx="ssa", state=None, ubas=None, vbas=None):
    """One assembly of the 2-D system in the reference's order
    (glissade_velo_higher.F90:2353-2725): stiffness (assemble_stiffness_matrix_2d),
    basal BC (basal_sliding_bc_2d; for DIVA beta_eff from omega and
    basal_sliding_bc_2d_diva), b = the level sum of the 3-D load (:2045-2058,
    :2636), Dirichlet rows with the basal-level masks (none for DIVA),
    halo of A and b, symmetry report and nonzero count.  The default 2-D
    velocity is the basal level of the 3-D guess (:1362).  approx "diva"
    updates state (efvs_qp_3d; btractx/y are read) in place.  ubas/vbas:
    the basal velocity a friction law reads (default the 2-D velocity; for
    DIVA the basal level of the last 3-D velocity, :2293-2296)."""
    L = lib()
    p = problem
    g = p.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    geo = geo or geometry(p)
    av = geo["active_vertex"]
    u2 = np.asfortranarray(p.uvel[nz - 1].copy() if uvel_2d is None else uvel_2d, dtype=np.float64)
    v2 = np.asfortranarray(p.vvel[nz - 1].copy() if vvel_2d is None else vvel_2d, dtype=np.float64)
    xV, yV = p.vertex_coords()
    # 3-D load (gravity + lateral), summed over levels in k order
    loadu = g.zeros_vec(); loadv = g.zeros_vec()
    from cism_amd.grid import stagwbndsigma
    swb = stagwbndsigma(p.sigma)
    gravity_load(p, geo, xV, yV, loadu, loadv)
    sigma = np.ascontiguousarray(p.sigma, dtype=np.float64)
    L.orc_load_vector_lateral_bc(nx, ny, nz, dp(sigma), 2, int(p.assemble_lateral), ip(geo["ocean_mask"]),
                                 ip(geo["active_cell"]), dp(xV), dp(yV), dp(geo["usrf"]), dp(p.thck),
                                 dp(geo["stagusrf_lateral"]), dp(geo["stagthck_lateral"]), dp(loadu), dp(loadv))
    lu2 = np.zeros((nx - 1, ny - 1), order="F"); lv2 = np.zeros((nx - 1, ny - 1), order="F")
    for k in range(nz):
        lu2 = lu2 + loadu[k]
        lv2 = lv2 + loadv[k]
    A = [np.zeros((nx - 1, ny - 1, 9), order="F") for _ in range(4)]
    efvs = np.zeros((nz - 1, nx, ny), order="F")
    extra = {}
    binternal = beta_internal(p, geo, u2 if ubas is None else ubas, v2 if vbas is None else vbas)
    if approx == "diva":
        dg = geo.get("diva") or geometry_diva(p, geo)
        geo["diva"] = dg
        state = state if state is not None else diva_state(p)
        stagsigma = np.ascontiguousarray((p.sigma[:-1] + p.sigma[1:]) / 2.0)
        omega_k = np.zeros((nz, nx, ny), order="F")
        omega = np.zeros((nx, ny), order="F")
        thck = np.asfortranarray(np.array(p.thck, dtype=np.float64))
        st = L.orc_assemble_stiffness_matrix_2d_diva(
            nx, ny, nz, dp(sigma), dp(stagsigma), 2, ip(geo["active_cell"]), dp(xV), dp(yV), dp(u2), dp(v2),
            dp(geo["stagthck"]), dp(thck), dp(geo["flwafact"]), dp(state["btractx"]), dp(state["btracty"]),
            dp(dg["slope_factor_x"]), dp(dg["slope_factor_y"]), p.whichefvs, p.efvs_constant, p.effstrain_min,
            p.n_glen, dp(state["efvs_qp_3d"]), dp(efvs), dp(omega_k), dp(omega), *[dp(a) for a in A])
        if st:
            raise RuntimeError(f"oracle 2-D DIVA assembly failed with status {st}")
        L.orc_cell_halo(nx, ny, g.global_bc, dp(omega))
        stag_omega = np.zeros((nx - 1, ny - 1), order="F")
        L.orc_stagger_masked(nx, ny, dp(omega), ip(dg["ice_plus_land_mask"]), dp(stag_omega))
        bex = np.zeros((nx - 1, ny - 1), order="F"); bey = np.zeros_like(bex)
        L.orc_diva_beta_eff(nx, ny, 1 if p.whichbabc == 6 else 0, dp(stag_omega), dp(binternal),
                            dp(dg["stag_slope_factor_x"]), dp(dg["stag_slope_factor_y"]),
                            dp(dg["stag_theta_slope_x"]), dp(dg["stag_theta_slope_y"]), dp(bex), dp(bey))
        if getattr(p, "assemble_beta", 1) == 0:     # standard option, curved Jacobian (:8535-8811)
            st = L.orc_basal_sliding_bc_standard(nx, ny, nz, 2, 1, 1, ip(geo["active_cell"]), dp(xV), dp(yV),
                                                 dp(stagger(geo["lsrf"])), dp(bex), dp(bey), dp(A[0]), dp(A[3]))
            if st:
                raise RuntimeError(f"oracle standard DIVA basal assembly failed with status {st}")
        else:
            L.orc_basal_sliding_bc_2d_diva_local(nx, ny, p.dx, p.dy, ip(av), dp(bex), dp(bey), dp(dg["theta_basal_x"]),
                                                 dp(dg["theta_basal_y"]), dp(A[0]), dp(A[3]))
        extra = dict(omega_k=omega_k, omega=omega, stag_omega=stag_omega, beta_eff_x=bex, beta_eff_y=bey,
                     state=state)
    elif approx == "l1l2":
        flwa = np.asfortranarray(np.array(p.flwa, dtype=np.float64))
        st = L.orc_assemble_stiffness_matrix_2d_l1l2(nx, ny, nz, dp(sigma), 2, ip(geo["active_cell"]), dp(xV), dp(yV),
                                                     dp(u2), dp(v2), dp(geo["stagthck"]), dp(geo["dusrf_dx"]),
                                                     dp(geo["dusrf_dy"]), dp(flwa), dp(geo["flwafact"]), p.whichefvs,
                                                     p.efvs_constant, p.effstrain_min, p.n_glen, dp(efvs),
                                                     *[dp(a) for a in A])
        if st:
            raise RuntimeError(f"oracle 2-D L1L2 assembly failed with status {st}")
        if p.whichbabc != 6:
            basal_2d(p, geo, xV, yV, av, binternal, A)
    else:
        st = L.orc_assemble_stiffness_matrix_2d_ssa(nx, ny, nz, dp(sigma), 2, ip(geo["active_cell"]), dp(xV),
                                                    dp(yV), dp(u2), dp(v2), dp(geo["stagthck"]),
                                                    dp(geo["flwafact"]), p.whichefvs, p.efvs_constant,
                                                    p.effstrain_min, p.n_glen, dp(efvs), *[dp(a) for a in A])
        if st:
            raise RuntimeError(f"oracle 2-D assembly failed with status {st}")
        if p.whichbabc != 6:
            basal_2d(p, geo, xV, yV, av, binternal, A)
    bu = np.asfortranarray(lu2.copy()); bv = np.asfortranarray(lv2.copy())
    umask = np.zeros((nx - 1, ny - 1), np.int32, order="F")
    if p.whichbabc == 6 and approx != "diva":     # DIVA: no-slip through beta_eff (:1423-1428)
        umask[:, :] = 1
    kin = getattr(p, "kinbcmask", None)       # kinematic-BC columns (:1444-1455)
    if kin is not None:
        umask[np.asarray(kin) == 1] = 1
    vmask = umask.copy(order="F")
    npu, npv = no_penetration_masks(g)        # (:1461-1471)
    umask[npu == 1] = 1
    vmask[npv == 1] = 1
    L.orc_dirichlet_bc_2d(nx, ny, 2, ip(av), ip(umask), ip(vmask), dp(u2), dp(v2), *[dp(a) for a in A],
                          dp(bu), dp(bv))
    for a in A:
        for m in range(9):
            sl = a[:, :, m]
            halo(g, sl)
    halo(g, bu)
    halo(g, bv)
    nwarn = L.orc_check_symmetry_assembled_2d(nx, ny, g.ilo, g.ihi, g.jlo, g.jhi, ip(av), *[dp(a) for a in A])
    nnz = L.orc_count_nonzeros_2d(nx, ny, g.ilo, g.ihi, g.jlo, g.jhi, ip(av), *[dp(a) for a in A])
    return dict(Auu=A[0], Auv=A[1], Avu=A[2], Avv=A[3], bu=bu, bv=bv, efvs=efvs, uvel_2d=u2, vvel_2d=v2,
                active_vertex=av, loadu_2d=lu2, loadv_2d=lv2, beta_internal=binternal, nwarn_asym=nwarn,
                nnz=nnz, geo=geo, **extra)


def diva_velocity(problem, s, u2, v2, beta=None):
    """compute_3d_velocity_diva after a 2-D DIVA solve: updates the basal
    traction in s["state"] and returns the 3-D velocity (nz, nx-1, ny-1).
    beta: the current beta_internal (default the one of assembly s; after a
    restored assembly the reference keeps the last calcbeta result)."""
    L = lib()
    p = problem
    g = p.grid
    dg = s["geo"]["diva"]
    u3 = g.zeros_vec(); v3 = g.zeros_vec()
    st = s["state"]
    L.orc_compute_3d_velocity_diva(g.nx, g.ny, g.nz, ip(s["active_vertex"]), ip(dg["ice_plus_land_mask"]),
                                   dp(s["omega_k"]), dp(s["beta_internal"] if beta is None else beta),
                                   dp(s["beta_eff_x"]), dp(s["beta_eff_y"]),
                                   dp(dg["stag_theta_slope_x"]), dp(dg["stag_theta_slope_y"]),
                                   dp(dg["stag_slope_factor_x"]), dp(dg["stag_slope_factor_y"]), dp(u2), dp(v2),
                                   dp(st["btractx"]), dp(st["btracty"]), dp(u3), dp(v3))
    return u3, v3


def l1l2_velocity(problem, s, u2, v2):
    """compute_3d_velocity_L1L2 after a 2-D L1L2 solve (glissade_velo_higher.F90:
    3698-3724): the basal level is the 2-D velocity, the levels above follow
    from the layer viscosities of the last assembly s; returns the 3-D
    velocity (nz, nx-1, ny-1) before its halo."""
    L = lib()
    p = problem
    g = p.grid
    geo = s["geo"]
    u3 = g.zeros_vec(); v3 = g.zeros_vec()
    u3[g.nz - 1] = u2
    v3[g.nz - 1] = v2
    mask = np.zeros((g.nx - 1, g.ny - 1), np.int32, order="F")
    if p.whichbabc == 6:
        mask[:, :] = 1
    xV, yV = p.vertex_coords()
    flwa = np.asfortranarray(np.array(p.flwa, dtype=np.float64))
    thck = np.asfortranarray(np.array(p.thck, dtype=np.float64))
    st = L.orc_compute_3d_velocity_l1l2(g.nx, g.ny, g.nz, g.ilo, g.ihi, g.jlo, g.jhi, g.global_bc, 2,
                                        dp(np.ascontiguousarray(p.sigma, dtype=np.float64)), p.dx, p.dy,
                                        ip(geo["ice_mask"]), ip(geo["active_cell"]), ip(s["active_vertex"]), ip(mask),
                                        ip(mask), dp(xV), dp(yV), dp(thck), dp(geo["stagthck"]), dp(geo["dusrf_dx"]),
                                        dp(geo["dusrf_dy"]), dp(flwa), dp(s["efvs"]), p.whichefvs, p.n_glen, dp(u3),
                                        dp(v3))
    if st:
        raise RuntimeError(f"oracle L1L2 3-D velocity failed with status {st}")
    return u3, v3


def residual_L2(grid, active_vertex, A, bu, bv, uvel, vvel):
    """compute_residual_vector_3d (glissade_velo_higher.F90:9693-9894): L2 of A u - b
    over owned active vertices, and relative to |b|."""
    yu, yv = matvec(grid, active_vertex, *A, uvel, vvel)
    own = grid.owned_slice()
    act = np.asarray(active_vertex, bool)[own[1], own[2]][None]
    ru = np.where(act, yu[own] - bu[own], 0.0)
    rv = np.where(act, yv[own] - bv[own], 0.0)
    wu = np.zeros(grid.vec_shape, order="F"); wv = np.zeros(grid.vec_shape, order="F")
    wu[own] = ru * ru; wv[own] = rv * rv
    L2 = np.sqrt(global_sum(grid, wu, wv))
    bu2 = np.zeros(grid.vec_shape, order="F"); bv2 = np.zeros(grid.vec_shape, order="F")
    bu2[own] = np.where(act, bu[own] * bu[own], 0.0); bv2[own] = np.where(act, bv[own] * bv[own], 0.0)
    rhs = np.sqrt(global_sum(grid, bu2, bv2))
    return L2, (L2 / rhs if rhs > 0.0 else 0.0)      # (:9881-9885)


def residual_L2_2d(grid, active_vertex, A, bu, bv, uvel, vvel):
    """compute_residual_vector_2d (glissade_velo_higher.F90:9502-9689): L2 of
    A u - b over owned active vertices, and relative to |b|."""
    import dataclasses
    g1 = dataclasses.replace(grid, nz=1)
    yu, yv = matvec_2d(grid, active_vertex, *A, uvel, vvel)
    il, jl = slice(grid.ilo - 1, grid.ihi), slice(grid.jlo - 1, grid.jhi)
    act = np.asarray(active_vertex, bool)[il, jl]
    w = [np.zeros((1, grid.nvx, grid.nvy), order="F") for _ in range(4)]
    w[0][0, il, jl] = np.where(act, (yu - bu)[il, jl] ** 2, 0.0)
    w[1][0, il, jl] = np.where(act, (yv - bv)[il, jl] ** 2, 0.0)
    w[2][0, il, jl] = np.where(act, bu[il, jl] * bu[il, jl], 0.0)
    w[3][0, il, jl] = np.where(act, bv[il, jl] * bv[il, jl], 0.0)
    L2 = np.sqrt(global_sum(g1, w[0], w[1]))
    rhs = np.sqrt(global_sum(g1, w[2], w[3]))
    return L2, (L2 / rhs if rhs > 0.0 else 0.0)


def picard(problem, *, maxiter=100, resid_target=1e-4, whichresid=0, algo="chrongear", precond=2, ncheck=5,
           linear_tol=1e-8, linear_maxiters=200, accelerate=False, two_d=False, approx="ssa"):
    """Outer Picard loop of glissade_velo_higher_solve for the BP path
    (glissade_velo_higher.F90:2163-3663), which_ho_nonlinear = 0, or with
    Picard acceleration (:2186-2263, :3108-3176, :3624-3630;
    evaluate_accelerated_picard_3d :10052-10184).  two_d: the SSA branch
    (solve_2d; assemble_2d, the 2-D solvers with which_ho_precond 2 taken as
    1, compute_residual_velocity_2d), started from the basal level of the
    3-D guess (:1362).  approx "l1l2" (two_d): the L1L2 viscosities, and
    compute_3d_velocity_L1L2 after the loop (:3698-3724).  approx "diva" (two_d): DIVA assembly with its state
    and compute_3d_velocity_diva after each solve (:3566-3590) and after each
    accelerated extension (:2188-2210), beta_eff and omega_k kept with the
    saved assembly (:9898-10048); the result adds the 3-D velocity (uvel3d, vvel3d)."""
    L = lib()
    p = problem
    g = p.grid
    geo = geometry(p)
    diva = two_d and approx == "diva"
    dstate = diva_state(p) if diva else None
    u3 = v3 = None
    if two_d:
        u = np.asfortranarray(p.uvel[g.nz - 1].copy())
        v = np.asfortranarray(p.vvel[g.nz - 1].copy())
        il, jl = slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi)
        inact = geo["active_vertex"][il, jl] == 0
        for a in (u, v):
            o = a[il, jl]; o[inact] = 0.0; a[il, jl] = o
        if precond == 2:
            precond = 1
    else:
        u = np.asfortranarray(p.uvel.copy())
        v = np.asfortranarray(p.vvel.copy())
        own = g.owned_slice()
        inact = geo["active_vertex"][own[1], own[2]] == 0
        for a in (u, v):
            o = a[own]; o[:, inact] = 0.0; a[own] = o
    npu, npv = no_penetration_masks(g)         # u, v = 0 on the no-penetration walls (:1461-1471)
    u[..., npu == 1] = 0.0
    v[..., npv == 1] = 0.0
    hist = dict(L2=[], resid_velo=[], linear_iters=[], linear_err=[], alpha=[])
    L2_large, gamma, alpha_max, thresh = 1.0e20, 0.40, 3.0, 0.85        # (:831, :1049-1056)
    if accelerate:
        u_old, v_old = u.copy(order="F"), v.copy(order="F")
        du, dv = np.zeros_like(u), np.zeros_like(v)
        alpha, L2_sav = 1.0, L2_large
        saved = None
    criterion = np.inf
    counter = 0
    s = None
    ub = vb = None          # DIVA: basal level of the last 3-D velocity (None: the 2-D guess)
    beta_cur = None         # DIVA: beta_internal of the last assembly (kept across a restore)
    while criterion >= resid_target and counter < maxiter:
        counter += 1
        done = False
        while not done:
            if accelerate and counter >= 2:
                u = np.asfortranarray(u_old + alpha * du)
                v = np.asfortranarray(v_old + alpha * dv)
                halo(g, u); halo(g, v)
                if diva:     # basal traction of the extended velocity (:2188-2210)
                    e3u, e3v = diva_velocity(p, s, u, v, beta=beta_cur)
                    ub, vb = e3u[g.nz - 1].copy(order="F"), e3v[g.nz - 1].copy(order="F")
            usav, vsav = u.copy(order="F"), v.copy(order="F")
            s = (assemble_2d(p, u, v, geo, approx=approx, state=dstate, ubas=ub, vbas=vb) if two_d
                 else assemble(p, u, v, geo))
            beta_cur = s.get("beta_internal")
            A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
            L2, L2r = (residual_L2_2d if two_d else residual_L2)(g, s["active_vertex"], A, s["bu"], s["bv"], u, v)
            hist["L2"].append(L2)
            if accelerate and counter >= 2:
                if L2 < thresh * L2_sav and alpha + gamma <= alpha_max:
                    saved = (u.copy(order="F"), v.copy(order="F"), s)
                    alpha = alpha + gamma
                    L2_sav = L2
                elif L2 < L2_sav:
                    u_old, v_old = u.copy(order="F"), v.copy(order="F")
                    alpha, L2_sav, done = 1.0, L2_large, True
                else:
                    u, v, s = saved[0].copy(order="F"), saved[1].copy(order="F"), saved[2]
                    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
                    u_old, v_old = u.copy(order="F"), v.copy(order="F")
                    alpha, L2_sav, done = 1.0, L2_large, True
                hist["alpha"].append(alpha)
            else:
                done = True
        st, err, nit = (pcg_solve_2d if two_d else pcg_solve)(g, s["active_vertex"], *A, s["bu"], s["bv"], u, v,
                                                              algo=algo, precond=precond, ncheck=ncheck,
                                                              tolerance=linear_tol, maxiters=linear_maxiters)
        if st:
            raise RuntimeError(f"oracle solver failed with status {st}")
        hist["linear_iters"].append(nit); hist["linear_err"].append(err)
        halo(g, u); halo(g, v)
        if diva:
            u3, v3 = diva_velocity(p, s, u, v, beta=beta_cur)
            halo(g, u3); halo(g, v3)
            ub, vb = u3[g.nz - 1].copy(order="F"), v3[g.nz - 1].copy(order="F")
        if two_d:
            rv = L.orc_residual_velocity_2d(g.nx, g.ny, g.ilo, g.ihi, g.jlo, g.jhi, whichresid,
                                            dp(usav), dp(vsav), dp(u), dp(v))
        else:
            rv = L.orc_residual_velocity_3d(g.nx, g.ny, g.nz, g.ilo, g.ihi, g.jlo, g.jhi, whichresid,
                                            dp(usav), dp(vsav), dp(u), dp(v))
        hist["resid_velo"].append(rv)
        if accelerate:
            du, dv = u - u_old, v - v_old
        criterion = L2 if whichresid == 3 else (L2r if whichresid == 4 else rv)
    if two_d and approx == "l1l2" and s is not None:
        u3, v3 = l1l2_velocity(p, s, u, v)
    return dict(uvel=u, vvel=v, efvs=s["efvs"] if s else None, iterations=counter,
                converged=criterion < resid_target, uvel3d=u3, vvel3d=v3, state=dstate, **hist)
