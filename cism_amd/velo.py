"""3-D Blatter-Pattyn velocity solve on the GPU.

Mirrors the 3-D BP branch of ``glissade_velo_higher_solve``
(glissade_velo_higher.F90:662-3918): host-side 2-D setup (masks, staggered
geometry, surface gradient, flow factor, beta), device assembly of the
linear system (stiffness, basal friction, gravity load, Dirichlet rows,
symmetry fix), device Krylov solve and the outer Picard loop on the
effective viscosity (:2163-3663) with the residual measures of
compute_residual_velocity_3d (:10314-10446) and compute_residual_vector_3d
(:9693-9894).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib, basal, geometry as ge
from .grid import stagwbndsigma
from .pcg import ALGO_CHRONGEAR, FP64, HO_PRECOND_SIA, System


class AssemblyIn(ctypes.Structure):
    _P = ctypes.c_void_p
    _fields_ = [("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("sigma", _P), ("stagwbndsigma", _P), ("active_cell", _P), ("active_vertex", _P),
                ("xVertex", _P), ("yVertex", _P), ("stagusrf", _P), ("stagthck", _P),
                ("dusrf_dx", _P), ("dusrf_dy", _P), ("flwafact", _P), ("beta_internal", _P),
                ("theta_basal", _P), ("umask_dirichlet", _P), ("vmask_dirichlet", _P),
                ("whichefvs", ctypes.c_int32), ("efvs_constant", ctypes.c_double),
                ("effstrain_min", ctypes.c_double), ("n_glen", ctypes.c_double),
                ("check_symmetry", ctypes.c_int32),
                ("ocean_mask", _P), ("usrf", _P), ("thck", _P), ("stagusrf_lateral", _P), ("stagthck_lateral", _P),
                ("which_assemble_lateral", ctypes.c_int32), ("flwa", _P), ("stag_lsrf", _P),
                ("assemble_beta_standard", ctypes.c_int32), ("assemble_taud_standard", ctypes.c_int32),
                ("beta_device", ctypes.c_int32)]


class DivaIn(ctypes.Structure):
    _P = ctypes.c_void_p
    _fields_ = [("slope_factor_x", _P), ("slope_factor_y", _P), ("stag_slope_factor_x", _P),
                ("stag_slope_factor_y", _P), ("stag_theta_slope_x", _P), ("stag_theta_slope_y", _P),
                ("theta_basal_x", _P), ("theta_basal_y", _P), ("ice_plus_land_mask", _P), ("efvs_qp_3d", _P),
                ("btractx", _P), ("btracty", _P), ("no_slip", ctypes.c_int32)]


class AssemblyInfo(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int64), ("asym_pairs", ctypes.c_int64), ("assemble_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double)]


def _declare(L):
    if getattr(L, "_velo_declared", False):
        return
    L.cismgpu_system_assemble_3d.argtypes = [ctypes.c_void_p, ctypes.POINTER(AssemblyIn), ctypes.POINTER(AssemblyInfo)]
    L.cismgpu_system_assemble_3d.restype = ctypes.c_int32
    L.cismgpu_system_assemble_2d.argtypes = [ctypes.c_void_p, ctypes.POINTER(AssemblyIn), ctypes.c_int32,
                                             ctypes.POINTER(AssemblyInfo)]
    L.cismgpu_system_assemble_2d.restype = ctypes.c_int32
    D = ctypes.POINTER(ctypes.c_double)
    L.cismgpu_system_get_efvs.argtypes = [ctypes.c_void_p, D]
    L.cismgpu_system_get_efvs.restype = ctypes.c_int32
    L.cismgpu_system_residual.argtypes = [ctypes.c_void_p, D, D, D, D]
    L.cismgpu_system_residual.restype = ctypes.c_int32
    L.cismgpu_system_halo_x.argtypes = [ctypes.c_void_p]
    L.cismgpu_system_halo_x.restype = ctypes.c_int32
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    L.cismgpu_system_save_assembly.argtypes = [vp]
    L.cismgpu_system_restore_assembly.argtypes = [vp]
    L.cismgpu_system_work_vectors.argtypes = [vp, i32]
    L.cismgpu_system_vector_op.argtypes = [vp, i32, i32, i32, i32, ctypes.c_double]
    L.cismgpu_system_residual_velocity.argtypes = [vp, i32, i32, D]
    L.cismgpu_system_diva_setup.argtypes = [vp, ctypes.POINTER(DivaIn)]
    L.cismgpu_system_diva_velocity.argtypes = [vp, D, D, D, D]
    L.cismgpu_system_l1l2_velocity.argtypes = [vp, ctypes.c_void_p, D, D]
    for n in ("cismgpu_system_save_assembly", "cismgpu_system_restore_assembly", "cismgpu_system_work_vectors",
              "cismgpu_system_vector_op", "cismgpu_system_residual_velocity", "cismgpu_system_diva_setup",
              "cismgpu_system_diva_velocity", "cismgpu_system_l1l2_velocity"):
        getattr(L, n).restype = ctypes.c_int32
    L._velo_declared = True


@dataclass
class PicardResult:
    uvel: np.ndarray
    vvel: np.ndarray
    efvs: np.ndarray
    iterations: int
    converged: bool
    resid_velo: list = field(default_factory=list)
    L2_norm: list = field(default_factory=list)
    linear_iters: list = field(default_factory=list)
    linear_err: list = field(default_factory=list)
    linear_status: list = field(default_factory=list)   # 0, or CISMGPU_ERR_NOT_CONVERGED per linear solve
    linear_outer: list = field(default_factory=list)    # mixed solves: refinement steps per linear solve
    assemble_ms: list = field(default_factory=list)
    solve_ms: list = field(default_factory=list)
    alpha: list = field(default_factory=list)      # accelerated Picard: alpha after each evaluation


NOT_CONVERGED = 12      # CISMGPU_ERR_NOT_CONVERGED: mixed refinement stopped before its tolerance

# accelerated Picard constants (glissade_velo_higher.F90:831, 1049-1056)
L2_NORM_LARGE = 1.0e20
GAMMA_ACCEL = 0.40
ALPHA_ACCEL_MAX = 3.0
RESID_REDUCTION_THRESHOLD = 0.85


class BPSolver:
    """Device-resident BP velocity solver for one :class:`~cism_amd.synthetic.Problem`."""

    def __init__(self, problem, system: System | None = None):
        self.p = problem
        self.grid = problem.grid
        self.sys = system or System(self.grid)
        self._L = _lib.load()
        _declare(self._L)
        self._setup()

    # ---------------------------------------------------------- host setup
    def _setup(self):
        p, g = self.p, self.grid
        self.lsrf = ge.calclsrf(p.thck, p.topg, p.eus)
        self.usrf = np.asfortranarray(np.maximum(0.0, p.thck + self.lsrf))
        self.ice_mask, self.floating_mask, self.ocean_mask, self.land_mask = ge.get_masks(
            p.thck, p.topg, p.eus, p.thklim)
        self.stagthck = ge.stagger(p.thck)
        self.stagusrf = ge.stagger(self.usrf)
        self.dusrf_dx, self.dusrf_dy = ge.surface_elevation_gradient(
            p.dx, p.dy, self.ice_mask, self.land_mask, self.usrf, p.thck, p.thklim, p.thck_gradient_ramp,
            p.gradient_margin, p.max_slope)
        ge.staggered_halo_2d(g, self.dusrf_dx)
        ge.staggered_halo_2d(g, self.dusrf_dy)
        self.active_cell, self.active_vertex = ge.vertex_masks(self.ice_mask)
        self.flwafact = ge.flwafact(np.asfortranarray(p.flwa), p.n_glen)
        self.theta_basal = ge.slope_angle(p.dx, p.dy, ge.stagger(self.lsrf))
        self.beta_internal = ge.beta_internal(p, self.ice_mask, self.floating_mask, self.land_mask)
        # velocity-dependent basal traction: glissade_calcbeta on the device before
        # every assembly (glissade_velo_higher.F90:2316-2332)
        self.friction = getattr(p, "friction", None)
        if self.friction is not None:
            if p.whichbabc == 6:
                raise ValueError("friction: a no-slip problem (which_ho_babc 6) has no basal friction term")
            if self.friction.whichbabc not in basal.DEVICE_OPTIONS:
                raise ValueError(f"friction: which_ho_babc {self.friction.whichbabc} is not a device sliding law")
            fground = ge.problem_f_ground(p, self.ice_mask, self.floating_mask, self.land_mask)
            basal.device_setup(self.sys, self.friction, f_ground=fground, topg=p.topg, thck=p.thck, eus=p.eus,
                               flwa_basal=np.asarray(p.flwa)[g.nz - 2], dew=p.dx, dns=p.dy)
        self.xV, self.yV = p.vertex_coords()
        self.sigma = np.ascontiguousarray(p.sigma, dtype=np.float64)
        self.swbs = np.ascontiguousarray(stagwbndsigma(p.sigma))
        no_slip = p.whichbabc == 6 and self.friction is None
        # Dirichlet masks (glissade_velo_higher.F90:1418-1474): the bed for
        # no-slip, whole columns where kinbcmask = 1 (u and v prescribed from
        # the input velocity)
        self.umask = np.zeros(g.vec_shape, np.int32, order="F")
        if no_slip:
            self.umask[g.nz - 1] = 1
        self.kinbc = self._kinbc_columns()
        if self.kinbc is not None:
            self.umask[:, self.kinbc] = 1
        self.vmask = self.umask.copy(order="F")
        self.nopen_u, self.nopen_v = self._no_penetration()
        if self.nopen_u is not None:
            self.umask[:, self.nopen_u] = 1
            self.vmask[:, self.nopen_v] = 1
        has_mask = bool(self.umask.any() or self.vmask.any())
        # lateral BC at calving fronts: staggered thickness/surface over ice-covered or
        # land cells (glissade_velo_higher.F90:1510-1515, 2006-2016)
        self.thck = np.asfortranarray(np.array(p.thck, dtype=np.float64))
        ipl = ((self.ice_mask == 1) | (self.land_mask == 1)).astype(np.int32)
        self.stagthck_lat = ge.stagger_masked(self.thck, ipl)
        self.stagusrf_lat = ge.stagger_masked(self.usrf, ipl)
        has_front = bool((self.ocean_mask == 1).any())
        self._ain = AssemblyIn(
            p.dx, p.dy, self.sigma.ctypes.data, self.swbs.ctypes.data, self.active_cell.ctypes.data,
            self.active_vertex.ctypes.data, self.xV.ctypes.data, self.yV.ctypes.data, self.stagusrf.ctypes.data,
            self.stagthck.ctypes.data, self.dusrf_dx.ctypes.data, self.dusrf_dy.ctypes.data,
            self.flwafact.ctypes.data, None if no_slip else self.beta_internal.ctypes.data,
            self.theta_basal.ctypes.data, self.umask.ctypes.data if has_mask else None,
            self.vmask.ctypes.data if has_mask else None, p.whichefvs, p.efvs_constant, p.effstrain_min,
            p.n_glen, 1,
            self.ocean_mask.ctypes.data if has_front else None, self.usrf.ctypes.data, self.thck.ctypes.data,
            self.stagusrf_lat.ctypes.data, self.stagthck_lat.ctypes.data, int(p.assemble_lateral))
        # quadrature (standard) options of the basal term and the gravity load
        # (which_ho_assemble_beta / _taud = 0; glissade_velo_higher.F90:8955-9111, 4350-4446)
        self.stag_lsrf = ge.stagger(self.lsrf)
        self._ain.stag_lsrf = self.stag_lsrf.ctypes.data
        self._ain.assemble_beta_standard = 1 if getattr(p, "assemble_beta", 1) == 0 else 0
        self._ain.assemble_taud_standard = 1 if getattr(p, "assemble_taud", 1) == 0 else 0
        if self.friction is not None:
            self._ain.beta_internal = None
            self._ain.beta_device = 1

    def _no_penetration(self):
        """No-penetration global BC (global_bc 2): u = 0 on the vertex column
        of the east/west global boundary and v = 0 on the row of the
        north/south one, i.e. the last owned column/row and its periodic
        image in the halo (staggered_no_penetration_mask, parallel_mpi.F90:
        10187-10231, applied at glissade_velo_higher.F90:1461-1471).
        Boolean (nx-1, ny-1) masks for u and v, or (None, None)."""
        g = self.grid
        if g.global_bc != 2:
            return None, None
        mu = np.zeros((g.nvx, g.nvy), bool, order="F")
        mv = np.zeros((g.nvx, g.nvy), bool, order="F")
        mu[[1, g.nvx - 2], :] = True
        mv[:, [1, g.nvy - 2]] = True
        return mu, mv

    def _apply_no_penetration(self, u, v):
        """Zero u and v where the no-penetration BC prescribes them (3-D or 2-D arrays)."""
        if self.nopen_u is None:
            return
        if u.ndim == 3:
            u[:, self.nopen_u] = 0.0
            v[:, self.nopen_v] = 0.0
        else:
            u[self.nopen_u] = 0.0
            v[self.nopen_v] = 0.0

    def _kinbc_columns(self):
        """Boolean (nx-1, ny-1) of the kinematic-BC columns (kinbcmask = 1), or None."""
        kin = getattr(self.p, "kinbcmask", None)
        if kin is None:
            return None
        k = np.asarray(kin)
        if k.shape != (self.grid.nvx, self.grid.nvy):
            raise ValueError("kinbcmask must be (nx-1, ny-1)")
        return (k == 1) if (k == 1).any() else None

    def _initial_velocity(self):
        u = np.asfortranarray(np.array(self.p.uvel, dtype=np.float64))
        v = np.asfortranarray(np.array(self.p.vvel, dtype=np.float64))
        return u, v

    def _zeros(self):
        return self.grid.zeros_vec()

    def _after_solve(self):
        """Work after each linear solve and the halo of x (none for BP and SSA)."""

    def _after_extend(self):
        """Work after the accelerated-Picard extension of x (none for BP and SSA)."""

    def zero_inactive(self, uvel, vvel):
        """Velocities at inactive owned vertices are zeroed (glissade_velo_higher.F90:1636-1647)."""
        g = self.grid
        own = g.owned_slice()
        inact = self.active_vertex[own[1], own[2]] == 0
        for a in (uvel, vvel):
            o = a[own]
            o[:, inact] = 0.0
            a[own] = o

    # ---------------------------------------------------------- device steps
    def calcbeta(self, want: bool = False):
        """glissade_calcbeta from the basal velocity of the resident x (a
        problem with a friction law; the assemblies call it); ``want``:
        return beta (nx-1, ny-1)."""
        if self.friction is None:
            raise ValueError("calcbeta: the problem has no friction law")
        return basal.device_calcbeta(self.sys, want=want)

    def assemble(self, uvel=None, vvel=None):
        """Assemble A, b for the viscosity of (uvel, vvel); None keeps the resident x."""
        if uvel is not None:
            u = np.asfortranarray(np.array(uvel, dtype=np.float64))
            v = np.asfortranarray(np.array(vvel, dtype=np.float64))
            self.zero_inactive(u, v)
            self._apply_no_penetration(u, v)
            self.sys.set_x(u, v)
        if self.friction is not None:
            basal.device_calcbeta(self.sys)
        info = AssemblyInfo()
        _lib.check(self._L.cismgpu_system_assemble_3d(self.sys.handle, ctypes.byref(self._ain), ctypes.byref(info)),
                   "assemble_3d")
        return dict(nnz=info.nnz, asym_pairs=info.asym_pairs, assemble_ms=info.assemble_ms, total_ms=info.total_ms)

    def internal_stress(self):
        """compute_internal_stress (glissade_velo_higher.F90:7046-7281) of the
        resident x with the last assembly's geometry: dict of tau_xz, tau_yz,
        tau_xx, tau_yy, tau_xy, tau_eff, each (nz-1, nx, ny) in Pa."""
        g = self.grid
        names = ("tau_xz", "tau_yz", "tau_xx", "tau_yy", "tau_xy", "tau_eff")
        out = {k: np.zeros((g.nz - 1, g.nx, g.ny), order="F") for k in names}
        D = ctypes.POINTER(ctypes.c_double)
        f = self._L.cismgpu_system_internal_stress
        f.argtypes = [ctypes.c_void_p] + [D] * 6
        f.restype = ctypes.c_int32
        _lib.check(f(self.sys.handle, *[out[k].ctypes.data_as(D) for k in names]), "internal_stress")
        return out

    def basal_friction(self, which_bfric=None, ubas=None, vbas=None, beta=None):
        """compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042)
        with its cell halo, and the basal traction beta u_b (:3817-3818):
        (bfricflx (nx, ny) W/m^2, btractx, btracty (nx-1, ny-1) Pa).
        which_bfric: which_ho_assemble_bfric (default: the problem's, else 1,
        local).  ubas/vbas/beta: None for the system's basal velocity and the
        last assembly's beta."""
        g = self.grid
        if which_bfric is None:
            which_bfric = getattr(self.p, "assemble_bfric", 1)
        if (ubas is None) != (vbas is None):
            raise ValueError("basal_friction: pass both basal velocity components or neither")
        D = ctypes.POINTER(ctypes.c_double)
        f64 = lambda a: None if a is None else np.asfortranarray(np.array(a, dtype=np.float64))
        ub, vb, b = f64(ubas), f64(vbas), f64(beta)
        for a in (ub, vb, b):
            if a is not None and a.shape != (g.nx - 1, g.ny - 1):
                raise ValueError(f"basal_friction: vertex fields are ({g.nx - 1}, {g.ny - 1}), got {a.shape}")
        bf = np.zeros((g.nx, g.ny), order="F")
        bx = np.zeros((g.nx - 1, g.ny - 1), order="F")
        by = np.zeros((g.nx - 1, g.ny - 1), order="F")
        f = self._L.cismgpu_system_basal_friction
        f.argtypes = [ctypes.c_void_p, ctypes.c_int32] + [D] * 6
        f.restype = ctypes.c_int32
        ptr = lambda a: None if a is None else a.ctypes.data_as(D)
        _lib.check(f(self.sys.handle, int(which_bfric), ptr(ub), ptr(vb), ptr(b), ptr(bf), ptr(bx), ptr(by)),
                   "basal_friction")
        return bf, bx, by

    def efvs(self):
        g = self.grid
        out = np.zeros((g.nz - 1, g.nx, g.ny), order="F")
        _lib.check(self._L.cismgpu_system_get_efvs(self.sys.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "get_efvs")
        return out

    def residual(self, want_vectors=False):
        g = self.grid
        D = ctypes.POINTER(ctypes.c_double)
        L2 = ctypes.c_double()
        L2r = ctypes.c_double()
        ru = rv = None
        if want_vectors:
            ru, rv = self._zeros(), self._zeros()
        _lib.check(self._L.cismgpu_system_residual(self.sys.handle, ru.ctypes.data_as(D) if ru is not None else None,
                                                   rv.ctypes.data_as(D) if rv is not None else None,
                                                   ctypes.byref(L2), ctypes.byref(L2r)), "residual")
        return (L2.value, L2r.value, ru, rv) if want_vectors else (L2.value, L2r.value)

    # ---------------------------------------------------------- Picard loop
    def save_assembly(self):
        _lib.check(self._L.cismgpu_system_save_assembly(self.sys.handle), "save_assembly")

    def restore_assembly(self):
        _lib.check(self._L.cismgpu_system_restore_assembly(self.sys.handle), "restore_assembly")

    # device-resident Picard vectors: slot -1 is the iterate x
    X, USAV, UOLD, DU, UBEST = -1, 0, 1, 2, 3

    def _vop(self, op, dst, a=-1, b=-1, alpha=0.0):
        _lib.check(self._L.cismgpu_system_vector_op(self.sys.handle, op, dst, a, b, float(alpha)), "vector_op")

    def picard(self, *, maxiter=100, resid_target=1e-4, whichresid=0, algo=ALGO_CHRONGEAR,
               precond=HO_PRECOND_SIA, ncheck=5, linear_tol=1e-8, linear_maxiters=200, precision=FP64,
               inner_tol=1e-3, max_outer=20, accelerate=False, verbose=False) -> PicardResult:
        """Outer Picard iteration on the viscosity (glissade_velo_higher.F90:2163-3663),
        with every velocity field kept on the device.  accelerate=False:
        which_ho_nonlinear = 0; True: Picard acceleration (Dukowicz 2019;
        :2186-2263, :3108-3176, evaluate_accelerated_picard_3d :10052-10184):
        from the second iteration the velocity is extended along the last
        Picard step, u = u_old + alpha du, and alpha grows by 0.4 (to at most
        3) while each extension cuts the residual |A(u) u - b| below 0.85 of
        the previous one; a rejected extension restores the best velocity and
        its assembled system.  Convergence: resid_velo < resid_target for
        which_ho_resid 0/1/2, L2 < resid_target for 3, L2 / |b| for 4
        (:3652-3661)."""
        u, v = self._initial_velocity()
        self.zero_inactive(u, v)
        self._apply_no_penetration(u, v)
        self.sys.set_x(u, v)
        _lib.check(self._L.cismgpu_system_work_vectors(self.sys.handle, 4), "work_vectors")
        X, USAV, UOLD, DU, UBEST = self.X, self.USAV, self.UOLD, self.DU, self.UBEST
        res = PicardResult(u, v, None, 0, False)
        if accelerate:
            self._vop(0, UOLD, X)
            self._vop(3, DU)
            alpha, L2_sav = 1.0, L2_NORM_LARGE
        criterion = np.inf
        counter = 0
        rvd = ctypes.c_double()
        while criterion >= resid_target and counter < maxiter:
            counter += 1
            done = False
            while not done:
                if accelerate and counter >= 2:
                    self._vop(1, X, UOLD, DU, alpha)
                    _lib.check(self._L.cismgpu_system_halo_x(self.sys.handle), "halo_x")
                    self._after_extend()
                self._vop(0, USAV, X)
                ai = self.assemble()
                res.assemble_ms.append(ai["total_ms"])
                if ai["nnz"] == 0:          # nothing to solve (:2937-3032)
                    res.uvel, res.vvel = self._zeros(), self._zeros()
                    res.converged = True
                    res.iterations = counter
                    return res
                L2, L2r = self.residual()
                res.L2_norm.append(L2)
                if accelerate and counter >= 2:
                    if L2 < RESID_REDUCTION_THRESHOLD * L2_sav and alpha + GAMMA_ACCEL <= ALPHA_ACCEL_MAX:
                        self._vop(0, UBEST, X)
                        self.save_assembly()
                        alpha = alpha + GAMMA_ACCEL
                        L2_sav = L2
                    elif L2 < L2_sav:
                        self._vop(0, UOLD, X)
                        alpha, L2_sav, done = 1.0, L2_NORM_LARGE, True
                    else:
                        self._vop(0, X, UBEST)
                        self.restore_assembly()
                        self._vop(0, UOLD, UBEST)
                        alpha, L2_sav, done = 1.0, L2_NORM_LARGE, True
                    res.alpha.append(alpha)
                else:
                    done = True
            si = self.sys.solve(algo=algo, precond=precond, ncheck=ncheck, maxiters=linear_maxiters,
                                tolerance=linear_tol, precision=precision, inner_tol=inner_tol, max_outer=max_outer,
                                raise_on_error=False)
            # A linear solve that stops short of its tolerance is not an error
            # (the reference's solvers return niters = maxiters silently,
            # glissade_velo_higher_pcg.F90:1196, :3047); every other status is.
            if si["status"] not in (0, NOT_CONVERGED):
                raise _lib.CismGpuError(si["status"], f"Picard iteration {counter}: linear solve")
            res.linear_status.append(si["status"])
            res.linear_outer.append(si["outer_iters"])
            res.linear_iters.append(si["niters"])
            res.linear_err.append(si["err"])
            res.solve_ms.append(si["solve_ms"])
            _lib.check(self._L.cismgpu_system_halo_x(self.sys.handle), "halo_x")
            self._after_solve()
            _lib.check(self._L.cismgpu_system_residual_velocity(self.sys.handle, whichresid, USAV, ctypes.byref(rvd)),
                       "residual_velocity")
            rv = rvd.value
            res.resid_velo.append(rv)
            if accelerate:
                self._vop(2, DU, X, UOLD)
            criterion = L2 if whichresid == 3 else (L2r if whichresid == 4 else rv)
            if verbose:
                print(f"picard {counter}: L2={L2:.4e} resid_velo={rv:.4e} linear iters={si['niters']} "
                      f"err={si['err']:.2e}", flush=True)
        res.iterations = counter
        res.converged = criterion < resid_target
        res.uvel, res.vvel = self.sys.get_x()
        res.efvs = self.efvs()
        return res


class SSASolver(BPSolver):
    """Device-resident 2-D shallow-shelf (SSA) velocity solver: the
    ``solve_2d`` branch of glissade_velo_higher_solve with which_ho_approx = 1
    (glissade_velo_higher.F90:2353-2725 assembly and BCs, :3245-3272 solve,
    compute_residual_velocity_2d :10188-10310).  The setup (masks, staggered
    geometry, flow factor, beta, load inputs) is the 3-D one; the unknowns
    are the (nx-1, ny-1) velocities, started from the basal level of the 3-D
    guess (:1362).  The Picard loop and its acceleration are those of
    :class:`BPSolver`."""

    APPROX = 1

    def __init__(self, problem, system: System | None = None):
        super().__init__(problem, system or System(problem.grid, two_d=True))
        if not self.sys.two_d:
            raise ValueError("SSASolver needs a 2-D system")

    def _initial_velocity(self):
        nz = self.grid.nz
        u = np.asfortranarray(np.array(self.p.uvel[nz - 1], dtype=np.float64))
        v = np.asfortranarray(np.array(self.p.vvel[nz - 1], dtype=np.float64))
        return u, v

    def _zeros(self):
        g = self.grid
        return np.zeros((g.nvx, g.nvy), order="F")

    def zero_inactive(self, uvel, vvel):
        """2-D velocities at inactive owned vertices are zeroed (glissade_velo_higher.F90:1636-1647)."""
        g = self.grid
        il, jl = slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi)
        inact = self.active_vertex[il, jl] == 0
        for a in (uvel, vvel):
            o = a[il, jl]
            o[inact] = 0.0
            a[il, jl] = o

    def internal_stress(self, uvel3d=None, vvel3d=None):
        """compute_internal_stress (glissade_velo_higher.F90:7046-7281) for
        the last assembly's approximation, of the 3-D velocity (nz, nx-1,
        ny-1) with that assembly's geometry: SSA reads the resident x at
        every level when no velocity is given; L1L2 and DIVA take
        :meth:`velocity_3d` of the resident x by default.  The same dict as
        :meth:`BPSolver.internal_stress`."""
        g = self.grid
        if uvel3d is None and self.APPROX != 1:
            uvel3d, vvel3d = self.velocity_3d()[:2]
        f64 = lambda a: None if a is None else np.asfortranarray(np.array(a, dtype=np.float64))
        u3, v3 = f64(uvel3d), f64(vvel3d)
        for a in (u3, v3):
            if a is not None and a.shape != (g.nz, g.nvx, g.nvy):
                raise ValueError(f"internal_stress: the 3-D velocity is ({g.nz}, {g.nvx}, {g.nvy}), got {a.shape}")
        names = ("tau_xz", "tau_yz", "tau_xx", "tau_yy", "tau_xy", "tau_eff")
        out = {k: np.zeros((g.nz - 1, g.nx, g.ny), order="F") for k in names}
        D = ctypes.POINTER(ctypes.c_double)
        f = self._L.cismgpu_system_internal_stress_2d
        f.argtypes = [ctypes.c_void_p] + [D] * 8
        f.restype = ctypes.c_int32
        ptr = lambda a: None if a is None else a.ctypes.data_as(D)
        _lib.check(f(self.sys.handle, ptr(u3), ptr(v3), *[out[k].ctypes.data_as(D) for k in names]),
                   "internal_stress_2d")
        return out

    def assemble(self, uvel=None, vvel=None):
        """Assemble the 2-D SSA A, b for the viscosity of (uvel, vvel) (nx-1, ny-1);
        None keeps the resident x."""
        if uvel is not None:
            u = np.asfortranarray(np.array(uvel, dtype=np.float64))
            v = np.asfortranarray(np.array(vvel, dtype=np.float64))
            self.zero_inactive(u, v)
            self._apply_no_penetration(u, v)
            self.sys.set_x(u, v)
        if self.friction is not None:
            basal.device_calcbeta(self.sys)
        info = AssemblyInfo()
        _lib.check(self._L.cismgpu_system_assemble_2d(self.sys.handle, ctypes.byref(self._ain), self.APPROX,
                                                      ctypes.byref(info)), "assemble_2d")
        return dict(nnz=info.nnz, asym_pairs=info.asym_pairs, assemble_ms=info.assemble_ms, total_ms=info.total_ms)

    def picard(self, *, precond=1, **kw) -> PicardResult:
        """Picard iteration of the 2-D solve; which_ho_precond 2 is replaced by
        the diagonal preconditioner for 2-D solves (glide_setup.F90:1363-1366)."""
        return super().picard(precond=1 if precond == HO_PRECOND_SIA else precond, **kw)


class L1L2Solver(SSASolver):
    """Device-resident L1L2 solver (Perego, Gunzburger and Burkardt 2012;
    which_ho_approx = 3): the ``solve_2d`` branch of glissade_velo_higher_solve
    with the layer viscosities of compute_effective_viscosity_L1L2
    (glissade_velo_higher.F90:7474-7726) weighting the 2-D elements through
    their vertical mean (:5228-5248); basal friction, Dirichlet rows, the
    Picard loop and its acceleration are those of the SSA solve; after the
    loop compute_3d_velocity_L1L2 (:5904-6315, called at :3698-3724) forms
    the 3-D velocity."""

    APPROX = 3

    def __init__(self, problem, system: System | None = None):
        super().__init__(problem, system)
        self.flwa = np.asfortranarray(np.array(self.p.flwa, dtype=np.float64))
        self._ain.flwa = self.flwa.ctypes.data

    def velocity_3d(self):
        """compute_3d_velocity_L1L2 from the resident 2-D x and the last
        assembly: (uvel, vvel) (nz, nx-1, ny-1), halo not applied."""
        g = self.grid
        D = ctypes.POINTER(ctypes.c_double)
        u3, v3 = g.zeros_vec(), g.zeros_vec()
        _lib.check(self._L.cismgpu_system_l1l2_velocity(self.sys.handle, self.ice_mask.ctypes.data,
                                                         u3.ctypes.data_as(D), v3.ctypes.data_as(D)), "l1l2_velocity")
        return u3, v3

    def picard(self, *, precond=1, **kw) -> PicardResult:
        """Picard iteration of the L1L2 solve (plain or accelerated); the result
        adds the 3-D velocity of the converged 2-D velocity (uvel3d, vvel3d)."""
        r = super().picard(precond=precond, **kw)
        r.uvel3d = r.vvel3d = None
        if r.efvs is not None:
            r.uvel3d, r.vvel3d = self.velocity_3d()
        return r


class DIVASolver(SSASolver):
    """Device-resident depth-integrated viscosity approximation (DIVA,
    Goldberg 2011; which_ho_approx = 4, diva_slope_correction on): the
    ``solve_2d`` branch of glissade_velo_higher_solve with the DIVA
    assembly (layer viscosities from the basal traction, their vertical
    mean in the element matrix; glissade_velo_higher.F90:5198-5300,
    compute_effective_viscosity_diva :7730-7997), omega and beta_eff
    (compute_integrals_diva :5649-5775, :2469-2610), basal_sliding_bc_2d_diva
    (:8535-8811) and compute_3d_velocity_diva (:5779-5900) after each solve
    (:3566-3590) and after each accelerated-Picard extension (:2188-2210),
    where beta_eff and omega_k are saved and restored with the matrix
    (:9898-10048).  No Dirichlet rows: no-slip enters through beta_eff
    (:1423-1428).  The state (efvs_qp_3d, btractx/y) stays on the device and
    starts at zero unless given.  A friction law reads the basal level of
    the last 3-D velocity (the 2-D guess before the first), as calcbeta
    reads uvel(nz,:,:) (:2293-2296)."""

    APPROX = 4

    def __init__(self, problem, system: System | None = None, state: dict | None = None):
        super().__init__(problem, system)
        p = self.p
        self.diva = ge.diva_geometry(self.grid, p.dx, p.dy, self.usrf, self.lsrf, self.ice_mask, self.land_mask)
        self.no_slip = p.whichbabc == 6
        if self.friction is None:
            self._ain.beta_internal = self.beta_internal.ctypes.data
        # no Dirichlet rows for no-slip (beta_eff carries it, :1423-1428); kinematic-BC columns stay
        if self.kinbc is None and self.nopen_u is None:
            self._ain.umask_dirichlet = None
            self._ain.vmask_dirichlet = None
        else:
            self.umask[:] = 0
            if self.kinbc is not None:
                self.umask[:, self.kinbc] = 1
            self.vmask = self.umask.copy(order="F")
            if self.nopen_u is not None:
                self.umask[:, self.nopen_u] = 1
                self.vmask[:, self.nopen_v] = 1
            self._ain.umask_dirichlet = self.umask.ctypes.data
            self._ain.vmask_dirichlet = self.vmask.ctypes.data
        self.set_state(state)

    def set_state(self, state: dict | None = None):
        """Load the DIVA inputs and state (efvs_qp_3d (nz-1, 4, nx, ny),
        btractx/y (nx-1, ny-1); None: zero)."""
        g = self.grid
        st = state or {}
        keep = {k: np.asfortranarray(np.array(st[k], dtype=np.float64))
                for k in ("efvs_qp_3d", "btractx", "btracty") if st.get(k) is not None}
        if "efvs_qp_3d" in keep and keep["efvs_qp_3d"].shape != (g.nz - 1, 4, g.nx, g.ny):
            raise ValueError("efvs_qp_3d must be (nz-1, 4, nx, ny)")
        dg = self.diva
        ptr = lambda a: None if a is None else a.ctypes.data
        din = DivaIn(*(ptr(dg[k]) for k in ("slope_factor_x", "slope_factor_y", "stag_slope_factor_x",
                                            "stag_slope_factor_y", "stag_theta_slope_x", "stag_theta_slope_y",
                                            "theta_basal_x", "theta_basal_y", "ice_plus_land_mask")),
                     ptr(keep.get("efvs_qp_3d")), ptr(keep.get("btractx")), ptr(keep.get("btracty")),
                     1 if self.no_slip else 0)
        _lib.check(self._L.cismgpu_system_diva_setup(self.sys.handle, ctypes.byref(din)), "diva_setup")

    def velocity_3d(self, want_velocity=True):
        """compute_3d_velocity_diva from the resident 2-D x: updates the basal
        traction state; returns (uvel, vvel) (nz, nx-1, ny-1) and (btractx, btracty)."""
        g = self.grid
        D = ctypes.POINTER(ctypes.c_double)
        bx, by = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
        u3 = v3 = None
        if want_velocity:
            u3, v3 = g.zeros_vec(), g.zeros_vec()
        _lib.check(self._L.cismgpu_system_diva_velocity(
            self.sys.handle, u3.ctypes.data_as(D) if u3 is not None else None,
            v3.ctypes.data_as(D) if v3 is not None else None, bx.ctypes.data_as(D), by.ctypes.data_as(D)),
            "diva_velocity")
        return u3, v3, bx, by

    def _after_solve(self):
        _lib.check(self._L.cismgpu_system_diva_velocity(self.sys.handle, None, None, None, None), "diva_velocity")

    def _after_extend(self):
        self._after_solve()

    def picard(self, *, precond=1, **kw) -> PicardResult:
        """Picard iteration of the DIVA solve (plain or accelerated); the result
        adds the 3-D velocity of the last iterate (uvel3d, vvel3d) and the
        basal traction."""
        r = super().picard(precond=precond, **kw)
        r.uvel3d, r.vvel3d, r.btractx, r.btracty = self.velocity_3d()
        return r
