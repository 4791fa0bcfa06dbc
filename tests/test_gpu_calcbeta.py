"""glissade_calcbeta on the device (cismgpu_system_calcbeta) against the
oracle's restatement, and Picard solves whose basal traction follows the
velocity.

Bar: beta within 1e-13 relative of the oracle at every vertex, halo
included (the device pow and tan are within an ulp or two of the host's
correctly rounded ones, the rest is rounded in the same order); the host
and resident basal-velocity paths give the same field bitwise; a NaN beta
is CISMGPU_ERR_BETA naming the oracle's vertex; Picard (BP and DIVA, plain
and accelerated; SSA; L1L2) with a velocity-dependent law reaches the
oracle's velocities (1e-7) in the same number of iterations.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from cism_amd import _lib, basal, synthetic
from cism_amd.basal import Friction
from cism_amd.grid import BC_NO_ICE, BC_PERIODIC, Grid
from cism_amd.pcg import System
from cism_amd.velo import AssemblyInfo, BPSolver, DIVASolver, L1L2Solver, SSASolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    shp, cshp = (g.nvx, g.nvy), (g.nx, g.ny)
    F = lambda lo, hi, s=shp: np.asfortranarray(rng.uniform(lo, hi, s))
    thck = F(0.0, 2000.0, cshp)
    thck[thck < 300.0] = 0.0
    fg = np.asfortranarray((rng.uniform(0, 1, shp) > 0.2) * rng.choice([1.0, 0.5, 0.25], shp))
    flwa = np.full((g.nz - 1, g.nx, g.ny), 1e-16, order="F")
    flwa[-1] = F(0.5e-16, 2e-16, cshp)
    return dict(N=F(1e4, 5e6), cc=F(0.05, 0.9), pc=F(1e3, 3e4), tauf=F(1e3, 1e5), csf=F(0.5, 1.0),
                bpmp=np.asfortranarray((rng.uniform(0, 1, shp) > 0.5).astype(np.int32)), ext=F(10.0, 1e4),
                topg=F(-800.0, 1200.0, cshp), thck=thck, flwa=flwa, fg=fg)


def _friction(opt, f, limit, **kw):
    base = dict(which_beta_limit=limit, beta_grounded_min=5.0, effecpress_stag=f["N"], coulomb_c=f["cc"],
                powerlaw_c=f["pc"], mintauf=f["tauf"], bpmp_mask=f["bpmp"], beta_external=f["ext"],
                c_space_factor_stag=f["csf"] if opt in (10, 11, 13) else None)
    base.update(kw)
    return Friction(opt, **base)


def _velocity(g, seed):
    rng = np.random.default_rng(seed)
    u = np.asfortranarray(rng.uniform(-600.0, 600.0, g.vec_shape))
    v = np.asfortranarray(rng.uniform(-600.0, 600.0, g.vec_shape))
    u[:, 3:7, 4:6] = 0.0                     # stagnant vertices: the smallnum floor
    v[:, 3:7, 4:6] = 0.0
    return u, v


def _oracle(g, fr, f, eus, ub, vb):
    p = SimpleNamespace(grid=g, friction=fr, eus=eus, topg=f["topg"], thck=f["thck"], flwa=f["flwa"], dx=DX, dy=DY)
    return po.calcbeta(p, {"f_ground": f["fg"]}, ub, vb)


def _device(sys3, g, fr, f, eus, **kw):
    basal.device_setup(sys3, fr, f_ground=f["fg"], topg=f["topg"], thck=f["thck"], eus=eus,
                       flwa_basal=f["flwa"][g.nz - 2], dew=DX, dns=DY)
    return basal.device_calcbeta(sys3, want=True, **kw)


DX, DY = 1500.0, 1500.0          # grid spacing (read by ISMIP-HOM C)
GRIDS = {"noice": Grid.from_owned(40, 36, 5, global_bc=BC_NO_ICE),
         "periodic": Grid.from_owned(33, 30, 4, global_bc=BC_PERIODIC)}
CASES = [(opt, limit, gname) for opt in basal.DEVICE_OPTIONS for limit in (0, 1) for gname in GRIDS
         if gname == "noice" or limit == 0]


@pytest.mark.parametrize("opt,limit,gname", CASES)
def test_calcbeta_matches_oracle(require_gpu, opt, limit, gname):
    g = GRIDS[gname]
    f = _fields(g, 100 + opt)
    u, v = _velocity(g, 200 + opt)
    kw = {}
    if opt in (2, 3):
        kw = dict(pseudo_plastic_q=0.3)
    if opt in (9, 11, 13):
        kw = dict(beta_powerlaw_umax=450.0)
    if opt == 7:
        kw = dict(zoet_iverson_nmax=2.0e6)
    fr = _friction(opt, f, limit, **kw)
    eus = -37.5
    ref = _oracle(g, fr, f, eus, u[g.nz - 1], v[g.nz - 1])
    with System(g) as s:
        s.set_x(u, v)
        got = _device(s, g, fr, f, eus)
        # the same field from host basal velocities
        got_host = basal.device_calcbeta(s, u[g.nz - 1], v[g.nz - 1], want=True)
    assert np.array_equal(got, got_host)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0.0)
    assert (got >= 0).all()


def test_calcbeta_on_a_two_d_system(require_gpu):
    """A 2-D system reads its x (the basal velocity of SSA and L1L2)."""
    g = GRIDS["noice"]
    f = _fields(g, 7)
    u, v = _velocity(g, 8)
    fr = _friction(11, f, 0)
    ref = _oracle(g, fr, f, 0.0, u[g.nz - 1], v[g.nz - 1])
    with System(g, two_d=True) as s:
        s.set_x(np.asfortranarray(u[g.nz - 1]), np.asfortranarray(v[g.nz - 1]))
        got = _device(s, g, fr, f, 0.0)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0.0)


def test_calcbeta_errors(require_gpu):
    import ctypes
    g = GRIDS["noice"]
    f = _fields(g, 9)
    u, v = _velocity(g, 10)
    L = _lib.load()
    basal._declare(L)
    with System(g) as s:
        s.set_x(u, v)
        assert L.cismgpu_system_calcbeta(s.handle, None, None, None) == 3       # no setup
        geo = dict(f_ground=f["fg"], topg=f["topg"], thck=f["thck"], eus=0.0, flwa_basal=f["flwa"][-1])
        for fr, status, text in ((Friction(6), 3, "no-slip"), (Friction(8), 3, "ISMIP-HOM"),
                                 (Friction(9), 3, "powerlaw_c"),
                                 (Friction(5, beta_external=-np.ones((g.nvx, g.nvy))), basal.CISMGPU_ERR_BETA, "< 0")):
            with pytest.raises(_lib.CismGpuError) as e:
                basal.device_setup(s, fr, **geo)
            assert e.value.status == status and text in str(e.value), str(e.value)
        # a NaN coefficient: the error names the vertex the oracle names
        cc = f["cc"].copy(order="F")
        cc[11, 17] = np.nan
        fr = _friction(2, {**f, "cc": cc}, 0)
        with pytest.raises(ValueError, match="i=12 j=18"):
            _oracle(g, fr, f, 0.0, u[g.nz - 1], v[g.nz - 1])
        with pytest.raises(_lib.CismGpuError) as e:
            _device(s, g, fr, f, 0.0)
        assert e.value.status == basal.CISMGPU_ERR_BETA and "i=12 j=18" in str(e.value)
    # an assembly that reads the device field before any calcbeta
    bp = BPSolver(_stream_with(_powerlaw))
    info = AssemblyInfo()
    assert bp._L.cismgpu_system_assemble_3d(bp.sys.handle, ctypes.byref(bp._ain), ctypes.byref(info)) == 3
    assert "beta_device" in _lib.last_error()


def _stream_with(friction_of):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    p.friction = friction_of(p)
    return p


def _powerlaw(p):
    g = p.grid
    return Friction(9, powerlaw_c=np.full((g.nvx, g.nvy), 2.0e4), powerlaw_m=3.0)


def _schoof(p):
    g = p.grid
    rng = np.random.default_rng(5)
    N = np.asfortranarray(910.0 * 9.81 * synthetic_stag(p.thck) * rng.uniform(0.05, 0.3, (g.nvx, g.nvy)))
    return Friction(11, powerlaw_c=np.full((g.nvx, g.nvy), 2.0e4), coulomb_c=np.full((g.nvx, g.nvy), 0.5),
                    effecpress_stag=N, powerlaw_m=3.0)


def synthetic_stag(h):
    return 0.25 * (h[:-1, 1:] + h[1:, 1:] + h[:-1, :-1] + h[1:, :-1])


def _compare(r, ref):
    assert r.converged and ref["converged"]
    assert r.iterations == ref["iterations"], (r.iterations, ref["iterations"])
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    rel = num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]]))
    assert rel < 1e-7, rel


@pytest.mark.parametrize("law", ["powerlaw", "schoof"])
@pytest.mark.parametrize("accelerate", [False, True])
def test_bp_picard_with_sliding_law_matches_oracle(require_gpu, law, accelerate):
    p = _stream_with(_powerlaw if law == "powerlaw" else _schoof)
    kw = dict(linear_tol=1e-10, resid_target=1e-4, maxiter=80, accelerate=accelerate)
    ref = po.picard(p, algo="chrongear", precond=2, linear_maxiters=2000, **kw)
    r = BPSolver(p).picard(algo=3, precond=2, linear_maxiters=2000, **kw)
    _compare(r, ref)
    if accelerate:
        assert r.alpha == ref["alpha"]


def test_ssa_picard_with_coulomb_law_matches_oracle(require_gpu):
    p = _stream_with(_schoof)
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=80)
    ref = po.picard(p, two_d=True, precond=1, **kw)
    r = SSASolver(p).picard(precond=1, **kw)
    _compare(r, ref)


@pytest.mark.parametrize("law", ["powerlaw", "schoof"])
@pytest.mark.parametrize("accelerate", [False, True])
def test_diva_picard_with_sliding_law_matches_oracle(require_gpu, law, accelerate):
    """DIVA: calcbeta reads the basal level of the last 3-D velocity
    (compute_3d_velocity_diva after each solve and each extension), and the
    3-D velocity after a restored assembly uses the latest beta."""
    p = _stream_with(_powerlaw if law == "powerlaw" else _schoof)
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=80, accelerate=accelerate)
    ref = po.picard(p, two_d=True, approx="diva", precond=1, **kw)
    r = DIVASolver(p).picard(precond=1, **kw)
    _compare(r, ref)
    if accelerate:
        assert r.alpha == pytest.approx(ref["alpha"])
    num = np.linalg.norm(np.stack([r.uvel3d - ref["uvel3d"], r.vvel3d - ref["vvel3d"]]))
    assert num / np.linalg.norm(np.stack([ref["uvel3d"], ref["vvel3d"]])) < 1e-7


def test_l1l2_picard_with_sliding_law_matches_oracle(require_gpu):
    p = _stream_with(_powerlaw)
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=80)
    ref = po.picard(p, two_d=True, approx="l1l2", precond=1, **kw)
    r = L1L2Solver(p).picard(precond=1, **kw)
    _compare(r, ref)


def test_device_beta_follows_the_velocity(require_gpu):
    """After a Picard solve the device beta is the law evaluated at the
    converged basal velocity (the oracle's calcbeta of the returned field)."""
    p = _stream_with(_powerlaw)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, linear_tol=1e-10, linear_maxiters=2000, resid_target=1e-4, maxiter=80)
    got = s.calcbeta(want=True)
    geo = po.geometry(p)
    ref = po.calcbeta(p, geo, r.uvel[p.grid.nz - 1], r.vvel[p.grid.nz - 1])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("approx", ["bp", "ssa", "l1l2", "diva"])
@pytest.mark.parametrize("kind", ["raymond", "schoof"])
def test_yield_stress_stream_matches_oracle_and_analytic(require_gpu, kind, approx):
    """The reference's ice-stream test (tests/stream): yield-stress till
    (which_ho_babc 15) recomputed on the device every iteration, a kinematic
    column, doubly periodic.  Device Picard = oracle Picard (iteration count,
    velocities 1e-7) and within 1.5 % of the Raymond / Schoof closed form."""
    p = synthetic.yield_stress_stream(kind)
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=300)
    if approx == "bp":
        ref = po.picard(p, algo="chrongear", precond=1, **kw)
        r = BPSolver(p).picard(algo=3, precond=1, **kw)
    else:
        ref = po.picard(p, two_d=True, approx=approx, precond=1, **kw)
        r = {"ssa": SSASolver, "l1l2": L1L2Solver, "diva": DIVASolver}[approx](p).picard(precond=1, **kw)
    _compare(r, ref)
    yv = p.meta["y_vertex"]
    ua = synthetic.stream_velocity(kind, yv)
    g = p.grid
    u = (r.uvel3d[0] if getattr(r, "uvel3d", None) is not None else (r.uvel[0] if approx == "bp" else r.uvel))
    u = u[g.ilo - 1 + 9]
    inside = np.abs(yv) < synthetic.STREAM_HALF_WIDTH
    assert np.max(np.abs(u[inside] - ua[inside])) / ua.max() < 1.5e-2


def test_yield_stress_stream_mixed_inner_solves(require_gpu):
    """The BP stream with FP32 inner solves inside FP64 refinement reaches the
    oracle's FP64 Picard velocities (1e-6) within one Picard iteration."""
    from cism_amd.pcg import MIXED_FP32
    p = synthetic.yield_stress_stream("raymond")
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=300)
    ref = po.picard(p, algo="chrongear", precond=1, **kw)
    r = BPSolver(p).picard(algo=3, precond=1, precision=MIXED_FP32, inner_tol=1e-6, **kw)
    assert r.converged and abs(r.iterations - ref["iterations"]) <= 1
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    assert num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]])) < 1e-6


@pytest.mark.parametrize("experiment,size_km", [("a", 80.0), ("c", 80.0), ("c", 20.0)])
def test_ismip_hom_matches_oracle(require_gpu, experiment, size_km):
    """ISMIP-HOM A and C as the reference runs them (tests/ismip-hom; C's
    friction built on the velocity grid by calcbeta option 8 on the device):
    device Picard = oracle Picard (iteration count, velocities 1e-7), and the
    device solution's basal drag balances the driving stress (the mean of
    beta u_b / cos(theta_b) = rho g mean(H) tan(alpha), 1e-6)."""
    p = synthetic.ismip_hom(experiment, size_km, n=20, nz=7)
    kw = dict(linear_tol=1e-10, linear_maxiters=2000, resid_target=1e-5, maxiter=100)
    ref = po.picard(p, algo="chrongear", precond=2, **kw)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, **kw)
    _compare(r, ref)
    g = p.grid
    geo = po.geometry(p)
    beta = s.calcbeta(want=True) if experiment == "c" else s.beta_internal
    if experiment == "c":      # the device field of the returned velocity = the oracle's
        np.testing.assert_allclose(beta, po.calcbeta(p, geo, r.uvel[-1], r.vvel[-1]), rtol=1e-13, atol=0.0)
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    alpha = (0.5 if experiment == "a" else 0.1) * np.pi / 180.0
    taud = 910.0 * 9.81 * geo["stagthck"][own].mean() * np.tan(alpha)
    c = np.cos(geo["theta_basal"])
    assert (beta * r.uvel[-1] / c)[own].mean() == pytest.approx(taud, rel=1e-6)


def _marine_glp(flotation, friction=False, ground=1):
    p = synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0)
    p.which_ho_ground = ground                  # GLP over the staggered cell (1) or its quadrants (2)
    p.which_ho_flotation_function = flotation
    if flotation == 3:   # a raised bed (pinning points) under every fourth cell
        raised = np.array(p.topg, dtype=np.float64)
        raised[::4, ::4] += 60.0
        p.topg_raised = np.asfortranarray(raised)
    if friction:
        g = p.grid
        N = np.asfortranarray(910.0 * 9.81 * synthetic_stag(p.thck) * 0.2)
        p.friction = Friction(11, powerlaw_c=np.full((g.nvx, g.nvy), 2.0e4), coulomb_c=np.full((g.nvx, g.nvy), 0.5),
                              effecpress_stag=N)
    return p


@pytest.mark.parametrize("flotation,friction,ground", [(2, False, 1), (0, False, 1), (2, True, 1), (2, False, 2),
                                                      (3, False, 1)])
def test_marine_picard_with_grounding_line_parameterisation(require_gpu, flotation, friction, ground):
    """Marine sheet with the GLP (fractional f_ground at the grounding line,
    host geometry = oracle to 1e-13): accelerated BP Picard (fixed beta, or
    the Schoof law recomputed on the device) = the oracle's (iteration count,
    alpha ladder, velocities 1e-7)."""
    p = _marine_glp(flotation, friction, ground)
    kw = dict(linear_tol=1e-10, resid_target=1e-4, maxiter=80, accelerate=True)
    ref = po.picard(p, algo="chrongear", precond=2, linear_maxiters=2000, **kw)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, linear_maxiters=2000, **kw)
    _compare(r, ref)
    assert r.alpha == ref["alpha"]
    fg = po.geometry(p)["f_ground"]
    assert ((fg > 0) & (fg < 1)).sum() > 20     # the parameterisation is active
