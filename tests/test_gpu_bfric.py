"""compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042) with
its cell halo (:3807) and the basal traction beta u_b (:3817-3818) on the
device (cismgpu_system_basal_friction), against the oracle's restatement.

Bar: the heat flux within 1e-14 of its largest value (every sum is rounded
in the oracle's order; the shape-function table is built by the same
formula) and the traction bit for bit, for both which_ho_assemble_bfric
options, after BP, SSA, L1L2 and DIVA Picard solves, with the system's own
basal velocity and beta and with host fields passed in.
"""
import numpy as np
import pytest

from cism_amd import _lib, synthetic
from cism_amd.basal import Friction
from cism_amd.grid import BC_PERIODIC
from cism_amd.velo import BPSolver, DIVASolver, L1L2Solver, SSASolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _check(s, p, which, u_ref, v_ref, beta_ref, **kw):
    got = s.basal_friction(which, **kw)
    ref = po.basal_friction(p.grid, which, s.active_cell, s.active_vertex, u_ref, v_ref, beta_ref)
    scale = np.abs(ref[0]).max()
    assert scale > 0
    assert np.abs(got[0] - ref[0]).max() <= 1e-14 * scale
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    return got


def _powerlaw(p):
    g = p.grid
    return Friction(9, powerlaw_c=np.full((g.nvx, g.nvy), 2.0e4), powerlaw_m=3.0)


KW = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-3, maxiter=10)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("name", ["stream", "marine"])
def test_bp_fixed_beta(require_gpu, which, name):
    """Fixed beta: the last assembly's beta is the host beta_internal."""
    p = synthetic.ice_stream(40, nz=7, seed=3) if name == "stream" else synthetic.marine_sheet(33, nz=6, seed=2)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, **KW)
    nz = p.grid.nz
    _check(s, p, which, r.uvel[nz - 1], r.vvel[nz - 1], s.beta_internal)


@pytest.mark.parametrize("which", [0, 1])
def test_bp_device_beta(require_gpu, which):
    """A sliding law: the device beta of the last calcbeta."""
    p = synthetic.ice_stream(40, nz=7, seed=3)
    p.friction = _powerlaw(p)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, **KW)
    beta = s.calcbeta(want=True)   # refresh the device beta from the converged velocity
    nz = p.grid.nz
    _check(s, p, which, r.uvel[nz - 1], r.vvel[nz - 1], beta)


@pytest.mark.parametrize("cls", [SSASolver, L1L2Solver])
def test_two_d_fixed_beta(require_gpu, cls):
    """SSA and L1L2: the basal velocity is the 2-D solution."""
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = cls(p)
    r = s.picard(**KW)
    for which in (0, 1):
        _check(s, p, which, r.uvel, r.vvel, s.beta_internal)


def test_diva_sliding_law(require_gpu):
    """DIVA: the bed level of the last compute_3d_velocity_diva."""
    p = synthetic.ice_stream(40, nz=7, seed=3)
    p.friction = _powerlaw(p)
    s = DIVASolver(p)
    r = s.picard(**KW)
    beta = s.calcbeta(want=True)
    nz = p.grid.nz
    for which in (0, 1):
        _check(s, p, which, r.uvel3d[nz - 1], r.vvel3d[nz - 1], beta)


def test_diva_without_friction_needs_velocity(require_gpu):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = DIVASolver(p)
    r = s.picard(**KW)
    with pytest.raises(_lib.CismGpuError, match="pass ubas"):
        s.basal_friction(1)
    nz = p.grid.nz
    ub, vb = r.uvel3d[nz - 1], r.vvel3d[nz - 1]
    _check(s, p, 1, ub, vb, s.beta_internal, ubas=ub, vbas=vb)


def test_host_fields_periodic(require_gpu):
    """Host velocity and beta on a periodic grid: the halo wraps."""
    p = synthetic.ismip_hom("c", size_km=40.0, n=20, nz=5)
    assert p.grid.global_bc == BC_PERIODIC
    s = BPSolver(p)
    s.assemble(s._zeros(), s._zeros())
    g = p.grid
    rng = np.random.default_rng(17)
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (g.nvx, g.nvy)))
    ub, vb, b = f(-300, 300), f(-300, 300), f(0, 5e3)
    for which in (0, 1):
        _check(s, p, which, ub, vb, b, ubas=ub, vbas=vb, beta=b)


def test_errors(require_gpu):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = BPSolver(p)
    with pytest.raises(_lib.CismGpuError, match="needs a system with an assembly"):
        s.basal_friction(1)
    s.assemble(s._zeros(), s._zeros())
    with pytest.raises(_lib.CismGpuError, match="which_ho_assemble_bfric 2"):
        s.basal_friction(2)
    with pytest.raises(ValueError, match="both"):
        s.basal_friction(1, ubas=np.zeros((p.grid.nvx, p.grid.nvy)))
