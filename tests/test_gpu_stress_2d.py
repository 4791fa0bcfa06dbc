"""compute_internal_stress (glissade_velo_higher.F90:7046-7281) for the
depth-integrated approximations on the device
(cismgpu_system_internal_stress_2d) against the oracle's branch of the same
approximation: SSA (horizontal strain rates of the 2-D velocity at every
level), L1L2 (every strain rate of the reconstructed 3-D velocity, the layer
viscosity of the last assembly) and DIVA (the 3-D branch).

Bar: each component within 1e-12 of its largest magnitude (the strain-rate
sums and the viscosity's pow are evaluated in the oracle's order).
"""
import numpy as np
import pytest

from cism_amd import _lib, synthetic
from cism_amd.velo import DIVASolver, L1L2Solver, SSASolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

KW = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-3, maxiter=8)


def _compare(got, ref):
    for k, v in ref.items():
        scale = np.abs(v).max()
        assert scale > 0 or k in ("tau_xz", "tau_yz"), k
        assert np.abs(got[k] - v).max() <= 1e-12 * max(scale, 1e-300), k


@pytest.mark.parametrize("name", ["stream", "marine"])
def test_ssa(require_gpu, name):
    p = synthetic.ice_stream(40, nz=7, seed=3) if name == "stream" else synthetic.marine_sheet(33, nz=6, seed=2)
    s = SSASolver(p)
    r = s.picard(**KW)
    g = p.grid
    u3 = np.asfortranarray(np.broadcast_to(r.uvel[None], (g.nz, g.nvx, g.nvy)))
    v3 = np.asfortranarray(np.broadcast_to(r.vvel[None], (g.nz, g.nvx, g.nvy)))
    ref = po.internal_stress(p, u3, v3, approx=1)
    assert np.abs(ref["tau_xz"]).max() == 0.0   # the SSA branch has no vertical shear
    _compare(s.internal_stress(), ref)          # the resident x at every level
    _compare(s.internal_stress(u3, v3), ref)    # the same field from the host


def test_l1l2(require_gpu):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = L1L2Solver(p)
    s.picard(**KW)
    u3, v3 = s.velocity_3d()
    ref = po.internal_stress(p, u3, v3, approx=3, efvs=s.efvs())
    assert np.abs(ref["tau_xz"]).max() > 0
    _compare(s.internal_stress(u3, v3), ref)
    _compare(s.internal_stress(), ref)          # velocity_3d of the resident x


def test_diva(require_gpu):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = DIVASolver(p)
    r = s.picard(**KW)
    ref = po.internal_stress(p, r.uvel3d, r.vvel3d, approx=4)
    _compare(s.internal_stress(r.uvel3d, r.vvel3d), ref)


def test_errors(require_gpu):
    p = synthetic.ice_stream(40, nz=7, seed=3)
    s = SSASolver(p)
    with pytest.raises(_lib.CismGpuError, match="needs a 2-D system"):
        s.internal_stress()
    s.assemble(s._zeros(), s._zeros())
    g = p.grid
    with pytest.raises(ValueError, match="3-D velocity"):
        s.internal_stress(np.zeros((g.nz - 1, g.nvx, g.nvy)), np.zeros((g.nz - 1, g.nvx, g.nvy)))
