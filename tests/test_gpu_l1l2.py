"""GPU L1L2 (which_ho_approx = 3) against the CPU restatement.

Bar: every owned 9-point row, the rhs and the layer viscosities of one
L1L2 assembly agree to 1e-12 relative to the largest entry of the row; the
3-D velocity of compute_3d_velocity_L1L2 to 1e-11 on owned vertices; the
L1L2 Picard loop (plain and accelerated) reaches the restatement's 2-D and
3-D velocities (1e-7) with the same iteration count.
"""
import ctypes

import numpy as np
import pytest

from cism_amd import synthetic
from cism_amd.velo import AssemblyInfo, L1L2Solver, SSASolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _problems():
    return {
        "stream": synthetic.ice_stream(40, nz=11, seed=3),
        "slab_periodic": synthetic.slab(n_cells=8, nz=9),
        "dome_noslip": synthetic.dome(),
        "marine": synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0),
    }


PROBLEMS = _problems()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("whichefvs", [2, 0])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_l1l2_assembly_and_velocity_match_oracle(require_gpu, name, whichefvs):
    p = PROBLEMS[name]
    g = p.grid
    p_efvs = p.whichefvs
    p.whichefvs = whichefvs
    try:
        rng = np.random.default_rng(11)
        u = np.asfortranarray(p.uvel[g.nz - 1] + rng.standard_normal((g.nvx, g.nvy)) * 5.0)
        v = np.asfortranarray(p.vvel[g.nz - 1] + rng.standard_normal((g.nvx, g.nvy)) * 5.0)
        s = L1L2Solver(p)
        s.zero_inactive(u, v)
        ref = po.assemble_2d(p, u, v, approx="l1l2")
        s.assemble(u, v)
        (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
        act = np.asarray(ref["active_vertex"], bool)[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi]
        for a, key in ((Auu, "Auu"), (Auv, "Auv"), (Avu, "Avu"), (Avv, "Avv")):
            mine = a[9:18, 0, g.ilo - 1:g.ihi, g.jlo - 1:g.jhi]
            want = np.moveaxis(ref[key], 2, 0)[:, g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] * act[None]
            scale = np.maximum(np.abs(want).max(axis=0, keepdims=True), 1e-300)
            assert float((np.abs(mine - want) / scale).max()) < 1e-12, key
        bscale = max(np.abs(ref["bu"]).max(), np.abs(ref["bv"]).max(), 1e-300)
        assert np.abs(bu - ref["bu"]).max() / bscale < 1e-13
        assert np.abs(bv - ref["bv"]).max() / bscale < 1e-13
        assert _rel(s.efvs(), ref["efvs"]) < 1e-12
        xu, xv = s.sys.get_x()
        u3, v3 = s.velocity_3d()
        ou3, ov3 = po.l1l2_velocity(p, ref, xu, xv)
        own = g.owned_slice()
        assert _rel(u3[own], ou3[own]) < 1e-11 and _rel(v3[own], ov3[own]) < 1e-11
        assert np.array_equal(u3[g.nz - 1], xu) and np.array_equal(v3[g.nz - 1], xv)
    finally:
        p.whichefvs = p_efvs


@pytest.mark.parametrize("accelerate", [False, True])
@pytest.mark.parametrize("name", ["stream", "marine"])
def test_l1l2_picard_matches_oracle(require_gpu, name, accelerate):
    p = PROBLEMS[name]
    g = p.grid
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=60, accelerate=accelerate)
    ref = po.picard(p, two_d=True, approx="l1l2", precond=1, **kw)
    r = L1L2Solver(p).picard(precond=1, **kw)
    assert r.converged and ref["converged"]
    assert r.iterations == ref["iterations"], (r.iterations, ref["iterations"])
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    assert num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]])) < 1e-7
    own = g.owned_slice()
    num3 = np.linalg.norm(np.stack([r.uvel3d[own] - ref["uvel3d"][own], r.vvel3d[own] - ref["vvel3d"][own]]))
    assert num3 / np.linalg.norm(np.stack([ref["uvel3d"][own], ref["vvel3d"][own]])) < 1e-7


def test_l1l2_guards(require_gpu):
    """L1L2 needs the rate factor; its 3-D velocity needs an L1L2 assembly."""
    from cism_amd import _lib
    p = PROBLEMS["stream"]
    s = SSASolver(p)
    st = s._L.cismgpu_system_assemble_2d(s.sys.handle, ctypes.byref(s._ain), 3, ctypes.byref(AssemblyInfo()))
    assert st == 3 and "flwa" in _lib.last_error()
    s.assemble()
    D = ctypes.POINTER(ctypes.c_double)
    u3, v3 = p.grid.zeros_vec(), p.grid.zeros_vec()
    st = s._L.cismgpu_system_l1l2_velocity(s.sys.handle, s.ice_mask.ctypes.data, u3.ctypes.data_as(D),
                                           v3.ctypes.data_as(D))
    assert st == 3 and "L1L2" in _lib.last_error()
