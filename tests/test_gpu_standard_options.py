"""Quadrature (standard) options of the gravity load and the basal term on
the GPU against the CPU restatement: which_ho_assemble_taud = 0
(glissade_velo_higher.F90:4350-4446) and which_ho_assemble_beta = 0
(3-D :8955-9111 with the curved basal Jacobian; 2-D :8386-8527).

Bar: owned rows 1e-12 relative to the row's largest entry, rhs 1e-13
relative to its largest entry; the SSA Picard solve with both options
reaches the restatement's iteration count and velocities (1e-7).
"""
import numpy as np
import pytest

from cism_amd import synthetic
from cism_amd.velo import BPSolver, SSASolver
from oracle import pyoracle as po

from test_gpu_assembly import _rowwise_close

pytestmark = pytest.mark.gpu


def _problems():
    return {
        "stream": synthetic.ice_stream(40, nz=11, seed=3),
        "slab_periodic": synthetic.slab(n_cells=8, nz=9),
        "marine": synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0),
        "dome": synthetic.dome(),
    }


PROBLEMS = _problems()
OPTIONS = [(0, 1), (1, 0), (0, 0)]      # (assemble_beta, assemble_taud)


@pytest.mark.parametrize("opts", OPTIONS)
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_bp_standard_options_match_oracle(require_gpu, name, opts):
    p = PROBLEMS[name]
    p.assemble_beta, p.assemble_taud = opts
    try:
        ref = po.assemble(p)
        s = BPSolver(p)
        s.assemble(p.uvel, p.vvel)
        (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
        for a, key in ((Auu, "Auu"), (Auv, "Auv"), (Avu, "Avu"), (Avv, "Avv")):
            e = _rowwise_close(a, ref[key], p.grid, ref["active_vertex"], 1e-12)
            assert e < 1e-12, (key, e)
        bscale = max(np.abs(ref["bu"]).max(), np.abs(ref["bv"]).max(), 1e-300)
        assert np.abs(bu - ref["bu"]).max() / bscale < 1e-13
        assert np.abs(bv - ref["bv"]).max() / bscale < 1e-13
    finally:
        p.assemble_beta, p.assemble_taud = 1, 1


@pytest.mark.parametrize("opts", OPTIONS)
@pytest.mark.parametrize("name", ["stream", "marine", "slab_periodic"])
def test_ssa_standard_options_match_oracle(require_gpu, name, opts):
    p = PROBLEMS[name]
    g = p.grid
    p.assemble_beta, p.assemble_taud = opts
    try:
        ref = po.assemble_2d(p)
        s = SSASolver(p)
        s.assemble(ref["uvel_2d"], ref["vvel_2d"])
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
    finally:
        p.assemble_beta, p.assemble_taud = 1, 1


def test_ssa_picard_with_standard_options(require_gpu):
    p = PROBLEMS["stream"]
    p.assemble_beta, p.assemble_taud = 0, 0
    try:
        kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=60)
        ref = po.picard(p, two_d=True, precond=1, **kw)
        r = SSASolver(p).picard(precond=1, **kw)
        assert r.converged and ref["converged"] and r.iterations == ref["iterations"]
        num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
        assert num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]])) < 1e-7
    finally:
        p.assemble_beta, p.assemble_taud = 1, 1


def test_standard_basal_needs_staggered_lower_surface(require_gpu):
    import ctypes
    from cism_amd import _lib
    from cism_amd.velo import AssemblyInfo
    p = PROBLEMS["stream"]
    s = BPSolver(p)
    s._ain.assemble_beta_standard = 1
    s._ain.stag_lsrf = None
    st = s._L.cismgpu_system_assemble_3d(s.sys.handle, ctypes.byref(s._ain), ctypes.byref(AssemblyInfo()))
    assert st == 3 and "stag_lsrf" in _lib.last_error()
