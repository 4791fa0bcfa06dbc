"""GPU assembly and Picard solve of the 2-D SSA system against the CPU
restatement.

Bar: every owned 9-point row, the rhs (the level-summed load after the
Dirichlet rows) and the element viscosities agree to 1e-12 relative to the
largest entry of the row; the nonzero pattern above round-off and the symmetry report equal
the restatement's; the assembled 2-D solve and the SSA Picard loop (plain
and accelerated) reach the restatement's velocities (1e-9 / 1e-7) with the
same Picard iteration count.
"""
import numpy as np
import pytest

from cism_amd import synthetic
from cism_amd.velo import SSASolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _local_front(p):
    p.assemble_lateral = 1
    return p


def _problems():
    return {
        "stream": synthetic.ice_stream(40, nz=11, seed=3),
        "slab_periodic": synthetic.slab(n_cells=8, nz=9),
        "dome_noslip": synthetic.dome(),
        "marine": synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0),
        "marine_local_front": _local_front(synthetic.marine_sheet(36, nz=5, dx=2500.0, seed=4, h_max=1000.0,
                                                                  h_shelf=200.0)),
    }


PROBLEMS = _problems()


def _rows(A, g):
    """(27, 1, nvx, nvy) device export -> owned part of the 9-point stencil (9, nxo, nyo)."""
    return A[9:18, 0, g.ilo - 1:g.ihi, g.jlo - 1:g.jhi]


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_assembly_2d_matches_oracle(require_gpu, name):
    p = PROBLEMS[name]
    g = p.grid
    ref = po.assemble_2d(p)
    s = SSASolver(p)
    info = s.assemble(ref["uvel_2d"], ref["vvel_2d"])
    (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
    act = np.asarray(ref["active_vertex"], bool)[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi]
    for a, key in ((Auu, "Auu"), (Auv, "Auv"), (Avu, "Avu"), (Avv, "Avv")):
        assert np.all(a[0:9] == 0.0) and np.all(a[18:27] == 0.0)
        mine = _rows(a, g)
        want = np.moveaxis(ref[key], 2, 0)[:, g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] * act[None]
        scale = np.maximum(np.abs(want).max(axis=0, keepdims=True), 1e-300)
        err = float((np.abs(mine - want) / scale).max())
        assert err < 1e-12, (key, err)
    bscale = max(np.abs(ref["bu"]).max(), np.abs(ref["bv"]).max(), 1e-300)
    assert np.abs(bu - ref["bu"]).max() / bscale < 1e-13
    assert np.abs(bv - ref["bv"]).max() / bscale < 1e-13
    ef = s.efvs()
    assert np.abs(ef - ref["efvs"]).max() / max(np.abs(ref["efvs"]).max(), 1e-300) < 1e-12
    # count_nonzeros_2d: the device count is the nonzeros of its own owned
    # active rows; the pattern above round-off is the restatement's (entries
    # that cancel exactly in one evaluation order can be ulps from zero in
    # the other, as on the uniform slab)
    blocks = [_rows(a, g) * act[None] for a in (Auu, Auv, Avu, Avv)]
    assert info["nnz"] == sum(int(np.count_nonzero(b)) for b in blocks)
    for bg, key in zip(blocks, ("Auu", "Auv", "Avu", "Avv")):
        br = np.moveaxis(ref[key], 2, 0)[:, g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] * act[None]
        sc = np.abs(br).max(axis=0, keepdims=True) * 1e-12
        assert np.array_equal(np.abs(bg) > sc, np.abs(br) > sc), key
    assert info["asym_pairs"] == ref["nwarn_asym"]


@pytest.mark.parametrize("name", ["stream", "marine", "slab_periodic"])
def test_assembled_2d_solve_matches_oracle(require_gpu, name):
    p = PROBLEMS[name]
    g = p.grid
    ref = po.assemble_2d(p)
    A = [ref["Auu"], ref["Auv"], ref["Avu"], ref["Avv"]]
    ou, ov = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
    st, oerr, onit = po.pcg_solve_2d(g, ref["active_vertex"], *A, ref["bu"], ref["bv"], ou, ov, precond=1,
                                     tolerance=1e-11, maxiters=20000)
    assert st == 0
    s = SSASolver(p)
    s.assemble(ref["uvel_2d"], ref["vvel_2d"])
    s.sys.set_x(np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F"))
    s.sys.solve(precond=1, tolerance=1e-11, maxiters=20000)
    xu, xv = s.sys.get_x()
    rel = np.linalg.norm(np.stack([xu - ou, xv - ov])) / np.linalg.norm(np.stack([ou, ov]))
    assert rel < 1e-9, rel


@pytest.mark.parametrize("accelerate", [False, True])
@pytest.mark.parametrize("name", ["stream", "marine"])
def test_ssa_picard_matches_oracle(require_gpu, name, accelerate):
    p = PROBLEMS[name]
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=60, accelerate=accelerate)
    ref = po.picard(p, two_d=True, precond=1, **kw)
    r = SSASolver(p).picard(precond=1, **kw)
    assert r.converged and ref["converged"]
    assert r.iterations == ref["iterations"], (r.iterations, ref["iterations"])
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    rel = num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]]))
    assert rel < 1e-7, rel
    if accelerate:
        assert r.alpha == pytest.approx(ref["alpha"])


def test_assemble_2d_guards(require_gpu):
    """The 2-D assembly refuses a 3-D system, approximations other than SSA,
    L1L2 and DIVA, L1L2 without the rate factor and DIVA without its set-up."""
    from cism_amd import _lib
    from cism_amd.velo import AssemblyInfo, BPSolver
    import ctypes
    p = PROBLEMS["stream"]
    b3 = BPSolver(p)
    st = b3._L.cismgpu_system_assemble_2d(b3.sys.handle, ctypes.byref(b3._ain), 1, ctypes.byref(AssemblyInfo()))
    assert st == 3
    s = SSASolver(p)
    for approx, code in ((0, 11), (2, 11), (3, 3), (4, 3)):
        st = s._L.cismgpu_system_assemble_2d(s.sys.handle, ctypes.byref(s._ain), approx, ctypes.byref(AssemblyInfo()))
        assert st == code
    st = s._L.cismgpu_system_assemble_3d(s.sys.handle, ctypes.byref(s._ain), ctypes.byref(AssemblyInfo()))
    assert st == 3
    assert "3-D system" in _lib.last_error()
