"""GPU assembly of the BP system against the CPU restatement.

Bar: every owned matrix row, the rhs and the element viscosities agree to
1e-12 relative to the largest entry of the row (the element kernel sums in
the reference's element order; FMA contraction is the remaining
difference); the Picard loop on the dome reaches the same iteration count
and velocities within 1e-7.
"""
import numpy as np
import pytest

from cism_amd import synthetic
from cism_amd.velo import BPSolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _problems():
    return {
        "dome": synthetic.dome(),
        "stream": synthetic.ice_stream(40, nz=11, seed=3),
        "slab_periodic": synthetic.slab(n_cells=8, nz=9),
        "marine": synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0),
        "marine_local_front": _local_front(synthetic.marine_sheet(36, nz=5, dx=2500.0, seed=4, h_max=1000.0,
                                                                  h_shelf=200.0)),
    }


def _local_front(p):
    p.assemble_lateral = 1
    return p


PROBLEMS = _problems()


def _rowwise_close(gpu, ref, grid, act, tol):
    own = grid.owned_slice()
    a = gpu[:, own[0], own[1], own[2]]
    b = ref[:, own[0], own[1], own[2]].copy()
    # coefficients that couple to a level outside 1..nz are not stored on the device
    b[0:9, 0] = 0.0
    b[18:27, grid.nz - 1] = 0.0
    b *= np.asarray(act, bool)[own[1], own[2]][None, None]
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-300)
    err = np.abs(a - b) / scale
    return float(err.max())


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_assembly_matches_oracle(require_gpu, name):
    p = PROBLEMS[name]
    ref = po.assemble(p)
    s = BPSolver(p)
    info = s.assemble(p.uvel, p.vvel)
    (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
    act = ref["active_vertex"]
    for a, key in ((Auu, "Auu"), (Auv, "Auv"), (Avu, "Avu"), (Avv, "Avv")):
        e = _rowwise_close(a, ref[key], p.grid, act, 1e-12)
        assert e < 1e-12, (key, e)
    bscale = max(np.abs(ref["bu"]).max(), np.abs(ref["bv"]).max(), 1e-300)
    assert np.abs(bu - ref["bu"]).max() / bscale < 1e-13
    assert np.abs(bv - ref["bv"]).max() / bscale < 1e-13
    ef = s.efvs()
    assert np.abs(ef - ref["efvs"]).max() / max(np.abs(ref["efvs"]).max(), 1e-300) < 1e-12
    # count_nonzeros_3d: the device count equals the nonzeros of the exported
    # owned active rows; the pattern of entries above round-off matches the
    # restatement's (entries that cancel exactly in one evaluation order can
    # be a few ulps from zero in the other)
    own = p.grid.owned_slice()
    actb = np.asarray(act, bool)[own[1], own[2]][None, None]
    blocks_gpu = [a[:, own[0], own[1], own[2]] * actb for a in (Auu, Auv, Avu, Avv)]
    assert info["nnz"] == sum(int(np.count_nonzero(b)) for b in blocks_gpu)
    for bg, key in zip(blocks_gpu, ("Auu", "Auv", "Avu", "Avv")):
        br = ref[key][:, own[0], own[1], own[2]] * actb
        sc = np.abs(br).max(axis=0, keepdims=True) * 1e-12
        assert np.array_equal(np.abs(bg) > sc, np.abs(br) > sc), key
    assert info["asym_pairs"] == 0 and ref["nwarn_asym"] == 0


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_rim_symmetry_check_matches_full_check(require_gpu, name):
    """The default symmetry check and nonzero count (rim rows only, the
    other rows counted as they are written) against the check of every
    row (CISMGPU_SYMCHECK=full): the same matrix bit for bit, the same
    nnz and asymmetric-pair count, and the same storage choice and FP64
    solve (half storage is taken only for an exactly symmetric matrix)."""
    import os
    p = PROBLEMS[name]
    out = []
    for mode in (None, "full"):
        if mode:
            os.environ["CISMGPU_SYMCHECK"] = mode
        try:
            s = BPSolver(p)
            info = s.assemble(p.uvel, p.vvel)
            (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
            s.sys.zero_x()
            r = s.sys.solve(precond=2, tolerance=1e-9, maxiters=2000)
            x = np.stack(s.sys.get_x())
            out.append((info, np.stack([Auu, Auv, Avu, Avv]), r, x))
            s.sys.close()
        finally:
            os.environ.pop("CISMGPU_SYMCHECK", None)
    (i0, A0, r0, x0), (i1, A1, r1, x1) = out
    assert np.array_equal(A0, A1)
    assert i0["nnz"] == i1["nnz"] and i0["asym_pairs"] == i1["asym_pairs"]
    assert r0["spmv_kind"] == r1["spmv_kind"] and r0["niters"] == r1["niters"]
    assert np.array_equal(x0, x1)


@pytest.mark.parametrize("name", ["dome", "stream", "marine"])
def test_assembled_solve_matches_oracle(require_gpu, name):
    p = PROBLEMS[name]
    ref = po.assemble(p)
    g = p.grid
    A = [ref["Auu"], ref["Auv"], ref["Avu"], ref["Avv"]]
    ou, ov = g.zeros_vec(), g.zeros_vec()
    st, oerr, onit = po.pcg_solve(g, ref["active_vertex"], *A, ref["bu"], ref["bv"], ou, ov, precond=2,
                                  tolerance=1e-11, maxiters=3000)
    s = BPSolver(p)
    s.assemble(p.uvel, p.vvel)
    s.sys.set_x(g.zeros_vec(), g.zeros_vec())
    r = s.sys.solve(precond=2, tolerance=1e-11, maxiters=3000)
    xu, xv = s.sys.get_x()
    rel = np.linalg.norm(np.stack([xu - ou, xv - ov])) / np.linalg.norm(np.stack([ou, ov]))
    assert rel < 1e-9, rel


def test_dome_picard_matches_oracle(require_gpu):
    """Dome, no-slip base, standard PCG with diagonal preconditioning to
    1e-8, Picard to 1e-4 maximum relative speed change."""
    p = PROBLEMS["dome"]
    ref = po.picard(p, algo="standard", precond=1, linear_tol=1e-8, linear_maxiters=200, resid_target=1e-4)
    s = BPSolver(p)
    r = s.picard(algo=2, precond=1, linear_tol=1e-8, linear_maxiters=200, resid_target=1e-4)
    assert r.converged and ref["converged"]
    assert r.iterations == ref["iterations"]
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    rel = num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]]))
    assert rel < 1e-7, rel
    # a relative residual landing within rounding of linear_tol can move the
    # exit by one convergence check (5 iterations)
    assert all(abs(a - b) <= 5 for a, b in zip(r.linear_iters, ref["linear_iters"]))


@pytest.mark.parametrize("name", ["dome", "marine"])
def test_accelerated_picard_matches_oracle(require_gpu, name):
    """Picard acceleration (evaluate_accelerated_picard_3d): the device driver
    takes the same accept / extend / restore decisions as the restatement
    (same alpha sequence and iteration count) and reaches the same velocity."""
    p = PROBLEMS[name]
    kw = dict(linear_tol=1e-10, resid_target=1e-4, maxiter=60)
    ref = po.picard(p, algo="chrongear", precond=2, linear_maxiters=2000, accelerate=True, **kw)
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, linear_maxiters=2000, accelerate=True, **kw)
    assert r.converged and ref["converged"]
    assert r.iterations == ref["iterations"]
    assert r.alpha == ref["alpha"]
    num = np.linalg.norm(np.stack([r.uvel - ref["uvel"], r.vvel - ref["vvel"]]))
    rel = num / np.linalg.norm(np.stack([ref["uvel"], ref["vvel"]]))
    assert rel < 1e-7, rel


@pytest.mark.parametrize("whichresid", [0, 1, 2])
def test_device_residual_velocity_and_vector_ops(require_gpu, whichresid):
    """compute_residual_velocity_3d on the device (work slot against the
    resident iterate) equals the restatement; vector_op forms a + alpha b,
    a - b, copies and zeros exactly as the host formulas do."""
    import ctypes
    p = PROBLEMS["stream"]
    g = p.grid
    s = BPSolver(p)
    L = s._L
    rng = np.random.default_rng(4)
    u0, v0 = (np.asfortranarray(rng.standard_normal(g.vec_shape)) for _ in range(2))
    u1, v1 = (np.asfortranarray(rng.standard_normal(g.vec_shape)) for _ in range(2))
    u1[:, 5, 5] = 0.0
    v1[:, 5, 5] = 0.0                     # zero speed is skipped
    assert L.cismgpu_system_work_vectors(s.sys.handle, 3) == 0
    s.sys.set_x(u0, v0)
    s._vop(0, 0, -1)                      # slot 0 = x
    s.sys.set_x(u1, v1)
    out = ctypes.c_double()
    assert L.cismgpu_system_residual_velocity(s.sys.handle, whichresid, 0, ctypes.byref(out)) == 0
    ref = po.lib().orc_residual_velocity_3d(g.nx, g.ny, g.nz, g.ilo, g.ihi, g.jlo, g.jhi, whichresid,
                                            po.dp(u0), po.dp(v0), po.dp(u1), po.dp(v1))
    if whichresid == 2:
        assert abs(out.value - ref) <= 1e-14 * abs(ref)
    else:
        assert out.value == ref
    # x = slot0 + 0.7 (x - slot0), built through slots 1 and 2
    s._vop(2, 1, -1, 0)                   # slot 1 = x - slot 0
    s._vop(1, -1, 0, 1, 0.7)              # x = slot 0 + 0.7 slot 1
    xu, xv = s.sys.get_x()
    own = g.owned_slice()
    assert np.array_equal(xu[own], (u0 + 0.7 * (u1 - u0))[own])
    assert np.array_equal(xv[own], (v0 + 0.7 * (v1 - v0))[own])
    s._vop(3, -1)
    xu, xv = s.sys.get_x()
    assert not xu.any() and not xv.any()


@pytest.mark.parametrize("name", ["dome", "stream", "marine"])
def test_internal_stress_matches_oracle(require_gpu, name):
    """compute_internal_stress (glissade_velo_higher.F90:7046-7281) on the
    device after a Picard solve = the oracle's of the same velocity, each
    component within 1e-12 of its largest magnitude."""
    p = PROBLEMS[name]
    s = BPSolver(p)
    r = s.picard(algo=3, precond=2, linear_tol=1e-10, linear_maxiters=2000, resid_target=1e-3, maxiter=8)
    got = s.internal_stress()
    ref = po.internal_stress(p, r.uvel, r.vvel)
    for k, v in ref.items():
        scale = np.abs(v).max()
        assert scale > 0, k
        assert np.abs(got[k] - v).max() <= 1e-12 * scale, k
