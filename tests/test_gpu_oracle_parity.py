"""GPU results against the oracle where earlier checks compared the GPU with
itself: the residual vector (compute_residual_vector_3d), the Picard solve
with mixed-precision inner solves on the marine sheet, ensemble members in
the larger autotuner size buckets, and a solve that starts from the
EISMINT-2 state the reference's tests hold (tests/golden/mound.nc).
"""
import ctypes
import os

import numpy as np
import pytest

from cism_amd import _lib, cismio, pcg, synthetic
from cism_amd.ensemble import Ensemble
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW
from cism_amd.velo import BPSolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _marine():
    return synthetic.marine_sheet(40, nz=7, dx=2000.0, seed=3, h_max=1000.0, h_shelf=200.0)


# ----------------------------------------------------------------- A13 residual vector
@pytest.mark.parametrize("make", [synthetic.dome, lambda: synthetic.ice_stream(40, nz=11, seed=3), _marine],
                         ids=["dome", "stream", "marine"])
def test_residual_vector_matches_oracle(require_gpu, make):
    """cismgpu_system_residual (compute_residual_vector_3d,
    glissade_velo_higher.F90:9693-9894): A u - b per owned active node, its
    L2 norm and L2 / |b| equal the oracle's on the same assembled system and
    velocity (1e-12, FMA contraction being the only difference)."""
    p = make()
    g = p.grid
    bp = BPSolver(p)
    bp.assemble(p.uvel, p.vvel)
    (Auu, Auv, Avu, Avv), (bu, bv) = bp.sys.get_matrix()
    rng = np.random.default_rng(5)
    u = np.asfortranarray(p.uvel * (1.0 + 0.1 * rng.standard_normal(g.vec_shape)))
    v = np.asfortranarray(p.vvel * (1.0 + 0.1 * rng.standard_normal(g.vec_shape)))
    bp.sys.set_x(u, v)
    L2, L2r, ru, rv = bp.residual(want_vectors=True)
    A = [Auu, Auv, Avu, Avv]
    oL2, oL2r = po.residual_L2(g, bp.active_vertex, A, bu, bv, u, v)
    assert L2 == pytest.approx(oL2, rel=1e-12) and L2r == pytest.approx(oL2r, rel=1e-12)
    yu, yv = po.matvec(g, bp.active_vertex, *A, u, v)
    own = g.owned_slice()
    act = np.zeros((g.nvx, g.nvy), bool)
    act[own[1], own[2]] = np.asarray(bp.active_vertex, bool)[own[1], own[2]]
    for got, y, b in ((ru, yu, bu), (rv, yv, bv)):
        want = np.where(act[None], y - b, 0.0)
        scale = max(np.abs(y).max(), np.abs(b).max())
        assert np.abs(got - want).max() <= 1e-12 * scale
        assert not got[:, ~act].any()


def test_residual_relative_to_zero_rhs_is_zero(require_gpu):
    """With b = 0 the relative residual is 0, as the reference sets it
    (glissade_velo_higher.F90:9881-9885), not a division by zero."""
    p = synthetic.ice_stream(24, nz=5, seed=2)
    g = p.grid
    bp = BPSolver(p)
    bp.assemble(p.uvel, p.vvel)
    (Auu, Auv, Avu, Avv), _ = bp.sys.get_matrix()
    z = g.zeros_vec()
    bp.sys.set_matrix(bp.active_vertex, Auu, Auv, Avu, Avv, z, z.copy(order="F"))
    bp.sys.set_x(p.uvel, p.vvel)
    L2, L2r = bp.residual()
    oL2, oL2r = po.residual_L2(g, bp.active_vertex, [Auu, Auv, Avu, Avv], z, z, p.uvel, p.vvel)
    assert L2 > 0 and L2r == 0.0 and oL2r == 0.0
    assert L2 == pytest.approx(oL2, rel=1e-12)


# ----------------------------------------------------------------- config 5: mixed inner solves in Picard
def test_accelerated_picard_mixed_inner_matches_oracle_fp64(require_gpu):
    """Accelerated Picard on the marine sheet (grounded interior, floating
    shelf, calving-front pressure) with FP32-stored mixed-precision inner
    solves and the local tridiagonal preconditioner, against the oracle's
    FP64 Picard (glissade_velo_higher.F90:2163-3663, solver call
    :3288-3300): the same number of Picard iterations (within one), the same
    acceleration ladder and velocities within 1e-6; every linear solve
    reaches its tolerance (status 0)."""
    p = _marine()
    kw = dict(linear_tol=1e-10, resid_target=1e-4, maxiter=60)
    ref = po.picard(p, algo="chrongear", precond=3, linear_maxiters=4000, accelerate=True, **kw)
    r = BPSolver(p).picard(algo=3, precond=3, linear_maxiters=4000, accelerate=True, precision=pcg.MIXED_FP32,
                           inner_tol=1e-3, max_outer=30, **kw)
    print(f"oracle: {ref['iterations']} Picard it, alpha {ref['alpha']}; GPU mixed: {r.iterations} it, "
          f"alpha {r.alpha}, linear iters {r.linear_iters}, outer {r.linear_outer}")
    assert r.converged and ref["converged"]
    assert all(st == 0 for st in r.linear_status), r.linear_status
    assert abs(r.iterations - ref["iterations"]) <= 1
    n = min(len(r.alpha), len(ref["alpha"]))
    assert r.alpha[:n] == ref["alpha"][:n]
    assert _rel(np.stack([r.uvel, r.vvel]), np.stack([ref["uvel"], ref["vvel"]])) <= 1e-6


def test_mixed_inner_budget_is_the_linear_maxiters(require_gpu):
    """maxiters bounds the inner iterations summed over the refinement
    steps of a mixed solve (as it bounds an FP64 solve's iterations): a
    budget too small for the tolerance stops the solve with status 12 and at
    most maxiters inner iterations; the Picard loop records the status and
    goes on, as the reference goes on after a linear solve that reaches its
    maxiters."""
    p = _marine()
    bp = BPSolver(p)
    bp.assemble(p.uvel, p.vvel)
    bp.sys.zero_x()
    r = bp.sys.solve(precond=3, tolerance=1e-12, maxiters=37, precision=pcg.MIXED_FP32, inner_tol=1e-6,
                     max_outer=30, raise_on_error=False)
    assert r["status"] == 12 and 0 < r["niters"] <= 37
    res = BPSolver(p).picard(precond=3, linear_tol=1e-12, linear_maxiters=37, precision=pcg.MIXED_FP32,
                             inner_tol=1e-6, max_outer=30, maxiter=3, resid_target=1e-30)
    assert res.iterations == 3 and res.linear_status == [12, 12, 12]
    assert all(n <= 37 for n in res.linear_iters)


# ----------------------------------------------------------------- ensemble: larger size buckets
def test_ensemble_large_buckets_match_oracle(require_gpu):
    """Members in the autotuner's second (81..136 owned vertices per side)
    and third (larger) size buckets, nz 11 and 15, FP32 and BF16 storage,
    against the oracle's FP64 solve of each member's system (1e-12).  With
    FP32 storage every member converges.  BF16 storage (8-bit mantissa) is
    not a contraction for every member: refinement around a BF16 operator
    converges only while the storage error times the condition number stays
    below one, and the stiffest member here (A = 1e-17, 172 x 172 x 15)
    does not converge; the launch reports it (status 12, its flags) and the
    members that converge meet the bar."""
    probs = [synthetic.ice_stream(100, nz=11, seed=21, flwa=3e-17),
             synthetic.ice_stream(104, nz=15, seed=22, flwa=2e-16),
             synthetic.ice_stream(168, nz=11, seed=23, flwa=5e-16),
             synthetic.ice_stream(172, nz=15, seed=24, flwa=1e-17)]
    E = Ensemble(probs, keep_members=True)
    try:
        nt = min(os.cpu_count() or 1, 16)
        refs = []
        for p, s in zip(probs, E.solvers):
            (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
            xu, xv = p.grid.zeros_vec(), p.grid.zeros_vec()
            st, err, nit = po.pcg_solve(p.grid, s.active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond=2,
                                        tolerance=1e-12, maxiters=20000, nthreads=nt)
            assert st == 0 and err < 1e-12, (st, err, nit)
            refs.append(np.stack([xu, xv]))
        split, _ = E.autotune()
        print(f"autotuned split per bucket: {split}")
        for bf16 in (False, True):
            info = E.solve(tolerance=1e-10, inner_tol=1e-3, inner_maxiters=1000, max_outer=30, bf16=bf16,
                           raise_on_error=False)
            mi = [E.member_info(m) for m in range(len(probs))]
            print(f"bf16={bf16}: {info}; members {mi}")
            assert info["barrier_error"] == 0, info
            if not bf16:
                assert info["status"] == 0 and info["converged"] == len(probs), _lib.last_error()
            else:
                assert info["converged"] >= 3 and info["status"] == (0 if info["converged"] == len(probs) else 12)
            for m, p in enumerate(probs):
                if not mi[m]["flags"] & 1:
                    continue
                own = p.grid.owned_slice()
                x = np.stack(E.get_x(m))
                e = _rel(x[(slice(None),) + own], refs[m][(slice(None),) + own])
                assert e < 1e-6, (m, bf16, e)
    finally:
        E.close()


# ----------------------------------------------------------------- a state from the reference's test data
def _mound_problem():
    """The EISMINT-2 mound bed (61 x 61 cells of 25 km) with a parabolic ice
    cap of 2500 m and radius 600 km centred on the mound, no-slip base."""
    p = cismio.read_cism_state(os.path.join(GOLDEN, "mound.nc"), nz=6, global_bc=BC_OUTFLOW, whichbabc=6)
    xc, yc = p.cell_centres()
    cx, cy = xc[p.grid.nx // 2], yc[p.grid.ny // 2]
    r = np.hypot(xc[:, None] - cx, yc[None, :] - cy)
    p.thck = np.asfortranarray(np.where(r < 600e3, 2500.0 * np.sqrt(np.clip(1.0 - (r / 600e3) ** 2, 0, 1)), 0.0))
    p.thck[:2, :] = p.thck[-2:, :] = 0.0
    p.thck[:, :2] = p.thck[:, -2:] = 0.0
    p.flwa[:] = 1e-16
    return p


def test_solve_from_eismint2_mound_state_matches_oracle(require_gpu):
    """A velocity problem read from the reference's mound.nc (bed) with an
    ice cap on it: the GPU assembly equals the oracle's rows (1e-12), the
    FP64 solve of the assembled system equals the oracle's solve (1e-9), and
    three Picard iterations give the oracle's velocities (1e-7)."""
    p = _mound_problem()
    g = p.grid
    ref = po.assemble(p)
    bp = BPSolver(p)
    bp.assemble(p.uvel, p.vvel)
    (Auu, Auv, Avu, Avv), (bu, bv) = bp.sys.get_matrix()
    own = g.owned_slice()
    act = np.asarray(ref["active_vertex"], bool)[own[1], own[2]]
    assert act.sum() > 100
    for a, key in ((Auu, "Auu"), (Auv, "Auv"), (Avu, "Avu"), (Avv, "Avv")):
        b = ref[key][:, own[0], own[1], own[2]].copy()
        b[0:9, 0] = 0.0
        b[18:27, g.nz - 1] = 0.0
        b *= act[None, None]
        sc = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-300)
        assert (np.abs(a[:, own[0], own[1], own[2]] - b) / sc).max() < 1e-12, key
    ou, ov = g.zeros_vec(), g.zeros_vec()
    A = [ref["Auu"], ref["Auv"], ref["Avu"], ref["Avv"]]
    st, oerr, onit = po.pcg_solve(g, ref["active_vertex"], *A, ref["bu"], ref["bv"], ou, ov, precond=2,
                                  tolerance=1e-11, maxiters=5000)
    assert st == 0
    bp.sys.zero_x()
    r = bp.sys.solve(precond=2, tolerance=1e-11, maxiters=5000)
    assert abs(r["niters"] - onit) <= 5
    assert _rel(np.stack(bp.sys.get_x()), np.stack([ou, ov])) < 1e-9
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, maxiter=3, resid_target=1e-30)
    pr = po.picard(p, algo="chrongear", precond=2, **kw)
    gr = BPSolver(p).picard(algo=3, precond=2, **kw)
    assert gr.iterations == pr["iterations"] == 3
    assert _rel(np.stack([gr.uvel, gr.vvel]), np.stack([pr["uvel"], pr["vvel"]])) < 1e-7
