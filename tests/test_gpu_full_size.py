"""Size-independent checks at the full 1024 x 1024 x 11 ice-stream size
(11.5 M owned nodes, 23 M unknowns), where the oracle cannot follow:

* the GPU-assembled matrix is exactly symmetric: its FP64 rows and their FP32
  rounding both take the half storage (SpMV kinds 4 and 3), and the FP64
  matvec gives (Ax, y) = (x, Ay) to rounding;
* the FP64 drop-in converges to 1e-10, and the true residual ||Ax - b|| / ||b||,
  recomputed by the separate residual kernel (compute_residual_vector_3d,
  glissade_velo_higher.F90:9693-9894), sits at that level;
* the mixed-precision solve is within 1e-6 of a tight FP64 solve, and
  repeating it gives the same answer bit for bit (fixed-order sums);
* the exchanging half-storage SpMV, the one that re-reads the partner rows
  and the full quad rows give bit-identical mixed solves.
"""
import os

import numpy as np
import pytest

from cism_amd import pcg, synthetic
from cism_amd.velo import BPSolver

pytestmark = pytest.mark.gpu

N, NZ, SEED = 1024, 11, 7


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def stream1024(require_gpu):
    print(f"\nfull-size ice stream {N}x{N}x{NZ} seed={SEED}", flush=True)
    prob = synthetic.ice_stream(N, nz=NZ, seed=SEED)
    bp = BPSolver(prob)
    bp.assemble(prob.uvel, prob.vvel)
    s = bp.sys
    s.zero_x()
    r = s.solve(precond=2, tolerance=1e-12, maxiters=50000)
    assert r["status"] == 0
    xref = np.stack(s.get_x())
    yield prob, bp, xref
    s.close()


def _mixed(s):
    s.zero_x()
    r = s.solve(precond=2, tolerance=1e-10, maxiters=50000, precision=pcg.MIXED_FP32, inner_tol=1e-6,
                max_outer=40)
    return r, np.stack(s.get_x())


def test_fp64_dropin_full_size(stream1024):
    prob, bp, xref = stream1024
    s = bp.sys
    s.zero_x()
    r = s.solve(precond=2, tolerance=1e-10, maxiters=50000)
    assert r["status"] == 0 and r["err"] <= 1e-10 and r["spmv_kind"] == 4
    x = np.stack(s.get_x())
    assert _rel(x, xref) < 1e-8
    L2, L2r = bp.residual()
    print(f"fp64: {r['niters']} it, err {r['err']:.3e}, true residual {L2r:.3e}, vs tight {_rel(x, xref):.3e}")
    assert L2r <= 2e-10


def test_matvec_symmetry_full_size(stream1024):
    """(Ax, y) = (x, Ay) for x, y supported three vertices inside the ice-free
    ring, where every row they touch is an active row."""
    prob, bp, _ = stream1024
    s, g = bp.sys, prob.grid
    rng = np.random.default_rng(11)
    sup = (slice(None), slice(g.ilo + 2, g.ihi - 3), slice(g.jlo + 2, g.jhi - 3))

    def rand():
        a = g.zeros_vec()
        a[sup] = rng.standard_normal(a[sup].shape)
        return a

    xu, xv, yu, yv = rand(), rand(), rand(), rand()
    s.set_x(xu, xv)
    axu, axv = s.matvec()
    s.set_x(yu, yv)
    ayu, ayv = s.matvec()
    lhs = float(np.sum(axu * yu) + np.sum(axv * yv))
    rhs = float(np.sum(xu * ayu) + np.sum(xv * ayv))
    scale = np.sqrt(np.sum(axu ** 2 + axv ** 2) * np.sum(yu ** 2 + yv ** 2))
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_mixed_within_1e6_and_reproducible(stream1024):
    prob, bp, xref = stream1024
    r1, x1 = _mixed(bp.sys)
    r2, x2 = _mixed(bp.sys)
    print(f"mixed: {r1['niters']} inner it, {r1['outer_iters']} outer, vs tight FP64 {_rel(x1, xref):.3e}")
    assert r1["status"] == 0 and r1["spmv_kind"] == 3
    assert _rel(x1, xref) <= 1e-6
    assert r2["niters"] == r1["niters"] and np.array_equal(x2, x1)


def test_fp64_and_mixed_against_the_oracle_full_size(stream1024):
    """The oracle's all-cores FP64 Chronopoulos-Gear solve of the exported
    1024^2 x 11 system to 1e-10 (about 20 s on 16 host threads): the GPU FP64
    drop-in takes the same number of iterations (within one convergence
    check) and both GPU solutions sit within the bar of the oracle's."""
    from oracle import pyoracle as po
    prob, bp, _ = stream1024
    s, g = bp.sys, prob.grid
    (Auu, Auv, Avu, Avv), (bu, bv) = s.get_matrix()
    ou, ov = g.zeros_vec(), g.zeros_vec()
    nt = min(os.cpu_count() or 1, 16)
    st, oerr, onit = po.pcg_solve(g, bp.active_vertex, Auu, Auv, Avu, Avv, bu, bv, ou, ov, precond=2,
                                  tolerance=1e-10, maxiters=5000, nthreads=nt)
    del Auu, Auv, Avu, Avv
    assert st == 0 and oerr <= 1e-10
    xo = np.stack([ou, ov])
    s.zero_x()
    r = s.solve(precond=2, tolerance=1e-10, maxiters=50000)
    x64 = np.stack(s.get_x())
    rm, xm = _mixed(s)
    e64, em = _rel(x64, xo), _rel(xm, xo)
    print(f"oracle ({nt} threads): {onit} it, err {oerr:.3e}; GPU fp64 {r['niters']} it, vs oracle {e64:.3e}; "
          f"mixed vs oracle {em:.3e}")
    assert abs(r["niters"] - onit) <= 5
    assert e64 < 1e-8 and em <= 1e-6


def test_spmv_variants_identical_full_size(stream1024):
    prob, bp, _ = stream1024
    r_x, x_x = _mixed(bp.sys)
    os.environ["CISMGPU_SPMV_SYMX"] = "0"
    try:
        r_s, x_s = _mixed(bp.sys)
    finally:
        os.environ.pop("CISMGPU_SPMV_SYMX", None)
    assert r_s["spmv_kind"] == 3 and r_s["niters"] == r_x["niters"] and np.array_equal(x_s, x_x)
    os.environ["CISMGPU_SPMV_FULL"] = "1"
    try:
        bq = BPSolver(prob)
        bq.assemble(prob.uvel, prob.vvel)
        r_q, x_q = _mixed(bq.sys)
    finally:
        os.environ.pop("CISMGPU_SPMV_FULL", None)
    bq.sys.close()
    assert r_q["spmv_kind"] == 1 and r_q["niters"] == r_x["niters"] and np.array_equal(x_q, x_x)
