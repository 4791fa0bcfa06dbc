"""Reproducible global sums on the GPU (reprosum = 1; cism_reprosum_calc,
cism_reprosum_mod.F90:237-1057, :1063-1818).

Bar: bit-exact.  The device sum of any set of doubles is the exact sum
rounded once to nearest-even: equal to math.fsum (Shewchuk's exactly
rounded sum) and to the oracle's integer-vector sum, for any order of the
summands.  FP64 solves with reprosum on follow the oracle's reprosum solves
(which use the same exact sums) to the drop-in's 1e-10 bar, with the same
iteration counts.
"""
import math

import numpy as np
import pytest

from cism_amd import _lib, pcg
from cism_amd.grid import BC_PERIODIC
from cism_amd.velo import BPSolver
from oracle import pyoracle as po
from systems import random_spd_system, stream_system

pytestmark = pytest.mark.gpu


def _sets():
    rng = np.random.default_rng(3)
    out = []
    for t in range(12):
        n = int(rng.integers(1, 5000))
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
        if t % 2:
            a = np.concatenate([a, -a[: n // 2], [5e-324, -3e-320, 2.0 ** -1022]])
        out.append(a)
    out.append(np.array([1e300, 1.0, -1e300]))
    out.append(np.array([1.0, 2.0 ** -53]))
    out.append(np.array([1.0, 2.0 ** -53, 2.0 ** -100]))
    out.append(np.array([]))
    # many summands of one magnitude: digits pile up in the same levels
    out.append(rng.uniform(0.0, 1.0, 3_000_000) * 1e-3)
    return out


def test_exact_sum_equals_fsum_and_oracle_in_any_order(require_gpu):
    rng = np.random.default_rng(4)
    for a in _sets():
        want = math.fsum(a) if a.size else 0.0
        got = _lib.reprosum_calc(a)[0]
        assert got == want, (a.size, got, want)
        assert got == po.reprosum(a)
        for _ in range(2):
            assert _lib.reprosum_calc(a[rng.permutation(a.size)])[0] == got
    # several fields (columns) in one call; non-finite summands give NaN
    cols = np.stack([np.r_[1e16, 1.0, -1e16, 3.0], np.r_[0.1, 0.2, 0.3, -0.6]], axis=1)
    out = _lib.reprosum_calc(cols)
    assert out[0] == 4.0 and out[1] == math.fsum(cols[:, 1])
    assert np.isnan(_lib.reprosum_calc(np.array([1.0, np.inf, -np.inf]))[0])


@pytest.mark.parametrize("algo", [pcg.ALGO_STANDARD, pcg.ALGO_CHRONGEAR])
@pytest.mark.parametrize("name", ["stream40", "random_bc0"])
def test_reprosum_solves_match_oracle(require_gpu, name, algo):
    if name == "stream40":
        p, s = stream_system()
        g, act, A, bu, bv = p.grid, s["active_vertex"], [s["Auu"], s["Auv"], s["Avu"], s["Avv"]], s["bu"], s["bv"]
    else:
        g, act, A, bu, bv = random_spd_system(bc=BC_PERIODIC)
    g.reprosum = 1
    try:
        ou, ov = g.zeros_vec(), g.zeros_vec()
        st, oerr, onit = po.pcg_solve(g, act, *A, bu, bv, ou, ov, algo="standard" if algo == 2 else "chrongear",
                                      precond=2, tolerance=1e-11, maxiters=3000)
        assert st == 0
        with pcg.System(g) as sys_:
            sys_.set_matrix(act, *A, bu, bv)
            sys_.set_x(g.zeros_vec(), g.zeros_vec())
            r = sys_.solve(algo=algo, precond=2, tolerance=1e-11, maxiters=3000)
            x = np.stack(sys_.get_x())
    finally:
        g.reprosum = 0
    assert r["niters"] == onit, (r["niters"], onit)
    xo = np.stack([ou, ov])
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10


def test_reprosum_residual_and_mixed_solve(require_gpu):
    """The residual vector's norms and the mixed solve's FP64 refinement
    with exact sums: norms equal the oracle's (exact sums on both sides, so
    only the residual vector's rounding differs), the mixed solve reaches
    its tolerance and the oracle's solution to the mixed bar."""
    p, s = stream_system()
    p.grid.reprosum = 1
    try:
        bp = BPSolver(p)
        bp.assemble(p.uvel, p.vvel)
        L2, L2r = bp.residual()
        (Auu, Auv, Avu, Avv), (bu, bv) = bp.sys.get_matrix()
        oL2, oL2r = po.residual_L2(p.grid, bp.active_vertex, [Auu, Auv, Avu, Avv], bu, bv, p.uvel, p.vvel)
        assert L2 == pytest.approx(oL2, rel=1e-13) and L2r == pytest.approx(oL2r, rel=1e-13)
        bp.sys.zero_x()
        r = bp.sys.solve(precond=2, tolerance=1e-10, maxiters=20000, precision=pcg.MIXED_FP32, inner_tol=1e-3,
                         max_outer=30)
        assert r["status"] == 0 and r["err"] < 1e-10
        xm = np.stack(bp.sys.get_x())
        ou, ov = p.grid.zeros_vec(), p.grid.zeros_vec()
        st, _, _ = po.pcg_solve(p.grid, bp.active_vertex, Auu, Auv, Avu, Avv, bu, bv, ou, ov, precond=2,
                                tolerance=1e-12, maxiters=20000)
        xo = np.stack([ou, ov])
        assert st == 0 and np.linalg.norm(xm - xo) / np.linalg.norm(xo) < 1e-6
    finally:
        p.grid.reprosum = 0
