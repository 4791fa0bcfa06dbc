"""Batched ensemble solve (one persistent launch for all members).

Bar: every member's solution agrees with the CPU restatement's FP64
Chronopoulos-Gear solve of the same assembled system (tolerance 1e-12) to
1e-6 in the relative 2-norm, for the FP32 and the BF16-matrix inner solve;
the launch is deterministic (graph replays give bit-identical solutions)
and the autotuned work split gives the same answers to the same bar.
"""
import numpy as np
import pytest

from cism_amd import synthetic
from cism_amd._lib import CismGpuError
from cism_amd.ensemble import Ensemble
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _members():
    return synthetic.ensemble(n_members=8, seed=5, nmin=20, nmax=44, nz_choices=(5, 7, 11, 15))


@pytest.fixture(scope="module")
def ens(require_gpu):
    probs = _members()
    E = Ensemble(probs, keep_members=True)
    refs = []
    for p, s in zip(probs, E.solvers):
        (Auu, Auv, Avu, Avv), (bu, bv) = s.sys.get_matrix()
        xu, xv = p.grid.zeros_vec(), p.grid.zeros_vec()
        st, err, nit = po.pcg_solve(p.grid, s.active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv,
                                    algo="chrongear", precond=2, ncheck=1, tolerance=1e-12, maxiters=5000)
        assert st == 0 and err < 1e-12, (st, err, nit)
        refs.append((xu, xv))
    yield E, refs
    E.close()


def _rel(E, refs, m):
    xu, xv = E.get_x(m)
    ru, rv = refs[m]
    own = E.problems[m].grid.owned_slice()
    num = np.sqrt(((xu - ru)[own] ** 2).sum() + ((xv - rv)[own] ** 2).sum())
    den = np.sqrt((ru[own] ** 2).sum() + (rv[own] ** 2).sum())
    return num / den


@pytest.mark.parametrize("bf16", [False, True])
def test_ensemble_matches_oracle(ens, bf16):
    E, refs = ens
    info = E.solve(tolerance=1e-10, inner_tol=1e-3, inner_maxiters=400, max_outer=30, bf16=bf16)
    assert info["converged"] == len(refs) and info["barrier_error"] == 0, info
    assert info["max_err"] < 1e-10
    for m in range(len(refs)):
        mi = E.member_info(m)
        assert mi["flags"] & 1 and mi["err"] < 1e-10, (m, mi)
        e = _rel(E, refs, m)
        assert e < TOL, (m, e)


def test_ensemble_deterministic_replay(ens):
    E, refs = ens
    E.solve()
    first = [E.get_x(m) for m in range(len(refs))]
    for _ in range(2):
        # eager work on the same device between graph replays
        E.member_info(0)
        E.solve()
        for m in range(len(refs)):
            xu, xv = E.get_x(m)
            assert np.array_equal(xu, first[m][0]) and np.array_equal(xv, first[m][1]), m


def test_ensemble_autotune(ens):
    E, refs = ens
    split, ms = E.autotune()
    assert all(s in (1, 2, 4) for s in split)
    assert ms.shape == (3, 3) and np.all(ms >= 0.0)
    info = E.solve()
    assert info["converged"] == len(refs)
    for m in range(len(refs)):
        assert _rel(E, refs, m) < TOL


def test_ensemble_half_storage_identical(require_gpu):
    """Every member's FP32 matrix is exactly symmetric, so the products keep
    each row's centre and 13 upper couplings (224 B per node instead of 432)
    and read the lower couplings transposed from the partner rows; they sum
    the same products in the same order, so the solve is bit-identical to the
    one on full quad rows (CISMGPU_SPMV_FULL=1)."""
    import os
    out = []
    for flag in ("1", "0"):
        os.environ["CISMGPU_SPMV_FULL"] = flag
        try:
            E = Ensemble(_members())
        finally:
            os.environ.pop("CISMGPU_SPMV_FULL", None)
        try:
            info = E.solve(tolerance=1e-10, inner_tol=1e-3, inner_maxiters=400, max_outer=30)
            xs = [np.stack(E.get_x(m)) for m in range(len(E.problems))]
            ib = E.solve(tolerance=1e-10, inner_tol=1e-3, inner_maxiters=400, max_outer=30, bf16=True)
            out.append((info, xs, ib["matrix_storage"]))
        finally:
            E.close()
    (i_full, x_full, b_full), (i_half, x_half, b_half) = out
    assert i_full["matrix_storage"] == 1 and i_half["matrix_storage"] == 3 and b_full == b_half == 2
    assert i_half["converged"] == i_full["converged"] == len(x_full)
    assert i_half["steps"] == i_full["steps"] and i_half["total_inner"] == i_full["total_inner"]
    for a, b in zip(x_half, x_full):
        assert np.array_equal(a, b)


def test_ensemble_rejects_deep_columns(require_gpu):
    p = synthetic.ice_stream(20, nz=17, seed=1)
    with pytest.raises(CismGpuError):
        Ensemble([p])
