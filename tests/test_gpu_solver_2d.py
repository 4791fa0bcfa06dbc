"""GPU parity of the structured 2-D solvers (pcg_solver_standard_2d,
pcg_solver_chrongear_2d, matvec_multiply_structured_2d) against the CPU
restatement.

Systems: random SPD 9-point systems under three halo rules, and the
depth-integrated (P^T A P) forms of the assembled dome and ice-stream
systems.  Bar as for the 3-D drop-ins: solution within 1e-10 relative of the
restated reference at a tight tolerance, iteration counts within one check
interval, k-th iterates equal, matvec within 1e-13.
"""
import numpy as np
import pytest

from cism_amd import pcg
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC
from oracle import pyoracle as po
from systems import collapse_to_2d, dome_system, random_spd_system_2d, stream_system

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _systems():
    out = {}
    for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW):
        out[f"random2d_bc{bc}"] = random_spd_system_2d(bc=bc)
    for name, (p, s) in (("dome2d", dome_system()), ("stream2d", stream_system())):
        act = s["active_vertex"]
        A2, bu, bv = collapse_to_2d(p.grid, act, [s["Auu"], s["Auv"], s["Avu"], s["Avv"]], s["bu"], s["bv"])
        out[name] = (p.grid, act, A2, bu, bv)
    return out


SYSTEMS = _systems()


def _zeros(g):
    return np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_matvec_2d_parity(require_gpu, name):
    g, act, A, bu, bv = SYSTEMS[name]
    rng = np.random.default_rng(2)
    xu = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    xv = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    yu, yv = pcg.matvec_multiply_structured_2d(g, act, *A, xu, xv)
    ou, ov = po.matvec_2d(g, act, *A, xu, xv)
    assert _rel(yu, ou) < 1e-13 and _rel(yv, ov) < 1e-13
    mask = np.zeros((g.nvx, g.nvy), bool)
    mask[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] = True
    mask &= np.asarray(act, bool)
    assert np.all(yu[~mask] == 0.0) and np.all(yv[~mask] == 0.0)


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("precond", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_solver_2d_parity_tight(require_gpu, name, algo, precond):
    g, act, A, bu, bv = SYSTEMS[name]
    tol = 1e-11
    ox, oy = _zeros(g)
    st, oerr, onit = po.pcg_solve_2d(g, act, *A, bu, bv, ox, oy, algo=algo, precond=precond, tolerance=tol,
                                     maxiters=4000)
    assert st == 0 and oerr < tol
    xu, xv = _zeros(g)
    f = pcg.pcg_solver_chrongear_2d if algo == "chrongear" else pcg.pcg_solver_standard_2d
    err, nit = f(g, act, *A, bu, bv, xu, xv, precond, 5, tol, 4000)
    assert err < tol
    assert abs(nit - onit) <= 5, (nit, onit)
    rel = _rel(np.stack([xu, xv]), np.stack([ox, oy]))
    assert rel < 1e-10, rel


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("precond", [1, 3, 4])
@pytest.mark.parametrize("name", ["random2d_bc0", "stream2d"])
def test_solver_2d_iterate_by_iterate(require_gpu, name, algo, precond):
    g, act, A, bu, bv = SYSTEMS[name]
    for maxit in (1, 2, 5, 9):
        ox, oy = _zeros(g)
        st, oerr, onit = po.pcg_solve_2d(g, act, *A, bu, bv, ox, oy, algo=algo, precond=precond, tolerance=0.0,
                                         maxiters=maxit)
        xu, xv = _zeros(g)
        f = pcg.pcg_solver_chrongear_2d if algo == "chrongear" else pcg.pcg_solver_standard_2d
        err, nit = f(g, act, *A, bu, bv, xu, xv, precond, 5, 0.0, maxit)
        assert nit == onit == maxit
        assert _rel(np.stack([xu, xv]), np.stack([ox, oy])) < 1e-11


def test_resident_2d_system_roundtrip_and_guards(require_gpu):
    g, act, A, bu, bv = SYSTEMS["stream2d"]
    with pcg.System(g, two_d=True) as s:
        s.set_matrix(act, *A, bu, bv)
        (Auu, Auv, Avu, Avv), (b1, b2) = s.get_matrix()
        il, jl = slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi)
        am = np.asarray(act, bool)
        for a, ref in zip((Auu, Auv, Avu, Avv), A):
            assert a.shape == (27, 1, g.nvx, g.nvy)
            assert np.all(a[0:9] == 0.0) and np.all(a[18:27] == 0.0)
            want = np.moveaxis(ref, 2, 0) * am[None]
            assert np.array_equal(a[9:18, 0, il, jl], want[:, il, jl])
        assert np.array_equal(b1, bu) and np.array_equal(b2, bv)
        # the 3-D loader refuses a 2-D system
        m = pcg._mask(act)
        st = s._L.cismgpu_system_set_matrix(s.handle, m.ctypes.data, *(a.ctypes.data for a in A), bu.ctypes.data,
                                            bv.ctypes.data, 0)
        assert st == 3


@pytest.mark.parametrize("precision", [pcg.MIXED_FP32, pcg.MIXED_BF16])
@pytest.mark.parametrize("precond", [1, 3])
def test_mixed_precision_2d_matches_oracle(require_gpu, precision, precond):
    """2-D mixed solve against the oracle's FP64 2-D C-G to 1e-12."""
    g, act, A, bu, bv = SYSTEMS["stream2d"]
    ou, ov = _zeros(g)
    st, oerr, _ = po.pcg_solve_2d(g, act, *A, bu, bv, ou, ov, algo="chrongear", precond=precond, tolerance=1e-12,
                                  maxiters=20000)
    assert st == 0 and oerr < 1e-12
    xo = np.stack([ou, ov])
    with pcg.System(g, two_d=True) as s:
        s.set_matrix(act, *A, bu, bv)
        s.set_x(*_zeros(g))
        r64 = s.solve(precond=precond, tolerance=1e-12, maxiters=5000)
        x64 = np.stack(s.get_x())
        s.set_x(*_zeros(g))
        rm = s.solve(precond=precond, tolerance=1e-10, maxiters=3000, precision=precision, inner_tol=1e-3,
                     max_outer=30)
        xm = np.stack(s.get_x())
    assert r64["err"] < 1e-12 and rm["err"] < 1e-10
    assert _rel(xm, xo) < 1e-8 and _rel(x64, xo) < 1e-10


def test_2d_spmv_reads_one_plane(require_gpu):
    """The 2-D SpMV streams 4 x 9 coefficients per vertex, not 4 x 27."""
    g, act, A, bu, bv = SYSTEMS["stream2d"]
    with pcg.System(g, two_d=True) as s:
        s.set_matrix(act, *A, bu, bv)
        ms, by = s.bench_spmv(kind=0, nrep=5)
    nodes = (g.ihi - g.ilo + 1) * (g.jhi - g.jlo + 1)
    assert by == nodes * (36 * 8 + 4 * 8)
    assert ms > 0
