"""GPU parity of the structured 3-D solvers against the CPU restatement.

Bar: the FP64 drop-in matches the restated reference solvers to 1e-10
relative in the solution on the same assembled system (tight tolerance),
with the same iteration count; the matvec matches to 1e-13 relative (FMA
contraction is the only difference).
"""
import numpy as np
import pytest

from cism_amd import pcg
from cism_amd._lib import CismGpuError
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC
from oracle import pyoracle as po
from systems import dome_system, random_spd_system, stream_system

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _systems():
    out = []
    for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW):
        g, act, A, bu, bv = random_spd_system(bc=bc)
        out.append((f"random_bc{bc}", g, act, A, bu, bv))
    p, s = dome_system()
    out.append(("dome", p.grid, s["active_vertex"], [s["Auu"], s["Auv"], s["Avu"], s["Avv"]], s["bu"], s["bv"]))
    p, s = stream_system()
    out.append(("stream40", p.grid, s["active_vertex"], [s["Auu"], s["Auv"], s["Avu"], s["Avv"]], s["bu"], s["bv"]))
    return out


SYSTEMS = {name: rest for name, *rest in _systems()}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_matvec_parity(require_gpu, name):
    g, act, A, bu, bv = SYSTEMS[name]
    rng = np.random.default_rng(1)
    xu = np.asfortranarray(rng.standard_normal(g.vec_shape))
    xv = np.asfortranarray(rng.standard_normal(g.vec_shape))
    yu, yv = pcg.matvec_multiply_structured_3d(g, act, *A, xu, xv)
    ou, ov = po.matvec(g, act, *A, xu, xv)
    assert _rel(yu, ou) < 1e-13 and _rel(yv, ov) < 1e-13
    # zero outside owned active vertices, exactly as the reference
    mask = np.zeros((g.nvx, g.nvy), bool)
    mask[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] = True
    mask &= np.asarray(act, bool)
    assert np.all(yu[:, ~mask] == 0.0) and np.all(yv[:, ~mask] == 0.0)


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("precond", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_solver_parity_tight(require_gpu, name, algo, precond):
    g, act, A, bu, bv = SYSTEMS[name]
    tol = 1e-11
    ox, oy = g.zeros_vec(), g.zeros_vec()
    st, oerr, onit = po.pcg_solve(g, act, *A, bu, bv, ox, oy, algo=algo, precond=precond, tolerance=tol,
                                  maxiters=2000)
    assert st == 0 and oerr < tol
    xu, xv = g.zeros_vec(), g.zeros_vec()
    f = pcg.pcg_solver_chrongear_3d if algo == "chrongear" else pcg.pcg_solver_standard_3d
    err, nit = f(g, act, *A, bu, bv, xu, xv, precond, 5, tol, 2000)
    assert err < tol
    assert abs(nit - onit) <= 5, (nit, onit)
    rel = _rel(np.stack([xu, xv]), np.stack([ox, oy]))
    assert rel < 1e-10, rel


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("name", ["random_bc0", "dome", "stream40"])
def test_solver_iterate_by_iterate(require_gpu, name, algo):
    """Fixed iteration counts (no convergence exit): the k-th iterate of the
    drop-in equals the restated reference's k-th iterate."""
    g, act, A, bu, bv = SYSTEMS[name]
    for maxit in (1, 2, 5, 7, 12):
        ox, oy = g.zeros_vec(), g.zeros_vec()
        st, oerr, onit = po.pcg_solve(g, act, *A, bu, bv, ox, oy, algo=algo, precond=2, tolerance=0.0,
                                      maxiters=maxit)
        xu, xv = g.zeros_vec(), g.zeros_vec()
        f = pcg.pcg_solver_chrongear_3d if algo == "chrongear" else pcg.pcg_solver_standard_3d
        err, nit = f(g, act, *A, bu, bv, xu, xv, 2, 5, 0.0, maxit)
        assert nit == onit == maxit
        assert _rel(np.stack([xu, xv]), np.stack([ox, oy])) < 1e-11
        if maxit >= 5:
            assert abs(err - oerr) <= 1e-9 * max(abs(oerr), 1e-300)
        else:
            assert np.isnan(err)   # no convergence check ran: err is left unassigned


def test_nan_alpha_is_an_error(require_gpu):
    g, act, A, bu, bv = SYSTEMS["random_bc0"]
    Z = [np.zeros_like(a) for a in A]
    xu, xv = g.zeros_vec(), g.zeros_vec()
    with pytest.raises(CismGpuError) as e:
        pcg.pcg_solver_chrongear_3d(g, act, *Z, bu, bv, xu, xv, 0, 5, 1e-8, 50)
    assert e.value.status == 1
    with pytest.raises(CismGpuError) as e:
        pcg.pcg_solver_standard_3d(g, act, *Z, bu, bv, xu, xv, 0, 5, 1e-8, 50)
    assert e.value.status == 1


def test_nonconvergence_is_silent(require_gpu):
    g, act, A, bu, bv = SYSTEMS["stream40"]
    xu, xv = g.zeros_vec(), g.zeros_vec()
    err, nit = pcg.pcg_solver_chrongear_3d(g, act, *A, bu, bv, xu, xv, 0, 5, 1e-14, 10)
    assert nit == 10 and err > 1e-14


def test_resident_system_roundtrip(require_gpu):
    g, act, A, bu, bv = SYSTEMS["stream40"]
    with pcg.System(g) as s:
        s.set_matrix(act, *A, bu, bv)
        (Auu, Auv, Avu, Avv), (b1, b2) = s.get_matrix()
        own = g.owned_slice()
        for a, ref in zip((Auu, Auv, Avu, Avv), A):
            # owned active rows are returned exactly; coefficients to levels outside 1..nz are dropped
            refm = ref.copy()
            refm[0:9, 0] = 0.0
            refm[18:27, g.nz - 1] = 0.0
            refm *= np.asarray(act, bool)[None, None]
            assert np.array_equal(a[:, own[0], own[1], own[2]], refm[:, own[0], own[1], own[2]])
        assert np.array_equal(b1, bu) and np.array_equal(b2, bv)


_TIGHT = {}


def _oracle_tight(name, precond=2):
    """The restated FP64 Chronopoulos-Gear solve of a test system to 1e-12."""
    key = (name, precond)
    if key not in _TIGHT:
        g, act, A, bu, bv = SYSTEMS[name]
        ou, ov = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *A, bu, bv, ou, ov, precond=precond, tolerance=1e-12, maxiters=20000)
        assert st == 0 and err < 1e-12, (name, st, err, nit)
        _TIGHT[key] = np.stack([ou, ov])
    return _TIGHT[key]


@pytest.mark.parametrize("precision", [pcg.MIXED_FP32, pcg.MIXED_BF16])
@pytest.mark.parametrize("name,precond", [("stream40", 2), ("dome", 2), ("random_bc0", 2), ("stream40", 3)])
def test_mixed_precision_matches_oracle(require_gpu, name, precond, precision):
    """Mixed solve (FP32 or BF16 storage, FP32 inner C-G, FP64 refinement to
    1e-10) against the oracle's FP64 solve to 1e-12: relative velocity error
    at most 1e-6 (the bar the mixed form is accepted on); the GPU FP64 solve
    of the same system sits within 1e-9 of the oracle."""
    g, act, A, bu, bv = SYSTEMS[name]
    xo = _oracle_tight(name, precond)
    with pcg.System(g) as s:
        s.set_matrix(act, *A, bu, bv)
        s.set_x(g.zeros_vec(), g.zeros_vec())
        r64 = s.solve(precond=precond, tolerance=1e-12, maxiters=20000)
        x64 = np.stack(s.get_x())
        s.set_x(g.zeros_vec(), g.zeros_vec())
        rm = s.solve(precond=precond, tolerance=1e-10, maxiters=20000, precision=precision, inner_tol=1e-3,
                     max_outer=30)
        xm = np.stack(s.get_x())
    assert rm["status"] == 0 and rm["err"] < 1e-10
    e_mixed, e64 = _rel(xm, xo), _rel(x64, xo)
    print(f"{name} precond {precond} precision {precision}: mixed vs oracle {e_mixed:.2e}, fp64 vs oracle {e64:.2e}")
    assert e_mixed < 1e-6
    assert e64 < 1e-9


def _mixed_run(g, act, A, bu, bv, precision, full):
    import os
    os.environ["CISMGPU_SPMV_FULL"] = "1" if full else "0"
    try:
        with pcg.System(g) as s:
            s.set_matrix(act, *A, bu, bv)
            s.set_x(g.zeros_vec(), g.zeros_vec())
            r = s.solve(precond=2, tolerance=1e-10, maxiters=3000, precision=precision, inner_tol=1e-3,
                        max_outer=30)
            return r, np.stack(s.get_x()), s.bench_spmv(kind=1 if precision == pcg.MIXED_FP32 else 2, nrep=3)
    finally:
        os.environ.pop("CISMGPU_SPMV_FULL", None)


@pytest.mark.parametrize("name", ["stream40", "dome", "random_bc0", "random_bc3"])
def test_half_storage_spmv_gives_identical_solves(require_gpu, name):
    """An exactly symmetric FP32 matrix is stored as its centre and 13 upper
    couplings; the half-storage SpMV sums the same products in the same
    order, so the mixed solve is bit-identical to the full-storage one and
    streams 224 matrix bytes per node instead of 432."""
    g, act, A, bu, bv = SYSTEMS[name]
    rf, xf, (msf, byf) = _mixed_run(g, act, A, bu, bv, pcg.MIXED_FP32, full=True)
    rh, xh, (msh, byh) = _mixed_run(g, act, A, bu, bv, pcg.MIXED_FP32, full=False)
    assert rf["spmv_kind"] == 1 and rh["spmv_kind"] == 3
    assert rh["niters"] == rf["niters"] and rh["outer_iters"] == rf["outer_iters"]
    assert np.array_equal(xh, xf)
    nodes = g.n_owned_nodes
    assert byf == nodes * (108 * 4 + 16) and byh == nodes * (56 * 4 + 16)


@pytest.mark.parametrize("name", ["stream40", "dome", "random_bc0", "random_bc3"])
def test_half_storage_exchange_kernel_identical(require_gpu, name):
    """The half-storage SpMV that passes the partner slots from level to level
    through LDS and lane shuffles gives the same solve, bit for bit, as the one
    that re-reads them from memory (CISMGPU_SPMV_SYMX=0)."""
    import os
    g, act, A, bu, bv = SYSTEMS[name]
    out = []
    for flag in ("0", "1"):
        os.environ["CISMGPU_SPMV_SYMX"] = flag
        try:
            out.append(_mixed_run(g, act, A, bu, bv, pcg.MIXED_FP32, full=False))
        finally:
            os.environ.pop("CISMGPU_SPMV_SYMX", None)
    (r0, x0, _), (r1, x1, _) = out
    assert r0["spmv_kind"] == 3 and r1["spmv_kind"] == 3
    assert r1["niters"] == r0["niters"] and r1["outer_iters"] == r0["outer_iters"]
    assert np.array_equal(x1, x0)


def test_half_storage_needs_exact_symmetry(require_gpu):
    """A coupling whose transposed partner differs after rounding keeps the
    full quad storage (spmv_kind 1)."""
    g, act, A, bu, bv = SYSTEMS["random_bc3"]
    B = [a.copy(order="F") for a in A]
    own = g.owned_slice()
    i0, j0 = g.ilo + 3, g.jlo + 3
    B[1][14, 2, i0 - 1, j0 - 1] += 1e-3 * abs(B[0][13, 2, i0 - 1, j0 - 1])   # Auv, neighbour (+1, 0, 0)
    r, x, _ = _mixed_run(g, act, B, bu, bv, pcg.MIXED_FP32, full=False)
    assert r["spmv_kind"] == 1


@pytest.mark.parametrize("algo", [pcg.ALGO_STANDARD, pcg.ALGO_CHRONGEAR])
@pytest.mark.parametrize("name", ["stream40", "random_bc0", "random_bc1"])
def test_fp64_half_storage_gives_identical_solves(require_gpu, name, algo):
    """The FP64 solves use half (symmetric) storage when the FP64 matrix is
    exactly symmetric: identical iterates, 448 instead of 864 matrix bytes
    per node."""
    import os
    g, act, A, bu, bv = SYSTEMS[name]
    out = {}
    for full in ("1", "0"):
        os.environ["CISMGPU_SPMV_FULL"] = full
        try:
            with pcg.System(g) as s:
                s.set_matrix(act, *A, bu, bv)
                s.set_x(g.zeros_vec(), g.zeros_vec())
                r = s.solve(algo=algo, precond=2, tolerance=1e-11, maxiters=3000)
                out[full] = (r, np.stack(s.get_x()), s.bench_spmv(kind=0, nrep=3), s.matvec())
        finally:
            os.environ.pop("CISMGPU_SPMV_FULL", None)
    (rf, xf, (_, byf), yf), (rh, xh, (_, byh), yh) = out["1"], out["0"]
    assert rh["niters"] == rf["niters"] and rh["err"] == rf["err"]
    assert np.array_equal(xh, xf)
    assert np.array_equal(yh[0], yf[0]) and np.array_equal(yh[1], yf[1])
    nodes = g.n_owned_nodes
    assert byf == nodes * (108 * 8 + 32) and byh == nodes * (56 * 8 + 32)
