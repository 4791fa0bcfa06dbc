"""Edge cases of the GPU solvers against the CPU restatement: deep columns
(the 16-, 32- and 64-level buckets of the fused column kernels, up to the
64-level limit), non-square and odd-sized owned blocks, two-level columns,
every global BC; a zero right-hand side (the reference's NaN abort,
status 1); and the nz limit.

Bar: FP64 solutions 1e-10 relative with the same iteration count; the
mixed solve 1e-6 of the FP64 solve.
"""
import ctypes

import numpy as np
import pytest

from cism_amd import _lib, pcg
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC, Grid
from oracle import pyoracle as po
from systems import random_spd_system

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = [(2, BC_NO_ICE, 9, 14), (20, BC_PERIODIC, 9, 7), (40, BC_OUTFLOW, 11, 6), (64, BC_NO_ICE, 7, 9),
         (33, BC_PERIODIC, 6, 6)]


@pytest.mark.parametrize("nz,bc,nx_own,ny_own", CASES)
@pytest.mark.parametrize("algo,precond", [("chrongear", 2), ("standard", 1), ("chrongear", 3)])
def test_fp64_solves_deep_and_ragged(require_gpu, nz, bc, nx_own, ny_own, algo, precond):
    g, act, A, bu, bv = random_spd_system(n=nx_own, ny=ny_own, nz=nz, bc=bc, seed=nz + bc)
    ou, ov = g.zeros_vec(), g.zeros_vec()
    st, oerr, onit = po.pcg_solve(g, act, *A, bu, bv, ou, ov, algo=algo, precond=precond, tolerance=1e-11,
                                  maxiters=3000)
    assert st == 0
    xu, xv = g.zeros_vec(), g.zeros_vec()
    fn = pcg.pcg_solver_chrongear_3d if algo == "chrongear" else pcg.pcg_solver_standard_3d
    err, nit = fn(g, act, *A, bu, bv, xu, xv, precond, 5, 1e-11, 3000)
    assert nit == onit, (nit, onit)
    assert _rel(np.stack([xu, xv]), np.stack([ou, ov])) < 1e-10


@pytest.mark.parametrize("nz,bc,nx_own,ny_own", CASES[1:4])
def test_mixed_solve_deep_columns(require_gpu, nz, bc, nx_own, ny_own):
    g, act, A, bu, bv = random_spd_system(n=nx_own, ny=ny_own, nz=nz, bc=bc, seed=7 * nz)
    ou, ov = g.zeros_vec(), g.zeros_vec()
    st, oerr, _ = po.pcg_solve(g, act, *A, bu, bv, ou, ov, precond=2, tolerance=1e-12, maxiters=20000)
    assert st == 0 and oerr < 1e-12
    with pcg.System(g) as s:
        s.set_matrix(act, *A, bu, bv)
        s.set_x(g.zeros_vec(), g.zeros_vec())
        r = s.solve(precond=2, tolerance=1e-10, maxiters=5000, precision=pcg.MIXED_FP32, inner_tol=1e-4,
                    max_outer=30)
        assert r["err"] < 1e-10
        assert _rel(np.stack(s.get_x()), np.stack([ou, ov])) < 1e-6


@pytest.mark.parametrize("algo", ["chrongear", "standard"])
def test_zero_rhs_is_the_reference_nan_abort(require_gpu, algo):
    """b = 0 makes the reference's first alpha 0/0; it aborts (GM_FATAL),
    here status 1, as the restatement reports."""
    g, act, A, bu, bv = random_spd_system(n=8, nz=4, bc=BC_NO_ICE, seed=1)
    z = g.zeros_vec()
    st, _, _ = po.pcg_solve(g, act, *A, z, z, g.zeros_vec(), g.zeros_vec(), algo=algo, precond=2,
                            tolerance=1e-10, maxiters=50)
    assert st == 1
    fn = pcg.pcg_solver_chrongear_3d if algo == "chrongear" else pcg.pcg_solver_standard_3d
    with pytest.raises(_lib.CismGpuError) as e:
        fn(g, act, *A, z, z.copy(order="F"), g.zeros_vec(), g.zeros_vec(), 2, 5, 1e-10, 50)
    assert e.value.status == 1


def test_nz_limit(require_gpu):
    g = Grid.from_owned(8, 8, 65)
    with pytest.raises(_lib.CismGpuError) as e:
        pcg.System(g)
    assert e.value.status == 3 and "2..64" in str(e.value)


@pytest.mark.parametrize("nz", [2, 33, 64])
def test_assembly_deep_and_shallow_columns(require_gpu, nz):
    """The BP assembly (stiffness, basal term, load, Dirichlet rows) at two
    levels and at 33 and 64 levels against the restatement (rows 1e-12)."""
    from cism_amd import synthetic
    from cism_amd.velo import BPSolver
    from test_gpu_assembly import _rowwise_close
    p = synthetic.ice_stream(20, nz=nz, seed=nz)
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


def test_dropin_device_system_is_reused_without_carrying_state(require_gpu):
    """The host-array drop-in keeps its device system for the last grid: a
    solve of another matrix on the same grid (other active mask, other rhs),
    a solve on another grid and a return to the first give bit for bit the
    results of fresh systems (cismgpu_dropin_release before each)."""
    L = _lib.load()
    L.cismgpu_dropin_release.restype = ctypes.c_int32

    def solve(sysdef, fresh):
        g, act, A, bu, bv = sysdef
        if fresh:
            assert L.cismgpu_dropin_release() == 0
        xu, xv = g.zeros_vec(), g.zeros_vec()
        err, nit = pcg.pcg_solver_chrongear_3d(g, act, *A, bu, bv, xu, xv, 2, 5, 1e-10, 500)
        return np.stack([xu, xv]), err, nit

    s1 = random_spd_system(n=12, nz=6, bc=BC_NO_ICE, seed=21)
    s2 = random_spd_system(n=12, nz=6, bc=BC_NO_ICE, seed=22)
    s3 = random_spd_system(n=10, nz=5, bc=BC_PERIODIC, seed=23)
    ref = [solve(s, True) for s in (s1, s2, s3, s1)]
    got = [solve(s1, True)] + [solve(s, False) for s in (s2, s3, s1)]
    for (x, e, n), (xr, er, nr) in zip(got, ref):
        assert np.array_equal(x, xr) and e == er and n == nr
    assert not np.array_equal(got[0][0], got[1][0])
    assert L.cismgpu_dropin_release() == 0
