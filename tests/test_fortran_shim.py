"""The Fortran drop-in (cism_amd/fortran/glissade_velo_higher_pcg.F90)
keeps the reference subroutine signatures and reaches the HIP library
through iso_c_binding.  CPU: it compiles and links with amdflang against
test doubles of the CISM modules it uses.  GPU: a Fortran driver calls
pcg_solver_{standard,chrongear}_3d as glissade_velo_higher.F90 does and the
solution matches the restated reference to 1e-10."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from systems import collapse_to_2d, random_spd_system, random_spd_system_2d, stream_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "tests", "fortran")
DRIVER = os.path.join(FDIR, "_build", "shim_driver")


def build_driver():
    subprocess.run(["make", "-s", "-C", FDIR], check=True, capture_output=True)
    return DRIVER


def test_shim_compiles_and_links():
    build_driver()
    assert os.path.exists(DRIVER)
    syms = subprocess.run(["nm", os.path.join(FDIR, "_build", "glissade_velo_higher_pcg.o")],
                          capture_output=True, text=True, check=True).stdout
    for name in ("pcg_solver_standard_3d", "pcg_solver_chrongear_3d", "matvec_multiply_structured_3d",
                 "pcg_solver_standard_2d", "pcg_solver_chrongear_2d", "matvec_multiply_structured_2d"):
        assert name in syms
    for cname in ("cismgpu_pcg_solver_standard_3d", "cismgpu_pcg_solver_chrongear_3d",
                  "cismgpu_matvec_multiply_structured_3d", "cismgpu_pcg_solver_standard_2d",
                  "cismgpu_pcg_solver_chrongear_2d", "cismgpu_matvec_multiply_structured_2d"):
        assert cname in syms


def run_driver(g, act, A, bu, bv, algo, precond, tol, maxiters, two_d=False):
    vshape = (g.nvx, g.nvy) if two_d else g.vec_shape
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        hdr = np.array([g.nx, g.ny, 0 if two_d else g.nz, g.ilo, g.ihi, g.jlo, g.jhi, g.global_bc, algo, precond,
                        5, maxiters], dtype=np.int32)
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(np.float64(tol).tobytes())
            f.write(np.asfortranarray(act, dtype=np.int32).tobytes(order="F"))
            for a in (*A, bu, bv, np.zeros(vshape), np.zeros(vshape)):
                f.write(np.asfortranarray(a).tobytes(order="F"))
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
        niters = int(np.frombuffer(raw[:4], np.int32)[0])
        err = float(np.frombuffer(raw[4:12], np.float64)[0])
        n = int(np.prod(vshape))
        x = np.frombuffer(raw[12:], np.float64)
        xu = x[:n].reshape(vshape, order="F")
        xv = x[n:].reshape(vshape, order="F")
        return xu, xv, err, niters


@pytest.mark.gpu
@pytest.mark.parametrize("algo", [2, 3])
def test_fortran_dropin_matches_oracle(require_gpu, algo):
    if not os.path.exists(DRIVER):
        build_driver()
    p, s = stream_system()
    g = p.grid
    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
    xu, xv, err, nit = run_driver(g, s["active_vertex"], A, s["bu"], s["bv"], algo, 2, 1e-11, 3000)
    ou, ov = g.zeros_vec(), g.zeros_vec()
    st, oerr, onit = po.pcg_solve(g, s["active_vertex"], *A, s["bu"], s["bv"], ou, ov,
                                  algo="chrongear" if algo == 3 else "standard", precond=2, tolerance=1e-11,
                                  maxiters=3000)
    assert err < 1e-11 and abs(nit - onit) <= 5
    rel = np.linalg.norm(np.stack([xu - ou, xv - ov])) / np.linalg.norm(np.stack([ou, ov]))
    assert rel < 1e-10


@pytest.mark.gpu
def test_fortran_dropin_periodic(require_gpu):
    if not os.path.exists(DRIVER):
        build_driver()
    g, act, A, bu, bv = random_spd_system(bc=0)
    xu, xv, err, nit = run_driver(g, act, A, bu, bv, 3, 3, 1e-11, 500)
    ou, ov = g.zeros_vec(), g.zeros_vec()
    po.pcg_solve(g, act, *A, bu, bv, ou, ov, algo="chrongear", precond=3, tolerance=1e-11, maxiters=500)
    rel = np.linalg.norm(np.stack([xu - ou, xv - ov])) / np.linalg.norm(np.stack([ou, ov]))
    assert rel < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("algo,precond", [(2, 1), (3, 1), (3, 3), (2, 4)])
def test_fortran_dropin_2d_matches_oracle(require_gpu, algo, precond):
    """pcg_solver_{standard,chrongear}_2d through the Fortran module on the
    depth-integrated ice-stream system and a periodic random system."""
    if not os.path.exists(DRIVER):
        build_driver()
    p, s = stream_system()
    A2, bu, bv = collapse_to_2d(p.grid, s["active_vertex"], [s["Auu"], s["Auv"], s["Avu"], s["Avv"]], s["bu"],
                                s["bv"])
    cases = [(p.grid, s["active_vertex"], A2, bu, bv), random_spd_system_2d(bc=0)]
    for g, act, A, bu, bv in cases:
        xu, xv, err, nit = run_driver(g, act, A, bu, bv, algo, precond, 1e-11, 3000, two_d=True)
        ou, ov = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
        st, oerr, onit = po.pcg_solve_2d(g, act, *A, bu, bv, ou, ov, algo="chrongear" if algo == 3 else "standard",
                                         precond=precond, tolerance=1e-11, maxiters=3000)
        assert err < 1e-11 and abs(nit - onit) <= 5
        rel = np.linalg.norm(np.stack([xu - ou, xv - ov])) / np.linalg.norm(np.stack([ou, ov]))
        assert rel < 1e-10
