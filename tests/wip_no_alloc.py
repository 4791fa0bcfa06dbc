"""cismgpu_residual_vector_3d_noalloc: the residual of compute_residual_vector_3d
on device arrays in the reference layout, with no device allocation (its
partial sums live in a library-held device array).  Two calls in a row, the
second on a smaller grid, each against the system entry point
cismgpu_system_residual on the same matrix, rhs and velocity."""
import ctypes

import numpy as np
import pytest

from cism_amd import _lib, synthetic
from cism_amd.velo import BPSolver

pytestmark = pytest.mark.gpu


def _dev(torch, a, dtype=None):
    """Device copy of a Fortran-ordered array, same memory order."""
    flat = np.ascontiguousarray(np.asfortranarray(a).ravel(order="F"))
    return torch.from_numpy(flat if dtype is None else flat.astype(dtype)).cuda()


def _case(torch, p, seed):
    g = p.grid
    bp = BPSolver(p)
    bp.assemble(p.uvel, p.vvel)
    (Auu, Auv, Avu, Avv), (bu, bv) = bp.sys.get_matrix()
    rng = np.random.default_rng(seed)
    u = np.asfortranarray(p.uvel * (1.0 + 0.1 * rng.standard_normal(g.vec_shape)))
    v = np.asfortranarray(p.vvel * (1.0 + 0.1 * rng.standard_normal(g.vec_shape)))
    bp.sys.set_x(u, v)
    L2, L2r, ru, rv = bp.residual(want_vectors=True)          # the system entry point
    bp.sys.close()
    d = {k: _dev(torch, a) for k, a in (("Auu", Auu), ("Auv", Auv), ("Avu", Avu), ("Avv", Avv), ("bu", bu),
                                          ("bv", bv), ("u", u), ("v", v))}
    d["act"] = _dev(torch, bp.active_vertex, np.int32)
    out_u = torch.full((g.nz * g.nvx * g.nvy,), float("nan"), dtype=torch.float64, device="cuda")
    out_v = torch.full_like(out_u, float("nan"))
    norms = torch.zeros(2, dtype=torch.float64, device="cuda")
    L = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    st = L.cismgpu_residual_vector_3d_noalloc(
        ctypes.byref(g.cstruct()), d["act"].data_ptr(), d["Auu"].data_ptr(), d["Auv"].data_ptr(),
        d["Avu"].data_ptr(), d["Avv"].data_ptr(), d["bu"].data_ptr(), d["bv"].data_ptr(), d["u"].data_ptr(),
        d["v"].data_ptr(), out_u.data_ptr(), out_v.data_ptr(), norms.data_ptr(), stream)
    assert st == 0, _lib.last_error()
    torch.cuda.synchronize()
    nl = norms.cpu().numpy()
    gu = out_u.cpu().numpy().reshape(g.vec_shape, order="F")
    gv = out_v.cpu().numpy().reshape(g.vec_shape, order="F")
    assert nl[0] == pytest.approx(L2, rel=1e-12) and nl[1] == pytest.approx(L2r, rel=1e-12)
    scale = max(np.abs(ru).max(), np.abs(rv).max())
    assert np.abs(gu - ru).max() <= 1e-12 * scale and np.abs(gv - rv).max() <= 1e-12 * scale
    return nl


def test_noalloc_residual_matches_system_residual_twice(require_gpu):
    torch = pytest.importorskip("torch")
    big = _case(torch, synthetic.ice_stream(48, nz=11, seed=3), 1)
    small = _case(torch, synthetic.dome(), 2)
    assert big[0] > 0 and small[0] > 0


def test_noalloc_rejects_grids_beyond_the_scratch(require_gpu):
    """A grid with more than 16,777,216 nodes is refused before any device
    work (the pointers are never dereferenced)."""
    from cism_amd.grid import Grid
    L = _lib.load()
    g = Grid.from_owned(1400, 1400, 9).cstruct()
    fake = ctypes.c_void_p(16)
    st = L.cismgpu_residual_vector_3d_noalloc(ctypes.byref(g), *([fake] * 12), None)
    assert st == 3 and "scratch" in _lib.last_error()
