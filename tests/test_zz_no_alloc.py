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


class _Hip:
    """Device buffers through the HIP runtime the library itself uses
    (libamdhip64, resolved by soname to the already loaded copy)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a, dtype=np.float64):
        flat = np.ascontiguousarray(np.asfortranarray(a).ravel(order="F"), dtype=dtype)
        p = self.alloc(flat.nbytes)
        assert self.rt.hipMemcpy(p, flat.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(flat.nbytes), 1) == 0
        return p

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 8))) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _case(p, seed):
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
    h = _Hip()
    try:
        args = [h.put(bp.active_vertex, np.int32)] + [h.put(a) for a in (Auu, Auv, Avu, Avv, bu, bv, u, v)]
        nvec = g.nz * g.nvx * g.nvy
        out_u, out_v, norms = h.alloc(8 * nvec), h.alloc(8 * nvec), h.alloc(16)
        L = _lib.load()
        st = L.cismgpu_residual_vector_3d_noalloc(ctypes.byref(g.cstruct()), *args, out_u, out_v, norms, None)
        assert st == 0, _lib.last_error()
        nl = h.get(norms, 2)
        gu = h.get(out_u, nvec).reshape(g.vec_shape, order="F")
        gv = h.get(out_v, nvec).reshape(g.vec_shape, order="F")
    finally:
        h.close()
    assert nl[0] == pytest.approx(L2, rel=1e-12) and nl[1] == pytest.approx(L2r, rel=1e-12)
    scale = max(np.abs(ru).max(), np.abs(rv).max())
    assert np.abs(gu - ru).max() <= 1e-12 * scale and np.abs(gv - rv).max() <= 1e-12 * scale
    return nl


def test_noalloc_residual_matches_system_residual_twice(require_gpu):
    big = _case(synthetic.ice_stream(48, nz=11, seed=3), 1)
    small = _case(synthetic.dome(), 2)
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
