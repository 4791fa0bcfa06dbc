"""Shared test systems assembled by the CPU restatement (oracle)."""
from __future__ import annotations

import functools

import numpy as np

from cism_amd import synthetic
from cism_amd.grid import BC_OUTFLOW, BC_PERIODIC, Grid
from oracle import pyoracle as po


@functools.lru_cache(maxsize=None)
def dome_system():
    p = synthetic.dome()
    return p, po.assemble(p)


@functools.lru_cache(maxsize=None)
def stream_system(n=40, nz=11, seed=3):
    p = synthetic.ice_stream(n, nz=nz, seed=seed)
    return p, po.assemble(p)


@functools.lru_cache(maxsize=None)
def random_spd_system(n=12, nz=6, bc=BC_PERIODIC, seed=0, ny=None):
    """Random symmetric positive definite 27-point 2x2-block system.

    Built from random SPD 16x16 "element" matrices on 2x2x2 node cubes, as
    the BP assembly builds its matrix.  Periodic grids wrap the elements on
    the torus of owned vertices, make every vertex active and fill the halo
    rows with the periodic halo rule (as glissade_velo_higher.F90:2893-2898
    does); other grids keep every element two vertices inside the owned
    block, so that the last owned row (which the no_ice halo rule zeroes)
    stays inactive.  ny (default n) makes the owned block n x ny."""
    from cism_amd.grid import indxA_3d
    g = Grid.from_owned(n, n if ny is None else ny, nz, global_bc=bc)
    rng = np.random.default_rng(seed)
    A = [g.zeros_mat() for _ in range(4)]
    periodic = bc in (0, 2)
    nox, noy = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1
    wrapi = lambda v: g.ilo + (v - g.ilo) % nox
    wrapj = lambda v: g.jlo + (v - g.jlo) % noy
    act = np.zeros((g.nvx, g.nvy), np.int32, order="F")
    if periodic:
        act[:, :] = 1
        icells, jcells = range(g.ilo, g.ihi + 1), range(g.jlo, g.jhi + 1)
    else:
        icells, jcells = range(g.ilo, g.ihi - 1), range(g.jlo, g.jhi - 1)
        act[g.ilo - 1:g.ihi - 1, g.jlo - 1:g.jhi - 1] = 1
    offs = [(di, dj, dk) for dk in (0, 1) for dj in (0, 1) for di in (0, 1)]
    for j in jcells:
        for i in icells:
            for k in range(1, nz):
                B = rng.standard_normal((16, 16)) * 0.3
                K = B.T @ B
                nodes = [(k + dk, i + di, j + dj) for (di, dj, dk) in offs]
                for r, (kr, ir, jr) in enumerate(nodes):
                    irw, jrw = (wrapi(ir), wrapj(jr)) if periodic else (ir, jr)
                    for c, (kc, ic, jc) in enumerate(nodes):
                        m = indxA_3d(ic - ir, jc - jr, kc - kr)
                        A[0][m - 1, kr - 1, irw - 1, jrw - 1] += K[r, c]
                        A[1][m - 1, kr - 1, irw - 1, jrw - 1] += K[r, 8 + c]
                        A[2][m - 1, kr - 1, irw - 1, jrw - 1] += K[8 + r, c]
                        A[3][m - 1, kr - 1, irw - 1, jrw - 1] += K[8 + r, 8 + c]
    d = indxA_3d(0, 0, 0) - 1
    own = np.zeros_like(act)
    own[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] = 1
    A[0][d] += 4.0 * (act * own)[None]
    A[3][d] += 4.0 * (act * own)[None]
    for a in A:
        po.halo(g, a.reshape((27 * nz, g.nvx, g.nvy), order="F"))
    bu = np.asfortranarray(rng.standard_normal(g.vec_shape) * (act * own)[None])
    bv = np.asfortranarray(rng.standard_normal(g.vec_shape) * (act * own)[None])
    return g, act, A, bu, bv


def random_spd_system_2d(n=14, bc=BC_PERIODIC, seed=0):
    """Random symmetric positive definite 9-point 2x2-block system in the
    layout of the 2-D solvers: A(nx-1, ny-1, 9), m = indxA_2d(iA, jA) last;
    vectors (nx-1, ny-1).  Built from random SPD 8x8 element matrices on
    2x2 vertex squares, with the same activity and halo rules as
    random_spd_system."""
    from cism_amd.grid import indxA_2d
    g = Grid.from_owned(n, n, 2, global_bc=bc)
    rng = np.random.default_rng(seed)
    A = [np.zeros((9, g.nvx, g.nvy), order="F") for _ in range(4)]      # m first while assembling
    periodic = bc in (0, 2)
    no = g.ihi - g.ilo + 1
    wrap = lambda v, lo: lo + (v - lo) % no
    act = np.zeros((g.nvx, g.nvy), np.int32, order="F")
    if periodic:
        act[:, :] = 1
        cells = range(g.ilo, g.ihi + 1)
    else:
        cells = range(g.ilo, g.ihi - 1)
        act[g.ilo - 1:g.ihi - 1, g.jlo - 1:g.jhi - 1] = 1
    offs = [(di, dj) for dj in (0, 1) for di in (0, 1)]
    for j in cells:
        for i in cells:
            B = rng.standard_normal((8, 8)) * 0.3
            K = B.T @ B
            nodes = [(i + di, j + dj) for (di, dj) in offs]
            for r, (ir, jr) in enumerate(nodes):
                irw, jrw = (wrap(ir, g.ilo), wrap(jr, g.jlo)) if periodic else (ir, jr)
                for c, (ic, jc) in enumerate(nodes):
                    m = indxA_2d(ic - ir, jc - jr) - 1
                    A[0][m, irw - 1, jrw - 1] += K[r, c]
                    A[1][m, irw - 1, jrw - 1] += K[r, 4 + c]
                    A[2][m, irw - 1, jrw - 1] += K[4 + r, c]
                    A[3][m, irw - 1, jrw - 1] += K[4 + r, 4 + c]
    own = np.zeros_like(act)
    own[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] = 1
    d = indxA_2d(0, 0) - 1
    A[0][d] += 2.0 * (act * own)
    A[3][d] += 2.0 * (act * own)
    for a in A:
        po.halo(g, a)
    A2 = [np.asfortranarray(np.moveaxis(a, 0, 2)) for a in A]
    bu = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * (act * own))
    bv = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * (act * own))
    return g, act, A2, bu, bv



def collapse_to_2d(g, act, A, bu, bv):
    """Depth-integrated 2-D form P^T A P of an assembled 3-D system (see
    cism_amd.synthetic.depth_integrated_system)."""
    from cism_amd.synthetic import depth_integrated_system
    return depth_integrated_system(*A, bu, bv)
