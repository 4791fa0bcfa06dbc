"""Host-side 2-D setup of the velocity solve (vectorised numpy).

These are the O(nx*ny) operators that ``glissade_velo_higher_solve`` runs on
the host before assembling the 3-D system (glissade_velo_higher.F90:1247-1847);
the 3-D work (assembly, boundary conditions, solve) runs on the GPU.  Arrays
are in the reference layouts: cell fields (nx, ny), vertex fields
(nx-1, ny-1), flow factors (nz-1, nx, ny), Fortran order.
"""
from __future__ import annotations

import numpy as np

from .grid import BC_OUTFLOW, BC_NO_ICE, Grid

RHOI = 910.0      # glimmer_physcon.F90:67
RHOO = 1028.0     # :68
GRAV = 9.81       # :69


def calclsrf(thck, topg, eus):
    """glide_calclsrf (glide_thck.F90:1053-1076)."""
    f = (-RHOI / RHOO) * thck
    return np.asfortranarray(np.where(topg - eus < f, f, topg))


def get_masks(thck, topg, eus, thklim):
    """ice/floating/ocean/land masks of glissade_get_masks (glissade_masks.F90:160-190)."""
    ice = (thck > thklim).astype(np.int32)
    ocean = ((topg < eus) & (ice == 0)).astype(np.int32)
    floating = ((topg - eus < (-RHOI / RHOO) * thck) & (ice == 1)).astype(np.int32)
    land = (topg >= eus).astype(np.int32)
    return (np.asfortranarray(ice), np.asfortranarray(floating), np.asfortranarray(ocean), np.asfortranarray(land))


def stagger(var):
    """glissade_stagger, stagger_margin = 0 (glissade_grid_operators.F90:125-133)."""
    return np.asfortranarray((var[:-1, 1:] + var[1:, 1:] + var[:-1, :-1] + var[1:, :-1]) / 4.0)


def surface_elevation_gradient(dx, dy, ice_mask, land_mask, usrf, thck, thklim, thck_gradient_ramp,
                               margin, max_slope):
    """glissade_surface_elevation_gradient with ho_gradient = centered
    (glissade_grid_operators.F90:834-1260)."""
    nx, ny = usrf.shape
    ex = np.zeros((nx, ny))
    ey = np.zeros((nx, ny))
    if margin == 0:
        ex[:-1, :] = (usrf[1:, :] - usrf[:-1, :]) / dx
        ey[:, :-1] = (usrf[:, 1:] - usrf[:, :-1]) / dy
    elif margin == 2:
        both = (ice_mask[:-1, :] == 1) & (ice_mask[1:, :] == 1)
        ex[:-1, :] = np.where(both, (usrf[1:, :] - usrf[:-1, :]) / dx, 0.0)
        both = (ice_mask[:, :-1] == 1) & (ice_mask[:, 1:] == 1)
        ey[:, :-1] = np.where(both, (usrf[:, 1:] - usrf[:, :-1]) / dy, 0.0)
    else:
        def edges(s0, s1, t0, t1, i0, i1, l0, l1, h):
            up_is0 = s0 > s1
            su = np.where(up_is0, s0, s1)
            sl = np.where(up_is0, s1, s0)
            sign = np.where(up_is0, -1.0, 1.0)
            tu = np.where(up_is0, t0, t1)
            iu = np.where(up_is0, i0, i1)
            il = np.where(up_is0, i1, i0)
            lu = np.where(up_is0, l0, l1)
            ll = np.where(up_is0, l1, l0)
            ef = np.ones_like(su)
            if thck_gradient_ramp > 0.0:
                ramp = np.maximum(np.minimum(1.0, (tu - thklim) / thck_gradient_ramp), 0.0)
                ef = np.where(lu == 1, ramp, 1.0)
            ok = ((iu == 1) & (il == 1)) | ((iu == 1) & (ll == 1))
            return np.where(ok, ef * sign * (su - sl) / h, 0.0)
        ex[:-1, :] = edges(usrf[:-1, :], usrf[1:, :], thck[:-1, :], thck[1:, :], ice_mask[:-1, :], ice_mask[1:, :],
                           land_mask[:-1, :], land_mask[1:, :], dx)
        ey[:, :-1] = edges(usrf[:, :-1], usrf[:, 1:], thck[:, :-1], thck[:, 1:], ice_mask[:, :-1], ice_mask[:, 1:],
                           land_mask[:, :-1], land_mask[:, 1:], dy)
    dsdx = 0.5 * (ex[:-1, :-1] + ex[:-1, 1:])
    dsdy = 0.5 * (ey[:-1, :-1] + ey[1:, :-1])
    if max_slope is not None and max_slope > 0.0:
        dsdx = np.where(dsdx > 0.0, np.minimum(dsdx, max_slope), np.maximum(dsdx, -max_slope))
        dsdy = np.where(dsdy > 0.0, np.minimum(dsdy, max_slope), np.maximum(dsdy, -max_slope))
    return np.asfortranarray(dsdx), np.asfortranarray(dsdy)


def slope_angle(dx, dy, zsfc):
    """glissade_slope_angle with a full slope mask (glissade_grid_operators.F90:
    1264-1395): vertex gradient (glissade_gradient, :309-468), unstagger
    (:226-305), theta = atan(|grad|); the outer ring is zero."""
    nx, ny = zsfc.shape
    north = (zsfc[1:, 1:] - zsfc[:-1, 1:]) / dx
    south = (zsfc[1:, :-1] - zsfc[:-1, :-1]) / dx
    gx = (north + south) / 2.0
    east = (zsfc[1:, 1:] - zsfc[1:, :-1]) / dy
    west = (zsfc[:-1, 1:] - zsfc[:-1, :-1]) / dy
    gy = (east + west) / 2.0
    theta = np.zeros((nx, ny))
    ux = (gx[1:, 1:] + gx[:-1, 1:] + gx[1:, :-1] + gx[:-1, :-1]) / 4.0
    uy = (gy[1:, 1:] + gy[:-1, 1:] + gy[1:, :-1] + gy[:-1, :-1]) / 4.0
    theta[1:-1, 1:-1] = np.arctan(np.sqrt(ux * ux + uy * uy))
    return np.asfortranarray(theta)


def _vertex_gradient(dx, dy, z):
    """glissade_gradient with a full mask (glissade_grid_operators.F90:309-468):
    the centred vertex gradient of a cell field."""
    gx = ((z[1:, 1:] - z[:-1, 1:]) / dx + (z[1:, :-1] - z[:-1, :-1]) / dx) / 2.0
    gy = ((z[1:, 1:] - z[1:, :-1]) / dy + (z[:-1, 1:] - z[:-1, :-1]) / dy) / 2.0
    return gx, gy


def slope_angle_xy(dx, dy, zsfc):
    """glissade_slope_angle with its optional x/y outputs (glissade_grid_operators.F90:
    1264-1395): theta = atan|grad|, theta_x = atan|dz/dx|, theta_y = atan|dz/dy|
    of the vertex gradient unstaggered to cell centres; the outer ring is zero."""
    z = np.asarray(zsfc, dtype=np.float64)
    nx, ny = z.shape
    gx, gy = _vertex_gradient(dx, dy, z)
    ux = (gx[1:, 1:] + gx[:-1, 1:] + gx[1:, :-1] + gx[:-1, :-1]) / 4.0
    uy = (gy[1:, 1:] + gy[:-1, 1:] + gy[1:, :-1] + gy[:-1, :-1]) / 4.0
    out = [np.zeros((nx, ny), order="F") for _ in range(3)]
    out[0][1:-1, 1:-1] = np.arctan(np.sqrt(ux * ux + uy * uy))
    out[1][1:-1, 1:-1] = np.arctan(np.abs(ux))
    out[2][1:-1, 1:-1] = np.arctan(np.abs(uy))
    return tuple(out)


def slope_angle_staggered(dx, dy, zsfc):
    """glissade_slope_angle_staggered (glissade_grid_operators.F90:1414-1552):
    the angles of the vertex gradient of a cell field, (nx-1, ny-1)."""
    gx, gy = _vertex_gradient(dx, dy, np.asarray(zsfc, dtype=np.float64))
    return (np.asfortranarray(np.arctan(np.sqrt(gx * gx + gy * gy))), np.asfortranarray(np.arctan(np.abs(gx))),
            np.asfortranarray(np.arctan(np.abs(gy))))


def cell_halo_2d(grid: Grid, a: np.ndarray) -> None:
    """parallel_halo for a cell field on one task, 2 halo cells (parallel_mpi.F90:
    7720-7912): periodic wrap (east/west over the owned rows, then north/south
    over every column); outflow zeroes the halo; no_ice zeroes 3 rows on each
    side."""
    nx, ny = a.shape
    if grid.global_bc in (0, 2):
        r = slice(2, ny - 2)
        w, e = a[nx - 4:nx - 2, r].copy(), a[2:4, r].copy()
        a[0:2, r] = w
        a[nx - 2:nx, r] = e
        s, n = a[:, ny - 4:ny - 2].copy(), a[:, 2:4].copy()
        a[:, 0:2] = s
        a[:, ny - 2:ny] = n
    else:
        w = 2 if grid.global_bc == BC_OUTFLOW else 3
        a[:w, :] = 0.0
        a[nx - w:, :] = 0.0
        a[:, :w] = 0.0
        a[:, ny - w:] = 0.0


def diva_geometry(grid: Grid, dx, dy, usrf, lsrf, ice_mask, land_mask):
    """DIVA inputs of glissade_velo_higher_solve with diva_slope_correction
    (glissade_velo_higher.F90:2078-2154): slope angles of z_mean = (usrf +
    lsrf)/2 at cells (with their halo) and vertices, the slope factors
    cos^2 + 4 sin^2, the basal slope angles (of the staggered lower surface,
    :8591-8605) and the ice-or-land cell mask (:1510-1515)."""
    zmean = 0.5 * (np.asarray(usrf) + np.asarray(lsrf))
    _, thx, thy = slope_angle_xy(dx, dy, zmean)
    cell_halo_2d(grid, thx)
    cell_halo_2d(grid, thy)
    _, sthx, sthy = slope_angle_staggered(dx, dy, zmean)
    _, tbx, tby = slope_angle_xy(dx, dy, stagger(lsrf))
    sf = lambda t: np.asfortranarray(np.cos(t) ** 2 + 4.0 * np.sin(t) ** 2)
    ipl = np.asfortranarray(((ice_mask == 1) | (land_mask == 1)).astype(np.int32))
    return dict(slope_factor_x=sf(thx), slope_factor_y=sf(thy), stag_slope_factor_x=sf(sthx),
                stag_slope_factor_y=sf(sthy), stag_theta_slope_x=sthx, stag_theta_slope_y=sthy,
                theta_basal_x=tbx, theta_basal_y=tby, ice_plus_land_mask=ipl)


def vertex_masks(ice_mask, nhalo=2):
    """active_cell / active_vertex of get_vertex_geometry (glissade_velo_higher.F90:4164-4183)."""
    nx, ny = ice_mask.shape
    acell = np.zeros((nx, ny), np.int32)
    acell[nhalo:nx - nhalo + 1, nhalo:ny - nhalo + 1] = ice_mask[nhalo:nx - nhalo + 1, nhalo:ny - nhalo + 1] == 1
    avert = np.zeros((nx - 1, ny - 1), np.int32)
    a = acell[nhalo:nx - nhalo + 1, nhalo:ny - nhalo + 1].astype(bool)
    # 1-based cell (i, j) activates vertices i-1..i, j-1..j
    for di in (0, 1):
        for dj in (0, 1):
            sl = avert[nhalo - 1 + di:nx - nhalo + di, nhalo - 1 + dj:ny - nhalo + dj]
            sl |= a
    return np.asfortranarray(acell), np.asfortranarray(avert)


def flwafact(flwa, n_glen, nhalo=2):
    """0.5 * A^(-1/n) over cells nhalo+1..n-nhalo+1 (glissade_velo_higher.F90:1838-1847)."""
    nzm1, nx, ny = flwa.shape
    out = np.zeros_like(flwa)
    sl = (slice(None), slice(nhalo, nx - nhalo + 1), slice(nhalo, ny - nhalo + 1))
    f = flwa[sl]
    with np.errstate(divide="ignore", invalid="ignore"):
        out[sl] = np.where(f > 0.0, 0.5 * f ** (-1.0 / n_glen), 0.0)
    return np.asfortranarray(out)


def staggered_halo_2d(grid: Grid, a: np.ndarray) -> None:
    """staggered_parallel_halo for a vertex field on one task (parallel_mpi.F90:10609-10769)."""
    nv, nw = a.shape
    if grid.global_bc in (0, 2):
        j0, j1 = grid.jlo - 1, grid.jhi
        w = a[2, j0:j1].copy()
        e = a[nv - 3:nv - 1, j0:j1].copy()
        a[nv - 1, j0:j1] = w
        a[0:2, j0:j1] = e
        so = a[:, 2].copy()
        no = a[:, nw - 3:nw - 1].copy()
        a[:, nw - 1] = so
        a[:, 0:2] = no
    elif grid.global_bc == BC_OUTFLOW:
        a[0, :] = 0.0
        a[nv - 1, :] = 0.0
        a[:, 0] = 0.0
        a[:, nw - 1] = 0.0
    elif grid.global_bc == BC_NO_ICE:
        a[0:2, :] = 0.0
        a[nv - 2:, :] = 0.0
        a[:, 0:2] = 0.0
        a[:, nw - 2:] = 0.0


def grounded_fraction(ice_mask, floating_mask, land_mask):
    """f_ground at vertices for the default options (no grounding-line
    parameterisation, grounded-neighbour rule; glissade_grounding_line.F90:59-674,
    glide_types.F90:1090, 1099): 1 where a neighbouring cell is land or grounded
    ice, 0 elsewhere.  The vertex mask tests the land mask of cell (i+1, j+1)
    twice and never that of (i+1, j), as the reference does."""
    ice = ice_mask == 1
    land = land_mask == 1
    cg = (ice & (floating_mask == 0)) | land
    vmask = (ice[:-1, 1:] | ice[1:, 1:] | ice[:-1, :-1] | ice[1:, :-1] |
             land[:-1, 1:] | land[1:, 1:] | land[:-1, :-1])
    anyg = cg[:-1, 1:] | cg[1:, 1:] | cg[:-1, :-1] | cg[1:, :-1]
    return np.asfortranarray(np.where(vmask & anyg, 1.0, 0.0))


def stagger_masked(var, mask):
    """glissade_stagger with stagger_margin = 1: mean over the cells whose mask
    is 1, zero where none is (glissade_grid_operators.F90:135-146)."""
    m = mask.astype(np.int64)
    sumvar = (m[:-1, 1:] * var[:-1, 1:] + m[1:, 1:] * var[1:, 1:]) + m[:-1, :-1] * var[:-1, :-1] + m[1:, :-1] * var[1:, :-1]
    summask = (m[:-1, 1:] + m[1:, 1:] + m[:-1, :-1] + m[1:, :-1]).astype(np.float64)
    out = np.zeros_like(sumvar, dtype=np.float64)
    np.divide(sumvar, summask, out=out, where=summask > 0)
    return np.asfortranarray(out)


def beta_internal(problem, ice_mask, floating_mask, land_mask):
    """glissade_calcbeta for constant / large / external beta (glissade_basal_traction.F90:
    202, 304, 399) times f_ground (:650-672, beta_grounded_min = 0), then its
    staggered halo."""
    g = problem.grid
    if problem.whichbabc == 0:
        b = np.full((g.nvx, g.nvy), problem.beta_const)
    elif problem.whichbabc == 4:
        b = np.full((g.nvx, g.nvy), 1.0e10)
    else:
        b = np.array(problem.beta, dtype=np.float64)
    b = np.asfortranarray(b * grounded_fraction(ice_mask, floating_mask, land_mask))
    staggered_halo_2d(g, b)                    # (:697)
    return b
