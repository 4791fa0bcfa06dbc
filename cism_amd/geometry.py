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


RHOO = 1028.0   # ocean density (glimmer_physcon.F90:68)


def flotation_function(thck, topg, eus, ice_mask, land_mask, which=2, global_bc=3, topg_raised=None):
    """Flotation function at the cells (glissade_grounding_line.F90:237-392):
    0 Pattyn (-rho_o (b - eus) / (rho_i H) - 1), 1 inverse Pattyn
    (1 - rho_i H / (-rho_o (b - eus)), at least -10), both -10 on land and 0
    in ice-free ocean, which then takes the smallest value of its ice or land
    edge neighbours (else corner neighbours) and the cell halo; 2 linear
    (-(b - eus) - rho_i/rho_o H, at least 1e-4 in magnitude; on land
    -max(b - eus, 50)); 3 the same with the raised bed topg_raised in place
    of b (:299-326; the reference stops without it).  Positive where the ice
    floats."""
    if which not in (0, 1, 2, 3):
        raise ValueError(f"which_ho_flotation_function {which}: 0, 1, 2 or 3")
    if which == 3:
        if topg_raised is None:
            raise ValueError("which_ho_flotation_function 3 needs topg_raised")
        topg = topg_raised
    H = np.asarray(thck, dtype=np.float64)
    d = np.asarray(topg, dtype=np.float64) - eus
    ice, land = ice_mask == 1, land_mask == 1
    if which in (0, 1):
        with np.errstate(divide="ignore", invalid="ignore"):
            if which == 0:
                fi = (-RHOO * d) / (RHOI * H) - 1.0
            else:
                fi = np.maximum(1.0 - (RHOI * H) / (-RHOO * d), -10.0)
        f = np.where(land, -10.0, np.where(ice, fi, 0.0))
        # ice-free ocean cells off the outermost ring: the smallest value of
        # their ice or land edge neighbours, else of their corner neighbours
        has = ice | land
        fin = np.where(has, f, np.inf)
        inner = (slice(1, -1), slice(1, -1))
        sh = lambda a, b: fin[1 + a:fin.shape[0] - 1 + a, 1 + b:fin.shape[1] - 1 + b]
        emin = np.minimum.reduce([sh(0, -1), sh(-1, 0), sh(1, 0), sh(0, 1)])
        cmin = np.minimum.reduce([sh(-1, -1), sh(1, -1), sh(-1, 1), sh(1, 1)])
        fill = np.where(np.isfinite(emin), emin, cmin)
        ex = f.copy()
        ex[inner] = np.where(~has[inner] & np.isfinite(fill), fill, f[inner])
        _cell_halo(ex, global_bc)
        return np.asfortranarray(ex)
    fl = -d - (RHOI / RHOO) * H
    fl = np.where(np.abs(fl) < 1.0e-4, np.where(fl < 0.0, -1.0e-4, 1.0e-4), fl)
    return np.asfortranarray(np.where(land, -np.maximum(d, 50.0), fl))


def _cell_halo(a, global_bc):
    """parallel_halo of a cell field on one task with two halo cells:
    periodic and no-penetration wrap, outflow and no-ice zero."""
    nx, ny = a.shape
    if global_bc in (0, 2):
        a[0, 2:ny - 2] = a[nx - 4, 2:ny - 2]; a[1, 2:ny - 2] = a[nx - 3, 2:ny - 2]
        a[nx - 2, 2:ny - 2] = a[2, 2:ny - 2]; a[nx - 1, 2:ny - 2] = a[3, 2:ny - 2]
        a[:, 0] = a[:, ny - 4]; a[:, 1] = a[:, ny - 3]
        a[:, ny - 2] = a[:, 2]; a[:, ny - 1] = a[:, 3]
    else:
        w = 2 if global_bc == 1 else 3
        a[:w, :] = 0.0; a[nx - w:, :] = 0.0; a[:, :w] = 0.0; a[:, ny - w:] = 0.0


def grounded_area(f1, f2, f3, f4):
    """compute_grounded_fraction (glissade_grounding_line.F90:678-1080),
    vectorised: the grounded (f <= 0) area of the unit square under the
    bilinear interpolant of corner values f1 (SW), f2 (SE), f3 (NE), f4 (NW),
    by the reference's closed forms for one, two adjacent, two opposite and
    three floating corners.  NaN where the reference stops (an opposite-
    corner pattern with bc - ad < 0 or (b + d)(c + d) < 0)."""
    F = np.stack([np.asarray(x, dtype=np.float64) for x in (f1, f2, f3, f4)])
    cf = F > 0.0
    nf = cf.sum(axis=0)
    out = np.zeros(F.shape[1:])
    out[nf == 0] = 1.0
    zero = ((nf == 4) | (F.min(axis=0) == 0.0)) & (nf != 0)     # no floating corner: grounded first

    def roll(r):
        return F[r % 4], F[(r + 1) % 4], F[(r + 2) % 4], F[(r + 3) % 4]

    def pick(rot):
        g = [np.zeros_like(F[0]) for _ in range(4)]
        for r in range(4):
            m = rot == r
            for q, v in enumerate(roll(r)):
                g[q] = np.where(m, v, g[q])
        return g

    with np.errstate(divide="ignore", invalid="ignore"):
        # one or three floating corners: the lone corner first
        lone = np.where(nf == 1, cf, ~cf)
        rot1 = np.where(lone[0], 0, np.where(lone[1], 1, np.where(lone[2], 2, 3)))
        a, f2r, f3r, f4r = pick(rot1)
        b, c, d = f2r - a, f4r - a, a + f3r - f2r - f4r
        t = (a * d) / (b * c)
        corner = np.where(np.abs(t) > 1.0e-6, ((b * c - a * d) * np.log(np.abs(1.0 - t)) + a * d) / (d * d),
                          (a * a) / (2.0 * b * c))
        m13 = ((nf == 1) | (nf == 3)) & ~zero
        out = np.where(m13, np.where(nf == 1, 1.0 - corner, corner), out)
        # two floating corners
        adj = (cf[0] & cf[1]) | (cf[1] & cf[2]) | (cf[2] & cf[3]) | (cf[3] & cf[0])
        rot2 = np.where(cf[0] & cf[1], 0, np.where(cf[1] & cf[2], 1, np.where(cf[2] & cf[3], 2, 3)))
        var = F[1] * F[3] - F[0] * F[2]
        rotated = ~adj & (var < 0.0)
        rot2 = np.where(adj, rot2, np.where(rotated, 1, 0))
        a, f2r, f3r, f4r = pick(rot2)
        b, c, d = f2r - a, f4r - a, a + f3r - f2r - f4r
        trap = np.where(np.abs(d / c) > 1.0e-6, ((b * c - a * d) * np.log(1.0 + d / c) - b * d) / (d * d),
                        -(2.0 * a + b) / (2.0 * c))
        bcad = b * c - a * d
        usual = np.abs(bcad) > 1.0e-6
        c1 = np.where(usual, (bcad * np.log(1.0 - (a * d) / (b * c)) + a * d) / (d * d), a * a / (b * c))
        c2 = np.where(usual, (bcad * np.log(bcad / ((b + d) * (c + d))) + d * (a + b + c + d)) / (d * d),
                      (a + b) * (a + c) / (b * c))
        grounded_corners = np.where(F[0] > 0.0, rotated, ~rotated)
        opp = np.where(grounded_corners, c1 + c2, 1.0 - (c1 + c2))
        bad = ~adj & ((bcad < 0.0) | ((b + d) * (c + d) < 0.0))
        opp = np.where(bad, np.nan, opp)
        m2 = (nf == 2) & ~zero
        out = np.where(m2, np.where(adj, 1.0 - trap, opp), out)
    out[zero] = 0.0
    return out


def grounded_fraction_options(thck, topg, eus, ice_mask, floating_mask, land_mask, which_ground=0,
                              which_flotation=2, which_no_glp=0, global_bc=3, topg_raised=None):
    """f_ground at the vertices for which_ho_ground 0 (no GLP: the
    grounded-neighbour rule, or with which_ho_fground_no_glp = 1 the sign of
    the staggered flotation function), 1 (GLP for basal friction: the
    grounded area of the bilinear flotation function over the staggered
    cell) and 2 (the mean over the cell's four quadrants, each bilinear
    between a cell value, two edge midpoints and the vertex)
    (glissade_grounding_line.F90:404-618).  Returns (f_ground, f_flotation)."""
    if which_ground not in (0, 1, 2):
        raise ValueError("which_ho_ground 0, 1 or 2")
    ff = flotation_function(thck, topg, eus, ice_mask, land_mask, which_flotation, global_bc, topg_raised)
    ice = ice_mask == 1
    land = land_mask == 1
    vmask = (ice[:-1, 1:] | ice[1:, 1:] | ice[:-1, :-1] | ice[1:, :-1] |
             land[:-1, 1:] | land[1:, 1:] | land[:-1, :-1])
    if which_ground == 0 and which_no_glp == 0:
        return grounded_fraction(ice_mask, floating_mask, land_mask), ff
    if which_ground == 0:
        stag = (ff[:-1, 1:] + ff[1:, 1:] + ff[:-1, :-1] + ff[1:, :-1]) / 4.0
        fg = np.where(vmask & (stag <= 0.0), 1.0, 0.0)
    else:
        sw, se, ne, nw = ff[:-1, :-1], ff[1:, :-1], ff[1:, 1:], ff[:-1, 1:]
        if which_ground == 1:
            area = grounded_area(sw, se, ne, nw)
        else:
            # the four quadrants of the staggered cell, each bilinear between a
            # cell value, two edge midpoints and the vertex (:521-618)
            ctr = 0.25 * (nw + ne + sw + se)
            q1 = grounded_area(sw, 0.5 * (sw + se), ctr, 0.5 * (sw + nw))
            q2 = grounded_area(0.5 * (se + sw), se, 0.5 * (se + ne), ctr)
            q3 = grounded_area(ctr, 0.5 * (ne + se), ne, 0.5 * (ne + nw))
            q4 = grounded_area(0.5 * (nw + sw), ctr, 0.5 * (nw + ne), nw)
            area = 0.25 * (q1 + q2 + q3 + q4)
        fg = np.where(vmask, area, 0.0)
        if np.isnan(fg[vmask]).any():
            raise ValueError("grounded fraction: a flotation pattern the reference stops on")
    return np.asfortranarray(fg), ff


def problem_f_ground(problem, ice_mask, floating_mask, land_mask):
    """f_ground for a problem's which_ho_ground, which_ho_flotation_function
    and which_ho_fground_no_glp (defaults 0, 2, 0: the grounded-neighbour
    rule)."""
    p = problem
    fg, _ = grounded_fraction_options(p.thck, p.topg, p.eus, ice_mask, floating_mask, land_mask,
                                      getattr(p, "which_ho_ground", 0), getattr(p, "which_ho_flotation_function", 2),
                                      getattr(p, "which_ho_fground_no_glp", 0), p.grid.global_bc,
                                      getattr(p, "topg_raised", None))
    return fg


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
    202, 304, 399) times f_ground (:650-672, beta_grounded_min = 0; f_ground
    by the problem's grounding-line options), then its staggered halo."""
    g = problem.grid
    if problem.whichbabc == 0:
        b = np.full((g.nvx, g.nvy), problem.beta_const)
    elif problem.whichbabc == 4:
        b = np.full((g.nvx, g.nvy), 1.0e10)
    else:
        b = np.array(problem.beta, dtype=np.float64)
    b = np.asfortranarray(b * problem_f_ground(problem, ice_mask, floating_mask, land_mask))
    staggered_halo_2d(g, b)                    # (:697)
    return b
