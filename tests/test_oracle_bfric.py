"""compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042) and
the basal traction beta u_b (:3817-3818) in the oracle, pinned by closed
forms.  The standard option integrates beta |u|^2 with the 2x2 Gauss rule,
which is exact for the cubics that bilinear beta times squared bilinear
velocity reduce to along one axis: with beta = bc + b1 xi and
u = uc + u1 xi on the reference element, the cell mean is
bc uc^2 + bc u1^2 / 3 + 2 b1 uc u1 / 3, which changes if a cell reads the
wrong vertices.  The local option is the plain mean of the four vertex
values over active vertices.  CPU only."""
from types import SimpleNamespace

import numpy as np

from oracle import pyoracle as po

SCYR = 31536000.0


def _grid(nx=10, ny=9, bc=1):
    return SimpleNamespace(nx=nx, ny=ny, global_bc=bc)


def _all_active(g):
    return np.ones((g.nx, g.ny), np.int32, order="F"), np.ones((g.nx - 1, g.ny - 1), np.int32, order="F")


def _owned(g):
    """owned cells (1-based nhalo+1 .. n-nhalo), as 0-based slices"""
    return slice(2, g.nx - 2), slice(2, g.ny - 2)


def _vertex_x(g, dx):
    """x of vertex (i, j) (1-based) = i dx, as an (nx-1, ny-1) field"""
    return np.asfortranarray(np.broadcast_to((np.arange(g.nx - 1)[:, None] + 1.0) * dx, (g.nx - 1, g.ny - 1)))


def _cell_x(g, dx):
    """x of the centre of cell (i, j) (1-based) = (i - 1/2) dx"""
    return np.broadcast_to((np.arange(g.nx)[:, None] + 0.5) * dx, (g.nx, g.ny))


def test_local_uniform():
    g = _grid()
    ac, av = _all_active(g)
    one = np.ones((g.nx - 1, g.ny - 1), order="F")
    bf, bx, by = po.basal_friction(g, 1, ac, av, 3.0 * one, 4.0 * one, 2.0 * one)
    own = _owned(g)
    np.testing.assert_allclose(bf[own], 2.0 * 25.0 / SCYR, rtol=1e-15)
    # outflow halo (global_bc 1): the two rings of halo cells are zero
    m = np.ones((g.nx, g.ny), bool)
    m[own] = False
    assert (bf[m] == 0).all()
    assert (bx == 6.0).all() and (by == 8.0).all()


def test_standard_equals_local_on_constant_fields():
    g = _grid()
    ac, av = _all_active(g)
    rng = np.random.default_rng(3)
    u0, v0, b0 = rng.uniform(-50, 50), rng.uniform(-50, 50), rng.uniform(1, 1e4)
    one = np.ones((g.nx - 1, g.ny - 1), order="F")
    s, _, _ = po.basal_friction(g, 0, ac, av, u0 * one, v0 * one, b0 * one)
    loc, _, _ = po.basal_friction(g, 1, ac, av, u0 * one, v0 * one, b0 * one)
    own = _owned(g)
    np.testing.assert_allclose(s[own], b0 * (u0 * u0 + v0 * v0) / SCYR, rtol=1e-14)
    np.testing.assert_allclose(s[own], loc[own], rtol=1e-14)


def test_standard_gauss_closed_form():
    """beta and u linear in x, v linear in x too: the Gauss mean of the cubic."""
    g = _grid()
    ac, av = _all_active(g)
    dx = 1000.0
    xv, xc = _vertex_x(g, dx), _cell_x(g, dx)
    b0, b1 = 500.0, 0.3          # beta = b0 + b1 x
    a0, a1 = 20.0, 0.01          # u = a0 + a1 x
    c0, c1 = -5.0, -0.004        # v = c0 + c1 x
    bf, _, _ = po.basal_friction(g, 0, ac, av, a0 + a1 * xv, c0 + c1 * xv, b0 + b1 * xv)
    h = dx / 2.0                 # xi = (x - xc) / h on the reference element
    bc, bs = b0 + b1 * xc, b1 * h
    uc, us = a0 + a1 * xc, a1 * h
    vc, vs = c0 + c1 * xc, c1 * h
    want = (bc * (uc * uc + vc * vc) + bc * (us * us + vs * vs) / 3.0 + 2.0 * bs * (uc * us + vc * vs) / 3.0) / SCYR
    own = _owned(g)
    np.testing.assert_allclose(bf[own], want[own], rtol=1e-13)
    # the local option averages the same cubic at the corners xi = +-1:
    # bc (uc^2 + us^2) + 2 bs uc us (no 1/3 factors)
    loc, _, _ = po.basal_friction(g, 1, ac, av, a0 + a1 * xv, c0 + c1 * xv, b0 + b1 * xv)
    want_loc = (bc * (uc * uc + vc * vc) + bc * (us * us + vs * vs) + 2.0 * bs * (uc * us + vc * vs)) / SCYR
    np.testing.assert_allclose(loc[own], want_loc[own], rtol=1e-13)


def test_standard_orthogonal_variation():
    """beta varying in y and the speed in x: the mean factorises as
    mean(beta) mean(u^2)."""
    g = _grid()
    ac, av = _all_active(g)
    dx = 2000.0
    xv = _vertex_x(g, dx)
    yv = np.asfortranarray(np.broadcast_to((np.arange(g.ny - 1)[None, :] + 1.0) * dx, (g.nx - 1, g.ny - 1)))
    xc = _cell_x(g, dx)
    yc = np.broadcast_to((np.arange(g.ny)[None, :] + 0.5) * dx, (g.nx, g.ny))
    bf, _, _ = po.basal_friction(g, 0, ac, av, 10.0 + 0.02 * xv, np.zeros_like(xv), 100.0 + 0.05 * yv)
    h = dx / 2.0
    uc, us = 10.0 + 0.02 * xc, 0.02 * h
    want = (100.0 + 0.05 * yc) * (uc * uc + us * us / 3.0) / SCYR
    own = _owned(g)
    np.testing.assert_allclose(bf[own], want[own], rtol=1e-13)


def test_inactive_cells_and_vertices():
    g = _grid()
    ac, av = _all_active(g)
    ac[4, 4] = 0                  # 0-based cell (4, 4) = 1-based (5, 5)
    av[4, 4] = 0                  # vertex (5, 5): NE corner of cell (5, 5)
    one = np.ones((g.nx - 1, g.ny - 1), order="F")
    s, _, _ = po.basal_friction(g, 0, ac, av, one, one, one)
    assert s[4, 4] == 0.0
    np.testing.assert_allclose(s[5, 5], 2.0 / SCYR, rtol=1e-14)  # the standard option ignores the vertex mask
    loc, _, _ = po.basal_friction(g, 1, ac, av, one, one, one)
    # the local option: the four cells around vertex (5, 5) lose a quarter
    for (i, j) in ((4, 4), (5, 4), (4, 5), (5, 5)):
        np.testing.assert_allclose(loc[i, j], 0.75 * 2.0 / SCYR, rtol=1e-15)
    np.testing.assert_allclose(loc[3, 3], 2.0 / SCYR, rtol=1e-15)


def test_periodic_halo():
    g = _grid(bc=0)
    ac, av = _all_active(g)
    rng = np.random.default_rng(7)
    f = lambda: np.asfortranarray(rng.uniform(0.5, 2.0, (g.nx - 1, g.ny - 1)))
    for which in (0, 1):
        bf, _, _ = po.basal_friction(g, which, ac, av, f(), f(), f())
        n, m = g.nx - 4, g.ny - 4
        np.testing.assert_array_equal(bf[:2, 2:g.ny - 2], bf[n:n + 2, 2:g.ny - 2])
        np.testing.assert_array_equal(bf[g.nx - 2:, 2:g.ny - 2], bf[2:4, 2:g.ny - 2])
        np.testing.assert_array_equal(bf[:, :2], bf[:, m:m + 2])
        np.testing.assert_array_equal(bf[:, g.ny - 2:], bf[:, 2:4])


def test_traction_is_beta_times_velocity():
    g = _grid()
    ac, av = _all_active(g)
    rng = np.random.default_rng(11)
    u, v, b = (np.asfortranarray(rng.normal(size=(g.nx - 1, g.ny - 1))) for _ in range(3))
    _, bx, by = po.basal_friction(g, 1, ac, av, u, v, b)
    np.testing.assert_array_equal(bx, b * u)
    np.testing.assert_array_equal(by, b * v)


def test_no_ice_and_no_penetration_halos():
    """global_bc 3 (no_ice) zeroes three rings of cells after the friction
    is computed; global_bc 2 (no_penetration) wraps like periodic cells."""
    rng = np.random.default_rng(13)
    for bc in (2, 3):
        g = _grid(bc=bc)
        ac, av = _all_active(g)
        f = lambda: np.asfortranarray(rng.uniform(0.5, 2.0, (g.nx - 1, g.ny - 1)))
        u, v, b = f(), f(), f()
        for which in (0, 1):
            bf, _, _ = po.basal_friction(g, which, ac, av, u, v, b)
            if bc == 3:
                m = np.zeros((g.nx, g.ny), bool)
                m[3:g.nx - 3, 3:g.ny - 3] = True
                assert (bf[~m] == 0).all() and (bf[m] > 0).all()
                ref, _, _ = po.basal_friction(_grid(bc=1), which, ac, av, u, v, b)
                np.testing.assert_array_equal(bf[m], ref[m])
            else:
                ref, _, _ = po.basal_friction(_grid(bc=0), which, ac, av, u, v, b)
                np.testing.assert_array_equal(bf, ref)
