"""The reference's ice-stream test case (tests/stream: runStream.py and
stream.config) against its closed-form solutions: till whose yield stress
varies across the flow (which_ho_babc 15, beta = tauf / |u|), a kinematic
column, doubly periodic.  The oracle's Picard loop (BP, SSA, L1L2 and DIVA)
must reproduce Raymond (2000) and Schoof (2006) across the stream, far from
the prescribed column, to within 1.5 % of the centreline speed (measured
0.01-0.95 %: the discretisation error of 2 km cells and, for BP, L1L2 and
DIVA, the vertical shear the plug-flow solutions leave out).  This pins the
yield-stress law, the kinematic Dirichlet columns, the periodic halos with
the sloping bed and the outer loop as a whole.  CPU only."""
import numpy as np
import pytest

from cism_amd import synthetic as sy
from oracle import pyoracle as po

KW = dict(precond=1, linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-4, maxiter=300)


def _far_column(p):
    g = p.grid
    return g.ilo - 1 + 9                 # 7 columns from the kinematic one, the farthest in the period


@pytest.mark.parametrize("approx", ["bp", "ssa", "l1l2", "diva"])
@pytest.mark.parametrize("kind", ["raymond", "schoof"])
def test_oracle_reproduces_the_analytic_stream(kind, approx):
    p = sy.yield_stress_stream(kind)
    yv = p.meta["y_vertex"]
    ua = sy.stream_velocity(kind, yv)
    if approx == "bp":
        r = po.picard(p, algo="chrongear", **KW)
        u = r["uvel"][0, _far_column(p)]
    else:
        r = po.picard(p, two_d=True, approx=approx, **KW)
        u = (r["uvel3d"][0] if r.get("uvel3d") is not None else r["uvel"])[_far_column(p)]
    assert r["converged"]
    inside = np.abs(yv) < sy.STREAM_HALF_WIDTH
    err = np.max(np.abs(u[inside] - ua[inside])) / ua.max()
    assert err < 1.5e-2, err
    # beyond one cell outside the stream the strong till holds the ice
    beyond = np.abs(yv) > sy.STREAM_HALF_WIDTH + p.dy
    assert np.max(np.abs(u[beyond])) < 1e-3 * ua.max()


def test_stream_geometry_follows_the_reference_script():
    """Grid sizes, vertex rows on the stream edges and the driving stress of
    runStream.py's defaults (25 stream cells, dx = dy = 2 km)."""
    p = sy.yield_stress_stream("raymond")
    g = p.grid
    assert (g.nx - 4, g.ny - 4) == (15, 31) and p.dx == p.dy == 2000.0
    yv = p.meta["y_vertex"]
    assert np.isclose(np.abs(yv), sy.STREAM_HALF_WIDTH).sum() == 2
    assert sy.stream_driving_stress() == pytest.approx(8927.1)
    q = sy.yield_stress_stream("schoof")
    assert q.grid.ny - 4 == 25 + 2 * 8
    # Schoof's yield stress equals the driving stress at |y| = L
    L = sy.STREAM_HALF_WIDTH / (sy.STREAM_SCHOOF_M + 1.0) ** (1.0 / sy.STREAM_SCHOOF_M)
    assert sy.stream_yield_stress("schoof", np.array([L]))[0] == pytest.approx(sy.stream_driving_stress())
