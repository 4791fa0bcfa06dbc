"""netCDF state interchange and the dense matrix dump (cism_amd/cismio.py).

Pins: the EISMINT-2 bed topographies the reference's tests hold
(tests/EISMINT/EISMINT-2/{trough,mound}.nc, copied under tests/golden/) read
with the right orientation and spacing; states written here read back
exactly with halos that follow the halo rules; the dense dump holds the
27-point operator (symmetric blocks, Auv = Avu^T) to the 8 significant
digits of its '(e16.8)' format and maps back to the structured layout.
"""
import os

import numpy as np
import pytest

from cism_amd import cismio, geometry as ge, synthetic
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC, Grid
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_eismint2_trough_orientation_and_spacing():
    """trough.nc: a bed that is level (1000 m) in its west part and drops
    eastwards towards 0 m, the same along every row (mirror-symmetric in y);
    61 x 61 cells of 25 km."""
    p = cismio.read_cism_state(os.path.join(GOLDEN, "trough.nc"), nz=5)
    g = p.grid
    assert (g.nx, g.ny) == (65, 65) and p.dx == 25000.0 and p.dy == 25000.0
    t = p.topg[3:62, 3:62]      # the file's cells, (x, y), less the outer ring the no_ice halo zeroes
    assert t[:29, :].min() == pytest.approx(1000.0) and t[:29, :].max() == pytest.approx(1000.0)
    east = t[35:, 29]
    assert np.all(np.diff(east) < 0) and east[-1] < 40.0
    assert np.array_equal(t, t[:, ::-1])      # symmetric in y
    assert not np.array_equal(t, t[::-1, :])  # but not in x
    assert np.all(p.thck == 0.0)
    # no_ice halo rule: three outer rows of cells are zero
    assert np.all(p.topg[:3] == 0.0) and np.all(p.topg[-3:] == 0.0)


def test_eismint2_mound_is_centred():
    """mound.nc: a bed mound rising to 500 m above a 0 m floor, symmetric in y."""
    p = cismio.read_cism_state(os.path.join(GOLDEN, "mound.nc"), nz=5, global_bc=BC_OUTFLOW)
    t = p.topg[2:63, 2:63]       # outflow: the halo rule zeroes only the two halo rows
    assert t.max() == pytest.approx(500.0) and t.min() == pytest.approx(0.0)
    assert np.abs(t - t[:, ::-1]).max() < 0.05


@pytest.mark.parametrize("bc", [BC_NO_ICE, BC_PERIODIC, BC_OUTFLOW])
def test_state_round_trip(tmp_path, bc):
    p = synthetic.ice_stream(24, nz=6, seed=4)
    p.grid.global_bc = bc
    g = p.grid
    path = str(tmp_path / "state.nc")
    cismio.write_cism_state(path, p)
    q = cismio.read_cism_state(path, global_bc=bc, whichbabc=p.whichbabc)
    assert q.grid == Grid.from_owned(g.nx - 4, g.ny - 4, g.nz, bc)
    assert np.array_equal(q.sigma, p.sigma) and (q.dx, q.dy) == (p.dx, p.dy)
    w = 3 if bc == BC_NO_ICE else 2     # no_ice: the halo rule also zeroes the outermost owned row
    own_c = (slice(w, g.nx - w), slice(w, g.ny - w))
    own_v = (slice(w, g.nx - 1 - w), slice(w, g.ny - 1 - w))
    assert np.array_equal(q.thck[own_c], p.thck[own_c]) and np.array_equal(q.topg[own_c], p.topg[own_c])
    assert np.array_equal(q.flwa[(slice(None),) + own_c], p.flwa[(slice(None),) + own_c])
    assert np.array_equal(q.beta[own_v], p.beta[own_v])
    assert np.array_equal(q.uvel[(slice(None),) + own_v], p.uvel[(slice(None),) + own_v])
    # halos follow the one-task halo rules of the global BC
    for a in (q.thck, q.topg):
        b = a.copy(order="F")
        ge.cell_halo_2d(q.grid, b)
        assert np.array_equal(a, b)
    b = q.beta.copy(order="F")
    ge.staggered_halo_2d(q.grid, b)
    assert np.array_equal(b, q.beta)


def test_velocity_output_round_trip(tmp_path):
    p = synthetic.dome()
    g = p.grid
    rng = np.random.default_rng(1)
    u = np.asfortranarray(rng.standard_normal(g.vec_shape))
    v = np.asfortranarray(rng.standard_normal(g.vec_shape))
    ef = np.asfortranarray(rng.random((g.nz - 1, g.nx, g.ny)))
    path = str(tmp_path / "out.nc")
    cismio.write_cism_velocity(path, p, u, v, efvs=ef)
    q = cismio.read_cism_state(path)
    own_v = (slice(None), slice(2, g.nx - 3), slice(2, g.ny - 3))
    assert np.array_equal(q.uvel[own_v], u[own_v]) and np.array_equal(q.vvel[own_v], v[own_v])
    own_c = (slice(None), slice(3, g.nx - 3), slice(3, g.ny - 3))
    assert np.array_equal(q.meta["efvs"][own_c], ef[own_c])


def test_netcdf4_is_refused(tmp_path):
    path = tmp_path / "x.nc"
    path.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ValueError, match="NETCDF3"):
        cismio.read_cism_state(str(path), nz=3)


def test_matrix_dump_round_trip(tmp_path):
    """write_matrix_elements_3d's dense dump of an assembled dome system:
    rows numbered by owned active vertex (j, then i) and level, the blocks
    symmetric (Auu, Avv) and transposed (Auv = Avu^T) like the operator,
    and reading it back gives the 27-point rows to 8 significant digits."""
    p = synthetic.dome(n_cells=8, nz=4)
    g = p.grid
    s = po.assemble(p)
    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
    act = s["active_vertex"]
    nn = cismio.write_matrix_elements_3d(g, act, *A, s["bu"], s["bv"], "dome", str(tmp_path))
    nodeID, idx = cismio.node_numbering(g, act)
    assert nn == len(idx) == int(np.count_nonzero(act[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi])) * g.nz
    assert tuple(idx[0]) == (idx[0][0], idx[0][1], 1) and idx[1][2] == 2
    d = cismio.read_matrix_elements_3d("dome", str(tmp_path))
    assert d["Auu"].shape == (nn, nn)
    sc = np.abs(d["Auu"]).max()
    # rows of the owned ring that the no_ice halo rule zeroes are dumped as
    # zero rows; among the others the blocks are symmetric / transposed
    live = np.flatnonzero(np.diag(d["Auu"]) != 0.0)
    assert 0 < len(live) < nn
    sub = lambda a: a[np.ix_(live, live)]
    assert np.abs(sub(d["Auu"]) - sub(d["Auu"]).T).max() <= 1e-7 * sc
    assert np.abs(sub(d["Avv"]) - sub(d["Avv"]).T).max() <= 1e-7 * sc
    assert np.abs(sub(d["Auv"]) - sub(d["Avu"]).T).max() <= 1e-7 * sc
    blocks, (bu, bv) = cismio.dense_to_structured(g, act, d)
    own = np.zeros(g.vec_shape, bool)
    for i, j, k in idx:
        own[k - 1, i - 1, j - 1] = True
    for got, want in zip(blocks, A):
        w = np.where(own[None], want, 0.0)
        # entries that couple to nodes without a row are not dumped
        for i, j, k in idx:
            for m in range(27):
                iA, jA, kA = m % 3 - 1, (m // 3) % 3 - 1, m // 9 - 1
                kk, ii, jj = k + kA, i + iA, j + jA
                if not (1 <= kk <= g.nz and nodeID[kk - 1, ii - 1, jj - 1] > 0):
                    w[m, k - 1, i - 1, j - 1] = 0.0
        assert np.all(np.abs(got - w) <= 5.1e-8 * np.abs(w) + 1e-300)
    assert np.all(np.abs(bu - np.where(own, s["bu"], 0.0)) <= 5.1e-8 * np.abs(s["bu"]) + 1e-300)


def test_dump_reads_fortran_three_digit_exponents(tmp_path):
    (tmp_path / "Auu.t").write_text("     1  0.10000000-299  0.20000000E+01\n     2  0.00000000E+00 -0.15000000E+03\n")
    for n in ("Auv", "Avu", "Avv"):
        (tmp_path / f"{n}.t").write_text("     1  0.0 0.0\n     2  0.0 0.0\n")
    for n in ("bu", "bv"):
        (tmp_path / f"{n}.t").write_text("     1  0.5E+00\n     2 -0.25E+01\n")
    d = cismio.read_matrix_elements_3d("t", str(tmp_path))
    assert d["Auu"][0, 0] == 1e-300 and d["Auu"][1, 1] == -150.0 and d["bv"][1] == -2.5
