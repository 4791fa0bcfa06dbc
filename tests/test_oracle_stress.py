"""compute_internal_stress (glissade_velo_higher.F90:7046-7281) in the
oracle, pinned where the finite-element strain rates are exact: on the
inclined slab every element is a parallelepiped, so velocity fields linear
in the physical coordinates are reproduced exactly.  Vertical shear
u = a z gives tau_xz = eta a and nothing else; pure strain u = e x,
v = -e y gives tau_xx = 2 eta e = -tau_yy and tau_eff = 2 eta e, with eta
the Glen viscosity of that strain rate.  CPU only."""
import numpy as np
import pytest

from cism_amd import synthetic as sy
from oracle import pyoracle as po


def _nodes(p, geo):
    g = p.grid
    xV, yV = p.vertex_coords()
    sig = np.asarray(p.sigma)
    z = geo["stagusrf"][None] - sig[:, None, None] * geo["stagthck"][None]
    return np.broadcast_to(xV, z.shape), np.broadcast_to(yV, z.shape), z


def _eta(p, geo, esq):
    n = p.n_glen
    return geo["flwafact"] * (p.effstrain_min ** 2 + esq) ** ((1.0 - n) / n / 2.0)


def _active(p, geo):
    g = p.grid
    m = np.zeros((g.nz - 1, g.nx, g.ny), bool)
    m[:, 2:g.nx - 1, 2:g.ny - 1] = (geo["active_cell"][2:g.nx - 1, 2:g.ny - 1] == 1)[None]
    return m


def test_vertical_shear():
    p = sy.slab(n_cells=6, nz=5)
    geo = po.geometry(p)
    x, y, z = _nodes(p, geo)
    a = 0.02
    tau = po.internal_stress(p, np.asfortranarray(a * z), np.zeros(z.shape, order="F"), geo)
    m = _active(p, geo)
    assert m.sum() > 0
    want = _eta(p, geo, 0.25 * a * a) * a
    np.testing.assert_allclose(tau["tau_xz"][m], want[m], rtol=1e-12)
    for k in ("tau_yz", "tau_xx", "tau_yy", "tau_xy"):
        assert np.abs(tau[k][m]).max() < 1e-10 * np.abs(want[m]).max()
    np.testing.assert_allclose(tau["tau_eff"][m], np.abs(want[m]), rtol=1e-12)
    assert (tau["tau_eff"][~m] == 0).all()


def test_pure_strain():
    p = sy.slab(n_cells=6, nz=5)
    geo = po.geometry(p)
    x, y, z = _nodes(p, geo)
    e = 1.0e-3
    tau = po.internal_stress(p, np.asfortranarray(e * x), np.asfortranarray(-e * y), geo)
    m = _active(p, geo)
    two_eta_e = 2.0 * _eta(p, geo, e * e) * e
    np.testing.assert_allclose(tau["tau_xx"][m], two_eta_e[m], rtol=1e-12)
    np.testing.assert_allclose(tau["tau_yy"][m], -two_eta_e[m], rtol=1e-12)
    np.testing.assert_allclose(tau["tau_eff"][m], two_eta_e[m], rtol=1e-12)
    for k in ("tau_xz", "tau_yz", "tau_xy"):
        assert np.abs(tau[k][m]).max() < 1e-10 * two_eta_e[m].max()


def _shear_plus_strain(p, geo, a, e):
    x, y, z = _nodes(p, geo)
    return np.asfortranarray(a * z + e * x), np.asfortranarray(-e * y)


def test_approximation_branches():
    """u = a z + e x, v = -e y on the slab, whose strain rates are exact:
    BP (and DIVA) keep both, with eta of e^2 + a^2/4; SSA drops the vertical
    shear from the strain rates and the viscosity; SIA keeps only it; L1L2
    uses the viscosity it is given."""
    p = sy.slab(n_cells=6, nz=5)
    geo = po.geometry(p)
    a, e = 0.02, 1.0e-3
    u, v = _shear_plus_strain(p, geo, a, e)
    m = _active(p, geo)
    for approx, esq, has_xz, has_xx in ((2, e * e + 0.25 * a * a, True, True), (4, e * e + 0.25 * a * a, True, True),
                                        (1, e * e, False, True), (0, 0.25 * a * a, True, False)):
        tau = po.internal_stress(p, u, v, geo, approx=approx)
        eta = _eta(p, geo, esq)
        want_xz = eta * a if has_xz else 0.0 * eta
        want_xx = 2.0 * eta * e if has_xx else 0.0 * eta
        np.testing.assert_allclose(tau["tau_xz"][m], want_xz[m], rtol=1e-12, atol=1e-10 * np.abs(eta * a).max())
        np.testing.assert_allclose(tau["tau_xx"][m], want_xx[m], rtol=1e-12, atol=1e-10 * np.abs(eta * e).max())
        np.testing.assert_allclose(tau["tau_yy"][m], -want_xx[m], rtol=1e-12, atol=1e-10 * np.abs(eta * e).max())
        np.testing.assert_allclose(tau["tau_eff"][m], np.sqrt(want_xx ** 2 + want_xz ** 2)[m], rtol=1e-12)
    # L1L2: the viscosity is the layer field passed in, read per element
    g = p.grid
    rng = np.random.default_rng(2)
    efvs = np.asfortranarray(rng.uniform(1e5, 1e7, (g.nz - 1, g.nx, g.ny)))
    tau = po.internal_stress(p, u, v, geo, approx=3, efvs=efvs)
    np.testing.assert_allclose(tau["tau_xz"][m], (efvs * a)[m], rtol=1e-12)
    np.testing.assert_allclose(tau["tau_xx"][m], (2.0 * efvs * e)[m], rtol=1e-12)
    with pytest.raises(ValueError, match="efvs"):
        po.internal_stress(p, u, v, geo, approx=3)
