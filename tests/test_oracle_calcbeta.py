"""Pins of the oracle's glissade_calcbeta (oracle/oracle_basal.c,
glissade_basal_traction.F90:69-719) by what the mathematics fixes: limits
where a sliding law reduces to a closed form, identities between two laws
(the bump-parameter Coulomb law equals the Schoof law when
lambda_max/m_max * A = (coulomb_c/powerlaw_c)^m), the stress each law
prescribes recovered as beta * speed, tan(22.5 deg) = sqrt(2) - 1 in the
old pseudo-plastic law, the f_ground limiting rules and the invalid-beta
error.  CPU only."""
from types import SimpleNamespace

import numpy as np
import pytest

from cism_amd.basal import Friction
from cism_amd.grid import BC_NO_ICE, BC_PERIODIC, Grid
from oracle import pyoracle as po

SMALL2 = 1.0e-2 * 1.0e-2


def _case(fr, u, v, *, grid=None, f_ground=None, topg=None, thck=None, flwa=None, eus=0.0):
    g = grid or Grid.from_owned(8, 6, 3)
    shp = (g.nvx, g.nvy)
    p = SimpleNamespace(grid=g, friction=fr, eus=eus,
                        topg=np.zeros((g.nx, g.ny)) if topg is None else topg,
                        thck=np.full((g.nx, g.ny), 1000.0) if thck is None else thck,
                        flwa=np.full((g.nz - 1, g.nx, g.ny), 1.0e-16) if flwa is None else flwa)
    geo = {"f_ground": np.ones(shp) if f_ground is None else f_ground}
    b = po.calcbeta(p, geo, np.broadcast_to(u, shp), np.broadcast_to(v, shp))
    return g, b


def _owned(g, a):
    """owned vertices but the last row and column (zeroed by the no-ice halo rule)"""
    return a[g.ilo - 1:g.ihi - 1, g.jlo - 1:g.jhi - 1]


def _speeds(g, seed=0, lo=0.0, hi=800.0):
    rng = np.random.default_rng(seed)
    shp = (g.nvx, g.nvy)
    return rng.uniform(lo, hi, shp), rng.uniform(-hi, hi, shp)


def _speed(u, v, umax=0.0):
    s = np.sqrt(u * u + v * v + SMALL2)
    return np.minimum(s, umax) if umax > 0 else s


G = Grid.from_owned(8, 6, 3)


def test_power_law_stress_and_linear_limit():
    u, v = _speeds(G, 1)
    pc = np.random.default_rng(2).uniform(1e3, 2e4, (G.nvx, G.nvy))
    _, b = _case(Friction(9, powerlaw_m=1.0, powerlaw_c=pc, beta_grounded_min=0.0), u, v)
    assert np.array_equal(_owned(G, b), _owned(G, pc))          # m = 1: beta = C exactly
    _, b = _case(Friction(9, powerlaw_m=3.0, powerlaw_c=pc, beta_grounded_min=0.0), u, v)
    s = _speed(u, v)
    np.testing.assert_allclose(_owned(G, b * s), _owned(G, pc * np.cbrt(s)), rtol=2e-15)
    # the speed cap: beyond umax, beta is that of umax
    _, b = _case(Friction(9, powerlaw_m=3.0, powerlaw_c=pc, beta_grounded_min=0.0, beta_powerlaw_umax=200.0), u, v)
    su = _speed(u, v, 200.0)
    np.testing.assert_allclose(_owned(G, b), _owned(G, pc * np.cbrt(su) / su), rtol=2e-15)


def test_pseudo_plastic_linear_and_plastic_limits():
    u, v = _speeds(G, 3)
    rng = np.random.default_rng(4)
    N, cc = rng.uniform(1e5, 1e7, (G.nvx, G.nvy)), rng.uniform(0.1, 0.9, (G.nvx, G.nvy))
    kw = dict(effecpress_stag=N, coulomb_c=cc, beta_grounded_min=0.0, pseudo_plastic_u0=137.0)
    _, b = _case(Friction(2, pseudo_plastic_q=1.0, **kw), u, v)      # linear: tau_c / u0
    assert np.array_equal(_owned(G, b), _owned(G, (N * cc) / 137.0))
    _, b = _case(Friction(2, pseudo_plastic_q=0.0, **kw), u, v)      # plastic: tau_b = tau_c
    np.testing.assert_allclose(_owned(G, b * _speed(u, v)), _owned(G, N * cc), rtol=2e-15)


def test_old_pseudo_plastic_angle_from_bed():
    u, v = _speeds(G, 5)
    N = np.full((G.nvx, G.nvy), 2.0e6)
    topg = np.full((G.nx, G.ny), 200.0)            # midway between bedmin -300 and bedmax 700: phi = 22.5 deg
    topg[4, 5] = -1000.0                           # below bedmin: phimin at the vertex with the cell's indices
    topg[6, 3] = 5000.0                            # above bedmax: phimax
    fr = Friction(3, pseudo_plastic_q=1.0, pseudo_plastic_u0=100.0, effecpress_stag=N, beta_grounded_min=0.0)
    _, b = _case(fr, u, v, topg=topg)
    tanphi = b * 100.0 / N
    mid = np.ones_like(tanphi, dtype=bool)
    mid[4, 5] = mid[6, 3] = False
    own = np.zeros_like(mid)
    own[G.ilo - 1:G.ihi - 1, G.jlo - 1:G.jhi - 1] = True
    np.testing.assert_allclose(tanphi[mid & own], np.sqrt(2.0) - 1.0, rtol=1e-14)
    np.testing.assert_allclose(tanphi[4, 5], np.tan(np.deg2rad(5.0)), rtol=1e-14)
    np.testing.assert_allclose(tanphi[6, 3], np.tan(np.deg2rad(40.0)), rtol=1e-14)
    # eus shifts the bed: topg 200 with eus 200 is bed 0, phi = 5 + 0.3 * 35 = 15.5 deg
    _, b = _case(fr, u, v, topg=np.full((G.nx, G.ny), 200.0), eus=200.0)
    np.testing.assert_allclose(_owned(G, b * 100.0 / N), np.tan(np.deg2rad(15.5)), rtol=1e-14)


def test_zoet_iverson_linear_form_and_cap():
    u, v = _speeds(G, 6)
    rng = np.random.default_rng(7)
    N, cc = rng.uniform(1e5, 1e7, (G.nvx, G.nvy)), rng.uniform(0.1, 0.9, (G.nvx, G.nvy))
    fr = Friction(7, powerlaw_m=1.0, zoet_iverson_ut=50.0, zoet_iverson_nmax=4.0e6, effecpress_stag=N, coulomb_c=cc,
                  beta_grounded_min=0.0)
    _, b = _case(fr, u, v)
    s = _speed(u, v)
    want = cc * np.minimum(N, 4.0e6) * s / (s + 50.0)       # tau_b = C N u / (u + u_t) for m = 1
    np.testing.assert_allclose(_owned(G, b * s), _owned(G, want), rtol=1e-14)


@pytest.mark.parametrize("m", [1.0, 3.0])
def test_bump_coulomb_law_equals_schoof_law(m):
    """Coulomb friction with lambda = (wavelength / max_slope) * A_stag is
    the Schoof law with powerlaw_c = coulomb_c * lambda^(-1/m)."""
    u, v = _speeds(G, 8)
    rng = np.random.default_rng(9)
    N, cc = rng.uniform(1e4, 1e6, (G.nvx, G.nvy)), rng.uniform(0.1, 0.9, (G.nvx, G.nvy))
    flwa = np.full((G.nz - 1, G.nx, G.ny), 1.0e-16)
    flwa[-1] = rng.uniform(0.5e-16, 2e-16, (G.nx, G.ny))     # the basal layer is the one read
    thck = np.full((G.nx, G.ny), 500.0)
    thck[3, 4] = thck[7, 2] = 0.0                            # cells left out of the masked stagger
    a = flwa[-1] * (thck > 0)
    msum = (thck > 0).astype(float)
    stag = lambda f: f[:-1, 1:] + f[1:, 1:] + f[:-1, :-1] + f[1:, :-1]
    a_stag = stag(a) / stag(msum)
    lam = (2.0 / 0.5) * a_stag
    pc = cc * lam ** (-1.0 / m)
    csf = rng.uniform(0.5, 1.0, (G.nvx, G.nvy))
    kw = dict(powerlaw_m=m, effecpress_stag=N, coulomb_c=cc, beta_grounded_min=0.0, c_space_factor_stag=csf)
    _, b10 = _case(Friction(10, **kw), u, v, thck=thck, flwa=flwa)
    _, b11 = _case(Friction(11, powerlaw_c=pc, **kw), u, v, thck=thck, flwa=flwa)
    np.testing.assert_allclose(_owned(G, b10), _owned(G, b11), rtol=1e-13)
    assert (_owned(G, b10) <= 1.0e8).all()
    # the cap of 1e8 (:501-503)
    _, b = _case(Friction(11, powerlaw_c=pc * 1e9, powerlaw_m=m, effecpress_stag=N * 1e9, coulomb_c=cc,
                          beta_grounded_min=0.0), np.zeros_like(u), np.zeros_like(v))
    assert (_owned(G, b) == 1.0e8).all()


def test_modified_schoof_and_tsai_stresses():
    u, v = _speeds(G, 10)
    rng = np.random.default_rng(11)
    shp = (G.nvx, G.nvy)
    N, cc, pc = rng.uniform(1e4, 1e6, shp), rng.uniform(0.1, 0.9, shp), rng.uniform(1e3, 2e4, shp)
    s = _speed(u, v)
    tp, tc = pc * np.cbrt(s), cc * N
    kw = dict(powerlaw_m=3.0, effecpress_stag=N, coulomb_c=cc, powerlaw_c=pc, beta_grounded_min=0.0)
    _, b = _case(Friction(12, **kw), u, v)
    np.testing.assert_allclose(1.0 / _owned(G, b * s), _owned(G, 1.0 / tp + 1.0 / tc), rtol=1e-14)
    _, b = _case(Friction(12, **{**kw, "effecpress_stag": np.zeros(shp)}), u, v)    # tau_c = 0: the power law
    np.testing.assert_allclose(_owned(G, b * s), _owned(G, tp), rtol=1e-14)
    _, b = _case(Friction(12, **{**kw, "powerlaw_c": np.zeros(shp)}), u, v)         # tau_p = 0: Coulomb
    np.testing.assert_allclose(_owned(G, b * s), _owned(G, tc), rtol=1e-14)
    _, b = _case(Friction(13, **kw), u, v)
    np.testing.assert_allclose(_owned(G, b * s), _owned(G, np.minimum(tp, tc)), rtol=1e-14)


def test_effective_pressure_power_law_and_yield_stress():
    u, v = _speeds(G, 12)
    rng = np.random.default_rng(13)
    shp = (G.nvx, G.nvy)
    N = rng.uniform(1e4, 1e6, shp)
    _, b = _case(Friction(14, effecpress_stag=N, beta_grounded_min=0.0), u, v)
    s = _speed(u, v)
    # u_b = k tau_b^3 N^-1 (Cuffey and Paterson eq. 7.17 with p = 3, q = 1)
    np.testing.assert_allclose(_owned(G, 8.4e-9 * (b * s) ** 3 / N), _owned(G, s), rtol=1e-13)
    tauf = rng.uniform(1e3, 1e5, shp)
    _, b = _case(Friction(15, mintauf=tauf, beta_grounded_min=0.0, beta_powerlaw_umax=1.0), u, v)
    np.testing.assert_allclose(_owned(G, b * _speed(u, v)), _owned(G, tauf), rtol=2e-15)   # no speed cap


def test_fixed_fields_match_the_fixed_beta_path():
    """Options 0, 1, 4, 5 with beta_grounded_min 0 give beta_internal of the
    fixed-beta path (beta times f_ground, then the halo)."""
    u, v = _speeds(G, 14)
    rng = np.random.default_rng(15)
    fg = (rng.uniform(0, 1, (G.nvx, G.nvy)) > 0.3).astype(float)
    ext = rng.uniform(10, 1e4, (G.nvx, G.nvy))
    for fr, want in ((Friction(0, ho_beta_const=321.0, beta_grounded_min=0.0), np.full_like(ext, 321.0)),
                     (Friction(4, beta_grounded_min=0.0), np.full_like(ext, 1.0e10)),
                     (Friction(5, beta_external=ext, beta_grounded_min=0.0), ext)):
        _, b = _case(fr, u, v, f_ground=fg)
        w = np.asfortranarray(want * fg)
        po.halo(G, w)
        assert np.array_equal(b, w)
    mask = (rng.uniform(0, 1, (G.nvx, G.nvy)) > 0.5).astype(np.int32)
    _, b = _case(Friction(1, ho_beta_small=7.0, ho_beta_large=9.0, bpmp_mask=mask, beta_grounded_min=0.0), u, v)
    assert np.array_equal(_owned(G, b), _owned(G, np.where(mask == 1, 7.0, 9.0)))


def test_grounded_minimum_before_and_after_f_ground():
    u, v = _speeds(G, 16)
    fg = np.full((G.nvx, G.nvy), 0.5)
    fg[5, 4] = 0.0
    _, b0 = _case(Friction(0, ho_beta_const=1.0, beta_grounded_min=3.0, which_beta_limit=0), u, v, f_ground=fg)
    _, b1 = _case(Friction(0, ho_beta_const=1.0, beta_grounded_min=3.0, which_beta_limit=1), u, v, f_ground=fg)
    assert b0[4, 3] == 3.0 and b1[4, 3] == 1.5          # floor after / before the weighting
    assert b0[5, 4] == 0.0 and b1[5, 4] == 0.0          # floating vertex: no floor


def test_invalid_beta_is_an_error_naming_the_vertex():
    u, v = _speeds(G, 17)
    cc = np.full((G.nvx, G.nvy), 0.5)
    cc[3, 6] = np.nan            # a negative value is floored at beta_grounded_min where f_ground > 0
    with pytest.raises(ValueError, match="i=4 j=7"):
        _case(Friction(2, effecpress_stag=np.full_like(cc, 1e6), coulomb_c=cc, beta_grounded_min=0.0), u, v)
    u[2, 2] = np.nan
    with pytest.raises(ValueError, match="i=3 j=3"):
        _case(Friction(9, powerlaw_c=np.full_like(cc, 1e4), beta_grounded_min=0.0), u, v)


def test_staggered_halo_is_applied():
    g = Grid.from_owned(8, 6, 3, global_bc=BC_PERIODIC)
    u, v = _speeds(g, 18)
    pc = np.random.default_rng(19).uniform(1e3, 2e4, (g.nvx, g.nvy))
    _, b = _case(Friction(9, powerlaw_c=pc, beta_grounded_min=0.0), u, v, grid=g)
    w = np.asfortranarray(b.copy())
    po.halo(g, w)
    assert np.array_equal(b, w)
    assert np.array_equal(b[0], b[g.ihi - 2])            # periodic copy of the owned edge
    g0 = Grid.from_owned(8, 6, 3, global_bc=BC_NO_ICE)
    _, b = _case(Friction(9, powerlaw_c=pc, beta_grounded_min=0.0), u, v, grid=g0)
    assert (b[:2] == 0).all() and (b[:, :2] == 0).all()


def test_ismip_hom_c_field():
    """Option 8 on a periodic 8 x 8-cell task: the field of ISMIP-HOM C at
    the vertices, 1000 + 1000 sin(w x) sin(w y) with x, y the vertex
    positions: mean 1000 over a period, 2000 and 0 at the quarter points,
    odd about 1000 under a half-period shift, periodic halo."""
    g = Grid.from_owned(8, 8, 3, global_bc=BC_PERIODIC)
    u, v = _speeds(g, 20)
    p = SimpleNamespace(grid=g, friction=Friction(8, beta_grounded_min=0.0), eus=0.0, dx=2500.0, dy=2500.0,
                        topg=np.zeros((g.nx, g.ny)), thck=np.full((g.nx, g.ny), 1000.0),
                        flwa=np.full((g.nz - 1, g.nx, g.ny), 1e-16))
    b = po.calcbeta(p, {"f_ground": np.ones((g.nvx, g.nvy))}, u, v)
    own = b[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi]            # vertex (i, j) at x = (i - 2) dx: quarter points at 2, 6
    assert own.mean() == pytest.approx(1000.0, rel=1e-14)
    assert own[1, 1] == pytest.approx(2000.0, rel=1e-14) and own[5, 5] == pytest.approx(2000.0, rel=1e-14)
    assert abs(own[1, 5]) < 1e-10 and abs(own[5, 1]) < 1e-10
    np.testing.assert_allclose(np.roll(own, 4, axis=0) - 1000.0, -(own - 1000.0), atol=1e-10)
    w = np.asfortranarray(b.copy())
    po.halo(g, w)
    assert np.array_equal(b, w)


@pytest.mark.parametrize("experiment", ["a", "c"])
def test_ismip_hom_force_balance(experiment):
    """ISMIP-HOM A and C (tests/ismip-hom): on a doubly periodic domain the
    stiffness of a uniform velocity is zero, so at a converged Picard step
    the basal drag balances the driving stress exactly: the mean over the
    owned vertices of beta u_b / cos(theta_b) (the local basal term is
    taken over the sloping bed's area) equals rho g mean(H) tan(alpha) to
    the Picard tolerance, and the mean of beta v_b / cos(theta_b)
    vanishes."""
    from cism_amd import synthetic as sy
    p = sy.ismip_hom(experiment, 80.0, n=20, nz=7)
    r = po.picard(p, algo="chrongear", precond=2, linear_tol=1e-10, linear_maxiters=2000, resid_target=1e-6,
                  maxiter=100)
    assert r["converged"]
    g = p.grid
    geo = po.geometry(p)
    ub, vb = r["uvel"][-1], r["vvel"][-1]
    beta = po.beta_internal(p, geo, ub, vb)
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    alpha = (0.5 if experiment == "a" else 0.1) * np.pi / 180.0
    taud = 910.0 * 9.81 * geo["stagthck"][own].mean() * np.tan(alpha)
    c = np.cos(geo["theta_basal"])
    assert (beta * ub / c)[own].mean() == pytest.approx(taud, rel=1e-8)
    assert abs((beta * vb / c)[own].mean()) < 1e-9 * taud
