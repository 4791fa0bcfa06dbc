"""Seeded synthetic ice geometries for the velocity solve.

Every field is generated here from a seed; nothing is read from disk.  A
:class:`Problem` holds the raw model state that ``glissade_velo_higher_solve``
receives (thickness, bed, flow factor, basal friction, velocity guess, in the
solver's yr-based units, glissade_velo_higher.F90:3922-3978) plus the options
that select the BP path.  Cell (i, j) (1-based, halos included) is centred
at ``((i - 2.5) dx, (j - 2.5) dy)``; vertex (i, j) is its NE corner.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import BC_NO_ICE, BC_PERIODIC, Grid, glide_sigma

RHOI_SYN, RHOO_SYN = 910.0, 1028.0   # ice and ocean densities (glimmer_physcon.F90:67-68)

# which_ho_babc values (glide_types.F90:253-268)
HO_BABC_BETA_CONSTANT = 0
HO_BABC_BETA_LARGE = 4
HO_BABC_BETA_EXTERNAL = 5
HO_BABC_NO_SLIP = 6


@dataclass
class Problem:
    grid: Grid
    dx: float
    dy: float
    sigma: np.ndarray            # (nz,)
    thck: np.ndarray             # (nx, ny) m
    topg: np.ndarray             # (nx, ny) m
    flwa: np.ndarray             # (nz-1, nx, ny) Pa^-n yr^-1
    beta: np.ndarray             # (nx-1, ny-1) Pa yr/m
    uvel: np.ndarray             # (nz, nx-1, ny-1) m/yr
    vvel: np.ndarray
    eus: float = 0.0
    thklim: float = 100.0        # glide_types.F90:2426
    max_slope: float = 1.0       # glide_types.F90:2569
    gradient_margin: int = 1     # which_ho_gradient_margin default (glide_types.F90:1021)
    thck_gradient_ramp: float = 0.0
    whichbabc: int = HO_BABC_BETA_EXTERNAL
    beta_const: float = 0.0
    n_glen: float = 3.0
    effstrain_min: float = 1.0e-8   # glide_types.F90:2567
    whichefvs: int = 2
    efvs_constant: float = 0.0
    assemble_lateral: int = 0    # which_ho_assemble_lateral (glide_types.F90:1067): 0 standard, 1 local
    assemble_beta: int = 1       # which_ho_assemble_beta (glide_types.F90:359-360, 1050): 0 standard, 1 local
    assemble_taud: int = 1       # which_ho_assemble_taud (glide_types.F90:362-363, 1042): 0 standard, 1 local
    kinbcmask: np.ndarray | None = None   # (nx-1, ny-1) int: 1 = u, v prescribed (from uvel/vvel) in the column
    which_ho_ground: int = 0     # glide_types.F90:378-380, 1090: 0 no GLP, 1 GLP for basal friction
    which_ho_flotation_function: int = 2   # :389-392, 1114: 0 Pattyn, 1 inverse Pattyn, 2 linear
    which_ho_fground_no_glp: int = 0       # :382-383, 1099: 0 grounded neighbour, 1 flotation function
    topg_raised: np.ndarray | None = None  # (nx, ny) m: the bed of flotation function 3 (glissade.F90:3330)
    friction: object = None      # basal.Friction: beta from the basal velocity in every Picard iteration
                                 # (glissade_calcbeta); None: the fixed beta of whichbabc 0, 4, 5 or no-slip
    meta: dict = field(default_factory=dict)

    def cell_centres(self):
        g = self.grid
        xc = (np.arange(1, g.nx + 1) - 2.5) * self.dx
        yc = (np.arange(1, g.ny + 1) - 2.5) * self.dy
        return xc, yc

    def vertex_coords(self):
        g = self.grid
        x0 = (np.arange(1, g.nx) - 2.0) * self.dx
        y0 = (np.arange(1, g.ny) - 2.0) * self.dy
        xV = np.asfortranarray(np.repeat(x0[:, None], g.ny - 1, axis=1))
        yV = np.asfortranarray(np.repeat(y0[None, :], g.nx - 1, axis=0))
        return xV, yV


def _smooth_field(rng, shape, corr_cells: float) -> np.ndarray:
    """Unit-variance Gaussian random field with a Gaussian spectrum (FFT filter)."""
    white = rng.standard_normal(shape)
    kx = np.fft.fftfreq(shape[0])[:, None]
    ky = np.fft.rfftfreq(shape[1])[None, :]
    filt = np.exp(-0.5 * (2 * np.pi * corr_cells) ** 2 * (kx ** 2 + ky ** 2))
    f = np.fft.irfft2(np.fft.rfft2(white) * filt, s=shape)
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def dome(n_cells: int = 31, nz: int = 10, dx: float = 2000.0, radius: float = 30.0e3,
         h0: float = 2000.0, flwa: float = 1.0e-16) -> Problem:
    """Parabolic dome, H(r) = h0 sqrt(1 - (r/R)^2) on a flat bed, no-slip base."""
    g = Grid.from_owned(n_cells, n_cells, nz, global_bc=BC_NO_ICE)
    xc = (np.arange(1, g.nx + 1) - 2.5) * dx
    yc = (np.arange(1, g.ny + 1) - 2.5) * dx
    xm = 0.5 * n_cells * dx
    r2 = ((xc[:, None] - xm) ** 2 + (yc[None, :] - xm) ** 2) / radius ** 2
    thck = np.where(r2 < 1.0, h0 * np.sqrt(np.clip(1.0 - r2, 0.0, None)), 0.0)
    p = Problem(
        grid=g, dx=dx, dy=dx, sigma=glide_sigma(nz),
        thck=np.asfortranarray(thck), topg=np.zeros((g.nx, g.ny), order="F"),
        flwa=np.full((nz - 1, g.nx, g.ny), flwa, order="F"),
        beta=np.zeros((g.nx - 1, g.ny - 1), order="F"),
        uvel=g.zeros_vec(), vvel=g.zeros_vec(),
        whichbabc=HO_BABC_NO_SLIP,
        meta=dict(workload="dome", n_cells=n_cells, nz=nz, dx=dx, radius=radius, h0=h0, flwa=flwa),
    )
    return p


def ice_stream(n_vert: int, nz: int = 11, dx: float = 1000.0, seed: int = 0,
               flwa: float = 1.0e-16, h_mean: float = 1000.0, h_amp: float = 200.0,
               wavelength: float = 20.0e3, bed_slope: float = 0.002, bump_sigma: float = 50.0,
               bump_corr: float = 5.0e3, beta_median: float = 1.0e3, beta_sigma_log: float = 1.0,
               beta_corr: float = 5.0e3, u_surf: float = 150.0, bed_amp: float = 0.0,
               bed_wavelength: float = 15.0e3) -> Problem:
    """Synthetic ice stream on an ``n_vert x n_vert`` owned vertex grid.

    Thickness ``h_mean + h_amp sin(2 pi x / wavelength)``; bed sloping down in
    +x at ``bed_slope`` plus seeded Gaussian bumps (std ``bump_sigma``,
    correlation length ``bump_corr``) on land; log-normal beta (median
    ``beta_median``, sigma_log ``beta_sigma_log``); the outermost owned cell
    ring and the halos are ice-free so that every halo value is zero.  The
    velocity guess is a plug-plus-shear profile that sets the viscosity of
    the single assembled system.
    """
    rng = np.random.default_rng(seed)
    g = Grid.from_owned(n_vert, n_vert, nz, global_bc=BC_NO_ICE)
    nx, ny = g.nx, g.ny
    xc = (np.arange(1, nx + 1) - 2.5) * dx
    yc = (np.arange(1, ny + 1) - 2.5) * dx
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    thck = h_mean + h_amp * np.sin(2 * np.pi * X / wavelength)
    ring = np.zeros((nx, ny), bool)
    ring[:3, :] = ring[-3:, :] = True
    ring[:, :3] = ring[:, -3:] = True
    thck[ring] = 0.0
    bumps = bump_sigma * _smooth_field(rng, (nx, ny), bump_corr / dx)
    topg = 2000.0 - bed_slope * X + bumps + bed_amp * np.sin(2 * np.pi * (X + 0.5 * Y) / bed_wavelength)
    lb = _smooth_field(rng, (nx - 1, ny - 1), beta_corr / dx)
    beta = beta_median * np.exp(beta_sigma_log * lb)
    # velocity guess: basal sliding + internal shear, varying across the stream
    sig = glide_sigma(nz)
    xv = (np.arange(1, nx) - 2.0) * dx
    yv = (np.arange(1, ny) - 2.0) * dx
    XV, YV = np.meshgrid(xv, yv, indexing="ij")
    Ly = max(yv[-1] - yv[0], dx)
    ub = 0.5 * u_surf * (1.0 + 0.5 * np.sin(np.pi * (YV - yv[0]) / Ly)) * (1.0 + 0.1 * np.cos(2 * np.pi * XV / wavelength))
    us = 0.5 * u_surf * (1.0 + 0.3 * np.cos(2 * np.pi * XV / (1.7 * wavelength)))
    shape = 1.0 - sig[:, None, None] ** 4          # 1 at the surface, 0 at the bed
    uvel = ub[None] + us[None] * shape
    vvel = 0.1 * uvel * np.sin(2 * np.pi * (YV[None] + 0.3 * XV[None]) / (2.3 * wavelength))
    p = Problem(
        grid=g, dx=dx, dy=dx, sigma=sig,
        thck=np.asfortranarray(thck), topg=np.asfortranarray(topg),
        flwa=np.full((nz - 1, nx, ny), flwa, order="F"),
        beta=np.asfortranarray(beta),
        uvel=np.asfortranarray(uvel), vvel=np.asfortranarray(vvel),
        whichbabc=HO_BABC_BETA_EXTERNAL,
        meta=dict(workload="ice_stream", n_vert=n_vert, nz=nz, dx=dx, seed=seed, flwa=flwa,
                  h_mean=h_mean, h_amp=h_amp, wavelength=wavelength, bed_slope=bed_slope,
                  bump_sigma=bump_sigma, bump_corr=bump_corr, beta_median=beta_median,
                  beta_sigma_log=beta_sigma_log, beta_corr=beta_corr),
    )
    return p


def slab(n_cells: int = 8, nz: int = 21, dx: float = 1000.0, thck: float = 1000.0,
         theta_deg: float = 0.5, beta: float = 2000.0, flwa: float = 1.0e-16,
         global_bc: int = BC_PERIODIC) -> Problem:
    """Doubly periodic inclined slab (tests/slab): uniform thickness on a bed
    tilted by ``theta`` in x, linear sliding with uniform beta.  Periodic BCs
    wrap the vertex grid, so the surface is shifted by the slab drop across
    one period through the cell elevations only (the solver sees the slope
    through the surface gradient and the node z coordinates).  global_bc 2
    (no-penetration) keeps the periodic halos and holds u = 0 on the
    east/west boundary column and v = 0 on the north/south boundary row."""
    if global_bc not in (0, 2):
        raise ValueError("slab: global_bc 0 (periodic) or 2 (no-penetration)")
    g = Grid.from_owned(n_cells, n_cells, nz, global_bc=global_bc)
    xc = (np.arange(1, g.nx + 1) - 2.5) * dx
    slope = np.tan(np.deg2rad(theta_deg))
    topg = np.asfortranarray(np.repeat((-slope * xc)[:, None], g.ny, axis=1) + 5000.0)
    p = Problem(
        grid=g, dx=dx, dy=dx, sigma=np.linspace(0.0, 1.0, nz),
        thck=np.full((g.nx, g.ny), thck, order="F"), topg=topg,
        flwa=np.full((nz - 1, g.nx, g.ny), flwa, order="F"),
        beta=np.full((g.nx - 1, g.ny - 1), beta, order="F"),
        uvel=g.zeros_vec(), vvel=g.zeros_vec(), whichbabc=HO_BABC_BETA_EXTERNAL,
        gradient_margin=0,
        meta=dict(workload="slab", theta_deg=theta_deg, thck=thck, beta=beta, flwa=flwa),
    )
    return p


def ensemble(n_members: int = 256, seed: int = 2024, nmin: int = 32, nmax: int = 192, nz_choices=(5, 7, 11, 15),
             dx: float = 1000.0):
    """Ensemble of independent ice-stream members with uneven sizes: side
    n ~ U{nmin..nmax} owned vertices, nz from ``nz_choices``, Glen A
    log-uniform in [1e-17, 1e-15] Pa^-3 yr^-1, log-normal beta and a
    sinusoid-plus-noise bed, every member seeded from ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        n = int(rng.integers(nmin, nmax + 1))
        nz = int(rng.choice(nz_choices))
        A = float(10.0 ** rng.uniform(-17.0, -15.0))
        p = ice_stream(n, nz=nz, dx=dx, seed=int(rng.integers(0, 2**31 - 1)), flwa=A,
                       bed_amp=float(rng.uniform(20.0, 80.0)),
                       beta_median=float(10.0 ** rng.uniform(2.5, 3.5)))
        p.meta.update(member=m, glen_A=A)
        out.append(p)
    return out


def marine_sheet(n_vert: int, nz: int = 11, dx: float = 1000.0, seed: int = 0, flwa: float = 1.0e-16,
                 h_max: float = 3000.0, h_shelf: float = 300.0, bed_front: float = -500.0,
                 bed_interior: float = -900.0, bump_sigma: float = 30.0, bump_corr: float = 8.0e3,
                 beta_median: float = 1.0e3, beta_sigma_log: float = 0.5, beta_corr: float = 8.0e3,
                 gl_frac: float = 0.55, front_frac: float = 0.85, gl_amp_frac: float = 0.05,
                 margin_frac: float = 0.25, u_front: float = 800.0) -> Problem:
    """Marine ice sheet on an ``n_vert x n_vert`` owned vertex grid.

    Ice flows east: a grounded interior (thickness up to ``h_max``) on a
    retrograde bed (``bed_interior`` in the west rising to ``bed_front`` in
    the east, plus seeded bumps) thins past a seeded, wavy grounding line
    (mean position ``gl_frac`` of the width) into a floating shelf that ends
    at a calving front (``front_frac``) facing open ocean.  A land strip
    borders the west, north and south edges; toward it the bed rises to sea
    level and the ice thins as sqrt(distance / (margin_frac * width)), so
    margins are grounded and every halo value is zero.  beta is log-normal;
    under floating ice it is removed by f_ground.
    """
    rng = np.random.default_rng(seed)
    g = Grid.from_owned(n_vert, n_vert, nz, global_bc=BC_NO_ICE)
    nx, ny = g.nx, g.ny
    xc = (np.arange(1, nx + 1) - 2.5) * dx
    yc = (np.arange(1, ny + 1) - 2.5) * dx
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    L = (n_vert - 1) * dx
    strip = 4
    x_w = (strip - 1.5) * dx                              # west edge of the ice
    y_s, y_n = (strip - 1.5) * dx, (ny - strip - 1.5) * dx
    d_land = np.maximum(np.minimum(np.minimum(X - x_w, Y - y_s), y_n - Y), 0.0)
    bed = bed_interior + (bed_front - bed_interior) * np.clip(X / L, 0.0, 1.0)
    bed = bed + bump_sigma * _smooth_field(rng, (nx, ny), bump_corr / dx)
    wb = 0.1 * L
    bed = bed + (20.0 - bed) * (1.0 - np.clip(d_land / wb, 0.0, 1.0)) ** 2
    # grounding line position varies along y with a seeded phase
    phase = rng.uniform(0, 2 * np.pi)
    x_gl = L * (gl_frac + gl_amp_frac * np.sin(2 * np.pi * Y / max(L, dx) * 1.5 + phase))
    x_front = front_frac * L
    rr = RHOO_SYN / RHOI_SYN
    h_float = np.maximum(-bed * rr, 0.0)                 # flotation thickness
    s = np.clip(X / x_gl, 0.0, 1.0)
    h_sheet = h_float * 1.15 + (h_max - h_float * 1.15) * np.sqrt(np.clip(1.0 - s ** 2, 0.0, 1.0))
    t = np.clip((X - x_gl) / np.maximum(x_front - x_gl, dx), 0.0, 1.0)
    h_sh = 0.85 * h_float * (1.0 - t) + h_shelf * t
    thck = np.where(X <= x_gl, h_sheet, np.minimum(h_sh, 0.95 * h_float))
    thck = thck * np.sqrt(np.clip(d_land / (margin_frac * L), 0.0, 1.0))
    thck = np.where(X > x_front, 0.0, thck)
    # land strip (west, north, south): bed above sea level, no ice
    land = np.zeros((nx, ny), bool)
    land[:strip, :] = True
    land[:, :strip] = True
    land[:, -strip:] = True
    bed = np.where(land, 50.0, bed)
    thck = np.where(land, 0.0, thck)
    thck[-3:, :] = 0.0
    lb = _smooth_field(rng, (nx - 1, ny - 1), beta_corr / dx)
    beta = beta_median * np.exp(beta_sigma_log * lb)
    sig = glide_sigma(nz)
    xv = (np.arange(1, nx) - 2.0) * dx
    yv = (np.arange(1, ny) - 2.0) * dx
    XV, YV = np.meshgrid(xv, yv, indexing="ij")
    ub = u_front * np.clip(XV / L, 0.0, 1.0) ** 2
    uvel = ub[None] * (0.8 + 0.2 * (1.0 - sig[:, None, None] ** 4))
    vvel = 0.05 * uvel * np.sin(2 * np.pi * YV[None] / max(L, dx))
    p = Problem(
        grid=g, dx=dx, dy=dx, sigma=sig,
        thck=np.asfortranarray(thck), topg=np.asfortranarray(bed),
        flwa=np.full((nz - 1, nx, ny), flwa, order="F"),
        beta=np.asfortranarray(beta),
        uvel=np.asfortranarray(uvel), vvel=np.asfortranarray(vvel),
        whichbabc=HO_BABC_BETA_EXTERNAL,
        meta=dict(workload="marine_sheet", n_vert=n_vert, nz=nz, dx=dx, seed=seed, flwa=flwa, h_max=h_max,
                  h_shelf=h_shelf, bed_front=bed_front, bed_interior=bed_interior, gl_frac=gl_frac,
                  front_frac=front_frac, margin_frac=margin_frac),
    )
    return p


# The reference's ice-stream test (tests/stream/runStream.py): a uniform slab
# (H = 1000 m, bed slope 1e-3 along x) over till whose yield stress varies
# across the flow only, which_ho_babc = 15 (beta = tauf / |u|), doubly
# periodic, with one across-flow column of prescribed velocity.
STREAM_HALF_WIDTH = 25000.0
STREAM_H = 1000.0
STREAM_DSDX = -1.0e-3
STREAM_A = 1.0e-16
STREAM_SCHOOF_M = 1.55


def stream_driving_stress():
    return RHOI_SYN * 9.81 * STREAM_H * (-STREAM_DSDX)


def stream_yield_stress(kind: str, y: np.ndarray) -> np.ndarray:
    """Till yield stress across the stream: Raymond (2000), 5.2 kPa inside
    and 70 kPa outside; Schoof (2006), tau_d |y / L|^m with
    L = W / (m + 1)^(1/m) (runStream.py:150-175)."""
    W = STREAM_HALF_WIDTH
    if kind == "raymond":
        t = np.full(np.shape(y), 5.2e3)
        t[np.abs(y) >= W] = 0.7e5
        return t
    m = STREAM_SCHOOF_M
    L = W / (m + 1.0) ** (1.0 / m)
    return stream_driving_stress() * np.abs(y / L) ** m


def stream_velocity(kind: str, y: np.ndarray) -> np.ndarray:
    """Analytic along-flow speed across the stream (m/yr): Raymond (2000)
    u = 2A/(n+1) ((tau_d - tau_0)/H)^n (W^(n+1) - |y|^(n+1)); Schoof (2006)
    eq. 4.3 with n = 3 (runStream.py:160-190)."""
    W, H, A, n = STREAM_HALF_WIDTH, STREAM_H, STREAM_A, 3.0
    taud = stream_driving_stress()
    y = np.asarray(y, dtype=np.float64)
    if kind == "raymond":
        t0 = np.minimum(stream_yield_stress(kind, y), taud)
        u = 2.0 * A / (n + 1.0) * ((taud - t0) / H) ** n * (W ** (n + 1) - np.abs(y) ** (n + 1))
        return np.maximum(u, 0.0)
    m = STREAM_SCHOOF_M
    L = W / (m + 1.0) ** (1.0 / m)
    B = A ** (-1.0 / n)
    yl = np.abs(y / L)
    u = -2.0 * taud ** 3 * L ** 4 / (B ** 3 * H ** 3) * (
        ((y / L) ** 4 - (m + 1.0) ** (4.0 / m)) / 4.0
        - 3.0 * (yl ** (m + 4.0) - (m + 1.0) ** (1.0 + 4.0 / m)) / ((m + 1.0) * (m + 4.0))
        + 3.0 * (yl ** (2.0 * m + 4.0) - (m + 1.0) ** (2.0 + 4.0 / m)) / ((m + 1.0) ** 2 * (2.0 * m + 4.0))
        - (yl ** (3.0 * m + 4.0) - (m + 1.0) ** (3.0 + 4.0 / m)) / ((m + 1.0) ** 3 * (3.0 * m + 4.0)))
    u[np.abs(y) >= W] = 0.0
    return u


def yield_stress_stream(kind: str = "raymond", n_stream: int = 25, nz: int = 2, along: float = 30000.0,
                        kin_column: int = 2) -> Problem:
    """The ice-stream test of the reference (tests/stream/runStream.py,
    stream.config): n_stream cells across the stream (dx = dy = 2W/n_stream),
    strips of 5H (Raymond) or 3L - W (Schoof) of strong till on each side,
    ``along`` metres along the flow, doubly periodic; the bed continues its
    slope through the halos (the periodic_offset_ew of the reference).  The
    yield stress is the problem's friction (which_ho_babc 15,
    beta_grounded_min 0); owned vertex column ``kin_column`` carries the
    analytic speed as a kinematic (Dirichlet) BC with v = 0, the rest of the
    velocity starts at zero.  Vertex rows sit at y = (jo - (ny-1)/2 + 1/2) dy
    so that the stream edges |y| = W fall on vertices; meta["y_vertex"] holds
    every vertex row's y (wrapped into one period)."""
    from .basal import Friction
    if kind not in ("raymond", "schoof"):
        raise ValueError("kind must be 'raymond' or 'schoof'")
    W, H = STREAM_HALF_WIDTH, STREAM_H
    dy = 2.0 * W / n_stream
    dx = dy
    if kind == "raymond":
        strong = 5.0 * H
    else:
        strong = 3.0 * (W / (STREAM_SCHOOF_M + 1.0) ** (1.0 / STREAM_SCHOOF_M)) - W
    ny_own = n_stream + 2 * int(round(strong / dy + 1.0e-10))
    nx_own = int(round(along / dx))
    g = Grid.from_owned(nx_own, ny_own, nz, global_bc=BC_PERIODIC)
    period = ny_own * dy
    yv = (np.arange(1, g.ny) - g.jlo) * dy - dy * (ny_own - 1) / 2.0 + dy / 2.0
    yv = (yv + period / 2.0) % period - period / 2.0
    xc = (np.arange(1, g.nx + 1) - 2.5) * dx
    topg = np.asfortranarray(np.repeat((1000.0 + STREAM_DSDX * xc)[:, None], g.ny, axis=1))
    tauf = np.asfortranarray(np.repeat(stream_yield_stress(kind, yv)[None, :], g.nvx, axis=0))
    uvel, vvel = g.zeros_vec(), g.zeros_vec()
    kin = np.zeros((g.nvx, g.nvy), np.int32, order="F")
    ic = g.ilo - 1 + kin_column
    kin[ic, :] = 1
    uvel[:, ic, :] = stream_velocity(kind, yv)[None, :]
    p = Problem(
        grid=g, dx=dx, dy=dy, sigma=glide_sigma(nz),
        thck=np.full((g.nx, g.ny), H, order="F"), topg=topg,
        flwa=np.full((nz - 1, g.nx, g.ny), STREAM_A, order="F"),
        beta=np.zeros((g.nvx, g.nvy), order="F"),
        uvel=np.asfortranarray(uvel), vvel=np.asfortranarray(vvel), whichbabc=HO_BABC_BETA_EXTERNAL,
        gradient_margin=0, kinbcmask=kin,
        meta=dict(workload="yield_stress_stream", kind=kind, n_stream=n_stream, nz=nz, y_vertex=yv),
    )
    p.friction = Friction(15, mintauf=tauf, beta_grounded_min=0.0)
    return p


def ismip_hom(experiment: str = "c", size_km: float = 80.0, n: int = 40, nz: int = 11,
              flwa: float = 1.0e-16) -> Problem:
    """ISMIP-HOM experiment A or C (Pattyn et al. 2008) as the reference sets
    it up (tests/ismip-hom/runISMIP_HOM.py, ismip-hom.config): n x n cells
    of size_km / n, doubly periodic, the surface sloping along x (0.5 deg
    for A, 0.1 deg for C) and continuing through the halos (the reference's
    periodic_offset_ew).  A: bed bumps, thickness 1000 - 500 sin(wx) sin(wy)
    under the surface 4000 - x tan(a), frozen bed (which_ho_babc 4);
    C: thickness 1000 on the bed 1000 - x tan(a), basal friction
    1000 + 1000 sin(wx) sin(wy) built on the velocity grid by calcbeta
    (which_ho_babc 8).  beta_grounded_min 0, velocity guess zero."""
    from .basal import Friction
    if experiment not in ("a", "c"):
        raise ValueError("experiment must be 'a' or 'c'")
    size = size_km * 1000.0
    dx = size / n
    g = Grid.from_owned(n, n, nz, global_bc=BC_PERIODIC)
    x = (np.arange(1, g.nx + 1) - 3 + 0.5) * dx          # cell centres, owned cell 3 at dx / 2
    y = (np.arange(1, g.ny + 1) - 3 + 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    w = 2.0 * np.pi / size
    if experiment == "a":
        alpha = 0.5 * np.pi / 180.0
        zz = 4000.0 - X * np.tan(alpha)
        thck = 1000.0 - 500.0 * np.sin(w * X) * np.sin(w * Y)
        topg = zz - thck
        whichbabc = HO_BABC_BETA_LARGE
    else:
        alpha = 0.1 * np.pi / 180.0
        thck = np.full((g.nx, g.ny), 1000.0)
        topg = 1000.0 - X * np.tan(alpha)
        whichbabc = HO_BABC_BETA_EXTERNAL
    p = Problem(
        grid=g, dx=dx, dy=dx, sigma=glide_sigma(nz),
        thck=np.asfortranarray(thck), topg=np.asfortranarray(topg),
        flwa=np.full((nz - 1, g.nx, g.ny), flwa, order="F"),
        beta=np.zeros((g.nvx, g.nvy), order="F"),
        uvel=g.zeros_vec(), vvel=g.zeros_vec(), whichbabc=whichbabc,
        meta=dict(workload="ismip_hom", experiment=experiment, size_km=size_km, n=n, nz=nz),
    )
    if experiment == "c":
        p.friction = Friction(8, beta_grounded_min=0.0)
    return p


def depth_integrated_system(Auu, Auv, Avu, Avv, bu, bv):
    """2-D (9-point) system P^T A P of an assembled 3-D system, with P x(i,j) =
    x on every level of column (i,j): block stencil entry (i,j; m2) = the sum
    over k and kA of A(m2 + 9(kA+1), k, i, j); b2(i,j) = sum_k b(k,i,j).  It
    is symmetric positive definite whenever A is, and carries the horizontal
    coupling and conditioning of the assembled ice-flow problem, so it serves
    as a synthetic depth-integrated (SSA-like) workload for the 2-D solvers.
    Inputs in the reference layouts (27, nz, nx-1, ny-1) and (nz, nx-1, ny-1);
    returns four (nx-1, ny-1, 9) stencils and two (nx-1, ny-1) vectors."""
    out = []
    for a in (Auu, Auv, Avu, Avv):
        s = a.sum(axis=1)
        s = s[0:9] + s[9:18] + s[18:27]
        out.append(np.asfortranarray(np.moveaxis(s, 0, 2)))
    return out, np.asfortranarray(bu.sum(axis=0)), np.asfortranarray(bv.sum(axis=0))
