"""Basal traction coefficient on the device: glissade_calcbeta
(glissade_basal_traction.F90:69-719).

:class:`Friction` holds what glissade_calcbeta reads from the basal physics
(glide_basal_physics, glide_types.F90:2190-2300; defaults are the
reference's) for one which_ho_babc option.  A solver whose problem carries a
``friction`` recomputes beta from the basal velocity before every assembly
(glissade_velo_higher.F90:2316-2332) with :func:`device_calcbeta`, and the
assembly reads that device field; the f_ground weighting with
which_ho_beta_limit (:650-685), the beta >= 0 check (:687-699) and the
staggered halo (:703) are part of the call.  A negative or NaN beta raises
:class:`~cism_amd._lib.CismGpuError` with status ``CISMGPU_ERR_BETA``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

# which_ho_babc (glide_types.F90:253-268)
HO_BABC_BETA_CONSTANT = 0
HO_BABC_BETA_BPMP = 1
HO_BABC_PSEUDO_PLASTIC = 2
HO_BABC_PSEUDO_PLASTIC_OLD = 3
HO_BABC_BETA_LARGE = 4
HO_BABC_BETA_EXTERNAL = 5
HO_BABC_NO_SLIP = 6
HO_BABC_ZOET_IVERSON = 7
HO_BABC_ISHOMC = 8            # ISMIP-HOM experiment C's prescribed field
HO_BABC_POWERLAW = 9
HO_BABC_COULOMB_FRICTION = 10
HO_BABC_SCHOOF = 11
HO_BABC_MODIFIED_SCHOOF = 12
HO_BABC_TSAI = 13
HO_BABC_POWERLAW_EFFECPRESS = 14
HO_BABC_YIELD_PICARD = 15
DEVICE_OPTIONS = (0, 1, 2, 3, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15)

CISMGPU_ERR_BETA = 6

# scalar parameters, in the order of cismgpu_calcbeta_in's doubles
PARAMS = ("ho_beta_const", "ho_beta_small", "ho_beta_large", "pseudo_plastic_q", "pseudo_plastic_u0",
          "pseudo_plastic_phimin", "pseudo_plastic_phimax", "pseudo_plastic_bedmin", "pseudo_plastic_bedmax",
          "eus", "powerlaw_m", "zoet_iverson_ut", "zoet_iverson_nmax", "coulomb_bump_max_slope",
          "coulomb_bump_wavelength", "friction_powerlaw_k", "beta_grounded_min", "beta_powerlaw_umax",
          "dew", "dns")
VERTEX_FIELDS = ("effecpress_stag", "coulomb_c", "powerlaw_c", "mintauf", "c_space_factor_stag")


@dataclass
class Friction:
    """Sliding law and basal physics of glissade_calcbeta.  Vertex fields are
    (nx-1, ny-1); ``flwa_basal`` (nx, ny) defaults to the basal layer of the
    problem's rate factor, as the reference passes flwa(nz-1,:,:) (:2323)."""
    whichbabc: int
    which_beta_limit: int = 0                 # which_ho_beta_limit (glide_types.F90:270-271)
    ho_beta_const: float = 1000.0             # glide_types.F90:2200-2202
    ho_beta_small: float = 1000.0
    ho_beta_large: float = 1.0e10
    pseudo_plastic_q: float = 0.5             # :2226-2236
    pseudo_plastic_u0: float = 100.0
    pseudo_plastic_phimin: float = 5.0
    pseudo_plastic_phimax: float = 40.0
    pseudo_plastic_bedmin: float = -300.0
    pseudo_plastic_bedmax: float = 700.0
    powerlaw_m: float = 3.0                   # :2255
    zoet_iverson_ut: float = 200.0            # :2217-2218
    zoet_iverson_nmax: float = 1.0e8
    coulomb_bump_max_slope: float = 0.5       # :2281-2282
    coulomb_bump_wavelength: float = 2.0
    friction_powerlaw_k: float = 8.4e-9       # :2239
    beta_grounded_min: float = 1.0            # :2198
    beta_powerlaw_umax: float = 0.0           # :2287
    effecpress_stag: np.ndarray | None = None
    coulomb_c: np.ndarray | None = None
    powerlaw_c: np.ndarray | None = None
    mintauf: np.ndarray | None = None
    c_space_factor_stag: np.ndarray | None = None
    bpmp_mask: np.ndarray | None = None
    beta_external: np.ndarray | None = None
    flwa_basal: np.ndarray | None = None

    def params(self, eus: float, dew: float = 0.0, dns: float = 0.0) -> np.ndarray:
        """The scalar parameters in the C-ABI order (eus and the grid spacing from the problem)."""
        grid = {"eus": eus, "dew": dew, "dns": dns}
        return np.array([grid[k] if k in grid else getattr(self, k) for k in PARAMS], dtype=np.float64)


_P = ctypes.c_void_p


class CalcbetaIn(ctypes.Structure):
    _fields_ = ([("whichbabc", ctypes.c_int32), ("which_beta_limit", ctypes.c_int32)] +
                [(k, ctypes.c_double) for k in PARAMS] +
                [("f_ground", _P)] + [(k, _P) for k in VERTEX_FIELDS] +
                [("bpmp_mask", _P), ("beta_external", _P), ("topg", _P), ("thck", _P), ("flwa_basal", _P)])


def _declare(L):
    if getattr(L, "_calcbeta_declared", False):
        return
    L.cismgpu_system_calcbeta_setup.argtypes = [ctypes.c_void_p, ctypes.POINTER(CalcbetaIn)]
    L.cismgpu_system_calcbeta_setup.restype = ctypes.c_int32
    D = ctypes.POINTER(ctypes.c_double)
    L.cismgpu_system_calcbeta.argtypes = [ctypes.c_void_p, D, D, D]
    L.cismgpu_system_calcbeta.restype = ctypes.c_int32
    L._calcbeta_declared = True


def _f64(a):
    return None if a is None else np.asfortranarray(np.array(a, dtype=np.float64))


def inputs(friction: Friction, *, f_ground, topg, thck, eus, flwa_basal, dew=0.0, dns=0.0):
    """(CalcbetaIn, arrays kept alive for it) for one friction and geometry
    (dew, dns: the grid spacing, read by ISMIP-HOM C)."""
    keep = dict(f_ground=_f64(f_ground), topg=_f64(topg), thck=_f64(thck),
                flwa_basal=_f64(friction.flwa_basal if friction.flwa_basal is not None else flwa_basal))
    for k in VERTEX_FIELDS + ("beta_external",):
        keep[k] = _f64(getattr(friction, k))
    keep["bpmp_mask"] = (None if friction.bpmp_mask is None
                         else np.asfortranarray(np.array(friction.bpmp_mask, dtype=np.int32)))
    ptr = lambda a: None if a is None else a.ctypes.data
    cin = CalcbetaIn(friction.whichbabc, friction.which_beta_limit,
                     *[float(x) for x in friction.params(eus, dew, dns)],
                     ptr(keep["f_ground"]), *[ptr(keep[k]) for k in VERTEX_FIELDS],
                     ptr(keep["bpmp_mask"]), ptr(keep["beta_external"]), ptr(keep["topg"]), ptr(keep["thck"]),
                     ptr(keep["flwa_basal"]))
    return cin, keep


def device_setup(system, friction: Friction, **geometry):
    """Upload the basal physics of ``friction`` to ``system`` (cismgpu_system_calcbeta_setup);
    geometry: f_ground, topg, thck, eus, flwa_basal, dew, dns (see :func:`inputs`)."""
    L = _lib.load()
    _declare(L)
    cin, keep = inputs(friction, **geometry)
    _lib.check(L.cismgpu_system_calcbeta_setup(system.handle, ctypes.byref(cin)), "calcbeta_setup")
    del keep


def device_calcbeta(system, ubas=None, vbas=None, want: bool = False):
    """beta from the basal velocity (the resident x when ubas/vbas are None)
    into the system's device field; ``want``: also return it (nx-1, ny-1)."""
    L = _lib.load()
    _declare(L)
    D = ctypes.POINTER(ctypes.c_double)
    g = system.grid
    u = v = None
    if ubas is not None:
        u, v = _f64(ubas), _f64(vbas)
        if u.shape != (g.nvx, g.nvy) or v.shape != (g.nvx, g.nvy):
            raise ValueError("ubas/vbas must be (nx-1, ny-1)")
    out = np.zeros((g.nvx, g.nvy), order="F") if want else None
    _lib.check(L.cismgpu_system_calcbeta(system.handle, u.ctypes.data_as(D) if u is not None else None,
                                         v.ctypes.data_as(D) if v is not None else None,
                                         out.ctypes.data_as(D) if out is not None else None), "calcbeta")
    return out
