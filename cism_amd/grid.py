"""Grid description shared by the host side and the C-ABI.

Mirrors the fields of CISM's ``parallel_type`` that the velocity solver reads
(glissade_velo_higher_pcg.F90:2888-2894; parallel_mpi.F90:2590-2653) for a
single task.  Horizontal sizes ``nx, ny`` count *cells* including the two
halo cells on each side; velocities live on the ``(nx-1, ny-1)`` vertex grid;
``nz`` is the number of sigma levels.  Owned-vertex bounds are 1-based, as in
the reference.

Arrays handed to the solver use the reference layouts (Fortran order):
vectors ``(nz, nx-1, ny-1)``, matrices ``(27, nz, nx-1, ny-1)``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

# global_bc values (glide_types.F90:442)
BC_PERIODIC = 0
BC_OUTFLOW = 1
BC_NO_PENETRATION = 2
BC_NO_ICE = 3

NHALO = 2
STAGGERED_LHALO = 2
STAGGERED_UHALO = 1


class CGrid(ctypes.Structure):
    """C mirror of ``cismgpu_grid`` / ``orc_grid`` (include/cismgpu.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ilo", ctypes.c_int32), ("ihi", ctypes.c_int32),
        ("jlo", ctypes.c_int32), ("jhi", ctypes.c_int32),
        ("global_bc", ctypes.c_int32), ("reprosum", ctypes.c_int32),
    ]


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    ilo: int
    ihi: int
    jlo: int
    jhi: int
    global_bc: int = BC_NO_ICE
    reprosum: int = 0

    @classmethod
    def from_owned(cls, n_own_x: int, n_own_y: int, nz: int, global_bc: int = BC_NO_ICE,
                   reprosum: int = 0) -> "Grid":
        """Single-task grid with ``n_own_x * n_own_y`` owned cells.

        One task holds ``global_ewn`` cells plus ``lhalo = uhalo = 2`` halo
        cells (parallel_mpi.F90:2590-2604); the owned vertex range is
        ``staggered_lhalo+1 .. nx-1-staggered_uhalo``, starting one vertex
        earlier on the west/south edge for outflow BCs (:2633-2653).
        """
        nx = n_own_x + 4
        ny = n_own_y + 4
        ilo = STAGGERED_LHALO if global_bc == BC_OUTFLOW else STAGGERED_LHALO + 1
        jlo = ilo
        return cls(nx, ny, nz, ilo, nx - 1 - STAGGERED_UHALO, jlo, ny - 1 - STAGGERED_UHALO,
                   global_bc, reprosum)

    @property
    def nvx(self) -> int:
        return self.nx - 1

    @property
    def nvy(self) -> int:
        return self.ny - 1

    @property
    def vec_shape(self):
        return (self.nz, self.nx - 1, self.ny - 1)

    @property
    def mat_shape(self):
        return (27, self.nz, self.nx - 1, self.ny - 1)

    @property
    def n_owned_nodes(self) -> int:
        return (self.ihi - self.ilo + 1) * (self.jhi - self.jlo + 1) * self.nz

    def cstruct(self) -> CGrid:
        return CGrid(self.nx, self.ny, self.nz, self.ilo, self.ihi, self.jlo, self.jhi,
                     self.global_bc, self.reprosum)

    def zeros_vec(self) -> np.ndarray:
        return np.zeros(self.vec_shape, dtype=np.float64, order="F")

    def zeros_mat(self) -> np.ndarray:
        return np.zeros(self.mat_shape, dtype=np.float64, order="F")

    def owned_slice(self):
        """numpy slice (0-based) of the owned vertex block of a (nz, nvx, nvy) array."""
        return (slice(None), slice(self.ilo - 1, self.ihi), slice(self.jlo - 1, self.jhi))


def indxA_3d(iA: int, jA: int, kA: int) -> int:
    """1-based neighbour index m (glissade_velo_higher.F90:607-627)."""
    return 1 + (iA + 1) + 3 * (jA + 1) + 9 * (kA + 1)


def indxA_2d(iA: int, jA: int) -> int:
    """1-based 2-D neighbour index (glissade_velo_higher.F90:629-640)."""
    return 1 + (iA + 1) + 3 * (jA + 1)


def glide_sigma(nz: int) -> np.ndarray:
    """Default Glide sigma levels, which_sigma = 0 (glide_setup.F90:271-275, :372-376)."""
    x = np.arange(nz, dtype=np.float64) / float(nz - 1)
    return (1.0 - (x + 1.0) ** (-2.0)) / (1.0 - 2.0 ** (-2.0))


def stagwbndsigma(sigma: np.ndarray) -> np.ndarray:
    """stagsigma augmented by 0 and 1, indexed 0..nz (glide_setup.F90:350-357)."""
    nz = sigma.size
    out = np.empty(nz + 1)
    out[1:nz] = (sigma[:-1] + sigma[1:]) / 2.0
    out[0] = 0.0
    out[nz] = 1.0
    return out
