"""On-disk interchange with CISM: netCDF model states and the dense matrix dump.

* ``read_cism_state`` builds a :class:`~cism_amd.synthetic.Problem` from a
  CISM input or restart file in the variable layout of glide_vars.def
  (``thk``/``topg`` on ``(time, y1, x1)`` :460-465, :667-673; ``beta`` on
  ``(time, y0, x0)`` :193-199; ``uvel``/``vvel`` on ``(time, level, y0, x0)``
  :1095-1111; ``flwa`` and ``efvs`` on the layer grid :1375-1400).  The file
  holds the global domain (``ewn x nsn`` cells); one task adds two halo cells
  on each side (parallel_mpi.F90:2590-2604) and fills them by the halo rules
  of the global BC, as CISM does after reading.
* ``write_cism_state`` / ``write_cism_velocity`` write the same layout
  (NETCDF3 classic), so states made here can be read by CISM and results
  handed back.
* ``write_matrix_elements_3d`` / ``read_matrix_elements_3d`` write and read
  the dense text dump of glissade_velo_higher.F90:11258-11400
  (``Auu.<label>`` ... ``bv.<label>``: one row per active node, numbered as
  get_vertex_geometry does, :4200-4216), and ``dense_to_structured`` maps a
  dump back to the 27-point layout.

netCDF files are read and written with scipy's NETCDF3 reader (classic and
64-bit-offset formats); a NETCDF4/HDF5 file raises with that message.
"""
from __future__ import annotations

import os

import numpy as np

from . import geometry as ge
from .grid import BC_NO_ICE, NHALO, Grid, glide_sigma, indxA_3d
from .synthetic import Problem

SCYR = 31536000.0
DEFAULT_FLWA = 1.0e-16     # Pa^-3 yr^-1, glide_types.F90 default_flwa


def _open(path, mode="r"):
    from scipy.io import netcdf_file
    if mode == "r":
        with open(path, "rb") as f:
            magic = f.read(4)
        if magic[:3] != b"CDF":
            raise ValueError(f"{path}: not a NETCDF3 (classic / 64-bit offset) file; convert NETCDF4/HDF5 "
                             "files with nccopy -k classic first")
        return netcdf_file(path, "r", mmap=False)
    return netcdf_file(path, "w", version=2)


def _var(f, name, t):
    """Variable `name` at time index t as float64, or None."""
    if name not in f.variables:
        return None
    v = f.variables[name]
    a = np.array(v[:], dtype=np.float64)
    if v.dimensions and v.dimensions[0] == "time":
        a = a[t]
    return a, v.dimensions[1:] if v.dimensions and v.dimensions[0] == "time" else v.dimensions


def _cells(grid: Grid, a_yx: np.ndarray) -> np.ndarray:
    """(nsn, ewn) file field -> (nx, ny) with the cell halo of the global BC."""
    out = np.zeros((grid.nx, grid.ny), order="F")
    ewn, nsn = a_yx.shape[1], a_yx.shape[0]
    out[NHALO:NHALO + ewn, NHALO:NHALO + nsn] = a_yx.T
    ge.cell_halo_2d(grid, out)
    return out


def _vertices(grid: Grid, a_yx: np.ndarray) -> np.ndarray:
    """(nsn-1, ewn-1) file field -> (nx-1, ny-1) with the staggered halo."""
    out = np.zeros((grid.nx - 1, grid.ny - 1), order="F")
    ewn1, nsn1 = a_yx.shape[1], a_yx.shape[0]
    out[NHALO:NHALO + ewn1, NHALO:NHALO + nsn1] = a_yx.T
    ge.staggered_halo_2d(grid, out)
    return out


def read_cism_state(path: str, *, nz: int | None = None, global_bc: int = BC_NO_ICE, time_index: int = -1,
                    flwa_default: float = DEFAULT_FLWA, **options) -> Problem:
    """A velocity problem from a CISM netCDF state.  Required: ``topg`` and
    ``x1``/``y1`` (the spacing); optional: ``thk`` (zero if absent), ``beta``,
    ``uvel``/``vvel`` (the initial guess), ``flwa`` on the layer grid
    (``staglevel``; else ``flwa_default``) and ``efvs``, kept in
    ``problem.meta['efvs']`` (the DIVA state starts from it,
    glissade_velo_higher.F90:1342-1352).  ``level`` gives sigma when present,
    else ``nz`` with the default spacing (glide_setup.F90:260-376).  Extra
    keyword options set Problem fields (whichbabc, n_glen, ...)."""
    with _open(path) as f:
        x1 = np.array(f.variables["x1"][:], dtype=np.float64)
        y1 = np.array(f.variables["y1"][:], dtype=np.float64)
        ewn, nsn = len(x1), len(y1)
        dx = float(x1[1] - x1[0]) if ewn > 1 else float(options.pop("dx"))
        dy = float(y1[1] - y1[0]) if nsn > 1 else float(options.pop("dy"))
        if "level" in f.variables and f.variables["level"].shape[0] > 1:
            sigma = np.array(f.variables["level"][:], dtype=np.float64)
            if nz is not None and nz != len(sigma):
                raise ValueError(f"{path}: the file has {len(sigma)} levels, nz={nz} was asked for")
            nz = len(sigma)
        else:
            if nz is None:
                raise ValueError(f"{path}: no sigma levels in the file; pass nz")
            sigma = glide_sigma(nz)
        grid = Grid.from_owned(ewn, nsn, nz, global_bc)
        thk = _var(f, "thk", time_index)
        topg = _var(f, "topg", time_index)
        if topg is None:
            raise ValueError(f"{path}: topg is required")
        topgl = _cells(grid, topg[0])
        thck = _cells(grid, thk[0]) if thk is not None else np.zeros((grid.nx, grid.ny), order="F")
        beta = np.zeros((grid.nvx, grid.nvy), order="F")
        b = _var(f, "beta", time_index)
        if b is not None:
            beta = _vertices(grid, b[0])
        uvel, vvel = grid.zeros_vec(), grid.zeros_vec()
        for name, dst in (("uvel", uvel), ("vvel", vvel)):
            a = _var(f, name, time_index)
            if a is None:
                continue
            if a[0].shape[0] != nz:
                raise ValueError(f"{path}: {name} has {a[0].shape[0]} levels, expected {nz}")
            for k in range(nz):
                dst[k] = _vertices(grid, a[0][k])
        flwa = np.full((nz - 1, grid.nx, grid.ny), flwa_default, order="F")
        a = _var(f, "flwa", time_index)
        if a is not None:
            if a[0].shape[0] != nz - 1:
                raise ValueError(f"{path}: flwa has {a[0].shape[0]} levels; the velocity solve needs the "
                                 f"{nz - 1} layer values (staglevel)")
            for k in range(nz - 1):
                flwa[k] = _cells(grid, a[0][k])
        meta = {"source": os.path.abspath(path), "ewn": ewn, "nsn": nsn}
        a = _var(f, "efvs", time_index)
        if a is not None and a[0].shape[0] == nz - 1:
            ef = np.zeros((nz - 1, grid.nx, grid.ny), order="F")
            for k in range(nz - 1):
                ef[k] = _cells(grid, a[0][k])
            meta["efvs"] = ef
    p = Problem(grid=grid, dx=dx, dy=dy, sigma=np.ascontiguousarray(sigma), thck=np.asfortranarray(thck),
                topg=np.asfortranarray(topgl), flwa=flwa, beta=beta, uvel=uvel, vvel=vvel, meta=meta)
    for k, v in options.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown Problem option {k}")
        setattr(p, k, v)
    return p


def _owned_cells(p: Problem, a: np.ndarray) -> np.ndarray:
    g = p.grid
    return np.ascontiguousarray(a[NHALO:g.nx - NHALO, NHALO:g.ny - NHALO].T)


def _owned_vertices(p: Problem, a: np.ndarray) -> np.ndarray:
    g = p.grid
    return np.ascontiguousarray(a[NHALO:g.nx - NHALO - 1, NHALO:g.ny - NHALO - 1].T)


def _write_axes(f, p: Problem, ewn: int, nsn: int):
    g = p.grid
    f.createDimension("time", None)
    f.createDimension("x1", ewn)
    f.createDimension("y1", nsn)
    f.createDimension("x0", ewn - 1)
    f.createDimension("y0", nsn - 1)
    f.createDimension("level", g.nz)
    f.createDimension("staglevel", g.nz - 1)
    x = p.dx * np.arange(ewn)
    y = p.dy * np.arange(nsn)
    for name, dim, val in (("x1", "x1", x), ("y1", "y1", y), ("x0", "x0", p.dx / 2 + x[:-1]),
                           ("y0", "y0", p.dy / 2 + y[:-1])):
        v = f.createVariable(name, "d", (dim,))
        v[:] = val
        v.units = "meter"
    v = f.createVariable("level", "d", ("level",))
    v[:] = p.sigma
    v = f.createVariable("staglevel", "d", ("staglevel",))
    v[:] = (p.sigma[:-1] + p.sigma[1:]) / 2.0


def write_cism_state(path: str, p: Problem, time: float = 0.0) -> None:
    """Write a problem's input fields in the layout read_cism_state reads
    (owned cells and vertices only, one time slice)."""
    g = p.grid
    ewn, nsn = g.nx - 2 * NHALO, g.ny - 2 * NHALO
    with _open(path, "w") as f:
        _write_axes(f, p, ewn, nsn)
        t = f.createVariable("time", "d", ("time",))
        t[0] = time
        for name, arr in (("thk", p.thck), ("topg", p.topg)):
            v = f.createVariable(name, "d", ("time", "y1", "x1"))
            v[0] = _owned_cells(p, np.asarray(arr, dtype=np.float64))
            v.units = "meter"
        v = f.createVariable("beta", "d", ("time", "y0", "x0"))
        v[0] = _owned_vertices(p, np.asarray(p.beta, dtype=np.float64))
        v.units = "Pa yr/m"
        v = f.createVariable("flwa", "d", ("time", "staglevel", "y1", "x1"))
        v[0] = np.stack([_owned_cells(p, np.asarray(p.flwa[k])) for k in range(g.nz - 1)])
        v.units = "pascal**(-n) year**(-1)"
        for name, arr in (("uvel", p.uvel), ("vvel", p.vvel)):
            v = f.createVariable(name, "d", ("time", "level", "y0", "x0"))
            v[0] = np.stack([_owned_vertices(p, np.asarray(arr[k])) for k in range(g.nz)])
            v.units = "meter/year"


def write_cism_velocity(path: str, p: Problem, uvel: np.ndarray, vvel: np.ndarray, efvs: np.ndarray | None = None,
                        time: float = 0.0) -> None:
    """Write a velocity solution (nz, nx-1, ny-1) -- and optionally the
    layer viscosities (nz-1, nx, ny) -- as uvel/vvel/efvs of a CISM output
    file, with thk and topg for reference."""
    g = p.grid
    ewn, nsn = g.nx - 2 * NHALO, g.ny - 2 * NHALO
    with _open(path, "w") as f:
        _write_axes(f, p, ewn, nsn)
        t = f.createVariable("time", "d", ("time",))
        t[0] = time
        for name, arr in (("thk", p.thck), ("topg", p.topg)):
            v = f.createVariable(name, "d", ("time", "y1", "x1"))
            v[0] = _owned_cells(p, np.asarray(arr, dtype=np.float64))
        for name, arr in (("uvel", uvel), ("vvel", vvel)):
            v = f.createVariable(name, "d", ("time", "level", "y0", "x0"))
            v[0] = np.stack([_owned_vertices(p, np.asarray(arr[k])) for k in range(g.nz)])
            v.units = "meter/year"
        if efvs is not None:
            v = f.createVariable("efvs", "d", ("time", "staglevel", "y1", "x1"))
            v[0] = np.stack([_owned_cells(p, np.asarray(efvs[k])) for k in range(g.nz - 1)])
            v.units = "Pascal * years"


# ---------------------------------------------------------------- dense matrix dump
def node_numbering(grid: Grid, active_vertex: np.ndarray):
    """nodeID (nz, nx-1, ny-1) and the (i, j, k) of each node, 1-based, in the
    order of get_vertex_geometry (glissade_velo_higher.F90:4200-4216): owned
    active vertices j-major, then i, each column k = 1..nz."""
    nodeID = np.zeros(grid.vec_shape, np.int64, order="F")
    idx = []
    n = 0
    for j in range(grid.jlo, grid.jhi + 1):
        for i in range(grid.ilo, grid.ihi + 1):
            if active_vertex[i - 1, j - 1]:
                for k in range(1, grid.nz + 1):
                    n += 1
                    nodeID[k - 1, i - 1, j - 1] = n
                    idx.append((i, j, k))
    return nodeID, np.array(idx, dtype=np.int64).reshape(-1, 3)


def structured_to_dense(grid: Grid, active_vertex, A):
    """Dense (nNodes, nNodes) matrix of one block of a 27-point matrix,
    as write_matrix_elements_3d assembles it (:11290-11330)."""
    nodeID, idx = node_numbering(grid, active_vertex)
    nn = len(idx)
    D = np.zeros((nn, nn))
    for row, (i, j, k) in enumerate(idx):
        for kA in (-1, 0, 1):
            for jA in (-1, 0, 1):
                for iA in (-1, 0, 1):
                    if 1 <= k + kA <= grid.nz and 1 <= i + iA <= grid.nx - 1 and 1 <= j + jA <= grid.ny - 1:
                        col = nodeID[k + kA - 1, i + iA - 1, j + jA - 1]
                        if col > 0:
                            D[row, col - 1] = A[indxA_3d(iA, jA, kA) - 1, k - 1, i - 1, j - 1]
    return D


def write_matrix_elements_3d(grid: Grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, label: str,
                             outdir: str = ".") -> int:
    """The dense text dump of write_matrix_elements_3d
    (glissade_velo_higher.F90:11258-11400): files Auu.<label>, Auv.<label>,
    Avu.<label>, Avv.<label> hold one line per node, its 1-based number
    ('(i6)') then the row ('(e16.8)' each); bu.<label>, bv.<label> hold
    'number value' lines.  Returns the node count."""
    nodeID, idx = node_numbering(grid, active_vertex)
    for name, A in (("Auu", Auu), ("Auv", Auv), ("Avu", Avu), ("Avv", Avv)):
        D = structured_to_dense(grid, active_vertex, A)
        with open(os.path.join(outdir, f"{name}.{label}"), "w") as f:
            for r in range(len(idx)):
                f.write(f"{r + 1:6d}" + "".join(_e16_8(x) for x in D[r]) + "  \n")
    for name, b in (("bu", bu), ("bv", bv)):
        with open(os.path.join(outdir, f"{name}.{label}"), "w") as f:
            for r, (i, j, k) in enumerate(idx):
                f.write(f"{r + 1:6d}{_e16_8(b[k - 1, i - 1, j - 1])}\n")
    return len(idx)


def _e16_8(x: float) -> str:
    """Fortran '(e16.8)': 0.dddddddd E+ee, right-justified in 16."""
    x = float(x)
    if x == 0.0:
        return f"{'0.00000000E+00':>16}"
    m, e = f"{abs(x):.7E}".split("E")
    # d.ddddddd E e  ->  0.dddddddd E (e+1), rounding to 8 significant digits
    digits = m.replace(".", "")
    exp = int(e) + 1
    s = ("-" if x < 0 else "") + "0." + digits + "E" + ("+" if exp >= 0 else "-") + f"{abs(exp):02d}"
    return f"{s:>16}"


def _ffloat(t: str) -> float:
    """A Fortran E-format number; a three-digit exponent is written without
    its 'E' (0.10000000-299)."""
    try:
        return float(t)
    except ValueError:
        k = max(t.rfind("+"), t.rfind("-"))
        return float(t[:k] + "E" + t[k:])


def read_matrix_elements_3d(label: str, indir: str = "."):
    """Read a dump written by write_matrix_elements_3d (the reference's or
    this module's): the four dense blocks and the two rhs vectors."""
    out = {}
    for name in ("Auu", "Auv", "Avu", "Avv"):
        rows = []
        with open(os.path.join(indir, f"{name}.{label}")) as f:
            for line in f:
                t = line.split()
                if t:
                    rows.append([_ffloat(v) for v in t[1:]])
        out[name] = np.array(rows)
    for name in ("bu", "bv"):
        vals = []
        with open(os.path.join(indir, f"{name}.{label}")) as f:
            for line in f:
                t = line.split()
                if t:
                    vals.append(_ffloat(t[1]))
        out[name] = np.array(vals)
    return out


def dense_to_structured(grid: Grid, active_vertex, dump):
    """27-point blocks and rhs vectors (zero off the active owned nodes) from
    a dense dump numbered as node_numbering; entries coupling nodes that are
    not 27-point neighbours must be zero and raise otherwise."""
    nodeID, idx = node_numbering(grid, active_vertex)
    nn = len(idx)
    blocks = []
    for name in ("Auu", "Auv", "Avu", "Avv"):
        D = dump[name]
        if D.shape != (nn, nn):
            raise ValueError(f"{name}: dump is {D.shape}, the grid has {nn} active owned nodes")
        A = grid.zeros_mat()
        used = np.zeros_like(D, dtype=bool)
        for row, (i, j, k) in enumerate(idx):
            for kA in (-1, 0, 1):
                for jA in (-1, 0, 1):
                    for iA in (-1, 0, 1):
                        if 1 <= k + kA <= grid.nz and 1 <= i + iA <= grid.nx - 1 and 1 <= j + jA <= grid.ny - 1:
                            col = nodeID[k + kA - 1, i + iA - 1, j + jA - 1]
                            if col > 0:
                                A[indxA_3d(iA, jA, kA) - 1, k - 1, i - 1, j - 1] = D[row, col - 1]
                                used[row, col - 1] = True
        if np.any(D[~used] != 0.0):
            raise ValueError(f"{name}: the dump couples nodes that are not 27-point neighbours")
        blocks.append(A)
    vecs = []
    for name in ("bu", "bv"):
        b = grid.zeros_vec()
        for r, (i, j, k) in enumerate(idx):
            b[k - 1, i - 1, j - 1] = dump[name][r]
        vecs.append(b)
    return blocks, vecs
