"""Pins of the CPU restatement (oracle/) against mathematics the reference
path must satisfy -- the reference ships no golden vectors for this path.

Each check is independent of the restated formula it pins: dense or sparse
linear algebra, closed-form integrals, exact rational sums, textbook PCG,
and the closed-form first-order slab solution of tests/slab/plotSlab.py.
"""
from fractions import Fraction
from functools import reduce
from math import cos, pi, sin, tan

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cism_amd import geometry as ge, synthetic
from cism_amd.grid import BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC, glide_sigma, indxA_3d
from oracle import pyoracle as po
from systems import dome_system, random_spd_system, stream_system


# ----------------------------------------------------------------- helpers
def owned_operator(g, act, A):
    """Sparse operator on owned-vertex unknowns implied by the structured
    matvec and the one-task halo rule (periodic: halo = owned image; other
    rules: halo = 0).  Unknown order: (comp, k, i, j) over owned vertices."""
    ni, nj, nz = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1, g.nz
    periodic = g.global_bc in (0, 2)

    def uid(c, k, i, j):
        return (((c * nz + k) * nj) + (j - g.jlo)) * ni + (i - g.ilo)

    rows, cols, vals = [], [], []
    for j in range(g.jlo, g.jhi + 1):
        for i in range(g.ilo, g.ihi + 1):
            if not act[i - 1, j - 1]:
                continue
            for k in range(1, nz + 1):
                for kA in (-1, 0, 1):
                    if not 1 <= k + kA <= nz:
                        continue
                    for jA in (-1, 0, 1):
                        for iA in (-1, 0, 1):
                            qi, qj = i + iA, j + jA
                            if not (g.ilo <= qi <= g.ihi and g.jlo <= qj <= g.jhi):
                                if not periodic:
                                    continue
                                qi = g.ilo + (qi - g.ilo) % ni
                                qj = g.jlo + (qj - g.jlo) % nj
                            m = indxA_3d(iA, jA, kA) - 1
                            for cr, blkpair in ((0, (0, 1)), (1, (2, 3))):
                                for cc, blk in enumerate(blkpair):
                                    v = A[blk][m, k - 1, i - 1, j - 1]
                                    if v != 0.0:
                                        rows.append(uid(cr, k - 1, i, j))
                                        cols.append(uid(cc, k + kA - 1, qi, qj))
                                        vals.append(v)
    n = 2 * nz * ni * nj
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return M, uid


def to_owned(g, u, v):
    o = g.owned_slice()
    return np.concatenate([np.transpose(u[o], (0, 2, 1)).ravel(), np.transpose(v[o], (0, 2, 1)).ravel()])


SPD = {bc: random_spd_system(bc=bc) for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW)}


# ----------------------------------------------------------------- FE tables
def test_fe_tables_identities():
    """Partition of unity and zero-sum gradients at every quadrature point
    (glissade_velo_higher.F90:398-401); the 2x2x2 rule integrates each
    basis function to 1/8 of the reference cube volume 8."""
    phi, dxr, dyr, dzr = po.fe_tables()
    assert np.allclose(phi.sum(axis=0), 1.0, atol=1e-15)
    for d in (dxr, dyr, dzr):
        assert np.allclose(d.sum(axis=0), 0.0, atol=1e-15)
    assert np.allclose(phi.sum(axis=1), 1.0, atol=1e-15)


# ----------------------------------------------------------------- tridiagonal
def test_thomas_matches_dense():
    rng = np.random.default_rng(0)
    for n in (2, 5, 11, 15):
        sb = rng.standard_normal(n)
        sp_ = rng.standard_normal(n)
        d = 4.0 + np.abs(sb) + np.abs(sp_)
        rhs = rng.standard_normal(n)
        T = np.diag(d) + np.diag(sb[1:], -1) + np.diag(sp_[:-1], 1)
        x = po.tridiag(sb, d, sp_, rhs)
        assert np.allclose(x, np.linalg.solve(T, rhs), rtol=1e-13, atol=1e-14)


# ----------------------------------------------------------------- matvec
@pytest.mark.parametrize("bc", list(SPD))
def test_matvec_matches_sparse_operator(bc):
    g, act, A, bu, bv = SPD[bc]
    M, _ = owned_operator(g, act, A)
    rng = np.random.default_rng(2)
    xu = np.asfortranarray(rng.standard_normal(g.vec_shape))
    xv = np.asfortranarray(rng.standard_normal(g.vec_shape))
    po.halo(g, xu)
    po.halo(g, xv)
    yu, yv = po.matvec(g, act, *A, xu, xv)
    y = M @ to_owned(g, xu, xv)
    assert np.allclose(to_owned(g, yu, yv), y, rtol=1e-12, atol=1e-12 * np.abs(y).max())


def test_matvec_dome_matches_sparse_operator():
    p, s = dome_system()
    g = p.grid
    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
    M, _ = owned_operator(g, s["active_vertex"], A)
    rng = np.random.default_rng(3)
    xu = np.asfortranarray(rng.standard_normal(g.vec_shape))
    xv = np.asfortranarray(rng.standard_normal(g.vec_shape))
    po.halo(g, xu)
    po.halo(g, xv)
    yu, yv = po.matvec(g, s["active_vertex"], *A, xu, xv)
    y = M @ to_owned(g, xu, xv)
    assert np.allclose(to_owned(g, yu, yv), y, rtol=1e-12, atol=1e-12 * np.abs(y).max())


# ----------------------------------------------------------------- solvers
def sia_block_inverse(g, act, A):
    """Dense-free application of the SIA preconditioner: per active column
    and component, the vertical tridiagonal part of Auu / Avv."""
    def apply(r):
        ni, nj, nz = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1, g.nz
        z = np.zeros_like(r)
        R = r.reshape(2, nz, nj, ni)
        Z = z.reshape(2, nz, nj, ni)
        for c, blk in ((0, 0), (1, 3)):
            for jo in range(nj):
                for io in range(ni):
                    i, j = g.ilo + io, g.jlo + jo
                    if not act[i - 1, j - 1]:
                        continue
                    col = A[blk][:, :, i - 1, j - 1]
                    T = np.diag(col[13]) + np.diag(col[4, 1:], -1) + np.diag(col[22, :-1], 1)
                    Z[c, :, jo, io] = np.linalg.solve(T, R[c, :, jo, io])
        return z
    return apply


def textbook_pcg(M, b, minv, iters):
    x = np.zeros_like(b)
    r = b.copy()
    z = minv(r)
    p = z.copy()
    rz = r @ z
    out = []
    for _ in range(iters):
        q = M @ p
        a = rz / (p @ q)
        x = x + a * p
        r = r - a * q
        z = minv(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        out.append(x.copy())
    return out


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_NO_ICE])
def test_iterates_equal_textbook_pcg(bc, algo):
    """Chronopoulos-Gear is algebraically standard PCG: the restated solvers'
    k-th iterates equal a textbook PCG with the same (SIA) preconditioner on
    the same operator.  Iterations 5 and 10 replace r by b - Ax, which is
    exact in exact arithmetic."""
    g, act, A, bu, bv = SPD[bc]
    M, _ = owned_operator(g, act, A)
    b = to_owned(g, bu, bv)
    ref = textbook_pcg(M, b, sia_block_inverse(g, act, A), 12)
    for k in (1, 2, 3, 5, 8, 12):
        xu, xv = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo=algo, precond=2, tolerance=0.0, maxiters=k)
        assert st == 0 and nit == k
        x = to_owned(g, xu, xv)
        assert np.linalg.norm(x - ref[k - 1]) <= 1e-10 * np.linalg.norm(ref[k - 1]), k


@pytest.mark.parametrize("precond", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("algo", ["standard", "chrongear"])
def test_converged_solution_matches_sparse_solve(algo, precond):
    for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW):
        g, act, A, bu, bv = SPD[bc]
        M, _ = owned_operator(g, act, A)
        b = to_owned(g, bu, bv)
        keep = np.asarray(M.getnnz(axis=1) > 0)
        xs = np.zeros_like(b)
        xs[keep] = spla.spsolve(M[keep][:, keep].tocsc(), b[keep])
        xu, xv = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo=algo, precond=precond, tolerance=1e-12,
                                    maxiters=3000)
        assert st == 0 and err < 1e-12
        x = to_owned(g, xu, xv)
        assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs), (bc, precond)


def test_stream_solution_matches_sparse_solve():
    p, s = stream_system()
    g = p.grid
    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
    M, _ = owned_operator(g, s["active_vertex"], A)
    b = to_owned(g, s["bu"], s["bv"])
    keep = np.asarray(M.getnnz(axis=1) > 0)
    xs = np.zeros_like(b)
    xs[keep] = spla.spsolve(M[keep][:, keep].tocsc(), b[keep])
    xu, xv = g.zeros_vec(), g.zeros_vec()
    st, err, nit = po.pcg_solve(g, s["active_vertex"], *A, s["bu"], s["bv"], xu, xv, precond=2, tolerance=1e-12,
                                maxiters=5000)
    assert st == 0
    assert np.linalg.norm(to_owned(g, xu, xv) - xs) <= 1e-9 * np.linalg.norm(xs)


def line_preconditioner(g, act, A):
    """precond 3 as an operator built here from the matrix: z = SIA^-1 r +
    the column broadcast of the line solves of the vertically summed system
    (u along i for each j, v along j for each i; inactive or zero-pivot
    points solve the identity)."""
    sia = sia_block_inverse(g, act, A)
    ni, nj, nz = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1, g.nz
    lines = []
    for c, blk, m_sub, m_sup in ((0, 0, 4, 6), (1, 3, 2, 8)):
        vs = [A[blk][[m - 1, m + 8, m + 17]].sum(axis=(0, 1)) for m in (m_sub, 5, m_sup)]
        sub, dia, sup = [v[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] for v in vs]
        a = act[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi].astype(bool)
        sub = np.where(a, sub, 0.0); sup = np.where(a, sup, 0.0); dia = np.where(a, dia, 1.0)
        bad = dia == 0.0
        dia = np.where(bad, 1.0, dia); sub = np.where(bad, 0.0, sub); sup = np.where(bad, 0.0, sup)
        if c == 0:
            lines.append([np.linalg.inv(np.diag(dia[:, jo]) + np.diag(sub[1:, jo], -1) + np.diag(sup[:-1, jo], 1))
                          for jo in range(nj)])
        else:
            lines.append([np.linalg.inv(np.diag(dia[io, :]) + np.diag(sub[io, 1:], -1) + np.diag(sup[io, :-1], 1))
                          for io in range(ni)])

    def apply(r):
        Z = sia(r).reshape(2, nz, nj, ni).copy()
        B = r.reshape(2, nz, nj, ni).sum(axis=1)
        for jo in range(nj):
            Z[0, :, jo, :] += (lines[0][jo] @ B[0, jo, :])[None, :]
        for io in range(ni):
            Z[1, :, :, io] += (lines[1][io] @ B[1, :, io])[None, :]
        return Z.ravel()
    return apply


def test_tridiagonal_preconditioner_first_step():
    """precond 3: z = SIA^-1 r + (column-broadcast of the line solves of the
    vertically summed system).  The first standard-PCG iterate is
    x1 = alpha z0 with alpha = (b,z0)/(z0,A z0), z0 = PC(b)."""
    g, act, A, bu, bv = SPD[BC_NO_ICE]
    M, uid = owned_operator(g, act, A)
    b = to_owned(g, bu, bv)
    z = line_preconditioner(g, act, A)(b)
    alpha = (b @ z) / (z @ (M @ z))
    for precond in (3, 4):
        xu, xv = g.zeros_vec(), g.zeros_vec()
        po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo="standard", precond=precond, tolerance=0.0, maxiters=1)
        assert np.linalg.norm(to_owned(g, xu, xv) - alpha * z) <= 1e-11 * np.linalg.norm(alpha * z)


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("bc", [BC_NO_ICE, BC_OUTFLOW, BC_PERIODIC])
def test_tridiagonal_preconditioner_iterates_equal_textbook_pcg(bc, algo):
    """precond 3 and 4 (one task: the same local line solves) over many
    iterations: the restated solvers' k-th iterates equal a textbook PCG
    whose preconditioner is the line-solve operator built above."""
    g, act, A, bu, bv = SPD[bc]
    M, _ = owned_operator(g, act, A)
    b = to_owned(g, bu, bv)
    ref = textbook_pcg(M, b, line_preconditioner(g, act, A), 12)
    for precond in (3, 4):
        for k in (2, 5, 8, 12):
            xu, xv = g.zeros_vec(), g.zeros_vec()
            st, err, nit = po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo=algo, precond=precond, tolerance=0.0,
                                        maxiters=k)
            assert st == 0 and nit == k
            x = to_owned(g, xu, xv)
            assert np.linalg.norm(x - ref[k - 1]) <= 1e-9 * np.linalg.norm(ref[k - 1]), (precond, k)


def test_nan_and_nonconvergence_semantics():
    g, act, A, bu, bv = SPD[BC_PERIODIC]
    Z = [np.zeros_like(a) for a in A]
    for algo in ("standard", "chrongear"):
        xu, xv = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *Z, bu, bv, xu, xv, algo=algo, precond=0, tolerance=1e-8, maxiters=10)
        assert st == 1      # alpha = NaN is fatal in the reference
        xu, xv = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo=algo, precond=2, tolerance=1e-30, maxiters=7)
        assert st == 0 and nit == 7 and err > 0   # non-convergence is silent
        xu, xv = g.zeros_vec(), g.zeros_vec()
        st, err, nit = po.pcg_solve(g, act, *A, bu, bv, xu, xv, algo=algo, precond=2, tolerance=1e-30, maxiters=3)
        assert st == 0 and nit == 3 and np.isnan(err)   # no check ran: err unassigned


# ----------------------------------------------------------------- parallel layer
def test_global_sums_exact_and_ordered():
    rng = np.random.default_rng(5)
    g, act, A, bu, bv = SPD[BC_NO_ICE]
    a1 = np.asfortranarray(rng.standard_normal(g.vec_shape) * 10.0 ** rng.integers(-8, 8, g.vec_shape))
    a2 = np.asfortranarray(rng.standard_normal(g.vec_shape) * 10.0 ** rng.integers(-8, 8, g.vec_shape))
    own = g.owned_slice()
    seq = []
    for j in range(g.jlo - 1, g.jhi):
        for i in range(g.ilo - 1, g.ihi):
            for k in range(g.nz):
                seq.append((a1[k, i, j], a2[k, i, j]))
    # non-reproducible order: local = (local + a1) + a2, loops j, i, k
    expect = reduce(lambda s, t: (s + t[0]) + t[1], seq, 0.0)
    assert po.global_sum(g, a1, a2) == expect
    # reproducible sum (integer vectors) of the packed a1 + a2: the exact sum, correctly rounded;
    # the DDPDD alternative is within 1 ulp of it
    exact = sum(Fraction(x + y) for x, y in seq)
    g.reprosum = 1
    try:
        r = po.global_sum(g, a1, a2)
    finally:
        g.reprosum = 0
    assert r == float(exact)
    hi, lo = po.ddpdd_sum(np.array([x + y for x, y in seq]))
    assert abs(Fraction(hi) - exact) <= abs(np.spacing(float(exact)))


def test_halo_rules():
    for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW):
        g = SPD[bc][0]
        a = np.asfortranarray(np.arange(np.prod(g.vec_shape), dtype=float).reshape(g.vec_shape, order="F") + 1.0)
        b = a.copy(order="F")
        po.halo(g, b)
        nv, nw = g.nvx, g.nvy
        if bc == BC_PERIODIC:
            # one-task periodic: vertex 1,2 <- nv-2,nv-1; vertex nv <- 3 (E/W on owned rows, then N/S)
            e = a.copy()
            e[:, 0:2, g.jlo - 1:g.jhi] = a[:, nv - 3:nv - 1, g.jlo - 1:g.jhi]
            e[:, nv - 1, g.jlo - 1:g.jhi] = a[:, 2, g.jlo - 1:g.jhi]
            e2 = e.copy()
            e2[:, :, 0:2] = e[:, :, nw - 3:nw - 1]
            e2[:, :, nw - 1] = e[:, :, 2]
            assert np.array_equal(b, e2)
        elif bc == BC_NO_ICE:
            e = a.copy()
            e[:, 0:2, :] = 0; e[:, nv - 2:, :] = 0; e[:, :, 0:2] = 0; e[:, :, nw - 2:] = 0
            assert np.array_equal(b, e)
            assert np.all(b[:, g.ihi - 1, :] == 0)   # the last owned column is zeroed too
        else:
            e = a.copy()
            e[:, 0, :] = 0; e[:, nv - 1, :] = 0; e[:, :, 0] = 0; e[:, :, nw - 1] = 0
            assert np.array_equal(b, e)


# ----------------------------------------------------------------- element matrix
def _brick_integrals(a, b, c):
    """Closed-form integrals of trilinear basis products on [0,a]x[0,b]x[0,c]
    from 1-D linear-element integrals."""
    sx = [-1, 1, 1, -1, -1, 1, 1, -1]
    sy = [-1, -1, 1, 1, -1, -1, 1, 1]
    sz = [-1, -1, -1, -1, 1, 1, 1, 1]

    def dd(L, s, t):   # int N_s' N_t'
        return (1.0 if s == t else -1.0) / L

    def mm(L, s, t):   # int N_s N_t
        return L * (1.0 / 3.0 if s == t else 1.0 / 6.0)

    def dm(s):         # int N_s' N_t (any t)
        return 0.5 * s

    I = {}
    for key in ("xx", "yy", "zz", "xy", "yx"):
        I[key] = np.zeros((8, 8))
    for r in range(8):
        for q in range(8):
            X = (sx[r], sx[q]); Y = (sy[r], sy[q]); Zs = (sz[r], sz[q])
            I["xx"][r, q] = dd(a, *X) * mm(b, *Y) * mm(c, *Zs)
            I["yy"][r, q] = mm(a, *X) * dd(b, *Y) * mm(c, *Zs)
            I["zz"][r, q] = mm(a, *X) * mm(b, *Y) * dd(c, *Zs)
            I["xy"][r, q] = dm(sx[r]) * dm(sy[q]) * mm(c, *Zs)   # int phi_r,x phi_q,y
            I["yx"][r, q] = dm(sy[r]) * dm(sx[q]) * mm(c, *Zs)
    return I


def test_element_matrix_brick_closed_form():
    """Constant viscosity on an axis-aligned brick: the BP element blocks
    are eta/vol0 times combinations of closed-form basis integrals."""
    a, b, c = 1000.0, 700.0, 90.0
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1])
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1])
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1])
    x = (sx + 1) / 2 * a
    y = (sy + 1) / 2 * b
    z = (sz + 1) / 2 * c
    eta = 3.7e6
    st, K, em = po.element_matrix(x, y, z, np.zeros(8), np.zeros(8), 0.0, whichefvs=0, efvs_constant=eta)
    assert st == 0 and em == eta
    I = _brick_integrals(a, b, c)
    f = eta / 1e9
    assert np.allclose(K[0], f * (4 * I["xx"] + I["yy"] + I["zz"]), rtol=1e-12, atol=0)
    assert np.allclose(K[3], f * (4 * I["yy"] + I["xx"] + I["zz"]), rtol=1e-12, atol=0)
    assert np.allclose(K[1], f * (2 * I["xy"] + I["yx"]), rtol=1e-12, atol=1e-20)
    assert np.allclose(K[2], f * (2 * I["yx"] + I["xy"]), rtol=1e-12, atol=1e-20)


def test_element_matrix_invariants_distorted():
    """Symmetry (Kuu, Kvv symmetric; Kuv = Kvu^T), rigid translation in the
    null space, positive semi-definiteness, and the viscosity of a linear
    field (trilinear elements reproduce it exactly at every point)."""
    rng = np.random.default_rng(7)
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1])
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1])
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1])
    x = (sx + 1) / 2 * 1000 + rng.uniform(-80, 80, 8)
    y = (sy + 1) / 2 * 1000 + rng.uniform(-80, 80, 8)
    z = (sz + 1) / 2 * 120 + rng.uniform(-20, 20, 8) + 0.01 * x
    u = rng.uniform(0, 100, 8)
    v = rng.uniform(-20, 20, 8)
    st, K, em = po.element_matrix(x, y, z, u, v, 1.2e5)
    assert st == 0
    Kuu, Kuv, Kvu, Kvv = K
    assert np.array_equal(Kuu, Kuu.T) and np.array_equal(Kvv, Kvv.T)
    assert np.allclose(Kuv, Kvu.T, rtol=1e-14, atol=0)
    big = np.block([[Kuu, Kuv], [Kvu, Kvv]])
    sc = np.abs(big).max()
    assert np.abs(big @ np.r_[np.ones(8), np.zeros(8)]).max() < 1e-12 * sc
    assert np.abs(big @ np.r_[np.zeros(8), np.ones(8)]).max() < 1e-12 * sc
    assert np.linalg.eigvalsh(big).min() > -1e-12 * sc
    # linear velocity u = s x: eps_xx = s, so efvs = flwafact * (emin^2 + s^2)^(-1/3)
    s = 3e-3
    st, K, em = po.element_matrix(x, y, z, s * x, np.zeros(8), 1.2e5)
    assert np.isclose(em, 1.2e5 * (1e-16 + s * s) ** (-1.0 / 3.0), rtol=1e-12)


def test_gravity_load_column_sum():
    """The local gravity load of a column sums to -rho g dx dy H ds/dx / vol0."""
    p, s = stream_system()
    g = p.grid
    geo = s["geo"]
    av = s["active_vertex"].astype(bool)
    tot = s["loadu"].sum(axis=0)
    expect = -910.0 * 9.81 * p.dx * p.dy * geo["stagthck"] / 1e9 * geo["dusrf_dx"]
    assert np.allclose(tot[av], expect[av], rtol=1e-12, atol=0)


def test_dirichlet_rows_and_columns():
    """No-slip: basal rows become identity rows, their columns are eliminated
    symmetrically (glissade_velo_higher.F90:9305-9498)."""
    p, s = dome_system()
    g = p.grid
    av = s["active_vertex"].astype(bool)
    own = np.zeros_like(av)
    own[g.ilo - 1:g.ihi, g.jlo - 1:g.jhi] = True
    sel = av & own
    k = g.nz - 1
    for blk, diag in ((0, 1.0), (1, 0.0), (2, 0.0), (3, 1.0)):
        rows = s[("Auu", "Auv", "Avu", "Avv")[blk]][:, k][:, sel]
        expect = np.zeros(27)
        expect[13] = diag
        # couplings to level nz+1 do not exist; the rest of the row is zero
        assert np.all(rows[:18] == expect[:18, None])
    # no row couples to a basal (prescribed) node except the basal rows themselves
    A = s["Auu"]
    for m in range(18, 27):
        assert np.all(A[m, k - 1][sel] == 0.0)


# ----------------------------------------------------------------- full BP: analytic slab
def _slab_error(nz):
    th = 5.0 * pi / 180
    gn, rhoi, grav, beta, Hp = 3.0, 910.0, 9.81, 2000.0, 1000.0
    mu_n_pwrn = 2.0 ** ((3.0 - gn) / 2.0) / (gn + 1.0) * sin(th) ** (gn - 1.0) * (rhoi * grav * Hp) ** (gn - 1.0) * 1e6
    flwa = 1.0 / (2.0 ** ((1.0 + gn) / 2.0) * mu_n_pwrn)
    p = synthetic.slab(n_cells=6, nz=nz, dx=50.0, thck=Hp / cos(th), theta_deg=5.0, beta=beta, flwa=flwa)
    r = po.picard(p, algo="chrongear", precond=2, linear_tol=1e-11, linear_maxiters=3000, resid_target=1e-9,
                  maxiter=300)
    assert r["converged"]
    g = p.grid
    u = r["uvel"][:, g.ilo + 1, g.jlo + 1]
    H = p.thck[0, 0] * cos(th)
    mu_n = 1.0 / (2.0 ** ((1.0 + gn) / (2.0 * gn)) * flwa ** (1.0 / gn))
    eta = (beta * H / mu_n ** gn) * (rhoi * grav * H) ** (gn - 1)
    velscale = (rhoi * grav * H / mu_n) ** gn * H
    zp = (1.0 - p.sigma) * p.thck[0, 0] / H * cos(th)
    ufo = tan(th) / eta - 2 ** ((1.0 - gn) / 2.0) * tan(th) ** gn / (
        (gn + 1) * (1.0 + 3.0 * sin(th) ** 2) ** ((gn + 1.0) / 2.0)) * ((1.0 - zp) ** (gn + 1) - 1.0)
    return float(np.abs(u / velscale - ufo).max() / np.abs(ufo).max()), abs(u[-1] / velscale / ufo[-1] - 1.0)


def test_slab_matches_first_order_analytic_solution():
    """Closed-form first-order (BP) velocity of a sliding, shearing slab
    (tests/slab/plotSlab.py:242-250, Dukowicz 2012): the restated assembly,
    basal friction and solver converge to it at second order in the level
    spacing, and the basal sliding velocity tan(theta)/eta is reproduced."""
    e1, b1 = _slab_error(11)
    e2, b2 = _slab_error(21)
    assert e2 < 1e-3
    assert 3.5 < e1 / e2 < 4.5
    assert b1 < 1e-7 and b2 < 1e-7


def _front_problem(H, topg, topg_ocean=-1000.0, eus=0.0, nz=6):
    """Uniform slab of thickness H on a bed at topg in the cells i <= 6 (all
    j), ice-free ocean (bed topg_ocean < eus) east of it."""
    g = synthetic.Grid.from_owned(8, 8, nz, global_bc=BC_NO_ICE)
    thck = np.zeros((g.nx, g.ny), order="F")
    thck[:6, :] = H
    bed = np.full((g.nx, g.ny), float(topg_ocean), order="F")
    bed[:6, :] = topg
    p = synthetic.Problem(grid=g, dx=1000.0, dy=750.0, sigma=glide_sigma(nz), thck=thck,
                          topg=bed,
                          flwa=np.full((nz - 1, g.nx, g.ny), 1e-16, order="F"),
                          beta=np.zeros((g.nx - 1, g.ny - 1), order="F"), uvel=g.zeros_vec(), vvel=g.zeros_vec(),
                          eus=eus)
    return p


@pytest.mark.parametrize("local", [0, 1])
@pytest.mark.parametrize("case", ["floating", "grounded_cliff"])
def test_lateral_bc_calving_front_force(local, case):
    """load_vector_lateral_bc (glissade_velo_higher.F90:4453-4957): the
    depth-integrated net pressure on a calving front is
    rho_i g H^2 / 2 - rho_o g D^2 / 2 (D = submerged depth) per unit
    width -- rho_i g H^2 (1 - rho_i/rho_o) / 2 for a floating shelf and
    rho_i g H^2 / 2 for a grounded cliff whose surface is above H -- and a
    uniform pressure loads each level with half the thickness of the layers
    on either side."""
    rhoi, rhoo, grav, vol0 = 910.0, 1028.0, 9.81, 1.0e9
    H = 400.0
    if case == "floating":
        p = _front_problem(H, topg=-1000.0)
        F = 0.5 * rhoi * grav * H * H * (1.0 - rhoi / rhoo)
    else:
        # ice on a bed above sea level facing open ocean: the surface is above H
        p = _front_problem(H, topg=5.0)
        F = 0.5 * rhoi * grav * H * H
    p.assemble_lateral = local
    geo = po.geometry(p)
    g = p.grid
    assert geo["ocean_mask"][6, 3] == 1 and geo["active_cell"][5, 3] == 1
    xV, yV = p.vertex_coords()
    lu, lv = g.zeros_vec(), g.zeros_vec()
    po.lib().orc_load_vector_lateral_bc(g.nx, g.ny, g.nz, po.dp(np.ascontiguousarray(p.sigma)), 2, local,
                                        po.ip(geo["ocean_mask"]), po.ip(geo["active_cell"]), po.dp(xV), po.dp(yV),
                                        po.dp(geo["usrf"]), po.dp(p.thck), po.dp(geo["stagusrf_lateral"]),
                                        po.dp(geo["stagthck_lateral"]), po.dp(lu), po.dp(lv))
    # east faces of the cells (6, j), j = 3 .. ny-1, load the vertex column i = 6
    nface = g.ny - 1 - 3 + 1
    east = lu[:, 5, :]
    assert np.isclose(east.sum() * vol0, nface * F * p.dy, rtol=1e-12)
    assert np.abs(lv).max() == 0.0
    # level distribution of the uniform pressure: (dsigma_below + dsigma_above)/2
    ds = np.diff(p.sigma)
    share = np.zeros(g.nz)
    share[:-1] += ds / 2
    share[1:] += ds / 2
    assert np.allclose(east.sum(axis=1) * vol0, nface * F * p.dy * share, rtol=1e-12)


def test_grounded_fraction_rules():
    """f_ground (no grounding-line parameterisation, grounded-neighbour rule):
    1 under and next to grounded ice or land, 0 at vertices surrounded by
    floating ice or open ocean only; beta_internal vanishes where f_ground does."""
    p = _front_problem(400.0, topg=-1000.0)      # floating shelf
    p.thck[:3, :] = 1200.0                      # grounded band in the west (1200 m > 1028/910 * 1000)
    geo = po.geometry(p)
    fg = geo["f_ground"]
    flt, ice = geo["floating_mask"], geo["ice_mask"]
    assert flt[4, 4] == 1 and flt[1, 4] == 0 and ice[1, 4] == 1
    assert fg[0, 4] == 1.0 and fg[2, 4] == 1.0        # touches grounded cells
    assert fg[3, 4] == 0.0 and fg[6, 4] == 0.0        # floating / ocean only
    p.whichbabc = synthetic.HO_BABC_BETA_CONSTANT
    p.beta_const = 1000.0
    b = po.beta_internal(p, geo)
    # (the no-ice halo rule also clears the last owned row and column)
    own = (slice(p.grid.ilo - 1, p.grid.ihi - 1), slice(p.grid.jlo - 1, p.grid.jhi - 1))
    assert np.array_equal(b[own] == 0.0, fg[own] == 0.0)


def test_accelerated_picard_same_fixed_point():
    """Picard acceleration (evaluate_accelerated_picard_3d,
    glissade_velo_higher.F90:10052-10184) changes the path, not the fixed
    point: driven to a tight tolerance it reaches the plain Picard solution
    in fewer iterations; every accepted extension has alpha in [1, 3] on the
    0.4 ladder, and a rejected one resets alpha to 1."""
    p = synthetic.dome(n_cells=15, nz=6, dx=2000.0, radius=14e3, h0=1500.0)
    kw = dict(algo="standard", precond=1, linear_tol=1e-12, linear_maxiters=2000, resid_target=1e-8, maxiter=200)
    r0 = po.picard(p, **kw)
    r1 = po.picard(p, accelerate=True, **kw)
    assert r0["converged"] and r1["converged"]
    assert r1["iterations"] < r0["iterations"]
    rel = np.linalg.norm(r1["uvel"] - r0["uvel"]) / np.linalg.norm(r0["uvel"])
    assert rel < 1e-6, rel
    ladder = [1.0 + 0.4 * m for m in range(6)]
    assert all(any(abs(a - l) < 1e-12 for l in ladder) for a in r1["alpha"])
    assert max(r1["alpha"]) <= 3.0 and 1.0 in r1["alpha"]


# ----------------------------------------------------------------- 2-D solvers
from systems import random_spd_system_2d   # noqa: E402

SPD2 = {bc: random_spd_system_2d(bc=bc) for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW)}


def owned_operator_2d(g, act, A):
    """Sparse operator of matvec_multiply_structured_2d (pcg.F90:5656-5773) on
    owned unknowns (comp, j, i), halo rule as in owned_operator."""
    ni, nj = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1
    periodic = g.global_bc in (0, 2)
    uid = lambda c, i, j: (c * nj + (j - g.jlo)) * ni + (i - g.ilo)
    rows, cols, vals = [], [], []
    for j in range(g.jlo, g.jhi + 1):
        for i in range(g.ilo, g.ihi + 1):
            if not act[i - 1, j - 1]:
                continue
            for jA in (-1, 0, 1):
                for iA in (-1, 0, 1):
                    qi, qj = i + iA, j + jA
                    if not (g.ilo <= qi <= g.ihi and g.jlo <= qj <= g.jhi):
                        if not periodic:
                            continue
                        qi = g.ilo + (qi - g.ilo) % ni
                        qj = g.jlo + (qj - g.jlo) % nj
                    m = 1 + (iA + 1) + 3 * (jA + 1) - 1
                    for cr, blkpair in ((0, (0, 1)), (1, (2, 3))):
                        for cc, blk in enumerate(blkpair):
                            v = A[blk][i - 1, j - 1, m]
                            if v != 0.0:
                                rows.append(uid(cr, i, j))
                                cols.append(uid(cc, qi, qj))
                                vals.append(v)
    n = 2 * ni * nj
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def to_owned_2d(g, u, v):
    s = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    return np.concatenate([u[s].T.ravel(), v[s].T.ravel()])


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW])
def test_matvec_2d_matches_sparse_operator(bc):
    g, act, A, bu, bv = SPD2[bc]
    M = owned_operator_2d(g, act, A)
    rng = np.random.default_rng(6)
    xu = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    xv = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    po.halo(g, xu)
    po.halo(g, xv)
    yu, yv = po.matvec_2d(g, act, *A, xu, xv)
    ref = M @ to_owned_2d(g, xu, xv)
    assert np.allclose(to_owned_2d(g, yu, yv), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("algo", ["standard", "chrongear"])
@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_NO_ICE])
def test_2d_iterates_equal_textbook_pcg(bc, algo):
    """The 2-D solvers' k-th iterates equal a textbook PCG with the diagonal
    preconditioner on the same operator."""
    g, act, A, bu, bv = SPD2[bc]
    M = owned_operator_2d(g, act, A)
    b = to_owned_2d(g, bu, bv)
    dg = to_owned_2d(g, A[0][:, :, 4], A[3][:, :, 4])
    ref = textbook_pcg(M, b, lambda r: np.where(dg != 0, r / np.where(dg != 0, dg, 1.0), 0.0), 10)
    for k in (1, 2, 5, 7, 10):
        xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
        st, err, nit = po.pcg_solve_2d(g, act, *A, bu, bv, xu, xv, algo=algo, precond=1, tolerance=0.0,
                                       maxiters=k)
        assert st == 0 and nit == k
        x = to_owned_2d(g, xu, xv)
        assert np.linalg.norm(x - ref[k - 1]) <= 1e-10 * np.linalg.norm(ref[k - 1]), k


@pytest.mark.parametrize("precond", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("algo", ["standard", "chrongear"])
def test_2d_converged_solution_matches_sparse_solve(algo, precond):
    for bc in (BC_PERIODIC, BC_NO_ICE, BC_OUTFLOW):
        g, act, A, bu, bv = SPD2[bc]
        M = owned_operator_2d(g, act, A)
        b = to_owned_2d(g, bu, bv)
        keep = np.asarray(M.getnnz(axis=1) > 0)
        xs = np.zeros_like(b)
        xs[keep] = spla.spsolve(M[keep][:, keep].tocsc(), b[keep])
        xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
        st, err, nit = po.pcg_solve_2d(g, act, *A, bu, bv, xu, xv, algo=algo, precond=precond, tolerance=1e-12,
                                       maxiters=3000)
        assert st == 0 and err < 1e-12
        assert np.linalg.norm(to_owned_2d(g, xu, xv) - xs) <= 1e-10 * np.linalg.norm(xs), (bc, precond)


def test_2d_tridiagonal_preconditioner_first_step():
    """2-D precond 3 is the row (u) and column (v) line solves of the stencil
    itself, without a column part: the first standard iterate is
    x1 = alpha z0, z0 = PC(b), alpha = (b, z0) / (z0, A z0)."""
    g, act, A, bu, bv = SPD2[BC_NO_ICE]
    M = owned_operator_2d(g, act, A)
    b = to_owned_2d(g, bu, bv)
    ni, nj = g.ihi - g.ilo + 1, g.jhi - g.jlo + 1
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    z = np.zeros_like(b)
    Z = z.reshape(2, nj, ni)
    Bv = b.reshape(2, nj, ni)
    for c, blk, m_sub, m_sup in ((0, 0, 3, 5), (1, 3, 1, 7)):
        sub, dia, sup = (A[blk][:, :, m][own] for m in (m_sub, 4, m_sup))
        dia = np.where(dia == 0, 1.0, dia)
        if c == 0:       # rows along i for each j
            for jo in range(nj):
                T = np.diag(dia[:, jo]) + np.diag(sub[1:, jo], -1) + np.diag(sup[:-1, jo], 1)
                Z[c, jo, :] = np.linalg.solve(T, Bv[c, jo, :])
        else:            # columns along j for each i
            for io in range(ni):
                T = np.diag(dia[io, :]) + np.diag(sub[io, 1:], -1) + np.diag(sup[io, :-1], 1)
                Z[c, :, io] = np.linalg.solve(T, Bv[c, :, io])
    alpha = (b @ z) / (z @ (M @ z))
    xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
    st, err, nit = po.pcg_solve_2d(g, act, *A, bu, bv, xu, xv, algo="standard", precond=3, tolerance=0.0,
                                   maxiters=1)
    assert st == 0
    x = to_owned_2d(g, xu, xv)
    assert np.linalg.norm(x - alpha * z) <= 1e-11 * np.linalg.norm(alpha * z)


def test_depth_integrated_system_is_projected_operator():
    """depth_integrated_system gives P^T A P: for x2 spread over every level,
    the level sum of the 3-D matvec equals the 2-D matvec of the collapsed
    stencils (both oracle matvecs are pinned above)."""
    from cism_amd.synthetic import depth_integrated_system
    p, s = dome_system()
    g, act = p.grid, s["active_vertex"]
    A = [s["Auu"], s["Auv"], s["Avu"], s["Avv"]]
    A2, bu2, bv2 = depth_integrated_system(*A, s["bu"], s["bv"])
    rng = np.random.default_rng(11)
    xu2 = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    xv2 = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)))
    xu3 = np.asfortranarray(np.broadcast_to(xu2, (g.nz, g.nvx, g.nvy)))
    xv3 = np.asfortranarray(np.broadcast_to(xv2, (g.nz, g.nvx, g.nvy)))
    yu3, yv3 = po.matvec(g, act, *A, xu3, xv3)
    yu2, yv2 = po.matvec_2d(g, act, *A2, xu2, xv2)
    sc = np.abs(yu3).sum(axis=0).max()
    assert np.allclose(yu3.sum(axis=0), yu2, rtol=0, atol=1e-12 * sc)
    assert np.allclose(yv3.sum(axis=0), yv2, rtol=0, atol=1e-12 * sc)
    assert np.array_equal(bu2, s["bu"].sum(axis=0))
    # symmetric: the (i,j)->(i+iA,j+jA) entry equals the transpose entry
    own = slice(g.ilo, g.ihi - 1)
    for iA, jA in ((1, 0), (0, 1), (1, 1), (1, -1)):
        m, mt = 4 + iA + 3 * jA, 4 - iA - 3 * jA
        a = A2[0][own, own, m]
        b = A2[0][own.start + iA:own.stop + iA, own.start + jA:own.stop + jA, mt]
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(A2[0]).max())


# ----------------------------------------------------------- 2-D SSA assembly
def _exact_rect_integrals(a, b, npts=6):
    """Exact integrals over [0,a]x[0,b] of products of bilinear-basis
    derivatives (nodes (0,0), (a,0), (a,b), (0,b)), by a 6-point
    Gauss-Legendre rule (exact for these polynomials)."""
    xg, wg = np.polynomial.legendre.leggauss(npts)
    X = 0.5 * a * (xg + 1.0)
    Y = 0.5 * b * (xg + 1.0)
    W = np.outer(wg, wg) * (0.25 * a * b)
    xx, yy = np.meshgrid(X, Y, indexing="ij")
    sx = np.array([-1, 1, 1, -1]) / a
    sy = np.array([-1, -1, 1, 1]) / b
    fx = [(xx if s > 0 else a - xx) / a for s in (-1, 1, 1, -1)]
    fy = [(yy if s > 0 else b - yy) / b for s in (-1, -1, 1, 1)]
    phix = [sx[n] * fy[n] for n in range(4)]     # d phi_n / dx
    phiy = [sy[n] * fx[n] for n in range(4)]     # d phi_n / dy
    I = lambda f, g_: float((f * g_ * W).sum())
    Ixx = np.array([[I(phix[r], phix[c]) for c in range(4)] for r in range(4)])
    Iyy = np.array([[I(phiy[r], phiy[c]) for c in range(4)] for r in range(4)])
    Ixy = np.array([[I(phix[r], phiy[c]) for c in range(4)] for r in range(4)])
    return Ixx, Iyy, Ixy


@pytest.mark.parametrize("a,b", [(1000.0, 1000.0), (2000.0, 500.0)])
def test_ssa_element_matrix_is_exact_weak_form_on_rectangles(a, b):
    """SSA weak form (MacAyeal 1989): K_uu = int eta h (4 phi_x phi_x + phi_y phi_y),
    K_uv = int eta h (2 phi_x phi_y' + phi_y phi_x'), ...; the 2x2 Gauss rule is
    exact for bilinear elements on rectangles, so the element matrix must equal
    the exact integrals (scaled by 1/vol0)."""
    x = np.array([0.0, a, a, 0.0]) + 3000.0
    y = np.array([0.0, 0.0, b, b]) - 7000.0
    h = np.full(4, 800.0)
    eta = 3.0e13
    st, (Kuu, Kuv, Kvu, Kvv), em = po.element_matrix_2d_ssa(x, y, np.zeros(4), np.zeros(4), h, 0.0, whichefvs=0,
                                                            efvs_constant=eta)
    assert st == 0 and em == eta
    Ixx, Iyy, Ixy = _exact_rect_integrals(a, b)
    f = eta * 800.0 / 1.0e9
    sc = f * max(np.abs(Ixx).max(), np.abs(Iyy).max())
    assert np.allclose(Kuu, f * (4 * Ixx + Iyy), rtol=0, atol=1e-13 * sc)
    assert np.allclose(Kvv, f * (4 * Iyy + Ixx), rtol=0, atol=1e-13 * sc)
    assert np.allclose(Kuv, f * (2 * Ixy + Ixy.T), rtol=0, atol=1e-13 * sc)
    assert np.allclose(Kvu, f * (2 * Ixy.T + Ixy), rtol=0, atol=1e-13 * sc)


def test_ssa_element_rotation_covariance_and_nonlinear_viscosity():
    """On a distorted quad with a linear velocity field: efvs is Glen's law of
    the SSA effective strain rate u_x^2 + v_y^2 + u_x v_y + (u_y + v_x)^2 / 4;
    rotating coordinates and velocities by 90 degrees maps the blocks as
    K'uu = Kvv, K'vv = Kuu, K'uv = -Kvu, K'vu = -Kuv."""
    x = np.array([0.0, 1100.0, 1250.0, -80.0])
    y = np.array([0.0, 60.0, 900.0, 1010.0])
    A_, B_, C_, D_ = 2.0e-3, -1.0e-3, 5.0e-4, -3.0e-3    # u = A x + B y, v = C x + D y (1/yr)
    u = A_ * x + B_ * y
    v = C_ * x + D_ * y
    h = np.array([900.0, 950.0, 1000.0, 1020.0])
    fl, emin, n = 1.7e5, 1.0e-8, 3.0
    st, K, em = po.element_matrix_2d_ssa(x, y, u, v, h, fl, whichefvs=2, effstrain_min=emin, n_glen=n)
    assert st == 0
    esq = emin ** 2 + A_ ** 2 + D_ ** 2 + A_ * D_ + 0.25 * (B_ + C_) ** 2
    assert em == pytest.approx(fl * esq ** ((1.0 - n) / (2.0 * n)), rel=1e-13)
    st, Kr, emr = po.element_matrix_2d_ssa(-y, x, -v, u, h, fl, whichefvs=2, effstrain_min=emin, n_glen=n)
    assert st == 0 and emr == pytest.approx(em, rel=1e-13)
    sc = max(np.abs(k).max() for k in K)
    assert np.allclose(Kr[0], K[3], atol=1e-12 * sc) and np.allclose(Kr[3], K[0], atol=1e-12 * sc)
    assert np.allclose(Kr[1], -K[2], atol=1e-12 * sc) and np.allclose(Kr[2], -K[1], atol=1e-12 * sc)
    # a degenerate (collapsed) element is a singular Jacobian
    st, _, _ = po.element_matrix_2d_ssa(np.zeros(4), y, u, v, h, fl)
    assert st == 4


def test_vertical_average_of_layer_field():
    sigma = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    var = np.zeros((4, 6, 6), order="F")
    var[:, :, :] = 2.5
    var[:, 2, 3] = [1.0, 2.0, 3.0, 4.0]
    out = po.vertical_average(sigma, var)
    assert np.allclose(out[0, 0], 2.5, rtol=1e-15)
    assert out[2, 3] == pytest.approx(1.0 * 0.1 + 2.0 * 0.2 + 3.0 * 0.3 + 4.0 * 0.4, rel=1e-15)


def test_ssa_slab_closed_form_and_patch_test():
    """Periodic inclined slab (uniform H, beta): the 2-D SSA solution is the
    uniform sliding velocity u = -rho_i g H ds/dx / beta (the stiffness
    annihilates constant fields; the level-summed gravity load is
    -rho_i g H ds/dx dx dy / vol0); and the assembled operator maps a linear
    velocity field to zero at interior vertices (consistency)."""
    p = synthetic.slab(n_cells=8, nz=5)
    s = po.assemble_2d(p)
    g = p.grid
    geo = s["geo"]
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    dsdx = geo["dusrf_dx"][own]
    assert np.ptp(dsdx) < 1e-12 * abs(dsdx).max()
    ucf = -910.0 * 9.81 * 1000.0 * dsdx.mean() / 2000.0
    xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
    st, err, nit = po.pcg_solve_2d(g, s["active_vertex"], s["Auu"], s["Auv"], s["Avu"], s["Avv"], s["bu"], s["bv"],
                                   xu, xv, precond=1, tolerance=1e-9, maxiters=20000)
    assert st == 0
    assert np.allclose(xu[own], ucf, rtol=1e-6) and np.abs(xv[own]).max() < 1e-6 * ucf
    # patch test with constant viscosity on the same grid
    p.whichefvs, p.efvs_constant = 0, 1.0e13
    s = po.assemble_2d(p)
    xV, yV = p.vertex_coords()
    lu = np.asfortranarray(3e-3 * xV - 1e-3 * yV)
    lv = np.asfortranarray(2e-3 * xV + 4e-3 * yV)
    Az = [s[k].copy(order="F") for k in ("Auu", "Auv", "Avu", "Avv")]
    Az[0][:, :, 4] -= s["beta_internal"] * p.dx * p.dy / 1e9 * (s["active_vertex"] != 0)
    Az[3][:, :, 4] -= s["beta_internal"] * p.dx * p.dy / 1e9 * (s["active_vertex"] != 0)
    yu, yv = po.matvec_2d(g, s["active_vertex"], *Az, lu, lv)
    inner = (slice(g.ilo, g.ihi - 1), slice(g.jlo, g.jhi - 1))
    scale = np.abs(Az[0]).max() * np.abs(lu).max()
    assert np.abs(yu[inner]).max() < 1e-11 * scale and np.abs(yv[inner]).max() < 1e-11 * scale


def test_ssa_basal_dirichlet_count_and_symmetry_report():
    """Basal term = dx dy beta / vol0 on the diagonal of active rows; with u
    and v prescribed on a set of vertices, the Dirichlet-modified system gives
    those values there and satisfies the unmodified equations elsewhere;
    count_nonzeros and the symmetry report against numpy."""
    p = synthetic.ice_stream(24, nz=5, seed=4)
    s = po.assemble_2d(p)
    g = p.grid
    av = s["active_vertex"]
    A = [s[k] for k in ("Auu", "Auv", "Avu", "Avv")]
    pb = synthetic.ice_stream(24, nz=5, seed=4)
    pb.beta = np.zeros_like(pb.beta)
    s0 = po.assemble_2d(pb)
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    d = (s["Auu"] - s0["Auu"])[:, :, 4][own]
    want = (p.dx * p.dy / 1e9 * s["beta_internal"] * (av != 0))[own]
    assert np.allclose(d, want, rtol=1e-9, atol=1e-12 * np.abs(s["Auu"]).max())
    # count and symmetry report
    nnz = sum(int(np.count_nonzero(a[own][np.asarray(av, bool)[own]])) for a in A)
    assert s["nnz"] == nnz and s["nwarn_asym"] == 0
    B = [a.copy(order="F") for a in A]
    i0, j0 = g.ilo + 5, g.jlo + 5
    B[0][i0 - 1, j0 - 1, 5] *= 1.5       # plant one asymmetric pair (row (i0,j0), neighbour i0+1)
    L = po.lib()
    nw = L.orc_check_symmetry_assembled_2d(g.nx, g.ny, g.ilo, g.ihi, g.jlo, g.jhi, po.ip(av), *[po.dp(b) for b in B])
    assert nw == 2
    # Dirichlet elimination
    rng = np.random.default_rng(9)
    mask = np.zeros((g.nvx, g.nvy), np.int32, order="F")
    pick = rng.random((g.nvx, g.nvy)) < 0.15
    mask[pick & (np.asarray(av) != 0)] = 1
    uval = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * 50.0)
    vval = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * 50.0)
    Ad = [a.copy(order="F") for a in A]
    bu, bv = s["bu"].copy(order="F"), s["bv"].copy(order="F")
    L.orc_dirichlet_bc_2d(g.nx, g.ny, 2, po.ip(av), po.ip(mask), po.ip(mask), po.dp(uval), po.dp(vval),
                          *[po.dp(a) for a in Ad], po.dp(bu), po.dp(bv))
    xu, xv = np.zeros((g.nvx, g.nvy), order="F"), np.zeros((g.nvx, g.nvy), order="F")
    st, err, nit = po.pcg_solve_2d(g, av, *Ad, bu, bv, xu, xv, precond=1, tolerance=1e-13, maxiters=20000)
    assert st == 0
    m_own = (mask[own] == 1) & (np.asarray(av)[own] != 0)
    assert np.allclose(xu[own][m_own], uval[own][m_own], rtol=1e-9) and np.allclose(xv[own][m_own], vval[own][m_own], rtol=1e-9)
    yu, yv = po.matvec_2d(g, av, *A, xu, xv)
    free = (mask[own] == 0) & (np.asarray(av)[own] != 0)
    # the halo rule (no_ice) zeroes the last owned row/column of x and b
    free[-1, :] = False
    free[:, -1] = False
    scale = np.abs(s["bu"][own]).max()
    assert np.abs((yu - s["bu"])[own][free]).max() < 1e-7 * scale
    assert np.abs((yv - s["bv"])[own][free]).max() < 1e-7 * scale


# ------------------------------------------------------------------- 2-D DIVA
def _diva_slab(nz):
    th = 5.0 * pi / 180
    gn, rhoi, grav, beta, Hp = 3.0, 910.0, 9.81, 2000.0, 1000.0
    mu_n_pwrn = 2.0 ** ((3.0 - gn) / 2.0) / (gn + 1.0) * sin(th) ** (gn - 1.0) * (rhoi * grav * Hp) ** (gn - 1.0) * 1e6
    flwa = 1.0 / (2.0 ** ((1.0 + gn) / 2.0) * mu_n_pwrn)
    p = synthetic.slab(n_cells=6, nz=nz, dx=50.0, thck=Hp / cos(th), theta_deg=5.0, beta=beta, flwa=flwa)
    r = po.picard(p, two_d=True, approx="diva", precond=1, linear_tol=1e-11, linear_maxiters=3000,
                  resid_target=1e-9, maxiter=300)
    assert r["converged"]
    g = p.grid
    H = p.thck[0, 0] * cos(th)
    mu_n = 1.0 / (2.0 ** ((1.0 + gn) / (2.0 * gn)) * flwa ** (1.0 / gn))
    eta = (beta * H / mu_n ** gn) * (rhoi * grav * H) ** (gn - 1)
    velscale = (rhoi * grav * H / mu_n) ** gn * H
    zp = (1.0 - p.sigma) * p.thck[0, 0] / H * cos(th)
    ufo = tan(th) / eta - 2 ** ((1.0 - gn) / 2.0) * tan(th) ** gn / (
        (gn + 1) * (1.0 + 3.0 * sin(th) ** 2) ** ((gn + 1.0) / 2.0)) * ((1.0 - zp) ** (gn + 1) - 1.0)
    u = r["uvel3d"][:, g.ilo + 1, g.jlo + 1]
    return float(np.abs(u / velscale - ufo).max() / np.abs(ufo).max()), abs(u[-1] / velscale / ufo[-1] - 1.0)


def test_diva_slab_matches_first_order_analytic_solution():
    """DIVA (Goldberg 2011) with the slope correction reproduces the
    first-order slab solution of tests/slab/plotSlab.py:242-250 through the
    whole path -- layer viscosities from the basal traction, the omega
    integrals, beta_eff, the basal term with the bed slope and the 3-D
    velocity reconstruction: basal velocity to rounding, the profile at
    second order in the level spacing."""
    e1, b1 = _diva_slab(11)
    e2, b2 = _diva_slab(21)
    assert e2 < 1e-3 and 3.5 < e1 / e2 < 4.5
    assert b1 < 1e-9 and b2 < 1e-9


def test_diva_omega_integrals_closed_form():
    """With a constant viscosity eta the single integral is omega_k(sigma) =
    H (1 - sigma^2) / (2 eta) exactly (midpoint rule on a linear integrand)
    and omega is its trapezoidal vertical mean."""
    p = synthetic.ice_stream(16, nz=7, seed=5)
    p.whichefvs, p.efvs_constant = 0, 2.5e13
    s = po.assemble_2d(p, approx="diva")
    g = p.grid
    acell = s["geo"]["active_cell"]
    sig = p.sigma
    for (i, j) in [(4, 5), (9, 7), (12, 12)]:
        if not acell[i - 1, j - 1]:
            continue
        H = p.thck[i - 1, j - 1]
        want = H * (1.0 - sig ** 2) / (2.0 * 2.5e13)
        assert np.allclose(s["omega_k"][:, i - 1, j - 1], want, rtol=1e-13, atol=1e-30)
        trap = sum(0.5 * (want[k] + want[k + 1]) * (sig[k + 1] - sig[k]) for k in range(g.nz - 1))
        assert s["omega"][i - 1, j - 1] == pytest.approx(trap, rel=1e-13)


def test_diva_without_basal_traction_is_ssa_stiffness():
    """With zero basal traction the DIVA layer viscosities have no vertical
    shear, so for a vertically uniform flow factor the DIVA stiffness equals
    the SSA stiffness (the vertical mean of the layer viscosities equals the
    viscosity of the vertically averaged flow factor)."""
    p = synthetic.ice_stream(16, nz=7, seed=6)
    p.beta = np.zeros_like(p.beta)
    g = p.grid
    rng = np.random.default_rng(3)
    u2 = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * 30.0)
    v2 = np.asfortranarray(rng.standard_normal((g.nvx, g.nvy)) * 30.0)
    ssa = po.assemble_2d(p, u2, v2)
    diva = po.assemble_2d(p, u2, v2, approx="diva")
    for key in ("Auu", "Auv", "Avu", "Avv"):
        a, b = ssa[key], diva[key]
        assert np.allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(a).max()), key


# ------------------------------------------------------------------- 2-D L1L2
@pytest.mark.parametrize("tau_perp", [0.0, 1.0e-3, 3.0e2, 1.0e4, 2.0e5, 1.0e6])
@pytest.mark.parametrize("effstrain", [1.0e-6, 1.0e-3, 0.3])
def test_l1l2_cubic_root_is_the_polynomial_root(effstrain, tau_perp):
    """tau_parallel of compute_effective_viscosity_L1L2 is the positive real
    root of x^3 + tau_perp^2 x - effstrain/A = 0 (numpy's companion-matrix
    roots), in both branches of the second cube root."""
    A = 1.0e-16
    got = po.lib().orc_l1l2_tau_parallel(effstrain, A, tau_perp)
    r = np.roots([1.0, 0.0, tau_perp * tau_perp, -effstrain / A])
    want = max(float(z.real) for z in r if abs(z.imag) <= 1e-9 * abs(z))
    assert got == pytest.approx(want, rel=1e-8)
    assert abs(got ** 3 + tau_perp ** 2 * got - effstrain / A) <= 1e-8 * (effstrain / A)


def _l1l2_velocity_direct(p, u2, v2, thck, dsdx, dsdy, eta):
    """compute_3d_velocity_L1L2 of the restatement on given fields (ice
    everywhere, viscosity eta in every layer)."""
    g = p.grid
    ice = np.ones((g.nx, g.ny), np.int32, order="F")
    acell, avert = ge.vertex_masks(ice)
    xV, yV = p.vertex_coords()
    u3, v3 = g.zeros_vec(), g.zeros_vec()
    u3[g.nz - 1] = u2
    v3[g.nz - 1] = v2
    efvs = np.full((g.nz - 1, g.nx, g.ny), eta, order="F")
    zero = np.zeros((g.nvx, g.nvy), np.int32, order="F")
    st = po.lib().orc_compute_3d_velocity_l1l2(
        g.nx, g.ny, g.nz, g.ilo, g.ihi, g.jlo, g.jhi, g.global_bc, 2, po.dp(np.ascontiguousarray(p.sigma)), p.dx,
        p.dy, po.ip(ice), po.ip(acell), po.ip(avert), po.ip(zero), po.ip(zero), po.dp(xV), po.dp(yV),
        po.dp(np.asfortranarray(thck)), po.dp(ge.stagger(thck)), po.dp(np.asfortranarray(dsdx)),
        po.dp(np.asfortranarray(dsdy)), po.dp(np.full((g.nz - 1, g.nx, g.ny), 1e-16, order="F")), po.dp(efvs), 0,
        3.0, po.dp(u3), po.dp(v3))
    assert st == 0
    return u3, v3


def test_l1l2_reconstruction_membrane_and_shear_closed_form():
    """With a constant viscosity eta, linear 2-D velocities u = a1 x + a2 y,
    v = b1 x + b2 y, thickness H = H0 + cx x + cy y and a uniform surface
    slope, the membrane terms of compute_3d_velocity_L1L2 are constant in
    each layer and the midpoint rule is exact, so
        u(sigma) - u_b = H (1 - sigma^2)/2 [2(2 a1 + b2) cx + (a2 + b1) cy - rho g H s_x / eta]
        v(sigma) - v_b = H (1 - sigma^2)/2 [(a2 + b1) cx + 2(2 b2 + a1) cy - rho g H s_y / eta]
    at every interior vertex (H at the vertex)."""
    p = synthetic.ice_stream(20, nz=9, seed=2)
    g = p.grid
    g.global_bc = 1
    a1, a2, b1, b2 = 3e-3, -1e-3, 2e-3, 5e-4
    cx, cy, H0, eta = 2e-3, -1e-3, 800.0, 3e13
    sx, sy = -2e-3, 1e-3
    xV, yV = p.vertex_coords()
    xc = (np.arange(g.nx) - 0.5) * p.dx
    yc = (np.arange(g.ny) - 0.5) * p.dy
    thck = np.asfortranarray(H0 + cx * xc[:, None] + cy * yc[None, :])
    u2 = np.asfortranarray(a1 * xV + a2 * yV)
    v2 = np.asfortranarray(b1 * xV + b2 * yV)
    dsdx = np.full((g.nvx, g.nvy), sx, order="F")
    dsdy = np.full((g.nvx, g.nvy), sy, order="F")
    u3, v3 = _l1l2_velocity_direct(p, u2, v2, thck, dsdx, dsdy, eta)
    Hv = ge.stagger(thck)
    rg = 910.0 * 9.81
    fu = 2 * (2 * a1 + b2) * cx + (a2 + b1) * cy - rg * Hv * sx / eta
    fv = (a2 + b1) * cx + 2 * (2 * b2 + a1) * cy - rg * Hv * sy / eta
    sig = p.sigma[:, None, None]
    wu = u2[None] + Hv[None] * (1 - sig ** 2) / 2 * fu[None]
    wv = v2[None] + Hv[None] * (1 - sig ** 2) / 2 * fv[None]
    inner = (slice(None), slice(g.ilo, g.ihi - 1), slice(g.jlo, g.jhi - 1))
    du, dv = (u3 - u2[None])[inner], (v3 - v2[None])[inner]
    assert np.abs(du - (wu - u2[None])[inner]).max() < 1e-9 * np.abs(du).max()
    assert np.abs(dv - (wv - v2[None])[inner]).max() < 1e-9 * np.abs(dv).max()


def _l1l2_slab(nz):
    th = 5.0 * pi / 180
    beta, Hp, A = 2000.0, 1000.0, 1.0e-16
    p = synthetic.slab(n_cells=6, nz=nz, dx=50.0, thck=Hp / cos(th), theta_deg=5.0, beta=beta, flwa=A)
    r = po.picard(p, two_d=True, approx="l1l2", precond=1, linear_tol=1e-11, linear_maxiters=3000,
                  resid_target=1e-9, maxiter=300)
    assert r["converged"]
    g = p.grid
    geo = po.geometry(p)
    i, j = g.ilo + 1, g.jlo + 1
    H = geo["stagthck"][i, j]
    sx = geo["dusrf_dx"][i, j]
    ub = -910.0 * 9.81 * H * sx / beta
    u = r["uvel3d"][:, i, j]
    want = ub + 2 * A * (910.0 * 9.81 * abs(sx)) ** 3 * H ** 4 * (1 - p.sigma ** 4) / 4
    return float(np.abs(u - want).max() / np.abs(want - ub).max()), abs(u[-1] / ub - 1.0)


def test_l1l2_slab_sia_profile_and_sliding():
    """On the periodic inclined slab the L1L2 solve slides at the closed-form
    u_b = -rho g H s_x / beta (membrane stresses vanish) and the 3-D velocity
    approaches the shallow-ice profile u_b + 2A (rho g |s_x|)^3 H^4 (1 -
    sigma^4)/4 at second order in the level spacing."""
    e1, b1 = _l1l2_slab(11)
    e2, b2 = _l1l2_slab(21)
    assert b1 < 1e-6 and b2 < 1e-6
    assert e2 < 2e-3 and 3.5 < e1 / e2 < 4.5


# ------------------------------------------------- surface gradient margin rules
def _margin_problem():
    """Ice with a planar surface s = s0 + a x + b y, cut by an ice-free stripe
    of three cell columns: land (bed 50 m above sea level) in the south half,
    ocean (bed 500 m below) in the north half.  The ice surface is higher
    than the stripe everywhere."""
    p = synthetic.ice_stream(24, nz=5, seed=1)
    g = p.grid
    p.dx = p.dy = 2000.0
    a, b, s0 = -4e-3, 1e-3, 1400.0
    xc = (np.arange(g.nx) + 0.5) * p.dx
    yc = (np.arange(g.ny) + 0.5) * p.dy
    s = s0 + a * xc[:, None] + b * yc[None, :]
    topg = np.full((g.nx, g.ny), -200.0)
    thck = s - topg
    c1, c2 = 12, 15                         # 0-based stripe columns c1..c2-1
    mid = g.ny // 2
    topg[c1:c2, :mid] = 50.0
    topg[c1:c2, mid:] = -500.0
    thck[c1:c2, :] = 0.0
    p.topg, p.thck = np.asfortranarray(topg), np.asfortranarray(thck)
    p.eus, p.thck_gradient_ramp, p.max_slope = 0.0, 0.0, 10.0
    return p, a, b, (c1, c2, mid)


@pytest.mark.parametrize("margin", [0, 1, 2])
def test_surface_gradient_margin_rules(margin):
    """glissade_surface_elevation_gradient (glissade_grid_operators.F90:834-1260),
    centred: every vertex surrounded by ice has the plane's slope (a, b)
    under every margin rule; at the stripe, an x-edge between ice and an
    ice-free cell counts for margin 0 (all edges), never for margin 2 (ice on
    both sides), and for margin 1 only when the lower cell is land -- so the
    vertex value is the mean of the counted edges' differences, or zero."""
    p, a, b, (c1, c2, mid) = _margin_problem()
    p.gradient_margin = margin
    geo = po.geometry(p)
    g = p.grid
    usrf, dsdx, dsdy = geo["usrf"], geo["dusrf_dx"], geo["dusrf_dy"]
    ice = geo["ice_mask"] == 1
    land = geo["land_mask"] == 1
    assert not ice[c1:c2].any() and ice[c1 - 1, 4] and ice[c2, 4]
    checked = 0
    for i in range(3, g.nx - 4):           # 0-based vertex (i, j): NE corner of cell (i, j)
        for j in range(3, g.ny - 4):
            cells = [(i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)]
            if all(ice[c] for c in cells):
                assert dsdx[i, j] == pytest.approx(a, rel=1e-9) and dsdy[i, j] == pytest.approx(b, rel=1e-9)
                checked += 1
                continue
            # x-edges south (j) and north (j+1) of the vertex
            vals = []
            for jj in (j, j + 1):
                lo, hi = (i, jj), (i + 1, jj)
                up, dn = (lo, hi) if usrf[lo] > usrf[hi] else (hi, lo)
                if margin == 0:
                    keep = True
                elif margin == 2:
                    keep = ice[lo] and ice[hi]
                else:
                    keep = (ice[lo] and ice[hi]) or (ice[up] and land[dn])
                vals.append((usrf[hi] - usrf[lo]) / p.dx if keep else 0.0)
            assert dsdx[i, j] == pytest.approx(0.5 * (vals[0] + vals[1]), rel=1e-12, abs=1e-15), (i, j)
            checked += 1
    assert checked > 200
    # the stripe is seen where it should be: hybrid keeps ice->land edges only
    i = c1 - 1
    if margin == 1:
        assert dsdx[i, 4] < -0.3 and dsdx[i, g.ny - 6] == 0.0
    if margin == 2:
        assert dsdx[i, 4] == 0.0 and dsdx[i, g.ny - 6] == 0.0


def test_symmetry_fix_averages_small_and_keeps_large_asymmetries():
    """check_symmetry_assembled_matrix_3d (glissade_velo_higher.F90:10854-11117):
    a coupling pair that differs by less than 1e-11 of the row's diagonal is
    replaced by its mean on both sides; a larger difference is reported and
    left; every other entry is untouched."""
    p = synthetic.ice_stream(16, nz=5, seed=2)
    s = po.assemble(p)
    g = p.grid
    A = [s[k].copy(order="F") for k in ("Auu", "Auv", "Avu", "Avv")]
    orig = [a.copy(order="F") for a in A]
    av = s["active_vertex"]
    rng = np.random.default_rng(5)
    planted = []
    while len(planted) < 12:
        i, j, k = (int(rng.integers(g.ilo + 1, g.ihi - 1)), int(rng.integers(g.jlo + 1, g.jhi - 1)),
                   int(rng.integers(2, g.nz)))
        m = int(rng.integers(0, 27))
        if m == 13 or not av[i - 1, j - 1] or any(t[:4] == (i, j, k, m) for t in planted):
            continue
        iA, jA, kA = m % 3 - 1, (m // 3) % 3 - 1, m // 9 - 1
        mm = 26 - m
        if any(t[:3] == (i + iA, j + jA, k + kA) and t[3] == mm for t in planted):
            continue
        blk = len(planted) % 2 * 3          # Auu or Avv
        big = len(planted) >= 9
        dg = min(abs(orig[blk][13, k - 1, i - 1, j - 1]), abs(orig[blk][13, k + kA - 1, i + iA - 1, j + jA - 1]))
        delta = (1e-6 if big else 1e-13) * dg
        A[blk][m, k - 1, i - 1, j - 1] += delta
        planted.append((i, j, k, m, blk, big, delta))
    before = [a.copy(order="F") for a in A]
    nwarn = po.lib().orc_check_symmetry_assembled_3d(g.nx, g.ny, g.nz, g.ilo, g.ihi, g.jlo, g.jhi, po.ip(av),
                                                    *[po.dp(a) for a in A])
    assert nwarn >= 3
    touched = [np.zeros(a.shape, bool) for a in A]
    for (i, j, k, m, blk, big, delta) in planted:
        iA, jA, kA = m % 3 - 1, (m // 3) % 3 - 1, m // 9 - 1
        r = (m, k - 1, i - 1, j - 1)
        c = (26 - m, k + kA - 1, i + iA - 1, j + jA - 1)
        v1, v2 = before[blk][r], before[blk][c]
        if big:
            assert A[blk][r] == v1 and A[blk][c] == v2
        else:
            assert A[blk][r] == A[blk][c] == 0.5 * (v1 + v2)
        touched[blk][r] = touched[blk][c] = True
    for a, b, t in zip(A, before, touched):
        assert np.array_equal(a[~t], b[~t])


# ------------------------------------------- standard (quadrature) load and basal term
def test_standard_gravity_load_slab_and_total_force():
    """load_vector_gravity, standard option: on the uniform slab every node
    whose eight elements are all ice gets the local option's load (the
    consistent load of a constant equals the lumped one on boxes); with the
    bed lowered and raised under a planar surface (thickness varying, ds/dx
    uniform) the loads sum to -rho_i g ds/dx V / vol0, V the ice volume of
    the active cells (bilinear thickness, integrated exactly)."""
    p = synthetic.slab(n_cells=8, nz=6)
    g = p.grid
    geo = po.geometry(p)
    xV, yV = p.vertex_coords()
    lu_s, lv_s, lu_l, lv_l = g.zeros_vec(), g.zeros_vec(), g.zeros_vec(), g.zeros_vec()
    p.assemble_taud = 0
    po.gravity_load(p, geo, xV, yV, lu_s, lv_s)
    p.assemble_taud = 1
    po.gravity_load(p, geo, xV, yV, lu_l, lv_l)
    inner = (slice(None), slice(3, g.nx - 3), slice(3, g.ny - 3))
    assert np.abs(lu_s[inner] - lu_l[inner]).max() <= 1e-12 * np.abs(lu_l[inner]).max()
    rng = np.random.default_rng(3)
    dh = rng.uniform(-150.0, 150.0, (g.nx, g.ny))
    p.topg = np.asfortranarray(p.topg + dh)
    p.thck = np.asfortranarray(p.thck - dh)
    geo = po.geometry(p)
    sx = geo["dusrf_dx"]
    assert np.ptp(sx) <= 1e-12 * abs(sx).max() and np.abs(geo["dusrf_dy"]).max() <= 1e-15
    lu, lv = g.zeros_vec(), g.zeros_vec()
    p.assemble_taud = 0
    po.gravity_load(p, geo, xV, yV, lu, lv)
    H = geo["stagthck"]
    vol = 0.0
    for j in range(3, g.ny):          # 1-based cells nhalo+1 .. ny-nhalo+1
        for i in range(3, g.nx):
            if geo["active_cell"][i - 1, j - 1]:
                vol += p.dx * p.dy * (H[i - 2, j - 2] + H[i - 1, j - 2] + H[i - 1, j - 1] + H[i - 2, j - 1]) / 4.0
    want = -910.0 * 9.81 * sx.mean() * vol / 1e9
    assert lu.sum() == pytest.approx(want, rel=1e-12)
    assert np.abs(lv).max() <= 1e-12 * np.abs(lu).max()


def _basal_rows(p, beta_opt, two_d=False):
    p.assemble_beta = beta_opt
    s = po.assemble_2d(p) if two_d else po.assemble(p)
    return s


@pytest.mark.parametrize("two_d", [False, True])
def test_standard_basal_term_row_sums_on_planar_bed(two_d):
    """basal_sliding_bc_3d/_2d, standard option, on the periodic slab (planar
    inclined bed, uniform beta): the consistent friction rows sum to the
    lumped local term -- in 3-D beta dx dy / (vol0 cos theta), the curved
    face's area factor being exactly 1/cos theta on a plane; in 2-D
    beta dx dy / vol0 (flat Jacobian) -- the rows above the bed are
    unchanged and the operator stays symmetric."""
    p = synthetic.slab(n_cells=8, nz=6)
    g = p.grid
    loc = _basal_rows(p, 1, two_d)
    std = _basal_rows(p, 0, two_d)
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    for key in ("Auu", "Avv"):
        if two_d:
            a, b = loc[key][own[0], own[1], :], std[key][own[0], own[1], :]
            assert np.abs(a.sum(axis=-1) - b.sum(axis=-1)).max() <= 1e-12 * np.abs(a[..., 4]).max()
            assert np.abs(b[..., [0, 1, 2, 3, 5, 6, 7, 8]]).max() > 0.0
        else:
            a, b = loc[key][:, :, own[0], own[1]], std[key][:, :, own[0], own[1]]
            assert np.abs(a[:, -1].sum(axis=0) - b[:, -1].sum(axis=0)).max() <= 1e-12 * np.abs(a[13, -1]).max()
            assert np.array_equal(a[:, :-1], b[:, :-1])
    assert std["nwarn_asym"] == 0


@pytest.mark.parametrize("make", [lambda: synthetic.ice_stream(40, nz=11, seed=3), lambda: synthetic.dome()])
def test_accelerated_diva_picard_same_fixed_point(make):
    """Accelerated Picard with DIVA (basal traction recomputed after each
    extension; beta_eff and omega_k kept with the saved assembly) reaches the
    plain DIVA Picard fixed point in fewer iterations."""
    p = make()
    kw = dict(linear_tol=1e-10, linear_maxiters=3000, resid_target=1e-6, maxiter=100)
    a = po.picard(p, two_d=True, approx="diva", precond=1, **kw)
    b = po.picard(p, two_d=True, approx="diva", precond=1, accelerate=True, **kw)
    assert a["converged"] and b["converged"] and b["iterations"] < a["iterations"]
    num = np.linalg.norm(np.stack([a["uvel"] - b["uvel"], a["vvel"] - b["vvel"]]))
    assert num / np.linalg.norm(np.stack([a["uvel"], a["vvel"]])) < 1e-5
    assert max(b["alpha"]) > 1.0


def test_standard_diva_basal_term_on_planar_bed():
    """basal_sliding_bc_2d_diva, standard option, on the periodic slab (bed
    inclined in x only): the consistent rows of Auu sum to the local
    beta_eff_x dx dy / (vol0 cos theta_x); those of Avv carry the face's area
    factor 1/cos theta where the local option divides by cos theta_y = 1."""
    p = synthetic.slab(n_cells=8, nz=6)
    g = p.grid
    out = {}
    for opt in (1, 0):
        p.assemble_beta = opt
        out[opt] = po.assemble_2d(p, approx="diva")
    p.assemble_beta = 1
    loc, std = out[1], out[0]
    own = (slice(g.ilo - 1, g.ihi), slice(g.jlo - 1, g.jhi))
    th = loc["geo"]["diva"]["theta_basal_x"][own]
    assert np.ptp(th) < 1e-12 and th.mean() > 1e-3
    au_l, au_s = loc["Auu"][own[0], own[1], :], std["Auu"][own[0], own[1], :]
    av_l, av_s = loc["Avv"][own[0], own[1], :], std["Avv"][own[0], own[1], :]
    assert np.abs(au_s.sum(axis=-1) - au_l.sum(axis=-1)).max() <= 1e-11 * np.abs(au_l[..., 4]).max()
    # the stiffness rows sum to ~0 on the uniform slab, so the row sums are the basal terms
    ratio = av_s.sum(axis=-1) / av_l.sum(axis=-1)
    assert np.allclose(ratio, 1.0 / np.cos(th.mean()), rtol=1e-9)
