/*
 * oracle_pcg.c -- CPU restatement of the structured 3-D solvers of
 * libglissade/glissade_velo_higher_pcg.F90 and of the one-rank semantics of
 * the parallel layer they call (libglimmer/parallel_mpi.F90,
 * libglimmer/cism_reprosum_mod.F90).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP product).  Loop orders,
 * summation orders and whole-array update ranges follow the reference
 * line by line; each function cites the lines it restates.  Built with
 * -ffp-contract=off so that every product and sum is rounded separately,
 * as in an x86-64 Fortran build without FMA contraction.
 */
#include "cism_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NVX (g->nx - 1)
#define NVY (g->ny - 1)
#define NZ  (g->nz)
/* 1-based (k,i,j) -> 0-based offset into an (nz, nx-1, ny-1) array */
#define V3(k, i, j) ((size_t)((k) - 1) + (size_t)NZ * ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1)))
#define V2(i, j)    ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1))
#define A4(m, k, i, j) ((size_t)((m) - 1) + 27 * V3(k, i, j))
#define IDX3(iA, jA, kA) (1 + ((iA) + 1) + 3 * ((jA) + 1) + 9 * ((kA) + 1))
#define IDX2(iA, jA) (1 + ((iA) + 1) + 3 * ((jA) + 1))

enum { PC_NONE = 0, PC_DIAG = 1, PC_SIA = 2, PC_TRIDIAG_LOCAL = 3, PC_TRIDIAG_GLOBAL = 4 };

/* ------------------------------------------------------------------ */
/* trace of the per-iteration scalars                                   */
/* ------------------------------------------------------------------ */
static double* g_trace = NULL;
static int32_t g_trace_cap = 0, g_trace_len = 0;
void orc_set_trace(double* buf, int32_t cap) { g_trace = buf; g_trace_cap = cap; g_trace_len = 0; }
int32_t orc_trace_len(void) { return g_trace_len; }
static void trace3(double a, double b, double c) {
    if (g_trace && g_trace_len < g_trace_cap) {
        g_trace[3 * g_trace_len + 0] = a;
        g_trace[3 * g_trace_len + 1] = b;
        g_trace[3 * g_trace_len + 2] = c;
        g_trace_len++;
    }
}

/* ------------------------------------------------------------------ */
/* staggered_parallel_halo_real8_3d/2d for a single rank               */
/* parallel_mpi.F90:10772-10936 (3-D), :10609-10769 (2-D).             */
/* staggered_lhalo = 2, staggered_uhalo = 1 (parallel_mpi.F90:66-67).  */
/* With one task every neighbour is this rank: periodic (and           */
/* no_penetration, which uses the periodic logic, :2499) exchanges     */
/* with itself; otherwise no message is sent and only the global-      */
/* boundary zeroing of :10886-10932 applies.                           */
/* ------------------------------------------------------------------ */
void orc_halo3d(const orc_grid* g, int nlev, double* a)
{
    const int nvx = NVX, nvy = NVY;
    const int lh = 2, uh = 1;
#define H(k, i, j) a[(size_t)((k) - 1) + (size_t)nlev * ((size_t)((i) - 1) + (size_t)nvx * (size_t)((j) - 1))]
    const int periodic = (g->global_bc == 0 || g->global_bc == 2);
    if (periodic) {
        /* E/W over owned rows jlo..jhi (:10820-10860).  The west-going
         * message a(:,1+lh:lh+uh,:) lands in the east halo; the east-going
         * message a(:,nvx-uh-lh+1:nvx-uh,:) lands in the west halo. */
        for (int j = g->jlo; j <= g->jhi; ++j)
            for (int k = 1; k <= nlev; ++k) {
                double wsend[1], esend[2];
                for (int h = 0; h < uh; ++h) wsend[h] = H(k, 1 + lh + h, j);
                for (int h = 0; h < lh; ++h) esend[h] = H(k, nvx - uh - lh + 1 + h, j);
                for (int h = 0; h < uh; ++h) H(k, nvx - uh + 1 + h, j) = wsend[h];
                for (int h = 0; h < lh; ++h) H(k, 1 + h, j) = esend[h];
            }
        /* N/S over all columns, after the E/W update (:10862-10884). */
        for (int i = 1; i <= nvx; ++i)
            for (int k = 1; k <= nlev; ++k) {
                double ssend[1], nsend[2];
                for (int h = 0; h < uh; ++h) ssend[h] = H(k, i, 1 + lh + h);
                for (int h = 0; h < lh; ++h) nsend[h] = H(k, i, nvy - uh - lh + 1 + h);
                for (int h = 0; h < uh; ++h) H(k, i, nvy - uh + 1 + h) = ssend[h];
                for (int h = 0; h < lh; ++h) H(k, i, 1 + h) = nsend[h];
            }
    } else if (g->global_bc == 1) {   /* outflow (:10886-10899) */
        for (int j = 1; j <= nvy; ++j)
            for (int k = 1; k <= nlev; ++k) {
                for (int i = 1; i <= lh - 1; ++i) H(k, i, j) = 0.0;
                for (int i = nvx - uh + 1; i <= nvx; ++i) H(k, i, j) = 0.0;
            }
        for (int j = 1; j <= nvy; ++j) {
            if (!(j <= lh - 1 || j >= nvy - uh + 1)) continue;
            for (int i = 1; i <= nvx; ++i)
                for (int k = 1; k <= nlev; ++k) H(k, i, j) = 0.0;
        }
    } else if (g->global_bc == 3) {   /* no_ice (:10900-10932): includes the last owned E/N row */
        for (int j = 1; j <= nvy; ++j)
            for (int k = 1; k <= nlev; ++k) {
                for (int i = 1; i <= lh; ++i) H(k, i, j) = 0.0;
                for (int i = nvx - uh; i <= nvx; ++i) H(k, i, j) = 0.0;
            }
        for (int j = 1; j <= nvy; ++j) {
            if (!(j <= lh || j >= nvy - uh)) continue;
            for (int i = 1; i <= nvx; ++i)
                for (int k = 1; k <= nlev; ++k) H(k, i, j) = 0.0;
        }
    }
#undef H
}

/* ------------------------------------------------------------------ */
/* DDPDD compensated accumulation: cism_reprosum_mod.F90:1952-1962     */
/* (Knuth TwoSum into a double-double), used when reprosum is on.       */
/* ------------------------------------------------------------------ */
static inline void ddpdd_add(double x, double* hi, double* lo)
{
    double t1 = x + *hi;
    double e = t1 - x;
    double t2 = ((*hi - e) + (x - (t1 - e))) + *lo;
    *hi = t1 + t2;
    *lo = t2 - ((t1 + t2) - t1);
}

void orc_ddpdd_sum(int64_t n, const double* arr, double* hi, double* lo)
{
    double h = 0.0, l = 0.0;
    for (int64_t s = 0; s < n; ++s) ddpdd_add(arr[s], &h, &l);
    *hi = h; *lo = l;
}

/* ------------------------------------------------------------------ */
/* cism_reprosum_calc, integer-vector algorithm (cism_reprosum_mod.F90: */
/* 237-1057 and :1063-1818), for one task and one field.  As in the     */
/* reference, the extreme exponents of the nonzero summands fix the     */
/* levels: level 0 holds the lowest bit any summand has, each level     */
/* takes RV_D bits of every summand's mantissa, and the levels are      */
/* 128-bit integers, so no level can overflow.  The integer vector is   */
/* then carried into digits of one sign and rounded once to the nearest */
/* double (ties to even): the exact sum, correctly rounded.            */
/* ------------------------------------------------------------------ */
#define RV_D 56
typedef __int128 i128;

double orc_reprosum_calc(int64_t n, const double* arr)
{
    int emin = 1 << 30, emax = -(1 << 30), bad = 0;
    for (int64_t s = 0; s < n; ++s) {
        const double x = arr[s];
        if (!isfinite(x)) { bad = 1; continue; }
        if (x == 0.0) continue;
        int e;
        frexp(x, &e);                       /* |x| = f 2^e, f in [0.5, 1) */
        int lsb = e - 53;                   /* weight of the last mantissa bit */
        if (lsb < -1074) lsb = -1074;       /* subnormals */
        if (lsb < emin) emin = lsb;
        if (e > emax) emax = e;
    }
    if (bad) return NAN;
    if (emin > emax) return 0.0;
    const int nlev = (emax - emin) / RV_D + 2;
    i128* lev = (i128*)calloc((size_t)nlev, sizeof(i128));
    for (int64_t s = 0; s < n; ++s) {
        const double x = arr[s];
        if (x == 0.0) continue;
        int e;
        const double f = frexp(fabs(x), &e);
        int lsb = e - 53;
        int64_t m;
        if (lsb < -1074) { m = (int64_t)ldexp(f, e + 1074); lsb = -1074; }
        else m = (int64_t)ldexp(f, 53);     /* exact: 53-bit integer mantissa */
        const int off = lsb - emin;         /* bit offset of m above level 0 */
        const int l = off / RV_D, sh = off % RV_D;
        const i128 v = (i128)m << sh;       /* < 2^(53 + 55) */
        const i128 mask = (((i128)1) << RV_D) - 1;
        const i128 sg = x < 0 ? -1 : 1;
        lev[l] += sg * (v & mask);
        lev[l + 1] += sg * (v >> RV_D);
    }
    /* sign of the total, then digits of that sign */
    i128 carry = 0;
    for (int l = 0; l < nlev; ++l) carry = (lev[l] + carry) >> RV_D;
    const int neg = carry < 0;
    if (neg)
        for (int l = 0; l < nlev; ++l) lev[l] = -lev[l];
    carry = 0;
    for (int l = 0; l < nlev; ++l) {
        const i128 t = lev[l] + carry;
        lev[l] = t & ((((i128)1) << RV_D) - 1);
        carry = t >> RV_D;
    }
    int h = nlev - 1;
    while (h >= 0 && lev[h] == 0) --h;
    double r = 0.0;
    if (carry > 0) {
        r = INFINITY;
    } else if (h >= 0) {
        /* the top two digits hold at least 57 bits once the top one is nonzero */
        const int lo = h >= 1 ? h - 1 : 0;
        unsigned __int128 T = 0;
        for (int l = h; l >= lo; --l) T = (T << RV_D) | (unsigned __int128)lev[l];
        int sticky = 0;
        for (int l = 0; l < lo; ++l) sticky |= lev[l] != 0;
        int nb = 0;
        for (unsigned __int128 u = T; u; u >>= 1) ++nb;
        const int ebase = emin + RV_D * lo;
        if (nb <= 53) {
            r = ldexp((double)(uint64_t)T, ebase);
        } else {
            const int sh = nb - 53;
            uint64_t q = (uint64_t)(T >> sh);
            const unsigned __int128 rem = T - ((unsigned __int128)q << sh);
            const unsigned __int128 half = ((unsigned __int128)1) << (sh - 1);
            if (rem > half || (rem == half && (sticky || (q & 1)))) q += 1;
            r = ldexp((double)q, ebase + sh);
        }
    }
    free(lev);
    return neg ? -r : r;
}

/* ------------------------------------------------------------------ */
/* parallel_global_sum_stagger_real8_3d (parallel_mpi.F90:6663-6794):  */
/* non-reproducible order local = (local + a1) + a2, loops j, i, k over */
/* owned vertices; reproducible variant packs a1+a2 (:6718-6731) and   */
/* sums the packed array with cism_reprosum_calc (integer vectors).    */
/* ------------------------------------------------------------------ */
double orc_global_sum(const orc_grid* g, const double* a1, const double* a2)
{
    if (g->reprosum) {
        const int64_t n = (int64_t)(g->ihi - g->ilo + 1) * (g->jhi - g->jlo + 1) * NZ;
        double* arr = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        int64_t c = 0;
        for (int j = g->jlo; j <= g->jhi; ++j)
            for (int i = g->ilo; i <= g->ihi; ++i)
                for (int k = 1; k <= NZ; ++k)
                    arr[c++] = a2 ? a1[V3(k, i, j)] + a2[V3(k, i, j)] : a1[V3(k, i, j)];
        const double r = orc_reprosum_calc(n, arr);
        free(arr);
        return r;
    }
    double s = 0.0;
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i)
            for (int k = 1; k <= NZ; ++k) {
                if (a2) s = s + a1[V3(k, i, j)] + a2[V3(k, i, j)];
                else    s = s + a1[V3(k, i, j)];
            }
    return s;
}

/* parallel_global_sum_stagger_real8_3d_nflds (parallel_mpi.F90:6928-7065),
 * nflds = 2, field loop outermost (:7031-7061). */
void orc_global_sum2(const orc_grid* g, const double* a1f1, const double* a2f1,
                     const double* a1f2, const double* a2f2, double out[2])
{
    out[0] = orc_global_sum(g, a1f1, a2f1);
    out[1] = orc_global_sum(g, a1f2, a2f2);
}

/* ------------------------------------------------------------------ */
/* matvec_multiply_structured_3d (pcg.F90:5528-5652)                   */
/* ------------------------------------------------------------------ */
void orc_matvec_3d(const orc_grid* g, const int32_t* active_vertex,
                   const double* Auu, const double* Auv, const double* Avu, const double* Avv,
                   const double* xu, const double* xv, double* yu, double* yv)
{
    const size_t n = (size_t)NZ * NVX * NVY;
    memset(yu, 0, n * sizeof(double));
    memset(yv, 0, n * sizeof(double));
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            for (int k = 1; k <= NZ; ++k) {
                double su = yu[V3(k, i, j)], sv = yv[V3(k, i, j)];
                for (int kA = -1; kA <= 1; ++kA)
                    for (int jA = -1; jA <= 1; ++jA)
                        for (int iA = -1; iA <= 1; ++iA) {
                            if (k + kA >= 1 && k + kA <= NZ && i + iA >= 1 && i + iA <= NVX &&
                                j + jA >= 1 && j + jA <= NVY) {
                                const int m = IDX3(iA, jA, kA);
                                const size_t a = A4(m, k, i, j);
                                const size_t x = V3(k + kA, i + iA, j + jA);
                                su = su + Auu[a] * xu[x] + Auv[a] * xv[x];
                                sv = sv + Avu[a] * xu[x] + Avv[a] * xv[x];
                            }
                        }
                yu[V3(k, i, j)] = su;
                yv[V3(k, i, j)] = sv;
            }
        }
}

/* ------------------------------------------------------------------ */
/* tridiag_solver (pcg.F90:5870-5922): Thomas algorithm                */
/* ------------------------------------------------------------------ */
void orc_tridiag_solver(int order, const double* sbdiag, const double* diag,
                        const double* spdiag, const double* rhs, double* soln)
{
    double gamma[64];
    double beta = diag[0];
    soln[0] = rhs[0] / beta;
    for (int k = 2; k <= order; ++k) {
        gamma[k - 1] = spdiag[k - 2] / beta;
        beta = diag[k - 1] - sbdiag[k - 1] * gamma[k - 1];
        soln[k - 1] = (rhs[k - 1] - sbdiag[k - 1] * soln[k - 2]) / beta;
    }
    for (int k = order - 1; k >= 1; --k) soln[k - 1] = soln[k - 1] - gamma[k] * soln[k];
}

/* easy_sia_solver (pcg.F90:5777-5866): per active column, inactive
 * columns left untouched.  M is (-1:1, nz, nx-1, ny-1). */
static void easy_sia_solver(const orc_grid* g, const int32_t* active_vertex,
                            const double* M, const double* b, double* x)
{
    double sb[64], d[64], sp[64], rhs[64], soln[64];
    const int nz = NZ;
    for (int j = 1; j <= NVY; ++j)
        for (int i = 1; i <= NVX; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            const size_t c = V3(1, i, j);
            for (int k = 0; k < nz; ++k) { rhs[k] = b[c + k]; soln[k] = x[c + k]; }
            for (int k = 1; k <= nz; ++k) {
                const double* mk = M + 3 * (c + (size_t)(k - 1));
                sb[k - 1] = (k == 1) ? 0.0 : mk[0];
                d[k - 1] = mk[1];
                sp[k - 1] = (k == nz) ? 0.0 : mk[2];
            }
            orc_tridiag_solver(nz, sb, d, sp, rhs, soln);
            for (int k = 0; k < nz; ++k) x[c + k] = soln[k];
        }
}

/* setup_preconditioner_sia_3d (pcg.F90:3647-3733) */
static void setup_sia(const orc_grid* g, const double* Auu, const double* Avv, double* Muu, double* Mvv)
{
    const size_t n = (size_t)NZ * NVX * NVY;
    for (size_t p = 0; p < n; ++p) {
        Muu[3 * p + 0] = Auu[27 * p + IDX3(0, 0, -1) - 1];
        Mvv[3 * p + 0] = Avv[27 * p + IDX3(0, 0, -1) - 1];
        Muu[3 * p + 1] = Auu[27 * p + IDX3(0, 0, 0) - 1];
        Mvv[3 * p + 1] = Avv[27 * p + IDX3(0, 0, 0) - 1];
        Muu[3 * p + 2] = Auu[27 * p + IDX3(0, 0, 1) - 1];
        Mvv[3 * p + 2] = Avv[27 * p + IDX3(0, 0, 1) - 1];
    }
}

/* ------------------------------------------------------------------ */
/* Horizontal tridiagonal preconditioners (precond 3 and one-task 4)   */
/* ------------------------------------------------------------------ */
typedef struct tri2d {
    /* vertex-grid arrays (nx-1, ny-1); for the v system the roles of i and j
     * are swapped inside the solve, as in the reference. */
    double *Adiag_u, *Asub_u, *Asup_u, *omega_u, *denom_u;
    double *Adiag_v, *Asub_v, *Asup_v, *omega_v, *denom_v;
} tri2d;

static void tri2d_alloc(const orc_grid* g, tri2d* t)
{
    const size_t n2 = (size_t)NVX * NVY;
    double* blk = (double*)calloc(10 * n2, sizeof(double));
    t->Adiag_u = blk + 0 * n2; t->Asub_u = blk + 1 * n2; t->Asup_u = blk + 2 * n2;
    t->omega_u = blk + 3 * n2; t->denom_u = blk + 4 * n2;
    t->Adiag_v = blk + 5 * n2; t->Asub_v = blk + 6 * n2; t->Asup_v = blk + 7 * n2;
    t->omega_v = blk + 8 * n2; t->denom_v = blk + 9 * n2;
}
static void tri2d_free(tri2d* t) { free(t->Adiag_u); }

/* vertical sums of the matrix into the 9-point 2-D stencil
 * (pcg.F90:3815-3840 local; :3996-4021 global) */
static void vsum_matrix(const orc_grid* g, const int32_t* active_vertex, const double* A,
                        double* vs /* (nx-1,ny-1,9) */, int zero_offdiag_inactive)
{
    const size_t n2 = (size_t)NVX * NVY;
    for (int j = 1; j <= NVY; ++j)
        for (int i = 1; i <= NVX; ++i)
            for (int m = 1; m <= 9; ++m) {
                double s = 0.0;
                for (int k = 1; k <= NZ; ++k)
                    s = s + A[A4(m, k, i, j)] + A[A4(m + 9, k, i, j)] + A[A4(m + 18, k, i, j)];
                vs[V2(i, j) + n2 * (m - 1)] = s;
            }
    if (zero_offdiag_inactive < 0) return;   /* 2-D solves take the stencil as given */
    for (int j = 1; j <= NVY; ++j)
        for (int i = 1; i <= NVX; ++i)
            if (!active_vertex[V2(i, j)]) {
                if (zero_offdiag_inactive)
                    for (int m = 1; m <= 9; ++m) vs[V2(i, j) + n2 * (m - 1)] = 0.0;
                vs[V2(i, j) + n2 * 4] = 1.0;
            }
}

/* setup_preconditioner_tridiag_local_2d (pcg.F90:4247-4392) */
static int setup_tridiag_local_2d(const orc_grid* g, const double* vsu, const double* vsv, tri2d* t)
{
    const size_t n2 = (size_t)NVX * NVY;
    for (size_t p = 0; p < n2; ++p) {
        t->Asub_u[p] = vsu[p + n2 * (IDX2(-1, 0) - 1)];
        t->Adiag_u[p] = vsu[p + n2 * (IDX2(0, 0) - 1)];
        t->Asup_u[p] = vsu[p + n2 * (IDX2(1, 0) - 1)];
        if (t->Adiag_u[p] == 0.0) { t->Adiag_u[p] = 1.0; t->Asub_u[p] = 0.0; t->Asup_u[p] = 0.0; }
        t->omega_u[p] = 0.0; t->denom_u[p] = 0.0;
    }
    for (int j = g->jlo; j <= g->jhi; ++j) {
        int i = g->ilo;
        t->omega_u[V2(i, j)] = t->Asup_u[V2(i, j)] / t->Adiag_u[V2(i, j)];
        for (i = g->ilo + 1; i <= g->ihi; ++i) {
            double d = t->Adiag_u[V2(i, j)] - t->Asub_u[V2(i, j)] * t->omega_u[V2(i - 1, j)];
            if (d == 0.0) return ORC_ERR_DIVZERO;
            t->denom_u[V2(i, j)] = 1.0 / d;
            t->omega_u[V2(i, j)] = t->Asup_u[V2(i, j)] * t->denom_u[V2(i, j)];
        }
    }
    for (size_t p = 0; p < n2; ++p) t->Adiag_u[p] = 1.0 / t->Adiag_u[p];

    for (size_t p = 0; p < n2; ++p) {
        t->Asub_v[p] = vsv[p + n2 * (IDX2(0, -1) - 1)];
        t->Adiag_v[p] = vsv[p + n2 * (IDX2(0, 0) - 1)];
        t->Asup_v[p] = vsv[p + n2 * (IDX2(0, 1) - 1)];
        if (t->Adiag_v[p] == 0.0) { t->Adiag_v[p] = 1.0; t->Asub_v[p] = 0.0; t->Asup_v[p] = 0.0; }
        t->omega_v[p] = 0.0; t->denom_v[p] = 0.0;
    }
    for (int i = g->ilo; i <= g->ihi; ++i) {
        int j = g->jlo;
        t->omega_v[V2(i, j)] = t->Asup_v[V2(i, j)] / t->Adiag_v[V2(i, j)];
        for (j = g->jlo + 1; j <= g->jhi; ++j) {
            double d = t->Adiag_v[V2(i, j)] - t->Asub_v[V2(i, j)] * t->omega_v[V2(i, j - 1)];
            if (d == 0.0) return ORC_ERR_DIVZERO;
            t->denom_v[V2(i, j)] = 1.0 / d;
            t->omega_v[V2(i, j)] = t->Asup_v[V2(i, j)] * t->denom_v[V2(i, j)];
        }
    }
    for (size_t p = 0; p < n2; ++p) t->Adiag_v[p] = 1.0 / t->Adiag_v[p];
    return ORC_OK;
}

/* setup_preconditioner_tridiag_global_2d (pcg.F90:4396-4520) with one task
 * per row/column: only the forward-elimination factors of the owned block
 * matter, because with tasks_rc = 1 the solve returns the particular
 * solution xr alone (pcg.F90:5495-5497).  The homogeneous-solution factors
 * (xuh, xlh) are still formed so that their zero-pivot check (:4474-4493)
 * fires exactly where the reference's does. */
static int setup_tridiag_global_owned(int ilocal, int jlocal, double* Adiag, double* Asub, double* Asup,
                                      double* omega, double* denom)
{
    const size_t n = (size_t)ilocal * jlocal;
    double* omega_uh = (double*)calloc((size_t)ilocal, sizeof(double));
    for (size_t p = 0; p < n; ++p) {
        if (Adiag[p] == 0.0) { Adiag[p] = 1.0; Asub[p] = 0.0; Asup[p] = 0.0; }
        omega[p] = 0.0; denom[p] = 0.0;
    }
#define L(i, j) ((size_t)((i) - 1) + (size_t)ilocal * (size_t)((j) - 1))
    for (int j = 1; j <= jlocal; ++j) {
        omega[L(1, j)] = Asup[L(1, j)] / Adiag[L(1, j)];
        for (int i = 2; i <= ilocal; ++i) {
            double d = Adiag[L(i, j)] - Asub[L(i, j)] * omega[L(i - 1, j)];
            if (d == 0.0) { free(omega_uh); return ORC_ERR_DIVZERO; }
            denom[L(i, j)] = 1.0 / d;
            omega[L(i, j)] = Asup[L(i, j)] * denom[L(i, j)];
        }
        omega_uh[ilocal - 1] = Asub[L(ilocal, j)] / Adiag[L(ilocal, j)];
        for (int i = ilocal - 1; i >= 1; --i) {
            double duh = Adiag[L(i, j)] - Asup[L(i, j)] * omega_uh[i];
            if (duh == 0.0) { free(omega_uh); return ORC_ERR_DIVZERO; }
            omega_uh[i - 1] = Asub[L(i, j)] / duh;
        }
    }
    for (size_t p = 0; p < n; ++p) Adiag[p] = 1.0 / Adiag[p];
#undef L
    free(omega_uh);
    return ORC_OK;
}

typedef struct pcstate {
    int precond;
    int two_d;                    /* 2-D solve: tridiagonal preconditioners without the SIA column part */
    double *Adiagu, *Adiagv;      /* precond 1 */
    double *Muu, *Mvv;            /* precond 2, 3, 4 */
    tri2d t;                      /* precond 3 */
    /* precond 4, owned-block arrays; the v arrays are (jlocal, ilocal) */
    int ilocal, jlocal;
    double *gu_diag, *gu_sub, *gu_sup, *gu_omega, *gu_denom;
    double *gv_diag, *gv_sub, *gv_sup, *gv_omega, *gv_denom;
    double *zu_vs, *zv_vs, *ru_vs, *rv_vs;   /* 2-D work */
} pcstate;

static int precond_setup(const orc_grid* g, const int32_t* active, const double* Auu, const double* Avv,
                         int precond, int two_d, pcstate* s)
{
    memset(s, 0, sizeof(*s));
    s->precond = precond;
    s->two_d = two_d;
    const size_t n = (size_t)NZ * NVX * NVY, n2 = (size_t)NVX * NVY;
    if (precond == PC_DIAG) {   /* setup_preconditioner_diag_3d (pcg.F90:3598-3643) */
        s->Adiagu = (double*)malloc(n * sizeof(double));
        s->Adiagv = (double*)malloc(n * sizeof(double));
        for (size_t p = 0; p < n; ++p) {
            s->Adiagu[p] = Auu[27 * p + IDX3(0, 0, 0) - 1];
            s->Adiagv[p] = Avv[27 * p + IDX3(0, 0, 0) - 1];
        }
    } else if (precond >= PC_SIA && precond <= PC_TRIDIAG_GLOBAL) {
        s->Muu = (double*)malloc(3 * n * sizeof(double));
        s->Mvv = (double*)malloc(3 * n * sizeof(double));
        s->zu_vs = (double*)calloc(4 * n2, sizeof(double));
        s->zv_vs = s->zu_vs + n2; s->ru_vs = s->zu_vs + 2 * n2; s->rv_vs = s->zu_vs + 3 * n2;
        if (precond == PC_TRIDIAG_LOCAL) {
            /* setup_preconditioner_tridiag_local_3d (pcg.F90:3737-3926) */
            double* vsu = (double*)malloc(9 * n2 * sizeof(double));
            double* vsv = (double*)malloc(9 * n2 * sizeof(double));
            vsum_matrix(g, active, Auu, vsu, two_d ? -1 : 1);
            vsum_matrix(g, active, Avv, vsv, two_d ? -1 : 1);
            tri2d_alloc(g, &s->t);
            int st = setup_tridiag_local_2d(g, vsu, vsv, &s->t);
            free(vsu); free(vsv);
            if (st) return st;
        } else if (precond == PC_TRIDIAG_GLOBAL) {
            /* setup_preconditioner_tridiag_global_3d (pcg.F90:3930-4195), one task */
            double* vsu = (double*)malloc(9 * n2 * sizeof(double));
            double* vsv = (double*)malloc(9 * n2 * sizeof(double));
            vsum_matrix(g, active, Auu, vsu, two_d ? -1 : 0);
            vsum_matrix(g, active, Avv, vsv, two_d ? -1 : 0);
            for (int m = 0; m < 9; ++m) { orc_halo3d(g, 1, vsu + n2 * m); orc_halo3d(g, 1, vsv + n2 * m); }
            const int il = g->ihi - g->ilo + 1, jl = g->jhi - g->jlo + 1;
            s->ilocal = il; s->jlocal = jl;
            const size_t nl = (size_t)il * jl;
            double* blk = (double*)calloc(10 * nl, sizeof(double));
            s->gu_diag = blk; s->gu_sub = blk + nl; s->gu_sup = blk + 2 * nl; s->gu_omega = blk + 3 * nl; s->gu_denom = blk + 4 * nl;
            s->gv_diag = blk + 5 * nl; s->gv_sub = blk + 6 * nl; s->gv_sup = blk + 7 * nl; s->gv_omega = blk + 8 * nl; s->gv_denom = blk + 9 * nl;
            for (int j = 1; j <= jl; ++j)
                for (int i = 1; i <= il; ++i) {
                    const size_t src = V2(i + g->ilo - 1, j + g->jlo - 1), dst = (size_t)(i - 1) + (size_t)il * (j - 1);
                    s->gu_sub[dst] = vsu[src + n2 * (IDX2(-1, 0) - 1)];
                    s->gu_diag[dst] = vsu[src + n2 * (IDX2(0, 0) - 1)];
                    s->gu_sup[dst] = vsu[src + n2 * (IDX2(1, 0) - 1)];
                    const size_t dv = (size_t)(j - 1) + (size_t)jl * (i - 1);
                    s->gv_sub[dv] = vsv[src + n2 * (IDX2(0, -1) - 1)];
                    s->gv_diag[dv] = vsv[src + n2 * (IDX2(0, 0) - 1)];
                    s->gv_sup[dv] = vsv[src + n2 * (IDX2(0, 1) - 1)];
                }
            free(vsu); free(vsv);
            int st = setup_tridiag_global_owned(il, jl, s->gu_diag, s->gu_sub, s->gu_sup, s->gu_omega, s->gu_denom);
            if (st) return st;
            st = setup_tridiag_global_owned(jl, il, s->gv_diag, s->gv_sub, s->gv_sup, s->gv_omega, s->gv_denom);
            if (st) return st;
        }
        setup_sia(g, Auu, Avv, s->Muu, s->Mvv);
    }
    return ORC_OK;
}

static void precond_free(pcstate* s)
{
    free(s->Adiagu); free(s->Adiagv); free(s->Muu); free(s->Mvv); free(s->zu_vs);
    if (s->precond == PC_TRIDIAG_LOCAL) tri2d_free(&s->t);
    free(s->gu_diag);
}

/* tridiag_solver_local_2d (pcg.F90:4917-5100): row solves for u along i,
 * column solves for v along j over the owned block; back-substitution
 * continues down to index 1 (:5023, :5074). */
static void tridiag_local_2d_solve(const orc_grid* g, const tri2d* t, const double* bu, const double* bv,
                                   double* xu, double* xv)
{
    const size_t n2 = (size_t)NVX * NVY;
    double* gam = (double*)calloc(n2, sizeof(double));
    memset(xu, 0, n2 * sizeof(double));
    for (int j = g->jlo; j <= g->jhi; ++j) {
        int i = g->ilo;
        gam[V2(i, j)] = bu[V2(i, j)] * t->Adiag_u[V2(i, j)];
        for (i = g->ilo + 1; i <= g->ihi; ++i)
            gam[V2(i, j)] = (bu[V2(i, j)] - t->Asub_u[V2(i, j)] * gam[V2(i - 1, j)]) * t->denom_u[V2(i, j)];
        i = g->ihi;
        xu[V2(i, j)] = gam[V2(i, j)];
        for (i = g->ihi - 1; i >= 1; --i) xu[V2(i, j)] = gam[V2(i, j)] - t->omega_u[V2(i, j)] * xu[V2(i + 1, j)];
    }
    memset(xv, 0, n2 * sizeof(double));
    memset(gam, 0, n2 * sizeof(double));
    for (int i = g->ilo; i <= g->ihi; ++i) {
        int j = g->jlo;
        gam[V2(i, j)] = bv[V2(i, j)] * t->Adiag_v[V2(i, j)];
        for (j = g->jlo + 1; j <= g->jhi; ++j)
            gam[V2(i, j)] = (bv[V2(i, j)] - t->Asub_v[V2(i, j)] * gam[V2(i, j - 1)]) * t->denom_v[V2(i, j)];
        j = g->jhi;
        xv[V2(i, j)] = gam[V2(i, j)];
        for (j = g->jhi - 1; j >= 1; --j) xv[V2(i, j)] = gam[V2(i, j)] - t->omega_v[V2(i, j)] * xv[V2(i, j + 1)];
    }
    free(gam);
}

/* particular solution xr of tridiag_solver_global_2d (pcg.F90:5228-5262)
 * on an (ilocal, jlocal) owned block; with one task xu = xr (:5495-5497). */
static void tridiag_global_owned_solve(int ilocal, int jlocal, const double* Adiag, const double* Asub,
                                       const double* omega, const double* denom, const double* b, double* x)
{
    double* gam = (double*)malloc((size_t)ilocal * sizeof(double));
#define L(i, j) ((size_t)((i) - 1) + (size_t)ilocal * (size_t)((j) - 1))
    for (int j = 1; j <= jlocal; ++j) {
        gam[0] = b[L(1, j)] * Adiag[L(1, j)];
        for (int i = 2; i <= ilocal; ++i) gam[i - 1] = (b[L(i, j)] - Asub[L(i, j)] * gam[i - 2]) * denom[L(i, j)];
        x[L(ilocal, j)] = gam[ilocal - 1];
        for (int i = ilocal - 1; i >= 1; --i) x[L(i, j)] = gam[i - 1] - omega[L(i, j)] * x[L(i + 1, j)];
    }
#undef L
    free(gam);
}

/* Apply z = PC(r); precond branches of pcg.F90:1236-1320 / :3301-3381. */
static void precond_apply(const orc_grid* g, const int32_t* active, pcstate* s,
                          const double* ru, const double* rv, double* zu, double* zv)
{
    const size_t n = (size_t)NZ * NVX * NVY, n2 = (size_t)NVX * NVY;
    if (s->precond == PC_NONE) {
        memcpy(zu, ru, n * sizeof(double));
        memcpy(zv, rv, n * sizeof(double));
    } else if (s->precond == PC_DIAG) {
        for (size_t p = 0; p < n; ++p) {
            zu[p] = (s->Adiagu[p] != 0.0) ? ru[p] / s->Adiagu[p] : 0.0;
            zv[p] = (s->Adiagv[p] != 0.0) ? rv[p] / s->Adiagv[p] : 0.0;
        }
    } else if (s->precond == PC_SIA) {
        easy_sia_solver(g, active, s->Muu, ru, zu);
        easy_sia_solver(g, active, s->Mvv, rv, zv);
    } else {
        /* tridiag_solver_local_3d (pcg.F90:4524-4658) / _global_3d (:4662-4913):
         * column sums of r, 2-D line solves, halo, SIA column solve, add. */
        for (int j = 1; j <= NVY; ++j)
            for (int i = 1; i <= NVX; ++i) {
                double su = 0.0, sv = 0.0;
                for (int k = 1; k <= NZ; ++k) { su += ru[V3(k, i, j)]; sv += rv[V3(k, i, j)]; }
                s->ru_vs[V2(i, j)] = su;
                s->rv_vs[V2(i, j)] = sv;
            }
        if (s->precond == PC_TRIDIAG_LOCAL) {
            tridiag_local_2d_solve(g, &s->t, s->ru_vs, s->rv_vs, s->zu_vs, s->zv_vs);
        } else {
            const int il = s->ilocal, jl = s->jlocal;
            double* bl = (double*)calloc((size_t)il * jl, sizeof(double));
            double* xl = (double*)malloc((size_t)il * jl * sizeof(double));
            for (int j = 1; j <= jl; ++j)
                for (int i = 1; i <= il; ++i)
                    bl[(size_t)(i - 1) + (size_t)il * (j - 1)] = s->ru_vs[V2(i + g->ilo - 1, j + g->jlo - 1)];
            tridiag_global_owned_solve(il, jl, s->gu_diag, s->gu_sub, s->gu_omega, s->gu_denom, bl, xl);
            memset(s->zu_vs, 0, n2 * sizeof(double));
            for (int j = 1; j <= jl; ++j)
                for (int i = 1; i <= il; ++i)
                    s->zu_vs[V2(i + g->ilo - 1, j + g->jlo - 1)] = xl[(size_t)(i - 1) + (size_t)il * (j - 1)];
            for (int i = 1; i <= il; ++i)
                for (int j = 1; j <= jl; ++j)
                    bl[(size_t)(j - 1) + (size_t)jl * (i - 1)] = s->rv_vs[V2(i + g->ilo - 1, j + g->jlo - 1)];
            tridiag_global_owned_solve(jl, il, s->gv_diag, s->gv_sub, s->gv_omega, s->gv_denom, bl, xl);
            memset(s->zv_vs, 0, n2 * sizeof(double));
            for (int i = 1; i <= il; ++i)
                for (int j = 1; j <= jl; ++j)
                    s->zv_vs[V2(i + g->ilo - 1, j + g->jlo - 1)] = xl[(size_t)(j - 1) + (size_t)jl * (i - 1)];
            free(bl); free(xl);
        }
        orc_halo3d(g, 1, s->zu_vs);
        orc_halo3d(g, 1, s->zv_vs);
        if (s->two_d) {   /* 2-D: z is the line solution (pcg.F90:1260-1320, :2352-2425) */
            memcpy(zu, s->zu_vs, n2 * sizeof(double));
            memcpy(zv, s->zv_vs, n2 * sizeof(double));
            return;
        }
        easy_sia_solver(g, active, s->Muu, ru, zu);
        easy_sia_solver(g, active, s->Mvv, rv, zv);
        for (int j = 1; j <= NVY; ++j)
            for (int i = 1; i <= NVX; ++i)
                for (int k = 1; k <= NZ; ++k) {
                    zu[V3(k, i, j)] = zu[V3(k, i, j)] + s->zu_vs[V2(i, j)];
                    zv[V3(k, i, j)] = zv[V3(k, i, j)] + s->zv_vs[V2(i, j)];
                }
    }
}

static int check_args(const orc_grid* g, int precond, int ncheck, int two_d)
{
    if (!g || g->nx < 6 || g->ny < 6 || g->nz < (two_d ? 1 : 2) || g->nz > 64) return ORC_ERR_ARG;
    if (two_d && g->nz != 1) return ORC_ERR_ARG;
    if (g->ilo < 2 || g->ihi > g->nx - 2 || g->jlo < 2 || g->jhi > g->ny - 2) return ORC_ERR_ARG;
    if (g->ilo > g->ihi || g->jlo > g->jhi) return ORC_ERR_ARG;
    if (precond < 0 || precond > 4 || ncheck < 1) return ORC_ERR_ARG;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* pcg_solver_standard_3d (pcg.F90:827-1537)                           */
/* ------------------------------------------------------------------ */
static int standard_core(const orc_grid* g, const int32_t* active,
                         const double* Auu, const double* Auv,
                         const double* Avu, const double* Avv,
                         const double* bu, const double* bv,
                         double* xu, double* xv,
                         int32_t precond, int32_t ncheck, double tolerance,
                         int32_t maxiters, double* err, int32_t* niters, int two_d)
{
    int st = check_args(g, precond, ncheck, two_d);
    if (st) return st;
    g_trace_len = 0;
    const size_t n = (size_t)NZ * NVX * NVY;
    double* buf = (double*)calloc(10 * n, sizeof(double));
    double *ru = buf, *rv = buf + n, *du = buf + 2 * n, *dv = buf + 3 * n, *qu = buf + 4 * n, *qv = buf + 5 * n;
    double *zu = buf + 6 * n, *zv = buf + 7 * n, *wu = buf + 8 * n, *wv = buf + 9 * n;
    pcstate pc;
    st = precond_setup(g, active, Auu, Avv, precond, two_d, &pc);
    if (st) { precond_free(&pc); free(buf); return st; }

    orc_halo3d(g, NZ, xu);
    orc_halo3d(g, NZ, xv);
    orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
    for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
    *niters = maxiters;
    double eta0 = 1.0, eta1, eta2, alpha, beta;
    memset(du, 0, n * sizeof(double)); memset(dv, 0, n * sizeof(double));
    memset(zu, 0, n * sizeof(double)); memset(zv, 0, n * sizeof(double));
    for (size_t p = 0; p < n; ++p) { wu[p] = bu[p] * bu[p]; wv[p] = bv[p] * bv[p]; }
    double L2_rhs = sqrt(orc_global_sum(g, wu, wv));

    for (int iter = 1; iter <= maxiters; ++iter) {
        precond_apply(g, active, &pc, ru, rv, zu, zv);
        for (size_t p = 0; p < n; ++p) { wu[p] = ru[p] * zu[p]; wv[p] = rv[p] * zv[p]; }
        eta1 = orc_global_sum(g, wu, wv);
        if (eta1 != eta1) { st = ORC_ERR_NAN; break; }
        beta = eta1 / eta0;
        for (size_t p = 0; p < n; ++p) { du[p] = zu[p] + beta * du[p]; dv[p] = zv[p] + beta * dv[p]; }
        orc_halo3d(g, NZ, du);
        orc_halo3d(g, NZ, dv);
        orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, du, dv, qu, qv);
        eta0 = eta1;
        for (size_t p = 0; p < n; ++p) { wu[p] = du[p] * qu[p]; wv[p] = dv[p] * qv[p]; }
        eta2 = orc_global_sum(g, wu, wv);
        alpha = eta1 / eta2;
        trace3(eta1, eta2, alpha);
        if (alpha != alpha) { st = ORC_ERR_NAN; break; }
        for (size_t p = 0; p < n; ++p) {
            xu[p] = xu[p] + alpha * du[p]; xv[p] = xv[p] + alpha * dv[p];
            ru[p] = ru[p] - alpha * qu[p]; rv[p] = rv[p] - alpha * qv[p];
        }
        if (iter % ncheck == 0 || iter == 5) {
            orc_halo3d(g, NZ, xu);
            orc_halo3d(g, NZ, xv);
            orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
            for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
            for (size_t p = 0; p < n; ++p) { wu[p] = ru[p] * ru[p]; wv[p] = rv[p] * rv[p]; }
            double L2_resid = sqrt(orc_global_sum(g, wu, wv));
            *err = L2_resid / L2_rhs;
            if (*err < tolerance) { *niters = iter; break; }
        }
    }
    precond_free(&pc);
    free(buf);
    return st;
}

/* ------------------------------------------------------------------ */
/* pcg_solver_chrongear_3d (pcg.F90:2632-3594)                          */
/* ------------------------------------------------------------------ */
static int chrongear_core(const orc_grid* g, const int32_t* active,
                          const double* Auu, const double* Auv,
                          const double* Avu, const double* Avv,
                          const double* bu, const double* bv,
                          double* xu, double* xv,
                          int32_t precond, int32_t ncheck, double tolerance,
                          int32_t maxiters, double* err, int32_t* niters, int two_d)
{
    int st = check_args(g, precond, ncheck, two_d);
    if (st) return st;
    g_trace_len = 0;
    const size_t n = (size_t)NZ * NVX * NVY;
    double* buf = (double*)calloc(16 * n, sizeof(double));
    double *ru = buf, *rv = buf + n, *du = buf + 2 * n, *dv = buf + 3 * n, *zu = buf + 4 * n, *zv = buf + 5 * n;
    double *qu = buf + 6 * n, *qv = buf + 7 * n, *Azu = buf + 8 * n, *Azv = buf + 9 * n;
    double *wu = buf + 10 * n, *wv = buf + 11 * n, *w2u1 = buf + 12 * n, *w2v1 = buf + 13 * n;
    double *w2u2 = buf + 14 * n, *w2v2 = buf + 15 * n;
    pcstate pc;
    st = precond_setup(g, active, Auu, Avv, precond, two_d, &pc);
    if (st) { precond_free(&pc); free(buf); return st; }

    *niters = maxiters;
    for (size_t p = 0; p < n; ++p) { wu[p] = bu[p] * bu[p]; wv[p] = bv[p] * bv[p]; }
    const double bb = orc_global_sum(g, wu, wv);
    const double L2_rhs = sqrt(bb);

    /* first pass, iter = 1 (:3085-3290) */
    orc_halo3d(g, NZ, xu);
    orc_halo3d(g, NZ, xv);
    orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
    for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
    orc_halo3d(g, NZ, ru);
    orc_halo3d(g, NZ, rv);
    precond_apply(g, active, &pc, ru, rv, zu, zv);
    for (size_t p = 0; p < n; ++p) { w2u1[p] = ru[p] * zu[p]; w2v1[p] = rv[p] * zv[p]; }
    memcpy(du, zu, n * sizeof(double));
    memcpy(dv, zv, n * sizeof(double));
    orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, du, dv, qu, qv);
    for (size_t p = 0; p < n; ++p) { w2u2[p] = du[p] * qu[p]; w2v2[p] = dv[p] * qv[p]; }
    double gsum[2];
    orc_global_sum2(g, w2u1, w2v1, w2u2, w2v2, gsum);
    orc_halo3d(g, NZ, qu);
    orc_halo3d(g, NZ, qv);
    double rho_old = gsum[0];
    double sigma = gsum[1];
    double alpha = rho_old / sigma;
    trace3(rho_old, sigma, alpha);
    if (alpha != alpha) { st = ORC_ERR_NAN; goto done; }
    for (size_t p = 0; p < n; ++p) {
        xu[p] = xu[p] + alpha * du[p]; xv[p] = xv[p] + alpha * dv[p];
        ru[p] = ru[p] - alpha * qu[p]; rv[p] = rv[p] - alpha * qv[p];
    }

    for (int iter = 2; iter <= maxiters; ++iter) {
        precond_apply(g, active, &pc, ru, rv, zu, zv);
        orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, zu, zv, Azu, Azv);
        for (size_t p = 0; p < n; ++p) {
            w2u1[p] = ru[p] * zu[p]; w2v1[p] = rv[p] * zv[p];
            w2u2[p] = Azu[p] * zu[p]; w2v2[p] = Azv[p] * zv[p];
        }
        orc_global_sum2(g, w2u1, w2v1, w2u2, w2v2, gsum);
        orc_halo3d(g, NZ, Azu);
        orc_halo3d(g, NZ, Azv);
        const double rho = gsum[0], delta = gsum[1];
        const double beta = rho / rho_old;
        sigma = delta - beta * rho / alpha;
        alpha = rho / sigma;
        rho_old = rho;
        trace3(rho, sigma, alpha);
        if (alpha != alpha) { st = ORC_ERR_NAN; break; }
        for (size_t p = 0; p < n; ++p) {
            du[p] = zu[p] + beta * du[p]; dv[p] = zv[p] + beta * dv[p];
            qu[p] = Azu[p] + beta * qu[p]; qv[p] = Azv[p] + beta * qv[p];
            xu[p] = xu[p] + alpha * du[p]; xv[p] = xv[p] + alpha * dv[p];
            ru[p] = ru[p] - alpha * qu[p]; rv[p] = rv[p] - alpha * qv[p];
        }
        if (iter % ncheck == 0 || iter == 5) {
            orc_matvec_3d(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
            for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
            for (size_t p = 0; p < n; ++p) { wu[p] = ru[p] * ru[p]; wv[p] = rv[p] * rv[p]; }
            const double rr = orc_global_sum(g, wu, wv);
            *err = sqrt(rr) / L2_rhs;
            if (*err < tolerance) { *niters = iter; break; }
            orc_halo3d(g, NZ, ru);
            orc_halo3d(g, NZ, rv);
        }
    }
done:
    precond_free(&pc);
    free(buf);
    return st;
}

/* ------------------------------------------------------------------ */
/* Public 3-D entry points                                              */
/* ------------------------------------------------------------------ */
int orc_pcg_solver_standard_3d(const orc_grid* g, const int32_t* active, const double* Auu, const double* Auv,
                               const double* Avu, const double* Avv, const double* bu, const double* bv,
                               double* xu, double* xv, int32_t precond, int32_t ncheck, double tolerance,
                               int32_t maxiters, double* err, int32_t* niters)
{
    return standard_core(g, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tolerance, maxiters, err,
                         niters, 0);
}

int orc_pcg_solver_chrongear_3d(const orc_grid* g, const int32_t* active, const double* Auu, const double* Auv,
                                const double* Avu, const double* Avv, const double* bu, const double* bv,
                                double* xu, double* xv, int32_t precond, int32_t ncheck, double tolerance,
                                int32_t maxiters, double* err, int32_t* niters)
{
    return chrongear_core(g, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tolerance, maxiters, err,
                          niters, 0);
}

/* ------------------------------------------------------------------ */
/* 2-D solvers: pcg_solver_standard_2d (pcg.F90:67-823),                */
/* pcg_solver_chrongear_2d (:1541-2628), matvec_multiply_structured_2d  */
/* (:5656-5773).  Their recurrences, check schedule and sums are those  */
/* of the 3-D solvers on one level; the 9-point stencil                 */
/* A(nx-1,ny-1,9) (m last, m = indxA_2d) is the kA = 0 plane of a       */
/* one-level 27-point row, the sums run over (j, i) as the 3-D ones over */
/* (j, i, k), and the preconditioners are diagonal or the line solves   */
/* alone (no SIA column part; the tridiagonal setup takes the stencil as */
/* given, :4247-4392).  which_ho_precond = 2 is replaced by the diagonal */
/* preconditioner, as glide_setup.F90:1355-1367 does for 2-D solves.    */
/* ------------------------------------------------------------------ */
static double* stencil_2d_to_rows(const orc_grid* g, const double* A2)
{
    const size_t n2 = (size_t)NVX * NVY;
    double* A3 = (double*)calloc(27 * n2, sizeof(double));
    for (size_t p = 0; p < n2; ++p)
        for (int m = 1; m <= 9; ++m) A3[27 * p + (m + 9) - 1] = A2[p + n2 * (m - 1)];
    return A3;
}

static int solve_2d(int algo, const orc_grid* g2, const int32_t* active, const double* Auu, const double* Auv,
                    const double* Avu, const double* Avv, const double* bu, const double* bv, double* xu, double* xv,
                    int32_t precond, int32_t ncheck, double tolerance, int32_t maxiters, double* err,
                    int32_t* niters)
{
    orc_grid g = *g2;
    g.nz = 1;
    const orc_grid* gp = &g;
    double* A3[4] = {stencil_2d_to_rows(gp, Auu), stencil_2d_to_rows(gp, Auv), stencil_2d_to_rows(gp, Avu),
                     stencil_2d_to_rows(gp, Avv)};
    if (precond == PC_SIA) precond = PC_DIAG;
    int st = algo == 2 ? standard_core(gp, active, A3[0], A3[1], A3[2], A3[3], bu, bv, xu, xv, precond, ncheck,
                                       tolerance, maxiters, err, niters, 1)
                       : chrongear_core(gp, active, A3[0], A3[1], A3[2], A3[3], bu, bv, xu, xv, precond, ncheck,
                                        tolerance, maxiters, err, niters, 1);
    for (int q = 0; q < 4; ++q) free(A3[q]);
    return st;
}

int orc_pcg_solver_standard_2d(const orc_grid* g, const int32_t* active, const double* Auu, const double* Auv,
                               const double* Avu, const double* Avv, const double* bu, const double* bv,
                               double* xu, double* xv, int32_t precond, int32_t ncheck, double tolerance,
                               int32_t maxiters, double* err, int32_t* niters)
{
    return solve_2d(2, g, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tolerance, maxiters, err,
                    niters);
}

int orc_pcg_solver_chrongear_2d(const orc_grid* g, const int32_t* active, const double* Auu, const double* Auv,
                                const double* Avu, const double* Avv, const double* bu, const double* bv,
                                double* xu, double* xv, int32_t precond, int32_t ncheck, double tolerance,
                                int32_t maxiters, double* err, int32_t* niters)
{
    return solve_2d(3, g, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tolerance, maxiters, err,
                    niters);
}

void orc_matvec_2d(const orc_grid* g2, const int32_t* active, const double* Auu, const double* Auv,
                   const double* Avu, const double* Avv, const double* xu, const double* xv, double* yu, double* yv)
{
    orc_grid g = *g2;
    g.nz = 1;
    double* A3[4] = {stencil_2d_to_rows(&g, Auu), stencil_2d_to_rows(&g, Auv), stencil_2d_to_rows(&g, Avu),
                     stencil_2d_to_rows(&g, Avv)};
    orc_matvec_3d(&g, active, A3[0], A3[1], A3[2], A3[3], xu, xv, yu, yv);
    for (int q = 0; q < 4; ++q) free(A3[q]);
}
