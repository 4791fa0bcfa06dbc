/*
 * oracle_omp.c -- all-cores variant of the CPU restatement of
 * pcg_solver_chrongear_3d (glissade_velo_higher_pcg.F90:2632-3594), used
 * only as the multi-core CPU baseline.  TEST INFRASTRUCTURE ONLY.
 *
 * Per-element arithmetic is identical to oracle_pcg.c.  Loops over
 * (j, i) columns are split across OpenMP threads; each global sum is formed
 * as per-row (fixed j) partial sums in the reference's (i, k) order, then
 * combined over j in order, so the result does not depend on the thread
 * count.  Preconditioners 0, 1 and 2 are supported.
 */
#include "cism_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define NVX (g->nx - 1)
#define NVY (g->ny - 1)
#define NZ  (g->nz)
#define V3(k, i, j) ((size_t)((k) - 1) + (size_t)NZ * ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1)))
#define V2(i, j)    ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1))
#define A4(m, k, i, j) ((size_t)((m) - 1) + 27 * V3(k, i, j))
#define IDX3(iA, jA, kA) (1 + ((iA) + 1) + 3 * ((jA) + 1) + 9 * ((kA) + 1))

static void matvec_omp(const orc_grid* g, const int32_t* active,
                       const double* Auu, const double* Auv, const double* Avu, const double* Avv,
                       const double* xu, const double* xv, double* yu, double* yv)
{
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NVY; ++j) {
        memset(yu + V3(1, 1, j), 0, (size_t)NZ * NVX * sizeof(double));
        memset(yv + V3(1, 1, j), 0, (size_t)NZ * NVX * sizeof(double));
        if (j < g->jlo || j > g->jhi) continue;
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!active[V2(i, j)]) continue;
            for (int k = 1; k <= NZ; ++k) {
                double su = 0.0, sv = 0.0;
                for (int kA = -1; kA <= 1; ++kA)
                    for (int jA = -1; jA <= 1; ++jA)
                        for (int iA = -1; iA <= 1; ++iA) {
                            if (k + kA >= 1 && k + kA <= NZ) {
                                const size_t a = A4(IDX3(iA, jA, kA), k, i, j);
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
}

/* sum over owned vertices of a1*b1 + a2*b2, per-row partials combined in j order */
static double dot_omp(const orc_grid* g, const double* a1, const double* b1, const double* a2, const double* b2,
                      double* rowbuf)
{
#pragma omp parallel for schedule(static)
    for (int j = g->jlo; j <= g->jhi; ++j) {
        double s = 0.0;
        for (int i = g->ilo; i <= g->ihi; ++i)
            for (int k = 1; k <= NZ; ++k) {
                const size_t p = V3(k, i, j);
                s = s + a1[p] * b1[p] + a2[p] * b2[p];
            }
        rowbuf[j - g->jlo] = s;
    }
    double s = 0.0;
    for (int j = g->jlo; j <= g->jhi; ++j) s += rowbuf[j - g->jlo];
    return s;
}

static void precond_omp(const orc_grid* g, const int32_t* active, int precond,
                        const double* Auu, const double* Avv,
                        const double* ru, const double* rv, double* zu, double* zv)
{
    const size_t n = (size_t)NZ * NVX * NVY;
    if (precond == 0) {
#pragma omp parallel for schedule(static)
        for (size_t p = 0; p < n; ++p) { zu[p] = ru[p]; zv[p] = rv[p]; }
    } else if (precond == 1) {
#pragma omp parallel for schedule(static)
        for (size_t p = 0; p < n; ++p) {
            const double du = Auu[27 * p + 13], dv = Avv[27 * p + 13];
            zu[p] = du != 0.0 ? ru[p] / du : 0.0;
            zv[p] = dv != 0.0 ? rv[p] / dv : 0.0;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= NVY; ++j)
            for (int i = 1; i <= NVX; ++i) {
                if (!active[V2(i, j)]) continue;
                for (int c = 0; c < 2; ++c) {
                    const double* A = c ? Avv : Auu;
                    const double* r = c ? rv : ru;
                    double* z = c ? zv : zu;
                    double sb[64], d[64], sp[64], rhs[64], soln[64];
                    for (int k = 1; k <= NZ; ++k) {
                        const size_t p = V3(k, i, j);
                        sb[k - 1] = (k == 1) ? 0.0 : A[27 * p + IDX3(0, 0, -1) - 1];
                        d[k - 1] = A[27 * p + IDX3(0, 0, 0) - 1];
                        sp[k - 1] = (k == NZ) ? 0.0 : A[27 * p + IDX3(0, 0, 1) - 1];
                        rhs[k - 1] = r[p];
                    }
                    orc_tridiag_solver(NZ, sb, d, sp, rhs, soln);
                    for (int k = 1; k <= NZ; ++k) z[V3(k, i, j)] = soln[k - 1];
                }
            }
    }
}

int orc_pcg_solver_chrongear_3d_omp(const orc_grid* g, const int32_t* active,
                                    const double* Auu, const double* Auv,
                                    const double* Avu, const double* Avv,
                                    const double* bu, const double* bv,
                                    double* xu, double* xv,
                                    int32_t precond, int32_t ncheck, double tolerance,
                                    int32_t maxiters, int32_t nthreads, double* err, int32_t* niters)
{
    if (precond < 0 || precond > 2 || ncheck < 1) return ORC_ERR_ARG;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const size_t n = (size_t)NZ * NVX * NVY;
    double* buf = (double*)calloc(10 * n, sizeof(double));
    double* rowbuf = (double*)malloc((size_t)(g->jhi - g->jlo + 1) * sizeof(double));
    double *ru = buf, *rv = buf + n, *du = buf + 2 * n, *dv = buf + 3 * n, *zu = buf + 4 * n, *zv = buf + 5 * n;
    double *qu = buf + 6 * n, *qv = buf + 7 * n, *Azu = buf + 8 * n, *Azv = buf + 9 * n;
    int st = ORC_OK;
    *niters = maxiters;
    const double L2_rhs = sqrt(dot_omp(g, bu, bu, bv, bv, rowbuf));
    orc_halo3d(g, NZ, xu);
    orc_halo3d(g, NZ, xv);
    matvec_omp(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
    orc_halo3d(g, NZ, ru);
    orc_halo3d(g, NZ, rv);
    precond_omp(g, active, precond, Auu, Avv, ru, rv, zu, zv);
    memcpy(du, zu, n * sizeof(double));
    memcpy(dv, zv, n * sizeof(double));
    matvec_omp(g, active, Auu, Auv, Avu, Avv, du, dv, qu, qv);
    double rho_old = dot_omp(g, ru, zu, rv, zv, rowbuf);
    double sigma = dot_omp(g, du, qu, dv, qv, rowbuf);
    orc_halo3d(g, NZ, qu);
    orc_halo3d(g, NZ, qv);
    double alpha = rho_old / sigma;
    if (alpha != alpha) { st = ORC_ERR_NAN; goto done; }
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n; ++p) {
        xu[p] = xu[p] + alpha * du[p]; xv[p] = xv[p] + alpha * dv[p];
        ru[p] = ru[p] - alpha * qu[p]; rv[p] = rv[p] - alpha * qv[p];
    }
    for (int iter = 2; iter <= maxiters; ++iter) {
        precond_omp(g, active, precond, Auu, Avv, ru, rv, zu, zv);
        matvec_omp(g, active, Auu, Auv, Avu, Avv, zu, zv, Azu, Azv);
        const double rho = dot_omp(g, ru, zu, rv, zv, rowbuf);
        const double delta = dot_omp(g, Azu, zu, Azv, zv, rowbuf);
        orc_halo3d(g, NZ, Azu);
        orc_halo3d(g, NZ, Azv);
        const double beta = rho / rho_old;
        sigma = delta - beta * rho / alpha;
        alpha = rho / sigma;
        rho_old = rho;
        if (alpha != alpha) { st = ORC_ERR_NAN; break; }
#pragma omp parallel for schedule(static)
        for (size_t p = 0; p < n; ++p) {
            du[p] = zu[p] + beta * du[p]; dv[p] = zv[p] + beta * dv[p];
            qu[p] = Azu[p] + beta * qu[p]; qv[p] = Azv[p] + beta * qv[p];
            xu[p] = xu[p] + alpha * du[p]; xv[p] = xv[p] + alpha * dv[p];
            ru[p] = ru[p] - alpha * qu[p]; rv[p] = rv[p] - alpha * qv[p];
        }
        if (iter % ncheck == 0 || iter == 5) {
            matvec_omp(g, active, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
#pragma omp parallel for schedule(static)
            for (size_t p = 0; p < n; ++p) { ru[p] = bu[p] - zu[p]; rv[p] = bv[p] - zv[p]; }
            *err = sqrt(dot_omp(g, ru, ru, rv, rv, rowbuf)) / L2_rhs;
            if (*err < tolerance) { *niters = iter; break; }
            orc_halo3d(g, NZ, ru);
            orc_halo3d(g, NZ, rv);
        }
    }
done:
    free(rowbuf);
    free(buf);
    return st;
}
