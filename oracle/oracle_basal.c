/*
 * oracle_basal.c -- CPU restatement of glissade_calcbeta
 * (libglissade/glissade_basal_traction.F90:69-719): the basal traction
 * coefficient beta at the vertices from the basal velocity, for every
 * which_ho_babc option but no-slip (:644, beta untouched), then the f_ground
 * weighting with which_ho_beta_limit (:650-685), the beta >= 0 check
 * (:687-699); the staggered halo that follows (:703) is orc_halo3d with
 * one level.
 *
 * TEST INFRASTRUCTURE ONLY: the tests and the benchmark's CPU leg use it as
 * the checker of the device kernel; nothing in the product calls it.
 *
 * Layouts: vertex fields (nx-1, ny-1) and cell fields (nx, ny), i fastest.
 * Built with -ffp-contract=off: every product and sum is rounded as the
 * Fortran expression orders it (left to right).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#include "cism_oracle_asm.h"

#define NVX (nx - 1)
#define C2(i, j) ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1))
#define V2(i, j) ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1))

/* glimmer_physcon.F90:57, 67-68 */
static const double ORC_PI = 3.14159265358979;
static const double RHOI_B = 910.0, RHOO_B = 1028.0;

/* speed with the floor smallnum = 1e-2 m/yr (:180), capped at
 * beta_powerlaw_umax when that is positive (:187-189) */
static double speed_at(double u, double v, double umax)
{
    const double smallnum = 1.0e-2;
    double s = sqrt(u * u + v * v + smallnum * smallnum);
    if (umax > 0.0 && s > umax) s = umax;
    return s;
}

/*
 * orc_calcbeta: p[] holds the scalar parameters in the order of
 * ORC_CB_* in cism_oracle_asm.h.  Fields that an option does not read may
 * be NULL; c_space_factor_stag NULL means 1 everywhere.  Returns 0, or
 * 1 + the vertex index (i-1 + (nx-1)(j-1)) of the first beta that is
 * negative or NaN (the reference stops there with GM_FATAL, :690-697),
 * or -1 for an option it does not cover.
 */
/* ISMIP-HOM C (:343-397) on one task (two halo cells): the global field
 * 1000 + 1000 sin(w dew ew) sin(w dns ns), w = 2 pi / (global_ewn dew),
 * placed on the owned cells (ew = i - 2, ns = j - 2), the cell halo rule
 * of parallel_halo (periodic: wrap; outflow / no-ice: zero), and each
 * vertex (i, j) takes cell (i, j). */
static double ishomc_at(int nx, int ny, int global_bc, double dew, double dns, int i, int j)
{
    const int n = nx - 4, m = ny - 4;
    if (global_bc == 0 || global_bc == 2) {
        if (i <= 2) i += n; else if (i >= nx - 1) i -= n;
        if (j <= 2) j += m; else if (j >= ny - 1) j -= m;
    } else {
        const int w = global_bc == 1 ? 2 : 3;
        if (i <= w || i >= nx - w + 1 || j <= w || j >= ny - w + 1) return 0.0;
    }
    const double ldomain = n * dew;
    const double omega = 2.0 * ORC_PI / ldomain;
    const double x = dew * (i - 2), y = dns * (j - 2);
    return 1000.0 + 1000.0 * sin(omega * x) * sin(omega * y);
}

int orc_calcbeta(int nx, int ny, int global_bc, int whichbabc, int which_beta_limit, const double* p,
                 const double* uvel, const double* vvel, const double* effecpress_stag, const double* coulomb_c,
                 const double* powerlaw_c, const double* mintauf, const double* c_space_factor_stag,
                 const int32_t* bpmp_mask, const double* beta_external, const double* f_ground, const double* topg,
                 const double* thck, const double* flwa_basal, double* beta)
{
    const double umax = p[ORC_CB_UMAX];
    const double m = p[ORC_CB_M];
    const size_t nv = (size_t)NVX * (size_t)(ny - 1);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const size_t v = V2(i, j);
            const double speed = speed_at(uvel[v], vvel[v], umax);
            double b = 0.0;
            switch (whichbabc) {
            case 0:   /* constant (:202-204) */
                b = p[ORC_CB_BETA_CONST];
                break;
            case 1:   /* bpmp (:206-212) */
                b = bpmp_mask[v] == 1 ? p[ORC_CB_BETA_SMALL] : p[ORC_CB_BETA_LARGE];
                break;
            case 2: { /* pseudo-plastic, tau_c = N coulomb_c (:214-245) */
                const double q = p[ORC_CB_PP_Q], u0 = p[ORC_CB_PP_U0];
                const double tau_c = effecpress_stag[v] * coulomb_c[v];
                b = tau_c / (pow(u0, q) * pow(speed, 1.0 - q));
                break;
            }
            case 3: { /* old pseudo-plastic: phi from the bed elevation at the
                       * cell with the vertex's indices (:247-289) */
                const double q = p[ORC_CB_PP_Q], u0 = p[ORC_CB_PP_U0];
                const double phimin = p[ORC_CB_PHIMIN], phimax = p[ORC_CB_PHIMAX];
                const double bedmin = p[ORC_CB_BEDMIN], bedmax = p[ORC_CB_BEDMAX];
                const double bed = topg[C2(i, j)] - p[ORC_CB_EUS];
                double phi;
                if (bed <= bedmin) phi = phimin;
                else if (bed >= bedmax) phi = phimax;
                else phi = phimin + ((bed - bedmin) / (bedmax - bedmin)) * (phimax - phimin);
                const double tanphi = tan(phi * ORC_PI / 180.0);
                const double tau_c = tanphi * effecpress_stag[v];
                b = tau_c / (pow(u0, q) * pow(speed, 1.0 - q));
                break;
            }
            case 4:   /* large (:304-309) */
                b = p[ORC_CB_BETA_LARGE];
                break;
            case 5:   /* external (:399-413) */
                b = beta_external[v];
                break;
            case 7: { /* Zoet-Iverson with N capped at nmax (:311-341) */
                const double nc = effecpress_stag[v] < p[ORC_CB_ZI_NMAX] ? effecpress_stag[v] : p[ORC_CB_ZI_NMAX];
                const double tau_c = coulomb_c[v] * nc;
                b = tau_c * pow(speed, 1.0 / m - 1.0) / pow(speed + p[ORC_CB_ZI_UT], 1.0 / m);
                break;
            }
            case 8:   /* ISMIP-HOM C (:343-397) */
                b = ishomc_at(nx, ny, global_bc, p[ORC_CB_DEW], p[ORC_CB_DNS], i, j);
                break;
            case 9:   /* power law (:415-433) */
                b = powerlaw_c[v] * pow(speed, 1.0 / m - 1.0);
                break;
            case 10: {/* Coulomb friction, Schoof (2005) with bump parameters (:456-507);
                       * basal flwa staggered over the cells with thck > 0 (:476-488) */
                int32_t mk[4];
                const size_t c01 = C2(i, j + 1), c11 = C2(i + 1, j + 1), c00 = C2(i, j), c10 = C2(i + 1, j);
                mk[0] = thck[c01] > 0.0; mk[1] = thck[c11] > 0.0; mk[2] = thck[c00] > 0.0; mk[3] = thck[c10] > 0.0;
                const double sumvar = mk[0] * flwa_basal[c01] + mk[1] * flwa_basal[c11] + mk[2] * flwa_basal[c00] +
                                      mk[3] * flwa_basal[c10];
                const double summask = (double)(mk[0] + mk[1] + mk[2] + mk[3]);
                const double fs = summask > 0.0 ? sumvar / summask : 0.0;
                const double big_lambda = (p[ORC_CB_BUMP_WAVELENGTH] / p[ORC_CB_BUMP_SLOPE]) * fs;
                const double N = effecpress_stag[v];
                b = coulomb_c[v] * N * pow(speed, 1.0 / m - 1.0) * pow(speed + pow(N, m) * big_lambda, -1.0 / m);
                if (c_space_factor_stag) b = b * c_space_factor_stag[v];
                if (b > 1.0e8) b = 1.0e8;
                break;
            }
            case 11: {/* Schoof, Asay-Davis et al. (2016) form (:509-579) */
                const double num = powerlaw_c[v] * coulomb_c[v] * effecpress_stag[v];
                const double den = pow(pow(powerlaw_c[v], m) * speed + pow(coulomb_c[v] * effecpress_stag[v], m), 1.0 / m);
                b = (num / den) * pow(speed, 1.0 / m - 1.0);
                if (c_space_factor_stag) b = b * c_space_factor_stag[v];
                if (b > 1.0e8) b = 1.0e8;
                break;
            }
            case 12: {/* modified Schoof: 1/tau_b = 1/tau_p + 1/tau_c (:581-614); with
                       * neither stress positive the reference leaves tau_b unset, here 0 */
                const double tp = powerlaw_c[v] * pow(speed, 1.0 / m);
                const double tc = coulomb_c[v] * effecpress_stag[v];
                double taub = 0.0;
                if (tc > 0.0 && tp > 0.0) taub = 1.0 / (1.0 / tp + 1.0 / tc);
                else if (tp > 0.0) taub = tp;
                else if (tc > 0.0) taub = tc;
                b = taub / speed;
                break;
            }
            case 13: {/* Tsai: the smaller of the two stresses (:616-642) */
                const double tp = powerlaw_c[v] * pow(speed, 1.0 / m);
                const double tc = coulomb_c[v] * effecpress_stag[v];
                b = (tc <= tp ? tc : tp) / speed;
                if (c_space_factor_stag) b = b * c_space_factor_stag[v];
                break;
            }
            case 14: {/* power law with effective pressure, p = 3, q = 1 (:435-454) */
                const double pp = 3.0, pq = 1.0;
                b = pow(p[ORC_CB_POWERLAW_K], -1.0 / pp) * pow(effecpress_stag[v], pq / pp) * pow(speed, 1.0 / pp - 1.0);
                break;
            }
            case 15: {/* yield stress for Picard, no speed cap (:291-302) */
                const double smallnum = 1.0e-2;
                b = mintauf[v] / sqrt(uvel[v] * uvel[v] + vvel[v] * vvel[v] + smallnum * smallnum);
                break;
            }
            default:
                return -1;
            }
            /* f_ground and beta_grounded_min (:667-685) */
            const double fg = f_ground[v], gmin = p[ORC_CB_BETA_GMIN];
            if (which_beta_limit == 0) {
                b = b * fg;
                if (fg > 0.0 && b < gmin) b = gmin;
            } else {
                if (fg > 0.0 && b < gmin) b = gmin;
                b = b * fg;
            }
            beta[v] = b;
        }
    for (size_t v = 0; v < nv; ++v)
        if (!(beta[v] >= 0.0)) return (int)v + 1;
    return 0;
}

/*
 * Grounded fraction at the vertices (glissade_grounded_fraction,
 * glissade_grounding_line.F90:59-674) for which_ho_ground 0 (no GLP: the
 * grounded-neighbour rule or the sign of the staggered flotation function,
 * which_ho_fground_no_glp 0/1), 1 (GLP for basal friction: the grounded
 * area of the bilinear interpolant of the flotation function over the
 * staggered cell, compute_grounded_fraction :678-1080) and 2 (the same
 * over the cell's four quadrants, :521-618), with the flotation
 * functions 0 (Pattyn), 1 (inverse Pattyn) and 2 (linear) and, for 0 and 1,
 * their extrapolation to ice-free ocean cells (:237-392).
 */

/* the grounded area of the unit square where the bilinear interpolant of
 * the corner values f[0..3] (SW, SE, NE, NW) is <= 0 (:678-1080); returns
 * 0, or -1 for the two patterns the reference stops on */
int orc_compute_grounded_fraction(const double* fq, double* f_ground)
{
    const double eps06 = 1.0e-6;
    int cfloat[4], nfloat = 0;
    double fmin = fq[0];
    for (int n = 0; n < 4; ++n) {
        cfloat[n] = fq[n] > 0.0;
        nfloat += cfloat[n];
        if (fq[n] < fmin) fmin = fq[n];
    }
    if (nfloat == 0) { *f_ground = 1.0; return 0; }
    if (nfloat == 4 || fmin == 0.0) { *f_ground = 0.0; return 0; }
    double f1, f2, f3, f4;
    /* the corner values rolled so that corner r (0-based) comes first */
#define ROLL(r) do { f1 = fq[(r) % 4]; f2 = fq[((r) + 1) % 4]; f3 = fq[((r) + 2) % 4]; f4 = fq[((r) + 3) % 4]; } while (0)
    if (nfloat == 1 || nfloat == 3) {
        int lv[4];
        for (int n = 0; n < 4; ++n) lv[n] = nfloat == 1 ? cfloat[n] : !cfloat[n];
        if (lv[0]) ROLL(0); else if (lv[1]) ROLL(1); else if (lv[2]) ROLL(2); else ROLL(3);
        const double a = f1, b = f2 - f1, c = f4 - f1, d = f1 + f3 - f2 - f4;
        double f_corner;
        if (fabs((a * d) / (b * c)) > eps06) f_corner = ((b * c - a * d) * log(fabs(1.0 - (a * d) / (b * c))) + a * d) / (d * d);
        else f_corner = (a * a) / (2.0 * b * c);
        *f_ground = nfloat == 1 ? 1.0 - f_corner : f_corner;
        return 0;
    }
    /* nfloat == 2 */
    int adjacent = 1, rotated = 0;
    if (cfloat[0] && cfloat[1]) ROLL(0);
    else if (cfloat[1] && cfloat[2]) ROLL(1);
    else if (cfloat[2] && cfloat[3]) ROLL(2);
    else if (cfloat[3] && cfloat[0]) ROLL(3);
    else {
        adjacent = 0;
        const double var = fq[1] * fq[3] - fq[0] * fq[2];
        if (var >= 0.0) ROLL(0);
        else { ROLL(1); rotated = 1; }
    }
#undef ROLL
    const double a = f1, b = f2 - f1, c = f4 - f1, d = f1 + f3 - f2 - f4;
    if (adjacent) {
        double f_trap;
        if (fabs(d / c) > eps06) f_trap = ((b * c - a * d) * log(1.0 + d / c) - b * d) / (d * d);
        else f_trap = -(2.0 * a + b) / (2.0 * c);
        *f_ground = 1.0 - f_trap;
        return 0;
    }
    if (b * c - a * d < 0.0 || (b + d) * (c + d) < 0.0) return -1;
    double c1, c2;
    if (fabs(b * c - a * d) > eps06) {
        c1 = ((b * c - a * d) * log(1.0 - (a * d) / (b * c)) + a * d) / (d * d);
        c2 = ((b * c - a * d) * log((b * c - a * d) / ((b + d) * (c + d))) + d * (a + b + c + d)) / (d * d);
    } else {
        c1 = a * a / (b * c);
        c2 = (a + b) * (a + c) / (b * c);
    }
    const int corners_grounded = (fq[0] > 0.0) ? rotated : !rotated;
    *f_ground = corners_grounded ? c1 + c2 : 1.0 - (c1 + c2);
    return 0;
}

/* flotation function at the cells (:237-392); ice-free ocean cells of
 * options 0 and 1 take the smallest value of their ice or land edge
 * neighbours (else corner neighbours), then the cell halo.  Option 3
 * (:299-326) is option 2's formula on the raised bed, which the caller
 * passes as topg. */
void orc_flotation_function(int nx, int ny, int global_bc, int which, const double* thck, const double* topg,
                            double eus, const int32_t* ice_mask, const int32_t* land_mask, double* ff)
{
    const double land_topg_min = 50.0, marine_min = 1.0e-4, land_pattyn = -10.0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t c = C2(i, j);
            double f;
            if (which == 0 || which == 1) {
                if (land_mask[c] == 1) f = land_pattyn;
                else if (ice_mask[c] == 1) {
                    if (which == 0) f = -RHOO_B * (topg[c] - eus) / (RHOI_B * thck[c]) - 1.0;
                    else {
                        f = 1.0 - RHOI_B * thck[c] / (-RHOO_B * (topg[c] - eus));
                        if (f < land_pattyn) f = land_pattyn;
                    }
                } else f = 0.0;
            } else {
                if (land_mask[c] == 1) {
                    const double dd = (topg[c] - eus) > land_topg_min ? (topg[c] - eus) : land_topg_min;
                    f = -dd;
                } else {
                    f = -(topg[c] - eus) - (RHOI_B / RHOO_B) * thck[c];
                    if (fabs(f) < marine_min) f = f < 0.0 ? -marine_min : marine_min;
                }
            }
            ff[c] = f;
        }
    if (which != 0 && which != 1) return;
    double* ex = (double*)malloc(sizeof(double) * (size_t)nx * (size_t)ny);
    for (size_t c = 0; c < (size_t)nx * (size_t)ny; ++c) ex[c] = ff[c];
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            if (ice_mask[C2(i, j)] != 0 || land_mask[C2(i, j)] != 0) continue;
            int filled = 0;
            for (int pass = 0; pass < 2 && !filled; ++pass)        /* edges, then corners */
                for (int jj = j - 1; jj <= j + 1; ++jj)
                    for (int ii = i - 1; ii <= i + 1; ++ii) {
                        const int edge = (ii == i || jj == j), corner = (abs(ii - i) == 1 && abs(jj - j) == 1);
                        if (!(pass == 0 ? edge : corner)) continue;
                        if (!(ice_mask[C2(ii, jj)] == 1 || land_mask[C2(ii, jj)] == 1)) continue;
                        const double v = ff[C2(ii, jj)];
                        if (!filled) { filled = 1; ex[C2(i, j)] = v; }
                        else if (v < ex[C2(i, j)]) ex[C2(i, j)] = v;
                    }
        }
    orc_cell_halo(nx, ny, global_bc, ex);
    for (size_t c = 0; c < (size_t)nx * (size_t)ny; ++c) ff[c] = ex[c];
    free(ex);
}

/* f_ground at the vertices (:404-674, which_ho_ground 0, 1, 2); returns 0,
 * -1 for an unsupported option, or 1 + the vertex index of a pattern the
 * reference stops on */
int orc_grounded_fraction(int nx, int ny, int global_bc, int which_ground, int which_flotation, int which_no_glp,
                          const double* thck, const double* topg, double eus, const int32_t* ice_mask,
                          const int32_t* floating_mask, const int32_t* land_mask, double* ff, double* f_ground)
{
    if (which_ground < 0 || which_ground > 2 || which_flotation < 0 || which_flotation > 3) return -1;
    orc_flotation_function(nx, ny, global_bc, which_flotation, thck, topg, eus, ice_mask, land_mask, ff);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            /* vertex mask: the reference tests land cell (i+1, j+1) twice (:218-229) */
            const int vmask = ice_mask[C2(i, j + 1)] == 1 || ice_mask[C2(i + 1, j + 1)] == 1 ||
                              ice_mask[C2(i, j)] == 1 || ice_mask[C2(i + 1, j)] == 1 ||
                              land_mask[C2(i, j + 1)] == 1 || land_mask[C2(i + 1, j + 1)] == 1 ||
                              land_mask[C2(i, j)] == 1 || land_mask[C2(i + 1, j + 1)] == 1;
            double f = 0.0;
            if (vmask) {
                if (which_ground == 0 && which_no_glp == 1) {
                    const double s = (ff[C2(i, j + 1)] + ff[C2(i + 1, j + 1)] + ff[C2(i, j)] + ff[C2(i + 1, j)]) / 4.0;
                    f = s <= 0.0 ? 1.0 : 0.0;
                } else if (which_ground == 0) {
#define CGR(a, b) ((ice_mask[C2(a, b)] == 1 && floating_mask[C2(a, b)] == 0) || land_mask[C2(a, b)] == 1)
                    f = (CGR(i, j + 1) || CGR(i + 1, j + 1) || CGR(i, j) || CGR(i + 1, j)) ? 1.0 : 0.0;
#undef CGR
                } else if (which_ground == 1) {
                    const double fq[4] = {ff[C2(i, j)], ff[C2(i + 1, j)], ff[C2(i + 1, j + 1)], ff[C2(i, j + 1)]};
                    if (orc_compute_grounded_fraction(fq, &f)) return 1 + (int)V2(i, j);
                } else {
                    /* GLP deluxe (:521-618): the staggered cell's four quadrants, each
                     * bilinear between a cell value, two edge midpoints and the vertex */
                    const double sw = ff[C2(i, j)], se = ff[C2(i + 1, j)], ne = ff[C2(i + 1, j + 1)], nw = ff[C2(i, j + 1)];
                    const double ctr = 0.25 * (nw + ne + sw + se);
                    const double q[4][4] = {{sw, 0.5 * (sw + se), ctr, 0.5 * (sw + nw)},
                                            {0.5 * (se + sw), se, 0.5 * (se + ne), ctr},
                                            {ctr, 0.5 * (ne + se), ne, 0.5 * (ne + nw)},
                                            {0.5 * (nw + sw), ctr, 0.5 * (nw + ne), nw}};
                    double fgq[4];
                    for (int r = 0; r < 4; ++r)
                        if (orc_compute_grounded_fraction(q[r], &fgq[r])) return 1 + (int)V2(i, j);
                    f = 0.25 * (fgq[0] + fgq[1] + fgq[2] + fgq[3]);
                }
            }
            f_ground[V2(i, j)] = f;
        }
    return 0;
}
