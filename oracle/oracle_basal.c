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

#include "cism_oracle_asm.h"

#define NVX (nx - 1)
#define C2(i, j) ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1))
#define V2(i, j) ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1))

/* glimmer_physcon.F90:57 */
static const double ORC_PI = 3.14159265358979;

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
