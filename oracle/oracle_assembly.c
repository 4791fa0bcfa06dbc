/*
 * oracle_assembly.c -- CPU restatement of the 3-D Blatter-Pattyn assembly
 * path of libglissade/glissade_velo_higher.F90 and of the 2-D input
 * operators it calls (glissade_grid_operators.F90, glissade_masks.F90,
 * glide_thck.F90).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP product).  Layouts as
 * in cism_oracle.h; each function cites the reference lines it follows.
 *
 * Pinned by tests/test_oracle_pins.py (closed forms, invariants, brute
 * force, the reference's slab solution and EISMINT-2 files) except, parity
 * unpinned: the vertex-mask rule of orc_grounded_fraction_no_glp (land
 * cell (i+1, j+1) tested twice, (i+1, j) never, as the reference) and the
 * thickness ramp branch (thck_gradient_ramp > 0) of
 * orc_surface_elevation_gradient.
 */
#include "cism_oracle.h"
#include "cism_oracle_asm.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NVX (nx - 1)
#define NVY (ny - 1)
#define C2(i, j) ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1))
#define V2(i, j) ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1))
#define V3(k, i, j) ((size_t)((k) - 1) + (size_t)nz * ((size_t)((i) - 1) + (size_t)NVX * (size_t)((j) - 1)))
#define A4(m, k, i, j) ((size_t)((m) - 1) + 27 * V3(k, i, j))
#define E3(k, i, j) ((size_t)((k) - 1) + (size_t)(nz - 1) * ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1)))
#define IDX3(iA, jA, kA) (1 + ((iA) + 1) + 3 * ((jA) + 1) + 9 * ((kA) + 1))

static const double RHOI = 910.0, RHOO = 1028.0, GRAV = 9.81, VOL0 = 1.0e9;
static const double EPS11 = 1.0e-11;

/* ------------------------------------------------------------------ */
/* glissade_velo_higher_init (glissade_velo_higher.F90:278-658)        */
/* Tables are stored [n][p] for Fortran (n,p).                          */
/* ------------------------------------------------------------------ */
static double phi3[8][8], dxr3[8][8], dyr3[8][8], dzr3[8][8];
static int node_di[8], node_dj[8], node_dk[8];   /* ishift/jshift/kshift(7,n) */
static int tables_ready = 0;

static void init_tables(void)
{
    if (tables_ready) return;
    const double r = 1.0 / sqrt(3.0);
    const double xq[8] = {-r, r, r, -r, -r, r, r, -r};
    const double yq[8] = {-r, -r, r, r, -r, -r, r, r};
    const double zq[8] = {-r, -r, -r, -r, r, r, r, r};
    /* node signs of the trilinear basis (:352-388) */
    const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int p = 0; p < 8; ++p)
        for (int n = 0; n < 8; ++n) {
            phi3[n][p] = (1.0 + sx[n] * xq[p]) * (1.0 + sy[n] * yq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
            dxr3[n][p] = sx[n] * (1.0 + sy[n] * yq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
            dyr3[n][p] = sy[n] * (1.0 + sx[n] * xq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
            dzr3[n][p] = sz[n] * (1.0 + sx[n] * xq[p]) * (1.0 + sy[n] * yq[p]) / 8.0;
        }
    /* offsets of node n from node 7 = (k,i,j) of the element (:461-494) */
    const int di[8] = {-1, 0, 0, -1, -1, 0, 0, -1};
    const int dj[8] = {-1, -1, 0, 0, -1, -1, 0, 0};
    const int dk[8] = {1, 1, 1, 1, 0, 0, 0, 0};
    for (int n = 0; n < 8; ++n) { node_di[n] = di[n]; node_dj[n] = dj[n]; node_dk[n] = dk[n]; }
    tables_ready = 1;
}

void orc_fe_tables(double* phi, double* dxr, double* dyr, double* dzr)
{
    init_tables();
    for (int n = 0; n < 8; ++n)
        for (int p = 0; p < 8; ++p) {
            phi[n + 8 * p] = phi3[n][p]; dxr[n + 8 * p] = dxr3[n][p];
            dyr[n + 8 * p] = dyr3[n][p]; dzr[n + 8 * p] = dzr3[n][p];
        }
}

/* ------------------------------------------------------------------ */
/* 2-D input operators                                                  */
/* ------------------------------------------------------------------ */

/* glide_calclsrf (glide_thck.F90:1053-1076) */
void orc_calclsrf(int nx, int ny, const double* thck, const double* topg, double eus, double* lsrf)
{
    for (size_t p = 0; p < (size_t)nx * ny; ++p)
        lsrf[p] = (topg[p] - eus < (-RHOI / RHOO) * thck[p]) ? (-RHOI / RHOO) * thck[p] : topg[p];
}

/* glissade_get_masks (glissade_masks.F90:70-210), the four masks used here */
void orc_get_masks(int nx, int ny, const double* thck, const double* topg, double eus, double thklim,
                   int32_t* ice_mask, int32_t* floating_mask, int32_t* ocean_mask, int32_t* land_mask)
{
    for (size_t p = 0; p < (size_t)nx * ny; ++p) {
        ice_mask[p] = thck[p] > thklim ? 1 : 0;
        if (ocean_mask) ocean_mask[p] = (topg[p] < eus && ice_mask[p] == 0) ? 1 : 0;
        if (floating_mask) floating_mask[p] = (topg[p] - eus < (-RHOI / RHOO) * thck[p] && ice_mask[p] == 1) ? 1 : 0;
        if (land_mask) land_mask[p] = topg[p] >= eus ? 1 : 0;
    }
}

/* glissade_stagger, stagger_margin = 0 (glissade_grid_operators.F90:125-133) */
void orc_stagger(int nx, int ny, const double* var, double* stagvar)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i)
            stagvar[V2(i, j)] = (var[C2(i, j + 1)] + var[C2(i + 1, j + 1)] + var[C2(i, j)] + var[C2(i + 1, j)]) / 4.0;
}

/* glissade_surface_elevation_gradient (glissade_grid_operators.F90:834-1260)
 * for ho_gradient = centered, all three margin options, max_slope clip. */
void orc_surface_elevation_gradient(int nx, int ny, double dx, double dy,
                                    const int32_t* ice_mask, const int32_t* land_mask,
                                    const double* usrf, const double* thck,
                                    double thklim, double thck_gradient_ramp, int margin,
                                    double max_slope, double* ds_dx, double* ds_dy)
{
    double* ex = (double*)calloc((size_t)nx * ny, sizeof(double));
    double* ey = (double*)calloc((size_t)nx * ny, sizeof(double));
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (margin == 0) {
                ex[C2(i, j)] = (usrf[C2(i + 1, j)] - usrf[C2(i, j)]) / dx;
            } else if (margin == 2) {
                if (ice_mask[C2(i, j)] == 1 && ice_mask[C2(i + 1, j)] == 1)
                    ex[C2(i, j)] = (usrf[C2(i + 1, j)] - usrf[C2(i, j)]) / dx;
            } else {
                int iu, il; double sgn;
                if (usrf[C2(i, j)] > usrf[C2(i + 1, j)]) { iu = i; il = i + 1; sgn = -1.0; }
                else { iu = i + 1; il = i; sgn = 1.0; }
                double ef = 1.0;
                if (land_mask[C2(iu, j)] == 1 && thck_gradient_ramp > 0.0) {
                    ef = fmin(1.0, (thck[C2(iu, j)] - thklim) / thck_gradient_ramp);
                    ef = fmax(ef, 0.0);
                }
                if ((ice_mask[C2(iu, j)] == 1 && ice_mask[C2(il, j)] == 1) ||
                    (ice_mask[C2(iu, j)] == 1 && land_mask[C2(il, j)] == 1))
                    ex[C2(i, j)] = ef * sgn * (usrf[C2(iu, j)] - usrf[C2(il, j)]) / dx;
            }
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx; ++i) {
            if (margin == 0) {
                ey[C2(i, j)] = (usrf[C2(i, j + 1)] - usrf[C2(i, j)]) / dy;
            } else if (margin == 2) {
                if (ice_mask[C2(i, j)] == 1 && ice_mask[C2(i, j + 1)] == 1)
                    ey[C2(i, j)] = (usrf[C2(i, j + 1)] - usrf[C2(i, j)]) / dy;
            } else {
                int ju, jl; double sgn;
                if (usrf[C2(i, j)] > usrf[C2(i, j + 1)]) { ju = j; jl = j + 1; sgn = -1.0; }
                else { ju = j + 1; jl = j; sgn = 1.0; }
                double ef = 1.0;
                if (land_mask[C2(i, ju)] == 1 && thck_gradient_ramp > 0.0) {
                    ef = fmin(1.0, (thck[C2(i, ju)] - thklim) / thck_gradient_ramp);
                    ef = fmax(ef, 0.0);
                }
                if ((ice_mask[C2(i, ju)] == 1 && ice_mask[C2(i, jl)] == 1) ||
                    (ice_mask[C2(i, ju)] == 1 && land_mask[C2(i, jl)] == 1))
                    ey[C2(i, j)] = ef * sgn * (usrf[C2(i, ju)] - usrf[C2(i, jl)]) / dy;
            }
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            ds_dx[V2(i, j)] = 0.5 * (ex[C2(i, j)] + ex[C2(i, j + 1)]);
            ds_dy[V2(i, j)] = 0.5 * (ey[C2(i, j)] + ey[C2(i + 1, j)]);
        }
    if (max_slope > 0.0)
        for (size_t p = 0; p < (size_t)(nx - 1) * (ny - 1); ++p) {
            ds_dx[p] = ds_dx[p] > 0.0 ? fmin(ds_dx[p], max_slope) : fmax(ds_dx[p], -max_slope);
            ds_dy[p] = ds_dy[p] > 0.0 ? fmin(ds_dy[p], max_slope) : fmax(ds_dy[p], -max_slope);
        }
    free(ex); free(ey);
}

/* glissade_gradient with gradient_margin = 1 and a full mask
 * (glissade_grid_operators.F90:309-468), as called by glissade_slope_angle:
 * with slope_mask = 1 everywhere the edge gradients are plain differences. */
static void gradient_full(int nx, int ny, double dx, double dy, const double* f, double* gx, double* gy)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const double n_ = (f[C2(i + 1, j + 1)] - f[C2(i, j + 1)]) / dx;
            const double s_ = (f[C2(i + 1, j)] - f[C2(i, j)]) / dx;
            gx[V2(i, j)] = (n_ + s_) / 2.0;
            const double e_ = (f[C2(i + 1, j + 1)] - f[C2(i + 1, j)]) / dy;
            const double w_ = (f[C2(i, j + 1)] - f[C2(i, j)]) / dy;
            gy[V2(i, j)] = (e_ + w_) / 2.0;
        }
}

/* glissade_slope_angle (glissade_grid_operators.F90:1264-1395): gradient at
 * vertices, unstagger to centres (:226-305, margin 0), theta = atan|grad|.
 * Called with the vertex grid as its "cell" grid by basal_sliding_bc_3d. */
void orc_slope_angle(int nx, int ny, double dx, double dy, const double* zsfc, double* theta)
{
    double* gx = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    double* gy = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    gradient_full(nx, ny, dx, dy, zsfc, gx, gy);
    for (size_t p = 0; p < (size_t)nx * ny; ++p) theta[p] = 0.0;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            double dzx = (gx[V2(i, j)] + gx[V2(i - 1, j)] + gx[V2(i, j - 1)] + gx[V2(i - 1, j - 1)]) / 4.0;
            double dzy = (gy[V2(i, j)] + gy[V2(i - 1, j)] + gy[V2(i, j - 1)] + gy[V2(i - 1, j - 1)]) / 4.0;
            double s = sqrt(dzx * dzx + dzy * dzy);
            theta[C2(i, j)] = (s >= 0.0) ? atan(s) : 0.0;
        }
    free(gx); free(gy);
}

/* get_vertex_geometry (glissade_velo_higher.F90:4052-4224), active masks */
void orc_vertex_masks(int nx, int ny, int nhalo, const int32_t* ice_mask,
                      int32_t* active_cell, int32_t* active_vertex)
{
    memset(active_cell, 0, (size_t)nx * ny * sizeof(int32_t));
    memset(active_vertex, 0, (size_t)(nx - 1) * (ny - 1) * sizeof(int32_t));
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i)
            if (ice_mask[C2(i, j)] == 1) active_cell[C2(i, j)] = 1;
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i)
            if (active_cell[C2(i, j)]) {
                active_vertex[V2(i - 1, j - 1)] = 1; active_vertex[V2(i, j - 1)] = 1;
                active_vertex[V2(i - 1, j)] = 1; active_vertex[V2(i, j)] = 1;
            }
}

/* flwafact = 0.5 * flwa^(-1/n) (glissade_velo_higher.F90:1838-1847);
 * cells outside the loop range keep 0. */
void orc_flwafact(int nx, int ny, int nz, int nhalo, double n_glen, const double* flwa, double* flwafact)
{
    memset(flwafact, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
    for (int j = 1 + nhalo; j <= ny - nhalo + 1; ++j)
        for (int i = 1 + nhalo; i <= nx - nhalo + 1; ++i)
            for (int k = 1; k <= nz - 1; ++k)
                if (flwa[E3(k, i, j)] > 0.0) flwafact[E3(k, i, j)] = 0.5 * pow(flwa[E3(k, i, j)], -1.0 / n_glen);
}

/* ------------------------------------------------------------------ */
/* load_vector_gravity, local option (glissade_velo_higher.F90:4311-4342) */
/* ------------------------------------------------------------------ */
void orc_load_vector_gravity_local(int nx, int ny, int nz, const double* stagwbndsigma /* 0:nz */,
                                   double dx, double dy, const int32_t* active_vertex,
                                   const double* stagthck, const double* dusrf_dx, const double* dusrf_dy,
                                   double* loadu, double* loadv)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            for (int k = 1; k <= nz; ++k) {
                double dz = stagthck[V2(i, j)] * (stagwbndsigma[k] - stagwbndsigma[k - 1]);
                loadu[V3(k, i, j)] = loadu[V3(k, i, j)] - RHOI * GRAV * dx * dy * dz / VOL0 * dusrf_dx[V2(i, j)];
                loadv[V3(k, i, j)] = loadv[V3(k, i, j)] - RHOI * GRAV * dx * dy * dz / VOL0 * dusrf_dy[V2(i, j)];
            }
        }
}

/* ------------------------------------------------------------------ */
/* Element kernels                                                      */
/* ------------------------------------------------------------------ */

/* get_basis_function_derivatives_3d (glissade_velo_higher.F90:6508-6732) */
static int basis_derivs(const double* x, const double* y, const double* z, int p,
                        double* dx, double* dy, double* dz, double* detJ)
{
    double J[3][3] = {{0}};
    for (int n = 0; n < 8; ++n) {
        J[0][0] = J[0][0] + dxr3[n][p] * x[n];
        J[0][1] = J[0][1] + dxr3[n][p] * y[n];
        J[0][2] = J[0][2] + dxr3[n][p] * z[n];
        J[1][0] = J[1][0] + dyr3[n][p] * x[n];
        J[1][1] = J[1][1] + dyr3[n][p] * y[n];
        J[1][2] = J[1][2] + dyr3[n][p] * z[n];
        J[2][0] = J[2][0] + dzr3[n][p] * x[n];
        J[2][1] = J[2][1] + dzr3[n][p] * y[n];
        J[2][2] = J[2][2] + dzr3[n][p] * z[n];
    }
    double c[3][3];
    c[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    c[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
    c[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    c[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
    c[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    c[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
    c[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    c[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
    c[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double d = J[0][0] * c[0][0] + J[0][1] * c[0][1] + J[0][2] * c[0][2];
    *detJ = d;
    if (!(fabs(d) > 0.0)) return ORC_ERR_SINGULAR;
    double Ji[3][3];
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) Ji[row][col] = c[col][row] / d;
    for (int n = 0; n < 8; ++n) {
        dx[n] = Ji[0][0] * dxr3[n][p] + Ji[0][1] * dyr3[n][p] + Ji[0][2] * dzr3[n][p];
        dy[n] = Ji[1][0] * dxr3[n][p] + Ji[1][1] * dyr3[n][p] + Ji[1][2] * dzr3[n][p];
        dz[n] = Ji[2][0] * dxr3[n][p] + Ji[2][1] * dyr3[n][p] + Ji[2][2] * dzr3[n][p];
    }
    return ORC_OK;
}

/* compute_effective_viscosity, BP branch (glissade_velo_higher.F90:7285-7470) */
static double effective_viscosity(int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                                  const double* dx, const double* dy, const double* dz,
                                  const double* u, const double* v, double flwafact)
{
    const double p_effstr = (1.0 - n_glen) / n_glen;
    if (whichefvs == 0) return efvs_constant;
    if (whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;   /* velo_scale/len_scale*scyr */
        return flwafact * pow(effstrain, p_effstr);
    }
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
    for (int n = 0; n < 8; ++n) {
        ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n]; uz = uz + dz[n] * u[n];
        vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n]; vz = vz + dz[n] * v[n];
    }
    double t1 = ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy));
    double t2 = 0.25 * (uz * uz + vz * vz);
    double esq = effstrain_min * effstrain_min + t1 + t2;
    return flwafact * pow(esq, p_effstr / 2.0);
}

/* One element's stiffness blocks: quadrature loop of
 * assemble_stiffness_matrix_3d (:5562-5602) with compute_element_matrix
 * (:8001-8123, BP branch).  K are [nr][nc]. */
int orc_element_matrix(const double* x, const double* y, const double* z, const double* u, const double* v,
                       double flwafact, int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                       double Kuu[8][8], double Kuv[8][8], double Kvu[8][8], double Kvv[8][8], double* efvs_mean)
{
    init_tables();
    double dx[8], dy[8], dz[8], detJ, efvs_qp[8];
    memset(Kuu, 0, 64 * sizeof(double)); memset(Kuv, 0, 64 * sizeof(double));
    memset(Kvu, 0, 64 * sizeof(double)); memset(Kvv, 0, 64 * sizeof(double));
    for (int p = 0; p < 8; ++p) {
        int st = basis_derivs(x, y, z, p, dx, dy, dz, &detJ);
        if (st) return st;
        efvs_qp[p] = effective_viscosity(whichefvs, efvs_constant, effstrain_min, n_glen, dx, dy, dz, u, v, flwafact);
        const double f = efvs_qp[p] * 1.0 * detJ / VOL0;
        for (int nc = 0; nc < 8; ++nc)
            for (int nr = 0; nr < 8; ++nr) {
                Kuu[nr][nc] = Kuu[nr][nc] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc] + dz[nr] * dz[nc]);
                Kuv[nr][nc] = Kuv[nr][nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                Kvu[nr][nc] = Kvu[nr][nc] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                Kvv[nr][nc] = Kvv[nr][nc] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc] + dz[nr] * dz[nc]);
            }
    }
    double s = 0.0;
    for (int p = 0; p < 8; ++p) s = s + efvs_qp[p];
    *efvs_mean = s / 8.0;
    /* check_symmetry_element_matrix (:10583-10644) */
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            if (fabs(Kuu[a][b] - Kuu[b][a]) > EPS11) return ORC_ERR_ASYM;
            if (fabs(Kvv[a][b] - Kvv[b][a]) > EPS11) return ORC_ERR_ASYM;
            if (fabs(Kuv[a][b] - Kvu[b][a]) > EPS11) return ORC_ERR_ASYM;
        }
    return ORC_OK;
}

/* compute_internal_stress (glissade_velo_higher.F90:7046-7281) for
 * whichapprox (HO_APPROX_SIA 0: vertical shear only; SSA 1: horizontal
 * strain rates only, viscosity of the SSA effective strain; BP 2 and DIVA 4:
 * every strain rate and the 3-D viscosity; L1L2 3: every strain rate, the
 * viscosity efvs(k,i,j) given): tau (6 fields of (nz-1, nx, ny): xz, yz,
 * xx, yy, xy, eff), zero outside the active cells of the assembly range.
 * The viscosity is compute_effective_viscosity's (:7285-7470) branch for
 * whichapprox, from the eight nodes. */
int orc_compute_internal_stress_approx(int nx, int ny, int nz, const double* sigma, int nhalo,
                                       const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                       const double* stagusrf, const double* stagthck, int whichapprox,
                                       const double* flwafact, int whichefvs, const double* efvs,
                                       double efvs_constant, double effstrain_min, double n_glen, const double* uvel,
                                       const double* vvel, double* tau)
{
    init_tables();
    const size_t nfl = (size_t)(nz - 1) * nx * ny;
    memset(tau, 0, 6 * nfl * sizeof(double));
    double x[8], y[8], z[8], u[8], v[8], dx[8], dy[8], dz[8], detJ;
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                for (int n = 0; n < 8; ++n) {
                    const int in = i + node_di[n], jn = j + node_dj[n], kn = k + node_dk[n];
                    x[n] = xVertex[V2(in, jn)];
                    y[n] = yVertex[V2(in, jn)];
                    z[n] = stagusrf[V2(in, jn)] - sigma[kn - 1] * stagthck[V2(in, jn)];
                    u[n] = uvel[V3(kn, in, jn)];
                    v[n] = vvel[V3(kn, in, jn)];
                }
                double txz = 0.0, tyz = 0.0, txx = 0.0, tyy = 0.0, txy = 0.0;
                for (int p = 0; p < 8; ++p) {
                    if (basis_derivs(x, y, z, p, dx, dy, dz, &detJ)) return ORC_ERR_SINGULAR;
                    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
                    for (int n = 0; n < 8; ++n) {
                        if (whichapprox != 1) { uz = uz + dz[n] * u[n]; vz = vz + dz[n] * v[n]; }
                        if (whichapprox != 0) {
                            ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
                            vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
                        }
                    }
                    double ef;
                    const double fl = flwafact[E3(k, i, j)];
                    if (whichapprox == 3) {
                        ef = efvs[E3(k, i, j)];
                    } else if (whichefvs == 2 && whichapprox == 0) {   /* :7406-7415 */
                        const double esq = effstrain_min * effstrain_min + 0.25 * (uz * uz + vz * vz);
                        ef = fl * pow(esq, (1.0 - n_glen) / n_glen / 2.0);
                    } else if (whichefvs == 2 && whichapprox == 1) {   /* :7416-7429 */
                        const double esq = effstrain_min * effstrain_min +
                                           (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
                        ef = fl * pow(esq, (1.0 - n_glen) / n_glen / 2.0);
                    } else {
                        ef = effective_viscosity(whichefvs, efvs_constant, effstrain_min, n_glen, dx, dy, dz, u, v, fl);
                    }
                    txz = txz + ef * uz;
                    tyz = tyz + ef * vz;
                    txx = txx + 2.0 * ef * ux;
                    tyy = tyy + 2.0 * ef * vy;
                    txy = txy + ef * (vx + uy);
                }
                txz = txz / 8.0; tyz = tyz / 8.0; txx = txx / 8.0; tyy = tyy / 8.0; txy = txy / 8.0;
                const size_t e = E3(k, i, j);
                tau[e] = txz; tau[nfl + e] = tyz; tau[2 * nfl + e] = txx; tau[3 * nfl + e] = tyy;
                tau[4 * nfl + e] = txy;
                tau[5 * nfl + e] = sqrt(txx * txx + tyy * tyy + txx * tyy + txy * txy + txz * txz + tyz * tyz);
            }
        }
    return ORC_OK;
}

/* the BP branch of compute_internal_stress */
int orc_compute_internal_stress(int nx, int ny, int nz, const double* sigma, int nhalo, const int32_t* active_cell,
                                const double* xVertex, const double* yVertex, const double* stagusrf,
                                const double* stagthck, const double* flwafact, int whichefvs, double efvs_constant,
                                double effstrain_min, double n_glen, const double* uvel, const double* vvel,
                                double* tau)
{
    return orc_compute_internal_stress_approx(nx, ny, nz, sigma, nhalo, active_cell, xVertex, yVertex, stagusrf,
                                              stagthck, 2, flwafact, whichefvs, NULL, efvs_constant, effstrain_min,
                                              n_glen, uvel, vvel, tau);
}

/* assemble_stiffness_matrix_3d (glissade_velo_higher.F90:5382-5643) with
 * element_to_global_matrix_3d (:8200-8268). */
int orc_assemble_stiffness_matrix_3d(int nx, int ny, int nz, const double* sigma, int nhalo,
                                     const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                     const double* uvel, const double* vvel,
                                     const double* stagusrf, const double* stagthck, const double* flwafact,
                                     int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                                     double* efvs, double* Auu, double* Auv, double* Avu, double* Avv)
{
    init_tables();
    const size_t nA = (size_t)27 * nz * (nx - 1) * (ny - 1);
    memset(efvs, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
    memset(Auu, 0, nA * sizeof(double)); memset(Auv, 0, nA * sizeof(double));
    memset(Avu, 0, nA * sizeof(double)); memset(Avv, 0, nA * sizeof(double));
    double Kuu[8][8], Kuv[8][8], Kvu[8][8], Kvv[8][8];
    double x[8], y[8], z[8], u[8], v[8];
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                for (int n = 0; n < 8; ++n) {
                    const int in = i + node_di[n], jn = j + node_dj[n], kn = k + node_dk[n];
                    x[n] = xVertex[V2(in, jn)];
                    y[n] = yVertex[V2(in, jn)];
                    z[n] = stagusrf[V2(in, jn)] - sigma[kn - 1] * stagthck[V2(in, jn)];
                    u[n] = uvel[V3(kn, in, jn)];
                    v[n] = vvel[V3(kn, in, jn)];
                }
                double em;
                int st = orc_element_matrix(x, y, z, u, v, flwafact[E3(k, i, j)], whichefvs, efvs_constant,
                                            effstrain_min, n_glen, Kuu, Kuv, Kvu, Kvv, &em);
                if (st) return st;
                efvs[E3(k, i, j)] = em;
                for (int nr = 0; nr < 8; ++nr) {
                    const int ir = i + node_di[nr], jr = j + node_dj[nr], kr = k + node_dk[nr];
                    for (int nc = 0; nc < 8; ++nc) {
                        const int m = IDX3(node_di[nc] - node_di[nr], node_dj[nc] - node_dj[nr], node_dk[nc] - node_dk[nr]);
                        const size_t a = A4(m, kr, ir, jr);
                        Auu[a] = Auu[a] + Kuu[nr][nc];
                        Auv[a] = Auv[a] + Kuv[nr][nc];
                        Avu[a] = Avu[a] + Kvu[nr][nc];
                        Avv[a] = Avv[a] + Kvv[nr][nc];
                    }
                }
            }
        }
    return ORC_OK;
}

/* basal_sliding_bc_3d, local option (glissade_velo_higher.F90:8921-8953).
 * theta_basal_slope is glissade_slope_angle of staglsrf on the vertex grid
 * (orc_slope_angle with nx-1, ny-1); its parallel_halo call is a no-op for
 * vertex-sized arrays (parallel_mpi.F90:7769). */
void orc_basal_sliding_bc_local(int nx, int ny, int nz, double dx, double dy,
                                const int32_t* active_vertex, const double* beta,
                                const double* theta_basal_slope, double* Auu, double* Avv)
{
    const int m = IDX3(0, 0, 0);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            const size_t a = A4(m, nz, i, j);
            const double c = cos(theta_basal_slope[V2(i, j)]);
            Auu[a] = Auu[a] + dx * dy / VOL0 * beta[V2(i, j)] / c;
            Avv[a] = Avv[a] + dx * dy / VOL0 * beta[V2(i, j)] / c;
        }
}

/* dirichlet_boundary_conditions_3d (glissade_velo_higher.F90:9305-9498) */
void orc_dirichlet_bc_3d(int nx, int ny, int nz, int nhalo, const int32_t* active_vertex,
                         const int32_t* umask, const int32_t* vmask, const double* uvel, const double* vvel,
                         double* Auu, double* Auv, double* Avu, double* Avv, double* bu, double* bv)
{
    const int md = IDX3(0, 0, 0);
    for (int j = nhalo; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo; i <= nx - nhalo + 1; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            for (int k = 1; k <= nz; ++k) {
                if (umask[V3(k, i, j)] == 1) {
                    bu[V3(k, i, j)] = uvel[V3(k, i, j)];
                    for (int kA = -1; kA <= 1; ++kA)
                        for (int jA = -1; jA <= 1; ++jA)
                            for (int iA = -1; iA <= 1; ++iA) {
                                if (!(k + kA >= 1 && k + kA <= nz && i + iA >= 1 && i + iA <= nx - 1 &&
                                      j + jA >= 1 && j + jA <= ny - 1)) continue;
                                if (iA == 0 && jA == 0 && kA == 0) {
                                    Auu[A4(md, k, i, j)] = 1.0; Auv[A4(md, k, i, j)] = 0.0; Avu[A4(md, k, i, j)] = 0.0;
                                } else {
                                    int m = IDX3(iA, jA, kA);
                                    Auu[A4(m, k, i, j)] = 0.0; Auv[A4(m, k, i, j)] = 0.0;
                                    m = IDX3(-iA, -jA, -kA);
                                    const size_t nb = V3(k + kA, i + iA, j + jA);
                                    if (umask[nb] != 1) {
                                        bu[nb] = bu[nb] - Auu[m - 1 + 27 * nb] * uvel[V3(k, i, j)];
                                        Auu[m - 1 + 27 * nb] = 0.0;
                                    }
                                    if (vmask[nb] != 1) {
                                        bv[nb] = bv[nb] - Avu[m - 1 + 27 * nb] * uvel[V3(k, i, j)];
                                        Avu[m - 1 + 27 * nb] = 0.0;
                                    }
                                }
                            }
                }
                if (vmask[V3(k, i, j)] == 1) {
                    bv[V3(k, i, j)] = vvel[V3(k, i, j)];
                    for (int kA = -1; kA <= 1; ++kA)
                        for (int jA = -1; jA <= 1; ++jA)
                            for (int iA = -1; iA <= 1; ++iA) {
                                if (!(k + kA >= 1 && k + kA <= nz && i + iA >= 1 && i + iA <= nx - 1 &&
                                      j + jA >= 1 && j + jA <= ny - 1)) continue;
                                if (iA == 0 && jA == 0 && kA == 0) {
                                    Auv[A4(md, k, i, j)] = 0.0; Avu[A4(md, k, i, j)] = 0.0; Avv[A4(md, k, i, j)] = 1.0;
                                } else {
                                    int m = IDX3(iA, jA, kA);
                                    Avu[A4(m, k, i, j)] = 0.0; Avv[A4(m, k, i, j)] = 0.0;
                                    m = IDX3(-iA, -jA, -kA);
                                    const size_t nb = V3(k + kA, i + iA, j + jA);
                                    if (umask[nb] != 1) {
                                        bu[nb] = bu[nb] - Auv[m - 1 + 27 * nb] * vvel[V3(k, i, j)];
                                        Auv[m - 1 + 27 * nb] = 0.0;
                                    }
                                    if (vmask[nb] != 1) {
                                        bv[nb] = bv[nb] - Avv[m - 1 + 27 * nb] * vvel[V3(k, i, j)];
                                        Avv[m - 1 + 27 * nb] = 0.0;
                                    }
                                }
                            }
                }
            }
        }
}

/* check_symmetry_assembled_matrix_3d (glissade_velo_higher.F90:10854-11117):
 * averages partner entries that differ by less than eps11*|diag|.
 * Returns the number of pairs left asymmetric (the reference warns). */
int64_t orc_check_symmetry_assembled_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi,
                                        const int32_t* active_vertex,
                                        double* Auu, double* Auv, double* Avu, double* Avv)
{
    int64_t nwarn = 0;
    const int md = IDX3(0, 0, 0);
    for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            for (int k = 1; k <= nz; ++k) {
                double diag = Auu[A4(md, k, i, j)];
                for (int jA = -1; jA <= 1; ++jA)
                    for (int iA = -1; iA <= 1; ++iA)
                        for (int kA = -1; kA <= 1; ++kA) {
                            if (!(k + kA >= 1 && k + kA <= nz)) continue;
                            const size_t a1 = A4(IDX3(iA, jA, kA), k, i, j);
                            const size_t a2 = A4(IDX3(-iA, -jA, -kA), k + kA, i + iA, j + jA);
                            double v1 = Auu[a1], v2 = Auu[a2];
                            if (v2 != v1) {
                                if (fabs(v2 - v1) < EPS11 * fabs(diag)) { double av = 0.5 * (v1 + v2); Auu[a1] = av; Auu[a2] = av; }
                                else nwarn++;
                            }
                            v1 = Auv[a1]; v2 = Avu[a2];
                            if (v2 != v1) {
                                if (fabs(v2 - v1) < EPS11 * fabs(diag)) { double av = 0.5 * (v1 + v2); Auv[a1] = av; Avu[a2] = av; }
                                else nwarn++;
                            }
                        }
                diag = Avv[A4(md, k, i, j)];
                for (int jA = -1; jA <= 1; ++jA)
                    for (int iA = -1; iA <= 1; ++iA)
                        for (int kA = -1; kA <= 1; ++kA) {
                            if (!(k + kA >= 1 && k + kA <= nz)) continue;
                            const size_t a1 = A4(IDX3(iA, jA, kA), k, i, j);
                            const size_t a2 = A4(IDX3(-iA, -jA, -kA), k + kA, i + iA, j + jA);
                            double v1 = Avv[a1], v2 = Avv[a2];
                            if (v2 != v1) {
                                if (fabs(v2 - v1) < EPS11 * fabs(diag)) { double av = 0.5 * (v1 + v2); Avv[a1] = av; Avv[a2] = av; }
                                else nwarn++;
                            }
                            v1 = Avu[a1]; v2 = Auv[a2];
                            if (v2 != v1) {
                                if (fabs(v2 - v1) < EPS11 * fabs(diag)) { double av = 0.5 * (v1 + v2); Avu[a1] = av; Auv[a2] = av; }
                                else nwarn++;
                            }
                        }
            }
        }
    return nwarn;
}

/* count_nonzeros_3d (glissade_velo_higher.F90:10513-10579): nonzero
 * entries of the four blocks over owned active rows. */
int64_t orc_count_nonzeros_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi,
                              const int32_t* active_vertex, const double* Auu, const double* Auv,
                              const double* Avu, const double* Avv)
{
    int64_t c = 0;
    for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            for (int k = 1; k <= nz; ++k)
                for (int m = 1; m <= 27; ++m) {
                    const size_t a = A4(m, k, i, j);
                    if (Auu[a] != 0.0) c++;
                    if (Auv[a] != 0.0) c++;
                    if (Avu[a] != 0.0) c++;
                    if (Avv[a] != 0.0) c++;
                }
        }
    return c;
}

/* compute_residual_velocity_3d (glissade_velo_higher.F90:10314-10446) */
double orc_residual_velocity_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi, int whichresid,
                                const double* usav, const double* vsav, const double* uvel, const double* vvel)
{
    double r = 0.0;
    int64_t count = 0;
    const int ktop = (whichresid == 1 || whichresid == 2) ? nz - 1 : nz;
    for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i)
            for (int k = 1; k <= ktop; ++k) {
                const size_t p = V3(k, i, j);
                double speed = sqrt(uvel[p] * uvel[p] + vvel[p] * vvel[p]);
                if (speed != 0.0) {
                    double old = sqrt(usav[p] * usav[p] + vsav[p] * vsav[p]);
                    double d = fabs((old - speed) / speed);
                    if (whichresid == 2) { count++; r = r + d; }
                    else if (d > r) r = d;
                }
            }
    if (whichresid == 2 && count > 0) r = r / (double)count;
    return r;
}

/* ------------------------------------------------------------------ */
/* Lateral (calving-front) boundary condition                           */
/* ------------------------------------------------------------------ */

/* glissade_stagger with stagger_margin = 1: average over the cells whose
 * mask is 1 (glissade_grid_operators.F90:135-146). */
void orc_stagger_masked(int nx, int ny, const double* var, const int32_t* mask, double* stagvar)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const double sumvar = mask[C2(i, j + 1)] * var[C2(i, j + 1)] + mask[C2(i + 1, j + 1)] * var[C2(i + 1, j + 1)]
                                + mask[C2(i, j)] * var[C2(i, j)] + mask[C2(i + 1, j)] * var[C2(i + 1, j)];
            const double summask = (double)(mask[C2(i, j + 1)] + mask[C2(i + 1, j + 1)] + mask[C2(i, j)] +
                                            mask[C2(i + 1, j)]);
            stagvar[V2(i, j)] = summask > 0.0 ? sumvar / summask : 0.0;
        }
}

/* lateral_shelf_bc (glissade_velo_higher.F90:4609-4957) for one face of
 * cell (ic, jc).  face: 0 west, 1 east, 2 south, 3 north.  The face is a
 * 2-D bilinear element per layer (nodes 1,2 at the lower level k+1, 3,4
 * above them at level k), integrated with the 2x2 Gauss rule; the net
 * outward pressure p_av = rho_i g h/2 - rho_o g h (1 - min(s/h,1))^2 / 2
 * is added to loadu (west/east faces) or loadv (south/north). */
static void lateral_face(int nx, int ny, int nz, const double* sigma, int local, int face, int ic, int jc,
                         const double* usrf, const double* thck, const double* stagusrf, const double* stagthck,
                         const double* xVertex, const double* yVertex, double* loadu, double* loadv)
{
    const double r = 1.0 / sqrt(3.0);
    const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r}, wq[4] = {1.0, 1.0, 1.0, 1.0};
    double phi[4][4], dxr[4][4], dyr[4][4];   /* [n][p] */
    for (int p = 0; p < 4; ++p) {
        phi[0][p] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
        phi[1][p] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
        phi[2][p] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
        phi[3][p] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
        dxr[0][p] = -(1.0 - yq[p]) / 4.0;
        dxr[1][p] = (1.0 - yq[p]) / 4.0;
        dxr[2][p] = (1.0 + yq[p]) / 4.0;
        dxr[3][p] = -(1.0 + yq[p]) / 4.0;
        dyr[0][p] = -(1.0 - xq[p]) / 4.0;
        dyr[1][p] = -(1.0 + xq[p]) / 4.0;
        dyr[2][p] = (1.0 + xq[p]) / 4.0;
        dyr[3][p] = (1.0 - xq[p]) / 4.0;
    }
    int iN[4], jN[4];
    double x[4], s[4], h[4], y[4];
    switch (face) {
    case 0: iN[0] = ic - 1; jN[0] = jc;     iN[1] = ic - 1; jN[1] = jc - 1; break;
    case 1: iN[0] = ic;     jN[0] = jc - 1; iN[1] = ic;     jN[1] = jc;     break;
    case 2: iN[0] = ic - 1; jN[0] = jc - 1; iN[1] = ic;     jN[1] = jc - 1; break;
    default: iN[0] = ic;    jN[0] = jc;     iN[1] = ic - 1; jN[1] = jc;     break;
    }
    const double* coord = (face <= 1) ? yVertex : xVertex;
    x[0] = coord[V2(iN[0], jN[0])];
    x[1] = coord[V2(iN[1], jN[1])];
    iN[2] = iN[1]; jN[2] = jN[1]; iN[3] = iN[0]; jN[3] = jN[0];
    x[2] = x[1];
    x[3] = x[0];
    if (local) {
        for (int n = 0; n < 4; ++n) { s[n] = usrf[C2(ic, jc)]; h[n] = thck[C2(ic, jc)]; }
    } else {
        s[0] = stagusrf[V2(iN[0], jN[0])];
        s[1] = stagusrf[V2(iN[1], jN[1])];
        s[2] = s[1];
        s[3] = s[0];
        h[0] = stagthck[V2(iN[0], jN[0])];
        h[1] = stagthck[V2(iN[1], jN[1])];
        h[2] = h[1];
        h[3] = h[0];
    }
    for (int k = 1; k <= nz - 1; ++k) {
        y[0] = s[0] - sigma[k] * h[0];
        y[1] = s[1] - sigma[k] * h[1];
        y[2] = s[2] - sigma[k - 1] * h[2];
        y[3] = s[3] - sigma[k - 1] * h[3];
        const int kN[4] = {k + 1, k + 1, k, k};
        for (int p = 0; p < 4; ++p) {
            double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
            for (int n = 0; n < 4; ++n) {
                J11 = J11 + dxr[n][p] * x[n];
                J12 = J12 + dxr[n][p] * y[n];
                J21 = J21 + dyr[n][p] * x[n];
                J22 = J22 + dyr[n][p] * y[n];
            }
            const double detJ = fabs(J11 * J22 - J12 * J21);
            double h_qp, s_qp;
            if (local) {
                h_qp = thck[C2(ic, jc)];
                s_qp = usrf[C2(ic, jc)];
            } else {
                h_qp = 0.0;
                s_qp = 0.0;
                for (int n = 0; n < 4; ++n) {
                    h_qp = h_qp + phi[n][p] * h[n];
                    s_qp = s_qp + phi[n][p] * s[n];
                }
            }
            const double f = 1.0 - fmin(s_qp / h_qp, 1.0);
            const double p_av = 0.5 * RHOI * GRAV * h_qp - 0.5 * RHOO * GRAV * h_qp * (f * f);
            for (int n = 0; n < 4; ++n) {
                const double t = p_av * wq[p] * detJ / VOL0 * phi[n][p];
                const size_t c = V3(kN[n], iN[n], jN[n]);
                if (face == 0) loadu[c] = loadu[c] - t;
                else if (face == 1) loadu[c] = loadu[c] + t;
                else if (face == 2) loadv[c] = loadv[c] - t;
                else loadv[c] = loadv[c] + t;
            }
        }
    }
}

/* load_vector_lateral_bc (glissade_velo_higher.F90:4453-4605): every active
 * cell of the loop range that shares an edge with an ice-free ocean cell
 * gets the face term on that edge; faces in the order west, east, south,
 * north.  local = which_ho_assemble_lateral (1: the cell's own usrf/thck;
 * 0: staggered values at the face's vertices). */
void orc_load_vector_lateral_bc(int nx, int ny, int nz, const double* sigma, int nhalo, int local,
                                const int32_t* ocean_mask, const int32_t* active_cell,
                                const double* xVertex, const double* yVertex, const double* usrf, const double* thck,
                                const double* stagusrf, const double* stagthck, double* loadu, double* loadv)
{
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            if (ocean_mask[C2(i - 1, j)] == 1)
                lateral_face(nx, ny, nz, sigma, local, 0, i, j, usrf, thck, stagusrf, stagthck, xVertex, yVertex, loadu, loadv);
            if (ocean_mask[C2(i + 1, j)] == 1)
                lateral_face(nx, ny, nz, sigma, local, 1, i, j, usrf, thck, stagusrf, stagthck, xVertex, yVertex, loadu, loadv);
            if (ocean_mask[C2(i, j - 1)] == 1)
                lateral_face(nx, ny, nz, sigma, local, 2, i, j, usrf, thck, stagusrf, stagthck, xVertex, yVertex, loadu, loadv);
            if (ocean_mask[C2(i, j + 1)] == 1)
                lateral_face(nx, ny, nz, sigma, local, 3, i, j, usrf, thck, stagusrf, stagthck, xVertex, yVertex, loadu, loadv);
        }
}

/* glissade_grounded_fraction with which_ho_ground = 0 (no GLP) and
 * which_ho_fground_no_glp = 0 (grounded neighbour), the defaults
 * (glissade_grounding_line.F90:59-674; glide_types.F90:1090, 1099):
 * vmask(i,j) = 1 if any of the four cells has ice, or one of the land
 * cells (i,j+1), (i+1,j+1), (i,j), (i+1,j+1) -- the reference tests
 * (i+1,j+1) twice and never (i+1,j); f_ground = 1 at vmask vertices with
 * a neighbour cell that is land or grounded ice, else 0. */
void orc_grounded_fraction_no_glp(int nx, int ny, const int32_t* ice_mask, const int32_t* floating_mask,
                                  const int32_t* land_mask, double* f_ground)
{
#define CG_(i, j) ((ice_mask[C2(i, j)] == 1 && floating_mask[C2(i, j)] == 0) || land_mask[C2(i, j)] == 1)
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const int vmask = ice_mask[C2(i, j + 1)] == 1 || ice_mask[C2(i + 1, j + 1)] == 1 ||
                              ice_mask[C2(i, j)] == 1 || ice_mask[C2(i + 1, j)] == 1 ||
                              land_mask[C2(i, j + 1)] == 1 || land_mask[C2(i + 1, j + 1)] == 1 ||
                              land_mask[C2(i, j)] == 1 || land_mask[C2(i + 1, j + 1)] == 1;
            double f = 0.0;
            if (vmask && (CG_(i, j + 1) || CG_(i + 1, j + 1) || CG_(i, j) || CG_(i + 1, j))) f = 1.0;
            f_ground[V2(i, j)] = f;
        }
#undef CG_
}

/* ------------------------------------------------------------------ */
/* 2-D (depth-integrated) assembly, SSA approximation                   */
/* ------------------------------------------------------------------ */

/* 2-D reference element (glissade_velo_higher.F90:528-570): 2x2 Gauss
 * points (-r,-r), (r,-r), (r,r), (-r,r), weights 1; bilinear phi_2d and its
 * reference derivatives, stored [n][p]. */
static double phi2[4][4], dxr2[4][4], dyr2[4][4];
static int tables2_ready = 0;
/* node n of the element in cell (i,j) sits at vertex (i + di2[n], j + dj2[n])
 * (ishift(3,n), jshift(3,n), :461-494) */
static const int di2[4] = {-1, 0, 0, -1}, dj2[4] = {-1, -1, 0, 0};
#define IDX2(iA, jA) (1 + ((iA) + 1) + 3 * ((jA) + 1))
#define A2(m, i, j) (V2(i, j) + (size_t)NVX * NVY * (size_t)((m) - 1))

static void init_tables_2d(void)
{
    if (tables2_ready) return;
    const double r = 1.0 / sqrt(3.0);
    const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r};
    for (int p = 0; p < 4; ++p) {
        phi2[0][p] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
        phi2[1][p] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
        phi2[2][p] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
        phi2[3][p] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
        dxr2[0][p] = -(1.0 - yq[p]) / 4.0;
        dxr2[1][p] = (1.0 - yq[p]) / 4.0;
        dxr2[2][p] = (1.0 + yq[p]) / 4.0;
        dxr2[3][p] = -(1.0 + yq[p]) / 4.0;
        dyr2[0][p] = -(1.0 - xq[p]) / 4.0;
        dyr2[1][p] = -(1.0 + xq[p]) / 4.0;
        dyr2[2][p] = (1.0 + xq[p]) / 4.0;
        dyr2[3][p] = (1.0 - xq[p]) / 4.0;
    }
    tables2_ready = 1;
}

/* get_basis_function_derivatives_2d (glissade_velo_higher.F90:6319-6504) */
static int basis_derivs_2d(const double* x, const double* y, int p, double* dx, double* dy, double* detJ)
{
    double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
    for (int n = 0; n < 4; ++n) {
        J11 = J11 + dxr2[n][p] * x[n];
        J12 = J12 + dxr2[n][p] * y[n];
        J21 = J21 + dyr2[n][p] * x[n];
        J22 = J22 + dyr2[n][p] * y[n];
    }
    const double d = J11 * J22 - J12 * J21;
    *detJ = d;
    if (!(fabs(d) > 0.0)) return ORC_ERR_SINGULAR;
    const double i11 = J22 / d, i12 = -J12 / d, i21 = -J21 / d, i22 = J11 / d;
    for (int n = 0; n < 4; ++n) {
        dx[n] = 0.0 + i11 * dxr2[n][p] + i12 * dyr2[n][p];
        dy[n] = 0.0 + i21 * dxr2[n][p] + i22 * dyr2[n][p];
    }
    return ORC_OK;
}

/* glissade_vertical_average, no mask (glissade_grid_operators.F90:1882-1946):
 * var (nz-1, nx, ny) layers -> var_2d (nx, ny) */
void orc_vertical_average(int nx, int ny, int nz, const double* sigma, const double* var, double* var_2d)
{
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double s = 0.0;
            for (int k = 1; k <= nz - 1; ++k) s = s + var[E3(k, i, j)] * (sigma[k] - sigma[k - 1]);
            var_2d[C2(i, j)] = s;
        }
}

/* compute_effective_viscosity, SSA branch (glissade_velo_higher.F90:7285-7470) */
static double effective_viscosity_ssa(int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                                      const double* dx, const double* dy, const double* u, const double* v,
                                      double flwafact)
{
    const double p_effstr = (1.0 - n_glen) / n_glen;
    if (whichefvs == 0) return efvs_constant;
    if (whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        return flwafact * pow(effstrain, p_effstr);
    }
    double ux = 0, uy = 0, vx = 0, vy = 0;
    for (int n = 0; n < 4; ++n) {
        ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
        vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
    }
    const double esq = effstrain_min * effstrain_min + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
    return flwafact * pow(esq, p_effstr / 2.0);
}

/* One SSA element: quadrature loop of assemble_stiffness_matrix_2d
 * (glissade_velo_higher.F90:5216-5337) with compute_element_matrix (SSA
 * branch, :8001-8123): factor = (h_qp efvs_qp) w detJ / vol0, h_qp =
 * sum_n phi_2d(n,p) h(n).  efvs_mean = mean over the 4 points (:5310-5318).
 * K are [nr][nc]. */
int orc_element_matrix_2d_ssa(const double* x, const double* y, const double* u, const double* v, const double* h,
                              double flwafact_2d, int whichefvs, double efvs_constant, double effstrain_min,
                              double n_glen, double Kuu[4][4], double Kuv[4][4], double Kvu[4][4], double Kvv[4][4],
                              double* efvs_mean)
{
    init_tables_2d();
    double dx[4], dy[4], detJ, efvs_qp[4];
    memset(Kuu, 0, 16 * sizeof(double)); memset(Kuv, 0, 16 * sizeof(double));
    memset(Kvu, 0, 16 * sizeof(double)); memset(Kvv, 0, 16 * sizeof(double));
    for (int p = 0; p < 4; ++p) {
        int st = basis_derivs_2d(x, y, p, dx, dy, &detJ);
        if (st) return st;
        efvs_qp[p] = effective_viscosity_ssa(whichefvs, efvs_constant, effstrain_min, n_glen, dx, dy, u, v,
                                             flwafact_2d);
        double h_qp = 0.0;
        for (int n = 0; n < 4; ++n) h_qp = h_qp + phi2[n][p] * h[n];
        const double f = (h_qp * efvs_qp[p]) * 1.0 * detJ / VOL0;
        for (int nc = 0; nc < 4; ++nc)
            for (int nr = 0; nr < 4; ++nr) {
                Kuu[nr][nc] = Kuu[nr][nc] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc]);
                Kuv[nr][nc] = Kuv[nr][nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                Kvu[nr][nc] = Kvu[nr][nc] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                Kvv[nr][nc] = Kvv[nr][nc] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc]);
            }
    }
    double s = 0.0;
    for (int p = 0; p < 4; ++p) s = s + efvs_qp[p];
    *efvs_mean = s / 4.0;
    /* check_symmetry_element_matrix (:10583-10644) */
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            if (fabs(Kuu[a][b] - Kuu[b][a]) > EPS11) return ORC_ERR_ASYM;
            if (fabs(Kvv[a][b] - Kvv[b][a]) > EPS11) return ORC_ERR_ASYM;
            if (fabs(Kuv[a][b] - Kvu[b][a]) > EPS11) return ORC_ERR_ASYM;
        }
    return ORC_OK;
}

/* assemble_stiffness_matrix_2d, SSA (glissade_velo_higher.F90:4961-5381)
 * with element_to_global_matrix_2d (:8127-8196).  flwafact_2d is the
 * vertical average of flwafact (:5187-5192); efvs(k,i,j) takes the element
 * mean for every layer k.  A are (nx-1, ny-1, 9), m last. */
int orc_assemble_stiffness_matrix_2d_ssa(int nx, int ny, int nz, const double* sigma, int nhalo,
                                         const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                         const double* uvel_2d, const double* vvel_2d, const double* stagthck,
                                         const double* flwafact, int whichefvs, double efvs_constant,
                                         double effstrain_min, double n_glen, double* efvs,
                                         double* Auu, double* Auv, double* Avu, double* Avv)
{
    init_tables_2d();
    const size_t nA = (size_t)9 * NVX * NVY;
    memset(efvs, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
    memset(Auu, 0, nA * sizeof(double)); memset(Auv, 0, nA * sizeof(double));
    memset(Avu, 0, nA * sizeof(double)); memset(Avv, 0, nA * sizeof(double));
    double* fl2 = (double*)malloc((size_t)nx * ny * sizeof(double));
    orc_vertical_average(nx, ny, nz, sigma, flwafact, fl2);
    double Kuu[4][4], Kuv[4][4], Kvu[4][4], Kvv[4][4];
    double x[4], y[4], u[4], v[4], h[4];
    int st = ORC_OK;
    for (int j = nhalo + 1; j <= ny - nhalo + 1 && !st; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1 && !st; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int n = 0; n < 4; ++n) {
                const int in = i + di2[n], jn = j + dj2[n];
                x[n] = xVertex[V2(in, jn)];
                y[n] = yVertex[V2(in, jn)];
                u[n] = uvel_2d[V2(in, jn)];
                v[n] = vvel_2d[V2(in, jn)];
                h[n] = stagthck[V2(in, jn)];
            }
            double em;
            st = orc_element_matrix_2d_ssa(x, y, u, v, h, fl2[C2(i, j)], whichefvs, efvs_constant, effstrain_min,
                                           n_glen, Kuu, Kuv, Kvu, Kvv, &em);
            if (st) break;
            for (int k = 1; k <= nz - 1; ++k) efvs[E3(k, i, j)] = em;
            for (int nr = 0; nr < 4; ++nr) {
                const int ir = i + di2[nr], jr = j + dj2[nr];
                for (int nc = 0; nc < 4; ++nc) {
                    const int m = IDX2(di2[nc] - di2[nr], dj2[nc] - dj2[nr]);
                    const size_t a = A2(m, ir, jr);
                    Auu[a] = Auu[a] + Kuu[nr][nc];
                    Auv[a] = Auv[a] + Kuv[nr][nc];
                    Avu[a] = Avu[a] + Kvu[nr][nc];
                    Avv[a] = Avv[a] + Kvv[nr][nc];
                }
            }
        }
    free(fl2);
    return st;
}

/* basal_sliding_bc_2d, local option (glissade_velo_higher.F90:8281-8531):
 * diag += dx dy / vol0 * beta on active vertices (no slope factor in 2-D) */
void orc_basal_sliding_bc_2d_local(int nx, int ny, double dx, double dy, const int32_t* active_vertex,
                                   const double* beta, double* Auu, double* Avv)
{
    const int m = IDX2(0, 0);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            const double inc = (dx * dy / VOL0) * beta[V2(i, j)];
            Auu[A2(m, i, j)] = Auu[A2(m, i, j)] + inc;
            Avv[A2(m, i, j)] = Avv[A2(m, i, j)] + inc;
        }
}

/* dirichlet_boundary_conditions_2d (glissade_velo_higher.F90:9133-9301),
 * loops (jA, iA) outside (j, i) as the reference; masks are the basal level
 * of the 3-D masks (:2652-2657).  A u-prescribed row sets the diagonal of
 * Auu, Auv and Avu to 1 (:9206-9210), as the reference does. */
void orc_dirichlet_bc_2d(int nx, int ny, int nhalo, const int32_t* active_vertex, const int32_t* umask,
                         const int32_t* vmask, const double* uvel, const double* vvel,
                         double* Auu, double* Auv, double* Avu, double* Avv, double* bu, double* bv)
{
    const int md = IDX2(0, 0);
    for (int jA = -1; jA <= 1; ++jA)
        for (int iA = -1; iA <= 1; ++iA) {
            const int m = IDX2(iA, jA), mm = IDX2(-iA, -jA);
            for (int j = nhalo; j <= ny - nhalo + 1; ++j)
                for (int i = nhalo; i <= nx - nhalo + 1; ++i) {
                    if (!active_vertex[V2(i, j)]) continue;
                    const int inb = (i + iA >= 1 && i + iA <= nx - 1) && (j + jA >= 1 && j + jA <= ny - 1);
                    if (umask[V2(i, j)] == 1) {
                        bu[V2(i, j)] = uvel[V2(i, j)];
                        if (inb) {
                            if (iA == 0 && jA == 0) {
                                Auu[A2(md, i, j)] = 1.0; Auv[A2(md, i, j)] = 1.0; Avu[A2(md, i, j)] = 1.0;
                            } else {
                                Auu[A2(m, i, j)] = 0.0; Auv[A2(m, i, j)] = 0.0;
                                const size_t nb = V2(i + iA, j + jA);
                                const size_t a = A2(mm, i + iA, j + jA);
                                if (umask[nb] != 1) { bu[nb] = bu[nb] - Auu[a] * uvel[V2(i, j)]; Auu[a] = 0.0; }
                                if (vmask[nb] != 1) { bv[nb] = bv[nb] - Avu[a] * uvel[V2(i, j)]; Avu[a] = 0.0; }
                            }
                        }
                    }
                    if (vmask[V2(i, j)] == 1) {
                        bv[V2(i, j)] = vvel[V2(i, j)];
                        if (inb) {
                            if (iA == 0 && jA == 0) {
                                Auv[A2(md, i, j)] = 0.0; Avu[A2(md, i, j)] = 0.0; Avv[A2(md, i, j)] = 1.0;
                            } else {
                                Avu[A2(m, i, j)] = 0.0; Avv[A2(m, i, j)] = 0.0;
                                const size_t nb = V2(i + iA, j + jA);
                                const size_t a = A2(mm, i + iA, j + jA);
                                if (umask[nb] != 1) { bu[nb] = bu[nb] - Auv[a] * vvel[V2(i, j)]; Auv[a] = 0.0; }
                                if (vmask[nb] != 1) { bv[nb] = bv[nb] - Avv[a] * vvel[V2(i, j)]; Avv[a] = 0.0; }
                            }
                        }
                    }
                }
        }
}

/* check_symmetry_assembled_matrix_2d (glissade_velo_higher.F90:10648-10850):
 * the 2-D check only reports (its averaging lines are commented out in the
 * reference); returns the number of pairs that differ by eps11 |diag| or
 * more.  The partner entry is read at (i+iA, j+jA) as the reference does. */
int64_t orc_check_symmetry_assembled_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi,
                                        const int32_t* active_vertex, const double* Auu, const double* Auv,
                                        const double* Avu, const double* Avv)
{
    int64_t nwarn = 0;
    for (int jA = -1; jA <= 1; ++jA)
        for (int iA = -1; iA <= 1; ++iA) {
            const int m = IDX2(iA, jA), mm = IDX2(-iA, -jA);
            for (int j = jlo; j <= jhi; ++j)
                for (int i = ilo; i <= ihi; ++i) {
                    if (!active_vertex[V2(i, j)]) continue;
                    const double du = Auu[A2(IDX2(0, 0), i, j)], dv = Avv[A2(IDX2(0, 0), i, j)];
                    const double pairs[4][3] = {
                        {Auu[A2(m, i, j)], Auu[A2(mm, i + iA, j + jA)], du},
                        {Auv[A2(m, i, j)], Avu[A2(mm, i + iA, j + jA)], du},
                        {Avv[A2(m, i, j)], Avv[A2(mm, i + iA, j + jA)], dv},
                        {Avu[A2(m, i, j)], Auv[A2(mm, i + iA, j + jA)], dv}};
                    for (int s = 0; s < 4; ++s)
                        if (pairs[s][1] != pairs[s][0] && !(fabs(pairs[s][1] - pairs[s][0]) < EPS11 * fabs(pairs[s][2])))
                            ++nwarn;
                }
        }
    return nwarn;
}

/* count_nonzeros_2d (glissade_velo_higher.F90:10450-10509) */
int64_t orc_count_nonzeros_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi, const int32_t* active_vertex,
                              const double* Auu, const double* Auv, const double* Avu, const double* Avv)
{
    int64_t n = 0;
    for (int m = 1; m <= 9; ++m)
        for (int j = jlo; j <= jhi; ++j)
            for (int i = ilo; i <= ihi; ++i) {
                if (!active_vertex[V2(i, j)]) continue;
                n += (Auu[A2(m, i, j)] != 0.0) + (Auv[A2(m, i, j)] != 0.0) + (Avu[A2(m, i, j)] != 0.0) +
                     (Avv[A2(m, i, j)] != 0.0);
            }
    return n;
}

/* compute_residual_velocity_2d (glissade_velo_higher.F90:10188-10310):
 * which_ho_resid 2 is the mean relative speed change, every other value
 * the maximum, over all owned vertices (no bed level to leave out). */
double orc_residual_velocity_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi, int whichresid,
                                const double* usav, const double* vsav, const double* uvel, const double* vvel)
{
    double r = 0.0;
    int64_t count = 0;
    for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i) {
            const size_t p = V2(i, j);
            const double speed = sqrt(uvel[p] * uvel[p] + vvel[p] * vvel[p]);
            if (speed != 0.0) {
                const double old = sqrt(usav[p] * usav[p] + vsav[p] * vsav[p]);
                const double d = fabs((old - speed) / speed);
                if (whichresid == 2) { count++; r = r + d; }
                else if (d > r) r = d;
            }
        }
    if (whichresid == 2 && count > 0) r = r / (double)count;
    return r;
}

/* ------------------------------------------------------------------ */
/* 2-D assembly, DIVA approximation (Goldberg 2011)                     */
/* ------------------------------------------------------------------ */

/* glissade_slope_angle with the x and y angles (glissade_grid_operators.F90:
 * 1264-1410): theta_x = atan|dz/dx|, theta_y = atan|dz/dy| of the
 * unstaggered gradient; zero on the outer ring. */
void orc_slope_angle_xy(int nx, int ny, double dx, double dy, const double* zsfc, double* theta, double* thx,
                        double* thy)
{
    double* gx = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    double* gy = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    gradient_full(nx, ny, dx, dy, zsfc, gx, gy);
    for (size_t p = 0; p < (size_t)nx * ny; ++p) { theta[p] = 0.0; thx[p] = 0.0; thy[p] = 0.0; }
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const double dzx = (gx[V2(i, j)] + gx[V2(i - 1, j)] + gx[V2(i, j - 1)] + gx[V2(i - 1, j - 1)]) / 4.0;
            const double dzy = (gy[V2(i, j)] + gy[V2(i - 1, j)] + gy[V2(i, j - 1)] + gy[V2(i - 1, j - 1)]) / 4.0;
            const double s = sqrt(dzx * dzx + dzy * dzy);
            theta[C2(i, j)] = (s >= 0.0) ? atan(s) : 0.0;
            thx[C2(i, j)] = atan(fabs(dzx));
            thy[C2(i, j)] = atan(fabs(dzy));
        }
    free(gx); free(gy);
}

/* glissade_slope_angle_staggered (glissade_grid_operators.F90:1414-1552):
 * angles of the vertex gradient (glissade_gradient, margin 1, full mask) */
void orc_slope_angle_staggered(int nx, int ny, double dx, double dy, const double* zsfc, double* theta,
                               double* thx, double* thy)
{
    double* gx = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    double* gy = (double*)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
    gradient_full(nx, ny, dx, dy, zsfc, gx, gy);
    for (size_t p = 0; p < (size_t)(nx - 1) * (ny - 1); ++p) {
        const double s = sqrt(gx[p] * gx[p] + gy[p] * gy[p]);
        theta[p] = (s >= 0.0) ? atan(s) : 0.0;
        thx[p] = atan(fabs(gx[p]));
        thy[p] = atan(fabs(gy[p]));
    }
    free(gx); free(gy);
}

/* parallel_halo_real8_2d for a cell field on one task, lhalo = uhalo = 2
 * (parallel_mpi.F90:7720-7912): periodic wrap (E/W over owned rows, then N/S
 * over all columns); outflow zeroes the 2 halo rows; no_ice zeroes 3 rows
 * on every side (the halo and the first/last owned row). */
void orc_cell_halo(int nx, int ny, int global_bc, double* a)
{
    if (global_bc == 0 || global_bc == 2) {
        for (int j = 3; j <= ny - 2; ++j) {
            a[C2(1, j)] = a[C2(nx - 3, j)]; a[C2(2, j)] = a[C2(nx - 2, j)];
            a[C2(nx - 1, j)] = a[C2(3, j)]; a[C2(nx, j)] = a[C2(4, j)];
        }
        for (int i = 1; i <= nx; ++i) {
            a[C2(i, 1)] = a[C2(i, ny - 3)]; a[C2(i, 2)] = a[C2(i, ny - 2)];
            a[C2(i, ny - 1)] = a[C2(i, 3)]; a[C2(i, ny)] = a[C2(i, 4)];
        }
    } else {
        const int w = (global_bc == 1) ? 2 : 3;
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                if (i <= w || i >= nx - w + 1 || j <= w || j >= ny - w + 1) a[C2(i, j)] = 0.0;
    }
}

/* compute_effective_viscosity_diva (glissade_velo_higher.F90:7730-7997),
 * the default (non-cubic) branch: layer k's vertical shear is estimated from
 * the basal traction and the previous efvs(k) (initialised from the minimum
 * strain rate when zero), efvs(k) is updated in place. */
static void efvs_diva(int whichefvs, double efvs_constant, double effstrain_min, double n_glen, int nz,
                      const double* stagsigma, int p, const double* dx, const double* dy, const double* u,
                      const double* v, const double* bx, const double* by, double sfx, double sfy, const double* h,
                      const double* flwafact, double* efvs)
{
    const double p_effstr = (1.0 - n_glen) / n_glen;
    if (whichefvs == 0) { for (int k = 0; k < nz - 1; ++k) efvs[k] = efvs_constant; return; }
    if (whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        for (int k = 0; k < nz - 1; ++k) efvs[k] = flwafact[k] * pow(effstrain, p_effstr);
        return;
    }
    double ux = 0, uy = 0, vx = 0, vy = 0, thck = 0, taux = 0, tauy = 0;
    for (int n = 0; n < 4; ++n) {
        ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
        vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
        taux = taux + phi2[n][p] * bx[n];
        tauy = tauy + phi2[n][p] * by[n];
        thck = thck + phi2[n][p] * h[n];
    }
    (void)thck;
    for (int k = 0; k < nz - 1; ++k) {
        if (efvs[k] == 0.0) efvs[k] = flwafact[k] * pow(effstrain_min, p_effstr);
        const double du_dz = taux * stagsigma[k] / (efvs[k] * sfx);
        const double dv_dz = tauy * stagsigma[k] / (efvs[k] * sfy);
        const double esq = effstrain_min * effstrain_min + ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)) +
                           0.25 * (du_dz * du_dz) * sfx + 0.25 * (dv_dz * dv_dz) * sfy;
        efvs[k] = flwafact[k] * pow(esq, p_effstr / 2.0);
    }
}

/* compute_integrals_diva (glissade_velo_higher.F90:5649-5775): omega_k(k) =
 * mean over the quadrature points of the integral from the bed up to level
 * k of depth/efvs dz (Goldberg 2011 eq. 32), omega = its vertical mean
 * (eq. 35, divided by H) */
static void integrals_diva(int nz, const double* sigma, double thck, const double* efvs_qp /* [p][k] */,
                           double* omega_k, double* omega)
{
    double okp[4][65];
    for (int p = 0; p < 4; ++p) {
        okp[p][nz - 1] = 0.0;
        for (int k = nz - 2; k >= 0; --k) {
            const double depth = 0.5 * (sigma[k] + sigma[k + 1]);
            const double dz = (sigma[k + 1] - sigma[k]) * thck;
            okp[p][k] = okp[p][k + 1] + depth / efvs_qp[p * (nz - 1) + k] * dz;
        }
    }
    for (int k = 0; k < nz; ++k) omega_k[k] = (((okp[0][k] + okp[1][k]) + okp[2][k]) + okp[3][k]) / 4.0;
    double om = 0.0;
    for (int k = 0; k < nz - 1; ++k) om = om + 0.5 * (omega_k[k] + omega_k[k + 1]) * (sigma[k + 1] - sigma[k]);
    *omega = om;
}

/* assemble_stiffness_matrix_2d, DIVA (glissade_velo_higher.F90:4961-5381):
 * per active cell and quadrature point the layer viscosities of the state
 * efvs_qp_3d (nz-1, 4, nx, ny) are updated from the 2-D velocity and the
 * basal traction (btractx/y at vertices), their vertical mean times the
 * interpolated thickness weights the element matrix (the BP form with
 * dphi_dz = 0), and omega_k (nz, nx, ny), omega (nx, ny) are formed from
 * the cell thickness thck (nx, ny).  efvs(k,i,j) = mean over the points. */
int orc_assemble_stiffness_matrix_2d_diva(int nx, int ny, int nz, const double* sigma, const double* stagsigma,
                                          int nhalo, const int32_t* active_cell, const double* xVertex,
                                          const double* yVertex, const double* uvel_2d, const double* vvel_2d,
                                          const double* stagthck, const double* thck, const double* flwafact,
                                          const double* btractx, const double* btracty, const double* sfx_cell,
                                          const double* sfy_cell, int whichefvs, double efvs_constant,
                                          double effstrain_min, double n_glen, double* efvs_qp_3d, double* efvs,
                                          double* omega_k, double* omega, double* Auu, double* Auv, double* Avu,
                                          double* Avv)
{
    init_tables_2d();
    const size_t nA = (size_t)9 * NVX * NVY;
    memset(efvs, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
    memset(omega_k, 0, (size_t)nz * nx * ny * sizeof(double));
    memset(omega, 0, (size_t)nx * ny * sizeof(double));
    memset(Auu, 0, nA * sizeof(double)); memset(Auv, 0, nA * sizeof(double));
    memset(Avu, 0, nA * sizeof(double)); memset(Avv, 0, nA * sizeof(double));
    double x[4], y[4], u[4], v[4], h[4], bx[4], by[4];
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int n = 0; n < 4; ++n) {
                const size_t q = V2(i + di2[n], j + dj2[n]);
                x[n] = xVertex[q]; y[n] = yVertex[q]; u[n] = uvel_2d[q]; v[n] = vvel_2d[q];
                h[n] = stagthck[q]; bx[n] = btractx[q]; by[n] = btracty[q];
            }
            double* st = efvs_qp_3d + (size_t)(nz - 1) * 4 * C2(i, j);   /* [p][k] */
            const double* fl = flwafact + (size_t)(nz - 1) * C2(i, j);
            double Kuu[4][4] = {{0}}, Kuv[4][4] = {{0}}, Kvu[4][4] = {{0}}, Kvv[4][4] = {{0}};
            for (int p = 0; p < 4; ++p) {
                double dx[4], dy[4], detJ;
                if (basis_derivs_2d(x, y, p, dx, dy, &detJ)) return ORC_ERR_SINGULAR;
                efvs_diva(whichefvs, efvs_constant, effstrain_min, n_glen, nz, stagsigma, p, dx, dy, u, v, bx, by,
                          sfx_cell[C2(i, j)], sfy_cell[C2(i, j)], h, fl, st + (size_t)p * (nz - 1));
                double va = 0.0;
                for (int k = 0; k < nz - 1; ++k) va = va + st[p * (nz - 1) + k] * (sigma[k + 1] - sigma[k]);
                double h_qp = 0.0;
                for (int n = 0; n < 4; ++n) h_qp = h_qp + phi2[n][p] * h[n];
                const double f = (h_qp * va) * 1.0 * detJ / VOL0;
                for (int nc = 0; nc < 4; ++nc)
                    for (int nr = 0; nr < 4; ++nr) {
                        Kuu[nr][nc] = Kuu[nr][nc] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc]);
                        Kuv[nr][nc] = Kuv[nr][nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                        Kvu[nr][nc] = Kvu[nr][nc] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                        Kvv[nr][nc] = Kvv[nr][nc] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc]);
                    }
            }
            integrals_diva(nz, sigma, thck[C2(i, j)], st, omega_k + (size_t)nz * C2(i, j), omega + C2(i, j));
            for (int k = 0; k < nz - 1; ++k) {
                double s = 0.0;
                for (int p = 0; p < 4; ++p) s = s + st[p * (nz - 1) + k];
                efvs[(size_t)k + (size_t)(nz - 1) * C2(i, j)] = s / 4.0;
            }
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    if (fabs(Kuu[a][b] - Kuu[b][a]) > EPS11 || fabs(Kvv[a][b] - Kvv[b][a]) > EPS11 ||
                        fabs(Kuv[a][b] - Kvu[b][a]) > EPS11)
                        return ORC_ERR_ASYM;
            for (int nr = 0; nr < 4; ++nr) {
                const int ir = i + di2[nr], jr = j + dj2[nr];
                for (int nc = 0; nc < 4; ++nc) {
                    const int m = IDX2(di2[nc] - di2[nr], dj2[nc] - dj2[nr]);
                    const size_t a = A2(m, ir, jr);
                    Auu[a] = Auu[a] + Kuu[nr][nc];
                    Auv[a] = Auv[a] + Kuv[nr][nc];
                    Avu[a] = Avu[a] + Kvu[nr][nc];
                    Avv[a] = Avv[a] + Kvv[nr][nc];
                }
            }
        }
    return ORC_OK;
}

/* beta_eff of the DIVA solve (glissade_velo_higher.F90:2469-2560), slope
 * correction on: slip  beta_eff_x = sfx beta / (sfx + beta stag_omega cos(theta_x));
 * no-slip (where stag_omega > 0)  beta_eff_x = sfx / (stag_omega cos(theta_x)). */
void orc_diva_beta_eff(int nx, int ny, int no_slip, const double* stag_omega, const double* beta,
                       const double* ssfx, const double* ssfy, const double* sthx, const double* sthy,
                       double* beta_eff_x, double* beta_eff_y)
{
    for (size_t p = 0; p < (size_t)NVX * NVY; ++p) {
        beta_eff_x[p] = 0.0; beta_eff_y[p] = 0.0;
        if (no_slip) {
            if (stag_omega[p] > 0.0) {
                beta_eff_x[p] = ssfx[p] / (stag_omega[p] * cos(sthx[p]));
                beta_eff_y[p] = ssfy[p] / (stag_omega[p] * cos(sthy[p]));
            }
        } else {
            beta_eff_x[p] = (ssfx[p] * beta[p]) / (ssfx[p] + beta[p] * stag_omega[p] * cos(sthx[p]));
            beta_eff_y[p] = (ssfy[p] * beta[p]) / (ssfy[p] + beta[p] * stag_omega[p] * cos(sthy[p]));
        }
    }
}

/* basal_sliding_bc_2d_diva, local option (glissade_velo_higher.F90:8535-8811):
 * diag += dx dy / vol0 * beta_eff / cos(theta_basal) per direction */
void orc_basal_sliding_bc_2d_diva_local(int nx, int ny, double dx, double dy, const int32_t* active_vertex,
                                        const double* beta_eff_x, const double* beta_eff_y, const double* tbx,
                                        const double* tby, double* Auu, double* Avv)
{
    const int m = IDX2(0, 0);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!active_vertex[V2(i, j)]) continue;
            Auu[A2(m, i, j)] = Auu[A2(m, i, j)] + (dx * dy / VOL0) * beta_eff_x[V2(i, j)] / cos(tbx[V2(i, j)]);
            Avv[A2(m, i, j)] = Avv[A2(m, i, j)] + (dx * dy / VOL0) * beta_eff_y[V2(i, j)] / cos(tby[V2(i, j)]);
        }
}

/* compute_3d_velocity_diva (glissade_velo_higher.F90:5779-5900): basal
 * traction btract = beta_eff u_2d (every vertex), then on active vertices
 * the basal velocity u_b = u_2d beta_eff / beta (u_2d where beta = 0) and
 * u(k) = u_b + btract stag_omega_k(k) cos(theta) / slope_factor, with
 * stag_omega_k the masked stagger (ice or land cells) of omega_k. */
void orc_compute_3d_velocity_diva(int nx, int ny, int nz, const int32_t* active_vertex, const int32_t* ipl_mask,
                                  const double* omega_k, const double* beta, const double* beta_eff_x,
                                  const double* beta_eff_y, const double* sthx, const double* sthy,
                                  const double* ssfx, const double* ssfy, const double* uvel_2d,
                                  const double* vvel_2d, double* btractx, double* btracty, double* uvel,
                                  double* vvel)
{
    const size_t n2 = (size_t)NVX * NVY, nc = (size_t)nx * ny;
    for (size_t p = 0; p < n2; ++p) {
        btractx[p] = beta_eff_x[p] * uvel_2d[p];
        btracty[p] = beta_eff_y[p] * vvel_2d[p];
    }
    double* ok = (double*)malloc(nc * sizeof(double));
    double* sok = (double*)malloc((size_t)nz * n2 * sizeof(double));
    for (int k = 0; k < nz; ++k) {
        for (size_t c = 0; c < nc; ++c) ok[c] = omega_k[(size_t)k + (size_t)nz * c];
        orc_stagger_masked(nx, ny, ok, ipl_mask, sok + (size_t)k * n2);
    }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const size_t p = V2(i, j);
            if (!active_vertex[p]) continue;
            double ub, vb;
            if (beta[p] > 0.0) {
                ub = uvel_2d[p] * beta_eff_x[p] / beta[p];
                vb = vvel_2d[p] * beta_eff_y[p] / beta[p];
            } else {
                ub = uvel_2d[p];
                vb = vvel_2d[p];
            }
            uvel[V3(nz, i, j)] = ub;
            vvel[V3(nz, i, j)] = vb;
            const double cx = cos(sthx[p]) / ssfx[p], cy = cos(sthy[p]) / ssfy[p];
            for (int k = 1; k <= nz - 1; ++k) {
                uvel[V3(k, i, j)] = uvel[V3(nz, i, j)] + btractx[p] * sok[(size_t)(k - 1) * n2 + p] * cx;
                vvel[V3(k, i, j)] = vvel[V3(nz, i, j)] + btracty[p] * sok[(size_t)(k - 1) * n2 + p] * cy;
            }
        }
    free(ok); free(sok);
}

/* ------------------------------------------------------------------ */
/* L1L2 (Perego, Gunzburger and Burkardt 2012)                         */
/* ------------------------------------------------------------------ */

/* compute_effective_viscosity_L1L2 (glissade_velo_higher.F90:7474-7726) at
 * quadrature point p: the membrane strain rate and surface slope at the
 * point give, per layer k, tau_perp = rhoi g (h sigma(k+1)) |grad s| and
 * tau_parallel, the real root of x^3 + tau_perp^2 x - effstrain/A = 0
 * (Cardano, with the first-order form of the second cube root when
 * a^3/27 <= 1e-6 b^2/4); efvs = 1/(2A) for n = 1, 1/(2A(tau_par^2 +
 * tau_perp^2)) for n = 3, an error otherwise. */
static int efvs_l1l2(int whichefvs, double efvs_constant, double effstrain_min, double n_glen, int nz,
                     const double* sigma, int p, const double* dx, const double* dy, const double* u,
                     const double* v, const double* h, const double* sx, const double* sy, const double* flwa,
                     const double* flwafact, double* efvs)
{
    const double p_effstr = (1.0 - n_glen) / n_glen;
    if (whichefvs == 0) { for (int k = 0; k < nz - 1; ++k) efvs[k] = efvs_constant; return ORC_OK; }
    if (whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        for (int k = 0; k < nz - 1; ++k) efvs[k] = flwafact[k] * pow(effstrain, p_effstr);
        return ORC_OK;
    }
    double ux = 0, uy = 0, vx = 0, vy = 0, dsx = 0, dsy = 0, thck = 0;
    for (int n = 0; n < 4; ++n) {
        ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
        vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
        dsx = dsx + phi2[n][p] * sx[n];
        dsy = dsy + phi2[n][p] * sy[n];
        thck = thck + phi2[n][p] * h[n];
    }
    const double esq = effstrain_min * effstrain_min + ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy));
    const double es = sqrt(esq);
    for (int k = 0; k < nz - 1; ++k) {
        const double depth = thck * sigma[k + 1];
        const double grads = sqrt(dsx * dsx + dsy * dsy);
        const double tau_perp = RHOI * GRAV * depth * grads;
        const double a = tau_perp * tau_perp;
        const double b = -es / flwa[k];
        const double c = sqrt(b * b / 4.0 + a * a * a / 27.0);
        const double rootA = pow(-b / 2.0 + c, 1.0 / 3.0);
        double rootB;
        if (a * a * a / 27.0 > 1.0e-6 * (b * b / 4.0)) rootB = -pow(b / 2.0 + c, 1.0 / 3.0);
        else rootB = -a / (3.0 * pow(fabs(b), 1.0 / 3.0));
        const double tau_par = rootA + rootB;
        if (fabs(n_glen - 1.0) < 1.0e-10) efvs[k] = 1.0 / (2.0 * flwa[k]);
        else if (fabs(n_glen - 3.0) < 1.0e-10) efvs[k] = 1.0 / (2.0 * flwa[k] * (tau_par * tau_par + tau_perp * tau_perp));
        else return ORC_ERR_ARG;
    }
    return ORC_OK;
}

/* One point of compute_effective_viscosity_L1L2's cubic for the pins:
 * tau_parallel for effstrain, A and tau_perp. */
double orc_l1l2_tau_parallel(double effstrain, double flwa, double tau_perp)
{
    const double a = tau_perp * tau_perp;
    const double b = -effstrain / flwa;
    const double c = sqrt(b * b / 4.0 + a * a * a / 27.0);
    const double rootA = pow(-b / 2.0 + c, 1.0 / 3.0);
    const double rootB = (a * a * a / 27.0 > 1.0e-6 * (b * b / 4.0)) ? -pow(b / 2.0 + c, 1.0 / 3.0)
                                                                     : -a / (3.0 * pow(fabs(b), 1.0 / 3.0));
    return rootA + rootB;
}

/* assemble_stiffness_matrix_2d, L1L2 (glissade_velo_higher.F90:5228-5248,
 * 5300-5346): per active cell and quadrature point the L1L2 layer
 * viscosities from the 2-D velocity, the vertex surface gradient and
 * thickness, and the layer rate factor flwa (nz-1, nx, ny); their vertical
 * mean times the interpolated thickness weights the element matrix;
 * efvs(k,i,j) = mean over the points. */
int orc_assemble_stiffness_matrix_2d_l1l2(int nx, int ny, int nz, const double* sigma, int nhalo,
                                          const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                          const double* uvel_2d, const double* vvel_2d, const double* stagthck,
                                          const double* dusrf_dx, const double* dusrf_dy, const double* flwa,
                                          const double* flwafact, int whichefvs, double efvs_constant,
                                          double effstrain_min, double n_glen, double* efvs, double* Auu,
                                          double* Auv, double* Avu, double* Avv)
{
    init_tables_2d();
    const size_t nA = (size_t)9 * NVX * NVY;
    memset(efvs, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
    memset(Auu, 0, nA * sizeof(double)); memset(Auv, 0, nA * sizeof(double));
    memset(Avu, 0, nA * sizeof(double)); memset(Avv, 0, nA * sizeof(double));
    double x[4], y[4], u[4], v[4], h[4], sx[4], sy[4];
    double* eq = (double*)malloc((size_t)4 * (nz - 1) * sizeof(double));   /* [p][k] */
    int st = ORC_OK;
    for (int j = nhalo + 1; j <= ny - nhalo + 1 && !st; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1 && !st; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int n = 0; n < 4; ++n) {
                const size_t q = V2(i + di2[n], j + dj2[n]);
                x[n] = xVertex[q]; y[n] = yVertex[q]; u[n] = uvel_2d[q]; v[n] = vvel_2d[q];
                h[n] = stagthck[q]; sx[n] = dusrf_dx[q]; sy[n] = dusrf_dy[q];
            }
            const double* fa = flwa + (size_t)(nz - 1) * C2(i, j);
            const double* fl = flwafact + (size_t)(nz - 1) * C2(i, j);
            double Kuu[4][4] = {{0}}, Kuv[4][4] = {{0}}, Kvu[4][4] = {{0}}, Kvv[4][4] = {{0}};
            for (int p = 0; p < 4 && !st; ++p) {
                double dx[4], dy[4], detJ;
                st = basis_derivs_2d(x, y, p, dx, dy, &detJ);
                if (st) break;
                double* e = eq + (size_t)p * (nz - 1);
                st = efvs_l1l2(whichefvs, efvs_constant, effstrain_min, n_glen, nz, sigma, p, dx, dy, u, v, h, sx, sy,
                               fa, fl, e);
                if (st) break;
                double va = 0.0;
                for (int k = 0; k < nz - 1; ++k) va = va + e[k] * (sigma[k + 1] - sigma[k]);
                double h_qp = 0.0;
                for (int n = 0; n < 4; ++n) h_qp = h_qp + phi2[n][p] * h[n];
                const double f = (h_qp * va) * 1.0 * detJ / VOL0;
                for (int nc = 0; nc < 4; ++nc)
                    for (int nr = 0; nr < 4; ++nr) {
                        Kuu[nr][nc] = Kuu[nr][nc] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc]);
                        Kuv[nr][nc] = Kuv[nr][nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                        Kvu[nr][nc] = Kvu[nr][nc] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                        Kvv[nr][nc] = Kvv[nr][nc] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc]);
                    }
            }
            if (st) break;
            for (int k = 0; k < nz - 1; ++k) {
                double s = 0.0;
                for (int p = 0; p < 4; ++p) s = s + eq[p * (nz - 1) + k];
                efvs[(size_t)k + (size_t)(nz - 1) * C2(i, j)] = s / 4.0;
            }
            for (int nr = 0; nr < 4; ++nr) {
                const int ir = i + di2[nr], jr = j + dj2[nr];
                for (int nc = 0; nc < 4; ++nc) {
                    const int m = IDX2(di2[nc] - di2[nr], dj2[nc] - dj2[nr]);
                    const size_t a = A2(m, ir, jr);
                    Auu[a] = Auu[a] + Kuu[nr][nc];
                    Auv[a] = Auv[a] + Kuv[nr][nc];
                    Avu[a] = Avu[a] + Kvu[nr][nc];
                    Avv[a] = Avv[a] + Kvv[nr][nc];
                }
            }
        }
    free(eq);
    return st;
}

/* glissade_gradient with gradient_margin 1 (glissade_grid_operators.F90:
 * 309-468): an edge difference counts where both cells of the edge have
 * ice; the vertex value is the mean of the two edges in a direction, the one
 * edge that counts, or zero. */
static void gradient_margin1(int nx, int ny, double dx, double dy, const double* f, const int32_t* ice,
                             double* gx, double* gy)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const int es = ice[C2(i, j)] == 1 && ice[C2(i + 1, j)] == 1;
            const int en = ice[C2(i, j + 1)] == 1 && ice[C2(i + 1, j + 1)] == 1;
            const int ew = ice[C2(i, j)] == 1 && ice[C2(i, j + 1)] == 1;
            const int ee = ice[C2(i + 1, j)] == 1 && ice[C2(i + 1, j + 1)] == 1;
            const double n_ = (f[C2(i + 1, j + 1)] - f[C2(i, j + 1)]) / dx;
            const double s_ = (f[C2(i + 1, j)] - f[C2(i, j)]) / dx;
            gx[V2(i, j)] = (es && en) ? (n_ + s_) / 2.0 : (es ? s_ : (en ? n_ : 0.0));
            const double e_ = (f[C2(i + 1, j + 1)] - f[C2(i + 1, j)]) / dy;
            const double w_ = (f[C2(i, j + 1)] - f[C2(i, j)]) / dy;
            gy[V2(i, j)] = (ew && ee) ? (e_ + w_) / 2.0 : (ew ? w_ : (ee ? e_ : 0.0));
        }
}

/* compute_3d_velocity_L1L2 (glissade_velo_higher.F90:5904-6315): from the
 * basal (2-D) velocity, uvel(nz) on input, and the layer viscosities efvs
 * (nz-1, nx, ny) of the last assembly.  Per active cell: the membrane
 * strain rate at the cell centre, tau_parallel(k) = 2 efvs eps_parallel and
 * the integral of efvs from the layer midpoint to the surface; halos; then
 * from the bed up at owned active vertices
 *     u(k) = u(k+1) + fact(k) tau_xz(k),
 * tau_xz = -rhoi g depth ds/dx + 2 d(work1)/dx + d(work2)/dy (membrane
 * terms included), fact = 2 A tau_eff^(n-1) dz (nonlinear viscosity) or
 * dz / efvs(k, i, j) (the cell of the vertex's index, as the reference), and
 * u(k) = u(nz) where the basal Dirichlet mask is set.  uvel/vvel (nz, nx-1,
 * ny-1): levels 1..nz-1 are overwritten (zero off the owned active
 * vertices); no halo. */
int orc_compute_3d_velocity_l1l2(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi, int global_bc,
                                 int nhalo, const double* sigma, double dx, double dy, const int32_t* ice_mask,
                                 const int32_t* active_cell, const int32_t* active_vertex, const int32_t* umask,
                                 const int32_t* vmask, const double* xVertex, const double* yVertex,
                                 const double* thck, const double* stagthck, const double* dusrf_dx,
                                 const double* dusrf_dy, const double* flwa, const double* efvs, int whichefvs,
                                 double n_glen, double* uvel, double* vvel)
{
    const size_t nc = (size_t)nx * ny, n2 = (size_t)NVX * NVY;
    const int nl = nz - 1;
    double* ss = (double*)malloc((size_t)nl * sizeof(double));
    for (int k = 0; k < nl; ++k) ss[k] = (sigma[k] + sigma[k + 1]) / 2.0;
    double* ux = (double*)calloc(nc, sizeof(double));
    double* uy = (double*)calloc(nc, sizeof(double));
    double* vx = (double*)calloc(nc, sizeof(double));
    double* vy = (double*)calloc(nc, sizeof(double));
    double* tp = (double*)calloc((size_t)nl * nc, sizeof(double));    /* [k][c] */
    double* ei = (double*)calloc((size_t)nl * nc, sizeof(double));    /* [k][c] */
    double* w = (double*)malloc(3 * nc * sizeof(double));
    double* g = (double*)malloc(8 * n2 * sizeof(double));
    const double cdx[4] = {-0.25, 0.25, 0.25, -0.25}, cdy[4] = {-0.25, -0.25, 0.25, 0.25};
    int st = ORC_OK;
    for (size_t p = 0; p < n2; ++p)
        for (int k = 1; k <= nz - 1; ++k) { uvel[(size_t)(k - 1) + (size_t)nz * p] = 0.0; vvel[(size_t)(k - 1) + (size_t)nz * p] = 0.0; }
    for (int j = nhalo + 1; j <= ny - nhalo + 1 && !st; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            const size_t c = C2(i, j);
            if (!active_cell[c]) continue;
            double x[4], y[4], u[4], v[4];
            for (int n = 0; n < 4; ++n) {
                const size_t q = V2(i + di2[n], j + dj2[n]);
                x[n] = xVertex[q]; y[n] = yVertex[q];
                u[n] = uvel[(size_t)(nz - 1) + (size_t)nz * q];
                v[n] = vvel[(size_t)(nz - 1) + (size_t)nz * q];
            }
            double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
            for (int n = 0; n < 4; ++n) {
                J11 = J11 + cdx[n] * x[n]; J12 = J12 + cdx[n] * y[n];
                J21 = J21 + cdy[n] * x[n]; J22 = J22 + cdy[n] * y[n];
            }
            const double d = J11 * J22 - J12 * J21;
            if (!(fabs(d) > 0.0)) { st = ORC_ERR_SINGULAR; break; }
            const double i11 = J22 / d, i12 = -J12 / d, i21 = -J21 / d, i22 = J11 / d;
            for (int n = 0; n < 4; ++n) {
                const double ddx = 0.0 + i11 * cdx[n] + i12 * cdy[n];
                const double ddy = 0.0 + i21 * cdx[n] + i22 * cdy[n];
                ux[c] = ux[c] + ddx * u[n]; uy[c] = uy[c] + ddy * u[n];
                vx[c] = vx[c] + ddx * v[n]; vy[c] = vy[c] + ddy * v[n];
            }
            const double eps = sqrt(ux[c] * ux[c] + vy[c] * vy[c] + ux[c] * vy[c] + 0.25 * ((vx[c] + uy[c]) * (vx[c] + uy[c])));
            const double* ef = efvs + (size_t)nl * c;
            for (int k = 0; k < nl; ++k) tp[(size_t)k * nc + c] = 2.0 * ef[k] * eps;
            ei[c] = ef[0] * ss[0] * thck[c];
            for (int k = 1; k < nl; ++k)
                ei[(size_t)k * nc + c] = ei[(size_t)(k - 1) * nc + c] + ef[k - 1] * (sigma[k] - ss[k - 1]) * thck[c] +
                                         ef[k] * (ss[k] - sigma[k]) * thck[c];
        }
    if (st) goto done;
    for (int k = 0; k < nl; ++k) orc_cell_halo(nx, ny, global_bc, tp + (size_t)k * nc);
    for (int k = nl - 1; k >= 0; --k) {
        double *sq = g, *sf = g + n2, *g1x = g + 2 * n2, *g1y = g + 3 * n2, *g2x = g + 4 * n2, *g2y = g + 5 * n2,
               *g3x = g + 6 * n2, *g3y = g + 7 * n2;
        orc_stagger_masked(nx, ny, tp + (size_t)k * nc, ice_mask, sq);
        for (size_t p = 0; p < n2; ++p) sq[p] = sq[p] * sq[p];
        double* fk = (double*)malloc(nc * sizeof(double));
        for (size_t c = 0; c < nc; ++c) fk[c] = flwa[(size_t)k + (size_t)nl * c];
        orc_stagger_masked(nx, ny, fk, ice_mask, sf);
        free(fk);
        const double* e = ei + (size_t)k * nc;
        for (size_t c = 0; c < nc; ++c) {
            w[c] = e[c] * (2.0 * ux[c] + vy[c]);
            w[nc + c] = e[c] * (uy[c] + vx[c]);
            w[2 * nc + c] = e[c] * (2.0 * vy[c] + ux[c]);
        }
        for (int q = 0; q < 3; ++q) orc_cell_halo(nx, ny, global_bc, w + (size_t)q * nc);
        gradient_margin1(nx, ny, dx, dy, w, ice_mask, g1x, g1y);
        gradient_margin1(nx, ny, dx, dy, w + nc, ice_mask, g2x, g2y);
        gradient_margin1(nx, ny, dx, dy, w + 2 * nc, ice_mask, g3x, g3y);
        for (int j = jlo; j <= jhi; ++j)
            for (int i = ilo; i <= ihi; ++i) {
                const size_t p = V2(i, j);
                if (!active_vertex[p]) continue;
                const double depth = ss[k] * stagthck[p];
                const double txs = -RHOI * GRAV * depth * dusrf_dx[p];
                const double tys = -RHOI * GRAV * depth * dusrf_dy[p];
                const double txz = txs + 2.0 * g1x[p] + g2y[p];
                const double tyz = tys + g2x[p] + 2.0 * g3y[p];
                const double te = sq[p] + txz * txz + tyz * tyz;
                double fact;
                if (whichefvs == 2) {
                    fact = 2.0 * sf[p] * pow(te, (n_glen - 1.0) / 2.0) * (sigma[k + 1] - sigma[k]) * stagthck[p];
                } else {
                    const double ev = efvs[(size_t)k + (size_t)nl * C2(i, j)];
                    fact = ev > 0.0 ? (sigma[k + 1] - sigma[k]) * stagthck[p] / ev : 0.0;
                }
                const size_t a = (size_t)k + (size_t)nz * p, b = (size_t)(k + 1) + (size_t)nz * p,
                             base = (size_t)(nz - 1) + (size_t)nz * p;
                uvel[a] = umask[p] == 1 ? uvel[base] : uvel[b] + fact * txz;
                vvel[a] = vmask[p] == 1 ? vvel[base] : vvel[b] + fact * tyz;
            }
    }
done:
    free(ss); free(ux); free(uy); free(vx); free(vy); free(tp); free(ei); free(w); free(g);
    return st;
}

/* ------------------------------------------------------------------ */
/* Standard (quadrature) options of the load and the basal term        */
/* ------------------------------------------------------------------ */

/* load_vector_gravity, standard option (glissade_velo_higher.F90:4350-4446):
 * per element of an active cell and per quadrature point, ds/dx (ds/dy)
 * interpolated from the vertices; load(n) -= rhoi g w detJ / vol0 *
 * ds/dx_qp * phi_n, accumulated in the loop order (cells j, i; layer k;
 * point p; node n). */
int orc_load_vector_gravity_standard(int nx, int ny, int nz, int nhalo, const double* sigma,
                                     const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                     const double* stagusrf, const double* stagthck, const double* dusrf_dx,
                                     const double* dusrf_dy, double* loadu, double* loadv)
{
    init_tables();
    double x[8], y[8], z[8], sx[8], sy[8];
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                for (int n = 0; n < 8; ++n) {
                    const size_t q = V2(i + node_di[n], j + node_dj[n]);
                    const int kn = k + node_dk[n];
                    x[n] = xVertex[q];
                    y[n] = yVertex[q];
                    z[n] = stagusrf[q] - sigma[kn - 1] * stagthck[q];
                    sx[n] = dusrf_dx[q];
                    sy[n] = dusrf_dy[q];
                }
                for (int p = 0; p < 8; ++p) {
                    double dx[8], dy[8], dz[8], detJ;
                    if (basis_derivs(x, y, z, p, dx, dy, dz, &detJ)) return ORC_ERR_SINGULAR;
                    double gx = 0.0, gy = 0.0;
                    for (int n = 0; n < 8; ++n) {
                        gx = gx + phi3[n][p] * sx[n];
                        gy = gy + phi3[n][p] * sy[n];
                    }
                    for (int n = 0; n < 8; ++n) {
                        const size_t a = V3(k + node_dk[n], i + node_di[n], j + node_dj[n]);
                        loadu[a] = loadu[a] - RHOI * GRAV * 1.0 * detJ / VOL0 * gx * phi3[n][p];
                        loadv[a] = loadv[a] - RHOI * GRAV * 1.0 * detJ / VOL0 * gy * phi3[n][p];
                    }
                }
            }
        }
    return ORC_OK;
}

/* basal_sliding_bc_3d standard option (glissade_velo_higher.F90:8955-9111)
 * and basal_sliding_bc_2d standard option (:8386-8527): per active cell and
 * 2x2 quadrature point, K(r,c) = beta_qp w detJ / vol0 phi_r phi_c is added
 * to the kA = 0 couplings of the basal rows.  curved (3-D, the reference's
 * curved_2d_basal_jacobian = .true., :8914): detJ is the area factor of
 * the face (x, y, staglsrf) (get_area_scale_factor_curved_2d :6736-6827);
 * otherwise the flat 2-D Jacobian.  two_d: A are (nx-1, ny-1, 9), else the
 * level nz of (27, nz, nx-1, ny-1).  beta_y (NULL: beta) weights Avv; DIVA
 * passes beta_eff_x / beta_eff_y with the curved Jacobian
 * (basal_sliding_bc_2d_diva standard option, :8535-8811). */
int orc_basal_sliding_bc_standard(int nx, int ny, int nz, int nhalo, int two_d, int curved,
                                  const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                  const double* staglsrf, const double* beta, const double* beta_y, double* Auu,
                                  double* Avv)
{
    init_tables_2d();
    double x[4], y[4], z[4], b[4], by[4];
    for (int j = nhalo + 1; j <= ny - nhalo + 1; ++j)
        for (int i = nhalo + 1; i <= nx - nhalo + 1; ++i) {
            if (!active_cell[C2(i, j)]) continue;
            for (int n = 0; n < 4; ++n) {
                const size_t q = V2(i + di2[n], j + dj2[n]);
                x[n] = xVertex[q];
                y[n] = yVertex[q];
                z[n] = staglsrf ? staglsrf[q] : 0.0;
                b[n] = beta[q];
                by[n] = beta_y ? beta_y[q] : beta[q];
            }
            for (int p = 0; p < 4; ++p) {
                double detJ;
                if (curved) {
                    double J11 = 0, J12 = 0, J13 = 0, J21 = 0, J22 = 0, J23 = 0;
                    for (int n = 0; n < 4; ++n) {
                        J11 = J11 + dxr2[n][p] * x[n]; J12 = J12 + dxr2[n][p] * y[n]; J13 = J13 + dxr2[n][p] * z[n];
                        J21 = J21 + dyr2[n][p] * x[n]; J22 = J22 + dyr2[n][p] * y[n]; J23 = J23 + dyr2[n][p] * z[n];
                    }
                    const double t1 = (J11 * J22 - J12 * J21);
                    const double t2 = -(J11 * J23 - J13 * J21);
                    const double t3 = (J12 * J23 - J13 * J22);
                    detJ = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
                } else {
                    double dx[4], dy[4];
                    if (basis_derivs_2d(x, y, p, dx, dy, &detJ)) return ORC_ERR_SINGULAR;
                }
                double bq = 0.0, bqy = 0.0;
                for (int n = 0; n < 4; ++n) {
                    bq = bq + phi2[n][p] * b[n];
                    bqy = bqy + phi2[n][p] * by[n];
                }
                double K[4][4], Ky[4][4];
                for (int nc = 0; nc < 4; ++nc)
                    for (int nr = 0; nr < 4; ++nr) {
                        K[nr][nc] = 0.0 + bq * 1.0 * detJ / VOL0 * phi2[nr][p] * phi2[nc][p];
                        Ky[nr][nc] = 0.0 + bqy * 1.0 * detJ / VOL0 * phi2[nr][p] * phi2[nc][p];
                    }
                for (int nr = 0; nr < 4; ++nr) {
                    const int ii = i + di2[nr], jj = j + dj2[nr];
                    for (int nc = 0; nc < 4; ++nc) {
                        const int iA = di2[nc] - di2[nr], jA = dj2[nc] - dj2[nr];
                        const size_t a = two_d ? A2(IDX2(iA, jA), ii, jj) : A4(IDX3(iA, jA, 0), nz, ii, jj);
                        Auu[a] = Auu[a] + K[nr][nc];
                        Avv[a] = Avv[a] + Ky[nr][nc];
                    }
                }
            }
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */

/* compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042, called
 * :3795-3807) and the basal traction beside it (:3817-3818), on one task.
 * bfricflx (nx, ny), W/m^2:
 *   which 0 (standard): on owned active cells, beta u.u at the 2x2 Gauss
 *     points from the bilinear interpolants of the four vertex values,
 *     weights 1, divided by (sum of weights) scyr (:6917-6960);
 *   which 1 (local): beta (u^2 + v^2) / scyr at active vertices, 0 elsewhere,
 *     then glissade_unstagger's plain mean of the four vertices onto cells
 *     2..nx-1, 2..ny-1 (:6962-6978; glissade_grid_operators.F90:277-285);
 * then the cell halo of global_bc (:3807).  btractx/btracty (nx-1, ny-1) =
 * beta times the basal velocity at every vertex.  Any output may be NULL. */
void orc_basal_friction(int nx, int ny, int nhalo, int global_bc, int which, const int32_t* active_cell,
                        const int32_t* active_vertex, const double* ubas, const double* vbas, const double* beta,
                        double* bfricflx, double* btractx, double* btracty)
{
    const double scyr = 31536000.0;   /* glimmer_physcon.F90:90 */
    init_tables_2d();
    if (bfricflx) {
        for (size_t c = 0; c < (size_t)nx * ny; ++c) bfricflx[c] = 0.0;
        if (which == 0) {
            for (int j = nhalo + 1; j <= ny - nhalo; ++j)
                for (int i = nhalo + 1; i <= nx - nhalo; ++i) {
                    if (!active_cell[C2(i, j)]) continue;
                    double u[4], v[4], b[4];
                    for (int n = 0; n < 4; ++n) {
                        const size_t q = V2(i + di2[n], j + dj2[n]);
                        u[n] = ubas[q];
                        v[n] = vbas[q];
                        b[n] = beta[q];
                    }
                    double f = 0.0, sum_w = 0.0;
                    for (int p = 0; p < 4; ++p) {
                        double uq = 0.0, vq = 0.0, bq = 0.0;
                        for (int n = 0; n < 4; ++n) {
                            uq = uq + phi2[n][p] * u[n];
                            vq = vq + phi2[n][p] * v[n];
                            bq = bq + phi2[n][p] * b[n];
                        }
                        f = f + 1.0 * bq * (uq * uq + vq * vq);
                        sum_w = sum_w + 1.0;
                    }
                    bfricflx[C2(i, j)] = f / (sum_w * scyr);
                }
        } else {
            double* st = calloc((size_t)NVX * NVY, sizeof(double));
            for (int j = 1; j <= NVY; ++j)
                for (int i = 1; i <= NVX; ++i)
                    if (active_vertex[V2(i, j)]) {
                        const size_t q = V2(i, j);
                        st[q] = beta[q] * (ubas[q] * ubas[q] + vbas[q] * vbas[q]);
                        st[q] = st[q] / scyr;
                    }
            for (int j = 2; j <= ny - 1; ++j)
                for (int i = 2; i <= nx - 1; ++i)
                    bfricflx[C2(i, j)] = (st[V2(i, j)] + st[V2(i - 1, j)] + st[V2(i, j - 1)] + st[V2(i - 1, j - 1)]) / 4.0;
            free(st);
        }
        orc_cell_halo(nx, ny, global_bc, bfricflx);
    }
    for (size_t q = 0; q < (size_t)NVX * NVY; ++q) {
        if (btractx) btractx[q] = beta[q] * ubas[q];
        if (btracty) btracty[q] = beta[q] * vbas[q];
    }
}
