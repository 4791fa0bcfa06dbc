// assembly.hip -- device assembly of the 3-D Blatter-Pattyn system
// (glissade_velo_higher.F90): element stiffness with Glen's-law viscosity
// (:5382-5643, :6508-6732, :7285-7470, :8001-8123), gravity load (:4311-4342),
// local basal friction (:8921-8953), Dirichlet rows/columns (:9305-9498),
// the symmetry fix (:10854-11117), nonzero count (:10513-10579) and the
// residual (:9693-9894).
//
// Owner-computes gather: one thread per owned node builds that node's whole
// matrix row from the (up to) 8 elements that contain it.  No two threads
// write the same row, so the scatter needs neither atomics nor colouring
// passes, every coefficient is written exactly once, and the result is
// deterministic.  The default path (k_elem, k_rows) evaluates each element's
// quadrature-point data once and gathers the rows from it; k_assemble
// evaluates every element once per row it contributes to, visiting them in
// the reference's element order (cells j-major then i, layer above before
// layer below), the order of element_to_global_matrix_3d.
#include "cismgpu_internal.h"

namespace cg {

// Reference-element tables [n][p] (glissade_velo_higher.F90:352-388)
__constant__ double c_dxr[8][8], c_dyr[8][8], c_dzr[8][8], c_phi3[8][8];

// node offsets from node 7 of an element (ishift/jshift/kshift(7,:), :461-494)
__host__ __device__ constexpr int nd_i(int n) { return (n == 0 || n == 3 || n == 4 || n == 7) ? -1 : 0; }
__host__ __device__ constexpr int nd_j(int n) { return (n == 0 || n == 1 || n == 4 || n == 5) ? -1 : 0; }
__host__ __device__ constexpr int nd_k(int n) { return n < 4 ? 1 : 0; }

struct AsmDev {
    const double *xV, *yV, *susrf, *sthck, *sigma, *swbs, *flwafact, *dsdx, *dsdy, *beta, *theta;
    const double* thck_cell;     // (nx, ny) ice thickness, when given
    const double* flwa;          // (nz-1, nx, ny) rate factor, when given
    const double* slsrf;         // (nvx, nvy) staggered lower surface (standard basal option)
    int beta_std, taud_std;      // quadrature (standard) options of the basal term and the load
    const int32_t *acell, *avert, *umask, *vmask;
    double* efvs;
    double dx, dy, efvs_constant, effstrain_min, n_glen;
    int whichefvs, nhalo;
};

struct Elem { double x[8], y[8], z[8], u[8], v[8]; };

__device__ __forceinline__ long long c2(const Dims& d, int i, int j) { return (long long)(i - 1) + (long long)d.nx * (j - 1); }
__device__ __forceinline__ long long v2(const Dims& d, int i, int j) { return (long long)(i - 1) + (long long)d.nvx * (j - 1); }

// Nodes of element (cell ci,cj; layer kk, 1-based) from vertex fields and
// the velocity in device layout.
__device__ __forceinline__ void load_elem(const Dims& d, const AsmDev& in, const double* __restrict__ uv,
                                          const double* __restrict__ vv, int ci, int cj, int kk, Elem& e) {
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int in_ = ci + nd_i(n), jn = cj + nd_j(n), kn = kk + nd_k(n);
        const long long p = v2(d, in_, jn);
        e.x[n] = in.xV[p];
        e.y[n] = in.yV[p];
        e.z[n] = in.susrf[p] - in.sigma[kn - 1] * in.sthck[p];
        const long long c = vidx(d, kn - 1, in_, jn);
        e.u[n] = uv[c];
        e.v[n] = vv[c];
    }
}

// get_basis_function_derivatives_3d at quadrature point p; false if singular
__device__ __forceinline__ bool basis_derivs(const Elem& e, int p, double* dx, double* dy, double* dz, double& detJ) {
    double J00 = 0, J01 = 0, J02 = 0, J10 = 0, J11 = 0, J12 = 0, J20 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double a = c_dxr[n][p], b = c_dyr[n][p], c = c_dzr[n][p];
        J00 = J00 + a * e.x[n]; J01 = J01 + a * e.y[n]; J02 = J02 + a * e.z[n];
        J10 = J10 + b * e.x[n]; J11 = J11 + b * e.y[n]; J12 = J12 + b * e.z[n];
        J20 = J20 + c * e.x[n]; J21 = J21 + c * e.y[n]; J22 = J22 + c * e.z[n];
    }
    const double c00 = J11 * J22 - J12 * J21;
    const double c01 = -(J10 * J22 - J12 * J20);
    const double c02 = J10 * J21 - J11 * J20;
    const double c10 = -(J01 * J22 - J02 * J21);
    const double c11 = J00 * J22 - J02 * J20;
    const double c12 = -(J00 * J21 - J01 * J20);
    const double c20 = J01 * J12 - J02 * J11;
    const double c21 = -(J00 * J12 - J02 * J10);
    const double c22 = J00 * J11 - J01 * J10;
    detJ = J00 * c00 + J01 * c01 + J02 * c02;
    if (!(fabs(detJ) > 0.0)) return false;
    // Jinv(row,col) = cofactor(col,row) / detJ
    const double i00 = c00 / detJ, i01 = c10 / detJ, i02 = c20 / detJ;
    const double i10 = c01 / detJ, i11 = c11 / detJ, i12 = c21 / detJ;
    const double i20 = c02 / detJ, i21 = c12 / detJ, i22 = c22 / detJ;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double a = c_dxr[n][p], b = c_dyr[n][p], c = c_dzr[n][p];
        dx[n] = i00 * a + i01 * b + i02 * c;
        dy[n] = i10 * a + i11 * b + i12 * c;
        dz[n] = i20 * a + i21 * b + i22 * c;
    }
    return true;
}

// compute_effective_viscosity, BP branch, from the point's velocity gradients
__device__ __forceinline__ double efvs_from_strain(const AsmDev& in, double ux, double uy, double uz, double vx,
                                                   double vy, double vz, double flwafact) {
    const double p_effstr = (1.0 - in.n_glen) / in.n_glen;
    if (in.whichefvs == 0) return in.efvs_constant;
    if (in.whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        return flwafact * pow(effstrain, p_effstr);
    }
    const double t1 = ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy));
    const double t2 = 0.25 * (uz * uz + vz * vz);
    const double esq = in.effstrain_min * in.effstrain_min + t1 + t2;
    return flwafact * pow(esq, p_effstr / 2.0);
}

// compute_effective_viscosity, BP branch
__device__ __forceinline__ double efvs_at(const AsmDev& in, const Elem& e, const double* dx, const double* dy,
                                          const double* dz, double flwafact) {
    if (in.whichefvs == 0 || in.whichefvs == 1) return efvs_from_strain(in, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, flwafact);
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        ux = ux + dx[n] * e.u[n]; uy = uy + dy[n] * e.u[n]; uz = uz + dz[n] * e.u[n];
        vx = vx + dx[n] * e.v[n]; vy = vy + dy[n] * e.v[n]; vz = vz + dz[n] * e.v[n];
    }
    return efvs_from_strain(in, ux, uy, uz, vx, vy, vz, flwafact);
}

__device__ __forceinline__ bool cell_in_range(const Dims& d, int nhalo, int ci, int cj) {
    return ci >= nhalo + 1 && ci <= d.nx - nhalo + 1 && cj >= nhalo + 1 && cj <= d.ny - nhalo + 1;
}

// Row NR of the element matrix of (ci,cj,kk), summed over the 8 quadrature
// points exactly as compute_element_matrix accumulates K, then added into
// the node row accumulators acc[blk][m].
template <int NR>
__device__ __forceinline__ bool add_element_row(const Dims& d, const AsmDev& in, const double* __restrict__ uv,
                                                const double* __restrict__ vv, int ci, int cj, int kk,
                                                double (&acc)[4][27]) {
    Elem e;
    load_elem(d, in, uv, vv, ci, cj, kk, e);
    const double fl = in.flwafact[(long long)(kk - 1) + (long long)(d.nz - 1) * c2(d, ci, cj)];
    double K[4][8];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 8; ++c) K[b][c] = 0.0;
    bool ok = true;
    for (int p = 0; p < 8; ++p) {
        double dx[8], dy[8], dz[8], detJ;
        if (!basis_derivs(e, p, dx, dy, dz, detJ)) { ok = false; break; }
        const double ef = efvs_at(in, e, dx, dy, dz, fl);
        const double f = ef * 1.0 * detJ / 1.0e9;
        const double ar = dx[NR], br = dy[NR], cr = dz[NR];
#pragma unroll
        for (int nc = 0; nc < 8; ++nc) {
            K[0][nc] = K[0][nc] + f * (4.0 * ar * dx[nc] + br * dy[nc] + cr * dz[nc]);
            K[1][nc] = K[1][nc] + f * (2.0 * ar * dy[nc] + br * dx[nc]);
            K[2][nc] = K[2][nc] + f * (2.0 * br * dx[nc] + ar * dy[nc]);
            K[3][nc] = K[3][nc] + f * (4.0 * br * dy[nc] + ar * dx[nc] + cr * dz[nc]);
        }
    }
#pragma unroll
    for (int nc = 0; nc < 8; ++nc) {
        constexpr int dummy = 0;
        (void)dummy;
        const int m = (nd_i(nc) - nd_i(NR) + 1) + 3 * (nd_j(nc) - nd_j(NR) + 1) + 9 * (nd_k(nc) - nd_k(NR) + 1);
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b][m] = acc[b][m] + K[b][nc];
    }
    return ok;
}

// One thread per owned node (io fastest).  Writes the full row (108
// coefficients) into the row-blocked matrix.  flag[0] = 1 on a singular
// Jacobian.
__global__ __launch_bounds__(64) void k_assemble(Dims d, AsmDev in, const double* __restrict__ uv,
                                                 const double* __restrict__ vv, double* __restrict__ A,
                                                 double* __restrict__ flag) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo), k = (int)(kj / d.nyo);   // k 0-based level
    double acc[4][27];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int m = 0; m < 27; ++m) acc[b][m] = 0.0;
    if (io < d.nxo) {
        const int i = d.ilo + io, j = d.jlo + jo;
        const int kf = k + 1;   // 1-based level of this node
        if (in.avert[v2(d, i, j)]) {
            bool ok = true;
            // cells in reference order: (i,j), (i+1,j), (i,j+1), (i+1,j+1); in each,
            // layer kf-1 (node on the lower face) then layer kf (node on the upper face)
#define CELL(DI, DJ, NL, NU)                                                                       \
    {                                                                                              \
        const int ci = i + DI, cj = j + DJ;                                                        \
        if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[c2(d, ci, cj)]) {                       \
            if (kf - 1 >= 1) ok = add_element_row<NL>(d, in, uv, vv, ci, cj, kf - 1, acc) && ok;   \
            if (kf <= d.nz - 1) ok = add_element_row<NU>(d, in, uv, vv, ci, cj, kf, acc) && ok;    \
        }                                                                                          \
    }
            CELL(0, 0, 2, 6)
            CELL(1, 0, 3, 7)
            CELL(0, 1, 1, 5)
            CELL(1, 1, 0, 4)
#undef CELL
            if (!ok) flag[0] = 1.0;
        }
    }
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int m = 0; m < 27; ++m) a[(b * 27 + m) * WAVE] = acc[b][m];
}

// ---------------------------------------------------------------------------
// Element-once assembly.  Pass 1 (k_elem): one thread per element computes,
// at each of its 8 quadrature points, the inverse Jacobian and the factor
// f = efvs * w * detJ / vol0 (the Jacobian, its cofactor inverse, the
// strain rate and the Glen-law pow are done once per element instead of
// once per element row), and the element viscosity efvs(k,i,j) as the mean
// over the points (:5604-5610).  Pass 2 (k_rows): one thread per owned node
// builds its row from the stored point data of the (up to) 8 elements that
// contain it: the basis-function derivatives from the inverse Jacobian and
// the row of the element matrix, added point by point into the row.  The
// node's elements are visited layer by layer (the layer above the node,
// then the layer below); within a layer, point by point, and for each point
// the cells in the reference's cell order (j-major, then i).  This is one
// global order of the terms, and every product is formed as in
// compute_element_matrix, so entry (P, Q) and entry (Q, P) are summed from
// equal terms in the same order: the assembled matrix is exactly symmetric.
// Keeping the couplings to the level above until the first layer is done,
// and those to the level below from the second layer on, holds 72 of the
// row's 108 coefficients in registers.
//
// Point data at E[(p * EQ + q) * ne + e], q = 0..8 the inverse Jacobian
// (row-major), q = 9 the factor f; e = (ci-1) + nx ((cj-1) + ny (kk-1)),
// so neighbouring lanes (neighbouring i) read neighbouring words.
// ---------------------------------------------------------------------------
constexpr int EQ = 10;

__device__ __forceinline__ long long elem_id(const Dims& d, int ci, int cj, int kk) {
    return (long long)(ci - 1) + (long long)d.nx * ((cj - 1) + (long long)d.ny * (kk - 1));
}

// Inverse Jacobian at point p (get_basis_function_derivatives_3d,
// :6576-6617); false if singular.
__device__ __forceinline__ bool jac_inverse(const Elem& e, int p, double (&Ji)[9], double& detJ) {
    double J00 = 0, J01 = 0, J02 = 0, J10 = 0, J11 = 0, J12 = 0, J20 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double a = c_dxr[n][p], b = c_dyr[n][p], c = c_dzr[n][p];
        J00 = J00 + a * e.x[n]; J01 = J01 + a * e.y[n]; J02 = J02 + a * e.z[n];
        J10 = J10 + b * e.x[n]; J11 = J11 + b * e.y[n]; J12 = J12 + b * e.z[n];
        J20 = J20 + c * e.x[n]; J21 = J21 + c * e.y[n]; J22 = J22 + c * e.z[n];
    }
    const double c00 = J11 * J22 - J12 * J21;
    const double c01 = -(J10 * J22 - J12 * J20);
    const double c02 = J10 * J21 - J11 * J20;
    const double c10 = -(J01 * J22 - J02 * J21);
    const double c11 = J00 * J22 - J02 * J20;
    const double c12 = -(J00 * J21 - J01 * J20);
    const double c20 = J01 * J12 - J02 * J11;
    const double c21 = -(J00 * J12 - J02 * J10);
    const double c22 = J00 * J11 - J01 * J10;
    detJ = J00 * c00 + J01 * c01 + J02 * c02;
    if (!(fabs(detJ) > 0.0)) return false;
    Ji[0] = c00 / detJ; Ji[1] = c10 / detJ; Ji[2] = c20 / detJ;
    Ji[3] = c01 / detJ; Ji[4] = c11 / detJ; Ji[5] = c21 / detJ;
    Ji[6] = c02 / detJ; Ji[7] = c12 / detJ; Ji[8] = c22 / detJ;
    return true;
}

__global__ __launch_bounds__(128) void k_elem(Dims d, AsmDev in, const double* __restrict__ uv,
                                              const double* __restrict__ vv, double* __restrict__ E,
                                              double* __restrict__ efvs, double* __restrict__ flag) {
    const long long ne = (long long)(d.nz - 1) * d.nx * d.ny;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= ne) return;
    const int ci = 1 + (int)(t % d.nx);
    const long long r = t / d.nx;
    const int cj = 1 + (int)(r % d.ny), kk = 1 + (int)(r / d.ny);
    const long long cell = (long long)(ci - 1) + (long long)d.nx * (cj - 1);
    double out = 0.0;
    if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[cell]) {
        Elem e;
        load_elem(d, in, uv, vv, ci, cj, kk, e);
        const double fl = in.flwafact[(long long)(kk - 1) + (long long)(d.nz - 1) * cell];
        double s = 0.0;
        for (int p = 0; p < 8; ++p) {
            double Ji[9], detJ;
            if (!jac_inverse(e, p, Ji, detJ)) {
                flag[0] = 1.0;
                s = NAN;
                break;
            }
            double dx[8], dy[8], dz[8];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const double a = c_dxr[n][p], b = c_dyr[n][p], c = c_dzr[n][p];
                dx[n] = Ji[0] * a + Ji[1] * b + Ji[2] * c;
                dy[n] = Ji[3] * a + Ji[4] * b + Ji[5] * c;
                dz[n] = Ji[6] * a + Ji[7] * b + Ji[8] * c;
            }
            const double ef = efvs_at(in, e, dx, dy, dz, fl);
            s = s + ef;
            double* q = E + (long long)p * EQ * ne + t;
#pragma unroll
            for (int k = 0; k < 9; ++k) q[k * ne] = Ji[k];
            q[9 * ne] = ef * 1.0 * detJ / 1.0e9;
        }
        out = s / 8.0;
    }
    efvs[(long long)(kk - 1) + (long long)(d.nz - 1) * cell] = out;
}

// Tile of k_rows: 64 consecutive vertices (io) of ROWS_J consecutive vertex
// rows (jo) of one level, one wavefront per vertex row.  Its rows touch the
// cells (i0 .. i0+64) x (j0 .. j0+ROWS_J) of an element layer.
constexpr int ROWS_J = 4;
constexpr int RC_I = WAVE + 1;
constexpr int RC = RC_I * (ROWS_J + 1);
constexpr int RQ = 25;          // per cell and point: 8 x 3 basis derivatives, f

// Point p's contribution of one cell (tile cell c) to row NR of the element
// matrix, from the staged derivatives: couplings to the element's
// upper-face nodes (kshift 0) go to acc_up, those to its lower-face nodes to
// acc_lo, at index (iA + 1) + 3 (jA + 1) of the 9 horizontal neighbours.
// Every rounding is spelled out (products, fused multiply-adds), so that the
// (row, column) term and the (column, row) term of each block pair are
// evaluated by the same operations on the same values whatever the
// compiler's contraction choices: uu(P,Q) = uu(Q,P), vv(P,Q) = vv(Q,P) and
// uv(P,Q) = vu(Q,P) bit for bit.
template <int NR>
__device__ __forceinline__ void add_point(const double (*__restrict__ sd)[RC], int c, double (&acc_up)[4][9],
                                          double (&acc_lo)[4][9]) {
    const double f = sd[24][c];
    const double ar = sd[3 * NR][c], br = sd[3 * NR + 1][c], cr = sd[3 * NR + 2][c];
    const double ar2 = 2.0 * ar, br2 = 2.0 * br, ar4 = 4.0 * ar, br4 = 4.0 * br;   // exact
#pragma unroll
    for (int nc = 0; nc < 8; ++nc) {
        const double dxn = sd[3 * nc][c], dyn = sd[3 * nc + 1][c], dzn = sd[3 * nc + 2][c];
        const int h = (nd_i(nc) - nd_i(NR) + 1) + 3 * (nd_j(nc) - nd_j(NR) + 1);
        double(&acc)[4][9] = nd_k(nc) == 0 ? acc_up : acc_lo;
        const double kuu = __fma_rn(cr, dzn, __fma_rn(br, dyn, __dmul_rn(ar4, dxn)));
        const double kuv = __fma_rn(ar2, dyn, __dmul_rn(br, dxn));
        const double kvu = __fma_rn(br2, dxn, __dmul_rn(ar, dyn));
        const double kvv = __fma_rn(cr, dzn, __fma_rn(ar, dxn, __dmul_rn(br4, dyn)));
        acc[0][h] = __fma_rn(f, kuu, acc[0][h]);
        acc[1][h] = __fma_rn(f, kuv, acc[1][h]);
        acc[2][h] = __fma_rn(f, kvu, acc[2][h]);
        acc[3][h] = __fma_rn(f, kvv, acc[3][h]);
    }
}

// The four cells of one element layer around a node at one point, in the
// reference's cell order: (i,j), (i+1,j), (i,j+1), (i+1,j+1) as nodes
// N0..N3 of their elements.
template <int N0, int N1, int N2, int N3>
__device__ __forceinline__ void add_cells(const double (*__restrict__ sd)[RC], int c0, unsigned on,
                                          double (&acc_up)[4][9], double (&acc_lo)[4][9]) {
    if (on & 1u) add_point<N0>(sd, c0, acc_up, acc_lo);
    if (on & 2u) add_point<N1>(sd, c0 + 1, acc_up, acc_lo);
    if (on & 4u) add_point<N2>(sd, c0 + RC_I, acc_up, acc_lo);
    if (on & 8u) add_point<N3>(sd, c0 + RC_I + 1, acc_up, acc_lo);
}

// Rows on the rim of the owned region (first or last owned vertex row or
// column, all levels): their partners may lie outside it.
__device__ __forceinline__ bool rim_row(const Dims& d, int io, int jo) {
    return io == 0 || io == d.nxo - 1 || jo == 0 || jo == d.nyo - 1;
}

// The rim rows per level: the first and the last owned vertex rows, then
// the first and last columns in between; rim_row_at maps t < nz * that
// count to the row-blocked row index.
__host__ __device__ inline long long rim_rows_per_level(const Dims& d) {
    return (long long)d.nxo * (d.nyo > 1 ? 2 : 1) + (long long)(d.nyo > 2 ? d.nyo - 2 : 0) * (d.nxo > 1 ? 2 : 1);
}

__device__ __forceinline__ long long rim_row_at(const Dims& d, long long t) {
    const long long nr = rim_rows_per_level(d);
    const int k = (int)(t / nr);
    const long long r = t % nr;
    int io, jo;
    if (r < d.nxo) {
        io = (int)r; jo = 0;
    } else if (d.nyo > 1 && r < 2LL * d.nxo) {
        io = (int)(r - d.nxo); jo = d.nyo - 1;
    } else {
        const long long q = r - (long long)d.nxo * (d.nyo > 1 ? 2 : 1);
        const int w = d.nxo > 1 ? 2 : 1;
        jo = 1 + (int)(q / w);
        io = (q % w == 0) ? 0 : d.nxo - 1;
    }
    return ((long long)k * d.nyo + jo) * d.nxp + io;
}

// One thread per owned node: the node's full row (108 coefficients) into
// the row-blocked matrix; one workgroup per tile.  For each element layer
// and quadrature point, the workgroup first computes the basis-function
// derivatives of every active cell of the tile from its stored inverse
// Jacobian (once per cell, not once per row) into LDS, then each thread adds
// the point's terms of its (up to) 4 cells.  Each coefficient sums its
// terms point by point and, within a point, cell by cell in one global
// element order (layer, then j, then i), and entry (P, Q) and entry (Q, P)
// are formed from the same staged values, so the assembled matrix is
// exactly symmetric.  The layer below the node first (couplings to the
// level above complete after it), then the layer above (the level below):
// 72 of the 108 coefficients live at a time.  With nnz given, the nonzero
// coefficients of the active rows off the rim are counted into nnz[0]
// (count_nonzeros_3d, :10513-10579; the rim rows are counted after the
// boundary passes by k_symmetry_rim).
__host__ __device__ inline long long rows_blocks(const Dims& d) {
    return (long long)(d.nxp / WAVE) * ((d.nyo + ROWS_J - 1) / ROWS_J) * d.nz;
}

__global__ __launch_bounds__(WAVE * ROWS_J) void k_rows(Dims d, AsmDev in, const double* __restrict__ E,
                                                        double* __restrict__ A, unsigned long long* __restrict__ nnz) {
    constexpr int NT = WAVE * ROWS_J;
    __shared__ double sd[RQ][RC];
    const int nbx = d.nxp / WAVE, nbj = (d.nyo + ROWS_J - 1) / ROWS_J;
    const long long b = blockIdx.x;
    const int bx = (int)(b % nbx);
    const long long rb = b / nbx;
    const int jt = (int)(rb % nbj) * ROWS_J;                    // first vertex row of the tile
    const int k = (int)(rb / nbj);                              // 0-based level (uniform)
    const int w = (int)(threadIdx.x / WAVE), lane = (int)(threadIdx.x % WAVE);
    const int jo = jt + w, io = bx * WAVE + lane;
    const bool rowok = jo < d.nyo;
    const long long ne = (long long)(d.nz - 1) * d.nx * d.ny;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int ci0 = d.ilo + bx * WAVE, cj0 = d.jlo + jt;        // tile cell (0, 0)
    const int kf = k + 1;
    const bool on = rowok && io < d.nxo && in.avert[v2(d, i, j)];
    auto cell_on = [&](int ci, int cj) { return cell_in_range(d, in.nhalo, ci, cj) && in.acell[c2(d, ci, cj)]; };
    const unsigned cells = on ? ((cell_on(i, j) ? 1u : 0u) | (cell_on(i + 1, j) ? 2u : 0u) |
                                 (cell_on(i, j + 1) ? 4u : 0u) | (cell_on(i + 1, j + 1) ? 8u : 0u))
                              : 0u;
    const int c0 = w * RC_I + lane;
    // derivatives of the tile's active cells of layer kk at point p
    auto stage = [&](int kk, int p) {
        for (int c = threadIdx.x; c < RC; c += NT) {
            const int ci = ci0 + c % RC_I, cj = cj0 + c / RC_I;
            if (!cell_on(ci, cj)) continue;
            const double* q = E + (long long)p * EQ * ne + elem_id(d, ci, cj, kk);
            double Ji[9];
#pragma unroll
            for (int m = 0; m < 9; ++m) Ji[m] = q[m * ne];
            sd[24][c] = q[9 * ne];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const double a = c_dxr[n][p], bb = c_dyr[n][p], cc = c_dzr[n][p];
                sd[3 * n][c] = Ji[0] * a + Ji[1] * bb + Ji[2] * cc;
                sd[3 * n + 1][c] = Ji[3] * a + Ji[4] * bb + Ji[5] * cc;
                sd[3 * n + 2][c] = Ji[6] * a + Ji[7] * bb + Ji[8] * cc;
            }
        }
    };
    const bool have0 = kf - 1 >= 1, have1 = kf <= d.nz - 1;     // uniform over the workgroup
    const bool count_row = nnz && on && !rim_row(d, io, jo);
    unsigned z = 0;
    const long long t = ((long long)k * d.nyo + (rowok ? jo : 0)) * d.nxp + io;
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    double mid[4][9];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int h = 0; h < 9; ++h) mid[bb][h] = 0.0;
    {
        // layer kf - 1: the node is on the elements' lower face
        double above[4][9];
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
#pragma unroll
            for (int h = 0; h < 9; ++h) above[bb][h] = 0.0;
        if (have0) {
#pragma unroll 1
            for (int p = 0; p < 8; ++p) {
                stage(kf - 1, p);
                __syncthreads();
                if (cells) add_cells<2, 3, 1, 0>(sd, c0, cells, above, mid);
                __syncthreads();
            }
        }
        if (rowok) {
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int h = 0; h < 9; ++h) a[(bb * 27 + h) * WAVE] = above[bb][h];
        }
        if (count_row) {
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int h = 0; h < 9; ++h) z += above[bb][h] != 0.0 ? 1u : 0u;
        }
    }
    double below[4][9];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int h = 0; h < 9; ++h) below[bb][h] = 0.0;
    if (have1) {
        // layer kf: the node is on the elements' upper face
#pragma unroll 1
        for (int p = 0; p < 8; ++p) {
            stage(kf, p);
            __syncthreads();
            if (cells) add_cells<6, 7, 5, 4>(sd, c0, cells, mid, below);
            __syncthreads();
        }
    }
    if (rowok) {
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                a[(bb * 27 + 9 + h) * WAVE] = mid[bb][h];
                a[(bb * 27 + 18 + h) * WAVE] = below[bb][h];
            }
    }
    if (nnz) {
        if (count_row) {
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int h = 0; h < 9; ++h) z += (mid[bb][h] != 0.0 ? 1u : 0u) + (below[bb][h] != 0.0 ? 1u : 0u);
        }
        unsigned long long zz = z;
        for (int off = WAVE / 2; off > 0; off >>= 1) zz += __shfl_down(zz, off, WAVE);
        if (lane == 0 && zz) atomicAdd(nnz, zz);
    }
}

// compute_internal_stress (glissade_velo_higher.F90:7046-7281): per active
// element in the assembly range, the mean over the 8 quadrature points of
// tau_xz = efvs du/dz, tau_yz = efvs dv/dz, tau_xx = 2 efvs du/dx,
// tau_yy = 2 efvs dv/dy, tau_xy = efvs (dv/dx + du/dy), and tau_eff =
// sqrt(xx^2 + yy^2 + xx yy + xy^2 + xz^2 + yz^2); zero elsewhere.  Six fields
// of (nz-1, nx, ny) at tau + q nfl; one thread per element.  approx is
// whichapprox: SIA (0) takes only the vertical shear, SSA (1) only the
// horizontal strain rates (the viscosity of those alone: the strain-rate
// sums with the dropped terms zero round the same), BP (2) and DIVA (4)
// every one; L1L2 (3) every one with the viscosity efvs_l1l2(k,i,j) of its
// assembly.  Velocity (vmode): 0 level k of a 3-D system vector, 1 the one
// level of a 2-D system at every k (SSA), 2 plain (nz, nx-1, ny-1) arrays.
__global__ void k_internal_stress(Dims d, AsmDev in, int approx, int vmode, const double* __restrict__ uv,
                                  const double* __restrict__ vv, const double* __restrict__ efvs_l1l2,
                                  double* __restrict__ tau) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)(d.nz - 1) * d.nx * d.ny;
    if (t >= n) return;
    const int kk = 1 + (int)(t % (d.nz - 1));
    const long long cell = t / (d.nz - 1);
    const int ci = 1 + (int)(cell % d.nx), cj = 1 + (int)(cell / d.nx);
    double txz = 0.0, tyz = 0.0, txx = 0.0, tyy = 0.0, txy = 0.0, teff = 0.0;
    if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[c2(d, ci, cj)]) {
        Elem e;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int in_ = ci + nd_i(q), jn = cj + nd_j(q), kn = kk + nd_k(q);
            const long long pv = v2(d, in_, jn);
            e.x[q] = in.xV[pv];
            e.y[q] = in.yV[pv];
            e.z[q] = in.susrf[pv] - in.sigma[kn - 1] * in.sthck[pv];
            const long long c = vmode == 0 ? vidx(d, kn - 1, in_, jn)
                              : vmode == 1 ? vidx(d, 0, in_, jn) : (long long)(kn - 1) + (long long)d.nz * pv;
            e.u[q] = uv[c];
            e.v[q] = vv[c];
        }
        const double fl = in.flwafact[t];
        const bool vert = approx != 1, horiz = approx != 0;
        for (int p = 0; p < 8; ++p) {
            double dx[8], dy[8], dz[8], detJ;
            if (!basis_derivs(e, p, dx, dy, dz, detJ)) { txz = tyz = txx = tyy = txy = NAN; break; }
            double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (vert) { uz = uz + dz[q] * e.u[q]; vz = vz + dz[q] * e.v[q]; }
                if (horiz) {
                    ux = ux + dx[q] * e.u[q]; uy = uy + dy[q] * e.u[q];
                    vx = vx + dx[q] * e.v[q]; vy = vy + dy[q] * e.v[q];
                }
            }
            const double ef = approx == 3 ? efvs_l1l2[t] : efvs_from_strain(in, ux, uy, uz, vx, vy, vz, fl);
            txz = txz + ef * uz;
            tyz = tyz + ef * vz;
            txx = txx + 2.0 * ef * ux;
            tyy = tyy + 2.0 * ef * vy;
            txy = txy + ef * (vx + uy);
        }
        txz = txz / 8.0; tyz = tyz / 8.0; txx = txx / 8.0; tyy = tyy / 8.0; txy = txy / 8.0;
        teff = sqrt(txx * txx + tyy * tyy + txx * tyy + txy * txy + txz * txz + tyz * tyz);
    }
    tau[t] = txz;
    tau[n + t] = tyz;
    tau[2 * n + t] = txx;
    tau[3 * n + t] = tyy;
    tau[4 * n + t] = txy;
    tau[5 * n + t] = teff;
}

// efvs(k,i,j) = mean over the quadrature points (:5604-5610), one thread
// per element in the assembly range.
__global__ void k_efvs(Dims d, AsmDev in, const double* __restrict__ uv, const double* __restrict__ vv,
                       double* __restrict__ efvs) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)(d.nz - 1) * d.nx * d.ny;
    if (t >= n) return;
    const int kk = 1 + (int)(t % (d.nz - 1));
    const long long cell = t / (d.nz - 1);
    const int ci = 1 + (int)(cell % d.nx), cj = 1 + (int)(cell / d.nx);
    double out = 0.0;
    if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[c2(d, ci, cj)]) {
        Elem e;
        load_elem(d, in, uv, vv, ci, cj, kk, e);
        const double fl = in.flwafact[t];
        double s = 0.0;
        for (int p = 0; p < 8; ++p) {
            double dx[8], dy[8], dz[8], detJ;
            if (!basis_derivs(e, p, dx, dy, dz, detJ)) { s = NAN; break; }
            s = s + efvs_at(in, e, dx, dy, dz, fl);
        }
        out = s / 8.0;
    }
    efvs[t] = out;
}

// b = load_vector_gravity, local option, over every active vertex
// (:4311-4342); b starts from zero.
__global__ void k_load_gravity(Dims d, AsmDev in, double* __restrict__ bu, double* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const double rhoi = 910.0, grav = 9.81, vol0 = 1.0e9;
    const bool act = in.avert[t] != 0;
    for (int k = 1; k <= d.nz; ++k) {
        const long long c = vidx(d, k - 1, i, j);
        double lu = 0.0, lv = 0.0;
        if (act) {
            const double dz = in.sthck[t] * (in.swbs[k] - in.swbs[k - 1]);
            lu = lu - rhoi * grav * in.dx * in.dy * dz / vol0 * in.dsdx[t];
            lv = lv - rhoi * grav * in.dx * in.dy * dz / vol0 * in.dsdy[t];
        }
        bu[c] = lu;
        bv[c] = lv;
    }
}

// load_vector_gravity, standard option (glissade_velo_higher.F90:4350-4446),
// gathered per node: node (k, i, j) takes the terms of the elements of the
// active cells (i, j), (i+1, j), (i, j+1), (i+1, j+1) -- the order in which
// the reference visits them -- layer k-1 before layer k, points in order:
// load -= rhoi g w detJ / vol0 * ds/dx_qp * phi_n.  Every node is written.
__global__ void k_load_gravity_std(Dims d, AsmDev in, double* __restrict__ bu, double* __restrict__ bv,
                                   double* __restrict__ flag) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const double rhoi = 910.0, grav = 9.81, vol0 = 1.0e9;
    for (int kn = 1; kn <= d.nz; ++kn) {
        double lu = 0.0, lv = 0.0;
        for (int cc = 0; cc < 4; ++cc) {
            const int ci = i + (cc & 1), cj = j + (cc >> 1);
            if (ci > d.nx || cj > d.ny || !cell_in_range(d, in.nhalo, ci, cj) || !in.acell[c2(d, ci, cj)]) continue;
            for (int ke = kn - 1; ke <= kn; ++ke) {
                if (ke < 1 || ke > d.nz - 1) continue;
                int nl = 0;
#pragma unroll
                for (int n = 0; n < 8; ++n)
                    if (nd_i(n) == i - ci && nd_j(n) == j - cj && nd_k(n) == kn - ke) nl = n;
                Elem e;
                double sx[8], sy[8];
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    const long long q = v2(d, ci + nd_i(n), cj + nd_j(n));
                    e.x[n] = in.xV[q];
                    e.y[n] = in.yV[q];
                    e.z[n] = in.susrf[q] - in.sigma[ke + nd_k(n) - 1] * in.sthck[q];
                    sx[n] = in.dsdx[q];
                    sy[n] = in.dsdy[q];
                }
                for (int p = 0; p < 8; ++p) {
                    double dxn[8], dyn[8], dzn[8], detJ;
                    if (!basis_derivs(e, p, dxn, dyn, dzn, detJ)) { flag[0] = 1.0; return; }
                    double gx = 0.0, gy = 0.0;
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        gx = gx + c_phi3[n][p] * sx[n];
                        gy = gy + c_phi3[n][p] * sy[n];
                    }
                    lu = lu - rhoi * grav * 1.0 * detJ / vol0 * gx * c_phi3[nl][p];
                    lv = lv - rhoi * grav * 1.0 * detJ / vol0 * gy * c_phi3[nl][p];
                }
            }
        }
        const long long c = vidx(d, kn - 1, i, j);
        bu[c] = lu;
        bv[c] = lv;
    }
}

// ---------------------------------------------------------------------------
// Lateral ocean-pressure BC at calving fronts (load_vector_lateral_bc /
// lateral_shelf_bc, glissade_velo_higher.F90:4453-4957), gathered per node.
// The reference visits active cells (j-major, loop range nhalo+1..n-nhalo+1)
// and their ocean-facing faces (west, east, south, north), and for each
// face layer k and 2x2 Gauss point p adds the term of every face node.  A
// node (k, i, j) is touched by at most one u-face and one v-face per
// neighbouring cell; visiting those cells in (j, i) order, then layer k-1
// (node below) before layer k (node above), then p, reproduces the
// reference's order of additions into each load entry.
// ---------------------------------------------------------------------------
struct LatDev {
    const int32_t* ocean;            // (nx, ny)
    const double *usrf, *thck;       // (nx, ny)
    const double *susrf, *sthck;     // (nvx, nvy): staggered over ice-covered or land cells
    int local;                       // which_ho_assemble_lateral: 1 cell values, 0 staggered
};

// Adds the terms of face `face` (0 W, 1 E, 2 S, 3 N) of cell (ic, jc) that
// load node (k, face node hn in {0, 1}) to acc.
__device__ void lateral_face_node(const Dims& d, const AsmDev& in, const LatDev& L, int face, int ic, int jc,
                                  int hn, int k, double& acc) {
#pragma clang fp contract(off)
    const double rhoi = 910.0, rhoo = 1028.0, grav = 9.81, vol0 = 1.0e9;
    const double r = 1.0 / sqrt(3.0);
    const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r};
    int iN0, jN0, iN1, jN1;
    if (face == 0) { iN0 = ic - 1; jN0 = jc; iN1 = ic - 1; jN1 = jc - 1; }
    else if (face == 1) { iN0 = ic; jN0 = jc - 1; iN1 = ic; jN1 = jc; }
    else if (face == 2) { iN0 = ic - 1; jN0 = jc - 1; iN1 = ic; jN1 = jc - 1; }
    else { iN0 = ic; jN0 = jc; iN1 = ic - 1; jN1 = jc; }
    const double* coord = face <= 1 ? in.yV : in.xV;
    const double x0 = coord[v2(d, iN0, jN0)], x1 = coord[v2(d, iN1, jN1)];
    const double x[4] = {x0, x1, x1, x0};
    double s[4], h[4];
    if (L.local) {
        const double sc = L.usrf[c2(d, ic, jc)], hc = L.thck[c2(d, ic, jc)];
        for (int n = 0; n < 4; ++n) { s[n] = sc; h[n] = hc; }
    } else {
        s[0] = L.susrf[v2(d, iN0, jN0)]; s[1] = L.susrf[v2(d, iN1, jN1)]; s[2] = s[1]; s[3] = s[0];
        h[0] = L.sthck[v2(d, iN0, jN0)]; h[1] = L.sthck[v2(d, iN1, jN1)]; h[2] = h[1]; h[3] = h[0];
    }
    const double sgn = (face == 0 || face == 2) ? -1.0 : 1.0;
    // layer k-1 (the node is a lower node n = hn), then layer k (upper node n = 3 - hn)
    for (int pass = 0; pass < 2; ++pass) {
        const int ke = k - 1 + pass;                  // element layer, 1-based
        if (ke < 1 || ke > d.nz - 1) continue;
        const int n = pass == 0 ? hn : 3 - hn;
        const double y[4] = {s[0] - in.sigma[ke] * h[0], s[1] - in.sigma[ke] * h[1],
                             s[2] - in.sigma[ke - 1] * h[2], s[3] - in.sigma[ke - 1] * h[3]};
        for (int p = 0; p < 4; ++p) {
            double phi[4], dxr[4], dyr[4];
            phi[0] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
            phi[1] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
            phi[2] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
            phi[3] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
            dxr[0] = -(1.0 - yq[p]) / 4.0;
            dxr[1] = (1.0 - yq[p]) / 4.0;
            dxr[2] = (1.0 + yq[p]) / 4.0;
            dxr[3] = -(1.0 + yq[p]) / 4.0;
            dyr[0] = -(1.0 - xq[p]) / 4.0;
            dyr[1] = -(1.0 + xq[p]) / 4.0;
            dyr[2] = (1.0 + xq[p]) / 4.0;
            dyr[3] = (1.0 - xq[p]) / 4.0;
            double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
            for (int q = 0; q < 4; ++q) {
                J11 = J11 + dxr[q] * x[q];
                J12 = J12 + dxr[q] * y[q];
                J21 = J21 + dyr[q] * x[q];
                J22 = J22 + dyr[q] * y[q];
            }
            const double detJ = fabs(J11 * J22 - J12 * J21);
            double hq, sq;
            if (L.local) {
                hq = h[0];
                sq = s[0];
            } else {
                hq = 0.0;
                sq = 0.0;
                for (int q = 0; q < 4; ++q) {
                    hq = hq + phi[q] * h[q];
                    sq = sq + phi[q] * s[q];
                }
            }
            const double f = 1.0 - fmin(sq / hq, 1.0);
            const double p_av = 0.5 * rhoi * grav * hq - 0.5 * rhoo * grav * hq * (f * f);
            const double t = p_av * 1.0 * detJ / vol0 * phi[n];
            acc = sgn < 0.0 ? acc - t : acc + t;
        }
    }
}

// One thread per vertex column; b already holds the gravity load.
__global__ void k_lateral_bc(Dims d, AsmDev in, LatDev L, double* __restrict__ bu, double* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const int lo = in.nhalo + 1, hix = d.nx - in.nhalo + 1, hiy = d.ny - in.nhalo + 1;
    auto face_on = [&](int ic, int jc, int face) -> bool {
        if (ic < lo || ic > hix || jc < lo || jc > hiy) return false;
        if (!in.acell[c2(d, ic, jc)]) return false;
        const int oi = ic + (face == 0 ? -1 : face == 1 ? 1 : 0), oj = jc + (face == 2 ? -1 : face == 3 ? 1 : 0);
        return L.ocean[c2(d, oi, oj)] == 1;
    };
    // faces touching vertex (i, j), in the reference's cell order; (cell, face, face node)
    const int ucell[4][4] = {{i, j, 1, 1}, {i + 1, j, 0, 0}, {i, j + 1, 1, 0}, {i + 1, j + 1, 0, 1}};
    const int vcell[4][4] = {{i, j, 3, 0}, {i + 1, j, 3, 1}, {i, j + 1, 2, 1}, {i + 1, j + 1, 2, 0}};
    bool uon[4], von[4], any = false;
    for (int q = 0; q < 4; ++q) {
        uon[q] = face_on(ucell[q][0], ucell[q][1], ucell[q][2]);
        von[q] = face_on(vcell[q][0], vcell[q][1], vcell[q][2]);
        any = any || uon[q] || von[q];
    }
    if (!any) return;
    for (int k = 1; k <= d.nz; ++k) {
        const long long c = vidx(d, k - 1, i, j);
        double au = bu[c], av = bv[c];
        for (int q = 0; q < 4; ++q)
            if (uon[q]) lateral_face_node(d, in, L, ucell[q][2], ucell[q][0], ucell[q][1], ucell[q][3], k, au);
        for (int q = 0; q < 4; ++q)
            if (von[q]) lateral_face_node(d, in, L, vcell[q][2], vcell[q][0], vcell[q][1], vcell[q][3], k, av);
        bu[c] = au;
        bv[c] = av;
    }
}

// basal_sliding_bc_3d, local option: diagonal of the basal level
// (:8921-8953), owned rows.
__global__ void k_basal_local(Dims d, AsmDev in, double* __restrict__ A) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nxo * d.nyo) return;
    const int io = (int)(t % d.nxo), jo = (int)(t / d.nxo);
    const int i = d.ilo + io, j = d.jlo + jo;
    const long long p = v2(d, i, j);
    if (!in.avert[p]) return;
    const long long r = ((long long)(d.nz - 1) * d.nyo + jo) * d.nxp + io;
    double* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
    const double add = in.dx * in.dy / 1.0e9 * in.beta[p] / cos(in.theta[p]);
    a[(0 * 27 + 13) * WAVE] = a[(0 * 27 + 13) * WAVE] + add;
    a[(3 * 27 + 13) * WAVE] = a[(3 * 27 + 13) * WAVE] + add;
}

// dirichlet_boundary_conditions_3d restated per owned row P: the row of a
// prescribed node becomes the identity; entries of P that couple to a
// prescribed active node Q in the loop range are moved to the rhs, in the
// reference's (j, i, k; u then v) order of Q.
__global__ void k_dirichlet(Dims d, AsmDev in, const double* __restrict__ uv, const double* __restrict__ vv,
                            double* __restrict__ A, double* __restrict__ bu, double* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo), k0 = (int)(kj / d.nyo);
    if (io >= d.nxo) return;
    const int i = d.ilo + io, j = d.jlo + jo, k = k0 + 1;
    const int nh = in.nhalo;
    auto vec3 = [&](int kk, int ii, int jj) { return (long long)(kk - 1) + (long long)d.nz * v2(d, ii, jj); };
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    const long long pc = vidx(d, k0, i, j);
    const bool inrange = i >= nh && i <= d.nx - nh + 1 && j >= nh && j <= d.ny - nh + 1;
    const bool pact = in.avert[v2(d, i, j)] != 0;
    const int um = in.umask[vec3(k, i, j)], vm = in.vmask[vec3(k, i, j)];
    double bU = bu[pc], bV = bv[pc];
    // column elimination from prescribed neighbours Q, in Q's loop order
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
            for (int dk = -1; dk <= 1; ++dk) {
                // Q = (k+dk, i+di, j+dj); loop order over Q is j, then i, then k
                const int qi = i + di, qj = j + dj, qk = k + dk;
                if (di == 0 && dj == 0 && dk == 0) continue;
                if (qk < 1 || qk > d.nz) continue;
                if (!(qi >= nh && qi <= d.nx - nh + 1 && qj >= nh && qj <= d.ny - nh + 1)) continue;
                if (!in.avert[v2(d, qi, qj)]) continue;
                const long long q = vec3(qk, qi, qj);
                const int m = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);   // entry of row P coupling to Q
                const long long qc = vidx(d, qk - 1, qi, qj);
                if (in.umask[q] == 1) {
                    if (um != 1) { bU = bU - a[(0 * 27 + m) * WAVE] * uv[qc]; a[(0 * 27 + m) * WAVE] = 0.0; }
                    if (vm != 1) { bV = bV - a[(2 * 27 + m) * WAVE] * uv[qc]; a[(2 * 27 + m) * WAVE] = 0.0; }
                }
                if (in.vmask[q] == 1) {
                    if (um != 1) { bU = bU - a[(1 * 27 + m) * WAVE] * vv[qc]; a[(1 * 27 + m) * WAVE] = 0.0; }
                    if (vm != 1) { bV = bV - a[(3 * 27 + m) * WAVE] * vv[qc]; a[(3 * 27 + m) * WAVE] = 0.0; }
                }
            }
    // own row (the reference does this when it visits P; entries of row P are
    // touched by no other node's visit, so the order is immaterial)
    if (inrange && pact) {
        if (um == 1) {
            bU = uv[pc];
            for (int m = 0; m < 27; ++m) {
                const int kA = m / 9 - 1;
                if (k + kA < 1 || k + kA > d.nz) continue;
                a[(0 * 27 + m) * WAVE] = (m == 13) ? 1.0 : 0.0;
                a[(1 * 27 + m) * WAVE] = 0.0;
                if (m == 13) a[(2 * 27 + m) * WAVE] = 0.0;
            }
        }
        if (vm == 1) {
            bV = vv[pc];
            for (int m = 0; m < 27; ++m) {
                const int kA = m / 9 - 1;
                if (k + kA < 1 || k + kA > d.nz) continue;
                a[(2 * 27 + m) * WAVE] = 0.0;
                a[(3 * 27 + m) * WAVE] = (m == 13) ? 1.0 : 0.0;
                if (m == 13) a[(1 * 27 + m) * WAVE] = 0.0;
            }
        }
    }
    bu[pc] = bU;
    bv[pc] = bV;
}

// Halo of the assembled matrix for the no_ice rule: rows of the last owned
// vertex row/column are zeroed (parallel_mpi.F90:10912-10921 applied by
// staggered_parallel_halo_real8_4d to A, glissade_velo_higher.F90:2893-2898).
// Those rows are rim rows: one thread per rim row.
__global__ void k_matrix_noice_edge(Dims d, double* __restrict__ A) {
    const long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (u >= rim_rows_per_level(d) * d.nz) return;
    const long long t = rim_row_at(d, u);
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo);
    if (io >= d.nxo) return;
    const int i = d.ilo + io, j = d.jlo + jo;
    if (i < d.nvx - 1 && j < d.nvy - 1) return;
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    for (int c = 0; c < NCOEF; ++c) a[c * WAVE] = 0.0;
}

// Row of a vertex after the halo update: owned -> itself; periodic halo ->
// its owned image; other rules -> none (zeroed halo rows).  Returns false
// when the row is zero.
__device__ __forceinline__ bool row_image(const Dims& d, int& i, int& j) {
    const bool periodic = (d.bc == 0 || d.bc == 2);
    const int n_i = d.ihi - d.ilo + 1, n_j = d.jhi - d.jlo + 1;
    if (i < d.ilo || i > d.ihi) {
        if (!periodic) return false;
        i = d.ilo + ((i - d.ilo) % n_i + n_i) % n_i;
    }
    if (j < d.jlo || j > d.jhi) {
        if (!periodic) return false;
        j = d.jlo + ((j - d.jlo) % n_j + n_j) % n_j;
    }
    return true;
}

__device__ __forceinline__ const double* row_ptr(const Dims& d, const double* A, int k0, int i, int j) {
    const long long r = ((long long)k0 * d.nyo + (j - d.jlo)) * d.nxp + (i - d.ilo);
    return A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
}

// check_symmetry_assembled_matrix_3d, read-only: the number of coupling
// pairs that differ (an entry and its transposed partner: the partner
// row's entry, or zero where the partner has no stored row).  Zero means
// the symmetry fix would change nothing and the matrix is exactly
// symmetric.  The same pass counts the nonzero entries of the owned active
// rows into nnz.
//
// rim_only: only the rim rows are checked (every pair from their side) and
// counted.  For the element-once rows with no pass after them that changes
// couplings (no Dirichlet nodes, local basal friction), a pair of owned
// nodes is summed in both rows from the same stored point data, with equal
// products, in one global element order, so it is equal bit for bit; only
// a pair with a node outside the owned region (a halo vertex or a periodic
// image, whose elements are other copies) or a rim row zeroed by the no-ice
// rule can differ, and each of those has a rim row.
// The checks and the nonzero count of row t (rows-blocked index); all
// pairs from this row's side when all_pairs, else each pair from one side.
__device__ __forceinline__ void sym_row(const Dims& d, const int32_t* __restrict__ avert,
                                        const double* __restrict__ A, long long t, bool all_pairs,
                                        unsigned long long& c, unsigned long long& z) {
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo), k0 = (int)(kj / d.nyo);
    const int i = d.ilo + io, j = d.jlo + jo;
    if (io >= d.nxo || !avert[(long long)(i - 1) + (long long)d.nvx * (j - 1)]) return;
    const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    // a lower coupling (m < 13) against its stored partner, the centre's uv
    // against its vu, and any coupling to a node without a stored row
    // against zero
    c += (a[(1 * 27 + 13) * WAVE] != a[(2 * 27 + 13) * WAVE]) ? 1ull : 0ull;
    for (int m = 0; m < 27; ++m) {
        if (m == 13) continue;
        const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
        if (k0 + kA < 0 || k0 + kA >= d.nz) continue;
        int qi = i + iA, qj = j + jA;
        const double* q = row_image(d, qi, qj) ? row_ptr(d, A, k0 + kA, qi, qj) : nullptr;
        // an active stored partner checks the pair from its side (its m < 13)
        if (!all_pairs && q && m > 13 && avert[(long long)(qi - 1) + (long long)d.nvx * (qj - 1)]) continue;
        const int mm = 26 - m;
        const int ob[4] = {0, 1, 3, 2}, pb[4] = {0, 2, 3, 1};
        for (int s = 0; s < 4; ++s) {
            const double v1 = a[(ob[s] * 27 + m) * WAVE];
            const double v2_ = q ? q[(pb[s] * 27 + mm) * WAVE] : 0.0;
            c += (v1 != v2_) ? 1ull : 0ull;
        }
    }
    // count_nonzeros_3d of the same row (:10513-10579)
    for (int cf = 0; cf < NCOEF; ++cf) z += (a[cf * WAVE] != 0.0) ? 1ull : 0ull;
}

__device__ __forceinline__ void sym_flush(unsigned long long c, unsigned long long z, unsigned long long* count,
                                          unsigned long long* nnz) {
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        c += __shfl_down(c, off, WAVE);
        z += __shfl_down(z, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0 && c) atomicAdd(count, c);
    if (threadIdx.x % WAVE == 0 && z) atomicAdd(nnz, z);
}

// check_symmetry_assembled_matrix_3d, read-only: the number of coupling
// pairs that differ (an entry and its transposed partner: the partner
// row's entry, or zero where the partner has no stored row).  Zero means
// the symmetry fix would change nothing and the matrix is exactly
// symmetric.  The same pass counts the nonzero entries of the owned active
// rows into nnz.
__global__ void k_symmetry_count(Dims d, const int32_t* __restrict__ avert, const double* __restrict__ A,
                                 unsigned long long* __restrict__ count, unsigned long long* __restrict__ nnz) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    unsigned long long c = 0, z = 0;
    if (t < d.nrows) sym_row(d, avert, A, t, false, c, z);
    sym_flush(c, z, count, nnz);
}

// The same check over the rim rows only, every pair from their side, one
// thread per rim row.  For the element-once rows with no pass after them
// that changes couplings (no Dirichlet nodes, local basal friction), a pair
// of owned nodes is summed in both rows from the same stored point data,
// with equal products, in one global element order, so it is equal bit for
// bit; only a pair with a node outside the owned region (a halo vertex or a
// periodic image, whose elements are other copies) or a rim row zeroed by
// the no-ice rule can differ, and each of those has a rim row.
__global__ void k_symmetry_rim(Dims d, const int32_t* __restrict__ avert, const double* __restrict__ A,
                               unsigned long long* __restrict__ count, unsigned long long* __restrict__ nnz) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    unsigned long long c = 0, z = 0;
    if (t < rim_rows_per_level(d) * d.nz) sym_row(d, avert, A, rim_row_at(d, t), true, c, z);
    sym_flush(c, z, count, nnz);
}

// check_symmetry_assembled_matrix_3d restated per entry: an entry and its
// transpose partner are replaced by their average when they differ by less
// than eps11 times the diagonal of either row that checks the pair (the row
// visited first checks with its own diagonal, the partner row -- when it is
// an owned active row -- with its diagonal).  Each thread writes only its
// own row, into a copy (Aout), so the update is race-free.  count[0] counts
// pairs left asymmetric (the reference warns for those).
__global__ void k_symmetry(Dims d, const int32_t* __restrict__ avert, const double* __restrict__ A,
                           double* __restrict__ Aout, unsigned long long* __restrict__ nwarn) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo), k0 = (int)(kj / d.nyo);
    const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    double* o = Aout + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    for (int c = 0; c < NCOEF; ++c) o[c * WAVE] = a[c * WAVE];
    if (io >= d.nxo) return;
    const int i = d.ilo + io, j = d.jlo + jo;
    const long long pv = (long long)(i - 1) + (long long)d.nvx * (j - 1);
    if (!avert[pv]) return;
    const double eps11 = 1.0e-11;
    const double diagu = a[13 * WAVE], diagv = a[(3 * 27 + 13) * WAVE];
    unsigned long long warns = 0;
    for (int m = 0; m < 27; ++m) {
        const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
        if (k0 + kA < 0 || k0 + kA >= d.nz) continue;
        const int mm = 26 - m;
        int qi = i + iA, qj = j + jA;
        const bool owned_q = qi >= d.ilo && qi <= d.ihi && qj >= d.jlo && qj <= d.jhi &&
                             avert[(long long)(qi - 1) + (long long)d.nvx * (qj - 1)];
        double qdu = 0.0, qdv = 0.0;
        const double* q = nullptr;
        if (row_image(d, qi, qj)) {
            q = row_ptr(d, A, k0 + kA, qi, qj);
            qdu = q[13 * WAVE];
            qdv = q[(3 * 27 + 13) * WAVE];
        }
        // (uu), (uv<->vu), (vv), (vu<->uv) pairs: own block, partner block, own diag, partner diag
        const int ob[4] = {0, 1, 3, 2}, pb[4] = {0, 2, 3, 1};
        for (int s = 0; s < 4; ++s) {
            const double v1 = a[(ob[s] * 27 + m) * WAVE];
            const double v2_ = q ? q[(pb[s] * 27 + mm) * WAVE] : 0.0;
            if (v2_ == v1) continue;
            const double dmine = (s == 0 || s == 1) ? diagu : diagv;
            const double dpart = (s == 0 || s == 3) ? qdu : qdv;
            const double diff = fabs(v2_ - v1);
            if (diff < eps11 * fabs(dmine) || (owned_q && diff < eps11 * fabs(dpart)))
                o[(ob[s] * 27 + m) * WAVE] = 0.5 * (v1 + v2_);
            else
                ++warns;
        }
    }
    if (warns) atomicAdd(nwarn, warns);
}

// count_nonzeros_3d over owned active rows
__global__ void k_count_nnz(Dims d, const int32_t* __restrict__ avert, const double* __restrict__ A,
                            unsigned long long* __restrict__ count) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    unsigned long long c = 0;
    if (t < d.nrows) {
        const int io = (int)(t % d.nxp);
        const long long kj = t / d.nxp;
        const int jo = (int)(kj % d.nyo);
        if (io < d.nxo) {
            const int i = d.ilo + io, j = d.jlo + jo;
            if (avert[(long long)(i - 1) + (long long)d.nvx * (j - 1)]) {
                const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
                for (int q = 0; q < NCOEF; ++q) c += (a[q * WAVE] != 0.0) ? 1ull : 0ull;
            }
        }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) c += __shfl_down(c, off, WAVE);
    if (threadIdx.x % WAVE == 0 && c) atomicAdd(count, c);
}

// SIA/diag preconditioner coefficients of the owned rows of the row-blocked
// matrix; the ring is filled afterwards with the halo rule.
__global__ void k_precond_from_rows(Dims d, const double* __restrict__ A, double* __restrict__ M,
                                    double* __restrict__ Dg) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp);
    const long long kj = t / d.nxp;
    const int jo = (int)(kj % d.nyo), k0 = (int)(kj / d.nyo);
    if (io >= d.nxo) return;
    const int i = d.ilo + io, j = d.jlo + jo;
    const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    const long long c = vidx(d, k0, i, j);
    M[0 * d.vlen + c] = a[4 * WAVE];
    M[1 * d.vlen + c] = a[13 * WAVE];
    M[2 * d.vlen + c] = a[22 * WAVE];
    M[3 * d.vlen + c] = a[(81 + 4) * WAVE];
    M[4 * d.vlen + c] = a[(81 + 13) * WAVE];
    M[5 * d.vlen + c] = a[(81 + 22) * WAVE];
    Dg[c] = a[13 * WAVE];
    Dg[d.vlen + c] = a[(81 + 13) * WAVE];
}

// Raw Auu/Avv staging for the tridiagonal preconditioners (reference layout
// of the vertical-sum inputs) from the owned rows plus the halo rule.
__global__ void k_rows_to_raw(Dims d, int blk, const double* __restrict__ A, double* __restrict__ raw) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long total = 27LL * d.nz * d.nvx * d.nvy;
    if (t >= total) return;
    const int m = (int)(t % 27);
    long long rest = t / 27;
    const int k0 = (int)(rest % d.nz);
    rest /= d.nz;
    int i = 1 + (int)(rest % d.nvx), j = 1 + (int)(rest / d.nvx);
    double v = 0.0;
    if (row_image(d, i, j)) v = row_ptr(d, A, k0, i, j)[(blk * 27 + m) * WAVE];
    raw[t] = v;
}

}  // namespace cg
