// assembly2d.hip -- device assembly of the 2-D (depth-integrated) SSA system
// (glissade_velo_higher.F90): element stiffness with the vertically averaged
// flow factor and the SSA effective strain rate (assemble_stiffness_matrix_2d
// :4961-5381, get_basis_function_derivatives_2d :6319-6504,
// compute_effective_viscosity :7285-7470, compute_element_matrix :8001-8123,
// element_to_global_matrix_2d :8127-8196), local basal friction
// (basal_sliding_bc_2d :8281-8531), the level-summed load (:2045-2058),
// Dirichlet rows/columns (dirichlet_boundary_conditions_2d :9133-9301) and
// the symmetry report (check_symmetry_assembled_matrix_2d :10648-10850).
//
// The matrix of a 2-D system is stored as one-level rows of the 27-point
// row-blocked layout: the 9-point stencil entry m2 = 1 + (iA+1) + 3(jA+1)
// lives at m = 9 + m2 - 1 (the kA = 0 plane).  As for the 3-D assembly, one
// thread per owned vertex gathers the (up to) 4 elements around it in the
// reference's element order, so every coefficient is summed in the order of
// element_to_global_matrix_2d without atomics.
#include "cismgpu_internal.h"

namespace cg {

// 2-D reference element [n][p] (glissade_velo_higher.F90:528-570)
__constant__ double c_phi2[4][4], c_dxr2[4][4], c_dyr2[4][4];

// node n of the element in cell (i,j) is vertex (i + nd2_i(n), j + nd2_j(n))
// (ishift(3,n), jshift(3,n), :461-494)
__host__ __device__ constexpr int nd2_i(int n) { return (n == 0 || n == 3) ? -1 : 0; }
__host__ __device__ constexpr int nd2_j(int n) { return (n == 0 || n == 1) ? -1 : 0; }

struct Elem2 { double x[4], y[4], u[4], v[4], h[4]; };

// Layer-field dims of a 2-D system's inputs: nz sigma levels
struct Lev { int nz; };

__device__ __forceinline__ void load_elem2(const Dims& d, const AsmDev& in, const double* __restrict__ uv,
                                           const double* __restrict__ vv, int ci, int cj, Elem2& e) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const int i = ci + nd2_i(n), j = cj + nd2_j(n);
        const long long p = v2(d, i, j);
        e.x[n] = in.xV[p];
        e.y[n] = in.yV[p];
        e.h[n] = in.sthck[p];
        const long long c = vidx(d, 0, i, j);
        e.u[n] = uv[c];
        e.v[n] = vv[c];
    }
}

// get_basis_function_derivatives_2d at quadrature point p; false if singular
__device__ __forceinline__ bool basis_derivs_2d(const Elem2& e, int p, double* dx, double* dy, double& detJ) {
    double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        J11 = J11 + c_dxr2[n][p] * e.x[n];
        J12 = J12 + c_dxr2[n][p] * e.y[n];
        J21 = J21 + c_dyr2[n][p] * e.x[n];
        J22 = J22 + c_dyr2[n][p] * e.y[n];
    }
    detJ = J11 * J22 - J12 * J21;
    if (!(fabs(detJ) > 0.0)) return false;
    const double i11 = J22 / detJ, i12 = -J12 / detJ, i21 = -J21 / detJ, i22 = J11 / detJ;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        dx[n] = 0.0 + i11 * c_dxr2[n][p] + i12 * c_dyr2[n][p];
        dy[n] = 0.0 + i21 * c_dxr2[n][p] + i22 * c_dyr2[n][p];
    }
    return true;
}

// compute_effective_viscosity, SSA branch
__device__ __forceinline__ double efvs_ssa(const AsmDev& in, const Elem2& e, const double* dx, const double* dy,
                                           double flwafact) {
    const double p_effstr = (1.0 - in.n_glen) / in.n_glen;
    if (in.whichefvs == 0) return in.efvs_constant;
    if (in.whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        return flwafact * pow(effstrain, p_effstr);
    }
    double ux = 0, uy = 0, vx = 0, vy = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        ux = ux + dx[n] * e.u[n]; uy = uy + dy[n] * e.u[n];
        vx = vx + dx[n] * e.v[n]; vy = vy + dy[n] * e.v[n];
    }
    const double esq =
        in.effstrain_min * in.effstrain_min + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
    return flwafact * pow(esq, p_effstr / 2.0);
}

// glissade_vertical_average of flwafact(:, ci, cj) (glissade_grid_operators.F90:1882-1946)
__device__ __forceinline__ double flwafact_2d(const Dims& d, const AsmDev& in, Lev L, int ci, int cj) {
    const double* f = in.flwafact + (long long)(L.nz - 1) * c2(d, ci, cj);
    double s = 0.0;
    for (int k = 1; k <= L.nz - 1; ++k) s = s + f[k - 1] * (in.sigma[k] - in.sigma[k - 1]);
    return s;
}

// Row NR of the SSA element matrix of cell (ci,cj), summed over the 4
// quadrature points as compute_element_matrix accumulates K, then added
// into the vertex row accumulators acc[blk][m2].
// efva != nullptr (DIVA): the vertical-mean viscosity of each quadrature
// point comes from k_diva_cell (efva[cell*4 + p]); otherwise the SSA
// viscosity of the vertically averaged flow factor.
template <int NR>
__device__ __forceinline__ bool add_element_row_2d(const Dims& d, const AsmDev& in, Lev L,
                                                   const double* __restrict__ uv, const double* __restrict__ vv,
                                                   const double* __restrict__ efva, int ci, int cj,
                                                   double (&acc)[4][9]) {
    Elem2 e;
    load_elem2(d, in, uv, vv, ci, cj, e);
    const double fl = efva ? 0.0 : flwafact_2d(d, in, L, ci, cj);
    double K[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 4; ++c) K[b][c] = 0.0;
    bool ok = true;
    for (int p = 0; p < 4; ++p) {
        double dx[4], dy[4], detJ;
        if (!basis_derivs_2d(e, p, dx, dy, detJ)) { ok = false; break; }
        const double ef = efva ? efva[4 * c2(d, ci, cj) + p] : efvs_ssa(in, e, dx, dy, fl);
        double hq = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) hq = hq + c_phi2[n][p] * e.h[n];
        const double f = (hq * ef) * 1.0 * detJ / 1.0e9;
        const double ar = dx[NR], br = dy[NR];
#pragma unroll
        for (int nc = 0; nc < 4; ++nc) {
            K[0][nc] = K[0][nc] + f * (4.0 * ar * dx[nc] + br * dy[nc]);
            K[1][nc] = K[1][nc] + f * (2.0 * ar * dy[nc] + br * dx[nc]);
            K[2][nc] = K[2][nc] + f * (2.0 * br * dx[nc] + ar * dy[nc]);
            K[3][nc] = K[3][nc] + f * (4.0 * br * dy[nc] + ar * dx[nc]);
        }
    }
#pragma unroll
    for (int nc = 0; nc < 4; ++nc) {
        const int m2 = (nd2_i(nc) - nd2_i(NR) + 1) + 3 * (nd2_j(nc) - nd2_j(NR) + 1);
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b][m2] = acc[b][m2] + K[b][nc];
    }
    return ok;
}

// One thread per owned vertex; writes the whole one-level row (the kA = +-1
// planes are zero).  flag[0] = 1 on a singular Jacobian.
__global__ __launch_bounds__(128) void k_assemble_2d(Dims d, AsmDev in, Lev L, const double* __restrict__ uv,
                                                     const double* __restrict__ vv, const double* __restrict__ efva,
                                                     double* __restrict__ A, double* __restrict__ flag) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp), jo = (int)(t / d.nxp);
    double acc[4][9];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int m = 0; m < 9; ++m) acc[b][m] = 0.0;
    if (io < d.nxo) {
        const int i = d.ilo + io, j = d.jlo + jo;
        if (in.avert[v2(d, i, j)]) {
            bool ok = true;
            // cells in reference order (j-major, then i); the vertex is node NR of each
#define CELL2(DI, DJ, NR)                                                                 \
    {                                                                                     \
        const int ci = i + DI, cj = j + DJ;                                               \
        if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[c2(d, ci, cj)])                \
            ok = add_element_row_2d<NR>(d, in, L, uv, vv, efva, ci, cj, acc) && ok;       \
    }
            CELL2(0, 0, 2)
            CELL2(1, 0, 3)
            CELL2(0, 1, 1)
            CELL2(1, 1, 0)
#undef CELL2
            if (!ok) flag[0] = 1.0;
        }
    }
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int m = 0; m < 27; ++m) a[(b * 27 + m) * WAVE] = (m >= 9 && m < 18) ? acc[b][m - 9] : 0.0;
}

// efvs(k,i,j) = element mean over the 4 quadrature points for every layer k
// (:5310-5318); one thread per cell.
__global__ void k_efvs_2d(Dims d, AsmDev in, Lev L, const double* __restrict__ uv, const double* __restrict__ vv,
                          double* __restrict__ efvs) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nx * d.ny) return;
    const int ci = 1 + (int)(t % d.nx), cj = 1 + (int)(t / d.nx);
    double out = 0.0;
    if (cell_in_range(d, in.nhalo, ci, cj) && in.acell[t]) {
        Elem2 e;
        load_elem2(d, in, uv, vv, ci, cj, e);
        const double fl = flwafact_2d(d, in, L, ci, cj);
        double ef[4];
        for (int p = 0; p < 4; ++p) {
            double dx[4], dy[4], detJ;
            ef[p] = basis_derivs_2d(e, p, dx, dy, detJ) ? efvs_ssa(in, e, dx, dy, fl) : NAN;
        }
        // per layer: efvs(k) = ((0 + e1) + e2 + e3 + e4) / 4
        double s = 0.0;
        for (int p = 0; p < 4; ++p) s = s + ef[p];
        out = s / 4.0;
    }
    for (int k = 0; k < L.nz - 1; ++k) efvs[(long long)k + (long long)(L.nz - 1) * t] = out;
}

// basal_sliding_bc_2d, local option: diagonal += dx dy / vol0 * beta on
// active owned rows (no slope factor in 2-D).
__global__ void k_basal_local_2d(Dims d, AsmDev in, double* __restrict__ A) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nxo * d.nyo) return;
    const int io = (int)(t % d.nxo), jo = (int)(t / d.nxo);
    const int i = d.ilo + io, j = d.jlo + jo;
    const long long p = v2(d, i, j);
    if (!in.avert[p]) return;
    const long long r = (long long)jo * d.nxp + io;
    double* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
    const double add = (in.dx * in.dy / 1.0e9) * in.beta[p];
    a[(0 * 27 + 13) * WAVE] = a[(0 * 27 + 13) * WAVE] + add;
    a[(3 * 27 + 13) * WAVE] = a[(3 * 27 + 13) * WAVE] + add;
}

// loadu_2d(i,j) = sum over k of the 3-D load, in k order (:2045-2058); d3
// describes the nz-level work vectors, d the 2-D system.
__global__ void k_sum_levels(Dims d3, const double* __restrict__ lu, const double* __restrict__ lv, Dims d,
                             double* __restrict__ bu, double* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    double su = 0.0, sv = 0.0;
    for (int k = 0; k < d3.nz; ++k) {
        const long long c = vidx(d3, k, i, j);
        su = su + lu[c];
        sv = sv + lv[c];
    }
    bu[vidx(d, 0, i, j)] = su;
    bv[vidx(d, 0, i, j)] = sv;
}

// dirichlet_boundary_conditions_2d restated per owned row P, with the
// basal level of the 3-D masks.  Column entries of P that couple to a
// prescribed active vertex Q in the loop range move to the rhs in the
// reference's order: its offset loop (jA outer, iA inner; Q = P - (iA,jA)),
// then u before v.  P's own row: the diagonal of a u-prescribed row sets
// Auu, Auv and Avu to 1 (:9206-9210), a v-prescribed row then sets Auv and
// Avu to 0 and Avv to 1; off-diagonal entries of the prescribed component
// are zeroed.
__global__ void k_dirichlet_2d(Dims d, AsmDev in, Lev L, const double* __restrict__ uv,
                               const double* __restrict__ vv, double* __restrict__ A, double* __restrict__ bu,
                               double* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= d.nrows) return;
    const int io = (int)(t % d.nxp), jo = (int)(t / d.nxp);
    if (io >= d.nxo) return;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int nh = in.nhalo;
    auto mask = [&](const int32_t* m, int ii, int jj) { return m[(long long)(L.nz - 1) + (long long)L.nz * v2(d, ii, jj)]; };
    double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
    const long long pc = vidx(d, 0, i, j);
    const bool pact = in.avert[v2(d, i, j)] != 0;
    const int um = mask(in.umask, i, j), vm = mask(in.vmask, i, j);
    double bU = bu[pc], bV = bv[pc];
    for (int jA = -1; jA <= 1; ++jA)
        for (int iA = -1; iA <= 1; ++iA) {
            if (iA == 0 && jA == 0) continue;
            const int qi = i - iA, qj = j - jA;
            if (!(qi >= nh && qi <= d.nx - nh + 1 && qj >= nh && qj <= d.ny - nh + 1)) continue;
            if (!in.avert[v2(d, qi, qj)]) continue;
            const int m = 9 + (-iA + 1) + 3 * (-jA + 1);     // entry of row P coupling to Q
            const long long qc = vidx(d, 0, qi, qj);
            if (mask(in.umask, qi, qj) == 1) {
                if (um != 1) { bU = bU - a[(0 * 27 + m) * WAVE] * uv[qc]; a[(0 * 27 + m) * WAVE] = 0.0; }
                if (vm != 1) { bV = bV - a[(2 * 27 + m) * WAVE] * uv[qc]; a[(2 * 27 + m) * WAVE] = 0.0; }
            }
            if (mask(in.vmask, qi, qj) == 1) {
                if (um != 1) { bU = bU - a[(1 * 27 + m) * WAVE] * vv[qc]; a[(1 * 27 + m) * WAVE] = 0.0; }
                if (vm != 1) { bV = bV - a[(3 * 27 + m) * WAVE] * vv[qc]; a[(3 * 27 + m) * WAVE] = 0.0; }
            }
        }
    if (pact) {   // owned vertices lie in the loop range nhalo..n-nhalo+1
        if (um == 1) {
            bU = uv[pc];
            for (int m = 9; m < 18; ++m) {
                if (m == 13) {
                    a[(0 * 27 + m) * WAVE] = 1.0; a[(1 * 27 + m) * WAVE] = 1.0; a[(2 * 27 + m) * WAVE] = 1.0;
                } else {
                    a[(0 * 27 + m) * WAVE] = 0.0; a[(1 * 27 + m) * WAVE] = 0.0;
                }
            }
        }
        if (vm == 1) {
            bV = vv[pc];
            for (int m = 9; m < 18; ++m) {
                if (m == 13) {
                    a[(1 * 27 + m) * WAVE] = 0.0; a[(2 * 27 + m) * WAVE] = 0.0; a[(3 * 27 + m) * WAVE] = 1.0;
                } else {
                    a[(2 * 27 + m) * WAVE] = 0.0; a[(3 * 27 + m) * WAVE] = 0.0;
                }
            }
        }
    }
    bu[pc] = bU;
    bv[pc] = bV;
}

// check_symmetry_assembled_matrix_2d: the 2-D check only reports; count[0]
// gets the pairs that differ by eps11 |diag| or more.  The partner row is
// read after the halo rule (row_image), as the reference reads it after the
// matrix halo update.
__global__ void k_symmetry_report_2d(Dims d, const int32_t* __restrict__ avert, const double* __restrict__ A,
                                     unsigned long long* __restrict__ count) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    unsigned long long warns = 0;
    if (t < d.nrows) {
        const int io = (int)(t % d.nxp), jo = (int)(t / d.nxp);
        if (io < d.nxo) {
            const int i = d.ilo + io, j = d.jlo + jo;
            if (avert[(long long)(i - 1) + (long long)d.nvx * (j - 1)]) {
                const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
                const double diagu = a[13 * WAVE], diagv = a[(3 * 27 + 13) * WAVE];
                for (int m2 = 0; m2 < 9; ++m2) {
                    const int iA = m2 % 3 - 1, jA = m2 / 3 - 1;
                    const int m = 9 + m2, mm = 9 + (8 - m2);
                    int qi = i + iA, qj = j + jA;
                    const double* q = row_image(d, qi, qj) ? row_ptr(d, A, 0, qi, qj) : nullptr;
                    const int ob[4] = {0, 1, 3, 2}, pb[4] = {0, 2, 3, 1};
                    for (int s = 0; s < 4; ++s) {
                        const double v1 = a[(ob[s] * 27 + m) * WAVE];
                        const double v2_ = q ? q[(pb[s] * 27 + mm) * WAVE] : 0.0;
                        const double dg = s < 2 ? diagu : diagv;
                        if (v2_ != v1 && !(fabs(v2_ - v1) < 1.0e-11 * fabs(dg))) ++warns;
                    }
                }
            }
        }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) warns += __shfl_down(warns, off, WAVE);
    if (threadIdx.x % WAVE == 0 && warns) atomicAdd(count, warns);
}

// ---------------------------------------------------------------------------
// DIVA (Goldberg 2011) for the 2-D system, slope correction on
// (glissade_velo_higher.F90:2078-2140, 2469-2560, 5649-5900, 7730-7997,
// 8535-8811).  State kept on the device between Picard iterations: the
// layer viscosities efvs_qp_3d (nz-1, 4, nx, ny) and the basal traction.
// ---------------------------------------------------------------------------
struct DivaDev {
    const double *sfx, *sfy;           // (nx, ny) slope factors at cells
    const double *ssfx, *ssfy;         // (nvx, nvy) slope factors at vertices
    const double *sthx, *sthy;         // (nvx, nvy) slope angles at vertices
    const double *tbx, *tby;           // (nvx, nvy) basal slope angles
    const int32_t* ipl;                // (nx, ny) ice or land cells
    double* efvs_qp;                   // (nz-1, 4, nx, ny) state
    double *btx, *bty;                 // (nvx, nvy) basal traction (state)
    double* omega_k;                   // (nz, nx, ny)
    double* omega;                     // (nx, ny)
    double* efva;                      // (4, nx, ny) vertical-mean viscosity per quadrature point
    double* somega;                    // (nvx, nvy) staggered omega
    double *bex, *bey;                 // (nvx, nvy) beta_eff
    int no_slip;
};

// compute_effective_viscosity_diva, default (non-cubic) branch, for point p;
// ef (nz-1 values, stride 1) updated in place
__device__ __forceinline__ void efvs_diva(const AsmDev& in, Lev L, const Elem2& e, const double* bx,
                                          const double* by, int p, const double* dx, const double* dy, double sfx,
                                          double sfy, const double* fl, double* ef) {
#pragma clang fp contract(off)
    const double p_effstr = (1.0 - in.n_glen) / in.n_glen;
    if (in.whichefvs == 0) { for (int k = 0; k < L.nz - 1; ++k) ef[k] = in.efvs_constant; return; }
    if (in.whichefvs == 1) {
        const double scyr = 31536000.0;
        const double effstrain = (500.0 / scyr) / 200.0e3 * scyr;
        for (int k = 0; k < L.nz - 1; ++k) ef[k] = fl[k] * pow(effstrain, p_effstr);
        return;
    }
    double ux = 0, uy = 0, vx = 0, vy = 0, taux = 0, tauy = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        ux = ux + dx[n] * e.u[n]; uy = uy + dy[n] * e.u[n];
        vx = vx + dx[n] * e.v[n]; vy = vy + dy[n] * e.v[n];
        taux = taux + c_phi2[n][p] * bx[n];
        tauy = tauy + c_phi2[n][p] * by[n];
    }
    for (int k = 0; k < L.nz - 1; ++k) {
        // sigma of layer k (the midpoint of levels k, k+1)
        const double ss = (in.sigma[k] + in.sigma[k + 1]) / 2.0;
        double efk = ef[k];
        if (efk == 0.0) efk = fl[k] * pow(in.effstrain_min, p_effstr);
        const double du_dz = taux * ss / (efk * sfx);
        const double dv_dz = tauy * ss / (efk * sfy);
        const double esq = in.effstrain_min * in.effstrain_min + ux * ux + vy * vy + ux * vy +
                           0.25 * ((vx + uy) * (vx + uy)) + 0.25 * (du_dz * du_dz) * sfx + 0.25 * (dv_dz * dv_dz) * sfy;
        ef[k] = fl[k] * pow(esq, p_effstr / 2.0);
    }
}

// One thread per cell: layer viscosities (state), their vertical means per
// quadrature point, omega_k / omega (compute_integrals_diva) and efvs.
__global__ void k_diva_cell(Dims d, AsmDev in, Lev L, DivaDev dv, const double* __restrict__ uv,
                            const double* __restrict__ vv, double* __restrict__ efvs, double* __restrict__ flag) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nx * d.ny) return;
    const int ci = 1 + (int)(t % d.nx), cj = 1 + (int)(t / d.nx);
    const int nl = L.nz - 1;
    double* st = dv.efvs_qp + (long long)nl * 4 * t;
    double* ok = dv.omega_k + (long long)L.nz * t;
    if (!(cell_in_range(d, in.nhalo, ci, cj) && in.acell[t])) {
        for (int p = 0; p < 4; ++p) dv.efva[4 * t + p] = 0.0;
        for (int k = 0; k < L.nz; ++k) ok[k] = 0.0;
        dv.omega[t] = 0.0;
        for (int k = 0; k < nl; ++k) efvs[(long long)k + (long long)nl * t] = 0.0;
        return;
    }
    Elem2 e;
    load_elem2(d, in, uv, vv, ci, cj, e);
    double bx[4], by[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const long long q = v2(d, ci + nd2_i(n), cj + nd2_j(n));
        bx[n] = dv.btx[q];
        by[n] = dv.bty[q];
    }
    const double* fl = in.flwafact + (long long)nl * t;
    for (int p = 0; p < 4; ++p) {
        double dx[4], dy[4], detJ;
        if (!basis_derivs_2d(e, p, dx, dy, detJ)) { flag[0] = 1.0; return; }
        efvs_diva(in, L, e, bx, by, p, dx, dy, dv.sfx[t], dv.sfy[t], fl, st + p * nl);
        double va = 0.0;
        for (int k = 0; k < nl; ++k) va = va + st[p * nl + k] * (in.sigma[k + 1] - in.sigma[k]);
        dv.efva[4 * t + p] = va;
    }
    // compute_integrals_diva with the cell thickness
    const double H = in.thck_cell[t];
    double okp[4];
    for (int p = 0; p < 4; ++p) okp[p] = 0.0;
    ok[L.nz - 1] = 0.0;
    for (int k = L.nz - 2; k >= 0; --k) {
        const double depth = 0.5 * (in.sigma[k] + in.sigma[k + 1]);
        const double dz = (in.sigma[k + 1] - in.sigma[k]) * H;
        for (int p = 0; p < 4; ++p) okp[p] = okp[p] + depth / st[p * nl + k] * dz;
        ok[k] = (((okp[0] + okp[1]) + okp[2]) + okp[3]) / 4.0;
    }
    double om = 0.0;
    for (int k = 0; k < L.nz - 1; ++k) om = om + 0.5 * (ok[k] + ok[k + 1]) * (in.sigma[k + 1] - in.sigma[k]);
    dv.omega[t] = om;
    for (int k = 0; k < nl; ++k) {
        double sm = 0.0;
        for (int p = 0; p < 4; ++p) sm = sm + st[p * nl + k];
        efvs[(long long)k + (long long)nl * t] = sm / 4.0;
    }
}

// parallel_halo_real8_2d for a cell field, one task, 2 halo cells: periodic
// wrap; outflow zeroes the halo; no_ice zeroes 3 rows on each side.  Run as
// two launches (E/W, then N/S) for the periodic rules, one otherwise.
__global__ void k_cell_halo(Dims d, int pass, double* __restrict__ a) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    if (d.bc == 0 || d.bc == 2) {
        if (pass == 0) {
            if (t >= ny - 4) return;
            const int j = 3 + (int)t;
            auto c = [&](int i) { return (long long)(i - 1) + (long long)nx * (j - 1); };
            const double a1 = a[c(nx - 3)], a2 = a[c(nx - 2)], a3 = a[c(3)], a4 = a[c(4)];
            a[c(1)] = a1; a[c(2)] = a2; a[c(nx - 1)] = a3; a[c(nx)] = a4;
        } else {
            if (t >= nx) return;
            const int i = 1 + (int)t;
            auto c = [&](int j) { return (long long)(i - 1) + (long long)nx * (j - 1); };
            const double a1 = a[c(ny - 3)], a2 = a[c(ny - 2)], a3 = a[c(3)], a4 = a[c(4)];
            a[c(1)] = a1; a[c(2)] = a2; a[c(ny - 1)] = a3; a[c(ny)] = a4;
        }
    } else {
        if (pass != 0 || t >= (long long)nx * ny) return;
        const int i = 1 + (int)(t % nx), j = 1 + (int)(t / nx);
        const int w = d.bc == 1 ? 2 : 3;
        if (i <= w || i >= nx - w + 1 || j <= w || j >= ny - w + 1) a[t] = 0.0;
    }
}

// masked stagger (glissade_stagger, stagger_margin 1) of a cell field
__device__ __forceinline__ double stagger_masked_at(const Dims& d, const int32_t* __restrict__ m,
                                                    const double* __restrict__ a, int i, int j) {
#pragma clang fp contract(off)
    const long long c00 = c2(d, i, j), c10 = c2(d, i + 1, j), c01 = c2(d, i, j + 1), c11 = c2(d, i + 1, j + 1);
    const double sv = (double)m[c01] * a[c01] + (double)m[c11] * a[c11] + (double)m[c00] * a[c00] + (double)m[c10] * a[c10];
    const double sm = (double)(m[c01] + m[c11] + m[c00] + m[c10]);
    return sm > 0.0 ? sv / sm : 0.0;
}

// stag_omega and beta_eff at every vertex
__global__ void k_diva_beta(Dims d, AsmDev in, DivaDev dv) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const double so = stagger_masked_at(d, dv.ipl, dv.omega, i, j);
    dv.somega[t] = so;
    double bex = 0.0, bey = 0.0;
    if (dv.no_slip) {
        if (so > 0.0) {
            bex = dv.ssfx[t] / (so * cos(dv.sthx[t]));
            bey = dv.ssfy[t] / (so * cos(dv.sthy[t]));
        }
    } else {
        const double b = in.beta[t];
        bex = (dv.ssfx[t] * b) / (dv.ssfx[t] + b * so * cos(dv.sthx[t]));
        bey = (dv.ssfy[t] * b) / (dv.ssfy[t] + b * so * cos(dv.sthy[t]));
    }
    dv.bex[t] = bex;
    dv.bey[t] = bey;
}

// basal_sliding_bc_2d_diva, local option, on active owned rows
__global__ void k_basal_diva_2d(Dims d, AsmDev in, DivaDev dv, double* __restrict__ A) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nxo * d.nyo) return;
    const int io = (int)(t % d.nxo), jo = (int)(t / d.nxo);
    const int i = d.ilo + io, j = d.jlo + jo;
    const long long p = v2(d, i, j);
    if (!in.avert[p]) return;
    const long long r = (long long)jo * d.nxp + io;
    double* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
    const double f = in.dx * in.dy / 1.0e9;
    a[(0 * 27 + 13) * WAVE] = a[(0 * 27 + 13) * WAVE] + f * dv.bex[p] / cos(dv.tbx[p]);
    a[(3 * 27 + 13) * WAVE] = a[(3 * 27 + 13) * WAVE] + f * dv.bey[p] / cos(dv.tby[p]);
}

// compute_3d_velocity_diva: basal traction at every vertex; the 3-D velocity
// (reference layout (nz, nvx, nvy), optional) on active vertices
// ubo/vbo (optional, (nvx, nvy)): the basal level at every vertex, zero at
// inactive ones (the basal velocity the next glissade_calcbeta reads)
__global__ void k_diva_velocity(Dims d, AsmDev in, Lev L, DivaDev dv, const double* __restrict__ uv,
                                const double* __restrict__ vv, double* __restrict__ u3, double* __restrict__ v3,
                                double* __restrict__ ubo, double* __restrict__ vbo) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const long long c = vidx(d, 0, i, j);
    const double u = uv[c], v = vv[c];
    const double btx = dv.bex[t] * u, bty = dv.bey[t] * v;
    dv.btx[t] = btx;
    dv.bty[t] = bty;
    double ub = 0.0, vb = 0.0;
    if (in.avert[t]) {
        const double b = in.beta[t];
        if (b > 0.0) {
            ub = u * dv.bex[t] / b;
            vb = v * dv.bey[t] / b;
        } else {
            ub = u;
            vb = v;
        }
    }
    if (ubo) {
        ubo[t] = ub;
        vbo[t] = vb;
    }
    if (!u3 || !in.avert[t]) return;
    const int nz = L.nz;
    u3[(long long)(nz - 1) + (long long)nz * t] = ub;
    v3[(long long)(nz - 1) + (long long)nz * t] = vb;
    const double cx = cos(dv.sthx[t]) / dv.ssfx[t], cy = cos(dv.sthy[t]) / dv.ssfy[t];
    for (int k = 0; k < nz - 1; ++k) {
        // masked stagger of omega_k(k, :, :)
        const long long c00 = c2(d, i, j), c10 = c2(d, i + 1, j), c01 = c2(d, i, j + 1), c11 = c2(d, i + 1, j + 1);
        const int32_t* m = dv.ipl;
        const double* o = dv.omega_k;
        const double sv = (double)m[c01] * o[(long long)k + nz * c01] + (double)m[c11] * o[(long long)k + nz * c11] +
                          (double)m[c00] * o[(long long)k + nz * c00] + (double)m[c10] * o[(long long)k + nz * c10];
        const double sm = (double)(m[c01] + m[c11] + m[c00] + m[c10]);
        const double sok = sm > 0.0 ? sv / sm : 0.0;
        u3[(long long)k + (long long)nz * t] = ub + btx * sok * cx;
        v3[(long long)k + (long long)nz * t] = vb + bty * sok * cy;
    }
}

// ------------------------------------------------------------------ L1L2
// compute_effective_viscosity_L1L2, layer k at one quadrature point of
// nonlinear viscosity: tau_perp = rhoi g (h sigma(k+1)) |grad s|;
// tau_parallel solves x^3 + tau_perp^2 x - es/A = 0 (Cardano, with the
// first-order form of the second cube root when a^3/27 <= 1e-6 b^2/4).
// n is 1 or 3 (checked on the host).
__device__ __forceinline__ double efvs_l1l2_layer(const AsmDev& in, double es, double thck, double grads,
                                                  double sig1, double A) {
#pragma clang fp contract(off)
    const double depth = thck * sig1;
    const double tau_perp = 910.0 * 9.81 * depth * grads;
    const double a = tau_perp * tau_perp;
    const double b = -es / A;
    const double c = sqrt(b * b / 4.0 + a * a * a / 27.0);
    const double rootA = pow(-b / 2.0 + c, 1.0 / 3.0);
    const double rootB = (a * a * a / 27.0 > 1.0e-6 * (b * b / 4.0)) ? -pow(b / 2.0 + c, 1.0 / 3.0)
                                                                     : -a / (3.0 * pow(fabs(b), 1.0 / 3.0));
    const double tp = rootA + rootB;
    if (fabs(in.n_glen - 1.0) < 1.0e-10) return 1.0 / (2.0 * A);
    return 1.0 / (2.0 * A * (tp * tp + tau_perp * tau_perp));
}

// One thread per cell: the L1L2 layer viscosities at the four quadrature
// points, their vertical means (efva) and the element means per layer (efvs).
__global__ void k_l1l2_cell(Dims d, AsmDev in, Lev L, const double* __restrict__ uv, const double* __restrict__ vv,
                            double* __restrict__ efva, double* __restrict__ efvs, double* __restrict__ flag) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nx * d.ny) return;
    const int ci = 1 + (int)(t % d.nx), cj = 1 + (int)(t / d.nx);
    const int nl = L.nz - 1;
    double* out = efvs + (long long)nl * t;
    if (!(cell_in_range(d, in.nhalo, ci, cj) && in.acell[t])) {
        for (int p = 0; p < 4; ++p) efva[4 * t + p] = 0.0;
        for (int k = 0; k < nl; ++k) out[k] = 0.0;
        return;
    }
    Elem2 e;
    load_elem2(d, in, uv, vv, ci, cj, e);
    double sx[4], sy[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const long long q = v2(d, ci + nd2_i(n), cj + nd2_j(n));
        sx[n] = in.dsdx[q];
        sy[n] = in.dsdy[q];
    }
    const double* fa = in.flwa + (long long)nl * t;
    const double* fl = in.flwafact + (long long)nl * t;
    const double p_effstr = (1.0 - in.n_glen) / in.n_glen;
    for (int p = 0; p < 4; ++p) {
        double dx[4], dy[4], detJ;
        if (!basis_derivs_2d(e, p, dx, dy, detJ)) { flag[0] = 1.0; return; }
        double es = 0.0, grads = 0.0, thck = 0.0;
        if (in.whichefvs == 2) {
            double ux = 0, uy = 0, vx = 0, vy = 0, dsx = 0, dsy = 0;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                ux = ux + dx[n] * e.u[n]; uy = uy + dy[n] * e.u[n];
                vx = vx + dx[n] * e.v[n]; vy = vy + dy[n] * e.v[n];
                dsx = dsx + c_phi2[n][p] * sx[n];
                dsy = dsy + c_phi2[n][p] * sy[n];
                thck = thck + c_phi2[n][p] * e.h[n];
            }
            const double esq = in.effstrain_min * in.effstrain_min + ux * ux + vy * vy + ux * vy +
                               0.25 * ((vx + uy) * (vx + uy));
            es = sqrt(esq);
            grads = sqrt(dsx * dsx + dsy * dsy);
        }
        double va = 0.0;
        for (int k = 0; k < nl; ++k) {
            double ek;
            if (in.whichefvs == 0) {
                ek = in.efvs_constant;
            } else if (in.whichefvs == 1) {
                const double scyr = 31536000.0;
                ek = fl[k] * pow((500.0 / scyr) / 200.0e3 * scyr, p_effstr);
            } else {
                ek = efvs_l1l2_layer(in, es, thck, grads, in.sigma[k + 1], fa[k]);
            }
            va = va + ek * (in.sigma[k + 1] - in.sigma[k]);
            out[k] = p == 0 ? ek : out[k] + ek;
        }
        efva[4 * t + p] = va;
    }
    for (int k = 0; k < nl; ++k) out[k] = out[k] / 4.0;
}

// parallel_halo of nlay cell fields stored one after another (blockIdx.y
// selects the field), with the rules of k_cell_halo.
__global__ void k_cell_halo_layers(Dims d, int pass, double* __restrict__ a) {
    const long long nc = (long long)d.nx * d.ny;
    double* f = a + (long long)blockIdx.y * nc;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    if (d.bc == 0 || d.bc == 2) {
        if (pass == 0) {
            if (t >= ny - 4) return;
            const int j = 3 + (int)t;
            auto c = [&](int i) { return (long long)(i - 1) + (long long)nx * (j - 1); };
            const double a1 = f[c(nx - 3)], a2 = f[c(nx - 2)], a3 = f[c(3)], a4 = f[c(4)];
            f[c(1)] = a1; f[c(2)] = a2; f[c(nx - 1)] = a3; f[c(nx)] = a4;
        } else {
            if (t >= nx) return;
            const int i = 1 + (int)t;
            auto c = [&](int j) { return (long long)(i - 1) + (long long)nx * (j - 1); };
            const double a1 = f[c(ny - 3)], a2 = f[c(ny - 2)], a3 = f[c(3)], a4 = f[c(4)];
            f[c(1)] = a1; f[c(2)] = a2; f[c(ny - 1)] = a3; f[c(ny)] = a4;
        }
    } else {
        if (pass != 0 || t >= nc) return;
        const int i = 1 + (int)(t % nx), j = 1 + (int)(t / nx);
        const int w = d.bc == 1 ? 2 : 3;
        if (i <= w || i >= nx - w + 1 || j <= w || j >= ny - w + 1) f[t] = 0.0;
    }
}

// compute_3d_velocity_L1L2, cell part: membrane strain rate at the cell
// centre, tau_parallel(k) = 2 efvs(k) eps_parallel and the efvs integral
// from the layer midpoint to the surface times the strain combinations
// (work1..3).  Layer-major outputs: f[(q*nl + k)*nc + c], q = 0 tau_parallel,
// 1..3 work1..3.
__global__ void k_l1l2_columns(Dims d, AsmDev in, Lev L, const double* __restrict__ uv, const double* __restrict__ vv,
                               const double* __restrict__ efvs, double* __restrict__ f, double* __restrict__ flag) {
#pragma clang fp contract(off)
    const long long nc = (long long)d.nx * d.ny;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= nc) return;
    const int ci = 1 + (int)(t % d.nx), cj = 1 + (int)(t / d.nx);
    const int nl = L.nz - 1;
    double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0, eps = 0.0;
    const bool act = cell_in_range(d, in.nhalo, ci, cj) && in.acell[t];
    const double* ef = efvs + (long long)nl * t;
    if (act) {
        Elem2 e;
        load_elem2(d, in, uv, vv, ci, cj, e);
        const double cdx[4] = {-0.25, 0.25, 0.25, -0.25}, cdy[4] = {-0.25, -0.25, 0.25, 0.25};
        double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            J11 = J11 + cdx[n] * e.x[n]; J12 = J12 + cdx[n] * e.y[n];
            J21 = J21 + cdy[n] * e.x[n]; J22 = J22 + cdy[n] * e.y[n];
        }
        const double det = J11 * J22 - J12 * J21;
        if (!(fabs(det) > 0.0)) { flag[0] = 1.0; return; }
        const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const double ddx = 0.0 + i11 * cdx[n] + i12 * cdy[n];
            const double ddy = 0.0 + i21 * cdx[n] + i22 * cdy[n];
            ux = ux + ddx * e.u[n]; uy = uy + ddy * e.u[n];
            vx = vx + ddx * e.v[n]; vy = vy + ddy * e.v[n];
        }
        eps = sqrt(ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
    }
    const double H = in.thck_cell[t];
    double ei = 0.0;
    for (int k = 0; k < nl; ++k) {
        double tp = 0.0;
        if (act) {
            const double ssk = (in.sigma[k] + in.sigma[k + 1]) / 2.0;
            tp = 2.0 * ef[k] * eps;
            if (k == 0) {
                ei = ef[0] * ssk * H;
            } else {
                const double ssm = (in.sigma[k - 1] + in.sigma[k]) / 2.0;
                ei = ei + ef[k - 1] * (in.sigma[k] - ssm) * H + ef[k] * (ssk - in.sigma[k]) * H;
            }
        }
        f[(0LL * nl + k) * nc + t] = tp;
        f[(1LL * nl + k) * nc + t] = ei * (2.0 * ux + vy);
        f[(2LL * nl + k) * nc + t] = ei * (uy + vx);
        f[(3LL * nl + k) * nc + t] = ei * (2.0 * vy + ux);
    }
}

// glissade_gradient, gradient_margin 1, of cell field f at vertex (i, j)
__device__ __forceinline__ void gradient_m1_at(const Dims& d, const int32_t* __restrict__ ice,
                                               const double* __restrict__ f, double dx, double dy, int i, int j,
                                               double& gx, double& gy) {
#pragma clang fp contract(off)
    const long long c00 = c2(d, i, j), c10 = c2(d, i + 1, j), c01 = c2(d, i, j + 1), c11 = c2(d, i + 1, j + 1);
    const bool es = ice[c00] == 1 && ice[c10] == 1, en = ice[c01] == 1 && ice[c11] == 1;
    const bool ew = ice[c00] == 1 && ice[c01] == 1, ee = ice[c10] == 1 && ice[c11] == 1;
    const double n_ = (f[c11] - f[c01]) / dx, s_ = (f[c10] - f[c00]) / dx;
    gx = (es && en) ? (n_ + s_) / 2.0 : (es ? s_ : (en ? n_ : 0.0));
    const double e_ = (f[c11] - f[c10]) / dy, w_ = (f[c01] - f[c00]) / dy;
    gy = (ew && ee) ? (e_ + w_) / 2.0 : (ew ? w_ : (ee ? e_ : 0.0));
}

// compute_3d_velocity_L1L2, vertex part: from the bed up at owned active
// vertices u(k) = u(k+1) + fact tau_xz(k) (u(nz) where the basal Dirichlet
// mask is set); the basal level is x at every vertex; output (nz, nvx, nvy).
__global__ void k_l1l2_velocity(Dims d, AsmDev in, Lev L, const double* __restrict__ uv,
                                const double* __restrict__ vv, const int32_t* __restrict__ ice,
                                const double* __restrict__ efvs, const double* __restrict__ f,
                                double* __restrict__ u3, double* __restrict__ v3) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const int nz = L.nz, nl = nz - 1;
    const long long nc = (long long)d.nx * d.ny;
    const long long b = vidx(d, 0, i, j);
    const double ub = uv[b], vb = vv[b];
    u3[(long long)(nz - 1) + (long long)nz * t] = ub;
    v3[(long long)(nz - 1) + (long long)nz * t] = vb;
    const bool own = i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi && in.avert[t];
    if (!own) {
        for (int k = 0; k < nl; ++k) {
            u3[(long long)k + (long long)nz * t] = 0.0;
            v3[(long long)k + (long long)nz * t] = 0.0;
        }
        return;
    }
    const bool um = in.umask[(long long)(nz - 1) + (long long)nz * t] == 1;
    const bool vm = in.vmask[(long long)(nz - 1) + (long long)nz * t] == 1;
    const long long c00 = c2(d, i, j), c10 = c2(d, i + 1, j), c01 = c2(d, i, j + 1), c11 = c2(d, i + 1, j + 1);
    const double sm = (double)(ice[c01] + ice[c11] + ice[c00] + ice[c10]);
    const double H = in.sthck[t];
    double uk = ub, vk = vb;
    for (int k = nl - 1; k >= 0; --k) {
        double sq = stagger_masked_at(d, ice, f + (long long)k * nc, i, j);
        sq = sq * sq;
        const double* fa = in.flwa;
        const double sv = (double)ice[c01] * fa[(long long)k + nl * c01] + (double)ice[c11] * fa[(long long)k + nl * c11] +
                          (double)ice[c00] * fa[(long long)k + nl * c00] + (double)ice[c10] * fa[(long long)k + nl * c10];
        const double sf = sm > 0.0 ? sv / sm : 0.0;
        double g1x, g1y, g2x, g2y, g3x, g3y;
        gradient_m1_at(d, ice, f + (1LL * nl + k) * nc, in.dx, in.dy, i, j, g1x, g1y);
        gradient_m1_at(d, ice, f + (2LL * nl + k) * nc, in.dx, in.dy, i, j, g2x, g2y);
        gradient_m1_at(d, ice, f + (3LL * nl + k) * nc, in.dx, in.dy, i, j, g3x, g3y);
        const double ssk = (in.sigma[k] + in.sigma[k + 1]) / 2.0;
        const double depth = ssk * H;
        const double txs = -(910.0 * 9.81 * depth * in.dsdx[t]);
        const double tys = -(910.0 * 9.81 * depth * in.dsdy[t]);
        const double txz = txs + 2.0 * g1x + g2y;
        const double tyz = tys + g2x + 2.0 * g3y;
        const double te = sq + txz * txz + tyz * tyz;
        double fact;
        if (in.whichefvs == 2) {
            fact = 2.0 * sf * pow(te, (in.n_glen - 1.0) / 2.0) * (in.sigma[k + 1] - in.sigma[k]) * H;
        } else {
            const double ev = efvs[(long long)k + (long long)nl * c2(d, i, j)];
            fact = ev > 0.0 ? (in.sigma[k + 1] - in.sigma[k]) * H / ev : 0.0;
        }
        uk = um ? ub : uk + fact * txz;
        vk = vm ? vb : vk + fact * tyz;
        u3[(long long)k + (long long)nz * t] = uk;
        v3[(long long)k + (long long)nz * t] = vk;
    }
}

// basal_sliding_bc_3d (glissade_velo_higher.F90:8955-9111) and
// basal_sliding_bc_2d (:8386-8527), standard option, gathered per owned row
// of level `level`: the cells (i, j), (i+1, j), (i, j+1), (i+1, j+1) in the
// reference's order, points p in order, K(r,c) = beta_qp w detJ / vol0
// phi_r phi_c added to the kA = 0 coupling of column node c.  curved: detJ
// is the area factor of the face (x, y, staggered lower surface)
// (get_area_scale_factor_curved_2d, :6736-6827), else the flat Jacobian.
// bx weights Auu, by Avv (DIVA: beta_eff_x / beta_eff_y with the curved
// Jacobian, basal_sliding_bc_2d_diva :8535-8811; otherwise both beta).
__global__ void k_basal_std(Dims d, AsmDev in, int level, int curved, const double* __restrict__ bx,
                            const double* __restrict__ by, double* __restrict__ A, double* __restrict__ flag) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nxo * d.nyo) return;
    const int io = (int)(t % d.nxo), jo = (int)(t / d.nxo);
    const int i = d.ilo + io, j = d.jlo + jo;
    const long long r = ((long long)level * d.nyo + jo) * d.nxp + io;
    double* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
    const double vol0 = 1.0e9;
    for (int cc = 0; cc < 4; ++cc) {
        const int ci = i + (cc & 1), cj = j + (cc >> 1);
        if (ci > d.nx || cj > d.ny || !cell_in_range(d, in.nhalo, ci, cj) || !in.acell[c2(d, ci, cj)]) continue;
        int nr = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n)
            if (nd2_i(n) == i - ci && nd2_j(n) == j - cj) nr = n;
        Elem2 e;
        double z[4], b[4], c[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const long long q = v2(d, ci + nd2_i(n), cj + nd2_j(n));
            e.x[n] = in.xV[q];
            e.y[n] = in.yV[q];
            z[n] = curved ? in.slsrf[q] : 0.0;
            b[n] = bx[q];
            c[n] = by[q];
        }
        for (int p = 0; p < 4; ++p) {
            double detJ;
            if (curved) {
                double J11 = 0, J12 = 0, J13 = 0, J21 = 0, J22 = 0, J23 = 0;
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    J11 = J11 + c_dxr2[n][p] * e.x[n]; J12 = J12 + c_dxr2[n][p] * e.y[n]; J13 = J13 + c_dxr2[n][p] * z[n];
                    J21 = J21 + c_dyr2[n][p] * e.x[n]; J22 = J22 + c_dyr2[n][p] * e.y[n]; J23 = J23 + c_dyr2[n][p] * z[n];
                }
                const double t1 = (J11 * J22 - J12 * J21);
                const double t2 = -(J11 * J23 - J13 * J21);
                const double t3 = (J12 * J23 - J13 * J22);
                detJ = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
            } else {
                double dxn[4], dyn[4];
                if (!basis_derivs_2d(e, p, dxn, dyn, detJ)) { flag[0] = 1.0; return; }
            }
            double bq = 0.0, cq = 0.0;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                bq = bq + c_phi2[n][p] * b[n];
                cq = cq + c_phi2[n][p] * c[n];
            }
#pragma unroll
            for (int nc = 0; nc < 4; ++nc) {
                const double Ku = 0.0 + bq * 1.0 * detJ / vol0 * c_phi2[nr][p] * c_phi2[nc][p];
                const double Kv = 0.0 + cq * 1.0 * detJ / vol0 * c_phi2[nr][p] * c_phi2[nc][p];
                const int m = (nd2_i(nc) - nd2_i(nr) + 1) + 3 * (nd2_j(nc) - nd2_j(nr) + 1) + 9;
                a[(0 * 27 + m) * WAVE] = a[(0 * 27 + m) * WAVE] + Ku;
                a[(3 * 27 + m) * WAVE] = a[(3 * 27 + m) * WAVE] + Kv;
            }
        }
    }
}

}  // namespace cg
