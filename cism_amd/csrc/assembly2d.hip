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
template <int NR>
__device__ __forceinline__ bool add_element_row_2d(const Dims& d, const AsmDev& in, Lev L,
                                                   const double* __restrict__ uv, const double* __restrict__ vv,
                                                   int ci, int cj, double (&acc)[4][9]) {
    Elem2 e;
    load_elem2(d, in, uv, vv, ci, cj, e);
    const double fl = flwafact_2d(d, in, L, ci, cj);
    double K[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 4; ++c) K[b][c] = 0.0;
    bool ok = true;
    for (int p = 0; p < 4; ++p) {
        double dx[4], dy[4], detJ;
        if (!basis_derivs_2d(e, p, dx, dy, detJ)) { ok = false; break; }
        const double ef = efvs_ssa(in, e, dx, dy, fl);
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
                                                     const double* __restrict__ vv, double* __restrict__ A,
                                                     double* __restrict__ flag) {
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
            ok = add_element_row_2d<NR>(d, in, L, uv, vv, ci, cj, acc) && ok;             \
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

}  // namespace cg
