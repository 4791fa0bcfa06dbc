// kernels.hip -- device kernels of the structured 3-D Krylov path.
//
// SpMV (matvec_multiply_structured_3d, glissade_velo_higher_pcg.F90:5528-5652),
// the SIA column-tridiagonal and diagonal preconditioners (:3598-3733,
// :5777-5922), the fused Chronopoulos-Gear / standard PCG vector updates
// (:3444-3459, :1395-1401), the one-rank halo rules of
// staggered_parallel_halo (parallel_mpi.F90:10772-10936) and compensated
// global sums (the role of cism_reprosum, cism_reprosum_mod.F90:1912-2036).
// Layouts are described in cismgpu_internal.h.
#include "cismgpu_internal.h"
#include "kernels.h"

namespace cg {

// ---------------------------------------------------------------------------
// XCD-aware block order: consecutive workgroups are dispatched round-robin
// over the 8 XCDs; give each XCD a contiguous range of (j, k) tiles so that
// the x rows a tile shares with its neighbours stay in that XCD's L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_remap(int lin, int nblk) {
    const int nx8 = nblk / 8;
    if (nx8 == 0 || lin >= nx8 * 8) return lin;
    return (lin % 8) * nx8 + lin / 8;
}

// ---------------------------------------------------------------------------
// SpMV: one thread per owned vertex column (io, jo); loop over levels.
// MODE 0: y = A x
// MODE 1: y = A x, partial (y, x) over owned            (delta, eta2)
// MODE 2: r = b - A x, partial (r, r) over owned        (convergence check)
// TA: matrix storage type; TV: vector type; accumulation in TV (double or
// float), partial sums in compensated FP64.
// ---------------------------------------------------------------------------
template <typename TA, typename TV, int MODE>
__global__ __launch_bounds__(256) void k_spmv(Dims d, const TA* __restrict__ A,
                                              const TV* __restrict__ xu, const TV* __restrict__ xv,
                                              TV* __restrict__ yu, TV* __restrict__ yv,
                                              const TV* __restrict__ bu, const TV* __restrict__ bv,
                                              double* __restrict__ partials) {
    const int nbx = d.nxp / WAVE;
    const int nblk = nbx * ((d.nyo + 3) / 4);
    const int lin = xcd_remap(blockIdx.x, nblk);
    const int bx = lin % nbx, by = lin / nbx;
    const int lane = threadIdx.x % WAVE;
    const int io = bx * WAVE + lane;
    const int jo = by * 4 + threadIdx.x / WAVE;
    const bool live = (io < d.nxo) && (jo < d.nyo);
    dd acc = {0.0, 0.0};
    if (live) {
        const int i = d.ilo + io, j = d.jlo + jo;
        const long long pl = d.plane;
        const int px = d.px;
        for (int k = 0; k < d.nz; ++k) {
            const long long r = ((long long)k * d.nyo + jo) * d.nxp + io;
            const TA* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
            const long long c0 = vidx(d, k, i, j);
            TV su = TV(0), sv = TV(0);
#pragma unroll
            for (int kA = -1; kA <= 1; ++kA)
#pragma unroll
                for (int jA = -1; jA <= 1; ++jA)
#pragma unroll
                    for (int iA = -1; iA <= 1; ++iA) {
                        const int m = (iA + 1) + 3 * (jA + 1) + 9 * (kA + 1);
                        const long long n = c0 + kA * pl + jA * px + iA;
                        const TV xun = xu[n], xvn = xv[n];
                        su = su + TV(ld_coef(a + (0 * 27 + m) * WAVE)) * xun + TV(ld_coef(a + (1 * 27 + m) * WAVE)) * xvn;
                        sv = sv + TV(ld_coef(a + (2 * 27 + m) * WAVE)) * xun + TV(ld_coef(a + (3 * 27 + m) * WAVE)) * xvn;
                    }
            if (MODE == 2) {
                const TV ru = bu[c0] - su, rv = bv[c0] - sv;
                yu[c0] = ru;
                yv[c0] = rv;
                acc = dd_add_d(acc, (double)ru * (double)ru);
                acc = dd_add_d(acc, (double)rv * (double)rv);
            } else {
                yu[c0] = su;
                yv[c0] = sv;
                if (MODE == 1) {
                    acc = dd_add_d(acc, (double)su * (double)xu[c0]);
                    acc = dd_add_d(acc, (double)sv * (double)xv[c0]);
                }
            }
        }
    }
    if (MODE != 0) {
        dd v[1] = {acc};
        block_reduce_dd<1, 256>(v);
        if (threadIdx.x == 0) {
            partials[2 * lin + 0] = v[0].hi;
            partials[2 * lin + 1] = v[0].lo;
        }
    }
}

// ---------------------------------------------------------------------------
// Low-precision SpMV on the quad layout (FP32 or BF16 storage, FP32 vectors).
// For owned row r and neighbour m the four block coefficients are adjacent,
//     Aq[((r/64)*27 + m)*256 + (r%64)*4 + blk]      (blk = uu, uv, vu, vv)
// so a lane fetches them with one 16-byte (FP32) or 8-byte (BF16) load and
// a wavefront reads 1 KiB (512 B) contiguous per neighbour.  Each thread
// owns one column and walks up the levels keeping the 3x3 neighbourhood of
// levels k-1, k, k+1 in registers, so x is loaded once per level instead of
// three times.  Same MODEs and partial sums as k_spmv.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ld_quad(const float* p, float& a, float& b, float& c, float& e) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    a = q.x; b = q.y; c = q.z; e = q.w;
}
__device__ __forceinline__ void ld_quad(const __hip_bfloat16* p, float& a, float& b, float& c, float& e) {
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    a = __uint_as_float(q.x << 16);
    b = __uint_as_float(q.x & 0xffff0000u);
    c = __uint_as_float(q.y << 16);
    e = __uint_as_float(q.y & 0xffff0000u);
}

template <typename TA, int MODE>
__global__ __launch_bounds__(256) void k_spmv_q(Dims d, const TA* __restrict__ A, const float* __restrict__ xu,
                                                const float* __restrict__ xv, float* __restrict__ yu,
                                                float* __restrict__ yv, const float* __restrict__ bu,
                                                const float* __restrict__ bv, double* __restrict__ partials) {
    const int nbx = d.nxp / WAVE;
    const int bx = blockIdx.x % nbx, by = blockIdx.x / nbx;
    const int lane = threadIdx.x % WAVE;
    const int io = bx * WAVE + lane;
    const int jo = by * 4 + threadIdx.x / WAVE;
    double acc = 0.0;     // per-column sum in FP64; the cross-column reduction is compensated
    if (io < d.nxo && jo < d.nyo) {
        const int i = d.ilo + io, j = d.jlo + jo;
        const long long pl = d.plane;
        const int px = d.px;
        float wu[3][9], wv[3][9];
        const long long cb = vidx(d, -1, i, j);
#pragma unroll
        for (int l = 0; l < 2; ++l)
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                const long long n = cb + l * pl + (h / 3 - 1) * px + (h % 3 - 1);
                wu[l][h] = xu[n];
                wv[l][h] = xv[n];
            }
        for (int k = 0; k < d.nz; ++k) {
            const long long cup = cb + (long long)(k + 2) * pl;
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                const long long n = cup + (h / 3 - 1) * px + (h % 3 - 1);
                wu[2][h] = xu[n];
                wv[2][h] = xv[n];
            }
            const long long r = ((long long)k * d.nyo + jo) * d.nxp + io;
            const TA* a = A + (r / WAVE) * (27 * 4 * WAVE) + (r % WAVE) * 4;
            float su = 0.f, sv = 0.f;
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                float cuu, cuv, cvu, cvv;
                ld_quad(a + m * (4 * WAVE), cuu, cuv, cvu, cvv);
                const float xun = wu[m / 9][m % 9], xvn = wv[m / 9][m % 9];
                su = su + cuu * xun + cuv * xvn;
                sv = sv + cvu * xun + cvv * xvn;
            }
            const long long c0 = cb + (long long)(k + 1) * pl;
            if (MODE == 2) {
                const float ru = bu[c0] - su, rv = bv[c0] - sv;
                yu[c0] = ru;
                yv[c0] = rv;
                acc += (double)ru * (double)ru + (double)rv * (double)rv;
            } else {
                yu[c0] = su;
                yv[c0] = sv;
                if (MODE == 1) {
                    acc += (double)su * (double)wu[1][4] + (double)sv * (double)wv[1][4];
                }
            }
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                wu[0][h] = wu[1][h]; wv[0][h] = wv[1][h];
                wu[1][h] = wu[2][h]; wv[1][h] = wv[2][h];
            }
        }
    }
    if (MODE != 0) {
        dd v[1] = {{acc, 0.0}};
        block_reduce_dd<1, 256>(v);
        if (threadIdx.x == 0) {
            partials[2 * blockIdx.x + 0] = v[0].hi;
            partials[2 * blockIdx.x + 1] = v[0].lo;
        }
    }
}

// ---------------------------------------------------------------------------
// One-level SpMV of the 2-D (depth-integrated) systems: the 9-point stencil
// is the kA = 0 plane (m = 9..17) of the 27-point row storage; the other two
// planes hold zeros (they couple outside the single level) and are not read.
// FP64 rows use the row-blocked SoA layout, FP32/BF16 rows the quad layout.
// Same MODEs and partial sums as k_spmv.
// ---------------------------------------------------------------------------
template <typename TA, typename TV, int MODE>
__global__ __launch_bounds__(256) void k_spmv_plane(Dims d, const TA* __restrict__ A,
                                                    const TV* __restrict__ xu, const TV* __restrict__ xv,
                                                    TV* __restrict__ yu, TV* __restrict__ yv,
                                                    const TV* __restrict__ bu, const TV* __restrict__ bv,
                                                    double* __restrict__ partials) {
    const int nbx = d.nxp / WAVE;
    const int nblk = nbx * ((d.nyo + 3) / 4);
    const int lin = xcd_remap(blockIdx.x, nblk);
    const int bx = lin % nbx, by = lin / nbx;
    const int io = bx * WAVE + threadIdx.x % WAVE;
    const int jo = by * 4 + threadIdx.x / WAVE;
    dd acc = {0.0, 0.0};
    if (io < d.nxo && jo < d.nyo) {
        const int i = d.ilo + io, j = d.jlo + jo;
        const long long c0 = vidx(d, 0, i, j);
        const long long r = (long long)jo * d.nxp + io;
        TV su = TV(0), sv = TV(0);
#pragma unroll
        for (int jA = -1; jA <= 1; ++jA)
#pragma unroll
            for (int iA = -1; iA <= 1; ++iA) {
                const int m = 9 + (iA + 1) + 3 * (jA + 1);
                const long long n = c0 + jA * d.px + iA;
                const TV xun = xu[n], xvn = xv[n];
                TV cuu, cuv, cvu, cvv;
                if constexpr (std::is_same<TA, double>::value) {
                    const TA* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
                    cuu = a[(0 * 27 + m) * WAVE]; cuv = a[(1 * 27 + m) * WAVE];
                    cvu = a[(2 * 27 + m) * WAVE]; cvv = a[(3 * 27 + m) * WAVE];
                } else {
                    float q0, q1, q2, q3;
                    ld_quad(A + (r / WAVE) * (27 * 4 * WAVE) + m * (4 * WAVE) + (r % WAVE) * 4, q0, q1, q2, q3);
                    cuu = q0; cuv = q1; cvu = q2; cvv = q3;
                }
                su = su + cuu * xun + cuv * xvn;
                sv = sv + cvu * xun + cvv * xvn;
            }
        if (MODE == 2) {
            const TV ru = bu[c0] - su, rv = bv[c0] - sv;
            yu[c0] = ru;
            yv[c0] = rv;
            acc = dd_add_d(acc, (double)ru * (double)ru);
            acc = dd_add_d(acc, (double)rv * (double)rv);
        } else {
            yu[c0] = su;
            yv[c0] = sv;
            if (MODE == 1) {
                acc = dd_add_d(acc, (double)su * (double)xu[c0]);
                acc = dd_add_d(acc, (double)sv * (double)xv[c0]);
            }
        }
    }
    if (MODE != 0) {
        dd v[1] = {acc};
        block_reduce_dd<1, 256>(v);
        if (threadIdx.x == 0) {
            partials[2 * lin + 0] = v[0].hi;
            partials[2 * lin + 1] = v[0].lo;
        }
    }
}

// Non-owned ring of the vertex grid: y = 0 after a matvec (the reference
// zeroes the whole output, pcg.F90:5605-5606), or r = b - 0 after a
// residual.  One thread per ring column.
template <typename TV>
__global__ void k_ring(Dims d, TV* __restrict__ yu, TV* __restrict__ yv, const TV* __restrict__ bu,
                       const TV* __restrict__ bv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ncol = (long long)d.nvx * d.nvy;
    if (t >= ncol) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    if (i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi) return;
    for (int k = 0; k < d.nz; ++k) {
        const long long c = vidx(d, k, i, j);
        yu[c] = bu ? bu[c] : TV(0);
        yv[c] = bv ? bv[c] : TV(0);
    }
}

// ---------------------------------------------------------------------------
// Halo rules of staggered_parallel_halo for one task (parallel_mpi.F90:
// 10772-10936): periodic/no_penetration copy owned edge rows into the halo
// (E/W over owned rows first, then N/S over all columns); outflow zeroes
// the outermost row; no_ice zeroes lhalo rows and uhalo+1 rows (which
// includes the last owned E/N row).  nlev levels per vertex.
// ---------------------------------------------------------------------------
template <typename TV>
__global__ void k_halo_ew(Dims d, TV* __restrict__ a, TV* __restrict__ b, int nlev) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // over (k, j in owned rows)
    const int nj = d.jhi - d.jlo + 1;
    if (t >= nlev * nj) return;
    const int k = t % nlev, j = d.jlo + t / nlev;
    const int nv = d.nvx;
    for (int s = 0; s < 2; ++s) {
        TV* v = s ? b : a;
        if (!v) continue;
        const TV w = v[vidx(d, k, 3, j)];
        const TV e0 = v[vidx(d, k, nv - 2, j)], e1 = v[vidx(d, k, nv - 1, j)];
        v[vidx(d, k, nv, j)] = w;
        v[vidx(d, k, 1, j)] = e0;
        v[vidx(d, k, 2, j)] = e1;
    }
}

template <typename TV>
__global__ void k_halo_ns(Dims d, TV* __restrict__ a, TV* __restrict__ b, int nlev) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // over (k, i all columns)
    if (t >= nlev * d.nvx) return;
    const int k = t % nlev, i = 1 + t / nlev;
    const int nv = d.nvy;
    for (int s = 0; s < 2; ++s) {
        TV* v = s ? b : a;
        if (!v) continue;
        const TV so = v[vidx(d, k, i, 3)];
        const TV n0 = v[vidx(d, k, i, nv - 2)], n1 = v[vidx(d, k, i, nv - 1)];
        v[vidx(d, k, i, nv)] = so;
        v[vidx(d, k, i, 1)] = n0;
        v[vidx(d, k, i, 2)] = n1;
    }
}

template <typename TV>
__global__ void k_halo_zero(Dims d, TV* __restrict__ a, TV* __restrict__ b, int nlev, int lw, int uw) {
    // zero columns i <= lw or i >= nvx-uw+1, and rows j <= lw or j >= nvy-uw+1
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ncol = (long long)d.nvx * d.nvy;
    if (t >= ncol) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    if (!(i <= lw || i >= d.nvx - uw + 1 || j <= lw || j >= d.nvy - uw + 1)) return;
    for (int k = 0; k < nlev; ++k) {
        const long long c = vidx(d, k, i, j);
        if (a) a[c] = TV(0);
        if (b) b[c] = TV(0);
    }
}

// ---------------------------------------------------------------------------
// Preconditioners, one thread per vertex column over the full extent.
// SIA: easy_sia_solver -- Thomas solve of the vertical tridiagonal part of
// Auu/Avv on active columns; inactive columns are left untouched.
// M holds 6 fields (sub, diag, sup for u; then for v) in vector layout.
// With sym set the matrix is known to be exactly symmetric, so the sub
// diagonal of level k equals the super diagonal of level k-1, which the
// elimination has just read: the sub-diagonal field is not read at all
// (the same values, 4 or 8 bytes per node less).
// ---------------------------------------------------------------------------
template <typename TV, int NZMAX>
__device__ __forceinline__ void thomas_column(const Dims& d, const TV* __restrict__ Ms, const TV* __restrict__ Md,
                                              const TV* __restrict__ Mp, const TV (&rhs)[NZMAX], TV (&z)[NZMAX],
                                              long long c0, bool sym) {
    TV gam[NZMAX];
    TV beta = Md[c0];
    z[0] = rhs[0] / beta;
#pragma unroll
    for (int k = 1; k < NZMAX; ++k) {
        if (k < d.nz) {
            const long long c = c0 + (long long)k * d.plane;
            const TV sp = Mp[c - d.plane];
            gam[k] = sp / beta;
            const TV sb = sym ? sp : Ms[c];
            beta = Md[c] - sb * gam[k];
            z[k] = (rhs[k] - sb * z[k - 1]) / beta;
        }
    }
#pragma unroll
    for (int k = NZMAX - 2; k >= 0; --k)
        if (k < d.nz - 1) z[k] = z[k] - gam[k + 1] * z[k + 1];
}

// Fused Chronopoulos-Gear update (pcg.F90:3444-3459) over all levels of one
// column, optionally followed by the SIA solve z = M^-1 r on the new
// residual (the next iteration's first step, :3301-3381) and the partial
// (r, z) over owned columns.  With PRE = 0 only the update is done.
// scal[S_ALPHA], scal[S_BETA] are read on the device.
template <typename TV, int NZMAX, int PRE, int STD, int SYM>
__global__ __launch_bounds__(256) void k_update(Dims d, const double* __restrict__ scal,
                                                TV* __restrict__ xu, TV* __restrict__ xv,
                                                TV* __restrict__ ru, TV* __restrict__ rv,
                                                TV* __restrict__ du, TV* __restrict__ dv,
                                                TV* __restrict__ qu, TV* __restrict__ qv,
                                                TV* __restrict__ zu, TV* __restrict__ zv,
                                                const TV* __restrict__ Azu, const TV* __restrict__ Azv,
                                                const TV* __restrict__ M, const int8_t* __restrict__ active,
                                                double* __restrict__ partials) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ncol = (long long)d.nvx * d.nvy;
    dd acc = {0.0, 0.0};
    dd acc2 = {0.0, 0.0};
    if (scal[S_STOP] == 0.0 && t < ncol) {
        const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
        const TV alpha = (TV)scal[S_ALPHA], beta = (TV)scal[S_BETA];
        const long long c0 = vidx(d, 0, i, j);
        TV rU[NZMAX], rV[NZMAX];
#pragma unroll
        for (int k = 0; k < NZMAX; ++k) {
            if (k < d.nz) {
                const long long c = c0 + (long long)k * d.plane;
                TV dU, dV;
                if (STD) {   // standard PCG: x += alpha d; r -= alpha q (pcg.F90:1395-1401)
                    dU = du[c]; dV = dv[c];
                    xu[c] = xu[c] + alpha * dU; xv[c] = xv[c] + alpha * dV;
                    rU[k] = ru[c] - alpha * qu[c]; rV[k] = rv[c] - alpha * qv[c];
                } else {
                    dU = zu[c] + beta * du[c]; dV = zv[c] + beta * dv[c];
                    const TV qU = Azu[c] + beta * qu[c], qV = Azv[c] + beta * qv[c];
                    du[c] = dU; dv[c] = dV; qu[c] = qU; qv[c] = qV;
                    xu[c] = xu[c] + alpha * dU; xv[c] = xv[c] + alpha * dV;
                    rU[k] = ru[c] - alpha * qU; rV[k] = rv[c] - alpha * qV;
                }
                ru[c] = rU[k]; rv[c] = rV[k];
            }
        }
        if (PRE) {
            const bool owned = i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi;
            if (active[t]) {
                TV zU[NZMAX], zV[NZMAX];
                const long long vl = d.vlen;
                thomas_column<TV, NZMAX>(d, M + 0 * vl, M + 1 * vl, M + 2 * vl, rU, zU, c0, SYM != 0);
                thomas_column<TV, NZMAX>(d, M + 3 * vl, M + 4 * vl, M + 5 * vl, rV, zV, c0, SYM != 0);
#pragma unroll
                for (int k = 0; k < NZMAX; ++k)
                    if (k < d.nz) {
                        const long long c = c0 + (long long)k * d.plane;
                        zu[c] = zU[k]; zv[c] = zV[k];
                        if (owned) {
                            acc = dd_add_d(acc, (double)rU[k] * (double)zU[k]);
                            acc = dd_add_d(acc, (double)rV[k] * (double)zV[k]);
                            acc2 = dd_add_d(acc2, (double)rU[k] * (double)rU[k]);
                            acc2 = dd_add_d(acc2, (double)rV[k] * (double)rV[k]);
                        }
                    }
            } else if (owned) {
                // inactive owned column: z keeps its value (zero on the solver path)
#pragma unroll
                for (int k = 0; k < NZMAX; ++k)
                    if (k < d.nz) {
                        const long long c = c0 + (long long)k * d.plane;
                        acc = dd_add_d(acc, (double)rU[k] * (double)zu[c]);
                        acc = dd_add_d(acc, (double)rV[k] * (double)zv[c]);
                        acc2 = dd_add_d(acc2, (double)rU[k] * (double)rU[k]);
                        acc2 = dd_add_d(acc2, (double)rV[k] * (double)rV[k]);
                    }
            }
        }
    }
    if (PRE) {
        dd v[2] = {acc, acc2};
        block_reduce_dd<2, 256>(v);
        if (threadIdx.x == 0) {
            partials[4 * blockIdx.x + 0] = v[0].hi;
            partials[4 * blockIdx.x + 1] = v[0].lo;
            partials[4 * blockIdx.x + 2] = v[1].hi;
            partials[4 * blockIdx.x + 3] = v[1].lo;
        }
    }
}

// Stand-alone preconditioner apply z = PC(r) with partial (r, z) and
// (r, r) over owned columns.  kind: 0 none (copy), 1 diag, 2 SIA.
// For kinds 3/4 the caller adds the horizontal line-solve correction; kind 5
// (one-level systems) takes that line solution alone.
template <typename TV, int NZMAX>
__global__ __launch_bounds__(256) void k_precond(Dims d, int kind, const double* __restrict__ scal,
                                                 const TV* __restrict__ ru, const TV* __restrict__ rv,
                                                 TV* __restrict__ zu, TV* __restrict__ zv,
                                                 const TV* __restrict__ M, const TV* __restrict__ Dg,
                                                 const int8_t* __restrict__ active,
                                                 const TV* __restrict__ zvs_u, const TV* __restrict__ zvs_v,
                                                 double* __restrict__ partials, int msym) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ncol = (long long)d.nvx * d.nvy;
    dd acc = {0.0, 0.0}, acc2 = {0.0, 0.0};
    if (scal[S_STOP] == 0.0 && t < ncol) {
        const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
        const bool owned = i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi;
        const long long c0 = vidx(d, 0, i, j);
        TV rU[NZMAX], rV[NZMAX], zU[NZMAX], zV[NZMAX];
#pragma unroll
        for (int k = 0; k < NZMAX; ++k)
            if (k < d.nz) {
                const long long c = c0 + (long long)k * d.plane;
                rU[k] = ru[c]; rV[k] = rv[c];
            }
        if (kind == 0) {
#pragma unroll
            for (int k = 0; k < NZMAX; ++k) { zU[k] = rU[k]; zV[k] = rV[k]; }
        } else if (kind == 1) {
#pragma unroll
            for (int k = 0; k < NZMAX; ++k)
                if (k < d.nz) {
                    const long long c = c0 + (long long)k * d.plane;
                    const TV gu = Dg[c], gv = Dg[d.vlen + c];
                    zU[k] = gu != TV(0) ? rU[k] / gu : TV(0);
                    zV[k] = gv != TV(0) ? rV[k] / gv : TV(0);
                }
        } else if (kind == 5) {
            // 2-D solves: z is the horizontal line solution itself on every
            // vertex (pcg.F90:456-620 standard, :2020-2140 and :2275-2420 C-G)
            const long long c2 = (long long)(j - 1) * d.nvx + (i - 1);
#pragma unroll
            for (int k = 0; k < NZMAX; ++k) { zU[k] = zvs_u[c2]; zV[k] = zvs_v[c2]; }
        } else {
            if (active[t]) {
                const long long vl = d.vlen;
                thomas_column<TV, NZMAX>(d, M + 0 * vl, M + 1 * vl, M + 2 * vl, rU, zU, c0, msym != 0);
                thomas_column<TV, NZMAX>(d, M + 3 * vl, M + 4 * vl, M + 5 * vl, rV, zV, c0, msym != 0);
            } else {
#pragma unroll
                for (int k = 0; k < NZMAX; ++k)
                    if (k < d.nz) {
                        const long long c = c0 + (long long)k * d.plane;
                        zU[k] = zu[c]; zV[k] = zv[c];
                    }
            }
            if (kind >= 3) {   // z = z_SIA + z_vsum(i,j) on every column (pcg.F90:4650-4656)
                const long long c2 = (long long)(j - 1) * d.nvx + (i - 1);
                const TV au = zvs_u[c2], av = zvs_v[c2];
#pragma unroll
                for (int k = 0; k < NZMAX; ++k) { zU[k] = zU[k] + au; zV[k] = zV[k] + av; }
            }
        }
#pragma unroll
        for (int k = 0; k < NZMAX; ++k)
            if (k < d.nz) {
                const long long c = c0 + (long long)k * d.plane;
                zu[c] = zU[k]; zv[c] = zV[k];
                if (owned) {
                    acc = dd_add_d(acc, (double)rU[k] * (double)zU[k]);
                    acc = dd_add_d(acc, (double)rV[k] * (double)zV[k]);
                    acc2 = dd_add_d(acc2, (double)rU[k] * (double)rU[k]);
                    acc2 = dd_add_d(acc2, (double)rV[k] * (double)rV[k]);
                }
            }
    }
    dd v[2] = {acc, acc2};
    block_reduce_dd<2, 256>(v);
    if (threadIdx.x == 0) {
        partials[4 * blockIdx.x + 0] = v[0].hi;
        partials[4 * blockIdx.x + 1] = v[0].lo;
        partials[4 * blockIdx.x + 2] = v[1].hi;
        partials[4 * blockIdx.x + 3] = v[1].lo;
    }
}

// Column sums of r for the horizontal tridiagonal preconditioners
// (pcg.F90:4585-4590): rvs(i,j) = sum_k r(k,i,j), reference order.
template <typename TV>
__global__ void k_vsum(Dims d, const TV* __restrict__ ru, const TV* __restrict__ rv, TV* __restrict__ su,
                       TV* __restrict__ sv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ncol = (long long)d.nvx * d.nvy;
    if (t >= ncol) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    TV a = TV(0), b = TV(0);
    for (int k = 0; k < d.nz; ++k) {
        const long long c = vidx(d, k, i, j);
        a = a + ru[c];
        b = b + rv[c];
    }
    su[t] = a;
    sv[t] = b;
}

// ---------------------------------------------------------------------------
// Line solves as parallel scans.  With the factors of k_tri_factor the
// forward sweep is the first-order recurrence g_s = (b_s - As_s g_{s-1}) de_s
// (g_lo = b_lo Ad_lo; g = 0 outside the owned block, where de = 0) and the
// back substitution x_s = g_s - om_s x_{s+1} (x = 0 past the line end, where
// g = om = 0).  Each step is an affine map y -> P + Q y; one wavefront per
// line composes the maps of 64 consecutive positions with a shuffle scan and
// carries the last value into the next 64, so a line of n positions takes
// n/64 short scans instead of n dependent steps.
// ---------------------------------------------------------------------------
template <typename TV>
__device__ __forceinline__ void affine_scan(TV& P, TV& Q, int lane) {
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const TV Pp = __shfl_up(P, off, WAVE), Qp = __shfl_up(Q, off, WAVE);
        if (lane >= off) {
            P = P + Q * Pp;
            Q = Q * Qp;
        }
    }
}

template <typename TV>
__global__ __launch_bounds__(256) void k_line_scan(Dims d, int dir, const TV* __restrict__ Ad,
                                                   const TV* __restrict__ As, const TV* __restrict__ de,
                                                   const TV* __restrict__ om, const TV* __restrict__ b,
                                                   TV* __restrict__ x, TV* __restrict__ gam) {
    const int lane = threadIdx.x % WAVE;
    const int line = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int nv = d.nvx;
    int fixed, n, lo;
    if (dir == 0) { fixed = d.jlo + line; if (fixed > d.jhi) return; n = d.nvx; lo = d.ilo; }
    else { fixed = d.ilo + line; if (fixed > d.ihi) return; n = d.nvy; lo = d.jlo; }
    auto id = [&](int s) -> long long {
        return dir == 0 ? (long long)(fixed - 1) * nv + (s - 1) : (long long)(s - 1) * nv + (fixed - 1);
    };
    // forward sweep over s = 1..n
    TV carry = TV(0);
    for (int s0 = 1; s0 <= n; s0 += WAVE) {
        const int s = s0 + lane;
        TV P = TV(0), Q = TV(1);
        if (s <= n) {
            const long long c = id(s);
            if (s == lo) { P = b[c] * Ad[c]; Q = TV(0); }
            else { const TV e = de[c]; P = b[c] * e; Q = -(As[c] * e); }
        }
        affine_scan<TV>(P, Q, lane);
        const TV y = P + Q * carry;
        if (s <= n) gam[id(s)] = y;
        carry = __shfl(y, WAVE - 1, WAVE);
    }
    // back substitution over s = n..1
    carry = TV(0);
    for (int s0 = n; s0 >= 1; s0 -= WAVE) {
        const int s = s0 - lane;
        TV P = TV(0), Q = TV(1);
        if (s >= 1) {
            const long long c = id(s);
            P = gam[c];
            Q = -om[c];
        }
        affine_scan<TV>(P, Q, lane);
        const TV y = P + Q * carry;
        if (s >= 1) x[id(s)] = y;
        carry = __shfl(y, WAVE - 1, WAVE);
    }
}

// 2-D halo of a vertex-grid field (nvx*nvy, i fastest), same rules as 3-D.
// Periodic: east/west columns first, then north/south rows (which carry the
// corners), as staggered_parallel_halo does.
template <typename TV>
__global__ void k_halo2d_ew(Dims d, TV* __restrict__ a, TV* __restrict__ b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = d.nvx;
    const int j = d.jlo + t / 2;
    if (j > d.jhi) return;
    TV* v = (t % 2) ? b : a;
    auto id = [&](int i) { return (long long)(j - 1) * nv + (i - 1); };
    const TV w = v[id(3)], e0 = v[id(nv - 2)], e1 = v[id(nv - 1)];
    v[id(nv)] = w; v[id(1)] = e0; v[id(2)] = e1;
}
template <typename TV>
__global__ void k_halo2d_ns(Dims d, TV* __restrict__ a, TV* __restrict__ b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = d.nvx, nw = d.nvy;
    const int i = 1 + t / 2;
    if (i > nv) return;
    TV* v = (t % 2) ? b : a;
    auto id = [&](int jj) { return (long long)(jj - 1) * nv + (i - 1); };
    const TV so = v[id(3)], n0 = v[id(nw - 2)], n1 = v[id(nw - 1)];
    v[id(nw)] = so; v[id(1)] = n0; v[id(2)] = n1;
}
template <typename TV>
__global__ void k_halo2d_zero(Dims d, TV* __restrict__ a, TV* __restrict__ b) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int nv = d.nvx, nw = d.nvy;
    if (t >= (long long)nv * nw) return;
    const int lw = d.bc == 1 ? 1 : 2, uw = d.bc == 1 ? 1 : 2;
    const int i = 1 + (int)(t % nv), j = 1 + (int)(t / nv);
    if (i <= lw || i >= nv - uw + 1 || j <= lw || j >= nw - uw + 1) { a[t] = TV(0); b[t] = TV(0); }
}

// Whole-array updates used by the standard PCG and the first C-G pass.
// op 0: d = z + beta d            (pcg.F90:1336-1337)
// op 1: d = z                      (pcg.F90:3240-3241)
// op 2: r = b - y                  (residual from a matvec output)
template <typename TV>
__global__ void k_axpy_ops(Dims d, int op, const double* __restrict__ scal, const TV* __restrict__ s1u,
                           const TV* __restrict__ s1v, TV* __restrict__ tu, TV* __restrict__ tv,
                           const TV* __restrict__ s2u, const TV* __restrict__ s2v) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long lo = d.plane, hi = (long long)(d.nz + 1) * d.plane;   // real levels only
    const long long c = lo + t;
    if (c >= hi || scal[S_STOP] != 0.0) return;
    if (op == 0) {
        const TV beta = (TV)scal[S_BETA];
        tu[c] = s1u[c] + beta * tu[c];
        tv[c] = s1v[c] + beta * tv[c];
    } else if (op == 1) {
        tu[c] = s1u[c];
        tv[c] = s1v[c];
    } else {
        tu[c] = s2u[c] - s1u[c];
        tv[c] = s2v[c] - s1v[c];
    }
}

// First Chronopoulos-Gear step x += alpha d; r -= alpha q (pcg.F90:3272-3276)
template <typename TV>
__global__ void k_first_update(Dims d, const double* __restrict__ scal, TV* __restrict__ xu, TV* __restrict__ xv,
                               TV* __restrict__ ru, TV* __restrict__ rv, const TV* __restrict__ du,
                               const TV* __restrict__ dv, const TV* __restrict__ qu, const TV* __restrict__ qv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long c = d.plane + t;
    if (c >= (long long)(d.nz + 1) * d.plane || scal[S_STOP] != 0.0) return;
    const TV a = (TV)scal[S_ALPHA];
    xu[c] = xu[c] + a * du[c];
    xv[c] = xv[c] + a * dv[c];
    ru[c] = ru[c] - a * qu[c];
    rv[c] = rv[c] - a * qv[c];
}

// ---------------------------------------------------------------------------
// Scalar stage: one block sums the per-block compensated partials in a fixed
// order and advances the Krylov scalars on the device.
// ---------------------------------------------------------------------------
// Sum of n compensated partials (stride s) by one thread of a
// SCAL_THREADS-thread workgroup: four independent loads in flight, then
// their adds, so the loop waits for memory once per four partials.
__device__ __forceinline__ dd sum_partials(const double* __restrict__ p, int n, int s) {
    dd v = {0.0, 0.0};
    for (int b0 = threadIdx.x; b0 < n; b0 += 4 * SCAL_THREADS) {
        double h[4], l[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = b0 + u * SCAL_THREADS;
            h[u] = b < n ? p[(long long)s * b] : 0.0;
            l[u] = b < n ? p[(long long)s * b + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b0 + u * SCAL_THREADS < n) v = dd_add(v, dd{h[u], l[u]});
    }
    return v;
}

__global__ __launch_bounds__(SCAL_THREADS) void k_scalars(int stage, const double* __restrict__ pa, int na, int sa,
                                                          const double* __restrict__ pb, int nb, int sb,
                                                          double* __restrict__ scal, double tolerance) {
    const dd va = sum_partials(pa, na, sa);
    const dd vb = pb ? sum_partials(pb, nb, sb) : dd{0.0, 0.0};
    dd v[2] = {va, vb};
    block_reduce_dd<2, SCAL_THREADS>(v);
    if (threadIdx.x != 0) return;
    const double A = v[0].hi + v[0].lo, B = v[1].hi + v[1].lo;
    if (scal[S_STOP] != 0.0 && stage != ST_BB && stage != ST_RR) return;
    switch (stage) {
    case ST_BB:   // L2_rhs = sqrt((b,b))
        scal[S_BB] = A;
        scal[S_L2RHS] = sqrt(A);
        break;
    case ST_RR: {  // err = sqrt((r,r)) / L2_rhs
        scal[S_RR] = A;
        const double err = sqrt(A) / scal[S_L2RHS];
        scal[S_ERR] = err;
        scal[S_CONV] = (err < tolerance) ? 1.0 : 0.0;
        break;
    }
    case ST_CG_FIRST: {   // rho_old = (r,z); sigma = (d,q); alpha = rho_old/sigma
        scal[S_RHO] = A;
        scal[S_SIGMA] = B;
        const double alpha = A / B;
        scal[S_ALPHA] = alpha;
        if (alpha != alpha) { scal[S_NAN] = 1.0; scal[S_STOP] = 1.0; }
        break;
    }
    case ST_CG_ITER: {    // beta = rho/rho_old; sigma = delta - beta*rho/alpha; alpha = rho/sigma
        const double rho = A, delta = B;
        const double beta = rho / scal[S_RHO];
        const double sigma = delta - beta * rho / scal[S_ALPHA];
        const double alpha = rho / sigma;
        scal[S_BETA] = beta;
        scal[S_SIGMA] = sigma;
        scal[S_ALPHA] = alpha;
        scal[S_RHO] = rho;
        if (alpha != alpha) { scal[S_NAN] = 1.0; scal[S_STOP] = 1.0; }
        break;
    }
    case ST_STD_ETA1: {   // eta1 = (r,z); beta = eta1/eta0
        const double eta1 = A;
        scal[S_RHO] = eta1;
        scal[S_BETA] = eta1 / scal[S_ETA0];
        if (eta1 != eta1) { scal[S_NAN] = 1.0; scal[S_STOP] = 1.0; }
        break;
    }
    case ST_STD_ETA2: {   // eta0 = eta1; eta2 = (d,q); alpha = eta1/eta2
        const double eta1 = scal[S_RHO];
        scal[S_ETA0] = eta1;
        scal[S_SIGMA] = A;
        const double alpha = eta1 / A;
        scal[S_ALPHA] = alpha;
        if (alpha != alpha) { scal[S_NAN] = 1.0; scal[S_STOP] = 1.0; }
        break;
    }
    case ST_DOT:          // plain dot product into S_DOT
        scal[S_DOT] = A;
        scal[S_DOT2] = B;
        break;
    }
}

// ---------------------------------------------------------------------------
// Layout conversion kernels
// ---------------------------------------------------------------------------

// Reference matrix A(27,nz,nvx,nvy) (m fastest) -> row-blocked SoA, one
// block; coefficients coupling outside 1..nz and rows of inactive columns
// become zero.
template <typename TA>
__global__ void k_pack_matrix(Dims d, int blk, const double* __restrict__ src, const int8_t* __restrict__ active,
                              TA* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;   // over rows*27
    const long long total = d.nrows * 27;
    if (t >= total) return;
    const int m = (int)(t % 27);
    const long long r = t / 27;
    const int io = (int)(r % d.nxp);
    const long long kj = r / d.nxp;
    const int jo = (int)(kj % d.nyo), k = (int)(kj / d.nyo);
    double v = 0.0;
    if (io < d.nxo) {
        const int i = d.ilo + io, j = d.jlo + jo;
        const int kA = m / 9 - 1;
        const bool in = (k + kA >= 0) && (k + kA < d.nz);
        if (in && active[(long long)(j - 1) * d.nvx + (i - 1)])
            v = src[(long long)m + 27LL * ((long long)k + (long long)d.nz * ((i - 1) + (long long)d.nvx * (j - 1)))];
    }
    dst[((r / WAVE) * NCOEF + blk * 27 + m) * WAVE + (r % WAVE)] = (TA)v;
}

// Row-blocked FP64 -> lower-precision copy (FP32 or BF16)
template <typename TA>
__global__ void k_convert_matrix(long long n, const double* __restrict__ src, TA* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (TA)src[t];
}

template <>
__global__ void k_convert_matrix<__hip_bfloat16>(long long n, const double* __restrict__ src,
                                                 __hip_bfloat16* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) dst[t] = __float2bfloat16((float)src[t]);
}

// Row-blocked FP64 -> quad-layout lower-precision copy; one thread per
// destination element, so the writes are contiguous.
template <typename TA>
__device__ __forceinline__ TA to_lowp(double v) { return (TA)v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_lowp<__hip_bfloat16>(double v) { return __float2bfloat16((float)v); }

template <typename TA>
__global__ void k_convert_quad(long long nrows, const double* __restrict__ src, TA* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= nrows * NCOEF) return;
    const int blk = (int)(t % 4);
    const long long u = t / 4;
    const int lane = (int)(u % WAVE);
    const long long v = u / WAVE;
    const int m = (int)(v % 27);
    const long long rb = v / 27;
    dst[t] = to_lowp<TA>(src[(rb * NCOEF + blk * 27 + m) * WAVE + lane]);
}

// Row-blocked -> reference layout for one block (export).
__global__ void k_unpack_matrix(Dims d, int blk, const double* __restrict__ A, double* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long total = 27LL * d.nz * d.nvx * d.nvy;
    if (t >= total) return;
    const int m = (int)(t % 27);
    long long rest = t / 27;
    const int k = (int)(rest % d.nz); rest /= d.nz;
    const int i = 1 + (int)(rest % d.nvx);
    const int j = 1 + (int)(rest / d.nvx);
    double v = 0.0;
    if (i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi) {
        const long long r = ((long long)k * d.nyo + (j - d.jlo)) * d.nxp + (i - d.ilo);
        v = A[((r / WAVE) * NCOEF + blk * 27 + m) * WAVE + (r % WAVE)];
    }
    dst[t] = v;
}

// 2-D stencil A(nvx, nvy, 9) (m last, m = 1 + (iA+1) + 3(jA+1)) -> the kA = 0
// plane of a one-level reference row A(27, 1, nvx, nvy); other planes zero.
__global__ void k_expand_stencil_2d(long long n2, const double* __restrict__ src, double* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;   // over 27 * n2
    if (t >= 27 * n2) return;
    const int m = (int)(t % 27);
    const long long p = t / 27;
    dst[t] = (m >= 9 && m < 18) ? src[p + n2 * (m - 9)] : 0.0;
}

// Reference vector (nz, nvx, nvy) <-> padded level-major layout.
template <typename TS, typename TD>
__global__ void k_vec_in(Dims d, const TS* __restrict__ src, TD* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)d.nz * d.nvx * d.nvy;
    if (t >= n) return;
    const int k = (int)(t % d.nz);
    const long long col = t / d.nz;
    const int i = 1 + (int)(col % d.nvx), j = 1 + (int)(col / d.nvx);
    dst[vidx(d, k, i, j)] = (TD)src[t];
}
template <typename TS, typename TD>
__global__ void k_vec_out(Dims d, const TS* __restrict__ src, TD* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)d.nz * d.nvx * d.nvy;
    if (t >= n) return;
    const int k = (int)(t % d.nz);
    const long long col = t / d.nz;
    const int i = 1 + (int)(col % d.nvx), j = 1 + (int)(col / d.nvx);
    dst[t] = (TD)src[vidx(d, k, i, j)];
}

// SIA / diagonal preconditioner coefficients from the reference-layout Auu,
// Avv over every vertex (setup_preconditioner_sia_3d, pcg.F90:3697-3726;
// setup_preconditioner_diag_3d, :3637-3640).
template <typename TV>
__global__ void k_setup_precond(Dims d, const double* __restrict__ Auu, const double* __restrict__ Avv,
                                TV* __restrict__ M, TV* __restrict__ Dg) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)d.nz * d.nvx * d.nvy;
    if (t >= n) return;
    const int k = (int)(t % d.nz);
    const long long col = t / d.nz;
    const int i = 1 + (int)(col % d.nvx), j = 1 + (int)(col / d.nvx);
    const long long c = vidx(d, k, i, j);
    const double* au = Auu + 27 * t;
    const double* av = Avv + 27 * t;
    M[0 * d.vlen + c] = (TV)au[4];
    M[1 * d.vlen + c] = (TV)au[13];
    M[2 * d.vlen + c] = (TV)au[22];
    M[3 * d.vlen + c] = (TV)av[4];
    M[4 * d.vlen + c] = (TV)av[13];
    M[5 * d.vlen + c] = (TV)av[22];
    Dg[c] = (TV)au[13];
    Dg[d.vlen + c] = (TV)av[13];
}

template <typename TS, typename TD>
__global__ void k_cast(long long n, const TS* __restrict__ s, TD* __restrict__ dd_) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) dd_[t] = (TD)s[t];
}

// x64 += (double) e32 over real levels (iterative refinement update)
__global__ void k_ir_update(Dims d, double* __restrict__ xu, double* __restrict__ xv, const float* __restrict__ eu,
                            const float* __restrict__ ev) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long c = d.plane + t;
    if (c >= (long long)(d.nz + 1) * d.plane) return;
    xu[c] = xu[c] + (double)eu[c];
    xv[c] = xv[c] + (double)ev[c];
}

// ---------------------------------------------------------------------------
// explicit instantiations used by the host drivers
// ---------------------------------------------------------------------------
#define INST_SPMV(TA, TV)                                                                                   \
    template __global__ void k_spmv<TA, TV, 0>(Dims, const TA*, const TV*, const TV*, TV*, TV*, const TV*,   \
                                               const TV*, double*);                                        \
    template __global__ void k_spmv<TA, TV, 1>(Dims, const TA*, const TV*, const TV*, TV*, TV*, const TV*,   \
                                               const TV*, double*);                                        \
    template __global__ void k_spmv<TA, TV, 2>(Dims, const TA*, const TV*, const TV*, TV*, TV*, const TV*,   \
                                               const TV*, double*);
INST_SPMV(double, double)
INST_SPMV(float, float)
INST_SPMV(__hip_bfloat16, float)
#define INST_SPMV_Q(TA)                                                                                      \
    template __global__ void k_spmv_q<TA, 0>(Dims, const TA*, const float*, const float*, float*, float*,     \
                                             const float*, const float*, double*);                          \
    template __global__ void k_spmv_q<TA, 1>(Dims, const TA*, const float*, const float*, float*, float*,     \
                                             const float*, const float*, double*);                          \
    template __global__ void k_spmv_q<TA, 2>(Dims, const TA*, const float*, const float*, float*, float*,     \
                                             const float*, const float*, double*);
INST_SPMV_Q(float)
INST_SPMV_Q(__hip_bfloat16)
#define INST_SPMV_PLANE(TA, TV)                                                                              \
    template __global__ void k_spmv_plane<TA, TV, 0>(Dims, const TA*, const TV*, const TV*, TV*, TV*,       \
                                                     const TV*, const TV*, double*);                        \
    template __global__ void k_spmv_plane<TA, TV, 1>(Dims, const TA*, const TV*, const TV*, TV*, TV*,       \
                                                     const TV*, const TV*, double*);                        \
    template __global__ void k_spmv_plane<TA, TV, 2>(Dims, const TA*, const TV*, const TV*, TV*, TV*,       \
                                                     const TV*, const TV*, double*);
INST_SPMV_PLANE(double, double)
INST_SPMV_PLANE(float, float)
INST_SPMV_PLANE(__hip_bfloat16, float)
template __global__ void k_convert_quad<float>(long long, const double*, float*);
template __global__ void k_convert_quad<__hip_bfloat16>(long long, const double*, __hip_bfloat16*);

}  // namespace cg
