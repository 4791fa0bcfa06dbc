// ensemble.hip -- batched solves of many independent velocity systems in one
// persistent launch.
//
// Each member is an assembled FP64 system (owned rows of A, b, the SIA
// column-tridiagonal preconditioner).  The solve is FP64 iterative
// refinement around an FP32 Chronopoulos-Gear inner solve (the same
// recurrence as pcg_solver_chrongear_3d, glissade_velo_higher_pcg.F90:
// 3296-3581, with beta = 0 and sigma = delta on the first inner step).
// All members advance in lock step through a sequence of "steps"; a step is
// a matrix phase (FP64 residual or FP32 SpMV, per member state) and a column
// phase (refinement bookkeeping or the fused C-G update + SIA solve), each
// ended by a grid-wide barrier.  Work units of a phase are taken from an
// atomic queue so that the members still iterating are spread over every
// workgroup.  Every per-member decision is a deterministic function of
// per-tile compensated partial sums reduced in a fixed order, so all
// workgroups reach the same decision without extra communication.
//
// Members must have an ice-free outer ring (no active vertex on a row that
// the halo rule writes), so halo updates are identities and are omitted;
// the host checks this when the ensemble is built.
#include "cismgpu_internal.h"

namespace cg {

constexpr int ENS_BS = 256;          // threads per workgroup
constexpr int ENS_WPB = ENS_BS / WAVE;

struct EnsMember {
    Dims d;                          // member grid (vector layout of the member)
    long long a_off;                 // offset (rows) of the member's compact matrix rows
    long long v_off;                 // offset (elements) of the member's vectors
    long long c_off;                 // offset of the member's vertex-column arrays (nvx*nvy)
    int ncol, ncolp;                 // owned columns nxo*nyo and padded to 64
    int split;                       // levels split across `split` lanes (1, 2, 4)
    int ntile_mat;                   // matrix-phase tiles (waves) of this member
    int ntile_col;                   // column-phase tiles (workgroups) of this member
    int t0_mat, t0_col;              // first tile indices in the global lists
    double bnorm;                    // ||b|| (FP64, compensated), set at build
};

// per-member state, double buffered by step parity
enum { M_OUTER = 0, M_INNER = 1, M_FOLD = 2, M_DONE = 3 };
struct EnsState {
    int phase;                       // M_*
    int inner_it;                    // inner iterations of the current refinement step
    int outer_it;                    // refinement steps done
    int total_inner;
    double rho, alpha;               // C-G scalars (rho_old, alpha_old)
    double r0;                       // ||r|| at the start of the inner solve
    double err;                      // last outer relative residual
    int flags;                       // 1 converged, 2 NaN, 4 inner limit
};

struct EnsParams {
    int nmem, ntile_mat, ntile_col;
    int maxsteps;
    int inner_maxiters, max_outer, ncheck;
    double tol, inner_tol;
    long long spin_limit;
};

struct EnsPtrs {
    const EnsMember* mem;
    const int* tile_mem_mat;         // global matrix tile -> member
    const int* tile_mem_col;         // global column tile -> member
    const double* A64;               // compact rows, row-blocked (108 coefs x 64)
    const float* A32;
    const __hip_bfloat16* A16;
    const float* M32;                // SIA coefficients per member vector layout (6 fields)
    const int8_t* active;
    const double* b64;               // 2 components per member block
    double* x64;
    double* r64;
    float *r32, *d32, *q32, *z32, *Az32, *e32;
    double* part_mat;                // [2][ntile_mat] dd
    double* part_col;                // [2][ntile_col][2] dd
    EnsState* st;                    // [2][nmem]
    unsigned* bar;                   // barrier: count, generation
    unsigned* queue;                 // 3 work counters
    int* err_flag;
    int use_bf16;
};

// ---------------------------------------------------------------- barrier
// Sense-reversing barrier for a grid whose workgroups are all resident (the
// launch is sized from the occupancy query).  Every wait is bounded: after
// `limit` polls the workgroup records an error and proceeds, so the grid
// always drains.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned nblocks, long long limit, int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived = atomicAdd(&bar[0], 1u) + 1u;
        if (arrived == nblocks) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long polls = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++polls > limit) { atomicExch(err, 1); break; }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// deterministic reduction of a member's tile partials (stride = dd per tile)
__device__ double member_sum(const double* part, int t0, int n, int stride, int which) {
    // one wave reduces; result broadcast through LDS by the caller
    dd acc = {0.0, 0.0};
    const int lane = threadIdx.x % WAVE;
    for (int t = lane; t < n; t += WAVE) {
        const double* p = part + (size_t)(t0 + t) * stride + 2 * which;
        acc = dd_add(acc, dd{p[0], p[1]});
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
    return __shfl(acc.hi + acc.lo, 0, WAVE);
}

// ---------------------------------------------------------------- matrix phase work
// One wave processes one matrix tile: 64/split flat owned columns, each lane
// a column and a contiguous share of the levels.
template <typename TA>
__device__ __forceinline__ void spmv_tile_f32(const EnsMember& M, int tile, const TA* __restrict__ A,
                                              const float* __restrict__ xu, const float* __restrict__ xv,
                                              float* __restrict__ yu, float* __restrict__ yv, dd& acc) {
    const Dims& d = M.d;
    const int lane = threadIdx.x % WAVE;
    const int cpw = WAVE / M.split;                      // columns per wave
    const int c = tile * cpw + lane % cpw;               // flat owned column
    const int part = lane / cpw;                         // level group
    if (c >= M.ncol) return;
    const int io = c % d.nxo, jo = c / d.nxo;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int kper = (d.nz + M.split - 1) / M.split;
    const int k0 = part * kper, k1 = min(d.nz, k0 + kper);
    const long long pl = d.plane;
    for (int k = k0; k < k1; ++k) {
        const long long r = (long long)k * M.ncolp + c + M.a_off;
        const TA* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
        const long long c0 = vidx(d, k, i, j);
        float su = 0.f, sv = 0.f;
#pragma unroll
        for (int m = 0; m < 27; ++m) {
            const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
            const long long n = c0 + kA * pl + jA * d.px + iA;
            const float xun = xu[n], xvn = xv[n];
            su = su + ld_coef(a + (0 * 27 + m) * WAVE) * xun + ld_coef(a + (1 * 27 + m) * WAVE) * xvn;
            sv = sv + ld_coef(a + (2 * 27 + m) * WAVE) * xun + ld_coef(a + (3 * 27 + m) * WAVE) * xvn;
        }
        yu[c0] = su;
        yv[c0] = sv;
        acc = dd_add_d(acc, (double)su * (double)xu[c0]);
        acc = dd_add_d(acc, (double)sv * (double)xv[c0]);
    }
}

__device__ __forceinline__ void resid_tile_f64(const EnsMember& M, int tile, const double* __restrict__ A,
                                               const double* __restrict__ xu, const double* __restrict__ xv,
                                               const double* __restrict__ bu, const double* __restrict__ bv,
                                               double* __restrict__ ru, double* __restrict__ rv, dd& acc) {
    const Dims& d = M.d;
    const int lane = threadIdx.x % WAVE;
    const int cpw = WAVE / M.split;
    const int c = tile * cpw + lane % cpw;
    const int part = lane / cpw;
    if (c >= M.ncol) return;
    const int io = c % d.nxo, jo = c / d.nxo;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int kper = (d.nz + M.split - 1) / M.split;
    const int k0 = part * kper, k1 = min(d.nz, k0 + kper);
    const long long pl = d.plane;
    for (int k = k0; k < k1; ++k) {
        const long long r = (long long)k * M.ncolp + c + M.a_off;
        const double* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
        const long long c0 = vidx(d, k, i, j);
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int m = 0; m < 27; ++m) {
            const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
            const long long n = c0 + kA * pl + jA * d.px + iA;
            const double xun = xu[n], xvn = xv[n];
            su = su + a[(0 * 27 + m) * WAVE] * xun + a[(1 * 27 + m) * WAVE] * xvn;
            sv = sv + a[(2 * 27 + m) * WAVE] * xun + a[(3 * 27 + m) * WAVE] * xvn;
        }
        const double tu = bu[c0] - su, tv = bv[c0] - sv;
        ru[c0] = tu;
        rv[c0] = tv;
        acc = dd_add_d(acc, tu * tu);
        acc = dd_add_d(acc, tv * tv);
    }
}

// ---------------------------------------------------------------- column phase work
// Thomas solve of the SIA column (easy_sia_solver) in FP32.
template <int NZMAX>
__device__ __forceinline__ void sia_col(const Dims& d, const float* __restrict__ Ms, const float* __restrict__ Md,
                                        const float* __restrict__ Mp, long long c0, const float (&rhs)[NZMAX],
                                        float (&z)[NZMAX]) {
    float gam[NZMAX];
    float beta = Md[c0];
    z[0] = rhs[0] / beta;
#pragma unroll
    for (int k = 1; k < NZMAX; ++k)
        if (k < d.nz) {
            const long long c = c0 + (long long)k * d.plane;
            gam[k] = Mp[c - d.plane] / beta;
            const float sb = Ms[c];
            beta = Md[c] - sb * gam[k];
            z[k] = (rhs[k] - sb * z[k - 1]) / beta;
        }
#pragma unroll
    for (int k = NZMAX - 2; k >= 0; --k)
        if (k < d.nz - 1) z[k] = z[k] - gam[k + 1] * z[k + 1];
}

__device__ __forceinline__ void col_of_tile(const EnsMember& M, int tile, int& i, int& j, long long& t2) {
    const int t = tile * ENS_BS + threadIdx.x;
    t2 = t;
    i = 1 + t % M.d.nvx;
    j = 1 + t / M.d.nvx;
}

template <int NZMAX>
__global__ __launch_bounds__(ENS_BS) void k_ensemble(EnsParams P, EnsPtrs Q) {
    __shared__ int s_unit;
    __shared__ double s_val[4];
    const unsigned G = gridDim.x;
    int step = 0;
    for (; step < P.maxsteps; ++step) {
        const int cur = step & 1, nxt = cur ^ 1;
        EnsState* S0 = Q.st + (size_t)cur * P.nmem;
        EnsState* S1 = Q.st + (size_t)nxt * P.nmem;
        // all members done?  (every workgroup reads the same state)
        {
            __shared__ int s_any;
            if (threadIdx.x == 0) s_any = 0;
            __syncthreads();
            for (int m = threadIdx.x; m < P.nmem; m += ENS_BS)
                if (S0[m].phase != M_DONE) s_any = 1;
            __syncthreads();
            if (!s_any) break;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) Q.queue[(step * 2 + 2) % 3] = 0u;   // counter of two phases ahead
        // ------------------------------------------------ matrix phase
        {
            unsigned* qc = Q.queue + (step * 2) % 3;
            const int nunits = (P.ntile_mat + ENS_WPB - 1) / ENS_WPB;
            while (true) {
                if (threadIdx.x == 0) s_unit = (int)atomicAdd(qc, 1u);
                __syncthreads();
                const int u = s_unit;
                __syncthreads();
                if (u >= nunits) break;
                const int tile = u * ENS_WPB + threadIdx.x / WAVE;
                if (tile < P.ntile_mat) {
                    const int m = Q.tile_mem_mat[tile];
                    const EnsMember& M = Q.mem[m];
                    const int ph = S0[m].phase;
                    dd acc = {0.0, 0.0};
                    const int lt = tile - M.t0_mat;
                    bool wrote = false;
                    if (ph == M_OUTER) {
                        const long long vo = M.v_off, vl = M.d.vlen;
                        resid_tile_f64(M, lt, Q.A64, Q.x64 + vo, Q.x64 + vo + vl, Q.b64 + vo, Q.b64 + vo + vl,
                                       Q.r64 + vo, Q.r64 + vo + vl, acc);
                        wrote = true;
                    } else if (ph == M_INNER) {
                        const long long vo = M.v_off, vl = M.d.vlen;
                        if (Q.use_bf16)
                            spmv_tile_f32<__hip_bfloat16>(M, lt, Q.A16, Q.z32 + vo, Q.z32 + vo + vl, Q.Az32 + vo,
                                                          Q.Az32 + vo + vl, acc);
                        else
                            spmv_tile_f32<float>(M, lt, Q.A32, Q.z32 + vo, Q.z32 + vo + vl, Q.Az32 + vo,
                                                 Q.Az32 + vo + vl, acc);
                        wrote = true;
                    }
                    if (wrote) {
                        for (int off = WAVE / 2; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
                        if (threadIdx.x % WAVE == 0) {
                            double* p = Q.part_mat + (size_t)tile * 2;
                            p[0] = acc.hi;
                            p[1] = acc.lo;
                        }
                    }
                }
            }
        }
        grid_barrier(Q.bar, G, P.spin_limit, Q.err_flag);
        // ------------------------------------------------ column phase
        {
            unsigned* qc = Q.queue + (step * 2 + 1) % 3;
            double* pc_old = Q.part_col + (size_t)cur * P.ntile_col * 4;
            double* pc_new = Q.part_col + (size_t)nxt * P.ntile_col * 4;
            while (true) {
                if (threadIdx.x == 0) s_unit = (int)atomicAdd(qc, 1u);
                __syncthreads();
                const int tile = s_unit;
                __syncthreads();
                if (tile >= P.ntile_col) break;
                const int m = Q.tile_mem_col[tile];
                const EnsMember& M = Q.mem[m];
                const Dims& d = M.d;
                const EnsState st = S0[m];
                const int lt = tile - M.t0_col;
                const bool lead = (lt == 0);
                // member-wide scalars from the previous phases (fixed-order reductions, wave 0)
                if (threadIdx.x < WAVE) {
                    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
                    if (st.phase == M_OUTER) {
                        v0 = member_sum(Q.part_mat, M.t0_mat, M.ntile_mat, 2, 0);    // (r,r) of b - Ax
                    } else if (st.phase == M_INNER) {
                        v0 = member_sum(pc_old, M.t0_col, M.ntile_col, 4, 0);        // (r,z)
                        v1 = member_sum(Q.part_mat, M.t0_mat, M.ntile_mat, 2, 0);    // (Az,z)
                        v2 = member_sum(pc_old, M.t0_col, M.ntile_col, 4, 1);        // (r,r)
                    }
                    if (threadIdx.x == 0) { s_val[0] = v0; s_val[1] = v1; s_val[2] = v2; }
                }
                __syncthreads();
                int i, j;
                long long t2;
                col_of_tile(M, lt, i, j, t2);
                const bool incol = t2 < (long long)d.nvx * d.nvy;
                const bool owned = incol && i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi;
                const long long vo = M.v_off, vl = d.vlen;
                const long long c0 = incol ? vidx(d, 0, i, j) : 0;
                const bool act = incol && Q.active[M.c_off + t2] != 0;
                dd a_rz = {0.0, 0.0}, a_rr = {0.0, 0.0};
                EnsState ns = st;
                if (st.phase == M_OUTER) {
                    const double err = sqrt(s_val[0]) / M.bnorm;
                    ns.err = err;
                    if (!(err == err)) { ns.phase = M_DONE; ns.flags |= 2; }
                    else if (err < P.tol) { ns.phase = M_DONE; ns.flags |= 1; }
                    else if (st.outer_it >= P.max_outer) { ns.phase = M_DONE; }
                    else {
                        // start an inner solve: r32 = r64, e = 0, z = M^-1 r, d = q = 0
                        ns.phase = M_INNER;
                        ns.inner_it = 0;
                        ns.outer_it = st.outer_it + 1;
                        if (incol) {
                            float rU[NZMAX], rV[NZMAX], zU[NZMAX], zV[NZMAX];
#pragma unroll
                            for (int k = 0; k < NZMAX; ++k)
                                if (k < d.nz) {
                                    const long long c = c0 + (long long)k * d.plane;
                                    rU[k] = (float)(owned ? Q.r64[vo + c] : 0.0);
                                    rV[k] = (float)(owned ? Q.r64[vo + vl + c] : 0.0);
                                }
                            if (act) {
                                const float* Mm = Q.M32 + 3 * vo;
                                sia_col<NZMAX>(d, Mm, Mm + vl, Mm + 2 * vl, c0, rU, zU);
                                sia_col<NZMAX>(d, Mm + 3 * vl, Mm + 4 * vl, Mm + 5 * vl, c0, rV, zV);
                            } else {
#pragma unroll
                                for (int k = 0; k < NZMAX; ++k) { zU[k] = 0.f; zV[k] = 0.f; }
                            }
#pragma unroll
                            for (int k = 0; k < NZMAX; ++k)
                                if (k < d.nz) {
                                    const long long c = c0 + (long long)k * d.plane;
                                    Q.r32[vo + c] = rU[k]; Q.r32[vo + vl + c] = rV[k];
                                    Q.z32[vo + c] = zU[k]; Q.z32[vo + vl + c] = zV[k];
                                    Q.e32[vo + c] = 0.f; Q.e32[vo + vl + c] = 0.f;
                                    Q.d32[vo + c] = 0.f; Q.d32[vo + vl + c] = 0.f;
                                    Q.q32[vo + c] = 0.f; Q.q32[vo + vl + c] = 0.f;
                                    if (owned) {
                                        a_rz = dd_add_d(a_rz, (double)rU[k] * (double)zU[k]);
                                        a_rz = dd_add_d(a_rz, (double)rV[k] * (double)zV[k]);
                                        a_rr = dd_add_d(a_rr, (double)rU[k] * (double)rU[k]);
                                        a_rr = dd_add_d(a_rr, (double)rV[k] * (double)rV[k]);
                                    }
                                }
                        }
                    }
                } else if (st.phase == M_INNER) {
                    const double rho = s_val[0], delta = s_val[1];
                    double beta = 0.0, sigma = delta;
                    if (st.inner_it > 0) {
                        beta = rho / st.rho;
                        sigma = delta - beta * rho / st.alpha;
                    }
                    const double alpha = rho / sigma;
                    if (st.inner_it == 0) ns.r0 = sqrt(s_val[2]);
                    ns.rho = rho;
                    ns.alpha = alpha;
                    ns.inner_it = st.inner_it + 1;
                    ns.total_inner = st.total_inner + 1;
                    if (!(alpha == alpha)) { ns.phase = M_DONE; ns.flags |= 2; }
                    if (incol && ns.phase == M_INNER) {
                        const float fa = (float)alpha, fb = (float)beta;
                        float rU[NZMAX], rV[NZMAX], zU[NZMAX], zV[NZMAX];
#pragma unroll
                        for (int k = 0; k < NZMAX; ++k)
                            if (k < d.nz) {
                                const long long c = vo + c0 + (long long)k * d.plane;
                                const float dU = Q.z32[c] + fb * Q.d32[c], dV = Q.z32[c + vl] + fb * Q.d32[c + vl];
                                const float qU = Q.Az32[c] + fb * Q.q32[c], qV = Q.Az32[c + vl] + fb * Q.q32[c + vl];
                                Q.d32[c] = dU; Q.d32[c + vl] = dV; Q.q32[c] = qU; Q.q32[c + vl] = qV;
                                Q.e32[c] = Q.e32[c] + fa * dU; Q.e32[c + vl] = Q.e32[c + vl] + fa * dV;
                                rU[k] = Q.r32[c] - fa * qU; rV[k] = Q.r32[c + vl] - fa * qV;
                                Q.r32[c] = rU[k]; Q.r32[c + vl] = rV[k];
                            }
                        if (act) {
                            const float* Mm = Q.M32 + 3 * vo;
                            sia_col<NZMAX>(d, Mm, Mm + vl, Mm + 2 * vl, c0, rU, zU);
                            sia_col<NZMAX>(d, Mm + 3 * vl, Mm + 4 * vl, Mm + 5 * vl, c0, rV, zV);
#pragma unroll
                            for (int k = 0; k < NZMAX; ++k)
                                if (k < d.nz) {
                                    const long long c = vo + c0 + (long long)k * d.plane;
                                    Q.z32[c] = zU[k]; Q.z32[c + vl] = zV[k];
                                }
                        } else {
#pragma unroll
                            for (int k = 0; k < NZMAX; ++k) { zU[k] = 0.f; zV[k] = 0.f; }
                        }
                        if (owned)
#pragma unroll
                            for (int k = 0; k < NZMAX; ++k)
                                if (k < d.nz) {
                                    a_rz = dd_add_d(a_rz, (double)rU[k] * (double)zU[k]);
                                    a_rz = dd_add_d(a_rz, (double)rV[k] * (double)zV[k]);
                                    a_rr = dd_add_d(a_rr, (double)rU[k] * (double)rU[k]);
                                    a_rr = dd_add_d(a_rr, (double)rV[k] * (double)rV[k]);
                                }
                    }
                    // inner convergence is decided on the next step from the new (r,r)
                    ns.phase = (ns.phase == M_DONE) ? M_DONE : M_INNER;
                } else if (st.phase == M_FOLD) {
                    // x64 += e; then the next outer residual
                    if (incol)
                        for (int k = 0; k < d.nz; ++k) {
                            const long long c = vo + c0 + (long long)k * d.plane;
                            Q.x64[c] = Q.x64[c] + (double)Q.e32[c];
                            Q.x64[c + vl] = Q.x64[c + vl] + (double)Q.e32[c + vl];
                        }
                    ns.phase = M_OUTER;
                }
                // column partials of this tile
                if (st.phase == M_OUTER || st.phase == M_INNER) {
                    dd v[2] = {a_rz, a_rr};
                    block_reduce_dd<2, ENS_BS>(v);
                    if (threadIdx.x == 0) {
                        double* p = pc_new + (size_t)tile * 4;
                        p[0] = v[0].hi; p[1] = v[0].lo; p[2] = v[1].hi; p[3] = v[1].lo;
                    }
                }
                if (lead && threadIdx.x == 0) S1[m] = ns;
                __syncthreads();
            }
        }
        grid_barrier(Q.bar, G, P.spin_limit, Q.err_flag);
        // ------------------------------------------------ inner convergence (same decision in every workgroup)
        // The lead workgroup of each member turns INNER into FOLD when the new
        // residual reached inner_tol (checked every ncheck steps) or the
        // inner iteration limit.  Decided here, before the next matrix phase.
        for (int m = blockIdx.x; m < P.nmem; m += G) {
            EnsState* s = Q.st + (size_t)nxt * P.nmem + m;
            const int ph = s->phase;
            __syncthreads();                 // every wave has read the phase before thread 0 may change it
            if (ph != M_INNER) continue;     // uniform across the workgroup
            const EnsMember& M = Q.mem[m];
            double rr = 0.0;
            if (threadIdx.x < WAVE) {
                double* pc_new = Q.part_col + (size_t)nxt * P.ntile_col * 4;
                rr = member_sum(pc_new, M.t0_col, M.ntile_col, 4, 1);
            }
            if (threadIdx.x == 0) {
                const bool chk = s->inner_it > 0 && ((s->inner_it % P.ncheck == 0) || s->inner_it >= P.inner_maxiters);
                const double ie = sqrt(rr) / s->r0;
                if (chk && (ie < P.inner_tol || s->inner_it >= P.inner_maxiters || !(ie == ie))) s->phase = M_FOLD;
            }
        }
        grid_barrier(Q.bar, G, P.spin_limit, Q.err_flag);
        if (*Q.err_flag) break;
    }
}

// ---------------------------------------------------------------- build helpers
// member row-blocked matrix (its own Dims) -> compact ensemble rows
template <typename TA>
__global__ void k_ens_pack(Dims d, long long a_off, int ncolp, const double* __restrict__ src, TA* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;   // over compact rows * 108
    const long long nrow = (long long)d.nz * ncolp;
    if (t >= nrow * NCOEF) return;
    const int cf = (int)(t % NCOEF);
    const long long rr = t / NCOEF;
    const int k = (int)(rr / ncolp), c = (int)(rr % ncolp);
    double v = 0.0;
    if (c < d.nxo * d.nyo) {
        const int io = c % d.nxo, jo = c / d.nxo;
        const long long r = ((long long)k * d.nyo + jo) * d.nxp + io;
        v = src[((r / WAVE) * NCOEF + cf) * WAVE + (r % WAVE)];
    }
    const long long R = rr + a_off;
    dst[((R / WAVE) * NCOEF + cf) * WAVE + (R % WAVE)] = (TA)v;
}

template <>
__global__ void k_ens_pack<__hip_bfloat16>(Dims d, long long a_off, int ncolp, const double* __restrict__ src,
                                           __hip_bfloat16* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long nrow = (long long)d.nz * ncolp;
    if (t >= nrow * NCOEF) return;
    const int cf = (int)(t % NCOEF);
    const long long rr = t / NCOEF;
    const int k = (int)(rr / ncolp), c = (int)(rr % ncolp);
    double v = 0.0;
    if (c < d.nxo * d.nyo) {
        const int io = c % d.nxo, jo = c / d.nxo;
        const long long r = ((long long)k * d.nyo + jo) * d.nxp + io;
        v = src[((r / WAVE) * NCOEF + cf) * WAVE + (r % WAVE)];
    }
    const long long R = rr + a_off;
    dst[((R / WAVE) * NCOEF + cf) * WAVE + (R % WAVE)] = __float2bfloat16((float)v);
}

// Stand-alone FP32 SpMV sweep over the members of a bucket (autotuner).
__global__ __launch_bounds__(ENS_BS) void k_ens_spmv_sweep(const EnsMember* __restrict__ mem,
                                                           const int* __restrict__ tiles, int ntiles,
                                                           const int* __restrict__ tile_mem, const float* __restrict__ A,
                                                           const float* __restrict__ z, float* __restrict__ Az,
                                                           double* __restrict__ part) {
    const int u = blockIdx.x * ENS_WPB + threadIdx.x / WAVE;
    if (u >= ntiles) return;
    const int tile = tiles[u];
    const EnsMember& M = mem[tile_mem[tile]];
    dd acc = {0.0, 0.0};
    const long long vo = M.v_off, vl = M.d.vlen;
    spmv_tile_f32<float>(M, tile - M.t0_mat, A, z + vo, z + vo + vl, Az + vo, Az + vo + vl, acc);
    for (int off = WAVE / 2; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
    if (threadIdx.x % WAVE == 0) { part[2 * tile] = acc.hi; part[2 * tile + 1] = acc.lo; }
}

}  // namespace cg
