// ensemble.hip -- batched solves of many independent velocity systems in one
// persistent launch.
//
// Each member is an assembled FP64 system (owned rows of A, b, the SIA
// column-tridiagonal preconditioner).  The solve is FP64 iterative
// refinement around an FP32 Chronopoulos-Gear inner solve (the recurrence of
// pcg_solver_chrongear_3d, glissade_velo_higher_pcg.F90:3296-3581, with
// beta = 0 and sigma = delta on the first inner step).
//
// All members advance in lock step through "steps".  A step is a matrix
// phase (per member: the FP64 residual b - A x, or the FP32 product A z) and
// a column phase (per member: refinement bookkeeping, or the fused C-G
// update + SIA column solve + partial sums), each closed by a grid-wide
// barrier.  Work is split into wavefront tiles -- 64 owned columns (a share
// of the levels) in the matrix phase, 64 vertex columns in the column phase --
// dealt round-robin over every wavefront of the grid, so members that are
// still iterating are spread over all CUs while finished members cost one
// state read per tile.  Every per-member decision is a deterministic
// function of per-tile compensated partial sums reduced in a fixed order,
// so all wavefronts reach the same decision without extra communication.
//
// Members must have an ice-free outer ring (no active vertex on a row that
// the halo rule writes), so halo updates are identities and are omitted;
// the host checks this when the ensemble is built.
#include "cismgpu_internal.h"

namespace cg {

constexpr int ENS_BS = 1024;         // threads per workgroup (16 wavefronts, one workgroup per CU)
constexpr int ENS_WPB = ENS_BS / WAVE;

struct EnsMember {
    Dims d;                          // member grid (vector layout of the member)
    long long a_off;                 // offset (rows) of the member's compact matrix rows
    long long v_off;                 // offset (elements) of the member's vectors
    long long c_off;                 // offset of the member's vertex-column arrays (nvx*nvy)
    int ncol, ncolp;                 // owned columns nxo*nyo and padded to 64
    int ncg;                         // groups of 64 owned columns
    int split;                       // level groups per column group (1, 2, 4): the tuned parameter
    int ntile_mat;                   // matrix-phase tiles = ncg * split
    int ntile_col;                   // column-phase tiles (64 vertex columns each)
    int t0_mat, t0_col;              // first tile indices in the global lists
    double bnorm;                    // ||b|| (FP64, compensated), set at build
};

// per-member state, double buffered by step parity
enum { M_OUTER = 0, M_INNER = 1, M_FOLD = 2, M_DONE = 3 };
struct EnsState {
    int phase;                       // M_OUTER, M_INNER or M_DONE (M_FOLD is decided at the start of a step)
    int inner_it;                    // inner iterations of the current refinement step
    int outer_it;                    // refinement steps done
    int total_inner;
    double rho, alpha;               // C-G scalars (rho_old, alpha_old)
    double r0;                       // ||r|| at the start of the inner solve
    double err;                      // last outer relative residual
    int flags;                       // 1 converged, 2 NaN
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
    const float* S32;                // half (symmetric) storage of A32: centre + 13 upper couplings
    int half;                        // 1: the FP32 products read S32, 0: A32
    const float* M32;                // SIA coefficients per member vector layout (6 fields)
    const int8_t* active;
    const double* b64;               // 2 components per member block
    double* x64;
    double* r64;
    float *r32, *d32, *q32, *z32, *Az32, *e32;
    double* part_mat;                // [ntile_mat] dd
    double* part_col;                // [2][ntile_col][2] dd
    EnsState* st;                    // [2][nmem]
    unsigned* bar;                   // barrier: count, generation
    int* err_flag;                   // set when a barrier wait exceeded its bound
    int* steps_out;                  // steps taken (written by workgroup 0)
    int use_bf16;
};

// ---------------------------------------------------------------- barrier
// Sense-reversing barrier for a grid whose workgroups are all resident (the
// launch is sized from the occupancy query).  Release before arriving, acquire
// after the generation flips.  Every wait is bounded: after `limit` polls the
// workgroup records an error and proceeds, so the grid always drains.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned nblocks, long long limit, int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == nblocks) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long polls = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (++polls > limit) { __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Deterministic reduction of a member's tile partials (stride = doubles per
// tile) by one wavefront; every lane returns the sum.
__device__ double member_sum(const double* part, int t0, int n, int stride, int which) {
    dd acc = {0.0, 0.0};
    const int lane = threadIdx.x % WAVE;
    for (int t = lane; t < n; t += WAVE) {
        const double* p = part + (size_t)(t0 + t) * stride + 2 * which;
        acc = dd_add(acc, dd{p[0], p[1]});
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
    return __shfl(acc.hi + acc.lo, 0, WAVE);
}

__device__ __forceinline__ dd wave_reduce(dd acc) {
    for (int off = WAVE / 2; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
    return acc;
}

// Phase a member runs in this step.  An inner solve turns into a fold once its
// residual reached inner_tol (tested every ncheck iterations) or at the
// iteration limit; the test reads the (r,r) partials of the previous step, so
// every wavefront evaluates it identically.
__device__ __forceinline__ int step_phase(const EnsState& st, const EnsMember& M, const double* pc_cur,
                                          const EnsParams& P) {
    if (st.phase != M_INNER || st.inner_it == 0) return st.phase;
    if (st.inner_it % P.ncheck != 0 && st.inner_it < P.inner_maxiters) return M_INNER;
    const double rr = member_sum(pc_cur, M.t0_col, M.ntile_col, 4, 1);
    const double ie = sqrt(rr) / st.r0;
    return (ie < P.inner_tol || st.inner_it >= P.inner_maxiters || !(ie == ie)) ? M_FOLD : M_INNER;
}

// ---------------------------------------------------------------- matrix phase work
// One wavefront: 64 consecutive owned columns, the levels of level group lg.
// Row (k, c) of a member sits at compact row k*ncolp + c; the 27 neighbour
// values of one plane are fetched together ahead of their FMAs.
template <typename TA, typename TV, bool RESID>
__device__ __forceinline__ dd matrix_tile(const EnsMember& M, int lt, const TA* __restrict__ A,
                                          const TV* __restrict__ xu, const TV* __restrict__ xv,
                                          const double* __restrict__ bu, const double* __restrict__ bv,
                                          TV* __restrict__ yu, TV* __restrict__ yv) {
    const Dims& d = M.d;
    const int lane = threadIdx.x % WAVE;
    const int cgi = lt % M.ncg, lg = lt / M.ncg;
    const int c = cgi * WAVE + lane;
    dd acc = {0.0, 0.0};
    if (c >= M.ncol) return acc;
    const int io = c % d.nxo, jo = c / d.nxo;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int kper = (d.nz + M.split - 1) / M.split;
    const int k0 = lg * kper, k1 = min(d.nz, k0 + kper);
    const long long pl = d.plane;
    for (int k = k0; k < k1; ++k) {
        const long long r = (long long)k * M.ncolp + c + M.a_off;
        const TA* a = A + ((r / WAVE) * NCOEF) * WAVE + (r % WAVE);
        const long long c0 = vidx(d, k, i, j);
        TV su = 0, sv = 0;
#pragma unroll
        for (int kA = 0; kA < 3; ++kA) {
            TV xa[9], xb[9], cuu[9], cuv[9], cvu[9], cvv[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const long long n = c0 + (kA - 1) * pl + (q / 3 - 1) * d.px + (q % 3 - 1);
                xa[q] = xu[n];
                xb[q] = xv[n];
                const int m = kA * 9 + q;
                cuu[q] = ld_coef(a + (0 * 27 + m) * WAVE);
                cuv[q] = ld_coef(a + (1 * 27 + m) * WAVE);
                cvu[q] = ld_coef(a + (2 * 27 + m) * WAVE);
                cvv[q] = ld_coef(a + (3 * 27 + m) * WAVE);
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                su = su + cuu[q] * xa[q] + cuv[q] * xb[q];
                sv = sv + cvu[q] * xa[q] + cvv[q] * xb[q];
            }
        }
        if (RESID) {
            const TV tu = (TV)bu[c0] - su, tv = (TV)bv[c0] - sv;
            yu[c0] = tu;
            yv[c0] = tv;
            acc = dd_add_d(acc, (double)tu * (double)tu);
            acc = dd_add_d(acc, (double)tv * (double)tv);
        } else {
            yu[c0] = su;
            yv[c0] = sv;
            acc = dd_add_d(acc, (double)su * (double)xu[c0]);
            acc = dd_add_d(acc, (double)sv * (double)xv[c0]);
        }
    }
    return acc;
}

// FP32 product with the half (symmetric) storage: row R holds its centre
// (uu, uv, vv, 0) and 13 upper couplings as quads at
//     S[((R/64)*14 + s)*256 + (R%64)*4 + c];
// the coupling to a lower neighbour (m < 13) is the transposed slot 13 - m
// of the neighbour's row.  Members keep an ice-free outer ring, so a
// neighbour outside the owned columns couples with zero.  The products are
// summed in the order and with the values of matrix_tile, so the results are
// identical when the FP32 matrix is exactly symmetric.
__device__ __forceinline__ dd matrix_tile_half(const EnsMember& M, int lt, const float* __restrict__ S,
                                               const float* __restrict__ xu, const float* __restrict__ xv,
                                               float* __restrict__ yu, float* __restrict__ yv) {
    const Dims& d = M.d;
    const int lane = threadIdx.x % WAVE;
    const int cgi = lt % M.ncg, lg = lt / M.ncg;
    const int c = cgi * WAVE + lane;
    dd acc = {0.0, 0.0};
    if (c >= M.ncol) return acc;
    const int io = c % d.nxo, jo = c / d.nxo;
    const int i = d.ilo + io, j = d.jlo + jo;
    const int kper = (d.nz + M.split - 1) / M.split;
    const int k0 = lg * kper, k1 = min(d.nz, k0 + kper);
    const long long pl = d.plane;
    // bit q: horizontal neighbour q (0..8) lies inside the owned columns
    unsigned inb = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int pi = io + q % 3 - 1, pj = jo + q / 3 - 1;
        inb |= (pi >= 0 && pi < d.nxo && pj >= 0 && pj < d.nyo) ? (1u << q) : 0u;
    }
    // raw buffer over this member's rows: 32-bit member-local byte offsets
    // (the host keeps every member's part below 2 GiB), loads past it read zero
    const float* Sm = S + (size_t)(M.a_off / WAVE) * (14 * 4 * WAVE);
    const int mbytes = d.nz * M.ncolp * 56 * (int)sizeof(float);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)Sm, (short)0, mbytes, 0x00020000);
    auto quad = [&](unsigned row, int sl, float& a, float& b, float& e, float& f) {
        const unsigned off = (((row / WAVE) * 14u + (unsigned)sl) * (4u * WAVE) + (row % WAVE) * 4u) * 4u;
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        a = __uint_as_float(v[0]); b = __uint_as_float(v[1]); e = __uint_as_float(v[2]); f = __uint_as_float(v[3]);
    };
    for (int k = k0; k < k1; ++k) {
        const int R = k * M.ncolp + c;     // member-local row
        const long long c0 = vidx(d, k, i, j);
        float su = 0.f, sv = 0.f;
        // rows of three neighbours (fixed kA, jA), unrolled by three: more
        // loads in flight would spill past the kernel's 128 VGPRs
#pragma unroll 3
        for (int rw = 0; rw < 9; ++rw) {
            const int kA = rw / 3, jq = rw % 3;
            float xa[3], xb[3], cuu[3], cuv[3], cvu[3], cvv[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int q = jq * 3 + t, m = kA * 9 + q;
                const long long n = c0 + (kA - 1) * pl + (jq - 1) * d.px + (t - 1);
                xa[t] = xu[n];
                xb[t] = xv[n];
                if (m < 13) {
                    cuu[t] = cuv[t] = cvu[t] = cvv[t] = 0.f;
                    if (((inb >> q) & 1u) && k + kA - 1 >= 0) {
                        const int P = R + (kA - 1) * M.ncolp + (jq - 1) * d.nxo + (t - 1);
                        float a, b, e, f;
                        quad((unsigned)P, 13 - m, a, b, e, f);
                        cuu[t] = a; cuv[t] = e; cvu[t] = b; cvv[t] = f;     // transposed
                    }
                } else {
                    const int sl = m == 13 ? 0 : m - 13;
                    float a, b, e, f;
                    quad((unsigned)R, sl, a, b, e, f);
                    if (m == 13) { cuu[t] = a; cuv[t] = b; cvu[t] = b; cvv[t] = e; }
                    else { cuu[t] = a; cuv[t] = b; cvu[t] = e; cvv[t] = f; }
                }
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                su = su + cuu[t] * xa[t] + cuv[t] * xb[t];
                sv = sv + cvu[t] * xa[t] + cvv[t] * xb[t];
            }
        }
        yu[c0] = su;
        yv[c0] = sv;
        acc = dd_add_d(acc, (double)su * (double)xu[c0]);
        acc = dd_add_d(acc, (double)sv * (double)xv[c0]);
    }
    return acc;
}

// ---------------------------------------------------------------- column phase work
// One vertex column, one velocity component: the C-G vector update (START:
// load r from the FP64 residual and clear e, d, q) followed by the SIA
// Thomas solve z = M^-1 r (easy_sia_solver), streamed level by level; the
// forward sweep leaves the partial solution in z, the back substitution
// finishes it.  Accumulates (r,z) and (r,r) over owned columns.
template <int NZMAX, bool START>
__device__ __forceinline__ void column_component(const Dims& d, const EnsPtrs& Q, long long vb,
                                                 const float* __restrict__ Mm, long long c0, bool owned, bool act,
                                                 float fa, float fb, dd& a_rz, dd& a_rr) {
    const long long pl = d.plane;
    const float *Ms = Mm, *Md = Mm + d.vlen, *Mp = Mm + 2 * d.vlen;
    float gam[NZMAX];
    float bet = 1.f, zprev = 0.f;
#pragma unroll
    for (int k = 0; k < NZMAX; ++k)
        if (k < d.nz) {
            const long long cl = c0 + (long long)k * pl;
            const long long c = vb + cl;
            float rk;
            if (START) {
                rk = owned ? (float)Q.r64[c] : 0.f;
                Q.e32[c] = 0.f;
                Q.d32[c] = 0.f;
                Q.q32[c] = 0.f;
            } else {
                const float dk = Q.z32[c] + fb * Q.d32[c];
                const float qk = Q.Az32[c] + fb * Q.q32[c];
                Q.d32[c] = dk;
                Q.q32[c] = qk;
                Q.e32[c] = Q.e32[c] + fa * dk;
                rk = Q.r32[c] - fa * qk;
            }
            Q.r32[c] = rk;
            if (act) {
                float zt;
                if (k == 0) {
                    bet = Md[cl];
                    zt = rk / bet;
                } else {
                    gam[k] = Mp[cl - pl] / bet;
                    const float sb = Ms[cl];
                    bet = Md[cl] - sb * gam[k];
                    zt = (rk - sb * zprev) / bet;
                }
                Q.z32[c] = zt;
                zprev = zt;
            } else if (START) {
                Q.z32[c] = 0.f;
            }
            if (owned) a_rr = dd_add_d(a_rr, (double)rk * (double)rk);
        }
    if (act) {
        float znext = zprev;
        if (owned) a_rz = dd_add_d(a_rz, (double)Q.r32[vb + c0 + (long long)(d.nz - 1) * pl] * (double)znext);
#pragma unroll
        for (int k = NZMAX - 2; k >= 0; --k)
            if (k < d.nz - 1) {
                const long long c = vb + c0 + (long long)k * pl;
                const float zk = Q.z32[c] - gam[k + 1] * znext;
                Q.z32[c] = zk;
                znext = zk;
                if (owned) a_rz = dd_add_d(a_rz, (double)Q.r32[c] * (double)zk);
            }
    }
}

template <int NZMAX>
__global__ __launch_bounds__(ENS_BS) void k_ensemble(EnsParams P, EnsPtrs Q) {
    const unsigned G = gridDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int gw = blockIdx.x * ENS_WPB + wave;          // global wavefront id
    const int W = G * ENS_WPB;
    const int lane = threadIdx.x % WAVE;
    __shared__ int s_any;
    int step = 0;
    for (; step < P.maxsteps; ++step) {
        const int cur = step & 1, nxt = cur ^ 1;
        const EnsState* S0 = Q.st + (size_t)cur * P.nmem;
        EnsState* S1 = Q.st + (size_t)nxt * P.nmem;
        const double* pc_cur = Q.part_col + (size_t)cur * P.ntile_col * 4;
        double* pc_nxt = Q.part_col + (size_t)nxt * P.ntile_col * 4;
        // all members done?  (every workgroup reads the same states)
        if (threadIdx.x == 0) s_any = 0;
        __syncthreads();
        for (int m = threadIdx.x; m < P.nmem; m += ENS_BS)
            if (S0[m].phase != M_DONE) s_any = 1;
        __syncthreads();
        if (!s_any) break;
        // ------------------------------------------------ matrix phase
        for (int tile = gw; tile < P.ntile_mat; tile += W) {
            const int m = Q.tile_mem_mat[tile];
            const EnsMember& M = Q.mem[m];
            const EnsState st = S0[m];
            if (st.phase == M_DONE) continue;
            const int ph = step_phase(st, M, pc_cur, P);
            if (ph != M_OUTER && ph != M_INNER) continue;
            const long long vo = M.v_off, vl = M.d.vlen;
            const int lt = tile - M.t0_mat;
            dd acc;
            if (ph == M_OUTER)
                acc = matrix_tile<double, double, true>(M, lt, Q.A64, Q.x64 + vo, Q.x64 + vo + vl, Q.b64 + vo,
                                                        Q.b64 + vo + vl, Q.r64 + vo, Q.r64 + vo + vl);
            else if (Q.use_bf16)
                acc = matrix_tile<__hip_bfloat16, float, false>(M, lt, Q.A16, Q.z32 + vo, Q.z32 + vo + vl, nullptr,
                                                                nullptr, Q.Az32 + vo, Q.Az32 + vo + vl);
            else if (Q.half)
                acc = matrix_tile_half(M, lt, Q.S32, Q.z32 + vo, Q.z32 + vo + vl, Q.Az32 + vo, Q.Az32 + vo + vl);
            else
                acc = matrix_tile<float, float, false>(M, lt, Q.A32, Q.z32 + vo, Q.z32 + vo + vl, nullptr, nullptr,
                                                       Q.Az32 + vo, Q.Az32 + vo + vl);
            acc = wave_reduce(acc);
            if (lane == 0) {
                Q.part_mat[2 * (size_t)tile] = acc.hi;
                Q.part_mat[2 * (size_t)tile + 1] = acc.lo;
            }
        }
        grid_barrier(Q.bar, G, P.spin_limit, Q.err_flag);
        // ------------------------------------------------ column phase
        for (int tile = gw; tile < P.ntile_col; tile += W) {
            const int m = Q.tile_mem_col[tile];
            const EnsMember& M = Q.mem[m];
            const Dims& d = M.d;
            const EnsState st = S0[m];
            const bool lead = (tile == M.t0_col);
            if (st.phase == M_DONE) {
                if (lead && lane == 0) S1[m] = st;
                continue;
            }
            const int ph = step_phase(st, M, pc_cur, P);
            const int lt = tile - M.t0_col;
            const long long t = (long long)lt * WAVE + lane;
            const bool incol = t < (long long)d.nvx * d.nvy;
            const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
            const bool owned = incol && i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi;
            const long long vo = M.v_off, vl = d.vlen;
            const long long c0 = incol ? vidx(d, 0, i, j) : 0;
            const bool act = incol && Q.active[M.c_off + t] != 0;
            const float* Mm = Q.M32 + 3 * vo;
            dd a_rz = {0.0, 0.0}, a_rr = {0.0, 0.0};
            EnsState ns = st;
            bool partials = false;
            if (ph == M_OUTER) {
                const double err = sqrt(member_sum(Q.part_mat, M.t0_mat, M.ntile_mat, 2, 0)) / M.bnorm;
                ns.err = err;
                if (!(err == err)) { ns.phase = M_DONE; ns.flags |= 2; }
                else if (err < P.tol) { ns.phase = M_DONE; ns.flags |= 1; }
                else if (st.outer_it >= P.max_outer) { ns.phase = M_DONE; }
                else {
                    // start an inner solve: r32 = r64, e = d = q = 0, z = M^-1 r
                    ns.phase = M_INNER;
                    ns.inner_it = 0;
                    ns.outer_it = st.outer_it + 1;
                    partials = true;
                    if (incol) {
                        column_component<NZMAX, true>(d, Q, vo, Mm, c0, owned, act, 0.f, 0.f, a_rz, a_rr);
                        column_component<NZMAX, true>(d, Q, vo + vl, Mm + 3 * vl, c0, owned, act, 0.f, 0.f, a_rz,
                                                      a_rr);
                    }
                }
            } else if (ph == M_INNER) {
                const double rho = member_sum(pc_cur, M.t0_col, M.ntile_col, 4, 0);
                const double delta = member_sum(Q.part_mat, M.t0_mat, M.ntile_mat, 2, 0);
                double beta = 0.0, sigma = delta;
                if (st.inner_it > 0) {
                    beta = rho / st.rho;
                    sigma = delta - beta * rho / st.alpha;
                } else {
                    ns.r0 = sqrt(member_sum(pc_cur, M.t0_col, M.ntile_col, 4, 1));
                }
                const double alpha = rho / sigma;
                ns.rho = rho;
                ns.alpha = alpha;
                ns.inner_it = st.inner_it + 1;
                ns.total_inner = st.total_inner + 1;
                partials = true;
                if (!(alpha == alpha)) {
                    ns.phase = M_DONE;
                    ns.flags |= 2;
                } else if (incol) {
                    const float fa = (float)alpha, fb = (float)beta;
                    column_component<NZMAX, false>(d, Q, vo, Mm, c0, owned, act, fa, fb, a_rz, a_rr);
                    column_component<NZMAX, false>(d, Q, vo + vl, Mm + 3 * vl, c0, owned, act, fa, fb, a_rz, a_rr);
                }
            } else {  // M_FOLD: x64 += e, then the next outer residual
                if (incol)
                    for (int k = 0; k < d.nz; ++k) {
                        const long long c = vo + c0 + (long long)k * d.plane;
                        Q.x64[c] = Q.x64[c] + (double)Q.e32[c];
                        Q.x64[c + vl] = Q.x64[c + vl] + (double)Q.e32[c + vl];
                    }
                ns.phase = M_OUTER;
            }
            if (partials) {
                a_rz = wave_reduce(a_rz);
                a_rr = wave_reduce(a_rr);
                if (lane == 0) {
                    double* p = pc_nxt + (size_t)tile * 4;
                    p[0] = a_rz.hi; p[1] = a_rz.lo; p[2] = a_rr.hi; p[3] = a_rr.lo;
                }
            }
            if (lead && lane == 0) S1[m] = ns;
        }
        grid_barrier(Q.bar, G, P.spin_limit, Q.err_flag);
        if (__hip_atomic_load(Q.err_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *Q.steps_out = step;
}

// ---------------------------------------------------------------- build helpers
// member row-blocked matrix (its own Dims) -> compact ensemble rows
template <typename TA>
__device__ __forceinline__ TA ens_cvt(double v) { return (TA)v; }
template <>
__device__ __forceinline__ __hip_bfloat16 ens_cvt<__hip_bfloat16>(double v) { return __float2bfloat16((float)v); }

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
    dst[((R / WAVE) * NCOEF + cf) * WAVE + (R % WAVE)] = ens_cvt<TA>(v);
}

// Half-storage rows of one member (compact rows, as k_ens_pack) from its
// FP64 row-blocked matrix, rounded to FP32.
__global__ void k_ens_pack_half(Dims d, long long a_off, int ncolp, const double* __restrict__ src,
                                float* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long nrow = (long long)d.nz * ncolp;
    if (t >= nrow * 56) return;
    const int cf = (int)(t % 56), sl = cf / 4, c4 = cf % 4;
    const long long rr = t / 56;
    const int k = (int)(rr / ncolp), c = (int)(rr % ncolp);
    double v = 0.0;
    if (c < d.nxo * d.nyo) {
        const int io = c % d.nxo, jo = c / d.nxo;
        const long long r = ((long long)k * d.nyo + jo) * d.nxp + io;
        const double* row = src + (r / WAVE) * NCOEF * WAVE + (r % WAVE);
        if (sl == 0) {
            const int blk = c4 == 0 ? 0 : (c4 == 1 ? 1 : 3);
            v = c4 == 3 ? 0.0 : row[(blk * 27 + 13) * WAVE];
        } else {
            v = row[(c4 * 27 + 13 + sl) * WAVE];
        }
    }
    const long long R = rr + a_off;
    dst[((R / WAVE) * 14 + sl) * (4 * WAVE) + (R % WAVE) * 4 + c4] = (float)v;
}

// Stand-alone FP32 products over a list of matrix tiles (autotuner).
__global__ __launch_bounds__(ENS_BS) void k_ens_spmv_sweep(const EnsMember* __restrict__ mem,
                                                           const int* __restrict__ tiles, int ntiles,
                                                           const int* __restrict__ tile_mem, const float* __restrict__ A,
                                                           const float* __restrict__ S32, int half, const float* __restrict__ z,
                                                           float* __restrict__ Az, double* __restrict__ part) {
    const int u = blockIdx.x * ENS_WPB + __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    if (u >= ntiles) return;
    const int tile = tiles[u];
    const EnsMember& M = mem[tile_mem[tile]];
    const long long vo = M.v_off, vl = M.d.vlen;
    dd acc = half ? matrix_tile_half(M, tile - M.t0_mat, S32, z + vo, z + vo + vl, Az + vo, Az + vo + vl)
                     : matrix_tile<float, float, false>(M, tile - M.t0_mat, A, z + vo, z + vo + vl, nullptr, nullptr,
                                                        Az + vo, Az + vo + vl);
    acc = wave_reduce(acc);
    if (threadIdx.x % WAVE == 0) { part[2 * (size_t)tile] = acc.hi; part[2 * (size_t)tile + 1] = acc.lo; }
}

}  // namespace cg
