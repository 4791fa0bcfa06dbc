// spmv_sym.hip -- half-storage SpMV for the inner FP32 iterations.
//
// The assembled operator is symmetric: Auu = Auu^T, Avv = Avv^T, Auv = Avu^T
// (glissade_velo_higher.F90:10859-10864), so of the 27 neighbour couplings
// of a node only the centre and the 13 "upper" ones (m = 14..27 in the
// reference numbering, 0-based 13..26) need storing; the coupling of node r
// to its lower neighbour q = r + off(m), m < 13, is the transposed 2x2
// block of q's upper coupling 26 - m.  Storage (per owned row r, slot
// s = m - 13 for m >= 13):
//     Sq[((r/64)*14 + s)*256 + (r%64)*4 + c]
// slot 0 holds (uu, uv, vv, 0) of the centre, slots 1..13 the quads
// (uu, uv, vu, vv) of the upper couplings, so the matrix streams 224 B per
// node instead of 432.
//
// k_spmv_sym walks a column up the levels like k_spmv_q and loads the 13
// lower quads from the partner rows, the neighbouring columns of the same
// level or of the previous level; k_spmv_symx (the default) takes the nine
// previous-level quads from the tile's own threads through LDS and lane
// shuffles instead (0.651 against 0.669 ms at 1024^2 x 11 on MI355X).  The
// products are summed in the same order (m = 0..26) with the same values as
// the full-storage kernel, so when the FP32 matrix is exactly symmetric
// (k_check_sym finds no mismatch) all three kernels give identical results;
// the solver uses this layout only then.  x is staged in LDS per 64 x 4
// tile with a one-vertex ring: a 3-level ring of slots in k_spmv_sym and
// k_spmv_sym64, a 4-level ring in k_spmv_symx (so that one barrier per
// level less is needed).  A 3-level x window in registers instead cost 54
// VGPRs and limited the kernel to 3 waves per SIMD (0.87 ms against 0.73 ms
// at 1024^2 x 11).
//
// Partner offsets within a level are 32-bit (po[] below): the half layouts
// are used only while one level of slots, nyo * nxp * 56 elements, stays
// below 2^31 (half_storage_fits); larger grids keep the full-row kernels.
#include "cismgpu_internal.h"

namespace cg {

constexpr int NSLOT = 14;

inline bool half_storage_fits(const Dims& d) {
    return (long long)d.nyo * d.nxp * NSLOT * 4 < (1LL << 31);
}

// partner column of row (i, j) at horizontal offset (iA, jA): its owned
// image (periodic rules) as jo*nxp + io, or -1 if it has no stored row
__device__ __forceinline__ int partner_col(const Dims& d, int i, int j) {
    if (!row_image(d, i, j)) return -1;
    return (j - d.jlo) * d.nxp + (i - d.ilo);
}

// Half-storage copy of the owned rows of the FP64 row-blocked matrix.
template <typename TA>
__global__ void k_convert_sym(long long nrows, const double* __restrict__ src, TA* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;   // over nrows * 14 * 4
    if (t >= nrows * NSLOT * 4) return;
    const int c = (int)(t % 4);
    const long long u = t / 4;
    const int lane = (int)(u % WAVE);
    const long long v = u / WAVE;
    const int s = (int)(v % NSLOT);
    const long long rb = v / NSLOT;
    const double* row = src + rb * NCOEF * WAVE + lane;
    double val;
    if (s == 0) {   // centre: uu, uv, vv, 0
        const int blk = c == 0 ? 0 : (c == 1 ? 1 : 3);
        val = c == 3 ? 0.0 : row[(blk * 27 + 13) * WAVE];
    } else {
        val = row[(c * 27 + 13 + s) * WAVE];
    }
    dst[t] = to_lowp<TA>(val);
}

// Counts the couplings that the half storage cannot reproduce exactly in
// precision TA (float or double): a lower coupling whose transposed partner differs, a
// nonzero coupling to a node without a stored row, or Auv != Avu at the
// centre.  count[0] += mismatches.
template <typename TA>
__global__ void k_check_sym(Dims d, const double* __restrict__ A, unsigned long long* __restrict__ count) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    unsigned long long bad = 0;
    if (t < d.nrows) {
        const int io = (int)(t % d.nxp);
        const long long kj = t / d.nxp;
        const int jo = (int)(kj % d.nyo), k0 = (int)(kj / d.nyo);
        if (io < d.nxo) {
            const int i = d.ilo + io, j = d.jlo + jo;
            const double* a = A + ((t / WAVE) * NCOEF) * WAVE + (t % WAVE);
            auto rd = [](double x) { return to_lowp<TA>(x); };
            if (rd(a[(1 * 27 + 13) * WAVE]) != rd(a[(2 * 27 + 13) * WAVE])) ++bad;
            for (int m = 0; m < 13; ++m) {
                const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
                if (k0 + kA < 0) continue;                   // stored as zero
                const double c0 = a[(0 * 27 + m) * WAVE], c1 = a[(1 * 27 + m) * WAVE];
                const double c2_ = a[(2 * 27 + m) * WAVE], c3 = a[(3 * 27 + m) * WAVE];
                int qi = i + iA, qj = j + jA;
                if (!row_image(d, qi, qj)) {
                    bad += (c0 != 0.0 || c1 != 0.0 || c2_ != 0.0 || c3 != 0.0) ? 1 : 0;
                    continue;
                }
                const double* q = row_ptr(d, A, k0 + kA, qi, qj);
                const int mm = 26 - m;
                if (rd(c0) != rd(q[(0 * 27 + mm) * WAVE]) || rd(c1) != rd(q[(2 * 27 + mm) * WAVE]) ||
                    rd(c2_) != rd(q[(1 * 27 + mm) * WAVE]) || rd(c3) != rd(q[(3 * 27 + mm) * WAVE]))
                    ++bad;
            }
        }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) bad += __shfl_down(bad, off, WAVE);
    if (threadIdx.x % WAVE == 0 && bad) atomicAdd(count, bad);
}

__device__ __forceinline__ void ld_slot(const float* p, float& a, float& b, float& c, float& e) { ld_quad(p, a, b, c, e); }

template <typename TA, int MODE>
__global__ __launch_bounds__(256) void k_spmv_sym(Dims d, const TA* __restrict__ S, const float* __restrict__ xu,
                                                  const float* __restrict__ xv, float* __restrict__ yu,
                                                  float* __restrict__ yv, const float* __restrict__ bu,
                                                  const float* __restrict__ bv, double* __restrict__ partials) {
    // x of the tile (64 x 4 columns) plus a one-column ring, for levels k-1,
    // k, k+1, staged in LDS (a 3-level ring)
    __shared__ float xs[2][3][6][WAVE + 2];
    const int nbx = d.nxp / WAVE;
    const int nblk = nbx * ((d.nyo + 3) / 4);
    const int lin = xcd_remap(blockIdx.x, nblk);     // neighbouring tiles share an XCD's L2
    const int bx = lin % nbx, by = lin / nbx;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int io = bx * WAVE + lane;
    const int jo = by * 4 + w;
    const bool live = io < d.nxo && jo < d.nyo;
    const int i0 = d.ilo + bx * WAVE - 1, j0 = d.jlo + by * 4 - 1;   // vertex of xs[.][.][0][0]
    const long long pl = d.plane;
    // stage level kk (-1..nz) of x into ring slot kk+1 mod 3
    auto stage = [&](int kk) {
        const int slot = (kk + 1) % 3;
        for (int e = threadIdx.x; e < 6 * (WAVE + 2); e += 256) {
            const int r = e / (WAVE + 2), c = e % (WAVE + 2);
            const int vi = i0 + c, vj = j0 + r;
            float a = 0.f, b = 0.f;
            if (vi >= 1 && vi <= d.nvx && vj >= 1 && vj <= d.nvy) {
                const long long n = vidx(d, kk, vi, vj);
                a = xu[n];
                b = xv[n];
            }
            xs[0][slot][r][c] = a;
            xs[1][slot][r][c] = b;
        }
    };
    double acc = 0.0;
    int po[9];
    const long long LS = (long long)d.nyo * d.nxp / WAVE * (NSLOT * 4 * WAVE);   // elements per level
    if (live) {
        const int i = d.ilo + io, j = d.jlo + jo;
        // element offsets of the partner rows' slot 0 within a level, -1 when
        // the neighbour has no stored row; rows per level are a multiple of 64
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            const int pc = partner_col(d, i + h % 3 - 1, j + h / 3 - 1);
            po[h] = pc < 0 ? -1 : (pc / WAVE) * (NSLOT * 4 * WAVE) + (pc % WAVE) * 4;
        }
    }
    stage(-1);
    stage(0);
    for (int k = 0; k < d.nz; ++k) {
        stage(k + 1);
        __syncthreads();
        if (live) {
            const long long ck = vidx(d, k, d.ilo + io, d.jlo + jo);
            const TA* Sk = S + (long long)k * LS;     // level k
            const TA* Sb = Sk - LS;                   // level k - 1
            float su = 0.f, sv = 0.f;
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                float cuu = 0.f, cuv = 0.f, cvu = 0.f, cvv = 0.f;
                bool have = true;
                if (m < 13) {
                    const int h = m % 9;
                    have = po[h] >= 0 && (m >= 9 || k > 0);
                    if (have) {
                        float a, b, c, e;
                        ld_slot((m < 9 ? Sb : Sk) + po[h] + (13 - m) * (4 * WAVE), a, b, c, e);
                        cuu = a; cuv = c; cvu = b; cvv = e;     // transposed 2x2 block
                    }
                } else if (m == 13) {
                    float a, b, c, e;
                    ld_slot(Sk + po[4], a, b, c, e);
                    cuu = a; cuv = b; cvu = b; cvv = c;
                } else {
                    ld_slot(Sk + po[4] + (m - 13) * (4 * WAVE), cuu, cuv, cvu, cvv);
                }
                if (have) {
                    const int slot = (k + m / 9) % 3;          // level k + kA lives in slot (k + kA + 1) % 3
                    const int r = w + 1 + ((m / 3) % 3 - 1), c = lane + 1 + (m % 3 - 1);
                    const float xun = xs[0][slot][r][c], xvn = xs[1][slot][r][c];
                    su = su + cuu * xun + cuv * xvn;
                    sv = sv + cvu * xun + cvv * xvn;
                }
            }
            if (MODE == 2) {
                const float ru = bu[ck] - su, rv = bv[ck] - sv;
                yu[ck] = ru;
                yv[ck] = rv;
                acc += (double)ru * (double)ru + (double)rv * (double)rv;
            } else {
                yu[ck] = su;
                yv[ck] = sv;
                if (MODE == 1) {
                    const int slot = (k + 1) % 3;
                    acc += (double)su * (double)xs[0][slot][w + 1][lane + 1] +
                           (double)sv * (double)xs[1][slot][w + 1][lane + 1];
                }
            }
        }
        __syncthreads();   // the slot of level k-1 is restaged next
    }
    if (MODE != 0) {
        dd v[1] = {{acc, 0.0}};
        block_reduce_dd<1, 256>(v);
        if (threadIdx.x == 0) {
            partials[2 * lin + 0] = v[0].hi;
            partials[2 * lin + 1] = v[0].lo;
        }
    }
}

// FP32 half storage with a level-to-level exchange of the partner slots.
// The lower couplings of row (k, i, j) to level k-1 (m = 0..8) are the
// transposed slots 13 - m of the rows (k-1, i+iA, j+jA), which the tile's
// own threads loaded one level earlier as their couplings to level k
// (slots 5..13).  Instead of re-reading them (they fall out of L2 between
// two levels: one level step of all resident tiles exceeds an XCD's 4 MB),
// each thread keeps its slots 8, 9, 10 in registers for itself and its two
// lane neighbours (cross-lane shuffles) and writes slots 5, 6, 7 and 11, 12,
// 13 to LDS for the threads of the neighbouring tile rows; only partners
// outside the 64 x 4 tile are read from memory.  The products are summed in
// the order and with the values of k_spmv_sym, so the results are identical.
template <int MODE>
__global__ __launch_bounds__(256) void k_spmv_symx(Dims d, const float* __restrict__ S, const float* __restrict__ xu,
                                                   const float* __restrict__ xv, float* __restrict__ yu,
                                                   float* __restrict__ yv, const float* __restrict__ bu,
                                                   const float* __restrict__ bv, double* __restrict__ partials) {
    __shared__ float xs[2][4][6][WAVE + 2];
    // ex[q][c][w][lane]: q 0..2 slots 11, 12, 13 (read by tile row w + 1),
    // q 3..5 slots 5, 6, 7 (read by tile row w - 1), of the previous level
    __shared__ float ex[6][4][4][WAVE];
    const int nbx = d.nxp / WAVE;
    const int nblk = nbx * ((d.nyo + 3) / 4);
    const int lin = xcd_remap(blockIdx.x, nblk);
    const int bx = lin % nbx, by = lin / nbx;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int io = bx * WAVE + lane;
    const int jo = by * 4 + w;
    const bool live = io < d.nxo && jo < d.nyo;
    const int i0 = d.ilo + bx * WAVE - 1, j0 = d.jlo + by * 4 - 1;
    auto stage = [&](int kk) {
        const int slot = (kk + 1) % 4;
        for (int e = threadIdx.x; e < 6 * (WAVE + 2); e += 256) {
            const int r = e / (WAVE + 2), c = e % (WAVE + 2);
            const int vi = i0 + c, vj = j0 + r;
            float a = 0.f, b = 0.f;
            if (vi >= 1 && vi <= d.nvx && vj >= 1 && vj <= d.nvy) {
                const long long n = vidx(d, kk, vi, vj);
                a = xu[n];
                b = xv[n];
            }
            xs[0][slot][r][c] = a;
            xs[1][slot][r][c] = b;
        }
    };
    double acc = 0.0;
    int po[9];
    unsigned inl = 0;   // bit h: partner h is the tile's own neighbouring thread
    const long long LS = (long long)d.nyo * d.nxp / WAVE * (NSLOT * 4 * WAVE);
    if (live) {
        const int i = d.ilo + io, j = d.jlo + jo;
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            const int iA = h % 3 - 1, jA = h / 3 - 1;
            const int pc = partner_col(d, i + iA, j + jA);
            po[h] = pc < 0 ? -1 : (pc / WAVE) * (NSLOT * 4 * WAVE) + (pc % WAVE) * 4;
            const int li = lane + iA, wj = w + jA;
            if (pc >= 0 && li >= 0 && li < WAVE && wj >= 0 && wj < 4 && pc == (jo + jA) * d.nxp + (io + iA))
                inl |= 1u << h;
        }
    }
    float p8[4] = {0.f, 0.f, 0.f, 0.f}, p9[4] = {0.f, 0.f, 0.f, 0.f}, p10[4] = {0.f, 0.f, 0.f, 0.f};
    stage(-1);
    stage(0);
    // x of level kk sits in ring slot (kk + 1) % 4.  stage(k + 1) overwrites
    // level k - 3, last read in step k - 2; every thread is past the mid
    // barrier of step k - 1 by then, so two barriers per level suffice.
    for (int k = 0; k < d.nz; ++k) {
        stage(k + 1);
        __syncthreads();   // x of levels k-1..k+1 staged; ex holds level k-1
        float l10[4], r8[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            l10[c] = __shfl_up(p10[c], 1, WAVE);     // slot 10 of lane - 1
            r8[c] = __shfl_down(p8[c], 1, WAVE);     // slot 8 of lane + 1
        }
        float su = 0.f, sv = 0.f;
        const float* Sk = S + (long long)k * LS;
        const float* Sb = Sk - LS;
        // this level's own slots and same-level partner slots, loaded before
        // the level k-1 part so that they are in flight across the barrier
        float own[NSLOT][4];
        if (live) {
#pragma unroll
            for (int q = 0; q < NSLOT; ++q) ld_slot(Sk + po[4] + q * (4 * WAVE), own[q][0], own[q][1], own[q][2], own[q][3]);
        }
        if (live && k > 0) {
#pragma unroll
            for (int m = 0; m < 9; ++m) {   // level k-1: transposed slots 13 - m of the partners
                if (po[m] < 0) continue;
                float a, b, c, e;
                if (inl & (1u << m)) {
                    const int li = lane + m % 3 - 1;
                    if (m == 4) {
                        a = p9[0]; b = p9[1]; c = p9[2]; e = p9[3];
                    } else if (m == 3) {
                        a = l10[0]; b = l10[1]; c = l10[2]; e = l10[3];
                    } else if (m == 5) {
                        a = r8[0]; b = r8[1]; c = r8[2]; e = r8[3];
                    } else if (m < 3) {     // row j-1 (wave w-1), slot 13 - m = 11 + (2 - m)
                        const int q = 2 - m;
                        a = ex[q][0][w - 1][li]; b = ex[q][1][w - 1][li];
                        c = ex[q][2][w - 1][li]; e = ex[q][3][w - 1][li];
                    } else {                // row j+1 (wave w+1), slot 13 - m = 5 + (8 - m)
                        const int q = 3 + (8 - m);
                        a = ex[q][0][w + 1][li]; b = ex[q][1][w + 1][li];
                        c = ex[q][2][w + 1][li]; e = ex[q][3][w + 1][li];
                    }
                } else {
                    ld_slot(Sb + po[m] + (13 - m) * (4 * WAVE), a, b, c, e);
                }
                const float cuu = a, cuv = c, cvu = b, cvv = e;     // transposed 2x2 block
                const int slot = k % 4;                              // level k-1
                const int r = w + 1 + ((m / 3) % 3 - 1), cc = lane + 1 + (m % 3 - 1);
                const float xun = xs[0][slot][r][cc], xvn = xs[1][slot][r][cc];
                su = su + cuu * xun + cuv * xvn;
                sv = sv + cvu * xun + cvv * xvn;
            }
        }
        __syncthreads();   // every read of ex (level k-1) is done
        if (live) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    ex[q][c][w][lane] = own[11 + q][c];
                    ex[3 + q][c][w][lane] = own[5 + q][c];
                }
                p8[c] = own[8][c];
                p9[c] = own[9][c];
                p10[c] = own[10][c];
            }
#pragma unroll
            for (int m = 9; m < 27; ++m) {
                float cuu, cuv, cvu, cvv;
                bool have = true;
                if (m < 13) {   // same level: transposed slots of the partners
                    const int h = m % 9;
                    have = po[h] >= 0;
                    cuu = cuv = cvu = cvv = 0.f;
                    if (have) {
                        float a, b, c, e;
                        ld_slot(Sk + po[h] + (13 - m) * (4 * WAVE), a, b, c, e);
                        cuu = a; cuv = c; cvu = b; cvv = e;
                    }
                } else if (m == 13) {
                    cuu = own[0][0]; cuv = own[0][1]; cvu = own[0][1]; cvv = own[0][2];
                } else {
                    cuu = own[m - 13][0]; cuv = own[m - 13][1]; cvu = own[m - 13][2]; cvv = own[m - 13][3];
                }
                if (have) {
                    const int slot = (k + m / 9) % 4;
                    const int r = w + 1 + ((m / 3) % 3 - 1), c = lane + 1 + (m % 3 - 1);
                    const float xun = xs[0][slot][r][c], xvn = xs[1][slot][r][c];
                    su = su + cuu * xun + cuv * xvn;
                    sv = sv + cvu * xun + cvv * xvn;
                }
            }
            const long long ck = vidx(d, k, d.ilo + io, d.jlo + jo);
            if (MODE == 2) {
                const float ru = bu[ck] - su, rv = bv[ck] - sv;
                yu[ck] = ru;
                yv[ck] = rv;
                acc += (double)ru * (double)ru + (double)rv * (double)rv;
            } else {
                yu[ck] = su;
                yv[ck] = sv;
                if (MODE == 1) {
                    const int slot = (k + 1) % 4;
                    acc += (double)su * (double)xs[0][slot][w + 1][lane + 1] +
                           (double)sv * (double)xs[1][slot][w + 1][lane + 1];
                }
            }
        }
    }
    if (MODE != 0) {
        dd v[1] = {{acc, 0.0}};
        block_reduce_dd<1, 256>(v);
        if (threadIdx.x == 0) {
            partials[2 * lin + 0] = v[0].hi;
            partials[2 * lin + 1] = v[0].lo;
        }
    }
}

// FP64 half storage: the same slots with double quads,
//     Sd[((r/64)*14 + s)*256 + (r%64)*4 + c]
// (448 B per node instead of 864), for the FP64 solves when the FP64 matrix
// is exactly symmetric.  Sums in the order and with the values of k_spmv,
// so the results are identical.
__global__ void k_convert_sym64(long long nrows, const double* __restrict__ src, double* __restrict__ dst) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= nrows * NSLOT * 4) return;
    const int c = (int)(t % 4);
    const long long u = t / 4;
    const int lane = (int)(u % WAVE);
    const long long v = u / WAVE;
    const int s = (int)(v % NSLOT);
    const long long rb = v / NSLOT;
    const double* row = src + rb * NCOEF * WAVE + lane;
    double val;
    if (s == 0) {
        const int blk = c == 0 ? 0 : (c == 1 ? 1 : 3);
        val = c == 3 ? 0.0 : row[(blk * 27 + 13) * WAVE];
    } else {
        val = row[(c * 27 + 13 + s) * WAVE];
    }
    dst[t] = val;
}

__device__ __forceinline__ void ld_slot64(const double* p, double& a, double& b, double& c, double& e) {
    const double2 q0 = *reinterpret_cast<const double2*>(p);
    const double2 q1 = *reinterpret_cast<const double2*>(p + 2);
    a = q0.x; b = q0.y; c = q1.x; e = q1.y;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_spmv_sym64(Dims d, const double* __restrict__ S,
                                                    const double* __restrict__ xu, const double* __restrict__ xv,
                                                    double* __restrict__ yu, double* __restrict__ yv,
                                                    const double* __restrict__ bu, const double* __restrict__ bv,
                                                    double* __restrict__ partials) {
    __shared__ double xs[2][3][6][WAVE + 2];
    const int nbx = d.nxp / WAVE;
    const int nblk = nbx * ((d.nyo + 3) / 4);
    const int lin = xcd_remap(blockIdx.x, nblk);
    const int bx = lin % nbx, by = lin / nbx;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int io = bx * WAVE + lane;
    const int jo = by * 4 + w;
    const bool live = io < d.nxo && jo < d.nyo;
    const int i0 = d.ilo + bx * WAVE - 1, j0 = d.jlo + by * 4 - 1;
    auto stage = [&](int kk) {
        const int slot = (kk + 1) % 3;
        for (int e = threadIdx.x; e < 6 * (WAVE + 2); e += 256) {
            const int r = e / (WAVE + 2), c = e % (WAVE + 2);
            const int vi = i0 + c, vj = j0 + r;
            double a = 0.0, b = 0.0;
            if (vi >= 1 && vi <= d.nvx && vj >= 1 && vj <= d.nvy) {
                const long long n = vidx(d, kk, vi, vj);
                a = xu[n];
                b = xv[n];
            }
            xs[0][slot][r][c] = a;
            xs[1][slot][r][c] = b;
        }
    };
    dd acc = {0.0, 0.0};
    int po[9];
    const long long LS = (long long)d.nyo * d.nxp / WAVE * (NSLOT * 4 * WAVE);
    if (live) {
        const int i = d.ilo + io, j = d.jlo + jo;
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            const int pc = partner_col(d, i + h % 3 - 1, j + h / 3 - 1);
            po[h] = pc < 0 ? -1 : (pc / WAVE) * (NSLOT * 4 * WAVE) + (pc % WAVE) * 4;
        }
    }
    stage(-1);
    stage(0);
    for (int k = 0; k < d.nz; ++k) {
        stage(k + 1);
        __syncthreads();
        if (live) {
            const long long ck = vidx(d, k, d.ilo + io, d.jlo + jo);
            const double* Sk = S + (long long)k * LS;
            const double* Sb = Sk - LS;
            double su = 0.0, sv = 0.0;
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                double cuu = 0.0, cuv = 0.0, cvu = 0.0, cvv = 0.0;
                bool have = true;
                if (m < 13) {
                    const int h = m % 9;
                    have = po[h] >= 0 && (m >= 9 || k > 0);
                    if (have) {
                        double a, b, c, e;
                        ld_slot64((m < 9 ? Sb : Sk) + po[h] + (13 - m) * (4 * WAVE), a, b, c, e);
                        cuu = a; cuv = c; cvu = b; cvv = e;
                    }
                } else if (m == 13) {
                    double a, b, c, e;
                    ld_slot64(Sk + po[4], a, b, c, e);
                    cuu = a; cuv = b; cvu = b; cvv = c;
                } else {
                    ld_slot64(Sk + po[4] + (m - 13) * (4 * WAVE), cuu, cuv, cvu, cvv);
                }
                if (have) {
                    const int slot = (k + m / 9) % 3;
                    const int r = w + 1 + ((m / 3) % 3 - 1), c = lane + 1 + (m % 3 - 1);
                    const double xun = xs[0][slot][r][c], xvn = xs[1][slot][r][c];
                    su = su + cuu * xun + cuv * xvn;
                    sv = sv + cvu * xun + cvv * xvn;
                }
            }
            if (MODE == 2) {
                const double ru = bu[ck] - su, rv = bv[ck] - sv;
                yu[ck] = ru;
                yv[ck] = rv;
                acc = dd_add_d(acc, ru * ru);
                acc = dd_add_d(acc, rv * rv);
            } else {
                yu[ck] = su;
                yv[ck] = sv;
                if (MODE == 1) {
                    const int slot = (k + 1) % 3;
                    acc = dd_add_d(acc, su * xs[0][slot][w + 1][lane + 1]);
                    acc = dd_add_d(acc, sv * xs[1][slot][w + 1][lane + 1]);
                }
            }
        }
        __syncthreads();
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

}  // namespace cg
