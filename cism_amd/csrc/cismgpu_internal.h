// cismgpu_internal.h -- device layouts and shared helpers of libcismgpu.
//
// Vector layout ("level-major, i fastest"): a field v(k,i,j) of the
// reference's (nz, nx-1, ny-1) staggered grid is stored at
//     vidx(k,i,j) = ((k+1)*nvy + (j-1))*px + (i-1)        k = 0..nz-1 (0-based level)
// with two zero padding levels (k = -1 and k = nz) so that the vertical
// neighbours of the top and bottom levels are always readable zeros, and a
// row pitch px = nvx rounded up to 8 doubles.
//
// Matrix layout ("row-blocked SoA"): only owned rows are stored.  Owned row
//     r = (k*nyo + jo)*nxp + io,   io = i-ilo, jo = j-jlo, nxp = nxo rounded to 64,
// coefficient c = blk*27 + (m-1) (blk 0..3 = uu, uv, vu, vv) lives at
//     A[((r/64)*108 + c)*64 + r%64]
// so one wavefront streams one contiguous 108*64-element block per level.
// Coefficients that couple to a level outside 1..nz and all rows of inactive
// columns are stored as zero; the SpMV therefore needs no masks or bounds.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/cismgpu.h"

#define CG_CHECK(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            cismgpu_set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return CISMGPU_ERR_HIP;                                                          \
        }                                                                                    \
    } while (0)

extern "C" void cismgpu_set_error(const char* fmt, ...);

namespace cg {

constexpr int WAVE = 64;
constexpr int NCOEF = 108;          // 4 blocks x 27 neighbours
constexpr int RED_THREADS = 256;    // block size of the reduction kernels
constexpr int SCAL_THREADS = 1024;  // block size of the one-workgroup scalar stage (k_scalars)

struct Dims {
    int nx, ny, nz;                 // cells incl. halos; levels
    int nvx, nvy;                   // vertex grid
    int ilo, ihi, jlo, jhi;         // owned vertex bounds (1-based)
    int nxo, nyo, nxp;              // owned extents; nxo padded to 64
    int px;                         // vector row pitch
    int bc;                         // global_bc
    long long vlen;                 // doubles per padded vector: (nz+2)*nvy*px
    long long nrows;                // padded owned rows: nz*nyo*nxp
    long long plane;                // nvy*px
};

__host__ __device__ inline long long vidx(const Dims& d, int k, int i, int j) {
    return ((long long)(k + 1) * d.nvy + (j - 1)) * d.px + (i - 1);
}

// ---------------------------------------------------------------- double-double
struct dd { double hi, lo; };

__device__ __forceinline__ dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ dd dd_add_d(dd a, double b) {
    dd s = two_sum(a.hi, b);
    double lo = s.lo + a.lo;
    double hi = s.hi + lo;
    return {hi, lo - (hi - s.hi)};
}
__device__ __forceinline__ dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    double lo = s.lo + (a.lo + b.lo);
    double hi = s.hi + lo;
    return {hi, lo - (hi - s.hi)};
}
__device__ __forceinline__ dd dd_shfl_down(dd v, int off) {
    return {__shfl_down(v.hi, off, WAVE), __shfl_down(v.lo, off, WAVE)};
}

// Block-wide compensated reduction of NQ quantities, deterministic tree.
// Result valid in thread 0.
template <int NQ, int BS>
__device__ __forceinline__ void block_reduce_dd(dd (&v)[NQ]) {
    __shared__ dd sh[NQ][BS / WAVE];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) v[q] = dd_add(v[q], dd_shfl_down(v[q], off));
        if (lane == 0) sh[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            dd s = sh[q][0];
            for (int i = 1; i < BS / WAVE; ++i) s = dd_add(s, sh[q][i]);
            v[q] = s;
        }
    }
}

// ---------------------------------------------------------------- storage types
__device__ __forceinline__ double ld_coef(const double* p) { return *p; }
__device__ __forceinline__ float ld_coef(const float* p) { return *p; }
__device__ __forceinline__ float ld_coef(const __hip_bfloat16* p) { return __bfloat162float(*p); }

}  // namespace cg
