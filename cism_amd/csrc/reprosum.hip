// reprosum.hip -- reproducible global sums: the integer-vector algorithm of
// cism_reprosum_calc (cism_reprosum_mod.F90:237-1057, :1063-1818) on the
// device.
//
// Every summand (a double) is split into integer digits of RS_D = 30 bits
// on a fixed ladder of RS_L levels, level l holding the bits of weight
// 2^(-1074 + 30 l) .. 2^(-1074 + 30 l + 29): the ladder covers every finite
// double, subnormals included, so no range pass is needed.  Digits are
// added as 64-bit integers (LDS atomics per workgroup, then per-level sums
// over the workgroups); integer addition is exact and associative, so the
// level sums do not depend on the order of the summands, the launch shape
// or the number of workgroups.  The 33 bits of headroom above a digit allow
// 2^33 summands.  The final conversion carries the levels into digits of
// one sign and rounds the resulting exact integer once, to nearest even:
// the result is the exact sum correctly rounded.
//
// The solver's global sums in this mode take, like
// parallel_global_sum_stagger with reprosum on (parallel_mpi.F90:6706-6731),
// the summand a1*b1 + a2*b2 per owned node, each product and the sum
// rounded separately (no fused multiply-add), over owned vertices, all
// levels.
#include "cismgpu_internal.h"

namespace cg {

constexpr int RS_D = 30;                       // bits per digit
constexpr int RS_L = 71;                       // levels: 71 x 30 bits from 2^-1074 up
constexpr int RS_EBASE = -1074;
constexpr int RS_THREADS = 256;
constexpr int RS_MAX_BLOCKS = 1024;

__global__ void k_fill_int(int* p, int n, int v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Splits x into (at most three) digits and adds them to the level counters.
__device__ __forceinline__ void rs_add(unsigned long long* lev, double x, int* bad) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    const int ex = (int)((bits >> 52) & 0x7FF);
    if (ex == 0x7FF) { *bad = 1; return; }                 // inf or NaN
    unsigned long long M = bits & ((1ULL << 52) - 1);
    int e = RS_EBASE;
    if (ex != 0) { M |= 1ULL << 52; e = ex - 1075; }
    if (M == 0) return;
    const int s = e - RS_EBASE;                             // weight of M's last bit above 2^-1074
    const int l0 = s / RS_D, o = s % RS_D;
    const int lowbits = RS_D - o;
    unsigned long long c0 = (M & ((1ULL << lowbits) - 1)) << o;
    const unsigned long long rest = M >> lowbits;
    unsigned long long c1 = rest & ((1ULL << RS_D) - 1);
    unsigned long long c2 = rest >> RS_D;
    if (bits >> 63) { c0 = 0ULL - c0; c1 = 0ULL - c1; c2 = 0ULL - c2; }   // two's complement digits
    if (c0) atomicAdd(lev + l0, c0);
    if (c1) atomicAdd(lev + l0 + 1, c1);
    if (c2) atomicAdd(lev + l0 + 2, c2);
}

// Level sums of one workgroup into part[blockIdx.x * RS_L + l]; the summands
// are the products a1*b1 + a2*b2 (b1, a2, b2 may be null: a1, a1*b1 or
// a1*b1 + a2*b2 as given) over owned vertices of the vector layout.
template <typename TV>
__global__ __launch_bounds__(RS_THREADS) void k_rs_levels(Dims d, const TV* __restrict__ a1, const TV* __restrict__ b1,
                                                          const TV* __restrict__ a2, const TV* __restrict__ b2,
                                                          unsigned long long* __restrict__ part, int* __restrict__ bad) {
    __shared__ unsigned long long lev[RS_L];
    __shared__ int sbad;
    for (int l = threadIdx.x; l < RS_L; l += RS_THREADS) lev[l] = 0ULL;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    int mybad = 0;
    const long long ncol = (long long)d.nxo * d.nyo;
    for (long long t = blockIdx.x * (long long)RS_THREADS + threadIdx.x; t < ncol;
         t += (long long)gridDim.x * RS_THREADS) {
        const int i = d.ilo + (int)(t % d.nxo), j = d.jlo + (int)(t / d.nxo);
        for (int k = 0; k < d.nz; ++k) {
            const long long c = vidx(d, k, i, j);
            double s = (double)a1[c];
            if (b1) s = __dmul_rn(s, (double)b1[c]);
            if (a2) s = __dadd_rn(s, b2 ? __dmul_rn((double)a2[c], (double)b2[c]) : (double)a2[c]);
            rs_add(lev, s, &mybad);
        }
    }
    if (mybad) sbad = 1;
    __syncthreads();
    for (int l = threadIdx.x; l < RS_L; l += RS_THREADS) part[(size_t)blockIdx.x * RS_L + l] = lev[l];
    if (threadIdx.x == 0 && sbad) *bad = 1;
}

// Level sums of a plain array of n summands.
__global__ __launch_bounds__(RS_THREADS) void k_rs_levels_flat(long long n, const double* __restrict__ a,
                                                               unsigned long long* __restrict__ part,
                                                               int* __restrict__ bad) {
    __shared__ unsigned long long lev[RS_L];
    __shared__ int sbad;
    for (int l = threadIdx.x; l < RS_L; l += RS_THREADS) lev[l] = 0ULL;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    int mybad = 0;
    for (long long t = blockIdx.x * (long long)RS_THREADS + threadIdx.x; t < n; t += (long long)gridDim.x * RS_THREADS)
        rs_add(lev, a[t], &mybad);
    if (mybad) sbad = 1;
    __syncthreads();
    for (int l = threadIdx.x; l < RS_L; l += RS_THREADS) part[(size_t)blockIdx.x * RS_L + l] = lev[l];
    if (threadIdx.x == 0 && sbad) *bad = 1;
}

// The exact integer vector (signed level sums V[l]) rounded once to the
// nearest double (ties to even).
__device__ double rs_to_double(const long long* V) {
    long long W[RS_L];
    // sign of the total: carry the levels upward
    long long carry = 0;
    for (int l = 0; l < RS_L; ++l) {
        const long long t = V[l] + carry;
        carry = t >> RS_D;                                  // floor division (arithmetic shift)
    }
    const bool neg = carry < 0;
    carry = 0;
    unsigned long long dg[RS_L];
    for (int l = 0; l < RS_L; ++l) W[l] = neg ? -V[l] : V[l];
    for (int l = 0; l < RS_L; ++l) {
        const long long t = W[l] + carry;
        dg[l] = (unsigned long long)(t & ((1LL << RS_D) - 1));
        carry = t >> RS_D;
    }
    if (carry > 0) return neg ? -INFINITY : INFINITY;       // beyond 2^(RS_EBASE + RS_L * RS_D)
    int h = RS_L - 1;
    while (h >= 0 && dg[h] == 0) --h;
    if (h < 0) return 0.0;
    const int lo = h >= 2 ? h - 2 : 0;
    unsigned __int128 T = 0;
    for (int l = h; l >= lo; --l) T = (T << RS_D) | dg[l];
    bool sticky = false;
    for (int l = 0; l < lo; ++l) sticky = sticky || dg[l] != 0;
    int nb = 0;
    for (unsigned __int128 u = T; u; u >>= 1) ++nb;
    const int ebase = RS_EBASE + RS_D * lo;
    double r;
    if (nb <= 53) {
        r = ldexp((double)(unsigned long long)T, ebase);   // exact (no digits below: lo = 0)
    } else {
        const int sh = nb - 53;
        unsigned long long q = (unsigned long long)(T >> sh);
        const unsigned __int128 rem = T & ((((unsigned __int128)1) << sh) - 1);
        const unsigned __int128 half = ((unsigned __int128)1) << (sh - 1);
        const bool up = rem > half || (rem == half && (sticky || (q & 1ULL)));
        q += up ? 1ULL : 0ULL;
        r = ldexp((double)q, ebase + sh);
    }
    return neg ? -r : r;
}

// Sums the level vectors of nblk workgroups and writes {sum, 0} to out
// (the (hi, lo) partial format of the solver's scalar stage); NaN when a
// summand was not finite.
__global__ __launch_bounds__(RS_THREADS) void k_rs_final(const unsigned long long* __restrict__ part, int nblk,
                                                         const int* __restrict__ bad, double* __restrict__ out) {
    __shared__ long long V[RS_L];
    for (int l = threadIdx.x; l < RS_L; l += RS_THREADS) {
        unsigned long long s = 0ULL;
        for (int b = 0; b < nblk; ++b) s += part[(size_t)b * RS_L + l];
        V[l] = (long long)s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = *bad ? NAN : rs_to_double(V);
        out[1] = 0.0;
    }
}

}  // namespace cg
