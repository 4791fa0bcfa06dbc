// noalloc.hip -- compute_residual_vector_3d on device arrays in the
// reference layout without any allocation: the variant of
// cismgpu_system_residual for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops).
//
// The per-workgroup partial sums live in a file-scope device array
// (`scratch`).  Host code never passes the array's name to a kernel: the
// name is only a host-side stand-in.  The first call on a device, with
// that device current, asks the runtime for the array's device address
// (hipGetSymbolAddress) and keeps it per device; every later call on that
// device passes that address.  All calls on one device share the array
// (each call overwrites every partial it reads), so calls that overlap on
// one device race; calls on different devices do not.
#include <mutex>

#include "cismgpu_internal.h"

namespace cg {

constexpr int NA_THREADS = 256;
constexpr int NA_MAX_BLOCKS = 1 << 16;                  // nz * (nx-1) * (ny-1) <= 16,777,216 nodes
constexpr int NA_PART = 4;                              // (r,r) and (b,b), double-double each

// 65536 x 4 doubles = 2 MiB per device, held until the process ends
static __device__ __align__(256) double scratch[NA_MAX_BLOCKS * NA_PART];

// r = A u - b per node of the reference layout (k fastest), zero outside the
// owned active vertices (glissade_velo_higher.F90:9768-9826); per-workgroup
// compensated partials of (r,r) and of (b,b) over owned active nodes.
__global__ __launch_bounds__(NA_THREADS) void k_resid_ref(Dims d, const int32_t* __restrict__ active,
                                                          const double* __restrict__ Auu, const double* __restrict__ Auv,
                                                          const double* __restrict__ Avu, const double* __restrict__ Avv,
                                                          const double* __restrict__ bu, const double* __restrict__ bv,
                                                          const double* __restrict__ u, const double* __restrict__ v,
                                                          double* __restrict__ ru, double* __restrict__ rv,
                                                          double* __restrict__ part) {
    const long long n = (long long)d.nz * d.nvx * d.nvy;
    const long long t = blockIdx.x * (long long)NA_THREADS + threadIdx.x;
    dd acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (t < n) {
        const int k = (int)(t % d.nz);
        const long long col = t / d.nz;
        const int i = 1 + (int)(col % d.nvx), j = 1 + (int)(col / d.nvx);
        double r0 = 0.0, r1 = 0.0;
        if (i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi && active[col]) {
            const double* au = Auu + 27 * t;
            const double* auv = Auv + 27 * t;
            const double* avu = Avu + 27 * t;
            const double* av = Avv + 27 * t;
            double su = 0.0, sv = 0.0;
            for (int kA = -1; kA <= 1; ++kA) {
                if (k + kA < 0 || k + kA >= d.nz) continue;
                for (int jA = -1; jA <= 1; ++jA)
                    for (int iA = -1; iA <= 1; ++iA) {
                        const int m = (iA + 1) + 3 * (jA + 1) + 9 * (kA + 1);
                        const long long q = (long long)(k + kA) + (long long)d.nz * ((i + iA - 1) + (long long)d.nvx * (j + jA - 1));
                        su = su + au[m] * u[q] + auv[m] * v[q];
                        sv = sv + avu[m] * u[q] + av[m] * v[q];
                    }
            }
            r0 = su - bu[t];
            r1 = sv - bv[t];
            acc[0] = dd_add_d(dd_add_d(acc[0], r0 * r0), r1 * r1);
            acc[1] = dd_add_d(dd_add_d(acc[1], bu[t] * bu[t]), bv[t] * bv[t]);
        }
        if (ru) ru[t] = r0;
        if (rv) rv[t] = r1;
    }
    block_reduce_dd<2, NA_THREADS>(acc);
    if (threadIdx.x == 0) {
        double* p = part + (size_t)blockIdx.x * NA_PART;
        p[0] = acc[0].hi; p[1] = acc[0].lo; p[2] = acc[1].hi; p[3] = acc[1].lo;
    }
}

// Fixed-order final sums: norms[0] = |A u - b|, norms[1] = |A u - b| / |b|
// (0 when b = 0, glissade_velo_higher.F90:9881-9885).
__global__ __launch_bounds__(NA_THREADS) void k_resid_ref_final(const double* __restrict__ part, int nblk,
                                                                double* __restrict__ norms) {
    dd acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int b = threadIdx.x; b < nblk; b += NA_THREADS) {
        const double* p = part + (size_t)b * NA_PART;
        acc[0] = dd_add(acc[0], dd{p[0], p[1]});
        acc[1] = dd_add(acc[1], dd{p[2], p[3]});
    }
    block_reduce_dd<2, NA_THREADS>(acc);
    if (threadIdx.x == 0) {
        const double l2 = sqrt(acc[0].hi + acc[0].lo);
        const double l2b = sqrt(acc[1].hi + acc[1].lo);
        norms[0] = l2;
        norms[1] = l2b > 0.0 ? l2 / l2b : 0.0;
    }
}

}  // namespace cg

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
double* g_na_scratch[NA_MAX_DEVICES] = {};   // device address of `scratch` on each device

// Address of this device's copy of the scratch array; looked up once per
// device, with that device current.
int na_scratch(double** out) {
    int dev = 0;
    CG_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        cismgpu_set_error("device %d beyond the %d devices the scratch table covers", dev, NA_MAX_DEVICES);
        return CISMGPU_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(cg::scratch));
        if (e != hipSuccess || !p) {
            cismgpu_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return CISMGPU_ERR_HIP;
        }
        g_na_scratch[dev] = (double*)p;
    }
    *out = g_na_scratch[dev];
    return CISMGPU_OK;
}

}  // namespace

extern "C" int32_t cismgpu_residual_vector_3d_noalloc(const cismgpu_grid* grid, const int32_t* active_vertex,
                                                      const double* Auu, const double* Auv, const double* Avu,
                                                      const double* Avv, const double* bu, const double* bv,
                                                      const double* uvel, const double* vvel, double* resid_u,
                                                      double* resid_v, double* norms, void* stream) {
    if (!grid || !active_vertex || !Auu || !Auv || !Avu || !Avv || !bu || !bv || !uvel || !vvel || !norms) {
        cismgpu_set_error("residual_vector_3d_noalloc: null argument");
        return CISMGPU_ERR_ARG;
    }
    if (grid->nx < 4 || grid->ny < 4 || grid->nz < 1 || grid->ilo < 2 || grid->jlo < 2 ||
        grid->ihi > grid->nx - 2 || grid->jhi > grid->ny - 2 || grid->ilo > grid->ihi || grid->jlo > grid->jhi) {
        cismgpu_set_error("residual_vector_3d_noalloc: bad grid nx=%d ny=%d nz=%d owned (%d:%d, %d:%d)", grid->nx,
                          grid->ny, grid->nz, grid->ilo, grid->ihi, grid->jlo, grid->jhi);
        return CISMGPU_ERR_ARG;
    }
    cg::Dims d{};
    d.nx = grid->nx; d.ny = grid->ny; d.nz = grid->nz;
    d.nvx = grid->nx - 1; d.nvy = grid->ny - 1;
    d.ilo = grid->ilo; d.ihi = grid->ihi; d.jlo = grid->jlo; d.jhi = grid->jhi;
    const long long n = (long long)d.nz * d.nvx * d.nvy;
    const long long nblk = (n + cg::NA_THREADS - 1) / cg::NA_THREADS;
    if (nblk > cg::NA_MAX_BLOCKS) {
        cismgpu_set_error("residual_vector_3d_noalloc: %lld nodes exceed the scratch limit of %d", n,
                          cg::NA_MAX_BLOCKS * cg::NA_THREADS);
        return CISMGPU_ERR_ARG;
    }
    double* part = nullptr;
    const int st = na_scratch(&part);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    cg::k_resid_ref<<<(unsigned)nblk, cg::NA_THREADS, 0, s>>>(d, active_vertex, Auu, Auv, Avu, Avv, bu, bv, uvel,
                                                             vvel, resid_u, resid_v, part);
    CG_CHECK(hipGetLastError());
    cg::k_resid_ref_final<<<1, cg::NA_THREADS, 0, s>>>(part, (int)nblk, norms);
    CG_CHECK(hipGetLastError());
    return CISMGPU_OK;
}
