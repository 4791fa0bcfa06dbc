// libcismgpu.hip -- host drivers and C ABI of the structured 3-D velocity
// solver (glissade_velo_higher_pcg.F90) on MI355X.
//
// One translation unit: the kernels are included so that every template
// instance the drivers launch is generated here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "cismgpu_internal.h"
#include "kernels.h"
#include "kernels.hip"
#include "assembly.hip"
#include "assembly2d.hip"
#include "spmv_sym.hip"
#include "ensemble.hip"
#include "noalloc.hip"
#include "reprosum.hip"
#include "calcbeta.hip"

using namespace cg;

static thread_local char g_err[1024] = "";
extern "C" void cismgpu_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
extern "C" const char* cismgpu_last_error(void) { return g_err; }

extern "C" int32_t cismgpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

namespace {

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct TriPC {               // horizontal tridiagonal preconditioner (precond 3/4), 2-D arrays
    double* a = nullptr;     // 13 fields of nvx*nvy doubles
    double *Adu, *Asu, *deu, *omu, *Adv, *Asv, *dev_, *omv, *rsu, *rsv, *zsu, *zsv, *gam;
    float* a32 = nullptr;
    float *Adu32, *Asu32, *deu32, *omu32, *Adv32, *Asv32, *dev32, *omv32, *rsu32, *rsv32, *zsu32, *zsv32, *gam32;
    int mode = 0;            // 3 or 4 when set up
};

}  // namespace

struct cismgpu_system {
    Dims d{};
    hipStream_t s = nullptr;
    int8_t* active = nullptr;      // nvx*nvy
    double* A64 = nullptr;
    float* A32 = nullptr;
    __hip_bfloat16* A16 = nullptr;
    double* stage = nullptr;       // 2 raw blocks (27*nz*nvx*nvy each), kept for precond setup
    double* M64 = nullptr;         // 6 vectors (SIA) + 2 (diag)
    float* M32 = nullptr;
    double* v64 = nullptr;         // vectors: 0 xu 1 xv 2 bu 3 bv 4 ru 5 rv 6 du 7 dv 8 qu 9 qv 10 zu 11 zv 12 Azu 13 Azv
    float* v32 = nullptr;          // inner: 0 eu 1 ev 2 r0u 3 r0v 4 ru 5 rv 6 du 7 dv 8 qu 9 qv 10 zu 11 zv 12 Azu 13 Azv
    double* partA = nullptr;
    double* partB = nullptr;
    double* scal = nullptr;
    double* scal_h = nullptr;      // pinned mirror
    TriPC tri;
    // assembly inputs on the device
    void* asm_buf = nullptr;
    size_t asm_bytes = 0;
    double* efvs = nullptr;
    double* Atmp = nullptr;      // second matrix buffer for the symmetry fix (only when a pair differs)
    double* elemq = nullptr;     // element-once assembly: inverse Jacobian and f per element and point
    bool raw_ready = false;      // stage holds the raw Auu/Avv of the current matrix
    bool saved_raw_ready = false;
    bool sym_exact = false;      // the assembled FP64 matrix is known to be exactly symmetric
    bool edge_zero = false;      // no-ice rule: the rows of the last owned vertex row/column are zero
    double* work = nullptr;      // Picard work vectors: slot q = (u, v) at work + 2*q*vlen
    int nwork = 0;
    double* saved = nullptr;     // saved assembly (A, b, preconditioner, raw diagonal blocks)
    size_t saved_n = 0;          // doubles in `saved`
    bool have_saved = false;
    bool have_matrix = false;
    bool two_d = false;          // one-level (depth-integrated) system: 9-point stencil, nz = 1
    int nz_levels = 0;           // 2-D system: sigma levels of the assembly inputs
    double* load3 = nullptr;     // 2-D system: nz_levels-level load work pair
    double* diva_buf = nullptr;  // 2-D system: DIVA inputs and state
    DivaDev dv{};
    AsmDev asm_in{};             // device inputs of the last assembly (2-D, or 3-D when asm3_ready)
    bool asm3_ready = false;     // asm_in holds the inputs of a 3-D assembly
    double* efva2 = nullptr;     // 2-D system: (4, nx, ny) vertical-mean viscosities (L1L2)
    int last_approx = 0;         // which_ho_approx of the last 2-D assembly
    bool diva_ready = false;
    bool m32_ready = false, a32_ready = false, a16_ready = false;
    bool a32_sym = false;        // FP32 copy in half (symmetric) storage
    double* A64h = nullptr;      // FP64 half (symmetric) storage, when the FP64 matrix is exactly symmetric
    bool a64h_ready = false, a64_sym = false;
    std::vector<hipEvent_t> ev;    // SpMV timing events (pairs)
    int ev_used = 0;
    bool timing = false;
    int spmv_kind_timed = 0;

    bool repro = false;          // grid->reprosum: exact (integer-vector) global sums in the FP64 solves
    unsigned long long* rs_part = nullptr;   // per-workgroup level sums of one exact sum
    int* rs_bad = nullptr;       // a summand was not finite
    double* rs_out = nullptr;    // exact sums as (value, 0) pairs, 4 slots

    // basal traction (glissade_calcbeta): uploaded basal physics, beta and
    // its halo partner, a host-velocity staging pair, the invalid-value slot
    void* cb_buf = nullptr;
    CBDev cb{};
    double* cb_beta = nullptr;   // (nvx, nvy), then a scratch field of the same size
    double* cb_vel = nullptr;    // 2 (nvx, nvy) fields
    double* cb_dvel = nullptr;   // DIVA: basal level of the last compute_3d_velocity_diva, 2 (nvx, nvy) fields
    bool cb_dvel_valid = false;
    unsigned long long* cb_bad = nullptr;
    bool cb_setup = false;       // calcbeta_setup done
    bool cb_valid = false;       // cb_beta holds a calcbeta result

    double* V(int q) { return v64 + (size_t)q * d.vlen; }
    float* W(int q) { return v32 + (size_t)q * d.vlen; }
};

namespace {

int check_grid(const cismgpu_grid* g, bool two_d = false) {
    if (!g) { cismgpu_set_error("null grid"); return CISMGPU_ERR_ARG; }
    if (g->nx < 6 || g->ny < 6 || g->nz < (two_d ? 1 : 2) || g->nz > (two_d ? 1 : 64)) {
        cismgpu_set_error("bad grid sizes nx=%d ny=%d nz=%d (nz must be 2..64)", g->nx, g->ny, g->nz);
        return CISMGPU_ERR_ARG;
    }
    if (g->ilo < 2 || g->ihi > g->nx - 2 || g->jlo < 2 || g->jhi > g->ny - 2 || g->ilo > g->ihi ||
        g->jlo > g->jhi) {
        cismgpu_set_error("owned bounds (%d:%d, %d:%d) outside the halo rules", g->ilo, g->ihi, g->jlo, g->jhi);
        return CISMGPU_ERR_ARG;
    }
    if (g->global_bc < 0 || g->global_bc > 3) { cismgpu_set_error("bad global_bc %d", g->global_bc); return CISMGPU_ERR_ARG; }
    if ((g->global_bc == 0 || g->global_bc == 2) && (g->ilo != 3 || g->jlo != 3 || g->ihi != g->nx - 2 || g->jhi != g->ny - 2)) {
        cismgpu_set_error("periodic halo rules need the one-task bounds 3..n-2");
        return CISMGPU_ERR_ARG;
    }
    return CISMGPU_OK;
}

Dims make_dims(const cismgpu_grid* g) {
    Dims d{};
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
    d.nvx = g->nx - 1; d.nvy = g->ny - 1;
    d.ilo = g->ilo; d.ihi = g->ihi; d.jlo = g->jlo; d.jhi = g->jhi;
    d.nxo = d.ihi - d.ilo + 1; d.nyo = d.jhi - d.jlo + 1;
    d.nxp = (d.nxo + WAVE - 1) / WAVE * WAVE;
    d.px = (d.nvx + 7) / 8 * 8;
    d.bc = g->global_bc;
    d.plane = (long long)d.nvy * d.px;
    d.vlen = (long long)(d.nz + 2) * d.plane;
    d.nrows = (long long)d.nz * d.nyo * d.nxp;
    return d;
}

inline int spmv_blocks(const Dims& d) { return (d.nxp / WAVE) * ((d.nyo + 3) / 4); }
inline int col_blocks(const Dims& d) { return (int)nblocks((long long)d.nvx * d.nvy, 256); }

template <typename TV>
void halo(cismgpu_system* S, TV* a, TV* b, int nlev) {
    const Dims& d = S->d;
    if (d.bc == 0 || d.bc == 2) {
        const int nj = d.jhi - d.jlo + 1;
        k_halo_ew<TV><<<nblocks((long long)nlev * nj, 256), 256, 0, S->s>>>(d, a, b, nlev);
        k_halo_ns<TV><<<nblocks((long long)nlev * d.nvx, 256), 256, 0, S->s>>>(d, a, b, nlev);
    } else if (d.bc == 1) {
        k_halo_zero<TV><<<col_blocks(d), 256, 0, S->s>>>(d, a, b, nlev, 1, 1);
    } else if (d.bc == 3) {
        k_halo_zero<TV><<<col_blocks(d), 256, 0, S->s>>>(d, a, b, nlev, 2, 2);
    }
}

// staggered halo of two 2-D vertex fields
template <typename TV>
void halo2d(cismgpu_system* S, TV* a, TV* b) {
    const Dims& d = S->d;
    if (d.bc == 0 || d.bc == 2) {
        k_halo2d_ew<TV><<<nblocks(2LL * (d.jhi - d.jlo + 1), 256), 256, 0, S->s>>>(d, a, b);
        k_halo2d_ns<TV><<<nblocks(2LL * d.nvx, 256), 256, 0, S->s>>>(d, a, b);
    } else if (d.bc == 1 || d.bc == 3) {
        k_halo2d_zero<TV><<<nblocks((long long)d.nvx * d.nvy, 256), 256, 0, S->s>>>(d, a, b);
    }
}

// FP32 half storage: the kernel with the level-to-level LDS exchange of the
// partner slots (default) or the one that re-reads them (CISMGPU_SPMV_SYMX=0)
static bool use_symx() {
    const char* e = getenv("CISMGPU_SPMV_SYMX");
    return !(e && e[0] == '0');
}

template <int MODE>
void spmv_sym32(cismgpu_system* S, const float* A, const float* xu, const float* xv, float* yu, float* yv,
                const float* bu, const float* bv, double* part) {
    if (use_symx())
        k_spmv_symx<MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, A, xu, xv, yu, yv, bu, bv, part);
    else
        k_spmv_sym<float, MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, A, xu, xv, yu, yv, bu, bv, part);
}

template <typename TA, typename TV, int MODE>
void spmv(cismgpu_system* S, const TA* A, const TV* xu, const TV* xv, TV* yu, TV* yv, const TV* bu, const TV* bv,
          double* part) {
    const bool t = S->timing && S->ev_used + 2 <= (int)S->ev.size();
    if (t) hipEventRecord(S->ev[S->ev_used], S->s);
    if (S->two_d)   // one level: only the kA = 0 plane of the rows is read
        k_spmv_plane<TA, TV, MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, A, xu, xv, yu, yv, bu, bv, part);
    else if constexpr (std::is_same<TA, double>::value) {
        if (S->a64_sym && (const double*)A == S->A64)   // FP64 half (symmetric) storage
            k_spmv_sym64<MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, S->A64h, xu, xv, yu, yv, bu, bv, part);
        else
            k_spmv<TA, TV, MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, A, xu, xv, yu, yv, bu, bv, part);
    }
    else if (std::is_same<TA, float>::value && S->a32_sym)   // FP32 half (symmetric) storage
        spmv_sym32<MODE>(S, (const float*)A, (const float*)xu, (const float*)xv, (float*)yu, (float*)yv,
                         (const float*)bu, (const float*)bv, part);
    else   // FP32 / BF16 copies in the quad layout
        k_spmv_q<TA, MODE><<<spmv_blocks(S->d), 256, 0, S->s>>>(S->d, A, xu, xv, yu, yv, bu, bv, part);
    if (t) { hipEventRecord(S->ev[S->ev_used + 1], S->s); S->ev_used += 2; }
}

// The preconditioner fields of an exactly symmetric assembled matrix: the
// SIA solve takes each sub diagonal from the super diagonal of the level
// below.  Only after an assembly that found the rows exactly symmetric:
// there the fields of the halo columns come from the owned ones by the halo
// rule (or are zero), so they are symmetric too; a matrix given by the
// caller may carry other halo values.
template <typename TV>
int precond_sym(const cismgpu_system* S) {
    return S->sym_exact ? 1 : 0;
}

template <typename TV, int PRE, int STD>
void launch_update(cismgpu_system* S, TV* const* v, const TV* M) {
    // v: xu xv ru rv du dv qu qv zu zv Azu Azv
    const Dims& d = S->d;
#define UPD(NZM, SY) k_update<TV, NZM, PRE, STD, SY><<<col_blocks(d), 256, 0, S->s>>>(d, S->scal, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11], M, S->active, S->partA)
    if (precond_sym<TV>(S)) {
        if (d.nz <= 16) UPD(16, 1);
        else if (d.nz <= 32) UPD(32, 1);
        else UPD(64, 1);
    } else {
        if (d.nz <= 16) UPD(16, 0);
        else if (d.nz <= 32) UPD(32, 0);
        else UPD(64, 0);
    }
#undef UPD
}

template <typename TV>
void launch_precond(cismgpu_system* S, int kind, const TV* ru, const TV* rv, TV* zu, TV* zv, const TV* M,
                    const TV* Dg, const TV* zsu, const TV* zsv) {
    const Dims& d = S->d;
    const int msym = precond_sym<TV>(S);
#define PRC(NZM) k_precond<TV, NZM><<<col_blocks(d), 256, 0, S->s>>>(d, kind, S->scal, ru, rv, zu, zv, M, Dg, S->active, zsu, zsv, S->partA, msym)
    if (d.nz <= 16) PRC(16);
    else if (d.nz <= 32) PRC(32);
    else PRC(64);
#undef PRC
}

void scalars(cismgpu_system* S, int stage, const double* pa, int na, int sa, const double* pb, int nb, int sb,
             double tol = 0.0) {
    k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(stage, pa, na, sa, pb, nb, sb, S->scal, tol);
}

// Exact (reproducible) sum over owned vertices of the summands a1*b1 + a2*b2
// into the pair S->rs_out[2 * slot] (value, 0), for the scalar stage.
template <typename TV>
void rs_dot(cismgpu_system* S, int slot, const TV* a1, const TV* b1, const TV* a2, const TV* b2) {
    const Dims& d = S->d;
    const int nb = std::max(1, std::min(RS_MAX_BLOCKS, (int)nblocks((long long)d.nxo * d.nyo, RS_THREADS)));
    k_fill_int<<<1, 64, 0, S->s>>>(S->rs_bad, 1, 0);
    k_rs_levels<TV><<<nb, RS_THREADS, 0, S->s>>>(d, a1, b1, a2, b2, S->rs_part, S->rs_bad);
    k_rs_final<<<1, RS_THREADS, 0, S->s>>>(S->rs_part, nb, S->rs_bad, S->rs_out + 2 * slot);
}

// dot over owned vertices of (a1,b1)+(a2,b2), compensated, into partB (stride 2)
template <typename TV>
__global__ __launch_bounds__(256) void k_dot(Dims d, const TV* __restrict__ a1, const TV* __restrict__ b1,
                                             const TV* __restrict__ a2, const TV* __restrict__ b2,
                                             double* __restrict__ partials) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    dd acc = {0.0, 0.0};
    if (t < (long long)d.nxo * d.nyo) {
        const int i = d.ilo + (int)(t % d.nxo), j = d.jlo + (int)(t / d.nxo);
        for (int k = 0; k < d.nz; ++k) {
            const long long c = vidx(d, k, i, j);
            acc = dd_add_d(acc, (double)a1[c] * (double)b1[c]);
            acc = dd_add_d(acc, (double)a2[c] * (double)b2[c]);
        }
    }
    dd v[1] = {acc};
    block_reduce_dd<1, 256>(v);
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = v[0].hi; partials[2 * blockIdx.x + 1] = v[0].lo; }
}

template <typename TV>
int dot_blocks(const Dims& d) { return (int)nblocks((long long)d.nxo * d.nyo, 256); }

int sync_scal(cismgpu_system* S) {
    CG_CHECK(hipMemcpyAsync(S->scal_h, S->scal, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

int set_scal(cismgpu_system* S, int slot, double v) {
    CG_CHECK(hipMemcpyAsync(S->scal + slot, &v, sizeof(double), hipMemcpyHostToDevice, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

int reset_scal(cismgpu_system* S) {
    CG_CHECK(hipMemsetAsync(S->scal, 0, S_NSLOTS * sizeof(double), S->s));
    return CISMGPU_OK;
}

// ---------------------------------------------------------------- tridiagonal PC setup
// Vertical sums of the 2-D stencil entries used by the line solves
// (pcg.F90:3815-3840 local; :3996-4021 global) from the raw reference blocks.
__global__ void k_tri_vsum(Dims d, int mode, int two_d, const double* __restrict__ Auu, const double* __restrict__ Avv,
                           const int8_t* __restrict__ active, double* __restrict__ su_s, double* __restrict__ su_d,
                           double* __restrict__ su_p, double* __restrict__ sv_s, double* __restrict__ sv_d,
                           double* __restrict__ sv_p) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    // 2-D indices: (-1,0)->4, (0,0)->5, (1,0)->6, (0,-1)->2, (0,1)->8 (1-based)
    const int mu[3] = {3, 4, 5}, mv[3] = {1, 4, 7};
    double u[3], v[3];
    for (int q = 0; q < 3; ++q) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < d.nz; ++k) {
            const double* pu = Auu + 27 * ((long long)k + (long long)d.nz * t);
            const double* pv = Avv + 27 * ((long long)k + (long long)d.nz * t);
            a = a + pu[mu[q]] + pu[mu[q] + 9] + pu[mu[q] + 18];
            b = b + pv[mv[q]] + pv[mv[q] + 9] + pv[mv[q] + 18];
        }
        u[q] = a; v[q] = b;
    }
    if (!active[t] && !two_d) {   // 2-D solves take the stencil as given (pcg.F90:287-297, :343-380)
        if (mode == 3) { u[0] = u[2] = v[0] = v[2] = 0.0; }
        u[1] = 1.0; v[1] = 1.0;
    }
    su_s[t] = u[0]; su_d[t] = u[1]; su_p[t] = u[2];
    sv_s[t] = v[0]; sv_d[t] = v[1]; sv_p[t] = v[2];
}

// Forward factors of the owned-block line systems (pcg.F90:4313-4390 local;
// :4446-4517 global, which also forms the reverse sweep whose zero pivots
// are fatal).  One thread per line.  In: sub, diag, sup (2-D).  Out: Ad =
// 1/diag, de = 1/denom, om; *flag = 1 on a zero pivot.
__global__ void k_tri_factor(Dims d, int dir, int mode, double* __restrict__ sub, double* __restrict__ diag,
                             double* __restrict__ sup, double* __restrict__ Ad, double* __restrict__ de,
                             double* __restrict__ om, double* __restrict__ flagp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = d.nvx;
    int lo, hi, fixed;
    if (dir == 0) { fixed = d.jlo + t; if (fixed > d.jhi) return; lo = d.ilo; hi = d.ihi; }
    else { fixed = d.ilo + t; if (fixed > d.ihi) return; lo = d.jlo; hi = d.jhi; }
    auto id = [&](int s) -> long long {
        return dir == 0 ? (long long)(fixed - 1) * nv + (s - 1) : (long long)(s - 1) * nv + (fixed - 1);
    };
    const int a0 = (mode == 3) ? 1 : lo, a1 = (mode == 3) ? (dir == 0 ? d.nvx : d.nvy) : hi;
    for (int s = a0; s <= a1; ++s) {
        if (diag[id(s)] == 0.0) { diag[id(s)] = 1.0; sub[id(s)] = 0.0; sup[id(s)] = 0.0; }
        om[id(s)] = 0.0; de[id(s)] = 0.0;
    }
    om[id(lo)] = sup[id(lo)] / diag[id(lo)];
    for (int s = lo + 1; s <= hi; ++s) {
        const double den = diag[id(s)] - sub[id(s)] * om[id(s - 1)];
        if (den == 0.0) { *flagp = 1.0; return; }
        de[id(s)] = 1.0 / den;
        om[id(s)] = sup[id(s)] * de[id(s)];
    }
    if (mode == 4) {
        double omuh = sub[id(hi)] / diag[id(hi)];
        for (int s = hi - 1; s >= lo; --s) {
            const double duh = diag[id(s)] - sup[id(s)] * omuh;
            if (duh == 0.0) { *flagp = 1.0; return; }
            omuh = sub[id(s)] / duh;
        }
    }
    for (int s = a0; s <= a1; ++s) Ad[id(s)] = 1.0 / diag[id(s)];
}

int setup_tri(cismgpu_system* S, int mode) {
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy;
    TriPC& T = S->tri;
    if (!T.a) {
        CG_CHECK(hipMalloc(&T.a, 13 * n2 * sizeof(double)));
        CG_CHECK(hipMalloc(&T.a32, 13 * n2 * sizeof(float)));
        double** f[13] = {&T.Adu, &T.Asu, &T.deu, &T.omu, &T.Adv, &T.Asv, &T.dev_, &T.omv, &T.rsu, &T.rsv, &T.zsu, &T.zsv, &T.gam};
        float** g[13] = {&T.Adu32, &T.Asu32, &T.deu32, &T.omu32, &T.Adv32, &T.Asv32, &T.dev32, &T.omv32, &T.rsu32, &T.rsv32, &T.zsu32, &T.zsv32, &T.gam32};
        for (int q = 0; q < 13; ++q) { *f[q] = T.a + q * n2; *g[q] = T.a32 + q * n2; }
    }
    CG_CHECK(hipMemsetAsync(T.a, 0, 13 * n2 * sizeof(double), S->s));
    // raw Auu/Avv of an assembled matrix: built from the owned rows and the
    // halo rule the first time a line-solve preconditioner needs them
    if (!S->raw_ready) {
        const long long nraw27 = 27LL * d.nz * n2;
        if (!S->stage) CG_CHECK(hipMalloc(&S->stage, 2 * nraw27 * sizeof(double)));
        k_rows_to_raw<<<nblocks(nraw27, 256), 256, 0, S->s>>>(d, 0, S->A64, S->stage);
        k_rows_to_raw<<<nblocks(nraw27, 256), 256, 0, S->s>>>(d, 3, S->A64, S->stage + nraw27);
        S->raw_ready = true;
    }
    // scratch for sums: use rsu.. as sub/diag/sup holders
    double *su_s = T.rsu, *su_d = T.rsv, *su_p = T.zsu, *sv_s = T.zsv, *sv_d = T.gam, *sv_p = T.Asv;
    // stage[0] holds Auu, stage[1] Avv (raw layout)
    const long long nraw = 27LL * d.nz * d.nvx * d.nvy;
    k_tri_vsum<<<nblocks(n2, 256), 256, 0, S->s>>>(d, mode, S->two_d ? 1 : 0, S->stage, S->stage + nraw, S->active, su_s, su_d, su_p,
                                                  sv_s, sv_d, sv_p);
    if (mode == 4) {
        halo2d<double>(S, su_s, su_d);
        halo2d<double>(S, su_p, sv_s);
        halo2d<double>(S, sv_d, sv_p);
    }
    // u lines: As_u = sub (copy), factor into Adu, deu, omu
    CG_CHECK(hipMemcpyAsync(T.Asu, su_s, n2 * sizeof(double), hipMemcpyDeviceToDevice, S->s));
    CG_CHECK(hipMemsetAsync(S->scal + S_DOT2, 0, sizeof(double), S->s));
    k_tri_factor<<<nblocks(d.nyo, 64), 64, 0, S->s>>>(d, 0, mode, T.Asu, su_d, su_p, T.Adu, T.deu, T.omu, S->scal + S_DOT2);
    // v lines (sv_p lives in T.Asv: move it out first)
    double* tmp_p = T.rsu;   // su_s no longer needed
    CG_CHECK(hipMemcpyAsync(tmp_p, sv_p, n2 * sizeof(double), hipMemcpyDeviceToDevice, S->s));
    CG_CHECK(hipMemcpyAsync(T.Asv, sv_s, n2 * sizeof(double), hipMemcpyDeviceToDevice, S->s));
    k_tri_factor<<<nblocks(d.nxo, 64), 64, 0, S->s>>>(d, 1, mode, T.Asv, sv_d, tmp_p, T.Adv, T.dev_, T.omv, S->scal + S_DOT2);
    {
        const int st0 = sync_scal(S);
        if (st0) return st0;
    }
    if (S->scal_h[S_DOT2] != 0.0) { cismgpu_set_error("zero pivot in tridiagonal preconditioner setup"); return CISMGPU_ERR_DIVZERO; }
    // FP32 copies of the factors
    k_cast<double, float><<<nblocks(8 * n2, 256), 256, 0, S->s>>>(8 * n2, T.a, T.a32);
    CG_CHECK(hipMemsetAsync(T.rsu, 0, 5 * n2 * sizeof(double), S->s));
    T.mode = mode;
    return CISMGPU_OK;
}

// precond 3/4 horizontal part: vsum(r) -> line solves -> 2-D halo -> zs
template <typename TV>
void tri_horizontal(cismgpu_system* S, const TV* ru, const TV* rv, TV** zsu_out, TV** zsv_out) {
    const Dims& d = S->d;
    TriPC& T = S->tri;
    TV *Adu, *Asu, *deu, *omu, *Adv, *Asv, *dev_, *omv, *rsu, *rsv, *zsu, *zsv, *gam;
    if constexpr (std::is_same<TV, double>::value) {
        Adu = T.Adu; Asu = T.Asu; deu = T.deu; omu = T.omu; Adv = T.Adv; Asv = T.Asv; dev_ = T.dev_; omv = T.omv;
        rsu = T.rsu; rsv = T.rsv; zsu = T.zsu; zsv = T.zsv; gam = T.gam;
    } else {
        Adu = T.Adu32; Asu = T.Asu32; deu = T.deu32; omu = T.omu32; Adv = T.Adv32; Asv = T.Asv32; dev_ = T.dev32;
        omv = T.omv32; rsu = T.rsu32; rsv = T.rsv32; zsu = T.zsu32; zsv = T.zsv32; gam = T.gam32;
    }
    k_vsum<TV><<<col_blocks(d), 256, 0, S->s>>>(d, ru, rv, rsu, rsv);
    k_line_scan<TV><<<nblocks(d.nyo, 4), 256, 0, S->s>>>(d, 0, Adu, Asu, deu, omu, rsu, zsu, gam);
    k_line_scan<TV><<<nblocks(d.nxo, 4), 256, 0, S->s>>>(d, 1, Adv, Asv, dev_, omv, rsv, zsv, gam);
    halo2d<TV>(S, zsu, zsv);
    *zsu_out = zsu;
    *zsv_out = zsv;
}

// z = PC(r) with partials (r,z), (r,r) in partA
template <typename TV>
void apply_precond(cismgpu_system* S, int precond, const TV* ru, const TV* rv, TV* zu, TV* zv) {
    const Dims& d = S->d;
    const TV* M;
    const TV* Dg;
    if constexpr (std::is_same<TV, double>::value) { M = S->M64; Dg = S->M64 + 6 * d.vlen; }
    else { M = S->M32; Dg = S->M32 + 6 * d.vlen; }
    TV *zsu = nullptr, *zsv = nullptr;
    if (precond >= 3) tri_horizontal<TV>(S, ru, rv, &zsu, &zsv);
    launch_precond<TV>(S, (S->two_d && precond >= 3) ? 5 : precond, ru, rv, zu, zv, M, Dg, zsu, zsv);
}

int finish_timing(cismgpu_system* S, cismgpu_solve_info* info) {
    if (!S->timing) return CISMGPU_OK;
    CG_CHECK(hipStreamSynchronize(S->s));
    double tot = 0.0;
    for (int e = 0; e + 1 < S->ev_used; e += 2) {
        float ms = 0.f;
        CG_CHECK(hipEventElapsedTime(&ms, S->ev[e], S->ev[e + 1]));
        tot += ms;
    }
    if (info) {
        info->spmv_ms += tot;
        info->spmv_count += S->ev_used / 2;
    }
    S->ev_used = 0;
    return CISMGPU_OK;
}

// Chronopoulos-Gear solve, reference semantics (pcg.F90:2632-3594), on
// vectors of type TV with matrix storage TA.  x, r etc. point into the
// system's vector pools.  With `given_r` the initial residual is supplied in
// r (x = 0), as in the inner solve of iterative refinement, and the
// convergence test uses the recurrence residual relative to ||r0||.
template <typename TA, typename TV>
int cg_solve(cismgpu_system* S, const TA* A, TV* const* v, const TV* bu, const TV* bv, int precond, int ncheck,
             double tol, int maxiters, bool given_r, double* err, int* niters, bool fuse_ok) {
    const Dims& d = S->d;
    TV *xu = v[0], *xv = v[1], *ru = v[2], *rv = v[3], *du = v[4], *dv = v[5], *qu = v[6], *qv = v[7];
    TV *zu = v[8], *zv = v[9], *Azu = v[10], *Azv = v[11];
    const TV* M = std::is_same<TV, double>::value ? (const TV*)S->M64 : (const TV*)S->M32;
    const int nsp = spmv_blocks(d), ncb = col_blocks(d);
    const long long nv = d.vlen * sizeof(TV);
    // reprosum: the global sums of the FP64 solve are exact integer-vector
    // sums (the fused compensated partials are then not used)
    const bool rs = S->repro && std::is_same<TV, double>::value && !given_r;
    int st = reset_scal(S);
    if (st) return st;
    for (TV* p : {du, dv, zu, zv, qu, qv, Azu, Azv}) CG_CHECK(hipMemsetAsync(p, 0, nv, S->s));
    *niters = maxiters;
    // L2 norm of the rhs (or of r0 for the inner solve)
    if (rs) {
        rs_dot<TV>(S, 0, bu, bu, bv, bv);
        scalars(S, ST_BB, S->rs_out, 1, 2, nullptr, 0, 0);
    } else {
        k_dot<TV><<<dot_blocks<TV>(d), 256, 0, S->s>>>(d, given_r ? ru : bu, given_r ? ru : bu, given_r ? rv : bv,
                                                       given_r ? rv : bv, S->partB);
        scalars(S, ST_BB, S->partB, dot_blocks<TV>(d), 2, nullptr, 0, 0);
    }
    if (!given_r) {
        CG_CHECK(hipMemsetAsync(ru, 0, nv, S->s));
        CG_CHECK(hipMemsetAsync(rv, 0, nv, S->s));
        halo<TV>(S, xu, xv, d.nz);
        spmv<TA, TV, 2>(S, A, xu, xv, ru, rv, bu, bv, S->partB);
        k_ring<TV><<<col_blocks(d), 256, 0, S->s>>>(d, ru, rv, bu, bv);
    }
    halo<TV>(S, ru, rv, d.nz);
    apply_precond<TV>(S, precond, ru, rv, zu, zv);
    k_axpy_ops<TV><<<nblocks(d.nz * d.plane, 256), 256, 0, S->s>>>(d, 1, S->scal, zu, zv, du, dv, nullptr, nullptr);
    spmv<TA, TV, 1>(S, A, du, dv, qu, qv, nullptr, nullptr, S->partB);
    if (rs) {
        rs_dot<TV>(S, 1, ru, zu, rv, zv);
        rs_dot<TV>(S, 2, du, qu, dv, qv);
        scalars(S, ST_CG_FIRST, S->rs_out + 2, 1, 2, S->rs_out + 4, 1, 2);
    } else {
        scalars(S, ST_CG_FIRST, S->partA, ncb, 4, S->partB, nsp, 2);
    }
    halo<TV>(S, qu, qv, d.nz);
    k_first_update<TV><<<nblocks(d.nz * d.plane, 256), 256, 0, S->s>>>(d, S->scal, xu, xv, ru, rv, du, dv, qu, qv);
    TV* vv[12] = {xu, xv, ru, rv, du, dv, qu, qv, zu, zv, Azu, Azv};
    bool z_ready = false;
    for (int iter = 2; iter <= maxiters; ++iter) {
        if (!z_ready) apply_precond<TV>(S, precond, ru, rv, zu, zv);
        spmv<TA, TV, 1>(S, A, zu, zv, Azu, Azv, nullptr, nullptr, S->partB);
        if (rs) {
            rs_dot<TV>(S, 1, ru, zu, rv, zv);
            rs_dot<TV>(S, 2, Azu, zu, Azv, zv);
            scalars(S, ST_CG_ITER, S->rs_out + 2, 1, 2, S->rs_out + 4, 1, 2);
        } else {
            scalars(S, ST_CG_ITER, S->partA, ncb, 4, S->partB, nsp, 2);
        }
        // The no-ice halo rule zeroes vertex columns 1, 2, nvx-1 and nvx
        // (rows likewise).  The product writes owned rows only, so columns
        // outside the owned range keep the zero they got at the start, and
        // an owned last column holds zero rows after the assembly's edge
        // pass, where A z is +0 already: then the halo pass of A z changes
        // nothing and is skipped.
        if (!(d.bc == 3 && d.ilo >= 3 && d.jlo >= 3 &&
              (S->edge_zero || (d.ihi <= d.nvx - 2 && d.jhi <= d.nvy - 2))))
            halo<TV>(S, Azu, Azv, d.nz);
        const bool check = (iter % ncheck == 0) || iter == 5;
        const bool fuse = fuse_ok && precond == 2 && !(check && !given_r);
        if (fuse) launch_update<TV, 1, 0>(S, vv, M);
        else launch_update<TV, 0, 0>(S, vv, M);
        z_ready = fuse;
        if (check) {
            if (given_r) {
                // recurrence residual (r,r) from the fused update's partials
                if (!fuse) {
                    apply_precond<TV>(S, precond, ru, rv, zu, zv);
                    z_ready = true;
                }
                scalars(S, ST_RR, S->partA + 2, ncb, 4, nullptr, 0, 0, tol);
            } else {
                spmv<TA, TV, 2>(S, A, xu, xv, ru, rv, bu, bv, S->partB);
                k_ring<TV><<<ncb, 256, 0, S->s>>>(d, ru, rv, bu, bv);
                if (rs) {
                    rs_dot<TV>(S, 3, ru, ru, rv, rv);
                    scalars(S, ST_RR, S->rs_out + 6, 1, 2, nullptr, 0, 0, tol);
                } else {
                    scalars(S, ST_RR, S->partB, nsp, 2, nullptr, 0, 0, tol);
                }
            }
            st = sync_scal(S);
            if (st) return st;
            if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("Chron-Gear PCG solver has failed, alpha = NaN"); return CISMGPU_ERR_NAN; }
            *err = S->scal_h[S_ERR];
            if (S->scal_h[S_CONV] != 0.0) { *niters = iter; break; }
            if (!given_r) halo<TV>(S, ru, rv, d.nz);
        }
    }
    st = sync_scal(S);
    if (st) return st;
    if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("Chron-Gear PCG solver has failed, alpha = NaN"); return CISMGPU_ERR_NAN; }
    return CISMGPU_OK;
}

// Standard PCG (pcg.F90:827-1537), FP64.
int std_solve(cismgpu_system* S, int precond, int ncheck, double tol, int maxiters, double* err, int* niters) {
    const Dims& d = S->d;
    double *xu = S->V(0), *xv = S->V(1), *bu = S->V(2), *bv = S->V(3), *ru = S->V(4), *rv = S->V(5);
    double *du = S->V(6), *dv = S->V(7), *qu = S->V(8), *qv = S->V(9), *zu = S->V(10), *zv = S->V(11);
    const int nsp = spmv_blocks(d), ncb = col_blocks(d);
    const long long nv = d.vlen * sizeof(double);
    int st = reset_scal(S);
    if (st) return st;
    for (double* p : {ru, rv, du, dv, zu, zv, qu, qv}) CG_CHECK(hipMemsetAsync(p, 0, nv, S->s));
    halo<double>(S, xu, xv, d.nz);
    spmv<double, double, 2>(S, S->A64, xu, xv, ru, rv, bu, bv, S->partB);
    k_ring<double><<<ncb, 256, 0, S->s>>>(d, ru, rv, bu, bv);
    *niters = maxiters;
    st = set_scal(S, S_ETA0, 1.0);
    if (st) return st;
    const bool rs = S->repro;
    if (rs) {
        rs_dot<double>(S, 0, bu, bu, bv, bv);
        scalars(S, ST_BB, S->rs_out, 1, 2, nullptr, 0, 0);
    } else {
        k_dot<double><<<dot_blocks<double>(d), 256, 0, S->s>>>(d, bu, bu, bv, bv, S->partB);
        scalars(S, ST_BB, S->partB, dot_blocks<double>(d), 2, nullptr, 0, 0);
    }
    double* vv[12] = {xu, xv, ru, rv, du, dv, qu, qv, zu, zv, nullptr, nullptr};
    bool z_ready = false;
    for (int iter = 1; iter <= maxiters; ++iter) {
        if (!z_ready) apply_precond<double>(S, precond, ru, rv, zu, zv);
        if (rs) {
            rs_dot<double>(S, 1, ru, zu, rv, zv);
            scalars(S, ST_STD_ETA1, S->rs_out + 2, 1, 2, nullptr, 0, 0);
        } else {
            scalars(S, ST_STD_ETA1, S->partA, ncb, 4, nullptr, 0, 0);
        }
        k_axpy_ops<double><<<nblocks(d.nz * d.plane, 256), 256, 0, S->s>>>(d, 0, S->scal, zu, zv, du, dv, nullptr, nullptr);
        halo<double>(S, du, dv, d.nz);
        spmv<double, double, 1>(S, S->A64, du, dv, qu, qv, nullptr, nullptr, S->partB);
        k_ring<double><<<ncb, 256, 0, S->s>>>(d, qu, qv, nullptr, nullptr);
        if (rs) {
            rs_dot<double>(S, 2, du, qu, dv, qv);
            scalars(S, ST_STD_ETA2, S->rs_out + 4, 1, 2, nullptr, 0, 0);
        } else {
            scalars(S, ST_STD_ETA2, S->partB, nsp, 2, nullptr, 0, 0);
        }
        const bool check = (iter % ncheck == 0) || iter == 5;
        const bool fuse = precond == 2 && !check;
        if (fuse) launch_update<double, 1, 1>(S, vv, S->M64);
        else launch_update<double, 0, 1>(S, vv, S->M64);
        z_ready = fuse;
        if (check) {
            halo<double>(S, xu, xv, d.nz);
            spmv<double, double, 2>(S, S->A64, xu, xv, ru, rv, bu, bv, S->partB);
            k_ring<double><<<ncb, 256, 0, S->s>>>(d, ru, rv, bu, bv);
            if (rs) {
                rs_dot<double>(S, 3, ru, ru, rv, rv);
                scalars(S, ST_RR, S->rs_out + 6, 1, 2, nullptr, 0, 0, tol);
            } else {
                scalars(S, ST_RR, S->partB, nsp, 2, nullptr, 0, 0, tol);
            }
            st = sync_scal(S);
            if (st) return st;
            if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("PCG solver has failed, alpha = NaN"); return CISMGPU_ERR_NAN; }
            *err = S->scal_h[S_ERR];
            if (S->scal_h[S_CONV] != 0.0) { *niters = iter; break; }
        }
    }
    st = sync_scal(S);
    if (st) return st;
    if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("PCG solver has failed, alpha = NaN"); return CISMGPU_ERR_NAN; }
    return CISMGPU_OK;
}

// FP64 half (symmetric) storage of the owned rows when the FP64 matrix is
// exactly symmetric (3-D systems; CISMGPU_SPMV_FULL=1 keeps full rows only).
int ensure_half64(cismgpu_system* S) {
    if (S->a64h_ready) return CISMGPU_OK;
    S->a64h_ready = true;
    S->a64_sym = false;
    const Dims& d = S->d;
    const char* env = getenv("CISMGPU_SPMV_FULL");
    if (S->two_d || (env && env[0] == '1') || !half_storage_fits(d)) return CISMGPU_OK;
    if (!S->sym_exact) {   // the assembly's symmetry count already showed it otherwise
        unsigned long long* cnt = nullptr;
        unsigned long long h = 1;
        CG_CHECK(hipMalloc(&cnt, sizeof(unsigned long long)));
        CG_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), S->s));
        k_check_sym<double><<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, S->A64, cnt);
        CG_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, S->s));
        CG_CHECK(hipStreamSynchronize(S->s));
        hipFree(cnt);
        if (h != 0) return CISMGPU_OK;
    }
    const long long ns = d.nrows * NSLOT * 4;
    if (!S->A64h) CG_CHECK(hipMalloc(&S->A64h, ns * sizeof(double)));
    k_convert_sym64<<<nblocks(ns, 256), 256, 0, S->s>>>(d.nrows, S->A64, S->A64h);
    S->a64_sym = true;
    return CISMGPU_OK;
}

int ensure_lowp(cismgpu_system* S, int precision) {
    const Dims& d = S->d;
    const long long na = d.nrows * NCOEF;
    if (!S->m32_ready) {
        k_cast<double, float><<<nblocks(8 * d.vlen, 256), 256, 0, S->s>>>(8 * d.vlen, S->M64, S->M32);
        S->m32_ready = true;
    }
    // FP32: half (symmetric) storage when the rounded matrix is exactly
    // symmetric (3-D systems; CISMGPU_SPMV_FULL=1 keeps the quad layout).
    // BF16 keeps the quad layout: its half-storage SpMV is issue-bound
    // (0.70 ms against 0.52 ms for the quad kernel at 1024^2 x 11).
    auto sym_ok = [&](bool& ok) -> int {
        ok = false;
        const char* env = getenv("CISMGPU_SPMV_FULL");
        if (S->two_d || (env && env[0] == '1') || !half_storage_fits(d)) return CISMGPU_OK;
        unsigned long long* cnt = nullptr;
        unsigned long long h = 1;
        CG_CHECK(hipMalloc(&cnt, sizeof(unsigned long long)));
        CG_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), S->s));
        k_check_sym<float><<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, S->A64, cnt);
        CG_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, S->s));
        CG_CHECK(hipStreamSynchronize(S->s));
        hipFree(cnt);
        ok = h == 0;
        return CISMGPU_OK;
    };
    const long long ns = d.nrows * NSLOT * 4;
    if (precision == CISMGPU_MIXED_FP32 && !S->a32_ready) {
        if (!S->A32) CG_CHECK(hipMalloc(&S->A32, na * sizeof(float)));
        const int st = sym_ok(S->a32_sym);
        if (st) return st;
        if (S->a32_sym) k_convert_sym<float><<<nblocks(ns, 256), 256, 0, S->s>>>(d.nrows, S->A64, S->A32);
        else k_convert_quad<float><<<nblocks(na, 256), 256, 0, S->s>>>(d.nrows, S->A64, S->A32);
        S->a32_ready = true;
    }
    if (precision == CISMGPU_MIXED_BF16 && !S->a16_ready) {
        if (!S->A16) CG_CHECK(hipMalloc(&S->A16, na * sizeof(__hip_bfloat16)));
        k_convert_quad<__hip_bfloat16><<<nblocks(na, 256), 256, 0, S->s>>>(d.nrows, S->A64, S->A16);
        S->a16_ready = true;
    }
    if (S->tri.mode && !S->tri.a32) return CISMGPU_ERR_ARG;
    return CISMGPU_OK;
}

// Mixed precision: FP64 iterative refinement around an FP32 (or BF16-matrix)
// Chronopoulos-Gear inner solve.  The outer residual r = b - A x uses the
// FP64 matrix and compensated FP64 sums; the correction e solves A e = r in
// FP32 to inner_tol relative (recurrence residual), then x += e.  The
// refinement stops at the tolerance, after max_outer corrections, or when
// the inner iterations of all corrections reach maxiters.
int mixed_solve(cismgpu_system* S, const cismgpu_solve_opts* o, cismgpu_solve_info* info) {
    const Dims& d = S->d;
    int st = ensure_lowp(S, o->precision);
    if (st) return st;
    double *xu = S->V(0), *xv = S->V(1), *bu = S->V(2), *bv = S->V(3), *ru = S->V(4), *rv = S->V(5);
    float* w[12] = {S->W(0), S->W(1), S->W(4), S->W(5), S->W(6), S->W(7), S->W(8), S->W(9), S->W(10), S->W(11), S->W(12), S->W(13)};
    const int nsp = spmv_blocks(d);
    const long long nvf = d.vlen * sizeof(float);
    // ||b|| once
    st = reset_scal(S);
    if (st) return st;
    if (S->repro) {
        rs_dot<double>(S, 0, bu, bu, bv, bv);
        scalars(S, ST_BB, S->rs_out, 1, 2, nullptr, 0, 0);
    } else {
        k_dot<double><<<dot_blocks<double>(d), 256, 0, S->s>>>(d, bu, bu, bv, bv, S->partB);
        scalars(S, ST_BB, S->partB, dot_blocks<double>(d), 2, nullptr, 0, 0);
    }
    st = sync_scal(S);
    if (st) return st;
    const double l2rhs = S->scal_h[S_L2RHS];
    int inner_total = 0, outer = 0;
    double err = INFINITY;
    const double inner_tol = o->inner_tol > 0 ? o->inner_tol : 1e-4;
    for (outer = 0; outer <= std::max(o->max_outer, 1); ++outer) {
        CG_CHECK(hipMemsetAsync(ru, 0, d.vlen * sizeof(double), S->s));
        CG_CHECK(hipMemsetAsync(rv, 0, d.vlen * sizeof(double), S->s));
        halo<double>(S, xu, xv, d.nz);
        const int tspmv = S->timing;
        S->timing = false;   // time the inner (dominant) SpMV only
        spmv<double, double, 2>(S, S->A64, xu, xv, ru, rv, bu, bv, S->partB);
        S->timing = tspmv;
        k_ring<double><<<col_blocks(d), 256, 0, S->s>>>(d, ru, rv, bu, bv);
        if (S->repro) {
            rs_dot<double>(S, 3, ru, ru, rv, rv);
            k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_DOT, S->rs_out + 6, 1, 2, nullptr, 0, 0, S->scal, 0.0);
        } else {
            k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_DOT, S->partB, nsp, 2, nullptr, 0, 0, S->scal, 0.0);
        }
        st = sync_scal(S);
        if (st) return st;
        err = std::sqrt(S->scal_h[S_DOT]) / l2rhs;
        if (!(err == err)) { cismgpu_set_error("refinement residual is NaN"); return CISMGPU_ERR_NAN; }
        if (err < o->tolerance || outer == std::max(o->max_outer, 1) || inner_total >= o->maxiters) break;
        // inner solve A e = r in FP32
        k_cast<double, float><<<nblocks(d.vlen, 256), 256, 0, S->s>>>(d.vlen, ru, w[2]);
        k_cast<double, float><<<nblocks(d.vlen, 256), 256, 0, S->s>>>(d.vlen, rv, w[3]);
        CG_CHECK(hipMemsetAsync(w[0], 0, nvf, S->s));
        CG_CHECK(hipMemsetAsync(w[1], 0, nvf, S->s));
        double ierr = 0.0;
        int nit = 0;
        // maxiters bounds the inner iterations of the whole solve, as it
        // bounds the iterations of the FP64 solve: each inner solve may use
        // what the earlier ones left
        const int budget = o->maxiters - inner_total;
        if (o->precision == CISMGPU_MIXED_BF16)
            st = cg_solve<__hip_bfloat16, float>(S, S->A16, w, nullptr, nullptr, o->precond, o->ncheck, inner_tol,
                                                 budget, true, &ierr, &nit, true);
        else
            st = cg_solve<float, float>(S, S->A32, w, nullptr, nullptr, o->precond, o->ncheck, inner_tol,
                                        budget, true, &ierr, &nit, true);
        if (st) return st;
        inner_total += nit;
        k_ir_update<<<nblocks(d.nz * d.plane, 256), 256, 0, S->s>>>(d, xu, xv, w[0], w[1]);
    }
    info->err = err;
    info->niters = inner_total;
    info->outer_iters = outer;
    info->spmv_kind = o->precision == CISMGPU_MIXED_BF16 ? 2 : (S->a32_sym ? 3 : 1);
    if (!(err < o->tolerance)) {
        cismgpu_set_error("iterative refinement stopped at relative residual %.3e > %.3e", err, o->tolerance);
        return CISMGPU_ERR_NOT_CONVERGED;
    }
    return CISMGPU_OK;
}

int check_opts(const cismgpu_solve_opts* o) {
    if (!o || o->ncheck < 1 || o->maxiters < 1 || o->precond < 0 || o->precond > 4 ||
        (o->algo != CISMGPU_ALGO_STANDARD && o->algo != CISMGPU_ALGO_CHRONGEAR) || o->precision < 0 ||
        o->precision > 2) {
        cismgpu_set_error("bad solve options");
        return CISMGPU_ERR_ARG;
    }
    if (o->precision != CISMGPU_FP64 && (o->algo != CISMGPU_ALGO_CHRONGEAR)) {
        cismgpu_set_error("mixed precision uses the Chronopoulos-Gear inner solver");
        return CISMGPU_ERR_ARG;
    }
    return CISMGPU_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

static int32_t system_create(const cismgpu_grid* grid, bool two_d, cismgpu_system** out) {
    if (!out) return CISMGPU_ERR_ARG;
    *out = nullptr;
    int st = check_grid(grid, two_d);
    if (st) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        cismgpu_set_error("no HIP device available");
        return CISMGPU_ERR_HIP;
    }
    cismgpu_system* S = new cismgpu_system();
    S->d = make_dims(grid);
    S->two_d = two_d;
    const Dims& d = S->d;
    auto fail = [&](hipError_t e, const char* what) {
        cismgpu_set_error("%s: %s", what, hipGetErrorString(e));
        cismgpu_system_destroy(S);
        return CISMGPU_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    const long long n2 = (long long)d.nvx * d.nvy;
    const long long na = d.nrows * NCOEF;
    const int nparts = std::max(spmv_blocks(d), std::max(col_blocks(d), dot_blocks<double>(d))) + 8;
    if ((e = hipMalloc(&S->active, n2)) != hipSuccess) return fail(e, "active");
    if ((e = hipMalloc(&S->A64, na * sizeof(double))) != hipSuccess) return fail(e, "A64");
    if ((e = hipMalloc(&S->M64, 8 * d.vlen * sizeof(double))) != hipSuccess) return fail(e, "M64");
    if ((e = hipMalloc(&S->M32, 8 * d.vlen * sizeof(float))) != hipSuccess) return fail(e, "M32");
    if ((e = hipMalloc(&S->v64, 14 * d.vlen * sizeof(double))) != hipSuccess) return fail(e, "v64");
    if ((e = hipMalloc(&S->v32, 14 * d.vlen * sizeof(float))) != hipSuccess) return fail(e, "v32");
    if ((e = hipMalloc(&S->partA, 4 * nparts * sizeof(double))) != hipSuccess) return fail(e, "partA");
    if ((e = hipMalloc(&S->partB, 4 * nparts * sizeof(double))) != hipSuccess) return fail(e, "partB");
    if ((e = hipMalloc(&S->scal, S_NSLOTS * sizeof(double))) != hipSuccess) return fail(e, "scal");
    if ((e = hipHostMalloc(&S->scal_h, S_NSLOTS * sizeof(double), hipHostMallocDefault)) != hipSuccess) return fail(e, "scal_h");
    S->repro = grid->reprosum != 0;
    if (S->repro) {
        if ((e = hipMalloc(&S->rs_part, (size_t)RS_MAX_BLOCKS * RS_L * sizeof(unsigned long long))) != hipSuccess)
            return fail(e, "rs_part");
        if ((e = hipMalloc(&S->rs_bad, sizeof(int))) != hipSuccess) return fail(e, "rs_bad");
        if ((e = hipMalloc(&S->rs_out, 8 * sizeof(double))) != hipSuccess) return fail(e, "rs_out");
    }
    hipMemsetAsync(S->v64, 0, 14 * d.vlen * sizeof(double), S->s);
    hipMemsetAsync(S->v32, 0, 14 * d.vlen * sizeof(float), S->s);
    hipMemsetAsync(S->M64, 0, 8 * d.vlen * sizeof(double), S->s);
    hipMemsetAsync(S->A64, 0, na * sizeof(double), S->s);
    hipMemsetAsync(S->active, 0, n2, S->s);
    if ((e = hipStreamSynchronize(S->s)) != hipSuccess) return fail(e, "init");
    *out = S;
    return CISMGPU_OK;
}

int32_t cismgpu_system_create(const cismgpu_grid* grid, cismgpu_system** out) {
    return system_create(grid, false, out);
}

int32_t cismgpu_system_create_2d(const cismgpu_grid* grid, cismgpu_system** out) {
    if (!grid) { cismgpu_set_error("null grid"); return CISMGPU_ERR_ARG; }
    cismgpu_grid g = *grid;
    g.nz = 1;
    const int st = system_create(&g, true, out);
    if (!st) (*out)->nz_levels = grid->nz;
    return st;
}

int32_t cismgpu_system_destroy(cismgpu_system* S) {
    if (!S) return CISMGPU_OK;
    if (S->s) hipStreamSynchronize(S->s);
    for (void* p : {(void*)S->active, (void*)S->A64, (void*)S->A32, (void*)S->A16, (void*)S->stage, (void*)S->M64,
                    (void*)S->M32, (void*)S->v64, (void*)S->v32, (void*)S->partA, (void*)S->partB, (void*)S->scal,
                    (void*)S->tri.a, (void*)S->tri.a32, S->asm_buf, (void*)S->efvs, (void*)S->Atmp, (void*)S->saved, (void*)S->work, (void*)S->load3, (void*)S->A64h, (void*)S->diva_buf, (void*)S->efva2, (void*)S->rs_part, (void*)S->rs_bad,
                    (void*)S->rs_out, (void*)S->elemq, S->cb_buf})
        if (p) hipFree(p);
    if (S->scal_h) hipHostFree(S->scal_h);
    for (auto& x : S->ev) hipEventDestroy(x);
    if (S->s) hipStreamDestroy(S->s);
    delete S;
    return CISMGPU_OK;
}

// Matrix, rhs and mask into the device layouts.  where_*: 0 host, 1 device.
static int set_matrix_impl(cismgpu_system* S, const int32_t* active_vertex, int where_act, const double* Auu,
                           const double* Auv, const double* Avu, const double* Avv, int where_A, const double* bu,
                           const double* bv, int where_b) {
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy;
    const long long nraw = 27LL * d.nz * n2;
    const long long nvec = (long long)d.nz * n2;
    const hipMemcpyKind kind = where_A ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kindb = where_b ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (!S->stage) CG_CHECK(hipMalloc(&S->stage, 2 * nraw * sizeof(double)));
    // active mask (int32 -> int8)
    std::vector<int8_t> act(n2);
    if (where_act) {
        std::vector<int32_t> tmp(n2);
        CG_CHECK(hipMemcpy(tmp.data(), active_vertex, n2 * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (long long i = 0; i < n2; ++i) act[i] = tmp[i] ? 1 : 0;
    } else {
        for (long long i = 0; i < n2; ++i) act[i] = active_vertex[i] ? 1 : 0;
    }
    CG_CHECK(hipMemcpyAsync(S->active, act.data(), n2, hipMemcpyHostToDevice, S->s));
    double* s0 = S->stage;
    double* s1 = S->stage + nraw;
    const unsigned gp = nblocks(d.nrows * 27, 256);
    // Auv, Avu first through the staging pair, then Auu/Avv which stay staged
    CG_CHECK(hipMemcpyAsync(s0, Auv, nraw * sizeof(double), kind, S->s));
    CG_CHECK(hipMemcpyAsync(s1, Avu, nraw * sizeof(double), kind, S->s));
    k_pack_matrix<double><<<gp, 256, 0, S->s>>>(d, 1, s0, S->active, S->A64);
    k_pack_matrix<double><<<gp, 256, 0, S->s>>>(d, 2, s1, S->active, S->A64);
    CG_CHECK(hipMemcpyAsync(s0, Auu, nraw * sizeof(double), kind, S->s));
    CG_CHECK(hipMemcpyAsync(s1, Avv, nraw * sizeof(double), kind, S->s));
    k_pack_matrix<double><<<gp, 256, 0, S->s>>>(d, 0, s0, S->active, S->A64);
    k_pack_matrix<double><<<gp, 256, 0, S->s>>>(d, 3, s1, S->active, S->A64);
    k_setup_precond<double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, s0, s1, S->M64, S->M64 + 6 * d.vlen);
    // rhs
    double* tmp = S->V(4);   // r vectors as staging for b
    CG_CHECK(hipMemcpyAsync(tmp, bu, nvec * sizeof(double), kindb, S->s));
    k_vec_in<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, tmp, S->V(2));
    CG_CHECK(hipMemcpyAsync(tmp, bv, nvec * sizeof(double), kindb, S->s));
    k_vec_in<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, tmp, S->V(3));
    CG_CHECK(hipMemsetAsync(tmp, 0, d.vlen * sizeof(double), S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    S->have_matrix = true;
    S->raw_ready = true;
    S->sym_exact = false;
    S->edge_zero = false;
    S->m32_ready = S->a32_ready = S->a16_ready = S->a64h_ready = false;
    S->a32_sym = S->a64_sym = false;
    S->tri.mode = 0;
    return CISMGPU_OK;
}

int32_t cismgpu_system_set_matrix(cismgpu_system* S, const int32_t* active_vertex, const double* Auu,
                                  const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                  const double* bv, int32_t where) {
    if (!S || !active_vertex || !Auu || !Auv || !Avu || !Avv || !bu || !bv) return CISMGPU_ERR_ARG;
    if (S->two_d) { cismgpu_set_error("a 2-D system takes its matrix through set_matrix_2d"); return CISMGPU_ERR_ARG; }
    return set_matrix_impl(S, active_vertex, where, Auu, Auv, Avu, Avv, where, bu, bv, where);
}

int32_t cismgpu_system_set_matrix_2d(cismgpu_system* S, const int32_t* active_vertex, const double* Auu,
                                     const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                     const double* bv, int32_t where) {
    if (!S || !active_vertex || !Auu || !Auv || !Avu || !Avv || !bu || !bv) return CISMGPU_ERR_ARG;
    if (!S->two_d) { cismgpu_set_error("set_matrix_2d needs a system made by cismgpu_system_create_2d"); return CISMGPU_ERR_ARG; }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy;
    // each 9-point block becomes the kA = 0 plane of a one-level 27-point row
    double* buf = nullptr;
    CG_CHECK(hipMalloc(&buf, (4 * 27 + 9) * n2 * sizeof(double)));
    double* in9 = buf + 4 * 27 * n2;
    const double* src[4] = {Auu, Auv, Avu, Avv};
    const hipMemcpyKind kind = where ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    int st = CISMGPU_OK;
    for (int b = 0; b < 4 && !st; ++b) {
        if (hipMemcpyAsync(in9, src[b], 9 * n2 * sizeof(double), kind, S->s) != hipSuccess) st = CISMGPU_ERR_HIP;
        else k_expand_stencil_2d<<<nblocks(27 * n2, 256), 256, 0, S->s>>>(n2, in9, buf + b * 27 * n2);
    }
    if (!st && hipStreamSynchronize(S->s) != hipSuccess) st = CISMGPU_ERR_HIP;
    if (st) {
        cismgpu_set_error("set_matrix_2d: device copy failed");
        hipFree(buf);
        return st;
    }
    // vectors (nvx, nvy) are the one-level (1, nvx, nvy) layout
    st = set_matrix_impl(S, active_vertex, where, buf, buf + 27 * n2, buf + 54 * n2, buf + 81 * n2, 1, bu, bv, where);
    hipFree(buf);
    return st;
}

int32_t cismgpu_system_get_matrix(cismgpu_system* S, double* Auu, double* Auv, double* Avu, double* Avv,
                                  double* bu, double* bv) {
    if (!S || !S->have_matrix) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy;
    const long long nraw = 27LL * d.nz * n2, nvec = (long long)d.nz * n2;
    double* tmp = nullptr;
    CG_CHECK(hipMalloc(&tmp, nraw * sizeof(double)));
    double* outs[4] = {Auu, Auv, Avu, Avv};
    int st = CISMGPU_OK;
    for (int b = 0; b < 4 && !st; ++b) {
        if (!outs[b]) continue;
        k_unpack_matrix<<<nblocks(nraw, 256), 256, 0, S->s>>>(d, b, S->A64, tmp);
        if (hipMemcpyAsync(outs[b], tmp, nraw * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
            hipStreamSynchronize(S->s) != hipSuccess)
            st = CISMGPU_ERR_HIP;
    }
    double* vo[2] = {bu, bv};
    for (int c = 0; c < 2 && !st; ++c) {
        if (!vo[c]) continue;
        k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, S->V(2 + c), tmp);
        if (hipMemcpyAsync(vo[c], tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
            hipStreamSynchronize(S->s) != hipSuccess)
            st = CISMGPU_ERR_HIP;
    }
    hipFree(tmp);
    if (st) cismgpu_set_error("get_matrix: device copy failed");
    return st;
}

int32_t cismgpu_system_set_x(cismgpu_system* S, const double* xu, const double* xv) {
    if (!S || !xu || !xv) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    const long long nvec = (long long)d.nz * d.nvx * d.nvy;
    double* tmp = S->V(12);
    CG_CHECK(hipMemsetAsync(S->V(0), 0, 2 * d.vlen * sizeof(double), S->s));
    CG_CHECK(hipMemcpyAsync(tmp, xu, nvec * sizeof(double), hipMemcpyHostToDevice, S->s));
    k_vec_in<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, tmp, S->V(0));
    CG_CHECK(hipMemcpyAsync(tmp, xv, nvec * sizeof(double), hipMemcpyHostToDevice, S->s));
    k_vec_in<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, tmp, S->V(1));
    CG_CHECK(hipMemsetAsync(tmp, 0, d.vlen * sizeof(double), S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

int32_t cismgpu_system_get_x(cismgpu_system* S, double* xu, double* xv) {
    if (!S || !xu || !xv) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    const long long nvec = (long long)d.nz * d.nvx * d.nvy;
    double* tmp = S->V(12);
    k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, S->V(0), tmp);
    CG_CHECK(hipMemcpyAsync(xu, tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, S->V(1), tmp);
    CG_CHECK(hipMemcpyAsync(xv, tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    CG_CHECK(hipMemsetAsync(tmp, 0, d.vlen * sizeof(double), S->s));
    return CISMGPU_OK;
}

int32_t cismgpu_system_solve(cismgpu_system* S, const cismgpu_solve_opts* o, cismgpu_solve_info* info) {
    if (!S || !S->have_matrix || !info) { cismgpu_set_error("system has no matrix"); return CISMGPU_ERR_ARG; }
    int st = check_opts(o);
    if (st) return st;
    cismgpu_solve_opts eff = *o;
    // one-level systems have no column part: which_ho_precond = 2 becomes the
    // diagonal preconditioner (glide_setup.F90:1363-1366)
    if (S->two_d && eff.precond == 2) eff.precond = 1;
    o = &eff;
    memset(info, 0, sizeof(*info));
    info->err = NAN;
    // per-assembly set-up, done by the first solve after an assembly or a
    // new matrix and cached until the next: the line-solve factors, the
    // exact-symmetry checks and the half-storage / low-precision copies
    {
        const bool fresh = ((o->precond == 3 || o->precond == 4) && S->tri.mode != o->precond) || !S->a64h_ready ||
                           (o->precision == CISMGPU_MIXED_FP32 && !S->a32_ready) ||
                           (o->precision == CISMGPU_MIXED_BF16 && !S->a16_ready) ||
                           (o->precision != CISMGPU_FP64 && !S->m32_ready);
        const auto t0 = std::chrono::steady_clock::now();
        if ((o->precond == 3 || o->precond == 4) && S->tri.mode != o->precond) {
            st = setup_tri(S, o->precond);
            if (st) return st;
        }
        st = ensure_half64(S);
        if (st) return st;
        if (o->precision != CISMGPU_FP64) {
            st = ensure_lowp(S, o->precision);
            if (st) return st;
        }
        if (fresh) {
            CG_CHECK(hipStreamSynchronize(S->s));
            info->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    S->timing = o->time_spmv != 0;
    if (S->timing && S->ev.empty()) {
        S->ev.resize(8192);
        for (auto& x : S->ev) hipEventCreate(&x);
    }
    S->ev_used = 0;
    hipEvent_t e0, e1;
    CG_CHECK(hipEventCreate(&e0));
    CG_CHECK(hipEventCreate(&e1));
    CG_CHECK(hipEventRecord(e0, S->s));
    if (o->precision == CISMGPU_FP64) {
        double err = NAN;
        int nit = 0;
        if (o->algo == CISMGPU_ALGO_CHRONGEAR) {
            double* v[12] = {S->V(0), S->V(1), S->V(4), S->V(5), S->V(6), S->V(7), S->V(8), S->V(9), S->V(10), S->V(11), S->V(12), S->V(13)};
            st = cg_solve<double, double>(S, S->A64, v, S->V(2), S->V(3), o->precond, o->ncheck, o->tolerance,
                                          o->maxiters, false, &err, &nit, true);
        } else {
            st = std_solve(S, o->precond, o->ncheck, o->tolerance, o->maxiters, &err, &nit);
        }
        info->err = err;
        info->niters = nit;
        info->spmv_kind = S->a64_sym ? 4 : 0;
    } else {
        st = mixed_solve(S, o, info);
    }
    hipEventRecord(e1, S->s);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    info->solve_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    int st2 = finish_timing(S, info);
    S->timing = false;
    return st ? st : st2;
}

int32_t cismgpu_system_matvec(cismgpu_system* S, double* yu, double* yv) {
    if (!S || !S->have_matrix || !yu || !yv) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    double *y0 = S->V(12), *y1 = S->V(13);
    spmv<double, double, 0>(S, S->A64, S->V(0), S->V(1), y0, y1, nullptr, nullptr, nullptr);
    k_ring<double><<<col_blocks(d), 256, 0, S->s>>>(d, y0, y1, nullptr, nullptr);
    const long long nvec = (long long)d.nz * d.nvx * d.nvy;
    double* tmp = S->V(10);
    k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, y0, tmp);
    CG_CHECK(hipMemcpyAsync(yu, tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, y1, tmp);
    CG_CHECK(hipMemcpyAsync(yv, tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    CG_CHECK(hipMemsetAsync(tmp, 0, d.vlen * sizeof(double), S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

int32_t cismgpu_system_bench_spmv(cismgpu_system* S, int32_t kind, int32_t nrep, double* avg_ms,
                                  double* bytes_per_launch) {
    if (!S || !S->have_matrix || nrep < 1 || kind < 0 || kind > 2) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    const long long nodes = (long long)d.nz * d.nxo * d.nyo;
    int st = CISMGPU_OK;
    if (kind > 0) {
        st = ensure_lowp(S, kind);
        if (st) return st;
        k_cast<double, float><<<nblocks(2 * d.vlen, 256), 256, 0, S->s>>>(2 * d.vlen, S->V(0), S->W(10));
    } else {
        st = ensure_half64(S);
        if (st) return st;
    }
    hipEvent_t e0, e1;
    CG_CHECK(hipEventCreate(&e0));
    CG_CHECK(hipEventCreate(&e1));
    // the form the solvers launch: y = A x with the fused (y, x) partial sums
    auto launch = [&]() {
        if (S->two_d) {
            if (kind == 0)
                k_spmv_plane<double, double, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A64, S->V(0), S->V(1), S->V(12), S->V(13), nullptr, nullptr, S->partB);
            else if (kind == 1)
                k_spmv_plane<float, float, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A32, S->W(10), S->W(11), S->W(12), S->W(13), nullptr, nullptr, S->partB);
            else
                k_spmv_plane<__hip_bfloat16, float, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A16, S->W(10), S->W(11), S->W(12), S->W(13), nullptr, nullptr, S->partB);
        } else if (kind == 0 && S->a64_sym)
            k_spmv_sym64<1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A64h, S->V(0), S->V(1), S->V(12), S->V(13), nullptr, nullptr, S->partB);
        else if (kind == 0)
            k_spmv<double, double, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A64, S->V(0), S->V(1), S->V(12), S->V(13), nullptr, nullptr, S->partB);
        else if (kind == 1 && S->a32_sym)
            spmv_sym32<1>(S, S->A32, S->W(10), S->W(11), S->W(12), S->W(13), nullptr, nullptr, S->partB);
        else if (kind == 1)
            k_spmv_q<float, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A32, S->W(10), S->W(11), S->W(12), S->W(13), nullptr, nullptr, S->partB);
        else
            k_spmv_q<__hip_bfloat16, 1><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A16, S->W(10), S->W(11), S->W(12), S->W(13), nullptr, nullptr, S->partB);
    };
    launch();   // warm-up
    CG_CHECK(hipEventRecord(e0, S->s));
    for (int r = 0; r < nrep; ++r) launch();
    CG_CHECK(hipEventRecord(e1, S->s));
    CG_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (avg_ms) *avg_ms = ms / nrep;
    const double esz = kind == 0 ? 8.0 : (kind == 1 ? 4.0 : 2.0);
    const double vsz = kind == 0 ? 8.0 : 4.0;
    const bool half = (kind == 1 && S->a32_sym) || (kind == 0 && S->a64_sym);
    // 4 blocks x 9 or x 27 neighbours; half storage: 14 slots of 4
    const double ncoef = S->two_d ? 36.0 : (half ? 4.0 * NSLOT : (double)NCOEF);
    if (bytes_per_launch) *bytes_per_launch = (double)nodes * (ncoef * esz + 4.0 * vsz);
    return CISMGPU_OK;
}

// cism_reprosum_calc for one task (cism_reprosum_mod.F90:237-1057): the
// exact, order-independent sum of each field of arr(nsummands, nflds),
// correctly rounded.  Host arrays; work on the device.
extern "C" int32_t cismgpu_reprosum_calc(const double* arr, int64_t nsummands, int32_t nflds, double* arr_gsum) {
    if (!arr || !arr_gsum || nsummands < 0 || nflds < 1) {
        cismgpu_set_error("reprosum_calc: bad arguments");
        return CISMGPU_ERR_ARG;
    }
    if (nsummands > (1LL << 32)) {
        cismgpu_set_error("reprosum_calc: %lld summands exceed the 2^32 the digit headroom allows", (long long)nsummands);
        return CISMGPU_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        cismgpu_set_error("no HIP device available");
        return CISMGPU_ERR_HIP;
    }
    hipStream_t s = nullptr;
    double *a = nullptr, *out = nullptr;
    unsigned long long* part = nullptr;
    int* bad = nullptr;
    int st = CISMGPU_OK;
    auto run = [&]() -> int {
        CG_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        CG_CHECK(hipMalloc(&a, std::max<int64_t>(nsummands, 1) * sizeof(double)));
        CG_CHECK(hipMalloc(&out, 2 * (size_t)nflds * sizeof(double)));
        CG_CHECK(hipMalloc(&part, (size_t)RS_MAX_BLOCKS * RS_L * sizeof(unsigned long long)));
        CG_CHECK(hipMalloc(&bad, sizeof(int)));
        const int nb = std::max(1, std::min(RS_MAX_BLOCKS, (int)nblocks(nsummands, RS_THREADS)));
        for (int f = 0; f < nflds; ++f) {
            if (nsummands)
                CG_CHECK(hipMemcpyAsync(a, arr + (size_t)f * nsummands, nsummands * sizeof(double),
                                        hipMemcpyHostToDevice, s));
            k_fill_int<<<1, 64, 0, s>>>(bad, 1, 0);
            k_rs_levels_flat<<<nb, RS_THREADS, 0, s>>>(nsummands, a, part, bad);
            k_rs_final<<<1, RS_THREADS, 0, s>>>(part, nb, bad, out + 2 * f);
        }
        std::vector<double> h(2 * (size_t)nflds);
        CG_CHECK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        CG_CHECK(hipStreamSynchronize(s));
        for (int f = 0; f < nflds; ++f) arr_gsum[f] = h[2 * f];
        return CISMGPU_OK;
    };
    st = run();
    if (s) hipStreamSynchronize(s);
    for (void* p : {(void*)a, (void*)out, (void*)part, (void*)bad})
        if (p) hipFree(p);
    if (s) hipStreamDestroy(s);
    return st;
}

extern "C" int32_t cismgpu_half_storage_supported(const cismgpu_grid* grid) {
    if (check_grid(grid)) return 0;
    const Dims d = make_dims(grid);
    return half_storage_fits(d) ? 1 : 0;
}

// ---------------------------------------------------------------- HBM probes (STREAM-style)
// Grid-stride kernels over double2 elements, sized to a few waves per SIMD:
// copy b = a, triad a = b + s c, read (sum of a into per-workgroup partials).
__global__ __launch_bounds__(256) void k_stream_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                                     long long n2) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n2; t += (long long)gridDim.x * blockDim.x)
        b[t] = a[t];
}
__global__ __launch_bounds__(256) void k_stream_triad(double2* __restrict__ a, const double2* __restrict__ b,
                                                      const double2* __restrict__ c, double s, long long n2) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n2; t += (long long)gridDim.x * blockDim.x) {
        const double2 x = b[t], y = c[t];
        a[t] = make_double2(x.x + s * y.x, x.y + s * y.y);
    }
}
__global__ __launch_bounds__(256) void k_stream_read(const double2* __restrict__ a, long long n2,
                                                     double* __restrict__ part) {
    double acc = 0.0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n2; t += (long long)gridDim.x * blockDim.x) {
        const double2 x = a[t];
        acc += x.x + x.y;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, WAVE);
    if (threadIdx.x % WAVE == 0) part[blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE] = acc;
}

extern "C" int32_t cismgpu_bench_stream(int64_t nbytes, int32_t nrep, double* copy_gbps, double* triad_gbps,
                                        double* read_gbps) {
    if (nbytes < (1 << 20) || nrep < 1) return CISMGPU_ERR_ARG;
    const long long n2 = nbytes / 16;
    const size_t bytes = (size_t)n2 * 16;
    double2 *a = nullptr, *b = nullptr, *c = nullptr;
    double* part = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0, ncu = 0;
    int rc = CISMGPU_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == CISMGPU_OK) {
            cismgpu_set_error("bench_stream: %s -> %s", what, hipGetErrorString(e));
            rc = CISMGPU_ERR_HIP;
        }
        return rc == CISMGPU_OK;
    };
    if (chk(hipGetDevice(&dev), "device") &&
        chk(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "attr") &&
        chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") && chk(hipMalloc(&a, bytes), "alloc") &&
        chk(hipMalloc(&b, bytes), "alloc") && chk(hipMalloc(&c, bytes), "alloc") &&
        chk(hipEventCreate(&e0), "event") && chk(hipEventCreate(&e1), "event")) {
        const unsigned grid = (unsigned)ncu * 8;
        if (chk(hipMalloc(&part, (size_t)grid * 4 * sizeof(double)), "alloc") &&
            chk(hipMemsetAsync(a, 0, bytes, s), "memset") && chk(hipMemsetAsync(b, 0, bytes, s), "memset") &&
            chk(hipMemsetAsync(c, 0, bytes, s), "memset")) {
            double* outs[3] = {copy_gbps, triad_gbps, read_gbps};
            const double moved[3] = {2.0 * bytes, 3.0 * bytes, 1.0 * bytes};
            for (int kind = 0; kind < 3 && rc == CISMGPU_OK; ++kind) {
                auto launch = [&]() {
                    if (kind == 0) k_stream_copy<<<grid, 256, 0, s>>>(a, b, n2);
                    else if (kind == 1) k_stream_triad<<<grid, 256, 0, s>>>(a, b, c, 3.0, n2);
                    else k_stream_read<<<grid, 256, 0, s>>>(a, n2, part);
                };
                launch();
                chk(hipEventRecord(e0, s), "record");
                for (int r = 0; r < nrep; ++r) launch();
                chk(hipEventRecord(e1, s), "record");
                if (!chk(hipEventSynchronize(e1), "sync")) break;
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                if (outs[kind]) *outs[kind] = moved[kind] / ((double)ms / nrep * 1e-3) / 1e9;
            }
        }
    }
    if (s) hipStreamSynchronize(s);
    for (void* p : {(void*)a, (void*)b, (void*)c, (void*)part})
        if (p) hipFree(p);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    return rc;
}

// ---------------------------------------------------------------- assembly
static bool g_tables_ready = false;

static int upload_fe_tables() {
    if (g_tables_ready) return CISMGPU_OK;
    const double r = 1.0 / std::sqrt(3.0);
    const double xq[8] = {-r, r, r, -r, -r, r, r, -r};
    const double yq[8] = {-r, -r, r, r, -r, -r, r, r};
    const double zq[8] = {-r, -r, -r, -r, r, r, r, r};
    const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double dxr[8][8], dyr[8][8], dzr[8][8];
    for (int p = 0; p < 8; ++p)
        for (int n = 0; n < 8; ++n) {
            dxr[n][p] = sx[n] * (1.0 + sy[n] * yq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
            dyr[n][p] = sy[n] * (1.0 + sx[n] * xq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
            dzr[n][p] = sz[n] * (1.0 + sx[n] * xq[p]) * (1.0 + sy[n] * yq[p]) / 8.0;
        }
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dxr), dxr, sizeof(dxr)));
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dyr), dyr, sizeof(dyr)));
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dzr), dzr, sizeof(dzr)));
    double phi3[8][8];
    for (int p = 0; p < 8; ++p)
        for (int n = 0; n < 8; ++n)
            phi3[n][p] = (1.0 + sx[n] * xq[p]) * (1.0 + sy[n] * yq[p]) * (1.0 + sz[n] * zq[p]) / 8.0;
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_phi3), phi3, sizeof(phi3)));
    // 2-D element (glissade_velo_higher.F90:528-570): points (-r,-r), (r,-r), (r,r), (-r,r)
    const double xq2[4] = {-r, r, r, -r}, yq2[4] = {-r, -r, r, r};
    double phi2[4][4], dxr2[4][4], dyr2[4][4];
    for (int p = 0; p < 4; ++p) {
        phi2[0][p] = (1.0 - xq2[p]) * (1.0 - yq2[p]) / 4.0;
        phi2[1][p] = (1.0 + xq2[p]) * (1.0 - yq2[p]) / 4.0;
        phi2[2][p] = (1.0 + xq2[p]) * (1.0 + yq2[p]) / 4.0;
        phi2[3][p] = (1.0 - xq2[p]) * (1.0 + yq2[p]) / 4.0;
        dxr2[0][p] = -(1.0 - yq2[p]) / 4.0;
        dxr2[1][p] = (1.0 - yq2[p]) / 4.0;
        dxr2[2][p] = (1.0 + yq2[p]) / 4.0;
        dxr2[3][p] = -(1.0 + yq2[p]) / 4.0;
        dyr2[0][p] = -(1.0 - xq2[p]) / 4.0;
        dyr2[1][p] = -(1.0 + xq2[p]) / 4.0;
        dyr2[2][p] = (1.0 + xq2[p]) / 4.0;
        dyr2[3][p] = (1.0 - xq2[p]) / 4.0;
    }
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_phi2), phi2, sizeof(phi2)));
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dxr2), dxr2, sizeof(dxr2)));
    CG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dyr2), dyr2, sizeof(dyr2)));
    g_tables_ready = true;
    return CISMGPU_OK;
}

// the assembly has a basal friction term (host field or the device calcbeta result)
static inline bool has_beta(const cismgpu_assembly_in* in) { return in->beta_internal || in->beta_device; }

// Inputs of an assembly into the system's device block.  d describes the
// input fields: the system's grid for a 3-D system, the grid with the
// input's sigma levels for a 2-D one.
static int upload_assembly_inputs(cismgpu_system* S, const Dims& d, const cismgpu_assembly_in* in, AsmDev& a,
                                  LatDev& lat, bool& lateral) {
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny;
    const long long nvec = (long long)d.nz * n2, nfl = (long long)(d.nz - 1) * nc;
    // device input block: doubles first, then int32s
    lateral = in->ocean_mask != nullptr;
    if (in->beta_device && !S->cb_valid) {
        cismgpu_set_error("assembly: beta_device needs a cismgpu_system_calcbeta result on this system");
        return CISMGPU_ERR_ARG;
    }
    if (lateral && (!in->usrf || !in->thck ||
                    (in->which_assemble_lateral == 0 && (!in->stagusrf_lateral || !in->stagthck_lateral)))) {
        cismgpu_set_error("assembly: lateral BC needs usrf, thck and (standard option) the staggered fields");
        return CISMGPU_ERR_ARG;
    }
    const size_t nd = (size_t)(d.nz + (d.nz + 1) + 8 * n2 + nfl + (lateral ? nc + 2 * n2 : 0) + (in->thck ? nc : 0) +
                             (in->flwa ? nfl : 0) + (in->stag_lsrf ? n2 : 0)),
                 ni = (size_t)(nc + n2 + 2 * nvec + (lateral ? nc : 0));
    const size_t bytes = nd * sizeof(double) + ni * sizeof(int32_t);
    if (S->asm_bytes < bytes) {
        if (S->asm_buf) hipFree(S->asm_buf);
        S->asm_buf = nullptr;
        CG_CHECK(hipMalloc(&S->asm_buf, bytes));
        S->asm_bytes = bytes;
    }
    if (!S->efvs) CG_CHECK(hipMalloc(&S->efvs, nfl * sizeof(double)));
    double* db = (double*)S->asm_buf;
    int32_t* ib = (int32_t*)(db + nd);
    double* p = db;
    hipError_t herr = hipSuccess;   // first failure of the copies below
    auto keep = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    auto putd = [&](const double* h, long long n, bool zero_if_null) -> const double* {
        double* dst = p;
        p += n;
        if (h) keep(hipMemcpyAsync(dst, h, n * sizeof(double), hipMemcpyHostToDevice, S->s));
        else if (zero_if_null) keep(hipMemsetAsync(dst, 0, n * sizeof(double), S->s));
        return dst;
    };
    a.sigma = putd(in->sigma, d.nz, true);
    a.swbs = putd(in->stagwbndsigma, d.nz + 1, true);
    a.xV = putd(in->xVertex, n2, true);
    a.yV = putd(in->yVertex, n2, true);
    a.susrf = putd(in->stagusrf, n2, true);
    a.sthck = putd(in->stagthck, n2, true);
    a.dsdx = putd(in->dusrf_dx, n2, true);
    a.dsdy = putd(in->dusrf_dy, n2, true);
    a.beta = putd(in->beta_device ? nullptr : in->beta_internal, n2, !in->beta_device);
    if (in->beta_device) a.beta = S->cb_beta;
    a.theta = putd(in->theta_basal, n2, true);
    a.flwafact = putd(in->flwafact, nfl, true);
    a.thck_cell = in->thck ? putd(in->thck, nc, true) : nullptr;
    a.flwa = in->flwa ? putd(in->flwa, nfl, true) : nullptr;
    a.slsrf = in->stag_lsrf ? putd(in->stag_lsrf, n2, true) : nullptr;
    a.beta_std = in->assemble_beta_standard ? 1 : 0;
    a.taud_std = in->assemble_taud_standard ? 1 : 0;
    if (lateral) {
        lat.usrf = putd(in->usrf, nc, true);
        lat.thck = a.thck_cell;
        lat.susrf = putd(in->stagusrf_lateral, n2, true);
        lat.sthck = putd(in->stagthck_lateral, n2, true);
        lat.local = in->which_assemble_lateral == 1 ? 1 : 0;
    }
    int32_t* q = ib;
    auto puti = [&](const int32_t* h, long long n) -> const int32_t* {
        int32_t* dst = q;
        q += n;
        if (h) keep(hipMemcpyAsync(dst, h, n * sizeof(int32_t), hipMemcpyHostToDevice, S->s));
        else keep(hipMemsetAsync(dst, 0, n * sizeof(int32_t), S->s));
        return dst;
    };
    a.acell = puti(in->active_cell, nc);
    a.avert = puti(in->active_vertex, n2);
    a.umask = puti(in->umask_dirichlet, nvec);
    a.vmask = puti(in->vmask_dirichlet, nvec);
    if (lateral) lat.ocean = puti(in->ocean_mask, nc);
    a.efvs = S->efvs;
    a.dx = in->dx; a.dy = in->dy; a.efvs_constant = in->efvs_constant; a.effstrain_min = in->effstrain_min;
    a.n_glen = in->n_glen; a.whichefvs = in->whichefvs; a.nhalo = 2;
    CG_CHECK(herr);
    // active mask (int8) for the solver
    {
        std::vector<int8_t> act(n2);
        for (long long t = 0; t < n2; ++t) act[t] = in->active_vertex[t] ? 1 : 0;
        CG_CHECK(hipMemcpyAsync(S->active, act.data(), n2, hipMemcpyHostToDevice, S->s));
        CG_CHECK(hipStreamSynchronize(S->s));
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_assemble_3d(cismgpu_system* S, const cismgpu_assembly_in* in,
                                              cismgpu_assembly_info* info) {
    if (!S || !in || !in->sigma || !in->stagwbndsigma || !in->active_cell || !in->active_vertex || !in->xVertex ||
        !in->yVertex || !in->stagusrf || !in->stagthck || !in->dusrf_dx || !in->dusrf_dy || !in->flwafact) {
        cismgpu_set_error("assemble_3d: missing input");
        return CISMGPU_ERR_ARG;
    }
    if (S->two_d) { cismgpu_set_error("assemble_3d needs a 3-D system"); return CISMGPU_ERR_ARG; }
    if (in->assemble_beta_standard && has_beta(in) && !in->stag_lsrf) {
        cismgpu_set_error("assemble_3d: the standard basal option needs stag_lsrf");
        return CISMGPU_ERR_ARG;
    }
    int st = upload_fe_tables();
    if (st) return st;
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny;
    const long long nfl = (long long)(d.nz - 1) * nc;
    AsmDev a{};
    LatDev lat{};
    bool lateral = false;
    st = upload_assembly_inputs(S, d, in, a, lat, lateral);
    if (st) return st;
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2);
    { const int st_ = reset_scal(S); if (st_) return st_; }
    const double* uv = S->V(0);
    const double* vv = S->V(1);
    // element-once assembly (CISMGPU_ASSEMBLY=gather: the per-row gather
    // that evaluates every element once per row)
    const char* aenv = getenv("CISMGPU_ASSEMBLY");
    const bool gather = aenv && !strcmp(aenv, "gather");
    if (!gather && !S->elemq) CG_CHECK(hipMalloc(&S->elemq, (size_t)8 * EQ * nfl * sizeof(double)));
    // counters: nnz, pairs left asymmetric by the fix, differing pairs
    unsigned long long* cnt = nullptr;
    CG_CHECK(hipMalloc(&cnt, 3 * sizeof(unsigned long long)));
    CG_CHECK(hipMemsetAsync(cnt, 0, 3 * sizeof(unsigned long long), S->s));
    // The element-once rows are symmetric by construction between owned
    // nodes; with no pass after them that changes couplings inside the
    // owned region (Dirichlet rows and columns, standard basal quadrature)
    // the symmetry check and the nonzero count visit only the rim rows, and
    // the other rows are counted as they are written (CISMGPU_SYMCHECK=full:
    // every row checked and counted).
    const char* senv = getenv("CISMGPU_SYMCHECK");
    const bool rim_only = !gather && !(senv && !strcmp(senv, "full")) && !in->umask_dirichlet &&
                          !in->vmask_dirichlet && !(has_beta(in) && a.beta_std);
    hipEventRecord(e0, S->s);
    if (gather) {
        k_assemble<<<nblocks(d.nrows, 64), 64, 0, S->s>>>(d, a, uv, vv, S->A64, S->scal + S_NAN);
        hipEventRecord(e1, S->s);
        k_efvs<<<nblocks(nfl, 128), 128, 0, S->s>>>(d, a, uv, vv, S->efvs);
    } else {
        k_elem<<<nblocks(nfl, 128), 128, 0, S->s>>>(d, a, uv, vv, S->elemq, S->efvs, S->scal + S_NAN);
        k_rows<<<(unsigned)rows_blocks(d), WAVE * ROWS_J, 0, S->s>>>(d, a, S->elemq, S->A64, rim_only ? cnt : nullptr);
        hipEventRecord(e1, S->s);
    }
    if (has_beta(in)) {
        if (a.beta_std)
            k_basal_std<<<nblocks((long long)d.nxo * d.nyo, 128), 128, 0, S->s>>>(d, a, d.nz - 1, 1, a.beta, a.beta, S->A64,
                                                                               S->scal + S_NAN);
        else
            k_basal_local<<<nblocks((long long)d.nxo * d.nyo, 256), 256, 0, S->s>>>(d, a, S->A64);
    }
    if (a.taud_std)
        k_load_gravity_std<<<nblocks(n2, 64), 64, 0, S->s>>>(d, a, S->V(2), S->V(3), S->scal + S_NAN);
    else
        k_load_gravity<<<nblocks(n2, 256), 256, 0, S->s>>>(d, a, S->V(2), S->V(3));
    if (lateral) k_lateral_bc<<<nblocks(n2, 128), 128, 0, S->s>>>(d, a, lat, S->V(2), S->V(3));
    // Dirichlet rows and columns (no prescribed nodes: nothing to do)
    if (in->umask_dirichlet || in->vmask_dirichlet)
        k_dirichlet<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a, uv, vv, S->A64, S->V(2), S->V(3));
    if (d.bc == 3) k_matrix_noice_edge<<<nblocks(rim_rows_per_level(d) * d.nz, 256), 256, 0, S->s>>>(d, S->A64);
    S->asm_in = a;
    S->asm3_ready = true;
    S->edge_zero = d.bc == 3;
    halo<double>(S, S->V(2), S->V(3), d.nz);
    // symmetry fix: count the differing pairs first; the copy-and-average
    // pass runs only when one exists
    S->sym_exact = false;
    {
        unsigned long long ndiff = 1;
        if (rim_only)
            k_symmetry_rim<<<nblocks(rim_rows_per_level(d) * d.nz, 256), 256, 0, S->s>>>(d, a.avert, S->A64, cnt + 2,
                                                                                      cnt);
        else
            k_symmetry_count<<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, a.avert, S->A64, cnt + 2, cnt);
        CG_CHECK(hipMemcpyAsync(&ndiff, cnt + 2, sizeof(ndiff), hipMemcpyDeviceToHost, S->s));
        CG_CHECK(hipStreamSynchronize(S->s));
        if (ndiff == 0) {
            S->sym_exact = true;
        } else if (in->check_symmetry) {
            if (!S->Atmp) CG_CHECK(hipMalloc(&S->Atmp, d.nrows * NCOEF * sizeof(double)));
            k_symmetry<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a.avert, S->A64, S->Atmp, cnt + 1);
            std::swap(S->A64, S->Atmp);
            S->edge_zero = false;   // averaging may fill entries of the zeroed rows
            // the averaged matrix is counted again
            CG_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), S->s));
            k_count_nnz<<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, a.avert, S->A64, cnt);
        }
    }
    // preconditioner coefficients: owned rows, then the halo rule on the ring
    CG_CHECK(hipMemsetAsync(S->M64, 0, 8 * d.vlen * sizeof(double), S->s));
    k_precond_from_rows<<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, S->A64, S->M64, S->M64 + 6 * d.vlen);
    for (int f = 0; f < 8; f += 2) halo<double>(S, S->M64 + f * d.vlen, S->M64 + (f + 1) * d.vlen, d.nz);
    // the raw Auu/Avv of the line-solve preconditioners are built when first needed
    S->raw_ready = false;
    hipEventRecord(e2, S->s);
    unsigned long long hc[2] = {0, 0};
    CG_CHECK(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(sync_scal(S) ? hipErrorUnknown : hipSuccess);
    hipFree(cnt);
    float ms1 = 0.f, ms2 = 0.f;
    hipEventElapsedTime(&ms1, e0, e1);
    hipEventElapsedTime(&ms2, e0, e2);
    hipEventDestroy(e0); hipEventDestroy(e1); hipEventDestroy(e2);
    if (info) {
        info->nnz = (int64_t)hc[0];
        info->asym_pairs = (int64_t)hc[1];
        info->assemble_ms = ms1;
        info->total_ms = ms2;
    }
    S->have_matrix = true;
    S->m32_ready = S->a32_ready = S->a16_ready = S->a64h_ready = false;
    S->a32_sym = S->a64_sym = false;
    S->tri.mode = 0;
    if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("Jacobian matrix is singular"); return CISMGPU_ERR_SINGULAR; }
    return CISMGPU_OK;
}

// 2-D (depth-integrated) assembly on a system from cismgpu_system_create_2d:
// the inputs are those of the 3-D assembly with the grid's nz sigma levels;
// the velocity is the resident 2-D x.
extern "C" int32_t cismgpu_system_assemble_2d(cismgpu_system* S, const cismgpu_assembly_in* in, int32_t whichapprox,
                                              cismgpu_assembly_info* info) {
    if (!S || !in || !in->sigma || !in->stagwbndsigma || !in->active_cell || !in->active_vertex || !in->xVertex ||
        !in->yVertex || !in->stagusrf || !in->stagthck || !in->dusrf_dx || !in->dusrf_dy || !in->flwafact) {
        cismgpu_set_error("assemble_2d: missing input");
        return CISMGPU_ERR_ARG;
    }
    if (!S->two_d) { cismgpu_set_error("assemble_2d needs a system made by cismgpu_system_create_2d"); return CISMGPU_ERR_ARG; }
    if (S->nz_levels < 2 || S->nz_levels > 64) {
        cismgpu_set_error("assemble_2d: the grid's nz (sigma levels of the inputs) must be 2..64, got %d", S->nz_levels);
        return CISMGPU_ERR_ARG;
    }
    const bool diva = whichapprox == CISMGPU_APPROX_DIVA, l1l2 = whichapprox == CISMGPU_APPROX_L1L2;
    if (whichapprox != CISMGPU_APPROX_SSA && !diva && !l1l2) {
        cismgpu_set_error("assemble_2d: which_ho_approx %d is not supported (SSA 1, L1L2 3, DIVA 4)", whichapprox);
        return CISMGPU_ERR_UNSUPPORTED;
    }
    if (l1l2 && (!in->flwa || (in->whichefvs == 2 && !(fabs(in->n_glen - 1.0) < 1.0e-10) &&
                               !(fabs(in->n_glen - 3.0) < 1.0e-10)))) {
        cismgpu_set_error("assemble_2d: L1L2 needs the rate factor (in->flwa) and n = 1 or 3 for the nonlinear viscosity");
        return CISMGPU_ERR_ARG;
    }
    if (diva && in->assemble_beta_standard && !in->stag_lsrf) {
        cismgpu_set_error("assemble_2d: the standard basal option of DIVA needs stag_lsrf");
        return CISMGPU_ERR_ARG;
    }
    if (diva && (!S->diva_ready || !in->thck)) {
        cismgpu_set_error("assemble_2d: DIVA needs cismgpu_system_diva_setup and the cell thickness (in->thck)");
        return CISMGPU_ERR_ARG;
    }
    int st = upload_fe_tables();
    if (st) return st;
    const Dims& d = S->d;
    Dims d3 = d;                       // the nz-level fields of the inputs
    d3.nz = S->nz_levels;
    d3.vlen = (long long)(d3.nz + 2) * d3.plane;
    d3.nrows = (long long)d3.nz * d3.nyo * d3.nxp;
    const Lev L{d3.nz};
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny;
    AsmDev a{};
    LatDev lat{};
    bool lateral = false;
    st = upload_assembly_inputs(S, d3, in, a, lat, lateral);
    if (st) return st;
    if (!S->load3) CG_CHECK(hipMalloc(&S->load3, 2 * d3.vlen * sizeof(double)));
    if (l1l2 && !S->efva2) CG_CHECK(hipMalloc(&S->efva2, 4 * nc * sizeof(double)));
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2);
    { const int st_ = reset_scal(S); if (st_) return st_; }
    const double* uv = S->V(0);
    const double* vv = S->V(1);
    hipEventRecord(e0, S->s);
    S->asm_in = a;
    S->last_approx = whichapprox;
    if (l1l2) {   // L1L2 layer viscosities and their vertical means
        k_l1l2_cell<<<nblocks(nc, 128), 128, 0, S->s>>>(d, a, L, uv, vv, S->efva2, S->efvs, S->scal + S_NAN);
        k_assemble_2d<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a, L, uv, vv, S->efva2, S->A64, S->scal + S_NAN);
    } else if (diva) {   // layer viscosities (state), omega integrals, efvs
        k_diva_cell<<<nblocks(nc, 128), 128, 0, S->s>>>(d, a, L, S->dv, uv, vv, S->efvs, S->scal + S_NAN);
        k_assemble_2d<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a, L, uv, vv, S->dv.efva, S->A64, S->scal + S_NAN);
    } else {
        k_assemble_2d<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a, L, uv, vv, nullptr, S->A64, S->scal + S_NAN);
    }
    hipEventRecord(e1, S->s);
    if (diva) {   // omega halo, beta_eff, basal_sliding_bc_2d_diva
        if (d.bc == 0 || d.bc == 2) {
            k_cell_halo<<<nblocks(d.ny, 256), 256, 0, S->s>>>(d, 0, S->dv.omega);
            k_cell_halo<<<nblocks(d.nx, 256), 256, 0, S->s>>>(d, 1, S->dv.omega);
        } else {
            k_cell_halo<<<nblocks(nc, 256), 256, 0, S->s>>>(d, 0, S->dv.omega);
        }
        k_diva_beta<<<nblocks(n2, 256), 256, 0, S->s>>>(d, a, S->dv);
        if (a.beta_std)   // standard option: curved Jacobian, beta_eff_x / beta_eff_y
            k_basal_std<<<nblocks((long long)d.nxo * d.nyo, 128), 128, 0, S->s>>>(d, a, 0, 1, S->dv.bex, S->dv.bey, S->A64,
                                                                               S->scal + S_NAN);
        else
            k_basal_diva_2d<<<nblocks((long long)d.nxo * d.nyo, 256), 256, 0, S->s>>>(d, a, S->dv, S->A64);
    } else {
        if (!l1l2) k_efvs_2d<<<nblocks(nc, 128), 128, 0, S->s>>>(d, a, L, uv, vv, S->efvs);
        if (has_beta(in)) {
            if (a.beta_std)
                k_basal_std<<<nblocks((long long)d.nxo * d.nyo, 128), 128, 0, S->s>>>(d, a, 0, 0, a.beta, a.beta, S->A64,
                                                                                   S->scal + S_NAN);
            else
                k_basal_local_2d<<<nblocks((long long)d.nxo * d.nyo, 256), 256, 0, S->s>>>(d, a, S->A64);
        }
    }
    // b = level sum of the 3-D load (gravity, then the lateral BC)
    double *lu = S->load3, *lv = S->load3 + d3.vlen;
    CG_CHECK(hipMemsetAsync(S->load3, 0, 2 * d3.vlen * sizeof(double), S->s));
    if (a.taud_std)
        k_load_gravity_std<<<nblocks(n2, 64), 64, 0, S->s>>>(d3, a, lu, lv, S->scal + S_NAN);
    else
        k_load_gravity<<<nblocks(n2, 256), 256, 0, S->s>>>(d3, a, lu, lv);
    if (lateral) k_lateral_bc<<<nblocks(n2, 128), 128, 0, S->s>>>(d3, a, lat, lu, lv);
    k_sum_levels<<<nblocks(n2, 256), 256, 0, S->s>>>(d3, lu, lv, d, S->V(2), S->V(3));
    k_dirichlet_2d<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a, L, uv, vv, S->A64, S->V(2), S->V(3));
    if (d.bc == 3) k_matrix_noice_edge<<<nblocks(rim_rows_per_level(d) * d.nz, 256), 256, 0, S->s>>>(d, S->A64);
    S->edge_zero = d.bc == 3;
    halo<double>(S, S->V(2), S->V(3), 1);
    unsigned long long* cnt = nullptr;
    CG_CHECK(hipMalloc(&cnt, 2 * sizeof(unsigned long long)));
    CG_CHECK(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), S->s));
    if (in->check_symmetry) k_symmetry_report_2d<<<nblocks(d.nrows, 128), 128, 0, S->s>>>(d, a.avert, S->A64, cnt + 1);
    k_count_nnz<<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, a.avert, S->A64, cnt);
    CG_CHECK(hipMemsetAsync(S->M64, 0, 8 * d.vlen * sizeof(double), S->s));
    k_precond_from_rows<<<nblocks(d.nrows, 256), 256, 0, S->s>>>(d, S->A64, S->M64, S->M64 + 6 * d.vlen);
    for (int f = 0; f < 8; f += 2) halo<double>(S, S->M64 + f * d.vlen, S->M64 + (f + 1) * d.vlen, 1);
    const long long nraw = 27LL * n2;
    if (!S->stage) CG_CHECK(hipMalloc(&S->stage, 2 * nraw * sizeof(double)));
    k_rows_to_raw<<<nblocks(nraw, 256), 256, 0, S->s>>>(d, 0, S->A64, S->stage);
    k_rows_to_raw<<<nblocks(nraw, 256), 256, 0, S->s>>>(d, 3, S->A64, S->stage + nraw);
    S->raw_ready = true;
    S->sym_exact = false;
    hipEventRecord(e2, S->s);
    unsigned long long hc[2] = {0, 0};
    CG_CHECK(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(sync_scal(S) ? hipErrorUnknown : hipSuccess);
    hipFree(cnt);
    float ms1 = 0.f, ms2 = 0.f;
    hipEventElapsedTime(&ms1, e0, e1);
    hipEventElapsedTime(&ms2, e0, e2);
    hipEventDestroy(e0); hipEventDestroy(e1); hipEventDestroy(e2);
    if (info) {
        info->nnz = (int64_t)hc[0];
        info->asym_pairs = (int64_t)hc[1];
        info->assemble_ms = ms1;
        info->total_ms = ms2;
    }
    S->have_matrix = true;
    S->m32_ready = S->a32_ready = S->a16_ready = S->a64h_ready = false;
    S->a32_sym = S->a64_sym = false;
    S->tri.mode = 0;
    if (S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("Jacobian matrix is singular"); return CISMGPU_ERR_SINGULAR; }
    return CISMGPU_OK;
}

// compute_3d_velocity_L1L2 (glissade_velo_higher.F90:5904-6315) from the
// resident 2-D x and the inputs and layer viscosities of the last L1L2
// assembly.
extern "C" int32_t cismgpu_system_l1l2_velocity(cismgpu_system* S, const int32_t* ice_mask, double* uvel,
                                                double* vvel) {
    if (!S || !ice_mask || !uvel || !vvel) { cismgpu_set_error("l1l2_velocity: null argument"); return CISMGPU_ERR_ARG; }
    if (!S->two_d || S->last_approx != CISMGPU_APPROX_L1L2 || !S->asm_in.thck_cell || !S->asm_in.flwa) {
        cismgpu_set_error("l1l2_velocity: the last 2-D assembly was not an L1L2 one with the cell thickness");
        return CISMGPU_ERR_ARG;
    }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny, nz = S->nz_levels, nl = nz - 1;
    if (4 * nl > 65535) { cismgpu_set_error("l1l2_velocity: too many levels"); return CISMGPU_ERR_ARG; }
    double* f = nullptr;
    CG_CHECK(hipMalloc(&f, (4 * nl * nc + 2 * nz * n2) * sizeof(double) + nc * sizeof(int32_t)));
    double* u3 = f + 4 * nl * nc;
    double* v3 = u3 + nz * n2;
    int32_t* ice = (int32_t*)(v3 + nz * n2);
    int st = CISMGPU_OK;
    const Lev L{(int)nz};
    const double* uv = S->V(0);
    const double* vv = S->V(1);
    if (hipMemcpyAsync(ice, ice_mask, nc * sizeof(int32_t), hipMemcpyHostToDevice, S->s) != hipSuccess) st = CISMGPU_ERR_HIP;
    if (!st) st = reset_scal(S);
    if (!st) {
        k_l1l2_columns<<<nblocks(nc, 128), 128, 0, S->s>>>(d, S->asm_in, L, uv, vv, S->efvs, f, S->scal + S_NAN);
        if (d.bc == 0 || d.bc == 2) {
            k_cell_halo_layers<<<dim3((unsigned)nblocks(d.ny, 256), (unsigned)(4 * nl)), 256, 0, S->s>>>(d, 0, f);
            k_cell_halo_layers<<<dim3((unsigned)nblocks(d.nx, 256), (unsigned)(4 * nl)), 256, 0, S->s>>>(d, 1, f);
        } else {
            k_cell_halo_layers<<<dim3((unsigned)nblocks(nc, 256), (unsigned)(4 * nl)), 256, 0, S->s>>>(d, 0, f);
        }
        k_l1l2_velocity<<<nblocks(n2, 128), 128, 0, S->s>>>(d, S->asm_in, L, uv, vv, ice, S->efvs, f, u3, v3);
        if (hipMemcpyAsync(uvel, u3, nz * n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
            hipMemcpyAsync(vvel, v3, nz * n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess)
            st = CISMGPU_ERR_HIP;
    }
    if (!st && sync_scal(S)) st = CISMGPU_ERR_HIP;
    hipFree(f);
    if (st == CISMGPU_ERR_HIP) { cismgpu_set_error("l1l2_velocity: device failure"); return st; }
    if (!st && S->scal_h[S_NAN] != 0.0) { cismgpu_set_error("Jacobian matrix is singular"); return CISMGPU_ERR_SINGULAR; }
    return st;
}

// DIVA inputs and state of a 2-D system (glissade_velo_higher.F90:2078-2140).
extern "C" int32_t cismgpu_system_diva_setup(cismgpu_system* S, const cismgpu_diva_in* in) {
    if (!S || !in || !S->two_d || !in->slope_factor_x || !in->slope_factor_y || !in->stag_slope_factor_x ||
        !in->stag_slope_factor_y || !in->stag_theta_slope_x || !in->stag_theta_slope_y || !in->theta_basal_x ||
        !in->theta_basal_y || !in->ice_plus_land_mask) {
        cismgpu_set_error("diva_setup: needs a 2-D system and every slope/mask field");
        return CISMGPU_ERR_ARG;
    }
    if (S->nz_levels < 2) { cismgpu_set_error("diva_setup: the grid's nz must be >= 2"); return CISMGPU_ERR_ARG; }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny, nz = S->nz_levels;
    // doubles: sfx, sfy (nc); ssfx, ssfy, sthx, sthy, tbx, tby, btx, bty, somega, bex, bey (n2);
    // efvs_qp (nz-1)*4*nc; omega_k nz*nc; omega nc; efva 4*nc; then the int mask (nc)
    const long long ndbl = 2 * nc + 11 * n2 + (nz - 1) * 4 * nc + nz * nc + nc + 4 * nc;
    if (!S->diva_buf) CG_CHECK(hipMalloc(&S->diva_buf, ndbl * sizeof(double) + nc * sizeof(int32_t)));
    double* p = S->diva_buf;
    DivaDev& dv = S->dv;
    auto take = [&](long long n) { double* q = p; p += n; return q; };
    hipError_t herr = hipSuccess;
    auto up = [&](const double* h, long long n, bool zero) -> double* {
        double* q = take(n);
        hipError_t e = hipSuccess;
        if (h) e = hipMemcpyAsync(q, h, n * sizeof(double), hipMemcpyHostToDevice, S->s);
        else if (zero) e = hipMemsetAsync(q, 0, n * sizeof(double), S->s);
        if (e != hipSuccess && herr == hipSuccess) herr = e;
        return q;
    };
    dv.sfx = up(in->slope_factor_x, nc, true);
    dv.sfy = up(in->slope_factor_y, nc, true);
    dv.ssfx = up(in->stag_slope_factor_x, n2, true);
    dv.ssfy = up(in->stag_slope_factor_y, n2, true);
    dv.sthx = up(in->stag_theta_slope_x, n2, true);
    dv.sthy = up(in->stag_theta_slope_y, n2, true);
    dv.tbx = up(in->theta_basal_x, n2, true);
    dv.tby = up(in->theta_basal_y, n2, true);
    dv.btx = up(in->btractx, n2, true);
    dv.bty = up(in->btracty, n2, true);
    dv.somega = take(n2);
    dv.bex = up(nullptr, n2, true);
    dv.bey = up(nullptr, n2, true);
    dv.efvs_qp = up(in->efvs_qp_3d, (nz - 1) * 4 * nc, true);
    dv.omega_k = up(nullptr, nz * nc, true);
    dv.omega = up(nullptr, nc, true);
    dv.efva = take(4 * nc);
    int32_t* m = (int32_t*)p;
    CG_CHECK(herr);
    CG_CHECK(hipMemcpyAsync(m, in->ice_plus_land_mask, nc * sizeof(int32_t), hipMemcpyHostToDevice, S->s));
    dv.ipl = m;
    dv.no_slip = in->no_slip ? 1 : 0;
    CG_CHECK(hipStreamSynchronize(S->s));
    S->diva_ready = true;
    S->cb_dvel_valid = false;
    return CISMGPU_OK;
}

// compute_3d_velocity_diva (glissade_velo_higher.F90:5779-5900) from the
// resident 2-D x and the beta_eff of the last DIVA assembly: updates the
// basal traction state; optionally returns the 3-D velocity (nz, nx-1, ny-1)
// and the traction (nx-1, ny-1) to the host.  a.beta is the last assembly's.
extern "C" int32_t cismgpu_system_diva_velocity(cismgpu_system* S, double* uvel, double* vvel, double* btractx,
                                                double* btracty) {
    if (!S || !S->diva_ready || !S->asm_buf) { cismgpu_set_error("diva_velocity: no DIVA assembly"); return CISMGPU_ERR_ARG; }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nz = S->nz_levels;
    double* u3 = nullptr;
    if (uvel && vvel) {
        CG_CHECK(hipMalloc(&u3, 2 * nz * n2 * sizeof(double)));
        CG_CHECK(hipMemsetAsync(u3, 0, 2 * nz * n2 * sizeof(double), S->s));   // inactive vertices stay zero
    }
    // the basal level goes to the calcbeta basal pair when the system has a friction law
    double* ub = S->cb_setup ? S->cb_dvel : nullptr;
    k_diva_velocity<<<nblocks(n2, 256), 256, 0, S->s>>>(d, S->asm_in, Lev{(int)nz}, S->dv, S->V(0), S->V(1), u3,
                                                        u3 ? u3 + nz * n2 : nullptr, ub, ub ? ub + n2 : nullptr);
    if (ub) S->cb_dvel_valid = true;
    int st = CISMGPU_OK;
    if (u3) {
        if (hipMemcpyAsync(uvel, u3, nz * n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
            hipMemcpyAsync(vvel, u3 + nz * n2, nz * n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess)
            st = CISMGPU_ERR_HIP;
    }
    if (!st && btractx && btracty &&
        (hipMemcpyAsync(btractx, S->dv.btx, n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
         hipMemcpyAsync(btracty, S->dv.bty, n2 * sizeof(double), hipMemcpyDeviceToHost, S->s) != hipSuccess))
        st = CISMGPU_ERR_HIP;
    if (hipStreamSynchronize(S->s) != hipSuccess) st = CISMGPU_ERR_HIP;
    if (u3) hipFree(u3);
    if (st) cismgpu_set_error("diva_velocity: device copy failed");
    return st;
}

// Saved-assembly snapshot for the accelerated Picard iteration
// (evaluate_accelerated_picard_3d, glissade_velo_higher.F90:10052-10184):
// the owned rows of A, b and the preconditioner coefficients derived from A.
// Saved assembly: A, b, the preconditioner, the raw rows and, for DIVA,
// beta_eff and omega_k (evaluate_accelerated_picard_2d saves and restores
// them with the matrix, glissade_velo_higher.F90:9898-10048).
constexpr int NSAVED = 7;

static void saved_sizes(const cismgpu_system* S, size_t sizes[NSAVED]) {
    const Dims& d = S->d;
    const size_t nraw = 27ull * d.nz * d.nvx * d.nvy;
    const size_t n2 = (size_t)d.nvx * d.nvy, nc = (size_t)d.nx * d.ny;
    sizes[0] = (size_t)d.nrows * NCOEF;
    sizes[1] = 2 * (size_t)d.vlen;
    sizes[2] = 8 * (size_t)d.vlen;
    sizes[3] = S->stage ? 2 * nraw : 0;
    sizes[4] = S->diva_ready ? n2 : 0;
    sizes[5] = S->diva_ready ? n2 : 0;
    sizes[6] = S->diva_ready ? (size_t)S->nz_levels * nc : 0;
}

static void saved_parts(const cismgpu_system* S, double* base, double* parts[NSAVED], size_t sizes[NSAVED]) {
    saved_sizes(S, sizes);
    parts[0] = base;
    for (int q = 1; q < NSAVED; ++q) parts[q] = parts[q - 1] + sizes[q - 1];
}

static size_t saved_doubles(const cismgpu_system* S) {
    size_t sizes[NSAVED], n = 0;
    saved_sizes(S, sizes);
    for (int q = 0; q < NSAVED; ++q) n += sizes[q];
    return n;
}

extern "C" int32_t cismgpu_system_save_assembly(cismgpu_system* S) {
    if (!S || !S->have_matrix) return CISMGPU_ERR_ARG;
    const size_t need = saved_doubles(S);
    if (S->saved && S->saved_n < need) {
        CG_CHECK(hipFree(S->saved));
        S->saved = nullptr;
    }
    if (!S->saved) {
        CG_CHECK(hipMalloc(&S->saved, need * sizeof(double)));
        S->saved_n = need;
    }
    double* parts[NSAVED];
    size_t sizes[NSAVED];
    saved_parts(S, S->saved, parts, sizes);
    double* src[NSAVED] = {S->A64, S->V(2), S->M64, S->stage, S->dv.bex, S->dv.bey, S->dv.omega_k};
    for (int q = 0; q < NSAVED; ++q)
        if (sizes[q]) CG_CHECK(hipMemcpyAsync(parts[q], src[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    S->have_saved = true;
    S->saved_raw_ready = S->raw_ready;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_restore_assembly(cismgpu_system* S) {
    if (!S || !S->have_saved || S->saved_n < saved_doubles(S)) {
        cismgpu_set_error("restore_assembly: nothing saved");
        return CISMGPU_ERR_ARG;
    }
    double* parts[NSAVED];
    size_t sizes[NSAVED];
    saved_parts(S, S->saved, parts, sizes);
    double* dst[NSAVED] = {S->A64, S->V(2), S->M64, S->stage, S->dv.bex, S->dv.bey, S->dv.omega_k};
    for (int q = 0; q < NSAVED; ++q)
        if (sizes[q]) CG_CHECK(hipMemcpyAsync(dst[q], parts[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    S->have_matrix = true;
    S->raw_ready = S->saved_raw_ready;
    S->sym_exact = false;
    S->edge_zero = false;
    S->m32_ready = S->a32_ready = S->a16_ready = S->a64h_ready = false;
    S->a32_sym = S->a64_sym = false;
    S->tri.mode = 0;
    return CISMGPU_OK;
}

// glissade_calcbeta: the fields each sliding law reads
// (glissade_basal_traction.F90:200-644)
static const char* calcbeta_missing(const cismgpu_calcbeta_in* in) {
    const int b = in->whichbabc;
    if (!in->f_ground) return "f_ground";
    if (b == 1 && !in->bpmp_mask) return "bpmp_mask";
    if ((b == 2 || b == 3 || b == 7 || (b >= 10 && b <= 14)) && !in->effecpress_stag) return "effecpress_stag";
    if ((b == 2 || b == 7 || (b >= 10 && b <= 13)) && !in->coulomb_c) return "coulomb_c";
    if ((b == 9 || (b >= 11 && b <= 13)) && !in->powerlaw_c) return "powerlaw_c";
    if (b == 3 && !in->topg) return "topg";
    if (b == 5 && !in->beta_external) return "beta_external";
    if (b == 10 && (!in->thck || !in->flwa_basal)) return "thck and flwa_basal";
    if (b == 15 && !in->mintauf) return "mintauf";
    return nullptr;
}

extern "C" int32_t cismgpu_system_calcbeta_setup(cismgpu_system* S, const cismgpu_calcbeta_in* in) {
    if (!S || !in) { cismgpu_set_error("calcbeta_setup: null argument"); return CISMGPU_ERR_ARG; }
    const int b = in->whichbabc;
    if (b < 0 || b > 15 || b == 6) {
        cismgpu_set_error("calcbeta_setup: which_ho_babc %d is not a sliding law (6 no-slip is a Dirichlet "
                          "condition)", b);
        return CISMGPU_ERR_ARG;
    }
    if (b == 8 && (!(in->dew > 0.0) || !(in->dns > 0.0))) {
        cismgpu_set_error("calcbeta_setup: ISMIP-HOM C (8) needs the grid spacing dew, dns");
        return CISMGPU_ERR_ARG;
    }
    if (in->which_beta_limit != 0 && in->which_beta_limit != 1) {
        cismgpu_set_error("calcbeta_setup: bad which_ho_beta_limit %d", in->which_beta_limit);
        return CISMGPU_ERR_ARG;
    }
    if (const char* miss = calcbeta_missing(in)) {
        cismgpu_set_error("calcbeta_setup: which_ho_babc %d needs %s", b, miss);
        return CISMGPU_ERR_ARG;
    }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny;
    if (b == 5) {   // beta could not be read (:405-411)
        double mx = -INFINITY;
        for (long long t = 0; t < n2; ++t) mx = std::max(mx, in->beta_external[t]);
        if (mx < 0.0) {
            cismgpu_set_error("calcbeta_setup: every external beta value is < 0 (beta not read from the input)");
            return CISMGPU_ERR_BETA;
        }
    }
    if (!S->cb_buf) {
        // doubles: beta + scratch, host and DIVA velocity pairs, 7 vertex fields, 2 cell fields;
        // then ints, the bad slot
        const size_t bytes = (size_t)(13 * n2 + 2 * nc) * sizeof(double) + (size_t)(n2 + nc) * sizeof(int32_t) + 64;
        CG_CHECK(hipMalloc(&S->cb_buf, bytes));
    }
    double* p = (double*)S->cb_buf;
    S->cb_beta = p; p += 2 * n2;
    S->cb_vel = p; p += 2 * n2;
    S->cb_dvel = p; p += 2 * n2;
    S->cb_dvel_valid = false;
    CBDev c{};
    c.babc = b;
    c.limit = in->which_beta_limit;
    const double prm[CB_NPARAM] = {in->ho_beta_const, in->ho_beta_small, in->ho_beta_large, in->pseudo_plastic_q,
                                   in->pseudo_plastic_u0, in->pseudo_plastic_phimin, in->pseudo_plastic_phimax,
                                   in->pseudo_plastic_bedmin, in->pseudo_plastic_bedmax, in->eus, in->powerlaw_m,
                                   in->zoet_iverson_ut, in->zoet_iverson_nmax, in->coulomb_bump_max_slope,
                                   in->coulomb_bump_wavelength, in->friction_powerlaw_k, in->beta_grounded_min,
                                   in->beta_powerlaw_umax, in->dew, in->dns};
    for (int q = 0; q < CB_NPARAM; ++q) c.p[q] = prm[q];
    hipError_t herr = hipSuccess;
    auto put = [&](const double* h, long long n) -> const double* {
        double* dst = p;
        p += n;
        if (!h) return nullptr;
        const hipError_t e = hipMemcpyAsync(dst, h, n * sizeof(double), hipMemcpyHostToDevice, S->s);
        if (e != hipSuccess && herr == hipSuccess) herr = e;
        return dst;
    };
    c.fg = put(in->f_ground, n2);
    c.N = put(in->effecpress_stag, n2);
    c.cc = put(in->coulomb_c, n2);
    c.pc = put(in->powerlaw_c, n2);
    c.mintauf = put(in->mintauf, n2);
    c.csf = put(in->c_space_factor_stag, n2);
    c.bext = put(in->beta_external, n2);
    c.topg = put(in->topg, nc);
    c.flwa = put(in->flwa_basal, nc);
    int32_t* ip = (int32_t*)p;
    if (in->bpmp_mask) {
        const hipError_t e = hipMemcpyAsync(ip, in->bpmp_mask, n2 * sizeof(int32_t), hipMemcpyHostToDevice, S->s);
        if (e != hipSuccess && herr == hipSuccess) herr = e;
        c.bpmp = ip;
    }
    ip += n2;
    std::vector<int32_t> imask;
    if (in->thck) {   // thck > 0 (:479-483)
        imask.resize(nc);
        for (long long t = 0; t < nc; ++t) imask[t] = in->thck[t] > 0.0 ? 1 : 0;
        const hipError_t e = hipMemcpyAsync(ip, imask.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, S->s);
        if (e != hipSuccess && herr == hipSuccess) herr = e;
        c.imask = ip;
    }
    ip += nc;
    S->cb_bad = (unsigned long long*)(((uintptr_t)ip + 7) & ~(uintptr_t)7);
    CG_CHECK(herr);
    CG_CHECK(hipStreamSynchronize(S->s));   // the host mask buffer goes out of scope
    S->cb = c;
    S->cb_setup = true;
    S->cb_valid = false;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_calcbeta(cismgpu_system* S, const double* ubas, const double* vbas, double* beta) {
    if (!S || !S->cb_setup) { cismgpu_set_error("calcbeta: no calcbeta_setup on this system"); return CISMGPU_ERR_ARG; }
    if ((ubas == nullptr) != (vbas == nullptr)) {
        cismgpu_set_error("calcbeta: pass both basal velocity components or neither");
        return CISMGPU_ERR_ARG;
    }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy;
    const double *u = S->V(0), *v = S->V(1);
    int lev = d.nz - 1;   // basal level of x (the only level of a 2-D system)
    if (!ubas && S->two_d && S->diva_ready && S->cb_dvel_valid) {
        // DIVA: the basal level of the last compute_3d_velocity_diva (the
        // reference's uvel(nz,:,:), glissade_velo_higher.F90:2293-2296)
        u = S->cb_dvel;
        v = S->cb_dvel + n2;
        lev = -1;
    }
    if (ubas) {
        CG_CHECK(hipMemcpyAsync(S->cb_vel, ubas, n2 * sizeof(double), hipMemcpyHostToDevice, S->s));
        CG_CHECK(hipMemcpyAsync(S->cb_vel + n2, vbas, n2 * sizeof(double), hipMemcpyHostToDevice, S->s));
        u = S->cb_vel;
        v = S->cb_vel + n2;
        lev = -1;
    }
    S->cb_valid = false;
    CG_CHECK(hipMemsetAsync(S->cb_bad, 0xff, sizeof(unsigned long long), S->s));
    k_calcbeta<<<nblocks(n2, 256), 256, 0, S->s>>>(d, S->cb, u, v, lev, S->cb_beta, S->cb_bad);
    CG_CHECK(hipGetLastError());
    unsigned long long bad = 0;
    CG_CHECK(hipMemcpyAsync(&bad, S->cb_bad, sizeof(bad), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    if (bad != ~0ULL) {
        double bv = 0.0;
        CG_CHECK(hipMemcpy(&bv, S->cb_beta + bad, sizeof(double), hipMemcpyDeviceToHost));
        cismgpu_set_error("calcbeta: invalid beta %g at vertex i=%lld j=%lld", bv, bad % d.nvx + 1, bad / d.nvx + 1);
        return CISMGPU_ERR_BETA;
    }
    halo2d<double>(S, S->cb_beta, S->cb_beta + n2);   // staggered halo (:703)
    if (beta) CG_CHECK(hipMemcpyAsync(beta, S->cb_beta, n2 * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    S->cb_valid = true;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_internal_stress(cismgpu_system* S, double* tau_xz, double* tau_yz, double* tau_xx,
                                                  double* tau_yy, double* tau_xy, double* tau_eff) {
    if (!S || S->two_d || !S->asm3_ready) {
        cismgpu_set_error("internal_stress: needs a 3-D system with an assembly");
        return CISMGPU_ERR_ARG;
    }
    const Dims& d = S->d;
    const long long nfl = (long long)(d.nz - 1) * d.nx * d.ny;
    double* tau = nullptr;
    CG_CHECK(hipMalloc(&tau, 6 * nfl * sizeof(double)));
    k_internal_stress<<<nblocks(nfl, 128), 128, 0, S->s>>>(d, S->asm_in, 2, 0, S->V(0), S->V(1), nullptr, tau);
    double* out[6] = {tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff};
    hipError_t e = hipGetLastError();
    for (int q = 0; q < 6 && e == hipSuccess; ++q)
        if (out[q]) e = hipMemcpyAsync(out[q], tau + q * nfl, nfl * sizeof(double), hipMemcpyDeviceToHost, S->s);
    if (e == hipSuccess) e = hipStreamSynchronize(S->s);
    hipFree(tau);
    if (e != hipSuccess) {
        cismgpu_set_error("internal_stress: %s", hipGetErrorString(e));
        return CISMGPU_ERR_HIP;
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_internal_stress_2d(cismgpu_system* S, const double* uvel, const double* vvel,
                                                     double* tau_xz, double* tau_yz, double* tau_xx, double* tau_yy,
                                                     double* tau_xy, double* tau_eff) {
    if (!S || !S->two_d || !S->asm_in.acell || S->last_approx == 0) {
        cismgpu_set_error("internal_stress_2d: needs a 2-D system with an SSA, L1L2 or DIVA assembly");
        return CISMGPU_ERR_ARG;
    }
    if ((uvel == nullptr) != (vvel == nullptr)) {
        cismgpu_set_error("internal_stress_2d: pass both velocity components or neither");
        return CISMGPU_ERR_ARG;
    }
    const int approx = S->last_approx;
    if (!uvel && approx != CISMGPU_APPROX_SSA) {
        cismgpu_set_error("internal_stress_2d: L1L2 and DIVA need the 3-D velocity (of l1l2_velocity or diva_velocity)");
        return CISMGPU_ERR_ARG;
    }
    Dims d = S->d;   // the element loop runs over the nz levels of the inputs
    d.nz = S->nz_levels;
    const long long n2 = (long long)d.nvx * d.nvy, nfl = (long long)(d.nz - 1) * d.nx * d.ny;
    const long long nv = uvel ? 2LL * d.nz * n2 : 0;
    double* buf = nullptr;
    CG_CHECK(hipMalloc(&buf, (6 * nfl + nv) * sizeof(double)));
    double *tau = buf, *u3 = buf + 6 * nfl;
    hipError_t e = hipSuccess;
    const double *u = S->V(0), *v = S->V(1);
    int vmode = 1;   // SSA: the 2-D velocity at every level
    if (uvel) {
        e = hipMemcpyAsync(u3, uvel, d.nz * n2 * sizeof(double), hipMemcpyHostToDevice, S->s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(u3 + d.nz * n2, vvel, d.nz * n2 * sizeof(double), hipMemcpyHostToDevice, S->s);
        u = u3;
        v = u3 + d.nz * n2;
        vmode = 2;
    }
    if (e == hipSuccess) {
        // DIVA takes the 3-D branch of compute_internal_stress (every strain rate)
        const int branch = approx == CISMGPU_APPROX_DIVA ? 2 : approx;
        k_internal_stress<<<nblocks(nfl, 128), 128, 0, S->s>>>(d, S->asm_in, branch, vmode, u, v, S->efvs, tau);
        e = hipGetLastError();
    }
    double* out[6] = {tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff};
    for (int q = 0; q < 6 && e == hipSuccess; ++q)
        if (out[q]) e = hipMemcpyAsync(out[q], tau + q * nfl, nfl * sizeof(double), hipMemcpyDeviceToHost, S->s);
    if (e == hipSuccess) e = hipStreamSynchronize(S->s);
    hipFree(buf);
    if (e != hipSuccess) {
        cismgpu_set_error("internal_stress_2d: %s", hipGetErrorString(e));
        return CISMGPU_ERR_HIP;
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_basal_friction(cismgpu_system* S, int32_t which_bfric, const double* ubas,
                                                 const double* vbas, const double* beta, double* bfricflx,
                                                 double* btractx, double* btracty) {
    if (!S || !S->asm_in.acell) {
        cismgpu_set_error("basal_friction: needs a system with an assembly");
        return CISMGPU_ERR_ARG;
    }
    if (which_bfric != 0 && which_bfric != 1) {
        cismgpu_set_error("basal_friction: bad which_ho_assemble_bfric %d (0 standard, 1 local)", which_bfric);
        return CISMGPU_ERR_ARG;
    }
    if ((ubas == nullptr) != (vbas == nullptr)) {
        cismgpu_set_error("basal_friction: pass both basal velocity components or neither");
        return CISMGPU_ERR_ARG;
    }
    const Dims& d = S->d;
    const long long n2 = (long long)d.nvx * d.nvy, nc = (long long)d.nx * d.ny;
    const bool diva = S->two_d && S->last_approx == CISMGPU_APPROX_DIVA;
    const double *u = S->V(0), *v = S->V(1);
    int lev = d.nz - 1;   // basal level of x (the only level of a 2-D system: SSA's and L1L2's basal velocity)
    if (!ubas && diva) {
        // DIVA's basal velocity is the bed level of compute_3d_velocity_diva
        if (!S->cb_dvel_valid) {
            cismgpu_set_error("basal_friction: a DIVA system keeps its basal velocity only with a friction law "
                              "and after diva_velocity; pass ubas, vbas");
            return CISMGPU_ERR_ARG;
        }
        u = S->cb_dvel;
        v = S->cb_dvel + n2;
        lev = -1;
    }
    double* tmp = nullptr;
    CG_CHECK(hipMalloc(&tmp, (nc + 5 * n2) * sizeof(double)));
    double *bf = tmp, *bt = tmp + nc, *hv = tmp + nc + 2 * n2, *hb = tmp + nc + 4 * n2;
    hipError_t e = hipSuccess;
    auto up = [&](double* dst, const double* h) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, h, n2 * sizeof(double), hipMemcpyHostToDevice, S->s);
    };
    if (ubas) {
        up(hv, ubas);
        up(hv + n2, vbas);
        u = hv;
        v = hv + n2;
        lev = -1;
    }
    const double* b = S->asm_in.beta;   // the last assembly's beta (calcbeta's on the device)
    if (beta) {
        up(hb, beta);
        b = hb;
    }
    if (e == hipSuccess) {
        k_basal_friction<<<nblocks(nc, 256), 256, 0, S->s>>>(d, which_bfric, 2, S->asm_in.acell, S->asm_in.avert, u, v,
                                                             lev, b, bfricflx ? bf : nullptr,
                                                             (btractx || btracty) ? bt : nullptr);
        if (bfricflx) {   // parallel_halo of the cell field (:3807)
            if (d.bc == 0 || d.bc == 2) {
                k_cell_halo<<<nblocks(d.ny, 256), 256, 0, S->s>>>(d, 0, bf);
                k_cell_halo<<<nblocks(d.nx, 256), 256, 0, S->s>>>(d, 1, bf);
            } else {
                k_cell_halo<<<nblocks(nc, 256), 256, 0, S->s>>>(d, 0, bf);
            }
        }
        e = hipGetLastError();
    }
    auto down = [&](double* h, const double* src, long long n) {
        if (h && e == hipSuccess) e = hipMemcpyAsync(h, src, n * sizeof(double), hipMemcpyDeviceToHost, S->s);
    };
    down(bfricflx, bf, nc);
    down(btractx, bt, n2);
    down(btracty, bt + n2, n2);
    if (e == hipSuccess) e = hipStreamSynchronize(S->s);
    hipFree(tmp);
    if (e != hipSuccess) {
        cismgpu_set_error("basal_friction: %s", hipGetErrorString(e));
        return CISMGPU_ERR_HIP;
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_get_efvs(cismgpu_system* S, double* efvs) {
    if (!S || !S->efvs || !efvs) return CISMGPU_ERR_ARG;
    const int nzl = S->two_d ? S->nz_levels : S->d.nz;
    const long long nfl = (long long)(nzl - 1) * S->d.nx * S->d.ny;
    CG_CHECK(hipMemcpyAsync(efvs, S->efvs, nfl * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_residual(cismgpu_system* S, double* ru, double* rv, double* L2, double* L2rel) {
    if (!S || !S->have_matrix) return CISMGPU_ERR_ARG;
    const Dims& d = S->d;
    double *r0 = S->V(12), *r1 = S->V(13);
    CG_CHECK(hipMemsetAsync(r0, 0, 2 * d.vlen * sizeof(double), S->s));
    { const int st_ = reset_scal(S); if (st_) return st_; }
    // r = b - A x over owned rows with partial (r,r); resid = -r on owned active rows
    k_spmv<double, double, 2><<<spmv_blocks(d), 256, 0, S->s>>>(d, S->A64, S->V(0), S->V(1), r0, r1, S->V(2), S->V(3), S->partB);
    if (S->repro) {
        rs_dot<double>(S, 3, r0, r0, r1, r1);
        k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_DOT, S->rs_out + 6, 1, 2, nullptr, 0, 0, S->scal, 0.0);
        rs_dot<double>(S, 0, S->V(2), S->V(2), S->V(3), S->V(3));
        k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_BB, S->rs_out, 1, 2, nullptr, 0, 0, S->scal, 0.0);
    } else {
        k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_DOT, S->partB, spmv_blocks(d), 2, nullptr, 0, 0, S->scal, 0.0);
        k_dot<double><<<dot_blocks<double>(d), 256, 0, S->s>>>(d, S->V(2), S->V(2), S->V(3), S->V(3), S->partA);
        k_scalars<<<1, SCAL_THREADS, 0, S->s>>>(ST_BB, S->partA, dot_blocks<double>(d), 2, nullptr, 0, 0, S->scal, 0.0);
    }
    int st = sync_scal(S);
    if (st) return st;
    if (L2) *L2 = std::sqrt(S->scal_h[S_DOT]);
    // |Ax - b| / |b|, and 0 when b = 0 (glissade_velo_higher.F90:9881-9885)
    if (L2rel) *L2rel = S->scal_h[S_L2RHS] > 0.0 ? std::sqrt(S->scal_h[S_DOT]) / S->scal_h[S_L2RHS] : 0.0;
    if (ru && rv) {
        const long long nvec = (long long)d.nz * d.nvx * d.nvy;
        std::vector<double> h(nvec);
        double* outs[2] = {ru, rv};
        double* srcs[2] = {r0, r1};
        for (int c = 0; c < 2; ++c) {
            k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, S->s>>>(d, srcs[c], S->V(10));
            CG_CHECK(hipMemcpyAsync(outs[c], S->V(10), nvec * sizeof(double), hipMemcpyDeviceToHost, S->s));
            CG_CHECK(hipStreamSynchronize(S->s));
        }
        // resid = A x - b on owned active vertices, zero elsewhere
        for (long long t = 0; t < nvec; ++t) {
            const long long col = t / d.nz;
            const int i = 1 + (int)(col % d.nvx), j = 1 + (int)(col / d.nvx);
            const bool own = i >= d.ilo && i <= d.ihi && j >= d.jlo && j <= d.jhi;
            ru[t] = own ? -ru[t] : 0.0;
            rv[t] = own ? -rv[t] : 0.0;
        }
        std::vector<int8_t> act(d.nvx * (size_t)d.nvy);
        CG_CHECK(hipMemcpy(act.data(), S->active, act.size(), hipMemcpyDeviceToHost));
        for (long long t = 0; t < nvec; ++t)
            if (!act[t / d.nz]) { ru[t] = 0.0; rv[t] = 0.0; }
        CG_CHECK(hipMemsetAsync(S->V(10), 0, d.vlen * sizeof(double), S->s));
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_halo_x(cismgpu_system* S) {
    if (!S) return CISMGPU_ERR_ARG;
    halo<double>(S, S->V(0), S->V(1), S->d.nz);
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

// ---------------------------------------------------------------- device-resident Picard vectors
// Elementwise operations on velocity pairs kept on the device: slot -1 is
// the resident iterate x, slots 0..n-1 are work pairs.  Written without
// contraction so the results equal the host formulas bit for bit.
__global__ void k_vec_op(long long n, int op, double alpha, double* dst, const double* a, const double* b) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v;
    if (op == 0) v = a[t];
    else if (op == 1) v = a[t] + alpha * b[t];
    else if (op == 2) v = a[t] - b[t];
    else v = 0.0;
    dst[t] = v;
}

// compute_residual_velocity_3d (glissade_velo_higher.F90:10314-10446):
// max (whichresid 0, 1, 3, 4) or mean (2) over owned vertices of
// |(|u_old| - |u|) / |u]| where |u| != 0; whichresid 1 and 2 skip the bed
// level.  Per-workgroup max / compensated sum and count; the host combines.
__global__ __launch_bounds__(256) void k_resid_velo(Dims d, int skip_bed, const double* __restrict__ u,
                                                    const double* __restrict__ v, const double* __restrict__ u0,
                                                    const double* __restrict__ v0, double* __restrict__ part) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    double mx = 0.0;
    dd sum = {0.0, 0.0};
    double cnt = 0.0;
    if (t < (long long)d.nxo * d.nyo) {
        const int i = d.ilo + (int)(t % d.nxo), j = d.jlo + (int)(t / d.nxo);
        const int kn = skip_bed ? d.nz - 1 : d.nz;
        for (int k = 0; k < kn; ++k) {
            const long long c = vidx(d, k, i, j);
            const double speed = sqrt(u[c] * u[c] + v[c] * v[c]);
            if (speed != 0.0) {
                const double old = sqrt(u0[c] * u0[c] + v0[c] * v0[c]);
                const double df = fabs((old - speed) / speed);
                mx = fmax(mx, df);
                sum = dd_add_d(sum, df);
                cnt += 1.0;
            }
        }
    }
    __shared__ double smx[256 / WAVE], scn[256 / WAVE];
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_down(mx, off, WAVE));
        cnt += __shfl_down(cnt, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) { smx[threadIdx.x / WAVE] = mx; scn[threadIdx.x / WAVE] = cnt; }
    dd vv[1] = {sum};
    block_reduce_dd<1, 256>(vv);
    if (threadIdx.x == 0) {
        double m = 0.0, c = 0.0;
        for (int w = 0; w < 256 / WAVE; ++w) { m = fmax(m, smx[w]); c += scn[w]; }
        part[4 * blockIdx.x + 0] = m;
        part[4 * blockIdx.x + 1] = vv[0].hi;
        part[4 * blockIdx.x + 2] = vv[0].lo;
        part[4 * blockIdx.x + 3] = c;
    }
}

static double* vec_slot(cismgpu_system* S, int q) {
    if (q == -1) return S->V(0);
    if (q < 0 || q >= S->nwork) return nullptr;
    return S->work + 2 * (size_t)q * S->d.vlen;
}

extern "C" int32_t cismgpu_system_work_vectors(cismgpu_system* S, int32_t nslots) {
    if (!S || nslots < 0 || nslots > 16) return CISMGPU_ERR_ARG;
    if (nslots <= S->nwork) return CISMGPU_OK;
    double* w = nullptr;
    CG_CHECK(hipMalloc(&w, 2 * (size_t)nslots * S->d.vlen * sizeof(double)));
    CG_CHECK(hipMemsetAsync(w, 0, 2 * (size_t)nslots * S->d.vlen * sizeof(double), S->s));
    if (S->work) {
        CG_CHECK(hipMemcpyAsync(w, S->work, 2 * (size_t)S->nwork * S->d.vlen * sizeof(double), hipMemcpyDeviceToDevice, S->s));
        CG_CHECK(hipStreamSynchronize(S->s));
        hipFree(S->work);
    }
    CG_CHECK(hipStreamSynchronize(S->s));
    S->work = w;
    S->nwork = nslots;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_vector_op(cismgpu_system* S, int32_t op, int32_t dst, int32_t a, int32_t b,
                                            double alpha) {
    if (!S || op < 0 || op > 3) return CISMGPU_ERR_ARG;
    double* D = vec_slot(S, dst);
    const double* A = op == 3 ? D : vec_slot(S, a);
    const double* B = (op == 1 || op == 2) ? vec_slot(S, b) : A;
    if (!D || !A || !B) { cismgpu_set_error("vector_op: bad slot"); return CISMGPU_ERR_ARG; }
    const long long n = 2 * S->d.vlen;   // both components are contiguous in every slot
    k_vec_op<<<nblocks(n, 256), 256, 0, S->s>>>(n, op, alpha, D, A, B);
    CG_CHECK(hipGetLastError());
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_residual_velocity(cismgpu_system* S, int32_t whichresid, int32_t old_slot,
                                                    double* resid) {
    if (!S || !resid) return CISMGPU_ERR_ARG;
    const double* O = vec_slot(S, old_slot);
    if (!O || old_slot == -1) { cismgpu_set_error("residual_velocity: bad slot"); return CISMGPU_ERR_ARG; }
    const Dims& d = S->d;
    const int nb = (int)nblocks((long long)d.nxo * d.nyo, 256);
    double* part = nullptr;
    CG_CHECK(hipMalloc(&part, 4 * (size_t)nb * sizeof(double)));
    // compute_residual_velocity_2d has no bed level to leave out (:10188-10310)
    const int skip_bed = (!S->two_d && (whichresid == 1 || whichresid == 2)) ? 1 : 0;
    k_resid_velo<<<nb, 256, 0, S->s>>>(d, skip_bed, S->V(0), S->V(1), O, O + d.vlen, part);
    std::vector<double> h(4 * (size_t)nb);
    CG_CHECK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    hipFree(part);
    double mx = 0.0, c = 0.0;
    dd sum = {0.0, 0.0};
    for (int q = 0; q < nb; ++q) {
        mx = std::max(mx, h[4 * q]);
        const double a = h[4 * q + 1], b = h[4 * q + 2];
        // host double-double add in block order
        double s1 = sum.hi + a, bb = s1 - sum.hi, e = (sum.hi - (s1 - bb)) + (a - bb);
        e += sum.lo + b;
        sum.hi = s1 + e;
        sum.lo = e - (sum.hi - s1);
        c += h[4 * q + 3];
    }
    *resid = (whichresid == 2) ? (c > 0.0 ? (sum.hi + sum.lo) / c : 0.0) : mx;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_system_zero_x(cismgpu_system* S) {
    if (!S) return CISMGPU_ERR_ARG;
    CG_CHECK(hipMemsetAsync(S->V(0), 0, 2 * S->d.vlen * sizeof(double), S->s));
    CG_CHECK(hipStreamSynchronize(S->s));
    return CISMGPU_OK;
}

// ---------------------------------------------------------------- ensemble
struct cismgpu_ensemble {
    int nmem = 0;
    hipStream_t s = nullptr;
    std::vector<EnsMember> mem;      // host copy
    std::vector<int> side;           // owned vertices per side (bucketing)
    EnsMember* d_mem = nullptr;
    int *d_tile_mem_mat = nullptr, *d_tile_mem_col = nullptr;
    int ntile_mat = 0, ntile_col = 0;
    double* A64 = nullptr;
    float* A32 = nullptr;
    __hip_bfloat16* A16 = nullptr;
    float* S32 = nullptr;            // half (symmetric) storage of A32, when every member is exactly symmetric
    int half = 0;                    // 1: the FP32 products use S32
    float* M32 = nullptr;
    int8_t* active = nullptr;
    double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr;
    float* f32 = nullptr;            // r, d, q, z, Az, e
    double *part_mat = nullptr, *part_col = nullptr;
    EnsState* st = nullptr;
    unsigned* sync = nullptr;        // [0..1] barrier, [5] barrier error, [6] steps taken
    long long nrows_total = 0, vtotal = 0, ctotal = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    EnsParams gparams{};
    int ggrid = 0;
    int gbf16 = -1;
};

__global__ void k_ens_init(long long nvec, double* __restrict__ x, int nmem, EnsState* __restrict__ st,
                           unsigned* __restrict__ sync) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (long long q = t; q < nvec; q += (long long)gridDim.x * blockDim.x) x[q] = 0.0;
    if (t < 2 * nmem) {
        EnsState z{};
        z.phase = M_OUTER;
        st[t] = z;
    }
    if (t < 8) sync[t] = 0u;
}

static int ens_tiles(cismgpu_ensemble* E) {
    std::vector<int> tm, tc;
    int tmat = 0, tcol = 0;
    for (int m = 0; m < E->nmem; ++m) {
        EnsMember& M = E->mem[m];
        M.ncg = (M.ncol + WAVE - 1) / WAVE;
        M.ntile_mat = M.ncg * M.split;
        M.ntile_col = (int)(((long long)M.d.nvx * M.d.nvy + WAVE - 1) / WAVE);
        M.t0_mat = tmat;
        M.t0_col = tcol;
        for (int t = 0; t < M.ntile_mat; ++t) tm.push_back(m);
        for (int t = 0; t < M.ntile_col; ++t) tc.push_back(m);
        tmat += M.ntile_mat;
        tcol += M.ntile_col;
    }
    E->ntile_mat = tmat;
    E->ntile_col = tcol;
    if (E->d_tile_mem_mat) hipFree(E->d_tile_mem_mat);
    if (E->d_tile_mem_col) hipFree(E->d_tile_mem_col);
    if (E->part_mat) hipFree(E->part_mat);
    if (E->part_col) hipFree(E->part_col);
    E->d_tile_mem_mat = E->d_tile_mem_col = nullptr;
    E->part_mat = E->part_col = nullptr;
    if (E->gexec) { hipGraphExecDestroy(E->gexec); E->gexec = nullptr; }
    if (E->graph) { hipGraphDestroy(E->graph); E->graph = nullptr; }
    CG_CHECK(hipMalloc(&E->d_tile_mem_mat, tm.size() * sizeof(int)));
    CG_CHECK(hipMalloc(&E->d_tile_mem_col, tc.size() * sizeof(int)));
    CG_CHECK(hipMemcpy(E->d_tile_mem_mat, tm.data(), tm.size() * sizeof(int), hipMemcpyHostToDevice));
    CG_CHECK(hipMemcpy(E->d_tile_mem_col, tc.data(), tc.size() * sizeof(int), hipMemcpyHostToDevice));
    CG_CHECK(hipMalloc(&E->part_mat, 2 * (size_t)tmat * sizeof(double)));
    CG_CHECK(hipMalloc(&E->part_col, 2 * 4 * (size_t)tcol * sizeof(double)));
    CG_CHECK(hipMemcpy(E->d_mem, E->mem.data(), E->nmem * sizeof(EnsMember), hipMemcpyHostToDevice));
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_ensemble_destroy(cismgpu_ensemble* E) {
    if (!E) return CISMGPU_OK;
    if (E->s) hipStreamSynchronize(E->s);
    if (E->gexec) hipGraphExecDestroy(E->gexec);
    if (E->graph) hipGraphDestroy(E->graph);
    for (void* p : {(void*)E->d_mem, (void*)E->d_tile_mem_mat, (void*)E->d_tile_mem_col, (void*)E->A64, (void*)E->A32, (void*)E->S32,
                    (void*)E->A16, (void*)E->M32, (void*)E->active, (void*)E->b64, (void*)E->x64, (void*)E->r64,
                    (void*)E->f32, (void*)E->part_mat, (void*)E->part_col, (void*)E->st, (void*)E->sync})
        if (p) hipFree(p);
    if (E->s) hipStreamDestroy(E->s);
    delete E;
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_ensemble_create(int32_t n, cismgpu_system* const* members, cismgpu_ensemble** out) {
    if (!members || n < 1 || !out) return CISMGPU_ERR_ARG;
    *out = nullptr;
    cismgpu_ensemble* E = new cismgpu_ensemble();
    E->nmem = n;
    auto fail = [&](int code, const char* msg) { cismgpu_set_error("ensemble: %s", msg); cismgpu_ensemble_destroy(E); return code; };
    if (hipStreamCreateWithFlags(&E->s, hipStreamNonBlocking) != hipSuccess) return fail(CISMGPU_ERR_HIP, "stream");
    long long rows = 0, voff = 0, coff = 0;
    for (int m = 0; m < n; ++m) {
        cismgpu_system* S = members[m];
        if (!S || !S->have_matrix) return fail(CISMGPU_ERR_ARG, "member without an assembled matrix");
        const Dims& d = S->d;
        if (S->two_d) return fail(CISMGPU_ERR_ARG, "members must be 3-D systems");
        if (d.nz > 16) return fail(CISMGPU_ERR_ARG, "members need nz <= 16");
        if (d.bc != 1 && d.bc != 3) return fail(CISMGPU_ERR_ARG, "members need the no_ice or outflow halo rule");
        // active vertices must avoid every row/column the halo rule writes
        std::vector<int8_t> act((size_t)d.nvx * d.nvy);
        if (hipMemcpy(act.data(), S->active, act.size(), hipMemcpyDeviceToHost) != hipSuccess) return fail(CISMGPU_ERR_HIP, "copy");
        const int lw = d.bc == 3 ? 2 : 1, uw = d.bc == 3 ? 2 : 1;
        for (int j = 1; j <= d.nvy; ++j)
            for (int i = 1; i <= d.nvx; ++i)
                if (act[(size_t)(j - 1) * d.nvx + (i - 1)] && (i <= lw || i >= d.nvx - uw + 1 || j <= lw || j >= d.nvy - uw + 1))
                    return fail(CISMGPU_ERR_ARG, "member has active vertices on halo-rule rows (needs an ice-free outer ring)");
        EnsMember M{};
        M.d = d;
        M.ncol = d.nxo * d.nyo;
        M.ncolp = (M.ncol + WAVE - 1) / WAVE * WAVE;
        M.split = 1;
        M.a_off = rows;
        M.v_off = voff;
        M.c_off = coff;
        rows += (long long)d.nz * M.ncolp;
        voff += 2 * d.vlen;
        coff += (long long)d.nvx * d.nvy;
        // ||b||
        int st = reset_scal(S);
        if (st) return fail(st, "scal");
        k_dot<double><<<dot_blocks<double>(d), 256, 0, S->s>>>(d, S->V(2), S->V(2), S->V(3), S->V(3), S->partB);
        scalars(S, ST_BB, S->partB, dot_blocks<double>(d), 2, nullptr, 0, 0);
        st = sync_scal(S);
        if (st) return fail(st, "norm");
        M.bnorm = S->scal_h[S_L2RHS];
        E->mem.push_back(M);
        E->side.push_back(std::max(d.nxo, d.nyo));
    }
    E->nrows_total = rows;
    E->vtotal = voff;   // doubles per vector pool (both components)
    E->ctotal = coff;
    const size_t na = (size_t)rows * NCOEF;
    hipError_t e = hipSuccess;
    // half (symmetric) storage for the FP32 products when every member's FP32
    // matrix is exactly symmetric (k_check_sym) and each member's part stays
    // under 2 GiB, unless CISMGPU_SPMV_FULL=1
    {
        const char* env = getenv("CISMGPU_SPMV_FULL");
        const size_t s_bytes = (size_t)rows * 56 * sizeof(float);
        bool sym = !(env && env[0] == '1');
        for (const EnsMember& M : E->mem)   // member-local 32-bit byte offsets
            if ((size_t)M.d.nz * M.ncolp * 56 * sizeof(float) >= ((size_t)1 << 31)) sym = false;
        unsigned long long* cnt = nullptr;
        if (sym && hipMalloc(&cnt, sizeof(unsigned long long)) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc sym");
        for (int m = 0; sym && m < n; ++m) {
            cismgpu_system* S = members[m];
            unsigned long long h = 1;
            if (hipMemsetAsync(cnt, 0, sizeof(h), S->s) != hipSuccess) { hipFree(cnt); return fail(CISMGPU_ERR_HIP, "sym"); }
            k_check_sym<float><<<nblocks(S->d.nrows, 256), 256, 0, S->s>>>(S->d, S->A64, cnt);
            if (hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, S->s) != hipSuccess ||
                hipStreamSynchronize(S->s) != hipSuccess) { hipFree(cnt); return fail(CISMGPU_ERR_HIP, "sym"); }
            sym = h == 0;
        }
        if (cnt) hipFree(cnt);
        if (sym) {
            if ((e = hipMalloc(&E->S32, s_bytes)) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc S32");
            for (int m = 0; m < n; ++m) {
                const EnsMember& M = E->mem[m];
                const long long nel = (long long)M.d.nz * M.ncolp * 56;
                k_ens_pack_half<<<nblocks(nel, 256), 256, 0, E->s>>>(M.d, M.a_off, M.ncolp, members[m]->A64, E->S32);
            }
            if ((e = hipStreamSynchronize(E->s)) != hipSuccess) return fail(CISMGPU_ERR_HIP, "pack half");
            E->half = 1;
        }
    }
    if ((e = hipMalloc(&E->d_mem, n * sizeof(EnsMember))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc");
    if ((e = hipMalloc(&E->A64, na * sizeof(double))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc A64");
    if (!E->half && (e = hipMalloc(&E->A32, na * sizeof(float))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc A32");
    if ((e = hipMalloc(&E->M32, 3 * (size_t)voff * sizeof(float))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc M");
    if ((e = hipMalloc(&E->active, coff)) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc active");
    if ((e = hipMalloc(&E->b64, voff * sizeof(double))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc b");
    if ((e = hipMalloc(&E->x64, voff * sizeof(double))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc x");
    if ((e = hipMalloc(&E->r64, voff * sizeof(double))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc r");
    if ((e = hipMalloc(&E->f32, 6 * (size_t)voff * sizeof(float))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc f32");
    if ((e = hipMalloc(&E->st, 2 * n * sizeof(EnsState))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc st");
    if ((e = hipMalloc(&E->sync, 8 * sizeof(unsigned))) != hipSuccess) return fail(CISMGPU_ERR_HIP, "alloc sync");
    hipMemsetAsync(E->f32, 0, 6 * (size_t)voff * sizeof(float), E->s);
    hipMemsetAsync(E->r64, 0, voff * sizeof(double), E->s);
    hipMemsetAsync(E->x64, 0, voff * sizeof(double), E->s);
    for (int m = 0; m < n; ++m) {
        cismgpu_system* S = members[m];
        const EnsMember& M = E->mem[m];
        const Dims& d = M.d;
        const long long nel = (long long)d.nz * M.ncolp * NCOEF;
        hipStreamSynchronize(S->s);
        k_ens_pack<double><<<nblocks(nel, 256), 256, 0, E->s>>>(d, M.a_off, M.ncolp, S->A64, E->A64);
        if (!E->half) k_ens_pack<float><<<nblocks(nel, 256), 256, 0, E->s>>>(d, M.a_off, M.ncolp, S->A64, E->A32);
        k_cast<double, float><<<nblocks(6 * d.vlen, 256), 256, 0, E->s>>>(6 * d.vlen, S->M64, E->M32 + 3 * M.v_off);
        hipMemcpyAsync(E->b64 + M.v_off, S->V(2), 2 * d.vlen * sizeof(double), hipMemcpyDeviceToDevice, E->s);
        hipMemcpyAsync(E->active + M.c_off, S->active, (size_t)d.nvx * d.nvy, hipMemcpyDeviceToDevice, E->s);
    }
    if ((e = hipStreamSynchronize(E->s)) != hipSuccess) return fail(CISMGPU_ERR_HIP, "build");
    if (ens_tiles(E) != CISMGPU_OK) return fail(CISMGPU_ERR_HIP, "tiles");
    *out = E;
    return CISMGPU_OK;
}

static int ens_bucket(int side) { return side <= 80 ? 0 : (side <= 136 ? 1 : 2); }

extern "C" int32_t cismgpu_ensemble_autotune(cismgpu_ensemble* E, int32_t* split_out, double* ms_out) {
    if (!E) return CISMGPU_ERR_ARG;
    const int splits[3] = {1, 2, 4};
    int best[3] = {1, 1, 1};
    double tbest[3] = {1e30, 1e30, 1e30};
    float* z = E->f32 + 3 * E->vtotal;
    float* Az = E->f32 + 4 * E->vtotal;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int si = 0; si < 3; ++si) {
        for (auto& M : E->mem) M.split = splits[si];
        { const int st_ = ens_tiles(E); if (st_) return st_; }
        for (int b = 0; b < 3; ++b) {
            std::vector<int> tl;
            for (int m = 0; m < E->nmem; ++m)
                if (ens_bucket(E->side[m]) == b)
                    for (int t = 0; t < E->mem[m].ntile_mat; ++t) tl.push_back(E->mem[m].t0_mat + t);
            double ms = 0.0;
            if (!tl.empty()) {
                int* dtl = nullptr;
                CG_CHECK(hipMalloc(&dtl, tl.size() * sizeof(int)));
                CG_CHECK(hipMemcpy(dtl, tl.data(), tl.size() * sizeof(int), hipMemcpyHostToDevice));
                const unsigned g = nblocks((long long)tl.size(), ENS_WPB);
                k_ens_spmv_sweep<<<g, ENS_BS, 0, E->s>>>(E->d_mem, dtl, (int)tl.size(), E->d_tile_mem_mat, E->A32, E->S32, E->half, z, Az, E->part_mat);
                hipEventRecord(e0, E->s);
                for (int r = 0; r < 5; ++r)
                    k_ens_spmv_sweep<<<g, ENS_BS, 0, E->s>>>(E->d_mem, dtl, (int)tl.size(), E->d_tile_mem_mat, E->A32, E->S32, E->half, z, Az, E->part_mat);
                hipEventRecord(e1, E->s);
                hipEventSynchronize(e1);
                float f = 0.f;
                hipEventElapsedTime(&f, e0, e1);
                ms = f / 5.0;
                hipFree(dtl);
            }
            if (ms_out) ms_out[b * 3 + si] = ms;
            if (ms < tbest[b]) { tbest[b] = ms; best[b] = splits[si]; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (int m = 0; m < E->nmem; ++m) E->mem[m].split = best[ens_bucket(E->side[m])];
    { const int st_ = ens_tiles(E); if (st_) return st_; }
    CG_CHECK(hipMemsetAsync(Az, 0, E->vtotal * sizeof(float), E->s));
    CG_CHECK(hipStreamSynchronize(E->s));
    if (split_out) for (int b = 0; b < 3; ++b) split_out[b] = best[b];
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_ensemble_solve(cismgpu_ensemble* E, const cismgpu_ensemble_opts* o,
                                          cismgpu_ensemble_info* info) {
    if (!E || !o || !info) return CISMGPU_ERR_ARG;
    if (o->ncheck < 1 || o->inner_maxiters < 1 || o->max_outer < 1) return CISMGPU_ERR_ARG;
    memset(info, 0, sizeof(*info));
    if (o->bf16 && !E->A16) {
        const size_t na = (size_t)E->nrows_total * NCOEF;
        CG_CHECK(hipMalloc(&E->A16, na * sizeof(__hip_bfloat16)));
        k_convert_matrix<__hip_bfloat16><<<nblocks((long long)na, 256), 256, 0, E->s>>>((long long)na, E->A64, E->A16);
    }
    int dev = 0, ncu = 0, per = 0;
    CG_CHECK(hipGetDevice(&dev));
    CG_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    CG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ensemble<16>, ENS_BS, 0));
    int bpc = o->blocks_per_cu > 0 ? std::min(o->blocks_per_cu, per) : per;
    if (bpc < 1) { cismgpu_set_error("ensemble kernel cannot be resident"); return CISMGPU_ERR_HIP; }
    const int grid = bpc * ncu;
    EnsParams P{};
    P.nmem = E->nmem; P.ntile_mat = E->ntile_mat; P.ntile_col = E->ntile_col;
    P.inner_maxiters = o->inner_maxiters; P.max_outer = o->max_outer; P.ncheck = o->ncheck;
    P.tol = o->tolerance; P.inner_tol = o->inner_tol;
    P.maxsteps = (o->inner_maxiters + 3) * (o->max_outer + 1) + 8;
    P.spin_limit = 50000000LL;
    EnsPtrs Q{};
    Q.mem = E->d_mem; Q.tile_mem_mat = E->d_tile_mem_mat; Q.tile_mem_col = E->d_tile_mem_col;
    Q.A64 = E->A64; Q.A32 = E->A32; Q.A16 = E->A16; Q.M32 = E->M32; Q.active = E->active;
    Q.b64 = E->b64; Q.x64 = E->x64; Q.r64 = E->r64;
    const long long V = E->vtotal;
    Q.r32 = E->f32; Q.d32 = E->f32 + V; Q.q32 = E->f32 + 2 * V; Q.z32 = E->f32 + 3 * V; Q.Az32 = E->f32 + 4 * V;
    Q.e32 = E->f32 + 5 * V;
    Q.part_mat = E->part_mat; Q.part_col = E->part_col; Q.st = E->st;
    Q.bar = E->sync; Q.err_flag = (int*)(E->sync + 5); Q.steps_out = (int*)(E->sync + 6);
    Q.use_bf16 = o->bf16 ? 1 : 0;
    Q.S32 = E->S32;
    Q.half = o->bf16 ? 0 : E->half;
    // (re)capture the graph: init kernel + persistent solve
    const bool same = E->gexec && E->ggrid == grid && E->gbf16 == Q.use_bf16 && !memcmp(&E->gparams, &P, sizeof(P));
    if (!same) {
        if (E->gexec) { hipGraphExecDestroy(E->gexec); E->gexec = nullptr; }
        if (E->graph) { hipGraphDestroy(E->graph); E->graph = nullptr; }
        CG_CHECK(hipStreamBeginCapture(E->s, hipStreamCaptureModeThreadLocal));
        k_ens_init<<<1024, 256, 0, E->s>>>(V, E->x64, E->nmem, E->st, E->sync);
        k_ensemble<16><<<grid, ENS_BS, 0, E->s>>>(P, Q);
        CG_CHECK(hipStreamEndCapture(E->s, &E->graph));
        CG_CHECK(hipGraphInstantiate(&E->gexec, E->graph, nullptr, nullptr, 0));
        E->gparams = P;
        E->ggrid = grid;
        E->gbf16 = Q.use_bf16;
    }
    hipEvent_t e0, e1;
    CG_CHECK(hipEventCreate(&e0));
    CG_CHECK(hipEventCreate(&e1));
    CG_CHECK(hipEventRecord(e0, E->s));
    CG_CHECK(hipGraphLaunch(E->gexec, E->s));
    CG_CHECK(hipEventRecord(e1, E->s));
    CG_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    info->solve_ms = ms;
    info->grid_blocks = grid;
    info->matrix_storage = o->bf16 ? 2 : (Q.half ? 3 : 1);
    unsigned sy[8];
    CG_CHECK(hipMemcpy(sy, E->sync, sizeof(sy), hipMemcpyDeviceToHost));
    info->barrier_error = ((int*)sy)[5] ? 1 : 0;
    info->steps = ((int*)sy)[6];
    std::vector<EnsState> st(2 * E->nmem);
    CG_CHECK(hipMemcpy(st.data(), E->st, st.size() * sizeof(EnsState), hipMemcpyDeviceToHost));
    // the final state is in the buffer of the parity after the last step: take, per member, the
    // buffer whose phase is DONE (both agree once a member is done)
    for (int m = 0; m < E->nmem; ++m) {
        const EnsState& a = st[m].phase == M_DONE ? st[m] : st[E->nmem + m];
        if (a.flags & 1) info->converged++;
        info->max_outer = std::max(info->max_outer, a.outer_it);
        info->max_inner = std::max(info->max_inner, a.total_inner);
        info->total_inner += a.total_inner;
        info->max_err = std::max(info->max_err, a.err);
    }
    if (info->barrier_error) { cismgpu_set_error("ensemble grid barrier timed out"); return CISMGPU_ERR_HIP; }
    if (info->converged != E->nmem) {
        cismgpu_set_error("%d of %d ensemble members reached the tolerance (largest residual %.3e after %d steps)",
                          info->converged, E->nmem, info->max_err, info->steps);
        return CISMGPU_ERR_NOT_CONVERGED;
    }
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_ensemble_get_x(cismgpu_ensemble* E, int32_t m, double* xu, double* xv) {
    if (!E || m < 0 || m >= E->nmem || !xu || !xv) return CISMGPU_ERR_ARG;
    const EnsMember& M = E->mem[m];
    const Dims& d = M.d;
    const long long nvec = (long long)d.nz * d.nvx * d.nvy;
    double* tmp = nullptr;
    CG_CHECK(hipMalloc(&tmp, nvec * sizeof(double)));
    double* outs[2] = {xu, xv};
    for (int c = 0; c < 2; ++c) {
        k_vec_out<double, double><<<nblocks(nvec, 256), 256, 0, E->s>>>(d, E->x64 + M.v_off + c * d.vlen, tmp);
        hipMemcpyAsync(outs[c], tmp, nvec * sizeof(double), hipMemcpyDeviceToHost, E->s);
        hipStreamSynchronize(E->s);
    }
    hipFree(tmp);
    return CISMGPU_OK;
}

extern "C" int32_t cismgpu_ensemble_member_info(cismgpu_ensemble* E, int32_t m, double* err, int32_t* outer,
                                                int32_t* total_inner, int32_t* flags) {
    if (!E || m < 0 || m >= E->nmem) return CISMGPU_ERR_ARG;
    EnsState a[2];
    CG_CHECK(hipMemcpy(&a[0], E->st + m, sizeof(EnsState), hipMemcpyDeviceToHost));
    CG_CHECK(hipMemcpy(&a[1], E->st + E->nmem + m, sizeof(EnsState), hipMemcpyDeviceToHost));
    const EnsState& s = a[0].phase == M_DONE ? a[0] : a[1];
    if (err) *err = s.err;
    if (outer) *outer = s.outer_it;
    if (total_inner) *total_inner = s.total_inner;
    if (flags) *flags = s.flags;
    return CISMGPU_OK;
}

// ---------------------------------------------------------------- drop-in entry points
// The host-array drop-ins keep one device system per (device, grid, 2-D)
// between calls: a Picard loop calls the solver with one grid every
// iteration, so its device buffers are allocated once (the reference's
// solvers allocate their work arrays on every call).  Nothing else carries
// over: the matrix, rhs and guess are uploaded by every call and every
// coefficient row is rewritten.  cismgpu_dropin_release frees the system.
struct DropinCache {
    int dev = -1;
    cismgpu_grid g{};
    bool two_d = false;
    cismgpu_system* S = nullptr;
};
static std::mutex g_dropin_mu;
static DropinCache g_dropin;

// the cached system for this grid (made if absent); g_dropin_mu held
static int dropin_system(const cismgpu_grid* grid, bool two_d, cismgpu_system** out) {
    if (!grid) { cismgpu_set_error("null grid"); return CISMGPU_ERR_ARG; }
    int dev = 0;
    CG_CHECK(hipGetDevice(&dev));
    if (g_dropin.S && g_dropin.dev == dev && g_dropin.two_d == two_d &&
        !memcmp(&g_dropin.g, grid, sizeof(cismgpu_grid))) {
        *out = g_dropin.S;
        return CISMGPU_OK;
    }
    if (g_dropin.S) {
        cismgpu_system_destroy(g_dropin.S);
        g_dropin.S = nullptr;
    }
    cismgpu_system* S = nullptr;
    const int st = two_d ? cismgpu_system_create_2d(grid, &S) : cismgpu_system_create(grid, &S);
    if (st) return st;
    g_dropin.dev = dev;
    g_dropin.g = *grid;
    g_dropin.two_d = two_d;
    g_dropin.S = S;
    *out = S;
    return CISMGPU_OK;
}

// after a HIP failure the cached system is not trusted again
static void dropin_drop_on(int st) {
    if (st == CISMGPU_ERR_HIP && g_dropin.S) {
        cismgpu_system_destroy(g_dropin.S);
        g_dropin.S = nullptr;
    }
}

int32_t cismgpu_dropin_release(void) {
    std::lock_guard<std::mutex> lock(g_dropin_mu);
    if (g_dropin.S) cismgpu_system_destroy(g_dropin.S);
    g_dropin.S = nullptr;
    return CISMGPU_OK;
}

static int dropin(int algo, const cismgpu_grid* grid, const int32_t* active, const double* Auu, const double* Auv,
                  const double* Avu, const double* Avv, const double* bu, const double* bv, double* xu, double* xv,
                  int32_t precond, int32_t ncheck, double tol, int32_t maxiters, double* err, int32_t* niters,
                  bool two_d = false) {
    if (!active || !Auu || !Auv || !Avu || !Avv || !bu || !bv || !xu || !xv || !err || !niters) {
        cismgpu_set_error("null argument");
        return CISMGPU_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_dropin_mu);
    cismgpu_system* S = nullptr;
    int st = dropin_system(grid, two_d, &S);
    if (st) return st;
    st = two_d ? cismgpu_system_set_matrix_2d(S, active, Auu, Auv, Avu, Avv, bu, bv, 0)
               : cismgpu_system_set_matrix(S, active, Auu, Auv, Avu, Avv, bu, bv, 0);
    if (!st) st = cismgpu_system_set_x(S, xu, xv);
    cismgpu_solve_info info{};
    if (!st) {
        cismgpu_solve_opts o{};
        o.algo = algo; o.precond = precond; o.ncheck = ncheck; o.maxiters = maxiters; o.tolerance = tol;
        o.precision = CISMGPU_FP64;
        st = cismgpu_system_solve(S, &o, &info);
    }
    if (!st || st == CISMGPU_ERR_NAN) {
        int st2 = cismgpu_system_get_x(S, xu, xv);
        if (!st) st = st2;
        if (info.err == info.err) *err = info.err;   // untouched when no check ran
        *niters = info.niters;
    }
    dropin_drop_on(st);
    return st;
}

int32_t cismgpu_pcg_solver_standard_3d(const cismgpu_grid* grid, const int32_t* active_vertex, const double* Auu,
                                       const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                       const double* bv, double* xu, double* xv, int32_t precond,
                                       int32_t linear_solve_ncheck, double tolerance, int32_t maxiters, double* err,
                                       int32_t* niters) {
    return dropin(CISMGPU_ALGO_STANDARD, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                  linear_solve_ncheck, tolerance, maxiters, err, niters);
}

int32_t cismgpu_pcg_solver_chrongear_3d(const cismgpu_grid* grid, const int32_t* active_vertex, const double* Auu,
                                        const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                        const double* bv, double* xu, double* xv, int32_t precond,
                                        int32_t linear_solve_ncheck, double tolerance, int32_t maxiters, double* err,
                                        int32_t* niters) {
    return dropin(CISMGPU_ALGO_CHRONGEAR, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                  linear_solve_ncheck, tolerance, maxiters, err, niters);
}

int32_t cismgpu_matvec_multiply_structured_3d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                              const double* Auu, const double* Auv, const double* Avu,
                                              const double* Avv, const double* xu, const double* xv, double* yu,
                                              double* yv) {
    std::lock_guard<std::mutex> lock(g_dropin_mu);
    cismgpu_system* S = nullptr;
    int st = dropin_system(grid, false, &S);
    if (st) return st;
    const Dims& d = S->d;
    std::vector<double> zero((size_t)d.nz * d.nvx * d.nvy, 0.0);
    st = cismgpu_system_set_matrix(S, active_vertex, Auu, Auv, Avu, Avv, zero.data(), zero.data(), 0);
    if (!st) st = cismgpu_system_set_x(S, xu, xv);
    if (!st) st = cismgpu_system_matvec(S, yu, yv);
    dropin_drop_on(st);
    return st;
}

int32_t cismgpu_pcg_solver_standard_2d(const cismgpu_grid* grid, const int32_t* active_vertex, const double* Auu,
                                       const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                       const double* bv, double* xu, double* xv, int32_t precond,
                                       int32_t linear_solve_ncheck, double tolerance, int32_t maxiters, double* err,
                                       int32_t* niters) {
    return dropin(CISMGPU_ALGO_STANDARD, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                  linear_solve_ncheck, tolerance, maxiters, err, niters, true);
}

int32_t cismgpu_pcg_solver_chrongear_2d(const cismgpu_grid* grid, const int32_t* active_vertex, const double* Auu,
                                        const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                        const double* bv, double* xu, double* xv, int32_t precond,
                                        int32_t linear_solve_ncheck, double tolerance, int32_t maxiters, double* err,
                                        int32_t* niters) {
    return dropin(CISMGPU_ALGO_CHRONGEAR, grid, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond,
                  linear_solve_ncheck, tolerance, maxiters, err, niters, true);
}

int32_t cismgpu_matvec_multiply_structured_2d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                              const double* Auu, const double* Auv, const double* Avu,
                                              const double* Avv, const double* xu, const double* xv, double* yu,
                                              double* yv) {
    std::lock_guard<std::mutex> lock(g_dropin_mu);
    cismgpu_system* S = nullptr;
    int st = dropin_system(grid, true, &S);
    if (st) return st;
    const Dims& d = S->d;
    std::vector<double> zero((size_t)d.nvx * d.nvy, 0.0);
    st = cismgpu_system_set_matrix_2d(S, active_vertex, Auu, Auv, Avu, Avv, zero.data(), zero.data(), 0);
    if (!st) st = cismgpu_system_set_x(S, xu, xv);
    if (!st) st = cismgpu_system_matvec(S, yu, yv);
    dropin_drop_on(st);
    return st;
}

}  // extern "C"
