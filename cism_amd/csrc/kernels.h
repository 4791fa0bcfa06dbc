// kernels.h -- declarations shared by kernels.hip and the host drivers.
#pragma once
#include "cismgpu_internal.h"

namespace cg {

// slots of the device scalar block
enum {
    S_ALPHA = 0, S_BETA, S_RHO, S_SIGMA, S_ETA0, S_BB, S_L2RHS, S_RR, S_ERR, S_CONV,
    S_NAN, S_STOP, S_DOT, S_DOT2, S_NSLOTS = 16
};
// stages of k_scalars
enum { ST_BB = 0, ST_RR, ST_CG_FIRST, ST_CG_ITER, ST_STD_ETA1, ST_STD_ETA2, ST_DOT };

template <typename TA, typename TV, int MODE>
__global__ void k_spmv(Dims d, const TA* A, const TV* xu, const TV* xv, TV* yu, TV* yv, const TV* bu,
                       const TV* bv, double* partials);
template <typename TA, int MODE>
__global__ void k_spmv_q(Dims d, const TA* A, const float* xu, const float* xv, float* yu, float* yv,
                         const float* bu, const float* bv, double* partials);
template <typename TA, typename TV, int MODE>
__global__ void k_spmv_plane(Dims d, const TA* A, const TV* xu, const TV* xv, TV* yu, TV* yv, const TV* bu,
                             const TV* bv, double* partials);
__global__ void k_expand_stencil_2d(long long n2, const double* src, double* dst);
template <typename TA>
__global__ void k_convert_quad(long long nrows, const double* src, TA* dst);
template <typename TV>
__global__ void k_ring(Dims d, TV* yu, TV* yv, const TV* bu, const TV* bv);
template <typename TV>
__global__ void k_halo_ew(Dims d, TV* a, TV* b, int nlev);
template <typename TV>
__global__ void k_halo_ns(Dims d, TV* a, TV* b, int nlev);
template <typename TV>
__global__ void k_halo_zero(Dims d, TV* a, TV* b, int nlev, int lw, int uw);
template <typename TV, int NZMAX, int PRE, int STD>
__global__ void k_update(Dims d, const double* scal, TV* xu, TV* xv, TV* ru, TV* rv, TV* du, TV* dv, TV* qu,
                         TV* qv, TV* zu, TV* zv, const TV* Azu, const TV* Azv, const TV* M, const int8_t* active,
                         double* partials);
template <typename TV, int NZMAX>
__global__ void k_precond(Dims d, int kind, const double* scal, const TV* ru, const TV* rv, TV* zu, TV* zv,
                          const TV* M, const TV* Dg, const int8_t* active, const TV* zvs_u, const TV* zvs_v,
                          double* partials);
template <typename TV>
__global__ void k_vsum(Dims d, const TV* ru, const TV* rv, TV* su, TV* sv);

template <typename TV>
__global__ void k_line_scan(Dims d, int dir, const TV* Ad, const TV* As, const TV* de, const TV* om, const TV* b,
                            TV* x, TV* gam);
template <typename TV>
__global__ void k_halo2d_ew(Dims d, TV* a, TV* b);
template <typename TV>
__global__ void k_halo2d_ns(Dims d, TV* a, TV* b);
template <typename TV>
__global__ void k_halo2d_zero(Dims d, TV* a, TV* b);
template <typename TV>
__global__ void k_axpy_ops(Dims d, int op, const double* scal, const TV* s1u, const TV* s1v, TV* tu, TV* tv,
                           const TV* s2u, const TV* s2v);
template <typename TV>
__global__ void k_first_update(Dims d, const double* scal, TV* xu, TV* xv, TV* ru, TV* rv, const TV* du,
                               const TV* dv, const TV* qu, const TV* qv);
__global__ void k_scalars(int stage, const double* pa, int na, int sa, const double* pb, int nb, int sb,
                          double* scal, double tolerance);
template <typename TA>
__global__ void k_pack_matrix(Dims d, int blk, const double* src, const int8_t* active, TA* dst);
template <typename TA>
__global__ void k_convert_matrix(long long n, const double* src, TA* dst);
__global__ void k_unpack_matrix(Dims d, int blk, const double* A, double* dst);
template <typename TS, typename TD>
__global__ void k_vec_in(Dims d, const TS* src, TD* dst);
template <typename TS, typename TD>
__global__ void k_vec_out(Dims d, const TS* src, TD* dst);
template <typename TV>
__global__ void k_setup_precond(Dims d, const double* Auu, const double* Avv, TV* M, TV* Dg);
template <typename TS, typename TD>
__global__ void k_cast(long long n, const TS* s, TD* d);
__global__ void k_ir_update(Dims d, double* xu, double* xv, const float* eu, const float* ev);

}  // namespace cg
