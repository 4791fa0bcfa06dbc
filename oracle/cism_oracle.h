/*
 * cism_oracle.h -- CPU restatement of CISM's Glissade higher-order velocity
 * hot path (Blatter-Pattyn assembly + structured 3-D PCG solvers).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in cism_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Layouts follow the reference exactly (Fortran column-major, 1-based
 * indices mapped to 0-based storage):
 *   vectors  x(nz, nx-1, ny-1)         -> x[(k-1) + nz*((i-1) + (nx-1)*(j-1))]
 *   matrices A(27, nz, nx-1, ny-1)     -> A[(m-1) + 27*vec_index(k,i,j)]
 *   cell fields f(nx, ny)              -> f[(i-1) + nx*(j-1)]
 *   vertex fields f(nx-1, ny-1)        -> f[(i-1) + (nx-1)*(j-1)]
 *   flwafact/efvs(nz-1, nx, ny)        -> f[(k-1) + (nz-1)*((i-1) + nx*(j-1))]
 * m = 1 + (iA+1) + 3(jA+1) + 9(kA+1)  (glissade_velo_higher.F90:607-627).
 */
#ifndef CISM_ORACLE_H
#define CISM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of the fields of parallel_type that the solver path reads
 * (glissade_velo_higher_pcg.F90:2888-2894; parallel_mpi.F90:2633-2653).
 * Bounds are 1-based, exactly as parallel%staggered_*. */
typedef struct orc_grid {
    int32_t nx, ny, nz;
    int32_t ilo, ihi, jlo, jhi;
    int32_t global_bc;   /* 0 periodic, 1 outflow, 2 no_penetration, 3 no_ice (glide_types.F90:442) */
    int32_t reprosum;    /* 1: compensated (DDPDD) global sums */
} orc_grid;

/* status codes shared with the product C-ABI */
#define ORC_OK              0
#define ORC_ERR_NAN         1   /* alpha or eta1 is NaN (pcg.F90:1331,1391,3263,3434) */
#define ORC_ERR_DIVZERO     2   /* zero pivot in tridiagonal setup (pcg.F90:4338,4474) */
#define ORC_ERR_ARG         3
#define ORC_ERR_SINGULAR    4   /* singular element Jacobian (glissade_velo_higher.F90:6611) */
#define ORC_ERR_ASYM        5   /* element matrix not symmetric (glissade_velo_higher.F90:10583) */

/* ---------------- parallel layer semantics (one rank) ---------------- */
void   orc_halo3d(const orc_grid* g, int nlev, double* a);
double orc_global_sum(const orc_grid* g, const double* a1, const double* a2);
void   orc_global_sum2(const orc_grid* g, const double* a1f1, const double* a2f1,
                       const double* a1f2, const double* a2f2, double out[2]);
void   orc_ddpdd_sum(int64_t n, const double* arr, double* hi, double* lo);
double orc_reprosum_calc(int64_t n, const double* arr);

/* ---------------- solver path ---------------- */
void orc_matvec_3d(const orc_grid* g, const int32_t* active_vertex,
                   const double* Auu, const double* Auv, const double* Avu, const double* Avv,
                   const double* xu, const double* xv, double* yu, double* yv);

void orc_tridiag_solver(int order, const double* sbdiag, const double* diag,
                        const double* spdiag, const double* rhs, double* soln);

/* 2-D solvers (nz of the grid is ignored): A(nx-1,ny-1,9), m last; vectors (nx-1,ny-1) */
int orc_pcg_solver_standard_2d(const orc_grid* g, const int32_t* active_vertex, const double* Auu,
                               const double* Auv, const double* Avu, const double* Avv, const double* bu,
                               const double* bv, double* xu, double* xv, int32_t precond, int32_t ncheck,
                               double tolerance, int32_t maxiters, double* err, int32_t* niters);
int orc_pcg_solver_chrongear_2d(const orc_grid* g, const int32_t* active_vertex, const double* Auu,
                                const double* Auv, const double* Avu, const double* Avv, const double* bu,
                                const double* bv, double* xu, double* xv, int32_t precond, int32_t ncheck,
                                double tolerance, int32_t maxiters, double* err, int32_t* niters);
void orc_matvec_2d(const orc_grid* g, const int32_t* active_vertex, const double* Auu, const double* Auv,
                   const double* Avu, const double* Avv, const double* xu, const double* xv, double* yu,
                   double* yv);
int orc_pcg_solver_standard_3d(const orc_grid* g, const int32_t* active_vertex,
                               const double* Auu, const double* Auv,
                               const double* Avu, const double* Avv,
                               const double* bu, const double* bv,
                               double* xu, double* xv,
                               int32_t precond, int32_t ncheck, double tolerance,
                               int32_t maxiters, double* err, int32_t* niters);

int orc_pcg_solver_chrongear_3d(const orc_grid* g, const int32_t* active_vertex,
                                const double* Auu, const double* Auv,
                                const double* Avu, const double* Avv,
                                const double* bu, const double* bv,
                                double* xu, double* xv,
                                int32_t precond, int32_t ncheck, double tolerance,
                                int32_t maxiters, double* err, int32_t* niters);

/* Per-iteration trace for iterate-by-iterate checks: if non-NULL, the solvers
 * store the (rho|eta1, sigma|eta2, alpha) scalars of each iteration here. */
void orc_set_trace(double* buf, int32_t cap);
int32_t orc_trace_len(void);

/* OpenMP variant of the Chronopoulos-Gear solve for the all-cores baseline:
 * identical arithmetic per element, global sums combined in a fixed order
 * over per-j partials.  nthreads<=0 uses the OpenMP default. */
int orc_pcg_solver_chrongear_3d_omp(const orc_grid* g, const int32_t* active_vertex,
                                    const double* Auu, const double* Auv,
                                    const double* Avu, const double* Avv,
                                    const double* bu, const double* bv,
                                    double* xu, double* xv,
                                    int32_t precond, int32_t ncheck, double tolerance,
                                    int32_t maxiters, int32_t nthreads,
                                    double* err, int32_t* niters);

#ifdef __cplusplus
}
#endif
#endif
