/*
 * cismgpu.h -- C ABI of libcismgpu, the MI355X (gfx950) implementation of
 * CISM Glissade's higher-order (Blatter-Pattyn) velocity hot path.
 *
 * Plain C types only.  Host arrays use the reference's Fortran layouts:
 *   vectors   (nz, nx-1, ny-1)        k fastest
 *   matrices  (27, nz, nx-1, ny-1)    neighbour index m fastest,
 *             m = 1 + (iA+1) + 3(jA+1) + 9(kA+1)
 *   vertex masks (nx-1, ny-1) as int32 (0/1) in place of Fortran LOGICAL
 * so that a Fortran caller passes its arrays with c_loc and no copies.
 *
 * Every entry point returns a status: 0 ok, otherwise one of CISMGPU_ERR_*.
 * cismgpu_last_error() gives a message for the last failure of this thread.
 */
#ifndef CISMGPU_H
#define CISMGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CISMGPU_OK                0
#define CISMGPU_ERR_NAN           1   /* alpha or eta1 is NaN: the reference's GM_FATAL (pcg.F90:1331,1391,3263,3434) */
#define CISMGPU_ERR_DIVZERO       2   /* zero pivot in tridiagonal preconditioner setup (pcg.F90:4338-4342, 4474-4493) */
#define CISMGPU_ERR_ARG           3   /* bad argument (sizes, bounds, options) */
#define CISMGPU_ERR_SINGULAR      4   /* singular element Jacobian (glissade_velo_higher.F90:6611-6626) */
#define CISMGPU_ERR_ASYM          5   /* element matrix not symmetric (glissade_velo_higher.F90:10583-10644) */
#define CISMGPU_ERR_BETA          6   /* negative or NaN basal traction (glissade_basal_traction.F90:687-699) */
#define CISMGPU_ERR_HIP          10   /* HIP runtime failure (message in cismgpu_last_error) */
#define CISMGPU_ERR_UNSUPPORTED  11   /* e.g. more than one task per row/column */
#define CISMGPU_ERR_NOT_CONVERGED 12  /* mixed-precision refinement stalled before its tolerance */

/* which_ho_sparse values (glide_types.F90:327-332) */
#define CISMGPU_ALGO_STANDARD  2
#define CISMGPU_ALGO_CHRONGEAR 3

/* precisions of the Krylov iterations */
#define CISMGPU_FP64        0   /* plain FP64: the drop-in */
#define CISMGPU_MIXED_FP32  1   /* FP32 matrix + FP32 inner C-G inside FP64 iterative refinement */
#define CISMGPU_MIXED_BF16  2   /* BF16 matrix storage, FP32 vectors/accumulation, FP64 refinement */

/* Fields of CISM's parallel_type read on the solver path
 * (glissade_velo_higher_pcg.F90:2888-2894; parallel_mpi.F90:2633-2653),
 * for a single task.  Bounds are 1-based like parallel%staggered_*. */
typedef struct cismgpu_grid {
    int32_t nx, ny, nz;          /* cells incl. 2-cell halos; sigma levels */
    int32_t ilo, ihi, jlo, jhi;  /* staggered_ilo/ihi/jlo/jhi */
    int32_t global_bc;           /* 0 periodic, 1 outflow, 2 no_penetration, 3 no_ice */
    int32_t reprosum;            /* parallel%reprosum.  0: global sums from per-workgroup compensated
                                    (double-double) partials in a fixed order; 1: the FP64 solves'
                                    global sums (and the residual's) are exact integer-vector sums
                                    of the packed summands a1*b1 + a2*b2, correctly rounded and
                                    independent of summation order (cism_reprosum_calc) */
} cismgpu_grid;

const char* cismgpu_last_error(void);
int32_t cismgpu_device_count(void);

/* ------------------------------------------------------------------ */
/* Drop-in solver entry points (host arrays in, host solution out).    */
/* ------------------------------------------------------------------ */

/* Replaces pcg_solver_standard_3d (glissade_velo_higher_pcg.F90:827-838);
 * called from glissade_velo_higher.F90:3304-3316.  xu, xv: initial guess in,
 * solution out.  err is written only when a convergence check ran (the
 * reference leaves it unassigned otherwise); niters = maxiters when the
 * tolerance is not reached (non-convergence is not an error). */
int32_t cismgpu_pcg_solver_standard_3d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                       const double* Auu, const double* Auv,
                                       const double* Avu, const double* Avv,
                                       const double* bu, const double* bv,
                                       double* xu, double* xv,
                                       int32_t precond, int32_t linear_solve_ncheck,
                                       double tolerance, int32_t maxiters,
                                       double* err, int32_t* niters);

/* Replaces pcg_solver_chrongear_3d (glissade_velo_higher_pcg.F90:2632-2643);
 * called from glissade_velo_higher.F90:3288-3300.  Same contract as above. */
int32_t cismgpu_pcg_solver_chrongear_3d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                        const double* Auu, const double* Auv,
                                        const double* Avu, const double* Avv,
                                        const double* bu, const double* bv,
                                        double* xu, double* xv,
                                        int32_t precond, int32_t linear_solve_ncheck,
                                        double tolerance, int32_t maxiters,
                                        double* err, int32_t* niters);

/* Replaces matvec_multiply_structured_3d (glissade_velo_higher_pcg.F90:5528-5652):
 * y = A x on owned active vertices, 0 elsewhere. */
int32_t cismgpu_matvec_multiply_structured_3d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                              const double* Auu, const double* Auv,
                                              const double* Avu, const double* Avv,
                                              const double* xu, const double* xv,
                                              double* yu, double* yv);

/* The host-array drop-ins above and below keep one device system between
 * calls, for the last (device, grid, 2-D) they were called with, so a
 * Picard loop allocates device memory once; every call still uploads the
 * matrix, rhs and guess.  No reference counterpart (the reference's
 * solvers hold no state): frees that system, for callers that want the
 * memory back. */
int32_t cismgpu_dropin_release(void);

/* 2-D (depth-integrated: SSA, L1L2, DIVA) solves.  Matrices are the
 * reference's (nx-1, ny-1, 9) stencils (m last, m = 1 + (iA+1) + 3(jA+1),
 * indxA_2d), vectors (nx-1, ny-1); grid->nz is ignored.  precond 0 none,
 * 1 diag, 3/4 tridiagonal line solves on the stencil as given; 2 (SIA) is
 * taken as 1, the replacement glide_setup.F90:1363-1366 makes for 2-D
 * solves. */

/* Replaces pcg_solver_standard_2d (glissade_velo_higher_pcg.F90:67-823);
 * called from glissade_velo_higher.F90:3260-3271.  Same contract as the
 * 3-D entry points. */
int32_t cismgpu_pcg_solver_standard_2d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                       const double* Auu, const double* Auv,
                                       const double* Avu, const double* Avv,
                                       const double* bu, const double* bv,
                                       double* xu, double* xv,
                                       int32_t precond, int32_t linear_solve_ncheck,
                                       double tolerance, int32_t maxiters,
                                       double* err, int32_t* niters);

/* Replaces pcg_solver_chrongear_2d (glissade_velo_higher_pcg.F90:1541-2628);
 * called from glissade_velo_higher.F90:3245-3256. */
int32_t cismgpu_pcg_solver_chrongear_2d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                        const double* Auu, const double* Auv,
                                        const double* Avu, const double* Avv,
                                        const double* bu, const double* bv,
                                        double* xu, double* xv,
                                        int32_t precond, int32_t linear_solve_ncheck,
                                        double tolerance, int32_t maxiters,
                                        double* err, int32_t* niters);

/* Replaces matvec_multiply_structured_2d (glissade_velo_higher_pcg.F90:5656-5773). */
int32_t cismgpu_matvec_multiply_structured_2d(const cismgpu_grid* grid, const int32_t* active_vertex,
                                              const double* Auu, const double* Auv,
                                              const double* Avu, const double* Avv,
                                              const double* xu, const double* xv,
                                              double* yu, double* yv);

/* ------------------------------------------------------------------ */
/* Device-resident system: the matrix, rhs and iterate stay in HBM     */
/* across calls (Picard iterations, benchmarks).  No reference         */
/* counterpart: the reference keeps everything in host arrays.         */
/* ------------------------------------------------------------------ */
typedef struct cismgpu_system cismgpu_system;

typedef struct cismgpu_solve_opts {
    int32_t algo;            /* CISMGPU_ALGO_STANDARD or CISMGPU_ALGO_CHRONGEAR */
    int32_t precond;         /* 0 none, 1 diag, 2 SIA column, 3 tridiag local, 4 tridiag global (one task) */
    int32_t ncheck;          /* linear_solve_ncheck */
    int32_t maxiters;        /* linear_maxiters (mixed mode: inner iterations summed over all corrections) */
    double  tolerance;       /* relative residual ||b-Ax||/||b|| target */
    int32_t precision;       /* CISMGPU_FP64 / CISMGPU_MIXED_FP32 / CISMGPU_MIXED_BF16 */
    double  inner_tol;       /* mixed: inner relative residual target per refinement step */
    int32_t max_outer;       /* mixed: maximum refinement steps */
    int32_t time_spmv;       /* 1: time every SpMV launch with HIP events (info.spmv_*) */
} cismgpu_solve_opts;

typedef struct cismgpu_solve_info {
    double  err;             /* final relative residual (FP64) */
    int32_t niters;          /* FP64: iterations; mixed: total inner iterations */
    int32_t outer_iters;     /* mixed: refinement steps (0 for FP64) */
    double  solve_ms;        /* device time of the solve (HIP events on the system stream) */
    double  spmv_ms;         /* summed duration of the timed SpMV launches */
    int32_t spmv_count;      /* number of timed SpMV launches */
    int32_t spmv_kind;       /* storage of the dominant SpMV: 0 fp64, 1 fp32, 2 bf16, 3 fp32 half (symmetric), 4 fp64 half */
    double  setup_ms;        /* host wall time of the per-assembly set-up this call did before solve_ms began
                                (symmetry checks, half-storage and low-precision copies, line-solve
                                factors); 0 when a previous solve of the same matrix did it */
} cismgpu_solve_info;

int32_t cismgpu_system_create(const cismgpu_grid* grid, cismgpu_system** out);
/* One-level system for the 2-D solves.  Its matrix is loaded with
 * cismgpu_system_set_matrix_2d or assembled by cismgpu_system_assemble_2d
 * (grid->nz: the sigma levels of the assembly inputs; the solves ignore it);
 * vectors are (nx-1, ny-1); the SpMV reads the 9-point stencil only.
 * Ensembles need 3-D systems. */
int32_t cismgpu_system_create_2d(const cismgpu_grid* grid, cismgpu_system** out);
int32_t cismgpu_system_destroy(cismgpu_system* sys);

/* Load A (reference layout) and b.  where = 0: host pointers; 1: device pointers. */
int32_t cismgpu_system_set_matrix(cismgpu_system* sys, const int32_t* active_vertex,
                                  const double* Auu, const double* Auv,
                                  const double* Avu, const double* Avv,
                                  const double* bu, const double* bv, int32_t where);
/* Load a 2-D stencil A(nx-1, ny-1, 9) and b(nx-1, ny-1) into a system from
 * cismgpu_system_create_2d.  where as above. */
int32_t cismgpu_system_set_matrix_2d(cismgpu_system* sys, const int32_t* active_vertex,
                                     const double* Auu, const double* Auv,
                                     const double* Avu, const double* Avv,
                                     const double* bu, const double* bv, int32_t where);
/* Copy A and b back to reference layout (host pointers); for a 2-D system
 * the rows are (27, 1, nx-1, ny-1) with the stencil in m = 10..18.  Rows that are not
 * owned are returned as zero. */
int32_t cismgpu_system_get_matrix(cismgpu_system* sys, double* Auu, double* Auv,
                                  double* Avu, double* Avv, double* bu, double* bv);
int32_t cismgpu_system_set_x(cismgpu_system* sys, const double* xu, const double* xv);
int32_t cismgpu_system_get_x(cismgpu_system* sys, double* xu, double* xv);

/* Solve with the device-resident x as initial guess; x is overwritten. */
int32_t cismgpu_system_solve(cismgpu_system* sys, const cismgpu_solve_opts* opts,
                             cismgpu_solve_info* info);

/* y = A x with the resident matrix (FP64) on the resident x; host result. */
int32_t cismgpu_system_matvec(cismgpu_system* sys, double* yu, double* yv);

/* Average device time of `nrep` back-to-back SpMV launches with storage
 * `kind` (0 fp64, 1 fp32, 2 bf16), measured with HIP events on the system
 * stream.  Also returns the algorithmic bytes of one launch. */
int32_t cismgpu_system_bench_spmv(cismgpu_system* sys, int32_t kind, int32_t nrep,
                                  double* avg_ms, double* bytes_per_launch);


/* Replaces cism_reprosum_calc (cism_reprosum_mod.F90:237-1057) for one
 * task: arr_gsum[f] = exact sum of arr(1:nsummands, f), f = 1..nflds
 * (Fortran layout, host arrays), rounded once to the nearest double.  The
 * summands are split into 30-bit integer digits on a fixed ladder covering
 * every finite double and added as 64-bit integers, so the result does not
 * depend on their order; NaN if a summand is not finite.  At most 2^32
 * summands per field. */
int32_t cismgpu_reprosum_calc(const double* arr, int64_t nsummands, int32_t nflds, double* arr_gsum);

/* 1 when the half (symmetric) storage layouts can hold this grid's owned
 * rows, 0 when the system keeps full rows whatever the symmetry (one level
 * of half-storage slots, nyo * nxp * 56 elements, must stay below 2^31).
 * Host-only query, no device work; no reference counterpart. */
int32_t cismgpu_half_storage_supported(const cismgpu_grid* grid);

/* HBM bandwidth probes (no reference counterpart; measurement only):
 * STREAM-style copy (2 x nbytes moved), triad (3 x nbytes) and read
 * (1 x nbytes) over arrays of nbytes each, average of nrep launches,
 * in GB/s.  Any output pointer may be NULL. */
int32_t cismgpu_bench_stream(int64_t nbytes, int32_t nrep, double* copy_gbps, double* triad_gbps,
                             double* read_gbps);

/* x = 0 on the device (initial guess for benchmarks) */
int32_t cismgpu_system_zero_x(cismgpu_system* sys);

/* staggered_parallel_halo of the resident x (glissade_velo_higher.F90:3598-3600) */
int32_t cismgpu_system_halo_x(cismgpu_system* sys);

/* ------------------------------------------------------------------ */
/* Device assembly of the linear BP system (3-D branch of              */
/* glissade_velo_higher_solve, glissade_velo_higher.F90:2815-2931):    */
/* assemble_stiffness_matrix_3d, basal_sliding_bc_3d (local), b = load  */
/* (load_vector_gravity, local), dirichlet_boundary_conditions_3d, the  */
/* halo of A and b, check_symmetry_assembled_matrix_3d.  The result is  */
/* left in the system (owned rows) for cismgpu_system_solve.  The       */
/* velocity that sets the viscosity is the system's resident x          */
/* (cismgpu_system_set_x).  Host arrays, reference layouts.             */
/* ------------------------------------------------------------------ */
typedef struct cismgpu_assembly_in {
    double dx, dy;
    const double*  sigma;            /* (nz) */
    const double*  stagwbndsigma;    /* (0:nz) */
    const int32_t* active_cell;      /* (nx, ny) */
    const int32_t* active_vertex;    /* (nx-1, ny-1) */
    const double*  xVertex;          /* (nx-1, ny-1) */
    const double*  yVertex;
    const double*  stagusrf;         /* (nx-1, ny-1) */
    const double*  stagthck;
    const double*  dusrf_dx;         /* (nx-1, ny-1), halo already applied */
    const double*  dusrf_dy;
    const double*  flwafact;         /* (nz-1, nx, ny): 0.5 * A^(-1/n) */
    const double*  beta_internal;    /* (nx-1, ny-1); NULL: no basal friction term (no-slip) */
    const double*  theta_basal;      /* (nx-1, ny-1) basal slope angle; NULL: flat */
    const int32_t* umask_dirichlet;  /* (nz, nx-1, ny-1); NULL: none */
    const int32_t* vmask_dirichlet;
    int32_t whichefvs;               /* 0 constant, 1 flow factor, 2 nonlinear (glide_types.F90:240-242) */
    double  efvs_constant;
    double  effstrain_min;
    double  n_glen;
    int32_t check_symmetry;          /* 1: check_symmetry_assembled_matrix_3d */
    /* lateral ocean-pressure BC at calving fronts (load_vector_lateral_bc,
     * glissade_velo_higher.F90:4453-4957); ocean_mask NULL: not applied */
    const int32_t* ocean_mask;       /* (nx, ny) */
    const double*  usrf;             /* (nx, ny) upper surface */
    const double*  thck;             /* (nx, ny) */
    const double*  stagusrf_lateral; /* (nx-1, ny-1): usrf, thck averaged over ice-covered or */
    const double*  stagthck_lateral; /*   land cells (glissade_velo_higher.F90:2006-2016)     */
    int32_t which_assemble_lateral;  /* 0 standard (staggered values), 1 local (cell values) */
    const double*  flwa;             /* (nz-1, nx, ny) rate factor A (Pa^-n yr^-1): the L1L2 viscosity */
    const double*  stag_lsrf;        /* (nx-1, ny-1) lower surface on the staggered grid: the curved basal
                                        Jacobian of the standard basal option (3-D) */
    int32_t assemble_beta_standard;  /* 1: which_ho_assemble_beta = 0, quadrature of the basal term
                                        (glissade_velo_higher.F90:8955-9111; 2-D :8386-8527); 0: local (default) */
    int32_t assemble_taud_standard;  /* 1: which_ho_assemble_taud = 0, quadrature of the gravity load
                                        (:4350-4446); 0: local (default) */
    int32_t beta_device;             /* 1: the basal friction field is the one the last
                                        cismgpu_system_calcbeta left on the device (beta_internal is
                                        not read); 0: beta_internal (host) */
} cismgpu_assembly_in;

typedef struct cismgpu_assembly_info {
    int64_t nnz;                     /* count_nonzeros_3d over owned active rows */
    int64_t asym_pairs;              /* entry pairs left asymmetric (the reference warns) */
    double  assemble_ms;             /* device time of the element kernel */
    double  total_ms;                /* device time of the whole assembly */
} cismgpu_assembly_info;

int32_t cismgpu_system_assemble_3d(cismgpu_system* sys, const cismgpu_assembly_in* in,
                                   cismgpu_assembly_info* info);
/* which_ho_approx values (glide_types.F90:335-339) */
#define CISMGPU_APPROX_SSA   1
#define CISMGPU_APPROX_L1L2  3
#define CISMGPU_APPROX_DIVA  4
/* 2-D (depth-integrated) assembly on a system from cismgpu_system_create_2d,
 * whose grid nz gives the sigma levels of the inputs (the same fields as
 * for the 3-D assembly; the Dirichlet masks' basal level is used).  Velocity:
 * the resident 2-D x.  Covers assemble_stiffness_matrix_2d
 * (glissade_velo_higher.F90:4961-5381) for which_ho_approx = SSA, the local
 * basal_sliding_bc_2d (:8281-8531), b = the level sum of the 3-D load
 * (:2045-2058, :2636), dirichlet_boundary_conditions_2d (:9133-9301), the
 * matrix and rhs halo, check_symmetry_assembled_matrix_2d (report only,
 * :10648-10850) and count_nonzeros_2d (:10450-10509).  DIVA (after
 * cismgpu_system_diva_setup, in->thck required) updates the layer
 * viscosities of the state from the basal traction, weights the elements
 * with their vertical mean (:5198-5300, compute_effective_viscosity_diva
 * :7730-7997), forms omega (compute_integrals_diva :5649-5775), beta_eff
 * (:2469-2610) and basal_sliding_bc_2d_diva (:8535-8811), with no Dirichlet
 * rows for no-slip (:1423-1428).  L1L2 (in->flwa required; n = 1 or 3 for
 * the nonlinear viscosity) takes the layer viscosities of
 * compute_effective_viscosity_L1L2 (:7474-7726) at every quadrature point
 * and weights the elements with their vertical mean (:5228-5248). */
int32_t cismgpu_system_assemble_2d(cismgpu_system* sys, const cismgpu_assembly_in* in, int32_t whichapprox,
                                   cismgpu_assembly_info* info);

/* DIVA inputs and state of a 2-D system (glissade_velo_higher.F90:2078-2154
 * slope factors cos^2 + 4 sin^2 of the slope angles of (usrf+lsrf)/2 at cells
 * and vertices; :8591-8605 basal slope angles; :1510-1515 ice-or-land
 * mask).  Cell fields are (nx, ny), vertex fields (nx-1, ny-1).  The state
 * (efvs_qp_3d (nz-1, 4, nx, ny), :1342-1352; btractx/y) stays on the device
 * and is updated by every DIVA assembly and cismgpu_system_diva_velocity;
 * NULL state arrays start at zero. */
typedef struct cismgpu_diva_in {
    const double*  slope_factor_x;        /* (nx, ny) diva_slope_factor_x */
    const double*  slope_factor_y;
    const double*  stag_slope_factor_x;   /* (nx-1, ny-1) */
    const double*  stag_slope_factor_y;
    const double*  stag_theta_slope_x;    /* (nx-1, ny-1) radians */
    const double*  stag_theta_slope_y;
    const double*  theta_basal_x;         /* (nx-1, ny-1) slope of the staggered lower surface */
    const double*  theta_basal_y;
    const int32_t* ice_plus_land_mask;    /* (nx, ny) */
    const double*  efvs_qp_3d;            /* state in, or NULL */
    const double*  btractx;               /* state in, or NULL */
    const double*  btracty;
    int32_t no_slip;                      /* which_ho_babc == no-slip: beta_eff = sf/(stag_omega cos) */
} cismgpu_diva_in;

int32_t cismgpu_system_diva_setup(cismgpu_system* sys, const cismgpu_diva_in* in);
/* compute_3d_velocity_diva (glissade_velo_higher.F90:5779-5900) from the
 * resident 2-D x and the beta_eff of the last DIVA assembly: updates the
 * basal traction state; uvel/vvel (nz, nx-1, ny-1) and btractx/y
 * (nx-1, ny-1) are optional host outputs. */
int32_t cismgpu_system_diva_velocity(cismgpu_system* sys, double* uvel, double* vvel, double* btractx,
                                     double* btracty);
/* compute_3d_velocity_L1L2 (glissade_velo_higher.F90:5904-6315, called at
 * :3698-3724) from the resident 2-D x (the basal velocity) and the layer
 * viscosities, surface gradient, thickness and rate factor of the last L1L2
 * assembly; ice_mask (nx, ny) gives the masked stagger and the one-sided
 * gradients.  uvel/vvel (nz, nx-1, ny-1): the basal level is x everywhere,
 * the levels above are set on owned active vertices and zero elsewhere;
 * the caller applies the halo (staggered_parallel_halo, :3723-3724). */
int32_t cismgpu_system_l1l2_velocity(cismgpu_system* sys, const int32_t* ice_mask, double* uvel, double* vvel);
/* ------------------------------------------------------------------ */
/* glissade_calcbeta (glissade_basal_traction.F90:69-719) on the device: */
/* the basal traction coefficient at every vertex from the basal        */
/* velocity, recomputed in each Picard iteration (called at             */
/* glissade_velo_higher.F90:2316-2332).  The fields of the basal        */
/* physics are uploaded once (calcbeta_setup); each calcbeta call       */
/* evaluates the sliding law, the f_ground weighting with               */
/* which_ho_beta_limit (:650-685), the beta >= 0 check (:687-699) and   */
/* the staggered halo (:703) into a device field that an assembly with  */
/* beta_device = 1 reads.  Units: Pa, m/yr, Pa yr/m, Pa^-n yr^-1.       */
/* ------------------------------------------------------------------ */
typedef struct cismgpu_calcbeta_in {
    int32_t whichbabc;               /* which_ho_babc (glide_types.F90:253-268): 0 constant, 1 bpmp,
                                        2 pseudo-plastic, 3 old pseudo-plastic, 4 large, 5 external,
                                        7 Zoet-Iverson, 9 power law, 10 Coulomb friction, 11 Schoof,
                                        8 ISMIP-HOM C, 12 modified Schoof, 13 Tsai, 14 power law
                                        with N, 15 yield stress; 6 (no-slip) is refused */
    int32_t which_beta_limit;        /* which_ho_beta_limit (:270-271): 0 multiply by f_ground, then
                                        floor at beta_grounded_min; 1 floor, then multiply */
    double ho_beta_const, ho_beta_small, ho_beta_large;                  /* Pa yr/m */
    double pseudo_plastic_q, pseudo_plastic_u0;                         /* -, m/yr */
    double pseudo_plastic_phimin, pseudo_plastic_phimax;                /* degrees */
    double pseudo_plastic_bedmin, pseudo_plastic_bedmax;                /* m */
    double eus;                                                         /* m */
    double powerlaw_m;
    double zoet_iverson_ut, zoet_iverson_nmax;                          /* m/yr, Pa */
    double coulomb_bump_max_slope, coulomb_bump_wavelength;             /* -, m */
    double friction_powerlaw_k;                                         /* m yr^-1 Pa^-2 */
    double beta_grounded_min;                                           /* Pa yr/m */
    double beta_powerlaw_umax;                                          /* m/yr; 0: no speed cap */
    double dew, dns;                                                    /* m: grid spacing, option 8 */
    const double*  f_ground;             /* (nx-1, ny-1) grounded fraction, always read */
    const double*  effecpress_stag;      /* (nx-1, ny-1) effective pressure N: options 2, 3, 7, 10-14 */
    const double*  coulomb_c;            /* (nx-1, ny-1): options 2, 7, 10-13 */
    const double*  powerlaw_c;           /* (nx-1, ny-1): options 9, 11-13 */
    const double*  mintauf;              /* (nx-1, ny-1) yield stress: option 15 */
    const double*  c_space_factor_stag;  /* (nx-1, ny-1) factor of options 10, 11, 13; NULL: 1 */
    const int32_t* bpmp_mask;            /* (nx-1, ny-1): option 1 */
    const double*  beta_external;        /* (nx-1, ny-1) Pa yr/m: option 5 */
    const double*  topg;                 /* (nx, ny): option 3 (read at the cell with the vertex's
                                            indices, as the reference does, :264-266) */
    const double*  thck;                 /* (nx, ny): option 10 (cells with thck > 0) */
    const double*  flwa_basal;           /* (nx, ny) basal-layer rate factor: option 10 */
} cismgpu_calcbeta_in;

int32_t cismgpu_system_calcbeta_setup(cismgpu_system* sys, const cismgpu_calcbeta_in* in);
/* beta from the basal velocity: ubas/vbas (nx-1, ny-1) host arrays, or
 * both NULL for the system's own: the basal level of x on a 3-D system, x
 * itself on a 2-D one, and on a DIVA system the basal level of the last
 * cismgpu_system_diva_velocity since diva_setup (x before the first).
 * beta (nx-1, ny-1, halo applied) is an optional host copy.  A negative or
 * NaN value returns CISMGPU_ERR_BETA naming the vertex. */
int32_t cismgpu_system_calcbeta(cismgpu_system* sys, const double* ubas, const double* vbas, double* beta);

/* Snapshot of the assembled system (A, b, preconditioner) kept on the
 * device, and its restoration: the matrix copies of the accelerated Picard
 * iteration (evaluate_accelerated_picard_3d, glissade_velo_higher.F90:10052-10184). */
int32_t cismgpu_system_save_assembly(cismgpu_system* sys);
int32_t cismgpu_system_restore_assembly(cismgpu_system* sys);
/* Device-resident Picard vectors (no reference counterpart: the reference
 * keeps them in host arrays).  Slot -1 is the resident iterate x; slots
 * 0..nslots-1 are work pairs (u, v) allocated by work_vectors (at most 16).
 * vector_op: op 0 dst = a; 1 dst = a + alpha b; 2 dst = a - b; 3 dst = 0,
 * over whole arrays (halo included), asynchronous on the system stream. */
int32_t cismgpu_system_work_vectors(cismgpu_system* sys, int32_t nslots);
int32_t cismgpu_system_vector_op(cismgpu_system* sys, int32_t op, int32_t dst, int32_t a, int32_t b,
                                 double alpha);
/* compute_residual_velocity_3d (glissade_velo_higher.F90:10314-10446) of the
 * resident x against work slot old_slot: max (which_ho_resid 0, 1, 3, 4) or
 * mean (2) relative speed change over owned vertices; 1 and 2 skip the bed. */
int32_t cismgpu_system_residual_velocity(cismgpu_system* sys, int32_t whichresid, int32_t old_slot,
                                         double* resid);
/* compute_internal_stress (glissade_velo_higher.F90:7046-7281, BP branch)
 * from the resident x and the inputs of the last 3-D assembly: tau_xz,
 * tau_yz, tau_xx, tau_yy, tau_xy (Pa; means over the quadrature points of
 * 2 efvs times the strain rates) and tau_eff, each (nz-1, nx, ny), zero
 * outside the active cells; any output may be NULL. */
int32_t cismgpu_system_internal_stress(cismgpu_system* sys, double* tau_xz, double* tau_yz, double* tau_xx,
                                       double* tau_yy, double* tau_xy, double* tau_eff);
/* compute_internal_stress (glissade_velo_higher.F90:7046-7281, called at
 * :3779) on a 2-D system, for its last assembly's approximation: SSA
 * (horizontal strain rates and their viscosity), L1L2 (every strain rate,
 * the layer viscosity of the last L1L2 assembly), DIVA (the 3-D branch).
 * uvel, vvel (nz, nx-1, ny-1): the 3-D velocity (of
 * cismgpu_system_l1l2_velocity or cismgpu_system_diva_velocity); both NULL
 * on an SSA system for the resident 2-D x at every level.  Outputs as
 * cismgpu_system_internal_stress; any may be NULL. */
int32_t cismgpu_system_internal_stress_2d(cismgpu_system* sys, const double* uvel, const double* vvel,
                                          double* tau_xz, double* tau_yz, double* tau_xx, double* tau_yy,
                                          double* tau_xy, double* tau_eff);
/* compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042, called
 * at :3795) with its cell halo (:3807), and the basal traction
 * btractx/btracty = beta u_b (:3817-3818).  which_bfric: 0 standard (2x2
 * Gauss points of the bilinear interpolants over owned active cells), 1 local
 * (active vertices, unstaggered).  ubas, vbas (nx-1, ny-1): the basal
 * velocity, or both NULL for the system's own (3-D: bed level of x; SSA and
 * L1L2: x; DIVA: the bed level of the last diva_velocity, kept with a
 * friction law).  beta (nx-1, ny-1): NULL for the last assembly's beta (the
 * calcbeta result when it read the device's).  Masks: the last assembly's.
 * Outputs bfricflx (nx, ny) in W/m^2, btractx, btracty (nx-1, ny-1) in Pa;
 * any may be NULL. */
int32_t cismgpu_system_basal_friction(cismgpu_system* sys, int32_t which_bfric, const double* ubas,
                                      const double* vbas, const double* beta, double* bfricflx, double* btractx,
                                      double* btracty);
/* efvs(nz-1, nx, ny) of the last assembly (mean over quadrature points) */
int32_t cismgpu_system_get_efvs(cismgpu_system* sys, double* efvs);
/* resid = A x - b on owned active vertices, L2 norm and norm relative to
 * |b| (compute_residual_vector_3d, glissade_velo_higher.F90:9693-9894). */
int32_t cismgpu_system_residual(cismgpu_system* sys, double* resid_u, double* resid_v,
                                double* L2_norm, double* L2_norm_relative);
/* The same residual without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops); no counterpart in
 * the reference.  Every pointer is a DEVICE pointer to the reference layout:
 * active_vertex (nx-1, ny-1) int32, A (27, nz, nx-1, ny-1), b, u and resid
 * (nz, nx-1, ny-1); resid_u / resid_v may be NULL; norms[0] = L2 norm,
 * norms[1] = L2 / |b| (0 when b = 0) are written on the device.  Work is
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) and not
 * waited for.  Partial sums use a library-held 2 MiB device array per GPU:
 * at most 16,777,216 nodes (nz * (nx-1) * (ny-1)), larger grids are
 * rejected; calls that overlap on one GPU race on that array. */
int32_t cismgpu_residual_vector_3d_noalloc(const cismgpu_grid* grid, const int32_t* active_vertex,
                                           const double* Auu, const double* Auv,
                                           const double* Avu, const double* Avv,
                                           const double* bu, const double* bv,
                                           const double* uvel, const double* vvel,
                                           double* resid_u, double* resid_v, double* norms, void* stream);

/* ------------------------------------------------------------------ */
/* Batched ensemble of independent systems (no reference counterpart:   */
/* the reference solves one system per call).  Members are assembled    */
/* systems; the ensemble copies their FP64 matrices, rhs and SIA         */
/* preconditioners into batched storage and solves all of them with FP64 */
/* iterative refinement around FP32 (or BF16-matrix) Chronopoulos-Gear   */
/* in ONE persistent launch, captured in a hipGraph.  Members must use   */
/* the no_ice or outflow halo rule with an ice-free outer ring.          */
/* ------------------------------------------------------------------ */
typedef struct cismgpu_ensemble cismgpu_ensemble;

typedef struct cismgpu_ensemble_opts {
    double  tolerance;        /* outer relative residual per member */
    double  inner_tol;        /* inner relative residual per refinement step */
    int32_t inner_maxiters;
    int32_t max_outer;
    int32_t ncheck;           /* inner convergence checked every ncheck iterations */
    int32_t bf16;             /* 1: BF16 matrix storage in the inner solve */
    int32_t blocks_per_cu;    /* persistent grid = blocks_per_cu x CUs (0: from occupancy) */
} cismgpu_ensemble_opts;

typedef struct cismgpu_ensemble_info {
    double  solve_ms;         /* device time of the graph launch */
    int32_t steps;            /* lock-step steps executed */
    int32_t converged;        /* members that reached tolerance */
    int32_t max_outer;        /* largest number of refinement steps */
    int32_t max_inner;        /* largest total inner iterations of a member */
    int64_t total_inner;      /* inner iterations summed over members */
    double  max_err;          /* largest final outer relative residual */
    int32_t barrier_error;    /* 1 if a grid barrier timed out (result invalid) */
    int32_t grid_blocks;      /* persistent workgroups launched */
    int32_t matrix_storage;   /* inner products: 1 FP32 quad rows, 2 BF16 quad rows, 3 FP32 half
                                 (symmetric) storage, used when every member is exactly symmetric */
} cismgpu_ensemble_info;

int32_t cismgpu_ensemble_create(int32_t nmembers, cismgpu_system* const* members, cismgpu_ensemble** out);
int32_t cismgpu_ensemble_destroy(cismgpu_ensemble* ens);
/* time the FP32 SpMV per size bucket for each level split and keep the
 * fastest; buckets: owned vertices per side <= 80, <= 136, larger.
 * split_out[3] receives the choice (1, 2 or 4 level groups per tile of
 * 64 owned columns). */
int32_t cismgpu_ensemble_autotune(cismgpu_ensemble* ens, int32_t* split_out, double* ms_out /* [3][3] */);
int32_t cismgpu_ensemble_solve(cismgpu_ensemble* ens, const cismgpu_ensemble_opts* opts,
                               cismgpu_ensemble_info* info);
/* member solution in reference layout; per-member statistics */
int32_t cismgpu_ensemble_get_x(cismgpu_ensemble* ens, int32_t member, double* xu, double* xv);
int32_t cismgpu_ensemble_member_info(cismgpu_ensemble* ens, int32_t member, double* err, int32_t* outer,
                                     int32_t* total_inner, int32_t* flags);

#ifdef __cplusplus
}
#endif
#endif
