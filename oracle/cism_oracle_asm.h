/*
 * cism_oracle_asm.h -- assembly-side entry points of the CPU restatement
 * (oracle_assembly.c, oracle_picard.c).  TEST INFRASTRUCTURE ONLY.
 * Layouts as documented in cism_oracle.h.
 */
#ifndef CISM_ORACLE_ASM_H
#define CISM_ORACLE_ASM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

void orc_fe_tables(double* phi, double* dxr, double* dyr, double* dzr);
void orc_calclsrf(int nx, int ny, const double* thck, const double* topg, double eus, double* lsrf);
void orc_get_masks(int nx, int ny, const double* thck, const double* topg, double eus, double thklim,
                   int32_t* ice_mask, int32_t* floating_mask, int32_t* ocean_mask, int32_t* land_mask);
void orc_stagger(int nx, int ny, const double* var, double* stagvar);
void orc_surface_elevation_gradient(int nx, int ny, double dx, double dy,
                                    const int32_t* ice_mask, const int32_t* land_mask,
                                    const double* usrf, const double* thck,
                                    double thklim, double thck_gradient_ramp, int margin,
                                    double max_slope, double* ds_dx, double* ds_dy);
void orc_slope_angle(int nx, int ny, double dx, double dy, const double* zsfc, double* theta);
void orc_vertex_masks(int nx, int ny, int nhalo, const int32_t* ice_mask,
                      int32_t* active_cell, int32_t* active_vertex);
void orc_flwafact(int nx, int ny, int nz, int nhalo, double n_glen, const double* flwa, double* flwafact);
void orc_load_vector_gravity_local(int nx, int ny, int nz, const double* stagwbndsigma,
                                   double dx, double dy, const int32_t* active_vertex,
                                   const double* stagthck, const double* dusrf_dx, const double* dusrf_dy,
                                   double* loadu, double* loadv);
int orc_element_matrix(const double* x, const double* y, const double* z, const double* u, const double* v,
                       double flwafact, int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                       double Kuu[8][8], double Kuv[8][8], double Kvu[8][8], double Kvv[8][8], double* efvs_mean);
int orc_assemble_stiffness_matrix_3d(int nx, int ny, int nz, const double* sigma, int nhalo,
                                     const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                     const double* uvel, const double* vvel,
                                     const double* stagusrf, const double* stagthck, const double* flwafact,
                                     int whichefvs, double efvs_constant, double effstrain_min, double n_glen,
                                     double* efvs, double* Auu, double* Auv, double* Avu, double* Avv);
void orc_basal_sliding_bc_local(int nx, int ny, int nz, double dx, double dy,
                                const int32_t* active_vertex, const double* beta,
                                const double* theta_basal_slope, double* Auu, double* Avv);
void orc_dirichlet_bc_3d(int nx, int ny, int nz, int nhalo, const int32_t* active_vertex,
                         const int32_t* umask, const int32_t* vmask, const double* uvel, const double* vvel,
                         double* Auu, double* Auv, double* Avu, double* Avv, double* bu, double* bv);
int64_t orc_check_symmetry_assembled_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi,
                                        const int32_t* active_vertex,
                                        double* Auu, double* Auv, double* Avu, double* Avv);
int64_t orc_count_nonzeros_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi,
                              const int32_t* active_vertex, const double* Auu, const double* Auv,
                              const double* Avu, const double* Avv);
void orc_stagger_masked(int nx, int ny, const double* var, const int32_t* mask, double* stagvar);
void orc_load_vector_lateral_bc(int nx, int ny, int nz, const double* sigma, int nhalo, int local,
                                const int32_t* ocean_mask, const int32_t* active_cell,
                                const double* xVertex, const double* yVertex, const double* usrf, const double* thck,
                                const double* stagusrf, const double* stagthck, double* loadu, double* loadv);
void orc_grounded_fraction_no_glp(int nx, int ny, const int32_t* ice_mask, const int32_t* floating_mask,
                                  const int32_t* land_mask, double* f_ground);
double orc_residual_velocity_3d(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi, int whichresid,
                                const double* usav, const double* vsav, const double* uvel, const double* vvel);
void orc_vertical_average(int nx, int ny, int nz, const double* sigma, const double* var, double* var_2d);
int orc_element_matrix_2d_ssa(const double* x, const double* y, const double* u, const double* v, const double* h,
                              double flwafact_2d, int whichefvs, double efvs_constant, double effstrain_min,
                              double n_glen, double Kuu[4][4], double Kuv[4][4], double Kvu[4][4], double Kvv[4][4],
                              double* efvs_mean);
int orc_assemble_stiffness_matrix_2d_ssa(int nx, int ny, int nz, const double* sigma, int nhalo,
                                         const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                         const double* uvel_2d, const double* vvel_2d, const double* stagthck,
                                         const double* flwafact, int whichefvs, double efvs_constant,
                                         double effstrain_min, double n_glen, double* efvs,
                                         double* Auu, double* Auv, double* Avu, double* Avv);
void orc_basal_sliding_bc_2d_local(int nx, int ny, double dx, double dy, const int32_t* active_vertex,
                                   const double* beta, double* Auu, double* Avv);
void orc_dirichlet_bc_2d(int nx, int ny, int nhalo, const int32_t* active_vertex, const int32_t* umask,
                         const int32_t* vmask, const double* uvel, const double* vvel,
                         double* Auu, double* Auv, double* Avu, double* Avv, double* bu, double* bv);
int64_t orc_check_symmetry_assembled_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi,
                                        const int32_t* active_vertex, const double* Auu, const double* Auv,
                                        const double* Avu, const double* Avv);
int64_t orc_count_nonzeros_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi, const int32_t* active_vertex,
                              const double* Auu, const double* Auv, const double* Avu, const double* Avv);
double orc_residual_velocity_2d(int nx, int ny, int ilo, int ihi, int jlo, int jhi, int whichresid,
                                const double* usav, const double* vsav, const double* uvel, const double* vvel);
void orc_slope_angle_xy(int nx, int ny, double dx, double dy, const double* zsfc, double* theta, double* thx,
                        double* thy);
void orc_slope_angle_staggered(int nx, int ny, double dx, double dy, const double* zsfc, double* theta,
                               double* thx, double* thy);
void orc_cell_halo(int nx, int ny, int global_bc, double* a);
int orc_assemble_stiffness_matrix_2d_diva(int nx, int ny, int nz, const double* sigma, const double* stagsigma,
                                          int nhalo, const int32_t* active_cell, const double* xVertex,
                                          const double* yVertex, const double* uvel_2d, const double* vvel_2d,
                                          const double* stagthck, const double* thck, const double* flwafact,
                                          const double* btractx, const double* btracty, const double* sfx_cell,
                                          const double* sfy_cell, int whichefvs, double efvs_constant,
                                          double effstrain_min, double n_glen, double* efvs_qp_3d, double* efvs,
                                          double* omega_k, double* omega, double* Auu, double* Auv, double* Avu,
                                          double* Avv);
void orc_diva_beta_eff(int nx, int ny, int no_slip, const double* stag_omega, const double* beta,
                       const double* ssfx, const double* ssfy, const double* sthx, const double* sthy,
                       double* beta_eff_x, double* beta_eff_y);
void orc_basal_sliding_bc_2d_diva_local(int nx, int ny, double dx, double dy, const int32_t* active_vertex,
                                        const double* beta_eff_x, const double* beta_eff_y, const double* tbx,
                                        const double* tby, double* Auu, double* Avv);
void orc_compute_3d_velocity_diva(int nx, int ny, int nz, const int32_t* active_vertex, const int32_t* ipl_mask,
                                  const double* omega_k, const double* beta, const double* beta_eff_x,
                                  const double* beta_eff_y, const double* sthx, const double* sthy,
                                  const double* ssfx, const double* ssfy, const double* uvel_2d,
                                  const double* vvel_2d, double* btractx, double* btracty, double* uvel,
                                  double* vvel);
double orc_l1l2_tau_parallel(double effstrain, double flwa, double tau_perp);
int orc_assemble_stiffness_matrix_2d_l1l2(int nx, int ny, int nz, const double* sigma, int nhalo,
                                          const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                          const double* uvel_2d, const double* vvel_2d, const double* stagthck,
                                          const double* dusrf_dx, const double* dusrf_dy, const double* flwa,
                                          const double* flwafact, int whichefvs, double efvs_constant,
                                          double effstrain_min, double n_glen, double* efvs, double* Auu,
                                          double* Auv, double* Avu, double* Avv);
int orc_compute_3d_velocity_l1l2(int nx, int ny, int nz, int ilo, int ihi, int jlo, int jhi, int global_bc,
                                 int nhalo, const double* sigma, double dx, double dy, const int32_t* ice_mask,
                                 const int32_t* active_cell, const int32_t* active_vertex, const int32_t* umask,
                                 const int32_t* vmask, const double* xVertex, const double* yVertex,
                                 const double* thck, const double* stagthck, const double* dusrf_dx,
                                 const double* dusrf_dy, const double* flwa, const double* efvs, int whichefvs,
                                 double n_glen, double* uvel, double* vvel);
int orc_load_vector_gravity_standard(int nx, int ny, int nz, int nhalo, const double* sigma,
                                     const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                     const double* stagusrf, const double* stagthck, const double* dusrf_dx,
                                     const double* dusrf_dy, double* loadu, double* loadv);
int orc_basal_sliding_bc_standard(int nx, int ny, int nz, int nhalo, int two_d, int curved,
                                  const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                  const double* staglsrf, const double* beta, const double* beta_y, double* Auu,
                                  double* Avv);

/* glissade_calcbeta (oracle_basal.c): scalar parameters p[ORC_CB_*] */
enum {
    ORC_CB_BETA_CONST, ORC_CB_BETA_SMALL, ORC_CB_BETA_LARGE, ORC_CB_PP_Q, ORC_CB_PP_U0, ORC_CB_PHIMIN,
    ORC_CB_PHIMAX, ORC_CB_BEDMIN, ORC_CB_BEDMAX, ORC_CB_EUS, ORC_CB_M, ORC_CB_ZI_UT, ORC_CB_ZI_NMAX,
    ORC_CB_BUMP_SLOPE, ORC_CB_BUMP_WAVELENGTH, ORC_CB_POWERLAW_K, ORC_CB_BETA_GMIN, ORC_CB_UMAX, ORC_CB_DEW,
    ORC_CB_DNS, ORC_CB_NPARAM
};
int orc_calcbeta(int nx, int ny, int global_bc, int whichbabc, int which_beta_limit, const double* p,
                 const double* uvel, const double* vvel, const double* effecpress_stag, const double* coulomb_c,
                 const double* powerlaw_c, const double* mintauf, const double* c_space_factor_stag,
                 const int32_t* bpmp_mask, const double* beta_external, const double* f_ground, const double* topg,
                 const double* thck, const double* flwa_basal, double* beta);

/* glissade_grounded_fraction (oracle_basal.c) */
int orc_compute_grounded_fraction(const double* f_flotation4, double* f_ground);
void orc_flotation_function(int nx, int ny, int global_bc, int which, const double* thck, const double* topg,
                            double eus, const int32_t* ice_mask, const int32_t* land_mask, double* f_flotation);
int orc_grounded_fraction(int nx, int ny, int global_bc, int which_ground, int which_flotation, int which_no_glp,
                          const double* thck, const double* topg, double eus, const int32_t* ice_mask,
                          const int32_t* floating_mask, const int32_t* land_mask, double* f_flotation,
                          double* f_ground);

int orc_compute_internal_stress(int nx, int ny, int nz, const double* sigma, int nhalo, const int32_t* active_cell,
                                const double* xVertex, const double* yVertex, const double* stagusrf,
                                const double* stagthck, const double* flwafact, int whichefvs, double efvs_constant,
                                double effstrain_min, double n_glen, const double* uvel, const double* vvel,
                                double* tau);
int orc_compute_internal_stress_approx(int nx, int ny, int nz, const double* sigma, int nhalo,
                                       const int32_t* active_cell, const double* xVertex, const double* yVertex,
                                       const double* stagusrf, const double* stagthck, int whichapprox,
                                       const double* flwafact, int whichefvs, const double* efvs,
                                       double efvs_constant, double effstrain_min, double n_glen, const double* uvel,
                                       const double* vvel, double* tau);

void orc_basal_friction(int nx, int ny, int nhalo, int global_bc, int which, const int32_t* active_cell,
                        const int32_t* active_vertex, const double* ubas, const double* vbas, const double* beta,
                        double* bfricflx, double* btractx, double* btracty);

#ifdef __cplusplus
}
#endif
#endif
