/*
 * diablo_oracle.h — CPU FP64 restatement of DIABLO's RKW3 + Crank–Nicolson step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline.  It is never linked into, loaded by, or
 * called from the product (diablo3_amd/, libdiablo_gpu.so).
 *
 * It takes the same dgpu_config as the GPU library and the same Fortran-layout
 * buffers, so a test can feed identical inputs to both.
 */
#ifndef DIABLO_ORACLE_H
#define DIABLO_ORACLE_H
#include "../include/diablo_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

int oracle_create(const dgpu_config *cfg, oracle_t **out);
void oracle_destroy(oracle_t *o);
void oracle_set_threads(int nthreads);
/* bit 0: back substitution multiplies by 1/b instead of dividing (round-off studies) */
void oracle_set_variant(int flags);
const char *oracle_last_error(void);

int oracle_upload_state(oracle_t *o, const double *cu1, const double *cu2,
                        const double *cu3, const double *cp, const double *cth);
int oracle_download_state(oracle_t *o, double *cu1, double *cu2, double *cu3,
                          double *cp, double *cth);
int oracle_upload_rk(oracle_t *o, const double *cf1, const double *cf2,
                     const double *cf3, const double *cfth);
int oracle_download_rk(oracle_t *o, double *cf1, double *cf2, double *cf3,
                       double *cfth);
/* nu_t and kappa_t in the physical Fortran layout (read-only view of the LES state) */
int oracle_download_les(oracle_t *o, double *nu_t, double *kappa_t);

int oracle_rk_substep(oracle_t *o, int rk_step, double delta_t, double time,
                      int time_step, int create_new_flow);
int oracle_rem_div(oracle_t *o);
int oracle_filter_chan(oracle_t *o, int n);
/* variable_dt: courant (for/flow.f90:746-857) runs inside rk_step 3 of every
 * update_dt-th step on the velocity before the projection (for/solver.f90:1241);
 * oracle_courant returns its delta_t (ran = 0 if it did not run since the last call). */
int oracle_set_variable_dt(oracle_t *o, int variable_dt, int update_dt, double cfl);
int oracle_courant(oracle_t *o, double time, double time_limit, double save_stats_time,
                   double save_flow_time, double *delta_t, int *ran);
double oracle_courant_schedule(double dt, double time, double time_limit, double save_stats_time,
                               double save_flow_time);
/* courant's limits (before the caps and scheduling) for the current state */
double oracle_courant_dt_state(oracle_t *o, double cfl);
/* apply_BC_vel_mpi_post + apply_BC_th_mpi_post (for/flow.f90:1161-1360, 1596-1732) */
int oracle_bc_post(oracle_t *o, double time);

int oracle_fft_xz_to_fourier(oracle_t *o, const double *phys, double *four);
int oracle_fft_xz_to_physical(oracle_t *o, const double *four, double *phys);

/* Restated helpers, exposed so tests can pin them one by one. */
void oracle_wavenumbers(int nx, int nz, double lx, double lz, double *kx, double *kz);
void oracle_grid_metrics(int ny, const double *gy_file, double *gy, double *gyf,
                         double *dy, double *dyf);
/* Builds the BC-augmented CN system of one physical column for field `which`
 * (0 = u1, 1 = u2, 2 = u3, 3+n = scalar n) at rk_step, with nu_t = 0;
 * matl/matd/matu have ny+2 entries (rows 0..Nyp+1). */
int oracle_cn_matrix(oracle_t *o, int which, int rk_step, double delta_t,
                     double *matl, double *matd, double *matu, double *vec_bc);
/* The same with an eddy-viscosity column nut(0:Nyp+1) (use_les handles only). */
int oracle_cn_matrix_nut(oracle_t *o, int which, int rk_step, double delta_t, const double *nut,
                         double *matl, double *matd, double *matu, double *vec_bc);
/* The physical CN right-hand side of field `which` left by the last substep
 * (physical Fortran layout), so tests can pin the explicit half on its own. */
int oracle_debug_rhs(oracle_t *o, int which, double *out);
/* Test hooks of the learned closure (les_model_type 6, no counterpart in the
 * reference), on an les_model_type 5 handle.  oracle_les_capture: the next LES
 * substeps write, per GY point of rows 2..Nyp, 16 doubles into cap
 * ([ny2][nz][nx][16], compact physical order): the scaled strain and rotation
 * entries a1..a6, o4..o6 of the type-5 GY coefficient, that coefficient, the
 * scaled scalar gradient (dx b, 2 dy, dz b) x grad th, th_y and the scalar
 * coefficient (slots 0-8, 9, 10-12, 13, 14).  oracle_les_override: the GY
 * velocity and scalar coefficients cgy, kgy ([ny2][nz][nx]) replace the type-5
 * ones as les_model_type 6 uses them: |c| at GY; at GYF the mean of the two GY
 * neighbours when neither has its sign bit set, else the type-5 GYF value.
 * NULL switches a hook off; the arrays stay the caller's. */
int oracle_les_capture(oracle_t *o, double *cap);
int oracle_les_override(oracle_t *o, const double *cgy, const double *kgy);
/* Thomas solves exactly as thomas_forward/backward_*_mpi at one rank. */
void oracle_thomas_real(int ny, double *a, double *b, const double *c, double *g);
void oracle_thomas_complex(int ny, double *a, double *b, const double *c, double *g);

#ifdef __cplusplus
}
#endif
#endif
