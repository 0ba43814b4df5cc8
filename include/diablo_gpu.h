/*
 * diablo_gpu.h — C ABI of the MI355X (gfx950) RKW3 + Crank–Nicolson channel time step.
 *
 * The library is a drop-in for DIABLO's time-advance call
 *     do rk_step = 1,3 : call rk_chan_1            (for/diablo.f90:68-71)
 * The reference routine takes no arguments and talks through module globals
 * (for/flow.f90:12-50, for/parameters.f90:8-65, for/domain.f90:35-85).  Here the
 * globals that rk_chan_1 reads become one dgpu_config passed at creation, the
 * state arrays cross the boundary through upload/download calls in the exact
 * Fortran declared layout, and every call returns an int status (0 = ok) where the
 * reference would `stop` (e.g. for/les.f90:1198, for/flow.f90:1198).
 *
 * Conventions shared by every entry point:
 *  - Code axes: x (periodic, Fourier), y (wall-normal, finite differences on the
 *    staggered stretched grid), z (periodic, Fourier).  Ny is the wall-normal count.
 *  - Fourier arrays cross the boundary as the reference declares them:
 *        complex(8) cu(0:Nxp, 0:Nz+1, 0:Nyp+1)      (for/fft.f90:71-96)
 *    column-major, leading dimensions Nxp+1 and Nz+2, passed as interleaved
 *    (re, im) doubles.  Only the active block i = 0..Nkx, k = 0..2Nkz is read on
 *    upload; on download the active block is written and the de-aliased columns
 *    i = Nkx+1..Nxp of rows k = 0..2Nkz are written as zeros; every other entry
 *    (the physical-alias padding) is left untouched.
 *  - Physical arrays cross as  real(8) u(0:Nx+1, 0:Nzp+1, 0:Nyp+1)  (same buffer size).
 *  - The library owns all device memory (allocated in dgpu_create, freed in
 *    dgpu_destroy) and keeps no host pointer past a call.  Work runs on one HIP
 *    stream; every call is synchronous on return.  A handle is not thread-safe.
 *  - Scalars live in arrays indexed n = 0..n_th-1 for the reference's 1..N_th.
 */
#ifndef DIABLO_GPU_H
#define DIABLO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGPU_MAX_TH 4

/* Precision of the step.  FP64: every array and operation in double.
 * MIXED: transforms, nonlinear products and physical-space intermediates in
 * float with float accumulation; state, RK memory, Fourier-space assembly,
 * tridiagonal CN solves, the pressure Poisson solve (including the kx=kz=0 mean
 * mode) and the thermal-wind terms stay double. */
enum { DGPU_FP64 = 0, DGPU_MIXED = 1 };

/* Boundary-condition codes, as read from input_chan.dat (for/parameters.f90:232-248). */
enum { DGPU_BC_DIRICHLET = 0, DGPU_BC_NEUMANN = 1 };

/* Everything rk_chan_1 reads from modules parameters/domain, flattened.
 * Field names follow the reference's variable names (code convention: u = u1,
 * v = u2 wall-normal, w = u3).  Pointer members are read during dgpu_create only. */
typedef struct dgpu_config {
  /* grid_def (for/domain.f90:18-32): one rank, so Nyp = Ny, Nxp = Nx/2, Nzp = Nz */
  int32_t nx, ny, nz, n_th;
  int32_t precision;        /* DGPU_FP64 or DGPU_MIXED */
  int32_t use_les;          /* input.dat USE_LES (for/parameters.f90:98) */
  int32_t les_model_type;   /* input_chan.dat LES_MODEL_TYPE (for/parameters.f90:217) */
  int32_t flavor_front;     /* flavor == 'Front' enables thermal-wind terms (for/solver.f90:47) */
  int32_t f_type;           /* forcing type (for/solver.f90:464-476, 2121) */
  /* time-stepping row bounds from init_chan_mpi (for/domain.f90:307-374) */
  int32_t jstart, jend;
  int32_t jstart_th[DGPU_MAX_TH], jend_th[DGPU_MAX_TH];
  /* BC codes: u = u1, v = u2 (wall-normal), w = u3 (for/domain.f90:69-84) */
  int32_t u_bc_ymin, v_bc_ymin, w_bc_ymin;
  int32_t u_bc_ymax, v_bc_ymax, w_bc_ymax;
  int32_t th_bc_ymin[DGPU_MAX_TH], th_bc_ymax[DGPU_MAX_TH];
  /* homogeneousX of init_parameters (for/parameters.f90:282-287): 0 for IC types
   * 5-8 (finite fronts), else 1; selects plane or z-line averages in dgpu_stats */
  int32_t homogeneous_x;
  int32_t reserved_i[3];

  double lx, ly, lz;        /* domain lengths (for/parameters.f90:101) */
  double nu, beta, nu_v_scale;
  double ro_inv, delta;     /* Coriolis f = ro_inv/delta (for/solver.f90:502-509) */
  double grav_x, grav_y, grav_z;
  double px0, omega0, amp_omega0, force_start;
  double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
  double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;  /* w_bc_ymax_c1 after IC compensation */
  double th_bc_ymin_c1[DGPU_MAX_TH], th_bc_ymax_c1[DGPU_MAX_TH];
  double pr[DGPU_MAX_TH], dthdx[DGPU_MAX_TH], dthdz[DGPU_MAX_TH];
  double reserved_d[4];

  /* grid metrics, each (0:Nyp+1) = ny+2 values (for/domain.f90:35-39, 276-300, 538-564) */
  const double *gy, *gyf, *dy, *dyf;
  /* wavenumbers exactly as init_fft computes them (for/fft.f90:209-225):
   * kx(0:Nxp-1) (nx/2 values), kz(0:2Nkz) (2*(nz/3)+1 values) */
  const double *kx, *kz;
} dgpu_config;

typedef struct dgpu_handle dgpu_handle;

/* Build a solver for one grid: validates the config, allocates device memory,
 * zero-initialises the state, the RK memory and nu_t/kappa_t (for/flow.f90:49-50).
 * Replaces the allocation half of init_flow (for/flow.f90:100-107). */
int dgpu_create(const dgpu_config *cfg, dgpu_handle **out);

/* Frees all device memory. Replaces nothing (the reference never frees, for/flow.f90:249-254). */
int dgpu_destroy(dgpu_handle *h);

/* State upload/download in the Fortran Fourier layout (see header comment).
 * cth holds n_th consecutive arrays (the reference's cth(...,1:N_th)).
 * Called by the host before/after save_flow, save_stats, create_flow/read_flow
 * (for/diablo.f90:87-123, for/flow.f90:118-141). */
int dgpu_upload_state(dgpu_handle *h, const double *cu1, const double *cu2,
                      const double *cu3, const double *cp, const double *cth);
int dgpu_download_state(dgpu_handle *h, double *cu1, double *cu2, double *cu3,
                        double *cp, double *cth);

/* RK memory cf1..cf3, cfth (for/flow.f90:12-18), carried between substeps
 * (for/solver.f90:413-429).  Needed only to resume between substeps. */
int dgpu_upload_rk(dgpu_handle *h, const double *cf1, const double *cf2,
                   const double *cf3, const double *cfth);
int dgpu_download_rk(dgpu_handle *h, double *cf1, double *cf2, double *cf3,
                     double *cfth);

/* Eddy viscosity nu_t and diffusivities kappa_t(.., 1:N_th) of the last LES
 * substep, in the physical Fortran layout real(8) (0:Nx+1, 0:Nzp+1, 0:Nyp+1)
 * (for/flow.f90:49-50; GY points).  Read-only view for stats
 * (save_stats_LES_OOL, for/les.f90:1030-1063).  Zero before the first LES substep. */
int dgpu_download_les(dgpu_handle *h, double *nu_t, double *kappa_t);

/* One RKW3 substep.  Replaces `call rk_chan_1` (for/solver.f90:366-1273) with the
 * module values rk_step, delta_t, time, time_step and create_new_flow passed in.
 * With use_les and the spin-up gate passed (not create_new_flow, or time_step >
 * 100; for/solver.f90:615) it runs the LES branch, les_model_type 5 (AMD). */
int dgpu_rk_substep(dgpu_handle *h, int32_t rk_step, double delta_t, double time,
                    int32_t time_step, int32_t create_new_flow);

/* One full time step: the three substeps of for/diablo.f90:68-71 queued on the
 * stream back to back (time is not advanced inside, as in the reference). */
int dgpu_step(dgpu_handle *h, double delta_t, double time, int32_t time_step,
              int32_t create_new_flow);

/* dgpu_step runs as a captured HIP graph of its three substeps (on by default;
 * no counterpart in the reference): captured on the first step, replayed while
 * delta_t and the LES spin-up gate are unchanged.  Steps whose launches depend
 * on the time (f_type 2 and 4 forcings), on variable_dt, or that record per-kernel
 * events run their launches directly.  enable = 0 turns the graph off (and frees
 * it); the results are the same either way. */
int dgpu_set_graph(dgpu_handle *h, int32_t enable);

/* Projection of the device velocity onto divergence-free fields: rem_div_chan
 * (for/solver.f90:1302-1379), called by create_flow_chan (for/flow.f90:487). */
int dgpu_rem_div(dgpu_handle *h);

/* 2D xz transforms of one field over all ny+2 planes, in the Fortran layouts:
 * fft_xz_to_fourier (for/fft.f90:242-307) and fft_xz_to_physical
 * (for/fft.f90:311-374).  Used by the create_th/create_flow hooks and by stats. */
int dgpu_fft_xz_to_fourier(dgpu_handle *h, const double *phys, double *four);
int dgpu_fft_xz_to_physical(dgpu_handle *h, const double *four, double *phys);

/* No counterpart in the reference: dgpu_fft_xz_to_fourier for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same transform, with phys and four in
 * the same Fortran layouts but in DEVICE memory (they may alias, as the
 * reference's u/cu views do); rows k >= 2 Nkz + 1 of four are left untouched.
 * It calls no hipMalloc or hipFree: the compact field, the x-spectra and the
 * packed spectrum live in a library-static device buffer of
 * DGPU_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and grids
 * needing more (8 Nx Nz (Ny+2) + 16 kxp (Nz + 2 Nkz + 1)(Ny+2) bytes, kxp = Nkx + 1
 * rounded up to a multiple of 8; 128x129x128 needs 38.6 MB) are rejected with
 * status 2.  The device current at the call must be the handle's.  Every call
 * on one GPU shares that buffer: calls that overlap on one GPU (two threads,
 * or two handles' streams) race. */
enum { DGPU_NOALLOC_SCRATCH_BYTES = 64 << 20 };
int dgpu_fft_xz_to_fourier_noalloc(dgpu_handle *h, const double *phys, double *four);

/* poisson_p_chan (for/solver.f90:1382-1560): the pressure cp from the velocity
 * (and the scalars' y gradients) of the device state, rows 1..Nyp, by the
 * pressure-Poisson solve of rem_div_chan's operator.  ReadHDF5 calls it for a
 * flow file without P (dgpu_read_flow does). */
int dgpu_poisson_p(dgpu_handle *h);

/* Blocks until all queued work on the handle's stream is done. */
int dgpu_synchronize(dgpu_handle *h);

/* Device-time of the last dgpu_step/dgpu_rk_substep in milliseconds (HIP events). */
int dgpu_last_step_ms(dgpu_handle *h, double *ms);

/* Per-kernel device timing, for measurement only (no counterpart in the
 * reference).  While enabled, every launch of the step's kernels is bracketed by
 * HIP events on the handle's stream.  Kernel ids: 0 = z inverse transform of the
 * state, 1 = the LES x pass (products), 2 = z forward transform of the products,
 * 3 = Fourier assembly + Crank-Nicolson solves, 4 = projection + pressure update;
 * LES substeps add 5 = BC posts, 6 = gradient fields, 7 = AMD model,
 * 8 = physical-column CN solves, 9 = right-hand-side and solution transforms,
 * 10 = the learned closure's coefficients (les_model_type 6, dgpu_closure.h),
 * 11 = the DNS x pass (x inverse transforms, products and x forward transforms
 * in one kernel marching in y).
 * dgpu_timing_read returns the device time and launch count accumulated since the
 * last enable, and the name.  Enabling resets the dgpu_step sampling to every call. */
int dgpu_timing_enable(dgpu_handle *h, int32_t enable);
/* Measurement only (no counterpart in the reference): with timing enabled, record
 * the kernel events on every `every`-th dgpu_step call only (1: every call, the
 * default), so that a timed loop carries the events' own cost on few steps. */
int dgpu_timing_every(dgpu_handle *h, int32_t every);
int dgpu_timing_read(dgpu_handle *h, int32_t kernel, double *total_ms, int64_t *launches,
                     const char **name);

/* variable_dt (input.dat VARIABLE_DT, UPDATE_DT, CFL; for/parameters.f90:53-58):
 * while enabled, rk_step 3 of every update_dt-th time step evaluates courant
 * (for/flow.f90:746-857) on the new velocity before the projection, as
 * for/solver.f90:1241-1244 does.  dgpu_courant then returns the delta_t courant
 * sets (ran = 1), with its event scheduling against time_limit, save_stats_time
 * and save_flow_time, or ran = 0 if courant has not run since the last call.
 * The caller passes the returned delta_t to the following steps (h_bar follows it). */
int dgpu_set_variable_dt(dgpu_handle *h, int32_t variable_dt, int32_t update_dt, double cfl);
int dgpu_courant(dgpu_handle *h, double time, double time_limit, double save_stats_time,
                 double save_flow_time, double *delta_t, int32_t *ran);
/* courant on a physical velocity the caller holds (FP64, the physical Fortran
 * layout u(0:Nx+1, 0:Nzp+1, 0:Nyp+1), u2 at GY points), as ReadHDF5 calls it on
 * the velocity it has just read (for/phdf5.f90:1437-1439): with variable_dt set
 * the CFL reduction runs on the device and the next dgpu_courant returns the
 * delta_t (ran = 1); without variable_dt it does nothing.  dgpu_read_flow
 * (dgpu_io.h) calls it. */
int dgpu_courant_physical(dgpu_handle *h, const double *u1, const double *u2, const double *u3);

/* Energy-budget statistics of save_stats_chan (for/diagnostics.f90:3-669, with
 * compute_averages, compute_TKE_diss, compute_MKE_diss, compute_TKE and
 * compute_TKE_Production, 676-1519), computed on the device in FP64 from the
 * current state, which first gets the BC posts save_stats_chan applies
 * (for/diagnostics.f90:26-27).  out holds dgpu_stats_doubles(cfg) doubles: profiles
 * of Ny+2 entries (index j = 0..Nyp+1 as the reference's arrays; entries it never
 * sets are 0), in this order (code variable names; the mean.h5 dataset names of
 * for/diagnostics.f90:140-480 swap v and w):
 *   0 ume  1 vme  2 wme  3 mke  4 urms  5 vrms  6 wrms  7 uw  8 uv  9 wv
 *   10 uu_dudx  11 vv_dvdy  12 wu_dwdx  13 vu_dvdx  14 uv_dudy  15 wv_dwdy
 *   16 dudy  17 dwdy  18 shear  19 epsilon  20 epsilon_m  21 cp (mean mode)
 *   22 u1y_left  23 omega_x  24 omega_y  25 omega_z (compute_Vorticity, 1525-1727:
 *   rms of the three vorticity components, omega_x less the thermal-wind shear)
 * then per scalar n: thme, thrms, thv, thv_m, dthdy, pe_diss; then one double,
 * u1z_x0 (the y-average of u1y_left, integrate_y_var, for/diagnostics.f90:1964);
 * then Nx/2 doubles, FTx_uu (641-663: per kx mode the y-integral of |u1(kz = 0)|^2
 * over Ly).
 * z* bins and the BPE come from dgpu_stats_zstar, the movie planes from
 * dgpu_stats_movie. */
int dgpu_stats(dgpu_handle *h, double time, double *out);

/* The z* statistics of save_stats_chan for scalar 1 (N_th >= 1), on the device in
 * FP64 from the current state after the same BC posts: compute_BPE
 * (for/diagnostics.f90:1733-1852) and Bin_Ystar_and_Write (1859-1953) with
 * Nbin = Nx bins of th + dTHdX(1) (gx - Lx/2) over rows jstart_th(1)..jend_th(1)
 * (uniform bins when x is homogeneous, else the cos-spaced bins of the finite
 * front).  out holds dgpu_stats_zstar_doubles(cfg) doubles:
 *   [0, Nbin)            th1PDF (normalised, dyf-weighted)
 *   [Nbin, 2 Nbin + 1)   th1bin, the Nbin + 1 bin edges (mean.h5 keeps the first Nbin)
 *   [2 Nbin + 1]         BPE
 *   then Nbin each, the z*-binned means (sum f dyf / sum dyf per bin; NaN in an
 *   empty bin, as the reference's 0/0): shear2_zstar, epsilon_zstar, TKE_zstar,
 *   KE_zstar, then per scalar n: thw_zstar(n), chi_zstar(n).
 * The bin sums use atomic additions, so their last bits vary from call to call.
 * Nx (5 + 2 N_th) must not exceed 8192 (the histogram lives in 64 KiB of LDS). */
int dgpu_stats_zstar(dgpu_handle *h, double time, double *out);
int64_t dgpu_stats_zstar_doubles(const dgpu_config *cfg);
int64_t dgpu_stats_doubles(const dgpu_config *cfg);

/* The movie output of save_stats_chan(movie = .true.) for the state at `time`
 * (the BC posts are applied first, as dgpu_stats does), in FP64.
 * planes (dgpu_stats_movie_sizes: n_planes doubles): the z-mean (x, y) planes of
 *   mean_xz.h5 that reduce_and_write_XYplane writes (for/diagnostics.f90:283-378,
 *   756-767, 1076-1087, 1238-1245, 1338-1346, 1452-1460), 16 planes of
 *   [Nyp][Nx] (rows 1..Nyp): ume, wme (u2 on GYF), vme (u3), thme_1, uu, ww
 *   (u2'u2'), vv (u3'u3'), uw (u1'u2' on GY), wv (u3'u2' on GY), uv (u1'u3'),
 *   epsilon, epsilon_m, thth, thu, thw, chi (scalar 1; zero without scalars).
 *   The group names are the reference's, whose v and w are u3 and u2.
 * slices (n_slices doubles): the movie.h5 slices (408-441, 485-617, 1041-1070,
 *   1566-1727) of 8 + N_th fields: u1, u2 averaged to GYF, u3, nu_t (zero
 *   without LES), epsilon, omega_x, omega_y, omega_z, th_1..th_N.  Per field:
 *   [Nyp][Nx] at z = nzm (the reference's *_xz), [Nz][Nx] at row y = nym (*_xy;
 *   zeros when nym = -1, the reference's rankyMovie = -1), [Nyp][Nz] at x = nxm
 *   (*_yz).  nxm, nym, nzm are NxMovie, NyMovie, NzMovie of init_chan_movie
 *   (for/domain.f90:380-401). */
int dgpu_stats_movie(dgpu_handle *h, double time, int32_t nxm, int32_t nym, int32_t nzm, double *planes,
                     double *slices);
int dgpu_stats_movie_sizes(const dgpu_config *cfg, int64_t *n_planes, int64_t *n_slices);

/* filter_chan(n) (for/solver.f90:1597-1685), as diablo.f90's time loop calls it every
 * filter_int(n) steps when filter_th(n) (for/diablo.f90:76-81): scalar n
 * (0-based) times the sharpened raised-cosine filter sigma(kx, kz) on rows
 * jstart_th..jend_th. */
int dgpu_filter_chan(dgpu_handle *h, int32_t n);

/* Grid sizes Nx, Ny, Nz and N_th of a handle's configuration (for buffers a
 * caller sizes after creation, e.g. the checkpoint library). */
int dgpu_get_sizes(const dgpu_handle *h, int32_t *nx, int32_t *ny, int32_t *nz, int32_t *nth);

/* Text of the last error on this thread ("" if none). */
const char *dgpu_last_error(void);

/* Sizes the host needs for Fortran-layout buffers: number of doubles in one
 * Fourier (= physical) array, i.e. 2*(Nx/2+1)*(Nz+2)*(Ny+2). */
int64_t dgpu_field_doubles(const dgpu_config *cfg);

#ifdef __cplusplus
}
#endif
#endif /* DIABLO_GPU_H */
