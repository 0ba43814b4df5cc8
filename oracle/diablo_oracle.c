/*
 * diablo_oracle.c — CPU FP64 restatement of DIABLO's rk_chan_1 path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP step and the timed
 * CPU baseline.  Nothing in the product links or loads it.
 *
 * Every routine restates the reference routine named beside it (file:line into
 * wienkers/diablo3).  The structure is the reference's: the state is advanced
 * with 3D transforms over all Ny+2 planes, the implicit Crank–Nicolson solves are
 * real Thomas solves per physical (x, z) column over rows 0..Nyp+1, and the
 * projection is a complex Thomas solve per (kx, kz) column over rows 1..Nyp.
 *
 * Differences of form (not of arithmetic):
 *  - The reference aliases each physical array with its Fourier array (one FFTW
 *    buffer, for/fft.f90:71-96).  rk_chan_1 never reads a Fourier array after
 *    transforming it to physical space (or the reverse) before rewriting it, so
 *    here the two views are separate arrays.  Planes that the implicit solve does
 *    not write keep the values of the inverse transform and are forward
 *    transformed again, as in the reference.
 *  - FFTW (third-party, not vendored, "v3.3+") is replaced by a radix-2
 *    transform with the FFTW sign and scaling conventions: forward exp(-i..),
 *    unnormalised; c2r ignores the imaginary part of the DC bin.  Powers of two only.
 *  - The per-column CN matrices are rebuilt from their initial values for every
 *    z-plane.  The reference builds them once and overwrites, per plane, every row
 *    that Thomas modified (rows jstart..jend by the build, the two end rows by the
 *    BC routines), so the two are the same.
 * The one rank of every config makes ghost_chan_mpi a no-op (for/flow.f90:885).
 */
#include "diablo_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

static __thread char g_err[512];
/* Arithmetic variant switch for round-off sensitivity studies: bit 0 makes the
 * back substitutions multiply by 1/b instead of dividing by b; bit 1 solves the
 * projection's complex systems with their rows in reverse order (elimination from
 * the top wall down).  0 = the reference. */
static int g_variant = 0;
void oracle_set_variant(int flags) { g_variant = flags; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *oracle_last_error(void) { return g_err; }

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* radix-2 complex FFT                                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n;
  zc *w;   /* exp(-2 pi i k / n), k < n/2, rounded from long double */
  int *rev;
} fftplan;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->w = (zc *)malloc(sizeof(zc) * (size_t)(n / 2 > 0 ? n / 2 : 1));
  p->rev = (int *)malloc(sizeof(int) * (size_t)n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int k = 0; k < n / 2; k++) {
    long double a = twopi * (long double)k / (long double)n;
    p->w[k] = (double)cosl(a) - I * (double)sinl(a);
  }
  int lg = 0;
  while ((1 << lg) < n) lg++;
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int b = 0; b < lg; b++)
      if (i & (1 << b)) r |= 1 << (lg - 1 - b);
    p->rev[i] = r;
  }
}
static void plan_free(fftplan *p) { free(p->w); free(p->rev); }

/* sign = -1: sum_j a_j exp(-2 pi i jk/n); sign = +1: exp(+...).  Unnormalised. */
static void fft_c(const fftplan *p, zc *a, int sign) {
  const int n = p->n;
  for (int i = 0; i < n; i++) {
    int r = p->rev[i];
    if (r > i) { zc t = a[i]; a[i] = a[r]; a[r] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len)
      for (int j = 0; j < half; j++) {
        zc w = p->w[j * step];
        if (sign > 0) w = conj(w);
        zc u = a[i + j], v = a[i + j + half] * w;
        a[i + j] = u + v;
        a[i + j + half] = u - v;
      }
  }
}

/* ------------------------------------------------------------------------- */
/* oracle state                                                               */
/* ------------------------------------------------------------------------- */
struct oracle {
  dgpu_config c;
  int nx, ny, nz, nth, nxp, nkx, nkz, twonkz, kzr, ny2;
  size_t nf, np;                  /* elements per Fourier / physical array */
  double *gy, *gyf, *dy, *dyf;    /* (0:Nyp+1) */
  double *kx, *kz, *kx2, *kz2;    /* kx(0:Nxp-1), kz(0:2Nkz) */
  zc *cu1, *cu2, *cu3, *cp, *cth; /* Fourier [j][k][i], i < Nxp, k <= 2Nkz */
  zc *cf1, *cf2, *cf3, *cfth, *cr1, *cr2, *cr3, *crth, *cs1;
  double *u1, *u2, *u3, *th, *r1, *r2, *r3, *rth, *s1; /* physical [j][z][x] */
  double *nu_t, *kappa_t;
  /* LES work arrays (allocated when use_les): physical [j][z][x] and a
   * Fourier scratch; the index names follow les.f90 */
  /* variable_dt (for/parameters.f90:53, for/solver.f90:1241-1244) */
  int variable_dt, update_dt, cfl_ran;
  double cfl, cfl_dt;
  double *les_p;   /* NLESP physical arrays, see les_phys() */
  zc *les_c;       /* one Fourier scratch array */
  fftplan px, pz;
  /* test hooks of the learned closure (les_model_type 6), caller-owned arrays:
   * cap [ny2][nz][nx][CAP_N] receives the type-5 GY quantities of the next LES
   * substep; ov_cgy, ov_kgy [ny2][nz][nx] replace the GY coefficients */
  double *cap;
  const double *ov_cgy, *ov_kgy;
};
/* capture slots per GY point: the scaled strain and rotation entries a1..a6,
 * o4..o6 the GY coefficient is formed from, that coefficient, the scaled scalar
 * gradient ax, ay, az, the unscaled th_y and the scalar coefficient */
enum { CAP_A1 = 0, CAP_O4 = 6, CAP_V = 9, CAP_BX = 10, CAP_TY = 13, CAP_VT = 14, CAP_N = 16 };
/* physical LES arrays: Sij1..6, SIij1..6, Oij4..6, du1dx du2dy du3dz du2dx du3dx
 * du1dz du2dz du1dy du3dy, temp, temp_th, s1_th, then dthetadx/dy/dz per scalar */
enum { L_S1 = 0, L_SI1 = 6, L_O4 = 12, L_DU1DX = 15, L_DU2DY, L_DU3DZ, L_DU2DX, L_DU3DX,
       L_DU1DZ, L_DU2DZ, L_DU1DY, L_DU3DY, L_TEMP, L_TEMPTH, L_S1TH, L_TH0 };
#define NLESP(nth) (L_TH0 + 3 * (nth))

#define FI(j, k, i) ((((size_t)(j)) * o->kzr + (size_t)(k)) * o->nxp + (size_t)(i))
#define PI(j, z, x) ((((size_t)(j)) * o->nz + (size_t)(z)) * o->nx + (size_t)(x))

/* for/fft.f90:147, 209-225 — pi is 4*atan(1.0) in single precision */
void oracle_wavenumbers(int nx, int nz, double lx, double lz, double *kx, double *kz) {
  const double pi = (double)(4.0f * atanf(1.0f));
  const int nxp = nx / 2, nkz = nz / 3, twonkz = 2 * nkz;
  for (int i = 0; i < nxp; i++) kx[i] = ((double)i * (2.0 * pi)) / lx;
  for (int k = 0; k <= nkz; k++) kz[k] = ((double)k * (2.0 * pi)) / lz;
  for (int k = 1; k <= nkz; k++) kz[twonkz + 1 - k] = -(((double)k * (2.0 * pi)) / lz);
}

/* ReadGridHDF5 (for/domain.f90:533-564) + create_grid_chan (for/domain.f90:276-300).
 * gy_file holds the ny+1 values of grid.h5:/grids/y; gy(0) stays 0 as in the
 * reference, where it is never read. */
void oracle_grid_metrics(int ny, const double *gy_file, double *gy, double *gyf,
                         double *dy, double *dyf) {
  const int n2 = ny + 2;
  for (int j = 0; j < n2; j++) gy[j] = gyf[j] = dy[j] = dyf[j] = 0.0;
  for (int j = 1; j <= ny + 1; j++) gy[j] = gy_file[j - 1];
  for (int j = 1; j <= ny; j++) gyf[j] = 0.5 * (gy[j] + gy[j + 1]);
  gyf[0] = 2.0 * gyf[1] - gyf[2];
  gyf[ny + 1] = 2.0 * gyf[ny] - gyf[ny - 1];
  for (int j = 1; j <= ny + 1; j++) dy[j] = gyf[j] - gyf[j - 1];
  for (int j = 1; j <= ny; j++) dyf[j] = gy[j + 1] - gy[j];
  dyf[0] = dyf[1];
  dyf[ny + 1] = dyf[ny];
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

int oracle_create(const dgpu_config *cfg, oracle_t **out) {
  if (!cfg || !out) { set_err("null argument"); return 1; }
  if (!is_pow2(cfg->nx) || !is_pow2(cfg->nz) || cfg->nx < 4 || cfg->nz < 4) {
    set_err("oracle: nx and nz must be powers of two >= 4");
    return 2;
  }
  if (cfg->ny < 4 || cfg->n_th < 0 || cfg->n_th > DGPU_MAX_TH) {
    set_err("oracle: bad ny or n_th");
    return 2;
  }
  oracle_t *o = (oracle_t *)xcalloc(1, sizeof *o);
  o->c = *cfg;
  o->nx = cfg->nx; o->ny = cfg->ny; o->nz = cfg->nz; o->nth = cfg->n_th;
  o->nxp = o->nx / 2; o->nkx = o->nx / 3; o->nkz = o->nz / 3;
  o->twonkz = 2 * o->nkz; o->kzr = o->twonkz + 1; o->ny2 = o->ny + 2;
  o->nf = (size_t)o->ny2 * o->kzr * o->nxp;
  o->np = (size_t)o->ny2 * o->nz * o->nx;
  const int n2 = o->ny2;
  o->gy = xcalloc(n2, 8); o->gyf = xcalloc(n2, 8); o->dy = xcalloc(n2, 8); o->dyf = xcalloc(n2, 8);
  memcpy(o->gy, cfg->gy, 8 * n2); memcpy(o->gyf, cfg->gyf, 8 * n2);
  memcpy(o->dy, cfg->dy, 8 * n2); memcpy(o->dyf, cfg->dyf, 8 * n2);
  o->kx = xcalloc(o->nxp, 8); o->kx2 = xcalloc(o->nxp, 8);
  o->kz = xcalloc(o->kzr, 8); o->kz2 = xcalloc(o->kzr, 8);
  memcpy(o->kx, cfg->kx, 8 * (size_t)o->nxp);
  memcpy(o->kz, cfg->kz, 8 * (size_t)o->kzr);
  for (int i = 0; i < o->nxp; i++) o->kx2[i] = o->kx[i] * o->kx[i];
  for (int k = 0; k < o->kzr; k++) o->kz2[k] = o->kz[k] * o->kz[k];
  const int nt = o->nth > 0 ? o->nth : 1;
  o->cu1 = xcalloc(o->nf, sizeof(zc)); o->cu2 = xcalloc(o->nf, sizeof(zc));
  o->cu3 = xcalloc(o->nf, sizeof(zc)); o->cp = xcalloc(o->nf, sizeof(zc));
  o->cth = xcalloc(o->nf * nt, sizeof(zc));
  o->cf1 = xcalloc(o->nf, sizeof(zc)); o->cf2 = xcalloc(o->nf, sizeof(zc));
  o->cf3 = xcalloc(o->nf, sizeof(zc)); o->cfth = xcalloc(o->nf * nt, sizeof(zc));
  o->cr1 = xcalloc(o->nf, sizeof(zc)); o->cr2 = xcalloc(o->nf, sizeof(zc));
  o->cr3 = xcalloc(o->nf, sizeof(zc)); o->crth = xcalloc(o->nf * nt, sizeof(zc));
  o->cs1 = xcalloc(o->nf, sizeof(zc));
  o->u1 = xcalloc(o->np, 8); o->u2 = xcalloc(o->np, 8); o->u3 = xcalloc(o->np, 8);
  o->th = xcalloc(o->np * nt, 8);
  o->r1 = xcalloc(o->np, 8); o->r2 = xcalloc(o->np, 8); o->r3 = xcalloc(o->np, 8);
  o->rth = xcalloc(o->np * nt, 8); o->s1 = xcalloc(o->np, 8);
  o->nu_t = xcalloc(o->np, 8); o->kappa_t = xcalloc(o->np * nt, 8);
  if (cfg->use_les) {
    if (cfg->les_model_type != 1 && cfg->les_model_type != 4 && cfg->les_model_type != 5) {
      set_err("oracle: les_model_type must be 1, 4 or 5 (2 and 3 stop in the reference)");
      oracle_destroy(o);
      return 2;
    }
    o->les_p = xcalloc(o->np * NLESP(nt), 8);
    o->les_c = xcalloc(o->nf * 6, sizeof(zc));
  }
  plan_init(&o->px, o->nx);
  plan_init(&o->pz, o->nz);
  *out = o;
  return 0;
}

void oracle_destroy(oracle_t *o) {
  if (!o) return;
  void *ps[] = {o->gy, o->gyf, o->dy, o->dyf, o->kx, o->kz, o->kx2, o->kz2,
                o->cu1, o->cu2, o->cu3, o->cp, o->cth, o->cf1, o->cf2, o->cf3,
                o->cfth, o->cr1, o->cr2, o->cr3, o->crth, o->cs1, o->u1, o->u2,
                o->u3, o->th, o->r1, o->r2, o->r3, o->rth, o->s1, o->nu_t, o->kappa_t,
                o->les_p, o->les_c};
  for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
  plan_free(&o->px);
  plan_free(&o->pz);
  free(o);
}

/* ------------------------------------------------------------------------- */
/* Fortran layout <-> oracle layout                                           */
/* ------------------------------------------------------------------------- */
static size_t ffort(const oracle_t *o, int i, int k, int j) {
  return ((size_t)j * (o->nz + 2) + (size_t)k) * (o->nxp + 1) + (size_t)i;
}
static size_t pfort(const oracle_t *o, int x, int z, int j) {
  return ((size_t)j * (o->nz + 2) + (size_t)z) * (o->nx + 2) + (size_t)x;
}
static void four_in(const oracle_t *o, const double *src, zc *dst) {
  const zc *s = (const zc *)src;
  for (int j = 0; j < o->ny2; j++)
    for (int k = 0; k < o->kzr; k++)
      for (int i = 0; i < o->nxp; i++)
        dst[FI(j, k, i)] = (i <= o->nkx) ? s[ffort(o, i, k, j)] : 0.0;
}
static void four_out(const oracle_t *o, const zc *src, double *dst) {
  zc *d = (zc *)dst;
  for (int j = 0; j < o->ny2; j++)
    for (int k = 0; k < o->kzr; k++)
      for (int i = 0; i <= o->nxp; i++)
        d[ffort(o, i, k, j)] = (i <= o->nkx) ? src[FI(j, k, i)] : 0.0;
}

int oracle_upload_state(oracle_t *o, const double *cu1, const double *cu2,
                        const double *cu3, const double *cp, const double *cth) {
  const size_t fd = (size_t)2 * (o->nxp + 1) * (o->nz + 2) * o->ny2;
  four_in(o, cu1, o->cu1); four_in(o, cu2, o->cu2); four_in(o, cu3, o->cu3);
  four_in(o, cp, o->cp);
  for (int n = 0; n < o->nth; n++) four_in(o, cth + n * fd, o->cth + n * o->nf);
  return 0;
}
int oracle_download_state(oracle_t *o, double *cu1, double *cu2, double *cu3,
                          double *cp, double *cth) {
  const size_t fd = (size_t)2 * (o->nxp + 1) * (o->nz + 2) * o->ny2;
  four_out(o, o->cu1, cu1); four_out(o, o->cu2, cu2); four_out(o, o->cu3, cu3);
  four_out(o, o->cp, cp);
  for (int n = 0; n < o->nth; n++) four_out(o, o->cth + n * o->nf, cth + n * fd);
  return 0;
}
int oracle_upload_rk(oracle_t *o, const double *cf1, const double *cf2,
                     const double *cf3, const double *cfth) {
  const size_t fd = (size_t)2 * (o->nxp + 1) * (o->nz + 2) * o->ny2;
  four_in(o, cf1, o->cf1); four_in(o, cf2, o->cf2); four_in(o, cf3, o->cf3);
  for (int n = 0; n < o->nth; n++) four_in(o, cfth + n * fd, o->cfth + n * o->nf);
  return 0;
}
int oracle_download_rk(oracle_t *o, double *cf1, double *cf2, double *cf3, double *cfth) {
  const size_t fd = (size_t)2 * (o->nxp + 1) * (o->nz + 2) * o->ny2;
  four_out(o, o->cf1, cf1); four_out(o, o->cf2, cf2); four_out(o, o->cf3, cf3);
  for (int n = 0; n < o->nth; n++) four_out(o, o->cfth + n * o->nf, cfth + n * fd);
  return 0;
}
int oracle_download_les(oracle_t *o, double *nu_t, double *kappa_t) {
  const size_t pd = (size_t)(o->nx + 2) * (o->nz + 2) * o->ny2;
  for (int j = 0; j < o->ny2; j++)
    for (int z = 0; z < o->nz; z++)
      for (int x = 0; x < o->nx; x++) {
        nu_t[pfort(o, x, z, j)] = o->nu_t[PI(j, z, x)];
        for (int n = 0; n < o->nth; n++)
          kappa_t[n * pd + pfort(o, x, z, j)] = o->kappa_t[n * o->np + PI(j, z, x)];
      }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* 2D xz transforms over all planes j = 0..Nyp+1                              */
/* ------------------------------------------------------------------------- */
/* fft_xz_to_fourier (for/fft.f90:242-307): x r2c, /Nx, zero i > Nkx; z c2c
 * forward; /Nz; pack kz < 0 into k = Nkz+1..2Nkz. */
static void to_fourier(oracle_t *o, const double *v, zc *vv) {
  const int nx = o->nx, nz = o->nz, nxp = o->nxp, nkx = o->nkx, nkz = o->nkz;
#pragma omp parallel
  {
    zc *line = (zc *)malloc(sizeof(zc) * (size_t)(nx > nz ? nx : nz));
    zc *half = (zc *)malloc(sizeof(zc) * (size_t)nz * nxp);
#pragma omp for schedule(static)
    for (int j = 0; j < o->ny2; j++) {
      /* x: two real lines per complex transform */
      for (int z = 0; z < nz; z += 2) {
        const double *a = v + PI(j, z, 0), *b = v + PI(j, z + 1, 0);
        for (int x = 0; x < nx; x++) line[x] = a[x] + I * b[x];
        fft_c(&o->px, line, -1);
        for (int i = 0; i < nxp; i++) {
          zc zk = line[i], zn = conj(line[(nx - i) % nx]);
          zc A = 0.5 * (zk + zn), B = -0.5 * I * (zk - zn);
          half[(size_t)z * nxp + i] = (i <= nkx) ? A / nx : 0.0;
          half[(size_t)(z + 1) * nxp + i] = (i <= nkx) ? B / nx : 0.0;
        }
      }
      /* z: one complex transform per kx column i = 0..Nxp-1 */
      for (int i = 0; i < nxp; i++) {
        for (int z = 0; z < nz; z++) line[z] = half[(size_t)z * nxp + i];
        fft_c(&o->pz, line, -1);
        for (int k = 0; k <= nkz; k++) vv[FI(j, k, i)] = line[k] / nz;
        for (int k = 1; k <= nkz; k++) vv[FI(j, nkz + k, i)] = line[nz - 1 + k - nkz] / nz;
      }
    }
    free(line);
    free(half);
  }
}

/* fft_xz_to_physical (for/fft.f90:311-374): unpack, zero Nkz<k<Nz-Nkz, z c2c
 * backward, zero i > Nkx, x c2r; no scaling. */
static void to_physical(oracle_t *o, const zc *vv, double *v) {
  const int nx = o->nx, nz = o->nz, nxp = o->nxp, nkx = o->nkx, nkz = o->nkz;
#pragma omp parallel
  {
    zc *line = (zc *)malloc(sizeof(zc) * (size_t)(nx > nz ? nx : nz));
    zc *half = (zc *)malloc(sizeof(zc) * (size_t)nz * (nxp + 1));
#pragma omp for schedule(static)
    for (int j = 0; j < o->ny2; j++) {
      for (int i = 0; i < nxp; i++) {
        for (int z = 0; z < nz; z++) line[z] = 0.0;
        for (int k = 0; k <= nkz; k++) line[k] = vv[FI(j, k, i)];
        for (int k = 1; k <= nkz; k++) line[nz - 1 + k - nkz] = vv[FI(j, nkz + k, i)];
        fft_c(&o->pz, line, +1);
        for (int z = 0; z < nz; z++) half[(size_t)z * (nxp + 1) + i] = line[z];
      }
      for (int z = 0; z < nz; z += 2) {
        const zc *A = half + (size_t)z * (nxp + 1), *B = half + (size_t)(z + 1) * (nxp + 1);
        for (int x = 0; x < nx; x++) line[x] = 0.0;
        /* Hermitian completion; bins i > Nkx (incl. Nyquist) are zero; the
         * imaginary part of the DC bin is not part of a real signal's spectrum */
        line[0] = creal(A[0]) + I * creal(B[0]);
        for (int i = 1; i <= nkx; i++) {
          line[i] = A[i] + I * B[i];
          line[nx - i] = conj(A[i]) + I * conj(B[i]);
        }
        fft_c(&o->px, line, +1);
        double *a = v + PI(j, z, 0), *b = v + PI(j, z + 1, 0);
        for (int x = 0; x < nx; x++) { a[x] = creal(line[x]); b[x] = cimag(line[x]); }
      }
    }
    free(line);
    free(half);
  }
}

int oracle_fft_xz_to_fourier(oracle_t *o, const double *phys, double *four) {
  double *v = (double *)xcalloc(o->np, 8);
  zc *vv = (zc *)xcalloc(o->nf, sizeof(zc));
  for (int j = 0; j < o->ny2; j++)
    for (int z = 0; z < o->nz; z++)
      for (int x = 0; x < o->nx; x++) v[PI(j, z, x)] = phys[pfort(o, x, z, j)];
  to_fourier(o, v, vv);
  four_out(o, vv, four);
  free(v); free(vv);
  return 0;
}
int oracle_fft_xz_to_physical(oracle_t *o, const double *four, double *phys) {
  double *v = (double *)xcalloc(o->np, 8);
  zc *vv = (zc *)xcalloc(o->nf, sizeof(zc));
  four_in(o, four, vv);
  to_physical(o, vv, v);
  for (int j = 0; j < o->ny2; j++)
    for (int z = 0; z < o->nz; z++)
      for (int x = 0; x < o->nx; x++) phys[pfort(o, x, z, j)] = v[PI(j, z, x)];
  free(v); free(vv);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Thomas solvers at one rank                                                 */
/* ------------------------------------------------------------------------- */
/* thomas_forward_real_mpi + thomas_backward_real_mpi (for/solver.f90:1691-1761,
 * 1816-1852), rankY = 0 = NprocY-1: rows 0..Nyp+1. */
void oracle_thomas_real(int ny, double *a, double *b, const double *c, double *g) {
  for (int j = 1; j <= ny + 1; j++) {
    a[j] = -a[j] / b[j - 1];
    b[j] = b[j] + a[j] * c[j - 1];
    g[j] = g[j] + a[j] * g[j - 1];
  }
  if (g_variant & 1) {
    g[ny + 1] = g[ny + 1] * (1.0 / b[ny + 1]);
    for (int j = ny; j >= 0; j--) g[j] = (g[j] - c[j] * g[j + 1]) * (1.0 / b[j]);
    return;
  }
  g[ny + 1] = g[ny + 1] / b[ny + 1];
  for (int j = ny; j >= 0; j--) g[j] = (g[j] - c[j] * g[j + 1]) / b[j];
}

/* thomas_forward/backward_complex_mpi (for/solver.f90:1764-1813, 1855-1896): rows 1..Nyp */
static void thomas_complex(int ny, double *a, double *b, const double *c, zc *g) {
  if (g_variant & 2) {
    /* the same system, eliminated upward: row j (1 <= j < ny) loses its
     * coefficient of x(j+1) against row j+1, then x(1) and forward substitution */
    for (int j = ny - 1; j >= 1; j--) {
      const double w = c[j] / b[j + 1];
      b[j] = b[j] - w * a[j + 1];
      g[j] = g[j] - w * g[j + 1];
    }
    g[1] = g[1] / b[1];
    for (int j = 2; j <= ny; j++) g[j] = (g[j] - a[j] * g[j - 1]) / b[j];
    return;
  }
  for (int j = 2; j <= ny; j++) {
    a[j] = -a[j] / b[j - 1];
    b[j] = b[j] + a[j] * c[j - 1];
    g[j] = g[j] + a[j] * g[j - 1];
  }
  if (g_variant & 1) {
    g[ny] = g[ny] * (1.0 / b[ny]);
    for (int j = ny - 1; j >= 1; j--) g[j] = (g[j] - c[j] * g[j + 1]) * (1.0 / b[j]);
    return;
  }
  g[ny] = g[ny] / b[ny];
  for (int j = ny - 1; j >= 1; j--) g[j] = (g[j] - c[j] * g[j + 1]) / b[j];
}
void oracle_thomas_complex(int ny, double *a, double *b, const double *c, double *g) {
  thomas_complex(ny, a, b, c, (zc *)g);
}

/* ------------------------------------------------------------------------- */
/* BC rows of the implicit systems (for/solver.f90:1901-2232), one column     */
/* ------------------------------------------------------------------------- */
static void set_row(double *l, double *d, double *u, double *v, int j, double a,
                    double b, double c, double val) {
  l[j] = a; d[j] = b; u[j] = c; v[j] = val;
}

/* apply_BC_1_lower/upper, apply_BC_3_lower/upper, apply_BC_th_lower/upper: the
 * three are identical in form for fields at GYF points. */
static void bc_gyf(const oracle_t *o, int bc_lo, double c1_lo, int bc_hi, double c1_hi,
                   double *l, double *d, double *u, double *v) {
  const int ny = o->ny;
  if (bc_lo == DGPU_BC_DIRICHLET) {
    set_row(l, d, u, v, 0, 0., 1., 0., 0.);
    set_row(l, d, u, v, 1, 0., 1., 0., c1_lo);
  } else {
    set_row(l, d, u, v, 0, 0., 1., 0., 0.);
    set_row(l, d, u, v, 1, 0., -1., 1., o->dy[2] * c1_lo);
  }
  if (bc_hi == DGPU_BC_DIRICHLET) {
    set_row(l, d, u, v, ny + 1, 0., 1., 0., 0.);
    set_row(l, d, u, v, ny, 0., 1., 0., c1_hi);
  } else {
    set_row(l, d, u, v, ny, -1., 1., 0., o->dy[ny] * c1_hi);
    set_row(l, d, u, v, ny + 1, 0., 1., 0., 0.);
  }
}

/* apply_BC_2_lower/upper (for/solver.f90:1989-2054) */
static void bc_u2(const oracle_t *o, double *l, double *d, double *u, double *v) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny;
  if (c->v_bc_ymin == DGPU_BC_DIRICHLET) {
    set_row(l, d, u, v, 1, 0., 1., 0., c->v_bc_ymin_c1);
    set_row(l, d, u, v, 2, 0., 1., 0., c->v_bc_ymin_c1);
  } else if (c->v_bc_ymin == DGPU_BC_NEUMANN) {
    set_row(l, d, u, v, 1, 0., -1., 1., o->dyf[1] * c->v_bc_ymin_c1);
  }
  set_row(l, d, u, v, 0, 0., 1., 0., 0.);
  if (c->v_bc_ymax == DGPU_BC_DIRICHLET) {
    set_row(l, d, u, v, ny + 1, 0., 1., 0., c->v_bc_ymax_c1);
    set_row(l, d, u, v, ny, 0., 1., 0., c->v_bc_ymax_c1);
  } else if (c->v_bc_ymax == DGPU_BC_NEUMANN) {
    set_row(l, d, u, v, ny + 1, -1., 1., 0., o->dyf[ny] * c->v_bc_ymax_c1);
  }
}

/* w_BC_Ymax_c1_transient (for/parameters.f90:279; for/solver.f90:2131) */
static double w_top_c1(const oracle_t *o, double time) {
  const dgpu_config *c = &o->c;
  if (c->f_type == 4)
    return c->w_bc_ymax_c1 +
           c->amp_omega0 * sin(c->omega0 * (c->ro_inv / c->delta * time - c->force_start));
  return c->w_bc_ymax_c1;
}

/* CN matrix of one column (rk_chan_1 matrix builds, for/solver.f90:1020-1045,
 * 1081-1105, 1136-1161, 1192-1217) with an explicit nu_t/kappa_t column. */
static void build_cn(const oracle_t *o, int which, double temp1, double temp2,
                     const double *nut, const double *rhs, size_t nstride, double time,
                     double *l, double *d, double *u, double *v) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny;
  const double *dy = o->dy, *dyf = o->dyf;
  const double nvs = c->nu_v_scale;
  for (int j = 0; j < o->ny2; j++) { l[j] = 0.; d[j] = 1.; u[j] = 0.; v[j] = 0.; }
#define NT(jj) (nut ? nut[(size_t)(jj) * nstride] : 0.0)
#define RHS(jj) (rhs ? rhs[(size_t)(jj) * nstride] : 0.0)
  if (which >= 3) {
    const int n = which - 3;
    const double t = temp1 / c->pr[n] * nvs;
    for (int j = c->jstart_th[n]; j <= c->jend_th[n]; j++) {
      l[j] = -t / (dy[j] * dyf[j]);
      d[j] = 1. + t / (dy[j + 1] * dyf[j]) + t / (dy[j] * dyf[j]);
      u[j] = -t / (dy[j + 1] * dyf[j]);
      v[j] = RHS(j);
      if (c->use_les) {
        l[j] = l[j] - temp2 * NT(j) / (dy[j] * dyf[j]);
        d[j] = d[j] + temp2 * NT(j + 1) / (dy[j + 1] * dyf[j]) + temp2 * NT(j) / (dy[j] * dyf[j]);
        u[j] = u[j] - temp2 * NT(j + 1) / (dy[j + 1] * dyf[j]);
      }
    }
    bc_gyf(o, c->th_bc_ymin[n], c->th_bc_ymin_c1[n], c->th_bc_ymax[n], c->th_bc_ymax_c1[n],
           l, d, u, v);
  } else if (which == 1) {
    for (int j = 2; j <= ny; j++) {
      l[j] = -temp1 * nvs / (dyf[j - 1] * dy[j]);
      d[j] = 1. + temp1 * nvs / (dyf[j] * dy[j]) + temp1 * nvs / (dyf[j - 1] * dy[j]);
      u[j] = -temp1 * nvs / (dyf[j] * dy[j]);
      v[j] = RHS(j);
      if (c->use_les) {
        l[j] = l[j] - temp2 * 0.5 * (NT(j) + NT(j - 1)) / (dyf[j - 1] * dy[j]);
        d[j] = d[j] + temp2 * 0.5 * (NT(j) + NT(j + 1)) / (dyf[j] * dy[j]) +
               temp2 * 0.5 * (NT(j) + NT(j - 1)) / (dyf[j - 1] * dy[j]);
        u[j] = u[j] - temp2 * 0.5 * (NT(j) + NT(j + 1)) / (dyf[j] * dy[j]);
      }
    }
    bc_u2(o, l, d, u, v);
  } else {
    for (int j = c->jstart; j <= c->jend; j++) {
      l[j] = -temp1 * nvs / (dy[j] * dyf[j]);
      d[j] = 1. - temp1 * nvs * (-1. / (dy[j + 1] * dyf[j]) - 1. / (dy[j] * dyf[j]));
      u[j] = -temp1 * nvs / (dy[j + 1] * dyf[j]);
      v[j] = RHS(j);
      if (c->use_les) {
        l[j] = l[j] - temp2 * NT(j) / (dy[j] * dyf[j]);
        d[j] = d[j] + temp2 * NT(j + 1) / (dy[j + 1] * dyf[j]) + temp2 * NT(j) / (dy[j] * dyf[j]);
        u[j] = u[j] - temp2 * NT(j + 1) / (dy[j + 1] * dyf[j]);
      }
    }
    if (which == 0)
      bc_gyf(o, c->u_bc_ymin, c->u_bc_ymin_c1, c->u_bc_ymax, c->u_bc_ymax_c1, l, d, u, v);
    else
      bc_gyf(o, c->w_bc_ymin, c->w_bc_ymin_c1, c->w_bc_ymax, w_top_c1(o, time), l, d, u, v);
  }
#undef NT
#undef RHS
}

int oracle_cn_matrix(oracle_t *o, int which, int rk_step, double delta_t, double *matl,
                     double *matd, double *matu, double *vec_bc) {
  return oracle_cn_matrix_nut(o, which, rk_step, delta_t, NULL, matl, matd, matu, vec_bc);
}

int oracle_cn_matrix_nut(oracle_t *o, int which, int rk_step, double delta_t, const double *nut,
                         double *matl, double *matd, double *matu, double *vec_bc) {
  const double h[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  if (rk_step < 1 || rk_step > 3 || which < 0 || which >= 3 + o->nth) return 1;
  if (nut && !o->c.use_les) { set_err("a nu_t column needs use_les"); return 1; }
  const double temp1 = 0.5 * o->c.nu * h[rk_step - 1], temp2 = 0.5 * h[rk_step - 1];
  build_cn(o, which, temp1, temp2, nut, NULL, 1, 0.0, matl, matd, matu, vec_bc);
  return 0;
}

/* The physical CN right-hand side r1, r2, r3 or rth(n) (which = 0, 1, 2, 3+n) as
 * the last substep left it: the inverse transform of cr plus the explicit half of
 * the vertical viscous terms (for/solver.f90:841-903, 967-1002). */
int oracle_debug_rhs(oracle_t *o, int which, double *out) {
  const double *src;
  if (which == 0) src = o->r1;
  else if (which == 1) src = o->r2;
  else if (which == 2) src = o->r3;
  else if (which >= 3 && which < 3 + o->nth) src = o->rth + (size_t)(which - 3) * o->np;
  else { set_err("which out of range"); return 1; }
  for (int j = 0; j < o->ny2; j++)
    for (int z = 0; z < o->nz; z++)
      for (int x = 0; x < o->nx; x++) out[pfort(o, x, z, j)] = src[PI(j, z, x)];
  return 0;
}

/* ------------------------------------------------------------------------- */
/* rem_div_chan (for/solver.f90:1302-1379) + apply_BC_rem_div_mpi (2330-2371) */
/* ------------------------------------------------------------------------- */
static void rem_div(oracle_t *o) {
  const int ny = o->ny, nxp = o->nxp;
  const double *dy = o->dy, *dyf = o->dyf;
#pragma omp parallel
  {
    double *a = (double *)malloc(8 * (size_t)o->ny2);
    double *b = (double *)malloc(8 * (size_t)o->ny2);
    double *c = (double *)malloc(8 * (size_t)o->ny2);
    zc *g = (zc *)malloc(sizeof(zc) * (size_t)o->ny2);
#pragma omp for schedule(static)
    for (int k = 0; k < o->kzr; k++) {
      for (int i = 0; i < nxp; i++) {
        for (int j = 0; j < o->ny2; j++) { a[j] = 0.; b[j] = 1.; c[j] = 0.; g[j] = 0.; }
        for (int j = 1; j <= ny; j++) {
          a[j] = 1. / (dy[j] * dyf[j]);
          b[j] = -o->kx2[i] - o->kz2[k] - 1. / (dy[j + 1] * dyf[j]) - 1. / (dy[j] * dyf[j]);
          c[j] = 1. / (dy[j + 1] * dyf[j]);
        }
        for (int j = 1; j <= ny; j++) {
          const zc ikx = I * o->kx[i], ikz = I * o->kz[k];
          g[j] = ikx * o->cu1[FI(j, k, i)] +
                 (o->cu2[FI(j + 1, k, i)] - o->cu2[FI(j, k, i)]) / dyf[j] +
                 ikz * o->cu3[FI(j, k, i)];
        }
        if (k == 0 && i == 0) { a[1] = 0.; b[1] = 1.; c[1] = 0.; g[1] = 0.; }
        else { a[1] = 0.; b[1] = 1.; c[1] = -1.; g[1] = 0.; }
        a[ny] = 1.; b[ny] = -1.; c[ny] = 0.; g[ny] = 0.;
        thomas_complex(ny, a, b, c, g);
        for (int j = 1; j <= ny; j++) o->cr1[FI(j, k, i)] = g[j];
      }
    }
    free(a); free(b); free(c); free(g);
  }
  for (int j = 1; j <= ny; j++)
    for (int k = 0; k < o->kzr; k++)
      for (int i = 0; i < nxp; i++) {
        const zc ikx = I * o->kx[i], ikz = I * o->kz[k];
        o->cu1[FI(j, k, i)] = o->cu1[FI(j, k, i)] - ikx * o->cr1[FI(j, k, i)];
        o->cu3[FI(j, k, i)] = o->cu3[FI(j, k, i)] - ikz * o->cr1[FI(j, k, i)];
      }
  for (int j = 2; j <= ny; j++)
    for (int k = 0; k < o->kzr; k++)
      for (int i = 0; i < nxp; i++)
        o->cu2[FI(j, k, i)] =
            o->cu2[FI(j, k, i)] - (o->cr1[FI(j, k, i)] - o->cr1[FI(j - 1, k, i)]) / dy[j];
}

int oracle_rem_div(oracle_t *o) {
  rem_div(o);
  return 0;
}

/* filter_chan (for/solver.f90:1597-1685): the sharpened raised-cosine filter of
 * scalar n in the horizontal, sigma = s^4 (35 - 84 s + 70 s^2 - 20 s^3) with
 * s = (1 + cos(sqrt((kx Lx/Nx)^2 + (kz Lz/Nz)^2)))/2, on rows jstart_th..jend_th
 * (the vertical compact filter of the reference is commented out there). */
int oracle_filter_chan(oracle_t *o, int n) {
  const dgpu_config *c = &o->c;
  if (n < 0 || n >= c->n_th) return 2;
  zc *cth = o->cth + n * o->nf;
  for (int k = 0; k < o->kzr; k++)
    for (int i = 0; i < o->nxp; i++) {
      const double ax = o->kx[i] * c->lx * 1.0 / (double)(float)c->nx;
      const double az = o->kz[k] * c->lz * 1.0 / (double)(float)c->nz;
      const double s0 = 0.5 * (1.0 + cos(sqrt(pow(ax, 2.0) + pow(az, 2.0))));
      const double sg = pow(s0, 4.0) * (35.0 - 84.0 * s0 + 70.0 * pow(s0, 2.0) - 20.0 * pow(s0, 3.0));
      for (int j = c->jstart_th[n]; j <= c->jend_th[n]; j++) cth[FI(j, k, i)] = cth[FI(j, k, i)] * sg;
    }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* rk_chan_1 (for/solver.f90:366-1273)                                        */
/* ------------------------------------------------------------------------- */
/* physical-space CN solve of one field over all z-planes: matrix build + BC +
 * Thomas per (x, z) column, result written on rows [jlo, jhi]. */
static void cn_solve_phys(oracle_t *o, int which, double temp1, double temp2,
                          const double *rhs, double *out, int jlo, int jhi,
                          const double *nut, double time) {
  const int nx = o->nx, ny2 = o->ny2;
  const size_t stride = (size_t)o->nz * nx;
#pragma omp parallel
  {
    double *l = (double *)malloc(8 * (size_t)ny2), *d = (double *)malloc(8 * (size_t)ny2);
    double *u = (double *)malloc(8 * (size_t)ny2), *v = (double *)malloc(8 * (size_t)ny2);
#pragma omp for schedule(static)
    for (int z = 0; z < o->nz; z++) {
      for (int x = 0; x < nx; x++) {
        const double *ncol = nut ? nut + PI(0, z, x) : NULL;
        build_cn(o, which, temp1, temp2, ncol, rhs + PI(0, z, x), stride, time, l, d, u, v);
        oracle_thomas_real(o->ny, l, d, u, v);
        for (int j = jlo; j <= jhi; j++) out[PI(j, z, x)] = v[j];
      }
    }
    free(l); free(d); free(u); free(v);
  }
}

/* ------------------------------------------------------------------------- */
/* LES: AMD closure (les_model_type 5), for/les.f90                           */
/* ------------------------------------------------------------------------- */
/* apply_BC_vel_mpi_post / apply_BC_th_mpi_post at one rank: the lower and the
 * upper routines both run (for/flow.f90:1161-1360, 1596-1732).  Loops run over
 * i < Nxp, k <= 2Nkz; the mean mode (0,0) is the rankZ == 0 special case. */
static void bc_post(oracle_t *o, double time) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny, kzr = o->kzr, nxp = o->nxp;
  const double *dy = o->dy, *dyf = o->dyf;
  zc *u1 = o->cu1, *u2 = o->cu2, *u3 = o->cu3;
#define PL(a, j) for (int k = 0; k < kzr; k++) for (int i = 0; i < nxp; i++)
  /* lower: u1, u3 (1168-1220) */
  zc *uu[2] = {u1, u3};
  const int bl[2] = {c->u_bc_ymin, c->w_bc_ymin};
  const double cl[2] = {c->u_bc_ymin_c1, c->w_bc_ymin_c1};
  for (int m = 0; m < 2; m++) {
    zc *a = uu[m];
    if (bl[m] == DGPU_BC_DIRICHLET) {
      PL(a, 0) { a[FI(0, k, i)] = 0.0; a[FI(1, k, i)] = 0.0; }
      a[FI(1, 0, 0)] = cl[m];
      a[FI(0, 0, 0)] = cl[m];
    } else {
      PL(a, 0) { a[FI(1, k, i)] = a[FI(2, k, i)]; a[FI(0, k, i)] = a[FI(1, k, i)]; }
      a[FI(1, 0, 0)] = a[FI(2, 0, 0)] - dy[2] * cl[m];
      a[FI(0, 0, 0)] = a[FI(1, 0, 0)] - dy[1] * cl[m];
    }
  }
  /* lower: u2 (1222-1243); the Dirichlet form adds 2 c1 to every mode */
  if (c->v_bc_ymin == DGPU_BC_DIRICHLET) {
    PL(u2, 0) { u2[FI(1, k, i)] = 2.0 * c->v_bc_ymin_c1 - u2[FI(2, k, i)]; u2[FI(0, k, i)] = u2[FI(1, k, i)]; }
  } else {
    PL(u2, 0) { u2[FI(1, k, i)] = u2[FI(2, k, i)]; u2[FI(0, k, i)] = u2[FI(1, k, i)]; }
    u2[FI(1, 0, 0)] = u2[FI(2, 0, 0)] - dyf[1] * c->v_bc_ymin_c1;
    u2[FI(0, 0, 0)] = u2[FI(1, 0, 0)] - dyf[1] * c->v_bc_ymin_c1;
  }
  /* upper: u1 (1266-1288), u3 (1290-1334) */
  const int bu[2] = {c->u_bc_ymax, c->w_bc_ymax};
  const double cu[2] = {c->u_bc_ymax_c1, w_top_c1(o, time)};
  for (int m = 0; m < 2; m++) {
    zc *a = uu[m];
    if (bu[m] == DGPU_BC_DIRICHLET) {
      PL(a, 0) { a[FI(ny, k, i)] = 0.0; a[FI(ny + 1, k, i)] = 0.0; }
      a[FI(ny, 0, 0)] = (m == 0) ? c->u_bc_ymax_c1 : c->w_bc_ymax_c1;
      a[FI(ny + 1, 0, 0)] = (m == 0) ? c->u_bc_ymax_c1 : c->w_bc_ymax_c1;
      if (m == 1) a[FI(ny + 1, 0, 0)] = 0.0; /* for/flow.f90:1311 */
    } else {
      PL(a, 0) { a[FI(ny, k, i)] = a[FI(ny - 1, k, i)]; a[FI(ny + 1, k, i)] = a[FI(ny, k, i)]; }
      a[FI(ny, 0, 0)] = a[FI(ny - 1, 0, 0)] + dy[ny] * cu[m];
      a[FI(ny + 1, 0, 0)] = a[FI(ny, 0, 0)] + dy[ny] * cu[m];
    }
  }
  /* upper: u2 (1336-1356); the Dirichlet form touches only the mean mode */
  if (c->v_bc_ymax == DGPU_BC_DIRICHLET) {
    u2[FI(ny + 1, 0, 0)] = 2.0 * c->v_bc_ymax_c1 - u2[FI(ny, 0, 0)];
  } else {
    PL(u2, 0) { u2[FI(ny, k, i)] = u2[FI(ny - 1, k, i)]; u2[FI(ny + 1, k, i)] = u2[FI(ny, k, i)]; }
    u2[FI(ny + 1, 0, 0)] = u2[FI(ny, 0, 0)] + dy[ny] * c->v_bc_ymax_c1;
  }
  /* scalars (1596-1689) */
  for (int n = 0; n < o->nth; n++) {
    zc *a = o->cth + n * o->nf;
    if (c->th_bc_ymin[n] == DGPU_BC_DIRICHLET) {
      PL(a, 0) { a[FI(0, k, i)] = 0.0; a[FI(1, k, i)] = 0.0; }
      a[FI(1, 0, 0)] = c->th_bc_ymin_c1[n];
      a[FI(0, 0, 0)] = c->th_bc_ymin_c1[n];
    } else {
      PL(a, 0) { a[FI(1, k, i)] = a[FI(2, k, i)]; a[FI(0, k, i)] = a[FI(1, k, i)]; }
      a[FI(1, 0, 0)] = a[FI(2, 0, 0)] - dy[2] * c->th_bc_ymin_c1[n];
      a[FI(0, 0, 0)] = a[FI(1, 0, 0)] - dy[1] * c->th_bc_ymin_c1[n];
    }
    if (c->th_bc_ymax[n] == DGPU_BC_DIRICHLET) {
      PL(a, 0) { a[FI(ny, k, i)] = 0.0; a[FI(ny + 1, k, i)] = 0.0; }
      a[FI(ny, 0, 0)] = c->th_bc_ymax_c1[n];
      a[FI(ny + 1, 0, 0)] = c->th_bc_ymax_c1[n];
    } else {
      PL(a, 0) { a[FI(ny, k, i)] = a[FI(ny - 1, k, i)]; a[FI(ny + 1, k, i)] = a[FI(ny, k, i)]; }
      a[FI(ny, 0, 0)] = a[FI(ny - 1, 0, 0)] + dy[ny] * c->th_bc_ymax_c1[n];
      a[FI(ny + 1, 0, 0)] = a[FI(ny, 0, 0)] + dy[ny] * c->th_bc_ymax_c1[n];
    }
  }
}

/* apply_BC_les / apply_BC_th_les (for/les.f90:1980-2079): with a Neumann
 * condition the two ghost planes copy the first interior plane, mean included */
static void bc_les(oracle_t *o) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny, kzr = o->kzr, nxp = o->nxp;
  zc *uu[2] = {o->cu1, o->cu3};
  const int bu[2] = {c->u_bc_ymax, c->w_bc_ymax}, bl[2] = {c->u_bc_ymin, c->w_bc_ymin};
  for (int m = 0; m < 2; m++) {
    zc *a = uu[m];
    if (bu[m] == DGPU_BC_NEUMANN)
      PL(a, 0) { a[FI(ny, k, i)] = a[FI(ny - 1, k, i)]; a[FI(ny + 1, k, i)] = a[FI(ny, k, i)]; }
  }
  for (int m = 0; m < 2; m++) {
    zc *a = uu[m];
    if (bl[m] == DGPU_BC_NEUMANN)
      PL(a, 0) { a[FI(1, k, i)] = a[FI(2, k, i)]; a[FI(0, k, i)] = a[FI(1, k, i)]; }
  }
  for (int n = 0; n < o->nth; n++) {
    zc *a = o->cth + n * o->nf;
    if (c->th_bc_ymax[n] == DGPU_BC_NEUMANN)
      PL(a, 0) { a[FI(ny, k, i)] = a[FI(ny - 1, k, i)]; a[FI(ny + 1, k, i)] = a[FI(ny, k, i)]; }
    if (c->th_bc_ymin[n] == DGPU_BC_NEUMANN)
      PL(a, 0) { a[FI(1, k, i)] = a[FI(2, k, i)]; a[FI(0, k, i)] = a[FI(1, k, i)]; }
  }
#undef PL
}

#define LOOP_PX(jlo, jhi) \
  for (int j = (jlo); j <= (jhi); j++) for (int z = 0; z < nz; z++) for (int x = 0; x < nx; x++)
#define LOOP_FX(jlo, jhi) \
  for (int j = (jlo); j <= (jhi); j++) for (int k = 0; k < kzr; k++) for (int i = 0; i < nxp; i++)
static double *les_phys(oracle_t *o, int slot) { return o->les_p + (size_t)slot * o->np; }

/* The filter-width factors of les_chan type 5 (for/les.f90:756-780, 1316-1338):
 * delta(j) = -2 C_AMD^2 * 3 / (1/(dx beta_sgs)^2 + 1/(2 h_j)^2 + 1/(dz beta_sgs)^2),
 * with h = dyf (GYF points) or dy (GY points). */
#define C_AMD 0.2887
#define BETA_SGS 3.0
static double amd_delta(const oracle_t *o, double h) {
  const double dx = o->c.lx / o->nx, dz = o->c.lz / o->nz;
  const double a = dx * BETA_SGS, b = h * 2.0, cz = dz * BETA_SGS;
  return -(2.0 * (C_AMD * C_AMD) * 3.0 / (1.0 / (a * a) + 1.0 / (b * b) + 1.0 / (cz * cz)));
}

/* Fourier field built per plane by f(j, k, i), then transformed to physical
 * space (the compute_* routines of for/les.f90:1443-1746).  Planes outside
 * [jlo, jhi] are zero here; in the reference they hold stale values that no
 * later statement reads. */
static void les_grad(oracle_t *o, int jlo, int jhi, zc (*f)(const oracle_t *, int, int, int, int),
                     int which, double *out) {
  zc *cs = o->les_c;
  memset(cs, 0, sizeof(zc) * o->nf);
#pragma omp parallel for schedule(static)
  for (int j = jlo; j <= jhi; j++)
    for (int k = 0; k < o->kzr; k++)
      for (int i = 0; i < o->nxp; i++) cs[FI(j, k, i)] = f(o, j, k, i, which);
  to_physical(o, cs, out);
}

/* Every spectral field of compute_strain_chan, compute_strain_chan_invariant,
 * compute_rotation_chan_invariant and compute_all_gradients_chan, by id. */
enum { G_S1, G_S2, G_S3, G_S4, G_S5, G_S6, G_SI1, G_SI2, G_SI3, G_SI4, G_SI5, G_SI6,
       G_O4, G_O5, G_O6, G_R4, G_R5, G_R6, G_DU1DX, G_DU2DY, G_DU3DZ, G_DU2DX, G_DU3DX, G_DU1DZ, G_DU2DZ,
       G_DU1DY, G_DU3DY, G_DTHDX, G_DTHDY, G_DTHDZ };
static zc les_spec(const oracle_t *o, int j, int k, int i, int which) {
  const zc ikx = I * o->kx[i], ikz = I * o->kz[k];
  const double dx = o->c.lx / o->nx * BETA_SGS, dz = o->c.lz / o->nz * BETA_SGS;
  const zc *u1 = o->cu1, *u2 = o->cu2, *u3 = o->cu3;
  const size_t p = FI(j, k, i);
  switch (which) {
    /* compute_strain_chan (1452-1474) */
    case G_S1: return ikx * u1[p];
    case G_S3: return ikz * u3[p];
    case G_S5: return 0.5 * (ikz * u1[p] + ikx * u3[p]);
    case G_S2: return (u2[FI(j + 1, k, i)] - u2[p]) / o->dyf[j];
    case G_S4: return 0.5 * ((u1[p] - u1[FI(j - 1, k, i)]) / o->dy[j] + ikx * u2[p]);
    case G_S6: return 0.5 * (ikz * u2[p] + (u3[p] - u3[FI(j - 1, k, i)]) / o->dy[j]);
    /* compute_strain_chan_invariant (1510-1542) */
    case G_SI1: return ikx * u1[p];
    case G_SI3: return ikz * u3[p];
    case G_SI5: return 0.5 * (ikz * u1[p] * (dz / dx) + ikx * u3[p] * (dx / dz));
    case G_SI2: return (u2[FI(j + 1, k, i)] - u2[p]) / o->dyf[j];
    case G_SI4: {
      const double dyl = o->dy[j] * 2.0;
      return 0.5 * ((u1[p] - u1[FI(j - 1, k, i)]) / o->dy[j] * (dyl / dx) + ikx * u2[p] * (dx / dyl));
    }
    case G_SI6: {
      const double dyl = o->dy[j] * 2.0;
      return 0.5 * (ikz * u2[p] * (dz / dyl) + (u3[p] - u3[FI(j - 1, k, i)]) / o->dy[j] * (dyl / dz));
    }
    /* compute_rotation_chan_invariant (1617-1645) */
    case G_O5: return 0.5 * (ikz * u1[p] * (dz / dx) - ikx * u3[p] * (dx / dz));
    case G_O4: {
      const double dyl = o->dy[j] * 2.0;
      return 0.5 * ((u1[p] - u1[FI(j - 1, k, i)]) / o->dy[j] * (dyl / dx) - ikx * u2[p] * (dx / dyl));
    }
    case G_O6: {
      const double dyl = o->dy[j] * 2.0;
      return 0.5 * (ikz * u2[p] * (dz / dyl) - (u3[p] - u3[FI(j - 1, k, i)]) / o->dy[j] * (dyl / dz));
    }
    /* compute_rotation_chan (1572-1589) */
    case G_R5: return 0.5 * (ikz * u1[p] - ikx * u3[p]);
    case G_R4: return 0.5 * ((u1[p] - u1[FI(j - 1, k, i)]) / o->dy[j] - ikx * u2[p]);
    case G_R6: return 0.5 * (ikz * u2[p] - (u3[p] - u3[FI(j - 1, k, i)]) / o->dy[j]);
    /* compute_all_gradients_chan (1663-1713) */
    case G_DU1DX: return ikx * u1[p];
    case G_DU1DZ: return ikz * u1[p];
    case G_DU2DX: return ikx * u2[p];
    case G_DU2DY: return (u2[FI(j + 1, k, i)] - u2[p]) / o->dyf[j];
    case G_DU2DZ: return ikz * u2[p];
    case G_DU3DX: return ikx * u3[p];
    case G_DU3DZ: return ikz * u3[p];
    case G_DU1DY: return (u1[p] - u1[FI(j - 1, k, i)]) / o->dy[j];
    case G_DU3DY: return (u3[p] - u3[FI(j - 1, k, i)]) / o->dy[j];
    default: return 0.0;
  }
}
/* the scalar gradients of compute_all_gradients_chan (1690-1706), scalar n */
static int g_les_n = 0;
static zc les_spec_th(const oracle_t *o, int j, int k, int i, int which) {
  const zc *t = o->cth + (size_t)g_les_n * o->nf;
  const size_t p = FI(j, k, i);
  if (which == G_DTHDX) return I * o->kx[i] * t[p];
  if (which == G_DTHDZ) return I * o->kz[k] * t[p];
  return (t[p] - t[FI(j - 1, k, i)]) / o->dy[j];
}

/* The part of les_chan common to every model (for/les.f90:800-901): tau_ij
 * (the coefficient times S_ij, in S) to Fourier space, the filter-width factors,
 * the stress divergence into cf1..cf3 on the one-rank rows, and the cross terms
 * -nu_t du/dy treated explicitly; temp is scratch. */
static void les_sgs_cf(oracle_t *o, double **S, double *const *dyfac, const double *u1, const double *u3,
                       double *temp) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny, nx = o->nx, nz = o->nz, nxp = o->nxp, kzr = o->kzr;
  const int js = c->jstart;
  const double *dy = o->dy, *dyf = o->dyf;
  /* tau_ij in Fourier space and its divergence into cf (for/les.f90:800-901) */
  zc *C1 = o->les_c + o->nf, *C3 = o->les_c + 2 * o->nf, *C4 = o->les_c + 3 * o->nf;
  zc *C5 = o->les_c + 4 * o->nf, *C6 = o->les_c + 5 * o->nf;
  to_fourier(o, S[0], C1);
  to_fourier(o, S[2], C3);
  to_fourier(o, S[3], C4);
  to_fourier(o, S[4], C5);
  to_fourier(o, S[5], C6);
  LOOP_FX(1, ny) {
    const size_t p = FI(j, k, i);
    C1[p] = dyfac[0][j] * C1[p];
    C5[p] = dyfac[0][j] * C5[p];
    C3[p] = dyfac[0][j] * C3[p];
  }
  LOOP_FX(1, ny + 1) {
    const size_t p = FI(j, k, i);
    C4[p] = dyfac[1][j] * C4[p];
    C6[p] = dyfac[1][j] * C6[p];
  }
  /* one rank: rankY == NprocY-1 sets j1 = jstart, j2 = Nyp-1 (for/les.f90:45-57) */
  LOOP_FX(js, ny - 1) {
    const size_t p = FI(j, k, i), pu = FI(j + 1, k, i);
    const zc ikx = I * o->kx[i], ikz = I * o->kz[k];
    o->cf1[p] = o->cf1[p] - ikx * C1[p] - (C4[pu] - C4[p]) / dyf[j] - ikz * C5[p];
    o->cf3[p] = o->cf3[p] - ikx * C5[p] - (C6[pu] - C6[p]) / dyf[j] - ikz * C3[p];
  }
  LOOP_FX(2, ny) {
    const size_t p = FI(j, k, i);
    o->cf2[p] = o->cf2[p] - I * o->kx[i] * C4[p] - I * o->kz[k] * C6[p];
  }
  /* the cross terms treated explicitly (for/les.f90:874-901) */
  for (int m = 0; m < 2; m++) {
    const double *uu = m == 0 ? u1 : u3;
    memset(temp, 0, 8 * o->np);
    LOOP_PX(2, ny) {
      const size_t p = PI(j, z, x);
      temp[p] = o->nu_t[p] * (uu[p] - uu[PI(j - 1, z, x)]) / dy[j];
    }
    to_fourier(o, temp, C6);
    LOOP_FX(2, ny) {
      const size_t p = FI(j, k, i);
      o->cf2[p] = o->cf2[p] + (m == 0 ? I * o->kx[i] : I * o->kz[k]) * C6[p];
    }
  }

}

#define C_SMAG 0.17
/* les_chan types 1 and 4 after compute_strain_chan (and compute_rotation_chan for
 * type 4), then les_chan_th types 1/4 (for/les.f90:1118-1193) */
static void les_chan_14(oracle_t *o, double **S, double **O) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny, nx = o->nx, nz = o->nz, nxp = o->nxp, kzr = o->kzr;
  const double *dy = o->dy, *dyf = o->dyf;
  const int model = c->les_model_type;
  to_physical(o, o->cu1, o->u1);
  to_physical(o, o->cu2, o->u2);
  to_physical(o, o->cu3, o->u3);
  const double *u1 = o->u1, *u3 = o->u3;
  double *s1 = o->s1, *temp = les_phys(o, L_TEMP);
  memset(s1, 0, 8 * o->np);
  memset(temp, 0, 8 * o->np);
  const double bx = c->lx / nx * BETA_SGS, bz = c->lz / nz * BETA_SGS;
  double *dyfac[2];
  dyfac[0] = (double *)xcalloc((size_t)o->ny2, 8);
  dyfac[1] = (double *)xcalloc((size_t)o->ny2, 8);
  if (model == 1) {
    /* |S| at GYF (for/les.f90:99-112) and at GY (114-132) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
          const double s4 = 0.5 * (S[3][q] + S[3][p]), s6 = 0.5 * (S[5][q] + S[5][p]);
          s1[p] = sqrt(2.0 * (S[0][p] * S[0][p]) + 4.0 * (s4 * s4) + 4.0 * (S[4][p] * S[4][p]) +
                       2.0 * (S[1][p] * S[1][p]) + 4.0 * (s6 * s6) + 2.0 * (S[2][p] * S[2][p]));
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny + 1; j++) {
      const double wc = dyf[j - 1], wm = dyf[j], den = 2.0 * dy[j];
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
          const double a1 = (S[0][p] * wc + S[0][m] * wm) / den, a5 = (S[4][p] * wc + S[4][m] * wm) / den;
          const double a2 = (S[1][p] * wc + S[1][m] * wm) / den, a3 = (S[2][p] * wc + S[2][m] * wm) / den;
          temp[p] = sqrt(2.0 * (a1 * a1) + 4.0 * (S[3][p] * S[3][p]) + 4.0 * (a5 * a5) + 2.0 * (a2 * a2) +
                         4.0 * (S[5][p] * S[5][p]) + 2.0 * (a3 * a3));
        }
    }
    /* (for/les.f90:134-158) */
    for (int j = 1; j <= ny; j++)
      for (size_t q = PI(j, 0, 0); q < PI(j + 1, 0, 0); q++) {
        S[0][q] = s1[q] * S[0][q];
        S[4][q] = s1[q] * S[4][q];
        S[1][q] = 0.0;
        S[2][q] = s1[q] * S[2][q];
      }
    for (int j = 1; j <= ny + 1; j++)
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
          S[3][p] = temp[p] * (S[3][p] - 0.5 * (u1[p] - u1[m]) / dy[j]);
          S[5][p] = temp[p] * (S[5][p] - 0.5 * (u3[p] - u3[m]) / dy[j]);
        }
    /* (for/les.f90:161-170); the GYF width uses dyf(j) * 1, the GY width dy(j) * 2 */
    for (int j = 1; j <= ny; j++)
      dyfac[0][j] = -(2.0 * (C_SMAG * C_SMAG) * pow(bx * dyf[j] * 1.0 * bz, 2.0 / 3.0));
    for (int j = 1; j <= ny + 1; j++)
      dyfac[1][j] = -(2.0 * (C_SMAG * C_SMAG) * pow(bx * dy[j] * 2.0 * bz, 2.0 / 3.0));
  } else {
    /* type 4 at GYF (for/les.f90:276-343): weights (dx beta)^2, (2 dy(j))^2, (dz beta)^2 */
    const double dX = bx * bx, dZ = bz * bz;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++) {
      const double dY = (dy[j] * 2.0) * (dy[j] * 2.0);
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
          const double a1 = S[0][p], a2 = S[1][p], a3 = S[2][p], a5 = S[4][p], o5 = O[1][p];
          const double a4 = 0.5 * (S[3][q] + S[3][p]), a6 = 0.5 * (S[5][q] + S[5][p]);
          const double o4 = 0.5 * (O[0][q] + O[0][p]), o6 = 0.5 * (O[2][q] + O[2][p]);
          double v = dX * (a1 * a1 * a1) + dY * (a2 * a2 * a2) + dZ * (a3 * a3 * a3) +
                     (2.0 * dX + dY) * a1 * (a4 * a4) + (2.0 * dX + dZ) * a1 * (a5 * a5) +
                     (2.0 * dY + dX) * a2 * (a4 * a4) + (2.0 * dY + dZ) * a2 * (a6 * a6) +
                     (2.0 * dZ + dX) * a3 * (a5 * a5) + (2.0 * dZ + dY) * a3 * (a6 * a6) +
                     2.0 * (dX + dY + dZ) * a4 * a5 * a6;
          v = -v - dY * a1 * (o4 * o4) - dZ * a1 * (o5 * o5) - dX * a2 * (o4 * o4) - dZ * a2 * (o6 * o6) -
              dX * a3 * (o5 * o5) - dY * a3 * (o6 * o6) - 2.0 * dZ * a4 * o5 * o6 + 2.0 * dY * a5 * o4 * o6 -
              2.0 * dX * a6 * o4 * o5;
          if (v <= 0.0) v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2.0 * (a4 * a4) + 2.0 * (o4 * o4) + 2.0 * (a5 * a5) +
                     2.0 * (o5 * o5) + 2.0 * (a6 * a6) + 2.0 * (o6 * o6));
          s1[p] = v;
        }
    }
    /* at GY (for/les.f90:345-440) */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny + 1; j++) {
      const double dY = (dy[j] * 2.0) * (dy[j] * 2.0);
      const double wc = dyf[j - 1], wm = dyf[j], den = 2.0 * dy[j];
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
          const double a1 = (S[0][p] * wc + S[0][m] * wm) / den, a2 = (S[1][p] * wc + S[1][m] * wm) / den;
          const double a3 = (S[2][p] * wc + S[2][m] * wm) / den, a5 = (S[4][p] * wc + S[4][m] * wm) / den;
          const double o5 = (O[1][p] * wc + O[1][m] * wm) / den;
          const double a4 = S[3][p], a6 = S[5][p], o4 = O[0][p], o6 = O[2][p];
          double v = dX * (a1 * a1 * a1) + dY * (a2 * a2 * a2) + dZ * (a3 * a3 * a3) +
                     (2.0 * dX + dY) * a1 * (a4 * a4) + (2.0 * dX + dZ) * a1 * (a5 * a5) +
                     (2.0 * dY + dX) * a2 * (a4 * a4) + (2.0 * dY + dZ) * a2 * (a6 * a6) +
                     (2.0 * dZ + dX) * a3 * (a5 * a5) + (2.0 * dZ + dY) * a3 * (a6 * a6) +
                     2.0 * (dX + dY + dZ) * a4 * a6 * a5;
          v = -v - dY * a1 * (o4 * o4) - dZ * a1 * (o5 * o5) - dX * a2 * (o4 * o4) - dZ * a2 * (o6 * o6) -
              dX * a3 * (o5 * o5) - dY * a3 * (o6 * o6) - 2.0 * dZ * a4 * o5 * o6 + 2.0 * dY * a5 * o4 * o6 -
              2.0 * dX * a6 * o4 * o5;
          if (v <= 0.0) v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2.0 * (a4 * a4) + 2.0 * (o4 * o4) + 2.0 * (a5 * a5) +
                     2.0 * (o5 * o5) + 2.0 * (a6 * a6) + 2.0 * (o6 * o6));
          temp[p] = v;
        }
    }
    /* (for/les.f90:444-463); s1(Nyp+1) is never set by the reference (0 here) */
    for (int j = 1; j <= ny + 1; j++)
      for (size_t q = PI(j, 0, 0); q < PI(j + 1, 0, 0); q++) {
        S[0][q] = s1[q] * S[0][q];
        S[4][q] = s1[q] * S[4][q];
        S[1][q] = 0.0;
        S[2][q] = s1[q] * S[2][q];
      }
    for (int j = 2; j <= ny + 1; j++)
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
          S[3][p] = temp[p] * (S[3][p] - 0.5 * (u1[p] - u1[m]) / dy[j]);
          S[5][p] = temp[p] * (S[5][p] - 0.5 * (u3[p] - u3[m]) / dy[j]);
        }
    /* (for/les.f90:464-469): a constant factor */
    for (int j = 1; j <= ny + 1; j++) dyfac[0][j] = dyfac[1][j] = -(2.0 * (C_AMD * C_AMD));
  }
  /* nu_t = -delta_y/2 * coefficient at GY on rows 2..Nyp, zero at 1, 2, Nyp, Nyp+1 */
  for (int j = 2; j <= ny; j++)
    for (size_t q = PI(j, 0, 0); q < PI(j + 1, 0, 0); q++) o->nu_t[q] = -0.5 * dyfac[1][j] * temp[q];
  for (size_t q = PI(1, 0, 0); q < PI(3, 0, 0); q++) o->nu_t[q] = 0.0;
  for (size_t q = PI(ny, 0, 0); q < PI(ny + 2, 0, 0); q++) o->nu_t[q] = 0.0;
  les_sgs_cf(o, S, dyfac, u1, u3, temp);
  free(dyfac[0]);
  free(dyfac[1]);
  /* les_chan_th types 1 and 4: kappa_t = nu_t, flux kappa_t grad(th) (for/les.f90:1118-1193) */
  zc *cs = o->les_c;
  double *f = o->s1;
  for (int n = 0; n < o->nth; n++) {
    double *kt = o->kappa_t + n * o->np;
    zc *cth = o->cth + n * o->nf, *cfth = o->cfth + n * o->nf;
    for (int j = 1; j <= ny + 1; j++)
      for (size_t q = PI(j, 0, 0); q < PI(j + 1, 0, 0); q++) kt[q] = 1.0 * o->nu_t[q];
    for (int d = 0; d < 2; d++) {
      memset(cs, 0, sizeof(zc) * o->nf);
      for (int j = 1; j <= ny + 1; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++)
            cs[FI(j, k, i)] = (d == 0 ? I * o->kx[i] : I * o->kz[k]) * cth[FI(j, k, i)];
      to_physical(o, cs, f);
      for (int j = 1; j <= ny + 1; j++)
        for (size_t q = PI(j, 0, 0); q < PI(j + 1, 0, 0); q++) f[q] = kt[q] * f[q];
      to_fourier(o, f, cs);
      for (int j = c->jstart_th[n]; j <= c->jend_th[n]; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++)
            cfth[FI(j, k, i)] = cfth[FI(j, k, i)] + (d == 0 ? I * o->kx[i] : I * o->kz[k]) * cs[FI(j, k, i)];
    }
  }
  for (int n = 0; n < o->nth; n++) to_physical(o, o->cth + n * o->nf, o->th + n * o->np);
}

/* les_chan (for/les.f90:3-75, 77-901): eddy viscosity nu_t at GY points and
 * the SGS stress divergence added to cf1..cf3, u1..u3 left in physical space, for
 * les_model_type 1 (constant Smagorinsky, 77-254), 4 (AMD with the filter widths
 * as weights, 260-550) and 5 (AMD of the scaled gradient, 551-797).  Then
 * les_chan_th (1072-1395): type 5 its own AMD diffusivity, types 1 and 4
 * kappa_t = nu_t; the SGS flux divergence added to cfth; th in physical space. */
static void les_amd(oracle_t *o, double time) {
  const dgpu_config *c = &o->c;
  const int ny = o->ny, nx = o->nx, nz = o->nz, nxp = o->nxp, kzr = o->kzr;
  const int js = c->jstart, je = c->jend;
  const double *dy = o->dy, *dyf = o->dyf;
  /* for/les.f90:63-75 */
  bc_post(o, time);
  bc_les(o);
  const int model = c->les_model_type;
  /* compute_strain_chan: Sij1..Sij6.  CSij2(Nyp+1) holds the -10000 sentinel of
   * ghost_CS2_mpi (for/les.f90:2509-2513); Sij2 is read by types 1 and 4 only. */
  double *S[6], *SI[6], *O[3];
  for (int m = 0; m < 6; m++) { S[m] = les_phys(o, L_S1 + m); SI[m] = les_phys(o, L_SI1 + m); }
  for (int m = 0; m < 3; m++) O[m] = les_phys(o, L_O4 + m);
  les_grad(o, 0, ny + 1, les_spec, G_S1, S[0]);
  les_grad(o, 0, ny + 1, les_spec, G_S3, S[2]);
  les_grad(o, 0, ny + 1, les_spec, G_S5, S[4]);
  les_grad(o, 1, ny + 1, les_spec, G_S4, S[3]);
  les_grad(o, 1, ny + 1, les_spec, G_S6, S[5]);
  if (model != 5) {
    zc *cs = o->les_c;
    memset(cs, 0, sizeof(zc) * o->nf);
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) cs[FI(j, k, i)] = les_spec(o, j, k, i, G_S2);
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) cs[FI(ny + 1, k, i)] = -10000.0;
    to_physical(o, cs, S[1]);
  }
  if (model == 4) {
    les_grad(o, 1, ny + 1, les_spec, G_R4, O[0]);
    les_grad(o, 0, ny + 1, les_spec, G_R5, O[1]);
    les_grad(o, 1, ny + 1, les_spec, G_R6, O[2]);
  }
  if (model == 1 || model == 4) {
    les_chan_14(o, S, O);
    return;
  }
  /* compute_strain_chan_invariant, with the sentinel plane of CSIij2 */
  les_grad(o, 0, ny + 1, les_spec, G_SI1, SI[0]);
  {
    zc *cs = o->les_c;
    memset(cs, 0, sizeof(zc) * o->nf);
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) cs[FI(j, k, i)] = les_spec(o, j, k, i, G_SI2);
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) cs[FI(ny + 1, k, i)] = -10000.0;
    to_physical(o, cs, SI[1]);
  }
  les_grad(o, 0, ny + 1, les_spec, G_SI3, SI[2]);
  les_grad(o, 1, ny + 1, les_spec, G_SI4, SI[3]);
  les_grad(o, 0, ny + 1, les_spec, G_SI5, SI[4]);
  les_grad(o, 1, ny + 1, les_spec, G_SI6, SI[5]);
  /* compute_rotation_chan_invariant */
  les_grad(o, 1, ny + 1, les_spec, G_O4, O[0]);
  les_grad(o, 0, ny + 1, les_spec, G_O5, O[1]);
  les_grad(o, 1, ny + 1, les_spec, G_O6, O[2]);
  /* compute_all_gradients_chan; du2dy at Nyp+1 would read cu2(Nyp+2), outside
   * the array, and feeds only temp_th(Nyp+1), which nothing reads */
  double *du1dx = les_phys(o, L_DU1DX), *du2dy = les_phys(o, L_DU2DY), *du3dz = les_phys(o, L_DU3DZ);
  double *du2dx = les_phys(o, L_DU2DX), *du3dx = les_phys(o, L_DU3DX), *du1dz = les_phys(o, L_DU1DZ);
  double *du2dz = les_phys(o, L_DU2DZ), *du1dy = les_phys(o, L_DU1DY), *du3dy = les_phys(o, L_DU3DY);
  les_grad(o, 0, ny + 1, les_spec, G_DU1DX, du1dx);
  les_grad(o, 0, ny, les_spec, G_DU2DY, du2dy);
  les_grad(o, 0, ny + 1, les_spec, G_DU3DZ, du3dz);
  les_grad(o, 0, ny + 1, les_spec, G_DU2DX, du2dx);
  les_grad(o, 0, ny + 1, les_spec, G_DU3DX, du3dx);
  les_grad(o, 0, ny + 1, les_spec, G_DU1DZ, du1dz);
  les_grad(o, 0, ny + 1, les_spec, G_DU2DZ, du2dz);
  les_grad(o, 1, ny + 1, les_spec, G_DU1DY, du1dy);
  les_grad(o, 1, ny + 1, les_spec, G_DU3DY, du3dy);
  for (int n = 0; n < o->nth; n++) {
    g_les_n = n;
    les_grad(o, 1, ny + 1, les_spec_th, G_DTHDX, les_phys(o, L_TH0 + 3 * n));
    les_grad(o, 1, ny + 1, les_spec_th, G_DTHDY, les_phys(o, L_TH0 + 3 * n + 1));
    les_grad(o, 1, ny + 1, les_spec_th, G_DTHDZ, les_phys(o, L_TH0 + 3 * n + 2));
  }
  to_physical(o, o->cu1, o->u1);
  to_physical(o, o->cu2, o->u2);
  to_physical(o, o->cu3, o->u3);
  const double *u1 = o->u1, *u3 = o->u3;
  double *s1 = o->s1, *temp = les_phys(o, L_TEMP);

  /* max{0,-(I3-I4)}/(I1-I2) at GYF points (for/les.f90:574-619) */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; j++)
    for (int z = 0; z < nz; z++)
      for (int x = 0; x < nx; x++) {
        const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
        const double a1 = SI[0][p], a2 = SI[1][p], a3 = SI[2][p], a5 = SI[4][p];
        const double a4 = 0.5 * (SI[3][q] + SI[3][p]), a6 = 0.5 * (SI[5][q] + SI[5][p]);
        const double o4 = 0.5 * (O[0][q] + O[0][p]), o5 = O[1][p], o6 = 0.5 * (O[2][q] + O[2][p]);
        double v = a1 * a1 * a1 + a2 * a2 * a2 + a3 * a3 * a3 + 3.0 * a1 * (a4 * a4) + 3.0 * a1 * (a5 * a5) +
                   3.0 * a2 * (a4 * a4) + 3.0 * a2 * (a6 * a6) + 3.0 * a3 * (a5 * a5) + 3.0 * a3 * (a6 * a6) +
                   6.0 * a4 * a5 * a6;
        v = -v - a1 * (o4 * o4) - a1 * (o5 * o5) - a2 * (o4 * o4) - a2 * (o6 * o6) - a3 * (o5 * o5) -
            a3 * (o6 * o6) - 2.0 * a4 * o5 * o6 + 2.0 * a5 * o4 * o6 - 2.0 * a6 * o4 * o5;
        if (v <= 0.0) v = 0.0;
        else
          v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2.0 * (a4 * a4) + 2.0 * (o4 * o4) + 2.0 * (a5 * a5) +
                   2.0 * (o5 * o5) + 2.0 * (a6 * a6) + 2.0 * (o6 * o6));
        s1[p] = v;
      }
  /* the same at GY points, GYF values interpolated (for/les.f90:625-740) */
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= ny; j++) {
    const double wc = dyf[j - 1], wm = dyf[j], den = 2.0 * dy[j];
    for (int z = 0; z < nz; z++)
      for (int x = 0; x < nx; x++) {
        const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
        const double a1 = (SI[0][p] * wc + SI[0][m] * wm) / den;
        const double a2 = (SI[1][p] * wc + SI[1][m] * wm) / den;
        const double a3 = (SI[2][p] * wc + SI[2][m] * wm) / den;
        const double a5 = (SI[4][p] * wc + SI[4][m] * wm) / den;
        const double o5 = (O[1][p] * wc + O[1][m] * wm) / den;
        const double a4 = SI[3][p], a6 = SI[5][p], o4 = O[0][p], o6 = O[2][p];
        double v = a1 * a1 * a1 + a2 * a2 * a2 + a3 * a3 * a3 + 3.0 * a1 * (a4 * a4) + 3.0 * a1 * (a5 * a5) +
                   3.0 * a2 * (a4 * a4) + 3.0 * a2 * (a6 * a6) + 3.0 * a3 * (a5 * a5) + 3.0 * a3 * (a6 * a6) +
                   6.0 * a4 * a6 * a5;
        v = -v - a1 * (o4 * o4) - a1 * (o5 * o5) - a2 * (o4 * o4) - a2 * (o6 * o6) - a3 * (o5 * o5) -
            a3 * (o6 * o6) - 2.0 * a4 * o5 * o6 + 2.0 * a5 * o4 * o6 - 2.0 * a6 * o4 * o5;
        if (v <= 0.0) v = 0.0;
        else
          v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2.0 * (a4 * a4) + 2.0 * (o4 * o4) + 2.0 * (a5 * a5) +
                   2.0 * (o5 * o5) + 2.0 * (a6 * a6) + 2.0 * (o6 * o6));
        temp[p] = v;
        if (o->cap) {
          double *cp = o->cap + p * CAP_N;
          cp[CAP_A1] = a1; cp[CAP_A1 + 1] = a2; cp[CAP_A1 + 2] = a3;
          cp[CAP_A1 + 3] = a4; cp[CAP_A1 + 4] = a5; cp[CAP_A1 + 5] = a6;
          cp[CAP_O4] = o4; cp[CAP_O4 + 1] = o5; cp[CAP_O4 + 2] = o6;
          cp[CAP_V] = v;
        }
      }
  }
  if (o->ov_cgy) {
    /* les_model_type 6: the GY coefficient |c|; at GYF the mean of the two GY
     * neighbours when neither carries the fallback sign, else the type-5 value */
    const double *cg = o->ov_cgy;
    LOOP_PX(js, je) {
      const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
      if (!signbit(cg[p]) && !signbit(cg[q])) s1[p] = 0.5 * (cg[p] + cg[q]);
    }
    LOOP_PX(2, ny) temp[PI(j, z, x)] = fabs(cg[PI(j, z, x)]);
  }
  /* coefficient times S_ij (for/les.f90:745-773); temp(Nyp+1) is never set by
   * the reference (a stale value that feeds only planes nothing reads): 0 here */
  for (size_t q = PI(ny + 1, 0, 0); q < PI(ny + 2, 0, 0); q++) temp[q] = 0.0;
  LOOP_PX(js, je) {
    const size_t p = PI(j, z, x);
    S[0][p] = s1[p] * S[0][p];
    S[4][p] = s1[p] * S[4][p];
    S[1][p] = 0.0;
    S[2][p] = s1[p] * S[2][p];
  }
  LOOP_PX(2, ny + 1) {
    const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
    S[3][p] = temp[p] * (S[3][p] - 0.5 * (u1[p] - u1[m]) / dy[j]);
    S[5][p] = temp[p] * (S[5][p] - 0.5 * (u3[p] - u3[m]) / dy[j]);
  }
  double *dyfac[2];
  dyfac[0] = (double *)xcalloc((size_t)o->ny2, 8);
  dyfac[1] = (double *)xcalloc((size_t)o->ny2, 8);
  for (int j = 1; j <= ny + 1; j++) { dyfac[0][j] = amd_delta(o, dyf[j]); dyfac[1][j] = amd_delta(o, dy[j]); }
  /* nu_t = -delta_y/2 * coefficient at GY, zero at j = 1, 2, Nyp, Nyp+1 (ghost_les_mpi) */
  LOOP_PX(2, ny + 1) o->nu_t[PI(j, z, x)] = -0.5 * dyfac[1][j] * temp[PI(j, z, x)];
  LOOP_PX(1, 2) o->nu_t[PI(j, z, x)] = 0.0;
  LOOP_PX(ny, ny + 1) o->nu_t[PI(j, z, x)] = 0.0;
  les_sgs_cf(o, S, dyfac, u1, u3, temp);

  /* ---- les_chan_th, type 5 ---- */
  zc *C1 = o->les_c + o->nf, *C5 = o->les_c + 4 * o->nf;
  const double dxs = c->lx / nx * BETA_SGS, dzs = c->lz / nz * BETA_SGS;
  double *s1th = les_phys(o, L_S1TH), *tth = les_phys(o, L_TEMPTH);
  for (int n = 0; n < o->nth; n++) to_physical(o, o->cth + n * o->nf, o->th + n * o->np);
  for (int n = 0; n < o->nth; n++) {
    double *tx = les_phys(o, L_TH0 + 3 * n), *ty = les_phys(o, L_TH0 + 3 * n + 1);
    double *tz = les_phys(o, L_TH0 + 3 * n + 2);
    const int jst = c->jstart_th[n], jet = c->jend_th[n];
    /* numerator / denominator at GYF (for/les.f90:1221-1256) */
#pragma omp parallel for schedule(static)
    for (int j = jst; j <= jet; j++) {
      const double dyl = dyf[j] * 2.0;
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
          const double tyc = 0.5 * (ty[q] + ty[p]);
          double v = -(dxs * dxs * du1dx[p] * (tx[p] * tx[p]) + dyl * dyl * du2dy[p] * (tyc * tyc) +
                       dzs * dzs * du3dz[p] * (tz[p] * tz[p]) +
                       (dyl * dyl * 0.5 * (du1dy[q] + du1dy[p]) + dxs * dxs * 0.5 * (du2dx[q] + du2dx[p])) *
                           tx[p] * tyc +
                       (dzs * dzs * du1dz[p] + dxs * dxs * du3dx[p]) * tx[p] * tz[p] +
                       (dzs * dzs * 0.5 * (du2dz[q] + du2dz[p]) + dyl * dyl * 0.5 * (du3dy[q] + du3dy[p])) *
                           tyc * tz[p]);
          if (v <= 0.0) v = 0.0;
          else {
            const double ax = dxs * tx[p], ay = dyl * tyc, az = dzs * tz[p];
            v = v / (ax * ax + ay * ay + az * az);
          }
          s1th[p] = v;
        }
    }
    /* at GY (for/les.f90:1259-1306), j = 2..Nyp: temp_th(Nyp+1) is not read */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny; j++) {
      const double dyl = dy[j] * 2.0, wc = dyf[j - 1], wm = dyf[j], den = 2.0 * dy[j];
      for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
          const size_t p = PI(j, z, x), m = PI(j - 1, z, x);
          /* evaluation order of the reference expression */
          const double S2 = tx[p] * wc + tx[m] * wm, S5 = tz[p] * wc + tz[m] * wm;
          const double X2 = dzs * dzs * (du1dz[p] * wc + du1dz[m] * wm) / den +
                            dxs * dxs * (du3dx[p] * wc + du3dx[m] * wm) / den;
          double v = -(dxs * dxs * (du1dx[p] * wc + du1dx[m] * wm) / den * ((S2 / den) * (S2 / den)) +
                       dyl * dyl * (du2dy[p] * wc + du2dy[m] * wm) / den * (ty[p] * ty[p]) +
                       dzs * dzs * (du3dz[p] * wc + du3dz[m] * wm) / den * ((S5 / den) * (S5 / den)) +
                       (dyl * dyl * du1dy[p] + dxs * dxs * du2dx[p]) * S2 / den * ty[p] +
                       X2 * S2 / den * S5 / den +
                       (dzs * dzs * du2dz[p] + dyl * dyl * du3dy[p]) * ty[p] * S5 / den);
          if (v <= 0.0) v = 0.0;
          else {
            const double ax = dxs * S2 / den, ay = dyl * ty[p], az = dzs * S5 / den;
            v = v / (ax * ax + ay * ay + az * az);
          }
          tth[p] = v;
          if (o->cap && n == 0) {
            double *cp = o->cap + p * CAP_N;
            cp[CAP_BX] = dxs * S2 / den; cp[CAP_BX + 1] = dyl * ty[p]; cp[CAP_BX + 2] = dzs * S5 / den;
            cp[CAP_TY] = ty[p];
            cp[CAP_VT] = v;
          }
        }
    }
    if (o->ov_kgy) {
      const double *kg = o->ov_kgy;
      LOOP_PX(jst, jet) {
        const size_t p = PI(j, z, x), q = PI(j + 1, z, x);
        if (!signbit(kg[p]) && !signbit(kg[q])) s1th[p] = 0.5 * (kg[p] + kg[q]);
      }
      LOOP_PX(2, ny) tth[PI(j, z, x)] = fabs(kg[PI(j, z, x)]);
    }
    LOOP_PX(jst, jet) {
      const size_t p = PI(j, z, x);
      tx[p] = s1th[p] * tx[p];
      tz[p] = s1th[p] * tz[p];
    }
    double *kt = o->kappa_t + n * o->np;
    LOOP_PX(2, ny) kt[PI(j, z, x)] = -0.5 * dyfac[1][j] * tth[PI(j, z, x)];
    LOOP_PX(1, 2) kt[PI(j, z, x)] = 0.0;
    LOOP_PX(ny, ny + 1) kt[PI(j, z, x)] = 0.0;
    to_fourier(o, tx, C1);
    to_fourier(o, tz, C5);
    LOOP_FX(1, ny + 1) {
      const size_t p = FI(j, k, i);
      C1[p] = 0.5 * dyfac[0][j] * C1[p];
      C5[p] = 0.5 * dyfac[0][j] * C5[p];
    }
    zc *cfth = o->cfth + n * o->nf;
    LOOP_FX(jst, jet) {
      const size_t p = FI(j, k, i);
      cfth[p] = cfth[p] - I * o->kx[i] * C1[p] - I * o->kz[k] * C5[p];
    }
  }
  free(dyfac[0]);
  free(dyfac[1]);
}
#undef LOOP_PX
#undef LOOP_FX

int oracle_les_capture(oracle_t *o, double *cap) {
  if (!o->c.use_les || o->c.les_model_type != 5) { set_err("capture needs an les_model_type 5 handle"); return 1; }
  o->cap = cap;
  return 0;
}

int oracle_les_override(oracle_t *o, const double *cgy, const double *kgy) {
  if (!o->c.use_les || o->c.les_model_type != 5) { set_err("override needs an les_model_type 5 handle"); return 1; }
  o->ov_cgy = cgy;
  o->ov_kgy = kgy;
  return 0;
}

/* apply_BC_vel_mpi_post + apply_BC_th_mpi_post on the Fourier state, as
 * save_stats_chan starts (for/diagnostics.f90:26-27) */
int oracle_bc_post(oracle_t *o, double time) {
  bc_post(o, time);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* courant (for/flow.f90:746-857)                                             */
/* ------------------------------------------------------------------------- */
/* The time-step limits of courant up to get_minimum_mpi: viscous, diffusive,
 * inertial and buoyancy periods, and the CFL limit of the physical velocity
 * (with the thermal wind for flavour 'Front'). */
static double courant_dt(const oracle_t *o) {
  const dgpu_config *c = &o->c;
  const double pi = 3.1415927410125732;      /* pi = 4.*atan(1.0) in real(4), for/fft.f90:147 */
  const double dx = c->lx / o->nx, dz = c->lz / o->nz;
  double dt = 1.0;
  dt = fmin(dt, 0.5 * pow(fmin(dx, dz), 2.0 * c->beta) / c->nu);
  for (int n = 0; n < o->nth; n++) dt = fmin(dt, dt * c->nu / (c->nu / c->pr[n]));
  if (c->ro_inv != 0.0) dt = fmin(dt, 2.0 * pi / fabs(c->ro_inv / c->delta) / 20.0);
  for (int n = 0; n < o->nth; n++) {
    const double nmax = sqrt(fabs(c->th_bc_ymin_c1[n]));
    dt = fmin(dt, (double)0.1f * 2.0 * pi / nmax);  /* the literal 0.1 is real(4) */
  }
  const double cfl = o->cfl;
  if (c->flavor_front) {
    for (int n = 0; n < o->nth; n++)
      for (int j = c->jstart; j <= c->jend; j++)
        for (int z = 0; z < o->nz; z++)
          for (int x = 0; x < o->nx; x++) {
            const size_t p = PI(j, z, x);
            const double dtx = cfl * dx / fabs(o->u1[p] - c->dthdz[n] * o->gyf[j] / (c->ro_inv / c->delta));
            const double dty = cfl * o->dy[j] / fabs(o->u2[p]);
            const double dtz = cfl * dz / fabs(o->u3[p] + (1.0 / (c->ro_inv / c->delta)) * c->dthdx[n] *
                                                           (o->gyf[j] - 0.5 * c->ly));
            dt = fmin(dt, fmin(dtx, fmin(dty, dtz)));
          }
  } else {
    for (int j = 1; j <= o->ny; j++)
      for (int z = 0; z < o->nz; z++)
        for (int x = 0; x < o->nx; x++) {
          const size_t p = PI(j, z, x);
          const double dtx = cfl * dx / fabs(o->u1[p]);
          const double dty = cfl * o->dy[j] / fabs(o->u2[p]);
          const double dtz = cfl * dz / fabs(o->u3[p]);
          dt = fmin(dt, fmin(dtx, fmin(dty, dtz)));
        }
  }
  return dt;
}

/* the rest of courant: caps and the scheduling towards the next event */
double oracle_courant_schedule(double dt, double time, double time_limit, double save_stats_time,
                               double save_flow_time) {
  double delta_t;
  if (dt <= 0) delta_t = 0.00001;
  else if (dt >= (double)1000.f) delta_t = 1000.0;
  else delta_t = dt;
  double next = fmin(time_limit, fmin(save_stats_time - time, fmin(save_flow_time - time, time_limit - time)));
  const double r_max = (double)(1.f - 0.025f);
  if (next < dt) delta_t = next + 1.e-14;
  else if (next < (1.0 + r_max) * dt) delta_t = 1.0 / (1.0 + r_max) * next;
  else delta_t = dt;
  return delta_t;
}

/* courant's limits for the current Fourier state (transformed to physical space) */
double oracle_courant_dt_state(oracle_t *o, double cfl) {
  to_physical(o, o->cu1, o->u1);
  to_physical(o, o->cu2, o->u2);
  to_physical(o, o->cu3, o->u3);
  const double keep = o->cfl;
  o->cfl = cfl;
  const double dt = courant_dt(o);
  o->cfl = keep;
  return dt;
}

int oracle_set_variable_dt(oracle_t *o, int variable_dt, int update_dt, double cfl) {
  if (variable_dt && update_dt < 1) { set_err("update_dt must be >= 1"); return 1; }
  o->variable_dt = variable_dt;
  o->update_dt = update_dt;
  o->cfl = cfl;
  o->cfl_ran = 0;
  return 0;
}

int oracle_courant(oracle_t *o, double time, double time_limit, double save_stats_time,
                   double save_flow_time, double *delta_t, int *ran) {
  *ran = o->cfl_ran;
  *delta_t = o->cfl_ran ? oracle_courant_schedule(o->cfl_dt, time, time_limit, save_stats_time, save_flow_time)
                        : 0.0;
  o->cfl_ran = 0;
  return 0;
}

int oracle_rk_substep(oracle_t *o, int rk_step, double delta_t, double time,
                      int time_step, int create_new_flow) {
  const dgpu_config *c = &o->c;
  if (rk_step < 1 || rk_step > 3) { set_err("rk_step must be 1..3"); return 1; }
  const int ny = o->ny, nxp = o->nxp, kzr = o->kzr, nx = o->nx, nz = o->nz, nth = o->nth;
  const int js = c->jstart, je = c->jend;
  const double *dy = o->dy, *dyf = o->dyf, *gyf = o->gyf, *gy = o->gy;
  /* RKW3 coefficients (for/parameters.f90:176-184) */
  const double h_bar[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  const double beta_bar[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0};
  const double zeta_bar[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0};
  const int r = rk_step - 1;
  const double temp1 = 0.5 * c->nu * h_bar[r];
  const double temp2 = 0.5 * h_bar[r];
  const double temp3 = zeta_bar[r] * h_bar[r];
  const double temp4 = h_bar[r];
  const double temp5 = beta_bar[r] * h_bar[r];
  const int les_on = c->use_les && ((!create_new_flow) || (time_step > 100));
  (void)gy;

#define CIKX(i) (I * o->kx[(i)])
#define CIKZ(k) (I * o->kz[(k)])
#define KPOW(i, k) pow(o->kx2[(i)] + o->kz2[(k)], c->beta)
  /* cr <- cu (396-410), + temp3*cf (413-429) */
  for (int j = js; j <= je; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) {
        o->cr1[FI(j, k, i)] = o->cu1[FI(j, k, i)];
        o->cr3[FI(j, k, i)] = o->cu3[FI(j, k, i)];
      }
  for (int j = 2; j <= ny; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) o->cr2[FI(j, k, i)] = o->cu2[FI(j, k, i)];
  if (rk_step > 1) {
    for (int j = js; j <= je; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) {
          o->cr1[FI(j, k, i)] = o->cr1[FI(j, k, i)] + temp3 * o->cf1[FI(j, k, i)];
          o->cr3[FI(j, k, i)] = o->cr3[FI(j, k, i)] + temp3 * o->cf3[FI(j, k, i)];
        }
    for (int j = 2; j <= ny; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++)
          o->cr2[FI(j, k, i)] = o->cr2[FI(j, k, i)] + temp3 * o->cf2[FI(j, k, i)];
  }
  /* pressure gradient (432-456) */
  for (int j = 2; j <= ny; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++)
        o->cs1[FI(j, k, i)] = (o->cp[FI(j, k, i)] - o->cp[FI(j - 1, k, i)]) / dy[j];
  for (int j = js; j <= je; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) {
        o->cr1[FI(j, k, i)] = o->cr1[FI(j, k, i)] - temp4 * (CIKX(i) * o->cp[FI(j, k, i)]);
        o->cr3[FI(j, k, i)] = o->cr3[FI(j, k, i)] - temp4 * (CIKZ(k) * o->cp[FI(j, k, i)]);
      }
  for (int j = 2; j <= ny; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++)
        o->cr2[FI(j, k, i)] = o->cr2[FI(j, k, i)] - temp4 * o->cs1[FI(j, k, i)];
  /* mean pressure-gradient forcing (464-476) */
  if (c->f_type == 1) {
    for (int j = js; j <= je; j++) o->cr1[FI(j, 0, 0)] = o->cr1[FI(j, 0, 0)] - temp4 * c->px0;
  } else if (c->f_type == 2) {
    for (int j = js; j <= je; j++)
      o->cr1[FI(j, 0, 0)] =
          o->cr1[FI(j, 0, 0)] - temp4 * (c->px0 + c->amp_omega0 * cos(c->omega0 * time));
  }
  /* horizontal viscous terms (484-498) and Coriolis (502-509) */
  for (int j = js; j <= je; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) {
        o->cf1[FI(j, k, i)] = -(c->nu * KPOW(i, k) * o->cu1[FI(j, k, i)]);
        o->cf3[FI(j, k, i)] = -(c->nu * KPOW(i, k) * o->cu3[FI(j, k, i)]);
      }
  for (int j = 2; j <= ny; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++)
        o->cf2[FI(j, k, i)] = -(c->nu * KPOW(i, k) * o->cu2[FI(j, k, i)]);
  for (int j = js; j <= je; j++)
    for (int k = 0; k < kzr; k++)
      for (int i = 0; i < nxp; i++) {
        o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] + (c->ro_inv / c->delta) * o->cu3[FI(j, k, i)];
        o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] - (c->ro_inv / c->delta) * o->cu1[FI(j, k, i)];
      }
  /* scalars: buoyancy (518-554), scalar RHS (565-602) */
  for (int n = 0; n < nth; n++) {
    zc *cth = o->cth + n * o->nf, *crth = o->crth + n * o->nf, *cfth = o->cfth + n * o->nf;
    for (int j = 2; j <= ny; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = (k == 0 ? 1 : 0); i < nxp; i++)
          o->cf2[FI(j, k, i)] = o->cf2[FI(j, k, i)] +
                                c->grav_y * (cth[FI(j, k, i)] * dyf[j - 1] + cth[FI(j - 1, k, i)] * dyf[j]) /
                                    (2.0 * dy[j]);
    for (int j = js; j <= je; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] + c->grav_x * cth[FI(j, k, i)];
    for (int j = js; j <= je; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] + c->grav_z * cth[FI(j, k, i)];
    const int jst = c->jstart_th[n], jet = c->jend_th[n];
    for (int j = jst; j <= jet; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++) crth[FI(j, k, i)] = cth[FI(j, k, i)];
    if (rk_step > 1)
      for (int j = jst; j <= jet; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++) crth[FI(j, k, i)] = crth[FI(j, k, i)] + temp3 * cfth[FI(j, k, i)];
    for (int j = jst; j <= jet; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++)
          cfth[FI(j, k, i)] = -((c->nu / c->pr[n]) * KPOW(i, k) * cth[FI(j, k, i)]);
    for (int j = jst; j <= jet; j++)
      for (int k = 0; k < kzr; k++)
        for (int i = 0; i < nxp; i++)
          cfth[FI(j, k, i)] = cfth[FI(j, k, i)] - o->cu1[FI(j, k, i)] * c->dthdx[n] -
                              o->cu3[FI(j, k, i)] * c->dthdz[n];
  }
  /* user_rhs_chan_fourier: thermal-wind advection (for/solver.f90:35-111) */
  if (c->flavor_front && c->ro_inv != 0.0) {
    for (int n = 0; n < nth; n++) {
      const double dtx = c->dthdx[n], dtz = c->dthdz[n], dl = c->delta, ro = c->ro_inv;
      const double ly = c->ly;
      for (int j = js; j <= je; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++) {
            const zc u2m = 0.5 * (o->cu2[FI(j, k, i)] + o->cu2[FI(j + 1, k, i)]);
            o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] -
                                  (dtx * (gyf[j] - 0.5 * ly) * dl / ro) * CIKZ(k) * o->cu1[FI(j, k, i)] -
                                  (-1.0 * dtz * (gyf[j] - 0.5 * ly) * dl / ro) * CIKX(i) * o->cu1[FI(j, k, i)] -
                                  (-1.0 * dtz * dl / ro) * u2m;
            o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] -
                                  (dtx * (gyf[j] - 0.5 * ly) * dl / ro) * CIKZ(k) * o->cu3[FI(j, k, i)] -
                                  (-1.0 * dtz * (gyf[j] - 0.5 * ly) * dl / ro) * CIKX(i) * o->cu3[FI(j, k, i)] -
                                  (dtx * dl / ro) * u2m;
          }
      for (int j = 2; j <= ny; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++)
            o->cf2[FI(j, k, i)] = o->cf2[FI(j, k, i)] -
                                  (dtx * (o->gy[j] - 0.5 * ly) * dl / ro) * CIKZ(k) * o->cu2[FI(j, k, i)] -
                                  (-1.0 * dtz * (o->gy[j] - 0.5 * ly) * dl / ro) * CIKX(i) * o->cu2[FI(j, k, i)];
      zc *cth = o->cth + n * o->nf, *cfth = o->cfth + n * o->nf;
      for (int j = c->jstart_th[n]; j <= c->jend_th[n]; j++)
        for (int k = 0; k < kzr; k++)
          for (int i = 0; i < nxp; i++)
            cfth[FI(j, k, i)] = cfth[FI(j, k, i)] -
                                (dl / ro) * dtx * (gyf[j] - 0.5 * ly) * CIKZ(k) * cth[FI(j, k, i)] -
                                (dl / ro) * (-1.0) * dtz * (gyf[j] - 0.5 * ly) * CIKX(i) * cth[FI(j, k, i)];
    }
  }
  if (les_on) {
    /* les_chan + les_chan_th (611-623): SGS terms into cf, state to physical space */
    les_amd(o, time);
  } else {
    /* state to physical space (635-643) */
    to_physical(o, o->cu1, o->u1);
    to_physical(o, o->cu2, o->u2);
    to_physical(o, o->cu3, o->u3);
    for (int n = 0; n < nth; n++) to_physical(o, o->cth + n * o->nf, o->th + n * o->np);
  }

  const double *u1 = o->u1, *u2 = o->u2, *u3 = o->u3;
  double *s1 = o->s1;
  zc *cs1 = o->cs1;
#define LOOP_P(jlo, jhi) \
  for (int j = (jlo); j <= (jhi); j++) for (int z = 0; z < nz; z++) for (int x = 0; x < nx; x++)
#define LOOP_F(jlo, jhi) \
  for (int j = (jlo); j <= (jhi); j++) for (int k = 0; k < kzr; k++) for (int i = 0; i < nxp; i++)
  /* nonlinear terms (655-815) */
  LOOP_P(js, je) s1[PI(j, z, x)] = u3[PI(j, z, x)] * u1[PI(j, z, x)];
  to_fourier(o, s1, cs1);
  LOOP_F(js, je) {
    o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] - CIKZ(k) * cs1[FI(j, k, i)];
    o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] - CIKX(i) * cs1[FI(j, k, i)];
  }
  LOOP_P(js, je) s1[PI(j, z, x)] = u1[PI(j, z, x)] * u1[PI(j, z, x)];
  to_fourier(o, s1, cs1);
  LOOP_F(js, je) o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] - CIKX(i) * cs1[FI(j, k, i)];
  LOOP_P(js, je) s1[PI(j, z, x)] = u3[PI(j, z, x)] * u3[PI(j, z, x)];
  to_fourier(o, s1, cs1);
  LOOP_F(js, je) o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] - CIKZ(k) * cs1[FI(j, k, i)];
  LOOP_P(2, ny) s1[PI(j, z, x)] =
      ((dyf[j] * u1[PI(j, z, x)] + dyf[j - 1] * u1[PI(j - 1, z, x)]) / (2. * dy[j])) * u2[PI(j, z, x)];
  to_fourier(o, s1, cs1);
  LOOP_F(2, ny) o->cf2[FI(j, k, i)] = o->cf2[FI(j, k, i)] - CIKX(i) * cs1[FI(j, k, i)];
  LOOP_P(2, ny) s1[PI(j, z, x)] =
      ((dyf[j] * u3[PI(j, z, x)] + dyf[j - 1] * u3[PI(j - 1, z, x)]) / (2. * dy[j])) * u2[PI(j, z, x)];
  to_fourier(o, s1, cs1);
  LOOP_F(2, ny) o->cf2[FI(j, k, i)] = o->cf2[FI(j, k, i)] - CIKZ(k) * cs1[FI(j, k, i)];
  LOOP_P(js, je) s1[PI(j, z, x)] =
      (u1[PI(j + 1, z, x)] * u2[PI(j + 1, z, x)] + u1[PI(j, z, x)] * u2[PI(j + 1, z, x)] -
       u1[PI(j, z, x)] * u2[PI(j, z, x)] - u1[PI(j - 1, z, x)] * u2[PI(j, z, x)]) / (2.0 * dyf[j]);
  to_fourier(o, s1, cs1);
  LOOP_F(js, je) o->cf1[FI(j, k, i)] = o->cf1[FI(j, k, i)] - cs1[FI(j, k, i)];
  LOOP_P(js, je) s1[PI(j, z, x)] =
      (u3[PI(j + 1, z, x)] * u2[PI(j + 1, z, x)] + u3[PI(j, z, x)] * u2[PI(j + 1, z, x)] -
       u3[PI(j, z, x)] * u2[PI(j, z, x)] - u3[PI(j - 1, z, x)] * u2[PI(j, z, x)]) / (2.0 * dyf[j]);
  to_fourier(o, s1, cs1);
  LOOP_F(js, je) o->cf3[FI(j, k, i)] = o->cf3[FI(j, k, i)] - cs1[FI(j, k, i)];
  LOOP_P(2, ny) {
    const double ap = u2[PI(j, z, x)] + u2[PI(j + 1, z, x)];
    const double am = u2[PI(j, z, x)] + u2[PI(j - 1, z, x)];
    s1[PI(j, z, x)] = (0.25 * (ap * ap) - 0.25 * (am * am)) / dy[j];
  }
  to_fourier(o, s1, cs1);
  LOOP_F(2, ny) o->cf2[FI(j, k, i)] = o->cf2[FI(j, k, i)] - cs1[FI(j, k, i)];

  /* cr += temp5*cf (824-838), cr -> r (841-843) */
  LOOP_F(js, je) {
    o->cr1[FI(j, k, i)] = o->cr1[FI(j, k, i)] + temp5 * o->cf1[FI(j, k, i)];
    o->cr3[FI(j, k, i)] = o->cr3[FI(j, k, i)] + temp5 * o->cf3[FI(j, k, i)];
  }
  LOOP_F(2, ny) o->cr2[FI(j, k, i)] = o->cr2[FI(j, k, i)] + temp5 * o->cf2[FI(j, k, i)];
  to_physical(o, o->cr1, o->r1);
  to_physical(o, o->cr2, o->r2);
  to_physical(o, o->cr3, o->r3);
  /* explicit half of CN (850-870) and eddy viscosity (875-903) */
  const double nvs = c->nu_v_scale;
  double *r1 = o->r1, *r2 = o->r2, *r3 = o->r3;
  const double *nt = o->nu_t;
  LOOP_P(js, je) {
    const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
    r1[p] = r1[p] + temp1 * nvs * (((u1[pu] - u1[p]) / dy[j + 1] - (u1[p] - u1[pd]) / dy[j]) / dyf[j]);
    r3[p] = r3[p] + temp1 * nvs * (((u3[pu] - u3[p]) / dy[j + 1] - (u3[p] - u3[pd]) / dy[j]) / dyf[j]);
  }
  LOOP_P(2, ny) {
    const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
    r2[p] = r2[p] + temp1 * nvs * (((u2[pu] - u2[p]) / dyf[j] - (u2[p] - u2[pd]) / dyf[j - 1]) / dy[j]);
  }
  if (c->use_les) {
    LOOP_P(js, je) {
      const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
      r1[p] = r1[p] + temp2 * ((nt[pu] * (u1[pu] - u1[p]) / dy[j + 1] - nt[p] * (u1[p] - u1[pd]) / dy[j]) / dyf[j]);
      r3[p] = r3[p] + temp2 * ((nt[pu] * (u3[pu] - u3[p]) / dy[j + 1] - nt[p] * (u3[p] - u3[pd]) / dy[j]) / dyf[j]);
    }
    LOOP_P(2, ny) {
      const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
      r2[p] = r2[p] + temp2 * ((0.5 * (nt[p] + nt[pu]) * (u2[pu] - u2[p]) / dyf[j] -
                                0.5 * (nt[p] + nt[pd]) * (u2[p] - u2[pd]) / dyf[j - 1]) / dy[j]);
    }
  }
  /* scalars: products, RHS, CN, solve (909-1066) */
  for (int n = 0; n < nth; n++) {
    const int jst = c->jstart_th[n], jet = c->jend_th[n];
    double *th = o->th + n * o->np, *rth = o->rth + n * o->np;
    zc *cfth = o->cfth + n * o->nf, *crth = o->crth + n * o->nf;
    const double *kt = o->kappa_t + n * o->np;
    LOOP_P(jst, jet) s1[PI(j, z, x)] = th[PI(j, z, x)] * u1[PI(j, z, x)];
    to_fourier(o, s1, cs1);
    LOOP_F(jst, jet) cfth[FI(j, k, i)] = cfth[FI(j, k, i)] - CIKX(i) * cs1[FI(j, k, i)];
    LOOP_P(jst, jet) s1[PI(j, z, x)] = th[PI(j, z, x)] * u3[PI(j, z, x)];
    to_fourier(o, s1, cs1);
    LOOP_F(jst, jet) cfth[FI(j, k, i)] = cfth[FI(j, k, i)] - CIKZ(k) * cs1[FI(j, k, i)];
    LOOP_P(jst, jet) s1[PI(j, z, x)] =
        (th[PI(j + 1, z, x)] * u2[PI(j + 1, z, x)] + th[PI(j, z, x)] * u2[PI(j + 1, z, x)] -
         th[PI(j, z, x)] * u2[PI(j, z, x)] - th[PI(j - 1, z, x)] * u2[PI(j, z, x)]) / (2.0 * dyf[j]);
    to_fourier(o, s1, cs1);
    LOOP_F(jst, jet) cfth[FI(j, k, i)] = cfth[FI(j, k, i)] - cs1[FI(j, k, i)];
    LOOP_F(jst, jet) crth[FI(j, k, i)] = crth[FI(j, k, i)] + temp5 * cfth[FI(j, k, i)];
    to_physical(o, crth, rth);
    LOOP_P(jst, jet) {
      const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
      rth[p] = rth[p] + (temp1 / c->pr[n]) * nvs *
                            (((th[pu] - th[p]) / dy[j + 1] - (th[p] - th[pd]) / dy[j]) / dyf[j]);
    }
    if (c->use_les) {
      LOOP_P(jst, jet) {
        const size_t p = PI(j, z, x), pu = PI(j + 1, z, x), pd = PI(j - 1, z, x);
        rth[p] = rth[p] + temp2 * ((kt[pu] * (th[pu] - th[p]) / dy[j + 1] - kt[p] * (th[p] - th[pd]) / dy[j]) / dyf[j]);
      }
    }
    cn_solve_phys(o, 3 + n, temp1, temp2, rth, th, jst - 1, jet + 1, c->use_les ? kt : NULL, time);
  }
  /* velocity solves: u2 (1081-1121), u1 (1136-1178), u3 (1192-1233) */
  cn_solve_phys(o, 1, temp1, temp2, r2, o->u2, 1, ny + 1, c->use_les ? nt : NULL, time);
  cn_solve_phys(o, 0, temp1, temp2, r1, o->u1, js - 1, je + 1, c->use_les ? nt : NULL, time);
  cn_solve_phys(o, 2, temp1, temp2, r3, o->u3, js - 1, je + 1, c->use_les ? nt : NULL, time);
  /* courant on the new velocity, before the projection (1241-1244) */
  if (o->variable_dt && rk_step == 3 && time_step % o->update_dt == 0) {
    o->cfl_dt = courant_dt(o);
    o->cfl_ran = 1;
  }
  /* back to Fourier space (1247-1252), projection (1257), pressure (1261-1267) */
  to_fourier(o, o->u1, o->cu1);
  to_fourier(o, o->u2, o->cu2);
  to_fourier(o, o->u3, o->cu3);
  for (int n = 0; n < nth; n++) to_fourier(o, o->th + n * o->np, o->cth + n * o->nf);
  rem_div(o);
  LOOP_F(js, je) o->cp[FI(j, k, i)] = o->cp[FI(j, k, i)] + o->cr1[FI(j, k, i)] / temp4;
#undef LOOP_P
#undef LOOP_F
#undef CIKX
#undef CIKZ
#undef KPOW
  return 0;
}
