/* dgpu_io.h — flow checkpoints and statistics series in the reference's HDF5
 * formats, for one rank (NprocY = NprocZ = 1), from the host.
 *
 * libdgpu_io.so (HDF5 C API, serial) sits beside libdiablo_gpu.so.  The file
 * layer (dgpu_h5_*) works on host arrays only; dgpu_write_flow / dgpu_read_flow
 * move the device state through it with the device transforms of
 * libdiablo_gpu.so, as WriteHDF5 / ReadHDF5 do with fft_xz_to_physical /
 * fft_xz_to_fourier.
 *
 * Physical arrays are in the reference's layout real(8) (0:Nx+1, 0:Nzp+1,
 * 0:Nyp+1), column-major (dgpu_field_doubles doubles each); scalar arrays th
 * hold N_th such arrays back to back.
 *
 * Every function returns 0 on success, non-zero on failure (then
 * dgpu_io_last_error() says why), where the reference would stop.
 */
#ifndef DGPU_IO_H
#define DGPU_IO_H

#include <stdint.h>
#include "diablo_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The module variables WriteHDF5 stores in group Timestep and ReadHDF5
 * restores (for/phdf5.f90:1049-1076, 1307-1329). */
typedef struct dgpu_flow_times {
  double time, save_flow_time, save_stats_time, save_movie_time;
  int32_t time_step;
} dgpu_flow_times;

/* ---- file layer (host arrays, no device) ---- */

/* WriteHDF5's file (for/phdf5.f90:918-1197) from physical fields: attributes
 * Resolution (Nx, Ny, Nz) and Date (20 characters, the format of time_string,
 * for/tools.f90:155-211; date20 = NULL takes the local time), group Timestep
 * with the times of t, datasets U (u1), W (u2f), V (u3), TH1..THn (th) and, if p
 * is not NULL, P: each the swapzy (for/tools.f90:216-232) of its field, FP64,
 * chunked one y plane per chunk.  u2f is u2 already on the fractional grid (the
 * g2gf of u2, as the reference writes it). */
int dgpu_h5_write_flow(const char *fname, int32_t nx, int32_t ny, int32_t nz, int32_t nth,
                       const double *u1, const double *u2f, const double *u3, const double *th,
                       const double *p, const dgpu_flow_times *t, const char *date20);

/* ReadHDF5's file reading (for/phdf5.f90:1204-1421): checks Resolution (a
 * mismatch is an error, where the reference stops), reads the Timestep times
 * into t and each dataset through swapyz (for/tools.f90:235-251: the arrays are
 * zeroed, then rows 1..Nyp of x < Nx, z < Nzp filled).  Scalar n is read only
 * if bit n of th_mask is set (the reference skips create_new_th scalars; their
 * arrays are left untouched).  p is filled and *has_p set to 1 if the file holds
 * P, else *has_p = 0 and p is untouched.  u2f receives W as stored (GYF). */
int dgpu_h5_read_flow(const char *fname, int32_t nx, int32_t ny, int32_t nz, int32_t nth,
                      uint32_t th_mask, double *u1, double *u2f, double *u3, double *th, double *p,
                      int32_t *has_p, dgpu_flow_times *t);

/* g2gf and gf2g at one rank (for/tools.f90:299-315 and 255-295), in place on a
 * physical array: u2 from the GY points to the fractional grid and back. */
void dgpu_g2gf(double *var, int32_t nx, int32_t ny, int32_t nz);
void dgpu_gf2g(double *var, int32_t nx, int32_t ny, int32_t nz);

/* WriteStatH5_Y (for/phdf5.f90:129-278) and WriteStatH5_X (285-431), which
 * write the same layout at one rank: appends the n values of diag (Nyp for a y
 * profile, Nbin for an x or z* profile) as dataset "%04d" (the next sample
 * number) of group gname, with attribute Time, and updates the group's SAMPLES
 * counter; the file is created if absent. */
int dgpu_h5_append_profile(const char *fname, const char *gname, const double *diag, int32_t ny,
                           double time);

/* WriteHDF5_XYplane, WriteHDF5_XZplane, WriteHDF5_ZYplane (for/phdf5.f90:438-589,
 * 596-740, 745-896) at one rank: the same series form for a plane var2d(n1, n2)
 * in Fortran order (n1 fastest), stored with C dimensions (n2, n1): (Nx, Nyp)
 * for an x-y plane, (Nx, Nz) for x-z, (Nz, Nyp) for z-y. */
int dgpu_h5_append_plane(const char *fname, const char *gname, const double *var2d, int32_t n1, int32_t n2,
                         double time);

/* WriteHDF5_real (for/phdf5.f90:7-128): the same series form for one value. */
int dgpu_h5_append_real(const char *fname, const char *gname, double value);

/* The grid file as proc/init/create_grid.m:84-87 writes it and ReadGridHDF5
 * (for/domain.f90:420-500) reads it: group /grids, dataset y of n = Ny + 1
 * doubles G(2..Ny+2) (diablo3_amd.grid.create_grid computes them).  The file is
 * created or truncated. */
int dgpu_h5_write_grid(const char *fname, const double *gy, int32_t n);
/* /grids/y of a grid file: *n receives its length, gy (may be NULL) the values. */
int dgpu_h5_read_grid(const char *fname, double *gy, int32_t *n);

/* ---- with the device state ---- */

/* WriteHDF5 (for/phdf5.f90:918-1197) of the handle's state: the fields are
 * transformed to physical space on the device, u2 is interpolated to the
 * fractional grid (g2gf), the file is written, and — as in the reference —
 * u2 is mapped back with gf2g and every field transformed back to Fourier
 * space into the state, so the state afterwards is the reference's (its u2
 * rows 1 and Nyp+1 replaced by gf2g's zero-mass-flux values).  With
 * save_pressure the pressure goes to P and makes the same round trip. */
int dgpu_write_flow(dgpu_handle *h, const char *fname, const dgpu_flow_times *t,
                    int32_t save_pressure);

/* ReadHDF5 (for/phdf5.f90:1204-1456) into the handle's state: scalars whose bit
 * in th_mask is clear keep their state; without P in the file the pressure is
 * recomputed from the velocity as the reference does (poisson_p_chan,
 * for/solver.f90:1382-1560, on the device: dgpu_poisson_p) and
 * *read_pressure = 0.  ghost_chan_mpi is empty at one rank. */
int dgpu_read_flow(dgpu_handle *h, const char *fname, uint32_t th_mask, dgpu_flow_times *t,
                   int32_t *read_pressure);

/* Text of the last error of this library on this thread ("" if none). */
const char *dgpu_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DGPU_IO_H */
