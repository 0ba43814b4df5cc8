// dgpu_les.hpp — kernels of the LES branch of rk_chan_1 (use_LES after the
// spin-up gate of for/solver.f90:615), les_model_type 5: the anisotropic
// minimum-dissipation (AMD) closure of les_chan / les_chan_th (for/les.f90).
//
// The reference forms 27 spectral gradient fields, transforms each to physical
// space, evaluates the model, and transforms 9 stress fields back (for/les.f90:
// 551-901, 1201-1395).  Here:
//  * the state's ghost planes get the BC posts of les_chan (k_bc_post);
//  * x-spectra of u, du/dz (k_inv_z with the ikz option) give, by one inverse x
//    transform each, u, du/dx (ikx option) and du/dz in physical space: 3 fields
//    per transported variable instead of 27 transforms (every strain, rotation
//    and invariant tensor entry is a pointwise combination of these, and
//    d/dy is the reference's own difference stencil applied in physical space);
//  * k_amd evaluates nu_t, kappa_t and the SGS stresses per point and writes them
//    as additions to the nonlinear product fields whose spectral derivatives the
//    reference applies to the stress (tau_11 joins u1u1, tau_13 joins u1u3, ...),
//    so the stress divergence rides on the product transforms of k_xprod;
//  * the Crank–Nicolson solve with nu_t runs per physical (x, z) column
//    (k_cn_phys), as in the reference, since nu_t varies in x and z.
// Transforms are linear, so moving a linear operation across one changes only
// round-off; the placement of every non-linear operation is the reference's.
#pragma once
#include "dgpu_kernels.hpp"

namespace dgpu {

constexpr int MAXA = 7 + 2 * MAXTH;  // SGS addition fields
constexpr double C_AMD = 0.2887;     // for/les.f90:25
constexpr double C_SMAG = 0.17;      // for/les.f90:24
constexpr double BETA_SGS = 3.0;     // for/les.f90:43

// ---------------------------------------------------------------------------
// apply_BC_vel_mpi_post + apply_BC_th_mpi_post at one rank (for/flow.f90:
// 1161-1360, 1596-1732), then apply_BC_les / apply_BC_th_les (for/les.f90:
// 1980-2079), on the Fourier state.  One thread per (kx, kz) column; the mean
// mode is column 0.
// ---------------------------------------------------------------------------
struct BcArgs {
  Geo g;
  int nth, les;  // les: also apply_BC_les / apply_BC_th_les
  int u_lo, v_lo, w_lo, u_hi, v_hi, w_hi, th_lo[MAXTH], th_hi[MAXTH];
  double u_lo_c1, v_lo_c1, w_lo_c1, u_hi_c1, v_hi_c1, w_hi_c1, w_hi_tr, th_lo_c1[MAXTH], th_hi_c1[MAXTH];
  double dy1, dy2, dyny, dyf1;
  cd* u[MAXF];
};

__global__ __launch_bounds__(256) void k_bc_post(BcArgs a) {
  const Geo& g = a.g;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= g.pc) return;
  if (c % g.kxp >= g.kxr) return;
  const bool mean = (c == 0);
  const int ny = g.ny;
  auto at = [&](cd* f, int j) -> cd& { return f[(long long)j * g.pc + c]; };
  const cd z0 = {0.0, 0.0};
  // GYF fields u1, u3 and every scalar share the form of the lower/upper posts
  auto gyf_lo = [&](cd* f, int bc, double c1) {
    if (bc == DGPU_BC_DIRICHLET) {
      at(f, 0) = z0; at(f, 1) = z0;
      if (mean) { at(f, 1) = cd{c1, 0.0}; at(f, 0) = cd{c1, 0.0}; }
    } else {
      at(f, 1) = at(f, 2); at(f, 0) = at(f, 1);
      if (mean) { at(f, 1) = at(f, 2) - cd{a.dy2 * c1, 0.0}; at(f, 0) = at(f, 1) - cd{a.dy1 * c1, 0.0}; }
    }
  };
  auto gyf_hi = [&](cd* f, int bc, double c1d, double c1n, bool w) {
    if (bc == DGPU_BC_DIRICHLET) {
      at(f, ny) = z0; at(f, ny + 1) = z0;
      if (mean) { at(f, ny) = cd{c1d, 0.0}; at(f, ny + 1) = cd{c1d, 0.0}; }
      if (w && mean) at(f, ny + 1) = z0;  // for/flow.f90:1311
    } else {
      at(f, ny) = at(f, ny - 1); at(f, ny + 1) = at(f, ny);
      if (mean) {
        at(f, ny) = at(f, ny - 1) + cd{a.dyny * c1n, 0.0};
        at(f, ny + 1) = at(f, ny) + cd{a.dyny * c1n, 0.0};
      }
    }
  };
  cd *u1 = a.u[0], *u2 = a.u[1], *u3 = a.u[2];
  // lower: u1, u3, then u2
  gyf_lo(u1, a.u_lo, a.u_lo_c1);
  gyf_lo(u3, a.w_lo, a.w_lo_c1);
  if (a.v_lo == DGPU_BC_DIRICHLET) {
    at(u2, 1) = cd{2.0 * a.v_lo_c1, 0.0} - at(u2, 2);
    at(u2, 0) = at(u2, 1);
  } else {
    at(u2, 1) = at(u2, 2); at(u2, 0) = at(u2, 1);
    if (mean) {
      at(u2, 1) = at(u2, 2) - cd{a.dyf1 * a.v_lo_c1, 0.0};
      at(u2, 0) = at(u2, 1) - cd{a.dyf1 * a.v_lo_c1, 0.0};
    }
  }
  // upper: u1, u3, then u2
  gyf_hi(u1, a.u_hi, a.u_hi_c1, a.u_hi_c1, false);
  gyf_hi(u3, a.w_hi, a.w_hi_c1, a.w_hi_tr, true);
  if (a.v_hi == DGPU_BC_DIRICHLET) {
    if (mean) at(u2, ny + 1) = cd{2.0 * a.v_hi_c1, 0.0} - at(u2, ny);
  } else {
    at(u2, ny) = at(u2, ny - 1); at(u2, ny + 1) = at(u2, ny);
    if (mean) at(u2, ny + 1) = at(u2, ny) + cd{a.dyny * a.v_hi_c1, 0.0};
  }
  for (int n = 0; n < a.nth; ++n) {
    gyf_lo(a.u[3 + n], a.th_lo[n], a.th_lo_c1[n]);
    gyf_hi(a.u[3 + n], a.th_hi[n], a.th_hi_c1[n], a.th_hi_c1[n], false);
  }
  if (!a.les) return;
  // apply_BC_les: Neumann ghosts copy the first interior plane, mean included
  if (a.u_hi == DGPU_BC_NEUMANN) { at(u1, ny) = at(u1, ny - 1); at(u1, ny + 1) = at(u1, ny); }
  if (a.w_hi == DGPU_BC_NEUMANN) { at(u3, ny) = at(u3, ny - 1); at(u3, ny + 1) = at(u3, ny); }
  if (a.u_lo == DGPU_BC_NEUMANN) { at(u1, 1) = at(u1, 2); at(u1, 0) = at(u1, 1); }
  if (a.w_lo == DGPU_BC_NEUMANN) { at(u3, 1) = at(u3, 2); at(u3, 0) = at(u3, 1); }
  for (int n = 0; n < a.nth; ++n) {
    cd* t = a.u[3 + n];
    if (a.th_hi[n] == DGPU_BC_NEUMANN) { at(t, ny) = at(t, ny - 1); at(t, ny + 1) = at(t, ny); }
    if (a.th_lo[n] == DGPU_BC_NEUMANN) { at(t, 1) = at(t, 2); at(t, 0) = at(t, 1); }
  }
}

// ---------------------------------------------------------------------------
// AMD model per point (for/les.f90:571-797, 1213-1346).
// Inputs (physical, [j][z][x], all planes): for each transported field v in
// (u1, u2, u3, th_n): v, dv/dx, dv/dz.  Outputs, on every plane (zero where the
// reference adds nothing):
//   nu_t, kappa_t_n            at GY points, zero at j = 0, 1, 2, Nyp, Nyp+1
//   A0 = tau_11  (joins u1u1)       A1 = tau_13 (joins u1u3)   rows jstart..Nyp-1
//   A2 = tau_33  (joins u3u3)       A3 = (tau_12(j+1)-tau_12(j))/dyf(j)  (u1 flux)
//   A4 = (tau_23(j+1)-tau_23(j))/dyf(j)  (u3 flux)                        rows jstart..Nyp-1
//   A5 = tau_12 - nu_t du1/dy (joins u1 u2 at GY)   A6 = tau_23 - nu_t du3/dy   rows 2..Nyp
//   A7+2n = flux_x of th_n (joins th u1)  A8+2n = flux_z (joins th u3)  rows jstart_th..jend_th
// with tau including the filter-width factor delta(j) of the reference.
// ---------------------------------------------------------------------------
struct AmdArgs {
  Geo g;
  int nth, jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  double dxs, dzs;                      // dx * beta_sgs, dz * beta_sgs
  double zx, xz;                        // dzs / dxs, dxs / dzs
  const double* __restrict__ rowc;      // [ny2][RC_COUNT] per-row grid quotients
  const double* __restrict__ dy;
  const double* __restrict__ dyf;
  const double* __restrict__ dlt;       // [2][ny2]: delta_yf(j), delta_y(j)
  const void* f[3 * MAXF];              // v, dv/dx, dv/dz per transported field
  void* nut;
  void* kap[MAXTH];
  void* aug[MAXA];
  // les_model_type 6: the learned closure's coefficients at GY (k_closure), in
  // the transform precision T, sign bit set where the point took the AMD value
  const void* cgy;   // velocity
  const void* kgy;   // scalar
};

// the invariant-scaled strain/rotation entries of one GYF plane
template <typename R>
struct Gyf {
  R a1, a2, a3, a5, o5;
};
// ... and of one GY plane
template <typename R>
struct Gy {
  R a4, o4, a6, o6;
};

// The inputs one point reads: every slot on rows j - 1 .. j + 2, loaded into
// registers before any arithmetic (the reads of the model's branches and of the
// neighbour rows are then all in flight at once instead of in dependent groups).
// at() is called with rows j + const, so the window index folds to a constant.
template <typename T>
struct AmdIn {
  T w[3 * MAXF][4];
  int j0;          // j - 1
  long long pp;
  __device__ T at(int f, int row, long long) const { return w[f][row - j0]; }
};
// NS slots; rows outside 0..Nyp+1 are clamped (no caller uses such a row's value)
template <typename T, int NS>
__device__ __forceinline__ void load_window(AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  const int last = a.g.ny + 1;
  in.j0 = j - 1;
  in.pp = a.g.pp;
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int row = j - 1 + r;
      row = row < 0 ? 0 : (row > last ? last : row);
      in.w[s][r] = ((const T*)a.f[s])[(long long)row * a.g.pp + q];
    }
}

// The reference divides by dy(j), dyf(j) and the filter widths at every point;
// the row table holds those quotients (RowCoef), so the kernels multiply (the
// results agree to rounding).
__device__ __forceinline__ const double* rq(const AmdArgs& a, int j) { return a.rowc + (long long)j * RC_COUNT; }
// the GY interpolation of GYF values c (row j) and m (row j-1):
// (c dyf(j-1) + m dyf(j)) / (2 dy(j))
template <typename R>
__device__ __forceinline__ R gy_interp(const AmdArgs& a, int j, R c, R m) {
  const double* rc = rq(a, j);
  return c * R(rc[RC_WTH]) + m * R(rc[RC_WTHM]);
}

// field slots: 3*f + {0: v, 1: dv/dx, 2: dv/dz}, f = 0 u1, 1 u2, 2 u3, 3+n th_n
template <typename T>
__device__ __forceinline__ Gyf<T> gyf_plane(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  Gyf<T> r;
  const R du1dz = in.at(2, j, q), du3dx = in.at(7, j, q);
  r.a1 = in.at(1, j, q);                                              // SI11 = du1/dx
  r.a2 = (in.at(3, j + 1, q) - in.at(3, j, q)) * R(rq(a, j)[RC_RDYF]);    // SI22 = du2/dy
  r.a3 = in.at(8, j, q);                                              // SI33 = du3/dz
  r.a5 = R(0.5) * (du1dz * R(a.zx) + du3dx * R(a.xz));                // SI13
  r.o5 = R(0.5) * (du1dz * R(a.zx) - du3dx * R(a.xz));                // O13
  return r;
}
template <typename T>
__device__ __forceinline__ Gy<T> gy_plane(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  Gy<T> r;
  // the width ratios 2 dy / (dx beta), ... of the scaled gradient, from the row table
  const double* rc = rq(a, j);
  const R rdy = R(rc[RC_RDY]), yx = R(rc[RC_YX]), xy = R(rc[RC_XY]), yz = R(rc[RC_YZ]), zy = R(rc[RC_ZY]);
  const R du1dy = (in.at(0, j, q) - in.at(0, j - 1, q)) * rdy;
  const R du3dy = (in.at(6, j, q) - in.at(6, j - 1, q)) * rdy;
  const R du2dx = in.at(4, j, q), du2dz = in.at(5, j, q);
  r.a4 = R(0.5) * (du1dy * yx + du2dx * xy);
  r.o4 = R(0.5) * (du1dy * yx - du2dx * xy);
  r.a6 = R(0.5) * (du2dz * zy + du3dy * yz);
  r.o6 = R(0.5) * (du2dz * zy - du3dy * yz);
  return r;
}

// max{0, -(I3 - I4)} / (I1 - I2) from the six strain and three rotation entries
template <typename R>
__device__ __forceinline__ R amd_coef(R a1, R a2, R a3, R a4, R a5, R a6,
                                           R o4, R o5, R o6) {
  R v = a1 * a1 * a1 + a2 * a2 * a2 + a3 * a3 * a3 + R(3.0) * a1 * (a4 * a4) + R(3.0) * a1 * (a5 * a5) +
             R(3.0) * a2 * (a4 * a4) + R(3.0) * a2 * (a6 * a6) + R(3.0) * a3 * (a5 * a5) + R(3.0) * a3 * (a6 * a6) +
             R(6.0) * a4 * a5 * a6;
  v = -v - a1 * (o4 * o4) - a1 * (o5 * o5) - a2 * (o4 * o4) - a2 * (o6 * o6) - a3 * (o5 * o5) -
      a3 * (o6 * o6) - R(2.0) * a4 * o5 * o6 + R(2.0) * a5 * o4 * o6 - R(2.0) * a6 * o4 * o5;
  if (v <= R(0.0)) return R(0.0);
  return v / (a1 * a1 + a2 * a2 + a3 * a3 + R(2.0) * (a4 * a4) + R(2.0) * (o4 * o4) + R(2.0) * (a5 * a5) +
              R(2.0) * (o5 * o5) + R(2.0) * (a6 * a6) + R(2.0) * (o6 * o6));
}

// coefficient at GYF j (GY entries averaged from j and j+1)
template <typename T>
__device__ __forceinline__ T coef_gyf(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  const Gyf<T> f = gyf_plane(in, a, j, q);
  const Gy<T> p = gy_plane(in, a, j + 1, q), m = gy_plane(in, a, j, q);
  return amd_coef(f.a1, f.a2, f.a3, R(0.5) * (p.a4 + m.a4), f.a5, R(0.5) * (p.a6 + m.a6), R(0.5) * (p.o4 + m.o4), f.o5,
                  R(0.5) * (p.o6 + m.o6));
}
// coefficient at GY j (GYF entries interpolated from j-1 and j)
template <typename T>
__device__ __forceinline__ T coef_gy(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  const Gyf<T> f = gyf_plane(in, a, j, q), fm = gyf_plane(in, a, j - 1, q);
  const Gy<T> p = gy_plane(in, a, j, q);
  auto w = [&](R c, R m) { return gy_interp(a, j, c, m); };
  return amd_coef(w(f.a1, fm.a1), w(f.a2, fm.a2), w(f.a3, fm.a3), p.a4, w(f.a5, fm.a5), p.a6, p.o4,
                  w(f.o5, fm.o5), p.o6);
}

// The strain S and rotation O of the unscaled gradient (compute_strain_chan,
// compute_rotation_chan, for/les.f90:1443-1596), per GYF and GY plane, for types 1, 4
template <typename R>
struct Sgyf {
  R s1, s2, s3, s5, o5;
};
template <typename R>
struct Sgy {
  R s4, s6, o4, o6;
};
template <typename T>
__device__ __forceinline__ Sgyf<T> sgyf_plane(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  const R du1dz = in.at(2, j, q), du3dx = in.at(7, j, q);
  return {in.at(1, j, q), (in.at(3, j + 1, q) - in.at(3, j, q)) * R(rq(a, j)[RC_RDYF]), in.at(8, j, q),
          R(0.5) * (du1dz + du3dx), R(0.5) * (du1dz - du3dx)};
}
template <typename T>
__device__ __forceinline__ Sgy<T> sgy_plane(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  const R rdy = R(rq(a, j)[RC_RDY]);
  const R du1dy = (in.at(0, j, q) - in.at(0, j - 1, q)) * rdy;
  const R du3dy = (in.at(6, j, q) - in.at(6, j - 1, q)) * rdy;
  const R du2dx = in.at(4, j, q), du2dz = in.at(5, j, q);
  return {R(0.5) * (du1dy + du2dx), R(0.5) * (du2dz + du3dy), R(0.5) * (du1dy - du2dx), R(0.5) * (du2dz - du3dy)};
}
// type 4 (for/les.f90:276-440): the AMD polynomial with the squared filter widths
// dX, dY, dZ as weights; at GYF (g6 = false) the last cubic term multiplies
// a4 a5 a6, at GY a4 a6 a5 (the reference's operand orders)
template <typename R>
__device__ __forceinline__ R amd4_coef(R dX, R dY, R dZ, R a1, R a2, R a3,
                                            R a4, R a5, R a6, R o4, R o5, R o6,
                                            bool gy) {
  R v = dX * (a1 * a1 * a1) + dY * (a2 * a2 * a2) + dZ * (a3 * a3 * a3) + (R(2.0) * dX + dY) * a1 * (a4 * a4) +
             (R(2.0) * dX + dZ) * a1 * (a5 * a5) + (R(2.0) * dY + dX) * a2 * (a4 * a4) + (R(2.0) * dY + dZ) * a2 * (a6 * a6) +
             (R(2.0) * dZ + dX) * a3 * (a5 * a5) + (R(2.0) * dZ + dY) * a3 * (a6 * a6) +
             (gy ? R(2.0) * (dX + dY + dZ) * a4 * a6 * a5 : R(2.0) * (dX + dY + dZ) * a4 * a5 * a6);
  v = -v - dY * a1 * (o4 * o4) - dZ * a1 * (o5 * o5) - dX * a2 * (o4 * o4) - dZ * a2 * (o6 * o6) -
      dX * a3 * (o5 * o5) - dY * a3 * (o6 * o6) - R(2.0) * dZ * a4 * o5 * o6 + R(2.0) * dY * a5 * o4 * o6 -
      R(2.0) * dX * a6 * o4 * o5;
  if (v <= R(0.0)) return R(0.0);
  return v / (a1 * a1 + a2 * a2 + a3 * a3 + R(2.0) * (a4 * a4) + R(2.0) * (o4 * o4) + R(2.0) * (a5 * a5) +
              R(2.0) * (o5 * o5) + R(2.0) * (a6 * a6) + R(2.0) * (o6 * o6));
}

// the model coefficient at GYF j (s1 of les_chan) and at GY j (temp)
template <int MODEL, typename T>
__device__ __forceinline__ T model_gyf(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  if constexpr (MODEL == 5) {
    return coef_gyf(in, a, j, q);
  } else if constexpr (MODEL == 6) {
    // the mean of the network's GY j and j+1 values, else the type-5 GYF coefficient
    const R c0 = (R)((const T*)a.cgy)[(long long)j * in.pp + q];
    const R c1 = (R)((const T*)a.cgy)[(long long)(j + 1) * in.pp + q];
    return (!signbit(c0) && !signbit(c1)) ? R(0.5) * (c0 + c1) : coef_gyf(in, a, j, q);
  } else {
    const Sgyf<T> f = sgyf_plane(in, a, j, q);
    const Sgy<T> p = sgy_plane(in, a, j + 1, q), m = sgy_plane(in, a, j, q);
    const R s4 = R(0.5) * (p.s4 + m.s4), s6 = R(0.5) * (p.s6 + m.s6);
    if constexpr (MODEL == 1) {  // |S| (for/les.f90:99-112)
      return sqrt(R(2.0) * (f.s1 * f.s1) + R(4.0) * (s4 * s4) + R(4.0) * (f.s5 * f.s5) + R(2.0) * (f.s2 * f.s2) +
                  R(4.0) * (s6 * s6) + R(2.0) * (f.s3 * f.s3));
    } else {
      const R dY = (R(a.dy[j]) * R(2.0)) * (R(a.dy[j]) * R(2.0));
      return amd4_coef(R(a.dxs) * R(a.dxs), dY, R(a.dzs) * R(a.dzs), f.s1, f.s2, f.s3, s4, f.s5, s6, R(0.5) * (p.o4 + m.o4),
                       f.o5, R(0.5) * (p.o6 + m.o6), false);
    }
  }
}
template <int MODEL, typename T>
__device__ __forceinline__ T model_gy(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  if constexpr (MODEL == 5) {
    return coef_gy(in, a, j, q);
  } else if constexpr (MODEL == 6) {
    return fabs((R)((const T*)a.cgy)[(long long)j * in.pp + q]);
  } else {
    const Sgyf<T> f = sgyf_plane(in, a, j, q), fm = sgyf_plane(in, a, j - 1, q);
    const Sgy<T> p = sgy_plane(in, a, j, q);
    auto w = [&](R c, R m) { return gy_interp(a, j, c, m); };
    const R s1 = w(f.s1, fm.s1), s2 = w(f.s2, fm.s2), s3 = w(f.s3, fm.s3), s5 = w(f.s5, fm.s5);
    if constexpr (MODEL == 1) {  // (for/les.f90:114-132)
      return sqrt(R(2.0) * (s1 * s1) + R(4.0) * (p.s4 * p.s4) + R(4.0) * (s5 * s5) + R(2.0) * (s2 * s2) + R(4.0) * (p.s6 * p.s6) +
                  R(2.0) * (s3 * s3));
    } else {
      const R dY = (R(a.dy[j]) * R(2.0)) * (R(a.dy[j]) * R(2.0));
      return amd4_coef(R(a.dxs) * R(a.dxs), dY, R(a.dzs) * R(a.dzs), s1, s2, s3, p.s4, s5, p.s6, p.o4, w(f.o5, fm.o5), p.o6,
                       true);
    }
  }
}

// The type-5 scalar coefficient max{0, v}/|b'|^2 (les_chan_th), with the scaled
// scalar gradient b' = (delta_x th_x, delta_y th_y, delta_z th_z) and the
// unscaled th_y it was formed from
template <typename R>
struct ThCoef {
  R c, bx, by, bz, ty;
};
// at GY j (for/les.f90:1259-1306)
template <typename T>
__device__ __forceinline__ ThCoef<T> amd_th_gy(const AmdIn<T>& in, const AmdArgs& a, int ft, int j, long long q) {
  using R = T;
  const R dxs2 = R(a.dxs) * R(a.dxs), dzs2 = R(a.dzs) * R(a.dzs);
  const R dyl = R(a.dy[j]) * R(2.0), dyl2 = dyl * dyl;
  const double* rc = rq(a, j);
  const R rdy = R(rc[RC_RDY]);
  auto w = [&](int s) { return gy_interp(a, j, in.at(s, j, q), in.at(s, j - 1, q)); };
  const R du1dx = w(1), du3dz = w(8), du1dz = w(2), du3dx = w(7);
  const R du2dy = gy_interp(a, j, (in.at(3, j + 1, q) - in.at(3, j, q)) * R(rc[RC_RDYF]),
                            (in.at(3, j, q) - in.at(3, j - 1, q)) * R(rc[RC_RDYFM]));
  const R tx = w(ft + 1), tz = w(ft + 2);
  const R ty = (in.at(ft, j, q) - in.at(ft, j - 1, q)) * rdy;
  const R du1dy = (in.at(0, j, q) - in.at(0, j - 1, q)) * rdy;
  const R du3dy = (in.at(6, j, q) - in.at(6, j - 1, q)) * rdy;
  const R du2dx = in.at(4, j, q), du2dz = in.at(5, j, q);
  const R v = -(dxs2 * du1dx * (tx * tx) + dyl2 * du2dy * (ty * ty) + dzs2 * du3dz * (tz * tz) +
                     (dyl2 * du1dy + dxs2 * du2dx) * tx * ty + (dzs2 * du1dz + dxs2 * du3dx) * tx * tz +
                     (dzs2 * du2dz + dyl2 * du3dy) * ty * tz);
  const R ax = R(a.dxs) * tx, ay = dyl * ty, az = R(a.dzs) * tz;
  return {v > R(0.0) ? v / (ax * ax + ay * ay + az * az) : R(0.0), ax, ay, az, ty};
}
// at GYF j (for/les.f90:1221-1256, 1310-1318)
template <typename T>
__device__ __forceinline__ ThCoef<T> amd_th_gyf(const AmdIn<T>& in, const AmdArgs& a, int ft, int j, long long q) {
  using R = T;
  const R dxs2 = R(a.dxs) * R(a.dxs), dzs2 = R(a.dzs) * R(a.dzs);
  const R dyl = R(a.dyf[j]) * R(2.0), dyl2 = dyl * dyl;
  auto gyv = [&](int s, int jj) { return (in.at(s, jj, q) - in.at(s, jj - 1, q)) * R(rq(a, jj)[RC_RDY]); };
  const R tx = in.at(ft + 1, j, q), tz = in.at(ft + 2, j, q);
  const R ty = R(0.5) * (gyv(ft, j + 1) + gyv(ft, j));
  const R du1dx = in.at(1, j, q), du3dz = in.at(8, j, q);
  const R du2dy = (in.at(3, j + 1, q) - in.at(3, j, q)) * R(rq(a, j)[RC_RDYF]);
  const R du1dy = R(0.5) * (gyv(0, j + 1) + gyv(0, j)), du3dy = R(0.5) * (gyv(6, j + 1) + gyv(6, j));
  const R du2dx = R(0.5) * (in.at(4, j + 1, q) + in.at(4, j, q));
  const R du2dz = R(0.5) * (in.at(5, j + 1, q) + in.at(5, j, q));
  const R du1dz = in.at(2, j, q), du3dx = in.at(7, j, q);
  const R v = -(dxs2 * du1dx * (tx * tx) + dyl2 * du2dy * (ty * ty) + dzs2 * du3dz * (tz * tz) +
                     (dyl2 * du1dy + dxs2 * du2dx) * tx * ty + (dzs2 * du1dz + dxs2 * du3dx) * tx * tz +
                     (dzs2 * du2dz + dyl2 * du3dy) * ty * tz);
  const R ax = R(a.dxs) * tx, ay = dyl * ty, az = R(a.dzs) * tz;
  return {v > R(0.0) ? v / (ax * ax + ay * ay + az * az) : R(0.0), ax, ay, az, ty};
}

// tau_12 (e = 0) or tau_23 (e = 1) at GY j: delta_y(j) coef(j) (S - du/dy / 2).
// At j = 1 types 4 and 5 scale the strain itself (their coefficient loops start
// at j = 2, for/les.f90:454, 762), type 1 multiplies it (for/les.f90:145); that
// row feeds only rows the BC rows overwrite.
// coef: model_gy at row j (unused on the row-1 special case).
template <int MODEL, typename T>
__device__ __forceinline__ T tau_gy(const AmdIn<T>& in, const AmdArgs& a, int e, int j, long long q, T coef) {
  using R = T;
  const R dlt = R(a.dlt[a.g.ny2 + j]);
  if (MODEL != 1 && j == 1) {
    const R dvdy = (in.at(e ? 6 : 0, 1, q) - in.at(e ? 6 : 0, 0, q)) * R(rq(a, 1)[RC_RDY]);
    return dlt * R(0.5) * (dvdy + in.at(e ? 5 : 4, 1, q));
  }
  return dlt * (coef * (R(0.5) * in.at(e ? 5 : 4, j, q)));
}

// Block b of a 1D grid over (256-point block qb, row j) pairs.  The dispatcher
// deals blocks to the 8 XCDs in turn, so with nqb a multiple of 8 XCD b % 8
// takes q-blocks 8 g + b % 8 (its pass g) through the rows in order, and the
// neighbour rows j - 1 .. j + 2 a block reads are in that XCD's L2 from the
// blocks of the rows around it.
__device__ __forceinline__ void xcd_rows(int b, int nqb, int nj, int& qb, int& j) {
  if (nqb % 8) {
    qb = b % nqb;
    j = b / nqb;
    return;
  }
  const int t = b >> 3;
  j = t % nj;
  qb = (t / nj) * 8 + (b & 7);
}

template <typename T, int NTH, int MODEL>
__global__ __launch_bounds__(256) void k_amd(AmdArgs a) {
  using R = T;
  const Geo& g = a.g;
  int qb, j;
  xcd_rows(blockIdx.x, (int)((g.pp + 255) / 256), g.ny2, qb, j);
  const long long q = (long long)qb * 256 + threadIdx.x;
  if (q >= g.pp) return;
  AmdIn<T> in;
  load_window<T, 3 * (3 + NTH)>(in, a, j, q);
  const int ny = g.ny;
  const long long o = (long long)j * g.pp + q;
  auto put = [&](void* p, R v) { ((T*)p)[o] = (T)v; };
  // the model coefficient at GY j and j + 1, each evaluated once: nu_t, the two
  // stresses at GY j and the GYF flux differences take them (row 1 of types 4
  // and 5 needs none)
  const bool need_j = j <= ny && (j >= 2 || (MODEL == 1 && j == 1));
  const R cj = need_j ? model_gy<MODEL>(in, a, j, q) : R(0.0);
  const R cj1 = (j >= a.jstart && j <= ny - 1) ? model_gy<MODEL>(in, a, j + 1, q) : R(0.0);
  const double* rc = rq(a, j);
  // nu_t at GY j (zero on the wall rows of ghost_les_mpi)
  R nut = R(0.0);
  if (j >= 3 && j <= ny - 1) nut = -R(0.5) * R(a.dlt[g.ny2 + j]) * cj;
  put(a.nut, nut);
  // GYF stresses, rows j1..j2 = jstart..Nyp-1
  R A0 = R(0.0), A1 = R(0.0), A2 = R(0.0), A3 = R(0.0), A4 = R(0.0);
  if (j >= a.jstart && j <= ny - 1) {
    const R dyf = R(a.dlt[j]) * model_gyf<MODEL>(in, a, j, q);
    const R du1dz = in.at(2, j, q), du3dx = in.at(7, j, q);
    A0 = dyf * in.at(1, j, q);
    A1 = dyf * (R(0.5) * (du1dz + du3dx));
    A2 = dyf * in.at(8, j, q);
    const R rdyf = R(rc[RC_RDYF]);
    A3 = (tau_gy<MODEL>(in, a, 0, j + 1, q, cj1) - tau_gy<MODEL>(in, a, 0, j, q, cj)) * rdyf;
    A4 = (tau_gy<MODEL>(in, a, 1, j + 1, q, cj1) - tau_gy<MODEL>(in, a, 1, j, q, cj)) * rdyf;
  }
  put(a.aug[0], A0); put(a.aug[1], A1); put(a.aug[2], A2); put(a.aug[3], A3); put(a.aug[4], A4);
  // GY stresses and the explicit cross terms, rows 2..Nyp
  R A5 = R(0.0), A6 = R(0.0);
  if (j >= 2 && j <= ny) {
    const R rdy = R(rc[RC_RDY]);
    A5 = tau_gy<MODEL>(in, a, 0, j, q, cj) - nut * ((in.at(0, j, q) - in.at(0, j - 1, q)) * rdy);
    A6 = tau_gy<MODEL>(in, a, 1, j, q, cj) - nut * ((in.at(6, j, q) - in.at(6, j - 1, q)) * rdy);
  }
  put(a.aug[5], A5); put(a.aug[6], A6);
  if constexpr (MODEL == 1 || MODEL == 4) {
    // les_chan_th types 1 and 4 (for/les.f90:1118-1193): kappa_t = nu_t (GY), and
    // cfth += ikx F(kappa_t dth/dx) + ikz F(kappa_t dth/dz) on rows jstart_th..jend_th
#pragma unroll
    for (int n = 0; n < NTH; ++n) {
      const int ft = 3 * (3 + n);
      put(a.kap[n], nut);
      const bool row = (j >= a.jstart_th[n] && j <= a.jend_th[n]);
      put(a.aug[7 + 2 * n], row ? -(nut * in.at(ft + 1, j, q)) : R(0.0));
      put(a.aug[8 + 2 * n], row ? -(nut * in.at(ft + 2, j, q)) : R(0.0));
    }
    return;
  }
  // scalars (les_chan_th type 5; type 6 reads the learned coefficients)
#pragma unroll
  for (int n = 0; n < NTH; ++n) {
    const int ft = 3 * (3 + n);
    R kap = R(0.0), fx = R(0.0), fz = R(0.0);
    // kappa_t at GY j (for/les.f90:1259-1306, 1344-1348)
    if (j >= 3 && j <= ny - 1) {
      R c;
      if constexpr (MODEL == 6) c = fabs((R)((const T*)a.kgy)[o]);
      else c = amd_th_gy(in, a, ft, j, q).c;
      kap = -R(0.5) * R(a.dlt[g.ny2 + j]) * c;
    }
    put(a.kap[n], kap);
    // SGS flux at GYF j (for/les.f90:1221-1256, 1310-1318, 1372-1393)
    if (j >= a.jstart_th[n] && j <= a.jend_th[n]) {
      R s1;
      if constexpr (MODEL == 6) {
        const R k0 = (R)((const T*)a.kgy)[o], k1 = (R)((const T*)a.kgy)[o + g.pp];
        s1 = (!signbit(k0) && !signbit(k1)) ? R(0.5) * (k0 + k1) : amd_th_gyf(in, a, ft, j, q).c;
      } else {
        s1 = amd_th_gyf(in, a, ft, j, q).c;
      }
      const R f = R(0.5) * R(a.dlt[j]) * s1;
      fx = f * in.at(ft + 1, j, q);
      fz = f * in.at(ft + 2, j, q);
    }
    put(a.aug[7 + 2 * n], fx);
    put(a.aug[8 + 2 * n], fz);
  }
}

// ---------------------------------------------------------------------------
// Crank–Nicolson solve per physical (x, z) column with the eddy viscosity /
// diffusivity (rk_chan_1, for/solver.f90:875-903, 990-1233): the explicit half
// of the nu_t term is added to the right-hand side r (the inverse transform of
// the Fourier right-hand side, which already holds the molecular part), the
// matrix gets the implicit half, the BC rows of apply_BC_* are set for every
// column, and Thomas (thomas_forward/backward_real_mpi at one rank,
// for/solver.f90:1691-1761, 1816-1852) solves rows 0..Nyp+1.
// One thread per (column, field): the fields' systems are independent, so
// blockIdx.y picks the field.  The thread keeps a register window of u and nu_t
// over rows j-1..j+1 and has the loads of the next rows in flight while row j
// is eliminated (the elimination chain is serial in j; the loads are not).
// The reciprocal pivot 1/b' and g' go to scratch; the back substitution reads
// them in reverse with the same look-ahead and recomputes c from nu_t (the
// same row function, so the same value).
//
// INC (the mixed-precision path): u is the state at the substep's start in
// float and r holds r - u (float).  The solve is then for the increment
// delta = u_new - u: right-hand side rhs - (l u_{j-1} + d u_j + c u_{j+1}) of the
// same rows, eliminated in FP64, and delta is written back to r (0 outside
// [wlo, whi]), so the state moves by a float transform of an O(dt) quantity.
// The scratch 1/b', g' of the increment system is stored in float (g' is an
// increment too; the rounding is of the order of the float transforms').
// Otherwise the solution overwrites u on rows [wlo, whi].
// ---------------------------------------------------------------------------
template <typename TU, typename TV>
struct PhysCnArgs {
  Geo g;
  int nf;
  int field[MAXF];                    // slot -> field (0 u1, 1 u2, 2 u3, 3+n th_n)
  int blo[MAXF], bhi[MAXF], wlo[MAXF], whi[MAXF];
  int lo_kind[MAXF], hi_kind[MAXF];   // BC code at each wall (0 Dirichlet, 1 Neumann)
  double lo_c1[MAXF], hi_c1[MAXF];
  double tcn[MAXF];                   // temp1 nu_v_scale (/Pr for scalars)
  double temp2;
  const double* __restrict__ dy;
  const double* __restrict__ dyf;
  const double* __restrict__ rowc;    // [ny2][RC_COUNT] per-row grid coefficients
  TU* u[MAXF];                        // physical field
  TU* r[MAXF];                        // right-hand side (INC: r - u, then delta)
  const TV* visc[MAXF];               // nu_t (velocity) or kappa_t_n (scalar)
  double* cp;                         // scratch, nf * 2 * ny2 planes: 1/b', g'
};

template <bool INC>
struct ScratchType { using type = double; };
template <>
struct ScratchType<true> { using type = float; };

template <typename TU, typename TV, bool INC>
__global__ __launch_bounds__(256) void k_cn_phys(PhysCnArgs<TU, TV> a) {
  const Geo& g = a.g;
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const int e = blockIdx.y;
  if (q >= g.pp || e >= a.nf) return;
  const int ny = g.ny, n2 = g.ny2, last = ny + 1;
  const long long pp = g.pp;
  const double* __restrict__ dy = a.dy;
  const double* __restrict__ dyf = a.dyf;
  const double* __restrict__ rowc = a.rowc;
  TU* __restrict__ u = a.u[e];
  TU* __restrict__ r = a.r[e];
  const TV* __restrict__ nt = a.visc[e];
  // scratch: FP64 for the FP64 solve, float for the increments (below)
  using TS = typename ScratchType<INC>::type;
  TS* __restrict__ sb = (TS*)a.cp + (long long)e * 2 * n2 * pp;  // 1/b'
  TS* __restrict__ sg = sb + (long long)n2 * pp;                 // g'
  const double T = a.tcn[e], t2 = a.temp2;
  const int fld = a.field[e];
  const bool v2 = (fld == 1);
  const int blo = a.blo[e], bhi = a.bhi[e], wlo = a.wlo[e], whi = a.whi[e];
  // row j of the system: (l, d, c | rhs) from the window values
  auto row = [&](int j, double um, double uc, double up, double nm, double nc, double np, double rv, double& l,
                 double& d, double& c, double& rhs) {
    l = 0.0; d = 1.0; c = 0.0; rhs = 0.0;
    if (j >= blo && j <= bhi) {
      rhs = INC ? rv + uc : rv;
      // The reference's rows (for/solver.f90:1136-1145 and the nu_t terms of
      // les_chan) divide T and temp2 nu by dy(j) dyf(j) and dy(j+1) dyf(j)
      // (GYF rows), dyf(j-1) dy(j) and dyf(j) dy(j) (u2's GY rows, nu_t averaged
      // to the faces); the host table holds those reciprocals, so a row costs
      // multiplications only (the quotients agree to rounding).
      const double* k = rowc + (long long)j * RC_COUNT;
      const double nlo = v2 ? 0.5 * (nc + nm) : nc, nhi = v2 ? 0.5 * (nc + np) : np;
      const double ilo = k[v2 ? RC_GLO : RC_FLO], ihi = k[v2 ? RC_GHI : RC_FHI];
      const double lo = (T + t2 * nlo) * ilo, hi = (T + t2 * nhi) * ihi;
      l = -lo;
      c = -hi;
      d = 1.0 + lo + hi;
      rhs = rhs + t2 * (nhi * (up - uc) * ihi - nlo * (uc - um) * ilo);
    }
    // BC rows (apply_BC_* of the reference, for/solver.f90:1901-2232)
    if (!v2) {
      if (j == 0) { l = 0.; d = 1.; c = 0.; rhs = 0.; }
      if (j == 1) {
        if (a.lo_kind[e] == DGPU_BC_DIRICHLET) { l = 0.; d = 1.; c = 0.; rhs = a.lo_c1[e]; }
        else { l = 0.; d = -1.; c = 1.; rhs = dy[2] * a.lo_c1[e]; }
      }
      if (j == ny) {
        if (a.hi_kind[e] == DGPU_BC_DIRICHLET) { l = 0.; d = 1.; c = 0.; rhs = a.hi_c1[e]; }
        else { l = -1.; d = 1.; c = 0.; rhs = dy[ny] * a.hi_c1[e]; }
      }
      if (j == ny + 1) { l = 0.; d = 1.; c = 0.; rhs = 0.; }
    } else {
      if (a.lo_kind[e] == DGPU_BC_DIRICHLET) {
        if (j == 1 || j == 2) { l = 0.; d = 1.; c = 0.; rhs = a.lo_c1[e]; }
      } else if (a.lo_kind[e] == DGPU_BC_NEUMANN) {
        if (j == 1) { l = 0.; d = -1.; c = 1.; rhs = dyf[1] * a.lo_c1[e]; }
      }
      if (j == 0) { l = 0.; d = 1.; c = 0.; rhs = 0.; }
      if (a.hi_kind[e] == DGPU_BC_DIRICHLET) {
        if (j == ny + 1 || j == ny) { l = 0.; d = 1.; c = 0.; rhs = a.hi_c1[e]; }
      } else if (a.hi_kind[e] == DGPU_BC_NEUMANN) {
        if (j == ny + 1) { l = -1.; d = 1.; c = 0.; rhs = dyf[ny] * a.hi_c1[e]; }
      }
    }
    if (INC) rhs = rhs - (l * um + d * uc + c * up);
  };
  // forward elimination; the loads run L + 1 rows (u, nu_t) and L rows (r)
  // ahead of the elimination, held raw in small register queues (float rows
  // carry half the bytes per load, so the increment solve keeps more in flight)
  constexpr int L = INC ? 4 : 2;
  auto LU = [&](const TU* p, int j) -> TU { return (j >= 0 && j <= last) ? p[(long long)j * pp + q] : (TU)0; };
  auto LV = [&](int j) -> TV { return (j >= 0 && j <= last) ? nt[(long long)j * pp + q] : (TV)0; };
  double um = 0.0, uc = (double)LU(u, 0), up = (double)LU(u, 1);
  double nm = 0.0, nc = (double)LV(0), np = (double)LV(1);
  double rv = (double)LU(r, 0);
  TU qu[L], qr[L];
  TV qv[L];
#pragma unroll
  for (int i = 0; i < L; ++i) { qu[i] = LU(u, i + 2); qv[i] = LV(i + 2); qr[i] = LU(r, i + 1); }
  double binv = 0.0, cprev = 0.0, gprev = 0.0;
  for (int j = 0; j <= last; ++j) {
    // here up = u(j+1), qu[i] = u(j+2+i) (nu_t alike), rv = r(j), qr[i] = r(j+1+i)
    double l, d, c, rhs;
    row(j, um, uc, up, nm, nc, np, rv, l, d, c, rhs);
    if (j > 0) {
      const double m = -l * binv;
      d = d + m * cprev;
      rhs = rhs + m * gprev;
    }
    binv = 1.0 / d;
    const long long o = (long long)j * pp + q;
    sb[o] = (TS)binv;
    sg[o] = (TS)rhs;
    cprev = c; gprev = rhs;
    um = uc; uc = up; up = (double)qu[0];
    nm = nc; nc = np; np = (double)qv[0];
    rv = (double)qr[0];
#pragma unroll
    for (int i = 0; i < L - 1; ++i) { qu[i] = qu[i + 1]; qv[i] = qv[i + 1]; qr[i] = qr[i + 1]; }
    qu[L - 1] = LU(u, j + L + 2); qv[L - 1] = LV(j + L + 2); qr[L - 1] = LU(r, j + L + 1);
  }
  // back substitution, rows last..0; c of row j from nu_t at j, j+1
  auto LS = [&](const TS* p, int j) -> TS { return j >= 0 ? p[(long long)j * pp + q] : (TS)0; };
  auto put = [&](int j, double v) {
    const long long o = (long long)j * pp + q;
    if (INC) r[o] = (TU)((j >= wlo && j <= whi) ? v : 0.0);
    else if (j >= wlo && j <= whi) u[o] = (TU)v;
  };
  double x = gprev * binv;
  put(last, x);
  double bc = (double)LS(sb, last - 1), gc = (double)LS(sg, last - 1);
  double vp = (double)LV(last), vc = (double)LV(last - 1);
  TS qb[L], qg[L];
#pragma unroll
  for (int i = 0; i < L; ++i) { qb[i] = LS(sb, last - 2 - i); qg[i] = LS(sg, last - 2 - i); qv[i] = LV(last - 2 - i); }
  for (int j = last - 1; j >= 0; --j) {
    // here bc, gc = row j's 1/b', g'; qb[i] = row j-1-i's (nu_t alike: vc = nu(j), vp = nu(j+1))
    double l, d, c, rhs;
    row(j, 0.0, 0.0, 0.0, 0.0, vc, vp, 0.0, l, d, c, rhs);
    x = (gc - c * x) * bc;
    put(j, x);
    bc = (double)qb[0]; gc = (double)qg[0];
    vp = vc; vc = (double)qv[0];
#pragma unroll
    for (int i = 0; i < L - 1; ++i) { qb[i] = qb[i + 1]; qg[i] = qg[i + 1]; qv[i] = qv[i + 1]; }
    qb[L - 1] = LS(sb, j - 1 - L); qg[L - 1] = LS(sg, j - 1 - L); qv[L - 1] = LV(j - 1 - L);
  }
}

}  // namespace dgpu
