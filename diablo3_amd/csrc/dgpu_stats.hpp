// dgpu_stats.hpp — energy-budget statistics of save_stats_chan on the device
// (for/diagnostics.f90:3-669, compute_averages 676-799, compute_TKE_diss 804-1092,
// compute_MKE_diss 1097-1249, compute_TKE 1254-1360, compute_TKE_Production
// 1366-1519), all in FP64.
//
// Inputs are the physical fields v, dv/dx, dv/dz of u1, u2, u3 and each scalar
// (the LES gradient layout, [j][z][x]) and the Fourier state (mean modes).
//  k_stat_mean   per (x, j): the z-line means of u1, u2, u3, th_n and of du_i/dx
//                (ume_xy, vme_xy, wme_xy, thme_xy, dudx_m, dvdx_m, dwdx_m); for an
//                x-homogeneous flow the plane means and zero gradients instead
//  k_stat_zsum   per (x, j): the z sums of every fluctuation product
//  k_stat_plane  per j: the x reductions and the x-only (mean-field) terms
//  k_stat_xz     the z-mean (x, y) planes of mean_xz.h5 (movie output)
//  k_stat_movie  the x-y, x-z and z-y slices of movie.h5
// The z and x sums replace the reference's sequential loops (round-off only).
#pragma once
#include "dgpu_kernels.hpp"

namespace dgpu {

// z-line mean slots
enum { SM_U1 = 0, SM_U2, SM_U3, SM_DXU1, SM_DXU2, SM_DXU3, SM_TH0 };
// z-sum slots
// (SZ_THU: th_1' u1'; SZ_EPSM: the per-x mean-field dissipation integrand, written
// by k_stat_plane)
enum { SZ_UU = 0, SZ_VV, SZ_WW, SZ_UW, SZ_UV, SZ_WV, SZ_EPS, SZ_SHEAR, SZ_U1, SZ_OX, SZ_OY, SZ_OZ,
       SZ_THU, SZ_EPSM, SZ_TH0 };
constexpr int SZ_PER_TH = 3;  // (th')^2, th' w', |grad th|^2
// output profiles (include/diablo_gpu.h: dgpu_stats)
enum {
  SP_UME = 0, SP_VME, SP_WME, SP_MKE, SP_URMS, SP_VRMS, SP_WRMS, SP_UW, SP_UV, SP_WV,
  SP_UU_DUDX, SP_VV_DVDY, SP_WU_DWDX, SP_VU_DVDX, SP_UV_DUDY, SP_WV_DWDY,
  SP_DUDY, SP_DWDY, SP_SHEAR, SP_EPS, SP_EPSM, SP_CP, SP_U1Y, SP_OMX, SP_OMY, SP_OMZ, SP_NBASE,
  SP_PER_TH = 6  // thme, thrms, thv, thv_m, dthdy, pe_diss
};

struct StatArgs {
  Geo g;
  int nth, hx;
  double nu, twf, ly, lx;
  double dthdx[MAXTH], dthdz[MAXTH];
  const double* __restrict__ dy;
  const double* __restrict__ dyf;
  const double* __restrict__ gy;
  const double* __restrict__ gyf;
  const double* f[3 * MAXF];  // v, dv/dx, dv/dz per field (u1, u2, u3, th_n)
  const cd* u[MAXF];          // Fourier state (mean modes)
  const cd* p;
  double* zm;                 // [6 + nth][ny2][nx]
  double* zs;                 // [SZ_TH0 + 3 nth][ny2][nx]
  double* prof;               // [SP_NBASE + 6 nth][ny2]
};

__global__ __launch_bounds__(256) void k_stat_mean(StatArgs a) {
  const Geo& g = a.g;
  const int x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (x >= g.nx) return;
  const long long pl = (long long)g.ny2 * g.nx;
  const long long o = (long long)j * g.nx + x;
  const int nm = 6 + a.nth;
  const int ny = g.ny;
  // slot -> (field slot in f, valid row range); vme_xy runs to Nyp+1
  for (int m = 0; m < nm; ++m) {
    double v = 0.0;
    const int fs = m < 3 ? 3 * m : (m < 6 ? 3 * (m - 3) + 1 : 3 * (3 + m - 6));
    const int jhi = (m == SM_U2) ? ny + 1 : ny;
    if (j >= 1 && j <= jhi) {
      if (a.hx) {
        // plane means from the mean mode; the x derivatives of the mean are zero
        // (compute_MKE_diss returns before setting them, for/diagnostics.f90:1117)
        if (m < 3 || m >= 6) {
          const cd* C = m < 3 ? a.u[m] : a.u[3 + m - 6];
          v = C[(long long)j * g.pc].x;
        }
      } else {
        const double* __restrict__ F = a.f[fs] + (long long)j * g.pp + x;
        for (int z = 0; z < g.nz; ++z) v += F[(long long)z * g.nx];
        v = v / (double)g.nz;
      }
    }
    a.zm[m * pl + o] = v;
  }
}

// The per-point quantities of save_stats_chan at (x, j, z), from the physical
// fields and the z-line means; shared by the z sums, the z* bins and the movie
// slices.
struct StatPoint {
  const StatArgs& a;
  const int x;
  const long long pl, pp;
  __device__ StatPoint(const StatArgs& a_, int x_)
      : a(a_), x(x_), pl((long long)a_.g.ny2 * a_.g.nx), pp(a_.g.pp) {}
  __device__ double M(int m, int j) const { return a.zm[m * pl + (long long)j * a.g.nx + x]; }
  __device__ double F(int s, int j, int z) const { return a.f[s][(long long)j * pp + (long long)z * a.g.nx + x]; }
  // the dissipation integrand of compute_TKE_diss on GY rows 2..Nyp, zero elsewhere
  // (for/diagnostics.f90:820-1016); when x is not homogeneous the kz = 0 row of
  // each x derivative (k_start = 1), i.e. its z-line mean, is left out
  __device__ double eps(int j, int z) const {
    if (j < 2 || j > a.g.ny) return 0.0;
    const double* dy = a.dy;
    const double* dyf = a.dyf;
    const double mdx1 = a.hx ? 0.0 : M(SM_DXU1, j), mdx1m = a.hx ? 0.0 : M(SM_DXU1, j - 1);
    const double mdx2 = a.hx ? 0.0 : M(SM_DXU2, j);
    const double mdx3 = a.hx ? 0.0 : M(SM_DXU3, j), mdx3m = a.hx ? 0.0 : M(SM_DXU3, j - 1);
    const double f1 = F(0, j, z) - M(SM_U1, j), f3 = F(6, j, z) - M(SM_U3, j);
    const double f1m = F(0, j - 1, z) - M(SM_U1, j - 1), f3m = F(6, j - 1, z) - M(SM_U3, j - 1);
    auto itp = [&](double c, double m) { return (dyf[j - 1] * c * c + dyf[j] * m * m) / (2.0 * dy[j]); };
    double e = itp(F(1, j, z) - mdx1, F(1, j - 1, z) - mdx1m);
    const double a2 = F(4, j, z) - mdx2;
    e += a2 * a2;
    const double t3 = (f1 - f1m) / dy[j];
    e += t3 * t3;
    e += itp(F(7, j, z) - mdx3, F(7, j - 1, z) - mdx3m);
    e += itp(F(2, j, z), F(2, j - 1, z));
    const double t6 = ((F(3, j + 1, z) - M(SM_U2, j + 1)) - (F(3, j - 1, z) - M(SM_U2, j - 1))) /
                      (a.gy[j + 1] - a.gy[j - 1]);
    e += t6 * t6;
    const double t7 = (f3 - f3m) / dy[j];
    e += t7 * t7;
    const double a8 = F(5, j, z);
    e += a8 * a8;
    e += itp(F(8, j, z), F(8, j - 1, z));
    return e;
  }
  // compute_Vorticity (for/diagnostics.f90:1534-1727): the three components at a
  // GYF row, omega_x with the thermal-wind shear removed
  __device__ void omega(int j, int z, double& ox, double& oy, double& oz) const {
    const double r2 = 1.0 / (2.0 * a.dyf[j]);
    ox = 0.5 * (F(5, j + 1, z) + F(5, j, z)) - (F(6, j + 1, z) - F(6, j - 1, z)) * r2 - a.twf;
    oy = F(7, j, z) - F(2, j, z);
    oz = (F(0, j + 1, z) - F(0, j - 1, z)) * r2 - 0.5 * (F(4, j + 1, z) + F(4, j, z));
  }
  // th_n' w' at a GYF row (w averaged to GYF, for/diagnostics.f90:296-307)
  __device__ double thw(int n, int j, int z) const {
    const double ft = F(3 * (3 + n), j, z) - M(SM_TH0 + n, j);
    return ft * (0.5 * (F(3, j, z) + F(3, j + 1, z)) - 0.5 * (M(SM_U2, j) + M(SM_U2, j + 1)));
  }
  // |grad th_n + background|^2 at a GYF row; the y difference is of scalar 1 for
  // every n (crth(.., 1) in for/diagnostics.f90:147, 355-361)
  __device__ double chi(int n, int j, int z) const {
    const int ts = 3 * (3 + n);
    const double gx_ = F(ts + 1, j, z) + a.dthdx[n];
    const double gy_ = (F(9, j + 1, z) - F(9, j - 1, z)) / (a.gyf[j + 1] - a.gyf[j - 1]);
    const double gz_ = F(ts + 2, j, z) + a.dthdz[n];
    return gx_ * gx_ + gy_ * gy_ + gz_ * gz_;
  }
};

__global__ __launch_bounds__(256) void k_stat_zsum(StatArgs a) {
  const Geo& g = a.g;
  const int x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (x >= g.nx) return;
  const int ny = g.ny, nx = g.nx;
  const long long pl = (long long)g.ny2 * nx;
  const StatPoint P(a, x);
  const double* dy = a.dy;
  const double* dyf = a.dyf;
  const bool gyf_row = (j >= 1 && j <= ny), gy_row = (j >= 2 && j <= ny);
  double uu = 0, vv = 0, ww = 0, uw = 0, uv = 0, wv = 0, eps = 0, sh = 0, s1 = 0, sox = 0, soy = 0, soz = 0;
  double thu = 0;
  double tt[MAXTH] = {}, tv[MAXTH] = {}, pe[MAXTH] = {};
  if (gyf_row || gy_row) {
    const double mu1 = P.M(SM_U1, j), mu2 = P.M(SM_U2, j), mu3 = P.M(SM_U3, j);
    for (int z = 0; z < g.nz; ++z) {
      const double u1 = P.F(0, j, z), u2 = P.F(3, j, z), u3 = P.F(6, j, z);
      const double f1 = u1 - mu1, f2 = u2 - mu2, f3 = u3 - mu3;
      if (gyf_row) {
        uu += f1 * f1; vv += f2 * f2; ww += f3 * f3; uw += f1 * f3;
        const double d1 = (P.F(0, j + 1, z) - P.F(0, j - 1, z)) / (2.0 * dyf[j]);
        const double d3 = (P.F(6, j + 1, z) - P.F(6, j - 1, z)) / (2.0 * dyf[j]);
        sh += d1 * d1 + d3 * d3;
        s1 += u1;
        double ox, oy, oz;
        P.omega(j, z, ox, oy, oz);
        sox += ox * ox; soy += oy * oy; soz += oz * oz;
        for (int n = 0; n < a.nth; ++n) {
          const double ft = P.F(3 * (3 + n), j, z) - P.M(SM_TH0 + n, j);
          tt[n] += ft * ft;
          tv[n] += P.thw(n, j, z);
          pe[n] += P.chi(n, j, z);
          if (n == 0) thu += ft * f1;  // (th_1 - thme_xy) (u1 - ume_xy), for/diagnostics.f90:301-302
        }
      }
      if (gy_row) {
        const double f1m = P.F(0, j - 1, z) - P.M(SM_U1, j - 1), f3m = P.F(6, j - 1, z) - P.M(SM_U3, j - 1);
        const double iu = (dyf[j - 1] * f1 + dyf[j] * f1m) / (2.0 * dy[j]);
        const double iw = (dyf[j - 1] * f3 + dyf[j] * f3m) / (2.0 * dy[j]);
        uv += iu * f2;
        wv += iw * f2;
        eps += P.eps(j, z);
      }
    }
  }
  const long long o = (long long)j * nx + x;
  a.zs[SZ_UU * pl + o] = uu; a.zs[SZ_VV * pl + o] = vv; a.zs[SZ_WW * pl + o] = ww;
  a.zs[SZ_UW * pl + o] = uw; a.zs[SZ_UV * pl + o] = uv; a.zs[SZ_WV * pl + o] = wv;
  a.zs[SZ_EPS * pl + o] = eps; a.zs[SZ_SHEAR * pl + o] = sh; a.zs[SZ_U1 * pl + o] = s1;
  a.zs[SZ_OX * pl + o] = sox; a.zs[SZ_OY * pl + o] = soy; a.zs[SZ_OZ * pl + o] = soz;
  a.zs[SZ_THU * pl + o] = thu;
  for (int n = 0; n < a.nth; ++n) {
    a.zs[(SZ_TH0 + SZ_PER_TH * n) * pl + o] = tt[n];
    a.zs[(SZ_TH0 + SZ_PER_TH * n + 1) * pl + o] = tv[n];
    a.zs[(SZ_TH0 + SZ_PER_TH * n + 2) * pl + o] = pe[n];
  }
}

// block-wide sum of 256 values (fixed tree: the result does not depend on timing)
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

constexpr int NSTAT_X = 19 + 4 * MAXTH;

__global__ __launch_bounds__(256) void k_stat_plane(StatArgs a) {
  __shared__ double sh[256];
  const Geo& g = a.g;
  const int j = blockIdx.x, t = threadIdx.x, nx = g.nx, ny = g.ny, nth = a.nth;
  const long long pl = (long long)g.ny2 * nx;
  auto M = [&](int m, int jj, int x) { return a.zm[m * pl + (long long)jj * nx + x]; };
  auto S = [&](int s, int x) { return a.zs[s * pl + (long long)j * nx + x]; };
  const double* dy = a.dy;
  const double* dyf = a.dyf;
  const double* gy = a.gy;
  const double* gyf = a.gyf;
  const double NxNz = (double)nx * (double)g.nz, Nz = (double)g.nz, Nx = (double)nx;
  const bool gyf_row = (j >= 1 && j <= ny), gy_row = (j >= 2 && j <= ny);
  // per-thread partial sums over x of every profile term
  double acc[NSTAT_X] = {};
  for (int x = t; x < nx; x += 256) {
    const double suu = S(SZ_UU, x), svv = S(SZ_VV, x), sww = S(SZ_WW, x), suw = S(SZ_UW, x);
    const double suv = S(SZ_UV, x), swv = S(SZ_WV, x);
    acc[0] += suu; acc[1] += svv; acc[2] += sww; acc[3] += suw; acc[4] += suv; acc[5] += swv;
    acc[6] += S(SZ_EPS, x); acc[7] += S(SZ_SHEAR, x);
    acc[16] += S(SZ_OX, x); acc[17] += S(SZ_OY, x); acc[18] += S(SZ_OZ, x);
    double em = 0.0;  // this x's mean-field dissipation integrand (the epsilon_m_xz plane)
    if (gyf_row) {
      acc[8] += (suu / Nz) * M(SM_DXU1, j, x);     // uu_dudx
      acc[9] += (suw / Nz) * M(SM_DXU3, j, x);     // wu_dwdx
      // compute_MKE_diss (x not homogeneous): interpolated and GY terms
      if (!a.hx) {
        const double m1 = M(SM_DXU1, j, x), m1m = M(SM_DXU1, j - 1, x);
        const double m3 = M(SM_DXU3, j, x), m3m = M(SM_DXU3, j - 1, x);
        const double m2 = M(SM_DXU2, j, x);
        double e = (dyf[j - 1] * m1 * m1 + dyf[j] * m1m * m1m) / (2.0 * dy[j]) + m2 * m2 +
                   (dyf[j - 1] * m3 * m3 + dyf[j] * m3m * m3m) / (2.0 * dy[j]);
        if (gy_row) {
          const double t3 = (M(SM_U1, j, x) - M(SM_U1, j - 1, x)) / dy[j];
          const double t5 = (M(SM_U2, j + 1, x) - M(SM_U2, j - 1, x)) / (gy[j + 1] - gy[j - 1]);
          const double t6 = (M(SM_U3, j, x) - M(SM_U3, j - 1, x)) / dy[j] + a.twf;
          e += t3 * t3 + t5 * t5 + t6 * t6;
        }
        acc[10] += e;
        em = e;
      }
    }
    a.zs[SZ_EPSM * pl + (long long)j * nx + x] = em;
    if (gy_row) {
      acc[11] += (suv / Nz) * M(SM_DXU2, j, x);    // vu_dvdx
      acc[12] += (M(SM_U2, j + 1, x) - M(SM_U2, j - 1, x)) / (gy[j + 1] - gy[j - 1]) * (svv / Nz);
      acc[13] += (M(SM_U1, j, x) - M(SM_U1, j - 1, x)) / dy[j] * (suv / Nz);
      acc[14] += ((M(SM_U3, j, x) - M(SM_U3, j - 1, x)) / dy[j] + a.twf) * (swv / Nz);
      // mke integrand (for/diagnostics.f90:44-61), same for every z
      const double um = M(SM_U1, j, x), umm = M(SM_U1, j - 1, x);
      const double wm = M(SM_U3, j, x) + a.twf * (gyf[j] - 0.5 * a.ly);
      const double wmm = M(SM_U3, j - 1, x) + a.twf * (gyf[j - 1] - 0.5 * a.ly);
      const double vm = M(SM_U2, j, x);
      acc[15] += (dyf[j - 1] * um * um + dyf[j] * umm * umm) / (2.0 * dy[j]) +
                 (dyf[j - 1] * wm * wm + dyf[j] * wmm * wmm) / (2.0 * dy[j]) + vm * vm;
    }
    for (int n = 0; n < nth; ++n) {
      acc[19 + 4 * n] += S(SZ_TH0 + SZ_PER_TH * n, x);
      acc[20 + 4 * n] += S(SZ_TH0 + SZ_PER_TH * n + 1, x);
      acc[21 + 4 * n] += S(SZ_TH0 + SZ_PER_TH * n + 2, x);
      if (gyf_row) {
        const double gxv = ((double)x * a.lx) / (double)nx;
        acc[22 + 4 * n] += (M(SM_TH0 + n, j, x) + a.dthdx[n] * (gxv - 0.5 * a.lx)) *
                           (0.5 * (M(SM_U2, j, x) + M(SM_U2, j + 1, x)));
      }
    }
  }
  double tot[NSTAT_X];
  for (int k = 0; k < 19 + 4 * nth; ++k) tot[k] = block_sum(acc[k], sh);
  if (t != 0) return;
  const int n2 = g.ny2;
  double* P = a.prof;
  auto put = [&](int p, double v) { P[(long long)p * n2 + j] = v; };
  const double ume = a.u[0][(long long)j * g.pc].x, vme = a.u[1][(long long)j * g.pc].x;
  const double wme = a.u[2][(long long)j * g.pc].x;
  put(SP_UME, ume); put(SP_VME, vme); put(SP_WME, wme);
  put(SP_CP, a.p[(long long)j * g.pc].x);
  double urms = 0, vrms = 0, wrms = 0, uw = 0, uv = 0, wv = 0, eps = 0, epsm = 0, shr = 0, mke = 0;
  double uud = 0, vvd = 0, wud = 0, vud = 0, uvd = 0, wvd = 0, dudy = 0, dwdy = 0, u1y = 0;
  double omx = 0, omy = 0, omz = 0;
  if (gyf_row) {
    urms = sqrt(tot[0] / NxNz); vrms = sqrt(tot[1] / NxNz); wrms = sqrt(tot[2] / NxNz);
    uw = tot[3] / NxNz;
    shr = tot[7] / NxNz;
    uud = tot[8] / Nx; wud = tot[9] / Nx;
    epsm = a.hx ? 0.0 : a.nu * tot[10] / Nx;
    const double um1 = a.u[0][(long long)(j - 1) * g.pc].x, wm1 = a.u[2][(long long)(j - 1) * g.pc].x;
    dudy = (ume - um1) / (gyf[j] - gyf[j - 1]);
    dwdy = (wme - wm1) / (gyf[j] - gyf[j - 1]);
    u1y = gyf[j] * S(SZ_U1, 0) / Nz;
    omx = sqrt(tot[16] / NxNz); omy = sqrt(tot[17] / NxNz); omz = sqrt(tot[18] / NxNz);
  }
  if (gy_row) {
    uv = tot[4] / NxNz; wv = tot[5] / NxNz;
    eps = a.nu * tot[6] / NxNz;
    vud = tot[11] / Nx; vvd = tot[12] / Nx; uvd = tot[13] / Nx; wvd = tot[14] / Nx;
    mke = 0.5 * (Nz * tot[15]) / NxNz;
  }
  put(SP_URMS, urms); put(SP_VRMS, vrms); put(SP_WRMS, wrms);
  put(SP_UW, uw); put(SP_UV, uv); put(SP_WV, wv);
  put(SP_UU_DUDX, uud); put(SP_VV_DVDY, vvd); put(SP_WU_DWDX, wud); put(SP_VU_DVDX, vud);
  put(SP_UV_DUDY, uvd); put(SP_WV_DWDY, wvd);
  put(SP_DUDY, dudy); put(SP_DWDY, dwdy); put(SP_SHEAR, shr); put(SP_EPS, eps); put(SP_EPSM, epsm);
  put(SP_MKE, mke); put(SP_U1Y, u1y);
  put(SP_OMX, omx); put(SP_OMY, omy); put(SP_OMZ, omz);
  for (int n = 0; n < nth; ++n) {
    const int b = SP_NBASE + SP_PER_TH * n;
    const cd* C = a.u[3 + n];
    const double thme = C[(long long)j * g.pc].x;
    double thrms = 0, thv = 0, thvm = 0, dthdy = 0, ped = 0;
    if (gyf_row) {
      thrms = sqrt(tot[19 + 4 * n] / NxNz);
      thv = tot[20 + 4 * n] / NxNz;
      ped = tot[21 + 4 * n] / NxNz;
      thvm = tot[22 + 4 * n] / Nx;
      dthdy = (thme - C[(long long)(j - 1) * g.pc].x) / (gyf[j] - gyf[j - 1]);
    }
    put(b + 0, thme); put(b + 1, thrms); put(b + 2, thv); put(b + 3, thvm); put(b + 4, dthdy); put(b + 5, ped);
  }
}

// The source of poisson_p_chan (for/solver.f90:1390-1530) on rows 2..Nyp, in
// physical space from the gradient fields: (du1/dx)^2 + ((u2(j+1) - u2(j))/dyf)^2
// + (du3/dz)^2 + 2 du1/dy dv/dx + 2 du3/dy dv/dz + 2 du1/dz du3/dx, with the
// reference's centred y differences and v averaged to GYF.  The reference
// transforms each product separately; the sum of the transforms is the transform
// of the sum.  Other rows are zero.
__global__ __launch_bounds__(256) void k_pp_source(StatArgs a, double* __restrict__ out) {
  const Geo& g = a.g;
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (q >= g.pp) return;
  const long long pp = g.pp, o = (long long)j * pp + q;
  double v = 0.0;
  if (j >= 2 && j <= g.ny) {
    auto F = [&](int s, int jj) { return a.f[s][(long long)jj * pp + q]; };
    const double r2f = 1.0 / (2.0 * a.dyf[j]);
    const double f1 = F(1, j), f2 = (F(3, j + 1) - F(3, j)) / a.dyf[j], f3 = F(8, j);
    const double g1 = (F(0, j + 1) - F(0, j - 1)) * r2f, g2 = 0.5 * (F(4, j) + F(4, j + 1));
    const double g3 = (F(6, j + 1) - F(6, j - 1)) * r2f, h2 = 0.5 * (F(5, j) + F(5, j + 1));
    const double h1 = F(2, j), h3 = F(7, j);
    v = f1 * f1 + f2 * f2 + f3 * f3 + 2.0 * g1 * g2 + 2.0 * g3 * h2 + 2.0 * h1 * h3;
  }
  out[o] = v;
}

// FTx_uu (for/diagnostics.f90:641-663): per kx mode i < Nx/2 the y integral
// (trapezoid over rows 1..Nyp) of |u1(i, kz = 0, j)|^2, over Ly; modes past
// Nkx are zero in the state
__global__ __launch_bounds__(256) void k_stat_ftx(StatArgs a, double* out, int nxp) {
  const Geo& g = a.g;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nxp) return;
  double v = 0.0;
  if (i < g.kxr) {
    const cd* U = a.u[0];
    for (int j = 2; j <= g.ny; ++j) {
      const cd c = U[(long long)j * g.pc + i], cm = U[(long long)(j - 1) * g.pc + i];
      const double p = c.x * c.x + c.y * c.y, pm = cm.x * cm.x + cm.y * cm.y;
      v = v + 0.5 * (p + pm) * a.dy[j];
    }
  }
  out[i] = v / a.ly;
}

// ---------------------------------------------------------------------------
// z* statistics: the buoyancy-surface binning of compute_BPE and
// Bin_Ystar_and_Write (for/diagnostics.f90:1733-1852, 1859-1953) for scalar 1.
// Each point of rows jstart_th(1)..jend_th(1) falls in the bin of
// s1 = th + dTHdX(1) (gx - Lx/2) (gx = x Lx / Nx): with x homogeneous the
// uniform bins of [thmin, thmin + Nbin dTH); else bin = int(acos(-clip(s1) /
// (Ro_inv + 1e-14)) Nbin / pi), pi the reference's single-precision value
// (for/fft.f90:147).  Slot 0 accumulates dyf(j) (the PDF), slot 1 + f the point
// field f times dyf(j): shear2 (for/diagnostics.f90:113-117), epsilon
// (nu times the dissipation integrand, 843-1016, rows 2..Nyp), TKE (1273-1275),
// KE (1354), and per scalar n th' w' (305-307) and |grad th + background|^2
// (359-361).  The fields are those k_stat_zsum sums for the profiles.  One block
// per (256 x, row): an LDS histogram (FP64 LDS atomics), added to the global one
// at the end (so the order of the additions, and the last bits, vary run to run).
struct ZstarArgs {
  int nbin, nf, jlo, jhi, hx;
  double thmin, thmax, dth, den, pi, dthdx0, lx;
  double* hist;  // [nf][nbin]
};

__global__ __launch_bounds__(256) void k_stat_zstar(StatArgs a, ZstarArgs z) {
  extern __shared__ double hs[];  // [nf][nbin]
  const Geo& g = a.g;
  const int t = threadIdx.x, x = blockIdx.x * 256 + t, j = blockIdx.y;
  if (j < z.jlo || j > z.jhi) return;  // the whole block: j is the block's row
  for (int i = t; i < z.nf * z.nbin; i += 256) hs[i] = 0.0;
  __syncthreads();
  const int ny = g.ny, nx = g.nx;
  if (x < nx) {
    const StatPoint P(a, x);
    const double* dyf = a.dyf;
    const bool gyf_row = (j >= 1 && j <= ny);
    const double wj = dyf[j];
    const double gx = ((double)x * z.lx) / (double)nx;
    const double bg = z.dthdx0 * (gx - 0.5 * z.lx);
    const double mu1 = P.M(SM_U1, j), mu2 = P.M(SM_U2, j), mu3 = P.M(SM_U3, j);
    for (int zz = 0; zz < g.nz; ++zz) {
      const double s1 = P.F(9, j, zz) + bg;
      int b;
      if (z.hx) {
        b = (int)((s1 - z.thmin) / z.dth);  // truncation toward zero, as int()
        b = b > z.nbin - 1 ? z.nbin - 1 : b;
        b = b < 0 ? 0 : b;  // below thmin (outside the reference's bins): the first bin
      } else {
        const double c = fmin(fmax(s1, z.thmin), z.thmax);
        b = (int)(acos(-c / z.den) * (double)z.nbin / z.pi);
        b = b > z.nbin - 1 ? z.nbin - 1 : b;
      }
      const double u1 = P.F(0, j, zz), u2 = P.F(3, j, zz), u3 = P.F(6, j, zz);
      const double f1 = u1 - mu1, f2 = u2 - mu2, f3 = u3 - mu3;
      double sh = 0.0, tke = 0.0;
      if (gyf_row) {
        const double d1 = (P.F(0, j + 1, zz) - P.F(0, j - 1, zz)) / (2.0 * dyf[j]);
        const double d3 = (P.F(6, j + 1, zz) - P.F(6, j - 1, zz)) / (2.0 * dyf[j]);
        sh = d1 * d1 + d3 * d3;
        tke = 0.5 * (f1 * f1) + 0.5 * (f2 * f2) + 0.5 * (f3 * f3);
      }
      const double eps = a.nu * P.eps(j, zz);
      const double ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
      atomicAdd(&hs[b], wj);
      atomicAdd(&hs[1 * z.nbin + b], sh * wj);
      atomicAdd(&hs[2 * z.nbin + b], eps * wj);
      atomicAdd(&hs[3 * z.nbin + b], tke * wj);
      atomicAdd(&hs[4 * z.nbin + b], ke * wj);
      for (int n = 0; n < a.nth; ++n) {
        const double tw = gyf_row ? P.thw(n, j, zz) : 0.0;
        const double chi = gyf_row ? P.chi(n, j, zz) : 0.0;
        atomicAdd(&hs[(5 + 2 * n) * z.nbin + b], tw * wj);
        atomicAdd(&hs[(6 + 2 * n) * z.nbin + b], chi * wj);
      }
    }
  }
  __syncthreads();
  for (int i = t; i < z.nf * z.nbin; i += 256)
    if (hs[i] != 0.0) atomicAdd(&z.hist[i], hs[i]);
}


// mean_xz.h5 (for/diagnostics.f90:283-378, 756-767, 1076-1087, 1238-1245, 1338-1346,
// 1452-1460): the z-mean (x, y) planes, rows 1..Nyp, each [Nyp][Nx] in the order
// ume, wme (v on GYF), vme (w), thme_1, uu, ww (v'v'), vv (w'w'), uw (u'v' on GY),
// wv (w'v' on GY), uv (u'w'), epsilon, epsilon_m, thth, thu, thw, chi (scalar 1;
// zero without a scalar).  The reference's plane names swap v and w.
constexpr int NXZ_PLANES = 16;
__global__ __launch_bounds__(256) void k_stat_xz(StatArgs a, double* __restrict__ out) {
  const Geo& g = a.g;
  const int x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y + 1;
  if (x >= g.nx) return;
  const int nx = g.nx;
  const long long pl = (long long)g.ny2 * nx, o = (long long)j * nx + x;
  const long long op = (long long)g.ny * nx, q = (long long)(j - 1) * nx + x;
  const double rz = 1.0 / (double)g.nz;
  auto Z = [&](int s) { return a.zs[s * pl + o] * rz; };
  auto M = [&](int m, int jj) { return a.zm[m * pl + (long long)jj * nx + x]; };
  const bool th = a.nth > 0;
  double v[NXZ_PLANES];
  v[0] = M(SM_U1, j);
  v[1] = 0.5 * (M(SM_U2, j) + M(SM_U2, j + 1));
  v[2] = M(SM_U3, j);
  v[3] = th ? M(SM_TH0, j) : 0.0;
  v[4] = Z(SZ_UU); v[5] = Z(SZ_VV); v[6] = Z(SZ_WW);
  v[7] = Z(SZ_UV); v[8] = Z(SZ_WV); v[9] = Z(SZ_UW);
  v[10] = a.nu * Z(SZ_EPS);
  v[11] = a.nu * a.zs[SZ_EPSM * pl + o];
  v[12] = th ? Z(SZ_TH0) : 0.0;
  v[13] = th ? Z(SZ_THU) : 0.0;
  v[14] = th ? Z(SZ_TH0 + 1) : 0.0;
  v[15] = th ? Z(SZ_TH0 + 2) : 0.0;
  for (int m = 0; m < NXZ_PLANES; ++m) out[m * op + q] = v[m];
}

// movie.h5 (for/diagnostics.f90:408-441, 485-617, 1041-1070, 1566-1727): per field
// three slices, x-y at z = nzm (reference name *_xz, [Nyp][Nx] of rows 1..Nyp),
// x-z at y = nym (*_xy, [Nz][Nx]; skipped when nym < 1) and z-y at x = nxm (*_yz,
// [Nyp][Nz]).  Fields: u1, u2 averaged to GYF, u3, nu_t, epsilon (nu times the
// integrand, zero on row 1), omega_x, omega_y, omega_z, then th_n.
struct MovieArgs {
  int nxm, nym, nzm, nfld;
  const void* nut;  // [j][z][x] in the handle's precision (null: zeros)
  int nut_f32;
  double* out;
};
constexpr int MOVIE_BASE = 8;
__global__ __launch_bounds__(256) void k_stat_movie(StatArgs a, MovieArgs m) {
  const Geo& g = a.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long n1 = (long long)ny * nx, n2 = (long long)nz * nx, n3 = (long long)ny * nz;
  const long long per = n1 + n2 + n3;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= per) return;
  int x, j, z;
  if (t < n1) { j = (int)(t / nx) + 1; x = (int)(t % nx); z = m.nzm; }
  else if (t < n1 + n2) { if (m.nym < 1) return; z = (int)((t - n1) / nx); x = (int)((t - n1) % nx); j = m.nym; }
  else { j = (int)((t - n1 - n2) / nz) + 1; z = (int)((t - n1 - n2) % nz); x = m.nxm; }
  const StatPoint P(a, x);
  const long long pt = (long long)j * g.pp + (long long)z * nx + x;
  double ox, oy, oz;
  P.omega(j, z, ox, oy, oz);
  double nt = 0.0;
  if (m.nut) nt = m.nut_f32 ? (double)((const float*)m.nut)[pt] : ((const double*)m.nut)[pt];
  double v[MOVIE_BASE];
  v[0] = P.F(0, j, z);
  v[1] = 0.5 * (P.F(3, j, z) + P.F(3, j + 1, z));
  v[2] = P.F(6, j, z);
  v[3] = nt;
  v[4] = a.nu * P.eps(j, z);
  v[5] = ox; v[6] = oy; v[7] = oz;
  for (int f = 0; f < MOVIE_BASE; ++f) m.out[f * per + t] = v[f];
  for (int n = 0; n < a.nth; ++n) m.out[(MOVIE_BASE + n) * per + t] = P.F(3 * (3 + n), j, z);
}

}  // namespace dgpu
