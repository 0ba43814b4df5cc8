// dgpu_boxfilter.hpp — the learned closure's training filter on the device: a DNS
// state box-filtered onto a coarser LES grid of the same domain, with the
// products the subgrid stress and scalar flux need.  (The reference filters DNS
// offline in MATLAB by block averages, proc/functions/coarsen.m:1-25; this is
// the same operation on the device, with the LES grid's own y cells.)
#pragma once
#include "dgpu_kernels.hpp"

namespace dgpu {

// Output fields, [f][nl + 2][nz / fz][nx / fx]: the filtered
//   0 u1, 1 u2 at GY (y weights mg), 2 u2 at GYF, 3 u3, 4 th,
//   5 u1u1, 6 u1u2, 7 u1u3, 8 u2u2, 9 u2u3, 10 u3u3, 11 u1th, 12 u2th, 13 u3th
// where u2 in the GYF fields and products is the DNS's own GYF interpolation
// (u2(j) + u2(j+1))/2 on rows 1..Ny (zero on rows 0 and Ny+1).
constexpr int BOX_FIELDS = 14;

struct BoxArgs {
  Geo g;                  // the DNS grid
  int fx, fz, nxl, nzl, nl2;
  const double* u[4];     // physical u1, u2, u3, th, [j][z][x]
  const double* mf;       // [nl2][ny2] y weights of the GYF fields (rows of the LES cells)
  const double* mg;       // [nl2][ny2] y weights of the GY field
  const int* rf;          // [nl2][2] first and last DNS row with a nonzero weight in mf
  const int* rg;          // same for mg
  double* out;
};

// One LES point per thread: the top-hat of width fx (fz) DNS cells centred on the
// point in x (z), trapezoid weights over its fx + 1 (fz + 1) DNS points,
// periodic; summed in x, then z, then over the DNS rows with the y weights.
__global__ __launch_bounds__(256) void k_box_filter(BoxArgs a) {
  const Geo& g = a.g;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= a.nxl * a.nzl) return;
  const int xl = t % a.nxl, zl = t / a.nxl, jl = blockIdx.y;
  const long long plane = (long long)a.nxl * a.nzl;
  const double rx = 1.0 / a.fx, rz = 1.0 / a.fz;
  auto wgt = [](int m, int f, double r) { return (m == 0 || m == f) ? 0.5 * r : r; };
  double acc[13];
  for (int q = 0; q < 13; ++q) acc[q] = 0.0;
  double accg = 0.0;
  const int k0 = a.rf[2 * jl], k1 = a.rf[2 * jl + 1];
  for (int k = k0; k <= k1; ++k) {
    const double wy = a.mf[(long long)jl * g.ny2 + k];
    if (wy == 0.0) continue;
    double sz[13];
    for (int q = 0; q < 13; ++q) sz[q] = 0.0;
    for (int mz = 0; mz <= a.fz; ++mz) {
      const int z = ((zl * a.fz + mz - a.fz / 2) % g.nz + g.nz) % g.nz;
      const long long row = ((long long)k * g.nz + z) * g.nx;
      const long long rowp = row + g.pp;  // row k+1 of u2
      double sx[13];
      for (int q = 0; q < 13; ++q) sx[q] = 0.0;
      for (int mx = 0; mx <= a.fx; ++mx) {
        const int x = ((xl * a.fx + mx - a.fx / 2) % g.nx + g.nx) % g.nx;
        const double w = wgt(mx, a.fx, rx);
        const double v1 = a.u[0][row + x];
        const double v2 = (k >= 1 && k <= g.ny) ? 0.5 * (a.u[1][row + x] + a.u[1][rowp + x]) : 0.0;
        const double v3 = a.u[2][row + x];
        const double th = a.u[3][row + x];
        const double v[13] = {v1, v2, v3, th, v1 * v1, v1 * v2, v1 * v3, v2 * v2, v2 * v3, v3 * v3,
                              v1 * th, v2 * th, v3 * th};
        for (int q = 0; q < 13; ++q) sx[q] = sx[q] + w * v[q];
      }
      const double wz = wgt(mz, a.fz, rz);
      for (int q = 0; q < 13; ++q) sz[q] = sz[q] + wz * sx[q];
    }
    for (int q = 0; q < 13; ++q) acc[q] = acc[q] + wy * sz[q];
  }
  const int g0 = a.rg[2 * jl], g1 = a.rg[2 * jl + 1];
  for (int k = g0; k <= g1; ++k) {
    const double wy = a.mg[(long long)jl * g.ny2 + k];
    if (wy == 0.0) continue;
    double sz = 0.0;
    for (int mz = 0; mz <= a.fz; ++mz) {
      const int z = ((zl * a.fz + mz - a.fz / 2) % g.nz + g.nz) % g.nz;
      const long long row = ((long long)k * g.nz + z) * g.nx;
      double sx = 0.0;
      for (int mx = 0; mx <= a.fx; ++mx) {
        const int x = ((xl * a.fx + mx - a.fx / 2) % g.nx + g.nx) % g.nx;
        sx = sx + wgt(mx, a.fx, rx) * a.u[1][row + x];
      }
      sz = sz + wgt(mz, a.fz, rz) * sx;
    }
    accg = accg + wy * sz;
  }
  const long long o = (long long)jl * plane + t;
  const long long fs = (long long)a.nl2 * plane;
  a.out[o] = acc[0];
  a.out[fs + o] = accg;
  a.out[2 * fs + o] = acc[1];
  a.out[3 * fs + o] = acc[2];
  a.out[4 * fs + o] = acc[3];
  for (int q = 4; q < 13; ++q) a.out[(long long)(q + 1) * fs + o] = acc[q];
}

}  // namespace dgpu
