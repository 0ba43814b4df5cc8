// dgpu_courant.hpp — the CFL limit of courant (for/flow.f90:746-806) on the
// device: min over the physical points of CFL dx/|u1 - Vx|, CFL dy/|u2|,
// CFL dz/|u3 + Vz| (V the thermal wind of flavour 'Front', per scalar), reduced
// per block; the host takes the min of the block results and of the
// viscous/inertial/buoyancy limits.  min is exact, so the result does not
// depend on the reduction order.
#pragma once
#include "dgpu_kernels.hpp"

namespace dgpu {

struct CourantArgs {
  Geo g;
  int jlo, jhi, nth, front;
  double cfl, dx, dz, ly, f;  // f = Ro_inv / delta
  double dthdx[MAXTH], dthdz[MAXTH];
  const double* __restrict__ dy;
  const double* __restrict__ gyf;
  const double* __restrict__ u1;
  const double* __restrict__ u2;
  const double* __restrict__ u3;
  double* part;               // one value per block
};

__global__ __launch_bounds__(256) void k_courant(CourantArgs a) {
  __shared__ double sh[256];
  const Geo& g = a.g;
  const long long npts = (long long)(a.jhi - a.jlo + 1) * g.pp;
  double m = INFINITY;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < npts; q += (long long)gridDim.x * 256) {
    const int j = a.jlo + (int)(q / g.pp);
    const long long o = (long long)j * g.pp + q % g.pp;
    const double u1 = a.u1[o], u2 = a.u2[o], u3 = a.u3[o];
    if (a.front) {
      for (int n = 0; n < a.nth; ++n) {
        const double dtx = a.cfl * a.dx / fabs(u1 - a.dthdz[n] * a.gyf[j] / a.f);
        const double dty = a.cfl * a.dy[j] / fabs(u2);
        const double dtz = a.cfl * a.dz / fabs(u3 + (1.0 / a.f) * a.dthdx[n] * (a.gyf[j] - 0.5 * a.ly));
        m = fmin(m, fmin(dtx, fmin(dty, dtz)));
      }
    } else {
      m = fmin(m, fmin(a.cfl * a.dx / fabs(u1), fmin(a.cfl * a.dy[j] / fabs(u2), a.cfl * a.dz / fabs(u3))));
    }
  }
  const int t = threadIdx.x;
  sh[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmin(sh[t], sh[t + s]);
    __syncthreads();
  }
  if (t == 0) a.part[blockIdx.x] = sh[0];
}

}  // namespace dgpu
