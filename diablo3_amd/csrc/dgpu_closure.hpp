// dgpu_closure.hpp — les_model_type 6: the learned closure evaluated per point.
// No counterpart in the reference; it replaces the AMD coefficient of
// les_chan / les_chan_th type 5 (for/les.f90:571-797, 1201-1348) by the
// network's prediction, and keeps the AMD coefficient wherever the point's
// features leave the training envelope.
//
// k_closure: per (j, x, z) point, the features at GY j and at GYF j from the
// same interpolated scaled-gradient entries the type-5 coefficients use
// (coef_gy, coef_gyf) and the type-5 scalar gradients (amd_th_gy, amd_th_gyf);
// each wave stages its 64 points x 2 locations in LDS and evaluates them as 8
// MFMA tiles of 16 samples.  It writes the velocity and scalar coefficients at
// GY and GYF (1/time; nu_t = C^2 D2 coefficient), which k_amd<..., 6> turns
// into nu_t, kappa_t and the SGS stresses and fluxes exactly as type 5 does.
// In feature mode it writes the features of the GYF points of rows jf0..Nyp-1
// instead (training data, dgpu_les_features).
#pragma once
#include "dgpu_les.hpp"
#include "dgpu_mlp.hpp"

namespace dgpu {

struct ClosureArgs {
  AmdArgs a;      // geometry, gradient fields, filter-width table
  MlpW w;
  int have_net;   // 0: no network set (every point takes the AMD coefficient)
  double nu, gyb; // molecular viscosity, grav_y
  float* cgyf;    // [ny2][pp] outputs (step mode)
  float* cgy;
  float* kgyf;
  float* kgy;
  int feat, jf0;  // feature mode and its first row
  float* X;       // [(ny - jf0) pp][16]
  float* scale;   // C^2 D2 sqrt(2 s2)
  float* amd;     // type-5 coefficient / sqrt(2 s2)
  int* inside;
};

// the 13 features (raw) of one point from the scaled strain (a1..a6), rotation
// (o4..o6), the scaled scalar gradient b, the unscaled th_y and |grad th|, the
// filter width D2 and the y width delta_y; returns s2 = S:S
__device__ __forceinline__ double closure_features(double a1, double a2, double a3, double a4, double a5, double a6,
                                                   double o4, double o5, double o6, double bx, double by, double bz,
                                                   double ty, double gth, double D2, double dyl, double dxs,
                                                   double dzs, double nu, double gyb, float* f) {
  const double s2 = a1 * a1 + a2 * a2 + a3 * a3 + 2.0 * (a4 * a4 + a5 * a5 + a6 * a6);
  const double w2 = 2.0 * (o4 * o4 + o5 * o5 + o6 * o6);
  const double e = s2 + w2;
  // S = [[a1 a4 a5][a4 a2 a6][a5 a6 a3]], O = [[0 o4 o5][-o4 0 o6][-o5 -o6 0]], P = O O^T
  const double S[3][3] = {{a1, a4, a5}, {a4, a2, a6}, {a5, a6, a3}};
  const double O[3][3] = {{0.0, o4, o5}, {-o4, 0.0, o6}, {-o5, -o6, 0.0}};
  double S2[3][3], P[3][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0, p = 0.0;
      for (int m = 0; m < 3; ++m) {
        s += S[i][m] * S[m][k];
        p += O[i][m] * O[k][m];
      }
      S2[i][k] = s;
      P[i][k] = p;
    }
  double trS3 = 0.0, trSP = 0.0, trS2P = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      trS3 += S2[i][k] * S[k][i];
      trSP += S[i][k] * P[k][i];
      trS2P += S2[i][k] * P[k][i];
    }
  const double b[3] = {bx, by, bz};
  const double b2 = bx * bx + by * by + bz * bz;
  double bSb = 0.0, bS2b = 0.0, bPb = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      bSb += b[i] * S[i][k] * b[k];
      bS2b += b[i] * S2[i][k] * b[k];
      bPb += b[i] * P[i][k] * b[k];
    }
  const double rs = sqrt(s2);
  auto q = [](double n, double d) { return d > 0.0 ? n / d : 0.0; };
  f[0] = (float)q(w2 - s2, e);
  f[1] = (float)q(trS3, s2 * rs);
  f[2] = (float)q(trSP, rs * w2);
  f[3] = (float)q(-(trS3 + trSP), rs * e);
  f[4] = (float)q(trS2P, s2 * w2);
  f[5] = (float)q(by * by, b2);
  f[6] = (float)q(bSb, b2 * rs);
  f[7] = (float)q(bS2b, b2 * s2);
  f[8] = (float)q(bPb, b2 * w2);
  f[9] = (float)tanh(q(gyb * ty, s2));
  f[10] = (float)tanh(q(fabs(gyb) * gth, s2));
  f[11] = (float)log1p(D2 * sqrt(2.0 * s2) / nu);
  f[12] = (float)log(dyl * dyl / (dxs * dzs));
  f[13] = f[14] = f[15] = 0.f;
  return s2;
}

__device__ __forceinline__ bool closure_inside(const float* norm, const float* f) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) in = in && f[k] >= norm[2 * MLP_INP + k] && f[k] <= norm[3 * MLP_INP + k];
  return in;
}

template <typename T, int NTH>
__global__ __launch_bounds__(256) void k_closure(ClosureArgs c) {
  __shared__ float sF[4][128][MLP_INP + 1];  // per wave: 64 GY + 64 GYF samples
  __shared__ float sY[4][128][2];
  const AmdArgs& a = c.a;
  const Geo& g = a.g;
  const long long q0 = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool live = q0 < g.pp;
  const long long q = live ? q0 : g.pp - 1;  // every lane takes part in the MFMA tiles
  const int j = blockIdx.y, ny = g.ny;
  AmdIn<T> in;
  for (int s = 0; s < 3 * (3 + NTH); ++s) in.v[s] = (const T*)a.f[s];
  in.pp = g.pp;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double C2 = C_AMD * C_AMD;
  // GYF rows: the velocity stresses jstart..Nyp-1, the scalar fluxes jstart_th..jend_th
  int jg0 = a.jstart, jg1 = ny - 1;
  for (int n = 0; n < NTH; ++n) {
    jg0 = a.jstart_th[n] < jg0 ? a.jstart_th[n] : jg0;
    jg1 = a.jend_th[n] > jg1 ? a.jend_th[n] : jg1;
  }
  const bool gy_on = live && !c.feat && j >= 2 && j <= ny;
  const bool gyf_on = live && (c.feat ? (j >= c.jf0 && j <= ny - 1) : (j >= jg0 && j <= jg1));
  float fa[MLP_INP], fb[MLP_INP];
  double s2a = 0.0, s2b = 0.0, amda = 0.0, amdb = 0.0, ka = 0.0, kb = 0.0;
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) fa[k] = fb[k] = 0.f;
  if (gy_on) {
    // the entries coef_gy passes to amd_coef
    const Gyf f = gyf_plane(in, a, j, q), fm = gyf_plane(in, a, j - 1, q);
    const Gy p = gy_plane(in, a, j, q);
    const double wc = a.dyf[j - 1], wm = a.dyf[j], den = 2.0 * a.dy[j];
    auto w = [&](double x, double y) { return (x * wc + y * wm) / den; };
    const double a1 = w(f.a1, fm.a1), a2 = w(f.a2, fm.a2), a3 = w(f.a3, fm.a3), a5 = w(f.a5, fm.a5),
                 o5 = w(f.o5, fm.o5);
    amda = amd_coef(a1, a2, a3, p.a4, a5, p.a6, p.o4, o5, p.o6);
    double bx = 0.0, by = 0.0, bz = 0.0, ty = 0.0, gth = 0.0;
    if constexpr (NTH > 0) {
      const ThCoef t = amd_th_gy(in, a, 9, j, q);
      bx = t.bx; by = t.by; bz = t.bz; ty = t.ty;
      ka = t.c;
      const double tx = bx / a.dxs, tz = bz / a.dzs;
      gth = sqrt(tx * tx + ty * ty + tz * tz);
    }
    const double D2 = -a.dlt[g.ny2 + j] / (2.0 * C2);
    s2a = closure_features(a1, a2, a3, p.a4, a5, p.a6, p.o4, o5, p.o6, bx, by, bz, ty, gth, D2, a.dy[j] * 2.0,
                           a.dxs, a.dzs, c.nu, c.gyb, fa);
  }
  if (gyf_on) {
    // the entries coef_gyf passes to amd_coef
    const Gyf f = gyf_plane(in, a, j, q);
    const Gy p = gy_plane(in, a, j + 1, q), m = gy_plane(in, a, j, q);
    const double a4 = 0.5 * (p.a4 + m.a4), a6 = 0.5 * (p.a6 + m.a6), o4 = 0.5 * (p.o4 + m.o4),
                 o6 = 0.5 * (p.o6 + m.o6);
    amdb = amd_coef(f.a1, f.a2, f.a3, a4, f.a5, a6, o4, f.o5, o6);
    double bx = 0.0, by = 0.0, bz = 0.0, ty = 0.0, gth = 0.0;
    if constexpr (NTH > 0) {
      const ThCoef t = amd_th_gyf(in, a, 9, j, q);
      bx = t.bx; by = t.by; bz = t.bz; ty = t.ty;
      kb = t.c;
      const double tx = bx / a.dxs, tz = bz / a.dzs;
      gth = sqrt(tx * tx + ty * ty + tz * tz);
    }
    const double D2 = -a.dlt[j] / (2.0 * C2);
    s2b = closure_features(f.a1, f.a2, f.a3, a4, f.a5, a6, o4, f.o5, o6, bx, by, bz, ty, gth, D2, a.dyf[j] * 2.0,
                           a.dxs, a.dzs, c.nu, c.gyb, fb);
  }
  // network: 8 tiles of 16 samples per wave (GY samples 0..63, GYF 64..127)
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) {
    sF[wave][lane][k] = fa[k];
    sF[wave][64 + lane][k] = fb[k];
  }
  __syncthreads();
  if (c.have_net) {
    const int r = lane & 15, gq = lane >> 4;
    for (int t = 0; t < 8; ++t) {
      const float* row = &sF[wave][16 * t + r][4 * gq];
      bf16x4 xb;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int fi = 4 * gq + v;
        xb[v] = (short)f2bf((row[v] - c.w.norm[fi]) * c.w.norm[MLP_INP + fi]);
      }
      MlpAct act;
      mlp_forward(c.w, xb, act);
      if (gq == 0) {
        sY[wave][16 * t + r][0] = act.y[0];
        sY[wave][16 * t + r][1] = act.y[1];
      }
    }
  }
  __syncthreads();
  const bool ina = c.have_net && s2a > 0.0 && closure_inside(c.w.norm, fa);
  const bool inb = c.have_net && s2b > 0.0 && closure_inside(c.w.norm, fb);
  const long long o = (long long)j * g.pp + q;
  if (c.feat) {
    if (gyf_on) {
      const long long s = (long long)(j - c.jf0) * g.pp + q;
#pragma unroll
      for (int k = 0; k < MLP_INP; ++k) c.X[s * MLP_INP + k] = fb[k];
      const double r2 = sqrt(2.0 * s2b);
      c.scale[s] = (float)(-0.5 * a.dlt[j] * r2);
      c.amd[s] = (float)(r2 > 0.0 ? amdb / r2 : 0.0);
      c.inside[s] = c.have_net ? (int)inb : 1;
    }
    return;
  }
  if (!live) return;
  // coefficients (1/time): the network's sqrt(2 s2) max(0, y), else AMD's
  const double ra = sqrt(2.0 * s2a), rb = sqrt(2.0 * s2b);
  c.cgy[o] = (float)(ina ? ra * fmax(0.0, (double)sY[wave][lane][0]) : amda);
  c.kgy[o] = (float)(ina ? ra * fmax(0.0, (double)sY[wave][lane][1]) : ka);
  c.cgyf[o] = (float)(inb ? rb * fmax(0.0, (double)sY[wave][64 + lane][0]) : amdb);
  c.kgyf[o] = (float)(inb ? rb * fmax(0.0, (double)sY[wave][64 + lane][1]) : kb);
}

}  // namespace dgpu
