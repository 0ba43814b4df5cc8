// dgpu_closure.hpp — les_model_type 6: the learned closure evaluated per point.
// No counterpart in the reference; it replaces the AMD coefficient of
// les_chan / les_chan_th type 5 (for/les.f90:571-797, 1201-1348) by the
// network's prediction, and keeps the AMD coefficient wherever the point's
// features leave the training envelope.
//
// k_closure: per (j, x, z) point, the features at GY j from the same
// interpolated scaled-gradient entries coef_gy gives amd_coef and the type-5
// scalar gradient of amd_th_gy; each wave stages its 64 points in LDS and
// evaluates them as 4 MFMA tiles of 16 samples.  It writes the velocity and
// scalar coefficients at GY (1/time; nu_t = C^2 D2 coefficient) with the sign
// bit marking a fallback point (the AMD value, stored negated).  k_amd<..., 6>
// takes |coefficient| at GY, and at GYF j the mean of GY j and j+1 when both
// are network points, else the type-5 GYF coefficient, so that a handle with
// no network (or an empty envelope) steps exactly as type 5.
// Feature mode (dgpu_les_features) evaluates the features at the GYF points of
// rows jf0..Nyp-1 instead (the training data).
#pragma once
#include "dgpu_les.hpp"
#include "dgpu_mlp.hpp"

namespace dgpu {

struct ClosureArgs {
  AmdArgs a;      // geometry, gradient fields, filter-width table
  MlpW w;
  int have_net;   // 0: no network set (every point takes the AMD coefficient)
  double nu, gyb; // molecular viscosity, grav_y
  void* cgy;      // [ny2][pp] outputs at GY (step mode), precision T; sign bit set
  void* kgy;      // where the point took the AMD value
  int feat, jf0;  // feature mode and its first row
  // per-launch constants: -1/(2 C^2), sqrt(2)/nu, 1/dxs, 1/dzs
  double md2;
  float rnu2, rdxs, rdzs;
  float* X;       // [(ny - jf0) pp][16]
  float* scale;   // C^2 D2 sqrt(2 s2)
  float* amd;     // [n][2]: type-5 velocity and scalar coefficients / sqrt(2 s2)
  int* inside;
};

// n / d where d > 0, else 0, with the hardware reciprocal (1 ulp): the features
// become BF16 network inputs, whose resolution is 2^-8
__device__ __forceinline__ float fq(float n, float d) { return d > 0.f ? n * __builtin_amdgcn_rcpf(d) : 0.f; }
// tanh from one exp2 and one reciprocal: (1 - e^-2a) / (1 + e^-2a), and a - a^3/3
// below a = 2^-8 (relative error under 1e-5 either way)
__device__ __forceinline__ float tanh_fast(float v) {
  const float a = fabsf(v);
  const float t = __builtin_amdgcn_exp2f(-2.8853900817779268f * a);
  const float r = a < 0.00390625f ? a * (1.f - 0.33333333f * a * a) : (1.f - t) * __builtin_amdgcn_rcpf(1.f + t);
  return copysignf(r, v);
}
// log(1 + y), y >= 0, from the hardware log2 (y (1 - y/2) below 2^-10)
__device__ __forceinline__ float log1p_fast(float y) {
  return y < 0.0009765625f ? y * (1.f - 0.5f * y) : __builtin_amdgcn_logf(1.f + y) * 0.69314718055994531f;
}

// the 13 features (raw) of one point from the scaled strain (a1..a6), rotation
// (o4..o6), the scaled scalar gradient b, the unscaled th_y and |grad th|, the
// filter width term red = D2 sqrt(2) / nu and the row's log(delta_y^2 / (dxs dzs)),
// in float; returns s2 = S:S
__device__ __forceinline__ float closure_features(float a1, float a2, float a3, float a4, float a5, float a6,
                                                  float o4, float o5, float o6, float bx, float by, float bz,
                                                  float ty, float gth, float red, float f12, float gyb, float* f) {
  const float s2 = a1 * a1 + a2 * a2 + a3 * a3 + 2.f * (a4 * a4 + a5 * a5 + a6 * a6);
  const float w2 = 2.f * (o4 * o4 + o5 * o5 + o6 * o6);
  const float e = s2 + w2;
  // S = [[a1 a4 a5][a4 a2 a6][a5 a6 a3]], O = [[0 o4 o5][-o4 0 o6][-o5 -o6 0]]:
  // S^2 and P = O O^T are symmetric, so every trace and quadratic form below is
  // its three diagonal terms plus twice its three upper ones
  const float q11 = a1 * a1 + a4 * a4 + a5 * a5, q22 = a4 * a4 + a2 * a2 + a6 * a6, q33 = a5 * a5 + a6 * a6 + a3 * a3;
  const float q12 = a1 * a4 + a4 * a2 + a5 * a6, q13 = a1 * a5 + a4 * a6 + a5 * a3, q23 = a4 * a5 + a2 * a6 + a6 * a3;
  const float p11 = o4 * o4 + o5 * o5, p22 = o4 * o4 + o6 * o6, p33 = o5 * o5 + o6 * o6;
  const float p12 = o5 * o6, p13 = -(o4 * o6), p23 = o4 * o5;
  auto sym = [](float x11, float x22, float x33, float x12, float x13, float x23,
                float y11, float y22, float y33, float y12, float y13, float y23) {
    return x11 * y11 + x22 * y22 + x33 * y33 + 2.f * (x12 * y12 + x13 * y13 + x23 * y23);
  };
  const float trS3 = sym(q11, q22, q33, q12, q13, q23, a1, a2, a3, a4, a5, a6);
  const float trSP = sym(a1, a2, a3, a4, a5, a6, p11, p22, p33, p12, p13, p23);
  const float trS2P = sym(q11, q22, q33, q12, q13, q23, p11, p22, p33, p12, p13, p23);
  const float bxx = bx * bx, byy = by * by, bzz = bz * bz, bxy = bx * by, bxz = bx * bz, byz = by * bz;
  const float b2 = bxx + byy + bzz;
  const float bSb = sym(a1, a2, a3, a4, a5, a6, bxx, byy, bzz, bxy, bxz, byz);
  const float bS2b = sym(q11, q22, q33, q12, q13, q23, bxx, byy, bzz, bxy, bxz, byz);
  const float bPb = sym(p11, p22, p33, p12, p13, p23, bxx, byy, bzz, bxy, bxz, byz);
  const float rs = sqrtf(s2);
  f[0] = fq(w2 - s2, e);
  f[1] = fq(trS3, s2 * rs);
  f[2] = fq(trSP, rs * w2);
  f[3] = fq(-(trS3 + trSP), rs * e);
  f[4] = fq(trS2P, s2 * w2);
  f[5] = fq(byy, b2);
  f[6] = fq(bSb, b2 * rs);
  f[7] = fq(bS2b, b2 * s2);
  f[8] = fq(bPb, b2 * w2);
  f[9] = tanh_fast(fq(gyb * ty, s2));
  f[10] = tanh_fast(fq(fabsf(gyb) * gth, s2));
  f[11] = log1p_fast(red * rs);
  f[12] = f12;
  f[13] = f[14] = f[15] = 0.f;
  return s2;
}

__device__ __forceinline__ bool closure_inside(const float* norm, const float* f) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) in = in && f[k] >= norm[2 * MLP_INP + k] && f[k] <= norm[3 * MLP_INP + k];
  return in;
}

// the entries coef_gy passes to amd_coef at GY j
template <typename T>
struct GyEntries {
  T a1, a2, a3, a4, a5, a6, o4, o5, o6;
};
template <typename T>
__device__ __forceinline__ GyEntries<T> gy_entries(const AmdIn<T>& in, const AmdArgs& a, int j, long long q) {
  using R = T;
  const Gyf<T> f = gyf_plane(in, a, j, q), fm = gyf_plane(in, a, j - 1, q);
  const Gy<T> p = gy_plane(in, a, j, q);
  auto w = [&](R x, R y) { return gy_interp(a, j, x, y); };
  return {w(f.a1, fm.a1), w(f.a2, fm.a2), w(f.a3, fm.a3), p.a4, w(f.a5, fm.a5), p.a6, p.o4, w(f.o5, fm.o5), p.o6};
}

template <typename T, int NTH>
__global__ __launch_bounds__(256) void k_closure(ClosureArgs c) {
  __shared__ float sF[4][64][MLP_INP + 1];  // per wave: 64 samples
  __shared__ float sY[4][64][2];
  const AmdArgs& a = c.a;
  const Geo& g = a.g;
  int qb, j;
  xcd_rows(blockIdx.x, (int)((g.pp + 255) / 256), g.ny2, qb, j);
  const long long q0 = (long long)qb * 256 + threadIdx.x;
  const bool live = q0 < g.pp;
  const long long q = live ? q0 : g.pp - 1;  // every lane takes part in the MFMA tiles
  const int ny = g.ny;
  AmdIn<T> in;
  load_window<T, 3 * (3 + NTH)>(in, a, j, q);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // step mode: GY rows 2..Nyp; feature mode: GYF rows jf0..Nyp-1
  const bool on = live && (c.feat ? (j >= c.jf0 && j <= ny - 1) : (j >= 2 && j <= ny));
  float fe[MLP_INP];
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) fe[k] = 0.f;
  // the AMD values in the transform precision T (as k_amd's type 5), the features in float
  T amd_v = T(0), amd_t = T(0);
  float s2 = 0.f;
  if (on) {
    T a1, a2, a3, a4, a5, a6, o4, o5, o6;
    double dyl, D2;
    ThCoef<T> t{T(0), T(0), T(0), T(0), T(0)};
    if (c.feat) {
      // the entries coef_gyf passes to amd_coef at GYF j
      const Gyf<T> f = gyf_plane(in, a, j, q);
      const Gy<T> p = gy_plane(in, a, j + 1, q), m = gy_plane(in, a, j, q);
      a1 = f.a1; a2 = f.a2; a3 = f.a3; a5 = f.a5; o5 = f.o5;
      a4 = T(0.5) * (p.a4 + m.a4); a6 = T(0.5) * (p.a6 + m.a6); o4 = T(0.5) * (p.o4 + m.o4);
      o6 = T(0.5) * (p.o6 + m.o6);
      if constexpr (NTH > 0) t = amd_th_gyf(in, a, 9, j, q);
      dyl = a.dyf[j] * 2.0;
      D2 = a.dlt[j] * c.md2;
    } else {
      const GyEntries<T> E = gy_entries(in, a, j, q);
      a1 = E.a1; a2 = E.a2; a3 = E.a3; a4 = E.a4; a5 = E.a5; a6 = E.a6; o4 = E.o4; o5 = E.o5; o6 = E.o6;
      if constexpr (NTH > 0) t = amd_th_gy(in, a, 9, j, q);
      dyl = a.dy[j] * 2.0;
      D2 = a.dlt[g.ny2 + j] * c.md2;
    }
    amd_v = amd_coef(a1, a2, a3, a4, a5, a6, o4, o5, o6);
    amd_t = t.c;
    const float tx = (float)t.bx * c.rdxs, tz = (float)t.bz * c.rdzs, ty = (float)t.ty;
    const float gth = sqrtf(tx * tx + ty * ty + tz * tz);
    // constant on a row, and an envelope bound of the training data is exactly
    // one row's value: the same float division and logf as the training features
    const float fy = (float)dyl;
    const float f12 = logf(fy * fy / ((float)a.dxs * (float)a.dzs));
    s2 = closure_features((float)a1, (float)a2, (float)a3, (float)a4, (float)a5, (float)a6, (float)o4, (float)o5,
                          (float)o6, (float)t.bx, (float)t.by, (float)t.bz, ty, gth, (float)D2 * c.rnu2, f12,
                          (float)c.gyb, fe);
  }
  // network: 4 tiles of 16 samples per wave, skipped by a wave none of whose
  // points is inside the envelope (the log width ratio is constant on a row, so
  // whole rows fall outside on stretched grids)
  const bool inside = c.have_net && s2 > 0.f && closure_inside(c.w.norm, fe);
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) sF[wave][lane][k] = fe[k];
  __syncthreads();
  if (c.have_net && !c.feat && __any(inside)) {
    const int r = lane & 15, gq = lane >> 4;
    MlpReg m;
    mlp_load(c.w, m);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const f32x4 y = mlp_infer(m, &sF[wave][16 * t + r][4 * gq]);
      if (gq == 0) {
        sY[wave][16 * t + r][0] = y[0];
        sY[wave][16 * t + r][1] = y[1];
      }
    }
  }
  __syncthreads();
  if (c.feat) {
    if (on) {
      const long long s = (long long)(j - c.jf0) * g.pp + q;
#pragma unroll
      for (int k = 0; k < MLP_INP; ++k) c.X[s * MLP_INP + k] = fe[k];
      const double r2 = sqrt(2.0 * (double)s2);
      c.scale[s] = (float)(-0.5 * a.dlt[j] * r2);
      c.amd[2 * s] = (float)(r2 > 0.0 ? (double)amd_v / r2 : 0.0);
      c.amd[2 * s + 1] = (float)(r2 > 0.0 ? (double)amd_t / r2 : 0.0);
      c.inside[s] = c.have_net ? (int)inside : 1;
    }
    return;
  }
  if (!live) return;
  // coefficients (1/time) at GY: the network's sqrt(2 s2) max(0, y), or AMD's stored negated
  const long long o = (long long)j * g.pp + q;
  const float r2 = sqrtf(2.f * s2);
  ((T*)c.cgy)[o] = inside ? (T)(r2 * fmaxf(0.f, sY[wave][lane][0])) : -amd_v;
  ((T*)c.kgy)[o] = inside ? (T)(r2 * fmaxf(0.f, sY[wave][lane][1])) : -amd_t;
}

// The GY points of rows 2..Nyp whose coefficient (velocity) came from the
// network: the sign bit of cgy is clear there (k_closure stores the AMD fallback
// negated).  One atomic per block into count.
template <typename T>
__global__ __launch_bounds__(256) void k_closure_count(const T* __restrict__ cgy, long long pp, int ny,
                                                        unsigned long long* count) {
  const long long n = (long long)(ny - 1) * pp, base = 2 * pp;
  int c = 0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    c += signbit(cgy[base + i]) ? 0 : 1;
  // block sum through LDS
  __shared__ int sc[256];
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sc[threadIdx.x] += sc[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && sc[0]) atomicAdd(count, (unsigned long long)sc[0]);
}

}  // namespace dgpu
