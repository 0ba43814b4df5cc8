// dgpu_fft.hpp — device-side complex type and batched LDS Stockham FFT for gfx950.
//
// The transform is the autosort Stockham formulation: stage s of radix R reads
// x[j + r*N/R] (r < R) for butterfly j, multiplies input r by w^(r*k) with
// k = j mod Ns, w = exp(-+2 pi i/(Ns*R)), applies a radix-R DFT and writes
// y[(j/Ns)*Ns*R + k + r*Ns]; Ns grows by R per stage and the output is in
// natural order.  Radix-16 stages (in-register 4x4 butterflies), with one radix
// 8 or 4 stage for leftover bits: 256 points take two passes through LDS.
// Lines live in LDS at s[l*ld + swz(n)] (an XOR swizzle of the index within the
// line that keeps the strided stage writes free of bank conflicts); every thread
// of the block takes part, and a stage reads all of its butterflies into
// registers before any write, so one LDS buffer suffices (two barriers per stage).
//
// Sign and scale follow FFTW: DIR = -1 is exp(-i...), DIR = +1 is exp(+i...),
// neither is normalised.  Twiddles come from a table tw[m] = exp(-2 pi i m/N),
// m < N, computed on the host in long double and rounded once.
#pragma once
#include <hip/hip_runtime.h>

namespace dgpu {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};
using cd = cx<double>;
using cf = cx<float>;

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__device__ __forceinline__ cx<T> conjg(cx<T> a) { return {a.x, -a.y}; }
// i*k*a for real k: (0 + i k)(x + i y) = (-k y) + i (k x), as Fortran's cikx*cu
template <typename T>
__device__ __forceinline__ cx<T> ik(T k, cx<T> a) { return {-(k * a.y), k * a.x}; }
// multiply by -i (DIR = -1) or +i (DIR = +1)
template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
  if constexpr (DIR < 0) return {a.y, -a.x};
  else return {-a.y, a.x};
}

template <typename TO, typename TI>
__device__ __forceinline__ cx<TO> cvt(cx<TI> a) { return {(TO)a.x, (TO)a.y}; }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cx<T>& v0, cx<T>& v1, cx<T>& v2, cx<T>& v3) {
  cx<T> a0 = v0 + v2, a1 = v0 - v2, a2 = v1 + v3, a3 = rot<DIR>(v1 - v3);
  v0 = a0 + a2;
  v2 = a0 - a2;
  v1 = a1 + a3;
  v3 = a1 - a3;
}

// exp(DIR * 2 pi i m / 16) for the m that the radix-8 and radix-16 butterflies use
template <int DIR, typename T>
__device__ __forceinline__ cx<T> w16(int m) {
  constexpr double C = 0.92387953251128675613;  // cos(pi/8)
  constexpr double S = 0.38268343236508977173;  // sin(pi/8)
  constexpr double H = 0.70710678118654752440;  // cos(pi/4)
  constexpr T sg = DIR < 0 ? T(-1) : T(1);
  switch (m) {
    case 1: return {T(C), sg * T(S)};
    case 2: return {T(H), sg * T(H)};
    case 3: return {T(S), sg * T(C)};
    case 6: return {T(-H), sg * T(H)};
    case 9: return {T(-C), -sg * T(S)};
    default: return {T(1), T(0)};
  }
}

// In-register DFTs of length 8 and 16 in natural order (Cooley-Tukey 2x4 and 4x4).
template <int DIR, typename T>
__device__ __forceinline__ void dft8(cx<T>* v) {
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  o1 = o1 * w16<DIR, T>(2);
  o2 = rot<DIR>(o2);
  o3 = o3 * w16<DIR, T>(6);
  v[0] = e0 + o0; v[4] = e0 - o0;
  v[1] = e1 + o1; v[5] = e1 - o1;
  v[2] = e2 + o2; v[6] = e2 - o2;
  v[3] = e3 + o3; v[7] = e3 - o3;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft16(cx<T>* v) {
  cx<T> y[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    y[r][0] = v[r]; y[r][1] = v[r + 4]; y[r][2] = v[r + 8]; y[r][3] = v[r + 12];
    dft4<DIR>(y[r][0], y[r][1], y[r][2], y[r][3]);
  }
  y[1][1] = y[1][1] * w16<DIR, T>(1);
  y[1][2] = y[1][2] * w16<DIR, T>(2);
  y[1][3] = y[1][3] * w16<DIR, T>(3);
  y[2][1] = y[2][1] * w16<DIR, T>(2);
  y[2][2] = rot<DIR>(y[2][2]);
  y[2][3] = y[2][3] * w16<DIR, T>(6);
  y[3][1] = y[3][1] * w16<DIR, T>(3);
  y[3][2] = y[3][2] * w16<DIR, T>(6);
  y[3][3] = y[3][3] * w16<DIR, T>(9);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cx<T> a = y[0][q], b = y[1][q], c = y[2][q], d = y[3][q];
    dft4<DIR>(a, b, c, d);
    v[q] = a; v[q + 4] = b; v[q + 8] = c; v[q + 12] = d;
  }
}

// Within-line LDS index swizzle: the low bits of the element index are XORed with
// the next bits so that the strided writes of the early Stockham stages (stride
// R*Ns elements) spread over all banks.  A bijection on [0, N) for every N.
template <typename T>
__device__ __forceinline__ int swz(int n) {
  constexpr int SH = sizeof(cx<T>) == 16 ? 4 : 5;  // 16 (complex double) or 32 (complex float) per 256 B
  constexpr int MASK = (1 << SH) - 1;
  return n ^ ((n >> SH) & MASK);
}

// One Stockham stage of radix R over LINES lines of length N; Ns (the length of
// the sub-transforms already done) is a compile-time constant, so the first stage
// (Ns = 1, every twiddle 1) has no twiddle loads or products.
template <typename T, int N, int LINES, int NT, int DIR, int R, int Ns>
__device__ __forceinline__ void fft_stage(cx<T>* __restrict__ s, int ld,
                                          const cx<T>* __restrict__ tw, int tid) {
  constexpr int NB = LINES * (N / R);
  constexpr int PER = (NB + NT - 1) / NT;
  cx<T> v[PER][R];
  int ob[PER], ol[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    ob[p] = -1;
    ol[p] = 0;
    if (NB % NT == 0 || b < NB) {
      const int l = b / (N / R);
      const int j = b - l * (N / R);
      const cx<T>* line = s + l * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = line[swz<T>(j + r * (N / R))];
      const int k = j & (Ns - 1);
      constexpr int tstep = N / (Ns * R);
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = tw[r * k * tstep];
          if constexpr (DIR > 0) w = conjg(w);
          v[p][r] = v[p][r] * w;
        }
      }
      if constexpr (R == 16) {
        dft16<DIR>(v[p]);
      } else if constexpr (R == 8) {
        dft8<DIR>(v[p]);
      } else if constexpr (R == 4) {
        dft4<DIR>(v[p][0], v[p][1], v[p][2], v[p][3]);
      } else {
        cx<T> a = v[p][0], c = v[p][1];
        v[p][0] = a + c;
        v[p][1] = a - c;
      }
      ob[p] = (j / Ns) * Ns * R + k;
      ol[p] = l * ld;
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    if (ob[p] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) s[ol[p] + swz<T>(ob[p] + r * Ns)] = v[p][r];
    }
  }
  __syncthreads();
}

// radix of stage st for a transform of 2^L points: radix-16 stages, then one
// stage of 2, 4 or 8 for the remaining bits (2 bits left over: a radix-4 stage;
// 1 bit: the last radix-16 becomes 8 x 4 rather than 16 x 2)
constexpr int stage_radix(int L, int st) {
  const int full = L / 4, rem = L % 4;
  if (rem == 0) return 16;
  if (rem == 1) {
    if (full == 0) return 2;
    if (st < full - 1) return 16;
    return st == full - 1 ? 8 : 4;
  }
  if (st < full) return 16;
  return 1 << rem;  // 4 or 8
}
constexpr int stage_count(int L) { return L % 4 == 0 ? L / 4 : L / 4 + 1; }

template <typename T, int N, int LINES, int NT, int DIR, int ST, int NST, int Ns>
__device__ __forceinline__ void fft_stages(cx<T>* __restrict__ s, int ld, const cx<T>* __restrict__ tw,
                                           int tid) {
  if constexpr (ST < NST) {
    constexpr int R = stage_radix(ilog2(N), ST);
    fft_stage<T, N, LINES, NT, DIR, R, Ns>(s, ld, tw, tid);
    fft_stages<T, N, LINES, NT, DIR, ST + 1, NST, Ns * R>(s, ld, tw, tid);
  }
}

// The same stage with every line owned by one wave (line l by wave l % NW): a
// wave's LDS reads and writes are processed in order, so the stage needs no
// barrier.  A wave runs the N / R butterflies of each of its lines, PER per
// lane.
template <typename T, int N, int LINES, int NT, int DIR, int R, int Ns>
__device__ __forceinline__ void fft_stage_wave(cx<T>* __restrict__ s, int ld,
                                               const cx<T>* __restrict__ tw, int tid) {
  constexpr int NW = NT / 64, BL = N / R;
  static_assert(NT % 64 == 0, "whole waves");
  constexpr int LW = (LINES + NW - 1) / NW;  // lines of the busiest wave
  constexpr int PER = (LW * BL + 63) / 64;
  const int w = tid >> 6, lane = tid & 63;
  cx<T> v[PER][R];
  int ob[PER], ol[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = lane + p * 64;
    const int li = b / BL, j = b - li * BL;
    const int l = w + li * NW;
    ob[p] = -1;
    ol[p] = 0;
    if (li < LW && l < LINES) {
      const cx<T>* line = s + l * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = line[swz<T>(j + r * (N / R))];
      const int k = j & (Ns - 1);
      constexpr int tstep = N / (Ns * R);
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> tw_ = tw[r * k * tstep];
          if constexpr (DIR > 0) tw_ = conjg(tw_);
          v[p][r] = v[p][r] * tw_;
        }
      }
      if constexpr (R == 16) {
        dft16<DIR>(v[p]);
      } else if constexpr (R == 8) {
        dft8<DIR>(v[p]);
      } else if constexpr (R == 4) {
        dft4<DIR>(v[p][0], v[p][1], v[p][2], v[p][3]);
      } else {
        cx<T> a = v[p][0], c = v[p][1];
        v[p][0] = a + c;
        v[p][1] = a - c;
      }
      ob[p] = (j / Ns) * Ns * R + k;
      ol[p] = l * ld;
    }
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    if (ob[p] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) s[ol[p] + swz<T>(ob[p] + r * Ns)] = v[p][r];
    }
  }
}

template <typename T, int N, int LINES, int NT, int DIR, int ST, int NST, int Ns>
__device__ __forceinline__ void fft_stages_wave(cx<T>* __restrict__ s, int ld, const cx<T>* __restrict__ tw,
                                                int tid) {
  if constexpr (ST < NST) {
    constexpr int R = stage_radix(ilog2(N), ST);
    fft_stage_wave<T, N, LINES, NT, DIR, R, Ns>(s, ld, tw, tid);
    fft_stages_wave<T, N, LINES, NT, DIR, ST + 1, NST, Ns * R>(s, ld, tw, tid);
  }
}

// Batched in-place FFT of LINES lines of length N (power of two, >= 2) in LDS,
// element n of line l at s[l*ld + swz<T>(n)], with the N twiddles tw (in LDS:
// see stage_twiddles).  Caller must __syncthreads() after filling s and tw; the
// routine ends with a barrier.
template <typename T, int N, int LINES, int NT, int DIR>
__device__ __forceinline__ void fft_lds(cx<T>* __restrict__ s, int ld,
                                        const cx<T>* __restrict__ tw, int tid) {
  fft_stages<T, N, LINES, NT, DIR, 0, stage_count(ilog2(N)), 1>(s, ld, tw, tid);
}

// The same transform with the lines shared out among the waves (line l to wave
// l % (NT / 64)) and no barrier inside.  Same arithmetic, so the same result as
// fft_lds.  Caller must __syncthreads() after filling s and tw, and before other
// waves read the lines.
template <typename T, int N, int LINES, int NT, int DIR>
__device__ __forceinline__ void fft_lds_wave(cx<T>* __restrict__ s, int ld,
                                             const cx<T>* __restrict__ tw, int tid) {
  fft_stages_wave<T, N, LINES, NT, DIR, 0, stage_count(ilog2(N)), 1>(s, ld, tw, tid);
}

// fft_lds_wave and one barrier where every wave has a line (LINES >= NT / 64)
// and a line fills a wave (N >= 64), else fft_lds: either way the lines are
// complete for every thread on return.  Above N = 256 only in FP64 (measured at
// 512x257x512: FP64 k_xprod 3.27 -> 3.12 ms and k_fwdz_nl 1.04 -> 0.91 ms, but
// float k_xprod 1.37 -> 1.41 and k_fwdz_nl 0.73 -> 0.82).
template <typename T, int N, int LINES, int NT, int DIR>
__device__ __forceinline__ void fft_lds_sync(cx<T>* __restrict__ s, int ld,
                                             const cx<T>* __restrict__ tw, int tid) {
  if constexpr (N >= 64 && NT % 64 == 0 && LINES >= NT / 64 && (N <= 256 || sizeof(T) == 8)) {
    fft_lds_wave<T, N, LINES, NT, DIR>(s, ld, tw, tid);
    __syncthreads();
  } else {
    fft_lds<T, N, LINES, NT, DIR>(s, ld, tw, tid);
  }
}

// copy the N twiddles exp(-2 pi i m / N) from global memory into the kernel's LDS
// table (before the caller's first barrier)
template <typename T, int N, int NT>
__device__ __forceinline__ void stage_twiddles(cx<T>* __restrict__ stw, const cx<T>* __restrict__ tw, int tid) {
  for (int m = tid; m < N; m += NT) stw[m] = tw[m];
}

}  // namespace dgpu
