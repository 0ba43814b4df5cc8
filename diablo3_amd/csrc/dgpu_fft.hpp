// dgpu_fft.hpp — device-side complex type and batched LDS Stockham FFT for gfx950.
//
// The transform is the autosort Stockham formulation: stage s of radix R reads
// x[j + r*N/R] (r < R) for butterfly j, multiplies input r by w^(r*k) with
// k = j mod Ns, w = exp(-+2 pi i/(Ns*R)), applies a radix-R DFT and writes
// y[(j/Ns)*Ns*R + k + r*Ns]; Ns grows by R per stage and the output is in
// natural order.  Radix-4 stages first, one radix-2 stage when log2(N) is odd.
// Lines live in LDS at s[l*ld + n]; every thread of the block takes part, and a
// stage reads all of its butterflies into registers before any write, so one LDS
// buffer suffices (two barriers per stage).
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

// One Stockham stage of radix R over LINES lines of length N.
template <typename T, int N, int LINES, int NT, int DIR, int R>
__device__ __forceinline__ void fft_stage(cx<T>* __restrict__ s, int ld,
                                          const cx<T>* __restrict__ tw, int Ns, int tid) {
  constexpr int NB = LINES * (N / R);
  constexpr int PER = (NB + NT - 1) / NT;
  cx<T> v[PER][R];
  int ob[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    ob[p] = -1;
    if (NB % NT == 0 || b < NB) {
      const int l = b / (N / R);
      const int j = b - l * (N / R);
      cx<T>* line = s + l * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = line[j + r * (N / R)];
      const int k = j & (Ns - 1);
      const int tstep = N / (Ns * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[r * k * tstep];
        if constexpr (DIR > 0) w = conjg(w);
        v[p][r] = v[p][r] * w;
      }
      if constexpr (R == 4) {
        dft4<DIR>(v[p][0], v[p][1], v[p][2], v[p][3]);
      } else {
        cx<T> a = v[p][0], c = v[p][1];
        v[p][0] = a + c;
        v[p][1] = a - c;
      }
      ob[p] = l * ld + (j / Ns) * Ns * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    if (ob[p] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) s[ob[p] + r * Ns] = v[p][r];
    }
  }
  __syncthreads();
}

// Batched in-place FFT of LINES lines of length N (power of two, >= 2) in LDS.
// Caller must __syncthreads() after filling s; the routine ends with a barrier.
template <typename T, int N, int LINES, int NT, int DIR>
__device__ __forceinline__ void fft_lds(cx<T>* __restrict__ s, int ld,
                                        const cx<T>* __restrict__ tw, int tid) {
  constexpr int L2 = ilog2(N);
  constexpr int N4 = L2 / 2;
  int Ns = 1;
#pragma unroll
  for (int st = 0; st < N4; ++st) {
    fft_stage<T, N, LINES, NT, DIR, 4>(s, ld, tw, Ns, tid);
    Ns *= 4;
  }
  if constexpr (L2 & 1) fft_stage<T, N, LINES, NT, DIR, 2>(s, ld, tw, Ns, tid);
}

}  // namespace dgpu
