// dgpu_kernels.hpp — HIP kernels of the RKW3 + CN channel step (gfx950).
//
// Device layouts (one rank, Nyp = ny):
//   Fourier  C[j][k][i]  complex double, i = 0..Nkx (kxr), k = 0..2Nkz (kzr),
//                        j = 0..ny+1; planes are contiguous (pc = kxr*kzr).
//   x-spectra H[j][z][i] complex T: the x transform of physical line (j, z)
//                        restricted to i = 0..Nkx (ph = nz*kxr).
//   physical R[j][z][x]  T (pp = nz*nx).
// Every kernel indexes only inside these extents; grids are sized by the host to
// the same extents (the host checks them before launch).
#pragma once
#include "dgpu_fft.hpp"

namespace dgpu {

constexpr int MAXTH = 4;
constexpr int MAXF = 3 + MAXTH;         // u1, u2, u3, th_1..th_n
constexpr int MAXO = 6 + 2 * MAXTH;     // fused product spectra

struct Geo {
  int nx, nz, ny, ny2, nkx, nkz, kxr, kzr;
  long long pc, ph, pp;
};

template <typename P, int M>
struct Ptrs {
  P p[M];
};

// packed kz index p -> row of the length-nz z transform (for/fft.f90:293-305, 316-333)
__device__ __forceinline__ int zpos(const Geo& g, int p) { return p <= g.nkz ? p : p + g.nz - g.kzr; }
// row z of the z transform -> packed index, or -1 for the de-aliased rows
__device__ __forceinline__ int zpack(const Geo& g, int z) {
  return z <= g.nkz ? z : (z >= g.nz - g.nkz ? z - g.nz + g.kzr : -1);
}

// ---------------------------------------------------------------------------
// z transforms: inverse (packed C -> H) and forward (H -> packed C)
// ---------------------------------------------------------------------------
template <typename T, int NZ, int TI, int NT>
__global__ __launch_bounds__(NT) void k_inv_z(Ptrs<const cd*, MAXO> src, Ptrs<cx<T>*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  constexpr int LD = NZ + 1;
  __shared__ cx<T> s[TI * LD];
  const int f = blockIdx.z, j = j0 + blockIdx.y, i0 = blockIdx.x * TI, tid = threadIdx.x;
  const cd* __restrict__ C = src.p[f] + (long long)j * g.pc;
  for (int idx = tid; idx < TI * NZ; idx += NT) {
    const int ii = idx % TI, z = idx / TI, i = i0 + ii, p = zpack(g, z);
    cx<T> v = {T(0), T(0)};
    if (i < g.kxr && p >= 0) v = cvt<T>(C[(long long)p * g.kxr + i]);
    s[ii * LD + z] = v;
  }
  __syncthreads();
  fft_lds<T, NZ, TI, NT, +1>(s, LD, tw, tid);
  cx<T>* __restrict__ H = dst.p[f] + (long long)j * g.ph;
  for (int idx = tid; idx < TI * NZ; idx += NT) {
    const int ii = idx % TI, z = idx / TI, i = i0 + ii;
    if (i < g.kxr) H[(long long)z * g.kxr + i] = s[ii * LD + z];
  }
}

template <typename T, int NZ, int TI, int NT>
__global__ __launch_bounds__(NT) void k_fwd_z(Ptrs<const cx<T>*, MAXO> src, Ptrs<cd*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  constexpr int LD = NZ + 1;
  __shared__ cx<T> s[TI * LD];
  const int f = blockIdx.z, j = j0 + blockIdx.y, i0 = blockIdx.x * TI, tid = threadIdx.x;
  const cx<T>* __restrict__ H = src.p[f] + (long long)j * g.ph;
  for (int idx = tid; idx < TI * NZ; idx += NT) {
    const int ii = idx % TI, z = idx / TI, i = i0 + ii;
    s[ii * LD + z] = (i < g.kxr) ? H[(long long)z * g.kxr + i] : cx<T>{T(0), T(0)};
  }
  __syncthreads();
  fft_lds<T, NZ, TI, NT, -1>(s, LD, tw, tid);
  cd* __restrict__ C = dst.p[f] + (long long)j * g.pc;
  const double rnz = 1.0 / NZ;
  for (int idx = tid; idx < TI * g.kzr; idx += NT) {
    const int ii = idx % TI, p = idx / TI, i = i0 + ii;
    if (i < g.kxr) {
      cd v = cvt<double>(s[ii * LD + zpos(g, p)]);
      C[(long long)p * g.kxr + i] = rnz * v;
    }
  }
}

// ---------------------------------------------------------------------------
// x transforms, two real lines (z, z+1) per complex transform
// ---------------------------------------------------------------------------
// Hermitian completion of the pair of half spectra A (line z), B (line z+1) into
// the length-NX spectrum of A + iB; bins i > Nkx are zero and the imaginary part
// of each DC bin is dropped (a real line's DC bin is real; FFTW's c2r ignores it).
template <typename T>
__device__ __forceinline__ cx<T> pair_spectrum(const cx<T>* __restrict__ A,
                                               const cx<T>* __restrict__ B, int n, int nx, int nkx) {
  if (n == 0) return {A[0].x, B[0].x};
  if (n <= nkx) {
    cx<T> a = A[n], b = B[n];
    return {a.x - b.y, a.y + b.x};
  }
  if (n >= nx - nkx) {
    cx<T> a = A[nx - n], b = B[nx - n];
    return {a.x + b.y, b.x - a.y};
  }
  return {T(0), T(0)};
}

// split the spectrum Z of a + i b into the half spectra of a and b at bin i
template <typename T>
__device__ __forceinline__ void pair_split(const cx<T>* __restrict__ Zl, int i, int nx, cx<T>& A, cx<T>& B) {
  const cx<T> zk = Zl[i], zm = conjg(Zl[(nx - i) & (nx - 1)]);
  const cx<T> sm = zk + zm, df = zk - zm;
  A = {T(0.5) * sm.x, T(0.5) * sm.y};
  B = {T(0.5) * df.y, T(-0.5) * df.x};
}

template <typename T, int NX, int ZP, int NT>
__global__ __launch_bounds__(NT) void k_inv_x(Ptrs<const cx<T>*, MAXO> src, Ptrs<T*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  __shared__ cx<T> s[ZP * NX];
  const int f = blockIdx.z, j = j0 + blockIdx.y, tid = threadIdx.x;
  const int zb = blockIdx.x * ZP * 2;
  const cx<T>* __restrict__ H = src.p[f] + (long long)j * g.ph;
  for (int idx = tid; idx < ZP * NX; idx += NT) {
    const int q = idx / NX, n = idx % NX, z = zb + 2 * q;
    s[idx] = pair_spectrum(H + (long long)z * g.kxr, H + (long long)(z + 1) * g.kxr, n, NX, g.nkx);
  }
  __syncthreads();
  fft_lds<T, NX, ZP, NT, +1>(s, NX, tw, tid);
  T* __restrict__ R = dst.p[f] + (long long)j * g.pp;
  for (int idx = tid; idx < ZP * NX; idx += NT) {
    const int q = idx / NX, n = idx % NX, z = zb + 2 * q;
    R[(long long)z * NX + n] = s[idx].x;
    R[(long long)(z + 1) * NX + n] = s[idx].y;
  }
}

template <typename T, int NX, int ZP, int NT>
__global__ __launch_bounds__(NT) void k_fwd_x(Ptrs<const T*, MAXO> src, Ptrs<cx<T>*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  __shared__ cx<T> s[ZP * NX];
  const int f = blockIdx.z, j = j0 + blockIdx.y, tid = threadIdx.x;
  const int zb = blockIdx.x * ZP * 2;
  const T* __restrict__ R = src.p[f] + (long long)j * g.pp;
  for (int idx = tid; idx < ZP * NX; idx += NT) {
    const int q = idx / NX, n = idx % NX, z = zb + 2 * q;
    s[idx] = {R[(long long)z * NX + n], R[(long long)(z + 1) * NX + n]};
  }
  __syncthreads();
  fft_lds<T, NX, ZP, NT, -1>(s, NX, tw, tid);
  cx<T>* __restrict__ H = dst.p[f] + (long long)j * g.ph;
  const T rnx = T(1) / T(NX);
  for (int idx = tid; idx < ZP * g.kxr; idx += NT) {
    const int q = idx / g.kxr, i = idx % g.kxr, z = zb + 2 * q;
    cx<T> A, B;
    pair_split(s + q * NX, i, NX, A, B);
    H[(long long)z * g.kxr + i] = rnx * A;
    H[(long long)(z + 1) * g.kxr + i] = rnx * B;
  }
}

// ---------------------------------------------------------------------------
// Fused DNS x pass: inverse x transform of u1, u2, u3, th at planes j-1, j, j+1,
// the nonlinear products of rk_chan_1 (for/solver.f90:655-815, 914-964) with their
// y stencils, forward x transform, and the x-derivative factors applied in
// x-spectral space.  Outputs, per plane j, the x-spectra of
//   N1 = -ikx F(u1u1) - F(flux1)     Z1 = F(u3u1)
//   N2 = -ikx F(u1u2i) - F(flux2)    Z2 = F(u3u2i)
//   N3 = -ikx F(u3u1) - F(flux3)     Z3 = F(u3u3)
//   Nt = -ikx F(th u1) - F(fluxth)   Zt = F(th u3)       (per scalar)
// so that the nonlinear tendency of equation e is Fz(Ne) - ikz Fz(Ze).
// ---------------------------------------------------------------------------
struct XProdArgs {
  Geo g;
  int jlo;
  int jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  const double* __restrict__ dy;
  const double* __restrict__ dyf;
  const double* __restrict__ kx;
  const void* hin[MAXF];   // x-spectra of u1, u2, u3, th_n (all planes)
  void* hout[MAXO];        // N1, N2, N3, Z1, Z2, Z3, (Nt, Zt) per scalar
  const void* tw;
};

template <typename T, int NX, int NT, int NTH>
__global__ __launch_bounds__(NT) void k_xprod(XProdArgs a) {
  constexpr int NF = 3 + NTH, NL = 3 * NF, NP = 8 + 3 * NTH, NO = 6 + 2 * NTH;
  static_assert(NP <= NL, "product lines overlay input lines");
  __shared__ cx<T> s[NL * NX];
  const Geo& g = a.g;
  const int tid = threadIdx.x, z = 2 * blockIdx.x, j = a.jlo + blockIdx.y;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  // inverse x transforms of (field f, plane j-1+d) into line 3f+d
  for (int idx = tid; idx < NL * NX; idx += NT) {
    const int l = idx / NX, n = idx % NX, f = l / 3, jj = j - 1 + (l - 3 * f);
    const cx<T>* H = (const cx<T>*)a.hin[f] + (long long)jj * g.ph;
    s[idx] = pair_spectrum(H + (long long)z * g.kxr, H + (long long)(z + 1) * g.kxr, n, NX, g.nkx);
  }
  __syncthreads();
  fft_lds<T, NX, NL, NT, +1>(s, NX, tw, tid);
  const bool vrow = (j >= a.jstart && j <= a.jend);
  const bool v2row = (j >= 2 && j <= g.ny);
  const T dyj = (T)a.dy[j], dyj1 = (T)a.dy[j + 1 <= g.ny + 1 ? j + 1 : j];
  const T dyfj = (T)a.dyf[j], dyfm = (T)a.dyf[j - 1];
  (void)dyj1;
  const T cint = T(1) / (T(2) * dyj);  // interpolation to GY(j): (dyf_j u_j + dyf_{j-1} u_{j-1})/(2 dy_j)
  for (int x = tid; x < NX; x += NT) {
    cx<T> U1[3], U2[3], U3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      U1[d] = s[(0 * 3 + d) * NX + x];
      U2[d] = s[(1 * 3 + d) * NX + x];
      U3[d] = s[(2 * 3 + d) * NX + x];
    }
    cx<T> TH[NTH > 0 ? NTH : 1][3];
#pragma unroll
    for (int n = 0; n < NTH; ++n)
#pragma unroll
      for (int d = 0; d < 3; ++d) TH[n][d] = s[((3 + n) * 3 + d) * NX + x];
    cx<T> P[NP];
    // the two real lines of the pair are the .x and .y components
#define DG_PROD(expr_x, expr_y) cx<T>{expr_x, expr_y}
#define C(v, c) (v).c
    if (vrow) {
      P[0] = DG_PROD(C(U3[1], x) * C(U1[1], x), C(U3[1], y) * C(U1[1], y));
      P[1] = DG_PROD(C(U1[1], x) * C(U1[1], x), C(U1[1], y) * C(U1[1], y));
      P[2] = DG_PROD(C(U3[1], x) * C(U3[1], x), C(U3[1], y) * C(U3[1], y));
      const T r2f = T(1) / (T(2) * dyfj);
      P[5] = DG_PROD((C(U1[2], x) * C(U2[2], x) + C(U1[1], x) * C(U2[2], x) - C(U1[1], x) * C(U2[1], x) -
                      C(U1[0], x) * C(U2[1], x)) * r2f,
                     (C(U1[2], y) * C(U2[2], y) + C(U1[1], y) * C(U2[2], y) - C(U1[1], y) * C(U2[1], y) -
                      C(U1[0], y) * C(U2[1], y)) * r2f);
      P[6] = DG_PROD((C(U3[2], x) * C(U2[2], x) + C(U3[1], x) * C(U2[2], x) - C(U3[1], x) * C(U2[1], x) -
                      C(U3[0], x) * C(U2[1], x)) * r2f,
                     (C(U3[2], y) * C(U2[2], y) + C(U3[1], y) * C(U2[2], y) - C(U3[1], y) * C(U2[1], y) -
                      C(U3[0], y) * C(U2[1], y)) * r2f);
    } else {
      P[0] = P[1] = P[2] = P[5] = P[6] = cx<T>{T(0), T(0)};
    }
    if (v2row) {
      P[3] = DG_PROD(((dyfj * C(U1[1], x) + dyfm * C(U1[0], x)) * cint) * C(U2[1], x),
                     ((dyfj * C(U1[1], y) + dyfm * C(U1[0], y)) * cint) * C(U2[1], y));
      P[4] = DG_PROD(((dyfj * C(U3[1], x) + dyfm * C(U3[0], x)) * cint) * C(U2[1], x),
                     ((dyfj * C(U3[1], y) + dyfm * C(U3[0], y)) * cint) * C(U2[1], y));
      const T apx = C(U2[1], x) + C(U2[2], x), amx = C(U2[1], x) + C(U2[0], x);
      const T apy = C(U2[1], y) + C(U2[2], y), amy = C(U2[1], y) + C(U2[0], y);
      P[7] = DG_PROD((T(0.25) * (apx * apx) - T(0.25) * (amx * amx)) / dyj,
                     (T(0.25) * (apy * apy) - T(0.25) * (amy * amy)) / dyj);
    } else {
      P[3] = P[4] = P[7] = cx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int n = 0; n < NTH; ++n) {
      const bool trow = (j >= a.jstart_th[n] && j <= a.jend_th[n]);
      if (trow) {
        const T r2f = T(1) / (T(2) * dyfj);
        P[8 + 3 * n] = DG_PROD(C(TH[n][1], x) * C(U1[1], x), C(TH[n][1], y) * C(U1[1], y));
        P[9 + 3 * n] = DG_PROD(C(TH[n][1], x) * C(U3[1], x), C(TH[n][1], y) * C(U3[1], y));
        P[10 + 3 * n] = DG_PROD(
            (C(TH[n][2], x) * C(U2[2], x) + C(TH[n][1], x) * C(U2[2], x) - C(TH[n][1], x) * C(U2[1], x) -
             C(TH[n][0], x) * C(U2[1], x)) * r2f,
            (C(TH[n][2], y) * C(U2[2], y) + C(TH[n][1], y) * C(U2[2], y) - C(TH[n][1], y) * C(U2[1], y) -
             C(TH[n][0], y) * C(U2[1], y)) * r2f);
      } else {
        P[8 + 3 * n] = P[9 + 3 * n] = P[10 + 3 * n] = cx<T>{T(0), T(0)};
      }
    }
#undef DG_PROD
#undef C
    // each thread rewrites only column x, which only it has read
#pragma unroll
    for (int p = 0; p < NP; ++p) s[p * NX + x] = P[p];
  }
  __syncthreads();
  fft_lds<T, NX, NP, NT, -1>(s, NX, tw, tid);
  const T rnx = T(1) / T(NX);
  for (int idx = tid; idx < NO * g.kxr; idx += NT) {
    const int o = idx / g.kxr, i = idx % g.kxr;
    const T kxi = (T)a.kx[i];
    cx<T> A0, B0, A1, B1;
    int pd, pf;  // derivative product (times -ikx) and flux product; pf < 0: plain copy of pd
    if (o == 0) { pd = 1; pf = 5; }
    else if (o == 1) { pd = 3; pf = 7; }
    else if (o == 2) { pd = 0; pf = 6; }
    else if (o == 3) { pd = 0; pf = -1; }
    else if (o == 4) { pd = 4; pf = -1; }
    else if (o == 5) { pd = 2; pf = -1; }
    else {
      const int n = (o - 6) >> 1;
      if (((o - 6) & 1) == 0) { pd = 8 + 3 * n; pf = 10 + 3 * n; }
      else { pd = 9 + 3 * n; pf = -1; }
    }
    pair_split(s + pd * NX, i, NX, A0, B0);
    cx<T> ra, rb;
    if (pf >= 0) {
      pair_split(s + pf * NX, i, NX, A1, B1);
      // -ikx F - Fflux
      ra = cx<T>{kxi * A0.y, -(kxi * A0.x)} - A1;
      rb = cx<T>{kxi * B0.y, -(kxi * B0.x)} - B1;
    } else {
      ra = A0;
      rb = B0;
    }
    cx<T>* H = (cx<T>*)a.hout[o] + (long long)j * g.ph;
    H[(long long)z * g.kxr + i] = rnx * ra;
    H[(long long)(z + 1) * g.kxr + i] = rnx * rb;
  }
}

// ---------------------------------------------------------------------------
// Fused z pass + Fourier-space assembly of rk_chan_1 (for/solver.f90:385-838,
// 909-973 and user_rhs_chan_fourier 35-111) for one plane j and a tile of kx.
// Writes the new RK memory cf and the full right-hand side cr including the
// explicit half of Crank–Nicolson (for/solver.f90:850-870, 981-1002).
// ---------------------------------------------------------------------------
struct AsmArgs {
  Geo g;
  int jlo, rk;
  int jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  int front, f_type;
  double nu, beta, nu_v_scale, ro_inv, delta, grav_x, grav_y, grav_z, ly, px0, omega0, amp_omega0, time;
  double pr[MAXTH], dthdx[MAXTH], dthdz[MAXTH];
  double temp1, temp3, temp4, temp5;
  const double *__restrict__ dy, *__restrict__ dyf, *__restrict__ gy, *__restrict__ gyf;
  const double *__restrict__ kx, *__restrict__ kz;
  const cd* u[MAXF];     // cu1, cu2, cu3, cth_n (state at the start of the substep)
  const cd* p;           // cp
  cd* cf[MAXF];          // RK memory (read when rk > 1, then overwritten)
  cd* cr[MAXF];          // right-hand side
  const void* h[MAXO];   // product x-spectra from k_xprod
  const void* tw;
};

template <typename T, int NZ, int TI, int NT, int NTH>
__global__ __launch_bounds__(NT) void k_fwdz_asm(AsmArgs a) {
  constexpr int LD = NZ + 1;
  __shared__ cx<T> s[2 * TI * LD];
  const Geo& g = a.g;
  const int tid = threadIdx.x, j = a.jlo + blockIdx.y, i0 = blockIdx.x * TI;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const long long pc = g.pc;
  const double rnz = 1.0 / NZ;
  const double dyj = a.dy[j], dyfj = a.dyf[j];
  const double cori = a.ro_inv / a.delta;
  constexpr int NE = 3 + NTH;
  for (int e = 0; e < NE; ++e) {
    // e: 0 = u1, 1 = u2, 2 = u3, 3+n = th_n ; product outputs per equation
    int oN, oZ;
    bool valid;
    if (e == 0) { oN = 0; oZ = 3; valid = (j >= a.jstart && j <= a.jend); }
    else if (e == 1) { oN = 1; oZ = 4; valid = (j >= 2 && j <= g.ny); }
    else if (e == 2) { oN = 2; oZ = 5; valid = (j >= a.jstart && j <= a.jend); }
    else {
      const int n = e - 3;
      oN = 6 + 2 * n; oZ = 7 + 2 * n;
      valid = (j >= a.jstart_th[n] && j <= a.jend_th[n]);
    }
    if (!valid) continue;  // uniform over the block
    const cx<T>* HN = (const cx<T>*)a.h[oN] + (long long)j * g.ph;
    const cx<T>* HZ = (const cx<T>*)a.h[oZ] + (long long)j * g.ph;
    for (int idx = tid; idx < 2 * TI * NZ; idx += NT) {
      const int ii = idx % TI, r = idx / TI, which = r / NZ, z = r - which * NZ, i = i0 + ii;
      const cx<T>* Hs = which ? HZ : HN;
      s[(which * TI + ii) * LD + z] = (i < g.kxr) ? Hs[(long long)z * g.kxr + i] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_lds<T, NZ, 2 * TI, NT, -1>(s, LD, tw, tid);
    for (int idx = tid; idx < TI * g.kzr; idx += NT) {
      const int ii = idx % TI, k = idx / TI, i = i0 + ii;
      if (i >= g.kxr) continue;
      const int zp = zpos(g, k);
      const cd fn = cvt<double>(s[ii * LD + zp]), fz = cvt<double>(s[(TI + ii) * LD + zp]);
      const double kxi = a.kx[i], kzk = a.kz[k];
      const cd nl = rnz * (fn - ik(kzk, fz));
      const double k2 = a.kx[i] * a.kx[i] + a.kz[k] * a.kz[k];
      const double kk = (a.beta == 1.0) ? k2 : pow(k2, a.beta);
      const long long o = (long long)j * pc + (long long)k * g.kxr + i;
      const long long om = o - pc, op = o + pc;
      cd cf, cr;
      if (e == 0 || e == 2) {
        const cd* U = a.u[e];
        const cd* V = a.u[2 - e];  // Coriolis partner
        const cd uc = U[o];
        cf = -1.0 * ((a.nu * kk) * uc);
        if (e == 0) cf = cf + cori * V[o];
        else cf = cf - cori * V[o];
        for (int n = 0; n < NTH; ++n) cf = cf + (e == 0 ? a.grav_x : a.grav_z) * a.u[3 + n][o];
        if (a.front && a.ro_inv != 0.0) {
          const cd u2m = 0.5 * (a.u[1][o] + a.u[1][op]);
          const double yo = a.gyf[j] - 0.5 * a.ly;
          for (int n = 0; n < NTH; ++n) {
            cf = cf - (a.dthdx[n] * yo * a.delta / a.ro_inv) * ik(kzk, uc) -
                 (-1.0 * a.dthdz[n] * yo * a.delta / a.ro_inv) * ik(kxi, uc);
            if (e == 0) cf = cf - (-1.0 * a.dthdz[n] * a.delta / a.ro_inv) * u2m;
            else cf = cf - (a.dthdx[n] * a.delta / a.ro_inv) * u2m;
          }
        }
        cf = cf + nl;
        cr = uc;
        if (a.rk > 1) cr = cr + a.temp3 * a.cf[e][o];
        cr = cr - a.temp4 * (e == 0 ? ik(kxi, a.p[o]) : ik(kzk, a.p[o]));
        if (e == 0 && k == 0 && i == 0) {
          if (a.f_type == 1) cr.x = cr.x - a.temp4 * a.px0;
          else if (a.f_type == 2) cr.x = cr.x - a.temp4 * (a.px0 + a.amp_omega0 * cos(a.omega0 * a.time));
        }
        cr = cr + a.temp5 * cf;
        const cd up = U[op], um = U[om];
        const double dyp = a.dy[j + 1];
        cr = cr + (a.temp1 * a.nu_v_scale) * ((1.0 / dyfj) * ((1.0 / dyp) * (up - uc) - (1.0 / dyj) * (uc - um)));
      } else if (e == 1) {
        const cd* U = a.u[1];
        const cd uc = U[o];
        cf = -1.0 * ((a.nu * kk) * uc);
        if (!(k == 0 && i == 0)) {
          for (int n = 0; n < NTH; ++n) {
            const cd* TH = a.u[3 + n];
            cf = cf + a.grav_y * ((1.0 / (2.0 * dyj)) * (a.dyf[j - 1] * TH[o] + dyfj * TH[om]));
          }
        }
        if (a.front && a.ro_inv != 0.0) {
          const double yo = a.gy[j] - 0.5 * a.ly;
          for (int n = 0; n < NTH; ++n)
            cf = cf - (a.dthdx[n] * yo * a.delta / a.ro_inv) * ik(kzk, uc) -
                 (-1.0 * a.dthdz[n] * yo * a.delta / a.ro_inv) * ik(kxi, uc);
        }
        cf = cf + nl;
        cr = uc;
        if (a.rk > 1) cr = cr + a.temp3 * a.cf[1][o];
        cr = cr - a.temp4 * ((1.0 / dyj) * (a.p[o] - a.p[om]));
        cr = cr + a.temp5 * cf;
        const cd up = U[op], um = U[om];
        cr = cr + (a.temp1 * a.nu_v_scale) *
                      ((1.0 / dyj) * ((1.0 / dyfj) * (up - uc) - (1.0 / a.dyf[j - 1]) * (uc - um)));
      } else {
        const int n = e - 3;
        const cd* TH = a.u[e];
        const cd tc = TH[o];
        cf = -1.0 * (((a.nu / a.pr[n]) * kk) * tc);
        cf = cf - a.dthdx[n] * a.u[0][o] - a.dthdz[n] * a.u[2][o];
        if (a.front && a.ro_inv != 0.0) {
          const double yo = a.gyf[j] - 0.5 * a.ly;
          cf = cf - ((a.delta / a.ro_inv) * a.dthdx[n] * yo) * ik(kzk, tc) -
               ((a.delta / a.ro_inv) * (-1.0) * a.dthdz[n] * yo) * ik(kxi, tc);
        }
        cf = cf + nl;
        cr = tc;
        if (a.rk > 1) cr = cr + a.temp3 * a.cf[e][o];
        cr = cr + a.temp5 * cf;
        const cd up = TH[op], um = TH[om];
        const double dyp = a.dy[j + 1];
        cr = cr + ((a.temp1 / a.pr[n]) * a.nu_v_scale) *
                      ((1.0 / dyfj) * ((1.0 / dyp) * (up - tc) - (1.0 / dyj) * (tc - um)));
      }
      a.cf[e][o] = cf;
      a.cr[e][o] = cr;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Column kernel: per (kx, kz) column, the Crank–Nicolson solves of every scalar
// and of u2, u1, u3 (for/solver.f90:1004-1233, in Fourier space: the DNS
// operator does not depend on x or z), then rem_div_chan (for/solver.f90:
// 1302-1379) and the pressure update (for/solver.f90:1261-1267).
// ---------------------------------------------------------------------------
struct SolveRows {
  int blo, bhi;      // rows filled from cr (the build range)
  int wlo, whi;      // rows written back
  int kind[5];       // rows 0, 1, 2, ny, ny+1: -1 = from the build range rule, 0 = zero, 1 = BC value
  double val[5];     // BC value of the mean mode at those rows
};

struct ColArgs {
  Geo g;
  int ncols;
  int jstart, jend;
  double temp4;
  SolveRows rows[MAXF];            // 0 = u1, 1 = u2, 2 = u3, 3+n = th_n
  const double* __restrict__ fac;  // [field][3][ny2]: l, c, 1/b of the LU factors
  const double *__restrict__ dy, *__restrict__ dyf, *__restrict__ kx, *__restrict__ kz;
  cd* u[MAXF];
  cd* p;
  const cd* cr[MAXF];
  int do_cn;                       // 0: projection only (rem_div_chan)
};

__device__ __forceinline__ int special_row(int j, int ny) {
  return j <= 2 ? j : (j == ny ? 3 : (j == ny + 1 ? 4 : -1));
}

template <int TC, int NTH>
__global__ __launch_bounds__(64) void k_columns(ColArgs a) {
  extern __shared__ unsigned char smem[];
  const Geo& g = a.g;
  const int ny2 = g.ny2, ny = g.ny, tid = threadIdx.x;
  const int c0 = blockIdx.x * TC;
  cd* sU1 = (cd*)smem;
  cd* sU2 = sU1 + TC * ny2;
  cd* sU3 = sU2 + TC * ny2;
  cd* sX = sU3 + TC * ny2;
  double* sB = (double*)(sX + TC * ny2);
  const long long pc = g.pc;
  cd* slots[3] = {sU1, sU2, sU3};

  auto load_rhs = [&](int F, cd* dst) {
    const SolveRows& R = a.rows[F];
    const cd* src = a.cr[F];
    for (int idx = tid; idx < TC * ny2; idx += 64) {
      const int cc = idx % TC, j = idx / TC, c = c0 + cc;
      cd v = {0.0, 0.0};
      if (c < a.ncols) {
        const int sr = special_row(j, ny);
        const int kd = sr >= 0 ? R.kind[sr] : -1;
        if (kd == 1) {
          if (c == 0) v = {R.val[sr], 0.0};
        } else if (kd == -1) {
          if (j >= R.blo && j <= R.bhi) v = src[(long long)j * pc + c];
        }
      }
      dst[cc * ny2 + j] = v;
    }
  };
  auto cn_sweep = [&](int F, cd* col) {
    const double* L = a.fac + (long long)F * 3 * ny2;
    const double* U = L + ny2;
    const double* RB = U + ny2;
    cd prev = col[0];
    for (int j = 1; j < ny2; ++j) {
      const cd v = col[j] + L[j] * prev;
      col[j] = v;
      prev = v;
    }
    cd nxt = RB[ny2 - 1] * col[ny2 - 1];
    col[ny2 - 1] = nxt;
    for (int j = ny2 - 2; j >= 0; --j) {
      const cd v = RB[j] * (col[j] - U[j] * nxt);
      col[j] = v;
      nxt = v;
    }
  };
  auto store_rows = [&](int F, const cd* src, cd* dst) {
    const SolveRows& R = a.rows[F];
    for (int idx = tid; idx < TC * ny2; idx += 64) {
      const int cc = idx % TC, j = idx / TC, c = c0 + cc;
      if (c < a.ncols && j >= R.wlo && j <= R.whi) dst[(long long)j * pc + c] = src[cc * ny2 + j];
    }
  };

  if (a.do_cn) {
    // scalars first, with the velocity of the previous substep (for/solver.f90:1004-1007)
    for (int n = 0; n < NTH; ++n) {
      load_rhs(3 + n, sX);
      __syncthreads();
      if (tid < TC) cn_sweep(3 + n, sX + tid * ny2);
      __syncthreads();
      store_rows(3 + n, sX, a.u[3 + n]);
      __syncthreads();
    }
    for (int F = 0; F < 3; ++F) load_rhs(F, slots[F]);
    __syncthreads();
    if (tid < 3 * TC) {
      const int F = tid / TC, cc = tid - F * TC;
      cn_sweep(F, slots[F] + cc * ny2);
    }
    __syncthreads();
  } else {
    for (int F = 0; F < 3; ++F)
      for (int idx = tid; idx < TC * ny2; idx += 64) {
        const int cc = idx % TC, j = idx / TC, c = c0 + cc;
        slots[F][cc * ny2 + j] = (c < a.ncols) ? a.u[F][(long long)j * pc + c] : cd{0.0, 0.0};
      }
    __syncthreads();
  }
  // rem_div_chan: right-hand side rows 1..ny, BC rows 1 and ny (for/solver.f90:1325-1350, 2330-2371)
  for (int idx = tid; idx < TC * ny2; idx += 64) {
    const int cc = idx % TC, j = idx / TC, c = c0 + cc;
    cd v = {0.0, 0.0};
    if (c < a.ncols && j >= 2 && j <= ny - 1) {
      const int i = c % g.kxr, k = c / g.kxr;
      const cd* u1 = sU1 + cc * ny2;
      const cd* u2 = sU2 + cc * ny2;
      const cd* u3 = sU3 + cc * ny2;
      v = ik(a.kx[i], u1[j]) + (1.0 / a.dyf[j]) * (u2[j + 1] - u2[j]) + ik(a.kz[k], u3[j]);
    }
    sX[cc * ny2 + j] = v;
  }
  __syncthreads();
  if (tid < TC && c0 + tid < a.ncols) {
    // complex Thomas over rows 1..ny (thomas_forward/backward_complex_mpi at one rank)
    const int c = c0 + tid, i = c % g.kxr, k = c / g.kxr;
    cd* phi = sX + tid * ny2;
    double* bp = sB + tid * ny2;
    const double k2 = -(a.kx[i] * a.kx[i]) - a.kz[k] * a.kz[k];
    // row 1
    double bprev = 1.0, cprev = (c == 0) ? 0.0 : -1.0;
    bp[1] = bprev;
    cd gprev = phi[1];
    for (int j = 2; j <= ny; ++j) {
      double aj, bj, cj;
      if (j < ny) {
        aj = 1.0 / (a.dy[j] * a.dyf[j]);
        bj = k2 - 1.0 / (a.dy[j + 1] * a.dyf[j]) - 1.0 / (a.dy[j] * a.dyf[j]);
        cj = 1.0 / (a.dy[j + 1] * a.dyf[j]);
      } else {
        aj = 1.0; bj = -1.0; cj = 0.0;
      }
      const double m = -aj / bprev;
      bj = bj + m * cprev;
      const cd gj = phi[j] + m * gprev;
      phi[j] = gj;
      bp[j] = bj;
      bprev = bj; cprev = cj; gprev = gj;
    }
    cd nxt = (1.0 / bp[ny]) * phi[ny];
    phi[ny] = nxt;
    for (int j = ny - 1; j >= 1; --j) {
      const double cj = (j == 1) ? ((c == 0) ? 0.0 : -1.0) : 1.0 / (a.dy[j + 1] * a.dyf[j]);
      const cd v = (1.0 / bp[j]) * (phi[j] - cj * nxt);
      phi[j] = v;
      nxt = v;
    }
  }
  __syncthreads();
  // velocity correction and pressure update
  for (int idx = tid; idx < TC * ny2; idx += 64) {
    const int cc = idx % TC, j = idx / TC, c = c0 + cc;
    if (c >= a.ncols) continue;
    const int i = c % g.kxr, k = c / g.kxr;
    const cd* phi = sX + cc * ny2;
    if (j >= 1 && j <= ny) {
      sU1[cc * ny2 + j] = sU1[cc * ny2 + j] - ik(a.kx[i], phi[j]);
      sU3[cc * ny2 + j] = sU3[cc * ny2 + j] - ik(a.kz[k], phi[j]);
    }
    if (j >= 2 && j <= ny) sU2[cc * ny2 + j] = sU2[cc * ny2 + j] - (1.0 / a.dy[j]) * (phi[j] - phi[j - 1]);
    if (a.do_cn && j >= a.jstart && j <= a.jend) {
      const long long o = (long long)j * pc + c;
      a.p[o] = a.p[o] + (1.0 / a.temp4) * phi[j];
    }
  }
  __syncthreads();
  if (a.do_cn) {
    for (int F = 0; F < 3; ++F) store_rows(F, slots[F], a.u[F]);
  } else {
    for (int F = 0; F < 3; ++F)
      for (int idx = tid; idx < TC * ny2; idx += 64) {
        const int cc = idx % TC, j = idx / TC, c = c0 + cc;
        if (c < a.ncols && j >= 1 && j <= ny) a.u[F][(long long)j * pc + c] = slots[F][cc * ny2 + j];
      }
  }
}

}  // namespace dgpu
