// dgpu_kernels.hpp — HIP kernels of the RKW3 + CN channel step (gfx950).
//
// Device layouts (one rank, Nyp = ny):
//   Fourier  C[j][k][i]  complex double, i = 0..Nkx active (kxr of them) in rows of
//                        pitch kxp = kxr rounded up to 8 (128-byte rows for complex
//                        double), k = 0..2Nkz (kzr), j = 0..ny+1; pc = kxp*kzr.
//   x-spectra H[j][z][i] complex T: the x transform of physical line (j, z)
//                        restricted to i = 0..Nkx, same pitch kxp (ph = nz*kxp).
//   physical R[j][z][x]  T (pp = nz*nx).
// Padding columns i = kxr..kxp-1 are zero and never read as modes.
// Every kernel indexes only inside these extents; grids are sized by the host to
// the same extents (the host checks them before launch).
#pragma once
#include "dgpu_fft.hpp"

namespace dgpu {

constexpr int MAXTH = 4;
constexpr int MAXF = 3 + MAXTH;         // u1, u2, u3, th_1..th_n
constexpr int MAXO = 3 * MAXF;         // largest transform batch (LES: v, dv/dx, dv/dz)

struct Geo {
  int nx, nz, ny, ny2, nkx, nkz, kxr, kzr, kxp;
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

// Tile order for the 8 XCDs: the dispatcher deals blocks (in linear order) to
// the XCDs in turn, so within each group of 8 K consecutive blocks XCD x is
// given the K consecutive tiles x K .. x K + K - 1.  Neighbouring column blocks
// of a z transform each read half of the same 128-byte line of every z row;
// ordered this way they run on one XCD together and the second half comes from
// its L2.  Blocks past the last whole group keep their own index.
__device__ __forceinline__ int xcd_tile(int b, int total, int K) {
  const int C = 8 * K;
  if (b >= (total / C) * C) return b;
  return (b / C) * C + (b & 7) * K + (b % C) / 8;
}
constexpr int XCD_K = 16;

// ---------------------------------------------------------------------------
// z transforms: inverse (packed C -> H) and forward (H -> packed C)
// ---------------------------------------------------------------------------
// Field f of a batch with bit f of dzmask set is multiplied by i kz before the
// transform (the x-spectrum of its z derivative).
template <typename T, int NZ, int TI, int NT>
__global__ __launch_bounds__(NT) void k_inv_z(Ptrs<const cd*, MAXO> src, Ptrs<cx<T>*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw,
                                              const double* __restrict__ kz, unsigned dzmask) {
  constexpr int LD = NZ + 1;
  __shared__ cx<T> s[TI * LD];
  __shared__ cx<T> stw[NZ];  // twiddles
  stage_twiddles<T, NZ, NT>(stw, tw, threadIdx.x);
  // grid (column blocks, planes, fields), taken in xcd_tile order
  const int nib = gridDim.x, nj = gridDim.y;
  const int tt = xcd_tile(blockIdx.x + nib * (blockIdx.y + nj * blockIdx.z), nib * nj * gridDim.z, XCD_K);
  const int f = tt / (nib * nj), j = j0 + (tt / nib) % nj, i0 = (tt % nib) * TI, tid = threadIdx.x;
  const cd* __restrict__ C = src.p[f] + (long long)j * g.pc;
  const bool dz = (dzmask >> f) & 1u;
  // every load of the thread first, then the LDS stores (the loads overlap)
  constexpr int PER = (TI * NZ + NT - 1) / NT;
  cx<T> v[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    v[q] = {T(0), T(0)};
    if (PER * NT == TI * NZ || idx < TI * NZ) {
      const int ii = idx % TI, z = idx / TI, i = i0 + ii, p = zpack(g, z);
      if (i < g.kxr && p >= 0) {
        const cd c = C[(long long)p * g.kxp + i];
        v[q] = cvt<T>(dz ? ik(kz[p], c) : c);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if (PER * NT == TI * NZ || idx < TI * NZ) s[(idx % TI) * LD + swz<T>(idx / TI)] = v[q];
  }
  __syncthreads();
  fft_lds<T, NZ, TI, NT, +1>(s, LD, stw, tid);
  cx<T>* __restrict__ H = dst.p[f] + (long long)j * g.ph;
  for (int idx = tid; idx < TI * NZ; idx += NT) {
    const int ii = idx % TI, z = idx / TI, i = i0 + ii;
    if (i < g.kxr) H[(long long)z * g.kxp + i] = s[ii * LD + swz<T>(z)];
  }
}

// ACC: add the transform into the destination (an increment of the state)
template <typename T, int NZ, int TI, int NT, bool ACC = false>
__global__ __launch_bounds__(NT) void k_fwd_z(Ptrs<const cx<T>*, MAXO> src, Ptrs<cd*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  constexpr int LD = NZ + 1;
  __shared__ cx<T> s[TI * LD];
  __shared__ cx<T> stw[NZ];  // twiddles
  stage_twiddles<T, NZ, NT>(stw, tw, threadIdx.x);
  // grid (column blocks, planes, fields), taken in xcd_tile order
  const int nib = gridDim.x, nj = gridDim.y;
  const int tt = xcd_tile(blockIdx.x + nib * (blockIdx.y + nj * blockIdx.z), nib * nj * gridDim.z, XCD_K);
  const int f = tt / (nib * nj), j = j0 + (tt / nib) % nj, i0 = (tt % nib) * TI, tid = threadIdx.x;
  const cx<T>* __restrict__ H = src.p[f] + (long long)j * g.ph;
  for (int idx = tid; idx < TI * NZ; idx += NT) {
    const int ii = idx % TI, z = idx / TI, i = i0 + ii;
    s[ii * LD + swz<T>(z)] = (i < g.kxr) ? H[(long long)z * g.kxp + i] : cx<T>{T(0), T(0)};
  }
  __syncthreads();
  fft_lds<T, NZ, TI, NT, -1>(s, LD, stw, tid);
  cd* __restrict__ C = dst.p[f] + (long long)j * g.pc;
  const double rnz = 1.0 / NZ;
  for (int idx = tid; idx < TI * g.kzr; idx += NT) {
    const int ii = idx % TI, p = idx / TI, i = i0 + ii;
    if (i < g.kxr) {
      cd v = cvt<double>(s[ii * LD + swz<T>(zpos(g, p))]);
      if constexpr (ACC) C[(long long)p * g.kxp + i] = C[(long long)p * g.kxp + i] + rnz * v;
      else C[(long long)p * g.kxp + i] = rnz * v;
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
  const cx<T> zk = Zl[swz<T>(i)], zm = conjg(Zl[swz<T>((nx - i) & (nx - 1))]);
  const cx<T> sm = zk + zm, df = zk - zm;
  A = {T(0.5) * sm.x, T(0.5) * sm.y};
  B = {T(0.5) * df.y, T(-0.5) * df.x};
}

// Field f of a batch with bit f of dxmask set yields its x derivative.
template <typename T, int NX, int ZP, int NT>
__global__ __launch_bounds__(NT) void k_inv_x(Ptrs<const cx<T>*, MAXO> src, Ptrs<T*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw,
                                              const double* __restrict__ kx, unsigned dxmask) {
  __shared__ cx<T> s[ZP * NX];
  __shared__ cx<T> stw[NX];  // twiddles
  stage_twiddles<T, NX, NT>(stw, tw, threadIdx.x);
  const int f = blockIdx.z, j = j0 + blockIdx.y, tid = threadIdx.x;
  const int zb = blockIdx.x * ZP * 2;
  const cx<T>* __restrict__ H = src.p[f] + (long long)j * g.ph;
  const bool dx = (dxmask >> f) & 1u;
  // Each half-spectrum bin m = 0..NX/2 of the two lines is loaded once and gives
  // both bins m and NX - m of the pair spectrum (pair_spectrum written out for the
  // two halves; for an x derivative the bins times i kx, DC zero); every load of
  // the thread first, then the LDS stores (the loads overlap).
  constexpr int HB = NX / 2 + 1;
  constexpr int PER = (ZP * HB + NT - 1) / NT;
  cx<T> lo[PER], hi[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int idx = tid + p * NT;
    if (idx < ZP * HB) {
      const int q = idx / HB, m = idx % HB, z = zb + 2 * q;
      const cx<T>* A = H + (long long)z * g.kxp;
      const cx<T>* B = H + (long long)(z + 1) * g.kxp;
      cx<T> l = {T(0), T(0)}, u = {T(0), T(0)};
      if (m == 0) {
        if (!dx) l = {A[0].x, B[0].x};  // DC: the imaginary parts dropped; d/dx of DC is 0
      } else if (m <= g.nkx) {
        cx<T> a = A[m], b = B[m];
        if (dx) {
          const T k = (T)kx[m];
          a = ik(k, a);
          b = ik(k, b);
        }
        l = {a.x - b.y, a.y + b.x};
        u = {a.x + b.y, b.x - a.y};
      }
      lo[p] = l;
      hi[p] = u;
    }
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int idx = tid + p * NT;
    if (idx < ZP * HB) {
      const int q = idx / HB, m = idx % HB;
      s[q * NX + swz<T>(m)] = lo[p];
      if (m > 0 && m < NX / 2) s[q * NX + swz<T>(NX - m)] = hi[p];
    }
  }
  __syncthreads();
  fft_lds_sync<T, NX, ZP, NT, +1>(s, NX, stw, tid);
  T* __restrict__ R = dst.p[f] + (long long)j * g.pp;
  for (int idx = tid; idx < ZP * NX; idx += NT) {
    const int q = idx / NX, n = idx % NX, z = zb + 2 * q;
    const cx<T> v = s[q * NX + swz<T>(n)];
    R[(long long)z * NX + n] = v.x;
    R[(long long)(z + 1) * NX + n] = v.y;
  }
}

template <typename T, int NX, int ZP, int NT>
__global__ __launch_bounds__(NT) void k_fwd_x(Ptrs<const T*, MAXO> src, Ptrs<cx<T>*, MAXO> dst,
                                              Geo g, int j0, const cx<T>* __restrict__ tw) {
  __shared__ cx<T> s[ZP * NX];
  __shared__ cx<T> stw[NX];  // twiddles
  stage_twiddles<T, NX, NT>(stw, tw, threadIdx.x);
  const int f = blockIdx.z, j = j0 + blockIdx.y, tid = threadIdx.x;
  const int zb = blockIdx.x * ZP * 2;
  const T* __restrict__ R = src.p[f] + (long long)j * g.pp;
  for (int idx = tid; idx < ZP * NX; idx += NT) {
    const int q = idx / NX, n = idx % NX, z = zb + 2 * q;
    s[q * NX + swz<T>(n)] = {R[(long long)z * NX + n], R[(long long)(z + 1) * NX + n]};
  }
  __syncthreads();
  fft_lds_sync<T, NX, ZP, NT, -1>(s, NX, stw, tid);
  cx<T>* __restrict__ H = dst.p[f] + (long long)j * g.ph;
  const T rnx = T(1) / T(NX);
  for (int idx = tid; idx < ZP * g.kxr; idx += NT) {
    const int q = idx / g.kxr, i = idx % g.kxr, z = zb + 2 * q;
    cx<T> A, B;
    pair_split(s + q * NX, i, NX, A, B);
    H[(long long)z * g.kxp + i] = rnx * A;
    H[(long long)(z + 1) * g.kxp + i] = rnx * B;
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
// Per-row coefficients of the assembly, computed once on the host from the grid
// (the reference evaluates them inline; the host table holds the same quotients).
enum RowCoef {
  RC_RDY = 0,    // 1/dy(j)
  RC_RDYP,       // 1/dy(j+1)
  RC_RDYF,       // 1/dyf(j)
  RC_RDYFM,      // 1/dyf(j-1)
  RC_WTH,        // dyf(j-1)/(2 dy(j)): weight of th(j) in the GY interpolation
  RC_WTHM,       // dyf(j)/(2 dy(j)):   weight of th(j-1)
  RC_FLO,        // 1/(dy(j) dyf(j)):    lower coupling of a GYF row (u1, u3, scalars)
  RC_FHI,        // 1/(dy(j+1) dyf(j)):  upper coupling of a GYF row
  RC_GLO,        // 1/(dyf(j-1) dy(j)):  lower coupling of a GY row (u2)
  RC_GHI,        // 1/(dyf(j) dy(j)):    upper coupling of a GY row
  RC_YX,         // 2 dy(j) / (dx beta_sgs):  the AMD closure's width ratios at GY j
  RC_XY,         // (dx beta_sgs) / (2 dy(j))
  RC_YZ,         // 2 dy(j) / (dz beta_sgs)
  RC_ZY,         // (dz beta_sgs) / (2 dy(j))
  RC_TW0,        // per scalar n, 4 entries at RC_TW0 + 4n:
  RC_COUNT = RC_TW0 + 4 * MAXTH
  // +0 dthdx (gyf-Ly/2) delta/Ro_inv, +1 -dthdz (gyf-Ly/2) delta/Ro_inv,
  // +2 dthdx (gy-Ly/2) delta/Ro_inv,  +3 -dthdz (gy-Ly/2) delta/Ro_inv
};

struct XProdArgs {
  Geo g;
  int jlo;
  int jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  const double* __restrict__ dy;
  const double* __restrict__ dyf;
  const double* __restrict__ rowc;  // [ny2][RC_COUNT] grid quotients
  const double* __restrict__ kx;
  const void* hin[MAXF];   // x-spectra of u1, u2, u3, th_n (all planes)
  void* hout[MAXO];        // N1, N2, N3, Z1, Z2, Z3, (Nt, Zt) per scalar
  const void* tw;
  // LES: physical SGS fields added to products 1, 0, 2, 5, 6, 3, 4, (8, 9) + 3n
  // (tau_11, tau_13, tau_33, the two stress fluxes, the two GY stresses, the
  // scalar fluxes); null for DNS
  const void* aug[7 + 2 * MAXTH];
  // the fields in physical space ([j][z][x], T), all planes, written by the
  // inverse x transform; the kernel reads the three planes of its x columns
  const void* phys[MAXF];
};

// The NP = 8 + 3 NTH product lines of one column x at plane j, from the physical
// fields of planes j-1, j, j+1 (index 0, 1, 2) of the line pair (the .x and .y
// components), with rk_chan_1's y stencils (for/solver.f90:655-815, 914-964):
//   0 u3u1, 1 u1u1, 2 u3u3 (GYF, rows jstart..jend), 3 u1u2i, 4 u3u2i (GY, rows
//   2..Nyp), 5 flux of u1, 6 flux of u3 (GYF), 7 the GY derivative of u2u2, and
//   per scalar n: 8+3n th u1, 9+3n th u3, 10+3n flux of th (rows jstart_th..jend_th).
template <typename T, int NTH>
__device__ __forceinline__ void xprod_products(const XProdArgs& a, int j, const cx<T>* U1, const cx<T>* U2,
                                               const cx<T>* U3, const cx<T> (*TH)[3], cx<T>* P) {
  const bool vrow = (j >= a.jstart && j <= a.jend);
  const bool v2row = (j >= 2 && j <= a.g.ny);
  const double* rc = a.rowc + (long long)j * RC_COUNT;
  const T rdy = (T)rc[RC_RDY], rdyf = (T)rc[RC_RDYF];
  const T dyfj = (T)a.dyf[j], dyfm = (T)a.dyf[j - 1];
  const T cint = T(0.5) * rdy;  // interpolation to GY(j): (dyf_j u_j + dyf_{j-1} u_{j-1})/(2 dy_j)
  // the two real lines of the pair are the .x and .y components
#define DG_PROD(expr_x, expr_y) cx<T>{expr_x, expr_y}
#define C(v, c) (v).c
  if (vrow) {
    P[0] = DG_PROD(C(U3[1], x) * C(U1[1], x), C(U3[1], y) * C(U1[1], y));
    P[1] = DG_PROD(C(U1[1], x) * C(U1[1], x), C(U1[1], y) * C(U1[1], y));
    P[2] = DG_PROD(C(U3[1], x) * C(U3[1], x), C(U3[1], y) * C(U3[1], y));
    const T r2f = T(0.5) * rdyf;
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
    P[7] = DG_PROD((T(0.25) * (apx * apx) - T(0.25) * (amx * amx)) * rdy,
                   (T(0.25) * (apy * apy) - T(0.25) * (amy * amy)) * rdy);
  } else {
    P[3] = P[4] = P[7] = cx<T>{T(0), T(0)};
  }
#pragma unroll
  for (int n = 0; n < NTH; ++n) {
    const bool trow = (j >= a.jstart_th[n] && j <= a.jend_th[n]);
    if (trow) {
      const T r2f = T(0.5) * rdyf;
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
}

// The product lines of the pair, forward-transformed in LDS (line p at s + p NX),
// to the 6 + 2 NTH x-spectra of plane j: the two real lines split apart, the
// derivative lines times -ikx less the flux lines, 1/Nx.  hout: the output
// pointers, kxt: kx (either may be a copy in LDS).
template <typename T, int NX, int NT, int NTH, typename KT>
__device__ __forceinline__ void xprod_store(const cx<T>* __restrict__ s, const Geo& g, void* const* hout,
                                            const KT* __restrict__ kxt, int z, int j, int tid) {
  constexpr int NO = 6 + 2 * NTH;
  const T rnx = T(1) / T(NX);
  for (int idx = tid; idx < NO * g.kxr; idx += NT) {
    const int o = idx / g.kxr, i = idx % g.kxr;
    const T kxi = (T)kxt[i];
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
    // device memory (the pointer may come from LDS: say so, for global stores)
    typedef T v2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) v2 gv2;
    gv2* H = (gv2*)((cx<T>*)hout[o] + (long long)j * g.ph);
    H[(long long)z * g.kxp + i] = v2{rnx * ra.x, rnx * ra.y};
    H[(long long)(z + 1) * g.kxp + i] = v2{rnx * rb.x, rnx * rb.y};
  }
}

// A separate inverse x transform has written the physical fields once per
// plane; each block reads the three planes j-1..j+1 of its line pair (z, z+1),
// forms the products and forward-transforms them in LDS.  (The LES x pass: the
// physical fields exist anyway for the closure, and the SGS fields are added to
// the products.)
template <typename T, int NX, int NT, int NTH, bool LES>
__global__ __launch_bounds__(NT) void k_xprod(XProdArgs a) {
  constexpr int NP = 8 + 3 * NTH;
  __shared__ cx<T> s[NP * NX];
  __shared__ cx<T> stw[NX];  // twiddles
  stage_twiddles<T, NX, NT>(stw, (const cx<T>*)a.tw, threadIdx.x);
  const Geo& g = a.g;
  const int tid = threadIdx.x, z = 2 * blockIdx.x, j = a.jlo + blockIdx.y;
  __syncthreads();  // the twiddle table
  // (field f, plane j-1+d) of the line pair (z, z+1) at column x
  auto field = [&](int f, int d, int x) -> cx<T> {
    const T* R = (const T*)a.phys[f] + (long long)(j - 1 + d) * g.pp + (long long)z * NX + x;
    return cx<T>{R[0], R[NX]};
  };
  for (int x = tid; x < NX; x += NT) {
    cx<T> U1[3], U2[3], U3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      U1[d] = field(0, d, x);
      U2[d] = field(1, d, x);
      U3[d] = field(2, d, x);
    }
    cx<T> TH[NTH > 0 ? NTH : 1][3];
#pragma unroll
    for (int n = 0; n < NTH; ++n)
#pragma unroll
      for (int d = 0; d < 3; ++d) TH[n][d] = field(3 + n, d, x);
    cx<T> P[NP];
    xprod_products<T, NTH>(a, j, U1, U2, U3, TH, P);
    if constexpr (LES) {
      const long long pz = (long long)j * g.pp + (long long)z * NX + x;
      auto A = [&](int m) {
        const T* p = (const T*)a.aug[m];
        return cx<T>{p[pz], p[pz + NX]};
      };
      P[1] = P[1] + A(0);
      P[0] = P[0] + A(1);
      P[2] = P[2] + A(2);
      P[5] = P[5] + A(3);
      P[6] = P[6] + A(4);
      P[3] = P[3] + A(5);
      P[4] = P[4] + A(6);
#pragma unroll
      for (int n = 0; n < NTH; ++n) {
        P[8 + 3 * n] = P[8 + 3 * n] + A(7 + 2 * n);
        P[9 + 3 * n] = P[9 + 3 * n] + A(8 + 2 * n);
      }
    }
    // each thread writes only column x of the product lines
#pragma unroll
    for (int p = 0; p < NP; ++p) s[p * NX + swz<T>(x)] = P[p];
  }
  __syncthreads();
  fft_lds_sync<T, NX, NP, NT, -1>(s, NX, stw, tid);
  xprod_store<T, NX, NT, NTH>(s, g, a.hout, a.kx, z, j, tid);
}

// The DNS x pass in one kernel, marching in y.  Block (z pair, segment) owns the
// line pair (z, z+1) over the planes j0..j1 of its segment and keeps the physical
// fields of planes j-1, j, j+1 at its columns in registers: each plane's
// x-spectra are read once (the segment's two halo planes aside), inverse-
// transformed, multiplied out and forward-transformed without the physical
// fields ever reaching HBM.  One LDS FFT per plane serves both directions: the
// NP product lines of plane j and, conjugated (the inverse transform is
// conj(F(conj(X))), the same arithmetic as the +i transform), the NF field lines
// of plane j+2; the raw spectra of plane j+3 are loaded into registers meanwhile
// (unconditional loads from per-thread addresses fixed for the whole block, so
// that nothing waits on them before they are used one plane later; the output
// pointers and kx sit in LDS, so the stores wait on no global load either).
template <typename T, int NX, int NT, int NTH>
__global__ __launch_bounds__(NT) void k_xmarch(XProdArgs a, int seg) {
  constexpr int NF = 3 + NTH, NP = 8 + 3 * NTH, NL = NP + NF, NO = 6 + 2 * NTH;
  static_assert(3 * NF <= NL, "the three first planes share one FFT batch");
  static_assert(NX <= NT, "one column per thread");
  constexpr int HB = NX / 2 + 1;                  // bins 0..NX/2 of a line pair
  constexpr int PREF = (NF * HB + NT - 1) / NT;   // loaded bins per thread
  constexpr bool ALLX = NX == NT;
  __shared__ cx<T> s[NL * NX];
  __shared__ cx<T> stw[NX];  // twiddles
  __shared__ T skx[NX / 2];
  __shared__ void* sout[NO];
  stage_twiddles<T, NX, NT>(stw, (const cx<T>*)a.tw, threadIdx.x);
  const Geo& g = a.g;
  const int tid = threadIdx.x, z = 2 * blockIdx.x, x = tid;
  for (int i = tid; i < g.kxr; i += NT) skx[i] = (T)a.kx[i];
  if (tid < NO) sout[tid] = a.hout[tid];
  const int j0 = a.jlo + blockIdx.y * seg;
  const int j1 = min(j0 + seg - 1, g.ny);
  // bin m of field f per prefetch slot; the slot reads bins min(m, Nkx) of the
  // two lines (in bounds for every slot; unused values are dropped at put time)
  const cx<T>* pb[PREF];
  int pm[PREF];
#pragma unroll
  for (int p = 0; p < PREF; ++p) {
    const int idx = tid + p * NT;
    const bool live = idx < NF * HB;
    const int f = live ? idx / HB : 0;
    const int m = live ? idx % HB : -1;
    pm[p] = m;
    pb[p] = (const cx<T>*)a.hin[f] + (long long)z * g.kxp + (m > 0 ? min(m, g.nkx) : 0);
  }
  cx<T> ra[PREF], rb[PREF];
  auto load_plane = [&](int jp) {
    const long long off = (long long)jp * g.ph;
#pragma unroll
    for (int p = 0; p < PREF; ++p) {
      ra[p] = pb[p][off];
      rb[p] = pb[p][off + g.kxp];
    }
  };
  // the pair spectrum (bins m and NX - m, as k_inv_x: i > Nkx zero, the DC bins'
  // imaginary parts dropped), conjugated, into lines line0 + f
  auto put_plane = [&](int line0) {
#pragma unroll
    for (int p = 0; p < PREF; ++p) {
      const int m = pm[p];
      if (m >= 0) {
        const int f = (tid + p * NT) / HB;
        const cx<T> u = ra[p], v = rb[p];
        cx<T> lo = {T(0), T(0)}, hi = {T(0), T(0)};
        if (m == 0) {
          lo = {u.x, v.x};
        } else if (m <= g.nkx) {
          lo = {u.x - v.y, u.y + v.x};
          hi = {u.x + v.y, v.x - u.y};
        }
        s[(line0 + f) * NX + swz<T>(m)] = conjg(lo);
        if (m > 0 && m < NX / 2) s[(line0 + f) * NX + swz<T>(NX - m)] = conjg(hi);
      }
    }
  };
  cx<T> W[3][NF];  // fields at column x, planes j-1, j, j+1
  auto get_plane = [&](int line0, cx<T>* w) {
#pragma unroll
    for (int f = 0; f < NF; ++f) w[f] = conjg(s[(line0 + f) * NX + swz<T>(x)]);
  };
  // the window's first three planes
  load_plane(j0 - 1);
  put_plane(0);
  load_plane(j0);
  put_plane(NF);
  load_plane(j0 + 1);
  put_plane(2 * NF);
  if (j0 + 2 <= j1 + 1) load_plane(j0 + 2);
  __syncthreads();
  fft_lds<T, NX, 3 * NF, NT, -1>(s, NX, stw, tid);
  if (ALLX || x < NX) {
    get_plane(0, W[0]);
    get_plane(NF, W[1]);
    get_plane(2 * NF, W[2]);
  }
  __syncthreads();
  for (int j = j0; j <= j1; ++j) {
    const bool next = j + 2 <= j1 + 1;  // plane j+2 enters the window
    if (ALLX || x < NX) {
      cx<T> U1[3] = {W[0][0], W[1][0], W[2][0]};
      cx<T> U2[3] = {W[0][1], W[1][1], W[2][1]};
      cx<T> U3[3] = {W[0][2], W[1][2], W[2][2]};
      cx<T> TH[NTH > 0 ? NTH : 1][3];
#pragma unroll
      for (int n = 0; n < NTH; ++n) {
        TH[n][0] = W[0][3 + n];
        TH[n][1] = W[1][3 + n];
        TH[n][2] = W[2][3 + n];
      }
      cx<T> P[NP];
      xprod_products<T, NTH>(a, j, U1, U2, U3, TH, P);
#pragma unroll
      for (int p = 0; p < NP; ++p) s[p * NX + swz<T>(x)] = P[p];
    }
    if (next) {
      put_plane(NP);
      if (j + 3 <= j1 + 1) load_plane(j + 3);
    }
    __syncthreads();
    fft_lds_sync<T, NX, NL, NT, -1>(s, NX, stw, tid);
    xprod_store<T, NX, NT, NTH>(s, g, sout, skx, z, j, tid);
    if (next && (ALLX || x < NX)) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        W[0][f] = W[1][f];
        W[1][f] = W[2][f];
      }
      get_plane(NP, W[2]);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// z forward transform of the fused product spectra: the nonlinear tendency of
// every equation, NL_e = Fz(N_e) - ikz Fz(Z_e), scaled by 1/Nz, on the rows the
// equation owns (u1, u3: jstart..jend; u2: 2..Nyp; th_n: jstart_th..jend_th).
// ---------------------------------------------------------------------------
struct NlArgs {
  Geo g;
  int jlo;
  int jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  const double* __restrict__ kz;
  const void* h[MAXO];
  void* nl[MAXF];
  const void* tw;
};

__device__ __forceinline__ void eq_rows(int e, int jstart, int jend, const int* jst, const int* jet,
                                        int ny, int& lo, int& hi) {
  if (e == 0 || e == 2) { lo = jstart; hi = jend; }
  else if (e == 1) { lo = 2; hi = ny; }
  else { lo = jst[e - 3]; hi = jet[e - 3]; }
}

// Persistent blocks: tile t = (column block, plane, equation) runs t = b,
// b + gridDim.x, ... with b the block's xcd_tile position (gridDim.x a multiple
// of 8: XCD x works on a contiguous range of each round's tiles); the next
// tile's spectra are loaded into registers while the current tile is
// transformed and stored, so the loads overlap the FFT and the stores (kz sits
// in LDS: a global load of kz in the store loop made every store wait for the
// whole prefetch).  One tile per block otherwise, in xcd_tile order.
template <typename T, int NZ, int TI, int NT, bool PERSIST>
__global__ __launch_bounds__(NT) void k_fwdz_nl(NlArgs a, int ntiles) {
  constexpr int LD = NZ + 1;
  constexpr int NE = 2 * TI * NZ;              // elements per tile
  constexpr int PER = (NE + NT - 1) / NT;      // per thread
  __shared__ cx<T> s[2 * TI * LD];
  __shared__ cx<T> stw[NZ];  // twiddles
  __shared__ T skz[NZ];      // kz of the packed rows (kzr < NZ)
  stage_twiddles<T, NZ, NT>(stw, (const cx<T>*)a.tw, threadIdx.x);
  const Geo& g = a.g;
  const int tid = threadIdx.x;
  for (int k = tid; k < g.kzr; k += NT) skz[k] = (T)a.kz[k];
  const int nib = (g.kxr + TI - 1) / TI, nj = g.ny - a.jlo + 1;
  const T rnz = T(1) / T(NZ);
  // tile -> (equation e, plane j, first column i0); live: j is one of e's rows
  auto coords = [&](int t, int& e, int& j, int& i0) -> bool {
    const int ib = t % nib, r = t / nib;
    j = a.jlo + r % nj;
    e = r / nj;
    i0 = ib * TI;
    int lo, hi;
    eq_rows(e, a.jstart, a.jend, a.jstart_th, a.jend_th, g.ny, lo, hi);
    return j >= lo && j <= hi;
  };
  cx<T> buf[PER];
  auto load = [&](int t) {
    if (t >= ntiles) return;
    int e, j, i0;
    const bool live = coords(t, e, j, i0);
    const int oN = e < 3 ? e : 6 + 2 * (e - 3);
    const int oZ = e < 3 ? 3 + e : 7 + 2 * (e - 3);
    const cx<T>* HN = (const cx<T>*)a.h[oN] + (long long)j * g.ph;
    const cx<T>* HZ = (const cx<T>*)a.h[oZ] + (long long)j * g.ph;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int idx = tid + p * NT;
      const int ii = idx % TI, r = idx / TI, which = r / NZ, z = r - which * NZ, i = i0 + ii;
      const cx<T>* Hs = which ? HZ : HN;
      buf[p] = (live && idx < NE && i < g.kxr) ? Hs[(long long)z * g.kxp + i] : cx<T>{T(0), T(0)};
    }
  };
  const int step = PERSIST ? (int)gridDim.x : ntiles;  // one tile per block unless persistent
  int t = PERSIST ? ((gridDim.x & 7) ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x, gridDim.x / 8))
                  : xcd_tile(blockIdx.x, ntiles, XCD_K);
  load(t);
  __syncthreads();  // the twiddle table
  for (; t < ntiles; t += step) {
    int e, j, i0;
    const bool live = coords(t, e, j, i0);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int idx = tid + p * NT;
      if (idx < NE) {
        const int ii = idx % TI, r = idx / TI, which = r / NZ, z = r - which * NZ;
        s[(which * TI + ii) * LD + swz<T>(z)] = buf[p];
      }
    }
    __syncthreads();
    if (PERSIST) load(t + step);
    fft_lds_sync<T, NZ, 2 * TI, NT, -1>(s, LD, stw, tid);
    if (live) {
      cx<T>* NL = (cx<T>*)a.nl[e] + (long long)j * g.pc;
      for (int idx = tid; idx < TI * g.kzr; idx += NT) {
        const int ii = idx % TI, k = idx / TI, i = i0 + ii;
        if (i >= g.kxr) continue;
        const int zp = zpos(g, k);
        const cx<T> fn = s[ii * LD + swz<T>(zp)], fz = s[(TI + ii) * LD + swz<T>(zp)];
        NL[(long long)k * g.kxp + i] = rnz * (fn - ik(skz[k], fz));
      }
    }
    __syncthreads();  // s is rewritten by the next tile
  }
}

// ---------------------------------------------------------------------------
// filter_chan (for/solver.f90:1597-1685): the horizontal filter of one scalar,
// cth *= sigma(kx, kz) on rows [jlo, jhi]; sigma [kzr][kxp] from the host
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_filter_th(Geo g, cd* __restrict__ th, const double* __restrict__ sigma,
                                                     int jlo, int jhi) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const int j = jlo + blockIdx.y;
  if (c >= g.pc || j > jhi) return;
  if (c % g.kxp >= g.kxr) return;
  const long long o = (long long)j * g.pc + c;
  th[o] = sigma[c] * th[o];
}

// ---------------------------------------------------------------------------
// Column kernel.  Per (kx, kz) column and all rows j = 0..Nyp+1:
//  1. Fourier-space assembly of rk_chan_1 (for/solver.f90:385-838, 909-973, and
//     user_rhs_chan_fourier 35-111): new RK memory cf and the right-hand side cr,
//     including the explicit half of Crank–Nicolson (for/solver.f90:850-870,
//     981-1002) and the BC rows of apply_BC_* (for/solver.f90:1901-2232);
//  2. the Crank–Nicolson solves of every scalar, then u2, u1, u3
//     (for/solver.f90:1004-1233), in Fourier space (the DNS operator does not
//     depend on x or z), by parallel cyclic reduction with coefficients that the
//     host precomputes from the reference's matrices;
//  3. rem_div_chan (for/solver.f90:1302-1379, BC rows 2330-2371) by parallel cyclic
//     reduction, then the velocity correction and the pressure update
//     (for/solver.f90:1261-1267).
// ---------------------------------------------------------------------------

struct SolveRows {
  int blo, bhi;      // build range: rows whose cr comes from the assembly
  int wlo, whi;      // rows written back
  int kind[5];       // rows 0, 1, 2, ny, ny+1: -1 = build-range rule, 0 = zero, 1 = BC value
  double val[5];     // BC value of the mean mode at those rows
};

struct ColArgs {
  Geo g;
  int rk, do_cn, plevels;
  int cr_minus_u;   // do_cn == 0: write the right-hand side minus the state (increment solves)
  int jstart, jend, jstart_th[MAXTH], jend_th[MAXTH];
  int front, f_type;
  double nu, beta, nu_v_scale, ro_inv, delta, grav_x, grav_y, grav_z, ly, px0, omega0, amp_omega0, time;
  double pr[MAXTH], dthdx[MAXTH], dthdz[MAXTH];
  double temp1, temp3, temp4, temp5;
  // host-precomputed scalars: Ro_inv/delta; per scalar dthdx delta/Ro_inv,
  // -dthdz delta/Ro_inv, nu/Pr, temp1 nu_v_scale/Pr; temp1 nu_v_scale
  double cori, twu_x[MAXTH], twu_z[MAXTH], nupr[MAXTH], t1pr[MAXTH], t1v;
  const double* __restrict__ rowc;  // [ny2][RC_COUNT]
  const double *__restrict__ dy, *__restrict__ dyf, *__restrict__ gy, *__restrict__ gyf;
  const double *__restrict__ kx, *__restrict__ kz;
  SolveRows rows[MAXF];
  // Thomas factorisation of the CN systems, [field][ny2][5]: mu_j (forward
  // multiplier), beta_j = 1/b'_j, gamma_j = -c_j/b'_j, F_j / G_j (products of mu /
  // gamma from the chunk start / to the chunk end); rows per chunk
  const double* __restrict__ thc;
  int chunk;
  cd* u[MAXF];
  cd* p;
  cd* cf[MAXF];
  const void* nl[MAXF];
  cd* crout[MAXF];  // do_cn == 0: right-hand side rows [blo, bhi] of the physical-space solve
  // k_project in its poisson_p_chan form: the transformed source (else null), the
  // scalar count for its buoyancy term; the solution goes to p
  const cd* prhs;
  int pnth;
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// The grid is padded to a multiple of 8 and block b takes tile
// (b % 8) * (grid / 8) + b / 8, so that every XCD works on one contiguous range of
// tiles and neighbouring tiles, which share cache lines, meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = (int)gridDim.x / 8;
  const int t = (b % 8) * per + b / 8;
  return t < nb ? t : -1;
}

__device__ __forceinline__ int special_row(int j, int ny) {
  return j <= 2 ? j : (j == ny ? 3 : (j == ny + 1 ? 4 : -1));
}

// Inputs of the assembly at one (row j, column) point.  Loaded before any store
// so that the loads of a thread's rows are independent of its cf stores.
template <int NTH>
struct RowIn {
  cd u1m, u1c, u1p, u2m, u2c, u2p, u3m, u3c, u3p, pm, pc_;
  cd thm[NTH > 0 ? NTH : 1], thc[NTH > 0 ? NTH : 1], thp[NTH > 0 ? NTH : 1];
  cd cfo[3 + NTH], nl[3 + NTH];
};

template <typename T, int NTH>
__device__ __forceinline__ void load_row(RowIn<NTH>& r, const ColArgs& a, long long o, long long pc,
                                         const cd* __restrict__ u1, const cd* __restrict__ u2,
                                         const cd* __restrict__ u3, const cd* __restrict__ t0,
                                         const cd* __restrict__ t1, const cd* __restrict__ p) {
  const long long om = o - pc, op = o + pc;
  r.u1m = u1[om]; r.u1c = u1[o]; r.u1p = u1[op];
  r.u2m = u2[om]; r.u2c = u2[o]; r.u2p = u2[op];
  r.u3m = u3[om]; r.u3c = u3[o]; r.u3p = u3[op];
  r.pm = p[om]; r.pc_ = p[o];
  if constexpr (NTH >= 1) { r.thm[0] = t0[om]; r.thc[0] = t0[o]; r.thp[0] = t0[op]; }
  if constexpr (NTH >= 2) { r.thm[1] = t1[om]; r.thc[1] = t1[o]; r.thp[1] = t1[op]; }
#pragma unroll
  for (int e = 0; e < 3 + NTH; ++e) {
    r.cfo[e] = (a.rk > 1) ? ((const cd* __restrict__)a.cf[e])[o] : cd{0.0, 0.0};
    r.nl[e] = cvt<double>(((const cx<T>* __restrict__)a.nl[e])[o]);
  }
}

// cf and cr of equation e at row j, column (i, k), from the loaded row inputs;
// kk = (kx^2 + kz^2)^beta of the column, rc the row's coefficient table
template <int NTH>
__device__ __forceinline__ void assemble(const ColArgs& a, const RowIn<NTH>& r, const double* __restrict__ rc,
                                         int e, int i, int k, double kxi, double kzk, double kk,
                                         cd& cf, cd& cr) {
  const bool tws = a.front && a.ro_inv != 0.0;
  if (e == 0 || e == 2) {
    const cd uc = (e == 0) ? r.u1c : r.u3c;
    const cd up = (e == 0) ? r.u1p : r.u3p;
    const cd um = (e == 0) ? r.u1m : r.u3m;
    cf = -1.0 * ((a.nu * kk) * uc);
    if (e == 0) cf = cf + a.cori * r.u3c;
    else cf = cf - a.cori * r.u1c;
    for (int n = 0; n < NTH; ++n) cf = cf + (e == 0 ? a.grav_x : a.grav_z) * r.thc[n];
    if (tws) {
      const cd u2m = 0.5 * (r.u2c + r.u2p);
      for (int n = 0; n < NTH; ++n) {
        cf = cf - rc[RC_TW0 + 4 * n] * ik(kzk, uc) - rc[RC_TW0 + 4 * n + 1] * ik(kxi, uc);
        if (e == 0) cf = cf - a.twu_z[n] * u2m;
        else cf = cf - a.twu_x[n] * u2m;
      }
    }
    cf = cf + r.nl[e];
    cr = uc;
    if (a.rk > 1) cr = cr + a.temp3 * r.cfo[e];
    cr = cr - a.temp4 * (e == 0 ? ik(kxi, r.pc_) : ik(kzk, r.pc_));
    if (e == 0 && k == 0 && i == 0) {
      if (a.f_type == 1) cr.x = cr.x - a.temp4 * a.px0;
      else if (a.f_type == 2) cr.x = cr.x - a.temp4 * (a.px0 + a.amp_omega0 * cos(a.omega0 * a.time));
    }
    cr = cr + a.temp5 * cf;
    cr = cr + a.t1v * (rc[RC_RDYF] * (rc[RC_RDYP] * (up - uc) - rc[RC_RDY] * (uc - um)));
  } else if (e == 1) {
    const cd uc = r.u2c;
    cf = -1.0 * ((a.nu * kk) * uc);
    if (!(k == 0 && i == 0)) {
      for (int n = 0; n < NTH; ++n) cf = cf + a.grav_y * (rc[RC_WTH] * r.thc[n] + rc[RC_WTHM] * r.thm[n]);
    }
    if (tws) {
      for (int n = 0; n < NTH; ++n)
        cf = cf - rc[RC_TW0 + 4 * n + 2] * ik(kzk, uc) - rc[RC_TW0 + 4 * n + 3] * ik(kxi, uc);
    }
    cf = cf + r.nl[1];
    cr = uc;
    if (a.rk > 1) cr = cr + a.temp3 * r.cfo[1];
    cr = cr - a.temp4 * (rc[RC_RDY] * (r.pc_ - r.pm));
    cr = cr + a.temp5 * cf;
    cr = cr + a.t1v * (rc[RC_RDY] * (rc[RC_RDYF] * (r.u2p - uc) - rc[RC_RDYFM] * (uc - r.u2m)));
  } else {
    const int n = e - 3;
    const cd tc = r.thc[n];
    cf = -1.0 * ((a.nupr[n] * kk) * tc);
    cf = cf - a.dthdx[n] * r.u1c - a.dthdz[n] * r.u3c;
    if (tws) cf = cf - rc[RC_TW0 + 4 * n] * ik(kzk, tc) - rc[RC_TW0 + 4 * n + 1] * ik(kxi, tc);
    cf = cf + r.nl[e];
    cr = tc;
    if (a.rk > 1) cr = cr + a.temp3 * r.cfo[e];
    cr = cr + a.temp5 * cf;
    cr = cr + a.t1pr[n] * (rc[RC_RDYF] * (rc[RC_RDYP] * (r.thp[n] - tc) - rc[RC_RDY] * (tc - r.thm[n])));
  }
}

// The assembly of k_cn alone (the LES branch, whose CN solve runs per physical
// column): cf and the right-hand side (or the right-hand side minus the state,
// cr_minus_u) of every equation at one point per thread.  The columns of a row
// are consecutive threads, so every load and store of a wave is one contiguous
// run of the plane, and no LDS is held (the column kernel reserves its solve's).
// Same rows, same load_row / assemble, same writes as k_cn with do_cn = 0.
template <typename T, int NTH>
__global__ __launch_bounds__(256) void k_assemble(ColArgs a) {
  constexpr int NF = 3 + NTH;
  const Geo& g = a.g;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (c >= g.pc) return;
  const int i = (int)(c % g.kxp), k = (int)(c / g.kxp);
  if (i >= g.kxr) return;  // padding columns: k_cn writes nothing there either
  const long long pc = g.pc;
  const double kxi = a.kx[i], kzk = a.kz[k];
  const double k2 = kxi * kxi + kzk * kzk;
  const double kk = (a.beta == 1.0) ? k2 : pow(k2, a.beta);
  const long long o = (long long)j * pc + c;
  const bool inner = j >= 1 && j <= g.ny;
  RowIn<NTH> r;
  if (inner)
    load_row<T, NTH>(r, a, o, pc, a.u[0], a.u[1], a.u[2], NTH >= 1 ? a.u[3] : a.u[0], NTH >= 2 ? a.u[4] : a.u[0],
                     a.p);
#pragma unroll
  for (int e = 0; e < NF; ++e) {
    const SolveRows& SR = a.rows[e];
    cd v = {0.0, 0.0};
    if (inner && j >= SR.blo && j <= SR.bhi) {
      cd cf;
      assemble<NTH>(a, r, a.rowc + (long long)j * RC_COUNT, e, i, k, kxi, kzk, kk, cf, v);
      a.cf[e][o] = cf;
      if (a.cr_minus_u) v = v - a.u[e][o];
    }
    a.crout[e][o] = v;
  }
}

// Assembly + Crank-Nicolson solves.  T: precision of the nonlinear terms;
// TC columns x RG row groups per block; RPT >= ceil(ny2 / RG) rows per thread.
// The CN solution overwrites the state rows [wlo, whi]; k_project follows.
template <typename T, int TC, int NT, int RPT, int NTH, int LMAX>
__global__ __launch_bounds__(NT) void k_cn(ColArgs a) {
  constexpr int NF = 3 + NTH;
  constexpr int RG = NT / TC;
  extern __shared__ unsigned char smem[];
  const Geo& g = a.g;
  const int ny2 = g.ny2, ny = g.ny, tid = threadIdx.x;
  const int cc = tid % TC, rg = tid / TC;
  const int tile = xcd_tile(blockIdx.x, (int)(g.pc / TC));
  if (tile < 0) return;
  const int c = tile * TC + cc;  // column in the padded plane: c = k*kxp + i
  const int i = c % g.kxp, k = c / g.kxp;
  const bool active = i < g.kxr;
  const long long pc = g.pc;
  cd* sR = (cd*)smem;
  // rows of the TC columns interleaved: lanes of one row group touch consecutive 16-byte slots
  auto R = [&](int f, int j) -> cd& { return sR[(f * ny2 + j) * TC + cc]; };
  const cd* __restrict__ u1 = a.u[0];
  const cd* __restrict__ u2 = a.u[1];
  const cd* __restrict__ u3 = a.u[2];
  const cd* __restrict__ t0 = NTH >= 1 ? a.u[3] : a.u[0];
  const cd* __restrict__ t1 = NTH >= 2 ? a.u[4] : a.u[0];
  const cd* __restrict__ pp = a.p;

  // ---- 1. assembly of cf and cr; BC rows of the right-hand side
  const double kxi = active ? a.kx[i] : 0.0, kzk = a.kz[k];
  const double k2 = kxi * kxi + kzk * kzk;
  const double kk = (a.beta == 1.0) ? k2 : pow(k2, a.beta);
  for (int m = 0; m < RPT; ++m) {
    const int j = rg + m * RG;
    if (j >= ny2) break;
    const long long o = (long long)j * pc + c;
    const bool inner = active && j >= 1 && j <= ny;
    RowIn<NTH> r;
    if (inner) load_row<T, NTH>(r, a, o, pc, u1, u2, u3, t0, t1, pp);
    cd crv[NF];
#pragma unroll
    for (int e = 0; e < NF; ++e) {
      const SolveRows& SR = a.rows[e];
      cd v = {0.0, 0.0};
      if (inner && j >= SR.blo && j <= SR.bhi) {
        cd cf;
        assemble<NTH>(a, r, a.rowc + (long long)j * RC_COUNT, e, i, k, kxi, kzk, kk, cf, v);
        a.cf[e][o] = cf;
      }
      if (a.do_cn) {
        const int sr = special_row(j, ny);
        const int kd = sr >= 0 ? SR.kind[sr] : -1;
        if (kd == 0) v = {0.0, 0.0};
        else if (kd == 1) v = (c == 0) ? cd{SR.val[sr], 0.0} : cd{0.0, 0.0};
        else if (!(j >= SR.blo && j <= SR.bhi)) v = {0.0, 0.0};
      } else if (!(inner && j >= SR.blo && j <= SR.bhi)) {
        v = {0.0, 0.0};  // physical-space solve: its BC rows are set per column there
      }
      crv[e] = v;
    }
    if (!a.do_cn) {
      if (active) {
#pragma unroll
        for (int e = 0; e < NF; ++e) {
          cd v = crv[e];
          if (a.cr_minus_u && inner && j >= a.rows[e].blo && j <= a.rows[e].bhi) v = v - a.u[e][o];
          a.crout[e][o] = v;
        }
      }
      continue;
    }
#pragma unroll
    for (int e = 0; e < NF; ++e) R(e, j) = crv[e];
  }
  if (!a.do_cn) return;  // uniform over the grid
  __syncthreads();

  // ---- 2. Thomas with the host's factorisation, in chunks.  One wave takes one
  // field of the TC columns: lane = p*TC + cc, p = chunk of L consecutive rows.
  // Per chunk: local elimination from zero, a shuffle scan of the chunk-boundary
  // carries (the affine maps y -> A y + B compose), and the fix-up with the
  // host's in-chunk products.  Same linear map as thomas_forward/backward_real_mpi
  // (for/solver.f90:1691-1761, 1816-1852); scalars first (for/solver.f90:1004-1007).
  constexpr int P = 64 / TC;
  const int lane = tid & 63, wave = tid / 64;
  const int p = lane / TC, ccl = lane % TC;
  const int L = a.chunk, j0 = p * L;
  const int cl = tile * TC + ccl;
  const bool act_l = (cl % g.kxp) < g.kxr;
  for (int e0 = wave; e0 < NF; e0 += NT / 64) {
    const int e = (e0 < NTH) ? 3 + e0 : e0 - NTH;
    const double* __restrict__ co = a.thc + (long long)e * ny2 * 5;  // mu, beta, gamma, F, G per row
    auto Rl = [&](int j) -> cd& { return sR[(e * ny2 + j) * TC + ccl]; };
    cd y[LMAX];
    // forward: y_j = g_j + mu_j y_{j-1}, from zero at the chunk start
    cd prev = {0.0, 0.0};
    double Af = 1.0;
    cd Bf = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < LMAX; ++m) {
      const int j = j0 + m;
      if (m < L && j < ny2) {
        y[m] = Rl(j) + co[5 * j] * prev;
        prev = y[m];
        Af = co[5 * j + 3];
        Bf = y[m];
      }
    }
    // inclusive scan of the maps over the chunks below (lanes p - d)
#pragma unroll
    for (int d = 1; d < P; d <<= 1) {
      const double Au = __shfl_up(Af, d * TC);
      const cd Bu = {__shfl_up(Bf.x, d * TC), __shfl_up(Bf.y, d * TC)};
      if (p >= d) {
        Bf = Af * Bu + Bf;
        Af = Af * Au;
      }
    }
    cd yin = {__shfl_up(Bf.x, TC), __shfl_up(Bf.y, TC)};
    if (p == 0) yin = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < LMAX; ++m) {
      const int j = j0 + m;
      if (m < L && j < ny2) y[m] = y[m] + co[5 * j + 3] * yin;
    }
    // backward: x_j = beta_j y_j + gamma_j x_{j+1}, from zero past the chunk end
    cd nxt = {0.0, 0.0};
    double Ab = 1.0;
    cd Bb = {0.0, 0.0};
#pragma unroll
    for (int m = LMAX - 1; m >= 0; --m) {
      const int j = j0 + m;
      if (m < L && j < ny2) {
        y[m] = co[5 * j + 1] * y[m] + co[5 * j + 2] * nxt;
        nxt = y[m];
        Ab = co[5 * j + 4];
        Bb = y[m];
      }
    }
#pragma unroll
    for (int d = 1; d < P; d <<= 1) {
      const double Ad = __shfl_down(Ab, d * TC);
      const cd Bd = {__shfl_down(Bb.x, d * TC), __shfl_down(Bb.y, d * TC)};
      if (p + d < P) {
        Bb = Ab * Bd + Bb;
        Ab = Ab * Ad;
      }
    }
    cd xin = {__shfl_down(Bb.x, TC), __shfl_down(Bb.y, TC)};
    if (p == P - 1) xin = {0.0, 0.0};
    const SolveRows& SR = a.rows[e];
#pragma unroll
    for (int m = 0; m < LMAX; ++m) {
      const int j = j0 + m;
      if (m < L && j < ny2 && act_l && j >= SR.wlo && j <= SR.whi)
        a.u[e][(long long)j * pc + cl] = y[m] + co[5 * j + 4] * xin;
    }
  }
}

// rem_div_chan (for/solver.f90:1302-1379, BC rows 2330-2371): Poisson rows 1..Nyp
// by PCR with per-column coefficients, velocity correction, and with
// update_p the pressure update of rk_chan_1 (for/solver.f90:1261-1267).
//
// Every row is kept with off-diagonals a, c >= 0 and diagonal b <= 0 (the bottom
// BC row is negated), and stored by its row sum s = a + b + c instead of b: the
// interior rows sum to exactly -(kx^2 + kz^2) and the BC rows to 0 (or -1 for
// the Dirichlet row of the (0,0) mode).  A reduction step then forms
//   s' = s + al s(j-s) + ga s(j+s),   b = s - a - c,
// sums of terms of one sign (al, ga >= 0), where the usual update
// b' = b + al c(j-s) + ga a(j+s) cancels almost all of b for the weakly dominant
// rows of small kx^2 + kz^2 (on a 257-row stretched grid plain PCR lost 3-4
// digits there against the Thomas sweep; this form stays within 10x of it).
template <int TC, int NT, int RPT>
__global__ __launch_bounds__(NT) void k_project(ColArgs a, int update_p) {
  constexpr int RG = NT / TC;
  extern __shared__ unsigned char smem[];
  const Geo& g = a.g;
  const int ny2 = g.ny2, ny = g.ny, tid = threadIdx.x;
  const int cc = tid % TC, rg = tid / TC;
  const int tile = xcd_tile(blockIdx.x, (int)(g.pc / TC));
  if (tile < 0) return;
  const int c = tile * TC + cc;
  const int i = c % g.kxp, k = c / g.kxp;
  const bool active = i < g.kxr;
  const long long pc = g.pc;
  cd* sD = (cd*)smem;
  double* sA = (double*)(sD + TC * ny2);
  double* sS = sA + TC * ny2;  // row sums a + b + c
  double* sC = sS + TC * ny2;
  cd* __restrict__ u1 = a.u[0];
  cd* __restrict__ u2 = a.u[1];
  cd* __restrict__ u3 = a.u[2];
  const double kxi = active ? a.kx[i] : 0.0, kzk = a.kz[k];
  const double k2 = -(kxi * kxi) - kzk * kzk;
  // the velocity of the thread's rows, read once for the divergence and kept for
  // the correction, on the tall columns (at 512x257x512 the second read came
  // back from HBM: 0.763 -> 0.712 ms); at 3 rows per thread the 56 more VGPRs
  // cost more occupancy than the reads (256x129x256: 0.084 -> 0.090 ms)
  constexpr bool KEEP = RPT >= 4 && RPT <= 8;
  cd ku1[KEEP ? RPT : 1], ku2[KEEP ? RPT : 1], ku3[KEEP ? RPT : 1];
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int j = rg + m * RG;
    if (j >= ny2) break;
    const int q = j * TC + cc;
    double aj = 0.0, sj = 1.0, cj = 0.0;
    cd d = {0.0, 0.0};
    if (j == 1) {
      // (1, -1) Neumann, or (1) for phi(1) = 0 in the (0,0) mode; negated
      cj = (c == 0) ? 0.0 : 1.0;
      sj = (c == 0) ? -1.0 : 0.0;
    } else if (j == ny) {
      aj = 1.0; sj = 0.0;  // (1, -1)
    } else if (j > 1 && j < ny) {
      aj = 1.0 / (a.dy[j] * a.dyf[j]);
      sj = k2;
      cj = 1.0 / (a.dy[j + 1] * a.dyf[j]);
      if (active) {
        const long long o = (long long)j * pc + c;
        if (a.prhs) {
          // poisson_p_chan (for/solver.f90:1382-1560): vec = -(source + the scalars'
          // y gradient, except in the (0, 0) mode)
          cd sv = a.prhs[o];
          if (c != 0)
            for (int n = 0; n < a.pnth; ++n) {
              const cd dt = a.u[3 + n][o + pc] - a.u[3 + n][o - pc];
              const double den = a.gyf[j + 1] - a.gyf[j - 1];
              sv = sv + cd{dt.x / den, dt.y / den};
            }
          d = -1.0 * sv;
        } else {
          const cd v1 = u1[o], v2 = u2[o], v3 = u3[o];
          d = ik(kxi, v1) + (1.0 / a.dyf[j]) * (u2[o + pc] - v2) + ik(kzk, v3);
          if constexpr (KEEP) { ku1[m] = v1; ku2[m] = v2; ku3[m] = v3; }
        }
      }
    }
    sA[q] = aj; sS[q] = sj; sC[q] = cj; sD[q] = d;
  }
  __syncthreads();
  for (int l = 0; l < a.plevels; ++l) {
    const int s = 1 << l;
    double na[RPT], ns[RPT], nc[RPT];
    cd nd[RPT];
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
      const int j = rg + m * RG;
      if (j >= 1 && j <= ny) {
        const int q = j * TC + cc, qm = q - s * TC, qp = q + s * TC;
        const int jm = j - s, jp = j + s;
        double a2 = 0.0, c2 = 0.0, s2 = sS[q];
        cd d2 = sD[q];
        if (jm >= 1) {
          const double bm = sS[qm] - sA[qm] - sC[qm];
          const double al = -sA[q] / bm;
          a2 = al * sA[qm];
          s2 = s2 + al * sS[qm];
          d2 = d2 + al * sD[qm];
        }
        if (jp <= ny) {
          const double bp = sS[qp] - sA[qp] - sC[qp];
          const double ga = -sC[q] / bp;
          c2 = ga * sC[qp];
          s2 = s2 + ga * sS[qp];
          d2 = d2 + ga * sD[qp];
        }
        na[m] = a2; ns[m] = s2; nc[m] = c2; nd[m] = d2;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
      const int j = rg + m * RG;
      if (j >= 1 && j <= ny) {
        const int q = j * TC + cc;
        sA[q] = na[m]; sS[q] = ns[m]; sC[q] = nc[m]; sD[q] = nd[m];
      }
    }
    __syncthreads();
  }
  for (int m = 0; m < RPT; ++m) {
    const int j = rg + m * RG;
    if (j >= ny2) break;
    const int q = j * TC + cc;
    sD[q] = (j >= 1 && j <= ny) ? (1.0 / (sS[q] - sA[q] - sC[q])) * sD[q] : cd{0.0, 0.0};
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int j = rg + m * RG;
    if (j >= ny2) break;
    if (j < 1 || j > ny) continue;
    const long long o = (long long)j * pc + c;
    const cd phi = sD[j * TC + cc];
    if (a.prhs) {  // cp(j) = vec(j), j = 1..Nyp
      a.p[o] = phi;
      continue;
    }
    // rows 1 and Nyp were not read above: their divergence row is a BC row
    const bool kept = KEEP && j > 1 && j < ny;
    const cd v1 = kept ? ku1[KEEP ? m : 0] : u1[o];
    const cd v2 = kept ? ku2[KEEP ? m : 0] : u2[o];
    const cd v3 = kept ? ku3[KEEP ? m : 0] : u3[o];
    u1[o] = v1 - ik(kxi, phi);
    u3[o] = v3 - ik(kzk, phi);
    if (j >= 2) u2[o] = v2 - (1.0 / a.dy[j]) * (phi - sD[(j - 1) * TC + cc]);
    if (update_p && j >= a.jstart && j <= a.jend) a.p[o] = a.p[o] + (1.0 / a.temp4) * phi;
  }
}

}  // namespace dgpu
