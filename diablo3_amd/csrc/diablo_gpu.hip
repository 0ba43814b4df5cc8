// diablo_gpu.hip — C ABI of the MI355X RKW3 + Crank–Nicolson channel step.
//
// One substep (rk_chan_1, for/solver.f90:366-1273) of DNS is four launches:
//   k_inv_z     packed Fourier state -> x-spectra per z row (z inverse transform)
//   k_xprod     x inverse transform, nonlinear products with their y stencils,
//               x forward transform, ikx factors (physical space stays in LDS)
//   k_fwdz_asm  z forward transform of the products + Fourier-space assembly of
//               cf (RK memory) and cr (right-hand side incl. explicit CN half)
//   k_columns   per (kx,kz) column: CN solves, rem_div_chan, pressure update
// The CN operator of DNS is independent of x and z, so solving it per Fourier
// column is the same linear map as the reference's per-physical-column solve
// followed by the forward transform; it saves the 8 transforms the reference
// spends on cr -> r and u -> cu.
#include "../../include/diablo_gpu.h"
#include "dgpu_kernels.hpp"

#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace dgpu;

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(x)                                                                 \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) return fail(10, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

template <typename T>
std::vector<cx<T>> twiddles(int n) {
  std::vector<cx<T>> t(n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int m = 0; m < n; ++m) {
    long double a = twopi * (long double)m / (long double)n;
    t[m] = cx<T>{(T)(double)cosl(a), (T)(double)(-sinl(a))};
  }
  return t;
}

constexpr int ti_for(int nz, int sz) {
  int t = 32768 / (nz * sz);
  return t < 2 ? 2 : (t > 16 ? 16 : t);
}

}  // namespace

struct dgpu_handle {
  dgpu_config c{};
  std::vector<double> gy, gyf, dy, dyf, kx, kz;
  Geo g{};
  int nth = 0, prec = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double *d_gy = nullptr, *d_gyf = nullptr, *d_dy = nullptr, *d_dyf = nullptr;
  double *d_kx = nullptr, *d_kz = nullptr;
  cd* U[MAXF] = {};
  cd* P = nullptr;
  cd* CF[MAXF] = {};
  cd* CR[MAXF] = {};
  void* Hin[MAXF] = {};
  void* Hout[MAXO] = {};
  void *twx64 = nullptr, *twz64 = nullptr, *twx32 = nullptr, *twz32 = nullptr;
  double* d_fac = nullptr;
  double fac_dt = NAN;
  double* d_R = nullptr;   // one physical FP64 field (transform entry points)
  cd* d_H = nullptr;       // one FP64 x-spectrum field
  cd* d_C = nullptr;       // one packed field
  double* hstage = nullptr;  // pinned staging, one Fortran array
  size_t fort_doubles = 0;
  std::vector<void*> allocs;
  // per-kernel timing (measurement only)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  struct Mark { int id; hipEvent_t a, b; };
  std::vector<Mark> marks;
  double t_ms[4] = {0, 0, 0, 0};
  long long t_n[4] = {0, 0, 0, 0};
};

namespace {

int dalloc(dgpu_handle* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(11, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->allocs.push_back(*p);
  e = hipMemsetAsync(*p, 0, bytes ? bytes : 16, h->st);
  if (e != hipSuccess) return fail(11, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return 0;
}

// --------------------------------------------------------------------------------
// Crank–Nicolson systems of one column (the DNS matrices do not depend on the
// column): rows 0..ny+1 as built in rk_chan_1 and overwritten by apply_BC_*.
// --------------------------------------------------------------------------------
void cn_matrix(const dgpu_handle* h, int F, double temp1, std::vector<double>& a,
               std::vector<double>& b, std::vector<double>& cc) {
  const dgpu_config& c = h->c;
  const int ny = c.ny, n2 = ny + 2;
  const double* dy = h->dy.data();
  const double* dyf = h->dyf.data();
  a.assign(n2, 0.0);
  b.assign(n2, 1.0);
  cc.assign(n2, 0.0);
  const double nvs = c.nu_v_scale;
  auto set = [&](int j, double l, double d, double u) { a[j] = l; b[j] = d; cc[j] = u; };
  auto gyf_bc = [&](int lo, int hi) {
    set(0, 0, 1, 0);
    if (lo == DGPU_BC_DIRICHLET) set(1, 0, 1, 0);
    else set(1, 0, -1, 1);
    if (hi == DGPU_BC_DIRICHLET) { set(ny + 1, 0, 1, 0); set(ny, 0, 1, 0); }
    else { set(ny, -1, 1, 0); set(ny + 1, 0, 1, 0); }
  };
  if (F >= 3) {
    const int n = F - 3;
    const double t = temp1 / c.pr[n] * nvs;
    for (int j = c.jstart_th[n]; j <= c.jend_th[n]; ++j)
      set(j, -t / (dy[j] * dyf[j]), 1. + t / (dy[j + 1] * dyf[j]) + t / (dy[j] * dyf[j]),
          -t / (dy[j + 1] * dyf[j]));
    gyf_bc(c.th_bc_ymin[n], c.th_bc_ymax[n]);
  } else if (F == 1) {
    for (int j = 2; j <= ny; ++j)
      set(j, -temp1 * nvs / (dyf[j - 1] * dy[j]),
          1. + temp1 * nvs / (dyf[j] * dy[j]) + temp1 * nvs / (dyf[j - 1] * dy[j]),
          -temp1 * nvs / (dyf[j] * dy[j]));
    if (c.v_bc_ymin == DGPU_BC_DIRICHLET) { set(1, 0, 1, 0); set(2, 0, 1, 0); }
    else if (c.v_bc_ymin == DGPU_BC_NEUMANN) set(1, 0, -1, 1);
    set(0, 0, 1, 0);
    if (c.v_bc_ymax == DGPU_BC_DIRICHLET) { set(ny + 1, 0, 1, 0); set(ny, 0, 1, 0); }
    else if (c.v_bc_ymax == DGPU_BC_NEUMANN) set(ny + 1, -1, 1, 0);
  } else {
    for (int j = c.jstart; j <= c.jend; ++j)
      set(j, -temp1 * nvs / (dy[j] * dyf[j]),
          1. - temp1 * nvs * (-1. / (dy[j + 1] * dyf[j]) - 1. / (dy[j] * dyf[j])),
          -temp1 * nvs / (dy[j + 1] * dyf[j]));
    if (F == 0) gyf_bc(c.u_bc_ymin, c.u_bc_ymax);
    else gyf_bc(c.w_bc_ymin, c.w_bc_ymax);
  }
}

int ensure_factors(dgpu_handle* h, double delta_t) {
  if (h->fac_dt == delta_t) return 0;
  const int n2 = h->c.ny + 2, nf = 3 + h->nth;
  std::vector<double> buf((size_t)3 * MAXF * 3 * n2, 0.0), a, b, cc;
  const double hb[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  for (int r = 0; r < 3; ++r) {
    const double temp1 = 0.5 * h->c.nu * hb[r];
    for (int F = 0; F < nf; ++F) {
      cn_matrix(h, F, temp1, a, b, cc);
      double* L = &buf[((size_t)r * MAXF + F) * 3 * n2];
      double* Uc = L + n2;
      double* RB = Uc + n2;
      // thomas_forward_real_mpi at one rank: rows 1..ny+1 eliminated against row j-1
      std::vector<double> bp(n2);
      bp[0] = b[0];
      L[0] = 0.0;
      for (int j = 1; j < n2; ++j) {
        L[j] = -a[j] / bp[j - 1];
        bp[j] = b[j] + L[j] * cc[j - 1];
      }
      for (int j = 0; j < n2; ++j) {
        if (bp[j] == 0.0) return fail(20, "singular Crank-Nicolson system");
        Uc[j] = cc[j];
        RB[j] = 1.0 / bp[j];
      }
    }
  }
  HIPCHK(hipMemcpyAsync(h->d_fac, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->fac_dt = delta_t;
  return 0;
}

// BC rows of the right-hand side, as apply_BC_* write vec (for/solver.f90:1901-2232)
void solve_rows(const dgpu_handle* h, double time, SolveRows* rows) {
  const dgpu_config& c = h->c;
  const int ny = c.ny;
  for (int F = 0; F < MAXF; ++F) {
    SolveRows& R = rows[F];
    for (int s = 0; s < 5; ++s) { R.kind[s] = -1; R.val[s] = 0.0; }
    R.blo = 1; R.bhi = 0; R.wlo = 1; R.whi = 0;
  }
  auto gyf_rows = [&](SolveRows& R, int lo, double c1lo, int hi, double c1hi) {
    R.kind[0] = 0;
    R.kind[1] = 1;
    R.val[1] = (lo == DGPU_BC_DIRICHLET) ? c1lo : h->dy[2] * c1lo;
    R.kind[4] = 0;
    R.kind[3] = 1;
    R.val[3] = (hi == DGPU_BC_DIRICHLET) ? c1hi : h->dy[ny] * c1hi;
  };
  double wtop = c.w_bc_ymax_c1;
  if (c.f_type == 4)
    wtop = c.w_bc_ymax_c1 + c.amp_omega0 * std::sin(c.omega0 * (c.ro_inv / c.delta * time - c.force_start));
  rows[0].blo = c.jstart; rows[0].bhi = c.jend; rows[0].wlo = c.jstart - 1; rows[0].whi = c.jend + 1;
  gyf_rows(rows[0], c.u_bc_ymin, c.u_bc_ymin_c1, c.u_bc_ymax, c.u_bc_ymax_c1);
  rows[2].blo = c.jstart; rows[2].bhi = c.jend; rows[2].wlo = c.jstart - 1; rows[2].whi = c.jend + 1;
  gyf_rows(rows[2], c.w_bc_ymin, c.w_bc_ymin_c1, c.w_bc_ymax, wtop);
  SolveRows& V = rows[1];
  V.blo = 2; V.bhi = ny; V.wlo = 1; V.whi = ny + 1;
  V.kind[0] = 0;
  if (c.v_bc_ymin == DGPU_BC_DIRICHLET) {
    V.kind[1] = 1; V.val[1] = c.v_bc_ymin_c1;
    V.kind[2] = 1; V.val[2] = c.v_bc_ymin_c1;
  } else if (c.v_bc_ymin == DGPU_BC_NEUMANN) {
    V.kind[1] = 1; V.val[1] = h->dyf[1] * c.v_bc_ymin_c1;
  } else {
    V.kind[1] = 0;
  }
  if (c.v_bc_ymax == DGPU_BC_DIRICHLET) {
    V.kind[3] = 1; V.val[3] = c.v_bc_ymax_c1;
    V.kind[4] = 1; V.val[4] = c.v_bc_ymax_c1;
  } else if (c.v_bc_ymax == DGPU_BC_NEUMANN) {
    V.kind[4] = 1; V.val[4] = h->dyf[ny] * c.v_bc_ymax_c1;
  } else {
    V.kind[4] = 0;
  }
  for (int n = 0; n < h->nth; ++n) {
    SolveRows& R = rows[3 + n];
    R.blo = c.jstart_th[n]; R.bhi = c.jend_th[n];
    R.wlo = c.jstart_th[n] - 1; R.whi = c.jend_th[n] + 1;
    gyf_rows(R, c.th_bc_ymin[n], c.th_bc_ymin_c1[n], c.th_bc_ymax[n], c.th_bc_ymax_c1[n]);
  }
}

// one FP64 field of each layout for the transform entry points, allocated on first use
int ensure_fft_scratch(dgpu_handle* h) {
  const Geo& g = h->g;
  void* p = nullptr;
  int rc;
  if (!h->d_R) {
    if ((rc = dalloc(h, &p, sizeof(double) * g.pp * g.ny2))) return rc;
    h->d_R = (double*)p;
  }
  if (!h->d_H) {
    if ((rc = dalloc(h, &p, sizeof(cd) * g.ph * g.ny2))) return rc;
    h->d_H = (cd*)p;
  }
  if (!h->d_C) {
    if ((rc = dalloc(h, &p, sizeof(cd) * g.pc * g.ny2))) return rc;
    h->d_C = (cd*)p;
  }
  return 0;
}

// --------------------------------------------------------------------------------
// per-kernel timing
// --------------------------------------------------------------------------------
const char* const kKernelNames[4] = {"k_inv_z", "k_xprod", "k_fwdz_asm", "k_columns"};

hipEvent_t pool_event(dgpu_handle* h) {
  if (!h->ev_pool.empty()) {
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
int tbeg(dgpu_handle* h, int id) {
  if (!h->timing) return 0;
  dgpu_handle::Mark m{id, pool_event(h), pool_event(h)};
  if (!m.a || !m.b) return fail(10, "hipEventCreate");
  HIPCHK(hipEventRecord(m.a, h->st));
  h->marks.push_back(m);
  return 0;
}
int tend(dgpu_handle* h) {
  if (!h->timing) return 0;
  HIPCHK(hipEventRecord(h->marks.back().b, h->st));
  return 0;
}
// after a stream synchronize: fold the recorded pairs into the totals
int harvest(dgpu_handle* h) {
  for (auto& m : h->marks) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m.a, m.b));
    h->t_ms[m.id] += ms;
    h->t_n[m.id] += 1;
    h->ev_pool.push_back(m.a);
    h->ev_pool.push_back(m.b);
  }
  h->marks.clear();
  return 0;
}

// --------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------
#define DG_CASE_N(VAL, ...) \
  case VAL: {               \
    constexpr int N_ = VAL; \
    __VA_ARGS__;            \
  } break;
#define DG_SWITCH_N(n, ...)                                                               \
  switch (n) {                                                                            \
    DG_CASE_N(16, __VA_ARGS__) DG_CASE_N(32, __VA_ARGS__) DG_CASE_N(64, __VA_ARGS__)      \
    DG_CASE_N(128, __VA_ARGS__) DG_CASE_N(256, __VA_ARGS__) DG_CASE_N(512, __VA_ARGS__)   \
    DG_CASE_N(1024, __VA_ARGS__)                                                          \
    default: return fail(2, "unsupported transform length");                              \
  }

template <typename T>
int launch_inv_z(dgpu_handle* h, int nf, const cd* const* src, void* const* dst, int j0, int nj) {
  Ptrs<const cd*, MAXO> s{};
  Ptrs<cx<T>*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = src[f]; d.p[f] = (cx<T>*)dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twz64 : h->twz32);
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI = ti_for(N_, sizeof(cx<T>));
    dim3 grid((h->g.kxr + TI - 1) / TI, nj, nf);
    k_inv_z<T, N_, TI, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_fwd_z(dgpu_handle* h, int nf, void* const* src, cd* const* dst, int j0, int nj) {
  Ptrs<const cx<T>*, MAXO> s{};
  Ptrs<cd*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = (const cx<T>*)src[f]; d.p[f] = dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twz64 : h->twz32);
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI = ti_for(N_, sizeof(cx<T>));
    dim3 grid((h->g.kxr + TI - 1) / TI, nj, nf);
    k_fwd_z<T, N_, TI, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
  });
  HIPCHK(hipGetLastError());
  return 0;
}

constexpr int zp_for(int nx) { return nx >= 2048 ? 1 : 2048 / nx; }

template <typename T>
int launch_inv_x(dgpu_handle* h, int nf, void* const* src, T* const* dst, int j0, int nj) {
  Ptrs<const cx<T>*, MAXO> s{};
  Ptrs<T*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = (const cx<T>*)src[f]; d.p[f] = dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twx64 : h->twx32);
  if (h->g.nz % 2) return fail(2, "nz must be even");
  DG_SWITCH_N(h->g.nx, {
    constexpr int ZP = zp_for(N_);
    const int pairs = h->g.nz / 2;
    if (pairs % ZP) {
      // fall back to one pair per block when nz/2 is smaller than the batch
      dim3 grid(pairs, nj, nf);
      k_inv_x<T, N_, 1, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
    } else {
      dim3 grid(pairs / ZP, nj, nf);
      k_inv_x<T, N_, ZP, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
    }
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_fwd_x(dgpu_handle* h, int nf, T* const* src, void* const* dst, int j0, int nj) {
  Ptrs<const T*, MAXO> s{};
  Ptrs<cx<T>*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = src[f]; d.p[f] = (cx<T>*)dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twx64 : h->twx32);
  DG_SWITCH_N(h->g.nx, {
    constexpr int ZP = zp_for(N_);
    const int pairs = h->g.nz / 2;
    if (pairs % ZP) {
      dim3 grid(pairs, nj, nf);
      k_fwd_x<T, N_, 1, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
    } else {
      dim3 grid(pairs / ZP, nj, nf);
      k_fwd_x<T, N_, ZP, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
    }
  });
  HIPCHK(hipGetLastError());
  return 0;
}

constexpr int xprod_nt(int nx) { return nx <= 32 ? 64 : (nx == 64 ? 128 : 256); }

template <typename T, int NTH>
int launch_xprod(dgpu_handle* h, XProdArgs& a, int nj) {
  a.tw = sizeof(T) == 8 ? h->twx64 : h->twx32;
  DG_SWITCH_N(h->g.nx, {
    constexpr size_t lds = sizeof(cx<T>) * 3 * (3 + NTH) * N_;
    if constexpr (lds > 163840) {
      return fail(2, "this nx needs the mixed-precision mode (x pass exceeds LDS in FP64)");
    } else {
      constexpr int NT = xprod_nt(N_);
      dim3 grid(h->g.nz / 2, nj, 1);
      k_xprod<T, N_, NT, NTH><<<grid, NT, 0, h->st>>>(a);
    }
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T, int NTH>
int launch_fwdz_asm(dgpu_handle* h, AsmArgs& a, int nj) {
  a.tw = sizeof(T) == 8 ? h->twz64 : h->twz32;
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI = ti_for(N_, sizeof(cx<T>));
    dim3 grid((h->g.kxr + TI - 1) / TI, nj, 1);
    k_fwdz_asm<T, N_, TI, 256, NTH><<<grid, 256, 0, h->st>>>(a);
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <int NTH>
int launch_columns(dgpu_handle* h, ColArgs& a) {
  const int ny2 = h->g.ny2;
  const int tc = (ny2 * 72 * 8 <= 96 * 1024) ? 8 : 4;
  const size_t lds = (size_t)tc * ny2 * (4 * sizeof(cd) + sizeof(double));
  if (lds > 160 * 1024) return fail(2, "ny too large for the column kernel");
  const int blocks = (a.ncols + tc - 1) / tc;
  if (tc == 8) {
    HIPCHK(hipFuncSetAttribute((const void*)&k_columns<8, NTH>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_columns<8, NTH><<<blocks, 64, lds, h->st>>>(a);
  } else {
    HIPCHK(hipFuncSetAttribute((const void*)&k_columns<4, NTH>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_columns<4, NTH><<<blocks, 64, lds, h->st>>>(a);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename F>
int by_nth(int nth, F&& f) {
  switch (nth) {
    case 0: return f(std::integral_constant<int, 0>{});
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    default: return fail(2, "n_th > 2 is not built in this library");
  }
}

int validate(const dgpu_config* c) {
  if (!c) return fail(1, "null config");
  if (!pow2_in(c->nx, 16, 1024) || !pow2_in(c->nz, 16, 1024))
    return fail(2, "nx and nz must be powers of two in [16, 1024]");
  if (c->ny < 4 || c->ny > 2045) return fail(2, "ny must be in [4, 2045]");
  if (c->n_th < 0 || c->n_th > 2) return fail(2, "n_th must be 0, 1 or 2");
  if (c->precision != DGPU_FP64 && c->precision != DGPU_MIXED) return fail(2, "bad precision");
  if (!c->gy || !c->gyf || !c->dy || !c->dyf || !c->kx || !c->kz) return fail(1, "null metric pointer");
  if (c->jstart < 1 || c->jstart > 2 || c->jend != c->ny) return fail(2, "jstart/jend do not match one rank");
  for (int n = 0; n < c->n_th; ++n) {
    if (c->jstart_th[n] < 1 || c->jstart_th[n] > 2 || c->jend_th[n] != c->ny)
      return fail(2, "jstart_th/jend_th do not match one rank");
    if (c->pr[n] == 0.0) return fail(2, "Pr must be non-zero");
  }
  if (c->delta == 0.0) return fail(2, "delta must be non-zero");
  for (int j = 1; j <= c->ny + 1; ++j)
    if (!(c->dy[j] > 0.0)) return fail(2, "dy must be positive");
  for (int j = 0; j <= c->ny + 1; ++j)
    if (!(c->dyf[j] > 0.0)) return fail(2, "dyf must be positive");
  return 0;
}

size_t fortran_doubles(const dgpu_config* c) {
  return (size_t)2 * (c->nx / 2 + 1) * (c->nz + 2) * (c->ny + 2);
}

// Fortran cu(0:Nxp, 0:Nz+1, 0:Nyp+1) <-> compact [j][k][i]
void pack_fourier(const dgpu_handle* h, const double* fort, cd* compact) {
  const Geo& g = h->g;
  const int ldi = g.nx / 2 + 1, ldk = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int k = 0; k < g.kzr; ++k) {
      const double* src = fort + 2 * (((size_t)j * ldk + k) * ldi);
      cd* dst = compact + (size_t)j * g.pc + (size_t)k * g.kxr;
      std::memcpy(dst, src, sizeof(cd) * g.kxr);
    }
}
void unpack_fourier(const dgpu_handle* h, const cd* compact, double* fort) {
  const Geo& g = h->g;
  const int ldi = g.nx / 2 + 1, ldk = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int k = 0; k < g.kzr; ++k) {
      double* dst = fort + 2 * (((size_t)j * ldk + k) * ldi);
      std::memcpy(dst, compact + (size_t)j * g.pc + (size_t)k * g.kxr, sizeof(cd) * g.kxr);
      std::memset(dst + 2 * g.kxr, 0, sizeof(cd) * (ldi - g.kxr));
    }
}

int upload_field(dgpu_handle* h, const double* fort, cd* dev) {
  if (!fort) return fail(1, "null array");
  cd* stage = (cd*)h->hstage;
  pack_fourier(h, fort, stage);
  HIPCHK(hipMemcpyAsync(dev, stage, sizeof(cd) * h->g.pc * h->g.ny2, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}
int download_field(dgpu_handle* h, const cd* dev, double* fort) {
  if (!fort) return fail(1, "null array");
  cd* stage = (cd*)h->hstage;
  HIPCHK(hipMemcpyAsync(stage, dev, sizeof(cd) * h->g.pc * h->g.ny2, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  unpack_fourier(h, stage, fort);
  return 0;
}

int substep(dgpu_handle* h, int rk, double delta_t, double time, int time_step, int create_new_flow) {
  const dgpu_config& c = h->c;
  if (rk < 1 || rk > 3) return fail(3, "rk_step must be 1, 2 or 3");
  if (c.use_les && (!create_new_flow || time_step > 100))
    return fail(4, "the LES branch (les_chan) is not available in this library build");
  int rc = ensure_factors(h, delta_t);
  if (rc) return rc;
  const double hb[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  const double bb[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0};
  const double zb[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0};
  const int r = rk - 1;
  const double temp1 = 0.5 * c.nu * hb[r], temp3 = zb[r] * hb[r], temp4 = hb[r], temp5 = bb[r] * hb[r];
  const Geo& g = h->g;
  const int nf = 3 + h->nth;
  int jlo = c.jstart < 2 ? c.jstart : 2;
  for (int n = 0; n < h->nth; ++n) jlo = c.jstart_th[n] < jlo ? c.jstart_th[n] : jlo;
  const int jhi = c.ny;
  const int nj = jhi - jlo + 1;
  const bool mixed = h->prec == DGPU_MIXED;

  // 1. z inverse of the state
  const cd* srcs[MAXF];
  for (int f = 0; f < nf; ++f) srcs[f] = h->U[f];
  if ((rc = tbeg(h, 0))) return rc;
  rc = mixed ? launch_inv_z<float>(h, nf, srcs, h->Hin, 0, g.ny2)
             : launch_inv_z<double>(h, nf, srcs, h->Hin, 0, g.ny2);
  if (rc || (rc = tend(h))) return rc;
  // 2. fused x pass
  XProdArgs xa{};
  xa.g = g; xa.jlo = jlo; xa.jstart = c.jstart; xa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { xa.jstart_th[n] = c.jstart_th[n]; xa.jend_th[n] = c.jend_th[n]; }
  xa.dy = h->d_dy; xa.dyf = h->d_dyf; xa.kx = h->d_kx;
  for (int f = 0; f < nf; ++f) xa.hin[f] = h->Hin[f];
  for (int o = 0; o < 6 + 2 * h->nth; ++o) xa.hout[o] = h->Hout[o];
  if ((rc = tbeg(h, 1))) return rc;
  rc = by_nth(h->nth, [&](auto NTHc) {
    constexpr int NTH = decltype(NTHc)::value;
    return mixed ? launch_xprod<float, NTH>(h, xa, nj) : launch_xprod<double, NTH>(h, xa, nj);
  });
  if (rc || (rc = tend(h))) return rc;
  // 3. z forward + assembly
  AsmArgs aa{};
  aa.g = g; aa.jlo = jlo; aa.rk = rk; aa.jstart = c.jstart; aa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) {
    aa.jstart_th[n] = c.jstart_th[n]; aa.jend_th[n] = c.jend_th[n];
    aa.pr[n] = c.pr[n]; aa.dthdx[n] = c.dthdx[n]; aa.dthdz[n] = c.dthdz[n];
  }
  aa.front = c.flavor_front; aa.f_type = c.f_type;
  aa.nu = c.nu; aa.beta = c.beta; aa.nu_v_scale = c.nu_v_scale; aa.ro_inv = c.ro_inv; aa.delta = c.delta;
  aa.grav_x = c.grav_x; aa.grav_y = c.grav_y; aa.grav_z = c.grav_z; aa.ly = c.ly;
  aa.px0 = c.px0; aa.omega0 = c.omega0; aa.amp_omega0 = c.amp_omega0; aa.time = time;
  aa.temp1 = temp1; aa.temp3 = temp3; aa.temp4 = temp4; aa.temp5 = temp5;
  aa.dy = h->d_dy; aa.dyf = h->d_dyf; aa.gy = h->d_gy; aa.gyf = h->d_gyf; aa.kx = h->d_kx; aa.kz = h->d_kz;
  for (int f = 0; f < nf; ++f) { aa.u[f] = h->U[f]; aa.cf[f] = h->CF[f]; aa.cr[f] = h->CR[f]; }
  aa.p = h->P;
  for (int o = 0; o < 6 + 2 * h->nth; ++o) aa.h[o] = h->Hout[o];
  if ((rc = tbeg(h, 2))) return rc;
  rc = by_nth(h->nth, [&](auto NTHc) {
    constexpr int NTH = decltype(NTHc)::value;
    return mixed ? launch_fwdz_asm<float, NTH>(h, aa, nj) : launch_fwdz_asm<double, NTH>(h, aa, nj);
  });
  if (rc || (rc = tend(h))) return rc;
  // 4. columns
  ColArgs ca{};
  ca.g = g; ca.ncols = (int)g.pc; ca.jstart = c.jstart; ca.jend = c.jend; ca.temp4 = temp4;
  solve_rows(h, time, ca.rows);
  ca.fac = h->d_fac + (size_t)r * MAXF * 3 * g.ny2;
  ca.dy = h->d_dy; ca.dyf = h->d_dyf; ca.kx = h->d_kx; ca.kz = h->d_kz;
  for (int f = 0; f < nf; ++f) { ca.u[f] = h->U[f]; ca.cr[f] = h->CR[f]; }
  ca.p = h->P;
  ca.do_cn = 1;
  if ((rc = tbeg(h, 3))) return rc;
  rc = by_nth(h->nth, [&](auto NTHc) { return launch_columns<decltype(NTHc)::value>(h, ca); });
  if (rc) return rc;
  return tend(h);
}

}  // namespace

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" {

const char* dgpu_last_error(void) { return g_err.c_str(); }

int64_t dgpu_field_doubles(const dgpu_config* cfg) { return cfg ? (int64_t)fortran_doubles(cfg) : -1; }

int dgpu_create(const dgpu_config* cfg, dgpu_handle** out) {
  if (!out) return fail(1, "null handle pointer");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  dgpu_handle* h = new dgpu_handle();
  h->c = *cfg;
  const int n2 = cfg->ny + 2;
  h->gy.assign(cfg->gy, cfg->gy + n2);
  h->gyf.assign(cfg->gyf, cfg->gyf + n2);
  h->dy.assign(cfg->dy, cfg->dy + n2);
  h->dyf.assign(cfg->dyf, cfg->dyf + n2);
  h->kx.assign(cfg->kx, cfg->kx + cfg->nx / 2);
  h->kz.assign(cfg->kz, cfg->kz + 2 * (cfg->nz / 3) + 1);
  h->c.gy = h->c.gyf = h->c.dy = h->c.dyf = h->c.kx = h->c.kz = nullptr;
  h->nth = cfg->n_th;
  h->prec = cfg->precision;
  Geo& g = h->g;
  g.nx = cfg->nx; g.nz = cfg->nz; g.ny = cfg->ny; g.ny2 = n2;
  g.nkx = cfg->nx / 3; g.nkz = cfg->nz / 3; g.kxr = g.nkx + 1; g.kzr = 2 * g.nkz + 1;
  g.pc = (long long)g.kxr * g.kzr;
  g.ph = (long long)g.nz * g.kxr;
  g.pp = (long long)g.nz * g.nx;
  auto bail = [&](int code) {
    dgpu_destroy(h);
    return code;
  };
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return bail(fail(10, "hipStreamCreate"));
  if (hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess)
    return bail(fail(10, "hipEventCreate"));
  const size_t cbytes = sizeof(cd) * g.pc * n2;
  const size_t hsz = (h->prec == DGPU_MIXED) ? sizeof(cf) : sizeof(cd);
  const size_t hbytes = hsz * g.ph * n2;
  const int nf = 3 + h->nth, no = 6 + 2 * h->nth;
  void* p;
#define ALLOC(dst, bytes)                          \
  do {                                             \
    if ((rc = dalloc(h, &p, (bytes)))) return bail(rc); \
    dst = decltype(dst)(p);                        \
  } while (0)
  ALLOC(h->d_gy, 8 * n2); ALLOC(h->d_gyf, 8 * n2); ALLOC(h->d_dy, 8 * n2); ALLOC(h->d_dyf, 8 * n2);
  ALLOC(h->d_kx, 8 * (size_t)g.kxr); ALLOC(h->d_kz, 8 * (size_t)g.kzr);
  for (int f = 0; f < nf; ++f) { ALLOC(h->U[f], cbytes); ALLOC(h->CF[f], cbytes); ALLOC(h->CR[f], cbytes); }
  ALLOC(h->P, cbytes);
  for (int f = 0; f < nf; ++f) ALLOC(h->Hin[f], hbytes);
  for (int o = 0; o < no; ++o) ALLOC(h->Hout[o], hbytes);
  ALLOC(h->twx64, sizeof(cd) * g.nx); ALLOC(h->twz64, sizeof(cd) * g.nz);
  ALLOC(h->twx32, sizeof(cf) * g.nx); ALLOC(h->twz32, sizeof(cf) * g.nz);
  ALLOC(h->d_fac, sizeof(double) * 3 * MAXF * 3 * n2);
#undef ALLOC
  h->fort_doubles = fortran_doubles(cfg);
  if (hipHostMalloc((void**)&h->hstage, sizeof(double) * h->fort_doubles, hipHostMallocDefault) != hipSuccess)
    return bail(fail(11, "hipHostMalloc"));
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->st) == hipSuccess;
  };
  auto tx64 = twiddles<double>(g.nx), tz64 = twiddles<double>(g.nz);
  auto tx32 = twiddles<float>(g.nx), tz32 = twiddles<float>(g.nz);
  bool ok = up(h->d_gy, h->gy.data(), 8 * n2) && up(h->d_gyf, h->gyf.data(), 8 * n2) &&
            up(h->d_dy, h->dy.data(), 8 * n2) && up(h->d_dyf, h->dyf.data(), 8 * n2) &&
            up(h->d_kx, h->kx.data(), 8 * (size_t)g.kxr) && up(h->d_kz, h->kz.data(), 8 * (size_t)g.kzr) &&
            up(h->twx64, tx64.data(), sizeof(cd) * g.nx) && up(h->twz64, tz64.data(), sizeof(cd) * g.nz) &&
            up(h->twx32, tx32.data(), sizeof(cf) * g.nx) && up(h->twz32, tz32.data(), sizeof(cf) * g.nz);
  if (!ok || hipStreamSynchronize(h->st) != hipSuccess) return bail(fail(10, "upload of constants failed"));
  *out = h;
  g_err.clear();
  return 0;
}

int dgpu_destroy(dgpu_handle* h) {
  if (!h) return 0;
  if (h->st) (void)hipStreamSynchronize(h->st);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->hstage) (void)hipHostFree(h->hstage);
  for (auto& m : h->marks) { h->ev_pool.push_back(m.a); h->ev_pool.push_back(m.b); }
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return 0;
}

int dgpu_upload_state(dgpu_handle* h, const double* cu1, const double* cu2, const double* cu3,
                      const double* cp, const double* cth) {
  if (!h) return fail(1, "null handle");
  int rc;
  if ((rc = upload_field(h, cu1, h->U[0])) || (rc = upload_field(h, cu2, h->U[1])) ||
      (rc = upload_field(h, cu3, h->U[2])) || (rc = upload_field(h, cp, h->P)))
    return rc;
  for (int n = 0; n < h->nth; ++n)
    if ((rc = upload_field(h, cth ? cth + (size_t)n * h->fort_doubles : nullptr, h->U[3 + n]))) return rc;
  return 0;
}

int dgpu_download_state(dgpu_handle* h, double* cu1, double* cu2, double* cu3, double* cp, double* cth) {
  if (!h) return fail(1, "null handle");
  int rc;
  if ((rc = download_field(h, h->U[0], cu1)) || (rc = download_field(h, h->U[1], cu2)) ||
      (rc = download_field(h, h->U[2], cu3)) || (rc = download_field(h, h->P, cp)))
    return rc;
  for (int n = 0; n < h->nth; ++n)
    if ((rc = download_field(h, h->U[3 + n], cth ? cth + (size_t)n * h->fort_doubles : nullptr))) return rc;
  return 0;
}

int dgpu_upload_rk(dgpu_handle* h, const double* cf1, const double* cf2, const double* cf3, const double* cfth) {
  if (!h) return fail(1, "null handle");
  int rc;
  if ((rc = upload_field(h, cf1, h->CF[0])) || (rc = upload_field(h, cf2, h->CF[1])) ||
      (rc = upload_field(h, cf3, h->CF[2])))
    return rc;
  for (int n = 0; n < h->nth; ++n)
    if ((rc = upload_field(h, cfth ? cfth + (size_t)n * h->fort_doubles : nullptr, h->CF[3 + n]))) return rc;
  return 0;
}

int dgpu_download_rk(dgpu_handle* h, double* cf1, double* cf2, double* cf3, double* cfth) {
  if (!h) return fail(1, "null handle");
  int rc;
  if ((rc = download_field(h, h->CF[0], cf1)) || (rc = download_field(h, h->CF[1], cf2)) ||
      (rc = download_field(h, h->CF[2], cf3)))
    return rc;
  for (int n = 0; n < h->nth; ++n)
    if ((rc = download_field(h, h->CF[3 + n], cfth ? cfth + (size_t)n * h->fort_doubles : nullptr))) return rc;
  return 0;
}

int dgpu_rk_substep(dgpu_handle* h, int32_t rk_step, double delta_t, double time, int32_t time_step,
                    int32_t create_new_flow) {
  if (!h) return fail(1, "null handle");
  HIPCHK(hipEventRecord(h->e0, h->st));
  int rc = substep(h, rk_step, delta_t, time, time_step, create_new_flow);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->e1, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return harvest(h);
}

int dgpu_step(dgpu_handle* h, double delta_t, double time, int32_t time_step, int32_t create_new_flow) {
  if (!h) return fail(1, "null handle");
  HIPCHK(hipEventRecord(h->e0, h->st));
  for (int rk = 1; rk <= 3; ++rk) {
    int rc = substep(h, rk, delta_t, time, time_step, create_new_flow);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->e1, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return harvest(h);
}

int dgpu_rem_div(dgpu_handle* h) {
  if (!h) return fail(1, "null handle");
  ColArgs ca{};
  ca.g = h->g; ca.ncols = (int)h->g.pc; ca.jstart = h->c.jstart; ca.jend = h->c.jend; ca.temp4 = 1.0;
  ca.dy = h->d_dy; ca.dyf = h->d_dyf; ca.kx = h->d_kx; ca.kz = h->d_kz;
  for (int f = 0; f < 3; ++f) ca.u[f] = h->U[f];
  ca.p = h->P;
  ca.do_cn = 0;
  int rc = launch_columns<0>(h, ca);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_fft_xz_to_fourier(dgpu_handle* h, const double* phys, double* four) {
  if (!h || !phys || !four) return fail(1, "null argument");
  const Geo& g = h->g;
  int rc = ensure_fft_scratch(h);
  if (rc) return rc;
  // physical Fortran layout u(0:Nx+1, 0:Nzp+1, 0:Nyp+1) -> [j][z][x]
  std::vector<double> tmp((size_t)g.pp * g.ny2);
  for (int j = 0; j < g.ny2; ++j)
    for (int z = 0; z < g.nz; ++z)
      std::memcpy(&tmp[((size_t)j * g.nz + z) * g.nx], phys + ((size_t)j * (g.nz + 2) + z) * (g.nx + 2),
                  sizeof(double) * g.nx);
  HIPCHK(hipMemcpyAsync(h->d_R, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->st));
  double* rs[1] = {h->d_R};
  void* hs[1] = {h->d_H};
  cd* cs[1] = {h->d_C};
  if ((rc = launch_fwd_x<double>(h, 1, rs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_fwd_z<double>(h, 1, hs, cs, 0, g.ny2))) return rc;
  return download_field(h, h->d_C, four);
}

int dgpu_fft_xz_to_physical(dgpu_handle* h, const double* four, double* phys) {
  if (!h || !phys || !four) return fail(1, "null argument");
  const Geo& g = h->g;
  int rc = ensure_fft_scratch(h);
  if (rc) return rc;
  if ((rc = upload_field(h, four, h->d_C))) return rc;
  const cd* cs[1] = {h->d_C};
  void* hs[1] = {h->d_H};
  double* rs[1] = {h->d_R};
  if ((rc = launch_inv_z<double>(h, 1, cs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_inv_x<double>(h, 1, hs, rs, 0, g.ny2))) return rc;
  std::vector<double> tmp((size_t)g.pp * g.ny2);
  HIPCHK(hipMemcpyAsync(tmp.data(), h->d_R, tmp.size() * 8, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int j = 0; j < g.ny2; ++j)
    for (int z = 0; z < g.nz; ++z)
      std::memcpy(phys + ((size_t)j * (g.nz + 2) + z) * (g.nx + 2), &tmp[((size_t)j * g.nz + z) * g.nx],
                  sizeof(double) * g.nx);
  return 0;
}

int dgpu_timing_enable(dgpu_handle* h, int32_t enable) {
  if (!h) return fail(1, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = harvest(h);
  if (rc) return rc;
  h->timing = enable != 0;
  for (int k = 0; k < 4; ++k) { h->t_ms[k] = 0.0; h->t_n[k] = 0; }
  return 0;
}

int dgpu_timing_read(dgpu_handle* h, int32_t kernel, double* total_ms, int64_t* launches, const char** name) {
  if (!h || !total_ms || !launches) return fail(1, "null argument");
  if (kernel < 0 || kernel > 3) return fail(2, "kernel id must be 0..3");
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = harvest(h);
  if (rc) return rc;
  *total_ms = h->t_ms[kernel];
  *launches = h->t_n[kernel];
  if (name) *name = kKernelNames[kernel];
  return 0;
}

int dgpu_synchronize(dgpu_handle* h) {
  if (!h) return fail(1, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_last_step_ms(dgpu_handle* h, double* ms) {
  if (!h || !ms) return fail(1, "null argument");
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, h->e0, h->e1));
  *ms = f;
  return 0;
}

}  // extern "C"
