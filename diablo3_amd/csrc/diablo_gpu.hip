// diablo_gpu.hip — C ABI of the MI355X RKW3 + Crank–Nicolson channel step.
//
// One DNS substep (rk_chan_1, for/solver.f90:366-1273) is five launches:
//   k_inv_z     packed Fourier state -> x-spectra per z row (z inverse transform)
//   k_xprod     x inverse transform, nonlinear products with their y stencils,
//               x forward transform, ikx factors (physical space stays in LDS)
//   k_fwdz_nl   z forward transform of the products -> nonlinear tendencies
//   k_cn        per (kx,kz) column: Fourier-space assembly of cf (RK memory) and
//               cr (right-hand side incl. explicit CN half), then the CN solves
//   k_project   per (kx,kz) column: rem_div_chan and the pressure update
// The CN operator of DNS is independent of x and z, so solving it per Fourier
// column is the same linear map as the reference's per-physical-column solve
// followed by the forward transform; it saves the 8 transforms the reference
// spends on cr -> r and u -> cu.
// The LES substep (use_LES past the spin-up gate) adds the AMD closure and solves
// CN per physical column (substep_les, kernels in dgpu_les.hpp).
#include "../../include/diablo_gpu.h"
#include "../../include/dgpu_closure.h"
#include "dgpu_kernels.hpp"
#include "dgpu_les.hpp"
#include "dgpu_closure.hpp"
#include "dgpu_stats.hpp"
#include "dgpu_courant.hpp"
#include "dgpu_boxfilter.hpp"
#include "dgpu_error.hpp"

#include <algorithm>
#include <cstdlib>
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

int dgpu::set_error(int code, const std::string& msg) { return fail(code, msg); }

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
  void* NL[MAXF] = {};     // nonlinear tendencies (complex T), packed layout
  void* Hin[MAXF] = {};
  // FP64 x-spectra, z-derivative x-spectra and physical v, dv/dx, dv/dz: the
  // physical-space CN stage of LES, statistics and courant.  On an FP64 handle
  // these alias Hin / HinZ / Pg; on a mixed one they are separate buffers.
  void* HinD[MAXF] = {};
  void* HinZD[MAXF] = {};
  void* PgD[3 * MAXF] = {};
  void* Hout[MAXO] = {};
  // persistent kernels: blocks to launch (a few per CU)
  int resident_blocks = 1024;
  void* d_box = nullptr;  // dgpu_box_filter: y weights, row ranges, output
  size_t box_bytes = 0;
  int cus = 256;        // compute units of the device
  int xm_resident = 0;  // k_xmarch blocks resident at once on the device
  int fz_resident = 0;  // persistent k_fwdz_nl blocks resident at once
  void *twx64 = nullptr, *twz64 = nullptr, *twx32 = nullptr, *twz32 = nullptr;
  double* d_fac = nullptr;
  double fac_dt = NAN;
  double* d_rowc = nullptr;  // [ny2][RC_COUNT] assembly row coefficients
  int dev = 0;             // the device current at dgpu_create
  unsigned char* noalloc = nullptr;  // that device's address of `scratch`
  double* d_R = nullptr;   // one physical FP64 field (transform entry points)
  cd* d_H = nullptr;       // one FP64 x-spectrum field
  cd* d_C = nullptr;       // one packed field
  // LES branch (use_les): physical fields of u, du/dx, du/dz per transported
  // variable, z-derivative x-spectra, eddy viscosity/diffusivity, SGS additions,
  // physical-space CN right-hand sides and scratch
  void* HinZ[MAXF] = {};    // complex T
  void* Pg[3 * MAXF] = {};  // T
  void* NuT = nullptr;
  void* KapT[MAXTH] = {};
  void* Aug[MAXA] = {};
  cd* CR[MAXF] = {};
  double* Rp[MAXF] = {};
  double* CPs = nullptr;
  double* d_dlt = nullptr;  // [2][ny2]: delta_yf, delta_y of the AMD filter width
  double* d_sigma = nullptr;  // filter_chan's horizontal filter, packed [kzr][kxp]
  // les_model_type 6: the network (BF16 forward weights w1t, w2t, w3t; FP32
  // parameter block for the biases; normaliser) and the per-point coefficients
  unsigned short* d_netw = nullptr;
  float* d_netp = nullptr;
  float* d_netnorm = nullptr;
  int have_net = 0;
  void *Cgy = nullptr, *Kgy = nullptr;  // coefficients at GY, precision T
  unsigned long long* d_ccount = nullptr;  // dgpu_closure_counts' sum
  // variable_dt: courant's CFL limit, reduced per block on the device at rk_step 3
  int variable_dt = 0, update_dt = 1, cfl_ran = 0;
  double cfl = 0.5;
  double* d_cfl = nullptr;  // block minima of the last courant evaluation
  // statistics (allocated on the first dgpu_stats call of a DNS handle)
  double *d_zm = nullptr, *d_zs = nullptr, *d_prof = nullptr;
  double* d_zhist = nullptr;  // z* histograms [5 + 2 MAXTH][nx]
  double* d_ftx = nullptr;    // FTx_uu [nx / 2]
  double* d_movie = nullptr;  // mean_xz planes, then movie slices
  double* hstage = nullptr;  // pinned staging, one Fortran array
  size_t fort_doubles = 0;
  std::vector<void*> allocs;
  // dgpu_step as a captured HIP graph: the three substeps' launches, replayed
  // while nothing that enters their arguments changes (delta_t, the LES gate)
  bool graph_on = true;
  hipGraphExec_t gexec = nullptr;
  double g_dt = NAN;
  int g_les = -1;
  // per-kernel timing (measurement only)
  bool timing = false;
  int timing_every = 1;       // record the kernel events on every n-th dgpu_step call
  long long timing_calls = 0;
  bool timing_now = true;     // this call records
  std::vector<hipEvent_t> ev_pool;
  struct Mark { int id; hipEvent_t a, b; };
  std::vector<Mark> marks;
  double t_ms[16] = {};
  long long t_n[16] = {};
};

// Transform scratch of dgpu_fft_xz_to_fourier_noalloc: the compact physical field,
// the x-spectra and the packed spectrum of one grid, in one file-scope device
// array (each GPU the library runs on holds its own copy until the process
// ends).  Host code never passes `scratch` itself to a kernel: its per-device
// address comes from hipGetSymbolAddress at dgpu_create.
constexpr size_t kNoallocScratchBytes = (size_t)DGPU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// The address of `scratch` on device `dev`, which must be current; resolved once
// per device.
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(2, "device ordinal out of range for the transform scratch");
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(10, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = (unsigned char*)ptr;
  }
  *out = g_scratch_addr[dev];
  return 0;
}

// Fortran physical array u(0:Nx+1, 0:Nzp+1, 0:Nyp+1) -> compact [j][z][x]
__global__ void k_pack_phys(const double* __restrict__ fort, double* __restrict__ out, Geo g) {
  const long long n = (long long)g.ny2 * g.nz * g.nx;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % g.nx);
    const long long r = t / g.nx;
    const int z = (int)(r % g.nz), j = (int)(r / g.nz);
    out[t] = fort[((long long)j * (g.nz + 2) + z) * (g.nx + 2) + x];
  }
}

// packed spectrum [j][k][kxp] -> Fortran cu(0:Nxp, 0:Nz+1, 0:Nyp+1) on the rows
// k < 2 Nkz + 1, the columns past Nkx zeroed (the host unpack_fourier's map)
__global__ void k_unpack_four(const cd* __restrict__ comp, cd* __restrict__ fort, Geo g) {
  const int ldi = g.nx / 2 + 1;
  const long long n = (long long)g.ny2 * g.kzr * ldi;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % ldi);
    const long long r = t / ldi;
    const int k = (int)(r % g.kzr), j = (int)(r / g.kzr);
    fort[((long long)j * (g.nz + 2) + k) * ldi + i] =
        i < g.kxr ? comp[(long long)j * g.pc + (long long)k * g.kxp + i] : cd{0.0, 0.0};
  }
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

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

int pcr_levels(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

struct ColShape;
int cn_chunk(const dgpu_handle* h);

// Thomas factorisation of one tridiagonal system (rows 0..n-1, a[0] = c[n-1] = 0)
// for the chunked solve of k_cn: per row j, out[5j..5j+4] = mu_j, beta_j, gamma_j,
// F_j, G_j with (thomas_forward/backward_real_mpi, for/solver.f90:1691-1852)
//   b'_0 = b_0, mu_j = -a_j / b'_{j-1}, b'_j = b_j + mu_j c_{j-1},
//   forward y_j = g_j + mu_j y_{j-1}, backward x_j = beta_j y_j + gamma_j x_{j+1},
//   beta_j = 1/b'_j, gamma_j = -c_j/b'_j,
// and, for chunks of L rows, F_j = prod of mu over [chunk start, j] and G_j = prod
// of gamma over [j, chunk end].  Long double, stored as double; self-checked by
// running the chunked algorithm against a long-double Thomas solve.
int thomas_chunk_coefficients(const std::vector<double>& a, const std::vector<double>& b,
                              const std::vector<double>& c, int L, double* out) {
  const int n = (int)b.size();
  std::vector<long double> mu(n), bp(n), be(n), ga(n);
  bp[0] = b[0];
  mu[0] = 0;
  for (int j = 1; j < n; ++j) {
    if (bp[j - 1] == 0) return fail(20, "zero pivot in the Crank-Nicolson system");
    mu[j] = -(long double)a[j] / bp[j - 1];
    bp[j] = b[j] + mu[j] * c[j - 1];
  }
  for (int j = 0; j < n; ++j) {
    if (bp[j] == 0) return fail(20, "singular Crank-Nicolson system");
    be[j] = 1.0L / bp[j];
    ga[j] = -(long double)c[j] / bp[j];
  }
  for (int j = 0; j < n; ++j) {
    const int s0 = j / L * L, e0 = std::min(s0 + L, n) - 1;
    long double F = 1, G = 1;
    for (int i = s0; i <= j; ++i) F *= mu[i];
    for (int i = j; i <= e0; ++i) G *= ga[i];
    out[5 * j + 0] = (double)mu[j];
    out[5 * j + 1] = (double)be[j];
    out[5 * j + 2] = (double)ga[j];
    out[5 * j + 3] = (double)F;
    out[5 * j + 4] = (double)G;
  }
  // self-check: the chunked algorithm in double against a long-double Thomas solve
  std::vector<double> d(n), yl(n);
  std::vector<long double> g(n), x(n);
  for (int j = 0; j < n; ++j) d[j] = std::sin(0.37 * j + 0.1) + 0.5;
  g[0] = d[0];
  for (int j = 1; j < n; ++j) g[j] = d[j] + mu[j] * g[j - 1];
  x[n - 1] = g[n - 1] * be[n - 1];
  for (int j = n - 2; j >= 0; --j) x[j] = be[j] * g[j] + ga[j] * x[j + 1];
  const int P = (n + L - 1) / L;
  std::vector<double> yin(P + 1, 0.0), xin(P + 1, 0.0);
  for (int q = 0; q < P; ++q) {  // forward: local + carries in chunk order
    double pr = 0.0;
    for (int j = q * L; j < std::min((q + 1) * L, n); ++j) { yl[j] = d[j] + out[5 * j] * pr; pr = yl[j]; }
    const int e0 = std::min((q + 1) * L, n) - 1;
    yin[q + 1] = yl[e0] + out[5 * e0 + 3] * yin[q];
  }
  for (int j = 0; j < n; ++j) yl[j] = yl[j] + out[5 * j + 3] * yin[j / L];
  std::vector<double> xl(n);
  for (int q = P - 1; q >= 0; --q) {
    double nx = 0.0;
    const int s0 = q * L, e0 = std::min((q + 1) * L, n) - 1;
    for (int j = e0; j >= s0; --j) { xl[j] = out[5 * j + 1] * yl[j] + out[5 * j + 2] * nx; nx = xl[j]; }
    xin[q] = xl[s0] + out[5 * s0 + 4] * xin[q + 1];
  }
  long double err = 0, nrm = 0;
  for (int j = 0; j < n; ++j) {
    const long double v = xl[j] + out[5 * j + 4] * xin[j / L + 1];
    err = std::max(err, std::fabs(v - x[j]));
    nrm = std::max(nrm, std::fabs(x[j]));
  }
  if (!(err <= 1e-12L * (nrm > 0 ? nrm : 1.0L))) return fail(21, "chunked Thomas disagrees with Thomas");
  return 0;
}

int ensure_factors(dgpu_handle* h, double delta_t) {
  if (h->fac_dt == delta_t) return 0;
  const int n2 = h->c.ny + 2, nf = 3 + h->nth;
  const size_t per_field = (size_t)5 * n2;
  std::vector<double> buf((size_t)3 * MAXF * per_field, 0.0), a, b, cc;
  const double hb[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  const int L = cn_chunk(h);
  if (L <= 0) return fail(2, "ny too large for the column kernels");
  for (int r = 0; r < 3; ++r) {
    const double temp1 = 0.5 * h->c.nu * hb[r];
    for (int F = 0; F < nf; ++F) {
      cn_matrix(h, F, temp1, a, b, cc);
      a[0] = 0.0;
      cc[n2 - 1] = 0.0;
      int rc = thomas_chunk_coefficients(a, b, cc, L, &buf[((size_t)r * MAXF + F) * per_field]);
      if (rc) return rc;
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
constexpr int NKERN = 12;
const char* const kKernelNames[NKERN] = {"k_inv_z",   "k_xprod",  "k_fwdz_nl", "k_cn",      "k_project",
                                         "k_bc_post", "les_grad", "k_amd",     "k_cn_phys", "les_cn_fft",
                                         "k_closure", "k_xmarch"};

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
  if (!h->timing || !h->timing_now) return 0;
  dgpu_handle::Mark m{id, pool_event(h), pool_event(h)};
  if (!m.a || !m.b) return fail(10, "hipEventCreate");
  HIPCHK(hipEventRecord(m.a, h->st));
  h->marks.push_back(m);
  return 0;
}
int tend(dgpu_handle* h) {
  if (!h->timing || !h->timing_now) return 0;
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
int launch_inv_z(dgpu_handle* h, int nf, const cd* const* src, void* const* dst, int j0, int nj,
                 unsigned dzmask = 0) {
  Ptrs<const cd*, MAXO> s{};
  Ptrs<cx<T>*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = src[f]; d.p[f] = (cx<T>*)dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twz64 : h->twz32);
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI = ti_for(N_, sizeof(cx<T>));
    dim3 grid((h->g.kxr + TI - 1) / TI, nj, nf);
    k_inv_z<T, N_, TI, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw, h->d_kz, dzmask);
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T, bool ACC = false>
int launch_fwd_z(dgpu_handle* h, int nf, void* const* src, cd* const* dst, int j0, int nj) {
  Ptrs<const cx<T>*, MAXO> s{};
  Ptrs<cd*, MAXO> d{};
  for (int f = 0; f < nf; ++f) { s.p[f] = (const cx<T>*)src[f]; d.p[f] = dst[f]; }
  const cx<T>* tw = (const cx<T>*)(sizeof(T) == 8 ? h->twz64 : h->twz32);
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI = ti_for(N_, sizeof(cx<T>));
    dim3 grid((h->g.kxr + TI - 1) / TI, nj, nf);
    k_fwd_z<T, N_, TI, 256, ACC><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw);
  });
  HIPCHK(hipGetLastError());
  return 0;
}

constexpr int zp_for(int nx) { return nx >= 2048 ? 1 : 2048 / nx; }

template <typename T>
int launch_inv_x(dgpu_handle* h, int nf, void* const* src, T* const* dst, int j0, int nj,
                 unsigned dxmask = 0) {
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
      k_inv_x<T, N_, 1, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw, h->d_kx, dxmask);
    } else {
      dim3 grid(pairs / ZP, nj, nf);
      k_inv_x<T, N_, ZP, 256><<<grid, 256, 0, h->st>>>(s, d, h->g, j0, tw, h->d_kx, dxmask);
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

// threads per k_xprod block: 256 for the larger lines; where the LDS tile
// (> 80 KiB) leaves room for one block per CU only, one thread per column of
// the line (up to 1024), so that a CU still holds 8 or 16 waves
constexpr int xprod_nt(int nx, size_t lds) {
  return nx <= 32 ? 64 : (nx == 64 ? 128 : (lds > 81920 ? (nx >= 1024 ? 1024 : 512) : 256));
}

template <typename T, int NTH, bool LES = false>
int launch_xprod(dgpu_handle* h, XProdArgs& a, int nj) {
  a.tw = sizeof(T) == 8 ? h->twx64 : h->twx32;
  DG_SWITCH_N(h->g.nx, {
    // product lines + twiddle table
    constexpr size_t lds = sizeof(cx<T>) * ((8 + 3 * NTH) + 1) * N_;
    if constexpr (lds > 163840) {
      return fail(2, "this nx needs the mixed-precision mode (x pass exceeds LDS in FP64)");
    } else {
      constexpr int NT = xprod_nt(N_, lds);
      dim3 grid(h->g.nz / 2, nj, 1);
      k_xprod<T, N_, NT, NTH, LES><<<grid, NT, 0, h->st>>>(a);
    }
  });
  HIPCHK(hipGetLastError());
  return 0;
}

// The DNS x pass: one thread per column (64 for the short lines).  The planes are
// split into segments of as few planes as keep every block resident at once
// (a block's planes run one after another; a segment costs two halo planes of
// inverse transforms): one plane per block on small grids, 32 at 256x129x256.
template <typename T, int NTH>
int launch_xmarch(dgpu_handle* h, XProdArgs& a, int nj) {
  a.tw = sizeof(T) == 8 ? h->twx64 : h->twx32;
  DG_SWITCH_N(h->g.nx, {
    constexpr int NT = N_ < 64 ? 64 : N_;
    // NP product lines + NF field lines + the twiddle table, kx, output pointers
    constexpr size_t lds = sizeof(cx<T>) * ((11 + 4 * NTH) + 1) * N_ + sizeof(T) * (N_ / 2) +
                           sizeof(void*) * (6 + 2 * NTH);
    if constexpr (lds > 163840) {
      return fail(2, "this nx needs the mixed-precision mode (x pass exceeds LDS in FP64)");
    } else {
      if (h->xm_resident <= 0) {
        int per_cu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_xmarch<T, N_, NT, NTH>, NT, 0));
        h->xm_resident = std::max(1, per_cu) * h->cus;
      }
      const int pairs = h->g.nz / 2;
      const long long work = (long long)pairs * nj;
      const int seg = (int)std::min<long long>(nj, std::max<long long>(1, (work + h->xm_resident - 1) / h->xm_resident));
      const int nseg = (nj + seg - 1) / seg;
      dim3 grid(pairs, nseg, 1);
      k_xmarch<T, N_, NT, NTH><<<grid, NT, 0, h->st>>>(a, seg);
    }
  });
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_fwdz_nl(dgpu_handle* h, NlArgs& a, int nj) {
  a.tw = sizeof(T) == 8 ? h->twz64 : h->twz32;
  DG_SWITCH_N(h->g.nz, {
    constexpr int TI0 = ti_for(N_, sizeof(cx<T>));
    auto go = [&](auto TIc, auto Pc) {
      constexpr int TI = decltype(TIc)::value;
      constexpr bool PERSIST = decltype(Pc)::value;
      const int ntiles = (h->g.kxr + TI - 1) / TI * nj * (3 + h->nth);
      // persistent: as many blocks as are resident at once (one per CU for the FP64
      // 72 KiB tile at nz = 256: 0.126 -> 0.118 ms against two per CU)
      if (PERSIST && h->fz_resident <= 0) {
        int per_cu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwdz_nl<T, N_, TI, 256, PERSIST>, 256, 0));
        h->fz_resident = std::max(1, per_cu) * h->cus;
      }
      const int nblk = PERSIST ? std::min(ntiles, h->fz_resident) : ntiles;
      k_fwdz_nl<T, N_, TI, 256, PERSIST><<<nblk, 256, 0, h->st>>>(a, ntiles);
      return 0;
    };
    using I0 = std::integral_constant<int, TI0>;
    using I1 = std::integral_constant<int, (TI0 > 1 ? TI0 / 2 : 1)>;
    using P0 = std::false_type;
    using P1 = std::true_type;
    // measured: FP64 best persistent with TI columns; float with TI/2 per block at
    // nz = 256 (256^3), persistent with TI columns at nz = 512 (512x257x512)
    const int rc = (sizeof(T) == 8 || h->g.nz >= 512) ? go(I0{}, P1{}) : go(I1{}, P0{});
    if (rc) return rc;
  });
  HIPCHK(hipGetLastError());
  return 0;
}

// columns per block (TC) and rows per thread (RPT) of the column kernels: the
// largest TC whose LDS lets at least two blocks share a CU (<= 80 KiB)
struct ColShape {
  int tc, rpt;
  size_t lds;
};
ColShape col_shape(const dgpu_handle* h, bool cn) {
  const int ny2 = h->g.ny2, nf = 3 + h->nth;
  ColShape cs{};
  if (cn) {
    // k_cn: TC columns x all fields in LDS; one wave solves one field of the TC
    // columns in 64/TC chunks of L <= 32 rows (LMAX = 4 RPT >= L); the widest TC
    // that fits and still gives the CUs 1.5 blocks each, else the narrowest that fits
    // (measured on MI355X: at 64x65x64, 946 columns, 8 columns per block 0.0228 ms,
    // 4 0.0184, 2 0.0160, 1 0.0231; at 256x129x256, 8 0.196, 4 0.243)
    for (int tc : {8, 4, 2, 1}) {
      const size_t lds = (size_t)tc * ny2 * 16 * nf;
      const int rg = 256 / tc, need = (ny2 + rg - 1) / rg;
      // exact rows per thread for 8 columns (the chunk bound LMAX = 4 RPT then
      // tracks the column: fewer unrolled Thomas steps), powers of two otherwise
      int rpt = 1;
      while (rpt < need) rpt *= 2;
      if (tc == 8 && need >= 1 && need <= 8) rpt = need;
      if (lds <= 80 * 1024 && rpt <= 8 && h->g.pc % tc == 0) {
        cs = {tc, rpt, lds};
        if (2 * (h->g.pc / tc) >= 3 * h->cus) break;
      }
    }
    return cs;
  }
  // k_project: 4 columns per block where that still gives two blocks per CU, else
  // 2 (measured on MI355X: at 256x129x256, 4 columns 0.080 ms, 2 0.093, 8 0.106;
  // at 64x65x64, 946 columns, 2 columns 0.011 ms, 4 0.013)
  const long long want = h->resident_blocks > 0 ? h->resident_blocks : 512;
  const int tc0 = h->g.pc / 4 >= want ? 4 : 2;
  for (int tc : {tc0, 2}) {
    const size_t lds = (size_t)tc * ny2 * 40;
    if (lds <= 80 * 1024 || (tc == 2 && lds <= 160 * 1024)) {
      const int rg = 256 / tc, need = (ny2 + rg - 1) / rg;
      const int rpt = need <= 8 ? need : 16;
      cs = {tc, rpt, lds};
      break;
    }
  }
  return cs;
}

// rows per chunk of the k_cn solve (0 if no shape fits)
int cn_chunk(const dgpu_handle* h) {
  const ColShape cs = col_shape(h, true);
  if (cs.tc == 0) return 0;
  const int P = 64 / cs.tc;
  return (h->g.ny2 + P - 1) / P;
}

int grid_xcd(long long tiles) { return (int)((tiles + 7) / 8 * 8); }

template <typename T, int NTH, int TC, int RPT>
int launch_cn_t(dgpu_handle* h, ColArgs& a, size_t lds) {
  constexpr int LMAX = 4 * RPT;
  if (a.chunk > LMAX || a.chunk * (64 / TC) < h->g.ny2) return fail(2, "column chunking mismatch");
  HIPCHK(hipFuncSetAttribute((const void*)&k_cn<T, TC, 256, RPT, NTH, LMAX>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  k_cn<T, TC, 256, RPT, NTH, LMAX><<<grid_xcd(h->g.pc / TC), 256, lds, h->st>>>(a);
  HIPCHK(hipGetLastError());
  return 0;
}

template <int TC, int RPT>
int launch_project_t(dgpu_handle* h, ColArgs& a, size_t lds, int update_p) {
  HIPCHK(hipFuncSetAttribute((const void*)&k_project<TC, 256, RPT>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  k_project<TC, 256, RPT><<<grid_xcd(h->g.pc / TC), 256, lds, h->st>>>(a, update_p);
  HIPCHK(hipGetLastError());
  return 0;
}

// k_project shapes: every rows-per-thread count up to 8 (a power-of-two round-up
// would leave up to 3 of 8 row slots idle in every PCR level), then 16 for TC = 2
#define DG_SHAPES(X) X(8, 1) X(8, 2) X(8, 3) X(8, 4) X(8, 5) X(8, 6) X(8, 7) X(8, 8) \
                     X(4, 1) X(4, 2) X(4, 3) X(4, 4) X(4, 5) X(4, 6) X(4, 7) X(4, 8) \
                     X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(2, 6) X(2, 7) X(2, 8) X(2, 16)

#define DG_CN_SHAPES(X) X(8, 1) X(8, 2) X(8, 3) X(8, 4) X(8, 5) X(8, 6) X(8, 7) X(8, 8) \
                        X(4, 1) X(4, 2) X(4, 4) X(4, 8) \
                        X(2, 1) X(2, 2) X(2, 4) X(2, 8) X(1, 1) X(1, 2) X(1, 4) X(1, 8)

template <typename T, int NTH>
int launch_cn(dgpu_handle* h, ColArgs& a) {
  const ColShape cs = col_shape(h, true);
  if (cs.tc == 0 || h->g.pc % cs.tc) return fail(2, "ny too large for the column kernels");
#define DG_X(TC_, RPT_) \
  if (cs.tc == TC_ && cs.rpt == RPT_) return launch_cn_t<T, NTH, TC_, RPT_>(h, a, cs.lds);
  DG_CN_SHAPES(DG_X)
#undef DG_X
  return fail(2, "ny too large for the column kernels");
}

int launch_project(dgpu_handle* h, ColArgs& a, int update_p) {
  const ColShape cs = col_shape(h, false);
  if (cs.tc == 0 || h->g.pc % cs.tc) return fail(2, "ny too large for the column kernels");
#define DG_X(TC_, RPT_) \
  if (cs.tc == TC_ && cs.rpt == RPT_) return launch_project_t<TC_, RPT_>(h, a, cs.lds, update_p);
  DG_SHAPES(DG_X)
#undef DG_X
  return fail(2, "ny too large for the column kernels");
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
  if (c->use_les) {
    if (c->les_model_type != 1 && c->les_model_type != 4 && c->les_model_type != 5 && c->les_model_type != 6)
      return fail(2, "les_model_type must be 1, 4, 5 or 6 (learned); 2 and 3 stop in the reference");
    if (c->les_model_type == 6 && c->n_th != 1) return fail(2, "les_model_type 6 needs exactly one scalar");
    // the Fourier terms that rk_chan_1 forms before les_chan's BC posts differ
    // from those after them only on rows that the BC rows of Dirichlet v overwrite
    if (c->v_bc_ymin != DGPU_BC_DIRICHLET || c->v_bc_ymax != DGPU_BC_DIRICHLET)
      return fail(2, "the LES branch needs Dirichlet v at both walls");
  }
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
      cd* dst = compact + (size_t)j * g.pc + (size_t)k * g.kxp;
      std::memcpy(dst, src, sizeof(cd) * g.kxr);
      std::memset((void*)(dst + g.kxr), 0, sizeof(cd) * (g.kxp - g.kxr));
    }
}
void unpack_fourier(const dgpu_handle* h, const cd* compact, double* fort) {
  const Geo& g = h->g;
  const int ldi = g.nx / 2 + 1, ldk = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int k = 0; k < g.kzr; ++k) {
      double* dst = fort + 2 * (((size_t)j * ldk + k) * ldi);
      std::memcpy(dst, compact + (size_t)j * g.pc + (size_t)k * g.kxp, sizeof(cd) * g.kxr);
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

void fill_col_args(dgpu_handle* h, ColArgs& ca, double time) {
  const dgpu_config& c = h->c;
  ca.g = h->g;
  ca.plevels = pcr_levels(h->g.ny);
  ca.jstart = c.jstart; ca.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) {
    ca.jstart_th[n] = c.jstart_th[n]; ca.jend_th[n] = c.jend_th[n];
    ca.pr[n] = c.pr[n]; ca.dthdx[n] = c.dthdx[n]; ca.dthdz[n] = c.dthdz[n];
  }
  ca.front = c.flavor_front; ca.f_type = c.f_type;
  ca.nu = c.nu; ca.beta = c.beta; ca.nu_v_scale = c.nu_v_scale; ca.ro_inv = c.ro_inv; ca.delta = c.delta;
  ca.grav_x = c.grav_x; ca.grav_y = c.grav_y; ca.grav_z = c.grav_z; ca.ly = c.ly;
  ca.px0 = c.px0; ca.omega0 = c.omega0; ca.amp_omega0 = c.amp_omega0; ca.time = time;
  ca.temp4 = 1.0;
  ca.cori = c.ro_inv / c.delta;
  for (int n = 0; n < MAXTH; ++n) {
    ca.twu_x[n] = (c.ro_inv != 0.0 && n < h->nth) ? c.dthdx[n] * c.delta / c.ro_inv : 0.0;
    ca.twu_z[n] = (c.ro_inv != 0.0 && n < h->nth) ? -1.0 * c.dthdz[n] * c.delta / c.ro_inv : 0.0;
    ca.nupr[n] = n < h->nth ? c.nu / c.pr[n] : 0.0;
  }
  ca.rowc = h->d_rowc;
  ca.dy = h->d_dy; ca.dyf = h->d_dyf; ca.gy = h->d_gy; ca.gyf = h->d_gyf; ca.kx = h->d_kx; ca.kz = h->d_kz;
  solve_rows(h, time, ca.rows);
  for (int f = 0; f < 3 + h->nth; ++f) { ca.u[f] = h->U[f]; ca.cf[f] = h->CF[f]; ca.nl[f] = h->NL[f]; }
  ca.p = h->P;
}

double w_top_c1(const dgpu_config& c, double time) {
  // w_BC_Ymax_c1_transient (for/parameters.f90:279, for/solver.f90:2131)
  if (c.f_type == 4)
    return c.w_bc_ymax_c1 + c.amp_omega0 * std::sin(c.omega0 * (c.ro_inv / c.delta * time - c.force_start));
  return c.w_bc_ymax_c1;
}

// apply_BC_vel_mpi_post + apply_BC_th_mpi_post on the device state, and with
// les also apply_BC_les / apply_BC_th_les
int launch_bc_post(dgpu_handle* h, double time, int les) {
  const dgpu_config& c = h->c;
  const Geo& g = h->g;
  const int ny = g.ny;
  BcArgs ba{};
  ba.g = g; ba.nth = h->nth; ba.les = les;
  ba.u_lo = c.u_bc_ymin; ba.v_lo = c.v_bc_ymin; ba.w_lo = c.w_bc_ymin;
  ba.u_hi = c.u_bc_ymax; ba.v_hi = c.v_bc_ymax; ba.w_hi = c.w_bc_ymax;
  ba.u_lo_c1 = c.u_bc_ymin_c1; ba.v_lo_c1 = c.v_bc_ymin_c1; ba.w_lo_c1 = c.w_bc_ymin_c1;
  ba.u_hi_c1 = c.u_bc_ymax_c1; ba.v_hi_c1 = c.v_bc_ymax_c1; ba.w_hi_c1 = c.w_bc_ymax_c1;
  ba.w_hi_tr = w_top_c1(c, time);
  for (int n = 0; n < h->nth; ++n) {
    ba.th_lo[n] = c.th_bc_ymin[n]; ba.th_hi[n] = c.th_bc_ymax[n];
    ba.th_lo_c1[n] = c.th_bc_ymin_c1[n]; ba.th_hi_c1[n] = c.th_bc_ymax_c1[n];
  }
  ba.dy1 = h->dy[1]; ba.dy2 = h->dy[2]; ba.dyny = h->dy[ny]; ba.dyf1 = h->dyf[1];
  for (int f = 0; f < 3 + h->nth; ++f) ba.u[f] = h->U[f];
  k_bc_post<<<(unsigned)((g.pc + 255) / 256), 256, 0, h->st>>>(ba);
  HIPCHK(hipGetLastError());
  return 0;
}

// physical v, dv/dx, dv/dz of every transported field on all planes, in
// precision T: x-spectra into H, z-derivative x-spectra into HZ, fields into P
template <typename T>
int physical_gradients(dgpu_handle* h, void* const* H, void* const* HZ, void* const* P) {
  const Geo& g = h->g;
  const int nf = 3 + h->nth;
  const cd* zs[2 * MAXF];
  void* zd[2 * MAXF];
  for (int f = 0; f < nf; ++f) {
    zs[f] = h->U[f]; zd[f] = H[f];
    zs[nf + f] = h->U[f]; zd[nf + f] = HZ[f];
  }
  int rc;
  if ((rc = launch_inv_z<T>(h, 2 * nf, zs, zd, 0, g.ny2, ((1u << nf) - 1u) << nf))) return rc;
  void* xs[3 * MAXF];
  T* xd[3 * MAXF];
  unsigned dxmask = 0;
  for (int f = 0; f < nf; ++f) {
    xs[3 * f] = H[f]; xd[3 * f] = (T*)P[3 * f];
    xs[3 * f + 1] = H[f]; xd[3 * f + 1] = (T*)P[3 * f + 1];
    xs[3 * f + 2] = HZ[f]; xd[3 * f + 2] = (T*)P[3 * f + 2];
    dxmask |= 1u << (3 * f + 1);
  }
  return launch_inv_x<T>(h, 3 * nf, xs, xd, 0, g.ny2, dxmask);
}

// the FP64 gradient buffers (aliases on an FP64 handle that has the LES buffers)
int ensure_grad_buffers(dgpu_handle* h) {
  const Geo& g = h->g;
  const int nf = 3 + h->nth;
  const size_t rbytes = sizeof(double) * g.pp * g.ny2;
  const size_t hbytes = sizeof(cd) * g.ph * g.ny2;
  const bool fp64 = h->prec == DGPU_FP64;
  void* p = nullptr;
  int rc;
  for (int f = 0; f < nf; ++f) {
    if (fp64) h->HinD[f] = h->Hin[f];
    else if (!h->HinD[f]) { if ((rc = dalloc(h, &p, hbytes))) return rc; h->HinD[f] = p; }
    if (fp64 && h->HinZ[f]) h->HinZD[f] = h->HinZ[f];
    if (!h->HinZD[f]) { if ((rc = dalloc(h, &p, hbytes))) return rc; h->HinZD[f] = p; }
    for (int m = 0; m < 3; ++m) {
      if (fp64 && h->Pg[3 * f + m]) h->PgD[3 * f + m] = h->Pg[3 * f + m];
      if (!h->PgD[3 * f + m]) { if ((rc = dalloc(h, &p, rbytes))) return rc; h->PgD[3 * f + m] = p; }
    }
  }
  return 0;
}

constexpr int COURANT_BLOCKS = 1024;

// courant's CFL minimum of the physical velocity u1, u2, u3 (FP64 [j][z][x]);
// the block minima stay on the device until dgpu_courant
int launch_courant(dgpu_handle* h, const double* u1, const double* u2, const double* u3) {
  const dgpu_config& c = h->c;
  if (!h->d_cfl) {
    void* p = nullptr;
    int rc = dalloc(h, &p, sizeof(double) * COURANT_BLOCKS);
    if (rc) return rc;
    h->d_cfl = (double*)p;
  }
  CourantArgs a{};
  a.g = h->g; a.nth = h->nth; a.front = c.flavor_front;
  a.jlo = c.flavor_front ? c.jstart : 1;
  a.jhi = c.flavor_front ? c.jend : c.ny;
  a.cfl = h->cfl; a.dx = c.lx / c.nx; a.dz = c.lz / c.nz; a.ly = c.ly;
  a.f = c.ro_inv / c.delta;
  for (int n = 0; n < MAXTH; ++n) { a.dthdx[n] = c.dthdx[n]; a.dthdz[n] = c.dthdz[n]; }
  a.dy = h->d_dy; a.gyf = h->d_gyf; a.u1 = u1; a.u2 = u2; a.u3 = u3; a.part = h->d_cfl;
  k_courant<<<COURANT_BLOCKS, 256, 0, h->st>>>(a);
  HIPCHK(hipGetLastError());
  h->cfl_ran = 1;
  return 0;
}

// the DNS path: the CN solution (Fourier rows of the state, before the
// projection) to physical space, then the reduction
int courant_dns(dgpu_handle* h) {
  int rc = ensure_grad_buffers(h);
  if (rc) return rc;
  const Geo& g = h->g;
  const cd* cs[3] = {h->U[0], h->U[1], h->U[2]};
  void* hs[3] = {h->HinD[0], h->HinD[1], h->HinD[2]};
  double* rs[3] = {(double*)h->PgD[0], (double*)h->PgD[3], (double*)h->PgD[6]};
  if ((rc = launch_inv_z<double>(h, 3, cs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_inv_x<double>(h, 3, hs, rs, 0, g.ny2))) return rc;
  return launch_courant(h, rs[0], rs[1], rs[2]);
}

// One substep of the LES branch: rk_chan_1 with les_chan / les_chan_th
// (for/solver.f90:611-623) and the physical-space Crank–Nicolson solves with the
// eddy viscosity (for/solver.f90:875-903, 990-1233).
// k_closure's per-launch constants
void closure_consts(dgpu_handle* h, ClosureArgs& k) {
  k.nu = h->c.nu;
  k.gyb = h->c.grav_y;
  k.md2 = -1.0 / (2.0 * C_AMD * C_AMD);
  k.rnu2 = (float)(std::sqrt(2.0) / h->c.nu);
  k.rdxs = (float)(1.0 / k.a.dxs);
  k.rdzs = (float)(1.0 / k.a.dzs);
}

// arguments of k_closure from the handle's network and an AmdArgs already filled
ClosureArgs closure_args(dgpu_handle* h, const AmdArgs& aa) {
  ClosureArgs k{};
  k.a = aa;
  MlpW& w = k.w;
  w.w1t = h->d_netw;
  w.w2t = h->d_netw + 64 * 16;
  w.w3t = h->d_netw + 64 * 16 + 64 * 64;
  w.w2 = nullptr;
  w.w3 = nullptr;
  w.b1 = h->d_netp + MLP_OB1;
  w.b2 = h->d_netp + MLP_OB2;
  w.b3 = h->d_netp + MLP_OB3;
  w.norm = h->d_netnorm;
  k.have_net = h->have_net;
  closure_consts(h, k);
  k.cgy = h->Cgy; k.kgy = h->Kgy;
  return k;
}

// the AmdArgs of the handle's current gradient fields (outputs left unset)
AmdArgs amd_args(dgpu_handle* h) {
  const dgpu_config& c = h->c;
  AmdArgs aa{};
  aa.g = h->g; aa.nth = h->nth; aa.jstart = c.jstart; aa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { aa.jstart_th[n] = c.jstart_th[n]; aa.jend_th[n] = c.jend_th[n]; }
  aa.dxs = c.lx / c.nx * BETA_SGS; aa.dzs = c.lz / c.nz * BETA_SGS;
  aa.zx = aa.dzs / aa.dxs; aa.xz = aa.dxs / aa.dzs;
  aa.dy = h->d_dy; aa.dyf = h->d_dyf; aa.dlt = h->d_dlt; aa.rowc = h->d_rowc;
  for (int s = 0; s < 3 * (3 + h->nth); ++s) aa.f[s] = h->Pg[s];
  return aa;
}

template <typename T, int NTH>
int substep_les(dgpu_handle* h, int rk, double delta_t, double time, int time_step) {
  const dgpu_config& c = h->c;
  int rc = ensure_factors(h, delta_t);
  if (rc) return rc;
  constexpr bool mixed = sizeof(T) == 4;
  const double hb[3] = {delta_t * (8.0 / 15.0), delta_t * (2.0 / 15.0), delta_t * (5.0 / 15.0)};
  const double bb[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0};
  const double zb[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0};
  const int r = rk - 1;
  const double temp1 = 0.5 * c.nu * hb[r], temp2 = 0.5 * hb[r], temp3 = zb[r] * hb[r], temp4 = hb[r],
               temp5 = bb[r] * hb[r];
  const Geo& g = h->g;
  constexpr int nf = 3 + NTH;
  const int ny = g.ny;
  int jlo = c.jstart < 2 ? c.jstart : 2;
  for (int n = 0; n < NTH; ++n) jlo = c.jstart_th[n] < jlo ? c.jstart_th[n] : jlo;
  const int nj = ny - jlo + 1;
  const double wtop = w_top_c1(c, time);

  // 0. BC posts of les_chan on the state's ghost planes
  if ((rc = tbeg(h, 5))) return rc;
  if ((rc = launch_bc_post(h, time, 1))) return rc;
  if ((rc = tend(h))) return rc;

  // 1. x-spectra of v and dv/dz, then physical v, dv/dx, dv/dz on all planes (T);
  //    a mixed handle also needs v in FP64 for the CN stage
  if ((rc = tbeg(h, 6))) return rc;
  if ((rc = physical_gradients<T>(h, h->Hin, h->HinZ, h->Pg))) return rc;
  if ((rc = tend(h))) return rc;

  // 2. AMD: nu_t, kappa_t and the SGS additions to the products
  AmdArgs aa{};
  aa.g = g; aa.nth = NTH; aa.jstart = c.jstart; aa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { aa.jstart_th[n] = c.jstart_th[n]; aa.jend_th[n] = c.jend_th[n]; }
  aa.dxs = c.lx / c.nx * BETA_SGS; aa.dzs = c.lz / c.nz * BETA_SGS;
  aa.zx = aa.dzs / aa.dxs; aa.xz = aa.dxs / aa.dzs;
  aa.dy = h->d_dy; aa.dyf = h->d_dyf; aa.dlt = h->d_dlt; aa.rowc = h->d_rowc;
  for (int s = 0; s < 3 * nf; ++s) aa.f[s] = h->Pg[s];
  aa.nut = h->NuT;
  for (int n = 0; n < NTH; ++n) aa.kap[n] = h->KapT[n];
  for (int m = 0; m < 7 + 2 * NTH; ++m) aa.aug[m] = h->Aug[m];
  // one block per (256 points, row): xcd_rows maps the 1D index
  const dim3 agrid((unsigned)(((g.pp + 255) / 256) * g.ny2));
  if constexpr (NTH == 1) {
    if (c.les_model_type == 6) {
      // the learned coefficients (AMD where a point leaves the envelope)
      ClosureArgs ka = closure_args(h, aa);
      if ((rc = tbeg(h, 10))) return rc;
      k_closure<T, NTH><<<agrid, 256, 0, h->st>>>(ka);
      HIPCHK(hipGetLastError());
      if ((rc = tend(h))) return rc;
      aa.cgy = h->Cgy; aa.kgy = h->Kgy;
    }
  }
  if ((rc = tbeg(h, 7))) return rc;
  if (c.les_model_type == 1) k_amd<T, NTH, 1><<<agrid, 256, 0, h->st>>>(aa);
  else if (c.les_model_type == 4) k_amd<T, NTH, 4><<<agrid, 256, 0, h->st>>>(aa);
  else if (c.les_model_type == 6 && NTH == 1) k_amd<T, NTH, 6><<<agrid, 256, 0, h->st>>>(aa);
  else k_amd<T, NTH, 5><<<agrid, 256, 0, h->st>>>(aa);
  HIPCHK(hipGetLastError());
  if ((rc = tend(h))) return rc;

  // 3. fused x pass: products + SGS stresses; 4. z forward -> nonlinear tendencies
  XProdArgs xa{};
  xa.g = g; xa.jlo = jlo; xa.jstart = c.jstart; xa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { xa.jstart_th[n] = c.jstart_th[n]; xa.jend_th[n] = c.jend_th[n]; }
  xa.dy = h->d_dy; xa.dyf = h->d_dyf; xa.kx = h->d_kx; xa.rowc = h->d_rowc;
  for (int f = 0; f < nf; ++f) xa.hin[f] = h->Hin[f];
  for (int o = 0; o < 6 + 2 * NTH; ++o) xa.hout[o] = h->Hout[o];
  for (int m = 0; m < 7 + 2 * NTH; ++m) xa.aug[m] = h->Aug[m];
  for (int f = 0; f < nf; ++f) {
    xa.phys[f] = h->Pg[3 * f];  // the gradient stage's physical fields
    if (!xa.phys[f]) return fail(12, "x pass: physical field buffer missing");
  }
  if ((rc = tbeg(h, 1))) return rc;
  if ((rc = launch_xprod<T, NTH, true>(h, xa, nj))) return rc;
  if ((rc = tend(h))) return rc;
  NlArgs na{};
  na.g = g; na.jlo = jlo; na.jstart = c.jstart; na.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { na.jstart_th[n] = c.jstart_th[n]; na.jend_th[n] = c.jend_th[n]; }
  na.kz = h->d_kz;
  for (int o = 0; o < 6 + 2 * NTH; ++o) na.h[o] = h->Hout[o];
  for (int f = 0; f < nf; ++f) na.nl[f] = h->NL[f];
  if ((rc = tbeg(h, 2))) return rc;
  if ((rc = launch_fwdz_nl<T>(h, na, nj))) return rc;
  if ((rc = tend(h))) return rc;

  // 5. Fourier assembly of cf and of the right-hand side (molecular explicit half included)
  ColArgs ca{};
  fill_col_args(h, ca, time);
  ca.rk = rk; ca.do_cn = 0;
  ca.temp1 = temp1; ca.temp3 = temp3; ca.temp4 = temp4; ca.temp5 = temp5;
  ca.t1v = temp1 * c.nu_v_scale;
  for (int n = 0; n < NTH; ++n) ca.t1pr[n] = temp1 / c.pr[n] * c.nu_v_scale;
  ca.thc = h->d_fac + (size_t)r * MAXF * 5 * g.ny2;
  ca.chunk = cn_chunk(h);
  // the mixed handle solves for the increment of the state in float transforms,
  // except where courant needs the new velocity in FP64 (rk 3 of an update step)
  const bool courant_now = h->variable_dt && rk == 3 && time_step % h->update_dt == 0;
  const bool inc = mixed && !courant_now;
  ca.cr_minus_u = inc ? 1 : 0;
  for (int f = 0; f < nf; ++f) ca.crout[f] = h->CR[f];
  if ((rc = tbeg(h, 3))) return rc;
  k_assemble<T, NTH><<<dim3((unsigned)((g.pc + 255) / 256), (unsigned)g.ny2), 256, 0, h->st>>>(ca);
  HIPCHK(hipGetLastError());
  if ((rc = tend(h))) return rc;

  // 6. right-hand sides to physical space: FP64, and on a mixed handle without
  //    increments also the FP64 state the solve starts from; with increments the
  //    float transform of r - u (the float state is the gradient stage's)
  const cd* cs[MAXF];
  double* rs[MAXF];
  for (int f = 0; f < nf; ++f) { cs[f] = h->CR[f]; rs[f] = h->Rp[f]; }
  double* ucn[MAXF];
  for (int f = 0; f < nf; ++f) ucn[f] = (double*)h->PgD[3 * f];
  float* rinc[MAXF];
  for (int f = 0; f < nf; ++f) rinc[f] = (float*)h->Rp[f];
  if ((rc = tbeg(h, 9))) return rc;
  if (inc) {
    if ((rc = launch_inv_z<float>(h, nf, cs, h->Hin, 0, g.ny2))) return rc;
    if ((rc = launch_inv_x<float>(h, nf, h->Hin, rinc, 0, g.ny2))) return rc;
  } else {
    if ((rc = launch_inv_z<double>(h, nf, cs, h->HinD, 0, g.ny2))) return rc;
    if ((rc = launch_inv_x<double>(h, nf, h->HinD, rs, 0, g.ny2))) return rc;
    if (mixed) {
      const cd* us0[MAXF];
      for (int f = 0; f < nf; ++f) us0[f] = h->U[f];
      if ((rc = launch_inv_z<double>(h, nf, us0, h->HinD, 0, g.ny2))) return rc;
      if ((rc = launch_inv_x<double>(h, nf, h->HinD, ucn, 0, g.ny2))) return rc;
    }
  }
  if ((rc = tend(h))) return rc;

  // 7. per-column CN with nu_t / kappa_t (the field systems are independent; the
  //    reference solves the scalars first, for/solver.f90:1004-1007)
  PhysCnArgs<double, T> pa{};
  pa.g = g; pa.nf = nf; pa.temp2 = temp2; pa.dy = h->d_dy; pa.dyf = h->d_dyf; pa.rowc = h->d_rowc; pa.cp = h->CPs;
  for (int e0 = 0; e0 < nf; ++e0) {
    const int e = e0 < NTH ? 3 + e0 : e0 - NTH;  // slot e0 solves field e
    const SolveRows& R = ca.rows[e];
    pa.blo[e0] = R.blo; pa.bhi[e0] = R.bhi; pa.wlo[e0] = R.wlo; pa.whi[e0] = R.whi;
    pa.u[e0] = ucn[e];
    pa.r[e0] = h->Rp[e];
    pa.visc[e0] = (const T*)(e < 3 ? h->NuT : h->KapT[e - 3]);
    pa.field[e0] = e;
    if (e == 0) { pa.lo_kind[e0] = c.u_bc_ymin; pa.hi_kind[e0] = c.u_bc_ymax; pa.lo_c1[e0] = c.u_bc_ymin_c1; pa.hi_c1[e0] = c.u_bc_ymax_c1; }
    else if (e == 1) { pa.lo_kind[e0] = c.v_bc_ymin; pa.hi_kind[e0] = c.v_bc_ymax; pa.lo_c1[e0] = c.v_bc_ymin_c1; pa.hi_c1[e0] = c.v_bc_ymax_c1; }
    else if (e == 2) { pa.lo_kind[e0] = c.w_bc_ymin; pa.hi_kind[e0] = c.w_bc_ymax; pa.lo_c1[e0] = c.w_bc_ymin_c1; pa.hi_c1[e0] = wtop; }
    else {
      const int n = e - 3;
      pa.lo_kind[e0] = c.th_bc_ymin[n]; pa.hi_kind[e0] = c.th_bc_ymax[n];
      pa.lo_c1[e0] = c.th_bc_ymin_c1[n]; pa.hi_c1[e0] = c.th_bc_ymax_c1[n];
    }
    pa.tcn[e0] = e < 3 ? temp1 * c.nu_v_scale : temp1 / c.pr[e - 3] * c.nu_v_scale;
  }
  if ((rc = tbeg(h, 8))) return rc;
  const dim3 pgrid((unsigned)((g.pp + 255) / 256), nf);
  if (inc) {
    PhysCnArgs<float, T> pf{};
    pf.g = pa.g; pf.nf = pa.nf; pf.temp2 = pa.temp2; pf.dy = pa.dy; pf.dyf = pa.dyf; pf.rowc = pa.rowc; pf.cp = pa.cp;
    for (int e0 = 0; e0 < nf; ++e0) {
      pf.field[e0] = pa.field[e0]; pf.blo[e0] = pa.blo[e0]; pf.bhi[e0] = pa.bhi[e0];
      pf.wlo[e0] = pa.wlo[e0]; pf.whi[e0] = pa.whi[e0]; pf.lo_kind[e0] = pa.lo_kind[e0];
      pf.hi_kind[e0] = pa.hi_kind[e0]; pf.lo_c1[e0] = pa.lo_c1[e0]; pf.hi_c1[e0] = pa.hi_c1[e0];
      pf.tcn[e0] = pa.tcn[e0]; pf.visc[e0] = pa.visc[e0];
      // slot e0 holds field pa.field[e0]
      pf.u[e0] = (float*)h->Pg[3 * pa.field[e0]];
      pf.r[e0] = rinc[pa.field[e0]];
    }
    k_cn_phys<float, T, true><<<pgrid, 256, 0, h->st>>>(pf);
  } else {
    k_cn_phys<double, T, false><<<pgrid, 256, 0, h->st>>>(pa);
  }
  HIPCHK(hipGetLastError());
  if ((rc = tend(h))) return rc;
  // courant on the new velocity, before the projection (for/solver.f90:1241-1244)
  if (courant_now)
    if ((rc = launch_courant(h, ucn[0], ucn[1], ucn[2]))) return rc;

  // 8. solved fields (or increments) back to Fourier space on every plane, as the
  //    reference, then rem_div_chan and the pressure update
  double* us[MAXF];
  cd* ud[MAXF];
  for (int f = 0; f < nf; ++f) { us[f] = ucn[f]; ud[f] = h->U[f]; }
  if ((rc = tbeg(h, 9))) return rc;
  if (inc) {
    if ((rc = launch_fwd_x<float>(h, nf, rinc, h->Hin, 0, g.ny2))) return rc;
    if ((rc = launch_fwd_z<float, true>(h, nf, h->Hin, ud, 0, g.ny2))) return rc;
  } else {
    if ((rc = launch_fwd_x<double>(h, nf, us, h->HinD, 0, g.ny2))) return rc;
    if ((rc = launch_fwd_z<double>(h, nf, h->HinD, ud, 0, g.ny2))) return rc;
  }
  if ((rc = tend(h))) return rc;
  if ((rc = tbeg(h, 4))) return rc;
  if ((rc = launch_project(h, ca, 1))) return rc;
  return tend(h);
}

int substep(dgpu_handle* h, int rk, double delta_t, double time, int time_step, int create_new_flow) {
  const dgpu_config& c = h->c;
  if (rk < 1 || rk > 3) return fail(3, "rk_step must be 1, 2 or 3");
  if (c.use_les && (!create_new_flow || time_step > 100))
    return by_nth(h->nth, [&](auto NTHc) {
      constexpr int NTH = decltype(NTHc)::value;
      return h->prec == DGPU_MIXED ? substep_les<float, NTH>(h, rk, delta_t, time, time_step)
                                   : substep_les<double, NTH>(h, rk, delta_t, time, time_step);
    });
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
  // 2. x pass marching in y: x inverse transforms, products, x forward transforms
  XProdArgs xa{};
  xa.g = g; xa.jlo = jlo; xa.jstart = c.jstart; xa.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { xa.jstart_th[n] = c.jstart_th[n]; xa.jend_th[n] = c.jend_th[n]; }
  xa.dy = h->d_dy; xa.dyf = h->d_dyf; xa.kx = h->d_kx; xa.rowc = h->d_rowc;
  for (int f = 0; f < nf; ++f) xa.hin[f] = h->Hin[f];
  for (int o = 0; o < 6 + 2 * h->nth; ++o) xa.hout[o] = h->Hout[o];
  if ((rc = tbeg(h, 11))) return rc;
  rc = by_nth(h->nth, [&](auto NTHc) {
    constexpr int NTH = decltype(NTHc)::value;
    return mixed ? launch_xmarch<float, NTH>(h, xa, nj) : launch_xmarch<double, NTH>(h, xa, nj);
  });
  if (rc || (rc = tend(h))) return rc;
  // 3. z forward transform of the products -> nonlinear tendencies
  NlArgs na{};
  na.g = g; na.jlo = jlo; na.jstart = c.jstart; na.jend = c.jend;
  for (int n = 0; n < MAXTH; ++n) { na.jstart_th[n] = c.jstart_th[n]; na.jend_th[n] = c.jend_th[n]; }
  na.kz = h->d_kz;
  for (int o = 0; o < 6 + 2 * h->nth; ++o) na.h[o] = h->Hout[o];
  for (int f = 0; f < nf; ++f) na.nl[f] = h->NL[f];
  if ((rc = tbeg(h, 2))) return rc;
  rc = mixed ? launch_fwdz_nl<float>(h, na, nj) : launch_fwdz_nl<double>(h, na, nj);
  if (rc || (rc = tend(h))) return rc;
  // 4. columns: assembly, CN, projection, pressure
  ColArgs ca{};
  fill_col_args(h, ca, time);
  ca.rk = rk; ca.do_cn = 1;
  ca.temp1 = temp1; ca.temp3 = temp3; ca.temp4 = temp4; ca.temp5 = temp5;
  ca.t1v = temp1 * c.nu_v_scale;
  for (int n = 0; n < h->nth; ++n) ca.t1pr[n] = temp1 / c.pr[n] * c.nu_v_scale;
  ca.thc = h->d_fac + (size_t)r * MAXF * 5 * g.ny2;
  ca.chunk = cn_chunk(h);
  if ((rc = tbeg(h, 3))) return rc;
  rc = by_nth(h->nth, [&](auto NTHc) {
    constexpr int NTH = decltype(NTHc)::value;
    return mixed ? launch_cn<float, NTH>(h, ca) : launch_cn<double, NTH>(h, ca);
  });
  if (rc || (rc = tend(h))) return rc;
  // courant on the new velocity, before the projection (for/solver.f90:1241-1244)
  if (h->variable_dt && rk == 3 && time_step % h->update_dt == 0)
    if ((rc = courant_dns(h))) return rc;
  if ((rc = tbeg(h, 4))) return rc;
  if ((rc = launch_project(h, ca, 1))) return rc;
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
  g.kxp = (g.kxr + 7) / 8 * 8;
  g.pc = (long long)g.kxp * g.kzr;
  g.ph = (long long)g.nz * g.kxp;
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
  const size_t nlbytes = hsz * g.pc * n2;
  for (int f = 0; f < nf; ++f) { ALLOC(h->U[f], cbytes); ALLOC(h->CF[f], cbytes); ALLOC(h->NL[f], nlbytes); }
  ALLOC(h->P, cbytes);
  for (int f = 0; f < nf; ++f) ALLOC(h->Hin[f], hbytes);
  for (int o = 0; o < no; ++o) ALLOC(h->Hout[o], hbytes);
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return bail(fail(10, "hipGetDevice"));
    h->dev = dev;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      h->resident_blocks = 2 * ncu, h->cus = ncu;
    if ((rc = scratch_address(dev, &h->noalloc))) return bail(rc);
  }
  ALLOC(h->twx64, sizeof(cd) * g.nx); ALLOC(h->twz64, sizeof(cd) * g.nz);
  ALLOC(h->twx32, sizeof(cf) * g.nx); ALLOC(h->twz32, sizeof(cf) * g.nz);
  ALLOC(h->d_fac, sizeof(double) * 3 * MAXF * 5 * n2);
  ALLOC(h->d_rowc, sizeof(double) * RC_COUNT * n2);
  if (cfg->use_les) {
    // gradients, eddy viscosity and SGS fields in the transform precision; the
    // right-hand sides and the solve scratch in FP64
    const size_t rbytes = sizeof(double) * g.pp * n2;
    const size_t tbytes = (h->prec == DGPU_MIXED ? sizeof(float) : sizeof(double)) * g.pp * n2;
    for (int f = 0; f < nf; ++f) {
      ALLOC(h->HinZ[f], hbytes);
      ALLOC(h->CR[f], cbytes);
      ALLOC(h->Rp[f], rbytes);
      for (int m = 0; m < 3; ++m) ALLOC(h->Pg[3 * f + m], tbytes);
    }
    ALLOC(h->NuT, tbytes);
    for (int n = 0; n < h->nth; ++n) ALLOC(h->KapT[n], tbytes);
    for (int m = 0; m < 7 + 2 * h->nth; ++m) ALLOC(h->Aug[m], tbytes);
    ALLOC(h->CPs, 2 * (size_t)nf * rbytes);
    ALLOC(h->d_dlt, sizeof(double) * 2 * n2);
    if (cfg->les_model_type == 6) {
      ALLOC(h->Cgy, tbytes);
      ALLOC(h->Kgy, tbytes);
      ALLOC(h->d_ccount, sizeof(unsigned long long));
      ALLOC(h->d_netw, sizeof(unsigned short) * (64 * 16 + 64 * 64 + 16 * 64));
      ALLOC(h->d_netp, sizeof(float) * MLP_NP);
      ALLOC(h->d_netnorm, sizeof(float) * MLP_NNORM);
    }
    if ((rc = ensure_grad_buffers(h))) return bail(rc);
  }
#undef ALLOC
  h->fort_doubles = fortran_doubles(cfg);
  const size_t stage_doubles = std::max<size_t>(h->fort_doubles, (size_t)2 * g.pc * n2);
  if (hipHostMalloc((void**)&h->hstage, sizeof(double) * stage_doubles, hipHostMallocDefault) != hipSuccess)
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
  std::vector<double> rowc((size_t)RC_COUNT * n2, 0.0);
  for (int j = 1; j <= cfg->ny; ++j) {
    double* r = &rowc[(size_t)j * RC_COUNT];
    const double dy = h->dy[j], dyp = h->dy[j + 1], dyf = h->dyf[j], dyfm = h->dyf[j - 1];
    r[RC_RDY] = 1.0 / dy;
    r[RC_RDYP] = 1.0 / dyp;
    r[RC_RDYF] = 1.0 / dyf;
    r[RC_RDYFM] = 1.0 / dyfm;
    r[RC_WTH] = dyfm / (2.0 * dy);
    r[RC_WTHM] = dyf / (2.0 * dy);
    r[RC_FLO] = 1.0 / (dy * dyf);
    r[RC_FHI] = 1.0 / (dyp * dyf);
    r[RC_GLO] = 1.0 / (dyfm * dy);
    r[RC_GHI] = 1.0 / (dyf * dy);
    if (cfg->ro_inv != 0.0)
      for (int n = 0; n < cfg->n_th; ++n) {
        const double yf = h->gyf[j] - 0.5 * cfg->ly, yg = h->gy[j] - 0.5 * cfg->ly;
        r[RC_TW0 + 4 * n + 0] = cfg->dthdx[n] * yf * cfg->delta / cfg->ro_inv;
        r[RC_TW0 + 4 * n + 1] = -1.0 * cfg->dthdz[n] * yf * cfg->delta / cfg->ro_inv;
        r[RC_TW0 + 4 * n + 2] = cfg->dthdx[n] * yg * cfg->delta / cfg->ro_inv;
        r[RC_TW0 + 4 * n + 3] = -1.0 * cfg->dthdz[n] * yg * cfg->delta / cfg->ro_inv;
      }
  }
  // the quotients the LES kernels take on the ghost rows too (dyf > 0 on rows
  // 0..Nyp+1, dy > 0 on 1..Nyp+1), and the AMD closure's width ratios
  rowc[RC_RDYF] = 1.0 / h->dyf[0];
  rowc[(size_t)(cfg->ny + 1) * RC_COUNT + RC_RDY] = 1.0 / h->dy[cfg->ny + 1];
  rowc[(size_t)(cfg->ny + 1) * RC_COUNT + RC_RDYF] = 1.0 / h->dyf[cfg->ny + 1];
  if (cfg->use_les) {
    const double ax = cfg->lx / cfg->nx * BETA_SGS, az = cfg->lz / cfg->nz * BETA_SGS;
    for (int j = 1; j <= cfg->ny + 1; ++j) {
      double* r = &rowc[(size_t)j * RC_COUNT];
      const double dyl = h->dy[j] * 2.0;
      r[RC_YX] = dyl / ax; r[RC_XY] = ax / dyl; r[RC_YZ] = dyl / az; r[RC_ZY] = az / dyl;
    }
  }
  ok = ok && up(h->d_rowc, rowc.data(), sizeof(double) * rowc.size());
  std::vector<double> dlt(2 * (size_t)n2, 0.0);
  if (cfg->use_les) {
    // the filter-width factors delta_yf(j) (GYF), delta_y(j) (GY) of les_chan
    const double ax = cfg->lx / cfg->nx * BETA_SGS, az = cfg->lz / cfg->nz * BETA_SGS;
    for (int j = 1; j <= cfg->ny + 1; ++j)
      for (int w = 0; w < 2; ++w) {
        double v;
        if (cfg->les_model_type == 1) {
          // -2 c_smag^2 (dx beta h dz beta)^(2/3), h = dyf(j) * 1 or dy(j) * 2 (for/les.f90:161-170)
          const double hh = (w == 0) ? h->dyf[j] * 1.0 : h->dy[j] * 2.0;
          v = -(2.0 * (C_SMAG * C_SMAG) * std::pow(ax * hh * az, 2.0 / 3.0));
        } else if (cfg->les_model_type == 4) {
          v = -(2.0 * (C_AMD * C_AMD));  // for/les.f90:464-469
        } else {
          // -2 C_AMD^2 * 3 / (1/(dx beta)^2 + 1/(2 h)^2 + 1/(dz beta)^2) (for/les.f90:756-780)
          const double b = (w == 0 ? h->dyf[j] : h->dy[j]) * 2.0;
          v = -(2.0 * (C_AMD * C_AMD) * 3.0 / (1.0 / (ax * ax) + 1.0 / (b * b) + 1.0 / (az * az)));
        }
        dlt[(size_t)w * n2 + j] = v;
      }
    ok = ok && up(h->d_dlt, dlt.data(), sizeof(double) * dlt.size());
  }
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
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->d_box) (void)hipFree(h->d_box);
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

int dgpu_download_les(dgpu_handle* h, double* nu_t, double* kappa_t) {
  if (!h || !nu_t) return fail(1, "null argument");
  const Geo& g = h->g;
  const size_t fd = fortran_doubles(&h->c);
  const int nfld = 1 + h->nth;
  // nu_t and kappa_t are held in the handle's working precision
  const bool f32 = h->prec == DGPU_MIXED;
  std::vector<double> tmp((size_t)g.pp * g.ny2);
  std::vector<float> tmpf(f32 ? tmp.size() : 0);
  for (int m = 0; m < nfld; ++m) {
    double* out = m == 0 ? nu_t : (kappa_t ? kappa_t + (size_t)(m - 1) * fd : nullptr);
    if (!out) return fail(1, "null kappa_t");
    const void* src = m == 0 ? h->NuT : h->KapT[m - 1];
    if (src && f32) {
      HIPCHK(hipMemcpyAsync(tmpf.data(), src, tmpf.size() * 4, hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      std::copy(tmpf.begin(), tmpf.end(), tmp.begin());
    } else if (src) {
      HIPCHK(hipMemcpyAsync(tmp.data(), src, tmp.size() * 8, hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
    } else {
      std::fill(tmp.begin(), tmp.end(), 0.0);
    }
    std::fill(out, out + fd, 0.0);
    for (int j = 0; j < g.ny2; ++j)
      for (int z = 0; z < g.nz; ++z)
        std::memcpy(out + ((size_t)j * (g.nz + 2) + z) * (g.nx + 2), &tmp[((size_t)j * g.nz + z) * g.nx],
                    sizeof(double) * g.nx);
  }
  return 0;
}

int dgpu_set_closure(dgpu_handle* h, const float* params, const float* norm) {
  if (!h || !params || !norm) return fail(1, "null argument");
  if (!h->c.use_les || h->c.les_model_type != 6) return fail(2, "dgpu_set_closure needs an LES handle of type 6");
  if (h->gexec) {  // the closure's branch changes the captured launches
    HIPCHK(hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  // BF16 copies of the forward weights: w1t [64][16], w2t [64][64], w3t [16][64]
  // (round to nearest even, as the device's own conversion)
  auto bf = [](float x) -> unsigned short {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  };
  std::vector<float> p(params, params + MLP_NP);
  for (int i = 0; i < MLP_NP; ++i) {
    const bool dead = (i < MLP_OB1 && (i - MLP_OW1) / MLP_H >= MLP_IN) ||
                      (i >= MLP_OW3 && i < MLP_OB3 && (i - MLP_OW3) % MLP_OUTP >= MLP_OUT) ||
                      (i >= MLP_OB3 && i - MLP_OB3 >= MLP_OUT);
    if (dead) p[i] = 0.f;
  }
  std::vector<unsigned short> w(64 * 16 + 64 * 64 + 16 * 64);
  for (int in = 0; in < MLP_INP; ++in)
    for (int out = 0; out < MLP_H; ++out) w[out * MLP_INP + in] = bf(p[MLP_OW1 + in * MLP_H + out]);
  for (int in = 0; in < MLP_H; ++in)
    for (int out = 0; out < MLP_H; ++out) w[64 * 16 + out * MLP_H + in] = bf(p[MLP_OW2 + in * MLP_H + out]);
  for (int in = 0; in < MLP_H; ++in)
    for (int out = 0; out < MLP_OUTP; ++out)
      w[64 * 16 + 64 * 64 + out * MLP_H + in] = bf(p[MLP_OW3 + in * MLP_OUTP + out]);
  HIPCHK(hipMemcpyAsync(h->d_netw, w.data(), sizeof(unsigned short) * w.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_netp, p.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_netnorm, norm, sizeof(float) * MLP_NNORM, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->have_net = 1;
  return 0;
}

int dgpu_closure_counts(dgpu_handle* h, int64_t* network, int64_t* points) {
  if (!h || !network || !points) return fail(1, "null argument");
  if (!h->d_ccount) return fail(2, "dgpu_closure_counts needs an les_model_type 6 handle");
  unsigned long long c = 0;
  HIPCHK(hipMemsetAsync(h->d_ccount, 0, sizeof(c), h->st));
  const Geo& g = h->g;
  if (h->prec == DGPU_MIXED)
    k_closure_count<float><<<1024, 256, 0, h->st>>>((const float*)h->Cgy, g.pp, g.ny, h->d_ccount);
  else
    k_closure_count<double><<<1024, 256, 0, h->st>>>((const double*)h->Cgy, g.pp, g.ny, h->d_ccount);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&c, h->d_ccount, sizeof(c), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  *network = (int64_t)c;
  *points = (int64_t)(g.ny - 1) * g.pp;
  return 0;
}

int dgpu_les_features(dgpu_handle* h, double time, float* X, float* scale, float* amd, int32_t* inside,
                      int64_t* n_out) {
  if (!h || !n_out) return fail(1, "null argument");
  if (!h->c.use_les || h->nth < 1) return fail(2, "dgpu_les_features needs an LES handle with a scalar");
  const Geo& g = h->g;
  const int jf0 = h->c.jstart, ny = g.ny;
  const long long n = (long long)(ny - jf0) * g.pp;
  *n_out = n;
  if (!X && !scale && !amd && !inside) return 0;
  // the BC posts and gradient fields of an LES substep, then the features
  int rc;
  if ((rc = launch_bc_post(h, time, 1))) return rc;
  const bool mixed = h->prec == DGPU_MIXED;
  if ((rc = mixed ? physical_gradients<float>(h, h->Hin, h->HinZ, h->Pg)
                  : physical_gradients<double>(h, h->Hin, h->HinZ, h->Pg)))
    return rc;
  float *dX = nullptr, *dS = nullptr, *dA = nullptr;
  int* dI = nullptr;
  auto cleanup = [&]() {
    for (void* q : {(void*)dX, (void*)dS, (void*)dA, (void*)dI})
      if (q) (void)hipFree(q);
  };
  if (hipMalloc((void**)&dX, sizeof(float) * MLP_INP * (size_t)n) != hipSuccess ||
      hipMalloc((void**)&dS, sizeof(float) * (size_t)n) != hipSuccess ||
      hipMalloc((void**)&dA, sizeof(float) * 2 * (size_t)n) != hipSuccess ||
      hipMalloc((void**)&dI, sizeof(int) * (size_t)n) != hipSuccess) {
    cleanup();
    return fail(11, "hipMalloc failed (feature buffers)");
  }
  const AmdArgs aa = amd_args(h);
  ClosureArgs k{};
  if (h->c.les_model_type == 6) {
    k = closure_args(h, aa);
  } else {
    k.a = aa;
    k.have_net = 0;
    closure_consts(h, k);
  }
  k.feat = 1;
  k.jf0 = jf0;
  k.X = dX; k.scale = dS; k.amd = dA; k.inside = dI;
  const dim3 grid((unsigned)(((g.pp + 255) / 256) * g.ny2));  // xcd_rows maps the 1D index
  bool ok = true;
  if (h->nth == 1) {
    if (mixed) k_closure<float, 1><<<grid, 256, 0, h->st>>>(k);
    else k_closure<double, 1><<<grid, 256, 0, h->st>>>(k);
  } else {
    cleanup();
    return fail(2, "dgpu_les_features supports one scalar");
  }
  ok = hipGetLastError() == hipSuccess;
  if (ok && X) ok = hipMemcpyAsync(X, dX, sizeof(float) * MLP_INP * (size_t)n, hipMemcpyDeviceToHost, h->st) == hipSuccess;
  if (ok && scale) ok = hipMemcpyAsync(scale, dS, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, h->st) == hipSuccess;
  if (ok && amd) ok = hipMemcpyAsync(amd, dA, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->st) == hipSuccess;
  if (ok && inside) ok = hipMemcpyAsync(inside, dI, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, h->st) == hipSuccess;
  ok = ok && hipStreamSynchronize(h->st) == hipSuccess;
  cleanup();
  if (!ok) return fail(10, "dgpu_les_features: kernel or download failed");
  return 0;
}

int dgpu_box_filter(dgpu_handle* h, int32_t fx, int32_t fz, int32_t nl, const double* mf, const double* mg,
                    double* out) {
  if (!h || !mf || !mg || !out) return fail(1, "null argument");
  const Geo& g = h->g;
  if (h->nth < 1) return fail(2, "dgpu_box_filter needs a scalar (n_th >= 1)");
  if (fx < 2 || fz < 2 || fx % 2 || fz % 2 || g.nx % fx || g.nz % fz)
    return fail(2, "dgpu_box_filter: fx, fz must be even divisors of nx, nz");
  if (nl < 2) return fail(2, "dgpu_box_filter: nl must be >= 2");
  const int nl2 = nl + 2, nxl = g.nx / fx, nzl = g.nz / fz;
  int rc = ensure_grad_buffers(h);
  if (rc) return rc;
  // physical u1, u2, u3, th of the state, FP64
  const cd* cs[4] = {h->U[0], h->U[1], h->U[2], h->U[3]};
  void* hs[4] = {h->HinD[0], h->HinD[1], h->HinD[2], h->HinD[3]};
  double* rs[4] = {(double*)h->PgD[0], (double*)h->PgD[3], (double*)h->PgD[6], (double*)h->PgD[9]};
  if ((rc = launch_inv_z<double>(h, 4, cs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_inv_x<double>(h, 4, hs, rs, 0, g.ny2))) return rc;
  // the y weights, their nonzero row ranges and the output in one device buffer
  const size_t mbytes = sizeof(double) * (size_t)nl2 * g.ny2;
  const size_t rbytes = sizeof(int) * 2 * (size_t)nl2;
  const size_t obytes = sizeof(double) * BOX_FIELDS * (size_t)nl2 * nzl * nxl;
  const size_t need = 2 * round256(mbytes) + 2 * round256(rbytes) + obytes;
  if (h->box_bytes < need) {
    HIPCHK(hipStreamSynchronize(h->st));
    if (h->d_box) HIPCHK(hipFree(h->d_box));
    h->d_box = nullptr;
    h->box_bytes = 0;
    HIPCHK(hipMalloc(&h->d_box, need));
    h->box_bytes = need;
  }
  std::vector<int> ranges(4 * (size_t)nl2);
  for (int m = 0; m < 2; ++m) {
    const double* M = m ? mg : mf;
    for (int j = 0; j < nl2; ++j) {
      int lo = 1, hi = 0;
      for (int k = 0; k < g.ny2; ++k)
        if (M[(size_t)j * g.ny2 + k] != 0.0) {
          if (lo > hi) lo = k;
          hi = k;
        }
      ranges[(size_t)m * 2 * nl2 + 2 * j] = lo;
      ranges[(size_t)m * 2 * nl2 + 2 * j + 1] = hi;
    }
  }
  unsigned char* b = (unsigned char*)h->d_box;
  BoxArgs a{};
  a.g = g; a.fx = fx; a.fz = fz; a.nxl = nxl; a.nzl = nzl; a.nl2 = nl2;
  for (int f = 0; f < 4; ++f) a.u[f] = rs[f];
  a.mf = (const double*)b;
  a.mg = (const double*)(b + round256(mbytes));
  a.rf = (const int*)(b + 2 * round256(mbytes));
  a.rg = (const int*)(b + 2 * round256(mbytes) + round256(rbytes));
  a.out = (double*)(b + 2 * round256(mbytes) + 2 * round256(rbytes));
  HIPCHK(hipMemcpyAsync((void*)a.mf, mf, mbytes, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync((void*)a.mg, mg, mbytes, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync((void*)a.rf, ranges.data(), rbytes, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync((void*)a.rg, ranges.data() + 2 * (size_t)nl2, rbytes, hipMemcpyHostToDevice, h->st));
  k_box_filter<<<dim3((unsigned)((nxl * nzl + 255) / 256), (unsigned)nl2), 256, 0, h->st>>>(a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, a.out, obytes, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_set_graph(dgpu_handle* h, int32_t enable) {
  if (!h) return fail(1, "null handle");
  h->graph_on = enable != 0;
  if (!h->graph_on && h->gexec) {
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  return 0;
}

int dgpu_set_variable_dt(dgpu_handle* h, int32_t variable_dt, int32_t update_dt, double cfl) {
  if (!h) return fail(1, "null handle");
  if (variable_dt && update_dt < 1) return fail(2, "update_dt must be >= 1");
  h->variable_dt = variable_dt != 0;
  h->update_dt = update_dt < 1 ? 1 : update_dt;
  h->cfl = cfl;
  h->cfl_ran = 0;
  return 0;
}

int dgpu_courant_physical(dgpu_handle* h, const double* u1, const double* u2, const double* u3) {
  if (!h || !u1 || !u2 || !u3) return fail(1, "null argument");
  if (!h->variable_dt) return 0;
  const Geo& g = h->g;
  int rc = ensure_grad_buffers(h);
  if (rc) return rc;
  // physical Fortran layout u(0:Nx+1, 0:Nzp+1, 0:Nyp+1) -> [j][z][x] in the
  // gradient stage's FP64 buffers
  std::vector<double> tmp((size_t)g.pp * g.ny2);
  const double* src[3] = {u1, u2, u3};
  double* dst[3] = {(double*)h->PgD[0], (double*)h->PgD[3], (double*)h->PgD[6]};
  for (int f = 0; f < 3; ++f) {
    for (int j = 0; j < g.ny2; ++j)
      for (int z = 0; z < g.nz; ++z)
        std::memcpy(&tmp[((size_t)j * g.nz + z) * g.nx], src[f] + ((size_t)j * (g.nz + 2) + z) * (g.nx + 2),
                    sizeof(double) * g.nx);
    HIPCHK(hipMemcpyAsync(dst[f], tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));  // tmp is refilled for the next field
  }
  if ((rc = launch_courant(h, dst[0], dst[1], dst[2]))) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_courant(dgpu_handle* h, double time, double time_limit, double save_stats_time, double save_flow_time,
                 double* delta_t, int32_t* ran) {
  if (!h || !delta_t || !ran) return fail(1, "null argument");
  *ran = h->cfl_ran;
  *delta_t = 0.0;
  if (!h->cfl_ran) return 0;
  h->cfl_ran = 0;
  std::vector<double> part(COURANT_BLOCKS);
  HIPCHK(hipMemcpyAsync(part.data(), h->d_cfl, sizeof(double) * COURANT_BLOCKS, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const dgpu_config& c = h->c;
  // the limits of courant before its velocity loop (for/flow.f90:757-773)
  const double pi = 3.1415927410125732;  // pi = 4.*atan(1.0) in real(4), for/fft.f90:147
  const double dx = c.lx / c.nx, dz = c.lz / c.nz;
  double dt = 1.0;
  dt = std::fmin(dt, 0.5 * std::pow(std::fmin(dx, dz), 2.0 * c.beta) / c.nu);
  for (int n = 0; n < c.n_th; ++n) dt = std::fmin(dt, dt * c.nu / (c.nu / c.pr[n]));
  if (c.ro_inv != 0.0) dt = std::fmin(dt, 2.0 * pi / std::fabs(c.ro_inv / c.delta) / 20.0);
  for (int n = 0; n < c.n_th; ++n)
    dt = std::fmin(dt, (double)0.1f * 2.0 * pi / std::sqrt(std::fabs(c.th_bc_ymin_c1[n])));
  for (double v : part) dt = std::fmin(dt, v);
  // caps and the scheduling towards the next event (for/flow.f90:810-851)
  double dtn;
  if (dt <= 0) dtn = 0.00001;
  else if (dt >= (double)1000.f) dtn = 1000.0;
  else dtn = dt;
  const double next = std::fmin(time_limit, std::fmin(save_stats_time - time,
                                std::fmin(save_flow_time - time, time_limit - time)));
  const double r_max = (double)(1.f - 0.025f);
  if (next < dt) dtn = next + 1.e-14;
  else if (next < (1.0 + r_max) * dt) dtn = 1.0 / (1.0 + r_max) * next;
  else dtn = dt;
  *delta_t = dtn;
  return 0;
}

int64_t dgpu_stats_doubles(const dgpu_config* cfg) {
  if (!cfg) return -1;
  return (int64_t)(SP_NBASE + SP_PER_TH * cfg->n_th) * (cfg->ny + 2) + 1 + cfg->nx / 2;
}

namespace {
// The first stage shared by the statistics entry points: the BC posts on the
// state, the physical gradient fields and the z-line means (k_stat_mean).
int stats_prepare(dgpu_handle* h, double time, StatArgs& a) {
  const dgpu_config& c = h->c;
  const Geo& g = h->g;
  const int nth = h->nth, n2 = g.ny2;
  int rc = ensure_grad_buffers(h);
  if (rc) return rc;
  void* p = nullptr;
  if (!h->d_zm) {
    if ((rc = dalloc(h, &p, sizeof(double) * (6 + MAXTH) * n2 * g.nx))) return rc;
    h->d_zm = (double*)p;
    if ((rc = dalloc(h, &p, sizeof(double) * (SZ_TH0 + SZ_PER_TH * MAXTH) * n2 * g.nx))) return rc;
    h->d_zs = (double*)p;
    if ((rc = dalloc(h, &p, sizeof(double) * (SP_NBASE + SP_PER_TH * MAXTH) * n2))) return rc;
    h->d_prof = (double*)p;
  }
  // save_stats_chan starts with the BC posts on the state (for/diagnostics.f90:26-27)
  if ((rc = launch_bc_post(h, time, 0))) return rc;
  if ((rc = physical_gradients<double>(h, h->HinD, h->HinZD, h->PgD))) return rc;
  a = StatArgs{};
  a.g = g; a.nth = nth;
  a.hx = c.homogeneous_x != 0;
  a.nu = c.nu; a.ly = c.ly; a.lx = c.lx;
  // the thermal-wind shear (1/f) dTHdX(1) that the stats add to dw/dy and wme
  a.twf = (nth > 0 && c.ro_inv != 0.0) ? (1.0 / (c.ro_inv / c.delta)) * c.dthdx[0] : 0.0;
  for (int n = 0; n < MAXTH; ++n) { a.dthdx[n] = c.dthdx[n]; a.dthdz[n] = c.dthdz[n]; }
  a.dy = h->d_dy; a.dyf = h->d_dyf; a.gy = h->d_gy; a.gyf = h->d_gyf;
  for (int s = 0; s < 3 * (3 + nth); ++s) a.f[s] = (const double*)h->PgD[s];
  for (int f = 0; f < 3 + nth; ++f) a.u[f] = h->U[f];
  a.p = h->P;
  a.zm = h->d_zm; a.zs = h->d_zs; a.prof = h->d_prof;
  const dim3 gxj((unsigned)((g.nx + 255) / 256), (unsigned)n2);
  k_stat_mean<<<gxj, 256, 0, h->st>>>(a);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace

int dgpu_stats(dgpu_handle* h, double time, double* out) {
  if (!h || !out) return fail(1, "null argument");
  const dgpu_config& c = h->c;
  const Geo& g = h->g;
  const int nth = h->nth, n2 = g.ny2;
  StatArgs a;
  int rc = stats_prepare(h, time, a);
  if (rc) return rc;
  const dim3 gxj((unsigned)((g.nx + 255) / 256), (unsigned)n2);
  k_stat_zsum<<<gxj, 256, 0, h->st>>>(a);
  k_stat_plane<<<(unsigned)n2, 256, 0, h->st>>>(a);
  const int nxp = g.nx / 2;
  const size_t np_ = (size_t)(SP_NBASE + SP_PER_TH * nth) * n2;
  // FTx_uu after the profiles (d_prof holds MAXTH scalars' room, enough for Nx/2 more)
  double* ftx = h->d_ftx;
  if (!ftx) {
    void* q = nullptr;
    if ((rc = dalloc(h, &q, sizeof(double) * (size_t)(nxp > 0 ? nxp : 1)))) return rc;
    ftx = h->d_ftx = (double*)q;
  }
  k_stat_ftx<<<(unsigned)((nxp + 255) / 256), 256, 0, h->st>>>(a, ftx, nxp);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, h->d_prof, sizeof(double) * np_, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(out + np_ + 1, ftx, sizeof(double) * nxp, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  // u1z_x0: integrate_y_var of u1y_left (for/diagnostics.f90:1964-1980)
  const double* u1y = out + (size_t)SP_U1Y * n2;
  double res = 0.0;
  for (int j = 2; j <= g.ny; ++j) res = res + 0.5 * (u1y[j] + u1y[j - 1]) * h->dy[j];
  out[np_] = res / c.ly;
  return 0;
}

int dgpu_stats_movie_sizes(const dgpu_config* cfg, int64_t* n_planes, int64_t* n_slices) {
  if (!cfg || !n_planes || !n_slices || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0) return 1;
  const int64_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  *n_planes = (int64_t)NXZ_PLANES * ny * nx;
  *n_slices = (int64_t)(MOVIE_BASE + cfg->n_th) * (ny * nx + nz * nx + ny * nz);
  return 0;
}

int dgpu_stats_movie(dgpu_handle* h, double time, int32_t nxm, int32_t nym, int32_t nzm, double* planes,
                     double* slices) {
  if (!h || !planes || !slices) return fail(1, "null argument");
  const Geo& g = h->g;
  if (nxm < 0 || nxm >= g.nx || nzm < 0 || nzm >= g.nz || nym == 0 || nym < -1 || nym > g.ny)
    return fail(2, "movie indices out of range (0 <= NxMovie < Nx, 0 <= NzMovie < Nz, "
                   "NyMovie in 1..Nyp or -1)");
  int64_t npl = 0, nsl = 0;
  dgpu_stats_movie_sizes(&h->c, &npl, &nsl);
  int rc;
  if (!h->d_movie) {
    void* p = nullptr;
    if ((rc = dalloc(h, &p, sizeof(double) * (size_t)(npl + nsl)))) return rc;
    h->d_movie = (double*)p;
  }
  StatArgs a;
  if ((rc = stats_prepare(h, time, a))) return rc;
  const int n2 = g.ny2;
  const dim3 gxj((unsigned)((g.nx + 255) / 256), (unsigned)n2);
  k_stat_zsum<<<gxj, 256, 0, h->st>>>(a);
  k_stat_plane<<<(unsigned)n2, 256, 0, h->st>>>(a);
  k_stat_xz<<<dim3((unsigned)((g.nx + 255) / 256), (unsigned)g.ny), 256, 0, h->st>>>(a, h->d_movie);
  MovieArgs m{};
  m.nxm = nxm; m.nym = nym; m.nzm = nzm; m.nfld = MOVIE_BASE + h->nth;
  m.nut = h->c.use_les ? h->NuT : nullptr;
  m.nut_f32 = h->prec == DGPU_MIXED;
  m.out = h->d_movie + npl;
  const long long per = (long long)g.ny * g.nx + (long long)g.nz * g.nx + (long long)g.ny * g.nz;
  if (nym < 1) HIPCHK(hipMemsetAsync(m.out, 0, sizeof(double) * (size_t)nsl, h->st));
  k_stat_movie<<<(unsigned)((per + 255) / 256), 256, 0, h->st>>>(a, m);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(planes, h->d_movie, sizeof(double) * (size_t)npl, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(slices, m.out, sizeof(double) * (size_t)nsl, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int64_t dgpu_stats_zstar_doubles(const dgpu_config* cfg) {
  if (!cfg || cfg->n_th < 1 || cfg->nx <= 0) return -1;
  const int64_t nb = cfg->nx;
  return nb + (nb + 1) + 1 + (4 + 2 * (int64_t)cfg->n_th) * nb;
}

int dgpu_stats_zstar(dgpu_handle* h, double time, double* out) {
  if (!h || !out) return fail(1, "null argument");
  const dgpu_config& c = h->c;
  const Geo& g = h->g;
  const int nth = h->nth, nbin = g.nx, nf = 5 + 2 * nth;
  if (nth < 1) return fail(2, "z* statistics need a scalar (N_th >= 1)");
  const size_t lds = sizeof(double) * (size_t)nf * nbin;
  if (lds > 65536) return fail(2, "z* statistics: Nx (5 + 2 N_th) bins exceed the 64 KiB LDS histogram");
  StatArgs a;
  int rc = stats_prepare(h, time, a);
  if (rc) return rc;
  if (!h->d_zhist) {
    void* p = nullptr;
    if ((rc = dalloc(h, &p, sizeof(double) * (5 + 2 * MAXTH) * (size_t)nbin))) return rc;
    h->d_zhist = (double*)p;
  }
  HIPCHK(hipMemsetAsync(h->d_zhist, 0, lds, h->st));
  ZstarArgs z{};
  z.nbin = nbin; z.nf = nf; z.jlo = c.jstart_th[0]; z.jhi = c.jend_th[0]; z.hx = a.hx;
  z.lx = c.lx; z.dthdx0 = c.dthdx[0];
  z.pi = (double)(4.0f * std::atan(1.0f));  // the reference's pi (for/fft.f90:147)
  if (z.hx) {
    // -dTHdX(1) * 0.5*Lx * 1.1 with the single-precision literal 1.1
    z.thmin = -(c.dthdx[0] * 0.5 * c.lx * (double)1.1f);
    z.thmax = c.dthdx[0] * 0.5 * c.lx * (double)1.1f;
    z.dth = (z.thmax - z.thmin) / nbin + 1.0e-14;
  } else {
    z.thmin = -c.ro_inv;
    z.thmax = c.ro_inv;
    z.den = c.ro_inv + 1.0e-14;
  }
  z.hist = h->d_zhist;
  k_stat_zstar<<<dim3((unsigned)((g.nx + 255) / 256), (unsigned)g.ny2), 256, lds, h->st>>>(a, z);
  HIPCHK(hipGetLastError());
  std::vector<double> hist((size_t)nf * nbin);
  HIPCHK(hipMemcpyAsync(hist.data(), h->d_zhist, lds, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  // compute_BPE (for/diagnostics.f90:1794-1830): PDF, bin edges, Y_r, BPE
  double* pdf = out;
  double* thb = out + nbin;
  double* bpe = thb + nbin + 1;
  double* fld = bpe + 1;
  double tot = 0.0;
  for (int i = 0; i < nbin; ++i) tot += hist[i];
  for (int i = 0; i < nbin; ++i) pdf[i] = hist[i] / tot;
  for (int i = 0; i <= nbin; ++i)
    thb[i] = z.hx ? z.thmin + i * z.dth : -(c.ro_inv + 1.0e-14) * std::cos(i * z.pi / nbin);
  std::vector<double> yr(nbin + 1, 0.0);
  for (int i = 1; i <= nbin; ++i) yr[i] = yr[i - 1] + pdf[i - 1];
  for (int i = 0; i <= nbin; ++i) yr[i] = yr[i] * c.ly;
  double b = 0.0;
  for (int i = 0; i < nbin; ++i)
    b = b - (0.5 * (thb[i + 1] + thb[i]) * 0.5 * (yr[i + 1] + yr[i])) * (yr[i + 1] - yr[i]) / c.ly;
  *bpe = b;
  // Bin_Ystar_and_Write (for/diagnostics.f90:1936): the binned sums over the raw PDF
  for (int f = 1; f < nf; ++f)
    for (int i = 0; i < nbin; ++i) fld[(size_t)(f - 1) * nbin + i] = hist[(size_t)f * nbin + i] / hist[i];
  return 0;
}

int dgpu_rk_substep(dgpu_handle* h, int32_t rk_step, double delta_t, double time, int32_t time_step,
                    int32_t create_new_flow) {
  if (!h) return fail(1, "null handle");
  h->timing_now = true;
  HIPCHK(hipEventRecord(h->e0, h->st));
  int rc = substep(h, rk_step, delta_t, time, time_step, create_new_flow);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->e1, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return harvest(h);
}

int dgpu_step(dgpu_handle* h, double delta_t, double time, int32_t time_step, int32_t create_new_flow) {
  if (!h) return fail(1, "null handle");
  h->timing_now = (h->timing_calls++ % h->timing_every) == 0;
  const dgpu_config& c = h->c;
  // The step's launches depend on the time only through the f_type 2 and 4
  // forcings, on host decisions inside the step only with variable_dt (courant
  // reads dt back in rk 3), and carry per-kernel events only while timing: every
  // other step replays one captured graph per (delta_t, LES gate).
  const bool les_on = c.use_les && (!create_new_flow || time_step > 100);
  if (h->graph_on && !h->variable_dt && c.f_type != 2 && c.f_type != 4 && !(h->timing && h->timing_now)) {
    int rc;
    if (!h->gexec || h->g_dt != delta_t || h->g_les != (int)les_on) {
      if (h->gexec) {
        HIPCHK(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
      }
      if ((rc = ensure_factors(h, delta_t))) return rc;  // host work and copies: before the capture
      HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
      for (int rk = 1; rk <= 3 && !rc; ++rk) rc = substep(h, rk, delta_t, time, time_step, create_new_flow);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(h->st, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (ec != hipSuccess) return fail(10, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        h->gexec = nullptr;
        return fail(10, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      h->g_dt = delta_t;
      h->g_les = (int)les_on;
    }
    HIPCHK(hipEventRecord(h->e0, h->st));
    HIPCHK(hipGraphLaunch(h->gexec, h->st));
    HIPCHK(hipEventRecord(h->e1, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    rc = harvest(h);
    h->timing_now = true;
    return rc;
  }
  HIPCHK(hipEventRecord(h->e0, h->st));
  for (int rk = 1; rk <= 3; ++rk) {
    int rc = substep(h, rk, delta_t, time, time_step, create_new_flow);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->e1, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const int rc = harvest(h);
  // calls between steps (rem_div, poisson_p, filter_chan) always record
  h->timing_now = true;
  return rc;
}

int dgpu_rem_div(dgpu_handle* h) {
  if (!h) return fail(1, "null handle");
  ColArgs ca{};
  fill_col_args(h, ca, 0.0);
  int rc = launch_project(h, ca, 0);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_filter_chan(dgpu_handle* h, int32_t n) {
  if (!h) return fail(1, "null handle");
  if (n < 0 || n >= h->nth) return fail(2, "dgpu_filter_chan: no such scalar");
  const dgpu_config& c = h->c;
  const Geo& g = h->g;
  if (!h->d_sigma) {
    // sigma(i, k) of filter_chan (for/solver.f90:1617-1626) in the packed layout
    std::vector<double> sg((size_t)g.pc, 0.0);
    for (int k = 0; k < g.kzr; ++k)
      for (int i = 0; i < g.kxr; ++i) {
        const double ax = h->kx[i] * c.lx * 1.0 / (double)(float)c.nx;
        const double az = h->kz[k] * c.lz * 1.0 / (double)(float)c.nz;
        const double s0 = 0.5 * (1.0 + std::cos(std::sqrt(std::pow(ax, 2.0) + std::pow(az, 2.0))));
        sg[(size_t)k * g.kxp + i] =
            std::pow(s0, 4.0) * (35.0 - 84.0 * s0 + 70.0 * std::pow(s0, 2.0) - 20.0 * std::pow(s0, 3.0));
      }
    void* p = nullptr;
    int rc = dalloc(h, &p, sizeof(double) * sg.size());
    if (rc) return rc;
    h->d_sigma = (double*)p;
    HIPCHK(hipMemcpyAsync(h->d_sigma, sg.data(), sizeof(double) * sg.size(), hipMemcpyHostToDevice, h->st));
  }
  const int jlo = c.jstart_th[n], jhi = c.jend_th[n];
  if (jhi >= jlo) {
    const dim3 grid((unsigned)((g.pc + 255) / 256), jhi - jlo + 1);
    k_filter_th<<<grid, 256, 0, h->st>>>(g, h->U[3 + n], h->d_sigma, jlo, jhi);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_poisson_p(dgpu_handle* h) {
  if (!h) return fail(1, "null handle");
  const Geo& g = h->g;
  int rc = ensure_grad_buffers(h);
  if (rc) return rc;
  if ((rc = ensure_fft_scratch(h))) return rc;
  // the FP64 physical gradient fields of the state, the source on rows 2..Nyp,
  // its transform (fft_xz_to_fourier), then the projection's Poisson solve with
  // that right-hand side (the same matrix and BC rows as rem_div_chan)
  if ((rc = physical_gradients<double>(h, h->HinD, h->HinZD, h->PgD))) return rc;
  StatArgs a{};
  a.g = g; a.nth = h->nth; a.dyf = h->d_dyf;
  for (int s = 0; s < 3 * (3 + h->nth); ++s) a.f[s] = (const double*)h->PgD[s];
  k_pp_source<<<dim3((unsigned)((g.pp + 255) / 256), (unsigned)g.ny2), 256, 0, h->st>>>(a, h->d_R);
  HIPCHK(hipGetLastError());
  double* rs[1] = {h->d_R};
  void* hs[1] = {h->d_H};
  cd* cs[1] = {h->d_C};
  if ((rc = launch_fwd_x<double>(h, 1, rs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_fwd_z<double>(h, 1, hs, cs, 0, g.ny2))) return rc;
  ColArgs ca{};
  fill_col_args(h, ca, 0.0);
  ca.prhs = h->d_C;
  ca.pnth = h->nth;
  if ((rc = launch_project(h, ca, 0))) return rc;
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

int dgpu_fft_xz_to_fourier_noalloc(dgpu_handle* h, const double* phys, double* four) {
  if (!h || !phys || !four) return fail(1, "null argument");
  const Geo& g = h->g;
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != h->dev) return fail(2, "dgpu_fft_xz_to_fourier_noalloc: the current device is not the handle's");
  const size_t rb = round256(sizeof(double) * g.pp * g.ny2);
  const size_t hb = round256(sizeof(cd) * g.ph * g.ny2);
  const size_t cb = round256(sizeof(cd) * g.pc * g.ny2);
  if (rb + hb + cb > kNoallocScratchBytes)
    return fail(2, "dgpu_fft_xz_to_fourier_noalloc: this grid needs " + std::to_string(rb + hb + cb) +
                       " bytes of transform scratch; the allocation-free entry point holds " +
                       std::to_string(kNoallocScratchBytes));
  double* R = (double*)h->noalloc;
  cd* H = (cd*)(h->noalloc + rb);
  cd* C = (cd*)(h->noalloc + rb + hb);
  const int nb = 1024;
  k_pack_phys<<<nb, 256, 0, h->st>>>(phys, R, g);
  HIPCHK(hipGetLastError());
  double* rs[1] = {R};
  void* hs[1] = {H};
  cd* cs[1] = {C};
  int rc;
  if ((rc = launch_fwd_x<double>(h, 1, rs, hs, 0, g.ny2))) return rc;
  if ((rc = launch_fwd_z<double>(h, 1, hs, cs, 0, g.ny2))) return rc;
  k_unpack_four<<<nb, 256, 0, h->st>>>(C, (cd*)four, g);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int dgpu_timing_enable(dgpu_handle* h, int32_t enable) {
  if (!h) return fail(1, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = harvest(h);
  if (rc) return rc;
  h->timing = enable != 0;
  h->timing_every = 1;
  h->timing_calls = 0;
  h->timing_now = true;
  for (int k = 0; k < NKERN; ++k) { h->t_ms[k] = 0.0; h->t_n[k] = 0; }
  return 0;
}

int dgpu_timing_every(dgpu_handle* h, int32_t every) {
  if (!h || every < 1) return fail(1, "dgpu_timing_every: every must be >= 1");
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = harvest(h);
  if (rc) return rc;
  h->timing_every = every;
  h->timing_calls = 0;
  h->timing_now = true;
  return 0;
}

int dgpu_timing_read(dgpu_handle* h, int32_t kernel, double* total_ms, int64_t* launches, const char** name) {
  if (!h || !total_ms || !launches) return fail(1, "null argument");
  if (kernel < 0 || kernel >= NKERN) return fail(2, "kernel id out of range");
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = harvest(h);
  if (rc) return rc;
  *total_ms = h->t_ms[kernel];
  *launches = h->t_n[kernel];
  if (name) *name = kKernelNames[kernel];
  return 0;
}

int dgpu_get_sizes(const dgpu_handle* h, int32_t* nx, int32_t* ny, int32_t* nz, int32_t* nth) {
  if (!h || !nx || !ny || !nz || !nth) return fail(1, "null argument");
  *nx = h->c.nx; *ny = h->c.ny; *nz = h->c.nz; *nth = h->c.n_th;
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
