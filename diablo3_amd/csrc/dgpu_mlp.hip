// dgpu_mlp.hip — C ABI of the learned closure network (include/dgpu_closure.h):
// parameters, data, full-batch gradients, Adam, minibatch training and inference
// on the device with the kernels of dgpu_mlp.hpp.  No counterpart in the
// reference (a further LES model beside les_chan's, for/les.f90).
#include "../../include/dgpu_closure.h"
#include "dgpu_error.hpp"
#include "dgpu_mlp.hpp"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace dgpu;

static_assert(DGPU_MLP_NPARAM == MLP_NP, "parameter block size");
static_assert(DGPU_MLP_NNORM == MLP_NNORM, "normaliser size");

namespace dgpu {

// ---------------------------------------------------------------------------
// Training: the gradient of the minibatch loss
//   L = (1/n) sum_s sum_o wo_o (y_so - t_so)^2
// One block (4 waves) accumulates its tiles in registers; the waves' sums are
// added in a fixed order in LDS and written as the block's partial gradient,
// so the result does not depend on scheduling (no atomics).
// ---------------------------------------------------------------------------
struct MlpGradArgs {
  MlpW w;
  const float* X;    // [N][16] raw features
  const float* T;    // [N][2] targets
  const int* idx;    // minibatch sample indices (null: samples 0..n-1)
  long long n;       // samples in the minibatch
  float wo0, wo1;    // output weights of the loss
  float* part;       // [gridDim.x][MLP_NP] partial gradients
  float* lpart;      // [gridDim.x] partial loss sums (sum of wo e^2, unscaled)
};

__device__ __forceinline__ long long mlp_row(const int* idx, long long s) { return idx ? (long long)idx[s] : s; }

__global__ __launch_bounds__(256) void k_mlp_grad(MlpGradArgs a) {
  __shared__ unsigned short sA[4][64 * 16];  // per wave: [neuron][sample] transposes
  __shared__ unsigned short sB[4][64 * 16];
  __shared__ float red[MLP_NP];
  __shared__ float sred[256];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  unsigned short* TA = sA[wave];
  unsigned short* TB = sB[wave];
  // weight-gradient accumulators: dW^T tiles [out][in] in the accumulator layout
  f32x4 gw1[4], gw2[4][4], gw3[4];
  f32x4 gb1[4], gb2[4], gb3;
  const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gw1[i] = zero4; gw3[i] = zero4; gb1[i] = zero4; gb2[i] = zero4;
#pragma unroll
    for (int k = 0; k < 4; ++k) gw2[i][k] = zero4;
  }
  gb3 = zero4;
  float lsum = 0.f;
  const long long ntile = (a.n + 15) / 16;
  const float inv_n = 1.0f / (float)a.n;
  for (long long t = (long long)blockIdx.x * 4 + wave; t < ntile; t += (long long)gridDim.x * 4) {
    const long long s = t * 16 + r;
    const bool live = s < a.n;
    const long long row = live ? mlp_row(a.idx, s) : 0;
    bf16x4 xb = norm_x(a.w.norm, a.X + row * MLP_INP + 4 * g, g);
    if (!live) xb = bf16x4{0, 0, 0, 0};
    MlpAct act;
    mlp_forward(a.w, xb, act);
    // dL/dy in the accumulator layout (rows = outputs 4g+v, column = sample r)
    f32x4 dy = zero4;
    if (g == 0 && live) {
      const float e0 = act.y[0] - a.T[row * 2], e1 = act.y[1] - a.T[row * 2 + 1];
      lsum += a.wo0 * e0 * e0 + a.wo1 * e1 * e1;
      dy[0] = 2.f * a.wo0 * e0 * inv_n;
      dy[1] = 2.f * a.wo1 * e1 * inv_n;
    }
    const bf16x4 dyb = tobf(dy);
    // layer 3: dW3^T[out][in] = dY^T H2 (K = samples): TA <- dY^T [16][16], TB <- H2^T [64][16]
#pragma unroll
    for (int v = 0; v < 4; ++v) TA[(4 * g + v) * 16 + r] = (unsigned short)dyb[v];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int v = 0; v < 4; ++v) TB[(kb * 16 + 4 * g + v) * 16 + r] = (unsigned short)act.h2[kb][v];
    __builtin_amdgcn_wave_barrier();
    {
      const bf16x4 A = ld4(TA + r * 16 + 4 * g);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gw3[ib] = mfma16(A, ld4(TB + (ib * 16 + r) * 16 + 4 * g), gw3[ib]);
    }
    gb3 += dy;
    __builtin_amdgcn_wave_barrier();
    // dH2^T = W3 dY^T; dZ2 = dH2 * gelu'(z2)
    f32x4 dz2[4];
    bf16x4 dz2b[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      f32x4 c = mfma16(ld4(a.w.w3 + (kb * 16 + r) * MLP_OUTP + 4 * g), dyb, zero4);
#pragma unroll
      for (int v = 0; v < 4; ++v) c[v] *= dgelu(act.z2[kb][v]);
      dz2[kb] = c;
      dz2b[kb] = tobf(c);
    }
    // layer 2: dW2^T[out][in] = dZ2^T H1: TA <- dZ2^T [64][16], TB <- H1^T [64][16]
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        TA[(ob * 16 + 4 * g + v) * 16 + r] = (unsigned short)dz2b[ob][v];
        TB[(ob * 16 + 4 * g + v) * 16 + r] = (unsigned short)act.h1[ob][v];
      }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const bf16x4 A = ld4(TA + (ob * 16 + r) * 16 + 4 * g);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gw2[ob][ib] = mfma16(A, ld4(TB + (ib * 16 + r) * 16 + 4 * g), gw2[ob][ib]);
      gb2[ob] += dz2[ob];
    }
    __builtin_amdgcn_wave_barrier();
    // dH1^T = W2 dZ2^T; dZ1 = dH1 * gelu'(z1)
    f32x4 dz1[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      f32x4 c = zero4;
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) c = mfma16(ld4(a.w.w2 + (kb * 16 + r) * MLP_H + ob * 16 + 4 * g), dz2b[ob], c);
#pragma unroll
      for (int v = 0; v < 4; ++v) c[v] *= dgelu(act.z1[kb][v]);
      dz1[kb] = c;
    }
    // layer 1: dW1^T[out][in] = dZ1^T X: TA <- dZ1^T [64][16], TB <- X^T [16][16]
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const bf16x4 b = tobf(dz1[ob]);
#pragma unroll
      for (int v = 0; v < 4; ++v) TA[(ob * 16 + 4 * g + v) * 16 + r] = (unsigned short)b[v];
      gb1[ob] += dz1[ob];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) TB[(4 * g + v) * 16 + r] = (unsigned short)xb[v];
    __builtin_amdgcn_wave_barrier();
    {
      const bf16x4 B = ld4(TB + r * 16 + 4 * g);
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) gw1[ob] = mfma16(ld4(TA + (ob * 16 + r) * 16 + 4 * g), B, gw1[ob]);
    }
    __builtin_amdgcn_wave_barrier();
  }
  // bias gradients: sum over the 16 sample columns (the lanes with equal g)
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) {
        gb1[ob][v] += __shfl_xor(gb1[ob][v], m, 16);
        gb2[ob][v] += __shfl_xor(gb2[ob][v], m, 16);
      }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) gb3[v] += __shfl_xor(gb3[v], m, 16);
  // the waves add their sums into red[] one after another (W^T tiles hold
  // [out][in]; the parameter block stores W as [in][out])
  for (int i = threadIdx.x; i < MLP_NP; i += 256) red[i] = 0.f;
  __syncthreads();
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[MLP_OW1 + r * MLP_H + ob * 16 + 4 * g + v] += gw1[ob][v];
#pragma unroll
      for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
          for (int v = 0; v < 4; ++v) red[MLP_OW2 + (ib * 16 + r) * MLP_H + ob * 16 + 4 * g + v] += gw2[ob][ib][v];
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[MLP_OW3 + (ib * 16 + r) * MLP_OUTP + 4 * g + v] += gw3[ib][v];
      if (r == 0) {
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            red[MLP_OB1 + ob * 16 + 4 * g + v] += gb1[ob][v];
            red[MLP_OB2 + ob * 16 + 4 * g + v] += gb2[ob][v];
          }
#pragma unroll
        for (int v = 0; v < 4; ++v) red[MLP_OB3 + 4 * g + v] += gb3[v];
      }
    }
    __syncthreads();
  }
  float* P = a.part + (long long)blockIdx.x * MLP_NP;
  for (int i = threadIdx.x; i < MLP_NP; i += 256) P[i] = red[i];
  sred[threadIdx.x] = lsum;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sred[threadIdx.x] += sred[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.lpart[blockIdx.x] = sred[0];
}

// Sum of the partial gradients (fixed order), then either the gradient out
// (mode 0) or one Adam step on the FP32 master parameters and their BF16
// copies (mode 1).  Padded rows/columns (inputs >= 13, outputs >= 2) stay zero.
// Thread 0 also adds the partial losses, times lscale, into *loss.
struct MlpAdamArgs {
  const float* part;
  const float* lpart;
  int nblk;
  int mode;
  float* grad;   // mode 0: the summed gradient
  float* p;      // FP32 master parameters
  float* m;
  float* v;
  unsigned short *w1t, *w2t, *w3t, *w2, *w3;
  float* loss;
  float lscale;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1/(1 - b1^t), c2 = 1/(1 - b2^t)
};

__device__ __forceinline__ bool mlp_param_live(int i) {
  if (i < MLP_OB1) return (i - MLP_OW1) / MLP_H < MLP_IN;
  if (i >= MLP_OW3 && i < MLP_OB3) return (i - MLP_OW3) % MLP_OUTP < MLP_OUT;
  if (i >= MLP_OB3) return (i - MLP_OB3) < MLP_OUT;
  return true;
}

// the BF16 copies of parameter i
__device__ __forceinline__ void mlp_store_bf16(const MlpAdamArgs& a, int i, float w) {
  const unsigned short b = f2bf(w);
  if (i < MLP_OB1) {
    const int in = (i - MLP_OW1) / MLP_H, out = (i - MLP_OW1) % MLP_H;
    a.w1t[out * MLP_INP + in] = b;
  } else if (i >= MLP_OW2 && i < MLP_OB2) {
    const int in = (i - MLP_OW2) / MLP_H, out = (i - MLP_OW2) % MLP_H;
    a.w2t[out * MLP_H + in] = b;
    a.w2[in * MLP_H + out] = b;
  } else if (i >= MLP_OW3 && i < MLP_OB3) {
    const int in = (i - MLP_OW3) / MLP_OUTP, out = (i - MLP_OW3) % MLP_OUTP;
    a.w3t[out * MLP_H + in] = b;
    a.w3[in * MLP_OUTP + out] = b;
  }
}

__global__ __launch_bounds__(256) void k_mlp_adam(MlpAdamArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0 && a.loss) {
    float s = 0.f;
    for (int b = 0; b < a.nblk; ++b) s += a.lpart[b];
    *a.loss += s * a.lscale;
  }
  if (i >= MLP_NP) return;
  float gr = 0.f;
  for (int b = 0; b < a.nblk; ++b) gr += a.part[(long long)b * MLP_NP + i];
  if (a.mode == 0) {
    a.grad[i] = gr;
    return;
  }
  float w = 0.f;
  if (mlp_param_live(i)) {
    const float m = a.b1 * a.m[i] + (1.f - a.b1) * gr;
    const float v = a.b2 * a.v[i] + (1.f - a.b2) * gr * gr;
    a.m[i] = m;
    a.v[i] = v;
    w = a.p[i] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
  }
  a.p[i] = w;
  mlp_store_bf16(a, i, w);
}

// the BF16 copies of the whole parameter block (after an upload)
__global__ __launch_bounds__(256) void k_mlp_cast(MlpAdamArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= MLP_NP) return;
  const float w = mlp_param_live(i) ? a.p[i] : 0.f;
  a.p[i] = w;
  mlp_store_bf16(a, i, w);
}

// inference on a batch of raw feature rows: Y [n][2] (weights held in registers
// across the wave's tiles)
__global__ __launch_bounds__(256) void k_mlp_predict(MlpW w, const float* X, long long n, float* Y) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const long long ntile = (n + 15) / 16;
  MlpReg m;
  mlp_load(w, m);
  for (long long t = (long long)blockIdx.x * 4 + wave; t < ntile; t += (long long)gridDim.x * 4) {
    const long long s = t * 16 + r;
    const bool live = s < n;
    float x[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = live ? X[s * MLP_INP + 4 * g + v] : m.shift[v];
    const f32x4 y = mlp_infer(m, x);
    if (g == 0 && live) {
      Y[s * 2] = y[0];
      Y[s * 2 + 1] = y[1];
    }
  }
}

}  // namespace dgpu

namespace {

#define MCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return set_error(10, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// offsets of the BF16 weight copies (in elements; each a multiple of 4, so
// every 4-element operand load is 8-byte aligned)
constexpr int WB_W1T = 0, WB_W2T = WB_W1T + 64 * 16, WB_W3T = WB_W2T + 64 * 64, WB_W2 = WB_W3T + 16 * 64,
              WB_W3 = WB_W2 + 64 * 64, WB_N = WB_W3 + 64 * 16;
constexpr int MAX_BLK = 512;

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct dgpu_mlp {
  hipStream_t st = nullptr;
  float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
  unsigned short* wb = nullptr;
  float *part = nullptr, *lpart = nullptr, *loss = nullptr;
  float *X = nullptr, *T = nullptr;
  int* idx = nullptr;
  long long n = 0;
  int loss_cap = 0;
  long long t = 0;  // Adam step count
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;

  MlpW weights() const {
    MlpW w;
    w.w1t = wb + WB_W1T; w.w2t = wb + WB_W2T; w.w3t = wb + WB_W3T; w.w2 = wb + WB_W2; w.w3 = wb + WB_W3;
    w.b1 = p + MLP_OB1; w.b2 = p + MLP_OB2; w.b3 = p + MLP_OB3;
    w.norm = norm;
    return w;
  }
  MlpAdamArgs adam_args() {
    MlpAdamArgs a{};
    a.part = part; a.lpart = lpart; a.nblk = 1; a.mode = 1;
    a.grad = grad; a.p = p; a.m = m; a.v = v;
    a.w1t = wb + WB_W1T; a.w2t = wb + WB_W2T; a.w3t = wb + WB_W3T; a.w2 = wb + WB_W2; a.w3 = wb + WB_W3;
    a.loss = nullptr; a.lscale = 0.f;
    a.b1 = b1; a.b2 = b2; a.eps = eps;
    return a;
  }
};

namespace {

int upload_params(dgpu_mlp* m, const float* params) {
  MCHK(hipMemcpyAsync(m->p, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice, m->st));
  MCHK(hipMemsetAsync(m->m, 0, sizeof(float) * MLP_NP, m->st));
  MCHK(hipMemsetAsync(m->v, 0, sizeof(float) * MLP_NP, m->st));
  m->t = 0;
  MlpAdamArgs a = m->adam_args();
  k_mlp_cast<<<(MLP_NP + 255) / 256, 256, 0, m->st>>>(a);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

// Glorot-uniform weights on the live rows/columns, zero biases, the output
// layer scaled by 0.1
std::vector<float> init_params(uint64_t seed) {
  std::vector<float> p(MLP_NP, 0.f);
  uint64_t s = seed ^ 0x5eed5eed12345678ull;
  auto fill = [&](int off, int nin, int nout, int live_in, int live_out, double gain) {
    const double lim = gain * std::sqrt(6.0 / (double)(live_in + live_out));
    for (int i = 0; i < live_in; ++i)
      for (int o = 0; o < live_out; ++o) p[off + i * nout + o] = (float)((2.0 * uniform01(s) - 1.0) * lim);
    (void)nin;
  };
  fill(MLP_OW1, MLP_INP, MLP_H, MLP_IN, MLP_H, 1.0);
  fill(MLP_OW2, MLP_H, MLP_H, MLP_H, MLP_H, 1.0);
  fill(MLP_OW3, MLP_H, MLP_OUTP, MLP_H, MLP_OUT, 0.1);
  return p;
}

int grad_blocks(long long n) {
  const long long tiles = (n + 15) / 16;
  long long b = (tiles + 3) / 4;
  return (int)(b < 1 ? 1 : (b > MAX_BLK ? MAX_BLK : b));
}

int launch_grad(dgpu_mlp* m, const int* idx, long long n, const float loss_w[2], int& nblk) {
  MlpGradArgs a{};
  a.w = m->weights();
  a.X = m->X; a.T = m->T; a.idx = idx; a.n = n;
  a.wo0 = loss_w ? loss_w[0] : 1.f;
  a.wo1 = loss_w ? loss_w[1] : 1.f;
  a.part = m->part; a.lpart = m->lpart;
  nblk = grad_blocks(n);
  k_mlp_grad<<<nblk, 256, 0, m->st>>>(a);
  MCHK(hipGetLastError());
  return 0;
}

int launch_adam(dgpu_mlp* m, int nblk, float lr, float* loss, float lscale) {
  MlpAdamArgs a = m->adam_args();
  a.nblk = nblk;
  a.mode = 1;
  a.loss = loss;
  a.lscale = lscale;
  m->t += 1;
  a.lr = lr;
  a.c1 = (float)(1.0 / (1.0 - std::pow((double)m->b1, (double)m->t)));
  a.c2 = (float)(1.0 / (1.0 - std::pow((double)m->b2, (double)m->t)));
  k_mlp_adam<<<(MLP_NP + 255) / 256, 256, 0, m->st>>>(a);
  MCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int dgpu_mlp_create(const float* params, uint64_t seed, dgpu_mlp** out) {
  if (!out) return set_error(1, "null output pointer");
  *out = nullptr;
  dgpu_mlp* m = new dgpu_mlp();
  auto bail = [&](int rc) {
    dgpu_mlp_destroy(m);
    return rc;
  };
#define MALLOC(p, bytes)                                                   \
  do {                                                                     \
    hipError_t e_ = hipMalloc((void**)&(p), (bytes));                      \
    if (e_ != hipSuccess) return bail(set_error(10, std::string("hipMalloc: ") + hipGetErrorString(e_))); \
  } while (0)
  if (hipStreamCreate(&m->st) != hipSuccess) return bail(set_error(10, "hipStreamCreate failed"));
  MALLOC(m->p, sizeof(float) * MLP_NP);
  MALLOC(m->m, sizeof(float) * MLP_NP);
  MALLOC(m->v, sizeof(float) * MLP_NP);
  MALLOC(m->grad, sizeof(float) * MLP_NP);
  MALLOC(m->norm, sizeof(float) * MLP_NNORM);
  MALLOC(m->wb, sizeof(unsigned short) * WB_N);
  MALLOC(m->part, sizeof(float) * (size_t)MAX_BLK * MLP_NP);
  MALLOC(m->lpart, sizeof(float) * MAX_BLK);
#undef MALLOC
  std::vector<float> nrm(MLP_NNORM, 0.f);
  for (int f = 0; f < MLP_INP; ++f) {
    nrm[MLP_INP + f] = f < MLP_IN ? 1.f : 0.f;
    nrm[2 * MLP_INP + f] = -INFINITY;
    nrm[3 * MLP_INP + f] = INFINITY;
  }
  int rc = dgpu_mlp_set_norm(m, nrm.data());
  if (rc) return bail(rc);
  if (params) {
    rc = upload_params(m, params);
  } else {
    const std::vector<float> p0 = init_params(seed);
    rc = upload_params(m, p0.data());
  }
  if (rc) return bail(rc);
  *out = m;
  return 0;
}

int dgpu_mlp_destroy(dgpu_mlp* m) {
  if (!m) return 0;
  if (m->st) (void)hipStreamSynchronize(m->st);
  for (void* q : {(void*)m->p, (void*)m->m, (void*)m->v, (void*)m->grad, (void*)m->norm, (void*)m->wb,
                  (void*)m->part, (void*)m->lpart, (void*)m->loss, (void*)m->X, (void*)m->T, (void*)m->idx})
    if (q) (void)hipFree(q);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
  return 0;
}

int dgpu_mlp_set_params(dgpu_mlp* m, const float* params) {
  if (!m || !params) return set_error(1, "null argument");
  return upload_params(m, params);
}

int dgpu_mlp_get_params(dgpu_mlp* m, float* params) {
  if (!m || !params) return set_error(1, "null argument");
  MCHK(hipMemcpyAsync(params, m->p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int dgpu_mlp_set_norm(dgpu_mlp* m, const float* norm) {
  if (!m || !norm) return set_error(1, "null argument");
  MCHK(hipMemcpyAsync(m->norm, norm, sizeof(float) * MLP_NNORM, hipMemcpyHostToDevice, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int dgpu_mlp_set_data(dgpu_mlp* m, const float* X, const float* T, int64_t n) {
  if (!m || !X || !T) return set_error(1, "null argument");
  if (n < 1 || n > 0x7fffffffLL) return set_error(2, "dgpu_mlp_set_data: n must be in [1, 2^31)");
  for (void* q : {(void*)m->X, (void*)m->T, (void*)m->idx})
    if (q) MCHK(hipFree(q));
  m->X = nullptr; m->T = nullptr; m->idx = nullptr; m->n = 0;
  MCHK(hipMalloc((void**)&m->X, sizeof(float) * MLP_INP * (size_t)n));
  MCHK(hipMalloc((void**)&m->T, sizeof(float) * 2 * (size_t)n));
  MCHK(hipMalloc((void**)&m->idx, sizeof(int) * (size_t)n));
  MCHK(hipMemcpyAsync(m->X, X, sizeof(float) * MLP_INP * (size_t)n, hipMemcpyHostToDevice, m->st));
  MCHK(hipMemcpyAsync(m->T, T, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, m->st));
  MCHK(hipStreamSynchronize(m->st));
  m->n = n;
  return 0;
}

int dgpu_mlp_grad(dgpu_mlp* m, int64_t n, const float loss_w[2], float* grad, float* loss) {
  if (!m || !grad) return set_error(1, "null argument");
  if (n < 1 || n > m->n) return set_error(2, "dgpu_mlp_grad: n must be in [1, samples set]");
  int nblk = 0, rc;
  if ((rc = launch_grad(m, nullptr, n, loss_w, nblk))) return rc;
  MlpAdamArgs a = m->adam_args();
  a.nblk = nblk;
  a.mode = 0;
  k_mlp_adam<<<(MLP_NP + 255) / 256, 256, 0, m->st>>>(a);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, m->st));
  std::vector<float> lp(nblk);
  MCHK(hipMemcpyAsync(lp.data(), m->lpart, sizeof(float) * nblk, hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  if (loss) {
    double s = 0.0;
    for (float x : lp) s += x;
    *loss = (float)(s / (double)n);
  }
  return 0;
}

int dgpu_mlp_adam_step(dgpu_mlp* m, const float* grad, float lr) {
  if (!m || !grad) return set_error(1, "null argument");
  // the given gradient as the single partial
  MCHK(hipMemcpyAsync(m->part, grad, sizeof(float) * MLP_NP, hipMemcpyHostToDevice, m->st));
  int rc = launch_adam(m, 1, lr, nullptr, 0.f);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int dgpu_mlp_train(dgpu_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, const float loss_w[2],
                   float* epoch_loss) {
  if (!m) return set_error(1, "null handle");
  if (m->n < 1) return set_error(2, "dgpu_mlp_train: no data (dgpu_mlp_set_data)");
  if (epochs < 1 || batch < 1) return set_error(2, "dgpu_mlp_train: epochs and batch must be >= 1");
  if (epochs > m->loss_cap) {
    if (m->loss) MCHK(hipFree(m->loss));
    m->loss = nullptr;
    MCHK(hipMalloc((void**)&m->loss, sizeof(float) * epochs));
    m->loss_cap = epochs;
  }
  MCHK(hipMemsetAsync(m->loss, 0, sizeof(float) * epochs, m->st));
  const long long n = m->n;
  std::vector<int> perm((size_t)n);
  uint64_t s = seed;
  for (int e = 0; e < epochs; ++e) {
    // Fisher-Yates permutation of the samples for this epoch
    for (long long i = 0; i < n; ++i) perm[(size_t)i] = (int)i;
    for (long long i = n - 1; i > 0; --i) {
      const long long j = (long long)(splitmix(s) % (uint64_t)(i + 1));
      std::swap(perm[(size_t)i], perm[(size_t)j]);
    }
    MCHK(hipMemcpyAsync(m->idx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->st));
    for (long long off = 0; off < n; off += batch) {
      const long long nb = std::min<long long>(batch, n - off);
      int nblk = 0, rc;
      if ((rc = launch_grad(m, m->idx + off, nb, loss_w, nblk))) return rc;
      if ((rc = launch_adam(m, nblk, lr, m->loss + e, (float)(1.0 / (double)n)))) return rc;
    }
    // the host buffer is rewritten next epoch: the copy must have completed
    MCHK(hipStreamSynchronize(m->st));
  }
  if (epoch_loss) MCHK(hipMemcpyAsync(epoch_loss, m->loss, sizeof(float) * epochs, hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int dgpu_mlp_predict(dgpu_mlp* m, const float* X, int64_t n, float* Y) {
  if (!m || !X || !Y) return set_error(1, "null argument");
  if (n < 1) return set_error(2, "dgpu_mlp_predict: n must be >= 1");
  float *dX = nullptr, *dY = nullptr;
  MCHK(hipMalloc((void**)&dX, sizeof(float) * MLP_INP * (size_t)n));
  if (hipMalloc((void**)&dY, sizeof(float) * 2 * (size_t)n) != hipSuccess) {
    (void)hipFree(dX);
    return set_error(10, "hipMalloc failed");
  }
  int rc = 0;
  if (hipMemcpyAsync(dX, X, sizeof(float) * MLP_INP * (size_t)n, hipMemcpyHostToDevice, m->st) != hipSuccess) {
    rc = set_error(10, "upload failed");
  } else {
    const long long tiles = (n + 15) / 16;
    long long nb = (tiles + 3) / 4;
    if (nb > 4096) nb = 4096;
    k_mlp_predict<<<(int)nb, 256, 0, m->st>>>(m->weights(), dX, n, dY);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(Y, dY, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
      rc = set_error(10, "dgpu_mlp_predict: kernel or download failed");
  }
  (void)hipFree(dX);
  (void)hipFree(dY);
  return rc;
}

}  // extern "C"
