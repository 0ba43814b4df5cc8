// dgpu_mlp.hpp — the learned SGS closure network: an MLP 13 -> 64 -> 64 -> 2
// with GELU activations (tanh form), BF16 operands on the MFMA units and FP32 accumulation
// (v_mfma_f32_16x16x16_bf16), with training (forward, backward, Adam) and
// inference.  It has no counterpart in the reference: it is a further
// les_model_type beside the AMD closure of les_chan (for/les.f90), which stays
// the fallback outside the network's training envelope.
//
// Orientation: a layer computes Out^T (neurons x samples) = W^T (out x in) *
// In^T (in x samples) + b, so the weights are the A operand and the activations
// the B operand.  The 16x16 accumulator a lane holds (rows 4*(lane/16)+v,
// column lane%16) is then exactly the B-operand layout of the next layer's
// 16-row K block: activations stay in registers from layer to layer.  The
// weight gradients reduce over samples, so their operands pass through LDS.
// One wave works on a tile of 16 samples.  The training and batch-inference
// kernels are in dgpu_mlp.hip; k_closure (dgpu_closure.hpp) runs mlp_forward
// inside the LES step.
#pragma once
#include <hip/hip_runtime.h>

namespace dgpu {

constexpr int MLP_IN = 13, MLP_INP = 16, MLP_H = 64, MLP_OUT = 2, MLP_OUTP = 16;
// parameter block (FP32 master copy), in this order:
//   W1 [in 16][out 64], b1 [64], W2 [64][64], b2 [64], W3 [64][out 16], b3 [16]
// (rows in >= 13 of W1 and columns out >= 2 of W3, b3 are held at zero)
constexpr int MLP_OW1 = 0, MLP_OB1 = MLP_OW1 + MLP_INP * MLP_H, MLP_OW2 = MLP_OB1 + MLP_H,
              MLP_OB2 = MLP_OW2 + MLP_H * MLP_H, MLP_OW3 = MLP_OB2 + MLP_H,
              MLP_OB3 = MLP_OW3 + MLP_H * MLP_OUTP, MLP_NP = MLP_OB3 + MLP_OUTP;
// input normaliser: shift[16], scale[16] (x' = (x - shift) * scale), and the
// training envelope lo[16], hi[16] of the raw features
constexpr int MLP_NNORM = 4 * MLP_INP;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU in its tanh form (Hendrycks & Gimpel, 2016):
//   gelu(x) = x (1 + tanh u)/2 = x s,  s = 1/(1 + exp(-2u)),  u = sqrt(2/pi) (x + 0.044715 x^3)
// and its derivative s + x s (1 - s) 2 du/dx
constexpr float GELU_K0 = 0.79788456080286536f, GELU_K1 = 0.044715f;
// exp(-2u) = exp2(x (G1 + G3 x^2)) with G1 = -2 K0 log2(e), G3 = G1 K1
constexpr float GELU_G1 = (float)(-2.0 * 0.79788456080286536 * 1.4426950408889634);
constexpr float GELU_G3 = (float)(-2.0 * 0.79788456080286536 * 1.4426950408889634 * 0.044715);
__device__ __forceinline__ float gelu_s(float x) {
  const float z = x * __builtin_fmaf(GELU_G3, x * x, GELU_G1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float dgelu(float x) {
  const float s = gelu_s(x);
  return s + x * s * (1.0f - s) * (2.0f * GELU_K0 * (1.0f + 3.0f * GELU_K1 * x * x));
}

// BF16 weights in the layouts the passes read: Wt = [out][in] (forward A
// operand), W = [in][out] (backward A operand); biases FP32; the normaliser
struct MlpW {
  const unsigned short* w1t;  // [64][16]
  const unsigned short* w2t;  // [64][64]
  const unsigned short* w3t;  // [16][64]
  const unsigned short* w2;   // [64][64]
  const unsigned short* w3;   // [64][16]
  const float* b1;
  const float* b2;
  const float* b3;
  const float* norm;          // [MLP_NNORM]
};

__device__ __forceinline__ bf16x4 ld4(const unsigned short* p) {
  const uint2 u = *(const uint2*)p;  // 8-byte aligned at every call site
  bf16x4 r;
  r[0] = (short)(u.x & 0xffffu); r[1] = (short)(u.x >> 16);
  r[2] = (short)(u.y & 0xffffu); r[3] = (short)(u.y >> 16);
  return r;
}
typedef __bf16 bfv2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// four values as two pairs, each one gfx950 v_cvt_pk_bf16_f32 (round to nearest
// even, the same result as f2bf)
__device__ __forceinline__ bf16x4 tobf(f32x4 v) {
  const bfv2 lo = __builtin_convertvector(f32x2{v[0], v[1]}, bfv2);
  const bfv2 hi = __builtin_convertvector(f32x2{v[2], v[3]}, bfv2);
  const unsigned long long ul = __builtin_bit_cast(unsigned int, lo), uh = __builtin_bit_cast(unsigned int, hi);
  return __builtin_bit_cast(bf16x4, ul | (uh << 32));
}

// activations of one 16-sample tile in the accumulator layout
struct MlpAct {
  f32x4 z1[4], z2[4];   // pre-activations (FP32)
  bf16x4 h1[4], h2[4];  // GELU outputs (BF16, the next layer's B operand)
  f32x4 y;              // outputs: lanes 0..15 hold outputs 0, 1 of sample lane
};

// xb: the normalised input tile in B layout (lane l holds features 4*(l/16)+v
// of sample l%16)
__device__ __forceinline__ void mlp_forward(const MlpW& w, bf16x4 xb, MlpAct& a) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 c;
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = w.b1[ob * 16 + 4 * g + v];
    c = mfma16(ld4(w.w1t + (ob * 16 + r) * MLP_INP + 4 * g), xb, c);
    a.z1[ob] = c;
    f32x4 h;
#pragma unroll
    for (int v = 0; v < 4; ++v) h[v] = gelu(c[v]);
    a.h1[ob] = tobf(h);
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 c;
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = w.b2[ob * 16 + 4 * g + v];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) c = mfma16(ld4(w.w2t + (ob * 16 + r) * MLP_H + kb * 16 + 4 * g), a.h1[kb], c);
    a.z2[ob] = c;
    f32x4 h;
#pragma unroll
    for (int v = 0; v < 4; ++v) h[v] = gelu(c[v]);
    a.h2[ob] = tobf(h);
  }
  f32x4 c;
#pragma unroll
  for (int v = 0; v < 4; ++v) c[v] = w.b3[4 * g + v];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) c = mfma16(ld4(w.w3t + r * MLP_H + kb * 16 + 4 * g), a.h2[kb], c);
  a.y = c;
}

// The forward weights' operand fragments and the biases and normaliser entries
// one lane needs, held in registers across the tiles of a wave (inference).
struct MlpReg {
  bf16x4 w1[4], w2[4][4], w3[4];
  f32x4 b1[4], b2[4], b3;
  float shift[4], scale[4];
};

__device__ __forceinline__ void mlp_load(const MlpW& w, MlpReg& m) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    m.w1[ob] = ld4(w.w1t + (ob * 16 + r) * MLP_INP + 4 * g);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) m.w2[ob][kb] = ld4(w.w2t + (ob * 16 + r) * MLP_H + kb * 16 + 4 * g);
    m.w3[ob] = ld4(w.w3t + r * MLP_H + ob * 16 + 4 * g);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      m.b1[ob][v] = w.b1[ob * 16 + 4 * g + v];
      m.b2[ob][v] = w.b2[ob * 16 + 4 * g + v];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    m.b3[v] = w.b3[4 * g + v];
    m.shift[v] = w.norm[4 * g + v];
    m.scale[v] = w.norm[MLP_INP + 4 * g + v];
  }
}

// the outputs of one tile (lanes 0..15: outputs 0, 1 of sample lane) from
// register-resident weights; x holds the lane's 4 raw features 4g..4g+3
__device__ __forceinline__ f32x4 mlp_infer(const MlpReg& m, const float* x) {
  f32x4 xn;
#pragma unroll
  for (int v = 0; v < 4; ++v) xn[v] = (x[v] - m.shift[v]) * m.scale[v];
  const bf16x4 xb = tobf(xn);
  bf16x4 h1[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    const f32x4 c = mfma16(m.w1[ob], xb, m.b1[ob]);
    f32x4 h;
#pragma unroll
    for (int v = 0; v < 4; ++v) h[v] = gelu(c[v]);
    h1[ob] = tobf(h);
  }
  bf16x4 h2[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 c = m.b2[ob];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) c = mfma16(m.w2[ob][kb], h1[kb], c);
    f32x4 h;
#pragma unroll
    for (int v = 0; v < 4; ++v) h[v] = gelu(c[v]);
    h2[ob] = tobf(h);
  }
  f32x4 y = m.b3;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) y = mfma16(m.w3[kb], h2[kb], y);
  return y;
}

// features 4g..4g+3 of one raw sample (p points at feature 4g), normalised and
// rounded to BF16
__device__ __forceinline__ bf16x4 norm_x(const float* norm, const float* p, int g) {
  bf16x4 r;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int f = 4 * g + v;
    r[v] = (short)f2bf((p[v] - norm[f]) * norm[MLP_INP + f]);
  }
  return r;
}

}  // namespace dgpu
