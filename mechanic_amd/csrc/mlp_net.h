// The learned warm-start network: 16 -> 64 -> 64 -> 3 with GELU (tanh form),
// its parameter layout and its MFMA inference for one group of 16 contacts.
//
// Inference uses v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32
// accumulation. Every layer is computed transposed (out^T = W^T in^T) so a
// layer's accumulator tile -- lane l holding rows 4*(l/16)..+3 of column l%16 --
// is already the B operand of the next layer and no data moves between lanes:
//   A operand: lane l holds A[l%16][4*(l/16)+i],  i = 0..3
//   B operand: lane l holds B[4*(l/16)+i][l%16]
//   D (C)    : lane l holds D[4*(l/16)+i][l%16]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mh {

// Parameter layout (floats): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3]
// (W[k][h] maps input k to output h), then the input normalisation (mean[16],
// 1/std[16]) and the output scaling (mean[3], std[3]): y = net(x) * std + mean.
constexpr int kIn = 16, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut, kParams = kOffB3 + kOut;
constexpr int kOffInMean = kParams, kOffInScale = kOffInMean + kIn, kOffOutMean = kOffInScale + kIn,
              kOffOutScale = kOffOutMean + kOut, kNetFloats = kOffOutScale + kOut;
static_assert(kParams == 5443, "16-64-64-3 network");
static_assert(kNetFloats == 5481, "parameters plus normalisation");

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

// gelu(x) = 0.5 x (1 + tanh(u)) = x * sigmoid(2u), u = sqrt(2/pi) (x + 0.044715 x^3).
__device__ __forceinline__ float gelu_exact(float x) {
  const float u = kGeluC * (x + kGeluA * x * x * x);
  return x / (1.0f + expf(-2.0f * u));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float u = kGeluC * (x + kGeluA * x * x * x);
  const float s = 1.0f / (1.0f + expf(-2.0f * u));
  const float du = kGeluC * (1.0f + 3.0f * kGeluA * x * x);
  return s + x * s * (1.0f - s) * 2.0f * du;
}
// Inference form: hardware exp and reciprocal (the result is rounded to BF16).
__device__ __forceinline__ float gelu_fast(float x) {
  const float u = kGeluC * (x + kGeluA * x * x * x);
  return x * __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// The network in MFMA operand layout, per lane.
struct MlpRegs {
  v4s w1[4];   // W1^T, per hidden tile
  v4s w2[16];  // W2^T, per (output tile, k chunk)
  v4s w3[4];   // W3^T with outputs padded to 16, per k chunk
  float b1[4][4], b2[4][4], b3[4];
  float in_mean[4], in_scale[4], out_mean[4], out_scale[4];
};

__device__ __forceinline__ void mlp_load(const float* net, uint32_t lane, MlpRegs& m) {
  const int r = lane % 16, q = lane / 16;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      m.w1[t][i] = to_bf16(net[kOffW1 + (4 * q + i) * kH + t * 16 + r]);
      m.w3[t][i] = r < kOut ? to_bf16(net[kOffW3 + (t * 16 + 4 * q + i) * kOut + r]) : (short)0;
      m.b1[t][i] = net[kOffB1 + t * 16 + 4 * q + i];
      m.b2[t][i] = net[kOffB2 + t * 16 + 4 * q + i];
    }
#pragma unroll
    for (int kc = 0; kc < 4; ++kc)
#pragma unroll
      for (int i = 0; i < 4; ++i) m.w2[t * 4 + kc][i] = to_bf16(net[kOffW2 + (kc * 16 + 4 * q + i) * kH + t * 16 + r]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int f = 4 * q + i;
    m.in_mean[i] = net[kOffInMean + f];
    m.in_scale[i] = net[kOffInScale + f];
    m.b3[i] = f < kOut ? net[kOffB3 + f] : 0.0f;
    m.out_mean[i] = f < kOut ? net[kOffOutMean + f] : 0.0f;
    m.out_scale[i] = f < kOut ? net[kOffOutScale + f] : 0.0f;
  }
}

// One group of 16 contacts. Lane l passes features 4*(l/16)..+3 of contact
// l%16 in x; lanes 0..15 receive that contact's three outputs in y[0..2].
__device__ __forceinline__ void mlp_group(const MlpRegs& m, const float x[4], float y[3]) {
  v4s xb;
#pragma unroll
  for (int i = 0; i < 4; ++i) xb[i] = to_bf16((x[i] - m.in_mean[i]) * m.in_scale[i]);
  v4s h1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m.w1[t], xb, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) h1[t][i] = to_bf16(gelu_fast(acc[i] + m.b1[t][i]));
  }
  v4s h2[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 4; ++kc)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m.w2[t * 4 + kc], h1[kc], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) h2[t][i] = to_bf16(gelu_fast(acc[i] + m.b2[t][i]));
  }
  v4f out = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m.w3[kc], h2[kc], out, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = (out[i] + m.b3[i]) * m.out_scale[i] + m.out_mean[i];
}

}  // namespace mh
