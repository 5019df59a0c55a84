// wf_mlp.hpp — the return-map surrogate: a 4 -> 32 -> 32 -> 1 ReLU MLP
// evaluated per wavefront on BF16 MFMA (v_mfma_f32_32x32x8bf16_1k) with FP32
// accumulation.
//
// Parameters (kMlpParams floats): W1[in 4][out 32], b1[32], W2[in 32][out 32],
// b2[32], W3[32], b3.  Features: s_trial/A, 10 eps_p, log10(max(rate,1e-3))/3,
// homologous temperature; output: the plastic increment as a fraction of its
// bracket end dg_max = (s_trial - s_y(eps_p))/(3G).
//
// The network runs transposed, H^T = W^T X^T, so that a layer's output
// registers are already the next layer's B operand.  MFMA 32x32x8 operand
// layout (lane l, c = l % 32, half = l / 32):
//   A (32 x 8):  lane holds A[c][4*half + j], j = 0..3
//   B (8 x 32):  lane holds B[4*half + j][c]
//   D (32 x 32): lane holds D[8*(r/4) + 4*half + r%4][c], r = 0..15
// so layer-1 output registers 4kb..4kb+3 are the B operand of layer-2 K-block
// kb.  One pass evaluates 32 samples; a wavefront evaluates its 64 lanes'
// samples in two passes.
#pragma once
#include "wf_device.hpp"

namespace wf {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// round-to-nearest-even FP32 -> BF16, two values per v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned bf16_pair(float lo, float hi) {
  const f32x2_t v = {lo, hi};
  const bf16x2_t b = __builtin_convertvector(v, bf16x2_t);
  unsigned u;
  __builtin_memcpy(&u, &b, 4);
  return u;
}
// four values as an MFMA operand
__device__ __forceinline__ bf16x4_t bf16_quad(float a, float b, float c, float d) {
  const unsigned lo = bf16_pair(a, b), hi = bf16_pair(c, d);
  bf16x4_t q;
  q[0] = (short)(lo & 0xFFFFu);
  q[1] = (short)(lo >> 16);
  q[2] = (short)(hi & 0xFFFFu);
  q[3] = (short)(hi >> 16);
  return q;
}

struct MlpWave {
  bf16x4_t a1;     // W1^T with the bias as a fifth input
  bf16x4_t a2[4];  // W2^T, four K-blocks of 8
  float b2r[16];   // b2 at this lane's D rows
  float w3r[16];   // W3 at this lane's D rows
  float b3;
};

__device__ __forceinline__ void mlp_wave_load(const float* __restrict__ w, MlpWave& mw) {
  const int l = lane_id(), half = l >> 5, c = l & 31;
  float v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int k = 4 * half + j;
    v[j] = k < 4 ? w[kMlpOffW1 + k * 32 + c] : (k == 4 ? w[kMlpOffB1 + c] : 0.f);
  }
  mw.a1 = bf16_quad(v[0], v[1], v[2], v[3]);
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    const float* r = w + kMlpOffW2 + (8 * kb + 4 * half) * 32 + c;
    mw.a2[kb] = bf16_quad(r[0], r[32], r[64], r[96]);
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int row = 8 * (r >> 2) + 4 * half + (r & 3);
    mw.b2r[r] = w[kMlpOffB2 + row];
    mw.w3r[r] = w[kMlpOffW3 + row];
  }
  mw.b3 = w[kMlpOffB3];
}

// mlp_wave_load's registers of one lane, stored after each weight update so
// that a wavefront loads its operands as kMlpPackWords coalesced 16-byte
// words (word q of lane l at pack[64 q + l]) instead of 53 scattered weights
// a lane
struct MlpLanePack {
  bf16x4_t a1;
  bf16x4_t a2[4];
  float b3, pad;
  float b2r[16];
  float w3r[16];
};
static_assert(sizeof(MlpLanePack) == 16 * kMlpPackWords, "16-byte words");

__device__ __forceinline__ void mlp_wave_load_packed(const uint4* __restrict__ pack,
                                                     MlpWave& mw) {
  const int l = lane_id();
  uint4 q[kMlpPackWords];
#pragma unroll
  for (int k = 0; k < kMlpPackWords; k++) q[k] = pack[64 * k + l];
  MlpLanePack lp;
  __builtin_memcpy(&lp, q, sizeof(lp));
  mw.a1 = lp.a1;
#pragma unroll
  for (int kb = 0; kb < 4; kb++) mw.a2[kb] = lp.a2[kb];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    mw.b2r[r] = lp.b2r[r];
    mw.w3r[r] = lp.w3r[r];
  }
  mw.b3 = lp.b3;
}

// Every lane passes the features of its own sample; every lane gets its own
// sample's output.  All 64 lanes must execute this.
__device__ __forceinline__ float mlp_wave_eval(const MlpWave& mw, float f0, float f1, float f2,
                                               float f3) {
  const int half = lane_id() >> 5;
  const bf16x4_t one = {(short)0x3F80, 0, 0, 0};  // BF16 1.0
  float out0 = 0.f, out1 = 0.f;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    float g0 = f0, g1 = f1, g2 = f2, g3 = f3;
    if (pass == 1) {  // samples of lanes 32..63 enter through lanes 0..31
      g0 = __shfl_xor(f0, 32, 64);
      g1 = __shfl_xor(f1, 32, 64);
      g2 = __shfl_xor(f2, 32, 64);
      g3 = __shfl_xor(f3, 32, 64);
    }
    bf16x4_t b1 = bf16_quad(g0, g1, g2, g3);
    if (half) b1 = one;
    f32x16_t zero;
#pragma unroll
    for (int r = 0; r < 16; r++) zero[r] = 0.f;
    const f32x16_t d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(mw.a1, b1, zero, 0, 0, 0);
    f32x16_t d2;
#pragma unroll
    for (int r = 0; r < 16; r++) d2[r] = mw.b2r[r];
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const bf16x4_t b2 = bf16_quad(fmaxf(d1[4 * kb], 0.f), fmaxf(d1[4 * kb + 1], 0.f),
                                    fmaxf(d1[4 * kb + 2], 0.f), fmaxf(d1[4 * kb + 3], 0.f));
      d2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(mw.a2[kb], b2, d2, 0, 0, 0);
    }
    float part = 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) part += mw.w3r[r] * fmaxf(d2[r], 0.f);
    const float y = part + __shfl_xor(part, 32, 64) + mw.b3;
    if (pass == 0) out0 = y;
    else out1 = y;
  }
  return half ? out1 : out0;
}

// features of one yielding particle
__device__ __forceinline__ void mlp_features(double s_trial, double A, double ep, double rate,
                                             double theta, float& f0, float& f1, float& f2,
                                             float& f3) {
  f0 = (float)s_trial * __builtin_amdgcn_rcpf((float)A);
  f1 = (float)(10.0 * ep);
  f2 = log10f(fmaxf((float)rate, 1e-3f)) * (1.f / 3.f);
  f3 = (float)theta;
}

}  // namespace wf
