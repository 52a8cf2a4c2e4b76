// Device side of the learned Kastaun bracket (include/pangu_mlp.h): BF16
// helpers, the gfx950 matrix-core product, and the network evaluated by one
// wavefront for its 64 cells inside the recovery kernel.
//
// v_mfma_f32_32x32x16_bf16 (one wavefront, C 32x32 += A 32x16 B 16x32):
// lane l holds row (l mod 32) of A and column (l mod 32) of B for the eight
// k of its half (8 (l / 32) + 0..7), and receives C rows
// 8 (v / 4) + 4 (l / 32) + (v mod 4) of column (l mod 32) in accumulator v.
//
// In-wave evaluation.  Activations are carried transposed, H^T [feature][cell],
// so that a layer is C^T = W^T H^T with the weights as the A operand and the
// 64 cells of the wave as two 32-column tiles g of the B operand.  Lane l
// then receives, for its column's cell, the output features
// f(ot, v, h) = 32 ot + 8 (v / 4) + 4 h + (v mod 4)    (h = l / 32),
// and those registers are fed unchanged as the B operand of the next layer:
// k-step s, slot q of half h stands for input feature
// phi(s, h, q) = f(s / 2, 8 (s mod 2) + q, h).  Any fixed ordering of the
// summation index is the same product, so the weights are packed in that
// order (pack_fragments) and no activation ever leaves its lane: no LDS, no
// cross-lane moves.  Layer 1 (5 inputs): for tile g the inputs sit in the
// k slots of half g, filled by the lanes that own those cells (lane = cell
// within the wave); layer 4 (one output) is a dot product on the vector ALU
// plus one exchange between lanes l and l ^ 32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace pangu {

constexpr int MLP_NH = 64;    // hidden width
constexpr int MLP_KIN = 32;   // padded input width of the training layout (5 features)
constexpr int MLP_NOUT = 32;  // padded output width of the training layout (1 output)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(unsigned short)))) unsigned short u16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Round to nearest even BF16 with the hardware conversion (gfx950
// v_cvt_pk_bf16_f32): the same bits as f2bf for every finite value, with
// no branch.
__device__ inline uint16_t f2bf_hw(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ inline f32x16 mfma32(u16x8 a, u16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Training-path SiLU (IEEE division, libm exp).
__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
// Inference-path SiLU: hardware exp2 and reciprocal (about 1 ulp each), far
// below the BF16 rounding the activation then takes.
__device__ inline float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// Output feature of accumulator v of output tile ot in lane half h.
__host__ __device__ constexpr int frag_feature(int ot, int v, int h) {
  return 32 * ot + 8 * (v >> 2) + 4 * h + (v & 3);
}

// Weights in fragment order (built by k_pack_frag from the BF16 copies):
// a1 [g 2][ot 2][lane 64][8], a2 and a3 [ot 2][s 4][lane 64][8]; per lane
// half h the values its accumulators need, in accumulator order
// (16 ot + v): bias[layer 0..2][h][32] and the output weights
// w4[h][32] (BF16-rounded, as FP32); b4 the output bias.
struct MlpFrag {
  const u16x8 *a1, *a2, *a3;
  const float *bias, *w4, *b4;
  const float *mean, *dev;
  int kind;  // 0: Kastaun bracket (logit mu); 1: 1Dvsq start (log u^2)
};

// The fragments staged in LDS by the workgroup before its wavefronts run
// the network: matrix-core operands then come from LDS (tens of cycles)
// rather than from L2 for every product.
struct MlpLds {
  u16x8 a1[2 * 2 * 64];
  u16x8 a2[2 * 4 * 64];
  u16x8 a3[2 * 4 * 64];
  float fb[4 * 64];
};

// Cooperative copy of the fragments into LDS by the nthreads threads of a
// workgroup (thread tid); the caller synchronises afterwards.
__device__ __forceinline__ void mlp_stage(const MlpFrag &W, MlpLds &L, int tid, int nthreads) {
  for (int e = tid; e < 2 * 2 * 64; e += nthreads) L.a1[e] = glb(W.a1)[e];
  for (int e = tid; e < 2 * 4 * 64; e += nthreads) {
    L.a2[e] = glb(W.a2)[e];
    L.a3[e] = glb(W.a3)[e];
  }
  for (int e = tid; e < 4 * 64; e += nthreads) L.fb[e] = glb(W.bias)[e];
}

// One hidden layer of tile g: B operands in[4] (k-steps) -> out[4] (the
// next layer's B operands), out = BF16(silu(W^T in + b)), with bias the
// lane half's 32 values.  With w4 (the lane half's 32 output weights), the
// layer is the last hidden one: out is not kept, and the lane's share of
// the output layer, sum of out * w4 over its 32 features, is returned.
__device__ __forceinline__ float wave_layer(const u16x8 *A, const float *bias, const u16x8 in[4],
                                            u16x8 out[4], const float *w4, int lane) {
  float p = 0.0f;
#pragma unroll
  for (int ot = 0; ot < 2; ++ot) {
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma32(A[(ot * 4 + s) * 64 + lane], in[s], acc);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const uint16_t a = f2bf_hw(silu_fast(acc[v] + bias[16 * ot + v]));
      if (w4)
        p += bf2f(a) * w4[16 * ot + v];
      else
        out[2 * ot + (v >> 3)][v & 7] = a;
    }
  }
  return p;
}

// Network output z (before the logistic) for this lane's cell, from its
// five standardised, BF16-rounded features x, with the fragments staged in
// L (mlp_stage).  Every lane of the wavefront must call it (the matrix
// products span all 64 lanes).
__device__ __forceinline__ float wave_mlp(const MlpLds &L, float b4, const float x[5], int lane) {
  const int h = lane >> 5;
  const float *bias = L.fb + 32 * h, *w4 = L.fb + 3 * 64 + 32 * h;
  float part[2];
#pragma unroll 1
  for (int g = 0; g < 2; ++g) {
    // keeps the fragment loads inside the loop: hoisted out of it, the
    // loop-invariant fragments of all four layers would stay live at once
    // (a compiler barrier only; no instruction is emitted)
    __asm__ volatile("" ::: "memory");
    u16x8 bx;
#pragma unroll
    for (int q = 0; q < 8; ++q) bx[q] = (h == g && q < 5) ? f2bf_hw(x[q < 5 ? q : 0]) : (uint16_t)0;
    u16x8 h1[4], h2[4];
#pragma unroll
    for (int ot = 0; ot < 2; ++ot) {
      f32x16 acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
      acc = mfma32(L.a1[(g * 2 + ot) * 64 + lane], bx, acc);
#pragma unroll
      for (int v = 0; v < 16; ++v) h1[2 * ot + (v >> 3)][v & 7] = f2bf_hw(silu_fast(acc[v] + bias[16 * ot + v]));
    }
    (void)wave_layer(L.a2, bias + 64, h1, h2, nullptr, lane);
    part[g] = wave_layer(L.a3, bias + 128, h2, h1, w4, lane);
  }
  // own cell: tile g = h; the other half of its features sits in lane l ^ 32
  const float mine = h ? part[1] : part[0];
  const float other = __shfl_xor(h ? part[0] : part[1], 32);
  return mine + other + b4;
}

}  // namespace pangu
