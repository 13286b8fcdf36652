// mlp_device.hpp — the learned return-map closure's network on MFMA.
//
// MLP 8 -> 64 -> 64 -> 64 -> 64 -> 4, SiLU on the hidden layers, BF16 weights
// and activations, FP32 accumulation and biases.  One wavefront evaluates 64
// samples at once (row = sample).  The 64-wide products run on gfx950's
// v_mfma_f32_16x16x32_bf16 (K = 32 per instruction); the first layer (8
// inputs padded to 16) on v_mfma_f32_16x16x16_bf16:
//   A operand (16 x K, M x K): lane l holds A[l % 16][k(l / 16, i)]
//   B operand (K x 16, K x N): lane l holds B[k(l / 16, i)][l % 16]
//   C/D       (16x16, M x N): lane l holds D[4 (l / 16) + i][l % 16]
// with i = 0..3 (K = 16) or 0..7 (K = 32).  The product sums A[m][k] B[k][n]
// over the K slots, so any assignment of logical k to (lane group, slot) gives
// the same result as long as A and B use the same one.  For K = 32 this code
// assigns slot i of group g to k = 4 g + i (i < 4) and k = 16 + 4 g + (i - 4)
// (i >= 4): each operand is then two of the K = 16 fragments (k-blocks 2m and
// 2m + 1) side by side, so weights, activations and the D -> B hand-over below
// keep their K = 16 layout.
// Weights are stored W[out][in] row-major, which is exactly the B operand
// layout of Z = A W^T (four consecutive `in` values per lane and fragment).
// Rows are padded (MLP_S0 / MLP_S bf16) so that the 16 rows a wavefront reads
// at once fall in different LDS banks.
//
// Rounding points (the CPU oracle of the network rounds at the same places):
// normalised inputs -> BF16; weights -> BF16; each hidden layer's
// SiLU(Z + b) -> BF16; the output layer stays FP32.
//
// Inference (mlp_wave_infer) runs the transposed product Z^T = W A^T instead:
// W is then the A operand (same loads), the activations the B operand, and the
// D tile of output block ob / sample block sb holds, per lane, four consecutive
// outputs of one sample -- exactly the B operand of k-block ob of the next
// layer.  The activations therefore stay in registers from layer to layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmx {

constexpr int MLP_IN = 8, MLP_HID = 64, MLP_OUT = 4;
constexpr int MLP_S0 = 20;   // bf16 row stride of W0 [64][16] (8 inputs padded to one k-block)
constexpr int MLP_S = 68;    // bf16 row stride of 64-wide rows (weights and activations)

// Device blob (bytes).  BF16 weights W0 [64][S0], W1..W3 [64][S], W4 [16][S]
// (4 outputs padded to one n-block, pad rows zero), FP32 biases b0..b3 [64] and
// b4 [16], then FP32 in_mean[8] in_std[8] out_mean[4] out_std[4].
constexpr int MLP_W0_OFF = 0;
constexpr int MLP_W1_OFF = MLP_W0_OFF + 64 * MLP_S0 * 2;
constexpr int MLP_W2_OFF = MLP_W1_OFF + 64 * MLP_S * 2;
constexpr int MLP_W3_OFF = MLP_W2_OFF + 64 * MLP_S * 2;
constexpr int MLP_W4_OFF = MLP_W3_OFF + 64 * MLP_S * 2;
constexpr int MLP_B_OFF = MLP_W4_OFF + 16 * MLP_S * 2;
constexpr int MLP_NORM_OFF = MLP_B_OFF + (4 * 64 + 16) * 4;
constexpr int MLP_BLOB_BYTES = MLP_NORM_OFF + 24 * 4;
static_assert(MLP_W1_OFF % 16 == 0 && MLP_W4_OFF % 16 == 0 && MLP_B_OFF % 16 == 0, "blob alignment");

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_bf16x2v __attribute__((ext_vector_type(2)));
typedef short mlp_s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  // round to nearest even (inputs here are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// two floats -> packed BF16 pair (v_cvt_pk_bf16_f32: round to nearest even), a in the low half
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
  const mlp_f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, mlp_bf16x2v));
}

__device__ __forceinline__ mlp_bf16x4 ld_bf16x4(const uint16_t* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  return __builtin_bit_cast(mlp_bf16x4, v);
}

// logistic sigmoid 1 / (1 + e^-z) from the hardware base-2 exponential and
// reciprocal (v_exp_f32, v_rcp_f32; about 1 ulp each): e^-z = 2^(-z log2 e).
// Saturates to 0 / 1 for large |z| (exp2 -> inf / 0).
__device__ __forceinline__ float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ mlp_f32x4 mfma16(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// K = 32: operands are the K = 16 fragments of k-blocks 2m (lo) and 2m + 1 (hi)
__device__ __forceinline__ mlp_f32x4 mfma32(mlp_bf16x4 a_lo, mlp_bf16x4 a_hi, mlp_bf16x4 b_lo, mlp_bf16x4 b_hi,
                                            mlp_f32x4 c) {
  const mlp_s16x8 a = __builtin_shufflevector(a_lo, a_hi, 0, 1, 2, 3, 4, 5, 6, 7);
  const mlp_s16x8 b = __builtin_shufflevector(b_lo, b_hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16x8, a), __builtin_bit_cast(mlp_bf16x8, b),
                                                 c, 0, 0, 0);
}
// Z += A B over K = 64 from LDS rows: a_row / b_row point at the lane's row of
// the A (M x K) and B^T (N x K) tiles, bf16, consecutive k.
__device__ __forceinline__ mlp_f32x4 mfma_k64(const uint16_t* a_row, const uint16_t* b_row, mlp_f32x4 c) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int m = 0; m < 2; ++m)
    c = mfma32(ld_bf16x4(a_row + 32 * m + 4 * g), ld_bf16x4(a_row + 32 * m + 16 + 4 * g),
               ld_bf16x4(b_row + 32 * m + 4 * g), ld_bf16x4(b_row + 32 * m + 16 + 4 * g), c);
  return c;
}

// The network blob (MLP_BLOB_BYTES; src and dst 16-byte aligned) into LDS by
// the NT threads of a workgroup: 16-byte loads, eight in flight per thread
// before their stores (a load-store loop waits on every load in turn).  No
// branches: the indices past the blob are clamped to its last word, which the
// threads concerned copy again with the same value (a branch would let the
// compiler sink each load into its store's block, one wait per load again).
template <int NT>
__device__ __forceinline__ void mlp_blob_to_lds(const uint8_t* __restrict__ src, uint8_t* dst) {
  static_assert(MLP_BLOB_BYTES % 16 == 0, "blob in whole 16-byte words");
  constexpr int NV = MLP_BLOB_BYTES / 16;
  constexpr int PER = (NV + NT - 1) / NT;
  constexpr int CH = PER < 8 ? PER : 8;
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* t = reinterpret_cast<uint4*>(dst);
  for (int c0 = 0; c0 < PER; c0 += CH) {
    uint4 v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) v[u] = s[min((int)threadIdx.x + NT * (c0 + u), NV - 1)];
#pragma unroll
    for (int u = 0; u < CH; ++u) t[min((int)threadIdx.x + NT * (c0 + u), NV - 1)] = v[u];
  }
}

struct MlpNet {
  const uint16_t* W[5];
  const float* b[5];
  const float* norm;   // in_mean[8] in_std[8] out_mean[4] out_std[4]
};

__device__ __forceinline__ MlpNet mlp_net(const uint8_t* blob) {
  MlpNet n;
  n.W[0] = reinterpret_cast<const uint16_t*>(blob + MLP_W0_OFF);
  n.W[1] = reinterpret_cast<const uint16_t*>(blob + MLP_W1_OFF);
  n.W[2] = reinterpret_cast<const uint16_t*>(blob + MLP_W2_OFF);
  n.W[3] = reinterpret_cast<const uint16_t*>(blob + MLP_W3_OFF);
  n.W[4] = reinterpret_cast<const uint16_t*>(blob + MLP_W4_OFF);
  const float* b = reinterpret_cast<const float*>(blob + MLP_B_OFF);
  n.b[0] = b; n.b[1] = b + 64; n.b[2] = b + 128; n.b[3] = b + 192; n.b[4] = b + 256;
  n.norm = reinterpret_cast<const float*>(blob + MLP_NORM_OFF);
  return n;
}

// ---------------------------------------------------------------- inference
// One layer of Z^T = W A^T + b for the wavefront's 4 sample blocks of 16:
// act[sb][kb] is the B fragment (bf16 x 4: samples sb*16 + lane%16, inputs
// kb*16 + 4 (lane/16) .. +3).  Hidden layers overwrite act[sb][ob] with
// BF16(SiLU(Z + b)); the output layer (OB = 1) leaves Z + b of outputs
// 4 (lane/16) .. +3 in out[sb].
template <int KB, int OB, bool HIDDEN>
__device__ __forceinline__ void mlp_layer_t(const uint16_t* W, int WS, const float* bias, mlp_bf16x4 (&act)[4][4],
                                            mlp_f32x4* out) {
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, g = lane >> 4;
  mlp_bf16x4 wf[OB][KB];
#pragma unroll
  for (int ob = 0; ob < OB; ++ob)
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) wf[ob][kb] = ld_bf16x4(W + (ob * 16 + r16) * WS + kb * 16 + 4 * g);
  mlp_f32x4 bz[OB];
#pragma unroll
  for (int ob = 0; ob < OB; ++ob) bz[ob] = *reinterpret_cast<const mlp_f32x4*>(bias + ob * 16 + 4 * g);
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) {
    mlp_f32x4 acc[OB];
#pragma unroll
    for (int ob = 0; ob < OB; ++ob) acc[ob] = bz[ob];   // Z = b + sum_k W A
    if constexpr (KB == 4) {
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int ob = 0; ob < OB; ++ob)
          acc[ob] = mfma32(wf[ob][2 * m], wf[ob][2 * m + 1], act[sb][2 * m], act[sb][2 * m + 1], acc[ob]);
    } else {
      static_assert(KB == 1, "64-wide (K = 32 pairs) or one 16-wide k-block");
#pragma unroll
      for (int ob = 0; ob < OB; ++ob) acc[ob] = mfma16(wf[ob][0], act[sb][0], acc[ob]);
    }
    if constexpr (HIDDEN) {
#pragma unroll
      for (int ob = 0; ob < OB; ++ob) {
        const mlp_f32x4 z = acc[ob];
        uint2 pk;
        pk.x = pack_bf16x2(silu(z[0]), silu(z[1]));
        pk.y = pack_bf16x2(silu(z[2]), silu(z[3]));
        act[sb][ob] = __builtin_bit_cast(mlp_bf16x4, pk);
      }
    } else {
      out[sb] = acc[0];
    }
  }
}

// Forward pass of the wavefront's 64 samples, activations in registers.
// stage: this wavefront's 64 x 16-byte LDS rows.  On entry row s holds sample
// s's 8 normalised inputs as BF16 (lane s wrote it; the caller fences).  On
// return y[0..3] are the normalised outputs of the lane's own sample (the rows
// have been reused to hand them back).
__device__ __forceinline__ void mlp_wave_infer(const MlpNet& net, uint4* stage, float y[4]) {
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, g = lane >> 4;
  mlp_bf16x4 act[4][4];
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) {
    const uint2 v = reinterpret_cast<const uint2*>(stage + sb * 16 + r16)[g & 1];
    act[sb][0] = g < 2 ? __builtin_bit_cast(mlp_bf16x4, v) : mlp_bf16x4{0, 0, 0, 0};
  }
  mlp_layer_t<1, 4, true>(net.W[0], MLP_S0, net.b[0], act, nullptr);
  mlp_layer_t<4, 4, true>(net.W[1], MLP_S, net.b[1], act, nullptr);
  mlp_layer_t<4, 4, true>(net.W[2], MLP_S, net.b[2], act, nullptr);
  mlp_layer_t<4, 4, true>(net.W[3], MLP_S, net.b[3], act, nullptr);
  mlp_f32x4 o[4];
  mlp_layer_t<4, 1, false>(net.W[4], MLP_S, net.b[4], act, o);
  // lanes 0..15 hold the 4 outputs of samples sb*16 + lane: hand them back
  // through the staging rows (all reads of the inputs are done)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  if (g == 0) {
#pragma unroll
    for (int sb = 0; sb < 4; ++sb)
      stage[sb * 16 + r16] = __builtin_bit_cast(uint4, o[sb]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const mlp_f32x4 r = __builtin_bit_cast(mlp_f32x4, stage[lane]);
  y[0] = r[0]; y[1] = r[1]; y[2] = r[2]; y[3] = r[3];
}

// Normalised inputs of the lane's sample -> its BF16 staging row.
__device__ __forceinline__ void mlp_stage_inputs(uint4* stage, const float xn[8]) {
  const int lane = threadIdx.x & 63;
  uint4 pk;
  pk.x = pack_bf16x2(xn[0], xn[1]);
  pk.y = pack_bf16x2(xn[2], xn[3]);
  pk.z = pack_bf16x2(xn[4], xn[5]);
  pk.w = pack_bf16x2(xn[6], xn[7]);
  stage[lane] = pk;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

}  // namespace mpmx
