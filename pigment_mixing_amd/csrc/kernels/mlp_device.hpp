// mlp_device.hpp — the learned return-map closure's network on MFMA.
//
// MLP 8 -> 64 -> 64 -> 64 -> 64 -> 4, SiLU on the hidden layers, BF16 weights
// and activations, FP32 accumulation and biases.  One wavefront evaluates 64
// samples at once (row = sample) with v_mfma_f32_16x16x16_bf16:
//   A operand (16x16, M x K): lane l holds A[l % 16][4 (l / 16) + i], i = 0..3
//   B operand (16x16, K x N): lane l holds B[4 (l / 16) + i][l % 16]
//   C/D       (16x16, M x N): lane l holds D[4 (l / 16) + i][l % 16]
// Weights are stored W[out][in] row-major, which is exactly the B operand
// layout of Z = A W^T (four consecutive `in` values per lane).  Rows are padded
// (MLP_S0 / MLP_S bf16) so that the 16 rows a wavefront reads at once fall in
// different LDS banks.
//
// Rounding points (the CPU oracle of the network rounds at the same places):
// normalised inputs -> BF16; weights -> BF16; each hidden layer's
// SiLU(Z + b) -> BF16; the output layer stays FP32.
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

// Per-wavefront activation buffer: 64 rows x MLP_S bf16 (8704 B).
constexpr int MLP_ACT_BYTES = 64 * MLP_S * 2;

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  // round to nearest even (inputs here are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ mlp_bf16x4 ld_bf16x4(const uint16_t* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  return __builtin_bit_cast(mlp_bf16x4, v);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ mlp_f32x4 mfma16(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
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

// One dense layer of the wavefront's 64 samples, in place in `act`:
// act[row][0 .. 16 KB) (bf16, stride MLP_S) -> act[row][0 .. 16 NB).  KB k-blocks
// in, NB n-blocks out, W with row stride WS.  HIDDEN: SiLU(Z + b) -> bf16 into
// act; otherwise Z + b (FP32) goes to out4[row * ostride + 0..3] (the last
// layer; out4 may alias act with ostride = MLP_S / 2: row r's outputs then
// overwrite row r's own, already consumed, inputs).
// Row block rb reads only rows 16 rb .. 16 rb + 15 and is written after all its
// reads (a wavefront's LDS operations complete in order), so in place is safe.
// Optionally stores the pre-activations Z (FP32, [64][64]) for training.
template <int KB, int NB, bool HIDDEN>
__device__ __forceinline__ void mlp_layer(uint16_t* act, const uint16_t* W, int WS, const float* bias,
                                          float* zsave, float* out4, int ostride) {
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, g = lane >> 4;
#pragma unroll 1
  for (int rb = 0; rb < 4; ++rb) {
    mlp_f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const mlp_bf16x4 a = ld_bf16x4(act + (rb * 16 + r16) * MLP_S + kb * 16 + 4 * g);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const mlp_bf16x4 b = ld_bf16x4(W + (nb * 16 + r16) * WS + kb * 16 + 4 * g);
        acc[nb] = mfma16(a, b, acc[nb]);
      }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int col = nb * 16 + r16;
      const float bc = bias[col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = rb * 16 + 4 * g + i;
        const float z = acc[nb][i] + bc;
        if constexpr (HIDDEN) {
          if (zsave) zsave[row * 64 + col] = z;
          act[row * MLP_S + col] = f32_to_bf16(silu(z));
        } else {
          if (col < MLP_OUT) out4[row * ostride + col] = z;
        }
      }
    }
  }
}

// Forward pass of the wavefront's 64 samples.  On entry act[row][0..15] holds
// the normalised inputs (bf16, cols 8..15 zero).  Normalised outputs go to
// out4[row * ostride + 0..3] (FP32).  zsave (optional, FP32 [4][64][64])
// receives the hidden layers' pre-activations.
__device__ __forceinline__ void mlp_wave_forward(const MlpNet& net, uint16_t* act, float* out4, int ostride,
                                                 float* zsave) {
  mlp_layer<1, 4, true>(act, net.W[0], MLP_S0, net.b[0], zsave, nullptr, 0);
  mlp_layer<4, 4, true>(act, net.W[1], MLP_S, net.b[1], zsave ? zsave + 4096 : nullptr, nullptr, 0);
  mlp_layer<4, 4, true>(act, net.W[2], MLP_S, net.b[2], zsave ? zsave + 8192 : nullptr, nullptr, 0);
  mlp_layer<4, 4, true>(act, net.W[3], MLP_S, net.b[3], zsave ? zsave + 12288 : nullptr, nullptr, 0);
  mlp_layer<4, 1, false>(act, net.W[4], MLP_S, net.b[4], nullptr, out4, ostride);
}

}  // namespace mpmx
