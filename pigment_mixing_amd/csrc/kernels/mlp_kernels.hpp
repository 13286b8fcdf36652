// mlp_kernels.hpp — launchers of the learned closure's training / evaluation
// kernels (network layout in mlp_device.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mlp_device.hpp"

namespace mpmx {

// Transposed BF16 weights for the backward pass, WT_l[in][out] for l = 1..4:
// rows of MLP_S (l = 1..3) or MLP_S0 (l = 4, 4 outputs padded to 16) elements.
constexpr int MLP_WT_OFF[5] = {0, 0, 64 * MLP_S, 2 * 64 * MLP_S, 3 * 64 * MLP_S};
#define MLP_WT_OFF_D(l) ((l) == 1 ? 0 : ((l) == 2 ? 64 * MLP_S : ((l) == 3 ? 2 * 64 * MLP_S : 3 * 64 * MLP_S)))
constexpr int MLP_WT_ELEMS = 3 * 64 * MLP_S + 64 * MLP_S0;

constexpr int MLP_N_PARAMS = (64 * 8 + 64) + 3 * (64 * 64 + 64) + (4 * 64 + 4);

// p[0..n) = 0 (a fill kernel: also valid where a caller captures the stream)
void launch_zero_words(uint32_t* p, int64_t n, hipStream_t st);
void launch_mlp_pack(const float* params, int n_params, uint8_t* blob, uint16_t* wt, const float* norm24,
                     hipStream_t st);
void launch_mlp_forward(const uint8_t* blob, const float* x8, float* y4, int64_t n, hipStream_t st);
void launch_mlp_grad(const uint8_t* blob, const uint16_t* wt, const float* xn, const float* tn, const int* idx,
                     int count, float inv_denom, float* grad, float* loss, float* zs, hipStream_t st);
void launch_mlp_adam(float* params, float* grad, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t st);
void launch_mlp_sample(int* idx, int batch, int64_t n, uint64_t seed, int64_t step, hipStream_t st);

}  // namespace mpmx
