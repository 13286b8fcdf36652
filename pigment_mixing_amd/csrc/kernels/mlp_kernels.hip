// mlp_kernels.hip — training and evaluation kernels of the learned return-map
// closure (network and MFMA conventions in mlp_device.hpp).
//
//   k_mlp_pack     FP32 master parameters -> BF16 blob (+ transposed copy for
//                  the backward pass) and FP32 biases / normalisation
//   k_mlp_forward  normalised outputs of n samples (inference check)
//   k_mlp_grad     forward + backward of 64 samples per wavefront: loss and
//                  parameter gradients (FP32 atomics into the gradient vector)
//   k_mlp_adam     Adam update of the FP32 master parameters
//
// Parameter vector layout (matches mpmx_closure_set_params): per layer
// W[out][in] then b[out], layers 8->64, 64->64 (x3), 64->4.
#include "mlp_kernels.hpp"

#include <algorithm>

namespace mpmx {

namespace {

__device__ __forceinline__ int layer_in(int l) { return l == 0 ? MLP_IN : MLP_HID; }
__device__ __forceinline__ int layer_out(int l) { return l == 4 ? MLP_OUT : MLP_HID; }
// offset of layer l's W in the parameter vector
__device__ __forceinline__ int param_off(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += layer_out(i) * layer_in(i) + layer_out(i);
  return o;
}
__device__ __forceinline__ int blob_w_off(int l) {
  return l == 0 ? MLP_W0_OFF : (l == 1 ? MLP_W1_OFF : (l == 2 ? MLP_W2_OFF : (l == 3 ? MLP_W3_OFF : MLP_W4_OFF)));
}
__device__ __forceinline__ int blob_w_stride(int l) { return l == 0 ? MLP_S0 : MLP_S; }

}  // namespace

// ---------------------------------------------------------------- packing
// One thread per parameter.  Padding entries of the blob stay zero (the blob
// is cleared before the first pack).  Transposed weights WT_l (l = 1..4) go to
// `wt`: WT_l[in][out] with row stride MLP_S (l = 1..3) or MLP_S0 (l = 4).
__global__ void k_mlp_pack(const float* __restrict__ params, int n_params, uint8_t* __restrict__ blob,
                           uint16_t* __restrict__ wt, const float* __restrict__ norm24) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < 24 && norm24) reinterpret_cast<float*>(blob + MLP_NORM_OFF)[p] = norm24[p];
  if (p >= n_params) return;
  int l = 0, o = p;
  while (l < 4 && o >= layer_out(l) * layer_in(l) + layer_out(l)) {
    o -= layer_out(l) * layer_in(l) + layer_out(l);
    ++l;
  }
  const int nin = layer_in(l), nout = layer_out(l);
  const float v = params[p];
  if (o < nout * nin) {
    const int r = o / nin, c = o % nin;
    const uint16_t b = f32_to_bf16(v);
    reinterpret_cast<uint16_t*>(blob + blob_w_off(l))[r * blob_w_stride(l) + c] = b;
    if (wt && l >= 1) wt[MLP_WT_OFF[l] + c * (l == 4 ? MLP_S0 : MLP_S) + r] = b;
  } else {
    reinterpret_cast<float*>(blob + MLP_B_OFF)[64 * l + (o - nout * nin)] = v;
  }
}

// ---------------------------------------------------------------- forward
// One wavefront per 64 samples; x8 are raw inputs (normalised here), y4 the
// raw outputs (denormalised).  Rows past n compute on zeros and are not stored.
__global__ void __launch_bounds__(64) k_mlp_forward(const uint8_t* __restrict__ blob, const float* __restrict__ x8,
                                                    float* __restrict__ y4, int64_t n) {
  __shared__ __align__(16) uint8_t net_lds[MLP_BLOB_BYTES];
  __shared__ uint4 stage[64];
  mlp_blob_to_lds<64>(blob, net_lds);
  const MlpNet net = mlp_net(net_lds);
  const int lane = threadIdx.x;
  const int64_t row = (int64_t)blockIdx.x * 64 + lane;
  __syncthreads();
  float xn[MLP_IN];
#pragma unroll
  for (int c = 0; c < MLP_IN; ++c) xn[c] = row < n ? (x8[row * MLP_IN + c] - net.norm[c]) / net.norm[8 + c] : 0.0f;
  mlp_stage_inputs(stage, xn);
  float y[MLP_OUT];
  mlp_wave_infer(net, stage, y);
  if (row < n) {
#pragma unroll
    for (int c = 0; c < MLP_OUT; ++c) y4[row * MLP_OUT + c] = y[c] * net.norm[20 + c] + net.norm[16 + c];
  }
}

// ---------------------------------------------------------------- gradient
// Training data are normalised FP32: xn [n][8], tn [n][4].  Sample s of
// wavefront w is idx[64 w + s] (or 64 w + s when idx is null); samples at or
// past `count` are masked (no loss, no gradient).  The loss is
// sum over samples and outputs of (y - t)^2 / (denom) with denom = 4 * count;
// gradients of that loss are added to grad (FP32 atomics).
//
// LDS: forward weights W_l (bf16, blob layout), transposed WT_l, biases; the
// activations' transposed copies AT_l [in][64] for every layer (dW operands),
// a row-major ping-pong pair for the forward, and the row-major / transposed
// gradient dZ.  Pre-activations Z go to global scratch (zs, 64 KB per wave).
struct GradLds {
  uint8_t w[MLP_B_OFF];
  uint16_t wt[MLP_WT_ELEMS];
  float bias[4 * 64 + 16];
  uint16_t at0[16 * MLP_S];          // AT_0: [16 in][64 samples]
  uint16_t at[4][64 * MLP_S];        // AT_1..4: [64 in][64 samples]
  uint16_t act[2][64 * MLP_S];       // row-major ping-pong
  uint16_t dz[64 * MLP_S];           // dZ row-major [sample][out]
  uint16_t dzt[64 * MLP_S];          // dZ transposed [out][sample]
  float y[64 * MLP_OUT];             // network outputs (normalised)
};

// dW_l += dZ^T A: M = out (NBo blocks), N = in (NBi blocks), K = 64 samples
template <int NBo, int NBi>
__device__ __forceinline__ void grad_w(const uint16_t* dzt, const uint16_t* at, float* gW, int nout, int nin) {
  const int lane = threadIdx.x & 63, r16 = lane & 15, g = lane >> 4;
#pragma unroll 1
  for (int mb = 0; mb < NBo; ++mb) {
#pragma unroll 1
    for (int nb = 0; nb < NBi; ++nb) {
      const mlp_f32x4 acc = mfma_k64(dzt + (mb * 16 + r16) * MLP_S, at + (nb * 16 + r16) * MLP_S,
                                     mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      const int k = nb * 16 + r16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = mb * 16 + 4 * g + i;
        if (o < nout && k < nin && acc[i] != 0.0f) atomicAdd(&gW[o * nin + k], acc[i]);
      }
    }
  }
}

// column sums over the 64 samples of a D tile's values (lanes with equal r16)
__device__ __forceinline__ float col_sum(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

__global__ void __launch_bounds__(64) k_mlp_grad(const uint8_t* __restrict__ blob, const uint16_t* __restrict__ wt,
                                                 const float* __restrict__ xn, const float* __restrict__ tn,
                                                 const int* __restrict__ idx, int count, float inv_denom,
                                                 float* __restrict__ grad, float* __restrict__ loss,
                                                 float* __restrict__ zs_all) {
  __shared__ __align__(16) GradLds S;
  const int lane = threadIdx.x, r16 = lane & 15, g = lane >> 4;
  for (int i = lane; i < MLP_B_OFF / 4; i += 64)
    reinterpret_cast<uint32_t*>(S.w)[i] = reinterpret_cast<const uint32_t*>(blob)[i];
  for (int i = lane; i < MLP_WT_ELEMS / 2; i += 64)
    reinterpret_cast<uint32_t*>(S.wt)[i] = reinterpret_cast<const uint32_t*>(wt)[i];
  for (int i = lane; i < 4 * 64 + 16; i += 64) S.bias[i] = reinterpret_cast<const float*>(blob + MLP_B_OFF)[i];
  const int s = blockIdx.x * 64 + lane;
  const bool live = s < count;
  const int64_t src = live ? (idx ? idx[s] : s) : 0;
  float* zs = zs_all + (size_t)blockIdx.x * 4 * 4096;
  // inputs: row-major act[0] and transposed at0
  for (int c = 0; c < 16; ++c) {
    const uint16_t b = f32_to_bf16(live && c < MLP_IN ? xn[src * MLP_IN + c] : 0.0f);
    S.act[0][lane * MLP_S + c] = b;
    S.at0[c * MLP_S + lane] = b;
  }
  __syncthreads();
  const uint16_t* W[5];
  for (int l = 0; l < 5; ++l) W[l] = reinterpret_cast<const uint16_t*>(S.w + blob_w_off(l));
  // ---- forward: hidden layers (row-major out + transposed copy + Z to scratch)
#pragma unroll 1
  for (int l = 0; l < 5; ++l) {
    const uint16_t* in = S.act[l & 1];
    uint16_t* out = S.act[(l + 1) & 1];
    const int KB = l == 0 ? 1 : 4;
    const int NB = l == 4 ? 1 : 4;
    const int WS = blob_w_stride(l);
#pragma unroll 1
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll 1
      for (int nb = 0; nb < NB; ++nb) {
        mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (KB == 4) {
          acc = mfma_k64(in + (rb * 16 + r16) * MLP_S, W[l] + (nb * 16 + r16) * WS, acc);
        } else {
          const mlp_bf16x4 a = ld_bf16x4(in + (rb * 16 + r16) * MLP_S + 4 * g);
          const mlp_bf16x4 b = ld_bf16x4(W[l] + (nb * 16 + r16) * WS + 4 * g);
          acc = mfma16(a, b, acc);
        }
        const int col = nb * 16 + r16;
        const float bc = S.bias[64 * l + col];
        if (l < 4) {
          uint16_t t4[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = rb * 16 + 4 * g + i;
            const float z = acc[i] + bc;
            zs[l * 4096 + row * 64 + col] = z;
            const uint16_t a16 = f32_to_bf16(silu(z));
            out[row * MLP_S + col] = a16;
            t4[i] = a16;
          }
          uint2 pk;
          pk.x = (uint32_t)t4[0] | ((uint32_t)t4[1] << 16);
          pk.y = (uint32_t)t4[2] | ((uint32_t)t4[3] << 16);
          *reinterpret_cast<uint2*>(&S.at[l][col * MLP_S + rb * 16 + 4 * g]) = pk;
        } else if (r16 < MLP_OUT) {
#pragma unroll
          for (int i = 0; i < 4; ++i) S.y[(rb * 16 + 4 * g + i) * MLP_OUT + r16] = acc[i] + bc;
        }
      }
    }
    __syncthreads();
  }
  // ---- output gradient dY = 2 (y - t) / denom; loss
  float lsum = 0.0f;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
    uint16_t t4[4];
    float bsum = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = rb * 16 + 4 * g + i;
      const int srow = blockIdx.x * 64 + row;
      float d = 0.0f;
      if (r16 < MLP_OUT && srow < count) {
        const int64_t sr = idx ? idx[srow] : srow;
        const float e = S.y[row * MLP_OUT + r16] - tn[sr * MLP_OUT + r16];
        lsum += e * e;
        d = 2.0f * e * inv_denom;
      }
      bsum += d;
      const uint16_t d16 = f32_to_bf16(d);
      S.dz[row * MLP_S + r16] = d16;
      t4[i] = d16;
    }
    uint2 pk;
    pk.x = (uint32_t)t4[0] | ((uint32_t)t4[1] << 16);
    pk.y = (uint32_t)t4[2] | ((uint32_t)t4[3] << 16);
    *reinterpret_cast<uint2*>(&S.dzt[r16 * MLP_S + rb * 16 + 4 * g]) = pk;
    bsum = col_sum(bsum);
    if (g == 0 && r16 < MLP_OUT && bsum != 0.0f) atomicAdd(&grad[param_off(4) + MLP_OUT * MLP_HID + r16], bsum);
  }
  // wave sum of the loss
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0 && lsum != 0.0f) atomicAdd(loss, lsum * inv_denom);
  __syncthreads();
  // ---- backward through layers 4 .. 0
#pragma unroll 1
  for (int l = 4; l >= 0; --l) {
    const int nout = layer_out(l), nin = layer_in(l);
    float* gW = grad + param_off(l);
    if (l == 4) grad_w<1, 4>(S.dzt, S.at[3], gW, nout, nin);
    else if (l == 0) grad_w<4, 1>(S.dzt, S.at0, gW, nout, nin);
    else grad_w<4, 4>(S.dzt, S.at[l - 1], gW, nout, nin);
    __syncthreads();
    if (l == 0) break;
    // dA_{l-1} = dZ_l W_l (M = samples, K = out, N = in), then dZ_{l-1} = dA * silu'(Z_{l-1})
    const int KB = l == 4 ? 1 : 4;
    const uint16_t* wtl = S.wt + MLP_WT_OFF_D(l);
    const int WTS = l == 4 ? MLP_S0 : MLP_S;
    float* gb = grad + param_off(l - 1) + MLP_HID * layer_in(l - 1);
#pragma unroll 1
    for (int rb = 0; rb < 4; ++rb) {
      mlp_f32x4 acc[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) acc[nb] = mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if (KB == 4) {
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[nb] = mfma_k64(S.dz + (rb * 16 + r16) * MLP_S, wtl + (nb * 16 + r16) * WTS,
                                                          acc[nb]);
      } else {
        const mlp_bf16x4 a = ld_bf16x4(S.dz + (rb * 16 + r16) * MLP_S + 4 * g);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
          const mlp_bf16x4 b = ld_bf16x4(wtl + (nb * 16 + r16) * WTS + 4 * g);
          acc[nb] = mfma16(a, b, acc[nb]);
        }
      }
      // rows rb of dz are fully read: overwrite them with dZ_{l-1}
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const int col = nb * 16 + r16;
        uint16_t t4[4];
        float bsum = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = rb * 16 + 4 * g + i;
          const float z = zs[(l - 1) * 4096 + row * 64 + col];
          const bool rl = blockIdx.x * 64 + row < count;
          const float dz = rl ? acc[nb][i] * silu_grad(z) : 0.0f;
          bsum += dz;
          const uint16_t d16 = f32_to_bf16(dz);
          S.dz[row * MLP_S + col] = d16;
          t4[i] = d16;
        }
        uint2 pk;
        pk.x = (uint32_t)t4[0] | ((uint32_t)t4[1] << 16);
        pk.y = (uint32_t)t4[2] | ((uint32_t)t4[3] << 16);
        *reinterpret_cast<uint2*>(&S.dzt[col * MLP_S + rb * 16 + 4 * g]) = pk;
        bsum = col_sum(bsum);
        if (g == 0 && bsum != 0.0f) atomicAdd(&gb[col], bsum);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- Adam
__global__ void k_mlp_adam(float* __restrict__ params, float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1,
                           float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.0f - b1) * g;
  const float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
  grad[p] = 0.0f;
}

// ---------------------------------------------------------------- minibatch
// idx[i] = hash(seed, step, i) mod n (counter-based, no state)
__global__ void k_mlp_sample(int* __restrict__ idx, int batch, int64_t n, uint64_t seed, int64_t step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)step * 0xD1B54A32D192ED03ull + (uint64_t)i;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[i] = (int)(z % (uint64_t)n);
}

// ---------------------------------------------------------------- launchers
__global__ void k_zero_words(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}
void launch_zero_words(uint32_t* p, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0,
                                st, p, n);
}
void launch_mlp_pack(const float* params, int n_params, uint8_t* blob, uint16_t* wt, const float* norm24,
                     hipStream_t st) {
  const int n = n_params > 24 ? n_params : 24;
  hipLaunchKernelGGL(k_mlp_pack, dim3((n + 255) / 256), dim3(256), 0, st, params, n_params, blob, wt, norm24);
}
void launch_mlp_forward(const uint8_t* blob, const float* x8, float* y4, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, blob, x8, y4, n);
}
size_t mlp_grad_lds_bytes() { return sizeof(GradLds); }
void launch_mlp_grad(const uint8_t* blob, const uint16_t* wt, const float* xn, const float* tn, const int* idx,
                     int count, float inv_denom, float* grad, float* loss, float* zs, hipStream_t st) {
  if (count > 0)
    hipLaunchKernelGGL(k_mlp_grad, dim3((count + 63) / 64), dim3(64), 0, st, blob, wt, xn, tn, idx,
                       count, inv_denom, grad, loss, zs);
}
void launch_mlp_adam(float* params, float* grad, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((n + 255) / 256), dim3(256), 0, st, params, grad, m, v, n, lr, b1, b2, eps,
                     bc1, bc2);
}
void launch_mlp_sample(int* idx, int batch, int64_t n, uint64_t seed, int64_t step, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_sample, dim3((batch + 255) / 256), dim3(256), 0, st, idx, batch, n, seed, step);
}

}  // namespace mpmx
