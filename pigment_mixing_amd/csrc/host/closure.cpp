// closure.cpp — learned return-map closure: parameter packing, the device
// trainer (Adam over harvested pairs) and handle-free forward / gradient entry
// points used to check the kernels against the CPU oracle of the network.
#include "closure.hpp"

#include "../../../include/mpmx.h"
#include "../kernels/mlp_kernels.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

namespace mpmx {

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

#define CK(x)                                          \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) {                            \
      err = std::string(#x) + ": " + hipGetErrorString(e_); \
      return MPMX_ERR_HIP;                             \
    }                                                  \
  } while (0)

int dmalloc(DevBuf& b, size_t bytes, hipStream_t s, std::string& err) {
  CK(hipMalloc(&b.p, bytes > 0 ? bytes : 1));
  CK(hipMemsetAsync(b.p, 0, bytes > 0 ? bytes : 1, s));
  return MPMX_OK;
}

// norm24 = in_mean[8] in_std[8] out_mean[4] out_std[4]
void fill_norm24(float* n24, const float* im, const float* is, const float* om, const float* os) {
  std::memcpy(n24, im, 32);
  std::memcpy(n24 + 8, is, 32);
  std::memcpy(n24 + 16, om, 16);
  std::memcpy(n24 + 20, os, 16);
}

// Pack FP32 parameters + normalisation into a zeroed device blob (and the
// transposed weights when wt != null).
int pack_to_device(const float* params, const float* norm24, uint8_t* blob, uint16_t* wt, hipStream_t s,
                   std::string& err) {
  DevBuf dp, dn;
  int rc;
  if ((rc = dmalloc(dp, sizeof(float) * MLP_N_PARAMS, s, err))) return rc;
  if ((rc = dmalloc(dn, sizeof(float) * 24, s, err))) return rc;
  CK(hipMemsetAsync(blob, 0, MLP_BLOB_BYTES, s));
  if (wt) CK(hipMemsetAsync(wt, 0, sizeof(uint16_t) * MLP_WT_ELEMS, s));
  CK(hipMemcpyAsync(dp.p, params, sizeof(float) * MLP_N_PARAMS, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dn.p, norm24, sizeof(float) * 24, hipMemcpyHostToDevice, s));
  launch_mlp_pack(dp.as<float>(), MLP_N_PARAMS, blob, wt, dn.as<float>(), s);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(s));
  return MPMX_OK;
}

// Normalised training data on the device: xn = (x - in_mean)/in_std, tn likewise.
int upload_normalised(const float* x8, const float* t4, int64_t n, const float* n24, DevBuf& xn, DevBuf& tn,
                      hipStream_t s, std::string& err) {
  std::vector<float> hx((size_t)n * 8), ht((size_t)n * 4);
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 8; ++c) hx[i * 8 + c] = (x8[i * 8 + c] - n24[c]) / n24[8 + c];
    for (int c = 0; c < 4; ++c) ht[i * 4 + c] = (t4[i * 4 + c] - n24[16 + c]) / n24[20 + c];
  }
  int rc;
  if ((rc = dmalloc(xn, sizeof(float) * hx.size(), s, err))) return rc;
  if ((rc = dmalloc(tn, sizeof(float) * ht.size(), s, err))) return rc;
  CK(hipMemcpyAsync(xn.p, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(tn.p, ht.data(), sizeof(float) * ht.size(), hipMemcpyHostToDevice, s));
  return MPMX_OK;
}

bool norm_ok(const float* n24) {
  for (int i = 0; i < 8; ++i)
    if (!(n24[8 + i] > 0) || !std::isfinite(n24[i])) return false;
  for (int i = 0; i < 4; ++i)
    if (!(n24[20 + i] > 0) || !std::isfinite(n24[16 + i])) return false;
  return true;
}

// Split harvest records (12 floats: 8 inputs, h1' h2' h3' dgamma) into network
// inputs and residual-form targets (h' - h_trial, dgamma).
void records_to_xt(const float* rec, int64_t n, std::vector<float>& x, std::vector<float>& t) {
  x.resize((size_t)n * 8);
  t.resize((size_t)n * 4);
  for (int64_t i = 0; i < n; ++i) {
    const float* r = rec + 12 * i;
    for (int c = 0; c < 8; ++c) x[i * 8 + c] = r[c];
    for (int c = 0; c < 3; ++c) t[i * 4 + c] = r[8 + c] - r[c];
    t[i * 4 + 3] = r[11];
  }
}

// mean / std per column (std of a constant column is 1, so it normalises to 0)
void column_stats(const std::vector<float>& a, int64_t n, int w, float* mean, float* sd) {
  for (int c = 0; c < w; ++c) {
    double s = 0, s2 = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i * w + c];
    const double m = n ? s / n : 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double d = a[i * w + c] - m;
      s2 += d * d;
    }
    const double v = n > 1 ? s2 / (n - 1) : 0.0;
    mean[c] = (float)m;
    sd[c] = v > 1e-30 * (m * m + 1e-30) ? (float)std::sqrt(v) : 1.0f;
  }
}

}  // namespace

int closure_upload(ClosureState& c, const std::vector<float>& params, const float* in_mean, const float* in_std,
                   const float* out_mean, const float* out_std, hipStream_t s, std::string& err) {
  if ((int64_t)params.size() != MLP_N_PARAMS) {
    err = "closure: wrong parameter count";
    return MPMX_ERR_ARG;
  }
  if (!c.d_blob) {
    CK(hipMalloc(&c.d_blob, MLP_BLOB_BYTES));
    c.blob_bytes = MLP_BLOB_BYTES;
  }
  float n24[24];
  fill_norm24(n24, in_mean, in_std, out_mean, out_std);
  return pack_to_device(params.data(), n24, static_cast<uint8_t*>(c.d_blob), nullptr, s, err);
}

void closure_free(ClosureState& c) {
  if (c.d_blob) (void)hipFree(c.d_blob);
  c.d_blob = nullptr;
  c.blob_bytes = 0;
  c.attached = 0;
}

// Xavier-uniform weights, zero biases (seeded)
void closure_init_params(std::vector<float>& p, uint64_t seed) {
  p.assign(MLP_N_PARAMS, 0.0f);
  std::mt19937_64 rng(seed);
  const int dims[6] = {MLP_IN, MLP_HID, MLP_HID, MLP_HID, MLP_HID, MLP_OUT};
  size_t k = 0;
  for (int l = 0; l < 5; ++l) {
    const int nin = dims[l], nout = dims[l + 1];
    const float a = std::sqrt(6.0f / (float)(nin + nout));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < nout * nin; ++i) p[k++] = u(rng);
    k += nout;
  }
}

int closure_train(std::vector<float>& params, float* n24_io, const float* records12, int64_t n,
                  const mpmx_train_config& tc, mpmx_train_report* rep, float* curve, int64_t curve_cap,
                  hipStream_t s, std::string& err) {
  if (n < 64 || tc.batch < 64 || tc.batch % 64 || tc.steps <= 0 || !(tc.lr > 0)) {
    err = "train: need >= 64 records, batch a positive multiple of 64, steps > 0, lr > 0";
    return MPMX_ERR_ARG;
  }
  if (n > INT32_MAX) {
    err = "train: too many records";
    return MPMX_ERR_CAPACITY;
  }
  std::vector<float> x, t;
  records_to_xt(records12, n, x, t);
  float n24[24];
  column_stats(x, n, 8, n24, n24 + 8);
  column_stats(t, n, 4, n24 + 16, n24 + 20);
  if (tc.init || (int64_t)params.size() != MLP_N_PARAMS) closure_init_params(params, tc.seed);
  DevBuf xn, tn, dpar, dgrad, dm, dv, blob, wt, idx, zs, loss;
  int rc;
  if ((rc = upload_normalised(x.data(), t.data(), n, n24, xn, tn, s, err))) return rc;
  if ((rc = dmalloc(dpar, sizeof(float) * MLP_N_PARAMS, s, err))) return rc;
  if ((rc = dmalloc(dgrad, sizeof(float) * MLP_N_PARAMS, s, err))) return rc;
  if ((rc = dmalloc(dm, sizeof(float) * MLP_N_PARAMS, s, err))) return rc;
  if ((rc = dmalloc(dv, sizeof(float) * MLP_N_PARAMS, s, err))) return rc;
  if ((rc = dmalloc(blob, MLP_BLOB_BYTES, s, err))) return rc;
  if ((rc = dmalloc(wt, sizeof(uint16_t) * MLP_WT_ELEMS, s, err))) return rc;
  if ((rc = dmalloc(idx, sizeof(int) * tc.batch, s, err))) return rc;
  const int64_t waves_full = (n + 63) / 64;
  const int64_t zs_waves = std::max<int64_t>(tc.batch / 64, waves_full);
  if ((rc = dmalloc(zs, sizeof(float) * 4 * 4096 * (size_t)zs_waves, s, err))) return rc;
  if ((rc = dmalloc(loss, sizeof(float), s, err))) return rc;
  CK(hipMemcpyAsync(dpar.p, params.data(), sizeof(float) * MLP_N_PARAMS, hipMemcpyHostToDevice, s));
  DevBuf dn;
  if ((rc = dmalloc(dn, sizeof(float) * 24, s, err))) return rc;
  CK(hipMemcpyAsync(dn.p, n24, sizeof(float) * 24, hipMemcpyHostToDevice, s));
  const float b1 = tc.beta1 > 0 ? tc.beta1 : 0.9f, b2 = tc.beta2 > 0 ? tc.beta2 : 0.999f;
  const float eps = tc.eps > 0 ? tc.eps : 1e-8f;
  const float lr_final = tc.lr_final > 0 ? tc.lr_final : tc.lr;
  const int every = tc.report_every > 0 ? tc.report_every : 100;
  auto t0 = std::chrono::steady_clock::now();
  float first = -1, last = -1;
  int64_t ncurve = 0;
  for (int64_t step = 0; step < tc.steps; ++step) {
    launch_mlp_pack(dpar.as<float>(), MLP_N_PARAMS, blob.as<uint8_t>(), wt.as<uint16_t>(), dn.as<float>(), s);
    launch_mlp_sample(idx.as<int>(), tc.batch, n, tc.seed, step, s);
    CK(hipMemsetAsync(loss.p, 0, sizeof(float), s));
    launch_mlp_grad(blob.as<uint8_t>(), wt.as<uint16_t>(), xn.as<float>(), tn.as<float>(), idx.as<int>(), tc.batch,
                    1.0f / (4.0f * tc.batch), dgrad.as<float>(), loss.as<float>(), zs.as<float>(), s);
    // cosine decay from lr to lr_final
    const double prog = (double)step / (double)std::max<int64_t>(1, tc.steps - 1);
    const float lr = (float)(lr_final + 0.5 * (tc.lr - lr_final) * (1.0 + std::cos(3.14159265358979 * prog)));
    const float bc1 = 1.0f - std::pow(b1, (float)(step + 1)), bc2 = 1.0f - std::pow(b2, (float)(step + 1));
    launch_mlp_adam(dpar.as<float>(), dgrad.as<float>(), dm.as<float>(), dv.as<float>(), MLP_N_PARAMS, lr, b1, b2,
                    eps, bc1, bc2, s);
    if (step == 0 || (step + 1) % every == 0 || step == tc.steps - 1) {
      float l = 0;
      CK(hipMemcpyAsync(&l, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      if (step == 0) first = l;
      last = l;
      if (curve && ncurve < curve_cap) curve[ncurve++] = l;
    }
  }
  CK(hipGetLastError());
  // full-data loss with the final parameters
  launch_mlp_pack(dpar.as<float>(), MLP_N_PARAMS, blob.as<uint8_t>(), wt.as<uint16_t>(), dn.as<float>(), s);
  CK(hipMemsetAsync(loss.p, 0, sizeof(float), s));
  CK(hipMemsetAsync(dgrad.p, 0, sizeof(float) * MLP_N_PARAMS, s));
  launch_mlp_grad(blob.as<uint8_t>(), wt.as<uint16_t>(), xn.as<float>(), tn.as<float>(), nullptr, (int)n,
                  1.0f / (4.0f * (float)n), dgrad.as<float>(), loss.as<float>(), zs.as<float>(), s);
  float full = 0;
  CK(hipMemcpyAsync(&full, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(params.data(), dpar.p, sizeof(float) * MLP_N_PARAMS, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::memcpy(n24_io, n24, sizeof n24);
  if (rep) {
    rep->loss_first = first;
    rep->loss_last = last;
    rep->loss_full = full;
    rep->steps = tc.steps;
    rep->n_curve = ncurve;
    rep->ms = ms;
    std::memcpy(rep->norm, n24, sizeof n24);
  }
  return MPMX_OK;
}

}  // namespace mpmx

// ---- forward pass that allocates nothing: device memory from a fixed array.
//
// `scratch` holds, 256-byte aligned: the packed network blob, the FP32
// parameter and normalisation staging read by the pack kernel, and the input
// and output rows of up to MPMX_MLP_NOALLOC_MAX samples.  Each device holds its
// own copy; its address comes from hipGetSymbolAddress, looked up once per
// device with that device current, and only that address is handed to copies
// and kernels.  Calls on one device share the array: overlapping calls race.
namespace mpmx {
namespace {
constexpr size_t NA_ALIGN = 256;
constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
constexpr size_t NA_BLOB = 0;
constexpr size_t NA_PAR = NA_BLOB + na_round(MLP_BLOB_BYTES);
constexpr size_t NA_NORM = NA_PAR + na_round(sizeof(float) * MLP_N_PARAMS);
constexpr size_t NA_X = NA_NORM + na_round(sizeof(float) * 24);
constexpr size_t NA_Y = NA_X + na_round(sizeof(float) * 8 * (size_t)MPMX_MLP_NOALLOC_MAX);
constexpr size_t NA_BYTES = NA_Y + na_round(sizeof(float) * 4 * (size_t)MPMX_MLP_NOALLOC_MAX);
constexpr int NA_MAX_DEVICES = 64;

__device__ __align__(256) uint8_t scratch[NA_BYTES];

std::mutex na_mutex;
uint8_t* na_ptr[NA_MAX_DEVICES] = {};

// Device address of this device's `scratch` (looked up on the first call on
// the device, with it current).
int na_scratch(uint8_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return MPMX_ERR_HIP;
  if (dev < 0 || dev >= NA_MAX_DEVICES) return MPMX_ERR_CAPACITY;
  std::lock_guard<std::mutex> lock(na_mutex);
  if (!na_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MPMX_ERR_HIP;
    na_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = na_ptr[dev];
  return MPMX_OK;
}
}  // namespace
}  // namespace mpmx


using namespace mpmx;

extern "C" {

// 8 -> 64 -> 64 -> 64 -> 64 -> 4: W[out][in] + b[out] per layer
int64_t mpmx_closure_n_params(void) { return MLP_N_PARAMS; }

int mpmx_mlp_forward(const float* params, const float* norm24, const float* x8, int64_t n, float* y4) {
  if (!params || !norm24 || n < 0 || (n > 0 && (!x8 || !y4)) || !norm_ok(norm24)) return MPMX_ERR_ARG;
  if (n == 0) return MPMX_OK;
  std::string err;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPMX_ERR_HIP;
  int rc = MPMX_OK;
  {
    DevBuf blob, dx, dy;
    if ((rc = dmalloc(blob, MLP_BLOB_BYTES, s, err)) == 0 && (rc = dmalloc(dx, sizeof(float) * 8 * n, s, err)) == 0 &&
        (rc = dmalloc(dy, sizeof(float) * 4 * n, s, err)) == 0 &&
        (rc = pack_to_device(params, norm24, blob.as<uint8_t>(), nullptr, s, err)) == 0) {
      if (hipMemcpyAsync(dx.p, x8, sizeof(float) * 8 * n, hipMemcpyHostToDevice, s) != hipSuccess) rc = MPMX_ERR_HIP;
      if (!rc) {
        launch_mlp_forward(blob.as<uint8_t>(), dx.as<float>(), dy.as<float>(), n, s);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(y4, dy.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
          rc = MPMX_ERR_HIP;
      }
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

int mpmx_mlp_forward_noalloc(const float* params, const float* norm24, const float* x8, int64_t n, float* y4) {
  if (!params || !norm24 || n < 0 || (n > 0 && (!x8 || !y4)) || !norm_ok(norm24)) return MPMX_ERR_ARG;
  if (n > MPMX_MLP_NOALLOC_MAX) return MPMX_ERR_CAPACITY;
  if (n == 0) return MPMX_OK;
  uint8_t* base = nullptr;
  int rc = na_scratch(&base);
  if (rc) return rc;
  uint8_t* blob = base + NA_BLOB;
  float* dpar = reinterpret_cast<float*>(base + NA_PAR);
  float* dnorm = reinterpret_cast<float*>(base + NA_NORM);
  float* dx = reinterpret_cast<float*>(base + NA_X);
  float* dy = reinterpret_cast<float*>(base + NA_Y);
  const hipStream_t s = nullptr;
  // the blob's padding must read zero: cleared by a kernel on every call
  launch_zero_words(reinterpret_cast<uint32_t*>(blob), MLP_BLOB_BYTES / 4, s);
  if (hipMemcpyAsync(dpar, params, sizeof(float) * MLP_N_PARAMS, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dnorm, norm24, sizeof(float) * 24, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dx, x8, sizeof(float) * 8 * n, hipMemcpyHostToDevice, s) != hipSuccess)
    return MPMX_ERR_HIP;
  launch_mlp_pack(dpar, MLP_N_PARAMS, blob, nullptr, dnorm, s);
  launch_mlp_forward(blob, dx, dy, n, s);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(y4, dy, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return MPMX_ERR_HIP;
  return MPMX_OK;
}

int mpmx_mlp_gradient(const float* params, const float* norm24, const float* x8, const float* t4, int64_t n,
                      float* grad, float* loss) {
  if (!params || !norm24 || !x8 || !t4 || !grad || !loss || n <= 0 || n > INT32_MAX || !norm_ok(norm24))
    return MPMX_ERR_ARG;
  std::string err;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPMX_ERR_HIP;
  int rc = MPMX_OK;
  {
    DevBuf blob, wt, xn, tn, dg, dl, zs;
    const int64_t waves = (n + 63) / 64;
    if ((rc = dmalloc(blob, MLP_BLOB_BYTES, s, err)) == 0 &&
        (rc = dmalloc(wt, sizeof(uint16_t) * MLP_WT_ELEMS, s, err)) == 0 &&
        (rc = dmalloc(dg, sizeof(float) * MLP_N_PARAMS, s, err)) == 0 && (rc = dmalloc(dl, sizeof(float), s, err)) == 0 &&
        (rc = dmalloc(zs, sizeof(float) * 4 * 4096 * (size_t)waves, s, err)) == 0 &&
        (rc = pack_to_device(params, norm24, blob.as<uint8_t>(), wt.as<uint16_t>(), s, err)) == 0 &&
        (rc = upload_normalised(x8, t4, n, norm24, xn, tn, s, err)) == 0) {
      launch_mlp_grad(blob.as<uint8_t>(), wt.as<uint16_t>(), xn.as<float>(), tn.as<float>(), nullptr, (int)n,
                      1.0f / (4.0f * (float)n), dg.as<float>(), dl.as<float>(), zs.as<float>(), s);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(grad, dg.p, sizeof(float) * MLP_N_PARAMS, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        rc = MPMX_ERR_HIP;
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

void mpmx_train_config_default(mpmx_train_config* tc) {
  if (!tc) return;
  std::memset(tc, 0, sizeof *tc);
  tc->steps = 4000;
  tc->batch = 4096;
  tc->lr = 2e-3f;
  tc->lr_final = 5e-5f;
  tc->beta1 = 0.9f;
  tc->beta2 = 0.999f;
  tc->eps = 1e-8f;
  tc->seed = 1234;
  tc->init = 1;
  tc->report_every = 100;
}

}  // extern "C"
