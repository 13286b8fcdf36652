// closure.cpp — learned return-map closure: parameter packing and upload.
#include "closure.hpp"

#include "../../../include/mpmx.h"

#include <cstdint>
#include <cstring>

namespace mpmx {

// The device blob: per layer BF16 W[out][in] (row-major) followed by FP32 b[out],
// then FP32 in_mean[8], in_std[8], out_mean[4], out_std[4].
int closure_upload(ClosureState& c, const std::vector<float>& params, const float* in_mean, const float* in_std,
                   const float* out_mean, const float* out_std, hipStream_t s, std::string& err) {
  const int dims[6] = {8, 64, 64, 64, 64, 4};
  std::vector<uint8_t> blob;
  size_t off = 0;
  auto put = [&](const void* p, size_t n) {
    blob.resize(off + n);
    std::memcpy(blob.data() + off, p, n);
    off += n;
  };
  size_t k = 0;
  for (int l = 0; l < 5; ++l) {
    const int nin = dims[l], nout = dims[l + 1];
    for (int i = 0; i < nout * nin; ++i) {
      uint32_t u;
      std::memcpy(&u, &params[k + i], 4);
      // round to nearest even BF16
      const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
      const uint16_t b = (uint16_t)(r >> 16);
      put(&b, 2);
    }
    k += (size_t)nout * nin;
    while (off % 4) put("\0", 1);
    put(&params[k], sizeof(float) * nout);
    k += nout;
  }
  put(in_mean, 32);
  put(in_std, 32);
  put(out_mean, 16);
  put(out_std, 16);
  if (c.blob_bytes < blob.size()) {
    if (c.d_blob) (void)hipFree(c.d_blob);
    c.d_blob = nullptr;
    if (hipMalloc(&c.d_blob, blob.size()) != hipSuccess) {
      err = "closure: hipMalloc failed";
      c.blob_bytes = 0;
      return MPMX_ERR_HIP;
    }
    c.blob_bytes = blob.size();
  }
  if (hipMemcpyAsync(c.d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "closure: upload failed";
    return MPMX_ERR_HIP;
  }
  return MPMX_OK;
}

void closure_free(ClosureState& c) {
  if (c.d_blob) (void)hipFree(c.d_blob);
  c.d_blob = nullptr;
  c.blob_bytes = 0;
  c.attached = 0;
}

}  // namespace mpmx

extern "C" {

// 8 -> 64 -> 64 -> 64 -> 64 -> 4: W[out][in] + b[out] per layer
int64_t mpmx_closure_n_params(void) {
  const int64_t d[6] = {8, 64, 64, 64, 64, 4};
  int64_t n = 0;
  for (int l = 0; l < 5; ++l) n += d[l + 1] * d[l] + d[l + 1];
  return n;
}

}  // extern "C"
