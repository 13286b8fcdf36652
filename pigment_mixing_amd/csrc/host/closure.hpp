// closure.hpp — learned return-map closure state owned by a handle: BF16
// weights + FP32 biases and normalisation on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../kernels/mpm_kernels.hpp"
#include "../../../include/mpmx.h"

namespace mpmx {

struct ClosureState {
  int attached = 0;
  void* d_blob = nullptr;      // device copy of the packed network (layout in mlp_kernels.hpp)
  size_t blob_bytes = 0;
};

int closure_upload(ClosureState& c, const std::vector<float>& params, const float* in_mean, const float* in_std,
                   const float* out_mean, const float* out_std, hipStream_t s, std::string& err);
inline int closure_attach(ClosureState& c, const DevPtrs&, hipStream_t, std::string&) {
  c.attached = 1;
  return 0;
}
void closure_free(ClosureState& c);
void closure_init_params(std::vector<float>& p, uint64_t seed);
int closure_train(std::vector<float>& params, float* n24_io, const float* records12, int64_t n,
                  const mpmx_train_config& tc, mpmx_train_report* rep, float* curve, int64_t curve_cap,
                  hipStream_t s, std::string& err);

}  // namespace mpmx
