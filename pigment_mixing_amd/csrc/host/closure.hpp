// closure.hpp — learned return-map closure state owned by a handle: BF16
// weights + FP32 biases and normalisation on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../kernels/mpm_kernels.hpp"

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

}  // namespace mpmx
