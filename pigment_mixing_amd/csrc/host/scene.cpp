// scene.cpp — seeded box-blob particle generator (synthetic benchmark inputs).
#include "../../../include/mpmx.h"

#include <cstdint>

namespace {

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline float uniform01(uint64_t seed, int64_t index, int axis) {
  const uint64_t h = splitmix64(seed * 0x100000001B3ull + 3ull * (uint64_t)index + (uint64_t)axis);
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

}  // namespace

extern "C" int mpmx_scene_boxes(const mpmx_box* boxes, int n_boxes, float dx, int per_axis, uint64_t seed,
                                float* x, float* pig, int64_t capacity, int64_t* n_out) {
  if (!boxes || n_boxes < 0 || !(dx > 0) || per_axis < 1 || !n_out) return MPMX_ERR_ARG;
  int64_t total = 0;
  for (int b = 0; b < n_boxes; ++b) {
    const mpmx_box& bx = boxes[b];
    if (bx.cells[0] < 0 || bx.cells[1] < 0 || bx.cells[2] < 0) return MPMX_ERR_ARG;
    total += (int64_t)bx.cells[0] * bx.cells[1] * bx.cells[2] * per_axis * per_axis * per_axis;
  }
  *n_out = total;
  if (!x) return MPMX_OK;
  if (capacity < total) return MPMX_ERR_CAPACITY;
  const float pa = (float)per_axis;
  int64_t id = 0;
  for (int b = 0; b < n_boxes; ++b) {
    const mpmx_box& bx = boxes[b];
    for (int ci = 0; ci < bx.cells[0]; ++ci)
      for (int cj = 0; cj < bx.cells[1]; ++cj)
        for (int ck = 0; ck < bx.cells[2]; ++ck)
          for (int si = 0; si < per_axis; ++si)
            for (int sj = 0; sj < per_axis; ++sj)
              for (int sk = 0; sk < per_axis; ++sk) {
                const int cell[3] = {ci, cj, ck};
                const int sub[3] = {si, sj, sk};
                for (int a = 0; a < 3; ++a) {
                  const float u = uniform01(seed, id, a);
                  x[3 * id + a] = bx.lo[a] + ((float)cell[a] + ((float)sub[a] + u) / pa) * dx;
                }
                if (pig)
                  for (int c = 0; c < 7; ++c) pig[7 * id + c] = bx.latent[c];
                ++id;
              }
  }
  return MPMX_OK;
}
