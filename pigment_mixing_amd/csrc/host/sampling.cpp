// sampling.cpp — particle sampling of blob objects for scene initialisation
// (host code; runs once before the step, like the reference's).
//
//   mpmx_sample_poisson_box   Poisson-disk points with minimum distance r in a
//                             box: Bridson's algorithm ("Fast Poisson disk
//                             sampling in arbitrary dimensions", SIGGRAPH 2007
//                             sketch), in place of thinks::PoissonDiskSampling
//                             (tph_poisson 0.3, fetched by the reference's build
//                             and absent here) as sampling_particles_vdb.hpp:33,
//                             113-117 calls it.
//   mpmx_sample_level_sets    the multi-object sampler of
//                             sampling_particles_vdb.hpp:86-160: per object, box
//                             samples over its bounds with seed + i, kept where
//                             the trilinearly interpolated level set is <= 0
//                             (object_vdb.hpp:73-85, bounds :107-118), tagged
//                             with the object index; dx from all box samples.
//
// The level sets come in as dense signed-distance grids (what a VDB grid's
// active bounding box holds after densification): reading .vdb files needs
// OpenVDB, which is not part of this build.
#include "mpmx.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

// PCG32 (O'Neill, "PCG: a family of simple fast space-efficient statistically
// good algorithms for random number generation", 2014), seeded through one
// splitmix64 round.
struct Pcg32 {
  uint64_t state;
  explicit Pcg32(uint64_t seed) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    state = z ^ (z >> 31);
  }
  uint32_t next() {
    const uint64_t old = state;
    state = old * 6364136223846793005ull + 1442695040888963407ull;
    const uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (x >> rot) | (x << ((32u - rot) & 31u));
  }
  double uniform() { return next() * (1.0 / 4294967296.0); }   // [0, 1)
};

using P3 = std::array<double, 3>;

// Bridson: background grid of cell r/sqrt(3) (at most one sample per cell), an
// active list seeded with one uniform point; each step takes a random active
// sample and tries `attempts` candidates uniform in the spherical shell
// [r, 2r] around it, keeping the first that lies in the box and no closer
// than r to any sample (neighbours within +-2 cells); a sample whose attempts
// all fail leaves the active list.
int poisson_box(const float lo_f[3], const float hi_f[3], float radius, int attempts, uint32_t seed,
                std::vector<P3>& pts) {
  pts.clear();
  if (!(radius > 0.0f) || attempts <= 0) return MPMX_ERR_ARG;
  P3 lo, hi;
  for (int a = 0; a < 3; ++a) {
    lo[a] = lo_f[a];
    hi[a] = hi_f[a];
    if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] >= lo[a])) return MPMX_ERR_ARG;
  }
  const double r = radius, r2 = r * r, cell = r / std::sqrt(3.0);
  int64_t gd[3];
  int64_t ncell = 1;
  for (int a = 0; a < 3; ++a) {
    gd[a] = std::max<int64_t>(1, (int64_t)std::ceil((hi[a] - lo[a]) / cell));
    ncell *= gd[a];
    if (ncell > ((int64_t)1 << 28)) return MPMX_ERR_ARG;   // radius too small for the box
  }
  std::vector<int32_t> grid((size_t)ncell, -1);
  auto cell_of = [&](const P3& p, int64_t c[3]) {
    for (int a = 0; a < 3; ++a) c[a] = std::min<int64_t>(gd[a] - 1, std::max<int64_t>(0, (int64_t)((p[a] - lo[a]) / cell)));
  };
  auto insert = [&](const P3& p) {
    int64_t c[3];
    cell_of(p, c);
    grid[(size_t)((c[0] * gd[1] + c[1]) * gd[2] + c[2])] = (int32_t)pts.size();
    pts.push_back(p);
  };
  Pcg32 rng(seed);
  std::vector<int32_t> active;
  P3 p0;
  for (int a = 0; a < 3; ++a) p0[a] = lo[a] + rng.uniform() * (hi[a] - lo[a]);
  active.push_back(0);
  insert(p0);
  const double two_pi = 6.283185307179586;
  while (!active.empty()) {
    const size_t ai = rng.next() % active.size();
    const P3 base = pts[(size_t)active[ai]];
    bool placed = false;
    for (int t = 0; t < attempts && !placed; ++t) {
      // radius with density ~ rho^2 on [r, 2r] (uniform in the shell's volume),
      // direction uniform on the sphere
      const double rr = r * std::cbrt(1.0 + 7.0 * rng.uniform());
      const double z = 2.0 * rng.uniform() - 1.0, ph = two_pi * rng.uniform();
      const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
      const P3 q = {base[0] + rr * s * std::cos(ph), base[1] + rr * s * std::sin(ph), base[2] + rr * z};
      if (q[0] < lo[0] || q[0] > hi[0] || q[1] < lo[1] || q[1] > hi[1] || q[2] < lo[2] || q[2] > hi[2]) continue;
      int64_t c[3];
      cell_of(q, c);
      bool ok = true;
      for (int64_t i = std::max<int64_t>(0, c[0] - 2); ok && i <= std::min(gd[0] - 1, c[0] + 2); ++i)
        for (int64_t j = std::max<int64_t>(0, c[1] - 2); ok && j <= std::min(gd[1] - 1, c[1] + 2); ++j)
          for (int64_t k = std::max<int64_t>(0, c[2] - 2); ok && k <= std::min(gd[2] - 1, c[2] + 2); ++k) {
            const int32_t o = grid[(size_t)((i * gd[1] + j) * gd[2] + k)];
            if (o < 0) continue;
            const P3& e = pts[(size_t)o];
            const double dx = e[0] - q[0], dy = e[1] - q[1], dz = e[2] - q[2];
            if (dx * dx + dy * dy + dz * dz < r2) ok = false;
          }
      if (!ok) continue;
      active.push_back((int32_t)pts.size());
      insert(q);
      placed = true;
    }
    if (!placed) {
      active[ai] = active.back();
      active.pop_back();
    }
  }
  return MPMX_OK;
}

// Trilinear value of a dense level set at world point p (after the object's
// offset and scale), +inf outside the grid (a level set's exterior).
double level_set_value(const mpmx_level_set& L, const P3& X) {
  double q[3];
  int64_t i0[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    const double p = (X[a] - (double)L.offset[a]) / (double)L.scale;
    q[a] = (p - (double)L.origin[a]) / (double)L.voxel;
    if (!(q[a] >= 0.0) || q[a] > (double)(L.dims[a] - 1)) return std::numeric_limits<double>::infinity();
    i0[a] = std::min<int64_t>(L.dims[a] - 2, (int64_t)std::floor(q[a]));
    f[a] = q[a] - (double)i0[a];
  }
  auto at = [&](int64_t i, int64_t j, int64_t k) {
    return (double)L.phi[(size_t)((i * L.dims[1] + j) * L.dims[2] + k)];
  };
  double v = 0.0;
  for (int c = 0; c < 8; ++c) {
    const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
    const double w = (di ? f[0] : 1.0 - f[0]) * (dj ? f[1] : 1.0 - f[1]) * (dk ? f[2] : 1.0 - f[2]);
    v += w * at(i0[0] + di, i0[1] + dj, i0[2] + dk);
  }
  return v;
}

}  // namespace

extern "C" int mpmx_sample_poisson_box(const float lo[3], const float hi[3], float radius, int32_t attempts,
                                       uint32_t seed, float* xyz, int64_t cap, int64_t* n_out) {
  if (!lo || !hi || !n_out || cap < 0 || (cap > 0 && !xyz)) return MPMX_ERR_ARG;
  std::vector<P3> pts;
  const int rc = poisson_box(lo, hi, radius, attempts, seed, pts);
  *n_out = rc == MPMX_OK ? (int64_t)pts.size() : 0;
  if (rc != MPMX_OK) return rc;
  if ((int64_t)pts.size() > cap) return MPMX_ERR_CAPACITY;
  for (size_t p = 0; p < pts.size(); ++p)
    for (int a = 0; a < 3; ++a) xyz[3 * p + a] = (float)pts[p][a];
  return MPMX_OK;
}

extern "C" int mpmx_sample_level_sets(const mpmx_level_set* objects, int32_t n_objects, float radius,
                                      int32_t attempts, uint32_t seed, float ppc, float* xyz, int32_t* object,
                                      int64_t cap, int64_t* n_out, float* dx_out) {
  if (!objects || n_objects <= 0 || !n_out || !dx_out || !(ppc > 0.0f) || cap < 0 ||
      (cap > 0 && (!xyz || !object)))
    return MPMX_ERR_ARG;
  *n_out = 0;
  *dx_out = 0.0f;
  for (int32_t i = 0; i < n_objects; ++i) {
    const mpmx_level_set& L = objects[i];
    if (!L.phi || L.dims[0] < 2 || L.dims[1] < 2 || L.dims[2] < 2 || !(L.voxel > 0.0f) || !(L.scale > 0.0f))
      return MPMX_ERR_ARG;
  }
  std::vector<P3> kept;
  std::vector<int32_t> tag;
  // the reference accumulates the box volumes and takes the cube root in
  // float (T = float, sampling_particles_vdb.hpp:100-141)
  float total_volume = 0.0f;
  int64_t total_samples = 0;
  std::vector<P3> pts;
  for (int32_t i = 0; i < n_objects; ++i) {
    const mpmx_level_set& L = objects[i];
    // bounds in FP32, one rounding per operation (no fused multiply-add), so
    // that a caller can reproduce them bit for bit
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
      const float o = L.origin[a] * L.scale;
      lo[a] = o + L.offset[a];
      const float e = (float)(L.dims[a] - 1) * L.voxel;
      const float top = (L.origin[a] + e) * L.scale;
      hi[a] = top + L.offset[a];
    }
    const int rc = poisson_box(lo, hi, radius, attempts, seed + (uint32_t)i, pts);
    if (rc != MPMX_OK) return rc;
    {
#pragma clang fp contract(off)
      const float l0 = hi[0] - lo[0], l1 = hi[1] - lo[1], l2 = hi[2] - lo[2];
      total_volume += (l0 * l1) * l2;
    }
    total_samples += (int64_t)pts.size();
    for (const P3& p : pts) {
      // the reference tests the float point it stores
      const P3 pf = {(double)(float)p[0], (double)(float)p[1], (double)(float)p[2]};
      if (level_set_value(L, pf) <= 0.0) {
        kept.push_back(pf);
        tag.push_back(i);
      }
    }
  }
  if (total_samples == 0) return MPMX_ERR_ARG;
  *n_out = (int64_t)kept.size();
  {
#pragma clang fp contract(off)
    *dx_out = std::cbrt(ppc / (float)total_samples * total_volume);
  }
  if ((int64_t)kept.size() > cap) return MPMX_ERR_CAPACITY;
  for (size_t p = 0; p < kept.size(); ++p) {
    for (int a = 0; a < 3; ++a) xyz[3 * p + a] = (float)kept[p][a];
    object[p] = tag[p];
  }
  return MPMX_OK;
}
