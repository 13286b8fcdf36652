// blobs.cpp — the paint app's initial scene (host code, run once before the
// step): Simulation::blobs (simulation_blobs.cpp:7-150) with getMorphPair and
// findTForExactVolume (:152-218) on dense level sets, and the spatula's
// animation file (ObjectSpatula::loadAnimation, object_spatula.hpp:235-250).
//
// For each colour ratio the blob model database (volume + calibration factor
// per model, simulation.hpp:272-279) gives a target volume; the two models
// bracketing it are blended as signed-distance fields, (1 - t) A + t B over the
// union of their extents, with t from a 15-step bisection on the voxel-counted
// volume; the blended blobs are placed touching each other (n = 1..4 by the
// reference's closed forms, a row beyond), centred in x-z, and sampled with
// the multi-object Poisson-disk sampler (mpmx_sample_level_sets) at the
// reference's radius 0.01 and ppc 27.  The reference holds its models as
// OpenVDB grids (blosc-compressed .vdb files); here they are dense grids on a
// common voxel lattice, values outside a grid's extent being the narrow-band
// background (3 voxels, positive).
#include "mpmx.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

namespace {

// A level set on the shared voxel lattice: index (i, j, k) sits at world
// (lo + (i, j, k)) * voxel.
struct Dense {
  int64_t lo[3];
  int dims[3];
  float voxel;
  float background;
  std::vector<float> phi;
  float at(int64_t i, int64_t j, int64_t k) const {
    const int64_t a = i - lo[0], b = j - lo[1], c = k - lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= dims[0] || b >= dims[1] || c >= dims[2]) return background;
    return phi[(size_t)((a * dims[1] + b) * dims[2] + c)];
  }
};

int to_dense(const mpmx_level_set& L, float voxel, Dense& d) {
  if (!L.phi || L.dims[0] < 2 || L.dims[1] < 2 || L.dims[2] < 2 || !(L.voxel > 0.0f)) return MPMX_ERR_ARG;
  if (std::fabs(L.voxel - voxel) > 1e-6f * voxel) return MPMX_ERR_ARG;   // one lattice for all models
  for (int a = 0; a < 3; ++a) {
    const double q = (double)L.origin[a] / voxel;
    const double r = std::nearbyint(q);
    if (std::fabs(q - r) > 1e-3) return MPMX_ERR_ARG;                   // origin on the lattice
    d.lo[a] = (int64_t)r;
    d.dims[a] = L.dims[a];
  }
  d.voxel = voxel;
  d.background = 3.0f * voxel;
  d.phi.assign(L.phi, L.phi + (size_t)L.dims[0] * L.dims[1] * L.dims[2]);
  return MPMX_OK;
}

// Simulation::getMorphPair (simulation_blobs.cpp:152-163)
int morph_pair(const mpmx_blob_model* db, int n_db, float target) {
  if (target <= db[0].volume) return 0;
  if (target >= db[n_db - 1].volume) return n_db - 2;
  for (int i = 0; i < n_db - 1; ++i)
    if (target >= db[i].volume && target <= db[i + 1].volume) return i;
  return n_db - 2;
}

void union_box(const Dense& A, const Dense& B, int64_t lo[3], int64_t hi[3]) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::min(A.lo[a], B.lo[a]);
    hi[a] = std::max(A.lo[a] + A.dims[a] - 1, B.lo[a] + B.dims[a] - 1);
  }
}

// Simulation::findTForExactVolume (simulation_blobs.cpp:165-218): voxels inside
// both models always count; voxels inside one of them count when the blend is
// <= 0; 15 bisection steps on t, the volume being count * voxel^3 * the
// blended calibration factor; returns the last midpoint.
float find_t(const Dense& A, const Dense& B, float fA, float fB, float target) {
  int64_t lo[3], hi[3];
  union_box(A, B, lo, hi);
  int64_t always = 0;
  std::vector<std::pair<float, float>> edge;
  for (int64_t k = lo[2]; k <= hi[2]; ++k)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        const float a = A.at(i, j, k), b = B.at(i, j, k);
        if (a <= 0.0f && b <= 0.0f) ++always;
        else if (a <= 0.0f || b <= 0.0f) edge.emplace_back(a, b);
      }
  const float vv = A.voxel * A.voxel * A.voxel;
  float t_low = 0.0f, t_high = 1.0f, t_mid = 0.5f;
  for (int it = 0; it < 15; ++it) {
    t_mid = (t_low + t_high) * 0.5f;
    int64_t inside = always;
    for (const auto& e : edge)
      if (e.first * (1.0f - t_mid) + e.second * t_mid <= 0.0f) ++inside;
    const float factor = fA * (1.0f - t_mid) + fB * t_mid;
    const float vol = (float)inside * vv * factor;
    if (vol < target) t_low = t_mid;
    else t_high = t_mid;
  }
  return t_mid;
}

// gridA->topologyUnion(gridB) then (1 - t) A + t B over it (simulation_blobs.cpp:31-40).
Dense blend(const Dense& A, const Dense& B, float t) {
  if (!(t > 1e-6f)) return A;
  int64_t lo[3], hi[3];
  union_box(A, B, lo, hi);
  Dense m;
  for (int a = 0; a < 3; ++a) {
    m.lo[a] = lo[a];
    m.dims[a] = (int)(hi[a] - lo[a] + 1);
  }
  m.voxel = A.voxel;
  m.background = A.background;
  m.phi.resize((size_t)m.dims[0] * m.dims[1] * m.dims[2]);
  size_t p = 0;
  for (int64_t i = lo[0]; i <= hi[0]; ++i)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t k = lo[2]; k <= hi[2]; ++k) m.phi[p++] = A.at(i, j, k) * (1.0f - t) + B.at(i, j, k) * t;
  return m;
}

}  // namespace

extern "C" int mpmx_blob_scene(const mpmx_blob_model* db, int32_t n_db, const float* ratios, int32_t n,
                               float radius, int32_t attempts, uint32_t seed, float ppc, float* xyz,
                               int32_t* object, int64_t cap, int64_t* n_out, float* dx_out,
                               mpmx_blob_placement* placement) {
  if (!db || n_db < 2 || !ratios || n <= 0 || !n_out || !dx_out) return MPMX_ERR_ARG;
  for (int i = 0; i + 1 < n_db; ++i)
    if (!(db[i].volume <= db[i + 1].volume)) return MPMX_ERR_ARG;   // ascending volumes
  const float voxel = db[0].shape.voxel;
  std::vector<Dense> models(n_db);
  for (int i = 0; i < n_db; ++i) {
    const int rc = to_dense(db[i].shape, voxel, models[i]);
    if (rc) return rc;
  }
  // ---- morphed blobs (simulation_blobs.cpp:8-59)
  const float min_ratio = 0.1f, max_ratio = 0.9f;
  const float min_vol = db[0].volume, max_vol = db[n_db - 1].volume;
  std::vector<Dense> blobs(n);
  std::vector<float> radii(n);
  std::vector<std::array<float, 3>> centers(n), bmin(n), bmax(n);
  std::vector<float> tv(n), targets(n);
  std::vector<int> pairs(n);
  for (int i = 0; i < n; ++i) {
    const float ratio = std::max(min_ratio, std::min(max_ratio, ratios[i]));
    float target = min_vol + (ratio - min_ratio) / (max_ratio - min_ratio) * (max_vol - min_vol);
    target = std::max(min_vol, std::min(max_vol, target));
    const int a = morph_pair(db, n_db, target);
    const float t = find_t(models[a], models[a + 1], db[a].calibration, db[a + 1].calibration, target);
    blobs[i] = blend(models[a], models[a + 1], t);
    for (int d = 0; d < 3; ++d) {   // ObjectVdb::bounds (object_vdb.hpp:107-118), offset 0, scale 1
      bmin[i][d] = (float)blobs[i].lo[d] * voxel;
      bmax[i][d] = (float)(blobs[i].lo[d] + blobs[i].dims[d] - 1) * voxel;
      centers[i][d] = (bmin[i][d] + bmax[i][d]) * 0.5f;
    }
    const float rx = (bmax[i][0] - bmin[i][0]) * 0.5f, rz = (bmax[i][2] - bmin[i][2]) * 0.5f;
    radii[i] = std::max(rx, rz) - 0.055f;   // the reference's padding allowance (:55-57)
    tv[i] = t;
    targets[i] = target;
    pairs[i] = a;
  }
  // ---- placement (simulation_blobs.cpp:61-129): blobs touching in x-z
  std::vector<std::array<float, 3>> off(n, {0.0f, 0.0f, 0.0f});
  if (n == 2) {
    off[0][0] = -radii[0];
    off[1][0] = radii[1];
  } else if (n == 3) {
    off[1][0] = radii[0] + radii[1];
    const float x1 = off[1][0], r0 = radii[0], r1 = radii[1], r2 = radii[2];
    const float x2 = ((r0 + r2) * (r0 + r2) - (r1 + r2) * (r1 + r2) + x1 * x1) / (2.0f * x1);
    const float z2 = std::sqrt(std::max(0.0f, (r0 + r2) * (r0 + r2) - x2 * x2));
    off[2][0] = x2;
    off[2][2] = z2;
  } else if (n == 4) {
    const float r0 = radii[0], r1 = radii[1], r2 = radii[2], r3 = radii[3];
    off[0][0] = -r0;
    off[1][0] = r1;
    off[2][0] = -r2;
    off[3][0] = r3;
    auto calc_z = [](float ddx, float dist) { return std::sqrt(std::max(0.0f, dist * dist - ddx * ddx)); };
    const float Z = std::max({calc_z(off[0][0] - off[2][0], r0 + r2), calc_z(off[1][0] - off[3][0], r1 + r3),
                              calc_z(off[0][0] - off[3][0], r0 + r3), calc_z(off[1][0] - off[2][0], r1 + r2)});
    off[2][2] = Z;
    off[3][2] = Z;
  } else if (n > 4) {
    float cx = 0.0f;
    for (int i = 0; i < n; ++i) {
      if (i > 0) cx += radii[i - 1] + radii[i];
      off[i][0] = cx;
    }
  }
  // centre the x-z bounds of the discs (y untouched: the blobs rest where they are)
  std::array<float, 3> lo3 = off[0], hi3 = off[0];
  for (int d : {0, 2}) {
    lo3[d] = off[0][d] - radii[0];
    hi3[d] = off[0][d] + radii[0];
  }
  for (int i = 1; i < n; ++i)
    for (int d : {0, 2}) {
      lo3[d] = std::min(lo3[d], off[i][d] - radii[i]);
      hi3[d] = std::max(hi3[d], off[i][d] + radii[i]);
    }
  std::array<float, 3> ctr = {(lo3[0] + hi3[0]) * 0.5f, 0.0f, (lo3[2] + hi3[2]) * 0.5f};
  std::vector<mpmx_level_set> sets(n);
  for (int i = 0; i < n; ++i) {
    float fo[3];
    for (int d = 0; d < 3; ++d) fo[d] = off[i][d] - ctr[d];
    fo[0] -= centers[i][0];
    fo[2] -= centers[i][2];
    mpmx_level_set& L = sets[i];
    L.phi = blobs[i].phi.data();
    for (int d = 0; d < 3; ++d) {
      L.dims[d] = blobs[i].dims[d];
      L.origin[d] = (float)blobs[i].lo[d] * voxel;
      L.offset[d] = fo[d];
    }
    L.voxel = voxel;
    L.scale = 1.0f;
    if (placement) {
      mpmx_blob_placement& P = placement[i];
      P.pair = pairs[i];
      P.t = tv[i];
      P.target_volume = targets[i];
      P.radius = radii[i];
      for (int d = 0; d < 3; ++d) {
        P.offset[d] = fo[d];
        P.bbox_min[d] = bmin[i][d] + fo[d];
        P.bbox_max[d] = bmax[i][d] + fo[d];
      }
    }
  }
  // ---- sampling (simulation_blobs.cpp:149 -> sampling_particles_vdb.hpp:86-160)
  return mpmx_sample_level_sets(sets.data(), n, radius, attempts, seed, ppc, xyz, object, cap, n_out, dx_out);
}

extern "C" int mpmx_load_keyframes(const char* path, float* frames7, int64_t cap, int64_t* n_out) {
  if (!path || !n_out || cap < 0 || (cap > 0 && !frames7)) return MPMX_ERR_ARG;
  *n_out = 0;
  FILE* f = std::fopen(path, "rb");
  if (!f) return MPMX_ERR_IO;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  const int64_t n = size > 0 ? (int64_t)(size / 28) : 0;   // whole FrameData records, as the reference reads
  *n_out = n;
  int rc = MPMX_OK;
  if (n > cap) {
    rc = MPMX_ERR_CAPACITY;
  } else if (n > 0 && std::fread(frames7, 28, (size_t)n, f) != (size_t)n) {
    rc = MPMX_ERR_IO;
  }
  std::fclose(f);
  return rc;
}
