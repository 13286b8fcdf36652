// mpmx_api.cpp — the C-ABI (include/mpmx.h): handle, device buffers, lattice
// setup, knife kinematics upload, and the per-step HIP graph.
#include "../../../include/mpmx.h"
#include "../kernels/mpm_kernels.hpp"
#include "closure.hpp"
#include "kinematics.hpp"
#include "json_min.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace mpmx;

namespace {

constexpr int MAX_CHUNK = 1024;  // knife states uploaded per chunk of steps
constexpr int MAX_SCENES = 64;   // scenes per batched handle

// Plate positions and velocities during a step (ObjectPlate::move runs at the
// end of every step, object_plate.hpp:136-162).
struct PlateRun {
  int64_t step = -1;           // the step whose state `cur` holds (-1: none yet)
  float time = 0.0f;           // Simulation::time at the start of that step
  std::vector<mpmx_plate> cur;
};

struct Scene {
  mpmx_config cfg;
  std::vector<mpmx_plate> plates;    // as set (construction state)
  std::vector<mpmx_plate> plates_base;  // their state at plates_step0 (moved up to a dt change)
  int64_t plates_step0 = 0;          // step from which they are replayed
  float plates_time0 = 0.0f;
  PlateRun prun;
  // the knife's clock: time(step) = knife_t0 + (step - knife_s0) dt (a dt
  // change re-bases it at the current step)
  double knife_t0 = 0.0;
  int64_t knife_s0 = 0;
  bool has_knife = false;
  mpmx_knife knife;
  std::vector<float> frames;
  float frame_fps = 60.0f;
  bool has_override = false;
  mpmx_knife_state override_state;
  int64_t p_begin = 0, p_count = 0;  // particle range in upload order
  std::vector<float> comp;           // batch staging: P_NCOMP x p_count, component-major
  std::vector<uint8_t> act;
  int dims[3] = {0, 0, 0};
  float origin[3] = {0, 0, 0};
  std::vector<float> xs, ys, zs;
  SceneDev dev;
};

}  // namespace

struct mpmx_handle {
  std::string err;
  int device = 0;
  int n_cu = 256;
  hipStream_t stream = nullptr;
  // the learned closure's stream: k_closure runs beside the sort (both follow
  // G2P, the scatter follows both); default priority, like the main stream
  hipStream_t stream_closure = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<Scene> scenes;
  // host particle staging (upload order)
  int64_t np = 0;
  std::vector<float> h_comp;   // P_NCOMP * np, component-major
  std::vector<uint8_t> h_active;
  std::vector<int> h_pscene;
  bool prepared = false;
  int64_t step = 0;            // completed steps
  double time = 0.0;
  float time_f = 0.0f;         // float accumulation as Simulation::time (simulation_units.cpp:228)
  int frame = 0;
  // restart from a checkpoint: the step, clocks and stresses the next
  // mpmx_prepare starts from (cleared by it and by every particle upload)
  bool resume = false;
  bool tau_given = false;      // staged Kirchhoff-stress rows are the saved ones: no recomputation
  int64_t resume_step = 0;
  double resume_time = 0.0;
  float resume_time_f = 0.0f;
  int resume_frame = 0;
  // device memory
  void* arena = nullptr;
  std::vector<void*> allocs;
  DevPtrs d{};
  int nkeys = 0;
  int n_node_blocks = 0;
  int nchunks = 0;
  SceneDev* d_scenes = nullptr;
  KnifeStateDev* d_kstate = nullptr;
  KnifeStateDev* h_kstate = nullptr;  // pinned
  int* h_counters = nullptr;          // pinned
  unsigned long long* h_stats = nullptr;
  uint32_t* h_err = nullptr;
  int* d_step_zero = nullptr;
  // graph of two steps (A->B, B->A)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool parity_swapped = false;        // true when the current state lives in the B buffers
  float ms_last = 0.0f;
  unsigned long long stats_prev[ST_N] = {0};
  unsigned long long stats_last[ST_N] = {0};
  int64_t steps_last = 0;
  ClosureState closure;
  // learned closure parameters (FP32 master copy) and normalisation
  std::vector<float> mlp_params;
  float in_mean[8] = {0}, in_std[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  float out_mean[4] = {0}, out_std[4] = {1, 1, 1, 1};
  // harvest of (trial, corrected) pairs
  float* d_harvest = nullptr;
  int64_t harvest_cap = 0;
  unsigned long long harvest_base = 0;
  // batched handle (mpmx_create_batch): particles staged per scene
  bool batch = false;
  // launch shapes (workgroups per CU), set by mpmx_autotune
  int wg_g2p = 4, wg_p2g = 12, wg_grid = 4;
  int p2g_resident = 12;   // k_p2g workgroups resident per CU (its occupancy)
  int g2p_resident = 8;    // k_g2p workgroups resident per CU
};

namespace {

// A scene carries the eps_pl_vol / eps_pl_vol_pradhana rows when its model reads
// them: DPVisc (plasticity.cpp:240-340) or a von Mises map with an active
// pressure cap (p_min > -1e19, plasticity.cpp:172-180).
bool tracks_volume(const mpmx_config& c) {
  return c.plastic_model == MPMX_PLASTIC_DPVISC || (c.plastic_model == MPMX_PLASTIC_VMVISC && c.p_min > -1e19f);
}

int fail(mpmx_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((h), MPMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

template <typename T>
int dalloc(mpmx_handle* h, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  HIPCHK(h, hipMalloc(&q, bytes));
  HIPCHK(h, hipMemsetAsync(q, 0, bytes, h->stream));
  h->allocs.push_back(q);
  *p = (T*)q;
  return MPMX_OK;
}

void free_device(mpmx_handle* h) {
  if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  h->d = DevPtrs{};
}

// remesh.cpp:45-180 origin rule + tools.hpp:325-330 float arange on a fixed lattice.
void make_lattice(Scene& s, const float* x, const uint8_t* act) {
  const mpmx_config& c = s.cfg;
  float mn[3] = {1e20f, 1e20f, 1e20f}, mx[3] = {-1e20f, -1e20f, -1e20f};
  bool any = false;
  for (int64_t p = 0; p < s.p_count; ++p) {
    if (act && !act[p]) continue;
    any = true;
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], x[3 * p + d]);
      mx[d] = std::max(mx[d], x[3 * p + d]);
    }
  }
  if (!any) for (int d = 0; d < 3; ++d) mn[d] = mx[d] = 0.0f;
  if (c.grid_reference_point[0] < 1e10f) {
    for (int d = 0; d < 3; ++d) {
      if (c.grid_reference_point[d] < mn[d]) mn[d] = c.grid_reference_point[d];
      else if (c.grid_reference_point[d] > mx[d]) mx[d] = c.grid_reference_point[d];
    }
  }
  std::vector<float>* ax[3] = {&s.xs, &s.ys, &s.zs};
  for (int d = 0; d < 3; ++d) {
    const float dx = c.dx;
    float low = mn[d] - dx * 2u;
    const float high = mx[d] + dx * 2u;
    if (c.origin_mode == 1) low = c.grid_origin[d];
    ax[d]->clear();
    if (c.grid_dims[d] > 0) {
      float val = low;
      for (int i = 0; i < c.grid_dims[d]; ++i) { ax[d]->push_back(val); val += dx; }
    } else {
      for (float val = low; val < high + dx; val += dx) ax[d]->push_back(val);
    }
    s.dims[d] = (int)ax[d]->size();
    s.origin[d] = (*ax[d])[0];
  }
}

void fill_scene_dev(Scene& s, int node_block_base, int axis_off) {
  const mpmx_config& c = s.cfg;
  SceneDev& d = s.dev;
  std::memset(&d, 0, sizeof d);
  for (int i = 0; i < 3; ++i) {
    d.dims[i] = s.dims[i];
    d.bdims[i] = (s.dims[i] + 3) / 4;
    d.origin[i] = s.origin[i];
    d.gravity[i] = c.gravity[i];
    d.pb_min[i] = c.particle_boundary_min[i];
    d.pb_max[i] = c.particle_boundary_max[i];
  }
  d.node_block_base = node_block_base;
  d.axis_off = axis_off;
  d.dx = c.dx;
  d.one_over_dx = (float)(1.0 / c.dx);
  // derived constants as simulation_units.cpp:151-183 (double arithmetic, float storage)
  const float lambda = (float)(c.nu * c.E / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu)));
  const float mu = (float)(c.E / (2.0 * (1.0 + c.nu)));
  d.E = c.E;
  d.nu = c.nu;
  d.lambda = lambda;
  d.mu = mu;
  d.K = (float)(lambda + 2.0 * mu / 3);
  const float d_prefac = 1 / c.q_prefac;
  d.inv_d_prefac = (float)(1.0 / d_prefac);
  d.e_mu_prefac = 2 * c.q_prefac * mu;
  d.f_mu_prefac = 2 * c.q_prefac * c.q_prefac * mu;
  d.q_min = c.q_min;
  d.q_max = c.q_max;
  d.xi = c.xi;
  d.p_min = c.p_min;
  d.perzyna_exp = c.perzyna_exp;
  d.perzyna_visc = c.perzyna_visc;
  d.flip_ratio = c.flip_ratio;
  d.apicDinv = (float)(4.0 / (c.dx * c.dx));
  d.particle_volume = c.dx * c.dx * c.dx / c.ppc;
  d.particle_mass = c.rho * d.particle_volume;
  d.dt = c.dt;
  d.D_max = c.D_max;
  d.D_edge0 = c.D_edge0;
  d.D_edge1 = c.D_edge1;
  d.start_boost = c.start_boost_time;
  d.end_boost = c.end_boost_time;
  d.boost_factor = c.boost_factor;
  d.plastic_model = c.plastic_model;
  d.M = c.M;
  d.q_cohesion = c.q_cohesion;
  d.d_prefac = d_prefac;
  d.use_pradhana = c.use_pradhana ? 1 : 0;
  d.track_vol = tracks_volume(c) ? 1 : 0;
  d.use_musl = c.use_musl;
  d.use_pb = c.use_particle_boundaries;
  d.closure = c.closure;
  d.closure_tol = c.closure_tol;
  d.n_plates = (int)std::min<size_t>(s.plates.size(), MAX_PLATES);
  for (int i = 0; i < d.n_plates; ++i) {
    const mpmx_plate& p = s.plates[i];
    PlateDev& q = d.plates[i];
    q.pos_object = p.pos_object;
    q.friction = p.friction;
    q.pos_lower = p.pos_lower;
    q.pos_upper = p.pos_upper;
    q.v[0] = p.v[0]; q.v[1] = p.v[1]; q.v[2] = p.v[2];
    q.plate_type = p.plate_type;
    q.bc = p.bc;
  }
  d.knife.present = s.has_knife ? 1 : 0;
  if (s.has_knife) {
    d.knife.shape = s.knife.shape;
    d.knife.bc = s.knife.bc;
    d.knife.half[0] = s.knife.half[0];
    d.knife.half[1] = s.knife.half[1];
    d.knife.half[2] = s.knife.half[2];
    d.knife.offset = s.knife.offset;
    d.knife.friction = s.knife.friction;
  }
}

// ObjectPlate::move (object_plate.hpp:136-162) at the end of the step that
// starts at r.time: the construction velocity, divided by vmin_factor during
// the first load_factor frames, and a float step of dt v along the plate's
// normal axis (pos_object) and its tangential one (pos_lower, pos_upper).
void plate_move(const Scene& s, PlateRun& r) {
  const float dt = s.cfg.dt;
  const float frame_dt = (float)(1.0 / s.cfg.fps);
  for (size_t i = 0; i < r.cur.size(); ++i) {
    const mpmx_plate& o = s.plates[i];
    mpmx_plate& p = r.cur[i];
    const float vmin = o.vmin_factor > 0.0f ? o.vmin_factor : 1.0f;
    const float factor = r.time < o.load_factor * frame_dt ? (float)(1.0 / vmin) : 1.0f;
    for (int c = 0; c < 3; ++c) p.v[c] = o.v[c] * factor;
    if (p.plate_type == MPMX_PLATE_LEFT || p.plate_type == MPMX_PLATE_RIGHT) {
      p.pos_object += dt * p.v[0];
      p.pos_upper += dt * p.v[1];
      p.pos_lower += dt * p.v[1];
    } else if (p.plate_type == MPMX_PLATE_BOTTOM || p.plate_type == MPMX_PLATE_TOP) {
      p.pos_object += dt * p.v[1];
      p.pos_upper += dt * p.v[0];
      p.pos_lower += dt * p.v[0];
    } else {
      p.pos_object += dt * p.v[2];
    }
  }
  r.time += dt;
  ++r.step;
}

// Plates during step `step`, replayed from the step at which they were set.
const std::vector<mpmx_plate>& plates_at(Scene& s, int64_t step) {
  PlateRun& r = s.prun;
  if (r.step < 0 || r.step > step || r.cur.size() != s.plates.size()) {
    r.step = s.plates_step0;
    r.time = s.plates_time0;
    r.cur = s.plates_base;
  }
  while (r.step < step) plate_move(s, r);
  return r.cur;
}

KnifeStateDev to_dev(const mpmx_knife_state& k, float time, const std::vector<mpmx_plate>& plates) {
  KnifeStateDev d;
  std::memset(&d, 0, sizeof d);
  std::memcpy(d.T, k.T, sizeof d.T);
  std::memcpy(d.Tinv, k.Tinv, sizeof d.Tinv);
  std::memcpy(d.N, k.Nmat, sizeof d.N);
  std::memcpy(d.v, k.v, sizeof d.v);
  std::memcpy(d.w, k.w, sizeof d.w);
  d.time = time;
  for (size_t i = 0; i < plates.size() && i < (size_t)MAX_PLATES; ++i) {
    const mpmx_plate& p = plates[i];
    d.plate[i][0] = p.pos_object;
    d.plate[i][1] = p.pos_lower;
    d.plate[i][2] = p.pos_upper;
    d.plate[i][3] = p.v[0];
    d.plate[i][4] = p.v[1];
    d.plate[i][5] = p.v[2];
  }
  return d;
}

mpmx_knife_state scene_knife_state(const Scene& s, int64_t step) {
  if (!s.has_knife) {
    mpmx_knife_state z;
    std::memset(&z, 0, sizeof z);
    return z;
  }
  const double dt = (double)s.cfg.dt;
  return knife_state_at_time(s.knife, s.frames, s.frame_fps, s.knife_t0 + (double)(step - s.knife_s0) * dt,
                             s.knife_t0 + (double)(step - 1 - s.knife_s0) * dt, dt);
}

// Upload knife states + step times for global steps [g0, g0 + n] and reset
// the chunk-local step counter.
int upload_kstates(mpmx_handle* h, int64_t g0, int n, float t0) {
  const int ns = (int)h->scenes.size();
  float t = t0;
  for (int i = 0; i <= n; ++i) {
    for (int s = 0; s < ns; ++s) {
      Scene& sc = h->scenes[s];
      mpmx_knife_state k;
      if (i == 0 && sc.has_override) {
        k = sc.override_state;
      } else {
        k = scene_knife_state(sc, g0 + i);
      }
      h->h_kstate[(size_t)i * ns + s] = to_dev(k, t, plates_at(sc, g0 + i));
    }
    t += h->scenes[0].cfg.dt;  // Simulation::time accumulates in float
  }
  HIPCHK(h, hipMemcpyAsync(h->d_kstate, h->h_kstate, sizeof(KnifeStateDev) * (size_t)(n + 1) * ns,
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(&h->d.counters[CNT_STEP], h->d_step_zero, sizeof(int), hipMemcpyDeviceToDevice,
                           h->stream));
  return MPMX_OK;
}

DevPtrs swapped(const DevPtrs& d) {
  DevPtrs e = d;
  std::swap(e.pa, e.pb);
  std::swap(e.pa_id, e.pb_id);
  std::swap(e.cq_cnt, e.cq_cnt_prev);
  return e;
}

// Persistent grids are a multiple of 8 workgroups: the block walkers split
// their lists into one contiguous segment per XCD (8 on gfx950).
int round8(int v) { return std::max(8, (v + 7) / 8 * 8); }
int grid_fused(mpmx_handle* h) { return round8(h->n_cu * h->wg_g2p); }
int grid_grid(mpmx_handle* h) { return round8(h->n_cu * h->wg_grid); }
// P2G runs one-wavefront workgroups
int grid_p2g(mpmx_handle* h) { return round8(h->n_cu * h->wg_p2g); }
// the closure kernel (network in LDS, activations in registers) runs three
// 256-thread workgroups per CU (its registers allow three waves per SIMD)
int grid_closure(mpmx_handle* h) { return round8(h->n_cu * 3); }

// G2P with the exact return map, or with the learned closure: G2P queues the
// plastic candidates and k_closure maps them.
bool learned(const mpmx_handle* h) { return h->scenes[0].cfg.closure != MPMX_CLOSURE_EXACT; }
bool musl(const mpmx_handle* h) { return h->scenes[0].cfg.use_musl != 0; }
// USL: one fused G2P.  MUSL (musl.cpp:6-98): first half, second scatter of
// momentum + grid velocity, second half.  The learned closure is forked onto
// its own stream (it reads only G2P's queue, and the sort reads nothing it
// writes) and joined before the scatter (enqueue_join_closure), unless
// before_closure asks for the kernels in sequence (for timing each).
void enqueue_g2p(mpmx_handle* h, const DevPtrs& d, hipEvent_t before_closure = nullptr) {
  const int g = grid_fused(h);
  const int lrn = learned(h) ? 1 : 0;
  if (musl(h)) {
    launch_g2p(d, g, 0, 1, h->stream);
    launch_musl(d, grid_p2g(h), grid_grid(h), h->stream);
    launch_g2p(d, g, lrn, 2, h->stream);
  } else {
    launch_g2p(d, g, lrn, 0, h->stream);
  }
  if (before_closure) {
    (void)hipEventRecord(before_closure, h->stream);
    if (lrn) launch_closure(d, h->closure.d_blob, grid_closure(h), h->stream);
  } else if (lrn) {
    (void)hipEventRecord(h->ev_fork, h->stream);
    (void)hipStreamWaitEvent(h->stream_closure, h->ev_fork, 0);
    launch_closure(d, h->closure.d_blob, grid_closure(h), h->stream_closure);
    (void)hipEventRecord(h->ev_join, h->stream_closure);
  }
}
// The scatter reads the closure's output (F, stress): wait for its stream.
void enqueue_join_closure(mpmx_handle* h) {
  if (learned(h)) (void)hipStreamWaitEvent(h->stream, h->ev_join, 0);
}

// One step: G2P (A -> B, counts for the new sort), sort, P2G from B, grid update.
void enqueue_step(mpmx_handle* h, const DevPtrs& d) {
  enqueue_g2p(h, d);
  launch_scan(d, h->nkeys, 1, h->stream);
  launch_perm(d, h->np, grid_grid(h), 1, h->stream);
  enqueue_join_closure(h);
  launch_p2g(d, grid_p2g(h), 1, h->stream);
  launch_grid(d, grid_grid(h), h->stream);
}

int build_graph(mpmx_handle* h) {
  if (h->graph_exec) return MPMX_OK;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(h, h->d);
  enqueue_step(h, swapped(h->d));
  hipError_t e = hipStreamEndCapture(h->stream, &h->graph);
  if (e != hipSuccess) return fail(h, MPMX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  HIPCHK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
  return MPMX_OK;
}

int read_stats(mpmx_handle* h) {
  HIPCHK(h, hipMemcpyAsync(h->h_stats, h->d.stats, sizeof(unsigned long long) * ST_N, hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_counters, h->d.counters, sizeof(int) * CNT_NCOUNTERS, hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_err, h->d.err, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return MPMX_OK;
}

// Device pointers with the current state in `pa`.
DevPtrs current(const mpmx_handle* h) { return h->parity_swapped ? swapped(h->d) : h->d; }

}  // namespace

// ============================================================== C entry points
extern "C" {

const char* mpmx_status_string(int s) {
  switch (s) {
    case MPMX_OK: return "ok";
    case MPMX_ERR_ARG: return "invalid argument";
    case MPMX_ERR_HIP: return "HIP runtime error";
    case MPMX_ERR_STATE: return "call out of order";
    case MPMX_ERR_DEVICE: return "device status word set";
    case MPMX_ERR_CAPACITY: return "capacity exceeded";
    case MPMX_ERR_IO: return "I/O error";
  }
  return "unknown";
}

void mpmx_config_default(mpmx_config* c) {
  std::memset(c, 0, sizeof *c);
  // Simulation constructor values (simulation.hpp:38-126)
  c->E = 1e5f;
  c->nu = 0.3f;
  c->rho = 1000.0f;
  c->flip_ratio = -0.95f;
  c->plastic_model = MPMX_PLASTIC_NONE;
  c->q_max = 100.0f;
  c->q_min = 100.0f;
  c->xi = 0.0f;
  c->p_min = -1e20f;
  c->perzyna_exp = 1.0f;
  c->perzyna_visc = 0.0f;
  c->q_prefac = (float)(1.0 / std::sqrt(2.0));
  c->cfl = 0.5f;
  c->M = 1.0f;
  c->q_cohesion = 0.0f;
  c->gravity[0] = 0.0f; c->gravity[1] = 0.0f; c->gravity[2] = 0.0f;
  c->D_max = 10e-2f;
  c->D_edge0 = 0.5f;
  c->D_edge1 = 0.8f;
  c->start_boost_time = 20.0f;
  c->end_boost_time = 40.0f;
  c->boost_factor = 2.0f;
  c->dx = 0.01f;
  c->ppc = 8.0f;
  c->origin_mode = 0;
  c->grid_reference_point[0] = c->grid_reference_point[1] = c->grid_reference_point[2] = 2e10f;
  c->dt = 1e-5f;
  c->fps = 1.0f;
  c->use_musl = 0;
  c->use_particle_boundaries = 0;
  for (int d = 0; d < 3; ++d) { c->particle_boundary_min[d] = -5.0f; c->particle_boundary_max[d] = 5.0f; }
  c->closure = MPMX_CLOSURE_EXACT;
  c->closure_tol = 1e-4f;
  c->n_threads = 1;
  c->use_pradhana = 1;
}

int mpmx_config_load_json(mpmx_config* c, const char* sim_path, const char* pig_path) {
  if (!c) return MPMX_ERR_ARG;
  int rc = MPMX_OK;
  if (sim_path) {
    JsonFlat j;
    if (j.load(sim_path)) {
      c->cfl = j.num("cfl", 0.2f);
      c->flip_ratio = j.num("flip_ratio", -0.8f);
      c->E = j.num("E", 500000.0f);
      c->nu = j.num("nu", 0.3f);
      c->rho = j.num("rho", 1000.0f);
      const float angle = j.num("friction_angle", 5.0f);
      c->M = (float)std::tan(angle * M_PI / 180.0);
      c->q_cohesion = j.num("q_cohesion", 500.0f);
      c->perzyna_exp = j.num("perzyna_exp", 1.0f);
      c->perzyna_visc = j.num("perzyna_visc", 0.1f);
      // keys beyond the reference's (BASELINE material spec)
      if (j.has("q_max")) c->q_max = j.num("q_max", c->q_max);
      if (j.has("q_min")) c->q_min = j.num("q_min", c->q_min);
      if (j.has("yield")) { c->q_max = c->q_min = j.num("yield", c->q_max); }
      if (j.has("viscosity") && c->q_max > 0) c->perzyna_visc = j.num("viscosity", 0.0f) / c->q_max;
      if (j.has("xi")) c->xi = j.num("xi", c->xi);
      if (j.has("dt")) c->dt = j.num("dt", c->dt);
      if (j.has("dx")) c->dx = j.num("dx", c->dx);
      if (j.has("ppc")) c->ppc = j.num("ppc", c->ppc);
      if (j.has("use_musl")) c->use_musl = (int)j.num("use_musl", 0.0f);
      if (j.has("gravity_y")) c->gravity[1] = j.num("gravity_y", c->gravity[1]);
      if (j.has("plastic_model")) c->plastic_model = (int)j.num("plastic_model", 0.0f);
      if (j.has("use_pradhana")) c->use_pradhana = (int)j.num("use_pradhana", 1.0f);
    } else {
      // simulation_units.cpp:69-80
      c->cfl = 0.2f; c->flip_ratio = -0.8f; c->E = 5e5f; c->nu = 0.3f; c->rho = 1000.0f;
      c->M = (float)std::tan(5 * M_PI / 180.0); c->q_cohesion = 500.0f; c->perzyna_exp = 1.0f;
      c->perzyna_visc = 0.1f;
      rc = MPMX_ERR_IO;
    }
  }
  if (pig_path) {
    JsonFlat j;
    if (j.load(pig_path)) {
      c->D_max = j.num("D_max", 10e-2f);
      c->D_edge0 = j.num("D_edge0", 0.5f);
      c->D_edge1 = j.num("D_edge1", 0.8f);
      c->start_boost_time = j.num("start_boost_time", 20.0f);
      c->end_boost_time = j.num("end_boost_time", 40.0f);
      c->boost_factor = j.num("boost_factor", 2.0f);
    } else {
      rc = MPMX_ERR_IO;
    }
  }
  return rc;
}

namespace {
int check_config(const mpmx_config* cfg);
}

int mpmx_create(const mpmx_config* cfg, mpmx_handle** out) {
  return mpmx_create_batch(cfg, 1, out);
}

int mpmx_create_batch(const mpmx_config* cfgs, int n_scenes, mpmx_handle** out) {
  if (!cfgs || !out || n_scenes < 1 || n_scenes > MAX_SCENES) return MPMX_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < n_scenes; ++i) {
    if (check_config(&cfgs[i])) return MPMX_ERR_ARG;
    // one clock and one G2P variant for the whole batch
    if (cfgs[i].dt != cfgs[0].dt || cfgs[i].closure != cfgs[0].closure || cfgs[i].use_musl != cfgs[0].use_musl)
      return MPMX_ERR_ARG;
  }
  const mpmx_config* cfg = cfgs;
  mpmx_handle* h = new mpmx_handle();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    delete h;
    return MPMX_ERR_HIP;
  }
  h->device = dev;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) h->n_cu = prop.multiProcessorCount;
  // P2G's first items per wavefront are static and assume every wavefront is
  // resident from the start: launch it at its occupancy
  h->wg_p2g = h->p2g_resident = p2g_resident_per_cu();
  h->g2p_resident = g2p_resident_per_cu();
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream_closure, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess) {
    mpmx_destroy(h);
    return MPMX_ERR_HIP;
  }
  h->scenes.resize(n_scenes);
  for (int i = 0; i < n_scenes; ++i) h->scenes[i].cfg = cfg[i];
  h->batch = n_scenes > 1;
  *out = h;
  return MPMX_OK;
}

void mpmx_destroy(mpmx_handle* h) {
  if (!h) return;
  if (h->d_harvest) (void)hipFree(h->d_harvest);
  free_device(h);
  closure_free(h->closure);
  if (h->h_kstate) (void)hipHostFree(h->h_kstate);
  if (h->h_counters) (void)hipHostFree(h->h_counters);
  if (h->h_stats) (void)hipHostFree(h->h_stats);
  if (h->h_err) (void)hipHostFree(h->h_err);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->stream_closure) (void)hipStreamDestroy(h->stream_closure);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char* mpmx_last_error(const mpmx_handle* h) { return h ? h->err.c_str() : "null handle"; }

namespace {
// Host SoA (per particle) -> component-major staging [P_NCOMP][n]
void stage(std::vector<float>& comp, std::vector<uint8_t>& actv, int64_t n, const float* x, const float* v,
           const float* F, const float* C, const float* pig, const float* diff, const float* eps,
           const uint8_t* active) {
  comp.assign((size_t)P_NCOMP * n, 0.0f);
  actv.assign(n, 1);
  for (int64_t p = 0; p < n; ++p) {
    for (int c = 0; c < 3; ++c) {
      comp[(P_X + c) * n + p] = x[3 * p + c];
      comp[(P_V + c) * n + p] = v ? v[3 * p + c] : 0.0f;
    }
    for (int c = 0; c < 9; ++c) {
      comp[(P_C + c) * n + p] = C ? C[9 * p + c] : 0.0f;
      comp[(P_F + c) * n + p] = F ? F[9 * p + c] : ((c == 0 || c == 4 || c == 8) ? 1.0f : 0.0f);
    }
    for (int c = 0; c < NPIG; ++c) comp[(P_PIG + c) * n + p] = pig ? pig[7 * p + c] : 0.0f;
    comp[P_DIFF * n + p] = diff ? diff[p] : 0.0f;
    comp[P_EPS * n + p] = eps ? eps[p] : 0.0f;
    if (active) actv[p] = active[p] ? 1 : 0;
  }
}

// Batched handle: concatenate the scenes' staging in scene order.
int gather_batch(mpmx_handle* h) {
  int64_t n = 0;
  for (Scene& s : h->scenes) {
    if (s.p_count <= 0) return fail(h, MPMX_ERR_STATE, "prepare: every scene of a batch needs particles");
    s.p_begin = n;
    n += s.p_count;
  }
  if (n > (int64_t)0x7fffffff / 2) return fail(h, MPMX_ERR_CAPACITY, "prepare: too many particles");
  h->np = n;
  h->h_comp.assign((size_t)P_NCOMP * n, 0.0f);
  h->h_active.resize(n);
  h->h_pscene.resize(n);
  for (int si = 0; si < (int)h->scenes.size(); ++si) {
    const Scene& s = h->scenes[si];
    for (int c = 0; c < P_NCOMP; ++c)
      std::copy(s.comp.begin() + (size_t)c * s.p_count, s.comp.begin() + (size_t)(c + 1) * s.p_count,
                h->h_comp.begin() + (size_t)c * n + s.p_begin);
    std::copy(s.act.begin(), s.act.end(), h->h_active.begin() + s.p_begin);
    std::fill(h->h_pscene.begin() + s.p_begin, h->h_pscene.begin() + s.p_begin + s.p_count, si);
  }
  return MPMX_OK;
}

int check_config(const mpmx_config* cfg) {
  if (!(cfg->dx > 0) || !(cfg->dt > 0) || !(cfg->ppc > 0) || !(cfg->rho > 0)) return MPMX_ERR_ARG;
  if (cfg->plastic_model != MPMX_PLASTIC_NONE && cfg->plastic_model != MPMX_PLASTIC_VMVISC &&
      cfg->plastic_model != MPMX_PLASTIC_DPVISC)
    return MPMX_ERR_ARG;
  if (cfg->plastic_model == MPMX_PLASTIC_DPVISC && !(cfg->M > 0)) return MPMX_ERR_ARG;
  // the learned closure replaces the von Mises/Perzyna map without the pressure cap
  if (cfg->closure != MPMX_CLOSURE_EXACT && tracks_volume(*cfg)) return MPMX_ERR_ARG;
  if (cfg->use_musl != 0 && cfg->use_musl != 1) return MPMX_ERR_ARG;
  if (cfg->closure != MPMX_CLOSURE_EXACT && cfg->closure != MPMX_CLOSURE_LEARNED &&
      cfg->closure != MPMX_CLOSURE_LEARNED_REFINED)
    return MPMX_ERR_ARG;
  if (cfg->closure != MPMX_CLOSURE_EXACT && !(cfg->closure_tol > 0)) return MPMX_ERR_ARG;
  return MPMX_OK;
}
}  // namespace

int mpmx_upload_particles(mpmx_handle* h, int64_t n, const float* x, const float* v, const float* F,
                          const float* C, const float* pig, const float* diff, const float* eps,
                          const uint8_t* active) {
  if (!h) return MPMX_ERR_ARG;
  if (h->batch) return fail(h, MPMX_ERR_STATE, "upload: batched handle, use mpmx_upload_scene");
  if (n <= 0 || !x) return fail(h, MPMX_ERR_ARG, "upload: need n > 0 and positions");
  if (n > (int64_t)0x7fffffff / 2) return fail(h, MPMX_ERR_CAPACITY, "upload: too many particles");
  h->np = n;
  h->resume = h->tau_given = false;
  stage(h->h_comp, h->h_active, n, x, v, F, C, pig, diff, eps, active);
  h->h_pscene.assign(n, 0);
  h->scenes[0].p_begin = 0;
  h->scenes[0].p_count = n;
  h->prepared = false;
  return MPMX_OK;
}

int mpmx_upload_scene(mpmx_handle* h, int scene, int64_t n, const float* x, const float* v, const float* F,
                      const float* C, const float* pig, const float* diff, const float* eps, const uint8_t* active) {
  if (!h) return MPMX_ERR_ARG;
  if (!h->batch) {
    if (scene != 0) return fail(h, MPMX_ERR_ARG, "upload_scene: no such scene");
    return mpmx_upload_particles(h, n, x, v, F, C, pig, diff, eps, active);
  }
  if (scene < 0 || scene >= (int)h->scenes.size()) return fail(h, MPMX_ERR_ARG, "upload_scene: no such scene");
  if (n <= 0 || !x) return fail(h, MPMX_ERR_ARG, "upload_scene: need n > 0 and positions");
  if (n > (int64_t)0x7fffffff / 2) return fail(h, MPMX_ERR_CAPACITY, "upload_scene: too many particles");
  Scene& s = h->scenes[scene];
  stage(s.comp, s.act, n, x, v, F, C, pig, diff, eps, active);
  s.p_count = n;
  h->np = 0;
  h->prepared = false;
  return MPMX_OK;
}

int mpmx_upload_plastic_state(mpmx_handle* h, int scene, const float* eps_pl_vol, const float* eps_pl_vol_pradhana) {
  if (!h) return MPMX_ERR_ARG;
  if (scene < 0 || scene >= (int)h->scenes.size()) return fail(h, MPMX_ERR_ARG, "upload_plastic_state: no such scene");
  std::vector<float>* comp;
  int64_t n;
  if (h->batch) {
    comp = &h->scenes[scene].comp;
    n = h->scenes[scene].p_count;
  } else {
    comp = &h->h_comp;
    n = h->np;
  }
  if (n <= 0 || comp->size() != (size_t)P_NCOMP * n)
    return fail(h, MPMX_ERR_STATE, "upload_plastic_state: upload the scene's particles first");
  for (int64_t p = 0; p < n; ++p) {
    (*comp)[(size_t)P_EPSV * n + p] = eps_pl_vol ? eps_pl_vol[p] : 0.0f;
    (*comp)[(size_t)P_PRAD * n + p] = eps_pl_vol_pradhana ? eps_pl_vol_pradhana[p] : 0.0f;
  }
  h->prepared = false;
  return MPMX_OK;
}

int mpmx_scene_range(mpmx_handle* h, int scene, int64_t* first, int64_t* count) {
  if (!h || !first || !count || scene < 0 || scene >= (int)h->scenes.size()) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "scene_range: prepare first");
  *first = h->scenes[scene].p_begin;
  *count = h->scenes[scene].p_count;
  return MPMX_OK;
}

int mpmx_set_objects(mpmx_handle* h, const mpmx_plate* plates, int n_plates, const mpmx_knife* knives,
                     int n_knives) {
  if (!h) return MPMX_ERR_ARG;
  if (h->batch) return fail(h, MPMX_ERR_STATE, "set_objects: batched handle, use mpmx_set_objects_scene");
  return mpmx_set_objects_scene(h, 0, plates, n_plates, knives, n_knives);
}

int mpmx_boundary_eval(const mpmx_config* cfg, const mpmx_plate* plates, int n_plates, const mpmx_knife* knife,
                       int64_t step, const float* X3, const float* v3, int64_t n, float* out3) {
  if (!cfg || n < 0 || (n > 0 && (!X3 || !v3 || !out3)) || step < 0 || !(cfg->dt > 0) || !(cfg->dx > 0))
    return MPMX_ERR_ARG;
  if (n_plates < 0 || n_plates > MAX_PLATES || (n_plates > 0 && !plates)) return MPMX_ERR_CAPACITY;
  for (int i = 0; i < n_plates; ++i)
    if (plates[i].bc < 0 || plates[i].bc > 2 || plates[i].plate_type < 0 || plates[i].plate_type > 5)
      return MPMX_ERR_ARG;
  if (knife && (knife->bc < 0 || knife->bc > 2 || knife->shape < 0 || knife->shape > 1)) return MPMX_ERR_ARG;
  if (n == 0) return MPMX_OK;
  Scene s;
  s.cfg = *cfg;
  if (n_plates > 0) s.plates.assign(plates, plates + n_plates);
  s.plates_base = s.plates;
  s.has_knife = knife != nullptr;
  if (knife) s.knife = *knife;
  fill_scene_dev(s, 0, 0);
  float t_step = 0.0f;
  for (int64_t i = 0; i < step; ++i) t_step += cfg->dt;   // Simulation::time accumulates in float
  const KnifeStateDev ks = to_dev(scene_knife_state(s, step), t_step, plates_at(s, step));
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return MPMX_ERR_HIP;
  SceneDev* d_sc = nullptr;
  KnifeStateDev* d_ks = nullptr;
  float* d_buf = nullptr;
  uint32_t* d_err = nullptr;
  uint32_t err = 0;
  const size_t bytes = sizeof(float) * 3 * (size_t)n;
  int rc = MPMX_OK;
  if (hipMalloc((void**)&d_sc, sizeof(SceneDev)) != hipSuccess || hipMalloc((void**)&d_ks, sizeof(KnifeStateDev)) != hipSuccess ||
      hipMalloc((void**)&d_buf, 3 * bytes) != hipSuccess || hipMalloc((void**)&d_err, sizeof(uint32_t)) != hipSuccess ||
      hipMemcpyAsync(d_sc, &s.dev, sizeof(SceneDev), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_ks, &ks, sizeof(KnifeStateDev), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_buf, X3, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_buf + 3 * n, v3, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(d_err, 0, sizeof(uint32_t), st) != hipSuccess) {
    rc = MPMX_ERR_HIP;
  } else {
    launch_boundary_eval(d_sc, d_ks, d_buf, d_buf + 3 * n, d_buf + 6 * n, n, d_err, st);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out3, d_buf + 6 * n, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&err, d_err, sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = MPMX_ERR_HIP;
  }
  (void)hipFree(d_sc);
  (void)hipFree(d_ks);
  (void)hipFree(d_buf);
  (void)hipFree(d_err);
  (void)hipStreamDestroy(st);
  if (rc == MPMX_OK && err) rc = MPMX_ERR_DEVICE;
  return rc;
}

int mpmx_set_objects_scene(mpmx_handle* h, int scene, const mpmx_plate* plates, int n_plates,
                           const mpmx_knife* knives, int n_knives) {
  if (!h) return MPMX_ERR_ARG;
  if (scene < 0 || scene >= (int)h->scenes.size()) return fail(h, MPMX_ERR_ARG, "set_objects: no such scene");
  if (n_plates < 0 || n_plates > MAX_PLATES || n_knives < 0 || n_knives > 1)
    return fail(h, MPMX_ERR_CAPACITY, "set_objects: at most 4 plates and 1 knife");
  for (int i = 0; i < n_plates; ++i)
    if (plates[i].bc < 0 || plates[i].bc > 2 || plates[i].plate_type < 0 || plates[i].plate_type > 5)
      return fail(h, MPMX_ERR_ARG, "set_objects: invalid plate bc/type");
  if (n_knives && (knives[0].bc < 0 || knives[0].bc > 2 || knives[0].shape < 0 || knives[0].shape > 1))
    return fail(h, MPMX_ERR_ARG, "set_objects: invalid knife bc/shape");
  Scene& s = h->scenes[scene];
  s.plates.assign(plates, plates + n_plates);
  s.plates_base = s.plates;
  // plates set on a running handle start from the given state at the current step
  s.plates_step0 = h->prepared ? h->step : 0;
  s.plates_time0 = h->prepared ? h->time_f : 0.0f;
  s.prun = PlateRun{};
  s.has_knife = n_knives > 0;
  if (s.has_knife) s.knife = knives[0];
  if (h->prepared) {
    fill_scene_dev(s, s.dev.node_block_base, s.dev.axis_off);
    HIPCHK(h, hipMemcpyAsync(h->d_scenes + scene, &s.dev, sizeof(SceneDev), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return MPMX_OK;
}

int mpmx_set_knife_keyframes(mpmx_handle* h, int k, const float* frames7, int n_frames, float fps) {
  if (!h || k < 0 || k >= (int)h->scenes.size() || n_frames < 0 || (n_frames > 0 && !frames7) || !(fps > 0))
    return MPMX_ERR_ARG;
  Scene& s = h->scenes[k];
  s.frames.assign(frames7, frames7 + 7 * (size_t)n_frames);
  s.frame_fps = fps;
  return MPMX_OK;
}

int mpmx_set_knife_state(mpmx_handle* h, int k, const mpmx_knife_state* st) {
  if (!h || k < 0 || k >= (int)h->scenes.size() || !st) return MPMX_ERR_ARG;
  h->scenes[k].has_override = true;
  h->scenes[k].override_state = *st;
  return MPMX_OK;
}

int mpmx_knife_state_at(mpmx_handle* h, int k, int64_t step, mpmx_knife_state* out) {
  if (!h || k < 0 || k >= (int)h->scenes.size() || !out) return MPMX_ERR_ARG;
  *out = scene_knife_state(h->scenes[k], step);
  return MPMX_OK;
}

int mpmx_set_dt(mpmx_handle* h, float dt) {
  if (!h || !(dt > 0.0f) || !std::isfinite(dt)) return MPMX_ERR_ARG;
  if (h->scenes.empty() || h->scenes[0].cfg.dt == dt) return MPMX_OK;
  const int64_t now = h->prepared ? h->step : 0;
  if (h->prepared) {
    // the grid velocities the next step gathers were updated at the end of
    // the last step (its P2G + grid update run ahead of the next G2P); they
    // are recomputed below with the new dt, from the merged tiles
    int rc = read_stats(h);
    if (rc) return rc;
    if (h->h_counters[CNT_PBLOCKS] > h->d.tile_cap)
      return fail(h, MPMX_ERR_STATE, "set_dt: the last scatter used the atomic path; its grid cannot be redone");
  }
  for (size_t i = 0; i < h->scenes.size(); ++i) {
    Scene& s = h->scenes[i];
    // the objects continue from their state at the current step: the knife's
    // clock and the plates' replay are re-based there with the old dt
    s.knife_t0 += (double)(now - s.knife_s0) * (double)s.cfg.dt;
    s.knife_s0 = now;
    if (h->prepared) {
      s.plates_base = plates_at(s, now);
      s.plates_time0 = s.prun.time;
      s.plates_step0 = now;
      s.prun = PlateRun{};
    }
    s.cfg.dt = dt;
    if (h->prepared) {
      fill_scene_dev(s, s.dev.node_block_base, s.dev.axis_off);
      HIPCHK(h, hipMemcpyAsync(h->d_scenes + i, &s.dev, sizeof(SceneDev), hipMemcpyHostToDevice, h->stream));
    }
  }
  if (h->prepared) {
    // the knife and plates of the next step as its grid update sees them
    // (entry 0, step counter 0), then the grid update again
    int rc = upload_kstates(h, h->step, 0, h->time_f);
    if (rc) return rc;
    launch_grid_redo(current(h), grid_grid(h), h->stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return MPMX_OK;
}

int mpmx_max_speed(mpmx_handle* h, float* vmax) {
  if (!h || !vmax) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "max_speed: prepare first");
  unsigned* d_out = nullptr;
  HIPCHK(h, hipMalloc(&d_out, sizeof(unsigned)));
  unsigned bits = 0u;
  int rc = MPMX_OK;
  if (hipMemsetAsync(d_out, 0, sizeof(unsigned), h->stream) != hipSuccess) rc = fail(h, MPMX_ERR_HIP, "max_speed");
  if (rc == MPMX_OK) {
    launch_max_speed(current(h), h->np, d_out, h->stream);
    if (hipMemcpyAsync(&bits, d_out, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(h, MPMX_ERR_HIP, "max_speed copy failed");
  }
  (void)hipFree(d_out);
  if (rc != MPMX_OK) return rc;
  float v2;
  std::memcpy(&v2, &bits, sizeof v2);
  *vmax = std::sqrt(v2);
  return MPMX_OK;
}

int mpmx_prepare(mpmx_handle* h) {
  if (!h) return MPMX_ERR_ARG;
  if (h->batch) {
    const int rc = gather_batch(h);
    if (rc) return rc;
  }
  if (h->np <= 0) return fail(h, MPMX_ERR_STATE, "prepare: upload particles first");
  free_device(h);
  HIPCHK(h, hipSetDevice(h->device));
  // ---- lattices (one per scene), node-block pool, coordinate tables
  int nb_total = 0;
  int axis_total = 0;
  std::vector<float> axis;
  // make_lattice needs interleaved positions of the scene's particles
  for (Scene& s : h->scenes) {
    std::vector<float> xi((size_t)3 * s.p_count);
    std::vector<uint8_t> ai(s.p_count);
    for (int64_t p = 0; p < s.p_count; ++p) {
      const int64_t q = s.p_begin + p;
      for (int c = 0; c < 3; ++c) xi[3 * p + c] = h->h_comp[(P_X + c) * h->np + q];
      ai[p] = h->h_active[q];
    }
    make_lattice(s, xi.data(), ai.data());
    for (int d = 0; d < 3; ++d)
      if (s.dims[d] < 3 || s.dims[d] > (1 << 20)) return fail(h, MPMX_ERR_ARG, "prepare: lattice too small/large");
    const int nbs = ((s.dims[0] + 3) / 4) * ((s.dims[1] + 3) / 4) * ((s.dims[2] + 3) / 4);
    if ((int64_t)nb_total + nbs > (int64_t)(1 << 30) / 64) return fail(h, MPMX_ERR_CAPACITY, "prepare: lattice too large");
    fill_scene_dev(s, nb_total, axis_total);
    nb_total += nbs;
    axis.insert(axis.end(), s.xs.begin(), s.xs.end());
    axis.insert(axis.end(), s.ys.begin(), s.ys.end());
    axis.insert(axis.end(), s.zs.begin(), s.zs.end());
    axis_total = (int)axis.size();
  }
  h->n_node_blocks = nb_total;
  h->nkeys = nb_total + 1;
  h->nchunks = scan_chunks(h->nkeys);
  const int ns = (int)h->scenes.size();
  const int64_t n = h->np;
  DevPtrs& d = h->d;
  d = DevPtrs{};
  d.cap = n;
  d.n_scenes = ns;
  d.dead_key = nb_total;
  d.nb_list_cap = nb_total;
  d.track_vol = 0;
  for (const Scene& sc : h->scenes) d.track_vol |= sc.dev.track_vol;
  int rc;
#define ALLOC(ptr, type, count)                         \
  if ((rc = dalloc(h, (type**)&(ptr), (count))) != 0) return rc;
  SceneDev* scn; ALLOC(scn, SceneDev, ns); d.scenes = scn; h->d_scenes = scn;
  KnifeStateDev* ks; ALLOC(ks, KnifeStateDev, (size_t)(MAX_CHUNK + 1) * ns); d.kstate = ks; h->d_kstate = ks;
  float* axd; ALLOC(axd, float, axis.size()); d.axis_coords = axd;
  ALLOC(d.pa, float, (size_t)P_NCOMP * n);
  ALLOC(d.pb, float, (size_t)P_NCOMP * n);
  ALLOC(d.pa_id, int, n);
  ALLOC(d.pb_id, int, n);
  int* psc; ALLOC(psc, int, n); d.pscene = psc;
  uint8_t* act; ALLOC(act, uint8_t, n); d.active_in = act;
  ALLOC(d.key_out, int, n);
  ALLOC(d.rank_out, int, n);
  ALLOC(d.cell_out, int, n);
  ALLOC(d.okey_out, int, n);
  ALLOC(d.perm, int, n);
  ALLOC(d.kslot, int, n);
  ALLOC(d.cmaj, int, n);
  ALLOC(d.cell_num, int, (size_t)h->nkeys * BLK3);
  ALLOC(d.pdesc, int4, h->nkeys);
  // merged P2G tiles: room for the listed blocks of dense paint (~355 particles
  // per block); blocks beyond it scatter with atomics.  MPMX_TILE_CAP lowers
  // the count (tests of the atomic path on small scenes).
  int64_t tiles = std::max<int64_t>(1024, n / 128);
  if (const char* env = std::getenv("MPMX_TILE_CAP")) tiles = std::max<int64_t>(1, std::atoll(env));
  d.tile_cap = (int)std::min<int64_t>(h->nkeys, tiles);
  ALLOC(d.tile_acc, float, (size_t)d.tile_cap * ACC_CH * PT3);
  ALLOC(d.list_pos, int, h->nkeys);
  if (learned(h)) {
    // closure queue: every live particle can be a candidate in one step
    ALLOC(d.cq, float4, (size_t)CQ_REC * ((n + CQ_GROUP - 1) / CQ_GROUP * CQ_GROUP));
    d.cq_cap = n;
    ALLOC(d.cq_cnt, int, CQ_SEGS * CQ_CNT_STRIDE);
    ALLOC(d.cq_cnt_prev, int, CQ_SEGS * CQ_CNT_STRIDE);
    ALLOC(d.cq_start, int, CQ_SEGS);
  }
  ALLOC(d.count_next, int, h->nkeys);
  ALLOC(d.count_cur, int, h->nkeys);
  ALLOC(d.offsets, int, h->nkeys);
  ALLOC(d.pb_list, int, h->nkeys);
  ALLOC(d.scan_partials, unsigned long long, h->nchunks);
  int* bsc; ALLOC(bsc, int, h->nkeys); d.blk_scene = bsc;
  ALLOC(d.acc, float, (size_t)nb_total * ACC_CH * BLK3);
  ALLOC(d.gv, float, (size_t)nb_total * GV_STRIDE * BLK3);
  ALLOC(d.nb_flag, int, nb_total);
  ALLOC(d.nb_list, int, (size_t)2 * nb_total);
  ALLOC(d.nb_live, int, nb_total);
  ALLOC(d.counters, int, CNT_NCOUNTERS);
  ALLOC(d.wq, int, 4 * CQ_SEGS * CQ_CNT_STRIDE);
  ALLOC(d.stats, unsigned long long, ST_N);
  ALLOC(d.err, uint32_t, 4);
  ALLOC(h->d_step_zero, int, 1);
#undef ALLOC
  if (!h->h_kstate) HIPCHK(h, hipHostMalloc((void**)&h->h_kstate, sizeof(KnifeStateDev) * (MAX_CHUNK + 1) * 64));
  if (!h->h_counters) HIPCHK(h, hipHostMalloc((void**)&h->h_counters, sizeof(int) * CNT_NCOUNTERS));
  if (!h->h_stats) HIPCHK(h, hipHostMalloc((void**)&h->h_stats, sizeof(unsigned long long) * ST_N));
  if (!h->h_err) HIPCHK(h, hipHostMalloc((void**)&h->h_err, sizeof(uint32_t) * 4));
  // ---- uploads
  std::vector<SceneDev> sd(ns);
  std::vector<int> bscene(h->nkeys, 0);
  for (int s = 0; s < ns; ++s) {
    sd[s] = h->scenes[s].dev;
    const int base = h->scenes[s].dev.node_block_base;
    const int nbs = h->scenes[s].dev.bdims[0] * h->scenes[s].dev.bdims[1] * h->scenes[s].dev.bdims[2];
    for (int b = 0; b < nbs; ++b) bscene[base + b] = s;
  }
  std::vector<int> ids(n);
  for (int64_t i = 0; i < n; ++i) ids[i] = (int)i;
  HIPCHK(h, hipMemcpyAsync(scn, sd.data(), sizeof(SceneDev) * ns, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(axd, axis.data(), sizeof(float) * axis.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.pa, h->h_comp.data(), sizeof(float) * P_NCOMP * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.pa_id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(psc, h->h_pscene.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(act, h->h_active.data(), n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(bsc, bscene.data(), sizeof(int) * h->nkeys, hipMemcpyHostToDevice, h->stream));
  // ---- knife state of the first step (0, or a checkpoint's), first scatter and grid update
  h->step = h->resume ? h->resume_step : 0;
  h->time = h->resume ? h->resume_time : 0.0;
  h->time_f = h->resume ? h->resume_time_f : 0.0f;
  h->frame = h->resume ? h->resume_frame : 0;
  h->parity_swapped = false;
  if ((rc = upload_kstates(h, h->step, 0, h->time_f)) != 0) return rc;
  if (!h->tau_given) launch_init_tau(d, n, h->stream);
  h->resume = h->tau_given = false;
  launch_keys(d, n, h->stream);
  launch_scan(d, h->nkeys, 0, h->stream);
  launch_perm(d, n, grid_grid(h), 0, h->stream);
  launch_p2g(d, grid_p2g(h), 0, h->stream);
  launch_grid(d, grid_grid(h), h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (Scene& s : h->scenes) s.has_override = false;
  if ((rc = read_stats(h)) != 0) return rc;
  std::memcpy(h->stats_prev, h->h_stats, sizeof h->stats_prev);
  h->prepared = true;
  if ((rc = closure_attach(h->closure, h->d, h->stream, h->err)) != 0) return rc;
  return MPMX_OK;
}

int mpmx_get_lattice(mpmx_handle* h, int* dims, float* origin, float* xs, float* ys, float* zs) {
  if (!h || !dims) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "get_lattice: prepare first");
  const Scene& s = h->scenes[0];
  for (int d = 0; d < 3; ++d) {
    dims[d] = s.dims[d];
    if (origin) origin[d] = s.origin[d];
  }
  if (xs) std::copy(s.xs.begin(), s.xs.end(), xs);
  if (ys) std::copy(s.ys.begin(), s.ys.end(), ys);
  if (zs) std::copy(s.zs.begin(), s.zs.end(), zs);
  return MPMX_OK;
}

int mpmx_step(mpmx_handle* h, int nsteps) {
  if (!h || nsteps < 0) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "step: prepare first");
  if (learned(h) && !h->closure.d_blob)
    return fail(h, MPMX_ERR_STATE, "step: learned closure selected but no network parameters set");
  if (nsteps == 0) return MPMX_OK;
  int rc;
  auto t0 = std::chrono::steady_clock::now();
  if ((rc = build_graph(h)) != 0) return rc;
  int done = 0;
  while (done < nsteps) {
    const int chunk = std::min(nsteps - done, MAX_CHUNK);
    if ((rc = upload_kstates(h, h->step, chunk, h->time_f)) != 0) return rc;
    for (Scene& s : h->scenes) s.has_override = false;
    int i = 0;
    // the captured graph always starts from the A side; use it in pairs
    while (i < chunk) {
      if (!h->parity_swapped && chunk - i >= 2) {
        HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
        i += 2;
      } else {
        enqueue_step(h, current(h));
        h->parity_swapped = !h->parity_swapped;
        i += 1;
      }
    }
    for (int k = 0; k < chunk; ++k) h->time_f += h->scenes[0].cfg.dt;
    h->step += chunk;
    done += chunk;
  }
  HIPCHK(h, hipGetLastError());
  if ((rc = read_stats(h)) != 0) return rc;
  auto t1 = std::chrono::steady_clock::now();
  h->ms_last = (float)(std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps);
  for (int i = 0; i < ST_N; ++i) {
    h->stats_last[i] = h->h_stats[i] - h->stats_prev[i];
    h->stats_prev[i] = h->h_stats[i];
  }
  h->steps_last = nsteps;
  h->time = h->time_f;
  if (*h->h_err) return fail(h, MPMX_ERR_DEVICE, "device status word set (reference exit = 1)");
  return MPMX_OK;
}

int mpmx_frame_finished(mpmx_handle* h, int* finished) {
  if (!h || !finished) return MPMX_ERR_ARG;
  const float frame_dt = 1.0f / h->scenes[0].cfg.fps;
  const float min_dt = 1e-14f;
  *finished = 0;
  if (frame_dt * (h->frame + 1) - h->time_f < min_dt * 1.1f) {
    h->frame++;
    *finished = 1;
  }
  return MPMX_OK;
}

int mpmx_synchronize(mpmx_handle* h) {
  if (!h) return MPMX_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return MPMX_OK;
}

int mpmx_download_particles(mpmx_handle* h, float* x, float* v, float* F, float* C, float* pig, float* diff,
                            float* eps, float* dg, uint8_t* active) {
  if (!h) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "download: prepare first");
  const int64_t n = h->np;
  const DevPtrs d = current(h);
  float* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, sizeof(float) * 9 * n));
  struct Item { float* out; int first, ncomp; } items[] = {
      {x, P_X, 3}, {v, P_V, 3}, {F, P_F, 9}, {C, P_C, 9}, {pig, P_PIG, 7}, {diff, P_DIFF, 1}, {eps, P_EPS, 1},
      {dg, P_DG, 1}};
  int rc = MPMX_OK;
  for (const Item& it : items) {
    if (!it.out) continue;
    launch_unsort(d, n, tmp, it.first, it.ncomp, h->stream);
    if (hipMemcpyAsync(it.out, tmp, sizeof(float) * it.ncomp * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      rc = fail(h, MPMX_ERR_HIP, "download copy failed");
      break;
    }
  }
  if (rc == MPMX_OK && active) {
    launch_active_out(d, n, (uint8_t*)tmp, h->stream);
    if (hipMemcpyAsync(active, tmp, n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(h, MPMX_ERR_HIP, "download copy failed");
  }
  (void)hipFree(tmp);
  return rc;
}

int mpmx_download_plastic_state(mpmx_handle* h, float* eps_pl_vol, float* eps_pl_vol_pradhana) {
  if (!h) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "download: prepare first");
  const int64_t n = h->np;
  const DevPtrs d = current(h);
  float* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, sizeof(float) * n));
  struct Item { float* out; int row; } items[] = {{eps_pl_vol, P_EPSV}, {eps_pl_vol_pradhana, P_PRAD}};
  int rc = MPMX_OK;
  for (const Item& it : items) {
    if (!it.out) continue;
    launch_unsort(d, n, tmp, it.row, 1, h->stream);
    if (hipMemcpyAsync(it.out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      rc = fail(h, MPMX_ERR_HIP, "download copy failed");
      break;
    }
  }
  (void)hipFree(tmp);
  return rc;
}

namespace {
// Particle rows [first, first + ncomp) in upload order (interleaved per particle).
int download_rows(mpmx_handle* h, int first, int ncomp, float* out) {
  const int64_t n = h->np;
  float* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, sizeof(float) * ncomp * n));
  launch_unsort(current(h), n, tmp, first, ncomp, h->stream);
  int rc = MPMX_OK;
  if (hipMemcpyAsync(out, tmp, sizeof(float) * ncomp * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    rc = fail(h, MPMX_ERR_HIP, "download copy failed");
  (void)hipFree(tmp);
  return rc;
}

// Checkpoint file: little-endian binary.
//   char[8] "MPMXCKP2"; u32 sizeof(mpmx_config); mpmx_config (origin_mode 1, the
//   lattice's origin and node counts); i64 n, step; f64 time; f32 time_f; i32 frame;
//   i32 n_plates; mpmx_plate[n_plates]; i32 has_knife; mpmx_knife (if any);
//   i32 n_frames; f32 fps; f32[7 n_frames];
//   f32 x[3n] v[3n] F[9n] C[9n] pig[7n] diff[n] eps_pl_dev[n] delta_gamma[n]
//   eps_pl_vol[n] eps_pl_vol_pradhana[n] tau[6n]; u8 active[n]   (upload order)
constexpr char CKPT_MAGIC[8] = {'M', 'P', 'M', 'X', 'C', 'K', 'P', '3'};

struct FileCloser {
  FILE* f;
  ~FileCloser() { if (f) std::fclose(f); }
};
}  // namespace

int mpmx_checkpoint_save(mpmx_handle* h, const char* path) {
  if (!h || !path) return MPMX_ERR_ARG;
  if (h->batch) return fail(h, MPMX_ERR_STATE, "checkpoint: one scene per handle");
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "checkpoint: prepare first");
  const int64_t n = h->np;
  std::vector<float> x(3 * n), v(3 * n), F(9 * n), C(9 * n), pig(7 * n), diff(n), eps(n), dg(n), epsv(n), prad(n),
      tau(6 * n);
  std::vector<uint8_t> act(n);
  int rc = mpmx_download_particles(h, x.data(), v.data(), F.data(), C.data(), pig.data(), diff.data(), eps.data(),
                                   dg.data(), act.data());
  if (rc == MPMX_OK) rc = mpmx_download_plastic_state(h, epsv.data(), prad.data());
  if (rc == MPMX_OK) rc = download_rows(h, P_TAU, 6, tau.data());
  if (rc != MPMX_OK) return rc;
  const Scene& s = h->scenes[0];
  mpmx_config cfg = s.cfg;
  cfg.origin_mode = 1;
  for (int d = 0; d < 3; ++d) { cfg.grid_origin[d] = s.origin[d]; cfg.grid_dims[d] = s.dims[d]; }
  FileCloser fc{std::fopen(path, "wb")};
  if (!fc.f) return fail(h, MPMX_ERR_IO, std::string("checkpoint: cannot write ") + path);
  bool ok = true;
  auto put = [&](const void* p, size_t bytes) { ok = ok && (bytes == 0 || std::fwrite(p, 1, bytes, fc.f) == bytes); };
  const uint32_t csize = sizeof(mpmx_config);
  const int32_t frame = h->frame, np = (int32_t)s.plates.size(), hk = s.has_knife ? 1 : 0;
  const int32_t nf = (int32_t)(s.frames.size() / 7);
  put(CKPT_MAGIC, 8); put(&csize, 4); put(&cfg, sizeof cfg);
  put(&n, 8); put(&h->step, 8); put(&h->time, 8); put(&h->time_f, 4); put(&frame, 4);
  put(&np, 4); put(s.plates.data(), sizeof(mpmx_plate) * np);
  put(&hk, 4); if (hk) put(&s.knife, sizeof(mpmx_knife));
  put(&nf, 4); put(&s.frame_fps, 4); put(s.frames.data(), sizeof(float) * s.frames.size());
  // the objects' clocks (re-based by dt changes)
  {
    Scene& sm = h->scenes[0];
    const std::vector<mpmx_plate>& base = sm.plates_base;
    put(&sm.knife_t0, 8); put(&sm.knife_s0, 8); put(&sm.plates_step0, 8); put(&sm.plates_time0, 4);
    put(base.data(), sizeof(mpmx_plate) * np);
  }
  for (const std::vector<float>* a : {&x, &v, &F, &C, &pig, &diff, &eps, &dg, &epsv, &prad, &tau})
    put(a->data(), sizeof(float) * a->size());
  put(act.data(), act.size());
  if (!ok) return fail(h, MPMX_ERR_IO, std::string("checkpoint: write failed: ") + path);
  return MPMX_OK;
}

int mpmx_checkpoint_load(const char* path, mpmx_handle** out) {
  if (!path || !out) return MPMX_ERR_ARG;
  *out = nullptr;
  FileCloser fc{std::fopen(path, "rb")};
  if (!fc.f) return MPMX_ERR_IO;
  bool ok = true;
  auto get = [&](void* p, size_t bytes) { ok = ok && (bytes == 0 || std::fread(p, 1, bytes, fc.f) == bytes); };
  char magic[8];
  uint32_t csize = 0;
  mpmx_config cfg;
  get(magic, 8); get(&csize, 4);
  if (!ok || std::memcmp(magic, CKPT_MAGIC, 8) != 0 || csize != sizeof(mpmx_config)) return MPMX_ERR_IO;
  int64_t n = 0, step = 0;
  double time = 0.0;
  float time_f = 0.0f;
  int32_t frame = 0, np = 0, hk = 0, nf = 0;
  float fps = 60.0f;
  get(&cfg, sizeof cfg); get(&n, 8); get(&step, 8); get(&time, 8); get(&time_f, 4); get(&frame, 4); get(&np, 4);
  if (!ok || n <= 0 || n > (int64_t)0x7fffffff / 2 || np < 0 || np > MAX_PLATES || step < 0) return MPMX_ERR_IO;
  std::vector<mpmx_plate> plates(np);
  mpmx_knife knife;
  get(plates.data(), sizeof(mpmx_plate) * np); get(&hk, 4);
  if (hk) get(&knife, sizeof knife);
  get(&nf, 4); get(&fps, 4);
  if (!ok || nf < 0 || nf > (1 << 24)) return MPMX_ERR_IO;
  std::vector<float> frames(7 * (size_t)nf);
  get(frames.data(), sizeof(float) * frames.size());
  double knife_t0 = 0.0;
  int64_t knife_s0 = 0, plates_step0 = 0;
  float plates_time0 = 0.0f;
  std::vector<mpmx_plate> plates_base(np);
  get(&knife_t0, 8); get(&knife_s0, 8); get(&plates_step0, 8); get(&plates_time0, 4);
  get(plates_base.data(), sizeof(mpmx_plate) * np);
  std::vector<float> x(3 * n), v(3 * n), F(9 * n), C(9 * n), pig(7 * n), diff(n), eps(n), dg(n), epsv(n), prad(n),
      tau(6 * n);
  std::vector<uint8_t> act(n);
  for (std::vector<float>* a : {&x, &v, &F, &C, &pig, &diff, &eps, &dg, &epsv, &prad, &tau})
    get(a->data(), sizeof(float) * a->size());
  get(act.data(), act.size());
  if (!ok) return MPMX_ERR_IO;
  mpmx_handle* h = nullptr;
  int rc = mpmx_create(&cfg, &h);
  if (rc != MPMX_OK) { mpmx_destroy(h); return rc; }
  if ((rc = mpmx_upload_particles(h, n, x.data(), v.data(), F.data(), C.data(), pig.data(), diff.data(), eps.data(),
                                  act.data())) != MPMX_OK ||
      (rc = mpmx_upload_plastic_state(h, 0, epsv.data(), prad.data())) != MPMX_OK ||
      (rc = mpmx_set_objects(h, plates.data(), np, hk ? &knife : nullptr, hk)) != MPMX_OK ||
      (nf > 0 && (rc = mpmx_set_knife_keyframes(h, 0, frames.data(), nf, fps)) != MPMX_OK)) {
    mpmx_destroy(h);
    return rc;
  }
  {
    Scene& sm = h->scenes[0];
    sm.knife_t0 = knife_t0;
    sm.knife_s0 = knife_s0;
    sm.plates_base = plates_base;
    sm.plates_step0 = plates_step0;
    sm.plates_time0 = plates_time0;
  }
  for (int64_t p = 0; p < n; ++p) {
    h->h_comp[(size_t)P_DG * n + p] = dg[p];
    for (int c = 0; c < 6; ++c) h->h_comp[(size_t)(P_TAU + c) * n + p] = tau[6 * p + c];
  }
  h->tau_given = true;
  h->resume = true;
  h->resume_step = step;
  h->resume_time = time;
  h->resume_time_f = time_f;
  h->resume_frame = frame;
  if ((rc = mpmx_prepare(h)) != MPMX_OK) { mpmx_destroy(h); return rc; }
  *out = h;
  return MPMX_OK;
}

namespace {
// Binary little-endian PLY with one `vertex` element of float properties,
// values interleaved per vertex (the layout tinyply writes for the
// reference's add_properties_to_element calls).
int write_ply(mpmx_handle* h, const char* path, const std::vector<std::string>& names, int64_t n,
              const std::vector<const float*>& cols, const std::vector<int>& strides) {
  FileCloser fc{std::fopen(path, "wb")};
  if (!fc.f) return fail(h, MPMX_ERR_IO, std::string("ply: cannot write ") + path);
  std::string hdr = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
  for (const std::string& nm : names) hdr += "property float " + nm + "\n";
  hdr += "end_header\n";
  bool ok = std::fwrite(hdr.data(), 1, hdr.size(), fc.f) == hdr.size();
  const size_t nc = names.size();
  std::vector<float> row(nc * 4096);
  for (int64_t p0 = 0; ok && p0 < n; p0 += 4096) {
    const int64_t m = std::min<int64_t>(4096, n - p0);
    for (int64_t p = 0; p < m; ++p)
      for (size_t c = 0; c < nc; ++c) row[p * nc + c] = cols[c][(p0 + p) * strides[c]];
    ok = std::fwrite(row.data(), sizeof(float) * nc, m, fc.f) == (size_t)m;
  }
  if (!ok) return fail(h, MPMX_ERR_IO, std::string("ply: write failed: ") + path);
  return MPMX_OK;
}
}  // namespace

int mpmx_write_particles_ply(mpmx_handle* h, const char* path) {
  if (!h || !path) return MPMX_ERR_ARG;
  if (h->batch) return fail(h, MPMX_ERR_STATE, "ply: one scene per handle");
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "ply: prepare first");
  const int64_t n = h->np;
  std::vector<float> x(3 * n), v(3 * n), F(9 * n), eps(n), dg(n), epsv(n), tau(6 * n);
  int rc = mpmx_download_particles(h, x.data(), v.data(), F.data(), nullptr, nullptr, nullptr, eps.data(), dg.data(),
                                   nullptr);
  if (rc == MPMX_OK) rc = mpmx_download_plastic_state(h, epsv.data(), nullptr);
  if (rc == MPMX_OK) rc = download_rows(h, P_TAU, 6, tau.data());
  if (rc != MPMX_OK) return rc;
  // pressure, deviatoric stress and elastic volume ratio (save_data.cpp:13-31);
  // tau is the Kirchhoff stress of the particle's current F
  const float q_prefac = h->scenes[0].cfg.q_prefac;
  std::vector<float> pr(n), q(n), je(n);
  for (int64_t p = 0; p < n; ++p) {
    const float* t = &tau[6 * p];   // xx yy zz xy xz yz
    const float pressure = -((t[0] + t[1]) + t[2]) / 3;
    const float a = t[0] + pressure, b = t[1] + pressure, c = t[2] + pressure;
    const float dd = a * a + t[3] * t[3] + t[4] * t[4] + t[3] * t[3] + b * b + t[5] * t[5] + t[4] * t[4] +
                     t[5] * t[5] + c * c;
    pr[p] = pressure;
    q[p] = q_prefac * std::sqrt(dd);
    const float* f = &F[9 * p];
    je[p] = f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) + f[2] * (f[3] * f[7] - f[4] * f[6]);
  }
  std::vector<std::string> names = {"x", "y", "z", "vx", "vy", "vz", "p", "q", "Je"};
  std::vector<const float*> cols = {&x[0], &x[1], &x[2], &v[0], &v[1], &v[2], pr.data(), q.data(), je.data()};
  std::vector<int> strides = {3, 3, 3, 3, 3, 3, 1, 1, 1};
  if (h->scenes[0].cfg.plastic_model != MPMX_PLASTIC_NONE) {
    for (const char* nm : {"eps_pl_vol", "eps_pl_dev", "delta_gamma"}) names.push_back(nm);
    for (const float* c : {epsv.data(), eps.data(), dg.data()}) { cols.push_back(c); strides.push_back(1); }
  }
  return write_ply(h, path, names, n, cols, strides);
}

int mpmx_write_grid_ply(mpmx_handle* h, const char* path) {
  if (!h || !path) return MPMX_ERR_ARG;
  if (h->batch) return fail(h, MPMX_ERR_STATE, "ply: one scene per handle");
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "ply: prepare first");
  const Scene& s = h->scenes[0];
  const int64_t nn = (int64_t)s.dims[0] * s.dims[1] * s.dims[2];
  std::vector<float> mass(nn), v(3 * nn), X(3 * nn);
  int rc = mpmx_get_grid(h, mass.data(), v.data(), nullptr, nullptr, nullptr);
  if (rc != MPMX_OK) return rc;
  int64_t c = 0;   // nodes in Simulation::ind order (i, j, k), k fastest (save_data.cpp:176-189)
  for (int i = 0; i < s.dims[0]; ++i)
    for (int j = 0; j < s.dims[1]; ++j)
      for (int k = 0; k < s.dims[2]; ++k, ++c) {
        X[3 * c] = s.xs[i];
        X[3 * c + 1] = s.ys[j];
        X[3 * c + 2] = s.zs[k];
      }
  return write_ply(h, path, {"x", "y", "z", "vx", "vy", "vz", "mass"}, nn,
                   {&X[0], &X[1], &X[2], &v[0], &v[1], &v[2], mass.data()}, {3, 3, 3, 3, 3, 3, 1});
}

int mpmx_pack_render_device(mpmx_handle* h, void* d_pos4, void* d_lat7, void* d_diff) {
  if (!h || !d_pos4 || !d_lat7 || !d_diff) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "pack_render: prepare first");
  launch_pack_render(current(h), h->np, (float*)d_pos4, (float*)d_lat7, (float*)d_diff, h->stream);
  HIPCHK(h, hipGetLastError());
  return MPMX_OK;
}

int mpmx_pack_render(mpmx_handle* h, float* pos4, float* lat7, float* diff) {
  if (!h || !pos4 || !lat7 || !diff) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "pack_render: prepare first");
  const int64_t n = h->np;
  float* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, sizeof(float) * 12 * n));
  launch_pack_render(current(h), n, tmp, tmp + 4 * n, tmp + 11 * n, h->stream);
  int rc = MPMX_OK;
  if (hipMemcpyAsync(pos4, tmp, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(lat7, tmp + 4 * n, sizeof(float) * 7 * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(diff, tmp + 11 * n, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    rc = fail(h, MPMX_ERR_HIP, "pack_render copy failed");
  (void)hipFree(tmp);
  return rc;
}

int mpmx_get_grid(mpmx_handle* h, float* mass, float* v, float* flip, float* pig, float* shear) {
  if (!h) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "get_grid: prepare first");
  const Scene& s = h->scenes[0];
  const int64_t nn = (int64_t)s.dims[0] * s.dims[1] * s.dims[2];
  float* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, sizeof(float) * 15 * nn));
  launch_grid_dense(current(h), 0, tmp, tmp + nn, tmp + 4 * nn, tmp + 7 * nn, tmp + 14 * nn, h->stream);
  int rc = MPMX_OK;
  struct Item { float* out; int64_t off, cnt; } items[] = {
      {mass, 0, nn}, {v, nn, 3 * nn}, {flip, 4 * nn, 3 * nn}, {pig, 7 * nn, 7 * nn}, {shear, 14 * nn, nn}};
  for (const Item& it : items) {
    if (!it.out) continue;
    if (hipMemcpyAsync(it.out, tmp + it.off, sizeof(float) * it.cnt, hipMemcpyDeviceToHost, h->stream) != hipSuccess) {
      rc = fail(h, MPMX_ERR_HIP, "get_grid copy failed");
      break;
    }
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(h, MPMX_ERR_HIP, "get_grid sync failed");
  (void)hipFree(tmp);
  return rc;
}

int mpmx_stats_get(mpmx_handle* h, mpmx_stats* o) {
  if (!h || !o) return MPMX_ERR_ARG;
  std::memset(o, 0, sizeof *o);
  o->step = h->step;
  o->time = h->time_f;
  o->n_particles = h->np;
  if (h->prepared) {
    int rc = read_stats(h);
    if (rc) return rc;
    o->n_active = h->h_counters[CNT_LIVE];
    const int64_t st = std::max<int64_t>(h->steps_last, 1);
    o->plastic_count = (int64_t)h->h_stats[ST_PLASTIC_LAST];
    o->fallback_count = (int64_t)h->h_stats[ST_FALLBACK_LAST];
    o->active_nodes = (int64_t)(h->stats_last[ST_ACTIVE_NODES] / st);
    o->active_blocks = (int64_t)(h->stats_last[ST_NBLOCKS] / st);
    o->stray_count = (int64_t)h->stats_last[ST_STRAY];
    o->device_error = *h->h_err;
    o->learned_count = (int64_t)h->h_stats[ST_LEARNED_LAST];
    o->newton_iters = (int64_t)h->stats_last[ST_NEWTON];
  }
  o->ms_last_step = h->ms_last;
  return MPMX_OK;
}


// ---------------------------------------------------------- learned closure
int mpmx_closure_set_params(mpmx_handle* h, const float* params, int64_t n) {
  if (!h || !params || n != mpmx_closure_n_params()) return MPMX_ERR_ARG;
  h->mlp_params.assign(params, params + n);
  return closure_upload(h->closure, h->mlp_params, h->in_mean, h->in_std, h->out_mean, h->out_std, h->stream,
                        h->err);
}

int mpmx_closure_get_params(mpmx_handle* h, float* params, int64_t n) {
  if (!h || !params || n != mpmx_closure_n_params()) return MPMX_ERR_ARG;
  if ((int64_t)h->mlp_params.size() != n) return fail(h, MPMX_ERR_STATE, "closure: no parameters set");
  std::copy(h->mlp_params.begin(), h->mlp_params.end(), params);
  return MPMX_OK;
}

int mpmx_closure_set_norm(mpmx_handle* h, const float* in_mean, const float* in_std, const float* out_mean,
                          const float* out_std) {
  if (!h || !in_mean || !in_std || !out_mean || !out_std) return MPMX_ERR_ARG;
  for (int i = 0; i < 8; ++i) {
    if (!(in_std[i] > 0)) return fail(h, MPMX_ERR_ARG, "closure: in_std must be > 0");
    h->in_mean[i] = in_mean[i];
    h->in_std[i] = in_std[i];
  }
  for (int i = 0; i < 4; ++i) {
    if (!(out_std[i] > 0)) return fail(h, MPMX_ERR_ARG, "closure: out_std must be > 0");
    h->out_mean[i] = out_mean[i];
    h->out_std[i] = out_std[i];
  }
  if (h->mlp_params.empty()) return MPMX_OK;
  return closure_upload(h->closure, h->mlp_params, h->in_mean, h->in_std, h->out_mean, h->out_std, h->stream,
                        h->err);
}

int mpmx_closure_train(mpmx_handle* h, const float* records12, int64_t n, const mpmx_train_config* tc,
                       mpmx_train_report* report, float* curve, int64_t curve_cap) {
  if (!h || !records12 || !tc || n <= 0 || curve_cap < 0 || (curve_cap > 0 && !curve)) return MPMX_ERR_ARG;
  float n24[24];
  const int rc = closure_train(h->mlp_params, n24, records12, n, *tc, report, curve, curve_cap, h->stream, h->err);
  if (rc) return rc;
  std::memcpy(h->in_mean, n24, 32);
  std::memcpy(h->in_std, n24 + 8, 32);
  std::memcpy(h->out_mean, n24 + 16, 16);
  std::memcpy(h->out_std, n24 + 20, 16);
  return closure_upload(h->closure, h->mlp_params, h->in_mean, h->in_std, h->out_mean, h->out_std, h->stream,
                        h->err);
}

int mpmx_closure_eval(mpmx_handle* h, const float* trial4, int64_t n, float* learned4, float* exact4,
                      uint8_t* accepted) {
  if (!h || n < 0 || (n > 0 && (!trial4 || !learned4 || !exact4 || !accepted))) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "closure_eval: prepare first");
  if (!learned(h) || !h->closure.d_blob)
    return fail(h, MPMX_ERR_STATE, "closure_eval: needs a learned closure mode and network parameters");
  if (n == 0) return MPMX_OK;
  float *dt = nullptr, *dl = nullptr, *de = nullptr;
  uint8_t* da = nullptr;
  int rc = MPMX_OK;
  if (hipMalloc(&dt, sizeof(float) * 4 * n) != hipSuccess || hipMalloc(&dl, sizeof(float) * 4 * n) != hipSuccess ||
      hipMalloc(&de, sizeof(float) * 4 * n) != hipSuccess || hipMalloc(&da, n) != hipSuccess) {
    rc = fail(h, MPMX_ERR_HIP, "closure_eval: hipMalloc failed");
  } else if (hipMemcpyAsync(dt, trial4, sizeof(float) * 4 * n, hipMemcpyHostToDevice, h->stream) != hipSuccess) {
    rc = fail(h, MPMX_ERR_HIP, "closure_eval: upload failed");
  } else {
    launch_closure_eval(h->d_scenes, h->closure.d_blob, dt, n, dl, de, da, h->n_cu * 2, h->stream);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(learned4, dl, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(exact4, de, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(accepted, da, n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(h, MPMX_ERR_HIP, "closure_eval: kernel or copy failed");
  }
  if (dt) (void)hipFree(dt);
  if (dl) (void)hipFree(dl);
  if (de) (void)hipFree(de);
  if (da) (void)hipFree(da);
  return rc;
}

int mpmx_harvest_begin(mpmx_handle* h, int64_t capacity) {
  if (!h || capacity <= 0) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "harvest: prepare first");
  if (h->d_harvest) { (void)hipFree(h->d_harvest); h->d_harvest = nullptr; }
  HIPCHK(h, hipMalloc(&h->d_harvest, sizeof(float) * 12 * (size_t)capacity));
  h->harvest_cap = capacity;
  int rc = read_stats(h);
  if (rc) return rc;
  h->harvest_base = h->h_stats[ST_HARVEST];
  // records are appended from slot 0: reset the device counter
  unsigned long long zero = 0;
  HIPCHK(h, hipMemcpy(&h->d.stats[ST_HARVEST], &zero, sizeof zero, hipMemcpyHostToDevice));
  h->d.harvest = h->d_harvest;
  h->d.harvest_cap = capacity;
  if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  return MPMX_OK;
}

int mpmx_harvest_end(mpmx_handle* h, float* out12, int64_t max_records, int64_t* n_out) {
  if (!h || !n_out || max_records < 0 || (max_records > 0 && !out12)) return MPMX_ERR_ARG;
  if (!h->d_harvest) return fail(h, MPMX_ERR_STATE, "harvest: not started");
  int rc = read_stats(h);
  if (rc) return rc;
  const int64_t got = std::min<int64_t>((int64_t)h->h_stats[ST_HARVEST], h->harvest_cap);
  const int64_t n = std::min(got, max_records);
  if (n > 0) HIPCHK(h, hipMemcpy(out12, h->d_harvest, sizeof(float) * 12 * (size_t)n, hipMemcpyDeviceToHost));
  *n_out = n;
  (void)hipFree(h->d_harvest);
  h->d_harvest = nullptr;
  h->harvest_cap = 0;
  h->d.harvest = nullptr;
  h->d.harvest_cap = 0;
  if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  return MPMX_OK;
}


int mpmx_profile_steps(mpmx_handle* h, int nsteps, mpmx_kernel_times* out) {
  if (!h || !out || nsteps <= 0) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "profile: prepare first");
  if (learned(h) && !h->closure.d_blob)
    return fail(h, MPMX_ERR_STATE, "profile: learned closure selected but no network parameters set");
  std::memset(out, 0, sizeof *out);
  int rc = read_stats(h);
  if (rc) return rc;
  const unsigned long long nodes0 = h->h_stats[ST_ACTIVE_NODES];
  const unsigned long long queued0 = h->h_stats[ST_QUEUED];
  hipEvent_t ev[6];
  for (auto& e : ev) HIPCHK(h, hipEventCreate(&e));
  double t_f = 0, t_c = 0, t_s = 0, t_p = 0, t_g = 0, t_all = 0, live = 0;
  // t_f: k_g2p (+ dead copy, MUSL halves), t_c: k_closure, t_s: sort (scans,
  // permutation, in-cell ranking), t_p: k_p2g, t_g: k_grid
  int done = 0;
  while (done < nsteps) {
    const int chunk = std::min(nsteps - done, MAX_CHUNK);
    if ((rc = upload_kstates(h, h->step, chunk, h->time_f)) != 0) return rc;
    for (int i = 0; i < chunk; ++i) {
      const DevPtrs d = current(h);
      HIPCHK(h, hipEventRecord(ev[0], h->stream));
      enqueue_g2p(h, d, ev[5]);
      HIPCHK(h, hipEventRecord(ev[1], h->stream));
      launch_scan(d, h->nkeys, 1, h->stream);
      launch_perm(d, h->np, grid_grid(h), 1, h->stream);
      HIPCHK(h, hipEventRecord(ev[2], h->stream));
      launch_p2g(d, grid_p2g(h), 1, h->stream);
      HIPCHK(h, hipEventRecord(ev[3], h->stream));
      launch_grid(d, grid_grid(h), h->stream);
      HIPCHK(h, hipEventRecord(ev[4], h->stream));
      HIPCHK(h, hipEventSynchronize(ev[4]));
      float a, b, c, e, all, cl;
      HIPCHK(h, hipEventElapsedTime(&a, ev[0], ev[5]));
      HIPCHK(h, hipEventElapsedTime(&cl, ev[5], ev[1]));
      HIPCHK(h, hipEventElapsedTime(&b, ev[1], ev[2]));
      HIPCHK(h, hipEventElapsedTime(&c, ev[2], ev[3]));
      HIPCHK(h, hipEventElapsedTime(&e, ev[3], ev[4]));
      HIPCHK(h, hipEventElapsedTime(&all, ev[0], ev[4]));
      t_f += a; t_c += cl; t_s += b; t_p += c; t_g += e; t_all += all;
      h->parity_swapped = !h->parity_swapped;
      h->time_f += h->scenes[0].cfg.dt;
      h->step += 1;
      HIPCHK(h, hipMemcpy(h->h_counters, h->d.counters, sizeof(int) * CNT_NCOUNTERS, hipMemcpyDeviceToHost));
      live += h->h_counters[CNT_LIVE];
    }
    done += chunk;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if ((rc = read_stats(h)) != 0) return rc;
  out->steps = nsteps;
  out->ms_g2p = (float)(t_f / nsteps);
  out->ms_closure = learned(h) ? (float)(t_c / nsteps) : 0.0f;
  out->ms_sort = (float)(t_s / nsteps);
  out->ms_p2g = (float)(t_p / nsteps);
  out->ms_grid = (float)(t_g / nsteps);
  out->ms_step = (float)(t_all / nsteps);
  out->particles_per_step = live / nsteps;
  out->active_nodes_per_step = (double)(h->h_stats[ST_ACTIVE_NODES] - nodes0) / nsteps;
  out->queued_per_step = (double)(h->h_stats[ST_QUEUED] - queued0) / nsteps;
  for (int i = 0; i < ST_N; ++i) h->stats_prev[i] = h->h_stats[i];
  h->time = h->time_f;
  return MPMX_OK;
}

}  // extern "C"

// Launch shapes set by the caller (e.g. the result of an earlier autotune of
// the same workload); the step graph is rebuilt at the next mpmx_step.
int mpmx_set_launch_shape(mpmx_handle* h, int wg_p2g, int wg_g2p, int wg_grid) {
  if (!h) return MPMX_ERR_ARG;
  if (wg_p2g < 1 || wg_p2g > 64 || wg_g2p < 1 || wg_g2p > 64 || wg_grid < 1 || wg_grid > 64)
    return fail(h, MPMX_ERR_ARG, "set_launch_shape: workgroups per CU must be in 1..64");
  h->wg_p2g = wg_p2g;
  h->wg_g2p = wg_g2p;
  h->wg_grid = wg_grid;
  if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  return MPMX_OK;
}

// ---------------------------------------------------------- launch-shape autotuner
// Coordinate descent over workgroups per CU: P2G (one-wavefront workgroups),
// then G2P (256-thread workgroups), then the grid update.  Each candidate is timed with HIP events over
// `steps_per_trial` ordinary steps (the simulation advances; the launch shape
// changes no result beyond FP32 atomic order), and the fastest is kept.  The
// step graph is rebuilt with the chosen shapes at the next mpmx_step.
int mpmx_autotune(mpmx_handle* h, int steps_per_trial, mpmx_tune_report* rep) {
  if (!h || steps_per_trial <= 0) return MPMX_ERR_ARG;
  if (!h->prepared) return fail(h, MPMX_ERR_STATE, "autotune: prepare first");
  // P2G: up to its occupancy (its static first items assume resident wavefronts)
  std::vector<int> cand_p2g;
  for (int c : {h->p2g_resident / 2, (2 * h->p2g_resident) / 3, h->p2g_resident})
    if (c >= 1 && (cand_p2g.empty() || c != cand_p2g.back())) cand_p2g.push_back(c);
  // G2P: up to its occupancy as well (static first blocks)
  std::vector<int> cand_g2p;
  for (int c : {2, 3, 4, 6, 8, 12, 16})
    if (c <= h->g2p_resident) cand_g2p.push_back(c);
  const std::vector<int> cand_grid = {2, 4, 8};
  int trials = 0;
  mpmx_kernel_times kt;
  auto sweep = [&](int* knob, const std::vector<int>& cands, float mpmx_kernel_times::*field, float* best_ms) -> int {
    int best = *knob;
    double best_t = 1e30;
    for (int c : cands) {
      *knob = c;
      int rc = mpmx_profile_steps(h, steps_per_trial, &kt);
      if (rc) return rc;
      ++trials;
      if (kt.*field < best_t) {
        best_t = kt.*field;
        best = c;
      }
    }
    *knob = best;
    if (best_ms) *best_ms = (float)best_t;
    return MPMX_OK;
  };
  float ms_p2g = 0, ms_g2p = 0, ms_grid = 0;
  int rc;
  if ((rc = sweep(&h->wg_p2g, cand_p2g, &mpmx_kernel_times::ms_p2g, &ms_p2g))) return rc;
  if ((rc = sweep(&h->wg_g2p, cand_g2p, &mpmx_kernel_times::ms_g2p, &ms_g2p))) return rc;
  if ((rc = sweep(&h->wg_grid, cand_grid, &mpmx_kernel_times::ms_grid, &ms_grid))) return rc;
  if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  if (rep) {
    rep->wg_p2g = h->wg_p2g;
    rep->wg_g2p = h->wg_g2p;
    rep->wg_grid = h->wg_grid;
    rep->ms_p2g = ms_p2g;
    rep->ms_g2p = ms_g2p;
    rep->ms_grid = ms_grid;
    rep->trials = trials;
    rep->steps = (int64_t)trials * steps_per_trial;
  }
  return MPMX_OK;
}
