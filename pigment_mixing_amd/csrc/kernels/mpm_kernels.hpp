// mpm_kernels.hpp — device pointer bundle and host-side launchers of the step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../common/mpmx_device.hpp"

namespace mpmx {

// particle component rows (each a contiguous array of `cap` floats)
enum {
  P_X = 0, P_V = 3, P_C = 6, P_F = 15, P_PIG = 24, P_DIFF = 31, P_EPS = 32, P_DG = 33,
  P_TAU = 34,  // symmetric Kirchhoff stress of the current F: xx, yy, zz, xy, xz, yz
  P_EPSV = 40,  // eps_pl_vol (data_structures.hpp:30), carried where a model tracks it
  P_PRAD = 41,  // eps_pl_vol_pradhana (data_structures.hpp:39), DPVisc's volume correction
  P_NCOMP = 42
};
constexpr int GV_STRIDE = 16;  // floats per node in the gather grid (v3 flip3 pig7 shear mass pad)

// int32 counters
enum {
  CNT_PBLOCKS = 0,      // non-empty live particle blocks (current sort)
  CNT_LIVE = 1,         // live particles
  CNT_STEP = 2,         // step index within the uploaded knife-state chunk
  CNT_GSTEP = 3,        // global step count (node-block list parity)
  CNT_NBLOCKS0 = 4,     // node-block list lengths (2 parities)
  CNT_NBLOCKS1 = 5,
  CNT_NCOUNTERS = 8
};
// uint64 cumulative statistics
enum {
  ST_PLASTIC = 0, ST_ACTIVE_NODES = 1, ST_STRAY = 2, ST_FALLBACK = 3, ST_NBLOCKS = 4, ST_HARVEST = 5,
  ST_PLASTIC_LAST = 6, ST_FALLBACK_LAST = 7, ST_PLASTIC_CUR = 8, ST_FALLBACK_CUR = 9,
  ST_NEWTON = 10,       // Newton iterations of the exact map (cumulative)
  ST_LEARNED = 11,      // learned-closure predictions accepted (cumulative)
  ST_LEARNED_LAST = 12, ST_LEARNED_CUR = 13,
  ST_QUEUED = 14,       // closure-queue records (cumulative)
  ST_N = 16
};

struct DevPtrs {
  const SceneDev* scenes;
  const KnifeStateDev* kstate;
  const float* axis_coords;
  float* pa;
  float* pb;
  int* pa_id;
  int* pb_id;
  const int* pscene;
  const uint8_t* active_in;
  int* key_out;
  int* rank_out;
  int* cell_out;           // cell (0..63) of the particle in its block, -1 if dead
  int* okey_out;           // order key inside the cell (cell_order_key), live particles
  int* perm;
  int* kslot;              // k_cellsort's bucket array for blocks past its LDS capacity (by slot)
  int* cmaj;               // live particles in their block's range by block rank (k_perm -> k_cellsort)
  int* cell_num;           // per (listed block, cell) counts of the current sort, list order
  int4* pdesc;             // per listed block: key, first slot, particle count, scene
  float* tile_acc;         // per listed block (< tile_cap): merged 6x6x6 node tile, [ACC_CH][216]
  int* list_pos;           // block key -> position in pb_list (listed blocks only)
  int* count_next;
  int* count_cur;
  int* offsets;
  int* pb_list;
  unsigned long long* scan_partials;
  const int* blk_scene;
  float* acc;
  float* gv;
  int* nb_flag;
  int* nb_list;
  int* nb_live;
  int* counters;
  unsigned long long* stats;
  uint32_t* err;
  float* harvest;          // (trial, corrected) records of 12 floats, or null
  int64_t harvest_cap;
  float4* cq;              // closure queue: CQ_REC float4 per deferred plastic candidate (cq_word), or null
  int64_t cq_cap;          // records it holds (the particle capacity)
  int* cq_cnt;             // records appended per XCD segment this step (CQ_SEGS, CQ_CNT_STRIDE apart)
  int* cq_cnt_prev;        // the same counts of the step before (the other of two buffers, which
                           // alternate with the particle buffers): the sort clears them
  int* cq_start;           // this step's first queue slot per XCD segment (CQ_SEGS), written by G2P
  int* wq;                 // work queues of P2G (items) and G2P (blocks): per XCD the next claim
                           // and the finished walkers (4 * CQ_SEGS counters, CQ_CNT_STRIDE apart;
                           // zero between launches)
  int64_t cap;
  int n_scenes;
  int nb_list_cap;
  int dead_key;
  int tile_cap;            // listed blocks with a tile; later ones scatter with atomics
  int track_vol;           // some scene carries the eps_pl_vol / pradhana rows (G2P variant)
};
// float4 per closure-queue record: the slot, scene, eps_pl_dev, the trial F and
// its principal frame from G2P (lam, V), so that the closure does not decompose
// F again: {slot, scene, eps, lam0} {lam1, lam2, F0, F1} {F2..F5} {F6, F7, F8, V0}
// {V1..V4} {V5..V8}
constexpr int CQ_REC = 6;
// The records are stored in groups of 64 slots, each float4 of the group's
// records contiguous: word f of slot s at ((s / 64) * CQ_REC + f) * 64 + s % 64.
// A wavefront's appends (consecutive slots) and the closure's reads (one slot
// per lane) then move whole runs of 16-byte words per instruction.
constexpr int CQ_GROUP = 64;
__host__ __device__ __forceinline__ int64_t cq_word(int64_t slot, int f) {   // slot >= 0
  const uint64_t u = (uint64_t)slot;
  return (int64_t)(((u / CQ_GROUP) * CQ_REC + (uint64_t)f) * CQ_GROUP + u % CQ_GROUP);
}
// The queue is split into one segment per XCD walker: segment x starts at the
// first particle slot of XCD x's contiguous share of the block list and can
// hold every particle of that share, so it never overflows.
constexpr int CQ_SEGS = 8;
constexpr int CQ_CNT_STRIDE = 32;      // ints: one 128-byte line per segment counter
constexpr int PT = 6;                  // scatter tile edge (nodes) = 4 bins + 2
constexpr int PT3 = PT * PT * PT;      // 216

// Stored order of a scatter tile: node (tx, ty, tz) of the 6x6x6 tile belongs
// to the node block at offset (ox, oy, oz) = (tx, ty, tz) >> 2 from the tile's
// particle block.  The 8 groups are stored one after another (offsets 000,
// 001, ..., 111), each its 4- or 2-node-per-axis sub-box in row-major order, so
// the grid kernel reads the group of every node block as one contiguous run.
MPMX_HD int tile_group_base(int ox, int oy, int oz) {
  const int ex = 4 - 2 * ox, ey = 4 - 2 * oy;
  return 144 * ox + 24 * ex * oy + 4 * ex * ey * oz;
}
MPMX_HD int tile_pos(int tx, int ty, int tz) {
  const int ox = tx >> 2, oy = ty >> 2, oz = tz >> 2;
  const int ey = 4 - 2 * oy, ez = 4 - 2 * oz;
  return tile_group_base(ox, oy, oz) + ((tx - 4 * ox) * ey + (ty - 4 * oy)) * ez + (tz - 4 * oz);
}
// Inverse of tile_pos for a stored position p in [0, 216).
MPMX_HD void tile_node_at(int p, int& tx, int& ty, int& tz) {
  tx = ty = tz = 0;
  for (int q = 0; q < 8; ++q) {
    const int ox = q >> 2, oy = (q >> 1) & 1, oz = q & 1;
    const int ex = 4 - 2 * ox, ey = 4 - 2 * oy, ez = 4 - 2 * oz;
    const int base = tile_group_base(ox, oy, oz);
    if (p >= base && p < base + ex * ey * ez) {
      const int l = p - base;
      tx = 4 * ox + l / (ey * ez);
      ty = 4 * oy + (l / ez) % ey;
      tz = 4 * oz + l % ez;
    }
  }
}

void launch_init_tau(const DevPtrs& d, int64_t n, hipStream_t st);
void launch_keys(const DevPtrs& d, int64_t n, hipStream_t st);
void launch_scan(const DevPtrs& d, int nkeys, int advance, hipStream_t st);
int scan_chunks(int nkeys);
// k_perm + k_cellsort (grid: workgroups of the per-block pass; src_is_b: the
// sorted slots index the B side's rows and ids, else the A side's)
void launch_perm(const DevPtrs& d, int64_t n, int grid, int src_is_b, hipStream_t st);
// phase 0: fused USL G2P (+ dead-particle copy); MUSL: phase 1 (first half),
// launch_musl (second scatter + grid velocity), phase 2 (second half + copy).
// learned (phases 0, 2): queue the closure's candidates for launch_closure.
void launch_g2p(const DevPtrs& d, int grid, int learned, int phase, hipStream_t st);
void launch_musl(const DevPtrs& d, int grid_p2g, int grid_grid, hipStream_t st);
// The learned closure over the candidates G2P queued (after each G2P of a step
// with the learned return map).
void launch_closure(const DevPtrs& d, const void* mlp_blob, int grid, hipStream_t st);
void launch_closure_eval(const SceneDev* scenes, const void* mlp_blob, const float* trial4, int64_t n,
                         float* learned4, float* exact4, uint8_t* accepted, int max_grid, hipStream_t st);
void launch_p2g(const DevPtrs& d, int grid, int src_is_b, hipStream_t st);
int p2g_resident_per_cu();
int g2p_resident_per_cu();
void launch_grid(const DevPtrs& d, int grid, hipStream_t st);
// the last step's grid update again from its tiles with the current scene
// constants (after a dt change; not after an atomic scatter)
void launch_grid_redo(const DevPtrs& d, int grid, hipStream_t st);
// largest squared speed of the live particles of the current side -> *out (float bits; clear it first)
void launch_max_speed(const DevPtrs& d, int64_t n, unsigned* out, hipStream_t st);
void launch_unsort(const DevPtrs& d, int64_t n, float* out, int first, int ncomp, hipStream_t st);
void launch_active_out(const DevPtrs& d, int64_t n, uint8_t* out, hipStream_t st);
void launch_pack_render(const DevPtrs& d, int64_t n, float* pos4, float* lat7, float* diff, hipStream_t st);
void launch_boundary_eval(const SceneDev* sc, const KnifeStateDev* ks, const float* X, const float* v, float* out,
                          int64_t n, uint32_t* err, hipStream_t st);
void launch_grid_dense(const DevPtrs& d, int s, float* mass, float* v, float* flip, float* pig, float* shear,
                       hipStream_t st);

}  // namespace mpmx
