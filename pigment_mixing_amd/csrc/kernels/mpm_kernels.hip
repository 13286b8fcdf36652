// mpm_kernels.hip — hand-written gfx950 kernels of the MPM step.
//
// Step pipeline (captured once as a HIP graph; all scenes of a batch at once):
//   k_g2p     lane = particle, one workgroup per particle block (4x4x4 stencil
//             bins): G2P gather of the updated grid (g2p.cpp:12-135), F update +
//             VMVisc return map (deformation_update.cpp:7-52, plasticity.cpp:163-238),
//             pigment mixing, advection and the particle guards
//             (position_update.cpp:6-91).  Writes the new state (with the Kirchhoff
//             stress of the new F for the next scatter) in sorted order, and counts
//             the particles' new blocks and cells for the next sort.
//   k_scan_*  prefix sums over block counts -> block offsets + block list.
//   k_perm    particles to their block's range by their rank in the block;
//   k_cellsort per block: cell counts and offsets, each cell's particles ordered
//             by position (canonical), and the permutation slot <- particle:
//             block-major, rank-major in a block.
//   k_p2g     lane = cell, wave = (block, channel group): each lane walks its
//             cell's particles and accumulates their 27-node contributions (mass,
//             APIC momentum, Kirchhoff force, pigment latent, shear intensity;
//             p2g.cpp:6-126, explicit_euler_update.cpp:12-55) in registers; the
//             lanes' stencils merge across lanes (DPP / shuffles, one axis at a
//             time) into the block's 6x6x6 node tile with one owner per node,
//             stored for the grid kernel (grouped by destination node block).
//   k_grid    one wavefront per active 4x4x4 node block: normalise, gravity + force,
//             boundary collision with plates and knife (explicit_euler_update.cpp:64-110,
//             boundary_collision.cpp:5-316), write the gather grid, clear accumulators.
#include "../common/mpmx_device.hpp"
#include "mpm_kernels.hpp"
#include "mlp_device.hpp"

#include <algorithm>

namespace mpmx {

// ------------------------------------------------------------------ helpers
// Direct-to-LDS loads (global_load_lds): lane l's dword lands at word l of
// the 64-word LDS line given (a wave-uniform address), without registers.
typedef __attribute__((address_space(1))) const void* GlobalPtr;
typedef __attribute__((address_space(3))) void* LdsPtr;

__device__ __forceinline__ void dma_dword(const void* row_base, int i, float* lds_line) {
  const char* p = (const char*)row_base + (uint64_t)((uint32_t)i * 4u);
  __builtin_amdgcn_global_load_lds((GlobalPtr)p, (LdsPtr)lds_line, 4, 0, 0);
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maximum): every vector
// memory operation of this wavefront has completed (direct-to-LDS loads have
// landed, atomics have been performed).
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (__lane_id() == 0) ? 0ull : ((~0ull) >> (64 - __lane_id()));
}

// Rank of this lane's key among equal keys, one atomic per distinct key in the
// wavefront (all issued by one instruction).  Only lanes with `valid`
// participate; all lanes must call it.
__device__ __forceinline__ int wave_rank(bool valid, int key, int* counts) {
  // group the lanes by key (ballots only), then one vector atomic in which
  // every group's first lane adds the group size, so the atomics' round
  // trips overlap instead of following one another
  uint64_t todo = __ballot(valid);
  int leader_of = 0, below = 0, size = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int k0 = __builtin_amdgcn_readlane(key, leader);
    const uint64_t same = __ballot(valid && key == k0);
    if (valid && key == k0) {
      leader_of = leader;
      below = (int)__popcll(same & lanemask_lt());
    }
    if ((int)__lane_id() == leader) size = (int)__popcll(same);
    todo &= ~same;
  }
  int base = 0;
  if (size > 0) base = atomicAdd(&counts[key], size);
  base = __shfl(base, leader_of);
  return valid ? base + below : 0;
}

__device__ __forceinline__ void activate_node_block(const DevPtrs& d, int nb, int parity) {
  if (atomicCAS(&d.nb_flag[nb], 0, 1) == 0) {
    const int idx = atomicAdd(&d.counters[CNT_NBLOCKS0 + parity], 1);
    atomicAdd(&d.stats[ST_NBLOCKS], 1ull);
    d.nb_list[parity * d.nb_list_cap + idx] = nb;
  }
}

// Particle key = node block holding the stencil base bin; cell = bin inside the
// block (0..63).  Dead if the 3x3x3 stencil leaves the lattice.
__device__ __forceinline__ int particle_key(const SceneDev& sc, const float x[3], int dead_key, int& cell) {
  const int b0 = stencil_base(x[0], sc.origin[0], sc.one_over_dx);
  const int b1 = stencil_base(x[1], sc.origin[1], sc.one_over_dx);
  const int b2 = stencil_base(x[2], sc.origin[2], sc.one_over_dx);
  cell = -1;
  if (b0 < 0 || b1 < 0 || b2 < 0 || b0 + 2 > sc.dims[0] - 1 || b1 + 2 > sc.dims[1] - 1 ||
      b2 + 2 > sc.dims[2] - 1)
    return dead_key;
  cell = node_local(b0, b1, b2);
  return node_block_of(sc, b0, b1, b2);
}

// Order key of a particle inside its cell (k_cellsort): its position in the
// cell, 10 bits per axis from the same scaled coordinate that bins it
// (stencil_base), x in the high bits.  Monotone in each coordinate, so ranking
// by the key and, where keys are equal, by the full position gives a total
// order that depends on the positions (and ids) alone.
__device__ __forceinline__ uint32_t cell_order_key(const SceneDev& sc, const float x[3]) {
  uint32_t k = 0u;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = __fsub_rn(__fmul_rn(__fsub_rn(x[a], sc.origin[a]), sc.one_over_dx), 0.5f);
    const int q = (int)((f - floorf(f)) * 1024.0f);
    k = (k << 10) | (uint32_t)min(max(q, 0), 1023);
  }
  return k;
}

__device__ __forceinline__ void block_coords(const SceneDev& sc, int key, int& bi, int& bj, int& bk) {
  const int lb = key - sc.node_block_base;
  bi = lb / (sc.bdims[1] * sc.bdims[2]);
  bj = (lb / sc.bdims[2]) % sc.bdims[1];
  bk = lb % sc.bdims[2];
}

// Persistent workgroups walk the block list XCD by XCD: workgroups are handed
// to the 8 XCDs round-robin (blockIdx.x % 8), so XCD x takes the x-th
// contiguous eighth of the (spatially ordered) block list and neighbouring
// blocks share halo nodes in that XCD's L2.
struct XcdRange {
  int e, end, stride;
  int first;   // first block of this XCD's share
};
__device__ __forceinline__ XcdRange xcd_range(int nblk) {
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int per = gridDim.x / nx;
  const int seg = (nblk + nx - 1) / nx;
  XcdRange r;
  r.first = xcd * seg;
  r.e = xcd * seg + (int)(blockIdx.x / nx);
  r.end = min(nblk, (xcd + 1) * seg);
  r.stride = per;
  return r;
}

// Count a particle for the next sort: its block (count and rank inside the
// block, one atomic per distinct key in the wavefront), cell and order key.
// All lanes call.  The order inside a block is fixed later by k_cellsort,
// from the cells and order keys alone, so the ranks' atomic timing does not
// reach any result.
__device__ __forceinline__ void count_for_sort(const DevPtrs& d, bool valid, int key, int cell, int64_t j,
                                               uint32_t okey = 0u) {
  const bool live = valid && cell >= 0;
  const int rb = wave_rank(valid, key, d.count_next);
  if (valid) {
    d.key_out[j] = key;
    d.cell_out[j] = live ? cell : -1;
    d.rank_out[j] = rb;
    if (live) d.okey_out[j] = okey;
  }
}

// ------------------------------------------------------------------ k_init_tau
// Kirchhoff stress of the uploaded F (HenckyPiola + tau = P F^T,
// simulation.hpp:287-294, explicit_euler_update.cpp:19-20).
__global__ void k_init_tau(DevPtrs d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SceneDev& sc = d.scenes[d.pscene[i]];
  float F[9], U[9], lam[3], V[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) F[c] = d.pa[(P_F + c) * d.cap + i];
  principal3(F, U, lam, V);
  float h[3] = {log_sigma_of(lam[0]), log_sigma_of(lam[1]), log_sigma_of(lam[2])};
  float tau[6];
  kirchhoff_from_principal(sc, U, h, tau);
#pragma unroll
  for (int c = 0; c < 6; ++c) d.pa[(P_TAU + c) * d.cap + i] = tau[c];
}

// ------------------------------------------------------------------ k_keys
// Bin every particle (upload order) for the first sort.
__global__ void k_keys(DevPtrs d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int key = d.dead_key, cell = -1;
  uint32_t okey = 0u;
  if (valid) {
    const SceneDev& sc = d.scenes[d.pscene[i]];
    const float x[3] = {d.pa[0 * d.cap + i], d.pa[1 * d.cap + i], d.pa[2 * d.cap + i]};
    if (d.pa_id[i] >= 0 && d.active_in[i]) key = particle_key(sc, x, d.dead_key, cell);
    okey = cell_order_key(sc, x);
    if (key == d.dead_key) {
      // inactive from the start: v = 0 (position_update.cpp:222-228 convention)
      d.pa[3 * d.cap + i] = 0.0f;
      d.pa[4 * d.cap + i] = 0.0f;
      d.pa[5 * d.cap + i] = 0.0f;
    }
  }
  count_for_sort(d, valid, key, cell, i, okey);
}

// ------------------------------------------------------------------ scans
// counts_next[0..nkeys] (last = DEAD) -> offsets (exclusive), count_cur, the
// non-empty live block list (ascending key order) and totals.  Two phases over
// chunks of SCAN_CHUNK keys; the scanned 64-bit value packs (count<<32 | nonempty).
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 2;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ unsigned long long scan_item(const DevPtrs& d, int k, int nkeys) {
  if (k >= nkeys) return 0ull;
  const unsigned long long c = (unsigned long long)d.count_next[k];
  const bool live = k < nkeys - 1;
  return (c << 32) | (unsigned long long)((c > 0 && live) ? 1 : 0);
}

// Hillis-Steele exclusive scan over the workgroup (deterministic).
__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long* sh,
                                                   unsigned long long& total) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const unsigned long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0ull;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  total = sh[SCAN_THREADS - 1];
  const unsigned long long incl = sh[threadIdx.x];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(DevPtrs d, int nkeys) {
  __shared__ unsigned long long sh[SCAN_THREADS];
  const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
  unsigned long long s = 0;
  for (int j = 0; j < SCAN_PER_THREAD; ++j) s += scan_item(d, base + j, nkeys);
  unsigned long long tot;
  block_exclusive_scan(s, sh, tot);
  if (threadIdx.x == 0) d.scan_partials[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(DevPtrs d, int nkeys, int nchunks, int advance) {
  __shared__ unsigned long long sh[SCAN_THREADS];
  unsigned long long pre = 0;
  for (int c = threadIdx.x; c < blockIdx.x; c += SCAN_THREADS) pre += d.scan_partials[c];
  unsigned long long chunk_base;
  block_exclusive_scan(pre, sh, chunk_base);
  const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
  unsigned long long items[SCAN_PER_THREAD];
  unsigned long long s = 0;
  for (int j = 0; j < SCAN_PER_THREAD; ++j) {
    items[j] = scan_item(d, base + j, nkeys);
    s += items[j];
  }
  unsigned long long ctot;
  unsigned long long run = chunk_base + block_exclusive_scan(s, sh, ctot);
  for (int j = 0; j < SCAN_PER_THREAD; ++j) {
    const int k = base + j;
    if (k < nkeys) {
      d.offsets[k] = (int)(run >> 32);
      d.count_cur[k] = (int)(items[j] >> 32);
      // list position of every key (INT_MAX for an unlisted one), so that a
      // lookup needs no count check first
      d.list_pos[k] = (items[j] & 1ull) ? (int)(run & 0xffffffffull) : 0x7fffffff;
      if (items[j] & 1ull) {
        const int e = (int)(run & 0xffffffffull);
        d.pb_list[e] = k;
        d.pdesc[e] = make_int4(k, (int)(run >> 32), (int)(items[j] >> 32), d.blk_scene[k]);
      }
      d.count_next[k] = 0;
    }
    run += items[j];
  }
  __syncthreads();
  if (blockIdx.x == nchunks - 1 && threadIdx.x == SCAN_THREADS - 1) {
    d.counters[CNT_PBLOCKS] = (int)(run & 0xffffffffull);
    d.counters[CNT_LIVE] = d.offsets[nkeys - 1];
    // the closure may still be reading this step's queue counts (it runs
    // beside the sort); the previous step's are consumed: clear them for the
    // step after this one, which appends to that buffer again
    if (d.cq_cnt_prev) {
      for (int x = 0; x < CQ_SEGS; ++x) d.cq_cnt_prev[x * CQ_CNT_STRIDE] = 0;
    }
    if (advance) {
      d.counters[CNT_STEP] += 1;
      d.counters[CNT_GSTEP] += 1;
    }
  }
}

// ------------------------------------------------------------------ k_perm
// Every particle to its block's range by its rank inside the block: live ones
// into `cmaj` (arrival order, which k_cellsort replaces by the block's
// rank-major layout), dead ones straight to their slot of the DEAD key's range.
__global__ void k_perm(DevPtrs d, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int pos = d.offsets[d.key_out[j]] + d.rank_out[j];
  if (d.cell_out[j] >= 0)
    d.cmaj[pos] = (int)j;
  else
    d.perm[pos] = (int)j;
}

// ------------------------------------------------------------------ k_cellsort
// Per listed block (one 256-thread workgroup; grid-stride over the list): the
// block's particles in their final, rank-major order.  Inside a block the
// particles are laid out rank-major: first rank 0 of every non-empty cell
// (cells in lane order), then rank 1 of every cell holding two or more, and so
// on.  The scatter walks a block with lane = cell and step k = rank, so at
// every step its active lanes read consecutive slots (coalesced), where a
// cell-major layout would spread one load over ~8 particles' stride.
//   1. the cells and order keys of the block's particles (cmaj, arrival
//      order) into LDS; cell counts by LDS atomics;
//   2. cell offsets (wave scan), the counts for the scatter (cell_num), and
//      per rank r the mask of cells holding more than r particles with the
//      slots of the ranks below: slot(c, r) = sum_{r' < r} #{cells with count
//      > r'} + #{cells below c with count > r};
//   3. the particles bucketed by cell;
//   4. each particle's rank inside its cell: ascending position, i.e. the
//      order key (cell_order_key, written by the key pass), on equal keys x,
//      y, z and then the particle id (pos_before); perm[slot] = particle.
// Ranking by position makes the order inside a cell, and with it every FP32
// sum of the scatter (mass, momentum, force, latents, shear), a function of
// the particles' state alone: independent of the upload order, of the layout
// the previous step left, of the atomics' timing, of the other scenes of a
// batch and of the launch shapes.  Blocks of more than CS_CAP particles keep
// the per-particle arrays in global memory (cmaj, the key pass's arrays, and
// `kslot` as the bucket array); ranks of CS_MAXR and above take their slot
// from a loop over the 64 counts.
constexpr int CS_CAP = 2048;
constexpr int CS_MAXR = 64;
// float -> unsigned with the same order (negative values reversed below the
// positive ones); -0 sorts just below +0
__device__ __forceinline__ uint32_t ord_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// slot a before slot b among particles with equal order keys: x, y, z, id
__device__ __forceinline__ bool pos_before(const float* __restrict__ X, const int* __restrict__ ids, int64_t cap,
                                           int a, int b) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const uint32_t ka = ord_key(X[(P_X + c) * cap + a]), kb = ord_key(X[(P_X + c) * cap + b]);
    if (ka != kb) return ka < kb;
  }
  return a != b && ids[a] < ids[b];
}
__global__ void __launch_bounds__(256) k_cellsort(DevPtrs d, const float* __restrict__ X, const int* __restrict__ ids) {
  __shared__ uint32_t s_key[CS_CAP];
  __shared__ unsigned short s_bucket[CS_CAP];
  __shared__ unsigned char s_cell[CS_CAP];
  __shared__ int s_cnt[BLK3], s_fill[BLK3], s_off[BLK3];
  __shared__ unsigned long long s_mask[CS_MAXR];
  __shared__ int s_base[CS_MAXR];
  const int nblk = d.counters[CNT_PBLOCKS];
  const int64_t cap = d.cap;
  const int lane = threadIdx.x & 63;
  for (int e = blockIdx.x; e < nblk; e += gridDim.x) {
    const int4 ds = d.pdesc[e];
    const int start = ds.y, cnt = ds.z;
    const bool staged = cnt <= CS_CAP;
    int* const gbucket = d.kslot + start;   // bucket array of an unstaged block
    __syncthreads();   // the previous block is done with the LDS
    if (threadIdx.x < BLK3) { s_cnt[threadIdx.x] = 0; s_fill[threadIdx.x] = 0; }
    __syncthreads();
    // 1. particles, cells, keys; cell counts
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int j = d.cmaj[start + t];
      const int c = d.cell_out[j];
      if (staged) {
        s_key[t] = (uint32_t)d.okey_out[j];
        s_cell[t] = (unsigned char)c;
      }
      atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
    // 2. offsets, counts for the scatter, rank masks (first wavefront)
    if (threadIdx.x < 64) {
      const int n = s_cnt[lane];
      int incl = n;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
      }
      s_off[lane] = incl - n;
      d.cell_num[e * BLK3 + lane] = n;
      int maxc = n;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) maxc = max(maxc, __shfl_xor(maxc, off));
      int run = 0;
      for (int r = 0; r < min(maxc, CS_MAXR); ++r) {
        const uint64_t m = __ballot(n > r);
        if (lane == 0) { s_mask[r] = m; s_base[r] = run; }
        run += (int)__popcll(m);
      }
    }
    __syncthreads();
    // 3. bucket by cell (order inside a bucket arbitrary)
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int c = staged ? (int)s_cell[t] : d.cell_out[d.cmaj[start + t]];
      const int q = s_off[c] + atomicAdd(&s_fill[c], 1);
      if (staged) s_bucket[q] = (unsigned short)t; else gbucket[q] = t;
    }
    __syncthreads();
    // 4. rank inside the cell, slot, permutation
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int j = d.cmaj[start + t];
      const int c = staged ? (int)s_cell[t] : d.cell_out[j];
      const uint32_t kr = staged ? s_key[t] : (uint32_t)d.okey_out[j];
      const int q0 = s_off[c], q1 = q0 + s_cnt[c];
      int rk = 0;
      bool tie = false;
      for (int q = q0; q < q1; ++q) {
        const int u = staged ? s_bucket[q] : gbucket[q];
        const uint32_t ku = staged ? s_key[u] : (uint32_t)d.okey_out[d.cmaj[start + u]];
        rk += ku < kr ? 1 : 0;
        tie |= u != t && ku == kr;
      }
      if (tie) {   // equal order keys (particles within 1/1024 of a cell of each other): full positions decide
        for (int q = q0; q < q1; ++q) {
          const int u = staged ? s_bucket[q] : gbucket[q];
          const int ju = d.cmaj[start + u];
          const uint32_t ku = staged ? s_key[u] : (uint32_t)d.okey_out[ju];
          if (u != t && ku == kr) rk += pos_before(X, ids, cap, ju, j) ? 1 : 0;
        }
      }
      int slot;
      if (rk < CS_MAXR) {
        slot = s_base[rk] + (int)__popcll(s_mask[rk] & ((1ull << c) - 1ull));
      } else {
        slot = 0;
        for (int c2 = 0; c2 < BLK3; ++c2) slot += min(s_cnt[c2], rk) + (c2 < c && s_cnt[c2] > rk ? 1 : 0);
      }
      d.perm[start + slot] = j;
    }
  }
}

// Dead particles keep their state: copy them to the output side and rank
// them.  Run by G2P's workgroups after their blocks (the dead slots are
// disjoint from the live ones), a share per workgroup: no launch of its own.
__device__ __forceinline__ void copy_dead_share(const DevPtrs& d) {
  const int dead_start = d.offsets[d.dead_key];
  const int ndead = d.count_cur[d.dead_key];
  for (int t0 = blockIdx.x * blockDim.x; t0 < ndead; t0 += gridDim.x * blockDim.x) {
    const int t = t0 + threadIdx.x;
    const bool valid = t < ndead;
    const int64_t j = dead_start + (valid ? t : 0);
    if (valid) {
      const int src = d.perm[j];
      for (int c = 0; c < P_NCOMP; ++c) d.pb[c * d.cap + j] = d.pa[c * d.cap + src];
      d.pb_id[j] = d.pa_id[src];
    }
    count_for_sort(d, valid, d.dead_key, -1, j);
  }
}

// ------------------------------------------------------------------ k_g2p
constexpr int G2P_STATIC = 4;   // blocks per workgroup assigned statically before the queue
constexpr int GT = 6;              // gather tile edge: stencils of bins [B0, B0+4) span nodes [B0, B0+6)
constexpr int GT3 = GT * GT * GT;  // 216 nodes
// LDS row of one node: 16 floats padded to 5 float4 (80 B), so that the nodes
// the cells of one wavefront read at once (node index differences 1, 2, 3, 6,
// 7, 8, 9, 36, ...) start in distinct 4-bank groups.
constexpr int GT_ROW = 5;

// Read-only inputs come in as __restrict__ kernel arguments (scene table, node
// coordinates, particle input side, permutation, grid values), so uniform reads
// of them become scalar loads that the particle stores cannot clobber.
//
// The learned return-map closure runs in its own kernel (k_closure) over a
// queue of the plastic candidates that G2P found: the network (BF16 weights,
// mlp_device.hpp) sits in LDS, each wavefront owns 64 16-byte staging rows
// and evaluates 64 queued candidates at once on MFMA (activations in
// registers), so the gather kernel keeps
// the exact variant's registers, LDS and occupancy.  A candidate's predicted
// state is accepted when its Perzyna consistency residual, its trace change
// (x K) and its flow-rule mismatch are all within closure_tol * q_y; otherwise
// the lane runs the exact map (fallback, counted).
struct G2PLearnedLds {
  uint8_t net[MLP_BLOB_BYTES];
  uint4 stage[4][64];
};

// Learned return map for the principal Hencky strains h (in/out) of this lane.
// Called by all 64 lanes of a wavefront together (MFMA).  `active`: the lane
// holds a particle.  cand_out: its trial state is plastic (q_tr > q_y, not the
// pressure cap).  Returns true when the predicted state was accepted and
// applied to h, eps, dg_rate; false leaves them untouched (the caller runs the
// exact map).
// Trial quantities of the return map (plasticity.cpp:163-170) from the principal
// Hencky strains: trace, unit deviator direction, ||dev h||, q_trial, q_yield.
struct TrialState {
  float htr, nv[3], dn, q_trial, q_yield;
  bool cand;   // plastic candidate of the closure: q_tr > q_y outside the pressure cap
};
__device__ __forceinline__ TrialState trial_state(const SceneDev& sc, const float h[3], float eps) {
  TrialState t;
  t.htr = (h[0] + h[1]) + h[2];
  const float m = fdiv(t.htr, 3.0f);
  t.nv[0] = h[0] - m; t.nv[1] = h[1] - m; t.nv[2] = h[2] - m;
  t.dn = sqrtf(t.nv[0] * t.nv[0] + t.nv[1] * t.nv[1] + t.nv[2] * t.nv[2]);
  if (t.dn > 0.0f) { const float inv = rcp_approx(t.dn); t.nv[0] *= inv; t.nv[1] *= inv; t.nv[2] *= inv; }
  t.q_trial = sc.e_mu_prefac * t.dn;
  t.q_yield = sc.q_min + (sc.q_max - sc.q_min) * expf(-sc.xi * eps);
  const float p_trial = -sc.K * t.htr;
  t.cand = sc.closure != 0 && sc.plastic_model == 5 && !(p_trial < sc.p_min) && t.q_trial > t.q_yield;
  return t;
}

__device__ bool learned_return_map(const SceneDev& sc, G2PLearnedLds& lrn, float h[3], float& eps, float& dg_rate,
                                   bool active, bool& cand_out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const TrialState ts = trial_state(sc, h, eps);
  const float htr = ts.htr, dn = ts.dn, q_trial = ts.q_trial, q_yield = ts.q_yield;
  const float nv[3] = {ts.nv[0], ts.nv[1], ts.nv[2]};
  const bool cand = active && ts.cand;
  cand_out = cand;
  if (!__ballot(cand)) return false;
  // ---- network inputs of all 64 lanes -> this wavefront's staging rows
  const MlpNet net = mlp_net(lrn.net);
  uint4* stage = lrn.stage[wave];
  const float x8[8] = {h[0], h[1], h[2], dn, q_yield, q_yield * sc.perzyna_visc, sc.dt * sc.E, sc.nu};
  float xn[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) xn[c] = (x8[c] - net.norm[c]) / net.norm[8 + c];
  mlp_stage_inputs(stage, xn);
  float y[4];
  mlp_wave_infer(net, stage, y);
#pragma unroll
  for (int c = 0; c < 4; ++c) y[c] = y[c] * net.norm[20 + c] + net.norm[16 + c];
  if (!cand) return false;
  // ---- predicted state
  float dg = fmaxf(y[3], 0.0f);
  float hp[3];
  if (sc.closure == 2) {
    // one Newton step on the consistency residual from the predicted increment,
    // strains by the flow rule
    const float r = perzyna_residual(sc, q_trial - sc.f_mu_prefac * dg, dg, eps);
    const float rd = perzyna_residual_slope(sc, q_trial, dg, eps);
    if (rd != 0.0f) dg = fmaxf(dg - fdiv(r, rd), 0.0f);
    const float fr = sc.inv_d_prefac * dg;
    hp[0] = h[0] - fr * nv[0]; hp[1] = h[1] - fr * nv[1]; hp[2] = h[2] - fr * nv[2];
  } else {
    hp[0] = h[0] + y[0]; hp[1] = h[1] + y[1]; hp[2] = h[2] + y[2];
  }
  // ---- acceptance, each residual within tol * q_y:
  //   yield consistency of (q(h'), dgamma); trace change x K; and the flow rule
  //   h' = h - (dgamma / d) n (direction and magnitude) x e_mu
  const float trp = (hp[0] + hp[1]) + hp[2];
  const float mp = trp / 3.0f;
  const float e0 = hp[0] - mp, e1 = hp[1] - mp, e2 = hp[2] - mp;
  const float qp = sc.e_mu_prefac * sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
  const float lim = sc.closure_tol * q_yield;
  const float ry = fabsf(perzyna_residual(sc, qp, dg, eps));
  const float rv = sc.K * fabsf(trp - htr);
  const float fac = sc.inv_d_prefac * dg;
  const float f0 = hp[0] - (h[0] - fac * nv[0]), f1 = hp[1] - (h[1] - fac * nv[1]), f2 = hp[2] - (h[2] - fac * nv[2]);
  const float rf = sc.e_mu_prefac * sqrtf(f0 * f0 + f1 * f1 + f2 * f2);
  if (!(ry <= lim && rv <= lim && rf <= lim)) return false;
  h[0] = hp[0]; h[1] = hp[1]; h[2] = hp[2];
  eps = eps + sc.inv_d_prefac * dg;
  dg_rate = fdiv(dg, sc.dt);
  return true;
}

// Harvest record of one plastic particle: h_trial[3], ||dev h_trial||, q_y,
// eta = q_y * perzyna_visc, dt * E, nu | corrected h[3], dgamma.
__device__ __forceinline__ void harvest_pair(const DevPtrs& d, const SceneDev& sc, const float h_trial[3],
                                             float eps_in, const float h[3], float dgr) {
  const unsigned long long slot = atomicAdd(&d.stats[ST_HARVEST], 1ull);
  if ((int64_t)slot >= d.harvest_cap) return;
  float* r = d.harvest + 12 * slot;
  const float m = ((h_trial[0] + h_trial[1]) + h_trial[2]) / 3.0f;
  const float e0 = h_trial[0] - m, e1 = h_trial[1] - m, e2 = h_trial[2] - m;
  const float qy = sc.q_min + (sc.q_max - sc.q_min) * expf(-sc.xi * eps_in);
  r[0] = h_trial[0]; r[1] = h_trial[1]; r[2] = h_trial[2];
  r[3] = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
  r[4] = qy;
  r[5] = qy * sc.perzyna_visc;
  r[6] = sc.dt * sc.E;
  r[7] = sc.nu;
  r[8] = h[0]; r[9] = h[1]; r[10] = h[2];
  r[11] = dgr * sc.dt;
}

// PHASE 0: the fused USL step.  MUSL (musl.cpp:6-98) splits it around a second
// scatter: PHASE 1 gathers from the updated grid (velocity, APIC B, FLIP,
// latent, shear), mixes the latent and sets v (AFLIP), and parks the PIC
// velocity in the output side's position rows; the MUSL scatter and grid
// normalisation then rebuild the grid velocity from the new v and B; PHASE 2
// gathers the velocity gradient from it and finishes the step (F update,
// return map, stress, advection with the parked PIC velocity, sort keys).
// LEARNED (phases 0 and 2): plastic candidates of the closure are appended to
// the closure queue with their trial F instead of being mapped here; k_closure
// writes their F, eps, dgamma and stress.
// Four wavefronts per SIMD (128 VGPRs; the fused phase spills a few dwords):
// the gather's latency hiding is worth more than the spills (measured ~20%
// faster than three wavefronts at 137 VGPRs).  Two-wavefront workgroups: a
// block's particles run in passes of 128, so that fewer wavefronts wait at
// the next block's barrier while the last pass finishes (blocks hold a few
// hundred particles; with four-wavefront workgroups G2P took 5 % longer).
template <bool LEARNED, int PHASE, bool VOL>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) k_g2p(DevPtrs d, const SceneDev* __restrict__ scenes,
                                             const float* __restrict__ axis, const float* __restrict__ A,
                                             const int* __restrict__ perm, const float* __restrict__ gv,
                                             const int4* __restrict__ pdesc) {
  __shared__ float4 gtile[GT3 * GT_ROW];
  __shared__ float sax[3][GT];       // node coordinates B0 .. B0+5 per axis
  __shared__ unsigned int plastic_wg, newton_wg;
  __shared__ int next_blk;
  const int step = d.counters[CNT_STEP];
  const int nblk = d.counters[CNT_PBLOCKS];
  float* __restrict__ Bf = d.pb;
  const int64_t cap = d.cap;
  if (threadIdx.x == 0) { plastic_wg = 0u; newton_wg = 0u; }
  const XcdRange xr = xcd_range(nblk);
  // this XCD's closure-queue segment (learned variant)
  int64_t cq_base = 0;
  int* cq_counter = nullptr;
  if constexpr (LEARNED) {
    cq_base = xr.first < nblk ? pdesc[xr.first].y : 0;
    cq_counter = d.cq_cnt + CQ_CNT_STRIDE * (blockIdx.x % CQ_SEGS);
    // the segment's start for the closure, which runs beside the next sort
    // and so cannot read the block list itself (the sort rewrites it)
    if (blockIdx.x < CQ_SEGS && threadIdx.x == 0) d.cq_start[blockIdx.x] = (int)cq_base;
  }
  // Blocks of this XCD's share: G2P_STATIC static ones per workgroup (walker
  // index w, w + W, ... for W walkers), then claims from the XCD's queue
  // (claim c is block G2P_STATIC * W + c of the share): block costs follow
  // their particle counts.  When a block starts, thread 0 settles the next
  // block and claims the one after it; the next block reaches the workgroup
  // through LDS (ordered by the staging barrier).  The last workgroup of the
  // XCD to finish resets the queue for the next launch.
  const int walkers = xr.stride;
  const int nshare = max(0, xr.end - xr.first);
  int* const q_next = d.wq + (2 * CQ_SEGS + (int)(blockIdx.x % CQ_SEGS)) * CQ_CNT_STRIDE;
  int* const q_done = d.wq + (3 * CQ_SEGS + (int)(blockIdx.x % CQ_SEGS)) * CQ_CNT_STRIDE;
  int claim = -1;     // thread 0: the claim made when the previous block ended
  int nstarted = 1;   // thread 0: blocks started so far
  for (int k = xr.e - xr.first; k < nshare;) {
    const int e = xr.first + k;
    const int4 ds = pdesc[e];       // key, first slot, count, scene
    const int key = ds.x;
    const int s = ds.w;
    // the first pass's source slots, loaded while the gather tile is staged
    // (each pass then loads the next pass's: the particle rows of a pass wait
    // on one round trip, not on two dependent ones)
    int src_next = perm[ds.y + min((int)threadIdx.x, max(ds.z - 1, 0))];
    const SceneDev& sc = scenes[s];
    int bi, bj, bk;
    block_coords(sc, key, bi, bj, bk);
    const int B0i = bi * 4, B0j = bj * 4, B0k = bk * 4;
    __syncthreads();  // previous block done with the tile (and with next_blk)
    // the next block: a static one, or the claim made when the previous block
    // started; every thread reads it after this block's passes
    if (threadIdx.x == 0) {
      next_blk = nstarted < G2P_STATIC ? k + walkers : G2P_STATIC * walkers + claim;
      ++nstarted;
      if (nstarted >= G2P_STATIC) claim = atomicAdd(q_next, 1);   // the block after next
    }
    // stage the gather grid of nodes [B0, B0+6)^3 (coalesced 16-byte rows)
    for (int q = threadIdx.x; q < GT3 * 4; q += blockDim.x) {
      const int t = q >> 2, part = q & 3;
      const int gi = min(B0i + t / (GT * GT), sc.dims[0] - 1);
      const int gj = min(B0j + (t / GT) % GT, sc.dims[1] - 1);
      const int gk = min(B0k + t % GT, sc.dims[2] - 1);
      const float4* g = reinterpret_cast<const float4*>(
          gv + ((size_t)node_block_of(sc, gi, gj, gk) * BLK3 + node_local(gi, gj, gk)) * GV_STRIDE);
      gtile[t * GT_ROW + part] = g[part];
    }
    if (threadIdx.x < 3 * GT) {
      const int a = threadIdx.x / GT, o = threadIdx.x % GT;
      const int B0a = a == 0 ? B0i : (a == 1 ? B0j : B0k);
      const int first = a == 0 ? 0 : (a == 1 ? sc.dims[0] : sc.dims[0] + sc.dims[1]);
      sax[a][o] = axis[sc.axis_off + first + min(B0a + o, sc.dims[a] - 1)];
    }
    __syncthreads();
    const float time = d.kstate[(size_t)step * d.n_scenes + s].time;
    // pigment boost (g2p.cpp:107-119)
    float tf = 1.0f;
    if (time > sc.start_boost && sc.end_boost > sc.start_boost) {
      float t = (time - sc.start_boost) / (sc.end_boost - sc.start_boost);
      t = fminf(fmaxf(t, 0.0f), 1.0f);
      tf = 1.0f + t * (sc.boost_factor - 1.0f);
    }
    const float dynD = sc.D_max * tf;
    const int start = ds.y, cnt = ds.z;
    const int iters = (cnt + blockDim.x - 1) / blockDim.x;
    for (int it = 0; it < iters; ++it) {
      // a wavefront whose lanes all lie past the block's count has nothing to
      // do in this pass: it skips to the next block's barrier and leaves the
      // SIMD to the other wavefronts
      if (it * (int)blockDim.x + (int)(threadIdx.x & ~63u) >= cnt) continue;
      const int t = it * blockDim.x + threadIdx.x;
      const bool valid = t < cnt;
      // Every lane computes; a lane past the block's count recomputes the block's
      // last particle and stores nothing.
      const int64_t j = start + (valid ? t : cnt - 1);
      const int src = src_next;
      if (it + 1 < iters) src_next = perm[start + min(t + (int)blockDim.x, cnt - 1)];
      uint32_t okey = 0u;
      int newkey = d.dead_key, newcell = -1, plastic = 0;
      int newton = 0;
      {
        float x[3], v[3], F[9], pig[NPIG], picB[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = A[(P_X + c) * cap + src];
        if constexpr (PHASE == 2) {
          // v and the PIC velocity come from the first half (output side, slot j)
#pragma unroll
          for (int c = 0; c < 3; ++c) { v[c] = Bf[(P_V + c) * cap + j]; picB[c] = Bf[(P_X + c) * cap + j]; }
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) { v[c] = A[(P_V + c) * cap + src]; picB[c] = 0.0f; }
        }
        if constexpr (PHASE != 1) {
#pragma unroll
          for (int c = 0; c < 9; ++c) F[c] = A[(P_F + c) * cap + src];
        }
        if constexpr (PHASE != 2) {
#pragma unroll
          for (int c = 0; c < NPIG; ++c) pig[c] = A[(P_PIG + c) * cap + src];
        }
        const float eps_in = PHASE == 1 ? 0.0f : A[P_EPS * cap + src];
        const int id = d.pa_id[src];
        // ---------------- G2P gather (g2p.cpp:25-91) over the 27-node stencil
        int b[3];
        AxisW w[3];
        b[0] = stencil_base(x[0], sc.origin[0], sc.one_over_dx);
        b[1] = stencil_base(x[1], sc.origin[1], sc.one_over_dx);
        b[2] = stencil_base(x[2], sc.origin[2], sc.one_over_dx);
        // stencil bases relative to the staged coordinates (0 .. 3)
        const int r0 = b[0] - B0i, r1 = b[1] - B0j, r2 = b[2] - B0k;
        axis_weights(x[0], sax[0], r0, sc.one_over_dx, w[0]);
        axis_weights(x[1], sax[1], r1, sc.one_over_dx, w[1]);
        axis_weights(x[2], sax[2], r2, sc.one_over_dx, w[2]);
        float pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, Bp[9], L[9], gp[NPIG], gs = 0.0f;
        float R[3][3];   // [ok][r]: sum over (oi, oj) of Nx Ny v_r at z offset ok
#pragma unroll
        for (int c = 0; c < 9; ++c) { Bp[c] = 0.0f; L[c] = 0.0f; R[c / 3][c % 3] = 0.0f; }
#pragma unroll
        for (int c = 0; c < NPIG; ++c) gp[c] = 0.0f;
        const float inv_dx = sc.one_over_dx;
        // The stencil weights are separable, w = Nx Ny Nz, and so are the
        // velocity moments: per row (oi, oj) the z-weighted sum A_r = sum_ok
        // Nz v_r gives the row's share of pic, of the x and y columns of B
        // (pd0, pd1 are constant along the row) and of L's x and y columns; the
        // z columns come from R_ok = sum over rows of Nx Ny v_r.  The flip
        // velocity, latent and shear use the full weight per node.
        // 9 rows, each with its 3 nodes along z unrolled: one row of gather loads
        // live at a time.
        const float zc[3] = {sax[2][r2] - x[2], sax[2][r2 + 1] - x[2], sax[2][r2 + 2] - x[2]};
#pragma unroll 1
        for (int row = 0; row < 9; ++row) {
          const int oi = row / 3, oj = row - 3 * oi;
          const float Nx = oi == 0 ? w[0].N[0] : (oi == 1 ? w[0].N[1] : w[0].N[2]);
          const float dNx = oi == 0 ? w[0].dN[0] : (oi == 1 ? w[0].dN[1] : w[0].dN[2]);
          const float Ny = oj == 0 ? w[1].N[0] : (oj == 1 ? w[1].N[1] : w[1].N[2]);
          const float dNy = oj == 0 ? w[1].dN[0] : (oj == 1 ? w[1].dN[1] : w[1].dN[2]);
          const float nij = Nx * Ny;
          const int trow = ((r0 + oi) * GT + (r1 + oj)) * GT + r2;
          float A[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int ok = 0; ok < 3; ++ok) {
            const float Nz = w[2].N[ok];
            const float4* gn = &gtile[(trow + ok) * GT_ROW];
            const float4 n0 = gn[0];
            const float gvv[3] = {n0.x, n0.y, n0.z};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              A[r] += Nz * gvv[r];
              R[ok][r] += nij * gvv[r];
            }
            if constexpr (PHASE != 2) {
              const float4 n1 = gn[1], n2 = gn[2], n3 = gn[3];
              const float gfl[3] = {n0.w, n1.x, n1.y};
              const float gpg[NPIG] = {n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, n3.x};
              const float wgt = nij * Nz;
#pragma unroll
              for (int r = 0; r < 3; ++r) fl[r] += gfl[r] * wgt;
#pragma unroll
              for (int c = 0; c < NPIG; ++c) gp[c] += gpg[c] * wgt;
              gs += n3.y * wgt;
            }
          }
          const float dij0 = dNx * Ny, dij1 = Nx * dNy;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            L[3 * r + 0] += dij0 * A[r];
            L[3 * r + 1] += dij1 * A[r];
          }
          if constexpr (PHASE != 2) {
            const float b0 = (sax[0][r0 + oi] - x[0]) * nij, b1 = (sax[1][r1 + oj] - x[1]) * nij;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              pic[r] += nij * A[r];
              Bp[3 * r + 0] += b0 * A[r];
              Bp[3 * r + 1] += b1 * A[r];
            }
          }
        }
#pragma unroll
        for (int ok = 0; ok < 3; ++ok) {
          const float b2 = zc[ok] * w[2].N[ok], l2 = w[2].dN[ok];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            if constexpr (PHASE != 2) Bp[3 * r + 2] += b2 * R[ok][r];
            L[3 * r + 2] += l2 * R[ok][r];
          }
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) L[c] *= inv_dx;
        if (sc.flip_ratio >= 0.0f) {
#pragma unroll
          for (int c = 0; c < 9; ++c) Bp[c] = 0.0f;  // PIC/FLIP: Bmat unused
        }
        if constexpr (PHASE == 2) {
#pragma unroll
          for (int c = 0; c < 3; ++c) pic[c] = picB[c];
        }
        // ---------------- shear intensity and pigment mixing (g2p.cpp:94-130)
        float sn = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float sv = 0.5f * (L[3 * r + c] + L[3 * c + r]);
            sn += sv * sv;
          }
        const float diff_new = sqrtf(sn);
        const float tt = fmaxf(0.0f, fminf(1.0f, fdiv(gs - sc.D_edge0, sc.D_edge1 - sc.D_edge0)));
        const float mixf = tt * tt * (3.0f - 2.0f * tt);
        const float smf = fminf(fmaxf(dynD * mixf * sc.dt, 0.0f), 1.0f);
        if constexpr (PHASE != 2) {
#pragma unroll
          for (int c = 0; c < NPIG; ++c) pig[c] = (1.0f - smf) * pig[c] + smf * gp[c];
          const float ps = ((pig[0] + pig[1]) + pig[2]) + pig[3];
          if (ps > 1e-6f) {
            const float ip = rcp_approx(ps);
            pig[0] *= ip; pig[1] *= ip; pig[2] *= ip; pig[3] *= ip;
          }
        }
        if constexpr (PHASE == 1) {
          // first half of MUSL (musl.cpp:13-26): particle velocity update; park
          // the PIC velocity for the advection of the second half
          if (valid) {
            const float a = -sc.flip_ratio;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              float vn;
              if (sc.flip_ratio < -1.0f) vn = pic[c];
              else if (sc.flip_ratio < 0.0f) vn = a * (v[c] + fl[c]) + (1.0f - a) * pic[c];
              else vn = sc.flip_ratio * (v[c] + fl[c]) + (1.0f - sc.flip_ratio) * pic[c];
              Bf[(P_V + c) * cap + j] = vn;
              Bf[(P_X + c) * cap + j] = pic[c];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) Bf[(P_C + c) * cap + j] = Bp[c];
#pragma unroll
            for (int c = 0; c < NPIG; ++c) Bf[(P_PIG + c) * cap + j] = pig[c];
            Bf[P_DIFF * cap + j] = diff_new;
            d.pb_id[j] = id;
          }
          continue;
        }
        // ---------------- F update + return map (deformation_update.cpp:33-45)
        float Ft[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            Ft[3 * r + c] = F[3 * r + c] + ((sc.dt * L[3 * r + 0]) * F[0 + c] + (sc.dt * L[3 * r + 1]) * F[3 + c] +
                                            (sc.dt * L[3 * r + 2]) * F[6 + c]);
        float U[9], lam[3], V[9];
        principal3(Ft, U, lam, V);
        float h[3] = {hencky_of(lam[0]), hencky_of(lam[1]), hencky_of(lam[2])};
        float eps = eps_in, dgr = 0.0f;
        uint32_t err = 0;
        const float h_trial[3] = {h[0], h[1], h[2]};
        bool deferred = false;
        uint64_t dm = 0;
        int cq_first = 0;   // lane 0: the wavefront's first slot in the queue segment
        if constexpr (LEARNED) {
          // append the wavefront's candidates to this XCD's queue segment (one
          // atomic per wavefront; slots in lane order).  The claim is issued
          // here and read when the records are written, after the stress and
          // the advection, which hide its round trip.
          deferred = valid && trial_state(sc, h, eps_in).cand;
          dm = __ballot(deferred);
          if (dm && __lane_id() == 0) cq_first = atomicAdd(cq_counter, __popcll(dm));
        }
        // VOL: the eps_pl_vol / eps_pl_vol_pradhana rows, read and written for
        // the scenes whose model tracks them (DPVisc, capped von Mises)
        float epsv = 0.0f, prad = 0.0f;
        if constexpr (VOL) {
          if (sc.track_vol) { epsv = A[P_EPSV * cap + src]; prad = A[P_PRAD * cap + src]; }
          if (sc.plastic_model == MPMX_PLASTIC_VMVISC)
            plastic = vmvisc_principal(sc, h, eps, dgr, err, &newton, sc.track_vol ? &epsv : nullptr);
          else if (sc.plastic_model == MPMX_PLASTIC_DPVISC)
            plastic = dpvisc_principal(sc, h, eps, epsv, prad, dgr, err, &newton);
        } else {
          if (!deferred && sc.plastic_model == MPMX_PLASTIC_VMVISC)
            plastic = vmvisc_principal(sc, h, eps, dgr, err, &newton);
        }
        if (!valid) newton = 0;
        if (err && valid) atomicOr(&d.err[0], err);
        if (plastic == 1 && valid && d.harvest) harvest_pair(d, sc, h_trial, eps_in, h, dgr);
        float Fn[9];
        if (plastic) {
          f_from_principal(Ft, U, V, lam, h_trial, h, Fn);
        } else {
#pragma unroll
          for (int c = 0; c < 9; ++c) Fn[c] = Ft[c];
        }
        float tau[6];
        kirchhoff_from_principal(sc, U, h, tau);
        // ---------------- position update (position_update.cpp:8-90)
        float xn[3], vn[3];
        const float a = -sc.flip_ratio;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          xn[c] = x[c] + sc.dt * pic[c];
          if (sc.use_musl) vn[c] = v[c];
          else if (sc.flip_ratio < -1.0f) vn[c] = pic[c];
          else if (sc.flip_ratio < 0.0f) vn[c] = a * (v[c] + fl[c]) + (1.0f - a) * pic[c];
          else vn[c] = sc.flip_ratio * (v[c] + fl[c]) + (1.0f - sc.flip_ratio) * pic[c];
        }
        bool alive = true;
        bool bad = false;
#pragma unroll
        for (int c = 0; c < 3; ++c) bad |= !isfinite(xn[c]) || !isfinite(vn[c]);
        if (bad) {
          alive = false;
#pragma unroll
          for (int c = 0; c < 3; ++c) { vn[c] = 0.0f; xn[c] = sc.use_pb ? sc.pb_min[c] : 0.0f; }
        } else {
          if (xn[1] < 0.0f) { xn[1] = 0.0f; if (vn[1] < 0.0f) vn[1] = 0.0f; }
          if (sc.use_pb) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              if (xn[c] < sc.pb_min[c]) { xn[c] = sc.pb_min[c]; alive = false; vn[c] = 0.0f; }
              else if (xn[c] > sc.pb_max[c]) { xn[c] = sc.pb_max[c]; alive = false; vn[c] = 0.0f; }
            }
          }
          if (alive) {
            newkey = particle_key(sc, xn, d.dead_key, newcell);
            if (newkey == d.dead_key) alive = false;
            okey = cell_order_key(sc, xn);
          }
          if (!alive) { vn[0] = vn[1] = vn[2] = 0.0f; }
        }
        if (!alive || !valid) { newkey = d.dead_key; newcell = -1; }
        if constexpr (LEARNED) {
          // the candidates' queue records.  A segment starts at the first
          // particle slot of its XCD's share of the block list and receives
          // at most one record per particle of that share, so it cannot run
          // into the next one or past the capacity; the check only guards.
          if (dm) {
            const int64_t slot = cq_base + __shfl(cq_first, 0) + __popcll(dm & ((1ull << __lane_id()) - 1ull));
            if (deferred && slot < d.cq_cap) {
              float4* r = d.cq + cq_word(slot, 0);
              r[0] = make_float4(__int_as_float((int)j), __int_as_float(s), eps_in, lam[0]);
              r[CQ_GROUP] = make_float4(lam[1], lam[2], Ft[0], Ft[1]);
              r[2 * CQ_GROUP] = make_float4(Ft[2], Ft[3], Ft[4], Ft[5]);
              r[3 * CQ_GROUP] = make_float4(Ft[6], Ft[7], Ft[8], V[0]);
              r[4 * CQ_GROUP] = make_float4(V[1], V[2], V[3], V[4]);
              r[5 * CQ_GROUP] = make_float4(V[5], V[6], V[7], V[8]);
            } else if (deferred) {
              atomicOr(&d.err[0], ERR_QUEUE_OVERFLOW);
            }
          }
        }
        // ---------------- write the new state in slot j
        if (valid) {
#pragma unroll
          for (int c = 0; c < 3; ++c) { Bf[(P_X + c) * cap + j] = xn[c]; Bf[(P_V + c) * cap + j] = vn[c]; }
          if (!deferred) {
#pragma unroll
            for (int c = 0; c < 9; ++c) Bf[(P_F + c) * cap + j] = Fn[c];
            Bf[P_EPS * cap + j] = eps;
            Bf[P_DG * cap + j] = dgr;
#pragma unroll
            for (int c = 0; c < 6; ++c) Bf[(P_TAU + c) * cap + j] = tau[c];
            if constexpr (VOL) {
              if (sc.track_vol) { Bf[P_EPSV * cap + j] = epsv; Bf[P_PRAD * cap + j] = prad; }
            }
          }
          if constexpr (PHASE == 0) {
#pragma unroll
            for (int c = 0; c < 9; ++c) Bf[(P_C + c) * cap + j] = Bp[c];
#pragma unroll
            for (int c = 0; c < NPIG; ++c) Bf[(P_PIG + c) * cap + j] = pig[c];
            Bf[P_DIFF * cap + j] = diff_new;
          }
          d.pb_id[j] = id;
        }
      }
      count_for_sort(d, valid, newkey, newcell, j, okey);
      const uint64_t pm = __ballot(valid && plastic == 1);
      if (__lane_id() == 0 && pm) atomicAdd(&plastic_wg, (unsigned)__popcll(pm));
      if (newton) atomicAdd(&newton_wg, (unsigned)newton);
    }
    k = next_blk;   // written by thread 0 before the staging barrier of this block
  }
  // the last (unused) claim has completed (vmcnt(0)) before this workgroup
  // counts itself done: a claim landing after the reset would skip a block of
  // the next launch
  wait_vm_all();
  if (threadIdx.x == 0) {
    if (atomicAdd(q_done, 1) == walkers - 1) {
      atomicExch(q_next, 0);
      atomicExch(q_done, 0);
    }
  }
  if constexpr (PHASE != 1) copy_dead_share(d);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (plastic_wg) atomicAdd(&d.stats[ST_PLASTIC_CUR], (unsigned long long)plastic_wg);
    if (newton_wg) atomicAdd(&d.stats[ST_NEWTON], (unsigned long long)newton_wg);
  }
}

// The learned closure over the queued candidates (one record per lane, 64 per
// wavefront on MFMA): the trial F's principal frame as G2P found it, learned map
// with the exact map as per-lane fallback, then F = U exp(h') V^T, eps_pl_dev,
// dgamma / dt and the Kirchhoff stress into the particle's output slot.  The
// queue's order (wavefront atomics) does not change any result: every record
// is mapped on its own (the network's rows are independent).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_closure(DevPtrs d, const SceneDev* __restrict__ scenes,
                                                 const uint8_t* __restrict__ mlp_blob) {
  __shared__ __align__(16) G2PLearnedLds lrn;
  __shared__ unsigned int plastic_wg, fallback_wg, learned_wg, newton_wg;
  // the XCD segments, concatenated: record q of the whole queue lies in segment
  // x with pre[x] <= q < pre[x + 1], at slot start[x] + q - pre[x].  Starts and
  // counts come from G2P (the block list is being rebuilt by the sort, which
  // runs beside this kernel).
  int pre[CQ_SEGS + 1], start[CQ_SEGS];   // slots and counts are below the particle capacity (int)
  pre[0] = 0;
#pragma unroll
  for (int x = 0; x < CQ_SEGS; ++x) {
    start[x] = d.cq_start[x];
    pre[x + 1] = pre[x] + d.cq_cnt[x * CQ_CNT_STRIDE];
  }
  const int n = pre[CQ_SEGS];
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) atomicAdd(&d.stats[ST_QUEUED], (unsigned long long)n);
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;   // whole workgroup: nothing queued for it
  if (threadIdx.x == 0) { plastic_wg = 0u; fallback_wg = 0u; learned_wg = 0u; newton_wg = 0u; }
  mlp_blob_to_lds<256>(mlp_blob, lrn.net);
  __syncthreads();
  float* __restrict__ Bf = d.pb;
  const int64_t cap = d.cap;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
    const int q0 = (int)base + (int)threadIdx.x;
    const int q = q0 < n ? q0 : n - 1;
    int slot = start[0] + q;   // segment by a select chain (pre, start stay uniform)
#pragma unroll
    for (int k = 1; k < CQ_SEGS; ++k)
      if (q >= pre[k]) slot = start[k] + (q - pre[k]);
    const bool valid = q0 < n && slot < d.cq_cap;      // (the capacity check only guards: G2P stays inside it)
    const float4* r = d.cq + cq_word(min((int64_t)slot, d.cq_cap - 1), 0);
    const float4 r0 = r[0];
    const float2 r1 = *reinterpret_cast<const float2*>(r + CQ_GROUP);
    const int64_t j = __float_as_int(r0.x);
    const SceneDev& sc = scenes[__float_as_int(r0.y)];
    const float eps_in = r0.z;
    const float lam[3] = {r0.w, r1.x, r1.y};
    float h[3] = {hencky_of(lam[0]), hencky_of(lam[1]), hencky_of(lam[2])};
    const float h_trial[3] = {h[0], h[1], h[2]};
    float eps = eps_in, dgr = 0.0f;
    uint32_t err = 0;
    int newton = 0, plastic = 0;
    bool cand = false;
    const bool ok = learned_return_map(sc, lrn, h, eps, dgr, valid, cand);
    // the trial F and V only after the network (its LDS fences keep these loads
    // below it), so that they do not hold registers across it
    const float4 r1b = r[CQ_GROUP], r2 = r[2 * CQ_GROUP], r3 = r[3 * CQ_GROUP], r4 = r[4 * CQ_GROUP],
                 r5 = r[5 * CQ_GROUP];
    const float Ft[9] = {r1b.z, r1b.w, r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z};
    const float V[9] = {r3.w, r4.x, r4.y, r4.z, r4.w, r5.x, r5.y, r5.z, r5.w};
    float U[9];
    principal_u(Ft, V, U);
    if (ok) plastic = 1;
    else if (sc.plastic_model == 5) plastic = vmvisc_principal(sc, h, eps, dgr, err, &newton);
    if (valid) {
      if (err) atomicOr(&d.err[0], err);
      if (plastic == 1 && d.harvest) harvest_pair(d, sc, h_trial, eps_in, h, dgr);
      float Fn[9];
      if (plastic) f_from_principal(Ft, U, V, lam, h_trial, h, Fn);
      else {
#pragma unroll
        for (int c = 0; c < 9; ++c) Fn[c] = Ft[c];
      }
      float tau[6];
      kirchhoff_from_principal(sc, U, h, tau);
#pragma unroll
      for (int c = 0; c < 9; ++c) Bf[(P_F + c) * cap + j] = Fn[c];
      Bf[P_EPS * cap + j] = eps;
      Bf[P_DG * cap + j] = dgr;
#pragma unroll
      for (int c = 0; c < 6; ++c) Bf[(P_TAU + c) * cap + j] = tau[c];
    }
    const uint64_t pm = __ballot(valid && plastic == 1);
    const uint64_t lm = __ballot(valid && ok);
    const uint64_t fm = __ballot(valid && cand && !ok);
    if (__lane_id() == 0) {
      if (pm) atomicAdd(&plastic_wg, (unsigned)__popcll(pm));
      if (lm) atomicAdd(&learned_wg, (unsigned)__popcll(lm));
      if (fm) atomicAdd(&fallback_wg, (unsigned)__popcll(fm));
    }
    if (valid && newton) atomicAdd(&newton_wg, (unsigned)newton);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (plastic_wg) atomicAdd(&d.stats[ST_PLASTIC_CUR], (unsigned long long)plastic_wg);
    if (newton_wg) atomicAdd(&d.stats[ST_NEWTON], (unsigned long long)newton_wg);
    if (learned_wg) atomicAdd(&d.stats[ST_LEARNED_CUR], (unsigned long long)learned_wg);
    if (fallback_wg) atomicAdd(&d.stats[ST_FALLBACK_CUR], (unsigned long long)fallback_wg);
  }
}

// Learned and exact return maps side by side on given trial states (principal
// Hencky strains h1 h2 h3 and eps_pl_dev per row), with scene 0's material: the
// closure's accuracy measured in principal space, where FP32 rounding of F
// does not hide it.  learned4 = (h', dgamma) when accepted, else NaN;
// exact4 = (h', dgamma) of the exact map.
__global__ void __launch_bounds__(256) k_closure_eval(const SceneDev* __restrict__ scenes,
                                                      const uint8_t* __restrict__ mlp_blob,
                                                      const float* __restrict__ trial4, int64_t n,
                                                      float* __restrict__ learned4, float* __restrict__ exact4,
                                                      uint8_t* __restrict__ accepted) {
  __shared__ __align__(16) G2PLearnedLds lrn;
  mlp_blob_to_lds<256>(mlp_blob, lrn.net);
  __syncthreads();
  const SceneDev& sc = scenes[0];
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n;
    const int64_t r = valid ? i : n - 1;
    float h[3] = {trial4[4 * r], trial4[4 * r + 1], trial4[4 * r + 2]};
    float eps = trial4[4 * r + 3], dgr = 0.0f;
    bool cand = false;
    const bool ok = learned_return_map(sc, lrn, h, eps, dgr, valid, cand);
    float he[3] = {trial4[4 * r], trial4[4 * r + 1], trial4[4 * r + 2]};
    float epse = trial4[4 * r + 3], dge = 0.0f;
    uint32_t err = 0;
    const int pl = sc.plastic_model == 5 ? vmvisc_principal(sc, he, epse, dge, err) : 0;
    if (valid) {
      const float nan = __uint_as_float(0x7FC00000u);
      learned4[4 * i] = ok ? h[0] : nan;
      learned4[4 * i + 1] = ok ? h[1] : nan;
      learned4[4 * i + 2] = ok ? h[2] : nan;
      learned4[4 * i + 3] = ok ? dgr * sc.dt : nan;
      exact4[4 * i] = he[0];
      exact4[4 * i + 1] = he[1];
      exact4[4 * i + 2] = he[2];
      exact4[4 * i + 3] = pl ? dge * sc.dt : 0.0f;
      accepted[i] = ok ? 1 : 0;
    }
  }
}

// ------------------------------------------------------------------ k_p2g
// Work item = (particle block of the new sort, channel group); one wavefront
// per item, lane = cell of the block (4x4x4 bins):
//   0: latent 6 + shear intensity + mass (sum of weights)
//   1: APIC momentum    2: Kirchhoff force    3: latent 0..2    4: latent 3..5
// Each lane walks its cell's particles (block-sorted, so the stencil base is
// the cell for all of them) and accumulates the 27 stencil nodes x the group's
// channels in registers.  The 64 lanes' stencils then merge into the block's
// 6x6x6 node tile in registers, one axis at a time (see p2g_merge), so that
// every tile node ends with exactly one owning lane; the tile leaves through a
// small LDS staging area (see p2g_group).  Items are independent, so there is
// no workgroup-wide phase: many small waves per CU overlap their particle
// loads with other waves' arithmetic.  Five groups of three FP32 channels keep
// the accumulators at 81 registers, so that three wavefronts fit per SIMD
// (168 VGPRs); four groups (a six-channel latent group, 250 VGPRs, two
// wavefronts per SIMD) measured 131 us against 110 us on the 1M-particle
// scene.
constexpr int P2G_GROUPS = 5;
constexpr int P2G_STATIC = 4;   // items per wavefront assigned statically before the queue
// Accumulator channels of group G (three each): latent 6 + shear + mass |
// momentum | force | latent 0..2 | latent 3..5 | (MUSL) momentum.  Every
// channel is an FP32 sum in a fixed order (the particles of a cell in the
// sort's position order, then the fixed merge and tile order), so it is
// reproducible bit for bit.
__host__ __device__ constexpr int p2g_nch(int) { return 3; }

// Per-item inputs shared by the channel-group bodies.
struct P2GItem {
  const float* P;      // particle side holding the sorted step's state
  const float* PB;     // MUSL: the first half's output side (indexed by sort slot)
  int64_t cap;
  const int* perm;     // permutation
  int start;           // the block's first slot
  int cnt, kmax;       // particles of this lane's cell, deepest cell of the wave
  const float* xs;     // node coordinates of the scene (x axis, then y, then z)
  int n0, n1, n2;      // lattice nodes per axis
  int ix, iy, iz;      // first node of the lane's stencil per axis
  float inv_dx, apicDinv;
};

// Per-particle inputs of channel group G: position, then the group's payload.
// Group 6 is MUSL's second scatter of momentum (musl.cpp:31-87): position from
// the step's input side (through the permutation), v and B from the first
// half's output (slot j of the sort).
template <int G>
struct P2GLoad {
  static constexpr int NQ = G == 0 ? 2 : (G == 1 || G == 6 ? 12 : (G == 2 ? 6 : 3));
  float x[3];
  float q[NQ > 0 ? NQ : 1];
};

// Row `row` of slot i: a wave-uniform row base plus a 32-bit byte offset per
// lane, which the compiler issues as a scalar-base global load with a 32-bit
// vector offset (no 64-bit per-lane address arithmetic).  Slots stay below
// 2^30, which the allocation checks.
__device__ __forceinline__ float ld_slot(const float* __restrict__ P, int64_t cap, int row, int i) {
  const char* base = (const char*)(P + row * cap);
  return *(const float*)(base + (uint64_t)((uint32_t)i * 4u));
}
__device__ __forceinline__ int ld_slot(const int* __restrict__ A, int i) {
  return *(const int*)((const char*)A + (uint64_t)((uint32_t)i * 4u));
}

// The particle stream of an item is prefetched one rank ahead into LDS with
// direct-to-LDS loads (global_load_lds: lane l's dword of a row lands at word l
// of the row's 64-word LDS line), so that the prefetch holds no registers.  A
// ring of two buffers, each of P2G_RING_ROWS lines (x, the group's payload, and
// the permutation entry of the rank after), lives in the first words of the
// wavefront's staging area, which is zeroed again once the item's loop is done.
constexpr int P2G_RING_ROWS = 16;
// Lines of a buffer: 0..2 x, 3..3+NQ-1 payload, 15 the permutation entry.
template <int G>
__device__ __forceinline__ void p2g_fetch(const float* __restrict__ P, const float* __restrict__ PB, int64_t cap,
                                          const int* __restrict__ perm, int src, int j, int j_after, float* buf) {
#pragma unroll
  for (int c = 0; c < 3; ++c) dma_dword(P + (P_X + c) * cap, src, buf + 64 * c);
  if constexpr (G == 6) {
#pragma unroll
    for (int c = 0; c < 3; ++c) dma_dword(PB + (P_V + c) * cap, j, buf + 64 * (3 + c));
#pragma unroll
    for (int c = 0; c < 9; ++c) dma_dword(PB + (P_C + c) * cap, j, buf + 64 * (6 + c));
  } else if constexpr (G == 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) dma_dword(P + (P_V + c) * cap, src, buf + 64 * (3 + c));
#pragma unroll
    for (int c = 0; c < 9; ++c) dma_dword(P + (P_C + c) * cap, src, buf + 64 * (6 + c));
  } else if constexpr (G == 2) {
#pragma unroll
    for (int c = 0; c < 6; ++c) dma_dword(P + (P_TAU + c) * cap, src, buf + 64 * (3 + c));
  } else if constexpr (G == 3 || G == 4) {
#pragma unroll
    for (int c = 0; c < 3; ++c) dma_dword(P + (P_PIG + 3 * (G - 3) + c) * cap, src, buf + 64 * (3 + c));
  } else if constexpr (G == 0) {
    dma_dword(P + (P_PIG + 6) * cap, src, buf + 64 * 3);
    dma_dword(P + P_DIFF * cap, src, buf + 64 * 4);
  }
  dma_dword(perm, j_after, buf + 64 * (P2G_RING_ROWS - 1));
}

template <int G>
__device__ __forceinline__ int p2g_read(const float* buf, int lane, P2GLoad<G>& o) {
#pragma unroll
  for (int c = 0; c < 3; ++c) o.x[c] = buf[64 * c + lane];
#pragma unroll
  for (int c = 0; c < P2GLoad<G>::NQ; ++c) o.q[c] = buf[64 * (3 + c) + lane];
  return __float_as_int(buf[64 * (P2G_RING_ROWS - 1) + lane]);
}

// Accumulate this lane's cell into acc[27][3] for channel group G.  The loop
// runs kmax (wave-uniform) times, step k = rank in cell; a lane past its own
// count contributes exact zeros through selects, so the accumulators never sit
// behind a branch.  With the block's rank-major layout the active lanes of a
// step read consecutive slots; every load is in bounds for every lane (a lane
// without rank k reads the block's first slot).  The next particle's inputs
// are loaded while the current one is accumulated.
//
// The per-node arithmetic uses the separable
// structure of the quadratic B-spline stencil: w = Nx Ny Nz with the 9 products
// Ny Nz precomputed, the APIC term v + C (X_node - x_p) split into per-axis
// parts (one add per axis per node), and the force gradients as products of
// per-axis factors.  (Rounding differs from the product order of the reference
// by O(ulp); the mass channel, which decides node activation, keeps the
// reference's product order and its tiny-weight cut.)
template <int G>
__device__ __forceinline__ void p2g_accumulate(const P2GItem& it, float (&acc)[27][p2g_nch(G)], float* ring) {
  const float* __restrict__ P = it.P;
  const int64_t cap = it.cap;
  const float inv_dx = it.inv_dx;
  const int lane = threadIdx.x;
  // Slot of this lane's rank k in the block's rank-major layout: the block
  // start + the slots of the ranks below k (kb, wave-uniform) + the cells
  // below this one that hold a rank k.  A lane without rank k reads the
  // block's first slot instead (in bounds; its contribution is zeroed).
  const uint64_t below = lanemask_lt();
  uint64_t m = __ballot(it.cnt > 0);
  const int j0 = it.cnt > 0 ? it.start + (int)__popcll(m & below) : it.start;
  int kb = (int)__popcll(m);
  m = __ballot(it.cnt > 1);
  int j_next = it.cnt > 1 ? it.start + kb + (int)__popcll(m & below) : it.start;   // rank k + 1
  kb += (int)__popcll(m);
  p2g_fetch<G>(P, it.PB, cap, it.perm, ld_slot(it.perm, j0), j0, j_next, ring);
  // node coordinates of this cell's stencil (clamped for empty edge cells),
  // loaded here so that they hold no registers outside the loop
  float X[3][3];
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    X[0][o] = it.xs[min(it.ix + o, it.n0 - 1)];
    X[1][o] = it.xs[it.n0 + min(it.iy + o, it.n1 - 1)];
    X[2][o] = it.xs[it.n0 + it.n1 + min(it.iz + o, it.n2 - 1)];
  }
  for (int k = 0; k < it.kmax; ++k) {
    const bool act = k < it.cnt;
    // rank k (fetched during the previous step) and the permutation entry of rank k + 1
    float* buf = ring + (k & 1) * (P2G_RING_ROWS * 64);
    wait_vm_all();
    P2GLoad<G> cur;
    const int src_next = p2g_read<G>(buf, lane, cur);
    if (k + 1 < it.kmax) {
      m = __ballot(it.cnt > k + 2);
      const int j_after = it.cnt > k + 2 ? it.start + kb + (int)__popcll(m & below) : it.start;   // rank k + 2
      kb += (int)__popcll(m);
      p2g_fetch<G>(P, it.PB, cap, it.perm, src_next, j_next, j_after, ring + ((k + 1) & 1) * (P2G_RING_ROWS * 64));
      j_next = j_after;
    }
    const float x0 = act ? cur.x[0] : X[0][1];
    const float x1 = act ? cur.x[1] : X[1][1];
    const float x2 = act ? cur.x[2] : X[2][1];
    float u[9], N[9];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      u[o] = spline_u(x0, X[0][o], inv_dx);
      u[3 + o] = spline_u(x1, X[1][o], inv_dx);
      u[6 + o] = spline_u(x2, X[2][o], inv_dx);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      N[3 * a] = spline_N_at<0>(u[3 * a]);
      N[3 * a + 1] = spline_N_at<1>(u[3 * a + 1]);
      N[3 * a + 2] = spline_N_at<2>(u[3 * a + 2]);
    }
    // group 0: whether the mass cut below can act in this step (see there),
    // from the factors before the x ones of lanes past their count are zeroed
    bool near_face = false;
    if constexpr (G == 0)
      near_face = fminf(fminf(fminf(N[0], N[2]), fminf(N[3], N[5])), fminf(N[6], N[8])) < 1e-8f;
#pragma unroll
    for (int o = 0; o < 3; ++o) N[o] = act ? N[o] : 0.0f;
    if constexpr (G == 0) {
      // latent 6 and shear intensity: weighted scalars, q * Nx folded per
      // offset; mass: the weight in the reference's product order (Nx Ny) Nz
      // and with its cut w > 1e-25 (p2g.cpp:36-43), which decides whether a
      // node becomes active
      float fq[2][3];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float q = act ? cur.q[c] : 0.0f;
#pragma unroll
        for (int o = 0; o < 3; ++o) fq[c][o] = q * N[o];
      }
#pragma unroll
      for (int n = 0; n < 27; ++n) {
        const float nyz = N[3 + (n / 3) % 3] * N[6 + n % 3];
        acc[n][0] += fq[0][n / 9] * nyz;
        acc[n][1] += fq[1][n / 9] * nyz;
      }
      // The cut acts only on a weight in (0, 1e-25], which needs a 1-D factor
      // below 1e-8: a particle within 1.5e-4 cells of a face of its cell (the
      // outer offsets' factors 0.5 t^2; the middle one is >= 0.5).  Lanes past
      // their count carry node 1's coordinates, whose factors are 0.125 and
      // 0.75.  When no lane of the wave is near a face, every nonzero weight
      // is at least 1e-24 and the compare is left out; the sums are the same.
      if (__ballot(near_face) == 0) {
#pragma unroll
        for (int n = 0; n < 27; ++n) acc[n][2] += (N[n / 9] * N[3 + (n / 3) % 3]) * N[6 + n % 3];
      } else {
#pragma unroll
        for (int n = 0; n < 27; ++n) {
          const float w = (N[n / 9] * N[3 + (n / 3) % 3]) * N[6 + n % 3];
          acc[n][2] += w > 1e-25f ? w : 0.0f;
        }
      }
    } else {
      float NyNz[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) NyNz[p] = N[3 + p / 3] * N[6 + p % 3];
      if constexpr (G == 1 || G == 6) {
        // v + C (X - x) = [v + C_.x (X_oi - x)] + C_.y (Y_oj - y) + C_.z (Z_ok - z)
        float ax[3][3], ay[3][3], az[3][3];   // [offset][component]
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          const float px = X[0][o] - x0, py = X[1][o] - x1, pz = X[2][o] - x2;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float v = act ? cur.q[r] : 0.0f;
            const float cx = act ? cur.q[3 + 3 * r] * it.apicDinv : 0.0f;
            const float cy = act ? cur.q[3 + 3 * r + 1] * it.apicDinv : 0.0f;
            const float cz = act ? cur.q[3 + 3 * r + 2] * it.apicDinv : 0.0f;
            ax[o][r] = v + cx * px;
            ay[o][r] = cy * py;
            az[o][r] = cz * pz;
          }
        }
        // w (a + b + c) = (Nx Ny) (Nz (a + b) + Nz c): two FMAs per node and component
        float nzaz[3][3];
#pragma unroll
        for (int ok = 0; ok < 3; ++ok)
#pragma unroll
          for (int r = 0; r < 3; ++r) nzaz[ok][r] = N[6 + ok] * az[ok][r];
#pragma unroll
        for (int oi = 0; oi < 3; ++oi)
#pragma unroll
          for (int oj = 0; oj < 3; ++oj) {
            const float wxy = N[oi] * N[3 + oj];
            float sxy[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) sxy[r] = ax[oi][r] + ay[oj][r];
#pragma unroll
            for (int ok = 0; ok < 3; ++ok)
#pragma unroll
              for (int r = 0; r < 3; ++r) {
                const int n = 9 * oi + 3 * oj + ok;
                acc[n][r] = fmaf(wxy, fmaf(N[6 + ok], sxy[r], nzaz[ok][r]), acc[n][r]);
              }
          }
      } else if constexpr (G == 2) {
        // f_r = tau_r0 dNx Ny Nz + tau_r1 Nx dNy Nz + tau_r2 Nx Ny dNz (gradient / dx)
        float dN[9];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          dN[3 * a] = spline_dN_at<0>(u[3 * a]) * inv_dx;
          dN[3 * a + 1] = spline_dN_at<1>(u[3 * a + 1]) * inv_dx;
          dN[3 * a + 2] = spline_dN_at<2>(u[3 * a + 2]) * inv_dx;
        }
        float tau[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) tau[c] = act ? cur.q[c] : 0.0f;
        const float T[3][3] = {{tau[0], tau[3], tau[4]}, {tau[3], tau[1], tau[5]}, {tau[4], tau[5], tau[2]}};
        // per (oj, ok): f_r = dNx (tau_r0 Ny Nz) + Nx (tau_r1 dNy Nz + tau_r2 Ny dNz),
        // two FMAs per node and component
#pragma unroll
        for (int p = 0; p < 9; ++p) {
          const float b = dN[3 + p / 3] * N[6 + p % 3], c = N[3 + p / 3] * dN[6 + p % 3];
          float Ar[3], Br[3];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            Ar[r] = T[r][0] * NyNz[p];
            Br[r] = fmaf(T[r][2], c, T[r][1] * b);
          }
#pragma unroll
          for (int oi = 0; oi < 3; ++oi)
#pragma unroll
            for (int r = 0; r < 3; ++r) acc[9 * oi + p][r] = fmaf(N[oi], Br[r], fmaf(dN[oi], Ar[r], acc[9 * oi + p][r]));
        }
      } else {
        // latent / shear: weighted scalars, q * Nx folded per offset
        constexpr int NC = p2g_nch(G);
        float fq[NC][3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const float q = act ? cur.q[c] : 0.0f;
#pragma unroll
          for (int o = 0; o < 3; ++o) fq[c][o] = q * N[o];
        }
#pragma unroll
        for (int n = 0; n < 27; ++n) {
          const int oi = n / 9, p = n % 9;
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[n][c] += fq[c][oi] * NyNz[p];
        }
      }
    }
  }
  wait_vm_all();
}

// ---- cross-lane merge of the 27-node stencils (lane = ci*16 + cj*4 + ck)
// The z and y axes merge in registers with DPP row shifts: the value of the
// lane 1 or 2 below within the 16-lane row (cells 1, 2 apart in z) or 4, 8
// below (cells 1, 2 apart in y).  A shift whose source lies outside the row
// reads zero (bound_ctrl), which for y is exactly the lanes without a source
// (cj < 1 or 2).  A z shift into a lane with ck < 1 or 2 reads the last cells
// of the previous z row instead, so it is scaled by a 0/1 lane mask: an FMA by
// 1 is the plain sum, an FMA by 0 leaves the accumulator as it is.  The x axis
// (cells 16 lanes apart) merges in the tile read-out of p2g_group, which sums
// the x sources of each tile node from the LDS staging area.  Every sum keeps
// the order own offset 0, then the cell below's offset 1, then the cell two
// below's offset 2.
template <int N>
__device__ __forceinline__ float row_shr(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + N, 0xF, 0xF, true));
}

// After the merge along an axis, slot 0 of the axis holds node `coord` of the
// lane's cell (all its contributions from this axis) and slots 1, 2 hold
// nodes 4 and 5, complete only on lanes with coord == 3.
__device__ __forceinline__ void merge_z(float (&v)[27], float m1, float m2) {
#pragma unroll
  for (int i = 0; i < 27; i += 3) {
    const float w0 = v[i], w1 = v[i + 1], w2 = v[i + 2];
    v[i] = fmaf(row_shr<2>(w2), m2, fmaf(row_shr<1>(w1), m1, w0));
    v[i + 1] = fmaf(row_shr<1>(w2), m1, w1);
  }
}
__device__ __forceinline__ void merge_y(float (&v)[27]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int i = 9 * a + c;
      const float w0 = v[i], w1 = v[i + 3], w2 = v[i + 6];
      v[i] = (w0 + row_shr<4>(w1)) + row_shr<8>(w2);
      v[i + 3] = w1 + row_shr<4>(w2);
    }
}

// Staging area of a channel, laid out by tile x: slot (xs, ys, zs) of the lane
// with cell (ci, cj, ck) holds node x = ci + xs, and goes to word
// ((9 xs + 3 ys + zs) * 6 + ci + xs) * 16 + 4 cj + ck.  In each slot's block of
// six x rows, the two rows no lane writes (x < xs or x > xs + 3) stay zero, so
// the read-out of tile node (tx, ty, tz) sums its three x sources, the words
// B + 864 xs with B = ((3 ys + zs) * 6 + tx) * 16 + 4 oj + ok (oj = min(ty, 3),
// ys = ty - oj, likewise for z), without masks: a source that does not exist
// reads an exact zero.
constexpr int P2G_STAGE_WORDS = 27 * 96;

// Per-lane constants of the read-out of tile positions p = lane + 64 r (r =
// 0..3, tile_pos order, p < 216): B of each position (two per word, 16 bits
// each) and, per position, the bit of its node block offset (tx, ty, tz) >> 2.
struct P2GReadout {
  unsigned b01, b23;
  unsigned bits;    // 8 bits per r: 1 << block offset index, 0 past 216
};

__device__ __forceinline__ P2GReadout p2g_readout(int lane) {
  P2GReadout ro;
  unsigned b[4];
  ro.bits = 0u;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int tx, ty, tz;
    tile_node_at(min(lane + 64 * r, PT3 - 1), tx, ty, tz);
    const int oj = min(ty, 3), ok = min(tz, 3);
    b[r] = (unsigned)((((ty - oj) * 3 + (tz - ok)) * 6 + tx) * 16 + oj * 4 + ok);
    if (lane + 64 * r < PT3) ro.bits |= (1u << ((((tx >> 2) * 2) + (ty >> 2)) * 2 + (tz >> 2))) << (8 * r);
  }
  ro.b01 = b[0] | (b[1] << 16);
  ro.b23 = b[2] | (b[3] << 16);
  return ro;
}

__device__ __forceinline__ int readout_base(const P2GReadout& ro, int r) {
  const unsigned w = r < 2 ? ro.b01 : ro.b23;
  return (int)((r & 1) ? w >> 16 : w & 0xFFFFu);
}

// Merged tile of block e: listed blocks below tile_cap store it (all 216
// nodes, zeros included) for the grid kernel to gather; later blocks add it to
// the grid accumulators with atomics.  Per channel, the lanes' z/y-merged
// stencils go to the wavefront's LDS staging area (stage[slot*64 + lane]); each
// lane then reads out tile positions lane, lane + 64, ..., summing their x
// sources, and the 216 nodes leave with four full-wavefront stores.
template <int G>
__device__ __forceinline__ void p2g_group(const P2GItem& it, const DevPtrs& d, const SceneDev& sc, int e, int B0i,
                                          int B0j, int B0k, unsigned& touched, float* stage_f,
                                          const P2GReadout* ro_tab) {
  const bool tiled = e < d.tile_cap;
  const int lane = threadIdx.x;
  float acc[27][p2g_nch(G)];
#pragma unroll
  for (int n = 0; n < 27; ++n)
#pragma unroll
    for (int c = 0; c < p2g_nch(G); ++c) acc[n][c] = 0.0f;
  p2g_accumulate<G>(it, acc, stage_f);
  // the ring's words become staging words again: zero them (rows that no lane
  // writes in the staging layout must read zero)
#pragma unroll
  for (int i = 0; i < 2 * P2G_RING_ROWS * 64 / 256; ++i)
    *(float4*)(stage_f + 256 * i + 4 * lane) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  // acc channel c goes to grid channel ch0 + c (group 0: latent 6, shear,
  // mass = A_PIG + 6, A_SHR, A_MASS)
  constexpr int ch0 = (G == 1 || G == 6) ? A_MOM : (G == 2 ? A_FRC : (G == 0 ? A_PIG + 6 : A_PIG + 3 * (G - 3)));
  static_assert(A_SHR == A_PIG + 7 && A_MASS == A_SHR + 1, "group 0's channels are contiguous");
  constexpr int nch = p2g_nch(G);
  const P2GReadout ro = ro_tab[lane];
  // z-merge masks: a z shift into a lane with ck < 1 or ck < 2 has no source in its row
  const float mz1 = (lane & 3) >= 1 ? 1.0f : 0.0f, mz2 = (lane & 3) >= 2 ? 1.0f : 0.0f;
  // all channels are read out before any of them is stored: LDS accesses
  // after a direct-to-LDS load wait for every outstanding vector memory
  // operation, so a store issued between two channels' staging would be
  // waited for
  float out[nch][4];
#pragma unroll
  for (int c = 0; c < nch; ++c) {
    float v[27];
#pragma unroll
    for (int n = 0; n < 27; ++n) v[n] = acc[n][c];
    merge_z(v, mz1, mz2);
    merge_y(v);
    // lane (ci, cj, ck) = 16 ci + 4 cj + ck writes slot n = (xs, ys, zs) at x
    // row ci + xs of the slot's block
#pragma unroll
    for (int n = 0; n < 27; ++n) stage_f[n * 96 + (n / 9) * 16 + lane] = v[n];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int B = readout_base(ro, r);
      out[c][r] = (stage_f[B] + stage_f[B + 864]) + stage_f[B + 1728];
    }
    __syncthreads();
  }
  if (G == 0) {
    // node blocks receiving mass from this block (channel 2 is the mass)
#pragma unroll
    for (int r = 0; r < 4; ++r) touched |= out[2][r] != 0.0f ? (ro.bits >> (8 * r)) & 0xFFu : 0u;
  }
#pragma unroll
  for (int c = 0; c < nch; ++c) {
    if (tiled) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (lane + 64 * r < PT3) d.tile_acc[((size_t)e * ACC_CH + ch0 + c) * PT3 + lane + 64 * r] = out[c][r];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (lane + 64 * r < PT3 && out[c][r] != 0.0f) {
          int tx, ty, tz;
          tile_node_at(lane + 64 * r, tx, ty, tz);
          const int gi = B0i + tx, gj = B0j + ty, gk = B0k + tz;
          atomicAdd(d.acc + ((size_t)node_block_of(sc, gi, gj, gk) * ACC_CH + ch0 + c) * BLK3 + node_local(gi, gj, gk),
                    out[c][r]);
        }
      }
    }
  }
}

// MUSL: only the momentum of group 6 (MUSL's second scatter), no activation.
template <bool MUSL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_p2g(DevPtrs d, const SceneDev* __restrict__ scenes,
                                            const float* __restrict__ axis, const float* __restrict__ P,
                                            const int* __restrict__ perm, const int4* __restrict__ pdesc,
                                            const int* __restrict__ cell_num, const float* __restrict__ PB) {
  constexpr int NG = MUSL ? 1 : P2G_GROUPS;
  // tile staging area of this workgroup's one wavefront (its rows that no lane
  // writes are zeroed once)
  __shared__ float stage[P2G_STAGE_WORDS];
  const int nblk = d.counters[CNT_PBLOCKS];
  const int parity = d.counters[CNT_GSTEP] & 1;
  const int lane = threadIdx.x;
  const int ci = lane >> 4, cj = (lane >> 2) & 3, ck = lane & 3;
  // read-out of the tile positions each lane stores, kept in LDS (no
  // registers held across the items)
  __shared__ P2GReadout ro_tab[64];
  ro_tab[lane] = p2g_readout(lane);
  for (int i = lane; i < P2G_STAGE_WORDS; i += 64) stage[i] = 0.0f;
  __syncthreads();
  // Items of this XCD: its contiguous segment of the block list x the groups.
  // Items differ several-fold in cost (cell depth, group), so after
  // P2G_STATIC static items (walker index w, w + W, ... for W walkers) a
  // wavefront takes its items from the XCD's queue: claim c is item
  // P2G_STATIC * W + c.  The claim for the item after next is made when an
  // item ends, so it has a whole item to return, and the wavefronts do not all
  // claim at once at the start.  The static items assume that every walker is
  // resident from the start: the launch is sized to the kernel's occupancy.
  // The last wavefront of the XCD to finish puts the queue back to zero for
  // the next launch (the grid is a multiple of 8 workgroups; workgroup w
  // belongs to walker w % 8).
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int walkers = gridDim.x / nx;
  const int seg = (nblk + nx - 1) / nx;
  const int e0 = xcd * seg;
  const int nitems = max(0, min(nblk, e0 + seg) - e0) * NG;
  int* const q_next = d.wq + xcd * CQ_CNT_STRIDE;
  int* const q_done = d.wq + (CQ_SEGS + xcd) * CQ_CNT_STRIDE;
  int item = (int)(blockIdx.x / nx);
  int claim = -1;   // lane 0: the claim made when the previous item ended
  int nstarted = 1;  // items started so far
  while (item < nitems) {
    const int e = e0 + item / NG;
    const int group = MUSL ? 6 : item % NG;
    // descriptor (key, first slot, count, scene) and this lane's cell, all by list position
    const int4 ds = pdesc[e];
    const int key = ds.x;
    const SceneDev& sc = scenes[ds.w];
    P2GItem it;
    it.cnt = cell_num[e * BLK3 + lane];
    int bi, bj, bk;
    block_coords(sc, key, bi, bj, bk);
    const int B0i = bi * 4, B0j = bj * 4, B0k = bk * 4;
    it.P = P;
    it.cap = d.cap;
    it.inv_dx = sc.one_over_dx;
    it.apicDinv = sc.apicDinv;
    it.perm = perm;
    it.start = ds.y;
    it.PB = PB;
    int kmax = it.cnt;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
    it.kmax = __builtin_amdgcn_readfirstlane(kmax);
    it.xs = axis + sc.axis_off;
    it.n0 = sc.dims[0];
    it.n1 = sc.dims[1];
    it.n2 = sc.dims[2];
    it.ix = B0i + ci;
    it.iy = B0j + cj;
    it.iz = B0k + ck;
    unsigned touched = 0u;
    if constexpr (MUSL) {
      p2g_group<6>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab);
    } else {
      switch (group) {
        case 0: p2g_group<0>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab); break;
        case 1: p2g_group<1>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab); break;
        case 2: p2g_group<2>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab); break;
        case 3: p2g_group<3>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab); break;
        default: p2g_group<4>(it, d, sc, e, B0i, B0j, B0k, touched, stage, ro_tab); break;
      }
    }
    if (group == 0) {
      // node blocks receiving mass from this block: OR of the lanes' bits
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) touched |= (unsigned)__shfl_xor((int)touched, off);
      if (lane < 8 && (touched & (1u << lane))) {
        const int nb = sc.node_block_base + ((bi + (lane >> 2)) * sc.bdims[1] + (bj + ((lane >> 1) & 1))) * sc.bdims[2] +
                       (bk + (lane & 1));
        activate_node_block(d, nb, parity);
      }
    }
    // the next item: a static one, or the claim made when the previous item ended
    const int prev = __builtin_amdgcn_readfirstlane(__shfl(claim, 0));
    const int next = nstarted < P2G_STATIC ? item + walkers : P2G_STATIC * walkers + prev;
    ++nstarted;
    if (nstarted >= P2G_STATIC && lane == 0) claim = atomicAdd(q_next, 1);   // the item after next
    item = next;
  }
  // this wavefront's last (unused) claim has completed (vmcnt(0)) before it
  // counts itself done: a claim landing after the reset would skip an item of
  // the next launch
  wait_vm_all();
  if (lane == 0 && atomicAdd(q_done, 1) == walkers - 1) {
    // every wavefront of this walker has drawn past the end: reset for the next launch
    atomicExch(q_next, 0);
    atomicExch(q_done, 0);
  }
}

// ------------------------------------------------------------------ k_grid
// One wavefront per active node block; lane = node.
// REDO: the grid update of the last step again from its merged tiles (which
// persist until the next scatter), with the scene constants as they are now
// (a changed dt); the list and statistics bookkeeping of the first pass is
// left alone.  Only for steps whose scatter used no atomics (the atomic
// accumulators are cleared by the first pass).
template <bool REDO>
__global__ void __launch_bounds__(256) k_grid(DevPtrs d) {
  __shared__ unsigned int active_wg;
  if (threadIdx.x == 0) active_wg = 0u;
  __syncthreads();
  const int step = d.counters[CNT_STEP];
  const int parity = d.counters[CNT_GSTEP] & 1;
  if (!REDO && blockIdx.x == 0 && threadIdx.x == 0) {
    d.counters[CNT_NBLOCKS0 + (parity ^ 1)] = 0;
    // per-step counters of the step's G2P and closure (both done: the grid
    // update follows the scatter, which follows both)
    d.stats[ST_PLASTIC_LAST] = d.stats[ST_PLASTIC_CUR];
    d.stats[ST_PLASTIC] += d.stats[ST_PLASTIC_CUR];
    d.stats[ST_PLASTIC_CUR] = 0;
    d.stats[ST_FALLBACK_LAST] = d.stats[ST_FALLBACK_CUR];
    d.stats[ST_FALLBACK] += d.stats[ST_FALLBACK_CUR];
    d.stats[ST_FALLBACK_CUR] = 0;
    d.stats[ST_LEARNED_LAST] = d.stats[ST_LEARNED_CUR];
    d.stats[ST_LEARNED] += d.stats[ST_LEARNED_CUR];
    d.stats[ST_LEARNED_CUR] = 0;
  }
  const int nblk = d.counters[CNT_NBLOCKS0 + parity];
  const bool scattered = d.counters[CNT_PBLOCKS] > d.tile_cap;   // P2G used its atomic path
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = threadIdx.x & 63;
  for (int e = wave; e < nblk; e += nwaves) {
    const int nb = d.nb_list[parity * d.nb_list_cap + e];
    const int s = d.blk_scene[nb];
    const SceneDev& sc = d.scenes[s];
    const KnifeStateDev& ks = d.kstate[(size_t)step * d.n_scenes + s];
    int bi, bj, bk;
    block_coords(sc, nb, bi, bj, bk);
    const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
    const int gi = bi * 4 + li, gj = bj * 4 + lj, gk = bk * 4 + lk;
    float* acc = d.acc + (size_t)nb * ACC_CH * BLK3;
    // this node's sums: atomically scattered part (blocks past tile_cap) plus,
    // in a fixed order, the merged tiles of the up to 8 particle blocks whose
    // stencils reach it (the block itself and its lower neighbours).  When
    // every listed block had a tile, nothing was scattered: the accumulators
    // are still zero and are neither read nor cleared.
    float a[ACC_CH];
#pragma unroll
    for (int c = 0; c < ACC_CH; ++c) a[c] = scattered ? acc[c * BLK3 + lane] : 0.0f;
    // the 8 source tiles' list positions, resolved up front (independent
    // loads) so that the tile reads below do not wait on each other's lookups
    int te[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ox = q >> 2, oy = (q >> 1) & 1, oz = q & 1;
      const int pi = bi - ox, pj = bj - oy, pk = bk - oz;
      const int key = sc.node_block_base + (pi * sc.bdims[1] + pj) * sc.bdims[2] + pk;
      const bool in = pi >= 0 && pj >= 0 && pk >= 0;
      te[q] = in ? d.list_pos[key] : d.tile_cap;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ox = q >> 2, oy = (q >> 1) & 1, oz = q & 1;
      const int e = te[q];
      if (e >= d.tile_cap) continue;
      const int tx = li + 4 * ox, ty = lj + 4 * oy, tz = lk + 4 * oz;
      if (tx >= PT || ty >= PT || tz >= PT) continue;
      const int t = tile_pos(tx, ty, tz);   // this block's group: one contiguous run
#pragma unroll
      for (int c = 0; c < ACC_CH; ++c) a[c] += d.tile_acc[((size_t)e * ACC_CH + c) * PT3 + t];
    }
    float out[GV_STRIDE];
#pragma unroll
    for (int c = 0; c < GV_STRIDE; ++c) out[c] = 0.0f;
    const float mw = a[A_MASS];   // sum of weights (= m_i / m_p, p2g.cpp:94)
    if (mw != 0.0f && gi < sc.dims[0] && gj < sc.dims[1] && gk < sc.dims[2]) {
      float vold[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) vold[c] = a[A_MOM + c] / mw;
      float pg[NPIG];
#pragma unroll
      for (int c = 0; c < NPIG; ++c) pg[c] = a[A_PIG + c] / mw;
      const float shear = a[A_SHR] / mw;
      const float ps = ((pg[0] + pg[1]) + pg[2]) + pg[3];
      if (ps > 1e-6f) { pg[0] /= ps; pg[1] /= ps; pg[2] /= ps; pg[3] /= ps; }
      const float mass = mw * sc.particle_mass;
      const float dtpv = sc.dt * sc.particle_volume;
      float vn[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float f = a[A_FRC + c];
        vn[c] = vold[c] + (((-dtpv) * f) / mass + sc.dt * sc.gravity[c]);
      }
      const float* xs = d.axis_coords + sc.axis_off;
      const float X[3] = {xs[gi], xs[sc.dims[0] + gj], xs[sc.dims[0] + sc.dims[1] + gk]};
      const uint32_t err = boundary_collision(sc, ks, X, vn);
      if (err) atomicOr(&d.err[0], err);
#pragma unroll
      for (int c = 0; c < 3; ++c) { out[G_V + c] = vn[c]; out[G_FLIP + c] = vn[c] - vold[c]; }
#pragma unroll
      for (int c = 0; c < NPIG; ++c) out[G_PIG + c] = pg[c];
      out[G_SHR] = shear;
      out[G_MASS] = mass;
    }
    const uint64_t act = __ballot(mw != 0.0f);
    if (lane == 0 && act) atomicAdd(&active_wg, (unsigned)__popcll(act));
    float4* g = reinterpret_cast<float4*>(d.gv + ((size_t)nb * BLK3 + lane) * GV_STRIDE);
    g[0] = make_float4(out[0], out[1], out[2], out[3]);
    g[1] = make_float4(out[4], out[5], out[6], out[7]);
    g[2] = make_float4(out[8], out[9], out[10], out[11]);
    g[3] = make_float4(out[12], out[13], out[14], out[15]);
    if (scattered) {
#pragma unroll
      for (int c = 0; c < ACC_CH; ++c) acc[c * BLK3 + lane] = 0.0f;
    }
    if (!REDO && lane == 0) d.nb_flag[nb] = 0;
  }
  __syncthreads();
  if (!REDO && threadIdx.x == 0 && active_wg) atomicAdd(&d.stats[ST_ACTIVE_NODES], (unsigned long long)active_wg);
}

// ------------------------------------------------------------------ k_grid_musl
// MUSL grid velocity (musl.cpp:90-96): the second scatter's momentum (merged
// tiles of the listed blocks in a fixed order + the atomic remainder), divided
// by m_i / m_p with m_i the step's node mass; no boundary condition.  Runs over
// the node blocks the grid update of this step processed.
__global__ void __launch_bounds__(256) k_grid_musl(DevPtrs d) {
  const int parity = d.counters[CNT_GSTEP] & 1;
  const int nblk = d.counters[CNT_NBLOCKS0 + parity];
  const bool scattered = d.counters[CNT_PBLOCKS] > d.tile_cap;   // the second scatter used atomics
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = threadIdx.x & 63;
  for (int e = wave; e < nblk; e += nwaves) {
    const int nb = d.nb_list[parity * d.nb_list_cap + e];
    const SceneDev& sc = d.scenes[d.blk_scene[nb]];
    int bi, bj, bk;
    block_coords(sc, nb, bi, bj, bk);
    const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
    float* acc = d.acc + (size_t)nb * ACC_CH * BLK3;
    float mom[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) mom[c] = scattered ? acc[(A_MOM + c) * BLK3 + lane] : 0.0f;
    for (int q = 0; q < 8; ++q) {
      const int ox = q >> 2, oy = (q >> 1) & 1, oz = q & 1;
      const int pi = bi - ox, pj = bj - oy, pk = bk - oz;
      if (pi < 0 || pj < 0 || pk < 0) continue;
      const int key = sc.node_block_base + (pi * sc.bdims[1] + pj) * sc.bdims[2] + pk;
      const int te = d.list_pos[key];
      if (te >= d.tile_cap) continue;
      const int tx = li + 4 * ox, ty = lj + 4 * oy, tz = lk + 4 * oz;
      if (tx >= PT || ty >= PT || tz >= PT) continue;
      const int t = tile_pos(tx, ty, tz);
#pragma unroll
      for (int c = 0; c < 3; ++c) mom[c] += d.tile_acc[((size_t)te * ACC_CH + A_MOM + c) * PT3 + t];
    }
    float* g = d.gv + ((size_t)nb * BLK3 + lane) * GV_STRIDE;
    const float mi = g[G_MASS];
    if (mi > 0.0f) {
      const float r = mi / sc.particle_mass;
#pragma unroll
      for (int c = 0; c < 3; ++c) g[G_V + c] = mom[c] / r;
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) g[G_V + c] = 0.0f;
    }
    if (scattered) {
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[(A_MOM + c) * BLK3 + lane] = 0.0f;
    }
  }
}

// --------------------------------------------------------------- downloads
__global__ void k_unsort(DevPtrs d, int64_t n, float* out, int ncomp_first, int ncomp) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t id = d.pa_id[j];
  for (int c = 0; c < ncomp; ++c) out[id * ncomp + c] = d.pa[(ncomp_first + c) * d.cap + j];
}

// Largest squared speed of the live particles (update_dt.cpp:7-15) as the
// bits of a non-negative float, whose order is that of unsigned integers.
__global__ void k_max_speed(DevPtrs d, int64_t n, unsigned* out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float v2 = 0.0f;
  if (j < n && d.cell_out[j] >= 0) {
    const float a = d.pa[(P_V + 0) * d.cap + j], b = d.pa[(P_V + 1) * d.cap + j], c = d.pa[(P_V + 2) * d.cap + j];
    v2 = (a * a + b * b) + c * c;
    if (!(v2 >= 0.0f)) v2 = 0.0f;   // NaN: not a speed
  }
  unsigned m = __float_as_uint(v2);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// Active flag of the particle in row slot j: the sort's record of its new
// cell (-1 = dead), written for that same slot by the key pass that follows
// every state update (k_keys at prepare, G2P each step).  The rows are in the
// order of the sort before it, so the new sort's dead range does not apply.
__global__ void k_active_out(DevPtrs d, int64_t n, uint8_t* out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[d.pa_id[j]] = d.cell_out[j] >= 0 ? 1 : 0;
}

__global__ void k_pack_render(DevPtrs d, int64_t n, float* pos4, float* lat7, float* diff) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t id = d.pa_id[j];
  pos4[4 * id + 0] = d.pa[(P_X + 0) * d.cap + j];
  pos4[4 * id + 1] = d.pa[(P_X + 1) * d.cap + j];
  pos4[4 * id + 2] = d.pa[(P_X + 2) * d.cap + j];
  pos4[4 * id + 3] = 1.0f;
  for (int c = 0; c < NPIG; ++c) lat7[7 * id + c] = d.pa[(P_PIG + c) * d.cap + j];
  diff[id] = d.pa[P_DIFF * d.cap + j];
}

// Dense grid export in Simulation::ind order (simulation.hpp:236-238).
__global__ void k_grid_dense(DevPtrs d, int s, float* mass, float* v, float* flip, float* pig, float* shear) {
  const SceneDev& sc = d.scenes[s];
  const int64_t n = (int64_t)sc.dims[0] * sc.dims[1] * sc.dims[2];
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(l / ((int64_t)sc.dims[1] * sc.dims[2]));
    const int j = (int)((l / sc.dims[2]) % sc.dims[1]);
    const int k = (int)(l % sc.dims[2]);
    const int nb = node_block_of(sc, i, j, k);
    const float* g = d.gv + ((size_t)nb * BLK3 + node_local(i, j, k)) * GV_STRIDE;
    const bool live = d.nb_live[nb] != 0;
    if (mass) mass[l] = live ? g[G_MASS] : 0.0f;
    for (int c = 0; c < 3; ++c) {
      if (v) v[3 * l + c] = live ? g[G_V + c] : 0.0f;
      if (flip) flip[3 * l + c] = live ? g[G_FLIP + c] : 0.0f;
    }
    for (int c = 0; c < NPIG; ++c)
      if (pig) pig[7 * l + c] = live ? g[G_PIG + c] : 0.0f;
    if (shear) shear[l] = live ? g[G_SHR] : 0.0f;
  }
}

// boundary_collision on given nodes (positions X, velocities v) with one
// scene's plates and knife state: the grid kernel's per-node rule, evaluated
// for checks of the device code.
__global__ void k_boundary_eval(const SceneDev* __restrict__ sc, const KnifeStateDev* __restrict__ ks,
                                const float* __restrict__ X, const float* __restrict__ v, float* __restrict__ out,
                                int64_t n, uint32_t* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float Xi[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
  float vi[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  const uint32_t e = boundary_collision(*sc, *ks, Xi, vi);
  if (e) atomicOr(err, e);
  out[3 * i] = vi[0];
  out[3 * i + 1] = vi[1];
  out[3 * i + 2] = vi[2];
}

// marks (or clears) the node blocks of the last grid update
__global__ void k_mark_live(DevPtrs d, int clear) {
  const int parity = d.counters[CNT_GSTEP] & 1;
  const int nblk = d.counters[CNT_NBLOCKS0 + parity];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nblk; e += gridDim.x * blockDim.x)
    d.nb_live[d.nb_list[parity * d.nb_list_cap + e]] = clear ? 0 : 1;
}

// ------------------------------------------------------------------ launchers
static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_init_tau(const DevPtrs& d, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_init_tau, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n);
}
void launch_keys(const DevPtrs& d, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n);
}
void launch_scan(const DevPtrs& d, int nkeys, int advance, hipStream_t st) {
  const int nchunks = (nkeys + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nchunks), dim3(SCAN_THREADS), 0, st, d, nkeys);
  hipLaunchKernelGGL(k_scan_down, dim3(nchunks), dim3(SCAN_THREADS), 0, st, d, nkeys, nchunks, advance);
}
int scan_chunks(int nkeys) { return (nkeys + SCAN_CHUNK - 1) / SCAN_CHUNK; }
void launch_perm(const DevPtrs& d, int64_t n, int grid, int src_is_b, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_perm, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n);
  hipLaunchKernelGGL(k_cellsort, dim3(grid * 4), dim3(256), 0, st, d, (const float*)(src_is_b ? d.pb : d.pa),
                     (const int*)(src_is_b ? d.pb_id : d.pa_id));
}
void launch_g2p(const DevPtrs& d, int grid, int learned, int phase, hipStream_t st) {
  const SceneDev* sc = d.scenes;
  const float* ax = d.axis_coords;
  const float* pa = d.pa;
  const int* pm = d.perm;
  const float* gv = d.gv;
  const int4* pd = d.pdesc;
  // the learned closure never runs with volume tracking (refused at create)
  if (phase == 1)
    hipLaunchKernelGGL((k_g2p<false, 1, false>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else if (phase == 2 && learned)
    hipLaunchKernelGGL((k_g2p<true, 2, false>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else if (phase == 2 && d.track_vol)
    hipLaunchKernelGGL((k_g2p<false, 2, true>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else if (phase == 2)
    hipLaunchKernelGGL((k_g2p<false, 2, false>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else if (learned)
    hipLaunchKernelGGL((k_g2p<true, 0, false>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else if (d.track_vol)
    hipLaunchKernelGGL((k_g2p<false, 0, true>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
  else
    hipLaunchKernelGGL((k_g2p<false, 0, false>), dim3(grid * 2), dim3(128), 0, st, d, sc, ax, pa, pm, gv, pd);
}
void launch_closure(const DevPtrs& d, const void* mlp_blob, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_closure, dim3(grid), dim3(256), 0, st, d, d.scenes, (const uint8_t*)mlp_blob);
}
void launch_closure_eval(const SceneDev* scenes, const void* mlp_blob, const float* trial4, int64_t n,
                         float* learned4, float* exact4, uint8_t* accepted, int max_grid, hipStream_t st) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, max_grid);
  hipLaunchKernelGGL(k_closure_eval, dim3((unsigned)blocks), dim3(256), 0, st, scenes, (const uint8_t*)mlp_blob,
                     trial4, n, learned4, exact4, accepted);
}
// Workgroups of k_g2p (two wavefronts each) resident per CU.
int g2p_resident_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_g2p<false, 0, false>, 128, 0) != hipSuccess || n <= 0)
    return 8;
  return n;
}
// Workgroups of k_p2g (one wavefront each) resident per CU: the P2G launch size.
int p2g_resident_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_p2g<false>, 64, 0) != hipSuccess || n <= 0) return 12;
  return n;
}
void launch_p2g(const DevPtrs& d, int grid, int src_is_b, hipStream_t st) {
  hipLaunchKernelGGL(k_p2g<false>, dim3(grid), dim3(64), 0, st, d, (const SceneDev*)d.scenes,
                     (const float*)d.axis_coords, (const float*)(src_is_b ? d.pb : d.pa), (const int*)d.perm,
                     (const int4*)d.pdesc, (const int*)d.cell_num, (const float*)nullptr);
}
void launch_musl(const DevPtrs& d, int grid_p2g, int grid_grid, hipStream_t st) {
  hipLaunchKernelGGL(k_p2g<true>, dim3(grid_p2g), dim3(64), 0, st, d, (const SceneDev*)d.scenes,
                     (const float*)d.axis_coords, (const float*)d.pa, (const int*)d.perm, (const int4*)d.pdesc,
                     (const int*)d.cell_num, (const float*)d.pb);
  hipLaunchKernelGGL(k_grid_musl, dim3(grid_grid), dim3(256), 0, st, d);
}
void launch_grid(const DevPtrs& d, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_grid<false>, dim3(grid), dim3(256), 0, st, d);
}
void launch_grid_redo(const DevPtrs& d, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_grid<true>, dim3(grid), dim3(256), 0, st, d);
}
void launch_max_speed(const DevPtrs& d, int64_t n, unsigned* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_max_speed, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n, out);
}
void launch_unsort(const DevPtrs& d, int64_t n, float* out, int first, int ncomp, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unsort, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n, out, first, ncomp);
}
void launch_active_out(const DevPtrs& d, int64_t n, uint8_t* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_active_out, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n, out);
}
void launch_pack_render(const DevPtrs& d, int64_t n, float* pos4, float* lat7, float* diff, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_pack_render, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, n, pos4, lat7, diff);
}
void launch_boundary_eval(const SceneDev* sc, const KnifeStateDev* ks, const float* X, const float* v, float* out,
                          int64_t n, uint32_t* err, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_boundary_eval, dim3(blocks_for(n, 256)), dim3(256), 0, st, sc, ks, X, v, out, n, err);
}
void launch_grid_dense(const DevPtrs& d, int s, float* mass, float* v, float* flip, float* pig, float* shear,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_mark_live, dim3(256), dim3(256), 0, st, d, 0);
  hipLaunchKernelGGL(k_grid_dense, dim3(1024), dim3(256), 0, st, d, s, mass, v, flip, pig, shear);
  hipLaunchKernelGGL(k_mark_live, dim3(256), dim3(256), 0, st, d, 1);
}

}  // namespace mpmx
