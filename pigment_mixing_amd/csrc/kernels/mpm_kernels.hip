// mpm_kernels.hip — hand-written gfx950 kernels of the MPM step.
//
// Step pipeline (one HIP graph per step, all scenes of a batch at once):
//   k_fused   : per particle block — G2P gather from the updated grid (g2p.cpp:12-135),
//               F update + VMVisc return map (deformation_update.cpp:7-52,
//               plasticity.cpp:163-238), pigment mixing, advection and the
//               particle guards (position_update.cpp:6-91), then the NEXT step's
//               P2G scatter of mass, APIC momentum, Kirchhoff force, pigment latent
//               and shear intensity (p2g.cpp:6-126, explicit_euler_update.cpp:12-55)
//               into an LDS tile flushed to the grid accumulators.  Particles are
//               written back physically sorted; their new bin keys are counted.
//   k_scan_*  : prefix sums over key counts -> offsets + non-empty block list.
//   k_perm    : permutation slot <- particle for the next k_fused.
//   k_grid    : per active node block — normalise, gravity + force, boundary
//               collision with plates and knife (explicit_euler_update.cpp:64-110,
//               boundary_collision.cpp:5-316), write the gather channels, clear
//               the accumulators.
// k_keys / k_p2g_init start the pipeline after an upload.
#include "../common/mpmx_device.hpp"
#include "mpm_kernels.hpp"

namespace mpmx {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint64_t lanemask_lt() {
  return (__lane_id() == 0) ? 0ull : ((~0ull) >> (64 - __lane_id()));
}

// Rank of this lane's key among equal keys, using one atomic per distinct key
// in the wavefront.  Only lanes with `valid` participate.
__device__ __forceinline__ int wave_rank(bool valid, int key, int* counts) {
  uint64_t todo = __ballot(valid);
  int rank = 0;
  while (todo) {
    int leader = __ffsll((long long)todo) - 1;
    int k0 = __shfl(key, leader);
    uint64_t same = __ballot(valid && key == k0);
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(&counts[k0], (int)__popcll(same));
    base = __shfl(base, leader);
    if (valid && key == k0) rank = base + (int)__popcll(same & lanemask_lt());
    todo &= ~same;
  }
  return rank;
}

__device__ __forceinline__ void activate_node_block(const DevPtrs& d, int nb, int parity) {
  if (atomicCAS(&d.nb_flag[nb], 0, 1) == 0) {
    int idx = atomicAdd(&d.counters[CNT_NBLOCKS0 + parity], 1);
    atomicAdd(&d.stats[ST_NBLOCKS], 1ull);
    d.nb_list[parity * d.nb_list_cap + idx] = nb;
  }
}

// Particle key: node block of the stencil base, or DEAD.
__device__ __forceinline__ int particle_key(const SceneDev& sc, const float x[3], int dead_key) {
  int b0 = stencil_base(x[0], sc.origin[0], sc.one_over_dx);
  int b1 = stencil_base(x[1], sc.origin[1], sc.one_over_dx);
  int b2 = stencil_base(x[2], sc.origin[2], sc.one_over_dx);
  if (b0 < 0 || b1 < 0 || b2 < 0 || b0 + 2 > sc.dims[0] - 1 || b1 + 2 > sc.dims[1] - 1 ||
      b2 + 2 > sc.dims[2] - 1)
    return dead_key;
  return node_block_of(sc, b0, b1, b2);
}

// ------------------------------------------------------------ P2G scatter
// Tile = nodes [B0-1, B0+7) per axis around particle block B (origin corner B0).
constexpr int TILE = 8;
constexpr int TILE3 = TILE * TILE * TILE;

struct P2GPayload {
  float x[3], v[3], C[9], tau[6], pig[NPIG], diff;
};

struct TileLDS {
  float acc[ACC_CH][TILE3];
  unsigned long long mass[TILE3];
  unsigned int touched;
};

// Scatter one particle (p2g.cpp:20-77 + explicit_euler_update.cpp:12-55).
__device__ __forceinline__ void scatter_particle(const DevPtrs& d, const SceneDev& sc, const P2GPayload& p,
                                                 TileLDS* lds, int t0i, int t0j, int t0k, int parity) {
  int b[3];
  AxisW w[3];
  const float* xs = d.axis_coords + sc.axis_off;
  const float* ys = xs + sc.dims[0];
  const float* zs = ys + sc.dims[1];
  b[0] = stencil_base(p.x[0], sc.origin[0], sc.one_over_dx);
  b[1] = stencil_base(p.x[1], sc.origin[1], sc.one_over_dx);
  b[2] = stencil_base(p.x[2], sc.origin[2], sc.one_over_dx);
  axis_weights(p.x[0], xs, b[0], sc.one_over_dx, w[0]);
  axis_weights(p.x[1], ys, b[1], sc.one_over_dx, w[1]);
  axis_weights(p.x[2], zs, b[2], sc.one_over_dx, w[2]);
  const int li = b[0] - t0i, lj = b[1] - t0j, lk = b[2] - t0k;
  const bool in_tile = lds != nullptr && li >= 0 && lj >= 0 && lk >= 0 && li + 2 < TILE && lj + 2 < TILE &&
                       lk + 2 < TILE;
  if (!in_tile) atomicAdd(&d.stats[ST_STRAY], 1ull);
  const float inv_dx = sc.one_over_dx;
#pragma unroll
  for (int oi = 0; oi < 3; ++oi) {
    const float xi = xs[b[0] + oi];
#pragma unroll
    for (int oj = 0; oj < 3; ++oj) {
      const float yi = ys[b[1] + oj];
#pragma unroll
      for (int ok = 0; ok < 3; ++ok) {
        const float zi = zs[b[2] + ok];
        const float wgt = (w[0].N[oi] * w[1].N[oj]) * w[2].N[ok];
        const float g0 = ((w[0].dN[oi] * w[1].N[oj]) * w[2].N[ok]) * inv_dx;
        const float g1 = ((w[0].N[oi] * w[1].dN[oj]) * w[2].N[ok]) * inv_dx;
        const float g2 = ((w[0].N[oi] * w[1].N[oj]) * w[2].dN[ok]) * inv_dx;
        if (!(wgt > 1e-25f)) continue;
        const float pd0 = xi - p.x[0], pd1 = yi - p.x[1], pd2 = zi - p.x[2];
        float val[ACC_CH];
        // APIC momentum (p2g.cpp:84-93)
        val[0] = p.v[0] * wgt + ((p.C[0] * pd0 + p.C[1] * pd1 + p.C[2] * pd2) * sc.apicDinv) * wgt;
        val[1] = p.v[1] * wgt + ((p.C[3] * pd0 + p.C[4] * pd1 + p.C[5] * pd2) * sc.apicDinv) * wgt;
        val[2] = p.v[2] * wgt + ((p.C[6] * pd0 + p.C[7] * pd1 + p.C[8] * pd2) * sc.apicDinv) * wgt;
        // force tau * grad w (explicit_euler_update.cpp:20-38), tau symmetric
        val[3] = p.tau[0] * g0 + p.tau[3] * g1 + p.tau[4] * g2;
        val[4] = p.tau[3] * g0 + p.tau[1] * g1 + p.tau[5] * g2;
        val[5] = p.tau[4] * g0 + p.tau[5] * g1 + p.tau[2] * g2;
#pragma unroll
        for (int c = 0; c < NPIG; ++c) val[6 + c] = p.pig[c] * wgt;
        val[13] = p.diff * wgt;
        const unsigned long long mfx = (unsigned long long)__float2uint_rn(wgt * MASS_FX_SCALE_F);
        if (in_tile) {
          const int t = ((li + oi) * TILE + (lj + oj)) * TILE + (lk + ok);
#pragma unroll
          for (int c = 0; c < ACC_CH; ++c) atomicAdd(&lds->acc[c][t], val[c]);
          atomicAdd(&lds->mass[t], mfx);
        } else {
          const int gi = b[0] + oi, gj = b[1] + oj, gk = b[2] + ok;
          const int nb = node_block_of(sc, gi, gj, gk);
          const int loc = node_local(gi, gj, gk);
#pragma unroll
          for (int c = 0; c < ACC_CH; ++c) atomicAdd(&d.acc[((size_t)nb * ACC_CH + c) * BLK3 + loc], val[c]);
          atomicAdd(&d.massfx[(size_t)nb * BLK3 + loc], mfx);
          activate_node_block(d, nb, parity);
        }
      }
    }
  }
}

__device__ void tile_zero(TileLDS* lds) {
  float* f = &lds->acc[0][0];
  for (int i = threadIdx.x; i < ACC_CH * TILE3; i += blockDim.x) f[i] = 0.0f;
  for (int i = threadIdx.x; i < TILE3; i += blockDim.x) lds->mass[i] = 0ull;
  if (threadIdx.x == 0) lds->touched = 0u;
}

// Flush the LDS tile into the grid accumulators (global atomics, only nodes
// that received mass) and activate the touched node blocks.
__device__ void tile_flush(const DevPtrs& d, const SceneDev& sc, TileLDS* lds, int t0i, int t0j, int t0k,
                           int parity) {
  const int bi0 = (t0i + 1) >> 2, bj0 = (t0j + 1) >> 2, bk0 = (t0k + 1) >> 2;
  for (int t = threadIdx.x; t < TILE3; t += blockDim.x) {
    const unsigned long long m = lds->mass[t];
    if (m == 0ull) continue;
    const int ti = t / (TILE * TILE), tj = (t / TILE) % TILE, tk = t % TILE;
    const int gi = t0i + ti, gj = t0j + tj, gk = t0k + tk;
    const int nb = node_block_of(sc, gi, gj, gk);
    const int loc = node_local(gi, gj, gk);
#pragma unroll
    for (int c = 0; c < ACC_CH; ++c) atomicAdd(&d.acc[((size_t)nb * ACC_CH + c) * BLK3 + loc], lds->acc[c][t]);
    atomicAdd(&d.massfx[(size_t)nb * BLK3 + loc], m);
    const int rb = (((gi >> 2) - bi0 + 1) * 3 + ((gj >> 2) - bj0 + 1)) * 3 + ((gk >> 2) - bk0 + 1);
    atomicOr(&lds->touched, 1u << rb);
  }
  __syncthreads();
  if (threadIdx.x < 27 && (lds->touched & (1u << threadIdx.x))) {
    const int r = threadIdx.x;
    const int bi = bi0 + r / 9 - 1, bj = bj0 + (r / 3) % 3 - 1, bk = bk0 + r % 3 - 1;
    const int nb = sc.node_block_base + (bi * sc.bdims[1] + bj) * sc.bdims[2] + bk;
    activate_node_block(d, nb, parity);
  }
}

// ------------------------------------------------------------------ k_keys
// Bin every particle (upload order) and count keys.
__global__ void k_keys(DevPtrs d, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = i < n;
  int key = d.dead_key;
  if (valid) {
    const int s = d.pscene[i];
    const SceneDev& sc = d.scenes[s];
    float x[3] = {d.pa[0 * d.cap + i], d.pa[1 * d.cap + i], d.pa[2 * d.cap + i]};
    if (d.pa_id[i] >= 0 && d.active_in[i]) key = particle_key(sc, x, d.dead_key);
    if (key == d.dead_key) {
      // inactive from the start: v = 0 (position_update.cpp:222-228 convention)
      d.pa[3 * d.cap + i] = 0.0f; d.pa[4 * d.cap + i] = 0.0f; d.pa[5 * d.cap + i] = 0.0f;
    }
    d.key_out[i] = key;
  }
  int r = wave_rank(valid, key, d.count_next);
  if (valid) d.rank_out[i] = r;
}

// ------------------------------------------------------------------ scans
// counts_next[0..nkeys] (last = DEAD) -> offsets (exclusive), count_cur,
// non-empty live block list (ascending key order), totals.  Two phases over
// chunks of SCAN_CHUNK keys; the 64-bit scanned value packs (count<<32 | nonempty).
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ unsigned long long scan_item(const DevPtrs& d, int k, int nkeys) {
  if (k >= nkeys) return 0ull;
  unsigned long long c = (unsigned long long)d.count_next[k];
  bool live = k < nkeys - 1;
  return (c << 32) | (unsigned long long)((c > 0 && live) ? 1 : 0);
}

__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long* sh, unsigned long long& total) {
  // Hillis-Steele over 256 threads in LDS (deterministic)
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    unsigned long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0ull;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  total = sh[SCAN_THREADS - 1];
  unsigned long long incl = sh[threadIdx.x];
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
  // prefix of previous chunks
  unsigned long long pre = 0;
  for (int c = threadIdx.x; c < blockIdx.x; c += SCAN_THREADS) pre += d.scan_partials[c];
  unsigned long long tot;
  unsigned long long dummy = block_exclusive_scan(pre, sh, tot);
  (void)dummy;
  const unsigned long long chunk_base = tot;
  const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
  unsigned long long items[SCAN_PER_THREAD];
  unsigned long long s = 0;
  for (int j = 0; j < SCAN_PER_THREAD; ++j) { items[j] = scan_item(d, base + j, nkeys); s += items[j]; }
  unsigned long long ctot;
  unsigned long long run = chunk_base + block_exclusive_scan(s, sh, ctot);
  for (int j = 0; j < SCAN_PER_THREAD; ++j) {
    const int k = base + j;
    if (k < nkeys) {
      const int cnt = (int)(items[j] >> 32);
      d.offsets[k] = (int)(run >> 32);
      d.count_cur[k] = cnt;
      if (items[j] & 1ull) d.pb_list[(int)(run & 0xffffffffull)] = k;
      d.count_next[k] = 0;
    }
    run += items[j];
  }
  __syncthreads();
  if (blockIdx.x == nchunks - 1 && threadIdx.x == SCAN_THREADS - 1) {
    // run now holds the grand total
    d.counters[CNT_PBLOCKS] = (int)(run & 0xffffffffull);
    d.counters[CNT_LIVE] = d.offsets[nkeys - 1];
    if (advance) {
      d.counters[CNT_STEP] += 1;
      d.counters[CNT_GSTEP] += 1;
    }
  }
}

// ------------------------------------------------------------------ k_perm
__global__ void k_perm(DevPtrs d, int64_t n) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int key = d.key_out[j];
  d.perm[d.offsets[key] + d.rank_out[j]] = (int)j;
}

// ------------------------------------------------------------- particle I/O
struct PState {
  float x[3], v[3], F[9], pig[NPIG], eps;
  int id;
};

__device__ __forceinline__ void load_state(const DevPtrs& d, const float* buf, const int* ids, int64_t i, PState& s) {
  const int64_t cap = d.cap;
#pragma unroll
  for (int c = 0; c < 3; ++c) { s.x[c] = buf[(P_X + c) * cap + i]; s.v[c] = buf[(P_V + c) * cap + i]; }
#pragma unroll
  for (int c = 0; c < 9; ++c) s.F[c] = buf[(P_F + c) * cap + i];
#pragma unroll
  for (int c = 0; c < NPIG; ++c) s.pig[c] = buf[(P_PIG + c) * cap + i];
  s.eps = buf[P_EPS * cap + i];
  s.id = ids[i];
}

// --------------------------------------------------------------- k_p2g_init
// First scatter after an upload: stress from F by SVD (HenckyPiola).
__global__ void __launch_bounds__(256) k_p2g_init(DevPtrs d) {
  __shared__ TileLDS lds;
  const int nblk = d.counters[CNT_PBLOCKS];
  const int parity = d.counters[CNT_GSTEP] & 1;
  for (int e = blockIdx.x; e < nblk; e += gridDim.x) {
    const int key = d.pb_list[e];
    const int s = d.blk_scene[key];
    const SceneDev& sc = d.scenes[s];
    const int lb = key - sc.node_block_base;
    const int bi = lb / (sc.bdims[1] * sc.bdims[2]), bj = (lb / sc.bdims[2]) % sc.bdims[1], bk = lb % sc.bdims[2];
    const int t0i = bi * 4 - 1, t0j = bj * 4 - 1, t0k = bk * 4 - 1;
    tile_zero(&lds);
    __syncthreads();
    const int start = d.offsets[key], cnt = d.count_cur[key];
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int src = d.perm[start + t];
      const float* A = d.pa;
      const int64_t cap = d.cap;
      P2GPayload p;
#pragma unroll
      for (int c = 0; c < 3; ++c) { p.x[c] = A[(P_X + c) * cap + src]; p.v[c] = A[(P_V + c) * cap + src]; }
#pragma unroll
      for (int c = 0; c < 9; ++c) p.C[c] = A[(P_C + c) * cap + src];
#pragma unroll
      for (int c = 0; c < NPIG; ++c) p.pig[c] = A[(P_PIG + c) * cap + src];
      p.diff = A[P_DIFF * cap + src];
      float F[9], U[9], sig[3], V[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) F[c] = A[(P_F + c) * cap + src];
      svd3(F, U, sig, V);
      float h[3] = {logf(fabsf(sig[0])), logf(fabsf(sig[1])), logf(fabsf(sig[2]))};
      kirchhoff_from_principal(sc, U, h, p.tau);
      scatter_particle(d, sc, p, &lds, t0i, t0j, t0k, parity);
    }
    __syncthreads();
    tile_flush(d, sc, &lds, t0i, t0j, t0k, parity);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ k_grid
// One wavefront per active node block; lane = node.
__global__ void __launch_bounds__(256) k_grid(DevPtrs d) {
  const int step = d.counters[CNT_STEP];
  const int parity = d.counters[CNT_GSTEP] & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) d.counters[CNT_NBLOCKS0 + (parity ^ 1)] = 0;
  const int nblk = d.counters[CNT_NBLOCKS0 + parity];
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = threadIdx.x & 63;
  for (int e = wave; e < nblk; e += nwaves) {
    const int nb = d.nb_list[parity * d.nb_list_cap + e];
    const int s = d.blk_scene[nb];
    const SceneDev& sc = d.scenes[s];
    const KnifeStateDev& ks = d.kstate[(size_t)step * d.n_scenes + s];
    const int lb = nb - sc.node_block_base;
    const int bi = lb / (sc.bdims[1] * sc.bdims[2]), bj = (lb / sc.bdims[2]) % sc.bdims[1], bk = lb % sc.bdims[2];
    const int gi = bi * 4 + (lane >> 4), gj = bj * 4 + ((lane >> 2) & 3), gk = bk * 4 + (lane & 3);
    float* acc = d.acc + (size_t)nb * ACC_CH * BLK3;
    float* gv = d.gv + (size_t)nb * GV_CH * BLK3;
    const unsigned long long mfx = d.massfx[(size_t)nb * BLK3 + lane];
    float out[GV_CH];
#pragma unroll
    for (int c = 0; c < GV_CH; ++c) out[c] = 0.0f;
    if (mfx != 0ull && gi < sc.dims[0] && gj < sc.dims[1] && gk < sc.dims[2]) {
      const float mw = (float)((double)mfx * MASS_FX_INV);  // sum of weights (= m_i / m_p)
      float vold[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) vold[c] = acc[(A_MOM + c) * BLK3 + lane] / mw;
      float pg[NPIG];
#pragma unroll
      for (int c = 0; c < NPIG; ++c) pg[c] = acc[(A_PIG + c) * BLK3 + lane] / mw;
      const float shear = acc[A_SHR * BLK3 + lane] / mw;
      const float ps = ((pg[0] + pg[1]) + pg[2]) + pg[3];
      if (ps > 1e-6f) { pg[0] /= ps; pg[1] /= ps; pg[2] /= ps; pg[3] /= ps; }
      const float mass = mw * sc.particle_mass;
      const float dtpv = sc.dt * sc.particle_volume;
      float vn[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float f = acc[(A_FRC + c) * BLK3 + lane];
        vn[c] = vold[c] + (((-dtpv) * f) / mass + sc.dt * sc.gravity[c]);
      }
      const float* xs = d.axis_coords + sc.axis_off;
      const float X[3] = {xs[gi], xs[sc.dims[0] + gj], xs[sc.dims[0] + sc.dims[1] + gk]};
      uint32_t err = boundary_collision(sc, ks, X, vn);
      if (err) atomicOr(&d.err[0], err);
#pragma unroll
      for (int c = 0; c < 3; ++c) { out[G_V + c] = vn[c]; out[G_FLIP + c] = vn[c] - vold[c]; }
#pragma unroll
      for (int c = 0; c < NPIG; ++c) out[G_PIG + c] = pg[c];
      out[G_SHR] = shear;
      out[G_MASS] = mass;
    }
    const uint64_t act = __ballot(mfx != 0ull);
    if (lane == 0 && act) atomicAdd(&d.stats[ST_ACTIVE_NODES], (unsigned long long)__popcll(act));
#pragma unroll
    for (int c = 0; c < GV_CH; ++c) gv[c * BLK3 + lane] = out[c];
#pragma unroll
    for (int c = 0; c < ACC_CH; ++c) acc[c * BLK3 + lane] = 0.0f;
    d.massfx[(size_t)nb * BLK3 + lane] = 0ull;
    if (lane == 0) d.nb_flag[nb] = 0;
  }
}

// ------------------------------------------------------------------ k_fused
__device__ __forceinline__ void gather_node(const DevPtrs& d, const SceneDev& sc, int gi, int gj, int gk,
                                            float out[GV_CH - 1]) {
  const int nb = node_block_of(sc, gi, gj, gk);
  const int loc = node_local(gi, gj, gk);
  const float* g = d.gv + (size_t)nb * GV_CH * BLK3 + loc;
#pragma unroll
  for (int c = 0; c < GV_CH - 1; ++c) out[c] = g[c * BLK3];
}

__global__ void __launch_bounds__(256) k_fused(DevPtrs d) {
  __shared__ TileLDS lds;
  const int step = d.counters[CNT_STEP];
  const int parity_next = (d.counters[CNT_GSTEP] + 1) & 1;
  const int nblk = d.counters[CNT_PBLOCKS];
  const float* A = d.pa;
  float* Bf = d.pb;
  const int64_t cap = d.cap;
  for (int e = blockIdx.x; e < nblk; e += gridDim.x) {
    const int key = d.pb_list[e];
    const int s = d.blk_scene[key];
    const SceneDev& sc = d.scenes[s];
    const float time = d.kstate[(size_t)step * d.n_scenes + s].time;
    const int lb = key - sc.node_block_base;
    const int bi = lb / (sc.bdims[1] * sc.bdims[2]), bj = (lb / sc.bdims[2]) % sc.bdims[1], bk = lb % sc.bdims[2];
    const int t0i = bi * 4 - 1, t0j = bj * 4 - 1, t0k = bk * 4 - 1;
    const float* xs = d.axis_coords + sc.axis_off;
    const float* ys = xs + sc.dims[0];
    const float* zs = ys + sc.dims[1];
    // pigment boost (g2p.cpp:107-119)
    float tf = 1.0f;
    if (time > sc.start_boost && sc.end_boost > sc.start_boost) {
      float t = (time - sc.start_boost) / (sc.end_boost - sc.start_boost);
      t = fminf(fmaxf(t, 0.0f), 1.0f);
      tf = 1.0f + t * (sc.boost_factor - 1.0f);
    }
    const float dynD = sc.D_max * tf;
    tile_zero(&lds);
    __syncthreads();
    const int start = d.offsets[key], cnt = d.count_cur[key];
    const int iters = (cnt + blockDim.x - 1) / blockDim.x;
    for (int it = 0; it < iters; ++it) {
      const int t = it * blockDim.x + threadIdx.x;
      const bool valid = t < cnt;
      const int64_t j = start + t;
      int newkey = d.dead_key;
      bool alive = false;
      int plastic = 0;
      P2GPayload p;
      if (valid) {
        const int src = d.perm[j];
        PState st;
        load_state(d, A, d.pa_id, src, st);
        // ---------------- G2P gather (g2p.cpp:25-91), 27-node stencil
        int b[3];
        AxisW w[3];
        b[0] = stencil_base(st.x[0], sc.origin[0], sc.one_over_dx);
        b[1] = stencil_base(st.x[1], sc.origin[1], sc.one_over_dx);
        b[2] = stencil_base(st.x[2], sc.origin[2], sc.one_over_dx);
        axis_weights(st.x[0], xs, b[0], sc.one_over_dx, w[0]);
        axis_weights(st.x[1], ys, b[1], sc.one_over_dx, w[1]);
        axis_weights(st.x[2], zs, b[2], sc.one_over_dx, w[2]);
        float pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, Bp[9], L[9], gp[NPIG], gs = 0.0f;
#pragma unroll
        for (int c = 0; c < 9; ++c) { Bp[c] = 0.0f; L[c] = 0.0f; }
#pragma unroll
        for (int c = 0; c < NPIG; ++c) gp[c] = 0.0f;
        const float inv_dx = sc.one_over_dx;
#pragma unroll
        for (int oi = 0; oi < 3; ++oi) {
          const float pd0 = xs[b[0] + oi] - st.x[0];
#pragma unroll
          for (int oj = 0; oj < 3; ++oj) {
            const float pd1 = ys[b[1] + oj] - st.x[1];
#pragma unroll
            for (int ok = 0; ok < 3; ++ok) {
              const float pd2 = zs[b[2] + ok] - st.x[2];
              const float wgt = (w[0].N[oi] * w[1].N[oj]) * w[2].N[ok];
              const float g0 = ((w[0].dN[oi] * w[1].N[oj]) * w[2].N[ok]) * inv_dx;
              const float g1 = ((w[0].N[oi] * w[1].dN[oj]) * w[2].N[ok]) * inv_dx;
              const float g2 = ((w[0].N[oi] * w[1].N[oj]) * w[2].dN[ok]) * inv_dx;
              float gn[GV_CH - 1];
              gather_node(d, sc, b[0] + oi, b[1] + oj, b[2] + ok, gn);
#pragma unroll
              for (int r = 0; r < 3; ++r) {
                pic[r] += gn[G_V + r] * wgt;
                Bp[3 * r + 0] += (gn[G_V + r] * pd0) * wgt;
                Bp[3 * r + 1] += (gn[G_V + r] * pd1) * wgt;
                Bp[3 * r + 2] += (gn[G_V + r] * pd2) * wgt;
                fl[r] += gn[G_FLIP + r] * wgt;
                L[3 * r + 0] += gn[G_V + r] * g0;
                L[3 * r + 1] += gn[G_V + r] * g1;
                L[3 * r + 2] += gn[G_V + r] * g2;
              }
#pragma unroll
              for (int c = 0; c < NPIG; ++c) gp[c] += gn[G_PIG + c] * wgt;
              gs += gn[G_SHR] * wgt;
            }
          }
        }
        if (sc.flip_ratio >= 0.0f) {
#pragma unroll
          for (int c = 0; c < 9; ++c) Bp[c] = 0.0f;  // PIC/FLIP: Bmat unused
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
        float tt = fmaxf(0.0f, fminf(1.0f, (gs - sc.D_edge0) / (sc.D_edge1 - sc.D_edge0)));
        const float mixf = tt * tt * (3.0f - 2.0f * tt);
        const float smf = fminf(fmaxf(dynD * mixf * sc.dt, 0.0f), 1.0f);
#pragma unroll
        for (int c = 0; c < NPIG; ++c) st.pig[c] = (1.0f - smf) * st.pig[c] + smf * gp[c];
        const float ps = ((st.pig[0] + st.pig[1]) + st.pig[2]) + st.pig[3];
        if (ps > 1e-6f) { st.pig[0] /= ps; st.pig[1] /= ps; st.pig[2] /= ps; st.pig[3] /= ps; }
        // ---------------- F update + return map (deformation_update.cpp:33-45)
        float Ft[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            Ft[3 * r + c] = st.F[3 * r + c] + ((sc.dt * L[3 * r + 0]) * st.F[0 + c] + (sc.dt * L[3 * r + 1]) * st.F[3 + c] +
                                               (sc.dt * L[3 * r + 2]) * st.F[6 + c]);
        float U[9], sig[3], V[9];
        svd3(Ft, U, sig, V);
        float h[3] = {logf(fmaxf(fabsf(sig[0]), 1e-4f)), logf(fmaxf(fabsf(sig[1]), 1e-4f)),
                      logf(fmaxf(fabsf(sig[2]), 1e-4f))};
        float eps = st.eps, dgr = 0.0f;
        uint32_t err = 0;
        float h_trial[3] = {h[0], h[1], h[2]};
        if (sc.plastic_model == 5) plastic = vmvisc_principal(sc, h, eps, dgr, err);
        if (plastic && d.harvest) {
          // record: h_trial[3], ||dev h_trial||, q_y, eta = q_y*perzyna_visc, dt*E, nu | h[3], dgamma
          const unsigned long long slot = atomicAdd(&d.stats[ST_HARVEST], 1ull);
          if ((int64_t)slot < d.harvest_cap) {
            float* r = d.harvest + 12 * slot;
            const float m = ((h_trial[0] + h_trial[1]) + h_trial[2]) / 3.0f;
            const float e0 = h_trial[0] - m, e1 = h_trial[1] - m, e2 = h_trial[2] - m;
            const float qy = sc.q_min + (sc.q_max - sc.q_min) * expf(-sc.xi * st.eps);
            r[0] = h_trial[0]; r[1] = h_trial[1]; r[2] = h_trial[2];
            r[3] = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
            r[4] = qy;
            r[5] = qy * sc.perzyna_visc;
            r[6] = sc.dt * sc.E;
            r[7] = sc.nu;
            r[8] = h[0]; r[9] = h[1]; r[10] = h[2];
            r[11] = dgr * sc.dt;
          }
        }
        if (err) atomicOr(&d.err[0], err);
        float Fn[9];
        if (plastic) {
          const float e0 = expf(h[0]), e1 = expf(h[1]), e2 = expf(h[2]);
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              Fn[3 * r + c] = (U[3 * r + 0] * e0) * V[3 * c + 0] + (U[3 * r + 1] * e1) * V[3 * c + 1] +
                              (U[3 * r + 2] * e2) * V[3 * c + 2];
        } else {
#pragma unroll
          for (int c = 0; c < 9; ++c) Fn[c] = Ft[c];
        }
        // ---------------- position update (position_update.cpp:8-90)
        float xn[3], vn[3];
        const float a = -sc.flip_ratio;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          xn[c] = st.x[c] + sc.dt * pic[c];
          if (sc.use_musl) vn[c] = st.v[c];
          else if (sc.flip_ratio < -1.0f) vn[c] = pic[c];
          else if (sc.flip_ratio < 0.0f) vn[c] = a * (st.v[c] + fl[c]) + (1.0f - a) * pic[c];
          else vn[c] = sc.flip_ratio * (st.v[c] + fl[c]) + (1.0f - sc.flip_ratio) * pic[c];
        }
        alive = true;
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
            newkey = particle_key(sc, xn, d.dead_key);
            if (newkey == d.dead_key) alive = false;
          }
          if (!alive) { vn[0] = vn[1] = vn[2] = 0.0f; }
        }
        // ---------------- write new state, sorted slot j
#pragma unroll
        for (int c = 0; c < 3; ++c) { Bf[(P_X + c) * cap + j] = xn[c]; Bf[(P_V + c) * cap + j] = vn[c]; }
#pragma unroll
        for (int c = 0; c < 9; ++c) { Bf[(P_C + c) * cap + j] = Bp[c]; Bf[(P_F + c) * cap + j] = Fn[c]; }
#pragma unroll
        for (int c = 0; c < NPIG; ++c) Bf[(P_PIG + c) * cap + j] = st.pig[c];
        Bf[P_DIFF * cap + j] = diff_new;
        Bf[P_EPS * cap + j] = eps;
        Bf[P_DG * cap + j] = dgr;
        d.pb_id[j] = st.id;
        if (alive) {
#pragma unroll
          for (int c = 0; c < 3; ++c) { p.x[c] = xn[c]; p.v[c] = vn[c]; }
#pragma unroll
          for (int c = 0; c < 9; ++c) p.C[c] = Bp[c];
#pragma unroll
          for (int c = 0; c < NPIG; ++c) p.pig[c] = st.pig[c];
          p.diff = diff_new;
          kirchhoff_from_principal(sc, U, h, p.tau);
        }
        d.key_out[j] = newkey;
      }
      const int r = wave_rank(valid, newkey, d.count_next);
      if (valid) d.rank_out[j] = r;
      const uint64_t pm = __ballot(valid && plastic);
      if (__lane_id() == 0 && pm) atomicAdd(&d.stats[ST_PLASTIC], (unsigned long long)__popcll(pm));
      if (valid && alive) scatter_particle(d, sc, p, &lds, t0i, t0j, t0k, parity_next);
    }
    __syncthreads();
    tile_flush(d, sc, &lds, t0i, t0j, t0k, parity_next);
    __syncthreads();
  }
}

// Dead particles keep their state: copy them to the output side.
__global__ void k_copy_dead(DevPtrs d) {
  const int dead_start = d.offsets[d.dead_key];
  const int ndead = d.count_cur[d.dead_key];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ndead; t += gridDim.x * blockDim.x) {
    const int64_t j = dead_start + t;
    const int src = d.perm[j];
    for (int c = 0; c < P_NCOMP; ++c) d.pb[c * d.cap + j] = d.pa[c * d.cap + src];
    d.pb_id[j] = d.pa_id[src];
    d.key_out[j] = d.dead_key;
  }
  // ranks of dead particles are assigned in a second pass (k_rank_dead)
}

__global__ void k_rank_dead(DevPtrs d) {
  const int dead_start = d.offsets[d.dead_key];
  const int ndead = d.count_cur[d.dead_key];
  for (int t0 = blockIdx.x * blockDim.x; t0 < ndead; t0 += gridDim.x * blockDim.x) {
    const int t = t0 + threadIdx.x;
    const bool valid = t < ndead;
    const int r = wave_rank(valid, d.dead_key, d.count_next);
    if (valid) d.rank_out[dead_start + t] = r;
  }
}

// --------------------------------------------------------------- downloads
// Sorted slot -> original id order, and the renderer format.
__global__ void k_unsort(DevPtrs d, int64_t n, float* out, int ncomp_first, int ncomp) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int id = d.pa_id[j];
  for (int c = 0; c < ncomp; ++c) out[(int64_t)id * ncomp + c] = d.pa[(ncomp_first + c) * d.cap + j];
}

__global__ void k_active_out(DevPtrs d, int64_t n, uint8_t* out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[d.pa_id[j]] = (j < d.offsets[d.dead_key]) ? 1 : 0;
}

__global__ void k_pack_render(DevPtrs d, int64_t n, float* pos4, float* lat7, float* diff) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
    const float* g = d.gv + (size_t)nb * GV_CH * BLK3 + node_local(i, j, k);
    const bool live = d.nb_live[nb] != 0;
    if (mass) mass[l] = live ? g[G_MASS * BLK3] : 0.0f;
    for (int c = 0; c < 3; ++c) {
      if (v) v[3 * l + c] = live ? g[(G_V + c) * BLK3] : 0.0f;
      if (flip) flip[3 * l + c] = live ? g[(G_FLIP + c) * BLK3] : 0.0f;
    }
    for (int c = 0; c < NPIG; ++c) if (pig) pig[7 * l + c] = live ? g[(G_PIG + c) * BLK3] : 0.0f;
    if (shear) shear[l] = live ? g[G_SHR * BLK3] : 0.0f;
  }
}

// marks the node blocks of the last grid update (for k_grid_dense)
__global__ void k_mark_live(DevPtrs d, int clear) {
  const int parity = d.counters[CNT_GSTEP] & 1;
  const int nblk = d.counters[CNT_NBLOCKS0 + parity];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nblk; e += gridDim.x * blockDim.x)
    d.nb_live[d.nb_list[parity * d.nb_list_cap + e]] = clear ? 0 : 1;
}

// ------------------------------------------------------------------ launchers
void launch_keys(const DevPtrs& d, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, n);
}
void launch_scan(const DevPtrs& d, int nkeys, int advance, hipStream_t st) {
  const int nchunks = (nkeys + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nchunks), dim3(SCAN_THREADS), 0, st, d, nkeys);
  hipLaunchKernelGGL(k_scan_down, dim3(nchunks), dim3(SCAN_THREADS), 0, st, d, nkeys, nchunks, advance);
}
int scan_chunks(int nkeys) { return (nkeys + SCAN_CHUNK - 1) / SCAN_CHUNK; }
void launch_perm(const DevPtrs& d, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_perm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, n);
}
void launch_p2g_init(const DevPtrs& d, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_p2g_init, dim3(grid), dim3(256), 0, st, d);
}
void launch_grid(const DevPtrs& d, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_grid, dim3(grid), dim3(256), 0, st, d);
}
void launch_fused(const DevPtrs& d, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_fused, dim3(grid), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_copy_dead, dim3(64), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_rank_dead, dim3(64), dim3(256), 0, st, d);
}
void launch_unsort(const DevPtrs& d, int64_t n, float* out, int first, int ncomp, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_unsort, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, n, out, first, ncomp);
}
void launch_active_out(const DevPtrs& d, int64_t n, uint8_t* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_active_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, n, out);
}
void launch_pack_render(const DevPtrs& d, int64_t n, float* pos4, float* lat7, float* diff, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pack_render, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, n, pos4, lat7, diff);
}
void launch_grid_dense(const DevPtrs& d, int s, float* mass, float* v, float* flip, float* pig, float* shear,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_mark_live, dim3(256), dim3(256), 0, st, d, 0);
  hipLaunchKernelGGL(k_grid_dense, dim3(1024), dim3(256), 0, st, d, s, mass, v, flip, pig, shear);
  hipLaunchKernelGGL(k_mark_live, dim3(256), dim3(256), 0, st, d, 1);
}

}  // namespace mpmx
