// phase_barrier — what a phase boundary of the step costs in the two forms a
// step can take on MI355X: a grid-wide barrier inside one persistent
// (cooperatively launched) kernel, or a kernel boundary inside a captured HIP
// graph (the form the library uses: one graph launch runs two whole steps).
//
//   phase_barrier [workgroups per CU = 4]
//
// Both forms run the same grid (workgroups of 256 threads, the given count per
// CU, capped by the kernel's occupancy so that a cooperative launch is
// possible).  Grid barrier (HIP's cooperative-groups barrier, and a direct
// one: an arrive atomic and a polled generation word): one launch with N =
// 400 barriers minus one with none, divided by N.  Kernel boundary: a graph of
// N = 400 launches of an (almost) empty kernel, replayed, divided by N.
// Prints one JSON line.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace cg = cooperative_groups;

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
      return 1;                                                              \
    }                                                                        \
  } while (0)

// n grid-wide barriers; each workgroup adds its count to *out at the end
__global__ void __launch_bounds__(256) k_barriers(int n, int* out) {
  cg::grid_group g = cg::this_grid();
  for (int i = 0; i < n; ++i) g.sync();
  if (threadIdx.x == 0) atomicAdd(out, n);
}

// n grid-wide barriers written directly: thread 0 of each workgroup arrives
// with one agent-scope atomic; the last to arrive resets the count and
// advances the generation word, which the others poll.  The polling has one
// budget for the whole launch: past it the workgroup sets *timed_out and
// stops waiting (the grid is co-resident by the cooperative launch, so the
// budget is never reached in a correct run).
__global__ void __launch_bounds__(256) k_spin_barriers(int n, unsigned* arrive, unsigned* gen, int* out,
                                                      unsigned* timed_out) {
  __shared__ int stop;
  if (threadIdx.x == 0) stop = 0;
  unsigned budget = 1u << 22;
  for (int i = 0; i < n; ++i) {
    __syncthreads();
    if (stop) break;
    if (threadIdx.x == 0) {
      const unsigned g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned a = __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (a == gridDim.x - 1) {
        __hip_atomic_store(arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(gen, g0 + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g0) {
          if (--budget == 0u) {
            atomicOr(timed_out, 1u);
            stop = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, n);
}

// one phase of nothing: the cost of a kernel boundary alone
__global__ void __launch_bounds__(256) k_phase(int* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(out, 1);
}

int main(int argc, char** argv) {
  const int per_cu_req = argc > 1 ? std::atoi(argv[1]) : 4;
  if (per_cu_req < 1 || per_cu_req > 16) {
    std::fprintf(stderr, "workgroups per CU must be 1..16\n");
    return 2;
  }
  constexpr int N = 400, REPS = 5;
  int dev = 0, n_cu = 0, coop = 0, occ = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_barriers, 256, 0));
  if (!coop || occ < 1) {
    std::fprintf(stderr, "no cooperative launch on this device (attribute %d, occupancy %d)\n", coop, occ);
    return 1;
  }
  const int per_cu = per_cu_req < occ ? per_cu_req : occ;
  const int blocks = per_cu * n_cu;
  int* out = nullptr;
  unsigned* words = nullptr;   // arrive count, generation, timed-out flag
  CHECK(hipMalloc(&out, sizeof(int)));
  CHECK(hipMemset(out, 0, sizeof(int)));
  CHECK(hipMalloc(&words, 3 * sizeof(unsigned)));
  CHECK(hipMemset(words, 0, 3 * sizeof(unsigned)));
  hipStream_t st;
  CHECK(hipStreamCreate(&st));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));

  // grid barriers: best of REPS launches with N barriers and with none
  unsigned* arrive = words;
  unsigned* gen = words + 1;
  unsigned* timed_out = words + 2;
  auto coop_ms = [&](const void* kernel, bool spin, int n, float* best) -> int {
    *best = 1e30f;
    for (int r = 0; r < REPS + 1; ++r) {
      void* args_cg[] = {&n, &out};
      void* args_spin[] = {&n, &arrive, &gen, &out, &timed_out};
      CHECK(hipEventRecord(e0, st));
      CHECK(hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(256), spin ? args_spin : args_cg, 0, st));
      CHECK(hipEventRecord(e1, st));
      CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0 && ms < *best) *best = ms;   // the first launch is a warm-up
    }
    return 0;
  };
  float ms_n = 0.0f, ms_0 = 0.0f, ms_sn = 0.0f, ms_s0 = 0.0f;
  if (coop_ms((const void*)k_barriers, false, N, &ms_n) || coop_ms((const void*)k_barriers, false, 0, &ms_0))
    return 1;
  if (coop_ms((const void*)k_spin_barriers, true, N, &ms_sn) ||
      coop_ms((const void*)k_spin_barriers, true, 0, &ms_s0))
    return 1;
  unsigned host_words[3] = {0, 0, 0};
  CHECK(hipMemcpy(host_words, words, sizeof host_words, hipMemcpyDeviceToHost));

  // kernel boundaries: a captured graph of N launches
  hipGraph_t graph;
  hipGraphExec_t exec;
  CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
  for (int i = 0; i < N; ++i) hipLaunchKernelGGL(k_phase, dim3(blocks), dim3(256), 0, st, out);
  CHECK(hipStreamEndCapture(st, &graph));
  CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  float ms_g = 1e30f;
  for (int r = 0; r < REPS + 1; ++r) {
    CHECK(hipEventRecord(e0, st));
    CHECK(hipGraphLaunch(exec, st));
    CHECK(hipEventRecord(e1, st));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (r > 0 && ms < ms_g) ms_g = ms;
  }
  int host_out = 0;
  CHECK(hipMemcpy(&host_out, out, sizeof(int), hipMemcpyDeviceToHost));
  const long long expect = 2LL * (REPS + 1) * blocks * N + (long long)(REPS + 1) * N;
  const bool ok = host_out == expect && host_words[2] == 0u;
  std::printf("{\"cus\": %d, \"workgroups\": %d, \"threads\": 256, \"grid_barrier_us\": %.3f, "
              "\"spin_barrier_us\": %.3f, \"graph_kernel_boundary_us\": %.3f, \"check\": %s}\n",
              n_cu, blocks, (ms_n - ms_0) * 1e3 / N, (ms_sn - ms_s0) * 1e3 / N, ms_g * 1e3 / N,
              ok ? "true" : "false");
  CHECK(hipGraphExecDestroy(exec));
  CHECK(hipGraphDestroy(graph));
  CHECK(hipFree(out));
  CHECK(hipFree(words));
  return ok ? 0 : 1;
}
