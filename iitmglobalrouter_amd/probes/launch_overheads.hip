// Launch-structure costs on one gfx950 GPU, measured with empty kernels.
//
// Not part of the routing library: it backs the choice of one hipGraph of
// dependent kernels per batch (costing, trees, refresh) over a single
// persistent launch whose stages are separated by grid-wide barriers.
// Prints, in microseconds:
//   1. a dependent kernel node in a replayed graph (1 and 8192 workgroups);
//   2. a fork to a second stream and the join back, per a -> (b || c) -> d
//      group of four empty kernels, captured and built node by node;
//   3. the same chain and fork/join launched eagerly;
//   4. one software grid barrier (an atomic arrival counter and a
//      generation word) inside a cooperative launch, whose workgroups are
//      all resident by construction.  The spin is bounded: on overrun the
//      kernel sets a flag and every workgroup leaves.
#include <hip/hip_runtime.h>
#include <cstdio>

#define CK(x)                                                                \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                              \
    }                                                                        \
  } while (0)

// Never true (the buffer is zeroed): keeps the kernels from being empty
// bodies while writing nothing.
__global__ void k_noop(int* p) {
  if (threadIdx.x == 0 && p[0] == -1) p[1 + blockIdx.x] = 1;
}

__device__ unsigned int g_arrived;
__device__ unsigned int g_generation;
__device__ int g_overrun;

// Returns false once any workgroup's wait has overrun its bound.
__device__ bool grid_barrier(unsigned int nblocks, unsigned int& gen) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int next = gen + 1;
    if (atomicAdd(&g_arrived, 1u) == nblocks - 1) {
      atomicExch(&g_arrived, 0u);
      __threadfence();
      atomicExch(&g_generation, next);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&g_generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != next) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1l << 26)) {
          atomicExch(&g_overrun, 1);
          break;
        }
      }
    }
    __threadfence();
    gen = next;
    s_ok = __hip_atomic_load(&g_overrun, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  }
  __syncthreads();
  return s_ok != 0;
}

__global__ void k_barriers(int iters) {
  unsigned int gen = 0;  // used by thread 0 only
  if (threadIdx.x == 0) gen = __hip_atomic_load(&g_generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int i = 0; i < iters; ++i)
    if (!grid_barrier(gridDim.x, gen)) break;
}

static float elapsed_us(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms * 1000.f;
}

int main() {
  int* d = nullptr;
  CK(hipMalloc(&d, 1 << 20));
  CK(hipMemset(d, 0, 1 << 20));
  hipStream_t s1, s2;
  CK(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
  CK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  hipEvent_t t0, t1, fork, join;
  CK(hipEventCreate(&t0));
  CK(hipEventCreate(&t1));
  CK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  const int reps = 10;

  // 1. dependent kernel nodes in one replayed graph
  for (int blocks : {1, 8192}) {
    const int N = 200;
    hipGraph_t g;
    hipGraphExec_t ge;
    CK(hipStreamBeginCapture(s1, hipStreamCaptureModeGlobal));
    for (int i = 0; i < N; ++i) hipLaunchKernelGGL(k_noop, dim3(blocks), dim3(256), 0, s1, d);
    CK(hipStreamEndCapture(s1, &g));
    CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    CK(hipGraphLaunch(ge, s1));
    CK(hipStreamSynchronize(s1));
    CK(hipEventRecord(t0, s1));
    for (int r = 0; r < reps; ++r) CK(hipGraphLaunch(ge, s1));
    CK(hipEventRecord(t1, s1));
    CK(hipEventSynchronize(t1));
    std::printf("graph chain, %5d workgroups per kernel: %6.2f us per node\n", blocks, elapsed_us(t0, t1) / (reps * N));
    CK(hipGraphExecDestroy(ge));
    CK(hipGraphDestroy(g));
  }

  // 2. a -> (b || c) -> d groups, captured and built node by node
  const int G = 50;
  for (int built : {0, 1}) {
    hipGraph_t g;
    hipGraphExec_t ge;
    if (!built) {
      CK(hipStreamBeginCapture(s1, hipStreamCaptureModeGlobal));
      for (int i = 0; i < G; ++i) {
        hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
        CK(hipEventRecord(fork, s1));
        CK(hipStreamWaitEvent(s2, fork, 0));
        hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
        hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s2, d);
        CK(hipEventRecord(join, s2));
        CK(hipStreamWaitEvent(s1, join, 0));
        hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
      }
      CK(hipStreamEndCapture(s1, &g));
    } else {
      CK(hipGraphCreate(&g, 0));
      void* args[] = {(void*)&d};
      hipKernelNodeParams kp = {};
      kp.func = (void*)k_noop;
      kp.gridDim = dim3(1);
      kp.blockDim = dim3(64);
      kp.kernelParams = args;
      hipGraphNode_t prev = nullptr;
      for (int i = 0; i < G; ++i) {
        hipGraphNode_t a, b, c, e;
        CK(hipGraphAddKernelNode(&a, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
        CK(hipGraphAddKernelNode(&b, g, &a, 1, &kp));
        CK(hipGraphAddKernelNode(&c, g, &a, 1, &kp));
        hipGraphNode_t deps[2] = {b, c};
        CK(hipGraphAddKernelNode(&e, g, deps, 2, &kp));
        prev = e;
      }
    }
    CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    CK(hipGraphLaunch(ge, s1));
    CK(hipStreamSynchronize(s1));
    CK(hipEventRecord(t0, s1));
    for (int r = 0; r < reps; ++r) CK(hipGraphLaunch(ge, s1));
    CK(hipEventRecord(t1, s1));
    CK(hipEventSynchronize(t1));
    std::printf("graph fork/join (%s): %6.2f us per group of 4 kernels\n", built ? "built" : "captured",
                elapsed_us(t0, t1) / (reps * G));
    CK(hipGraphExecDestroy(ge));
    CK(hipGraphDestroy(g));
  }

  // 3. eager launches
  {
    const int N = 200;
    CK(hipStreamSynchronize(s1));
    CK(hipEventRecord(t0, s1));
    for (int i = 0; i < N; ++i) hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
    CK(hipEventRecord(t1, s1));
    CK(hipEventSynchronize(t1));
    std::printf("eager chain: %6.2f us per kernel\n", elapsed_us(t0, t1) / N);
    CK(hipEventRecord(t0, s1));
    for (int i = 0; i < G; ++i) {
      hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
      CK(hipEventRecord(fork, s1));
      CK(hipStreamWaitEvent(s2, fork, 0));
      hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
      hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s2, d);
      CK(hipEventRecord(join, s2));
      CK(hipStreamWaitEvent(s1, join, 0));
      hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, s1, d);
    }
    CK(hipEventRecord(t1, s1));
    CK(hipEventSynchronize(t1));
    std::printf("eager fork/join: %6.2f us per group of 4 kernels\n", elapsed_us(t0, t1) / G);
  }

  // 4. software grid barrier in a cooperative launch
  int ncu = 0;
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  for (int threads : {256, 1024}) {
    int per_cu = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_barriers, threads, 0));
    for (int bpc : {1, 2}) {
      if (bpc > per_cu) continue;
      const int nb = ncu * bpc;
      int iters = 2000;
      int zero = 0;
      CK(hipMemcpyToSymbol(HIP_SYMBOL(g_overrun), &zero, sizeof(int)));
      void* args[] = {(void*)&iters};
      CK(hipEventRecord(t0, s1));
      CK(hipLaunchCooperativeKernel((const void*)k_barriers, dim3(nb), dim3(threads), args, 0, s1));
      CK(hipEventRecord(t1, s1));
      CK(hipEventSynchronize(t1));
      int overrun = 0;
      CK(hipMemcpyFromSymbol(&overrun, HIP_SYMBOL(g_overrun), sizeof(int)));
      std::printf("grid barrier, %d workgroups x %d threads (%d CUs): %6.2f us per barrier%s\n", nb, threads, ncu,
                  elapsed_us(t0, t1) / iters, overrun ? " (spin bound reached)" : "");
    }
  }
  CK(hipFree(d));
  return 0;
}
