// Vector-ALU fused multiply-add throughput probe (measurement only, not part
// of the collision library's C ABI).
//
// Each thread runs CHAINS independent FMA chains for `iters` rounds of UNROLL
// steps, so the issue rate, not the FMA latency, bounds the loop.  The result
// is the sustained FMA rate of the vector ALU at the given type with the build
// flags the collision library uses (no packed FP32 instructions), i.e. the
// ceiling for kernels whose work is scalar FP32 / FP64 arithmetic.
#include <hip/hip_runtime.h>

namespace {

constexpr int THREADS = 256;
constexpr int CHAINS = 8;
constexpr int UNROLL = 16;

template <typename T>
__global__ void __launch_bounds__(THREADS) k_fma_loop(T* out, int iters, T a, T b) {
  T x[CHAINS];
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) x[c] = (T)(threadIdx.x + c) * (T)1e-3;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) x[c] = fma(x[c], a, b);
    }
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) s += x[c];
  out[(size_t)blockIdx.x * THREADS + threadIdx.x] = s;
}

template <typename T>
int run(long long blocks, int iters, int reps, double* flops, double* ms) {
  T* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * THREADS * sizeof(T)) != hipSuccess) return -5;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  // a < 1 keeps the chains bounded: x -> b / (1 - a)
  const T a = (T)0.999, b = (T)1e-3;
  hipLaunchKernelGGL(k_fma_loop<T>, dim3((unsigned)blocks), dim3(THREADS), 0, 0, out, iters, a, b);
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_fma_loop<T>, dim3((unsigned)blocks), dim3(THREADS), 0, 0, out, iters, a, b);
  hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(out);
  if (err != hipSuccess) return -5;
  *ms = (double)t / reps;
  *flops = 2.0 * CHAINS * UNROLL * (double)iters * (double)THREADS * (double)blocks;
  return 0;
}

}  // namespace

// dtype 0: FP32, 1: FP64.  Writes the FLOPs of one launch and its average time.
extern "C" int ibp_fma_peak(int dtype, long long blocks, int iters, int reps, double* flops,
                            double* ms) {
  if (blocks < 1 || blocks > (1LL << 30) || iters < 1 || reps < 1 || !flops || !ms) return -1;
  return dtype == 0 ? run<float>(blocks, iters, reps, flops, ms)
                    : dtype == 1 ? run<double>(blocks, iters, reps, flops, ms) : -1;
}
