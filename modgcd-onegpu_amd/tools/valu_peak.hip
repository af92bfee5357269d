// valu_peak.hip -- the achievable VALU issue rate of this device, the
// denominator of the gcd kernel's roofline fraction (bench.py prices it at
// 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.32 T lane-ops/s).
//
// Each kernel runs many independent chains per lane (no dependency stalls),
// with enough waves per SIMD to fill every issue slot, over a grid of many
// workgroups per CU, and reports lane-ops per second:
//   fma_f64   v_fma_f64 (the residue, elimination and replay arithmetic);
//   mulmod    the kernel's FP64 mulmod sequence (mul, fma, mul, rndne, fma,
//             add per product), as the elimination issues it;
//   add_u32   v_add_u32 (integer bookkeeping);
//   rcp_f32   v_rcp_f32 (the inverse's quotient estimate; a quarter of the
//             32-bit rate).
// Each kernel also times itself with s_memtime (shader clock) against
// s_memrealtime (100 MHz): the clock the device ran at under that load.
// Results are written to memory, so no chain is optimised away.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

namespace {

constexpr int kChains = 8;
constexpr int kThreads = 256;

// Shader-clock and real-time deltas of workgroup 0, thread 0.
struct Clock {
  unsigned long long mem, real;
};

__device__ __forceinline__ void clock_start(uint64_t& t0, uint64_t& r0) {
  t0 = __builtin_amdgcn_s_memtime();
  r0 = __builtin_amdgcn_s_memrealtime();
}
__device__ __forceinline__ void clock_stop(uint64_t t0, uint64_t r0, Clock* clk) {
  const uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    clk->mem = t1 - t0;
    clk->real = r1 - r0;
  }
}

__global__ __launch_bounds__(kThreads) void fma_f64(int iters, double* out, Clock* clk) {
  uint64_t t0, r0;
  clock_start(t0, r0);
  double a[kChains];
  const double x = 1.0 + 1e-9 * threadIdx.x, y = 1e-12;
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = 1.0 + c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int c = 0; c < kChains; ++c) a[c] = __builtin_fma(a[c], x, y);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  if (s == 12345.678) out[blockIdx.x * kThreads + threadIdx.x] = s;
  clock_stop(t0, r0, clk);
}

// The elimination's product of two exact integers below 2^33 modulo m < 2^32
// (one FP64 two-product and a rounded quotient), chains kept below 2^33.
__device__ __forceinline__ double mulmod(double a, double b, double m, double minv) {
  const double h = a * b;
  const double l = __builtin_fma(a, b, -h);
  const double q = __builtin_rint(h * minv);
  return __builtin_fma(-q, m, h) + l;
}

__global__ __launch_bounds__(kThreads) void mulmod_f64(int iters, double* out, Clock* clk) {
  uint64_t t0, r0;
  clock_start(t0, r0);
  const double m = 4294967291.0 - 2.0 * threadIdx.x, minv = 1.0 / m;
  double a[kChains];
  const double b = 3141592653.0 - threadIdx.x;
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = 12345.0 + 1000.0 * c + blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < kChains; ++c) a[c] = mulmod(a[c], b, m, minv);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  if (s == 0.5) out[blockIdx.x * kThreads + threadIdx.x] = s;
  clock_stop(t0, r0, clk);
}

__global__ __launch_bounds__(kThreads) void add_u32(int iters, double* out, Clock* clk) {
  uint64_t t0, r0;
  clock_start(t0, r0);
  uint32_t a[kChains];
  const uint32_t x = threadIdx.x * 2654435761u + 1u;
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = blockIdx.x + 7u * c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        a[c] += x;
        asm volatile("" : "+v"(a[c]));  // one v_add_u32 per step, not a folded multiply
      }
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s ^= a[c];
  if (s == 0x12345678u) out[blockIdx.x * kThreads + threadIdx.x] = s;
  clock_stop(t0, r0, clk);
}

__global__ __launch_bounds__(kThreads) void rcp_f32(int iters, double* out, Clock* clk) {
  uint64_t t0, r0;
  clock_start(t0, r0);
  float a[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = 1.5f + 0.25f * c + 1e-6f * threadIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int c = 0; c < kChains; ++c) a[c] = __builtin_amdgcn_rcpf(a[c]);
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  if (s == 1234.5f) out[blockIdx.x * kThreads + threadIdx.x] = s;
  clock_stop(t0, r0, clk);
}

using Kernel = void (*)(int, double*, Clock*);

// lanes: lane-ops per clock per SIMD the instruction can issue at (FP64 16,
// 32-bit ALU 32, v_rcp_f32 8); the clock-adjusted fraction divides by
// 256 CUs x 4 SIMD x lanes x the measured shader clock.
int run(const char* name, Kernel k, double ops_per_iter, int lanes, int blocks, int iters, double* out,
        Clock* dclk) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, 0, iters / 8, out, dclk);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  Clock clk{};
  for (int rep = 0; rep < 5; ++rep) {
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, 0, iters, out, dclk);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      CHECK(hipMemcpy(&clk, dclk, sizeof clk, hipMemcpyDeviceToHost));
    }
  }
  const double lane_ops = (double)blocks * kThreads * iters * ops_per_iter;
  const double rate = lane_ops / (best * 1e-3);
  const double ghz = clk.real ? (double)clk.mem / ((double)clk.real * 10.0) : 0.0;
  const double at_clock = 256.0 * 4 * lanes * ghz * 1e9;
  std::printf(
      "{\"kernel\": \"%s\", \"lane_ops\": %.4e, \"ms\": %.3f, \"lane_ops_per_s\": %.4e, \"frac_of_39.32T\": %.4f, "
      "\"shader_clock_ghz\": %.3f, \"lanes_per_clock_per_simd\": %d, \"frac_of_peak_at_that_clock\": %.4f}\n",
      name, lane_ops, best, rate, rate / 39.3216e12, ghz, lanes, at_clock > 0 ? rate / at_clock : 0.0);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  std::printf("{\"device\": \"%s\", \"gcn_arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d}\n", prop.name,
              prop.gcnArchName, prop.multiProcessorCount, prop.clockRate);
  const int blocks = prop.multiProcessorCount * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
  double* out;
  Clock* clk;
  CHECK(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(double)));
  CHECK(hipMalloc(&clk, sizeof(Clock)));
  // ops per lane per outer iteration
  // about 0.2 s per timed launch: long enough for the clock to settle under the load
  if (run("fma_f64", fma_f64, 16.0 * kChains, 16, blocks, 100000, out, clk)) return 1;
  if (run("mulmod_f64", mulmod_f64, 4.0 * kChains * 6, 16, blocks, 60000, out, clk)) return 1;
  if (run("add_u32", add_u32, 16.0 * kChains, 32, blocks, 150000, out, clk)) return 1;
  if (run("rcp_f32", rcp_f32, 16.0 * kChains, 8, blocks, 50000, out, clk)) return 1;
  CHECK(hipFree(out));
  CHECK(hipFree(clk));
  return 0;
}
