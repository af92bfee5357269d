// exchange_latency.hip -- the latency floor of the gcd kernel's cross-workgroup
// exchange on this device: agent-scope relaxed atomics (what the kernel's
// exchange slots use) between workgroups.
//
//   ping-pong:  two workgroups, one 8-byte flag each way per round trip,
//               on the same XCD (workgroups 0 and 8 under round-robin
//               dispatch) and on two XCDs (0 and 1); agent and system scope,
//               with and without s_sleep between polls;
//   all-reduce: G workgroups (16 .. 256) post one tagged value per round into
//               a two-row slot ring and take the minimum: flat (every
//               workgroup polls all G slots), relay (workgroup 0 polls them
//               and posts the minimum into a relay slot), and by XCD (a
//               leader per XCD polls its XCD's slots, every workgroup polls
//               the eight leaders' slots).  Every result is checked.
//   XCD-local:  the same two exchanges through the XCD's L2 instead of past
//               it: stores that write through the CU's L1 to the L2, polls
//               with loads that skip the L1 (sc0) and are served by the L2.
//               Only workgroups of one XCD share that L2, so the ping-pong
//               runs between workgroups 0 and 8 (one XCD) and 0 and 1 (two
//               XCDs: expected to give up), and the all-reduce among the
//               workgroups of XCD 0 only.
// Every wait is bounded (2^22 polls) and reported if it gives up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

namespace {

constexpr uint32_t kSpin = 1u << 22;

// SYS 0: agent scope, 1: system scope, 2: XCD-local (through the L2), 3 and
// 4: XCD-local with the CU's L1 invalidated before each poll (buffer_inv at
// workgroup scope, sc0, or at agent scope, sc1) and a plain load, 5: agent
// scope among the workgroups of one XCD (the comparison for 2-4).
template <int SYS>
__device__ __forceinline__ uint64_t ld(const uint64_t* p) {
  if constexpr (SYS == 2) {
    uint64_t v;  // the load completes inside the asm block
    asm volatile("global_load_dwordx2 %0, %1, off sc0\n\ts_waitcnt vmcnt(0)" : "=v"(v) : "v"(p) : "memory");
    return v;
  } else if constexpr (SYS == 3) {
    uint64_t v;
    asm volatile("buffer_inv sc0\n\tglobal_load_dwordx2 %0, %1, off\n\ts_waitcnt vmcnt(0)" : "=v"(v) : "v"(p) : "memory");
    return v;
  } else if constexpr (SYS == 4) {
    uint64_t v;
    asm volatile("buffer_inv sc1\n\tglobal_load_dwordx2 %0, %1, off\n\ts_waitcnt vmcnt(0)" : "=v"(v) : "v"(p) : "memory");
    return v;
  } else if constexpr (SYS == 1) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  } else if constexpr (SYS == 5) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
template <int SYS>
__device__ __forceinline__ void st(uint64_t* p, uint64_t v) {
  if constexpr (SYS >= 2 && SYS <= 4) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else if constexpr (SYS == 1) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }

// Workgroups a and b (all others return at once) trade flags n times;
// out[0, 1] real-time ticks, out[2, 3] waits that gave up, out[4, 5] XCDs.
template <int SYS, int SLEEP>
__global__ void pingpong(uint64_t* flags, int a, int b, int n, uint64_t* out) {
  if (((int)blockIdx.x != a && (int)blockIdx.x != b) || threadIdx.x != 0) return;
  const bool isA = (int)blockIdx.x == a;
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  uint32_t fails = 0;
  for (int i = 1; i <= n && !fails; ++i) {
    const uint64_t* wait_on = isA ? flags + 16 : flags;
    if (isA) st<SYS>(flags, (uint64_t)i);
    uint32_t s = 0;
    while (ld<SYS>(wait_on) != (uint64_t)i) {
      if (++s > kSpin) {
        fails++;
        break;
      }
      if (SLEEP) __builtin_amdgcn_s_sleep(1);
    }
    if (!isA) st<SYS>(flags + 16, (uint64_t)i);
  }
  const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
  out[isA ? 0 : 1] = t1 - t0;
  out[isA ? 2 : 3] = fails;
  out[isA ? 4 : 5] = xcc_id();
}

__host__ __device__ inline uint64_t value_of(uint64_t wg, uint64_t r) { return (wg * 2654435761ull + r * 40503ull) & 0xFFFFFull; }

// MODE 0 flat, 1 relay, 2 by XCD.  slots [2][G], relay [2], xs [2][8][G], gs [2][8],
// reg [9] counters; out[wg] = xor of the round minima (or ~0 after a give-up),
// out[G + wg] real-time ticks, out[2G + wg] its XCD.
// SCOPE as for ld/st; with SCOPE 2 only every eighth workgroup (one XCD
// under round-robin dispatch) takes part, as workgroup blockIdx / 8.
template <int MODE, int SCOPE = 0>
__global__ void allreduce(uint64_t* slots, uint64_t* relay, uint64_t* xs, uint64_t* gs, uint32_t* reg, int n,
                          uint64_t* out) {
  constexpr uint32_t STRIDE = SCOPE >= 2 ? 8 : 1;
  if (blockIdx.x % STRIDE) return;
  const uint32_t G = gridDim.x / STRIDE, wg = blockIdx.x / STRIDE, lane = threadIdx.x;
  __shared__ uint32_t s_li, s_nx[8], s_fail;
  const uint32_t x = xcc_id() & 7;
  if (lane == 0) {
    s_fail = 0;
    if (MODE == 2) {
      s_li = __hip_atomic_fetch_add(reg + x, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(reg + 8, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t s = 0;
      while (__hip_atomic_load(reg + 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < G) {
        if (++s > kSpin) {
          s_fail = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      for (int i = 0; i < 8; ++i) s_nx[i] = __hip_atomic_load(reg + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (s_fail) {
    if (lane == 0) out[wg] = ~0ull;
    return;
  }
  uint64_t acc = 0;
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  bool fail = false;
  for (uint32_t r = 1; r <= (uint32_t)n && !fail; ++r) {
    const uint64_t tag = (uint64_t)r << 32;
    const uint32_t row = r & 1;
    uint64_t m = ~0ull;
    auto wait_value = [&](const uint64_t* p) -> uint64_t {
      uint64_t v = ld<SCOPE>(p);
      uint32_t s = 0;
      while ((v >> 32) != r) {
        if (++s > kSpin) {
          fail = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ld<SCOPE>(p);
      }
      return v & 0xFFFFFFFFull;
    };
    if (MODE == 0 || MODE == 1) {
      if (lane == 0) st<SCOPE>(slots + row * G + wg, tag | value_of(wg, r));
      if (MODE == 0 || wg == 0) {
        for (uint32_t j = lane; j < G; j += 64) m = std::min(m, wait_value(slots + row * G + j));
        for (int o = 32; o >= 1; o >>= 1) m = std::min(m, (uint64_t)__shfl_xor((unsigned long long)m, o));
        if (MODE == 1 && lane == 0) st<SCOPE>(relay + row, tag | m);
      } else {
        m = wait_value(relay + row);
      }
    } else {
      if (lane == 0) st<0>(xs + (row * 8 + x) * G + s_li, tag | value_of(wg, r));
      if (s_li == 0) {
        for (uint32_t j = lane; j < s_nx[x]; j += 64) m = std::min(m, wait_value(xs + (row * 8 + x) * G + j));
        for (int o = 32; o >= 1; o >>= 1) m = std::min(m, (uint64_t)__shfl_xor((unsigned long long)m, o));
        if (lane == 0) st<0>(gs + row * 8 + x, tag | m);
      }
      m = ~0ull;
      if (lane < 8 && s_nx[lane] > 0) m = wait_value(gs + row * 8 + lane);
      for (int o = 32; o >= 1; o >>= 1) m = std::min(m, (uint64_t)__shfl_xor((unsigned long long)m, o));
    }
    fail = __any(fail);
    acc ^= m + r;
  }
  const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
  if (lane == 0) {
    out[wg] = fail ? ~0ull : acc;
    out[G + wg] = t1 - t0;
    out[2 * G + wg] = x;
  }
}

}  // namespace

int main() {
  constexpr int G = 256;
  uint64_t *flags, *out, *slots, *relay, *xs, *gs;
  uint32_t* reg;
  CHECK(hipMalloc(&flags, 4096));
  CHECK(hipMalloc(&out, 3 * G * 8));
  CHECK(hipMalloc(&slots, 2 * G * 8));
  CHECK(hipMalloc(&relay, 2 * 8));
  CHECK(hipMalloc(&xs, 2 * 8 * G * 8));
  CHECK(hipMalloc(&gs, 2 * 8 * 8));
  CHECK(hipMalloc(&reg, 64));
  std::vector<uint64_t> h(3 * G);
  const int n = 20000;
  auto pp = [&](auto kern, const char* name, int a, int b) -> int {
    CHECK(hipMemset(flags, 0, 4096));
    CHECK(hipMemset(out, 0, 3 * G * 8));
    hipLaunchKernelGGL(kern, dim3(16), dim3(64), 0, 0, flags, a, b, n, out);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(h.data(), out, 6 * 8, hipMemcpyDeviceToHost));
    std::printf("ping-pong %-20s workgroups %d (XCD %llu) and %d (XCD %llu): %7.1f ns per round trip, gave up %llu/%llu\n",
                name, a, (unsigned long long)h[4], b, (unsigned long long)h[5], h[0] * 10.0 / n,
                (unsigned long long)h[2], (unsigned long long)h[3]);
    return 0;
  };
  for (int b : {8, 1}) {
    if (pp(pingpong<0, 1>, "agent, s_sleep", 0, b)) return 1;
    if (pp(pingpong<0, 0>, "agent, no sleep", 0, b)) return 1;
    if (pp(pingpong<1, 0>, "system, no sleep", 0, b)) return 1;
    if (pp(pingpong<2, 1>, "XCD-local, s_sleep", 0, b)) return 1;
    if (pp(pingpong<2, 0>, "XCD-local, no sleep", 0, b)) return 1;
    if (pp(pingpong<3, 1>, "L1-inv sc0, s_sleep", 0, b)) return 1;
    if (pp(pingpong<3, 0>, "L1-inv sc0, no sleep", 0, b)) return 1;
    if (pp(pingpong<4, 1>, "L1-inv sc1, s_sleep", 0, b)) return 1;
  }
  auto ar = [&](auto kern, const char* name, int g, int stride = 1) -> int {
    CHECK(hipMemset(slots, 0, 2 * G * 8));
    CHECK(hipMemset(relay, 0, 16));
    CHECK(hipMemset(xs, 0, 2 * 8 * G * 8));
    CHECK(hipMemset(gs, 0, 2 * 8 * 8));
    CHECK(hipMemset(reg, 0, 64));
    CHECK(hipMemset(out, 0, 3 * G * 8));
    const int nr = 5000;
    hipLaunchKernelGGL(kern, dim3(g * stride), dim3(64), 0, 0, slots, relay, xs, gs, reg, nr, out);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(h.data(), out, 3 * G * 8, hipMemcpyDeviceToHost));
    uint64_t acc = 0;
    for (uint32_t r = 1; r <= (uint32_t)nr; ++r) {
      uint64_t m = ~0ull;
      for (int w = 0; w < g; ++w) m = std::min<uint64_t>(m, value_of((uint64_t)w, r));
      acc ^= m + r;
    }
    int bad = 0;
    double tmax = 0;
    int perx[8] = {0};
    for (int w = 0; w < g; ++w) {
      bad += h[w] != acc;
      tmax = std::max(tmax, (double)h[g + w]);
      perx[h[2 * g + w] & 7]++;
    }
    std::printf("all-reduce %-15s G = %3d: %7.1f ns per round, wrong %d, workgroups per XCD %d %d %d %d %d %d %d %d\n",
                name, g, tmax * 10.0 / nr, bad, perx[0], perx[1], perx[2], perx[3], perx[4], perx[5], perx[6],
                perx[7]);
    return 0;
  };
  for (int g : {16, 40, 80, 160, 256}) {
    if (ar(allreduce<0>, "flat", g)) return 1;
    if (ar(allreduce<1>, "relay", g)) return 1;
    if (ar(allreduce<2>, "by XCD", g)) return 1;
  }
  for (int g : {4, 10, 16, 32}) {
    if (ar(allreduce<0, 2>, "XCD-local flat", g, 8)) return 1;
    if (ar(allreduce<1, 2>, "XCD-local relay", g, 8)) return 1;
    if (ar(allreduce<0, 3>, "L1-inv sc0 flat", g, 8)) return 1;
    if (ar(allreduce<0, 4>, "L1-inv sc1 flat", g, 8)) return 1;
    if (ar(allreduce<1, 4>, "L1-inv sc1 relay", g, 8)) return 1;
    if (ar(allreduce<0, 5>, "one-XCD agent flat", g, 8)) return 1;
    if (ar(allreduce<1, 5>, "one-XCD agent relay", g, 8)) return 1;
  }
  return 0;
}
