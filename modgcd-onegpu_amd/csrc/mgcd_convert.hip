// mgcd_convert.hip -- signed mixed-radix digits -> positional limbs on the
// device, by a product tree (the role of the host Horner loop at
// GmpCudaDevice-gcd.cu:113-129, which is O(k^2) and sequential).
//
// A pair's k digits (m_i, g_i), least significant first, stand for
//   G = g_0 + m_0 (g_1 + m_1 (g_2 + ... + m_(k-2) g_(k-1))).
// For a range [a, b) of digits let
//   V(a, b) = sum_{a <= i < b} g_i prod_{a <= j < i} m_j,   P(a, b) = prod_{a <= j < b} m_j.
// Two adjacent ranges combine as
//   V(a, c) = V(a, b) + P(a, b) V(b, c),   P(a, c) = P(a, b) P(b, c),
// and G = V(0, k).  Level j of the tree holds nodes of 2^j digits (the last
// one shorter); level 0 is the digits themselves.  Since |g_i| <= (m_i+1)/2,
// |V(a, b)| < P(a, b) < 2^(32 s) for a node of s digits, so V fits s + 1
// limbs in two's complement and P fits s limbs; every operation is exact
// modulo 2^(32 (s + 1)).
//
// Layout of level j >= 1 in a level buffer: node n of a pair at
// buf_off + n (2^(j+1) + 1): V in 2^j + 1 limbs, then P in 2^j limbs (the
// node's own s + 1 and s limbs are meaningful).  Levels alternate between
// two buffers; a pair's region is 8k + 8 limbs in each.
#include "mgcd_kernels.h"

namespace mgcd {
namespace {

struct Node {
  const uint32_t* v;  // nullptr: one digit (g, m), or an empty node (0, 1)
  const uint32_t* p;
  int32_t g;          // the digit, when v == nullptr
  uint32_t m;
  uint32_t s;         // digits in the node
};

// Node n of the pair at level a.j.  A node past the last digit is empty:
// V = 0, P = 1, s = 0.
__device__ __forceinline__ Node node_at(const MrArgs& a, const MrPair& P, uint32_t n) {
  Node x;
  const uint32_t w = 1u << a.j;
  const uint32_t first = n * w;
  x.s = first < P.k ? min(w, P.k - first) : 0u;
  x.v = x.p = nullptr;
  x.g = 0;
  x.m = 1u;
  if (x.s == 0) return x;
  if (a.j == 0) {
    const mgcd_pair_t d = a.digits[P.dig_off + first];
    x.g = d.value;
    x.m = d.modulus;
  } else {
    const uint64_t base = P.buf_off + (uint64_t)n * (2u * w + 1u);
    x.v = a.in + base;
    x.p = a.in + base + w + 1u;
  }
  return x;
}
// Limb l of V, sign-extended beyond its s + 1 limbs.
__device__ __forceinline__ uint32_t v_limb(const Node& x, uint32_t l) {
  if (!x.v) return l == 0 ? (uint32_t)x.g : (x.g < 0 ? 0xFFFFFFFFu : 0u);
  if (l <= x.s) return x.v[l];
  return (x.v[x.s] >> 31) ? 0xFFFFFFFFu : 0u;
}
// Limb l of P (zero beyond its s limbs).
__device__ __forceinline__ uint32_t p_limb(const Node& x, uint32_t l) {
  if (!x.v) return l == 0 ? x.m : 0u;
  return l < x.s ? x.p[l] : 0u;
}

// Column c of V' = V_lo + P_lo V_hi before carries: A = the low halves
// of the column's limb products, plus the V_lo limb and the sign
// correction; B = their high halves, which belong to column c + 1.
__device__ __forceinline__ void v_column(const Node& lo, const Node& hi, uint32_t c, int64_t& A, uint64_t& B) {
  uint64_t a = 0, b = 0;
  const uint32_t i0 = c > hi.s ? c - hi.s : 0u;
  const uint32_t i1 = min(c + 1, lo.s);
  for (uint32_t i = i0; i < i1; ++i) {
    const uint64_t prod = (uint64_t)p_limb(lo, i) * (uint64_t)v_limb(hi, c - i);
    a += (uint32_t)prod;
    b += prod >> 32;
  }
  // V_hi < 0: its unsigned reading is V_hi + 2^(32 (s_hi + 1)), so P_lo is
  // subtracted once at limb s_hi + 1.
  int64_t e = (int64_t)v_limb(lo, c);
  const bool neg_hi = (v_limb(hi, hi.s) >> 31) != 0;
  if (neg_hi && c >= hi.s + 1 && c - (hi.s + 1) < lo.s) e -= (int64_t)p_limb(lo, c - (hi.s + 1));
  A = (int64_t)a + e;
  B = b;
}
// Column c of P' = P_lo P_hi.
__device__ __forceinline__ void p_column(const Node& lo, const Node& hi, uint32_t c, int64_t& A, uint64_t& B) {
  if (hi.s == 0) {  // P_hi = 1
    A = (int64_t)p_limb(lo, c);
    B = 0;
    return;
  }
  uint64_t a = 0, b = 0;
  const uint32_t i0 = c + 1 > hi.s ? c + 1 - hi.s : 0u;
  const uint32_t i1 = min(c + 1, lo.s);
  for (uint32_t i = i0; i < i1; ++i) {
    const uint64_t prod = (uint64_t)p_limb(lo, i) * (uint64_t)p_limb(hi, c - i);
    a += (uint32_t)prod;
    b += prod >> 32;
  }
  A = (int64_t)a;
  B = b;
}

// Task t of the level -> (pair, output node).
__device__ __forceinline__ void find_task(const MrArgs& a, uint32_t t, uint32_t& p, uint32_t& n) {
  uint32_t lo = 0, hi = a.npairs;  // cum[lo] <= t < cum[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (a.cum[mid] <= t) lo = mid;
    else hi = mid;
  }
  p = lo;
  n = t - a.cum[lo];
}

// Level j -> j + 1, one task per thread (small nodes).
__global__ void mr_level_small(MrArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.cum[a.npairs]) return;
  uint32_t p, n;
  find_task(a, t, p, n);
  const MrPair P = a.pairs[p];
  const Node lo = node_at(a, P, 2 * n), hi = node_at(a, P, 2 * n + 1);
  const uint32_t w2 = 2u << a.j;  // output node capacity in digits
  uint32_t* ov = a.out + P.buf_off + (uint64_t)n * (2u * w2 + 1u);
  uint32_t* op = ov + w2 + 1u;
  const uint32_t s = lo.s + hi.s;
  const bool need_p = P.k > w2;  // the output level is not this pair's root
  int64_t carry = 0;
  uint64_t bprev = 0;
  for (uint32_t c = 0; c <= s; ++c) {
    int64_t A;
    uint64_t B;
    v_column(lo, hi, c, A, B);
    const int64_t x = A + (int64_t)bprev + carry;
    ov[c] = (uint32_t)x;
    carry = x >> 32;
    bprev = B;
  }
  if (need_p) {
    carry = 0;
    bprev = 0;
    for (uint32_t c = 0; c < s; ++c) {
      int64_t A;
      uint64_t B;
      p_column(lo, hi, c, A, B);
      const int64_t x = A + (int64_t)bprev + carry;
      op[c] = (uint32_t)x;
      carry = x >> 32;
      bprev = B;
    }
  }
}

// Level j -> j + 1, one workgroup of 256 per task: columns in chunks of
// 256, one per thread, then thread 0 resolves the chunk's carries.
constexpr int MR_T = 256;
__global__ __launch_bounds__(MR_T) void mr_level_large(MrArgs a) {
  __shared__ int64_t colA[MR_T];
  __shared__ uint64_t colB[MR_T];
  uint32_t p, n;
  find_task(a, blockIdx.x, p, n);
  const MrPair P = a.pairs[p];
  const Node lo = node_at(a, P, 2 * n), hi = node_at(a, P, 2 * n + 1);
  const uint32_t w2 = 2u << a.j;
  uint32_t* ov = a.out + P.buf_off + (uint64_t)n * (2u * w2 + 1u);
  uint32_t* op = ov + w2 + 1u;
  const uint32_t s = lo.s + hi.s;
  const bool need_p = P.k > w2;
  for (int which = 0; which < (need_p ? 2 : 1); ++which) {
    const uint32_t ncol = which == 0 ? s + 1 : s;
    uint32_t* o = which == 0 ? ov : op;
    int64_t carry = 0;
    uint64_t bprev = 0;
    for (uint32_t c0 = 0; c0 < ncol; c0 += MR_T) {
      const uint32_t c = c0 + threadIdx.x;
      int64_t A = 0;
      uint64_t B = 0;
      if (c < ncol) {
        if (which == 0) v_column(lo, hi, c, A, B);
        else p_column(lo, hi, c, A, B);
      }
      colA[threadIdx.x] = A;
      colB[threadIdx.x] = B;
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint32_t m = min((uint32_t)MR_T, ncol - c0);
        for (uint32_t q = 0; q < m; ++q) {
          const int64_t x = colA[q] + (int64_t)bprev + carry;
          o[c0 + q] = (uint32_t)x;
          carry = x >> 32;
          bprev = colB[q];
        }
      }
      __syncthreads();
    }
  }
}

// |G| of every pair from its root (level root_level of the tree; the digit
// itself when k == 1): magnitude limbs at out_off, trimmed length in lens.
__global__ void mr_emit(MrArgs a, const uint32_t* buf0, const uint32_t* buf1, uint32_t* res, uint32_t* lens) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npairs) return;
  const MrPair P = a.pairs[p];
  uint32_t* o = res + P.out_off;
  if (P.k == 0) {
    lens[p] = 0;
    return;
  }
  const uint32_t* v = nullptr;
  uint32_t tmp[2];
  if (P.root_level == 0) {
    const int32_t g = a.digits[P.dig_off].value;
    tmp[0] = (uint32_t)g;
    tmp[1] = g < 0 ? 0xFFFFFFFFu : 0u;
    v = tmp;
  } else {
    v = ((P.root_level & 1) ? buf1 : buf0) + P.buf_off;
  }
  const bool neg = (v[P.k] >> 31) != 0;
  uint32_t len = 0;
  uint64_t c = 1;
  for (uint32_t l = 0; l <= P.k; ++l) {
    uint32_t x = v[l];
    if (neg) {
      const uint64_t y = (uint64_t)(~x) + c;
      x = (uint32_t)y;
      c = y >> 32;
    }
    o[l] = x;
    if (x) len = l + 1;
  }
  lens[p] = len;
}

}  // namespace

hipError_t launch_mr_level(const MrArgs& a, uint32_t ntasks, hipStream_t s) {
  if (ntasks == 0) return hipSuccess;
  if (a.j < kMrSmallLevels) {
    hipLaunchKernelGGL(mr_level_small, dim3((ntasks + 255) / 256), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL(mr_level_large, dim3(ntasks), dim3(MR_T), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_mr_emit(const MrArgs& a, const uint32_t* buf0, const uint32_t* buf1, uint32_t* res, uint32_t* lens,
                          hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  hipLaunchKernelGGL(mr_emit, dim3((a.npairs + 255) / 256), dim3(256), 0, s, a, buf0, buf1, res, lens);
  return hipGetLastError();
}

}  // namespace mgcd
