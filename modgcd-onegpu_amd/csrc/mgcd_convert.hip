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

// find_task for a whole wave that shares t: each step every lane probes
// one of 64 evenly spaced cum entries and the ballot keeps the slice that
// holds t, so 4,096 pairs take two dependent loads instead of twelve.
// Probes are non-decreasing in the lane, hence so is cum[pos] <= t, and the
// highest such lane L has cum[pos_L] <= t < cum[pos_(L+1)] (pos_(L+1) >
// pos_L, or cum there would be <= t too).
__device__ __forceinline__ void find_task_wave(const MrArgs& a, uint32_t t, uint32_t lane, uint32_t& p, uint32_t& n) {
  uint32_t lo = 0, hi = a.npairs;  // cum[lo] <= t < cum[hi]
  while (hi - lo > 1) {
    const uint64_t span = hi - lo;
    const uint32_t pos = lo + (uint32_t)(((uint64_t)lane * span) / kMrChunk);
    const uint64_t mask = __ballot(a.cum[pos] <= t);  // lane 0 (pos = lo) always set
    const uint32_t L = 63u - (uint32_t)__clzll((unsigned long long)mask);
    const uint32_t nlo = lo + (uint32_t)(((uint64_t)L * span) / kMrChunk);
    if (L + 1 < kMrChunk) hi = lo + (uint32_t)(((uint64_t)(L + 1) * span) / kMrChunk);
    lo = nlo;
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

// Level j -> j + 1 for larger nodes: one wave per unit, a unit being 64
// columns (a chunk) of one output node's V or P.  The wave stages the
// operands through LDS in tiles of 64 limbs of P_lo and the 127 limbs of
// the other factor those 64 columns meet, zero outside its range, so every
// lane runs the same 64 products per tile with no bounds tests.  The chunk
// is then resolved with no carry in; what it passes on goes to a carry
// record that mr_fix applies (levels of one chunk per node need neither:
// the carry out of a node's top limb is dropped, as every value is exact
// modulo 2^(32 (s + 1))).
__device__ __forceinline__ uint32_t factor_limb(const Node& hi, bool vpart, int64_t d) {
  if (vpart) return d >= 0 && d <= (int64_t)hi.s ? v_limb(hi, (uint32_t)d) : 0u;
  const uint32_t len = hi.s ? hi.s : 1u;  // an empty P_hi is 1
  return d >= 0 && d < (int64_t)len ? p_limb(hi, (uint32_t)d) : 0u;
}

__global__ __launch_bounds__(kMrChunk) void mr_level_wave(MrArgs a, uint64_t units) {
  __shared__ uint32_t lp[kMrChunk];
  __shared__ uint32_t hv[2 * kMrChunk];
  __shared__ int64_t colA[kMrChunk];
  __shared__ uint64_t colB[kMrChunk];
  const uint32_t lane = threadIdx.x;
  const uint64_t cpt = mr_chunks(a.j);
  const uint64_t w2 = 2ull << a.j;
  for (uint64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint64_t task = u / (2 * cpt);
    const bool vpart = ((u / cpt) & 1) == 0;
    const uint64_t q = u % cpt;
    uint32_t p, n;
    find_task_wave(a, (uint32_t)task, lane, p, n);
    const MrPair P = a.pairs[p];
    const Node lo = node_at(a, P, 2 * n), hi = node_at(a, P, 2 * n + 1);
    const uint64_t s = (uint64_t)lo.s + hi.s;
    const bool need_p = P.k > w2;  // the output level is not this pair's root
    const uint64_t ncol = vpart ? s + 1 : s;
    const uint64_t c0 = q * kMrChunk;
    if ((!vpart && !need_p) || c0 >= ncol) continue;  // uniform over the block
    const uint64_t c = c0 + lane;
    // Column c sums P_lo[i] * F[c - i] over i with 0 <= c - i <= hmax.
    const int64_t hmax = vpart ? (int64_t)hi.s : (int64_t)(hi.s ? hi.s : 1u) - 1;
    const uint64_t i_begin = (int64_t)c0 > hmax ? c0 - (uint64_t)hmax : 0;
    const uint64_t i_end = min(c0 + kMrChunk, (uint64_t)lo.s);
    uint64_t acc_lo = 0, acc_hi = 0;
    for (uint64_t i0 = i_begin; i0 < i_end; i0 += kMrChunk) {
      lp[lane] = i0 + lane < lo.s ? p_limb(lo, (uint32_t)(i0 + lane)) : 0u;
      // hv[x] = F[c0 - i0 - 63 + x]: lane l, product ii reads hv[l + 63 - ii].
      const int64_t d0 = (int64_t)c0 - (int64_t)i0 - (int64_t)(kMrChunk - 1);
      hv[lane] = factor_limb(hi, vpart, d0 + lane);
      hv[lane + kMrChunk] = lane + kMrChunk < 2 * kMrChunk - 1 ? factor_limb(hi, vpart, d0 + kMrChunk + lane) : 0u;
      __syncthreads();
#pragma unroll 16
      for (uint32_t ii = 0; ii < kMrChunk; ++ii) {
        const uint64_t prod = (uint64_t)lp[ii] * (uint64_t)hv[lane + kMrChunk - 1 - ii];
        acc_lo += (uint32_t)prod;
        acc_hi += prod >> 32;
      }
      __syncthreads();
    }
    int64_t e = 0;
    if (vpart && c < ncol) {
      // + V_lo (sign-extended); V_hi < 0 was read as V_hi + 2^(32 (s_hi + 1)),
      // so P_lo is subtracted once at limb s_hi + 1.
      e = (int64_t)v_limb(lo, (uint32_t)c);
      const bool neg_hi = (v_limb(hi, hi.s) >> 31) != 0;
      if (neg_hi && c >= (uint64_t)hi.s + 1 && c - (hi.s + 1) < lo.s) e -= (int64_t)p_limb(lo, (uint32_t)(c - (hi.s + 1)));
    }
    colA[lane] = (int64_t)acc_lo + e;
    colB[lane] = acc_hi;
    __syncthreads();
    const uint32_t m = (uint32_t)min((uint64_t)kMrChunk, ncol - c0);
    uint32_t* o = a.out + P.buf_off + (uint64_t)n * (2 * w2 + 1) + (vpart ? 0 : w2 + 1) + c0;
    if (lane == 0) {
      int64_t carry = 0;
      uint64_t bprev = 0;
      for (uint32_t k = 0; k < m; ++k) {
        const int64_t x = colA[k] + (int64_t)bprev + carry;
        colA[k] = (int64_t)(uint32_t)x;
        carry = x >> 32;
        bprev = colB[k];
      }
      colB[0] = (uint64_t)(carry + (int64_t)bprev);
    }
    __syncthreads();
    const uint32_t limb = (uint32_t)colA[lane];
    if (lane < m) o[lane] = limb;
    const bool ones = lane < 2 || lane >= m || limb == 0xFFFFFFFFu;
    const bool zeros = lane < 2 || lane >= m || limb == 0u;
    const bool all_ones = __all(ones), all_zeros = __all(zeros);
    const uint32_t limb1 = __shfl(limb, 1);
    if (lane == 0 && cpt > 1) {
      MrCarry r;
      r.out = (int64_t)colB[0];
      r.limb0 = limb;
      r.limb1 = m > 1 ? limb1 : 0u;
      r.flags = (all_ones ? 1u : 0u) | (all_zeros ? 2u : 0u);
      r.pad = 0;
      a.carry[u] = r;
    }
    __syncthreads();
  }
}

// Carries across the chunks of one output V or P (units of 2 per task):
// chunk q's carry in is what chunk q - 1 passes on, plus the +-1 that a
// carry in to q - 1 ripples out of it (only when its limbs above the two
// lowest are all ones, or all zeros for a borrow).  Lane 0 runs that chain
// over 64 chunks at a time; each lane then adds its chunk's carry in.
__global__ __launch_bounds__(kMrChunk) void mr_fix(MrArgs a, uint64_t pairs_of_units) {
  __shared__ MrCarry rec[kMrChunk];
  __shared__ int64_t cin[kMrChunk];
  const uint32_t lane = threadIdx.x;
  const uint64_t cpt = mr_chunks(a.j);
  const uint64_t w2 = 2ull << a.j;
  for (uint64_t u2 = blockIdx.x; u2 < pairs_of_units; u2 += gridDim.x) {
    const uint64_t task = u2 / 2;
    const bool vpart = (u2 & 1) == 0;
    uint32_t p, n;
    find_task_wave(a, (uint32_t)task, lane, p, n);
    const MrPair P = a.pairs[p];
    const Node lo = node_at(a, P, 2 * n), hi = node_at(a, P, 2 * n + 1);
    const uint64_t s = (uint64_t)lo.s + hi.s;
    const bool need_p = P.k > w2;
    const uint64_t ncol = vpart ? s + 1 : s;
    if (!vpart && !need_p) continue;
    const uint64_t nch = (ncol + kMrChunk - 1) / kMrChunk;
    if (nch <= 1) continue;
    const MrCarry* r = a.carry + (task * 2 + (vpart ? 0 : 1)) * cpt;
    uint32_t* o = a.out + P.buf_off + (uint64_t)n * (2 * w2 + 1) + (vpart ? 0 : w2 + 1);
    int64_t chain = 0;  // carry in of the next chunk to enter the chain
    for (uint64_t qb = 0; qb < nch; qb += kMrChunk) {
      const uint32_t nb = (uint32_t)min((uint64_t)kMrChunk, nch - qb);
      if (lane < nb) rec[lane] = r[qb + lane];
      __syncthreads();
      if (lane == 0) {
        for (uint32_t k = 0; k < nb; ++k) {
          cin[k] = chain;
          const int64_t x0 = (int64_t)rec[k].limb0 + chain;
          const int64_t x1 = (int64_t)rec[k].limb1 + (x0 >> 32);
          const int64_t up = x1 >> 32;  // -1, 0 or 1
          int64_t rip = 0;
          if (up == 1 && (rec[k].flags & 1u)) rip = 1;
          if (up == -1 && (rec[k].flags & 2u)) rip = -1;
          chain = rec[k].out + rip;
        }
      }
      __syncthreads();
      if (lane < nb && cin[lane] != 0) {
        const uint64_t q = qb + lane;
        const uint32_t m = (uint32_t)min((uint64_t)kMrChunk, ncol - q * kMrChunk);
        uint32_t* oc = o + q * kMrChunk;
        int64_t x = (int64_t)oc[0] + cin[lane];
        oc[0] = (uint32_t)x;
        int64_t cy = x >> 32;
        for (uint32_t k = 1; cy != 0 && k < m; ++k) {
          x = (int64_t)oc[k] + cy;
          oc[k] = (uint32_t)x;
          cy = x >> 32;
        }
      }
      __syncthreads();
    }
  }
}

// |G| of every pair from its root (level root_level of the tree; the digit
// itself when k == 1), one wave per pair: magnitude limbs at out_off,
// trimmed length in lens.  For V < 0, |V| = ~V + 1, whose carry stops at
// the lowest nonzero limb f: limbs below f stay 0, limb f is negated and
// those above are complemented.
__global__ __launch_bounds__(kMrChunk) void mr_emit(MrArgs a, const uint32_t* buf0, const uint32_t* buf1, uint32_t* res,
                                                    uint32_t* lens) {
  const uint32_t lane = threadIdx.x;
  for (uint32_t p = blockIdx.x; p < a.npairs; p += gridDim.x) {
    const MrPair P = a.pairs[p];
    uint32_t* o = res + P.out_off;
    if (P.k == 0) {
      if (lane == 0) lens[p] = 0;
      continue;
    }
    if (P.root_level == 0) {
      if (lane == 0) {
        const int32_t g = a.digits[P.dig_off].value;
        const uint32_t x = g < 0 ? 0u - (uint32_t)g : (uint32_t)g;
        o[0] = x;
        o[1] = 0;
        lens[p] = x ? 1 : 0;
      }
      continue;
    }
    const uint32_t* v = ((P.root_level & 1) ? buf1 : buf0) + P.buf_off;
    const uint64_t nl = (uint64_t)P.k + 1;
    const bool neg = (v[P.k] >> 31) != 0;
    uint64_t f = nl;
    if (neg) {
      for (uint64_t b = 0; b < nl; b += kMrChunk) {
        const uint64_t mask = __ballot(b + lane < nl && v[b + lane] != 0u);
        if (mask) {
          f = b + (uint64_t)__ffsll((unsigned long long)mask) - 1;
          break;
        }
      }
    }
    uint64_t len = 0;
    for (uint64_t b = 0; b < nl; b += kMrChunk) {
      const uint64_t l = b + lane;
      uint32_t x = 0;
      if (l < nl) {
        x = v[l];
        if (neg) x = l < f ? 0u : (l == f ? 0u - x : ~x);
        o[l] = x;
      }
      const uint64_t mask = __ballot(x != 0u);
      if (mask) len = b + 64 - (uint64_t)__clzll((unsigned long long)mask);
    }
    if (lane == 0) lens[p] = (uint32_t)len;
  }
}

constexpr uint64_t kMrMaxGrid = 1u << 20;

}  // namespace

hipError_t launch_mr_level(const MrArgs& a, uint32_t ntasks, hipStream_t s) {
  if (ntasks == 0) return hipSuccess;
  if (a.j < kMrSmallLevels) {
    hipLaunchKernelGGL(mr_level_small, dim3((ntasks + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
  }
  const uint64_t units = mr_units(a.j, ntasks);
  hipLaunchKernelGGL(mr_level_wave, dim3((uint32_t)min(units, kMrMaxGrid)), dim3(kMrChunk), 0, s, a, units);
  if (mr_chunks(a.j) > 1) {
    const uint64_t u2 = 2ull * ntasks;
    hipLaunchKernelGGL(mr_fix, dim3((uint32_t)min(u2, kMrMaxGrid)), dim3(kMrChunk), 0, s, a, u2);
  }
  return hipGetLastError();
}

hipError_t launch_mr_emit(const MrArgs& a, const uint32_t* buf0, const uint32_t* buf1, uint32_t* res, uint32_t* lens,
                          hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  hipLaunchKernelGGL(mr_emit, dim3((uint32_t)min((uint64_t)a.npairs, kMrMaxGrid)), dim3(kMrChunk), 0, s, a, buf0, buf1,
                     res, lens);
  return hipGetLastError();
}

}  // namespace mgcd
