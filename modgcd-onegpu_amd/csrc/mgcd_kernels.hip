// mgcd_kernels.hip -- gfx950 kernels of the modular GCD (Weber's algorithm).
//
// One persistent kernel computes whole GCDs.  A work queue hands each
// workgroup a work item = (pair, part): the pair's N_Q moduli are split into
// G contiguous parts, one per workgroup, and inside a workgroup thread t,
// slot k owns modulus index base + k*T + t (one modulus per lane per slot,
// slots unrolled into registers).  For each item the workgroup runs
//   (a) residues: limbs staged through LDS in tiles, read from HBM
//       coalesced, folded by a Horner step per limb and slot
//       (modMP, getGcdKernel.cu:297-324);
//   (b) Weber elimination (getGcdKernel.cu:363-402) with the pivot chosen as
//       the minimum key |t|<<32 | (m - [t>=0]) over all moduli (:143-220);
//   (c) signed mixed-radix recovery (:404-433), digit = the nonzero residue
//       of the lowest modulus index (:57-116).
//
// Elimination state.  The reference keeps canonical uq, vq and computes
// tq = (uq - b vq)/p' and t = vq/tq: two modular inverses per modulus per
// iteration.  Here each lane keeps (a, c, s) with uq = a/s, vq = c/s (mod m):
//   c' = a - b c,  a' = c p',  s' = s p',   t = a'/c'
// which are the same residues and the same t (s cancels), so the pivot
// sequence is identical, with one inverse per modulus per iteration.  At
// the end uq = a/s and recovery keeps u = n/d the same way: a lane only
// inverts when it holds the selected digit.
//
// Cross-workgroup reduction (G > 1).  Each workgroup posts its 52-bit key
// with a 12-bit round tag into its slot of a two-row ring (agent-scope
// relaxed atomics, one 8-byte store per workgroup per round); wave 0 of
// every workgroup polls all G slots of the row until every tag matches and
// takes the minimum.  Rows alternate by round: a workgroup can be at most
// one round ahead of any other, so the row it overwrites has been read by
// everyone.  Waits are bounded by spin_limit polls and then report a
// timeout instead of hanging.
#include "mgcd_arith.h"
#include "mgcd_kernels.h"

namespace mgcd {
using namespace dev;

namespace {

constexpr int TILE = 1024;                      // limbs per LDS tile
constexpr uint32_t IDX_MAX = (1u << 19) - 1;    // moduli indices < 2^19
constexpr uint64_t PAYLOAD = (1ull << 52) - 1;  // key bits below the tag
constexpr uint64_t NONE = ~0ull;                // "no candidate" / timeout
constexpr uint32_t NONE_IDX = 0xFFFFFu;

// Elimination key: |t| (32 bits) | (IDX_MAX - idx) (19) | [t < 0] (1).  For
// moduli in descending index order this orders exactly like the
// reference's (|t| << 32) | (m - [t >= 0]) (getGcdKernel.cu:154).  The
// infinity value INT32_MIN gives |t| = 2^31, above every finite |t|.
__device__ __forceinline__ uint64_t elim_key(int32_t t, uint32_t idx) {
  const uint32_t a = t < 0 ? 0u - (uint32_t)t : (uint32_t)t;
  return ((uint64_t)a << 20) | ((uint64_t)(IDX_MAX - idx) << 1) | (uint64_t)(t < 0);
}
__device__ __forceinline__ void elim_decode(uint64_t key, uint32_t& idx, int32_t& val) {
  const uint32_t a = (uint32_t)(key >> 20);
  idx = IDX_MAX - (uint32_t)((key >> 1) & IDX_MAX);
  val = (key & 1) ? (int32_t)(0u - a) : (int32_t)a;
}

// __ffs(abs(x)) as at getGcdKernel.cu:391 (abs(INT32_MIN) stays INT32_MIN).
__device__ __forceinline__ int ffs_abs(int32_t x) {
  const uint32_t a = x < 0 ? 0u - (uint32_t)x : (uint32_t)x;
  return a ? __builtin_ctz(a) + 1 : 0;
}

__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return b < a ? b : a; }

// A value known to be the same in every lane, moved to scalar registers so
// that everything derived from it (pivot index, modulus, bookkeeping) stays
// scalar too.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Wave-wide minimum, result uniform.  Within each row of 16 lanes DPP
// quad_perm [1,0,3,2], [2,3,0,1] and row_ror 4, 8 leave the row minimum in
// every lane; the four rows are then combined through readlane in SGPRs.
// Must be called with all 64 lanes active.
__device__ __forceinline__ uint32_t dpp_min_u32(uint32_t x) {
  x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0xB1, 0xF, 0xF, false));
  x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x4E, 0xF, 0xF, false));
  x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x124, 0xF, 0xF, false));
  x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x128, 0xF, 0xF, false));
  const uint32_t r0 = __builtin_amdgcn_readlane(x, 0), r1 = __builtin_amdgcn_readlane(x, 16);
  const uint32_t r2 = __builtin_amdgcn_readlane(x, 32), r3 = __builtin_amdgcn_readlane(x, 48);
  return min(min(r0, r1), min(r2, r3));
}

__device__ __forceinline__ uint64_t wave_min(uint64_t x) {
  const uint32_t hi = (uint32_t)(x >> 32);
  const uint32_t mh = dpp_min_u32(hi);
  const uint32_t ml = dpp_min_u32(hi == mh ? (uint32_t)x : 0xFFFFFFFFu);
  return ((uint64_t)mh << 32) | ml;
}

struct Ctx {
  uint64_t* red;             // LDS [2][B][NW]
  uint64_t* bc;              // LDS [2][B]
  int nb;                    // B: per-parity stride of red (B*NW) and bc (B)
  int par;                   // LDS double-buffer parity
  uint32_t G, g;             // parts of the group, this part
  uint32_t rnd;              // cross-workgroup round counter
  unsigned long long* bar;   // this group's 2*G*B slots
  uint32_t spin_limit;
};

// Minima of x[0..N) over the workgroup; every thread gets them.  One
// barrier; the LDS buffer alternates so the next call cannot overwrite
// values still being read.
template <int T, int N>
__device__ __forceinline__ void wg_min(uint64_t (&x)[N], Ctx& C) {
  constexpr int NW = T / 64;
  uint64_t* red = C.red + C.par * C.nb * NW;  // fixed stride: calls with other N never overlap
#pragma unroll
  for (int b = 0; b < N; ++b) {
    const uint64_t w = wave_min(x[b]);
    if ((threadIdx.x & 63) == 0) red[b * NW + (threadIdx.x >> 6)] = w;
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < N; ++b) {
    uint64_t r = red[b * NW];
#pragma unroll
    for (int w = 1; w < NW; ++w) r = min64(r, red[b * NW + w]);
    x[b] = uniform64(r);
  }
  C.par ^= 1;
}

// Minima of x[0..N) over the G workgroups of the group (x: this workgroup's
// minima, uniform).  Slot layout of a row: [G][N].  Sets every x[b] to NONE
// on timeout.
template <int N>
__device__ void grid_min(uint64_t (&x)[N], Ctx& C) {
  C.rnd += 1;
  const uint64_t tag = (uint64_t)((C.rnd - 1) % 4095u + 1u);
  unsigned long long* row = C.bar + (size_t)(C.rnd & 1) * C.G * N;
  if (threadIdx.x < N) {
    uint64_t v = x[0];
#pragma unroll
    for (int b = 1; b < N; ++b)
      if ((int)threadIdx.x == b) v = x[b];
    __hip_atomic_store(row + (size_t)C.g * N + threadIdx.x, (unsigned long long)((tag << 52) | (v & PAYLOAD)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 64) {
    uint64_t m[N];
#pragma unroll
    for (int b = 0; b < N; ++b) m[b] = NONE;
    bool late = false;
    const uint32_t nslots = C.G * N;
    for (uint32_t j = threadIdx.x; j < nslots; j += 64) {
      uint64_t v;
      uint32_t spins = 0;
      for (;;) {
        v = __hip_atomic_load(row + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((v >> 52) == tag) break;
        if (++spins > C.spin_limit) { late = true; break; }
        __builtin_amdgcn_s_sleep(1);
      }
      const uint32_t b = j % N;
#pragma unroll
      for (int q = 0; q < N; ++q)
        if ((uint32_t)q == b) m[q] = min64(m[q], v & PAYLOAD);
    }
    const bool any_late = __ballot(late) != 0;
#pragma unroll
    for (int b = 0; b < N; ++b) {
      const uint64_t r = wave_min(m[b]);
      if (threadIdx.x == 0) C.bc[C.par * C.nb + b] = any_late ? NONE : r;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < N; ++b) x[b] = uniform64(C.bc[C.par * C.nb + b]);
  C.par ^= 1;
}

// (a) Residues of x (n limbs) for every slot row in use.
template <int T, int K>
__device__ void residues(const uint32_t* __restrict__ x, uint32_t n, double (&r)[K], const ModD (&M)[K],
                         int rows, double* tile) {
#pragma unroll
  for (int k = 0; k < K; ++k) r[k] = 0.0;
  for (int hi = (int)n; hi > 0; hi -= TILE) {
    const int lo = hi > TILE ? hi - TILE : 0;
    const int len = hi - lo;
    __syncthreads();
    for (int j = threadIdx.x; j < len; j += T) tile[j] = (double)x[lo + j];
    __syncthreads();
    for (int j = len - 1; j >= 0; --j) {
      const double limb = tile[j];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k < rows) r[k] = horner_step(r[k], limb, M[k]);
    }
  }
}

// y_b^-1 for y[0..B) (each canonical or signed reduced, nonzero mod m) with
// one modular inverse: prefix products p_b = y_0 ... y_b, inv = p_{B-1}^-1,
// then y_b^-1 = inv * p_{b-1} and inv <- inv * y_b going down (Montgomery's
// batch-inversion trick).  Results signed reduced.
template <int B, class A>
__device__ __forceinline__ void batch_inverse(const double (&y)[B], double (&inv)[B], ModD M) {
  if (B == 1) {
    inv[0] = A::modinv(canon(y[0], M), M);
    return;
  }
  double p[B];
  p[0] = y[0];
#pragma unroll
  for (int b = 1; b < B; ++b) p[b] = A::mulmod(p[b - 1], y[b], M);
  double r = A::modinv(canon(p[B - 1], M), M);
#pragma unroll
  for (int b = B - 1; b >= 1; --b) {
    inv[b] = A::mulmod(r, p[b - 1], M);
    r = A::mulmod(r, y[b], M);
  }
  inv[0] = r;
}

// Everything one workgroup does for part C.g of a group of up to B pairs.
// Pairs of a group run the elimination in lockstep (a pair that has left
// the loop only skips its arithmetic); recovery then runs pair by pair.
template <int T, int K, int B, class A>
__device__ void process_group(const KernelArgs& args, const GroupDesc& Gd, double* tile, Ctx& C,
                              uint64_t* sel_lds) {
  const uint32_t t = threadIdx.x;
  const uint32_t g = C.g;
  const uint32_t nq = Gd.nq;
  // Part g owns moduli [g*nq/G, (g+1)*nq/G): sizes differ by at most one,
  // none is empty (G <= nq), none exceeds T*K (checked by the host).
  const uint32_t base = (uint32_t)(((uint64_t)g * nq) / C.G);
  const uint32_t cnt = (uint32_t)(((uint64_t)(g + 1) * nq) / C.G) - base;
  const int rows = (int)((cnt + T - 1) / T);
  const bool leader = (g == 0) && (t == 0);
  const int L = args.L;
  const uint32_t npairs = Gd.count;

  ModD M[K];
  uint32_t valid = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t j = (uint32_t)k * T + t;
    if (j < cnt) {
      M[k].m = args.mod_d[base + j];
      M[k].minv = args.mod_inv[base + j];
      valid |= 1u << k;
    } else {
      M[k].m = 3.0;  // unused lane of a partial row: any modulus keeps it harmless
      M[k].minv = 1.0 / 3.0;
    }
  }

  double a[B][K], c[B][K], s[B][K];
  uint32_t active[B];
  int32_t status[B], remaining[B], ubits[B], vbits[B];
  uint32_t ntrace[B], ndig[B];
  bool live[B];  // still in the elimination loop
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const bool here = (uint32_t)b < npairs;
    const PairDesc P = args.pairs[Gd.first + (here ? b : 0)];
    if (here) {
      residues<T, K>(args.limbs + P.u_off, P.nu, a[b], M, rows, tile);
      residues<T, K>(args.limbs + P.v_off, P.nv, c[b], M, rows, tile);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) a[b][k] = c[b][k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) s[b][k] = 1.0;
    active[b] = here ? valid : 0u;
    status[b] = here ? KST_OK : KST_NOTRUN;
    remaining[b] = (int32_t)nq;                  // totalModuliRemaining (:350)
    ubits[b] = (int32_t)((P.nu + 1) * 32);       // :351
    vbits[b] = (int32_t)((P.nv + 1) * 32);       // :352
    ntrace[b] = ndig[b] = 0;
    live[b] = here;
  }

  // ---- (b) elimination ----------------------------------------------------
  // Initial keys t = u/v (infinity where v == 0).
  uint64_t key[B];
#pragma unroll
  for (int b = 0; b < B; ++b) key[b] = NONE;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < rows && ((valid >> k) & 1)) {
      const uint32_t idx = base + (uint32_t)k * T + t;
      double y[B], inv[B];
#pragma unroll
      for (int b = 0; b < B; ++b) y[b] = (live[b] && c[b][k] != 0.0) ? c[b][k] : 1.0;
      batch_inverse<B, A>(y, inv, M[k]);
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (!live[b]) continue;
        int32_t tv = INT32_MIN;
        if (c[b][k] != 0.0) tv = to_signed(A::mulmod(a[b][k], inv[b], M[k]), M[k]);
        key[b] = min64(key[b], elim_key(tv, idx));
      }
    }
  }
  wg_min<T, B>(key, C);
  if (C.G > 1) grid_min<B>(key, C);
  uint64_t piv[B];
  bool any_live = false;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    piv[b] = key[b];
    if (live[b] && piv[b] == NONE) { status[b] = KST_TIMEOUT; live[b] = false; }
    any_live |= live[b];
  }

  while (any_live) {
    uint32_t pidx[B];
    int32_t pval[B];
    double pm[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      pidx[b] = IDX_MAX + 1;  // matches no modulus
      pval[b] = 0;
      pm[b] = 0.0;
      if (!live[b]) continue;
      elim_decode(piv[b], pidx[b], pval[b]);
      const PairDesc& P = args.pairs[Gd.first + b];
      if (leader && ntrace[b] < P.trace_cap) {
        mgcd_pair_t tp;
        tp.modulus = args.mod_u32[pidx[b]];
        tp.value = pval[b];
        args.trace[P.trace_off + ntrace[b]] = tp;
      }
      ntrace[b]++;
      pm[b] = args.mod_d[pidx[b]];
      key[b] = NONE;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k < rows && ((valid >> k) & 1)) {
        const uint32_t idx = base + (uint32_t)k * T + t;
        // Forward: new state in place, prefix products of the values to
        // invert (1 where there is nothing to invert).
        double p[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          double y = 1.0;
          if (live[b]) {
            if (idx == pidx[b]) active[b] &= ~(1u << k);
            if ((active[b] >> k) & 1) {
              const double Pp = pm[b] - M[k].m;  // p mod m, signed (p, m in [2^(L-1), 2^L))
              const double c2 = reduce(a[b][k] - A::mulmod((double)pval[b], c[b][k], M[k]), M[k]);
              a[b][k] = A::mulmod(c[b][k], Pp, M[k]);
              s[b][k] = A::mulmod(s[b][k], Pp, M[k]);
              c[b][k] = c2;
              if (c2 != 0.0) y = c2;
            }
          }
          p[b] = (b == 0) ? y : A::mulmod(p[b - 1], y, M[k]);
        }
        // One inverse of the product, then back down: inv_b = r * p_{b-1},
        // r <- r * y_b (Montgomery's batch inversion).
        double r = A::modinv(canon(p[B - 1], M[k]), M[k]);
#pragma unroll
        for (int b = B - 1; b >= 0; --b) {
          if (!live[b]) continue;
          const bool has = ((active[b] >> k) & 1) && c[b][k] != 0.0;
          int32_t tv = INT32_MIN;
          if (has) {
            const double inv = (b == 0) ? r : A::mulmod(r, p[b - 1], M[k]);
            tv = to_signed(A::mulmod(a[b][k], inv, M[k]), M[k]);
            if (b > 0) r = A::mulmod(r, c[b][k], M[k]);
          }
          key[b] = min64(key[b], elim_key(tv, idx));
        }
      }
    }
    wg_min<T, B>(key, C);
    if (C.G > 1) grid_min<B>(key, C);
    any_live = false;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      if (!live[b]) continue;
      remaining[b] -= 1;                                             // :390
      const int32_t tbits = ubits[b] - (L - 1) + ffs_abs(pval[b]);  // :391
      ubits[b] = vbits[b];
      vbits[b] = tbits;
      if (remaining[b] * (L - 2) <= ubits[b]) {                     // :393-399
        status[b] = KST_REDUX;
        live[b] = false;
      } else if (key[b] == NONE) {
        status[b] = KST_TIMEOUT;
        live[b] = false;
      } else {
        piv[b] = key[b];
        if ((piv[b] & 1) && (uint32_t)(piv[b] >> 20) == 0x80000000u) live[b] = false;  // infinity: done
      }
      any_live |= live[b];
    }
  }

  // ---- (c) recovery, pair by pair ------------------------------------------
#pragma unroll
  for (int b = 0; b < B; ++b) {
    if (status[b] != KST_OK) continue;
    const PairDesc& P = args.pairs[Gd.first + b];
    // Candidate: the lowest index holding a nonzero residue (:57-116).
    auto select = [&](void) -> uint64_t {
      uint32_t best = NONE_IDX;
#pragma unroll
      for (int k = K - 1; k >= 0; --k)
        if (k < rows && ((active[b] >> k) & 1) && a[b][k] != 0.0) best = base + (uint32_t)k * T + t;
      uint64_t w[1] = {(uint64_t)best};
      wg_min<T, 1>(w, C);
      uint64_t sel[1] = {(uint64_t)NONE_IDX << 32};
      if (w[0] != NONE_IDX) {
        if (best == w[0]) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if (base + (uint32_t)k * T + t == best) {
              const int32_t v =
                  to_signed(A::mulmod(a[b][k], A::modinv(canon(s[b][k], M[k]), M[k]), M[k]), M[k]);
              *sel_lds = (w[0] << 32) | (uint32_t)v;
            }
          }
        }
        __syncthreads();
        sel[0] = uniform64(*sel_lds);
        __syncthreads();
      }
      if (C.G > 1) grid_min<1>(sel, C);
      return sel[0];
    };

    uint64_t sel = select();
    if (sel == NONE) status[b] = KST_TIMEOUT;
    while (sel != NONE) {
      uint32_t pidx = (uint32_t)(sel >> 32);
      int32_t pval = (int32_t)(uint32_t)sel;
      if (pidx == NONE_IDX) { pidx = 0; pval = 0; }  // no nonzero residue: thread 0's pair (:57-64)
      const double pm = args.mod_d[pidx];
      const double gv = (double)pval;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k < rows && ((valid >> k) & 1)) {
          const uint32_t idx = base + (uint32_t)k * T + t;
          if (idx == pidx) active[b] &= ~(1u << k);
          if ((active[b] >> k) & 1) {
            a[b][k] = reduce(a[b][k] - A::mulmod(gv, s[b][k], M[k]), M[k]);  // u' = (u - g) / p'
            s[b][k] = A::mulmod(s[b][k], pm - M[k].m, M[k]);
          }
        }
      }
      const uint64_t next = select();
      if (leader) {
        mgcd_pair_t dp;
        dp.modulus = args.mod_u32[pidx];
        dp.value = pval;
        args.digits[P.dig_off + ndig[b]] = dp;
      }
      ndig[b]++;
      remaining[b] -= 1;                                                                // :428
      if (remaining[b] <= 0) { status[b] = pval != 0 ? KST_RECOVERY : KST_OK; break; }  // :429-430, :440
      if (next == NONE) { status[b] = KST_TIMEOUT; break; }
      sel = next;
      if ((int32_t)(uint32_t)sel == 0) break;  // value 0: done (:433)
    }
  }

  if (leader) {
#pragma unroll
    for (int b = 0; b < B; ++b) {
      if ((uint32_t)b >= npairs) continue;
      PairOut o;
      o.status = status[b];
      o.ndig = ndig[b];
      o.ntrace = ntrace[b];
      o.pad = 0;
      args.out[Gd.first + b] = o;
    }
  }
}

template <int T, int K, int B, class A>
__global__ __launch_bounds__(T) void gcd_kernel(KernelArgs args) {
  __shared__ double tile[TILE];
  __shared__ uint64_t red[2 * B * (T / 64)];
  __shared__ uint64_t bc[2 * B];
  __shared__ uint64_t sel_lds;
  __shared__ uint32_t item_lds;
  Ctx C;
  C.red = red;
  C.bc = bc;
  C.nb = B;
  C.par = 0;
  C.spin_limit = args.spin_limit;
  for (;;) {
    if (threadIdx.x == 0) item_lds = atomicAdd(args.queue, 1u);
    __syncthreads();
    const uint32_t it = item_lds;
    __syncthreads();
    if (it >= args.nitems) break;
    const WorkItem W = args.items[it];
    const GroupDesc Gd = args.groups[W.group];
    C.G = Gd.G;
    C.g = W.part;
    C.rnd = 0;
    C.bar = args.bar + Gd.bar_off;
    process_group<T, K, B, A>(args, Gd, tile, C, &sel_lds);
  }
}

// ---------------------------------------------------------------------------
// Small helpers.
__global__ void fill_u64_kernel(unsigned long long* p, size_t n, unsigned long long v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Certifier of quasi_quo_f (the role of certifyQuoRemQuasi.cu:25-48).
// Postcondition for integers x >= 0, y >= 1 with Q = floor(x/y):
//   q~ <= Q always, and q~ >= Q - 1 whenever Q < 2^19.
// Inputs reach quasi_quo_f the way modinv passes them: fl32 of an exact
// integer (round to nearest even).
__device__ __forceinline__ bool qq_bad(uint64_t x, uint64_t y) {
  const float q = quasi_quo_f((float)(double)x, (float)(double)y);
  const int64_t r = (int64_t)x - (int64_t)q * (int64_t)y;
  if (r < 0) return true;
  return r >= (int64_t)(2 * y) && x < (y << 19);
}

// Mode 0: every x in [1, 2^23), y in [1, 2^22), point by point.
__global__ void certify_float_kernel(unsigned long long* fail, unsigned long long* checked) {
  unsigned long long bad = 0, cnt = 0;
  for (uint32_t y = blockIdx.x * blockDim.x + threadIdx.x + 1; y < (1u << 22); y += gridDim.x * blockDim.x) {
    for (uint32_t x = 1; x < (1u << 23); ++x) bad += qq_bad(x, y);
    cnt += (1u << 23) - 1;
  }
  if (bad) atomicAdd(fail, bad);
  atomicAdd(checked, cnt);
}

// Mode 1: width L, y in [y_lo, y_hi), every quotient k of x in [0, 2^L):
// endpoints x = k*y and x = min((k+1)*y - 1, 2^L - 1).  q~(x, y) is
// non-decreasing in x (conversion, multiplication by a positive constant
// and rounding are monotone), so the endpoints bound every x of class k.
// A launch covers one octave of y; unit u is (y = y_lo + (u >> cshift),
// chunk = u & (2^cshift - 1)) of 4096 quotients; chunks past kmax(y) skip.
__global__ void certify_endpoints_kernel(int L, uint64_t y_lo, uint64_t y_hi, int cshift, unsigned long long* fail,
                                         unsigned long long* checked) {
  const uint64_t xmax = (1ull << L) - 1;
  const uint64_t CH = 4096;
  const uint64_t units = (y_hi - y_lo) << cshift;
  const uint64_t cmask = (1ull << cshift) - 1;
  unsigned long long bad = 0, cnt = 0;
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < units;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t y = y_lo + (u >> cshift);
    const uint64_t k0 = (u & cmask) * CH;
    const uint64_t kmax = xmax / y;
    if (k0 > kmax) continue;
    const uint64_t k1 = min(kmax, k0 + CH - 1);
    for (uint64_t k = k0; k <= k1; ++k) {
      if (k > 0) bad += qq_bad(k * y, y);
      const uint64_t xh = min((k + 1) * y - 1, xmax);
      if (xh > 0) bad += qq_bad(xh, y);
      cnt += 2;
    }
  }
  if (bad) atomicAdd(fail, bad);
  if (cnt) atomicAdd(checked, cnt);
}

}  // namespace

// ---------------------------------------------------------------------------
// Variant table.
#define MGCD_VARIANT(T, K, B, A, W) {T, K, B, W, reinterpret_cast<const void*>(&gcd_kernel<T, K, B, A>)}
#define MGCD_VARIANTS(A, W)                                                                              \
  MGCD_VARIANT(256, 1, 1, A, W), MGCD_VARIANT(256, 2, 1, A, W), MGCD_VARIANT(256, 4, 1, A, W),           \
      MGCD_VARIANT(256, 8, 1, A, W), MGCD_VARIANT(1024, 1, 1, A, W), MGCD_VARIANT(1024, 2, 1, A, W),     \
      MGCD_VARIANT(1024, 4, 1, A, W), MGCD_VARIANT(1024, 8, 1, A, W), MGCD_VARIANT(256, 2, 2, A, W),     \
      MGCD_VARIANT(256, 4, 2, A, W), MGCD_VARIANT(256, 1, 4, A, W), MGCD_VARIANT(256, 2, 4, A, W),       \
      MGCD_VARIANT(256, 4, 4, A, W), MGCD_VARIANT(256, 1, 8, A, W),      \
      MGCD_VARIANT(256, 2, 8, A, W)
static const KernelVariant kVariants[] = {MGCD_VARIANTS(Arith32, 1), MGCD_VARIANTS(Arith22, 0)};
#undef MGCD_VARIANTS
#undef MGCD_VARIANT

int kernel_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
const KernelVariant& kernel_variant(int i) { return kVariants[i]; }
int find_kernel_variant(int T, int K, int B, int wide) {
  for (int i = 0; i < kernel_variant_count(); ++i)
    if (kVariants[i].T == T && kVariants[i].K == K && kVariants[i].B == B && kVariants[i].wide == wide) return i;
  return -1;
}

hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_certify_float(unsigned long long* fail, unsigned long long* checked, int grid, hipStream_t s) {
  hipLaunchKernelGGL(certify_float_kernel, dim3(grid), dim3(256), 0, s, fail, checked);
  return hipGetLastError();
}
hipError_t launch_certify_endpoints(int L, uint64_t y_lo, uint64_t y_hi, unsigned long long* fail,
                                    unsigned long long* checked, int grid, hipStream_t s) {
  // chunks of 4096 quotients per y: enough for kmax(y_lo) = (2^L - 1) / y_lo
  const uint64_t kmax = ((1ull << L) - 1) / y_lo;
  int cshift = 0;
  while ((4096ull << cshift) <= kmax) ++cshift;
  hipLaunchKernelGGL(certify_endpoints_kernel, dim3(grid), dim3(256), 0, s, L, y_lo, y_hi, cshift, fail, checked);
  return hipGetLastError();
}

}  // namespace mgcd
