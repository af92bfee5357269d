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

__device__ __forceinline__ uint64_t wave_min(uint64_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = min64(x, __shfl_xor(x, off, 64));
  return x;
}

struct Ctx {
  uint64_t* red;             // LDS [2][NW]
  uint64_t* bc;              // LDS [2]
  int par;                   // LDS double-buffer parity
  uint32_t G, g;             // parts of the pair, this part
  uint32_t rnd;              // cross-workgroup round counter
  unsigned long long* bar;   // this pair's 2*G slots
  uint32_t spin_limit;
};

// Minimum of x over the workgroup; every thread gets it.  One barrier; the
// LDS buffer alternates so the next call cannot overwrite values still read.
template <int T>
__device__ __forceinline__ uint64_t wg_min(uint64_t x, Ctx& C) {
  constexpr int NW = T / 64;
  x = wave_min(x);
  uint64_t* red = C.red + C.par * NW;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t r = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) r = min64(r, red[w]);
  C.par ^= 1;
  return r;
}

// Minimum over the G workgroups of the pair (x: this workgroup's minimum,
// the same in every thread).  Returns NONE on timeout.
__device__ uint64_t grid_min(uint64_t x, Ctx& C) {
  C.rnd += 1;
  const uint64_t tag = (uint64_t)((C.rnd - 1) % 4095u + 1u);
  unsigned long long* row = C.bar + (size_t)(C.rnd & 1) * C.G;
  if (threadIdx.x == 0)
    __hip_atomic_store(row + C.g, (unsigned long long)((tag << 52) | (x & PAYLOAD)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 64) {
    uint64_t m = NONE;
    bool late = false;
    for (uint32_t j = threadIdx.x; j < C.G; j += 64) {
      uint64_t v;
      uint32_t spins = 0;
      for (;;) {
        v = __hip_atomic_load(row + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((v >> 52) == tag) break;
        if (++spins > C.spin_limit) { late = true; break; }
        __builtin_amdgcn_s_sleep(1);
      }
      m = min64(m, v & PAYLOAD);
    }
    m = wave_min(m);
    const bool any_late = __ballot(late) != 0;
    if (threadIdx.x == 0) C.bc[C.par] = any_late ? NONE : m;
  }
  __syncthreads();
  const uint64_t r = C.bc[C.par];
  C.par ^= 1;
  return r;
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

template <int T, int K, class A>
__device__ void process_item(const KernelArgs& args, const PairDesc& P, uint32_t pair, double* tile, Ctx& C,
                             uint64_t* sel_lds) {
  const uint32_t t = threadIdx.x;
  const uint32_t g = C.g;
  // Part g owns moduli [g*nq/G, (g+1)*nq/G): sizes differ by at most one,
  // none is empty (G <= nq), none exceeds T*K (checked by the host).
  const uint32_t base = (uint32_t)(((uint64_t)g * P.nq) / C.G);
  const uint32_t cnt = (uint32_t)(((uint64_t)(g + 1) * P.nq) / C.G) - base;
  const int rows = (int)((cnt + T - 1) / T);
  const bool leader = (g == 0) && (t == 0);
  const int L = args.L;

  ModD M[K];
  double a[K], c[K], s[K];
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
    s[k] = 1.0;
  }
  uint32_t active = valid;

  residues<T, K>(args.limbs + P.u_off, P.nu, a, M, rows, tile);
  residues<T, K>(args.limbs + P.v_off, P.nv, c, M, rows, tile);

  int32_t status = KST_OK;
  uint32_t ntrace = 0, ndig = 0;
  int32_t remaining = (int32_t)P.nq;             // totalModuliRemaining (:350)
  int32_t ubits = (int32_t)((P.nu + 1) * 32);    // :351
  int32_t vbits = (int32_t)((P.nv + 1) * 32);    // :352

  // ---- (b) elimination ----------------------------------------------------
  uint64_t lmin = NONE;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < rows && ((valid >> k) & 1)) {
      int32_t tv = INT32_MIN;
      if (c[k] != 0.0) tv = to_signed(A::mulmod(a[k], A::modinv(canon(c[k], M[k]), M[k]), M[k]), M[k]);
      lmin = min64(lmin, elim_key(tv, base + (uint32_t)k * T + t));
    }
  }
  uint64_t piv = wg_min<T>(lmin, C);
  if (C.G > 1) piv = grid_min(piv, C);

  bool done = (piv == NONE);
  if (done) status = KST_TIMEOUT;
  while (!done) {
    uint32_t pidx;
    int32_t pval;
    elim_decode(piv, pidx, pval);
    if (leader && ntrace < P.trace_cap) {
      mgcd_pair_t tp;
      tp.modulus = args.mod_u32[pidx];
      tp.value = pval;
      args.trace[P.trace_off + ntrace] = tp;
    }
    ntrace++;
    const double pm = args.mod_d[pidx];
    const double B = (double)pval;
    lmin = NONE;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k < rows && ((valid >> k) & 1)) {
        const uint32_t idx = base + (uint32_t)k * T + t;
        if (idx == pidx) active &= ~(1u << k);
        int32_t tv = INT32_MIN;
        if ((active >> k) & 1) {
          const double Pp = pm - M[k].m;  // p mod m, signed (p, m in [2^(L-1), 2^L))
          const double c2 = reduce(a[k] - A::mulmod(B, c[k], M[k]), M[k]);
          const double a2 = A::mulmod(c[k], Pp, M[k]);
          s[k] = A::mulmod(s[k], Pp, M[k]);
          a[k] = a2;
          c[k] = c2;
          if (c2 != 0.0) tv = to_signed(A::mulmod(a2, A::modinv(canon(c2, M[k]), M[k]), M[k]), M[k]);
        }
        lmin = min64(lmin, elim_key(tv, idx));
      }
    }
    uint64_t next = wg_min<T>(lmin, C);
    if (C.G > 1) next = grid_min(next, C);
    remaining -= 1;                                        // :390
    const int32_t tbits = ubits - (L - 1) + ffs_abs(pval);  // :391
    ubits = vbits;
    vbits = tbits;
    if (remaining * (L - 2) <= ubits) { status = KST_REDUX; done = true; break; }  // :393-399
    if (next == NONE) { status = KST_TIMEOUT; done = true; break; }
    piv = next;
    if ((piv & 1) && (uint32_t)(piv >> 20) == 0x80000000u) break;  // pivot value is infinity
  }

  // ---- (c) recovery ---------------------------------------------------------
  if (!done) {
    // Candidate of this lane: the lowest index holding a nonzero residue.
    auto select = [&](void) -> uint64_t {
      uint32_t best = NONE_IDX;
#pragma unroll
      for (int k = K - 1; k >= 0; --k)
        if (k < rows && ((active >> k) & 1) && a[k] != 0.0) best = base + (uint32_t)k * T + t;
      const uint64_t wbest = wg_min<T>((uint64_t)best, C);
      uint64_t key = (uint64_t)NONE_IDX << 32;
      if (wbest != NONE_IDX) {
        if (best == wbest) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if (base + (uint32_t)k * T + t == best) {
              const int32_t v = to_signed(A::mulmod(a[k], A::modinv(canon(s[k], M[k]), M[k]), M[k]), M[k]);
              *sel_lds = (wbest << 32) | (uint32_t)v;
            }
          }
        }
        __syncthreads();
        key = *sel_lds;
        __syncthreads();
      }
      if (C.G > 1) key = grid_min(key, C);
      return key;
    };

    uint64_t sel = select();
    if (sel == NONE) status = KST_TIMEOUT;
    while (sel != NONE) {
      uint32_t pidx = (uint32_t)(sel >> 32);
      int32_t pval = (int32_t)(uint32_t)sel;
      if (pidx == NONE_IDX) { pidx = 0; pval = 0; }  // no nonzero residue: thread 0's pair (:57-64)
      const double pm = args.mod_d[pidx];
      const double Gd = (double)pval;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k < rows && ((valid >> k) & 1)) {
          const uint32_t idx = base + (uint32_t)k * T + t;
          if (idx == pidx) active &= ~(1u << k);
          if ((active >> k) & 1) {
            a[k] = reduce(a[k] - A::mulmod(Gd, s[k], M[k]), M[k]);  // u' = (u - g) / p'
            s[k] = A::mulmod(s[k], pm - M[k].m, M[k]);
          }
        }
      }
      const uint64_t next = select();
      if (leader) {
        mgcd_pair_t dp;
        dp.modulus = args.mod_u32[pidx];
        dp.value = pval;
        args.digits[P.dig_off + ndig] = dp;
      }
      ndig++;
      remaining -= 1;                                                        // :428
      if (remaining <= 0) { status = pval != 0 ? KST_RECOVERY : KST_OK; break; }  // :429-430, :440
      if (next == NONE) { status = KST_TIMEOUT; break; }
      sel = next;
      if ((int32_t)(uint32_t)sel == 0) break;  // value 0: done (:433)
    }
  }

  if (leader) {
    PairOut o;
    o.status = status;
    o.ndig = ndig;
    o.ntrace = ntrace;
    o.pad = 0;
    args.out[pair] = o;
  }
}

template <int T, int K, class A>
__global__ __launch_bounds__(T) void gcd_kernel(KernelArgs args) {
  __shared__ double tile[TILE];
  __shared__ uint64_t red[2 * (T / 64)];
  __shared__ uint64_t bc[2];
  __shared__ uint64_t sel_lds;
  __shared__ uint32_t item_lds;
  Ctx C;
  C.red = red;
  C.bc = bc;
  C.par = 0;
  C.spin_limit = args.spin_limit;
  for (;;) {
    if (threadIdx.x == 0) item_lds = atomicAdd(args.queue, 1u);
    __syncthreads();
    const uint32_t it = item_lds;
    __syncthreads();
    if (it >= args.nitems) break;
    const WorkItem W = args.items[it];
    const PairDesc P = args.pairs[W.pair];
    C.G = P.G;
    C.g = W.part;
    C.rnd = 0;
    C.bar = args.bar + P.bar_off;
    process_item<T, K, A>(args, P, W.pair, tile, C, &sel_lds);
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
#define MGCD_VARIANT(T, K, A, W) {T, K, W, reinterpret_cast<const void*>(&gcd_kernel<T, K, A>)}
static const KernelVariant kVariants[] = {
    MGCD_VARIANT(256, 1, Arith32, 1),  MGCD_VARIANT(256, 2, Arith32, 1),  MGCD_VARIANT(256, 4, Arith32, 1),
    MGCD_VARIANT(256, 8, Arith32, 1),  MGCD_VARIANT(256, 16, Arith32, 1), MGCD_VARIANT(1024, 1, Arith32, 1),
    MGCD_VARIANT(1024, 2, Arith32, 1), MGCD_VARIANT(1024, 4, Arith32, 1), MGCD_VARIANT(1024, 8, Arith32, 1),
    MGCD_VARIANT(256, 1, Arith22, 0),  MGCD_VARIANT(256, 2, Arith22, 0),  MGCD_VARIANT(256, 4, Arith22, 0),
    MGCD_VARIANT(256, 8, Arith22, 0),  MGCD_VARIANT(256, 16, Arith22, 0), MGCD_VARIANT(1024, 1, Arith22, 0),
    MGCD_VARIANT(1024, 2, Arith22, 0), MGCD_VARIANT(1024, 4, Arith22, 0), MGCD_VARIANT(1024, 8, Arith22, 0),
};
#undef MGCD_VARIANT

int kernel_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
const KernelVariant& kernel_variant(int i) { return kVariants[i]; }
int find_kernel_variant(int T, int K, int wide) {
  for (int i = 0; i < kernel_variant_count(); ++i)
    if (kVariants[i].T == T && kVariants[i].K == K && kVariants[i].wide == wide) return i;
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
