// mgcd_kernels.hip -- gfx950 kernels of the modular GCD (Weber's algorithm).
//
// One persistent kernel computes whole GCDs.  Pairs with the same number of
// moduli N_Q form groups of up to B pairs; a group's N_Q moduli are split
// into G contiguous parts, one per workgroup, and inside a workgroup thread
// t, slot k owns modulus index base + k*T + t for every pair of the group
// (one modulus per lane per slot, slots unrolled into registers).  A work
// queue hands out (group, part) items.  For each item the workgroup runs
//   (a) residues: limbs staged through LDS in tiles, read from HBM
//       coalesced, folded by a Horner step per limb and slot
//       (modMP, getGcdKernel.cu:297-324); kept in HBM for (c);
//   (b) Weber elimination (getGcdKernel.cu:363-402), the B pairs in
//       lockstep, the pivot being the minimum key |t|, modulus, sign over
//       all moduli (:143-220);
//   (c) replay of the pivot sequence on the surviving moduli, then signed
//       mixed-radix recovery (:404-433), digit = the nonzero residue of the
//       lowest modulus index (:57-116), the B pairs in lockstep.
//
// Elimination state.  The reference keeps canonical uq, vq and computes
// tq = (uq - b vq)/p' and t = vq/tq: two modular inverses per modulus per
// iteration.  The pivot choice only needs t = uq/vq, which obeys
//   t' = p' / (t - b)        (1/0 = infinity, 1/infinity = 0)
// as long as (uq, vq) != (0, 0) mod m (the map (u, v) -> (v, (u - b v)/p')
// is invertible, so that state never changes).  Each lane therefore keeps
// only t per (pair, slot), with sentinels TINF (vq == 0), TZERO (uq == vq
// == 0, never a pivot) and TRET (retired: was a pivot); one inverse per
// modulus and iteration remains, and Montgomery's trick shares it between
// the B pairs of the group, which hold the same modulus in each lane.
// uq, vq themselves are needed only on the moduli that survive the
// elimination; (c) recomputes them there from the stored residues by
// replaying the recorded pivots on (a, c, s) with uq = a/s, vq = c/s:
//   c' = a - b c,  a' = c p',  s' = s p'.
//
// Cross-workgroup reduction (G > 1).  Wave 0 of each workgroup posts its B
// per-pair minima with a 12-bit round tag into its slots of a two-row ring
// (agent-scope relaxed atomics, one 8-byte store per pair per workgroup
// per round), then polls all G*B slots of the row until every tag matches.
// Rows alternate by round: a workgroup can be at most one round ahead of
// any other, so the row it overwrites has been read by everyone.  Waits
// are bounded by spin_limit polls and then report a timeout instead of
// hanging.
#include <type_traits>

#include "mgcd_kernels.h"  // first: the HIP runtime header, which the two below rely on
#include "mgcd_arith.h"
#include "mgcd_keys.h"

namespace mgcd {
using namespace dev;
using namespace keys;

namespace {

constexpr int TILE = 1024;  // doubles of LDS staging (limbs, or 512 pivots) of the one-pair
                            // tilings; batch tilings stage more
// t sentinels.  Finite t are the reference's toSigned values, |t| < 2^31.
constexpr double TINF = 2147483648.0;    // vq == 0: the reference's MOD_INFINITY
constexpr double TZERO = -2147483648.0;  // uq == vq == 0
constexpr double TRET = 4294967296.0;    // retired modulus
constexpr double TFIN = 2147483648.0;    // |t| < TFIN: finite
// Within one round only: a finite t equal to its pair's pivot (t - b == 0
// mod m), whose next state is TINF.  Same high word as TINF, low word 1.
constexpr double TPOLE = 2147483648.0 + 0x1p-21;

// The reference's toSigned (getGcdKernel.cu:278-281) of an exact integer
// r, |r| < 2^40: the representative in [-(m+1)/2, (m-3)/2].  rint((r+1)/m)
// has no ties (m is odd) and lies 1/(2m) away from them, far above the
// rounding error.
__device__ __forceinline__ double tsigned(double r, ModD M) {
  const double q = __builtin_rint(__builtin_fma(r, M.minv, M.minv));
  return __builtin_fma(-q, M.m, r);
}

// s_setprio with a level known once loops are unrolled (below 0: 0).
__device__ __forceinline__ void set_prio(int p) {
  switch (p) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
  }
}

// Phase clocks of the instrumented build (-DMGCD_PHASE_CLOCKS): shader-clock
// cycles (s_memtime) per phase of a work item, as wave 0 sees them, summed
// into args.clocks at the end of each item; the 100 MHz s_memrealtime total
// calibrates the clock.  Without the macro every call compiles to nothing.
enum : int {
  PH_RESIDUES = 0, PH_COMPUTE, PH_BARRIER, PH_EXCHANGE, PH_BOOK, PH_BARRIER2, PH_TAIL, PH_REPLAY, PH_RECOVERY,
  PH_ROUNDS, PH_ITEMS, PH_REALTIME, PH_MEMTIME,
  // one-pass collects (groups of up to 256 slots): shader cycles from issuing
  // a round's slot loads to their return, the collects whose first loads
  // found every slot posted, and the collects
  PH_FIRST_LOAD, PH_COLLECT_READY, PH_COLLECTS,
  // every wave's own compute and barrier laps, by wave (0-7): the two waves
  // of a SIMD need not finish a round's compute together
  PH_WAVE_COMPUTE = 16, PH_WAVE_BARRIER = 24
};
static_assert(PH_WAVE_BARRIER + 8 <= kPhaseClocks, "phase clock slots");
struct PhaseClock {
#ifdef MGCD_PHASE_CLOCKS
  uint64_t acc[PH_ROUNDS] = {};
  uint64_t extra[3] = {};  // PH_FIRST_LOAD .. PH_COLLECTS
  __device__ void acc_extra(int p, uint64_t v) { extra[p - PH_FIRST_LOAD] += v; }
  uint64_t t = 0, t0 = 0, r0 = 0, rounds = 0;
  __device__ void start() {
    t = t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  __device__ void lap(int p) {
    const uint64_t n = __builtin_amdgcn_s_memtime();
    acc[p] += n - t;
    t = n;
  }
  __device__ void round() { ++rounds; }
  __device__ void flush(unsigned long long* out, bool writer) {
    const uint64_t n = __builtin_amdgcn_s_memtime(), r = __builtin_amdgcn_s_memrealtime();
    if (!writer || !out) return;
    for (int p = 0; p < PH_ROUNDS; ++p) atomicAdd(out + p, (unsigned long long)acc[p]);
    atomicAdd(out + PH_ROUNDS, (unsigned long long)rounds);
    atomicAdd(out + PH_ITEMS, 1ull);
    atomicAdd(out + PH_REALTIME, (unsigned long long)(r - r0));
    atomicAdd(out + PH_MEMTIME, (unsigned long long)(n - t0));
    for (int p = 0; p < 3; ++p) atomicAdd(out + PH_FIRST_LOAD + p, (unsigned long long)extra[p]);
  }
  __device__ void flush_waves(unsigned long long* out) {
    const uint32_t w = threadIdx.x >> 6;
    if (!out || (threadIdx.x & 63) != 0 || w >= 8) return;
    atomicAdd(out + PH_WAVE_COMPUTE + w, (unsigned long long)acc[PH_COMPUTE]);
    atomicAdd(out + PH_WAVE_BARRIER + w, (unsigned long long)acc[PH_BARRIER]);
  }
#else
  __device__ void start() {}
  __device__ void lap(int) {}
  __device__ void round() {}
  __device__ void flush(unsigned long long*, bool) {}
  __device__ void flush_waves(unsigned long long*) {}
#endif
};

// __ffs(abs(x)) as at getGcdKernel.cu:391 (abs(INT32_MIN) stays INT32_MIN).
__device__ __forceinline__ int ffs_abs(int32_t x) {
  const uint32_t a = x < 0 ? 0u - (uint32_t)x : (uint32_t)x;
  return a ? __builtin_ctz(a) + 1 : 0;
}

// Reads p[i] after `after` has been computed: the empty asm makes the
// address depend on it, so the compiler cannot hoist the load (and keep
// the value in registers across a long stretch of code).  The cursor is
// updated in place, so a chain of such reads keeps one address register.
template <class V>
using LdsPtr = const __attribute__((address_space(3))) V*;
template <class V>
__device__ __forceinline__ LdsPtr<V> lds_ptr(const V* x) {
  return (LdsPtr<V>)x;
}
template <class V>
__device__ __forceinline__ V load_after(LdsPtr<V>& p, int i, double after) {
  asm volatile("" : "+v"(p) : "v"(after));
  return p[i];
}

// Values every lane holds alike, moved into scalar registers.  process_group
// is a called function whose arguments reach it through memory: reading
// such a value once per work item and keeping it here keeps the per-round
// paths (bookkeeping, exchange) free of loads of the kernel's arguments.
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uniform(uint64_t x) {
  return (uint64_t)uniform((uint32_t)x) | ((uint64_t)uniform((uint32_t)(x >> 32)) << 32);
}
template <class V>
__device__ __forceinline__ V* uniform(V* p) {
  return reinterpret_cast<V*>(uniform((uint64_t)reinterpret_cast<uintptr_t>(p)));
}
// A kernel argument array as a global pointer held in scalar registers.
// (The host pass only parses device code; there the pointer stays generic,
// since its C++ copies of structs cannot read an address-space-qualified
// source.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class V>
using GPtr = __attribute__((address_space(1))) V*;
#else
template <class V>
using GPtr = V*;
#endif
template <class V>
__device__ __forceinline__ GPtr<V> gptr(V* p) {
  return (GPtr<V>)uniform(p);
}
// A copy of a global pointer laundered into vector registers.  Where scalar
// registers run short the compiler re-reads a kernel argument with a scalar
// load, and waits for it, at each use; the laundered copy stays in place.
template <class V>
__device__ __forceinline__ GPtr<V> vptr(GPtr<V> p) {
  uint64_t x = (uint64_t)(uintptr_t)p;
  asm volatile("" : "+v"(x));
  return (GPtr<V>)(uintptr_t)x;
}
// Exchange slots as global memory, so their atomics compile to global (not
// flat) loads and stores.
using SlotPtr = __attribute__((address_space(1))) unsigned long long*;
__device__ __forceinline__ SlotPtr slots(unsigned long long* p) { return (SlotPtr)p; }

// ---------------------------------------------------------------------------
// Cross-lane minima of doubles.
__device__ __forceinline__ uint32_t lo32(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
__device__ __forceinline__ uint32_t hi32(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
__device__ __forceinline__ double mk64(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int CTRL>
__device__ __forceinline__ double dpp64(double x) {
  // Every control used here is a permutation within a row, so every lane
  // has a source and the moves need no old value.
  return mk64((uint32_t)__builtin_amdgcn_mov_dpp((int)lo32(x), CTRL, 0xF, 0xF, false),
              (uint32_t)__builtin_amdgcn_mov_dpp((int)hi32(x), CTRL, 0xF, 0xF, false));
}
// v_permlane32_swap: the upper 32 lanes of a trade places with the lower
// 32 lanes of b.  v_permlane16_swap: within each half, the odd row (16
// lanes) of a with the even row of b.
template <int S>
__device__ __forceinline__ void swap64(double& a, double& b) {
  if constexpr (S == 0) {
    const auto l = __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false, false);
    a = mk64(l[0], h[0]);
    b = mk64(l[1], h[1]);
  } else {
    const auto l = __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false, false);
    a = mk64(l[0], h[0]);
    b = mk64(l[1], h[1]);
  }
}
// Stage S of a 64-lane butterfly; partner lanes differ in lane bit 5 - S:
// S = 0, 1 permlane32/16 swaps, S = 2 row_mirror, 3 row_half_mirror,
// 4 quad_perm [2,3,0,1], 5 quad_perm [1,0,3,2].
template <int S>
constexpr int stage_dpp() {
  return S == 2 ? 0x140 : S == 3 ? 0x141 : S == 4 ? 0x4E : 0xB1;
}
// The lanes whose index has bit b set, as a wave mask (a constant: a
// select on it needs no per-lane test).
constexpr uint64_t lane_bit_mask(int b) {
  uint64_t m = 0;
  for (int l = 0; l < 64; ++l)
    if ((l >> b) & 1) m |= 1ull << l;
  return m;
}
// Halving step: a lane whose bit is 0 keeps the minimum of the first
// halves (a) of itself and its partner, a lane whose bit is 1 of the
// second halves (b); the result is left in a.
template <int S>
__device__ __forceinline__ void half_step(double& a, double& b) {
  if constexpr (S <= 1) {
    swap64<S>(a, b);
    a = __builtin_fmin(a, b);
  } else {
    const bool hi = __builtin_amdgcn_inverse_ballot_w64(lane_bit_mask(5 - S));
    const double send = hi ? a : b;
    const double keep = hi ? b : a;
    a = __builtin_fmin(keep, dpp64<stage_dpp<S>()>(send));
  }
}
template <int S>
__device__ __forceinline__ void full_step(double& x) {
  if constexpr (S <= 1) {
    double y = x;
    swap64<S>(x, y);
    x = __builtin_fmin(x, y);
  } else {
    x = __builtin_fmin(x, dpp64<stage_dpp<S>()>(x));
  }
}
template <int S, int N, int B>
__device__ __forceinline__ void scatter_min(double (&v)[B]) {
  if constexpr (S < 6) {
    if constexpr (N >= 2) {
#pragma unroll
      for (int j = 0; j < N / 2; ++j) half_step<S>(v[j], v[j + N / 2]);
      scatter_min<S + 1, N / 2, B>(v);
    } else {
      full_step<S>(v[0]);
      scatter_min<S + 1, 1, B>(v);
    }
  }
}
constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }
// Wave minima of B values per lane (B a power of two <= 64), as a
// reduce-scatter: about B/2 + 6 exchanges instead of 6 per value.
// Afterwards lane l holds the minimum of value l >> (6 - log2 B).
template <int B>
__device__ __forceinline__ double wave_min_scatter(double (&v)[B]) {
  scatter_min<0, B, B>(v);
  return v[0];
}
// Minimum over the lanes l with the same l % B (B a power of two <= 32).
template <int B>
__device__ __forceinline__ double class_min(double x) {
  if constexpr (B <= 1) x = __builtin_fmin(x, dpp64<0x121>(x));  // row_ror 1
  if constexpr (B <= 2) x = __builtin_fmin(x, dpp64<0x122>(x));  // row_ror 2
  if constexpr (B <= 4) x = __builtin_fmin(x, dpp64<0x124>(x));  // row_ror 4
  if constexpr (B <= 8) x = __builtin_fmin(x, dpp64<0x128>(x));  // row_ror 8
  if constexpr (B <= 16) full_step<1>(x);
  full_step<0>(x);
  return x;
}

// ---------------------------------------------------------------------------
// Workgroup state: the small per-round fields of one group (Small), and the
// staging area that the groups a workgroup runs share (Big).
template <int T, int K, int B>
struct Small {
  static constexpr int NW = T / 64;
  // The small per-round fields come first in LDS: their addresses stay
  // within the 16-bit immediate offset of ds_read/ds_write.
  double red[B][NW];          // per-wave minima
  double bb[B], pp[B];        // current pivot of each pair: value, modulus
  uint32_t pidx[B];           // its modulus index (NONE_IDX: none)
  uint32_t flag;              // some pair still running
  uint64_t actw[B][K][NW];    // surviving moduli: ballot words
  int32_t status[B];
  uint32_t ntrace[B];
  double sel[B];              // recovery: the owner's (index, value) key
  uint32_t retire[T * K];     // bit b: this part's modulus is pair b's pivot
  uint32_t retw[K][NW];       // per slot row and wave: OR of its retire words
  uint32_t item;
};
// Staging for (a) and (c), and during (b) the group's moduli (when N_Q
// fits), so that a pivot's modulus is an LDS read instead of a global one
// on the path between two rounds.  Group tilings stage 16 doubles per thread
// (fewer barriers and global round trips per staged chunk).
template <int T, int K, int B>
struct Big {
  static constexpr int TL = B >= 8 ? 16 * T : TILE;
  static constexpr uint32_t MODS = (TL * 8 + B * T * K * 2) / 4;
  union {
    struct {
      double tile[TL];           // limb tile / pivot tile (value, modulus)
      uint16_t surv[B * T * K + 4 * B];  // survivors of the group: pair << 11 | local offset
                                          // (each pair's run padded to four entries in group tilings)
    } st;
    uint32_t mods[MODS];         // (b): modulus of every index < N_Q
  } u;
};
template <int T, int K, int B>
struct Shared {
  Small<T, K, B> s;
  Big<T, K, B> b;
};

struct Ctx {
  uint32_t G, g;             // parts of the group, this part
  uint32_t rnd;              // cross-workgroup round counter
  uint32_t phase;            // 0: elimination rounds; 1: the rounds after them
  unsigned long long* bar;   // this group's 2*(G+1)*B exchange slots, then G residue-job words
  uint32_t spin_limit;
  bool relay;                // exchange through part 0 (large G)
  bool abort;                // a wait of this part timed out, or a peer's did: post ABORT from now on
#ifdef MGCD_PHASE_CLOCKS
  uint64_t first_load = 0, collect_ready = 0, collects = 0;  // see PH_FIRST_LOAD
#endif
};

__device__ __forceinline__ Ctx uniform_ctx(const Ctx& c) {
  Ctx u;
  u.G = uniform(c.G);
  u.g = uniform(c.g);
  u.rnd = uniform(c.rnd);
  u.phase = uniform(c.phase);
  u.bar = uniform(c.bar);
  u.spin_limit = uniform(c.spin_limit);
  u.relay = uniform((uint32_t)c.relay) != 0;
  u.abort = uniform((uint32_t)c.abort) != 0;
  return u;
}

// Cross-workgroup minima, called by wave 0 (all 64 lanes): lane b < B
// brings this workgroup's value for pair b; lane l gets the group's minimum
// for pair l % B, KNONE for every pair when a wait timed out or a peer
// posted ABORT.  Slot layout of a row: [G][B], so lane l polls only slots of
// pair l % B.  Every part polls the whole row (one hop), or, for groups of
// many parts, part 0 polls it and posts the minima into a relay row of B
// slots that the others poll (two hops, but G*B + (G-1)*B polled slots per
// round instead of G*G*B).
//
// Tag: an 11-bit round number (never 0, so cleared slots never match) and
// the phase bit.  Once a part has given up it posts ABORT in every later
// round, whatever it would have posted, so a peer still waiting for it
// gives up too instead of reading a value of another kind.
//
// grid_post starts a round (this part's slot), grid_collect finishes it.  A
// part posts round r + 1 only after collecting round r, so no part is more
// than one round ahead of another, and the two rows of the ring suffice.
__device__ __forceinline__ uint64_t round_tag(const Ctx& C) {
  return (uint64_t)(((C.rnd - 1) % 2047u + 1u) | (C.phase << 11));
}

template <int B>
__device__ __forceinline__ void grid_post(double x, Ctx& C) {
  const uint32_t lane = __lane_id();
  C.rnd += 1;
  const uint64_t tag = round_tag(C);
  SlotPtr row = slots(C.bar) + (size_t)(C.rnd & 1) * C.G * B;
  if (lane < (uint32_t)B)
    __hip_atomic_store(row + (size_t)C.g * B + lane,
                       (unsigned long long)((tag << 52) | (C.abort ? ABORT : key_bits(x))), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// s_sleep units (64 cycles each) between a part's post and its first slot
// loads (see grid_collect): group tilings, one-pair tilings polling every
// part, and one-pair parts waiting for the relay (two hops: the leader's
// collect and its relay post).
constexpr int kFirstPollSleep = 9;
constexpr int kFirstPollSleepOnePair = 12;
constexpr int kRelayPollSleep = 48;

template <int B>
__device__ double grid_collect(Ctx& C) {
  const uint32_t lane = __lane_id();
  const uint64_t tag = round_tag(C);
  // A slot is ready when its round matches; a matching round of the other
  // phase means a peer has left the sequence, and counts as ABORT.
  auto ready = [&](uint64_t v) { return ((v >> 52) & 0x7FFu) == (tag & 0x7FFu); };
  auto aborted = [&](uint64_t v) { return (v >> 52) != tag || (v & PAYLOAD) == ABORT; };
  SlotPtr row = slots(C.bar) + (size_t)(C.rnd & 1) * C.G * B;
  SlotPtr relay = slots(C.bar) + (size_t)2 * C.G * B + (size_t)(C.rnd & 1) * B;
  if (C.relay && C.g != 0) {
    // The relay slot cannot change before the leader has collected every
    // part's post and posted the minimum: about 1.3 us before the first
    // load (measured on one ~1M-bit pair, 317 parts: 253-263 ms without,
    // 237-252 / 224-249 / 237-242 / 252 ms with 20 / 40 / 60 / 80 units).
    __builtin_amdgcn_s_sleep(kRelayPollSleep);
    uint64_t v = __hip_atomic_load(relay + lane % B, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    bool late = false;
    while (!ready(v)) {
      if (++spins > C.spin_limit) {
        late = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      v = __hip_atomic_load(relay + lane % B, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const bool none = !late && aborted(v);
    if (__ballot(late) || __ballot(none)) {
      C.abort = true;
      return KNONE;
    }
    return key_from_bits(v & PAYLOAD);
  }
  double m = KNONE;
  bool late = false, none = false;
  const uint32_t nslots = C.G * B;
  // Up to PL slots per lane per pass: their loads are in flight together.
  // Several passes (groups of many parts): every slot not yet posted is
  // reloaded in each round trip (a pass's late slots would otherwise cost a
  // round trip each; measured 12 % faster on 317-part groups).  One pass (up
  // to PL * 64 slots): a lane waits for its slots in turn and reloads only
  // the one it waits for (the joint form measured 0.5 % slower on 10-part
  // groups).  (Eight slots per lane, one pass for up to 512 slots, measured
  // slower: 10 % on 45-part groups.)  One-pair tilings, short of nothing
  // but the exchange, poll up to PJ = 8 slots per lane jointly: the relay
  // leader of a 317-part pair reads all of them in one pass.
  constexpr int PL = 4;
  constexpr int PJ = B == 1 ? 8 : PL;
  if (nslots > PL * 64) {
    // Group tilings wait before the first loads as in the one-pass case
    // below (16,384 mixed 4k/500k-bit pairs, 45-part groups: 4,081 -> 4,151
    // GCD/s; 4,096 mixed-size pairs 9,249 -> 9,340).
    if constexpr (B > 1) __builtin_amdgcn_s_sleep(kFirstPollSleep);
    for (uint32_t j0 = lane; j0 < nslots && !late; j0 += PJ * 64) {
      uint64_t v[PJ];
      bool pend[PJ];
#pragma unroll
      for (int q = 0; q < PJ; ++q) {
        const uint32_t j = j0 + 64 * q;
        pend[q] = j < nslots;
        v[q] = pend[q] ? __hip_atomic_load(row + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag << 52) | NEUTRAL;
      }
      uint32_t spins = 0;
      for (;;) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < PJ; ++q) {
          pend[q] = pend[q] && !ready(v[q]);
          any |= pend[q];
        }
        if (!any) break;
        if (++spins > C.spin_limit) {
          late = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int q = 0; q < PJ; ++q)
          if (pend[q]) v[q] = __hip_atomic_load(row + j0 + 64 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int q = 0; q < PJ; ++q) {
        none |= aborted(v[q]);
        m = __builtin_fmin(m, key_from_bits(v[q] & PAYLOAD));
      }
    }
  } else {
    const uint32_t j0 = lane;
    uint64_t v[PL];
    // Group tilings wait about 0.24 us (9 x 64 cycles) before the first
    // loads: the parts of a group post within a few tenths of a microsecond
    // of each other, and a load that arrives before the latest post costs a
    // whole further round trip (past the L2, about 0.7 us) -- without the
    // wait 80 % of the rounds needed one.  Measured on 4,096 x 100 kbit:
    // 6,091 GCD/s without, 6,224 / 6,334 / 6,317 / 6,290 / 6,210 / 6,080 /
    // 5,880 with 6 / 9 / 12 / 15 / 20 / 28 / 40 units.  One-pair tilings
    // (20 parts at 100 kbit): 20.3 ms without, 19.0-19.3 ms with 9 or 12.
    if constexpr (B > 1) __builtin_amdgcn_s_sleep(kFirstPollSleep);
    else __builtin_amdgcn_s_sleep(kFirstPollSleepOnePair);
#ifdef MGCD_PHASE_CLOCKS
    const uint64_t tl0 = __builtin_amdgcn_s_memtime();
#endif
#pragma unroll
    for (int q = 0; q < PL; ++q) {
      const uint32_t j = j0 + 64 * q;
      v[q] = j < nslots ? __hip_atomic_load(row + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag << 52) | NEUTRAL;
    }
#ifdef MGCD_PHASE_CLOCKS
    {
      bool all = true;
#pragma unroll
      for (int q = 0; q < PL; ++q) all &= ready(v[q]);
      C.first_load += __builtin_amdgcn_s_memtime() - tl0;
      C.collect_ready += __ballot(!all) == 0;
      C.collects += 1;
    }
#endif
    // (Pipelined polls, several generations of loads in flight so that a
    // late slot is seen sooner, measured slower: the extra poll traffic of
    // every part lengthens every load's round trip.)
#pragma unroll
    for (int q = 0; q < PL; ++q) {
      uint32_t spins = 0;
      while (!ready(v[q])) {
        if (++spins > C.spin_limit) {
          late = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        v[q] = __hip_atomic_load(row + j0 + 64 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      none |= aborted(v[q]);
      m = __builtin_fmin(m, key_from_bits(v[q] & PAYLOAD));
    }
  }
  m = class_min<B>(m);
  const bool give_up = (__ballot(late) | __ballot(none)) != 0;
  if (C.relay && lane < (uint32_t)B)
    __hip_atomic_store(relay + lane, (unsigned long long)((tag << 52) | (give_up ? ABORT : key_bits(m))),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (give_up) {
    C.abort = true;
    return KNONE;
  }
  return m;
}

template <int B>
__device__ __forceinline__ double grid_min(double x, Ctx& C) {
  grid_post<B>(x, C);
  return grid_collect<B>(C);
}

// (a) Residues of NOP operands at once (x[o], n[o] limbs) for every slot
// row in use: NOP * K independent Horner chains per lane keep the FP64
// pipeline busy.  The operands are aligned at their least significant limb
// and padded with leading zero limbs, which leave a Horner residue
// unchanged, so every chain takes the same steps.  Limbs are staged
// through LDS, TL / NOP per operand at a time, read from HBM coalesced.
template <int T, int K, int NOP, int TL>
__device__ void residues(GPtr<const uint32_t> (&x)[NOP], const uint32_t (&n)[NOP], double (&r)[NOP][K],
                         const ModD (&M)[K], int rows, double* tile) {
  constexpr int TS = TL / NOP;
  uint32_t nmax = 0;
  double ls[NOP];  // the last limb of each operand, times 2^32, not yet added (horner_step)
#pragma unroll
  for (int o = 0; o < NOP; ++o) {
    nmax = max(nmax, n[o]);
    ls[o] = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) r[o][k] = 0.0;
  }
  // Step s handles limb index nmax - 1 - s of every operand.
  for (uint32_t s0 = 0; s0 < nmax; s0 += TS) {
    const uint32_t len = min((uint32_t)TS, nmax - s0);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < len * NOP; j += T) {
      const uint32_t o = j / len, jj = j - o * len;
      const uint32_t i = nmax - 1 - (s0 + jj);
#pragma unroll
      for (int q = 0; q < NOP; ++q)
        if ((uint32_t)q == o) tile[q * TS + jj] = i < n[q] ? (double)x[q][i] * 4294967296.0 : 0.0;
    }
    __syncthreads();
    // Unrolled by eight with the group tilings' larger tiles (the limbs'
    // LDS reads issue under the previous steps' arithmetic; four measured
    // 2 % slower, sixteen the same).  The number of slot rows in use is
    // uniform: one branch here, not a select per step and row.
    constexpr int RU = TL > TILE ? 8 : 1;
    auto steps = [&](auto all_rows) {
#pragma unroll RU
      for (uint32_t j = 0; j < len; ++j) {
#pragma unroll
        for (int o = 0; o < NOP; ++o) {
          const double next = tile[o * TS + j];
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (decltype(all_rows)::value || k < rows) r[o][k] = horner_step(r[o][k], ls[o], M[k]);
          ls[o] = next;
        }
      }
    };
    if (rows == K) steps(std::true_type{});
    else steps(std::false_type{});
  }
  // The last limb, then once to signed reduced (|s + L| < m/2 + 2^13 + 2^32).
#pragma unroll
  for (int o = 0; o < NOP; ++o)
#pragma unroll
    for (int k = 0; k < K; ++k) r[o][k] = reduce(r[o][k] + ls[o] * 0x1p-32, M[k]);
}

// Per-pair bookkeeping, one lane per pair in wave 0 (getGcdKernel.cu:350-352,
// 389-402 and 417-441).
struct PairState {
  int32_t live, status, remaining, ubits, vbits, pval;
  uint32_t niter, ndig;
  uint32_t trace_cap;  // the pair's trace slots, kept here so that the
  uint64_t trace_off;  // per-iteration path loads nothing from its descriptor
};

// Everything one workgroup does for part C.g of a group of up to B pairs,
// phase by phase: (a) residues, (b) elimination rounds, (c) replay and
// recovery, run in order by process_group.  S: the group's small LDS fields;
// U: the staging area (and, during (b), the moduli table).
template <int T, int K, int B, class A>
struct GroupRun {
  static constexpr int NW = T / 64;
  static constexpr int SH = 6 - ilog2(B);  // lane l of a reduce-scatter holds pair l >> SH
  using SmallT = Small<T, K, B>;
  using BigT = Big<T, K, B>;

  const KernelArgs& args;
  const GroupDesc Gd;
  SmallT& S;
  BigT& U;
  Ctx C;  // this work item's exchange state, in scalar registers
  PhaseClock& clk;
  // The kernel's arrays as global pointers in scalar registers (read once
  // per work item): global, not flat, loads and stores with a scalar base.
  const GPtr<const uint32_t> limbs;
  const GPtr<const PairDesc> pairs;
  const GPtr<const uint32_t> mod_u32;
  const GPtr<const double> mod_d, mod_inv;
  const GPtr<uint32_t> res;
  const GPtr<mgcd_pair_t> trace, digits;
  const GPtr<PairOut> out;
  // The two arrays the per-round bookkeeping reads and writes, in vector
  // registers in one-pair tilings (short of scalar registers there).
  const GPtr<const uint32_t> mod_u32_r;
  const GPtr<mgcd_pair_t> trace_r;
  const uint32_t tid, lane, wave, nq, npairs;
  // Part g owns moduli [g*nq/G, (g+1)*nq/G): sizes differ by at most one,
  // none is empty (G <= nq), none exceeds T*K (checked by the host).
  const uint32_t base, cnt;
  const int rows;
  const bool lead;
  const int L;
  bool lds_mods = false;
  ModD M[K];
  double ck[K];
  double t[B][K];
  double wkey = KNONE;
  uint32_t infw = 0;  // bit k: some lane of this wave holds TINF in slot row k (uniform)
  PairState ps;

  __device__ __forceinline__ GroupRun(const KernelArgs& a, const GroupDesc& gd, SmallT& s, BigT& u, Ctx& c, PhaseClock& ck_)
      : args(a), Gd(gd), S(s), U(u), C(uniform_ctx(c)), clk(ck_), limbs(gptr(a.limbs)), pairs(gptr(a.pairs)),
        mod_u32(gptr(a.mod_u32)), mod_d(gptr(a.mod_d)), mod_inv(gptr(a.mod_inv)), res(gptr(a.res)),
        trace(gptr(a.trace)), digits(gptr(a.digits)), out(gptr(a.out)),
        mod_u32_r(B == 1 ? vptr(gptr(a.mod_u32)) : gptr(a.mod_u32)),
        trace_r(B == 1 ? vptr(gptr(a.trace)) : gptr(a.trace)), tid(threadIdx.x), lane(threadIdx.x & 63),
        wave(threadIdx.x >> 6), nq(gd.nq), npairs(gd.count),
        base((uint32_t)(((uint64_t)c.g * gd.nq) / c.G)),
        cnt((uint32_t)(((uint64_t)(c.g + 1) * gd.nq) / c.G) - (uint32_t)(((uint64_t)c.g * gd.nq) / c.G)),
        rows((int)(((uint32_t)(((uint64_t)(c.g + 1) * gd.nq) / c.G) - (uint32_t)(((uint64_t)c.g * gd.nq) / c.G) + T - 1) / T)),
        lead(c.g == 0), L(a.L) {}

  // ---- (a) residues of u and v, stored canonical for (c) -------------------
  // The residues of part p (PP pairs, 2 PP operands, at a time).
  __device__ __forceinline__ void residues_of_part(uint32_t p) {
    const uint32_t pb = (uint32_t)(((uint64_t)p * nq) / C.G);
    const uint32_t pc = (uint32_t)(((uint64_t)(p + 1) * nq) / C.G) - pb;
    ModD PM[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t j = (uint32_t)k * T + tid;
      PM[k].m = j < pc ? mod_d[pb + j] : 3.0;
      PM[k].minv = j < pc ? mod_inv[pb + j] : 1.0 / 3.0;
    }
    constexpr int PP = (B >= 2 && K <= 2) ? 2 : 1;
    for (uint32_t b0 = 0; b0 < npairs; b0 += PP) {
      GPtr<const uint32_t> x[2 * PP];
      uint32_t nx[2 * PP];
      uint64_t ro[2 * PP];
#pragma unroll
      for (int q = 0; q < PP; ++q) {
        const bool here = b0 + q < npairs;
        const PairDesc P = pairs[Gd.first + (here ? b0 + q : b0)];
        x[2 * q] = limbs + P.u_off;
        x[2 * q + 1] = limbs + P.v_off;
        nx[2 * q] = here ? P.nu : 0u;
        nx[2 * q + 1] = here ? P.nv : 0u;
        ro[2 * q] = P.res_off;
        ro[2 * q + 1] = P.res_off + nq;
      }
      double r[2 * PP][K];
      residues<T, K, 2 * PP, BigT::TL>(x, nx, r, PM, (int)((pc + T - 1) / T), U.u.st.tile);
#pragma unroll
      for (int o = 0; o < 2 * PP; ++o) {
        if (b0 + o / 2 >= npairs) break;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if ((uint32_t)k * T + tid < pc) res[ro[o] + pb + k * T + tid] = (uint32_t)canon(r[o][k], PM[k]);
      }
    }
  }

  __device__ __forceinline__ void residues_phase() {
    if (npairs == 0) return;
    if (C.G == 1) {
      residues_of_part(0);
      return;
    }
    // Every part's residue job has a state word after the group's exchange
    // slots: 0 free, 1 taken, 2 done.  A part takes its own job, or waits
    // for the part that took it (its results are the same either way, so
    // after a timed-out wait it computes them itself).  Then, instead of
    // waiting at its first exchange for the parts not yet started, it takes
    // their jobs, which they will find done.
    unsigned long long* job = C.bar + (size_t)2 * (C.G + 1) * B;
    auto finish_job = [&](uint32_t p) {
      __syncthreads();
      if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(job + p, 2ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    };
    if (tid == 0) {
      unsigned long long z = 0;
      const bool mine = __hip_atomic_compare_exchange_strong(job + C.g, &z, 1ull, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t st = mine ? 1u : 0u;
      if (!mine) {
        uint32_t spins = 0;
        while (__hip_atomic_load(job + C.g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 2ull) {
          if (++spins > C.spin_limit) {
            st = 1u;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      S.flag = st;
    }
    __syncthreads();
    if (S.flag) {
      residues_of_part(C.g);
      finish_job(C.g);
    } else {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    for (;;) {
      __syncthreads();
      if (wave == 0) {
        uint32_t take = ~0u;
        for (uint32_t p0 = 0; p0 < C.G && take == ~0u; p0 += 64) {
          const uint32_t p = p0 + lane;
          const bool free_job =
              p < C.G && p != C.g && __hip_atomic_load(job + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull;
          uint64_t m = __ballot(free_job);
          while (m && take == ~0u) {
            const uint32_t q = p0 + (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            uint32_t got = ~0u;
            if (lane == 0) {
              unsigned long long z = 0;
              if (__hip_atomic_compare_exchange_strong(job + q, &z, 1ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT))
                got = q;
            }
            take = __shfl(got, 0);
          }
        }
        if (lane == 0) S.item = take;
      }
      __syncthreads();
      const uint32_t p = S.item;
      if (p == ~0u) break;
      residues_of_part(p);
      finish_job(p);
    }
  }

  // ---- (b) elimination -------------------------------------------------------
  __device__ __forceinline__ void elim_setup() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t j = (uint32_t)k * T + tid;
      if (j < cnt) {
        M[k].m = mod_d[base + j];
        M[k].minv = mod_inv[base + j];
      } else {
        // No modulus here: the lane's values are all sentinels, which enter
        // the elimination's products as they are, so it takes a table
        // modulus (a small one would make such a product 0 often).
        M[k].m = mod_d[0];
        M[k].minv = mod_inv[0];
      }
      ck[k] = 0.5 + 2.0 * (double)(IDX_MAX - (base + j));
    }
    lds_mods = nq <= BigT::MODS;
  }

  // The moduli table of the group in LDS, over (a)'s staging tile.
  __device__ __forceinline__ void load_mods() {
    __syncthreads();
    if (lds_mods)
      for (uint32_t i = tid; i < nq; i += T) U.u.mods[i] = mod_u32[i];
  }

  // Initial t = u/v (getGcdKernel.cu:369), one batched inverse per slot.
  // Keys are reduced over the wave slot by slot (reduce-scatter), so only one
  // slot's B keys are ever held: wkey in lane l is the wave's minimum for
  // pair l >> SH.  Also the per-pair state (lane b of wave 0) and the retire
  // lists.
  __device__ __forceinline__ void elim_init() {
    wkey = KNONE;
    infw = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t j = (uint32_t)k * T + tid;
      const bool here = j < cnt;
      double P[B], key[B];
      bool anyinf = false;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        double v = 0.0;
        if (here && (uint32_t)b < npairs) v = (double)res[pairs[Gd.first + b].res_off + nq + base + j];
        const double y = v != 0.0 ? v : 1.0;
        P[b] = b == 0 ? y : A::mulmod(P[b - 1], y, M[k]);
      }
      double r = A::modinv(P[B - 1], M[k]);
#pragma unroll
      for (int b = B - 1; b >= 0; --b) {
        double u = 0.0, v = 0.0;
        if (here && (uint32_t)b < npairs) {
          const uint64_t ro = pairs[Gd.first + b].res_off + base + j;
          u = (double)res[ro];
          v = (double)res[ro + nq];
        }
        const double inv = b == 0 ? r : A::mulmod(r, P[b - 1], M[k]);
        if (b > 0) r = A::mulmod(r, v != 0.0 ? v : 1.0, M[k]);
        double tv = v != 0.0 ? tsigned(A::mulmod(u, inv, M[k]), M[k]) : (u != 0.0 ? TINF : TZERO);
        if (!here) tv = TRET;
        anyinf |= tv == TINF;
        t[b][k] = tv;
        key[b] = elim_key(tv, ck[k]);
      }
      if (__ballot(anyinf)) infw |= 1u << k;
      wkey = __builtin_fmin(wkey, wave_min_scatter<B>(key));
    }

    ps.live = 0;
    ps.status = KST_NOTRUN;
    ps.remaining = ps.ubits = ps.vbits = ps.pval = 0;
    ps.niter = ps.ndig = 0;
    ps.trace_cap = 0;
    ps.trace_off = 0;
    if (wave == 0 && lane < npairs) {
      const PairDesc Pl = pairs[Gd.first + lane];
      ps.live = 1;
      ps.status = KST_OK;
      ps.remaining = (int32_t)nq;                   // totalModuliRemaining (:350)
      ps.ubits = (int32_t)((Pl.nu + 1) * 32);       // :351
      ps.vbits = (int32_t)((Pl.nv + 1) * 32);       // :352
      ps.trace_cap = Pl.trace_cap;
      ps.trace_off = Pl.trace_off;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) S.retire[k * T + tid] = 0;
    if (tid < (uint32_t)(K * NW)) (&S.retw[0][0])[tid] = 0;
  }

  // This wave's minima into LDS (before the barrier that ends a round's compute).
  __device__ __forceinline__ void stage_keys() {
    clk.lap(PH_COMPUTE);
    clk.round();
    if ((lane & ((1u << SH) - 1)) == 0) S.red[lane >> SH][wave] = wkey;
  }
  // After that barrier, wave 0: the workgroup's minimum of pair lane % B... in lane b < B.
  __device__ __forceinline__ double workgroup_keys() {
    double x = KNONE;
    if (lane < (uint32_t)B) {
      x = S.red[lane][0];
#pragma unroll
      for (int w2 = 1; w2 < NW; ++w2) x = __builtin_fmin(x, S.red[lane][w2]);
      x = __builtin_fmin(x, KINF);
    }
    return x;
  }
  // Reduces the wave minima over the group; wave 0 lane b < B gets pair b's.
  __device__ __forceinline__ double reduce_keys() {
    stage_keys();
    __syncthreads();
    clk.lap(PH_BARRIER);
    double x = KNONE;
    if (wave == 0) {
      // The reduction, the exchange and the bookkeeping after it are the
      // serial part of a round: in group tilings wave 0 runs them ahead of
      // the waves of any other workgroup on the CU.  (One-pair tilings,
      // whose exchange dominates the round, measured slower with it.)
      if constexpr (B > 1) set_prio(3);
      x = workgroup_keys();
      if (C.G > 1) x = grid_min<B>(x, C);
    }
    clk.lap(PH_EXCHANGE);
    return x;
  }

  // Marks pair b's pivot modulus for retirement if it lies in this part.
  __device__ __forceinline__ void mark_retire(uint32_t b, uint32_t idx) {
    const uint32_t off = idx - base;
    if (off < cnt) {
      atomicOr(&S.retire[off], 1u << b);
      atomicOr(&S.retw[off / T][(off % T) >> 6], 1u << b);
    }
  }

  // Records pivot (idx, val) as iteration niter of the pair: trace entry
  // (leader workgroup), and the next round's pivot in LDS.
  __device__ __forceinline__ void publish_pivot(uint32_t b, uint32_t idx, int32_t val) {
    // Address spaces explicit: an LDS read, or a global one (the group's
    // table did not fit LDS), never a flat load through a merged pointer.
    uint32_t mu;
    if (lds_mods) mu = lds_ptr(U.u.mods)[idx];
    else mu = mod_u32_r[idx];
    S.bb[b] = (double)val;
    S.pp[b] = (double)mu;
    S.pidx[b] = idx;
    // The trace store goes last: waits for mu must not include it (stores
    // count in the same counter as global loads).
    asm volatile("" ::: "memory");
    if (lead && ps.niter < ps.trace_cap) {
      mgcd_pair_t e;
      e.modulus = mu;
      e.value = val;
      trace_r[ps.trace_off + ps.niter] = e;
    }
    ps.niter++;
    ps.pval = val;
  }

  // Wave 0, after a round's minima kr (lane b: pair b's) are known: the
  // first round's pivot (the loop body runs once even for an infinite one,
  // do-while :374-402) or a later one with the size bookkeeping.
  __device__ __forceinline__ void book(double kr, bool first) {
    if (lane < npairs && ps.live) {
      bool take = true;
      if (!first) {
        ps.remaining -= 1;                                          // :390
        const int32_t tbits = ps.ubits - (L - 1) + ffs_abs(ps.pval);  // :391
        ps.ubits = ps.vbits;
        ps.vbits = tbits;
        if (ps.remaining * (L - 2) <= ps.ubits) {                   // :393-399
          ps.status = KST_REDUX;
          ps.live = 0;
          take = false;
        }
      }
      if (take) {
        uint32_t idx;
        int32_t val;
        // A malformed key (only ever after a peer gave up) ends the pair
        // like a timeout.
        const int kind = kr == KNONE ? KEY_BAD : decode_pivot(kr, nq, idx, val);
        if (kind == KEY_BAD) {
          ps.status = KST_TIMEOUT;
          ps.live = 0;
        } else if (kind == KEY_INF && !first) {
          ps.live = 0;  // infinity: done (:402)
        } else {
          publish_pivot(lane, idx, val);
          mark_retire(lane, idx);
        }
      }
    }
    if (lane < (uint32_t)B && !ps.live) {
      // A pair that is done (or absent) keeps being updated with t' =
      // 0 / (t - 1): its finite t settle at 0, whose y = -1 is never 0
      // (mod m), so it sends no wave through the full forward pass.
      S.bb[lane] = 1.0;
      S.pp[lane] = 0.0;
      S.pidx[lane] = NONE_IDX;
    }
    const bool any = __ballot(ps.live != 0) != 0;
    if (lane == 0) S.flag = any;
  }

  // One round's arithmetic on every slot row: retire the pivots' moduli,
  // then t' = p' / (t - b) for every pair (:376-385); leaves the wave's
  // minima in wkey.
  __device__ __forceinline__ void compute_round() {
    // Retire each pair's pivot modulus (:376-377).  Few waves hold one; a
    // wave that does visits only the pairs named in its summary word
    // (uniform), with one select per pair.
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint32_t sw = __builtin_amdgcn_readfirstlane(S.retw[k][wave]);
      if (sw) {
        const uint32_t rw = S.retire[k * T + tid];
        S.retire[k * T + tid] = 0;
        if (lane == 0) S.retw[k][wave] = 0;
        while (sw) {
          const uint32_t b = (uint32_t)__builtin_ctz(sw);
          sw &= sw - 1;
          // A uniform branch per candidate pair (the empty asm keeps the
          // compiler from turning the B branches into B selects).
#pragma unroll
          for (int c = 0; c < B; ++c)
            if ((uint32_t)c == b) {
              asm volatile("" ::: "memory");
              if ((rw >> c) & 1) t[c][k] = TRET;
            }
        }
      }
    }
    // Prefix products of the values to invert, one inverse, then back down
    // (Montgomery's batch inversion).
    wkey = KNONE;
    LdsPtr<double> bbp = lds_ptr(S.bb), ppp = lds_ptr(S.pp);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= rows) break;
      // Forward: prefix products of the values to invert, y = t - b.  Where
      // mulmod takes operands up to 2^33, every lane's y enters as it is,
      // also where t is a sentinel (an integer below 2^33 like the others;
      // that lane's result is replaced below); else y = 1 there.  A y == 0
      // (mod m) makes the product 0 (mod m); that is rare (about one value
      // in m), so only then does the wave redo the pass with y = 1 wherever
      // t is a sentinel or t == b (mod m), and marks the latter TPOLE.
      double P[B], Y[B], key[B];
      bool pole_any = false;
      auto forward = [&](auto full) {
        // Every y first: the pivot values' LDS reads all issue at the row's
        // start (each address depends only on its own t, so the compiler
        // cannot keep the values across rows), and the chain of products
        // below then runs without waiting on LDS.
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const double tv = t[b][k];
          const double d = tv - load_after(bbp, b, tv);
          double y = d;
          if constexpr (decltype(full)::value) {
            const bool fin = __builtin_fabs(tv) < TFIN;
            const bool pole = fin & ((d == 0.0) | (__builtin_fabs(d) == M[k].m));
            if (pole) t[b][k] = TPOLE;
            pole_any |= pole;
            y = (fin & !pole) ? d : 1.0;
          } else if constexpr (!A::kWideOperands) {
            y = __builtin_fabs(tv) < TFIN ? d : 1.0;
          }
          Y[b] = y;
        }
#pragma unroll
        for (int b = 0; b < B; ++b) P[b] = b == 0 ? Y[0] : A::mulmod(P[b - 1], Y[b], M[k]);
        if constexpr (B == 1) P[0] = reduce(P[0], M[k]);
      };
      // Issue priority falls as a wave moves through the round's phases
      // (forward pass and first half of the backward pass, then its second
      // half; row by row): a wave that has fallen behind the other wave on
      // its SIMD gets the issue slots until it catches up, so the two finish
      // together instead of one running alone at the end.
      if constexpr (K == 2 && B >= 4) {
        if (k == 0) set_prio(3);
      } else {
        set_prio(3 - 2 * k);
      }
      forward(std::false_type{});
      // |P[B-1]| < m/2 + 2^15 (a product, or for B == 1 the reduced y), so it
      // is 0 (mod m) only as 0; the inverse takes it signed.
      const bool redo = __ballot(P[B - 1] == 0.0) != 0;
      if (redo) forward(std::true_type{});
      const bool poled = redo && __ballot(pole_any) != 0;
      double r = A::modinv(P[B - 1], M[k]);
      // Backward: t' = p' / y where t is finite.  A sentinel stays, except
      // that TINF gives 0 and TPOLE gives TINF; all of them have a zero low
      // word, so those two take one high-word select, in the rare rounds
      // (uniform per wave and row) where the row holds one of them.
      auto backward = [&](auto special) {
        double ppn = load_after(ppp, B - 1, r);
#pragma unroll
        for (int b = B - 1; b >= 0; --b) {
          // The priority steps down halfway through the backward pass: the
          // forward pass and the inverse are the shorter part of a row.
          // (Two rows: levels 3 | 2 from halfway through the first backward
          // pass | 1 from the second | 0 for its last quarter, which keeps
          // the final, equal-priority stretch short.)
          if constexpr (K == 2 && B >= 4) {
            if (k == 0 && b == B / 2 - 1) set_prio(2);
            if (k == 1 && b == B - 1) set_prio(1);
            if (k == 1 && b == B / 4 - 1) set_prio(0);
          } else {
            if (b == B / 2 - 1 || (B == 1 && b == 0)) set_prio(2 - 2 * k);
          }
          const double tv = t[b][k];
          const double ppv = ppn;
          if (b > 0) ppn = load_after(ppp, b - 1, r);  // read a step ahead of its use
          const double inv = b == 0 ? r : A::mulmod(r, P[b - 1], M[k]);
          if (b > 0) r = A::mulmod(r, Y[b], M[k]);
          const double tn = tsigned(A::mulmod(ppv, inv, M[k]), M[k]);
          const bool fin = __builtin_fabs(tv) < TFIN;
          double tv2;
          if constexpr (decltype(special)::value) {
            const uint32_t alt_hi = tv == TINF ? 0u : hi32(tv);
            tv2 = mk64(fin ? lo32(tn) : 0u, fin ? hi32(tn) : alt_hi);
          } else {
            tv2 = fin ? tn : tv;
          }
          t[b][k] = tv2;
          key[b] = elim_key(tv2, ck[k]);
        }
      };
      // (One-pair tilings, whose round is mostly the exchange, measured
      // slower with the two versions: they keep the general one.)
      if (B == 1 || (((infw >> k) & 1) | poled)) backward(std::true_type{});
      else backward(std::false_type{});
      // TINF lanes become 0 this round; the poles' TINF stay one round.
      infw = (infw & ~(1u << k)) | ((uint32_t)poled << k);
      wkey = __builtin_fmin(wkey, wave_min_scatter<B>(key));
    }
    __builtin_amdgcn_s_setprio(0);
  }

  // The whole elimination of one group, round after round.
  __device__ __forceinline__ void elimination() {
    elim_init();
    {
      const double kr = reduce_keys();
      if (wave == 0) {
        book(kr, true);
        if constexpr (B > 1) set_prio(0);
      }
      clk.lap(PH_BOOK);
      __syncthreads();
      clk.lap(PH_BARRIER2);
    }
    while (S.flag) {
      compute_round();
      const double kr = reduce_keys();
      if (wave == 0) {
        book(kr, false);
        if constexpr (B > 1) set_prio(0);
      }
      clk.lap(PH_BOOK);
      __syncthreads();
      clk.lap(PH_BARRIER2);
    }
  }

  // Surviving moduli, status and iteration counts into LDS; the pivots the
  // leader workgroup recorded become visible to all parts.
  __device__ __forceinline__ void finish_elimination() {
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t w = __ballot(t[b][k] != TRET);
        if (lane == 0) S.actw[b][k][wave] = w;
      }
    if (wave == 0 && lane < (uint32_t)B) {
      S.status[lane] = lane < npairs ? ps.status : KST_NOTRUN;
      S.ntrace[lane] = ps.niter;
    }
    C.phase = 1;
    if (C.G > 1 && wave == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      (void)grid_min<B>(0.0, C);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    clk.lap(PH_TAIL);
  }

  // ---- (c1) replay of the pivots on the survivors ---------------------------
  // The survivors of every pair of the group form one compacted list
  // (pair-major; entry = pair << 11 | slot offset); the pivots of all pairs
  // are staged through LDS together, RC steps at a time.  Group tilings
  // (QUADS) pad each pair's run to a multiple of four entries and give a
  // lane four consecutive entries, all of one pair, so that one LDS read of
  // a step's pivot serves all four; one-pair tilings give lane t entries
  // t, t + T, ... (up to K).
  static constexpr bool QUADS = B >= 4;
  static constexpr uint16_t PAD_ENTRY = 0xFFFF;  // a padding entry (pair 31, offset 2047: never a survivor)
  static_assert(B < 32 || T * K < 2048, "padding entry collides with a survivor");
  __device__ __forceinline__ void replay() {
    uint32_t nsurv = 0, maxR = 0;
    for (uint32_t b = 0; b < npairs; ++b) {
      if (S.status[b] != KST_OK) continue;
      maxR = max(maxR, S.ntrace[b]);
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const uint64_t word = S.actw[b][k][w];
          if ((uint32_t)w == wave && ((word >> lane) & 1))
            U.u.st.surv[nsurv + __builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0))] =
                (uint16_t)((b << 11) | (uint32_t)(k * T + tid));
          nsurv += (uint32_t)__builtin_popcountll(word);
        }
      if constexpr (QUADS) {
        const uint32_t pad = (0u - nsurv) & 3u;
        if (tid < pad) U.u.st.surv[nsurv + tid] = PAD_ENTRY;
        nsurv += pad;
      }
    }
    __syncthreads();
    constexpr int E = QUADS ? 4 : K;  // entries per lane per round
    constexpr int RC = BigT::TL / (2 * B);  // pivots per pair per staging
    for (uint32_t e0 = 0; e0 < nsurv; e0 += E * T) {
      ModD RM[E];
      double ra[E], rc[E], rs[E];
      uint32_t eb[E], eR[E];
      uint64_t ro[E];
      bool ok[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t j = QUADS ? e0 + 4 * tid + (uint32_t)e : e0 + (uint32_t)e * T + tid;
        RM[e].m = 3.0;
        RM[e].minv = 1.0 / 3.0;
        ra[e] = rc[e] = 0.0;
        rs[e] = 1.0;
        ro[e] = 0;
        eb[e] = 0;
        eR[e] = 0;
        ok[e] = false;
        const uint32_t ent = j < nsurv ? (uint32_t)U.u.st.surv[j] : (uint32_t)PAD_ENTRY;
        if (ent != PAD_ENTRY) {
          const uint32_t idx = base + (ent & 0x7FFu);
          ok[e] = true;
          eb[e] = ent >> 11;
          eR[e] = S.ntrace[eb[e]];
          RM[e].m = mod_d[idx];
          RM[e].minv = mod_inv[idx];
          ro[e] = pairs[Gd.first + eb[e]].res_off + idx;
          ra[e] = (double)res[ro[e]];
          rc[e] = (double)res[ro[e] + nq];
        }
      }
      // Entries in use by this wave (uniform; in a scalar register): with
      // QUADS all four or none (a quad's first entry is never padding).
      int ew = 0;
      if constexpr (QUADS) {
        if (e0 + 256 * wave < nsurv) ew = E;
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e)
          if (e0 + (uint32_t)e * T + wave * 64 < nsurv) ew = e + 1;
      }
      ew = __builtin_amdgcn_readfirstlane(ew);
      for (uint32_t c0 = 0; c0 < maxR; c0 += RC) {
        const uint32_t len = min((uint32_t)RC, maxR - c0);
        __syncthreads();
        for (uint32_t j = tid; j < len * npairs; j += T) {
          const uint32_t b = j / len, jj = j - b * len;
          double bv = 0.0, pv = 1.0;
          if (S.status[b] == KST_OK && c0 + jj < S.ntrace[b]) {
            const mgcd_pair_t tr = trace[pairs[Gd.first + b].trace_off + c0 + jj];
            bv = (double)tr.value;
            pv = (double)tr.modulus;
          }
          U.u.st.tile[2 * (b * RC + jj)] = bv;
          U.u.st.tile[2 * (b * RC + jj) + 1] = pv;
        }
        __syncthreads();
        // Unrolled by eight in group tilings: the next pivots' LDS reads
        // issue under the current ones' multiplications (the one-pair tiling
        // measured slower unrolled).  Every entry takes every step: past its
        // pair's last pivot the staged step is (b, p') = (0, 1), which swaps
        // a and c exactly (c' = a - 0 c, a' = c, s' = s), so no per-step test
        // is needed; the parity is settled once, below.  The entries in use
        // (uniform per wave) select one of E versions of the loop.
        auto steps = [&](auto ne) {
          constexpr int NE = decltype(ne)::value;
          constexpr int RU = B > 1 ? 8 : 1;
#pragma unroll RU
          for (uint32_t j = 0; j < len; ++j) {
            double bq = 0.0, pq = 1.0;  // QUADS: the lane's pair's step, read once
            if constexpr (QUADS) {
              bq = U.u.st.tile[2 * (eb[0] * RC + j)];
              pq = U.u.st.tile[2 * (eb[0] * RC + j) + 1];
            }
#pragma unroll
            for (int e = 0; e < NE; ++e) {
              const double bv = QUADS ? bq : U.u.st.tile[2 * (eb[e] * RC + j)];
              const double pv = QUADS ? pq : U.u.st.tile[2 * (eb[e] * RC + j) + 1];
              const double c2 = ra[e] - A::mulmod(bv, rc[e], RM[e]);  // |c2| <= m + 2^14
              ra[e] = A::mulmod(rc[e], pv, RM[e]);
              rs[e] = A::mulmod(rs[e], pv, RM[e]);
              rc[e] = c2;
            }
          }
        };
        if (ew == E) steps(std::integral_constant<int, E>{});
        else if constexpr (E > 1 && !QUADS) {
          if (ew == 1) steps(std::integral_constant<int, 1>{});
          else if constexpr (E > 2) {
            if (ew == 2) steps(std::integral_constant<int, 2>{});
            else if constexpr (E > 3) {
              if (ew == 3) steps(std::integral_constant<int, 3>{});
            }
          }
        }
      }
      // u = a/s on the survivors, kept as (a, s) in place of (u, v).  After
      // an odd number of padding swaps a sits in c (reduced: a swap moves
      // the reduced a there).
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (ok[e]) {
          const double af = ((maxR - eR[e]) & 1u) ? rc[e] : ra[e];
          res[ro[e]] = (uint32_t)canon(af, RM[e]);
          res[ro[e] + nq] = (uint32_t)canon(rs[e], RM[e]);
        }
      }
    }
    __syncthreads();
    clk.lap(PH_REPLAY);
  }

  // ---- (c2) recovery, RB pairs at a time in lockstep (:404-433) -------------
  // u = a/s per surviving modulus; a lane inverts s only for a digit it owns.
  // Chunk c holds pairs c*RB + j; wave 0 lane c*RB + j keeps pair c*RB + j's
  // bookkeeping and gets its minima from the slots of class j.
  __device__ __forceinline__ void recovery() {
    constexpr int RB = B < 8 ? B : 8;
    if (wave == 0 && lane < npairs) ps.live = ps.status == KST_OK;
    for (uint32_t c0 = 0; c0 < npairs; c0 += RB) {
      double ua[RB][K], us[RB][K];
      uint32_t am[RB];  // surviving slots (bits)
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const uint32_t b = c0 + j;
        am[j] = 0;
        const bool run = b < npairs && S.status[b] == KST_OK;
        const uint64_t ro = run ? pairs[Gd.first + b].res_off : 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          ua[j][k] = 0.0;
          us[j][k] = 1.0;
          if (run && ((S.actw[b][k][wave] >> lane) & 1)) {
            am[j] |= 1u << k;
            const uint32_t idx = base + (uint32_t)k * T + tid;
            ua[j][k] = (double)res[ro + idx];
            us[j][k] = (double)res[ro + nq + idx];
          }
        }
      }
      const bool mine = wave == 0 && lane >= c0 && lane < c0 + RB;  // this lane's pair is in the chunk

      // Selects each pair's next digit: the lowest index holding a nonzero u;
      // wave 0 lane c0 + j gets (index, value) of pair c0 + j as a recovery key.
      auto select = [&]() -> double {
        constexpr int SHR = 6 - ilog2(RB);
        double cand[RB];
#pragma unroll
        for (int j = 0; j < RB; ++j) {
          cand[j] = (double)NONE_IDX;
#pragma unroll
          for (int k = K - 1; k >= 0; --k)
            if (((am[j] >> k) & 1) && ua[j][k] != 0.0) cand[j] = (double)(base + (uint32_t)k * T + tid);
        }
        const double w = wave_min_scatter<RB>(cand);
        if ((lane & ((1u << SHR) - 1)) == 0) S.red[lane >> SHR][wave] = w;
        __syncthreads();
        // Owners compute toSigned(u) (:424); one batched inverse per slot.
        uint32_t own[K];  // bit j: this slot holds pair c0 + j's candidate
#pragma unroll
        for (int k = 0; k < K; ++k) own[k] = 0;
#pragma unroll
        for (int j = 0; j < RB; ++j) {
          double x = S.red[j][0];
#pragma unroll
          for (int w2 = 1; w2 < NW; ++w2) x = __builtin_fmin(x, S.red[j][w2]);
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (((am[j] >> k) & 1) && (double)(base + (uint32_t)k * T + tid) == x) own[k] |= 1u << j;
          // No nonzero residue in this part: (NONE_IDX, 0), which the reference
          // resolves to its thread 0 pair with value 0 (:57-64).
          if ((uint32_t)j == tid && x == (double)NONE_IDX) S.sel[j] = KNODIGIT;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (__ballot(own[k] != 0) == 0) continue;
          const double idx = (double)(base + (uint32_t)k * T + tid);
          double P[RB];
#pragma unroll
          for (int j = 0; j < RB; ++j) {
            const double y = ((own[k] >> j) & 1) ? us[j][k] : 1.0;
            P[j] = j == 0 ? y : A::mulmod(P[j - 1], y, M[k]);
          }
          double r = A::modinv(P[RB - 1], M[k]);
#pragma unroll
          for (int j = RB - 1; j >= 0; --j) {
            const bool owns = (own[k] >> j) & 1;
            const double inv = j == 0 ? r : A::mulmod(r, P[j - 1], M[k]);
            if (j > 0) r = A::mulmod(r, owns ? us[j][k] : 1.0, M[k]);
            if (owns) {
              const double v = tsigned(A::mulmod(ua[j][k], inv, M[k]), M[k]);
              S.sel[j] = idx * 0x1p32 + (v + 0x1p31);
            }
          }
        }
        __syncthreads();
        double x = KNONE;
        if (wave == 0) {
          x = S.sel[lane % RB];
          if (C.G > 1) x = grid_min<RB>(x, C);
        }
        return x;
      };

      // Publishes digit (idx, val) of pair b (this lane's) as the next update.
      auto set_digit = [&](uint32_t idx, int32_t val) {
        S.bb[lane - c0] = (double)val;
        S.pp[lane - c0] = mod_d[idx];
        S.pidx[lane - c0] = idx;
        ps.pval = val;
      };
      // Pairs not (or no longer) in recovery publish an update that matches
      // no modulus.
      auto finish_round = [&]() {
        if (wave == 0) {
          const uint32_t b = c0 + lane % RB;
          const bool running = __shfl(ps.live, (int)(b & 63)) != 0 && b < npairs;
          if (lane < (uint32_t)RB) {
            if (!running) {
              S.bb[lane] = 0.0;
              S.pp[lane] = 1.0;
              S.pidx[lane] = NONE_IDX;
            }
          }
          const bool any = __ballot(mine && ps.live != 0) != 0;
          if (lane == 0) S.flag = any;
        }
        __syncthreads();
      };

      {
        const double x = select();
        if (mine && lane < npairs && ps.live) {
          uint32_t idx;
          int32_t val;
          if (x == KNONE || decode_digit(x, nq, idx, val) == KEY_BAD) {
            ps.status = KST_TIMEOUT;
            ps.live = 0;
          } else {
            set_digit(idx, val);
          }
        }
        finish_round();
      }
      while (S.flag) {
        // u' = (u - g) / p' on every other surviving modulus (:417-425).
#pragma unroll
        for (int j = 0; j < RB; ++j) {
          const uint32_t pidx = __builtin_amdgcn_readfirstlane(S.pidx[j]);
          const double gv = S.bb[j], pv = S.pp[j];
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if ((am[j] >> k) & 1) {
              if (base + (uint32_t)k * T + tid == pidx) {
                am[j] &= ~(1u << k);
              } else {
                ua[j][k] = reduce(ua[j][k] - A::mulmod(gv, us[j][k], M[k]), M[k]);
                us[j][k] = A::mulmod(us[j][k], pv, M[k]);
              }
            }
          }
        }
        const double x = select();
        if (mine && lane < npairs && ps.live) {
          if (lead) {
            mgcd_pair_t dp;
            dp.modulus = mod_u32[S.pidx[lane - c0]];
            dp.value = ps.pval;
            digits[pairs[Gd.first + lane].dig_off + ps.ndig] = dp;
          }
          ps.ndig++;
          ps.remaining -= 1;                                       // :428
          if (ps.remaining <= 0) {                                 // :429-430, :440
            ps.status = ps.pval != 0 ? KST_RECOVERY : KST_OK;
            ps.live = 0;
          } else {
            uint32_t idx;
            int32_t val;
            if (x == KNONE || decode_digit(x, nq, idx, val) == KEY_BAD) {
              ps.status = KST_TIMEOUT;
              ps.live = 0;
            } else if (val == 0) {
              ps.live = 0;  // :433
            } else {
              set_digit(idx, val);
            }
          }
        }
        finish_round();
      }
    }

    if (lead && wave == 0 && lane < npairs) {
      PairOut o;
      o.status = ps.status;
      o.ndig = ps.ndig;
      o.ntrace = ps.niter;
      o.pad = 0;
      out[Gd.first + lane] = o;
    }
    clk.lap(PH_RECOVERY);
  }
};

// Everything one workgroup does for part C.g of a group of up to B pairs.
template <int T, int K, int B, class A>
__device__ void process_group(const KernelArgs& args, const GroupDesc& Gd, Shared<T, K, B>& S, Ctx& C) {
  PhaseClock clk;
  clk.start();
  GroupRun<T, K, B, A> g(args, Gd, S.s, S.b, C, clk);
  g.residues_phase();
  clk.lap(PH_RESIDUES);
  g.elim_setup();
  g.load_mods();
  g.elimination();
  g.finish_elimination();
  g.replay();
  g.recovery();
#ifdef MGCD_PHASE_CLOCKS
  clk.acc_extra(PH_FIRST_LOAD, g.C.first_load);
  clk.acc_extra(PH_COLLECT_READY, g.C.collect_ready);
  clk.acc_extra(PH_COLLECTS, g.C.collects);
#endif
  clk.flush(args.clocks, threadIdx.x == 0);
  clk.flush_waves(args.clocks);
}

// Next work item for this workgroup (thread 0): queue 0, then the queue of
// the XCD the workgroup runs on, then the others; exhausted queues are
// remembered in *done.
__device__ uint32_t next_item(const KernelArgs& args, uint32_t own, uint32_t* done) {
  for (uint32_t s = 0; s < args.nqueues; ++s) {
    const uint32_t q = s == 0 ? 0u : 1u + (own + s - 1) % (args.nqueues - 1);
    if ((*done >> q) & 1) continue;
    const uint32_t i = atomicAdd(args.queue + q * kQueueStride, 1u);
    if (i < args.qoff[q + 1] - args.qoff[q]) return args.qoff[q] + i;
    *done |= 1u << q;
  }
  return ~0u;
}

__device__ __forceinline__ void start_ctx(Ctx& C, const KernelArgs& args, const GroupDesc& Gd, uint32_t part) {
  C.spin_limit = args.spin_limit;
  C.G = Gd.G;
  C.g = part;
  C.rnd = 0;
  C.phase = 0;
  C.abort = false;
  C.relay = Gd.G >= args.relay_min_g;
  C.bar = args.bar + Gd.bar_off;
}

// Minimum waves per SIMD the compiler must leave room for.  Group tilings of
// at most four waves per workgroup and one slot row otherwise spill into
// AGPRs past 256 registers and run one wave per SIMD; held to two they fit
// 255 VGPRs with at most 16 bytes of spill (measured at (256, 1, 16): 395k
// against 234k GCD/s on 8,192 pairs of 8 kbit).
constexpr int min_waves_per_simd(int T, int K, int B) { return (B > 1 && B <= 16 && K == 1 && T <= 256) ? 2 : 1; }

template <int T, int K, int B, class A>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(min_waves_per_simd(T, K, B)))) void gcd_kernel(
    KernelArgs args) {
  __shared__ Shared<T, K, B> S;
  Ctx C;
  const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // hwreg(HW_REG_XCC_ID, 0, 4)
  const uint32_t own = args.nqueues > 1 ? xcc % (args.nqueues - 1) : 0;
  uint32_t done = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) S.s.item = next_item(args, own, &done);
    __syncthreads();
    const uint32_t it = S.s.item;
    if (it == ~0u) break;
    const WorkItem W = args.items[it];
    const GroupDesc Gd = args.groups[W.group];
    start_ctx(C, args, Gd, W.part);
    process_group<T, K, B, A>(args, Gd, S, C);
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
// Certifier of the quotient estimate near_quo_f that the modular inverse
// uses (the role of certifyQuoRemQuasi.cu:25-48).  Postcondition for
// integers x >= 0, y >= 1 with q = near_quo_f(fl32(x), fl32(y)) and
// r = x - q*y:
//   |r| < y whenever x < 2^20 y,   |r| < x whenever x >= 2^20 y,
// which is what makes |x| + |y| of the inverse's Euclid fall; the sign
// symmetry of the estimate (mode 2) extends it to negative x and y.
// Inputs reach near_quo_f the way modinv passes them: fl32 of an exact
// integer (round to nearest even).
__device__ __forceinline__ bool nq_bad(uint64_t x, uint64_t y) {
  const float q = near_quo_f((float)(double)x, (float)(double)y);
  const int64_t r = (int64_t)x - (int64_t)q * (int64_t)y;
  const uint64_t ar = (uint64_t)(r < 0 ? -r : r);
  return x < (y << 20) ? ar >= y : ar >= x;
}

// Quotient class k of y, x in [k*y, xh] with xh < (k+1)*y: q is
// non-decreasing in x (conversion, multiplication by a positive constant
// and rounding are monotone), so the estimates at the two ends bound every
// x of the class.  k < 2^20: q(k*y) = k and q(xh) in {k, k+1} give |r| < y
// throughout; k >= 2^20: 1 <= q <= 2k - 1 gives 0 < q*y < 2x, i.e. |r| < x.
__device__ __forceinline__ bool class_bad(uint64_t k, uint64_t y, uint64_t xh) {
  const int64_t qa = (int64_t)near_quo_f((float)(double)(k * y), (float)(double)y);
  const int64_t qb = (int64_t)near_quo_f((float)(double)xh, (float)(double)y);
  const int64_t ki = (int64_t)k;
  if (k < (1u << 20)) return qa != ki || qb < ki || qb > ki + 1;
  return qa < 1 || qb > 2 * ki - 1;
}

// Mode 0: every x in [1, 2^23), y in [1, 2^22), point by point.
__global__ void certify_float_kernel(unsigned long long* fail, unsigned long long* checked) {
  unsigned long long bad = 0, cnt = 0;
  for (uint32_t y = blockIdx.x * blockDim.x + threadIdx.x + 1; y < (1u << 22); y += gridDim.x * blockDim.x) {
    for (uint32_t x = 1; x < (1u << 23); ++x) bad += nq_bad(x, y);
    cnt += (1u << 23) - 1;
  }
  if (bad) atomicAdd(fail, bad);
  atomicAdd(checked, cnt);
}

// Mode 1: width L, y in [y_lo, y_hi), every quotient class k of x in
// [0, 2^L) through its endpoints x = k*y and x = min((k+1)*y - 1, 2^L - 1)
// (class_bad).  A launch covers one octave of y; unit u is (y = y_lo + (u >> cshift),
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
      bad += class_bad(k, y, min((k + 1) * y - 1, xmax));
      cnt += 2;
    }
  }
  if (bad) atomicAdd(fail, bad);
  if (cnt) atomicAdd(checked, cnt);
}

// Mode 2: every float f in [1, 2^34) (every |y| the inverse divides by):
// |rcp(f) * f - 1| <= 2^-23, evaluated exactly as one FP64 FMA (two 24-bit
// significands), and rcp(-f) == -rcp(f).
__global__ void certify_rcp_kernel(unsigned long long* fail, unsigned long long* checked) {
  constexpr uint32_t kLo = 0x3f800000u;  // 1.0f
  constexpr uint32_t kHi = 0x50800000u;  // 2^34
  unsigned long long bad = 0, cnt = 0;
  for (uint32_t b = kLo + blockIdx.x * blockDim.x + threadIdx.x; b < kHi; b += gridDim.x * blockDim.x) {
    const float f = __uint_as_float(b);
    const float r = __builtin_amdgcn_rcpf(f);
    const float rn = __builtin_amdgcn_rcpf(-f);
    const double e = __builtin_fma((double)r, (double)f, -1.0);
    bad += (__builtin_fabs(e) > 0x1p-23) | (rn != -r);
    ++cnt;
  }
  if (bad) atomicAdd(fail, bad);
  atomicAdd(checked, cnt);
}

}  // namespace

// ---------------------------------------------------------------------------
// Variant table.
#define MGCD_VARIANT(T, K, B, A, W) {T, K, B, W, reinterpret_cast<const void*>(&gcd_kernel<T, K, B, A>)}
#ifdef MGCD_PHASE_CLOCKS
// The instrumented build: the one-pair tilings and the default batch tiling.
#define MGCD_VARIANTS(A, W)                                                                      \
  MGCD_VARIANT(256, 1, 1, A, W), MGCD_VARIANT(1024, 1, 1, A, W), MGCD_VARIANT(512, 2, 16, A, W), \
      MGCD_VARIANT(512, 1, 1, A, W)
#else
#define MGCD_VARIANTS(A, W)                                                                              \
  MGCD_VARIANT(256, 1, 1, A, W), MGCD_VARIANT(256, 2, 1, A, W), MGCD_VARIANT(256, 4, 1, A, W),           \
      MGCD_VARIANT(256, 8, 1, A, W), MGCD_VARIANT(1024, 1, 1, A, W), MGCD_VARIANT(1024, 2, 1, A, W),     \
      MGCD_VARIANT(256, 1, 8, A, W), MGCD_VARIANT(256, 2, 8, A, W), MGCD_VARIANT(256, 1, 16, A, W),      \
      MGCD_VARIANT(256, 2, 16, A, W), MGCD_VARIANT(256, 1, 32, A, W), MGCD_VARIANT(128, 1, 16, A, W),    \
      MGCD_VARIANT(512, 2, 16, A, W), MGCD_VARIANT(512, 1, 16, A, W), MGCD_VARIANT(512, 1, 1, A, W)
#endif
static const KernelVariant kVariants[] = {MGCD_VARIANTS(Arith32, 1), MGCD_VARIANTS(Arith22, 0)};
#undef MGCD_VARIANTS
#undef MGCD_VARIANT

bool phase_clocks_built() {
#ifdef MGCD_PHASE_CLOCKS
  return true;
#else
  return false;
#endif
}

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
hipError_t launch_certify_rcp(unsigned long long* fail, unsigned long long* checked, int grid, hipStream_t s) {
  hipLaunchKernelGGL(certify_rcp_kernel, dim3(grid), dim3(256), 0, s, fail, checked);
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
