// mgcd_kernels.h -- host-visible declarations shared by the kernels
// (mgcd_kernels.hip) and the runtime (mgcd_runtime.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/mgcd.h"

namespace mgcd {

// Per-pair descriptor in device memory.  The larger operand is u.
struct PairDesc {
  uint64_t u_off;      // limb offset of u
  uint64_t v_off;      // limb offset of v
  uint32_t nu, nv;     // limb counts (no leading zero limbs)
  uint32_t nq;         // N_Q: moduli the pair starts with
  uint32_t pad0;
  uint64_t dig_off;    // first digit slot in the digit buffer
  uint64_t trace_off;  // first pivot-trace slot
  uint32_t trace_cap;  // pivot-trace slots (N_Q + 1: every pivot is recorded)
  uint32_t pad;
  uint64_t res_off;    // residue scratch: u mod m_i at [res_off + i], v at [res_off + N_Q + i]
};

// Pairs that share N_Q and are processed together by the same workgroups:
// pairs [first, first + count) of the descriptor array, count <= B of the
// kernel variant.  Every lane then holds the same modulus for all of them,
// which lets it invert all their values with one modular inverse.
struct GroupDesc {
  uint32_t first;
  uint32_t count;
  uint32_t nq;         // N_Q of every pair of the group
  uint32_t G;          // workgroups the group is split over
  uint64_t bar_off;    // first barrier slot (2*(G+1)*B exchange slots, then G residue-job words), if G > 1
};

// One unit of the persistent work queue: part `part` of group `group`.
struct WorkItem {
  uint32_t group;
  uint32_t part;
};

// Per-pair result header.
struct PairOut {
  int32_t status;   // KST_*
  uint32_t ndig;    // digits written
  uint32_t ntrace;  // elimination iterations (pivots used)
  uint32_t pad;
};

enum : int32_t { KST_OK = 0, KST_REDUX = 1, KST_RECOVERY = 2, KST_TIMEOUT = 3, KST_NOTRUN = 4 };

// Work queues: queue 0 is popped first by every workgroup (groups spanning
// many workgroups); queues 1..NQ-1 belong to the XCDs (all parts of a group
// on one XCD: the cross-workgroup exchange stays within one L2); a
// workgroup whose queues are empty takes items from the other XCDs'.
constexpr int kMaxQueues = 9;
constexpr int kQueueStride = 16;  // counters 64 bytes apart

struct KernelArgs {
  const uint32_t* limbs;
  const PairDesc* pairs;
  const GroupDesc* groups;
  const WorkItem* items;       // queue q: items [qoff[q], qoff[q+1])
  uint32_t nqueues;
  uint32_t qoff[kMaxQueues + 1];
  int L;
  uint32_t* queue;             // [q * kQueueStride]: next item of queue q
  const uint32_t* mod_u32;     // moduli table
  const double* mod_d;         // moduli as double
  const double* mod_inv;       // fl(1 / m)
  unsigned long long* bar;     // barrier slots (zeroed before each launch)
  mgcd_pair_t* digits;
  mgcd_pair_t* trace;          // pivots, written by part 0, replayed by every part
  uint32_t* res;               // residue scratch (2 N_Q per pair)
  PairOut* out;
  uint32_t spin_limit;         // polls before a barrier wait gives up
  uint32_t relay_min_g;        // groups of at least this many parts exchange through part 0
  unsigned long long* clocks;  // kPhaseClocks phase sums (instrumented build), else unused
};

// Phase sums of the instrumented build (-DMGCD_PHASE_CLOCKS, lib/libmgcd_phases.so):
// see mgcd_ctx_phase_clocks.
constexpr int kPhaseClocks = 32;
// True in the instrumented build.
bool phase_clocks_built();

// Kernel variants: threads per workgroup T, moduli slots per thread K,
// pairs per group B, for the two arithmetic classes (L <= 22: FP32
// inverse; L <= 32).
struct KernelVariant {
  int T, K, B, wide;  // wide = 1 for the 32-bit class
  const void* fn;     // __global__ function pointer for hipLaunchCooperativeKernel
};

int kernel_variant_count();
const KernelVariant& kernel_variant(int i);
// Finds the variant with exactly (T, K, B, wide), or -1.
int find_kernel_variant(int T, int K, int B, int wide);

// Small helper kernels.
hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

// Mixed-radix digits -> positional limbs on the device (mgcd_convert.hip):
// a product tree over each pair's digits, level j holding nodes of 2^j
// digits in two alternating level buffers.
struct MrPair {
  uint64_t dig_off;     // the pair's digits (least significant first) in the digit buffer
  uint64_t buf_off;     // its region (8k + 8 limbs) in each level buffer
  uint64_t out_off;     // |G| limbs, k + 1 of room, in the result buffer
  uint32_t k;           // digits
  uint32_t root_level;  // ceil(log2 k): the level that holds V(0, k)
};
struct MrArgs {
  const mgcd_pair_t* digits;
  const MrPair* pairs;
  uint32_t npairs;
  uint32_t j;           // input level (0: the digits)
  const uint32_t* cum;  // output nodes of pairs [0, p) at level j + 1; npairs + 1 entries
  const uint32_t* in;   // level j (j >= 1)
  uint32_t* out;        // level j + 1
  struct MrCarry* carry;  // per-chunk carry records (levels of one wave per chunk)
};
// Carry record of one 64-column chunk of an output node, resolved with no
// carry in: what it passes on (it holds the last column's high-half sum, so
// it can exceed 2^32), its two lowest limbs, and whether its limbs above
// those are all ones (bit 0) or all zeros (bit 1).  A carry in c with
// |c| < 2^63 leaves a carry of -1, 0 or 1 above the second limb, which
// passes through the chunk only over all ones (1) or all zeros (-1).
struct MrCarry {
  int64_t out;
  uint32_t limb0, limb1;
  uint32_t flags, pad;
};
// Levels below this (nodes of at most 4 digits) combine one node per
// thread; above it one wave computes 64 columns of a node's V or P.
constexpr uint32_t kMrSmallLevels = 3;
constexpr uint32_t kMrChunk = 64;
// Chunks per output V or P at level j + 1 (V has at most 2^(j+1) + 1 limbs).
constexpr uint64_t mr_chunks(uint32_t j) { return ((2ull << j) + kMrChunk) / kMrChunk; }
// Work units of level j -> j + 1 over ntasks output nodes: V and P chunks
// (one carry record each).
constexpr uint64_t mr_units(uint32_t j, uint64_t ntasks) { return j < kMrSmallLevels ? 0 : ntasks * 2 * mr_chunks(j); }
// Carry records a level of one pair of k digits needs at most (levels of
// one chunk per node write none).
constexpr uint64_t mr_max_records(uint64_t k) {
  uint64_t m = 0;
  for (uint32_t j = kMrSmallLevels; j < 63 && (1ull << j) < k; ++j) {
    if (mr_chunks(j) < 2) continue;
    const uint64_t u = mr_units(j, (k + (2ull << j) - 1) >> (j + 1));
    m = u > m ? u : m;
  }
  return m;
}
hipError_t launch_mr_level(const MrArgs& a, uint32_t ntasks, hipStream_t s);
// |G| per pair from its root: magnitude limbs at out_off, length in lens[p].
hipError_t launch_mr_emit(const MrArgs& a, const uint32_t* buf0, const uint32_t* buf1, uint32_t* res, uint32_t* lens,
                          hipStream_t s);

// Certifier launches (see mgcd_certify_quasi_quotient).
hipError_t launch_certify_float(unsigned long long* fail, unsigned long long* checked, int grid, hipStream_t s);
hipError_t launch_certify_rcp(unsigned long long* fail, unsigned long long* checked, int grid, hipStream_t s);
hipError_t launch_certify_endpoints(int L, uint64_t y_lo, uint64_t y_hi, unsigned long long* fail,
                                    unsigned long long* checked, int grid, hipStream_t s);

}  // namespace mgcd
