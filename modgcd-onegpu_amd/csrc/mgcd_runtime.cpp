// mgcd_runtime.cpp -- host runtime behind the C ABI of include/mgcd.h.
//
// Owns the device context (moduli table in HBM, stream, pooled buffers,
// occupancy of each kernel variant), plans launches (how many workgroups a
// pair is split over, which (T, K) variant runs a batch), launches the
// persistent kernel cooperatively so that all of its workgroups are
// resident (the cross-workgroup waits need that), and converts the signed
// mixed-radix digits back to positional limbs (the host Horner of
// GmpCudaDevice-gcd.cu:116-129, without GMP).
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <deque>
#include <exception>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mgcd_kernels.h"
#include "mgcd_table.h"

using namespace mgcd;

namespace {

constexpr uint32_t kMaxModuli = (1u << 19) - 1;  // kernel key field width
constexpr uint32_t kSpinLimit = 1u << 24;        // polls before a wait reports a timeout
// Default relay thresholds (parts): groups of pairs relay from 48 parts (the
// quadratic polling of ~160-part groups slowed every kernel on the device);
// one pair's parts poll one slot each, and relaying paid only from ~128 parts
// (measured: flat faster at 31-80 parts, the relay faster at 317).
constexpr uint32_t kRelayGroups = 48;
constexpr uint32_t kRelayOnePair = 128;

const char* kMessages[] = {
    "OK",
    "Invalid device number.",
    "Not enough moduli available for given input.",
    "Cannot allocate enough threads to support computation.",
    "Ran out of moduli in the reduction loop.",
    "Ran out of moduli in the recovery loop.",
    "Unknown error in the gcd kernel.",
    "HIP runtime error.",
    "Invalid argument.",
    "Moduli table missing or invalid.",
    "Output buffer too small.",
    "Out of memory.",
    "A device-side wait timed out.",
    "Admission budget exhausted.",
    "Input exceeds the no-allocation scratch limit.",
    "Request still pending.",
};

thread_local std::string g_last_error;

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      g_last_error = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return MGCD_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

// No C++ exception crosses the C ABI: a host allocation that fails (or any
// other exception the standard library raises, such as a thread that cannot
// be started) becomes MGCD_ERR_NOMEM, with the detail in g_last_error.
#define MGCD_CATCH                                   \
  catch (const std::exception& e_) {                 \
    g_last_error = std::string("host: ") + e_.what(); \
    return MGCD_ERR_NOMEM;                           \
  }                                                  \
  catch (...) {                                      \
    g_last_error = "host: unknown exception";        \
    return MGCD_ERR_NOMEM;                           \
  }

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  // retired: where the old allocation goes instead of hipFree (which waits
  // for the whole device), for callers that must not block while other
  // streams' kernels run; the caller frees them later.
  hipError_t ensure(size_t bytes, std::vector<void*>* retired = nullptr) {
    if (bytes <= cap) return hipSuccess;
    if (p && retired) retired->push_back(p);
    else if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    // Headroom: a reallocation stays rare.
    size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Grow-only pinned host buffer: copies to and from it are asynchronous.
struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes, std::vector<void*>* retired = nullptr) {
    if (bytes <= cap) return hipSuccess;
    if (p && retired) retired->push_back(p);
    else if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 2, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

size_t trim(const uint32_t* x, size_t n) {
  while (n > 0 && x[n - 1] == 0) --n;
  return n;
}
uint64_t bit_length(const uint32_t* x, size_t n) {
  n = trim(x, n);
  if (n == 0) return 0;
  return 32 * (uint64_t)(n - 1) + (32 - __builtin_clz(x[n - 1]));
}
int cmp_mag(const uint32_t* a, size_t na, const uint32_t* b, size_t nb) {
  na = trim(a, na);
  nb = trim(b, nb);
  if (na != nb) return na < nb ? -1 : 1;
  for (size_t i = na; i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

// Buffers of the device conversion (grow-only).
struct MrBufs {
  DevBuf pairs, cum, buf0, buf1, res, lens, carry;
  PinnedBuf h_res, h_lens;  // |G| limbs and lengths come back through pinned memory
  std::vector<MrPair> h_pairs;
  std::vector<uint32_t> h_cum;
  void release() {
    for (DevBuf* d : {&pairs, &cum, &buf0, &buf1, &res, &lens, &carry}) d->release();
    h_res.release();
    h_lens.release();
  }
};

// One-pair calls convert this many digits or fewer on the host: the
// device's launch sequence and copies cost about 0.27 ms whatever the
// size, which the quadratic Horner loop reaches near 700 digits (measured
// on MI355X: 0.16 ms at 313 digits, 0.36 ms at 1,001).
constexpr uint32_t kMrHostDigits = 640;

uint32_t ceil_log2(uint32_t k) {
  uint32_t j = 0;
  while ((1ull << j) < k) ++j;
  return j;
}

// |G| of every (dig_off, k) listed, from the digits in device memory at
// d_digits, by the product tree of mgcd_convert.hip on stream s; waits for
// the stream.  out[i] receives pair i's limbs.  Returns an MGCD_ status.
// retired, retired_host: as in DevBuf::ensure and PinnedBuf::ensure (callers whose device
// runs other streams' kernels).
int mr_convert(hipStream_t s, const mgcd_pair_t* d_digits, const std::vector<uint64_t>& dig_off,
               const std::vector<uint32_t>& ks, MrBufs& mb, std::vector<std::vector<uint32_t>*>& out,
               std::vector<void*>* retired = nullptr, std::vector<void*>* retired_host = nullptr) {
  const uint32_t n = (uint32_t)ks.size();
  if (n == 0) return MGCD_OK;
  mb.h_pairs.resize(n);
  uint64_t buf = 0, res = 0;
  uint32_t levels = 0;
  for (uint32_t i = 0; i < n; ++i) {
    MrPair& p = mb.h_pairs[i];
    p.dig_off = dig_off[i];
    p.k = ks[i];
    p.buf_off = buf;
    p.out_off = res;
    p.root_level = ceil_log2(p.k);
    buf += 8ull * p.k + 8;
    res += (uint64_t)p.k + 1;
    levels = std::max(levels, p.root_level);
  }
  // Level j -> j + 1 combines, per pair, the ceil(k / 2^(j+1)) output nodes
  // of every pair not yet reduced to its root.
  mb.h_cum.assign((size_t)std::max(levels, 1u) * (n + 1), 0);
  for (uint32_t j = 0; j < levels; ++j) {
    uint32_t* c = mb.h_cum.data() + (size_t)j * (n + 1);
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t k = ks[i];
      const uint32_t t = (uint64_t)k > (1ull << j) ? (uint32_t)((k + (2ull << j) - 1) >> (j + 1)) : 0u;
      c[i + 1] = c[i] + t;
    }
  }
  uint64_t records = 1;
  for (uint32_t j = kMrSmallLevels; j < levels; ++j)
    if (mr_chunks(j) > 1) records = std::max<uint64_t>(records, mr_units(j, mb.h_cum[(size_t)j * (n + 1) + n]));
  if (mb.pairs.ensure(n * sizeof(MrPair), retired) != hipSuccess ||
      mb.cum.ensure(mb.h_cum.size() * 4, retired) != hipSuccess || mb.buf0.ensure(buf * 4, retired) != hipSuccess ||
      mb.buf1.ensure(buf * 4, retired) != hipSuccess || mb.res.ensure(res * 4, retired) != hipSuccess ||
      mb.lens.ensure(n * 4, retired) != hipSuccess || mb.carry.ensure(records * sizeof(MrCarry), retired) != hipSuccess ||
      mb.h_res.ensure(res * 4, retired_host) != hipSuccess || mb.h_lens.ensure(n * 4, retired_host) != hipSuccess)
    return MGCD_ERR_NOMEM;
  HIP_TRY(hipMemcpyAsync(mb.pairs.p, mb.h_pairs.data(), n * sizeof(MrPair), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(mb.cum.p, mb.h_cum.data(), mb.h_cum.size() * 4, hipMemcpyHostToDevice, s));
  uint32_t* lv[2] = {mb.buf0.as<uint32_t>(), mb.buf1.as<uint32_t>()};
  MrArgs a{};
  a.digits = d_digits;
  a.pairs = mb.pairs.as<MrPair>();
  a.npairs = n;
  a.carry = mb.carry.as<MrCarry>();
  for (uint32_t j = 0; j < levels; ++j) {
    a.j = j;
    a.cum = mb.cum.as<uint32_t>() + (size_t)j * (n + 1);
    a.in = lv[j & 1];
    a.out = lv[(j + 1) & 1];
    HIP_TRY(launch_mr_level(a, mb.h_cum[(size_t)j * (n + 1) + n], s));
  }
  HIP_TRY(launch_mr_emit(a, lv[0], lv[1], mb.res.as<uint32_t>(), mb.lens.as<uint32_t>(), s));
  HIP_TRY(hipMemcpyAsync(mb.h_res.p, mb.res.p, res * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(mb.h_lens.p, mb.lens.p, n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t* r = mb.h_res.as<uint32_t>() + mb.h_pairs[i].out_off;
    out[i]->assign(r, r + mb.h_lens.as<uint32_t>()[i]);
  }
  return MGCD_OK;
}

// Signed mixed-radix digits (least significant first) -> |G| as limbs:
// G = g_k; G = G * m_i + g_i for i = k-1 .. 1 (GmpCudaDevice-gcd.cu:116-129).
// When G != 0, |G * m_i| >= m_i > |g_i|, so the sign never flips mid-way.
void digits_to_limbs(const mgcd_pair_t* d, size_t k, std::vector<uint32_t>& mag) {
  mag.clear();
  if (k == 0) return;
  bool neg = d[k - 1].value < 0;
  uint32_t top = neg ? 0u - (uint32_t)d[k - 1].value : (uint32_t)d[k - 1].value;
  if (top) mag.push_back(top);
  for (size_t i = k - 1; i-- > 0;) {
    const uint64_t p = d[i].modulus;
    uint64_t carry = 0;
    for (uint32_t& w : mag) {
      const uint64_t x = (uint64_t)w * p + carry;
      w = (uint32_t)x;
      carry = x >> 32;
    }
    if (carry) mag.push_back((uint32_t)carry);
    const int32_t g = d[i].value;
    if (g == 0) continue;
    const bool gneg = g < 0;
    const uint32_t ga = gneg ? 0u - (uint32_t)g : (uint32_t)g;
    if (mag.empty()) {
      mag.push_back(ga);
      neg = gneg;
    } else if (gneg == neg) {
      uint64_t c = ga;
      for (size_t j = 0; c && j < mag.size(); ++j) {
        const uint64_t x = (uint64_t)mag[j] + c;
        mag[j] = (uint32_t)x;
        c = x >> 32;
      }
      if (c) mag.push_back((uint32_t)c);
    } else {
      uint64_t b = ga;  // mag > ga
      for (size_t j = 0; b && j < mag.size(); ++j) {
        const uint64_t x = (uint64_t)mag[j] - b;
        mag[j] = (uint32_t)x;
        b = (x >> 63) & 1;
      }
      while (!mag.empty() && mag.back() == 0) mag.pop_back();
    }
  }
}

std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&library_dir), &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t s = p.rfind('/');
    if (s != std::string::npos) return p.substr(0, s);
  }
  return ".";
}

int kst_to_status(int32_t k) {
  switch (k) {
    case KST_OK: return MGCD_OK;
    case KST_REDUX: return MGCD_ERR_REDUX;
    case KST_RECOVERY: return MGCD_ERR_RECOVERY;
    case KST_TIMEOUT: return MGCD_ERR_TIMEOUT;
    default: return MGCD_ERR_KERNEL;
  }
}

// ---------------------------------------------------------------------------
// The one-pair path without allocation (mgcd_gcd_limbs_noalloc).  Every
// device buffer of such a call lives in `scratch`, a file-scope device array
// that each device the library runs on holds until the process ends.  Its
// name is a host-side stand-in, not a device address: host code uses only
// the address hipGetSymbolAddress returns for the current device, taken once
// per device when a context is created there (scratch_init).  Calls share
// the array of their device and keep nothing in it between calls; calls that
// overlap on one device race.

constexpr uint32_t kScratchMaxModuli = 16384;  // N_Q (~176 kbit operands at L = 32, ~153 kbit at L = 22)
constexpr size_t kScratchMaxLimbs = 12288;     // nu + nv
constexpr uint32_t kScratchMaxParts = 64;      // workgroups of one pair (N_Q / 256)
constexpr uint32_t kScratchMaxLevels = 32;     // product-tree levels of the conversion
constexpr size_t kScratchBytes = (size_t)2 << 20;

struct ScratchLayout {
  size_t limbs, pair, group, items, queue, bar, digits, trace, res, out, mrpair, mrcum, buf0, buf1, mres, mlen, mrcarry,
      end;
};
constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr ScratchLayout scratch_layout() {
  ScratchLayout l{};
  size_t o = 0;
  l.limbs = o;  o = al256(o + kScratchMaxLimbs * 4);
  l.pair = o;   o = al256(o + sizeof(PairDesc));
  l.group = o;  o = al256(o + sizeof(GroupDesc));
  l.items = o;  o = al256(o + kScratchMaxParts * sizeof(WorkItem));
  l.queue = o;  o = al256(o + (size_t)kMaxQueues * kQueueStride * 4);
  l.bar = o;    o = al256(o + ((size_t)2 * (kScratchMaxParts + 1) + kScratchMaxParts) * 8);  // one pair per group
  l.digits = o; o = al256(o + (size_t)kScratchMaxModuli * 8);
  l.trace = o;  o = al256(o + ((size_t)kScratchMaxModuli + 32) * 8);
  l.res = o;    o = al256(o + ((size_t)2 * kScratchMaxModuli + 64) * 4);
  l.out = o;    o = al256(o + sizeof(PairOut));
  l.mrpair = o; o = al256(o + sizeof(MrPair));
  l.mrcum = o;  o = al256(o + (size_t)kScratchMaxLevels * 2 * 4);
  l.buf0 = o;   o = al256(o + (8 * (size_t)kScratchMaxModuli + 8) * 4);
  l.buf1 = o;   o = al256(o + (8 * (size_t)kScratchMaxModuli + 8) * 4);
  l.mres = o;   o = al256(o + ((size_t)kScratchMaxModuli + 1) * 4);
  l.mlen = o;   o = al256(o + 4);
  l.mrcarry = o; o = al256(o + std::max<uint64_t>(1, mr_max_records(kScratchMaxModuli)) * sizeof(MrCarry));
  l.end = o;
  return l;
}
constexpr ScratchLayout kScratch = scratch_layout();
static_assert(kScratch.end <= kScratchBytes, "the scratch layout must fit the array");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on `device`, which must be current.
int scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices) return MGCD_ERR_INVALID_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MGCD_OK;
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) {
    g_last_error = "hipGetSymbolAddress gave no address for the scratch array";
    return MGCD_ERR_HIP;
  }
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  return MGCD_OK;
}

unsigned char* scratch_of(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace

struct mgcd_ctx {
  int device = 0;
  int L = 32;
  int wide = 1;
  int num_cus = 0;
  std::vector<uint32_t> moduli;
  DevBuf d_mod_u32, d_mod_d, d_mod_inv;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<int> max_resident;  // per kernel variant: co-resident workgroups
  // pooled buffers of the single-pair path
  DevBuf limbs, pairs, groups, items, queue, bar, digits, trace, res, out;
  double last_kernel_s = 0.0;
  // the last one-pair launch: N_Q, iterations, digits, tiling (mgcd_ctx_last_pair_stats)
  uint32_t last_nq = 0, last_iters = 0, last_digits = 0;
  int last_T = 0, last_K = 0, last_G = 0;
  uint32_t force_nq = 0;  // MGCD_OPT_FORCE_MODULI
  bool retry = true;      // MGCD_OPT_RETRY
  int single_T = 0;       // MGCD_OPT_SINGLE_WORKGROUP (0: automatic)
  int single_K = 0;       // MGCD_OPT_SINGLE_SLOTS (0: automatic)
  bool autotune = false;  // MGCD_OPT_AUTOTUNE
  bool batch_graph = false;  // MGCD_OPT_BATCH_GRAPH
  bool tune_loaded = false;
  uint32_t relay_min_g = 0;   // MGCD_OPT_RELAY_PARTS (0: kRelayGroups / kRelayOnePair)
  bool device_convert = true;  // MGCD_OPT_DEVICE_CONVERT
  uint32_t spin_limit = kSpinLimit;  // MGCD_OPT_SPIN_LIMIT
  bool cooperative = true;     // MGCD_OPT_COOPERATIVE (default: $MGCD_COOPERATIVE, else 1)
  std::map<std::string, int> tuned;  // size class -> kernel variant
  MrBufs mr;                   // the one-pair path's conversion buffers
  DevBuf clocks;               // phase sums (instrumented build only)
};

// A prepared launch over device-resident pairs.
struct LaunchSet {
  int variant = -1;
  uint32_t nitems = 0;
  uint32_t nqueues = 1;
  uint32_t qoff[kMaxQueues + 1] = {0};
  uint32_t grid = 0;
  size_t nbar = 0;       // barrier slots
  size_t ndigits = 0;    // digit slots
  size_t ntrace = 0;     // trace slots
  size_t nres = 0;       // residue scratch words
};

namespace {

// Plans a launch of variant `vi`: pairs with equal N_Q are grouped, at most
// B per group; a group is split over G = ceil(N_Q / (T*K)) workgroups.
// Groups go to the work queue in decreasing order of work (N_Q^2) with the
// G parts of a group consecutive, so groups spanning several workgroups are
// popped first.  pd is reordered into group order; order[j] is the caller's
// index of kernel pair j.
// grid_cap (0: none) bounds the grid below the variant's co-resident limit
// (a serving slot's share of the device); the queue plan is made for it.
int plan_items(mgcd_ctx* ctx, int vi, std::vector<PairDesc>& pd, std::vector<uint32_t>& order,
               std::vector<GroupDesc>& groups, std::vector<WorkItem>& items, LaunchSet& ls, uint32_t grid_cap = 0) {
  const KernelVariant& kv = kernel_variant(vi);
  const uint32_t per_wg = (uint32_t)(kv.T * kv.K);
  const uint32_t maxres =
      grid_cap ? std::min<uint32_t>(grid_cap, (uint32_t)ctx->max_resident[vi]) : (uint32_t)ctx->max_resident[vi];
  if (maxres == 0) return MGCD_ERR_TOO_MANY_THREADS;
  order.resize(pd.size());
  for (uint32_t i = 0; i < pd.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return pd[x].nq > pd[y].nq; });
  std::vector<PairDesc> sorted;
  sorted.reserve(pd.size());
  groups.clear();
  size_t bar = 0, dig = 0, tr = 0, res = 0;
  uint32_t maxG = 1;
  for (uint32_t j = 0; j < order.size(); ++j) {
    PairDesc p = pd[order[j]];
    p.dig_off = dig;
    dig += p.nq;
    // Per-pair regions start on 256-byte boundaries, so no cache line is
    // shared between pairs that different workgroups write and read.
    p.trace_off = tr;
    p.trace_cap = p.nq + 1;
    tr += (p.trace_cap + 31) & ~(size_t)31;
    p.res_off = res;
    res += (2 * (size_t)p.nq + 63) & ~(size_t)63;
    if (groups.empty() || groups.back().nq != p.nq || groups.back().count == (uint32_t)kv.B) {
      GroupDesc gd;
      std::memset(&gd, 0, sizeof gd);
      gd.first = j;
      gd.nq = p.nq;
      gd.G = (p.nq + per_wg - 1) / per_wg;
      if (gd.G > maxres) return MGCD_ERR_TOO_MANY_THREADS;
      maxG = std::max(maxG, gd.G);
      gd.bar_off = bar;
      if (gd.G > 1) bar += 2 * ((size_t)gd.G + 1) * kv.B + gd.G;  // exchange rows, relay row, residue jobs
      groups.push_back(gd);
    }
    groups.back().count++;
    sorted.push_back(p);
  }
  pd.swap(sorted);
  // Queues: groups of more than maxres/16 parts go to the shared queue 0;
  // the others are spread over the 8 XCD queues, largest work first onto
  // the least loaded.  At any time each queue has at most one group whose
  // parts are not all taken; with the shared queue popped first this
  // cannot tie up every workgroup as long as the shared queue's groups
  // need at most maxres/2 workgroups.  Otherwise one queue holds everything.
  const uint32_t nitems_all = [&] {
    uint32_t n = 0;
    for (const GroupDesc& gd : groups) n += gd.G;
    return n;
  }();
  const uint32_t grid = std::min<uint32_t>(nitems_all, maxres);
  const uint32_t small = grid / 16;
  std::vector<std::vector<uint32_t>> qg(kMaxQueues);
  uint32_t nqueues = maxG <= grid / 2 && grid >= 64 ? (uint32_t)kMaxQueues : 1u;
  if (nqueues == 1) {
    for (uint32_t gi = 0; gi < groups.size(); ++gi) qg[0].push_back(gi);
  } else {
    std::vector<double> load(kMaxQueues, 0.0);
    for (uint32_t gi = 0; gi < groups.size(); ++gi) {  // groups are in decreasing N_Q order
      const GroupDesc& gd = groups[gi];
      if (gd.G > small) {
        qg[0].push_back(gi);
        continue;
      }
      const double w = (double)gd.count * gd.nq * gd.nq;
      const int q = 1 + (int)(std::min_element(load.begin() + 1, load.end()) - load.begin() - 1);
      load[q] += w;
      qg[q].push_back(gi);
    }
  }
  items.clear();
  for (uint32_t q = 0; q < nqueues; ++q) {
    ls.qoff[q] = (uint32_t)items.size();
    for (uint32_t gi : qg[q])
      for (uint32_t g = 0; g < groups[gi].G; ++g) items.push_back(WorkItem{gi, g});
  }
  ls.qoff[nqueues] = (uint32_t)items.size();
  ls.nqueues = nqueues;
  ls.variant = vi;
  ls.nitems = (uint32_t)items.size();
  ls.grid = std::min<uint32_t>(ls.nitems, maxres);
  if (ls.grid < maxG) return MGCD_ERR_TOO_MANY_THREADS;
  ls.nbar = bar;
  ls.ndigits = dig;
  ls.ntrace = tr;
  ls.nres = res;
  return MGCD_OK;
}

// Launches a persistent kernel whose grid fits the device's resident
// capacity for it.  Cooperatively when asked and possible: that guarantees
// that all workgroups are resident.  Where the runtime refuses a
// cooperative launch (another process holds the device's cooperative
// queue, or a graph is being captured) the kernel is launched plainly with
// the same grid; on an otherwise idle device its workgroups are then all
// resident as well, and a workgroup that is not only delays the others'
// bounded waits, which end in a timeout status that the host retries.
hipError_t launch_resident(const void* fn, uint32_t grid, int T, void** params, hipStream_t s, bool cooperative) {
  if (cooperative) {
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(T), params, 0, s);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
  }
  return hipLaunchKernel(fn, dim3(grid), dim3(T), params, 0, s);
}

// Clears the queue counter and barrier slots, then launches the persistent
// kernel cooperatively on ctx->stream; records events around the kernel.
int launch(mgcd_ctx* ctx, const LaunchSet& ls, const KernelArgs& ka) {
  const KernelVariant& kv = kernel_variant(ls.variant);
  HIP_TRY(launch_fill_u32(ka.queue, kMaxQueues * kQueueStride, 0u, ctx->stream));
  if (ls.nbar) HIP_TRY(launch_fill_u64(ka.bar, ls.nbar, 0ull, ctx->stream));
  KernelArgs args = ka;
  args.nqueues = ls.nqueues;
  std::memcpy(args.qoff, ls.qoff, sizeof args.qoff);
  void* params[] = {&args};
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  HIP_TRY(launch_resident(kv.fn, ls.grid, kv.T, params, ctx->stream, ctx->cooperative));
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  return MGCD_OK;
}

// The relay threshold a launch of variant vi uses.
uint32_t relay_parts(const mgcd_ctx* ctx, int vi) {
  if (ctx->relay_min_g) return ctx->relay_min_g;
  return kernel_variant(vi).B == 1 ? kRelayOnePair : kRelayGroups;
}

int choose_single_variant(mgcd_ctx* ctx, uint32_t nq) {
  const int w = ctx->wide;
  if (ctx->single_T > 0 && ctx->single_K > 0) {
    const int vi = find_kernel_variant(ctx->single_T, ctx->single_K, 1, w);
    if (vi >= 0 && (nq + ctx->single_T * ctx->single_K - 1) / (ctx->single_T * ctx->single_K) <=
                       (uint32_t)ctx->max_resident[vi])
      return vi;
  }
  if (nq <= 1024) return find_kernel_variant(1024, 1, 1, w);
  // From 8 parts of 256 threads up to 128 parts of 512 threads the
  // 512-thread tiling (two waves per SIMD, half the parts, no relay) is
  // faster: measured 2-3 % at 16-48 Kibit, 4-7 % at 64-100 kbit, 10-15 % at
  // 128-256 Kibit, 7-12 % at 352 Kibit (1,024 threads, a quarter of the
  // parts, is 10-20 % slower throughout).  Below 8 parts (12 Kibit) the
  // 256-thread tiling is 2 % faster; with more parts (1 Mbit) too, relayed.
  {
    const int v512 = find_kernel_variant(512, 1, 1, w);
    const uint32_t g256 = (nq + 255) / 256, g512 = (nq + 511) / 512;
    if (v512 >= 0 && g256 >= 8 && g512 < kRelayOnePair && g512 <= (uint32_t)ctx->max_resident[v512])
      return v512;
  }
  for (int K : {1, 2, 4, 8}) {
    const int vi = find_kernel_variant(256, K, 1, w);
    if (vi < 0) continue;
    const uint32_t G = (nq + 256 * K - 1) / (256 * K);
    if (G <= (uint32_t)ctx->max_resident[vi]) return vi;
  }
  return -1;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* mgcd_strerror(int status) {
  if (status < 0 || status >= (int)(sizeof(kMessages) / sizeof(kMessages[0]))) return "Unknown status.";
  return kMessages[status];
}

int mgcd_device_count(int* count) {
  if (!count) return MGCD_ERR_ARG;
  HIP_TRY(hipGetDeviceCount(count));
  return MGCD_OK;
}

int mgcd_set_device(int dev) {
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (dev < 0 || dev >= n) return MGCD_ERR_INVALID_DEVICE;
  HIP_TRY(hipSetDevice(dev));
  return MGCD_OK;
}

int mgcd_num_moduli_needed(uint64_t bits, int L) {
  // GmpCudaDevice-gcd.cu:49-55, in float as the reference computes it.
  // ln(1) = 0 leaves the formula undefined below two bits: such inputs are
  // sized as two-bit ones.  Counts beyond 2^30 saturate there (no table
  // comes near it).
  if (bits < 2) bits = 2;
  const float C_L = (float)(1.6 - 0.015 * L);
  float x = C_L * (float)bits;
  x = x / logf((float)bits);
  if (!(x < 0x1p30f)) return 1 << 30;
  const int n = (int)ceilf(x);
  return ((n - 1) / 256 + 1) * 256;
}

int mgcd_ctx_create(mgcd_ctx** out, int device, int L, const char* table_path) try {
  if (!out || (L != 22 && L != 27 && L != 32)) return MGCD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return MGCD_ERR_INVALID_DEVICE;
  HIP_TRY(hipSetDevice(device));

  std::string path;
  if (table_path && *table_path) {
    path = table_path;
  } else {
    const char* dir = std::getenv("MGCD_TABLE_DIR");
    std::string d = dir && *dir ? std::string(dir) : library_dir() + "/../data";
    path = d + "/moduli_L" + std::to_string(L) + ".bin";
  }
  int fileL = 0;
  std::vector<uint32_t> mod;
  const std::string err = readTable(path, &fileL, &mod);
  if (!err.empty()) {
    g_last_error = err;
    return MGCD_ERR_TABLE;
  }
  if (fileL != L) {
    g_last_error = "table width mismatch in " + path;
    return MGCD_ERR_TABLE;
  }
  if (mod.size() > kMaxModuli) mod.resize(kMaxModuli);

  mgcd_ctx* ctx = new mgcd_ctx();
  ctx->device = device;
  ctx->L = L;
  ctx->wide = L > 22 ? 1 : 0;
  if (const char* co = std::getenv("MGCD_COOPERATIVE")) ctx->cooperative = std::atoi(co) != 0;
  ctx->moduli = std::move(mod);
  hipDeviceProp_t props;
  hipError_t e = hipGetDeviceProperties(&props, device);
  if (e != hipSuccess || props.warpSize != 64) {
    delete ctx;
    g_last_error = "device properties unavailable or wavefront size != 64";
    return MGCD_ERR_HIP;
  }
  ctx->num_cus = props.multiProcessorCount;
  {
    const int src = scratch_init(device);
    if (src != MGCD_OK) {
      delete ctx;
      return src;
    }
  }
  const size_t nm = ctx->moduli.size();
  std::vector<double> md(nm), mi(nm);
  for (size_t i = 0; i < nm; ++i) {
    md[i] = (double)ctx->moduli[i];
    mi[i] = 1.0 / md[i];
  }
  int rc = MGCD_OK;
  auto fail = [&](int code) {
    rc = code;
    return code;
  };
  if (ctx->d_mod_u32.ensure(nm * 4) != hipSuccess || ctx->d_mod_d.ensure(nm * 8) != hipSuccess ||
      ctx->d_mod_inv.ensure(nm * 8) != hipSuccess) {
    fail(MGCD_ERR_NOMEM);
  } else if (hipMemcpy(ctx->d_mod_u32.p, ctx->moduli.data(), nm * 4, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(ctx->d_mod_d.p, md.data(), nm * 8, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(ctx->d_mod_inv.p, mi.data(), nm * 8, hipMemcpyHostToDevice) != hipSuccess ||
             hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    fail(MGCD_ERR_HIP);
  }
  if (rc == MGCD_OK && phase_clocks_built()) {
    if (ctx->clocks.ensure(kPhaseClocks * 8) != hipSuccess)
      fail(MGCD_ERR_NOMEM);
    else if (hipMemset(ctx->clocks.p, 0, kPhaseClocks * 8) != hipSuccess)
      fail(MGCD_ERR_HIP);
  }
  if (rc == MGCD_OK) {
    ctx->max_resident.assign(kernel_variant_count(), 0);
    for (int i = 0; i < kernel_variant_count(); ++i) {
      const KernelVariant& kv = kernel_variant(i);
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kv.fn, kv.T, 0) != hipSuccess) per_cu = 0;
      ctx->max_resident[i] = per_cu * ctx->num_cus;
    }
  }
  if (rc != MGCD_OK) {
    mgcd_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return MGCD_OK;
} MGCD_CATCH

int mgcd_ctx_destroy(mgcd_ctx* ctx) {
  if (!ctx) return MGCD_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (DevBuf* b : {&ctx->d_mod_u32, &ctx->d_mod_d, &ctx->d_mod_inv, &ctx->limbs, &ctx->pairs, &ctx->groups, &ctx->items,
                    &ctx->queue, &ctx->bar, &ctx->digits, &ctx->trace, &ctx->res, &ctx->out, &ctx->clocks})
    b->release();
  ctx->mr.release();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MGCD_OK;
}

int mgcd_ctx_L(const mgcd_ctx* ctx) { return ctx ? ctx->L : 0; }
int mgcd_ctx_num_moduli(const mgcd_ctx* ctx) { return ctx ? (int)ctx->moduli.size() : 0; }
int mgcd_ctx_max_grid(const mgcd_ctx* ctx) {
  if (!ctx) return 0;
  const int vi = find_kernel_variant(256, 1, 1, ctx->wide);
  return vi < 0 ? 0 : ctx->max_resident[vi];
}
int mgcd_ctx_moduli(const mgcd_ctx* ctx, uint32_t* out, size_t cap) {
  if (!ctx || !out || cap < ctx->moduli.size()) return MGCD_ERR_CAPACITY;
  std::memcpy(out, ctx->moduli.data(), ctx->moduli.size() * 4);
  return MGCD_OK;
}

const char* mgcd_last_error_detail(void) { return g_last_error.c_str(); }

int mgcd_ctx_set_option(mgcd_ctx* ctx, int option, int64_t value) {
  if (!ctx) return MGCD_ERR_ARG;
  switch (option) {
    case MGCD_OPT_FORCE_MODULI:
      if (value < 0 || (uint64_t)value > ctx->moduli.size()) return MGCD_ERR_ARG;
      ctx->force_nq = (uint32_t)value;
      return MGCD_OK;
    case MGCD_OPT_RETRY:
      ctx->retry = value != 0;
      return MGCD_OK;
    case MGCD_OPT_SINGLE_WORKGROUP:
      ctx->single_T = (int)value;
      return MGCD_OK;
    case MGCD_OPT_SINGLE_SLOTS:
      ctx->single_K = (int)value;
      return MGCD_OK;
    case MGCD_OPT_AUTOTUNE:
      ctx->autotune = value != 0;
      return MGCD_OK;
    case MGCD_OPT_BATCH_GRAPH:
      ctx->batch_graph = value != 0;
      return MGCD_OK;
    case MGCD_OPT_RELAY_PARTS:
      if (value < 0 || value == 1) return MGCD_ERR_ARG;
      ctx->relay_min_g = (uint32_t)std::min<int64_t>(value, 1 << 30);  // 0: the defaults
      return MGCD_OK;
    case MGCD_OPT_DEVICE_CONVERT:
      ctx->device_convert = value != 0;
      return MGCD_OK;
    case MGCD_OPT_SPIN_LIMIT:
      if (value < 1 || value > (int64_t)1 << 31) return MGCD_ERR_ARG;
      ctx->spin_limit = (uint32_t)value;
      return MGCD_OK;
    case MGCD_OPT_COOPERATIVE:
      ctx->cooperative = value != 0;
      return MGCD_OK;
    default:
      return MGCD_ERR_ARG;
  }
}

}  // extern "C"

namespace {

// N_Q for an operand of `bits` bits: the reference's sizing, or the forced
// count.  A 1-bit operand is sized as a 2-bit one: the reference's formula
// divides by ln(1) = 0 there (GmpCudaDevice-gcd.cu:54).
int moduli_for(const mgcd_ctx* ctx, uint64_t bits) {
  if (ctx->force_nq) return (int)ctx->force_nq;
  return mgcd_num_moduli_needed(std::max<uint64_t>(bits, 2), ctx->L);
}

// The last-launch figures of a one-pair call that may not launch (v = 0).
void clear_last_pair(mgcd_ctx* ctx) {
  ctx->last_kernel_s = 0.0;
  ctx->last_nq = ctx->last_iters = ctx->last_digits = 0;
  ctx->last_T = ctx->last_K = ctx->last_G = 0;
}

// One pair, u taken as given, with N_Q = nq moduli.
int gcd_single(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, int nqi, uint32_t* g,
               size_t cap, size_t* ng, mgcd_pair_t* digits, size_t dcap, size_t* nd, mgcd_pair_t* pivots,
               size_t tcap, size_t* nt) {
  if (nd) *nd = 0;
  if (nt) *nt = 0;
  *ng = 0;
  if (nqi <= 0 || (size_t)nqi > ctx->moduli.size()) return MGCD_ERR_NOT_ENOUGH_MODULI;
  const uint32_t nq = (uint32_t)nqi;
  const int vi = choose_single_variant(ctx, nq);
  if (vi < 0) return MGCD_ERR_TOO_MANY_THREADS;

  std::vector<PairDesc> pd(1);
  std::memset(pd.data(), 0, sizeof(PairDesc));
  pd[0].u_off = 0;
  pd[0].v_off = nu;
  pd[0].nu = (uint32_t)nu;
  pd[0].nv = (uint32_t)nv;
  pd[0].nq = nq;
  std::vector<WorkItem> items;
  std::vector<GroupDesc> groups;
  std::vector<uint32_t> order;
  LaunchSet ls;
  const size_t trace_cap = (pivots && tcap) ? std::min<size_t>(tcap, nq) : 0;
  int rc = plan_items(ctx, vi, pd, order, groups, items, ls);
  if (rc != MGCD_OK) return rc;

  if (ctx->limbs.ensure((nu + nv + 1) * 4) != hipSuccess || ctx->pairs.ensure(sizeof(PairDesc)) != hipSuccess ||
      ctx->groups.ensure(sizeof(GroupDesc)) != hipSuccess ||
      ctx->items.ensure(items.size() * sizeof(WorkItem)) != hipSuccess || ctx->queue.ensure(kMaxQueues * kQueueStride * 4) != hipSuccess ||
      ctx->bar.ensure((ls.nbar + 1) * 8) != hipSuccess || ctx->digits.ensure(ls.ndigits * 8) != hipSuccess ||
      ctx->trace.ensure((ls.ntrace + 1) * 8) != hipSuccess || ctx->res.ensure((ls.nres + 1) * 4) != hipSuccess ||
      ctx->out.ensure(sizeof(PairOut)) != hipSuccess)
    return MGCD_ERR_NOMEM;
  hipStream_t s = ctx->stream;
  if (nu) HIP_TRY(hipMemcpyAsync(ctx->limbs.p, u, nu * 4, hipMemcpyHostToDevice, s));
  if (nv) HIP_TRY(hipMemcpyAsync(ctx->limbs.as<uint32_t>() + nu, v, nv * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->pairs.p, pd.data(), sizeof(PairDesc), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->groups.p, groups.data(), sizeof(GroupDesc), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->items.p, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s));

  KernelArgs ka;
  ka.limbs = ctx->limbs.as<uint32_t>();
  ka.pairs = ctx->pairs.as<PairDesc>();
  ka.groups = ctx->groups.as<GroupDesc>();
  ka.items = ctx->items.as<WorkItem>();
  ka.L = ctx->L;
  ka.queue = ctx->queue.as<uint32_t>();
  ka.mod_u32 = ctx->d_mod_u32.as<uint32_t>();
  ka.mod_d = ctx->d_mod_d.as<double>();
  ka.mod_inv = ctx->d_mod_inv.as<double>();
  ka.bar = ctx->bar.as<unsigned long long>();
  ka.digits = ctx->digits.as<mgcd_pair_t>();
  ka.trace = ctx->trace.as<mgcd_pair_t>();
  ka.res = ctx->res.as<uint32_t>();
  ka.out = ctx->out.as<PairOut>();
  ka.spin_limit = ctx->spin_limit;
  ka.relay_min_g = relay_parts(ctx, vi);
  ka.clocks = ctx->clocks.as<unsigned long long>();
  rc = launch(ctx, ls, ka);
  if (rc != MGCD_OK) return rc;
  PairOut po;
  HIP_TRY(hipMemcpyAsync(&po, ctx->out.p, sizeof(PairOut), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_s = ms * 1e-3;
  ctx->last_nq = nq;
  ctx->last_iters = po.ntrace;
  ctx->last_digits = po.ndig;
  ctx->last_T = kernel_variant(vi).T;
  ctx->last_K = kernel_variant(vi).K;
  ctx->last_G = (int)groups[0].G;
  if (nt) *nt = po.ntrace;
  if (nd) *nd = po.ndig;
  if (trace_cap && po.ntrace)
    HIP_TRY(hipMemcpy(pivots, ctx->trace.p, std::min<size_t>(po.ntrace, trace_cap) * 8, hipMemcpyDeviceToHost));
  if (po.status != KST_OK) return kst_to_status(po.status);
  if (po.ndig > nq) return MGCD_ERR_KERNEL;
  if (ctx->device_convert && po.ndig > kMrHostDigits && !(digits && dcap)) {
    std::vector<uint32_t> mag;
    std::vector<std::vector<uint32_t>*> outs{&mag};
    const int crc = mr_convert(s, ctx->digits.as<mgcd_pair_t>(), {0}, {po.ndig}, ctx->mr, outs);
    if (crc != MGCD_OK) return crc;
    *ng = mag.size();
    if (mag.size() > cap) return MGCD_ERR_CAPACITY;
    if (!mag.empty()) std::memcpy(g, mag.data(), mag.size() * 4);
    return MGCD_OK;
  }
  std::vector<mgcd_pair_t> dig(po.ndig);
  if (po.ndig) HIP_TRY(hipMemcpy(dig.data(), ctx->digits.p, po.ndig * 8, hipMemcpyDeviceToHost));
  if (digits && dcap) std::memcpy(digits, dig.data(), std::min<size_t>(dcap, dig.size()) * 8);
  std::vector<uint32_t> mag;
  digits_to_limbs(dig.data(), dig.size(), mag);
  *ng = mag.size();
  if (mag.size() > cap) return MGCD_ERR_CAPACITY;
  if (!mag.empty()) std::memcpy(g, mag.data(), mag.size() * 4);
  return MGCD_OK;
}

// gcd_single with every device buffer in the scratch array and no host
// container: one pair, its G parts, one queue; digits converted on the
// device by the product tree, |G| copied straight into g.
int gcd_single_noalloc(mgcd_ctx* ctx, unsigned char* base, const uint32_t* u, size_t nu, const uint32_t* v,
                       size_t nv, uint32_t nq, uint32_t* g, size_t cap, size_t* ng) {
  *ng = 0;
  if (nq == 0 || nq > ctx->moduli.size()) return MGCD_ERR_NOT_ENOUGH_MODULI;
  if (nq > kScratchMaxModuli || nu + nv > kScratchMaxLimbs) return MGCD_ERR_SCRATCH;
  const int vi = choose_single_variant(ctx, nq);
  if (vi < 0) return MGCD_ERR_TOO_MANY_THREADS;
  const KernelVariant& kv = kernel_variant(vi);
  const uint32_t G = (nq + (uint32_t)(kv.T * kv.K) - 1) / (uint32_t)(kv.T * kv.K);
  if (kv.B != 1 || G > kScratchMaxParts) return MGCD_ERR_SCRATCH;
  if (G > (uint32_t)ctx->max_resident[vi]) return MGCD_ERR_TOO_MANY_THREADS;

  PairDesc pd;
  std::memset(&pd, 0, sizeof pd);
  pd.u_off = 0;
  pd.v_off = nu;
  pd.nu = (uint32_t)nu;
  pd.nv = (uint32_t)nv;
  pd.nq = nq;
  pd.trace_cap = nq + 1;
  GroupDesc gd;
  std::memset(&gd, 0, sizeof gd);
  gd.count = 1;
  gd.nq = nq;
  gd.G = G;
  WorkItem items[kScratchMaxParts];
  for (uint32_t i = 0; i < G; ++i) items[i] = WorkItem{0, i};
  LaunchSet ls;
  ls.variant = vi;
  ls.nitems = G;
  ls.nqueues = 1;
  ls.qoff[0] = 0;
  ls.qoff[1] = G;
  ls.grid = G;
  ls.nbar = G > 1 ? 2 * ((size_t)G + 1) + G : 0;

  hipStream_t s = ctx->stream;
  uint32_t* d_limbs = reinterpret_cast<uint32_t*>(base + kScratch.limbs);
  if (nu) HIP_TRY(hipMemcpyAsync(d_limbs, u, nu * 4, hipMemcpyHostToDevice, s));
  if (nv) HIP_TRY(hipMemcpyAsync(d_limbs + nu, v, nv * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(base + kScratch.pair, &pd, sizeof pd, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(base + kScratch.group, &gd, sizeof gd, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(base + kScratch.items, items, G * sizeof(WorkItem), hipMemcpyHostToDevice, s));

  KernelArgs ka;
  std::memset(&ka, 0, sizeof ka);
  ka.limbs = d_limbs;
  ka.pairs = reinterpret_cast<const PairDesc*>(base + kScratch.pair);
  ka.groups = reinterpret_cast<const GroupDesc*>(base + kScratch.group);
  ka.items = reinterpret_cast<const WorkItem*>(base + kScratch.items);
  ka.L = ctx->L;
  ka.queue = reinterpret_cast<uint32_t*>(base + kScratch.queue);
  ka.mod_u32 = ctx->d_mod_u32.as<uint32_t>();
  ka.mod_d = ctx->d_mod_d.as<double>();
  ka.mod_inv = ctx->d_mod_inv.as<double>();
  ka.bar = reinterpret_cast<unsigned long long*>(base + kScratch.bar);
  ka.digits = reinterpret_cast<mgcd_pair_t*>(base + kScratch.digits);
  ka.trace = reinterpret_cast<mgcd_pair_t*>(base + kScratch.trace);
  ka.res = reinterpret_cast<uint32_t*>(base + kScratch.res);
  ka.out = reinterpret_cast<PairOut*>(base + kScratch.out);
  ka.spin_limit = ctx->spin_limit;
  ka.relay_min_g = relay_parts(ctx, vi);
  int rc = launch(ctx, ls, ka);  // clears the queue counters and exchange slots with fill kernels
  if (rc != MGCD_OK) return rc;
  PairOut po;
  HIP_TRY(hipMemcpyAsync(&po, ka.out, sizeof po, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_s = ms * 1e-3;
  ctx->last_nq = nq;
  ctx->last_iters = po.ntrace;
  ctx->last_digits = po.ndig;
  ctx->last_T = kv.T;
  ctx->last_K = kv.K;
  ctx->last_G = (int)G;
  if (po.status != KST_OK) return kst_to_status(po.status);
  if (po.ndig > nq) return MGCD_ERR_KERNEL;
  const uint32_t k = po.ndig;
  if (k == 0) return MGCD_OK;

  // Digits -> |G| (the product tree of mr_convert, for one pair).
  MrPair mp;
  std::memset(&mp, 0, sizeof mp);
  mp.k = k;
  mp.root_level = ceil_log2(k);
  const uint32_t levels = mp.root_level;
  if (levels > kScratchMaxLevels) return MGCD_ERR_SCRATCH;
  uint32_t cum[kScratchMaxLevels * 2];
  for (uint32_t j = 0; j < levels; ++j) {
    cum[2 * j] = 0;
    cum[2 * j + 1] = (uint64_t)k > (1ull << j) ? (uint32_t)((k + (2ull << j) - 1) >> (j + 1)) : 0u;
  }
  HIP_TRY(hipMemcpyAsync(base + kScratch.mrpair, &mp, sizeof mp, hipMemcpyHostToDevice, s));
  if (levels) HIP_TRY(hipMemcpyAsync(base + kScratch.mrcum, cum, levels * 2 * 4, hipMemcpyHostToDevice, s));
  uint32_t* lv[2] = {reinterpret_cast<uint32_t*>(base + kScratch.buf0), reinterpret_cast<uint32_t*>(base + kScratch.buf1)};
  MrArgs a{};
  a.digits = ka.digits;
  a.pairs = reinterpret_cast<const MrPair*>(base + kScratch.mrpair);
  a.npairs = 1;
  a.carry = reinterpret_cast<MrCarry*>(base + kScratch.mrcarry);
  for (uint32_t j = 0; j < levels; ++j) {
    a.j = j;
    a.cum = reinterpret_cast<const uint32_t*>(base + kScratch.mrcum) + 2 * j;
    a.in = lv[j & 1];
    a.out = lv[(j + 1) & 1];
    HIP_TRY(launch_mr_level(a, cum[2 * j + 1], s));
  }
  uint32_t* d_res = reinterpret_cast<uint32_t*>(base + kScratch.mres);
  uint32_t* d_len = reinterpret_cast<uint32_t*>(base + kScratch.mlen);
  HIP_TRY(launch_mr_emit(a, lv[0], lv[1], d_res, d_len, s));
  uint32_t len = 0;
  HIP_TRY(hipMemcpyAsync(&len, d_len, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (len > k + 1) return MGCD_ERR_KERNEL;
  *ng = len;
  if (len > cap) return MGCD_ERR_CAPACITY;
  if (len) {
    HIP_TRY(hipMemcpyAsync(g, d_res, (size_t)len * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  return MGCD_OK;
}

}  // namespace

extern "C" {

int mgcd_gcd_limbs_trace(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, uint32_t* g,
                         size_t cap, size_t* ng, mgcd_pair_t* digits, size_t dcap, size_t* nd,
                         mgcd_pair_t* pivots, size_t tcap, size_t* nt) try {
  if (!ctx || (nu && !u) || (nv && !v) || !ng) return MGCD_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  nu = trim(u, nu);
  nv = trim(v, nv);
  const uint64_t bits = bit_length(u, nu);
  if (bits == 0) return MGCD_ERR_ARG;
  clear_last_pair(ctx);
  return gcd_single(ctx, u, nu, v, nv, moduli_for(ctx, bits), g, cap, ng, digits, dcap, nd, pivots, tcap, nt);
} MGCD_CATCH

int mgcd_mixed_radix_to_limbs(mgcd_ctx* ctx, const mgcd_pair_t* digits, size_t k, uint32_t* g, size_t cap,
                              size_t* ng) try {
  if (!ctx || (k && !digits) || !ng || k > 0xFFFFFFFFu) return MGCD_ERR_ARG;
  *ng = 0;
  if (k == 0) return MGCD_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->digits.ensure(k * 8) != hipSuccess) return MGCD_ERR_NOMEM;
  HIP_TRY(hipMemcpyAsync(ctx->digits.p, digits, k * 8, hipMemcpyHostToDevice, ctx->stream));
  std::vector<uint32_t> mag;
  std::vector<std::vector<uint32_t>*> outs{&mag};
  const int crc = mr_convert(ctx->stream, ctx->digits.as<mgcd_pair_t>(), {0}, {(uint32_t)k}, ctx->mr, outs);
  if (crc != MGCD_OK) return crc;
  *ng = mag.size();
  if (mag.size() > cap) return MGCD_ERR_CAPACITY;
  if (!mag.empty()) std::memcpy(g, mag.data(), mag.size() * 4);
  return MGCD_OK;
} MGCD_CATCH

int mgcd_gcd_limbs(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, uint32_t* g,
                   size_t cap, size_t* ng) try {
  if (!ctx || (nu && !u) || (nv && !v) || !ng) return MGCD_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  clear_last_pair(ctx);
  nu = trim(u, nu);
  nv = trim(v, nv);
  if (cmp_mag(u, nu, v, nv) < 0) {  // larger operand first (GmpCudaDevice-gcd.cu:61-62)
    std::swap(u, v);
    std::swap(nu, nv);
  }
  if (nv == 0) {  // gcd(u, 0) = |u|; the reference kernel cannot take v = 0
    *ng = nu;
    if (nu > cap) return MGCD_ERR_CAPACITY;
    if (nu) std::memmove(g, u, nu * 4);  // g may alias u
    return MGCD_OK;
  }
  // The reference's attempt first; where it would run out of moduli, retry
  // with a quarter more (up to the whole table) when retries are enabled.
  int nq = moduli_for(ctx, bit_length(u, nu));
  for (;;) {
    const int rc = gcd_single(ctx, u, nu, v, nv, nq, g, cap, ng, nullptr, 0, nullptr, nullptr, 0, nullptr);
    const int table = (int)ctx->moduli.size();
    if ((rc != MGCD_ERR_REDUX && rc != MGCD_ERR_RECOVERY) || !ctx->retry || ctx->force_nq || nq >= table) return rc;
    nq = std::min(table, ((nq + nq / 4 + 255) / 256) * 256);
  }
} MGCD_CATCH

int mgcd_gcd_limbs_noalloc(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, uint32_t* g,
                           size_t cap, size_t* ng) {
  if (!ctx || (nu && !u) || (nv && !v) || !ng) return MGCD_ERR_ARG;
  *ng = 0;
  HIP_TRY(hipSetDevice(ctx->device));
  unsigned char* base = scratch_of(ctx->device);
  if (!base) return MGCD_ERR_HIP;  // a context on this device took the address; none did
  clear_last_pair(ctx);
  nu = trim(u, nu);
  nv = trim(v, nv);
  if (cmp_mag(u, nu, v, nv) < 0) {  // larger operand first (GmpCudaDevice-gcd.cu:61-62)
    std::swap(u, v);
    std::swap(nu, nv);
  }
  if (nv == 0) {
    *ng = nu;
    if (nu > cap) return MGCD_ERR_CAPACITY;
    if (nu) std::memmove(g, u, nu * 4);  // g may alias u
    return MGCD_OK;
  }
  if (nu + nv > kScratchMaxLimbs) return MGCD_ERR_SCRATCH;
  uint32_t nq = (uint32_t)moduli_for(ctx, bit_length(u, nu));
  if (nq > kScratchMaxModuli) return MGCD_ERR_SCRATCH;
  // Retries as mgcd_gcd_limbs does, within the scratch limit.
  const uint32_t top = std::min<uint32_t>((uint32_t)ctx->moduli.size(), kScratchMaxModuli);
  for (;;) {
    const int rc = gcd_single_noalloc(ctx, base, u, nu, v, nv, nq, g, cap, ng);
    if ((rc != MGCD_ERR_REDUX && rc != MGCD_ERR_RECOVERY) || !ctx->retry || ctx->force_nq || nq >= top) return rc;
    nq = std::min(top, ((nq + nq / 4 + 255) / 256) * 256);
  }
}

double mgcd_ctx_last_kernel_seconds(const mgcd_ctx* ctx) { return ctx ? ctx->last_kernel_s : 0.0; }

int mgcd_ctx_phase_clocks(mgcd_ctx* ctx, uint64_t* out, size_t n, int reset) {
  if (!ctx || (n && !out)) return MGCD_ERR_ARG;
  if (!phase_clocks_built() || !ctx->clocks.p) {
    g_last_error = "not the instrumented build (lib/libmgcd_phases.so)";
    return MGCD_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  uint64_t h[kPhaseClocks];
  HIP_TRY(hipMemcpy(h, ctx->clocks.p, sizeof h, hipMemcpyDeviceToHost));
  std::memcpy(out, h, std::min<size_t>(n, kPhaseClocks) * 8);
  if (reset) HIP_TRY(hipMemset(ctx->clocks.p, 0, sizeof h));
  return MGCD_OK;
}

int mgcd_ctx_last_pair_stats(const mgcd_ctx* ctx, uint32_t* nq, uint32_t* iters, uint32_t* digits, int* workgroup,
                             int* slots, int* parts) {
  if (!ctx) return MGCD_ERR_ARG;
  if (nq) *nq = ctx->last_nq;
  if (iters) *iters = ctx->last_iters;
  if (digits) *digits = ctx->last_digits;
  if (workgroup) *workgroup = ctx->last_T;
  if (slots) *slots = ctx->last_K;
  if (parts) *parts = ctx->last_G;
  return MGCD_OK;
}

int mgcd_ctx_synchronize(mgcd_ctx* ctx) {
  if (!ctx) return MGCD_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  return MGCD_OK;
}

}  // extern "C"

// ===========================================================================
// Batches.
//
// A batch is split into size classes by N_Q; each class is one launch of
// the persistent kernel with its own tiling (explicit, tuned, or the
// default), all on the context's stream.  With MGCD_OPT_BATCH_GRAPH the
// launch sequence (queue and barrier fills, kernels) is captured into a
// hipGraph on the first run and replayed afterwards.

namespace {

// Upper N_Q bound of each size class (the last class is unbounded).
constexpr uint32_t kClassMaxNq[] = {1024, 4096, 16384, 0xFFFFFFFFu};
constexpr int kNumClasses = 4;

int class_of(uint32_t nq) {
  int c = 0;
  while (nq > kClassMaxNq[c]) ++c;
  return c;
}

// Batch tilings the tuner tries (threads, slots, pairs per group).
constexpr int kTuneCandidates[][3] = {{512, 2, 16}, {256, 2, 16}, {256, 1, 16}, {512, 1, 16},
                                     {256, 2, 8},  {256, 1, 8},  {256, 1, 32}, {128, 1, 16}};

// The built-in tiling of a size class whose pairs have these N_Q: (512, 2,
// 16) when its work items fill that tiling's co-resident grid (throughput:
// 16 pairs share each inverse, one workgroup per CU, few parts), else
// (256, 1, 16), whose four times smaller parts spread a small population
// over more CUs (latency).  Measured at 8 kbit: 256 and 1,024 pairs 54k / 169k
// against 37k / 143k GCD/s, 4,096 pairs 388k against 534k; at 30 kbit: 256
// pairs 17.1k against 11.2k, 1,024 pairs 30.0k against 41.3k.  Pairs above
// 16,384 moduli keep (512, 2, 16).
int default_variant(const mgcd_ctx* ctx, int cls, const std::map<uint32_t, size_t>& nq_count) {
  const int big = find_kernel_variant(512, 2, 16, ctx->wide);
  const int small = find_kernel_variant(256, 1, 16, ctx->wide);
  if (cls == kNumClasses - 1 || small < 0 || big < 0) return big;
  size_t items = 0;
  for (const auto& e : nq_count) items += (e.second + 15) / 16 * ((e.first + 1023) / 1024);
  return items >= (size_t)ctx->max_resident[big] ? big : small;
}

}  // namespace

struct BatchLaunch {
  int variant = -1;
  LaunchSet ls;
  size_t pair_first = 0, pair_count = 0;
  size_t item_first = 0;
  size_t queue_first = 0;  // counters of this launch's queues
  size_t bar_first = 0, nbar = 0;
  uint32_t nq_min = 0, nq_max = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double device_s = 0.0;
};

struct mgcd_batch {
  mgcd_ctx* ctx = nullptr;
  size_t npairs = 0;
  std::vector<PairDesc> pd;          // kernel pairs, launch by launch, in group order
  std::vector<int64_t> kidx;         // pair i -> kernel pair index, or -1 (handled on host)
  std::vector<std::vector<uint32_t>> host_result;  // results of pairs handled on host / converted
  std::vector<int> status;
  std::vector<BatchLaunch> launches;
  std::vector<GroupDesc> groups;
  std::vector<WorkItem> items;
  size_t ndigits = 0, ntrace = 0, nres = 0, nbar = 0;
  DevBuf limbs, pairs, groups_d, items_d, queue, bar, digits, trace, res, out;
  std::vector<PairOut> h_out;
  std::vector<mgcd_pair_t> h_digits;
  double device_s = 0.0, host_s = 0.0;
  bool ran = false;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint32_t> packed;      // host copy of the kernel pairs' limbs (for retries)
  MrBufs mr;                         // device conversion of the digits
};

namespace {

KernelArgs batch_args(const mgcd_batch* b, const BatchLaunch& l) {
  const mgcd_ctx* ctx = b->ctx;
  KernelArgs ka;
  std::memset(&ka, 0, sizeof ka);
  ka.limbs = b->limbs.as<uint32_t>();
  ka.pairs = b->pairs.as<PairDesc>();
  ka.groups = b->groups_d.as<GroupDesc>();
  ka.items = b->items_d.as<WorkItem>() + l.item_first;
  ka.L = ctx->L;
  ka.queue = b->queue.as<uint32_t>() + l.queue_first;
  ka.mod_u32 = ctx->d_mod_u32.as<uint32_t>();
  ka.mod_d = ctx->d_mod_d.as<double>();
  ka.mod_inv = ctx->d_mod_inv.as<double>();
  ka.bar = b->bar.as<unsigned long long>();
  ka.digits = b->digits.as<mgcd_pair_t>();
  ka.trace = b->trace.as<mgcd_pair_t>();
  ka.res = b->res.as<uint32_t>();
  ka.out = b->out.as<PairOut>();
  ka.spin_limit = ctx->spin_limit;
  ka.relay_min_g = relay_parts(ctx, l.variant);
  ka.clocks = ctx->clocks.as<unsigned long long>();
  ka.nqueues = l.ls.nqueues;
  std::memcpy(ka.qoff, l.ls.qoff, sizeof ka.qoff);
  return ka;
}

// Queues the fills and the kernel of one launch on ctx->stream.  A
// cooperative launch guarantees that every workgroup is resident; inside a
// graph capture the kernel is launched plainly with the same grid, which
// never exceeds the device's resident capacity for the variant.
int enqueue_launch(mgcd_batch* b, const BatchLaunch& l, bool cooperative) {
  mgcd_ctx* ctx = b->ctx;
  KernelArgs ka = batch_args(b, l);
  HIP_TRY(launch_fill_u32(ka.queue, kMaxQueues * kQueueStride, 0u, ctx->stream));
  if (l.nbar) HIP_TRY(launch_fill_u64(ka.bar + l.bar_first, l.nbar, 0ull, ctx->stream));
  const KernelVariant& kv = kernel_variant(l.variant);
  void* params[] = {&ka};
  HIP_TRY(launch_resident(kv.fn, l.ls.grid, kv.T, params, ctx->stream, cooperative));
  return MGCD_OK;
}

// Plans the batch's launches: one per size class, or a single one for an
// explicit variant.  b->pd is reordered launch by launch.
int plan_batch(mgcd_batch* b, int forced_variant, const std::vector<int>& class_variant) {
  mgcd_ctx* ctx = b->ctx;
  std::vector<std::vector<uint32_t>> byclass(forced_variant >= 0 ? 1 : kNumClasses);
  for (uint32_t j = 0; j < b->pd.size(); ++j) byclass[forced_variant >= 0 ? 0 : class_of(b->pd[j].nq)].push_back(j);
  // Adjacent classes with the same tiling share one launch: its work queue
  // then balances them against each other.
  std::vector<std::vector<uint32_t>> members;
  std::vector<int> set_variant;
  for (size_t c = 0; c < byclass.size(); ++c) {
    if (byclass[c].empty()) continue;
    const int v = forced_variant >= 0 ? forced_variant : class_variant[c];
    if (v < 0) return MGCD_ERR_ARG;
    if (!set_variant.empty() && set_variant.back() == v) {
      members.back().insert(members.back().end(), byclass[c].begin(), byclass[c].end());
    } else {
      members.push_back(byclass[c]);
      set_variant.push_back(v);
    }
  }
  std::vector<PairDesc> all;
  std::vector<int64_t> pos(b->pd.size(), -1);
  size_t dig = 0, tr = 0, rs = 0, bar = 0;
  for (size_t c = 0; c < members.size(); ++c) {
    BatchLaunch l;
    l.variant = set_variant[c];
    std::vector<PairDesc> cpd;
    for (uint32_t j : members[c]) cpd.push_back(b->pd[j]);
    std::vector<uint32_t> order;
    std::vector<GroupDesc> cg;
    std::vector<WorkItem> ci;
    const int rc = plan_items(ctx, l.variant, cpd, order, cg, ci, l.ls);
    if (rc != MGCD_OK) return rc;
    l.pair_first = all.size();
    l.pair_count = cpd.size();
    l.item_first = b->items.size();
    l.queue_first = b->launches.size() * (size_t)kMaxQueues * kQueueStride;
    l.bar_first = bar;
    l.nbar = l.ls.nbar;
    l.nq_min = 0xFFFFFFFFu;
    for (PairDesc& p : cpd) {
      p.dig_off += dig;
      p.trace_off += tr;
      p.res_off += rs;
      l.nq_min = std::min(l.nq_min, p.nq);
      l.nq_max = std::max(l.nq_max, p.nq);
    }
    const size_t group_first = b->groups.size();
    for (GroupDesc& g : cg) {
      g.first += (uint32_t)l.pair_first;
      g.bar_off += bar;
      b->groups.push_back(g);
    }
    for (WorkItem& w : ci) {
      w.group += (uint32_t)group_first;
      b->items.push_back(w);
    }
    for (size_t j = 0; j < order.size(); ++j) pos[members[c][order[j]]] = (int64_t)(l.pair_first + j);
    all.insert(all.end(), cpd.begin(), cpd.end());
    dig += l.ls.ndigits;
    tr += l.ls.ntrace;
    rs += l.ls.nres;
    bar += l.ls.nbar;
    b->launches.push_back(l);
  }
  b->pd.swap(all);
  for (auto& k : b->kidx)
    if (k >= 0) k = pos[k];
  b->ndigits = dig;
  b->ntrace = tr;
  b->nres = rs;
  b->nbar = bar;
  return MGCD_OK;
}

// Key of a tuned tiling: width L, size class and the class's population
// (power of two, up to 2^12): how many groups a class fills the device
// with decides between tilings as much as the pairs' size does.
std::string tune_key(int L, int cls, size_t count) {
  int p = 0;
  while (p < 12 && (size_t(2) << p) <= count) ++p;
  return std::to_string(L) + ":" + std::to_string(cls) + ":" + std::to_string(p);
}

void load_tune_cache(mgcd_ctx* ctx) {
  const char* path = std::getenv("MGCD_TUNE_CACHE");
  if (!path || !*path || ctx->tune_loaded) return;
  ctx->tune_loaded = true;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  char key[64];
  int T, K, B;
  while (std::fscanf(f, "%63s %d %d %d", key, &T, &K, &B) == 4) {
    const int vi = find_kernel_variant(T, K, B, ctx->wide);
    if (vi >= 0) ctx->tuned[key] = vi;
  }
  std::fclose(f);
}

void save_tune_cache(const mgcd_ctx* ctx) {
  const char* path = std::getenv("MGCD_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "w");
  if (!f) return;
  for (const auto& kv : ctx->tuned) {
    const KernelVariant& v = kernel_variant(kv.second);
    std::fprintf(f, "%s %d %d %d\n", kv.first.c_str(), v.T, v.K, v.B);
  }
  std::fclose(f);
}

}  // namespace

extern "C" {

int mgcd_batch_create(mgcd_ctx* ctx, const uint32_t* limbs, const uint64_t* u_off, const uint32_t* nu,
                      const uint64_t* v_off, const uint32_t* nv, size_t npairs, int opts_workgroup, int opts_slots,
                      int opts_group, mgcd_batch** out) try {
  if (!ctx || !out || (npairs && (!limbs || !u_off || !nu || !v_off || !nv))) return MGCD_ERR_ARG;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  mgcd_batch* b = new mgcd_batch();
  b->ctx = ctx;
  b->npairs = npairs;
  b->use_graph = ctx->batch_graph;
  b->kidx.assign(npairs, -1);
  b->host_result.resize(npairs);
  b->status.assign(npairs, MGCD_OK);
  // Pack the kernel pairs' limbs, larger operand first.
  std::vector<uint32_t>& packed = b->packed;
  for (size_t i = 0; i < npairs; ++i) {
    const uint32_t* a = limbs + u_off[i];
    const uint32_t* c = limbs + v_off[i];
    size_t na = trim(a, nu[i]), nc = trim(c, nv[i]);
    if (cmp_mag(a, na, c, nc) < 0) {
      std::swap(a, c);
      std::swap(na, nc);
    }
    if (nc == 0) {  // gcd(u, 0) = |u|
      b->host_result[i].assign(a, a + na);
      continue;
    }
    const int nqi = mgcd_num_moduli_needed(std::max<uint64_t>(bit_length(a, na), 2), ctx->L);
    if (nqi <= 0 || (size_t)nqi > ctx->moduli.size()) {
      b->status[i] = MGCD_ERR_NOT_ENOUGH_MODULI;
      continue;
    }
    PairDesc p;
    std::memset(&p, 0, sizeof p);
    p.u_off = packed.size();
    packed.insert(packed.end(), a, a + na);
    p.v_off = packed.size();
    packed.insert(packed.end(), c, c + nc);
    p.nu = (uint32_t)na;
    p.nv = (uint32_t)nc;
    p.nq = (uint32_t)nqi;
    b->kidx[i] = (int64_t)b->pd.size();
    b->pd.push_back(p);
  }
  int rc = MGCD_OK;
  int forced = -1;
  std::vector<int> class_variant(kNumClasses, -1);
  if (opts_workgroup > 0 && opts_slots > 0 && opts_group > 0) {
    forced = find_kernel_variant(opts_workgroup, opts_slots, opts_group, ctx->wide);
    if (forced < 0) rc = MGCD_ERR_ARG;
  } else if (opts_workgroup || opts_slots || opts_group) {
    rc = MGCD_ERR_ARG;
  } else {
    load_tune_cache(ctx);
    std::vector<std::vector<size_t>> byclass(kNumClasses);
    for (size_t i = 0; i < npairs; ++i)
      if (b->kidx[i] >= 0) byclass[class_of(b->pd[b->kidx[i]].nq)].push_back(i);
    bool tuned_any = false;
    auto pick = [&](int c) {
      const std::string key = tune_key(ctx->L, c, byclass[c].size());
      const auto it = ctx->tuned.find(key);
      if (it != ctx->tuned.end()) {
        class_variant[c] = it->second;
      } else if (ctx->autotune) {
        // The class itself, or an evenly spaced sample of 4096 of its pairs:
        // a smaller sample fills the device with fewer groups than the class
        // does, which favours the small-group tilings.
        const size_t want = std::min<size_t>(byclass[c].size(), 4096);
        std::vector<uint64_t> so, to;
        std::vector<uint32_t> sn, tn;
        for (size_t s = 0; s < want; ++s) {
          const size_t i = byclass[c][s * byclass[c].size() / want];
          so.push_back(u_off[i]);
          sn.push_back(nu[i]);
          to.push_back(v_off[i]);
          tn.push_back(nv[i]);
        }
        int T, K, Bq;
        rc = mgcd_autotune(ctx, limbs, so.data(), sn.data(), to.data(), tn.data(), want, &T, &K, &Bq);
        if (rc == MGCD_OK) {
          class_variant[c] = find_kernel_variant(T, K, Bq, ctx->wide);
          ctx->tuned[key] = class_variant[c];
          tuned_any = true;
        }
      } else {
        std::map<uint32_t, size_t> nq_count;
        for (size_t i : byclass[c]) nq_count[b->pd[b->kidx[i]].nq]++;
        class_variant[c] = default_variant(ctx, c, nq_count);
      }
    };
    // Elimination work of each class, about N_Q^2 per pair.  A class with
    // under a tenth of the largest one's work takes that class's tiling: it
    // then shares its launch, filling the workgroups that the large items
    // leave idle, instead of running in a launch of its own.
    std::vector<double> work(kNumClasses, 0.0);
    int cmax = -1;
    for (int c = 0; c < kNumClasses; ++c) {
      for (size_t i : byclass[c]) work[c] += (double)b->pd[b->kidx[i]].nq * (double)b->pd[b->kidx[i]].nq;
      if (!byclass[c].empty() && (cmax < 0 || work[c] > work[cmax])) cmax = c;
    }
    // (Only where that tiling can split the class's largest pair over
    // co-resident workgroups.)
    auto fits = [&](int c, int vi) {
      if (vi < 0) return false;
      const KernelVariant& kv = kernel_variant(vi);
      uint32_t nqmax = 0;
      for (size_t i : byclass[c]) nqmax = std::max(nqmax, b->pd[b->kidx[i]].nq);
      const uint32_t per_wg = (uint32_t)(kv.T * kv.K);
      return (nqmax + per_wg - 1) / per_wg <= (uint32_t)ctx->max_resident[vi];
    };
    if (cmax >= 0) pick(cmax);
    for (int c = 0; c < kNumClasses && rc == MGCD_OK; ++c) {
      if (byclass[c].empty() || c == cmax) continue;
      if (work[c] < 0.1 * work[cmax] && fits(c, class_variant[cmax])) class_variant[c] = class_variant[cmax];
      else pick(c);
    }
    if (tuned_any) save_tune_cache(ctx);
  }
  if (rc == MGCD_OK && !b->pd.empty()) rc = plan_batch(b, forced, class_variant);
  if (rc == MGCD_OK && !b->pd.empty()) {
    const size_t np = b->pd.size();
    if (b->limbs.ensure(packed.size() * 4 + 4) != hipSuccess || b->pairs.ensure(np * sizeof(PairDesc)) != hipSuccess ||
        b->groups_d.ensure(b->groups.size() * sizeof(GroupDesc)) != hipSuccess ||
        b->items_d.ensure(b->items.size() * sizeof(WorkItem)) != hipSuccess ||
        b->queue.ensure(b->launches.size() * kMaxQueues * kQueueStride * 4) != hipSuccess ||
        b->bar.ensure((b->nbar + 1) * 8) != hipSuccess || b->digits.ensure(b->ndigits * 8) != hipSuccess ||
        b->trace.ensure(b->ntrace * 8) != hipSuccess || b->res.ensure(b->nres * 4) != hipSuccess ||
        b->out.ensure(np * sizeof(PairOut)) != hipSuccess) {
      rc = MGCD_ERR_NOMEM;
    } else if (hipMemcpy(b->limbs.p, packed.data(), packed.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->pairs.p, b->pd.data(), np * sizeof(PairDesc), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->groups_d.p, b->groups.data(), b->groups.size() * sizeof(GroupDesc),
                         hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(b->items_d.p, b->items.data(), b->items.size() * sizeof(WorkItem), hipMemcpyHostToDevice) !=
                   hipSuccess) {
      rc = MGCD_ERR_HIP;
    }
    for (BatchLaunch& l : b->launches)
      if (rc == MGCD_OK && (hipEventCreate(&l.e0) != hipSuccess || hipEventCreate(&l.e1) != hipSuccess))
        rc = MGCD_ERR_HIP;
  }
  if (rc != MGCD_OK) {
    mgcd_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MGCD_OK;
} MGCD_CATCH

int mgcd_batch_run(mgcd_batch* b) try {
  if (!b) return MGCD_ERR_ARG;
  mgcd_ctx* ctx = b->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t np = b->pd.size();
  b->device_s = 0.0;
  b->host_s = 0.0;
  if (np) {
    hipStream_t s = ctx->stream;
    if (b->use_graph) {
      if (!b->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = MGCD_OK;
        for (const BatchLaunch& l : b->launches)
          if (rc == MGCD_OK) rc = enqueue_launch(b, l, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != MGCD_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
      }
      HIP_TRY(hipEventRecord(ctx->ev0, s));
      HIP_TRY(hipGraphLaunch(b->graph_exec, s));
      HIP_TRY(hipEventRecord(ctx->ev1, s));
    } else {
      HIP_TRY(hipEventRecord(ctx->ev0, s));
      for (const BatchLaunch& l : b->launches) {
        HIP_TRY(hipEventRecord(l.e0, s));
        const int rc = enqueue_launch(b, l, ctx->cooperative);
        if (rc != MGCD_OK) return rc;
        HIP_TRY(hipEventRecord(l.e1, s));
      }
      HIP_TRY(hipEventRecord(ctx->ev1, s));
    }
    b->h_out.resize(np);
    HIP_TRY(hipMemcpyAsync(b->h_out.data(), b->out.p, np * sizeof(PairOut), hipMemcpyDeviceToHost, s));
    if (!ctx->device_convert) {
      b->h_digits.resize(b->ndigits);
      HIP_TRY(hipMemcpyAsync(b->h_digits.data(), b->digits.p, b->ndigits * 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) b->device_s = ms * 1e-3;
    for (BatchLaunch& l : b->launches) {
      l.device_s = 0.0;
      if (!b->use_graph && hipEventElapsedTime(&ms, l.e0, l.e1) == hipSuccess) l.device_s = ms * 1e-3;
    }
    ctx->last_kernel_s = b->device_s;
    // Digit conversion: on the device (only |G| comes back), or on host
    // threads from the copied digits.
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<size_t> todo;
    for (size_t i = 0; i < b->npairs; ++i)
      if (b->kidx[i] >= 0) todo.push_back(i);
    if (ctx->device_convert) {
      std::vector<uint64_t> doff;
      std::vector<uint32_t> ks;
      std::vector<std::vector<uint32_t>*> outs;
      for (size_t i : todo) {
        const PairDesc& p = b->pd[b->kidx[i]];
        const PairOut& o = b->h_out[b->kidx[i]];
        if (o.status != KST_OK || o.ndig > p.nq) {
          b->status[i] = o.status != KST_OK ? kst_to_status(o.status) : MGCD_ERR_KERNEL;
          b->host_result[i].clear();
          continue;
        }
        b->status[i] = MGCD_OK;
        doff.push_back(p.dig_off);
        ks.push_back(o.ndig);
        outs.push_back(&b->host_result[i]);
      }
      const int crc = mr_convert(s, b->digits.as<mgcd_pair_t>(), doff, ks, b->mr, outs);
      if (crc != MGCD_OK) return crc;
    }

    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (;;) {
        const size_t j = next.fetch_add(1);
        if (j >= todo.size()) break;
        const size_t i = todo[j];
        const PairDesc& p = b->pd[b->kidx[i]];
        const PairOut& o = b->h_out[b->kidx[i]];
        if (o.status != KST_OK || o.ndig > p.nq) {
          b->status[i] = o.status != KST_OK ? kst_to_status(o.status) : MGCD_ERR_KERNEL;
          b->host_result[i].clear();
          continue;
        }
        b->status[i] = MGCD_OK;
        digits_to_limbs(b->h_digits.data() + p.dig_off, o.ndig, b->host_result[i]);
      }
    };
    if (!ctx->device_convert) {
      unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      if (todo.size() < 64) nth = 1;
      std::vector<std::thread> th;
      for (unsigned k = 1; k < nth; ++k) th.emplace_back(work);
      work();
      for (auto& x : th) x.join();
    }
    b->host_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // Pairs the reference's sizing left short of moduli: retry one by one
    // with a quarter more moduli each time (MGCD_OPT_RETRY).  A pair whose
    // in-kernel wait timed out is run again on the one-pair path.
    for (size_t i : todo) {
      const int st = b->status[i];
      const bool more = ctx->retry && !ctx->force_nq && (st == MGCD_ERR_REDUX || st == MGCD_ERR_RECOVERY);
      if (!more && st != MGCD_ERR_TIMEOUT) continue;
      const PairDesc& p = b->pd[b->kidx[i]];
      const uint32_t* a = b->packed.data() + p.u_off;
      const uint32_t* c = b->packed.data() + p.v_off;
      const int table = (int)ctx->moduli.size();
      int nq = (int)p.nq, rc = st;
      std::vector<uint32_t> gl(p.nv + 1);
      size_t ng = 0;
      if (st == MGCD_ERR_TIMEOUT)
        rc = gcd_single(ctx, a, p.nu, c, p.nv, nq, gl.data(), gl.size(), &ng, nullptr, 0, nullptr, nullptr, 0,
                        nullptr);
      while (ctx->retry && !ctx->force_nq && (rc == MGCD_ERR_REDUX || rc == MGCD_ERR_RECOVERY) && nq < table) {
        nq = std::min(table, ((nq + nq / 4 + 255) / 256) * 256);
        rc = gcd_single(ctx, a, p.nu, c, p.nv, nq, gl.data(), gl.size(), &ng, nullptr, 0, nullptr, nullptr, 0,
                        nullptr);
      }
      b->status[i] = rc;
      b->host_result[i].assign(gl.begin(), gl.begin() + (rc == MGCD_OK ? ng : 0));
    }
  }
  b->ran = true;
  return MGCD_OK;
} MGCD_CATCH

int mgcd_batch_status(const mgcd_batch* b, size_t i) {
  if (!b || i >= b->npairs) return MGCD_ERR_ARG;
  if (!b->ran && b->kidx[i] >= 0) return MGCD_ERR_KERNEL;
  return b->status[i];
}

int mgcd_batch_result(const mgcd_batch* b, size_t i, uint32_t* g, size_t cap, size_t* ng) {
  if (!b || i >= b->npairs || !ng) return MGCD_ERR_ARG;
  const int st = mgcd_batch_status(b, i);
  if (st != MGCD_OK) return st;
  const auto& r = b->host_result[i];
  *ng = r.size();
  if (r.size() > cap) return MGCD_ERR_CAPACITY;
  if (!r.empty()) std::memcpy(g, r.data(), r.size() * 4);
  return MGCD_OK;
}

int mgcd_batch_timing(const mgcd_batch* b, double* device_s, double* host_s) {
  if (!b) return MGCD_ERR_ARG;
  if (device_s) *device_s = b->device_s;
  if (host_s) *host_s = b->host_s;
  return MGCD_OK;
}

int mgcd_batch_pivots(const mgcd_batch* b, size_t i, mgcd_pair_t* pivots, size_t cap, size_t* n) try {
  if (!b || i >= b->npairs || !n || (cap && !pivots)) return MGCD_ERR_ARG;
  *n = 0;
  if (!b->ran || b->kidx[i] < 0) return MGCD_ERR_ARG;
  const PairDesc& p = b->pd[b->kidx[i]];
  const size_t nt = std::min<size_t>(b->h_out[b->kidx[i]].ntrace, p.trace_cap);
  *n = nt;
  const size_t c = std::min(nt, cap);
  if (c) {
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipMemcpy(pivots, b->trace.as<mgcd_pair_t>() + p.trace_off, c * sizeof(mgcd_pair_t),
                      hipMemcpyDeviceToHost));
  }
  return MGCD_OK;
} MGCD_CATCH

int mgcd_batch_pair_stats(const mgcd_batch* b, size_t i, uint32_t* nq, uint32_t* iters, uint32_t* digits) {
  if (!b || i >= b->npairs) return MGCD_ERR_ARG;
  uint32_t q = 0, r = 0, k = 0;
  if (b->kidx[i] >= 0) {
    q = b->pd[b->kidx[i]].nq;
    if (b->ran) {
      r = b->h_out[b->kidx[i]].ntrace;
      k = b->h_out[b->kidx[i]].ndig;
    }
  }
  if (nq) *nq = q;
  if (iters) *iters = r;
  if (digits) *digits = k;
  return MGCD_OK;
}

int mgcd_batch_tiling(const mgcd_batch* b, int* workgroup, int* slots, int* group) {
  if (!b) return MGCD_ERR_ARG;
  // The launch with the most pairs.
  const BatchLaunch* best = nullptr;
  for (const BatchLaunch& l : b->launches)
    if (!best || l.pair_count > best->pair_count) best = &l;
  const KernelVariant* kv = best ? &kernel_variant(best->variant) : nullptr;
  if (workgroup) *workgroup = kv ? kv->T : 0;
  if (slots) *slots = kv ? kv->K : 0;
  if (group) *group = kv ? kv->B : 0;
  return MGCD_OK;
}

int mgcd_batch_launch_count(const mgcd_batch* b, size_t* count) {
  if (!b || !count) return MGCD_ERR_ARG;
  *count = b->launches.size();
  return MGCD_OK;
}

int mgcd_batch_launch_info(const mgcd_batch* b, size_t i, int* workgroup, int* slots, int* group, size_t* npairs,
                           uint32_t* nq_min, uint32_t* nq_max, double* device_s) {
  if (!b || i >= b->launches.size()) return MGCD_ERR_ARG;
  const BatchLaunch& l = b->launches[i];
  const KernelVariant& kv = kernel_variant(l.variant);
  if (workgroup) *workgroup = kv.T;
  if (slots) *slots = kv.K;
  if (group) *group = kv.B;
  if (npairs) *npairs = l.pair_count;
  if (nq_min) *nq_min = l.nq_min;
  if (nq_max) *nq_max = l.nq_max;
  if (device_s) *device_s = l.device_s;
  return MGCD_OK;
}

int mgcd_batch_destroy(mgcd_batch* b) {
  if (!b) return MGCD_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (BatchLaunch& l : b->launches) {
    if (l.e0) (void)hipEventDestroy(l.e0);
    if (l.e1) (void)hipEventDestroy(l.e1);
  }
  for (DevBuf* d : {&b->limbs, &b->pairs, &b->groups_d, &b->items_d, &b->queue, &b->bar, &b->digits, &b->trace, &b->res,
                    &b->out})
    d->release();
  b->mr.release();
  delete b;
  return MGCD_OK;
}

// Measures each candidate tiling on the given pairs (one warm-up and one
// timed run of a batch with that tiling) and returns the fastest.
int mgcd_autotune(mgcd_ctx* ctx, const uint32_t* limbs, const uint64_t* u_off, const uint32_t* nu,
                  const uint64_t* v_off, const uint32_t* nv, size_t npairs, int* workgroup, int* slots, int* group) try {
  if (!ctx || !workgroup || !slots || !group || npairs == 0) return MGCD_ERR_ARG;
  double best_t = 0.0;
  int best = -1;
  const bool graph = ctx->batch_graph;
  ctx->batch_graph = false;
  for (const auto& c : kTuneCandidates) {
    const int vi = find_kernel_variant(c[0], c[1], c[2], ctx->wide);
    if (vi < 0) continue;
    mgcd_batch* b = nullptr;
    if (mgcd_batch_create(ctx, limbs, u_off, nu, v_off, nv, npairs, c[0], c[1], c[2], &b) != MGCD_OK) continue;
    double t = 0.0;
    bool ok = mgcd_batch_run(b) == MGCD_OK;
    if (ok) ok = mgcd_batch_run(b) == MGCD_OK;
    if (ok) {
      t = b->device_s;
      for (size_t i = 0; i < npairs; ++i) ok = ok && b->status[i] == MGCD_OK;
    }
    mgcd_batch_destroy(b);
    // Earlier candidates are kept unless a later one is clearly (3 %) faster.
    if (ok && (best < 0 || t < 0.97 * best_t)) {
      best_t = t;
      best = vi;
    }
  }
  ctx->batch_graph = graph;
  if (best < 0) return MGCD_ERR_KERNEL;
  const KernelVariant& kv = kernel_variant(best);
  *workgroup = kv.T;
  *slots = kv.K;
  *group = kv.B;
  return MGCD_OK;
} MGCD_CATCH

// ---------------------------------------------------------------------------
int mgcd_certify_quasi_quotient(mgcd_ctx* ctx, int mode, int L, uint64_t* failures, uint64_t* checked) try {
  if (!ctx || !failures || !checked || mode < 0 || mode > 2 || (mode == 1 && (L < 2 || L > 32)))
    return MGCD_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  DevBuf cnt;
  if (cnt.ensure(16) != hipSuccess) return MGCD_ERR_NOMEM;
  unsigned long long* d = cnt.as<unsigned long long>();
  hipStream_t s = ctx->stream;
  int rc = MGCD_OK;
  auto run = [&]() -> int {
    HIP_TRY(launch_fill_u64(d, 2, 0ull, s));
    const int grid = ctx->num_cus * 16;
    if (mode == 0) {
      HIP_TRY(launch_certify_float(d, d + 1, grid, s));
    } else if (mode == 2) {
      HIP_TRY(launch_certify_rcp(d, d + 1, grid, s));
    } else {
      for (int b = 0; b < L; ++b) {
        const uint64_t lo = 1ull << b;           // octave [2^b, 2^(b+1)) of y
        const uint64_t hi = 1ull << (b + 1);      // b + 1 <= L: all y < 2^L
        HIP_TRY(launch_certify_endpoints(L, lo, hi, d, d + 1, grid, s));
      }
    }
    unsigned long long h[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *failures = h[0];
    *checked = h[1];
    return MGCD_OK;
  };
  rc = run();
  cnt.release();
  return rc;
} MGCD_CATCH

}  // extern "C"

// ===========================================================================
// Serving: a request loop that admits GCD requests of mixed sizes against a
// device-memory budget, batches whatever is waiting into the next launch,
// and keeps several launches in flight on their own streams so that work
// admitted while others run overlaps with them on the device.

namespace {

struct ServerReq {
  uint64_t ticket = 0;
  std::vector<uint32_t> u, v;  // trimmed, |u| >= |v| > 0
  uint32_t nq = 0;
  uint64_t bytes = 0;          // device footprint while resident
  std::chrono::steady_clock::time_point t_submit;
  double latency = 0.0;
  int status = MGCD_OK;
  std::vector<uint32_t> result;
  bool done = false;
};

// One launch slot: a stream, its device buffers and the requests in flight.
struct ServerSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  DevBuf limbs, pairs, groups, items, queue, bar, digits, trace, res, out;
  // Pinned staging, so that neither direction of copy blocks the worker
  // while other slots' kernels run.
  PinnedBuf h_in, h_out;
  MrBufs mr;  // conversion of the digits on the device
  // Buffers outgrown while serving, freed with the server: hipFree and
  // hipHostFree wait for every stream, which would stall this slot behind
  // the other slots' kernels.
  std::vector<void*> retired_dev, retired_host;
  std::vector<ServerReq*> reqs;
  std::vector<PairDesc> pd;
  std::vector<uint32_t> order;
  LaunchSet ls;
  uint64_t bytes = 0;
  bool busy = false;
  int cls = 0;  // size class it serves: 0 small, 1 large
  int variant = -1;
  bool own_stream = false;
  std::chrono::steady_clock::time_point t_launch;
};

// Requests of more than this many moduli (above ~200 kbit) are served by
// their own slot, so that small requests never wait behind them in a launch.
constexpr uint32_t kServeLargeNq = 16384;

// Device bytes a pair holds while resident: limbs, residues, pivots,
// digits, descriptors.
uint64_t footprint(size_t nu, size_t nv, uint32_t nq) {
  return 4ull * (nu + nv) + 24ull * (nq + 64) + sizeof(PairDesc) + sizeof(PairOut) + sizeof(GroupDesc) +
         32 * sizeof(WorkItem);
}

}  // namespace

struct mgcd_server {
  mgcd_ctx* ctx = nullptr;
  uint64_t budget = 0, used = 0, peak = 0;
  int variant = -1;
  std::vector<ServerSlot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<ServerReq*> pending[2];  // by size class
  std::unordered_map<uint64_t, ServerReq*> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  bool trace = false;  // $MGCD_SERVER_TRACE: one stderr line per launch and completion
  std::thread worker;
  std::vector<double> latencies;
  uint64_t completed = 0;
  std::chrono::steady_clock::time_point t_first, t_last;
  bool any = false;
};

namespace {

// Launches slot `sl` with the requests in sl.reqs (worker thread, no lock held).
int server_launch(mgcd_server* s, ServerSlot& sl) {
  mgcd_ctx* ctx = s->ctx;
  std::vector<uint32_t> packed;
  sl.pd.clear();
  // Kernel pair j is request order[j]; until the plan reorders them, the
  // identity (a launch that fails before planning answers every request).
  sl.order.resize(sl.reqs.size());
  for (uint32_t j = 0; j < sl.order.size(); ++j) sl.order[j] = j;
  for (ServerReq* r : sl.reqs) {
    PairDesc p;
    std::memset(&p, 0, sizeof p);
    p.u_off = packed.size();
    packed.insert(packed.end(), r->u.begin(), r->u.end());
    p.v_off = packed.size();
    packed.insert(packed.end(), r->v.begin(), r->v.end());
    p.nu = (uint32_t)r->u.size();
    p.nv = (uint32_t)r->v.size();
    p.nq = r->nq;
    sl.pd.push_back(p);
  }
  std::vector<GroupDesc> groups;
  std::vector<WorkItem> items;
  sl.ls = LaunchSet();
  // A few large requests alone: the one-pair tiling with the fewest moduli
  // per thread whose workgroup count fits this slot's share of the device
  // (lower latency per iteration); otherwise groups of 16 sharing their
  // inverses.
  int vi = s->variant;
  if (sl.cls == 1 && sl.reqs.size() < 4) {
    uint32_t nq_max = 0;
    for (const ServerReq* r : sl.reqs) nq_max = std::max(nq_max, r->nq);
    for (int K : {1, 2, 4, 8}) {
      const int v = find_kernel_variant(256, K, 1, ctx->wide);
      if (v < 0) continue;
      const uint32_t G = (nq_max + 256 * K - 1) / (256 * K);
      if (G <= (uint32_t)(ctx->max_resident[v] / (int)s->slots.size())) {
        vi = v;
        break;
      }
    }
  }
  sl.variant = vi;
  // Every slot's grid fits the device at the same time as the others': the
  // plan (grid, queues) is made for this slot's share.
  const uint32_t share = (uint32_t)(ctx->max_resident[vi] / (int)s->slots.size());
  if (share == 0) return MGCD_ERR_TOO_MANY_THREADS;
  int rc = plan_items(ctx, vi, sl.pd, sl.order, groups, items, sl.ls, share);
  if (rc != MGCD_OK) return rc;
  const size_t np = sl.pd.size();
  const size_t in_bytes = packed.size() * 4 + np * sizeof(PairDesc) + groups.size() * sizeof(GroupDesc) +
                          items.size() * sizeof(WorkItem) + 64;
  std::vector<void*>* rh = &sl.retired_host;
  std::vector<void*>* rd = &sl.retired_dev;
  if (sl.h_in.ensure(in_bytes, rh) != hipSuccess || sl.h_out.ensure(np * sizeof(PairOut), rh) != hipSuccess)
    return MGCD_ERR_NOMEM;
  if (sl.limbs.ensure(packed.size() * 4 + 4, rd) != hipSuccess ||
      sl.pairs.ensure(np * sizeof(PairDesc), rd) != hipSuccess ||
      sl.groups.ensure(groups.size() * sizeof(GroupDesc), rd) != hipSuccess ||
      sl.items.ensure(items.size() * sizeof(WorkItem), rd) != hipSuccess ||
      sl.queue.ensure(kMaxQueues * kQueueStride * 4, rd) != hipSuccess ||
      sl.bar.ensure((sl.ls.nbar + 1) * 8, rd) != hipSuccess || sl.digits.ensure(sl.ls.ndigits * 8, rd) != hipSuccess ||
      sl.trace.ensure(sl.ls.ntrace * 8, rd) != hipSuccess || sl.res.ensure(sl.ls.nres * 4, rd) != hipSuccess ||
      sl.out.ensure(np * sizeof(PairOut), rd) != hipSuccess)
    return MGCD_ERR_NOMEM;
  hipStream_t st = sl.stream;
  char* hin = sl.h_in.as<char>();
  auto stage = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    std::memcpy(hin, src, bytes);
    const hipError_t e = hipMemcpyAsync(dst, hin, bytes, hipMemcpyHostToDevice, st);
    hin += (bytes + 15) & ~(size_t)15;
    return e;
  };
  HIP_TRY(stage(sl.limbs.p, packed.data(), packed.size() * 4));
  HIP_TRY(stage(sl.pairs.p, sl.pd.data(), np * sizeof(PairDesc)));
  HIP_TRY(stage(sl.groups.p, groups.data(), groups.size() * sizeof(GroupDesc)));
  HIP_TRY(stage(sl.items.p, items.data(), items.size() * sizeof(WorkItem)));
  KernelArgs ka;
  std::memset(&ka, 0, sizeof ka);
  ka.limbs = sl.limbs.as<uint32_t>();
  ka.pairs = sl.pairs.as<PairDesc>();
  ka.groups = sl.groups.as<GroupDesc>();
  ka.items = sl.items.as<WorkItem>();
  ka.L = ctx->L;
  ka.queue = sl.queue.as<uint32_t>();
  ka.mod_u32 = ctx->d_mod_u32.as<uint32_t>();
  ka.mod_d = ctx->d_mod_d.as<double>();
  ka.mod_inv = ctx->d_mod_inv.as<double>();
  ka.bar = sl.bar.as<unsigned long long>();
  ka.digits = sl.digits.as<mgcd_pair_t>();
  ka.trace = sl.trace.as<mgcd_pair_t>();
  ka.res = sl.res.as<uint32_t>();
  ka.out = sl.out.as<PairOut>();
  ka.spin_limit = ctx->spin_limit;
  ka.relay_min_g = relay_parts(ctx, vi);
  ka.nqueues = sl.ls.nqueues;
  std::memcpy(ka.qoff, sl.ls.qoff, sizeof ka.qoff);
  HIP_TRY(launch_fill_u32(ka.queue, kMaxQueues * kQueueStride, 0u, st));
  if (sl.ls.nbar) HIP_TRY(launch_fill_u64(ka.bar, sl.ls.nbar, 0ull, st));
  // Plain launch: the slots' grids together fit the device, so all of their
  // workgroups are resident at once while slots overlap.
  const KernelVariant& kv = kernel_variant(vi);
  void* params[] = {&ka};
  HIP_TRY(launch_resident(kv.fn, sl.ls.grid, kv.T, params, st, false));
  HIP_TRY(hipMemcpyAsync(sl.h_out.p, sl.out.p, np * sizeof(PairOut), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(sl.done_ev, st));
  return MGCD_OK;
}

// Results of a finished slot (worker thread, no lock held).  Pairs the
// batch could not finish (moduli short, a wait timed out) go through the
// one-pair path with the retry policy.
void server_finish(mgcd_server* s, ServerSlot& sl, int launch_rc) {
  mgcd_ctx* ctx = s->ctx;
  // The digits of the finished pairs become |G| on the device (the product
  // tree batches use); only |G| comes back.
  // A launch that failed answers every request (its descriptors may be
  // incomplete); a finished one, kernel pair j = request order[j].
  const bool ran = launch_rc == MGCD_OK && sl.pd.size() == sl.reqs.size() && sl.order.size() == sl.reqs.size();
  std::vector<int> sts(sl.reqs.size(), ran ? MGCD_OK : (launch_rc != MGCD_OK ? launch_rc : MGCD_ERR_KERNEL));
  std::vector<uint64_t> offs;
  std::vector<uint32_t> ks;
  std::vector<std::vector<uint32_t>*> outs;
  for (size_t j = 0; j < sl.pd.size() && ran; ++j) {
    const PairOut& o = sl.h_out.as<PairOut>()[j];
    ServerReq* r = sl.reqs[sl.order[j]];
    sts[j] = o.status != KST_OK ? kst_to_status(o.status) : (o.ndig > sl.pd[j].nq ? MGCD_ERR_KERNEL : MGCD_OK);
    r->result.clear();
    if (sts[j] == MGCD_OK && o.ndig > 0) {
      offs.push_back(sl.pd[j].dig_off);
      ks.push_back(o.ndig);
      outs.push_back(&r->result);
    }
  }
  if (!ks.empty()) {
    const int crc = mr_convert(sl.stream, sl.digits.as<mgcd_pair_t>(), offs, ks, sl.mr, outs, &sl.retired_dev,
                                &sl.retired_host);
    if (crc != MGCD_OK)
      for (int& st : sts)
        if (st == MGCD_OK) st = crc;
  }
  for (size_t j = 0; j < sl.reqs.size(); ++j) {
    ServerReq* r = sl.reqs[ran ? sl.order[j] : j];
    int st = sts[j];
    // Short of moduli, a timed-out wait, or a launch whose parts do not fit
    // this slot's share of the device (requests of well over 1 Mbit, or many
    // slots): the request alone through the one-pair path, whose tiling may
    // use the whole device.
    if (st == MGCD_ERR_REDUX || st == MGCD_ERR_RECOVERY || st == MGCD_ERR_TIMEOUT || st == MGCD_ERR_TOO_MANY_THREADS) {
      std::vector<uint32_t> g(r->v.size() + 1);
      size_t ng = 0;
      st = mgcd_gcd_limbs(ctx, r->u.data(), r->u.size(), r->v.data(), r->v.size(), g.data(), g.size(), &ng);
      if (st == MGCD_OK) r->result.assign(g.begin(), g.begin() + ng);
    }
    r->status = st;
  }
}

// The worker thread lets no exception out either: a launch whose host
// bookkeeping throws answers its requests with MGCD_ERR_NOMEM.
int server_launch_safe(mgcd_server* s, ServerSlot& sl) {
  try {
    return server_launch(s, sl);
  } catch (...) {
    return MGCD_ERR_NOMEM;
  }
}
void server_finish_safe(mgcd_server* s, ServerSlot& sl, int launch_rc) {
  try {
    server_finish(s, sl, launch_rc);
  } catch (...) {
    for (ServerReq* r : sl.reqs) {
      r->result.clear();
      r->status = MGCD_ERR_NOMEM;
    }
  }
}

void server_loop(mgcd_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progress = false;
    // Completions.
    for (ServerSlot& sl : s->slots) {
      if (!sl.busy || hipEventQuery(sl.done_ev) != hipSuccess) continue;
      lk.unlock();
      server_finish_safe(s, sl, MGCD_OK);
      lk.lock();
      const auto now = std::chrono::steady_clock::now();
      if (s->trace)
        std::fprintf(stderr, "mgcd_server: slot %d done after %.3f ms\n", (int)(&sl - s->slots.data()),
                     std::chrono::duration<double>(now - sl.t_launch).count() * 1e3);
      for (ServerReq* r : sl.reqs) {
        r->latency = std::chrono::duration<double>(now - r->t_submit).count();
        r->done = true;
        s->latencies.push_back(r->latency);
        s->completed++;
      }
      s->t_last = now;
      s->used -= sl.bytes;
      sl.bytes = 0;
      sl.reqs.clear();
      sl.busy = false;
      progress = true;
      s->cv.notify_all();
    }
    // Admission: waiting requests of the slot's size class in arrival
    // order, while they fit the budget, into a free slot.
    for (ServerSlot& sl : s->slots) {
      std::deque<ServerReq*>& q = s->pending[sl.cls];
      if (sl.busy || q.empty()) continue;
      while (!q.empty() && s->used + sl.bytes + q.front()->bytes <= s->budget && sl.reqs.size() < 8192) {
        ServerReq* r = q.front();
        q.pop_front();
        sl.reqs.push_back(r);
        sl.bytes += r->bytes;
      }
      if (sl.reqs.empty()) break;  // the head waits for memory
      s->used += sl.bytes;
      s->peak = std::max(s->peak, s->used);
      sl.busy = true;
      sl.t_launch = std::chrono::steady_clock::now();
      lk.unlock();
      const int rc = server_launch_safe(s, sl);
      lk.lock();
      if (s->trace)
        std::fprintf(stderr, "mgcd_server: slot %d (class %d) launch of %zu requests, variant %d, rc %d\n",
                     (int)(&sl - s->slots.data()), sl.cls, sl.reqs.size(), sl.variant, rc);
      if (rc != MGCD_OK) {  // nothing is in flight on this slot: answer now
        (void)hipStreamSynchronize(sl.stream);
        lk.unlock();
        server_finish_safe(s, sl, rc);
        lk.lock();
        const auto now = std::chrono::steady_clock::now();
        for (ServerReq* r : sl.reqs) {
          r->latency = std::chrono::duration<double>(now - r->t_submit).count();
          r->done = true;
          s->completed++;
        }
        s->used -= sl.bytes;
        sl.bytes = 0;
        sl.reqs.clear();
        sl.busy = false;
        s->cv.notify_all();
      }
      progress = true;
    }
    bool busy = false;
    for (const ServerSlot& sl : s->slots) busy |= sl.busy;
    const bool waiting = !s->pending[0].empty() || !s->pending[1].empty();
    if (s->stop && !busy && !waiting) break;
    if (!progress) {
      if (busy)
        s->cv.wait_for(lk, std::chrono::microseconds(100));
      else
        s->cv.wait(lk, [&] { return s->stop || !s->pending[0].empty() || !s->pending[1].empty(); });
    }
  }
}

}  // namespace

extern "C" {

int mgcd_server_create(mgcd_ctx* ctx, uint64_t budget_bytes, int slots, mgcd_server** out) try {
  if (!ctx || !out || budget_bytes == 0 || slots < 1 || slots > 8) return MGCD_ERR_ARG;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  mgcd_server* s = new mgcd_server();
  s->ctx = ctx;
  s->budget = budget_bytes;
  s->variant = find_kernel_variant(256, 2, 16, ctx->wide);
  if (s->variant < 0 || ctx->max_resident[s->variant] < slots) {
    delete s;
    return MGCD_ERR_TOO_MANY_THREADS;
  }
  const char* tr = std::getenv("MGCD_SERVER_TRACE");
  s->trace = tr && *tr && *tr != '0';
  s->slots.resize(slots);
  // With two or more slots the last one serves the large requests.
  if (slots > 1) s->slots.back().cls = 1;
  // Slot 0 runs on the context's own stream (the server owns the context):
  // one hardware queue fewer, so the slots' streams do not end up sharing a
  // queue, where a long kernel would hold back the others.
  for (size_t i = 0; i < s->slots.size(); ++i) {
    ServerSlot& sl = s->slots[i];
    sl.own_stream = i > 0;
    if (i == 0)
      sl.stream = ctx->stream;
    else if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess)
      sl.stream = nullptr;
    if (!sl.stream || hipEventCreateWithFlags(&sl.done_ev, hipEventDisableTiming) != hipSuccess) {
      mgcd_server_destroy(s);
      return MGCD_ERR_HIP;
    }
  }
  s->worker = std::thread(server_loop, s);
  *out = s;
  return MGCD_OK;
} MGCD_CATCH

int mgcd_server_submit(mgcd_server* s, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, uint64_t* ticket) try {
  if (!s || !ticket || (nu && !u) || (nv && !v)) return MGCD_ERR_ARG;
  nu = trim(u, nu);
  nv = trim(v, nv);
  if (cmp_mag(u, nu, v, nv) < 0) {
    std::swap(u, v);
    std::swap(nu, nv);
  }
  ServerReq* r = new ServerReq();
  r->t_submit = std::chrono::steady_clock::now();
  r->u.assign(u, u + nu);
  r->v.assign(v, v + nv);
  bool immediate = false;
  if (nv == 0) {  // gcd(u, 0) = |u|, answered on the host
    r->result = r->u;
    immediate = true;
  } else {
    const int nqi = mgcd_num_moduli_needed(std::max<uint64_t>(bit_length(u, nu), 2), s->ctx->L);
    if (nqi <= 0 || (size_t)nqi > s->ctx->moduli.size()) {
      r->status = MGCD_ERR_NOT_ENOUGH_MODULI;
      immediate = true;
    } else {
      r->nq = (uint32_t)nqi;
      r->bytes = footprint(nu, nv, r->nq);
      if (r->bytes > s->budget) {
        delete r;
        return MGCD_ERR_BUSY;
      }
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->reqs[r->ticket] = r;
  if (!s->any) {
    s->any = true;
    s->t_first = r->t_submit;
  }
  if (immediate) {
    r->done = true;
    s->completed++;
    s->latencies.push_back(0.0);
  } else {
    s->pending[s->slots.size() > 1 && r->nq > kServeLargeNq ? 1 : 0].push_back(r);
  }
  s->cv.notify_all();
  return MGCD_OK;
} MGCD_CATCH

int mgcd_server_wait(mgcd_server* s, uint64_t ticket, double timeout_s, uint32_t* g, size_t cap, size_t* ng,
                     double* latency_s) {
  if (!s || !ng) return MGCD_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MGCD_ERR_ARG;
  ServerReq* r = it->second;
  const auto pred = [&] { return r->done; };
  if (timeout_s < 0)
    s->cv.wait(lk, pred);
  else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
    return MGCD_ERR_PENDING;  // ticket kept: ask again
  if (latency_s) *latency_s = r->latency;
  *ng = r->result.size();
  const int st = r->status;
  if (st == MGCD_OK && r->result.size() > cap) return MGCD_ERR_CAPACITY;  // ticket kept: ask again
  if (st == MGCD_OK && !r->result.empty()) std::memcpy(g, r->result.data(), r->result.size() * 4);
  s->reqs.erase(it);
  delete r;
  return st;
}

int mgcd_server_stats(mgcd_server* s, uint64_t* completed, double* p50_s, double* p99_s, double* gcds_per_s,
                      uint64_t* peak_bytes) try {
  if (!s) return MGCD_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto q = [&](double f) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t)std::ceil(f * l.size()) - 1)]; };
  if (completed) *completed = s->completed;
  if (p50_s) *p50_s = q(0.50);
  if (p99_s) *p99_s = q(0.99);
  if (gcds_per_s) {
    const double span = s->any ? std::chrono::duration<double>(s->t_last - s->t_first).count() : 0.0;
    *gcds_per_s = span > 0 ? s->completed / span : 0.0;
  }
  if (peak_bytes) *peak_bytes = s->peak;
  return MGCD_OK;
} MGCD_CATCH

int mgcd_server_destroy(mgcd_server* s) {
  if (!s) return MGCD_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->ctx->device);
  for (ServerSlot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (DevBuf* d : {&sl.limbs, &sl.pairs, &sl.groups, &sl.items, &sl.queue, &sl.bar, &sl.digits, &sl.trace, &sl.res,
                      &sl.out})
      d->release();
    for (PinnedBuf* h : {&sl.h_in, &sl.h_out}) h->release();
    sl.mr.release();
    for (void* p : sl.retired_dev) (void)hipFree(p);
    for (void* p : sl.retired_host) (void)hipHostFree(p);
    if (sl.done_ev) (void)hipEventDestroy(sl.done_ev);
    if (sl.stream && sl.own_stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto& kv : s->reqs) delete kv.second;
  delete s;
  return MGCD_OK;
}

}  // extern "C"
