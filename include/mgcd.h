/* mgcd.h -- C ABI of the MI355X modular integer GCD library (libmgcd.so).
 *
 * This is the thin, GMP-free boundary that the drop-in C++ class
 * GmpCuda::GmpCudaDevice (GmpCuda.h) calls.  Integers cross it as arrays of
 * 32-bit limbs, least significant limb first (the layout mpz_export(order=-1,
 * size=4, endian=0) produces, GmpCudaDevice-gcd.cu:71-73).  No C++ exception
 * and no HIP type crosses it; every entry point returns an mgcd_status.
 *
 * Threading: a context may be used by one host thread at a time.  Batches
 * and servers belong to the context that created them.
 */
#ifndef MGCD_H
#define MGCD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The first six map one-to-one to the std::runtime_error
 * messages of the reference (GmpCudaDevice.cu:101, GmpCudaDevice-gcd.cu:78,
 * :82, :107, :108, :109); mgcd_strerror returns those exact strings. */
typedef enum {
  MGCD_OK = 0,
  MGCD_ERR_INVALID_DEVICE = 1,     /* "Invalid device number." */
  MGCD_ERR_NOT_ENOUGH_MODULI = 2,  /* "Not enough moduli available for given input." */
  MGCD_ERR_TOO_MANY_THREADS = 3,   /* "Cannot allocate enough threads to support computation." */
  MGCD_ERR_REDUX = 4,              /* "Ran out of moduli in the reduction loop." */
  MGCD_ERR_RECOVERY = 5,           /* "Ran out of moduli in the recovery loop." */
  MGCD_ERR_KERNEL = 6,             /* "Unknown error in the gcd kernel." */
  MGCD_ERR_HIP = 7,                /* a HIP runtime call failed */
  MGCD_ERR_ARG = 8,                /* invalid argument */
  MGCD_ERR_TABLE = 9,              /* moduli table file missing or invalid */
  MGCD_ERR_CAPACITY = 10,          /* caller's output buffer too small */
  MGCD_ERR_NOMEM = 11,             /* host or device allocation failed */
  MGCD_ERR_TIMEOUT = 12,           /* a bounded in-kernel wait expired */
  MGCD_ERR_BUSY = 13,              /* server: admission budget exhausted */
  MGCD_ERR_SCRATCH = 14,           /* input beyond the limit of mgcd_gcd_limbs_noalloc */
  MGCD_ERR_PENDING = 15            /* server: the wait expired before the request finished */
} mgcd_status;

/* One signed mixed-radix digit: the pair_t the reference kernel hands back
 * to the host (GmpCuda.h:112).  Digit i of a result G satisfies
 * G = g_1 + m_1 (g_2 + m_2 (g_3 + ...)), g_i in [-(m_i+1)/2, (m_i-3)/2]. */
typedef struct {
  uint32_t modulus;
  int32_t value;
} mgcd_pair_t;

typedef struct mgcd_ctx mgcd_ctx;
typedef struct mgcd_batch mgcd_batch;
typedef struct mgcd_server mgcd_server;

/* Human-readable message for a status. */
const char* mgcd_strerror(int status);
/* Detail of the last failure on the calling thread (HIP call, table path). */
const char* mgcd_last_error_detail(void);

/* Device selection (replaces GmpCudaDevice::setDevice, GmpCudaDevice.cu:96-103):
 * MGCD_ERR_INVALID_DEVICE if dev is out of range. */
int mgcd_device_count(int* count);
int mgcd_set_device(int dev);

/* Number of moduli the algorithm is started with for an input whose larger
 * operand has `bits` bits, at modulus width L (numModuliNeededFor,
 * GmpCudaDevice-gcd.cu:49-55): roundUp(ceil(C_L*bits/ln(bits)), 256), C_L =
 * 1.6 - 0.015 L, evaluated in float as the reference does.  bits < 2 is
 * taken as 2 (the formula divides by ln(bits)); results saturate at 2^30. */
int mgcd_num_moduli_needed(uint64_t bits, int L);

/* Context: a device, a moduli table of width L in {22, 27, 32} uploaded to
 * it, streams and pooled device buffers (replaces the GmpCudaDevice
 * constructor, GmpCudaDevice.cu:107-138).  table_path may be NULL: the
 * library then loads moduli_L<L>.bin from the directory named by
 * $MGCD_TABLE_DIR, else from ../data next to libmgcd.so. */
int mgcd_ctx_create(mgcd_ctx** out, int device, int L, const char* table_path);
int mgcd_ctx_destroy(mgcd_ctx* ctx);
int mgcd_ctx_L(const mgcd_ctx* ctx);
int mgcd_ctx_num_moduli(const mgcd_ctx* ctx);           /* table size */
int mgcd_ctx_max_grid(const mgcd_ctx* ctx);             /* getMaxGridSize (GmpCuda.h:118) */
int mgcd_ctx_moduli(const mgcd_ctx* ctx, uint32_t* out, size_t cap); /* copy of the table */
/* Context options (mgcd_ctx_set_option). */
enum {
  /* N_Q for every following pair, 0 (default) for the reference's sizing. */
  MGCD_OPT_FORCE_MODULI = 1,
  /* 1 (default): where the reference's run stops with "Ran out of moduli"
   * (reduction or recovery), mgcd_gcd_limbs and batches retry the pair with
   * a quarter more moduli, up to the whole table; 0: report the error.
   * mgcd_gcd_limbs_trace never retries. */
  MGCD_OPT_RETRY = 2,
  /* Tiling of the one-pair path: threads per workgroup and moduli slots per
   * thread (both 0, the default: chosen from N_Q). */
  MGCD_OPT_SINGLE_WORKGROUP = 3,
  MGCD_OPT_SINGLE_SLOTS = 4,
  /* 1: batches created with tiling (0, 0, 0) measure the candidate tilings
   * on each size class (up to 4096 of its pairs; mgcd_autotune) unless a
   * tuned tiling for that class and population is cached in the context or
   * in the file named by $MGCD_TUNE_CACHE, which receives the new results.
   * 0 (default): cached or built-in tilings.  Either way a class with under
   * a tenth of the largest class's work (sum of N_Q^2) runs with that
   * class's tiling, in its launch. */
  MGCD_OPT_AUTOTUNE = 5,
  /* 1: mgcd_batch_run captures the batch's launch sequence into a hipGraph
   * on its first run and replays the graph afterwards.  0 (default):
   * cooperative launches. */
  MGCD_OPT_BATCH_GRAPH = 6,
  /* Pairs split over at least this many workgroups (>= 2; 0, the default:
   * 48 for groups of pairs, 128 for one-pair launches) exchange their per-iteration
   * minima through part 0 (two hops, traffic linear in the parts) instead of
   * every part reading every part's slot (one hop, traffic quadratic in the
   * parts). */
  MGCD_OPT_RELAY_PARTS = 7,
  /* 1 (default): the signed mixed-radix digits become positional limbs on
   * the device, by a product tree (batches; one-pair calls above 640 digits),
   * and only |G| is copied back.  0: the digits are copied back and
   * converted by the host Horner loop (GmpCudaDevice-gcd.cu:113-129). */
  MGCD_OPT_DEVICE_CONVERT = 8,
  /* Polls before an in-kernel wait (cross-workgroup exchange, residue job)
   * gives up and the pair reports MGCD_ERR_TIMEOUT: default 2^24, at least
   * 1.  Every wait is bounded by it, so a small value makes the kernel end
   * early with timeouts instead of results (used to test that path). */
  MGCD_OPT_SPIN_LIMIT = 9,
  /* 1: the persistent kernel is launched cooperatively where the runtime
   * allows it (all of its workgroups resident); 0: a plain launch of the same
   * co-resident grid (on an otherwise idle device its workgroups are all
   * resident too; a straggler only delays the bounded waits).  Default: the
   * environment variable MGCD_COOPERATIVE if set, else 1.  Profilers that
   * fail on cooperative launches (rocprofv3's teardown) run with 0. */
  MGCD_OPT_COOPERATIVE = 10
};
int mgcd_ctx_set_option(mgcd_ctx* ctx, int option, int64_t value);
/* Signed mixed-radix digits (least significant first, the layout the kernel
 * writes: G = g_0 + m_0 (g_1 + m_1 (...))) -> |G| as limbs, converted on the
 * device by the product tree that MGCD_OPT_DEVICE_CONVERT enables (the role
 * of the host Horner loop, GmpCudaDevice-gcd.cu:113-129).  Each |g_i| must
 * be at most (m_i + 1) / 2, as the kernel's digits are. */
int mgcd_mixed_radix_to_limbs(mgcd_ctx* ctx, const mgcd_pair_t* digits, size_t k, uint32_t* g, size_t cap,
                              size_t* ng);
/* Blocks until all work queued on the context's device has finished. */
int mgcd_ctx_synchronize(mgcd_ctx* ctx);
/* Seconds the last gcd kernel launch of this context took (HIP events). */
double mgcd_ctx_last_kernel_seconds(const mgcd_ctx* ctx);
/* Work of the last one-pair launch of this context (mgcd_gcd_limbs,
 * mgcd_gcd_limbs_trace): N_Q, elimination iterations R, recovery digits k,
 * and its tiling (threads per workgroup, moduli slots per thread,
 * workgroups).  Any pointer may be NULL.  No counterpart in the reference,
 * whose harness times the call only (testmodgcd.cpp:265-275). */
int mgcd_ctx_last_pair_stats(const mgcd_ctx* ctx, uint32_t* nq, uint32_t* iters, uint32_t* digits, int* workgroup,
                             int* slots, int* parts);

/* Phase sums of the instrumented build of this library (lib/libmgcd_phases.so,
 * built with -DMGCD_PHASE_CLOCKS; the default build returns MGCD_ERR_ARG).
 * Over every work item the gcd kernel ran since the last reset, as wave 0 of
 * each workgroup sees them: out[0..8] shader-clock cycles spent in residues,
 * elimination compute, the barrier after it, the cross-workgroup exchange,
 * the per-pair bookkeeping, the barrier ending a round, the step from
 * elimination to replay, replay, recovery; out[9] elimination rounds;
 * out[10] work items; out[11] 100 MHz real-time ticks and out[12] shader
 * cycles over whole items (their ratio gives the clock); out[13..15] for the
 * one-pass collects (groups of up to 256 exchange slots): cycles from issuing
 * a round's first slot loads to their return, the collects whose first loads
 * found every slot posted, and the collects; out[16 + w] and out[24 + w]
 * wave w's own elimination compute and its wait at the barrier after it
 * (waves 0-7).  Copies min(n, 32) values; reset != 0 clears the sums
 * afterwards.  No counterpart in the reference. */
int mgcd_ctx_phase_clocks(mgcd_ctx* ctx, uint64_t* out, size_t n, int reset);

/* One pair, one launch sequence: the path of GmpCudaDevice::gcd
 * (GmpCudaDevice-gcd.cu:57-132) without GMP.  u, v may come in either order
 * and may be zero.  The gcd is written to g (cap limbs), its limb count to
 * *ng (0 for a zero gcd).  Blocking. */
int mgcd_gcd_limbs(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                   uint32_t* g, size_t cap, size_t* ng);

/* As mgcd_gcd_limbs, without allocating: no hipMalloc, hipFree or other
 * allocating call.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  Every device buffer of the call lives
 * in a fixed 2 MiB array of the library that each device holds until the
 * process ends; its address is taken when a context is first created on the
 * device.  Limits: N_Q <= 16384 moduli (operands up to about 176 kbit at
 * L = 32, 153 kbit at L = 22; retries stay within it) and nu + nv <= 12288
 * limbs; larger inputs return MGCD_ERR_SCRATCH.  Calls on one device share
 * that array: calls that overlap on one device (from several threads or
 * contexts) race, so they must be serialised by the caller. */
int mgcd_gcd_limbs_noalloc(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                           uint32_t* g, size_t cap, size_t* ng);

/* As mgcd_gcd_limbs, and also returns the kernel's signed mixed-radix
 * digits (least significant first) and the elimination pivots (p, b) in
 * the order the elimination loop used them.  Any of the extra buffers may
 * be NULL with capacity 0; *nd / *nt always receive the full counts.  The
 * operands are taken in the order given (u must have at least as many bits
 * as v) so the kernel sees exactly what the reference kernel would. */
int mgcd_gcd_limbs_trace(mgcd_ctx* ctx, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                         uint32_t* g, size_t cap, size_t* ng,
                         mgcd_pair_t* digits, size_t dcap, size_t* nd,
                         mgcd_pair_t* pivots, size_t tcap, size_t* nt);

/* Batch of pairs for the persistent kernel.  Limbs of all operands are
 * packed in `limbs`; pair i is (limbs + u_off[i], nu[i]) and
 * (limbs + v_off[i], nv[i]).  The inputs are copied to device memory once,
 * at creation; mgcd_batch_run computes every gcd from the resident copy.
 * Tiling: threads per workgroup, moduli slots per thread, pairs per group
 * (pairs with equal N_Q that share one modular inverse per modulus); all
 * three 0 lets the library choose. */
int mgcd_batch_create(mgcd_ctx* ctx, const uint32_t* limbs, const uint64_t* u_off, const uint32_t* nu,
                      const uint64_t* v_off, const uint32_t* nv, size_t npairs,
                      int opts_workgroup, int opts_slots, int opts_group, mgcd_batch** out);
int mgcd_batch_run(mgcd_batch* b);
int mgcd_batch_status(const mgcd_batch* b, size_t i);
int mgcd_batch_result(const mgcd_batch* b, size_t i, uint32_t* g, size_t cap, size_t* ng);
/* Seconds the last run spent in its device launch sequence (HIP events on
 * the context's stream), and in host-side digit conversion. */
int mgcd_batch_timing(const mgcd_batch* b, double* device_s, double* host_s);
/* Work done for pair i in the last run: N_Q, elimination iterations R and
 * recovery digits k (0 for pairs answered on the host). */
int mgcd_batch_pair_stats(const mgcd_batch* b, size_t i, uint32_t* nq, uint32_t* iters, uint32_t* digits);
/* The elimination pivots (p, b) pair i used in the last run, in order, as
 * the kernel recorded them (the reference's collect points,
 * GmpCudaDevice-getGcdKernel.cu:371, :400): copies min(cap, count) of them,
 * *n receives the count.  MGCD_ERR_ARG for a pair answered on the host
 * (v = 0) or before the first run. */
int mgcd_batch_pivots(const mgcd_batch* b, size_t i, mgcd_pair_t* pivots, size_t cap, size_t* n);
/* The tiling of the launch with the most pairs: threads per workgroup,
 * moduli slots per thread, pairs per group. */
int mgcd_batch_tiling(const mgcd_batch* b, int* workgroup, int* slots, int* group);
/* A batch created with tiling (0, 0, 0) is split into size classes of N_Q
 * (up to 1024, 4096, 16384, more), each with its own tiling; adjacent
 * classes with the same tiling share a kernel launch.  An explicit tiling
 * gives one launch.  Launch i: tiling, pairs,
 * N_Q range and the seconds its kernel took in the last run (0 when the run
 * replayed a graph). */
int mgcd_batch_launch_count(const mgcd_batch* b, size_t* count);
int mgcd_batch_launch_info(const mgcd_batch* b, size_t i, int* workgroup, int* slots, int* group, size_t* npairs,
                           uint32_t* nq_min, uint32_t* nq_max, double* device_s);
int mgcd_batch_destroy(mgcd_batch* b);

/* Autotuner: runs the batch kernel with each candidate tiling on these
 * pairs (one warm-up and one timed run each) and returns the fastest.
 * No counterpart in the reference, which has one fixed launch shape
 * (GmpCudaDevice-gcd.cu:80-84). */
int mgcd_autotune(mgcd_ctx* ctx, const uint32_t* limbs, const uint64_t* u_off, const uint32_t* nu,
                  const uint64_t* v_off, const uint32_t* nv, size_t npairs, int* workgroup, int* slots, int* group);

/* Serving loop.  Requests of any size are submitted one by one; a worker
 * thread admits waiting requests in arrival order while their device
 * footprint (limbs, residues, pivots, digits: about 24 N_Q + 4 (nu + nv)
 * bytes each) fits budget_bytes, packs them into the next launch of the
 * persistent kernel, and keeps up to `slots` launches in flight on their
 * own streams (each with a 1/slots share of the device's workgroups), so
 * requests that arrive while others run are started as soon as a slot and
 * memory free up and overlap with them on the device.  With two or more
 * slots the last one takes only requests of more than 16,384 moduli (above
 * ~200 kbit at L = 32), so small requests never wait behind a large one in
 * the same launch.  The context belongs to the server's worker until
 * mgcd_server_destroy.  No counterpart in the
 * reference (one blocking call per pair, GmpCudaDevice-gcd.cu:57-132).
 *   submit: copies the operands; MGCD_ERR_BUSY if the request alone
 *           exceeds the budget.
 *   wait:   blocks until ticket's gcd is known (timeout_s < 0: no limit;
 *           after timeout_s seconds otherwise MGCD_ERR_PENDING, and the
 *           ticket stays valid), then returns its status (MGCD_ERR_TIMEOUT
 *           only when an in-kernel wait of the request's launch expired),
 *           its limbs in g and the latency from submit to completion, and
 *           releases the ticket (kept on MGCD_ERR_CAPACITY).
 *   stats:  completed requests, p50 and p99 latency, completions per
 *           second since the first submit, peak resident bytes. */
int mgcd_server_create(mgcd_ctx* ctx, uint64_t budget_bytes, int slots, mgcd_server** out);
int mgcd_server_submit(mgcd_server* s, const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                       uint64_t* ticket);
int mgcd_server_wait(mgcd_server* s, uint64_t ticket, double timeout_s, uint32_t* g, size_t cap, size_t* ng,
                     double* latency_s);
int mgcd_server_stats(mgcd_server* s, uint64_t* completed, double* p50_s, double* p99_s, double* gcds_per_s,
                      uint64_t* peak_bytes);
int mgcd_server_destroy(mgcd_server* s);

/* Exhaustive certification of the FP32-reciprocal quotient estimate that the
 * modular inverse uses, on this device (the role of certifyQuoRemQuasi.cu).
 * Postcondition, q = rint(fl32(x) * rcp(fl32(y))), r = x - q*y: |r| < y when
 * x < 2^20 y, |r| < x otherwise (the inverse's Euclid then terminates).
 *   mode 0: every x in [1, 2^23), y in [1, 2^22)  (the float domain);
 *   mode 1: for width L, every y in [1, 2^L) and every quotient class k, the
 *           two endpoints x = k*y and x = (k+1)*y - 1 (the estimate is
 *           monotone in x, so the endpoints decide every x in between);
 *   mode 2: every float f in [1, 2^34): |rcp(f)*f - 1| <= 2^-23 and
 *           rcp(-f) == -rcp(f) (the error bound and sign symmetry that the
 *           inverse's termination and exactness arguments use).
 * *failures receives the number of violations, *checked the number of cases. */
int mgcd_certify_quasi_quotient(mgcd_ctx* ctx, int mode, int L, uint64_t* failures, uint64_t* checked);

#ifdef __cplusplus
}
#endif

#endif /* MGCD_H */
