/* mgcd_oracle.c -- CPU restatement of the reference's modular GCD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and then only to compute expected answers or to time GMP.  The product
 * (the modgcd-onegpu_amd package) never links, loads or calls it.
 *
 * Each function restates, in plain sequential C, one piece of the reference
 * (mountunion/ModGCD-OneGPU); the file:line it follows is named in its
 * comment.  The kernel restatement runs the N_Q "threads" of the reference
 * kernel one after another and reproduces its pivot sequence, mixed-radix
 * digits and error codes exactly.  Integer arithmetic uses uint64_t and
 * unsigned __int128; GMP is used only where the reference uses it on the
 * host (mpz Horner, mpz_gcd).
 *
 * How each part is pinned (tests/test_oracle.py):
 *  - oracle_create_moduli: equal to the reference's own createModuli built
 *    from its sources (oracle/_ref) for L = 32; primality and the CW
 *    condition re-checked with Python integers for all widths.
 *  - oracle_cw_inverse / oracle_mod / oracle_mod_mul: exhaustive and random
 *    comparison with x % m.
 *  - oracle_mod_inv: v * inv == 1 (mod m), exhaustively for small moduli and
 *    against Python pow(v, -1, m) on random 22/27/32-bit inputs.
 *  - oracle_mod_mp: equal to Python int % m.
 *  - oracle_gcd_kernel: pivots checked against a brute-force big-integer
 *    Weber elimination on small inputs (exact U, V as Python ints); digits
 *    reconstruct the gcd; final result equals mpz_gcd and math.gcd.
 *  - oracle_gcd_kernel_exact (the same steps with x % m, for "all primes"
 *    tables): equal to oracle_gcd_kernel on CW tables; over all-primes
 *    moduli (some outside the CW condition), pivots and digits equal the
 *    brute-force run.
 *  - oracle_num_moduli_needed: the reference's float formula; closed-form
 *    values at the configs' sizes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>
#include <gmp.h>

#define ORACLE_W 64
#define MOD_INFINITY INT32_MIN /* getGcdKernel.cu:46 */

typedef struct { uint32_t modulus; int32_t value; } oracle_pair_t; /* GmpCuda.h:112 */

enum {
  ORACLE_OK = 0,
  ORACLE_REDUX_ERROR = 1,    /* getGcdKernel.cu:393-399 */
  ORACLE_RECOVERY_ERROR = 2, /* getGcdKernel.cu:440-441 */
  ORACLE_CAPACITY = 3,       /* output buffer too small (not in the reference) */
  ORACLE_ARG = 4
};

/* ---------------------------------------------------------------------------
 * Sizing: numModuliNeededFor and roundUp (GmpCudaDevice-gcd.cu:38-55).  The
 * reference computes in float: C_L is a constexpr float, numBits is converted
 * to float, logf is the natural log.  Result rounded up to GCD_BLOCK_SZ=256.
 */
int oracle_num_moduli_needed(uint64_t num_bits, int L) {
  const float C_L = (float)(1.6 - 0.015 * L);
  float x = C_L * (float)num_bits;
  x = x / logf((float)num_bits);
  int n = (int)ceilf(x);
  return ((n - 1) / 256 + 1) * 256;
}

/* ---------------------------------------------------------------------------
 * Moduli (createModuli.cpp:104-179): odd integers walked downward from
 * 2^L - 1; keep primes D with the Cavagnino-Werbrouck condition
 * Ncr = ceil(J / (D*J - FC)) * D - 1 >= 2^W, FC = 2^(W+L-1), J = FC/D + 1.
 * Primality here is trial division (a different method from the product's
 * segmented sieve).
 */
static int is_prime_u32(uint32_t n) {
  if (n < 2) return 0;
  if ((n & 1) == 0) return n == 2;
  for (uint32_t d = 3; (uint64_t)d * d <= n; d += 2)
    if (n % d == 0) return 0;
  return 1;
}

int oracle_cw_usable(uint32_t D, int L) {
  typedef unsigned __int128 u128;
  u128 FC = (u128)1 << (ORACLE_W + L - 1);
  u128 J = FC / D + 1;
  u128 DJ_FC = (u128)D * J - FC;
  u128 Qcr = J / DJ_FC + (J % DJ_FC != 0);
  u128 Ncr = Qcr * D - 1;
  return (Ncr >> ORACLE_W) != 0;
}

size_t oracle_create_moduli(int L, uint32_t* out, size_t count) {
  size_t n = 0;
  uint64_t lo = (uint64_t)1 << (L - 1);
  for (uint64_t x = ((uint64_t)1 << L) - 1; x > lo && n < count; x -= 2)
    if (is_prime_u32((uint32_t)x) && oracle_cw_usable((uint32_t)x, L)) out[n++] = (uint32_t)x;
  return n;
}

/* ---------------------------------------------------------------------------
 * Modular arithmetic (getGcdKernel.cu:222-292, 332-340).
 */
typedef struct { uint32_t modulus; uint64_t inverse; int L; } oracle_modulus_t;

/* getModulus (getGcdKernel.cu:332-340): inverse = floor(2^(W+L-1)/m) + 1. */
uint64_t oracle_cw_inverse(uint32_t m, int L) {
  uint64_t D = m;
  const uint64_t FC_hi = (uint64_t)1 << (ORACLE_W - 1);
  uint64_t q = FC_hi / D, r = FC_hi % D;
  return 1 + (q << L) + (r << L) / D;
}

/* mod (getGcdKernel.cu:248-251): x - m * (umulhi64(inv, x) >> (L-1)).
 * inverse == 0 selects exact reduction (x % m): the reference's reduction
 * is exact only for moduli that satisfy the Cavagnino-Werbrouck condition,
 * and "all primes" tables hold others. */
static inline uint32_t mod64(uint64_t x, const oracle_modulus_t* m) {
  if (m->inverse == 0) return (uint32_t)(x % m->modulus);
  uint64_t hi = (uint64_t)(((unsigned __int128)m->inverse * x) >> 64);
  return (uint32_t)(x - (uint64_t)m->modulus * (hi >> (m->L - 1)));
}

uint32_t oracle_mod(uint64_t x, uint32_t m, int L) {
  oracle_modulus_t q = {m, oracle_cw_inverse(m, L), L};
  return mod64(x, &q);
}

static inline uint32_t modSub(uint32_t a, uint32_t b, const oracle_modulus_t* m) {
  return a - b + ((uint32_t)(-(int32_t)(a < b)) & m->modulus); /* getGcdKernel.cu:238-241 */
}
static inline uint32_t modMul(uint32_t a, uint32_t b, const oracle_modulus_t* m) {
  return mod64((uint64_t)a * b, m); /* getGcdKernel.cu:259-262 */
}
uint32_t oracle_mod_mul(uint32_t a, uint32_t b, uint32_t m, int L) {
  oracle_modulus_t q = {m, oracle_cw_inverse(m, L), L};
  return modMul(a, b, &q);
}
static inline uint32_t fromSigned(int32_t x, const oracle_modulus_t* m) {
  return (x < 0) ? (uint32_t)x + m->modulus : (uint32_t)x; /* getGcdKernel.cu:268-271 */
}
static inline int32_t toSigned(uint32_t x, const oracle_modulus_t* m) {
  return (x >= m->modulus / 2) ? (int32_t)(x - m->modulus) : (int32_t)x; /* getGcdKernel.cu:278-281 */
}

/* modInv (modInv.h:141-188): Knuth's Algorithm X with non-negative cofactors
 * that alternate roles.  The reference divides with FP32 quasi-quotients;
 * every variant it certifies returns the same (unique) inverse, so this
 * restatement divides exactly.  Precondition gcd(m, v) == 1, m > v > 0. */
uint32_t oracle_mod_inv(uint32_t v, uint32_t m) {
  uint32_t u2 = 0, u3 = m, v2 = 1, v3 = v;
  for (;;) {
    if (v3 == 0) return m - u2;          /* u3 == 1: result -u2 */
    uint32_t q = u3 / v3;
    u3 -= q * v3; u2 += q * v2;
    if (u3 == 0) return v2;              /* v3 == 1: result +v2 */
    q = v3 / u3;
    v3 -= q * u3; v2 += q * u2;
  }
}

static inline uint32_t modDiv(uint32_t u, uint32_t v, const oracle_modulus_t* m) {
  return modMul(u, oracle_mod_inv(v, m->modulus), m); /* getGcdKernel.cu:289-292 */
}

/* modMP (getGcdKernel.cu:297-324): Horner over 32-bit limbs, most significant
 * first (the reference stages 32 limbs at a time in shared memory; the order
 * of the mod steps is the same). */
static uint32_t mod_mp(const uint32_t* x, size_t n, const oracle_modulus_t* q) {
  uint64_t r = 0;
  for (size_t i = n; i-- != 0;) r = mod64(r << 32 | x[i], q);
  return (uint32_t)r;
}
uint32_t oracle_mod_mp(const uint32_t* x, size_t n, uint32_t m, int L) {
  oracle_modulus_t q = {m, oracle_cw_inverse(m, L), L};
  return mod_mp(x, n, &q);
}

/* ---------------------------------------------------------------------------
 * The kernel (getGcdKernel.cu:343-444), all N_Q threads run in turn.
 *
 * Inputs: u (nu limbs), v (nv limbs), moduli[0..nq).  Outputs:
 *  - digits[0..*ndigits): the pairs the reference writes to buf[1..] (:427),
 *    least significant first;
 *  - trace[0..*ntrace): the pivot pair collected before each elimination
 *    iteration (:371, :400), i.e. the (p, b) used by iteration i;
 *  - return value: ORACLE_OK / ORACLE_REDUX_ERROR / ORACLE_RECOVERY_ERROR.
 */

/* Pivot key of postMinPair (getGcdKernel.cu:154) and its decoding
 * (:214-219).  modAbs(INT32_MIN) converts the int -2^31 to uint64, whose
 * high bits are shifted out by the << 32, leaving 0x80000000 in the high
 * word (:130-137). */
static inline uint64_t min_key(oracle_pair_t p) {
  uint64_t a = (p.value < 0) ? (uint64_t)(uint32_t)(-(int64_t)p.value) : (uint64_t)p.value;
  if (p.value == MOD_INFINITY) a = 0x80000000u;
  return (a << 32) | (uint64_t)(p.modulus - (p.value >= 0));
}
static inline oracle_pair_t key_pair(uint64_t x) {
  oracle_pair_t p;
  p.modulus = (uint32_t)x;
  p.value = (int32_t)(uint32_t)(x >> 32);
  if (p.modulus & 1) p.value = (int32_t)(~(uint32_t)p.value + 1u);
  p.modulus |= 1;
  return p;
}

/* __ffs(abs(x)) of getGcdKernel.cu:391, with abs(INT32_MIN) == INT32_MIN. */
static inline int ffs_abs(int32_t x) {
  uint32_t a = (x < 0) ? (uint32_t)0 - (uint32_t)x : (uint32_t)x;
  return a ? __builtin_ctz(a) + 1 : 0;
}

static int gcd_kernel(const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                      const uint32_t* moduli, int nq, int L, int exact,
                      oracle_pair_t* digits, size_t dcap, size_t* ndigits,
                      oracle_pair_t* trace, size_t tcap, size_t* ntrace) {
  if (nq <= 0 || L < 3 || L > 32) return ORACLE_ARG;
  oracle_modulus_t* q = (oracle_modulus_t*)malloc(sizeof(oracle_modulus_t) * nq);
  uint32_t* uq = (uint32_t*)malloc(sizeof(uint32_t) * nq);
  uint32_t* vq = (uint32_t*)malloc(sizeof(uint32_t) * nq);
  uint32_t* tq = (uint32_t*)malloc(sizeof(uint32_t) * nq);
  char* active = (char*)malloc(nq);
  oracle_pair_t* my = (oracle_pair_t*)malloc(sizeof(oracle_pair_t) * nq);
  size_t nt = 0, nd = 0;
  int status = ORACLE_OK;

  int totalModuliRemaining = nq;                 /* :350 */
  int ubits = (int)((nu + 1) * 32);              /* :351 */
  int vbits = (int)((nv + 1) * 32);              /* :352 */

  for (int i = 0; i < nq; ++i) {                 /* MGCD1, MGCD2 (:355-361) */
    q[i].modulus = moduli[i];
    q[i].inverse = exact ? 0 : oracle_cw_inverse(moduli[i], L);
    q[i].L = L;
    uq[i] = mod_mp(u, nu, &q[i]);
    vq[i] = mod_mp(v, nv, &q[i]);
    active[i] = 1;
    my[i].modulus = moduli[i];                   /* :368-369 */
    my[i].value = (vq[i] == 0) ? MOD_INFINITY : toSigned(modDiv(uq[i], vq[i], &q[i]), &q[i]);
  }
  uint64_t best = UINT64_MAX;                    /* postMinPair/collectMinPair (:370-371) */
  for (int i = 0; i < nq; ++i) { uint64_t k = min_key(my[i]); if (k < best) best = k; }
  oracle_pair_t pair = key_pair(best);

  do {                                           /* MGCD3 (:373-402) */
    if (nt < tcap) trace[nt] = pair;
    nt++;
    uint64_t next = UINT64_MAX;
    for (int i = 0; i < nq; ++i) {
      if (pair.modulus == q[i].modulus) active[i] = 0, my[i].value = MOD_INFINITY;
      if (active[i]) {
        uint32_t p = pair.modulus;
        if (p > q[i].modulus) p -= q[i].modulus;
        tq[i] = modDiv(modSub(uq[i], modMul(fromSigned(pair.value, &q[i]), vq[i], &q[i]), &q[i]), p, &q[i]);
        my[i].value = (tq[i] == 0) ? MOD_INFINITY : toSigned(modDiv(vq[i], tq[i], &q[i]), &q[i]);
      }
      uint64_t k = min_key(my[i]);
      if (k < next) next = k;
    }
    for (int i = 0; i < nq; ++i)
      if (active[i]) uq[i] = vq[i], vq[i] = tq[i];
    totalModuliRemaining -= 1;
    int tbits = ubits - (L - 1) + ffs_abs(pair.value);
    ubits = vbits, vbits = tbits;
    if (totalModuliRemaining * (L - 2) <= ubits) { status = ORACLE_REDUX_ERROR; goto done; }
    pair = key_pair(next);
  } while (pair.value != MOD_INFINITY);

  /* MGCD4 (:406-433): signed mixed-radix digits; the "any" pair is the
   * nonzero value of the lowest thread index, else thread 0's pair
   * (:57-64, :69-116). */
  {
    int first = -1;
    for (int i = 0; i < nq; ++i) {
      my[i].value = active[i] ? toSigned(uq[i], &q[i]) : 0;
      if (first < 0 && my[i].value != 0) first = i;
    }
    pair = my[first < 0 ? 0 : first];
    do {
      first = -1;
      for (int i = 0; i < nq; ++i) {
        if (pair.modulus == q[i].modulus) active[i] = 0, my[i].value = 0;
        if (active[i]) {
          uint32_t p = pair.modulus;
          if (pair.modulus > q[i].modulus) p -= q[i].modulus;
          uq[i] = modDiv(modSub(uq[i], fromSigned(pair.value, &q[i]), &q[i]), p, &q[i]);
          my[i].value = toSigned(uq[i], &q[i]);
        }
        if (first < 0 && my[i].value != 0) first = i;
      }
      if (nd < dcap) digits[nd] = pair;
      nd++;
      totalModuliRemaining -= 1;
      if (totalModuliRemaining <= 0) break;
      pair = my[first < 0 ? 0 : first];
    } while (pair.value != 0);
    if (pair.value != 0) status = ORACLE_RECOVERY_ERROR; /* :440-441 */
  }
done:
  if (ndigits) *ndigits = nd;
  if (ntrace) *ntrace = nt;
  if (status == ORACLE_OK && nd > dcap) status = ORACLE_CAPACITY;
  free(q); free(uq); free(vq); free(tq); free(active); free(my);
  return status;
}

/* The reference kernel's steps with its reduction (CW moduli). */
int oracle_gcd_kernel(const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                      const uint32_t* moduli, int nq, int L,
                      oracle_pair_t* digits, size_t dcap, size_t* ndigits,
                      oracle_pair_t* trace, size_t tcap, size_t* ntrace) {
  return gcd_kernel(u, nu, v, nv, moduli, nq, L, 0, digits, dcap, ndigits, trace, tcap, ntrace);
}
/* The same steps with exact reduction: any odd moduli ("all primes" tables).
 * Equal to oracle_gcd_kernel wherever every modulus satisfies the CW
 * condition. */
int oracle_gcd_kernel_exact(const uint32_t* u, size_t nu, const uint32_t* v, size_t nv,
                            const uint32_t* moduli, int nq, int L,
                            oracle_pair_t* digits, size_t dcap, size_t* ndigits,
                            oracle_pair_t* trace, size_t tcap, size_t* ntrace) {
  return gcd_kernel(u, nu, v, nv, moduli, nq, L, 1, digits, dcap, ndigits, trace, tcap, ntrace);
}

/* ---------------------------------------------------------------------------
 * Host side (GmpCudaDevice-gcd.cu:116-129): Horner over the signed digits,
 * most significant first, then |G|.  Output as 32-bit limbs, least
 * significant first.
 */
static void limbs_to_mpz(mpz_t z, const uint32_t* x, size_t n) {
  mpz_import(z, n, -1, sizeof(uint32_t), 0, 0, x);
}
static size_t mpz_to_limbs(uint32_t* out, size_t cap, const mpz_t z) {
  size_t need = (mpz_sizeinbase(z, 2) + 31) / 32;
  if (mpz_sgn(z) == 0) return 0;
  if (need > cap) return need;
  size_t cnt = 0;
  mpz_export(out, &cnt, -1, sizeof(uint32_t), 0, 0, z);
  return cnt;
}

size_t oracle_horner(const oracle_pair_t* digits, size_t k, uint32_t* out, size_t cap) {
  mpz_t g;
  mpz_init(g);
  if (k > 0) {
    const oracle_pair_t* p = digits + k - 1;
    mpz_set_si(g, p->value);
    while (p-- != digits) {
      mpz_mul_ui(g, g, p->modulus);
      if (p->value < 0) mpz_sub_ui(g, g, (unsigned long)(-(int64_t)p->value));
      else mpz_add_ui(g, g, (unsigned long)p->value);
    }
  }
  mpz_abs(g, g);
  size_t n = mpz_to_limbs(out, cap, g);
  mpz_clear(g);
  return n;
}

/* mpz_gcd on limb arrays: the reference's CPU path (testmodgcd.cpp:261-263). */
size_t oracle_mpz_gcd(const uint32_t* u, size_t nu, const uint32_t* v, size_t nv, uint32_t* out, size_t cap) {
  mpz_t a, b, g;
  mpz_inits(a, b, g, NULL);
  limbs_to_mpz(a, u, nu);
  limbs_to_mpz(b, v, nv);
  mpz_gcd(g, a, b);
  size_t n = mpz_to_limbs(out, cap, g);
  mpz_clears(a, b, g, NULL);
  return n;
}

const char* oracle_gmp_version(void) { return gmp_version; }

/* ---------------------------------------------------------------------------
 * CPU baseline timer: mpz_gcd over a set of pairs, `threads` workers pulling
 * pairs from a shared counter, timed with CLOCK_MONOTONIC_RAW as at
 * testmodgcd.cpp:70-75.  Input conversion happens before the clock starts.
 * Returns elapsed seconds; writes each gcd's bit length to gbits[i].
 */
typedef struct {
  mpz_t* u; mpz_t* v; mpz_t* g; size_t n; volatile size_t next; pthread_mutex_t lock;
} baseline_job_t;

static void* baseline_worker(void* arg) {
  baseline_job_t* job = (baseline_job_t*)arg;
  for (;;) {
    pthread_mutex_lock(&job->lock);
    size_t i = job->next++;
    pthread_mutex_unlock(&job->lock);
    if (i >= job->n) break;
    mpz_gcd(job->g[i], job->u[i], job->v[i]);
  }
  return NULL;
}

static double now_raw(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC_RAW, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* limbs: all pairs' limbs packed; offsets: u_off[i], nu[i], v_off[i], nv[i]. */
double oracle_time_mpz_gcd(const uint32_t* limbs, const uint64_t* u_off, const uint32_t* nu,
                           const uint64_t* v_off, const uint32_t* nv, size_t npairs, int threads,
                           uint64_t* gbits) {
  baseline_job_t job;
  job.n = npairs;
  job.next = 0;
  pthread_mutex_init(&job.lock, NULL);
  job.u = (mpz_t*)malloc(sizeof(mpz_t) * npairs);
  job.v = (mpz_t*)malloc(sizeof(mpz_t) * npairs);
  job.g = (mpz_t*)malloc(sizeof(mpz_t) * npairs);
  for (size_t i = 0; i < npairs; ++i) {
    mpz_inits(job.u[i], job.v[i], job.g[i], NULL);
    limbs_to_mpz(job.u[i], limbs + u_off[i], nu[i]);
    limbs_to_mpz(job.v[i], limbs + v_off[i], nv[i]);
  }
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  double t0 = now_raw();
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, baseline_worker, &job);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  double t1 = now_raw();
  for (size_t i = 0; i < npairs; ++i) {
    if (gbits) gbits[i] = mpz_sgn(job.g[i]) ? mpz_sizeinbase(job.g[i], 2) : 0;
    mpz_clears(job.u[i], job.v[i], job.g[i], NULL);
  }
  free(th); free(job.u); free(job.v); free(job.g);
  pthread_mutex_destroy(&job.lock);
  return t1 - t0;
}
