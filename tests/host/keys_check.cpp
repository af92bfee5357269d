// keys_check.cpp -- CPU check of the exchange-key decoders (csrc/mgcd_keys.h).
// Prints one line "name violations cases" per check; every count of
// violations must be 0.
#include <cstdio>
#include <cstdint>

#include "mgcd_keys.h"

using namespace mgcd::keys;

static uint64_t sm_state = 0x1234567ull;
static uint64_t splitmix() {
  uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main() {
  const uint32_t nqs[] = {256, 1024, 9984, 81152, 262144, IDX_MAX};
  // 1. Any payload (a foreign key of either kind, ABORT, NEUTRAL, 0): the
  //    decoders never hand out an index outside the pair's moduli.
  uint64_t bad = 0, cases = 0;
  const uint64_t special[] = {0, 1, 2, ABORT, NEUTRAL, key_bits(KINF), key_bits(KNODIGIT), (uint64_t)IDX_MAX << 1,
                              (1ull << 51), (1ull << 51) - 1, (uint64_t)NONE_IDX << 32};
  for (uint32_t nq : nqs) {
    for (int i = 0; i < 2000000 + 11; ++i) {
      const uint64_t p = i < 11 ? special[i] : (splitmix() & PAYLOAD);
      const double k = key_from_bits(p);
      uint32_t idx;
      int32_t val;
      const int a = decode_pivot(k, nq, idx, val);
      if (a != KEY_BAD && idx >= nq) ++bad;
      const int b = decode_digit(k, nq, idx, val);
      if (b != KEY_BAD && idx >= nq) ++bad;
      cases += 2;
    }
  }
  std::printf("foreign_payloads %llu %llu\n", (unsigned long long)bad, (unsigned long long)cases);

  // 2. Key 0 (what a workgroup that gave up used to post), KINF and the
  //    recovery keys decode as elimination keys to KEY_BAD or in range.
  bad = cases = 0;
  {
    uint32_t idx;
    int32_t val;
    bad += decode_pivot(0.0, 81152, idx, val) != KEY_BAD;  // idx IDX_MAX
    bad += decode_pivot(KINF, 81152, idx, val) != KEY_INF || idx != 0 || val != INT32_MIN;
    bad += decode_digit(KNODIGIT, 81152, idx, val) != KEY_NODIGIT || idx != 0 || val != 0;
    cases += 3;
  }
  std::printf("special_keys %llu %llu\n", (unsigned long long)bad, (unsigned long long)cases);

  // 3. Round trips: elimination keys of every kind of t at indices < nq,
  //    recovery keys of every digit value at indices < nq.
  bad = cases = 0;
  for (uint32_t nq : nqs) {
    for (int i = 0; i < 1000000; ++i) {
      const uint32_t idx0 = (uint32_t)(splitmix() % nq);
      const int64_t t0 = i % 7 == 0 ? (int64_t)(i % 3) - 1 : (int64_t)(splitmix() % 0xFFFFFFFFull) - 0x7FFFFFFF;
      const double ck = 0.5 + 2.0 * (double)(IDX_MAX - idx0);
      uint32_t idx;
      int32_t val;
      const int a = decode_pivot(elim_key((double)t0, ck), nq, idx, val);
      bad += a != KEY_PIVOT || idx != idx0 || val != (int32_t)t0;
      const int b = decode_pivot(elim_key(2147483648.0, ck), nq, idx, val);  // TINF
      bad += b != KEY_INF || idx != idx0 || val != INT32_MIN;
      const int32_t g0 = (int32_t)(uint32_t)splitmix();
      const double rk = (double)idx0 * 0x1p32 + ((double)g0 + 0x1p31);
      const int c = decode_digit(rk, nq, idx, val);
      bad += c != KEY_DIGIT || idx != idx0 || val != g0;
      cases += 3;
    }
  }
  std::printf("round_trips %llu %llu\n", (unsigned long long)bad, (unsigned long long)cases);

  // 4. Order: the elimination key orders like the reference's
  //    (|t| << 32) | (m - [t >= 0]) for descending moduli (smaller index =
  //    larger modulus): by |t|, then by index, then negative before positive.
  bad = cases = 0;
  for (int i = 0; i < 1000000; ++i) {
    const uint32_t i1 = (uint32_t)(splitmix() % 4096), i2 = (uint32_t)(splitmix() % 4096);
    const int64_t t1 = (int64_t)(splitmix() % 64) - 32, t2 = (int64_t)(splitmix() % 64) - 32;
    const double k1 = elim_key((double)t1, 0.5 + 2.0 * (IDX_MAX - i1));
    const double k2 = elim_key((double)t2, 0.5 + 2.0 * (IDX_MAX - i2));
    const uint64_t a1 = (uint64_t)(t1 < 0 ? -t1 : t1), a2 = (uint64_t)(t2 < 0 ? -t2 : t2);
    // modulus of index i: any strictly decreasing map, e.g. 2^32 - 1 - 2 i
    const uint64_t r1 = (a1 << 32) | ((0xFFFFFFFFull - 2 * i1) - (t1 >= 0));
    const uint64_t r2 = (a2 << 32) | ((0xFFFFFFFFull - 2 * i2) - (t2 >= 0));
    if (!(i1 == i2 && t1 == t2)) bad += (k1 < k2) != (r1 < r2);
    ++cases;
  }
  std::printf("order %llu %llu\n", (unsigned long long)bad, (unsigned long long)cases);
  return 0;
}
