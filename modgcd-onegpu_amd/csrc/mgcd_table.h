// mgcd_table.h -- moduli tables: generation and the on-disk file format.
//
// A table holds the largest L-bit primes D, in descending order, that satisfy
// the Cavagnino-Werbrouck condition for the DBM_a(N, J) division-by-multiply
// method at word width W = 64 (createModuli.cpp:143-164): with
// FC = 2^(W+L-1), J = FC/D + 1, Qcr = ceil(J / (D*J - FC)), Ncr = Qcr*D - 1,
// D is usable iff Ncr >= 2^W.  The walk order (odd integers downward from
// 2^L - 1, createModuli.cpp:40-44,150-154) and the filter are the same as the
// reference generator's, so a table's first entries equal the reference's
// compiled-in GmpCuda::moduli[] for the same L.  What differs: the sieve is
// segmented (only the top of the range is visited), the criterion uses
// unsigned __int128 instead of GMP, the count is not capped at 65,536, and
// the result is written to a binary file that the runtime loads.
//
// An "all primes" table (CW section 2.4) keeps every L-bit prime in the
// same descending walk: this build's FP64 arithmetic is exact for any odd
// modulus below 2^32 and needs no division criterion, so the filter only
// keeps tables (and pivot sequences) equal to the reference's; without it,
// L = 22 has 140,336 moduli instead of 97,358.
//
// File layout (little endian):
//   MgcdTableHeader (48 bytes) followed by uint32_t moduli[count].
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mgcd {

struct MgcdTableHeader {
  char magic[8];        // "MGCDMODS"
  uint32_t version;     // 1
  uint32_t L;           // modulus width in bits (every modulus is in [2^(L-1), 2^L))
  uint32_t W;           // 64: word width of the division criterion; 0: every prime (no criterion)
  uint32_t count;       // number of moduli that follow
  uint64_t checksum;    // FNV-1a 64 over the moduli array bytes
  uint64_t reserved[2];
};
static_assert(sizeof(MgcdTableHeader) == 48, "table header layout");

inline uint64_t fnv1a64(const void* data, size_t n) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// Cavagnino-Werbrouck usability test for divisor D at widths (W, L)
// (createModuli.cpp:143-164, restated with 128-bit integers).
inline bool cwUsable(uint32_t D, int L, int W = 64) {
  typedef unsigned __int128 u128;
  const u128 FC = u128(1) << (W + L - 1);
  const u128 J = FC / D + 1;
  const u128 DJ_FC = u128(D) * J - FC;                 // in [1, D]
  const u128 Qcr = (J + DJ_FC - 1) / DJ_FC;            // ceil(J / DJ_FC)
  const u128 Ncr = Qcr * D - 1;
  return (Ncr >> W) != 0;                              // sizeinbase(Ncr,2) > W
}

// Generate up to `count` usable L-bit primes in descending order (every
// prime when all_primes).  Returns the moduli found (fewer than `count` if
// the range is exhausted).
inline std::vector<uint32_t> generateModuli(int L, size_t count, bool all_primes = false) {
  std::vector<uint32_t> out;
  if (L < 3 || L > 32 || count == 0) return out;
  const uint64_t top = (uint64_t(1) << L) - 1;          // 2^L - 1 (odd)
  const uint64_t bottom = uint64_t(1) << (L - 1);       // 2^(L-1)
  // Small odd primes up to sqrt(2^L).
  uint32_t lim = 1;
  while (uint64_t(lim) * lim < (uint64_t(1) << L)) ++lim;
  std::vector<char> small(lim + 1, 1);
  std::vector<uint32_t> primes;
  for (uint32_t i = 3; i <= lim; i += 2) {
    if (!small[i]) continue;
    primes.push_back(i);
    for (uint64_t j = uint64_t(i) * i; j <= lim; j += 2 * i) small[j] = 0;
  }
  // Segments of SEG odd numbers walking downward: segment covers odd x in
  // [hi - 2*(SEG-1), hi].
  const uint64_t SEG = 1u << 20;
  std::vector<char> comp(SEG);
  uint64_t hi = top;
  while (out.size() < count && hi >= bottom) {
    uint64_t lo = (hi >= bottom + 2 * (SEG - 1)) ? hi - 2 * (SEG - 1) : ((bottom | 1));
    if (lo < bottom) lo = bottom | 1;
    const uint64_t n = (hi - lo) / 2 + 1;              // odd numbers lo, lo+2, ..., hi
    std::fill(comp.begin(), comp.begin() + n, 0);
    for (uint32_t p : primes) {
      if (uint64_t(p) * p > hi) break;
      // first odd multiple of p that is >= lo
      uint64_t m = (lo + p - 1) / p * p;
      if ((m & 1) == 0) m += p;
      if (m < uint64_t(p) * p) m = uint64_t(p) * p;
      for (; m <= hi; m += 2 * p) comp[(m - lo) / 2] = 1;
    }
    for (uint64_t k = n; k-- > 0 && out.size() < count;) {
      if (comp[k]) continue;
      const uint32_t D = uint32_t(lo + 2 * k);
      if (all_primes || cwUsable(D, L)) out.push_back(D);
    }
    if (lo <= (bottom | 1)) break;
    hi = lo - 2;
  }
  return out;
}

inline bool writeTable(const std::string& path, int L, const std::vector<uint32_t>& moduli, bool all_primes = false) {
  MgcdTableHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, "MGCDMODS", 8);
  h.version = 1;
  h.L = uint32_t(L);
  h.W = all_primes ? 0 : 64;
  h.count = uint32_t(moduli.size());
  h.checksum = fnv1a64(moduli.data(), moduli.size() * sizeof(uint32_t));
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
            std::fwrite(moduli.data(), sizeof(uint32_t), moduli.size(), f) == moduli.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok;
}

// Reads and validates a table file.  Returns an empty string on success,
// otherwise a message naming what is wrong.
inline std::string readTable(const std::string& path, int* L, std::vector<uint32_t>* moduli) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return "cannot open moduli table " + path;
  MgcdTableHeader h;
  if (std::fread(&h, sizeof h, 1, f) != 1) { std::fclose(f); return "short moduli table header in " + path; }
  if (std::memcmp(h.magic, "MGCDMODS", 8) != 0 || h.version != 1) { std::fclose(f); return "bad moduli table magic/version in " + path; }
  if (h.L < 3 || h.L > 32 || (h.W != 64 && h.W != 0)) { std::fclose(f); return "bad moduli width in " + path; }
  moduli->resize(h.count);
  size_t got = std::fread(moduli->data(), sizeof(uint32_t), h.count, f);
  std::fclose(f);
  if (got != h.count) return "truncated moduli table " + path;
  if (fnv1a64(moduli->data(), moduli->size() * sizeof(uint32_t)) != h.checksum) return "checksum mismatch in " + path;
  const uint64_t lo = uint64_t(1) << (h.L - 1), hi = uint64_t(1) << h.L;
  for (size_t i = 0; i < moduli->size(); ++i) {
    const uint32_t m = (*moduli)[i];
    if (m < lo || m >= hi || (m & 1) == 0 || (i && m >= (*moduli)[i - 1])) return "moduli table not descending odd L-bit values: " + path;
  }
  *L = int(h.L);
  return "";
}

}  // namespace mgcd
