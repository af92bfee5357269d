// horner_check.cpp -- CPU run of the library's residue Horner step
// (horner_step in csrc/mgcd_arith.h, the role of modMP at
// getGcdKernel.cu:297-324) and the reduction after the last step, as the
// residue loop of csrc/mgcd_kernels.hip runs them.
//
//   horner_check <in>
// in: int64 records (m, n, limb[n-1], ..., limb[0]) little-endian, most
// significant limb first.  For every record the chain s = 0, ls = 0, then
// s = horner_step(s, ls), ls = limb * 2^32 over the limbs, then the last
// limb added, reduce and canon, is compared with the residue computed in
// 128-bit integers; every intermediate s is also checked against the step's
// bound |s| <= m/2 + 2^13 and for being an integer.  Prints
// "checked N wrong W bound B".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "mgcd_arith.h"

using namespace mgcd::dev;

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int64_t> in;
  int64_t buf[4096];
  size_t got;
  while ((got = std::fread(buf, 8, 4096, f)) > 0) in.insert(in.end(), buf, buf + got);
  std::fclose(f);
  uint64_t checked = 0, wrong = 0, bound = 0;
  for (size_t i = 0; i + 2 <= in.size();) {
    const int64_t m = in[i], n = in[i + 1];
    i += 2;
    if (n < 0 || i + (size_t)n > in.size()) return 3;
    ModD M;
    M.m = (double)m;
    M.minv = 1.0 / M.m;
    const double lim = (double)m / 2 + 8192.0;
    double s = 0.0, ls = 0.0;
    unsigned __int128 exact = 0;
    for (int64_t j = 0; j < n; ++j) {
      const uint32_t limb = (uint32_t)in[i + (size_t)j];
      s = horner_step(s, ls, M);
      ls = (double)limb * 4294967296.0;
      if (!(std::fabs(s) <= lim) || s != std::rint(s)) ++bound;
      exact = ((exact << 32) + limb) % (unsigned __int128)m;
    }
    i += (size_t)n;
    const double c = canon(reduce(s + ls * 0x1p-32, M), M);
    ++checked;
    if (!(c >= 0.0 && c < (double)m) || (uint64_t)c != (uint64_t)exact) ++wrong;
  }
  std::printf("checked %llu wrong %llu bound %llu\n", (unsigned long long)checked, (unsigned long long)wrong,
              (unsigned long long)bound);
  return 0;
}
