// modinv_check.cpp -- CPU run of the library's modular inverses
// (Arith32::modinv, Arith22::modinv in csrc/mgcd_arith.h) under the v_rcp_f32
// model of emu.h.
//
//   modinv_check <in> <out> <rcp_mode>
// in:  int64 triples (L, m, y), little-endian; out: int64 inverses in [0, m).
// Prints "checked N wrong W": W counts results with y * inv != 1 (mod m),
// checked with 128-bit integers.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mgcd_arith.h"

using namespace mgcd::dev;

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  g_rcp_mode = std::atoi(argv[3]);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int64_t> in;
  int64_t buf[3];
  while (std::fread(buf, 8, 3, f) == 3) in.insert(in.end(), buf, buf + 3);
  std::fclose(f);
  const size_t n = in.size() / 3;
  std::vector<int64_t> out(n);
  uint64_t wrong = 0;
  for (size_t i = 0; i < n; ++i) {
    const int L = (int)in[3 * i];
    const int64_t m = in[3 * i + 1], y = in[3 * i + 2];
    ModD M;
    M.m = (double)m;
    M.minv = 1.0 / M.m;
    const double r = L <= 22 ? Arith22::modinv((double)y, M) : Arith32::modinv((double)y, M);
    int64_t v = (int64_t)r % m;
    if (v < 0) v += m;
    out[i] = v;
    int64_t yr = y % m;
    if (yr < 0) yr += m;
    if ((unsigned __int128)yr * (unsigned __int128)v % (unsigned __int128)m != 1) ++wrong;
  }
  FILE* g = std::fopen(argv[2], "wb");
  if (!g) return 2;
  std::fwrite(out.data(), 8, n, g);
  std::fclose(g);
  std::printf("checked %zu wrong %llu\n", n, (unsigned long long)wrong);
  return 0;
}
