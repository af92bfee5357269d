// createModuli -- writes a moduli table file for one modulus width L.
//
//   createModuli L COUNT OUT.bin [--all-primes]   binary table loaded by the runtime
//   createModuli L COUNT --cpp                     C++ array GmpCuda::moduli[] on stdout
//
// --all-primes keeps every L-bit prime instead of those that satisfy the
// Cavagnino-Werbrouck condition (CW section 2.4; the runtime's arithmetic
// does not need it).
//
// Counterpart of createModuli.cpp:104-179 (which prints a C++ array capped at
// NUM_MODULI for the compile-time L).  Exit status 2 when fewer than COUNT
// usable moduli exist, as at createModuli.cpp:171-178.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../csrc/mgcd_table.h"

int main(int argc, char** argv) {
  const bool all_primes = argc == 5 && std::strcmp(argv[4], "--all-primes") == 0;
  if (argc != 4 && !all_primes) {
    std::fprintf(stderr, "usage: %s L COUNT OUT.bin [--all-primes] | %s L COUNT --cpp\n", argv[0], argv[0]);
    return 1;
  }
  const int L = std::atoi(argv[1]);
  const long count = std::atol(argv[2]);
  if (L < 3 || L > 32 || count <= 0) {
    std::fprintf(stderr, "L = %d is invalid (or COUNT <= 0).\n", L);
    return 1;
  }
  std::vector<uint32_t> mod = mgcd::generateModuli(L, size_t(count), all_primes);
  if (std::strcmp(argv[3], "--cpp") == 0 && !all_primes) {
    std::printf("//  AUTOMATICALLY GENERATED by createModuli: do not edit\n\n");
    std::printf("//  %zu %d-bit primes, descending, each satisfying the Cavagnino-Werbrouck\n", mod.size(), L);
    std::printf("//  condition for DBM_a(N, J) with W = 64.\n\n#include \"GmpCuda.h\"\n");
    std::printf("const uint32_t GmpCuda::moduli[] =\n{\n");
    for (uint32_t m : mod) std::printf("\t%u,\n", m);
    std::printf("};\n");
  } else if (std::strcmp(argv[3], "--cpp") == 0 || !mgcd::writeTable(argv[3], L, mod, all_primes)) {
    std::fprintf(stderr, "cannot write %s\n", argv[3]);
    return 1;
  }
  if (mod.size() < size_t(count)) {
    std::fprintf(stderr, "Only %zu usable %d-bit moduli exist; %ld were called for.\n", mod.size(), L, count);
    return 2;
  }
  std::fprintf(stderr, "%zu moduli were generated (L = %d).\n", mod.size(), L);
  return 0;
}
