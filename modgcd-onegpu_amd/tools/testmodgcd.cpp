// testmodgcd -- command-line check and timing of GmpCudaDevice::gcd against
// GMP mpz_gcd, with the interface of the reference harness
// (testmodgcd.cpp:32-51):
//
//   testmodgcd [ r ] [ i increment ] [ f final_bits ] num_bits [ gcd_bits ]
//
// Ten pairs per size.  Inputs come from GMP's Mersenne-Twister state with
// its default seed, drawn in the same order as the reference harness
// (three mpz_urandomb calls per pair, testmodgcd.cpp:217-230), so the same
// command line produces the same pairs.  Without `r` the pairs are cached
// in tests/<bits>-<gcd_bits> as base-62 lines (testmodgcd.cpp:173-258) and
// replayed from there on later runs, which reads files the reference
// harness wrote too.  The environment variable MGCD_L selects the modulus
// width (22, 27 or 32; default 32).  Unlike the reference, the exit status
// is 1 when any result differs from mpz_gcd.
//
// Built with -DMGCD_NO_GPU (bin/testmodgcd-nogpu) it times mpz_gcd alone on
// the same pairs and needs neither a GPU nor the library, like the
// reference's testmodgcd-nogpu (Makefile:61-62, testmodgcd.cpp:132-133).
#include <sys/stat.h>
#include <time.h>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>

#include "GmpCuda.h"

static long long monotonicNs(void) {
  struct timespec now;
  clock_gettime(CLOCK_MONOTONIC_RAW, &now);
  return now.tv_sec * 1000000000LL + now.tv_nsec;
}

static const int NUM_REPS = 10;

static void drawPair(mpz_t u, mpz_t v, mpz_t g, gmp_randstate_t st, unsigned bits, unsigned gbits) {
  do mpz_urandomb(u, st, bits - gbits); while (mpz_sgn(u) == 0);
  do mpz_urandomb(v, st, bits - gbits); while (mpz_sgn(v) == 0);
  do mpz_urandomb(g, st, gbits); while (mpz_sgn(g) == 0);
  mpz_mul(u, g, u);
  mpz_mul(v, g, v);
}

int main(int argc, char* argv[]) {
  unsigned num_bits = 0, num_g_bits = 1, increment = 0, final_bits = 0;
  bool random = false;
  if (argc == 1) {
    std::cout << "Usage: testmodgcd [ r ] [ i increment ] [ f final ] num_bits [ gcd_size ]" << std::endl;
    return 0;
  }
  if (argv[1][0] == 'r') { random = true; argc -= 1; argv += 1; }
  if (argc > 2 && argv[1][0] == 'i') { increment = std::strtoul(argv[2], nullptr, 10); argc -= 2; argv += 2; final_bits = UINT_MAX; }
  if (argc > 2 && argv[1][0] == 'f') { final_bits = std::strtoul(argv[2], nullptr, 10); argc -= 2; argv += 2; }
  if (argc == 3) num_g_bits = std::strtoul(argv[2], nullptr, 10);
  if (argc == 2 || argc == 3) num_bits = std::strtoul(argv[1], nullptr, 10);
  else return 1;
  if (num_bits <= num_g_bits) { std::cerr << "num_bits must exceed gcd_bits" << std::endl; return 1; }
  const char* envL = std::getenv("MGCD_L");
  const int L = envL ? std::atoi(envL) : 32;

  std::cout << "GMP version " << gmp_version << "." << std::endl;
#if defined(MGCD_NO_GPU)
  (void)L;
  std::cout << "Executing tests only on CPU." << std::endl;
#else
  long long ttime;
  try {
    GmpCuda::GmpCudaDevice::setDevice(0);
  } catch (std::runtime_error& e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  ttime = -monotonicNs();
  GmpCuda::GmpCudaDevice dev(L);
  ttime += monotonicNs();
  std::cout << "Max grid size = " << dev.getMaxGridSize() << std::endl
            << "Modulus width L = " << L << ", moduli available = " << dev.getNumModuli() << std::endl
            << "Initialization time: " << ttime / 1e6 << " ms." << std::endl;
#endif

  mpz_t u, v, g, mod_g;
  mpz_inits(u, v, g, mod_g, NULL);
  gmp_randstate_t state;
  gmp_randinit_mt(state);
  int totalErrs = 0;
  do {
    std::cout << "***************************** Input size = " << num_bits << " (" << num_bits / 1024.0
              << " Kibit); gcd bits = " << num_g_bits << " *****************************" << std::endl;
    long long mtime = 0;
#if !defined(MGCD_NO_GPU)
    long long atime = 0;
#endif
    int numErrs = 0;
    FILE* pFile = nullptr;
    bool newFile = false;
    if (!random) {
      struct stat sb;
      if (!(stat("tests", &sb) == 0 && S_ISDIR(sb.st_mode)) && mkdir("tests", 0775) != 0) {
        std::cerr << "Error creating directory!" << std::endl;
        return 3;
      }
      const std::string file = "tests/" + std::to_string(num_bits) + "-" + std::to_string(num_g_bits);
      if (!(stat(file.c_str(), &sb) == 0 && S_ISREG(sb.st_mode))) {
        pFile = std::fopen(file.c_str(), "a+");
        newFile = true;
      } else {
        pFile = std::fopen(file.c_str(), "r");
      }
      if (!pFile) { std::cerr << "Error opening " << file << std::endl; return 4; }
    }
#if !defined(MGCD_NO_GPU)
    bool warmup = true;
#endif
    for (int i = 0; i < NUM_REPS; ++i) {
      if (random || newFile) {
        drawPair(u, v, g, state, num_bits, num_g_bits);
        if (newFile) {
          mpz_out_str(pFile, 62, u);
          std::fputc('\n', pFile);
          mpz_out_str(pFile, 62, v);
          std::fputc('\n', pFile);
          std::fflush(pFile);
        }
      } else if (mpz_inp_str(u, pFile, 62) == 0 || mpz_inp_str(v, pFile, 62) == 0) {
        std::cerr << "Error reading cached pair " << i << std::endl;
        return 4;
      }
      mtime -= monotonicNs();
      mpz_gcd(g, u, v);
      mtime += monotonicNs();
#if !defined(MGCD_NO_GPU)
      try {
        if (warmup) { dev.gcd(mod_g, u, v); warmup = false; }
        atime -= monotonicNs();
        dev.gcd(mod_g, u, v);
        atime += monotonicNs();
      } catch (std::runtime_error& e) {
        std::cout << e.what() << std::endl;
        return 1;
      }
      if (mpz_cmp(g, mod_g) != 0) {
        std::cout << "Test " << i << " error\n";
        gmp_printf("GMP_GCD = %Zx\nMOD_GCD = %Zx, size = %zu\n", g, mod_g, mpz_size(mod_g));
        numErrs += 1;
      }
#endif
    }
    if (pFile) std::fclose(pFile);
    std::cout << std::fixed << std::setprecision(3);
#if !defined(MGCD_NO_GPU)
    if (numErrs == 0) std::cout << "Modular gcd correct\n";
    std::cout << "MOD time(avg): " << atime / NUM_REPS / 1e6 << " ms\n";
#endif
    std::cout << "GMP time     : " << mtime / NUM_REPS / 1e6 << " ms\n";
#if !defined(MGCD_NO_GPU)
    std::cout << "MOD/GMP ratio: " << (double)atime / (double)mtime << std::endl;
#endif
    totalErrs += numErrs;
    num_bits += increment;
  } while (increment && num_bits <= final_bits);
  std::cout << std::endl;
  gmp_randclear(state);
  mpz_clears(u, v, g, mod_g, NULL);
  return totalErrs ? 1 : 0;
}
