// test_dropin -- exercises GmpCuda::GmpCudaDevice the way reference code
// uses it (testmodgcd.cpp:139-150, 268-275) and checks each result against
// mpz_gcd.  Prints one line per check; exit status 0 when all pass.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "GmpCuda.h"

static int failures = 0;

static void check(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  std::fflush(stdout);
  if (!ok) ++failures;
}

static bool sameAsGmp(GmpCuda::GmpCudaDevice& dev, mpz_t u, mpz_t v) {
  mpz_t g, want, uc, vc;
  mpz_inits(g, want, uc, vc, NULL);
  mpz_set(uc, u);
  mpz_set(vc, v);
  mpz_gcd(want, u, v);
  dev.gcd(g, uc, vc);
  const bool ok = mpz_cmp(g, want) == 0;
  mpz_clears(g, want, uc, vc, NULL);
  return ok;
}

int main() {
  // Invalid device number (GmpCudaDevice.cu:99-102).
  try {
    GmpCuda::GmpCudaDevice::setDevice(9999);
    check(false, "setDevice(9999) throws");
  } catch (std::runtime_error& e) {
    check(std::strcmp(e.what(), "Invalid device number.") == 0, "setDevice(9999) throws \"Invalid device number.\"");
  }
  GmpCuda::GmpCudaDevice::setDevice(0);
  GmpCuda::GmpCudaDevice dev;  // default: 32-bit moduli
  check(dev.getMaxGridSize() > 0, "getMaxGridSize() > 0");
  check(dev.getModulusBits() == 32 && dev.getNumModuli() >= 65536, "default table is 32-bit, >= 65536 moduli");
  check(GmpCuda::moduli[0] == 4294967291u && GmpCuda::moduli[1] == 4294967279u, "GmpCuda::moduli[] head");

  gmp_randstate_t st;
  gmp_randinit_mt(st);
  mpz_t u, v, g, x;
  mpz_inits(u, v, g, x, NULL);

  // u = g a, v = g b at several sizes.
  bool all = true;
  for (unsigned bits : {100u, 1000u, 8192u, 30000u, 100000u}) {
    std::printf("  gcd of %u-bit pair\n", bits);
    std::fflush(stdout);
    mpz_urandomb(g, st, bits / 8 + 1);
    mpz_urandomb(u, st, bits);
    mpz_urandomb(v, st, bits);
    mpz_mul(u, u, g);
    mpz_mul(v, v, g);
    all = all && sameAsGmp(dev, u, v);
  }
  check(all, "gcd equals mpz_gcd for 100..100000-bit pairs");

  // Side effects kept from the reference (GmpCudaDevice-gcd.cu:61-62): u and
  // v swap when u < v; g may alias an input.
  mpz_set_ui(u, 12);
  mpz_set_ui(v, 18);
  dev.gcd(g, u, v);
  check(mpz_cmp_ui(g, 6) == 0 && mpz_cmp_ui(u, 18) == 0 && mpz_cmp_ui(v, 12) == 0, "gcd(12, 18) = 6, u and v swapped");
  mpz_set_ui(u, 1071);
  mpz_set_ui(v, 462);
  dev.gcd(u, u, v);
  check(mpz_cmp_ui(u, 21) == 0, "g aliases u");

  // Cases the reference kernel cannot take; results follow mpz_gcd.
  mpz_urandomb(u, st, 5000);
  mpz_set_ui(v, 0);
  dev.gcd(g, u, v);
  check(mpz_cmp(g, u) == 0, "gcd(u, 0) = |u|");
  mpz_set_ui(u, 0);
  dev.gcd(g, u, v);
  check(mpz_sgn(g) == 0, "gcd(0, 0) = 0");
  mpz_urandomb(u, st, 3000);
  mpz_urandomb(v, st, 3000);
  mpz_neg(u, u);
  mpz_gcd(x, u, v);
  dev.gcd(g, u, v);
  check(mpz_cmp(g, x) == 0, "negative operand: gcd(-u, v) = mpz_gcd");

  // Other modulus widths.
  for (int L : {22, 27}) {
    GmpCuda::GmpCudaDevice d(L);
    mpz_urandomb(g, st, 2000);
    mpz_urandomb(u, st, 20000);
    mpz_urandomb(v, st, 20000);
    mpz_mul(u, u, g);
    mpz_mul(v, v, g);
    check(sameAsGmp(d, u, v), L == 22 ? "22-bit table: gcd equals mpz_gcd" : "27-bit table: gcd equals mpz_gcd");
  }

  // A moduli table that does not exist.
  try {
    GmpCuda::GmpCudaDevice bad(32, "/nonexistent/moduli.bin");
    check(false, "missing table throws");
  } catch (std::runtime_error&) {
    check(true, "missing table throws");
  }

  mpz_clears(u, v, g, x, NULL);
  gmp_randclear(st);
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
