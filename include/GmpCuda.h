/*  GmpCuda.h -- drop-in host interface of the MI355X modular GCD.
 *
 *  Same namespace, class, constants and entry points as the reference's
 *  GmpCuda.h (GmpCuda.h:22-127): code written against
 *      GmpCuda::GmpCudaDevice::setDevice(0);
 *      GmpCuda::GmpCudaDevice dev;
 *      dev.gcd(g, u, v);
 *  compiles unchanged against this header and links against libgmpcuda.so
 *  (which calls the HIP library through the C ABI of mgcd.h).  This header
 *  is host-only: it includes <gmp.h> and no HIP header.
 *
 *  Differences, all additive:
 *   - WARP_SZ is 64 (the wavefront of gfx950; the reference asserts 32).
 *   - An extra constructor selects the modulus width (22, 27 or 32 bits)
 *     and optionally the moduli table file; the default constructor uses
 *     the 32-bit table, as the reference does.
 *   - The table is not capped at 65,536 moduli, so inputs up to about
 *     2.3 Mbit (L = 32) run instead of throwing.
 *   - gcd(g, u, 0) returns |u| and gcd(g, 0, 0) returns 0, the mpz_gcd
 *     answers, where the reference kernel cannot handle v = 0.
 */
#ifndef GMPCUDA_H
#define GMPCUDA_H

#include <gmp.h>

#include <cstdint>
#include <stdexcept>

struct mgcd_ctx;

namespace GmpCuda {
constexpr int WARP_SZ = 64;

class GmpCudaDevice {
 public:
  static constexpr int GCD_BLOCK_SZ = 256;  // moduli counts are rounded up to this
  static constexpr int MAX_THREADS = 262144;
  //  Error codes of the reference kernel's result header (GmpCuda.h:106-109).
  static constexpr uint32_t GCD_KERNEL_ERROR = 0;
  static constexpr uint32_t GCD_REDUX_ERROR = 0;
  static constexpr uint32_t GCD_RECOVERY_ERROR = 1;
  //  One signed mixed-radix digit, as the kernel hands it back (GmpCuda.h:112).
  typedef struct alignas(8) {
    uint32_t modulus;
    int32_t value;
  } pair_t;

  static void setDevice(int devNo = 0);  // throws "Invalid device number."
  GmpCudaDevice(void);
  explicit GmpCudaDevice(int modulusBits, const char* tablePath = nullptr);
  ~GmpCudaDevice();
  GmpCudaDevice(const GmpCudaDevice&) = delete;
  GmpCudaDevice& operator=(const GmpCudaDevice&) = delete;

  //  g <- gcd(u, v).  As in the reference, u and v are swapped in place when
  //  mpz_cmp(u, v) < 0, and g may alias u or v.  Throws std::runtime_error
  //  with the reference's messages.
  void gcd(mpz_t g, mpz_t u, mpz_t v);

  int getMaxGridSize(void) const { return maxGridSize; }
  int getModulusBits(void) const { return modulusBits; }
  int getNumModuli(void) const;
  double lastKernelSeconds(void) const;
  mgcd_ctx* handle(void) const { return ctx; }

 private:
  mgcd_ctx* ctx;
  int deviceNum;
  int maxGridSize;
  int modulusBits;
};

constexpr int L = 32;
constexpr int W = 64;
constexpr int NUM_MODULI = GmpCudaDevice::MAX_THREADS;

//  The 32-bit moduli table, descending (generated by createModuli).
extern const uint32_t moduli[];
}  // namespace GmpCuda

#endif  // GMPCUDA_H
