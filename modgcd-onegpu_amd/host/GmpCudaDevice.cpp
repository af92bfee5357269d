// GmpCudaDevice.cpp -- the drop-in GmpCuda::GmpCudaDevice over the C ABI.
//
// Replaces GmpCudaDevice.cu (setDevice, constructor, destructor) and the
// host half of GmpCudaDevice-gcd.cu (operand staging, error decoding, mpz
// reconstruction).  All device work happens behind mgcd.h.
#include "GmpCuda.h"

#include <string>
#include <vector>

#include "mgcd.h"

using namespace GmpCuda;

namespace {

int g_device = 0;  // selected by setDevice, used by the constructors

[[noreturn]] void raise(int status) {
  std::string msg = mgcd_strerror(status);
  if (status >= MGCD_ERR_HIP) {
    const char* d = mgcd_last_error_detail();
    if (d && *d) msg += std::string(" (") + d + ")";
  }
  throw std::runtime_error(msg);
}

void exportLimbs(std::vector<uint32_t>& out, const mpz_t x) {
  size_t n = (mpz_sizeinbase(x, 2) + 31) / 32;
  out.assign(n + 1, 0);
  size_t cnt = 0;
  if (mpz_sgn(x) != 0) mpz_export(out.data(), &cnt, -1, sizeof(uint32_t), 0, 0, x);
  out.resize(cnt);
}

}  // namespace

// GmpCudaDevice.cu:96-103
void GmpCudaDevice::setDevice(int devNo) {
  const int st = mgcd_set_device(devNo);
  if (st != MGCD_OK) raise(st);
  g_device = devNo;
}

// GmpCudaDevice.cu:107-138
GmpCudaDevice::GmpCudaDevice(void) : GmpCudaDevice(32, nullptr) {}

GmpCudaDevice::GmpCudaDevice(int bits, const char* tablePath)
    : ctx(nullptr), deviceNum(g_device), maxGridSize(0), modulusBits(bits) {
  const int st = mgcd_ctx_create(&ctx, deviceNum, bits, tablePath);
  if (st != MGCD_OK) raise(st);
  maxGridSize = mgcd_ctx_max_grid(ctx);
}

GmpCudaDevice::~GmpCudaDevice() { mgcd_ctx_destroy(ctx); }

int GmpCudaDevice::getNumModuli(void) const { return mgcd_ctx_num_moduli(ctx); }
double GmpCudaDevice::lastKernelSeconds(void) const { return mgcd_ctx_last_kernel_seconds(ctx); }

// GmpCudaDevice-gcd.cu:57-132
void GmpCudaDevice::gcd(mpz_t g, mpz_t u, mpz_t v) {
  if (mpz_cmp(u, v) < 0) mpz_swap(u, v);  // caller-visible side effect kept (:61-62)
  std::vector<uint32_t> a, b;
  exportLimbs(a, u);
  exportLimbs(b, v);
  std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1);
  size_t nr = 0;
  const int st = mgcd_gcd_limbs(ctx, a.data(), a.size(), b.data(), b.size(), r.data(), r.size(), &nr);
  if (st != MGCD_OK) raise(st);
  mpz_import(g, nr, -1, sizeof(uint32_t), 0, 0, r.data());  // g written last: may alias u or v
}
