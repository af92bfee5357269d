// mgcd_arith.h -- exact modular arithmetic for one modulus per lane, gfx950.
//
// Residues are held as doubles in a "signed reduced" form: an integer r with
// r == x (mod m) and |r| <= m/2 + 2^13.  Every value fits a double exactly
// (m < 2^32), and the full-rate FP64 FMA of CDNA4 does the reductions:
//
//   mulmod(a, b): h = fl(a*b), l = fma(a, b, -h) (exact low part),
//                 q = rint(h * minv), r = fma(-q, m, h) + l.
//     For |a|, |b| <= m + 2^14 the product is < 2^64.5, so h differs from
//     a*b by at most 2^11, |h*minv - a*b/m| < 2^-19, and |a*b - q*m| <=
//     m/2 + 2^13.  fma(-q, m, h) is then an integer below 2^53 and exact,
//     and adding l is exact.
//   reduce(x):    q = rint(x * minv), r = fma(-q, m, x)  (|x| < 2^52).
//   canon(x):     x + m if x < 0 else x  -> [0, m).
//
// The only inverse computed is modinv (Knuth's Algorithm X with remainders
// and cofactors kept exactly, modInv.h:141-188), whose quotients come from
// the FP32 reciprocal v_rcp_f32 as quasi-quotients:
//
//   q~(x, y) = floor( fl32(x) * (rcp(fl32(y)) * QBIAS) ),  QBIAS = 1 - 2^-20
//
// The bias makes q~ never exceed floor(x/y) (the combined relative error of
// the conversions, v_rcp_f32 and two roundings is below 2^-21), so every
// remainder x - q~ y is non-negative; it is computed exactly with an FMA,
// and one exact integer correction (r >= y: r -= y) makes the quotient
// exact whenever q~ >= floor(x/y) - 1.  The correction is what guarantees
// progress: without it, x and y that round to the same float give q~ = 0
// and the pair would only swap, forever.  When floor(x/y) >= 2^19 the
// estimate may fall short by more; the remainder then stays >= y, the next
// step has quotient 0 (a swap) and division resumes, still exact.
// The "never exceeds" postcondition, and "short by at most one when
// floor(x/y) < 2^19", are certified exhaustively on the device by
// mgcd_certify_quasi_quotient (float domain x < 2^23, y < 2^22 point by
// point, and all widths up to 32 bits through the monotone endpoints).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mgcd {
namespace dev {

constexpr float QBIAS = 1.0f - 0x1p-20f;

struct ModD {
  double m;     // modulus
  double minv;  // fl(1/m)
};

__device__ __forceinline__ double reduce(double x, ModD M) {
  const double q = __builtin_rint(x * M.minv);
  return __builtin_fma(-q, M.m, x);
}

__device__ __forceinline__ double canon(double x, ModD M) { return x < 0.0 ? x + M.m : x; }

// Signed reduced -> the reference's toSigned range [-(m+1)/2, (m-3)/2]
// (getGcdKernel.cu:278-281).
__device__ __forceinline__ int32_t to_signed(double r, ModD M) {
  const double x = canon(r, M);
  const double half = __builtin_fma(M.m, 0.5, -0.5);  // (m-1)/2 exactly
  return (int32_t)(x >= half ? x - M.m : x);
}

// Quasi-quotient from the FP32 reciprocal (see file comment).
__device__ __forceinline__ float quasi_quo_f(float xf, float yf) {
  const float c = __builtin_amdgcn_rcpf(yf) * QBIAS;
  return __builtin_floorf(xf * c);
}

// Moduli up to 32 bits: products need the two-product low part.
struct Arith32 {
  static __device__ __forceinline__ double mulmod(double a, double b, ModD M) {
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    const double q = __builtin_rint(h * M.minv);
    return __builtin_fma(-q, M.m, h) + l;
  }
  // x <- x mod y (exactly, see quasi_quo_f), sx <- sx - q sy, fx <- fl32(x).
  static __device__ __forceinline__ void eea_step(double& x, float& fx, double& sx, double y, float fy, double sy) {
    double q = (double)quasi_quo_f(fx, fy);
    const double r = __builtin_fma(-q, y, x);
    q += (r >= y) ? 1.0 : 0.0;  // exact integer correction: q~ was floor(x/y) - 1
    x = __builtin_fma(-q, y, x);
    sx = __builtin_fma(-q, sy, sx);
    fx = (float)x;
  }
  // y in [1, m) canonical; returns y^-1 mod m in [0, m).
  // Invariant r_i == s_i * y (mod m); unrolled by two so the pair swaps
  // roles instead of moving registers; stops when a remainder is <= 1.
  static __device__ double modinv(double y, ModD M) {
    double r0 = M.m, r1 = y, s0 = 0.0, s1 = 1.0, s;
    float f0 = (float)M.m, f1 = (float)y;
    for (;;) {
      eea_step(r0, f0, s0, r1, f1, s1);
      if (!(r0 > 1.0)) { s = (r0 == 1.0) ? s0 : s1; break; }
      eea_step(r1, f1, s1, r0, f0, s0);
      if (!(r1 > 1.0)) { s = (r1 == 1.0) ? s1 : s0; break; }
    }
    return s < 0.0 ? s + M.m : s;
  }
};

// Moduli below 2^22: residues, products (< 2^44) and the whole extended
// Euclid fit FP32/FP64 exactly, so the inverse runs in FP32 only.
struct Arith22 {
  static __device__ __forceinline__ double mulmod(double a, double b, ModD M) {
    const double h = a * b;  // exact: |a*b| < 2^46
    const double q = __builtin_rint(h * M.minv);
    return __builtin_fma(-q, M.m, h);
  }
  static __device__ __forceinline__ void eea_step(float& x, float& sx, float y, float sy) {
    float q = quasi_quo_f(x, y);
    const float r = __builtin_fmaf(-q, y, x);
    q += (r >= y) ? 1.0f : 0.0f;  // exact integer correction (q < 2^22: the sum is exact)
    x = __builtin_fmaf(-q, y, x);
    sx = __builtin_fmaf(-q, sy, sx);
  }
  static __device__ double modinv(double y, ModD M) {
    const float m = (float)M.m;
    float r0 = m, r1 = (float)y, s0 = 0.0f, s1 = 1.0f, s;
    for (;;) {
      eea_step(r0, s0, r1, s1);
      if (!(r0 > 1.0f)) { s = (r0 == 1.0f) ? s0 : s1; break; }
      eea_step(r1, s1, r0, s0);
      if (!(r1 > 1.0f)) { s = (r1 == 1.0f) ? s1 : s0; break; }
    }
    return (double)(s < 0.0f ? s + m : s);
  }
};

// One Horner step of the residue reduction: (r * 2^32 + limb) mod m, signed
// reduced in and out (modMP, getGcdKernel.cu:297-324).  t = r*2^32 is exact;
// x = t + limb is rounded but only feeds the quotient estimate; t - q*m is
// an integer below 2^34 and exact in the FMA.
__device__ __forceinline__ double horner_step(double r, double limb, ModD M) {
  const double t = r * 4294967296.0;
  const double x = t + limb;
  const double q = __builtin_rint(x * M.minv);
  return __builtin_fma(-q, M.m, t) + limb;
}

}  // namespace dev
}  // namespace mgcd
