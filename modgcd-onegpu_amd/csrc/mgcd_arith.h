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
// The only inverse computed is modinv, an extended Euclid with remainders
// and cofactors kept exactly (the role of modInv.h:141-188), in its
// least-absolute-remainder form: each step subtracts the multiple of y
// nearest to x, so remainders are signed and |r| <= |y|/2 when the quotient
// is exact.  That takes about ln(phi)/ln(2) = 0.69 of the steps of the
// floor-quotient form (12.7 against 18.2 on average for m near 2^32;
// Knuth, TAOCP vol. 2, 4.5.3, on least remainders).  Quotients come from the
// FP32 reciprocal v_rcp_f32:
//
//   q(x, y) = rint( fl32(x) * rcp(fl32(y)) )
//
// Any integer q keeps r == s*y (mod m) exact, so the estimate only decides
// how fast |r| falls.  With rcp's relative error at most 2^-23 and three
// roundings of 2^-24, |q - x/y| <= 1/2 + |x/y| * 2^-21.3, hence
//   |r| < |y| when |x/y| < 2^20,  and  |r| < |x| otherwise,
// and with |x| < |y| the step is a swap (q = 0) or leaves |r| <= |x|; so
// |x| + |y| falls at least every second step and the loop ends.  These
// postconditions of the actual device estimate are certified exhaustively
// by mgcd_certify_quasi_quotient (modes 0 and 1), and the rcp error bound
// and its sign symmetry over every float in [1, 2^34) by mode 2.
// Below 2^22 (Arith22) x and y are exact floats and the estimate is off by
// less than 1/4 (exact when y is a power of two), so |r| < 3|y|/4; with
// |r_i s_(i+1) - r_(i+1) s_i| = m the cofactors then stay below 1.6 m, and
// every FP32 FMA of the loop is exact.
//
// The header includes no HIP header: its includer provides the device
// qualifiers (hip_runtime.h), and the CPU tests build it with g++ and a
// model of v_rcp_f32.
#pragma once

#include <cstdint>

namespace mgcd {
namespace dev {

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

// Nearest quotient estimate from the FP32 reciprocal (see file comment).
__device__ __forceinline__ float near_quo_f(float xf, float yf) {
  return __builtin_rintf(xf * __builtin_amdgcn_rcpf(yf));
}

// Moduli up to 32 bits: products need the two-product low part.
struct Arith32 {
  // mulmod is exact for operands below 2^33 in magnitude, not only reduced ones.
  static constexpr bool kWideOperands = true;
  static __device__ __forceinline__ double mulmod(double a, double b, ModD M) {
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    const double q = __builtin_rint(h * M.minv);
    return __builtin_fma(-q, M.m, h) + l;
  }
  // x <- x - q y, sx <- sx - q sy with q the nearest quotient estimate
  // (exact: both are integers below 2^34), fx <- fl32(x).
  static __device__ __forceinline__ void eea_step(double& x, float& fx, double& sx, double y, float fy, double sy) {
    const double q = (double)near_quo_f(fx, fy);
    x = __builtin_fma(-q, y, x);
    sx = __builtin_fma(-q, sy, sx);
    fx = (float)x;
  }
  // 0 < |y| < m, either sign (q(x, -y) = -q(x, y): the steps mirror those
  // for |y|); returns y^-1 mod m, signed reduced (|r| < m/2 + 1: every
  // caller only multiplies with it).
  // Invariant r_i == s_i * y (mod m); unrolled by two so the pair swaps
  // roles instead of moving registers; stops when |r| <= 1.  The last
  // nonzero remainder is +-1 (gcd(m, y) = 1): the inverse is its cofactor
  // times its sign.
  static __device__ double modinv(double y, ModD M) {
    double r0 = M.m, r1 = y, s0 = 0.0, s1 = 1.0, s;
    float f0 = (float)M.m, f1 = (float)y;
    // The exit tests read the FP32 copies: for an integer r, |fl32(r)| > 1
    // exactly when |r| > 1 (rounding is monotone and 1, 2 are floats), and
    // the 32-bit compare issues at twice the FP64 rate.  One test per two
    // steps: where the first step already reached |r0| <= 1, the second
    // divides by r0 = +-1 (never 0: the remainders keep gcd(m, y) = 1, and
    // the first step cannot give 0 since an odd m > 2^24 is not a float)
    // and leaves r0, s0 alone, which the selection below then takes.
    for (;;) {
      eea_step(r0, f0, s0, r1, f1, s1);
      eea_step(r1, f1, s1, r0, f0, s0);
      if (!(__builtin_fabsf(f0) > 1.0f) || !(__builtin_fabsf(f1) > 1.0f)) break;
    }
    // The first remainder to reach |r| <= 1 is r0's if r0 is there (the
    // step after it left r0 as it was), else r1's; the other holds the one
    // before it.
    const bool last0 = !(__builtin_fabsf(f0) > 1.0f);
    const double rl = last0 ? r0 : r1, sl = last0 ? s0 : s1;
    const double rp = last0 ? r1 : r0, sp = last0 ? s1 : s0;
    s = rl != 0.0 ? rl * sl : rp * sp;
    return reduce(s, M);
  }
};

// Moduli below 2^22: residues, products (< 2^44) and the whole extended
// Euclid fit FP32/FP64 exactly, so the inverse runs in FP32 only.
struct Arith22 {
  // mulmod needs |a * b| < 2^53: operands below 2^31 in magnitude only.
  static constexpr bool kWideOperands = false;
  static __device__ __forceinline__ double mulmod(double a, double b, ModD M) {
    const double h = a * b;  // exact: |a*b| < 2^46
    const double q = __builtin_rint(h * M.minv);
    return __builtin_fma(-q, M.m, h);
  }
  static __device__ __forceinline__ void eea_step(float& x, float& sx, float y, float sy) {
    const float q = near_quo_f(x, y);
    x = __builtin_fmaf(-q, y, x);
    sx = __builtin_fmaf(-q, sy, sx);
  }
  static __device__ double modinv(double y, ModD M) {
    float r0 = (float)M.m, r1 = (float)y, s0 = 0.0f, s1 = 1.0f, s;
    for (;;) {
      eea_step(r0, s0, r1, s1);
      if (!(__builtin_fabsf(r0) > 1.0f)) break;
      eea_step(r1, s1, r0, s0);
      if (!(__builtin_fabsf(r1) > 1.0f)) break;
    }
    const bool last0 = !(__builtin_fabsf(r0) > 1.0f);  // as in Arith32::modinv
    const float rl = last0 ? r0 : r1, sl = last0 ? s0 : s1;
    const float rp = last0 ? r1 : r0, sp = last0 ? s1 : s0;
    s = rl != 0.0f ? rl * sl : rp * sp;
    return reduce((double)s, M);
  }
};

// One Horner step of the residue reduction, r * 2^32 + limb (mod m) (modMP,
// getGcdKernel.cu:297-324), with the limb added one step late: the running
// residue is s + L for the state s and the last limb L, and the step takes
// s and ls = L * 2^32 (the tile holds limbs scaled so):
//   t = s * 2^32 + ls = (s + L) * 2^32   exact: s + L is an integer < 2^34,
//   q = rint(t * minv),  s' = t - q*m    an integer, exact in the FMA.
// t < 2^65 puts fl(t * minv) within t/m * 2^-52 <= 2^13 / m of t / m, so
// |s'| <= m/2 + 2^13 (signed reduced, as above) whatever the step count;
// after the last step the caller adds the last limb and reduces once.  Four FP64 operations
// (reducing r * 2^32 + limb to signed reduced every step takes six).
__device__ __forceinline__ double horner_step(double s, double ls, ModD M) {
  const double t = __builtin_fma(s, 4294967296.0, ls);
  const double q = __builtin_rint(t * M.minv);
  return __builtin_fma(-q, M.m, t);
}

}  // namespace dev
}  // namespace mgcd
