// mgcd_keys.h -- the keys the kernel reduces and exchanges between
// workgroups, and their decoding.  Host- and device-callable: the device
// build compiles it as HIP, the CPU tests build it with g++ (defining
// __host__, __device__ and __forceinline__ away).
//
// Keys are doubles holding exact integers, compared with v_min_f64:
//   elimination  |t| * 2^20 + 2 (IDX_MAX - idx) + [t < 0]
//   recovery     idx * 2^32 + (value + 2^31)
// For moduli in descending index order the elimination key orders exactly
// like the reference's (|t| << 32) | (m - [t >= 0]) (getGcdKernel.cu:154);
// |t| = 2^31 (TINF, TZERO) is its infinity.  KINF caps the keys of retired
// lanes so that every key fits the 52-bit payload of an exchange slot.
//
// An exchange slot holds (tag << 52) | payload.  Two payloads are not keys:
// ABORT (all ones), posted by a workgroup whose wait timed out, which makes
// every reader give up too, and NEUTRAL, the filler of unused poll lanes.
// The decoders below turn anything that is not a well-formed key for the
// pair (an index outside its N_Q moduli, |t| beyond the infinity) into
// KEY_BAD, so that no decoded value ever indexes memory out of bounds.
#pragma once

#include <cstdint>

namespace mgcd {
namespace keys {

constexpr uint32_t IDX_MAX = (1u << 19) - 1;    // moduli indices < 2^19
constexpr uint64_t PAYLOAD = (1ull << 52) - 1;  // slot bits below the tag
constexpr uint64_t ABORT = PAYLOAD;             // payload of a workgroup that gave up
constexpr uint64_t NEUTRAL = PAYLOAD - 1;       // above every key, below ABORT
constexpr uint32_t NONE_IDX = 0xFFFFFu;         // recovery: no nonzero residue in the part
constexpr uint64_t EXP52 = 0x4330000000000000ull;  // bit pattern of 2^52
constexpr double KINF = 2251799814733823.0;        // 2^51 + 2^20 - 1
constexpr double KNONE = __builtin_huge_val();     // no result (a wait timed out or aborted)

enum KeyKind : int { KEY_PIVOT = 0, KEY_INF = 1, KEY_DIGIT = 0, KEY_NODIGIT = 1, KEY_BAD = 2 };

// |t - 2^-21| * 2^20 is |t| * 2^20 - 1/2 for t > 0 and |t| * 2^20 + 1/2
// for t <= 0 (exact: t is an integer below 2^32); ck = 1/2 + 2 (IDX_MAX -
// idx) completes the key.  t = 0 gets sign bit 1, which decodes to 0 too.
__host__ __device__ __forceinline__ double elim_key(double t, double ck) {
  return __builtin_fma(__builtin_fabs(t - 0x1p-21), 0x1p20, ck);
}
__host__ __device__ __forceinline__ uint64_t key_bits(double k) {  // exact integer of a key < 2^52
  return __builtin_bit_cast(uint64_t, k + 0x1p52) & PAYLOAD;
}
__host__ __device__ __forceinline__ double key_from_bits(uint64_t x) {
  return __builtin_bit_cast(double, x | EXP52) - 0x1p52;
}

// Elimination key -> (modulus index, toSigned value) of a pair with nq
// moduli.  KEY_INF: |t| is the infinity 2^31 (val = INT32_MIN, the
// reference's MOD_INFINITY); idx is still the index it came from (the
// first pivot is recorded even when infinite, getGcdKernel.cu:374-402).
// KEY_BAD: |t| above 2^31 or idx >= nq.
__host__ __device__ __forceinline__ int decode_pivot(double key, uint32_t nq, uint32_t& idx, int32_t& val) {
  const uint64_t kb = key_bits(key);
  const uint64_t a = kb >> 20;
  idx = IDX_MAX - (uint32_t)((kb >> 1) & IDX_MAX);
  val = (kb & 1) ? (int32_t)(0u - (uint32_t)a) : (int32_t)(uint32_t)a;  // |t| = 2^31 -> INT32_MIN
  if (a > 0x80000000ull || idx >= nq) return KEY_BAD;
  return a == 0x80000000ull ? KEY_INF : KEY_PIVOT;
}

// Recovery key -> (modulus index, signed digit).  KEY_NODIGIT: no part holds
// a nonzero residue; the reference then takes its thread 0's pair with
// value 0 (getGcdKernel.cu:57-64), i.e. (0, 0).  KEY_BAD: idx >= nq.
__host__ __device__ __forceinline__ int decode_digit(double key, uint32_t nq, uint32_t& idx, int32_t& val) {
  const uint64_t kb = key_bits(key);
  idx = (uint32_t)(kb >> 32);
  val = (int32_t)((uint32_t)kb - 0x80000000u);
  if (idx == NONE_IDX) {
    idx = 0;
    val = 0;
    return KEY_NODIGIT;
  }
  return idx < nq ? KEY_DIGIT : KEY_BAD;
}
// The recovery key of a part without a nonzero residue: (NONE_IDX, 0).
constexpr double KNODIGIT = (double)NONE_IDX * 0x1p32 + 0x1p31;

}  // namespace keys
}  // namespace mgcd
