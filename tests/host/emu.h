// emu.h -- host build of the library's device arithmetic for the CPU tests.
// Included (g++ -include) before csrc/mgcd_arith.h or csrc/mgcd_keys.h:
// the device qualifiers become plain inline functions, and v_rcp_f32
// (__builtin_amdgcn_rcpf) becomes a model with the error the device
// reciprocal is certified to keep (|rcp(f) f - 1| <= 2^-23, rcp(-f) =
// -rcp(f)): the correctly rounded reciprocal moved by -1, 0 or +1 ulp,
// wherever the moved value still meets that bound.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#define __host__
#define __device__
#define __forceinline__ inline
#define __builtin_amdgcn_rcpf mgcd_emu_rcpf

// 0: no move; 1: -1/0/+1 ulp chosen by a hash of |f|; 2: always toward 0;
// 3: always away from 0.
inline int g_rcp_mode = 0;

inline float mgcd_emu_rcpf(float f) {
  const float a = std::fabs(f);
  float r = (float)(1.0 / (double)a);
  int d = 0;
  if (g_rcp_mode == 1) {
    uint32_t b;
    std::memcpy(&b, &a, 4);
    d = (int)((b * 2654435761u) >> 30) % 3 - 1;
  } else if (g_rcp_mode == 2) {
    d = -1;
  } else if (g_rcp_mode == 3) {
    d = 1;
  }
  if (d) {
    const float r2 = std::nextafter(r, d > 0 ? INFINITY : 0.0f);
    if (std::fabs((double)r2 * (double)a - 1.0) <= 0x1p-23) r = r2;
  }
  return f < 0 ? -r : r;
}
