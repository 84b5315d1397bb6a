// Cross-lane sums within one 64-lane wave (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace qca {

// v of lane (lane pattern CTRL) through a DPP move of both 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum of v over the 64 lanes of the wave, the same value in every lane.  All
// 64 lanes must be active.  Butterfly within each row of 16 lanes through DPP
// moves (lane ^ 1, lane ^ 2, mirror within 8, mirror within 16: after each
// step both partners hold the same sum), then the four row sums read as
// scalars and added in a fixed order.  Replaces a ds_bpermute chain of six
// dependent LDS-unit round trips.
__device__ __forceinline__ double wave_sum64(double v) {
  v += dpp_f64<0xB1>(v);    // quad_perm(1,0,3,2)
  v += dpp_f64<0x4E>(v);    // quad_perm(2,3,0,1)
  v += dpp_f64<0x141>(v);   // row_half_mirror
  v += dpp_f64<0x140>(v);   // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

}  // namespace qca
