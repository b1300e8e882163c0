// Fixed-point conversions shared by the brick scatter and the binned scatter.
// A scatter multiplies its values by a power of two (which commutes with
// FP32 rounding), rounds each weighted contribution to an integer and adds
// the integers in LDS: the sums do not depend on the order of the additions,
// so every scatter is bitwise reproducible.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tt {

typedef unsigned long long u64;

constexpr float kWScale = 1073741824.0f;  // 2^30: fixed-point scale of the weights

__device__ __forceinline__ int cvt_floor_i32(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));  // floor and convert in one VALU op
  return r;
}

// round(w * v) (half up) as a 32-bit integer (|w v| < 2^31).
__device__ __forceinline__ int to_fixed32(float w, float v) { return cvt_floor_i32(__fmaf_rn(w, v, 0.5f)); }

// round(p) (half to even) as a 64-bit integer, |p| < 2^62.  hi = trunc(p 2^-31)
// and lo = p - hi 2^31 are exact in FP32: below 2^31 hi is 0 and lo = p;
// above, p's last bit weighs at least 2^8, so lo is a multiple of 2^8 below
// 2^31 (at most 23 significant bits).
__device__ __forceinline__ long long to_fixed64(float p) {
  const float hi = truncf(p * 4.656612873077393e-10f);  // 2^-31
  const float lo = __fmaf_rn(hi, -2147483648.0f, p);
  return (long long)((u64)(long long)(int)hi << 31) + (long long)(int)rintf(lo);
}

// Split accumulators of the 32-bit contributions.  Each accumulator of a
// tile is a plane of low words followed by a plane of high words, and a
// voxel's 64-bit fixed-point sum is hi 2^32 + lo - kLoBias.  A contribution c
// (|c| <= 2^30) is added to the low word alone with a returning 32-bit LDS
// atomic (at random addresses it costs about 60 % of a 64-bit one); the high
// word changes only when the low word wraps, which needs the low word it
// replaced outside [2^30, 3 2^30).  Low words start at kLoBias = 2^31, the
// middle of that window, so wraps are rare: every returned low word is
// tested, and a lane that saw one outside the window adds the exact carries
// to the high words.  The sums are integers, so the result does not depend
// on the order of the atomics, and equals the 64-bit accumulation's.
constexpr unsigned kLoBias = 0x80000000u;

// Whether the N low words a lane's atomics returned all lay inside the
// window [2^30, 3 2^30), so that none of its additions can have carried: the
// smallest and the largest word are tested, and the compiler forms them with
// three-input v_min3_u32 / v_max3_u32, about one VALU operation per word.
template <int N>
__device__ __forceinline__ bool lo_all_in_window(const unsigned *old) {
  unsigned mn = old[0], mx = old[0];
#pragma unroll
  for (int k = 1; k < N; ++k) {
    mn = min(mn, old[k]);
    mx = max(mx, old[k]);
  }
  return mn >= 0x40000000u && mx < 0xC0000000u;
}

// Change of the high word when c is added to a low word that held old.
__device__ __forceinline__ unsigned hi_step(unsigned old, int c) {
  const bool carry = old + (unsigned)c < old;
  return carry ? (c < 0 ? 0u : 1u) : (c < 0 ? 0xffffffffu : 0u);
}

__device__ __forceinline__ u64 split_value(unsigned lo, unsigned hi) {
  return (((u64)hi << 32) | lo) - (u64)kLoBias;
}

// ceil(log2(n)) for n >= 1 (0 for n <= 1).
__host__ __device__ inline int ceil_log2(int64_t n) {
  int b = 0;
  while (b < 62 && ((int64_t)1 << b) < n) ++b;
  return b;
}

}  // namespace tt
