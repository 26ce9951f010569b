// Runs of float ray-length steps taken at once.  The raycaster steps a ray
// over stretches whose samples it knows in advance with
//   while (len < lim) { last = len; len += step; }
// every addition rounded to nearest.  advance_to() returns the same len and
// last bit for bit, without one loop trip per step.  The header compiles for
// the host too, so the CPU tests check it against the plain loop.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define KF_STEPS_HD __host__ __device__
#else
#define KF_STEPS_HD
#endif

namespace kf {

// Inside a binade [2^e, 2^(e+1)) every float is an integer multiple of the
// ulp u = 2^(e-23).  For len = m*u there, the exact sum len + step is
// (m + S)*u with S = step / u, and rounding it to the u grid gives
// (m + round(S))*u whatever m is, unless frac(S) is exactly 1/2 (a tie,
// decided by the parity of m).  So, without a tie, every addition whose
// result stays below 2^(e+1) adds the same d = round(S)*u, and k of them add
// k*d exactly.  Ties and the addition that leaves the binade are taken one at
// a time, as are lengths outside the normal range this covers.
KF_STEPS_HD inline float advance_to(float len, float step, float lim, float& last) {
  while (len < lim) {
    const float next = len + step;  // the loop's own rounded addition
    const uint32_t bits = __builtin_bit_cast(uint32_t, len);
    const int ex = static_cast<int>(bits >> 23);  // biased exponent; len > 0 keeps the sign bit clear
    // u = 2^(ex - 150) and 1/u = 2^(150 - ex), both normal floats
    const bool normal = len > 0.0F && ex >= 24 && ex <= 253;
    if (!normal || !(next > len)) {
      last = len;
      len = next;
      continue;
    }
    const float u = __builtin_bit_cast(float, static_cast<uint32_t>(ex - 23) << 23);
    const float inv_u = __builtin_bit_cast(float, static_cast<uint32_t>(277 - ex) << 23);
    const float top = __builtin_bit_cast(float, static_cast<uint32_t>(ex + 1) << 23);
    const float s = step * inv_u;  // exact: a power-of-two scaling, s < 2^24
    if (s - floorf(s) == 0.5F || !(next < top)) {
      last = len;
      len = next;
      continue;
    }
    const int m = static_cast<int>((bits & 0x7FFFFFU) | 0x800000U);
    const int dm = static_cast<int>((next - len) * inv_u);  // next - len is exact (next < 2 len)
    const int kmax = ((1 << 24) - 1 - m) / dm;             // steps whose results stay below top (>= 1)
    int k = kmax + 1;                                       // lim at or beyond top: all of them, then on
    if (lim < top) {
      const int lm = static_cast<int>(lim * inv_u);  // lim lies in (len, top): a multiple of u
      k = (lm - m + dm - 1) / dm;                    // the first k with m + k*dm >= lm
    }
    const int take = k <= kmax ? k : kmax;
    last = static_cast<float>(m + (take - 1) * dm) * u;
    len = static_cast<float>(m + take * dm) * u;
  }
  return len;
}

}  // namespace kf
