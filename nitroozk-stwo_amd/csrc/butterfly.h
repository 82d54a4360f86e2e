// The M31 butterflies of the circle FFT, shared by the transform kernels (fft.hip) and the
// measurement of their issue-rate ceiling (measure.hip), so that both run the same code.
//
// Twiddles are read from doubled trees (every entry 2t, below 2^32): a t (mod P) for a <= P is
// then hi + lo / 2 with a (2t) = hi 2^32 + lo, since a t = hi 2^31 + lo / 2 and 2^31 = 1 (mod P).
// The sum is below 2P, so one conditional subtraction (a minimum) makes it canonical: 5 operations
// against 6 for the product with t itself.
#pragma once
#include "field.h"

namespace sh {

__device__ __forceinline__ uint32_t mul_tw(uint32_t a, uint32_t t2) {
  const uint64_t p = (uint64_t)a * t2;
  const uint32_t s = (uint32_t)(p >> 32) + ((uint32_t)p >> 1);
  const uint32_t r = s - P;
  return r < s ? r : s;
}
// The negation of a doubled twiddle, again doubled: 2P - t2 (0 stays 0).
__device__ __forceinline__ uint32_t neg_tw(uint32_t t2) { return t2 ? 2 * P - t2 : 0u; }

// NEG: the twiddle is the negation of t2, folded into the butterfly's additions.
template <bool INV, bool NEG = false>
__device__ __forceinline__ void bfly(uint32_t& v0, uint32_t& v1, uint32_t t2) {
  const uint32_t a = v0, b = v1;
  if (INV) {  // ibutterfly(v0, v1, t) = (v0 + v1, (v0 - v1) t); with -t: (v1 - v0) t
    v0 = m31_add(a, b);
    v1 = mul_tw(NEG ? m31_sub(b, a) : m31_sub(a, b), t2);
  } else {  // butterfly(v0, v1, t) = (v0 + v1 t, v0 - v1 t); with -t the two swap
    const uint32_t m = mul_tw(b, t2);
    v0 = NEG ? m31_sub(a, m) : m31_add(a, m);
    v1 = NEG ? m31_add(a, m) : m31_sub(a, m);
  }
}

}  // namespace sh
