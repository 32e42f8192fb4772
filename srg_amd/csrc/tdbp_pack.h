// tdbp_pack.h -- the storage conversion of range-compressed samples shared by
// the pack kernel and the fused range-compression store.  Per pulse, scale =
// the power of two 2^e with max |component| / scale in [8192, 16384) (0 for
// an all-zero pulse); each component is multiplied by 1/scale (exact) and
// rounded to FP16 (hardware conversion, nearest even) or BF16 (nearest even
// on the FP32 bits).  k_unify_scales then gives every pulse within 2^16 of
// the largest scale of the call that largest scale (an exact power-of-two
// rescale of its words; all-zero pulses just take it), so the pulses of a
// burst normally share one scale and the backprojection applies it once per
// stage of pulses instead of once per pixel-pulse pair.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tdbp.h"

namespace tdbp {

__device__ __forceinline__ float pulse_scale(float maxabs) {
  if (!(maxabs > 0.f)) return 0.f;
  const int E = (int)((__float_as_uint(maxabs) >> 23) & 0xffu);
  if (E == 0xff) return 1.f;                      // inf/NaN: no scaling
  return __uint_as_float((unsigned)max(E - 13, 1) << 23);
}
// 1/scale (exact for a power of two); 0 for the zero scale of an all-zero pulse.
__device__ __forceinline__ float pulse_inv(float sc) { return sc > 0.f ? 1.0f / sc : 0.f; }

__device__ __forceinline__ uint32_t pack_word(float re, float im, int dtype) {
  uint32_t w;
  if (dtype == TDBP_CF16) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    h2 h;
    h.x = (_Float16)re;
    h.y = (_Float16)im;
    __builtin_memcpy(&w, &h, 4);
  } else {
    uint32_t a = __float_as_uint(re), c = __float_as_uint(im);
    a = (a + 0x7fffu + ((a >> 16) & 1u)) >> 16;
    c = (c + 0x7fffu + ((c >> 16) & 1u)) >> 16;
    w = (a & 0xffffu) | (c << 16);
  }
  return w;
}

}  // namespace tdbp
