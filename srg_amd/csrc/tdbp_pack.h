// tdbp_pack.h -- the storage conversion of range-compressed samples shared by
// the pack kernel and the fused range-compression store: per pulse, scale =
// max |component| / 2^14 (1 for an all-zero pulse); each component is
// multiplied by 1/scale and rounded to FP16 (hardware conversion, nearest
// even) or BF16 (nearest even on the FP32 bits).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tdbp.h"

namespace tdbp {

__device__ __forceinline__ float pulse_scale(float maxabs) {
  return maxabs > 0.f ? maxabs * (1.0f / 16384.0f) : 1.0f;
}

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
