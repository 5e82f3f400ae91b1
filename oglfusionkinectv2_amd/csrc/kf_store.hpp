// kf_store.hpp -- TSDF storage traits shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace kf {

// ---------------------------------------------------------------------------
// TSDF storage traits.  The volume is two planes (value, weight), x fastest.
// ---------------------------------------------------------------------------
struct StoreRG16I {  // integrate.cs:37-50 (RG16I texture)
    using V = int16_t;
    static constexpr bool kScaled = true;
    __device__ static float value(V v) { return (float)v * 0.00003051944088f; }
    __device__ static float weight(V v) { return (float)v; }
    __device__ static float raw(V v) { return (float)v; }
    __device__ static V from_value(float t) { return (V)(int)(t * 32766.0f); }
    __device__ static V from_weight(float w) { return (V)(int)w; }
    // value that can end a ray (< 0) or reach the 0.8 large-step threshold
    __device__ static bool interesting(V v) { return v < 0 || v >= (V)25886; }  // 0.79 * 32766
    __device__ static bool zero_bits(V v) { return v == 0; }
    __device__ static uint32_t bits(V v) { return (uint32_t)(uint16_t)v; }
    __device__ static V from_bits(uint32_t b) { return (V)(uint16_t)b; }
};
struct StoreF32 {
    using V = float;
    static constexpr bool kScaled = false;
    __device__ static float value(V v) { return v; }
    __device__ static float weight(V v) { return v; }
    __device__ static float raw(V v) { return v; }
    __device__ static V from_value(float t) { return t; }
    __device__ static V from_weight(float w) { return w; }
    __device__ static bool interesting(V v) { return v < 0.0f || v >= 0.79f; }
    __device__ static bool zero_bits(V v) { return __float_as_uint(v) == 0u; }
    __device__ static uint32_t bits(V v) { return __float_as_uint(v); }
    __device__ static V from_bits(uint32_t b) { return __uint_as_float(b); }
};
// binary16 values held as their bit patterns: 16-bit loads then land in a
// register of their own (no packing of two halves into one register, which
// ties independent loads together)
struct StoreF16 {
    using V = uint16_t;
    static constexpr bool kScaled = false;
    __device__ static float half(V v) { return (float)__builtin_bit_cast(_Float16, v); }
    __device__ static float value(V v) { return half(v); }
    __device__ static float weight(V v) { return half(v); }
    __device__ static float raw(V v) { return half(v); }
    __device__ static V from_value(float t) { return __builtin_bit_cast(uint16_t, (_Float16)t); }
    __device__ static V from_weight(float w) { return __builtin_bit_cast(uint16_t, (_Float16)w); }
    __device__ static bool interesting(V v) { return half(v) < 0.0f || half(v) >= 0.79f; }
    __device__ static bool zero_bits(V v) { return v == 0; }
    __device__ static uint32_t bits(V v) { return (uint32_t)v; }
    __device__ static V from_bits(uint32_t b) { return (V)b; }
};

// One volume plane addressed by 32-bit byte offsets through a buffer
// resource (the base address lives in scalar registers, so an access costs
// one offset computation instead of a 64-bit address; the hardware range
// check reads 0 beyond the plane and drops stores there).  For planes whose
// size in bytes fits in 32 bits.
template <class S>
struct BufPlane {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ BufPlane(const void *p, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000))
    {
    }
    // stored bits of the element at byte offset voff + soff (soff uniform:
    // it rides in the instruction's scalar offset), zero-extended
    __device__ __forceinline__ uint32_t bits(uint32_t voff, uint32_t soff) const
    {
        if constexpr (sizeof(typename S::V) == 4)
            return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
        else
            return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
    }
    __device__ __forceinline__ void store(uint32_t voff, uint32_t soff, typename S::V v) const
    {
        if constexpr (sizeof(typename S::V) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, (int)voff, (int)soff, 0);
    }
};

}  // namespace kf
