// Raw buffer loads for the z-march kernels: a scalar resource (base, byte
// bound) per plane and a per-thread 32-bit byte offset; loads past the bound
// return zero, so a clamped or stale offset can never fault.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgx {

using Rsrc = __amdgpu_buffer_rsrc_t;

// dword 3 of a raw (unformatted, stride 0) buffer resource on gfx9
constexpr int kRawBufferWord3 = 0x00020000;

__device__ __forceinline__ Rsrc make_rsrc(const char* base, int64_t bytes) {
    const int64_t b = bytes < 0 ? 0 : (bytes > 0xffffffffLL ? 0xffffffffLL : bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, (int)(uint32_t)b, kRawBufferWord3);
}

// Resource over a fixed number of bytes (a plane or two of an array): no
// per-call clamp arithmetic; the caller guarantees the bytes exist.
__device__ __forceinline__ Rsrc make_rsrc_n(const char* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, (int)bytes, kRawBufferWord3);
}

template <class T>
__device__ __forceinline__ T bld(Rsrc r, uint32_t off) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// with a scalar byte offset soff added (a plane of a whole-array resource)
template <class T>
__device__ __forceinline__ T bld(Rsrc r, uint32_t off, uint32_t soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}

// Two consecutive elements (off, off + sizeof(T)) in one load; off needs only
// the element's alignment.
template <class T>
struct Two {
    T x, y;
};
template <class T>
__device__ __forceinline__ Two<T> bld2(Rsrc r, uint32_t off) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(Two<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    else
        return __builtin_bit_cast(Two<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <class T>
__device__ __forceinline__ Two<T> bld2(Rsrc r, uint32_t off, uint32_t soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(Two<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
    else
        return __builtin_bit_cast(Two<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
// store of two consecutive elements
template <class T>
__device__ __forceinline__ void bst2(Rsrc r, uint32_t off, uint32_t soff, T a, T b) {
    const Two<T> v{a, b};
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r,
                                               off, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                              off, soff, 0);
}

}  // namespace mgx
