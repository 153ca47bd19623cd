// storage.h -- storage types of the weighted projections (FP32 / FP16 /
// BF16) and their FP32 load/store.  Only storage changes; every sum is FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace atr {

struct bf16_t {
    uint16_t bits;
};

template <typename S>
__device__ __forceinline__ float ld(const S* p);
template <>
__device__ __forceinline__ float ld<float>(const float* p) { return *p; }
template <>
__device__ __forceinline__ float ld<_Float16>(const _Float16* p) { return (float)*p; }
template <>
__device__ __forceinline__ float ld<bf16_t>(const bf16_t* p)
{
    return __uint_as_float((uint32_t)p->bits << 16);
}

template <typename S>
__device__ __forceinline__ void st(S* p, float v);
template <>
__device__ __forceinline__ void st<float>(float* p, float v) { *p = v; }
// FP16 stores saturate at the largest finite half on both sides.  Sentinels
// (-1e30) then become -65504, which the kernels still classify as "empty"
// through the storage-aware threshold; a value above 65504 stays finite
// (65504) instead of becoming +inf, which would turn every difference and
// interpolation pair that reads it into inf or NaN.
__device__ __forceinline__ float half_saturate(float v)
{
    return fminf(fmaxf(v, -65504.0f), 65504.0f);
}
template <>
__device__ __forceinline__ void st<_Float16>(_Float16* p, float v)
{
    *p = (_Float16)half_saturate(v);
}
template <>
__device__ __forceinline__ void st<bf16_t>(bf16_t* p, float v)
{
    uint32_t b = __float_as_uint(v);
    b += 0x7FFFu + ((b >> 16) & 1u);  // round to nearest even
    p->bits = (uint16_t)(b >> 16);
}

// Address-space-qualified pointers.  Tables a whole wave reads alike are
// reached through constant-space pointers (scalar loads); per-lane data
// through global-space pointers (global_load / global_store, never flat).
#define ATR_CONST_AS __attribute__((address_space(4)))
#define ATR_GLOBAL_AS __attribute__((address_space(1)))
template <typename T>
using cptr = const ATR_CONST_AS T*;
template <typename T>
using gptr = ATR_GLOBAL_AS T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_cptr(const T* p) { return (cptr<T>)p; }
template <typename T>
__device__ __forceinline__ gptr<T> as_gptr(T* p) { return (gptr<T>)p; }

// Whole-object loads through qualified pointers (the address space is
// recovered by the compiler from the cast's source).
template <typename T>
__device__ __forceinline__ T ldc(cptr<T> p) { return *(const T*)p; }
template <typename T>
__device__ __forceinline__ T ldgo(const ATR_GLOBAL_AS T* p) { return *(const T*)p; }

__device__ __forceinline__ float ldg(const ATR_GLOBAL_AS float* p) { return *p; }
__device__ __forceinline__ float ldg(const ATR_GLOBAL_AS _Float16* p) { return (float)*p; }
__device__ __forceinline__ float ldg(const ATR_GLOBAL_AS bf16_t* p)
{
    return __uint_as_float((uint32_t)p->bits << 16);
}
__device__ __forceinline__ void stg(ATR_GLOBAL_AS float* p, float v) { *p = v; }
__device__ __forceinline__ void stg(ATR_GLOBAL_AS _Float16* p, float v)
{
    *p = (_Float16)half_saturate(v);
}
__device__ __forceinline__ void stg(ATR_GLOBAL_AS bf16_t* p, float v)
{
    uint32_t b = __float_as_uint(v);
    b += 0x7FFFu + ((b >> 16) & 1u);
    p->bits = (uint16_t)(b >> 16);
}

// Raw buffer access (a wave-uniform resource over one array, a per-lane byte
// offset and a scalar byte offset): no VALU for the address.  The resource's
// range is the array, so a lane offset past it reads 0 instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
template <typename S>
__device__ __forceinline__ float ldb(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ float ldb<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float ldb<_Float16>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float ldb<bf16_t>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0) << 16);
}
__device__ __forceinline__ void stb(__amdgpu_buffer_rsrc_t r, float v, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// Threshold below which a stored value counts as the reference's "empty"
// (<= -1e10).  FP16 cannot hold -1e10, so its sentinel is the most negative
// finite half; real data never gets there after weighting by cos/VolZ.
template <typename S>
__device__ __forceinline__ float empty_threshold() { return -1e10f; }
template <>
__device__ __forceinline__ float empty_threshold<_Float16>() { return -65000.0f; }

}  // namespace atr
