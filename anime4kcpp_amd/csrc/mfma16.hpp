// The 16-bit MFMA helpers shared by the fused ACNetLegacy kernel (acnet_mfma.hip) and the
// ACNet<8> 16-bit layer kernels (acnet8_mfma16.hip): v_mfma_f32_16x16x32_{f16,bf16} with FP32
// accumulation, the 16-bit operand vectors, and the round-to-16-bit packing of results.
#pragma once

#include "acnet_device.hpp"

namespace achip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

template <typename T> struct Mma;
template <> struct Mma<_Float16> {
    using v8 = f16x8;
    using v2 = f16x2;
    __device__ static f32x4 mma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct Mma<__bf16> {
    using v8 = bf16x8;
    using v2 = bf16x2;
    __device__ static f32x4 mma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

__device__ inline f32x4 mma_f32(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// relu(round16(lo)), relu(round16(hi)) packed: v_cvt_pk_{f16,bf16}_f32 (RNE), then
// v_pk_max_i16 with 0 (negative values, -0 included, have the sign bit set).
template <typename T16>
__device__ inline uint32_t relu_pack(float lo, float hi)
{
    using v2 = typename Mma<T16>::v2;
    const v2 h = __builtin_convertvector((f32x2){ lo, hi }, v2);
    const i16x2 r = __builtin_elementwise_max(__builtin_bit_cast(i16x2, h), (i16x2){ 0, 0 });
    return __builtin_bit_cast(uint32_t, r);
}

// round16(lo), round16(hi) packed (RNE)
template <typename T16>
__device__ inline uint32_t pack16(float lo, float hi)
{
    using v2 = typename Mma<T16>::v2;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){ lo, hi }, v2));
}

} // namespace achip
