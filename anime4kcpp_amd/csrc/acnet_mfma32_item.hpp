// The FP32 MFMA work item shared by the fused ACNetLegacy kernel (acnet_mfma32.hip) and the
// ACNet<8> layer kernels (acnet8.hip): 16 consecutive pixels x 2 output rows of an 8 -> 8
// (or 8 -> 4, zero-padded) conv3x3 from a 4 x 3 input patch on v_mfma_f32_16x16x4_f32.
// D[16 x 16] = A[16 x 96] B[96 x 16]: lane l holds, in D, rows 4(l/16) .. 4(l/16)+3 of
// column l%16 -- output channels 4*(g & 1) .. +3 of output row g >> 1 (g = l/16); in B,
// patch row g; one MFMA covers one (patch column t, channel c).  The A operands are per-lane
// weight blobs laid out as in acnet_device.hpp.
#pragma once

#include "acnet_device.hpp"

namespace achip {

typedef float f32x4m __attribute__((ext_vector_type(4)));

__device__ inline f32x4m mma32(float a, float b, f32x4m c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// B operands of one item: the 8 channels of this lane's pixel in patch columns t = 0..2
struct Item32 {
    float4 lo[3], hi[3];
};

__device__ __forceinline__ Item32 load_item(const float4* __restrict__ pa, const float4* __restrict__ pb, int p)
{
    Item32 q;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        q.lo[t] = pa[p + t];
        q.hi[t] = pb[p + t];
    }
    return q;
}

// the 24 MFMAs of one item on two accumulator chains, the first starting from the bias
__device__ __forceinline__ f32x4m item24(const Item32& q, const float (&a)[24], f32x4m bias)
{
    f32x4m c0 = bias, c1 = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int t = 0; t < 3; t++) {
        c0 = mma32(a[t * 8 + 0], q.lo[t].x, c0);
        c1 = mma32(a[t * 8 + 1], q.lo[t].y, c1);
        c0 = mma32(a[t * 8 + 2], q.lo[t].z, c0);
        c1 = mma32(a[t * 8 + 3], q.lo[t].w, c1);
        c0 = mma32(a[t * 8 + 4], q.hi[t].x, c0);
        c1 = mma32(a[t * 8 + 5], q.hi[t].y, c1);
        c0 = mma32(a[t * 8 + 6], q.hi[t].z, c0);
        c1 = mma32(a[t * 8 + 7], q.hi[t].w, c1);
    }
    return c0 + c1;
}

} // namespace achip
