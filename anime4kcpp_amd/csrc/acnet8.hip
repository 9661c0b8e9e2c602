// ACNet<8> 2x pass (core/src/processor/cpu/CPUProcessor.cpp:1417-1436) on gfx950, layer by
// layer, FP32 throughout:
//   L0     conv3x3 1 -> 8 + PReLU            (Backend.hpp:58-75)        VALU, one pixel per thread
//   L1..LB conv3x3 8 -> 8 + PReLU            (Backend.hpp:95-98)        FP32 MFMA items
//   final  conv3x3 8 -> 4, pixel shuffle 2x, + the source pixel  (Backend.hpp:116-133,
//          Common.hpp:290-341)                                          FP32 MFMA items
// Features between layers are 8-channel f32 maps in HBM (32 bytes per pixel, rows padded to
// 256 bytes).  With a halo of one pixel per layer there is no recompute, and at 8 TB/s the
// 64 bytes a pixel moves per layer cost less than the layer's MFMA time; the depth (4, 8
// or 18 blocks) is a loop on the host.  Borders replicate the edge pixel at every layer,
// as the reference's clamped taps (Common.hpp:120-160) do.
#include <utility>

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "acnet_mfma32_item.hpp"

namespace achip {

#pragma clang fp contract(off)

// PReLU (Common.hpp:42-49): max(v, 0) + alpha * min(v, 0), each step rounded
__device__ inline float prelu(float v, float alpha)
{
    return __fadd_rn(fmaxf(v, 0.0f), __fmul_rn(alpha, fminf(v, 0.0f)));
}

// ---- L0: one pixel per thread, 64 x 4 pixels per workgroup ----
template <typename TIN>
__global__ __launch_bounds__(256) void acnet8_l0(const void* __restrict__ src, long ss, int W, int H,
                                                 float4* __restrict__ feat, long fp, const float* __restrict__ wts)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    float r[9];
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
        const TIN* row = row_ptr<TIN>(src, ss, clampi(y + dy - 1, 0, H - 1));
#pragma unroll
        for (int dx = 0; dx < 3; dx++) r[dy * 3 + dx] = to_float<TIN>(row[clampi(x + dx - 1, 0, W - 1)]);
    }
    float o[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        float s = 0.0f;
#pragma unroll
        for (int p = 0; p < 9; p++) s = __fadd_rn(s, __fmul_rn(r[p], wts[kA8OffW0 + p * 8 + n]));
        o[n] = prelu(__fadd_rn(s, wts[kA8OffB0 + n]), wts[kA8OffAlpha0 + n]);
    }
    float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(feat) + static_cast<long>(y) * fp) + 2 * x;
    px[0] = make_float4(o[0], o[1], o[2], o[3]);
    px[1] = make_float4(o[4], o[5], o[6], o[7]);
}

// ---- conv3x3 8 -> 8 (+ PReLU), or the final 8 -> 4 with pixel shuffle and residual ----
// Tile: 64 x 16 output pixels, 256 threads (4 waves); the input region (66 x 18 pixels with
// the one-pixel halo) staged in LDS as two float4 planes (channels 0-3, 4-7) with a row pitch
// of 80 pixels (a multiple of 16: the 16-byte item reads run at full rate).  Items: 4
// segments of 16 pixels x 8 row pairs = 32 per tile, 8 per wave.
constexpr int kA8TW = 64, kA8TH = 16, kA8PW = 80, kA8PH = kA8TH + 2;

template <bool FINAL, typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void acnet8_conv(const float4* __restrict__ fin, long fp, int W, int H,
                                                   float4* __restrict__ fout, const float* __restrict__ wts,
                                                   const float* __restrict__ wa, int layer,
                                                   const void* __restrict__ src, long ss, void* __restrict__ dst, long ds)
{
    __shared__ float4 pa[kA8PW * kA8PH];
    __shared__ float4 pb[kA8PW * kA8PH];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kA8TW, y0 = ty * kA8TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;

    // stage rows y0-1 .. y0+16, columns x0-1 .. x0+64 (replicated at the image edges)
    for (int e = tid; e < kA8PH * (kA8TW + 2); e += 256) {
        const int r = e / (kA8TW + 2), c = e - r * (kA8TW + 2);
        const int sy = clampi(y0 - 1 + r, 0, H - 1), sx = clampi(x0 - 1 + c, 0, W - 1);
        const float4* px = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(fin) + static_cast<long>(sy) * fp) + 2 * sx;
        pa[r * kA8PW + c] = px[0];
        pb[r * kA8PW + c] = px[1];
    }
    // this layer's A operands, bias and PReLU slopes
    float a[24];
    const float* al = wa + static_cast<long>(layer - 1) * kMfma32LayerFloats + lane;
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = al[s * 64];
    f32x4m bias;
    float alpha[4];
    if constexpr (FINAL) {
        const float* bl = wts + kA8OffFinalB;
        bias = half ? (f32x4m){ 0.f, 0.f, 0.f, 0.f } : (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
    } else {
        const float* bl = wts + kA8OffLayer + (layer - 1) * 16 + 4 * half;
        bias = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
#pragma unroll
        for (int r = 0; r < 4; r++) alpha[r] = bl[8 + r];
    }
    __syncthreads();

    const int gx1 = min(x0 + kA8TW, W), gy1 = min(y0 + kA8TH, H);
    const int nseg = (gx1 - x0 + 15) >> 4;
    const int nitems = nseg * ((gy1 - y0 + 1) >> 1);
    const int lane_rd = g * kA8PW + j;
    for (int item = wave; item < nitems; item += 4) {
        const int pair = item / nseg, seg = item - pair * nseg;
        // patch row g of output rows (y0 + 2 pair, +1) is staged row 2 pair + g; column j + t
        const Item32 q = load_item(pa, pb, 2 * pair * kA8PW + 16 * seg + lane_rd);
        const f32x4m acc = item24(q, a, bias);
        const int x = x0 + 16 * seg + j, y = y0 + 2 * pair + ro;
        if constexpr (!FINAL) {
            if (x < gx1 && y < gy1) {
                float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + static_cast<long>(y) * fp) + 2 * x + half;
                *px = make_float4(prelu(acc[0], alpha[0]), prelu(acc[1], alpha[1]), prelu(acc[2], alpha[2]),
                                  prelu(acc[3], alpha[3]));
            }
        } else {
            // lanes with half = 0 hold the four sub-pixels n = 0..3 (output (2x + n % 2,
            // 2y + n / 2)); half = 1 lanes take row 2y + 1 from them
            const float n2 = __shfl_xor(acc[2], 16), n3 = __shfl_xor(acc[3], 16);
            if (x < gx1 && y < gy1) {
                const float res = to_float<TIN>(row_ptr<TIN>(src, ss, y)[x]);
                const float s0 = half ? n2 : acc[0], s1 = half ? n3 : acc[1];
                TOUT* orow = row_ptr<TOUT>(dst, ds, 2 * y + half);
                store_pair<TOUT>(orow + 2 * x, __fadd_rn(__fmul_rn(s0, 1.0f), res), __fadd_rn(__fmul_rn(s1, 1.0f), res));
            }
        }
    }
}

// ---- launch ----

template <typename TIN, typename TOUT>
static hipError_t acnet8_pass_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                const float* wts, const float* wa, float4* f0, float4* f1, long fp, hipStream_t stream)
{
    hipLaunchKernelGGL((acnet8_l0<TIN>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H, f0, fp, wts);
    const dim3 grid((W + kA8TW - 1) / kA8TW, (H + kA8TH - 1) / kA8TH), block(256);
    float4* in = f0;
    float4* out = f1;
    for (int l = 1; l <= blocks; l++) {
        hipLaunchKernelGGL((acnet8_conv<false, float, float>), grid, block, 0, stream, in, fp, W, H, out, wts, wa, l,
                           src, ss, dst, ds);
        std::swap(in, out);
    }
    hipLaunchKernelGGL((acnet8_conv<true, TIN, TOUT>), grid, block, 0, stream, in, fp, W, H, out, wts, wa,
                       blocks + 1, src, ss, dst, ds);
    return hipGetLastError();
}

long acnet8_feature_pitch(int w) { return (static_cast<long>(w) * 32 + 255) / 256 * 256; }

hipError_t launch_acnet8(int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst, long ds,
                         const float* wts, const float* wa, void* workspace, hipStream_t stream)
{
    const long fp = acnet8_feature_pitch(W);
    float4* f0 = static_cast<float4*>(workspace);
    float4* f1 = reinterpret_cast<float4*>(static_cast<char*>(workspace) + fp * H);
#define CALL(TI, TO) acnet8_pass_t<TI, TO>(blocks, src, ss, W, H, dst, ds, wts, wa, f0, f1, fp, stream)
    switch (stype) {
    case U8: switch (dtype) {
        case U8: return CALL(uint8_t, uint8_t); case U16: return CALL(uint8_t, uint16_t);
        case F16: return CALL(uint8_t, __half); case F32: return CALL(uint8_t, float);
        default: return hipErrorInvalidValue; }
    case U16: switch (dtype) {
        case U8: return CALL(uint16_t, uint8_t); case U16: return CALL(uint16_t, uint16_t);
        case F16: return CALL(uint16_t, __half); case F32: return CALL(uint16_t, float);
        default: return hipErrorInvalidValue; }
    case F16: switch (dtype) {
        case U8: return CALL(__half, uint8_t); case U16: return CALL(__half, uint16_t);
        case F16: return CALL(__half, __half); case F32: return CALL(__half, float);
        default: return hipErrorInvalidValue; }
    case F32: switch (dtype) {
        case U8: return CALL(float, uint8_t); case U16: return CALL(float, uint16_t);
        case F16: return CALL(float, __half); case F32: return CALL(float, float);
        default: return hipErrorInvalidValue; }
    default: return hipErrorInvalidValue;
    }
#undef CALL
}

} // namespace achip
