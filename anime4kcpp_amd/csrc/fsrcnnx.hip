// FSRCNNX<F> 2x pass, F = 8 or 16 (core/src/processor/cpu/CPUProcessor.cpp:1629-1659 and the
// F = 16 twin at :1687-1717), layer by layer, FP32 throughout:
//   head   conv5x5 1 -> F, identity                  -> feat         VALU, one pixel per thread
//   body   B - 1 x conv3x3 F -> F + PReLU                            conv layer <kPrelu>
//          conv3x3 F -> F + PReLU, conv1x1 F -> F, * 1 + feat, PReLU conv layer <kPreluConv1x1>
//   tail   conv3x3 F -> 4, identity, pixel shuffle 2x -> dst         conv layer <kShuffle>
// (Backend.hpp:225-305; Common.hpp:116-164, 199-341.)  F = 8 runs the 2-row MFMA item
// (conv8_layer), F = 16 the channel-tile mapping (convf_layer); both fuse the conv1x1 into the
// last block's epilogue as F/4 x F/16 extra MFMAs fed from the conv3x3's own D registers.
// F-channel f32 maps live in HBM between layers, as for ACNet<8> and ArtCNN.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "conv_layers.hpp"

namespace achip {

#pragma clang fp contract(off)

// ---- head: conv5x5 1 -> F, identity; one pixel per thread, 64 x 4 per workgroup ----
// Taps clamp to the image (Common.hpp:199-225).  The sum follows the reference's FMA order
// (AVX.hpp:93-124): eight lanes accumulate taps l, l + 8, l + 16 with fused multiply-adds,
// their tree sum (AVX.hpp:24-30), then tap 24's rounded product, then the bias.
template <typename TIN, int F>
__global__ __launch_bounds__(256) void fsrcnnx_head(const void* __restrict__ src, long ss, int W, int H,
                                                    float4* __restrict__ feat, long fp, const float* __restrict__ k,
                                                    const float* __restrict__ b)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    float r[25];
#pragma unroll
    for (int dy = 0; dy < 5; dy++) {
        const TIN* row = row_ptr<TIN>(src, ss, clampi(y + dy - 2, 0, H - 1));
#pragma unroll
        for (int dx = 0; dx < 5; dx++) r[dy * 5 + dx] = to_float<TIN>(row[clampi(x + dx - 2, 0, W - 1)]);
    }
    float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(feat) + static_cast<long>(y) * fp) + (F / 4) * x;
#pragma unroll
    for (int q = 0; q < F / 4; q++) {
        float o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float* kn = k + (4 * q + i) * 25;
            float s[8];
#pragma unroll
            for (int l = 0; l < 8; l++) s[l] = __fmul_rn(r[l], kn[l]);
#pragma unroll
            for (int l = 0; l < 8; l++) s[l] = fmaf(r[8 + l], kn[8 + l], s[l]);
#pragma unroll
            for (int l = 0; l < 8; l++) s[l] = fmaf(r[16 + l], kn[16 + l], s[l]);
            const float a0 = __fadd_rn(s[0], s[4]), a1 = __fadd_rn(s[1], s[5]), a2 = __fadd_rn(s[2], s[6]),
                        a3 = __fadd_rn(s[3], s[7]);
            float sum = __fadd_rn(__fadd_rn(a0, a2), __fadd_rn(a1, a3));
            sum = __fadd_rn(sum, __fmul_rn(r[24], kn[24]));
            o[i] = __fadd_rn(sum, b[4 * q + i]);
        }
        px[q] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// ---- launch ----

long fsrcnnx_feature_pitch(int f, int w) { return (static_cast<long>(w) * f * 4 + 255) / 256 * 256; }

template <int F, int MODE, typename TOUT>
static void layer(int W, int H, const float4* in, long fp, float4* out, const LayerArgs& la, hipStream_t stream)
{
    if constexpr (F == 8)
        hipLaunchKernelGGL((conv8_layer<MODE, float, TOUT>), conv8_grid(W, H), dim3(256), 0, stream, in, fp, W, H, out, la);
    else
        hipLaunchKernelGGL((convf_layer<F, MODE, TOUT>), convf_grid<F>(W, H), dim3(ConvTile<F>::NT), 0, stream, in, fp, W, H,
                           out, la);
}

template <int F, typename TIN, typename TOUT>
static hipError_t fsrcnnx_pass_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                 const float* kernels, const float* biases, const float* alphas, const float* wa,
                                 const float* w1, char* ws, hipStream_t stream)
{
    const long fp = fsrcnnx_feature_pitch(F, W);
    float4* feat = reinterpret_cast<float4*>(ws);
    float4* m0 = reinterpret_cast<float4*>(ws + fp * H);
    float4* m1 = reinterpret_cast<float4*>(ws + 2 * fp * H);
    hipLaunchKernelGGL((fsrcnnx_head<TIN, F>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H, feat,
                       fp, kernels, biases);
    const long wl = static_cast<long>(conv_operand_floats(F));
    auto body_args = [&](int l) {
        LayerArgs la{};
        la.wa = wa + (l - 1) * wl;
        la.bias = biases + F * l;
        la.alpha = alphas + F * (l - 1);
        if (l == blocks) {   // layer `blocks` fused with the conv1x1 (layer blocks + 1) and the skip
            la.w1 = w1;
            la.b1 = biases + F * (blocks + 1);
            la.alpha1 = alphas + F * blocks;
            la.feat = feat;
        }
        return la;
    };
    const float4* in = feat;
    float4* out = m0;
    auto advance = [&]() {
        in = out;
        out = out == m0 ? m1 : m0;
    };
    int l = 1;
    if constexpr (F == 8) {   // two layers per launch where the first of them is a plain PReLU block
        for (; l + 1 <= blocks; l += 2) {
            if (l + 1 < blocks)
                hipLaunchKernelGGL((conv8_pair<kPrelu, float, float>), conv8_pair_grid(W, H), dim3(kP8NT), 0, stream, in, fp, W,
                                   H, out, body_args(l), body_args(l + 1));
            else
                hipLaunchKernelGGL((conv8_pair<kPreluConv1x1, float, float>), conv8_pair_grid(W, H), dim3(kP8NT), 0, stream, in,
                                   fp, W, H, out, body_args(l), body_args(l + 1));
            advance();
        }
    }
    for (; l <= blocks; l++) {
        if (l < blocks) layer<F, kPrelu, float>(W, H, in, fp, out, body_args(l), stream);
        else layer<F, kPreluConv1x1, float>(W, H, in, fp, out, body_args(l), stream);
        advance();
    }
    LayerArgs lt{};
    lt.wa = wa + blocks * wl;
    lt.bias = biases + F * (blocks + 2);
    lt.dst = dst;
    lt.ds = ds;
    layer<F, kShuffle, TOUT>(W, H, in, fp, nullptr, lt, stream);
    return hipGetLastError();
}

template <int F>
static hipError_t fsrcnnx_types(int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                                long ds, const float* kernels, const float* biases, const float* alphas,
                                const float* wa, const float* w1, char* ws, hipStream_t stream)
{
#define CALL(TI, TO) fsrcnnx_pass_t<F, TI, TO>(blocks, src, ss, W, H, dst, ds, kernels, biases, alphas, wa, w1, ws, stream)
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

hipError_t launch_fsrcnnx(int f, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                          long ds, const float* kernels, const float* biases, const float* alphas, const float* wa,
                          const float* w1, void* workspace, hipStream_t stream)
{
    char* ws = static_cast<char*>(workspace);
    if (f == 8)
        return fsrcnnx_types<8>(blocks, stype, dtype, src, ss, W, H, dst, ds, kernels, biases, alphas, wa, w1, ws, stream);
    if (f == 16)
        return fsrcnnx_types<16>(blocks, stype, dtype, src, ss, W, H, dst, ds, kernels, biases, alphas, wa, w1, ws, stream);
    return hipErrorInvalidValue;
}

// ---- host-side operand packing (layouts read by conv_layers.hpp) ----

void pack_conv8_operands(const float* k, int cout, float* out)
{
    // lane l: D/A row i = l % 16 is (output row i / 8, channel i % 8); patch row g = l / 16
    // feeds output row i / 8 through kernel row ky = g - i / 8
    for (int t = 0; t < 3; t++)
        for (int c = 0; c < 8; c++)
            for (int lane = 0; lane < 64; lane++) {
                const int i = lane % 16, g = lane / 16, n = i % 8, ky = g - i / 8;
                out[(t * 8 + c) * 64 + lane] = (n < cout && ky >= 0 && ky <= 2) ? k[n * 72 + (ky * 3 + t) * 8 + c] : 0.0f;
            }
}

void pack_conv8_1x1(const float* k1, float* out)
{
    // K step r of lane group g: channel 4 (g & 1) + r of output row g >> 1
    for (int r = 0; r < 4; r++)
        for (int lane = 0; lane < 64; lane++) {
            const int m = lane % 16, g = lane / 16;
            out[r * 64 + lane] = (m / 8 == (g >> 1)) ? k1[(m % 8) * 8 + 4 * (g & 1) + r] : 0.0f;
        }
}

void pack_convf_operands(const float* k, int f, int cout, float* out)
{
    // output channel 16 h + lane % 16, input channel (f / 4) (lane / 16) + q; the F -> 4 tail
    // has one output tile
    const int kq = f / 4, mt = cout <= 16 ? 1 : f / 16;
    for (int p = 0; p < 9; p++)
        for (int lane = 0; lane < 64; lane++)
            for (int q = 0; q < kq; q++)
                for (int h = 0; h < mt; h++) {
                    const int o = 16 * h + lane % 16, c = kq * (lane / 16) + q;
                    out[(static_cast<size_t>(p) * 64 + lane) * (kq * mt) + q * mt + h] =
                        o < cout ? k[static_cast<size_t>(o) * 9 * f + p * f + c] : 0.0f;
                }
}

void pack_convf_1x1(const float* k1, int f, float* out)
{
    // K step 4 h + r of lane group g: input channel 16 h + 4 g + r; output channel 16 h2 + lane % 16
    const int mt = f / 16;
    for (int h = 0; h < mt; h++)
        for (int r = 0; r < 4; r++)
            for (int lane = 0; lane < 64; lane++)
                for (int h2 = 0; h2 < mt; h2++)
                    out[((static_cast<size_t>(4 * h + r)) * 64 + lane) * mt + h2] =
                        k1[(16 * h2 + lane % 16) * f + 16 * h + 4 * (lane / 16) + r];
}

} // namespace achip
