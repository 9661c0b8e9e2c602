// ArtCNN<F> 2x pass, F = 16 or 32 (core/src/processor/cpu/CPUProcessor.cpp:1519-1546 and the
// F = 32 twin), layer by layer, FP32 throughout:
//   head   conv3x3 1 -> F, identity                   -> feat        VALU, one pixel per thread
//   body   B x conv3x3 F -> F + ReLU                                  convf_layer<kRelu>
//          conv3x3 F -> F, identity, + feat                           convf_layer<kAddFeat>
//   tail   conv3x3 F -> 4, identity, pixel shuffle 2x -> dst          convf_layer<kShuffle>
// (Backend.hpp:135-215; Common.hpp:116-164, 290-341.)  F-channel f32 maps live in HBM between
// layers, as for ACNet<8>; the MFMA mapping is convf_layer's (conv_layers.hpp): no MFMA row
// is padding except in the F -> 4 tail.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "conv_layers.hpp"

namespace achip {

#pragma clang fp contract(off)

// ---- head: conv3x3 1 -> F, identity; one pixel per thread, 64 x 4 per workgroup ----
template <typename TIN, int F>
__global__ __launch_bounds__(256) void artcnn_head(const void* __restrict__ src, long ss, int W, int H,
                                                   float4* __restrict__ feat, long fp, const float* __restrict__ k,
                                                   const float* __restrict__ b)
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
    float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(feat) + static_cast<long>(y) * fp) + (F / 4) * x;
#pragma unroll
    for (int q = 0; q < F / 4; q++) {
        float o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int n = 4 * q + i;
            float s = 0.0f;
#pragma unroll
            for (int p = 0; p < 9; p++) s = __fadd_rn(s, __fmul_rn(r[p], k[n * 9 + p]));
            o[i] = __fadd_rn(s, b[n]);
        }
        px[q] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// ---- launch ----

long artcnn_feature_pitch(int f, int w) { return (static_cast<long>(w) * f * 4 + 255) / 256 * 256; }

template <int F, typename TIN, typename TOUT>
static hipError_t artcnn_pass_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                const float* kernels, const float* biases, const float* wa, char* ws, hipStream_t stream)
{
    const long fp = artcnn_feature_pitch(F, W);
    float4* feat = reinterpret_cast<float4*>(ws);
    float4* m0 = reinterpret_cast<float4*>(ws + fp * H);
    float4* m1 = reinterpret_cast<float4*>(ws + 2 * fp * H);
    hipLaunchKernelGGL((artcnn_head<TIN, F>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H, feat,
                       fp, kernels, biases);
    const dim3 grid = convf_grid<F>(W, H), block(ConvTile<F>::NT);
    const long wl = static_cast<long>(conv_operand_floats(F));   // A-operand floats per layer
    const float4* in = feat;
    float4* out = m0;
    for (int l = 1; l <= blocks; l++) {         // body: F -> F + ReLU
        LayerArgs la{};
        la.wa = wa + (l - 1) * wl;
        la.bias = biases + F * l;
        hipLaunchKernelGGL((convf_layer<F, kRelu, float>), grid, block, 0, stream, in, fp, W, H, out, la);
        in = out;
        out = out == m0 ? m1 : m0;
    }
    LayerArgs la{};
    la.wa = wa + blocks * wl;
    la.bias = biases + F * (blocks + 1);
    la.feat = feat;
    hipLaunchKernelGGL((convf_layer<F, kAddFeat, float>), grid, block, 0, stream, in, fp, W, H, out, la);
    LayerArgs lt{};
    lt.wa = wa + (blocks + 1) * wl;
    lt.bias = biases + F * (blocks + 2);
    lt.dst = dst;
    lt.ds = ds;
    hipLaunchKernelGGL((convf_layer<F, kShuffle, TOUT>), grid, block, 0, stream, out, fp, W, H, nullptr, lt);
    return hipGetLastError();
}

template <int F>
static hipError_t artcnn_types(int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst, long ds,
                               const float* kernels, const float* biases, const float* wa, char* ws, hipStream_t stream)
{
#define CALL(TI, TO) artcnn_pass_t<F, TI, TO>(blocks, src, ss, W, H, dst, ds, kernels, biases, wa, ws, stream)
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

hipError_t launch_artcnn(int f, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                         long ds, const float* kernels, const float* biases, const float* wa, void* workspace,
                         hipStream_t stream)
{
    char* ws = static_cast<char*>(workspace);
    if (f == 16) return artcnn_types<16>(blocks, stype, dtype, src, ss, W, H, dst, ds, kernels, biases, wa, ws, stream);
    if (f == 32) return artcnn_types<32>(blocks, stype, dtype, src, ss, W, H, dst, ds, kernels, biases, wa, ws, stream);
    return hipErrorInvalidValue;
}

} // namespace achip
