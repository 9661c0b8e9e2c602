// ArtCNN<F> 2x pass, F = 16 or 32 (core/src/processor/cpu/CPUProcessor.cpp:1519-1546 and the
// F = 32 twin), layer by layer, FP32 throughout:
//   head   conv3x3 1 -> F, identity                   -> feat        VALU, one pixel per thread
//   body   B x conv3x3 F -> F + ReLU                                  FP32 MFMA
//          conv3x3 F -> F, identity, + feat                           FP32 MFMA
//   tail   conv3x3 F -> 4, identity, pixel shuffle 2x -> dst          FP32 MFMA
// (Backend.hpp:135-215; Common.hpp:116-164, 290-341.)  F-channel f32 maps live in HBM between
// layers, as for ACNet<8>.
//
// MFMA mapping (v_mfma_f32_16x16x4_f32): an item is 16 consecutive pixels of one row; D rows
// are 16 output channels (two 16-row tiles for F = 32), columns the pixels, K runs over the 9
// taps x F input channels.  Lane group g supplies input channels g*F/4 .. g*F/4 + F/4 - 1 of
// its pixel -- one (F = 16) or two (F = 32) 16-byte LDS reads per tap from channel-quad
// planes -- and K step q of a tap takes the q-th of them.  No MFMA row is padding except in
// the F -> 4 tail.  A wave computes 4 items at once, loading each tap's weights (per-lane A
// operands) once for all 4.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "acnet_mfma32_item.hpp"

namespace achip {

#pragma clang fp contract(off)

template <int F>
struct ArtTile {
    static constexpr int TW = F == 16 ? 64 : 32, TH = 16, NT = 512, NW = NT / 64;
    static constexpr int PW = (TW + 2 + 15) / 16 * 16, PH = TH + 2;    // plane pitch: 16-pixel multiple
    static constexpr int NQ = F / 4;                                    // channel-quad planes
    static constexpr int KQ = F / 4;                                    // K steps per tap (channels per lane group)
    static constexpr int MT = F / 16;                                   // 16-channel output tiles
    static constexpr int NI = 4;                                        // items per wave batch
    static constexpr int SEGS = TW / 16, ITEMS = SEGS * TH;
    static constexpr int BATCHES = (ITEMS / NW + NI - 1) / NI;
};

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

enum ArtMode : int { kArtRelu = 0, kArtAdd = 1, kArtTail = 2 };

// ---- conv3x3 F -> F (ReLU, or identity + feat) or the F -> 4 tail ----
// wa: this layer's A operands, [tap][lane][KQ * MT] floats; bias: F (or 4) floats.
template <int F, int MODE, typename TOUT>
__global__ __launch_bounds__(512) void artcnn_conv(const float4* __restrict__ fin, long fp, int W, int H,
                                                   float4* __restrict__ fout, const float4* __restrict__ feat,
                                                   const float* __restrict__ wa, const float* __restrict__ bias,
                                                   void* __restrict__ dst, long ds)
{
    using C = ArtTile<F>;
    __shared__ float4 planes[C::NQ][C::PH * C::PW];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * C::TW, y0 = ty * C::TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15;

    // stage rows y0-1 .. y0+TH, columns x0-1 .. x0+TW (replicated at the image edges)
    for (int e = tid; e < C::PH * (C::TW + 2); e += C::NT) {
        const int r = e / (C::TW + 2), c = e - r * (C::TW + 2);
        const int sy = clampi(y0 - 1 + r, 0, H - 1), sx = clampi(x0 - 1 + c, 0, W - 1);
        const float4* px = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(fin) + static_cast<long>(sy) * fp) + C::NQ * sx;
#pragma unroll
        for (int q = 0; q < C::NQ; q++) planes[q][r * C::PW + c] = px[q];
    }
    f32x4m b0[C::MT];
#pragma unroll
    for (int h = 0; h < C::MT; h++) {
        if (MODE == kArtTail)   // output channels 0..3 only: rows 0-3 of the first tile
            b0[h] = (h == 0 && g == 0) ? (f32x4m){ bias[0], bias[1], bias[2], bias[3] } : (f32x4m){ 0.f, 0.f, 0.f, 0.f };
        else
            b0[h] = (f32x4m){ bias[16 * h + 4 * g], bias[16 * h + 4 * g + 1], bias[16 * h + 4 * g + 2], bias[16 * h + 4 * g + 3] };
    }
    __syncthreads();

    const int gx1 = min(x0 + C::TW, W), gy1 = min(y0 + C::TH, H);
    for (int batch = 0; batch < C::BATCHES; batch++) {
        f32x4m acc[C::NI][C::MT];
        int base[C::NI];
#pragma unroll
        for (int i = 0; i < C::NI; i++) {
            const int item = wave + C::NW * (batch * C::NI + i);
            const int row = item / C::SEGS, seg = item - row * C::SEGS;
            // item pixels (x0 + 16 seg + j, y0 + row): staged at (16 seg + j + 1, row + 1)
            base[i] = row * C::PW + 16 * seg + j;
#pragma unroll
            for (int h = 0; h < C::MT; h++) acc[i][h] = b0[h];
        }
#pragma unroll 1
        for (int p = 0; p < 9; p++) {
            const int dy = p / 3, dx = p - dy * 3;
            float a[C::KQ * C::MT];
            const float4* ap = reinterpret_cast<const float4*>(wa + (static_cast<long>(p) * 64 + lane) * (C::KQ * C::MT));
#pragma unroll
            for (int v = 0; v < C::KQ * C::MT / 4; v++) {
                const float4 f = ap[v];
                a[4 * v] = f.x; a[4 * v + 1] = f.y; a[4 * v + 2] = f.z; a[4 * v + 3] = f.w;
            }
#pragma unroll
            for (int i = 0; i < C::NI; i++) {
                const int off = base[i] + dy * C::PW + dx;
                float bv[C::KQ];
#pragma unroll
                for (int v = 0; v < C::KQ / 4; v++) {
                    const float4 f = planes[(C::KQ / 4) * g + v][off];
                    bv[4 * v] = f.x; bv[4 * v + 1] = f.y; bv[4 * v + 2] = f.z; bv[4 * v + 3] = f.w;
                }
#pragma unroll
                for (int q = 0; q < C::KQ; q++)
#pragma unroll
                    for (int h = 0; h < C::MT; h++) acc[i][h] = mma32(a[q * C::MT + h], bv[q], acc[i][h]);
            }
        }
        // epilogue: lane (g, j) holds output channels 16 h + 4 g .. + 3 of pixel j
#pragma unroll
        for (int i = 0; i < C::NI; i++) {
            const int item = wave + C::NW * (batch * C::NI + i);
            const int row = item / C::SEGS, seg = item - row * C::SEGS;
            const int x = x0 + 16 * seg + j, y = y0 + row;
            if (x >= gx1 || y >= gy1) continue;
            if constexpr (MODE == kArtTail) {
                if (g == 0) {   // the 4 sub-pixels n: output (2x + n % 2, 2y + n / 2)
                    const f32x4m v = acc[i][0];
                    store_pair<TOUT>(row_ptr<TOUT>(dst, ds, 2 * y) + 2 * x, v[0], v[1]);
                    store_pair<TOUT>(row_ptr<TOUT>(dst, ds, 2 * y + 1) + 2 * x, v[2], v[3]);
                }
            } else {
                float4* po = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + static_cast<long>(y) * fp) + C::NQ * x;
                const float4* pf = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(feat) + static_cast<long>(y) * fp) + C::NQ * x;
#pragma unroll
                for (int h = 0; h < C::MT; h++) {
                    const f32x4m v = acc[i][h];
                    float4 o;
                    if constexpr (MODE == kArtRelu) {
                        o = make_float4(fmaxf(v[0], 0.f), fmaxf(v[1], 0.f), fmaxf(v[2], 0.f), fmaxf(v[3], 0.f));
                    } else {   // identity, then * 1 + feat (Common.hpp:151-155)
                        const float4 r = pf[4 * h + g];
                        o = make_float4(__fadd_rn(__fmul_rn(v[0], 1.0f), r.x), __fadd_rn(__fmul_rn(v[1], 1.0f), r.y),
                                        __fadd_rn(__fmul_rn(v[2], 1.0f), r.z), __fadd_rn(__fmul_rn(v[3], 1.0f), r.w));
                    }
                    po[4 * h + g] = o;
                }
            }
        }
    }
}

// ---- launch ----

long artcnn_feature_pitch(int f, int w) { return (static_cast<long>(w) * f * 4 + 255) / 256 * 256; }

template <int F, typename TIN, typename TOUT>
static hipError_t artcnn_pass_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                const float* kernels, const float* biases, const float* wa, char* ws, hipStream_t stream)
{
    using C = ArtTile<F>;
    const long fp = artcnn_feature_pitch(F, W);
    float4* feat = reinterpret_cast<float4*>(ws);
    float4* m0 = reinterpret_cast<float4*>(ws + fp * H);
    float4* m1 = reinterpret_cast<float4*>(ws + 2 * fp * H);
    hipLaunchKernelGGL((artcnn_head<TIN, F>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H, feat,
                       fp, kernels, biases);
    const dim3 grid((W + C::TW - 1) / C::TW, (H + C::TH - 1) / C::TH), block(C::NT);
    const long wl = 9L * 64 * C::KQ * C::MT;   // A-operand floats per F -> F layer
    const float4* in = feat;
    float4* out = m0;
    for (int l = 1; l <= blocks; l++) {         // body: F -> F + ReLU
        hipLaunchKernelGGL((artcnn_conv<F, kArtRelu, float>), grid, block, 0, stream, in, fp, W, H, out, feat,
                           wa + (l - 1) * wl, biases + F * l, nullptr, 0L);
        in = out;
        out = out == m0 ? m1 : m0;
    }
    hipLaunchKernelGGL((artcnn_conv<F, kArtAdd, float>), grid, block, 0, stream, in, fp, W, H, out, feat,
                       wa + blocks * wl, biases + F * (blocks + 1), nullptr, 0L);
    hipLaunchKernelGGL((artcnn_conv<F, kArtTail, TOUT>), grid, block, 0, stream, out, fp, W, H, nullptr, feat,
                       wa + (blocks + 1) * wl, biases + F * (blocks + 2), dst, ds);
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
