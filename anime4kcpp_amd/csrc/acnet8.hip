// ACNet<8> 2x pass (core/src/processor/cpu/CPUProcessor.cpp:1417-1436) on gfx950, layer by
// layer, FP32 throughout:
//   L0     conv3x3 1 -> 8 + PReLU            (Backend.hpp:58-75)        VALU; with an even block
//          count per tile in LDS together with block 1 (conv8_head), else one pixel per thread
//   L1..LB conv3x3 8 -> 8 + PReLU            (Backend.hpp:95-98)        conv8_pair: two per launch
//   final  conv3x3 8 -> 4, pixel shuffle 2x, + the source pixel  (Backend.hpp:116-133,
//          Common.hpp:290-341)                  with the last block: conv8_pair<kShuffleSrc>
// Features between launches are 8-channel f32 maps in HBM (32 bytes per pixel, rows padded
// to 256 bytes); the blocks run two per launch (conv8_pair), so each map round trip covers
// two layers at the cost of a one-pixel recompute ring; the depth (4, 8 or 18 blocks) is a
// loop on the host.  Borders replicate the edge pixel at every layer,
// as the reference's clamped taps (Common.hpp:120-160) do.
#include <utility>

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "conv_layers.hpp"

namespace achip {

#pragma clang fp contract(off)

// L0 of one pixel from its 9 clamped taps (row-major): conv3x3 1 -> 8 + PReLU with the blob's
// slopes, each product and sum rounded in the reference's order
__device__ inline void l0_pixel8(const float (&r)[9], const float* __restrict__ wts, float (&o)[8])
{
#pragma unroll
    for (int n = 0; n < 8; n++) {
        float s = 0.0f;
#pragma unroll
        for (int p = 0; p < 9; p++) s = __fadd_rn(s, __fmul_rn(r[p], wts[kA8OffW0 + p * 8 + n]));
        o[n] = prelu(__fadd_rn(s, wts[kA8OffB0 + n]), wts[kA8OffAlpha0 + n]);
    }
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
    l0_pixel8(r, wts, o);
    float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(feat) + static_cast<long>(y) * fp) + 2 * x;
    px[0] = make_float4(o[0], o[1], o[2], o[3]);
    px[1] = make_float4(o[4], o[5], o[6], o[7]);
}

// ---- L0 and the first block in one launch (62 x 16 outputs, kP8NT threads) ----
// The source tile (66 x 20 pixels from (x0 - 2, y0 - 2), clamped, as f32) gives L0 at plane
// pixels (1..64, 1..18) with l0_pixel8's arithmetic; the first block then runs as conv8_pair's
// second layer, so the map between them never goes through HBM.
constexpr int kH8SW = kP8TW + 4, kH8SH = kP8TH + 4;

template <typename TIN>
__global__ __launch_bounds__(kP8NT) void conv8_head(const void* __restrict__ src, long ss, int W, int H,
                                                  float4* __restrict__ fout, long fp, const float* __restrict__ wts,
                                                  const LayerArgs la)
{
    __shared__ float4 pa[kP8PW * kP8PH];
    __shared__ float4 pb[kP8PW * kP8PH];
    __shared__ float img[kH8SH * kH8SW];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kP8TW, y0 = ty * kP8TH;
    const int tid = threadIdx.x, lane = tid & 63;
    const int half = (lane >> 4) & 1;
    // img pixel (c, r) is image pixel (x0 - 2 + c, y0 - 2 + r), as plane pixel (c, r)
    for (int e = tid; e < kH8SH * kH8SW; e += kP8NT) {
        const int r = e / kH8SW, c = e - r * kH8SW;
        img[e] = to_float<TIN>(row_ptr<TIN>(src, ss, clampi(y0 - 2 + r, 0, H - 1))[clampi(x0 - 2 + c, 0, W - 1)]);
    }
    float a[24];
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = la.wa[s * 64 + lane];
    Epi8<kPrelu, float, float> epi;
    epi.load(la, lane, half);
    __syncthreads();
    for (int e = tid; e < (kP8TH + 2) * 64; e += kP8NT) {
        const int r = 1 + e / 64, c = 1 + (e & 63);
        float q[9];
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int dx = 0; dx < 3; dx++) q[dy * 3 + dx] = img[(r - 1 + dy) * kH8SW + c - 1 + dx];
        float o[8];
        l0_pixel8(q, wts, o);
        pa[r * kP8PW + c] = make_float4(o[0], o[1], o[2], o[3]);
        pb[r * kP8PW + c] = make_float4(o[4], o[5], o[6], o[7]);
    }
    __syncthreads();
    pair8_second<kPrelu, float, float>(pa, pb, W, H, x0, y0, a, epi, fp, fout, la);
}

// ---- launch ----

template <typename TIN, typename TOUT>
static hipError_t acnet8_pass_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                const float* wts, const float* wa, float4* f0, float4* f1, long fp, hipStream_t stream)
{
    const dim3 pgrid = conv8_pair_grid(W, H), block(kP8NT);
    auto block_args = [&](int l) {
        LayerArgs la{};
        la.wa = wa + static_cast<long>(l - 1) * kMfma32LayerFloats;
        la.bias = wts + kA8OffLayer + (l - 1) * 16;   // 8 biases, then 8 slopes
        la.alpha = la.bias + 8;
        return la;
    };
    float4* in = f0;
    float4* out = f1;
    int l = 1;
    if (blocks % 2 == 0) {   // an even count: L0 with block 1 (conv8_head)
        hipLaunchKernelGGL((conv8_head<TIN>), pgrid, block, 0, stream, src, ss, W, H, in, fp, wts, block_args(1));
        l = 2;
    } else {
        hipLaunchKernelGGL((acnet8_l0<TIN>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H, in, fp,
                           wts);
    }
    // an odd number of blocks is left: two per launch, the last one with the final layer
    for (; l < blocks; l += 2) {
        hipLaunchKernelGGL((conv8_pair<kPrelu, float, float>), pgrid, block, 0, stream, in, fp, W, H, out, block_args(l),
                           block_args(l + 1));
        std::swap(in, out);
    }
    LayerArgs la{};
    la.wa = wa + static_cast<long>(blocks) * kMfma32LayerFloats;
    la.bias = wts + kA8OffFinalB;
    la.src = src;
    la.ss = ss;
    la.dst = dst;
    la.ds = ds;
    hipLaunchKernelGGL((conv8_pair<kShuffleSrc, TIN, TOUT>), pgrid, block, 0, stream, in, fp, W, H, out, block_args(blocks), la);
    return hipGetLastError();
}

long acnet8_feature_pitch(int w) { return (static_cast<long>(w) * 32 + 255) / 256 * 256; }

// ---- ACNetLegacy, layer by layer (the layer-split form of the fused FP32 pass) ----
// Layer 0 and layers 1-2 in one launch (legacy_head), then layers 3..8 two per launch on
// conv8_pair (ReLU; the last pair ends in the deconv 2x2 8 -> 1 straight to the output):
// four launches, each 8-channel f32 map written and read once through HBM, a one-pixel
// recompute ring per layer instead of the fused kernel's eight-pixel halo.  Every layer
// runs the fused kernel's own instructions on the same operands in the same order (layer 0:
// three MFMAs per item from the bias; layers 1-8: item24; the deconv epilogue), so this
// form gives the fused pass's bits.
//
// legacy_head, 62 x 16 outputs per workgroup of kP8NT threads: the source tile, 82 x 22 pixels from
// (x0 - 3, y0 - 3), clamped, in the source type, gives layer 0 at plane pixels (0..79, 0..19)
// (five 16-pixel segments x 10 row pairs; plane pixel (c, r) is image pixel
// (x0 - 2 + c, y0 - 2 + r)); plane pixels outside the image then take the nearest pixel
// inside, as every layer's clamped taps see them (Common.hpp:120-143), and layers 1 and 2 run
// as conv8_pair's two layers.
constexpr int kLHW = 16 * 5 + 2, kLHH = kP8PH + 2;

template <typename TIN>
__device__ __forceinline__ void legacy_head_tile(const void* __restrict__ src, long ss, int W, int H,
                                                 float4* __restrict__ fout, long fp, const float* __restrict__ wts,
                                                 const float* __restrict__ wa32, const LayerArgs& la0, const LayerArgs& la,
                                                 int x0, int y0, float4* __restrict__ pa, float4* __restrict__ pb,
                                                 TIN* __restrict__ img)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    // img pixel (c, r) is image pixel (x0 - 3 + c, y0 - 3 + r)
    for (int e = tid; e < kLHH * kLHW; e += kP8NT) {
        const int r = e / kLHW, c = e - r * kLHW;
        img[e] = row_ptr<TIN>(src, ss, clampi(y0 - 3 + r, 0, H - 1))[clampi(x0 - 3 + c, 0, W - 1)];
    }
    const float a0 = wa32[lane], a1 = wa32[64 + lane], a2 = wa32[128 + lane];
    const float* bl = wts + kOffB0 + 4 * half;
    const f32x4m bias0 = { bl[0], bl[1], bl[2], bl[3] };
    float a[24];
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = la0.wa[s * 64 + lane];
    Epi8<kRelu, float, float> epa;
    epa.load(la0, lane, half);
    __syncthreads();
    // layer 0: item (pair, seg) writes plane pixels (16 seg + j, 2 pair + ro); its B operands
    // are img row 2 pair + g (image row y0 - 3 + 2 pair + g), columns 16 seg + j + t
    {
        float4* const pw = half ? pb : pa;
#pragma unroll
        for (int i0 = 0; i0 < (50 + kP8NW - 1) / kP8NW; i0++) {
            const int item = wave + kP8NW * i0, pair = item / 5, seg = item - 5 * pair;
            if (item >= 50) break;   // uniform across the wave
            const TIN* p = img + (2 * pair + g) * kLHW + 16 * seg + j;
            f32x4m acc = mma32(a0, to_float<TIN>(p[0]), bias0);
            acc = mma32(a1, to_float<TIN>(p[1]), acc);
            acc = mma32(a2, to_float<TIN>(p[2]), acc);
            pw[(2 * pair + ro) * kP8PW + 16 * seg + j] =
                make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f), fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
        }
    }
    __syncthreads();
    // layer 1 reads plane pixels (0..65, 0..19): those outside the image take the nearest inside
    if (x0 - 2 < 0 || y0 - 2 < 0 || x0 + 63 >= W || y0 + 17 >= H) {
        for (int e = tid; e < kP8PH * 66; e += kP8NT) {
            const int r = e / 66, c = e - r * 66;
            const int gy = y0 - 2 + r, gx = x0 - 2 + c;
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) continue;
            const int s = (clampi(gy, 0, H - 1) - (y0 - 2)) * kP8PW + clampi(gx, 0, W - 1) - (x0 - 2);
            pa[r * kP8PW + c] = pa[s];
            pb[r * kP8PW + c] = pb[s];
        }
        __syncthreads();
    }
    Epi8<kRelu, float, float> epi;
    pair8_first<kRelu>(pa, pb, a, epa, epi, la);
    pair8_second<kRelu, float, float>(pa, pb, W, H, x0, y0, a, epi, fp, fout, la);
}


template <typename TIN>
__global__ __launch_bounds__(kP8NT) void legacy_head(const void* __restrict__ src, long ss, int W, int H,
                                                   float4* __restrict__ fout, long fp, const float* __restrict__ wts,
                                                   const float* __restrict__ wa32, const LayerArgs la0, const LayerArgs la)
{
    __shared__ float4 pa[kP8PW * kP8PH];
    __shared__ float4 pb[kP8PW * kP8PH];
    __shared__ TIN img[kLHH * kLHW];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    legacy_head_tile<TIN>(src, ss, W, H, fout, fp, wts, wa32, la0, la, tx * kP8TW, ty * kP8TH, pa, pb, img);
}

// ---- the layer-split form over a batch of images: every launch takes the tiles of all of
// them (numbered image by image, XCD-aware), each image with its own two maps ----
template <typename TIN>
__global__ __launch_bounds__(kP8NT) void legacy_head_batch(const SplitDesc* __restrict__ desc, int n, int total,
                                                         const float* __restrict__ wts, const float* __restrict__ wa32,
                                                         const LayerArgs la0, const LayerArgs la)
{
    __shared__ float4 pa[kP8PW * kP8PH];
    __shared__ float4 pb[kP8PW * kP8PH];
    __shared__ TIN img[kLHH * kLHW];
    const int t = xcd_tile(blockIdx.x, total);
    const SplitDesc d = find_tile(desc, n, t);
    const int local = t - d.first_tile, ty = local / d.tiles_x, tx = local - ty * d.tiles_x;
    legacy_head_tile<TIN>(d.src, d.sstride, d.w, d.h, d.m0, d.fp, wts, wa32, la0, la, tx * kP8TW, ty * kP8TH, pa, pb, img);
}

// from_m1: the launch reads map 1 and writes map 0 (else the other way); kReluDeconv writes
// the image's output instead of a map
template <int MODE, typename TOUT>
__global__ __launch_bounds__(kP8NT) void pair8_batch(const SplitDesc* __restrict__ desc, int n, int total, int from_m1,
                                                   const LayerArgs la0, const LayerArgs la)
{
    __shared__ float4 pa[kP8PW * kP8PH];
    __shared__ float4 pb[kP8PW * kP8PH];
    const int t = xcd_tile(blockIdx.x, total);
    const SplitDesc d = find_tile(desc, n, t);
    const int local = t - d.first_tile, ty = local / d.tiles_x, tx = local - ty * d.tiles_x;
    LayerArgs lb = la;
    if constexpr (MODE == kReluDeconv) {
        lb.dst = d.dst;
        lb.ds = d.dstride;
    }
    pair8_tile<MODE, float, TOUT, kRelu>(from_m1 ? d.m1 : d.m0, d.fp, d.w, d.h, from_m1 ? d.m0 : d.m1, la0, lb,
                                         tx * kP8TW, ty * kP8TH, pa, pb);
}

template <typename TIN, typename TOUT>
static hipError_t legacy_split_t(const void* src, long ss, int W, int H, void* dst, long ds, const float* wts,
                                 const float* wa32, float4* f0, float4* f1, long fp, hipStream_t stream)
{
    const dim3 pgrid = conv8_pair_grid(W, H), block(kP8NT);
    auto layer = [&](int l) {
        LayerArgs la{};
        la.wa = wa32 + kMfma32L0Floats + static_cast<long>(l - 1) * kMfma32LayerFloats;
        la.bias = wts + offB(l);
        return la;
    };
    hipLaunchKernelGGL((legacy_head<TIN>), pgrid, block, 0, stream, src, ss, W, H, f0, fp, wts, wa32, layer(1), layer(2));
    float4* in = f0;
    float4* out = f1;
    for (int l = 3; l <= 5; l += 2) {
        hipLaunchKernelGGL((conv8_pair<kRelu, float, float, kRelu>), pgrid, block, 0, stream, in, fp, W, H, out, layer(l),
                           layer(l + 1));
        std::swap(in, out);
    }
    LayerArgs last = layer(8);
    last.k2 = wts + kOffK2;
    last.dst = dst;
    last.ds = ds;
    hipLaunchKernelGGL((conv8_pair<kReluDeconv, float, TOUT, kRelu>), pgrid, block, 0, stream, in, fp, W, H, out, layer(7), last);
    return hipGetLastError();
}

hipError_t launch_legacy_split(int stype, int dtype, const void* src, long ss, int W, int H, void* dst, long ds,
                               const float* wts, const float* wa32, void* workspace, hipStream_t stream)
{
    const long fp = acnet8_feature_pitch(W);
    float4* f0 = static_cast<float4*>(workspace);
    float4* f1 = reinterpret_cast<float4*>(static_cast<char*>(workspace) + fp * H);
#define CALL(TI, TO) legacy_split_t<TI, TO>(src, ss, W, H, dst, ds, wts, wa32, f0, f1, fp, stream)
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

template <typename TIN, typename TOUT>
static hipError_t legacy_split_batch_t(const SplitDesc* desc, int n, int total, const float* wts, const float* wa32,
                                       hipStream_t stream)
{
    const dim3 grid(total), block(kP8NT);
    auto layer = [&](int l) {
        LayerArgs la{};
        la.wa = wa32 + kMfma32L0Floats + static_cast<long>(l - 1) * kMfma32LayerFloats;
        la.bias = wts + offB(l);
        return la;
    };
    hipLaunchKernelGGL((legacy_head_batch<TIN>), grid, block, 0, stream, desc, n, total, wts, wa32, layer(1), layer(2));
    hipLaunchKernelGGL((pair8_batch<kRelu, float>), grid, block, 0, stream, desc, n, total, 0, layer(3), layer(4));
    hipLaunchKernelGGL((pair8_batch<kRelu, float>), grid, block, 0, stream, desc, n, total, 1, layer(5), layer(6));
    LayerArgs last = layer(8);
    last.k2 = wts + kOffK2;
    hipLaunchKernelGGL((pair8_batch<kReluDeconv, TOUT>), grid, block, 0, stream, desc, n, total, 0, layer(7), last);
    return hipGetLastError();
}

hipError_t launch_legacy_split_batch(int stype, int dtype, const SplitDesc* desc, int n, int total, const float* wts,
                                     const float* wa32, hipStream_t stream)
{
    if (n <= 0 || total <= 0) return hipSuccess;
#define CALL(TI, TO) legacy_split_batch_t<TI, TO>(desc, n, total, wts, wa32, stream)
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
