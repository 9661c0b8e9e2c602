// ACNet<8> 2x pass (core/src/processor/cpu/CPUProcessor.cpp:1417-1436) with the conv3x3 layers
// as an FP16 or BF16 implicit GEMM on gfx950 MFMA (v_mfma_f32_16x16x32_{f16,bf16}), FP32
// accumulation, and 8-channel 16-bit feature maps between launches (16 bytes per pixel: half
// the HBM traffic of the FP32 path's maps):
//   L0     conv3x3 1 -> 8 + PReLU in FP32 (the FP32 path's arithmetic), rounded to 16 bits;
//          with an even block count computed per tile in LDS together with block 1
//          (conv16_head), else written as a map (acnet8_l0_16)
//   L1..LB conv3x3 8 -> 8, bias and PReLU in FP32, rounded to 16 bits; two layers per launch
//          with the map between them in LDS (conv16_pair); the last block runs with the final
//          layer in one conv16_pair, or alone before it (conv16_layer) when the pairs use all
//   final  conv3x3 8 -> 4, pixel shuffle 2x, + the source pixel, all in FP32 from the FP32
//          accumulators, then fromFloat to the output type
// The MFMA work item is the ACNetLegacy 16-bit kernel's (acnet_mfma.hip): 16 consecutive pixels
// x 2 output rows from a 4 x 3 input patch, K = 96 over three MFMAs (one per patch column),
// lane group g = lane/16 holding patch row g, so each lane's B operand is the 8 channels of one
// pixel (one 16-byte LDS read); lane l's D registers are 4 channels (4 (g & 1) ..) of pixel
// l % 16 in output row g >> 1.  A operands per layer: [t][lane] x 8 values (the final layer's
// output rows 4-7 zero).  Borders replicate the edge pixel at every layer, as the reference's
// clamped taps (Common.hpp:120-160) do.
#include <utility>

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "conv_layers.hpp"
#include "mfma16.hpp"

namespace achip {

#pragma clang fp contract(off)

long acnet8_map16_pitch(int w) { return (static_cast<long>(w) * 16 + 255) / 256 * 256; }

// L0 of one pixel from its 9 clamped taps (row-major), in FP32, rounded to the 16-bit map
template <typename T16>
__device__ inline uint4 l0_pixel16(const float (&r)[9], const float* __restrict__ wts)
{
    float o[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        float s = 0.0f;
#pragma unroll
        for (int p = 0; p < 9; p++) s = __fadd_rn(s, __fmul_rn(r[p], wts[kA8OffW0 + p * 8 + n]));
        o[n] = prelu(__fadd_rn(s, wts[kA8OffB0 + n]), wts[kA8OffAlpha0 + n]);
    }
    return make_uint4(pack16<T16>(o[0], o[1]), pack16<T16>(o[2], o[3]), pack16<T16>(o[4], o[5]), pack16<T16>(o[6], o[7]));
}

// ---- L0: one pixel per thread, 64 x 4 pixels per workgroup ----
template <typename T16, typename TIN>
__global__ __launch_bounds__(256) void acnet8_l0_16(const void* __restrict__ src, long ss, int W, int H,
                                                    uint4* __restrict__ feat, long fp, const float* __restrict__ wts)
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
    uint4* px = reinterpret_cast<uint4*>(reinterpret_cast<char*>(feat) + static_cast<long>(y) * fp) + x;
    *px = l0_pixel16<T16>(r, wts);
}

// Arguments of one 16-bit layer: A operands (uint4 per lane and patch column), FP32 bias and
// PReLU slopes from the ACNet<8> blob; for the final layer the source and output images.
struct Layer16 {
    const uint4* wa;
    const float* bias;
    const float* alpha;
    const void* src;
    long ss;
    void* dst;
    long ds;
};

// This lane's bias and slopes (channels 4 half .. + 3), and its three A operands.
template <typename T16>
struct Ops16 {
    typename Mma<T16>::v8 a[3];
    f32x4 bias;
    float alpha[4];
    __device__ void load(const Layer16& la, int lane, int half, bool final_layer)
    {
#pragma unroll
        for (int t = 0; t < 3; t++) a[t] = __builtin_bit_cast(typename Mma<T16>::v8, la.wa[t * 64 + lane]);
        if (final_layer) {   // four output channels, all in the half = 0 lanes
            bias = half ? (f32x4){ 0.f, 0.f, 0.f, 0.f } : (f32x4){ la.bias[0], la.bias[1], la.bias[2], la.bias[3] };
        } else {
            const float* bl = la.bias + 4 * half;
            bias = (f32x4){ bl[0], bl[1], bl[2], bl[3] };
#pragma unroll
            for (int r = 0; r < 4; r++) alpha[r] = la.alpha[4 * half + r];
        }
    }
    // the 3 MFMAs of one item: B operands at p[0], p[1], p[2] (patch columns)
    __device__ f32x4 item(const uint4* __restrict__ p) const
    {
        using M = Mma<T16>;
        using v8 = typename M::v8;
        f32x4 acc = M::mma(a[0], __builtin_bit_cast(v8, p[0]), bias);
        acc = M::mma(a[1], __builtin_bit_cast(v8, p[1]), acc);
        return M::mma(a[2], __builtin_bit_cast(v8, p[2]), acc);
    }
    __device__ uint2 prelu16(const f32x4& acc) const
    {
        return make_uint2(pack16<T16>(prelu(acc[0], alpha[0]), prelu(acc[1], alpha[1])),
                          pack16<T16>(prelu(acc[2], alpha[2]), prelu(acc[3], alpha[3])));
    }
};

// The output side of a layer: a 16-bit map (PReLU) or, for the final layer, the shuffled 2x
// image plus the source pixel.  Every lane of the wave calls this (the shuffle reads partners).
template <typename T16, bool FINAL, typename TIN, typename TOUT>
__device__ inline void store16(const Ops16<T16>& op, const f32x4& acc, int x, int y, bool ok, int half, uint4* fout,
                               long fp, const Layer16& la)
{
    if constexpr (!FINAL) {
        if (ok) {
            uint2* px = reinterpret_cast<uint2*>(reinterpret_cast<char*>(fout) + static_cast<long>(y) * fp) + 2 * x + half;
            *px = op.prelu16(acc);
        }
    } else {
        // half = 0 lanes hold the four sub-pixels n = 0..3 (output (2x + n % 2, 2y + n / 2));
        // half = 1 lanes take row 2y + 1 from them
        const float n2 = __shfl_xor(acc[2], 16), n3 = __shfl_xor(acc[3], 16);
        if (ok) {
            const float res = to_float<TIN>(row_ptr<TIN>(la.src, la.ss, y)[x]);
            const float s0 = add_skip(half ? n2 : acc[0], res), s1 = add_skip(half ? n3 : acc[1], res);
            store_pair<TOUT>(row_ptr<TOUT>(la.dst, la.ds, 2 * y + half) + 2 * x, s0, s1);
        }
    }
}

// Stage ROWS x CW pixels of a 16-bit map whose top-left is image pixel (cx0, ry0), clamped to
// the image, into LDS plane[r * PW + c]; all of a thread's loads in flight before its stores.
template <int NT, int CW, int ROWS, int PW>
__device__ __forceinline__ void stage16(const uint4* __restrict__ fin, long fp, int W, int H, int cx0, int ry0,
                                        uint4* __restrict__ plane)
{
    constexpr int N = ROWS * CW, K = (N + NT - 1) / NT;
    uint4 v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = min(static_cast<int>(threadIdx.x) + k * NT, N - 1);
        const int r = e / CW, c = e - r * CW;
        const int sy = clampi(ry0 + r, 0, H - 1), sx = clampi(cx0 + c, 0, W - 1);
        v[k] = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(fin) + static_cast<long>(sy) * fp)[sx];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = threadIdx.x + k * NT;
        if (k < K - 1 || e < N) {
            const int r = e / CW, c = e - r * CW;
            plane[r * PW + c] = v[k];
        }
    }
}

// ---- one layer: 64 x 16 outputs per workgroup of 256 threads ----
constexpr int kL16TW = 64, kL16TH = 16, kL16PW = 80, kL16PH = kL16TH + 2;

template <typename T16, bool FINAL, typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void conv16_layer(const uint4* __restrict__ fin, long fp, int W, int H,
                                                    uint4* __restrict__ fout, const Layer16 la)
{
    __shared__ uint4 plane[kL16PW * kL16PH];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kL16TW, y0 = ty * kL16TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    // plane pixel (c, r) is image pixel (x0 - 1 + c, y0 - 1 + r)
    stage16<256, kL16TW + 2, kL16PH, kL16PW>(fin, fp, W, H, x0 - 1, y0 - 1, plane);
    Ops16<T16> op;
    op.load(la, lane, half, FINAL);
    __syncthreads();
    const int gx1 = min(x0 + kL16TW, W), gy1 = min(y0 + kL16TH, H);
    const int nseg = (gx1 - x0 + 15) >> 4;
    const int nitems = nseg * ((gy1 - y0 + 1) >> 1);
    const int lane_rd = g * kL16PW + j;
    for (int item = wave; item < nitems; item += 4) {
        const int pair = item / nseg, seg = item - pair * nseg;
        const f32x4 acc = op.item(plane + 2 * pair * kL16PW + 16 * seg + lane_rd);
        const int x = x0 + 16 * seg + j, y = y0 + 2 * pair + ro;
        store16<T16, FINAL, TIN, TOUT>(op, acc, x, y, x < gx1 && y < gy1, half, fout, fp, la);
    }
}

// ---- two layers: layer A (PReLU) into LDS, then layer B; 62 x 32 outputs, 512 threads ----
// Layer A covers the tile grown by one pixel per side, 64 x 34 = 4 segments x 17 row pairs;
// the staged input is 66 x 36 at an 80-pixel pitch (46 KB, three workgroups per CU).  After
// layer A, tiles at the image edge overwrite A's pixels outside the image with the nearest
// pixel inside, which is the replicate padding of the next layer.
constexpr int kP16TW = 62, kP16TH = 32, kP16PW = 80, kP16PH = kP16TH + 4, kP16NT = 512;

// The second layer of a two-layer launch, once layer A fills plane pixels (1..64, 1..34):
// the replicate ring, then layer B (operands in op) over the 62 x 32 tile to fout or the
// output image.
template <typename T16, bool FINAL, typename TIN, typename TOUT>
__device__ __forceinline__ void pair_second(uint4* __restrict__ plane, int W, int H, int x0, int y0, const Ops16<T16>& op,
                                            uint4* __restrict__ fout, long fp, const Layer16& la)
{
    constexpr int NW = kP16NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    const int lane_rd = g * kP16PW + j;
    // replicate ring: A's pixels outside the image take the nearest pixel inside
    if (x0 - 1 < 0 || y0 - 1 < 0 || x0 + kP16TW >= W || y0 + kP16TH >= H) {
        for (int e = tid; e < (kP16TH + 2) * 64; e += kP16NT) {
            const int r = 1 + e / 64, c = 1 + (e & 63);
            const int gy = y0 - 2 + r, gx = x0 - 2 + c;
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) continue;
            plane[r * kP16PW + c] = plane[(clampi(gy, 0, H - 1) - (y0 - 2)) * kP16PW + clampi(gx, 0, W - 1) - (x0 - 2)];
        }
        __syncthreads();
    }
    // layer B over the 62 x 32 tile: patch rows 2 pair + 1 + g, columns 16 seg + j + t + 1
    // (columns past A's region feed only lanes past the tile)
    const int gx1 = min(x0 + kP16TW, W), gy1 = min(y0 + kP16TH, H);
    for (int item = wave; item < 4 * (kP16TH / 2); item += NW) {
        const int pair = item >> 2, seg = item & 3;
        const f32x4 acc = op.item(plane + (2 * pair + 1) * kP16PW + 16 * seg + 1 + lane_rd);
        const int x = x0 + 16 * seg + j, y = y0 + 2 * pair + ro;
        store16<T16, FINAL, TIN, TOUT>(op, acc, x, y, x < gx1 && y < gy1, half, fout, fp, la);
    }
}

template <typename T16, bool FINAL, typename TIN, typename TOUT>
__global__ __launch_bounds__(kP16NT) void conv16_pair(const uint4* __restrict__ fin, long fp, int W, int H,
                                                      uint4* __restrict__ fout, const Layer16 la0, const Layer16 la)
{
    constexpr int NW = kP16NT / 64, ITEMS_A = 4 * (kP16TH / 2 + 1), KA = (ITEMS_A + NW - 1) / NW;
    __shared__ uint4 plane[kP16PW * kP16PH];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kP16TW, y0 = ty * kP16TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    // plane pixel (c, r) is image pixel (x0 - 2 + c, y0 - 2 + r)
    stage16<kP16NT, kP16TW + 4, kP16PH, kP16PW>(fin, fp, W, H, x0 - 2, y0 - 2, plane);
    Ops16<T16> op;
    op.load(la0, lane, half, false);
    __syncthreads();

    // layer A over plane pixels (1..64, 1..34): item (pair, seg) reads patch rows 2 pair + g,
    // columns 16 seg + j + t
    const int lane_rd = g * kP16PW + j;
    uint2 res[KA];
#pragma unroll
    for (int ii = 0; ii < KA; ii++) {
        const int item = wave + NW * ii;
        if (item < ITEMS_A) {
            const int pair = item >> 2, seg = item & 3;
            res[ii] = op.prelu16(op.item(plane + 2 * pair * kP16PW + 16 * seg + lane_rd));
        }
    }
    op.load(la, lane, half, FINAL);   // layer B's operands while the other waves finish
    __syncthreads();
    {
        uint2* const pw = reinterpret_cast<uint2*>(plane);
#pragma unroll
        for (int ii = 0; ii < KA; ii++) {
            const int item = wave + NW * ii;
            if (item < ITEMS_A) {
                const int pair = item >> 2, seg = item & 3;
                pw[((1 + 2 * pair + ro) * kP16PW + 1 + 16 * seg + j) * 2 + half] = res[ii];
            }
        }
    }
    __syncthreads();
    pair_second<T16, FINAL, TIN, TOUT>(plane, W, H, x0, y0, op, fout, fp, la);
}


// ---- L0 and the first block in one launch (62 x 32 outputs, 512 threads) ----
// The source tile (66 x 36 pixels from (x0 - 2, y0 - 2), clamped to the image, as f32) gives
// L0 at plane pixels (1..64, 1..34) with l0_pixel16's arithmetic, then the first block runs as
// conv16_pair's second layer: the map acnet8_l0_16 writes and conv16_pair reads never goes
// through HBM.
constexpr int kH16SW = kP16TW + 4, kH16SH = kP16TH + 4;

template <typename T16, typename TIN>
__global__ __launch_bounds__(kP16NT) void conv16_head(const void* __restrict__ src, long ss, int W, int H,
                                                      uint4* __restrict__ fout, long fp, const float* __restrict__ wts,
                                                      const Layer16 la)
{
    __shared__ uint4 plane[kP16PW * kP16PH];
    __shared__ float img[kH16SH * kH16SW];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kP16TW, y0 = ty * kP16TH;
    const int tid = threadIdx.x, lane = tid & 63;
    const int half = (lane >> 4) & 1;
    // img pixel (c, r) is image pixel (x0 - 2 + c, y0 - 2 + r), as plane pixel (c, r)
    for (int e = tid; e < kH16SH * kH16SW; e += kP16NT) {
        const int r = e / kH16SW, c = e - r * kH16SW;
        const int sy = clampi(y0 - 2 + r, 0, H - 1), sx = clampi(x0 - 2 + c, 0, W - 1);
        img[e] = to_float<TIN>(row_ptr<TIN>(src, ss, sy)[sx]);
    }
    Ops16<T16> op;
    op.load(la, lane, half, false);
    __syncthreads();
    for (int e = tid; e < (kP16TH + 2) * (kP16TW + 2); e += kP16NT) {
        const int r = 1 + e / (kP16TW + 2), c = 1 + e % (kP16TW + 2);
        float q[9];
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int dx = 0; dx < 3; dx++) q[dy * 3 + dx] = img[(r - 1 + dy) * kH16SW + c - 1 + dx];
        plane[r * kP16PW + c] = l0_pixel16<T16>(q, wts);
    }
    __syncthreads();
    pair_second<T16, false, float, float>(plane, W, H, x0, y0, op, fout, fp, la);
}

// ---- launch ----

template <typename T16, typename TIN, typename TOUT>
static hipError_t acnet8_pass16_t(int blocks, const void* src, long ss, int W, int H, void* dst, long ds,
                                  const float* wts, const uint4* wa, uint4* f0, uint4* f1, long fp, hipStream_t stream)
{
    const dim3 grid((W + kL16TW - 1) / kL16TW, (H + kL16TH - 1) / kL16TH);
    const dim3 pgrid((W + kP16TW - 1) / kP16TW, (H + kP16TH - 1) / kP16TH);
    constexpr long kLayerVecs = 3 * 64;   // uint4 A operands per layer
    auto block_args = [&](int l) {
        Layer16 la{};
        la.wa = wa + (l - 1) * kLayerVecs;
        la.bias = wts + kA8OffLayer + (l - 1) * 16;   // 8 biases, then 8 slopes
        la.alpha = la.bias + 8;
        return la;
    };
    Layer16 fin{};
    fin.wa = wa + static_cast<long>(blocks) * kLayerVecs;
    fin.bias = wts + kA8OffFinalB;
    fin.src = src;
    fin.ss = ss;
    fin.dst = dst;
    fin.ds = ds;
    uint4* in = f0;
    uint4* out = f1;
    int l = 1;
    if (blocks % 2 == 0) {
        // an even count: L0 with block 1, then blocks two per launch, the last with the final
        // layer
        hipLaunchKernelGGL((conv16_head<T16, TIN>), pgrid, dim3(kP16NT), 0, stream, src, ss, W, H, in, fp, wts, block_args(1));
        l = 2;
    } else {
        hipLaunchKernelGGL((acnet8_l0_16<T16, TIN>), dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, stream, src, ss, W, H,
                           in, fp, wts);
    }
    // blocks two per launch; the last one left over pairs with the final layer
    for (; l + 1 <= blocks; l += 2) {
        hipLaunchKernelGGL((conv16_pair<T16, false, float, float>), pgrid, dim3(kP16NT), 0, stream, in, fp, W, H, out,
                           block_args(l), block_args(l + 1));
        std::swap(in, out);
    }
    if (l == blocks)
        hipLaunchKernelGGL((conv16_pair<T16, true, TIN, TOUT>), pgrid, dim3(kP16NT), 0, stream, in, fp, W, H, out,
                           block_args(l), fin);
    else
        hipLaunchKernelGGL((conv16_layer<T16, true, TIN, TOUT>), grid, dim3(256), 0, stream, in, fp, W, H, out, fin);
    return hipGetLastError();
}

template <typename T16>
static hipError_t acnet8_pass16_types(int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                                      long ds, const float* wts, const uint4* wa, uint4* f0, uint4* f1, long fp,
                                      hipStream_t stream)
{
#define CALL(TI, TO) acnet8_pass16_t<T16, TI, TO>(blocks, src, ss, W, H, dst, ds, wts, wa, f0, f1, fp, stream)
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

hipError_t launch_acnet8_16(int bf16, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                            long ds, const float* wts, const void* wa16, void* workspace, hipStream_t stream)
{
    const long fp = acnet8_map16_pitch(W);
    uint4* f0 = static_cast<uint4*>(workspace);
    uint4* f1 = reinterpret_cast<uint4*>(static_cast<char*>(workspace) + fp * H);
    const uint4* wa = static_cast<const uint4*>(wa16);
    if (bf16) return acnet8_pass16_types<__bf16>(blocks, stype, dtype, src, ss, W, H, dst, ds, wts, wa, f0, f1, fp, stream);
    return acnet8_pass16_types<_Float16>(blocks, stype, dtype, src, ss, W, H, dst, ds, wts, wa, f0, f1, fp, stream);
}

} // namespace achip
