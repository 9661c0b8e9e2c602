// Fused ACNetLegacy 2x pass with the 8->8 layers as an FP16 or BF16 implicit GEMM on
// gfx950 MFMA (v_mfma_f32_16x16x32_{f16,bf16}), FP32 accumulation.  Layer 0 (1 -> 8) is an
// FP32 MFMA (v_mfma_f32_16x16x4_f32) on the f32 input tile; bias, ReLU and the deconv are
// FP32.  Same tiling, halo and per-layer border clamps as the fused FP32 kernel
// (acnet_fp32.hip); the difference is how the layers are computed.
//
// The GEMM of one work item (16 pixels x 2 rows):
//   D[16 x 16] = A[16 x 96] * B[96 x 16]
//   rows i of D/A : i = 8*ro + n, output channel n of output row y+ro (ro = 0, 1);
//   columns j     : the item's 16 pixels;
//   K = 96        : a 4 x 3 input patch (rows y-1..y+2, columns x-1..x+1) x 8 channels.
// Both output rows are linear in the same 4 x 3 patch, so all 16 MFMA rows are useful;
// 72 of each row's 96 weights are nonzero.  K is split over three MFMAs, one per patch
// column t; inside one MFMA lane group g = lane/16 holds patch row g, so each lane's B
// operand is the 8 channels of one pixel: a single 16-byte LDS read.  A (the weights,
// arranged per lane on the host) stays in registers for a whole layer.
//
// Layers 0..7 work on UNITS of 16*kPh columns x 2 rows (kPh = 5 for 64-wide tiles: 80
// columns, the widest region).  Item o of a unit (o < kPh) holds the columns x + o + kPh*j,
// so lane j of items o = 0..kPh-1 reads the pixels x + kPh*j - 1 .. x + kPh*j + kPh: kPh+2
// LDS reads feed 3*kPh MFMAs (instead of 3*kPh reads for 16 contiguous pixels per item),
// and the 80-byte lane stride keeps the 16-lane reads and writes free of bank conflicts.
// Each wave takes whole row pairs.  Layer 8 uses contiguous 16-pixel items so that the
// deconv's stores are contiguous.
//
// D lands as lane l holding rows 4*(l/16)..+3 of column l%16: 4 channels of one pixel.
// The bias is the accumulator's initial value; the sum is rounded to 16 bits (RNE) and
// ReLU'd on the packed bits (max with +0 as signed 16-bit integers, exact for both
// formats), then written back in place; after layer 8 the FP32 sums feed the deconv.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "mfma16.hpp"

namespace achip {

template <int TW, int TH, int NT>
struct MfmaTile {
    static constexpr int NW = NT / 64;
    static constexpr int BW = TW + 16, BH = TH + 16;                // tile + margin 8
    static constexpr int kPh = (TW + 16 + 15) / 16;                 // phases (items) of a unit
    static constexpr int kSeg8 = (TW + 15) / 16;                    // layer-8 items per row pair
    // Feature plane, origin (x0 - 8, y0 - 8), 8 x 16-bit channels per pixel.  A unit starts
    // at local column <= 8 and reads up to 16*kPh columns right of it; row pairs read one
    // row past a layer's region.  The pitch is a multiple of 16 pixels (256 bytes): the two
    // 16-lane groups of a half-wave read rows one pitch apart, and 16-byte LDS reads only
    // run at full rate when those addresses differ by a multiple of 256 bytes.
    static constexpr int PW = (8 + 16 * kPh + 1 + 15) / 16 * 16;
    static constexpr int PH = BH + 1;
    // input tile in the source element type, origin (x0 - 9, y0 - 9): plane-local (r, c) is
    // tile (r + 1, c + 1)
    static constexpr int IW = PW + 1, IH = TH + 18;
    static constexpr int kStage = (IW * IH + NT - 1) / NT;          // input elements per thread
    static constexpr int kMaxPairs = ((BH + 1) / 2 + NW - 1) / NW;  // row pairs per wave and layer
    static constexpr int plane_bytes = PW * PH * 16;
    static constexpr int lds_bytes(int elem) { return plane_bytes + (IW * IH * elem + 15) / 16 * 16; }
    static_assert(16 * kSeg8 + 8 <= PW - 1, "layer-8 reads stay in the plane");
    static_assert(kPh % 2 == 1, "an odd unit stride keeps 16-lane LDS accesses conflict-free");
};

// Replicate ring.  After a layer of margin m has written its in-image region, copy the
// image's edge pixels one pixel outward wherever the margin-m band reaches past the image
// border.  The next layer then reads every tap without clamping and still sees the
// reference's per-layer replicate border (Common.hpp:122-143).  Interior tiles need none.
// (A row pair whose second row lies outside a cut region still reads input row y+2 for
// it, with zero weights for row y -- but 0 x NaN is NaN.  In a tile whose band touches the
// image border such a row may never be written, so those tiles zero the plane first;
// interior tiles only read unwritten pixels for columns whose results are discarded.)
template <int TW, int TH, int NT>
__device__ inline void fill_ring(uint4* __restrict__ plane, int x0, int y0, int W, int H, int m)
{
    using C = MfmaTile<TW, TH, NT>;
    const bool L = x0 - m <= -1, R = x0 + TW + m >= W + 1;
    const bool T = y0 - m <= -1, B = y0 + TH + m >= H + 1;
    if (!(L || R || T || B)) return;
    const int bx = x0 - 8, by = y0 - 8;
    const int gx0 = max(x0 - m, 0), gx1 = min(x0 + TW + m, W);
    const int gy0 = max(y0 - m, 0), gy1 = min(y0 + TH + m, H);
    const int ry0 = gy0 - (T ? 1 : 0), nrow = gy1 + (B ? 1 : 0) - ry0;
    const int rx0 = gx0 - (L ? 1 : 0), ncol = gx1 + (R ? 1 : 0) - rx0;
    const int nL = L ? nrow : 0, nR = R ? nrow : 0, nT = T ? ncol : 0, nB = B ? ncol : 0;
    const int total = nL + nR + nT + nB;
    for (int i = threadIdx.x; i < total; i += NT) {
        int gx, gy;
        if (i < nL) { gx = -1; gy = ry0 + i; }
        else if (i < nL + nR) { gx = W; gy = ry0 + i - nL; }
        else if (i < nL + nR + nT) { gx = rx0 + i - nL - nR; gy = -1; }
        else { gx = rx0 + i - nL - nR - nT; gy = H; }
        const int sx = clampi(gx, 0, W - 1), sy = clampi(gy, 0, H - 1);
        plane[(gy - by) * C::PW + gx - bx] = plane[(sy - by) * C::PW + sx - bx];
    }
}

// The whole network on one tile whose top-left input pixel is (x0, y0); plane/img are the
// workgroup's LDS arrays.  wa: per-lane 16-bit A operands of layers 1..8; wa0: per-lane
// FP32 A operands of layer 0.  Ends without a barrier.
template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__device__ __forceinline__ void mfma_tile(const void* __restrict__ src, long sstride, int W, int H,
                                          void* __restrict__ dst, long dstride,
                                          const float* __restrict__ wts, const uint4* __restrict__ wa,
                                          const float* __restrict__ wa0,
                                          int x0, int y0, uint4* __restrict__ plane, TIN* __restrict__ img)
{
    using C = MfmaTile<TW, TH, NT>;
    using M = Mma<T16>;
    using v8 = typename M::v8;
    constexpr int P = C::kPh;
    // The thread index passes through an empty asm statement, so nothing derived from it
    // (LDS offsets, operand loads) is hoisted out of a persistent kernel's tile loop: held
    // across the loop, those values pushed the 1024-thread batch kernel past its 128 VGPRs
    // into scratch spills (18 VGPRs) and the 512-thread one to 160 VGPRs, one workgroup per
    // CU (126 VGPRs for both and no spills with this).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);    // wave-uniform
    const int bx = x0 - 8, by = y0 - 8;
    const int j = lane & 15;          // column of the item
    const int g = lane >> 4;          // lane group: patch row (B), D row block
    const int half = g & 1;           // channels 0-3 or 4-7 of D
    const int ro = g >> 1;            // output row y or y+1 of D

    // ---- border tiles (the margin-8 band reaches the image edge) start from a zeroed
    // plane and keep a replicate ring (see fill_ring) ----
    const bool border = x0 - 8 <= 0 || y0 - 8 <= 0 || x0 + TW + 8 >= W || y0 + TH + 8 >= H;
    if (border)
        for (int i = tid; i < C::PW * C::PH; i += NT) plane[i] = make_uint4(0u, 0u, 0u, 0u);

    // ---- stage the input tile over its whole extent (replicate clamp), all loads in
    // flight before the first LDS store ----
    {
        constexpr int DX = NT % C::IW, DY = NT / C::IW;
        int iy = tid / C::IW, ix = tid - (tid / C::IW) * C::IW;
        TIN v[C::kStage];
#pragma unroll
        for (int k = 0; k < C::kStage; k++) {
            if (k < C::kStage - 1 || tid + k * NT < C::IW * C::IH) {
                const int gy = clampi(y0 - 9 + iy, 0, H - 1);
                const int gx = clampi(x0 - 9 + ix, 0, W - 1);
                v[k] = row_ptr<TIN>(src, sstride, gy)[gx];
            }
            ix += DX;
            iy += DY;
            if (ix >= C::IW) { ix -= C::IW; iy++; }
        }
#pragma unroll
        for (int k = 0; k < C::kStage; k++)
            if (k < C::kStage - 1 || tid + k * NT < C::IW * C::IH) img[tid + k * NT] = v[k];
    }
    __syncthreads();

    // ---- layers 0..7 on units, in place ----
    // The next layer's A operands and bias are loaded at the end of each compute phase, so
    // that their latency hides behind the barrier and the write-back.
    const int lane_px = g * C::PW + P * j;                  // B read (pixels), plane
    const int lane_im = g * C::IW + P * j;                  // B read, input tile
    const int lane_wr = (ro * C::PW + P * j) * 2 + half;    // write-back, 8-byte units
    const float a0f = wa0[lane], a1f = wa0[64 + lane], a2f = wa0[128 + lane];
    v8 a0 = {}, a1 = {}, a2 = {};
    f32x4 bias;
    {
        const float* bl = wts + kOffB0 + 4 * half;
        bias = (f32x4){ bl[0], bl[1], bl[2], bl[3] };
    }
#pragma nounroll
    for (int l = 0; l <= 7; l++) {
        const int m = 8 - l;
        const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
        const int gx1 = min(x0 + TW + m, W), gy1 = min(y0 + TH + m, H);
        const int lx0 = gx0 - bx, ly0 = gy0 - by;
        const int npairs = (gy1 - gy0 + 1) >> 1;
        uint2 res[C::kMaxPairs][P];
        if (l == 0) {
#pragma unroll
            for (int k = 0; k < C::kMaxPairs; k++) {
                const int pair = wave + k * C::NW;
                if (pair < npairs) {
                    const TIN* p = img + (ly0 + 2 * pair) * C::IW + lx0 + lane_im;
                    float q[P + 2];
#pragma unroll
                    for (int d = 0; d < P + 2; d++) q[d] = to_float<TIN>(p[d]);
#pragma unroll
                    for (int o = 0; o < P; o++) {
                        f32x4 acc = mma_f32(a0f, q[o], bias);
                        acc = mma_f32(a1f, q[o + 1], acc);
                        acc = mma_f32(a2f, q[o + 2], acc);
                        res[k][o] = make_uint2(relu_pack<T16>(acc[0], acc[1]), relu_pack<T16>(acc[2], acc[3]));
                    }
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < C::kMaxPairs; k++) {
                const int pair = wave + k * C::NW;
                if (pair < npairs) {
                    const uint4* p = plane + (ly0 - 1 + 2 * pair) * C::PW + lx0 - 1 + lane_px;
                    v8 q[P + 2];
#pragma unroll
                    for (int d = 0; d < P + 2; d++) q[d] = __builtin_bit_cast(v8, p[d]);
#pragma unroll
                    for (int o = 0; o < P; o++) {
                        f32x4 acc = M::mma(a0, q[o], bias);
                        acc = M::mma(a1, q[o + 1], acc);
                        acc = M::mma(a2, q[o + 2], acc);
                        res[k][o] = make_uint2(relu_pack<T16>(acc[0], acc[1]), relu_pack<T16>(acc[2], acc[3]));
                    }
                }
            }
        }
        {   // operands of layer l + 1 (1..8)
            const uint4* al = wa + l * 3 * 64 + lane;
            a0 = __builtin_bit_cast(v8, al[0]);
            a1 = __builtin_bit_cast(v8, al[64]);
            a2 = __builtin_bit_cast(v8, al[128]);
            const float* bl = wts + offB(l + 1) + 4 * half;
            bias = (f32x4){ bl[0], bl[1], bl[2], bl[3] };
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < C::kMaxPairs; k++) {
            const int pair = wave + k * C::NW;
            if (pair < npairs && gy0 + 2 * pair + ro < gy1) {
                uint2* w = reinterpret_cast<uint2*>(plane) + ((ly0 + 2 * pair) * C::PW + lx0) * 2 + lane_wr;
#pragma unroll
                for (int o = 0; o < P; o++)
                    if (gx0 + o + P * j < gx1) w[2 * o] = res[k][o];
            }
        }
        __syncthreads();
        if (border) {
            fill_ring<TW, TH, NT>(plane, x0, y0, W, H, m);
            __syncthreads();
        }
    }

    // ---- layer 8 (contiguous 16-pixel items), ReLU, then the deconv 2x2 stride 2
    // (8 -> 1) in FP32 ----
    {
        const int gx1 = min(x0 + TW, W), gy1 = min(y0 + TH, H);
        const int npairs = (gy1 - y0 + 1) >> 1;
        // this lane's 4 channels x 4 output sub-pixels of the deconv, k2[c][q]
        const float* k2 = wts + kOffK2 + 16 * half;
        float kd[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) kd[c][q] = k2[c * 4 + q];
        const int lane_rd = g * C::PW + j;
        for (int pair = wave; pair < npairs; pair += C::NW) {
            // region origin is plane-local (8, 8); B row y-1+g, column x+j+t-1
            const uint4* p = plane + (7 + 2 * pair) * C::PW + 7 + lane_rd;
            v8 q[C::kSeg8][3];
#pragma unroll
            for (int s = 0; s < C::kSeg8; s++)
#pragma unroll
                for (int t = 0; t < 3; t++) q[s][t] = __builtin_bit_cast(v8, p[16 * s + t]);
            const int oy = y0 + 2 * pair + ro;
#pragma unroll
            for (int s = 0; s < C::kSeg8; s++) {
                f32x4 acc = M::mma(a0, q[s][0], bias);
                acc = M::mma(a1, q[s][1], acc);
                acc = M::mma(a2, q[s][2], acc);
                const float v0 = fmaxf(acc[0], 0.f), v1 = fmaxf(acc[1], 0.f);
                const float v2 = fmaxf(acc[2], 0.f), v3 = fmaxf(acc[3], 0.f);
                // this lane's 4 channels, then the partner lane (lane ^ 16) holds the other 4
                float sum[4];
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++)
                    sum[q4] = fmaf(v3, kd[3][q4], fmaf(v2, kd[2][q4], fmaf(v1, kd[1][q4], v0 * kd[0][q4])));
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++) sum[q4] += __shfl_xor(sum[q4], 16);
                const int gx = x0 + 16 * s + j;
                if (gx < gx1 && oy < gy1) {
                    TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * oy + half);
                    store_pair<TOUT>(orow + 2 * gx, half ? sum[2] : sum[0], half ? sum[3] : sum[1]);
                }
            }
        }
    }
}

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                       void* __restrict__ dst, long dstride,
                                                       const float* __restrict__ wts, const uint4* __restrict__ wa,
                                                       const float* __restrict__ wa0)
{
    using C = MfmaTile<TW, TH, NT>;
    __shared__ uint4 plane[C::PW * C::PH];        // 8 x 16-bit channels per pixel
    __shared__ TIN img[C::IW * C::IH];
    const int t = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
    mfma_tile<T16, TIN, TOUT, TW, TH, NT>(src, sstride, W, H, dst, dstride, wts, wa, wa0, tx * TW, ty * TH, plane, img);
}

// Persistent batched form (see acnet_fp32_batch).
template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma_batch(const BatchDesc* __restrict__ desc, int n, int total,
                                                       int* __restrict__ counter,
                                                       const float* __restrict__ wts, const uint4* __restrict__ wa,
                                                       const float* __restrict__ wa0)
{
    using C = MfmaTile<TW, TH, NT>;
    __shared__ uint4 plane[C::PW * C::PH];
    __shared__ TIN img[C::IW * C::IH];
    __shared__ int s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    int t = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: keeps the tile state scalar
    while (t < total) {
        // The next tile is drawn as this one starts and read once it is done, so the atomic's
        // round trip overlaps the tile instead of leaving the CU idle between tiles (one or
        // two workgroups per CU).  Each workgroup still draws exactly one index past its last
        // tile.
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        const BatchDesc d = find_tile(desc, n, t);
        const int local = t - d.first_tile;
        const int ly = local / d.tiles_x;
        const int tx = local - ly * d.tiles_x;
        const int ty = d.ty0 + ly;
        mfma_tile<T16, TIN, TOUT, TW, TH, NT>(d.src, d.sstride, d.w, d.h, d.dst, d.dstride, wts, wa, wa0,
                                              tx * TW, ty * TH, plane, img);
        __syncthreads();                    // the tile's LDS and s_tile are free again
        if (threadIdx.x == 0) s_tile = next;
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(s_tile);
    }
}

// ---- launch table ---------------------------------------------------------------------

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_one(const void* src, long ss, int W, int H, void* dst, long ds,
                             const float* wts, const uint4* wa, const float* wa0, hipStream_t stream)
{
    dim3 grid((W + TW - 1) / TW, (H + TH - 1) / TH);
    hipLaunchKernelGGL((acnet_mfma_fused<T16, TIN, TOUT, TW, TH, NT>), grid, dim3(NT), 0, stream,
                       src, ss, W, H, dst, ds, wts, wa, wa0);
    return hipGetLastError();
}

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_batch_one(const BatchDesc* desc, int n, int total, int* counter, const float* wts,
                                   const uint4* wa, const float* wa0, int max_wg, hipStream_t stream)
{
    auto kern = acnet_mfma_batch<T16, TIN, TOUT, TW, TH, NT>;
    const int grid = persistent_grid(reinterpret_cast<const void*>(kern), NT, total, max_wg);
    if (grid <= 0) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, stream, desc, n, total, counter, wts, wa, wa0);
    return hipGetLastError();
}

template <typename T16, int TW, int TH, int NT>
static hipError_t launch_batch_types(int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                                     const float* wts, const uint4* wa, const float* wa0, int max_wg, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                                \
    switch (dtype) {                                                                                                  \
    case U8: return launch_batch_one<T16, TIN, uint8_t, TW, TH, NT>(desc, n, total, counter, wts, wa, wa0, max_wg, stream);        \
    case U16: return launch_batch_one<T16, TIN, uint16_t, TW, TH, NT>(desc, n, total, counter, wts, wa, wa0, max_wg, stream);      \
    case F16: return launch_batch_one<T16, TIN, __half, TW, TH, NT>(desc, n, total, counter, wts, wa, wa0, max_wg, stream);        \
    case F32: return launch_batch_one<T16, TIN, float, TW, TH, NT>(desc, n, total, counter, wts, wa, wa0, max_wg, stream);         \
    default: return hipErrorInvalidValue;                                                                             \
    }
    switch (stype) {
    case U8: ACHIP_OUT(uint8_t)
    case U16: ACHIP_OUT(uint16_t)
    case F16: ACHIP_OUT(__half)
    case F32: ACHIP_OUT(float)
    default: return hipErrorInvalidValue;
    }
#undef ACHIP_OUT
}

template <typename T16, int TW, int TH, int NT>
static hipError_t launch_types(int stype, int dtype, const void* src, long ss, int W, int H,
                               void* dst, long ds, const float* wts, const uint4* wa, const float* wa0, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                        \
    switch (dtype) {                                                                                          \
    case U8: return launch_one<T16, TIN, uint8_t, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, wa0, stream);       \
    case U16: return launch_one<T16, TIN, uint16_t, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, wa0, stream);     \
    case F16: return launch_one<T16, TIN, __half, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, wa0, stream);       \
    case F32: return launch_one<T16, TIN, float, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, wa0, stream);        \
    default: return hipErrorInvalidValue;                                                                     \
    }
    switch (stype) {
    case U8: ACHIP_OUT(uint8_t)
    case U16: ACHIP_OUT(uint16_t)
    case F16: ACHIP_OUT(__half)
    case F32: ACHIP_OUT(float)
    default: return hipErrorInvalidValue;
    }
#undef ACHIP_OUT
}

// Config 0, the default for u8 input where none is chosen (untuned sizes, the serving loop's
// steps): 64x32 tiles at 512 threads, two workgroups per CU, so one tile's serial phases
// (staging, barriers, write-backs) overlap the other's MFMAs.  Measured on 1080p u8 frames:
// 58.7 us per launch against 63.6 for 64x64 at 1024 threads (49.5 against 55.2 us per frame
// over two streams), 1.71 against 1.98 ms for the 64-image batch.  With wider inputs only
// one such workgroup fits a CU's LDS, and config 1 is the default.
const Fp32Config kMfmaConfigs[] = {
    { 64, 32, 512, 0, MfmaTile<64, 32, 512>::lds_bytes(4) },     // LDS for f32 input; less for narrower types
    { 64, 64, 1024, 0, MfmaTile<64, 64, 1024>::lds_bytes(4) },
};
const int kMfmaConfigCount = sizeof(kMfmaConfigs) / sizeof(kMfmaConfigs[0]);

hipError_t launch_mfma_batch(int bf16, int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total,
                             int* counter, const float* wts, const void* wa, const float* wa0, int max_wg,
                             hipStream_t stream)
{
    const uint4* a = static_cast<const uint4*>(wa);
    if (bf16) {
        switch (cfg) {
        case 0: return launch_batch_types<__bf16, 64, 32, 512>(stype, dtype, desc, n, total, counter, wts, a, wa0, max_wg, stream);
        case 1: return launch_batch_types<__bf16, 64, 64, 1024>(stype, dtype, desc, n, total, counter, wts, a, wa0, max_wg, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (cfg) {
    case 0: return launch_batch_types<_Float16, 64, 32, 512>(stype, dtype, desc, n, total, counter, wts, a, wa0, max_wg, stream);
    case 1: return launch_batch_types<_Float16, 64, 64, 1024>(stype, dtype, desc, n, total, counter, wts, a, wa0, max_wg, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mfma(int bf16, int cfg, int stype, int dtype, const void* src, long ss, int W, int H,
                       void* dst, long ds, const float* wts, const void* wa, const float* wa0, hipStream_t stream)
{
    const uint4* a = static_cast<const uint4*>(wa);
    if (bf16) {
        switch (cfg) {
        case 0: return launch_types<__bf16, 64, 32, 512>(stype, dtype, src, ss, W, H, dst, ds, wts, a, wa0, stream);
        case 1: return launch_types<__bf16, 64, 64, 1024>(stype, dtype, src, ss, W, H, dst, ds, wts, a, wa0, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (cfg) {
    case 0: return launch_types<_Float16, 64, 32, 512>(stype, dtype, src, ss, W, H, dst, ds, wts, a, wa0, stream);
    case 1: return launch_types<_Float16, 64, 64, 1024>(stype, dtype, src, ss, W, H, dst, ds, wts, a, wa0, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace achip
