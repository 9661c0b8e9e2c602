// Fused ACNetLegacy 2x pass with the 8->8 layers as an FP16 or BF16 implicit GEMM on
// gfx950 MFMA (v_mfma_f32_16x16x32_{f16,bf16}), FP32 accumulation; layer 0, bias, ReLU
// and the deconv stay in FP32.  Same tiling, halo and per-layer border clamps as the
// fused FP32 kernel (acnet_fp32.hip); the difference is how layers 1..8 are computed.
//
// The GEMM for one wave and one work item (16 consecutive pixels x 2 rows):
//   D[16 x 16] = A[16 x 96] * B[96 x 16]
//   rows i of D/A : i = 8*ro + n, output channel n of output row y+ro (ro = 0, 1);
//   columns j     : the 16 pixels x0+j;
//   K = 96        : a 4 x 3 input patch (rows y-1..y+2, columns x-1..x+1) x 8 channels.
// Both output rows are linear in the same 4 x 3 patch, so all 16 MFMA rows are useful
// (the 8 output channels alone would leave half of them empty); 72 of each row's 96
// weights are nonzero.  K is split over three MFMAs, one per patch column t, and inside
// one MFMA lane group g = lane/16 holds patch row g, so each lane's B operand is the 8
// channels of one pixel: a single 16-byte LDS read.  A (the weights, arranged per lane on
// the host) stays in registers for a whole layer.
//
// D lands as lane l holding rows 4*(l/16)..+3 of column l%16: 4 channels of one pixel.
// Those are biased, ReLU'd in FP32, rounded to 16 bits and written back in place; after
// layer 8 they feed the FP32 deconv directly.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Mma;
template <> struct Mma<_Float16> {
    using v8 = f16x8;
    __device__ static f32x4 mma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    __device__ static uint32_t pack2(float lo, float hi)
    {
        const _Float16 a = static_cast<_Float16>(lo), b = static_cast<_Float16>(hi);   // RNE
        return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, a)) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, b)) << 16);
    }
};
template <> struct Mma<__bf16> {
    using v8 = bf16x8;
    __device__ static f32x4 mma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
    __device__ static uint32_t pack2(float lo, float hi)
    {
        const __bf16 a = static_cast<__bf16>(lo), b = static_cast<__bf16>(hi);          // RNE
        return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, a)) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, b)) << 16);
    }
};

template <int TW, int TH, int NT>
struct MfmaTile {
    static constexpr int NW = NT / 64;
    static constexpr int BW = TW + 16, BH = TH + 16;
    static constexpr int IW = TW + 18, IH = TH + 18;
    static constexpr int items(int m) { return ((TW + 2 * m + 15) / 16) * ((TH + 2 * m + 1) / 2); }
    static constexpr int kMaxItems = (items(7) + NW - 1) / NW;     // per wave, layers 1..7
    static constexpr int kL0Runs = 4;                               // layer-0 pixels per thread run
    static constexpr int lds_bytes = BW * BH * 16 + IW * IH * 4;
};

// The whole network on one tile whose top-left input pixel is (x0, y0); plane/img are the
// workgroup's LDS arrays.  Ends without a barrier.
template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__device__ __forceinline__ void mfma_tile(const void* __restrict__ src, long sstride, int W, int H,
                                          void* __restrict__ dst, long dstride,
                                          const float* __restrict__ wts, const uint4* __restrict__ wa,
                                          int x0, int y0, uint4* __restrict__ plane, float* __restrict__ img)
{
    using C = MfmaTile<TW, TH, NT>;
    using M = Mma<T16>;
    using v8 = typename M::v8;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int bx = x0 - 8, by = y0 - 8;

    // ---- stage the input tile: toFloat, replicate clamp ----
    for (int i = tid; i < C::IW * C::IH; i += NT) {
        const int iy = i / C::IW;
        const int ix = i - iy * C::IW;
        const int gy = clampi(y0 - 9 + iy, 0, H - 1);
        const int gx = clampi(x0 - 9 + ix, 0, W - 1);
        img[i] = to_float<TIN>(row_ptr<TIN>(src, sstride, gy)[gx]);
    }
    __syncthreads();

    // ---- layer 0 in FP32 over margin 8, rounded to 16 bits into the plane ----
    {
        constexpr int P = C::kL0Runs;
        const int gx0 = max(x0 - 8, 0), gy0 = max(y0 - 8, 0);
        const int cols = min(x0 + TW + 8, W) - gx0;
        const int gy1 = min(y0 + TH + 8, H);
        const int nitems = cols * ((gy1 - gy0 + P - 1) / P);
        for (int item = tid; item < nitems; item += NT) {
            const int run = item / cols;
            const int col = item - run * cols;
            const int gx = gx0 + col, gy = gy0 + run * P;
            const int lx = gx - bx, ly = gy - by;
            float v[P + 2][3];
#pragma unroll
            for (int r = 0; r < P + 2; r++) {
                const int iy = min(ly + r, C::IH - 1);
#pragma unroll
                for (int k = 0; k < 3; k++) v[r][k] = img[iy * C::IW + lx + k];
            }
            float acc[P][8];
#pragma unroll
            for (int p = 0; p < P; p++)
#pragma unroll
                for (int n = 0; n < 8; n++) acc[p][n] = wts[kOffB0 + n];
#pragma unroll
            for (int ky = 0; ky < 3; ky++)
#pragma unroll
                for (int kx = 0; kx < 3; kx++)
#pragma unroll
                    for (int n = 0; n < 8; n++) {
                        const float w = wts[kOffW0 + (ky * 3 + kx) * 8 + n];
#pragma unroll
                        for (int p = 0; p < P; p++) acc[p][n] = fmaf(v[p + ky][kx], w, acc[p][n]);
                    }
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (gy + p < gy1) {
                    uint4 q;
                    q.x = M::pack2(fmaxf(acc[p][0], 0.f), fmaxf(acc[p][1], 0.f));
                    q.y = M::pack2(fmaxf(acc[p][2], 0.f), fmaxf(acc[p][3], 0.f));
                    q.z = M::pack2(fmaxf(acc[p][4], 0.f), fmaxf(acc[p][5], 0.f));
                    q.w = M::pack2(fmaxf(acc[p][6], 0.f), fmaxf(acc[p][7], 0.f));
                    plane[(ly + p) * C::BW + lx] = q;
                }
            }
        }
    }
    __syncthreads();

    // ---- layers 1..8 on MFMA ----
    const int j = lane & 15;          // pixel column inside the item
    const int g = lane >> 4;          // lane group: patch row (B), D row block
    const int half = g & 1;           // channels 0-3 or 4-7 of D
    const int ro = g >> 1;            // output row y or y+1 of D
    const float* k2 = wts + kOffK2;
#pragma nounroll
    for (int l = 1; l <= 8; l++) {
        const int m = 8 - l;
        const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
        const int gx1 = min(x0 + TW + m, W), gy1 = min(y0 + TH + m, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        // A operands of this layer (per lane, arranged on the host) and this lane's 4 biases
        const uint4* al = wa + (l - 1) * 3 * 64 + lane;
        const v8 a0 = __builtin_bit_cast(v8, al[0]);
        const v8 a1 = __builtin_bit_cast(v8, al[64]);
        const v8 a2 = __builtin_bit_cast(v8, al[128]);
        const float* bl = wts + offB(l) + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];

        uint2 res[C::kMaxItems];
#pragma unroll
        for (int ii = 0; ii < C::kMaxItems; ii++) {
            const int item = wave + ii * C::NW;           // wave-uniform
            if (item < nitems) {
                const int pair = item / nseg;
                const int seg = item - pair * nseg;
                const int gx = gx0 + seg * 16 + j;
                const int gy = gy0 + 2 * pair;
                // B: pixel gx + t - 1 of input row gy - 1 + g, clamped to the image, then
                // to the buffer (lanes past the region read in bounds and are discarded)
                int row = clampi(gy - 1 + g, 0, H - 1) - by;
                row = min(row, C::BH - 1);
                const int cc = min(gx, W - 1);
                const int c0 = min(max(cc - 1, 0) - bx, C::BW - 1);
                const int c1 = min(cc - bx, C::BW - 1);
                const int c2 = min(min(cc + 1, W - 1) - bx, C::BW - 1);
                const uint4* prow = plane + row * C::BW;
                const v8 q0 = __builtin_bit_cast(v8, prow[c0]);
                const v8 q1 = __builtin_bit_cast(v8, prow[c1]);
                const v8 q2 = __builtin_bit_cast(v8, prow[c2]);
                f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
                acc = M::mma(a0, q0, acc);
                acc = M::mma(a1, q1, acc);
                acc = M::mma(a2, q2, acc);
                const float v0 = fmaxf(acc[0] + b0, 0.f), v1 = fmaxf(acc[1] + b1, 0.f);
                const float v2 = fmaxf(acc[2] + b2, 0.f), v3 = fmaxf(acc[3] + b3, 0.f);
                if (l < 8) {
                    res[ii].x = M::pack2(v0, v1);
                    res[ii].y = M::pack2(v2, v3);
                } else {
                    // deconv 2x2 stride 2 (8 -> 1) in FP32: this lane's 4 channels, then the
                    // partner lane (lane ^ 16) holds the other 4
                    float s[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const float w0 = half ? k2[16 + q] : k2[q];
                        const float w1 = half ? k2[20 + q] : k2[4 + q];
                        const float w2 = half ? k2[24 + q] : k2[8 + q];
                        const float w3 = half ? k2[28 + q] : k2[12 + q];
                        s[q] = fmaf(v3, w3, fmaf(v2, w2, fmaf(v1, w1, v0 * w0)));
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) s[q] += __shfl_xor(s[q], 16);
                    const int oy = gy + ro;
                    if (gx < gx1 && oy < gy1) {
                        TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * oy + half);
                        orow[2 * gx] = from_float<TOUT>(half ? s[2] : s[0]);
                        orow[2 * gx + 1] = from_float<TOUT>(half ? s[3] : s[1]);
                    }
                }
            }
        }
        if (l < 8) {
            __syncthreads();
#pragma unroll
            for (int ii = 0; ii < C::kMaxItems; ii++) {
                const int item = wave + ii * C::NW;
                if (item < nitems) {
                    const int pair = item / nseg;
                    const int seg = item - pair * nseg;
                    const int gx = gx0 + seg * 16 + j;
                    const int oy = gy0 + 2 * pair + ro;
                    if (gx < gx1 && oy < gy1)
                        reinterpret_cast<uint2*>(plane)[2 * ((oy - by) * C::BW + (gx - bx)) + half] = res[ii];
                }
            }
            __syncthreads();
        }
    }
}

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                       void* __restrict__ dst, long dstride,
                                                       const float* __restrict__ wts, const uint4* __restrict__ wa)
{
    using C = MfmaTile<TW, TH, NT>;
    __shared__ uint4 plane[C::BW * C::BH];        // 8 x 16-bit channels per pixel
    __shared__ float img[C::IW * C::IH];
    mfma_tile<T16, TIN, TOUT, TW, TH, NT>(src, sstride, W, H, dst, dstride, wts, wa,
                                          blockIdx.x * TW, blockIdx.y * TH, plane, img);
}

// Persistent batched form (see acnet_fp32_batch).
template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma_batch(const BatchDesc* __restrict__ desc, int n, int total,
                                                       int* __restrict__ counter,
                                                       const float* __restrict__ wts, const uint4* __restrict__ wa)
{
    using C = MfmaTile<TW, TH, NT>;
    __shared__ uint4 plane[C::BW * C::BH];
    __shared__ float img[C::IW * C::IH];
    __shared__ int s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
        __syncthreads();
        const int t = s_tile;
        if (t >= total) break;
        const BatchDesc d = find_tile(desc, n, t);
        const int local = t - d.first_tile;
        const int ty = local / d.tiles_x;
        const int tx = local - ty * d.tiles_x;
        mfma_tile<T16, TIN, TOUT, TW, TH, NT>(d.src, d.sstride, d.w, d.h, d.dst, d.dstride, wts, wa,
                                              tx * TW, ty * TH, plane, img);
        __syncthreads();
    }
}

// ---- launch table ---------------------------------------------------------------------

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_one(const void* src, long ss, int W, int H, void* dst, long ds,
                             const float* wts, const uint4* wa, hipStream_t stream)
{
    dim3 grid((W + TW - 1) / TW, (H + TH - 1) / TH);
    hipLaunchKernelGGL((acnet_mfma_fused<T16, TIN, TOUT, TW, TH, NT>), grid, dim3(NT), 0, stream,
                       src, ss, W, H, dst, ds, wts, wa);
    return hipGetLastError();
}

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_batch_one(const BatchDesc* desc, int n, int total, int* counter, const float* wts,
                                   const uint4* wa, hipStream_t stream)
{
    auto kern = acnet_mfma_batch<T16, TIN, TOUT, TW, TH, NT>;
    const int grid = persistent_grid(reinterpret_cast<const void*>(kern), NT, total);
    if (grid <= 0) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, stream, desc, n, total, counter, wts, wa);
    return hipGetLastError();
}

template <typename T16, int TW, int TH, int NT>
static hipError_t launch_batch_types(int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                                     const float* wts, const uint4* wa, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                                \
    switch (dtype) {                                                                                                  \
    case U8: return launch_batch_one<T16, TIN, uint8_t, TW, TH, NT>(desc, n, total, counter, wts, wa, stream);        \
    case U16: return launch_batch_one<T16, TIN, uint16_t, TW, TH, NT>(desc, n, total, counter, wts, wa, stream);      \
    case F16: return launch_batch_one<T16, TIN, __half, TW, TH, NT>(desc, n, total, counter, wts, wa, stream);        \
    case F32: return launch_batch_one<T16, TIN, float, TW, TH, NT>(desc, n, total, counter, wts, wa, stream);         \
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
                               void* dst, long ds, const float* wts, const uint4* wa, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                        \
    switch (dtype) {                                                                                          \
    case U8: return launch_one<T16, TIN, uint8_t, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, stream);       \
    case U16: return launch_one<T16, TIN, uint16_t, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, stream);     \
    case F16: return launch_one<T16, TIN, __half, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, stream);       \
    case F32: return launch_one<T16, TIN, float, TW, TH, NT>(src, ss, W, H, dst, ds, wts, wa, stream);        \
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

const Fp32Config kMfmaConfigs[] = {
    { 64, 64, 1024, 2, MfmaTile<64, 64, 1024>::lds_bytes },
    { 64, 32, 512, 2, MfmaTile<64, 32, 512>::lds_bytes },
};
const int kMfmaConfigCount = sizeof(kMfmaConfigs) / sizeof(kMfmaConfigs[0]);

hipError_t launch_mfma_batch(int bf16, int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total,
                             int* counter, const float* wts, const void* wa, hipStream_t stream)
{
    const uint4* a = static_cast<const uint4*>(wa);
    if (bf16) {
        switch (cfg) {
        case 0: return launch_batch_types<__bf16, 64, 64, 1024>(stype, dtype, desc, n, total, counter, wts, a, stream);
        case 1: return launch_batch_types<__bf16, 64, 32, 512>(stype, dtype, desc, n, total, counter, wts, a, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (cfg) {
    case 0: return launch_batch_types<_Float16, 64, 64, 1024>(stype, dtype, desc, n, total, counter, wts, a, stream);
    case 1: return launch_batch_types<_Float16, 64, 32, 512>(stype, dtype, desc, n, total, counter, wts, a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mfma(int bf16, int cfg, int stype, int dtype, const void* src, long ss, int W, int H,
                       void* dst, long ds, const float* wts, const void* wa, hipStream_t stream)
{
    const uint4* a = static_cast<const uint4*>(wa);
    if (bf16) {
        switch (cfg) {
        case 0: return launch_types<__bf16, 64, 64, 1024>(stype, dtype, src, ss, W, H, dst, ds, wts, a, stream);
        case 1: return launch_types<__bf16, 64, 32, 512>(stype, dtype, src, ss, W, H, dst, ds, wts, a, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (cfg) {
    case 0: return launch_types<_Float16, 64, 64, 1024>(stype, dtype, src, ss, W, H, dst, ds, wts, a, stream);
    case 1: return launch_types<_Float16, 64, 32, 512>(stype, dtype, src, ss, W, H, dst, ds, wts, a, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace achip
