// Fused ACNetLegacy 2x pass in FP32 on gfx950 matrix cores (v_mfma_f32_16x16x4_f32):
// every conv layer, layer 0 included, is an FP32 implicit GEMM with FP32 accumulation;
// bias, ReLU and the deconv run in FP32 on the vector units.  Features stay f32 in LDS,
// so the result stays within the FP32 tolerance of the reference CPU path.
//
// Same work item as the 16-bit MFMA kernel (acnet_mfma.hip): 16 consecutive pixels x 2
// output rows from a 4 x 3 input patch, D[16 x 16] = A[16 x 96] B[96 x 16], all 16 MFMA
// rows useful (8 channels x 2 rows).  With K = 4 per instruction, lane group g = lane/16
// holds patch row g and one MFMA covers one (patch column t, channel c): 24 MFMAs per
// item, each lane's B operand one float of the pixel (row y-1+g, column x+j+t-1).  The
// 8 channels of a pixel come from two 16-byte LDS reads (feature planes 0-3 and 4-7).
// Layer 0 (one channel) is 3 MFMAs per item, B read from the staged input tile.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

typedef float f32x4m __attribute__((ext_vector_type(4)));

template <int TW, int TH, int NT>
struct Mfma32Tile {
    static constexpr int NW = NT / 64;
    static constexpr int BW = TW + 16, BH = TH + 16;
    // padded so that 16-lane segments and row pairs may overhang a layer's region
    static constexpr int PW = 8 + 16 * ((TW + 14 + 15) / 16) + 1;
    static constexpr int PH = BH + 1;
    // input tile, origin (x0 - 9, y0 - 9), padded the same way for layer 0's items
    static constexpr int IW = 16 * ((TW + 16 + 15) / 16) + 2 + 8;
    static constexpr int IH = TH + 18 + 2;
    static constexpr int items(int m) { return ((TW + 2 * m + 15) / 16) * ((TH + 2 * m + 1) / 2); }
    static constexpr int kMaxItems = (items(7) + NW - 1) / NW;
    static constexpr int lds_bytes = 2 * PW * PH * 16 + IW * IH * 4;
};

// Replicate ring and zeroed border tiles: see fill_ring in acnet_mfma.hip.
template <int TW, int TH, int NT>
__device__ inline void fill_ring32(float4* __restrict__ pa, float4* __restrict__ pb, int x0, int y0, int W, int H, int m)
{
    using C = Mfma32Tile<TW, TH, NT>;
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
        const int d = (gy - by) * C::PW + gx - bx, s = (sy - by) * C::PW + sx - bx;
        pa[d] = pa[s];
        pb[d] = pb[s];
    }
}

__device__ inline f32x4m mma32(float a, float b, f32x4m c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// the 24 MFMAs of one item on two accumulator chains; p: this lane's B row start
__device__ __forceinline__ f32x4m item24(const float4* __restrict__ pa, const float4* __restrict__ pb, int p,
                                         const float (&a)[24])
{
    f32x4m c0 = { 0.f, 0.f, 0.f, 0.f }, c1 = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const float4 lo = pa[p + t];
        const float4 hi = pb[p + t];
        c0 = mma32(a[t * 8 + 0], lo.x, c0);
        c1 = mma32(a[t * 8 + 1], lo.y, c1);
        c0 = mma32(a[t * 8 + 2], lo.z, c0);
        c1 = mma32(a[t * 8 + 3], lo.w, c1);
        c0 = mma32(a[t * 8 + 4], hi.x, c0);
        c1 = mma32(a[t * 8 + 5], hi.y, c1);
        c0 = mma32(a[t * 8 + 6], hi.z, c0);
        c1 = mma32(a[t * 8 + 7], hi.w, c1);
    }
    return c0 + c1;
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
__device__ __forceinline__ void mfma32_tile(const void* __restrict__ src, long sstride, int W, int H,
                                            void* __restrict__ dst, long dstride,
                                            const float* __restrict__ wts, const float* __restrict__ wa,
                                            int x0, int y0, float4* __restrict__ pa, float4* __restrict__ pb,
                                            float* __restrict__ img)
{
    using C = Mfma32Tile<TW, TH, NT>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int bx = x0 - 8, by = y0 - 8;
    const int j = lane & 15, g = lane >> 4, half = g & 1, ro = g >> 1;

    if (x0 - 8 <= 0 || y0 - 8 <= 0 || x0 + TW + 8 >= W || y0 + TH + 8 >= H)
        for (int i = tid; i < C::PW * C::PH; i += NT) {
            pa[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            pb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    // input tile with replicate clamp over its whole padded extent (every value finite)
    for (int i = tid; i < C::IW * C::IH; i += NT) {
        const int iy = i / C::IW;
        const int ix = i - iy * C::IW;
        const int gy = clampi(y0 - 9 + iy, 0, H - 1);
        const int gx = clampi(x0 - 9 + ix, 0, W - 1);
        img[i] = to_float<TIN>(row_ptr<TIN>(src, sstride, gy)[gx]);
    }
    __syncthreads();

    // ---- layer 0: 3 MFMAs per item (K = 4 patch rows, one per patch column) ----
    {
        const int gx0 = max(x0 - 8, 0), gy0 = max(y0 - 8, 0);
        const int gx1 = min(x0 + TW + 8, W), gy1 = min(y0 + TH + 8, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        const float a0 = wa[lane], a1 = wa[64 + lane], a2 = wa[128 + lane];
        const float* bl = wts + kOffB0 + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];
        const int lane_rd = g * C::IW + j;
        int pair = wave / nseg, seg = wave - (wave / nseg) * nseg;
        for (int item = wave; item < nitems; item += C::NW) {
            // img holds pixel (x, y) at ((y - y0 + 9) * IW + x - x0 + 9); B row y-1+g, column x+j+t-1
            const int base = (gy0 + 2 * pair - 1 - (y0 - 9)) * C::IW + (gx0 + 16 * seg - 1 - (x0 - 9));
            const float* p = img + base + lane_rd;
            f32x4m acc = { 0.f, 0.f, 0.f, 0.f };
            acc = mma32(a0, p[0], acc);
            acc = mma32(a1, p[1], acc);
            acc = mma32(a2, p[2], acc);
            const int gx = gx0 + 16 * seg + j;
            const int oy = gy0 + 2 * pair + ro;
            if (gx < gx1 && oy < gy1) {
                const float4 v = make_float4(fmaxf(acc[0] + b0, 0.f), fmaxf(acc[1] + b1, 0.f),
                                             fmaxf(acc[2] + b2, 0.f), fmaxf(acc[3] + b3, 0.f));
                const int d = (oy - by) * C::PW + gx - bx;
                if (half) pb[d] = v; else pa[d] = v;
            }
            seg += C::NW;
            while (seg >= nseg) { seg -= nseg; pair++; }
        }
    }
    __syncthreads();
    fill_ring32<TW, TH, NT>(pa, pb, x0, y0, W, H, 8);
    __syncthreads();

    const int lane_rd = g * C::PW + j;
    // ---- layers 1..7, in place ----
#pragma nounroll
    for (int l = 1; l <= 7; l++) {
        const int m = 8 - l;
        const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
        const int gx1 = min(x0 + TW + m, W), gy1 = min(y0 + TH + m, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        float a[24];
        const float* al = wa + 192 + (l - 1) * 24 * 64 + lane;
#pragma unroll
        for (int s = 0; s < 24; s++) a[s] = al[s * 64];
        const float* bl = wts + offB(l) + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];

        float4 res[C::kMaxItems];
        int pair = wave / nseg, seg = wave - (wave / nseg) * nseg;
#pragma unroll
        for (int ii = 0; ii < C::kMaxItems; ii++) {
            const int item = wave + ii * C::NW;
            if (item < nitems) {
                const int base = (gy0 + 2 * pair - 1 - by) * C::PW + (gx0 + 16 * seg - 1 - bx);
                const f32x4m acc = item24(pa, pb, base + lane_rd, a);
                res[ii] = make_float4(fmaxf(acc[0] + b0, 0.f), fmaxf(acc[1] + b1, 0.f),
                                      fmaxf(acc[2] + b2, 0.f), fmaxf(acc[3] + b3, 0.f));
            }
            seg += C::NW;
            while (seg >= nseg) { seg -= nseg; pair++; }
        }
        __syncthreads();
        pair = wave / nseg;
        seg = wave - pair * nseg;
#pragma unroll
        for (int ii = 0; ii < C::kMaxItems; ii++) {
            const int item = wave + ii * C::NW;
            if (item < nitems) {
                const int gx = gx0 + 16 * seg + j;
                const int oy = gy0 + 2 * pair + ro;
                if (gx < gx1 && oy < gy1) {
                    const int d = (oy - by) * C::PW + gx - bx;
                    if (half) pb[d] = res[ii]; else pa[d] = res[ii];
                }
            }
            seg += C::NW;
            while (seg >= nseg) { seg -= nseg; pair++; }
        }
        __syncthreads();
        fill_ring32<TW, TH, NT>(pa, pb, x0, y0, W, H, m);
        __syncthreads();
    }

    // ---- layer 8, ReLU, deconv 2x2 stride 2 (8 -> 1) ----
    {
        const int gx0 = x0, gy0 = y0;
        const int gx1 = min(x0 + TW, W), gy1 = min(y0 + TH, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        float a[24];
        const float* al = wa + 192 + 7 * 24 * 64 + lane;
#pragma unroll
        for (int s = 0; s < 24; s++) a[s] = al[s * 64];
        const float* bl = wts + offB(8) + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];
        const float* k2 = wts + kOffK2 + 16 * half;
        float kd[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) kd[c][q] = k2[c * 4 + q];
        int pair = wave / nseg, seg = wave - (wave / nseg) * nseg;
        for (int item = wave; item < nitems; item += C::NW) {
            const int base = (gy0 + 2 * pair - 1 - by) * C::PW + (gx0 + 16 * seg - 1 - bx);
            const f32x4m acc = item24(pa, pb, base + lane_rd, a);
            const float v0 = fmaxf(acc[0] + b0, 0.f), v1 = fmaxf(acc[1] + b1, 0.f);
            const float v2 = fmaxf(acc[2] + b2, 0.f), v3 = fmaxf(acc[3] + b3, 0.f);
            float s[4];
#pragma unroll
            for (int q = 0; q < 4; q++) s[q] = fmaf(v3, kd[3][q], fmaf(v2, kd[2][q], fmaf(v1, kd[1][q], v0 * kd[0][q])));
#pragma unroll
            for (int q = 0; q < 4; q++) s[q] += __shfl_xor(s[q], 16);
            const int gx = gx0 + 16 * seg + j;
            const int oy = gy0 + 2 * pair + ro;
            if (gx < gx1 && oy < gy1) {
                TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * oy + half);
                store_pair<TOUT>(orow + 2 * gx, half ? s[2] : s[0], half ? s[3] : s[1]);
            }
            seg += C::NW;
            while (seg >= nseg) { seg -= nseg; pair++; }
        }
    }
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma32_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                         void* __restrict__ dst, long dstride,
                                                         const float* __restrict__ wts, const float* __restrict__ wa)
{
    using C = Mfma32Tile<TW, TH, NT>;
    __shared__ float4 pa[C::PW * C::PH];
    __shared__ float4 pb[C::PW * C::PH];
    __shared__ float img[C::IW * C::IH];
    mfma32_tile<TIN, TOUT, TW, TH, NT>(src, sstride, W, H, dst, dstride, wts, wa,
                                       blockIdx.x * TW, blockIdx.y * TH, pa, pb, img);
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma32_batch(const BatchDesc* __restrict__ desc, int n, int total,
                                                         int* __restrict__ counter,
                                                         const float* __restrict__ wts, const float* __restrict__ wa)
{
    using C = Mfma32Tile<TW, TH, NT>;
    __shared__ float4 pa[C::PW * C::PH];
    __shared__ float4 pb[C::PW * C::PH];
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
        mfma32_tile<TIN, TOUT, TW, TH, NT>(d.src, d.sstride, d.w, d.h, d.dst, d.dstride, wts, wa,
                                           tx * TW, ty * TH, pa, pb, img);
        __syncthreads();
    }
}

// ---- launch table ---------------------------------------------------------------------

template <typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_one(const void* src, long ss, int W, int H, void* dst, long ds,
                             const float* wts, const float* wa, hipStream_t stream)
{
    dim3 grid((W + TW - 1) / TW, (H + TH - 1) / TH);
    hipLaunchKernelGGL((acnet_mfma32_fused<TIN, TOUT, TW, TH, NT>), grid, dim3(NT), 0, stream,
                       src, ss, W, H, dst, ds, wts, wa);
    return hipGetLastError();
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
static hipError_t launch_batch_one(const BatchDesc* desc, int n, int total, int* counter, const float* wts,
                                   const float* wa, hipStream_t stream)
{
    auto kern = acnet_mfma32_batch<TIN, TOUT, TW, TH, NT>;
    const int grid = persistent_grid(reinterpret_cast<const void*>(kern), NT, total);
    if (grid <= 0) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, stream, desc, n, total, counter, wts, wa);
    return hipGetLastError();
}

#define ACHIP_TYPES(CALL)                                                              \
    switch (stype) {                                                                   \
    case U8: switch (dtype) {                                                          \
        case U8: return CALL(uint8_t, uint8_t); case U16: return CALL(uint8_t, uint16_t);  \
        case F16: return CALL(uint8_t, __half); case F32: return CALL(uint8_t, float);     \
        default: return hipErrorInvalidValue; }                                        \
    case U16: switch (dtype) {                                                         \
        case U8: return CALL(uint16_t, uint8_t); case U16: return CALL(uint16_t, uint16_t); \
        case F16: return CALL(uint16_t, __half); case F32: return CALL(uint16_t, float);    \
        default: return hipErrorInvalidValue; }                                        \
    case F16: switch (dtype) {                                                         \
        case U8: return CALL(__half, uint8_t); case U16: return CALL(__half, uint16_t);     \
        case F16: return CALL(__half, __half); case F32: return CALL(__half, float);        \
        default: return hipErrorInvalidValue; }                                        \
    case F32: switch (dtype) {                                                         \
        case U8: return CALL(float, uint8_t); case U16: return CALL(float, uint16_t);       \
        case F16: return CALL(float, __half); case F32: return CALL(float, float);          \
        default: return hipErrorInvalidValue; }                                        \
    default: return hipErrorInvalidValue;                                              \
    }

hipError_t launch_mfma32(int stype, int dtype, const void* src, long ss, int W, int H,
                         void* dst, long ds, const float* wts, const float* wa, hipStream_t stream)
{
#define CALL(TI, TO) launch_one<TI, TO, 64, 32, 512>(src, ss, W, H, dst, ds, wts, wa, stream)
    ACHIP_TYPES(CALL)
#undef CALL
}

hipError_t launch_mfma32_batch(int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                               const float* wts, const float* wa, hipStream_t stream)
{
#define CALL(TI, TO) launch_batch_one<TI, TO, 64, 32, 512>(desc, n, total, counter, wts, wa, stream)
    ACHIP_TYPES(CALL)
#undef CALL
}
#undef ACHIP_TYPES

const int kMfma32LdsBytes = Mfma32Tile<64, 32, 512>::lds_bytes;

} // namespace achip
