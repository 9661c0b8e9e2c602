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
//
// Layers 0..7 take their items from the layer's region as one flat sequence: the w x np
// (row pair, column) positions of a region of w columns and np row pairs, in row-pair-major
// order, 16 consecutive positions per item.  An item may end one row pair and begin the
// next; every lane finds its own (row pair, column).  So a region of 66..80 columns costs
// ceil(w * np / 16) items rather than np * ceil(w / 16): 8.5% fewer MFMAs per 64 x 32 tile.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "acnet_mfma32_item.hpp"

namespace achip {

// ES: element size of the source type (the staged input tile's bytes per pixel)
template <int TW, int TH, int NT, int ES>
struct Mfma32Tile {
    static constexpr int NW = NT / 64;
    static constexpr int BW = TW + 16, BH = TH + 16;
    static constexpr int kSeg = (TW + 16 + 15) / 16;                // 16-pixel items per row pair, at most
    // padded so that 16-lane segments and row pairs may overhang a layer's region; the two
    // 16-lane groups of a half-wave read rows one pitch apart, and 16-byte LDS reads run at
    // full rate only when that is a multiple of 256 bytes (16 pixels) -- which fits the LDS
    // for inputs of up to 2 bytes per pixel
    static constexpr int PWmin = 8 + 16 * ((TW + 14 + 15) / 16) + 1;
    static constexpr int PW = ES <= 2 ? (PWmin + 15) / 16 * 16 : PWmin;
    static constexpr int PH = BH + 1;
    // input tile in the source element type, origin (x0 - 9, y0 - 9), padded the same way
    // for layer 0's items
    static constexpr int IW = 16 * kSeg + 2 + 8;
    static constexpr int IH = TH + 18 + 2;
    static constexpr int kStage = (IW * IH + NT - 1) / NT;          // input elements per thread
    static constexpr int items(int m) { return ((TW + 2 * m + 15) / 16) * ((TH + 2 * m + 1) / 2); }
    static constexpr int kMaxItems = (items(7) + NW - 1) / NW;      // items per wave, layers 1..7
    // with 512 threads (2 waves per SIMD) each wave loads its next item's operands before
    // the current item's MFMAs; with 1024 the other waves hide the latency and the
    // registers are needed elsewhere
    static constexpr bool kPrefetch = NT <= 512;
    static constexpr int lds_bytes(int elem) { return 2 * PW * PH * 16 + (IW * IH * elem + 15) / 16 * 16; }
};

// The items of one wave in a layer: wave, wave + NW, ... over (row pair, 16-pixel segment)
// in row-pair-major order, stepped without divisions.
struct ItemWalk {
    int pair, seg, dq, dr, nseg;
    __device__ ItemWalk(int wave, int nw, int nseg_) : nseg(nseg_)
    {
        pair = wave / nseg;
        seg = wave - pair * nseg;
        dq = nw / nseg;
        dr = nw - dq * nseg;
    }
    __device__ void next()
    {
        pair += dq;
        seg += dr;
        if (seg >= nseg) {
            seg -= nseg;
            pair++;
        }
    }
};

// A layer region of w columns x np row pairs as one sequence of n = w * np positions, 16 per
// item.  Lane j of item i sits at position min(16 i + j, n - 1): lanes past the end repeat the
// last position (their reads stay in the region) and write nothing.  The row pair is
// floor((p + 1/2) / w) in single precision, exact here: p < 2^13 and w <= 96, so the quotient
// lies at least 1/(2w) from an integer while its rounding error stays below 1e-5.
struct FlatRegion {
    int w, n;
    float inv_w;
    __device__ FlatRegion(int w_, int np) : w(w_), n(w_ * np), inv_w(1.0f / static_cast<float>(w_)) {}
    __device__ int items() const { return (n + 15) >> 4; }
    __device__ void locate(int item, int j, int& pair, int& col, bool& valid) const
    {
        const int raw = item * 16 + j;
        valid = raw < n;
        const int p = min(raw, n - 1);
        pair = static_cast<int>((static_cast<float>(p) + 0.5f) * inv_w);
        col = p - pair * w;
    }
};

// Replicate ring and zeroed border tiles: see fill_ring in acnet_mfma.hip.
template <int TW, int TH, int NT, int ES>
__device__ inline void fill_ring32(float4* __restrict__ pa, float4* __restrict__ pb, int x0, int y0, int W, int H, int m)
{
    using C = Mfma32Tile<TW, TH, NT, ES>;
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


template <typename TIN, typename TOUT, int TW, int TH, int NT>
__device__ __forceinline__ void mfma32_tile(const void* __restrict__ src, long sstride, int W, int H,
                                            void* __restrict__ dst, long dstride,
                                            const float* __restrict__ wts, const float* __restrict__ wa,
                                            int x0, int y0, float4* __restrict__ pa, float4* __restrict__ pb,
                                            TIN* __restrict__ img)
{
    using C = Mfma32Tile<TW, TH, NT, sizeof(TIN)>;
    // The thread index passes through an empty asm statement, so nothing derived from it
    // (LDS offsets, operand loads) is hoisted out of a persistent kernel's tile loop: held
    // across the loop, those values pushed the 1024-thread batch kernel past its 128 VGPRs
    // into scratch spills (12 VGPRs; 114 VGPRs and no spills with this).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int bx = x0 - 8, by = y0 - 8;
    const int j = lane & 15, g = lane >> 4, half = g & 1, ro = g >> 1;

    const bool border = x0 - 8 <= 0 || y0 - 8 <= 0 || x0 + TW + 8 >= W || y0 + TH + 8 >= H;
    if (border)
        for (int i = tid; i < C::PW * C::PH; i += NT) {
            pa[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            pb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    // input tile with replicate clamp over its whole padded extent, all loads in flight
    // before the first LDS store
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

    // Items (16 pixels x 2 rows) are dealt to the waves round-robin (ItemWalk).
    // this lane's write target: channels 0-3 (pa) or 4-7 (pb) of pixel (column j, row ro)
    float4* const pw = half ? pb : pa;

    // ---- layer 0: 3 MFMAs per item (K = 4 patch rows, one per patch column) ----
    {
        const int gx0 = max(x0 - 8, 0), gy0 = max(y0 - 8, 0);
        const int gx1 = min(x0 + TW + 8, W), gy1 = min(y0 + TH + 8, H);
        const FlatRegion fr(gx1 - gx0, (gy1 - gy0 + 1) >> 1);
        const int nitems = fr.items();
        const float a0 = wa[lane], a1 = wa[64 + lane], a2 = wa[128 + lane];
        const float* bl = wts + kOffB0 + 4 * half;
        const f32x4m bias = { bl[0], bl[1], bl[2], bl[3] };
        // img holds pixel (x, y) at ((y - y0 + 9) * IW + x - x0 + 9); B row y-1+g, column x+t-1
        const TIN* base = img + (gy0 - 1 - (y0 - 9) + g) * C::IW + (gx0 - 1 - (x0 - 9));
        for (int item = wave; item < nitems; item += C::NW) {
            int pair, col;
            bool valid;
            fr.locate(item, j, pair, col, valid);
            const TIN* p = base + 2 * pair * C::IW + col;
            f32x4m acc = mma32(a0, to_float<TIN>(p[0]), bias);
            acc = mma32(a1, to_float<TIN>(p[1]), acc);
            acc = mma32(a2, to_float<TIN>(p[2]), acc);
            const int oy = gy0 + 2 * pair + ro;
            if (valid && oy < gy1)
                pw[(oy - by) * C::PW + gx0 + col - bx] = make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f),
                                                                   fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
        }
    }
    // layer 1's operands
    float a[24];
    f32x4m bias;
    {
        const float* al = wa + kMfma32L0Floats + lane;
#pragma unroll
        for (int s = 0; s < 24; s++) a[s] = al[s * 64];
        const float* bl = wts + offB(1) + 4 * half;
        bias = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
    }
    __syncthreads();
    if (border) {
        fill_ring32<TW, TH, NT, sizeof(TIN)>(pa, pb, x0, y0, W, H, 8);
        __syncthreads();
    }

    // ---- layers 1..7, in place ----
#pragma nounroll
    for (int l = 1; l <= 7; l++) {
        const int m = 8 - l;
        const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
        const int gx1 = min(x0 + TW + m, W), gy1 = min(y0 + TH + m, H);
        const FlatRegion fr(gx1 - gx0, (gy1 - gy0 + 1) >> 1);
        const int nitems = fr.items();
        // B row y-1+g, column x+t-1 of the item's pixel (x, y)
        const int base0 = (gy0 - 1 - by + g) * C::PW + (gx0 - 1 - bx);
        auto item_base = [&](int item) {
            int pair, col;
            bool valid;
            fr.locate(item, j, pair, col, valid);
            return base0 + 2 * pair * C::PW + col;
        };
        float4 res[C::kMaxItems];
        {
            Item32 q;
            if constexpr (C::kPrefetch)
                if (wave < nitems) q = load_item(pa, pb, item_base(wave));
#pragma unroll
            for (int ii = 0; ii < C::kMaxItems; ii++) {
                const int item = wave + ii * C::NW;
                if (item < nitems) {
                    if constexpr (C::kPrefetch) {
                        Item32 nq;
                        if (item + C::NW < nitems) nq = load_item(pa, pb, item_base(item + C::NW));
                        const f32x4m acc = item24(q, a, bias);
                        res[ii] = make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f), fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
                        q = nq;
                    } else {
                        const f32x4m acc = item24(load_item(pa, pb, item_base(item)), a, bias);
                        res[ii] = make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f), fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
                    }
                }
            }
        }
        {   // operands of layer l + 1
            const float* al = wa + kMfma32L0Floats + l * kMfma32LayerFloats + lane;
#pragma unroll
            for (int s = 0; s < 24; s++) a[s] = al[s * 64];
            const float* bl = wts + offB(l + 1) + 4 * half;
            bias = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
        }
        __syncthreads();
#pragma unroll
        for (int ii = 0; ii < C::kMaxItems; ii++) {
            const int item = wave + ii * C::NW;
            if (item < nitems) {
                int pair, col;
                bool valid;
                fr.locate(item, j, pair, col, valid);
                const int oy = gy0 + 2 * pair + ro;
                if (valid && oy < gy1) pw[(oy - by) * C::PW + gx0 + col - bx] = res[ii];
            }
        }
        __syncthreads();
        if (border) {
            fill_ring32<TW, TH, NT, sizeof(TIN)>(pa, pb, x0, y0, W, H, m);
            __syncthreads();
        }
    }

    const int lane_rd = g * C::PW + j;
    // ---- layer 8, ReLU, deconv 2x2 stride 2 (8 -> 1) ----
    {
        const int gx1 = min(x0 + TW, W), gy1 = min(y0 + TH, H);
        const int nseg = (gx1 - x0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - y0 + 1) >> 1);
        const float* k2 = wts + kOffK2 + 16 * half;
        float kd[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) kd[c][q] = k2[c * 4 + q];
        // region origin is plane-local (8, 8)
        const int base0 = 7 * C::PW + 7 + lane_rd;
        ItemWalk it(wave, C::NW, nseg);
        for (int item = wave; item < nitems; item += C::NW, it.next()) {
            const f32x4m acc = item24(load_item(pa, pb, base0 + 2 * it.pair * C::PW + 16 * it.seg), a, bias);
            const float v0 = fmaxf(acc[0], 0.f), v1 = fmaxf(acc[1], 0.f);
            const float v2 = fmaxf(acc[2], 0.f), v3 = fmaxf(acc[3], 0.f);
            float sum[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++)
                sum[q4] = fmaf(v3, kd[3][q4], fmaf(v2, kd[2][q4], fmaf(v1, kd[1][q4], v0 * kd[0][q4])));
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) sum[q4] += __shfl_xor(sum[q4], 16);
            const int gx = x0 + 16 * it.seg + j;
            const int oy = y0 + 2 * it.pair + ro;
            if (gx < gx1 && oy < gy1) {
                TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * oy + half);
                store_pair<TOUT>(orow + 2 * gx, half ? sum[2] : sum[0], half ? sum[3] : sum[1]);
            }
        }
    }
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma32_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                         void* __restrict__ dst, long dstride,
                                                         const float* __restrict__ wts, const float* __restrict__ wa)
{
    using C = Mfma32Tile<TW, TH, NT, sizeof(TIN)>;
    __shared__ float4 pa[C::PW * C::PH];
    __shared__ float4 pb[C::PW * C::PH];
    __shared__ TIN img[C::IW * C::IH];
    const int t = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
    mfma32_tile<TIN, TOUT, TW, TH, NT>(src, sstride, W, H, dst, dstride, wts, wa, tx * TW, ty * TH, pa, pb, img);
}

template <typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma32_batch(const BatchDesc* __restrict__ desc, int n, int total,
                                                         int* __restrict__ counter,
                                                         const float* __restrict__ wts, const float* __restrict__ wa)
{
    using C = Mfma32Tile<TW, TH, NT, sizeof(TIN)>;
    __shared__ float4 pa[C::PW * C::PH];
    __shared__ float4 pb[C::PW * C::PH];
    __shared__ TIN img[C::IW * C::IH];
    __shared__ int s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    int t = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: keeps the tile state scalar
    while (t < total) {
        // The next tile is drawn as this one starts and read once it is done, so the atomic's
        // round trip overlaps the tile instead of leaving the CU idle between tiles (one
        // 1024-thread workgroup per CU).  Each workgroup still draws exactly one index past
        // its last tile.
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        const BatchDesc d = find_tile(desc, n, t);
        const int local = t - d.first_tile;
        const int ly = local / d.tiles_x;
        const int tx = local - ly * d.tiles_x;
        const int ty = d.ty0 + ly;
        mfma32_tile<TIN, TOUT, TW, TH, NT>(d.src, d.sstride, d.w, d.h, d.dst, d.dstride, wts, wa,
                                           tx * TW, ty * TH, pa, pb, img);
        __syncthreads();                    // the tile's LDS and s_tile are free again
        if (threadIdx.x == 0) s_tile = next;
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(s_tile);
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
                                   const float* wa, int max_wg, hipStream_t stream)
{
    auto kern = acnet_mfma32_batch<TIN, TOUT, TW, TH, NT>;
    const int grid = persistent_grid(reinterpret_cast<const void*>(kern), NT, total, max_wg);
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

hipError_t launch_mfma32(int cfg, int stype, int dtype, const void* src, long ss, int W, int H,
                         void* dst, long ds, const float* wts, const float* wa, hipStream_t stream)
{
    if (cfg == 0) {
#define CALL(TI, TO) launch_one<TI, TO, 64, 32, 1024>(src, ss, W, H, dst, ds, wts, wa, stream)
        ACHIP_TYPES(CALL)
#undef CALL
    }
#define CALL(TI, TO) launch_one<TI, TO, 64, 32, 512>(src, ss, W, H, dst, ds, wts, wa, stream)
    ACHIP_TYPES(CALL)
#undef CALL
}

hipError_t launch_mfma32_batch(int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                               const float* wts, const float* wa, int max_wg, hipStream_t stream)
{
    if (cfg == 0) {
#define CALL(TI, TO) launch_batch_one<TI, TO, 64, 32, 1024>(desc, n, total, counter, wts, wa, max_wg, stream)
        ACHIP_TYPES(CALL)
#undef CALL
    }
#define CALL(TI, TO) launch_batch_one<TI, TO, 64, 32, 512>(desc, n, total, counter, wts, wa, max_wg, stream)
    ACHIP_TYPES(CALL)
#undef CALL
}
#undef ACHIP_TYPES

// f32 input; less LDS for narrower types
const Fp32Config kMfma32Configs[] = {
    { 64, 32, 1024, 0, Mfma32Tile<64, 32, 1024, 4>::lds_bytes(4) },
    { 64, 32, 512, 0, Mfma32Tile<64, 32, 512, 4>::lds_bytes(4) },
};
const int kMfma32ConfigCount = sizeof(kMfma32Configs) / sizeof(kMfma32Configs[0]);

} // namespace achip
