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
    static constexpr int BW = TW + 16, BH = TH + 16;                // tile + margin 8
    // The plane is padded so that the 16-lane segments and 2-row pairs of a layer may
    // overhang its region (they read in bounds and are discarded): columns up to
    // 8 + 16 * ceil((TW + 14) / 16), rows up to BH.
    static constexpr int PW = 8 + 16 * ((TW + 14 + 15) / 16) + 1;
    static constexpr int PH = BH + 1;
    static constexpr int IW = TW + 18, IH = TH + 18;
    static constexpr int items(int m) { return ((TW + 2 * m + 15) / 16) * ((TH + 2 * m + 1) / 2); }
    static constexpr int kMaxItems = (items(7) + NW - 1) / NW;     // per wave, layers 1..7
    static constexpr int kL0Runs = 4;                               // layer-0 pixels per thread run
    static constexpr int lds_bytes = PW * PH * 16 + IW * IH * 4;
};

// Replicate ring.  After a layer of margin m has written its in-image region, copy the
// image's edge pixels one pixel outward wherever the margin-m band reaches past the image
// border.  The next layer then reads every tap without clamping and still sees the
// reference's per-layer replicate border (Common.hpp:122-143).  Interior tiles need none.
// (A row pair whose second row lies outside a cut region still reads input row y+2 for
// it, with zero weights for row y -- but 0 x NaN is NaN.  In a tile whose band touches the
// image border such a row may never be written, so those tiles zero the plane first;
// interior tiles only read rows their earlier layers wrote.)
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
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);    // wave-uniform
    const int bx = x0 - 8, by = y0 - 8;

    // ---- border tiles (the margin-8 band reaches the image edge) start from a zeroed
    // plane and keep a replicate ring (see fill_ring) ----
    const bool border = x0 - 8 <= 0 || y0 - 8 <= 0 || x0 + TW + 8 >= W || y0 + TH + 8 >= H;
    if (border)
        for (int i = tid; i < C::PW * C::PH; i += NT) plane[i] = make_uint4(0u, 0u, 0u, 0u);

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
                    plane[(ly + p) * C::PW + lx] = q;
                }
            }
        }
    }
    __syncthreads();

    if (border) {
        fill_ring<TW, TH, NT>(plane, x0, y0, W, H, 8);
        __syncthreads();
    }

    // ---- layers 1..8 on MFMA ----
    // Work item = 16 pixels x 2 rows of a layer's region, items numbered row pair by row
    // pair and dealt to the waves round-robin.  Per lane, B for patch column t is the pixel
    // (row gy - 1 + g, column x + j + t - 1): one scalar item base plus a per-lane constant,
    // and the three reads differ by 16 bytes.
    const int j = lane & 15;          // pixel column inside the item
    const int g = lane >> 4;          // lane group: patch row (B), D row block
    const int half = g & 1;           // channels 0-3 or 4-7 of D
    const int ro = g >> 1;            // output row y or y+1 of D
    const int lane_rd = g * C::PW + j;                         // B read, pixels
    const int lane_wr = (ro * C::PW + j) * 2 + half;           // write-back, 8-byte units
#pragma nounroll
    for (int l = 1; l <= 7; l++) {
        const int m = 8 - l;
        const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
        const int gx1 = min(x0 + TW + m, W), gy1 = min(y0 + TH + m, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        const uint4* al = wa + (l - 1) * 3 * 64 + lane;
        const v8 a0 = __builtin_bit_cast(v8, al[0]);
        const v8 a1 = __builtin_bit_cast(v8, al[64]);
        const v8 a2 = __builtin_bit_cast(v8, al[128]);
        const float* bl = wts + offB(l) + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];

        // items in pairs: one uniform branch per pair, so the two items' LDS reads and MFMA
        // chains share a scheduling region and overlap (a missing second item repeats the
        // first, and is not written back)
        uint2 res[C::kMaxItems + 1];
        int pair = wave / nseg, seg = wave - (wave / nseg) * nseg;   // item = wave
#pragma unroll
        for (int ii = 0; ii < C::kMaxItems; ii += 2) {
            if (wave + ii * C::NW < nitems) {
                int pb = pair, sb = seg + C::NW;
                while (sb >= nseg) { sb -= nseg; pb++; }
                const int baseA = (gy0 + 2 * pair - 1 - by) * C::PW + (gx0 + 16 * seg - 1 - bx);
                const int baseB = (wave + (ii + 1) * C::NW < nitems)
                                      ? (gy0 + 2 * pb - 1 - by) * C::PW + (gx0 + 16 * sb - 1 - bx) : baseA;
                const uint4* pA = plane + baseA + lane_rd;
                const uint4* pB = plane + baseB + lane_rd;
                const v8 qa0 = __builtin_bit_cast(v8, pA[0]), qb0 = __builtin_bit_cast(v8, pB[0]);
                const v8 qa1 = __builtin_bit_cast(v8, pA[1]), qb1 = __builtin_bit_cast(v8, pB[1]);
                const v8 qa2 = __builtin_bit_cast(v8, pA[2]), qb2 = __builtin_bit_cast(v8, pB[2]);
                f32x4 ca = { 0.f, 0.f, 0.f, 0.f }, cb = { 0.f, 0.f, 0.f, 0.f };
                ca = M::mma(a0, qa0, ca);
                cb = M::mma(a0, qb0, cb);
                ca = M::mma(a1, qa1, ca);
                cb = M::mma(a1, qb1, cb);
                ca = M::mma(a2, qa2, ca);
                cb = M::mma(a2, qb2, cb);
                res[ii].x = M::pack2(fmaxf(ca[0] + b0, 0.f), fmaxf(ca[1] + b1, 0.f));
                res[ii].y = M::pack2(fmaxf(ca[2] + b2, 0.f), fmaxf(ca[3] + b3, 0.f));
                res[ii + 1].x = M::pack2(fmaxf(cb[0] + b0, 0.f), fmaxf(cb[1] + b1, 0.f));
                res[ii + 1].y = M::pack2(fmaxf(cb[2] + b2, 0.f), fmaxf(cb[3] + b3, 0.f));
            }
#pragma unroll
            for (int k = 0; k < 2; k++) {
                seg += C::NW;
                while (seg >= nseg) { seg -= nseg; pair++; }
            }
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
                    const int base = ((gy0 + 2 * pair - by) * C::PW + (gx0 + 16 * seg - bx)) * 2;
                    reinterpret_cast<uint2*>(plane)[base + lane_wr] = res[ii];
                }
            }
            seg += C::NW;
            while (seg >= nseg) { seg -= nseg; pair++; }
        }
        __syncthreads();
        if (border) {
            fill_ring<TW, TH, NT>(plane, x0, y0, W, H, m);
            __syncthreads();
        }
    }

    // ---- layer 8 on MFMA, ReLU, then the deconv 2x2 stride 2 (8 -> 1) in FP32 ----
    {
        const int gx0 = x0, gy0 = y0;
        const int gx1 = min(x0 + TW, W), gy1 = min(y0 + TH, H);
        const int nseg = (gx1 - gx0 + 15) >> 4;
        const int nitems = nseg * ((gy1 - gy0 + 1) >> 1);
        const uint4* al = wa + 7 * 3 * 64 + lane;
        const v8 a0 = __builtin_bit_cast(v8, al[0]);
        const v8 a1 = __builtin_bit_cast(v8, al[64]);
        const v8 a2 = __builtin_bit_cast(v8, al[128]);
        const float* bl = wts + offB(8) + 4 * half;
        const float b0 = bl[0], b1 = bl[1], b2 = bl[2], b3 = bl[3];
        // this lane's 4 channels x 4 output sub-pixels of the deconv, k2[c][q]
        const float* k2 = wts + kOffK2 + 16 * half;
        float kd[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) kd[c][q] = k2[c * 4 + q];
        int pair = wave / nseg, seg = wave - (wave / nseg) * nseg;
        for (int item = wave; item < nitems; item += 2 * C::NW) {
            int pb = pair, sb = seg + C::NW;
            while (sb >= nseg) { sb -= nseg; pb++; }
            const bool hasB = item + C::NW < nitems;
            const int baseA = (gy0 + 2 * pair - 1 - by) * C::PW + (gx0 + 16 * seg - 1 - bx);
            const int baseB = hasB ? (gy0 + 2 * pb - 1 - by) * C::PW + (gx0 + 16 * sb - 1 - bx) : baseA;
            const uint4* pA = plane + baseA + lane_rd;
            const uint4* pB = plane + baseB + lane_rd;
            const v8 qa0 = __builtin_bit_cast(v8, pA[0]), qb0 = __builtin_bit_cast(v8, pB[0]);
            const v8 qa1 = __builtin_bit_cast(v8, pA[1]), qb1 = __builtin_bit_cast(v8, pB[1]);
            const v8 qa2 = __builtin_bit_cast(v8, pA[2]), qb2 = __builtin_bit_cast(v8, pB[2]);
            f32x4 acc[2] = { { 0.f, 0.f, 0.f, 0.f }, { 0.f, 0.f, 0.f, 0.f } };
            acc[0] = M::mma(a0, qa0, acc[0]);
            acc[1] = M::mma(a0, qb0, acc[1]);
            acc[0] = M::mma(a1, qa1, acc[0]);
            acc[1] = M::mma(a1, qb1, acc[1]);
            acc[0] = M::mma(a2, qa2, acc[0]);
            acc[1] = M::mma(a2, qb2, acc[1]);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const float v0 = fmaxf(acc[u][0] + b0, 0.f), v1 = fmaxf(acc[u][1] + b1, 0.f);
                const float v2 = fmaxf(acc[u][2] + b2, 0.f), v3 = fmaxf(acc[u][3] + b3, 0.f);
                // this lane's 4 channels, then the partner lane (lane ^ 16) holds the other 4
                float s[4];
#pragma unroll
                for (int q = 0; q < 4; q++) s[q] = fmaf(v3, kd[3][q], fmaf(v2, kd[2][q], fmaf(v1, kd[1][q], v0 * kd[0][q])));
#pragma unroll
                for (int q = 0; q < 4; q++) s[q] += __shfl_xor(s[q], 16);
                const int gx = gx0 + 16 * (u ? sb : seg) + j;
                const int oy = gy0 + 2 * (u ? pb : pair) + ro;
                if ((u == 0 || hasB) && gx < gx1 && oy < gy1) {
                    TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * oy + half);
                    store_pair<TOUT>(orow + 2 * gx, half ? s[2] : s[0], half ? s[3] : s[1]);
                }
            }
#pragma unroll
            for (int k = 0; k < 2; k++) {
                seg += C::NW;
                while (seg >= nseg) { seg -= nseg; pair++; }
            }
        }
    }
}

template <typename T16, typename TIN, typename TOUT, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void acnet_mfma_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                       void* __restrict__ dst, long dstride,
                                                       const float* __restrict__ wts, const uint4* __restrict__ wa)
{
    using C = MfmaTile<TW, TH, NT>;
    __shared__ uint4 plane[C::PW * C::PH];        // 8 x 16-bit channels per pixel
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
    __shared__ uint4 plane[C::PW * C::PH];
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
