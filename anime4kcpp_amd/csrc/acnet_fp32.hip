// Fused FP32 ACNetLegacy 2x pass for gfx950: one launch runs the whole network.
//
// The reference computes the pass layer by layer over the full image
// (core/src/processor/cpu/CPUProcessor.cpp:1371-1390; CUDA: 9 launches with FP16
// intermediates in HBM, core/src/processor/cuda/CUDAProcessor.cpp:406-413).  Here a
// workgroup owns one TW x TH tile of input-resolution pixels and never writes a
// feature map to HBM:
//   - the input tile plus a 9-pixel halo is staged once in LDS as f32 (toFloat applied),
//     with coordinates clamped to the image, which is the replicate border of layer 0;
//   - the 8-channel f32 features of layers 0..7 live in one LDS plane of
//     (TW+16) x (TH+16) pixels, stored as two float4 planes (channels 0-3 and 4-7) so
//     consecutive lanes read consecutive 16-byte words;
//   - layer l (margin m = 8 - l) is computed over the tile grown by m pixels, cut to the
//     image.  Every layer clamps its own taps to the image (Common.hpp:122-143), so a
//     tile at an image edge gives the same values as the full-image reference;
//   - layers 1..7 update the plane in place: results go to registers, a barrier, then
//     the write-back;
//   - layer 8, the ReLU and the 2x2 stride-2 deconv (Common.hpp:344-393) run in the
//     epilogue and store the 2x2 output block of each pixel straight to HBM.
// Each thread computes a vertical run of P pixels: the (P+2) x 3 input taps are read
// from LDS once, and each weight (a uniform scalar load) feeds P FMAs.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

template <int TW, int TH, int NT, int P>
struct Fp32Tile {
    static constexpr int BW = TW + 16, BH = TH + 16;     // feature plane, margin 8
    static constexpr int IW = TW + 18, IH = TH + 18;     // input tile, margin 9
    static constexpr int NPIX = BW * BH;
    static constexpr int items(int m) { return (TW + 2 * m) * ((TH + 2 * m + P - 1) / P); }
    static constexpr int rounds(int m) { return (items(m) + NT - 1) / NT; }
    static constexpr int kMidRounds = rounds(7);
    static constexpr int lds_bytes = 2 * NPIX * 16 + IW * IH * 4;
};

struct Region {
    int gx0, gy0, cols, nitems;
    int gy1;   // exclusive end row in the image
};

template <int TW, int TH, int P>
__device__ inline Region layer_region(int x0, int y0, int W, int H, int m)
{
    Region r;
    r.gx0 = max(x0 - m, 0);
    r.gy0 = max(y0 - m, 0);
    const int gx1 = min(x0 + TW + m, W);
    r.gy1 = min(y0 + TH + m, H);
    r.cols = gx1 - r.gx0;
    const int rows = r.gy1 - r.gy0;
    r.nitems = r.cols * ((rows + P - 1) / P);
    return r;
}

// One middle layer (8 -> 8, ReLU) for a vertical run of P pixels at buffer column lx,
// rows starting at buffer row ly (global gx, gy).  Reads the plane, returns results.
template <int TW, int TH, int NT, int P>
__device__ __forceinline__ void mid_item(const float4* __restrict__ fa, const float4* __restrict__ fb,
                                         const float* __restrict__ wl, const float* __restrict__ bl,
                                         int gx, int gy, int lx, int ly, int W, int H, int by,
                                         float (&out)[P][8])
{
    using C = Fp32Tile<TW, TH, NT, P>;
    // taps clamped to the image, then to the buffer (rows past a partial run are
    // read from inside the buffer and discarded)
    const int cx[3] = { gx > 0 ? lx - 1 : lx, lx, gx < W - 1 ? lx + 1 : lx };
    float v[P + 2][3][8];
#pragma unroll
    for (int r = 0; r < P + 2; r++) {
        int row = clampi(gy - 1 + r, 0, H - 1) - by;
        row = row < C::BH - 1 ? row : C::BH - 1;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int idx = row * C::BW + cx[k];
            const float4 a = fa[idx];
            const float4 b = fb[idx];
            v[r][k][0] = a.x; v[r][k][1] = a.y; v[r][k][2] = a.z; v[r][k][3] = a.w;
            v[r][k][4] = b.x; v[r][k][5] = b.y; v[r][k][6] = b.z; v[r][k][7] = b.w;
        }
    }
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int n = 0; n < 8; n++) out[p][n] = bl[n];
#pragma unroll
    for (int ky = 0; ky < 3; ky++)
#pragma unroll
        for (int kx = 0; kx < 3; kx++)
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    const float w = wl[((ky * 3 + kx) * 8 + c) * 8 + n];
#pragma unroll
                    for (int p = 0; p < P; p++) out[p][n] = fmaf(v[p + ky][kx][c], w, out[p][n]);
                }
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int n = 0; n < 8; n++) out[p][n] = fmaxf(out[p][n], 0.0f);
}

// The whole network on one tile whose top-left input pixel is (x0, y0).  fa/fb/img are the
// workgroup's LDS arrays.  Ends without a barrier: a caller reusing the LDS for another
// tile synchronises first.
template <typename TIN, typename TOUT, int TW, int TH, int NT, int P>
__device__ __forceinline__ void fp32_tile(const void* __restrict__ src, long sstride, int W, int H,
                                          void* __restrict__ dst, long dstride,
                                          const float* __restrict__ wts, int zero, int x0, int y0,
                                          float4* __restrict__ fa, float4* __restrict__ fb, float* __restrict__ img)
{
    using C = Fp32Tile<TW, TH, NT, P>;
    const int tid = threadIdx.x;
    const int bx = x0 - 8, by = y0 - 8;      // buffer origin; input tile origin is (bx-1, by-1)

    // ---- stage the input tile: toFloat, replicate clamp ----
    for (int i = tid; i < C::IW * C::IH; i += NT) {
        const int iy = i / C::IW;
        const int ix = i - iy * C::IW;
        const int gy = clampi(y0 - 9 + iy, 0, H - 1);
        const int gx = clampi(x0 - 9 + ix, 0, W - 1);
        img[i] = to_float<TIN>(row_ptr<TIN>(src, sstride, gy)[gx]);
    }
    __syncthreads();

    // ---- layer 0: conv3x3 1 -> 8 + ReLU over margin 8 ----
    {
        const Region R = layer_region<TW, TH, P>(x0, y0, W, H, 8);
#pragma unroll
        for (int rr = 0; rr < C::rounds(8); rr++) {
            const int item = tid + rr * NT;
            if (item < R.nitems) {
                const int run = item / R.cols;
                const int col = item - run * R.cols;
                const int gx = R.gx0 + col, gy = R.gy0 + run * P;
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
                    if (gy + p < R.gy1) {
                        const int idx = (ly + p) * C::BW + lx;
                        fa[idx] = make_float4(fmaxf(acc[p][0], 0.f), fmaxf(acc[p][1], 0.f), fmaxf(acc[p][2], 0.f), fmaxf(acc[p][3], 0.f));
                        fb[idx] = make_float4(fmaxf(acc[p][4], 0.f), fmaxf(acc[p][5], 0.f), fmaxf(acc[p][6], 0.f), fmaxf(acc[p][7], 0.f));
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- layers 1..8: conv3x3 8 -> 8 + ReLU.  Layers 1..7 update the plane in place;
    // layer 8 feeds the deconv epilogue.  One runtime loop for all eight keeps the
    // weight reads runtime-addressed, so they stream through two scalar buffers
    // instead of being hoisted together.
#pragma nounroll
    for (int l = 1; l <= 8; l++) {
        const int m = 8 - l;
        const Region R = layer_region<TW, TH, P>(x0, y0, W, H, m);
        const float* wl = wts + offW(l);
        const float* bl = wts + offB(l);
        float res[C::kMidRounds][P][8];
#pragma unroll
        for (int rr = 0; rr < C::kMidRounds; rr++) {
            const int item = tid + rr * NT;
            if (item < R.nitems) {
                const int run = item / R.cols;
                const int col = item - run * R.cols;
                const int gx = R.gx0 + col, gy = R.gy0 + run * P;
                // each round reads its weights at an offset the compiler cannot see is zero,
                // so the rounds do not share (and keep live) 584 scalar weight registers
                const float* wr = wl + rr * zero;
                const float* br = bl + rr * zero;
                mid_item<TW, TH, NT, P>(fa, fb, wr, br, gx, gy, gx - bx, gy - by, W, H, by, res[rr]);
            }
            // keep the rounds apart: overlapping them would need two sets of taps live
            __builtin_amdgcn_sched_barrier(0);
        }
        if (l < 8) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < C::kMidRounds; rr++) {
                const int item = tid + rr * NT;
                if (item < R.nitems) {
                    const int run = item / R.cols;
                    const int col = item - run * R.cols;
                    const int gx = R.gx0 + col, gy = R.gy0 + run * P;
                    const int lx = gx - bx, ly = gy - by;
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        if (gy + p < R.gy1) {
                            const int idx = (ly + p) * C::BW + lx;
                            fa[idx] = make_float4(res[rr][p][0], res[rr][p][1], res[rr][p][2], res[rr][p][3]);
                            fb[idx] = make_float4(res[rr][p][4], res[rr][p][5], res[rr][p][6], res[rr][p][7]);
                        }
                    }
                }
            }
            __syncthreads();
        } else {
            // ReLU'd layer-8 features -> deconv 2x2 stride 2 (8 -> 1) -> fromFloat -> HBM
            const float* k2 = wts + kOffK2;
#pragma unroll
            for (int rr = 0; rr < C::kMidRounds; rr++) {
                const int item = tid + rr * NT;
                if (item < R.nitems) {
                    const int run = item / R.cols;
                    const int col = item - run * R.cols;
                    const int gx = R.gx0 + col, gy = R.gy0 + run * P;
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        if (gy + p < R.gy1) {
#pragma unroll
                            for (int dy = 0; dy < 2; dy++) {
                                float o[2];
#pragma unroll
                                for (int dx = 0; dx < 2; dx++) {
                                    float acc = 0.0f;
#pragma unroll
                                    for (int c = 0; c < 8; c++) acc = fmaf(res[rr][p][c], k2[c * 4 + dy * 2 + dx], acc);
                                    o[dx] = acc;
                                }
                                TOUT* orow = row_ptr<TOUT>(dst, dstride, 2 * (gy + p) + dy);
                                store_pair<TOUT>(orow + 2 * gx, o[0], o[1]);
                            }
                        }
                    }
                }
            }
        }
    }
}

template <typename TIN, typename TOUT, int TW, int TH, int NT, int P>
__global__ __launch_bounds__(NT) void acnet_fp32_fused(const void* __restrict__ src, long sstride, int W, int H,
                                                       void* __restrict__ dst, long dstride,
                                                       const float* __restrict__ wts, int zero)
{
    using C = Fp32Tile<TW, TH, NT, P>;
    __shared__ float4 fa[C::NPIX];
    __shared__ float4 fb[C::NPIX];
    __shared__ float img[C::IW * C::IH];
    const int t = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
    fp32_tile<TIN, TOUT, TW, TH, NT, P>(src, sstride, W, H, dst, dstride, wts, zero, tx * TW, ty * TH, fa, fb, img);
}

// Persistent batched form: every workgroup takes the next tile of the batch from a
// counter in global memory until the batch is done; tiles are numbered image by image
// (BatchDesc::first_tile is the prefix sum).  Every workgroup reaches the exit when
// the counter passes `total`.
template <typename TIN, typename TOUT, int TW, int TH, int NT, int P>
__global__ __launch_bounds__(NT) void acnet_fp32_batch(const BatchDesc* __restrict__ desc, int n, int total,
                                                       int* __restrict__ counter,
                                                       const float* __restrict__ wts, int zero)
{
    using C = Fp32Tile<TW, TH, NT, P>;
    __shared__ float4 fa[C::NPIX];
    __shared__ float4 fb[C::NPIX];
    __shared__ float img[C::IW * C::IH];
    __shared__ int s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: keeps the tile state scalar
        if (t >= total) break;
        const BatchDesc d = find_tile(desc, n, t);
        const int local = t - d.first_tile;
        const int ly = local / d.tiles_x;
        const int tx = local - ly * d.tiles_x;
        const int ty = d.ty0 + ly;
        fp32_tile<TIN, TOUT, TW, TH, NT, P>(d.src, d.sstride, d.w, d.h, d.dst, d.dstride, wts, zero,
                                            tx * TW, ty * TH, fa, fb, img);
        __syncthreads();
    }
}

// ---- launch table ---------------------------------------------------------------------

template <typename TIN, typename TOUT, int TW, int TH, int NT, int P>
static hipError_t launch_one(const void* src, long ss, int W, int H, void* dst, long ds,
                             const float* wts, hipStream_t stream)
{
    dim3 grid((W + TW - 1) / TW, (H + TH - 1) / TH);
    hipLaunchKernelGGL((acnet_fp32_fused<TIN, TOUT, TW, TH, NT, P>), grid, dim3(NT), 0, stream,
                       src, ss, W, H, dst, ds, wts, 0);
    return hipGetLastError();
}

template <typename TIN, typename TOUT, int TW, int TH, int NT, int P>
static hipError_t launch_batch_one(const BatchDesc* desc, int n, int total, int* counter, const float* wts, int max_wg,
                                   hipStream_t stream)
{
    auto kern = acnet_fp32_batch<TIN, TOUT, TW, TH, NT, P>;
    const int grid = persistent_grid(reinterpret_cast<const void*>(kern), NT, total, max_wg);
    if (grid <= 0) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, stream, desc, n, total, counter, wts, 0);
    return hipGetLastError();
}

template <int TW, int TH, int NT, int P>
static hipError_t launch_batch_types(int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                                     const float* wts, int max_wg, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                          \
    switch (dtype) {                                                                                            \
    case U8: return launch_batch_one<TIN, uint8_t, TW, TH, NT, P>(desc, n, total, counter, wts, max_wg, stream);        \
    case U16: return launch_batch_one<TIN, uint16_t, TW, TH, NT, P>(desc, n, total, counter, wts, max_wg, stream);      \
    case F16: return launch_batch_one<TIN, __half, TW, TH, NT, P>(desc, n, total, counter, wts, max_wg, stream);        \
    case F32: return launch_batch_one<TIN, float, TW, TH, NT, P>(desc, n, total, counter, wts, max_wg, stream);         \
    default: return hipErrorInvalidValue;                                                                       \
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

template <int TW, int TH, int NT, int P>
static hipError_t launch_types(int stype, int dtype, const void* src, long ss, int W, int H,
                               void* dst, long ds, const float* wts, hipStream_t stream)
{
#define ACHIP_OUT(TIN)                                                                                 \
    switch (dtype) {                                                                                   \
    case U8: return launch_one<TIN, uint8_t, TW, TH, NT, P>(src, ss, W, H, dst, ds, wts, stream);      \
    case U16: return launch_one<TIN, uint16_t, TW, TH, NT, P>(src, ss, W, H, dst, ds, wts, stream);    \
    case F16: return launch_one<TIN, __half, TW, TH, NT, P>(src, ss, W, H, dst, ds, wts, stream);      \
    case F32: return launch_one<TIN, float, TW, TH, NT, P>(src, ss, W, H, dst, ds, wts, stream);       \
    default: return hipErrorInvalidValue;                                                              \
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

const Fp32Config kFp32Configs[] = {
    { 64, 32, 512, 4, Fp32Tile<64, 32, 512, 4>::lds_bytes },
    { 32, 32, 256, 4, Fp32Tile<32, 32, 256, 4>::lds_bytes },
    { 16, 16, 256, 4, Fp32Tile<16, 16, 256, 4>::lds_bytes },   // the small end of the tile sweep
};
const int kFp32ConfigCount = sizeof(kFp32Configs) / sizeof(kFp32Configs[0]);

hipError_t launch_fp32_batch(int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                             const float* wts, int max_wg, hipStream_t stream)
{
    switch (cfg) {
    case 0: return launch_batch_types<64, 32, 512, 4>(stype, dtype, desc, n, total, counter, wts, max_wg, stream);
    case 1: return launch_batch_types<32, 32, 256, 4>(stype, dtype, desc, n, total, counter, wts, max_wg, stream);
    case 2: return launch_batch_types<16, 16, 256, 4>(stype, dtype, desc, n, total, counter, wts, max_wg, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fp32(int cfg, int stype, int dtype, const void* src, long ss, int W, int H,
                       void* dst, long ds, const float* wts, hipStream_t stream)
{
    switch (cfg) {
    case 0: return launch_types<64, 32, 512, 4>(stype, dtype, src, ss, W, H, dst, ds, wts, stream);
    case 1: return launch_types<32, 32, 256, 4>(stype, dtype, src, ss, W, H, dst, ds, wts, stream);
    case 2: return launch_types<16, 16, 256, 4>(stype, dtype, src, ss, W, H, dst, ds, wts, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace achip
