// The FP32 MFMA conv3x3 layer kernels of the layer-by-layer networks (ACNet<8>, ArtCNN<F>,
// FSRCNNX<F>): one launch per layer (two for conv8_pair), reading an F-channel f32 map from
// HBM (channel quads interleaved per pixel: pixel x at float4 index F/4 * x + q, rows of `fp`
// bytes), and writing the next map or, for the last layer, the 2x output image.  The layer's epilogue is
// chosen at compile time (LayerMode); every form keeps the reference's separately rounded
// steps after the conv sum (Common.hpp:116-164, 227-288, 290-341):
//   kRelu          max(v, 0)                                        ArtCNN body
//   kPrelu         max(v, 0) + a * min(v, 0)                        ACNet<8>, FSRCNNX body
//   kAddFeat       v * 1 + feat                                     ArtCNN's last F -> F
//   kPreluConv1x1  PReLU, then conv1x1 F -> F + bias, * 1 + feat, PReLU
//                                                                   FSRCNNX's last block
//   kShuffle       conv F -> 4, pixel shuffle to 2x                 ArtCNN, FSRCNNX tail
//   kShuffleSrc    the same, + the source pixel                     ACNet<8> final
//   kReluDeconv    ReLU, then the 2x2 stride-2 deconv 8 -> 1 to 2x  ACNetLegacy's last layer
//                  (Common.hpp:344-393), layer by layer
// Two MFMA mappings (v_mfma_f32_16x16x4_f32):
//   conv8_layer, conv8_pair  F = 8: the 2-row item of acnet_mfma32_item.hpp, 16 pixels x 2
//                output rows per item, all 16 MFMA rows useful;
//   convf_layer  F = 16 or 32: M = 16 output channels per tile (F/16 tiles), N = 16 pixels of
//                one row, K = 9 taps x F; lane group g supplies input channels g*F/4 ..
//                g*F/4 + F/4 - 1 from channel-quad LDS planes; NI items per wave batch share
//                each tap's A operands.
// In both, a lane's D registers hold 4 consecutive channels of one pixel, which is exactly
// the B operand a conv1x1 MFMA needs (K step r of lane group g = the r-th of them), so the
// fused conv1x1 costs F/4 x F/16 more MFMAs and no data movement.
#pragma once

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"
#include "acnet_mfma32_item.hpp"

namespace achip {

#pragma clang fp contract(off)

enum LayerMode : int { kRelu = 0, kPrelu = 1, kAddFeat = 2, kPreluConv1x1 = 3, kShuffle = 4, kShuffleSrc = 5,
                       kReluDeconv = 6 };

// Arguments of one layer launch; a mode leaves the fields it does not read null.
struct LayerArgs {
    const float* wa;        // this layer's per-lane A operands
    const float* bias;      // F biases (4 for the shuffle modes)
    const float* alpha;     // F PReLU slopes
    const float* w1;        // conv1x1 A operands
    const float* b1;        // conv1x1 biases (F)
    const float* alpha1;    // F PReLU slopes after the skip connection
    const float4* feat;     // skip-connection map, same pitch as the input map
    const float* k2;        // deconv weights k2[c][q] (kReluDeconv)
    const void* src;        // source image (kShuffleSrc)
    long ss;
    void* dst;              // output image (shuffle modes)
    long ds;
};

// PReLU (Common.hpp:42-49): max(v, 0) + alpha * min(v, 0), each step rounded
__device__ inline float prelu(float v, float alpha)
{
    return __fadd_rn(fmaxf(v, 0.0f), __fmul_rn(alpha, fminf(v, 0.0f)));
}

__device__ inline float add_skip(float v, float f) { return __fadd_rn(__fmul_rn(v, 1.0f), f); }

// Stage ROWS x CW pixels of an f32 map (NQ channel quads per pixel, rows of fp bytes) whose
// top-left is image pixel (cx0, ry0), clamped to the image, into LDS: quad q of staged pixel
// (c, r) goes to planes[q][r * PW + c].  Each thread issues all of its loads before its first
// LDS store, so a workgroup waits for one load latency rather than one per pass over the tile.
template <int NT, int CW, int ROWS, int NQ, int PW>
__device__ __forceinline__ void stage_map(const float4* __restrict__ fin, long fp, int W, int H, int cx0, int ry0,
                                          float4* const* planes)
{
    constexpr int N = ROWS * CW, K = (N + NT - 1) / NT;
    float4 v[K][NQ];
#pragma unroll
    for (int k = 0; k < K; k++) {
        // threads past the last pixel load it again (no branch around the loads) and skip the store
        const int e = min(static_cast<int>(threadIdx.x) + k * NT, N - 1);
        const int r = e / CW, c = e - r * CW;
        const int sy = clampi(ry0 + r, 0, H - 1), sx = clampi(cx0 + c, 0, W - 1);
        const float4* px = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(fin) + static_cast<long>(sy) * fp) + NQ * sx;
#pragma unroll
        for (int q = 0; q < NQ; q++) v[k][q] = px[q];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = threadIdx.x + k * NT;
        if (k < K - 1 || e < N) {
            const int r = e / CW, c = e - r * CW;
#pragma unroll
            for (int q = 0; q < NQ; q++) planes[q][r * PW + c] = v[k][q];
        }
    }
}

// ---- F = 8: the 2-row item ----
// Tile: TW x TH output pixels, NT threads; the input region (with the one-pixel halo) staged
// in LDS as two float4 planes (channels 0-3, 4-7) with a row pitch that is a multiple of 16
// pixels (the 16-byte item reads run at full rate).  Items: 16-pixel segments x row pairs,
// dealt to the waves round-robin.  Default 64 x 16, 256 threads: 46 KB, three workgroups
// per CU.
template <int TW_ = 64, int TH_ = 16, int NT_ = 256>
struct Conv8Tile {
    static constexpr int TW = TW_, TH = TH_, NT = NT_, NW = NT / 64;
    static constexpr int PW = (TW + 2 + 15) / 16 * 16, PH = TH + 2;
};

// The epilogue of an 8-channel item: this lane's bias, slopes and conv1x1 operands (lane
// (g, j) owns channels 4 half .. + 3 of output row ro of pixel j), and what it does with the
// item's D registers.  `ok`: the lane's pixel (x, y) is an output of this tile.
template <int MODE, typename TIN, typename TOUT>
struct Epi8 {
    static_assert(MODE == kRelu || MODE == kPrelu || MODE == kPreluConv1x1 || MODE == kShuffle || MODE == kShuffleSrc ||
                      MODE == kReluDeconv,
                  "8-channel epilogues");
    static constexpr bool SHUFFLE = MODE == kShuffle || MODE == kShuffleSrc;
    f32x4m bias, bias1 = { 0.f, 0.f, 0.f, 0.f };
    float alpha[4] = { 0.f, 0.f, 0.f, 0.f }, alpha1[4] = { 0.f, 0.f, 0.f, 0.f }, w1[4] = { 0.f, 0.f, 0.f, 0.f };
    float kd[4][4];   // kReluDeconv: this lane's 4 channels x 4 sub-pixels of k2

    __device__ void load(const LayerArgs& la, int lane, int half)
    {
        if constexpr (MODE == kRelu || MODE == kReluDeconv) {
            const float* bl = la.bias + 4 * half;
            bias = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
            if constexpr (MODE == kReluDeconv)
#pragma unroll
                for (int c = 0; c < 4; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++) kd[c][q] = la.k2[(4 * half + c) * 4 + q];
        } else if constexpr (SHUFFLE) {
            bias = half ? (f32x4m){ 0.f, 0.f, 0.f, 0.f } : (f32x4m){ la.bias[0], la.bias[1], la.bias[2], la.bias[3] };
        } else {
            const float* bl = la.bias + 4 * half;
            bias = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
#pragma unroll
            for (int r = 0; r < 4; r++) alpha[r] = la.alpha[4 * half + r];
        }
        if constexpr (MODE == kPreluConv1x1) {
            const float* bl = la.b1 + 4 * half;
            bias1 = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
#pragma unroll
            for (int r = 0; r < 4; r++) {
                alpha1[r] = la.alpha1[4 * half + r];
                w1[r] = la.w1[r * 64 + lane];
            }
        }
    }

    // every lane of the wave calls this (the conv1x1 MFMAs and the shuffle take all lanes)
    __device__ void apply(const f32x4m& acc, int x, int y, bool ok, int half, long fp, float4* __restrict__ fout,
                          const LayerArgs& la) const
    {
        if constexpr (MODE == kRelu) {
            if (ok) {
                float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + static_cast<long>(y) * fp) + 2 * x + half;
                *px = make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f), fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
            }
        } else if constexpr (MODE == kReluDeconv) {
            // this lane's 4 channels' share of the 4 sub-pixels, plus the partner lane's
            // (lane ^ 16 holds the other 4 channels of the same pixel and row); half = 0
            // lanes write output row 2y (q = 0, 1), half = 1 lanes row 2y + 1 (q = 2, 3)
            const float v0 = fmaxf(acc[0], 0.f), v1 = fmaxf(acc[1], 0.f), v2 = fmaxf(acc[2], 0.f), v3 = fmaxf(acc[3], 0.f);
            float sum[4];
#pragma unroll
            for (int q = 0; q < 4; q++) sum[q] = fmaf(v3, kd[3][q], fmaf(v2, kd[2][q], fmaf(v1, kd[1][q], v0 * kd[0][q])));
#pragma unroll
            for (int q = 0; q < 4; q++) sum[q] += __shfl_xor(sum[q], 16);
            if (ok) store_pair<TOUT>(row_ptr<TOUT>(la.dst, la.ds, 2 * y + half) + 2 * x, half ? sum[2] : sum[0], half ? sum[3] : sum[1]);
        } else if constexpr (MODE == kPrelu) {
            if (ok) {
                float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + static_cast<long>(y) * fp) + 2 * x + half;
                *px = make_float4(prelu(acc[0], alpha[0]), prelu(acc[1], alpha[1]), prelu(acc[2], alpha[2]),
                                  prelu(acc[3], alpha[3]));
            }
        } else if constexpr (MODE == kPreluConv1x1) {
            // conv1x1 over (output row, channel): K step r of lane group g is channel
            // 4 half + r of row ro -- this lane's own D register r
            f32x4m d = bias1;
#pragma unroll
            for (int r = 0; r < 4; r++) d = mma32(w1[r], prelu(acc[r], alpha[r]), d);
            if (ok) {
                const long off = static_cast<long>(y) * fp;
                const float4 f = *(reinterpret_cast<const float4*>(reinterpret_cast<const char*>(la.feat) + off) + 2 * x + half);
                float4* px = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + off) + 2 * x + half;
                *px = make_float4(prelu(add_skip(d[0], f.x), alpha1[0]), prelu(add_skip(d[1], f.y), alpha1[1]),
                                  prelu(add_skip(d[2], f.z), alpha1[2]), prelu(add_skip(d[3], f.w), alpha1[3]));
            }
        } else {
            // lanes with half = 0 hold the four sub-pixels n = 0..3 (output (2x + n % 2,
            // 2y + n / 2)); half = 1 lanes take row 2y + 1 from them
            const float n2 = __shfl_xor(acc[2], 16), n3 = __shfl_xor(acc[3], 16);
            if (ok) {
                float s0 = half ? n2 : acc[0], s1 = half ? n3 : acc[1];
                if constexpr (MODE == kShuffleSrc) {
                    const float res = to_float<TIN>(row_ptr<TIN>(la.src, la.ss, y)[x]);
                    s0 = add_skip(s0, res);
                    s1 = add_skip(s1, res);
                }
                store_pair<TOUT>(row_ptr<TOUT>(la.dst, la.ds, 2 * y + half) + 2 * x, s0, s1);
            }
        }
    }
};

template <int MODE, typename TIN, typename TOUT, class C = Conv8Tile<>>
__global__ __launch_bounds__(C::NT) void conv8_layer(const float4* __restrict__ fin, long fp, int W, int H,
                                                     float4* __restrict__ fout, const LayerArgs la)
{
    constexpr int kC8TW = C::TW, kC8TH = C::TH, kC8PW = C::PW, kC8PH = C::PH;
    __shared__ float4 pa[kC8PW * kC8PH];
    __shared__ float4 pb[kC8PW * kC8PH];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * kC8TW, y0 = ty * kC8TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;

    // stage rows y0-1 .. y0+TH, columns x0-1 .. x0+TW (replicated at the image edges)
    {
        float4* const pl[2] = { pa, pb };
        stage_map<C::NT, kC8TW + 2, kC8PH, 2, kC8PW>(fin, fp, W, H, x0 - 1, y0 - 1, pl);
    }
    float a[24];
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = la.wa[s * 64 + lane];
    Epi8<MODE, TIN, TOUT> epi;
    epi.load(la, lane, half);
    __syncthreads();

    const int gx1 = min(x0 + kC8TW, W), gy1 = min(y0 + kC8TH, H);
    const int nseg = (gx1 - x0 + 15) >> 4;
    const int nitems = nseg * ((gy1 - y0 + 1) >> 1);
    const int lane_rd = g * kC8PW + j;
    for (int item = wave; item < nitems; item += C::NW) {
        const int pair = item / nseg, seg = item - pair * nseg;
        // patch row g of output rows (y0 + 2 pair, +1) is staged row 2 pair + g; column j + t
        const Item32 q = load_item(pa, pb, 2 * pair * kC8PW + 16 * seg + lane_rd);
        const f32x4m acc = item24(q, a, epi.bias);
        const int x = x0 + 16 * seg + j, y = y0 + 2 * pair + ro;
        epi.apply(acc, x, y, x < gx1 && y < gy1, half, fp, fout, la);
    }
}

// Two consecutive 8-channel layers in one launch: layer A (conv3x3 + PReLU or ReLU, arguments la0)
// and layer B (conv3x3 with the MODE epilogue, arguments la), so the map between them stays
// in LDS.  Output tile 62 x 16: layer A's region grows by one pixel per side to 64 x 18,
// exactly 4 segments x 9 row pairs; the staged input is 66 x 20 (80-pixel pitch, 51 KB,
// three workgroups per CU).  After layer A, tiles at the image edge overwrite A's pixels
// outside the image with the nearest pixel inside, so layer B sees the replicate padding
// every layer of the reference applies (Common.hpp:120-143).
constexpr int kP8TW = 62, kP8TH = 16, kP8PW = 80, kP8PH = kP8TH + 4;
// threads per two-layer workgroup, its waves, and layer A's items per wave
constexpr int kP8NT = 256, kP8NW = kP8NT / 64, kP8AI = (36 + kP8NW - 1) / kP8NW;

// The second layer of a two-layer launch, once layer A fills plane pixels (1..64, 1..18)
// of pa/pb: the replicate ring, then layer B (A operands a, epilogue epi) over the 62 x 16 tile.
template <int MODE, typename TIN, typename TOUT>
__device__ __forceinline__ void pair8_second(float4* __restrict__ pa, float4* __restrict__ pb, int W, int H, int x0, int y0,
                                             const float (&a)[24], const Epi8<MODE, TIN, TOUT>& epi, long fp,
                                             float4* __restrict__ fout, const LayerArgs& la)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    const int lane_rd = g * kP8PW + j;
    // replicate ring: A's pixels outside the image take the nearest pixel inside
    if (x0 - 1 < 0 || y0 - 1 < 0 || x0 + kP8TW >= W || y0 + kP8TH >= H) {
        for (int e = tid; e < 18 * 64; e += kP8NT) {
            const int r = 1 + e / 64, c = 1 + (e & 63);
            const int gy = y0 - 2 + r, gx = x0 - 2 + c;
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) continue;
            const int s = (clampi(gy, 0, H - 1) - (y0 - 2)) * kP8PW + clampi(gx, 0, W - 1) - (x0 - 2);
            pa[r * kP8PW + c] = pa[s];
            pb[r * kP8PW + c] = pb[s];
        }
        __syncthreads();
    }
    // layer B over the 62 x 16 tile: 32 items dealt round-robin to the waves; patch rows 2 pair + 1 + g,
    // columns 16 seg + j + t + 1 (columns past A's region feed only lanes past the tile)
    const int gx1 = min(x0 + kP8TW, W), gy1 = min(y0 + kP8TH, H);
    static_assert(32 % kP8NW == 0, "layer B's items split evenly over the waves");
#pragma unroll
    for (int ib = 0; ib < 32 / kP8NW; ib++) {   // unrolled: the next item's LDS reads issue early
        const int item = wave + kP8NW * ib, pair = item >> 2, seg = item & 3;
        const f32x4m acc = item24(load_item(pa, pb, (2 * pair + 1) * kP8PW + 16 * seg + 1 + lane_rd), a, epi.bias);
        const int x = x0 + 16 * seg + j, y = y0 + 2 * pair + ro;
        epi.apply(acc, x, y, x < gx1 && y < gy1, half, fp, fout, la);
    }
}


// Layer A of a two-layer launch: its input at plane pixels (0..65, 0..19) of pa/pb, its
// output written back to (1..64, 1..18) -- 36 items, round-robin over the waves; item (pair, seg) reads
// patch rows 2 pair + g, columns 16 seg + j + t.  Takes layer A's A operands in `a` and its
// epilogue in epa; returns with layer B's (la) in `a` and epi, loaded while the other waves
// finish layer A.
template <int MODE_A, int MODE, typename TIN, typename TOUT>
__device__ __forceinline__ void pair8_first(float4* __restrict__ pa, float4* __restrict__ pb, float (&a)[24],
                                            const Epi8<MODE_A, float, float>& epa, Epi8<MODE, TIN, TOUT>& epi,
                                            const LayerArgs& la)
{
    static_assert(MODE_A == kPrelu || MODE_A == kRelu, "layer A: conv3x3 + PReLU or ReLU");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15, half = g & 1, ro = g >> 1;
    const int lane_rd = g * kP8PW + j;
    float4 res[kP8AI];
#pragma unroll
    for (int ii = 0; ii < kP8AI; ii++) {
        const int item = wave + kP8NW * ii, pair = item >> 2, seg = item & 3;
        if (36 % kP8NW == 0 || item < 36) {   // uniform across the wave
            const f32x4m acc = item24(load_item(pa, pb, 2 * pair * kP8PW + 16 * seg + lane_rd), a, epa.bias);
            if constexpr (MODE_A == kPrelu)
                res[ii] = make_float4(prelu(acc[0], epa.alpha[0]), prelu(acc[1], epa.alpha[1]), prelu(acc[2], epa.alpha[2]),
                                      prelu(acc[3], epa.alpha[3]));
            else
                res[ii] = make_float4(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f), fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
        }
    }
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = la.wa[s * 64 + lane];
    epi.load(la, lane, half);
    __syncthreads();
    {
        float4* const pw = half ? pb : pa;
#pragma unroll
        for (int ii = 0; ii < kP8AI; ii++) {
            const int item = wave + kP8NW * ii, pair = item >> 2, seg = item & 3;
            if (36 % kP8NW == 0 || item < 36) pw[(1 + 2 * pair + ro) * kP8PW + 1 + 16 * seg + j] = res[ii];
        }
    }
    __syncthreads();
}

// Both layers of a two-layer launch on the tile whose output starts at (x0, y0); pa/pb are the
// workgroup's LDS planes.
template <int MODE, typename TIN, typename TOUT, int MODE_A>
__device__ __forceinline__ void pair8_tile(const float4* __restrict__ fin, long fp, int W, int H, float4* __restrict__ fout,
                                           const LayerArgs& la0, const LayerArgs& la, int x0, int y0,
                                           float4* __restrict__ pa, float4* __restrict__ pb)
{
    const int lane = threadIdx.x & 63, half = (lane >> 4) & 1;
    // plane pixel (c, r) is image pixel (x0 - 2 + c, y0 - 2 + r); stage 66 x 20, clamped
    {
        float4* const pl[2] = { pa, pb };
        stage_map<kP8NT, kP8TW + 4, kP8PH, 2, kP8PW>(fin, fp, W, H, x0 - 2, y0 - 2, pl);
    }
    float a[24];
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = la0.wa[s * 64 + lane];
    Epi8<MODE_A, float, float> epa;
    epa.load(la0, lane, half);
    __syncthreads();
    Epi8<MODE, TIN, TOUT> epi;
    pair8_first<MODE_A>(pa, pb, a, epa, epi, la);
    pair8_second<MODE, TIN, TOUT>(pa, pb, W, H, x0, y0, a, epi, fp, fout, la);
}

template <int MODE, typename TIN, typename TOUT, int MODE_A = kPrelu>
__global__ __launch_bounds__(kP8NT) void conv8_pair(const float4* __restrict__ fin, long fp, int W, int H,
                                                  float4* __restrict__ fout, const LayerArgs la0, const LayerArgs la)
{
    __shared__ float4 pa[kP8PW * kP8PH];
    __shared__ float4 pb[kP8PW * kP8PH];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    pair8_tile<MODE, TIN, TOUT, MODE_A>(fin, fp, W, H, fout, la0, la, tx * kP8TW, ty * kP8TH, pa, pb);
}

inline dim3 conv8_pair_grid(int W, int H) { return dim3((W + kP8TW - 1) / kP8TW, (H + kP8TH - 1) / kP8TH); }

// ---- F = 16 or 32: channel tiles ----
// TW x 16 output pixels per workgroup of NT threads; NI items per wave batch.  The defaults
// are the fastest of a sweep over 1080p maps on one MI355X: 32 x 16 (F = 16, 55 KB of LDS)
// and 16 x 16 (F = 32, 74 KB) tiles let two workgroups share a CU, so one stages its tile
// while the other computes.
template <int F, int TW_ = (F == 16 ? 32 : 16), int NT_ = 512, int NI_ = (F == 16 ? 4 : 2), int PF_ = 1>
struct ConvTile {
    static constexpr bool PF = PF_;                                     // load tap p + 1's A operands during tap p
    static constexpr int TW = TW_, TH = 16, NT = NT_, NW = NT / 64;
    static constexpr int PW = (TW + 2 + 15) / 16 * 16, PH = TH + 2;    // plane pitch: 16-pixel multiple
    static constexpr int NQ = F / 4;                                    // channel-quad planes
    static constexpr int KQ = F / 4;                                    // K steps per tap (channels per lane group)
    static constexpr int MT = F / 16;                                   // 16-channel output tiles
    static constexpr int NI = NI_;                                      // items per wave batch
    static constexpr int SEGS = TW / 16, ITEMS = SEGS * TH;
    static constexpr int BATCHES = (ITEMS / NW + NI - 1) / NI;
    static_assert(ITEMS % (NW * NI) == 0, "whole batches");
};

template <int F, int MODE, typename TOUT, class C = ConvTile<F>>
__global__ __launch_bounds__(C::NT) void convf_layer(const float4* __restrict__ fin, long fp, int W, int H,
                                                     float4* __restrict__ fout, const LayerArgs la)
{
    static_assert(MODE != kShuffleSrc, "convf_layer epilogues");
    __shared__ float4 planes[C::NQ][C::PH * C::PW];
    const int tiles_x = gridDim.x;
    const int t = xcd_tile(blockIdx.y * tiles_x + blockIdx.x, tiles_x * gridDim.y);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int x0 = tx * C::TW, y0 = ty * C::TH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15;

    // stage rows y0-1 .. y0+TH, columns x0-1 .. x0+TW (replicated at the image edges)
    {
        float4* pl[C::NQ];
#pragma unroll
        for (int q = 0; q < C::NQ; q++) pl[q] = planes[q];
        stage_map<C::NT, C::TW + 2, C::PH, C::NQ, C::PW>(fin, fp, W, H, x0 - 1, y0 - 1, pl);
    }
    // lane (g, j) of output tile h owns channels 16 h + 4 g .. + 3
    f32x4m b0[C::MT];
#pragma unroll
    for (int h = 0; h < C::MT; h++) {
        if constexpr (MODE == kShuffle)   // output channels 0..3 only: rows 0-3 of the first tile
            b0[h] = (h == 0 && g == 0) ? (f32x4m){ la.bias[0], la.bias[1], la.bias[2], la.bias[3] } : (f32x4m){ 0.f, 0.f, 0.f, 0.f };
        else {
            const float* bl = la.bias + 16 * h + 4 * g;
            b0[h] = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
        }
    }
    // the F -> 4 tail needs only the first output tile (channels 0-3)
    constexpr int MTU = MODE == kShuffle ? 1 : C::MT;
    // slopes and conv1x1 operands of this lane's channels, in registers for the whole tile
    constexpr bool PRELU = MODE == kPrelu || MODE == kPreluConv1x1;
    constexpr int NA = PRELU ? C::MT : 1, N1 = MODE == kPreluConv1x1 ? C::MT : 1;
    float al[NA][4], al1[N1][4], w1[4 * N1][N1];
    f32x4m b1[N1];
    if constexpr (PRELU) {
#pragma unroll
        for (int h = 0; h < C::MT; h++)
#pragma unroll
            for (int r = 0; r < 4; r++) al[h][r] = la.alpha[16 * h + 4 * g + r];
    }
    if constexpr (MODE == kPreluConv1x1) {
#pragma unroll
        for (int h = 0; h < C::MT; h++) {
            const float* bl = la.b1 + 16 * h + 4 * g;
            b1[h] = (f32x4m){ bl[0], bl[1], bl[2], bl[3] };
#pragma unroll
            for (int r = 0; r < 4; r++) {
                al1[h][r] = la.alpha1[16 * h + 4 * g + r];
#pragma unroll
                for (int h2 = 0; h2 < C::MT; h2++) w1[4 * h + r][h2] = la.w1[((4 * h + r) * 64 + lane) * C::MT + h2];
            }
        }
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
        // A operands [tap][lane][q * MTU + h]: KQ * MTU floats per tap and lane
        constexpr int AW = C::KQ * MTU;
        float4 an[AW / 4];
        const float4* const ap0 = reinterpret_cast<const float4*>(la.wa + static_cast<long>(lane) * AW);
        constexpr long kTapF4 = 64L * AW / 4;   // float4s per tap
        if constexpr (C::PF) {
#pragma unroll
            for (int v = 0; v < AW / 4; v++) an[v] = ap0[v];
        }
#pragma unroll 1
        for (int p = 0; p < 9; p++) {
            const int dy = p / 3, dx = p - dy * 3;
            float a[AW];
            if constexpr (C::PF) {
#pragma unroll
                for (int v = 0; v < AW / 4; v++) {
                    a[4 * v] = an[v].x; a[4 * v + 1] = an[v].y; a[4 * v + 2] = an[v].z; a[4 * v + 3] = an[v].w;
                }
                if (p < 8) {
#pragma unroll
                    for (int v = 0; v < AW / 4; v++) an[v] = ap0[(p + 1) * kTapF4 + v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < AW / 4; v++) {
                    const float4 f = ap0[p * kTapF4 + v];
                    a[4 * v] = f.x; a[4 * v + 1] = f.y; a[4 * v + 2] = f.z; a[4 * v + 3] = f.w;
                }
            }
            // B operands of every item first, then the MFMAs with the K step outermost, so
            // that consecutive MFMAs update different accumulators (NI x MTU chains)
            float bv[C::NI][C::KQ];
#pragma unroll
            for (int i = 0; i < C::NI; i++) {
                const int off = base[i] + dy * C::PW + dx;
#pragma unroll
                for (int v = 0; v < C::KQ / 4; v++) {
                    const float4 f = planes[(C::KQ / 4) * g + v][off];
                    bv[i][4 * v] = f.x; bv[i][4 * v + 1] = f.y; bv[i][4 * v + 2] = f.z; bv[i][4 * v + 3] = f.w;
                }
            }
#pragma unroll
            for (int q = 0; q < C::KQ; q++)
#pragma unroll
                for (int i = 0; i < C::NI; i++)
#pragma unroll
                    for (int h = 0; h < MTU; h++) acc[i][h] = mma32(a[q * MTU + h], bv[i][q], acc[i][h]);
        }
        if constexpr (MODE == kPreluConv1x1) {
            // PReLU, then conv1x1: K step 4 h + r of lane group g is channel 16 h + 4 g + r,
            // this lane's D register r of tile h; A operands [4 h + r][lane][h2]
#pragma unroll
            for (int i = 0; i < C::NI; i++) {
                f32x4m d[C::MT];
#pragma unroll
                for (int h2 = 0; h2 < C::MT; h2++) d[h2] = b1[h2];
#pragma unroll
                for (int h = 0; h < C::MT; h++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float bv = prelu(acc[i][h][r], al[h][r]);
#pragma unroll
                        for (int h2 = 0; h2 < C::MT; h2++) d[h2] = mma32(w1[4 * h + r][h2], bv, d[h2]);
                    }
#pragma unroll
                for (int h2 = 0; h2 < C::MT; h2++) acc[i][h2] = d[h2];
            }
        }
        // epilogue: lane (g, j) holds output channels 16 h + 4 g .. + 3 of pixel j
#pragma unroll
        for (int i = 0; i < C::NI; i++) {
            const int item = wave + C::NW * (batch * C::NI + i);
            const int row = item / C::SEGS, seg = item - row * C::SEGS;
            const int x = x0 + 16 * seg + j, y = y0 + row;
            if (x >= gx1 || y >= gy1) continue;
            if constexpr (MODE == kShuffle) {
                if (g == 0) {   // the 4 sub-pixels n: output (2x + n % 2, 2y + n / 2)
                    const f32x4m v = acc[i][0];
                    store_pair<TOUT>(row_ptr<TOUT>(la.dst, la.ds, 2 * y) + 2 * x, v[0], v[1]);
                    store_pair<TOUT>(row_ptr<TOUT>(la.dst, la.ds, 2 * y + 1) + 2 * x, v[2], v[3]);
                }
            } else {
                const long off = static_cast<long>(y) * fp;
                float4* po = reinterpret_cast<float4*>(reinterpret_cast<char*>(fout) + off) + C::NQ * x;
#pragma unroll
                for (int h = 0; h < C::MT; h++) {
                    const f32x4m v = acc[i][h];
                    float4 o;
                    if constexpr (MODE == kRelu) {
                        o = make_float4(fmaxf(v[0], 0.f), fmaxf(v[1], 0.f), fmaxf(v[2], 0.f), fmaxf(v[3], 0.f));
                    } else if constexpr (MODE == kPrelu) {
                        o = make_float4(prelu(v[0], al[h][0]), prelu(v[1], al[h][1]), prelu(v[2], al[h][2]), prelu(v[3], al[h][3]));
                    } else {   // kAddFeat, kPreluConv1x1: * 1 + feat (then PReLU)
                        const float4 r = *(reinterpret_cast<const float4*>(reinterpret_cast<const char*>(la.feat) + off) + C::NQ * x + 4 * h + g);
                        o = make_float4(add_skip(v[0], r.x), add_skip(v[1], r.y), add_skip(v[2], r.z), add_skip(v[3], r.w));
                        if constexpr (MODE == kPreluConv1x1)
                            o = make_float4(prelu(o.x, al1[h][0]), prelu(o.y, al1[h][1]), prelu(o.z, al1[h][2]), prelu(o.w, al1[h][3]));
                    }
                    po[4 * h + g] = o;
                }
            }
        }
    }
}

// Grids of the two mappings for a W x H map.
template <class C = Conv8Tile<>>
inline dim3 conv8_grid(int W, int H) { return dim3((W + C::TW - 1) / C::TW, (H + C::TH - 1) / C::TH); }
template <int F, class C = ConvTile<F>>
inline dim3 convf_grid(int W, int H) { return dim3((W + C::TW - 1) / C::TW, (H + C::TH - 1) / C::TH); }

} // namespace achip
