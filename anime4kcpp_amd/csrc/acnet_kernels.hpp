// Host-visible launch entry points of the ACNetLegacy HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace achip {

// One image of a batched launch (device memory).  Tiles of the batch are numbered image by
// image: image i owns tiles [first_tile, first_tile + tiles_x * rows), the tile rows ty0,
// ty0 + 1, ... of the image (a launch may cover part of an image's rows; borders are
// those of the whole w x h image).
struct BatchDesc {
    const void* src;
    void* dst;
    long sstride, dstride;
    int w, h;
    int tiles_x, first_tile;
    int ty0;
};

// One image of a batched layer-split pass: its images, its two 8-channel f32 maps (rows fp
// bytes apart) and its 62 x 16-pixel tiles, numbered image by image from first_tile.
struct SplitDesc {
    const void* src;
    void* dst;
    float4* m0;
    float4* m1;
    long sstride, dstride, fp;
    int w, h;
    int tiles_x, first_tile;
};

// The image that owns tile t (binary search over first_tile; desc sorted by it).
template <typename D>
__device__ inline D find_tile(const D* __restrict__ desc, int n, int t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (desc[mid].first_tile <= t) lo = mid;
        else hi = mid - 1;
    }
    return desc[lo];
}

// Tile of a non-persistent launch for workgroup b of T (row-major tile numbering).  The
// hardware hands consecutive workgroups to the 8 XCDs in turn; this permutation gives the
// workgroups that land on one XCD a contiguous band of tiles, so the input rows a band
// shares (halo) are fetched into that XCD's L2 once instead of into every XCD's.
__device__ inline int xcd_tile(int b, int T)
{
    constexpr int kXcd = 8;
    const int q = T / kXcd, r = T - q * kXcd;
    const int x = b % kXcd, i = b / kXcd;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Workgroups of a persistent launch: resident workgroups per CU x CUs, at most `tiles`, and
// at most max_wg when max_wg > 0 (leaves CUs to work on other streams).
int persistent_grid(const void* kernel, int threads, int tiles, int max_wg);

// One tile shape of the fused FP32 kernel: TW x TH input-resolution pixels per
// workgroup, NT threads, P-pixel vertical runs per thread.
struct Fp32Config {
    int tw, th, nt, p;
    int lds_bytes;
};
extern const Fp32Config kFp32Configs[];
extern const int kFp32ConfigCount;

// Fused FP32 2x pass.  src/dst are device pointers with byte strides; stype/dtype are
// element types (0x001 u8, 0x002 u16, 0x202 f16, 0x204 f32); wts is the device weight
// blob (acnet_device.hpp layout).  The caller has validated every argument.
hipError_t launch_fp32(int cfg, int stype, int dtype, const void* src, long sstride, int W, int H,
                       void* dst, long dstride, const float* wts, hipStream_t stream);

// Fused pass with layers 1..8 on 16-bit MFMA and layer 0 on FP32 MFMA.  bf16 = 0: FP16
// operands, 1: BF16 operands; wa is the per-lane A-operand blob of that type, wa0 the
// FP32 blob's layer-0 part (acnet_device.hpp layouts).
extern const Fp32Config kMfmaConfigs[];
extern const int kMfmaConfigCount;
hipError_t launch_mfma(int bf16, int cfg, int stype, int dtype, const void* src, long sstride, int W, int H,
                       void* dst, long dstride, const float* wts, const void* wa, const float* wa0, hipStream_t stream);

// Fused pass with every conv layer on FP32 MFMA (tile 64 x 32; configuration 0: 1024
// threads, 1: 512); wa32 is the per-lane FP32 A-operand blob (acnet_device.hpp layout).
extern const Fp32Config kMfma32Configs[];
extern const int kMfma32ConfigCount;
hipError_t launch_mfma32(int cfg, int stype, int dtype, const void* src, long sstride, int W, int H,
                         void* dst, long dstride, const float* wts, const float* wa32, hipStream_t stream);
hipError_t launch_mfma32_batch(int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                               const float* wts, const float* wa32, int max_wg, hipStream_t stream);

// ACNet<8> 2x pass, layer by layer through two feature maps in `workspace` (acnet8.hip):
// blocks 4, 8 or 18; wts the acnet_device.hpp ACNet<8> blob, wa the A operands of the
// blocks and the final layer.  Workspace: 2 * acnet8_feature_pitch(W) * H bytes.
long acnet8_feature_pitch(int w);
hipError_t launch_acnet8(int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst, long ds,
                         const float* wts, const float* wa, void* workspace, hipStream_t stream);

// ACNetLegacy's FP32 pass layer by layer (acnet8.hip): layer 0 into an f32 map, layers 1..8
// two per launch, the last pair ending in the deconv; workspace as launch_acnet8 (two maps of
// acnet8_feature_pitch(W) * H bytes).
hipError_t launch_legacy_split(int stype, int dtype, const void* src, long ss, int W, int H, void* dst, long ds,
                               const float* wts, const float* wa32, void* workspace, hipStream_t stream);
// The same for n images (one element type pair) in four launches over the tiles of all of
// them, `total` in all; desc is device memory.
hipError_t launch_legacy_split_batch(int stype, int dtype, const SplitDesc* desc, int n, int total, const float* wts,
                                     const float* wa32, hipStream_t stream);

// ACNet<8> 2x pass with the conv3x3 layers on 16-bit MFMA (acnet8_mfma16.hip): bf16 = 0 FP16,
// 1 BF16 operands; wa16 that type's per-lane A operands of the blocks and the final layer
// (3 x 64 x 8 values per layer, final layer rows 4-7 zero).  Workspace: two 16-bit maps of
// acnet8_map16_pitch(W) * H bytes (fits in the FP32 path's workspace).
long acnet8_map16_pitch(int w);
hipError_t launch_acnet8_16(int bf16, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                            long ds, const float* wts, const void* wa16, void* workspace, hipStream_t stream);

// ArtCNN<F> 2x pass, F = 16 or 32 (artcnn.hip): kernels/biases in the reference layout
// (device copies), wa the per-lane A operands of the B body layers, the residual layer and
// the F -> 4 tail, 9 * 64 * (F/4) * (F/16) floats each ([tap][lane][K step][channel tile]).
// Workspace: 3 * artcnn_feature_pitch(F, W) * H bytes (head features and two maps).
long artcnn_feature_pitch(int f, int w);
hipError_t launch_artcnn(int f, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                         long ds, const float* kernels, const float* biases, const float* wa, void* workspace,
                         hipStream_t stream);

// FSRCNNX<F> 2x pass, F = 8 or 16 (fsrcnnx.hip): kernels/biases/alphas in the reference
// layout (device copies); wa the per-lane A operands of the B conv3x3 layers and the F -> 4
// tail (pack_conv8_operands for F = 8, pack_convf_operands for F = 16, each layer
// conv_operand_floats(F) floats), w1 those of the conv1x1 (pack_conv8_1x1 / pack_convf_1x1).
// Workspace: 3 * fsrcnnx_feature_pitch(F, W) * H bytes (head features and two maps).
long fsrcnnx_feature_pitch(int f, int w);
hipError_t launch_fsrcnnx(int f, int blocks, int stype, int dtype, const void* src, long ss, int W, int H, void* dst,
                          long ds, const float* kernels, const float* biases, const float* alphas, const float* wa,
                          const float* w1, void* workspace, hipStream_t stream);

// Host-side packing of the per-lane A operands the layer kernels (conv_layers.hpp) read.
// F = 8 (2-row items), conv3x3 8 -> cout (8 or 4): [t * 8 + c][lane] for patch column t,
// channel c; conv1x1 8 -> 8: [r][lane], block-diagonal over the item's two output rows.
// F = 16 or 32 (channel tiles), conv3x3 F -> F: [tap][lane][q * F/16 + h]; F -> 4 (the tail,
// one output tile): [tap][lane][q];
// conv1x1 F -> F: [4 h + r][lane][h2].  k in the reference layout k[n][tap][c].
void pack_conv8_operands(const float* k, int cout, float* out);
void pack_conv8_1x1(const float* k1, float* out);
void pack_convf_operands(const float* k, int f, int cout, float* out);
void pack_convf_1x1(const float* k1, int f, float* out);
inline size_t conv_operand_floats(int f) { return f == 8 ? 24 * 64 : 9 * 64 * static_cast<size_t>(f / 4) * (f / 16); }
inline size_t conv_1x1_floats(int f) { return f == 8 ? 4 * 64 : static_cast<size_t>(f / 4) * 64 * (f / 16); }

// ---- colour conversion and resize (acnet_color.hip) ----

// One colour component of an image in device memory: component (x, y) is the element at
// ptr + y * pitch bytes + (x * step + off) elements.
struct PlaneView {
    void* ptr;
    long pitch;
    int step, off;
};
// Components 0..3 in (R, G, B[, A] or Y, U, V[, A]) and out, in that order.
struct ColourPlanes {
    PlaneView in[4], out[4];
};
// ImageProcess.cpp: rgb2yuv, rgba2yuva, yuv2rgb, yuva2rgba
enum ColourKind : int { kRgb2Yuv = 0, kRgba2Yuva = 1, kYuv2Rgb = 2, kYuva2Rgba = 3 };
hipError_t launch_colour(int type, int kind, const ColourPlanes& planes, int W, int H, hipStream_t stream);

// Resize modes: the reference's ResizeModes numbering (core/include/AC/Core/Image.hpp:276-295).
enum ResizeMode : int {
    kResizePoint = 0, kResizeCatmullRom, kResizeMitchell, kResizeBicubic0_60, kResizeBicubic0_75,
    kResizeBicubic0_100, kResizeBicubic20_50, kResizeSoftCubic50, kResizeSoftCubic75, kResizeSoftCubic100,
    kResizeLanczos2, kResizeLanczos3, kResizeLanczos4, kResizeSpline16, kResizeSpline36, kResizeSpline64,
    kResizeBilinear, kResizeModeCount
};
// filter support radius in source pixels at scale 1 (ImageResize.cpp:193-263)
__host__ __device__ inline float resize_support(int mode)
{
    switch (mode) {
    case kResizePoint: return 0.5f;
    case kResizeLanczos3: case kResizeSpline36: return 3.0f;
    case kResizeLanczos4: case kResizeSpline64: return 4.0f;
    case kResizeBilinear: return 1.0f;
    default: return 2.0f;
    }
}
// Upper bound of the taps of one output coordinate (source n_in -> n_out pixels); a resize
// whose bound on either axis exceeds kResizeMaxTaps is rejected (downscales below about
// 1/47 for support-2 filters, 1/23 for Lanczos4).
constexpr int kResizeMaxTaps = 192;
int resize_max_taps(int mode, int n_in, int n_out);

// Colour split RGB(A) -> Y + UV(A) and merge back (interleaved channels, one element type,
// byte strides), and resize; hipErrorInvalidValue for a mode, size or tap count out of range.
hipError_t launch_color_split(int type, int channels, const void* src, long ss, int W, int H,
                              void* y, long ys, void* uv, long uvs, hipStream_t stream);
hipError_t launch_color_merge(int type, int channels, const void* y, long ys, const void* uv, long uvs, int W, int H,
                              void* dst, long ds, hipStream_t stream);
hipError_t launch_resize(int type, int channels, int mode, const void* src, long ss, int sw, int sh,
                         void* dst, long ds, int dw, int dh, hipStream_t stream);
// UV(A) (channels - 1 interleaved, sw x sh) resized to W x H, rounded to the element type,
// and merged with Y (W x H) into RGB(A): launch_resize + launch_color_merge in one pass.
hipError_t launch_resize_merge(int type, int channels, int mode, const void* y, long ys, const void* uv, long uvs, int sw,
                               int sh, int W, int H, void* dst, long ds, hipStream_t stream);

// ---- element-wise image helpers (acnet_image.hip) ----
// n = elements per row (width x channels); byte strides.
// dst = fromFloat<dtype>(toFloat(src)) (ImageProcess.cpp:339-355)
hipError_t launch_convert(int stype, int dtype, const void* src, long ss, void* dst, long ds, int n, int h,
                          hipStream_t stream);
// unsigned integer shift by `bits` (left when left != 0), truncated to the element type
hipError_t launch_shift(int type, int left, int bits, const void* src, long ss, void* dst, long ds, int n, int h,
                        hipStream_t stream);
// pixelShuffle (ImageProcess.cpp:358-385): src sw x sh with sc = dc * up * up channels ->
// dst (sw * up) x (sh * up) with dc channels
hipError_t launch_pixel_shuffle(int stype, int dtype, const void* src, long ss, int sw, int sh, int sc, int up,
                                void* dst, long ds, hipStream_t stream);
// rows x row_bytes bytes from src (rows ss bytes apart) to dst (ds apart) by a kernel; either
// pointer may be the device mapping of page-locked host memory
hipError_t launch_copy_rows(const void* src, long ss, void* dst, long ds, long row_bytes, int rows, hipStream_t stream);
// for every pixel of a w x h grid, copy len bytes from src + y*ss + x*spx + soff to
// dst + y*ds + x*dpx + doff (channel extract / insert)
hipError_t launch_copy_runs(const void* src, long ss, int spx, int soff, void* dst, long ds, int dpx, int doff, int len,
                            int w, int h, hipStream_t stream);

// Persistent batched launches: `counter` (device int, zero at launch) is the tile queue.
hipError_t launch_fp32_batch(int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                             const float* wts, int max_wg, hipStream_t stream);
hipError_t launch_mfma_batch(int bf16, int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total,
                             int* counter, const float* wts, const void* wa, const float* wa0, int max_wg,
                             hipStream_t stream);

} // namespace achip
