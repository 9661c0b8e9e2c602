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

// The image that owns tile t (binary search over first_tile; desc sorted by it).
__device__ inline BatchDesc find_tile(const BatchDesc* __restrict__ desc, int n, int t)
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

// Colour split RGB(A) -> Y + UV(A), merge back, and resize (mode 0 bilinear, 1 Catmull-Rom);
// interleaved channels, one element type, byte strides (acnet_color.hip).
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

// Persistent batched launches: `counter` (device int, zero at launch) is the tile queue.
hipError_t launch_fp32_batch(int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total, int* counter,
                             const float* wts, int max_wg, hipStream_t stream);
hipError_t launch_mfma_batch(int bf16, int cfg, int stype, int dtype, const BatchDesc* desc, int n, int total,
                             int* counter, const float* wts, const void* wa, const float* wa0, int max_wg,
                             hipStream_t stream);

} // namespace achip
