// Device kernels around the 2x passes: colour split and merge, and resize.  These are the
// HBM-bound steps of Processor::process (core/src/processor/Processor.cpp:199-276) for
// multi-channel input and non-power-of-two factors.
//
// Colour conversion follows core/src/ImageProcess.cpp:38-61 / :114-139 (split) and
// :191-215 / :276-308 (merge): BT.601-like float constants, each product and sum rounded
// separately in the order written there (no contraction), RGBA premultiplied by alpha
// before the split and divided by it after the merge, every output through fromFloat.
//
// Resize replaces core/src/ImageResize.cpp:136-273 (stb_image_resize2): separable filter,
// edge clamp, pixel-centre alignment (source x = (dst x + 0.5) / scale - 0.5); Catmull-Rom
// (a = -0.5) or bilinear; when downscaling the filter is stretched by 1/scale and its
// weights normalised.  Horizontal taps are summed first for each source row, then the
// rows vertically, in FP32.
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

#pragma clang fp contract(off)

template <typename T, int C>
__global__ __launch_bounds__(256) void color_split(const void* __restrict__ src, long ss, int W, int H,
                                                  void* __restrict__ y, long ys, void* __restrict__ uv, long uvs)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || r0 >= H) return;
    const T* in = row_ptr<T>(src, ss, r0) + static_cast<long>(x) * C;
    float r = to_float<T>(in[0]), g = to_float<T>(in[1]), b = to_float<T>(in[2]);
    float a = 1.0f;
    if (C == 4) {
        a = to_float<T>(in[3]);
        r = __fmul_rn(r, a);
        g = __fmul_rn(g, a);
        b = __fmul_rn(b, a);
    }
    const float yy = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, r), __fmul_rn(0.587f, g)), __fmul_rn(0.114f, b));
    const float u = __fadd_rn(__fmul_rn(0.564f, __fsub_rn(b, yy)), 0.5f);
    const float v = __fadd_rn(__fmul_rn(0.713f, __fsub_rn(r, yy)), 0.5f);
    row_ptr<T>(y, ys, r0)[x] = from_float<T>(yy);
    T* o = row_ptr<T>(uv, uvs, r0) + static_cast<long>(x) * (C - 1);
    o[0] = from_float<T>(u);
    o[1] = from_float<T>(v);
    if (C == 4) o[2] = from_float<T>(a);
}

template <typename T, int C>
__global__ __launch_bounds__(256) void color_merge(const void* __restrict__ y, long ys, const void* __restrict__ uv, long uvs,
                                                  int W, int H, void* __restrict__ dst, long ds)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || r0 >= H) return;
    const T* ui = row_ptr<T>(uv, uvs, r0) + static_cast<long>(x) * (C - 1);
    const float yy = to_float<T>(row_ptr<T>(y, ys, r0)[x]);
    const float u = __fsub_rn(to_float<T>(ui[0]), 0.5f);
    const float v = __fsub_rn(to_float<T>(ui[1]), 0.5f);
    float r = __fadd_rn(yy, __fmul_rn(1.403f, v));
    float g = __fsub_rn(__fsub_rn(yy, __fmul_rn(0.344f, u)), __fmul_rn(0.714f, v));
    float b = __fadd_rn(yy, __fmul_rn(1.773f, u));
    T* o = row_ptr<T>(dst, ds, r0) + static_cast<long>(x) * C;
    if (C == 4) {
        const float a = to_float<T>(ui[2]);
        if (a > 1e-6f) {
            r = __fdiv_rn(r, a);
            g = __fdiv_rn(g, a);
            b = __fdiv_rn(b, a);
        } else {
            r = g = b = 0.0f;
        }
        o[3] = from_float<T>(a);
    }
    o[0] = from_float<T>(r);
    o[1] = from_float<T>(g);
    o[2] = from_float<T>(b);
}

// filter kernels: mode 0 bilinear (support 1), 1 Catmull-Rom (support 2)
__device__ inline float filter_weight(int mode, float d)
{
    d = fabsf(d);
    if (mode == 0) return d < 1.0f ? 1.0f - d : 0.0f;
    if (d < 1.0f) return __fadd_rn(__fmul_rn(__fmul_rn(1.5f, d), __fmul_rn(d, d)), __fadd_rn(__fmul_rn(-2.5f, __fmul_rn(d, d)), 1.0f));
    if (d < 2.0f)
        return __fadd_rn(__fadd_rn(__fmul_rn(-0.5f, __fmul_rn(d, __fmul_rn(d, d))), __fmul_rn(2.5f, __fmul_rn(d, d))),
                         __fadd_rn(__fmul_rn(-4.0f, d), 2.0f));
    return 0.0f;
}

// Taps of one output coordinate: first source index and normalised weights (at most kMaxTaps).
constexpr int kMaxTaps = 12;
__device__ inline int resize_taps(int mode, int o, int n_in, int n_out, float* w, int& first)
{
    const float scale = static_cast<float>(n_out) / static_cast<float>(n_in);
    const float center = (static_cast<float>(o) + 0.5f) / scale - 0.5f;
    const float support = (mode == 0 ? 1.0f : 2.0f) * (scale < 1.0f ? 1.0f / scale : 1.0f);
    const float stretch = scale < 1.0f ? scale : 1.0f;
    int lo = static_cast<int>(ceilf(center - support));
    int hi = static_cast<int>(floorf(center + support));
    if (hi - lo + 1 > kMaxTaps) {                 // only for scales below 1/3 (not used by the pipeline)
        lo = static_cast<int>(center) - kMaxTaps / 2 + 1;
        hi = lo + kMaxTaps - 1;
    }
    float sum = 0.0f;
    int n = 0;
    for (int i = lo; i <= hi; i++, n++) {
        w[n] = filter_weight(mode, (static_cast<float>(i) - center) * stretch);
        sum += w[n];
    }
    if (sum != 0.0f)
        for (int i = 0; i < n; i++) w[i] /= sum;
    first = lo;
    return n;
}

// A workgroup computes a 64 x 16 block of outputs (thread: one column, four rows).  The taps
// of the block's 64 columns and 16 rows are computed once into LDS (resize_taps is the
// costly part: filter, normalising division), then every output sums its taps in the order
// above.
constexpr int kRsW = 64, kRsH = 16;

struct ResizeTaps {
    float wx[kRsW][kMaxTaps];
    float wy[kRsH][kMaxTaps];
    int fx[kRsW], nx[kRsW], fy[kRsH], ny[kRsH];
};

__device__ inline void block_taps(ResizeTaps& t, int mode, int sw, int sh, int dw, int dh)
{
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kRsW, y0 = blockIdx.y * kRsH;
    if (tid < kRsW) {
        const int x = min(x0 + tid, dw - 1);
        t.nx[tid] = resize_taps(mode, x, sw, dw, t.wx[tid], t.fx[tid]);
    } else if (tid < kRsW + kRsH) {
        const int r = tid - kRsW;
        const int y = min(y0 + r, dh - 1);
        t.ny[r] = resize_taps(mode, y, sh, dh, t.wy[r], t.fy[r]);
    }
    __syncthreads();
}

// resized value of channel block [C] at output (column tx, row ty) of the block
template <typename T, int C>
__device__ inline void resample(const ResizeTaps& t, const void* __restrict__ src, long ss, int sw, int sh, int tx,
                                int ty, float (&acc)[C])
{
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0f;
    const int nx = t.nx[tx], fx = t.fx[tx];
    for (int j = 0; j < t.ny[ty]; j++) {
        const T* row = row_ptr<T>(src, ss, clampi(t.fy[ty] + j, 0, sh - 1));
        float h[C];
#pragma unroll
        for (int c = 0; c < C; c++) h[c] = 0.0f;
        for (int i = 0; i < nx; i++) {
            const T* p = row + static_cast<long>(clampi(fx + i, 0, sw - 1)) * C;
#pragma unroll
            for (int c = 0; c < C; c++) h[c] = __fadd_rn(h[c], __fmul_rn(t.wx[tx][i], to_float<T>(p[c])));
        }
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] = __fadd_rn(acc[c], __fmul_rn(t.wy[ty][j], h[c]));
    }
}

template <typename T, int C>
__global__ __launch_bounds__(256) void resize_kernel(const void* __restrict__ src, long ss, int sw, int sh,
                                                    void* __restrict__ dst, long ds, int dw, int dh, int mode)
{
    __shared__ ResizeTaps t;
    block_taps(t, mode, sw, sh, dw, dh);
    const int tx = threadIdx.x & 63;
    const int x = blockIdx.x * kRsW + tx;
    if (x >= dw) return;
#pragma unroll
    for (int k = 0; k < kRsH / 4; k++) {
        const int ty = (threadIdx.x >> 6) + 4 * k;
        const int y = blockIdx.y * kRsH + ty;
        if (y >= dh) break;
        float acc[C];
        resample<T, C>(t, src, ss, sw, sh, tx, ty, acc);
        T* o = row_ptr<T>(dst, ds, y) + static_cast<long>(x) * C;
#pragma unroll
        for (int c = 0; c < C; c++) o[c] = from_float<T>(acc[c]);
    }
}

// Chroma resize fused with the merge (Processor.cpp:258-274 with dst allocated: resize UV(A)
// to the luma's size, then yuv2rgb): each output pixel's UV(A) is resampled, rounded to the
// element type exactly as the resized image would hold it, and merged with Y at once, so
// the full-size UV image is never written or read back.
template <typename T, int C>
__global__ __launch_bounds__(256) void resize_merge_kernel(const void* __restrict__ y, long ys, const void* __restrict__ uv,
                                                          long uvs, int sw, int sh, int W, int H, void* __restrict__ dst,
                                                          long ds, int mode)
{
    __shared__ ResizeTaps t;
    block_taps(t, mode, sw, sh, W, H);
    const int tx = threadIdx.x & 63;
    const int x = blockIdx.x * kRsW + tx;
    if (x >= W) return;
#pragma unroll
    for (int k = 0; k < kRsH / 4; k++) {
        const int ty = (threadIdx.x >> 6) + 4 * k;
        const int r0 = blockIdx.y * kRsH + ty;
        if (r0 >= H) break;
        float acc[C - 1];
        resample<T, C - 1>(t, uv, uvs, sw, sh, tx, ty, acc);
        const float yy = to_float<T>(row_ptr<T>(y, ys, r0)[x]);
        const float u = __fsub_rn(to_float<T>(from_float<T>(acc[0])), 0.5f);
        const float v = __fsub_rn(to_float<T>(from_float<T>(acc[1])), 0.5f);
        float r = __fadd_rn(yy, __fmul_rn(1.403f, v));
        float g = __fsub_rn(__fsub_rn(yy, __fmul_rn(0.344f, u)), __fmul_rn(0.714f, v));
        float b = __fadd_rn(yy, __fmul_rn(1.773f, u));
        T* o = row_ptr<T>(dst, ds, r0) + static_cast<long>(x) * C;
        if (C == 4) {
            const float a = to_float<T>(from_float<T>(acc[C - 2]));
            if (a > 1e-6f) {
                r = __fdiv_rn(r, a);
                g = __fdiv_rn(g, a);
                b = __fdiv_rn(b, a);
            } else {
                r = g = b = 0.0f;
            }
            o[3] = from_float<T>(a);
        }
        o[0] = from_float<T>(r);
        o[1] = from_float<T>(g);
        o[2] = from_float<T>(b);
    }
}

template <typename T>
static hipError_t split_t(int channels, const void* src, long ss, int W, int H, void* y, long ys, void* uv, long uvs,
                          hipStream_t stream)
{
    const dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
    if (channels == 4) hipLaunchKernelGGL((color_split<T, 4>), grid, block, 0, stream, src, ss, W, H, y, ys, uv, uvs);
    else hipLaunchKernelGGL((color_split<T, 3>), grid, block, 0, stream, src, ss, W, H, y, ys, uv, uvs);
    return hipGetLastError();
}

template <typename T>
static hipError_t merge_t(int channels, const void* y, long ys, const void* uv, long uvs, int W, int H, void* dst, long ds,
                          hipStream_t stream)
{
    const dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
    if (channels == 4) hipLaunchKernelGGL((color_merge<T, 4>), grid, block, 0, stream, y, ys, uv, uvs, W, H, dst, ds);
    else hipLaunchKernelGGL((color_merge<T, 3>), grid, block, 0, stream, y, ys, uv, uvs, W, H, dst, ds);
    return hipGetLastError();
}

template <typename T>
static hipError_t resize_t(int channels, int mode, const void* src, long ss, int sw, int sh, void* dst, long ds, int dw,
                           int dh, hipStream_t stream)
{
    const dim3 grid((dw + kRsW - 1) / kRsW, (dh + kRsH - 1) / kRsH), block(256);
    switch (channels) {
    case 1: hipLaunchKernelGGL((resize_kernel<T, 1>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode); break;
    case 2: hipLaunchKernelGGL((resize_kernel<T, 2>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode); break;
    case 3: hipLaunchKernelGGL((resize_kernel<T, 3>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode); break;
    case 4: hipLaunchKernelGGL((resize_kernel<T, 4>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t resize_merge_t(int channels, int mode, const void* y, long ys, const void* uv, long uvs, int sw, int sh,
                                 int W, int H, void* dst, long ds, hipStream_t stream)
{
    const dim3 grid((W + kRsW - 1) / kRsW, (H + kRsH - 1) / kRsH), block(256);
    if (channels == 4)
        hipLaunchKernelGGL((resize_merge_kernel<T, 4>), grid, block, 0, stream, y, ys, uv, uvs, sw, sh, W, H, dst, ds, mode);
    else if (channels == 3)
        hipLaunchKernelGGL((resize_merge_kernel<T, 3>), grid, block, 0, stream, y, ys, uv, uvs, sw, sh, W, H, dst, ds, mode);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_resize_merge(int type, int channels, int mode, const void* y, long ys, const void* uv, long uvs, int sw,
                               int sh, int W, int H, void* dst, long ds, hipStream_t stream)
{
    switch (type) {
    case U8: return resize_merge_t<uint8_t>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, stream);
    case U16: return resize_merge_t<uint16_t>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, stream);
    case F16: return resize_merge_t<__half>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, stream);
    case F32: return resize_merge_t<float>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_color_split(int type, int channels, const void* src, long ss, int W, int H,
                              void* y, long ys, void* uv, long uvs, hipStream_t stream)
{
    switch (type) {
    case U8: return split_t<uint8_t>(channels, src, ss, W, H, y, ys, uv, uvs, stream);
    case U16: return split_t<uint16_t>(channels, src, ss, W, H, y, ys, uv, uvs, stream);
    case F16: return split_t<__half>(channels, src, ss, W, H, y, ys, uv, uvs, stream);
    case F32: return split_t<float>(channels, src, ss, W, H, y, ys, uv, uvs, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_color_merge(int type, int channels, const void* y, long ys, const void* uv, long uvs, int W, int H,
                              void* dst, long ds, hipStream_t stream)
{
    switch (type) {
    case U8: return merge_t<uint8_t>(channels, y, ys, uv, uvs, W, H, dst, ds, stream);
    case U16: return merge_t<uint16_t>(channels, y, ys, uv, uvs, W, H, dst, ds, stream);
    case F16: return merge_t<__half>(channels, y, ys, uv, uvs, W, H, dst, ds, stream);
    case F32: return merge_t<float>(channels, y, ys, uv, uvs, W, H, dst, ds, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_resize(int type, int channels, int mode, const void* src, long ss, int sw, int sh,
                         void* dst, long ds, int dw, int dh, hipStream_t stream)
{
    switch (type) {
    case U8: return resize_t<uint8_t>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, stream);
    case U16: return resize_t<uint16_t>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, stream);
    case F16: return resize_t<__half>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, stream);
    case F32: return resize_t<float>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace achip
