// Device kernels around the 2x passes and behind the image helpers: colour conversion
// and resize.  These are the HBM-bound steps of Processor::process
// (core/src/processor/Processor.cpp:199-276) for multi-channel input and non-power-of-two
// factors, and of ac::core::{rgb2yuv, rgba2yuva, yuv2rgb, yuva2rgba, resize}.
//
// Colour conversion follows core/src/ImageProcess.cpp:15-308: BT.601-like float constants,
// each product and sum rounded separately in the order written there (no contraction),
// RGBA premultiplied by alpha before the forward conversion and divided by it after the
// inverse, every output through fromFloat.  One kernel serves the packed, two-plane and
// per-channel forms: each component is read and written through a plane view.
//
// Resize replaces core/src/ImageResize.cpp:136-273 (stb_image_resize2 with the reference's
// filter callbacks): separable filter, edge clamp, pixel-centre alignment (source
// x = (dst x + 0.5) / scale - 0.5); when downscaling the filter is stretched by 1/scale;
// weights normalised.  Horizontal taps are summed first for each source row, then the
// rows vertically, in FP32.
#include <type_traits>

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

#pragma clang fp contract(off)

// ---- colour -------------------------------------------------------------------------------

// ImageProcess.cpp:26-28 (forward) and :186-188 (inverse; u, v already centred)
__device__ inline void rgb_to_yuv(float r, float g, float b, float& y, float& u, float& v)
{
    y = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, r), __fmul_rn(0.587f, g)), __fmul_rn(0.114f, b));
    u = __fadd_rn(__fmul_rn(0.564f, __fsub_rn(b, y)), 0.5f);
    v = __fadd_rn(__fmul_rn(0.713f, __fsub_rn(r, y)), 0.5f);
}
__device__ inline void yuv_to_rgb(float y, float u, float v, float& r, float& g, float& b)
{
    r = __fadd_rn(y, __fmul_rn(1.403f, v));
    g = __fsub_rn(__fsub_rn(y, __fmul_rn(0.344f, u)), __fmul_rn(0.714f, v));
    b = __fadd_rn(y, __fmul_rn(1.773f, u));
}
// yuva2rgba's un-premultiply (ImageProcess.cpp:290-296)
__device__ inline void unpremultiply(float a, float& r, float& g, float& b)
{
    if (a > 1e-6f) {
        r = __fdiv_rn(r, a);
        g = __fdiv_rn(g, a);
        b = __fdiv_rn(b, a);
    } else {
        r = g = b = 0.0f;
    }
}

template <typename T>
__device__ inline const T& comp_at(const PlaneView& p, int x, int y)
{
    return row_ptr<T>(p.ptr, p.pitch, y)[static_cast<long>(x) * p.step + p.off];
}
template <typename T>
__device__ inline T& comp_at_mut(const PlaneView& p, int x, int y)
{
    return row_ptr<T>(p.ptr, p.pitch, y)[static_cast<long>(x) * p.step + p.off];
}

// One pixel's conversion on float components (in and out in the component order of
// ColourPlanes), shared by the kernels below.
template <int KIND>
__device__ inline void colour_px(const float (&in)[4], float (&out)[4])
{
    if constexpr (KIND == kRgb2Yuv || KIND == kRgba2Yuva) {
        float r = in[0], g = in[1], b = in[2];
        if constexpr (KIND == kRgba2Yuva) {
            r = __fmul_rn(r, in[3]);
            g = __fmul_rn(g, in[3]);
            b = __fmul_rn(b, in[3]);
            out[3] = in[3];
        }
        rgb_to_yuv(r, g, b, out[0], out[1], out[2]);
    } else {
        const float u = __fsub_rn(in[1], 0.5f), v = __fsub_rn(in[2], 0.5f);
        yuv_to_rgb(in[0], u, v, out[0], out[1], out[2]);
        if constexpr (KIND == kYuva2Rgba) {
            unpremultiply(in[3], out[0], out[1], out[2]);
            out[3] = in[3];
        }
    }
}

template <int KIND>
constexpr int colour_components() { return KIND == kRgba2Yuva || KIND == kYuva2Rgba ? 4 : 3; }

// One pixel per thread, 64 x 4 per workgroup; any element type and plane layout.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void colour_kernel(ColourPlanes pl, int W, int H)
{
    constexpr int NC = colour_components<KIND>();
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || yy >= H) return;
    float in[4] = { 0.f, 0.f, 0.f, 1.f }, out[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int k = 0; k < NC; k++) in[k] = to_float<T>(comp_at<T>(pl.in[k], x, yy));
    colour_px<KIND>(in, out);
#pragma unroll
    for (int k = 0; k < NC; k++) comp_at_mut<T>(pl.out[k], x, yy) = from_float<T>(out[k]);
}

// u8 packed pixels <-> Y + interleaved chroma, four pixels per thread: the packed side is
// NC dwords, the Y side one dword, the chroma side NC - 1 dwords (every base and pitch a
// multiple of 4 bytes).  SPLIT: packed -> planes.  Pixels past the last multiple of four in
// a row are done one by one by the row's last thread.  Same per-pixel steps as colour_kernel.
template <int KIND, bool SPLIT>
__global__ __launch_bounds__(256) void colour_u8x4_kernel(const uint8_t* __restrict__ pk, long ps, uint8_t* __restrict__ yp,
                                                          long yps, uint8_t* __restrict__ cp, long cps, int W, int H)
{
    constexpr int NC = colour_components<KIND>();
    const int q = blockIdx.x * 256 + threadIdx.x;   // pixel quad
    const int x = 4 * q;
    if (x >= W) return;
    for (int yy = blockIdx.y; yy < H; yy += gridDim.y) {
        uint8_t* prow = const_cast<uint8_t*>(pk) + static_cast<long>(yy) * ps;
        uint8_t* yrow = yp + static_cast<long>(yy) * yps;
        uint8_t* crow = cp + static_cast<long>(yy) * cps;
        const int n = min(4, W - x);
        uint8_t a[4][4], b[4][4];   // [pixel][component] of the source and the result
        if (n == 4) {
            if constexpr (SPLIT) {
                uint32_t w[NC];
#pragma unroll
                for (int k = 0; k < NC; k++) w[k] = reinterpret_cast<const uint32_t*>(prow + x * NC)[k];
#pragma unroll
                for (int i = 0; i < 4 * NC; i++) a[i / NC][i % NC] = static_cast<uint8_t>(w[i / 4] >> (8 * (i % 4)));
            } else {
                const uint32_t wy = *reinterpret_cast<const uint32_t*>(yrow + x);
                uint32_t wc[NC - 1];
#pragma unroll
                for (int k = 0; k < NC - 1; k++) wc[k] = reinterpret_cast<const uint32_t*>(crow + x * (NC - 1))[k];
#pragma unroll
                for (int p = 0; p < 4; p++) a[p][0] = static_cast<uint8_t>(wy >> (8 * p));
#pragma unroll
                for (int i = 0; i < 4 * (NC - 1); i++)
                    a[i / (NC - 1)][1 + i % (NC - 1)] = static_cast<uint8_t>(wc[i / 4] >> (8 * (i % 4)));
            }
        } else {
            for (int p = 0; p < n; p++)
#pragma unroll
                for (int k = 0; k < NC; k++)
                    a[p][k] = SPLIT ? prow[(x + p) * NC + k] : (k == 0 ? yrow[x + p] : crow[(x + p) * (NC - 1) + k - 1]);
        }
#pragma unroll
        for (int p = 0; p < 4; p++) {
            float in[4] = { 0.f, 0.f, 0.f, 1.f }, out[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int k = 0; k < NC; k++) in[k] = to_float<uint8_t>(a[p][k]);
            colour_px<KIND>(in, out);
#pragma unroll
            for (int k = 0; k < NC; k++) b[p][k] = from_float<uint8_t>(out[k]);
        }
        if (n == 4) {
            if constexpr (SPLIT) {
                uint32_t wy = 0, wc[NC - 1];
#pragma unroll
                for (int p = 0; p < 4; p++) wy |= static_cast<uint32_t>(b[p][0]) << (8 * p);
#pragma unroll
                for (int k = 0; k < NC - 1; k++) wc[k] = 0;
#pragma unroll
                for (int i = 0; i < 4 * (NC - 1); i++)
                    wc[i / 4] |= static_cast<uint32_t>(b[i / (NC - 1)][1 + i % (NC - 1)]) << (8 * (i % 4));
                *reinterpret_cast<uint32_t*>(yrow + x) = wy;
#pragma unroll
                for (int k = 0; k < NC - 1; k++) reinterpret_cast<uint32_t*>(crow + x * (NC - 1))[k] = wc[k];
            } else {
                uint32_t w[NC];
#pragma unroll
                for (int k = 0; k < NC; k++) w[k] = 0;
#pragma unroll
                for (int i = 0; i < 4 * NC; i++) w[i / 4] |= static_cast<uint32_t>(b[i / NC][i % NC]) << (8 * (i % 4));
#pragma unroll
                for (int k = 0; k < NC; k++) reinterpret_cast<uint32_t*>(prow + x * NC)[k] = w[k];
            }
        } else {
            for (int p = 0; p < n; p++)
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    if constexpr (SPLIT) {
                        if (k == 0) yrow[x + p] = b[p][0];
                        else crow[(x + p) * (NC - 1) + k - 1] = b[p][k];
                    } else {
                        prow[(x + p) * NC + k] = b[p][k];
                    }
                }
        }
    }
}

// ---- resize filters -----------------------------------------------------------------------
// Mode numbers are the reference's ResizeModes (core/include/AC/Core/Image.hpp:276-295);
// the kernels are its filter callbacks (ImageResize.cpp:53-134, 193-263): the (b, c)
// cubic family of Mitchell and Netravali, Lanczos, the spline16/36/64 piecewise cubics
// and the tent, each evaluated in float as a Horner polynomial.

__device__ inline float poly3(float x, float c0, float c1, float c2, float c3)
{
    return __fadd_rn(c0, __fmul_rn(x, __fadd_rn(c1, __fmul_rn(x, __fadd_rn(c2, __fmul_rn(x, c3))))));
}

// (b, c) cubic with b = bn/bd, c = cn/cd; coefficients rounded to float step by step
struct CubicBC {
    float p0, p2, p3, q0, q1, q2, q3;
};
__host__ __device__ constexpr CubicBC cubic_bc(int bn, int bd, int cn, int cd)
{
    const float b = static_cast<float>(bn) / static_cast<float>(bd);
    const float c = static_cast<float>(cn) / static_cast<float>(cd);
    return { (6.0f - 2.0f * b) / 6.0f,
             (-18.0f + 12.0f * b + 6.0f * c) / 6.0f,
             (12.0f - 9.0f * b - 6.0f * c) / 6.0f,
             (8.0f * b + 24.0f * c) / 6.0f,
             (-12.0f * b - 48.0f * c) / 6.0f,
             (6.0f * b + 30.0f * c) / 6.0f,
             (-b - 6.0f * c) / 6.0f };
}
__device__ inline float cubic(const CubicBC& k, float x)
{
    if (x < 1.0f) return poly3(x, k.p0, 0.0f, k.p2, k.p3);
    if (x < 2.0f) return poly3(x, k.q0, k.q1, k.q2, k.q3);
    return 0.0f;
}

__device__ inline float sinc(float x)
{
    constexpr float pi = 3.14159265358979f;
    return x == 0.0f ? 1.0f : __fdiv_rn(sinf(__fmul_rn(x, pi)), __fmul_rn(x, pi));
}
__device__ inline float lanczos(float x, float taps)
{
    return x < taps ? __fmul_rn(sinc(x), sinc(__fdiv_rn(x, taps))) : 0.0f;
}

// Spline16/36/64 (Panorama Tools' interpolating spline kernels), piece k on [k, k+1)
__device__ inline float spline16(float x)
{
    if (x < 1.0f) return poly3(x, 1.0f, -1.0f / 5.0f, -9.0f / 5.0f, 1.0f);
    if (x < 2.0f) return poly3(__fsub_rn(x, 1.0f), 0.0f, -7.0f / 15.0f, 4.0f / 5.0f, -1.0f / 3.0f);
    return 0.0f;
}
__device__ inline float spline36(float x)
{
    if (x < 1.0f) return poly3(x, 1.0f, -3.0f / 209.0f, -453.0f / 209.0f, 13.0f / 11.0f);
    if (x < 2.0f) return poly3(__fsub_rn(x, 1.0f), 0.0f, -156.0f / 209.0f, 270.0f / 209.0f, -6.0f / 11.0f);
    if (x < 3.0f) return poly3(__fsub_rn(x, 2.0f), 0.0f, 26.0f / 209.0f, -45.0f / 209.0f, 1.0f / 11.0f);
    return 0.0f;
}
__device__ inline float spline64(float x)
{
    if (x < 1.0f) return poly3(x, 1.0f, -3.0f / 2911.0f, -6387.0f / 2911.0f, 49.0f / 41.0f);
    if (x < 2.0f) return poly3(__fsub_rn(x, 1.0f), 0.0f, -2328.0f / 2911.0f, 4032.0f / 2911.0f, -24.0f / 41.0f);
    if (x < 3.0f) return poly3(__fsub_rn(x, 2.0f), 0.0f, 582.0f / 2911.0f, -1008.0f / 2911.0f, 6.0f / 41.0f);
    if (x < 4.0f) return poly3(__fsub_rn(x, 3.0f), 0.0f, -97.0f / 2911.0f, 168.0f / 2911.0f, -1.0f / 41.0f);
    return 0.0f;
}

__device__ inline float filter_weight(int mode, float d)
{
    const float x = fabsf(d);
    switch (mode) {
    case kResizePoint: return 1.0f;
    case kResizeCatmullRom: { constexpr CubicBC k = cubic_bc(0, 1, 1, 2); return cubic(k, x); }
    case kResizeMitchell: { constexpr CubicBC k = cubic_bc(1, 3, 1, 3); return cubic(k, x); }
    case kResizeBicubic0_60: { constexpr CubicBC k = cubic_bc(0, 1, 3, 5); return cubic(k, x); }
    case kResizeBicubic0_75: { constexpr CubicBC k = cubic_bc(0, 1, 3, 4); return cubic(k, x); }
    case kResizeBicubic0_100: { constexpr CubicBC k = cubic_bc(0, 1, 1, 1); return cubic(k, x); }
    case kResizeBicubic20_50: { constexpr CubicBC k = cubic_bc(1, 5, 1, 2); return cubic(k, x); }
    case kResizeSoftCubic50: { constexpr CubicBC k = cubic_bc(1, 2, 1, 2); return cubic(k, x); }
    case kResizeSoftCubic75: { constexpr CubicBC k = cubic_bc(3, 4, 1, 4); return cubic(k, x); }
    case kResizeSoftCubic100: { constexpr CubicBC k = cubic_bc(1, 1, 0, 1); return cubic(k, x); }
    case kResizeLanczos2: return lanczos(x, 2.0f);
    case kResizeLanczos3: return lanczos(x, 3.0f);
    case kResizeLanczos4: return lanczos(x, 4.0f);
    case kResizeSpline16: return spline16(x);
    case kResizeSpline36: return spline36(x);
    case kResizeSpline64: return spline64(x);
    default: return x < 1.0f ? __fsub_rn(1.0f, x) : 0.0f;   // bilinear
    }
}

// Taps of one output coordinate: the source pixels whose centres lie in
// (center - support, center + support], their filter weights normalised to sum 1.
// At most kmax taps are written (the launcher sizes kmax from resize_max_taps).
__device__ inline int resize_taps(int mode, int o, int n_in, int n_out, float* w, int kmax, int& first)
{
    const float scale = static_cast<float>(n_out) / static_cast<float>(n_in);
    const float center = __fsub_rn(__fdiv_rn(__fadd_rn(static_cast<float>(o), 0.5f), scale), 0.5f);
    const float support = __fmul_rn(resize_support(mode), scale < 1.0f ? __fdiv_rn(1.0f, scale) : 1.0f);
    const float stretch = scale < 1.0f ? scale : 1.0f;
    const int lo = static_cast<int>(floorf(__fsub_rn(center, support))) + 1;
    const int hi = static_cast<int>(floorf(__fadd_rn(center, support)));
    const int n = min(max(hi - lo + 1, 0), kmax);
    float sum = 0.0f;
    for (int i = 0; i < n; i++) {
        w[i] = filter_weight(mode, __fmul_rn(__fsub_rn(static_cast<float>(lo + i), center), stretch));
        sum = __fadd_rn(sum, w[i]);
    }
    if (sum != 0.0f)
        for (int i = 0; i < n; i++) w[i] = __fdiv_rn(w[i], sum);
    first = lo;
    return n;
}

// A workgroup computes a 64 x 16 block of outputs (thread: one column, four rows).  The taps
// of the block's 64 columns and 16 rows are computed once into LDS (resize_taps is the
// costly part: filter, normalising division), then every output sums its taps in the order
// above.  LDS: wx[64][kp], wy[16][kp] floats, then fx, nx[64], fy, ny[16] ints; kp = kmax | 1
// (odd row pitch: the 64 columns' reads of one tap fall in different banks).
constexpr int kRsW = 64, kRsH = 16;

struct ResizeTaps {
    float* wx;
    float* wy;
    int *fx, *nx, *fy, *ny;
    int kp;
};

__device__ inline ResizeTaps block_taps(int mode, int sw, int sh, int dw, int dh, int kmax)
{
    extern __shared__ float smem[];
    ResizeTaps t;
    t.kp = kmax | 1;
    t.wx = smem;
    t.wy = smem + kRsW * t.kp;
    t.fx = reinterpret_cast<int*>(t.wy + kRsH * t.kp);
    t.nx = t.fx + kRsW;
    t.fy = t.nx + kRsW;
    t.ny = t.fy + kRsH;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kRsW, y0 = blockIdx.y * kRsH;
    if (tid < kRsW) {
        const int x = min(x0 + tid, dw - 1);
        t.nx[tid] = resize_taps(mode, x, sw, dw, t.wx + tid * t.kp, kmax, t.fx[tid]);
    } else if (tid < kRsW + kRsH) {
        const int r = tid - kRsW;
        const int y = min(y0 + r, dh - 1);
        t.ny[r] = resize_taps(mode, y, sh, dh, t.wy + r * t.kp, kmax, t.fy[r]);
    }
    __syncthreads();
    return t;
}

// resized value of channel block [C] at output (column tx, row ty) of the block
template <typename T, int C>
__device__ inline void resample(const ResizeTaps& t, const void* __restrict__ src, long ss, int sw, int sh, int tx,
                                int ty, float (&acc)[C])
{
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0f;
    const int nx = t.nx[tx], fx = t.fx[tx];
    const float* wx = t.wx + tx * t.kp;
    const float* wy = t.wy + ty * t.kp;
    for (int j = 0; j < t.ny[ty]; j++) {
        const T* row = row_ptr<T>(src, ss, clampi(t.fy[ty] + j, 0, sh - 1));
        float h[C];
#pragma unroll
        for (int c = 0; c < C; c++) h[c] = 0.0f;
        for (int i = 0; i < nx; i++) {
            const T* p = row + static_cast<long>(clampi(fx + i, 0, sw - 1)) * C;
#pragma unroll
            for (int c = 0; c < C; c++) h[c] = __fadd_rn(h[c], __fmul_rn(wx[i], to_float<T>(p[c])));
        }
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] = __fadd_rn(acc[c], __fmul_rn(wy[j], h[c]));
    }
}

template <typename T, int C>
__global__ __launch_bounds__(256) void resize_kernel(const void* __restrict__ src, long ss, int sw, int sh,
                                                    void* __restrict__ dst, long ds, int dw, int dh, int mode, int kmax)
{
    const ResizeTaps t = block_taps(mode, sw, sh, dw, dh, kmax);
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
                                                          long ds, int mode, int kmax)
{
    const ResizeTaps t = block_taps(mode, sw, sh, W, H, kmax);
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
        float r, g, b;
        yuv_to_rgb(yy, u, v, r, g, b);
        T* o = row_ptr<T>(dst, ds, r0) + static_cast<long>(x) * C;
        if constexpr (C == 4) {
            const float a = to_float<T>(from_float<T>(acc[C - 2]));
            unpremultiply(a, r, g, b);
            o[3] = from_float<T>(a);
        }
        o[0] = from_float<T>(r);
        o[1] = from_float<T>(g);
        o[2] = from_float<T>(b);
    }
}

// ---- exact 2x / 4x upscale ---------------------------------------------------------------
// When dst is exactly K = 2 or 4 times src on both axes, every output of phase p = x mod K
// has the same normalised taps, shifted by x div K: the centre (x + 0.5) / K - 0.5 and every
// tap distance are exact in float, so resize_taps gives bit-identical weights for all of them.
// A workgroup then stages its source region once in LDS as float, runs the horizontal taps
// for every staged row, then the vertical taps -- the same sums in the same order as
// resample(), so the results equal the general kernel's exactly, with each source pixel
// loaded and converted once instead of once per tap.  Taps of an upscale reach at most
// 4 source pixels from x div K (support <= 4); the staged region has kUpR on each side.
constexpr int kUpW = 64, kUpH = 32, kUpR = 5, kUpMaxTaps = 12;

template <int K>
struct UpTile {
    static constexpr int IW = kUpW / K + 2 * kUpR;   // staged source columns
    static constexpr int IH = kUpH / K + 2 * kUpR;   // staged source rows
};

struct UpTaps {
    float w[2][4][kUpMaxTaps];   // [axis][phase][tap]
    int off[2][4], n[2][4];      // first tap relative to x div K, tap count
};

template <int K>
__device__ inline void up_taps(UpTaps& t, int mode, int sw, int sh, int dw, int dh)
{
    const int tid = threadIdx.x;
    if (tid < 2 * K) {
        const int axis = tid / K, p = tid % K;
        int first = 0;
        t.n[axis][p] = resize_taps(mode, p, axis ? sh : sw, axis ? dh : dw, t.w[axis][p], kUpMaxTaps, first);
        t.off[axis][p] = first;   // output p has source index 0 as its x div K
    }
}

// the staged region's source rows/columns -> LDS index, kept inside the region
__device__ inline int up_local(int first, int base, int size) { return clampi(first - base, 0, size - 1); }

// One pass over a K-times upscale: channels [C] of src into acc for outputs of this tile;
// MERGE: the resampled chroma is rounded to T and merged with Y into RGB(A) (C = chroma
// channels, the output has C + 1).
template <typename T, int C, int K, bool MERGE>
__global__ __launch_bounds__(256) void resize_up_kernel(const void* __restrict__ src, long ss, int sw, int sh,
                                                       void* __restrict__ dst, long ds, int dw, int dh, int mode,
                                                       const void* __restrict__ y, long ys)
{
    using Tile = UpTile<K>;
    __shared__ UpTaps taps;
    __shared__ float in[Tile::IH][Tile::IW][C];
    __shared__ float hs[Tile::IH][kUpW][C];
    const int tid = threadIdx.x;
    const int ox0 = blockIdx.x * kUpW, oy0 = blockIdx.y * kUpH;
    const int ix0 = ox0 / K - kUpR, iy0 = oy0 / K - kUpR;   // staged region origin (source pixels)
    up_taps<K>(taps, mode, sw, sh, dw, dh);
    for (int e = tid; e < Tile::IH * Tile::IW; e += 256) {
        const int r = e / Tile::IW, c = e - r * Tile::IW;
        const T* p = row_ptr<T>(src, ss, clampi(iy0 + r, 0, sh - 1)) + static_cast<long>(clampi(ix0 + c, 0, sw - 1)) * C;
#pragma unroll
        for (int ch = 0; ch < C; ch++) in[r][c][ch] = to_float<T>(p[ch]);
    }
    __syncthreads();
    // horizontal taps of every staged row for the tile's 64 output columns
    {
        const int X = tid & 63, x = ox0 + X, p = x % K;
        const int nx = taps.n[0][p];
        const int base = up_local(x / K + taps.off[0][p], ix0, Tile::IW);
        for (int r = tid >> 6; r < Tile::IH; r += 4) {
            float h[C];
#pragma unroll
            for (int ch = 0; ch < C; ch++) h[ch] = 0.0f;
            for (int i = 0; i < nx; i++) {
                const int c = min(base + i, Tile::IW - 1);
#pragma unroll
                for (int ch = 0; ch < C; ch++) h[ch] = __fadd_rn(h[ch], __fmul_rn(taps.w[0][p][i], in[r][c][ch]));
            }
#pragma unroll
            for (int ch = 0; ch < C; ch++) hs[r][X][ch] = h[ch];
        }
    }
    __syncthreads();
    // vertical taps: thread = one output column, rows tid/64, +4, ...
    const int X = tid & 63, x = ox0 + X;
    if (x >= dw) return;
    for (int Y = tid >> 6; Y < kUpH; Y += 4) {
        const int yy = oy0 + Y;
        if (yy >= dh) break;
        const int p = yy % K, ny = taps.n[1][p];
        const int base = up_local(yy / K + taps.off[1][p], iy0, Tile::IH);
        float acc[C];
#pragma unroll
        for (int ch = 0; ch < C; ch++) acc[ch] = 0.0f;
        for (int j = 0; j < ny; j++) {
            const int r = min(base + j, Tile::IH - 1);
#pragma unroll
            for (int ch = 0; ch < C; ch++) acc[ch] = __fadd_rn(acc[ch], __fmul_rn(taps.w[1][p][j], hs[r][X][ch]));
        }
        if constexpr (!MERGE) {
            T* o = row_ptr<T>(dst, ds, yy) + static_cast<long>(x) * C;
#pragma unroll
            for (int ch = 0; ch < C; ch++) o[ch] = from_float<T>(acc[ch]);
        } else {
            const float lum = to_float<T>(row_ptr<T>(y, ys, yy)[x]);
            const float u = __fsub_rn(to_float<T>(from_float<T>(acc[0])), 0.5f);
            const float v = __fsub_rn(to_float<T>(from_float<T>(acc[1])), 0.5f);
            float r, g, b;
            yuv_to_rgb(lum, u, v, r, g, b);
            T* o = row_ptr<T>(dst, ds, yy) + static_cast<long>(x) * (C + 1);
            if constexpr (C == 3) {
                const float a = to_float<T>(from_float<T>(acc[2]));
                unpremultiply(a, r, g, b);
                o[3] = from_float<T>(a);
            }
            o[0] = from_float<T>(r);
            o[1] = from_float<T>(g);
            o[2] = from_float<T>(b);
        }
    }
}

// ---- exact 2x / 4x upscale of u8 planes, four adjacent outputs per thread ----------------
// The same sums as resize_up_kernel (horizontal taps of each source row first, then the
// vertical taps, each product and sum rounded separately, taps in increasing order), for the
// support-1 and support-2 filters (NT = 2 or 4 taps per output, every phase) on u8 data:
// - a thread owns four adjacent output columns and walks kU8Rows output rows of its wave's
//   band, keeping only the horizontal sums of the source rows still ahead in registers (the
//   loops are unrolled, so which rows are live is known at compile time);
// - a sum starts from its first product instead of 0 + the first product: the two differ
//   only in the sign of an all-zero sum, which fromFloat maps to the same code;
// - the source region is staged in LDS as float once (36 x 132 pixels at K = 2);
// - MERGE: the resampled chroma codes index 256-entry tables of the products the inverse
//   colour conversion forms from them (c / 255 - 0.5 times each coefficient, computed with
//   the same operations), and Y's toFloat comes from a table the same way;
// - outputs leave as whole dwords (a wave writes 768 contiguous bytes of RGB per row) when
//   the planes' bases and pitches allow it, byte by byte otherwise.
// Phase p of K (output x with x mod K = p) has its first tap at x div K - S when
// (p + 1/2) / K - 1/2 < 0, else at x div K - S + 1 (S = NT / 2: floor(centre - S) + 1 with a
// non-integer centre in (-1/2, 1/2)), so the staged region needs NT source pixels beyond the
// tile's own on each axis in total.
constexpr int kU8Cols = 256;   // output columns per workgroup (64 lanes x 4)

template <int C, int K, int NT, int ROWS, int NWV>
struct U8Up {
    static constexpr int S = NT / 2;
    static constexpr int R = NWV * ROWS;           // output rows per workgroup (NWV waves)
    static constexpr int SW = kU8Cols / K + NT;    // staged source columns
    static constexpr int SH = R / K + NT;          // staged source rows
    static constexpr int HR = ROWS / K + NT;       // source rows one wave's band reads
    // first staged row (relative to the wave's) of output row r's taps
    static constexpr int q0(int r) { return r / K + (2 * (r % K) >= K ? 1 : 0); }
    // staged column (relative to 4 * lane * (1 / K)) of column i's first tap
    static constexpr int c0(int i) { return i / K + (2 * (i % K) >= K ? 1 : 0); }
    static constexpr int HC = 4 / K + NT;          // staged columns one thread reads per row
};

// fromFloat<uint8_t> for finite v as a 32-bit code: the median of (v, 0, 1) is the
// reference's saturation for every finite v except that -0 may come out as +0, and both
// give code 0.  (The sums below are of finite u8 inputs, so no NaN reaches it.)  The
// multiply and add are one FMA: see u8_scale_code.
__device__ inline unsigned u8_code(float v)
{
    return u8_scale_code(__builtin_amdgcn_fmed3f(v, 0.0f, 1.0f));
}

template <int C, bool MERGE>
struct U8Lut {
    float y[256];
    float2 u[256], v[256];   // (0.344 u, 1.773 u), (1.403 v, 0.714 v) with u = c / 255 - 0.5
};
template <int C>
struct U8Lut<C, false> {};

template <int C, int K, int NT, int ROWS, int NWV, bool MERGE>
__global__ __launch_bounds__(64 * NWV) void resize_up_u8_kernel(const uint8_t* __restrict__ src, long ss, int sw, int sh,
                                                          uint8_t* __restrict__ dst, long ds, int dw, int dh, int mode,
                                                          const uint8_t* __restrict__ yp, long ys, int vec)
{
    using G = U8Up<C, K, NT, ROWS, NWV>;
    constexpr int CO = MERGE ? C + 1 : C;   // output bytes per pixel
    __shared__ float in[G::SH][G::SW][C];
    __shared__ float wt[2][K][NT];          // [axis][phase][tap]
    __shared__ U8Lut<C, MERGE> lut;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * kU8Cols, y0 = blockIdx.y * G::R;
    const int ix0 = x0 / K - G::S, iy0 = y0 / K - G::S;
    if (tid < 2 * K) {
        const int axis = tid / K, p = tid % K;
        float w[NT];
        int first = 0;
        const int n = resize_taps(mode, p, axis ? sh : sw, axis ? dh : dw, w, NT, first);
#pragma unroll
        for (int j = 0; j < NT; j++) wt[axis][p][j] = j < n ? w[j] : 0.0f;
    }
    if constexpr (MERGE)
        for (int i = tid; i < 256; i += 64 * NWV) {
            const float f = to_float<uint8_t>(static_cast<uint8_t>(i)), d = __fsub_rn(f, 0.5f);
            lut.y[i] = f;
            lut.u[i] = make_float2(__fmul_rn(0.344f, d), __fmul_rn(1.773f, d));
            lut.v[i] = make_float2(__fmul_rn(1.403f, d), __fmul_rn(0.714f, d));
        }
    for (int r = wave; r < G::SH; r += NWV) {
        const uint8_t* row = src + static_cast<long>(clampi(iy0 + r, 0, sh - 1)) * ss;
        for (int c = lane; c < G::SW; c += 64) {
            const uint8_t* p = row + static_cast<long>(clampi(ix0 + c, 0, sw - 1)) * C;
#pragma unroll
            for (int ch = 0; ch < C; ch++) in[r][c][ch] = to_float<uint8_t>(p[ch]);
        }
    }
    __syncthreads();

    float wx[K][NT], wy[K][NT];
#pragma unroll
    for (int p = 0; p < K; p++)
#pragma unroll
        for (int j = 0; j < NT; j++) {
            wx[p][j] = wt[0][p][j];
            wy[p][j] = wt[1][p][j];
        }
    const int xs = lane * (4 / K);      // staged column of this thread's first tap, less c0
    const int qb = wave * (ROWS / K);   // staged row of this wave's first tap, less q0
    const int xo = x0 + 4 * lane;       // first output column
    const bool full = xo + 4 <= dw;
    const bool fast = full && vec;

    // horizontal taps of staged row q for this thread's four columns
    auto hrow = [&](int q, float (&h)[4][C]) {
        float s[G::HC][C];
#pragma unroll
        for (int k = 0; k < G::HC; k++)
#pragma unroll
            for (int ch = 0; ch < C; ch++) s[k][ch] = in[qb + q][xs + k][ch];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int p = i % K, b = G::c0(i);
#pragma unroll
            for (int ch = 0; ch < C; ch++) {
                float a = __fmul_rn(wx[p][0], s[b][ch]);
#pragma unroll
                for (int j = 1; j < NT; j++) a = __fadd_rn(a, __fmul_rn(wx[p][j], s[b + j][ch]));
                h[i][ch] = a;
            }
        }
    };

    float h[G::HR][4][C];
#pragma unroll
    for (int q = 0; q < NT; q++) hrow(q, h[q]);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        if (r > 0 && G::q0(r) > G::q0(r - 1)) hrow(G::q0(r) + NT - 1, h[G::q0(r) + NT - 1]);
        const int oy = y0 + wave * ROWS + r;
        if (oy >= dh) break;
        const int p = r % K;
        unsigned code[4][C];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int ch = 0; ch < C; ch++) {
                float a = __fmul_rn(wy[p][0], h[G::q0(r)][i][ch]);
#pragma unroll
                for (int j = 1; j < NT; j++) a = __fadd_rn(a, __fmul_rn(wy[p][j], h[G::q0(r) + j][i][ch]));
                code[i][ch] = u8_code(a);
            }
        uint8_t* o = dst + static_cast<long>(oy) * ds + static_cast<long>(xo) * CO;
        unsigned px[4][CO];
        if constexpr (MERGE) {
            const uint8_t* yr = yp + static_cast<long>(oy) * ys + xo;
            unsigned yc[4];
            if (fast) {
                const unsigned w = *reinterpret_cast<const unsigned*>(yr);
#pragma unroll
                for (int i = 0; i < 4; i++) yc[i] = (w >> (8 * i)) & 0xffu;
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) yc[i] = xo + i < dw ? yr[i] : 0;
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float yy = lut.y[yc[i]];
                const float2 lu = lut.u[code[i][0]], lv = lut.v[code[i][1]];
                float r_ = __fadd_rn(yy, lv.x);
                float g_ = __fsub_rn(__fsub_rn(yy, lu.x), lv.y);
                float b_ = __fadd_rn(yy, lu.y);
                if constexpr (C == 3) {
                    const float a = lut.y[code[i][2]];
                    unpremultiply(a, r_, g_, b_);
                    px[i][3] = u8_code(a);
                }
                px[i][0] = u8_code(r_);
                px[i][1] = u8_code(g_);
                px[i][2] = u8_code(b_);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int ch = 0; ch < C; ch++) px[i][ch] = code[i][ch];
        }
        if (fast) {
            // 4 * CO bytes as CO dwords, byte k of the run in byte k % 4 of dword k / 4
            unsigned wds[CO];
#pragma unroll
            for (int d = 0; d < CO; d++) {
                unsigned v = px[(4 * d) / CO][(4 * d) % CO];
#pragma unroll
                for (int k = 1; k < 4; k++) {
                    const int e = 4 * d + k;
                    v |= px[e / CO][e % CO] << (8 * k);
                }
                wds[d] = v;
            }
            unsigned* od = reinterpret_cast<unsigned*>(o);
#pragma unroll
            for (int d = 0; d < CO; d++) od[d] = wds[d];
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (xo + i < dw)
#pragma unroll
                    for (int ch = 0; ch < CO; ch++) o[i * CO + ch] = static_cast<uint8_t>(px[i][ch]);
        }
    }
}

// the u8 kernel above handles this (type, channel count, filter) combination
template <typename T, int C, bool MERGE>
constexpr bool u8_up_kind() { return std::is_same_v<T, uint8_t> && (MERGE ? (C == 2 || C == 3) : (C == 1 || C == 2)); }

// 8-row bands per wave, or 4-row bands when that would leave the grid under two
// workgroups per CU (a 1080p frame's chroma planes: 272 workgroups of 8-row bands)
template <int C, int K, int NT, bool MERGE>
static void launch_up_u8(const void* src, long ss, int sw, int sh, void* dst, long ds, int dw, int dh, int mode, const void* y,
                         long ys, int vec, hipStream_t stream)
{
    const int gx = (dw + kU8Cols - 1) / kU8Cols;
    const auto* s8 = static_cast<const uint8_t*>(src);
    const auto* y8 = static_cast<const uint8_t*>(y);
    auto* d8 = static_cast<uint8_t*>(dst);
    if (static_cast<long>(gx) * ((dh + 31) / 32) >= 512) {
        const dim3 grid(gx, (dh + 31) / 32);
        hipLaunchKernelGGL((resize_up_u8_kernel<C, K, NT, 8, 4, MERGE>), grid, dim3(256), 0, stream, s8, ss, sw, sh, d8, ds, dw,
                           dh, mode, y8, ys, vec);
    } else {
        const dim3 grid(gx, (dh + 15) / 16);
        hipLaunchKernelGGL((resize_up_u8_kernel<C, K, NT, 4, 4, MERGE>), grid, dim3(256), 0, stream, s8, ss, sw, sh, d8, ds, dw,
                           dh, mode, y8, ys, vec);
    }
}

// K when dst is exactly 2 or 4 times src on both axes (and small enough for exact centres)
static int upscale_factor(int sw, int sh, int dw, int dh)
{
    if (sw > (1 << 20) || sh > (1 << 20)) return 0;
    for (int k : { 2, 4 })
        if (dw == k * sw && dh == k * sh) return k;
    return 0;
}

static bool aligned4(const void* p, long pitch) { return (reinterpret_cast<uintptr_t>(p) % 4) == 0 && pitch % 4 == 0; }

template <typename T, int C, bool MERGE>
static hipError_t resize_up_t(int k, int mode, const void* src, long ss, int sw, int sh, void* dst, long ds, int dw, int dh,
                              const void* y, long ys, hipStream_t stream)
{
    if constexpr (u8_up_kind<T, C, MERGE>()) {
        const float sup = resize_support(mode);
        if (sup == 1.0f || sup == 2.0f) {
            const int vec = aligned4(dst, ds) && (!MERGE || aligned4(y, ys)) ? 1 : 0;
            if (sup == 1.0f) {
                if (k == 2) launch_up_u8<C, 2, 2, MERGE>(src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys, vec, stream);
                else launch_up_u8<C, 4, 2, MERGE>(src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys, vec, stream);
            } else {
                if (k == 2) launch_up_u8<C, 2, 4, MERGE>(src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys, vec, stream);
                else launch_up_u8<C, 4, 4, MERGE>(src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys, vec, stream);
            }
            return hipGetLastError();
        }
    }
    const dim3 grid((dw + kUpW - 1) / kUpW, (dh + kUpH - 1) / kUpH), block(256);
    if (k == 2)
        hipLaunchKernelGGL((resize_up_kernel<T, C, 2, MERGE>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys);
    else
        hipLaunchKernelGGL((resize_up_kernel<T, C, 4, MERGE>), grid, block, 0, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, y, ys);
    return hipGetLastError();
}

// ---- launchers ----------------------------------------------------------------------------

int resize_max_taps(int mode, int n_in, int n_out)
{
    const float scale = static_cast<float>(n_out) / static_cast<float>(n_in);
    const float support = resize_support(mode) * (scale < 1.0f ? 1.0f / scale : 1.0f);
    return static_cast<int>(2.0f * support) + 2;
}

static size_t resize_lds_bytes(int kmax)
{
    return (static_cast<size_t>(kRsW + kRsH) * (kmax | 1) + 2 * (kRsW + kRsH)) * sizeof(float);
}

// Taps per output for this resize, or -1 when the tables would not fit in LDS.
static int resize_kmax(int mode, int sw, int sh, int dw, int dh)
{
    const int k = max(resize_max_taps(mode, sw, dw), resize_max_taps(mode, sh, dh));
    return k <= kResizeMaxTaps ? k : -1;
}

template <typename T, int KIND>
static hipError_t colour_t(const ColourPlanes& pl, int W, int H, hipStream_t stream)
{
    const dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
    hipLaunchKernelGGL((colour_kernel<T, KIND>), grid, block, 0, stream, pl, W, H);
    return hipGetLastError();
}

template <typename T>
static hipError_t colour_kind(int kind, const ColourPlanes& pl, int W, int H, hipStream_t stream)
{
    switch (kind) {
    case kRgb2Yuv: return colour_t<T, kRgb2Yuv>(pl, W, H, stream);
    case kRgba2Yuva: return colour_t<T, kRgba2Yuva>(pl, W, H, stream);
    case kYuv2Rgb: return colour_t<T, kYuv2Rgb>(pl, W, H, stream);
    case kYuva2Rgba: return colour_t<T, kYuva2Rgba>(pl, W, H, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_colour(int type, int kind, const ColourPlanes& pl, int W, int H, hipStream_t stream)
{
    if (W <= 0 || H <= 0) return hipSuccess;
    switch (type) {
    case U8: return colour_kind<uint8_t>(kind, pl, W, H, stream);
    case U16: return colour_kind<uint16_t>(kind, pl, W, H, stream);
    case F16: return colour_kind<__half>(kind, pl, W, H, stream);
    case F32: return colour_kind<float>(kind, pl, W, H, stream);
    default: return hipErrorInvalidValue;
    }
}

// Packed RGB(A) <-> Y plane + interleaved UV(A) plane: the colour kernel's views.
static ColourPlanes split_planes(int channels, const void* rgb, long rs, void* y, long ys, void* uv, long uvs)
{
    ColourPlanes pl{};
    for (int k = 0; k < channels; k++) pl.in[k] = { const_cast<void*>(rgb), rs, channels, k };
    pl.out[0] = { y, ys, 1, 0 };
    for (int k = 1; k < channels; k++) pl.out[k] = { uv, uvs, channels - 1, k - 1 };
    return pl;
}


template <int KIND, bool SPLIT>
static hipError_t colour_u8x4(const void* pk, long ps, void* y, long ys, void* c, long cs, int W, int H, hipStream_t stream)
{
    const dim3 grid((W + 1023) / 1024, H < 32768 ? H : 32768), block(256);
    hipLaunchKernelGGL((colour_u8x4_kernel<KIND, SPLIT>), grid, block, 0, stream, static_cast<const uint8_t*>(pk), ps,
                       static_cast<uint8_t*>(y), ys, static_cast<uint8_t*>(c), cs, W, H);
    return hipGetLastError();
}

hipError_t launch_color_split(int type, int channels, const void* src, long ss, int W, int H,
                              void* y, long ys, void* uv, long uvs, hipStream_t stream)
{
    if (channels != 3 && channels != 4) return hipErrorInvalidValue;
    if (W <= 0 || H <= 0) return hipSuccess;
    if (type == U8 && aligned4(src, ss) && aligned4(y, ys) && aligned4(uv, uvs))
        return channels == 4 ? colour_u8x4<kRgba2Yuva, true>(src, ss, y, ys, uv, uvs, W, H, stream)
                             : colour_u8x4<kRgb2Yuv, true>(src, ss, y, ys, uv, uvs, W, H, stream);
    return launch_colour(type, channels == 4 ? kRgba2Yuva : kRgb2Yuv, split_planes(channels, src, ss, y, ys, uv, uvs), W, H,
                         stream);
}

hipError_t launch_color_merge(int type, int channels, const void* y, long ys, const void* uv, long uvs, int W, int H,
                              void* dst, long ds, hipStream_t stream)
{
    if (channels != 3 && channels != 4) return hipErrorInvalidValue;
    if (W <= 0 || H <= 0) return hipSuccess;
    if (type == U8 && aligned4(dst, ds) && aligned4(y, ys) && aligned4(uv, uvs))
        return channels == 4 ? colour_u8x4<kYuva2Rgba, false>(dst, ds, const_cast<void*>(y), ys, const_cast<void*>(uv), uvs, W, H, stream)
                             : colour_u8x4<kYuv2Rgb, false>(dst, ds, const_cast<void*>(y), ys, const_cast<void*>(uv), uvs, W, H, stream);
    ColourPlanes pl{};
    pl.in[0] = { const_cast<void*>(y), ys, 1, 0 };
    for (int k = 1; k < channels; k++) pl.in[k] = { const_cast<void*>(uv), uvs, channels - 1, k - 1 };
    for (int k = 0; k < channels; k++) pl.out[k] = { dst, ds, channels, k };
    return launch_colour(type, channels == 4 ? kYuva2Rgba : kYuv2Rgb, pl, W, H, stream);
}

template <typename T>
static hipError_t resize_t(int channels, int mode, const void* src, long ss, int sw, int sh, void* dst, long ds, int dw,
                           int dh, int kmax, hipStream_t stream)
{
    if (const int k = upscale_factor(sw, sh, dw, dh)) {
        switch (channels) {
        case 1: return resize_up_t<T, 1, false>(k, mode, src, ss, sw, sh, dst, ds, dw, dh, nullptr, 0, stream);
        case 2: return resize_up_t<T, 2, false>(k, mode, src, ss, sw, sh, dst, ds, dw, dh, nullptr, 0, stream);
        case 3: return resize_up_t<T, 3, false>(k, mode, src, ss, sw, sh, dst, ds, dw, dh, nullptr, 0, stream);
        case 4: return resize_up_t<T, 4, false>(k, mode, src, ss, sw, sh, dst, ds, dw, dh, nullptr, 0, stream);
        default: return hipErrorInvalidValue;
        }
    }
    const dim3 grid((dw + kRsW - 1) / kRsW, (dh + kRsH - 1) / kRsH), block(256);
    const size_t lds = resize_lds_bytes(kmax);
    switch (channels) {
    case 1: hipLaunchKernelGGL((resize_kernel<T, 1>), grid, block, lds, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, kmax); break;
    case 2: hipLaunchKernelGGL((resize_kernel<T, 2>), grid, block, lds, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, kmax); break;
    case 3: hipLaunchKernelGGL((resize_kernel<T, 3>), grid, block, lds, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, kmax); break;
    case 4: hipLaunchKernelGGL((resize_kernel<T, 4>), grid, block, lds, stream, src, ss, sw, sh, dst, ds, dw, dh, mode, kmax); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t resize_merge_t(int channels, int mode, const void* y, long ys, const void* uv, long uvs, int sw, int sh,
                                 int W, int H, void* dst, long ds, int kmax, hipStream_t stream)
{
    if (const int k = upscale_factor(sw, sh, W, H)) {
        if (channels == 4) return resize_up_t<T, 3, true>(k, mode, uv, uvs, sw, sh, dst, ds, W, H, y, ys, stream);
        if (channels == 3) return resize_up_t<T, 2, true>(k, mode, uv, uvs, sw, sh, dst, ds, W, H, y, ys, stream);
        return hipErrorInvalidValue;
    }
    const dim3 grid((W + kRsW - 1) / kRsW, (H + kRsH - 1) / kRsH), block(256);
    const size_t lds = resize_lds_bytes(kmax);
    if (channels == 4)
        hipLaunchKernelGGL((resize_merge_kernel<T, 4>), grid, block, lds, stream, y, ys, uv, uvs, sw, sh, W, H, dst, ds, mode, kmax);
    else if (channels == 3)
        hipLaunchKernelGGL((resize_merge_kernel<T, 3>), grid, block, lds, stream, y, ys, uv, uvs, sw, sh, W, H, dst, ds, mode, kmax);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_resize_merge(int type, int channels, int mode, const void* y, long ys, const void* uv, long uvs, int sw,
                               int sh, int W, int H, void* dst, long ds, hipStream_t stream)
{
    if (mode < 0 || mode >= kResizeModeCount || sw <= 0 || sh <= 0) return hipErrorInvalidValue;
    if (W <= 0 || H <= 0) return hipSuccess;
    const int kmax = resize_kmax(mode, sw, sh, W, H);
    if (kmax < 0) return hipErrorInvalidValue;
    switch (type) {
    case U8: return resize_merge_t<uint8_t>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, kmax, stream);
    case U16: return resize_merge_t<uint16_t>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, kmax, stream);
    case F16: return resize_merge_t<__half>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, kmax, stream);
    case F32: return resize_merge_t<float>(channels, mode, y, ys, uv, uvs, sw, sh, W, H, dst, ds, kmax, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_resize(int type, int channels, int mode, const void* src, long ss, int sw, int sh,
                         void* dst, long ds, int dw, int dh, hipStream_t stream)
{
    if (mode < 0 || mode >= kResizeModeCount || sw <= 0 || sh <= 0) return hipErrorInvalidValue;
    if (dw <= 0 || dh <= 0) return hipSuccess;
    const int kmax = resize_kmax(mode, sw, sh, dw, dh);
    if (kmax < 0) return hipErrorInvalidValue;
    switch (type) {
    case U8: return resize_t<uint8_t>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, kmax, stream);
    case U16: return resize_t<uint16_t>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, kmax, stream);
    case F16: return resize_t<__half>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, kmax, stream);
    case F32: return resize_t<float>(channels, mode, src, ss, sw, sh, dst, ds, dw, dh, kmax, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace achip
