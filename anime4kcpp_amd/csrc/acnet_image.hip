// Element-wise image kernels behind the ac::core image helpers
// (core/src/ImageProcess.cpp:310-797): type conversion (astype, copy), integer shifts
// (shl, shr), pixelShuffle and per-pixel byte runs (extract, insert).  All are one pass
// over HBM; a thread handles one element (or one pixel's run), rows in blockIdx.y.
#include <algorithm>

#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace achip {

#pragma clang fp contract(off)

// ImageProcess.cpp:339-355: same type is a plain copy, otherwise fromFloat<OUT>(toFloat(in))
template <typename TI, typename TO>
__device__ inline TO convert_one(TI v)
{
    if constexpr (std::is_same_v<TI, TO>) return v;
    else return from_float<TO>(to_float<TI>(v));
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void convert_kernel(const void* __restrict__ src, long ss, void* __restrict__ dst, long ds,
                                                     int n, int h)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int y = blockIdx.y; y < h; y += gridDim.y)
        row_ptr<TO>(dst, ds, y)[i] = convert_one<TI, TO>(row_ptr<TI>(src, ss, y)[i]);
}

// a << n / a >> n on the integer promoted from an unsigned element (ImageProcess.cpp:
// 464-479, 515-530), stored back truncated to the element type; shifts of 32 or more give 0.
// src and dst may be the same buffer (each element is read before it is written).
template <typename T>
__global__ __launch_bounds__(256) void shift_kernel(const void* src, long ss, void* dst, long ds,
                                                   int n, int h, int left, int bits)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int y = blockIdx.y; y < h; y += gridDim.y) {
        const uint32_t v = row_ptr<T>(src, ss, y)[i];
        const uint32_t r = bits >= 32 ? 0u : (left ? v << bits : v >> bits);
        row_ptr<T>(dst, ds, y)[i] = static_cast<T>(r);
    }
}

// ImageProcess.cpp:358-385: output pixel (x*up + p%up, y*up + p/up), channel n, takes input
// channel n*up*up + p of pixel (x, y); integer copies of one type stay raw, everything
// else goes through toFloat/fromFloat (so float images are clamped to [0, 1]).
template <typename TI, typename TO>
__device__ inline TO shuffle_one(TI v)
{
    if constexpr (std::is_same_v<TI, TO> && std::is_integral_v<TI>) return v;
    else return from_float<TO>(to_float<TI>(v));
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void pixel_shuffle_kernel(const void* __restrict__ src, long ss, int sc, int up, int dw,
                                                           int dh, int dc, void* __restrict__ dst, long ds)
{
    const int X = blockIdx.x * 256 + threadIdx.x;
    if (X >= dw) return;
    const int x = X / up, dx = X - x * up;
    const int group = up * up;
    for (int Y = blockIdx.y; Y < dh; Y += gridDim.y) {
        const int y = Y / up, dy = Y - y * up;
        const int p = dy * up + dx;
        const TI* in = row_ptr<TI>(src, ss, y) + static_cast<long>(x) * sc;
        TO* out = row_ptr<TO>(dst, ds, Y) + static_cast<long>(X) * dc;
        for (int n = 0; n < dc; n++) out[n] = shuffle_one<TI, TO>(in[n * group + p]);
    }
}

__global__ __launch_bounds__(256) void copy_runs_kernel(const unsigned char* __restrict__ src, long ss, int spx, int soff,
                                                       unsigned char* __restrict__ dst, long ds, int dpx, int doff, int len,
                                                       int w, int h)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= w) return;
    for (long y = blockIdx.y; y < h; y += gridDim.y) {
        const unsigned char* s = src + y * ss + static_cast<long>(x) * spx + soff;
        unsigned char* d = dst + y * ds + static_cast<long>(x) * dpx + doff;
        for (int i = 0; i < len; i++) d[i] = s[i];
    }
}

// rows x units elements of T (the widest unit that every address, stride and row length is a
// multiple of); either side may be page-locked host memory seen through its device mapping.
// A small grid walks the elements with four loads in flight per thread: over PCIe the copy
// is latency-bound, so few workgroups keep the link busy and leave the CUs to the passes.
template <typename T>
__global__ __launch_bounds__(256) void copy_rows_kernel(const unsigned char* __restrict__ src, long ss,
                                                       unsigned char* __restrict__ dst, long ds, unsigned units,
                                                       unsigned total)
{
    const unsigned stride = gridDim.x * 256u;
    for (unsigned i0 = blockIdx.x * 256u + threadIdx.x; i0 < total; i0 += 4u * stride) {
        T v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned i = i0 + k * stride;
            if (i < total) {
                const unsigned y = i / units, u = i - y * units;
                v[k] = reinterpret_cast<const T*>(src + static_cast<long>(y) * ss)[u];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned i = i0 + k * stride;
            if (i < total) {
                const unsigned y = i / units, u = i - y * units;
                reinterpret_cast<T*>(dst + static_cast<long>(y) * ds)[u] = v[k];
            }
        }
    }
}

// ---- launchers ----------------------------------------------------------------------------

// rows beyond the grid's y extent are strided over inside the kernels
static dim3 row_grid(int n, int h) { return dim3((n + 255) / 256, h < 32768 ? h : 32768); }

template <typename TI>
static hipError_t convert_to(int dtype, const void* src, long ss, void* dst, long ds, int n, int h, hipStream_t stream)
{
    const dim3 g = row_grid(n, h), b(256);
    switch (dtype) {
    case U8: hipLaunchKernelGGL((convert_kernel<TI, uint8_t>), g, b, 0, stream, src, ss, dst, ds, n, h); break;
    case U16: hipLaunchKernelGGL((convert_kernel<TI, uint16_t>), g, b, 0, stream, src, ss, dst, ds, n, h); break;
    case F16: hipLaunchKernelGGL((convert_kernel<TI, __half>), g, b, 0, stream, src, ss, dst, ds, n, h); break;
    case F32: hipLaunchKernelGGL((convert_kernel<TI, float>), g, b, 0, stream, src, ss, dst, ds, n, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_convert(int stype, int dtype, const void* src, long ss, void* dst, long ds, int n, int h,
                          hipStream_t stream)
{
    if (n <= 0 || h <= 0) return hipSuccess;
    switch (stype) {
    case U8: return convert_to<uint8_t>(dtype, src, ss, dst, ds, n, h, stream);
    case U16: return convert_to<uint16_t>(dtype, src, ss, dst, ds, n, h, stream);
    case F16: return convert_to<__half>(dtype, src, ss, dst, ds, n, h, stream);
    case F32: return convert_to<float>(dtype, src, ss, dst, ds, n, h, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_shift(int type, int left, int bits, const void* src, long ss, void* dst, long ds, int n, int h,
                        hipStream_t stream)
{
    if (n <= 0 || h <= 0) return hipSuccess;
    if (bits < 0) return hipErrorInvalidValue;
    const dim3 g = row_grid(n, h), b(256);
    switch (type) {
    case U8: hipLaunchKernelGGL((shift_kernel<uint8_t>), g, b, 0, stream, src, ss, dst, ds, n, h, left, bits); break;
    case U16: hipLaunchKernelGGL((shift_kernel<uint16_t>), g, b, 0, stream, src, ss, dst, ds, n, h, left, bits); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename TI>
static hipError_t shuffle_to(int dtype, const void* src, long ss, int sc, int up, int dw, int dh, int dc, void* dst,
                             long ds, hipStream_t stream)
{
    const dim3 g = row_grid(dw, dh), b(256);
    switch (dtype) {
    case U8: hipLaunchKernelGGL((pixel_shuffle_kernel<TI, uint8_t>), g, b, 0, stream, src, ss, sc, up, dw, dh, dc, dst, ds); break;
    case U16: hipLaunchKernelGGL((pixel_shuffle_kernel<TI, uint16_t>), g, b, 0, stream, src, ss, sc, up, dw, dh, dc, dst, ds); break;
    case F16: hipLaunchKernelGGL((pixel_shuffle_kernel<TI, __half>), g, b, 0, stream, src, ss, sc, up, dw, dh, dc, dst, ds); break;
    case F32: hipLaunchKernelGGL((pixel_shuffle_kernel<TI, float>), g, b, 0, stream, src, ss, sc, up, dw, dh, dc, dst, ds); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pixel_shuffle(int stype, int dtype, const void* src, long ss, int sw, int sh, int sc, int up,
                                void* dst, long ds, hipStream_t stream)
{
    if (up <= 0 || sc <= 0 || sc % (up * up)) return hipErrorInvalidValue;
    const long dw = static_cast<long>(sw) * up, dh = static_cast<long>(sh) * up;
    if (dw <= 0 || dh <= 0) return hipSuccess;
    if (dh > (1L << 30) || dw > (1L << 30)) return hipErrorInvalidValue;
    const int dc = sc / (up * up);
    const int w = static_cast<int>(dw), h = static_cast<int>(dh);
    switch (stype) {
    case U8: return shuffle_to<uint8_t>(dtype, src, ss, sc, up, w, h, dc, dst, ds, stream);
    case U16: return shuffle_to<uint16_t>(dtype, src, ss, sc, up, w, h, dc, dst, ds, stream);
    case F16: return shuffle_to<__half>(dtype, src, ss, sc, up, w, h, dc, dst, ds, stream);
    case F32: return shuffle_to<float>(dtype, src, ss, sc, up, w, h, dc, dst, ds, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_copy_rows(const void* src, long ss, void* dst, long ds, long row_bytes, int rows, hipStream_t stream)
{
    if (row_bytes <= 0 || rows <= 0) return hipSuccess;
    const auto* s8 = static_cast<const unsigned char*>(src);
    auto* d8 = static_cast<unsigned char*>(dst);
    if (ss == row_bytes && ds == row_bytes) {   // dense planes: one row, so wide units fit more often
        row_bytes *= rows;
        rows = 1;
        ss = ds = row_bytes;
    }
    if (static_cast<unsigned long>(row_bytes) * rows >= (1ul << 31)) return hipErrorInvalidValue;   // keeps the unsigned index walk from wrapping
    auto go = [&](auto kernel, long unit, const unsigned char* s0, unsigned char* d0, long bytes) {
        const unsigned units = static_cast<unsigned>(bytes / unit), total = units * static_cast<unsigned>(rows);
        if (!total) return;
        const unsigned blocks = std::min<unsigned>(64u, (total + 1023u) / 1024u);
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, stream, s0, ss, d0, ds, units, total);
    };
    const unsigned long bases = reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst);
    const unsigned long steps = static_cast<unsigned long>(ss) | static_cast<unsigned long>(ds);
    if (rows == 1 && bases % 16 == 0) {   // one row: 16-byte units, then the last bytes
        const long head = row_bytes / 16 * 16;
        go(copy_rows_kernel<uint4>, 16, s8, d8, head);
        if (head < row_bytes) go(copy_rows_kernel<unsigned char>, 1, s8 + head, d8 + head, row_bytes - head);
    } else if ((bases | steps | static_cast<unsigned long>(row_bytes)) % 16 == 0) {
        go(copy_rows_kernel<uint4>, 16, s8, d8, row_bytes);
    } else if ((bases | steps | static_cast<unsigned long>(row_bytes)) % 4 == 0) {
        go(copy_rows_kernel<unsigned int>, 4, s8, d8, row_bytes);
    } else {
        go(copy_rows_kernel<unsigned char>, 1, s8, d8, row_bytes);
    }
    return hipGetLastError();
}

hipError_t launch_copy_runs(const void* src, long ss, int spx, int soff, void* dst, long ds, int dpx, int doff, int len,
                            int w, int h, hipStream_t stream)
{
    if (w <= 0 || h <= 0 || len <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_runs_kernel, row_grid(w, h), dim3(256), 0, stream, static_cast<const unsigned char*>(src), ss,
                       spx, soff, static_cast<unsigned char*>(dst), ds, dpx, doff, len, w, h);
    return hipGetLastError();
}

} // namespace achip
