// Device-side helpers shared by the ACNetLegacy HIP kernels (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace achip {

// Element types, encoded like ac::core::Image::ElementType
// (core/include/AC/Core/Image.hpp:367-371): kind << 8 | bytes.
enum ElemType : int { U8 = 0x001, U16 = 0x002, F16 = 0x202, F32 = 0x204 };

// Device weight blob: the reference arrays (ACNet.hpp:34-60) repacked on upload so
// the kernels' uniform weight reads are contiguous scalar loads.
//   L0     w[p][n]    72 floats, then bias[8]
//   L1..L8 w[p][c][n] 576 floats, then bias[8]          (p = 3*ky + kx)
//   L9     k2[c][q]   32 floats                          (q = 2*dy + dx)
constexpr int kOffW0 = 0;
constexpr int kOffB0 = 72;
constexpr int kLayerStride = 584;
__host__ __device__ constexpr int offW(int l) { return 80 + (l - 1) * kLayerStride; }   // l in 1..8
__host__ __device__ constexpr int offB(int l) { return offW(l) + 576; }
constexpr int kOffK2 = 80 + 8 * kLayerStride;     // 4752
constexpr int kBlobFloats = kOffK2 + 32;         // 4784

// MFMA A operands for layers 1..8, one blob per 16-bit type: [layer 1..8][t 0..2][lane 0..63]
// x 8 values.  Lane l of MFMA t holds row i = l % 16 (output channel i % 8 of output row
// y + i / 8) and K block g = l / 16 (patch row g, patch column t, channels 0..7):
// the weight of tap (ky = g - i / 8, kx = t) when 0 <= ky <= 2, else 0.
constexpr int kMfmaLayerHalves = 3 * 64 * 8;
constexpr int kMfmaBlobBytes = 8 * kMfmaLayerHalves * 2;   // 24576

// FP32 MFMA (16x16x4) A operands: layer 0 as [t 0..2][lane] (K = patch row g = lane/16,
// one channel), then layers 1..8 as [layer][s = 8*t + c][lane]: lane l holds row
// i = l % 16 (channel i % 8 of output row y + i / 8) at K = patch row g, patch column t,
// channel c -- the weight of tap (ky = g - i / 8, kx = t) when 0 <= ky <= 2, else 0.
constexpr int kMfma32L0Floats = 3 * 64;
constexpr int kMfma32LayerFloats = 24 * 64;
constexpr int kMfma32BlobFloats = kMfma32L0Floats + 8 * kMfma32LayerFloats;   // 12480

// ACNet<8> device blob (model arrays of core/include/AC/Core/Model/ACNet.hpp:63-118,
// repacked): L0 weights w0[p][n] (72), bias (8), PReLU slopes (8); the final layer's 4
// biases (padded to 8); then per block l = 1..B its 8 biases and 8 slopes.  The conv weights
// of the blocks and of the final 8 -> 4 layer (output channels 4-7 zero) are FP32 MFMA
// per-lane A operands, kMfma32LayerFloats per layer (layout of the legacy layers 1..8).
constexpr int kA8OffW0 = 0, kA8OffB0 = 72, kA8OffAlpha0 = 80, kA8OffFinalB = 88, kA8OffLayer = 96;
__host__ __device__ constexpr int acnet8_blob_floats(int blocks) { return kA8OffLayer + 16 * blocks; }

// toFloat (core/internal/AC/Core/Internal/Util.hpp:49-61): unsigned -> v / max, the
// correctly rounded quotient as the host's IEEE divide.  Computed as q0 = v * RN(1/max),
// one FMA residual and one FMA correction: for every one of the 256 (u8) and 65536 (u16)
// inputs this equals RN(v / max) exactly, at a fraction of a full-precision divide.
__device__ inline float div_exact(float v, float d, float r)
{
    const float q0 = __fmul_rn(v, r);
    return __fmaf_rn(__fmaf_rn(-q0, d, v), r, q0);
}
template <typename T> __device__ inline float to_float(T v);
template <> __device__ inline float to_float<uint8_t>(uint8_t v) { return div_exact(static_cast<float>(v), 255.0f, 1.0f / 255.0f); }
template <> __device__ inline float to_float<uint16_t>(uint16_t v) { return div_exact(static_cast<float>(v), 65535.0f, 1.0f / 65535.0f); }
template <> __device__ inline float to_float<__half>(__half v) { return __half2float(v); }
template <> __device__ inline float to_float<float>(float v) { return v; }

// fromFloat (Util.hpp:63-78): saturate to [0,1]; unsigned types then (T)(s*max + 0.5f)
// with the multiply and the add rounded separately and the conversion truncating.
template <typename T> __device__ inline T from_float(float v);
__device__ inline float saturate01(float v) { return v < 0.0f ? 0.0f : (v < 1.0f ? v : 1.0f); }
// (unsigned)(s * 255 + 0.5f) for s in [0, 1], the multiply and the add rounded separately,
// computed as one FMA: the truncated results are equal for every float s in [0, 1] (all
// 1,065,353,217 of them compared on the CPU; the windows where the two roundings could
// part are checked in tests/test_device_math.py).
__device__ inline unsigned u8_scale_code(float s) { return static_cast<unsigned>(__fmaf_rn(s, 255.0f, 0.5f)); }
template <> __device__ inline uint8_t from_float<uint8_t>(float v)
{
    return static_cast<uint8_t>(u8_scale_code(saturate01(v)));
}
template <> __device__ inline uint16_t from_float<uint16_t>(float v)
{
    return static_cast<uint16_t>(static_cast<unsigned>(__fadd_rn(__fmul_rn(saturate01(v), 65535.0f), 0.5f)));
}
template <> __device__ inline __half from_float<__half>(float v) { return __float2half_rn(saturate01(v)); }
template <> __device__ inline float from_float<float>(float v) { return saturate01(v); }

template <typename T>
__device__ inline const T* row_ptr(const void* base, long stride, int y)
{
    return reinterpret_cast<const T*>(static_cast<const unsigned char*>(base) + static_cast<long>(y) * stride);
}
template <typename T>
__device__ inline T* row_ptr(void* base, long stride, int y)
{
    return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + static_cast<long>(y) * stride);
}

// Two horizontally adjacent output pixels: one store of twice the element size when the
// pair is aligned for it (whole cache lines per wave instead of byte-interleaved partial
// writes), else two stores.
template <typename T> struct Bits;
template <> struct Bits<uint8_t> { using one = uint8_t; using two = uint16_t; };
template <> struct Bits<uint16_t> { using one = uint16_t; using two = uint32_t; };
template <> struct Bits<__half> { using one = uint16_t; using two = uint32_t; };
template <> struct Bits<float> { using one = uint32_t; using two = uint64_t; };
template <typename T>
__device__ inline void store_pair(T* p, float a, float b)
{
    using One = typename Bits<T>::one;
    using Two = typename Bits<T>::two;
    const T va = from_float<T>(a), vb = from_float<T>(b);
    if ((reinterpret_cast<uintptr_t>(p) & (2 * sizeof(T) - 1)) == 0) {
        *reinterpret_cast<Two*>(p) = static_cast<Two>(__builtin_bit_cast(One, va)) |
                                     (static_cast<Two>(__builtin_bit_cast(One, vb)) << (8 * sizeof(T)));
    } else {
        p[0] = va;
        p[1] = vb;
    }
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

} // namespace achip
