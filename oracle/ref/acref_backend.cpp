// Reference-kernel driver for oracle/_ref -- test infrastructure only.
//
// Compiled once per ISA (BACKEND_HEADER / BACKEND_OP / BACKEND_SUFFIX set by
// oracle/Makefile) directly against the reference's own headers where they
// lie: core/internal/AC/Core/Internal/Processor/CPU/{Generic.hpp,X86/*.hpp}
// (the per-ISA dot/conv micro-kernels) and core/internal/AC/Core/Internal/Util.hpp
// (toFloat/fromFloat).  Those headers need nothing that is absent here.
//
// The per-pixel replicate-clamped loops of Common.hpp:116-197,344-393 cannot be
// included (Common.hpp pulls in Image.hpp, which needs a CMake-generated export
// header), so the three loops below restate them and call the reference kernels.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "AC/Core/Internal/Util.hpp"
#include BACKEND_HEADER

namespace {
using Op = ac::core::cpu::BACKEND_OP;

template <typename IN>
void layer0(const void* src, long sstride, int w, int h, float* out, const float* k, const float* b)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        auto row = [&](int y) { return reinterpret_cast<const IN*>(static_cast<const std::uint8_t*>(src) + y * sstride); };
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            alignas(64) const float r[] = {
                ac::core::toFloat(row(i - tp)[j - lp]), ac::core::toFloat(row(i - tp)[j]), ac::core::toFloat(row(i - tp)[j + rp]),
                ac::core::toFloat(row(i)[j - lp]),      ac::core::toFloat(row(i)[j]),      ac::core::toFloat(row(i)[j + rp]),
                ac::core::toFloat(row(i + bp)[j - lp]), ac::core::toFloat(row(i + bp)[j]), ac::core::toFloat(row(i + bp)[j + rp]),
            };
            float* o = out + (static_cast<long>(i) * w + j) * 8;
            Op::conv_cin1<8, 9>(r, o, k, b);
            for (int n = 0; n < 8; n++) o[n] = o[n] > 0.0f ? o[n] : 0.0f;   // ReLU: std::max(v, 0.0f)
        }
    }
}

void layer8(const float* in, int w, int h, float* out, const float* k, const float* b)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            auto px = [&](int y, int x) { return in + (static_cast<long>(y) * w + x) * 8; };
            const float* rptr[] = { px(i - tp, j - lp), px(i - tp, j), px(i - tp, j + rp),
                                    px(i, j - lp),      px(i, j),      px(i, j + rp),
                                    px(i + bp, j - lp), px(i + bp, j), px(i + bp, j + rp) };
            alignas(64) float sum[8]{};
            Op::conv<8, 8, 9>(rptr, sum, k, b);
            float* o = out + (static_cast<long>(i) * w + j) * 8;
            for (int n = 0; n < 8; n++) o[n] = sum[n] > 0.0f ? sum[n] : 0.0f;
        }
    }
}

template <typename OUT>
void last(const float* in, int w, int h, void* dst, long dstride, const float* k, const float* b, const float* k2)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            auto px = [&](int y, int x) { return in + (static_cast<long>(y) * w + x) * 8; };
            const float* rptr[] = { px(i - tp, j - lp), px(i - tp, j), px(i - tp, j + rp),
                                    px(i, j - lp),      px(i, j),      px(i, j + rp),
                                    px(i + bp, j - lp), px(i + bp, j), px(i + bp, j + rp) };
            alignas(64) float sum[8]{};
            Op::conv<8, 8, 9>(rptr, sum, k, b);
            for (int n = 0; n < 8; n++) sum[n] = sum[n] > 0.0f ? sum[n] : 0.0f;
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    auto o = reinterpret_cast<OUT*>(static_cast<std::uint8_t*>(dst) + (2 * i + dy) * dstride) + 2 * j + dx;
                    *o = ac::core::fromFloat<OUT>(Op::dot<8>(sum, k2 + 8 * (dy * 2 + dx)));
                }
        }
    }
}
// ---- ACNet<8> (CPUProcessor.cpp:1417-1436) ----

inline float prelu(float v, float a) { return std::max(v, 0.0f) + a * std::min(v, 0.0f); }   // Common.hpp:42-49

template <typename IN>
void a8_layer0(const void* src, long sstride, int w, int h, float* out, const float* k, const float* b, const float* al)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        auto row = [&](int y) { return reinterpret_cast<const IN*>(static_cast<const std::uint8_t*>(src) + y * sstride); };
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            alignas(64) const float r[] = {
                ac::core::toFloat(row(i - tp)[j - lp]), ac::core::toFloat(row(i - tp)[j]), ac::core::toFloat(row(i - tp)[j + rp]),
                ac::core::toFloat(row(i)[j - lp]),      ac::core::toFloat(row(i)[j]),      ac::core::toFloat(row(i)[j + rp]),
                ac::core::toFloat(row(i + bp)[j - lp]), ac::core::toFloat(row(i + bp)[j]), ac::core::toFloat(row(i + bp)[j + rp]),
            };
            float* o = out + (static_cast<long>(i) * w + j) * 8;
            Op::conv_cin1<8, 9>(r, o, k, b);
            for (int n = 0; n < 8; n++) o[n] = prelu(o[n], al[n]);
        }
    }
}

void a8_block(const float* in, int w, int h, float* out, const float* k, const float* b, const float* al)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            auto px = [&](int y, int x) { return in + (static_cast<long>(y) * w + x) * 8; };
            const float* rptr[] = { px(i - tp, j - lp), px(i - tp, j), px(i - tp, j + rp),
                                    px(i, j - lp),      px(i, j),      px(i, j + rp),
                                    px(i + bp, j - lp), px(i + bp, j), px(i + bp, j + rp) };
            alignas(64) float sum[8]{};
            Op::conv<8, 8, 9>(rptr, sum, k, b);
            float* o = out + (static_cast<long>(i) * w + j) * 8;
            for (int n = 0; n < 8; n++) o[n] = prelu(sum[n], al[n]);
        }
    }
}

// Common.hpp:290-341 with Identity and the source pixel as the nearest residual
template <typename IN, typename OUT>
void a8_last(const float* in, int w, int h, const void* src, long sstride, void* dst, long dstride, const float* k,
             const float* b)
{
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            auto px = [&](int y, int x) { return in + (static_cast<long>(y) * w + x) * 8; };
            const float* rptr[] = { px(i - tp, j - lp), px(i - tp, j), px(i - tp, j + rp),
                                    px(i, j - lp),      px(i, j),      px(i, j + rp),
                                    px(i + bp, j - lp), px(i + bp, j), px(i + bp, j + rp) };
            alignas(64) float sum[4]{};
            Op::conv<8, 4, 9>(rptr, sum, k, b);
            const float res = ac::core::toFloat(reinterpret_cast<const IN*>(static_cast<const std::uint8_t*>(src) + i * sstride)[j]);
            for (int n = 0; n < 4; n++) {
                const float v = sum[n] * 1.0f + res;
                auto o = reinterpret_cast<OUT*>(static_cast<std::uint8_t*>(dst) + (2 * i + (n >> 1)) * dstride) + 2 * j + (n & 1);
                *o = ac::core::fromFloat<OUT>(v);
            }
        }
    }
}
// ---- ArtCNN<F> (CPUProcessor.cpp:1519-1602) ----

template <int F>
struct ArtLayers {
    static const float** taps(const float* in, int w, int i, int j, int h, const float** rptr)
    {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0, lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
        auto px = [&](int y, int x) { return in + (static_cast<long>(y) * w + x) * F; };
        const float* t[] = { px(i - tp, j - lp), px(i - tp, j), px(i - tp, j + rp), px(i, j - lp), px(i, j), px(i, j + rp),
                             px(i + bp, j - lp), px(i + bp, j), px(i + bp, j + rp) };
        for (int p = 0; p < 9; p++) rptr[p] = t[p];
        return rptr;
    }
    template <typename IN>
    static void head(const void* src, long sstride, int w, int h, float* out, const float* k, const float* b)
    {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            auto row = [&](int y) { return reinterpret_cast<const IN*>(static_cast<const std::uint8_t*>(src) + y * sstride); };
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                alignas(64) const float r[] = {
                    ac::core::toFloat(row(i - tp)[j - lp]), ac::core::toFloat(row(i - tp)[j]), ac::core::toFloat(row(i - tp)[j + rp]),
                    ac::core::toFloat(row(i)[j - lp]),      ac::core::toFloat(row(i)[j]),      ac::core::toFloat(row(i)[j + rp]),
                    ac::core::toFloat(row(i + bp)[j - lp]), ac::core::toFloat(row(i + bp)[j]), ac::core::toFloat(row(i + bp)[j + rp]),
                };
                Op::conv_cin1<F, 9>(r, out + (static_cast<long>(i) * w + j) * F, k, b);
            }
        }
    }
    static void body(const float* in, int w, int h, float* out, const float* k, const float* b, const float* feat)
    {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                const float* rptr[9];
                alignas(64) float sum[F]{};
                Op::conv<F, F, 9>(taps(in, w, i, j, h, rptr), sum, k, b);
                float* o = out + (static_cast<long>(i) * w + j) * F;
                const float* fe = feat ? feat + (static_cast<long>(i) * w + j) * F : nullptr;
                for (int n = 0; n < F; n++) o[n] = fe ? sum[n] * 1.0f + fe[n] : std::max(sum[n], 0.0f);
            }
    }
    template <typename OUT>
    static void tail(const float* in, int w, int h, void* dst, long dstride, const float* k, const float* b)
    {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                const float* rptr[9];
                alignas(64) float sum[4]{};
                Op::conv<F, 4, 9>(taps(in, w, i, j, h, rptr), sum, k, b);
                for (int n = 0; n < 4; n++) {
                    auto o = reinterpret_cast<OUT*>(static_cast<std::uint8_t*>(dst) + (2 * i + (n >> 1)) * dstride) + 2 * j + (n & 1);
                    *o = ac::core::fromFloat<OUT>(sum[n]);
                }
            }
    }
    static int run(const void* src, long sstride, int w, int h, int type, void* dst, long dstride, const float* kernels,
                   const float* biases, int blocks)
    {
        std::vector<float> feat(static_cast<size_t>(w) * h * F), t1(feat.size()), t2(feat.size());
        switch (type) {
        case 0x001: head<std::uint8_t>(src, sstride, w, h, feat.data(), kernels, biases); break;
        case 0x002: head<std::uint16_t>(src, sstride, w, h, feat.data(), kernels, biases); break;
        case 0x204: head<float>(src, sstride, w, h, feat.data(), kernels, biases); break;
        default: return -22;
        }
        const float* in = feat.data();
        float* out = t1.data();
        for (int l = 1; l <= blocks + 1; l++) {
            body(in, w, h, out, kernels + F * 9 + static_cast<size_t>(F) * F * 9 * (l - 1), biases + F * l,
                 l == blocks + 1 ? feat.data() : nullptr);
            in = out;
            out = out == t1.data() ? t2.data() : t1.data();
        }
        const float* kt = kernels + F * 9 + static_cast<size_t>(F) * F * 9 * (blocks + 1);
        const float* bt = biases + F * (blocks + 2);
        switch (type) {
        case 0x001: tail<std::uint8_t>(in, w, h, dst, dstride, kt, bt); break;
        case 0x002: tail<std::uint16_t>(in, w, h, dst, dstride, kt, bt); break;
        case 0x204: tail<float>(in, w, h, dst, dstride, kt, bt); break;
        default: return -22;
        }
        return 0;
    }
};
// ---- FSRCNNX<F> (CPUProcessor.cpp:1629-1659) ----

template <int F>
struct FsrLayers {
    static float prelu(float v, float a) { return std::max(v, 0.0f) + a * std::min(v, 0.0f); }
    template <typename IN>
    static void head(const void* src, long sstride, int w, int h, float* out, const float* k, const float* b)
    {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < h; i++) {
            auto row = [&](int y) { return reinterpret_cast<const IN*>(static_cast<const std::uint8_t*>(src) + y * sstride); };
            for (int j = 0; j < w; j++) {
                alignas(64) float r[25];
                for (int a = 0; a < 5; a++)
                    for (int c = 0; c < 5; c++)
                        r[a * 5 + c] = ac::core::toFloat(row(std::clamp(i + a - 2, 0, h - 1))[std::clamp(j + c - 2, 0, w - 1)]);
                Op::conv_cin1<F, 25>(r, out + (static_cast<long>(i) * w + j) * F, k, b);
            }
        }
    }
    static void body(const float* in, int w, int h, float* out, const float* k, const float* b, const float* al,
                     const float* k1, const float* b1, const float* a1, const float* feat)
    {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                const float* rptr[9];
                alignas(64) float buf[F]{};
                Op::conv<F, F, 9>(ArtLayers<F>::taps(in, w, i, j, h, rptr), buf, k, b);
                for (int n = 0; n < F; n++) buf[n] = prelu(buf[n], al[n]);
                float* o = out + (static_cast<long>(i) * w + j) * F;
                if (!feat) {
                    for (int n = 0; n < F; n++) o[n] = buf[n];
                    continue;
                }
                rptr[0] = buf;
                alignas(64) float sum[F]{};
                Op::conv<F, F, 1>(rptr, sum, k1, b1);
                const float* fe = feat + (static_cast<long>(i) * w + j) * F;
                for (int n = 0; n < F; n++) o[n] = prelu(sum[n] * 1.0f + fe[n], a1[n]);
            }
    }
    static int run(const void* src, long sstride, int w, int h, int type, void* dst, long dstride, const float* kernels,
                   const float* biases, const float* alphas, int blocks)
    {
        std::vector<float> feat(static_cast<size_t>(w) * h * F), t1(feat.size()), t2(feat.size());
        switch (type) {
        case 0x001: head<std::uint8_t>(src, sstride, w, h, feat.data(), kernels, biases); break;
        case 0x002: head<std::uint16_t>(src, sstride, w, h, feat.data(), kernels, biases); break;
        case 0x204: head<float>(src, sstride, w, h, feat.data(), kernels, biases); break;
        default: return -22;
        }
        const float* in = feat.data();
        float* out = t1.data();
        const float* k1 = kernels + F * 25 + static_cast<size_t>(F) * F * 9 * blocks;
        for (int l = 1; l <= blocks; l++) {
            body(in, w, h, out, kernels + F * 25 + static_cast<size_t>(F) * F * 9 * (l - 1), biases + F * l,
                 alphas + F * (l - 1), k1, biases + F * (blocks + 1), alphas + F * blocks,
                 l == blocks ? feat.data() : nullptr);
            in = out;
            out = out == t1.data() ? t2.data() : t1.data();
        }
        const float* kt = k1 + F * F;
        const float* bt = biases + F * (blocks + 2);
        switch (type) {
        case 0x001: ArtLayers<F>::template tail<std::uint8_t>(in, w, h, dst, dstride, kt, bt); break;
        case 0x002: ArtLayers<F>::template tail<std::uint16_t>(in, w, h, dst, dstride, kt, bt); break;
        case 0x204: ArtLayers<F>::template tail<float>(in, w, h, dst, dstride, kt, bt); break;
        default: return -22;
        }
        return 0;
    }
};
} // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

// src_type/dst_type: 0x001 u8, 0x002 u16, 0x204 f32 (core/include/AC/Core/Image.hpp:367-371)
extern "C" int CAT(acref_pass_, BACKEND_SUFFIX)(const void* src, long sstride, int w, int h, int stype,
                                                void* dst, long dstride, int dtype,
                                                const float* kernels, const float* biases)
{
    std::vector<float> t1(static_cast<size_t>(w) * h * 8), t2(static_cast<size_t>(w) * h * 8);
    switch (stype) {   // CPUProcessor.cpp:1384: conv3x3_1to8_relu(src, tmpO, kernel(0), bias(0))
    case 0x001: layer0<std::uint8_t>(src, sstride, w, h, t1.data(), kernels, biases); break;
    case 0x002: layer0<std::uint16_t>(src, sstride, w, h, t1.data(), kernels, biases); break;
    case 0x204: layer0<float>(src, sstride, w, h, t1.data(), kernels, biases); break;
    default: return -22;
    }
    float* ti = t1.data();
    float* to = t2.data();
    for (int l = 1; l < 8; l++) {   // :1386-1390 blocks()-1 middle layers
        layer8(ti, w, h, to, kernels + 72 + 576 * (l - 1), biases + 8 * l);
        std::swap(ti, to);
    }
    const float* k8 = kernels + 72 + 576 * 7;
    const float* b8 = biases + 64;
    const float* k9 = kernels + 72 + 576 * 8;
    switch (dtype) {
    case 0x001: last<std::uint8_t>(ti, w, h, dst, dstride, k8, b8, k9); break;
    case 0x002: last<std::uint16_t>(ti, w, h, dst, dstride, k8, b8, k9); break;
    case 0x204: last<float>(ti, w, h, dst, dstride, k8, b8, k9); break;
    default: return -22;
    }
    return 0;
}

// ACNet<8>: same element types as acref_pass; src and dst of one type (the residual is read
// in that type, as Common.hpp:330 reads it in the output type)
extern "C" int CAT(acref_pass8_, BACKEND_SUFFIX)(const void* src, long sstride, int w, int h, int type, void* dst,
                                                 long dstride, const float* kernels, const float* biases,
                                                 const float* alphas, int blocks)
{
    std::vector<float> t1(static_cast<size_t>(w) * h * 8), t2(static_cast<size_t>(w) * h * 8);
    switch (type) {
    case 0x001: a8_layer0<std::uint8_t>(src, sstride, w, h, t1.data(), kernels, biases, alphas); break;
    case 0x002: a8_layer0<std::uint16_t>(src, sstride, w, h, t1.data(), kernels, biases, alphas); break;
    case 0x204: a8_layer0<float>(src, sstride, w, h, t1.data(), kernels, biases, alphas); break;
    default: return -22;
    }
    float* ti = t1.data();
    float* to = t2.data();
    for (int l = 1; l <= blocks; l++) {
        a8_block(ti, w, h, to, kernels + 72 + 576 * (l - 1), biases + 8 * l, alphas + 8 * l);
        std::swap(ti, to);
    }
    const float* kf = kernels + 72 + 576 * blocks;
    const float* bf = biases + 8 + 8 * blocks;
    switch (type) {
    case 0x001: a8_last<std::uint8_t, std::uint8_t>(ti, w, h, src, sstride, dst, dstride, kf, bf); break;
    case 0x002: a8_last<std::uint16_t, std::uint16_t>(ti, w, h, src, sstride, dst, dstride, kf, bf); break;
    case 0x204: a8_last<float, float>(ti, w, h, src, sstride, dst, dstride, kf, bf); break;
    default: return -22;
    }
    return 0;
}

extern "C" int CAT(acref_artcnn_, BACKEND_SUFFIX)(const void* src, long sstride, int w, int h, int type, void* dst,
                                                  long dstride, const float* kernels, const float* biases, int f,
                                                  int blocks)
{
    if (f == 16) return ArtLayers<16>::run(src, sstride, w, h, type, dst, dstride, kernels, biases, blocks);
    if (f == 32) return ArtLayers<32>::run(src, sstride, w, h, type, dst, dstride, kernels, biases, blocks);
    return -22;
}

extern "C" int CAT(acref_fsrcnnx_, BACKEND_SUFFIX)(const void* src, long sstride, int w, int h, int type, void* dst,
                                                   long dstride, const float* kernels, const float* biases,
                                                   const float* alphas, int f, int blocks)
{
    if (f == 8) return FsrLayers<8>::run(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, blocks);
    if (f == 16) return FsrLayers<16>::run(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, blocks);
    return -22;
}
