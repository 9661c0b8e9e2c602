/*
 * ORACLE -- test infrastructure only (see acnet_oracle.h for the pins).
 *
 * CPU restatement of the reference ACNetLegacy 2x pass:
 *   CPUProcessor<ACNetLegacy>::process   core/src/processor/cpu/CPUProcessor.cpp:1371-1390
 *   conv3x3_1to8_relu                    core/internal/.../CPU/Backend.hpp:17-34, Common.hpp:166-197
 *   conv3x3_8to8_relu                    Backend.hpp:35-38, Common.hpp:116-164
 *   conv3x3_8to8_relu_deconv2x2_8to1     Backend.hpp:39-56, Common.hpp:344-393
 *   summation orders                     Generic.hpp:12-82, X86/AVX.hpp:14-150 (FMA),
 *                                        X86/AVX512.hpp:56-110 (falls back to the 256-bit
 *                                        FMA path for 8 channels and 9 taps)
 *   toFloat / fromFloat                  core/internal/AC/Core/Internal/Util.hpp:49-78
 *   weight offsets                       core/include/AC/Core/Model/ACNet.hpp:34-60
 *
 * Build with -ffp-contract=off: every multiply-add that is not an explicit
 * fmaf() below is rounded twice, as in the reference's intrinsics.
 */
#include "acnet_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NB 8          /* blocks (core/src/Model.cpp:10) */
#define CH 8          /* feature channels */

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---- element conversion (Util.hpp:49-78) ------------------------------------------ */

float orc_half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1f;
    uint32_t man = h & 0x3ffu;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) bits = sign;
        else {                       /* subnormal: value = man * 2^-24 */
            float f = (float)man * 5.9604644775390625e-08f;
            memcpy(&bits, &f, 4);
            bits |= sign;
        }
    } else if (exp == 31) bits = sign | 0x7f800000u | (man << 13);
    else bits = sign | ((exp + 112u) << 23) | (man << 13);
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* IEEE round-to-nearest-even float -> binary16 */
uint16_t orc_float_to_half(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          /* rounds to inf */
    if (ax < 0x38800000u) {                                             /* subnormal / zero */
        float a;
        memcpy(&a, &ax, 4);
        /* a * 2^24 is exact in float range; round half to even */
        float s = a * 16777216.0f;
        float r = nearbyintf(s);
        return (uint16_t)(sign | (uint32_t)r);
    }
    uint32_t mant = ax & 0x7fffffu;
    uint32_t e = (ax >> 23) - 112u;
    uint32_t hv = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (hv & 1u))) hv++;
    return (uint16_t)(sign | hv);
}

float orc_to_float(const void* p, int type)
{
    switch (type) {
    case ORC_U8: return (float)(*(const uint8_t*)p) / 255.0f;
    case ORC_U16: return (float)(*(const uint16_t*)p) / 65535.0f;
    case ORC_F16: return orc_half_to_float(*(const uint16_t*)p);
    default: return *(const float*)p;
    }
}

void orc_from_float(float v, void* p, int type)
{
    float s = v < 0.0f ? 0.0f : (v < 1.0f ? v : 1.0f);
    switch (type) {
    case ORC_U8: *(uint8_t*)p = (uint8_t)(s * 255.0f + 0.5f); break;
    case ORC_U16: *(uint16_t*)p = (uint16_t)(s * 65535.0f + 0.5f); break;
    case ORC_F16: *(uint16_t*)p = orc_float_to_half(s); break;
    default: *(float*)p = s; break;
    }
}

static inline int elem_size(int type) { return type & 0xff; }

/* ---- the two reference summation orders ------------------------------------------- */

/* AVX.hpp:24-30: ((s0+s4)+(s2+s6)) + ((s1+s5)+(s3+s7)) */
static inline float hsum8(const float* s)
{
    float a0 = s[0] + s[4], a1 = s[1] + s[5], a2 = s[2] + s[6], a3 = s[3] + s[7];
    float b0 = a0 + a2, b1 = a1 + a3;
    return b0 + b1;
}

/* L0: OpImplFMA::conv_cin1<8,9> (AVX.hpp:93-124): fma over taps 0..7, hsum, then
 * the 9th tap as a separate multiply and add, then + bias. */
static inline void conv_cin1_fma(const float* r, const float* k, const float* b, float* out)
{
    for (int n = 0; n < CH; n++) {
        const float* kp = k + n * 9;
        float s[8];
        for (int p = 0; p < 8; p++) s[p] = fmaf(r[p], kp[p], 0.0f);
        float sum = hsum8(s);
        float t = r[8] * kp[8];
        sum = sum + t;
        out[n] = sum + b[n];
    }
}

/* L0: OpImplGeneric::conv_cin1 (Generic.hpp:60-69): left fold of 9 products, + bias */
static inline void conv_cin1_generic(const float* r, const float* k, const float* b, float* out)
{
    for (int n = 0; n < CH; n++) {
        const float* kp = k + n * 9;
        float sum = r[0] * kp[0];
        for (int p = 1; p < 9; p++) { float t = r[p] * kp[p]; sum = sum + t; }
        out[n] = sum + b[n];
    }
}

/* 8->8: OpImplFMA::conv<8,8,9> (AVX.hpp:32-59,126-146): per output, one 8-lane
 * accumulator over channels, fma over taps 0..8, hsum, added to the bias. */
static inline void convn_fma(const float* const* rp, const float* k, const float* b, float* out, int cout)
{
    for (int n = 0; n < cout; n++) {
        const float* kn = k + n * 72;
        float s[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int p = 0; p < 9; p++)
            for (int c = 0; c < 8; c++) s[c] = fmaf(rp[p][c], kn[p * 8 + c], s[c]);
        out[n] = b[n] + hsum8(s);
    }
}

/* 8->8: OpImplGeneric::conv (Generic.hpp:71-81): sum over channels (outer) of
 * the left-folded 9-tap block sums, + bias. */
static inline void convn_generic(const float* const* rp, const float* k, const float* b, float* out, int cout)
{
    for (int n = 0; n < cout; n++) {
        const float* kn = k + n * 72;
        float sum = 0.0f;
        for (int c = 0; c < 8; c++) {
            float bs = rp[0][c] * kn[c];
            for (int p = 1; p < 9; p++) { float t = rp[p][c] * kn[p * 8 + c]; bs = bs + t; }
            sum = sum + bs;
        }
        out[n] = sum + b[n];
    }
}

static inline void conv8_fma(const float* const* rp, const float* k, const float* b, float* out) { convn_fma(rp, k, b, out, CH); }
static inline void conv8_generic(const float* const* rp, const float* k, const float* b, float* out) { convn_generic(rp, k, b, out, CH); }

/* deconv dot<8>: FMA path (AVX.hpp:62-91) = 0 + hsum(products); Generic = left fold */
static inline float dot8_fma(const float* v, const float* k)
{
    float s[8];
    for (int c = 0; c < 8; c++) s[c] = fmaf(v[c], k[c], 0.0f);
    return 0.0f + hsum8(s);
}
static inline float dot8_generic(const float* v, const float* k)
{
    float sum = v[0] * k[0];
    for (int c = 1; c < 8; c++) { float t = v[c] * k[c]; sum = sum + t; }
    return sum;
}

/* ---- the pass (CPUProcessor.cpp:1371-1390) ---------------------------------------- */

int orc_acnet_pass(const void* src, long src_stride, int w, int h, int src_type,
                   void* dst, long dst_stride, int dst_type,
                   const float* kernels, const float* biases, int order, int threads)
{
    if (!src || !dst || w <= 0 || h <= 0 || !kernels || !biases) return -22;
    size_t plane = (size_t)w * h * CH;
    float* t1 = (float*)malloc(plane * sizeof(float));
    float* t2 = (float*)malloc(plane * sizeof(float));
    if (!t1 || !t2) { free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    const int fma_order = (order == ORC_ORDER_FMA);
    const unsigned char* sb = (const unsigned char*)src;
    const int ses = elem_size(src_type);

    /* L0: conv3x3 1->8 + ReLU, replicate-clamped taps (Common.hpp:166-197) */
    {
        const float* k = kernels;          /* kernelOffset(0) = 0 */
        const float* b = biases;           /* biasOffset(0) = 0 */
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
                float r[9];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++)
                        r[a * 3 + c] = orc_to_float(sb + ys[a] * src_stride + (long)xs[c] * ses, src_type);
                float* out = t1 + ((size_t)i * w + j) * CH;
                if (fma_order) conv_cin1_fma(r, k, b, out); else conv_cin1_generic(r, k, b, out);
                for (int n = 0; n < CH; n++) out[n] = out[n] > 0.0f ? out[n] : 0.0f;   /* max(v, 0) */
            }
        }
    }

    float* tin = t1;
    float* tout = t2;
    /* L1..L7: conv3x3 8->8 + ReLU (Common.hpp:116-164) */
    for (int l = 1; l < NB; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        const float* b = biases + 8 * l;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
                const float* rpp[9];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++)
                        rpp[a * 3 + c] = tin + ((size_t)ys[a] * w + xs[c]) * CH;
                float* out = tout + ((size_t)i * w + j) * CH;
                if (fma_order) conv8_fma(rpp, k, b, out); else conv8_generic(rpp, k, b, out);
                for (int n = 0; n < CH; n++) out[n] = out[n] > 0.0f ? out[n] : 0.0f;
            }
        }
        float* tmp = tin; tin = tout; tout = tmp;
    }

    /* L8 + L9: conv3x3 8->8 + ReLU, then deconv 2x2 stride 2 8->1 (Common.hpp:344-393) */
    {
        const float* k = kernels + 72 + 576 * (NB - 1);     /* layer 8 */
        const float* b = biases + 8 * NB;
        const float* k2 = kernels + 72 + 576 * NB;          /* layer 9: offset 4680 */
        unsigned char* db = (unsigned char*)dst;
        const int des = elem_size(dst_type);
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
                const float* rpp[9];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++)
                        rpp[a * 3 + c] = tin + ((size_t)ys[a] * w + xs[c]) * CH;
                float s[CH];
                if (fma_order) conv8_fma(rpp, k, b, s); else conv8_generic(rpp, k, b, s);
                for (int n = 0; n < CH; n++) s[n] = s[n] > 0.0f ? s[n] : 0.0f;
                for (int dy = 0; dy < 2; dy++)
                    for (int dx = 0; dx < 2; dx++) {
                        const float* kk = k2 + 8 * (dy * 2 + dx);
                        float v = fma_order ? dot8_fma(s, kk) : dot8_generic(s, kk);
                        orc_from_float(v, db + (long)(2 * i + dy) * dst_stride + (long)(2 * j + dx) * des, dst_type);
                    }
            }
        }
    }
    free(t1);
    free(t2);
    return 0;
}

/* ---- ACNet<8> (CPUProcessor.cpp:1417-1436) ---------------------------------------- */

/* PReLU (Common.hpp:42-49): max(v, 0) + alpha * min(v, 0) */
static inline float prelu_(float v, float a) { return (v > 0.0f ? v : 0.0f) + a * (v < 0.0f ? v : 0.0f); }

int orc_acnet8_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                    const float* kernels, const float* biases, const float* alphas, int blocks, int order, int threads)
{
    if (!src || !dst || w <= 0 || h <= 0 || !kernels || !biases || !alphas || blocks < 1) return -22;
    size_t plane = (size_t)w * h * CH;
    float* t1 = (float*)malloc(plane * sizeof(float));
    float* t2 = (float*)malloc(plane * sizeof(float));
    if (!t1 || !t2) { free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    const int fma_order = (order == ORC_ORDER_FMA);
    const unsigned char* sb = (const unsigned char*)src;
    const int es = elem_size(type);

    /* L0: conv3x3 1->8 + PReLU (Backend.hpp:58-75, Common.hpp:166-197) */
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++) {
        int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
        for (int j = 0; j < w; j++) {
            int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
            int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
            float r[9];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++)
                    r[a * 3 + c] = orc_to_float(sb + ys[a] * src_stride + (long)xs[c] * es, type);
            float* out = t1 + ((size_t)i * w + j) * CH;
            if (fma_order) conv_cin1_fma(r, kernels, biases, out); else conv_cin1_generic(r, kernels, biases, out);
            for (int n = 0; n < CH; n++) out[n] = prelu_(out[n], alphas[n]);
        }
    }
    float* tin = t1;
    float* tout = t2;
    /* blocks: conv3x3 8->8 + PReLU (Backend.hpp:95-98, Common.hpp:116-164) */
    for (int l = 1; l <= blocks; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        const float* b = biases + 8 * l;
        const float* al = alphas + 8 * l;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
                const float* rpp[9];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) rpp[a * 3 + c] = tin + ((size_t)ys[a] * w + xs[c]) * CH;
                float* out = tout + ((size_t)i * w + j) * CH;
                if (fma_order) convn_fma(rpp, k, b, out, CH); else convn_generic(rpp, k, b, out, CH);
                for (int n = 0; n < CH; n++) out[n] = prelu_(out[n], al[n]);
            }
        }
        float* tmp = tin; tin = tout; tout = tmp;
    }
    /* final: conv3x3 8->4, identity, + the source pixel, pixel shuffle 2x (Backend.hpp:116-133,
     * Common.hpp:290-341); output channel n goes to (2x + n % 2, 2y + n / 2) */
    {
        const float* k = kernels + 72 + 576 * blocks;
        const float* b = biases + 8 + 8 * blocks;
        unsigned char* db = (unsigned char*)dst;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0;
            for (int j = 0; j < w; j++) {
                int lp = j > 0 ? 1 : 0, rp = j < w - 1 ? 1 : 0;
                int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rp };
                const float* rpp[9];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) rpp[a * 3 + c] = tin + ((size_t)ys[a] * w + xs[c]) * CH;
                float s[4];
                if (fma_order) convn_fma(rpp, k, b, s, 4); else convn_generic(rpp, k, b, s, 4);
                const float res = orc_to_float(sb + (long)i * src_stride + (long)j * es, type);
                for (int n = 0; n < 4; n++) {
                    const float v = s[n] * 1.0f + res;
                    orc_from_float(v, db + (long)(2 * i + (n >> 1)) * dst_stride + (long)(2 * j + (n & 1)) * es, type);
                }
            }
        }
    }
    free(t1);
    free(t2);
    return 0;
}

/* ---- ArtCNN<F> (CPUProcessor.cpp:1519-1602) --------------------------------------- */

/* conv_cin1<cout, cpos> (cpos = 9 or 25 taps of one channel).  FMA order (AVX.hpp:93-124):
 * an 8-lane accumulator over the first cpos - cpos % 8 taps, its hsum, then the remaining
 * taps added one by one, then the bias; Generic order (Generic.hpp:60-69): a left fold of
 * the tap products, then the bias. */
static void cin1_fma_n(const float* r, const float* k, const float* b, float* out, int cout, int cpos)
{
    const int count = cpos / 8, remain = cpos % 8;
    for (int n = 0; n < cout; n++) {
        const float* kp = k + n * cpos;
        float s[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int idx = 0; idx < count; idx++)
            for (int l = 0; l < 8; l++) s[l] = fmaf(r[idx * 8 + l], kp[idx * 8 + l], s[l]);
        float sum = hsum8(s);
        for (int i = 0; i < remain; i++) { float t = r[count * 8 + i] * kp[count * 8 + i]; sum = sum + t; }
        out[n] = sum + b[n];
    }
}
static void cin1_generic_n(const float* r, const float* k, const float* b, float* out, int cout, int cpos)
{
    for (int n = 0; n < cout; n++) {
        const float* kp = k + n * cpos;
        float sum = r[0] * kp[0];
        for (int p = 1; p < cpos; p++) { float t = r[p] * kp[p]; sum = sum + t; }
        out[n] = sum + b[n];
    }
}
/* conv<cin, cout, cpos> (cin a multiple of 8; cpos = 9 taps or 1): FMA order (AVX.hpp:32-59,
 * 126-146): an 8-lane accumulator, lane l taking channels l, l + 8, ... of every tap in tap
 * order, its hsum added to the bias; Generic order (Generic.hpp:71-81): per channel the
 * left-folded tap sum, channels summed in order, then the bias */
static void conv_fma_cin(const float* const* rp, const float* k, const float* b, float* out, int cin, int cout, int cpos)
{
    for (int n = 0; n < cout; n++) {
        const float* kn = k + (long)n * cin * cpos;
        float s[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int p = 0; p < cpos; p++)
            for (int idx = 0; idx < cin / 8; idx++)
                for (int l = 0; l < 8; l++) s[l] = fmaf(rp[p][idx * 8 + l], kn[cin * p + idx * 8 + l], s[l]);
        out[n] = b[n] + hsum8(s);
    }
}
static void conv_generic_cin(const float* const* rp, const float* k, const float* b, float* out, int cin, int cout, int cpos)
{
    for (int n = 0; n < cout; n++) {
        const float* kn = k + (long)n * cin * cpos;
        float sum = 0.0f;
        for (int c = 0; c < cin; c++) {
            float bs = rp[0][c] * kn[c];
            for (int p = 1; p < cpos; p++) { float t = rp[p][c] * kn[cin * p + c]; bs = bs + t; }
            sum = sum + bs;
        }
        out[n] = sum + b[n];
    }
}

/* one conv3x3 layer over an F-channel map with the clamped taps of Common.hpp:120-143 */
#define ORC_TAPS(MAP, F)                                                                   \
    int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0, lp = j > 0 ? 1 : 0, rpp_ = j < w - 1 ? 1 : 0; \
    int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rpp_ };                    \
    const float* rp9[9];                                                                   \
    for (int a = 0; a < 3; a++)                                                            \
        for (int c = 0; c < 3; c++) rp9[a * 3 + c] = (MAP) + ((size_t)ys[a] * w + xs[c]) * (F);

int orc_artcnn_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                    const float* kernels, const float* biases, int f, int blocks, int order, int threads)
{
    if (!src || !dst || w <= 0 || h <= 0 || !kernels || !biases || (f != 16 && f != 32) || blocks < 1) return -22;
    const size_t plane = (size_t)w * h * f;
    float* feat = (float*)malloc(plane * sizeof(float));
    float* t1 = (float*)malloc(plane * sizeof(float));
    float* t2 = (float*)malloc(plane * sizeof(float));
    if (!feat || !t1 || !t2) { free(feat); free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    const int fma_order = (order == ORC_ORDER_FMA);
    const unsigned char* sb = (const unsigned char*)src;
    const int es = elem_size(type);
    /* head: conv3x3 1 -> F, identity (Backend.hpp:135-156) */
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            int tp = i > 0 ? 1 : 0, bp = i < h - 1 ? 1 : 0, lp = j > 0 ? 1 : 0, rq = j < w - 1 ? 1 : 0;
            int ys[3] = { i - tp, i, i + bp }, xs[3] = { j - lp, j, j + rq };
            float r[9];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) r[a * 3 + c] = orc_to_float(sb + ys[a] * src_stride + (long)xs[c] * es, type);
            float* out = feat + ((size_t)i * w + j) * f;
            if (fma_order) cin1_fma_n(r, kernels, biases, out, f, 9); else cin1_generic_n(r, kernels, biases, out, f, 9);
        }
    /* body: `blocks` x conv3x3 F -> F + ReLU; then F -> F identity + feat (Backend.hpp:157-160) */
    const float* in = feat;
    float* out = t1;
    for (int l = 1; l <= blocks + 1; l++) {
        const float* k = kernels + f * 9 + (size_t)f * f * 9 * (l - 1);
        const float* b = biases + f * l;
        const int add = l == blocks + 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                ORC_TAPS(in, f)
                float* o = out + ((size_t)i * w + j) * f;
                if (fma_order) conv_fma_cin(rp9, k, b, o, f, f, 9); else conv_generic_cin(rp9, k, b, o, f, f, 9);
                const float* fe = feat + ((size_t)i * w + j) * f;
                for (int n = 0; n < f; n++) o[n] = add ? o[n] * 1.0f + fe[n] : (o[n] > 0.0f ? o[n] : 0.0f);
            }
        in = out;
        out = out == t1 ? t2 : t1;
    }
    /* tail: conv3x3 F -> 4, identity, pixel shuffle (Backend.hpp:161-178, Common.hpp:290-341) */
    {
        const float* k = kernels + f * 9 + (size_t)f * f * 9 * (blocks + 1);
        const float* b = biases + f * (blocks + 2);
        unsigned char* db = (unsigned char*)dst;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                ORC_TAPS(in, f)
                float s4[4];
                if (fma_order) conv_fma_cin(rp9, k, b, s4, f, 4, 9); else conv_generic_cin(rp9, k, b, s4, f, 4, 9);
                for (int n = 0; n < 4; n++)
                    orc_from_float(s4[n], db + (long)(2 * i + (n >> 1)) * dst_stride + (long)(2 * j + (n & 1)) * es, type);
            }
    }
    free(feat);
    free(t1);
    free(t2);
    return 0;
}

/* ---- FSRCNNX<F> (CPUProcessor.cpp:1629-1659 and the F = 16 twin) -------------------- */

int orc_fsrcnnx_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                     const float* kernels, const float* biases, const float* alphas, int f, int blocks, int order,
                     int threads)
{
    if (!src || !dst || w <= 0 || h <= 0 || !kernels || !biases || !alphas || (f != 8 && f != 16) || blocks < 2)
        return -22;
    const size_t plane = (size_t)w * h * f;
    float* feat = (float*)malloc(plane * sizeof(float));
    float* t1 = (float*)malloc(plane * sizeof(float));
    float* t2 = (float*)malloc(plane * sizeof(float));
    if (!feat || !t1 || !t2) { free(feat); free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    const int fma_order = (order == ORC_ORDER_FMA);
    const unsigned char* sb = (const unsigned char*)src;
    const int es = elem_size(type);
    /* head: conv5x5 1 -> F, identity, taps clamped to the image (Common.hpp:199-225) */
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            float r[25];
            for (int a = 0; a < 5; a++)
                for (int c = 0; c < 5; c++)
                    r[a * 5 + c] = orc_to_float(sb + clampi(i + a - 2, 0, h - 1) * src_stride + (long)clampi(j + c - 2, 0, w - 1) * es, type);
            float* out = feat + ((size_t)i * w + j) * f;
            if (fma_order) cin1_fma_n(r, kernels, biases, out, f, 25); else cin1_generic_n(r, kernels, biases, out, f, 25);
        }
    /* body: layers 1 .. blocks - 1 conv3x3 F -> F + PReLU; layer `blocks` conv3x3 + PReLU
     * fused with layer blocks + 1, conv1x1 F -> F, * 1 + feat, PReLU (Backend.hpp:243-254,
     * Common.hpp:227-288) */
    const float* in = feat;
    float* out = t1;
    for (int l = 1; l <= blocks; l++) {
        const float* k = kernels + f * 25 + (size_t)f * f * 9 * (l - 1);
        const float* b = biases + f * l;
        const float* al = alphas + f * (l - 1);
        const int fused = l == blocks;
        const float* k1 = kernels + f * 25 + (size_t)f * f * 9 * blocks;
        const float* b1 = biases + f * (blocks + 1);
        const float* a1 = alphas + f * blocks;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                ORC_TAPS(in, f)
                float buf[16];
                if (fma_order) conv_fma_cin(rp9, k, b, buf, f, f, 9); else conv_generic_cin(rp9, k, b, buf, f, f, 9);
                for (int n = 0; n < f; n++) buf[n] = prelu_(buf[n], al[n]);
                float* o = out + ((size_t)i * w + j) * f;
                if (!fused) {
                    for (int n = 0; n < f; n++) o[n] = buf[n];
                    continue;
                }
                const float* bp1[1] = { buf };
                float sum[16];
                if (fma_order) conv_fma_cin(bp1, k1, b1, sum, f, f, 1); else conv_generic_cin(bp1, k1, b1, sum, f, f, 1);
                const float* fe = feat + ((size_t)i * w + j) * f;
                for (int n = 0; n < f; n++) o[n] = prelu_(sum[n] * 1.0f + fe[n], a1[n]);
            }
        in = out;
        out = out == t1 ? t2 : t1;
    }
    /* tail: conv3x3 F -> 4, identity, pixel shuffle (Common.hpp:290-341) */
    {
        const float* k = kernels + f * 25 + (size_t)f * f * 9 * blocks + f * f;
        const float* b = biases + f * (blocks + 2);
        unsigned char* db = (unsigned char*)dst;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                ORC_TAPS(in, f)
                float s4[4];
                if (fma_order) conv_fma_cin(rp9, k, b, s4, f, 4, 9); else conv_generic_cin(rp9, k, b, s4, f, 4, 9);
                for (int n = 0; n < 4; n++)
                    orc_from_float(s4[n], db + (long)(2 * i + (n >> 1)) * dst_stride + (long)(2 * j + (n & 1)) * es, type);
            }
    }
    free(feat);
    free(t1);
    free(t2);
    return 0;
}

/* ---- FP64 restatement (error reference only) --------------------------------------- */

int orc_acnet_pass_f64(const float* src, int w, int h, double* dst,
                       const float* kernels, const float* biases, int threads)
{
    if (!src || !dst || w <= 0 || h <= 0) return -22;
    size_t plane = (size_t)w * h * CH;
    double* t1 = (double*)malloc(plane * sizeof(double));
    double* t2 = (double*)malloc(plane * sizeof(double));
    if (!t1 || !t2) { free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++)
            for (int n = 0; n < CH; n++) {
                double acc = biases[n];
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int y = clampi(i + dy, 0, h - 1), x = clampi(j + dx, 0, w - 1);
                        acc += (double)src[(size_t)y * w + x] * (double)kernels[n * 9 + (dy + 1) * 3 + (dx + 1)];
                    }
                t1[((size_t)i * w + j) * CH + n] = acc > 0.0 ? acc : 0.0;
            }
    double* tin = t1;
    double* tout = t2;
    for (int l = 1; l <= NB; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        const float* b = biases + 8 * l;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++)
                for (int n = 0; n < CH; n++) {
                    double acc = b[n];
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int y = clampi(i + dy, 0, h - 1), x = clampi(j + dx, 0, w - 1);
                            const double* r = tin + ((size_t)y * w + x) * CH;
                            const float* kk = k + n * 72 + ((dy + 1) * 3 + (dx + 1)) * 8;
                            for (int c = 0; c < CH; c++) acc += r[c] * (double)kk[c];
                        }
                    tout[((size_t)i * w + j) * CH + n] = acc > 0.0 ? acc : 0.0;
                }
        double* tmp = tin; tin = tout; tout = tmp;
    }
    const float* k2 = kernels + 72 + 576 * NB;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const double* s = tin + ((size_t)i * w + j) * CH;
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    double v = 0.0;
                    for (int c = 0; c < CH; c++) v += s[c] * (double)k2[8 * (dy * 2 + dx) + c];
                    v = v < 0.0 ? 0.0 : (v < 1.0 ? v : 1.0);
                    dst[(size_t)(2 * i + dy) * (2 * w) + (2 * j + dx)] = v;
                }
        }
    free(t1);
    free(t2);
    return 0;
}

/* ACNet<8> 2x pass in double precision (the same network as orc_acnet8_pass: replicate-padded
 * conv3x3 layers, PReLU, the final conv 8 -> 4, pixel shuffle, + the source pixel, saturated to
 * [0,1]); src is the toFloat'ed input.  Pinned against torch float64 conv2d / prelu /
 * pixel_shuffle (tests/test_oracle_acnet8.py). */
int orc_acnet8_pass_f64(const float* src, int w, int h, double* dst, const float* kernels, const float* biases,
                        const float* alphas, int blocks, int threads)
{
    if (!src || !dst || w <= 0 || h <= 0 || !kernels || !biases || !alphas || blocks < 1) return -22;
    size_t plane = (size_t)w * h * CH;
    double* t1 = (double*)malloc(plane * sizeof(double));
    double* t2 = (double*)malloc(plane * sizeof(double));
    if (!t1 || !t2) { free(t1); free(t2); return -12; }
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++)
            for (int n = 0; n < CH; n++) {
                double acc = biases[n];
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int y = clampi(i + dy, 0, h - 1), x = clampi(j + dx, 0, w - 1);
                        acc += (double)src[(size_t)y * w + x] * (double)kernels[n * 9 + (dy + 1) * 3 + (dx + 1)];
                    }
                t1[((size_t)i * w + j) * CH + n] = acc > 0.0 ? acc : (double)alphas[n] * acc;
            }
    double* tin = t1;
    double* tout = t2;
    for (int l = 1; l <= blocks + 1; l++) {
        const int cout = l <= blocks ? CH : 4;
        const float* k = kernels + 72 + 576 * (l - 1);
        const float* b = biases + 8 * l;
        const float* al = alphas + 8 * l;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nt)
#endif
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++)
                for (int n = 0; n < cout; n++) {
                    double acc = b[n];
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int y = clampi(i + dy, 0, h - 1), x = clampi(j + dx, 0, w - 1);
                            const double* r = tin + ((size_t)y * w + x) * CH;
                            const float* kk = k + n * 72 + ((dy + 1) * 3 + (dx + 1)) * 8;
                            for (int c = 0; c < CH; c++) acc += r[c] * (double)kk[c];
                        }
                    if (l <= blocks) {
                        tout[((size_t)i * w + j) * CH + n] = acc > 0.0 ? acc : (double)al[n] * acc;
                    } else {
                        /* sub-pixel n goes to (2j + n % 2, 2i + n / 2), plus the source pixel */
                        double v = acc + (double)src[(size_t)i * w + j];
                        v = v < 0.0 ? 0.0 : (v < 1.0 ? v : 1.0);
                        dst[(size_t)(2 * i + (n >> 1)) * (2 * w) + (2 * j + (n & 1))] = v;
                    }
                }
        double* tmp = tin; tin = tout; tout = tmp;
    }
    free(t1);
    free(t2);
    return 0;
}

/* ---- colour split / merge (ImageProcess.cpp:38-61, :114-139, :191-215, :276-308) ---- */

int orc_rgb2yuv(const void* rgb, long rgb_stride, int w, int h, int channels, int type,
                void* y, long y_stride, void* uv, long uv_stride)
{
    if (channels != 3 && channels != 4) return -22;
    const int es = elem_size(type);
    const int uvc = channels - 1;
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const unsigned char* in = (const unsigned char*)rgb + i * rgb_stride + (long)j * channels * es;
            unsigned char* yo = (unsigned char*)y + i * y_stride + (long)j * es;
            unsigned char* uo = (unsigned char*)uv + i * uv_stride + (long)j * uvc * es;
            float r = orc_to_float(in, type), g = orc_to_float(in + es, type), b = orc_to_float(in + 2 * es, type);
            float alpha = 1.0f;
            if (channels == 4) {   /* rgba2yuva premultiplies (ImageProcess.cpp:114-139) */
                alpha = orc_to_float(in + 3 * es, type);
                r = r * alpha;
                g = g * alpha;
                b = b * alpha;
            }
            float a0 = 0.299f * r, a1 = 0.587f * g, a2 = 0.114f * b;
            float yy = (a0 + a1) + a2;
            float u = 0.564f * (b - yy) + 0.5f;
            float v = 0.713f * (r - yy) + 0.5f;
            orc_from_float(yy, yo, type);
            orc_from_float(u, uo, type);
            orc_from_float(v, uo + es, type);
            if (channels == 4) orc_from_float(alpha, uo + 2 * es, type);
        }
    return 0;
}

int orc_yuv2rgb(const void* y, long y_stride, const void* uv, long uv_stride,
                int w, int h, int channels, int type, void* rgb, long rgb_stride)
{
    if (channels != 3 && channels != 4) return -22;
    const int es = elem_size(type);
    const int uvc = channels - 1;
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const unsigned char* yi = (const unsigned char*)y + i * y_stride + (long)j * es;
            const unsigned char* ui = (const unsigned char*)uv + i * uv_stride + (long)j * uvc * es;
            unsigned char* out = (unsigned char*)rgb + i * rgb_stride + (long)j * channels * es;
            float yy = orc_to_float(yi, type);
            float u = orc_to_float(ui, type) - 0.5f;
            float v = orc_to_float(ui + es, type) - 0.5f;
            float r = yy + 1.403f * v;
            float g = (yy - 0.344f * u) - 0.714f * v;
            float b = yy + 1.773f * u;
            if (channels == 4) {   /* yuva2rgba un-premultiplies (ImageProcess.cpp:276-308) */
                float alpha = orc_to_float(ui + 2 * es, type);
                if (alpha > 1e-6f) { r = r / alpha; g = g / alpha; b = b / alpha; }
                else r = g = b = 0.0f;
                orc_from_float(alpha, out + 3 * es, type);
            }
            orc_from_float(r, out, type);
            orc_from_float(g, out + es, type);
            orc_from_float(b, out + 2 * es, type);
        }
    return 0;
}

/* ---- resize (parity with stb_image_resize2 unpinned) -------------------------------- */

/* Filter kernels of ImageResize.cpp:53-134 and the mode switch of :193-263 (mode numbers are
 * the ResizeModes of core/include/AC/Core/Image.hpp:272-291): the (b, c) cubic family of
 * Mitchell and Netravali as Horner polynomials (:62-84), Lanczos (:85-91), the spline16/36/64
 * piecewise cubics (:92-125) and the tent (:126-134).  Float throughout. */
static float poly3(float x, float c0, float c1, float c2, float c3)
{
    return c0 + x * (c1 + x * (c2 + x * c3));
}
static float cubic_bc(float v, float b, float c)
{
    const float p0 = (6.0f - 2.0f * b) / 6.0f;
    const float p1 = 0.0f;
    const float p2 = (-18.0f + 12.0f * b + 6.0f * c) / 6.0f;
    const float p3 = (12.0f - 9.0f * b - 6.0f * c) / 6.0f;
    const float q0 = (8.0f * b + 24.0f * c) / 6.0f;
    const float q1 = (-12.0f * b - 48.0f * c) / 6.0f;
    const float q2 = (6.0f * b + 30.0f * c) / 6.0f;
    const float q3 = (-b - 6.0f * c) / 6.0f;
    const float x = fabsf(v);
    if (x < 1.0f) return poly3(x, p0, p1, p2, p3);
    if (x < 2.0f) return poly3(x, q0, q1, q2, q3);
    return 0.0f;
}
static float sincf_(float x)
{
    const float pi = 3.14159265358979f;
    return x == 0.0f ? 1.0f : sinf(x * pi) / (x * pi);
}
static float lanczos_(float v, float taps)
{
    const float x = fabsf(v);
    return x < taps ? sincf_(x) * sincf_(x / taps) : 0.0f;
}
static float spline16_(float v)
{
    const float x = fabsf(v);
    if (x < 1.0f) return poly3(x, 1.0f, -1.0f / 5.0f, -9.0f / 5.0f, 1.0f);
    if (x < 2.0f) return poly3(x - 1.0f, 0.0f, -7.0f / 15.0f, 4.0f / 5.0f, -1.0f / 3.0f);
    return 0.0f;
}
static float spline36_(float v)
{
    const float x = fabsf(v);
    if (x < 1.0f) return poly3(x, 1.0f, -3.0f / 209.0f, -453.0f / 209.0f, 13.0f / 11.0f);
    if (x < 2.0f) return poly3(x - 1.0f, 0.0f, -156.0f / 209.0f, 270.0f / 209.0f, -6.0f / 11.0f);
    if (x < 3.0f) return poly3(x - 2.0f, 0.0f, 26.0f / 209.0f, -45.0f / 209.0f, 1.0f / 11.0f);
    return 0.0f;
}
static float spline64_(float v)
{
    const float x = fabsf(v);
    if (x < 1.0f) return poly3(x, 1.0f, -3.0f / 2911.0f, -6387.0f / 2911.0f, 49.0f / 41.0f);
    if (x < 2.0f) return poly3(x - 1.0f, 0.0f, -2328.0f / 2911.0f, 4032.0f / 2911.0f, -24.0f / 41.0f);
    if (x < 3.0f) return poly3(x - 2.0f, 0.0f, 582.0f / 2911.0f, -1008.0f / 2911.0f, 6.0f / 41.0f);
    if (x < 4.0f) return poly3(x - 3.0f, 0.0f, -97.0f / 2911.0f, 168.0f / 2911.0f, -1.0f / 41.0f);
    return 0.0f;
}

float orc_filter_weight(int mode, float x)
{
    switch (mode) {
    case 0: return 1.0f;                                   /* POINT */
    case 1: return cubic_bc(x, 0.0f, 1.0f / 2.0f);         /* CATMULL_ROM */
    case 2: return cubic_bc(x, 1.0f / 3.0f, 1.0f / 3.0f);  /* MITCHELL_NETRAVALI */
    case 3: return cubic_bc(x, 0.0f, 3.0f / 5.0f);         /* BICUBIC_0_60 */
    case 4: return cubic_bc(x, 0.0f, 3.0f / 4.0f);         /* BICUBIC_0_75 */
    case 5: return cubic_bc(x, 0.0f, 1.0f);                /* BICUBIC_0_100 */
    case 6: return cubic_bc(x, 1.0f / 5.0f, 1.0f / 2.0f);  /* BICUBIC_20_50 */
    case 7: return cubic_bc(x, 1.0f / 2.0f, 1.0f / 2.0f);  /* SOFTCUBIC50 */
    case 8: return cubic_bc(x, 3.0f / 4.0f, 1.0f / 4.0f);  /* SOFTCUBIC75 */
    case 9: return cubic_bc(x, 1.0f, 0.0f);                /* SOFTCUBIC100 */
    case 10: return lanczos_(x, 2.0f);
    case 11: return lanczos_(x, 3.0f);
    case 12: return lanczos_(x, 4.0f);
    case 13: return spline16_(x);
    case 14: return spline36_(x);
    case 15: return spline64_(x);
    default: { const float a = fabsf(x); return a < 1.0f ? 1.0f - a : 0.0f; }   /* BILINEAR */
    }
}

float orc_filter_support(int mode)
{
    switch (mode) {
    case 0: return 0.5f;
    case 11: case 14: return 3.0f;
    case 12: case 15: return 4.0f;
    case 16: return 1.0f;
    default: return 2.0f;
    }
}

/* Taps of output coordinate o: the source pixels whose centres lie in
 * (center - support, center + support] (stb's pixel range), pixel-centre mapping, the
 * filter stretched by 1/scale when downscaling, weights normalised to sum 1. */
#define ORC_MAX_TAPS 256
static int resize_taps(int mode, int o, int n_in, int n_out, float* w, int* first)
{
    const float scale = (float)n_out / (float)n_in;
    const float center = ((float)o + 0.5f) / scale - 0.5f;
    const float support = orc_filter_support(mode) * (scale < 1.0f ? 1.0f / scale : 1.0f);
    const float stretch = scale < 1.0f ? scale : 1.0f;
    const int lo = (int)floorf(center - support) + 1;
    const int hi = (int)floorf(center + support);
    int n = hi - lo + 1;
    if (n < 0) n = 0;
    if (n > ORC_MAX_TAPS) n = ORC_MAX_TAPS;
    float sum = 0.0f;
    for (int i = 0; i < n; i++) {
        w[i] = orc_filter_weight(mode, ((float)(lo + i) - center) * stretch);
        sum += w[i];
    }
    if (sum != 0.0f)
        for (int i = 0; i < n; i++) w[i] /= sum;
    *first = lo;
    return n;
}

int orc_resize(const void* src, long src_stride, int sw, int sh, int channels, int type,
               void* dst, long dst_stride, int dw, int dh, int mode)
{
    if (sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0 || channels <= 0 || channels > 4) return -22;
    if (mode < 0 || mode > 16) mode = 16;
    const int es = elem_size(type);
    for (int y = 0; y < dh; y++) {
        float wy[ORC_MAX_TAPS], wx[ORC_MAX_TAPS];
        int fy, fx;
        const int ny = resize_taps(mode, y, sh, dh, wy, &fy);
        for (int x = 0; x < dw; x++) {
            const int nx = resize_taps(mode, x, sw, dw, wx, &fx);
            float acc[4] = { 0, 0, 0, 0 };
            for (int j = 0; j < ny; j++) {
                const unsigned char* row = (const unsigned char*)src + (long)clampi(fy + j, 0, sh - 1) * src_stride;
                float h[4] = { 0, 0, 0, 0 };
                for (int i = 0; i < nx; i++) {
                    const unsigned char* p = row + (long)clampi(fx + i, 0, sw - 1) * channels * es;
                    for (int c = 0; c < channels; c++) {
                        float m = wx[i] * orc_to_float(p + c * es, type);
                        h[c] = h[c] + m;
                    }
                }
                for (int c = 0; c < channels; c++) {
                    float m = wy[j] * h[c];
                    acc[c] = acc[c] + m;
                }
            }
            for (int c = 0; c < channels; c++)
                orc_from_float(acc[c], (unsigned char*)dst + (long)y * dst_stride + ((long)x * channels + c) * es, type);
        }
    }
    return 0;
}

/* 2x closed form kept for the tests' tap check: even outputs sample i - 0.25, odd i + 0.25 */
int orc_resize2x(const void* src, long src_stride, int w, int h, int channels, int type,
                 void* dst, long dst_stride, int mode)
{
    return orc_resize(src, src_stride, w, h, channels, type, dst, dst_stride, 2 * w, 2 * h, mode);
}
