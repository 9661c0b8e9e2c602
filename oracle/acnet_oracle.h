/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C CPU restatement of the reference's ACNetLegacy 2x pass and of the
 * colour/resize helpers around it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline; the product library never links or calls it.
 *
 * Parity pins (each function below names its own):
 *   - orc_acnet_pass (FMA and Generic orders): bit-exact against oracle/_ref,
 *     which compiles the reference's own per-ISA dot/conv kernels
 *     (core/internal/AC/Core/Internal/Processor/CPU/{Generic.hpp,X86/*.hpp});
 *     against torch conv2d/conv_transpose2d (library routine); the constant
 *     image closed form; brute force on tiny images; translation equivariance.
 *   - orc_acnet_pass_f64: against torch float64 conv2d (library routine).
 *   - orc_acnet8_pass_f64: against torch float64 conv2d / prelu / pixel_shuffle.
 *   - orc_from_float / orc_to_float: exhaustive closed-form checks.
 *   - orc_rgb2yuv / orc_yuv2rgb: closed-form per-pixel formulas and the
 *     reference's own round-trip bound (tests/core/src/ImageProcessTest.cpp:52-79).
 *   - orc_resize2x_*: parity unpinned (stb_image_resize2 is absent); pinned
 *     only against closed-form 2x taps and invariants.
 */
#ifndef ACNET_ORACLE_H
#define ACNET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types, encoded like ac::core::Image (core/include/AC/Core/Image.hpp:367-371) */
enum { ORC_U8 = 0x001, ORC_U16 = 0x002, ORC_F16 = 0x202, ORC_F32 = 0x204 };

/* summation orders of the reference CPU backends */
enum {
    ORC_ORDER_FMA = 0,      /* X86/AVX.hpp OpImplFMA, also what AVX512 runs for 8 channels */
    ORC_ORDER_GENERIC = 1   /* Generic.hpp OpImplGeneric */
};

float orc_to_float(const void* p, int type);
void orc_from_float(float v, void* p, int type);
float orc_half_to_float(uint16_t h);
uint16_t orc_float_to_half(float f);

/* One 2x pass: src w x h (1 channel, byte stride) -> dst 2w x 2h.
 * kernels: 4712 floats, biases: 72 floats (ACNet.hpp:34-60 layout).
 * threads <= 0 uses all cores.  Returns 0 on success. */
int orc_acnet_pass(const void* src, long src_stride, int w, int h, int src_type,
                   void* dst, long dst_stride, int dst_type,
                   const float* kernels, const float* biases, int order, int threads);

/* ACNet<8> 2x pass (core/src/processor/cpu/CPUProcessor.cpp:1417-1436): conv3x3 1->8 +
 * PReLU, `blocks` x conv3x3 8->8 + PReLU, conv3x3 8->4 + the source pixel, pixel shuffle;
 * src and dst of one element type; weights in the reference layout
 * (core/include/AC/Core/Model/ACNet.hpp:77-117).  order as orc_acnet_pass. */
int orc_acnet8_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                    const float* kernels, const float* biases, const float* alphas, int blocks, int order, int threads);

/* ArtCNN<F> 2x pass, F = 16 or 32 (core/src/processor/cpu/CPUProcessor.cpp:1519-1602): head
 * conv3x3 1->F, `blocks` x conv3x3 F->F + ReLU, conv3x3 F->F + head features, conv3x3 F->4 +
 * pixel shuffle; src and dst of one element type.  order as orc_acnet_pass (FMA = the AVX/FMA
 * kernels' order; AVX512 sums 16 lanes and may differ in the last bit). */
int orc_artcnn_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                    const float* kernels, const float* biases, int f, int blocks, int order, int threads);

/* FSRCNNX<F> 2x pass, F = 8 or 16 (core/src/processor/cpu/CPUProcessor.cpp:1629-1659):
 * conv5x5 1->F, `blocks` - 1 x conv3x3 F->F + PReLU, conv3x3 F->F + PReLU then conv1x1 F->F
 * + head features + PReLU, conv3x3 F->4 + pixel shuffle; src and dst of one element type. */
int orc_fsrcnnx_pass(const void* src, long src_stride, int w, int h, int type, void* dst, long dst_stride,
                     const float* kernels, const float* biases, const float* alphas, int f, int blocks, int order,
                     int threads);

/* The same network in double precision.  src is the toFloat'ed input
 * (w*h floats, dense); dst receives the saturated 2w x 2h output in double. */
/* ACNet<8> in double precision, for the error of the 16-bit forms: src toFloat'ed, dst 2h x 2w
 * saturated to [0,1]. */
int orc_acnet8_pass_f64(const float* src, int w, int h, double* dst, const float* kernels, const float* biases,
                        const float* alphas, int blocks, int threads);
int orc_acnet_pass_f64(const float* src, int w, int h, double* dst,
                       const float* kernels, const float* biases, int threads);

/* Colour split/merge, core/src/ImageProcess.cpp:15-61 and :168-215.
 * rgb: interleaved 3 (or 4 with alpha) channels; y: 1 channel; uv: 2 (or 3 with
 * alpha) interleaved channels.  All of the same element type. */
int orc_rgb2yuv(const void* rgb, long rgb_stride, int w, int h, int channels, int type,
                void* y, long y_stride, void* uv, long uv_stride);
int orc_yuv2rgb(const void* y, long y_stride, const void* uv, long uv_stride,
                int w, int h, int channels, int type, void* rgb, long rgb_stride);

/* Separable resize with edge clamp and pixel-centre alignment, per channel, to dw x dh,
 * with the filter of ResizeModes `mode` (core/include/AC/Core/Image.hpp:272-291; out of range
 * means bilinear, as ImageResize.cpp:260 does); the filter is stretched by 1/scale when
 * downscaling.  f32 math, horizontal then vertical, output through fromFloat.
 * Parity with stb_image_resize2 is unpinned; pinned against closed-form 2x taps, the
 * filters' published forms and their interpolation / partition-of-unity properties. */
int orc_resize(const void* src, long src_stride, int sw, int sh, int channels, int type,
               void* dst, long dst_stride, int dw, int dh, int mode);
/* the filter kernel of `mode` at distance x, and its support radius */
float orc_filter_weight(int mode, float x);
float orc_filter_support(int mode);
int orc_resize2x(const void* src, long src_stride, int w, int h, int channels, int type,
                 void* dst, long dst_stride, int mode);

#ifdef __cplusplus
}
#endif

#endif
