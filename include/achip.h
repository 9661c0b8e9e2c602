/*
 * achip.h -- the thin device-level C ABI under the HIP processor.
 *
 * Plain pointers and sizes only.  Images are device pointers with byte strides;
 * element types use the ac::core::Image encoding (0x001 u8, 0x002 u16, 0x202 f16,
 * 0x204 f32).  Every call validates its arguments on the host (sizes, strides,
 * alignment, and that each buffer is device memory of the net's device large enough
 * for the image) before anything is launched; an invalid call returns an error code
 * and launches nothing.  Error codes are negative errno values:
 * -22 invalid argument, -19 no device, -12 out of memory, -5 HIP runtime error.
 * achip_last_error() returns the calling thread's last message.
 *
 * It replaces, for the ACNetLegacy path, the reference's per-backend 2x pass:
 *   CPUProcessor<ACNetLegacy>::process   core/src/processor/cpu/CPUProcessor.cpp:1371-1390
 *   CUDAProcessor<ACNetLegacy>::process  core/src/processor/cuda/CUDAProcessor.cpp:383-422
 * behind the private virtual Processor::process(src, dst) (core/include/AC/Core/Processor.hpp:37).
 */
#ifndef ACHIP_H
#define ACHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define ACHIP_API __attribute__((visibility("default")))
#else
#define ACHIP_API
#endif

/* arithmetic of the 8->8 layers */
#define ACHIP_FP32 0   /* fused FP32 VALU kernel */
#define ACHIP_FP16 1   /* FP16 MFMA implicit GEMM, FP32 accumulate */
#define ACHIP_BF16 2   /* BF16 MFMA implicit GEMM, FP32 accumulate */

#define ACHIP_OK 0
#define ACHIP_EIO (-5)
#define ACHIP_ENOMEM (-12)
#define ACHIP_ENODEV (-19)
#define ACHIP_EINVAL (-22)

typedef struct ACHipNet ACHipNet;

/* Number of HIP devices; 0 when none (never an error). */
ACHIP_API int achip_device_count(void);

/* Upload one ACNetLegacy weight set (kernels: 4712 floats, biases: 72 floats, the
 * reference layout of core/include/AC/Core/Model/ACNet.hpp:34-60) to a device.
 * Replaces the weight upload in the reference's GPU processor constructor
 * (core/src/processor/cuda/CUDAProcessor.cpp:291-326).  device < 0 or out of range
 * picks the device with the most compute units.  Returns 0 and *out on success. */
ACHIP_API int achip_net_create(int device, const float* kernels, const float* biases, ACHipNet** out);
ACHIP_API void achip_net_destroy(ACHipNet* net);
ACHIP_API int achip_net_device(const ACHipNet* net);

/* One 2x pass of the whole network on device memory: src w x h (1 channel) ->
 * dst 2w x 2h (1 channel), on `stream` (a hipStream_t; NULL = the null stream).
 * Asynchronous: returns after the launch.  Output element type follows dst_type
 * (Common.hpp:344-393 writes fromFloat<OUT>). */
ACHIP_API int achip_pass(ACHipNet* net, int precision,
                         const void* src, long src_stride, int src_type, int w, int h,
                         void* dst, long dst_stride, int dst_type, void* stream);

/* The same with an explicit kernel configuration (autotuning and tests).  config
 * ACHIP_AUTO uses the autotuner's choice for the image's size class, or the default
 * configuration when that class was not tuned (with ACHIP_AUTOTUNE=1 in the
 * environment an untuned class is tuned on first use). */
#define ACHIP_AUTO (-1)
ACHIP_API int achip_pass_config(ACHipNet* net, int precision, int config,
                                const void* src, long src_stride, int src_type, int w, int h,
                                void* dst, long dst_stride, int dst_type, void* stream);

/* One image of a batched pass (device pointers; 1 channel; dst is 2w x 2h). */
typedef struct ACHipImage {
    const void* src;
    long src_stride;
    void* dst;
    long dst_stride;
    int w;
    int h;
} ACHipImage;

/* Device workspace a batch of n images needs (the tile-queue counter and the image
 * descriptors). */
ACHIP_API size_t achip_batch_workspace_bytes(int n);

/* 2x passes of n images of uneven sizes (all of src_type -> dst_type) in ONE persistent
 * launch: resident workgroups pull tiles from a queue until the batch is done.
 * `workspace` is device memory of at least achip_batch_workspace_bytes(n) bytes, owned by
 * the caller and not touched by any other call until this one has completed on `stream`.
 * No counterpart in the reference, whose callers loop over images one pass at a time
 * (tools/benchmark/src/Benchmark.cpp:58-63, cli/src/Main.cpp:112-116). */
ACHIP_API int achip_pass_batch(ACHipNet* net, int precision, int config, int src_type, int dst_type,
                               const ACHipImage* images, int n, void* workspace, size_t workspace_bytes,
                               void* stream);

/* Kernel configurations of a precision: tile width/height (input pixels), threads per
 * workgroup, pixels per thread run, LDS bytes.  Returns the count, or fills config i. */
ACHIP_API int achip_config_count(int precision);
ACHIP_API int achip_config_info(int precision, int i, int* tw, int* th, int* nt, int* p, int* lds_bytes);

/* Autotuner (project-written; no counterpart in the reference): time every configuration
 * of `precision` on w x h images (src_type -> dst_type) on the net's device, on temporary
 * buffers, a few launches each, and remember the fastest for that size class (log2 of
 * width and height).  Fills *best and, if ms is not NULL, the mean launch time of each
 * configuration (achip_config_count(precision) floats). */
ACHIP_API int achip_autotune(ACHipNet* net, int precision, int src_type, int dst_type, int w, int h,
                             int* best, float* ms);

/* Calling thread's last error message ("" if none). */
ACHIP_API const char* achip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
