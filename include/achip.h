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
#define ACHIP_FP32 0   /* fused FP32 kernel: every conv on FP32 MFMA (default) or on FP32 VALU */
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

/* ACNetLegacy's FP32 2x pass layer by layer instead of fused: layers 0-2 in one launch, then
 * layers 3..8 two per launch through 8-channel f32 maps in `workspace`
 * (achip_pass8_workspace_bytes(w, h) bytes, 256-byte aligned, device memory of the net's
 * device, not overlapping src or dst): four launches, each map through HBM once, a one-pixel
 * recompute ring per layer instead of the fused kernel's eight-pixel halo.  Every layer runs
 * the fused FP32 kernel's instructions on the same operands, so the results are its bits.
 * The reference's CPU pass runs the layers one by one (CPUProcessor.cpp:1371-1390). */
ACHIP_API int achip_pass_layers(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h, void* dst,
                                long dst_stride, int dst_type, void* workspace, size_t workspace_bytes, void* stream);

/* The layer-split form over a batch of n images of uneven sizes (all src_type -> dst_type),
 * in four launches that each take the tiles of every image (no counterpart in the
 * reference, whose callers run one pass per image: tools/benchmark/src/Benchmark.cpp:58-63).
 * `workspace`: device memory of achip_pass_batch_layers_bytes(images, n) bytes, 256-byte
 * aligned, the descriptors and every image's two 8-channel f32 maps, not overlapping any
 * image; no image's dst may overlap another image's src or dst.  Outputs are the fused
 * kernel's bits, as achip_pass_layers' are. */
ACHIP_API size_t achip_pass_batch_layers_bytes(const ACHipImage* images, int n);
ACHIP_API int achip_pass_batch_layers(ACHipNet* net, int src_type, int dst_type, const ACHipImage* images, int n,
                                      void* workspace, size_t workspace_bytes, void* stream);

/* The 2x pass in the faster of its forms for the image's size, given a workspace.  Replaces
 * achip_pass where the caller can lend device memory (the processor's per-thread buffers do;
 * the reference's processors run one form, CUDAProcessor.cpp:383-422).  For FP32 ACNetLegacy
 * passes the forms are the fused kernel and the layer-split form of achip_pass_layers, which
 * give the same bits; the choice is the autotuner's for the size class (achip_autotune times
 * both), else a rule measured on MI355X: the fused kernel where its 64 x 32 tiles fit in one
 * round of the device's compute units on an image of more than 256 x 256 pixels, the
 * layer-split form elsewhere (1080p u8: 299 against 328 us).  achip_pass_ws_bytes gives the
 * workspace bytes the chosen form needs, 0 for the fused kernel and every other net or
 * precision; with a smaller workspace (or none) the pass runs the fused kernel, as achip_pass
 * does.  Workspace rules as for achip_pass_layers. */
ACHIP_API size_t achip_pass_ws_bytes(ACHipNet* net, int precision, int src_type, int dst_type, int w, int h);
ACHIP_API int achip_pass_ws(ACHipNet* net, int precision, const void* src, long src_stride, int src_type, int w, int h,
                            void* dst, long dst_stride, int dst_type, void* workspace, size_t workspace_bytes, void* stream);

/* achip_pass_batch without a workspace argument and without any allocation, host or device.
 * No counterpart in the reference: it is for callers that cannot allocate (code under HIP
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The tile-queue
 * counter and the descriptors live in a fixed device array of the library
 * (achip_noalloc_scratch_bytes() bytes on each device, held until the process ends), so a
 * call takes at most ACHIP_NOALLOC_MAX_IMAGES images (more: ACHIP_EINVAL).  Calls on one
 * device share that array: calls that overlap in time on one device race, so serialise them
 * (one stream, or an event between them).  Same images, results and errors as
 * achip_pass_batch otherwise. */
#define ACHIP_NOALLOC_MAX_IMAGES 256
ACHIP_API size_t achip_noalloc_scratch_bytes(void);
ACHIP_API int achip_pass_batch_noalloc(ACHipNet* net, int precision, int config, int src_type, int dst_type,
                                       const ACHipImage* images, int n, void* stream);

/* Kernel configurations of a precision: tile width/height (input pixels), threads per
 * workgroup, pixels per thread run of the FP32 VALU kernel (0 for the MFMA kernels), LDS
 * bytes for f32 input.  FP32: the FP32-MFMA kernel first (configuration 0 is the default),
 * then the FP32 VALU kernel.  FP16/BF16: the 64x32 tile at 512 threads (configuration 0, the
 * default for u8 input) and the 64x64 tile at 1024 threads (configuration 1, the default for
 * wider inputs).  Returns the count, or fills config i. */
ACHIP_API int achip_config_count(int precision);
ACHIP_API int achip_config_info(int precision, int i, int* tw, int* th, int* nt, int* p, int* lds_bytes);

/* Autotuner (project-written; no counterpart in the reference): time every configuration
 * of `precision` on w x h images (src_type -> dst_type) on the net's device, on temporary
 * buffers, a few launches each, and remember the fastest for that size class (log2 of
 * width and height).  Fills *best and, if ms is not NULL, the mean launch time of each
 * configuration (achip_config_count(precision) floats).  For an FP32 ACNetLegacy net it also
 * times the layer-split form and remembers for achip_pass_ws whether it beat the fastest
 * configuration. */
ACHIP_API int achip_autotune(ACHipNet* net, int precision, int src_type, int dst_type, int w, int h,
                             int* best, float* ms);

/* Calling thread's last error message ("" if none). */
ACHIP_API const char* achip_last_error(void);

/* ---- ACNet<8> ----
 * The reference's default model family (core/include/AC/Core/Model/ACNet.hpp:63-118; its CPU
 * pass core/src/processor/cpu/CPUProcessor.cpp:1417-1436): conv3x3 1 -> 8 + PReLU, `blocks`
 * x conv3x3 8 -> 8 + PReLU, conv3x3 8 -> 4, pixel shuffle to 2x, plus the source pixel.
 * kernels F*9 + F*F*9*blocks + F*4*9, biases F + F*blocks + 4, alphas F*(blocks+1) floats
 * (F = 8), in the reference's layout.  FP32 throughout.  Such a net runs through
 * achip_pass8 only (achip_pass, achip_pass_batch and the server take ACNetLegacy nets). */
ACHIP_API int achip_net8_create(int device, int blocks, const float* kernels, const float* biases,
                                const float* alphas, ACHipNet** out);
/* block count of an ACNet<8> net, 0 for an ACNetLegacy net */
ACHIP_API int achip_net_blocks(const ACHipNet* net);
/* device workspace (two 8-channel f32 feature maps) one w x h pass needs */
ACHIP_API size_t achip_pass8_workspace_bytes(int w, int h);
/* one ACNet<8> 2x pass, device pointers as achip_pass; workspace: 256-byte aligned device
 * memory of at least achip_pass8_workspace_bytes(w, h), not overlapping src or dst.  The
 * residual reads the source in its own element type (the reference reads it as the output
 * type, which only agrees when both types are the same). */
ACHIP_API int achip_pass8(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h,
                          void* dst, long dst_stride, int dst_type, void* workspace, size_t workspace_bytes,
                          void* stream);
/* achip_pass8 in a given arithmetic: ACHIP_FP32 (the same as achip_pass8), or ACHIP_FP16 /
 * ACHIP_BF16 -- the conv3x3 layers on 16-bit MFMA with FP32 accumulation, bias, PReLU and the
 * final shuffle + residual in FP32, the feature maps between launches rounded to 16 bits; the
 * tests bound its error against a double-precision restatement of the network.  Same
 * workspace as achip_pass8.  The reference's ACNet<8> backends compute in FP32 on the CPU
 * (CPUProcessor.cpp:1417-1436) and keep FP16 maps on CUDA (CUDAProcessor.cpp:463-464). */
ACHIP_API int achip_pass8_precision(ACHipNet* net, int precision, const void* src, long src_stride, int src_type, int w,
                                    int h, void* dst, long dst_stride, int dst_type, void* workspace,
                                    size_t workspace_bytes, void* stream);

/* ---- ArtCNN<F> ----
 * core/include/AC/Core/Model/ArtCNN.hpp (F = 16 or 32; CPU pass CPUProcessor.cpp:1519-1602):
 * conv3x3 1 -> F (identity) = feat, `blocks` x conv3x3 F -> F + ReLU, conv3x3 F -> F + feat,
 * conv3x3 F -> 4 and pixel shuffle to 2x.  kernels F*9 + F*F*9*(blocks+1) + F*4*9, biases
 * F + F*(blocks+1) + 4 floats in the reference's layout.  FP32; runs through achip_pass_net. */
ACHIP_API int achip_net_artcnn_create(int device, int channels, int blocks, const float* kernels,
                                      const float* biases, ACHipNet** out);

/* ---- FSRCNNX<F> ----
 * core/include/AC/Core/Model/FSRCNNX.hpp (F = 8 or 16; CPU pass CPUProcessor.cpp:1629-1659):
 * conv5x5 1 -> F (identity) = feat, blocks - 1 x conv3x3 F -> F + PReLU, conv3x3 F -> F +
 * PReLU then conv1x1 F -> F + feat + PReLU, conv3x3 F -> 4 and pixel shuffle to 2x.  kernels
 * F*25 + F*F*9*blocks + F*F + F*4*9, biases F + F*(blocks+1) + 4, alphas F*(blocks+1) floats
 * in the reference's layout.  FP32; runs through achip_pass_net. */
ACHIP_API int achip_net_fsrcnnx_create(int device, int channels, int blocks, const float* kernels,
                                       const float* biases, const float* alphas, ACHipNet** out);

/* ---- any net ----
 * Device workspace one w x h pass of `net` needs (0 for ACNetLegacy), and that pass: an
 * ACNetLegacy net runs achip_pass in FP32, the others their layer-by-layer forms. */
ACHIP_API size_t achip_net_workspace_bytes(const ACHipNet* net, int w, int h);
ACHIP_API int achip_pass_net(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h,
                             void* dst, long dst_stride, int dst_type, void* workspace, size_t workspace_bytes,
                             void* stream);

/* ---- Device-memory colour and resize stages ----
 * The steps of Processor::process around the 2x passes (core/src/processor/Processor.cpp:
 * 207-213, 242-274) on device pointers of the current HIP device: the colour conversions of
 * core/src/ImageProcess.cpp:15-308 and the resize of core/src/ImageResize.cpp:136-273.  The
 * reference only has host-image forms of these (ac::core::rgb2yuv, resize; declared in
 * include/AC/Core/Image.hpp here); these take HBM-resident planes, for callers whose frames
 * stay on the device.  All planes of one element type (0x001 u8, 0x002 u16, 0x202 f16,
 * 0x204 f32), byte strides, asynchronous on `stream`.  channels: 3 (RGB) or 4 (RGBA,
 * premultiplied as the reference does); the UV plane interleaves channels - 1 components. */
ACHIP_API int achip_rgb2yuv(int type, int channels, const void* rgb, long rgb_stride, int w, int h,
                            void* y, long y_stride, void* uv, long uv_stride, void* stream);
ACHIP_API int achip_yuv2rgb(int type, int channels, const void* y, long y_stride, const void* uv, long uv_stride,
                            int w, int h, void* rgb, long rgb_stride, void* stream);
/* resize of a 1-4 channel plane (mode: ResizeModes of core/include/AC/Core/Image.hpp:272-291) */
ACHIP_API int achip_resize(int type, int channels, int mode, const void* src, long src_stride, int sw, int sh,
                           void* dst, long dst_stride, int dw, int dh, void* stream);
/* UV(A) (uv_w x uv_h) resized to w x h and merged with Y (w x h) into RGB(A) in one pass:
 * achip_resize then achip_yuv2rgb, with the resized chroma rounded to the element type */
ACHIP_API int achip_resize_yuv2rgb(int type, int channels, int mode, const void* y, long y_stride, const void* uv,
                                   long uv_stride, int uv_w, int uv_h, int w, int h, void* rgb, long rgb_stride,
                                   void* stream);

/* ---- Serving loop -----------------------------------------------------------------------
 * No single counterpart in the reference: it replaces the frame/image dispatch that the
 * reference's callers do one processor call at a time per thread (video
 * video/src/Filter.cpp:33-121 filterParallel, the CLI's per-image loop cli/src/Main.cpp:
 * 91-199, both through Processor::process, core/src/processor/Processor.cpp:199-276).
 *
 * A server owns one scheduler thread.  Requests of mixed kind, size and factor are admitted
 * in arrival order as long as their device buffers fit an HBM budget; admitted one-off
 * requests have their 2x passes batched, step by step, into persistent launches with a
 * tile work-queue (a big pass is spread over several steps by tile rows, so that newly
 * admitted work joins within one step); video-stream frames (fixed geometry per stream)
 * replay a captured hipGraph on the stream's own HIP stream and overlap with the batched
 * work on the device.  Outputs are bit-identical to processing the request alone (same
 * kernels, same tiles). */
typedef struct ACHipServer ACHipServer;

enum {
    ACHIP_REQ_LUMA = 0,     /* 1 channel u8 (ACHIP_REQ_YUV with the format { 1, 0, 0, 0x001, 0, 0 }) */
    ACHIP_REQ_YUV420 = 1,   /* u8 planes Y (W x H), U and V (W/2 x H/2), W and H even; chroma bilinear
                               (ACHIP_REQ_YUV with { 3, 1, 1, 0x001, 0, ACHIP_CHROMA_BILINEAR }) */
    ACHIP_REQ_RGB = 2,      /* 3 interleaved u8 channels */
    ACHIP_REQ_YUV = 3       /* a video frame laid out as `format` says */
};

enum { ACHIP_CHROMA_BILINEAR = 0, ACHIP_CHROMA_CATMULL_ROM = 1 };

/* Plane layout of an ACHIP_REQ_YUV frame: the formats of the reference's video module
 * (video/src/Pipeline.cpp:495-566) and the handling of its frame callback (cli/src/Main.cpp:
 * 183-205: luma shifted up by `shift` bits, through the passes, shifted back; chroma planes
 * resized to the output size).  Examples:
 *   gray8 { 1, 0, 0, 0x001, 0, - }      gray10le { 1, 0, 0, 0x002, 6, - }
 *   yuv420p { 3, 1, 1, 0x001, 0, f }    yuv420p10le { 3, 1, 1, 0x002, 6, f }   yuv444p16 { 3, 0, 0, 0x002, 0, f }
 *   nv12 { 2, 1, 1, 0x001, 0, f }       p010 { 2, 1, 1, 0x002, 0, f }          nv16 { 2, 1, 0, 0x001, 0, f } */
typedef struct ACHipFrameFormat {
    int planes;             /* 1: Y only; 2: Y + one plane of interleaved U, V; 3: Y, U, V */
    int chroma_shift_x;     /* log2 horizontal chroma subsampling: 1 (4:2:0, 4:2:2) or 0 (4:4:4) */
    int chroma_shift_y;     /* log2 vertical chroma subsampling: 1 (4:2:0) or 0 */
    int element_type;       /* 0x001: 8-bit samples; 0x002: 16-bit words (10/12/16-bit video) */
    int shift;              /* luma bits below the top of the word (LSB-aligned 10/12-bit: 6/4); 0 else */
    int chroma_filter;      /* ACHIP_CHROMA_BILINEAR or ACHIP_CHROMA_CATMULL_ROM (the reference CLI's) */
} ACHipFrameFormat;

typedef struct ACHipRequest {
    int kind;               /* ACHIP_REQ_* */
    int width, height;      /* input size (luma) */
    int factor;             /* 2, 4 or 8: log2(factor) chained 2x passes on luma */
    int stream;             /* >= 0: frame of video stream `stream` (LUMA/YUV420/YUV, factor 2, one
                               geometry and format per stream, hipGraph path); -1: one-off */
    const void* src[3];     /* planes: LUMA/RGB src[0]; YUV420 Y, U, V; YUV as `format` says.  Host memory
                               (page-locked for an asynchronous upload) or device memory */
    long src_stride[3];     /* byte strides */
    void* dst[3];           /* output planes (host or device), or all NULL to keep the output in the
                               server until achip_server_release (read it with achip_server_output) */
    long dst_stride[3];
    ACHipFrameFormat format;   /* ACHIP_REQ_YUV only */
} ACHipRequest;

/* Output geometry: luma fW x fH; RGB fW x fH x 3; chroma planes f times their input size.
 * Chroma of RGB requests is resized by Catmull-Rom (the reference's multi-channel path). */

typedef struct ACHipServerConfig {
    int precision;          /* ACHIP_FP32 / FP16 / BF16 */
    size_t hbm_budget;      /* bytes of device memory for request buffers (0: 16 GiB) */
    int step_tiles;         /* tile budget of one batched step (0: a default for the precision) */
    int max_workgroups;     /* grid cap of a batched step (0: a default that leaves CUs for frames) */
} ACHipServerConfig;

typedef struct ACHipServerStats {
    long completed_frames, completed_oneoffs, failed;
    double frame_p50_ms, frame_p99_ms, oneoff_p50_ms, oneoff_p99_ms;   /* submit -> output ready */
    double output_mpix;     /* output pixels of completed requests, millions */
    double busy_s;          /* wall time from the first submit to the last completion */
    long steps, graph_replays;
} ACHipServerStats;

/* A server runs one net.  ACNetLegacy nets run in any precision, their one-off passes
 * batched into persistent tile-queue launches; ACNet<8> nets in any precision and ArtCNN and
 * FSRCNNX nets in ACHIP_FP32 only (ACHIP_EINVAL otherwise) run each step's admitted one-offs'
 * next passes whole, with a feature-map workspace per request (and per video stream) counted
 * against the HBM budget. */
ACHIP_API int achip_server_create(ACHipNet* net, const ACHipServerConfig* config, ACHipServer** out);
/* Waits for all admitted work, then frees everything. */
ACHIP_API void achip_server_destroy(ACHipServer* server);
/* Queue a request (the server copies the descriptor; buffers must stay valid until done).
 * Fills *ticket.  Returns ACHIP_EINVAL for malformed requests and ACHIP_ENOMEM when the
 * request alone exceeds the budget. */
ACHIP_API int achip_server_submit(ACHipServer* server, const ACHipRequest* request, unsigned long long* ticket);
/* Block until the request is done (timeout_ms < 0: no limit).  Returns ACHIP_OK, the
 * request's error, or ACHIP_EIO on time-out; *latency_ms (optional) gets submit -> done. */
ACHIP_API int achip_server_wait(ACHipServer* server, unsigned long long ticket, double timeout_ms, double* latency_ms);
/* Device address and byte stride of output plane `plane` of a finished request kept in the
 * server (dst all NULL), valid until achip_server_release. */
ACHIP_API int achip_server_output(ACHipServer* server, unsigned long long ticket, int plane, void** ptr, long* stride);
/* Forget a finished request and free what it holds. */
ACHIP_API int achip_server_release(ACHipServer* server, unsigned long long ticket);
/* Statistics since creation or the last reset (reset != 0 clears them after reading). */
ACHIP_API int achip_server_stats(ACHipServer* server, ACHipServerStats* stats, int reset);

#ifdef __cplusplus
}
#endif

#endif
