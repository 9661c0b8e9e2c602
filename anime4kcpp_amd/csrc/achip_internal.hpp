// Library-internal pieces shared by the device C ABI (achip.cpp) and the serving loop
// (achip_server.cpp): the net object and unvalidated launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "achip.h"
#include "acnet_kernels.hpp"

struct ACHipNet {
    int device;
    float* wts;      // device FP32 weight blob, acnet_device.hpp layout (ACNetLegacy)
    void* wa16;      // MFMA A operands, FP16
    void* wabf16;    // MFMA A operands, BF16
    float* wa32;     // FP32 MFMA A operands
    int blocks = 0;          // ACNet<8>: its block count (the fields above are then unused)
    float* a8wts = nullptr;  // ACNet<8> blob, acnet_device.hpp layout (one allocation with a8wa)
    float* a8wa = nullptr;   // ACNet<8> A operands: blocks 1..B, then the final layer
    void* a8wa16 = nullptr;  // the same layers' 16-bit MFMA A operands, FP16 then BF16
    void* a8wabf16 = nullptr;
    // ArtCNN<F> and FSRCNNX<F>: family, F channels, B blocks; the reference-layout kernels,
    // biases and slopes and the per-lane A operands of the conv3x3 layers and the conv1x1,
    // one allocation starting at lk
    int lfamily = 0;         // kNetArtCNN or kNetFSRCNNX; 0 for the nets above
    int lf = 0;
    int lblocks = 0;
    float* lk = nullptr;
    float* lb = nullptr;
    float* la = nullptr;
    float* lwa = nullptr;
    float* lw1 = nullptr;
};

constexpr int kNetArtCNN = 1, kNetFSRCNNX = 2;

namespace achip {

// One 2x pass with an explicit configuration; arguments already validated, device current.
hipError_t launch_pass(const ACHipNet* net, int precision, int config, int stype, int dtype, const void* src,
                       long sstride, int w, int h, void* dst, long dstride, hipStream_t stream);

// One persistent batched pass over `total` tiles described by n device descriptors;
// counter (device int) must be zero when the kernel starts.  max_wg > 0 caps the grid.
hipError_t launch_batch(const ACHipNet* net, int precision, int config, int stype, int dtype, const BatchDesc* desc,
                        int n, int total, int* counter, int max_wg, hipStream_t stream);

// The configuration ACHIP_AUTO resolves to for this shape (autotuner cache, else 0).
int auto_config(const ACHipNet* net, int precision, int stype, int dtype, int w, int h);

// A batch workspace: the tile-queue counter, padded to this many bytes, then one
// BatchDesc per image.
constexpr size_t kBatchHeaderBytes = 256;

// A byte range [lo, hi) that a batched launch reads or writes; image -1 is the workspace.
struct BatchSpan {
    uintptr_t lo, hi;
    bool written;
    int image;
};

// Checks a batch (achip_pass_batch's arguments; the workspace only for size, alignment and
// overlap) and writes its header (counter 0) and descriptors into host[0 ..
// achip_batch_workspace_bytes(n)); spans has room for 2n + 1 entries.  No allocation on the
// success path.  Returns ACHIP_OK or an ACHIP_* code (recorded for achip_last_error).
int batch_prepare(ACHipNet* net, int precision, int config, int src_type, int dst_type, const ACHipImage* images, int n,
                  const void* workspace, size_t ws_bytes, char* host, BatchSpan* spans, int* total_tiles);

// The host-to-device copy of a prepared batch into `workspace` and the persistent launch, on
// `stream`, with the net's device current.
int batch_launch(ACHipNet* net, int precision, int config, int src_type, int dst_type, int n, int total, void* workspace,
                 const char* host, hipStream_t stream);

// Records msg for achip_last_error; returns code.
int record_error(int code, const char* msg);

} // namespace achip
