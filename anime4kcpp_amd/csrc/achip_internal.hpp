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

} // namespace achip
