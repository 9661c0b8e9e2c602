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
    int artf = 0;            // ArtCNN<F>: F (16 or 32), with artblocks body blocks
    int artblocks = 0;
    float* artk = nullptr;   // ArtCNN kernels and biases (reference layout) and A operands, one allocation
    float* artb = nullptr;
    float* artwa = nullptr;
};

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
