// Host-visible launch entry points of the ACNetLegacy HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace achip {

// One tile shape of the fused FP32 kernel: TW x TH input-resolution pixels per
// workgroup, NT threads, P-pixel vertical runs per thread.
struct Fp32Config {
    int tw, th, nt, p;
    int lds_bytes;
};
extern const Fp32Config kFp32Configs[];
extern const int kFp32ConfigCount;

// Fused FP32 2x pass.  src/dst are device pointers with byte strides; stype/dtype are
// element types (0x001 u8, 0x002 u16, 0x202 f16, 0x204 f32); wts is the device weight
// blob (acnet_device.hpp layout).  The caller has validated every argument.
hipError_t launch_fp32(int cfg, int stype, int dtype, const void* src, long sstride, int W, int H,
                       void* dst, long dstride, const float* wts, hipStream_t stream);

} // namespace achip
