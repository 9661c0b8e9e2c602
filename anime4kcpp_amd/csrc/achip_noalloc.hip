// achip_pass_batch_noalloc: the batched persistent pass without a caller workspace and
// without any device allocation, for callers that cannot allocate (code under stream
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The tile-queue counter and the image descriptors live in `scratch`, a file-scope device
// array of fixed size (ACHIP_NOALLOC_MAX_IMAGES descriptors).  Host code never passes the
// array's name to the device: on the first call on each device, with that device current,
// hipGetSymbolAddress gives that device's own copy's address, which is kept per device and
// used for the descriptor copy and as the kernel's workspace argument.  The host side stages
// the descriptors in arrays on the calling thread's stack, so nothing is allocated on the
// success path, host or device.  Each call rewrites the whole header (the counter starts at
// zero) and its descriptors; calls that overlap in time on one device share the array and
// race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "achip.h"
#include "achip_internal.hpp"

namespace {

constexpr size_t kScratchBytes = achip::kBatchHeaderBytes + ACHIP_NOALLOC_MAX_IMAGES * sizeof(achip::BatchDesc);
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mtx;
void* g_scratch_ptr[kMaxDevices];

// This device's address of `scratch`; the device must be current.
int scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return achip::record_error(ACHIP_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return achip::record_error(ACHIP_EIO, "hipGetSymbolAddress(scratch) failed");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return ACHIP_OK;
}

// Makes `device` current for the scope of one call and restores the caller's device after.
struct CurrentDevice {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit CurrentDevice(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) err = hipSetDevice(dev);
    }
    ~CurrentDevice()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

} // namespace

extern "C" {

size_t achip_noalloc_scratch_bytes(void) { return kScratchBytes; }

int achip_pass_batch_noalloc(ACHipNet* net, int precision, int config, int src_type, int dst_type,
                             const ACHipImage* images, int n, void* stream)
{
    if (!net) return achip::record_error(ACHIP_EINVAL, "net is NULL");
    if (n <= 0) return achip::record_error(ACHIP_EINVAL, "empty batch");
    if (n > ACHIP_NOALLOC_MAX_IMAGES)
        return achip::record_error(ACHIP_EINVAL,
                                   "achip_pass_batch_noalloc takes at most ACHIP_NOALLOC_MAX_IMAGES (256) images per call");
    const int device = achip_net_device(net);
    CurrentDevice g(device);
    if (g.err != hipSuccess) return achip::record_error(ACHIP_EIO, "hipSetDevice failed");
    void* ws = nullptr;
    int r = scratch_address(device, &ws);
    if (r) return r;
    alignas(16) char host[kScratchBytes];
    achip::BatchSpan spans[2 * ACHIP_NOALLOC_MAX_IMAGES + 1];
    int total = 0;
    r = achip::batch_prepare(net, precision, config, src_type, dst_type, images, n, ws, kScratchBytes, host, spans, &total);
    if (r) return r;
    return achip::batch_launch(net, precision, config, src_type, dst_type, n, total, ws, host,
                               static_cast<hipStream_t>(stream));
}

} // extern "C"
