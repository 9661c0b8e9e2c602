// Thin device-level C ABI (include/achip.h): weight upload, argument validation and
// kernel dispatch for the ACNetLegacy HIP path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <cstring>
#include <string>
#include <vector>

#include "achip.h"
#include "achip_internal.hpp"
#include "acnet_device.hpp"
#include "acnet_kernels.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(ACHIP_EIO, "%s: %s", what, hipGetErrorString(e));
}

// Restores the calling thread's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) err = hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// IEEE binary32 -> binary16 / bfloat16 bits, round to nearest even (finite inputs)
uint16_t f32_to_f16_bits(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);
    if (ax < 0x38800000u) {
        float a;
        std::memcpy(&a, &ax, 4);
        return static_cast<uint16_t>(sign | static_cast<uint32_t>(std::nearbyint(a * 16777216.0f)));
    }
    uint32_t hv = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (hv & 1u))) hv++;
    return static_cast<uint16_t>(sign | hv);
}
uint16_t f32_to_bf16_bits(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    x += 0x7fffu + ((x >> 16) & 1u);
    return static_cast<uint16_t>(x >> 16);
}

// Per-lane MFMA A operands (acnet_device.hpp) for one 16-bit type.
std::vector<uint16_t> mfma_operands(const float* kernels, bool bf16)
{
    std::vector<uint16_t> a(static_cast<size_t>(8) * achip::kMfmaLayerHalves);
    for (int l = 1; l <= 8; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        for (int t = 0; t < 3; t++)
            for (int lane = 0; lane < 64; lane++) {
                const int i = lane % 16, g = lane / 16;
                const int n = i % 8, ro = i / 8, ky = g - ro, kx = t;
                for (int c = 0; c < 8; c++) {
                    const float w = (ky >= 0 && ky <= 2) ? k[n * 72 + (ky * 3 + kx) * 8 + c] : 0.0f;
                    a[(static_cast<size_t>(l - 1) * 3 * 64 + t * 64 + lane) * 8 + c] = bf16 ? f32_to_bf16_bits(w) : f32_to_f16_bits(w);
                }
            }
    }
    return a;
}

// Per-lane 16-bit MFMA A operands of one conv3x3 8 -> cout layer (k in the reference layout
// k[n][tap][c]), [t][lane] x 8 values, output rows n >= cout zero.
void mfma16_conv8_operands(const float* k, int cout, bool bf16, uint16_t* out)
{
    for (int t = 0; t < 3; t++)
        for (int lane = 0; lane < 64; lane++) {
            const int i = lane % 16, g = lane / 16;
            const int n = i % 8, ro = i / 8, ky = g - ro;
            for (int c = 0; c < 8; c++) {
                const float w = (n < cout && ky >= 0 && ky <= 2) ? k[n * 72 + (ky * 3 + t) * 8 + c] : 0.0f;
                out[(t * 64 + lane) * 8 + c] = bf16 ? f32_to_bf16_bits(w) : f32_to_f16_bits(w);
            }
        }
}

// Per-lane FP32 MFMA A operands (acnet_device.hpp).
std::vector<float> mfma32_operands(const float* kernels)
{
    std::vector<float> a(achip::kMfma32BlobFloats, 0.0f);
    for (int t = 0; t < 3; t++)
        for (int lane = 0; lane < 64; lane++) {
            const int i = lane % 16, g = lane / 16, n = i % 8, ky = g - i / 8;
            if (ky >= 0 && ky <= 2) a[t * 64 + lane] = kernels[n * 9 + ky * 3 + t];
        }
    for (int l = 1; l <= 8; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        float* al = a.data() + achip::kMfma32L0Floats + (l - 1) * achip::kMfma32LayerFloats;
        for (int t = 0; t < 3; t++)
            for (int c = 0; c < 8; c++)
                for (int lane = 0; lane < 64; lane++) {
                    const int i = lane % 16, g = lane / 16, n = i % 8, ky = g - i / 8;
                    if (ky >= 0 && ky <= 2) al[(t * 8 + c) * 64 + lane] = k[n * 72 + (ky * 3 + t) * 8 + c];
                }
    }
    return a;
}

int elem_size(int type)
{
    switch (type) {
    case achip::U8: return 1;
    case achip::U16: return 2;
    case achip::F16: return 2;
    case achip::F32: return 4;
    default: return 0;
    }
}

// The buffer [ptr, ptr + bytes) must be device (or managed) memory of `device`, inside
// one allocation.  A failed query leaves a sticky runtime error, which is cleared here
// so that it is not reported by the next launch check.
int check_device_buffer(const void* ptr, size_t bytes, int device, const char* name)
{
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ACHIP_EINVAL, "%s: %p is not HIP device memory", name, ptr);
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
        return fail(ACHIP_EINVAL, "%s: %p is not device memory (memory type %d)", name, ptr, (int)attr.type);
    if (attr.type == hipMemoryTypeDevice && attr.device != device)
        return fail(ACHIP_EINVAL, "%s: %p belongs to device %d, the net is on device %d", name, ptr, attr.device, device);
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    e = hipMemGetAddressRange(&base, &size, const_cast<void*>(ptr));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ACHIP_EINVAL, "%s: cannot query the allocation of %p", name, ptr);
    }
    const char* b = static_cast<const char*>(base);
    const char* p = static_cast<const char*>(ptr);
    if (p < b || static_cast<size_t>(p - b) + bytes > size)
        return fail(ACHIP_EINVAL, "%s: image needs %zu bytes at %p, allocation holds %zu from %p",
                    name, bytes, ptr, size - static_cast<size_t>(p - b), base);
    return ACHIP_OK;
}

int validate(const ACHipNet* net, const void* src, long sstride, int stype, int w, int h,
             const void* dst, long dstride, int dtype)
{
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    const int ies = elem_size(stype), oes = elem_size(dtype);
    if (!ies) return fail(ACHIP_EINVAL, "unsupported source element type 0x%x", stype);
    if (!oes) return fail(ACHIP_EINVAL, "unsupported destination element type 0x%x", dtype);
    if (w <= 0 || h <= 0) return fail(ACHIP_EINVAL, "empty image %dx%d", w, h);
    if (w > (1 << 29) || h > (1 << 29)) return fail(ACHIP_EINVAL, "image %dx%d too large", w, h);
    if (!src || !dst) return fail(ACHIP_EINVAL, "NULL image pointer");
    if (sstride < static_cast<long>(w) * ies || sstride % ies)
        return fail(ACHIP_EINVAL, "source stride %ld invalid for width %d x %d bytes", sstride, w, ies);
    if (dstride < 2L * w * oes || dstride % oes)
        return fail(ACHIP_EINVAL, "destination stride %ld invalid for width %d x %d bytes", dstride, 2 * w, oes);
    if (reinterpret_cast<uintptr_t>(src) % ies || reinterpret_cast<uintptr_t>(dst) % oes)
        return fail(ACHIP_EINVAL, "image pointer not aligned to its element size");
    const size_t sbytes = static_cast<size_t>(h - 1) * sstride + static_cast<size_t>(w) * ies;
    const size_t dbytes = static_cast<size_t>(2 * h - 1) * dstride + static_cast<size_t>(2 * w) * oes;
    int r = check_device_buffer(src, sbytes, net->device, "src");
    if (r) return r;
    r = check_device_buffer(dst, dbytes, net->device, "dst");
    if (r) return r;
    const char* s0 = static_cast<const char*>(src);
    const char* d0 = static_cast<const char*>(dst);
    if (s0 < d0 + dbytes && d0 < s0 + sbytes) return fail(ACHIP_EINVAL, "src and dst overlap");
    return ACHIP_OK;
}

constexpr size_t kBatchHeader = achip::kBatchHeaderBytes;

// autotuner cache: (device, precision, src type, dst type, log2 w, log2 h) -> config
std::mutex g_tune_mtx;
std::map<std::tuple<int, int, int, int, int, int>, int> g_tune;

int ceil_log2i(int v)
{
    int l = 0;
    while ((1 << l) < v && l < 30) l++;
    return l;
}

std::tuple<int, int, int, int, int, int> tune_key(int device, int precision, int st, int dt, int w, int h)
{
    return { device, precision, st, dt, ceil_log2i(w), ceil_log2i(h) };
}

// the autotuner's form for achip_pass_ws, same key: true = the layer-split form
std::map<std::tuple<int, int, int, int, int, int>, bool> g_split;

int compute_units(int device)
{
    static std::mutex mtx;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cache.find(device);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        n = 256;
    }
    cache[device] = n;
    return n;
}

// The form achip_pass_ws runs: the layer-split one only for FP32 ACNetLegacy passes, and there
// the autotuner's verdict for the size class, else the rule measured on MI355X (u8 1080p:
// fused 328 us, split 299 us; 4K: 1307 / 1042 us; 256 x 256: 82 / 52 us; 960 x 540: 84 /
// 109 us): the fused kernel wins only where its grid of 64 x 32 tiles fits in one round of
// the device's compute units and the image is larger than 256 x 256.
bool split_preferred(const ACHipNet* net, int precision, int st, int dt, int w, int h)
{
    if (precision != ACHIP_FP32 || net->blocks || net->lfamily) return false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        auto it = g_split.find(tune_key(net->device, precision, st, dt, w, h));
        if (it != g_split.end()) return it->second;
    }
    const long tiles = static_cast<long>((w + 63) / 64) * ((h + 31) / 32);
    return !(tiles <= compute_units(net->device) && static_cast<long>(w) * h > 256L * 256);
}

} // namespace

namespace achip {

int persistent_grid(const void* kernel, int threads, int tiles, int max_wg)
{
    static std::mutex mtx;
    static std::map<std::pair<const void*, int>, int> cache;   // (kernel, device) -> workgroups
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    int per_device = 0;
    {
        std::lock_guard<std::mutex> lk(mtx);
        auto it = cache.find({ kernel, dev });
        if (it != cache.end()) per_device = it->second;
    }
    if (!per_device) {
        int blocks = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess) return -1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
        per_device = std::max(1, blocks) * cus;
        std::lock_guard<std::mutex> lk(mtx);
        cache[{ kernel, dev }] = per_device;
    }
    if (max_wg > 0) per_device = std::min(per_device, max_wg);
    return std::min(per_device, std::max(tiles, 1));
}

hipError_t launch_pass(const ACHipNet* net, int precision, int config, int stype, int dtype, const void* src,
                       long sstride, int w, int h, void* dst, long dstride, hipStream_t stream)
{
    if (precision == ACHIP_FP32 && config < kMfma32ConfigCount)
        return launch_mfma32(config, stype, dtype, src, sstride, w, h, dst, dstride, net->wts, net->wa32, stream);
    if (precision == ACHIP_FP32)
        return launch_fp32(config - kMfma32ConfigCount, stype, dtype, src, sstride, w, h, dst, dstride, net->wts, stream);
    return launch_mfma(precision == ACHIP_BF16, config, stype, dtype, src, sstride, w, h, dst, dstride, net->wts,
                       precision == ACHIP_BF16 ? net->wabf16 : net->wa16, net->wa32, stream);
}

hipError_t launch_batch(const ACHipNet* net, int precision, int config, int stype, int dtype, const BatchDesc* desc,
                        int n, int total, int* counter, int max_wg, hipStream_t stream)
{
    if (precision == ACHIP_FP32 && config < kMfma32ConfigCount)
        return launch_mfma32_batch(config, stype, dtype, desc, n, total, counter, net->wts, net->wa32, max_wg, stream);
    if (precision == ACHIP_FP32)
        return launch_fp32_batch(config - kMfma32ConfigCount, stype, dtype, desc, n, total, counter, net->wts, max_wg,
                                 stream);
    return launch_mfma_batch(precision == ACHIP_BF16, config, stype, dtype, desc, n, total, counter, net->wts,
                             precision == ACHIP_BF16 ? net->wabf16 : net->wa16, net->wa32, max_wg, stream);
}

// Without a tuned entry: config 0, except that the 16-bit forms take their 64x64 tile at 1024
// threads (config 1) for inputs wider than u8.  Config 0's 64x32 tile at 512 threads wins by
// running two workgroups per CU, and only with u8 input do two fit the CU's 160 KB of LDS
// (80 KB each; 85 KB with 16-bit input, 95 KB with f32).  Measured: 4K f32 -> 8K -> 16K in FP16
// took 1.55 ms with config 0 against 1.28 ms with config 1.
int auto_config(const ACHipNet* net, int precision, int stype, int dtype, int w, int h)
{
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    auto it = g_tune.find(tune_key(net->device, precision, stype, dtype, w, h));
    if (it != g_tune.end()) return it->second;
    return precision != ACHIP_FP32 && stype != U8 ? 1 : 0;
}

int record_error(int code, const char* msg)
{
    return fail(code, "%s", msg);
}

// Sweep in address order: a span overlaps an earlier one iff it starts below that one's end.
// A written span may overlap nothing; a read span may overlap other read spans only.
static int check_disjoint_writes(BatchSpan* spans, int count)
{
    std::sort(spans, spans + count, [](const BatchSpan& a, const BatchSpan& b) { return a.lo < b.lo; });
    uintptr_t end_any = 0, end_written = 0;
    const BatchSpan* last_any = nullptr;
    const BatchSpan* last_written = nullptr;
    for (int i = 0; i < count; i++) {
        const BatchSpan& s = spans[i];
        const BatchSpan* hit = s.written ? (s.lo < end_any ? last_any : nullptr) : (s.lo < end_written ? last_written : nullptr);
        if (hit) {
            auto name = [](const BatchSpan& x) {
                return x.image < 0 ? std::string("the workspace")
                                   : "image " + std::to_string(x.image) + (x.written ? " dst" : " src");
            };
            return fail(ACHIP_EINVAL, "%s overlaps %s", name(*hit).c_str(), name(s).c_str());
        }
        if (s.hi > end_any) { end_any = s.hi; last_any = &s; }
        if (s.written && s.hi > end_written) { end_written = s.hi; last_written = &s; }
    }
    return ACHIP_OK;
}

int batch_prepare(ACHipNet* net, int precision, int config, int src_type, int dst_type, const ACHipImage* images, int n,
                  const void* workspace, size_t ws_bytes, char* host, BatchSpan* spans, int* total_tiles)
{
    if (precision != ACHIP_FP32 && precision != ACHIP_FP16 && precision != ACHIP_BF16)
        return fail(ACHIP_EINVAL, "precision %d not available", precision);
    if (config < 0 || config >= achip_config_count(precision)) return fail(ACHIP_EINVAL, "no config %d", config);
    if (!images || n <= 0) return fail(ACHIP_EINVAL, "empty batch");
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (net->blocks || net->lfamily) return fail(ACHIP_EINVAL, "batched passes take an ACNetLegacy net");
    const size_t need = achip_batch_workspace_bytes(n);
    if (!workspace || ws_bytes < need)
        return fail(ACHIP_EINVAL, "workspace of %zu bytes, the batch needs %zu", ws_bytes, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 256) return fail(ACHIP_EINVAL, "workspace not 256-byte aligned");
    int tw = 0, th = 0;
    achip_config_info(precision, config, &tw, &th, nullptr, nullptr, nullptr);
    std::memset(host, 0, kBatchHeaderBytes);
    auto* desc = reinterpret_cast<BatchDesc*>(host + kBatchHeaderBytes);
    long total = 0;
    int ns = 0;
    for (int i = 0; i < n; i++) {
        const ACHipImage& im = images[i];
        const int r = validate(net, im.src, im.src_stride, src_type, im.w, im.h, im.dst, im.dst_stride, dst_type);
        if (r) {
            std::string msg = g_error;
            return fail(r, "image %d: %s", i, msg.c_str());
        }
        const uintptr_t s0 = reinterpret_cast<uintptr_t>(im.src), d0 = reinterpret_cast<uintptr_t>(im.dst);
        spans[ns++] = { s0, s0 + static_cast<size_t>(im.h - 1) * im.src_stride + static_cast<size_t>(im.w) * (src_type & 0xff), false, i };
        spans[ns++] = { d0, d0 + static_cast<size_t>(2 * im.h - 1) * im.dst_stride + static_cast<size_t>(2 * im.w) * (dst_type & 0xff), true, i };
        const int tx = (im.w + tw - 1) / tw, ty = (im.h + th - 1) / th;
        desc[i] = BatchDesc{ im.src, im.dst, im.src_stride, im.dst_stride, im.w, im.h, tx, static_cast<int>(total), 0 };
        total += static_cast<long>(tx) * ty;
        if (total > (1L << 30)) return fail(ACHIP_EINVAL, "batch too large (%ld tiles)", total);
    }
    // Every launch writes the workspace and each dst while it reads each src: no written span
    // may overlap any other span (an image's own src/dst pair was checked above).
    spans[ns++] = { reinterpret_cast<uintptr_t>(workspace), reinterpret_cast<uintptr_t>(workspace) + need, true, -1 };
    const int r = check_disjoint_writes(spans, ns);
    if (r) return r;
    *total_tiles = static_cast<int>(total);
    return ACHIP_OK;
}

int batch_launch(ACHipNet* net, int precision, int config, int src_type, int dst_type, int n, int total, void* workspace,
                 const char* host, hipStream_t st)
{
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    // counter (= 0) and descriptors in one copy; from pageable memory the copy is staged
    // before the call returns, so `host` may go out of scope
    hipError_t e = hipMemcpyAsync(workspace, host, achip_batch_workspace_bytes(n), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(batch descriptors)");
    auto* ddesc = reinterpret_cast<const BatchDesc*>(static_cast<char*>(workspace) + kBatchHeaderBytes);
    e = launch_batch(net, precision, config, src_type, dst_type, ddesc, n, total, static_cast<int*>(workspace), 0, st);
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
    return ACHIP_OK;
}

} // namespace achip

extern "C" {

int achip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int achip_net_create(int device, const float* kernels, const float* biases, ACHipNet** out)
{
    if (!out) return fail(ACHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!kernels || !biases) return fail(ACHIP_EINVAL, "NULL weights");
    const int n = achip_device_count();
    if (n <= 0) return fail(ACHIP_ENODEV, "no HIP device");
    if (device < 0 || device >= n) {
        // best device by compute units x clock, as the reference picks its GPU when the
        // index is out of range (CUDAProcessor.cpp:233-251)
        long best = -1;
        for (int i = 0; i < n; i++) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
            long score = static_cast<long>(prop.multiProcessorCount) * prop.clockRate;
            if (score > best) { best = score; device = i; }
        }
        if (device < 0 || device >= n) device = 0;
    }
    // repack the reference layout into the kernels' layout
    std::vector<float> blob(achip::kBlobFloats);
    for (int p = 0; p < 9; p++)
        for (int o = 0; o < 8; o++) blob[achip::kOffW0 + p * 8 + o] = kernels[o * 9 + p];
    for (int o = 0; o < 8; o++) blob[achip::kOffB0 + o] = biases[o];
    for (int l = 1; l <= 8; l++) {
        const float* k = kernels + 72 + 576 * (l - 1);
        for (int p = 0; p < 9; p++)
            for (int c = 0; c < 8; c++)
                for (int o = 0; o < 8; o++) blob[achip::offW(l) + (p * 8 + c) * 8 + o] = k[o * 72 + p * 8 + c];
        for (int o = 0; o < 8; o++) blob[achip::offB(l) + o] = biases[8 * l + o];
    }
    const float* k2 = kernels + 72 + 576 * 8;
    for (int c = 0; c < 8; c++)
        for (int q = 0; q < 4; q++) blob[achip::kOffK2 + c * 4 + q] = k2[q * 8 + c];

    const std::vector<uint16_t> a16 = mfma_operands(kernels, false);
    const std::vector<uint16_t> abf = mfma_operands(kernels, true);
    const std::vector<float> a32 = mfma32_operands(kernels);

    DeviceGuard g(device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    // one allocation: FP32 blob, then the two A-operand blobs, each 256-byte aligned
    const size_t off16 = (blob.size() * sizeof(float) + 255) / 256 * 256;
    const size_t offbf = off16 + achip::kMfmaBlobBytes;
    const size_t off32 = offbf + achip::kMfmaBlobBytes;
    const size_t total = off32 + a32.size() * sizeof(float);
    char* d = nullptr;
    hipError_t e = hipMalloc(&d, total);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(weights)");
    e = hipMemcpy(d, blob.data(), blob.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + off16, a16.data(), achip::kMfmaBlobBytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + offbf, abf.data(), achip::kMfmaBlobBytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + off32, a32.data(), a32.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "hipMemcpy(weights)");
    }
    *out = new ACHipNet{ device, reinterpret_cast<float*>(d), d + off16, d + offbf, reinterpret_cast<float*>(d + off32) };
    g_error.clear();
    return ACHIP_OK;
}

void achip_net_destroy(ACHipNet* net)
{
    if (!net) return;
    {
        DeviceGuard g(net->device);
        if (net->wts) (void)hipFree(net->wts);
        if (net->a8wts) (void)hipFree(net->a8wts);
        if (net->lk) (void)hipFree(net->lk);
    }
    delete net;
}

// ---- ACNet<8> ----


int achip_net8_create(int device, int blocks, const float* kernels, const float* biases, const float* alphas, ACHipNet** out)
{
    if (!out) return fail(ACHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!kernels || !biases || !alphas) return fail(ACHIP_EINVAL, "NULL weights");
    if (blocks < 1 || blocks > 64) return fail(ACHIP_EINVAL, "blocks %d out of range", blocks);
    const int n = achip_device_count();
    if (n <= 0) return fail(ACHIP_ENODEV, "no HIP device");
    if (device < 0 || device >= n) device = 0;
    std::vector<float> blob(achip::acnet8_blob_floats(blocks), 0.0f);
    for (int p = 0; p < 9; p++)
        for (int o = 0; o < 8; o++) blob[achip::kA8OffW0 + p * 8 + o] = kernels[o * 9 + p];
    for (int o = 0; o < 8; o++) {
        blob[achip::kA8OffB0 + o] = biases[o];
        blob[achip::kA8OffAlpha0 + o] = alphas[o];
    }
    for (int o = 0; o < 4; o++) blob[achip::kA8OffFinalB + o] = biases[8 + 8 * blocks + o];
    for (int l = 1; l <= blocks; l++)
        for (int o = 0; o < 8; o++) {
            blob[achip::kA8OffLayer + 16 * (l - 1) + o] = biases[8 * l + o];
            blob[achip::kA8OffLayer + 16 * (l - 1) + 8 + o] = alphas[8 * l + o];
        }
    std::vector<float> wa(static_cast<size_t>(blocks + 1) * achip::kMfma32LayerFloats);
    for (int l = 1; l <= blocks + 1; l++)
        achip::pack_conv8_operands(kernels + 72 + 576 * (l - 1), l <= blocks ? 8 : 4, wa.data() + (l - 1) * achip::kMfma32LayerFloats);

    // the same layers as 16-bit operands: FP16 for all layers, then BF16
    const size_t per16 = static_cast<size_t>(3) * 64 * 8;
    std::vector<uint16_t> wa16(2 * static_cast<size_t>(blocks + 1) * per16);
    for (int b16 = 0; b16 < 2; b16++)
        for (int l = 1; l <= blocks + 1; l++)
            mfma16_conv8_operands(kernels + 72 + 576 * (l - 1), l <= blocks ? 8 : 4, b16 != 0,
                                  wa16.data() + (static_cast<size_t>(b16) * (blocks + 1) + (l - 1)) * per16);

    DeviceGuard g(device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const size_t offa = (blob.size() * sizeof(float) + 255) / 256 * 256;
    const size_t off16 = offa + (wa.size() * sizeof(float) + 255) / 256 * 256;
    const size_t half16 = static_cast<size_t>(blocks + 1) * per16 * sizeof(uint16_t);
    char* d = nullptr;
    hipError_t e = hipMalloc(&d, off16 + wa16.size() * sizeof(uint16_t));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(weights)");
    e = hipMemcpy(d, blob.data(), blob.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + offa, wa.data(), wa.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + off16, wa16.data(), wa16.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "hipMemcpy(weights)");
    }
    auto* net = new ACHipNet{ device, nullptr, nullptr, nullptr, nullptr };
    net->blocks = blocks;
    net->a8wts = reinterpret_cast<float*>(d);
    net->a8wa = reinterpret_cast<float*>(d + offa);
    net->a8wa16 = d + off16;
    net->a8wabf16 = d + off16 + half16;
    *out = net;
    g_error.clear();
    return ACHIP_OK;
}

int achip_net_blocks(const ACHipNet* net) { return net ? (net->lfamily ? net->lblocks : net->blocks) : ACHIP_EINVAL; }

// ---- ArtCNN<F> ----

namespace {
// Uploads a layered net (ArtCNN, FSRCNNX): kernels, biases, slopes (may be empty), the
// conv3x3 A operands and the conv1x1 ones (may be empty), each 256-byte aligned, in one
// allocation.
int upload_layered(int device, int family, int f, int blocks, const float* kernels, size_t klen, const float* biases,
                   size_t blen, const float* alphas, size_t alen, const std::vector<float>& wa,
                   const std::vector<float>& w1, ACHipNet** out)
{
    auto up = [](size_t floats) { return (floats * sizeof(float) + 255) / 256 * 256; };
    const size_t offb = up(klen), offa = offb + up(blen), offwa = offa + up(alen), offw1 = offwa + up(wa.size());
    DeviceGuard g(device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    char* d = nullptr;
    hipError_t e = hipMalloc(&d, offw1 + up(w1.size()));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(weights)");
    e = hipMemcpy(d, kernels, klen * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + offb, biases, blen * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && alen) e = hipMemcpy(d + offa, alphas, alen * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + offwa, wa.data(), wa.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && !w1.empty()) e = hipMemcpy(d + offw1, w1.data(), w1.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "hipMemcpy(weights)");
    }
    auto* net = new ACHipNet{ device, nullptr, nullptr, nullptr, nullptr };
    net->lfamily = family;
    net->lf = f;
    net->lblocks = blocks;
    net->lk = reinterpret_cast<float*>(d);
    net->lb = reinterpret_cast<float*>(d + offb);
    net->la = alen ? reinterpret_cast<float*>(d + offa) : nullptr;
    net->lwa = reinterpret_cast<float*>(d + offwa);
    net->lw1 = w1.empty() ? nullptr : reinterpret_cast<float*>(d + offw1);
    *out = net;
    g_error.clear();
    return ACHIP_OK;
}

// conv3x3 A operands of layers F -> F ... F -> 4 laid end to end (cout of each given)
static std::vector<float> pack_layers(int f, const float* k0, size_t kstride, const std::vector<int>& couts)
{
    const size_t wl = achip::conv_operand_floats(f);
    std::vector<float> wa(wl * couts.size(), 0.0f);
    for (size_t i = 0; i < couts.size(); i++) {
        if (f == 8) achip::pack_conv8_operands(k0 + kstride * i, couts[i], wa.data() + wl * i);
        else achip::pack_convf_operands(k0 + kstride * i, f, couts[i], wa.data() + wl * i);
    }
    return wa;
}
} // namespace

// ---- ArtCNN<F> ----

int achip_net_artcnn_create(int device, int channels, int blocks, const float* kernels, const float* biases, ACHipNet** out)
{
    if (!out) return fail(ACHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!kernels || !biases) return fail(ACHIP_EINVAL, "NULL weights");
    if (channels != 16 && channels != 32) return fail(ACHIP_EINVAL, "ArtCNN has 16 or 32 channels, not %d", channels);
    if (blocks < 1 || blocks > 64) return fail(ACHIP_EINVAL, "blocks %d out of range", blocks);
    const int n = achip_device_count();
    if (n <= 0) return fail(ACHIP_ENODEV, "no HIP device");
    if (device < 0 || device >= n) device = 0;
    const int F = channels;
    const size_t klen = static_cast<size_t>(F) * 9 + static_cast<size_t>(F) * F * 9 * (blocks + 1) + F * 4 * 9;   // ArtCNN.hpp:28
    const size_t blen = static_cast<size_t>(F) + static_cast<size_t>(F) * (blocks + 1) + 4;
    // the B body layers and the residual layer (F -> F), then the tail (F -> 4)
    std::vector<int> couts(blocks + 1, F);
    couts.push_back(4);
    const std::vector<float> wa = pack_layers(F, kernels + F * 9, static_cast<size_t>(F) * F * 9, couts);
    return upload_layered(device, kNetArtCNN, F, blocks, kernels, klen, biases, blen, nullptr, 0, wa, {}, out);
}

// ---- FSRCNNX<F> ----

int achip_net_fsrcnnx_create(int device, int channels, int blocks, const float* kernels, const float* biases,
                             const float* alphas, ACHipNet** out)
{
    if (!out) return fail(ACHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!kernels || !biases || !alphas) return fail(ACHIP_EINVAL, "NULL weights");
    if (channels != 8 && channels != 16) return fail(ACHIP_EINVAL, "FSRCNNX has 8 or 16 channels, not %d", channels);
    if (blocks < 2 || blocks > 64) return fail(ACHIP_EINVAL, "blocks %d out of range", blocks);
    const int n = achip_device_count();
    if (n <= 0) return fail(ACHIP_ENODEV, "no HIP device");
    if (device < 0 || device >= n) device = 0;
    const int F = channels;
    const size_t k3 = static_cast<size_t>(F) * F * 9;
    const size_t klen = static_cast<size_t>(F) * 25 + k3 * blocks + F * F + F * 4 * 9;   // FSRCNNX.hpp:29
    const size_t blen = static_cast<size_t>(F) + static_cast<size_t>(F) * (blocks + 1) + 4;
    const size_t alen = static_cast<size_t>(F) * (blocks + 1);
    // the B conv3x3 layers (F -> F), then the tail (F -> 4), which follows the conv1x1
    std::vector<float> wa = pack_layers(F, kernels + F * 25, k3, std::vector<int>(blocks, F));
    const float* k1 = kernels + F * 25 + k3 * blocks;
    const std::vector<float> wt = pack_layers(F, k1 + F * F, 0, { 4 });
    wa.insert(wa.end(), wt.begin(), wt.end());
    std::vector<float> w1(achip::conv_1x1_floats(F));
    if (F == 8) achip::pack_conv8_1x1(k1, w1.data());
    else achip::pack_convf_1x1(k1, F, w1.data());
    return upload_layered(device, kNetFSRCNNX, F, blocks, kernels, klen, biases, blen, alphas, alen, wa, w1, out);
}

size_t achip_net_workspace_bytes(const ACHipNet* net, int w, int h)
{
    if (!net || w <= 0 || h <= 0) return 0;
    if (net->lfamily) {
        const long fp = net->lfamily == kNetArtCNN ? achip::artcnn_feature_pitch(net->lf, w) : achip::fsrcnnx_feature_pitch(net->lf, w);
        return 3 * static_cast<size_t>(fp) * static_cast<size_t>(h);
    }
    if (net->blocks) return achip_pass8_workspace_bytes(w, h);
    return 0;
}

int achip_pass_net(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h, void* dst, long dst_stride,
                   int dst_type, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (!net->lfamily && !net->blocks)
        return achip_pass(net, ACHIP_FP32, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, stream);
    if (net->blocks)
        return achip_pass8(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, workspace, workspace_bytes, stream);
    int r = validate(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type);
    if (r) return r;
    const size_t need = achip_net_workspace_bytes(net, w, h);
    if (!workspace || workspace_bytes < need)
        return fail(ACHIP_EINVAL, "workspace of %zu bytes, a %dx%d pass needs %zu", workspace_bytes, w, h, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 256) return fail(ACHIP_EINVAL, "workspace not 256-byte aligned");
    r = check_device_buffer(workspace, need, net->device, "workspace");
    if (r) return r;
    const char* ws = static_cast<const char*>(workspace);
    const char* s0 = static_cast<const char*>(src);
    const char* d0 = static_cast<const char*>(dst);
    const size_t sbytes = static_cast<size_t>(h - 1) * src_stride + static_cast<size_t>(w) * (src_type & 0xff);
    const size_t dbytes = static_cast<size_t>(2 * h - 1) * dst_stride + static_cast<size_t>(2 * w) * (dst_type & 0xff);
    if ((ws < s0 + sbytes && s0 < ws + need) || (ws < d0 + dbytes && d0 < ws + need))
        return fail(ACHIP_EINVAL, "the workspace overlaps src or dst");
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const hipError_t e =
        net->lfamily == kNetArtCNN
            ? achip::launch_artcnn(net->lf, net->lblocks, src_type, dst_type, src, src_stride, w, h, dst, dst_stride, net->lk,
                                   net->lb, net->lwa, workspace, st)
            : achip::launch_fsrcnnx(net->lf, net->lblocks, src_type, dst_type, src, src_stride, w, h, dst, dst_stride,
                                    net->lk, net->lb, net->la, net->lwa, net->lw1, workspace, st);
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "kernel launch");
}

size_t achip_pass8_workspace_bytes(int w, int h)
{
    if (w <= 0 || h <= 0) return 0;
    return 2 * static_cast<size_t>(achip::acnet8_feature_pitch(w)) * static_cast<size_t>(h);
}

int achip_pass8(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h, void* dst, long dst_stride,
                int dst_type, void* workspace, size_t workspace_bytes, void* stream)
{
    return achip_pass8_precision(net, ACHIP_FP32, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, workspace,
                                 workspace_bytes, stream);
}

int achip_pass8_precision(ACHipNet* net, int precision, const void* src, long src_stride, int src_type, int w, int h,
                          void* dst, long dst_stride, int dst_type, void* workspace, size_t workspace_bytes, void* stream)
{
    if (precision != ACHIP_FP32 && precision != ACHIP_FP16 && precision != ACHIP_BF16)
        return fail(ACHIP_EINVAL, "precision %d not available", precision);
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (!net->blocks) return fail(ACHIP_EINVAL, "achip_pass8 takes an ACNet<8> net");
    int r = validate(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type);
    if (r) return r;
    const size_t need = achip_pass8_workspace_bytes(w, h);
    if (!workspace || workspace_bytes < need)
        return fail(ACHIP_EINVAL, "workspace of %zu bytes, a %dx%d pass needs %zu", workspace_bytes, w, h, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 256) return fail(ACHIP_EINVAL, "workspace not 256-byte aligned");
    r = check_device_buffer(workspace, need, net->device, "workspace");
    if (r) return r;
    const char* ws = static_cast<const char*>(workspace);
    const char* s0 = static_cast<const char*>(src);
    const char* d0 = static_cast<const char*>(dst);
    const size_t sbytes = static_cast<size_t>(h - 1) * src_stride + static_cast<size_t>(w) * (src_type & 0xff);
    const size_t dbytes = static_cast<size_t>(2 * h - 1) * dst_stride + static_cast<size_t>(2 * w) * (dst_type & 0xff);
    if ((ws < s0 + sbytes && s0 < ws + need) || (ws < d0 + dbytes && d0 < ws + need))
        return fail(ACHIP_EINVAL, "the workspace overlaps src or dst");
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const hipError_t e =
        precision == ACHIP_FP32
            ? achip::launch_acnet8(net->blocks, src_type, dst_type, src, src_stride, w, h, dst, dst_stride, net->a8wts,
                                   net->a8wa, workspace, st)
            : achip::launch_acnet8_16(precision == ACHIP_BF16, net->blocks, src_type, dst_type, src, src_stride, w, h, dst,
                                      dst_stride, net->a8wts, precision == ACHIP_BF16 ? net->a8wabf16 : net->a8wa16,
                                      workspace, st);
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "kernel launch");
}

int achip_pass_layers(ACHipNet* net, const void* src, long src_stride, int src_type, int w, int h, void* dst,
                      long dst_stride, int dst_type, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (net->blocks || net->lfamily) return fail(ACHIP_EINVAL, "achip_pass_layers takes an ACNetLegacy net");
    int r = validate(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type);
    if (r) return r;
    const size_t need = achip_pass8_workspace_bytes(w, h);
    if (!workspace || workspace_bytes < need)
        return fail(ACHIP_EINVAL, "workspace of %zu bytes, a %dx%d pass needs %zu", workspace_bytes, w, h, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 256) return fail(ACHIP_EINVAL, "workspace not 256-byte aligned");
    r = check_device_buffer(workspace, need, net->device, "workspace");
    if (r) return r;
    const char* ws = static_cast<const char*>(workspace);
    const char* s0 = static_cast<const char*>(src);
    const char* d0 = static_cast<const char*>(dst);
    const size_t sbytes = static_cast<size_t>(h - 1) * src_stride + static_cast<size_t>(w) * (src_type & 0xff);
    const size_t dbytes = static_cast<size_t>(2 * h - 1) * dst_stride + static_cast<size_t>(2 * w) * (dst_type & 0xff);
    if ((ws < s0 + sbytes && s0 < ws + need) || (ws < d0 + dbytes && d0 < ws + need))
        return fail(ACHIP_EINVAL, "the workspace overlaps src or dst");
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const hipError_t e = achip::launch_legacy_split(src_type, dst_type, src, src_stride, w, h, dst, dst_stride, net->wts,
                                                    net->wa32, workspace, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "kernel launch");
}

int achip_net_device(const ACHipNet* net) { return net ? net->device : ACHIP_EINVAL; }

size_t achip_pass_ws_bytes(ACHipNet* net, int precision, int src_type, int dst_type, int w, int h)
{
    if (!net || w <= 0 || h <= 0 || !elem_size(src_type) || !elem_size(dst_type)) return 0;
    return split_preferred(net, precision, src_type, dst_type, w, h) ? achip_pass8_workspace_bytes(w, h) : 0;
}

int achip_pass_ws(ACHipNet* net, int precision, const void* src, long src_stride, int src_type, int w, int h, void* dst,
                  long dst_stride, int dst_type, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (net->blocks || net->lfamily) return achip_pass(net, precision, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, stream);
    const char* env = std::getenv("ACHIP_AUTOTUNE");
    if (precision == ACHIP_FP32 && env && env[0] == '1' && w > 0 && h > 0) {   // tune an untuned class on first use
        bool found;
        {
            std::lock_guard<std::mutex> lk(g_tune_mtx);
            found = g_split.count(tune_key(net->device, precision, src_type, dst_type, w, h)) != 0;
        }
        int best = 0;
        if (!found) (void)achip_autotune(net, precision, src_type, dst_type, w, h, &best, nullptr);
    }
    const size_t need = achip_pass_ws_bytes(net, precision, src_type, dst_type, w, h);
    if (need && workspace && workspace_bytes >= need)
        return achip_pass_layers(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, workspace, workspace_bytes, stream);
    return achip_pass(net, precision, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, stream);
}

int achip_config_count(int precision)
{
    switch (precision) {
    case ACHIP_FP32: return achip::kMfma32ConfigCount + achip::kFp32ConfigCount;   // FP32 MFMA, then VALU
    case ACHIP_FP16:
    case ACHIP_BF16: return achip::kMfmaConfigCount;
    default: return 0;
    }
}

int achip_config_info(int precision, int i, int* tw, int* th, int* nt, int* p, int* lds_bytes)
{
    if (i < 0 || i >= achip_config_count(precision)) return fail(ACHIP_EINVAL, "no config %d for precision %d", i, precision);
    const achip::Fp32Config& c = precision != ACHIP_FP32        ? achip::kMfmaConfigs[i]
                               : i < achip::kMfma32ConfigCount ? achip::kMfma32Configs[i]
                                                               : achip::kFp32Configs[i - achip::kMfma32ConfigCount];
    if (tw) *tw = c.tw;
    if (th) *th = c.th;
    if (nt) *nt = c.nt;
    if (p) *p = c.p;
    if (lds_bytes) *lds_bytes = c.lds_bytes;
    return ACHIP_OK;
}

int achip_pass_config(ACHipNet* net, int precision, int config,
                      const void* src, long src_stride, int src_type, int w, int h,
                      void* dst, long dst_stride, int dst_type, void* stream)
{
    if (precision != ACHIP_FP32 && precision != ACHIP_FP16 && precision != ACHIP_BF16)
        return fail(ACHIP_EINVAL, "precision %d not available", precision);
    if (config == ACHIP_AUTO && net) {
        bool found = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mtx);
            found = g_tune.count(tune_key(net->device, precision, src_type, dst_type, w, h)) != 0;
        }
        const char* env = std::getenv("ACHIP_AUTOTUNE");
        if (!found && env && env[0] == '1') {
            int best = 0;
            (void)achip_autotune(net, precision, src_type, dst_type, w, h, &best, nullptr);
        }
        config = achip::auto_config(net, precision, src_type, dst_type, w, h);
    }
    if (config < 0 || config >= achip_config_count(precision)) return fail(ACHIP_EINVAL, "no config %d", config);
    if (net && net->blocks) return fail(ACHIP_EINVAL, "an ACNet<8> net runs through achip_pass8");
    if (net && net->lfamily) return fail(ACHIP_EINVAL, "an ArtCNN or FSRCNNX net runs through achip_pass_net");
    int r = validate(net, src, src_stride, src_type, w, h, dst, dst_stride, dst_type);
    if (r) return r;
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const hipError_t e = achip::launch_pass(net, precision, config, src_type, dst_type, src, src_stride, w, h, dst,
                                            dst_stride, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return ACHIP_OK;
}

// ---- device-memory colour and resize stages ----

namespace {
// a plane of w x h pixels of `comps` elements: stride, alignment and allocation on `device`
int check_plane(const void* p, long stride, int w, int h, int comps, int type, int device, const char* name)
{
    const int es = elem_size(type);
    if (!es) return fail(ACHIP_EINVAL, "unsupported element type 0x%x", type);
    if (!p) return fail(ACHIP_EINVAL, "%s is NULL", name);
    if (w <= 0 || h <= 0 || w > (1 << 29) || h > (1 << 29)) return fail(ACHIP_EINVAL, "%s: bad size %dx%d", name, w, h);
    const long row = static_cast<long>(w) * comps * es;
    if (stride < row || stride % es) return fail(ACHIP_EINVAL, "%s: stride %ld invalid for %ld-byte rows", name, stride, row);
    if (reinterpret_cast<uintptr_t>(p) % es) return fail(ACHIP_EINVAL, "%s: pointer not aligned to its element size", name);
    return check_device_buffer(p, static_cast<size_t>(h - 1) * stride + row, device, name);
}
int current_device(int* dev)
{
    const hipError_t e = hipGetDevice(dev);
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "hipGetDevice");
}
} // namespace

int achip_rgb2yuv(int type, int channels, const void* rgb, long rgb_stride, int w, int h, void* y, long y_stride,
                  void* uv, long uv_stride, void* stream)
{
    if (channels != 3 && channels != 4) return fail(ACHIP_EINVAL, "channels must be 3 or 4");
    int dev = 0, r = current_device(&dev);
    if (!r) r = check_plane(rgb, rgb_stride, w, h, channels, type, dev, "rgb");
    if (!r) r = check_plane(y, y_stride, w, h, 1, type, dev, "y");
    if (!r) r = check_plane(uv, uv_stride, w, h, channels - 1, type, dev, "uv");
    if (r) return r;
    const hipError_t e = achip::launch_color_split(type, channels, rgb, rgb_stride, w, h, y, y_stride, uv, uv_stride,
                                                   static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "colour kernel");
}

int achip_yuv2rgb(int type, int channels, const void* y, long y_stride, const void* uv, long uv_stride, int w, int h,
                  void* rgb, long rgb_stride, void* stream)
{
    if (channels != 3 && channels != 4) return fail(ACHIP_EINVAL, "channels must be 3 or 4");
    int dev = 0, r = current_device(&dev);
    if (!r) r = check_plane(y, y_stride, w, h, 1, type, dev, "y");
    if (!r) r = check_plane(uv, uv_stride, w, h, channels - 1, type, dev, "uv");
    if (!r) r = check_plane(rgb, rgb_stride, w, h, channels, type, dev, "rgb");
    if (r) return r;
    const hipError_t e = achip::launch_color_merge(type, channels, y, y_stride, uv, uv_stride, w, h, rgb, rgb_stride,
                                                   static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "colour kernel");
}

int achip_resize(int type, int channels, int mode, const void* src, long src_stride, int sw, int sh, void* dst,
                 long dst_stride, int dw, int dh, void* stream)
{
    if (channels < 1 || channels > 4) return fail(ACHIP_EINVAL, "channels must be 1 to 4");
    if (mode < 0 || mode >= achip::kResizeModeCount) return fail(ACHIP_EINVAL, "resize mode %d unknown", mode);
    int dev = 0, r = current_device(&dev);
    if (!r) r = check_plane(src, src_stride, sw, sh, channels, type, dev, "src");
    if (!r) r = check_plane(dst, dst_stride, dw, dh, channels, type, dev, "dst");
    if (r) return r;
    if (std::max(achip::resize_max_taps(mode, sw, dw), achip::resize_max_taps(mode, sh, dh)) > achip::kResizeMaxTaps)
        return fail(ACHIP_EINVAL, "downscale %dx%d -> %dx%d needs more than %d filter taps", sw, sh, dw, dh, achip::kResizeMaxTaps);
    const hipError_t e = achip::launch_resize(type, channels, mode, src, src_stride, sw, sh, dst, dst_stride, dw, dh,
                                              static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "resize kernel");
}

int achip_resize_yuv2rgb(int type, int channels, int mode, const void* y, long y_stride, const void* uv, long uv_stride,
                         int uv_w, int uv_h, int w, int h, void* rgb, long rgb_stride, void* stream)
{
    if (channels != 3 && channels != 4) return fail(ACHIP_EINVAL, "channels must be 3 or 4");
    if (mode < 0 || mode >= achip::kResizeModeCount) return fail(ACHIP_EINVAL, "resize mode %d unknown", mode);
    int dev = 0, r = current_device(&dev);
    if (!r) r = check_plane(y, y_stride, w, h, 1, type, dev, "y");
    if (!r) r = check_plane(uv, uv_stride, uv_w, uv_h, channels - 1, type, dev, "uv");
    if (!r) r = check_plane(rgb, rgb_stride, w, h, channels, type, dev, "rgb");
    if (r) return r;
    if (std::max(achip::resize_max_taps(mode, uv_w, w), achip::resize_max_taps(mode, uv_h, h)) > achip::kResizeMaxTaps)
        return fail(ACHIP_EINVAL, "chroma downscale needs more than %d filter taps", achip::kResizeMaxTaps);
    const hipError_t e = achip::launch_resize_merge(type, channels, mode, y, y_stride, uv, uv_stride, uv_w, uv_h, w, h,
                                                    rgb, rgb_stride, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "resize kernel");
}

int achip_pass(ACHipNet* net, int precision, const void* src, long src_stride, int src_type, int w, int h,
               void* dst, long dst_stride, int dst_type, void* stream)
{
    return achip_pass_config(net, precision, ACHIP_AUTO, src, src_stride, src_type, w, h, dst, dst_stride, dst_type, stream);
}

int achip_autotune(ACHipNet* net, int precision, int src_type, int dst_type, int w, int h, int* best, float* ms)
{
    if (!net || !best) return fail(ACHIP_EINVAL, "NULL argument");
    const int n = achip_config_count(precision);
    if (n <= 0) return fail(ACHIP_EINVAL, "precision %d not available", precision);
    const int ies = elem_size(src_type), oes = elem_size(dst_type);
    if (!ies || !oes || w <= 0 || h <= 0 || w > (1 << 15) || h > (1 << 15)) return fail(ACHIP_EINVAL, "bad tuning shape");
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const long sp = static_cast<long>(w) * ies, dp = 2L * w * oes;
    void* s = nullptr;
    void* d = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ACHIP_OK;
    auto cleanup = [&]() {
        if (st) (void)hipStreamSynchronize(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        if (s) (void)hipFree(s);
        if (d) (void)hipFree(d);
    };
    hipError_t e = hipMalloc(&s, static_cast<size_t>(sp) * h);
    if (e == hipSuccess) e = hipMalloc(&d, static_cast<size_t>(dp) * 2 * h);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(s, 0x40, static_cast<size_t>(sp) * h, st);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "autotune setup");
    }
    // Two rounds over the configurations, each one's faster round: in a single round the
    // configuration timed first read 5-15% slow.
    std::vector<float> cfg_ms(static_cast<size_t>(n), 1e30f);
    for (int round = 0; round < 2 && rc == ACHIP_OK; round++)
        for (int c = 0; c < n && rc == ACHIP_OK; c++) {
            rc = achip_pass_config(net, precision, c, s, sp, src_type, w, h, d, dp, dst_type, st);   // warm-up
            if (rc) break;
            const int reps = 3;
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps && rc == ACHIP_OK; r++)
                rc = achip_pass_config(net, precision, c, s, sp, src_type, w, h, d, dp, dst_type, st);
            (void)hipEventRecord(e1, st);
            e = hipEventSynchronize(e1);
            if (e != hipSuccess) {
                cleanup();
                return hip_fail(e, "autotune run");
            }
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            cfg_ms[static_cast<size_t>(c)] = std::min(cfg_ms[static_cast<size_t>(c)], t / reps);
        }
    float best_ms = 1e30f;
    int best_cfg = 0;
    for (int c = 0; c < n && rc == ACHIP_OK; c++) {
        if (ms) ms[c] = cfg_ms[static_cast<size_t>(c)];
        if (cfg_ms[static_cast<size_t>(c)] < best_ms) { best_ms = cfg_ms[static_cast<size_t>(c)]; best_cfg = c; }
    }
    // FP32 ACNetLegacy: the layer-split form too (same bits), for achip_pass_ws
    int split = -1;
    if (rc == ACHIP_OK && precision == ACHIP_FP32 && !net->blocks && !net->lfamily) {
        const size_t wsb = achip_pass8_workspace_bytes(w, h);
        void* ws = nullptr;
        e = hipMalloc(&ws, wsb);
        if (e != hipSuccess) {
            (void)hipGetLastError();   // no room for the maps: the fused kernel it is
            split = 0;
        } else {
            rc = achip_pass_layers(net, s, sp, src_type, w, h, d, dp, dst_type, ws, wsb, st);
            const int reps = 3;
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps && rc == ACHIP_OK; r++)
                rc = achip_pass_layers(net, s, sp, src_type, w, h, d, dp, dst_type, ws, wsb, st);
            (void)hipEventRecord(e1, st);
            e = hipEventSynchronize(e1);
            float t = 0.0f;
            if (e == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
            (void)hipFree(ws);
            if (e != hipSuccess) {
                cleanup();
                return hip_fail(e, "autotune run");
            }
            split = t / reps < best_ms ? 1 : 0;
        }
    }
    cleanup();
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        g_tune[tune_key(net->device, precision, src_type, dst_type, w, h)] = best_cfg;
        if (split >= 0) g_split[tune_key(net->device, precision, src_type, dst_type, w, h)] = split == 1;
    }
    *best = best_cfg;
    return ACHIP_OK;
}

size_t achip_batch_workspace_bytes(int n)
{
    return n <= 0 ? kBatchHeader : kBatchHeader + static_cast<size_t>(n) * sizeof(achip::BatchDesc);
}

int achip_pass_batch(ACHipNet* net, int precision, int config, int src_type, int dst_type,
                     const ACHipImage* images, int n, void* workspace, size_t workspace_bytes, void* stream)
{
    if (n <= 0 || n > (1 << 24)) return fail(ACHIP_EINVAL, "batch of %d images", n);
    if (workspace) {
        const int r = check_device_buffer(workspace, std::min(workspace_bytes, achip_batch_workspace_bytes(n)),
                                          net ? net->device : 0, "workspace");
        if (r) return r;
    }
    std::vector<char> host(achip_batch_workspace_bytes(n));
    std::vector<achip::BatchSpan> spans(2 * static_cast<size_t>(n) + 1);
    int total = 0;
    const int r = achip::batch_prepare(net, precision, config, src_type, dst_type, images, n, workspace, workspace_bytes,
                                       host.data(), spans.data(), &total);
    if (r) return r;
    return achip::batch_launch(net, precision, config, src_type, dst_type, n, total, workspace, host.data(),
                               static_cast<hipStream_t>(stream));
}

namespace {

// workspace of achip_pass_batch_layers: the descriptors, then each image's two maps
size_t split_batch_layout(const ACHipImage* images, int n, std::vector<size_t>* offsets)
{
    const size_t header = (static_cast<size_t>(n) * sizeof(achip::SplitDesc) + 255) / 256 * 256;
    size_t off = header;
    for (int i = 0; i < n; i++) {
        if (offsets) offsets->push_back(off);
        const ACHipImage& im = images[i];
        if (im.w > 0 && im.h > 0) off += (achip_pass8_workspace_bytes(im.w, im.h) + 255) / 256 * 256;
    }
    return off;
}

} // namespace

size_t achip_pass_batch_layers_bytes(const ACHipImage* images, int n)
{
    if (!images || n <= 0 || n > (1 << 24)) return 0;
    return split_batch_layout(images, n, nullptr);
}

int achip_pass_batch_layers(ACHipNet* net, int src_type, int dst_type, const ACHipImage* images, int n, void* workspace,
                            size_t workspace_bytes, void* stream)
{
    if (!net) return fail(ACHIP_EINVAL, "net is NULL");
    if (net->blocks || net->lfamily) return fail(ACHIP_EINVAL, "achip_pass_batch_layers takes an ACNetLegacy net");
    if (!images || n <= 0 || n > (1 << 24)) return fail(ACHIP_EINVAL, "batch of %d images", n);
    std::vector<size_t> offsets;
    const size_t need = split_batch_layout(images, n, &offsets);
    if (!workspace || workspace_bytes < need)
        return fail(ACHIP_EINVAL, "workspace of %zu bytes, the batch needs %zu", workspace_bytes, need);
    if (reinterpret_cast<uintptr_t>(workspace) % 256) return fail(ACHIP_EINVAL, "workspace not 256-byte aligned");
    int r = check_device_buffer(workspace, need, net->device, "workspace");
    if (r) return r;
    std::vector<achip::SplitDesc> desc(static_cast<size_t>(n));
    std::vector<achip::BatchSpan> spans;
    spans.reserve(2 * static_cast<size_t>(n) + 1);
    char* ws = static_cast<char*>(workspace);
    long total = 0;
    for (int i = 0; i < n; i++) {
        const ACHipImage& im = images[i];
        r = validate(net, im.src, im.src_stride, src_type, im.w, im.h, im.dst, im.dst_stride, dst_type);
        if (r) {
            std::string msg = g_error;
            return fail(r, "image %d: %s", i, msg.c_str());
        }
        const uintptr_t s0 = reinterpret_cast<uintptr_t>(im.src), d0 = reinterpret_cast<uintptr_t>(im.dst);
        spans.push_back({ s0, s0 + static_cast<size_t>(im.h - 1) * im.src_stride + static_cast<size_t>(im.w) * (src_type & 0xff), false, i });
        spans.push_back({ d0, d0 + static_cast<size_t>(2 * im.h - 1) * im.dst_stride + static_cast<size_t>(2 * im.w) * (dst_type & 0xff), true, i });
        const long fp = achip::acnet8_feature_pitch(im.w);
        auto* m0 = reinterpret_cast<float4*>(ws + offsets[i]);
        auto* m1 = reinterpret_cast<float4*>(ws + offsets[i] + static_cast<size_t>(fp) * im.h);
        const int tx = (im.w + 61) / 62, ty = (im.h + 15) / 16;
        desc[i] = achip::SplitDesc{ im.src, im.dst, m0, m1, im.src_stride, im.dst_stride, fp, im.w, im.h, tx, static_cast<int>(total) };
        total += static_cast<long>(tx) * ty;
        if (total > (1L << 30)) return fail(ACHIP_EINVAL, "batch too large (%ld tiles)", total);
    }
    // the maps and descriptors are written by the launches: no image may overlap them, and no
    // image's dst another image's src or dst
    spans.push_back({ reinterpret_cast<uintptr_t>(workspace), reinterpret_cast<uintptr_t>(workspace) + need, true, -1 });
    r = achip::check_disjoint_writes(spans.data(), static_cast<int>(spans.size()));
    if (r) return r;
    DeviceGuard g(net->device);
    if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // from pageable memory the copy is staged before the call returns, so `desc` may go out
    // of scope
    hipError_t e = hipMemcpyAsync(workspace, desc.data(), static_cast<size_t>(n) * sizeof(achip::SplitDesc),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = achip::launch_legacy_split_batch(src_type, dst_type, reinterpret_cast<const achip::SplitDesc*>(workspace), n,
                                             static_cast<int>(total), net->wts, net->wa32, st);
    return e == hipSuccess ? ACHIP_OK : hip_fail(e, "kernel launch");
}

const char* achip_last_error(void) { return g_error.c_str(); }

} // extern "C"
