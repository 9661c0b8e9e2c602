// The HIP processor: ACNetLegacy 2x passes on an MI355X through the thin C ABI
// (include/achip.h).  It takes the place of the reference's GPU processors
// (core/src/processor/cuda/CUDAProcessor.cpp:363-428) behind ac::core::Processor:
//   - weights are uploaded once, in the constructor;
//   - every calling thread gets its own HIP stream (default priority), device buffers
//     and error state, so one processor can serve many threads at once, as the
//     reference's callers (CLI pools, VapourSynth fmParallel, the video filter) use it;
//   - the whole of process(src, dst, factor) runs on the device (processAll): colour
//     split, the chained 2x passes (factor 4, 8, ...), the resize of Y by the
//     non-power-of-two remainder and of UV by the full factor, and the merge.  Only the
//     input and the final output cross PCIe, pageable caller memory through two
//     page-locked chunks of the thread's own.  Intermediate images keep the input's
//     element type, as the reference's pass loop does (Processor.cpp:219-236).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>

#include "AC/Core/Processor.hpp"
#include "achip.h"
#include "acnet_kernels.hpp"
#include "core_internal.hpp"
#include "host_staging.hpp"

namespace ac::core::hip
{
    struct DeviceBuffer
    {
        void* ptr = nullptr;
        size_t bytes = 0;
    };

    class HIPProcessorACNetLegacy final : public Processor
    {
    public:
        HIPProcessorACNetLegacy(int device, const model::ACNetLegacy& model, int precision) noexcept;
        // ACNet<8>, ArtCNN and FSRCNNX: the same processor around an achip_net8_create,
        // achip_net_artcnn_create or achip_net_fsrcnnx_create net (FP32 only)
        HIPProcessorACNetLegacy(int device, const model::ACNet<8>& model, int precision) noexcept;
        template<int F>
        HIPProcessorACNetLegacy(int device, const model::ArtCNN<F>& model, int precision) noexcept;
        template<int F>
        HIPProcessorACNetLegacy(int device, const model::FSRCNNX<F>& model, int precision) noexcept;
        ~HIPProcessorACNetLegacy() noexcept override;

        bool ok() noexcept override;
        const char* error() noexcept override;
        const char* name() const noexcept override { return devName.c_str(); }
        int type() const noexcept override { return Processor::HIP; }
        const char* typeName() const noexcept override { return "HIP"; }

    private:
        void process(const Image& src, Image& dst) override;

    protected:
        bool processAll(const Image& src, Image& dst, double factor) override;
        bool processChain(const Image& src, Image& dst, int passes) override;
        void setError(const char* message) noexcept override;

    private:
        struct ThreadState
        {
            hipStream_t stream = nullptr;
            DeviceBuffer buf[8];
            DeviceBuffer ws;          // feature maps: ACNet<8>, ArtCNN, FSRCNNX, ACNetLegacy's layer-split form
            detail::HostStaging staging;   // page-locked chunks for pageable caller images
            std::string error;
        };
        ThreadState& local();
        bool fail(const std::string& what);
        bool ensure(DeviceBuffer& b, size_t bytes);
        bool run(const Image& src, Image& dst, int passes);
        bool pipeline(const Image& src, Image& dst, double factor);
        bool stream_ready(ThreadState& st);
        // caller image rows <-> device rows on this thread's stream (through st.staging when
        // the host memory is pageable); toHost returns once the rows are in host memory
        bool toDevice(ThreadState& st, void* d, size_t dpitch, const void* h, size_t hpitch, size_t row, int rows);
        bool toHost(ThreadState& st, void* h, size_t hpitch, const void* d, size_t dpitch, size_t row, int rows);
        // one 2x pass of the net on device buffers (ACNetLegacy: achip_pass_ws, lending the
        // workspace to the layer-split form; ACNet<8>, ArtCNN and FSRCNNX: achip_pass8_precision
        // or achip_pass_net with this thread's workspace)
        bool pass(ThreadState& st, const void* src, long sp, int stype, int w, int h, void* dst, long dp, int dtype);
        void finishInit();

    private:
        ACHipNet* net = nullptr;
        int precision;
        bool layered = false;     // ACNet<8> or ArtCNN: feature maps in a workspace
        bool acnet8 = false;      // ACNet<8>: achip_pass8_precision in this processor's arithmetic
        std::string devName{ "no HIP device" };
        std::string initError;
        std::shared_mutex mtx;
        std::unordered_map<std::thread::id, std::unique_ptr<ThreadState>> states;
    };

    struct CurrentDevice
    {
        int prev = -1;
        explicit CurrentDevice(int dev) { if (hipGetDevice(&prev) != hipSuccess) prev = -1; if (prev != dev) (void)hipSetDevice(dev); }
        ~CurrentDevice() { int cur = -1; if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev); }
    };

    inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

    const char* precisionName(const int precision) noexcept
    {
        return precision == 1 ? "FP16" : (precision == 2 ? "BF16" : "FP32");
    }
}

ac::core::hip::HIPProcessorACNetLegacy::HIPProcessorACNetLegacy(const int device, const model::ACNetLegacy& model, const int precision) noexcept
    : precision(precision)
{
    const int rc = achip_net_create(device, model.kernel(0), model.bias(0), &net);
    if (rc != ACHIP_OK)
    {
        initError = std::string("HIP processor: ") + achip_last_error();
        net = nullptr;
        return;
    }
    finishInit();
    if (achip_config_count(precision) <= 0)
        initError = "HIP processor: precision " + std::to_string(precision) + " is not available in this build";
}

ac::core::hip::HIPProcessorACNetLegacy::HIPProcessorACNetLegacy(const int device, const model::ACNet<8>& model, const int precision) noexcept
    : precision(precision), layered(true), acnet8(true)
{
    const int rc = achip_net8_create(device, model.blocks(), model.kernel(0), model.bias(0), model.alpha(0), &net);
    if (rc != ACHIP_OK)
    {
        initError = std::string("HIP processor: ") + achip_last_error();
        net = nullptr;
        return;
    }
    finishInit();
}

template<int F>
ac::core::hip::HIPProcessorACNetLegacy::HIPProcessorACNetLegacy(const int device, const model::ArtCNN<F>& model, const int precision) noexcept
    : precision(precision), layered(true)
{
    if (precision != 0)
    {
        initError = std::string("HIP processor: ArtCNN runs in FP32 only (the FP16/BF16 forms are ACNetLegacy's); the type string asked for ") + precisionName(precision);
        return;
    }
    const int rc = achip_net_artcnn_create(device, F, model.blocks(), model.kernel(0), model.bias(0), &net);
    if (rc != ACHIP_OK)
    {
        initError = std::string("HIP processor: ") + achip_last_error();
        net = nullptr;
        return;
    }
    finishInit();
}

template<int F>
ac::core::hip::HIPProcessorACNetLegacy::HIPProcessorACNetLegacy(const int device, const model::FSRCNNX<F>& model, const int precision) noexcept
    : precision(precision), layered(true)
{
    if (precision != 0)
    {
        initError = std::string("HIP processor: FSRCNNX runs in FP32 only (the FP16/BF16 forms are ACNetLegacy's); the type string asked for ") + precisionName(precision);
        return;
    }
    const int rc = achip_net_fsrcnnx_create(device, F, model.blocks(), model.kernel(0), model.bias(0), model.alpha(0), &net);
    if (rc != ACHIP_OK)
    {
        initError = std::string("HIP processor: ") + achip_last_error();
        net = nullptr;
        return;
    }
    finishInit();
}

void ac::core::hip::HIPProcessorACNetLegacy::finishInit()
{
    idx = achip_net_device(net);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, idx) == hipSuccess) devName = prop.name;
}

bool ac::core::hip::HIPProcessorACNetLegacy::pass(ThreadState& st, const void* src, const long sp, const int stype, const int w,
                                                  const int h, void* dst, const long dp, const int dtype)
{
    int rc;
    if (layered)
    {
        if (!ensure(st.ws, achip_net_workspace_bytes(net, w, h))) return false;
        rc = acnet8 ? achip_pass8_precision(net, precision, src, sp, stype, w, h, dst, dp, dtype, st.ws.ptr, st.ws.bytes, st.stream)
                    : achip_pass_net(net, src, sp, stype, w, h, dst, dp, dtype, st.ws.ptr, st.ws.bytes, st.stream);
    }
    else
    {
        // the faster form for the size; the layer-split one borrows this thread's map buffer
        // (same bits as the fused kernel, which runs when that buffer cannot be had)
        const size_t need = achip_pass_ws_bytes(net, precision, stype, dtype, w, h);
        if (need && st.ws.bytes < need)
        {
            if (st.ws.ptr) { (void)hipFree(st.ws.ptr); st.ws = DeviceBuffer{}; }
            if (hipMalloc(&st.ws.ptr, need) == hipSuccess) st.ws.bytes = need;
            else { st.ws.ptr = nullptr; (void)hipGetLastError(); }
        }
        rc = achip_pass_ws(net, precision, src, sp, stype, w, h, dst, dp, dtype, st.ws.ptr, st.ws.bytes, st.stream);
    }
    return rc == ACHIP_OK ? true : fail(std::string(layered ? "achip_pass_net: " : "achip_pass_ws: ") + achip_last_error());
}

ac::core::hip::HIPProcessorACNetLegacy::~HIPProcessorACNetLegacy() noexcept
{
    if (net)
    {
        CurrentDevice g(idx);
        for (auto& kv : states)
        {
            auto& st = *kv.second;
            if (st.stream) (void)hipStreamSynchronize(st.stream);
            for (auto& b : st.buf) if (b.ptr) (void)hipFree(b.ptr);
            if (st.ws.ptr) (void)hipFree(st.ws.ptr);
            st.staging.release();
            if (st.stream) (void)hipStreamDestroy(st.stream);
        }
    }
    achip_net_destroy(net);
}

ac::core::hip::HIPProcessorACNetLegacy::ThreadState& ac::core::hip::HIPProcessorACNetLegacy::local()
{
    const auto id = std::this_thread::get_id();
    {
        std::shared_lock<std::shared_mutex> lk(mtx);
        auto it = states.find(id);
        if (it != states.end()) return *it->second;
    }
    std::unique_lock<std::shared_mutex> lk(mtx);
    auto& slot = states[id];
    if (!slot) slot = std::make_unique<ThreadState>();
    return *slot;
}

bool ac::core::hip::HIPProcessorACNetLegacy::ok() noexcept
{
    return initError.empty() && local().error.empty();
}

const char* ac::core::hip::HIPProcessorACNetLegacy::error() noexcept
{
    if (!initError.empty()) return initError.c_str();
    auto& e = local().error;
    return e.empty() ? "NO ERROR" : e.c_str();
}

void ac::core::hip::HIPProcessorACNetLegacy::setError(const char* message) noexcept
{
    local().error = message ? message : "unknown error";
}

bool ac::core::hip::HIPProcessorACNetLegacy::fail(const std::string& what)
{
    local().error = what;
    return false;
}

bool ac::core::hip::HIPProcessorACNetLegacy::ensure(DeviceBuffer& b, const size_t bytes)
{
    if (b.bytes >= bytes) return true;
    if (b.ptr) { (void)hipFree(b.ptr); b.ptr = nullptr; b.bytes = 0; }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) { b.ptr = nullptr; return fail(std::string("hipMalloc: ") + hipGetErrorString(e)); }
    b.bytes = bytes;
    return true;
}

// Upload src, run `passes` 2x passes in device memory, download into dst.
bool ac::core::hip::HIPProcessorACNetLegacy::run(const Image& src, Image& dst, const int passes)
{
    if (!initError.empty()) return false;
    auto& st = local();
    st.error.clear();
    if (src.empty() || src.channels() != 1) return fail("the 2x pass takes a non-empty 1-channel image");
    CurrentDevice g(idx);
    if (!st.stream)
    {
        hipError_t e = hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking);
        if (e != hipSuccess) { st.stream = nullptr; return fail(std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    }
    const int es = src.elementSize();
    int w = src.width(), h = src.height();
    const long fw = static_cast<long>(w) << passes, fh = static_cast<long>(h) << passes;
    if (fw > (1L << 30) || fh > (1L << 30)) return fail("output too large");
    if (dst.empty()) dst.create(static_cast<int>(fw), static_cast<int>(fh), 1, src.type());
    if (dst.width() != fw || dst.height() != fh || dst.channels() != 1)
        return fail("destination has the wrong size for this factor");
    if (dst.empty()) return fail("cannot allocate the destination image");
    const int des = dst.elementSize();

    // buffers: [0] input, then ping-pong [1]/[2]; rows padded to 256 bytes
    size_t pitch = alignUp(static_cast<size_t>(w) * es, 256);
    if (!ensure(st.buf[0], pitch * h)) return false;
    if (!toDevice(st, st.buf[0].ptr, pitch, src.ptr(), src.stride(), static_cast<size_t>(w) * es, h)) return false;
    void* cur = st.buf[0].ptr;
    size_t curPitch = pitch;
    for (int i = 0; i < passes; i++)
    {
        const bool last = (i == passes - 1);
        const int ots = last ? dst.type() : src.type();
        const size_t opitch = alignUp(static_cast<size_t>(2 * w) * (last ? des : es), 256);
        DeviceBuffer& ob = st.buf[1 + (i & 1)];
        if (!ensure(ob, opitch * 2 * h)) return false;
        if (!pass(st, cur, static_cast<long>(curPitch), src.type(), w, h, ob.ptr, static_cast<long>(opitch), ots)) return false;
        cur = ob.ptr;
        curPitch = opitch;
        w *= 2;
        h *= 2;
    }
    return toHost(st, dst.ptr(), dst.stride(), cur, curPitch, static_cast<size_t>(w) * des, h);
}

void ac::core::hip::HIPProcessorACNetLegacy::process(const Image& src, Image& dst)
{
    run(src, dst, 1);
}

bool ac::core::hip::HIPProcessorACNetLegacy::processChain(const Image& src, Image& dst, const int passes)
{
    return run(src, dst, passes);
}

bool ac::core::hip::HIPProcessorACNetLegacy::processAll(const Image& src, Image& dst, const double factor)
{
    pipeline(src, dst, factor);
    return true;
}

bool ac::core::hip::HIPProcessorACNetLegacy::stream_ready(ThreadState& st)
{
    if (st.stream) return true;
    hipError_t e = hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking);
    if (e != hipSuccess) { st.stream = nullptr; return fail(std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    return true;
}

bool ac::core::hip::HIPProcessorACNetLegacy::toDevice(ThreadState& st, void* d, const size_t dpitch, const void* h,
                                                      const size_t hpitch, const size_t row, const int rows)
{
    return st.staging.toDevice(d, dpitch, h, hpitch, row, rows, st.stream) ? true : fail(st.staging.error());
}

bool ac::core::hip::HIPProcessorACNetLegacy::toHost(ThreadState& st, void* h, const size_t hpitch, const void* d,
                                                    const size_t dpitch, const size_t row, const int rows)
{
    return st.staging.toHost(h, hpitch, d, dpitch, row, rows, st.stream) ? true : fail(st.staging.error());
}

// Processor.cpp:199-276 on the device.  Buffers: [0] input, [1] Y, [2] UV, [3]/[4] pass
// ping-pong, [5] resized Y, [6] resized UV, [7] merged output.
bool ac::core::hip::HIPProcessorACNetLegacy::pipeline(const Image& src, Image& dst, const double factor)
{
    if (!initError.empty()) return false;
    auto& st = local();
    st.error.clear();
    const int C = src.channels();
    if (src.empty()) return true;
    if (C != 1 && C != 3 && C != 4) return fail("input must have 1, 3 or 4 channels");
    if (!(factor > 0.0) || factor > 256.0) return fail("factor must be in (0, 256]");
    const int power = factor > 2.0 ? detail::ceilLog2(factor) : 1;
    const double fxy = factor / static_cast<double>(1 << power);
    const int T = src.type(), es = src.elementSize();
    const long w = src.width(), h = src.height();
    const long pw = w << power, ph = h << power;
    long fw, fh;
    if (!dst.empty() && C == 1 && fxy != 1.0) { fw = dst.width(); fh = dst.height(); }     // resize(out, dst, 0, 0)
    else if (fxy == 1.0) { fw = pw; fh = ph; }
    else { fw = static_cast<long>(static_cast<double>(pw) * fxy); fh = static_cast<long>(static_cast<double>(ph) * fxy); }
    if (pw > (1L << 30) || ph > (1L << 30) || fw <= 0 || fh <= 0 || fw > (1L << 30) || fh > (1L << 30))
        return fail("output size out of range");
    if (!dst.empty()) {
        if (dst.width() != fw || dst.height() != fh || dst.channels() != C)
            return fail("the preallocated destination has the wrong size or channel count");
        if ((C > 1 || fxy != 1.0) && dst.type() != T)
            return fail("a preallocated destination of another element type needs a 1-channel input and a power-of-two factor");
    }
    const int outT = dst.empty() ? T : dst.type();
    CurrentDevice g(idx);
    if (!stream_ready(st)) return false;
    const hipStream_t s = st.stream;
    auto pitch_of = [](long width, int bytes) { return static_cast<long>(alignUp(static_cast<size_t>(width) * bytes, 256)); };
    auto hipok = [&](hipError_t e, const char* what) { return e == hipSuccess ? true : fail(std::string(what) + ": " + hipGetErrorString(e)); };

    // upload
    const long inPitch = pitch_of(w * C, es);
    if (!ensure(st.buf[0], inPitch * h)) return false;
    if (!toDevice(st, st.buf[0].ptr, inPitch, src.ptr(), src.stride(), w * C * es, h)) return false;
    // colour split
    void* y = st.buf[0].ptr;
    long yPitch = inPitch;
    const long uvPitch = pitch_of(w * (C - 1), es);
    if (C > 1) {
        yPitch = pitch_of(w, es);
        if (!ensure(st.buf[1], yPitch * h) || !ensure(st.buf[2], uvPitch * h)) return false;
        if (!hipok(achip::launch_color_split(T, C, st.buf[0].ptr, inPitch, static_cast<int>(w), static_cast<int>(h),
                                             st.buf[1].ptr, yPitch, st.buf[2].ptr, uvPitch, s), "colour split")) return false;
        y = st.buf[1].ptr;
    }
    // 2x passes on Y; the last writes the output type when nothing follows it
    long cw = w, ch = h;
    for (int i = 0; i < power; i++) {
        const bool last = i == power - 1;
        const int ot = (last && C == 1 && fxy == 1.0) ? outT : T;
        const long op = pitch_of(2 * cw, ot & 0xff);
        DeviceBuffer& ob = st.buf[3 + (i & 1)];
        if (!ensure(ob, op * 2 * ch)) return false;
        if (!pass(st, y, yPitch, T, static_cast<int>(cw), static_cast<int>(ch), ob.ptr, op, ot)) return false;
        y = ob.ptr;
        yPitch = op;
        cw *= 2;
        ch *= 2;
    }
    int yT = (C == 1 && fxy == 1.0) ? outT : T;
    // the remainder of a non-power-of-two factor (Catmull-Rom, as ImageResize.cpp defaults to)
    if (fxy != 1.0) {
        const long rp = pitch_of(fw, es);
        if (!ensure(st.buf[5], rp * fh)) return false;
        if (!hipok(achip::launch_resize(T, 1, achip::kResizeCatmullRom, y, yPitch, static_cast<int>(cw), static_cast<int>(ch), st.buf[5].ptr, rp,
                                        static_cast<int>(fw), static_cast<int>(fh), s), "resize")) return false;
        y = st.buf[5].ptr;
        yPitch = rp;
    }
    void* out = y;
    long outPitch = yPitch;
    if (C > 1) {
        // UV by the full factor and the merge, in one pass (the resized UV is rounded to the
        // element type as the reference's intermediate image holds it)
        const long mp = pitch_of(fw * C, es);
        if (!ensure(st.buf[7], mp * fh)) return false;
        if (!hipok(achip::launch_resize_merge(T, C, achip::kResizeCatmullRom, y, yPitch, st.buf[2].ptr, uvPitch, static_cast<int>(w), static_cast<int>(h),
                                              static_cast<int>(fw), static_cast<int>(fh), st.buf[7].ptr, mp, s),
                   "resize and merge")) return false;
        out = st.buf[7].ptr;
        outPitch = mp;
        yT = T;
    }
    // download
    if (dst.empty()) dst.create(static_cast<int>(fw), static_cast<int>(fh), C, yT);
    if (dst.empty()) return fail("cannot allocate the destination image");
    return toHost(st, dst.ptr(), dst.stride(), out, outPitch, fw * C * dst.elementSize(), fh);
}

std::shared_ptr<ac::core::Processor> ac::core::detail::createHIP(const int device, const model::ACNetLegacy& model, const int precision)
{
    return std::make_shared<hip::HIPProcessorACNetLegacy>(device, model, precision);
}

std::shared_ptr<ac::core::Processor> ac::core::detail::createHIP8(const int device, const model::ACNet<8>& model, const int precision)
{
    return std::make_shared<hip::HIPProcessorACNetLegacy>(device, model, precision);
}

template<int F>
std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPArt(const int device, const model::ArtCNN<F>& model, const int precision)
{
    return std::make_shared<hip::HIPProcessorACNetLegacy>(device, model, precision);
}
template std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPArt<16>(int, const model::ArtCNN<16>&, int);
template std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPArt<32>(int, const model::ArtCNN<32>&, int);

template<int F>
std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPFsr(const int device, const model::FSRCNNX<F>& model, const int precision)
{
    return std::make_shared<hip::HIPProcessorACNetLegacy>(device, model, precision);
}
template std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPFsr<8>(int, const model::FSRCNNX<8>&, int);
template std::shared_ptr<ac::core::Processor> ac::core::detail::createHIPFsr<16>(int, const model::FSRCNNX<16>&, int);

template<>
AC_CORE_EXPORT const char* ac::core::Processor::info<ac::core::Processor::HIP>()
{
    static const std::string text = []() {
        std::ostringstream s;
        s << "HIP:\n";
        int n = achip_device_count();
        for (int i = 0; i < n; i++)
        {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
            s << "  [" << i << "] " << prop.name << " (" << prop.gcnArchName << ", " << prop.multiProcessorCount
              << " CUs, " << (prop.totalGlobalMem >> 30) << " GiB)\n";
        }
        if (n == 0) s << "  no HIP device\n";
        return s.str();
    }();
    return text.c_str();
}
