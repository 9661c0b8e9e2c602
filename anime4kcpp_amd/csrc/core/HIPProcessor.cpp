// The HIP processor: ACNetLegacy 2x passes on an MI355X through the thin C ABI
// (include/achip.h).  It takes the place of the reference's GPU processors
// (core/src/processor/cuda/CUDAProcessor.cpp:363-428) behind ac::core::Processor:
//   - weights are uploaded once, in the constructor;
//   - every calling thread gets its own HIP stream (default priority), device buffers
//     and error state, so one processor can serve many threads at once, as the
//     reference's callers (CLI pools, VapourSynth fmParallel, the video filter) use it;
//   - chained passes (factor 4, 8, ...) stay in device memory between passes; only the
//     input and the final output cross PCIe.  Intermediate images keep the input's
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
#include "core_internal.hpp"

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
        ~HIPProcessorACNetLegacy() noexcept override;

        bool ok() noexcept override;
        const char* error() noexcept override;
        const char* name() const noexcept override { return devName.c_str(); }
        int type() const noexcept override { return Processor::HIP; }
        const char* typeName() const noexcept override { return "HIP"; }

    private:
        void process(const Image& src, Image& dst) override;

    protected:
        bool processChain(const Image& src, Image& dst, int passes) override;
        void setError(const char* message) noexcept override;

    private:
        struct ThreadState
        {
            hipStream_t stream = nullptr;
            DeviceBuffer buf[3];
            std::string error;
        };
        ThreadState& local();
        bool fail(const std::string& what);
        bool ensure(DeviceBuffer& b, size_t bytes);
        bool run(const Image& src, Image& dst, int passes);

    private:
        ACHipNet* net = nullptr;
        int precision;
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
    idx = achip_net_device(net);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, idx) == hipSuccess) devName = prop.name;
    if (achip_config_count(precision) <= 0)
        initError = "HIP processor: precision " + std::to_string(precision) + " is not available in this build";
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
    hipError_t e = hipMemcpy2DAsync(st.buf[0].ptr, pitch, src.ptr(), src.stride(), static_cast<size_t>(w) * es, h,
                                    hipMemcpyHostToDevice, st.stream);
    if (e != hipSuccess) return fail(std::string("hipMemcpy2DAsync(H2D): ") + hipGetErrorString(e));
    void* cur = st.buf[0].ptr;
    size_t curPitch = pitch;
    for (int i = 0; i < passes; i++)
    {
        const bool last = (i == passes - 1);
        const int ots = last ? dst.type() : src.type();
        const size_t opitch = alignUp(static_cast<size_t>(2 * w) * (last ? des : es), 256);
        DeviceBuffer& ob = st.buf[1 + (i & 1)];
        if (!ensure(ob, opitch * 2 * h)) return false;
        const int rc = achip_pass(net, precision, cur, static_cast<long>(curPitch), src.type(), w, h,
                                  ob.ptr, static_cast<long>(opitch), ots, st.stream);
        if (rc != ACHIP_OK) return fail(std::string("achip_pass: ") + achip_last_error());
        cur = ob.ptr;
        curPitch = opitch;
        w *= 2;
        h *= 2;
    }
    e = hipMemcpy2DAsync(dst.ptr(), dst.stride(), cur, curPitch, static_cast<size_t>(w) * des, h,
                         hipMemcpyDeviceToHost, st.stream);
    if (e != hipSuccess) return fail(std::string("hipMemcpy2DAsync(D2H): ") + hipGetErrorString(e));
    e = hipStreamSynchronize(st.stream);
    if (e != hipSuccess) return fail(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return true;
}

void ac::core::hip::HIPProcessorACNetLegacy::process(const Image& src, Image& dst)
{
    run(src, dst, 1);
}

bool ac::core::hip::HIPProcessorACNetLegacy::processChain(const Image& src, Image& dst, const int passes)
{
    return run(src, dst, passes);
}

std::shared_ptr<ac::core::Processor> ac::core::detail::createHIP(const int device, const model::ACNetLegacy& model, const int precision)
{
    return std::make_shared<hip::HIPProcessorACNetLegacy>(device, model, precision);
}

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
