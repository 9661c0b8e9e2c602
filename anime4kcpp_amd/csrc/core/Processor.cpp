// ac::core::Processor: type/model string parsing, factory, and the process(src, dst,
// factor) orchestration around the 2x passes (core/src/processor/Processor.cpp:12-333).
//
// This library builds one processor type, HIP (MI355X, ACNetLegacy).  The reference's
// CPU, OpenCL and CUDA types and its other model families are not part of it: asking for
// them yields a processor whose ok() is false and whose error() says so -- never a
// silent substitute.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "AC/Core/Model.hpp"
#include "AC/Core/Processor.hpp"
#include "core_internal.hpp"

namespace ac::core::detail
{
    // Processor.cpp:12-25, plus "hip": unknown strings mean CPU, "auto" is -1.
    int findProcessorType(const char* type) noexcept
    {
        if (type)
        {
            std::string s = type;
            for (char& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
            if (s == "auto") return -1;
            if (s == "opencl") return Processor::OpenCL;
            if (s == "cuda") return Processor::CUDA;
            if (s.rfind("hip", 0) == 0) return Processor::HIP;
        }
        return Processor::CPU;
    }

    // Arithmetic of the HIP path's 8->8 layers from the type string: "hip" (FP32),
    // "hip-fp16" / "hip:fp16" / "hip-f16" (FP16), "hip-bf16" / "hip:bf16" (BF16).  The
    // suffix after the first '-' or ':' is matched whole, so "bf16" (which contains
    // "f16") can never select FP16; any other suffix means FP32.
    int findPrecision(const char* type) noexcept
    {
        if (!type) return 0;
        std::string s = type;
        for (char& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        const auto cut = s.find_first_of("-:");
        if (cut == std::string::npos) return 0;
        const std::string suffix = s.substr(cut + 1);
        if (suffix == "bf16") return 2;
        if (suffix == "fp16" || suffix == "f16") return 1;
        return 0;
    }

    // Processor.cpp:26-187 restricted to what this library implements: ACNetLegacy
    // ("acnet-legacy-{gan,hdn0..3}"; "hdn" without a digit is HDN0, the first digit 0-3 after
    // it selects the variant; no model string means GAN), ACNet<8> ("acnet" without
    // "legacy": box / hdn select the family member, b4 / b8 / b18 the depth, B8 by default)
    // ArtCNN ("artcnn": f32 selects 32 channels, else 16; dn, else ds, the weight set;
    // Processor.cpp:59-77) and FSRCNNX ("fsrcnnx": f16 selects 16 channels, else 8;
    // "distort" or "dp" the distort-plus weights; Processor.cpp:40-58).  ARNet reports its
    // name.
    ModelChoice findModel(const char* model) noexcept
    {
        ModelChoice m;
        if (!model) return m;
        std::string s = model;
        for (char& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        auto has = [&](const char* t) { return s.find(t) != std::string::npos; };
        if (has("fsrcnnx"))
        {
            using V = model::FSRCNNX<8>::Variant;
            m.family = has("f16") ? 5 : 4;
            m.fsr = has("distort") || has("dp") ? V::B4_DISTORT_PLUS : V::B4_NORMAL;
            return m;
        }
        if (has("artcnn"))
        {
            using V = model::ArtCNN<16>::Variant;
            m.family = has("f32") ? 3 : 2;
            m.art = has("dn") ? V::C4_DN : (has("ds") ? V::C4_DS : V::C4_NORMAL);
            return m;
        }
        if (has("arnet")) { m.family = -1; m.other = "ARNet"; return m; }
        if (!has("acnet")) return m;
        auto& variant = m.legacy;
        if (has("legacy"))
        {
            if (has("hdn"))
            {
                variant = model::ACNetLegacy::Variant::HDN0;
                for (char ch : s)
                {
                    if (ch == '0') variant = model::ACNetLegacy::Variant::HDN0;
                    else if (ch == '1') variant = model::ACNetLegacy::Variant::HDN1;
                    else if (ch == '2') variant = model::ACNetLegacy::Variant::HDN2;
                    else if (ch == '3') variant = model::ACNetLegacy::Variant::HDN3;
                    else continue;
                    break;
                }
            }
            return m;
        }
        using V = model::ACNet<8>::Variant;
        const bool box = has("box"), hdn = has("hdn");
        const int member = box ? (hdn ? 3 : 2) : (hdn ? 1 : 0);   // normal, hdn, box, box_hdn
        int depth = 1;                                            // b8 unless b4 / b18 (b8 tested first)
        if (has("b4")) depth = 0;
        else if (has("b8")) depth = 1;
        else if (has("b18")) depth = 2;
        m.family = 1;
        m.acnet8 = static_cast<V>(depth * 4 + member);
        return m;
    }

    // ceil(log2(v)) from the IEEE exponent and mantissa (core/internal/.../Util.hpp:80-89)
    int ceilLog2(const double v) noexcept
    {
        std::uint64_t bits = 0;
        std::memcpy(&bits, &v, sizeof(bits));
        return static_cast<int>((((bits >> 52) & 0x7ff) - 1023) + ((bits << 12) != 0));
    }

    const char* typeNameOf(int type) noexcept
    {
        switch (type)
        {
        case Processor::CPU: return "CPU";
        case Processor::OpenCL: return "OpenCL";
        case Processor::CUDA: return "CUDA";
        case Processor::HIP: return "HIP";
        default: return "unknown";
        }
    }

    // A processor this library cannot provide: ok() is false everywhere, process() does
    // nothing.
    class UnavailableProcessor final : public Processor
    {
    public:
        UnavailableProcessor(const int type, std::string message) noexcept : t(type), msg(std::move(message)) {}
        bool ok() noexcept override { return false; }
        const char* error() noexcept override { return msg.c_str(); }
        const char* name() const noexcept override { return "unavailable"; }
        int type() const noexcept override { return t; }
        const char* typeName() const noexcept override { return typeNameOf(t); }

    private:
        void process(const Image&, Image&) override {}
        bool processChain(const Image&, Image&, int) override { return false; }
        bool processAll(const Image&, Image&, double) override { return true; }

    private:
        int t;
        std::string msg;
    };
}

ac::core::Processor::Processor() noexcept : idx(0) {}
ac::core::Processor::~Processor() = default;

ac::core::Image ac::core::Processor::process(const Image& src, const double factor)
{
    Image dst{};
    process(src, dst, factor);
    return dst;
}

// Processor.cpp:199-276.  power = factor > 2 ? ceil(log2(factor)) : 1 passes, then a resize
// by factor / 2^power; multi-channel input is split into Y (through the passes) and UV
// (resized by the full factor) and merged back.
void ac::core::Processor::process(const Image& src, Image& dst, const double factor)
{
    if (src.empty()) return;
    if (processAll(src, dst, factor)) return;
    const int power = factor > 2.0 ? detail::ceilLog2(factor) : 1;
    const double fxy = factor / static_cast<double>(1 << power);

    Image y = src, uv{};
    if (src.channels() > 1 && !colorSplit(src, y, uv)) return;

    if (!dst.empty())
    {
        if (src.channels() == 1)
        {
            if (fxy == 1.0) processChain(y, dst, power);
            else
            {
                Image out{};
                if (processChain(y, out, power)) resizeTo(out, dst, 0.0, 0.0);
            }
        }
        else
        {
            Image out{};
            if (!processChain(y, out, power)) return;
            if (fxy != 1.0 && !resizeTo(out, out, fxy, fxy)) return;
            Image uvs{};
            if (!resizeTo(uv, uvs, factor, factor)) return;
            colorMerge(out, uvs, dst);
        }
    }
    else
    {
        Image out{};
        if (!processChain(y, out, power)) return;
        if (!resizeTo(out, dst, fxy, fxy)) return;
        if (src.channels() > 1)
        {
            Image uvs{}, rgb{};
            if (!resizeTo(uv, uvs, factor, factor)) return;
            if (!colorMerge(dst, uvs, rgb)) return;
            dst = rgb;
        }
    }
}

// The reference's pass loop: intermediate images keep the input's element type; the last
// pass writes into dst (allocated with the input's type when empty).
bool ac::core::Processor::processChain(const Image& src, Image& dst, const int passes)
{
    Image in{}, out{ src };
    for (int i = 0; i < passes - 1; i++)
    {
        in = out;
        out = Image{};
        out.create(in.width() * 2, in.height() * 2, 1, in.type());
        process(in, out);
        if (!ok()) return false;
    }
    if (dst.empty()) dst.create(out.width() * 2, out.height() * 2, 1, out.type());
    process(out, dst);
    return ok();
}

bool ac::core::Processor::processAll(const Image&, Image&, double)
{
    return false;
}

bool ac::core::Processor::colorSplit(const Image&, Image&, Image&)
{
    setError("multi-channel input is not supported by this processor");
    return false;
}

bool ac::core::Processor::colorMerge(const Image&, const Image&, Image&)
{
    setError("multi-channel output is not supported by this processor");
    return false;
}

// ImageResize.cpp:136-163 semantics for the identity cases (dst = src, no copy); other
// sizes need a processor that implements resizing.
bool ac::core::Processor::resizeTo(const Image& src, Image& dst, const double fx, const double fy)
{
    if (src.empty()) return false;
    if ((fx > 0.0 && fy > 0.0 && fx == 1.0 && fy == 1.0) ||
        (!(fx > 0.0 && fy > 0.0) && !dst.empty() && dst.width() == src.width() && dst.height() == src.height()))
    {
        dst = src;
        return true;
    }
    setError("non power-of-two factors need a resize, which this processor does not implement");
    return false;
}

void ac::core::Processor::setError(const char*) noexcept {}

bool ac::core::Processor::ok() noexcept
{
    return true;
}
const char* ac::core::Processor::error() noexcept
{
    return "NO ERROR";
}

// Processor.cpp:286-317: "auto" tries the GPU first.  Here it is HIP and only HIP.
std::shared_ptr<ac::core::Processor> ac::core::Processor::create(const char* const type, const int device, const char* const model)
{
    int t = detail::findProcessorType(type);
    if (t == -1) t = Processor::HIP;
    if (t != Processor::HIP)
        return std::make_shared<detail::UnavailableProcessor>(t, std::string("processor type '") + detail::typeNameOf(t) +
            "' is not part of this library, which provides the HIP processor only");
    const detail::ModelChoice m = detail::findModel(model);
    const int precision = detail::findPrecision(type);
    switch (m.family)
    {
    case 0: return detail::createHIP(device, model::ACNetLegacy{ m.legacy }, precision);
    case 1: return detail::createHIP8(device, model::ACNet<8>{ m.acnet8 }, precision);
    case 2: return detail::createHIPArt<16>(device, model::ArtCNN<16>{ m.art }, precision);
    case 3:
        return detail::createHIPArt<32>(device, model::ArtCNN<32>{ static_cast<model::ArtCNN<32>::Variant>(m.art) }, precision);
    case 4: return detail::createHIPFsr<8>(device, model::FSRCNNX<8>{ m.fsr }, precision);
    case 5:
        return detail::createHIPFsr<16>(device, model::FSRCNNX<16>{ static_cast<model::FSRCNNX<16>::Variant>(m.fsr) }, precision);
    default:
        return std::make_shared<detail::UnavailableProcessor>(t, "model family " + m.other +
            " is not part of this library, which provides ACNetLegacy, ACNet<8>, ArtCNN and FSRCNNX");
    }
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::ACNetLegacy>(const int idx, const model::ACNetLegacy& model)
{
    return detail::createHIP(idx, model, 0);
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::ACNet<8>>(const int idx, const model::ACNet<8>& model)
{
    return detail::createHIP8(idx, model, 0);
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::ArtCNN<16>>(const int idx, const model::ArtCNN<16>& model)
{
    return detail::createHIPArt<16>(idx, model, 0);
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::ArtCNN<32>>(const int idx, const model::ArtCNN<32>& model)
{
    return detail::createHIPArt<32>(idx, model, 0);
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::FSRCNNX<8>>(const int idx, const model::FSRCNNX<8>& model)
{
    return detail::createHIPFsr<8>(idx, model, 0);
}

template<>
AC_CORE_EXPORT std::shared_ptr<ac::core::Processor> ac::core::Processor::create<ac::core::Processor::HIP, ac::core::model::FSRCNNX<16>>(const int idx, const model::FSRCNNX<16>& model)
{
    return detail::createHIPFsr<16>(idx, model, 0);
}

const char* ac::core::Processor::listInfo()
{
    return info<Processor::HIP>();
}
