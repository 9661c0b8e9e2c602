// C++ checks of the drop-in core API, written like the reference's doctest suites
// (tests/core/src/{ImageTest,ModelTest,ProcessorTest}.cpp).  Built with g++ against
// include/ and lib/libac_hip.so by tests/test_cpp_core.py.  With argument "gpu" it also
// runs the processor cases that need a device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "AC/Core.hpp"

static int failures = 0, checks = 0;
#define CHECK(cond) do { checks++; if (!(cond)) { failures++; std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); } } while (0)
#define TEST_CASE(name) static void name()

TEST_CASE(image_default)
{
    ac::core::Image img{};
    CHECK(img.empty());
    CHECK(img.width() == 0 && img.height() == 0 && img.channels() == 0);
    CHECK(img.size() == 0 && img.data() == nullptr && !img.ownership());
}

TEST_CASE(image_sized)
{
    ac::core::Image img{ 4, 3, 3, ac::core::Image::UInt8 };
    CHECK(!img.empty() && img.width() == 4 && img.height() == 3 && img.channels() == 3);
    CHECK(img.elementSize() == 1 && img.pixelSize() == 3 && img.isUint() && !img.isFloat());
    CHECK(img.ownership() && img.stride() >= 12 && img.size() == img.height() * img.stride());
    CHECK(img.stride() % 4 == 0);   // create() aligns rows to 4 bytes
    ac::core::Image s{ 4, 3, 3, ac::core::Image::UInt8, 16 };
    CHECK(s.stride() == 16 && s.size() == 48);
    ac::core::Image f{ 5, 2, 1, ac::core::Image::Float32 };
    CHECK(f.stride() == 20 && f.isFloat() && f.elementSize() == 4);
    ac::core::Image h{ 5, 2, 1, ac::core::Image::Float16 };
    CHECK(h.stride() == 12 && h.elementSize() == 2);
}

TEST_CASE(image_external_and_views)
{
    std::uint8_t buffer[4 * 3 * 4]{};
    ac::core::Image img{ 4, 3, 4, ac::core::Image::UInt8, buffer };
    CHECK(img.ptr() == static_cast<void*>(buffer) && !img.ownership() && img.stride() == 16);
    ac::core::Image v = img.view(1, 1, 10, 10);
    CHECK(v.width() == 3 && v.height() == 2 && v.ptr() == img.ptr(1, 1) && v.stride() == 16);
    ac::core::Image none = img.view(5, 5, 2, 2);
    CHECK(none.empty());
    ac::core::Image inv{};
    inv.create(0, 0, 0, ac::core::Image::UInt8);
    CHECK(inv.empty());
}

TEST_CASE(image_from_to_clone)
{
    std::uint8_t src[3 * 5];
    for (int i = 0; i < 15; i++) src[i] = static_cast<std::uint8_t>(i * 7);
    ac::core::Image a{};
    a.from(5, 3, 1, ac::core::Image::UInt8, src);
    CHECK(a.ownership() && a.stride() == 8);
    CHECK(a.data()[8] == src[5] && a.data()[2 * 8 + 4] == src[14]);
    std::uint8_t out[3 * 6]{};
    a.to(out, 6);
    CHECK(out[6] == src[5] && out[5] == 0);
    ac::core::Image b = a.clone();
    CHECK(b != a && b.data()[8] == src[5]);
    ac::core::Image c = a;
    CHECK(c == a);
}

TEST_CASE(model_acnet_legacy)
{
    using M = ac::core::model::ACNetLegacy;
    for (auto v : { M::Variant::GAN, M::Variant::HDN0, M::Variant::HDN1, M::Variant::HDN2, M::Variant::HDN3 })
    {
        M model{ v };
        CHECK(model.blocks() == 8 && model.kernels() == 10 && model.biases() == 9 && model.alphas() == 0);
        CHECK(model.kernel(0) != nullptr && model.bias(0) != nullptr);
        CHECK(model.kernelOffset(0) == 0 && model.biasOffset(0) == 0);
        const int lk = model.kernels() - 1, lb = model.biases() - 1;
        CHECK(model.kernelOffset(lk) + model.kernelLength(lk) == model.kernelLength());
        CHECK(model.biasOffset(lb) + model.biasLength(lb) == model.biasLength());
        int ks = 0, bs = 0;
        for (int i = 0; i < model.kernels(); i++) ks += model.kernelLength(i);
        for (int i = 0; i < model.biases(); i++) bs += model.biasLength(i);
        CHECK(ks == 4712 && ks == model.kernelLength() && bs == 72 && bs == model.biasLength());
        CHECK(model.kernelOffset(9) == 4680 && model.kernelOffset(8) == 4104);
    }
    // first weights of HDN0 and GAN (ACNet.p lines 606 and 2)
    CHECK(std::fabs(M{ M::Variant::HDN0 }.kernel(0)[0] - 0.0609167f) < 1e-9f);
    CHECK(std::fabs(M{ M::Variant::GAN }.kernel(0)[0] - 0.0441616f) < 1e-9f);
}

TEST_CASE(artcnn_layer_algebra)
{
    // ArtCNN.hpp:28-55: head F*9, blockNum + 1 F->F layers, tail F*4*9
    using A16 = ac::core::model::ArtCNN<16>;
    using A32 = ac::core::model::ArtCNN<32>;
    A16 m{ A16::Variant::C4_DS };
    CHECK(m.blocks() == 4 && m.kernels() == 7 && m.biases() == 7);
    CHECK(m.kernelLength() == 144 + 2304 * 5 + 576 && m.biasLength() == 16 + 80 + 4);
    CHECK(m.kernelOffset(1) == 144 && m.kernelOffset(6) == 144 + 2304 * 5 && m.kernelLength(6) == 576);
    CHECK(m.biasOffset(6) == 96 && m.biasLength(6) == 4 && m.biasLength(5) == 16);
    A32 n{ A32::Variant::C4_DN };
    CHECK(n.kernelLength() == 288 + 9216 * 5 + 1152 && n.biasLength() == 32 + 160 + 4);
    CHECK(A16{ A16::Variant::C4_NORMAL }.kernel(0) != m.kernel(0) && A16{ A16::Variant::C4_DN }.kernel(0) != m.kernel(0));
}

TEST_CASE(fsrcnnx_layer_algebra)
{
    // FSRCNNX.hpp:29-78: head F*25, blockNum conv3x3 F*F*9, conv1x1 F*F, tail F*4*9; slopes
    // for layers 1..blockNum+1 (alpha(0) is alpha(1))
    using X8 = ac::core::model::FSRCNNX<8>;
    using X16 = ac::core::model::FSRCNNX<16>;
    X8 m{ X8::Variant::B4_NORMAL };
    CHECK(m.blocks() == 4 && m.kernels() == 7 && m.biases() == 7 && m.alphas() == 5);
    CHECK(m.kernelLength() == 200 + 576 * 4 + 64 + 288 && m.biasLength() == 8 + 40 + 4 && m.alphaLength() == 40);
    CHECK(m.kernelOffset(1) == 200 && m.kernelOffset(5) == 200 + 576 * 4 && m.kernelOffset(6) == 200 + 576 * 4 + 64);
    CHECK(m.kernelLength(5) == 64 && m.kernelLength(6) == 288 && m.biasLength(6) == 4);
    CHECK(m.alphaOffset(0) == 0 && m.alphaOffset(1) == 0 && m.alphaOffset(5) == 32 && m.alpha(0) == m.alpha(1));
    X16 n{ X16::Variant::B4_DISTORT_PLUS };
    CHECK(n.kernelLength() == 400 + 2304 * 4 + 256 + 576 && n.biasLength() == 16 + 80 + 4 && n.alphaLength() == 80);
    CHECK(X16{ X16::Variant::B4_NORMAL }.kernel(0) != n.kernel(0));
}

// gpu: a device is present, so the families this library provides must come up ok
static void processor_factory(const bool gpu)
{
    auto cpu = ac::core::Processor::create("cpu", 0, "acnet-legacy-hdn0");
    CHECK(static_cast<bool>(cpu) && !cpu->ok() && cpu->type() == ac::core::Processor::CPU);
    CHECK(std::strstr(cpu->error(), "HIP processor only") != nullptr);
    auto fam = ac::core::Processor::create("hip", 0, "arnet-b8");
    CHECK(!fam->ok() && std::strstr(fam->error(), "ARNet") != nullptr);
    // provided families: without a device they fail on the device only
    for (const char* model : { "fsrcnnx-f16-dp", "artcnn-c4f16", "acnet-b4" })
    {
        auto p = ac::core::Processor::create("hip", 0, model);
        if (gpu) CHECK(p->ok());
        else CHECK(!p->ok() && std::strstr(p->error(), "no HIP device") != nullptr);
    }
    // precision suffixes, parsed whole: "bf16" contains "f16" and must still mean BF16.
    // The FP32-only families (ArtCNN, FSRCNNX) reject 16-bit types before touching a device
    // and name the precision they were asked for, so the parse is visible without a GPU.
    const struct { const char* type; const char* want; } suffixes[] = {
        { "hip-bf16", "BF16" }, { "hip:bf16", "BF16" }, { "HIP-BF16", "BF16" },
        { "hip-fp16", "FP16" }, { "hip:fp16", "FP16" }, { "hip-f16", "FP16" } };
    for (const auto& c : suffixes)
    {
        auto q = ac::core::Processor::create(c.type, 0, "artcnn-c4f16");
        CHECK(!q->ok() && std::strstr(q->error(), "FP32 only") != nullptr);
        const char* tail = std::strstr(q->error(), "asked for ");
        CHECK(tail != nullptr && std::strcmp(tail + 10, c.want) == 0);
        // ACNet<8> and ACNetLegacy run every precision
        for (const char* model : { "acnet-b8", "acnet-legacy-hdn0" })
        {
            auto r = ac::core::Processor::create(c.type, 0, model);
            if (gpu) CHECK(r->ok());
            else CHECK(!r->ok() && std::strstr(r->error(), "no HIP device") != nullptr);
        }
    }
    for (const char* fp32 : { "hip", "hip-", "hip-fp32", "hip-bf16x" })
    {
        auto q = ac::core::Processor::create(fp32, 0, "acnet-b8");
        if (gpu) CHECK(q->ok());
        else CHECK(!q->ok() && std::strstr(q->error(), "no HIP device") != nullptr);
    }
    auto p = ac::core::Processor::create("HIP", 0, "acnet-legacy-hdn2");
    CHECK(static_cast<bool>(p) && p->type() == ac::core::Processor::HIP && std::strcmp(p->typeName(), "HIP") == 0);
    auto a = ac::core::Processor::create("auto", -1, nullptr);
    CHECK(a->type() == ac::core::Processor::HIP);
    CHECK(ac::core::Processor::listInfo() != nullptr && std::strlen(ac::core::Processor::listInfo()) > 0);
}

TEST_CASE(processor_no_device)
{
    auto p = ac::core::Processor::create("hip", 0, "acnet-legacy-hdn0");
    CHECK(!p->ok() && std::strstr(p->error(), "no HIP device") != nullptr);
    ac::core::Image src{ 8, 8, 1, ac::core::Image::UInt8 };
    auto dst = p->process(src, 2.0);
    CHECK(dst.empty() && !p->ok());
}

// ProcessorTest.cpp:45-104 on the HIP processor, 1-channel (the path this library builds)
TEST_CASE(processor_gpu)
{
    auto proc = ac::core::Processor::create("hip", 0, "acnet-legacy-hdn0");
    CHECK(proc->ok());
    ac::core::Image src{ 32, 32, 1, ac::core::Image::UInt8 };
    std::memset(src.ptr(), 128, src.size());
    auto dst = proc->process(src, 2.0);
    CHECK(proc->ok() && dst.width() == 64 && dst.height() == 64 && dst.channels() == 1 && dst.type() == src.type());
    ac::core::Image pre{ 64, 64, 1, ac::core::Image::UInt8 };
    proc->process(src, pre, 2.0);
    CHECK(proc->ok() && pre.width() == 64);
    CHECK(std::memcmp(pre.ptr(), dst.ptr(), 64) == 0);
    ac::core::Image s16{ 16, 16, 1, ac::core::Image::UInt8 };
    std::memset(s16.ptr(), 40, s16.size());
    auto d4 = proc->process(s16, 4.0);
    CHECK(proc->ok() && d4.width() == 64 && d4.height() == 64);
    // a constant image stays constant (up to the 2x2 deconv pattern) under replicate padding
    const std::uint8_t* d = d4.data();
    CHECK(d[0] == d[2 * d4.stride() + 2] && d[1] == d[2 * d4.stride() + 3]);
    // the layer-by-layer families through the same factory, from a model object
    auto art = ac::core::Processor::create<ac::core::Processor::HIP>(0, ac::core::model::ArtCNN<32>{ ac::core::model::ArtCNN<32>::Variant::C4_DN });
    auto fsr = ac::core::Processor::create<ac::core::Processor::HIP>(0, ac::core::model::FSRCNNX<8>{ ac::core::model::FSRCNNX<8>::Variant::B4_NORMAL });
    for (auto* q : { art.get(), fsr.get() })
    {
        auto o = q->process(src, 2.0);
        CHECK(q->ok() && o.width() == 64 && o.height() == 64 && o.type() == src.type());
        // constant in, constant out away from the borders (replicate padding)
        const std::uint8_t* e = o.data();
        CHECK(e[20 * o.stride() + 20] == e[30 * o.stride() + 32]);   // same 2x2 sub-pixel phase
    }
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
    image_default();
    image_sized();
    image_external_and_views();
    image_from_to_clone();
    model_acnet_legacy();
    artcnn_layer_algebra();
    fsrcnnx_layer_algebra();
    processor_factory(gpu);
    if (gpu) processor_gpu();
    else processor_no_device();
    std::printf("%d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
