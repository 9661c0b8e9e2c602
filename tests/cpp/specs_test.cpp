// AC/Specs.hpp as its callers use it: the CLI's listing loops (cli/src/Main.cpp:47-62),
// the GUI's combo boxes (gui/src/MainWindow.cpp:118-126) and the benchmark's default
// model (tools/benchmark/src/Benchmark.cpp:76), compiled against include/ and run.  The
// listed parameter counts must equal the model classes' layer algebra, and every listed
// model and processor name must reach a processor this library provides.  With argument
// "gpu" every listed model also runs one small pass.
#include <cstdint>
#include <cstdio>
#include <iterator>
#include <cstring>
#include <string>

#include "AC/Core.hpp"
#include "AC/Specs.hpp"

static int failures = 0, checks = 0;
#define CHECK(cond) do { checks++; if (!(cond)) { failures++; std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); } } while (0)

namespace M = ac::core::model;

// kernels + biases + slopes of the network a ModelList name stands for
static int paramsOf(const std::string& name)
{
    auto has = [&](const char* t) { return name.find(t) != std::string::npos; };
    if (has("legacy"))
    {
        M::ACNetLegacy m{ M::ACNetLegacy::Variant::GAN };
        return m.kernelLength() + m.biasLength();
    }
    if (has("artcnn"))
    {
        if (has("f32")) { M::ArtCNN<32> m{ M::ArtCNN<32>::Variant::C4_NORMAL }; return m.kernelLength() + m.biasLength(); }
        M::ArtCNN<16> m{ M::ArtCNN<16>::Variant::C4_NORMAL };
        return m.kernelLength() + m.biasLength();
    }
    if (has("fsrcnnx"))
    {
        if (has("f16")) { M::FSRCNNX<16> m{ M::FSRCNNX<16>::Variant::B4_NORMAL }; return m.kernelLength() + m.biasLength() + m.alphaLength(); }
        M::FSRCNNX<8> m{ M::FSRCNNX<8>::Variant::B4_NORMAL };
        return m.kernelLength() + m.biasLength() + m.alphaLength();
    }
    using V = M::ACNet<8>::Variant;
    const V v = has("b18") ? V::B18_NORMAL : (has("b4") ? V::B4_NORMAL : V::B8_NORMAL);
    M::ACNet<8> m{ v };
    return m.kernelLength() + m.biasLength() + m.alphaLength();
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";

    // the CLI's --list output, same loops over the same members
    std::printf("Processors:\n");
    for (auto&& processor : ac::specs::ProcessorList) std::printf("  %-16s  %s\n", processor.name, processor.description);
    std::printf("Models:\n");
    for (auto&& model : ac::specs::ModelList)
    {
        std::printf("  %s:\n", model.name);
        std::printf("    parameter count: %d\n", model.parameterCount);
        if (model.version) std::printf("    version: %s\n", model.version);
        if (model.author) std::printf("    author: %s\n", model.author);
        if (model.homepage) std::printf("    homepage: %s\n", model.homepage);
        std::printf("    description: %s\n", model.description);
    }

    CHECK(std::size(ac::specs::ModelList) == 27);
    CHECK(std::strcmp(ac::specs::ModelList[0].name, "acnet-legacy-gan") == 0);
    // published sizes of the networks (the reference lists the same counts)
    CHECK(paramsOf("acnet-legacy-gan") == 4784 && paramsOf("acnet-f8b8") == 5116 && paramsOf("artcnn-c4f32") == 47716 &&
          paramsOf("fsrcnnx-f8b4") == 2948);

    ac::core::Image src{ 24, 16, 1, ac::core::Image::UInt8 };
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 24; x++) *static_cast<std::uint8_t*>(src.ptr(x, y)) = static_cast<std::uint8_t>((x * 11 + y * 7) & 0xff);

    for (auto&& model : ac::specs::ModelList)
    {
        CHECK(model.parameterCount == paramsOf(model.name));
        CHECK(model.description && std::strlen(model.description) > 0);
        auto p = ac::core::Processor::create("hip", 0, model.name);
        if (!gpu)
        {
            // a family this library provides: it fails on the missing device only
            CHECK(!p->ok() && std::strstr(p->error(), "no HIP device") != nullptr);
            continue;
        }
        CHECK(p->ok());
        auto out = p->process(src, 2.0);
        CHECK(p->ok() && out.width() == 48 && out.height() == 32);
    }

    for (auto&& processor : ac::specs::ProcessorList)
    {
        auto p = ac::core::Processor::create(processor.name, 0, ac::specs::ModelList[0].name);
        CHECK(p->type() == ac::core::Processor::HIP);
        if (gpu)
        {
            CHECK(p->ok());
            auto out = p->process(src, 2.0);
            CHECK(p->ok() && out.width() == 48);
        }
        else CHECK(!p->ok() && std::strstr(p->error(), "no HIP device") != nullptr);
    }

    std::printf("%d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
