// ac::specs -- the model and processor listings that callers iterate
// (core/include/AC/Specs.hpp:4-291: the Model / Processor structs, ModelList and
// ProcessorList; read by cli/src/Main.cpp:50-62, gui/src/MainWindow.cpp:118-126,
// filter/directshow/src/Filter.cpp:533-584, binding/python/src/Binding.cpp:159-190 and
// tools/benchmark/src/Benchmark.cpp:76).
//
// The two struct shapes are the interface and keep the reference's members and order.
// The lists hold what this library runs:
//   - ModelList: every weight set of ACNetLegacy, ACNet<8>, ArtCNN and FSRCNNX, named by
//     the model strings Processor::create parses.  ARNet is absent: its weight table is
//     not part of the reference snapshot, so the library cannot run it.
//     parameterCount = kernels + biases + PReLU slopes of the network, the sum the model
//     classes in AC/Core/Model/*.hpp give (kernelLength() + biasLength() + alphaLength()).
//   - ProcessorList: "auto" (the first HIP device) and the HIP processor in its three
//     arithmetic forms.  There is no CPU, OpenCL or CUDA processor in this library.
// ModelList[0] is the model a caller gets when it names none (Benchmark.cpp:76).
#ifndef AC_SPECS_HPP
#define AC_SPECS_HPP

namespace ac::specs
{
    struct Model
    {
        const char* name;
        const char* description;
        int parameterCount;
        const char* version = nullptr;
        const char* author = nullptr;
        const char* homepage = nullptr;
    };

    struct Processor
    {
        const char* name;
        const char* description;
    };

    namespace detail
    {
        // network sizes from the layer algebra of AC/Core/Model/*.hpp
        constexpr int acnetLegacyParams() { return (8 * 9 + 8 * 8 * 9 * 8 + 8 * 4) + (8 + 8 * 8); }
        constexpr int acnet8Params(const int blocks)
        {
            return (8 * 9 + 8 * 8 * 9 * blocks + 8 * 4 * 9) + (8 + 8 * blocks + 4) + (8 + 8 * blocks);
        }
        constexpr int artcnnParams(const int f, const int blocks)
        {
            return (f * 9 + f * f * 9 * (blocks + 1) + f * 4 * 9) + (f + f * (blocks + 1) + 4);
        }
        constexpr int fsrcnnxParams(const int f, const int blocks)
        {
            return (f * 25 + f * f * 9 * blocks + f * f + f * 4 * 9) + (f + f * (blocks + 1) + 4) + f * (blocks + 1);
        }

        constexpr const char* artcnnVersion = "v1.6.2";
        constexpr const char* artcnnAuthor = "Artoriuz";
        constexpr const char* artcnnHome = "https://github.com/Artoriuz/ArtCNN";
        constexpr const char* fsrcnnxHome = "https://github.com/igv/FSRCNN-TensorFlow";
        constexpr const char* fsrcnnxDpHome = "https://github.com/nessotrin/FSRCNN-TensorFlow";
    }

    constexpr Model ModelList[] = {
        { "acnet-legacy-gan", "ACNet (legacy): 8-layer CNN 2x upscaler, GAN-trained for sharper detail.", detail::acnetLegacyParams() },
        { "acnet-legacy-hdn0", "ACNet (legacy): 8-layer CNN 2x upscaler, light denoising.", detail::acnetLegacyParams() },
        { "acnet-legacy-hdn1", "ACNet (legacy): 8-layer CNN 2x upscaler, medium denoising.", detail::acnetLegacyParams() },
        { "acnet-legacy-hdn2", "ACNet (legacy): 8-layer CNN 2x upscaler, heavy denoising.", detail::acnetLegacyParams() },
        { "acnet-legacy-hdn3", "ACNet (legacy): 8-layer CNN 2x upscaler, strongest denoising.", detail::acnetLegacyParams() },

        { "acnet-f8b4", "ACNet 8 channels x 4 blocks (PReLU, pixel shuffle), neutral.", detail::acnet8Params(4) },
        { "acnet-f8b4-hdn", "ACNet 8 channels x 4 blocks, mild denoising.", detail::acnet8Params(4) },
        { "acnet-f8b4-box", "ACNet 8 channels x 4 blocks, trained on box-filtered input (cleaner lines, slightly softer).", detail::acnet8Params(4) },
        { "acnet-f8b4-box-hdn", "ACNet 8 channels x 4 blocks, box-trained with mild denoising.", detail::acnet8Params(4) },
        { "acnet-f8b8", "ACNet 8 channels x 8 blocks (PReLU, pixel shuffle), neutral.", detail::acnet8Params(8) },
        { "acnet-f8b8-hdn", "ACNet 8 channels x 8 blocks, mild denoising.", detail::acnet8Params(8) },
        { "acnet-f8b8-box", "ACNet 8 channels x 8 blocks, trained on box-filtered input (cleaner lines, slightly softer).", detail::acnet8Params(8) },
        { "acnet-f8b8-box-hdn", "ACNet 8 channels x 8 blocks, box-trained with mild denoising.", detail::acnet8Params(8) },
        { "acnet-f8b18", "ACNet 8 channels x 18 blocks (PReLU, pixel shuffle), neutral.", detail::acnet8Params(18) },
        { "acnet-f8b18-hdn", "ACNet 8 channels x 18 blocks, mild denoising.", detail::acnet8Params(18) },
        { "acnet-f8b18-box", "ACNet 8 channels x 18 blocks, trained on box-filtered input (cleaner lines, slightly softer).", detail::acnet8Params(18) },
        { "acnet-f8b18-box-hdn", "ACNet 8 channels x 18 blocks, box-trained with mild denoising.", detail::acnet8Params(18) },

        { "artcnn-c4f16", "ArtCNN C4F16 by Artoriuz, unmodified weights, neutral.", detail::artcnnParams(16, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },
        { "artcnn-c4f16-dn", "ArtCNN C4F16_DN by Artoriuz, unmodified weights, denoise and soften.", detail::artcnnParams(16, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },
        { "artcnn-c4f16-ds", "ArtCNN C4F16_DS by Artoriuz, unmodified weights, denoise and sharpen.", detail::artcnnParams(16, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },
        { "artcnn-c4f32", "ArtCNN C4F32 by Artoriuz, unmodified weights, neutral.", detail::artcnnParams(32, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },
        { "artcnn-c4f32-dn", "ArtCNN C4F32_DN by Artoriuz, unmodified weights, denoise and soften.", detail::artcnnParams(32, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },
        { "artcnn-c4f32-ds", "ArtCNN C4F32_DS by Artoriuz, unmodified weights, denoise and sharpen.", detail::artcnnParams(32, 4),
          detail::artcnnVersion, detail::artcnnAuthor, detail::artcnnHome },

        { "fsrcnnx-f8b4", "FSRCNNX_x2_8-0-4-1 by igv, unmodified weights, light denoising.", detail::fsrcnnxParams(8, 4),
          "v1.1", "igv", detail::fsrcnnxHome },
        { "fsrcnnx-f8b4-distort-plus", "FSRCNNX-x2-2-8-0-4-1.v1.fastv2 by nessotrin, unmodified weights, strong denoising.", detail::fsrcnnxParams(8, 4),
          "v1.3_distort_plus", "nessotrin", detail::fsrcnnxDpHome },
        { "fsrcnnx-f16b4", "FSRCNNX_x2_16-0-4-1 by igv, unmodified weights, light denoising.", detail::fsrcnnxParams(16, 4),
          "v1.1", "igv", detail::fsrcnnxHome },
        { "fsrcnnx-f16b4-distort-plus", "FSRCNNX-x2-2-16-0-4-1.v1.fastv2 by nessotrin, unmodified weights, strong denoising.", detail::fsrcnnxParams(16, 4),
          "v1.3_distort_plus", "nessotrin", detail::fsrcnnxDpHome },
    };

    constexpr Processor ProcessorList[] = {
        { "auto", "The first HIP device (this library has no CPU processor)." },
        { "hip", "AMD Instinct (gfx950) HIP processor, FP32 arithmetic as the reference CPU path." },
        { "hip-fp16", "HIP processor, ACNet (legacy) 8->8 layers on FP16 matrix cores, FP32 accumulate." },
        { "hip-bf16", "HIP processor, ACNet (legacy) 8->8 layers on BF16 matrix cores, FP32 accumulate." },
    };
}

#endif
