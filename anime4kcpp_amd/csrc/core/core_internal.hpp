// Internal declarations shared by the C++ core sources.
#pragma once

#include <memory>
#include <string>

#include "AC/Core/Model.hpp"
#include "AC/Core/Processor.hpp"

namespace ac::core::detail
{
    int findProcessorType(const char* type) noexcept;
    int findPrecision(const char* type) noexcept;
    // What a model string names: family 0 ACNetLegacy (`legacy`), 1 ACNet<8> (`acnet8`),
    // 2 ArtCNN<16> and 3 ArtCNN<32> (`art`), 4 FSRCNNX<8> and 5 FSRCNNX<16> (`fsr`), -1 a
    // family this library does not provide (its name in `other`).
    struct ModelChoice
    {
        int family = 0;
        model::ACNetLegacy::Variant legacy = model::ACNetLegacy::Variant::GAN;
        model::ACNet<8>::Variant acnet8 = model::ACNet<8>::Variant::B8_NORMAL;
        model::ArtCNN<16>::Variant art = model::ArtCNN<16>::Variant::C4_NORMAL;
        model::FSRCNNX<8>::Variant fsr = model::FSRCNNX<8>::Variant::B4_NORMAL;
        std::string other;
    };
    ModelChoice findModel(const char* model) noexcept;
    int ceilLog2(double v) noexcept;
    const char* typeNameOf(int type) noexcept;

    // The calling thread's last image-helper failure (empty when the last helper call
    // succeeded or none failed since the previous take); clears it.
    std::string takeImageError() noexcept;

    // HIP processor for ACNetLegacy; precision 0 FP32, 1 FP16-MFMA, 2 BF16-MFMA
    std::shared_ptr<Processor> createHIP(int device, const model::ACNetLegacy& model, int precision);
    // HIP processor for ACNet<8> (FP32; precision must be 0)
    std::shared_ptr<Processor> createHIP8(int device, const model::ACNet<8>& model, int precision);
    // HIP processor for ArtCNN<16> / ArtCNN<32> (FP32; precision must be 0)
    template<int F>
    std::shared_ptr<Processor> createHIPArt(int device, const model::ArtCNN<F>& model, int precision);
    // HIP processor for FSRCNNX<8> / FSRCNNX<16> (FP32; precision must be 0)
    template<int F>
    std::shared_ptr<Processor> createHIPFsr(int device, const model::FSRCNNX<F>& model, int precision);
}
