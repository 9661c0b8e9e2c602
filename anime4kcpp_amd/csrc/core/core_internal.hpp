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
    // model string -> 0 ACNetLegacy (`variant`), 1 ACNet<8> (`variant8`), -1 another family
    // (its name in `family`)
    int findModel(const char* model, model::ACNetLegacy::Variant& variant, model::ACNet<8>::Variant& variant8,
                  std::string& family) noexcept;
    int ceilLog2(double v) noexcept;
    const char* typeNameOf(int type) noexcept;

    // The calling thread's last image-helper failure (empty when the last helper call
    // succeeded or none failed since the previous take); clears it.
    std::string takeImageError() noexcept;

    // HIP processor for ACNetLegacy; precision 0 FP32, 1 FP16-MFMA, 2 BF16-MFMA
    std::shared_ptr<Processor> createHIP(int device, const model::ACNetLegacy& model, int precision);
    // HIP processor for ACNet<8> (FP32; precision must be 0)
    std::shared_ptr<Processor> createHIP8(int device, const model::ACNet<8>& model, int precision);
}
