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
    bool findModel(const char* model, model::ACNetLegacy::Variant& variant, std::string& family) noexcept;
    int ceilLog2(double v) noexcept;
    const char* typeNameOf(int type) noexcept;

    // The calling thread's last image-helper failure (empty when the last helper call
    // succeeded or none failed since the previous take); clears it.
    std::string takeImageError() noexcept;

    // HIP processor for ACNetLegacy; precision 0 FP32, 1 FP16-MFMA, 2 BF16-MFMA
    std::shared_ptr<Processor> createHIP(int device, const model::ACNetLegacy& model, int precision);
}
