// ac::core::model::FSRCNNX<F> -- the network of core/include/AC/Core/Model/FSRCNNX.hpp:11-79
// (F = 8 or 16 channels): conv5x5 1->F, blockNum (4) conv3x3 F->F with PReLU, a conv1x1
// F->F added to the first layer's output then PReLU, conv3x3 F->4 + pixel shuffle; two
// weight sets each (normal, distort-plus).  Same layer -> offset algebra; weights are
// embedded in the library.
#ifndef AC_CORE_MODEL_FSRCNNX_HPP
#define AC_CORE_MODEL_FSRCNNX_HPP

#include <algorithm>
#include <cstddef>

#include "ACCoreExport.hpp"

namespace ac::core::model
{
    template<int F>
    class FSRCNNX;
}

template<int F>
class ac::core::model::FSRCNNX
{
public:
    enum class Variant
    {
        B4_NORMAL, B4_DISTORT_PLUS
    };

public:
    AC_CORE_EXPORT explicit FSRCNNX(Variant v) noexcept;

public:
    Variant variant() const noexcept { return v; }
    int blocks() const noexcept { return blockNum; }
    int channels() const noexcept { return F; }

    // layers: 0 the head (F*25), 1..blockNum conv3x3 (F*F*9), blockNum+1 conv1x1 (F*F),
    // blockNum+2 the tail (F*4*9); slopes for layers 1..blockNum+1
    int kernelLength() const noexcept { return F * 25 + F * F * 9 * blockNum + F * F + F * 4 * 9; }
    int kernelLength(const int layer) const noexcept
    {
        if (layer == 0) return F * 25;
        if (layer > 0 && layer < blockNum + 1) return F * F * 9;
        if (layer == blockNum + 1) return F * F;
        if (layer == blockNum + 2) return F * 4 * 9;
        return 0;
    }
    int biasLength() const noexcept { return F + F * (blockNum + 1) + 4; }
    int biasLength(const int layer) const noexcept { return layer >= 0 && layer < blockNum + 2 ? F : (layer == blockNum + 2 ? 4 : 0); }
    int alphaLength() const noexcept { return F * (blockNum + 1); }
    int alphaLength(const int layer) const noexcept { return layer >= 1 && layer < blockNum + 2 ? F : 0; }
    std::size_t kernelSize() const noexcept { return kernelLength() * sizeof(float); }
    std::size_t biasSize() const noexcept { return biasLength() * sizeof(float); }
    std::size_t alphaSize() const noexcept { return alphaLength() * sizeof(float); }

    int kernels() const noexcept { return blockNum + 3; }
    int biases() const noexcept { return blockNum + 3; }
    int alphas() const noexcept { return blockNum + 1; }
    int kernelIndex(const int layer) const noexcept { return std::clamp(layer, 0, kernels() - 1); }
    int biasIndex(const int layer) const noexcept { return std::clamp(layer, 0, biases() - 1); }
    int alphaIndex(const int layer) const noexcept { return std::clamp(layer, 1, alphas()) - 1; }
    int kernelLayer(const int idx) const noexcept { return std::clamp(idx, 0, kernels() - 1); }
    int biasLayer(const int idx) const noexcept { return std::clamp(idx, 0, biases() - 1); }
    int alphaLayer(const int idx) const noexcept { return std::clamp(idx, 0, alphas() - 1) + 1; }

    int kernelOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer <= blockNum + 1) return F * 25 + F * F * 9 * (layer - 1);
        if (layer < kernels()) return F * 25 + F * F * 9 * blockNum + F * F;
        return kernelLength();
    }
    int biasOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer < biases()) return F * layer;
        return biasLength();
    }
    int alphaOffset(const int layer) const noexcept
    {
        const int idx = alphaIndex(layer);
        if (idx <= 0) return 0;
        if (idx < alphas()) return F * idx;
        return alphaLength();
    }

    const float* kernel(const int layer = 0) const noexcept { return kptr + kernelOffset(layer); }
    const float* bias(const int layer = 0) const noexcept { return bptr + biasOffset(layer); }
    const float* alpha(const int layer = 0) const noexcept { return aptr + alphaOffset(layer); }

private:
    Variant v;
    int blockNum;
    const float* kptr;
    const float* bptr;
    const float* aptr;
};

#endif
