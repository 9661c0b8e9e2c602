// ac::core::model::ArtCNN<F> -- the network of core/include/AC/Core/Model/ArtCNN.hpp:11-56
// (F = 16 or 32 channels): conv3x3 1->F, blockNum (4) conv3x3 F->F with ReLU, conv3x3 F->F
// added to the first layer's output, conv3x3 F->4 + pixel shuffle; three weight sets each
// (normal, DN, DS).  Same layer -> offset algebra; weights are embedded in the library.
#ifndef AC_CORE_MODEL_ARTCNN_HPP
#define AC_CORE_MODEL_ARTCNN_HPP

#include <algorithm>
#include <cstddef>

#include "ACCoreExport.hpp"

namespace ac::core::model
{
    template<int F>
    class ArtCNN;
}

template<int F>
class ac::core::model::ArtCNN
{
public:
    enum class Variant
    {
        C4_NORMAL, C4_DN, C4_DS
    };

public:
    AC_CORE_EXPORT explicit ArtCNN(Variant v) noexcept;

public:
    Variant variant() const noexcept { return v; }
    int blocks() const noexcept { return blockNum; }
    int channels() const noexcept { return F; }

    // layers: 0 the head (F*9), 1..blockNum+1 F->F (F*F*9), blockNum+2 the tail (F*4*9)
    int kernelLength() const noexcept { return F * 9 + F * F * 9 * (blockNum + 1) + F * 4 * 9; }
    int kernelLength(const int layer) const noexcept { return layer == 0 ? F * 9 : (layer > 0 && layer < blockNum + 2 ? F * F * 9 : (layer == blockNum + 2 ? F * 4 * 9 : 0)); }
    int biasLength() const noexcept { return F + F * (blockNum + 1) + 4; }
    int biasLength(const int layer) const noexcept { return layer >= 0 && layer < blockNum + 2 ? F : (layer == blockNum + 2 ? 4 : 0); }
    std::size_t kernelSize() const noexcept { return kernelLength() * sizeof(float); }
    std::size_t biasSize() const noexcept { return biasLength() * sizeof(float); }

    int kernels() const noexcept { return blockNum + 3; }
    int biases() const noexcept { return blockNum + 3; }
    int kernelIndex(const int layer) const noexcept { return std::clamp(layer, 0, kernels() - 1); }
    int biasIndex(const int layer) const noexcept { return std::clamp(layer, 0, biases() - 1); }
    int kernelLayer(const int idx) const noexcept { return std::clamp(idx, 0, kernels() - 1); }
    int biasLayer(const int idx) const noexcept { return std::clamp(idx, 0, biases() - 1); }

    int kernelOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer < kernels()) return F * 9 + (F * F * 9) * (layer - 1);
        return kernelLength();
    }
    int biasOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer < biases()) return F * layer;
        return biasLength();
    }

    const float* kernel(const int layer = 0) const noexcept { return kptr + kernelOffset(layer); }
    const float* bias(const int layer = 0) const noexcept { return bptr + biasOffset(layer); }

private:
    Variant v;
    int blockNum;
    const float* kptr;
    const float* bptr;
};

#endif
