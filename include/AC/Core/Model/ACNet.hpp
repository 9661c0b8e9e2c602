// ac::core::model::ACNetLegacy -- the network of core/include/AC/Core/Model/ACNet.hpp:18-61:
// conv3x3 1->8, blockNum (8) conv3x3 8->8 with ReLU, deconv 2x2 stride 2 8->1; five weight
// sets.  ac::core::model::ACNet<8> -- ACNet.hpp:63-118: conv3x3 1->8 + PReLU, blockNum (4,
// 8 or 18) conv3x3 8->8 + PReLU, conv3x3 8->4 + pixel shuffle + the source pixel; twelve
// weight sets.  Same layer -> offset algebra; weights are embedded in the library.
#ifndef AC_CORE_MODEL_ACNET_HPP
#define AC_CORE_MODEL_ACNET_HPP

#include <algorithm>
#include <cstddef>

#include "ACCoreExport.hpp"

namespace ac::core::model
{
    class ACNetLegacy;

    template<int F>
    class ACNet;
}

class ac::core::model::ACNetLegacy
{
public:
    enum class Variant
    {
        GAN,
        HDN0,
        HDN1,
        HDN2,
        HDN3
    };

public:
    AC_CORE_EXPORT explicit ACNetLegacy(Variant v) noexcept;

public:
    Variant variant() const noexcept { return v; }
    int blocks() const noexcept { return blockNum; }

    int kernelLength() const noexcept { return 8 * 9 + 8 * 8 * 9 * blockNum + 8 * 4; }
    int kernelLength(const int layer) const noexcept { return (layer == 0) ? 8 * 9 : ((layer > 0 && layer < (blockNum + 1)) ? 8 * 8 * 9 : ((layer == (blockNum + 1)) ? 8 * 4 : 0)); }
    int biasLength() const noexcept { return 8 + 8 * blockNum; }
    int biasLength(const int layer) const noexcept { return (layer >= 0 && layer < (blockNum + 1)) ? 8 : 0; }
    std::size_t kernelSize() const noexcept { return kernelLength() * sizeof(float); }
    std::size_t kernelSize(const int layer) const noexcept { return kernelLength(layer) * sizeof(float); }
    std::size_t biasSize() const noexcept { return biasLength() * sizeof(float); }
    std::size_t biasSize(const int layer) const noexcept { return biasLength(layer) * sizeof(float); }

    int alphas() const noexcept { return 0; }   // no PReLU in this network
    int kernels() const noexcept { return blockNum + 2; }
    int biases() const noexcept { return blockNum + 1; }
    int kernelIndex(const int layer) const noexcept { return std::clamp(layer, 0, kernels() - 1); }
    int biasIndex(const int layer) const noexcept { return std::clamp(layer, 0, biases() - 1); }
    int kernelLayer(const int idx) const noexcept { return std::clamp(idx, 0, kernels() - 1); }
    int biasLayer(const int idx) const noexcept { return std::clamp(idx, 0, biases() - 1); }

    int kernelOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer <= 1) return 8 * 9;
        if (layer < kernels()) return 8 * 9 + (8 * 8 * 9) * (layer - 1);
        return kernelLength();
    }
    int biasOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer < biases()) return 8 * layer;
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

template<int F>
class ac::core::model::ACNet
{
public:
    enum class Variant
    {
        B4_NORMAL, B4_HDN, B4_BOX, B4_BOX_HDN,
        B8_NORMAL, B8_HDN, B8_BOX, B8_BOX_HDN,
        B18_NORMAL, B18_HDN, B18_BOX, B18_BOX_HDN
    };

public:
    AC_CORE_EXPORT explicit ACNet(Variant v) noexcept;

public:
    Variant variant() const noexcept { return v; }
    int blocks() const noexcept { return blockNum; }

    int kernelLength() const noexcept { return F * 9 + F * F * 9 * blockNum + F * 4 * 9; }
    int kernelLength(const int layer) const noexcept { return (layer == 0) ? F * 9 : ((layer > 0 && layer < (blockNum + 1)) ? F * F * 9 : ((layer == (blockNum + 1)) ? F * 4 * 9 : 0)); }
    int biasLength() const noexcept { return F + F * blockNum + 4; }
    int biasLength(const int layer) const noexcept { return (layer >= 0 && layer < (blockNum + 1)) ? F : (layer == (blockNum + 1) ? 4 : 0); }
    int alphaLength() const noexcept { return F * (blockNum + 1); }
    int alphaLength(const int layer) const noexcept { return (layer >= 0 && layer < (blockNum + 1)) ? F : 0; }
    std::size_t kernelSize() const noexcept { return kernelLength() * sizeof(float); }
    std::size_t biasSize() const noexcept { return biasLength() * sizeof(float); }
    std::size_t alphaSize() const noexcept { return alphaLength() * sizeof(float); }

    int kernels() const noexcept { return blockNum + 1 + 1; }
    int biases() const noexcept { return blockNum + 1 + 1; }
    int alphas() const noexcept { return blockNum + 1; }
    int kernelIndex(const int layer) const noexcept { return std::clamp(layer, 0, kernels() - 1); }
    int biasIndex(const int layer) const noexcept { return std::clamp(layer, 0, biases() - 1); }
    int alphaIndex(const int layer) const noexcept { return std::clamp(layer, 0, alphas() - 1); }

    int kernelOffset(const int layer) const noexcept
    {
        if (layer <= 0) return 0;
        if (layer <= 1) return F * 9;
        if (layer < kernels()) return F * 9 + (F * F * 9) * (layer - 1);
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
        if (layer <= 0) return 0;
        if (layer < alphas()) return F * layer;
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
