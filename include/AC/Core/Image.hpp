// ac::core::Image -- strided, reference-counted 2-D image, interface-compatible with
// core/include/AC/Core/Image.hpp:359-419 (same members, element-type encoding and
// ownership rules: create() owns and aligns the row stride to 4 bytes, map() wraps
// foreign memory without copying, view() shares the buffer).
#ifndef AC_CORE_IMAGE_HPP
#define AC_CORE_IMAGE_HPP

#include <cstdint>
#include <memory>

#include "ACCoreExport.hpp"

namespace ac::core
{
    class Image;
}

class ac::core::Image
{
private:
    struct ImageData;

public:
    /// Format: 2 (float) or 1 (int) or 0 (uint) << 8 | sizeof(type)
    using ElementType = int;
    static constexpr ElementType UInt8   = 0 << 8 | 1;
    static constexpr ElementType UInt16  = 0 << 8 | 2;
    static constexpr ElementType Float16 = 2 << 8 | 2;
    static constexpr ElementType Float32 = 2 << 8 | 4;

public:
    AC_CORE_EXPORT Image() noexcept;
    AC_CORE_EXPORT Image(int w, int h, int c, ElementType elementType, int stride = 0);
    AC_CORE_EXPORT Image(int w, int h, int c, ElementType elementType, void* data, int stride = 0);
    AC_CORE_EXPORT Image(const Image&) noexcept;
    AC_CORE_EXPORT Image(Image&&) noexcept;
    AC_CORE_EXPORT ~Image() noexcept;
    AC_CORE_EXPORT Image& operator=(const Image&) noexcept;
    AC_CORE_EXPORT Image& operator=(Image&&) noexcept;

    AC_CORE_EXPORT void create(int w, int h, int c, ElementType elementType, int stride = 0);
    AC_CORE_EXPORT void map(int w, int h, int c, ElementType elementType, void* data, int stride = 0) noexcept;
    AC_CORE_EXPORT void from(int w, int h, int c, ElementType elementType, const void* data, int stride = 0);
    AC_CORE_EXPORT void to(void* data, int stride = 0) const noexcept;
    AC_CORE_EXPORT Image view(int x, int y, int w, int h) const noexcept;
    AC_CORE_EXPORT Image clone() const;

public:
    int width() const noexcept { return w; }
    int height() const noexcept { return h; }
    int channels() const noexcept { return c; }
    int stride() const noexcept { return pitch; }
    int size() const noexcept { return h * pitch; }
    int elementSize() const noexcept { return elementType & 0xff; }
    int pixelSize() const noexcept { return c * elementSize(); }
    ElementType type() const noexcept { return elementType; }
    std::uint8_t* data() const noexcept { return static_cast<std::uint8_t*>(pixels); }
    std::uint8_t* line(const int y) const noexcept { return data() + static_cast<long>(y) * pitch; }
    std::uint8_t* pixel(const int x, const int y) const noexcept { return line(y) + x * pixelSize(); }
    void* ptr() const noexcept { return pixels; }
    void* ptr(const int y) const noexcept { return line(y); }
    void* ptr(const int x, const int y) const noexcept { return pixel(x, y); }
    bool empty() const noexcept { return pixels == nullptr; }
    bool isUint() const noexcept { return (elementType >> 8) == 0; }
    bool isInt() const noexcept { return (elementType >> 8) == 1; }
    bool isFloat() const noexcept { return (elementType >> 8) == 2; }
    bool ownership() const noexcept { return dptr != nullptr; }
    /// Whether the images share the same buffer.
    bool operator==(const Image& other) const noexcept { return (ownership() && other.ownership()) ? (dptr == other.dptr) : (pixels == other.pixels); }
    bool operator!=(const Image& other) const noexcept { return !operator==(other); }

private:
    int w, h, c;
    ElementType elementType;
    int pitch;
    void* pixels;
    std::shared_ptr<ImageData> dptr;
};

#endif
