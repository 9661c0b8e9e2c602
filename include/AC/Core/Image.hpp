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

// Image helpers, the same declarations as core/include/AC/Core/Image.hpp:14-319 (line of
// each given below).  Every pixel operation runs on the calling thread's current HIP
// device: the helper uploads its host images, runs one kernel and downloads the result
// (crop only builds a view).  Output images that are not empty must already have the shape
// and element type the helper writes.  On a device failure the output is left untouched
// and the reason is printed to stderr.  Image file I/O (imread, imwrite, imdecode) is not
// part of this library.
namespace ac::core
{
    AC_CORE_EXPORT void rgb2yuv(const Image& rgb, Image& yuv);                                  // :21
    AC_CORE_EXPORT void rgb2yuv(const Image& rgb, Image& y, Image& uv);                         // :31
    AC_CORE_EXPORT void rgb2yuv(const Image& rgb, Image& y, Image& u, Image& v);                // :42
    AC_CORE_EXPORT void rgba2yuva(const Image& rgba, Image& yuva);                              // :52
    AC_CORE_EXPORT void rgba2yuva(const Image& rgba, Image& y, Image& uva);                     // :62
    AC_CORE_EXPORT void rgba2yuva(const Image& rgba, Image& y, Image& u, Image& v, Image& a);   // :74
    AC_CORE_EXPORT void yuv2rgb(const Image& yuv, Image& rgb);                                  // :84
    AC_CORE_EXPORT void yuv2rgb(const Image& y, const Image& uv, Image& rgb);                   // :94
    AC_CORE_EXPORT void yuv2rgb(const Image& y, const Image& u, const Image& v, Image& rgb);    // :105
    AC_CORE_EXPORT void yuva2rgba(const Image& yuva, Image& rgba);                              // :115
    AC_CORE_EXPORT void yuva2rgba(const Image& y, const Image& uva, Image& rgba);               // :125
    AC_CORE_EXPORT void yuva2rgba(const Image& y, const Image& u, const Image& v, const Image& a, Image& rgba); // :137

    AC_CORE_EXPORT Image unpadding(const Image& src) noexcept;                                  // :144
    AC_CORE_EXPORT void shl(Image& image, int n) noexcept;                                      // :156
    AC_CORE_EXPORT void shl(const Image& src, Image& dst, int n) noexcept;                      // :168
    AC_CORE_EXPORT void shr(Image& image, int n) noexcept;                                      // :179
    AC_CORE_EXPORT void shr(const Image& src, Image& dst, int n) noexcept;                      // :191
    AC_CORE_EXPORT Image astype(const Image& src, int type) noexcept;                           // :202
    AC_CORE_EXPORT void copy(const Image& src, Image& dst) noexcept;                            // :217
    AC_CORE_EXPORT Image crop(const Image& src, int x, int y, int w, int h) noexcept;           // :228
    AC_CORE_EXPORT Image extract(const Image& src, int channel, int n) noexcept;                // :240
    AC_CORE_EXPORT Image insert(const Image& src, const Image& image, int channel) noexcept;    // :251
    AC_CORE_EXPORT void pixelShuffle(const Image& src, Image& dst, int upscale) noexcept;       // :269

    /// Resize interpolation modes (:272-291), same names and values.
    enum ResizeModes
    {
        RESIZE_POINT,               ///< nearest when upscaling, box when downscaling
        RESIZE_CATMULL_ROM,         ///< cubic b = 0, c = 1/2
        RESIZE_MITCHELL_NETRAVALI,  ///< cubic b = 1/3, c = 1/3
        RESIZE_BICUBIC_0_60,        ///< cubic b = 0, c = 0.6
        RESIZE_BICUBIC_0_75,        ///< cubic b = 0, c = 0.75
        RESIZE_BICUBIC_0_100,       ///< cubic b = 0, c = 1
        RESIZE_BICUBIC_20_50,       ///< cubic b = 0.2, c = 0.5
        RESIZE_SOFTCUBIC50,         ///< cubic b = 1/2, c = 1/2
        RESIZE_SOFTCUBIC75,         ///< cubic b = 3/4, c = 1/4
        RESIZE_SOFTCUBIC100,        ///< cubic b = 1, c = 0 (cubic B-spline)
        RESIZE_LANCZOS2,
        RESIZE_LANCZOS3,
        RESIZE_LANCZOS4,
        RESIZE_SPLINE16,
        RESIZE_SPLINE36,
        RESIZE_SPLINE64,
        RESIZE_BILINEAR,
    };
    /// fx, fy > 0: dst becomes (int)(w * fx) x (int)(h * fy); otherwise dst's own size is the
    /// target.  Downscales below about 1/47 (support-2 filters) are refused.
    AC_CORE_EXPORT void resize(const Image& src, Image& dst, double fx, double fy, int mode = RESIZE_CATMULL_ROM) noexcept; // :307
    AC_CORE_EXPORT Image resize(const Image& src, double fx, double fy, int mode = RESIZE_CATMULL_ROM) noexcept;            // :319
}

#endif
