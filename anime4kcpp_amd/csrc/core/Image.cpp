// ac::core::Image (include/AC/Core/Image.hpp).  Behaviour follows core/src/Image.cpp:13-110:
// create() owns a buffer whose row stride is the given stride if it covers the row, else
// the row size rounded up to 4 bytes; map() wraps caller memory; from() copies; to()
// copies out; view() shares the buffer of the intersection; clone() deep-copies.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "AC/Core/Image.hpp"

namespace {
constexpr int kStrideAlign = 4;
constexpr std::size_t kBufferAlign = 64;
inline int alignUp(int v, int a) { return (v + a - 1) / a * a; }
}

struct ac::core::Image::ImageData
{
    void* data;
    explicit ImageData(const std::size_t size) noexcept
        : data(std::aligned_alloc(kBufferAlign, (size + kBufferAlign - 1) / kBufferAlign * kBufferAlign)) {}
    ~ImageData() noexcept { std::free(data); }
};

ac::core::Image::Image() noexcept : w(0), h(0), c(0), elementType(UInt8), pitch(0), pixels(nullptr), dptr(nullptr) {}
ac::core::Image::Image(const int w, const int h, const int c, const ElementType elementType, const int stride) : Image()
{
    create(w, h, c, elementType, stride);
}
ac::core::Image::Image(const int w, const int h, const int c, const ElementType elementType, void* const data, const int stride) : Image()
{
    if (data) map(w, h, c, elementType, data, stride);
    else create(w, h, c, elementType, stride);
}
ac::core::Image::Image(const Image&) noexcept = default;
ac::core::Image::Image(Image&&) noexcept = default;
ac::core::Image::~Image() noexcept = default;
ac::core::Image& ac::core::Image::operator=(const Image&) noexcept = default;
ac::core::Image& ac::core::Image::operator=(Image&&) noexcept = default;

void ac::core::Image::create(const int w, const int h, const int c, const ElementType elementType, const int stride)
{
    const int lineSize = w * c * (elementType & 0xff);
    if (!(h > 0 && lineSize > 0)) return;
    const int p = (stride >= lineSize) ? stride : alignUp(lineSize, kStrideAlign);
    auto data = std::make_shared<ImageData>(static_cast<std::size_t>(h) * p);
    if (!data->data) return;
    this->w = w;
    this->h = h;
    this->c = c;
    this->elementType = elementType;
    this->pitch = p;
    this->dptr = data;
    this->pixels = data->data;
}
void ac::core::Image::map(const int w, const int h, const int c, const ElementType elementType, void* const data, const int stride) noexcept
{
    const int lineSize = w * c * (elementType & 0xff);
    if (!(h > 0 && lineSize > 0) || !data) return;
    this->w = w;
    this->h = h;
    this->c = c;
    this->elementType = elementType;
    this->pitch = (stride >= lineSize) ? stride : lineSize;
    this->dptr = nullptr;
    this->pixels = data;
}
void ac::core::Image::from(const int w, const int h, const int c, const ElementType elementType, const void* const data, const int stride)
{
    const int lineSize = w * c * (elementType & 0xff);
    if (!(h > 0 && lineSize > 0) || !data) return;
    const int srcPitch = (stride >= lineSize) ? stride : lineSize;
    create(w, h, c, elementType);
    if (empty()) return;
    auto s = static_cast<const std::uint8_t*>(data);
    for (int i = 0; i < h; i++) std::memcpy(line(i), s + static_cast<long>(i) * srcPitch, lineSize);
}
void ac::core::Image::to(void* const data, const int stride) const noexcept
{
    const int lineSize = width() * pixelSize();
    if (!(height() > 0 && lineSize > 0) || !data) return;
    const int dstPitch = (stride >= lineSize) ? stride : lineSize;
    auto d = static_cast<std::uint8_t*>(data);
    for (int i = 0; i < height(); i++) std::memcpy(d + static_cast<long>(i) * dstPitch, line(i), lineSize);
}
ac::core::Image ac::core::Image::view(const int x, const int y, const int w, const int h) const noexcept
{
    Image v{};
    const int ix = std::max(x, 0), iy = std::max(y, 0);
    const int iw = std::min(width(), x + w) - ix, ih = std::min(height(), y + h) - iy;
    if (iw > 0 && ih > 0) {
        v.w = iw;
        v.h = ih;
        v.c = channels();
        v.elementType = type();
        v.pitch = stride();
        v.pixels = ptr(ix, iy);
        v.dptr = dptr;
    }
    return v;
}
ac::core::Image ac::core::Image::clone() const
{
    Image d{};
    d.from(width(), height(), channels(), type(), ptr(), stride());
    return d;
}
