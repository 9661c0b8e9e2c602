// C++ checks of the image helpers (include/AC/Core/Image.hpp free functions), written like
// the reference's doctest suites tests/core/src/{ImageProcessTest,ImageResizeTest}.cpp and
// extended with numeric checks against the oracle's toFloat/fromFloat (oracle/liboracle.so).
// Built with g++ by tests/test_cpp_image_ops.py.  With argument "gpu" the helpers run on
// the device; without it, the test checks that they fail loudly and leave outputs alone.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "AC/Core.hpp"

extern "C" {
float orc_to_float(const void* p, int type);
void orc_from_float(float v, void* p, int type);
}

static int failures = 0, checks = 0;
#define CHECK(cond) do { checks++; if (!(cond)) { failures++; std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); } } while (0)
#define TEST_CASE(name) static void name()

using ac::core::Image;
static const int kTypes[] = { Image::UInt8, Image::UInt16, Image::Float16, Image::Float32 };

// random content; float images get values slightly outside [0, 1] to exercise the clamps
static void fill(Image& img, unsigned seed)
{
    std::mt19937 rng(seed);
    for (int y = 0; y < img.height(); y++)
        for (int i = 0; i < img.width() * img.channels(); i++)
        {
            std::uint8_t* p = img.line(y) + i * img.elementSize();
            if (img.type() == Image::UInt8) *p = static_cast<std::uint8_t>(rng());
            else if (img.type() == Image::UInt16) { std::uint16_t v = static_cast<std::uint16_t>(rng()); std::memcpy(p, &v, 2); }
            else
            {
                const float f = static_cast<float>(rng() % 12001) / 10000.0f - 0.1f;
                orc_from_float(f, p, img.type() == Image::Float32 ? 0x204 : 0x202);   // fromFloat clamps; store raw for f32
                if (img.type() == Image::Float32) std::memcpy(p, &f, 4);
                else if (rng() % 8 == 0) { const std::uint16_t v = 0x3c66; std::memcpy(p, &v, 2); }   // 1.1 in f16
            }
        }
}

static bool sameBytes(const Image& a, const Image& b)
{
    if (a.width() != b.width() || a.height() != b.height() || a.pixelSize() != b.pixelSize()) return false;
    for (int y = 0; y < a.height(); y++)
        if (std::memcmp(a.line(y), b.line(y), static_cast<size_t>(a.width()) * a.pixelSize())) return false;
    return true;
}

// fromFloat<dst>(toFloat(src)) element by element, through the oracle
static bool convertedFrom(const Image& src, const Image& dst)
{
    for (int y = 0; y < src.height(); y++)
        for (int i = 0; i < src.width() * src.channels(); i++)
        {
            std::uint8_t want[4] = {};
            orc_from_float(orc_to_float(src.line(y) + i * src.elementSize(), src.type()), want, dst.type());
            if (std::memcmp(want, dst.line(y) + i * dst.elementSize(), dst.elementSize())) return false;
        }
    return true;
}

// ---- ImageProcessTest.cpp ----

TEST_CASE(rgb2yuv_dimensions)   // :7-50
{
    Image rgb{ 4, 4, 3, Image::UInt8 };
    std::memset(rgb.ptr(), 100, rgb.size());
    Image yuv{};
    ac::core::rgb2yuv(rgb, yuv);
    CHECK(yuv.width() == 4 && yuv.height() == 4 && yuv.channels() == 3 && yuv.type() == Image::UInt8);
    Image y{}, uv{};
    ac::core::rgb2yuv(rgb, y, uv);
    CHECK(y.channels() == 1 && uv.channels() == 2 && uv.width() == 4 && uv.height() == 4);
    Image y3{}, u{}, v{};
    ac::core::rgb2yuv(rgb, y3, u, v);
    CHECK(y3.channels() == 1 && u.channels() == 1 && v.channels() == 1);
}

TEST_CASE(gray_round_trips)   // :52-137: gray 100 comes back within +-1 in all three forms
{
    for (int form = 0; form < 3; form++)
    {
        Image rgb{ 4, 4, 3, Image::UInt8 };
        std::memset(rgb.ptr(), 100, rgb.size());
        Image back{};
        if (form == 0) { Image yuv{}; ac::core::rgb2yuv(rgb, yuv); ac::core::yuv2rgb(yuv, back); }
        if (form == 1) { Image y{}, uv{}; ac::core::rgb2yuv(rgb, y, uv); ac::core::yuv2rgb(y, uv, back); }
        if (form == 2) { Image y{}, u{}, v{}; ac::core::rgb2yuv(rgb, y, u, v); ac::core::yuv2rgb(y, u, v, back); }
        bool ok = !back.empty();
        for (int i = 0; ok && i < 4; i++)
            for (int j = 0; j < 4; j++)
                for (int c = 0; c < 3; c++) ok = ok && back.pixel(j, i)[c] >= 99 && back.pixel(j, i)[c] <= 101;
        CHECK(ok);
    }
}

TEST_CASE(shl_shr)   // :139-177
{
    Image img{ 2, 1, 1, Image::UInt8 };
    img.pixel(0, 0)[0] = 0x0f;
    img.pixel(1, 0)[0] = 0x0f;
    ac::core::shl(img, 4);
    CHECK(img.pixel(0, 0)[0] == 0xf0 && img.pixel(1, 0)[0] == 0xf0);
    ac::core::shr(img, 4);
    CHECK(img.pixel(0, 0)[0] == 0x0f && img.pixel(1, 0)[0] == 0x0f);
    Image src{ 1, 1, 1, Image::UInt8 }, dst{};
    src.pixel(0, 0)[0] = 0xff;
    ac::core::shl(src, dst, 6);
    CHECK(!dst.empty() && dst.pixel(0, 0)[0] == 0xc0 && src.pixel(0, 0)[0] == 0xff);
    Image dst2{};
    ac::core::shr(src, dst2, 6);
    CHECK(!dst2.empty() && dst2.pixel(0, 0)[0] == 0x03);
    // src == dst makes a new image (the CLI shifts a decoded frame it must not modify)
    Image frame{ 3, 2, 1, Image::UInt16 };
    fill(frame, 5);
    Image alias = frame;
    ac::core::shl(alias, alias, 2);
    CHECK(alias != frame && !alias.empty());
    bool ok = true;
    for (int y = 0; y < 2; y++)
        for (int x = 0; x < 3; x++)
        {
            std::uint16_t a, b;
            std::memcpy(&a, frame.pixel(x, y), 2);
            std::memcpy(&b, alias.pixel(x, y), 2);
            ok = ok && b == static_cast<std::uint16_t>(a << 2);
        }
    CHECK(ok);
    Image f{ 2, 2, 1, Image::Float32 };
    fill(f, 6);
    Image fcopy = f.clone();
    ac::core::shl(f, 3);   // float images are left alone
    CHECK(sameBytes(f, fcopy));
}

TEST_CASE(astype_and_copy)   // :179-234
{
    Image src{ 2, 2, 1, Image::UInt8 };
    src.pixel(0, 0)[0] = 0;
    src.pixel(1, 0)[0] = 255;
    src.pixel(0, 1)[0] = 128;
    src.pixel(1, 1)[0] = 64;
    auto dst = ac::core::astype(src, Image::Float32);
    CHECK(dst.type() == Image::Float32 && dst.width() == 2 && dst.height() == 2 && dst.channels() == 1);
    if (!dst.empty())
    {
        CHECK(*static_cast<const float*>(dst.ptr(0, 0)) == 0.0f);
        CHECK(*static_cast<const float*>(dst.ptr(1, 0)) == 1.0f);
        CHECK(*static_cast<const float*>(dst.ptr(0, 1)) == 128.0f / 255.0f);
        CHECK(*static_cast<const float*>(dst.ptr(1, 1)) == 64.0f / 255.0f);
    }
    CHECK(ac::core::astype(src, Image::UInt8) == src);
    Image s{ 4, 4, 1, Image::UInt8 };
    fill(s, 9);
    Image d{};
    ac::core::copy(s, d);
    CHECK(sameBytes(s, d) && d != s);
    Image f{ 2, 1, 1, Image::Float32 };
    Image c8{ 2, 1, 1, Image::UInt8 };
    c8.pixel(0, 0)[0] = 128;
    c8.pixel(1, 0)[0] = 128;
    ac::core::copy(c8, f);
    CHECK(f.type() == Image::Float32 && *static_cast<const float*>(f.ptr(0, 0)) == 128.0f / 255.0f);
}

TEST_CASE(every_conversion_matches_oracle)
{
    for (int a : kTypes)
        for (int b : kTypes)
        {
            if (a == b) continue;
            Image src{ 37, 5, 3, a };
            fill(src, static_cast<unsigned>(a * 7 + b));
            Image dst = ac::core::astype(src, b);
            CHECK(!dst.empty() && dst.type() == b && convertedFrom(src, dst));
            Image pre{ 37, 5, 3, b };   // copy into a preallocated image of another type
            ac::core::copy(src, pre);
            CHECK(convertedFrom(src, pre));
        }
}

TEST_CASE(crop_extract_insert)   // :236-294
{
    Image src{ 4, 4, 1, Image::UInt8 };
    auto c = ac::core::crop(src, 1, 1, 2, 2);
    CHECK(c.width() == 2 && c.height() == 2 && c.channels() == 1 && c == src);
    auto n = ac::core::crop(src, 3, 3, -2, -2);
    CHECK(n.width() == 2 && n.height() == 2);
    CHECK(ac::core::crop(Image{}, 0, 0, 1, 1).empty());

    Image rgba{ 5, 3, 4, Image::UInt16 };
    fill(rgba, 21);
    auto e = ac::core::extract(rgba, 1, 2);
    CHECK(e.width() == 5 && e.height() == 3 && e.channels() == 2 && e.type() == Image::UInt16);
    bool ok = !e.empty();
    for (int y = 0; ok && y < 3; y++)
        for (int x = 0; x < 5; x++) ok = ok && !std::memcmp(e.pixel(x, y), rgba.pixel(x, y) + 2, 4);
    CHECK(ok);
    CHECK(ac::core::extract(rgba, 4, 1).empty() && ac::core::extract(rgba, 0, 0).empty());
    CHECK(ac::core::extract(rgba, 0, 4) == rgba);

    Image rgb{ 2, 2, 2, Image::UInt8 }, alpha{ 2, 2, 1, Image::UInt8 };
    fill(rgb, 3);
    fill(alpha, 4);
    auto r = ac::core::insert(rgb, alpha, 1);
    CHECK(r.width() == 2 && r.height() == 2 && r.channels() == 3 && r.type() == Image::UInt8);
    ok = !r.empty();
    for (int y = 0; ok && y < 2; y++)
        for (int x = 0; x < 2; x++)
            ok = ok && r.pixel(x, y)[0] == rgb.pixel(x, y)[0] && r.pixel(x, y)[1] == alpha.pixel(x, y)[0] &&
                 r.pixel(x, y)[2] == rgb.pixel(x, y)[1];
    CHECK(ok);
    CHECK(ac::core::insert(rgb, alpha, 3) == rgb);   // channel beyond the end: src back
}

TEST_CASE(pixel_shuffle)   // :296-314, and the channel order of the shuffle
{
    Image src{ 4, 4, 16, Image::Float32 }, dst{};
    fill(src, 8);
    ac::core::pixelShuffle(src, dst, 2);
    CHECK(dst.width() == 8 && dst.height() == 8 && dst.channels() == 4);
    Image bad{ 4, 4, 3, Image::UInt8 }, none{};
    ac::core::pixelShuffle(bad, none, 2);
    CHECK(none.empty());

    // output (2x + dx, 2y + dy), channel n = input channel 4n + 2dy + dx
    Image one{ 1, 1, 4, Image::UInt8 }, o{};
    for (int i = 0; i < 4; i++) one.pixel(0, 0)[i] = static_cast<std::uint8_t>(i + 1);
    ac::core::pixelShuffle(one, o, 2);
    CHECK(!o.empty() && o.pixel(0, 0)[0] == 1 && o.pixel(1, 0)[0] == 2 && o.pixel(0, 1)[0] == 3 && o.pixel(1, 1)[0] == 4);

    // every element through toFloat/fromFloat unless both are one integer type
    bool ok = !dst.empty();
    for (int Y = 0; ok && Y < 8; Y++)
        for (int X = 0; X < 8; X++)
            for (int n = 0; n < 4; n++)
            {
                const float* in = static_cast<const float*>(src.ptr(X / 2, Y / 2));
                float want;
                orc_from_float(in[n * 4 + (Y % 2) * 2 + X % 2], &want, 0x204);
                ok = ok && static_cast<const float*>(dst.ptr(X, Y))[n] == want;
            }
    CHECK(ok);
    Image u{ 3, 2, 4, Image::UInt16 }, uf{ 6, 4, 1, Image::Float32 };
    fill(u, 12);
    ac::core::pixelShuffle(u, uf, 2);   // preallocated f32 output: converted
    ok = uf.type() == Image::Float32;
    for (int Y = 0; ok && Y < 4; Y++)
        for (int X = 0; X < 6; X++)
        {
            const auto* in = reinterpret_cast<const std::uint16_t*>(u.ptr(X / 2, Y / 2));
            ok = ok && *static_cast<const float*>(uf.ptr(X, Y)) == orc_to_float(&in[(Y % 2) * 2 + X % 2], 0x002);
        }
    CHECK(ok);
}

TEST_CASE(unpadding)   // :316-336
{
    Image src{ 4, 4, 1, Image::UInt8 };
    CHECK(ac::core::unpadding(src) == src);
    std::vector<std::uint8_t> buf(4 * 16);
    for (size_t i = 0; i < buf.size(); i++) buf[i] = static_cast<std::uint8_t>(i * 7);
    Image padded{ 4, 4, 1, Image::UInt8, buf.data(), 16 };
    auto u = ac::core::unpadding(padded);
    CHECK(u.width() == 4 && u.height() == 4 && u.channels() == 1 && u.stride() == 4 && sameBytes(u, padded));
}

TEST_CASE(colour_forms_agree)
{
    // the packed, two-plane and per-channel forms give the same components, for every type,
    // RGB and RGBA, both directions, and in place
    for (int t : kTypes)
    {
        Image rgb{ 19, 7, 3, t };
        fill(rgb, static_cast<unsigned>(t));
        Image yuv{}, y{}, uv{}, y3{}, u{}, v{};
        ac::core::rgb2yuv(rgb, yuv);
        ac::core::rgb2yuv(rgb, y, uv);
        ac::core::rgb2yuv(rgb, y3, u, v);
        CHECK(sameBytes(ac::core::extract(yuv, 0, 1), y) && sameBytes(ac::core::extract(yuv, 1, 2), uv));
        CHECK(sameBytes(y3, y) && sameBytes(ac::core::extract(yuv, 1, 1), u) && sameBytes(ac::core::extract(yuv, 2, 1), v));
        Image b1{}, b2{}, b3{};
        ac::core::yuv2rgb(yuv, b1);
        ac::core::yuv2rgb(y, uv, b2);
        ac::core::yuv2rgb(y3, u, v, b3);
        CHECK(!b1.empty() && sameBytes(b1, b2) && sameBytes(b1, b3));
        Image inplace = rgb.clone();
        ac::core::rgb2yuv(inplace, inplace);
        CHECK(sameBytes(inplace, yuv));

        Image rgba{ 19, 7, 4, t };
        fill(rgba, static_cast<unsigned>(t) + 1);
        Image yuva{}, ya{}, uva{}, y4{}, u4{}, v4{}, a4{};
        ac::core::rgba2yuva(rgba, yuva);
        ac::core::rgba2yuva(rgba, ya, uva);
        ac::core::rgba2yuva(rgba, y4, u4, v4, a4);
        CHECK(sameBytes(ac::core::extract(yuva, 0, 1), ya) && sameBytes(ac::core::extract(yuva, 1, 3), uva));
        CHECK(sameBytes(y4, ya) && sameBytes(ac::core::extract(yuva, 3, 1), a4) && sameBytes(ac::core::extract(yuva, 1, 1), u4));
        Image c1{}, c2{}, c3{};
        ac::core::yuva2rgba(yuva, c1);
        ac::core::yuva2rgba(ya, uva, c2);
        ac::core::yuva2rgba(y4, u4, v4, a4, c3);
        CHECK(!c1.empty() && sameBytes(c1, c2) && sameBytes(c1, c3));
        // an RGBA image given to rgb2yuv: alpha ignored, not premultiplied
        Image yuvFromRgba{};
        ac::core::rgb2yuv(rgba, yuvFromRgba);
        Image rgbPart = ac::core::extract(rgba, 0, 3), yuvRef{};
        ac::core::rgb2yuv(rgbPart, yuvRef);
        CHECK(yuvFromRgba.channels() == 3 && sameBytes(yuvFromRgba, yuvRef));
    }
}

// ---- ImageResizeTest.cpp ----

TEST_CASE(resize_cases)
{
    Image src{ 4, 4, 3, Image::UInt8 };
    fill(src, 1);
    CHECK(ac::core::resize(src, 1.0, 1.0) == src);   // :5-10
    auto d2 = ac::core::resize(src, 2.0, 2.0);       // :12-22
    CHECK(d2.width() == 8 && d2.height() == 8 && d2.channels() == 3 && d2.type() == Image::UInt8);
    auto dh = ac::core::resize(src, 0.5, 0.5);       // :24-33
    CHECK(dh.width() == 2 && dh.height() == 2 && dh.channels() == 3);
    auto dxy = ac::core::resize(src, 2.0, 1.0);      // :35-42
    CHECK(dxy.width() == 8 && dxy.height() == 4);

    Image p{ 2, 2, 1, Image::UInt8 };                // :44-79
    p.pixel(0, 0)[0] = 10; p.pixel(1, 0)[0] = 20; p.pixel(0, 1)[0] = 30; p.pixel(1, 1)[0] = 40;
    auto q = ac::core::resize(p, 2.0, 2.0, ac::core::RESIZE_POINT);
    const std::uint8_t want[4][4] = { { 10, 10, 20, 20 }, { 10, 10, 20, 20 }, { 30, 30, 40, 40 }, { 30, 30, 40, 40 } };
    bool ok = q.width() == 4 && q.height() == 4;
    for (int y = 0; ok && y < 4; y++)
        for (int x = 0; x < 4; x++) ok = ok && q.pixel(x, y)[0] == want[y][x];
    CHECK(ok);

    Image pre{ 8, 8, 3, Image::UInt8 };              // :81-89
    ac::core::resize(src, pre, 0.0, 0.0);
    CHECK(pre.width() == 8 && pre.height() == 8);
    CHECK(sameBytes(pre, d2));                       // same default mode (Catmull-Rom) either way

    for (int mode = ac::core::RESIZE_CATMULL_ROM; mode <= ac::core::RESIZE_BILINEAR; mode++)   // :91-113
    {
        auto m = ac::core::resize(src, 2.0, 2.0, mode);
        CHECK(m.width() == 8 && m.height() == 8);
    }
    CHECK(ac::core::resize(src, 0.0, 2.0) == src);   // :115-123
    CHECK(ac::core::resize(src, -1.0, -1.0) == src);

    Image inplace = src.clone(), keep = inplace;     // src == dst resizes through a temporary
    ac::core::resize(inplace, inplace, 2.0, 2.0);
    CHECK(inplace.width() == 8 && keep.width() == 4);
}

// ---- without a device: every helper fails loudly and leaves its output alone ----

TEST_CASE(no_device)
{
    Image rgb{ 4, 4, 3, Image::UInt8 };
    fill(rgb, 2);
    Image yuv{ 4, 4, 3, Image::UInt8 };
    std::memset(yuv.ptr(), 7, yuv.size());
    ac::core::rgb2yuv(rgb, yuv);
    CHECK(yuv.pixel(0, 0)[0] == 7 && yuv.pixel(3, 3)[2] == 7);
    CHECK(ac::core::resize(rgb, 2.0, 2.0).empty());
    Image pre{ 8, 8, 3, Image::UInt8 }, keep = pre;
    ac::core::resize(rgb, pre, 0.0, 0.0);
    CHECK(pre == keep);
    Image fresh{};
    ac::core::rgb2yuv(rgb, fresh);                    // allocated for the call, then put back
    CHECK(fresh.empty());
    CHECK(ac::core::astype(rgb, Image::Float32).empty());
    CHECK(ac::core::extract(rgb, 1, 1).empty());
    Image img{ 2, 1, 1, Image::UInt8 };
    img.pixel(0, 0)[0] = 1;
    ac::core::shl(img, 1);
    CHECK(img.pixel(0, 0)[0] == 1);
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
    if (gpu)
    {
        rgb2yuv_dimensions();
        gray_round_trips();
        shl_shr();
        astype_and_copy();
        every_conversion_matches_oracle();
        crop_extract_insert();
        pixel_shuffle();
        unpadding();
        colour_forms_agree();
        resize_cases();
    }
    else no_device();
    std::printf("%d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
