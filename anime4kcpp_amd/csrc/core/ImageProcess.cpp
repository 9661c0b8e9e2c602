// The ac::core image helpers (core/include/AC/Core/Image.hpp:14-319; behaviour of
// core/src/ImageProcess.cpp and core/src/ImageResize.cpp) with every pixel operation on the
// HIP device: the helper uploads its host images, runs one of the kernels of
// acnet_color.hip / acnet_image.hip and downloads the result (pageable host memory through
// two page-locked chunks of the engine).  crop() only builds a view.
//
// Each call runs on the calling thread's current HIP device, on an engine (a stream and
// grow-only device buffers) leased from a process-wide pool for the duration of the call,
// so concurrent callers never share a stream.  The reference's helpers cannot fail except
// by allocation; these can (no device, a device error), and then they leave their output
// untouched, print the reason to stderr and record it for the C binding (detail::takeImageError).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "AC/Core/Image.hpp"
#include "acnet_kernels.hpp"
#include "core_internal.hpp"
#include "host_staging.hpp"

namespace ac::core::detail::ops
{
    constexpr int kSlots = 6;
    constexpr size_t kPitchAlign = 256;

    struct Engine
    {
        int device = -1;
        hipStream_t stream = nullptr;
        void* buf[kSlots] = {};
        size_t bytes[kSlots] = {};
        HostStaging staging;   // page-locked chunks for pageable images
    };

    // Engines live until the process ends (no teardown against the HIP runtime's own exit).
    struct Pool
    {
        std::mutex m;
        std::vector<Engine*> idle;
    };
    Pool& pool()
    {
        static Pool* p = new Pool;
        return *p;
    }

    thread_local std::string lastError;

    // One device image: w x h pixels of c channels of `type`, rows `pitch` bytes apart.
    struct Dev
    {
        void* p = nullptr;
        long pitch = 0;
        int w = 0, h = 0, c = 0, type = 0;
        int es() const { return type & 0xff; }
        long line() const { return static_cast<long>(w) * c * es(); }
    };

    class Session
    {
    public:
        explicit Session(const char* fn) : fn(fn)
        {
            int n = 0;
            if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { fail("no HIP device"); return; }
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess) { fail("hipGetDevice failed"); return; }
            {
                std::lock_guard<std::mutex> lk(pool().m);
                auto& idle = pool().idle;
                for (size_t i = 0; i < idle.size(); i++)
                    if (idle[i]->device == dev) { e = idle[i]; idle.erase(idle.begin() + static_cast<long>(i)); break; }
            }
            if (!e) { e = new Engine; e->device = dev; }
            if (!e->stream && !check(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate")) return;
        }
        ~Session()
        {
            if (e)
            {
                if (e->stream) (void)hipStreamSynchronize(e->stream);
                std::lock_guard<std::mutex> lk(pool().m);
                pool().idle.push_back(e);
            }
            if (!failed) lastError.clear();
        }
        bool ok() const { return !failed; }
        hipStream_t stream() const { return e->stream; }

        bool fail(const std::string& msg)
        {
            if (!failed)
            {
                failed = true;
                lastError = std::string("ac::core::") + fn + ": " + msg;
                std::fprintf(stderr, "%s\n", lastError.c_str());
            }
            return false;
        }
        bool check(hipError_t err, const char* what)
        {
            return err == hipSuccess ? true : fail(std::string(what) + ": " + hipGetErrorString(err));
        }

        Dev alloc(const int w, const int h, const int c, const int type, const int slot)
        {
            Dev d;
            if (failed) return d;
            d.w = w; d.h = h; d.c = c; d.type = type;
            d.pitch = static_cast<long>((static_cast<size_t>(d.line()) + kPitchAlign - 1) / kPitchAlign * kPitchAlign);
            const size_t need = static_cast<size_t>(d.pitch) * static_cast<size_t>(h);
            if (e->bytes[slot] < need)
            {
                if (e->buf[slot])
                {
                    (void)hipStreamSynchronize(e->stream);
                    (void)hipFree(e->buf[slot]);
                    e->buf[slot] = nullptr;
                    e->bytes[slot] = 0;
                }
                if (!check(hipMalloc(&e->buf[slot], need), "hipMalloc")) { e->buf[slot] = nullptr; return Dev{}; }
                e->bytes[slot] = need;
            }
            d.p = e->buf[slot];
            return d;
        }
        Dev upload(const Image& img, const int slot)
        {
            Dev d = alloc(img.width(), img.height(), img.channels(), img.type(), slot);
            if (failed) return d;
            if (!e->staging.toDevice(d.p, d.pitch, img.ptr(), img.stride(), d.line(), d.h, e->stream)) fail(e->staging.error());
            return d;
        }
        bool download(const Dev& d, Image& img)
        {
            if (failed) return false;
            return e->staging.toHost(img.ptr(), img.stride(), d.p, d.pitch, d.line(), d.h, e->stream) ? true : fail(e->staging.error());
        }
        bool finish()
        {
            if (failed) return false;
            return check(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
        }

    private:
        const char* fn;
        Engine* e = nullptr;
        bool failed = false;
    };

    bool validType(const int t)
    {
        return t == Image::UInt8 || t == Image::UInt16 || t == Image::Float16 || t == Image::Float32;
    }

    // The colour helpers: `in` images supply the kind's components in order (the last one
    // may hold more channels than it supplies, as an RGBA image given to rgb2yuv does);
    // `out` images take them, each with exactly its channel count.
    bool colour(const char* fn, const int kind, std::initializer_list<const Image*> in, std::initializer_list<Image*> out)
    {
        Session s(fn);
        if (!s.ok()) return false;
        const int comps = (kind == achip::kRgba2Yuva || kind == achip::kYuva2Rgba) ? 4 : 3;
        const Image& first = **in.begin();
        const int w = first.width(), h = first.height(), t = first.type();
        if (!validType(t)) return s.fail("unsupported element type");
        for (auto p : in)
            if (p->empty() || p->width() != w || p->height() != h || p->type() != t) return s.fail("input planes differ in size or type");
        int have = 0;
        for (auto p : in) have += p->channels();
        if (have < comps) return s.fail("the input has fewer channels than the conversion reads");
        int give = 0;
        for (auto p : out)
        {
            if (p->empty() || p->width() != w || p->height() != h || p->type() != t)
                return s.fail("an output plane has the wrong size or element type");
            give += p->channels();
        }
        if (give != comps) return s.fail("the output planes' channels do not match the conversion");

        achip::ColourPlanes pl{};
        int slot = 0, k = 0;
        for (auto p : in)
        {
            Dev d = s.upload(*p, slot++);
            for (int c = 0; c < p->channels() && k < comps; c++) pl.in[k++] = { d.p, d.pitch, p->channels(), c };
        }
        std::vector<Dev> dout;
        k = 0;
        for (auto p : out)
        {
            Dev d = s.alloc(w, h, p->channels(), t, slot++);
            dout.push_back(d);
            for (int c = 0; c < p->channels(); c++) pl.out[k++] = { d.p, d.pitch, p->channels(), c };
        }
        if (!s.ok()) return false;
        if (!s.check(achip::launch_colour(t, kind, pl, w, h, s.stream()), "colour kernel")) return false;
        size_t i = 0;
        for (auto p : out) s.download(dout[i++], *p);
        return s.finish();
    }

    // Snapshot of output image handles: put back unless the helper succeeded, so a failed
    // call leaves its outputs as they were (including the ones it had just allocated).
    class Outputs
    {
    public:
        Outputs(std::initializer_list<Image*> list) { for (auto p : list) saved.emplace_back(p, *p); }
        void commit(const bool ok) noexcept { if (!ok) for (auto& kv : saved) *kv.first = kv.second; }

    private:
        std::vector<std::pair<Image*, Image>> saved;
    };

    // ImageProcess.cpp:427-462: a new image unless dst already has src's shape and type and
    // is not src itself; float images are left alone.
    void shift(const char* fn, const Image& src, Image& dst, const int n, const bool left, const bool inPlace) noexcept
    {
        if (src.empty() || src.isFloat()) return;
        try
        {
            Image tmp{};
            if (inPlace) tmp = dst;
            else if (dst.empty() || src == dst || dst.width() != src.width() || dst.height() != src.height() ||
                     dst.channels() != src.channels() || dst.type() != src.type())
                tmp.create(src.width(), src.height(), src.channels(), src.type());
            else tmp = dst;
            if (tmp.empty()) return;
            Session s(fn);
            if (n < 0) { s.fail("negative shift"); return; }
            Dev d = s.upload(src, 0);
            if (!s.ok()) return;
            if (!s.check(achip::launch_shift(src.type(), left ? 1 : 0, n, d.p, d.pitch, d.p, d.pitch, d.w * d.c, d.h, s.stream()),
                         "shift kernel")) return;
            s.download(d, tmp);
            if (!s.finish()) return;
            if (dst != tmp) dst = tmp;
        }
        catch (...) { std::fprintf(stderr, "ac::core::%s: out of memory\n", fn); }
    }

    // dst (already shaped, any type) = src converted element by element
    bool convertInto(const char* fn, const Image& src, Image& dst)
    {
        Session s(fn);
        if (!validType(src.type()) || !validType(dst.type())) return s.fail("unsupported element type");
        Dev a = s.upload(src, 0);
        if (!s.ok()) return false;
        if (src.type() == dst.type()) { s.download(a, dst); return s.finish(); }
        Dev b = s.alloc(src.width(), src.height(), src.channels(), dst.type(), 1);
        if (!s.ok()) return false;
        if (!s.check(achip::launch_convert(src.type(), dst.type(), a.p, a.pitch, b.p, b.pitch, a.w * a.c, a.h, s.stream()),
                     "convert kernel")) return false;
        s.download(b, dst);
        return s.finish();
    }

    // stb_image_resize2's behaviour as ImageResize.cpp:136-273 drives it (resize kernels)
    bool resizeInto(const Image& src, Image& dst, const double fx, const double fy, int mode);
    void resize(const Image& src, Image& dst, const double fx, const double fy, const int mode) noexcept
    {
        Outputs o{ &dst };
        bool ok = false;
        try { ok = resizeInto(src, dst, fx, fy, mode); }
        catch (...) { std::fprintf(stderr, "ac::core::resize: out of memory\n"); }
        o.commit(ok);
    }
    // false only when the resize should have happened and did not
    bool resizeInto(const Image& src, Image& dst, const double fx, const double fy, int mode)
    {
        if (src.empty()) return true;
        {
            if (fx > 0.0 && fy > 0.0)
            {
                if (fx == 1.0 && fy == 1.0) { dst = src; return true; }
                const int dstW = static_cast<int>(src.width() * fx);
                const int dstH = static_cast<int>(src.height() * fy);
                if (dst.width() != dstW || dst.height() != dstH || dst.channels() != src.channels() || dst.type() != src.type())
                    dst.create(dstW, dstH, src.channels(), src.type());
            }
            else
            {
                if (dst.empty()) return true;
                if (dst.width() == src.width() && dst.height() == src.height()) { dst = src; return true; }
                if (dst.channels() != src.channels() || dst.type() != src.type())
                    dst.create(dst.width(), dst.height(), src.channels(), src.type());
            }
            if (dst.empty()) return false;
            if (mode < 0 || mode >= achip::kResizeModeCount) mode = achip::kResizeBilinear;   // the reference's default case
            Session s("resize");
            if (src.channels() < 1 || src.channels() > 4) return s.fail("resize takes 1 to 4 channels");
            if (!validType(src.type())) return s.fail("unsupported element type");
            Dev a = s.upload(src, 0);
            Dev b = s.alloc(dst.width(), dst.height(), dst.channels(), dst.type(), 1);
            if (!s.ok()) return false;
            const hipError_t err = achip::launch_resize(src.type(), src.channels(), mode, a.p, a.pitch, a.w, a.h, b.p, b.pitch,
                                                        b.w, b.h, s.stream());
            if (err == hipErrorInvalidValue) return s.fail("the downscale needs more filter taps than the kernel holds");
            if (!s.check(err, "resize kernel")) return false;
            s.download(b, dst);
            return s.finish();
        }
    }
}

std::string ac::core::detail::takeImageError() noexcept
{
    std::string e;
    e.swap(ops::lastError);
    return e;
}

// ---- colour (ImageProcess.cpp:217-418) ----

void ac::core::rgb2yuv(const Image& rgb, Image& yuv)
{
    if (rgb.empty()) return;
    detail::ops::Outputs o{ &yuv };
    if (yuv.empty()) yuv.create(rgb.width(), rgb.height(), 3, rgb.type());
    o.commit(detail::ops::colour("rgb2yuv", achip::kRgb2Yuv, { &rgb }, { &yuv }));
}
void ac::core::rgb2yuv(const Image& rgb, Image& y, Image& uv)
{
    if (rgb.empty()) return;
    detail::ops::Outputs o{ &y, &uv };
    if (y.empty()) y.create(rgb.width(), rgb.height(), 1, rgb.type());
    if (uv.empty()) uv.create(rgb.width(), rgb.height(), 2, rgb.type());
    o.commit(detail::ops::colour("rgb2yuv", achip::kRgb2Yuv, { &rgb }, { &y, &uv }));
}
void ac::core::rgb2yuv(const Image& rgb, Image& y, Image& u, Image& v)
{
    if (rgb.empty()) return;
    detail::ops::Outputs o{ &y, &u, &v };
    if (y.empty()) y.create(rgb.width(), rgb.height(), 1, rgb.type());
    if (u.empty()) u.create(rgb.width(), rgb.height(), 1, rgb.type());
    if (v.empty()) v.create(rgb.width(), rgb.height(), 1, rgb.type());
    o.commit(detail::ops::colour("rgb2yuv", achip::kRgb2Yuv, { &rgb }, { &y, &u, &v }));
}
void ac::core::rgba2yuva(const Image& rgba, Image& yuva)
{
    if (rgba.empty()) return;
    detail::ops::Outputs o{ &yuva };
    if (yuva.empty()) yuva.create(rgba.width(), rgba.height(), 4, rgba.type());
    o.commit(detail::ops::colour("rgba2yuva", achip::kRgba2Yuva, { &rgba }, { &yuva }));
}
void ac::core::rgba2yuva(const Image& rgba, Image& y, Image& uva)
{
    if (rgba.empty()) return;
    detail::ops::Outputs o{ &y, &uva };
    if (y.empty()) y.create(rgba.width(), rgba.height(), 1, rgba.type());
    if (uva.empty()) uva.create(rgba.width(), rgba.height(), 3, rgba.type());
    o.commit(detail::ops::colour("rgba2yuva", achip::kRgba2Yuva, { &rgba }, { &y, &uva }));
}
void ac::core::rgba2yuva(const Image& rgba, Image& y, Image& u, Image& v, Image& a)
{
    if (rgba.empty()) return;
    detail::ops::Outputs o{ &y, &u, &v, &a };
    if (y.empty()) y.create(rgba.width(), rgba.height(), 1, rgba.type());
    if (u.empty()) u.create(rgba.width(), rgba.height(), 1, rgba.type());
    if (v.empty()) v.create(rgba.width(), rgba.height(), 1, rgba.type());
    if (a.empty()) a.create(rgba.width(), rgba.height(), 1, rgba.type());
    o.commit(detail::ops::colour("rgba2yuva", achip::kRgba2Yuva, { &rgba }, { &y, &u, &v, &a }));
}
void ac::core::yuv2rgb(const Image& yuv, Image& rgb)
{
    if (yuv.empty()) return;
    detail::ops::Outputs o{ &rgb };
    if (rgb.empty()) rgb.create(yuv.width(), yuv.height(), 3, yuv.type());
    o.commit(detail::ops::colour("yuv2rgb", achip::kYuv2Rgb, { &yuv }, { &rgb }));
}
void ac::core::yuv2rgb(const Image& y, const Image& uv, Image& rgb)
{
    if (y.empty() || uv.empty()) return;
    detail::ops::Outputs o{ &rgb };
    if (rgb.empty()) rgb.create(y.width(), y.height(), 3, y.type());
    o.commit(detail::ops::colour("yuv2rgb", achip::kYuv2Rgb, { &y, &uv }, { &rgb }));
}
void ac::core::yuv2rgb(const Image& y, const Image& u, const Image& v, Image& rgb)
{
    if (y.empty() || u.empty() || v.empty()) return;
    detail::ops::Outputs o{ &rgb };
    if (rgb.empty()) rgb.create(y.width(), y.height(), 3, y.type());
    o.commit(detail::ops::colour("yuv2rgb", achip::kYuv2Rgb, { &y, &u, &v }, { &rgb }));
}
void ac::core::yuva2rgba(const Image& yuva, Image& rgba)
{
    if (yuva.empty()) return;
    detail::ops::Outputs o{ &rgba };
    if (rgba.empty()) rgba.create(yuva.width(), yuva.height(), 4, yuva.type());
    o.commit(detail::ops::colour("yuva2rgba", achip::kYuva2Rgba, { &yuva }, { &rgba }));
}
void ac::core::yuva2rgba(const Image& y, const Image& uva, Image& rgba)
{
    if (y.empty() || uva.empty()) return;
    detail::ops::Outputs o{ &rgba };
    if (rgba.empty()) rgba.create(y.width(), y.height(), 4, y.type());
    o.commit(detail::ops::colour("yuva2rgba", achip::kYuva2Rgba, { &y, &uva }, { &rgba }));
}
void ac::core::yuva2rgba(const Image& y, const Image& u, const Image& v, const Image& a, Image& rgba)
{
    if (y.empty() || u.empty() || v.empty() || a.empty()) return;
    detail::ops::Outputs o{ &rgba };
    if (rgba.empty()) rgba.create(y.width(), y.height(), 4, y.type());
    o.commit(detail::ops::colour("yuva2rgba", achip::kYuva2Rgba, { &y, &u, &v, &a }, { &rgba }));
}

// ---- element helpers (ImageProcess.cpp:419-797) ----

ac::core::Image ac::core::unpadding(const Image& src) noexcept
{
    const int lineSize = src.width() * src.pixelSize();
    if (src.empty() || src.stride() == lineSize) return src;
    try
    {
        Image dst{ src.width(), src.height(), src.channels(), src.type(), lineSize };
        if (dst.empty()) return Image{};
        detail::ops::Session s("unpadding");
        detail::ops::Dev d = s.upload(src, 0);
        s.download(d, dst);
        return s.finish() ? dst : Image{};
    }
    catch (...) { return Image{}; }
}
void ac::core::shl(Image& image, const int n) noexcept { detail::ops::shift("shl", image, image, n, true, true); }
void ac::core::shl(const Image& src, Image& dst, const int n) noexcept { detail::ops::shift("shl", src, dst, n, true, false); }
void ac::core::shr(Image& image, const int n) noexcept { detail::ops::shift("shr", image, image, n, false, true); }
void ac::core::shr(const Image& src, Image& dst, const int n) noexcept { detail::ops::shift("shr", src, dst, n, false, false); }

ac::core::Image ac::core::astype(const Image& src, const int type) noexcept
{
    if (src.empty() || src.type() == type) return src;
    try
    {
        if (!detail::ops::validType(type))
        {
            detail::ops::Session s("astype");
            s.fail("unsupported element type");
            return Image{};
        }
        Image dst{ src.width(), src.height(), src.channels(), type };
        if (dst.empty()) return Image{};
        return detail::ops::convertInto("astype", src, dst) ? dst : Image{};
    }
    catch (...) { return Image{}; }
}

void ac::core::copy(const Image& src, Image& dst) noexcept
{
    if (src.empty()) { dst = src; return; }
    const bool sameShape = dst.width() == src.width() && dst.height() == src.height() && dst.channels() == src.channels();
    if (dst.ptr() == src.ptr() && dst.type() == src.type() && sameShape) return;
    try
    {
        Image tmp{};
        if (dst.empty() || src == dst || !sameShape)
            tmp.create(src.width(), src.height(), src.channels(), dst.empty() ? src.type() : dst.type());
        else tmp = dst;
        if (tmp.empty()) return;
        if (!detail::ops::convertInto("copy", src, tmp)) return;
        if (dst != tmp) dst = tmp;
    }
    catch (...) {}
}

ac::core::Image ac::core::crop(const Image& src, const int x, const int y, const int w, const int h) noexcept
{
    if (src.empty()) return src;
    const int rectX = w < 0 ? x + w : x;
    const int rectY = h < 0 ? y + h : y;
    return src.view(rectX, rectY, std::abs(w), std::abs(h));
}

ac::core::Image ac::core::extract(const Image& src, const int channel, const int n) noexcept
{
    if (src.empty() || channel >= src.channels() || channel < 0 || n <= 0) return Image{};
    if (channel == 0 && n >= src.channels()) return src;
    try
    {
        Image dst{ src.width(), src.height(), std::min(src.channels() - channel, n), src.type() };
        if (dst.empty()) return Image{};
        detail::ops::Session s("extract");
        auto a = s.upload(src, 0);
        auto b = s.alloc(dst.width(), dst.height(), dst.channels(), dst.type(), 1);
        if (!s.ok()) return Image{};
        if (!s.check(achip::launch_copy_runs(a.p, a.pitch, src.pixelSize(), channel * src.elementSize(), b.p, b.pitch,
                                             dst.pixelSize(), 0, dst.pixelSize(), a.w, a.h, s.stream()), "copy kernel"))
            return Image{};
        s.download(b, dst);
        return s.finish() ? dst : Image{};
    }
    catch (...) { return Image{}; }
}

ac::core::Image ac::core::insert(const Image& src, const Image& image, const int channel) noexcept
{
    if (src.empty() || image.empty() || channel > src.channels() || channel < 0 ||
        image.width() != src.width() || image.height() != src.height() || image.type() != src.type())
        return src;
    try
    {
        Image dst{ src.width(), src.height(), src.channels() + image.channels(), src.type() };
        if (dst.empty()) return Image{};
        detail::ops::Session s("insert");
        auto a = s.upload(src, 0);
        auto b = s.upload(image, 1);
        auto o = s.alloc(dst.width(), dst.height(), dst.channels(), dst.type(), 2);
        if (!s.ok()) return Image{};
        const int head = channel * src.elementSize(), mid = image.pixelSize(), tail = src.pixelSize() - head;
        const int ps = src.pixelSize(), pd = dst.pixelSize();
        hipError_t err = achip::launch_copy_runs(a.p, a.pitch, ps, 0, o.p, o.pitch, pd, 0, head, a.w, a.h, s.stream());
        if (err == hipSuccess) err = achip::launch_copy_runs(b.p, b.pitch, mid, 0, o.p, o.pitch, pd, head, mid, a.w, a.h, s.stream());
        if (err == hipSuccess) err = achip::launch_copy_runs(a.p, a.pitch, ps, head, o.p, o.pitch, pd, head + mid, tail, a.w, a.h, s.stream());
        if (!s.check(err, "copy kernel")) return Image{};
        s.download(o, dst);
        return s.finish() ? dst : Image{};
    }
    catch (...) { return Image{}; }
}

void ac::core::pixelShuffle(const Image& src, Image& dst, const int upscale) noexcept
{
    const int group = upscale * upscale;
    if (src.empty() || upscale <= 0 || (src.channels() % group)) return;
    detail::ops::Outputs o{ &dst };
    bool ok = false;
    try
    {
        if (dst.empty() || dst.width() != src.width() * upscale || dst.height() != src.height() * upscale ||
            dst.channels() != src.channels() / group)
            dst.create(src.width() * upscale, src.height() * upscale, src.channels() / group, src.type());
        if (dst.empty()) return;
        detail::ops::Session s("pixelShuffle");
        if (!detail::ops::validType(src.type()) || !detail::ops::validType(dst.type())) s.fail("unsupported element type");
        auto a = s.upload(src, 0);
        auto b = s.alloc(dst.width(), dst.height(), dst.channels(), dst.type(), 1);
        if (s.ok() && s.check(achip::launch_pixel_shuffle(src.type(), dst.type(), a.p, a.pitch, a.w, a.h, a.c, upscale, b.p,
                                                          b.pitch, s.stream()), "pixel shuffle kernel"))
        {
            s.download(b, dst);
            ok = s.finish();
        }
    }
    catch (...) {}
    o.commit(ok);
}

// ---- resize (ImageResize.cpp:275-293) ----

void ac::core::resize(const Image& src, Image& dst, const double fx, const double fy, const int mode) noexcept
{
    if (src == dst)
    {
        Image tmp{};
        detail::ops::resize(src, tmp, fx, fy, mode);
        if (!tmp.empty()) dst = tmp;
    }
    else detail::ops::resize(src, dst, fx, fy, mode);
}
ac::core::Image ac::core::resize(const Image& src, const double fx, const double fy, const int mode) noexcept
{
    if (fx <= 0.0 || fy <= 0.0) return src;
    Image dst{};
    detail::ops::resize(src, dst, fx, fy, mode);
    return dst;
}
