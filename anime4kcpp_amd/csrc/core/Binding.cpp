// C API of include/AC/Core/{Image,Processor}.h: each function replaces the reference
// function of the same name in binding/c/src/Binding.cpp:33-247 (same arguments, return
// codes and ownership rules; no exception crosses the boundary).
//
// Design.  Both C structs are plain records that the library keeps in step with a C++
// object behind `hptr`:
//   - a Slot<T> holds that object (an Image, or a shared Processor);
//   - `bind` puts a value into a struct's slot, creating the slot on first use;
//   - `mirror` rewrites an ACImage's public fields from its Image;
//   - `imageIn` / `processorIn` turn a struct pointer into its object, or null when the
//     struct (or its slot) is missing, which every entry point answers with AC_EINVAL.
// Every image entry point is one of three shapes: build an image in place from the
// struct's own fields (create, map, from), derive one image from another (ref, view,
// clone), or run a device helper from one image into another (resize, colour).
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <utility>

#include "AC/Core.h"
#include "AC/Core.hpp"
#include "core_internal.hpp"

namespace
{
    template<typename T>
    struct Slot
    {
        T object{};
    };
}

struct ACImageHandle : Slot<ac::core::Image> {};
struct ACProcessorHandle : Slot<std::shared_ptr<ac::core::Processor>> {};

namespace
{
    constexpr int kInvalid = AC_ERROR(AC_EINVAL);

    // Put `value` into the struct's slot; false only when a new slot cannot be allocated.
    template<typename Handle, typename Value>
    bool bind(Handle*& slot, Value&& value) noexcept
    {
        if (slot == nullptr)
        {
            slot = new (std::nothrow) Handle{};
            if (slot == nullptr) return false;
        }
        slot->object = std::forward<Value>(value);
        return true;
    }

    void mirror(ACImage* const s) noexcept
    {
        const ac::core::Image& im = s->hptr->object;
        *s = ACImage{ im.width(), im.height(), im.channels(), im.stride(), im.type(), im.ptr(), s->hptr };
    }

    ac::core::Image* imageIn(const ACImage* const s) noexcept
    {
        return (s != nullptr && s->hptr != nullptr) ? &s->hptr->object : nullptr;
    }

    ac::core::Processor* processorIn(const ACProcessor* const s) noexcept
    {
        return (s != nullptr && s->hptr != nullptr) ? s->hptr->object.get() : nullptr;
    }

    // create / map / from: an (empty) slot filled by `fill(image, struct)` from the
    // struct's own width, height, channels, type, stride (and pointer).
    template<typename Fill>
    int buildInPlace(ACImage* const s, Fill&& fill) noexcept
    {
        if (s == nullptr || !bind(s->hptr, ac::core::Image{})) return kInvalid;
        fill(s->hptr->object, *s);
        mirror(s);
        return AC_SUCCESS;
    }

    // ref / view / clone: dst's slot takes `derive(src image)`.
    template<typename Derive>
    int deriveFrom(const ACImage* const src, ACImage* const dst, Derive&& derive) noexcept
    {
        const ac::core::Image* from = imageIn(src);
        if (from == nullptr || dst == nullptr || !bind(dst->hptr, derive(*from))) return kInvalid;
        mirror(dst);
        return AC_SUCCESS;
    }

    // resize / colour conversions: both structs must carry images; the helper's device
    // error (if any) becomes AC_EPROCESSOR, the reference's helpers cannot fail.
    using Helper = void (*)(const ac::core::Image&, ac::core::Image&);
    int runHelper(const ACImage* const in, ACImage* const out, const Helper helper)
    {
        const ac::core::Image* a = imageIn(in);
        ac::core::Image* b = imageIn(out);
        if (a == nullptr || b == nullptr) return kInvalid;
        (void)ac::core::detail::takeImageError();
        helper(*a, *b);
        mirror(out);
        return ac::core::detail::takeImageError().empty() ? AC_SUCCESS : AC_ERROR(AC_EPROCESSOR);
    }

    // a query on a created processor, `fallback` when there is none
    template<typename R, typename Query>
    R ask(const ACProcessor* const s, const R fallback, Query&& query) noexcept
    {
        ac::core::Processor* p = processorIn(s);
        return p != nullptr ? query(*p) : fallback;
    }

    template<typename S>
    S* zeroed() noexcept
    {
        return static_cast<S*>(std::calloc(1, sizeof(S)));
    }

    // unref: the slot goes, the struct returns to all-zero
    template<typename S>
    void clear(S* const s) noexcept
    {
        if (s == nullptr || s->hptr == nullptr) return;
        delete s->hptr;
        *s = S{};
    }

    template<typename S>
    void release(S** const s) noexcept
    {
        if (s == nullptr || *s == nullptr) return;
        clear(*s);
        std::free(*s);
        *s = nullptr;
    }
}

// ---- images ---------------------------------------------------------------------------

ACImage* ac_image_alloc(void) { return zeroed<ACImage>(); }
void ac_image_free(ACImage** const image) { release(image); }
void ac_image_unref(ACImage* const image) { clear(image); }

int ac_image_create(ACImage* const image)
{
    return buildInPlace(image, [](ac::core::Image& im, const ACImage& s) {
        im.create(s.width, s.height, s.channels, s.element_type, s.stride);
    });
}

int ac_image_map(ACImage* const image)
{
    return buildInPlace(image, [](ac::core::Image& im, const ACImage& s) {
        im.map(s.width, s.height, s.channels, s.element_type, s.ptr, s.stride);
    });
}

int ac_image_from(ACImage* const image, const void* const data)
{
    return buildInPlace(image, [data](ac::core::Image& im, const ACImage& s) {
        im.from(s.width, s.height, s.channels, s.element_type, data, s.stride);
    });
}

int ac_image_ref(const ACImage* const src, ACImage* const dst)
{
    return deriveFrom(src, dst, [](const ac::core::Image& im) { return im; });
}

int ac_image_view(const ACImage* const src, ACImage* const dst, const int x, const int y, const int w, const int h)
{
    return deriveFrom(src, dst, [=](const ac::core::Image& im) { return im.view(x, y, w, h); });
}

int ac_image_clone(const ACImage* const src, ACImage* const dst)
{
    return deriveFrom(src, dst, [](const ac::core::Image& im) { return im.clone(); });
}

int ac_image_to(const ACImage* const image, void* const data, const int stride)
{
    const ac::core::Image* im = imageIn(image);
    if (im == nullptr || data == nullptr) return kInvalid;
    im->to(data, stride);
    return AC_SUCCESS;
}

int ac_resize(const ACImage* const src, ACImage* const dst, const double fx, const double fy, const int mode)
{
    const ac::core::Image* a = imageIn(src);
    ac::core::Image* b = imageIn(dst);
    if (a == nullptr || b == nullptr) return kInvalid;
    (void)ac::core::detail::takeImageError();
    ac::core::resize(*a, *b, fx, fy, mode);
    mirror(dst);
    return ac::core::detail::takeImageError().empty() ? AC_SUCCESS : AC_ERROR(AC_EPROCESSOR);
}

int ac_rgb2yuv(const ACImage* const rgb, ACImage* const yuv) { return runHelper(rgb, yuv, ac::core::rgb2yuv); }
int ac_rgba2yuva(const ACImage* const rgba, ACImage* const yuva) { return runHelper(rgba, yuva, ac::core::rgba2yuva); }
int ac_yuv2rgb(const ACImage* const yuv, ACImage* const rgb) { return runHelper(yuv, rgb, ac::core::yuv2rgb); }
int ac_yuva2rgba(const ACImage* const yuva, ACImage* const rgba) { return runHelper(yuva, rgba, ac::core::yuva2rgba); }

// ---- processors -----------------------------------------------------------------------

ACProcessor* ac_processor_alloc(void) { return zeroed<ACProcessor>(); }
void ac_processor_free(ACProcessor** const processor) { release(processor); }
void ac_processor_unref(ACProcessor* const processor) { clear(processor); }

int ac_processor_ref(const ACProcessor* const src, ACProcessor* const dst)
{
    if (src == nullptr || src->hptr == nullptr || dst == nullptr || !bind(dst->hptr, src->hptr->object)) return kInvalid;
    *dst = ACProcessor{ src->device, src->type, src->model, dst->hptr };
    return AC_SUCCESS;
}

int ac_processor_create(ACProcessor* const processor)
{
    if (processor == nullptr ||
        !bind(processor->hptr, ac::core::Processor::create(processor->type, processor->device, processor->model)))
        return kInvalid;
    return ac_processor_ok(processor);
}

int ac_processor_process(ACProcessor* const processor, const ACImage* const src, ACImage* const dst, const double factor)
{
    ac::core::Processor* p = processorIn(processor);
    const ac::core::Image* in = imageIn(src);
    ac::core::Image* out = imageIn(dst);
    if (p == nullptr || in == nullptr || out == nullptr) return kInvalid;
    p->process(*in, *out, factor);
    mirror(dst);
    return ac_processor_ok(processor);
}

int ac_processor_ok(const ACProcessor* const processor)
{
    return ask(processor, kInvalid, [](ac::core::Processor& p) { return p.ok() ? AC_SUCCESS : AC_ERROR(AC_EPROCESSOR); });
}

const char* ac_processor_error(const ACProcessor* const processor)
{
    return ask(processor, static_cast<const char*>(nullptr), [](ac::core::Processor& p) { return p.error(); });
}

const char* ac_processor_name(const ACProcessor* const processor)
{
    return ask(processor, static_cast<const char*>(nullptr), [](ac::core::Processor& p) { return p.name(); });
}

int ac_processor_type(const ACProcessor* const processor)
{
    return ask(processor, kInvalid, [](ac::core::Processor& p) { return p.type(); });
}

const char* ac_processor_type_name(const ACProcessor* const processor)
{
    return ask(processor, static_cast<const char*>(nullptr), [](ac::core::Processor& p) { return p.typeName(); });
}

// Only the HIP type exists in this library; every other type has no device list.
const char* ac_processor_info(const int processor_type)
{
    return processor_type == AC_PROCESSOR_HIP ? ac::core::Processor::info<ac::core::Processor::HIP>() : "unsupported processor";
}

const char* ac_processor_list_info(void) { return ac::core::Processor::listInfo(); }
