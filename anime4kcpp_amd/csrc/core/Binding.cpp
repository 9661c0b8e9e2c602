// C API (include/AC/Core/{Image,Processor}.h): malloc'd POD structs holding a handle to
// the C++ object, int error codes, no exceptions across the boundary.  Semantics follow
// binding/c/src/Binding.cpp:11-247 function by function.
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>

#include "AC/Core.h"
#include "AC/Core.hpp"
#include "core_internal.hpp"

struct ACImageHandle
{
    ac::core::Image object{};
};
struct ACProcessorHandle
{
    std::shared_ptr<ac::core::Processor> object{};
};

namespace
{
    void copyProp(const ac::core::Image& src, ACImage* const dst)
    {
        dst->width = src.width();
        dst->height = src.height();
        dst->channels = src.channels();
        dst->stride = src.stride();
        dst->element_type = src.type();
        dst->ptr = src.ptr();
    }
    template <typename H, typename... A>
    H* makeHandle(A&&... a) noexcept
    {
        return new (std::nothrow) H{ std::forward<A>(a)... };
    }
}

ACImage* ac_image_alloc(void)
{
    return static_cast<ACImage*>(std::calloc(1, sizeof(ACImage)));
}
void ac_image_free(ACImage** const image)
{
    if (!(image && *image)) return;
    ac_image_unref(*image);
    std::free(*image);
    *image = nullptr;
}
int ac_image_ref(const ACImage* const src, ACImage* const dst)
{
    if (!(src && dst && src->hptr)) return AC_ERROR(AC_EINVAL);
    if (dst->hptr) dst->hptr->object = src->hptr->object;
    else if (!(dst->hptr = makeHandle<ACImageHandle>(src->hptr->object))) return AC_ERROR(AC_EINVAL);
    copyProp(dst->hptr->object, dst);
    return AC_SUCCESS;
}
void ac_image_unref(ACImage* const image)
{
    if (!(image && image->hptr)) return;
    delete image->hptr;
    std::memset(image, 0, sizeof(ACImage));
}
int ac_image_create(ACImage* const image)
{
    if (!image) return AC_ERROR(AC_EINVAL);
    if (!image->hptr && !(image->hptr = makeHandle<ACImageHandle>())) return AC_ERROR(AC_EINVAL);
    image->hptr->object.create(image->width, image->height, image->channels, image->element_type, image->stride);
    copyProp(image->hptr->object, image);
    return AC_SUCCESS;
}
int ac_image_map(ACImage* const image)
{
    if (!image) return AC_ERROR(AC_EINVAL);
    if (!image->hptr && !(image->hptr = makeHandle<ACImageHandle>())) return AC_ERROR(AC_EINVAL);
    image->hptr->object.map(image->width, image->height, image->channels, image->element_type, image->ptr, image->stride);
    copyProp(image->hptr->object, image);
    return AC_SUCCESS;
}
int ac_image_from(ACImage* const image, const void* const data)
{
    if (!image) return AC_ERROR(AC_EINVAL);
    if (!image->hptr && !(image->hptr = makeHandle<ACImageHandle>())) return AC_ERROR(AC_EINVAL);
    image->hptr->object.from(image->width, image->height, image->channels, image->element_type, data, image->stride);
    copyProp(image->hptr->object, image);
    return AC_SUCCESS;
}
int ac_image_view(const ACImage* const src, ACImage* const dst, const int x, const int y, const int w, const int h)
{
    if (!(src && dst && src->hptr)) return AC_ERROR(AC_EINVAL);
    if (dst->hptr) dst->hptr->object = src->hptr->object.view(x, y, w, h);
    else if (!(dst->hptr = makeHandle<ACImageHandle>(src->hptr->object.view(x, y, w, h)))) return AC_ERROR(AC_EINVAL);
    copyProp(dst->hptr->object, dst);
    return AC_SUCCESS;
}
int ac_image_clone(const ACImage* const src, ACImage* const dst)
{
    if (!(src && dst && src->hptr)) return AC_ERROR(AC_EINVAL);
    if (dst->hptr) dst->hptr->object = src->hptr->object.clone();
    else if (!(dst->hptr = makeHandle<ACImageHandle>(src->hptr->object.clone()))) return AC_ERROR(AC_EINVAL);
    copyProp(dst->hptr->object, dst);
    return AC_SUCCESS;
}
int ac_image_to(const ACImage* const image, void* const data, const int stride)
{
    if (!(image && image->hptr && data)) return AC_ERROR(AC_EINVAL);
    image->hptr->object.to(data, stride);
    return AC_SUCCESS;
}

// Binding.cpp:125-159.  The reference returns AC_SUCCESS once its arguments are valid; here
// the helpers run on the device and can fail, which returns AC_ERROR(AC_EPROCESSOR).
namespace
{
    template <typename F>
    int imageOp(ACImage* const out, F&& f)
    {
        (void)ac::core::detail::takeImageError();
        f();
        copyProp(out->hptr->object, out);
        return ac::core::detail::takeImageError().empty() ? AC_SUCCESS : AC_ERROR(AC_EPROCESSOR);
    }
}
int ac_resize(const ACImage* const src, ACImage* const dst, const double fx, const double fy, const int mode)
{
    if (!(src && src->hptr && dst && dst->hptr)) return AC_ERROR(AC_EINVAL);
    return imageOp(dst, [&] { ac::core::resize(src->hptr->object, dst->hptr->object, fx, fy, mode); });
}
int ac_rgb2yuv(const ACImage* const rgb, ACImage* const yuv)
{
    if (!(rgb && rgb->hptr && yuv && yuv->hptr)) return AC_ERROR(AC_EINVAL);
    return imageOp(yuv, [&] { ac::core::rgb2yuv(rgb->hptr->object, yuv->hptr->object); });
}
int ac_rgba2yuva(const ACImage* const rgba, ACImage* const yuva)
{
    if (!(rgba && rgba->hptr && yuva && yuva->hptr)) return AC_ERROR(AC_EINVAL);
    return imageOp(yuva, [&] { ac::core::rgba2yuva(rgba->hptr->object, yuva->hptr->object); });
}
int ac_yuv2rgb(const ACImage* const yuv, ACImage* const rgb)
{
    if (!(yuv && yuv->hptr && rgb && rgb->hptr)) return AC_ERROR(AC_EINVAL);
    return imageOp(rgb, [&] { ac::core::yuv2rgb(yuv->hptr->object, rgb->hptr->object); });
}
int ac_yuva2rgba(const ACImage* const yuva, ACImage* const rgba)
{
    if (!(yuva && yuva->hptr && rgba && rgba->hptr)) return AC_ERROR(AC_EINVAL);
    return imageOp(rgba, [&] { ac::core::yuva2rgba(yuva->hptr->object, rgba->hptr->object); });
}

ACProcessor* ac_processor_alloc(void)
{
    return static_cast<ACProcessor*>(std::calloc(1, sizeof(ACProcessor)));
}
void ac_processor_free(ACProcessor** const processor)
{
    if (!(processor && *processor)) return;
    ac_processor_unref(*processor);
    std::free(*processor);
    *processor = nullptr;
}
int ac_processor_ref(const ACProcessor* const src, ACProcessor* const dst)
{
    if (!(src && dst && src->hptr)) return AC_ERROR(AC_EINVAL);
    if (dst->hptr) dst->hptr->object = src->hptr->object;
    else if (!(dst->hptr = makeHandle<ACProcessorHandle>(src->hptr->object))) return AC_ERROR(AC_EINVAL);
    dst->type = src->type;
    dst->device = src->device;
    dst->model = src->model;
    return AC_SUCCESS;
}
void ac_processor_unref(ACProcessor* const processor)
{
    if (!(processor && processor->hptr)) return;
    delete processor->hptr;
    std::memset(processor, 0, sizeof(ACProcessor));
}
int ac_processor_create(ACProcessor* const processor)
{
    if (!processor) return AC_ERROR(AC_EINVAL);
    if (!processor->hptr && !(processor->hptr = makeHandle<ACProcessorHandle>())) return AC_ERROR(AC_EINVAL);
    processor->hptr->object = ac::core::Processor::create(processor->type, processor->device, processor->model);
    return ac_processor_ok(processor);
}
int ac_processor_process(ACProcessor* const processor, const ACImage* const src, ACImage* const dst, const double factor)
{
    if (!(processor && processor->hptr && processor->hptr->object && src && src->hptr && dst && dst->hptr)) return AC_ERROR(AC_EINVAL);
    processor->hptr->object->process(src->hptr->object, dst->hptr->object, factor);
    copyProp(dst->hptr->object, dst);
    return ac_processor_ok(processor);
}
int ac_processor_ok(const ACProcessor* const processor)
{
    if (!(processor && processor->hptr && processor->hptr->object)) return AC_ERROR(AC_EINVAL);
    return processor->hptr->object->ok() ? AC_SUCCESS : AC_ERROR(AC_EPROCESSOR);
}
const char* ac_processor_error(const ACProcessor* const processor)
{
    if (!(processor && processor->hptr && processor->hptr->object)) return nullptr;
    return processor->hptr->object->error();
}
const char* ac_processor_name(const ACProcessor* const processor)
{
    if (!(processor && processor->hptr && processor->hptr->object)) return nullptr;
    return processor->hptr->object->name();
}
int ac_processor_type(const ACProcessor* const processor)
{
    if (!(processor && processor->hptr && processor->hptr->object)) return AC_ERROR(AC_EINVAL);
    return processor->hptr->object->type();
}
const char* ac_processor_type_name(const ACProcessor* const processor)
{
    if (!(processor && processor->hptr && processor->hptr->object)) return nullptr;
    return processor->hptr->object->typeName();
}
const char* ac_processor_info(const int processor_type)
{
    switch (processor_type)
    {
    case AC_PROCESSOR_HIP: return ac::core::Processor::info<ac::core::Processor::HIP>();
    default: return "unsupported processor";
    }
}
const char* ac_processor_list_info(void)
{
    return ac::core::Processor::listInfo();
}
