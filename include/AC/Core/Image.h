/*
 * C image API -- the same structs and functions as the reference's C binding
 * (binding/c/include/AC/Core/Image.h:15-83).  Each function below replaces the
 * reference function of the same name (binding/c/src/Binding.cpp:33-160).
 * Image file I/O (ac_imread/ac_imwrite) is not part of this library.
 */
#ifndef AC_BINDING_C_CORE_IMAGE_H
#define AC_BINDING_C_CORE_IMAGE_H

#include <stdint.h>
#include <stddef.h>

#include "ACCExport.h"

#ifdef __cplusplus
#   define AC_C_API extern "C" AC_C_EXPORT
#else
#   define AC_C_API AC_C_EXPORT
#endif

#define AC_CORE_DISABLE_IMAGE_IO 1

enum ACImageElementType
{
    AC_IMAGE_UINT8   = 0 << 8 | 1,
    AC_IMAGE_UINT16  = 0 << 8 | 2,
    AC_IMAGE_FLOAT16 = 2 << 8 | 2,   /* the core supports f16; the reference's C enum omits it */
    AC_IMAGE_FLOAT32 = 2 << 8 | 4
};

enum ACResizeModes
{
    AC_RESIZE_POINT,
    AC_RESIZE_CATMULL_ROM,
    AC_RESIZE_MITCHELL_NETRAVALI,
    AC_RESIZE_BICUBIC_0_60,
    AC_RESIZE_BICUBIC_0_75,
    AC_RESIZE_BICUBIC_0_100,
    AC_RESIZE_BICUBIC_20_50,
    AC_RESIZE_SOFTCUBIC50,
    AC_RESIZE_SOFTCUBIC75,
    AC_RESIZE_SOFTCUBIC100,
    AC_RESIZE_LANCZOS2,
    AC_RESIZE_LANCZOS3,
    AC_RESIZE_LANCZOS4,
    AC_RESIZE_SPLINE16,
    AC_RESIZE_SPLINE36,
    AC_RESIZE_SPLINE64,
    AC_RESIZE_BILINEAR,
};

typedef struct ACImage
{
    int width;
    int height;
    int channels;
    int stride;
    int element_type;
    void* ptr;
    struct ACImageHandle* hptr;
} ACImage;

/* Binding.cpp:33-38 */
AC_C_API ACImage* ac_image_alloc(void);
/* Binding.cpp:39-45 */
AC_C_API void ac_image_free(ACImage** image);
/* Binding.cpp:46-53 */
AC_C_API int ac_image_ref(const ACImage* src, ACImage* dst);
/* Binding.cpp:54-59 */
AC_C_API void ac_image_unref(ACImage* image);
/* Binding.cpp:60-67 */
AC_C_API int ac_image_create(ACImage* image);
/* Binding.cpp:68-75 */
AC_C_API int ac_image_map(ACImage* image);
/* Binding.cpp:76-83 */
AC_C_API int ac_image_from(ACImage* image, const void* data);
/* Binding.cpp:84-91 */
AC_C_API int ac_image_view(const ACImage* src, ACImage* dst, int x, int y, int w, int h);
/* Binding.cpp:92-99 */
AC_C_API int ac_image_clone(const ACImage* src, ACImage* dst);
/* Binding.cpp:100-105 */
AC_C_API int ac_image_to(const ACImage* image, void* data, int stride);

/* Binding.cpp:125-159.  These run on the calling thread's current HIP device; a device
 * failure returns AC_ERROR(AC_EPROCESSOR) (the reference's never fail once their arguments
 * are valid).  mode is an ACResizeModes value. */
/* Binding.cpp:125-131 */
AC_C_API int ac_resize(const ACImage* src, ACImage* dst, double fx, double fy, int mode);
/* Binding.cpp:132-138 */
AC_C_API int ac_rgb2yuv(const ACImage* rgb, ACImage* yuv);
/* Binding.cpp:139-145 */
AC_C_API int ac_rgba2yuva(const ACImage* rgba, ACImage* yuva);
/* Binding.cpp:146-152 */
AC_C_API int ac_yuv2rgb(const ACImage* yuv, ACImage* rgb);
/* Binding.cpp:153-159 */
AC_C_API int ac_yuva2rgba(const ACImage* yuva, ACImage* rgba);

static inline int ac_image_size(const ACImage* image) { return image->height * image->stride; }
static inline int ac_image_element_size(const ACImage* image) { return image->element_type & 0xff; }
static inline int ac_image_pixel_size(const ACImage* image) { return image->channels * ac_image_element_size(image); }
static inline uint8_t* ac_image_data(const ACImage* image) { return (uint8_t*)image->ptr; }
static inline uint8_t* ac_image_line(const ACImage* image, int y) { return ac_image_data(image) + image->stride * y; }
static inline uint8_t* ac_image_pixel(const ACImage* image, int x, int y) { return ac_image_line(image, y) + x * ac_image_pixel_size(image); }
static inline int ac_image_empty(const ACImage* image) { return image->ptr == NULL; }
static inline int ac_image_is_uint(const ACImage* image) { return (image->element_type >> 8) == 0; }
static inline int ac_image_is_int(const ACImage* image) { return (image->element_type >> 8) == 1; }
static inline int ac_image_is_float(const ACImage* image) { return (image->element_type >> 8) == 2; }

#endif
