/* Export macro of this library's C API (binding/c/include/AC/Core/Image.h includes it by this name). */
#ifndef AC_C_EXPORT_H
#define AC_C_EXPORT_H
#if defined(__GNUC__)
#define AC_C_EXPORT __attribute__((visibility("default")))
#else
#define AC_C_EXPORT
#endif
#endif
