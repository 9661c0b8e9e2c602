// Export macro of this library's C++ API (core/include/AC/Core/*.hpp include it by this name).
#ifndef AC_CORE_EXPORT_HPP
#define AC_CORE_EXPORT_HPP
#if defined(__GNUC__)
#define AC_CORE_EXPORT __attribute__((visibility("default")))
#else
#define AC_CORE_EXPORT
#endif
#endif
