/* Error codes of the C API (binding/c/include/AC/Error.h:1-11). */
#ifndef AC_BINDING_C_ERROR_H
#define AC_BINDING_C_ERROR_H

#define AC_SUCCESS 0
#define AC_ERROR(e) (-(e))

#define AC_EIO          5
#define AC_EINVAL      22
#define AC_EPROCESSOR 256

#endif
