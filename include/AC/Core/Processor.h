/*
 * C processor API -- the same structs and functions as the reference's C binding
 * (binding/c/include/AC/Core/Processor.h:8-36), with AC_PROCESSOR_HIP added.  Each
 * function replaces the reference function of the same name (binding/c/src/Binding.cpp:161-247).
 * type strings: "hip" (FP32), "hip-fp16", "hip-bf16", or "auto" (= HIP here).
 */
#ifndef AC_BINDING_C_CORE_PROCESSOR_H
#define AC_BINDING_C_CORE_PROCESSOR_H

#include <stdint.h>

#include "AC/Core/Image.h"

enum ACProcessorType
{
    AC_PROCESSOR_CPU    = 0,
    AC_PROCESSOR_OPENCL = 1,
    AC_PROCESSOR_CUDA   = 2,
    AC_PROCESSOR_HIP    = 3
};

typedef struct ACProcessor
{
    int device;
    const char* type;
    const char* model;
    struct ACProcessorHandle* hptr;
} ACProcessor;

/* Binding.cpp:161-166 */
AC_C_API ACProcessor* ac_processor_alloc(void);
/* Binding.cpp:167-173 */
AC_C_API void ac_processor_free(ACProcessor** processor);
/* Binding.cpp:174-183 */
AC_C_API int ac_processor_ref(const ACProcessor* src, ACProcessor* dst);
/* Binding.cpp:184-189 */
AC_C_API void ac_processor_unref(ACProcessor* processor);
/* Binding.cpp:190-196 */
AC_C_API int ac_processor_create(ACProcessor* processor);
/* Binding.cpp:197-203 */
AC_C_API int ac_processor_process(ACProcessor* processor, const ACImage* src, ACImage* dst, double factor);
/* Binding.cpp:204-208 */
AC_C_API int ac_processor_ok(const ACProcessor* processor);
/* Binding.cpp:209-213 */
AC_C_API const char* ac_processor_error(const ACProcessor* processor);
/* Binding.cpp:214-218 */
AC_C_API const char* ac_processor_name(const ACProcessor* processor);
/* Binding.cpp:219-223 */
AC_C_API int ac_processor_type(const ACProcessor* processor);
/* Binding.cpp:224-228 */
AC_C_API const char* ac_processor_type_name(const ACProcessor* processor);
/* Binding.cpp:230-243 */
AC_C_API const char* ac_processor_info(int processor_type);
/* Binding.cpp:244-247 */
AC_C_API const char* ac_processor_list_info(void);

#endif
