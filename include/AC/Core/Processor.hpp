// ac::core::Processor -- interface-compatible with core/include/AC/Core/Processor.hpp:15-48,
// with one more processor type: HIP = 3 (the MI355X-native ACNetLegacy path).
#ifndef AC_CORE_PROCESSOR_HPP
#define AC_CORE_PROCESSOR_HPP

#include <memory>

#include "AC/Core/Image.hpp"

#include "ACCoreExport.hpp"

namespace ac::core
{
    class Processor;
}

class ac::core::Processor
{
public:
    static constexpr int CPU = 0;
    static constexpr int OpenCL = 1;
    static constexpr int CUDA = 2;
    static constexpr int HIP = 3;

public:
    AC_CORE_EXPORT Processor() noexcept;
    AC_CORE_EXPORT virtual ~Processor();

    AC_CORE_EXPORT Image process(const Image& src, double factor);
    // If `dst` is not empty, it is assumed to be correctly allocated and the result is
    // stored in that preallocated buffer.
    AC_CORE_EXPORT void process(const Image& src, Image& dst, double factor);

    AC_CORE_EXPORT virtual bool ok() noexcept;
    AC_CORE_EXPORT virtual const char* error() noexcept;
    AC_CORE_EXPORT virtual const char* name() const noexcept = 0;
    AC_CORE_EXPORT virtual int type() const noexcept = 0;
    AC_CORE_EXPORT virtual const char* typeName() const noexcept = 0;
private:
    // one 2x pass on a 1-channel image; input type from src, output type from dst
    AC_CORE_EXPORT virtual void process(const Image& src, Image& dst) = 0;

protected:
    // A processor that runs all of process(src, dst, factor) itself (colour split, passes,
    // resize, merge) returns true; errors are reported through ok()/error().  The default
    // returns false and the steps below run one at a time.
    AC_CORE_EXPORT virtual bool processAll(const Image& src, Image& dst, double factor);
    // Steps of process(src, dst, factor) around the 2x passes.  The defaults run the
    // reference's sequence one pass at a time; a processor may keep chained passes and
    // colour work on its device instead.  Return false after recording an error.
    AC_CORE_EXPORT virtual bool processChain(const Image& src, Image& dst, int passes);
    AC_CORE_EXPORT virtual bool colorSplit(const Image& src, Image& y, Image& uv);
    AC_CORE_EXPORT virtual bool colorMerge(const Image& y, const Image& uv, Image& dst);
    AC_CORE_EXPORT virtual bool resizeTo(const Image& src, Image& dst, double fx, double fy);
    AC_CORE_EXPORT virtual void setError(const char* message) noexcept;

public:
    AC_CORE_EXPORT static std::shared_ptr<Processor> create(const char* type, int device, const char* model);
    AC_CORE_EXPORT static const char* listInfo();

    template<int type, typename Model> static std::shared_ptr<Processor> create(int idx, const Model& model);
    template<int type> static const char* info();

protected:
    int idx;
};

#endif
