#ifndef AC_CORE_HPP
#define AC_CORE_HPP

#include "AC/Core/Image.hpp"
#include "AC/Core/Model.hpp"
#include "AC/Core/Processor.hpp"

#endif
