// prodm.hpp -- umbrella header of the host C++ layer over libprodm_mdr.so.
//
// Link: -lprodm_mdr -lamdhip64 -l:libzstd.so.1 -pthread, compile with
// -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include and this include directory.
#ifndef PRODM_HPP
#define PRODM_HPP

#include "components.hpp"
#include "composed.hpp"
#include "compressors.hpp"
#include "estimators.hpp"
#include "gpu_mdr.hpp"
#include "interpreters.hpp"
#include "io.hpp"

#endif
