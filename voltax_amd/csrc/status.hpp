// Status codes shared by every C-ABI entry point (defined in the public header).
#pragma once
#include "../../include/voltax_mi355x.h"

#include <new>

// Every extern "C" body runs inside VXM_TRY ... VXM_CATCH: no C++ exception
// (std::bad_alloc from the host-side tables above all) crosses the C ABI.
#define VXM_TRY try {
#define VXM_CATCH                          \
  }                                        \
  catch (const std::bad_alloc&) {          \
    return VXM_ERR_HOST_MEMORY;            \
  }                                        \
  catch (...) {                            \
    return VXM_ERR_INVALID;                \
  }
