// Exception boundary of the C-ABI: every int-returning entry point runs its
// body through guarded(), so that no C++ exception reaches a C, Go or Python
// caller.  The exceptions host code here can raise all come from obtaining
// memory or threads (std::bad_alloc, std::length_error for an impossible size,
// std::system_error from std::thread), and each becomes DMF_ENOMEM.
#pragma once

#include "../../include/dmf.h"

template <class F>
int guarded(F&& body) noexcept {
  try {
    return body();
  } catch (...) {
    return DMF_ENOMEM;
  }
}
