// Errors inside the library: thrown as LfError and turned into an LF_ERR_*
// code at the C ABI.
#pragma once

#include <string>

#include "../../include/latticeflow_amd.h"

namespace lfamd {

struct LfError {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw LfError{code, msg}; }

}  // namespace lfamd
