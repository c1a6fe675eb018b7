// capi_internal.h -- definitions shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "../../include/bsql/interp.h"
#include "program.hpp"

// The opaque program handle: the program plus its cache of lowered plans.
struct bsql_program : bsql::Program {
  std::mutex cache_mu;
  std::map<std::string, bsql::Lowered> cache;
};

namespace bsql {
// Records the message of the calling thread's last error; returns code.
int set_error(int code, const std::string& msg);
// Maps a HIP error to BSQL_ERR_OOM / BSQL_ERR_DEVICE with a message.
int device_error(hipError_t e, const char* where);
}  // namespace bsql
