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

// Phases of a filter launch: the selection (row ids + count into *d_count)
// and the output ops over the selected rows. A caller that sizes its
// outputs from the count runs PHASE_SELECT, reads the count, then
// PHASE_OUTPUT on the same stream with nothing else on it in between (the
// row ids stay in the stream's scratch). In PHASE_SELECT `out` only gives
// the output types; its buffers are not touched.
enum { PHASE_SELECT = 1, PHASE_OUTPUT = 2, PHASE_ALL = 3 };
int run_compact(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in, bsql_column* out,
                int n_out, int64_t num_rows, int64_t capacity, int64_t* d_count, void* stream, int phases);
}  // namespace bsql

// Body of a C entry point: exceptions become return codes + message.
#define BSQL_TRY(...)                                               \
  try {                                                             \
    __VA_ARGS__                                                     \
  } catch (const bsql::Error& e) {                                  \
    return bsql::set_error(e.code, e.what());                       \
  } catch (const std::bad_alloc&) {                                 \
    return bsql::set_error(BSQL_ERR_OOM, "host allocation failed"); \
  } catch (const std::exception& e) {                               \
    return bsql::set_error(BSQL_ERR_INVALID, e.what());             \
  }
