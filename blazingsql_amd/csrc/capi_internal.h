// capi_internal.h -- definitions shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bsql/interp.h"
#include "program.hpp"

namespace bsql {
// True once the process has begun to exit (a C atexit handler of this
// library sets it): destructors then leave device memory to the process
// teardown instead of calling into a runtime that may be shutting down.
bool process_exiting();
}  // namespace bsql

// The opaque program handle: the program plus its cache of lowered plans.
// `uid` is unique per handle (copies of the Program part keep their own), so
// that a lowering cached for a pair of programs never outlives either.
struct bsql_program : bsql::Program {
  const uint64_t uid = bsql::next_program_id();
  std::mutex cache_mu;
  std::map<std::string, bsql::Lowered> cache;
  // copies whose step columns are numbered from another table width (a
  // projection's steps placed after a predicate's), by that width
  std::map<int, std::unique_ptr<bsql_program>> rebased;
  ~bsql_program() {
    if (d_strings && !bsql::process_exiting()) (void)hipFree(d_strings);
  }
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
// Filter + project for any projection: one selection, then the output phase
// per part of a split projection and per group of at most
// BSQL_DEV_MAX_OUT outputs.
int run_filter_project(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in, bsql_column* out,
                       int n_out, int64_t num_rows, int64_t capacity, int64_t* d_count, void* stream);
// bsql_interp_execute for any program: its parts, steps and output groups.
int run_execute(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                int64_t num_rows, void* stream);
// Fails with BSQL_ERR_INVALID if `stream` is capturing a graph (the entry
// points that allocate scratch or copy host memory cannot be captured).
void reject_capture(void* stream, const char* entry);

// The body of bsql_interp_execute for a program without steps; a private
// scratch buffer (allocated and freed on the stream) leaves the stream's
// shared scratch, and the row ids a filter keeps there, untouched.
int execute_program(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                    int64_t num_rows, void* stream, bool private_scratch);

// The device allocator installed with bsql_set_allocator (stream-ordered
// hipMallocAsync / hipFreeAsync when none is): alloc returns 0, 1 for out of
// memory, 2 for another failure.
struct AllocRef {
  int (*alloc)(void* ctx, void** ptr, size_t bytes, void* stream);
  void (*release)(void* ctx, void* ptr, size_t bytes, void* stream);
  void* ctx;
};
AllocRef installed_allocator();

// Columns of the table followed by the columns a program's steps build.
// Owned buffers are released on the stream when this is destroyed.
struct StepColumns {
  std::vector<bsql_column> cols;
  std::vector<std::pair<void*, size_t>> owned;  // freed with `alloc` on `stream`
  hipStream_t stream = nullptr;
  AllocRef alloc = installed_allocator();
  StepColumns() = default;
  StepColumns(const StepColumns&) = delete;
  StepColumns& operator=(const StepColumns&) = delete;
  ~StepColumns();
};
// Appends the columns of p's steps to sc.cols (which holds p's table).
void build_steps(bsql_program& p, StepColumns& sc, int64_t num_rows);
// p with its step columns numbered from table width n_table.
bsql_program* rebased_program(bsql_program* p, int n_table);
// Chars a string output over these columns (table + steps) can need.
int64_t string_chars_bound(const bsql_program& p, const bsql_column* in, int n_in, int64_t num_rows);
// Types of the table's columns followed by the steps' columns.
std::vector<int32_t> extended_types(const Program& p, const int32_t* types, int n);
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
