// capi.cpp -- the C ABI (include/bsql/interp.h) over the host program
// representation and the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bsql/interp.h"
#include "program.hpp"

#include "capi_internal.h"

namespace bsql {

hipError_t launch_execute(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt,
                          hipStream_t stream);
int64_t select_scratch_words(const DevPlan& plan, int nreg, int rpt);
int64_t string_tiles(int64_t capacity);
int batch_tile_rows(int nreg, int rpt);
hipError_t launch_batches(const DevPlan* plans, const int64_t* tile_start, int n_batches, int64_t n_items,
                          int nreg, int rpt, hipStream_t stream);
hipError_t launch_string_output(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                                const uint8_t* src_chars, uint8_t* dst_chars, const int64_t* d_count,
                                int64_t count, int64_t capacity, uint64_t* status, hipStream_t stream);
hipError_t launch_compact(const DevPlan& plan, const DevPlan* d_plan, int nreg, int sel_rpt, int rpt,
                          uint64_t* sel, uint32_t* idx, int64_t* d_count, int64_t capacity,
                          hipStream_t stream, int phases);

namespace {
thread_local std::string g_last_error;
}  // namespace

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int device_error(hipError_t e, const char* where) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(BSQL_ERR_OOM, std::string(where) + ": out of device memory");
  return set_error(BSQL_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

namespace {

// Per-stream scratch: the device copy of the launch plan, then the compaction
// tile status words + ticket. One buffer per stream, so that concurrent calls
// on different streams never share it.
constexpr int64_t kPlanWords = (int64_t)((sizeof(DevPlan) + 255) / 256 * 256 / 8);
struct Scratch {
  uint64_t* p = nullptr;
  int64_t words = 0;
  std::vector<char> plan;  // the plan bytes last copied to p (stream order)
};
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;

hipError_t get_scratch(hipStream_t stream, int64_t words, uint64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Scratch& s = g_scratch[{dev, stream}];
  if (!s.p || s.words < words) {
    if (s.p) {
      // the previous buffer may still be read by queued work on this stream
      e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return e;
      (void)hipFree(s.p);
      s.plan.clear();
      s.p = nullptr;
      s.words = 0;
    }
    const int64_t n = (words < 4096 ? 4096 : words + words / 4) + kPlanWords;
    e = hipMalloc((void**)&s.p, (size_t)n * sizeof(uint64_t));
    if (e != hipSuccess) return e;
    s.words = n - kPlanWords;
  }
  *out = s.p;
  return hipSuccess;
}

// Copies the launch plan into scratch (stream-ordered). The stream's own
// scratch remembers the plan last copied there: every later copy to it is
// queued on the same stream, so an identical plan (the same program over the
// same columns again) is already in place and is not copied again.
hipError_t upload_plan(const DevPlan& plan, uint64_t* scratch, hipStream_t stream) {
  const char* bytes = (const char*)&plan;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find({dev, stream});
  Scratch* own = it != g_scratch.end() && it->second.p == scratch ? &it->second : nullptr;
  if (own && own->plan.size() == sizeof(DevPlan) && std::memcmp(own->plan.data(), bytes, sizeof(DevPlan)) == 0)
    return hipSuccess;
  const hipError_t e = hipMemcpyAsync(scratch, &plan, sizeof(DevPlan), hipMemcpyHostToDevice, stream);
  if (own) {
    if (e == hipSuccess) own->plan.assign(bytes, bytes + sizeof(DevPlan));
    else own->plan.clear();
  }
  return e;
}

// string literals -> one device buffer of {int32 len, bytes} records (LIKE
// patterns: their compiled form)
hipError_t ensure_strings(Program& p) {
  if (p.d_strings || !p.str_offsets.empty()) return hipSuccess;
  std::vector<char> blob;
  std::vector<int64_t> offs(2 * p.ops.size(), -1);
  for (size_t i = 0; i < p.ops.size(); ++i) {
    for (int side = 0; side < 2; ++side) {
      const Scalar& s = side ? p.rsc[i] : p.lsc[i];
      const int pos = side ? p.right[i] : p.left[i];
      if (pos != BSQL_SCALAR_INDEX || s.type_id != BSQL_STRING) continue;
      while (blob.size() % 8) blob.push_back(0);
      offs[2 * i + side] = (int64_t)blob.size();
      if (side == 1 && p.ops[i] == BSQL_OP_STR_LIKE) {
        const std::vector<int32_t> units = like_compile(s.str);
        blob.insert(blob.end(), (const char*)units.data(), (const char*)(units.data() + units.size()));
        continue;
      }
      if (side == 1 && (p.ops[i] == BSQL_OP_TO_DATE || p.ops[i] == BSQL_OP_TO_TIMESTAMP)) {
        const std::vector<int32_t> items = timestamp_format_compile(s.str);
        blob.insert(blob.end(), (const char*)items.data(), (const char*)(items.data() + items.size()));
        continue;
      }
      int32_t len = (int32_t)s.str.size();
      blob.insert(blob.end(), (const char*)&len, (const char*)&len + 4);
      blob.insert(blob.end(), s.str.begin(), s.str.end());
    }
  }
  if (blob.empty()) {
    p.str_offsets = offs;
    return hipSuccess;
  }
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, blob.size());
  if (e != hipSuccess) return e;
  e = hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) { (void)hipFree(d); return e; }
  p.d_strings = d;
  p.str_offsets = offs;
  return hipSuccess;
}

DevInCol to_dev_in(const bsql_column& c) {
  DevInCol d{};
  d.type = (int16_t)c.type_id;
  if (c.type_id == BSQL_STRING) {
    d.data = c.data;
    d.offsets = c.offsets ? c.offsets + c.offset : nullptr;
  } else {
    d.data = (const char*)c.data + c.offset * type_size(c.type_id);
  }
  if (c.validity) {
    d.valid = c.validity + (c.offset >> 5);
    d.vbit = (int32_t)(c.offset & 31);
  }
  d.aligned16 = ((((uintptr_t)d.data) & 15) == 0) ? 1 : 0;
  return d;
}

void check_inputs(const bsql_column* in, int n_in, int64_t num_rows) {
  if (n_in < 0 || (n_in > 0 && !in)) fail("invalid input table");
  if (num_rows < 0) fail("negative row count");
  for (int k = 0; k < n_in; ++k) {
    const bsql_column& c = in[k];
    if (c.type_id == BSQL_STRING) {
      if (num_rows > 0 && (!c.offsets || !c.data)) fail("string column without offsets/chars");
    } else if (type_size(c.type_id) == 0) {
      fail("unsupported input column type " + std::to_string(c.type_id));
    } else if (num_rows > 0 && !c.data) {
      fail("input column " + std::to_string(k) + " has no data");
    }
    if (c.size < num_rows) fail("input column " + std::to_string(k) + " is shorter than num_rows");
  }
}

std::string cache_key(int mode, const bsql_program* pred, const bsql_program* proj,
                      const bsql_column* in, int n_in, const bsql_column* out, int n_out) {
  std::string k = std::to_string(mode) + "|" + std::to_string((uintptr_t)pred) + "|" +
                  std::to_string((uintptr_t)proj) + "|";
  for (int i = 0; i < n_in; ++i) k += std::to_string(in[i].type_id) + ",";
  k += "|";
  for (int i = 0; i < n_out; ++i) k += std::to_string(out ? out[i].type_id : 0) + ",";
  if (pred) {
    k += "|o";
    for (int32_t c : pred->conj_order) k += std::to_string(c) + ",";
    k += "|f" + std::to_string(pred->launch.flags);
  }
  return k;
}

// Output column checks shared by the launching entries (rows = capacity).
void check_outputs(const bsql_column* out, int n_out, int64_t rows) {
  for (int k = 0; k < n_out; ++k) {
    const bsql_column& c = out[k];
    if (c.offset != 0) fail("output columns must have offset 0");
    if (c.size < rows) fail("output column " + std::to_string(k) + " holds fewer rows than requested");
    if (c.type_id == BSQL_STRING) {
      if (!c.offsets) fail("string output column without offsets");
      if (c.chars_size < 0 || (c.chars_size > 0 && !c.data)) fail("string output column without chars");
    } else {
      if (type_size(c.type_id) == 0) fail("unsupported output column type");
      if (rows > 0 && !c.data) fail("output column has no data");
    }
  }
}

// String outputs: the VM stores slices into scratch; each then needs the
// chars buffer of its source column to fit in the output's chars.
int count_string_outputs(const Lowered& L, const bsql_column* in, const bsql_column* out, int n_out,
                         bool check = true) {
  int n = 0;
  for (int k = 0; k < n_out; ++k) {
    const int src = L.out_str_col[(size_t)k];
    if (src < 0) continue;
    if (check && out[k].chars_size < in[src].chars_size)
      fail("string output " + std::to_string(k) + ": chars_size " + std::to_string(out[k].chars_size) +
           " is smaller than its source column's " + std::to_string(in[src].chars_size));
    ++n;
  }
  return n;
}

hipError_t finish_string_outputs(const Lowered& L, const bsql_column* in, const bsql_column* out, int n_out,
                                 uint64_t* slices, const int64_t* d_count, int64_t count, int64_t capacity,
                                 uint64_t* status, hipStream_t stream) {
  int j = 0;
  for (int k = 0; k < n_out; ++k) {
    const int src = L.out_str_col[(size_t)k];
    if (src < 0) continue;
    hipError_t e = launch_string_output(slices + (size_t)j * (size_t)capacity, out[k].validity, out[k].offsets,
                                        (const uint8_t*)in[src].data, (uint8_t*)out[k].data, d_count, count,
                                        capacity, status, stream);
    if (e != hipSuccess) return e;
    ++j;
  }
  return hipSuccess;
}

const Lowered& lowered_for(bsql_program* owner, const LowerOptions& opt, const bsql_column* in,
                           int n_in, const bsql_column* out, int n_out) {
  std::string key = cache_key((int)opt.mode, (const bsql_program*)opt.pred,
                              (const bsql_program*)opt.proj, in, n_in, out, n_out);
  std::lock_guard<std::mutex> lk(owner->cache_mu);
  auto it = owner->cache.find(key);
  if (it != owner->cache.end()) return it->second;
  std::vector<int32_t> in_types(n_in), out_types(n_out);
  for (int i = 0; i < n_in; ++i) in_types[i] = in[i].type_id;
  for (int i = 0; i < n_out; ++i) out_types[i] = out ? out[i].type_id : 0;
  Lowered l = lower(opt, in_types.data(), n_in, out_types.data(), n_out);
  return owner->cache.emplace(key, l).first->second;
}

int rows_per_thread(const bsql_program* p) { return p ? p->launch.rows_per_thread : 0; }

}  // namespace
}  // namespace bsql

using namespace bsql;

namespace bsql {

// Filter (+ project) launches; `phases` as launch_compact's. Scratch is
// sized for num_rows whatever the capacity, so that a selection phase and a
// later output phase on the same stream find the same layout.
int run_compact(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in,
                bsql_column* out, int n_out, int64_t num_rows, int64_t capacity,
                int64_t* d_count, void* stream, int phases) {
  if (!pred) fail("null predicate");
  if (!d_count) fail("null count pointer");
  check_inputs(in, n_in, num_rows);
  // step columns: the predicate's after the table, the projection's after those
  StepColumns sc;
  if (!pred->steps.empty() || (proj && !proj->steps.empty())) {
    sc.stream = (hipStream_t)stream;
    sc.cols.assign(in, in + n_in);
    build_steps(*pred, sc, num_rows);
    if (proj && !proj->steps.empty()) {
      proj = rebased_program(proj, (int)sc.cols.size());
      build_steps(*proj, sc, num_rows);
    }
    in = sc.cols.data();
    n_in = (int)sc.cols.size();
  }
  // selected rows are listed as 32-bit row ids
  if (num_rows > (int64_t)UINT32_MAX) fail("a filter handles at most 2^32-1 rows per call");
  const bool outputs = (phases & PHASE_OUTPUT) != 0;
  if (outputs) check_outputs(out, n_out, capacity);
  hipError_t e = ensure_strings(*pred);
  if (e == hipSuccess && proj) e = ensure_strings(*proj);
  if (e != hipSuccess) return device_error(e, "filter");
  LowerOptions opt;
  opt.mode = proj ? LowerOptions::FILTER_PROJECT
                  : (n_out == 0 && !out ? LowerOptions::FILTER_COUNT : LowerOptions::FILTER_GATHER);
  opt.pred = pred;
  opt.proj = proj;
  opt.conj_order = &pred->conj_order;
  opt.ranges = !(pred->launch.flags & BSQL_LAUNCH_NO_RANGES);
  opt.lazy_loads = (pred->launch.flags & BSQL_LAUNCH_LAZY_LOADS) != 0;
  const Lowered& L = lowered_for(pred, opt, in, n_in, out, n_out);
  DevPlan plan = L.plan;
  plan.num_rows = num_rows;
  for (int k = 0; k < n_in; ++k) plan.in[k] = to_dev_in(in[k]);
  const int n_str = count_string_outputs(L, in, out, n_out, outputs);
  // the predicate's launch shape drives the selection kernel, the
  // projection's (or the predicate's, for a plain filter) the output kernel
  const int sel_rpt = rows_per_thread(pred);
  const int rpt = rows_per_thread(proj ? proj : pred);
  const int64_t sel_words = select_scratch_words(plan, L.nreg, sel_rpt);
  if (sel_words < 0) fail("no kernel variant for this program");
  // scratch: plan | selection ballots + slice counts/offsets | string
  // assembly tile status + ticket | selected row ids (uint32) | string
  // output slices
  const int64_t st_words = n_str ? string_tiles(num_rows) + 1 : 0;
  const int64_t id_words = (num_rows + 1) / 2;
  uint64_t* scratch = nullptr;
  e = get_scratch((hipStream_t)stream, sel_words + st_words + id_words + (int64_t)n_str * num_rows, &scratch);
  if (e != hipSuccess) return device_error(e, "filter scratch");
  uint64_t* sel = scratch + kPlanWords;
  uint64_t* status = sel + sel_words;
  uint32_t* ids = (uint32_t*)(status + st_words);
  uint64_t* slices = status + st_words + id_words;
  for (int k = 0, j = 0; k < n_out; ++k) {
    const bool str = L.out_str_col[(size_t)k] >= 0;
    plan.out[k].data = str ? (void*)(slices + (size_t)(j++) * (size_t)capacity) : (outputs ? out[k].data : nullptr);
    plan.out[k].valid = outputs ? out[k].validity : nullptr;
    plan.out[k].type = out[k].type_id;
  }
  e = upload_plan(plan, scratch, (hipStream_t)stream);
  if (e == hipSuccess)
    e = launch_compact(plan, (const DevPlan*)scratch, L.nreg, sel_rpt, rpt, sel, ids, d_count, capacity,
                       (hipStream_t)stream, phases);
  if (e == hipSuccess && outputs)
    e = finish_string_outputs(L, in, out, n_out, slices, d_count, 0, capacity, status, (hipStream_t)stream);
  if (e != hipSuccess) return device_error(e, "filter");
  return BSQL_OK;
}

int execute_program(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                    int64_t num_rows, void* stream, bool private_scratch) {
  if (!prog) fail("null program");
  if (!prog->steps.empty() && n_in == prog->n_table) fail("the program's step columns were not built");
  check_inputs(in, n_in, num_rows);
  if (n_out != (int)prog->final_out.size()) fail("number of outputs does not match the program");
  check_outputs(out, n_out, num_rows);
  hipError_t e = ensure_strings(*prog);
  if (e != hipSuccess) return device_error(e, "bsql_interp_execute");
  LowerOptions opt;
  opt.mode = LowerOptions::EXECUTE;
  opt.proj = prog;
  const Lowered& L = lowered_for(prog, opt, in, n_in, out, n_out);
  DevPlan plan = L.plan;
  plan.num_rows = num_rows;
  for (int k = 0; k < n_in; ++k) plan.in[k] = to_dev_in(in[k]);
  const int n_str = count_string_outputs(L, in, out, n_out);
  // scratch: plan | string assembly status + ticket | one slice column per string output
  const int64_t st_words = n_str ? string_tiles(num_rows) + 1 : 0;
  const int64_t words = st_words + (int64_t)n_str * num_rows;
  uint64_t* scratch = nullptr;
  if (private_scratch) {
    e = hipMallocAsync((void**)&scratch, (size_t)(kPlanWords + std::max<int64_t>(words, 1)) * 8,
                       (hipStream_t)stream);
  } else {
    e = get_scratch((hipStream_t)stream, words, &scratch);
  }
  if (e != hipSuccess) return device_error(e, "bsql_interp_execute scratch");
  struct Release {
    uint64_t* p;
    hipStream_t s;
    ~Release() {
      if (p) (void)hipFreeAsync(p, s);
    }
  } release{private_scratch ? scratch : nullptr, (hipStream_t)stream};
  uint64_t* status = scratch + kPlanWords;
  uint64_t* slices = status + st_words;
  for (int k = 0, j = 0; k < n_out; ++k) {
    const bool str = L.out_str_col[(size_t)k] >= 0;
    plan.out[k].data = str ? (void*)(slices + (size_t)(j++) * (size_t)num_rows) : out[k].data;
    plan.out[k].valid = out[k].validity;
    plan.out[k].type = out[k].type_id;
  }
  if (num_rows == 0 && n_str == 0) return BSQL_OK;
  e = upload_plan(plan, scratch, (hipStream_t)stream);
  if (e == hipSuccess && num_rows > 0)
    e = launch_execute(plan, (const DevPlan*)scratch, L.nreg, rows_per_thread(prog), (hipStream_t)stream);
  if (e == hipSuccess)
    e = finish_string_outputs(L, in, out, n_out, slices, nullptr, num_rows, num_rows, status,
                              (hipStream_t)stream);
  if (e != hipSuccess) return device_error(e, "bsql_interp_execute");
  return BSQL_OK;
}

}  // namespace bsql

extern "C" {

const char* bsql_last_error(void) { return g_last_error.c_str(); }
int bsql_version(void) { return 1; }

int bsql_program_from_arrays(const int16_t* left, const int16_t* right, const int16_t* out,
                             const int32_t* ops, int n_ops, const int16_t* final_out, int n_final,
                             const bsql_scalar* left_scalars, const bsql_scalar* right_scalars,
                             bsql_program** prog) {
  BSQL_TRY({
    if (!prog || n_ops < 0 || n_final <= 0 || !final_out) fail("invalid arguments");
    if (n_ops > 0 && (!left || !right || !out || !ops)) fail("missing program arrays");
    std::unique_ptr<bsql_program> p(new bsql_program());
    p->left.assign(left, left + n_ops);
    p->right.assign(right, right + n_ops);
    p->out.assign(out, out + n_ops);
    p->ops.assign(ops, ops + n_ops);
    p->final_out.assign(final_out, final_out + n_final);
    p->lsc.resize(n_ops);
    p->rsc.resize(n_ops);
    for (int i = 0; i < n_ops; ++i) {
      for (int side = 0; side < 2; ++side) {
        const int pos = side ? right[i] : left[i];
        if (pos != BSQL_SCALAR_INDEX) continue;
        const bsql_scalar* arr = side ? right_scalars : left_scalars;
        if (!arr) fail("op " + std::to_string(i) + " has a scalar operand but no scalar array");
        const bsql_scalar& s = arr[i];
        Scalar& d = side ? p->rsc[i] : p->lsc[i];
        d.type_id = s.type_id;
        d.is_null = false;
        d.i64 = s.v.i64;
        d.f64 = s.v.f64;
        if (s.type_id == BSQL_STRING) d.str.assign(s.str ? s.str : "", s.str ? (size_t)s.str_len : 0);
      }
    }
    validate(*p, 0);
    *prog = p.release();
    return BSQL_OK;
  })
}

int bsql_program_compile(const char* const* exprs, int n_exprs, const int32_t* in_types, int n_in,
                         bsql_program** prog) {
  BSQL_TRY({
    if (!prog || !exprs || n_exprs <= 0 || n_in < 0 || (n_in > 0 && !in_types))
      fail("invalid arguments");
    std::vector<std::string> v;
    for (int i = 0; i < n_exprs; ++i) {
      if (!exprs[i]) fail("null expression");
      v.emplace_back(exprs[i]);
    }
    std::unique_ptr<Program> p = compile_expressions(v, in_types, n_in);
    std::unique_ptr<bsql_program> bp(new bsql_program());
    static_cast<Program&>(*bp) = std::move(*p);
    *prog = bp.release();
    return BSQL_OK;
  })
}

void bsql_program_free(bsql_program* prog) { delete prog; }

int bsql_program_num_ops(const bsql_program* p) { return p ? (int)p->ops.size() : -1; }
int bsql_program_num_outputs(const bsql_program* p) { return p ? (int)p->final_out.size() : -1; }
int bsql_program_num_inputs(const bsql_program* p) { return p ? p->n_inputs : -1; }

int bsql_program_arrays(const bsql_program* p, int16_t* left, int16_t* right, int16_t* out,
                        int32_t* ops, int16_t* final_out) {
  if (!p) return set_error(BSQL_ERR_INVALID, "null program");
  const size_t n = p->ops.size();
  if (left) std::memcpy(left, p->left.data(), n * 2);
  if (right) std::memcpy(right, p->right.data(), n * 2);
  if (out) std::memcpy(out, p->out.data(), n * 2);
  if (ops) std::memcpy(ops, p->ops.data(), n * 4);
  if (final_out) std::memcpy(final_out, p->final_out.data(), p->final_out.size() * 2);
  return BSQL_OK;
}

int bsql_program_scalar(const bsql_program* p, int op, int side, bsql_scalar* out) {
  if (!p || !out || op < 0 || op >= (int)p->ops.size())
    return set_error(BSQL_ERR_INVALID, "invalid arguments");
  const Scalar& s = side ? p->rsc[op] : p->lsc[op];
  const int pos = side ? p->right[op] : p->left[op];
  std::memset(out, 0, sizeof(*out));
  out->type_id = pos == BSQL_SCALAR_INDEX ? s.type_id : BSQL_EMPTY;
  if (is_float_type(s.type_id)) out->v.f64 = s.f64; else out->v.i64 = s.i64;
  if (s.type_id == BSQL_STRING) {
    out->str = s.str.c_str();
    out->str_len = (int32_t)s.str.size();
  }
  return BSQL_OK;
}

int bsql_program_output_type(const bsql_program* p, const int32_t* in_types, int n_in, int k,
                             int32_t* type_id) {
  BSQL_TRY({
    if (!p || !type_id || k < 0 || k >= (int)p->final_out.size()) fail("invalid arguments");
    const std::vector<int32_t> ext = extended_types(*p, in_types, n_in);
    std::vector<int32_t> st = slot_types_for(*p, ext.data(), (int)ext.size());
    OpTypes t = propagate_types(*p, st.data(), (int)st.size());
    const int slot = p->final_out[k];
    if (slot < 0) {  // literal root
      *type_id = p->root_literals[(size_t)k].type_id;
      return BSQL_OK;
    }
    int32_t ty = slot < (int)st.size() ? st[(size_t)slot] : BSQL_EMPTY;
    for (size_t i = 0; i < p->ops.size(); ++i)
      if (p->out[i] == slot) ty = t.ot[i];
    *type_id = ty;
    return BSQL_OK;
  })
}

int bsql_program_used_columns(const bsql_program* p, int32_t* cols, int cap) {
  if (!p) return set_error(BSQL_ERR_INVALID, "null program");
  const int n = (int)p->used_columns.size();
  for (int i = 0; i < n && i < cap; ++i) cols[i] = p->used_columns[(size_t)i];
  return n;
}

int bsql_program_num_conjuncts(const bsql_program* pred) {
  if (!pred) return set_error(BSQL_ERR_INVALID, "null program") * -1;
  return count_conjuncts(*pred);
}

int bsql_program_set_conjunct_order(bsql_program* pred, const int32_t* order, int n_conj) {
  BSQL_TRY({
    if (!pred) fail("null program");
    if (!order) { pred->conj_order.clear(); return BSQL_OK; }
    const int n = count_conjuncts(*pred);
    if (n_conj != n) fail("the predicate has " + std::to_string(n) + " conjuncts");
    std::vector<int> seen(n, 0);
    for (int i = 0; i < n_conj; ++i) {
      if (order[i] < 0 || order[i] >= n || seen[order[i]]++) fail("order is not a permutation");
    }
    pred->conj_order.assign(order, order + n_conj);
    return BSQL_OK;
  })
}

int bsql_program_set_launch(bsql_program* p, const bsql_launch_config* cfg) {
  if (!p || !cfg) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  const int r = cfg->rows_per_thread;
  if (r != 0 && r != 1 && r != 2 && r != 4 && r != 8 && r != 16 && r != 32)
    return set_error(BSQL_ERR_INVALID, "rows_per_thread must be 0, 1, 2, 4, 8, 16 or 32");
  if (cfg->block_size != 0 && cfg->block_size != 256)
    return set_error(BSQL_ERR_INVALID, "block_size must be 0 or 256");
  if (cfg->flags & ~(BSQL_LAUNCH_NO_RANGES | BSQL_LAUNCH_LAZY_LOADS))
    return set_error(BSQL_ERR_INVALID, "unknown launch flags");
  std::lock_guard<std::mutex> lk(p->cache_mu);
  p->launch = *cfg;
  return BSQL_OK;
}

int bsql_interp_execute(const bsql_program* cprog, const bsql_column* in, int n_in,
                        bsql_column* out, int n_out, int64_t num_rows, void* stream) {
  BSQL_TRY({
    bsql_program* prog = const_cast<bsql_program*>(cprog);
    if (!prog) fail("null program");
    if (prog->steps.empty()) return execute_program(prog, in, n_in, out, n_out, num_rows, stream, false);
    check_inputs(in, n_in, num_rows);
    StepColumns sc;
    sc.stream = (hipStream_t)stream;
    sc.cols.assign(in, in + n_in);
    build_steps(*prog, sc, num_rows);
    return execute_program(prog, sc.cols.data(), (int)sc.cols.size(), out, n_out, num_rows, stream, false);
  })
}

int bsql_filter(const bsql_program* pred, const bsql_column* in, int n_in, bsql_column* out,
                int64_t num_rows, int64_t* d_count, void* stream) {
  BSQL_TRY({
    return run_compact(const_cast<bsql_program*>(pred), nullptr, in, n_in, out, n_in, num_rows,
                       num_rows, d_count, stream, PHASE_ALL);
  })
}

int bsql_filter_count(const bsql_program* pred, const bsql_column* in, int n_in, int64_t num_rows,
                      int64_t* d_count, void* stream) {
  BSQL_TRY({
    return run_compact(const_cast<bsql_program*>(pred), nullptr, in, n_in, nullptr, 0, num_rows, num_rows,
                       d_count, stream, PHASE_ALL);
  })
}

int bsql_filter_project_sized(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                              int n_in, bsql_column* out, int n_out, int64_t num_rows, int64_t capacity,
                              int64_t* d_count, void* stream) {
  BSQL_TRY({
    if (!proj) fail("null projection");
    if (capacity < 0 || capacity > num_rows) fail("capacity must be in [0, num_rows]");
    return run_compact(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in, n_in, out, n_out,
                       num_rows, capacity, d_count, stream, PHASE_ALL);
  })
}

int bsql_filter_project(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                        int n_in, bsql_column* out, int n_out, int64_t num_rows, int64_t* d_count,
                        void* stream) {
  BSQL_TRY({
    if (!proj) fail("null projection");
    return run_compact(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in, n_in,
                       out, n_out, num_rows, num_rows, d_count, stream, PHASE_ALL);
  })
}

}  // extern "C"

// ------------------------------------------------------------------ batches

struct bsql_batch {
  std::vector<DevPlan> plans;  // host copies (columns filled in)
  std::vector<int64_t> rows;
  int nreg = 1;
  bsql_launch_config launch{0, 0, 0};
  DevPlan* d_plans = nullptr;
  int64_t* d_tile_start = nullptr;
  int64_t n_items = 0;
  ~bsql_batch() {
    if (d_plans) (void)hipFree(d_plans);
    if (d_tile_start) (void)hipFree(d_tile_start);
  }
};

namespace {

// Work map for the current launch shape: prefix counts of tiles per batch.
hipError_t upload_work_map(bsql_batch* b) {
  const int tile = batch_tile_rows(b->nreg, b->launch.rows_per_thread);
  if (tile <= 0) fail("no kernel variant for these programs");
  std::vector<int64_t> start(b->plans.size() + 1, 0);
  for (size_t i = 0; i < b->plans.size(); ++i) start[i + 1] = start[i] + (b->rows[i] + tile - 1) / tile;
  b->n_items = start.back();
  if (!b->d_tile_start) {
    hipError_t e = hipMalloc((void**)&b->d_tile_start, start.size() * sizeof(int64_t));
    if (e != hipSuccess) return e;
  }
  return hipMemcpy(b->d_tile_start, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice);
}

}  // namespace

extern "C" {

int bsql_batch_prepare(const bsql_batch_desc* descs, int n, bsql_batch** out) {
  BSQL_TRY({
    if (!descs || n <= 0 || !out) fail("invalid arguments");
    auto b = std::make_unique<bsql_batch>();
    for (int i = 0; i < n; ++i) {
      const bsql_batch_desc& d = descs[i];
      bsql_program* prog = const_cast<bsql_program*>(d.prog);
      if (!prog) fail("batch " + std::to_string(i) + " has no program");
      if (!prog->steps.empty())
        fail("batch " + std::to_string(i) + ": programs that build string columns do not run batched");
      check_inputs(d.in, d.n_in, d.num_rows);
      if (d.n_out != (int)prog->final_out.size()) fail("batch " + std::to_string(i) + ": outputs do not match");
      check_outputs(d.out, d.n_out, d.num_rows);
      for (int k = 0; k < d.n_out; ++k)
        if (d.out[k].type_id == BSQL_STRING) fail("batched launches write fixed-width outputs only");
      hipError_t e = ensure_strings(*prog);
      if (e != hipSuccess) return device_error(e, "bsql_batch_prepare");
      LowerOptions opt;
      opt.mode = LowerOptions::EXECUTE;
      opt.proj = prog;
      const Lowered& L = lowered_for(prog, opt, d.in, d.n_in, d.out, d.n_out);
      DevPlan plan = L.plan;
      plan.num_rows = d.num_rows;
      for (int k = 0; k < d.n_in; ++k) plan.in[k] = to_dev_in(d.in[k]);
      for (int k = 0; k < d.n_out; ++k) {
        plan.out[k].data = d.out[k].data;
        plan.out[k].valid = d.out[k].validity;
        plan.out[k].type = d.out[k].type_id;
      }
      b->plans.push_back(plan);
      b->rows.push_back(d.num_rows);
      b->nreg = std::max(b->nreg, L.nreg);
    }
    hipError_t e = hipMalloc((void**)&b->d_plans, b->plans.size() * sizeof(DevPlan));
    if (e == hipSuccess)
      e = hipMemcpy(b->d_plans, b->plans.data(), b->plans.size() * sizeof(DevPlan), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = upload_work_map(b.get());
    if (e != hipSuccess) return device_error(e, "bsql_batch_prepare");
    *out = b.release();
    return BSQL_OK;
  })
}

int bsql_batch_launch(const bsql_batch* b, void* stream) {
  BSQL_TRY({
    if (!b) fail("null batch");
    hipError_t e = launch_batches(b->d_plans, b->d_tile_start, (int)b->plans.size(), b->n_items, b->nreg,
                                  b->launch.rows_per_thread, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_batch_launch");
    return BSQL_OK;
  })
}

int bsql_batch_set_launch(bsql_batch* b, const bsql_launch_config* cfg) {
  BSQL_TRY({
    if (!b || !cfg) fail("invalid arguments");
    const int r = cfg->rows_per_thread;
    if (r != 0 && r != 1 && r != 2 && r != 4) fail("rows_per_thread must be 0, 1, 2 or 4");
    b->launch = *cfg;
    hipError_t e = upload_work_map(b);
    if (e != hipSuccess) return device_error(e, "bsql_batch_set_launch");
    return BSQL_OK;
  })
}

int64_t bsql_batch_total_rows(const bsql_batch* b) {
  if (!b) return -1;
  int64_t t = 0;
  for (int64_t r : b->rows) t += r;
  return t;
}

void bsql_batch_free(bsql_batch* b) { delete b; }

}  // extern "C"

// ----------------------------------------------------------------- autotuner

namespace {

// Average milliseconds of `run` on `stream` over a few launches after one
// warm-up, timed with hipEvents on that stream.
template <class F>
hipError_t time_launches(hipStream_t stream, F run, float* ms) {
  hipEvent_t a, z;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&z);
  if (e != hipSuccess) { (void)hipEventDestroy(a); return e; }
  constexpr int REPS = 3;
  e = run();
  if (e == hipSuccess) e = hipEventRecord(a, stream);
  for (int i = 0; i < REPS && e == hipSuccess; ++i) e = run();
  if (e == hipSuccess) e = hipEventRecord(z, stream);
  if (e == hipSuccess) e = hipEventSynchronize(z);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, z);
  *ms = t / REPS;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
  return e;
}

const int kShapes[] = {1, 2, 4};  // rows per lane the VM variants cover

}  // namespace

extern "C" {

int bsql_batch_autotune(bsql_batch* b, void* stream, bsql_launch_config* chosen) {
  BSQL_TRY({
    if (!b) fail("null batch");
    float best = 1e30f;
    bsql_launch_config best_cfg = b->launch;
    for (int r : kShapes) {
      if (batch_tile_rows(b->nreg, r) != 256 * r) continue;  // no variant of this shape
      bsql_launch_config cfg{0, r, 0};
      b->launch = cfg;
      hipError_t e = upload_work_map(b);
      float ms = 0.f;
      if (e == hipSuccess)
        e = time_launches((hipStream_t)stream, [&] {
          return launch_batches(b->d_plans, b->d_tile_start, (int)b->plans.size(), b->n_items, b->nreg, r,
                                (hipStream_t)stream);
        }, &ms);
      if (e != hipSuccess) return device_error(e, "bsql_batch_autotune");
      if (ms < best) { best = ms; best_cfg = cfg; }
    }
    b->launch = best_cfg;
    hipError_t e = upload_work_map(b);
    if (e != hipSuccess) return device_error(e, "bsql_batch_autotune");
    if (chosen) *chosen = best_cfg;
    return BSQL_OK;
  })
}

int bsql_program_autotune(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                          int64_t num_rows, void* stream, bsql_launch_config* chosen) {
  BSQL_TRY({
    if (!prog) fail("null program");
    float best = 1e30f;
    bsql_launch_config best_cfg = prog->launch;
    for (int r : kShapes) {
      bsql_launch_config cfg = prog->launch;
      cfg.rows_per_thread = r;
      prog->launch = cfg;
      float ms = 0.f;
      int rc = BSQL_OK;
      hipError_t e = time_launches((hipStream_t)stream, [&] {
        rc = bsql_interp_execute(prog, in, n_in, out, n_out, num_rows, stream);
        return rc == BSQL_OK ? hipSuccess : hipErrorUnknown;
      }, &ms);
      if (rc != BSQL_OK) return rc;
      if (e != hipSuccess) return device_error(e, "bsql_program_autotune");
      if (ms < best) { best = ms; best_cfg = cfg; }
    }
    prog->launch = best_cfg;
    if (chosen) *chosen = best_cfg;
    return BSQL_OK;
  })
}

}  // extern "C"
