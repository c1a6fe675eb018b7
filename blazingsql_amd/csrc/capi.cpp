// capi.cpp -- the C ABI (include/bsql/interp.h) over the host program
// representation and the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
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

hipError_t launch_execute(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt, int block,
                          hipStream_t stream, bool lf);
int64_t select_scratch_words(const DevPlan& plan, int nreg, int rpt, int block, bool lf);
int64_t string_tiles(int64_t capacity);
int batch_tile_rows(int nreg, int rpt, int block, bool lf);
bool variant_fits(int nreg, int rpt, int block, int n_spill, bool lf = false);
int variant_rows(int nreg, int rpt, int block, int n_spill, bool lf = false);
hipError_t launch_batches(const DevPlan* plans, const int64_t* tile_start, int n_batches, int64_t n_items,
                          int nreg, int rpt, int block, int n_spill, bool lf, hipStream_t stream);
hipError_t launch_string_output(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                                const uint8_t* src_chars, uint8_t* dst_chars, const int64_t* d_count,
                                int64_t count, int64_t capacity, uint64_t* status, hipStream_t stream);
hipError_t launch_compact(const DevPlan& plan, const DevPlan* d_plan, int nreg, int sel_rpt, int rpt, int block,
                          uint64_t* sel, uint32_t* idx, int64_t* d_count, int64_t capacity,
                          hipStream_t stream, int phases, bool fused, bool lf);
bool fused_eligible(const DevPlan& plan, int nreg);
int64_t fused_scratch_words(int64_t num_rows);
// kernel_gen.cpp
bool generated_filter_project_eligible(const DevPlan& plan, const std::vector<DevOp>& ops);
hipError_t launch_generated_filter_project_batch(const std::vector<DevPlan>& plans,
                                                 const std::vector<std::vector<DevOp>>& ops,
                                                 const std::vector<std::vector<int>>& in_maps,
                                                 const std::vector<const DevPlan*>& d_plans, uint64_t* status,
                                                 int64_t* d_counts, const std::vector<int64_t>& capacity,
                                                 hipStream_t stream);
hipError_t launch_generated_filter_project(const DevPlan& plan, const std::vector<DevOp>& ops,
                                           const DevPlan* d_plan, uint64_t* status, int64_t* d_count,
                                           int64_t capacity, hipStream_t stream);
int64_t generated_filter_project_scratch_words(int64_t num_rows);
bool generated_execute_eligible(const DevPlan& plan, const std::vector<DevOp>& ops);
hipError_t launch_generated_execute(const DevPlan& plan, const std::vector<DevOp>& ops, const DevPlan* d_plan,
                                    int nreg, int rpt, int block, hipStream_t stream);
void generated_stats(int64_t* kernels, int64_t* launches, double* compile_seconds);
void set_generated_check(const char* arch);
bool generated_check_mode();
int64_t generated_check_bytes();

namespace {
thread_local std::string g_last_error;
std::atomic<uint64_t> g_next_id{1};
}  // namespace

uint64_t next_program_id() { return g_next_id.fetch_add(1); }

namespace {
std::atomic<bool> g_exiting{false};
void mark_exiting() { g_exiting = true; }
// Registered when the library is loaded, after the HIP runtime it links
// against, so it runs before that runtime's own exit handlers.
__attribute__((constructor)) void register_exit_flag() { std::atexit(mark_exiting); }
}  // namespace

bool process_exiting() { return g_exiting.load(); }

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int device_error(hipError_t e, const char* where) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(BSQL_ERR_OOM, std::string(where) + ": out of device memory");
  return set_error(BSQL_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

void reject_capture(void* stream, const char* entry) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing((hipStream_t)stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
    fail(std::string(entry) +
         " cannot be captured in a graph (it may grow its scratch and copies its launch plan from host "
         "memory); bsql_filter_project_noalloc can");
}

namespace {

// Per-stream scratch, one buffer per (device, stream) so that concurrent
// calls on different streams never share it. Layout of a filter call:
// [selection ballots, slice counts, block totals | selected row ids (uint32)
// | string assembly status | string output slices | launch plan + ops]; the
// first two depend only on the predicate and the row count, so a selection
// and several output launches of the same call find them in place.
// Growing is stream-ordered (hipFreeAsync / hipMallocAsync on the stream):
// work still queued on the stream keeps reading the old buffer, and the host
// never waits.
struct Scratch {
  uint64_t* p = nullptr;
  int64_t words = 0;
  std::vector<char> plan;  // the plan blob last copied into this buffer (stream order)
  int64_t plan_at = -1;    // where (words)
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
      e = hipFreeAsync(s.p, stream);
      if (e != hipSuccess) return e;
      s.p = nullptr;
      s.words = 0;
      s.plan.clear();
      s.plan_at = -1;
    }
    const int64_t n = words < 4096 ? 4096 : words + words / 4;
    e = hipMallocAsync((void**)&s.p, (size_t)n * sizeof(uint64_t), stream);
    if (e != hipSuccess) return e;
    s.words = n;
  }
  *out = s.p;
  return hipSuccess;
}

// Words of a plan blob: the DevPlan, then its ops.
int64_t plan_words(size_t n_ops) {
  return (int64_t)((sizeof(DevPlan) + n_ops * sizeof(DevOp) + 255) / 256 * 256 / 8);
}

// The plan blob (DevPlan with its op pointer aimed at `at`, then the ops).
std::vector<char> plan_blob(const DevPlan& plan, const std::vector<DevOp>& ops, const void* at) {
  std::vector<char> blob(sizeof(DevPlan) + ops.size() * sizeof(DevOp));
  DevPlan p = plan;
  p.ops = (const DevOp*)((const char*)at + sizeof(DevPlan));
  std::memcpy(blob.data(), &p, sizeof(DevPlan));
  std::memcpy(blob.data() + sizeof(DevPlan), ops.data(), ops.size() * sizeof(DevOp));
  return blob;
}

// Copies the plan blob to `dst` in the stream's scratch (stream-ordered). The
// scratch remembers the blob last copied there: every later copy to it is
// queued on the same stream, so an identical blob (the same program over the
// same columns again) is already in place and is not copied again.
hipError_t upload_plan(const DevPlan& plan, const std::vector<DevOp>& ops, uint64_t* scratch, int64_t at,
                       hipStream_t stream) {
  uint64_t* dst = scratch + at;
  std::vector<char> blob = plan_blob(plan, ops, dst);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find({dev, stream});
  Scratch* own = it != g_scratch.end() && it->second.p == scratch ? &it->second : nullptr;
  if (own && own->plan_at == at && own->plan == blob) return hipSuccess;
  const hipError_t e = hipMemcpyAsync(dst, blob.data(), blob.size(), hipMemcpyHostToDevice, stream);
  if (own) {
    if (e == hipSuccess) {
      own->plan = std::move(blob);
      own->plan_at = at;
    } else {
      own->plan.clear();
      own->plan_at = -1;
    }
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

}  // namespace

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

namespace {

uint64_t uid_of(const Program* p) { return p ? static_cast<const bsql_program*>(p)->uid : 0; }

std::string cache_key(const LowerOptions& opt, const bsql_column* in, int n_in, const bsql_column* out,
                      int n_out) {
  std::string k = std::to_string((int)opt.mode) + "|" + std::to_string(uid_of(opt.pred)) + "|" +
                  std::to_string(uid_of(opt.proj)) + "|b" + std::to_string(opt.reg_budget) + "|";
  for (int i = 0; i < n_in; ++i) k += std::to_string(in[i].type_id) + ",";
  k += "|";
  for (int i = 0; i < n_out; ++i) k += std::to_string(out ? out[i].type_id : 0) + ",";
  if (opt.out_subset) {
    k += "|s";
    for (int c : *opt.out_subset) k += std::to_string(c) + ",";
  }
  if (opt.pred) {
    k += "|o";
    for (int32_t c : opt.pred->conj_order) k += std::to_string(c) + ",";
    k += "|f" + std::to_string(opt.pred->launch.flags);
  }
  return k;
}

}  // namespace

// Output column checks shared by the launching entries (rows = capacity).
void check_outputs(const bsql_column* out, int n_out, int64_t rows) {
  if (n_out < 0 || (n_out > 0 && !out)) fail("invalid output columns");
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

namespace {

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

}  // namespace

// The default register budget tries the 8-register file first (it runs 4
// rows per lane) and takes the 16-register one only if the program would
// spill there; an explicit budget (the wide spill layout) is kept.
Lowered lower_best(const LowerOptions& opt, const int32_t* in_types, int n_in, const int32_t* out_types,
                   int n_out) {
  if (opt.reg_budget >= BSQL_DEV_MAX_REG - 1) {
    LowerOptions small = opt;
    small.reg_budget = 7;
    Lowered l = lower(small, in_types, n_in, out_types, n_out);
    if (l.n_spill == 0) return l;
  }
  return lower(opt, in_types, n_in, out_types, n_out);
}

const Lowered& lowered_for(bsql_program* owner, const LowerOptions& opt, const bsql_column* in, int n_in,
                           const bsql_column* out, int n_out) {
  std::string key = cache_key(opt, in, n_in, out, n_out);
  std::lock_guard<std::mutex> lk(owner->cache_mu);
  auto it = owner->cache.find(key);
  if (it != owner->cache.end()) return it->second;
  std::vector<int32_t> in_types(n_in), out_types(n_out);
  for (int i = 0; i < n_in; ++i) in_types[i] = in[i].type_id;
  for (int i = 0; i < n_out; ++i) out_types[i] = out ? out[i].type_id : 0;
  return owner->cache.emplace(key, lower_best(opt, in_types.data(), n_in, out_types.data(), n_out)).first->second;
}

namespace {

int rows_per_thread(const bsql_program* p) { return p ? p->launch.rows_per_thread : 0; }
int block_size(const bsql_program* p) { return p && p->launch.block_size ? p->launch.block_size : 256; }
// The register file in LDS for a launch these programs drive
// (BSQL_LAUNCH_LDS_REGS on either).
bool lds_regs(const bsql_program* a, const bsql_program* b) {
  return (a && (a->launch.flags & BSQL_LAUNCH_LDS_REGS)) || (b && (b->launch.flags & BSQL_LAUNCH_LDS_REGS));
}

int reg_budget(const bsql_program* p) {
  return p && (p->launch.flags & BSQL_LAUNCH_SPILL_WIDE) ? 7 : BSQL_DEV_MAX_REG - 1;
}

// The launch shape a lowered plan runs with: the requested block size and
// rows per lane if a kernel variant holds its registers and spill slots,
// else the closest one that does (a spilling plan may need smaller
// workgroups for its LDS).
void launch_shape(const Lowered& L, int rpt_req, int block_req, int* rpt, int* block) {
  const int blocks[2] = {block_req, block_req == 256 ? 64 : 256};
  const int rpts[4] = {rpt_req, 0, 1, 2};
  for (int b : blocks)
    for (int r : rpts)
      if (variant_fits(L.nreg, r, b, L.n_spill)) {
        *rpt = r;
        *block = b;
        return;
      }
  fail("program needs " + std::to_string(L.n_spill) + " spill slots, more than a workgroup's LDS holds");
}

}  // namespace
}  // namespace bsql

using namespace bsql;

namespace bsql {

// Filter (+ project) launches; `phases` as launch_compact's. Scratch is
// sized for num_rows whatever the capacity, so that a selection phase and a
// later output phase on the same stream find the same layout.
int run_compact(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in,
                bsql_column* out, int n_out, int64_t num_rows, int64_t capacity,
                int64_t* d_count, void* stream, int phases, const std::vector<int>* out_subset) {
  if (!pred) fail("null predicate");
  if (!d_count) fail("null count pointer");
  check_inputs(in, n_in, num_rows);
  // step columns: the predicate's after the table, the projection's after those
  StepColumns sc;
  if (!pred->steps.empty() || (proj && !proj->steps.empty())) {
    if (generated_check_mode()) fail("a program with step columns has no generated kernel to check");
    sc.stream = (hipStream_t)stream;
    sc.cols.assign(in, in + n_in);
    if (!pred->steps.empty()) {
      if (phases & PHASE_SELECT) {
        build_steps(*pred, sc, num_rows);
      } else {
        // the output phase does not read the predicate's step columns:
        // placeholders keep the projection's numbering
        for (const Step& st : pred->steps) {
          bsql_column c;
          std::memset(&c, 0, sizeof(c));
          c.type_id = st.out_type;
          c.size = num_rows;
          sc.cols.push_back(c);
        }
      }
    }
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
  hipError_t e = hipSuccess;
  if (!generated_check_mode()) {
    e = ensure_strings(*pred);
    if (e == hipSuccess && proj) e = ensure_strings(*proj);
    if (e != hipSuccess) return device_error(e, "filter");
  }
  LowerOptions opt;
  opt.mode = proj ? LowerOptions::FILTER_PROJECT
                  : (n_out == 0 && !out ? LowerOptions::FILTER_COUNT : LowerOptions::FILTER_GATHER);
  opt.pred = pred;
  opt.proj = proj;
  opt.conj_order = &pred->conj_order;
  opt.ranges = !(pred->launch.flags & BSQL_LAUNCH_NO_RANGES);
  opt.lazy_loads = (pred->launch.flags & BSQL_LAUNCH_LAZY_LOADS) != 0;
  opt.out_subset = out_subset;
  opt.reg_budget = std::min(reg_budget(pred), reg_budget(proj));
  const Lowered& L = lowered_for(pred, opt, in, n_in, out, n_out);
  DevPlan plan = L.plan;
  plan.num_rows = num_rows;
  for (size_t j = 0; j < L.in_cols.size(); ++j) plan.in[j] = to_dev_in(in[L.in_cols[j]]);
  const int n_str = count_string_outputs(L, in, out, n_out, outputs);
  // the predicate's launch shape drives the selection kernel, the
  // projection's (or the predicate's, for a plain filter) the output kernel
  int sel_rpt = rows_per_thread(pred), sel_block = 256;
  int rpt = 0, block = 256;
  launch_shape(L, rows_per_thread(proj ? proj : pred), block_size(proj ? proj : pred), &rpt, &block);
  if (plan.n_ranges == 0) launch_shape(L, sel_rpt, block_size(pred), &sel_rpt, &sel_block);
  if (sel_block != block) block = sel_block;  // one block size per plan (the spill layout is per block)
  // the one-pass kernels serve a whole call over fixed-width outputs
  const bool one_pass = phases == PHASE_ALL && n_str == 0;
  const bool generated = one_pass && ((pred->launch.flags & BSQL_LAUNCH_GENERATED) || generated_check_mode()) &&
                         generated_filter_project_eligible(plan, L.ops);
  if (generated_check_mode()) {
    if (!generated) fail("this filter does not run as a generated kernel");
    for (int k = 0; k < n_out; ++k) {
      plan.out[k].valid = out[k].validity;
      plan.out[k].type = out[k].type_id;
    }
    e = launch_generated_filter_project(plan, L.ops, nullptr, nullptr, d_count, capacity, nullptr);
    return e == hipSuccess ? BSQL_OK : device_error(e, "generated kernel check");
  }
  const int64_t sel_words = std::max({select_scratch_words(plan, L.nreg, sel_rpt, block, lds_regs(pred, proj)),
                                      (int64_t)fused_scratch_words(num_rows),
                                      generated_filter_project_scratch_words(num_rows)});
  if (sel_words < 0) fail("no kernel variant for this program");
  const bool fused = !generated && one_pass && (pred->launch.flags & BSQL_LAUNCH_FUSED) &&
                     fused_eligible(plan, L.nreg);
  const int64_t id_words = fused || generated ? 0 : (num_rows + 1) / 2;
  const int64_t st_words = n_str ? string_tiles(num_rows) + 1 : 0;
  const int64_t pw = plan_words(L.ops.size());
  const int64_t at_plan = sel_words + id_words + st_words + (int64_t)n_str * num_rows;
  uint64_t* scratch = nullptr;
  e = get_scratch((hipStream_t)stream, at_plan + pw, &scratch);
  if (e != hipSuccess) return device_error(e, "filter scratch");
  uint64_t* sel = scratch;
  uint32_t* ids = (uint32_t*)(scratch + sel_words);
  uint64_t* status = scratch + sel_words + id_words;
  uint64_t* slices = status + st_words;
  for (int k = 0, j = 0; k < n_out; ++k) {
    const bool str = L.out_str_col[(size_t)k] >= 0;
    plan.out[k].data = str ? (void*)(slices + (size_t)(j++) * (size_t)capacity) : (outputs ? out[k].data : nullptr);
    plan.out[k].valid = outputs ? out[k].validity : nullptr;
    plan.out[k].type = out[k].type_id;
  }
  e = upload_plan(plan, L.ops, scratch, at_plan, (hipStream_t)stream);
  if (e == hipSuccess && generated)
    e = launch_generated_filter_project(plan, L.ops, (const DevPlan*)(scratch + at_plan), sel, d_count, capacity,
                                        (hipStream_t)stream);
  else if (e == hipSuccess)
    e = launch_compact(plan, (const DevPlan*)(scratch + at_plan), L.nreg, sel_rpt, rpt, block, sel, ids, d_count,
                       capacity, (hipStream_t)stream, phases, fused, lds_regs(pred, proj));
  if (e == hipSuccess && outputs)
    e = finish_string_outputs(L, in, out, n_out, slices, d_count, 0, capacity, status, (hipStream_t)stream);
  if (e != hipSuccess) return device_error(e, "filter");
  return BSQL_OK;
}

int run_compact(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in, bsql_column* out,
                int n_out, int64_t num_rows, int64_t capacity, int64_t* d_count, void* stream, int phases) {
  return run_compact(pred, proj, in, n_in, out, n_out, num_rows, capacity, d_count, stream, phases, nullptr);
}

// Several (predicate, projection) queries over the same table: one launch of
// the kernel generated for all of them when every query can run as a
// generated one-pass filter (range-test predicate, fixed-width outputs, no
// step columns), else one bsql_filter_project call per query. Query q writes
// out[q][0 .. n_out[q]) up to capacity[q] rows and its count to d_counts[q].
int run_filter_project_batch(bsql_program* const* preds, bsql_program* const* projs, int nq, const bsql_column* in,
                             int n_in, bsql_column* const* out, const int32_t* n_out, int64_t num_rows,
                             const int64_t* capacity, int64_t* d_counts, void* stream) {
  if (nq < 0 || nq > BSQL_MAX_BATCH_QUERIES) fail("a batch holds at most 16 queries");
  if (nq > 0 && (!preds || !projs || !out || !n_out || !capacity)) fail("null batch arrays");
  if (!d_counts) fail("null count pointer");
  check_inputs(in, n_in, num_rows);
  if (num_rows > (int64_t)UINT32_MAX) fail("a filter handles at most 2^32-1 rows per call");
  std::vector<DevPlan> plans;
  std::vector<std::vector<DevOp>> ops;
  std::vector<std::vector<int>> in_maps;
  std::vector<int64_t> caps;
  bool together = nq > 0;
  for (int q = 0; q < nq; ++q) {
    bsql_program* pred = preds[q];
    bsql_program* proj = projs[q];
    if (!pred || !proj) fail("null program in batch");
    if (capacity[q] < 0 || capacity[q] > num_rows) fail("capacity must be in [0, num_rows]");
    if (n_out[q] != num_roots(*proj) && proj->parts.empty()) fail("number of outputs does not match the program");
    check_outputs(out[q], n_out[q], capacity[q]);
    if (!together) continue;
    if (!pred->steps.empty() || !proj->steps.empty() || !proj->parts.empty() || n_out[q] > BSQL_DEV_MAX_OUT) {
      together = false;
      continue;
    }
    if (!generated_check_mode()) {
      hipError_t e = ensure_strings(*pred);
      if (e == hipSuccess) e = ensure_strings(*proj);
      if (e != hipSuccess) return device_error(e, "filter batch");
    }
    LowerOptions opt;
    opt.mode = LowerOptions::FILTER_PROJECT;
    opt.pred = pred;
    opt.proj = proj;
    opt.conj_order = &pred->conj_order;
    opt.ranges = !(pred->launch.flags & BSQL_LAUNCH_NO_RANGES);
    opt.lazy_loads = (pred->launch.flags & BSQL_LAUNCH_LAZY_LOADS) != 0;
    opt.reg_budget = std::min(reg_budget(pred), reg_budget(proj));
    const Lowered& L = lowered_for(pred, opt, in, n_in, out[q], n_out[q]);
    DevPlan plan = L.plan;
    plan.num_rows = num_rows;
    for (size_t j = 0; j < L.in_cols.size(); ++j) plan.in[j] = to_dev_in(in[L.in_cols[j]]);
    if (count_string_outputs(L, in, out[q], n_out[q]) != 0 || !generated_filter_project_eligible(plan, L.ops)) {
      together = false;
      continue;
    }
    for (int k = 0; k < n_out[q]; ++k) {
      plan.out[k].data = out[q][k].data;
      plan.out[k].valid = out[q][k].validity;
      plan.out[k].type = out[q][k].type_id;
    }
    plans.push_back(plan);
    ops.push_back(L.ops);
    in_maps.emplace_back(L.in_cols.begin(), L.in_cols.end());
    caps.push_back(capacity[q]);
  }
  if (!together) {
    if (generated_check_mode()) fail("this batch does not run as one generated kernel");
    for (int q = 0; q < nq; ++q) {
      const int rc = run_filter_project(preds[q], projs[q], in, n_in, out[q], n_out[q], num_rows, capacity[q],
                                        d_counts + q, stream);
      if (rc != BSQL_OK) return rc;
    }
    return BSQL_OK;
  }
  if (generated_check_mode()) {
    const std::vector<const DevPlan*> none((size_t)nq, nullptr);
    const hipError_t e =
        launch_generated_filter_project_batch(plans, ops, in_maps, none, nullptr, d_counts, caps, nullptr);
    return e == hipSuccess ? BSQL_OK : device_error(e, "generated batch kernel check");
  }
  // scratch: every query's tile status words, then the plan blobs
  const int64_t tiles = generated_filter_project_scratch_words(num_rows);
  int64_t at = tiles * nq;
  std::vector<int64_t> plan_at;
  for (int q = 0; q < nq; ++q) {
    plan_at.push_back(at);
    at += plan_words(ops[(size_t)q].size());
  }
  uint64_t* scratch = nullptr;
  hipError_t e = get_scratch((hipStream_t)stream, at, &scratch);
  if (e != hipSuccess) return device_error(e, "filter batch scratch");
  // one copy of all the plan blobs (their op pointers aimed at the scratch)
  std::vector<char> blob((size_t)(at - plan_at[0]) * sizeof(uint64_t), 0);
  std::vector<const DevPlan*> d_plans;
  for (int q = 0; q < nq; ++q) {
    uint64_t* dst = scratch + plan_at[(size_t)q];
    const std::vector<char> b = plan_blob(plans[(size_t)q], ops[(size_t)q], dst);
    std::memcpy(blob.data() + (size_t)(plan_at[(size_t)q] - plan_at[0]) * sizeof(uint64_t), b.data(), b.size());
    d_plans.push_back((const DevPlan*)dst);
  }
  {
    // the stream's scratch no longer holds the single plan it remembers
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find({dev, (hipStream_t)stream});
    if (it != g_scratch.end()) {
      it->second.plan.clear();
      it->second.plan_at = -1;
    }
  }
  e = hipMemcpyAsync(scratch + plan_at[0], blob.data(), blob.size(), hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e == hipSuccess)
    e = launch_generated_filter_project_batch(plans, ops, in_maps, d_plans, scratch, d_counts, caps,
                                              (hipStream_t)stream);
  if (e != hipSuccess) return device_error(e, "filter batch");
  return BSQL_OK;
}

size_t noalloc_scratch_bytes();
hipError_t noalloc_scratch(uint64_t** out);
hipError_t write_words_async(uint64_t* dst, const uint64_t* src, size_t n, hipStream_t stream);

namespace {

bool has_string_literals(const Program& p) {
  for (size_t i = 0; i < p.ops.size(); ++i)
    if ((p.left[i] == BSQL_SCALAR_INDEX && p.lsc[i].type_id == BSQL_STRING) ||
        (p.right[i] == BSQL_SCALAR_INDEX && p.rsc[i].type_id == BSQL_STRING))
      return true;
  return false;
}

}  // namespace

// bsql_filter_project with every buffer beyond the inputs and outputs in
// the fixed device scratch: no allocation, no host-memory copy (the plan is
// written by kernels that carry it as arguments), nothing to clear (the
// selection writes every ballot and count it later reads).
int run_filter_project_noalloc(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in,
                               bsql_column* out, int n_out, int64_t num_rows, int64_t* d_count, void* stream) {
  if (!pred || !proj) fail("null program");
  if (!d_count) fail("null count pointer");
  for (bsql_program* p : {pred, proj}) {
    if (!p->steps.empty()) fail("bsql_filter_project_noalloc: programs that build string columns allocate them");
    if (!p->parts.empty()) fail("bsql_filter_project_noalloc: a projection split into parts runs in several calls");
    if (has_string_literals(*p))
      fail("bsql_filter_project_noalloc: string literals need a device copy the library allocates");
  }
  check_inputs(in, n_in, num_rows);
  if (n_out != num_roots(*proj)) fail("number of outputs does not match the program");
  if (n_out > BSQL_DEV_MAX_OUT) fail("bsql_filter_project_noalloc writes at most 64 outputs");
  check_outputs(out, n_out, num_rows);
  for (int k = 0; k < n_out; ++k)
    if (out[k].type_id == BSQL_STRING) fail("bsql_filter_project_noalloc writes fixed-width outputs only");
  if (num_rows > (int64_t)UINT32_MAX) fail("a filter handles at most 2^32-1 rows per call");
  LowerOptions opt;
  opt.mode = LowerOptions::FILTER_PROJECT;
  opt.pred = pred;
  opt.proj = proj;
  opt.conj_order = &pred->conj_order;
  opt.ranges = !(pred->launch.flags & BSQL_LAUNCH_NO_RANGES);
  opt.lazy_loads = (pred->launch.flags & BSQL_LAUNCH_LAZY_LOADS) != 0;
  opt.reg_budget = std::min(reg_budget(pred), reg_budget(proj));
  const Lowered& L = lowered_for(pred, opt, in, n_in, out, n_out);
  DevPlan plan = L.plan;
  plan.num_rows = num_rows;
  for (size_t j = 0; j < L.in_cols.size(); ++j) plan.in[j] = to_dev_in(in[L.in_cols[j]]);
  int sel_rpt = rows_per_thread(pred), sel_block = 256;
  int rpt = 0, block = 256;
  launch_shape(L, rows_per_thread(proj), block_size(proj), &rpt, &block);
  if (plan.n_ranges == 0) launch_shape(L, sel_rpt, block_size(pred), &sel_rpt, &sel_block);
  if (sel_block != block) block = sel_block;
  const int64_t sel_words = std::max(select_scratch_words(plan, L.nreg, sel_rpt, block, lds_regs(pred, proj)),
                                     (int64_t)fused_scratch_words(num_rows));
  if (sel_words < 0) fail("no kernel variant for this program");
  const bool fused = (pred->launch.flags & BSQL_LAUNCH_FUSED) && fused_eligible(plan, L.nreg);
  const int64_t id_words = fused ? 0 : (num_rows + 1) / 2;
  const int64_t at_plan = sel_words + id_words;
  const int64_t words = at_plan + plan_words(L.ops.size());
  const int64_t cap = (int64_t)(noalloc_scratch_bytes() / 8);
  if (words > cap)
    fail("bsql_filter_project_noalloc: " + std::to_string(num_rows) + " rows need " + std::to_string(words * 8) +
         " bytes of scratch, more than its " + std::to_string(cap * 8) + " (split the table into batches)");
  uint64_t* scratch = nullptr;
  hipError_t e = noalloc_scratch(&scratch);
  if (e != hipSuccess) return device_error(e, "bsql_filter_project_noalloc scratch");
  for (int k = 0; k < n_out; ++k) {
    plan.out[k].data = out[k].data;
    plan.out[k].valid = out[k].validity;
    plan.out[k].type = out[k].type_id;
  }
  std::vector<char> blob = plan_blob(plan, L.ops, scratch + at_plan);
  blob.resize((blob.size() + 7) / 8 * 8, 0);
  e = write_words_async(scratch + at_plan, (const uint64_t*)blob.data(), blob.size() / 8, (hipStream_t)stream);
  if (e == hipSuccess)
    e = launch_compact(plan, (const DevPlan*)(scratch + at_plan), L.nreg, sel_rpt, rpt, block, scratch,
                       (uint32_t*)(scratch + sel_words), d_count, num_rows, (hipStream_t)stream, PHASE_ALL, fused,
                       lds_regs(pred, proj));
  if (e != hipSuccess) return device_error(e, "bsql_filter_project_noalloc");
  return BSQL_OK;
}

namespace {

// Output groups of a program without parts: consecutive runs of at most
// BSQL_DEV_MAX_OUT outputs.
std::vector<std::vector<int>> output_groups(int n_out) {
  std::vector<std::vector<int>> g;
  for (int k = 0; k < n_out; k += BSQL_DEV_MAX_OUT) {
    g.emplace_back();
    for (int j = k; j < std::min(n_out, k + BSQL_DEV_MAX_OUT); ++j) g.back().push_back(j);
  }
  return g;
}

}  // namespace

int run_filter_project(bsql_program* pred, bsql_program* proj, const bsql_column* in, int n_in, bsql_column* out,
                       int n_out, int64_t num_rows, int64_t capacity, int64_t* d_count, void* stream) {
  if (!proj) fail("null projection");
  if (n_out != num_roots(*proj) && proj->parts.empty()) fail("number of outputs does not match the program");
  if (proj->parts.empty() && n_out <= BSQL_DEV_MAX_OUT)
    return run_compact(pred, proj, in, n_in, out, n_out, num_rows, capacity, d_count, stream, PHASE_ALL, nullptr);
  // one selection, then one output launch per part and output group (the
  // row ids stay in the stream's scratch in between)
  std::vector<std::pair<bsql_program*, std::vector<int>>> launches;  // program, outputs of `out`
  if (!proj->parts.empty()) {
    for (size_t j = 0; j < proj->parts.size(); ++j) {
      bsql_program* part = static_cast<bsql_program*>(proj->parts[j].get());
      for (const std::vector<int>& g : output_groups(num_roots(*part))) {
        std::vector<int> outs;
        for (int k : g) outs.push_back(proj->part_outputs[j][(size_t)k]);
        launches.push_back({part, outs});
      }
    }
  } else {
    for (const std::vector<int>& g : output_groups(n_out)) launches.push_back({proj, g});
  }
  bool first = true;
  for (auto& l : launches) {
    std::vector<bsql_column> o;
    std::vector<int> local;  // outputs of l.first this launch writes
    for (size_t i = 0; i < l.second.size(); ++i) {
      if (l.second[i] < 0 || l.second[i] >= n_out) fail("output index out of range");
      o.push_back(out[l.second[i]]);
    }
    // the part's outputs, in the part's own numbering
    if (l.first == proj) {
      local = l.second;
    } else {
      const std::vector<int>& po = proj->part_outputs[(size_t)(std::find_if(proj->parts.begin(), proj->parts.end(),
          [&](const std::shared_ptr<Program>& q) { return q.get() == l.first; }) - proj->parts.begin())];
      for (int k : l.second) local.push_back((int)(std::find(po.begin(), po.end(), k) - po.begin()));
    }
    const int rc = run_compact(pred, l.first, in, n_in, o.data(), (int)o.size(), num_rows, capacity, d_count,
                               stream, first ? PHASE_ALL : PHASE_OUTPUT, &local);
    if (rc != BSQL_OK) return rc;
    first = false;
  }
  return BSQL_OK;
}

int execute_program(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                    int64_t num_rows, void* stream, bool private_scratch) {
  if (!prog) fail("null program");
  if (!prog->parts.empty()) fail("internal: a split program runs part by part");
  if (!prog->steps.empty() && n_in == prog->n_table) fail("the program's step columns were not built");
  check_inputs(in, n_in, num_rows);
  if (n_out != num_roots(*prog)) fail("number of outputs does not match the program");
  check_outputs(out, n_out, num_rows);
  hipError_t e = generated_check_mode() ? hipSuccess : ensure_strings(*prog);
  if (e != hipSuccess) return device_error(e, "bsql_interp_execute");
  for (const std::vector<int>& group : output_groups(n_out)) {
    std::vector<bsql_column> gout;
    for (int k : group) gout.push_back(out[k]);
    const int g_out = (int)gout.size();
    LowerOptions opt;
    opt.mode = LowerOptions::EXECUTE;
    opt.proj = prog;
    opt.reg_budget = reg_budget(prog);
    const bool all = g_out == n_out;
    opt.out_subset = all ? nullptr : &group;
    const Lowered& L = lowered_for(prog, opt, in, n_in, gout.data(), g_out);
    DevPlan plan = L.plan;
    plan.num_rows = num_rows;
    for (size_t j = 0; j < L.in_cols.size(); ++j) plan.in[j] = to_dev_in(in[L.in_cols[j]]);
    const int n_str = count_string_outputs(L, in, gout.data(), g_out);
    int rpt = 0, block = 256;
    launch_shape(L, rows_per_thread(prog), block_size(prog), &rpt, &block);
    // scratch: string assembly status + ticket | one slice column per string output | plan + ops
    const int64_t st_words = n_str ? string_tiles(num_rows) + 1 : 0;
    const int64_t at_plan = st_words + (int64_t)n_str * num_rows;
    const int64_t words = at_plan + plan_words(L.ops.size());
    if (generated_check_mode()) {
      if (!generated_execute_eligible(plan, L.ops)) fail("this program does not run as a generated kernel");
      for (int k = 0; k < g_out; ++k) {
        plan.out[k].valid = gout[k].validity;
        plan.out[k].type = gout[k].type_id;
      }
      e = launch_generated_execute(plan, L.ops, nullptr, L.nreg, variant_rows(L.nreg, rpt, block, L.n_spill), block,
                                   nullptr);
      if (e != hipSuccess) return device_error(e, "generated kernel check");
      continue;
    }
    uint64_t* scratch = nullptr;
    if (private_scratch) {
      e = hipMallocAsync((void**)&scratch, (size_t)words * 8, (hipStream_t)stream);
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
    uint64_t* status = scratch;
    uint64_t* slices = status + st_words;
    for (int k = 0, j = 0; k < g_out; ++k) {
      const bool str = L.out_str_col[(size_t)k] >= 0;
      plan.out[k].data = str ? (void*)(slices + (size_t)(j++) * (size_t)num_rows) : gout[k].data;
      plan.out[k].valid = gout[k].validity;
      plan.out[k].type = gout[k].type_id;
    }
    if (num_rows == 0 && n_str == 0) continue;
    if (private_scratch) {
      // a private buffer is fresh: the plan goes in without the reuse check
      const std::vector<char> blob = plan_blob(plan, L.ops, scratch + at_plan);
      e = hipMemcpyAsync(scratch + at_plan, blob.data(), blob.size(), hipMemcpyHostToDevice, (hipStream_t)stream);
      // the host blob must outlive the copy: pageable copies are staged before the call returns
    } else {
      e = upload_plan(plan, L.ops, scratch, at_plan, (hipStream_t)stream);
    }
    if (e == hipSuccess && num_rows > 0) {
      if ((prog->launch.flags & BSQL_LAUNCH_GENERATED) && generated_execute_eligible(plan, L.ops))
        e = launch_generated_execute(plan, L.ops, (const DevPlan*)(scratch + at_plan), L.nreg,
                                     variant_rows(L.nreg, rpt, block, L.n_spill), block, (hipStream_t)stream);
      else
        e = launch_execute(plan, (const DevPlan*)(scratch + at_plan), L.nreg, rpt, block, (hipStream_t)stream,
                           lds_regs(prog, nullptr));
    }
    if (e == hipSuccess)
      e = finish_string_outputs(L, in, gout.data(), g_out, slices, nullptr, num_rows, num_rows, status,
                                (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_interp_execute");
  }
  return BSQL_OK;
}

int run_execute(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                int64_t num_rows, void* stream) {
  if (!prog) fail("null program");
  if (!prog->parts.empty()) {
    if (n_out != num_roots(*prog)) fail("number of outputs does not match the program");
    for (size_t j = 0; j < prog->parts.size(); ++j) {
      std::vector<bsql_column> o;
      for (int k : prog->part_outputs[j]) {
        if (k < 0 || k >= n_out) fail("output index out of range");
        o.push_back(out[k]);
      }
      const int rc = run_execute(static_cast<bsql_program*>(prog->parts[j].get()), in, n_in, o.data(),
                                 (int)o.size(), num_rows, stream);
      if (rc != BSQL_OK) return rc;
    }
    return BSQL_OK;
  }
  if (prog->steps.empty()) return execute_program(prog, in, n_in, out, n_out, num_rows, stream, false);
  check_inputs(in, n_in, num_rows);
  StepColumns sc;
  sc.stream = (hipStream_t)stream;
  sc.cols.assign(in, in + n_in);
  build_steps(*prog, sc, num_rows);
  return execute_program(prog, sc.cols.data(), (int)sc.cols.size(), out, n_out, num_rows, stream, false);
}

}  // namespace bsql

extern "C" {

const char* bsql_last_error(void) { return g_last_error.c_str(); }
int bsql_version(void) { return 2; }

int bsql_expression_rewrite(const char* expr, int stage, const int32_t* types, int n_types, char* out,
                            size_t cap, size_t* len) {
  BSQL_TRY({
    if (!expr || (n_types > 0 && !types) || (cap > 0 && !out)) fail("invalid arguments");
    const std::string r = rewrite_expression(expr, stage, types, n_types);
    if (cap > 0) {
      const size_t n = std::min(cap - 1, r.size());
      std::memcpy(out, r.data(), n);
      out[n] = 0;
    }
    if (len) *len = r.size();
    return BSQL_OK;
  })
}

int bsql_set_current_timestamp(const char* ts) {
  BSQL_TRY({
    set_current_timestamp(ts ? ts : "");
    return BSQL_OK;
  })
}

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
int bsql_program_num_outputs(const bsql_program* p) { return p ? num_roots(*p) : -1; }
int bsql_program_num_inputs(const bsql_program* p) { return p ? p->n_inputs : -1; }
int bsql_program_num_interpreted(const bsql_program* p) { return p ? (int)p->final_out.size() : -1; }

int bsql_program_root(const bsql_program* p, int k, int32_t* kind, int32_t* index, bsql_scalar* literal) {
  if (!p || k < 0 || k >= num_roots(*p) || !p->parts.empty())
    return set_error(BSQL_ERR_INVALID, "invalid arguments");
  const Root r = root_of(*p, k);
  if (kind) *kind = r.kind == Root::VM ? BSQL_ROOT_INTERPRETED : r.kind == Root::COLUMN ? BSQL_ROOT_COLUMN
                                                                                       : BSQL_ROOT_LITERAL;
  if (index) *index = r.kind == Root::LITERAL ? -1 : r.index;
  if (literal) {
    std::memset(literal, 0, sizeof(*literal));
    if (r.kind == Root::LITERAL) {
      literal->type_id = r.lit.type_id;
      if (is_float_type(r.lit.type_id)) literal->v.f64 = r.lit.f64; else literal->v.i64 = r.lit.i64;
      // a null literal: str_len -1
      if (r.lit.is_null) literal->str_len = -1;
    }
  }
  return BSQL_OK;
}

int bsql_program_num_parts(const bsql_program* p) { return p ? (int)p->parts.size() : -1; }

int bsql_program_part(const bsql_program* p, int j, const bsql_program** part, int32_t* outputs, int cap) {
  if (!p || j < 0 || j >= (int)p->parts.size() || !part) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  *part = static_cast<const bsql_program*>(p->parts[(size_t)j].get());
  const std::vector<int>& o = p->part_outputs[(size_t)j];
  for (size_t i = 0; i < o.size() && (int)i < cap && outputs; ++i) outputs[i] = o[i];
  return (int)o.size();
}

int bsql_program_arrays(const bsql_program* p, int16_t* left, int16_t* right, int16_t* out,
                        int32_t* ops, int16_t* final_out) {
  if (!p) return set_error(BSQL_ERR_INVALID, "null program");
  if (!p->parts.empty())
    return set_error(BSQL_ERR_INVALID, "the expression list was split into " + std::to_string(p->parts.size()) +
                                           " programs (bsql_program_part)");
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
    if (!p || !type_id || k < 0 || k >= num_roots(*p)) fail("invalid arguments");
    if (!p->parts.empty()) {
      for (size_t j = 0; j < p->parts.size(); ++j)
        for (size_t i = 0; i < p->part_outputs[j].size(); ++i)
          if (p->part_outputs[j][i] == k)
            return bsql_program_output_type(static_cast<const bsql_program*>(p->parts[j].get()), in_types, n_in,
                                            (int)i, type_id);
      fail("output " + std::to_string(k) + " is in no part");
    }
    const std::vector<int32_t> ext = extended_types(*p, in_types, n_in);
    const Root r = root_of(*p, k);
    if (r.kind == Root::LITERAL) {
      *type_id = r.lit.type_id;
      return BSQL_OK;
    }
    if (r.kind == Root::COLUMN) {
      if (r.index < 0 || r.index >= (int)ext.size()) fail("output column not in the table");
      *type_id = ext[(size_t)r.index];
      return BSQL_OK;
    }
    std::vector<int32_t> st = slot_types_for(*p, ext.data(), (int)ext.size());
    OpTypes t = propagate_types(*p, st.data(), (int)st.size());
    const int slot = p->final_out[(size_t)r.index];
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

int bsql_program_plan_info(const bsql_program* cp, const int32_t* in_types, int n_in, bsql_plan_info* info) {
  BSQL_TRY({
    bsql_program* p = const_cast<bsql_program*>(cp);
    if (!p || !info || (n_in > 0 && !in_types)) fail("invalid arguments");
    if (!p->parts.empty()) fail("the expression list was split into parts (bsql_program_part)");
    const std::vector<int32_t> ext = extended_types(*p, in_types, n_in);
    const int n_out = num_roots(*p);
    if (n_out > BSQL_DEV_MAX_OUT) fail("more outputs than one launch writes");
    std::vector<int32_t> out_types((size_t)n_out);
    for (int k = 0; k < n_out; ++k) {
      const int rc = bsql_program_output_type(p, in_types, n_in, k, &out_types[(size_t)k]);
      if (rc != BSQL_OK) return rc;
    }
    LowerOptions opt;
    opt.mode = LowerOptions::EXECUTE;
    opt.proj = p;
    opt.reg_budget = reg_budget(p);
    const Lowered L = lower_best(opt, ext.data(), (int)ext.size(), out_types.data(), n_out);
    std::memset(info, 0, sizeof(*info));
    info->n_ops = L.plan.n_ops;
    info->n_regs = L.nreg;
    info->n_spill = L.n_spill;
    info->n_inputs = (int32_t)L.in_cols.size();
    int in_flight = 0;
    for (const DevOp& d : L.ops) {
      if (d.code == D_LOADN) {
        info->n_load_issues += 1;
        in_flight = 0;
      }
      if (d.code == D_COMMIT) info->max_loads_in_flight = std::max<int32_t>(info->max_loads_in_flight, ++in_flight);
      if (d.code == D_SPILL) info->n_spill_ops += 1;
      if (d.code == D_FILL) info->n_fill_ops += 1;
    }
    return BSQL_OK;
  })
}

int bsql_program_set_launch(bsql_program* p, const bsql_launch_config* cfg) {
  if (!p || !cfg) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  const int r = cfg->rows_per_thread;
  if (r != 0 && r != 1 && r != 2 && r != 4 && r != 8 && r != 16 && r != 32)
    return set_error(BSQL_ERR_INVALID, "rows_per_thread must be 0, 1, 2, 4, 8, 16 or 32");
  if (cfg->block_size != 0 && cfg->block_size != 64 && cfg->block_size != 256)
    return set_error(BSQL_ERR_INVALID, "block_size must be 0, 64 or 256");
  if (cfg->flags & ~(BSQL_LAUNCH_NO_RANGES | BSQL_LAUNCH_LAZY_LOADS | BSQL_LAUNCH_SPILL_WIDE | BSQL_LAUNCH_FUSED |
                     BSQL_LAUNCH_GENERATED | BSQL_LAUNCH_LDS_REGS))
    return set_error(BSQL_ERR_INVALID, "unknown launch flags");
  std::lock_guard<std::mutex> lk(p->cache_mu);
  p->launch = *cfg;
  return BSQL_OK;
}

int bsql_generated_kernel_stats(int64_t* kernels, int64_t* launches, double* compile_seconds) {
  generated_stats(kernels, launches, compile_seconds);
  return BSQL_OK;
}

int bsql_generated_kernel_check(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                                int n_in, const bsql_column* out, int n_out, int64_t num_rows, const char* arch,
                                int64_t* code_bytes) {
  if (!proj || !arch || !*arch || num_rows <= 0) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  BSQL_TRY({
    struct Reset {
      ~Reset() { set_generated_check(nullptr); }
    } reset;
    set_generated_check(arch);
    int64_t count = 0;
    const int rc = pred ? run_filter_project(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in,
                                             n_in, const_cast<bsql_column*>(out), n_out, num_rows, num_rows,
                                             &count, nullptr)
                        : run_execute(const_cast<bsql_program*>(proj), in, n_in, const_cast<bsql_column*>(out),
                                      n_out, num_rows, nullptr);
    if (rc != BSQL_OK) return rc;
    if (code_bytes) *code_bytes = generated_check_bytes();
    return BSQL_OK;
  })
}

int bsql_interp_execute(const bsql_program* cprog, const bsql_column* in, int n_in,
                        bsql_column* out, int n_out, int64_t num_rows, void* stream) {
  BSQL_TRY({
    bsql_program* prog = const_cast<bsql_program*>(cprog);
    if (!prog) fail("null program");
    reject_capture(stream, "bsql_interp_execute");
    return run_execute(prog, in, n_in, out, n_out, num_rows, stream);
  })
}

int bsql_filter(const bsql_program* pred, const bsql_column* in, int n_in, bsql_column* out,
                int64_t num_rows, int64_t* d_count, void* stream) {
  BSQL_TRY({
    reject_capture(stream, "bsql_filter");
    return run_compact(const_cast<bsql_program*>(pred), nullptr, in, n_in, out, n_in, num_rows,
                       num_rows, d_count, stream, PHASE_ALL);
  })
}

int bsql_filter_count(const bsql_program* pred, const bsql_column* in, int n_in, int64_t num_rows,
                      int64_t* d_count, void* stream) {
  BSQL_TRY({
    reject_capture(stream, "bsql_filter_count");
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
    reject_capture(stream, "bsql_filter_project_sized");
    return run_filter_project(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in, n_in, out,
                              n_out, num_rows, capacity, d_count, stream);
  })
}

int bsql_filter_project_noalloc(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                                int n_in, bsql_column* out, int n_out, int64_t num_rows, int64_t* d_count,
                                void* stream) {
  BSQL_TRY({
    return run_filter_project_noalloc(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in, n_in,
                                      out, n_out, num_rows, d_count, stream);
  })
}

int bsql_filter_project_batch(const bsql_program* const* preds, const bsql_program* const* projs, int n_queries,
                              const bsql_column* in, int n_in, bsql_column* const* out, const int32_t* n_out,
                              int64_t num_rows, const int64_t* capacity, int64_t* d_counts, void* stream) {
  BSQL_TRY({
    reject_capture(stream, "bsql_filter_project_batch");
    return run_filter_project_batch(const_cast<bsql_program* const*>(preds), const_cast<bsql_program* const*>(projs),
                                    n_queries, in, n_in, out, n_out, num_rows, capacity, d_counts, stream);
  })
}

int bsql_generated_batch_kernel_check(const bsql_program* const* preds, const bsql_program* const* projs,
                                      int n_queries, const bsql_column* in, int n_in, bsql_column* const* out,
                                      const int32_t* n_out, int64_t num_rows, const char* arch,
                                      int64_t* code_bytes) {
  if (!arch || !*arch || num_rows <= 0 || n_queries <= 0) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  BSQL_TRY({
    struct Reset {
      ~Reset() { set_generated_check(nullptr); }
    } reset;
    set_generated_check(arch);
    std::vector<int64_t> caps((size_t)n_queries, num_rows), counts((size_t)n_queries, 0);
    const int rc = run_filter_project_batch(const_cast<bsql_program* const*>(preds),
                                            const_cast<bsql_program* const*>(projs), n_queries, in, n_in, out, n_out,
                                            num_rows, caps.data(), counts.data(), nullptr);
    if (rc != BSQL_OK) return rc;
    if (code_bytes) *code_bytes = generated_check_bytes();
    return BSQL_OK;
  })
}

int bsql_filter_project(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                        int n_in, bsql_column* out, int n_out, int64_t num_rows, int64_t* d_count,
                        void* stream) {
  BSQL_TRY({
    if (!proj) fail("null projection");
    reject_capture(stream, "bsql_filter_project");
    return run_filter_project(const_cast<bsql_program*>(pred), const_cast<bsql_program*>(proj), in, n_in,
                              out, n_out, num_rows, num_rows, d_count, stream);
  })
}

}  // extern "C"

// ------------------------------------------------------------------ batches

// The batch's programs lowered one way: every plan (columns filled in, ops
// aimed at d_ops) and the register file and spill slots they need.
struct BatchForm {
  std::vector<DevPlan> plans;
  int nreg = 1;
  int n_spill = 0;
  DevPlan* d_plans = nullptr;
  DevOp* d_ops = nullptr;
};

struct bsql_batch {
  std::vector<int64_t> rows;
  // forms[0]: each program with the register file it needs (16 registers
  // for the whole batch if one program needs more than 8); forms[1]: at most
  // 7 live values in registers and the rest in LDS spill slots, so that the
  // 8-register file runs every program
  BatchForm forms[2];
  int form = 0;
  bsql_launch_config launch{0, 0, 0};
  int block = 256;
  int rpt = 0;
  bool lf = false;  // register file in LDS
  // one work map per launch shape (block, rows per lane), written once when
  // the shape is first used and never rewritten, so changing the shape never
  // touches a map that a queued launch may still read
  std::map<std::pair<int, int>, std::pair<int64_t*, int64_t>> maps;
  int64_t* d_tile_start = nullptr;
  int64_t n_items = 0;
  const BatchForm& f() const { return forms[form]; }
  ~bsql_batch() {
    if (process_exiting()) return;
    for (BatchForm& x : forms) {
      if (x.d_plans) (void)hipFree(x.d_plans);
      if (x.d_ops) (void)hipFree(x.d_ops);
    }
    for (auto& m : maps) (void)hipFree(m.second.first);
  }
};

namespace {

// Selects the work map of a launch shape (prefix counts of tiles per batch),
// building it on first use.
hipError_t upload_work_map(bsql_batch* b, int rpt, int block, hipStream_t) {
  const int tile = batch_tile_rows(b->f().nreg, rpt, block, b->lf);
  if (tile <= 0) fail("no kernel variant for these programs");
  auto it = b->maps.find({block, rpt});
  if (it == b->maps.end()) {
    std::vector<int64_t> start(b->rows.size() + 1, 0);
    for (size_t i = 0; i < b->rows.size(); ++i) start[i + 1] = start[i] + (b->rows[i] + tile - 1) / tile;
    int64_t* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, start.size() * sizeof(int64_t));
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(d); return e; }
    it = b->maps.emplace(std::make_pair(block, rpt), std::make_pair(d, start.back())).first;
  }
  b->d_tile_start = it->second.first;
  b->n_items = it->second.second;
  b->rpt = rpt;
  b->block = block;
  return hipSuccess;
}

// The launch shape closest to the requested one that holds the batch's
// registers and spill slots.
void batch_shape(const bsql_batch* b, int rpt_req, int block_req, int* rpt, int* block) {
  const int blocks[2] = {block_req ? block_req : 256, block_req == 64 ? 256 : 64};
  const int rpts[4] = {rpt_req, 0, 1, 2};
  for (int bl : blocks)
    for (int r : rpts)
      if (variant_fits(b->f().nreg, r, bl, b->f().n_spill, b->lf)) {
        *rpt = variant_rows(b->f().nreg, r, bl, b->f().n_spill, b->lf);  // concrete: the work map's tile size
        *block = bl;
        return;
      }
  fail("the batch's programs need more spill slots than a workgroup's LDS holds");
}

}  // namespace

extern "C" {

int bsql_batch_prepare(const bsql_batch_desc* descs, int n, bsql_batch** out) {
  BSQL_TRY({
    if (!descs || n <= 0 || !out) fail("invalid arguments");
    auto b = std::make_unique<bsql_batch>();
    std::vector<DevOp> all_ops[2];
    std::vector<size_t> op_at[2];
    for (int i = 0; i < n; ++i) {
      const bsql_batch_desc& d = descs[i];
      bsql_program* prog = const_cast<bsql_program*>(d.prog);
      if (!prog) fail("batch " + std::to_string(i) + " has no program");
      if (!prog->steps.empty())
        fail("batch " + std::to_string(i) + ": programs that build string columns do not run batched");
      if (!prog->parts.empty() || num_roots(*prog) > BSQL_DEV_MAX_OUT)
        fail("batch " + std::to_string(i) + ": a program split over several launches does not run batched");
      check_inputs(d.in, d.n_in, d.num_rows);
      if (d.n_out != num_roots(*prog)) fail("batch " + std::to_string(i) + ": outputs do not match");
      check_outputs(d.out, d.n_out, d.num_rows);
      for (int k = 0; k < d.n_out; ++k)
        if (d.out[k].type_id == BSQL_STRING) fail("batched launches write fixed-width outputs only");
      hipError_t e = ensure_strings(*prog);
      if (e != hipSuccess) return device_error(e, "bsql_batch_prepare");
      b->rows.push_back(d.num_rows);
      for (int form = 0; form < 2; ++form) {
        LowerOptions opt;
        opt.mode = LowerOptions::EXECUTE;
        opt.proj = prog;
        opt.reg_budget = form == 0 ? reg_budget(prog) : 7;
        const Lowered& L = lowered_for(prog, opt, d.in, d.n_in, d.out, d.n_out);
        DevPlan plan = L.plan;
        plan.num_rows = d.num_rows;
        for (size_t j = 0; j < L.in_cols.size(); ++j) plan.in[j] = to_dev_in(d.in[L.in_cols[j]]);
        for (int k = 0; k < d.n_out; ++k) {
          plan.out[k].data = d.out[k].data;
          plan.out[k].valid = d.out[k].validity;
          plan.out[k].type = d.out[k].type_id;
        }
        BatchForm& F = b->forms[form];
        op_at[form].push_back(all_ops[form].size());
        all_ops[form].insert(all_ops[form].end(), L.ops.begin(), L.ops.end());
        F.plans.push_back(plan);
        F.nreg = std::max(F.nreg, L.nreg);
        F.n_spill = std::max(F.n_spill, L.n_spill);
      }
    }
    // one register file size per form: a program that spills even with 15
    // registers fixes the first form at 16
    if (b->forms[0].n_spill > 0) b->forms[0].nreg = BSQL_DEV_MAX_REG;
    hipError_t e = hipSuccess;
    for (int form = 0; form < 2 && e == hipSuccess; ++form) {
      BatchForm& F = b->forms[form];
      e = hipMalloc((void**)&F.d_ops, all_ops[form].size() * sizeof(DevOp));
      for (size_t i = 0; i < F.plans.size() && e == hipSuccess; ++i) {
        F.plans[i].ops = F.d_ops + op_at[form][i];
        F.plans[i].n_spill = F.n_spill;
      }
      if (e == hipSuccess)
        e = hipMemcpy(F.d_ops, all_ops[form].data(), all_ops[form].size() * sizeof(DevOp), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMalloc((void**)&F.d_plans, F.plans.size() * sizeof(DevPlan));
      if (e == hipSuccess)
        e = hipMemcpy(F.d_plans, F.plans.data(), F.plans.size() * sizeof(DevPlan), hipMemcpyHostToDevice);
    }
    // default: the register file in LDS and every program on the 8-register
    // file (the fastest on every batch measured); the VGPR file and each
    // program's own form where that does not fit a workgroup's LDS
    int rpt = 0, block = 256;
    b->lf = true;
    b->form = 1;
    try {
      batch_shape(b.get(), 0, 256, &rpt, &block);
    } catch (const Error&) {
      b->lf = false;
      b->form = 0;
      batch_shape(b.get(), 0, 256, &rpt, &block);
    }
    b->launch = bsql_launch_config{block, rpt, (b->lf ? BSQL_LAUNCH_LDS_REGS : 0) | (b->form ? BSQL_LAUNCH_SPILL_WIDE : 0)};
    if (e == hipSuccess) e = upload_work_map(b.get(), rpt, block, nullptr);
    if (e != hipSuccess) return device_error(e, "bsql_batch_prepare");
    *out = b.release();
    return BSQL_OK;
  })
}

int bsql_batch_launch(const bsql_batch* b, void* stream) {
  BSQL_TRY({
    if (!b) fail("null batch");
    const BatchForm& F = b->f();
    hipError_t e = launch_batches(F.d_plans, b->d_tile_start, (int)F.plans.size(), b->n_items, F.nreg, b->rpt,
                                  b->block, F.n_spill, b->lf, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_batch_launch");
    return BSQL_OK;
  })
}

int bsql_batch_set_launch(bsql_batch* b, const bsql_launch_config* cfg) {
  BSQL_TRY({
    if (!b || !cfg) fail("invalid arguments");
    const int r = cfg->rows_per_thread;
    if (r != 0 && r != 1 && r != 2 && r != 4) fail("rows_per_thread must be 0, 1, 2 or 4");
    if (cfg->block_size != 0 && cfg->block_size != 64 && cfg->block_size != 256)
      fail("block_size must be 0, 64 or 256");
    if (cfg->flags & ~(BSQL_LAUNCH_LDS_REGS | BSQL_LAUNCH_SPILL_WIDE))
      fail("batched launches take only BSQL_LAUNCH_LDS_REGS and BSQL_LAUNCH_SPILL_WIDE");
    const bool lf_was = b->lf;
    const int form_was = b->form;
    b->lf = (cfg->flags & BSQL_LAUNCH_LDS_REGS) != 0;
    b->form = (cfg->flags & BSQL_LAUNCH_SPILL_WIDE) ? 1 : 0;
    int rpt = 0, block = 256;
    try {
      batch_shape(b, r, cfg->block_size, &rpt, &block);
    } catch (...) {
      b->lf = lf_was;
      b->form = form_was;
      throw;
    }
    b->launch = *cfg;
    hipError_t e = upload_work_map(b, rpt, block, nullptr);
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

// launch shapes the autotuners try: (block size, rows per lane)
const int kShapes[][2] = {{256, 1}, {256, 2}, {256, 4}, {64, 1}, {64, 2}, {64, 4}};

}  // namespace

extern "C" {

int bsql_batch_autotune(bsql_batch* b, void* stream, bsql_launch_config* chosen) {
  BSQL_TRY({
    if (!b) fail("null batch");
    float best = 1e30f;
    int best_r = b->rpt, best_b = b->block, best_form = b->form;
    bool best_lf = b->lf;
    // the second form only where it differs (some program needs more than 8 registers)
    const int n_forms = b->forms[0].nreg > 8 ? 2 : 1;
    for (int form = 0; form < n_forms; ++form)
      for (const bool lf : {false, true})
        for (const auto& sh : kShapes) {
          const int bl = sh[0], r = sh[1];
          const BatchForm& F = b->forms[form];
          if (!variant_fits(F.nreg, r, bl, F.n_spill, lf)) continue;  // no variant of this shape
          b->lf = lf;
          b->form = form;
          hipError_t e = upload_work_map(b, r, bl, (hipStream_t)stream);
          float ms = 0.f;
          if (e == hipSuccess)
            e = time_launches((hipStream_t)stream, [&] {
              return launch_batches(F.d_plans, b->d_tile_start, (int)F.plans.size(), b->n_items, F.nreg, r, bl,
                                    F.n_spill, lf, (hipStream_t)stream);
            }, &ms);
          if (e != hipSuccess) return device_error(e, "bsql_batch_autotune");
          if (ms < best) { best = ms; best_r = r; best_b = bl; best_lf = lf; best_form = form; }
        }
    b->lf = best_lf;
    b->form = best_form;
    hipError_t e = upload_work_map(b, best_r, best_b, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_batch_autotune");
    b->launch = bsql_launch_config{best_b, best_r, (best_lf ? BSQL_LAUNCH_LDS_REGS : 0) |
                                                       (best_form ? BSQL_LAUNCH_SPILL_WIDE : 0)};
    if (chosen) *chosen = b->launch;
    return BSQL_OK;
  })
}

// Tries each launch shape, the register file in LDS or in VGPRs and, for a
// program that needs more registers than the small file holds, both spill
// layouts (16 registers, or 8 registers with more values in LDS slots and
// more rows per lane).
int bsql_program_autotune(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                          int64_t num_rows, void* stream, bsql_launch_config* chosen) {
  BSQL_TRY({
    if (!prog) fail("null program");
    float best = 1e30f;
    const bsql_launch_config base = prog->launch;
    bsql_launch_config best_cfg = prog->launch;
    const int lds_opts[2] = {(int)BSQL_LAUNCH_LDS_REGS, 0};
    for (int wide = 0; wide < 2; ++wide)
    for (const int lds : lds_opts) {
      for (const auto& sh : kShapes) {
        bsql_launch_config cfg = base;
        cfg.block_size = sh[0];
        cfg.rows_per_thread = sh[1];
        cfg.flags = (base.flags & ~(BSQL_LAUNCH_SPILL_WIDE | BSQL_LAUNCH_LDS_REGS)) |
                    (wide ? BSQL_LAUNCH_SPILL_WIDE : 0) | lds;
        prog->launch = cfg;
        float ms = 0.f;
        int rc = BSQL_OK;
        hipError_t e = time_launches((hipStream_t)stream, [&] {
          rc = bsql_interp_execute(prog, in, n_in, out, n_out, num_rows, stream);
          return rc == BSQL_OK ? hipSuccess : hipErrorUnknown;
        }, &ms);
        if (rc != BSQL_OK) {
          prog->launch = base;
          return rc;
        }
        if (e != hipSuccess) {
          prog->launch = base;
          return device_error(e, "bsql_program_autotune");
        }
        if (ms < best) { best = ms; best_cfg = cfg; }
      }
    }
    prog->launch = best_cfg;
    if (chosen) *chosen = best_cfg;
    return BSQL_OK;
  })
}

}  // extern "C"
