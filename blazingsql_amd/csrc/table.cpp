// table.cpp -- owning tables and the processor functions of the filter and
// project operators (include/bsql/table.h), over the filter and interpreter
// launches of capi.cpp.
//
//   process_filter        engine/src/execution_kernels/LogicalFilter.cpp:31-52
//   applyBooleanFilter    engine/src/execution_kernels/LogicalFilter.cpp:22-29
//   process_project       engine/src/execution_kernels/LogicalProject.cpp:943-969
//   evaluate_expressions  engine/src/execution_kernels/LogicalProject.cpp:796-941
//   bypassingProject      engine/src/parser/expression_utils.cpp:1115-1150
//   fix_column_aliases    engine/src/parser/expression_utils.cpp:512-535
//   get_projections       engine/src/parser/expression_utils.cpp:549-580
//                         engine/src/execution_kernels/BatchProcessing.cpp:30-46
#include <algorithm>
#include <cstring>
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bsql/table.h"
#include "capi_internal.h"
#include "program.hpp"

namespace {

using Allocator = bsql::AllocRef;

int default_alloc(void*, void** p, size_t bytes, void* stream) {
  const hipError_t e = hipMallocAsync(p, bytes, (hipStream_t)stream);
  if (e == hipSuccess) return 0;
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? 1 : 2;
}

void default_free(void*, void* p, size_t, void* stream) { (void)hipFreeAsync(p, (hipStream_t)stream); }

std::mutex g_alloc_mu;
Allocator g_alloc{default_alloc, default_free, nullptr};

Allocator current_allocator() {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  return g_alloc;
}

}  // namespace

bsql::AllocRef bsql::installed_allocator() { return current_allocator(); }

// Columns with their names; each buffer the table allocated is freed with
// the allocator it came from, on the table's stream.
struct bsql_table {
  struct Col {
    bsql_column view{};
    std::string name;
    void* own[3] = {nullptr, nullptr, nullptr};  // data, validity, offsets
    size_t bytes[3] = {0, 0, 0};
  };
  std::vector<Col> cols;
  int64_t num_rows = 0;
  void* stream = nullptr;
  Allocator alloc = current_allocator();

  bsql_table() = default;
  bsql_table(const bsql_table&) = delete;
  bsql_table& operator=(const bsql_table&) = delete;
  ~bsql_table() {
    for (Col& c : cols)
      for (int i = 0; i < 3; ++i)
        if (c.own[i]) alloc.release(alloc.ctx, c.own[i], c.bytes[i], stream);
  }
};

namespace bsql {
namespace {

using TablePtr = std::unique_ptr<bsql_table>;

void* device_alloc(const Allocator& a, size_t bytes, void* stream) {
  void* p = nullptr;
  const int rc = a.alloc(a.ctx, &p, bytes, stream);
  if (rc == 1) throw Error(BSQL_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  if (rc != 0 || !p) throw Error(BSQL_ERR_DEVICE, "device allocation of " + std::to_string(bytes) + " bytes failed");
  return p;
}

// validity words of `rows` rows, padded to 64 bytes as cudf pads its masks
size_t validity_bytes(int64_t rows) { return (size_t)(((rows + 31) / 32 * 4 + 63) / 64 * 64); }

// Appends an uninitialized column of `rows` rows to t.
bsql_table::Col& add_column(bsql_table& t, int32_t type, const std::string& name, int64_t rows, bool nullable,
                            int64_t chars) {
  if (type != BSQL_STRING && type_size(type) == 0) fail("unsupported column type " + std::to_string(type));
  t.cols.emplace_back();
  bsql_table::Col& c = t.cols.back();
  c.name = name;
  c.view.type_id = type;
  c.view.size = rows;
  c.view.offset = 0;
  const auto take = [&](int i, size_t bytes) -> void* {
    if (bytes == 0) return nullptr;
    c.own[i] = device_alloc(t.alloc, bytes, t.stream);
    c.bytes[i] = bytes;
    return c.own[i];
  };
  if (type == BSQL_STRING) {
    if (chars < 0 || chars > INT32_MAX) fail("string column chars must fit int32 offsets");
    c.view.chars_size = (int32_t)chars;
    c.view.data = take(0, (size_t)chars);
    c.view.offsets = rows > 0 ? (int32_t*)take(2, (size_t)(rows + 1) * 4) : nullptr;
  } else {
    c.view.data = take(0, (size_t)rows * (size_t)type_size(type));
  }
  if (nullable && rows > 0) c.view.validity = (uint32_t*)take(1, validity_bytes(rows));
  return c;
}

TablePtr new_table(void* stream) {
  TablePtr t(new bsql_table());
  t->stream = stream;
  return t;
}

// cudf::empty_like: the same schema, no rows.
TablePtr empty_like(const bsql_table& in, void* stream) {
  TablePtr t = new_table(stream);
  for (const bsql_table::Col& c : in.cols) add_column(*t, c.view.type_id, c.name, 0, false, 0);
  return t;
}

std::vector<int32_t> types_of(const bsql_table& t) {
  std::vector<int32_t> types;
  for (const bsql_table::Col& c : t.cols) types.push_back(c.view.type_id);
  return types;
}

std::vector<bsql_column> views_of(const bsql_table& t) {
  std::vector<bsql_column> v;
  for (const bsql_table::Col& c : t.cols) v.push_back(c.view);
  return v;
}

struct ProgramFree {
  void operator()(bsql_program* p) const { bsql_program_free(p); }
};
using ProgramPtr = std::unique_ptr<bsql_program, ProgramFree>;

ProgramPtr compile(const std::vector<std::string>& exprs, const std::vector<int32_t>& types) {
  std::vector<const char*> ptrs;
  for (const std::string& e : exprs) ptrs.push_back(e.c_str());
  bsql_program* p = nullptr;
  const int rc = bsql_program_compile(ptrs.data(), (int)ptrs.size(), types.data(), (int)types.size(), &p);
  if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
  return ProgramPtr(p);
}

int32_t output_type(const bsql_program* p, const std::vector<int32_t>& types, int k) {
  int32_t t = BSQL_EMPTY;
  const int rc = bsql_program_output_type(p, types.data(), (int)types.size(), k, &t);
  if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
  return t;
}

// A device int64 from the table allocator, freed on the stream on scope exit.
struct DeviceCount {
  Allocator a;
  void* stream;
  int64_t* p;
  DeviceCount(const Allocator& al, void* s) : a(al), stream(s), p((int64_t*)device_alloc(al, 8, s)) {}
  ~DeviceCount() { a.release(a.ctx, p, 8, stream); }
};

// Filter `in` by a one-output BOOL8 program over its columns: selection,
// count to the host, outputs sized to the count, then the gather.
TablePtr filter_table(const bsql_table& in, bsql_program* pred, const std::vector<bsql_column>& extra,
                      void* stream) {
  std::vector<bsql_column> cols = views_of(in);
  cols.insert(cols.end(), extra.begin(), extra.end());
  const int n = (int)in.cols.size();
  TablePtr res = new_table(stream);
  DeviceCount d_count(res->alloc, stream);
  std::vector<bsql_column> outs(cols.size());
  for (int k = 0; k < n; ++k) outs[(size_t)k].type_id = cols[(size_t)k].type_id;
  // the gather writes the table's own columns (the first n of the program's inputs)
  int rc = run_compact(pred, nullptr, cols.data(), (int)cols.size(), outs.data(), n, in.num_rows, in.num_rows,
                       d_count.p, stream, PHASE_SELECT);
  if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
  int64_t count = 0;
  hipError_t e = hipMemcpyAsync(&count, d_count.p, 8, hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) throw Error(device_error(e, "filter count"), bsql_last_error());
  for (int k = 0; k < n; ++k) {
    const bsql_table::Col& c = in.cols[(size_t)k];
    // a string column's selected chars fit in its whole chars buffer
    add_column(*res, c.view.type_id, c.name, count, c.view.validity != nullptr,
               c.view.type_id == BSQL_STRING ? c.view.chars_size : 0);
    outs[(size_t)k] = res->cols[(size_t)k].view;
  }
  res->num_rows = count;
  if (count > 0) {
    rc = run_compact(pred, nullptr, cols.data(), (int)cols.size(), outs.data(), n, in.num_rows, count,
                     d_count.p, stream, PHASE_OUTPUT);
    if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
  }
  return res;
}

// How output k of a compiled list is produced (through its part if the list
// was split).
Root output_root(const Program& p, int k) {
  for (size_t j = 0; j < p.parts.size(); ++j)
    for (size_t i = 0; i < p.part_outputs[j].size(); ++i)
      if (p.part_outputs[j][i] == k) return output_root(*p.parts[j], (int)i);
  return root_of(p, k);
}

// Whether output k of the projection over `in` carries a validity mask:
// interpreted outputs always do (make_fixed_width_column(..., UNINITIALIZED));
// a copied table column keeps its own nullability; a literal column has a
// mask only for a null literal (make_column_from_scalar)
// (LogicalProject.cpp:822-839).
bool output_nullable(const bsql_program& prog, const bsql_table& in, int k) {
  const Root r = output_root(prog, k);
  if (r.kind == Root::COLUMN && r.index < (int)in.cols.size())
    return in.cols[(size_t)r.index].view.validity != nullptr;
  if (r.kind == Root::LITERAL) return r.lit.is_null;
  return true;
}

TablePtr evaluate(const bsql_table& in, const std::vector<std::string>& exprs,
                  const std::vector<std::string>& names, void* stream) {
  const std::vector<int32_t> types = types_of(in);
  ProgramPtr prog = compile(exprs, types);
  std::vector<bsql_column> views = views_of(in);
  // a string output is a slice of one column: the table's or a step's
  const int64_t chars_bound = string_chars_bound(*prog, views.data(), (int)views.size(), in.num_rows);
  TablePtr res = new_table(stream);
  res->num_rows = in.num_rows;
  std::vector<bsql_column> outs;
  for (size_t k = 0; k < exprs.size(); ++k) {
    const int32_t t = output_type(prog.get(), types, (int)k);
    add_column(*res, t, k < names.size() ? names[k] : std::string(), in.num_rows,
               output_nullable(*prog, in, (int)k), t == BSQL_STRING ? chars_bound : 0);
    outs.push_back(res->cols.back().view);
  }
  if (in.num_rows > 0 && !outs.empty()) {
    const int rc = bsql_interp_execute(prog.get(), views.data(), (int)views.size(), outs.data(), (int)outs.size(),
                                       in.num_rows, stream);
    if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
  }
  return res;
}

struct BatchFree {
  void operator()(bsql_batch* b) const { bsql_batch_free(b); }
};

// The projection over several batches of one schema as ONE persistent launch
// (bsql_batch_*): the program is compiled once, each batch's outputs are
// allocated as evaluate allocates them, and the batches' tiles are the work
// items of one kernel, so many small batches cost one launch instead of one
// each. Programs the batched launch does not run (string steps or outputs,
// programs split over several launches) and batches of different schemas
// are evaluated batch by batch.
std::vector<TablePtr> evaluate_batches(const std::vector<const bsql_table*>& ins,
                                       const std::vector<std::string>& exprs,
                                       const std::vector<std::string>& names, void* stream) {
  const std::vector<int32_t> types = types_of(*ins[0]);
  bool batchable = true;
  for (const bsql_table* t : ins) batchable = batchable && types_of(*t) == types;
  std::vector<TablePtr> res;
  ProgramPtr prog = compile(exprs, types);
  batchable = batchable && prog->steps.empty() && prog->parts.empty() && num_roots(*prog) <= BSQL_DEV_MAX_OUT;
  std::vector<int32_t> ot;
  for (size_t k = 0; k < exprs.size(); ++k) {
    ot.push_back(output_type(prog.get(), types, (int)k));
    batchable = batchable && ot.back() != BSQL_STRING;
  }
  if (!batchable) {
    for (const bsql_table* t : ins) res.push_back(evaluate(*t, exprs, names, stream));
    return res;
  }
  std::vector<std::vector<bsql_column>> views(ins.size()), outs(ins.size());
  std::vector<bsql_batch_desc> descs;
  for (size_t i = 0; i < ins.size(); ++i) {
    const bsql_table& in = *ins[i];
    TablePtr r = new_table(stream);
    r->num_rows = in.num_rows;
    for (size_t k = 0; k < exprs.size(); ++k)
      add_column(*r, ot[k], k < names.size() ? names[k] : std::string(), in.num_rows,
                 output_nullable(*prog, in, (int)k), 0);
    views[i] = views_of(in);
    outs[i] = views_of(*r);
    if (in.num_rows > 0 && !exprs.empty())
      descs.push_back(bsql_batch_desc{prog.get(), views[i].data(), (int)views[i].size(), outs[i].data(),
                                      (int)outs[i].size(), in.num_rows});
    res.push_back(std::move(r));
  }
  if (!descs.empty()) {
    bsql_batch* b = nullptr;
    int rc = bsql_batch_prepare(descs.data(), (int)descs.size(), &b);
    if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
    std::unique_ptr<bsql_batch, BatchFree> hold(b);
    rc = bsql_batch_launch(b, stream);
    if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
    // the batch's device plans must outlive the launch that reads them
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) throw Error(device_error(e, "batched projection"), "batched projection failed");
  }
  return res;
}

bool bypass_kind(const std::string& query_part, const std::vector<std::string>& names, bool* with_aliases,
                 std::vector<std::string>* aliases) {
  std::vector<std::string> exprs;
  project_items(query_part, aliases, &exprs);
  *with_aliases = false;
  if (names.size() != aliases->size()) return false;
  for (size_t i = 0; i < exprs.size(); ++i)
    if (exprs[i] != "$" + std::to_string(i)) return false;
  for (size_t i = 0; i < aliases->size(); ++i)
    if ((*aliases)[i] != names[i]) *with_aliases = true;
  return true;
}

std::vector<std::string> names_of(const bsql_table& t) {
  std::vector<std::string> n;
  for (const bsql_table::Col& c : t.cols) n.push_back(c.name);
  return n;
}

}  // namespace
}  // namespace bsql

using namespace bsql;

extern "C" {

int bsql_set_allocator(bsql_alloc_fn alloc, bsql_free_fn release, void* ctx) {
  if ((alloc == nullptr) != (release == nullptr))
    return set_error(BSQL_ERR_INVALID, "allocator needs both functions or neither");
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_alloc = alloc ? Allocator{alloc, release, ctx} : Allocator{default_alloc, default_free, nullptr};
  return BSQL_OK;
}

int bsql_table_create(const int32_t* types, const char* const* names, const uint8_t* nullable,
                      const int64_t* chars_bytes, int n_cols, int64_t num_rows, void* stream, bsql_table** out) {
  BSQL_TRY({
    if (!out || n_cols < 0 || (n_cols > 0 && !types) || num_rows < 0) fail("invalid arguments");
    *out = nullptr;
    TablePtr t = new_table(stream);
    t->num_rows = num_rows;
    for (int k = 0; k < n_cols; ++k) {
      if (types[k] == BSQL_STRING && !chars_bytes) fail("string column without chars_bytes");
      add_column(*t, types[k], names && names[k] ? names[k] : "", num_rows, nullable && nullable[k],
                 types[k] == BSQL_STRING ? chars_bytes[k] : 0);
    }
    *out = t.release();
    return BSQL_OK;
  })
}

int bsql_table_from_views(const bsql_column* cols, const char* const* names, int n_cols, int64_t num_rows,
                          bsql_table** out) {
  BSQL_TRY({
    if (!out || n_cols < 0 || (n_cols > 0 && !cols) || num_rows < 0) fail("invalid arguments");
    *out = nullptr;
    TablePtr t = new_table(nullptr);
    t->num_rows = num_rows;
    for (int k = 0; k < n_cols; ++k) {
      if (cols[k].size < num_rows) fail("column " + std::to_string(k) + " is shorter than the table");
      t->cols.emplace_back();
      t->cols.back().view = cols[k];
      t->cols.back().name = names && names[k] ? names[k] : "";
    }
    *out = t.release();
    return BSQL_OK;
  })
}

void bsql_table_free(bsql_table* table) { delete table; }

int64_t bsql_table_num_rows(const bsql_table* t) { return t ? t->num_rows : -1; }

int bsql_table_num_columns(const bsql_table* t) { return t ? (int)t->cols.size() : -1; }

int bsql_table_column(const bsql_table* t, int k, bsql_column* out) {
  if (!t || !out || k < 0 || k >= (int)t->cols.size()) return set_error(BSQL_ERR_INVALID, "no such column");
  *out = t->cols[(size_t)k].view;
  return BSQL_OK;
}

const char* bsql_table_name(const bsql_table* t, int k) {
  if (!t || k < 0 || k >= (int)t->cols.size()) return nullptr;
  return t->cols[(size_t)k].name.c_str();
}

int bsql_table_set_names(bsql_table* t, const char* const* names, int n) {
  if (!t || !names || n != (int)t->cols.size()) return set_error(BSQL_ERR_INVALID, "one name per column");
  for (int k = 0; k < n; ++k) t->cols[(size_t)k].name = names[k] ? names[k] : "";
  return BSQL_OK;
}

int bsql_table_record_stream(bsql_table* t, void* stream) {
  if (!t) return set_error(BSQL_ERR_INVALID, "null table");
  t->stream = stream;
  return BSQL_OK;
}

int bsql_process_filter(const bsql_table* in, const char* query_part, bsql_table** out, void* stream) {
  BSQL_TRY({
    if (!in || !query_part || !out) fail("null argument");
    *out = nullptr;
    const std::string qp(query_part);
    if (in->num_rows == 0) {
      *out = empty_like(*in, stream).release();
      return BSQL_OK;
    }
    std::string cond = named_expression(qp, "condition");
    if (cond.empty()) cond = named_expression(qp, "filters");
    if (cond.empty()) fail("query part has neither condition=[...] nor filters=[[...]]: " + qp);
    const std::vector<int32_t> types = types_of(*in);
    ProgramPtr pred = compile({cond}, types);
    if (output_type(pred.get(), types, 0) != BSQL_BOOL8) fail("Expression does not evaluate to a boolean mask");
    *out = filter_table(*in, pred.get(), {}, stream).release();
    const_cast<bsql_table*>(in)->stream = stream;
    return BSQL_OK;
  })
}

int bsql_apply_boolean_filter(const bsql_table* in, const bsql_column* mask, bsql_table** out, void* stream) {
  BSQL_TRY({
    if (!in || !mask || !out) fail("null argument");
    *out = nullptr;
    if (mask->type_id != BSQL_BOOL8) fail("the mask must be a BOOL8 column");
    if (mask->size < in->num_rows) fail("the mask is shorter than the table");
    if (in->num_rows == 0) {
      *out = empty_like(*in, stream).release();
      return BSQL_OK;
    }
    // the mask is the predicate's one input, appended after the table's
    std::vector<int32_t> types = types_of(*in);
    types.push_back(BSQL_BOOL8);
    ProgramPtr pred = compile({"$" + std::to_string(in->cols.size())}, types);
    *out = filter_table(*in, pred.get(), {*mask}, stream).release();
    const_cast<bsql_table*>(in)->stream = stream;
    return BSQL_OK;
  })
}

int bsql_evaluate_expressions(const bsql_table* in, const char* const* exprs, int n_exprs, bsql_table** out,
                              void* stream) {
  BSQL_TRY({
    if (!in || !out || n_exprs < 0 || (n_exprs > 0 && !exprs)) fail("invalid arguments");
    *out = nullptr;
    std::vector<std::string> e;
    for (int i = 0; i < n_exprs; ++i) {
      if (!exprs[i]) fail("null expression");
      e.emplace_back(exprs[i]);
    }
    *out = evaluate(*in, e, {}, stream).release();
    const_cast<bsql_table*>(in)->stream = stream;
    return BSQL_OK;
  })
}

int bsql_process_project(bsql_table* in, const char* query_part, bsql_table** out, void* stream) {
  BSQL_TRY({
    if (!in || !query_part || !out) fail("null argument");
    *out = nullptr;
    bool with_aliases = false;
    std::vector<std::string> aliases;
    if (bypass_kind(query_part, names_of(*in), &with_aliases, &aliases)) {
      for (size_t k = 0; k < aliases.size(); ++k) in->cols[k].name = aliases[k];
      in->stream = stream;
      *out = in;
      return BSQL_OK;
    }
    std::vector<std::string> names, exprs;
    project_items(query_part, &names, &exprs);
    TablePtr res = evaluate(*in, exprs, names, stream);
    in->stream = stream;
    delete in;
    *out = res.release();
    return BSQL_OK;
  })
}

int64_t bsql_table_size_in_bytes(const bsql_table* t) {
  if (!t) return -1;
  int64_t total = 0;
  for (const bsql_table::Col& c : t->cols) {
    const int64_t rows = c.view.size;
    if (c.view.type_id == BSQL_STRING)
      total += (int64_t)c.view.chars_size + (rows > 0 ? (rows + 1) * 4 : 0);
    else
      total += rows * type_size(c.view.type_id);
    if (c.view.validity) total += (int64_t)validity_bytes(rows);
  }
  return total;
}

int bsql_process_project_batches(bsql_table* const* ins, int n, const char* query_part, bsql_table** outs,
                                 void* stream) {
  BSQL_TRY({
    if (!ins || n <= 0 || !query_part || !outs) fail("invalid arguments");
    for (int i = 0; i < n; ++i) {
      if (!ins[i]) fail("batch " + std::to_string(i) + " is null");
      outs[i] = nullptr;
    }
    // an identity projection moves every batch through (renamed), as
    // bsql_process_project does batch by batch
    bool all_bypass = true;
    for (int i = 0; i < n && all_bypass; ++i) {
      bool wa = false;
      std::vector<std::string> al;
      all_bypass = bypass_kind(query_part, names_of(*ins[i]), &wa, &al);
    }
    if (all_bypass) {
      for (int i = 0; i < n; ++i) {
        bool wa = false;
        std::vector<std::string> aliases;
        bypass_kind(query_part, names_of(*ins[i]), &wa, &aliases);
        for (size_t k = 0; k < aliases.size(); ++k) ins[i]->cols[k].name = aliases[k];
        ins[i]->stream = stream;
        outs[i] = ins[i];
      }
      return BSQL_OK;
    }
    std::vector<std::string> names, exprs;
    project_items(query_part, &names, &exprs);
    std::vector<const bsql_table*> v(ins, ins + n);
    std::vector<TablePtr> res = evaluate_batches(v, exprs, names, stream);
    for (int i = 0; i < n; ++i) {
      ins[i]->stream = stream;
      delete ins[i];
      outs[i] = res[(size_t)i].release();
    }
    return BSQL_OK;
  })
}

int bsql_project_bypass(const char* query_part, const char* const* names, int n_names, int* bypass,
                        int* with_aliases) {
  BSQL_TRY({
    if (!query_part || !bypass || !with_aliases || n_names < 0 || (n_names > 0 && !names))
      fail("invalid arguments");
    std::vector<std::string> nm;
    for (int i = 0; i < n_names; ++i) nm.emplace_back(names[i] ? names[i] : "");
    bool wa = false;
    std::vector<std::string> aliases;
    *bypass = bypass_kind(query_part, nm, &wa, &aliases) ? 1 : 0;
    *with_aliases = wa ? 1 : 0;
    return BSQL_OK;
  })
}

int bsql_table_set_project_aliases(bsql_table* t, const char* query_part) {
  BSQL_TRY({
    if (!t || !query_part) fail("null argument");
    std::vector<std::string> names, exprs;
    project_items(query_part, &names, &exprs);
    if (names.size() != t->cols.size()) fail("the projection names a different number of columns");
    for (size_t k = 0; k < names.size(); ++k) t->cols[k].name = names[k];
    return BSQL_OK;
  })
}

int bsql_table_fix_column_aliases(bsql_table* t, const char* query_part) {
  BSQL_TRY({
    if (!t || !query_part) fail("null argument");
    const std::vector<std::string> aliases = expression_list(named_expression(query_part, "aliases"));
    // more aliases than columns (window functions): names unchanged
    if (aliases.size() > t->cols.size()) return BSQL_OK;
    for (size_t k = 0; k < aliases.size(); ++k) t->cols[k].name = aliases[k];
    return BSQL_OK;
  })
}

int bsql_scan_projections(const char* query_part, int num_columns, int32_t* cols, int cap) {
  try {
    if (!query_part || num_columns < 0 || (cap > 0 && !cols)) fail("invalid arguments");
    const std::string qp(query_part);
    std::vector<int> proj;
    if (qp.find(" projects=[[]]") != std::string::npos) {
      // count(*): an empty projection list with one alias loads column 0
      if (expression_list(named_expression(qp, "aliases")).size() == 1) proj.push_back(0);
    } else {
      for (const std::string& s : expression_list(named_expression(qp, "projects"))) {
        size_t used = 0;
        const int v = std::stoi(s, &used);
        if (used != s.size() || v < 0 || v >= num_columns) fail("invalid projection index: " + s);
        proj.push_back(v);
      }
    }
    if (proj.empty())
      for (int i = 0; i < num_columns; ++i) proj.push_back(i);
    for (int i = 0; i < (int)proj.size() && i < cap; ++i) cols[i] = proj[(size_t)i];
    return (int)proj.size();
  } catch (const Error& e) {
    return -set_error(e.code, e.what());
  } catch (const std::exception& e) {
    return -set_error(BSQL_ERR_INVALID, e.what());
  }
}

int bsql_aggregation_input(const bsql_table* in, const char* expr, int kind, bsql_table** out, void* stream) {
  BSQL_TRY({
    if (!in || !expr || !out) fail("null argument");
    *out = nullptr;
    std::string e(expr);
    // one column of `in`, borrowed
    auto column = [&](int k) {
      if (k < 0 || k >= (int)in->cols.size()) fail("aggregation input column " + std::to_string(k) + " out of range");
      TablePtr t = new_table(stream);
      t->num_rows = in->num_rows;
      t->cols.emplace_back();
      t->cols.back().view = in->cols[(size_t)k].view;
      t->cols.back().name = in->cols[(size_t)k].name;
      return t;
    };
    if (e.empty()) {
      if (kind != BSQL_AGG_COUNT_ALL) fail("an aggregation without an input expression must be COUNT(*)");
      *out = column(0).release();  // COUNT(*): the first column stands for the rows
      return BSQL_OK;
    }
    const size_t dollars = (size_t)std::count(e.begin(), e.end(), '$');
    if (dollars > 1 && kind == BSQL_AGG_COUNT_VALID) e = multi_column_count_expression(e);
    // is_number (expression_utils.cpp:358-361) / get_index
    // (CalciteExpressionParsing.cpp:48-52)
    const bool number = !e.empty() && std::all_of(e.begin(), e.end(), [](char c) { return c >= '0' && c <= '9'; });
    if (dollars == 1 || number) {
      const std::string idx = e[0] == '$' ? e.substr(1) : e;
      size_t used = 0;
      const int k = std::stoi(idx, &used);
      if (used != idx.size()) fail("aggregation input is not a column: " + e);
      *out = column(k).release();
      return BSQL_OK;
    }
    *out = evaluate(*in, {e}, {}, stream).release();
    const_cast<bsql_table*>(in)->stream = stream;
    return BSQL_OK;
  })
}

int bsql_skip_data_filter(const char* filter, const uint8_t* has_stats, int n_cols, char* out, size_t cap,
                          size_t* len) {
  BSQL_TRY({
    if (!filter || n_cols < 0 || (n_cols > 0 && !has_stats)) fail("invalid arguments");
    std::vector<bool> st;
    for (int i = 0; i < n_cols; ++i) st.push_back(has_stats[i] != 0);
    const std::string r = skip_data_filter(filter, st);
    if (len) *len = r.size();
    if (out && cap > 0) {
      const size_t n = std::min(cap - 1, r.size());
      std::memcpy(out, r.data(), n);
      out[n] = '\0';
    }
    return BSQL_OK;
  })
}

int bsql_process_skipdata_for_table(const bsql_table* metadata, const char* const* names, int n_names,
                                    const char* table_scan, bsql_table** out, int* read_all, void* stream) {
  BSQL_TRY({
    if (!metadata || !table_scan || !out || !read_all || n_names < 0 || (n_names > 0 && !names))
      fail("invalid arguments");
    *out = nullptr;
    *read_all = 1;
    if (metadata->cols.size() < 2) fail("the statistics table needs the file and row-group id columns");
    const std::string scan(table_scan);
    std::string filter = named_expression(scan, "condition");
    if (filter.empty()) filter = named_expression(scan, "filters");
    if (filter.empty()) return BSQL_OK;
    // the projected columns (projects=[[..]], else all of them)
    std::vector<int> proj;
    const std::string projects = named_expression(scan, "projects");
    if (projects.empty()) {
      for (int i = 0; i < n_names; ++i) proj.push_back(i);
    } else {
      for (const std::string& p : expression_list(projects)) proj.push_back(std::stoi(p));
    }
    // their statistics: min and max columns, or placeholders the rewritten
    // filter never reads
    std::vector<bsql_column> cols;
    std::vector<bool> has;
    auto find = [&](const std::string& n) -> int {
      for (size_t k = 0; k < metadata->cols.size(); ++k)
        if (metadata->cols[k].name == n) return (int)k;
      return -1;
    };
    for (int c : proj) {
      if (c < 0 || c >= n_names || !names[c]) fail("projected column " + std::to_string(c) + " has no name");
      const std::string suffix = std::to_string(c) + "_" + names[c];
      const int lo = find("min_" + suffix), hi = find("max_" + suffix);
      if (lo >= 0 && hi >= 0) {
        has.push_back(true);
        cols.push_back(metadata->cols[(size_t)lo].view);
        cols.push_back(metadata->cols[(size_t)hi].view);
      } else {
        has.push_back(false);
        bsql_column none = metadata->cols[0].view;
        none.type_id = BSQL_INT8;
        none.validity = nullptr;
        cols.push_back(none);
        cols.push_back(none);
      }
    }
    const std::string expr = skip_data_filter(filter, has);
    if (expr.empty()) return BSQL_OK;
    TablePtr stats = new_table(stream);
    stats->num_rows = metadata->num_rows;
    for (const bsql_column& c : cols) {
      stats->cols.emplace_back();
      stats->cols.back().view = c;
    }
    TablePtr mask = evaluate(*stats, {expr}, {}, stream);
    if (mask->cols.size() != 1 || mask->cols[0].view.type_id != BSQL_BOOL8)
      fail("Expression in skip_data processing did not evaluate to a boolean mask");
    // the file and row-group ids of the row groups to read
    TablePtr ids = new_table(stream);
    ids->num_rows = metadata->num_rows;
    for (size_t k = metadata->cols.size() - 2; k < metadata->cols.size(); ++k) {
      ids->cols.emplace_back();
      ids->cols.back().view = metadata->cols[k].view;
      ids->cols.back().name = metadata->cols[k].name;
    }
    bsql_table* kept = nullptr;
    const int rc = bsql_apply_boolean_filter(ids.get(), &mask->cols[0].view, &kept, stream);
    if (rc != BSQL_OK) return rc;
    mask->stream = stream;
    *out = kept;
    *read_all = 0;
    return BSQL_OK;
  })
}

}  // extern "C"
