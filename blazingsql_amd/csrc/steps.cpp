// steps.cpp -- builds the columns a compiled program needs before the
// interpreter runs (string functions that make new strings, and expressions
// they take as arguments), the role function_evaluator_transformer plays in
// the reference (engine/src/execution_kernels/LogicalProject.cpp:670-731).
//
// Every buffer is allocated and released with the installed allocator
// (bsql_set_allocator; stream-ordered hipMallocAsync by default) on the call's
// stream, so a step column lives exactly as long as the launches that read it
// and an allocator with a pool limit sees the step columns' memory too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "strfn.h"

namespace bsql {

int64_t string_tiles(int64_t capacity);
hipError_t launch_strfn(StrFnArgs f, uint64_t* status, hipStream_t stream);

namespace {

// A step column's buffer from the installed allocator, on the call's stream.
void* stream_alloc(StepColumns& sc, int64_t bytes) {
  void* p = nullptr;
  const size_t n = (size_t)std::max<int64_t>(bytes, 8);
  const int rc = sc.alloc.alloc(sc.alloc.ctx, &p, n, sc.stream);
  if (rc == 1) throw Error(BSQL_ERR_OOM, "step column: device allocation of " + std::to_string(n) + " bytes failed");
  if (rc != 0 || !p) throw Error(BSQL_ERR_DEVICE, "step column: device allocation failed");
  sc.owned.emplace_back(p, n);
  return p;
}

int64_t validity_bytes(int64_t n) { return (n + 63) / 64 * 8; }

int64_t column_chars(const bsql_column& c) { return c.type_id == BSQL_STRING ? c.chars_size : 0; }

int64_t max_string_chars(const std::vector<bsql_column>& cols) {
  int64_t m = 0;
  for (const bsql_column& c : cols) m = std::max(m, column_chars(c));
  return m;
}

int64_t arg_chars(const StepArg& a, const std::vector<bsql_column>& cols, int64_t n) {
  if (a.kind == StepArg::LITERAL) return (int64_t)a.str.size() * n;
  if (a.kind == StepArg::NULL_LITERAL) return 0;
  return column_chars(cols[(size_t)a.col]);
}

// Upper bound of a string step's chars: the sum over its inputs, scaled by
// the most a function can grow a string (case mapping: 3x; REPLACE: one
// replacement per target length; a number or timestamp: 32 bytes a row).
int64_t step_chars(const Step& st, const std::vector<bsql_column>& cols, int64_t n) {
  if (st.kind == Step::EVAL) return st.out_type == BSQL_STRING ? max_string_chars(cols) : 0;
  const int64_t a0 = st.args.empty() ? 0 : arg_chars(st.args[0], cols, n);
  switch (st.op) {
    case SF_CONCAT: return a0 + arg_chars(st.args[1], cols, n);
    case SF_UPPER: case SF_LOWER: case SF_INITCAP: return 3 * a0;
    case SF_REVERSE: return a0;
    case SF_REPLACE: {
      const int64_t tn = (int64_t)st.args[1].str.size(), rn = (int64_t)st.args[2].str.size();
      return rn > tn ? a0 + a0 / tn * (rn - tn) : a0;
    }
    case SF_CAST: {
      const int32_t t = cols[(size_t)st.args[0].col].type_id;
      if (t == BSQL_STRING) return a0;
      return (t == BSQL_BOOL8 ? 5 : t <= BSQL_UINT64 ? 20 : 32) * n;
    }
    case SF_CASE: return arg_chars(st.args[1], cols, n) + arg_chars(st.args[2], cols, n);
    case SF_REGEXP_REPLACE: {
      // a row of b bytes has at most 2b + 1 matches (empty ones included),
      // each giving the replacement's L bytes plus, for each of its r
      // backreferences, at most the match's own bytes
      const std::string& rp = st.args[2].str;
      int64_t r = 0;
      for (size_t k = 0; k + 1 < rp.size(); ++k)
        if (rp[k] == '\\' && rp[k + 1] >= '1' && rp[k + 1] <= '9') ++r, ++k;
      const int64_t L = (int64_t)rp.size();
      return a0 * (1 + 2 * L + r) + n * L;
    }
    case SF_FILL: return a0;
    default: return 0;
  }
}

// Device copy of a step's literal arguments, one {bytes} record per argument.
const uint8_t* step_literals(bsql_program& p, Step& st, std::vector<int32_t>* at) {
  std::vector<uint8_t> blob;
  for (size_t k = 0; k < st.args.size(); ++k) {
    const StepArg& a = st.args[k];
    while (blob.size() % 8) blob.push_back(0);
    at->push_back((int32_t)blob.size());
    if (a.kind != StepArg::LITERAL) continue;
    if (st.op == SF_REGEXP_REPLACE && k == 1) {  // the pattern, compiled for the matcher
      const std::vector<int32_t> prog = regex_compile(a.str);
      blob.insert(blob.end(), (const uint8_t*)prog.data(), (const uint8_t*)(prog.data() + prog.size()));
    } else {
      blob.insert(blob.end(), a.str.begin(), a.str.end());
    }
  }
  std::lock_guard<std::mutex> lk(p.cache_mu);
  if (!st.d_lits && !blob.empty()) {
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, blob.size());
    if (e == hipSuccess) e = hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (d) (void)hipFree(d);
      throw Error(device_error(e, "step literals"), "step literals");
    }
    st.d_lits = std::shared_ptr<void>(d, [](void* q) {
      if (!process_exiting()) (void)hipFree(q);
    });
  }
  return (const uint8_t*)st.d_lits.get();
}

StrSrc source_of(const StepArg& a, const std::vector<bsql_column>& cols, const uint8_t* lits, int32_t at) {
  StrSrc s;
  std::memset(&s, 0, sizeof(s));
  if (a.kind == StepArg::LITERAL) {
    s.kind = SRC_LITERAL;
    s.type = BSQL_STRING;
    s.lit = lits + at;
    s.lit_len = (int32_t)a.str.size();
    return s;
  }
  if (a.kind == StepArg::NULL_LITERAL) {
    s.kind = SRC_NULL;
    return s;
  }
  const bsql_column& c = cols[(size_t)a.col];
  s.type = c.type_id;
  s.valid = c.validity ? c.validity + (c.offset >> 5) : nullptr;
  s.vbit = (int32_t)(c.offset & 31);
  if (c.type_id == BSQL_STRING) {
    s.kind = SRC_STRING;
    s.offsets = c.offsets + c.offset;
    s.chars = (const uint8_t*)c.data;
  } else {
    s.kind = SRC_FIXED;
    s.data = (const char*)c.data + c.offset * type_size(c.type_id);
  }
  return s;
}

}  // namespace

StepColumns::~StepColumns() {
  for (const auto& b : owned) alloc.release(alloc.ctx, b.first, b.second, stream);
}

std::vector<int32_t> extended_types(const Program& p, const int32_t* types, int n) {
  std::vector<int32_t> t(types, types + std::max(n, 0));
  for (const Step& st : p.steps) t.push_back(st.out_type);
  return t;
}

void build_steps(bsql_program& p, StepColumns& sc, int64_t num_rows) {
  if (p.steps.empty()) return;
  if ((int)sc.cols.size() != p.n_table)
    fail("the program was compiled for a table of " + std::to_string(p.n_table) + " columns, not " +
         std::to_string(sc.cols.size()));
  const int64_t n = num_rows;
  for (Step& st : p.steps) {
    bsql_column c;
    std::memset(&c, 0, sizeof(c));
    c.type_id = st.out_type;
    c.size = n;
    c.validity = (uint32_t*)stream_alloc(sc, validity_bytes(n));
    const int64_t chars = step_chars(st, sc.cols, n);
    if (st.out_type == BSQL_STRING) {
      if (chars > INT32_MAX) fail("a string column built by a step would exceed 2^31-1 chars");
      c.chars_size = (int32_t)chars;
      c.offsets = (int32_t*)stream_alloc(sc, (n + 1) * 4);
      c.data = stream_alloc(sc, chars + 1);
    } else {
      c.data = stream_alloc(sc, type_size(st.out_type) * n);
    }
    if (st.kind == Step::EVAL) {
      bsql_program* sub = static_cast<bsql_program*>(st.sub.get());
      const int rc = execute_program(sub, sc.cols.data(), (int)sc.cols.size(), &c, 1, n, sc.stream, true);
      if (rc != BSQL_OK) throw Error(rc, bsql_last_error());
    } else {
      std::vector<int32_t> at;
      const uint8_t* lits = step_literals(p, st, &at);
      StrFnArgs f;
      std::memset(&f, 0, sizeof(f));
      f.op = st.op;
      f.n_args = (int32_t)st.args.size();
      for (size_t k = 0; k < st.args.size(); ++k) f.a[k] = source_of(st.args[k], sc.cols, lits, at[k]);
      f.num_rows = n;
      f.start = st.start;
      f.lens = (uint64_t*)stream_alloc(sc, n * 8);
      f.out_valid = c.validity;
      f.out_offsets = c.offsets;
      f.out_chars = (uint8_t*)c.data;
      f.chars_cap = chars;
      uint64_t* status = (uint64_t*)stream_alloc(sc, (string_tiles(n) + 1) * 8);
      const hipError_t e = launch_strfn(f, status, sc.stream);
      if (e != hipSuccess) throw Error(device_error(e, "string step"), bsql_last_error());
    }
    sc.cols.push_back(c);
  }
}

bsql_program* rebased_program(bsql_program* p, int n_table) {
  if (p->steps.empty() || p->n_table == n_table) return p;
  std::lock_guard<std::mutex> lk(p->cache_mu);
  auto it = p->rebased.find(n_table);
  if (it != p->rebased.end()) return it->second.get();
  const int shift = n_table - p->n_table;
  const int base = p->n_table;
  auto move_cols = [&](Program& q) {
    for (int32_t& c : q.used_columns)
      if (c >= base) c += shift;
    for (Root& rt : q.roots)
      if (rt.kind == Root::COLUMN && rt.index >= base) rt.index += shift;
  };
  std::unique_ptr<bsql_program> r(new bsql_program());
  static_cast<Program&>(*r) = static_cast<const Program&>(*p);
  r->d_strings = nullptr;
  r->str_offsets.clear();
  r->n_table = n_table;
  move_cols(*r);
  for (Step& st : r->steps) {
    for (StepArg& a : st.args)
      if (a.kind == StepArg::COLUMN && a.col >= base) a.col += shift;
    if (st.sub) {
      auto copy = std::make_unique<Program>(static_cast<const Program&>(*st.sub));
      copy->d_strings = nullptr;
      copy->str_offsets.clear();
      move_cols(*copy);
      copy->n_table += shift;
      st.sub = make_program_handle(std::move(copy));
    }
    st.d_lits.reset();
  }
  bsql_program* out = r.get();
  p->rebased.emplace(n_table, std::move(r));
  return out;
}

int64_t string_chars_bound(const bsql_program& p, const bsql_column* in, int n_in, int64_t num_rows) {
  std::vector<bsql_column> cols(in, in + std::max(n_in, 0));
  for (const Step& st : p.steps) {
    bsql_column c;
    std::memset(&c, 0, sizeof(c));
    c.type_id = st.out_type;
    c.chars_size = (int32_t)std::min<int64_t>(step_chars(st, cols, num_rows), INT32_MAX);
    cols.push_back(c);
  }
  return max_string_chars(cols);
}

std::shared_ptr<Program> make_program_handle(std::unique_ptr<Program> p) {
  bsql_program* bp = new bsql_program();
  static_cast<Program&>(*bp) = *p;
  return std::shared_ptr<Program>(static_cast<Program*>(bp),
                                  [](Program* q) { delete static_cast<bsql_program*>(q); });
}

}  // namespace bsql

extern "C" {

int bsql_program_num_steps(const bsql_program* p) { return p ? (int)p->steps.size() : -1; }

int bsql_program_table_columns(const bsql_program* p) { return p ? p->n_table : -1; }

int bsql_program_step(const bsql_program* p, int k, bsql_step_info* out) {
  using namespace bsql;
  if (!p || !out || k < 0 || k >= (int)p->steps.size()) return set_error(BSQL_ERR_INVALID, "invalid arguments");
  const Step& st = p->steps[(size_t)k];
  std::memset(out, 0, sizeof(*out));
  out->kind = st.kind == Step::EVAL ? 0 : 1;
  out->op = st.op;
  out->out_type = st.out_type;
  out->n_args = (int32_t)st.args.size();
  for (size_t i = 0; i < st.args.size() && i < 3; ++i) {
    const StepArg& a = st.args[i];
    out->arg_col[i] = a.kind == StepArg::COLUMN ? a.col : (a.kind == StepArg::LITERAL ? -1 : -2);
    out->arg_str[i] = a.str.c_str();
    out->arg_len[i] = (int32_t)a.str.size();
  }
  out->sub = static_cast<const bsql_program*>(st.sub.get());
  out->start = st.start;
  return BSQL_OK;
}

int bsql_program_output_chars(const bsql_program* p, const bsql_column* in, int n_in, int64_t num_rows,
                              int64_t* chars) {
  BSQL_TRY({
    if (!p || !chars || n_in < 0 || (n_in > 0 && !in)) bsql::fail("invalid arguments");
    *chars = bsql::string_chars_bound(*p, in, n_in, num_rows);
    return BSQL_OK;
  })
}

}  // extern "C"
