// interp_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// CPU restatement of BlazingSQL's row-parallel expression interpreter and of
// the filter (boolean-mask compaction) path, used only as the correctness
// checker of the HIP implementation and as the host-core CPU baseline.
// Nothing in blazingsql_amd/ links or calls this file; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// It follows the reference op by op:
//   row loop, slot buffer and validity word handling
//       engine/src/Interpreter/interpreter_ops.cuh:204-240, 397-426
//   input conversion (integral -> int64, float -> double, ticks)
//       interpreter_ops.cuh:267-324
//   output conversion (static_cast to the output column type)
//       interpreter_ops.cuh:329-395
//   operand type dispatch (LeftType/RightType = double iff FLOAT32/FLOAT64)
//       interpreter_ops.cuh:479-498
//   per-op semantics (binary, unary, nullary branches)
//       interpreter_ops.cuh:500-925
//   scalar value conversion get_scalar_value<T>      interpreter_ops.cuh:428-477
//   magic numbers of the CASE encoding               interpreter_ops.cuh:35-48
//   EXTRACT components                               interpreter_ops.cuh:69-99
//   per-op operand type propagation                  engine/src/Interpreter/interpreter_cpp.cu:502-550
//   output type rules get_output_type                engine/src/parser/expression_utils.cpp:115-266
//   filter: null/false drops the row, stable order  engine/src/execution_kernels/LogicalFilter.cpp:22-52
//   string functions inside programs                 engine/src/execution_kernels/LogicalProject.cpp:
//       LIKE :31-47 and :113-131, SUBSTRING/LEFT/RIGHT :183-306, CAST(varchar)
//       :379-468 (libcudf 21.08 to_integers / to_floats), TRIM :566-579 (strip)
//
// CAST(varchar AS FLOAT/DOUBLE) multiplies by exp10(e); here that factor is
// the double nearest 10^e, so its last bit is parity unpinned where the
// reference's exp10 is not correctly rounded.
//
// Values the reference leaves undefined are given one fixed definition here
// and in the HIP kernels alike (they are "parity unpinned" against the
// reference): integer x/0 (result is null, data 0), x%0 -> 0,
// INT64_MIN/-1 -> INT64_MIN, INT64_MIN%-1 -> 0, double->int64 of NaN or
// out-of-range values -> INT64_MIN (the x86 cvttsd2si result), data under null
// results, and validity bits at or beyond num_rows (written as 0).
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fno-fast-math -fPIC -shared
//        -pthread (see oracle/build_oracle.py).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <regex>
#include <string>
#include <thread>
#include <vector>

#include "../include/bsql/interp.h"

namespace {

typedef int16_t slot_t;

// ---------------------------------------------------------------- type helpers
bool is_float_type(int t) { return t == BSQL_FLOAT32 || t == BSQL_FLOAT64; }
bool is_timestamp_type(int t) { return t >= BSQL_TIMESTAMP_DAYS && t <= BSQL_TIMESTAMP_NANOSECONDS; }
bool is_duration_type(int t) { return t >= BSQL_DURATION_DAYS && t <= BSQL_DURATION_NANOSECONDS; }
bool is_string_type(int t) { return t == BSQL_STRING; }
bool is_integral_type(int t) { return (t >= BSQL_INT8 && t <= BSQL_UINT64) || t == BSQL_BOOL8; }

int type_size(int t) {
  switch (t) {
    case BSQL_INT8: case BSQL_UINT8: case BSQL_BOOL8: return 1;
    case BSQL_INT16: case BSQL_UINT16: return 2;
    case BSQL_INT32: case BSQL_UINT32: case BSQL_FLOAT32: case BSQL_TIMESTAMP_DAYS:
    case BSQL_DURATION_DAYS: case BSQL_DECIMAL32: case BSQL_DICTIONARY32: return 4;
    case BSQL_INT64: case BSQL_UINT64: case BSQL_FLOAT64:
    case BSQL_TIMESTAMP_SECONDS: case BSQL_TIMESTAMP_MILLISECONDS:
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_TIMESTAMP_NANOSECONDS:
    case BSQL_DURATION_SECONDS: case BSQL_DURATION_MILLISECONDS:
    case BSQL_DURATION_MICROSECONDS: case BSQL_DURATION_NANOSECONDS:
    case BSQL_DECIMAL64: return 8;
    default: return 0;
  }
}

// ticks per day and ticks->ns factor of each timestamp/duration unit
int64_t ticks_per_day(int t) {
  switch (t) {
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return 1;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_DURATION_SECONDS: return 86400LL;
    case BSQL_TIMESTAMP_MILLISECONDS: case BSQL_DURATION_MILLISECONDS: return 86400000LL;
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_DURATION_MICROSECONDS: return 86400000000LL;
    default: return 86400000000000LL;
  }
}
int64_t ns_per_tick(int t) {
  switch (t) {
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return 86400000000000LL;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_DURATION_SECONDS: return 1000000000LL;
    case BSQL_TIMESTAMP_MILLISECONDS: case BSQL_DURATION_MILLISECONDS: return 1000000LL;
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_DURATION_MICROSECONDS: return 1000LL;
    default: return 1LL;
  }
}

int64_t wrap_mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
int64_t wrap_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
int64_t wrap_sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }

int64_t f64_to_i64(double v) {
  if (std::isnan(v) || v >= 9223372036854775808.0 || v < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)v;
}

int64_t idiv(int64_t a, int64_t b) {
  if (b == 0) return 0;
  if (b == -1) return wrap_sub(0, a);
  return a / b;
}
int64_t imod(int64_t a, int64_t b) {
  if (b == 0 || b == -1) return 0;
  return a % b;
}

// a value held in one 8-byte slot: int64 or the bits of a double
struct Val {
  bool is_f;
  int64_t i;
  double d;
};
double as_f(const Val& v) { return v.is_f ? v.d : (double)v.i; }
bool truthy(const Val& v) { return v.is_f ? (v.d != 0.0) : (v.i != 0); }

int64_t bits_of(double d) { int64_t b; std::memcpy(&b, &d, 8); return b; }
double dbl_of(int64_t b) { double d; std::memcpy(&d, &b, 8); return d; }

const int64_t MAGIC_I64 = INT64_MAX - 13LL;   // interpreter_ops.cuh:41-43
const double MAGIC_F64 = 1.7976931348623123e+308;  // interpreter_ops.cuh:46-48

// ------------------------------------------------------------------ calendar
// days since 1970-01-01 -> civil date (proleptic Gregorian), Hinnant's
// civil_from_days, as cuda::std::chrono::year_month_day computes it.
void civil_from_days(int64_t z, int64_t* y, unsigned* m, unsigned* d) {
  z += 719468;
  const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const int64_t yy = (int64_t)yoe + era * 400;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  *d = doy - (153 * mp + 2) / 5 + 1;
  *m = mp < 10 ? mp + 3 : mp - 9;
  *y = yy + (*m <= 2);
}
int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

// EXTRACT (interpreter_ops.cuh:69-99): component of a timestamp of type t
int64_t extract(int op, int t, int64_t ticks) {
  const int64_t tpd = ticks_per_day(t);
  int64_t days = floor_div(ticks, tpd);
  int64_t tsm = ticks - days * tpd;  // time since midnight, in ticks
  if (tsm < 0) tsm += tpd;
  const int64_t tph = tpd / 24, tpm = tpd / 1440, tps = tpd / 86400;
  int64_t hrs = tph ? tsm / tph : 0;
  int64_t mins = tpm ? (tsm - hrs * tph) / tpm : 0;
  int64_t secs = tps ? (tsm - hrs * tph - mins * tpm) / tps : 0;
  int64_t y; unsigned m, d;
  switch (op) {
    case BSQL_OP_YEAR: civil_from_days(days, &y, &m, &d); return (int16_t)y;
    case BSQL_OP_MONTH: civil_from_days(days, &y, &m, &d); return (int16_t)m;
    case BSQL_OP_DAY: civil_from_days(days, &y, &m, &d); return (int16_t)d;
    case BSQL_OP_DAYOFWEEK: {
      int64_t wd = days >= -4 ? (days + 4) % 7 : (days + 5) % 7 + 6;  // 0 = Sunday
      return wd == 0 ? 7 : wd;
    }
    case BSQL_OP_HOUR: return (int16_t)hrs;
    case BSQL_OP_MINUTE: return (int16_t)mins;
    case BSQL_OP_SECOND: return (int16_t)secs;
  }
  return 0;
}

// ------------------------------------------------------------------- strings
struct StrView { const char* p; int32_t n; };
int str_cmp(StrView a, StrView b) {  // cudf::string_view::compare: bytewise
  int32_t n = std::min(a.n, b.n);
  int c = n ? std::memcmp(a.p, b.p, (size_t)n) : 0;
  if (c != 0) return c < 0 ? -1 : 1;
  return a.n < b.n ? -1 : (a.n > b.n ? 1 : 0);
}
int64_t utf8_length(StrView s) {
  int64_t n = 0;
  for (int32_t i = 0; i < s.n; ++i) n += ((unsigned char)s.p[i] & 0xC0) != 0x80;
  return n;
}

// UTF-8 <-> code points (the reference's string functions count characters)
std::u32string utf8_decode(StrView s) {
  std::u32string out;
  for (int32_t i = 0; i < s.n;) {
    unsigned char c = (unsigned char)s.p[i];
    int len = c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : 4;
    char32_t cp = len == 1 ? c : len == 2 ? (c & 0x1F) : len == 3 ? (c & 0x0F) : (c & 0x07);
    for (int k = 1; k < len && i + k < s.n; ++k) cp = (cp << 6) | ((unsigned char)s.p[i + k] & 0x3F);
    out.push_back(cp);
    i += len;
  }
  return out;
}
std::string utf8_encode(const std::u32string& s) {
  std::string out;
  for (char32_t cp : s) {
    if (cp < 0x80) out.push_back((char)cp);
    else if (cp < 0x800) { out.push_back((char)(0xC0 | (cp >> 6))); out.push_back((char)(0x80 | (cp & 0x3F))); }
    else if (cp < 0x10000) {
      out.push_back((char)(0xE0 | (cp >> 12))); out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
      out.push_back((char)(0x80 | (cp & 0x3F)));
    } else {
      out.push_back((char)(0xF0 | (cp >> 18))); out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
      out.push_back((char)(0x80 | ((cp >> 6) & 0x3F))); out.push_back((char)(0x80 | (cp & 0x3F)));
    }
  }
  return out;
}

// like_expression_to_regex_str (LogicalProject.cpp:31-47): '%' and '_' not
// escaped by a backslash become (?:.*?) and (?:.); anchors unless the pattern
// starts/ends with '%'. The regex is searched as cudf::strings::contains_re does.
std::string like_to_regex(const std::string& like) {
  if (like.empty()) return like;
  const bool match_start = like.front() != '%';
  const bool match_end = like.back() != '%';
  static const std::regex any_string_re{R"(([^\\]?|\\{2})%)"};
  static const std::regex any_char_re{R"(([^\\]?|\\{2})_)"};
  std::string re = std::regex_replace(like, any_string_re, "$1(?:.*?)");
  re = std::regex_replace(re, any_char_re, "$1(?:.)");
  return (match_start ? "^" : "") + re + (match_end ? "$" : "");
}
std::wstring to_wide(const std::u32string& s) { return std::wstring(s.begin(), s.end()); }

// slice_strings on characters [start, stop) with stop < 0 meaning the end
std::string slice_chars(StrView s, int64_t start, int64_t stop) {
  std::u32string cps = utf8_decode(s);
  const int64_t n = (int64_t)cps.size();
  if (stop < 0 || stop > n) stop = n;
  if (start < 0) start = 0;
  if (start >= stop) return std::string();
  return utf8_encode(cps.substr((size_t)start, (size_t)(stop - start)));
}

// ------------------------------------------------------------------- program
struct Prog {
  std::vector<slot_t> left, right, out, final_out;
  std::vector<int32_t> ops;
  std::vector<bsql_scalar> lsc, rsc;
  std::vector<int32_t> ltype, rtype, otype;  // interpreter_cpp.cu:502-550
};

int out_type_unary(int op, int lt) {  // expression_utils.cpp:126-199
  switch (op) {
    case BSQL_OP_CAST_TINYINT: return BSQL_INT8;
    case BSQL_OP_CAST_SMALLINT: return BSQL_INT16;
    case BSQL_OP_CAST_INTEGER: return BSQL_INT32;
    case BSQL_OP_CAST_BIGINT: return BSQL_INT64;
    case BSQL_OP_CAST_FLOAT: return BSQL_FLOAT32;
    case BSQL_OP_CAST_DOUBLE: return BSQL_FLOAT64;
    case BSQL_OP_CAST_DATE: return BSQL_TIMESTAMP_DAYS;
    case BSQL_OP_CAST_TIMESTAMP_SECONDS: return BSQL_TIMESTAMP_SECONDS;
    case BSQL_OP_CAST_TIMESTAMP_MILLISECONDS: return BSQL_TIMESTAMP_MILLISECONDS;
    case BSQL_OP_CAST_TIMESTAMP_MICROSECONDS: return BSQL_TIMESTAMP_MICROSECONDS;
    case BSQL_OP_CAST_TIMESTAMP: return BSQL_TIMESTAMP_NANOSECONDS;
    case BSQL_OP_STR_LOWER: case BSQL_OP_STR_UPPER: case BSQL_OP_STR_INITCAP:
    case BSQL_OP_STR_REVERSE: case BSQL_OP_CAST_VARCHAR: return BSQL_STRING;
    case BSQL_OP_YEAR: case BSQL_OP_MONTH: case BSQL_OP_DAY: case BSQL_OP_DAYOFWEEK:
    case BSQL_OP_HOUR: case BSQL_OP_MINUTE: case BSQL_OP_SECOND: return BSQL_INT16;
    case BSQL_OP_SIN: case BSQL_OP_COS: case BSQL_OP_ASIN: case BSQL_OP_ACOS:
    case BSQL_OP_TAN: case BSQL_OP_COTAN: case BSQL_OP_ATAN: case BSQL_OP_LN:
    case BSQL_OP_LOG: case BSQL_OP_FLOOR: case BSQL_OP_CEIL:
      return is_float_type(lt) ? lt : BSQL_FLOAT64;
    case BSQL_OP_ABS: return lt;
    case BSQL_OP_NOT: case BSQL_OP_IS_TRUE: case BSQL_OP_IS_NULL: case BSQL_OP_IS_NOT_NULL:
    case BSQL_OP_IS_NOT_TRUE: case BSQL_OP_IS_NOT_FALSE: return BSQL_BOOL8;
    case BSQL_OP_CHAR_LENGTH: return BSQL_INT32;
  }
  return BSQL_EMPTY;
}
int out_type_binary(int op, int lt, int rt) {  // expression_utils.cpp:201-266
  if (lt == BSQL_EMPTY) lt = rt; else if (rt == BSQL_EMPTY) rt = lt;
  switch (op) {
    case BSQL_OP_ADD: case BSQL_OP_SUB: case BSQL_OP_MUL: case BSQL_OP_DIV: case BSQL_OP_MOD:
      if (is_float_type(lt) && is_float_type(rt)) return type_size(lt) >= type_size(rt) ? lt : rt;
      if (is_float_type(lt)) return lt;
      if (is_float_type(rt)) return rt;
      return type_size(lt) >= type_size(rt) ? lt : rt;
    case BSQL_OP_EQUAL: case BSQL_OP_NOT_EQUAL: case BSQL_OP_LESS: case BSQL_OP_GREATER:
    case BSQL_OP_LESS_EQUAL: case BSQL_OP_GREATER_EQUAL: case BSQL_OP_LOGICAL_AND:
    case BSQL_OP_LOGICAL_OR: case BSQL_OP_STR_LIKE: return BSQL_BOOL8;
    case BSQL_OP_POW: case BSQL_OP_ROUND: return type_size(lt) <= 4 ? BSQL_FLOAT32 : BSQL_FLOAT64;
    case BSQL_OP_MAGIC_IF_NOT: return rt;
    case BSQL_OP_FIRST_NON_MAGIC: return type_size(lt) >= type_size(rt) ? lt : rt;
    case BSQL_OP_STR_SUBSTRING: case BSQL_OP_STR_LEFT: case BSQL_OP_STR_RIGHT: case BSQL_OP_STR_TRIM:
      return BSQL_STRING;
    case BSQL_OP_TO_DATE: return BSQL_TIMESTAMP_DAYS;
    case BSQL_OP_TO_TIMESTAMP: return BSQL_TIMESTAMP_NANOSECONDS;
  }
  return BSQL_EMPTY;  // IS_NOT_DISTINCT_FROM and others fall to the default branch
}

bool is_unary_op(int op) { return op >= BSQL_OP_NOT && op <= BSQL_OP_STR_REVERSE; }

void propagate_types(Prog& p, const int32_t* in_types, int n_in) {
  const size_t n = p.ops.size();
  p.ltype.assign(n, 0); p.rtype.assign(n, 0); p.otype.assign(n, 0);
  int32_t slot_type[512] = {0};  // output_map_type (default EMPTY)
  for (size_t i = 0; i < n; ++i) {
    int l = p.left[i], r = p.right[i];
    if (l >= 0 && l < n_in) p.ltype[i] = in_types[l];
    else if (l == BSQL_SCALAR_NULL_INDEX) p.ltype[i] = BSQL_EMPTY;
    else if (l == BSQL_SCALAR_INDEX) p.ltype[i] = p.lsc[i].type_id;
    else if (l >= 0) p.ltype[i] = slot_type[l];
    if (r >= 0 && r < n_in) p.rtype[i] = in_types[r];
    else if (r == BSQL_SCALAR_NULL_INDEX || r == BSQL_UNARY_INDEX) p.rtype[i] = BSQL_EMPTY;
    else if (r == BSQL_SCALAR_INDEX) p.rtype[i] = p.rsc[i].type_id;
    else if (r >= 0) p.rtype[i] = slot_type[r];
    if (r == BSQL_UNARY_INDEX) p.otype[i] = out_type_unary(p.ops[i], p.ltype[i]);
    else if (r == BSQL_NULLARY_INDEX) p.otype[i] = BSQL_FLOAT64;
    else p.otype[i] = out_type_binary(p.ops[i], p.ltype[i], p.rtype[i]);
    if (p.out[i] >= 0) slot_type[p.out[i]] = p.otype[i];
  }
}

// --------------------------------------------------------------------- table
struct Table {
  const bsql_column* cols;
  int n;
};

bool col_valid(const bsql_column& c, int64_t row) {
  if (!c.validity) return true;
  int64_t b = row + c.offset;
  return (c.validity[b >> 5] >> (b & 31)) & 1u;
}

StrView col_str(const bsql_column& c, int64_t row) {
  int64_t r = row + c.offset;
  int32_t s = c.offsets[r], e = c.offsets[r + 1];
  return StrView{(const char*)c.data + s, e - s};
}

// read_data (interpreter_ops.cuh:267-324): returns false for types not loaded
bool read_elem(const bsql_column& c, int64_t row, int64_t* bits) {
  const char* base = (const char*)c.data;
  int64_t r = row + c.offset;
  switch (c.type_id) {
    case BSQL_INT8: *bits = ((const int8_t*)base)[r]; return true;
    case BSQL_INT16: *bits = ((const int16_t*)base)[r]; return true;
    case BSQL_INT32: *bits = ((const int32_t*)base)[r]; return true;
    case BSQL_INT64: *bits = ((const int64_t*)base)[r]; return true;
    case BSQL_UINT8: *bits = ((const uint8_t*)base)[r]; return true;
    case BSQL_UINT16: *bits = ((const uint16_t*)base)[r]; return true;
    case BSQL_UINT32: *bits = ((const uint32_t*)base)[r]; return true;
    case BSQL_UINT64: *bits = (int64_t)((const uint64_t*)base)[r]; return true;
    case BSQL_BOOL8: *bits = ((const uint8_t*)base)[r] != 0; return true;
    case BSQL_FLOAT32: *bits = bits_of((double)((const float*)base)[r]); return true;
    case BSQL_FLOAT64: *bits = bits_of(((const double*)base)[r]); return true;
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
      *bits = ((const int32_t*)base)[r]; return true;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_TIMESTAMP_MILLISECONDS:
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_TIMESTAMP_NANOSECONDS:
    case BSQL_DURATION_SECONDS: case BSQL_DURATION_MILLISECONDS:
    case BSQL_DURATION_MICROSECONDS: case BSQL_DURATION_NANOSECONDS:
      *bits = ((const int64_t*)base)[r]; return true;
  }
  return false;
}

// write_data (interpreter_ops.cuh:329-395): static_cast from the slot bits
void write_elem(bsql_column& c, int64_t row, int64_t bits) {
  char* base = (char*)c.data;
  int64_t r = row + c.offset;
  switch (c.type_id) {
    case BSQL_INT8: ((int8_t*)base)[r] = (int8_t)bits; break;
    case BSQL_INT16: ((int16_t*)base)[r] = (int16_t)bits; break;
    case BSQL_INT32: ((int32_t*)base)[r] = (int32_t)bits; break;
    case BSQL_INT64: ((int64_t*)base)[r] = bits; break;
    case BSQL_UINT8: ((uint8_t*)base)[r] = (uint8_t)bits; break;
    case BSQL_UINT16: ((uint16_t*)base)[r] = (uint16_t)bits; break;
    case BSQL_UINT32: ((uint32_t*)base)[r] = (uint32_t)bits; break;
    case BSQL_UINT64: ((uint64_t*)base)[r] = (uint64_t)bits; break;
    case BSQL_BOOL8: ((uint8_t*)base)[r] = bits != 0; break;
    case BSQL_FLOAT32: ((float*)base)[r] = (float)dbl_of(bits); break;
    case BSQL_FLOAT64: ((double*)base)[r] = dbl_of(bits); break;
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: ((int32_t*)base)[r] = (int32_t)bits; break;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_TIMESTAMP_MILLISECONDS:
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_TIMESTAMP_NANOSECONDS:
    case BSQL_DURATION_SECONDS: case BSQL_DURATION_MILLISECONDS:
    case BSQL_DURATION_MICROSECONDS: case BSQL_DURATION_NANOSECONDS:
      ((int64_t*)base)[r] = bits; break;
    default: break;
  }
}

// get_scalar_value<T> (interpreter_ops.cuh:428-477)
Val scalar_val(const bsql_scalar& s, bool want_f) {
  Val v{want_f, 0, 0.0};
  if (is_float_type(s.type_id)) {
    if (want_f) v.d = s.v.f64; else v.i = f64_to_i64(s.v.f64);
  } else {
    if (want_f) v.d = (double)s.v.i64; else v.i = s.v.i64;
  }
  return v;
}

// Philox4x32-10 counter-based generator: one uniform double in (0, 1] per
// (seed, row). RAND has no pinned stream in the reference (curand seeded
// from std::random_device, interpreter_cpp.cu:577-585).
double philox_uniform(uint64_t seed, uint64_t row) {
  uint32_t c0 = (uint32_t)row, c1 = (uint32_t)(row >> 32), c2 = 0, c3 = 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int i = 0; i < 10; ++i) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  uint64_t hi = ((uint64_t)c0 << 32 | c1) >> 11;  // 53 random bits
  return ((double)hi + 1.0) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------ row evaluation
struct RowCtx {
  int64_t buf[512];
  uint64_t valids[8];  // bit per slot (512 slots)
  std::string str[512];  // string results of SUBSTRING/LEFT/RIGHT
  void clear() {
    std::memset(buf, 0, sizeof(buf));
    std::memset(valids, 0, sizeof(valids));
  }
};
bool getv(const RowCtx& c, int s) { return (c.valids[s >> 6] >> (s & 63)) & 1u; }
void setv(RowCtx& c, int s, bool v) {
  if (v) c.valids[s >> 6] |= (uint64_t)1 << (s & 63);
  else c.valids[s >> 6] &= ~((uint64_t)1 << (s & 63));
}

struct Exec {
  const Prog* p;
  Table in;
  uint64_t seed;
  std::vector<std::shared_ptr<std::wregex>> like_re;  // per op: LIKE's regex
};

void prepare_like(Exec& ex) {
  ex.like_re.assign(ex.p->ops.size(), nullptr);
  for (size_t i = 0; i < ex.p->ops.size(); ++i) {
    if (ex.p->ops[i] != BSQL_OP_STR_LIKE || ex.p->right[i] != BSQL_SCALAR_INDEX) continue;
    const bsql_scalar& pat = ex.p->rsc[i];
    std::string like(pat.str, (size_t)pat.str_len);
    std::string re = like_to_regex(like);
    ex.like_re[i] = std::make_shared<std::wregex>(to_wide(utf8_decode(StrView{re.data(), (int32_t)re.size()})));
  }
}

// string operand: an input column, or a string an earlier op produced
// cudf::strings::strip (libcudf 21.08, LogicalProject.cpp:567-580): drops
// the leading (side bit 0) and trailing (bit 1) characters that are in
// `set`, or for an empty set those whose encoded value is <= ' '.
std::string strip_chars(StrView s, StrView set, int side) {
  const std::vector<std::string> chars = [&] {
    std::vector<std::string> v;
    for (int32_t i = 0; i < s.n;) {
      int32_t j = i + 1;
      while (j < s.n && (((uint8_t)s.p[j]) & 0xC0) == 0x80) ++j;
      v.emplace_back(s.p + i, (size_t)(j - i));
      i = j;
    }
    return v;
  }();
  std::vector<std::string> strip;
  for (int32_t i = 0; i < set.n;) {
    int32_t j = i + 1;
    while (j < set.n && (((uint8_t)set.p[j]) & 0xC0) == 0x80) ++j;
    strip.emplace_back(set.p + i, (size_t)(j - i));
    i = j;
  }
  auto is_strip = [&](const std::string& ch) {
    if (strip.empty()) {
      uint32_t v = 0;
      for (size_t k = 0; k < ch.size(); ++k) v |= (uint32_t)(uint8_t)ch[k] << (8 * k);
      return v <= (uint32_t)' ';
    }
    for (const std::string& x : strip)
      if (x == ch) return true;
    return false;
  };
  size_t lo = 0, hi = chars.size();
  if (side & 1) while (lo < hi && is_strip(chars[lo])) ++lo;
  if (side & 2) while (hi > lo && is_strip(chars[hi - 1])) --hi;
  std::string out;
  for (size_t k = lo; k < hi; ++k) out += chars[k];
  return out;
}

// cudf::strings::to_integers (libcudf 21.08): optional sign, then digits up
// to the first other character, in wrapping arithmetic.
int64_t parse_int(StrView s) {
  int32_t i = 0;
  uint64_t sign = 1;
  if (s.n > 0 && (s.p[0] == '-' || s.p[0] == '+')) { sign = s.p[0] == '-' ? ~0ULL : 1ULL; i = 1; }
  uint64_t v = 0;
  for (; i < s.n; ++i) {
    const char c = s.p[i];
    if (c < '0' || c > '9') break;
    v = v * 10ULL + (uint64_t)(c - '0');
  }
  return (int64_t)(v * sign);
}

// cudf::strings::to_timestamps (libcudf 21.08), which LogicalProject.cpp:469-533
// calls for CAST(varchar AS DATE/TIMESTAMP) ("%Y-%m-%d" / "%Y-%m-%d %H:%M:%S")
// and TO_DATE / TO_TIMESTAMP (the given format): the format is walked left to
// right; %Y takes the next 4 characters, %m %d %H %M %S the next 2, each read
// as the decimal digits before the first non-digit among them; %% and any
// other character skip one character of the string; parsing stops where the
// string ends, the fields not reached staying 1970-01-01 00:00:00. The date
// is counted in days from 1970-01-01 on the proleptic Gregorian calendar;
// sub-day types add the time of day, in their ticks.
int64_t parse_timestamp(StrView s, const std::string& format, int type) {
  long y = 1970, mo = 1, d = 1, h = 0, mi = 0, se = 0;
  size_t at = 0;
  const size_t n = (size_t)s.n;
  for (size_t f = 0; f < format.size() && at < n; ++f) {
    char spec = 0;
    if (format[f] == '%' && f + 1 < format.size() && format[f + 1] != '%') spec = format[++f];
    else if (format[f] == '%') ++f;  // "%%" is one literal character
    if (!spec) { ++at; continue; }
    const size_t width = spec == 'Y' ? 4 : 2;
    long v = 0;
    for (size_t k = at; k < std::min(n, at + width); ++k) {
      const char c = s.p[k];
      if (!std::isdigit((unsigned char)c)) break;
      v = v * 10 + (c - '0');
    }
    if (spec == 'Y') y = v;
    else if (spec == 'm') mo = v;
    else if (spec == 'd') d = v;
    else if (spec == 'H') h = v;
    else if (spec == 'M') mi = v;
    else se = v;
    at += width;
  }
  // days from civil: shift the year to start in March, count 400-year eras
  const long yy = mo <= 2 ? y - 1 : y;
  const long era = (yy >= 0 ? yy : yy - 399) / 400;
  const long yoe = yy - era * 400;
  const long mp = mo > 2 ? mo - 3 : mo + 9;
  const long doe = yoe * 365 + yoe / 4 - yoe / 100 + (153 * mp + 2) / 5 + d - 1;
  const int64_t days = (int64_t)(era * 146097 + doe - 719468);
  if (type == BSQL_TIMESTAMP_DAYS) return days;
  const uint64_t secs = (uint64_t)days * 86400u + (uint64_t)(h * 3600 + mi * 60 + se);
  const uint64_t per = type == BSQL_TIMESTAMP_MILLISECONDS ? 1000u
                       : type == BSQL_TIMESTAMP_MICROSECONDS ? 1000000u
                       : type == BSQL_TIMESTAMP_NANOSECONDS ? 1000000000u : 1u;
  return (int64_t)(secs * per);
}

// The double nearest 10^e, from the C library's correctly rounded strtod.
double pow10_nearest(int e) {
  static const std::vector<double> table = [] {
    std::vector<double> t;
    for (int k = -307; k <= 308; ++k) t.push_back(std::strtod(("1e" + std::to_string(k)).c_str(), nullptr));
    return t;
  }();
  return table[(size_t)(e + 307)];
}

// cudf::strings::to_floats (libcudf 21.08), which LogicalProject.cpp:439-468
// calls for CAST(varchar AS DOUBLE/FLOAT): the string exactly "NaN", "Inf" or
// "-Inf"; else an optional sign, then digits and '.' (any '.' puts the digits
// after it in the fraction), digits kept while the value stays within 52 bits
// (later integer digits only raise the exponent), then an optional exponent
// 'e'/'E' with a sign. A decimal exponent above 308 gives +-infinity, below
// -307 gives +0, otherwise sign * digits * exp10(e). exp10 is taken as the
// double nearest 10^e.
double parse_double(StrView s) {
  auto at = [&](int32_t k) { return (uint8_t)s.p[k]; };
  const std::string str(s.p, (size_t)s.n);
  if (s.n == 0) return 0.0;
  if (str == "NaN") return std::numeric_limits<double>::quiet_NaN();
  if (str == "Inf") return std::numeric_limits<double>::infinity();
  if (str == "-Inf") return -std::numeric_limits<double>::infinity();
  int32_t i = 0;
  double sign = 1.0;
  if (at(0) == '-' || at(0) == '+') { sign = at(0) == '-' ? -1.0 : 1.0; ++i; }
  const uint64_t limit = (1ULL << 52) - 1;
  uint64_t mant = 0;
  int64_t shift = 0;
  bool frac = false;
  for (; i < s.n; ++i) {
    const uint8_t c = at(i);
    if (c == '.') { frac = true; continue; }
    if (!std::isdigit(c)) break;
    if (mant > limit) { if (!frac) ++shift; continue; }
    mant = mant * 10 + (uint64_t)(c - '0');
    if (mant > limit) { mant /= 10; if (!frac) ++shift; }
    else if (frac) --shift;
  }
  int64_t ex = 0;
  if (i < s.n) {
    const uint8_t c = at(i++);
    if ((c == 'e' || c == 'E') && i < s.n) {
      bool neg = false;
      if (at(i) == '-' || at(i) == '+') neg = at(i++) == '-';
      uint32_t v = 0;  // the reference accumulates in a 32-bit int, wrapping
      while (i < s.n) {
        const uint8_t d = at(i++);
        if (!std::isdigit(d)) break;
        v = v * 10u + (uint32_t)(d - '0');
      }
      ex = (int32_t)v;
      if (neg) ex = -(int64_t)(int32_t)v;
    }
  }
  const int32_t e = (int32_t)(uint32_t)(ex + shift);
  if (e > 308) return sign > 0 ? std::numeric_limits<double>::infinity() : -std::numeric_limits<double>::infinity();
  if (e < -307) return 0.0;
  double base = sign * (double)mant;
  if (e) base *= pow10_nearest(e);
  return base;
}

StrView get_str(const Exec& ex, const RowCtx& c, int pos, int64_t row) {
  if (pos < ex.in.n) return col_str(ex.in.cols[pos], row);
  return StrView{c.str[pos].data(), (int32_t)c.str[pos].size()};
}

void store(RowCtx& c, int s, const Val& v) { c.buf[s] = v.is_f ? bits_of(v.d) : v.i; }
void store_i(RowCtx& c, int s, int64_t v) { c.buf[s] = v; }
void store_f(RowCtx& c, int s, double v) { c.buf[s] = bits_of(v); }

Val slot_val(const RowCtx& c, int s, bool is_f) {
  Val v{is_f, 0, 0.0};
  if (is_f) v.d = dbl_of(c.buf[s]); else v.i = c.buf[s];
  return v;
}

template <class F>
int64_t cmp_generic(const Val& l, const Val& r, F f) {
  if (!l.is_f && !r.is_f) return f(l.i, r.i) ? 1 : 0;
  return f(as_f(l), as_f(r)) ? 1 : 0;
}

int64_t compare(int op, int c3) {  // c3: -1, 0, 1 three-way
  switch (op) {
    case BSQL_OP_EQUAL: return c3 == 0;
    case BSQL_OP_NOT_EQUAL: return c3 != 0;
    case BSQL_OP_LESS: return c3 < 0;
    case BSQL_OP_GREATER: return c3 > 0;
    case BSQL_OP_LESS_EQUAL: return c3 <= 0;
    case BSQL_OP_GREATER_EQUAL: return c3 >= 0;
  }
  return 0;
}
template <class T>
int three_way(T a, T b) { return a < b ? -1 : (b < a ? 1 : (a == b ? 0 : 2)); }
// 2 encodes "unordered" (NaN): every comparison but != is false
int64_t compare_u(int op, int c3) {
  if (c3 == 2) return op == BSQL_OP_NOT_EQUAL;
  return compare(op, c3);
}

void eval_op(const Exec& ex, size_t i, int64_t row, RowCtx& c) {
  const Prog& p = *ex.p;
  const int op = p.ops[i];
  const int lpos = p.left[i], rpos = p.right[i], opos = p.out[i];
  const int lt = p.ltype[i], rt = p.rtype[i];
  const bool lf = is_float_type(lt), rf = is_float_type(rt);

  if (rpos != BSQL_UNARY_INDEX && rpos != BSQL_NULLARY_INDEX) {
    // ---------------- binary (interpreter_ops.cuh:508-727)
    Val L{lf, 0, 0.0}, R{rf, 0, 0.0};
    StrView ls{nullptr, 0}, rs{nullptr, 0};
    bool lv, rv;
    if (lpos >= 0) {
      if (is_string_type(lt)) ls = get_str(ex, c, lpos, row); else L = slot_val(c, lpos, lf);
      lv = getv(c, lpos);
    } else if (lpos == BSQL_SCALAR_INDEX) {
      if (is_string_type(lt)) ls = StrView{p.lsc[i].str, p.lsc[i].str_len}; else L = scalar_val(p.lsc[i], lf);
      lv = true;
    } else {
      lv = false;
    }
    if (rpos >= 0) {
      if (is_string_type(rt)) rs = get_str(ex, c, rpos, row); else R = slot_val(c, rpos, rf);
      rv = getv(c, rpos);
    } else if (rpos == BSQL_SCALAR_INDEX) {
      if (is_string_type(rt)) rs = StrView{p.rsc[i].str, p.rsc[i].str_len}; else R = scalar_val(p.rsc[i], rf);
      rv = true;
    } else {
      rv = false;
    }

    if (op == BSQL_OP_MAGIC_IF_NOT) {
      if (lv && truthy(L)) { store(c, opos, R); setv(c, opos, rv); }
      else {
        if (rf) store_f(c, opos, MAGIC_F64); else store_i(c, opos, MAGIC_I64);
        setv(c, opos, false);  // reference leaves the bit stale; defined as null here
      }
    } else if (op == BSQL_OP_FIRST_NON_MAGIC) {
      bool is_magic = lf ? (L.d == MAGIC_F64) : (L.i == MAGIC_I64);
      if (is_magic) { store(c, opos, R); setv(c, opos, rv); }
      else { store(c, opos, L); setv(c, opos, lv); }
    } else if (op == BSQL_OP_LOGICAL_OR) {
      if (lv && rv) { store_i(c, opos, (truthy(L) || truthy(R)) ? 1 : 0); setv(c, opos, true); }
      else if (lv) { store(c, opos, L); setv(c, opos, truthy(L)); }
      else if (rv) { store(c, opos, R); setv(c, opos, true); }
      else setv(c, opos, false);
    } else if (op == BSQL_OP_STR_SUBSTRING || op == BSQL_OP_STR_LEFT || op == BSQL_OP_STR_RIGHT) {
      // positions as LogicalProject.cpp:183-306 prepares them for
      // slice_strings: a literal SUBSTRING start is max(s, 1) - 1, a computed
      // one s - 1; a null position reads as 0; LEFT/RIGHT take max(n, 0)
      const int64_t v = rv ? (rf ? f64_to_i64(R.d) : R.i) : 0;
      std::string res;
      if (lv) {
        if (op == BSQL_OP_STR_SUBSTRING) {
          const int64_t start = rpos == BSQL_SCALAR_INDEX ? std::max<int64_t>(v, 1) - 1 : (rv ? v - 1 : 0);
          res = slice_chars(ls, start, -1);
        } else if (op == BSQL_OP_STR_LEFT) {
          res = slice_chars(ls, 0, std::max<int64_t>(v, 0));
        } else {
          const int64_t n = std::max<int64_t>(v, 0);
          const int64_t len = (int64_t)utf8_decode(ls).size();
          res = n == 0 ? std::string() : slice_chars(ls, std::max<int64_t>(len - n, 0), -1);
        }
      }
      c.str[opos] = res;
      setv(c, opos, lv);
    } else if (op == BSQL_OP_STR_TRIM) {
      c.str[opos] = lv ? strip_chars(ls, rs, (int)p.rsc[i].v.i64) : std::string();
      setv(c, opos, lv);
    } else if (op == BSQL_OP_TO_DATE || op == BSQL_OP_TO_TIMESTAMP) {
      // to_timestamps(column, type, format), LogicalProject.cpp:510-533
      const std::string fmt(rs.p, (size_t)rs.n);
      store_i(c, opos, lv ? parse_timestamp(ls, fmt, op == BSQL_OP_TO_DATE ? BSQL_TIMESTAMP_DAYS
                                                                           : BSQL_TIMESTAMP_NANOSECONDS) : 0);
      setv(c, opos, lv);
    } else if (op == BSQL_OP_IS_NOT_DISTINCT_FROM) {
      bool res;
      if (!lv && !rv) res = true;
      else if (!lv || !rv) res = false;
      else res = cmp_generic(L, R, [](auto a, auto b) { return a == b; }) != 0;
      store_i(c, opos, res ? 1 : 0);
      setv(c, opos, true);
    } else if (lv && rv) {
      bool valid = true;
      switch (op) {
        case BSQL_OP_LOGICAL_AND: store_i(c, opos, (truthy(L) && truthy(R)) ? 1 : 0); break;
        case BSQL_OP_STR_LIKE: {
          const std::wstring w = to_wide(utf8_decode(ls));
          store_i(c, opos, std::regex_search(w, *ex.like_re[i]) ? 1 : 0);
          break;
        }
        case BSQL_OP_ADD:
          if (!lf && !rf) store_i(c, opos, wrap_add(L.i, R.i)); else store_f(c, opos, as_f(L) + as_f(R));
          break;
        case BSQL_OP_SUB:
          if (!lf && !rf) store_i(c, opos, wrap_sub(L.i, R.i)); else store_f(c, opos, as_f(L) - as_f(R));
          break;
        case BSQL_OP_MUL:
          if (!lf && !rf) store_i(c, opos, wrap_mul(L.i, R.i)); else store_f(c, opos, as_f(L) * as_f(R));
          break;
        case BSQL_OP_DIV:
          if (!lf && !rf) store_i(c, opos, idiv(L.i, R.i)); else store_f(c, opos, as_f(L) / as_f(R));
          valid = rf ? !(R.d == 0.0) : (R.i != 0);  // interpreter_ops.cuh:721-724
          break;
        case BSQL_OP_MOD:
          if (!lf && !rf) store_i(c, opos, imod(L.i, R.i)); else store_f(c, opos, std::fmod(as_f(L), as_f(R)));
          break;
        case BSQL_OP_POW: store_f(c, opos, std::pow(as_f(L), as_f(R))); break;
        case BSQL_OP_ROUND: {
          double factor = std::pow(10.0, as_f(R));
          store_f(c, opos, std::round(as_f(L) * factor) / factor);
          break;
        }
        case BSQL_OP_EQUAL: case BSQL_OP_NOT_EQUAL: case BSQL_OP_LESS:
        case BSQL_OP_GREATER: case BSQL_OP_LESS_EQUAL: case BSQL_OP_GREATER_EQUAL: {
          int64_t res;
          if (is_string_type(lt) && is_string_type(rt)) {
            res = compare(op, str_cmp(ls, rs));
          } else if (is_timestamp_type(lt) && is_timestamp_type(rt)) {
            int64_t a = wrap_mul(L.i, ns_per_tick(lt)), b = wrap_mul(R.i, ns_per_tick(rt));
            res = compare(op, three_way(a, b));
          } else if (!lf && !rf) {
            res = compare(op, three_way(L.i, R.i));
          } else {
            res = compare_u(op, three_way(as_f(L), as_f(R)));
          }
          store_i(c, opos, res);
          break;
        }
        default: break;  // BITWISE_* and string ops never reach the VM
      }
      setv(c, opos, valid);
    } else {
      setv(c, opos, false);
    }
  } else if (rpos != BSQL_NULLARY_INDEX) {
    // ---------------- unary (interpreter_ops.cuh:728-917)
    Val L{lf, 0, 0.0};
    StrView ls{nullptr, 0};
    if (is_string_type(lt)) ls = get_str(ex, c, lpos, row); else L = slot_val(c, lpos, lf);
    bool lv = getv(c, lpos);
    if (op == BSQL_OP_IS_NOT_TRUE) {
      bool v = (lv && (lf ? L.d == 1.0 : L.i == 1)) ? false : true;
      lv = true; store_i(c, opos, v);
    } else if (op == BSQL_OP_IS_NOT_FALSE) {
      bool v = (lv && (lf ? L.d == 0.0 : L.i == 0)) ? false : true;
      lv = true; store_i(c, opos, v);
    } else if (op == BSQL_OP_IS_NULL) {
      store_i(c, opos, !lv);
    } else if (op == BSQL_OP_IS_NOT_NULL) {
      store_i(c, opos, lv);
    } else if (lv) {
      const double x = as_f(L);
      switch (op) {
        case BSQL_OP_FLOOR: store_f(c, opos, std::floor(x)); break;
        case BSQL_OP_CEIL: store_f(c, opos, std::ceil(x)); break;
        case BSQL_OP_SIN: store_f(c, opos, std::sin(x)); break;
        case BSQL_OP_COS: store_f(c, opos, std::cos(x)); break;
        case BSQL_OP_ASIN: store_f(c, opos, std::asin(x)); break;
        case BSQL_OP_ACOS: store_f(c, opos, std::acos(x)); break;
        case BSQL_OP_TAN: store_f(c, opos, std::tan(x)); break;
        case BSQL_OP_COTAN: store_f(c, opos, std::cos(x) / std::sin(x)); break;
        case BSQL_OP_ATAN: store_f(c, opos, std::atan(x)); break;
        case BSQL_OP_ABS:
          if (lf) store_f(c, opos, std::fabs(L.d));
          else store_i(c, opos, L.i < 0 ? wrap_sub(0, L.i) : L.i);
          break;
        case BSQL_OP_NOT: store_i(c, opos, !truthy(L)); break;
        case BSQL_OP_IS_TRUE: store_i(c, opos, lf ? f64_to_i64(L.d) : L.i); break;
        case BSQL_OP_LN: store_f(c, opos, std::log(x)); break;
        case BSQL_OP_LOG: store_f(c, opos, std::log10(x)); break;
        case BSQL_OP_YEAR: case BSQL_OP_MONTH: case BSQL_OP_DAY: case BSQL_OP_DAYOFWEEK:
        case BSQL_OP_HOUR: case BSQL_OP_MINUTE: case BSQL_OP_SECOND:
          store_i(c, opos, is_timestamp_type(lt) ? extract(op, lt, L.i) : 0);
          break;
        case BSQL_OP_CAST_TINYINT: case BSQL_OP_CAST_SMALLINT:
        case BSQL_OP_CAST_INTEGER: case BSQL_OP_CAST_BIGINT:
          if (is_string_type(lt)) store_i(c, opos, parse_int(ls));  // LogicalProject.cpp:379-438
          else store_i(c, opos, lf ? f64_to_i64(L.d) : L.i);
          break;
        case BSQL_OP_CAST_FLOAT: case BSQL_OP_CAST_DOUBLE:
          store_f(c, opos, is_string_type(lt) ? parse_double(ls) : x);  // :439-468
          break;
        case BSQL_OP_CAST_DATE: {
          if (is_string_type(lt)) {  // LogicalProject.cpp:469-482
            store_i(c, opos, parse_timestamp(ls, "%Y-%m-%d", BSQL_TIMESTAMP_DAYS));
            break;
          }
          int64_t v = lf ? f64_to_i64(L.d) : L.i;
          if (is_timestamp_type(lt) && lt != BSQL_TIMESTAMP_DAYS) v = v / ticks_per_day(lt);  // truncation
          store_i(c, opos, v);
          break;
        }
        case BSQL_OP_CAST_TIMESTAMP: case BSQL_OP_CAST_TIMESTAMP_SECONDS:
        case BSQL_OP_CAST_TIMESTAMP_MILLISECONDS: case BSQL_OP_CAST_TIMESTAMP_MICROSECONDS: {
          if (is_string_type(lt)) {  // :484-508
            const int to = op == BSQL_OP_CAST_TIMESTAMP_SECONDS ? BSQL_TIMESTAMP_SECONDS
                           : op == BSQL_OP_CAST_TIMESTAMP_MILLISECONDS ? BSQL_TIMESTAMP_MILLISECONDS
                           : op == BSQL_OP_CAST_TIMESTAMP_MICROSECONDS ? BSQL_TIMESTAMP_MICROSECONDS
                                                                       : BSQL_TIMESTAMP_NANOSECONDS;
            store_i(c, opos, parse_timestamp(ls, "%Y-%m-%d %H:%M:%S", to));
            break;
          }
          int64_t v = lf ? f64_to_i64(L.d) : L.i;
          if (is_timestamp_type(lt)) v = wrap_mul(v, ns_per_tick(lt));
          store_i(c, opos, v);
          break;
        }
        case BSQL_OP_CHAR_LENGTH: store_i(c, opos, utf8_length(ls)); break;
        default: break;
      }
    }
    bool ov = (op == BSQL_OP_IS_NULL || op == BSQL_OP_IS_NOT_NULL) ? true : lv;
    setv(c, opos, ov);
  } else {
    // ---------------- nullary (interpreter_ops.cuh:918-924)
    if (op == BSQL_OP_RAND) store_f(c, opos, philox_uniform(ex.seed, (uint64_t)row));
    setv(c, opos, true);
  }
}

void eval_row(const Exec& ex, int64_t row, RowCtx& c) {
  for (int k = 0; k < ex.in.n; ++k) {
    const bsql_column& col = ex.in.cols[k];
    setv(c, k, col_valid(col, row));
    int64_t bits;
    if (read_elem(col, row, &bits)) c.buf[k] = bits;
  }
  for (size_t i = 0; i < ex.p->ops.size(); ++i) eval_op(ex, i, row, c);
}

int nthreads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class F>
void parallel_chunks(int64_t n, int threads, F f) {
  // chunks aligned to 32 rows so that validity words never straddle threads
  threads = nthreads(threads);
  int64_t words = (n + 31) / 32;
  int64_t per = (words + threads - 1) / threads;
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    int64_t b = (int64_t)t * per * 32, e = std::min(n, (int64_t)(t + 1) * per * 32);
    if (b >= e) break;
    ts.emplace_back([=] { f(t, b, e); });
  }
  for (auto& th : ts) th.join();
}

bool load_prog(const bsql_scalar* lsc, const bsql_scalar* rsc, const int16_t* left,
               const int16_t* right, const int16_t* out, const int32_t* ops, int n_ops,
               const int16_t* fo, int n_final, Prog* p) {
  p->left.assign(left, left + n_ops);
  p->right.assign(right, right + n_ops);
  p->out.assign(out, out + n_ops);
  p->ops.assign(ops, ops + n_ops);
  p->final_out.assign(fo, fo + n_final);
  bsql_scalar zero{};
  p->lsc.assign(n_ops, zero);
  p->rsc.assign(n_ops, zero);
  for (int i = 0; i < n_ops; ++i) {
    if (left[i] == BSQL_SCALAR_INDEX) { if (!lsc) return false; p->lsc[i] = lsc[i]; }
    if (right[i] == BSQL_SCALAR_INDEX) { if (!rsc) return false; p->rsc[i] = rsc[i]; }
    if (left[i] >= 512 || right[i] >= 512 || out[i] >= 512) return false;
  }
  return true;
}

// String outputs (string-valued expression roots, as evaluate_expressions
// returns them) are assembled after the row pass: offsets then chars, in row
// order; a null row has length 0. Returns false if a chars buffer is short.
bool assemble_strings(bsql_column& out, const std::vector<std::string>& vals, int64_t num_rows) {
  int64_t pos = 0;
  out.offsets[0] = 0;
  for (int64_t r = 0; r < num_rows; ++r) {
    const std::string& v = vals[(size_t)r];
    if (pos + (int64_t)v.size() > out.chars_size) return false;
    if (!v.empty()) std::memcpy((char*)out.data + pos, v.data(), v.size());
    pos += (int64_t)v.size();
    out.offsets[r + 1] = (int32_t)pos;
  }
  return true;
}

bool execute_rows(const Prog& p, const bsql_column* in, int n_in, bsql_column* out, int n_out,
                  int64_t num_rows, int threads, uint64_t seed) {
  Exec ex{&p, Table{in, n_in}, seed, {}};
  prepare_like(ex);
  std::vector<std::vector<std::string>> strs((size_t)n_out);
  for (int k = 0; k < n_out; ++k)
    if (out[k].type_id == BSQL_STRING) strs[(size_t)k].assign((size_t)num_rows, std::string());
  parallel_chunks(num_rows, threads, [&](int, int64_t b, int64_t e) {
    RowCtx c;
    for (int64_t w = b; w < e; w += 32) {
      c.clear();
      uint32_t vw[64] = {0};
      int64_t we = std::min(e, w + 32);
      for (int64_t row = w; row < we; ++row) {
        eval_row(ex, row, c);
        for (int k = 0; k < n_out; ++k) {
          int fs = p.final_out[k];
          const bool v = getv(c, fs);
          if (out[k].type_id == BSQL_STRING) {
            if (v) {
              StrView sv = get_str(ex, c, fs, row);
              strs[(size_t)k][(size_t)row].assign(sv.p, (size_t)sv.n);
            }
          } else {
            write_elem(out[k], row, c.buf[fs]);
          }
          if (v) vw[k] |= 1u << (row - w);
        }
      }
      for (int k = 0; k < n_out; ++k)
        if (out[k].validity) out[k].validity[w >> 5] = vw[k];
    }
  });
  for (int k = 0; k < n_out; ++k)
    if (out[k].type_id == BSQL_STRING && !assemble_strings(out[k], strs[(size_t)k], num_rows)) return false;
  return true;
}

void gather_column(const bsql_column& src, bsql_column& dst, const std::vector<int64_t>& rows) {
  const int64_t n = (int64_t)rows.size();
  if (src.type_id == BSQL_STRING) {
    int32_t pos = 0;
    dst.offsets[0] = 0;
    for (int64_t j = 0; j < n; ++j) {
      StrView s = col_str(src, rows[j]);
      std::memcpy((char*)dst.data + pos, s.p, (size_t)s.n);
      pos += s.n;
      dst.offsets[j + 1] = pos;
    }
  } else {
    const int w = type_size(src.type_id);
    for (int64_t j = 0; j < n; ++j)
      std::memcpy((char*)dst.data + j * w, (const char*)src.data + (rows[j] + src.offset) * w, (size_t)w);
  }
  if (dst.validity) {
    const int64_t words = (n + 31) / 32;
    for (int64_t wi = 0; wi < words; ++wi) {
      uint32_t word = 0;
      for (int64_t j = wi * 32; j < std::min(n, wi * 32 + 32); ++j)
        if (col_valid(src, rows[j])) word |= 1u << (j - wi * 32);
      dst.validity[wi] = word;
    }
  }
}

std::vector<int64_t> select_rows(const Prog& pred, const bsql_column* in, int n_in,
                                 int64_t num_rows, int threads, uint64_t seed) {
  std::vector<uint8_t> mask((size_t)num_rows);
  Exec ex{&pred, Table{in, n_in}, seed, {}};
  prepare_like(ex);
  const int fs = pred.final_out[0];
  parallel_chunks(num_rows, threads, [&](int, int64_t b, int64_t e) {
    RowCtx c;
    c.clear();
    for (int64_t row = b; row < e; ++row) {
      eval_row(ex, row, c);
      // the BOOL8 output is static_cast<bool>(slot) and a null drops the row
      mask[(size_t)row] = getv(c, fs) && c.buf[fs] != 0;
    }
  });
  std::vector<int64_t> rows;
  for (int64_t r = 0; r < num_rows; ++r)
    if (mask[(size_t)r]) rows.push_back(r);
  return rows;
}

}  // namespace

extern "C" {

typedef struct oracle_program {
  const int16_t* left;
  const int16_t* right;
  const int16_t* out;
  const int32_t* ops;
  int32_t n_ops;
  int32_t n_final;
  const int16_t* final_out;
  const bsql_scalar* left_scalars;
  const bsql_scalar* right_scalars;
} oracle_program;

static bool build(const oracle_program* op, const bsql_column* in, int n_in, Prog* p) {
  if (!load_prog(op->left_scalars, op->right_scalars, op->left, op->right, op->out, op->ops,
                 op->n_ops, op->final_out, op->n_final, p))
    return false;
  std::vector<int32_t> types(n_in);
  for (int k = 0; k < n_in; ++k) types[k] = in[k].type_id;
  propagate_types(*p, types.data(), n_in);
  return true;
}

// perform_interpreter_operation (interpreter_cpp.cu:440-594) on host columns.
int oracle_execute(const oracle_program* prog, const bsql_column* in, int n_in,
                   bsql_column* out, int n_out, int64_t num_rows, int threads, uint64_t seed) {
  Prog p;
  if (!build(prog, in, n_in, &p)) return BSQL_ERR_INVALID;
  if ((int)p.final_out.size() != n_out) return BSQL_ERR_INVALID;
  if (!execute_rows(p, in, n_in, out, n_out, num_rows, threads, seed)) return BSQL_ERR_INVALID;
  return BSQL_OK;
}

// Output type of each op, as perform_interpreter_operation propagates it.
int oracle_op_types(const oracle_program* prog, const int32_t* in_types, int n_in,
                    int32_t* ltypes, int32_t* rtypes, int32_t* otypes) {
  Prog p;
  if (!load_prog(prog->left_scalars, prog->right_scalars, prog->left, prog->right, prog->out,
                 prog->ops, prog->n_ops, prog->final_out, prog->n_final, &p))
    return BSQL_ERR_INVALID;
  propagate_types(p, in_types, n_in);
  for (int i = 0; i < prog->n_ops; ++i) {
    if (ltypes) ltypes[i] = p.ltype[i];
    if (rtypes) rtypes[i] = p.rtype[i];
    if (otypes) otypes[i] = p.otype[i];
  }
  return BSQL_OK;
}

// process_filter + apply_boolean_mask (LogicalFilter.cpp:22-52): every input
// column is gathered into out[k] (strings included).
int oracle_filter(const oracle_program* pred, const bsql_column* in, int n_in, bsql_column* out,
                  int64_t num_rows, int64_t* count, int threads, uint64_t seed) {
  Prog p;
  if (!build(pred, in, n_in, &p) || p.final_out.size() != 1) return BSQL_ERR_INVALID;
  std::vector<int64_t> rows = select_rows(p, in, n_in, num_rows, threads, seed);
  *count = (int64_t)rows.size();
  for (int k = 0; k < n_in; ++k) gather_column(in[k], out[k], rows);
  return BSQL_OK;
}

// Filter then project over the filtered table: the reference's Filter kernel
// followed by its Projection kernel (BatchProcessing.cpp:454-539, 365-448).
int oracle_filter_project(const oracle_program* pred, const oracle_program* proj,
                          const bsql_column* in, int n_in, bsql_column* out, int n_out,
                          int64_t num_rows, int64_t* count, int threads, uint64_t seed) {
  Prog pp, pj;
  if (!build(pred, in, n_in, &pp) || pp.final_out.size() != 1) return BSQL_ERR_INVALID;
  if (!build(proj, in, n_in, &pj) || (int)pj.final_out.size() != n_out) return BSQL_ERR_INVALID;
  std::vector<int64_t> rows = select_rows(pp, in, n_in, num_rows, threads, seed);
  const int64_t n = (int64_t)rows.size();
  *count = n;
  // materialize the filtered input table, as the Filter kernel's output cache does
  std::vector<std::vector<char>> data(n_in), chars(n_in);
  std::vector<std::vector<uint32_t>> valid(n_in);
  std::vector<std::vector<int32_t>> offs(n_in);
  std::vector<bsql_column> fin(n_in);
  for (int k = 0; k < n_in; ++k) {
    fin[k] = in[k];
    fin[k].offset = 0;
    fin[k].size = n;
    if (in[k].type_id == BSQL_STRING) {
      int64_t total = 0;
      for (int64_t r : rows) total += col_str(in[k], r).n;
      chars[k].resize((size_t)total + 1);
      offs[k].resize((size_t)n + 1);
      fin[k].data = chars[k].data();
      fin[k].offsets = offs[k].data();
      fin[k].chars_size = (int32_t)total;
    } else {
      data[k].resize((size_t)(n * type_size(in[k].type_id)) + 8);
      fin[k].data = data[k].data();
    }
    if (in[k].validity) {
      valid[k].resize((size_t)((n + 31) / 32) + 1);
      fin[k].validity = valid[k].data();
    }
    gather_column(in[k], fin[k], rows);
  }
  if (!execute_rows(pj, fin.data(), n_in, out, n_out, n, threads, seed)) return BSQL_ERR_INVALID;
  return BSQL_OK;
}

int oracle_version(void) { return 1; }

}  // extern "C"
