// program.cpp -- type rules, validation and lowering of interpreter programs
// to the gfx950 device plan.
#include "program.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <map>
#include <tuple>
#include <set>

namespace bsql {

bool is_float_type(int t) { return t == BSQL_FLOAT32 || t == BSQL_FLOAT64; }
bool is_timestamp_type(int t) { return t >= BSQL_TIMESTAMP_DAYS && t <= BSQL_TIMESTAMP_NANOSECONDS; }
bool is_duration_type(int t) { return t >= BSQL_DURATION_DAYS && t <= BSQL_DURATION_NANOSECONDS; }
bool is_string_type(int t) { return t == BSQL_STRING; }

int type_size(int t) {
  switch (t) {
    case BSQL_INT8: case BSQL_UINT8: case BSQL_BOOL8: return 1;
    case BSQL_INT16: case BSQL_UINT16: return 2;
    case BSQL_INT32: case BSQL_UINT32: case BSQL_FLOAT32: case BSQL_TIMESTAMP_DAYS:
    case BSQL_DURATION_DAYS: case BSQL_DECIMAL32: case BSQL_DICTIONARY32: return 4;
    case BSQL_INT64: case BSQL_UINT64: case BSQL_FLOAT64: case BSQL_TIMESTAMP_SECONDS:
    case BSQL_TIMESTAMP_MILLISECONDS: case BSQL_TIMESTAMP_MICROSECONDS:
    case BSQL_TIMESTAMP_NANOSECONDS: case BSQL_DURATION_SECONDS:
    case BSQL_DURATION_MILLISECONDS: case BSQL_DURATION_MICROSECONDS:
    case BSQL_DURATION_NANOSECONDS: case BSQL_DECIMAL64: return 8;
    default: return 0;
  }
}

int64_t ticks_per_day(int t) {
  switch (t) {
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return 1;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_DURATION_SECONDS: return 86400LL;
    case BSQL_TIMESTAMP_MILLISECONDS: case BSQL_DURATION_MILLISECONDS: return 86400000LL;
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_DURATION_MICROSECONDS: return 86400000000LL;
    default: return 86400000000000LL;
  }
}

int64_t ns_per_tick(int t) { return 86400000000000LL / ticks_per_day(t); }

bool is_nullary_op(int op) { return op == BSQL_OP_RAND; }
bool is_unary_op(int op) { return op >= BSQL_OP_NOT && op <= BSQL_OP_STR_REVERSE; }
bool is_binary_op(int op) { return op >= BSQL_OP_ADD && op <= BSQL_OP_IS_NOT_DISTINCT_FROM; }

int output_type_nullary(int op) { return op == BSQL_OP_RAND ? BSQL_FLOAT64 : BSQL_EMPTY; }

int output_type_unary(int op, int lt) {
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
    case BSQL_OP_SIN: case BSQL_OP_COS: case BSQL_OP_ASIN: case BSQL_OP_ACOS: case BSQL_OP_TAN:
    case BSQL_OP_COTAN: case BSQL_OP_ATAN: case BSQL_OP_LN: case BSQL_OP_LOG: case BSQL_OP_FLOOR:
    case BSQL_OP_CEIL:
      return is_float_type(lt) ? lt : BSQL_FLOAT64;
    case BSQL_OP_ABS: return lt;
    case BSQL_OP_NOT: case BSQL_OP_IS_TRUE: case BSQL_OP_IS_NULL: case BSQL_OP_IS_NOT_NULL:
    case BSQL_OP_IS_NOT_TRUE: case BSQL_OP_IS_NOT_FALSE: return BSQL_BOOL8;
    case BSQL_OP_CHAR_LENGTH: return BSQL_INT32;
    default: return BSQL_EMPTY;
  }
}

int output_type_binary(int op, int lt, int rt) {
  if (lt == BSQL_EMPTY && rt == BSQL_EMPTY)
    fail("In get_output_type function: both operands types are empty");
  if (lt == BSQL_EMPTY) lt = rt;
  else if (rt == BSQL_EMPTY) rt = lt;
  switch (op) {
    case BSQL_OP_ADD: case BSQL_OP_SUB: case BSQL_OP_MUL: case BSQL_OP_DIV: case BSQL_OP_MOD:
      if (is_float_type(lt) && is_float_type(rt)) return type_size(lt) >= type_size(rt) ? lt : rt;
      if (is_float_type(lt)) return lt;
      if (is_float_type(rt)) return rt;
      return type_size(lt) >= type_size(rt) ? lt : rt;
    case BSQL_OP_EQUAL: case BSQL_OP_NOT_EQUAL: case BSQL_OP_LESS: case BSQL_OP_GREATER:
    case BSQL_OP_LESS_EQUAL: case BSQL_OP_GREATER_EQUAL: case BSQL_OP_LOGICAL_AND:
    case BSQL_OP_LOGICAL_OR: case BSQL_OP_STR_LIKE:
      return BSQL_BOOL8;
    case BSQL_OP_POW: case BSQL_OP_ROUND: return type_size(lt) <= 4 ? BSQL_FLOAT32 : BSQL_FLOAT64;
    case BSQL_OP_MAGIC_IF_NOT: return rt;
    case BSQL_OP_FIRST_NON_MAGIC: return type_size(lt) >= type_size(rt) ? lt : rt;
    case BSQL_OP_STR_SUBSTRING: case BSQL_OP_STR_REPLACE: case BSQL_OP_STR_REGEXP_REPLACE:
    case BSQL_OP_STR_CONCAT: case BSQL_OP_STR_TRIM: case BSQL_OP_STR_LEFT: case BSQL_OP_STR_RIGHT:
      return BSQL_STRING;
    case BSQL_OP_TO_DATE: return BSQL_TIMESTAMP_DAYS;
    case BSQL_OP_TO_TIMESTAMP: return BSQL_TIMESTAMP_NANOSECONDS;
    default:
      // IS_NOT_DISTINCT_FROM is listed only in the unary rule set of the
      // reference, so the binary rule falls through to EMPTY as it does there.
      return BSQL_EMPTY;
  }
}

OpTypes propagate_types(const Program& p, const int32_t* in_types, int n_in) {
  const size_t n = p.ops.size();
  OpTypes t;
  t.lt.assign(n, BSQL_EMPTY);
  t.rt.assign(n, BSQL_EMPTY);
  t.ot.assign(n, BSQL_EMPTY);
  std::map<int, int32_t> slot_type;  // output_map_type: missing slots read as EMPTY
  auto type_of = [&](int s) -> int32_t {
    auto it = slot_type.find(s);
    return it == slot_type.end() ? BSQL_EMPTY : it->second;
  };
  for (size_t i = 0; i < n; ++i) {
    const int l = p.left[i], r = p.right[i];
    if (l >= 0 && l < n_in) t.lt[i] = in_types[l];
    else if (l == BSQL_SCALAR_INDEX) t.lt[i] = p.lsc[i].type_id;
    else if (l >= 0) t.lt[i] = type_of(l);
    if (r >= 0 && r < n_in) t.rt[i] = in_types[r];
    else if (r == BSQL_SCALAR_INDEX) t.rt[i] = p.rsc[i].type_id;
    else if (r >= 0) t.rt[i] = type_of(r);
    if (r == BSQL_UNARY_INDEX) t.ot[i] = output_type_unary(p.ops[i], t.lt[i]);
    else if (r == BSQL_NULLARY_INDEX) t.ot[i] = output_type_nullary(p.ops[i]);
    else t.ot[i] = output_type_binary(p.ops[i], t.lt[i], t.rt[i]);
    slot_type[p.out[i]] = t.ot[i];
  }
  return t;
}

void validate(const Program& p, int n_in) {
  const size_t n = p.ops.size();
  if (p.left.size() != n || p.right.size() != n || p.out.size() != n || p.lsc.size() != n ||
      p.rsc.size() != n)
    fail("program arrays have different lengths");
  if (p.final_out.empty() && p.roots.empty() && p.parts.empty()) fail("program has no outputs");
  int mx = -1;
  for (size_t i = 0; i < n; ++i) {
    mx = std::max<int>(mx, std::max<int>(p.left[i], std::max<int>(p.right[i], p.out[i])));
    const int op = p.ops[i];
    if (op <= BSQL_OP_INVALID || op >= BSQL_OP_COUNT) fail("invalid operator " + std::to_string(op));
    const int l = p.left[i], r = p.right[i];
    if (p.out[i] < 0) fail("op " + std::to_string(i) + " has no output slot");
    if (r == BSQL_NULLARY_INDEX) {
      if (!is_nullary_op(op)) fail("operator " + std::to_string(op) + " is not nullary");
    } else if (r == BSQL_UNARY_INDEX) {
      if (!is_unary_op(op)) fail("operator " + std::to_string(op) + " is not unary");
      if (l < 0) fail("Unary operations on literals is not supported");
    } else {
      if (!is_binary_op(op)) fail("operator " + std::to_string(op) + " is not binary");
      if (l < 0 && r < 0) fail("Operations between literals is not supported");
      if (op == BSQL_OP_STR_LIKE && r == BSQL_SCALAR_INDEX && p.rsc[i].type_id == BSQL_STRING)
        (void)like_compile(p.rsc[i].str);  // rejects patterns the matcher does not take
      if ((op == BSQL_OP_TO_DATE || op == BSQL_OP_TO_TIMESTAMP) && r == BSQL_SCALAR_INDEX &&
          p.rsc[i].type_id == BSQL_STRING)
        (void)timestamp_format_compile(p.rsc[i].str);
      if (l < BSQL_SCALAR_NULL_INDEX || r < BSQL_SCALAR_NULL_INDEX) fail("invalid operand index");
    }
  }
  for (size_t k = 0; k < p.final_out.size(); ++k) {
    const int16_t f = p.final_out[k];
    if (f < 0) fail("output " + std::to_string(k) + " has no slot");
    mx = std::max<int>(mx, f);
  }
  if (mx > BSQL_MAX_SLOT)
    fail("Interops does not support plans with an input or output index greater than 63");
  (void)n_in;
}

// One regex_replace pass of like_expression_to_regex_str
// (LogicalProject.cpp:39-43): matches of ([^\\]?|\\{2})W are found left to
// right and their W becomes a wildcard (here the marker `mark`). A character
// other than a backslash right before W is consumed with it, so in "WW" the
// first W stays a literal character, as it does in the reference.
std::string like_pass(const std::string& in, char w, char mark) {
  std::string out;
  size_t i = 0;
  while (i < in.size()) {
    if (in[i] != '\\' && i + 1 < in.size() && in[i + 1] == w) {
      out += in[i];
      out += mark;
      i += 2;
    } else if (in[i] == w) {
      out += mark;
      i += 1;
    } else {
      out += in[i];
      i += 1;
    }
  }
  return out;
}

// The format language of cudf::strings::to_timestamps (libcudf 21.08) that
// the reference's casts and TO_DATE / TO_TIMESTAMP use
// (LogicalProject.cpp:469-533): %Y (4 digits), %m, %d, %H, %M, %S (2 each)
// read as fixed-width fields, %% a literal '%', any other character a
// literal the parser steps over; other specifiers are rejected.
std::vector<int32_t> timestamp_format_compile(const std::string& f) {
  std::vector<int32_t> items;
  int32_t lit = 0;
  auto flush = [&] {
    while (lit > 0) {
      const int32_t w = lit > 255 ? 255 : lit;
      items.push_back(w);
      lit -= w;
    }
  };
  for (size_t i = 0; i < f.size(); ++i) {
    if (f[i] != '%') { ++lit; continue; }
    if (i + 1 >= f.size()) fail("timestamp format ends in '%'");
    const char c = f[++i];
    if (c == '%') { ++lit; continue; }
    int32_t width = 0;
    switch (c) {
      case 'Y': width = 4; break;
      case 'm': case 'd': case 'H': case 'M': case 'S': width = 2; break;
      default: fail(std::string("timestamp format specifier %") + c + " is not supported");
    }
    flush();
    items.push_back(((int32_t)c << 8) | width);
  }
  flush();
  items.insert(items.begin(), (int32_t)items.size());
  return items;
}

// LIKE pattern -> {n_units, flags, units...} for the device matcher: the
// reference's translation to a regex (LogicalProject.cpp:31-47), anchored at
// the start unless the pattern's first character is '%' and at the end unless
// its last is. Without regex syntax in the pattern the regex is reduced to
// literal bytes (unit >= 0), '(?:.)' (-1) and '(?:.*?)' (-2) for the direct
// matcher; otherwise the pattern's characters keep the regex meaning the
// reference gives them ('a.b%' matches "axb..."), and the translated regex is
// compiled for the bitset matcher (flags LIKE_REGEX, units = regex_compile).
std::vector<int32_t> like_compile(const std::string& pat) {
  enum { ANCHOR_START = 1, ANCHOR_END = 2, MATCH_ALL = 4, REGEX = 8 };
  const char ANY_RUN = '\x01', ANY_CHAR = '\x02';
  std::vector<int32_t> units;
  int32_t flags = 0;
  if (pat.empty()) {
    flags = MATCH_ALL;  // the empty regex is found in every string
  } else {
    bool syntax = false;
    for (char c : pat) {
      if (c == ANY_RUN || c == ANY_CHAR) fail("LIKE pattern with a control character \\x01 or \\x02");
      if (c != 0 && std::strchr(".^$*+?()[]{}|\\", c)) syntax = true;
    }
    if (pat.front() != '%') flags |= ANCHOR_START;
    if (pat.back() != '%') flags |= ANCHOR_END;
    const std::string marked = like_pass(like_pass(pat, '%', ANY_RUN), '_', ANY_CHAR);
    if (syntax) {
      std::string re = (flags & ANCHOR_START) ? "^" : "";
      for (char c : marked) re += c == ANY_RUN ? "(?:.*?)" : (c == ANY_CHAR ? "(?:.)" : std::string(1, c));
      if (flags & ANCHOR_END) re += "$";
      flags |= REGEX;
      units = regex_compile(re);
    } else {
      for (char c : marked) units.push_back(c == ANY_RUN ? -2 : (c == ANY_CHAR ? -1 : (int32_t)(unsigned char)c));
    }
  }
  std::vector<int32_t> out{(int32_t)units.size(), flags};
  out.insert(out.end(), units.begin(), units.end());
  return out;
}

// ------------------------------------------------------------------ lowering

namespace {

// One SSA value of the lowered program.
struct Value {
  int def = -1;        // index in the emitted list
  int last_use = -1;
  int reg = -1;
  bool is_load = false;
  int column = -1;     // is_load: input column
  int str_col = -1;    // string values: the column whose chars the slice points into
  bool emitted = false;
};

// An op of the lowered program before register assignment.
struct LOp {
  DevOp d{};
  int va = -1, vb = -1;  // operand values (-1: literal or none)
  int vdst = -1;         // defined value
};

struct SourceOp {  // one reference op in SSA form
  int prog_op = -1;
  int va = -1, vb = -1, vdst = -1;
  DevOp d{};
};

struct Lowering {
  const int32_t* in_types;
  int n_in;
  std::vector<Value> values;
  std::vector<int> load_value;  // per input column
  std::vector<LOp> out;

  int new_value() { values.push_back(Value{}); return (int)values.size() - 1; }

  int load_of(int col) {
    if (load_value[col] < 0) {
      int v = new_value();
      values[v].is_load = true;
      values[v].column = col;
      if (is_string_type(in_types[col])) values[v].str_col = col;
      load_value[col] = v;
    }
    return load_value[col];
  }

  void emit_value_load(int v) {
    if (v < 0 || !values[v].is_load || values[v].emitted) return;
    LOp l;
    const int col = values[v].column;
    l.d.code = is_string_type(in_types[col]) ? D_LOADS : D_LOAD;
    l.d.a = (uint8_t)col;
    l.d.aux = (uint8_t)in_types[col];
    l.vdst = v;
    values[v].emitted = true;
    values[v].def = (int)out.size();
    out.push_back(l);
  }

  void emit(const LOp& op) {
    emit_value_load(op.va);
    emit_value_load(op.vb);
    LOp o = op;
    if (o.vdst >= 0) {
      values[o.vdst].def = (int)out.size();
      values[o.vdst].emitted = true;
    }
    out.push_back(o);
  }
};

uint8_t unit_code(int t) {  // 0 days .. 4 ns
  switch (t) {
    case BSQL_TIMESTAMP_DAYS: return 0;
    case BSQL_TIMESTAMP_SECONDS: return 1;
    case BSQL_TIMESTAMP_MILLISECONDS: return 2;
    case BSQL_TIMESTAMP_MICROSECONDS: return 3;
    default: return 4;
  }
}

int64_t f64_bits(double d) { int64_t b; std::memcpy(&b, &d, 8); return b; }

// slot bits of a literal, as get_scalar_value<T> yields it for the kind of its type
int64_t scalar_bits(const Scalar& s) {
  if (is_float_type(s.type_id)) return f64_bits(s.f64);
  return s.i64;
}

// The column whose chars a string value slices.
int str_col_of(const Lowering& L, int v) {
  if (v < 0 || L.values[(size_t)v].str_col < 0) fail("string operand without a source column");
  return L.values[(size_t)v].str_col;
}

// Converts reference op i of program p into SSA form (without emitting).
SourceOp translate(Lowering& L, const Program& p, const OpTypes& t, size_t i,
                   std::vector<int>& slot_val, int64_t str_base_a, int64_t str_base_b,
                   uint64_t seed, int n_slots_in, const std::vector<int>& col_of) {
  SourceOp s;
  s.prog_op = (int)i;
  DevOp& d = s.d;
  const int op = p.ops[i];
  const int l = p.left[i], r = p.right[i];
  const int lt = t.lt[i], rt = t.rt[i];
  const bool lf = is_float_type(lt), rf = is_float_type(rt);
  auto operand = [&](int pos) -> int {
    if (pos < 0) return -1;
    if (pos < n_slots_in && slot_val[pos] < 0) slot_val[pos] = L.load_of(col_of[(size_t)pos]);
    if (slot_val[pos] < 0) fail("slot " + std::to_string(pos) + " is read before it is written");
    return slot_val[pos];
  };
  if (r != BSQL_UNARY_INDEX && r != BSQL_NULLARY_INDEX) {
    s.va = operand(l);
    s.vb = operand(r);
    if (l == BSQL_SCALAR_INDEX) { d.flags |= F_A_IMM; d.imm = scalar_bits(p.lsc[i]); }
    if (l == BSQL_SCALAR_NULL_INDEX) d.flags |= F_A_NULL;
    if (r == BSQL_SCALAR_INDEX) { d.flags |= F_B_IMM; d.imm = scalar_bits(p.rsc[i]); }
    if (r == BSQL_SCALAR_NULL_INDEX) d.flags |= F_B_NULL;
    if (lf) d.flags |= F_A_F;
    if (rf) d.flags |= F_B_F;
    const bool ii = !lf && !rf;
    switch (op) {
      case BSQL_OP_ADD: d.code = ii ? D_ADD_II : D_ADD_FF; break;
      case BSQL_OP_SUB: d.code = ii ? D_SUB_II : D_SUB_FF; break;
      case BSQL_OP_MUL: d.code = ii ? D_MUL_II : D_MUL_FF; break;
      case BSQL_OP_DIV: d.code = ii ? D_DIV_II : D_DIV_FF; break;
      case BSQL_OP_MOD: d.code = ii ? D_MOD_II : D_MOD_FF; break;
      case BSQL_OP_POW: d.code = D_POW; break;
      case BSQL_OP_ROUND: d.code = D_ROUND; break;
      case BSQL_OP_EQUAL: case BSQL_OP_NOT_EQUAL: case BSQL_OP_LESS: case BSQL_OP_GREATER:
      case BSQL_OP_LESS_EQUAL: case BSQL_OP_GREATER_EQUAL: {
        const int k = op - BSQL_OP_EQUAL;
        if (is_string_type(lt) || is_string_type(rt)) {
          if (!(is_string_type(lt) && is_string_type(rt)))
            fail("comparison between a string and a non-string operand");
          d.code = (uint8_t)(D_EQ_SS + k);
          if (l >= 0) d.aux = (uint8_t)str_col_of(L, s.va); else { d.aux = 0xff; d.imm = str_base_a; }
          if (r >= 0) d.aux2 = (uint8_t)str_col_of(L, s.vb); else { d.aux2 = 0xff; d.imm = str_base_b; }
          if (l < 0 && r < 0) fail("Operations between literals is not supported");
        } else if (is_timestamp_type(lt) && is_timestamp_type(rt)) {
          d.code = (uint8_t)(D_EQ_TS + k);
          d.aux2 = (uint8_t)(unit_code(lt) | (unit_code(rt) << 4));
        } else {
          d.code = (uint8_t)((ii ? D_EQ_II : D_EQ_FF) + k);
        }
        break;
      }
      case BSQL_OP_LOGICAL_AND: d.code = D_AND; break;
      case BSQL_OP_LOGICAL_OR: d.code = D_OR; break;
      case BSQL_OP_IS_NOT_DISTINCT_FROM: d.code = ii ? D_INDF_II : D_INDF_FF; break;
      case BSQL_OP_MAGIC_IF_NOT: d.code = D_MAGIC_IF_NOT; break;
      case BSQL_OP_FIRST_NON_MAGIC: d.code = D_FIRST_NON_MAGIC; break;
      // string functions (evaluated ahead of the interpreter in the reference,
      // LogicalProject.cpp:100-330): here they work on slices in registers
      case BSQL_OP_STR_LIKE:
        if (!is_string_type(lt) || l < 0) fail("LIKE needs a string column or expression");
        if (r != BSQL_SCALAR_INDEX || !is_string_type(rt)) fail("LIKE needs a string literal pattern");
        d.code = D_LIKE;
        d.aux = (uint8_t)str_col_of(L, s.va);
        d.imm = str_base_b;  // the pattern, compiled when the literals are uploaded
        {
          // a literal prefix ('abc%') or a whole literal ('abc') of at most 7
          // bytes: compared in place (the same strings match as under the
          // matcher: anchored at the start, the trailing '%' takes any rest)
          enum { ANCHOR_START = 1, ANCHOR_END = 2, MATCH_ALL = 4, REGEX = 8 };
          const std::vector<int32_t> u = like_compile(p.rsc[i].str);
          const int32_t nu = u[0], fl = u[1];
          int32_t k = 0;
          while (k < nu && u[(size_t)(2 + k)] >= 0) ++k;
          const bool plain = (fl & ANCHOR_START) && !(fl & (MATCH_ALL | REGEX));
          const bool prefix = plain && !(fl & ANCHOR_END) && k == nu - 1 && u[(size_t)(2 + k)] == -2;
          const bool whole = plain && (fl & ANCHOR_END) && k == nu;
          if ((prefix || whole) && k <= 7) {
            uint64_t packed = (uint64_t)k << 56;
            for (int32_t j = 0; j < k; ++j) packed |= (uint64_t)(uint8_t)u[(size_t)(2 + j)] << (8 * j);
            d.code = D_LIKE_PREFIX;
            d.imm = (int64_t)packed;
            d.aux2 = whole ? 1 : 0;
          }
        }
        break;
      case BSQL_OP_STR_SUBSTRING: case BSQL_OP_STR_LEFT: case BSQL_OP_STR_RIGHT: {
        if (!is_string_type(lt) || l < 0) fail("string function on a non-string operand");
        if (is_string_type(rt) || rf) fail("string position arguments must be integers");
        d.code = D_SLICE;
        d.aux = (uint8_t)str_col_of(L, s.va);
        d.aux2 = (uint8_t)(op == BSQL_OP_STR_SUBSTRING ? 0 : (op == BSQL_OP_STR_LEFT ? 1 : 2));
        if (r == BSQL_SCALAR_INDEX) {
          // literal arguments are clamped as the reference does before
          // slice_strings: SUBSTRING start max(s, 1) - 1, LEFT/RIGHT max(n, 0)
          const int64_t v = p.rsc[i].i64;
          d.imm = op == BSQL_OP_STR_SUBSTRING ? std::max<int64_t>(v, 1) - 1 : std::max<int64_t>(v, 0);
          d.flags |= F_SLICE_LIT;
        }
        break;
      }
      case BSQL_OP_TO_DATE: case BSQL_OP_TO_TIMESTAMP: {
        // to_timestamps(column, type, format), LogicalProject.cpp:510-533
        if (!is_string_type(lt) || l < 0) fail(std::string(op == BSQL_OP_TO_DATE ? "TO_DATE" : "TO_TIMESTAMP") +
                                               " first argument must be a column of type string");
        if (r != BSQL_SCALAR_INDEX || !is_string_type(rt))
          fail("TO_DATE / TO_TIMESTAMP take a string literal format");
        d.code = D_STR_TO_TS;
        d.aux = (uint8_t)str_col_of(L, s.va);
        d.aux2 = (uint8_t)(op == BSQL_OP_TO_DATE ? BSQL_TIMESTAMP_DAYS : BSQL_TIMESTAMP_NANOSECONDS);
        d.imm = str_base_b;  // the format, compiled when the literals are uploaded
        break;
      }
      case BSQL_OP_STR_TRIM: {
        // the characters to strip and the side come in the right scalar
        if (!is_string_type(lt) || l < 0) fail("TRIM needs a string column or expression");
        if (r != BSQL_SCALAR_INDEX || !is_string_type(rt)) fail("TRIM needs a literal set of characters");
        const int64_t side = p.rsc[i].i64;
        if (side < 1 || side > 3) fail("TRIM needs a side (LEADING, TRAILING or BOTH)");
        d.code = D_TRIM;
        d.aux = (uint8_t)str_col_of(L, s.va);
        d.aux2 = (uint8_t)side;
        d.imm = str_base_b;
        break;
      }
      default: fail("operator " + std::to_string(op) + " is not supported by the interpreter");
    }
    if (is_string_type(lt) && !((d.code >= D_EQ_SS && d.code <= D_GE_SS) || d.code == D_LIKE || d.code == D_LIKE_PREFIX ||
                                d.code == D_SLICE || d.code == D_TRIM || d.code == D_STR_TO_TS))
      fail("string operand to a non-string operator");
    if (is_string_type(rt) && !((d.code >= D_EQ_SS && d.code <= D_GE_SS) || d.code == D_LIKE || d.code == D_LIKE_PREFIX ||
                                d.code == D_TRIM || d.code == D_STR_TO_TS))
      fail("string operand to a non-string operator");
  } else if (r == BSQL_UNARY_INDEX) {
    s.va = operand(l);
    if (lf) d.flags |= F_A_F;
    switch (op) {
      case BSQL_OP_NOT: d.code = D_NOT; break;
      case BSQL_OP_ABS: d.code = lf ? D_ABS_F : D_ABS_I; break;
      case BSQL_OP_FLOOR: d.code = D_FLOOR; break;
      case BSQL_OP_CEIL: d.code = D_CEIL; break;
      case BSQL_OP_SIN: d.code = D_SIN; break;
      case BSQL_OP_COS: d.code = D_COS; break;
      case BSQL_OP_ASIN: d.code = D_ASIN; break;
      case BSQL_OP_ACOS: d.code = D_ACOS; break;
      case BSQL_OP_TAN: d.code = D_TAN; break;
      case BSQL_OP_COTAN: d.code = D_COTAN; break;
      case BSQL_OP_ATAN: d.code = D_ATAN; break;
      case BSQL_OP_LN: d.code = D_LN; break;
      case BSQL_OP_LOG: d.code = D_LOG; break;
      case BSQL_OP_YEAR: case BSQL_OP_MONTH: case BSQL_OP_DAY: case BSQL_OP_DAYOFWEEK:
      case BSQL_OP_HOUR: case BSQL_OP_MINUTE: case BSQL_OP_SECOND:
        d.code = D_EXTRACT; d.aux = (uint8_t)op; d.aux2 = (uint8_t)lt; break;
      case BSQL_OP_IS_NULL: d.code = D_IS_NULL; break;
      case BSQL_OP_IS_NOT_NULL: d.code = D_IS_NOT_NULL; break;
      case BSQL_OP_IS_TRUE: d.code = D_IS_TRUE; break;
      case BSQL_OP_IS_NOT_TRUE: d.code = D_IS_NOT_TRUE; break;
      case BSQL_OP_IS_NOT_FALSE: d.code = D_IS_NOT_FALSE; break;
      // a string operand is parsed (cudf::strings::to_integers / to_floats,
      // LogicalProject.cpp:379-468)
      case BSQL_OP_CAST_TINYINT: case BSQL_OP_CAST_SMALLINT: case BSQL_OP_CAST_INTEGER:
      case BSQL_OP_CAST_BIGINT:
        d.code = is_string_type(lt) ? D_STR_TO_I64 : D_TO_I64;
        if (is_string_type(lt)) d.aux = (uint8_t)str_col_of(L, s.va);
        break;
      case BSQL_OP_CAST_FLOAT: case BSQL_OP_CAST_DOUBLE:
        d.code = is_string_type(lt) ? D_STR_TO_F64 : D_TO_F64;
        if (is_string_type(lt)) d.aux = (uint8_t)str_col_of(L, s.va);
        break;
      // a string operand is parsed (to_timestamps with "%Y-%m-%d" or
      // "%Y-%m-%d %H:%M:%S", LogicalProject.cpp:469-508)
      case BSQL_OP_CAST_DATE: case BSQL_OP_CAST_TIMESTAMP: case BSQL_OP_CAST_TIMESTAMP_SECONDS:
      case BSQL_OP_CAST_TIMESTAMP_MILLISECONDS: case BSQL_OP_CAST_TIMESTAMP_MICROSECONDS:
        if (is_string_type(lt)) {
          d.code = D_STR_TO_TS;
          d.aux = (uint8_t)str_col_of(L, s.va);
          d.aux2 = (uint8_t)output_type_unary(op, lt);
          d.imm = 0;
        } else {
          d.code = op == BSQL_OP_CAST_DATE ? D_CAST_DATE : D_CAST_TS;
          d.aux2 = (uint8_t)lt;
        }
        break;
      case BSQL_OP_CHAR_LENGTH:
        if (!is_string_type(lt)) fail("CHAR_LENGTH needs a string operand");
        d.code = D_CHAR_LENGTH; d.aux = (uint8_t)str_col_of(L, s.va); break;
      default: fail("operator " + std::to_string(op) + " is not supported by the interpreter");
    }
    if (is_string_type(lt) && d.code != D_CHAR_LENGTH && d.code != D_IS_NULL &&
        d.code != D_IS_NOT_NULL && d.code != D_STR_TO_I64 && d.code != D_STR_TO_F64 && d.code != D_STR_TO_TS)
      fail("string operand to a non-string operator");
  } else {
    d.code = D_RAND;
    d.imm = (int64_t)seed;
  }
  return s;
}

// Result value of each op and the ops each value depends on.
struct SsaProgram {
  std::vector<SourceOp> ops;
  std::vector<int> finals;  // value per final output
  OpTypes types;
  std::vector<int32_t> slot_types;
};

// Interpreter input slot -> table column: programs compiled from text refer to
// the used columns only (col_idx_map, LogicalProject.cpp:859-881); programs
// built from arrays address the table directly.
std::vector<int> slot_columns(const Program& p, int n_table) {
  std::vector<int> col_of;
  if (!p.used_columns.empty() || p.n_inputs == 0) {
    for (int32_t c : p.used_columns) {
      if (c < 0 || c >= n_table) fail("program refers to column " + std::to_string(c) + " outside the table");
      col_of.push_back(c);
    }
  } else {
    for (int k = 0; k < n_table; ++k) col_of.push_back(k);
  }
  return col_of;
}

}  // namespace

std::vector<int32_t> slot_types_for(const Program& p, const int32_t* table_types, int n_table) {
  std::vector<int32_t> t;
  for (int c : slot_columns(p, n_table)) t.push_back(table_types[c]);
  return t;
}

namespace {

bool is_float_code(uint8_t c) {
  switch (c) {
    case D_ADD_FF: case D_SUB_FF: case D_MUL_FF: case D_DIV_FF: case D_MOD_FF: case D_POW: case D_ROUND:
    case D_EQ_FF: case D_NE_FF: case D_LT_FF: case D_GT_FF: case D_LE_FF: case D_GE_FF: case D_INDF_FF:
    case D_FLOOR: case D_CEIL: case D_SIN: case D_COS: case D_ASIN: case D_ACOS: case D_TAN:
    case D_COTAN: case D_ATAN: case D_LN: case D_LOG:
      return true;
    default:
      return false;
  }
}

bool is_unary_code(uint8_t c) { return c >= D_NOT; }

// Float ops read both operands as doubles on the device, the conversion the
// reference applies when it reads an integer slot into a double operand
// (interpreter_ops.cuh:600-727): int literals are converted here, int values
// by the op itself (F_A_CVT / F_B_CVT: (double)(int64) of the slot), so a
// conversion costs no op of its own.
void float_operands(Lowering&, SsaProgram& s) {
  for (SourceOp& so : s.ops) {
    if (!is_float_code(so.d.code)) continue;
    auto fix = [&](int v, uint8_t fflag, uint8_t immflag, uint8_t nullflag, uint8_t cvtflag) {
      DevOp& d = so.d;
      if (d.flags & fflag) return;
      if (d.flags & immflag) { d.imm = f64_bits((double)d.imm); d.flags |= fflag; return; }
      if (d.flags & nullflag) { d.flags |= fflag; return; }
      if (v < 0) return;
      d.flags |= fflag | cvtflag;
    };
    fix(so.va, F_A_F, F_A_IMM, F_A_NULL, F_A_CVT);
    if (!is_unary_code(so.d.code)) fix(so.vb, F_B_F, F_B_IMM, F_B_NULL, F_B_CVT);
  }
}

// Common subexpressions, then fused pairs.
// CASE WHEN x IS NOT NULL THEN x ... (COALESCE, and the reference's CASE
// lowering) reaches the interpreter as IS_NOT_NULL(x) feeding
// MAGIC_IF_NOT(_, x); when nothing else reads the IS_NOT_NULL the pair is
// one op, MAGIC_IF_NULL(x), with the same result and validity
// (interpreter_ops.cuh:559-566 with a condition that is valid(x)).
void fuse_ops(Lowering& L, SsaProgram& s) {
  // common subexpressions: a text that repeats an expression (Calcite's CASE
  // and COALESCE repeat their operands) computes it once; every op but RAND
  // is a pure function of its operands and fields
  {
    std::vector<int> repl(L.values.size());
    for (size_t v = 0; v < repl.size(); ++v) repl[v] = (int)v;
    std::map<std::tuple<int, int, int, int, int64_t, int, int>, int> seen;
    std::vector<SourceOp> kept;
    for (SourceOp so : s.ops) {
      if (so.va >= 0) so.va = repl[(size_t)so.va];
      if (so.vb >= 0) so.vb = repl[(size_t)so.vb];
      if (so.d.code != D_RAND) {
        const auto key = std::make_tuple((int)so.d.code, (int)so.d.flags, (int)so.d.aux, (int)so.d.aux2, so.d.imm,
                                         so.va, so.vb);
        auto it = seen.find(key);
        if (it != seen.end()) {
          repl[(size_t)so.vdst] = it->second;
          continue;
        }
        seen.emplace(key, so.vdst);
      }
      kept.push_back(so);
    }
    for (int& v : s.finals) v = repl[(size_t)v];
    s.ops = kept;
  }
  std::vector<int> uses(L.values.size(), 0);
  for (const SourceOp& o : s.ops) {
    if (o.va >= 0) ++uses[(size_t)o.va];
    if (o.vb >= 0) ++uses[(size_t)o.vb];
  }
  for (int v : s.finals) ++uses[(size_t)v];
  std::vector<int> prod(L.values.size(), -1);
  for (size_t i = 0; i < s.ops.size(); ++i) prod[(size_t)s.ops[i].vdst] = (int)i;
  std::vector<char> dead(s.ops.size(), 0);
  const uint8_t literal = F_A_IMM | F_A_NULL | F_B_IMM | F_B_NULL;
  for (SourceOp& m : s.ops) {
    if (m.d.code != D_MAGIC_IF_NOT || (m.d.flags & literal) || m.va < 0 || m.vb < 0) continue;
    if (uses[(size_t)m.va] != 1) continue;
    const int p = prod[(size_t)m.va];
    if (p < 0) continue;
    const SourceOp& t = s.ops[(size_t)p];
    if (t.d.code != D_IS_NOT_NULL || (t.d.flags & literal) || t.va != m.vb) continue;
    dead[(size_t)p] = 1;
    m.d.code = D_MAGIC_IF_NULL;
    m.d.flags = (m.d.flags & F_B_F) ? F_A_F : 0;
    m.va = m.vb;
    m.vb = -1;
  }
  // FIRST_NON_MAGIC(MAGIC_IF_NULL(x), y), the MAGIC_IF_NULL read by nothing
  // else: COALESCE(x, y) (y may be a literal); both ops must agree on x's kind
  for (size_t i = 0; i < s.ops.size(); ++i) {
    if (dead[i]) continue;
    SourceOp& f = s.ops[i];
    if (f.d.code != D_FIRST_NON_MAGIC || (f.d.flags & (F_A_IMM | F_A_NULL)) || f.va < 0) continue;
    if (uses[(size_t)f.va] != 1) continue;
    const int p = prod[(size_t)f.va];
    if (p < 0 || dead[(size_t)p]) continue;
    const SourceOp& m = s.ops[(size_t)p];
    if (m.d.code != D_MAGIC_IF_NULL || ((m.d.flags ^ f.d.flags) & F_A_F)) continue;
    dead[(size_t)p] = 1;
    f.d.code = D_COALESCE;
    f.va = m.va;
  }
  std::vector<SourceOp> out;
  for (size_t i = 0; i < s.ops.size(); ++i)
    if (!dead[i]) out.push_back(s.ops[i]);
  s.ops = out;
}

SsaProgram to_ssa(Lowering& L, const Program& p, uint64_t seed) {
  const std::vector<int> col_of = slot_columns(p, L.n_in);
  const int n_slots_in = (int)col_of.size();
  std::vector<int32_t> slot_types(col_of.size());
  for (size_t k = 0; k < col_of.size(); ++k) slot_types[k] = L.in_types[col_of[k]];
  validate(p, n_slots_in);
  OpTypes t = propagate_types(p, slot_types.data(), n_slots_in);
  SsaProgram s;
  std::vector<int> slot_val(512, -1);
  for (size_t i = 0; i < p.ops.size(); ++i) {
    int64_t sa = -1, sb = -1;
    if (p.str_offsets.size() == 2 * p.ops.size()) { sa = p.str_offsets[2 * i]; sb = p.str_offsets[2 * i + 1]; }
    int64_t base = (int64_t)(intptr_t)p.d_strings;
    SourceOp so = translate(L, p, t, i, slot_val, sa >= 0 ? base + sa : 0, sb >= 0 ? base + sb : 0,
                            seed, n_slots_in, col_of);
    so.vdst = L.new_value();
    if (so.d.code == D_SLICE || so.d.code == D_TRIM)
      L.values[(size_t)so.vdst].str_col = L.values[(size_t)so.va].str_col;
    slot_val[p.out[i]] = so.vdst;
    s.ops.push_back(so);
  }
  float_operands(L, s);
  for (int k = 0; k < num_roots(p); ++k) {
    const Root rt = root_of(p, k);
    if (rt.kind == Root::LITERAL) {
      // literal root: an identity conversion of the literal (null: no value)
      const Scalar& lit = rt.lit;
      SourceOp c;
      c.prog_op = -1;
      c.d.code = is_float_type(lit.type_id) ? D_TO_F64 : D_TO_I64;
      c.d.flags = (lit.is_null ? F_A_NULL : F_A_IMM) | (is_float_type(lit.type_id) ? F_A_F : 0);
      c.d.imm = lit.is_null ? 0 : scalar_bits(lit);
      c.vdst = L.new_value();
      s.ops.push_back(c);
      s.finals.push_back(c.vdst);
      continue;
    }
    if (rt.kind == Root::COLUMN) {
      // bare column root: the column itself (a table column or a step's)
      if (rt.index < 0 || rt.index >= L.n_in) fail("output column $" + std::to_string(rt.index) + " not in the table");
      s.finals.push_back(L.load_of(rt.index));
      continue;
    }
    if (rt.index < 0 || rt.index >= (int)p.final_out.size()) fail("output " + std::to_string(k) + " has no slot");
    const int16_t f = p.final_out[(size_t)rt.index];
    if (f < 0) fail("output " + std::to_string(k) + " has no slot");
    int v = f < n_slots_in && slot_val[f] < 0 ? L.load_of(col_of[(size_t)f]) : slot_val[f];
    if (v < 0) fail("output slot " + std::to_string(f) + " is never written");
    s.finals.push_back(v);
  }
  s.types = t;
  s.slot_types = slot_types;
  fuse_ops(L, s);
  return s;
}

// producer op (index into ssa.ops) of each value, -1 for loads
std::vector<int> producers(const SsaProgram& s, size_t nvalues) {
  std::vector<int> prod(nvalues, -1);
  for (size_t i = 0; i < s.ops.size(); ++i) prod[s.ops[i].vdst] = (int)i;
  return prod;
}

void subtree(const SsaProgram& s, const std::vector<int>& prod, int v, std::set<int>& acc) {
  if (v < 0) return;
  int o = prod[v];
  if (o < 0 || acc.count(o)) return;
  acc.insert(o);
  subtree(s, prod, s.ops[o].va, acc);
  subtree(s, prod, s.ops[o].vb, acc);
}

// Flattens the AND tree rooted at value v into its conjunct values.
void conjuncts_of(const Program& p, const SsaProgram& s, const std::vector<int>& prod, int v,
                  std::vector<int>& acc) {
  int o = v >= 0 ? prod[v] : -1;
  if (o >= 0 && s.ops[o].prog_op >= 0 && p.ops[s.ops[o].prog_op] == BSQL_OP_LOGICAL_AND && s.ops[o].va >= 0 &&
      s.ops[o].vb >= 0) {
    conjuncts_of(p, s, prod, s.ops[o].va, acc);
    conjuncts_of(p, s, prod, s.ops[o].vb, acc);
  } else {
    acc.push_back(v);
  }
}

// The conjunct with value v as a range test on one column, if it is a
// comparison of a column with a literal in the int64 or double domain (the
// domain the comparison op itself would use).
bool as_range(const Lowering& L, const SsaProgram& s, const std::vector<int>& prod, int v,
              DevRange& rg) {
  const int o = v >= 0 ? prod[(size_t)v] : -1;
  if (o < 0) return false;
  const DevOp& d = s.ops[(size_t)o].d;
  int base;
  if (d.code >= D_EQ_II && d.code <= D_GE_II) base = D_EQ_II;
  else if (d.code >= D_EQ_FF && d.code <= D_GE_FF) base = D_EQ_FF;
  else if (d.code >= D_EQ_TS && d.code <= D_GE_TS && (d.aux2 & 15) == (d.aux2 >> 4)) base = D_EQ_TS;
  else return false;
  if (d.flags & (F_A_NULL | F_B_NULL)) return false;
  const bool lit_left = (d.flags & F_A_IMM) != 0, lit_right = (d.flags & F_B_IMM) != 0;
  if (lit_left == lit_right) return false;
  int cv = lit_left ? s.ops[(size_t)o].vb : s.ops[(size_t)o].va;
  if (cv >= 0 && !L.values[(size_t)cv].is_load && prod[(size_t)cv] >= 0) {
    // an int column converted for a float comparison: the range test converts
    const SourceOp& c = s.ops[(size_t)prod[(size_t)cv]];
    if (c.d.code == D_TO_F64 && !(c.d.flags & (F_A_F | F_A_IMM | F_A_NULL))) cv = c.va;
  }
  if (cv < 0 || !L.values[(size_t)cv].is_load) return false;
  const int col = L.values[(size_t)cv].column;
  if (is_string_type(L.in_types[col])) return false;
  int k = d.code - base;  // EQ NE LT GT LE GE
  if (lit_left) {
    static const int flip[6] = {0, 1, 3, 2, 5, 4};
    k = flip[k];
  }
  rg = DevRange{};
  rg.col = col;
  rg.mul = 1;
  if (base == D_EQ_FF) {
    // double domain; the literal is already double bits (float_operands)
    rg.kind = RK_F;
    double lit;
    std::memcpy(&lit, &d.imm, 8);
    const double inf = std::numeric_limits<double>::infinity();
    double lo = -inf, hi = inf;
    bool empty = std::isnan(lit);  // only NE is true against NaN
    switch (k) {
      case 0: lo = hi = lit; break;
      case 1: rg.kind |= RK_NE; rg.lo = f64_bits(lit); return true;
      case 2: if (lit == -inf) empty = true; else hi = std::nextafter(lit, -inf); break;  // x < lit
      case 3: if (lit == inf) empty = true; else lo = std::nextafter(lit, inf); break;    // x > lit
      case 4: hi = lit; break;
      default: lo = lit; break;
    }
    if (empty) rg.kind |= RK_EMPTY;
    rg.lo = f64_bits(lo);
    rg.hi = f64_bits(hi);
    return true;
  }
  // int64 domain; timestamps compare as nanoseconds (wrapping, like D_*_TS)
  if (base == D_EQ_TS) {
    static const int64_t ns[5] = {86400000000000LL, 1000000000LL, 1000000LL, 1000LL, 1LL};
    rg.mul = ns[d.aux2 & 15];
  }
  const int64_t lit = (int64_t)((uint64_t)d.imm * (uint64_t)rg.mul);
  int64_t lo = INT64_MIN, hi = INT64_MAX;
  bool empty = false;
  switch (k) {
    case 0: lo = hi = lit; break;
    case 1: rg.kind = RK_NE; rg.lo = lit; return true;
    case 2: if (lit == INT64_MIN) empty = true; else hi = lit - 1; break;
    case 3: if (lit == INT64_MAX) empty = true; else lo = lit + 1; break;
    case 4: hi = lit; break;
    default: lo = lit; break;
  }
  if (empty) rg.kind |= RK_EMPTY;
  rg.lo = lo;
  rg.hi = hi;
  return true;
}

// Intersects range b into a (same column and domain, neither NE).
void intersect(DevRange& a, const DevRange& b) {
  if (b.kind & RK_EMPTY) a.kind |= RK_EMPTY;
  if (a.kind & RK_F) {
    double alo, ahi, blo, bhi;
    std::memcpy(&alo, &a.lo, 8); std::memcpy(&ahi, &a.hi, 8);
    std::memcpy(&blo, &b.lo, 8); std::memcpy(&bhi, &b.hi, 8);
    if (blo > alo) a.lo = b.lo;
    if (bhi < ahi) a.hi = b.hi;
    std::memcpy(&alo, &a.lo, 8); std::memcpy(&ahi, &a.hi, 8);
    if (alo > ahi) a.kind |= RK_EMPTY;
  } else {
    a.lo = std::max(a.lo, b.lo);
    a.hi = std::min(a.hi, b.hi);
    if (a.lo > a.hi) a.kind |= RK_EMPTY;
  }
}

LOp to_lop(const SourceOp& so) {
  LOp l;
  l.d = so.d;
  l.va = so.va;
  l.vb = so.vb;
  l.vdst = so.vdst;
  return l;
}

}  // namespace

int count_conjuncts(const Program& p) {
  // The tree shape does not depend on the column types: a neutral schema suffices.
  int n_table = BSQL_MAX_SLOT + 1;
  for (int32_t c : p.used_columns) n_table = std::max(n_table, c + 1);
  std::vector<int32_t> types((size_t)n_table, BSQL_INT64);
  Lowering L{types.data(), n_table, {}, {}, {}};
  L.load_value.assign((size_t)n_table, -1);
  try {
    SsaProgram s = to_ssa(L, p, 0);
    if (s.finals.size() != 1) return 0;
    std::vector<int> prod = producers(s, L.values.size());
    std::vector<int> conj;
    conjuncts_of(p, s, prod, s.finals[0], conj);
    return (int)conj.size();
  } catch (const Error&) {
    return 0;
  }
}

std::vector<DevRange> conjunct_ranges(const Program& p, const int32_t* in_types, int n_in) {
  Lowering L{in_types, n_in, {}, {}, {}};
  L.load_value.assign((size_t)n_in, -1);
  SsaProgram s = to_ssa(L, p, 0);
  if (s.finals.size() != 1) fail("a predicate has one output");
  std::vector<int> prod = producers(s, L.values.size());
  std::vector<int> conj;
  conjuncts_of(p, s, prod, s.finals[0], conj);
  std::vector<DevRange> out;
  for (int v : conj) {
    DevRange rg;
    if (!as_range(L, s, prod, v, rg)) fail("conjunct is not a column-vs-literal comparison");
    out.push_back(rg);
  }
  return out;
}

namespace {

// Ops of `s` that the values `roots` depend on (dead-code elimination), as
// a flag per op.
std::vector<char> needed_ops(const SsaProgram& s, const std::vector<int>& prod, const std::vector<int>& roots) {
  std::vector<char> need(s.ops.size(), 0);
  std::vector<int> stack(roots.begin(), roots.end());
  while (!stack.empty()) {
    const int v = stack.back();
    stack.pop_back();
    const int o = v >= 0 ? prod[(size_t)v] : -1;
    if (o < 0 || need[(size_t)o]) continue;
    need[(size_t)o] = 1;
    stack.push_back(s.ops[(size_t)o].va);
    stack.push_back(s.ops[(size_t)o].vb);
  }
  return need;
}

// Column loads issued as early as registers allow. `ops` holds the op list
// with every D_LOAD right before the first op that reads it; within each
// segment [bounds[k], bounds[k+1]) (the selection phase or, with lazy loads,
// one conjunct; the output phase) a load moves up past earlier ops while
// the values live there, plus it, leave room for one result within `budget`
// registers, and never above an earlier load (so runs of loads stay in
// first-use order and merge into multi-column LOADN issues).
std::vector<LOp> place_loads(const std::vector<LOp>& ops, const std::vector<int>& bounds, size_t n_values,
                             int budget) {
  std::vector<LOp> res;
  for (size_t sg = 0; sg + 1 < bounds.size(); ++sg) {
    std::vector<LOp> body, loads;
    std::vector<int> at;  // first use of each load, as an index into body
    for (int i = bounds[sg]; i < bounds[sg + 1]; ++i) {
      const LOp& o = ops[(size_t)i];
      if (o.d.code == D_LOAD) {
        loads.push_back(o);
        at.push_back(-1);
      } else {
        body.push_back(o);
      }
    }
    const int m = (int)body.size();
    // live-in pressure before each body op, loads not counted
    std::vector<int> def(n_values, -1), last(n_values, -1);
    for (int j = 0; j < m; ++j) {
      const LOp& o = body[(size_t)j];
      for (int v : {o.va, o.vb})
        if (v >= 0) last[(size_t)v] = std::max(last[(size_t)v], j);
      if (o.vdst >= 0) def[(size_t)o.vdst] = j;
    }
    std::vector<int> pressure((size_t)m + 1, 0);
    for (int j = 0; j < m; ++j) {
      const int v = body[(size_t)j].vdst;
      if (v >= 0 && last[(size_t)v] > j)
        for (int q = j + 1; q <= last[(size_t)v]; ++q) ++pressure[(size_t)q];
    }
    for (size_t k = 0; k < loads.size(); ++k) {
      const int v = loads[k].vdst;
      for (int j = 0; j < m && at[k] < 0; ++j)
        if (body[(size_t)j].va == v || body[(size_t)j].vb == v) at[k] = j;
      if (at[k] < 0) at[k] = m;  // unused: placed at the end (never happens for emitted loads)
    }
    std::vector<size_t> order(loads.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return at[x] < at[y]; });
    std::vector<std::vector<LOp>> before((size_t)m + 1);
    int floor_pos = 0;
    for (size_t k : order) {
      const int u = at[k];
      const int lu = std::max(u, last[(size_t)loads[k].vdst]);
      int p = u;
      while (p > floor_pos && pressure[(size_t)(p - 1)] + 2 <= budget) --p;
      for (int q = p; q <= std::min(lu, m); ++q) ++pressure[(size_t)q];
      before[(size_t)p].push_back(loads[k]);
      floor_pos = p;
    }
    for (int j = 0; j <= m; ++j) {
      for (const LOp& l : before[(size_t)j]) res.push_back(l);
      if (j < m) res.push_back(body[(size_t)j]);
    }
  }
  return res;
}

// Linear-scan register assignment with spilling: values take registers
// 0..budget-1; when none is free, the value whose next use is furthest away
// goes to an LDS slot (D_SPILL, once per value) and comes back with D_FILL
// before its next use. Returns the op list with registers filled in.
std::vector<LOp> assign_registers(const std::vector<LOp>& ops, size_t n_values, int budget, int* regs_used,
                                  int* slots_used) {
  std::vector<std::vector<int>> uses(n_values);
  for (size_t i = 0; i < ops.size(); ++i)
    for (int v : {ops[i].va, ops[i].vb})
      if (v >= 0 && (uses[(size_t)v].empty() || uses[(size_t)v].back() != (int)i)) uses[(size_t)v].push_back((int)i);
  std::vector<int> reg_of(n_values, -1), slot_of(n_values, -1), holder((size_t)budget, -1);
  std::vector<int> free_slots;
  int n_slots = 0, used = 0;
  std::vector<LOp> res;
  auto next_use = [&](int v, int i) -> int {
    const std::vector<int>& u = uses[(size_t)v];
    auto it = std::upper_bound(u.begin(), u.end(), i);
    return it == u.end() ? INT32_MAX : *it;
  };
  auto last_use = [&](int v) -> int { return uses[(size_t)v].empty() ? -1 : uses[(size_t)v].back(); };
  auto take_reg = [&](int i, int keep_a, int keep_b) -> int {
    for (int r = 0; r < budget; ++r)
      if (holder[(size_t)r] < 0) {
        used = std::max(used, r + 1);
        return r;
      }
    int victim = -1, far = -1;
    for (int r = 0; r < budget; ++r) {
      const int v = holder[(size_t)r];
      if (v == keep_a || v == keep_b) continue;
      const int nu = next_use(v, i - 1);
      if (nu > far) { far = nu; victim = v; }
    }
    if (victim < 0) fail("internal: no register to spill");
    if (slot_of[(size_t)victim] < 0) {
      if (!free_slots.empty()) {
        slot_of[(size_t)victim] = free_slots.back();
        free_slots.pop_back();
      } else {
        slot_of[(size_t)victim] = n_slots++;
      }
      LOp sp;
      sp.d.code = D_SPILL;
      sp.d.a = (uint8_t)reg_of[(size_t)victim];
      sp.d.dst = 0xff;
      sp.d.imm = slot_of[(size_t)victim];
      res.push_back(sp);
    }
    const int r = reg_of[(size_t)victim];
    reg_of[(size_t)victim] = -1;
    holder[(size_t)r] = -1;
    return r;
  };
  auto release = [&](int v) {
    if (reg_of[(size_t)v] >= 0) holder[(size_t)reg_of[(size_t)v]] = -1;
    reg_of[(size_t)v] = -1;
    if (slot_of[(size_t)v] >= 0) free_slots.push_back(slot_of[(size_t)v]);
    slot_of[(size_t)v] = -1;
  };
  for (size_t i = 0; i < ops.size(); ++i) {
    LOp o = ops[i];
    // operands come back from their slots first
    for (int v : {o.va, o.vb}) {
      if (v < 0 || reg_of[(size_t)v] >= 0) continue;
      if (slot_of[(size_t)v] < 0) fail("internal: operand without a register");
      const int r = take_reg((int)i, o.va, o.vb);
      LOp f;
      f.d.code = D_FILL;
      f.d.dst = (uint8_t)r;
      f.d.imm = slot_of[(size_t)v];
      res.push_back(f);
      reg_of[(size_t)v] = r;
      holder[(size_t)r] = v;
    }
    if (o.va >= 0) o.d.a = (uint8_t)reg_of[(size_t)o.va];
    if (o.vb >= 0) o.d.b = (uint8_t)reg_of[(size_t)o.vb];
    // operands whose last use is this op free their registers (the result may reuse one)
    for (int v : {o.va, o.vb})
      if (v >= 0 && last_use(v) == (int)i && (reg_of[(size_t)v] >= 0 || slot_of[(size_t)v] >= 0)) release(v);
    if (o.vdst >= 0) {
      const int r = take_reg((int)i, -1, -1);
      o.d.dst = (uint8_t)r;
      if (last_use(o.vdst) > (int)i) {
        reg_of[(size_t)o.vdst] = r;
        holder[(size_t)r] = o.vdst;
      }
    } else {
      o.d.dst = 0xff;
    }
    res.push_back(o);
  }
  *regs_used = used;
  *slots_used = n_slots;
  return res;
}

// Table columns an op reads (loads and the chars of string operands).
void op_columns(const DevOp& d, std::vector<int>& cols) {
  if (d.code == D_LOAD || d.code == D_LOADS) cols.push_back(d.a);
  const bool str_aux = (d.code >= D_EQ_SS && d.code <= D_GE_SS) || d.code == D_LIKE || d.code == D_LIKE_PREFIX ||
                       d.code == D_SLICE ||
                       d.code == D_TRIM || d.code == D_STR_TO_I64 || d.code == D_STR_TO_F64 ||
                       d.code == D_STR_TO_TS || d.code == D_CHAR_LENGTH;
  if (str_aux && d.aux != 0xff) cols.push_back(d.aux);
  if (d.code >= D_EQ_SS && d.code <= D_GE_SS && d.aux2 != 0xff) cols.push_back(d.aux2);
}

void remap_columns(DevOp& d, const std::vector<int>& map) {
  if (d.code == D_LOAD || d.code == D_LOADS) d.a = (uint8_t)map[d.a];
  const bool str_aux = (d.code >= D_EQ_SS && d.code <= D_GE_SS) || d.code == D_LIKE || d.code == D_LIKE_PREFIX ||
                       d.code == D_SLICE ||
                       d.code == D_TRIM || d.code == D_STR_TO_I64 || d.code == D_STR_TO_F64 ||
                       d.code == D_STR_TO_TS || d.code == D_CHAR_LENGTH;
  if (str_aux && d.aux != 0xff) d.aux = (uint8_t)map[d.aux];
  if (d.code >= D_EQ_SS && d.code <= D_GE_SS && d.aux2 != 0xff) d.aux2 = (uint8_t)map[d.aux2];
}

}  // namespace

Lowered lower(const LowerOptions& opt, const int32_t* in_types, int n_in,
              const int32_t* out_types, int n_out) {
  if (n_in > 255) fail("a launch reads columns of a table of at most 255 columns");
  if (n_out > BSQL_DEV_MAX_OUT) fail("too many output columns for one launch");
  const int budget = std::max(3, std::min(opt.reg_budget, BSQL_DEV_MAX_REG - 1));
  Lowering L{in_types, n_in, {}, {}, {}};
  L.load_value.assign(n_in, -1);
  Lowered res;
  std::vector<int> store_value;  // value stored into output k

  // ---- phase 1: row selection
  if (opt.mode != LowerOptions::EXECUTE) {
    const Program& pred = *opt.pred;
    if (num_roots(pred) != 1) fail("a filter predicate must have exactly one output");
    SsaProgram s = to_ssa(L, pred, pred.seed);
    const OpTypes& t = s.types;
    const Root rt0 = root_of(pred, 0);
    int root_type = BSQL_EMPTY;
    if (rt0.kind == Root::LITERAL) {
      root_type = rt0.lit.type_id;
    } else if (rt0.kind == Root::COLUMN) {
      root_type = in_types[rt0.index];
    } else {
      const int root = pred.final_out[(size_t)rt0.index];
      root_type = root < (int)s.slot_types.size() ? s.slot_types[(size_t)root] : BSQL_EMPTY;
      for (size_t i = 0; i < pred.ops.size(); ++i)
        if (pred.out[i] == root) root_type = t.ot[i];
    }
    if (root_type != BSQL_BOOL8)
      fail("Expression does not evaluate to a boolean mask");
    std::vector<int> prod = producers(s, L.values.size());
    std::vector<int> conj;
    if (opt.conjuncts) conjuncts_of(pred, s, prod, s.finals[0], conj);
    res.n_conjuncts = (int)conj.size();
    std::vector<int32_t> order(conj.size());
    for (size_t j = 0; j < conj.size(); ++j) order[j] = (int32_t)j;
    if (opt.conj_order && opt.conj_order->size() == conj.size()) order = *opt.conj_order;
    // every conjunct a column-vs-literal comparison: range tests, no ops.
    // Bounds on one column (same domain) merge into one test; the selected
    // rows do not depend on the order in which conjuncts are applied.
    std::vector<DevRange> ranges;
    if (opt.ranges && !conj.empty()) {
      for (int32_t j : order) {
        DevRange rg;
        if (!as_range(L, s, prod, conj[(size_t)j], rg)) { ranges.clear(); break; }
        bool merged = false;
        if (!(rg.kind & RK_NE)) {
          for (DevRange& pr : ranges) {
            if (pr.col == rg.col && !(pr.kind & RK_NE) && (pr.kind & RK_F) == (rg.kind & RK_F) &&
                pr.mul == rg.mul) {
              intersect(pr, rg);
              merged = true;
              break;
            }
          }
        }
        if (!merged) ranges.push_back(rg);
      }
      if (ranges.size() > BSQL_DEV_MAX_RANGES) ranges.clear();
    }
    if (!ranges.empty()) {
      res.plan_ranges = ranges;
    } else if (conj.size() >= 2) {
      std::set<int> done;
      for (int32_t j : order) {
        std::set<int> sub;
        subtree(s, prod, conj[(size_t)j], sub);
        for (int o : sub)  // std::set iterates in program order
          if (!done.count(o)) { L.emit(to_lop(s.ops[(size_t)o])); done.insert(o); }
        LOp k;
        k.d.code = D_KILL;
        k.va = conj[(size_t)j];
        // truthiness follows the kind of the conjunct's value
        int o = prod[(size_t)k.va];
        int ty = o < 0 ? L.in_types[L.values[(size_t)k.va].column]
                 : s.ops[(size_t)o].prog_op >= 0 ? t.ot[(size_t)s.ops[(size_t)o].prog_op]
                 : (s.ops[(size_t)o].d.code == D_TO_F64 ? BSQL_FLOAT64 : BSQL_INT64);
        if (is_float_type(ty)) k.d.flags |= F_A_F;
        L.emit(k);
      }
    } else {
      const std::vector<char> need = needed_ops(s, prod, s.finals);
      for (size_t i = 0; i < s.ops.size(); ++i)
        if (need[i]) L.emit(to_lop(s.ops[i]));
      LOp k;
      k.d.code = D_SEL;
      k.va = s.finals[0];
      L.emit(k);
    }
  }
  const int split = (int)L.out.size();
  // The output phase runs over the selected rows only, packed densely: it
  // reloads what it reads instead of reusing selection-phase registers.
  if (opt.mode != LowerOptions::EXECUTE) L.load_value.assign((size_t)n_in, -1);

  // ---- phase 2: outputs
  std::vector<int> out_index;  // projection output written by launch output k
  if (opt.mode == LowerOptions::FILTER_COUNT) {
    if (n_out != 0) fail("a count has no output columns");
  } else if (opt.mode == LowerOptions::FILTER_GATHER) {
    // the first n_out input columns are gathered (the rest only feed the
    // predicate, e.g. an appended mask column)
    if (n_out > n_in) fail("a filter gathers input columns only");
    for (int k = 0; k < n_out; ++k) store_value.push_back(L.load_of(k));
  } else {
    const Program& proj = *opt.proj;
    const int n_roots = num_roots(proj);
    if (opt.out_subset) {
      out_index = *opt.out_subset;
    } else {
      for (int k = 0; k < n_roots; ++k) out_index.push_back(k);
    }
    if ((int)out_index.size() != n_out) fail("number of outputs does not match the program");
    SsaProgram s = to_ssa(L, proj, proj.seed);
    std::vector<int> finals;
    for (int k : out_index) {
      if (k < 0 || k >= n_roots) fail("output index out of range");
      finals.push_back(s.finals[(size_t)k]);
    }
    const std::vector<int> prod = producers(s, L.values.size());
    const std::vector<char> need = needed_ops(s, prod, finals);
    for (size_t i = 0; i < s.ops.size(); ++i)
      if (need[i]) L.emit(to_lop(s.ops[i]));
    store_value = finals;
  }
  res.out_str_col.assign((size_t)n_out, -1);
  for (int k = 0; k < n_out; ++k) {
    // a string value is stored as its slice; the chars are copied after the
    // kernel from the column the slice points into
    L.emit_value_load(store_value[(size_t)k]);
    const int sc = L.values[(size_t)store_value[(size_t)k]].str_col;
    if ((sc >= 0) != is_string_type(out_types[k]))
      fail(sc >= 0 ? "a string expression needs a STRING output column"
                   : "a STRING output column needs a string expression");
    res.out_str_col[(size_t)k] = sc;
    LOp st;
    st.d.code = D_STORE;
    st.va = store_value[(size_t)k];
    st.d.aux = (uint8_t)k;
    L.emit(st);
  }
  // Place each store right after the definition of its value, but never
  // inside the selection phase (compacted positions exist only after it).
  {
    std::vector<LOp> body, stores;
    for (const LOp& o : L.out) (o.d.code == D_STORE ? stores : body).push_back(o);
    std::vector<int> def_at(L.values.size(), -1);
    for (size_t i = 0; i < body.size(); ++i)
      if (body[i].vdst >= 0) def_at[(size_t)body[i].vdst] = (int)i;
    std::vector<std::vector<LOp>> after(body.size() + 1);
    for (const LOp& st : stores) {
      int at = std::max(def_at[(size_t)st.va], split - 1);
      after[(size_t)(at + 1)].push_back(st);
    }
    std::vector<LOp> final_list;
    for (size_t i = 0; i <= body.size(); ++i) {
      for (const LOp& st : after[i]) final_list.push_back(st);
      if (i < body.size()) final_list.push_back(body[i]);
    }
    L.out = final_list;
  }
  const int new_split = opt.mode == LowerOptions::EXECUTE ? 0 : split;

  // ---- load placement: each segment (the whole selection phase, or each
  // conjunct with lazy loads; the output phase) issues its column loads as
  // early as the registers allow, so that several are in flight together.
  {
    std::vector<int> bounds{0};
    if (opt.lazy_loads)
      for (int i = 0; i < new_split; ++i)
        if (L.out[(size_t)i].d.code == D_KILL) bounds.push_back(i + 1);
    if (new_split > 0 && bounds.back() != new_split) bounds.push_back(new_split);
    if (bounds.back() != (int)L.out.size()) bounds.push_back((int)L.out.size());
    L.out = place_loads(L.out, bounds, L.values.size(), budget);
  }

  // ---- registers, with spill slots beyond the budget
  int regs_used = 0, slots = 0;
  std::vector<LOp> alloc = assign_registers(L.out, L.values.size(), budget, &regs_used, &slots);
  // loads keep their column index in `a`
  for (LOp& o : alloc)
    if (o.d.code == D_LOAD || o.d.code == D_LOADS) o.d.a = (uint8_t)L.values[(size_t)o.vdst].column;

  // ---- dense input numbering: the distinct columns this launch reads
  {
    std::vector<int> cols;
    for (const LOp& o : alloc) op_columns(o.d, cols);
    for (const DevRange& rg : res.plan_ranges) cols.push_back(rg.col);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    if ((int)cols.size() > BSQL_DEV_MAX_IN) fail("a launch reads at most 128 distinct columns");
    std::vector<int> map((size_t)std::max(n_in, 1), -1);
    for (size_t j = 0; j < cols.size(); ++j) map[(size_t)cols[j]] = (int)j;
    for (LOp& o : alloc) remap_columns(o.d, map);
    for (DevRange& rg : res.plan_ranges) rg.col = map[(size_t)rg.col];
    res.in_cols = cols;
  }

  // runs of column loads (same segment) become one LOADN, which issues up to
  // four columns' loads together, and one COMMIT per column
  int new_split_final = new_split;
  {
    std::vector<LOp> merged;
    for (size_t i = 0; i < alloc.size();) {
      if (alloc[i].d.code != D_LOAD) { merged.push_back(alloc[i]); ++i; continue; }
      LOp issue;
      issue.d.code = D_LOADN;
      issue.d.dst = 0xff;
      issue.d.imm = 0;
      std::vector<LOp> commits;
      int n = 0;
      size_t j = i;
      while (j < alloc.size() && alloc[j].d.code == D_LOAD && n < 4) {
        issue.d.imm |= (int64_t)((uint64_t)alloc[j].d.a << (8 * n));
        LOp c;
        c.d.code = D_COMMIT;
        c.d.aux = (uint8_t)n;
        c.d.dst = alloc[j].d.dst;
        commits.push_back(c);
        ++n;
        ++j;
      }
      issue.d.aux = (uint8_t)n;
      merged.push_back(issue);
      for (const LOp& c : commits) merged.push_back(c);
      i = j;
    }
    alloc = merged;
    if (opt.mode != LowerOptions::EXECUTE) {
      // the selection phase ends after its last KILL/SEL
      int last_sel = -1;
      for (size_t i = 0; i < alloc.size(); ++i)
        if (alloc[i].d.code == D_KILL || alloc[i].d.code == D_SEL) last_sel = (int)i;
      new_split_final = last_sel + 1;
    }
  }

  if (alloc.size() > ((size_t)1 << 20)) fail("program too long for one launch");
  std::memset(&res.plan, 0, sizeof(res.plan));
  res.plan.n_ops = (int32_t)alloc.size();
  res.plan.split = new_split_final;
  res.plan.n_in = (int32_t)res.in_cols.size();
  res.plan.n_out = n_out;
  for (const LOp& o : alloc) res.ops.push_back(o.d);
  if (res.ops.empty()) res.ops.push_back(DevOp{});  // one NOP so that the list is never empty
  res.plan.n_ranges = (int32_t)res.plan_ranges.size();
  res.plan.range_eager = opt.lazy_loads ? 0 : 1;
  res.plan.n_spill = slots;
  for (size_t i = 0; i < res.plan_ranges.size(); ++i) res.plan.ranges[i] = res.plan_ranges[i];
  res.n_spill = slots;
  res.nreg = slots > 0 ? BSQL_DEV_MAX_REG : regs_used + 1;  // + the spare register
  if (slots > 0 && budget < BSQL_DEV_MAX_REG - 1) res.nreg = budget + 1;
  return res;
}

Program::~Program() = default;

}  // namespace bsql
