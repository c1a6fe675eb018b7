// parser.cpp -- Calcite expression text -> interpreter program.
//
// Host front-end of the interpreter path, following the reference:
//   replace_calcite_regex        engine/src/parser/expression_utils.cpp:1068-1113
//   expand_if_logical_op         engine/src/parser/CalciteExpressionParsing.cpp:330-398
//   fill_minus_op_with_zero      engine/src/parser/expression_utils.cpp:1155-1164
//   lexer / expr_parser          engine/src/parser/expression_tree.cpp:33-279
//   literal typing               engine/src/parser/expression_tree.cpp:281-375
//   custom_op_transformer        engine/src/parser/expression_tree.hpp:181-258
//   map_to_operator_type         engine/src/parser/expression_utils.cpp:268-352
//   get_scalar_from_string       engine/src/parser/CalciteExpressionParsing.cpp:138-262
//   used-column remap            engine/src/execution_kernels/LogicalProject.cpp:859-881
//   expr_to_plan_visitor         engine/src/Interpreter/interpreter_cpp.cu:274-438
//   get_named_expression         engine/src/parser/expression_utils.cpp:537-547
//   get_query_part               engine/src/parser/expression_utils.cpp:764-771
// The slot numbering is the reference's: inputs 0..n-1 (used columns in table
// order), final outputs n..n+e-1, temporaries from n+e with first-free reuse.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <mutex>
#include <map>
#include <memory>
#include <regex>
#include <string>
#include <vector>

#include "program.hpp"
#include "strfn.h"

namespace bsql {
namespace {

// ------------------------------------------------------------ text rewrites

void replace_all(std::string& data, const std::string& what, const std::string& with) {
  size_t pos = data.find(what);
  while (pos != std::string::npos) {
    data.replace(pos, what.size(), with);
    pos = data.find(what, pos + with.size());
  }
}

std::string replace_calcite_regex(const std::string& expression) {
  std::string ret = expression;
  static const std::regex char_collate_re{R"((?:\(\d+\))? CHARACTER SET ".+?" COLLATE ".+?")",
                                          std::regex_constants::icase};
  ret = std::regex_replace(ret, char_collate_re, "");
  static const std::regex timestamp_re{R"(TIMESTAMP\(\d+\))", std::regex_constants::icase};
  ret = std::regex_replace(ret, timestamp_re, "TIMESTAMP");
  static const std::regex decimal_re{R"(DECIMAL\(\d+, \d+\))", std::regex_constants::icase};
  ret = std::regex_replace(ret, decimal_re, "DOUBLE");
  static const std::regex char_re{R"(:CHAR\(\d+\))", std::regex_constants::icase};
  ret = std::regex_replace(ret, char_re, ":VARCHAR");
  replace_all(ret, "IS NOT DISTINCT FROM", "IS_NOT_DISTINCT_FROM");
  replace_all(ret, "IS NOT NULL", "IS_NOT_NULL");
  replace_all(ret, "IS NULL", "IS_NULL");
  replace_all(ret, " NOT NULL", "");
  replace_all(ret, "RAND()", "BLZ_RND()");
  replace_all(ret, "EXTRACT(FLAG(YEAR), ", "BL_YEAR(");
  replace_all(ret, "EXTRACT(FLAG(MONTH), ", "BL_MONTH(");
  replace_all(ret, "EXTRACT(FLAG(DAY), ", "BL_DAY(");
  replace_all(ret, "EXTRACT(FLAG(DOW), ", "BL_DOW(");
  replace_all(ret, "EXTRACT(FLAG(HOUR), ", "BL_HOUR(");
  replace_all(ret, "EXTRACT(FLAG(MINUTE), ", "BL_MINUTE(");
  replace_all(ret, "EXTRACT(FLAG(SECOND), ", "BL_SECOND(");
  // TRIM's side flag (expression_utils.cpp:1102-1105 rewrites it to a quoted
  // word); here the side goes into the function name
  replace_all(ret, "TRIM(FLAG(BOTH), ", "BLZ_TRIM_BOTH(");
  replace_all(ret, "TRIM(FLAG(LEADING), ", "BLZ_TRIM_LEADING(");
  replace_all(ret, "TRIM(FLAG(TRAILING), ", "BLZ_TRIM_TRAILING(");
  replace_all(ret, "DAY TO SECOND", "SECOND");
  replace_all(ret, "HOUR TO SECOND", "SECOND");
  replace_all(ret, "MINUTE TO SECOND", "SECOND");
  replace_all(ret, "/INT(", "/(");
  return ret;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// characters inside single-quoted literals ('' is an escaped quote)
std::vector<bool> quoted_vector(const std::string& s) {
  std::vector<bool> q(s.size(), false);
  bool in = false;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '\'') {
      if (in && i + 1 < s.size() && s[i + 1] == '\'') { q[i] = q[i + 1] = true; ++i; continue; }
      in = !in;
      q[i] = true;
      continue;
    }
    q[i] = in;
  }
  return q;
}

int find_not_in_quotes(const std::string& hay, const std::string& needle, size_t pos) {
  std::vector<bool> q = quoted_vector(hay);
  while (pos < hay.size()) {
    pos = hay.find(needle, pos);
    if (pos == std::string::npos) return -1;
    if (!q[pos]) return (int)pos;
    pos += needle.size();
  }
  return -1;
}

int find_closing(const std::string& e, int start) {
  const char open = e[(size_t)start];
  const char close = open == '(' ? ')' : (open == '[' ? ']' : 0);
  if (!close) return -1;
  int depth = 1;
  bool in_q = false;
  for (size_t i = (size_t)start + 1; i < e.size(); ++i) {
    if (in_q) {
      if (e[i] == '\'' && !(i + 1 < e.size() && e[i + 1] == '\'')) in_q = false;
    } else if (e[i] == '\'') {
      in_q = true;
    } else if (e[i] == open) {
      ++depth;
    } else if (e[i] == close) {
      if (--depth == 0) return (int)i;
    }
  }
  return -1;
}

}  // namespace

// splits a comma separated expression list at depth 0 (expression_utils.cpp:1012-1066)
std::vector<std::string> split_expression_list(const std::string& combined, bool do_trim) {
  std::vector<std::string> out;
  size_t cur = 0, start = 0;
  bool in_q = false;
  int paren = 0, sq = 0;
  while (cur < combined.size()) {
    const char c = combined[cur];
    if (in_q) {
      if (c == '\'' && !(cur + 1 < combined.size() && combined[cur + 1] == '\'')) in_q = false;
    } else if (c == '\'') {
      in_q = true;
    } else if (c == '(') {
      ++paren;
    } else if (c == ')') {
      --paren;
    } else if (c == '[') {
      ++sq;
    } else if (c == ']') {
      --sq;
    } else if (c == ',' && paren == 0 && sq == 0) {
      std::string e = combined.substr(start, cur - start);
      if (do_trim) {
        size_t b = 0;
        while (b < e.size() && std::isspace((unsigned char)e[b])) ++b;
        e = e.substr(b);
      }
      out.push_back(e);
      start = cur + 1;
    }
    ++cur;
  }
  if (start < combined.size()) {
    std::string e = combined.substr(start);
    out.push_back(do_trim ? trim(e) : e);
  }
  return out;
}

namespace {

// N-ary AND/OR -> left-deep binary chain in plan order
std::string expand_if_logical_op(std::string expression) {
  std::string output = expression;
  size_t start_pos = 0;
  while (start_pos < expression.size()) {
    const int first_and = find_not_in_quotes(expression, "AND(", start_pos);
    int first_or = -1;
    if (expression.find("FLOOR") == std::string::npos)
      first_or = find_not_in_quotes(expression, "OR(", start_pos);
    int first;
    std::string op;
    if (first_and >= 0) {
      if (first_or >= 0 && first_or < first_and) { first = first_or; op = "OR("; }
      else { first = first_and; op = "AND("; }
    } else {
      first = first_or;
      op = "OR(";
    }
    if (first < 0) return output;
    const int ex_start = first + (int)op.size() - 1;
    const int ex_end = find_closing(expression, ex_start);
    if (ex_end < 0) fail("unbalanced parentheses in expression");
    const std::string rest = expression.substr((size_t)ex_start + 1, (size_t)(ex_end - ex_start - 1));
    std::vector<std::string> parts = split_expression_list(rest, false);
    if (parts.size() == 2) {
      start_pos = (size_t)ex_start;
      continue;
    }
    if (parts.size() < 2) fail("logical operator with fewer than two operands");
    start_pos = (size_t)first;
    output = expression.substr(0, (size_t)first);
    for (size_t i = 0; i + 1 < parts.size(); ++i) {
      output += op;
      start_pos += op.size();
    }
    output += parts[0] + ",";
    for (size_t i = 1; i + 1 < parts.size(); ++i) output += parts[i] + "),";
    output += parts.back() + ")";
    if ((size_t)ex_end < expression.size() - 1) output += expression.substr((size_t)ex_end + 1);
    expression = output;
  }
  return output;
}

std::string fill_minus_op_with_zero(std::string e) {
  const size_t vars = (size_t)std::count(e.begin(), e.end(), '$');
  if (!e.empty() && e[0] == '-' && vars == 1 && e.find(',') == std::string::npos)
    e = e.substr(0, 2) + "0, " + e.substr(2);
  return e;
}

std::string replaced(std::string s, const std::string& what, const std::string& with) {
  replace_all(s, what, with);
  return s;
}

// n-ary CONCAT(a, b, c, ...) -> CONCAT(CONCAT(CONCAT(a, b), c), ...), each
// CONCAT found left to right with its argument list taken up to the bracket
// that closes it (quotes are not considered there, as in the reference);
// arguments split at commas outside quotes and brackets and rejoined with
// ',' (nary_to_nestedbinary, expression_utils.cpp:1520-1622). A CONCAT inside
// another's arguments is left as written.
std::string nested_concat(const std::string& e) {
  const std::string name = "CONCAT";
  std::string outp;
  size_t from = 0;
  size_t at = e.find(name);
  while (at != std::string::npos) {
    const size_t open = at + name.size();
    // the position after the bracket closing the one at `open`
    size_t end = open;
    int depth = 0;
    do {
      if (end >= e.size()) fail("end expression reached");
      if (e[end] == '(') ++depth;
      else if (e[end] == ')') --depth;
      ++end;
    } while (end < e.size() && depth != 0);
    if (depth != 0) fail("end expression reached");
    outp += e.substr(from, at - from);
    const std::string args = e.substr(open + 1, end - open - 2);
    std::vector<std::string> parts;
    std::string cur;
    int br = 0;
    bool in_str = false;
    for (char c : args) {
      if (c == '(') ++br;
      else if (c == ')') --br;
      else if (c == '\'') in_str = !in_str;
      if (!in_str && br == 0 && c == ',') {
        parts.push_back(cur);
        cur.clear();
      } else {
        cur += c;
      }
    }
    if (!cur.empty()) parts.push_back(cur);
    if (parts.size() < 2) fail("CONCAT operator must have at least two children, as CONCAT($0, $1) .");
    for (size_t k = 0; k + 1 < parts.size(); ++k) outp += name + "(";
    outp += parts[0];
    for (size_t k = 1; k < parts.size(); ++k) outp += "," + parts[k] + ")";
    from = end;
    at = e.find(name, end);
  }
  return outp + e.substr(from);
}

// CURRENT_DATE / CURRENT_TIMESTAMP / CURRENT_TIME -> the query's timestamp
// (get_current_date_or_timestamp, expression_utils.cpp:1494-1518).
std::string g_current_timestamp;  // "YYYY-MM-DD hh:mm:ss.mmm"; empty: the clock
std::mutex g_current_mu;

std::string current_timestamp_text() {
  std::lock_guard<std::mutex> lk(g_current_mu);
  if (!g_current_timestamp.empty()) return g_current_timestamp;
  const auto now = std::chrono::system_clock::now();
  const int64_t ms = std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count();
  const std::time_t secs = (std::time_t)(ms / 1000);
  std::tm tm{};
  gmtime_r(&secs, &tm);
  char buf[96];
  std::snprintf(buf, sizeof(buf), "%04d-%02d-%02d %02d:%02d:%02d.%03d", tm.tm_year + 1900, tm.tm_mon + 1,
                tm.tm_mday, tm.tm_hour, tm.tm_min, tm.tm_sec, (int)(ms % 1000));
  return buf;
}

std::string current_date_or_timestamp(std::string e) {
  if (e.find("CURRENT_TIME") != std::string::npos && e.find("CURRENT_TIMESTAMP") == std::string::npos)
    e = replaced(e, "CURRENT_TIME", "CURRENT_TIMESTAMP");
  const bool date = e.find("CURRENT_DATE") != std::string::npos;
  if (!date && e.find("CURRENT_TIMESTAMP") == std::string::npos) return e;
  std::string ts = current_timestamp_text().substr(0, 23);
  if (date) return replaced(e, "CURRENT_DATE", ts.substr(0, 10));
  return replaced(e, "CURRENT_TIMESTAMP", ts);
}

// Calcite states timestamp differences and interval additions in
// milliseconds; the interpreter's timestamps here are nanoseconds, so the
// unit factors gain six zeros (convert_ms_to_ns_units,
// expression_utils.cpp:1316-1350).
std::string ms_to_ns_units(const std::string& e) {
  const bool cast_ts = e.find("CAST(") != std::string::npos && e.find(":TIMESTAMP") != std::string::npos;
  const bool cast_date = e.find("CAST(") != std::string::npos && e.find(":DATE") != std::string::npos;
  if (!cast_ts && !cast_date) return e;
  const std::string zeros = "000000";
  if (e.find("Reinterpret(") != std::string::npos) {
    for (const char* unit : {"86400000", "3600000", "60000", "1000"})
      if (e.find(unit) != std::string::npos) return replaced(e, unit, std::string(unit) + zeros);
  }
  const size_t iv = e.find(":INTERVAL");
  if (iv != std::string::npos) {
    const size_t start = e.find(", ") + 2;
    const std::string value = e.substr(start, iv - start);
    return replaced(e, value, value + zeros);
  }
  return e;
}

// A timestamp literal token (date with '-' or '/', a time, an optional
// fraction of 3, 6 or 9 digits; expression_utils.cpp:401-467).
bool is_timestamp_token(const std::string& t) {
  static const std::regex re{
      R"(^[0-9]{4}([-/])[0-9]{2}\1([0-9]{2})?[ T]?[0-9]{2}:[0-9]{2}:([0-9]{2})?(.[0-9]{3}|.[0-9]{6}|.[0-9]{9})?[Z]?$)"};
  return std::regex_match(t, re);
}

// Column index of "$k" or of the first "$k" inside e.g. "CAST($k):TIMESTAMP".
int column_of_operand(const std::string& e) {
  if (!e.empty() && e[0] == '$') return std::stoi(e.substr(1));
  const size_t a = e.find('$') + 1;
  const size_t b = e.find(')');
  return std::stoi(e.substr(a, b - a));
}

// Reinterpret(-(<timestamp literal>, $k)): the literal takes the unit of
// column k (reinterpret_timestamp, expression_utils.cpp:1371-1421).
std::string reinterpret_timestamp(std::string e, const int32_t* types, int n) {
  if (n <= 0) return e;
  const std::string key = "Reinterpret(-(";
  const size_t at = e.find(key);
  if (at == std::string::npos) return e;
  std::string rest = e.substr(at + key.size());
  rest = rest.substr(0, rest.find("))"));
  const std::vector<std::string> ops = split_expression_list(rest, true);
  if (ops.size() != 2) return e;
  std::string ts;
  int col;
  if (is_timestamp_token(ops[0])) {
    ts = ops[0];
    col = column_of_operand(ops[1]);
  } else if (is_timestamp_token(ops[1])) {
    ts = ops[1];
    col = column_of_operand(ops[0]);
  } else {
    return e;
  }
  if (col < 0 || col >= n) fail("Reinterpret refers to column $" + std::to_string(col) + " outside the table");
  const char* unit = types[col] == BSQL_TIMESTAMP_SECONDS        ? ":TIMESTAMP_SECONDS"
                     : types[col] == BSQL_TIMESTAMP_MILLISECONDS ? ":TIMESTAMP_MILLISECONDS"
                     : types[col] == BSQL_TIMESTAMP_MICROSECONDS ? ":TIMESTAMP_MICROSECONDS"
                                                                 : ":TIMESTAMP";
  return replaced(e, ts, "CAST(" + ts + ")" + unit);
}

// Interval literals come in milliseconds; they take the unit of the
// duration column they are added to (apply_interval_conversion,
// expression_utils.cpp:1425-1469).
std::string interval_conversion(const std::string& e, const int32_t* types, int n) {
  const std::string iv = ":INTERVAL";
  if (n <= 0 || e.find(iv) == std::string::npos) return e;
  if (e.find('+') != std::string::npos || e.find('-') != std::string::npos) {
    size_t start = 0;
    if (e.find("+($") != std::string::npos) start = e.find("+($") + 3;
    else if (e.find("-($") != std::string::npos) start = e.find("-($") + 3;
    std::string idx = e.substr(start);
    idx = idx.substr(0, idx.find(", "));
    int col = -1;
    try {
      col = std::stoi(idx);
    } catch (const std::exception&) {
      fail("interval arithmetic without a column operand: " + e);
    }
    if (col < 0 || col >= n) fail("interval arithmetic on column $" + std::to_string(col) + " outside the table");
    if (types[col] == BSQL_DURATION_SECONDS) return replaced(e, "000" + iv, iv);
    if (types[col] == BSQL_DURATION_MICROSECONDS) return replaced(e, "000" + iv, "000000" + iv);
    if (types[col] == BSQL_DURATION_NANOSECONDS) return replaced(e, "000" + iv, "000000000" + iv);
    return e;
  }
  if (e.find('$') == std::string::npos && e.find("000:INTERVAL") != std::string::npos)
    return replaced(e, "000" + iv, iv);
  return e;
}

// ------------------------------------------------------------------- lexer

enum class Tok {
  Open, Close, Comma, Colon, Variable, Null, Boolean, Number, Duration_s, Duration_ms,
  Duration_us, Duration_ns, Timestamp_d, Timestamp_s, Timestamp_ms, Timestamp_us, Timestamp_ns,
  String, Identifier, Eof
};

struct Token {
  Tok type;
  std::string value;
};

class Lexer {
 public:
  explicit Lexer(const std::string& s) : text_(trim(s)), pos_(0) {}

  Token next() {
    if (pos_ >= text_.size()) return {Tok::Eof, ""};
    while (pos_ < text_.size() && text_[pos_] == ' ') ++pos_;
    if (pos_ >= text_.size()) return {Tok::Eof, ""};
    const char c = text_[pos_];
    if (c == '(') { ++pos_; return {Tok::Open, "("}; }
    if (c == ')') { ++pos_; return {Tok::Close, ")"}; }
    if (c == ',') { ++pos_; return {Tok::Comma, ","}; }
    if (c == ':') { ++pos_; return {Tok::Colon, ":"}; }
    static const std::regex variable_re{R"(^\$\d+)"};
    static const std::regex null_re{R"(^null)"};
    static const std::regex boolean_re{R"(^(true|false))"};
    static const std::regex number_re{R"(^[-+]?\d*\.?\d+([eE][-+]?\d+)?)"};
    static const std::regex dur_s_re{R"(^\d{2}:\d{2}:\d{2})"};
    static const std::regex dur_ms_re{R"(^\d{2}:\d{2}:\d{2}.\d{3})"};
    static const std::regex ts_d_re{R"(^\d{4}-\d{2}-\d{2})"};
    static const std::regex ts_s_re{R"(^\d{4}-\d{2}-\d{2}(?:[ T]?\d{2}:\d{2}:\d{2}))"};
    static const std::regex ts_ms_re{R"(^\d{4}-\d{2}-\d{2}(?:[ T]?\d{2}:\d{2}:\d{2}.\d{3}))"};
    static const std::regex ts_us_re{R"(^\d{4}-\d{2}-\d{2}(?:[ T]?\d{2}:\d{2}:\d{2}.\d{6}))"};
    static const std::regex string_re{R"(^(["'])(?:(?!\1|\\).|\\.)*?\1)"};
    const std::string rem = text_.substr(pos_);
    std::smatch m;
    auto take = [&](const std::regex& re, Tok t, Token* out) {
      if (!std::regex_search(rem, m, re)) return false;
      pos_ += (size_t)m.length(0);
      *out = {t, m.str(0)};
      return true;
    };
    Token t;
    // same order and patterns as the reference lexer, including its reuse of
    // the millisecond duration pattern for us/ns and of the us timestamp
    // pattern for ns (expression_tree.hpp:319-332)
    if (take(variable_re, Tok::Variable, &t)) return t;
    if (take(null_re, Tok::Null, &t)) return t;
    if (take(boolean_re, Tok::Boolean, &t)) return t;
    if (take(dur_s_re, Tok::Duration_s, &t)) return t;
    if (take(dur_ms_re, Tok::Duration_ms, &t)) return t;
    if (take(dur_ms_re, Tok::Duration_us, &t)) return t;
    if (take(dur_ms_re, Tok::Duration_ms, &t)) return t;
    if (take(ts_us_re, Tok::Timestamp_ns, &t)) return t;
    if (take(ts_us_re, Tok::Timestamp_us, &t)) return t;
    if (take(ts_ms_re, Tok::Timestamp_ms, &t)) return t;
    if (take(ts_s_re, Tok::Timestamp_s, &t)) return t;
    if (take(ts_d_re, Tok::Timestamp_d, &t)) return t;
    if (take(number_re, Tok::Number, &t)) return t;
    if (take(string_re, Tok::String, &t)) return t;
    size_t len = 0;
    do {
      const char ch = text_[pos_ + len];
      if (ch == '(' || ch == ')' || ch == ',' || ch == ':') break;
    } while (pos_ + (++len) < text_.size());
    std::string v = text_.substr(pos_, len);
    pos_ += len;
    return {Tok::Identifier, v};
  }

 private:
  std::string text_;
  size_t pos_;
};

// ------------------------------------------------------------------- tree

enum class NodeType { Operator, Literal, Variable };

struct Node {
  NodeType type;
  std::string value;
  int32_t lit_type = BSQL_EMPTY;  // literals
  int var_index = -1;             // variables
  std::vector<std::unique_ptr<Node>> children;
};

int32_t infer_literal_type(const Token& t) {
  switch (t.type) {
    case Tok::Null: return BSQL_EMPTY;
    case Tok::Boolean: return BSQL_BOOL8;
    case Tok::Number:
      if (t.value.find_first_of(".eE") != std::string::npos) {
        const double d = std::stod(t.value);
        return d == (double)(float)d ? BSQL_FLOAT32 : BSQL_FLOAT64;
      }
      return std::stoll(t.value) > INT_MAX ? BSQL_INT64 : BSQL_INT32;
    case Tok::Duration_s: return BSQL_DURATION_SECONDS;
    case Tok::Duration_ms: return BSQL_DURATION_MILLISECONDS;
    case Tok::Duration_us: return BSQL_DURATION_MICROSECONDS;
    case Tok::Duration_ns: return BSQL_DURATION_NANOSECONDS;
    case Tok::Timestamp_ns: return BSQL_TIMESTAMP_NANOSECONDS;
    case Tok::Timestamp_us: return BSQL_TIMESTAMP_MICROSECONDS;
    case Tok::Timestamp_ms: return BSQL_TIMESTAMP_MILLISECONDS;
    case Tok::Timestamp_s: return BSQL_TIMESTAMP_SECONDS;
    case Tok::Timestamp_d: return BSQL_TIMESTAMP_DAYS;
    default: return BSQL_STRING;
  }
}

int32_t type_from_type_token(const std::string& v) {
  if (v == "NULL" || v == "BOOLEAN") return BSQL_BOOL8;
  if (v == "TINYINT") return BSQL_INT8;
  if (v == "SMALLINT") return BSQL_INT16;
  if (v == "INTEGER") return BSQL_INT32;
  if (v == "INTERVAL SECOND" || v == "INTERVAL MINUTE" || v == "INTERVAL HOUR" ||
      v == "INTERVAL DAY")
    return BSQL_DURATION_SECONDS;
  if (v == "INTERVAL MONTH" || v == "INTERVAL YEAR")
    fail("operation over MONTH or YEAR units are not currently supported.");
  if (v == "BIGINT") return BSQL_INT64;
  if (v == "FLOAT") return BSQL_FLOAT32;
  if (v == "DOUBLE") return BSQL_FLOAT64;
  if (v == "DATE") return BSQL_TIMESTAMP_DAYS;
  if (v == "TIMESTAMP_SECONDS") return BSQL_TIMESTAMP_SECONDS;
  if (v == "TIMESTAMP_MILLISECONDS") return BSQL_TIMESTAMP_MILLISECONDS;
  if (v == "TIMESTAMP_MICROSECONDS") return BSQL_TIMESTAMP_MICROSECONDS;
  if (v == "TIMESTAMP") return BSQL_TIMESTAMP_NANOSECONDS;
  if (v == "VARCHAR") return BSQL_STRING;
  fail("Invalid literal cast type");
}

class Parser {
 public:
  explicit Parser(const std::string& s) : lex_(s) {}
  std::unique_ptr<Node> parse() {
    tok_ = lex_.next();
    auto n = expr();
    return n;
  }

 private:
  bool accept(Tok t) {
    if (tok_.type == t) { tok_ = lex_.next(); return true; }
    return false;
  }
  std::unique_ptr<Node> expr() {
    auto r = term();
    if (!r) r = func();
    if (!r) fail("Couldn't parse calcite expression");
    return r;
  }
  std::unique_ptr<Node> term() {
    Token v = tok_;
    if (accept(Tok::Variable)) {
      auto n = std::make_unique<Node>();
      n->type = NodeType::Variable;
      n->value = v.value;
      n->var_index = std::stoi(v.value.substr(1));
      return n;
    }
    return literal();
  }
  std::unique_ptr<Node> func() {
    Token name = tok_;
    if (!accept(Tok::Identifier)) return nullptr;
    accept(Tok::Open);
    std::vector<std::unique_ptr<Node>> args;
    if (!accept(Tok::Close)) {
      args.push_back(expr());
      while (accept(Tok::Comma)) args.push_back(expr());
      accept(Tok::Close);
    }
    std::string id = name.value;
    if (accept(Tok::Colon)) {
      Token rt = tok_;
      accept(Tok::Identifier);
      id += "_" + rt.value;
    }
    auto n = std::make_unique<Node>();
    n->type = NodeType::Operator;
    n->value = id;
    n->children = std::move(args);
    return n;
  }
  std::unique_ptr<Node> literal() {
    Token lit = tok_;
    const Tok t = tok_.type;
    if (t == Tok::Null || t == Tok::Boolean || t == Tok::Number || t == Tok::Duration_s ||
        t == Tok::Duration_ms || t == Tok::Duration_us || t == Tok::Duration_ns ||
        t == Tok::Timestamp_d || t == Tok::Timestamp_s || t == Tok::Timestamp_ms ||
        t == Tok::Timestamp_us || t == Tok::Timestamp_ns || t == Tok::String) {
      accept(t);
      auto n = std::make_unique<Node>();
      n->type = NodeType::Literal;
      n->value = lit.value;
      if (accept(Tok::Colon)) {
        Token tt = tok_;
        accept(Tok::Identifier);
        n->lit_type = type_from_type_token(tt.value);
      } else {
        n->lit_type = infer_literal_type(lit);
      }
      return n;
    }
    return nullptr;
  }
  Lexer lex_;
  Token tok_;
};

// ------------------------------------------------------------- operators

int map_to_operator(const std::string& tok) {
  static const std::map<std::string, int> M = {
      {"BLZ_RND", BSQL_OP_RAND},
      {"NOT", BSQL_OP_NOT}, {"IS NOT TRUE", BSQL_OP_IS_NOT_TRUE},
      {"IS NOT FALSE", BSQL_OP_IS_NOT_FALSE}, {"IS TRUE", BSQL_OP_IS_TRUE},
      {"SIN", BSQL_OP_SIN}, {"ASIN", BSQL_OP_ASIN}, {"COS", BSQL_OP_COS}, {"ACOS", BSQL_OP_ACOS},
      {"TAN", BSQL_OP_TAN}, {"ATAN", BSQL_OP_ATAN}, {"FLOOR", BSQL_OP_FLOOR},
      {"CEIL", BSQL_OP_CEIL}, {"ABS", BSQL_OP_ABS}, {"LOG10", BSQL_OP_LOG}, {"LN", BSQL_OP_LN},
      {"BL_YEAR", BSQL_OP_YEAR}, {"BL_MONTH", BSQL_OP_MONTH}, {"BL_DAY", BSQL_OP_DAY},
      {"BL_DOW", BSQL_OP_DAYOFWEEK}, {"BL_HOUR", BSQL_OP_HOUR}, {"BL_MINUTE", BSQL_OP_MINUTE},
      {"BL_SECOND", BSQL_OP_SECOND}, {"IS_NULL", BSQL_OP_IS_NULL},
      {"IS_NOT_NULL", BSQL_OP_IS_NOT_NULL}, {"CAST_TINYINT", BSQL_OP_CAST_TINYINT},
      {"CAST_SMALLINT", BSQL_OP_CAST_SMALLINT}, {"CAST_INTEGER", BSQL_OP_CAST_INTEGER},
      {"CAST_BIGINT", BSQL_OP_CAST_BIGINT}, {"CAST_FLOAT", BSQL_OP_CAST_FLOAT},
      {"CAST_DOUBLE", BSQL_OP_CAST_DOUBLE}, {"CAST_DATE", BSQL_OP_CAST_DATE},
      {"CAST_TIMESTAMP_SECONDS", BSQL_OP_CAST_TIMESTAMP_SECONDS},
      {"CAST_TIMESTAMP_MILLISECONDS", BSQL_OP_CAST_TIMESTAMP_MILLISECONDS},
      {"CAST_TIMESTAMP_MICROSECONDS", BSQL_OP_CAST_TIMESTAMP_MICROSECONDS},
      {"CAST_TIMESTAMP", BSQL_OP_CAST_TIMESTAMP}, {"CAST_VARCHAR", BSQL_OP_CAST_VARCHAR},
      {"CAST_CHAR", BSQL_OP_CAST_VARCHAR}, {"CHAR_LENGTH", BSQL_OP_CHAR_LENGTH},
      {"LOWER", BSQL_OP_STR_LOWER}, {"UPPER", BSQL_OP_STR_UPPER},
      {"INITCAP", BSQL_OP_STR_INITCAP}, {"REVERSE", BSQL_OP_STR_REVERSE},
      {"=", BSQL_OP_EQUAL}, {"<>", BSQL_OP_NOT_EQUAL}, {">", BSQL_OP_GREATER},
      {">=", BSQL_OP_GREATER_EQUAL}, {"<", BSQL_OP_LESS}, {"<=", BSQL_OP_LESS_EQUAL},
      {"+", BSQL_OP_ADD}, {"-", BSQL_OP_SUB}, {"*", BSQL_OP_MUL}, {"/", BSQL_OP_DIV},
      {"POWER", BSQL_OP_POW}, {"ROUND", BSQL_OP_ROUND}, {"MOD", BSQL_OP_MOD},
      {"AND", BSQL_OP_LOGICAL_AND}, {"OR", BSQL_OP_LOGICAL_OR},
      {"FIRST_NON_MAGIC", BSQL_OP_FIRST_NON_MAGIC}, {"MAGIC_IF_NOT", BSQL_OP_MAGIC_IF_NOT},
      {"LIKE", BSQL_OP_STR_LIKE}, {"SUBSTRING", BSQL_OP_STR_SUBSTRING},
      {"REGEXP_REPLACE", BSQL_OP_STR_REGEXP_REPLACE}, {"REPLACE", BSQL_OP_STR_REPLACE},
      {"TO_DATE", BSQL_OP_TO_DATE}, {"TO_TIMESTAMP", BSQL_OP_TO_TIMESTAMP},
      {"||", BSQL_OP_STR_CONCAT}, {"CONCAT", BSQL_OP_STR_CONCAT}, {"TRIM", BSQL_OP_STR_TRIM},
      {"BLZ_TRIM_BOTH", BSQL_OP_STR_TRIM}, {"BLZ_TRIM_LEADING", BSQL_OP_STR_TRIM},
      {"BLZ_TRIM_TRAILING", BSQL_OP_STR_TRIM},
      {"LEFT", BSQL_OP_STR_LEFT}, {"RIGHT", BSQL_OP_STR_RIGHT},
      {"IS_NOT_DISTINCT_FROM", BSQL_OP_IS_NOT_DISTINCT_FROM},
  };
  auto it = M.find(tok);
  if (it == M.end()) fail("Unsupported operator: " + tok);
  return it->second;
}

// CASE / Reinterpret / ROUND / CAST(literal) rewrites, bottom-up
std::unique_ptr<Node> transform_custom(std::unique_ptr<Node> n) {
  if (n->type != NodeType::Operator) return n;
  for (auto& c : n->children) c = transform_custom(std::move(c));
  if (n->value == "CASE") {
    if (n->children.size() < 3 || n->children.size() % 2 == 0) fail("malformed CASE");
    // CASE(c1, v1, c2, v2, ..., e) -> FIRST_NON_MAGIC(MAGIC_IF_NOT(c1, v1), ...)
    std::unique_ptr<Node> acc = std::move(n->children.back());
    for (int i = (int)n->children.size() - 3; i >= 0; i -= 2) {
      auto mif = std::make_unique<Node>();
      mif->type = NodeType::Operator;
      mif->value = "MAGIC_IF_NOT";
      mif->children.push_back(std::move(n->children[(size_t)i]));
      mif->children.push_back(std::move(n->children[(size_t)i + 1]));
      auto fnm = std::make_unique<Node>();
      fnm->type = NodeType::Operator;
      fnm->value = "FIRST_NON_MAGIC";
      fnm->children.push_back(std::move(mif));
      fnm->children.push_back(std::move(acc));
      acc = std::move(fnm);
    }
    return acc;
  }
  if (n->value == "Reinterpret") {
    if (n->children.size() != 1) fail("malformed Reinterpret");
    return std::move(n->children[0]);
  }
  if (n->value == "ROUND") {
    if (n->children.size() == 1) {
      auto z = std::make_unique<Node>();
      z->type = NodeType::Literal;
      z->value = "0";
      z->lit_type = BSQL_INT8;
      n->children.push_back(std::move(z));
    }
    return n;
  }
  if (n->value.rfind("CAST", 0) == 0) {
    if (n->children.size() != 1) fail("malformed CAST");
    const int cast_op = map_to_operator(n->value);
    if (n->children[0]->type == NodeType::Literal) {
      std::unique_ptr<Node> lit = std::move(n->children[0]);
      lit->lit_type = output_type_unary(cast_op, lit->lit_type);
      return lit;
    }
  }
  return n;
}

// ---------------------------------------------------------- literal values

bool is_null_literal(const std::string& v) { return v == "null"; }

int64_t days_from_civil(int64_t y, unsigned m, unsigned d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const unsigned yoe = (unsigned)(y - era * 400);
  const unsigned doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + (int64_t)doe - 719468;
}

// "%Y-%m-%d[ %H:%M:%S[.%f]]" (or with '/') -> ticks of the timestamp type
int64_t parse_timestamp(const std::string& raw, int32_t type) {
  std::string s = raw;
  if (!s.empty() && s.front() == '\'') s = s.substr(1);
  if (!s.empty() && s.back() == '\'') s.pop_back();
  static const std::regex re{
      R"(^(\d{4})[-/](\d{2})[-/](\d{2})(?:[ T](\d{2}):(\d{2}):(\d{2})(?:\.(\d{1,9}))?)?Z?$)"};
  std::smatch m;
  if (!std::regex_match(s, m, re)) fail("not recognized timestamp literal: " + raw);
  const int64_t days = days_from_civil(std::stoll(m[1]), (unsigned)std::stoi(m[2]),
                                       (unsigned)std::stoi(m[3]));
  int64_t secs = 0, nanos = 0;
  if (m[4].matched)
    secs = std::stoll(m[4]) * 3600 + std::stoll(m[5]) * 60 + std::stoll(m[6]);
  if (m[7].matched) {
    std::string f = m[7].str();
    f.resize(9, '0');
    nanos = std::stoll(f);
  }
  switch (type) {
    case BSQL_TIMESTAMP_DAYS: return days;
    case BSQL_TIMESTAMP_SECONDS: return days * 86400 + secs;
    case BSQL_TIMESTAMP_MILLISECONDS: return (days * 86400 + secs) * 1000 + nanos / 1000000;
    case BSQL_TIMESTAMP_MICROSECONDS: return (days * 86400 + secs) * 1000000 + nanos / 1000;
    default: return (days * 86400 + secs) * 1000000000LL + nanos;
  }
}

Scalar scalar_from_string(const std::string& v, int32_t type) {
  Scalar s;
  s.type_id = type;
  s.is_null = false;
  switch (type) {
    case BSQL_BOOL8: s.i64 = v == "true"; break;
    case BSQL_INT8: s.i64 = (int8_t)std::stoi(v); break;
    case BSQL_INT16: s.i64 = (int16_t)std::stoi(v); break;
    case BSQL_INT32: s.i64 = (int32_t)std::stoi(v); break;
    case BSQL_INT64: s.i64 = (int64_t)std::stoll(v); break;
    case BSQL_UINT8: s.i64 = (uint8_t)std::stoul(v); break;
    case BSQL_UINT16: s.i64 = (uint16_t)std::stoul(v); break;
    case BSQL_UINT32: s.i64 = (uint32_t)std::stoul(v); break;
    case BSQL_UINT64: s.i64 = (int64_t)std::stoull(v); break;
    case BSQL_FLOAT32: s.f64 = (double)std::stof(v); break;
    case BSQL_FLOAT64: s.f64 = std::stod(v); break;
    case BSQL_STRING: s.str = v.size() >= 2 ? v.substr(1, v.size() - 2) : v; break;
    case BSQL_DURATION_SECONDS: case BSQL_DURATION_MILLISECONDS:
    case BSQL_DURATION_MICROSECONDS: case BSQL_DURATION_NANOSECONDS:
      s.i64 = std::stoll(v); break;
    case BSQL_TIMESTAMP_DAYS: case BSQL_TIMESTAMP_SECONDS: case BSQL_TIMESTAMP_MILLISECONDS:
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_TIMESTAMP_NANOSECONDS:
      s.i64 = parse_timestamp(v, type); break;
    default: fail("unsupported literal type " + std::to_string(type));
  }
  return s;
}

// ---------------------------------------------------------------- encoder

// TRIM's side: 1 leading, 2 trailing, 3 both (0: not a side-flagged TRIM).
int trim_side(const std::string& v) {
  if (v == "BLZ_TRIM_LEADING") return 1;
  if (v == "BLZ_TRIM_TRAILING") return 2;
  if (v == "BLZ_TRIM_BOTH") return 3;
  return 0;
}

struct Encoder {
  Program& p;
  const std::map<int, int>& col_map;
  int start;
  std::vector<bool> free_;
  std::map<const Node*, int> pos_;

  Encoder(Program& prog, const std::map<int, int>& m, int start_pos)
      : p(prog), col_map(m), start(start_pos), free_(512, true) {
    std::fill_n(free_.begin(), start_pos, false);
  }

  int first_free() {
    for (size_t i = (size_t)start; i < free_.size(); ++i)
      if (free_[i]) { free_[i] = false; return (int)i; }
    return -1;
  }

  void push(int l, int r, const Scalar& ls, const Scalar& rs, int op) {
    p.left.push_back((int16_t)l);
    p.right.push_back((int16_t)r);
    p.lsc.push_back(ls);
    p.rsc.push_back(rs);
    p.ops.push_back(op);
  }

  void visit(const Node* n) {
    for (auto& c : n->children) visit(c.get());
    if (n->type == NodeType::Literal) { pos_[n] = BSQL_SCALAR_INDEX; return; }
    if (n->type == NodeType::Variable) {
      auto it = col_map.find(n->var_index);
      if (it == col_map.end()) fail("column $" + std::to_string(n->var_index) + " not in the table");
      pos_[n] = it->second;
      return;
    }
    const int op = map_to_operator(n->value);
    const Scalar none;
    auto literal_scalar = [&](const Node* lit, int* code) {
      if (is_null_literal(lit->value)) { *code = BSQL_SCALAR_NULL_INDEX; return Scalar(); }
      *code = BSQL_SCALAR_INDEX;
      return scalar_from_string(lit->value, lit->lit_type);
    };
    if (is_binary_op(op)) {
      if (n->children.size() != 2) fail("operator " + n->value + " expects two operands");
      const Node* L = n->children[0].get();
      const Node* R = n->children[1].get();
      const int lp = pos_.at(L), rp = pos_.at(R);
      if (L->type != NodeType::Literal && lp >= start) free_[(size_t)lp] = true;
      if (R->type != NodeType::Literal && rp >= start) free_[(size_t)rp] = true;
      if (L->type == NodeType::Literal && R->type == NodeType::Literal)
        fail("Operations between literals is not supported");
      if (L->type == NodeType::Literal) {
        int code;
        Scalar s = literal_scalar(L, &code);
        push(code, rp, s, none, op);
      } else if (R->type == NodeType::Literal) {
        int code;
        Scalar s = literal_scalar(R, &code);
        if (op == BSQL_OP_STR_TRIM) s.i64 = trim_side(n->value);  // with the characters
        push(lp, code, none, s, op);
      } else {
        push(lp, rp, none, none, op);
      }
    } else if (is_unary_op(op)) {
      if (n->children.size() != 1) fail("operator " + n->value + " expects one operand");
      const Node* L = n->children[0].get();
      if (L->type == NodeType::Literal) fail("Unary operations on literals is not supported");
      const int lp = pos_.at(L);
      if (lp >= start) free_[(size_t)lp] = true;
      push(lp, BSQL_UNARY_INDEX, none, none, op);
    } else {
      if (!n->children.empty()) fail("operator " + n->value + " takes no operands");
      push(BSQL_NULLARY_INDEX, BSQL_NULLARY_INDEX, none, none, op);
    }
    const int out = first_free();
    pos_[n] = out;
    p.out.push_back((int16_t)out);
  }
};

void collect_vars(const Node* n, std::vector<int>& acc) {
  if (n->type == NodeType::Variable) acc.push_back(n->var_index);
  for (auto& c : n->children) collect_vars(c.get(), acc);
}

// String functions this build does not evaluate (the reference computes them
// with libcudf before the interpreter, LogicalProject.cpp:100-742). LIKE,
// SUBSTRING, LEFT, RIGHT, TRIM and the string-to-number/timestamp parses
// run inside the interpreter on string slices; the functions that build new
// strings become steps (below).
bool is_string_function(const std::string& v) {
  static const char* names[] = {"TRIM"};
  // TRIM with its side flag is interpreted (BLZ_TRIM_*); a bare TRIM is not
  for (const char* s : names)
    if (v == s) return true;
  return false;
}

// SUBSTRING(s, a, b) -> LEFT(SUBSTRING(s, a), b). The reference slices
// [a0, a0 + b) with a0 = max(a, 1) - 1 (LogicalProject.cpp:217-223, :240-306);
// only its all-literal form over a plain column (:217) reads a negative
// literal b as "to the end", which stays SUBSTRING(s, a).
std::unique_ptr<Node> rewrite_substring(std::unique_ptr<Node> n) {
  for (auto& c : n->children) c = rewrite_substring(std::move(c));
  if (n->type != NodeType::Operator || n->value != "SUBSTRING") return n;
  if (n->children.size() == 2) return n;
  if (n->children.size() != 3) fail("SUBSTRING takes two or three arguments");
  std::unique_ptr<Node> len = std::move(n->children[2]);
  n->children.pop_back();
  const bool literal_form = n->children[0]->type == NodeType::Variable &&
                            n->children[1]->type == NodeType::Literal && len->type == NodeType::Literal;
  if (literal_form && !is_null_literal(len->value) && std::stoll(len->value) < 0) return n;
  auto left = std::make_unique<Node>();
  left->type = NodeType::Operator;
  left->value = "LEFT";
  left->children.push_back(std::move(n));
  left->children.push_back(std::move(len));
  return left;
}

// BLZ_TRIM_<side>('chars', s) -> the string first and the characters to
// strip second, as the interpreter's string ops take them; the side is kept
// in the name and goes into the literal's scalar when it is encoded.
std::unique_ptr<Node> rewrite_trim(std::unique_ptr<Node> n) {
  for (auto& c : n->children) c = rewrite_trim(std::move(c));
  if (n->type != NodeType::Operator || !trim_side(n->value)) return n;
  if (n->children.size() != 2 || n->children[0]->type != NodeType::Literal)
    fail("TRIM takes a side flag, a literal set of characters and a string");
  std::swap(n->children[0], n->children[1]);
  return n;
}

// Result type of a node over the columns' types (get_output_type,
// expression_utils.cpp:115-266); a null literal is EMPTY.
int32_t node_type(const Node* n, const std::vector<int32_t>& types) {
  if (n->type == NodeType::Literal) return is_null_literal(n->value) ? BSQL_EMPTY : n->lit_type;
  if (n->type == NodeType::Variable) return types.at((size_t)n->var_index);
  const int op = map_to_operator(n->value);
  if (is_binary_op(op) && n->children.size() == 2)
    return output_type_binary(op, node_type(n->children[0].get(), types), node_type(n->children[1].get(), types));
  if (is_unary_op(op) && n->children.size() == 1)
    return output_type_unary(op, node_type(n->children[0].get(), types));
  return output_type_nullary(op);
}

std::unique_ptr<Node> var_node(int col) {
  auto v = std::make_unique<Node>();
  v->type = NodeType::Variable;
  v->value = "$" + std::to_string(col);
  v->var_index = col;
  return v;
}

std::unique_ptr<Program> compile_trees(std::vector<std::unique_ptr<Node>> trees, int n_cols);
int max_slot(const Program& p);

// Replaces, bottom-up, every function that builds new strings by a column
// built before the launch, as function_evaluator_transformer does
// (LogicalProject.cpp:670-731): CONCAT / ||, UPPER, LOWER, INITCAP, REVERSE,
// REPLACE, CAST(x AS VARCHAR) and a CASE whose values are strings
// (evaluate_string_case_when_else, :612-659). An argument that is an
// expression is first evaluated into a column of its own (the reference's
// evaluate_expressions calls for such arguments).
struct StepExtractor {
  Program& prog;
  std::vector<int32_t> types;  // the table's columns, then the steps'
  int n_table;

  int add_step(Step st) {
    prog.steps.push_back(std::move(st));
    types.push_back(prog.steps.back().out_type);
    return (int)types.size() - 1;
  }

  int eval_column(std::unique_ptr<Node> e) {
    if (e->type == NodeType::Variable) return e->var_index;
    Step st;
    st.kind = Step::EVAL;
    st.out_type = node_type(e.get(), types);
    std::vector<std::unique_ptr<Node>> one;
    one.push_back(std::move(e));
    std::unique_ptr<Program> sub = compile_trees(std::move(one), (int)types.size());
    if (max_slot(*sub) > BSQL_MAX_SLOT)
      fail("Interops does not support plans with an input or output index greater than 63");
    validate(*sub, sub->n_inputs);
    st.sub = make_program_handle(std::move(sub));
    return add_step(std::move(st));
  }

  // a string argument: a column (evaluated first when it is an expression),
  // a literal or null
  StepArg string_arg(std::unique_ptr<Node>& c, const char* fn, bool literal_ok) {
    StepArg a;
    if (c->type == NodeType::Literal) {
      if (is_null_literal(c->value)) { a.kind = StepArg::NULL_LITERAL; return a; }
      if (!is_string_type(c->lit_type)) fail(std::string(fn) + " needs string arguments");
      if (!literal_ok) fail(std::string(fn) + " operator not supported for string literals");
      a.kind = StepArg::LITERAL;
      a.str = scalar_from_string(c->value, BSQL_STRING).str;
      return a;
    }
    if (!is_string_type(node_type(c.get(), types))) fail(std::string(fn) + " needs string arguments");
    a.col = eval_column(std::move(c));
    return a;
  }

  StepArg literal_arg(const Node* c, const char* fn) {
    if (c->type != NodeType::Literal || is_null_literal(c->value) || !is_string_type(c->lit_type))
      fail(std::string(fn) + " needs string literals for its search and replacement");
    StepArg a;
    a.kind = StepArg::LITERAL;
    a.str = scalar_from_string(c->value, BSQL_STRING).str;
    return a;
  }

  std::unique_ptr<Node> visit(std::unique_ptr<Node> n) {
    if (n->type != NodeType::Operator) return n;
    for (auto& c : n->children) c = visit(std::move(c));
    const std::string& v = n->value;
    Step st;
    if (v == "||" || v == "CONCAT") {
      if (n->children.size() != 2) fail("CONCAT takes two arguments");
      if (n->children[0]->type == NodeType::Literal && n->children[1]->type == NodeType::Literal)
        fail("CONCAT operator between literals is not supported");
      st.op = SF_CONCAT;
      st.args.push_back(string_arg(n->children[0], "CONCAT", true));
      st.args.push_back(string_arg(n->children[1], "CONCAT", true));
    } else if (v == "UPPER" || v == "LOWER" || v == "INITCAP" || v == "REVERSE") {
      if (n->children.size() != 1) fail(v + " takes one argument");
      st.op = v == "UPPER" ? SF_UPPER : v == "LOWER" ? SF_LOWER : v == "INITCAP" ? SF_INITCAP : SF_REVERSE;
      st.args.push_back(string_arg(n->children[0], v.c_str(), false));
    } else if (v == "REPLACE") {
      if (n->children.size() != 3) fail("REPLACE takes a string, a search string and a replacement");
      st.op = SF_REPLACE;
      st.args.push_back(string_arg(n->children[0], "REPLACE", false));
      st.args.push_back(literal_arg(n->children[1].get(), "REPLACE"));
      st.args.push_back(literal_arg(n->children[2].get(), "REPLACE"));
      // cudf::strings::replace rejects an empty target
      if (st.args[1].str.empty()) fail("Parameter target must not be empty string.");
    } else if (v == "REGEXP_REPLACE") {
      // column, pattern, replacement[, position] (LogicalProject.cpp:148-182)
      if (n->children.size() < 3 || n->children.size() > 6)
        fail("REGEXP_REPLACE takes a string, a pattern, a replacement and an optional position");
      if (n->children.size() > 4) fail("Optional parameters occurrence and match_type are not yet supported.");
      st.op = SF_REGEXP_REPLACE;
      st.args.push_back(string_arg(n->children[0], "REGEXP_REPLACE", false));
      st.args.push_back(literal_arg(n->children[1].get(), "REGEXP_REPLACE"));
      st.args.push_back(literal_arg(n->children[2].get(), "REGEXP_REPLACE"));
      (void)regex_compile(st.args[1].str);  // rejects what the device matcher does not take
      if (n->children.size() == 4) {
        const Node* pos = n->children[3].get();
        if (pos->type != NodeType::Literal || is_null_literal(pos->value) || is_string_type(pos->lit_type) ||
            is_float_type(pos->lit_type))
          fail("REGEXP_REPLACE position must be an integer literal");
        const long long v1 = std::stoll(pos->value) - 1;
        if (v1 < 0) fail("Position must be greater than zero.");
        st.start = (int32_t)std::min<long long>(v1, INT32_MAX);
      }
    } else if (v == "CAST_VARCHAR" || v == "CAST_CHAR") {
      if (n->children.size() != 1) fail("malformed CAST");
      if (n->children[0]->type == NodeType::Literal) fail("CAST operator not supported for literals");
      const int32_t t = node_type(n->children[0].get(), types);
      if (is_duration_type(t) || t == BSQL_EMPTY || (type_size(t) == 0 && !is_string_type(t)))
        fail("CAST to VARCHAR is not supported for type " + std::to_string(t));
      st.op = SF_CAST;
      StepArg a;
      a.col = eval_column(std::move(n->children[0]));
      st.args.push_back(a);
    } else if (v == "FIRST_NON_MAGIC" && n->children.size() == 2 && n->children[0]->type == NodeType::Operator &&
               n->children[0]->value == "MAGIC_IF_NOT" && n->children[0]->children.size() == 2 &&
               (is_string_type(node_type(n->children[0]->children[1].get(), types)) ||
                is_string_type(node_type(n->children[1].get(), types)))) {
      std::unique_ptr<Node>& cond = n->children[0]->children[0];
      if (cond->type == NodeType::Literal) fail("CASE operator not supported for condition expression literals");
      if (node_type(cond.get(), types) != BSQL_BOOL8) fail("Expression does not evaluate to a boolean mask");
      st.op = SF_CASE;
      StepArg c;
      c.col = eval_column(std::move(cond));
      st.args.push_back(c);
      st.args.push_back(string_arg(n->children[0]->children[1], "CASE", true));
      st.args.push_back(string_arg(n->children[1], "CASE", true));
    } else {
      return n;
    }
    return var_node(add_step(std::move(st)));
  }
};

void check_interpretable(const Node* n) {
  if (n->type == NodeType::Operator && is_string_function(n->value))
    fail("string function " + n->value +
         " is evaluated before the interpreter (see the string column functions)");
  for (auto& c : n->children) check_interpretable(c.get());
}

}  // namespace

namespace {

// One program over n_cols columns for already transformed trees (each tree
// one output), as evaluate_expressions builds it (LogicalProject.cpp:
// 810-904): a literal root fills its column, a bare column root is copied,
// and only the other trees are encoded; their columns become the inputs, in
// column order, and interpreted output j has slot n_inputs + j.
std::unique_ptr<Program> compile_trees(std::vector<std::unique_ptr<Node>> trees, int n_cols) {
  std::vector<bool> used((size_t)std::max(n_cols, 0), false);
  const int n_exprs = (int)trees.size();
  auto prog = std::make_unique<Program>();
  prog->n_table = n_cols;
  std::vector<const Node*> interp;
  for (int i = 0; i < n_exprs; ++i) {
    const Node* root = trees[(size_t)i].get();
    Root r;
    if (root->type == NodeType::Literal) {
      r.kind = Root::LITERAL;
      if (is_null_literal(root->value)) {
        r.lit.type_id = root->lit_type;
        r.lit.is_null = true;
      } else {
        r.lit = scalar_from_string(root->value, root->lit_type);
      }
      if (is_string_type(r.lit.type_id)) fail("a string literal expression is not produced by the interpreter");
    } else if (root->type == NodeType::Variable) {
      r.kind = Root::COLUMN;
      r.index = root->var_index;
      if (r.index < 0 || r.index >= n_cols) fail("column $" + std::to_string(r.index) + " not in the table");
    } else {
      check_interpretable(root);
      std::vector<int> vars;
      collect_vars(root, vars);
      for (int v : vars) {
        if (v < 0 || v >= n_cols) fail("column $" + std::to_string(v) + " not in the table");
        used[(size_t)v] = true;
      }
      r.kind = Root::VM;
      r.index = (int)interp.size();
      interp.push_back(root);
    }
    prog->roots.push_back(r);
  }
  std::map<int, int> col_map;
  for (int i = 0; i < n_cols; ++i)
    if (used[(size_t)i]) {
      const int slot = (int)col_map.size();
      col_map[i] = slot;
      prog->used_columns.push_back(i);
    }
  const int n_inputs = (int)col_map.size();
  const int n_interp = (int)interp.size();
  prog->n_inputs = n_inputs;
  for (int j = 0; j < n_interp; ++j) {
    const int final_pos = n_inputs + j;
    prog->final_out.push_back((int16_t)final_pos);
    Encoder enc(*prog, col_map, n_inputs + n_interp);
    enc.visit(interp[(size_t)j]);
    prog->out.back() = (int16_t)final_pos;
  }
  return prog;
}

// Largest slot of an encoded program.
int max_slot(const Program& p) {
  int mx = -1;
  for (size_t i = 0; i < p.ops.size(); ++i)
    mx = std::max<int>(mx, std::max<int>(p.left[i], std::max<int>(p.right[i], p.out[i])));
  for (int16_t f : p.final_out) mx = std::max<int>(mx, f);
  return mx;
}

}  // namespace

namespace {

// The program of one expression list: the string-building functions become
// steps, the remaining trees are encoded; an encoding that reaches slot 64
// is replaced by the two halves of the list, recursively
// (LogicalProject.cpp:906-922). A single expression over the limit fails
// (the reference recurses without end).
std::unique_ptr<Program> compile_list(const std::vector<std::string>& exprs, const int32_t* in_types, int n_in) {
  auto holder = std::make_unique<Program>();
  StepExtractor ex{*holder, std::vector<int32_t>(in_types, in_types + std::max(n_in, 0)), n_in};
  std::vector<std::unique_ptr<Node>> trees;
  for (const std::string& raw : exprs) {
    std::string e = replace_calcite_regex(preprocess_expression(raw, in_types, n_in));
    e = expand_if_logical_op(e);
    Parser parser(e);
    std::unique_ptr<Node> t = rewrite_trim(rewrite_substring(transform_custom(parser.parse())));
    std::vector<int> vars;
    collect_vars(t.get(), vars);
    for (int v : vars)
      if (v < 0 || v >= n_in) fail("column $" + std::to_string(v) + " not in the table");
    t = ex.visit(std::move(t));
    // a string literal root fills its column (make_column_from_scalar, :822-832)
    if (t->type == NodeType::Literal && !is_null_literal(t->value) && is_string_type(t->lit_type)) {
      Step st;
      st.op = SF_FILL;
      StepArg a;
      a.kind = StepArg::LITERAL;
      a.str = scalar_from_string(t->value, BSQL_STRING).str;
      st.args.push_back(a);
      t = var_node(ex.add_step(std::move(st)));
    }
    trees.push_back(std::move(t));
  }
  std::unique_ptr<Program> prog = compile_trees(std::move(trees), (int)ex.types.size());
  if (max_slot(*prog) > BSQL_MAX_SLOT) {
    if (exprs.size() < 2)
      fail("Interops does not support plans with an input or output index greater than 63");
    const size_t half = exprs.size() / 2;
    const std::vector<std::string> lo(exprs.begin(), exprs.begin() + (long)half);
    const std::vector<std::string> hi(exprs.begin() + (long)half, exprs.end());
    auto split = std::make_unique<Program>();
    split->n_table = n_in;
    split->n_inputs = 0;
    for (int h = 0; h < 2; ++h) {
      std::unique_ptr<Program> part = compile_list(h ? hi : lo, in_types, n_in);
      const int base = h ? (int)half : 0;
      auto add = [&](std::shared_ptr<Program> q, const std::vector<int>& outs) {
        split->parts.push_back(std::move(q));
        split->part_outputs.push_back(outs);
      };
      if (part->parts.empty()) {
        std::vector<int> outs;
        for (int k = 0; k < num_roots(*part); ++k) outs.push_back(base + k);
        add(make_program_handle(std::move(part)), outs);
      } else {
        // flatten: the halves' own parts
        for (size_t j = 0; j < part->parts.size(); ++j) {
          std::vector<int> outs;
          for (int k : part->part_outputs[j]) outs.push_back(base + k);
          add(part->parts[j], outs);
        }
      }
    }
    Root r;
    for (size_t k = 0; k < exprs.size(); ++k) split->roots.push_back(r);  // counted; produced by the parts
    return split;
  }
  validate(*prog, prog->n_inputs);
  prog->steps = std::move(holder->steps);
  prog->n_table = n_in;
  return prog;
}

}  // namespace

std::unique_ptr<Program> compile_expressions(const std::vector<std::string>& exprs,
                                             const int32_t* in_types, int n_in) {
  return compile_list(exprs, in_types, n_in);
}

std::string preprocess_expression(const std::string& e, const int32_t* types, int n) {
  std::string r = fill_minus_op_with_zero(e);
  r = nested_concat(r);
  r = current_date_or_timestamp(r);
  r = ms_to_ns_units(r);
  r = reinterpret_timestamp(r, types, n);
  return interval_conversion(r, types, n);
}

namespace {
std::string rebuild(const Node* n) {
  if (n->type != NodeType::Operator) return n->value;
  std::string args;
  for (const auto& c : n->children) args += (args.empty() ? "" : ", ") + rebuild(c.get());
  return n->value + "(" + args + ")";
}
}  // namespace

// modify_multi_column_count_expression (expression_utils.cpp:1467-1492):
// the input of COUNT over several columns, "$a, $b, .." (the text between
// the aggregation's parentheses), as the sum of their INTEGER casts, so that
// a row is null when any of the columns is; one column stays as it is.
std::string multi_column_count_expression(const std::string& e) {
  const size_t first = e.find('$');
  if (first == std::string::npos) return e;
  std::string reduced = e.substr(first + 1);
  reduced.erase(std::remove(reduced.begin(), reduced.end(), '$'), reduced.end());
  const std::vector<std::string> idx = split_expression_list(reduced, true);
  if (idx.size() <= 1) return e;
  std::string out = "+(CAST($" + std::to_string(std::stoi(idx[0])) + "):INTEGER, CAST($" +
                    std::to_string(std::stoi(idx[1])) + "):INTEGER)";
  for (size_t i = 2; i < idx.size(); ++i)
    out = "+(" + out + ", CAST($" + std::to_string(std::stoi(idx[i])) + "):INTEGER)";
  return out;
}

std::string rewrite_expression(const std::string& e, int stage, const int32_t* types, int n) {
  switch (stage) {
    case 0: return fill_minus_op_with_zero(e);
    case 1: return nested_concat(e);
    case 2: return ms_to_ns_units(e);
    case 3: return reinterpret_timestamp(e, types, n);
    case 4: return interval_conversion(e, types, n);
    case 5: return preprocess_expression(e, types, n);
    case 6: return replace_calcite_regex(e);
    case 7: return expand_if_logical_op(replace_calcite_regex(e));
    case 8: {
      Parser parser(e);
      return rebuild(transform_custom(parser.parse()).get());
    }
    case 9: return multi_column_count_expression(e);
    default: fail("unknown rewrite stage " + std::to_string(stage));
  }
}

// ------------------------------------------------------------ skip data
//
// ral::skip_data (SkipDataProcessor.cpp:23-232, utils.cpp:12-34): the filter
// of a table scan rewritten over the table of row-group statistics, in which
// projected column i has its minimum at column 2i and its maximum at 2i + 1.
// Three post-order passes over the parsed filter, as the reference's
// transformers: columns without statistics become NONE; operators other than
// = > >= < <= + - AND OR become NONE, and a NONE (or unary) operand makes an
// AND its other operand and anything else NONE; then each comparison is
// taken over the ranges (= : AND(<=(a_min, b_max), >=(a_max, b_min)); < <= :
// (a_min, b_max); > >= : (a_max, b_min)) and + - as the pair (a_min op
// b_min, a_max op b_max), as the reference writes them.

namespace {

std::unique_ptr<Node> sd_operator(const std::string& v) {
  auto n = std::make_unique<Node>();
  n->type = NodeType::Operator;
  n->value = v;
  return n;
}

std::unique_ptr<Node> sd_clone(const Node* n) {
  auto c = std::make_unique<Node>();
  c->type = n->type;
  c->value = n->value;
  c->lit_type = n->lit_type;
  c->var_index = n->var_index;
  for (const auto& k : n->children) c->children.push_back(sd_clone(k.get()));
  return c;
}

template <class F>
std::unique_ptr<Node> sd_post_order(std::unique_ptr<Node> n, F& f) {
  if (n->type == NodeType::Operator)
    for (auto& c : n->children) c = sd_post_order(std::move(c), f);
  return f(std::move(n));
}

bool sd_supported(const std::string& op) {
  static const char* const ok[] = {"=", ">", ">=", "<", "<=", "+", "-", "AND", "OR"};
  for (const char* o : ok)
    if (op == o) return true;
  return false;
}

// utils.cpp:22-34: NONE, or a unary operator
bool sd_excluded(const Node& n) {
  if (n.type != NodeType::Operator) return false;
  if (n.value == "NONE") return true;
  return is_unary_op(map_to_operator(n.value));
}

// the minimum (k = 0) or maximum (k = 1) side of an operand
std::unique_ptr<Node> sd_side(const Node* n, int k) {
  if (n->type == NodeType::Operator) {
    if (n->value == "&&&" && n->children.size() == 2) return sd_clone(n->children[(size_t)k].get());
    return sd_clone(n);
  }
  if (n->type == NodeType::Variable) {
    auto v = sd_clone(n);
    v->var_index = 2 * n->var_index + k;
    v->value = "$" + std::to_string(v->var_index);
    return v;
  }
  return sd_clone(n);
}

std::unique_ptr<Node> sd_binary(const std::string& op, std::unique_ptr<Node> a, std::unique_ptr<Node> b) {
  auto n = sd_operator(op);
  n->children.push_back(std::move(a));
  n->children.push_back(std::move(b));
  return n;
}

}  // namespace

std::string skip_data_filter(const std::string& filter, const std::vector<bool>& has_stats) {
  Parser parser(expand_if_logical_op(replace_calcite_regex(filter)));
  std::unique_ptr<Node> root = parser.parse();
  auto drop = [&](std::unique_ptr<Node> n) -> std::unique_ptr<Node> {
    if (n->type == NodeType::Variable &&
        (n->var_index < 0 || n->var_index >= (int)has_stats.size() || !has_stats[(size_t)n->var_index]))
      return sd_operator("NONE");
    return n;
  };
  root = sd_post_order(std::move(root), drop);
  auto reduce = [&](std::unique_ptr<Node> n) -> std::unique_ptr<Node> {
    if (n->type != NodeType::Operator) return n;
    if (!sd_supported(n->value)) return sd_operator("NONE");
    if (n->children.size() != 2) fail("skip data: operator " + n->value + " without two operands");
    const bool l = sd_excluded(*n->children[0]), r = sd_excluded(*n->children[1]);
    if (l && r) return sd_operator("NONE");
    if (l) return n->value == "AND" ? std::move(n->children[1]) : sd_operator("NONE");
    if (r) return n->value == "AND" ? std::move(n->children[0]) : sd_operator("NONE");
    return n;
  };
  root = sd_post_order(std::move(root), reduce);
  if (root->type == NodeType::Operator && root->value == "NONE") return "";
  auto ranges = [&](std::unique_ptr<Node> n) -> std::unique_ptr<Node> {
    if (n->type != NodeType::Operator || n->children.size() != 2) return n;
    const Node* a = n->children[0].get();
    const Node* b = n->children[1].get();
    const std::string& op = n->value;
    if (op == "=")
      return sd_binary("AND", sd_binary("<=", sd_side(a, 0), sd_side(b, 1)), sd_binary(">=", sd_side(a, 1), sd_side(b, 0)));
    if (op == "<" || op == "<=") return sd_binary(op, sd_side(a, 0), sd_side(b, 1));
    if (op == ">" || op == ">=") return sd_binary(op, sd_side(a, 1), sd_side(b, 0));
    if (op == "+" || op == "-")
      return sd_binary("&&&", sd_binary(op, sd_side(a, 0), sd_side(b, 0)), sd_binary(op, sd_side(a, 1), sd_side(b, 1)));
    return n;
  };
  root = sd_post_order(std::move(root), ranges);
  return rebuild(root.get());
}

void set_current_timestamp(const std::string& ts) {
  std::lock_guard<std::mutex> lk(g_current_mu);
  g_current_timestamp = ts;
}

// ------------------------------------------------------------ query parts

std::string named_expression(const std::string& query_part, const std::string& name) {
  const size_t single = query_part.find(name + "=[");
  if (single == std::string::npos) return "";
  const size_t dbl = query_part.find(name + "=[[");
  const size_t start = dbl != std::string::npos ? dbl + name.size() + 3 : single + name.size() + 2;
  const size_t end = query_part.find(']', start);
  return query_part.substr(start, end == std::string::npos ? std::string::npos : end - start);
}

std::string query_part_body(const std::string& logical_plan) {
  const size_t open = logical_plan.find('(');
  const size_t close = logical_plan.rfind(')');
  if (open == std::string::npos || close == std::string::npos || close <= open)
    fail("query part without a parenthesized body: " + logical_plan);
  return logical_plan.substr(open + 1, close - open - 1);
}

std::vector<std::string> expression_list(const std::string& combined) {
  return split_expression_list(replace_calcite_regex(combined), true);
}

void project_items(const std::string& query_part, std::vector<std::string>* names,
                   std::vector<std::string>* exprs) {
  for (const std::string& item : expression_list(query_part_body(query_part))) {
    const size_t eq = item.find("=[");
    if (eq == std::string::npos || item.size() < eq + 3 || item.back() != ']')
      fail("projection item is not name=[expression]: " + item);
    names->push_back(item.substr(0, eq));
    exprs->push_back(item.substr(eq + 2, item.size() - eq - 3));
  }
}

}  // namespace bsql
