// program.hpp -- host-side program representation of the row-parallel
// interpreter, and its lowering to the MI355X device plan.
//
// A Program holds the reference encoding (engine/src/Interpreter/interpreter_cpp.h:50-86):
// parallel int16 vectors left/right/out, the operator per op, the final output
// slots, and the literal scalars. lower() turns it into a DevPlan: operand
// types resolved per op (interpreter_cpp.cu:502-550), each op specialised to
// its int64/double operand kinds, column loads inserted before first use, and
// the slot numbering replaced by a dense register numbering found by liveness,
// so that the whole per-row state lives in VGPRs on gfx950.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bsql/interp.h"
#include "devplan.h"

namespace bsql {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(const std::string& m) { throw Error(BSQL_ERR_INVALID, m); }

// literal scalar with owned string storage
struct Scalar {
  int32_t type_id = BSQL_EMPTY;
  bool is_null = true;
  int64_t i64 = 0;
  double f64 = 0.0;
  std::string str;
};

struct Program;

// A column built before the interpreter runs and appended after the table's
// columns (the reference builds these with libcudf and appends them as $N+k,
// LogicalProject.cpp:670-731). EVAL runs a one-output sub-program over the
// columns so far (an argument of a string function that is an expression, or
// a string CASE's condition); STRFN builds a string column (strfn.h's ops).
struct StepArg {
  enum Kind { COLUMN, LITERAL, NULL_LITERAL } kind = COLUMN;
  int col = -1;     // COLUMN: index among the table's columns, then earlier steps'
  std::string str;  // LITERAL
};
struct Step {
  enum Kind { EVAL, STRFN } kind = STRFN;
  std::shared_ptr<Program> sub;  // EVAL
  int op = 0;                    // STRFN: StrFnOp
  std::vector<StepArg> args;
  int32_t out_type = BSQL_STRING;
  int32_t start = 0;             // REGEXP_REPLACE: characters kept before the search
  std::shared_ptr<void> d_lits;  // device copy of the literal arguments, made on first use
};

// One expression of a compiled list (evaluate_expressions, LogicalProject.cpp:
// 810-852): interpreted (output `index` of the encoded program), a bare column
// (the reference copies table column `index`, or moves a column a step built),
// or a literal (make_column_from_scalar; lit.is_null for a null literal).
struct Root {
  enum Kind { VM, COLUMN, LITERAL } kind = VM;
  int index = 0;
  Scalar lit;
};

struct Program {
  // the reference encoding of the interpreted expressions only
  std::vector<int16_t> left, right, out, final_out;
  std::vector<int32_t> ops;
  std::vector<Scalar> lsc, rsc;
  // every expression of the list, in order (empty: output k is final_out[k],
  // as for programs built from arrays)
  std::vector<Root> roots;
  // an expression list whose encoding reaches slot 64 is evaluated as halves
  // (LogicalProject.cpp:906-922): each part is a program of its own over the
  // same table, and part_outputs[j] lists the outputs part j produces
  std::vector<std::shared_ptr<Program>> parts;
  std::vector<std::vector<int>> part_outputs;
  // schema hints from the compiler: number of interpreter inputs and the
  // table columns they come from (col_idx_map of LogicalProject.cpp:859-881)
  int n_inputs = -1;
  std::vector<int32_t> used_columns;
  // conjunct order override (empty = plan order) and launch override
  std::vector<int32_t> conj_order;
  // (the interpreter's register file in LDS unless the caller clears it)
  bsql_launch_config launch{0, 0, BSQL_LAUNCH_LDS_REGS};
  uint64_t seed = 0x5eed5eed12345678ULL;
  // columns built before the launch: step k is column n_table + k
  std::vector<Step> steps;
  int n_table = -1;
  // device copies of string literals, one buffer per program
  void* d_strings = nullptr;
  std::vector<int64_t> str_offsets;  // per (op, side): offset into d_strings or -1
  ~Program();
};

// A program allocated as a C ABI handle (it carries the lowering cache), for
// the sub-programs of steps and the parts of split lists.
std::shared_ptr<Program> make_program_handle(std::unique_ptr<Program> p);

// A fresh unique program id.
uint64_t next_program_id();

// Number of expressions (outputs) of a program.
inline int num_roots(const Program& p) {
  return p.roots.empty() ? (int)p.final_out.size() : (int)p.roots.size();
}
// Root k of a program (VM k for programs without a root list).
inline Root root_of(const Program& p, int k) {
  if (!p.roots.empty()) return p.roots[(size_t)k];
  Root r;
  r.index = k;
  return r;
}

// type helpers (cudf type_id semantics)
bool is_float_type(int t);
bool is_timestamp_type(int t);
bool is_duration_type(int t);
bool is_string_type(int t);
int type_size(int t);
int64_t ticks_per_day(int t);
int64_t ns_per_tick(int t);

bool is_nullary_op(int op);
bool is_unary_op(int op);
bool is_binary_op(int op);
// get_output_type (engine/src/parser/expression_utils.cpp:115-266)
int output_type_nullary(int op);
int output_type_unary(int op, int lt);
int output_type_binary(int op, int lt, int rt);

// Operand and result types of every op for a given input schema.
struct OpTypes {
  std::vector<int32_t> lt, rt, ot;
};
OpTypes propagate_types(const Program& p, const int32_t* in_types, int n_in);

// LIKE pattern compiled for the device matcher: {n_units, flags, units...}.
std::vector<int32_t> like_compile(const std::string& pattern);

// A regular expression compiled for the device's bitset matcher (regex.cpp).
std::vector<int32_t> regex_compile(const std::string& re);

// to_timestamps format compiled for the device parser: {n_items, items...},
// an item (spec << 8) | width with spec 0 for literal characters.
std::vector<int32_t> timestamp_format_compile(const std::string& format);

// Types of the interpreter input slots of `p` for a table schema: programs
// compiled from text address their used columns, array-built ones the table.
std::vector<int32_t> slot_types_for(const Program& p, const int32_t* table_types, int n_table);

// Checks the reference's program limits (interpreter_cpp.cu:329-378, 466).
void validate(const Program& p, int n_in);

// Which register-file size / rows-per-thread a lowered plan needs.
struct LowerOptions {
  enum Mode { EXECUTE, FILTER_PROJECT, FILTER_GATHER, FILTER_COUNT } mode = EXECUTE;
  const Program* pred = nullptr;   // FILTER_*: predicate (1 BOOL8 output)
  const Program* proj = nullptr;   // EXECUTE / FILTER_PROJECT: outputs
  bool conjuncts = true;           // wave-level early exit per conjunct
  bool lazy_loads = false;         // load each conjunct's columns only when it is reached
  bool ranges = true;              // column-vs-literal conjuncts become range tests
  const std::vector<int32_t>* conj_order = nullptr;
  // outputs of the projection this launch writes (empty: all of them)
  const std::vector<int>* out_subset = nullptr;
  // registers the allocator may use before it spills (the file minus the
  // spare): 15 for a 16-register file, 7 for an 8-register file
  int reg_budget = BSQL_DEV_MAX_REG - 1;
};

struct Lowered {
  DevPlan plan;          // columns are filled by the launcher; plan.ops is not set
  std::vector<DevOp> ops;  // the op list (plan.n_ops entries, at least one)
  int nreg = 0;          // registers used (+ the spare)
  int n_spill = 0;       // LDS spill slots
  int n_conjuncts = 0;   // conjuncts found in the predicate (0 = not an AND chain)
  std::vector<DevRange> plan_ranges;  // the selection as range tests (empty: ops)
  std::vector<int> out_str_col;       // per output: source table column of a string output, else -1
  std::vector<int> in_cols;           // device input j reads table column in_cols[j]
};
Lowered lower(const LowerOptions& opt, const int32_t* in_types, int n_in,
              const int32_t* out_types, int n_out);

// Conjunct decomposition of a predicate whose root is a left-deep AND chain.
int count_conjuncts(const Program& p);

// The range test of each conjunct of a predicate, in plan order (not merged);
// fails unless every conjunct is a column-vs-literal comparison.
std::vector<DevRange> conjunct_ranges(const Program& p, const int32_t* in_types, int n_in);

// Text front-end (parser.cpp): Calcite expression strings -> Program.
std::unique_ptr<Program> compile_expressions(const std::vector<std::string>& exprs,
                                             const int32_t* in_types, int n_in);

// preprocess_expression (expression_utils.cpp:1624-1632): unary minus,
// n-ary CONCAT, CURRENT_DATE/TIMESTAMP, ms -> ns units, Reinterpret
// timestamp units and interval units, for a table schema.
std::string preprocess_expression(const std::string& e, const int32_t* types, int n);
// One text stage of the front end (for tests): 0 fill_minus_op_with_zero,
// 1 n-ary CONCAT, 2 ms -> ns, 3 Reinterpret, 4 intervals, 5 all of
// preprocess_expression, 6 replace_calcite_regex, 7 + expand_if_logical_op,
// 8 parse + custom-op rewrites, rebuilt as text (rebuildExpression).
std::string rewrite_expression(const std::string& e, int stage, const int32_t* types, int n);
// The timestamp CURRENT_DATE / CURRENT_TIMESTAMP stand for ("" = the clock).
void set_current_timestamp(const std::string& ts);

// Query-part text of the operators (parser.cpp).
// `name=[[...]]` or `name=[...]` of a query part; "" when absent.
std::string named_expression(const std::string& query_part, const std::string& name);
// A table scan's filter over the row-group statistics table (min of
// projected column i at 2i, max at 2i + 1); has_stats[i]: column i has
// statistics. "" when no row group can be skipped.
std::string skip_data_filter(const std::string& filter, const std::vector<bool>& has_stats);
std::string multi_column_count_expression(const std::string& e);
// The text between the first '(' and the last ')'.
std::string query_part_body(const std::string& logical_plan);
// Comma-separated list at bracket/paren depth 0 outside quotes, after the
// Calcite regex rewrites; items trimmed.
std::vector<std::string> expression_list(const std::string& combined);
// `name=[expr]` items of a LogicalProject: (names, expressions).
void project_items(const std::string& query_part, std::vector<std::string>* names,
                   std::vector<std::string>* exprs);

}  // namespace bsql
