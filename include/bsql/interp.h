/*
 * bsql/interp.h -- C ABI of the MI355X row-parallel expression interpreter,
 * filter (boolean-mask stream compaction) and fused filter+project path.
 *
 * Drop-in boundary for BlazingSQL's interpreter/filter path. Each entry point
 * names the reference interface it replaces (paths relative to the
 * BlazingSQL 21.08 source tree).
 *
 * Conventions
 *   - Plain C: no exceptions cross the ABI, no torch/cudf types.
 *   - Column buffers are device pointers (hipMalloc'd or torch-owned).
 *   - Return codes: BSQL_OK (0); BSQL_ERR_OOM (1, retryable: inputs untouched,
 *     the operator layer maps it to task_status::RETRY); BSQL_ERR_INVALID (2,
 *     invalid or unsupported program/arguments: FAIL); BSQL_ERR_DEVICE (3).
 *     bsql_last_error() returns the message of the calling thread's last error.
 *   - Every launching call takes an explicit hipStream_t (passed as void*), is
 *     reentrant per stream and never synchronizes the device.
 */
#ifndef BSQL_INTERP_H
#define BSQL_INTERP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  BSQL_OK = 0,
  BSQL_ERR_OOM = 1,
  BSQL_ERR_INVALID = 2,
  BSQL_ERR_DEVICE = 3
};

/* ---- type ids: numbering of cudf::type_id (libcudf 21.08) ----------------
 * The interpreter dispatches on these exactly as the reference does
 * (engine/src/Interpreter/interpreter_ops.cuh:141-159, 267-395). */
enum {
  BSQL_EMPTY = 0,
  BSQL_INT8 = 1, BSQL_INT16 = 2, BSQL_INT32 = 3, BSQL_INT64 = 4,
  BSQL_UINT8 = 5, BSQL_UINT16 = 6, BSQL_UINT32 = 7, BSQL_UINT64 = 8,
  BSQL_FLOAT32 = 9, BSQL_FLOAT64 = 10, BSQL_BOOL8 = 11,
  BSQL_TIMESTAMP_DAYS = 12, BSQL_TIMESTAMP_SECONDS = 13,
  BSQL_TIMESTAMP_MILLISECONDS = 14, BSQL_TIMESTAMP_MICROSECONDS = 15,
  BSQL_TIMESTAMP_NANOSECONDS = 16,
  BSQL_DURATION_DAYS = 17, BSQL_DURATION_SECONDS = 18,
  BSQL_DURATION_MILLISECONDS = 19, BSQL_DURATION_MICROSECONDS = 20,
  BSQL_DURATION_NANOSECONDS = 21,
  BSQL_DICTIONARY32 = 22, BSQL_STRING = 23, BSQL_LIST = 24,
  BSQL_DECIMAL32 = 25, BSQL_DECIMAL64 = 26, BSQL_STRUCT = 27
};

/* ---- operand codes of the postfix program --------------------------------
 * engine/src/Interpreter/interpreter_cpp.h:19-26 (column_index). */
enum {
  BSQL_UNARY_INDEX = -1,
  BSQL_SCALAR_INDEX = -2,
  BSQL_SCALAR_NULL_INDEX = -3,
  BSQL_NULLARY_INDEX = -4
};

/* ---- opcodes: numeric values of operator_type ------------------------------
 * engine/src/parser/expression_utils.hpp:11-93 (enumerator order). */
enum {
  BSQL_OP_INVALID = 0,
  BSQL_OP_RAND = 1,
  BSQL_OP_NOT = 2, BSQL_OP_ABS = 3, BSQL_OP_FLOOR = 4, BSQL_OP_CEIL = 5,
  BSQL_OP_SIN = 6, BSQL_OP_COS = 7, BSQL_OP_ASIN = 8, BSQL_OP_ACOS = 9,
  BSQL_OP_TAN = 10, BSQL_OP_COTAN = 11, BSQL_OP_ATAN = 12, BSQL_OP_LN = 13,
  BSQL_OP_LOG = 14, BSQL_OP_YEAR = 15, BSQL_OP_MONTH = 16, BSQL_OP_DAY = 17,
  BSQL_OP_DAYOFWEEK = 18, BSQL_OP_HOUR = 19, BSQL_OP_MINUTE = 20,
  BSQL_OP_SECOND = 21, BSQL_OP_IS_NULL = 22, BSQL_OP_IS_NOT_NULL = 23,
  BSQL_OP_IS_TRUE = 24, BSQL_OP_IS_NOT_TRUE = 25, BSQL_OP_IS_NOT_FALSE = 26,
  BSQL_OP_CAST_TINYINT = 27, BSQL_OP_CAST_SMALLINT = 28,
  BSQL_OP_CAST_INTEGER = 29, BSQL_OP_CAST_BIGINT = 30,
  BSQL_OP_CAST_FLOAT = 31, BSQL_OP_CAST_DOUBLE = 32, BSQL_OP_CAST_DATE = 33,
  BSQL_OP_CAST_TIMESTAMP_SECONDS = 34, BSQL_OP_CAST_TIMESTAMP_MILLISECONDS = 35,
  BSQL_OP_CAST_TIMESTAMP_MICROSECONDS = 36, BSQL_OP_CAST_TIMESTAMP = 37,
  BSQL_OP_CAST_VARCHAR = 38, BSQL_OP_CHAR_LENGTH = 39,
  BSQL_OP_STR_LOWER = 40, BSQL_OP_STR_UPPER = 41, BSQL_OP_STR_INITCAP = 42,
  BSQL_OP_STR_REVERSE = 43,
  BSQL_OP_ADD = 44, BSQL_OP_SUB = 45, BSQL_OP_MUL = 46, BSQL_OP_DIV = 47,
  BSQL_OP_MOD = 48, BSQL_OP_POW = 49, BSQL_OP_ROUND = 50,
  BSQL_OP_EQUAL = 51, BSQL_OP_NOT_EQUAL = 52, BSQL_OP_LESS = 53,
  BSQL_OP_GREATER = 54, BSQL_OP_LESS_EQUAL = 55, BSQL_OP_GREATER_EQUAL = 56,
  BSQL_OP_BITWISE_AND = 57, BSQL_OP_BITWISE_OR = 58, BSQL_OP_BITWISE_XOR = 59,
  BSQL_OP_LOGICAL_AND = 60, BSQL_OP_LOGICAL_OR = 61,
  BSQL_OP_FIRST_NON_MAGIC = 62, BSQL_OP_MAGIC_IF_NOT = 63,
  BSQL_OP_STR_LIKE = 64, BSQL_OP_STR_SUBSTRING = 65,
  BSQL_OP_STR_REGEXP_REPLACE = 66, BSQL_OP_STR_CONCAT = 67,
  BSQL_OP_STR_REPLACE = 68, BSQL_OP_STR_LEFT = 69, BSQL_OP_STR_RIGHT = 70,
  BSQL_OP_TO_DATE = 71, BSQL_OP_TO_TIMESTAMP = 72, BSQL_OP_STR_TRIM = 73,
  BSQL_OP_IS_NOT_DISTINCT_FROM = 74,
  BSQL_OP_COUNT = 75
};

/* Largest slot index a program may use (interpreter_cpp.cu:466). */
#define BSQL_MAX_SLOT 63

/* ---- columns: Arrow/cudf layout -------------------------------------------
 * Mirrors cudf::column_view as used by BlazingColumn/BlazingTable
 * (engine/src/execution_kernels/LogicPrimitives.h:23-89).
 *   data        fixed-width elements (STRING: the chars buffer)
 *   validity    32-bit words, LSB-first, 1 = valid; NULL means all valid
 *   offsets     STRING only: size+1 int32 offsets into data
 *   offset      first row of the view (row units; bit units into validity)
 *   chars_size  STRING only: bytes in the chars buffer (an output's
 *               capacity; it must hold the chars of the column it is made
 *               from, which bounds any slice or selection of it)          */
typedef struct bsql_column {
  int32_t type_id;
  int32_t chars_size;
  int64_t size;
  int64_t offset;
  void* data;
  uint32_t* validity;
  int32_t* offsets;
} bsql_column;

/* ---- scalars (literal operands) --------------------------------------------
 * One per program operand that is BSQL_SCALAR_INDEX. The typed value follows
 * get_scalar_from_string (engine/src/parser/CalciteExpressionParsing.cpp:138-262):
 *   integral/bool/timestamp/duration -> i64 (already narrowed to the type)
 *   FLOAT32/FLOAT64                  -> f64 (FLOAT32 holds the float value)
 *   STRING                           -> str/str_len (host memory)            */
typedef struct bsql_scalar {
  int32_t type_id;
  int32_t str_len;
  union { int64_t i64; double f64; } v;
  const char* str;
} bsql_scalar;

typedef struct bsql_program bsql_program;

/* ---- program construction ---------------------------------------------------
 * bsql_program_from_arrays: the reference encoding as built by
 * add_expression_to_interpreter_plan (engine/src/Interpreter/interpreter_cpp.h:50-59)
 * and consumed by perform_interpreter_operation (interpreter_cpp.h:77-86):
 * four parallel int16 vectors (left, right, out per op, final_out per output),
 * the operator per op, and one left/right scalar per op (consulted only where
 * the operand is BSQL_SCALAR_INDEX; may be NULL when no op has a scalar). */
int bsql_program_from_arrays(const int16_t* left, const int16_t* right,
                             const int16_t* out, const int32_t* ops, int n_ops,
                             const int16_t* final_out, int n_final,
                             const bsql_scalar* left_scalars,
                             const bsql_scalar* right_scalars,
                             bsql_program** prog);

/* bsql_program_compile: Calcite expression text -> program, replacing the
 * per-batch front-end of evaluate_expressions
 * (engine/src/execution_kernels/LogicalProject.cpp:796-941): regex rewrite,
 * AND/OR expansion, parse, CASE/ROUND/CAST rewrites, slot allocation.
 * `in_types[n_in]` is the schema of the table the expressions refer to ($k).
 * As in the reference, only expressions with an operator at the root are
 * encoded for the interpreter; a bare column is copied and a literal fills
 * its column (bsql_program_root). The launches write every output.        */
int bsql_program_compile(const char* const* exprs, int n_exprs,
                         const int32_t* in_types, int n_in,
                         bsql_program** prog);

void bsql_program_free(bsql_program* prog);

/* Introspection (reference encoding, after compile): counts and arrays.
 * num_outputs counts every expression of the list; the encoded program holds
 * only the interpreted ones (num_interpreted of them, final_out[j] = n_in + j
 * as add_expression_to_interpreter_plan numbers them,
 * LogicalProject.cpp:883-904). */
int bsql_program_num_ops(const bsql_program* prog);
int bsql_program_num_outputs(const bsql_program* prog);
int bsql_program_num_inputs(const bsql_program* prog);
int bsql_program_num_interpreted(const bsql_program* prog);
/* Fills caller arrays of length num_ops / num_interpreted; any may be NULL.
 * Fails for a program split into parts (below). */
int bsql_program_arrays(const bsql_program* prog, int16_t* left, int16_t* right,
                        int16_t* out, int32_t* ops, int16_t* final_out);
/* How expression k is produced (evaluate_expressions, LogicalProject.cpp:
 * 810-852): by the interpreter (index = its position among the interpreted
 * outputs), as a copy of a column (index = table column; a column a step
 * builds is moved, the table's copied), or filled with a literal (index -1;
 * *literal gets its type and value, str_len -1 for a null literal). */
enum { BSQL_ROOT_INTERPRETED = 0, BSQL_ROOT_COLUMN = 1, BSQL_ROOT_LITERAL = 2 };
int bsql_program_root(const bsql_program* prog, int k, int32_t* kind, int32_t* index,
                      bsql_scalar* literal);
/* An expression list whose encoding would use a slot >= 64 is compiled as
 * halves, recursively, each a program of its own over the same table
 * (LogicalProject.cpp:906-922); launches run the parts one after another.
 * num_parts is 0 for a program in one piece. bsql_program_part gives part j
 * (owned by prog) and which outputs of prog it produces (up to cap written);
 * returns their number. */
int bsql_program_num_parts(const bsql_program* prog);
int bsql_program_part(const bsql_program* prog, int j, const bsql_program** part,
                      int32_t* outputs, int cap);
/* Scalar of op i (side 0 = left, 1 = right); string pointer stays owned by prog. */
int bsql_program_scalar(const bsql_program* prog, int op, int side,
                        bsql_scalar* out);
/* Output type of output k, for the given input schema (in_types[n_in]);
 * follows get_output_type (engine/src/parser/expression_utils.cpp:115-266). */
int bsql_program_output_type(const bsql_program* prog, const int32_t* in_types,
                             int n_in, int k, int32_t* type_id);
/* Input columns referenced by a compiled program, in slot order (the
 * col_idx_map of LogicalProject.cpp:859-881); fills up to cap entries. */
int bsql_program_used_columns(const bsql_program* prog, int32_t* cols, int cap);

/* ---- columns built before the interpreter -------------------------------------
 * <=> function_evaluator_transformer + evaluate_string_functions +
 * evaluate_string_case_when_else (engine/src/execution_kernels/
 * LogicalProject.cpp:100-742): functions that build new strings are computed
 * before the interpreter runs and appended to the table, step k of a compiled
 * program being column n_table + k (n_table = the table it was compiled for;
 * the launching calls take that table and build the steps themselves, on the
 * call's stream). A step is an expression evaluated by the interpreter into a
 * column (an argument of a string function, a string CASE's condition) or a
 * string function over columns and literals. */
enum {
  BSQL_STRFN_CONCAT = 1,  /* a || b, null if either is null (cudf::strings::concatenate) */
  BSQL_STRFN_UPPER = 2,   /* full case mapping (cudf::strings::to_upper / to_lower) */
  BSQL_STRFN_LOWER = 3,
  BSQL_STRFN_INITCAP = 4, /* cudf::strings::title */
  BSQL_STRFN_REVERSE = 5, /* characters reversed (slice_strings with step -1) */
  BSQL_STRFN_REPLACE = 6, /* every occurrence of arg 1 by arg 2 (cudf::strings::replace) */
  BSQL_STRFN_CAST = 7,    /* CAST(x AS VARCHAR): from_integers/booleans/floats/timestamps */
  BSQL_STRFN_CASE = 8,    /* arg 0 (BOOL8) true ? arg 1 : arg 2 (cudf::copy_if_else) */
  BSQL_STRFN_FILL = 9,    /* the literal in every row (make_column_from_scalar) */
  BSQL_STRFN_REGEXP_REPLACE = 10 /* every match of arg 1 by arg 2 with \1..\9 (replace_with_backrefs) */
};
typedef struct bsql_step_info {
  int32_t kind;       /* 0: expression (sub, one output); 1: string function (op) */
  int32_t op;         /* BSQL_STRFN_* */
  int32_t out_type;   /* type id of the column built */
  int32_t n_args;
  int32_t arg_col[3]; /* column index, or -1 for a literal, -2 for a null literal */
  const char* arg_str[3];
  int32_t arg_len[3];
  const bsql_program* sub; /* kind 0: compiled over columns [0, n_table + k) */
  int32_t start;           /* REGEXP_REPLACE: characters kept before the search (position - 1) */
} bsql_step_info;
int bsql_program_num_steps(const bsql_program* prog);
int bsql_program_table_columns(const bsql_program* prog);
int bsql_program_step(const bsql_program* prog, int k, bsql_step_info* out);
/* Chars a string output of `prog` over this table can need at most (size
 * string output columns' chars with it). */
int bsql_program_output_chars(const bsql_program* prog, const bsql_column* in, int n_in,
                              int64_t num_rows, int64_t* chars);

/* ---- execution ---------------------------------------------------------------
 * bsql_interp_execute  <=> interops::perform_interpreter_operation
 * (engine/src/Interpreter/interpreter_cpp.cu:440-594). `in` is the interpreter
 * input table (slot k = in[k]); `out[k]` receives final output k, written with
 * the column's own type, and its validity words (outputs must have a validity
 * buffer, offset 0). Rows [0, num_rows).                                      */
int bsql_interp_execute(const bsql_program* prog, const bsql_column* in,
                        int n_in, bsql_column* out, int n_out, int64_t num_rows,
                        void* stream);

/* bsql_filter <=> ral::processor::process_filter + applyBooleanFilter
 * (engine/src/execution_kernels/LogicalFilter.cpp:22-52): `pred` must have one
 * BOOL8 output. Every column of `in` is compacted stably into `out[k]`
 * (same type; capacity >= num_rows rows; validity written iff in[k] has one).
 * A null or false predicate drops the row. The selected row count is written
 * to the device int64 `*d_count` (no host sync).                              */
int bsql_filter(const bsql_program* pred, const bsql_column* in, int n_in,
                bsql_column* out, int64_t num_rows, int64_t* d_count,
                void* stream);

/* bsql_filter_project: fused WHERE + SELECT in one pass. Equivalent to
 * process_filter followed by process_project on the filtered table
 * (LogicalFilter.cpp:31-52 then LogicalProject.cpp:943-969), whose
 * expressions `proj` refer to the same input slots. Output k (capacity >=
 * num_rows, validity buffer required) receives projection output k for the
 * selected rows, in row order; *d_count receives the selected row count.   */
int bsql_filter_project(const bsql_program* pred, const bsql_program* proj,
                        const bsql_column* in, int n_in, bsql_column* out,
                        int n_out, int64_t num_rows, int64_t* d_count,
                        void* stream);

/* bsql_filter_project for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments and results as
 * bsql_filter_project; every buffer beyond the inputs and outputs lives in a
 * fixed 128 MiB device array per GPU (held until the process ends), so a
 * call allocates nothing, copies nothing from host memory and can be
 * captured in a graph. Limits: the scratch must hold the selection and 4
 * bytes of row id per input row (about 32M rows; more fails with
 * BSQL_ERR_INVALID; a predicate launched with BSQL_LAUNCH_FUSED needs no row
 * ids), at most 64 fixed-width outputs, no programs with string literals or
 * string-building functions. Calls that overlap on one GPU share the array
 * and race. */
int bsql_filter_project_noalloc(const bsql_program* pred, const bsql_program* proj,
                                const bsql_column* in, int n_in, bsql_column* out,
                                int n_out, int64_t num_rows, int64_t* d_count,
                                void* stream);

/* Row count of the filter only (the selection kernel; nothing is gathered):
 * the count process_filter's output would have. */
int bsql_filter_count(const bsql_program* pred, const bsql_column* in, int n_in, int64_t num_rows,
                      int64_t* d_count, void* stream);

/* bsql_filter_project into outputs pre-sized for `capacity` rows (e.g. from
 * a predicted selectivity): one pass writes the first min(count, capacity)
 * rows and *d_count receives the full count, so count > capacity tells the
 * caller to run again with outputs of that size. */
int bsql_filter_project_sized(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                              int n_in, bsql_column* out, int n_out, int64_t num_rows, int64_t capacity,
                              int64_t* d_count, void* stream);

/* Several filter + project queries over the same table in one pass; no
 * counterpart in the reference, whose executor runs each query's Filter and
 * Projection kernels on its own. For callers with many predicates over one
 * table (the learned form of several queries, dashboards): query q is
 * preds[q] then projs[q], writes out[q][0 .. n_out[q]) up to capacity[q]
 * rows and its full count to d_counts[q], exactly as bsql_filter_project_sized
 * would. When every query is a range-test predicate with fixed-width outputs
 * and no string-building steps, one kernel generated for all of them runs
 * tile by tile over the table, each query reading the tile's columns soon
 * after the queries before it (hipRTC compile on first use of the set's
 * shape); otherwise the queries run one call each. At most 16 queries. */
int bsql_filter_project_batch(const bsql_program* const* preds, const bsql_program* const* projs,
                              int n_queries, const bsql_column* in, int n_in, bsql_column* const* out,
                              const int32_t* n_out, int64_t num_rows, const int64_t* capacity,
                              int64_t* d_counts, void* stream);

/* Compile-only check of bsql_filter_project_batch's generated kernel for
 * `arch` (no GPU needed; nothing is launched): BSQL_OK and the code size, or
 * BSQL_ERR_INVALID when the batch would not run as one generated kernel. */
int bsql_generated_batch_kernel_check(const bsql_program* const* preds, const bsql_program* const* projs,
                                      int n_queries, const bsql_column* in, int n_in,
                                      bsql_column* const* out, const int32_t* n_out, int64_t num_rows,
                                      const char* arch, int64_t* code_bytes);

/* Conjunct order: for a predicate whose root is a left-deep AND chain (the
 * shape expand_if_logical_op produces, CalciteExpressionParsing.cpp:330-398),
 * evaluate its conjuncts in `order` (a permutation of 0..n_conj-1) with a
 * wave-uniform early exit after each conjunct. The selected rows are the same
 * for every order. order == NULL restores the static (plan) order.         */
int bsql_program_num_conjuncts(const bsql_program* pred);
int bsql_program_set_conjunct_order(bsql_program* pred, const int32_t* order,
                                    int n_conj);

/* ---- launch configuration (autotuner hook) ------------------------------ */
enum {
  BSQL_LAUNCH_NO_RANGES = 1,  /* predicates: evaluate column-vs-literal conjuncts
                                 with the general ops instead of range tests */
  BSQL_LAUNCH_LAZY_LOADS = 2, /* predicates: load a conjunct's columns only when
                                 the wave reaches it (default: as early as
                                 registers allow) */
  BSQL_LAUNCH_SPILL_WIDE = 4, /* spill layout: keep at most 7 live values in
                                 registers (the 8-register file, up to 4 rows
                                 per lane) and the rest in LDS slots; default
                                 15 in registers (16-register file) */
  BSQL_LAUNCH_FUSED = 8,      /* predicates: a filter whose conjuncts are range
                                 tests, with at most 8 fixed-width outputs,
                                 runs as one pass (selection, prefix and
                                 outputs in one kernel, no row ids); default
                                 selection, row ids and output pass */
  BSQL_LAUNCH_GENERATED = 16, /* run through a kernel generated for the
                                 program: its op fields, column types and
                                 validity presence compiled in as constants
                                 (hipRTC, at the first launch of each program
                                 shape on each device; kept for the process),
                                 literals, columns and range bounds passed at
                                 launch. bsql_interp_execute: any program;
                                 filters with a projection (set on the
                                 predicate): a range-test predicate with
                                 fixed-width outputs and no RAND runs as one
                                 pass. Other calls run on the interpreter. */
  BSQL_LAUNCH_LDS_REGS = 32   /* the interpreter's register file in LDS (one
                                 address add and one LDS access per operand)
                                 instead of VGPRs (indexed register moves):
                                 batched launches, bsql_interp_execute and a
                                 filter's output phase (set on either
                                 program). A compiled program starts with
                                 this flag set; a configuration without it
                                 puts the file in VGPRs. */
};
typedef struct bsql_launch_config {
  int32_t block_size;     /* threads per workgroup: 0 (auto), 64 or 256 */
  int32_t rows_per_thread;/* 0 (auto), 1, 2, 4; range-test selection: 8, 16, 32 */
  int32_t flags;          /* BSQL_LAUNCH_* */
} bsql_launch_config;
/* Sets the configuration used for launches with this program (a predicate's
 * flags apply to filters it drives; 0 fields = auto). No counterpart in the
 * reference, which sizes its launch with the occupancy API
 * (interpreter_cpp.cu:76-270, :473). */
int bsql_program_set_launch(bsql_program* prog, const bsql_launch_config* cfg);

/* Kernels generated so far in this process (BSQL_LAUNCH_GENERATED), launches
 * that ran one, and the seconds spent compiling them. No counterpart in the
 * reference, which interprets every program. */
int bsql_generated_kernel_stats(int64_t* kernels, int64_t* launches, double* compile_seconds);
/* Generates and compiles, for architecture `arch` (e.g. "gfx950"), the
 * kernels that bsql_filter_project(pred, proj, ...) (pred non-NULL) or
 * bsql_interp_execute(proj, ...) (pred NULL) would run with
 * BSQL_LAUNCH_GENERATED over these columns (only their types, validity
 * presence and alignment are read), without a device: loads and launches
 * nothing. BSQL_ERR_INVALID if the call would run on the interpreter;
 * *code_bytes = size of the code objects. */
int bsql_generated_kernel_check(const bsql_program* pred, const bsql_program* proj, const bsql_column* in,
                                int n_in, const bsql_column* out, int n_out, int64_t num_rows, const char* arch,
                                int64_t* code_bytes);

/* Shape of the device plan a program lowers to for bsql_interp_execute over
 * a table of these column types (its lowering runs on the host, so this
 * needs no device): device ops, register file (registers + the spare), LDS
 * spill slots and the spill/fill ops, distinct input columns, load issues
 * and the most columns one issue keeps in flight. No counterpart in the
 * reference (its slots are LDS cells of the program's own numbering). */
typedef struct bsql_plan_info {
  int32_t n_ops, n_regs, n_spill, n_spill_ops, n_fill_ops, n_inputs, n_load_issues, max_loads_in_flight;
} bsql_plan_info;
int bsql_program_plan_info(const bsql_program* prog, const int32_t* in_types, int n_in, bsql_plan_info* info);

/* ---- batched execution ----------------------------------------------------
 * Many batches, each with its own program and columns (the uneven batches an
 * execution graph hands to ProjectKernel one task at a time,
 * BatchProcessing.cpp:365-448), run as ONE persistent launch: workgroups
 * stride over every batch's tiles. bsql_batch_prepare lowers each program
 * and uploads all plans (synchronous, allocates); bsql_batch_launch then
 * allocates and copies nothing, so it can be captured in a hipGraph and
 * replayed. Outputs must be fixed-width. Replaces one
 * perform_interpreter_operation call per batch (interpreter_cpp.h:77-86).
 * The prepared launch keeps the register file in LDS with every program on
 * the 8-register file (BSQL_LAUNCH_LDS_REGS | BSQL_LAUNCH_SPILL_WIDE) where
 * the programs' spill slots fit, else the VGPR file; bsql_batch_set_launch
 * and bsql_batch_autotune change it. */
typedef struct bsql_batch_desc {
  const bsql_program* prog;
  const bsql_column* in;
  int n_in;
  const bsql_column* out;
  int n_out;
  int64_t num_rows;
} bsql_batch_desc;
typedef struct bsql_batch bsql_batch;
int bsql_batch_prepare(const bsql_batch_desc* descs, int n, bsql_batch** batch);
int bsql_batch_launch(const bsql_batch* batch, void* stream);
/* Launch shape of the batched launch: block size, rows per lane and the
 * flags BSQL_LAUNCH_LDS_REGS (register file in LDS) and
 * BSQL_LAUNCH_SPILL_WIDE (every program lowered for the 8-register file,
 * values beyond 7 live in LDS spill slots; without it one program that needs
 * more than 8 registers puts the whole batch on the 16-register file).
 * Selects the work map of that shape (built on first use; not for use while
 * a graph holding the launch may run). */
int bsql_batch_set_launch(bsql_batch* batch, const bsql_launch_config* cfg);
/* Autotuner: times each launch shape, register file place and (where they
 * differ) lowering form on `stream` (a few launches each, hipEvents), keeps
 * the fastest and returns it. */
int bsql_batch_autotune(bsql_batch* batch, void* stream, bsql_launch_config* chosen);
int64_t bsql_batch_total_rows(const bsql_batch* batch);
void bsql_batch_free(bsql_batch* batch);

/* Autotuner for one program: times bsql_interp_execute on these columns for
 * each launch shape, sets the fastest on the program and returns it. */
int bsql_program_autotune(bsql_program* prog, const bsql_column* in, int n_in, bsql_column* out,
                          int n_out, int64_t num_rows, void* stream, bsql_launch_config* chosen);

/* ---- learned conjunct ordering -------------------------------------------
 * No counterpart in the reference, which evaluates conjuncts in plan order.
 * A small FP32 MLP predicts, for each range conjunct of a predicate, its
 * selectivity (sigmoid of output 0) and its log cost (output 1) from the
 * sketch of the conjunct's column; bsql_order_by_prediction then evaluates
 * cheap, selective conjuncts first. The rows selected do not depend on the
 * order; the plan order stays the default. */
#define BSQL_SKETCH_DIM 19    /* null fraction, min, max, 16-bin histogram */
#define BSQL_FEATURE_DIM 24
#define BSQL_MLP_HIDDEN 64
/* Sketch of each column over a 1024-row evenly spaced sample (device output
 * of n_cols * BSQL_SKETCH_DIM floats; synchronizes the stream). */
int bsql_sketch(const bsql_column* cols, int n_cols, int64_t num_rows, float* d_sketch, void* stream);
typedef struct bsql_mlp bsql_mlp;   /* BSQL_FEATURE_DIM -> 64 -> 64 -> 2, ReLU */
int bsql_mlp_create(uint64_t seed, bsql_mlp** mlp);
int bsql_mlp_num_params(const bsql_mlp* mlp);
int bsql_mlp_get_params(const bsql_mlp* mlp, float* host_params);
int bsql_mlp_set_params(bsql_mlp* mlp, const float* host_params);
/* d_x: batch x BSQL_FEATURE_DIM, d_y: batch x 2 (device) */
int bsql_mlp_forward(bsql_mlp* mlp, const float* d_x, int batch, float* d_y, void* stream);
/* Gradient of the mean loss (binary cross-entropy of output 0 against the
 * observed selectivity + cost_weight * squared error of output 1 against the
 * observed log cost; targets batch x 2) into d_grads, per-sample losses in
 * d_loss. */
int bsql_mlp_gradients(bsql_mlp* mlp, const float* d_x, const float* d_target, int batch, float cost_weight,
                       float* d_grads, float* d_loss, void* stream);
/* One Adam step on that gradient. */
int bsql_mlp_train_step(bsql_mlp* mlp, const float* d_x, const float* d_target, int batch, float lr,
                        float cost_weight, float* d_loss, void* stream);
void bsql_mlp_free(bsql_mlp* mlp);
/* Features (host, cap x BSQL_FEATURE_DIM) of each conjunct of a predicate
 * whose conjuncts are all column-vs-literal comparisons, from the host copy
 * of the table's sketch; returns the number of conjuncts (written only if
 * cap is large enough), or minus an error code. */
int bsql_conjunct_features(const bsql_program* pred, const int32_t* in_types, int n_in, const float* sketch,
                           float* features, int cap);
/* Sets the conjunct order: ascending cost / (1 - selectivity). */
int bsql_order_by_prediction(bsql_program* pred, const float* selectivity, const float* cost, int n_conj);
/* Sets the conjunct order of least expected cost for the range-test
 * selection, over the table schema in_types: a conjunct after others that
 * pass fraction p of the rows costs cost * (fraction of its 128-byte lines
 * with a live row + 0.2 * fraction of 1024-row slices alive), selectivities
 * taken as independent; exact over all orders (up to 16 conjuncts; more, or
 * conjuncts that are not all range tests: the rank rule above). The order
 * is also written to order_out (n_conj entries) unless NULL. No counterpart
 * in the reference (its plans keep Calcite's order). */
int bsql_order_conjuncts(bsql_program* pred, const int32_t* in_types, int n_in, const float* selectivity,
                         const float* cost, int n_conj, int32_t* order_out);

/* ---- front-end text stages ------------------------------------------------
 * bsql_expression_rewrite returns one text stage of the expression front end
 * (the rewrites bsql_program_compile applies), for tests and tooling:
 *   0 fill_minus_op_with_zero        engine/src/parser/expression_utils.cpp:1155-1164
 *   1 convert_nary_to_binary_concat  expression_utils.cpp:1520-1622
 *   2 convert_ms_to_ns_units         expression_utils.cpp:1316-1350
 *   3 reinterpret_timestamp          expression_utils.cpp:1371-1421 (types: schema)
 *   4 apply_interval_conversion      expression_utils.cpp:1425-1469 (types: schema)
 *   5 preprocess_expression (0..4 and CURRENT_DATE)   expression_utils.cpp:1624-1632
 *   6 replace_calcite_regex          expression_utils.cpp:1068-1113
 *   7 6 then expand_if_logical_op    CalciteExpressionParsing.cpp:330-398
 *   8 parse_tree::build + transform_to_custom_op + rebuildExpression
 *                                    engine/src/parser/expression_tree.hpp:147-162, 181-258
 *   9 modify_multi_column_count_expression (COUNT over several columns)
 *                                    expression_utils.cpp:1467-1492
 * Writes at most cap-1 bytes and a NUL, and the full length to *len. */
int bsql_expression_rewrite(const char* expr, int stage, const int32_t* types, int n_types, char* out,
                            size_t cap, size_t* len);
/* The query's timestamp that CURRENT_DATE / CURRENT_TIMESTAMP / CURRENT_TIME
 * stand for, "YYYY-MM-DD hh:mm:ss.mmm" (Context::getCurrentTimestamp in the
 * reference); NULL or "" uses the clock at compile time. */
int bsql_set_current_timestamp(const char* ts);

/* ---- diagnostics --------------------------------------------------------- */
/* Message of the last error on the calling thread. */
const char* bsql_last_error(void);
int bsql_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BSQL_INTERP_H */
