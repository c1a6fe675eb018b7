// strfn.h -- launch arguments of the string-building kernels (strfn.hip),
// shared by the host launcher and the device code.
//
// These kernels build the string columns that the reference computes with
// libcudf before its interpreter runs (engine/src/execution_kernels/
// LogicalProject.cpp:100-742, appended to the table as $N+k by
// function_evaluator_transformer, :670-731): CONCAT, UPPER, LOWER, INITCAP,
// REVERSE, REPLACE, CAST(x AS VARCHAR), the string CASE (copy_if_else) and a
// string literal filling a column (make_column_from_scalar, :822-832).
#pragma once

#include <stdint.h>

#include "../../include/bsql/interp.h"

namespace bsql {

enum StrFnOp : int32_t {
  SF_CONCAT = BSQL_STRFN_CONCAT,    // a || b; null if either is null
  SF_UPPER = BSQL_STRFN_UPPER,      // full case mapping of the BMP, other characters unchanged
  SF_LOWER = BSQL_STRFN_LOWER,
  SF_INITCAP = BSQL_STRFN_INITCAP,  // first letter of each run of letters upper, the rest lower
  SF_REVERSE = BSQL_STRFN_REVERSE,  // characters in reverse order
  SF_REPLACE = BSQL_STRFN_REPLACE,  // every non-overlapping occurrence of a[1] replaced by a[2]
  SF_CAST = BSQL_STRFN_CAST,        // a[0] (integer, bool, float, timestamp or string) as text
  SF_CASE = BSQL_STRFN_CASE,        // a[0] (BOOL8) valid and true ? a[1] : a[2]
  SF_FILL = BSQL_STRFN_FILL,        // the literal a[0] in every row
  SF_REGEXP_REPLACE = BSQL_STRFN_REGEXP_REPLACE,  // a[1] (compiled regex) -> a[2], from char `start`
};

enum StrSrcKind : int32_t { SRC_STRING = 0, SRC_LITERAL = 1, SRC_NULL = 2, SRC_FIXED = 3 };

// One argument: a string column, a literal, a null literal or a fixed-width
// column. Column pointers are already advanced to the column's row offset;
// validity bit of row r is bit (r + vbit) of valid[].
struct StrSrc {
  int32_t kind;
  int32_t type;
  const int32_t* offsets;
  const uint8_t* chars;
  const void* data;
  const uint32_t* valid;
  int32_t vbit;
  int32_t lit_len;
  const uint8_t* lit;
};

struct StrFnArgs {
  int32_t op;
  int32_t n_args;
  StrSrc a[3];
  int64_t num_rows;
  int32_t start;               // REGEXP_REPLACE: characters kept before the search
  uint64_t* lens;              // length pass: row length << 32 (null rows: 0)
  uint32_t* out_valid;         // length pass: validity words of the result
  const int32_t* out_offsets;  // write pass
  uint8_t* out_chars;          // write pass
  int64_t chars_cap;           // bytes of out_chars
};

}  // namespace bsql
