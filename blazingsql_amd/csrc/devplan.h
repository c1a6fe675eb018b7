// devplan.h -- the lowered program the gfx950 VM kernel executes.
//
// Shared by the host lowering (program.cpp) and the HIP kernels
// (interp_kernels.hip). The plan and its op list live in device memory and
// are read through the scalar cache with wave-uniform loads.
#pragma once

#include <stdint.h>

// Columns one launch reads (the distinct table columns its ops and range
// tests refer to, densely renumbered) and writes. The reference's programs
// address at most 64 slots (inputs, outputs and temporaries together), so a
// predicate and a projection part fit one launch; longer output lists run as
// several output launches over the same selection.
#define BSQL_DEV_MAX_IN 128
#define BSQL_DEV_MAX_OUT 64
// Register file sizes of the kernel variants (the last register is the spare
// that result-less ops write); live values beyond the file are spilled to
// LDS slots.
#define BSQL_DEV_MAX_REG 16

// Device opcodes: reference operators specialised by operand kind
// (I = int64 slot, F = double slot), plus the VM's own load/store/select ops.
enum DevCode : uint8_t {
  D_NOP = 0,
  D_LOAD,        // reg[dst] <- input column a (converted to the slot kind), validity
  D_LOADS,       // reg[dst] <- slice of string column a: start byte | byte length << 32
  D_LOADN,       // issue loads of aux (1..4) columns (imm byte j = column j) into temporaries
  D_COMMIT,      // reg[dst] <- loaded column aux of the last LOADN (data and validity)
  D_STORE,       // output column aux <- reg[a] (static_cast to its type)
  D_KILL,        // alive &= valid(a) && truthy(a); wave-uniform early exit
  D_SEL,         // alive &= valid(a) && bits(a) != 0 (BOOL8 write of the slot)
  D_ADD_II, D_SUB_II, D_MUL_II, D_DIV_II, D_MOD_II,
  D_ADD_FF, D_SUB_FF, D_MUL_FF, D_DIV_FF, D_MOD_FF,
  D_POW, D_ROUND,
  D_EQ_II, D_NE_II, D_LT_II, D_GT_II, D_LE_II, D_GE_II,
  D_EQ_FF, D_NE_FF, D_LT_FF, D_GT_FF, D_LE_FF, D_GE_FF,
  D_EQ_TS, D_NE_TS, D_LT_TS, D_GT_TS, D_LE_TS, D_GE_TS,
  D_EQ_SS, D_NE_SS, D_LT_SS, D_GT_SS, D_LE_SS, D_GE_SS,
  D_AND, D_OR, D_INDF_II, D_INDF_FF, D_MAGIC_IF_NOT, D_FIRST_NON_MAGIC,
  D_COALESCE,    // FIRST_NON_MAGIC(MAGIC_IF_NULL(a), b): COALESCE(a, b) with the magic-value rule
  D_NOT,
  D_MAGIC_IF_NULL,  // valid(a) ? a : the magic value of a's kind; valid(a) (IS_NOT_NULL + MAGIC_IF_NOT fused)
  D_ABS_I, D_ABS_F, D_FLOOR, D_CEIL, D_SIN, D_COS, D_ASIN, D_ACOS,
  D_TAN, D_COTAN, D_ATAN, D_LN, D_LOG, D_EXTRACT,
  D_IS_NULL, D_IS_NOT_NULL, D_IS_TRUE, D_IS_NOT_TRUE, D_IS_NOT_FALSE,
  D_TO_I64, D_TO_F64, D_CAST_DATE, D_CAST_TS, D_CHAR_LENGTH, D_RAND,
  // string slices (a = slice of the chars of column aux): LIKE against the
  // compiled pattern at imm; SLICE with aux2 = 0 from char b (0-based), 1 the
  // first b chars, 2 the last b chars
  D_LIKE, D_SLICE,
  // LIKE 'abc%' (aux2 = 0) or LIKE 'abc' (aux2 = 1) with a literal of at most
  // 7 bytes: imm holds the bytes (byte j at bits 8j) and their count at bits
  // 56-63; a = slice of the chars of column aux
  D_LIKE_PREFIX,
  // TRIM: the slice without the leading (aux2 bit 0) / trailing (bit 1)
  // characters found in the literal at imm ({int32 len, bytes}; empty: those
  // <= ' '); STR_TO_I64 / STR_TO_F64: the slice parsed as a number
  D_TRIM, D_STR_TO_I64, D_STR_TO_F64,
  // STR_TO_TS: the slice parsed as a timestamp of type aux2 by the format
  // items at imm ({n, (spec << 8 | width)...}; 0: "%Y-%m-%d" for days,
  // "%Y-%m-%d %H:%M:%S" otherwise)
  D_STR_TO_TS,
  // register <-> LDS spill slot imm (data of the lane's rows and their validity)
  D_SPILL,       // slot imm <- reg[a]
  D_FILL,        // reg[dst] <- slot imm
  D_NUM_CODES
};

// DevOp.flags
enum : uint8_t {
  F_A_IMM = 1,    // operand a is the literal in imm (already in the op's domain)
  F_B_IMM = 2,    // operand b is the literal in imm
  F_A_NULL = 4,   // operand a is a null literal
  F_B_NULL = 8,   // operand b is a null literal
  F_A_F = 16,     // operand a holds double bits (else int64)
  F_B_F = 32,     // operand b holds double bits
  F_STORE_VALID = 64,  // D_STORE: the output column has a validity buffer
  F_SLICE_LIT = 128,   // D_SLICE: the position is the (already clamped) literal in imm
  // ops on doubles (the *_FF codes, POW, ROUND, FLOOR, CEIL and the
  // transcendentals): operand a / b is an int64 slot read as a double
  F_A_CVT = 64,
  F_B_CVT = 128
};

typedef struct DevOp {
  uint8_t code;
  uint8_t dst;   // destination register (0xff: no result)
  uint8_t a;     // register of operand a (D_LOAD/D_LOADS: input column)
  uint8_t b;     // register of operand b
  uint8_t flags;
  uint8_t aux;   // D_STORE: output column; D_EXTRACT: operator; string ops: column whose
                 // chars a slices (0xff: a is a literal)
  uint8_t aux2;  // *_TS: (unit of a) | (unit of b) << 4; *_SS: column of b or 0xff;
                 // D_EXTRACT/D_CAST_DATE/D_CAST_TS: source type id; D_SLICE: kind
  uint8_t pad;
  int64_t imm;   // literal bits; *_SS: device pointer to {int32 len, bytes};
                 // D_LIKE: device pointer to the compiled pattern (like_compile)
} DevOp;

// Every field the kernels read is 32 or 64 bits wide: gfx950 has no scalar
// byte loads, and a narrow field read by a whole wave would become a vector
// load (and a wait on every load in flight).

// Input column as the kernel sees it: data already advanced to the view's
// first row; validity pointer advanced by whole words, vbit = remaining bit.
typedef struct DevInCol {
  const void* data;
  const uint32_t* valid;  // NULL = all valid
  const int32_t* offsets; // strings
  int32_t vbit;
  int32_t type;
  int32_t aligned16;      // data is 16-byte aligned
  int32_t pad;
} DevInCol;

typedef struct DevOutCol {
  void* data;
  uint32_t* valid;
  int32_t type;
  int32_t pad;
} DevOutCol;

// One conjunct (or several merged conjuncts on one column) of a predicate
// made of column-vs-literal comparisons: column `col` must be valid and lie
// in [lo, hi] (inclusive bounds, computed on the host from the comparison
// operators), or with RK_NE differ from lo. RK_F: the comparison domain is
// double (the column converted as the float ops convert it), else int64.
// RK_EMPTY: no value passes. In the int64 domain the column value is first
// multiplied by `mul` (wrapping): timestamps compare as nanoseconds.
#define BSQL_DEV_MAX_RANGES 16
enum : int32_t { RK_F = 1, RK_NE = 2, RK_EMPTY = 4 };
typedef struct DevRange {
  int32_t col;
  int32_t kind;
  int64_t lo;  // int64 or double bits (RK_F)
  int64_t hi;
  int64_t mul;
} DevRange;

typedef struct DevPlan {
  int32_t n_ops;
  int32_t split;     // ops [0, split) decide the row selection, [split, n_ops) write outputs
  int32_t n_in;
  int32_t n_out;
  int64_t num_rows;
  int32_t n_ranges;  // > 0: the selection is these range conjuncts (ops [0, split) unused)
  int32_t range_eager;  // load every range column before the first test
  int32_t n_spill;   // LDS spill slots per lane (0: the program fits the register file)
  int32_t pad;
  const DevOp* ops;  // n_ops ops (at least one entry), device memory
  DevInCol in[BSQL_DEV_MAX_IN];
  DevOutCol out[BSQL_DEV_MAX_OUT];
  DevRange ranges[BSQL_DEV_MAX_RANGES];
} DevPlan;

// Arguments of the one-pass filter + project kernels (the tile status words
// and tickets live in the call's scratch).
typedef struct FusedArgs {
  uint64_t* status;   // num_tiles words, zeroed before the launch
  uint32_t* ticket;   // zeroed before the launch (kernels that take tiles from tickets)
  int64_t* d_count;   // selected rows (written by the last tile)
  int64_t capacity;   // rows the outputs hold; positions beyond are not written
  int64_t num_tiles;
} FusedArgs;

// The one-pass filter + project kernel (filter_project_rows) stages up to
// `cap` selected rows of its tile in LDS: one 8-byte slot per row and
// output, and per output fp_null_words(cap) words of null bits (one word
// more than the rows need, so a 32-bit window at any row offset reads two
// words). Its dynamic LDS: n_out x cap slots, n_out x fp_null_words(cap)
// null words, then `waves` x n_out x vw validity words (tiles that selected
// more rows than `cap` write them from the registers).
static constexpr inline int fp_null_words(int cap) { return cap > 0 ? cap / 32 + 1 : 0; }
static constexpr inline int64_t fp_lds_bytes(int n_out, int cap, int waves, int vw) {
  return (int64_t)n_out * cap * 8 + (int64_t)n_out * fp_null_words(cap) * 4 + (int64_t)waves * n_out * vw * 4;
}

// Arguments of the launch that runs several filter + project queries over
// one table (bsql_filter_project_batch): each query's plan and one-pass
// arguments.
#define BSQL_MAX_BATCH_QUERIES 16
typedef struct MultiArgs {
  const DevPlan* plan[BSQL_MAX_BATCH_QUERIES];
  FusedArgs fa[BSQL_MAX_BATCH_QUERIES];
} MultiArgs;

// LDS bytes of the spill slots of one workgroup: per wave and slot, the
// 64 lanes x R rows of 8-byte data, then one validity byte per lane.
static inline int64_t spill_lds_bytes(int n_spill, int block, int rows) {
  return (int64_t)n_spill * (block / 64) * (64 * 8 * rows + 64);
}
