// interp_kernels.hip -- gfx950 kernels of the row-parallel expression
// interpreter, its boolean-mask stream compaction, and the fused
// filter+project pass.
//
// Execution model (replaces transformKernel / InterpreterFunctor,
// engine/src/Interpreter/interpreter_ops.cuh:165-964):
//   * a workgroup of 256 lanes owns a tile of 256*R consecutive rows; lane t
//     owns rows [t*R, t*R+R) of the tile, so every column load is one
//     contiguous 4R- or 8R-byte vector per lane (128-bit loads for R >= 4
//     int32 or R >= 2 int64/double) and a wave reads one contiguous span;
//   * the program is wave-uniform: it is read from the kernel argument block
//     through the scalar cache and dispatched with scalar branches;
//   * the slot buffer of the reference (8 bytes per slot per row, in LDS) is a
//     register file reg[NREG][R] in VGPRs, indexed with the wave-uniform
//     register number (s_set_gpr_idx relative addressing); validity is one bit
//     per register per row;
//   * compaction: per-row selection bits -> wave ballots -> mbcnt prefix ->
//     workgroup prefix in LDS -> decoupled look-back across tiles (tiles are
//     taken from an atomic ticket so every predecessor is resident or done);
//     output validity words are packed in LDS per wave and written with plain
//     stores (interior words) or disjoint-mask atomics (boundary words).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devplan.h"
#include "../../include/bsql/interp.h"

namespace bsql_dev {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

constexpr int64_t MAGIC_I64 = INT64_MAX - 13LL;       // interpreter_ops.cuh:41-43
constexpr double MAGIC_F64 = 1.7976931348623123e+308;  // interpreter_ops.cuh:46-48

__device__ __forceinline__ double as_f(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t bits(double d) { return (uint64_t)__double_as_longlong(d); }

// operand as double: converts int64 slots, reinterprets double slots
__device__ __forceinline__ double to_f(uint64_t b, bool is_f) {
  return is_f ? as_f(b) : (double)(int64_t)b;
}
__device__ __forceinline__ bool truthy(uint64_t b, bool is_f) {
  return is_f ? (as_f(b) != 0.0) : (b != 0);
}
// double -> int64 with NaN/out-of-range -> INT64_MIN (one definition for
// the host oracle and the device; the reference leaves it undefined)
__device__ __forceinline__ int64_t f64_to_i64(double v) {
  if (!(v >= -9223372036854775808.0 && v < 9223372036854775808.0)) return INT64_MIN;
  return (int64_t)v;
}

__device__ __forceinline__ int64_t ticks_per_day(int t) {
  switch (t) {
    case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return 1;
    case BSQL_TIMESTAMP_SECONDS: case BSQL_DURATION_SECONDS: return 86400LL;
    case BSQL_TIMESTAMP_MILLISECONDS: case BSQL_DURATION_MILLISECONDS: return 86400000LL;
    case BSQL_TIMESTAMP_MICROSECONDS: case BSQL_DURATION_MICROSECONDS: return 86400000000LL;
    default: return 86400000000000LL;
  }
}
__device__ __forceinline__ int64_t ns_per_unit(int u) {  // unit codes 0..4 = D,s,ms,us,ns
  switch (u) {
    case 0: return 86400000000000LL;
    case 1: return 1000000000LL;
    case 2: return 1000000LL;
    case 3: return 1000LL;
    default: return 1LL;
  }
}
__device__ __forceinline__ bool is_ts(int t) {
  return t >= BSQL_TIMESTAMP_DAYS && t <= BSQL_TIMESTAMP_NANOSECONDS;
}

__device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

// EXTRACT (interpreter_ops.cuh:69-99) with Hinnant's civil_from_days
__device__ int64_t extract(int op, int t, int64_t ticks) {
  const int64_t tpd = ticks_per_day(t);
  const int64_t days = floor_div(ticks, tpd);
  int64_t tsm = ticks - days * tpd;
  if (tsm < 0) tsm += tpd;
  if (op == BSQL_OP_HOUR || op == BSQL_OP_MINUTE || op == BSQL_OP_SECOND) {
    const int64_t tph = tpd / 24, tpm = tpd / 1440, tps = tpd / 86400;
    const int64_t hrs = tph ? tsm / tph : 0;
    const int64_t mins = tpm ? (tsm - hrs * tph) / tpm : 0;
    const int64_t secs = tps ? (tsm - hrs * tph - mins * tpm) / tps : 0;
    return op == BSQL_OP_HOUR ? (int16_t)hrs : op == BSQL_OP_MINUTE ? (int16_t)mins : (int16_t)secs;
  }
  if (op == BSQL_OP_DAYOFWEEK) {
    const int64_t wd = days >= -4 ? (days + 4) % 7 : (days + 5) % 7 + 6;
    return wd == 0 ? 7 : wd;
  }
  int64_t z = days + 719468;
  const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const int64_t yy = (int64_t)yoe + era * 400;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  const unsigned d = doy - (153 * mp + 2) / 5 + 1;
  const unsigned m = mp < 10 ? mp + 3 : mp - 9;
  if (op == BSQL_OP_DAY) return (int16_t)d;
  if (op == BSQL_OP_MONTH) return (int16_t)m;
  return (int16_t)(yy + (m <= 2));
}

__device__ __forceinline__ int64_t idiv(int64_t a, int64_t b) {
  if (b == 0) return 0;
  if (b == -1) return (int64_t)(0ULL - (uint64_t)a);
  return a / b;
}
__device__ __forceinline__ int64_t imod(int64_t a, int64_t b) {
  if (b == 0 || b == -1) return 0;
  return a % b;
}

// Philox4x32-10: uniform double in (0, 1] per (seed, row)
__device__ double philox_uniform(uint64_t seed, uint64_t row) {
  uint32_t c0 = (uint32_t)row, c1 = (uint32_t)(row >> 32), c2 = 0, c3 = 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  const uint64_t hi = (((uint64_t)c0 << 32) | c1) >> 11;
  return ((double)hi + 1.0) * (1.0 / 9007199254740992.0);
}

// string compare (cudf::string_view::compare: bytewise, shorter prefix first)
__device__ int str_cmp(const uint8_t* a, int32_t na, const uint8_t* b, int32_t nb) {
  const int32_t n = na < nb ? na : nb;
  for (int32_t i = 0; i < n; ++i) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return na < nb ? -1 : (na > nb ? 1 : 0);
}
__device__ __forceinline__ int64_t cmp_result(int k, int c3) {  // k = op - EQ
  switch (k) {
    case 0: return c3 == 0;
    case 1: return c3 != 0;
    case 2: return c3 < 0;
    case 3: return c3 > 0;
    case 4: return c3 <= 0;
    default: return c3 >= 0;
  }
}

// ------------------------------------------------------------- column I/O

// Loads R consecutive elements of column c starting at row r0 into slot bits.
// `n` = number of those rows that exist (< R only in the last tile).
template <int R>
__device__ __forceinline__ void load_rows(const DevInCol& c, int64_t r0, int n, uint64_t (&v)[R]) {
  const int t = c.type;
  const bool vec = (n == R) && c.aligned16;
#define BSQL_LOAD_CASE(TYPE_ID, CT, EXPR)                                          \
  case TYPE_ID: {                                                                  \
    const CT* p = (const CT*)c.data + r0;                                          \
    CT tmp[R];                                                                     \
    if (vec && (sizeof(CT) * R) % 16 == 0) {                                       \
      _Pragma("unroll") for (int k = 0; k < (int)(sizeof(CT) * R / 16); ++k)       \
        ((v4u32*)tmp)[k] = __builtin_nontemporal_load((const v4u32*)p + k);        \
    } else {                                                                       \
      _Pragma("unroll") for (int k = 0; k < R; ++k) tmp[k] = k < n ? p[k] : CT(0); \
    }                                                                              \
    _Pragma("unroll") for (int k = 0; k < R; ++k) { CT x = tmp[k]; v[k] = (EXPR); } \
    break;                                                                         \
  }
  switch (t) {
    BSQL_LOAD_CASE(BSQL_INT8, int8_t, (uint64_t)(int64_t)x)
    BSQL_LOAD_CASE(BSQL_INT16, int16_t, (uint64_t)(int64_t)x)
    BSQL_LOAD_CASE(BSQL_INT32, int32_t, (uint64_t)(int64_t)x)
    BSQL_LOAD_CASE(BSQL_INT64, int64_t, (uint64_t)x)
    BSQL_LOAD_CASE(BSQL_UINT8, uint8_t, (uint64_t)x)
    BSQL_LOAD_CASE(BSQL_UINT16, uint16_t, (uint64_t)x)
    BSQL_LOAD_CASE(BSQL_UINT32, uint32_t, (uint64_t)x)
    BSQL_LOAD_CASE(BSQL_UINT64, uint64_t, (uint64_t)x)
    BSQL_LOAD_CASE(BSQL_BOOL8, uint8_t, (uint64_t)(x != 0))
    BSQL_LOAD_CASE(BSQL_FLOAT32, float, bits((double)x))
    BSQL_LOAD_CASE(BSQL_FLOAT64, double, bits(x))
    BSQL_LOAD_CASE(BSQL_TIMESTAMP_DAYS, int32_t, (uint64_t)(int64_t)x)
    BSQL_LOAD_CASE(BSQL_DURATION_DAYS, int32_t, (uint64_t)(int64_t)x)
    default: {  // 8-byte ticks: timestamps s/ms/us/ns, durations s/ms/us/ns
      const int64_t* p = (const int64_t*)c.data + r0;
      int64_t tmp[R];
      if (vec && (8 * R) % 16 == 0) {
#pragma unroll
        for (int k = 0; k < 8 * R / 16; ++k)
          ((v4u32*)tmp)[k] = __builtin_nontemporal_load((const v4u32*)p + k);
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) tmp[k] = k < n ? p[k] : 0;
      }
#pragma unroll
      for (int k = 0; k < R; ++k) v[k] = (uint64_t)tmp[k];
      break;
    }
  }
#undef BSQL_LOAD_CASE
}

// validity bits of R consecutive rows (bit k = row r0+k), all-ones without a mask
template <int R>
__device__ __forceinline__ uint32_t load_valid(const DevInCol& c, int64_t r0, int n) {
  if (n <= 0) return 0u;
  if (!c.valid) return n >= R ? (1u << R) - 1u : (1u << n) - 1u;
  const int64_t b = (int64_t)c.vbit + r0;
  const int64_t w = b >> 5;
  const int sh = (int)(b & 31);
  uint64_t word = c.valid[w];
  if (sh + R > 32 && n > 32 - sh) word |= (uint64_t)c.valid[w + 1] << 32;
  uint32_t m = (uint32_t)(word >> sh) & ((1u << R) - 1u);
  return n >= R ? m : (m & ((1u << n) - 1u));
}

// static_cast of slot bits to the output column type (interpreter_ops.cuh:329-395)
__device__ __forceinline__ void store_elem(const DevOutCol& c, int64_t row, uint64_t b) {
  switch (c.type) {
    case BSQL_INT8: case BSQL_UINT8: ((uint8_t*)c.data)[row] = (uint8_t)b; break;
    case BSQL_BOOL8: ((uint8_t*)c.data)[row] = b != 0; break;
    case BSQL_INT16: case BSQL_UINT16: ((uint16_t*)c.data)[row] = (uint16_t)b; break;
    case BSQL_INT32: case BSQL_UINT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
      ((uint32_t*)c.data)[row] = (uint32_t)b; break;
    case BSQL_FLOAT32: ((float*)c.data)[row] = (float)as_f(b); break;
    default: ((uint64_t*)c.data)[row] = b; break;  // int64, uint64, float64, 8-byte ticks
  }
}

// R consecutive rows, vectorised when the tile is full and aligned
template <int R>
__device__ __forceinline__ void store_rows(const DevOutCol& c, int64_t r0, int n, const uint64_t (&v)[R]) {
  if (n == R && (c.type == BSQL_FLOAT64 || c.type == BSQL_INT64) && (R % 2 == 0) &&
      ((((uintptr_t)c.data) & 15) == 0)) {
    uint64_t* p = (uint64_t*)c.data + r0;
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      v4u32 q;
      q.x = (uint32_t)v[2 * k]; q.y = (uint32_t)(v[2 * k] >> 32);
      q.z = (uint32_t)v[2 * k + 1]; q.w = (uint32_t)(v[2 * k + 1] >> 32);
      __builtin_nontemporal_store(q, (v4u32*)p + k);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (k < n) store_elem(c, r0 + k, v[k]);
}

// ------------------------------------------------------------ the VM core
//
// The register file: rv[k][reg] holds the slot bits of register reg for
// this lane's row k (one array per row, each indexed by the register number); rvalid[k] has one validity bit per register. One VM step
// reads the register file at exactly two sites (operands a and b) and writes
// it at one (the result), all with the wave-uniform register number, so the
// compiler keeps the file in VGPRs; the op itself works on plain locals.

// One register array per row handled by the lane: separate allocas, each
// indexed only by the wave-uniform register number, are what the compiler
// promotes to VGPRs (relative addressing) instead of scratch.
template <int NREG, int R>
struct RegFile {
  uint64_t (&r0)[NREG];
  uint64_t (&r1)[NREG];
  uint64_t (&r2)[NREG];
  uint64_t (&r3)[NREG];
  uint64_t (&r4)[NREG];
  uint64_t (&r5)[NREG];
  uint64_t (&r6)[NREG];
  uint64_t (&r7)[NREG];
  __device__ __forceinline__ uint64_t& at(int k, int reg) const {
    switch (k) {
      case 0: return r0[reg];
      case 1: return r1[reg];
      case 2: return r2[reg];
      case 3: return r3[reg];
      case 4: return r4[reg];
      case 5: return r5[reg];
      case 6: return r6[reg];
      default: return r7[reg];
    }
  }
};
#define BSQL_REGFILE(NAME)                                                               \
  uint64_t NAME##0[NREG], NAME##1[NREG], NAME##2[NREG], NAME##3[NREG], NAME##4[NREG],      \
      NAME##5[NREG], NAME##6[NREG], NAME##7[NREG];                                         \
  RegFile<NREG, R> NAME{NAME##0, NAME##1, NAME##2, NAME##3, NAME##4, NAME##5, NAME##6, NAME##7}

struct Operands {
  uint64_t a, b;
  uint32_t va, vb;
};

#define FA ((op.flags & F_A_F) != 0)
#define FB ((op.flags & F_B_F) != 0)

// Transcendental and calendar ops: kept out of line so that their large
// register footprint does not lower the occupancy of every program.
__device__ __noinline__ uint64_t compute_heavy(uint8_t code, uint8_t aux, uint8_t aux2, int64_t imm,
                                               double af, double bf, uint64_t a, int64_t row) {
  switch (code) {
    case D_MOD_FF: return bits(fmod(af, bf));
    case D_POW: return bits(pow(af, bf));
    case D_ROUND: {
      const double factor = pow(10.0, bf);
      return bits(__ddiv_rn(round(__dmul_rn(af, factor)), factor));
    }
    case D_SIN: return bits(sin(af));
    case D_COS: return bits(cos(af));
    case D_ASIN: return bits(asin(af));
    case D_ACOS: return bits(acos(af));
    case D_TAN: return bits(tan(af));
    case D_COTAN: return bits(__ddiv_rn(cos(af), sin(af)));
    case D_ATAN: return bits(atan(af));
    case D_LN: return bits(log(af));
    case D_LOG: return bits(log10(af));
    case D_EXTRACT: return is_ts(aux2) ? (uint64_t)extract(aux, aux2, (int64_t)a) : 0;
    case D_RAND: return bits(philox_uniform((uint64_t)imm, (uint64_t)row));
    default: return 0;
  }
}

__device__ __forceinline__ bool is_heavy(uint8_t code) {
  return code == D_MOD_FF || code == D_POW || code == D_ROUND || (code >= D_SIN && code <= D_EXTRACT) ||
         code == D_RAND;
}

// Computes the result bits/validity of `op` on one row (binary and unary ops
// that only need the operand values). Returns false if the op has no result.
__device__ __forceinline__ void compute(const DevOp& op, const Operands& o, uint64_t& d, uint32_t& vd,
                                        int64_t row) {
  const uint64_t a = o.a, b = o.b;
  const bool both = o.va && o.vb;
  const double af = to_f(a, FA), bf = to_f(b, FB);
  if (is_heavy(op.code)) {
    d = compute_heavy(op.code, op.aux, op.aux2, op.imm, af, bf, a, row);
    vd = op.code == D_RAND ? 1u : (op.code >= D_SIN ? o.va : (uint32_t)both);
    return;
  }
  switch (op.code) {
    // binary ops valid only if both operands are (interpreter_ops.cuh:599-727)
    case D_ADD_II: d = a + b; vd = both; return;
    case D_SUB_II: d = a - b; vd = both; return;
    case D_MUL_II: d = a * b; vd = both; return;
    case D_DIV_II: d = (uint64_t)idiv((int64_t)a, (int64_t)b); vd = both && b != 0; return;
    case D_MOD_II: d = (uint64_t)imod((int64_t)a, (int64_t)b); vd = both; return;
    case D_ADD_FF: d = bits(__dadd_rn(af, bf)); vd = both; return;
    case D_SUB_FF: d = bits(__dsub_rn(af, bf)); vd = both; return;
    case D_MUL_FF: d = bits(__dmul_rn(af, bf)); vd = both; return;
    case D_DIV_FF:
      d = bits(__ddiv_rn(af, bf));
      vd = both && (FB ? !(as_f(b) == 0.0) : (b != 0));
      return;
    case D_EQ_II: d = (int64_t)a == (int64_t)b; vd = both; return;
    case D_NE_II: d = (int64_t)a != (int64_t)b; vd = both; return;
    case D_LT_II: d = (int64_t)a < (int64_t)b; vd = both; return;
    case D_GT_II: d = (int64_t)a > (int64_t)b; vd = both; return;
    case D_LE_II: d = (int64_t)a <= (int64_t)b; vd = both; return;
    case D_GE_II: d = (int64_t)a >= (int64_t)b; vd = both; return;
    case D_EQ_FF: d = af == bf; vd = both; return;
    case D_NE_FF: d = af != bf; vd = both; return;
    case D_LT_FF: d = af < bf; vd = both; return;
    case D_GT_FF: d = af > bf; vd = both; return;
    case D_LE_FF: d = af <= bf; vd = both; return;
    case D_GE_FF: d = af >= bf; vd = both; return;
    case D_EQ_TS: case D_NE_TS: case D_LT_TS: case D_GT_TS: case D_LE_TS: case D_GE_TS: {
      const int64_t x = (int64_t)(a * (uint64_t)ns_per_unit(op.aux2 & 15));
      const int64_t y = (int64_t)(b * (uint64_t)ns_per_unit(op.aux2 >> 4));
      d = (uint64_t)cmp_result(op.code - D_EQ_TS, x < y ? -1 : (x > y ? 1 : 0));
      vd = both;
      return;
    }
    case D_AND: d = truthy(a, FA) && truthy(b, FB); vd = both; return;
    case D_OR:  // interpreter_ops.cuh:575-587
      if (both) { d = truthy(a, FA) || truthy(b, FB); vd = 1; }
      else if (o.va) { d = a; vd = truthy(a, FA); }
      else if (o.vb) { d = b; vd = 1; }
      else { d = 0; vd = 0; }
      return;
    case D_INDF_II: case D_INDF_FF:  // interpreter_ops.cuh:588-598
      if (!o.va && !o.vb) d = 1;
      else if (!o.va || !o.vb) d = 0;
      else d = op.code == D_INDF_II ? (a == b) : (af == bf);
      vd = 1;
      return;
    case D_MAGIC_IF_NOT:  // interpreter_ops.cuh:559-566
      if (o.va && truthy(a, FA)) { d = b; vd = o.vb; }
      else { d = FB ? bits(MAGIC_F64) : (uint64_t)MAGIC_I64; vd = 0; }
      return;
    case D_FIRST_NON_MAGIC: {  // interpreter_ops.cuh:567-574
      const bool magic = FA ? (as_f(a) == MAGIC_F64) : ((int64_t)a == MAGIC_I64);
      if (magic) { d = b; vd = o.vb; } else { d = a; vd = o.va; }
      return;
    }
    // unary ops: computed only for a valid operand, validity passes through
    // (interpreter_ops.cuh:728-917)
    case D_NOT: d = !truthy(a, FA); vd = o.va; return;
    case D_ABS_I: d = (int64_t)a < 0 ? 0ULL - a : a; vd = o.va; return;
    case D_ABS_F: d = bits(fabs(as_f(a))); vd = o.va; return;
    case D_FLOOR: d = bits(floor(af)); vd = o.va; return;
    case D_CEIL: d = bits(ceil(af)); vd = o.va; return;
    case D_IS_NULL: d = !o.va; vd = 1; return;
    case D_IS_NOT_NULL: d = o.va; vd = 1; return;
    case D_IS_TRUE: d = FA ? (uint64_t)f64_to_i64(as_f(a)) : a; vd = o.va; return;
    case D_IS_NOT_TRUE: d = !(o.va && (FA ? as_f(a) == 1.0 : a == 1)); vd = 1; return;
    case D_IS_NOT_FALSE: d = !(o.va && (FA ? as_f(a) == 0.0 : a == 0)); vd = 1; return;
    case D_TO_I64: d = FA ? (uint64_t)f64_to_i64(as_f(a)) : a; vd = o.va; return;
    case D_TO_F64: d = bits(af); vd = o.va; return;
    case D_CAST_DATE: {
      int64_t v = FA ? f64_to_i64(as_f(a)) : (int64_t)a;
      if (is_ts(op.aux2) && op.aux2 != BSQL_TIMESTAMP_DAYS) v = v / ticks_per_day(op.aux2);
      d = (uint64_t)v;
      vd = o.va;
      return;
    }
    case D_CAST_TS: {
      uint64_t v = FA ? (uint64_t)f64_to_i64(as_f(a)) : a;
      if (is_ts(op.aux2)) v = v * (uint64_t)(86400000000000LL / ticks_per_day(op.aux2));
      d = v;
      vd = o.va;
      return;
    }
    default: d = a; vd = 0; return;
  }
}

__device__ __forceinline__ void str_at(const DevPlan& plan, uint8_t col, int64_t imm, int64_t row,
                                       const uint8_t** p, int32_t* n) {
  if (col != 0xff) {
    const DevInCol& c = plan.in[col];
    const int32_t s = c.offsets[row], e = c.offsets[row + 1];
    *p = (const uint8_t*)c.data + s;
    *n = e - s;
  } else {
    const int32_t* lit = (const int32_t*)imm;
    *n = lit[0];
    *p = (const uint8_t*)(lit + 1);
  }
}

// One VM step for this lane's R rows (rows row0 .. row0+nrows-1 exist).
template <int NREG, int R>
__device__ __forceinline__ void vm_step(const DevPlan& plan, const DevOp& op,
                                        const RegFile<NREG, R>& rv, uint32_t (&rvalid)[R],
                                        uint32_t& alive, int64_t row0, int nrows) {
  const uint8_t code = op.code;
  uint64_t D[R];
  uint32_t VD = 0;  // bit k = result validity of row k
  if (code == D_LOADN) {
    // issue every column's loads before the first register write, so that
    // all of them are in flight together
    uint64_t T[4][R];
    uint32_t TV[4] = {0, 0, 0, 0};
    const int n = op.aux;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < n) {
        const DevInCol& c = plan.in[(op.imm >> (16 * j)) & 0xff];
        load_rows<R>(c, row0, nrows, T[j]);
        TV[j] = load_valid<R>(c, row0, nrows);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < n) {
        const int reg = (int)((op.imm >> (16 * j + 8)) & 0xff);
        const uint32_t keep = ~(1u << reg);
#pragma unroll
        for (int k = 0; k < R; ++k) {
          rv.at(k, reg) = T[j][k];
          rvalid[k] = (rvalid[k] & keep) | (((TV[j] >> k) & 1u) << reg);
        }
      }
    }
    return;
  }
  if (code == D_LOAD) {
    const DevInCol& c = plan.in[op.a];
    load_rows<R>(c, row0, nrows, D);
    VD = load_valid<R>(c, row0, nrows);
  } else if (code == D_LOADV) {
    VD = load_valid<R>(plan.in[op.a], row0, nrows);
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = 0;
  } else {
    // operand fetch: the only reads of the register file
    uint64_t A[R], B[R];
    uint32_t VA = 0, VB = 0;
    const bool a_imm = op.flags & F_A_IMM, b_imm = op.flags & F_B_IMM;
    const bool has_b = code < D_NOT && code != D_KILL && code != D_SEL;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      A[k] = a_imm ? (uint64_t)op.imm : rv.at(k, op.a);
      VA |= (a_imm ? 1u : (op.flags & F_A_NULL) ? 0u : ((rvalid[k] >> op.a) & 1u)) << k;
    }
    if (has_b) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        B[k] = b_imm ? (uint64_t)op.imm : rv.at(k, op.b);
        VB |= (b_imm ? 1u : (op.flags & F_B_NULL) ? 0u : ((rvalid[k] >> op.b) & 1u)) << k;
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) B[k] = 0;
    }
    if (code == D_KILL || code == D_SEL) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const bool t = code == D_KILL ? truthy(A[k], FA) : (A[k] != 0);
        if (!(((VA >> k) & 1u) && t)) alive &= ~(1u << k);
      }
      return;
    }
    if (code >= D_EQ_SS && code <= D_GE_SS) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (k < nrows && ((VA >> k) & 1u) && ((VB >> k) & 1u)) {
          const uint8_t *pa, *pb;
          int32_t na, nb;
          str_at(plan, op.aux, op.imm, row0 + k, &pa, &na);
          str_at(plan, op.aux2, op.imm, row0 + k, &pb, &nb);
          D[k] = (uint64_t)cmp_result(code - D_EQ_SS, str_cmp(pa, na, pb, nb));
          VD |= 1u << k;
        }
      }
    } else if (code == D_CHAR_LENGTH) {
      const DevInCol& c = plan.in[op.aux];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (k < nrows && ((VA >> k) & 1u)) {
          const int32_t s = c.offsets[row0 + k], e = c.offsets[row0 + k + 1];
          const uint8_t* p = (const uint8_t*)c.data;
          int64_t n = 0;
          for (int32_t i = s; i < e; ++i) n += (p[i] & 0xC0) != 0x80;
          D[k] = (uint64_t)n;
          VD |= 1u << k;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        Operands o{A[k], B[k], (VA >> k) & 1u, (VB >> k) & 1u};
        uint32_t vd;
        compute(op, o, D[k], vd, row0 + k);
        VD |= vd << k;
      }
    }
  }
  // result write: the only write of the register file
  const uint32_t keep = ~(1u << op.dst);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    rv.at(k, op.dst) = D[k];
    rvalid[k] = (rvalid[k] & keep) | (((VD >> k) & 1u) << op.dst);
  }
}

// ------------------------------------------------------------- execute path

// perform_interpreter_operation: outputs at the input row positions.
template <int NREG, int R>
__global__ __launch_bounds__(BLOCK) void vm_execute(DevPlan plan) {
  BSQL_REGFILE(rv);
  uint32_t rvalid[R];
#pragma unroll
  for (int k = 0; k < R; ++k) rvalid[k] = 0;
  const int64_t tile_rows = (int64_t)BLOCK * R;
  const int64_t row0 = (int64_t)blockIdx.x * tile_rows + (int64_t)threadIdx.x * R;
  const int64_t left = plan.num_rows - row0;
  const int nrows = left <= 0 ? 0 : (left >= R ? R : (int)left);
  uint32_t alive = (1u << nrows) - 1u;
  const int lane = threadIdx.x & 63;

  for (int q = 0; q < plan.n_ops; ++q) {
    const DevOp op = plan.ops[q];
    if (op.code == D_STORE) {
      const DevOutCol& c = plan.out[op.aux];
      uint64_t vals[R];
      uint32_t vm = 0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        vals[k] = rv.at(k, op.a);
        vm |= ((rvalid[k] >> op.a) & 1u) << k;
      }
      if (nrows > 0) store_rows<R>(c, row0, nrows, vals);
      if (c.valid) {
        // lanes covering one 32-row word: 32/R of them; OR their bits together
        constexpr int LPW = 32 / R;
        uint32_t word = (vm & alive) << ((lane % LPW) * R);
#pragma unroll
        for (int s = 1; s < LPW; s <<= 1) word |= __shfl_xor(word, s, 64);
        if ((lane % LPW) == 0 && nrows > 0) c.valid[row0 >> 5] = word;
      }
      continue;
    }
    if (nrows == 0) continue;
    vm_step<NREG, R>(plan, op, rv, rvalid, alive, row0, nrows);
  }
}

// --------------------------------------------------------- compaction path

// Tile status words: flag in bits 62-63, value (row count) in bits 0-61.
constexpr uint64_t ST_AGG = 1ULL << 62;
constexpr uint64_t ST_PREFIX = 2ULL << 62;
constexpr uint64_t ST_VALUE = (1ULL << 62) - 1;

struct CompactArgs {
  uint64_t* status;      // one word per tile, zeroed before the launch
  uint32_t* ticket;      // zeroed before the launch
  int64_t* d_count;      // total selected rows (written by the last tile)
  int64_t capacity;      // rows the outputs can hold; rows beyond are not written
  int64_t num_tiles;
};

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

// Exclusive prefix of this tile across all earlier tiles (wave 0 only).
__device__ int64_t look_back(const CompactArgs& ca, int64_t tile, int64_t agg, int lane) {
  if (tile == 0) {
    if (lane == 0) st_status(&ca.status[0], ST_PREFIX | (uint64_t)agg);
    return 0;
  }
  if (lane == 0) st_status(&ca.status[tile], ST_AGG | (uint64_t)agg);
  int64_t excl = 0;
  int64_t look = tile - 1;
  while (true) {
    const int64_t t = look - lane;
    const uint64_t s = t >= 0 ? ld_status(&ca.status[t]) : ST_PREFIX;
    const uint64_t flag = s & ~ST_VALUE;
    const uint64_t not_ready = __ballot(flag == 0);
    const uint64_t is_p = __ballot(flag == ST_PREFIX);
    const uint64_t upto = is_p ? ((is_p & (~is_p + 1)) << 1) - 1 : ~0ULL;  // lanes [0, first prefix]
    if (not_ready & upto) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const int64_t mine = ((upto >> lane) & 1) ? (int64_t)(s & ST_VALUE) : 0;
    excl += wave_sum(mine);
    if (is_p) break;
    look -= 64;
  }
  if (lane == 0) st_status(&ca.status[tile], ST_PREFIX | (uint64_t)(excl + agg));
  return excl;
}

// filter / filter+project: selection (phase 1), compaction, outputs (phase 2)
template <int NREG, int R>
__global__ __launch_bounds__(BLOCK) void vm_compact(DevPlan plan, CompactArgs ca) {
  __shared__ uint32_t s_ticket;
  __shared__ int64_t s_wave_total[WAVES];
  __shared__ int64_t s_tile_prefix;
  constexpr int WORDS = 2 * R + 2;  // validity words one wave's 64*R outputs can touch
  __shared__ uint32_t s_words[WAVES][WORDS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  if (tid == 0) s_ticket = atomicAdd(ca.ticket, 1u);
  __syncthreads();
  const int64_t tile = s_ticket;

  BSQL_REGFILE(rv);
  uint32_t rvalid[R];
#pragma unroll
  for (int k = 0; k < R; ++k) rvalid[k] = 0;
  const int64_t tile_rows = (int64_t)BLOCK * R;
  const int64_t row0 = tile * tile_rows + (int64_t)tid * R;
  const int64_t left = plan.num_rows - row0;
  const int nrows = left <= 0 ? 0 : (left >= R ? R : (int)left);
  uint32_t alive = (1u << nrows) - 1u;

  // ---- phase 1: selection with wave-uniform early exit
  for (int q = 0; q < plan.split; ++q) {
    const DevOp op = plan.ops[q];
    vm_step<NREG, R>(plan, op, rv, rvalid, alive, row0, nrows);
    if (op.code == D_KILL && __ballot(alive != 0) == 0) break;
  }

  // ---- wave / workgroup prefix of the selected rows
  int lane_excl = 0, wave_total = 0;
  uint64_t sel_ballot[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t b = __ballot((alive >> k) & 1u);
    sel_ballot[k] = b;
    lane_excl += __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
    wave_total += __popcll(b);
  }
  if (lane == 0) s_wave_total[wave] = wave_total;
  __syncthreads();
  int64_t wave_off = 0, tile_total = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    const int64_t t = s_wave_total[w];
    if (w < wave) wave_off += t;
    tile_total += t;
  }
  if (wave == 0) {
    const int64_t excl = look_back(ca, tile, tile_total, lane);
    if (lane == 0) {
      s_tile_prefix = excl;
      if (tile == ca.num_tiles - 1) *ca.d_count = excl + tile_total;
    }
  }
  __syncthreads();
  if (wave_total == 0) return;  // nothing selected in this wave: no phase 2

  const int64_t wave_pos = s_tile_prefix + wave_off;
  int64_t pos[R];
  {
    int64_t p = wave_pos + lane_excl;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      pos[k] = p;
      p += (alive >> k) & 1u;
    }
  }
  (void)sel_ballot;

  // ---- phase 2: projection ops and compacted stores
  for (int q = plan.split; q < plan.n_ops; ++q) {
    const DevOp op = plan.ops[q];
    if (op.code != D_STORE) {
      vm_step<NREG, R>(plan, op, rv, rvalid, alive, row0, nrows);
      continue;
    }
    const DevOutCol& c = plan.out[op.aux];
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (((alive >> k) & 1u) && pos[k] < ca.capacity) store_elem(c, pos[k], rv.at(k, op.a));
    if (!c.valid) continue;
    // pack this wave's output validity words: [wave_pos, wave_pos + wave_total)
    const int64_t w0 = wave_pos >> 5;
    const int nwords = (int)(((wave_pos + wave_total + 31) >> 5) - w0);  // <= WORDS
    if (lane < WORDS) s_words[wave][lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (((alive >> k) & 1u) && ((rvalid[k] >> op.a) & 1u)) {
        const int64_t rel = pos[k] - (w0 << 5);
        atomicOr(&s_words[wave][rel >> 5], 1u << (rel & 31));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < nwords) {
      const int64_t wi = w0 + lane;
      const int64_t lo = wave_pos > (wi << 5) ? wave_pos : (wi << 5);
      int64_t hi = wave_pos + wave_total < ((wi + 1) << 5) ? wave_pos + wave_total : ((wi + 1) << 5);
      if (hi > ca.capacity) hi = ca.capacity;
      if (hi > lo) {
        const int a = (int)(lo - (wi << 5)), b = (int)(hi - (wi << 5));
        const uint32_t mask = (b >= 32 ? 0xffffffffu : ((1u << b) - 1u)) & ~((1u << a) - 1u);
        const uint32_t word = s_words[wave][lane] & mask;
        if (mask == 0xffffffffu) {
          c.valid[wi] = word;
        } else {
          // boundary word shared with neighbouring waves: disjoint bit ranges
          atomicAnd(&c.valid[wi], ~mask | word);
          atomicOr(&c.valid[wi], word);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace bsql_dev

// ------------------------------------------------------------------ launchers

namespace bsql {

using namespace bsql_dev;

template <int NREG, int R>
static hipError_t launch_execute_t(const DevPlan& plan, hipStream_t stream) {
  const int64_t tile = (int64_t)BLOCK * R;
  const int64_t grid = (plan.num_rows + tile - 1) / tile;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((vm_execute<NREG, R>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, plan);
  return hipGetLastError();
}

template <int NREG, int R>
static hipError_t launch_compact_t(const DevPlan& plan, const CompactArgs& ca, hipStream_t stream) {
  hipLaunchKernelGGL((vm_compact<NREG, R>), dim3((unsigned)ca.num_tiles), dim3(BLOCK), 0, stream,
                     plan, ca);
  return hipGetLastError();
}

// (NREG, R) variants compiled: register file size x rows per lane.
#define BSQL_VARIANTS(X) X(8, 2) X(8, 4) X(16, 1) X(16, 2) X(32, 1)

int tile_rows_for(int nreg, int rows_per_thread) {
  (void)nreg;
  return BLOCK * rows_per_thread;
}

// Picks the variant: the smallest register file that holds the program, and
// the rows per lane requested (0 = largest that keeps the register file small).
static bool pick(int nreg, int rpt_req, int* NR, int* RR) {
  struct V { int n, r; };
  // default rows per lane for each register-file size first
  const V vs[] = {{8, 4}, {16, 2}, {32, 1}, {8, 2}, {16, 1}};
  for (const V& v : vs) {
    if (v.n < nreg) continue;
    if (rpt_req > 0 && v.r != rpt_req) continue;
    *NR = v.n;
    *RR = v.r;
    return true;
  }
  if (rpt_req > 0) return pick(nreg, 0, NR, RR);
  return false;
}

hipError_t launch_execute(const DevPlan& plan, int nreg, int rpt, hipStream_t stream) {
  int NR, RR;
  if (!pick(nreg, rpt, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT) if (NR == N && RR == RPT) return launch_execute_t<N, RPT>(plan, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

int64_t compact_tiles(int64_t num_rows, int nreg, int rpt) {
  int NR, RR;
  if (!pick(nreg, rpt, &NR, &RR)) return -1;
  const int64_t tile = (int64_t)BLOCK * RR;
  return (num_rows + tile - 1) / tile;
}

// `status` must hold compact_tiles(...) + 1 words (the last one is the ticket).
hipError_t launch_compact(const DevPlan& plan, int nreg, int rpt, uint64_t* status,
                          int64_t* d_count, int64_t capacity, hipStream_t stream) {
  int NR, RR;
  if (!pick(nreg, rpt, &NR, &RR)) return hipErrorInvalidValue;
  const int64_t tile = (int64_t)BLOCK * RR;
  const int64_t tiles = (plan.num_rows + tile - 1) / tile;
  CompactArgs ca{status, (uint32_t*)(status + tiles), d_count, capacity, tiles};
  // clear tile status, ticket and the count (fill kernels, so the sequence can
  // be captured in a hipGraph)
  const int64_t nclear = tiles + 1;
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, stream,
                     status, nclear, 0ULL);
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, stream, (uint64_t*)d_count, (int64_t)1, 0ULL);
  if (tiles == 0) return hipGetLastError();
#define X(N, RPT) if (NR == N && RR == RPT) return launch_compact_t<N, RPT>(plan, ca, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace bsql
