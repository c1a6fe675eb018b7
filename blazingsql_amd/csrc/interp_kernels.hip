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
//   * compaction (filters): a selection kernel writes 64-bit ballots and a
//     count per wave slice and a total per workgroup block, with no order
//     between workgroups; slice_ids turns them into the selected row ids in
//     row order; the output pass gathers those rows, 64 per instruction.
//     Optionally (BSQL_LAUNCH_FUSED) a range-test filter runs as one kernel:
//     wave tiles taken from ticket counters, decoupled look-back for the
//     output position, output validity assembled in LDS and written as whole
//     words, with an AND/OR atomic pair for the two words a tile shares.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "devplan.h"
#include "pow10.h"
#include "../../include/bsql/interp.h"

namespace bsql_dev {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));

// Column pointers arrive through the plan in memory, so the compiler cannot
// tell that they address global memory and would emit flat loads/stores;
// every column access goes through a global (address space 1) pointer.
#define GLOBAL __attribute__((address_space(1)))
#define LDS __attribute__((address_space(3)))
// The launch plan and its op list are never written while a kernel runs:
// read through the constant address space, their wave-uniform loads are
// scalar loads (a generic pointer would make them vector loads, and every
// op field derived from them divergent).
#define CONSTANT __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const GLOBAL T* gp(const void* p) { return (const GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ GLOBAL T* gpw(void* p) { return (GLOBAL T*)p; }

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
__device__ int str_cmp(const GLOBAL uint8_t* a, int32_t na, const GLOBAL uint8_t* b, int32_t nb) {
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
    const GLOBAL CT* p = gp<CT>(c.data) + r0;                                      \
    CT tmp[R];                                                                     \
    if (vec && (sizeof(CT) * R) % 16 == 0) {                                       \
      _Pragma("unroll") for (int k = 0; k < (int)(sizeof(CT) * R / 16); ++k)       \
        ((v4u32*)tmp)[k] = __builtin_nontemporal_load((const GLOBAL v4u32*)p + k); \
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
      const GLOBAL int64_t* p = gp<int64_t>(c.data) + r0;
      int64_t tmp[R];
      if (vec && (8 * R) % 16 == 0) {
#pragma unroll
        for (int k = 0; k < 8 * R / 16; ++k)
          ((v4u32*)tmp)[k] = __builtin_nontemporal_load((const GLOBAL v4u32*)p + k);
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
  const GLOBAL uint32_t* vp = gp<uint32_t>(c.valid);
  uint64_t word = vp[w];
  if (sh + R > 32 && n > 32 - sh) word |= (uint64_t)vp[w + 1] << 32;
  uint32_t m = (uint32_t)(word >> sh) & ((1u << R) - 1u);
  return n >= R ? m : (m & ((1u << n) - 1u));
}

// static_cast of slot bits to the output column type (interpreter_ops.cuh:329-395);
// NT: 8-byte elements with a non-temporal store
template <bool NT = false>
__device__ __forceinline__ void store_elem(const DevOutCol& c, int64_t row, uint64_t b) {
  switch (c.type) {
    case BSQL_INT8: case BSQL_UINT8: gpw<uint8_t>(c.data)[row] = (uint8_t)b; break;
    case BSQL_BOOL8: gpw<uint8_t>(c.data)[row] = b != 0; break;
    case BSQL_INT16: case BSQL_UINT16: gpw<uint16_t>(c.data)[row] = (uint16_t)b; break;
    case BSQL_INT32: case BSQL_UINT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
      gpw<uint32_t>(c.data)[row] = (uint32_t)b; break;
    case BSQL_FLOAT32: gpw<float>(c.data)[row] = (float)as_f(b); break;
    default:  // int64, uint64, float64, 8-byte ticks
      if (NT) __builtin_nontemporal_store(b, gpw<uint64_t>(c.data) + row);
      else gpw<uint64_t>(c.data)[row] = b;
      break;
  }
}

// R consecutive rows, vectorised when the tile is full and aligned
template <int R>
__device__ __forceinline__ void store_rows(const DevOutCol& c, int64_t r0, int n, const uint64_t (&v)[R]) {
  if (n == R && (c.type == BSQL_FLOAT64 || c.type == BSQL_INT64) && (R % 2 == 0) &&
      ((((uintptr_t)((uint64_t*)c.data + r0)) & 15) == 0)) {
    GLOBAL uint64_t* p = gpw<uint64_t>(c.data) + r0;
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      v4u32 q;
      q.x = (uint32_t)v[2 * k]; q.y = (uint32_t)(v[2 * k] >> 32);
      q.z = (uint32_t)v[2 * k + 1]; q.w = (uint32_t)(v[2 * k + 1] >> 32);
      __builtin_nontemporal_store(q, (GLOBAL v4u32*)p + k);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (k < n) store_elem(c, r0 + k, v[k]);
}

// ------------------------------------------------------------ the VM core
//

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

// String values in registers are slices of a column's chars: start byte in
// the low 32 bits, byte length in the high 32 bits. A string literal operand
// holds the device address of its {int32 len, bytes} record instead.
__device__ __forceinline__ uint64_t slice_pack(uint32_t start, uint32_t len) {
  return (uint64_t)start | ((uint64_t)len << 32);
}
__device__ __forceinline__ void str_at(const DevPlan& plan, uint8_t col, uint64_t v, const GLOBAL uint8_t** p,
                                       int32_t* n) {
  if (col != 0xff) {
    *p = gp<uint8_t>(plan.in[col].data) + (uint32_t)v;
    *n = (int32_t)(v >> 32);
  } else {
    const GLOBAL int32_t* lit = gp<int32_t>((const void*)v);
    *n = lit[0];
    *p = (const GLOBAL uint8_t*)(lit + 1);
  }
}

__device__ __forceinline__ bool utf8_lead(uint8_t c) { return (c & 0xC0) != 0x80; }
// bytes of the UTF-8 character starting at s[i] (i < n)
__device__ __forceinline__ int32_t utf8_step(const GLOBAL uint8_t* s, int32_t i, int32_t n) {
  int32_t j = i + 1;
  while (j < n && !utf8_lead(s[j])) ++j;
  return j - i;
}
// byte offset of character `chars` of s (n bytes if s has fewer characters)
__device__ int32_t utf8_offset(const GLOBAL uint8_t* s, int32_t n, int64_t chars) {
  int64_t c = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (utf8_lead(s[i])) {
      if (c == chars) return i;
      ++c;
    }
  }
  return n;
}
__device__ int64_t utf8_chars(const GLOBAL uint8_t* s, int32_t n) {
  int64_t c = 0;
  for (int32_t i = 0; i < n; ++i) c += utf8_lead(s[i]);
  return c;
}

// The UTF-8 character (its encoded bytes packed little-endian) at s[i].
__device__ __forceinline__ uint32_t utf8_char_at(const GLOBAL uint8_t* s, int32_t i, int32_t len) {
  uint32_t c = 0;
  for (int32_t k = 0; k < len; ++k) c |= (uint32_t)s[i + k] << (8 * k);
  return c;
}

// Whether character c is one to strip: one of the characters of `set`, or
// for an empty set a byte <= ' ' (cudf::strings::strip's rule).
__device__ bool strip_char(uint32_t c, const GLOBAL uint8_t* set, int32_t nset) {
  if (nset == 0) return c <= (uint32_t)' ';
  for (int32_t i = 0; i < nset;) {
    const int32_t l = utf8_step(set, i, nset);  // >= 1
    if (utf8_char_at(set, i, l) == c) return true;
    i += l;
  }
  return false;
}

// [b0, b1) of s[0, n) after stripping (side bit 0: leading, bit 1: trailing).
__device__ void strip_range(const GLOBAL uint8_t* s, int32_t n, const GLOBAL uint8_t* set, int32_t nset, int side,
                            int32_t* b0, int32_t* b1) {
  int32_t lo = 0, hi = n;
  if (side & 1) {
    while (lo < hi) {
      const int32_t l = utf8_step(s, lo, hi);  // >= 1
      if (!strip_char(utf8_char_at(s, lo, l), set, nset)) break;
      lo += l;
    }
  }
  if (side & 2) {
    while (hi > lo) {
      int32_t st = hi - 1;
      while (st > lo && !utf8_lead(s[st])) --st;
      if (!strip_char(utf8_char_at(s, st, hi - st), set, nset)) break;
      hi = st;
    }
  }
  *b0 = lo;
  *b1 = hi;
}

// cudf::strings::to_integers (libcudf 21.08, string_to_integer): an optional
// sign, then decimal digits up to the first other character; wraps.
__device__ int64_t str_to_i64(const GLOBAL uint8_t* s, int32_t n) {
  int32_t i = 0;
  uint64_t sign = 1;
  if (n > 0 && (s[0] == '-' || s[0] == '+')) {
    sign = s[0] == '-' ? ~0ULL : 1ULL;
    i = 1;
  }
  uint64_t v = 0;
  for (; i < n; ++i) {
    const uint8_t c = s[i];
    if (c < '0' || c > '9') break;
    v = v * 10ULL + (uint64_t)(c - '0');
  }
  return (int64_t)(v * sign);
}

__constant__ double POW10[BSQL_POW10_MAX - BSQL_POW10_MIN + 1] = {BSQL_POW10_VALUES};

// cudf::strings::to_floats (libcudf 21.08, the parser LogicalProject.cpp:439-468
// hands CAST(varchar AS DOUBLE/FLOAT) to): exactly "NaN", "Inf" or "-Inf";
// else an optional sign, digits and '.'s (every '.' marks the fraction) kept
// while they fit 52 bits, an optional exponent; a decimal exponent above 308
// is infinity and below -307 is +0; else sign * digits * 10^e, with 10^e the
// double nearest it.
__device__ double str_to_f64(const GLOBAL uint8_t* s, int32_t n) {
  if (n == 0) return 0.0;
  if (n == 3 && s[0] == 'N' && s[1] == 'a' && s[2] == 'N') return __longlong_as_double(0x7ff8000000000000LL);
  if (n == 3 && s[0] == 'I' && s[1] == 'n' && s[2] == 'f') return __longlong_as_double(0x7ff0000000000000LL);
  if (n == 4 && s[0] == '-' && s[1] == 'I' && s[2] == 'n' && s[3] == 'f')
    return __longlong_as_double((long long)0xfff0000000000000ULL);
  int32_t i = 0;
  double sign = 1.0;
  if (s[0] == '-' || s[0] == '+') {
    sign = s[0] == '-' ? -1.0 : 1.0;
    i = 1;
  }
  const uint64_t max_mantissa = 0x0FFFFFFFFFFFFFULL;
  uint64_t digits = 0;
  uint32_t exp_off = 0;  // int32 arithmetic, wrapping
  bool decimal = false;
  for (; i < n; ++i) {
    const uint8_t c = s[i];
    if (c == '.') { decimal = true; continue; }
    if (c < '0' || c > '9') break;
    if (digits > max_mantissa) {
      exp_off += !decimal;
    } else {
      digits = digits * 10ULL + (uint64_t)(c - '0');
      if (digits > max_mantissa) {
        digits /= 10ULL;
        exp_off += !decimal;
      } else {
        exp_off -= decimal;
      }
    }
  }
  uint32_t exp_ten = 0;
  bool exp_neg = false;
  if (i < n) {
    const uint8_t c = s[i++];
    if ((c == 'e' || c == 'E') && i < n) {
      if (s[i] == '-' || s[i] == '+') exp_neg = s[i++] == '-';
      while (i < n) {
        const uint8_t d = s[i++];
        if (d < '0' || d > '9') break;
        exp_ten = exp_ten * 10u + (uint32_t)(d - '0');
      }
    }
  }
  const int32_t e = (int32_t)((exp_neg ? 0u - exp_ten : exp_ten) + exp_off);
  if (e > BSQL_POW10_MAX) return sign > 0 ? __longlong_as_double(0x7ff0000000000000LL)
                                          : __longlong_as_double((long long)0xfff0000000000000ULL);
  if (e < BSQL_POW10_MIN) return 0.0;
  double base = __dmul_rn(sign, (double)digits);
  if (e) base = __dmul_rn(base, POW10[e - BSQL_POW10_MIN]);
  return base;
}

// to_timestamps' default formats as compiled items ((spec << 8) | width,
// spec 0 for literal characters): "%Y-%m-%d" and "%Y-%m-%d %H:%M:%S".
__constant__ int32_t TS_FMT_DATE[6] = {5, ('Y' << 8) | 4, 1, ('m' << 8) | 2, 1, ('d' << 8) | 2};
__constant__ int32_t TS_FMT_DATETIME[12] = {11, ('Y' << 8) | 4, 1, ('m' << 8) | 2, 1, ('d' << 8) | 2, 1,
                                            ('H' << 8) | 2, 1, ('M' << 8) | 2, 1, ('S' << 8) | 2};

// Days since 1970-01-01 of a proleptic Gregorian date (Hinnant's days_from_civil).
__device__ int64_t days_from_ymd(int64_t y, int64_t m, int64_t d) {
  y -= m <= 2 ? 1 : 0;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t doy = (153 * (m > 2 ? m - 3 : m + 9) + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

// cudf::strings::to_timestamps (libcudf 21.08): each field of the format is
// read from a fixed number of characters (decimal digits up to the first
// other character), literal characters are stepped over, fields the string
// does not reach keep 1970-01-01 00:00:00; the result is the date in days or
// the date and time in the ticks of `type`.
__device__ int64_t str_to_ticks(const GLOBAL uint8_t* s, int32_t n, const int32_t* fmt, int type) {
  int64_t y = 1970, mo = 1, d = 1, h = 0, mi = 0, se = 0;
  const int32_t n_items = fmt[0];
  int32_t pos = 0;
  for (int32_t k = 1; k <= n_items && pos < n; ++k) {
    const int32_t it = fmt[k];
    const int32_t spec = it >> 8, width = it & 0xFF;
    if (spec) {
      const int32_t take = width < n - pos ? width : n - pos;
      int64_t v = 0;
      for (int32_t i = 0; i < take; ++i) {
        const uint8_t c = s[pos + i];
        if (c < '0' || c > '9') break;
        v = v * 10 + (c - '0');
      }
      switch (spec) {
        case 'Y': y = v; break;
        case 'm': mo = v; break;
        case 'd': d = v; break;
        case 'H': h = v; break;
        case 'M': mi = v; break;
        default: se = v; break;
      }
    }
    pos += width;
  }
  const int64_t days = days_from_ymd(y, mo, d);
  if (type == BSQL_TIMESTAMP_DAYS) return days;
  const uint64_t secs = (uint64_t)days * 86400ULL + (uint64_t)(h * 3600 + mi * 60 + se);
  switch (type) {
    case BSQL_TIMESTAMP_MILLISECONDS: return (int64_t)(secs * 1000ULL);
    case BSQL_TIMESTAMP_MICROSECONDS: return (int64_t)(secs * 1000000ULL);
    case BSQL_TIMESTAMP_NANOSECONDS: return (int64_t)(secs * 1000000000ULL);
    default: return (int64_t)secs;
  }
}

// SQL LIKE through the reference's translation to a regex searched with
// contains_re (LogicalProject.cpp:31-47, :113-131): '%' = (?:.*?) and '_' =
// (?:.), neither matching '\n'; anchored at the start unless the pattern
// begins with '%', at the end unless it ends with '%'. The compiled pattern is
// {n_units, flags, units...}: unit >= 0 a literal byte, -1 any character,
// -2 any run of characters.
enum : int32_t { LIKE_ANCHOR_START = 1, LIKE_ANCHOR_END = 2, LIKE_MATCH_ALL = 4, LIKE_REGEX = 8 };

// Whether the regex compiled by regex_compile (its layout is described
// there) is found in s[0, n): a Thompson simulation over 128-bit state sets,
// restarted at every character (a search), '^' only at 0, '$' only at n.
// Linear in n times the states alive: no backtracking.
__device__ bool regex_char_in_class(const GLOBAL int32_t* prog, int32_t at, uint32_t cp) {
  const int32_t nr = prog[at];
  bool in = false;
  for (int32_t k = 0; k < nr && !in; ++k)
    in = cp >= (uint32_t)prog[at + 2 + 2 * k] && cp <= (uint32_t)prog[at + 3 + 2 * k];
  return prog[at + 1] ? !in : in;
}

__device__ bool regex_contains(const GLOBAL uint8_t* s, int32_t n, const GLOBAL int32_t* prog) {
  const int32_t P = prog[0];
  const GLOBAL int32_t* kind = prog + 4;
  const GLOBAL int32_t* mid = prog + 4 + P;
  const GLOBAL int32_t* end = prog + 4 + 5 * P;
  const GLOBAL int32_t* start = prog + 4 + 9 * P;
  const int32_t match = P - 1;
  uint32_t cur[4];
  {
    const GLOBAL int32_t* m = start + 4 * (2 + (n == 0 ? 1 : 0));
    for (int w = 0; w < 4; ++w) cur[w] = (uint32_t)m[w];
  }
  if ((cur[match >> 5] >> (match & 31)) & 1u) return true;
  int32_t i = 0;
  while (i < n) {
    int32_t l = utf8_step(s, i, n);
    uint32_t cp;
    const uint8_t c0 = s[i];
    if (l == 1) cp = c0;
    else if (l == 2) cp = ((uint32_t)(c0 & 0x1F) << 6) | (s[i + 1] & 0x3F);
    else if (l == 3) cp = ((uint32_t)(c0 & 0x0F) << 12) | ((uint32_t)(s[i + 1] & 0x3F) << 6) | (s[i + 2] & 0x3F);
    else cp = ((uint32_t)(c0 & 0x07) << 18) | ((uint32_t)(s[i + 1] & 0x3F) << 12) |
              ((uint32_t)(s[i + 2] & 0x3F) << 6) | (s[i + 3] & 0x3F);
    i += l;
    const bool at_end = i >= n;
    const GLOBAL int32_t* after = at_end ? end : mid;
    uint32_t nxt[4];
    {
      const GLOBAL int32_t* m = start + 4 * (at_end ? 1 : 0);
      for (int w = 0; w < 4; ++w) nxt[w] = (uint32_t)m[w];
    }
    for (int w = 0; w < 4; ++w) {
      uint32_t bits = cur[w];
      while (bits) {
        const int q = w * 32 + __builtin_ctz(bits);
        bits &= bits - 1u;
        const uint32_t k = (uint32_t)kind[q];
        const uint32_t arg = k >> 3;
        bool ok;
        switch (k & 7u) {
          case 0: ok = cp == arg; break;
          case 1: ok = cp != '\n'; break;
          case 2: ok = regex_char_in_class(prog, (int32_t)arg, cp); break;
          default: ok = false; break;
        }
        if (ok) {
          const GLOBAL int32_t* m = after + 4 * (q + 1);
          for (int v = 0; v < 4; ++v) nxt[v] |= (uint32_t)m[v];
        }
      }
    }
    for (int w = 0; w < 4; ++w) cur[w] = nxt[w];
    if ((cur[match >> 5] >> (match & 31)) & 1u) return true;
  }
  return false;
}
__device__ bool like_from(const GLOBAL uint8_t* s, int32_t n, int32_t st, const GLOBAL int32_t* u, int32_t nu,
                          bool anchor_end) {
  int32_t i = st, q = 0, star_q = -1, star_i = 0;
  while (true) {
    if (q == nu) {
      if (!anchor_end || i == n) return true;
    } else if (u[q] == -2) {
      star_q = q++;
      star_i = i;
      continue;
    } else if (i < n && (u[q] == -1 ? s[i] != '\n' : (int32_t)s[i] == u[q])) {
      i += u[q] == -1 ? utf8_step(s, i, n) : 1;
      ++q;
      continue;
    }
    // let the last '%' take one more character
    if (star_q < 0 || star_i >= n || s[star_i] == '\n') return false;
    star_i += utf8_step(s, star_i, n);
    i = star_i;
    q = star_q + 1;
  }
}
__device__ bool like_match(const GLOBAL uint8_t* s, int32_t n, const GLOBAL int32_t* pat) {
  const int32_t nu = pat[0], flags = pat[1];
  if (flags & LIKE_MATCH_ALL) return true;
  const GLOBAL int32_t* u = pat + 2;
  if (flags & LIKE_REGEX) return regex_contains(s, n, u);
  const bool anchor_end = (flags & LIKE_ANCHOR_END) != 0;
  if (flags & LIKE_ANCHOR_START) return like_from(s, n, 0, u, nu, anchor_end);
  if (nu > 0 && u[0] == -2) {
    // a leading run wildcard covers every start on the first line; a match
    // may also start after any newline
    if (like_from(s, n, 0, u, nu, anchor_end)) return true;
    for (int32_t i = 0; i < n; ++i)
      if (s[i] == '\n' && like_from(s, n, i + 1, u, nu, anchor_end)) return true;
    return false;
  }
  // unanchored with a literal first unit ("%%..." leaves the first '%'
  // literal): try every character start
  for (int32_t i = 0; i <= n; i += (i < n ? utf8_step(s, i, n) : 1))
    if (like_from(s, n, i, u, nu, anchor_end)) return true;
  return false;
}

#define FA ((op.flags & F_A_F) != 0)
#define FB ((op.flags & F_B_F) != 0)
// float ops: the lowering hands them both operands as doubles
#define AF as_f(a)
#define BF as_f(b)
// one result per row: D[k] and validity bit k of VD
#define ROWS(EXPR_D, EXPR_V)                                                   \
  _Pragma("unroll") for (int k = 0; k < R; ++k) {                              \
    const uint64_t a = A[k], b = B[k];                                         \
    const uint32_t va = (VA >> k) & 1u, vb = (VB >> k) & 1u;                   \
    (void)a; (void)b; (void)va; (void)vb;                                      \
    D[k] = (uint64_t)(EXPR_D);                                                 \
    VD |= ((uint32_t)(EXPR_V) & 1u) << k;                                      \
  }                                                                            \
  return;

// Result of one op for this lane's R rows. Reference semantics per op in
// engine/src/Interpreter/interpreter_ops.cuh:500-925.
template <int R> struct Rows;

template <int R>
__device__ __forceinline__ void compute_rows(const DevPlan& plan, const DevOp& op, const uint64_t (&A)[R],
                                             const uint64_t (&B)[R], uint32_t VA, uint32_t VB,
                                             uint64_t (&D)[R], uint32_t& VD, const Rows<R>& rs) {
  VD = 0;
  switch (op.code) {
    // binary ops: valid only if both operands are (:599-727); x/0 is null (:721-724)
    case D_ADD_II: ROWS(a + b, va & vb)
    case D_SUB_II: ROWS(a - b, va & vb)
    case D_MUL_II: ROWS(a * b, va & vb)
    case D_DIV_II: ROWS(idiv((int64_t)a, (int64_t)b), va & vb & (b != 0))
    case D_MOD_II: ROWS(imod((int64_t)a, (int64_t)b), va & vb)
    case D_ADD_FF: ROWS(bits(__dadd_rn(AF, BF)), va & vb)
    case D_SUB_FF: ROWS(bits(__dsub_rn(AF, BF)), va & vb)
    case D_MUL_FF: ROWS(bits(__dmul_rn(AF, BF)), va & vb)
    case D_DIV_FF: ROWS(bits(__ddiv_rn(AF, BF)), va & vb & !(BF == 0.0))
    case D_EQ_II: ROWS((int64_t)a == (int64_t)b, va & vb)
    case D_NE_II: ROWS((int64_t)a != (int64_t)b, va & vb)
    case D_LT_II: ROWS((int64_t)a < (int64_t)b, va & vb)
    case D_GT_II: ROWS((int64_t)a > (int64_t)b, va & vb)
    case D_LE_II: ROWS((int64_t)a <= (int64_t)b, va & vb)
    case D_GE_II: ROWS((int64_t)a >= (int64_t)b, va & vb)
    case D_EQ_FF: ROWS(AF == BF, va & vb)
    case D_NE_FF: ROWS(AF != BF, va & vb)
    case D_LT_FF: ROWS(AF < BF, va & vb)
    case D_GT_FF: ROWS(AF > BF, va & vb)
    case D_LE_FF: ROWS(AF <= BF, va & vb)
    case D_GE_FF: ROWS(AF >= BF, va & vb)
    case D_EQ_TS: case D_NE_TS: case D_LT_TS: case D_GT_TS: case D_LE_TS: case D_GE_TS: {
      const uint64_t fa = (uint64_t)ns_per_unit(op.aux2 & 15), fb = (uint64_t)ns_per_unit(op.aux2 >> 4);
      const int kind = op.code - D_EQ_TS;
      ROWS(cmp_result(kind, (int64_t)(a * fa) < (int64_t)(b * fb) ? -1
                            : ((int64_t)(a * fa) > (int64_t)(b * fb) ? 1 : 0)), va & vb)
    }
    case D_AND: ROWS(truthy(a, FA) && truthy(b, FB), va & vb)
    case D_OR:  // :575-587
      ROWS((va & vb) ? (uint64_t)(truthy(a, FA) || truthy(b, FB)) : (va ? a : b),
           (va & vb) | (va & (uint32_t)truthy(a, FA)) | ((va ^ 1u) & vb))
    case D_INDF_II:  // :588-598
      ROWS((!va && !vb) ? 1 : ((!va || !vb) ? 0 : (a == b)), 1u)
    case D_INDF_FF:
      ROWS((!va && !vb) ? 1 : ((!va || !vb) ? 0 : (AF == BF)), 1u)
    case D_MAGIC_IF_NOT:  // :559-566
      ROWS((va && truthy(a, FA)) ? b : (FB ? bits(MAGIC_F64) : (uint64_t)MAGIC_I64),
           (va && truthy(a, FA)) ? vb : 0u)
    case D_FIRST_NON_MAGIC: {  // :567-574
      ROWS((FA ? (as_f(a) == MAGIC_F64) : ((int64_t)a == MAGIC_I64)) ? b : a,
           (FA ? (as_f(a) == MAGIC_F64) : ((int64_t)a == MAGIC_I64)) ? vb : va)
    }
    // unary ops: computed for valid inputs, validity passes through (:728-917)
    case D_NOT: ROWS(!truthy(a, FA), va)
    case D_ABS_I: ROWS((int64_t)a < 0 ? 0ULL - a : a, va)
    case D_ABS_F: ROWS(bits(fabs(as_f(a))), va)
    case D_FLOOR: ROWS(bits(floor(AF)), va)
    case D_CEIL: ROWS(bits(ceil(AF)), va)
    case D_IS_NULL: ROWS(va ^ 1u, 1u)
    case D_IS_NOT_NULL: ROWS(va, 1u)
    case D_IS_TRUE: ROWS(FA ? (uint64_t)f64_to_i64(as_f(a)) : a, va)
    case D_IS_NOT_TRUE: ROWS(!(va && (FA ? as_f(a) == 1.0 : a == 1)), 1u)
    case D_IS_NOT_FALSE: ROWS(!(va && (FA ? as_f(a) == 0.0 : a == 0)), 1u)
    case D_TO_I64: ROWS(FA ? (uint64_t)f64_to_i64(as_f(a)) : a, va)
    case D_TO_F64: ROWS(bits(to_f(a, FA)), va)
    case D_CAST_DATE: {
      const bool sub_day = is_ts(op.aux2) && op.aux2 != BSQL_TIMESTAMP_DAYS;
      const int64_t tpd = ticks_per_day(op.aux2);
      ROWS(sub_day ? (uint64_t)((FA ? f64_to_i64(as_f(a)) : (int64_t)a) / tpd)
                   : (uint64_t)(FA ? f64_to_i64(as_f(a)) : (int64_t)a), va)
    }
    case D_CAST_TS: {
      const uint64_t f = is_ts(op.aux2) ? (uint64_t)(86400000000000LL / ticks_per_day(op.aux2)) : 1ULL;
      ROWS((FA ? (uint64_t)f64_to_i64(as_f(a)) : a) * f, va)
    }
    case D_MOD_FF: case D_POW: case D_ROUND:
      ROWS(compute_heavy(op.code, op.aux, op.aux2, op.imm, AF, BF, a, rs.row[k]), va & vb)
    case D_SIN: case D_COS: case D_ASIN: case D_ACOS: case D_TAN: case D_COTAN: case D_ATAN:
    case D_LN: case D_LOG: case D_EXTRACT:
      ROWS(compute_heavy(op.code, op.aux, op.aux2, op.imm, AF, BF, a, rs.row[k]), va)
    case D_RAND:
      ROWS(compute_heavy(op.code, op.aux, op.aux2, op.imm, 0.0, 0.0, 0, rs.pos0 + (int64_t)k * rs.pstride), 1u)
    case D_EQ_SS: case D_NE_SS: case D_LT_SS: case D_GT_SS: case D_LE_SS: case D_GE_SS: {
      const int kind = op.code - D_EQ_SS;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u) && ((VB >> k) & 1u)) {
          const GLOBAL uint8_t *pa, *pb;
          int32_t na, nb;
          str_at(plan, op.aux, A[k], &pa, &na);
          str_at(plan, op.aux2, B[k], &pb, &nb);
          D[k] = (uint64_t)cmp_result(kind, str_cmp(pa, na, pb, nb));
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_CHAR_LENGTH: {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          D[k] = (uint64_t)utf8_chars(p, n);
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_LIKE: {
      const GLOBAL int32_t* pat = gp<int32_t>((const void*)op.imm);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          D[k] = like_match(p, n, pat) ? 1 : 0;
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_SLICE: {
      // slice_strings positions (LogicalProject.cpp:183-306): a literal
      // arrives clamped; a computed SUBSTRING start is made 0-based, a null
      // start or count reads as 0, and a negative start as 0
      const bool lit = (op.flags & F_SLICE_LIT) != 0;
      const int kind = op.aux2;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = A[k];
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          int64_t pos = ((VB >> k) & 1u) ? (int64_t)B[k] : 0;
          if (!lit) pos = kind == 0 ? pos - (((VB >> k) & 1u) ? 1 : 0) : pos;
          if (pos < 0) pos = 0;
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          const uint32_t st = (uint32_t)A[k];
          if (kind == 0) {
            const int32_t b0 = utf8_offset(p, n, pos);
            D[k] = slice_pack(st + (uint32_t)b0, (uint32_t)(n - b0));
          } else if (kind == 1) {
            D[k] = slice_pack(st, (uint32_t)utf8_offset(p, n, pos));
          } else {
            const int64_t total = utf8_chars(p, n);
            const int32_t b0 = total > pos ? utf8_offset(p, n, total - pos) : 0;
            D[k] = slice_pack(st + (uint32_t)b0, (uint32_t)(n - b0));
          }
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_TRIM: {
      const GLOBAL int32_t* lit = gp<int32_t>((const void*)op.imm);
      const int32_t nset = lit[0];
      const GLOBAL uint8_t* set = (const GLOBAL uint8_t*)(lit + 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = A[k];
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          int32_t b0, b1;
          strip_range(p, n, set, nset, op.aux2, &b0, &b1);
          D[k] = slice_pack((uint32_t)A[k] + (uint32_t)b0, (uint32_t)(b1 - b0));
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_STR_TO_TS: {
      const int32_t* fmt = op.imm ? (const int32_t*)op.imm
                                  : (op.aux2 == BSQL_TIMESTAMP_DAYS ? TS_FMT_DATE : TS_FMT_DATETIME);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          D[k] = (uint64_t)str_to_ticks(p, n, fmt, op.aux2);
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_STR_TO_I64: case D_STR_TO_F64: {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          D[k] = op.code == D_STR_TO_I64 ? (uint64_t)str_to_i64(p, n) : bits(str_to_f64(p, n));
          VD |= 1u << k;
        }
      }
      return;
    }
    default:
#pragma unroll
      for (int k = 0; k < R; ++k) D[k] = A[k];
      return;
  }
}
#undef ROWS

// --------------------------------------------------------- compaction state

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
  return __hip_atomic_load((const GLOBAL uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
  __hip_atomic_store((GLOBAL uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

// Exclusive prefix of this tile across all earlier tiles (wave 0 only):
// decoupled look-back over a 64-tile window per iteration.
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

// One op of the plan, read as two dwords and the literal so that the whole
// wave reads it with scalar loads (the byte fields are unpacked in SGPRs).
__device__ __forceinline__ DevOp load_op(const DevPlan& plan, int q) {
  const CONSTANT uint32_t* p = (const CONSTANT uint32_t*)(plan.ops + q);
  const uint32_t w0 = p[0], w1 = p[1];
  DevOp op;
  op.code = (uint8_t)w0;
  op.dst = (uint8_t)(w0 >> 8);
  op.a = (uint8_t)(w0 >> 16);
  op.b = (uint8_t)(w0 >> 24);
  op.flags = (uint8_t)w1;
  op.aux = (uint8_t)(w1 >> 8);
  op.aux2 = (uint8_t)(w1 >> 16);
  op.pad = 0;
  op.imm = *(const CONSTANT int64_t*)(p + 2);
  return op;
}

// -------------------------------------------------------------- rows

// The rows one lane works on: row[k] for k < R where bit k of `exist` is set.
// `contig` means row[k] == row0 + k (vector loads and stores allowed).
// pos0 + k * pstride is the row's position in the table the ops see (RAND's
// sequence number): the input row, or the output row after a filter.
template <int R>
struct Rows {
  int64_t row[R];
  uint32_t exist;
  int64_t row0;
  int64_t pos0;
  int nrows;
  int pstride;
  bool contig;
};

// gathers R elements of column c at arbitrary rows (converted to slot bits)
template <int R, bool NT = true>
__device__ __forceinline__ void gather_rows(const DevInCol& c, const Rows<R>& rs, uint64_t (&v)[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    v[k] = 0;
    if (!((rs.exist >> k) & 1u)) continue;
    const int64_t r = rs.row[k];
    switch (c.type) {
      case BSQL_INT8: v[k] = (uint64_t)(int64_t)gp<int8_t>(c.data)[r]; break;
      case BSQL_INT16: v[k] = (uint64_t)(int64_t)gp<int16_t>(c.data)[r]; break;
      case BSQL_INT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
        v[k] = (uint64_t)(int64_t)gp<int32_t>(c.data)[r]; break;
      case BSQL_UINT8: v[k] = gp<uint8_t>(c.data)[r]; break;
      case BSQL_BOOL8: v[k] = gp<uint8_t>(c.data)[r] != 0; break;
      case BSQL_UINT16: v[k] = gp<uint16_t>(c.data)[r]; break;
      case BSQL_UINT32: v[k] = gp<uint32_t>(c.data)[r]; break;
      case BSQL_FLOAT32: v[k] = bits((double)gp<float>(c.data)[r]); break;
      default:  // 8-byte types
        v[k] = NT ? __builtin_nontemporal_load(gp<uint64_t>(c.data) + r) : gp<uint64_t>(c.data)[r];
        break;
    }
  }
}

template <int R>
__device__ __forceinline__ uint32_t gather_valid(const DevInCol& c, const Rows<R>& rs) {
  if (!c.valid) return rs.exist;
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (!((rs.exist >> k) & 1u)) continue;
    const int64_t b = (int64_t)c.vbit + rs.row[k];
    m |= ((gp<uint32_t>(c.valid)[b >> 5] >> (b & 31)) & 1u) << k;
  }
  return m;
}

template <int R, bool NT = true>
__device__ __forceinline__ void col_load(const DevInCol& c, const Rows<R>& rs, uint64_t (&v)[R], uint32_t& vm) {
  if (rs.contig) {
    load_rows<R>(c, rs.row0, rs.nrows, v);
    vm = load_valid<R>(c, rs.row0, rs.nrows);
  } else {
    gather_rows<R, NT>(c, rs, v);
    vm = gather_valid<R>(c, rs);
  }
}

// -------------------------------------------------------------- VM step
//
// Register file: one vector of NREG slot values per row the lane owns,
// f[k][reg], and validity bit-planes vb (bit reg*R + k). Every op reads its
// two operands with the wave-uniform register numbers and writes exactly one
// register (result-less ops write the spare register NREG-1), so the file
// stays in VGPRs and is read and written with relative (s_set_gpr_idx) moves.
// Column loads are split into LOADN, which issues up to four columns' loads
// into fixed temporaries so that they are all in flight together, and one
// COMMIT per column.

template <int N>
struct RegVecT {
  typedef uint64_t type __attribute__((ext_vector_type(N)));
};

template <int R>
struct LoadTemps {
  uint64_t t0[R], t1[R], t2[R], t3[R];
  uint32_t v0, v1, v2, v3;
};

// Executes one non-store op. Returns true if it was a KILL that left the
// wave without a selected row.
// Spill slot q of the calling wave: 64 lanes x R rows of data (row k of lane
// t at k*64 + t), then one validity byte per lane (R bits).
template <int R>
__device__ __forceinline__ LDS uint64_t* spill_slot(LDS uint64_t* sp, int64_t q) {
  return sp + q * (64 * R + 8);
}

template <int NREG, int R, bool NT = true>
__device__ __forceinline__ bool vm_step(const DevPlan& plan, const DevOp& op,
                                        typename RegVecT<NREG>::type (&f)[R], uint64_t& vb,
                                        LoadTemps<R>& T, const Rows<R>& rs, uint32_t& alive,
                                        LDS uint64_t* sp) {
  constexpr uint32_t RMASK = (1u << R) - 1u;
  const uint8_t code = op.code;
  uint64_t A[R], B[R], D[R];
  const int ia = op.a & (NREG - 1), ib = op.b & (NREG - 1);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    A[k] = f[k][ia];
    B[k] = f[k][ib];
  }
  uint32_t VA = (uint32_t)(vb >> (ia * R)) & RMASK, VB = (uint32_t)(vb >> (ib * R)) & RMASK, VD = 0;
  if (op.flags & F_A_IMM) {
#pragma unroll
    for (int k = 0; k < R; ++k) A[k] = (uint64_t)op.imm;
    VA = RMASK;
  }
  if (op.flags & F_B_IMM) {
#pragma unroll
    for (int k = 0; k < R; ++k) B[k] = (uint64_t)op.imm;
    VB = RMASK;
  }
  if (op.flags & F_A_NULL) VA = 0;
  if (op.flags & F_B_NULL) VB = 0;
  int dst = op.dst == 0xff ? NREG - 1 : (op.dst & (NREG - 1));
  bool dead = false;
  if (code == D_LOADN) {
    const int n = op.aux;
    col_load<R, NT>(plan.in[op.imm & 0xff], rs, T.t0, T.v0);
    if (n > 1) col_load<R, NT>(plan.in[(op.imm >> 8) & 0xff], rs, T.t1, T.v1);
    if (n > 2) col_load<R, NT>(plan.in[(op.imm >> 16) & 0xff], rs, T.t2, T.v2);
    if (n > 3) col_load<R, NT>(plan.in[(op.imm >> 24) & 0xff], rs, T.t3, T.v3);
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = A[k];
    dst = NREG - 1;
  } else if (code == D_COMMIT) {
    const int j = op.aux;
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = j == 0 ? T.t0[k] : (j == 1 ? T.t1[k] : (j == 2 ? T.t2[k] : T.t3[k]));
    VD = j == 0 ? T.v0 : (j == 1 ? T.v1 : (j == 2 ? T.v2 : T.v3));
  } else if (code == D_LOADS) {
    const DevInCol& c = plan.in[op.a];
    const GLOBAL int32_t* off = gp<int32_t>(c.offsets);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      D[k] = 0;
      if ((rs.exist >> k) & 1u) {
        const int32_t s0 = off[rs.row[k]], s1 = off[rs.row[k] + 1];
        D[k] = slice_pack((uint32_t)s0, (uint32_t)(s1 - s0));
      }
    }
    VD = rs.contig ? load_valid<R>(c, rs.row0, rs.nrows) : gather_valid<R>(c, rs);
  } else if (code == D_SPILL) {
    const int lane = threadIdx.x & 63;
    LDS uint64_t* slot = spill_slot<R>(sp, op.imm);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      slot[k * 64 + lane] = A[k];
      D[k] = A[k];
    }
    ((LDS uint8_t*)(slot + 64 * R))[lane] = (uint8_t)VA;
    dst = NREG - 1;
  } else if (code == D_FILL) {
    const int lane = threadIdx.x & 63;
    const LDS uint64_t* slot = spill_slot<R>(sp, op.imm);
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = slot[k * 64 + lane];
    VD = ((const LDS uint8_t*)(slot + 64 * R))[lane];
  } else if (code == D_KILL || code == D_SEL) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const bool t = code == D_KILL ? truthy(A[k], FA) : (A[k] != 0);
      if (!(((VA >> k) & 1u) && t)) alive &= ~(1u << k);
      D[k] = A[k];
    }
    dst = NREG - 1;
    dead = code == D_KILL && __ballot(alive != 0) == 0;
  } else {
    compute_rows<R>(plan, op, A, B, VA, VB, D, VD, rs);
  }
  // result write: the only write of the register file
#pragma unroll
  for (int k = 0; k < R; ++k) f[k][dst] = D[k];
  vb = (vb & ~((uint64_t)RMASK << (dst * R))) | ((uint64_t)(VD & RMASK) << (dst * R));
  return dead;
}

// -------------------------------------------------------------- the kernels

// Stores the R results of one STORE op for rows at output positions
// [pos0, pos0 + n) and their validity (`vbits`, bit k = row k). pos0 is a
// multiple of R and a tile starts at a multiple of 32, so the 32/R lanes of
// an aligned lane group own one whole validity word: it is assembled with
// shuffles and written with one plain store.
template <int R>
__device__ __forceinline__ void store_out(const DevOutCol& c, int64_t pos0, int n, const uint64_t (&A)[R],
                                          uint32_t vbits, int lane) {
  if (n > 0) store_rows<R>(c, pos0, n, A);
  if (c.valid) {
    constexpr int LPW = 32 / R;
    uint32_t word = (vbits & ((1u << R) - 1u)) << ((lane % LPW) * R);
#pragma unroll
    for (int s = 1; s < LPW; s <<= 1) word |= __shfl_xor(word, s, 64);
    if ((lane % LPW) == 0 && n > 0) gpw<uint32_t>(c.valid)[pos0 >> 5] = word;
  }
}

// Stores one STORE op's results for the output positions pos0 + 64*k (the
// lane-strided rows of a gather tile; bit k of `exist` = position k is
// below the row count) and their validity (`vbits`): for each k the wave's
// 64 positions are one aligned 64-bit validity span, built by one ballot
// and written as two words by lanes 0 and 32.
template <int R>
__device__ __forceinline__ void store_out_strided(const DevOutCol& c, int64_t pos0, uint32_t exist,
                                                  const uint64_t (&A)[R], uint32_t vbits, int lane) {
#pragma unroll
  for (int k = 0; k < R; ++k)
    if ((exist >> k) & 1u) store_elem<true>(c, pos0 + 64 * k, A[k]);
  if (c.valid) {
    const int64_t w0 = (pos0 - lane) >> 5;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const uint64_t b = __ballot((vbits >> k) & 1u);
      if ((lane & 31) == 0 && ((exist >> k) & 1u))
        gpw<uint32_t>(c.valid)[w0 + 2 * k + (lane >> 5)] = (uint32_t)(b >> (lane & 32));
    }
  }
}

// Ops [plan.split, n_ops) over one tile of B*R positions [tile*TILE,
// count): positions are rows (or, GATHER, indexes into the row-id list idx).
// Rows: lane t of wave w owns positions tile*TILE + w*64*R + t*R + k
// (consecutive, vector loads and stores) or, GATHER, tile*TILE + w*64*R +
// 64*k + t, so that each gather instruction of the wave reads 64 consecutive
// selected rows: neighbouring lanes share 128-byte lines, which are fetched
// once per instruction rather than once per lane.
// The register file is local to the tile: every register is written before
// it is read, so nothing carries over from one tile to the next. `sp` is the
// wave's spill area (plan.n_spill slots).
template <int NREG, int R, bool GATHER, int B>
__device__ __forceinline__ void vm_tile(const DevPlan& plan, int64_t tile, int64_t count,
                                        const uint32_t* __restrict__ idx, int lane, LDS uint64_t* sp) {
  typedef typename RegVecT<NREG>::type RegVec;
  constexpr int TILE = B * R;
  RegVec f[R];
#pragma unroll
  for (int k = 0; k < R; ++k) f[k] = (RegVec)(0);
  uint64_t vb = 0;
  LoadTemps<R> T;
  Rows<R> rs;
  int n;
  if (GATHER) {
    rs.pos0 = tile * TILE + (int64_t)(threadIdx.x >> 6) * (64 * R) + lane;
    rs.pstride = 64;
    const int64_t left = count - rs.pos0;  // positions pos0 + 64k < count
    n = left <= 0 ? 0 : (left >= 64 * (R - 1) + 1 ? R : (int)((left + 63) >> 6));
    rs.nrows = n;
    rs.exist = (1u << n) - 1u;
    rs.contig = false;
    rs.row0 = 0;
#pragma unroll
    for (int k = 0; k < R; ++k)
      rs.row[k] = k < n ? (int64_t)__builtin_nontemporal_load(gp<uint32_t>(idx) + rs.pos0 + 64 * k) : 0;
  } else {
    rs.pos0 = tile * TILE + (int64_t)threadIdx.x * R;
    rs.pstride = 1;
    const int64_t left = count - rs.pos0;
    n = left <= 0 ? 0 : (left >= R ? R : (int)left);
    rs.nrows = n;
    rs.exist = (1u << n) - 1u;
    rs.contig = true;
    rs.row0 = rs.pos0;
#pragma unroll
    for (int k = 0; k < R; ++k) rs.row[k] = rs.pos0 + k;
  }
  uint32_t alive = rs.exist;
  // the next op's scalar loads are issued before this op runs, so their
  // latency hides behind it
  const int n_ops = plan.n_ops;
  DevOp next = load_op(plan, plan.split < n_ops ? plan.split : 0);
  for (int q = plan.split; q < n_ops; ++q) {
    const DevOp op = next;
    next = load_op(plan, q + 1 < n_ops ? q + 1 : q);
    if (op.code != D_STORE) {
      vm_step<NREG, R>(plan, op, f, vb, T, rs, alive, sp);
      continue;
    }
    const int ia = op.a & (NREG - 1);
    uint64_t A[R];
#pragma unroll
    for (int k = 0; k < R; ++k) A[k] = f[k][ia];
    const uint32_t VA = (uint32_t)(vb >> (ia * R));
    if (GATHER) store_out_strided<R>(plan.out[op.aux], rs.pos0, rs.exist, A, VA & rs.exist, lane);
    else store_out<R>(plan.out[op.aux], rs.pos0, n, A, VA & rs.exist, lane);
  }
}

// The calling wave's spill area in the workgroup's dynamic LDS.
template <int R>
__device__ __forceinline__ LDS uint64_t* wave_spill(LDS uint64_t* base, int n_spill) {
  return base + (int64_t)(threadIdx.x >> 6) * n_spill * (64 * R + 8);
}

extern __shared__ uint64_t s_spill_area[];

// perform_interpreter_operation (rows are the positions [0, num_rows)) or
// the output phase of a filter (GATHER: rows are the selected row ids
// idx[0, count), outputs at the positions in that list). Persistent:
// workgroups stride over tiles of B*R positions.
template <int NREG, int R, bool GATHER, int B>
__global__ __launch_bounds__(B) void vm_rows(const DevPlan* __restrict__ pplan,
                                             const uint32_t* __restrict__ idx,
                                             const int64_t* __restrict__ d_count, int64_t capacity) {
  constexpr int TILE = B * R;
  const DevPlan& plan = *pplan;  // device memory: ops are scalar-cache loads
  const int lane = threadIdx.x & 63;
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plan.n_spill);
  int64_t count = plan.num_rows;
  if (GATHER) {
    count = *d_count;
    if (count > capacity) count = capacity;
  }
  for (int64_t tile = blockIdx.x; tile * TILE < count; tile += gridDim.x)
    vm_tile<NREG, R, GATHER, B>(plan, tile, count, idx, lane, sp);
}

// Batched interpreter: every batch has its own plan (program and columns);
// the work items are all batches' tiles, batch b owning items
// [tile_start[b], tile_start[b+1]). Workgroups stride over the items, so
// uneven batch sizes spread over the whole chip in one persistent launch.
template <int NREG, int R, int B>
__global__ __launch_bounds__(B) void vm_batches(const DevPlan* __restrict__ plans,
                                                const int64_t* __restrict__ tile_start, int n_batches) {
  const int lane = threadIdx.x & 63;
  const int64_t n_items = tile_start[n_batches];
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plans[0].n_spill);
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    int lo = 0, hi = n_batches - 1;  // last batch whose first item <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const DevPlan& plan = plans[lo];
    vm_tile<NREG, R, false, B>(plan, item - tile_start[lo], plan.num_rows, nullptr, lane, sp);
  }
}

// Selection tiles: a wave slice holds S chunks of 64*R rows, and lane t of a
// wave owns rows [t*R, t*R+R) of each chunk, so every load instruction of a
// wave reads one contiguous span; a workgroup's WAVES slices make a tile.
template <int R, int S>
struct SelectTile {
  static constexpr int CHUNK = 64 * R;
  static constexpr int WAVE_ROWS = S * CHUNK;
  static constexpr int TILE = WAVES * WAVE_ROWS;
  static_assert(R * S <= 32, "selection bits per lane");
};

template <int R>
__device__ __forceinline__ void chunk_rows(Rows<R>& rs, int64_t row0, int64_t num_rows) {
  rs.row0 = row0;
  rs.pos0 = row0;
  rs.pstride = 1;
  const int64_t left = num_rows - row0;
  rs.nrows = left <= 0 ? 0 : (left >= R ? R : (int)left);
  rs.exist = (1u << rs.nrows) - 1u;
  rs.contig = true;
#pragma unroll
  for (int k = 0; k < R; ++k) rs.row[k] = row0 + k;
}

// ------------------------------------------------- selection output
//
// The selection kernels impose no order between workgroups: workgroup g owns
// the contiguous block of "slices" [g*per, (g+1)*per) (a slice is one wave's
// S*R*64 rows). For each slice a wave writes S*R 64-bit ballots (ballot
// q = k*R + j, bit t = row k*64*R + t*R + j of the slice) and the selected
// count; the workgroup writes its block's total. slice_ids sums the block
// totals before its block, rebuilds the slice offsets of its block in LDS and
// writes the selected row ids in row order; its last workgroup writes the row
// total. Without output ops, block_scan (one workgroup) writes the row total.
// No workgroup waits on another: no ticket, no look-back.
struct SelArgs {
  uint64_t* bits;        // n_slices * S * R ballots
  uint32_t* counts;      // n_slices
  uint32_t* block_sums;  // grid: selected rows per workgroup block
  uint32_t* block_base;  // grid: exclusive prefix of block_sums
  int64_t n_slices;
  int64_t per;           // slices per workgroup block
};

template <int R, int S>
__device__ __forceinline__ uint32_t emit_slice(const SelArgs& sa, int64_t slice, uint32_t alive) {
  const int lane = threadIdx.x & 63;
  uint64_t mine = 0;
  uint32_t total = 0;
#pragma unroll
  for (int q = 0; q < S * R; ++q) {
    const uint64_t b = __ballot((alive >> q) & 1u);
    total += (uint32_t)__popcll(b);
    mine = lane == q ? b : mine;
  }
  if (lane < S * R) gpw<uint64_t>(sa.bits)[slice * (S * R) + lane] = mine;
  if (lane == 0) gpw<uint32_t>(sa.counts)[slice] = total;
  return total;
}

// The block total of a selection workgroup (every thread calls it once,
// after its wave's last slice).
template <int B = BLOCK>
__device__ __forceinline__ void emit_block(const SelArgs& sa, uint32_t wave_total) {
  constexpr int NW = B / 64;
  __shared__ uint32_t s_sum[NW];
  if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = wave_total;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += s_sum[w];
    gpw<uint32_t>(sa.block_sums)[blockIdx.x] = t;
  }
}

// Exclusive prefix of the block totals, and the row total (one workgroup).
constexpr int SCAN_THREADS = 1024;
__global__ __launch_bounds__(SCAN_THREADS) void block_scan(const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ base, int64_t n,
                                                           int64_t* __restrict__ d_count) {
  __shared__ uint64_t s_part[SCAN_THREADS / 64];
  uint64_t carry = 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t i0 = 0; i0 < n; i0 += SCAN_THREADS) {
    const int64_t i = i0 + threadIdx.x;
    const uint64_t v = i < n ? gp<uint32_t>(sums)[i] : 0;
    uint64_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    if (lane == 63) s_part[wave] = incl;
    __syncthreads();
    uint64_t before = carry;
    for (int w = 0; w < wave; ++w) before += s_part[w];
    if (i < n) gpw<uint32_t>(base)[i] = (uint32_t)(before + incl - v);
    uint64_t all = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) all += s_part[w];
    __syncthreads();
    carry += all;
  }
  if (threadIdx.x == 0) *gpw<int64_t>(d_count) = (int64_t)carry;
}

// Row ids of the selected rows, in row order. Workgroup g takes the same
// block of slices as in the selection. Its first id's position is the sum of
// the block totals before it (at most a few thousand words, read from L2);
// the last workgroup writes the row total. The block's slice offsets are
// scanned in LDS, BLOCK slices at a time.
template <int R, int S>
__global__ __launch_bounds__(BLOCK) void slice_ids(SelArgs sa, uint32_t* __restrict__ idx,
                                                   int64_t* __restrict__ d_count) {
  constexpr int CHUNK = 64 * R;
  constexpr int SLICE = S * CHUNK;
  __shared__ uint32_t s_off[BLOCK];
  __shared__ uint32_t s_cnt[BLOCK];
  __shared__ uint32_t s_part[WAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const GLOBAL uint32_t* sums = gp<uint32_t>(sa.block_sums);
  uint32_t before_block = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += BLOCK) before_block += sums[b];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) before_block += __shfl_xor(before_block, d, 64);
  if (lane == 0) s_part[wave] = before_block;
  __syncthreads();
  uint32_t carry = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) carry += s_part[w];
  __syncthreads();  // s_part is reused below
  const uint32_t mine_total = sums[blockIdx.x];
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *gpw<int64_t>(d_count) = (int64_t)carry + mine_total;
  // a block that selected nothing has no ids to write (workgroup-uniform)
  if (mine_total == 0) return;
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  GLOBAL uint32_t* ids = gpw<uint32_t>(idx);
  const int64_t s0 = (int64_t)blockIdx.x * sa.per;
  const int64_t s1 = s0 + sa.per < sa.n_slices ? s0 + sa.per : sa.n_slices;
  for (int64_t c0 = s0; c0 < s1; c0 += BLOCK) {
    // offsets of slices [c0, c0 + BLOCK) within the block
    const int64_t mys = c0 + threadIdx.x;
    const uint32_t v = mys < s1 ? gp<uint32_t>(sa.counts)[mys] : 0u;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    if (lane == 63) s_part[wave] = incl;
    __syncthreads();
    uint32_t before = carry;
    for (int w = 0; w < wave; ++w) before += s_part[w];
    s_off[threadIdx.x] = before + incl - v;
    s_cnt[threadIdx.x] = v;
    uint32_t all = 0;
    for (int w = 0; w < WAVES; ++w) all += s_part[w];
    __syncthreads();
    const int64_t c1 = c0 + BLOCK < s1 ? c0 + BLOCK : s1;
    for (int64_t slice = c0 + wave; slice < c1; slice += WAVES) {
      if (s_cnt[slice - c0] == 0) continue;  // wave-uniform: no rows, no ballots to read
      const uint64_t mine = lane < S * R ? gp<uint64_t>(sa.bits)[slice * (S * R) + lane] : 0ULL;
      uint32_t off = s_off[slice - c0];
      const uint32_t row0 = (uint32_t)(slice * SLICE) + (uint32_t)(lane * R);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        uint64_t b[R];
        uint32_t before_k = 0, chunk_total = 0;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)mine, k * R + j);
          const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(mine >> 32), k * R + j);
          b[j] = ((uint64_t)hi << 32) | lo;
          before_k += (uint32_t)__popcll(b[j] & below);
          chunk_total += (uint32_t)__popcll(b[j]);
        }
        uint32_t pos = off + before_k;
#pragma unroll
        for (int j = 0; j < R; ++j)
          if ((b[j] >> lane) & 1ULL) ids[pos++] = row0 + (uint32_t)(k * CHUNK + j);
        off += chunk_total;
      }
    }
    __syncthreads();  // s_off is rewritten for the next BLOCK slices
    carry += all;
  }
}

// Counts of one chunk's selection bits `a` (R bits per lane): the lane's
// exclusive prefix within the chunk and the chunk's total over the wave.
template <int R>
__device__ __forceinline__ void chunk_counts(uint32_t a, int& lane_excl, int& total) {
  lane_excl = 0;
  total = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint64_t b = __ballot((a >> j) & 1u);
    lane_excl += __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
    total += __popcll(b);
  }
}

// Column element kinds of the range tests: the raw element type and its
// value in the int64 slot domain (as the VM's column loads convert it).
struct K_I8 { typedef int8_t T; static constexpr bool F = false; };
struct K_I16 { typedef int16_t T; static constexpr bool F = false; };
struct K_I32 { typedef int32_t T; static constexpr bool F = false; };
struct K_I64 { typedef int64_t T; static constexpr bool F = false; };
struct K_U8 { typedef uint8_t T; static constexpr bool F = false; };
struct K_U16 { typedef uint16_t T; static constexpr bool F = false; };
struct K_U32 { typedef uint32_t T; static constexpr bool F = false; };
struct K_B8 { typedef uint8_t T; static constexpr bool F = false; };
struct K_F32 { typedef float T; static constexpr bool F = true; };
struct K_F64 { typedef double T; static constexpr bool F = true; };

template <class K>
__device__ __forceinline__ int64_t kind_i64(typename K::T x) {
  if constexpr (std::is_same<K, K_B8>::value) return x != 0;
  else if constexpr (K::F) return 0;  // float columns are only compared as doubles
  else return (int64_t)x;
}
template <class K>
__device__ __forceinline__ double kind_f64(typename K::T x) {
  if constexpr (K::F) return (double)x;
  else return (double)kind_i64<K>(x);
}

// Raw elements of R consecutive rows starting at r0 (n of them exist), with
// the widest aligned loads the element size allows.
template <class CT, int R, bool NT = true>
__device__ __forceinline__ void load_raw(const DevInCol& c, int64_t r0, int n, CT (&v)[R]) {
  const GLOBAL CT* p = gp<CT>(c.data) + r0;
  constexpr int BYTES = (int)sizeof(CT) * R;
  if (n == R && c.aligned16 && BYTES >= 4 && BYTES % 4 == 0) {
    if constexpr (BYTES >= 16) {
#pragma unroll
      for (int k = 0; k < BYTES / 16; ++k)
        ((v4u32*)v)[k] = NT ? __builtin_nontemporal_load((const GLOBAL v4u32*)p + k) : ((const GLOBAL v4u32*)p)[k];
    } else if constexpr (BYTES == 8) {
      *(v2u32*)v = NT ? __builtin_nontemporal_load((const GLOBAL v2u32*)p) : *(const GLOBAL v2u32*)p;
    } else {
      *(uint32_t*)v = NT ? __builtin_nontemporal_load((const GLOBAL uint32_t*)p) : *(const GLOBAL uint32_t*)p;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = k < n ? p[k] : CT(0);
  }
}

// In-range bits of the R elements v (bit j) for range rg.
template <class K, int R>
__device__ __forceinline__ uint32_t range_bits(const DevRange& rg, const typename K::T (&v)[R]) {
  uint32_t pass = 0;
  if (rg.kind & RK_F) {
    const double lo = as_f((uint64_t)rg.lo), hi = as_f((uint64_t)rg.hi);
    if (rg.kind & RK_NE) {
#pragma unroll
      for (int j = 0; j < R; ++j) pass |= (uint32_t)(kind_f64<K>(v[j]) != lo) << j;
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const double x = kind_f64<K>(v[j]);
        pass |= (uint32_t)(x >= lo && x <= hi) << j;
      }
    }
  } else if constexpr (!K::F) {
    const uint64_t mul = (uint64_t)rg.mul;
    if (rg.kind & RK_NE) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        pass |= (uint32_t)((int64_t)((uint64_t)kind_i64<K>(v[j]) * mul) != rg.lo) << j;
    } else if (mul == 1) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int64_t x = kind_i64<K>(v[j]);
        pass |= (uint32_t)(x >= rg.lo && x <= rg.hi) << j;
      }
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int64_t x = (int64_t)((uint64_t)kind_i64<K>(v[j]) * mul);
        pass |= (uint32_t)(x >= rg.lo && x <= rg.hi) << j;
      }
    }
  }
  return pass;
}

// One range test over this lane's S chunks: returns alive & valid & in range
// (bit k*R + j = row j of chunk k). A full tile issues every chunk's loads
// before the first test; the last, partial tile goes chunk by chunk with
// guarded loads.
template <class K, int R, int S, bool FULL, bool NT = true>
__device__ __forceinline__ uint32_t range_chunks(const DevInCol& c, const DevRange& rg, int64_t row0,
                                                 int64_t num_rows, uint32_t alive) {
  typedef typename K::T CT;
  constexpr uint32_t RMASK = (1u << R) - 1u;
  uint32_t pass = 0;
  if constexpr (FULL) {
    // a chunk whose rows are all dead in the whole wave is not read: the
    // earlier (more selective) conjuncts spare the later columns' bytes
    uint32_t live = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) live |= (__ballot(((alive >> (k * R)) & RMASK) != 0) != 0 ? 1u : 0u) << k;
    CT v[S][R];
    uint32_t vm = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      // only lanes with a live row load (exec-masked loads): a cache line
      // whose rows are all dead is not fetched
      if (((live >> k) & 1u) && ((alive >> (k * R)) & RMASK)) {
        const int64_t r0 = row0 + (int64_t)k * 64 * R;
        load_raw<CT, R, NT>(c, r0, R, v[k]);
        vm |= load_valid<R>(c, r0, R) << (k * R);
      } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[k][j] = CT(0);
      }
    }
#pragma unroll
    for (int k = 0; k < S; ++k) pass |= range_bits<K, R>(rg, v[k]) << (k * R);
    return alive & vm & pass;
  } else {
    uint32_t vm = 0;
#pragma unroll 1
    for (int k = 0; k < S; ++k) {
      const int64_t r0 = row0 + (int64_t)k * 64 * R;
      const int64_t left = num_rows - r0;
      const int n = ((alive >> (k * R)) & RMASK) == 0 ? 0 : (left <= 0 ? 0 : (left >= R ? R : (int)left));
      CT v[R];
      load_raw<CT, R, NT>(c, r0, n, v);
      vm |= load_valid<R>(c, r0, n) << (k * R);
      pass |= range_bits<K, R>(rg, v) << (k * R);
    }
    return alive & vm & pass;
  }
}

template <int R, int S, bool FULL, bool NT = true>
__device__ __forceinline__ uint32_t range_test(const DevInCol& c, const DevRange& rg, int64_t row0,
                                               int64_t num_rows, uint32_t alive) {
  switch (c.type) {
    case BSQL_INT8: return range_chunks<K_I8, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_INT16: return range_chunks<K_I16, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_INT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
      return range_chunks<K_I32, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_UINT8: return range_chunks<K_U8, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_UINT16: return range_chunks<K_U16, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_UINT32: return range_chunks<K_U32, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_BOOL8: return range_chunks<K_B8, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_FLOAT32: return range_chunks<K_F32, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    case BSQL_FLOAT64: return range_chunks<K_F64, R, S, FULL, NT>(c, rg, row0, num_rows, alive);
    default: return range_chunks<K_I64, R, S, FULL, NT>(c, rg, row0, num_rows, alive);  // 8-byte ticks/ints
  }
}

// Row selection by range tests (plan.ranges): one range at a time over the
// wave's whole slice, so each column is loaded with every chunk in flight
// and its type is dispatched once per slice; no register file. Waves stride
// over slices independently.
template <int R, int S>
__global__ __launch_bounds__(BLOCK) void select_ranges(const DevPlan* __restrict__ pplan, SelArgs sa) {
  typedef SelectTile<R, S> T;
  const DevPlan& plan = *pplan;
  const int lane = threadIdx.x & 63;
  const int64_t num_rows = plan.num_rows;
  const int nr = plan.n_ranges;
  const int64_t s0 = (int64_t)blockIdx.x * sa.per;
  const int64_t s1 = s0 + sa.per < sa.n_slices ? s0 + sa.per : sa.n_slices;
  uint32_t wave_total = 0;
  for (int64_t slice = s0 + (threadIdx.x >> 6); slice < s1; slice += WAVES) {
    const int64_t slice_row0 = slice * T::WAVE_ROWS;
    const int64_t row0 = slice_row0 + lane * R;
    uint32_t alive = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int64_t left = num_rows - (row0 + (int64_t)k * T::CHUNK);
      const int n = left <= 0 ? 0 : (left >= R ? R : (int)left);
      alive |= ((1u << n) - 1u) << (k * R);
    }
    const bool full = slice_row0 + T::WAVE_ROWS <= num_rows;
    for (int r = 0; r < nr; ++r) {
      if (__ballot(alive != 0) == 0) break;
      const DevRange rg = plan.ranges[r];
      if (rg.kind & RK_EMPTY) {
        alive = 0;
        break;
      }
      if (full)
        alive = range_test<R, S, true>(plan.in[rg.col], rg, row0, num_rows, alive);
      else
        alive = range_test<R, S, false>(plan.in[rg.col], rg, row0, num_rows, alive);
    }
    wave_total += emit_slice<R, S>(sa, slice, alive);
  }
  emit_block(sa, wave_total);
}

// ------------------------------------------------- fused filter + project
//
// One pass for a predicate of range tests (opt-in: BSQL_LAUNCH_FUSED). A
// workgroup is one wave, and a tile is one wave slice of S chunks of 64
// lanes x R rows, taken from ticket counters, so every earlier tile is
// running or done when a tile looks back. The wave tests its slice's ranges
// with the loads of SH chunks in flight, writes the 16-bit offsets of its
// selected rows, in row order, to LDS, publishes its count and finds its
// output position by decoupled look-back. It then runs the output ops over
// its selected rows, 64 consecutive ones per instruction (the predicate's
// columns come back from cache), and writes its outputs at consecutive
// positions. Output validity is assembled in LDS and written as whole
// words, except the two words a tile shares with its neighbours, whose bits
// it sets with an AND/OR atomic pair that leaves the other bits alone.
// Nothing but the tile status words and the tickets needs clearing before
// the launch. Measured on 100M rows (5 conjuncts, 2 outputs) this pass is
// slower than selection + row ids + output pass (1.29 ms against 0.80 ms):
// one wave per workgroup at 205 VGPRs keeps too few loads in flight.
constexpr int FUSED_R = 4;            // rows per lane per chunk
constexpr int FUSED_S = 8;            // chunks per wave slice
constexpr int FUSED_SLICE = 64 * FUSED_R * FUSED_S;  // 2048 rows
constexpr int FUSED_TILE = FUSED_SLICE;              // one wave's tile
constexpr int FUSED_MAX_OUT = 8;
constexpr int FUSED_VWORDS = FUSED_SLICE / 32 + 2;   // validity words a wave's outputs can touch

struct FusedArgs {
  uint64_t* status;   // num_tiles words, zeroed before the launch
  uint32_t* ticket;   // zeroed before the launch
  int64_t* d_count;   // selected rows (written by the last tile)
  int64_t capacity;   // rows the outputs hold; positions beyond are not written
  int64_t num_tiles;
};

// Sets validity bit `pos` (wave-local numbering) of output o in LDS.
__device__ __forceinline__ void lds_set_bit(LDS uint32_t* bits, int pos) {
  __hip_atomic_fetch_or(bits + (pos >> 5), 1u << (pos & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Validity word g of an output: the bits in `own` become those of `val`,
// the others are left as they are (two atomics, each changing only `own`).
__device__ __forceinline__ void merge_word(uint32_t* words, int64_t g, uint32_t own, uint32_t val) {
  GLOBAL uint32_t* w = gpw<uint32_t>(words) + g;
  if (own == 0xffffffffu) {
    *w = val;
    return;
  }
  __hip_atomic_fetch_and(w, val | ~own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_or(w, val & own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The output ops over a wave's selected rows (slice row0 + offs[i] for
// i < cnt), written at positions base + i (base + i < capacity). vbits: the
// wave's LDS validity words, FUSED_VWORDS per output, bit (base & 31) + i.
// Column loads are ordinary (cached) loads: the predicate's columns were
// just read by the range tests.
template <int NREG, int R>
__device__ __forceinline__ void vm_dense(const DevPlan& plan, const LDS uint16_t* offs, int64_t row0, int cnt,
                                         int64_t base, int64_t capacity, int lane, LDS uint32_t* vbits) {
  typedef typename RegVecT<NREG>::type RegVec;
  const int shift = (int)(base & 31);
  for (int c0 = 0; c0 < cnt; c0 += 64 * R) {
    RegVec f[R];
#pragma unroll
    for (int k = 0; k < R; ++k) f[k] = (RegVec)(0);
    uint64_t vb = 0;
    LoadTemps<R> T;
    Rows<R> rs;
    rs.pos0 = base + c0 + lane;
    rs.pstride = 64;
    rs.contig = false;
    rs.row0 = 0;
    rs.exist = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int i = c0 + 64 * k + lane;
      const bool e = i < cnt && base + i < capacity;
      rs.row[k] = e ? row0 + (int64_t)offs[i] : 0;
      rs.exist |= (e ? 1u : 0u) << k;
    }
    rs.nrows = __popc(rs.exist);
    uint32_t alive = rs.exist;
    const int n_ops = plan.n_ops;
    DevOp next = load_op(plan, plan.split < n_ops ? plan.split : 0);
    for (int q = plan.split; q < n_ops; ++q) {
      const DevOp op = next;
      next = load_op(plan, q + 1 < n_ops ? q + 1 : q);
      if (op.code != D_STORE) {
        vm_step<NREG, R, false>(plan, op, f, vb, T, rs, alive, nullptr);
        continue;
      }
      const int ia = op.a & (NREG - 1);
      const uint32_t VA = (uint32_t)(vb >> (ia * R)) & rs.exist;
      const DevOutCol& oc = plan.out[op.aux];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if ((rs.exist >> k) & 1u) store_elem<true>(oc, rs.pos0 + 64 * k, f[k][ia]);
        if ((VA >> k) & 1u) lds_set_bit(vbits + op.aux * FUSED_VWORDS, shift + c0 + 64 * k + lane);
      }
    }
  }
}

// LDS of the fused kernel per workgroup (one wave): the validity words of
// its outputs (dynamic, n_out of FUSED_VWORDS); the 16-bit row offsets of
// its selected rows are static.
int64_t fused_lds_bytes(int n_out) { return (int64_t)n_out * FUSED_VWORDS * 4; }

// Ticket counters: tile 8*t + c is the t-th ticket of counter c, served to
// the workgroups whose index is c mod 8, so that no single address takes
// every ticket. Progress: a tile waits only on smaller tiles; if the
// smallest unfinished tile u (counter c) were not taken, every tile counter
// c has handed out is smaller than u, hence finished, so its workgroups are
// free to take u.
constexpr int FUSED_COUNTERS = 8;

template <int NREG, int R, int SH>
__global__ __launch_bounds__(64) void select_project_fused(const DevPlan* __restrict__ pplan, FusedArgs fa) {
  constexpr int RR = FUSED_R, S = FUSED_S;
  constexpr int CHUNK = 64 * RR;
  static_assert(S % SH == 0, "range loads in whole groups of chunks");
  __shared__ uint16_t s_offs[FUSED_SLICE];
  const DevPlan& plan = *pplan;
  const int lane = threadIdx.x;
  const int64_t num_rows = plan.num_rows;
  const int nr = plan.n_ranges;
  const int n_out = plan.n_out;
  const int counter = (int)(blockIdx.x % FUSED_COUNTERS);
  LDS uint32_t* vbits = (LDS uint32_t*)s_spill_area;
  LDS uint16_t* offs = (LDS uint16_t*)s_offs;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(fa.ticket + counter, 1u);
    const int64_t tile = (int64_t)__builtin_amdgcn_readfirstlane(t) * FUSED_COUNTERS + counter;
    if (tile >= fa.num_tiles) break;  // wave-uniform
    // ---- range tests over the tile, SH chunks' loads in flight at a time
    const int64_t slice_row0 = tile * FUSED_SLICE;
    uint32_t alive = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int64_t left = num_rows - (slice_row0 + lane * RR + (int64_t)k * CHUNK);
      const int n = left <= 0 ? 0 : (left >= RR ? RR : (int)left);
      alive |= ((1u << n) - 1u) << (k * RR);
    }
#pragma unroll 1
    for (int h = 0; h < S / SH; ++h) {
      const int64_t row0 = slice_row0 + (int64_t)h * SH * CHUNK + lane * RR;
      const bool full = slice_row0 + (int64_t)(h + 1) * SH * CHUNK <= num_rows;
      constexpr uint32_t HMASK = SH * RR >= 32 ? 0xffffffffu : ((1u << (SH * RR)) - 1u);
      uint32_t a = (alive >> (h * SH * RR)) & HMASK;
      for (int r = 0; r < nr; ++r) {
        if (__ballot(a != 0) == 0) break;
        const DevRange rg = plan.ranges[r];
        if (rg.kind & RK_EMPTY) {
          a = 0;
          break;
        }
        if (full)
          a = range_test<RR, SH, true, false>(plan.in[rg.col], rg, row0, num_rows, a);
        else
          a = range_test<RR, SH, false, false>(plan.in[rg.col], rg, row0, num_rows, a);
      }
      alive = (alive & ~(HMASK << (h * SH * RR))) | (a << (h * SH * RR));
    }
    // ---- tile offsets of the selected rows, in row order, into LDS
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const uint32_t a = (alive >> (k * RR)) & ((1u << RR) - 1u);
      int lane_excl, total;
      chunk_counts<RR>(a, lane_excl, total);
      int pos = cnt + lane_excl;
#pragma unroll
      for (int j = 0; j < RR; ++j)
        if ((a >> j) & 1u) offs[pos++] = (uint16_t)(k * CHUNK + lane * RR + j);
      cnt += total;
    }
    for (int i = lane; i < n_out * FUSED_VWORDS; i += 64) vbits[i] = 0u;
    // ---- the tile's output position (decoupled look-back)
    CompactArgs ca{fa.status, fa.ticket, fa.d_count, fa.capacity, fa.num_tiles};
    const int64_t base = look_back(ca, tile, cnt, lane);
    if (lane == 0 && tile == fa.num_tiles - 1) *gpw<int64_t>(fa.d_count) = base + cnt;
    __syncthreads();  // one wave: the LDS offsets and cleared words are visible to every lane
    // ---- outputs of the selected rows
    if (cnt > 0) {
      vm_dense<NREG, R>(plan, offs, slice_row0, cnt, base, fa.capacity, lane, vbits);
      __syncthreads();
      // validity: whole words inside the tile's range, the boundary words merged
      int64_t end = base + cnt;
      if (end > fa.capacity) end = fa.capacity;
      if (end > base) {
        const int64_t w0 = base >> 5, w1 = (end - 1) >> 5;
        for (int o = 0; o < n_out; ++o) {
          uint32_t* words = plan.out[o].valid;
          if (!words) continue;
          for (int64_t g = w0 + lane; g <= w1; g += 64) {
            uint32_t own = 0xffffffffu;
            if (g == w0) own &= 0xffffffffu << (base & 31);
            if (g == w1 && (end & 31)) own &= 0xffffffffu >> (32 - (end & 31));
            merge_word(words, g, own, vbits[o * FUSED_VWORDS + (int)(g - w0)] & own);
          }
        }
      }
    }
    __syncthreads();  // LDS is rewritten by the next tile
  }
}

// Row selection by ops [0, split) (KILL per conjunct with a wave-uniform
// early exit, or one SEL), chunk by chunk over the wave's slice.
template <int NREG, int R, int S, int B>
__global__ __launch_bounds__(B) void select_vm(const DevPlan* __restrict__ pplan, SelArgs sa) {
  typedef typename RegVecT<NREG>::type RegVec;
  typedef SelectTile<R, S> T;
  constexpr int NW = B / 64;
  const DevPlan& plan = *pplan;
  const int lane = threadIdx.x & 63;
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plan.n_spill);
  const int64_t s0 = (int64_t)blockIdx.x * sa.per;
  const int64_t s1 = s0 + sa.per < sa.n_slices ? s0 + sa.per : sa.n_slices;
  uint32_t wave_total = 0;
  const int split = plan.split;
  for (int64_t slice = s0 + (threadIdx.x >> 6); slice < s1; slice += NW) {
    RegVec f[R];
#pragma unroll
    for (int k = 0; k < R; ++k) f[k] = (RegVec)(0);
    uint64_t vb = 0;
    LoadTemps<R> L;
    uint32_t alive_all = 0;
#pragma unroll 1
    for (int k = 0; k < S; ++k) {
      Rows<R> rs;
      chunk_rows<R>(rs, slice * T::WAVE_ROWS + k * T::CHUNK + lane * R, plan.num_rows);
      uint32_t alive = rs.exist;
      if (__ballot(alive != 0) != 0) {
        DevOp next = load_op(plan, 0);
        for (int q = 0; q < split; ++q) {
          const DevOp op = next;
          next = load_op(plan, q + 1 < split ? q + 1 : q);
          if (vm_step<NREG, R>(plan, op, f, vb, L, rs, alive, sp)) break;
        }
      }
      alive_all |= alive << (k * R);
    }
    wave_total += emit_slice<R, S>(sa, slice, alive_all);
  }
  emit_block<B>(sa, wave_total);
}

// ------------------------------------------------------- string outputs
//
// A string output is written by the VM as one slice per row (start byte |
// length << 32, into the chars of its source column). Two kernels turn the
// slices into an Arrow string column: offsets (exclusive scan of the lengths
// of valid rows, decoupled look-back across 4096-row tiles, offsets[count] =
// total), then the chars. Null rows get length 0.
constexpr int STR_R = 16;
constexpr int STR_TILE = BLOCK * STR_R;

__device__ __forceinline__ bool row_valid(const uint32_t* valid, int64_t row) {
  return !valid || ((gp<uint32_t>(valid)[row >> 5] >> (row & 31)) & 1u);
}

__global__ __launch_bounds__(BLOCK) void str_offsets_kernel(const uint64_t* __restrict__ slices,
                                                            const uint32_t* __restrict__ valid,
                                                            int32_t* __restrict__ offsets,
                                                            const int64_t* __restrict__ d_count,
                                                            int64_t count_h, CompactArgs ca) {
  __shared__ uint32_t s_ticket;
  __shared__ int64_t s_wave_total[WAVES];
  __shared__ int64_t s_prefix;
  if (threadIdx.x == 0) s_ticket = atomicAdd(ca.ticket, 1u);
  __syncthreads();
  const int64_t tile = s_ticket;
  int64_t count = count_h;
  if (d_count) {
    count = *d_count;
    if (count > ca.capacity) count = ca.capacity;
  }
  const int64_t tiles = count > 0 ? (count + STR_TILE - 1) / STR_TILE : 1;
  if (tile >= tiles) return;  // rows beyond the count: nothing to do
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = tile * STR_TILE + (int64_t)tid * STR_R;
  const GLOBAL uint64_t* sl = gp<uint64_t>(slices);
  uint32_t len[STR_R];
  int64_t mine = 0;
#pragma unroll
  for (int k = 0; k < STR_R; ++k) {
    const int64_t row = r0 + k;
    len[k] = (row < count && row_valid(valid, row)) ? (uint32_t)(sl[row] >> 32) : 0u;
    mine += len[k];
  }
  // inclusive scan of the lane sums over the wave
  int64_t incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  if (lane == 63) s_wave_total[wave] = incl;
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
    if (lane == 0) s_prefix = excl;
  }
  __syncthreads();
  int64_t pos = s_prefix + wave_off + incl - mine;
  GLOBAL int32_t* off = gpw<int32_t>(offsets);
#pragma unroll
  for (int k = 0; k < STR_R; ++k) {
    const int64_t row = r0 + k;
    if (row < count) off[row] = (int32_t)pos;
    pos += len[k];
  }
  if (tile == tiles - 1 && tid == BLOCK - 1) off[count] = (int32_t)(s_prefix + tile_total);
}

__global__ __launch_bounds__(BLOCK) void str_copy_kernel(const uint64_t* __restrict__ slices,
                                                         const uint32_t* __restrict__ valid,
                                                         const int32_t* __restrict__ offsets,
                                                         const uint8_t* __restrict__ src,
                                                         uint8_t* __restrict__ dst,
                                                         const int64_t* __restrict__ d_count,
                                                         int64_t count_h, int64_t capacity) {
  int64_t count = count_h;
  if (d_count) {
    count = *d_count;
    if (count > capacity) count = capacity;
  }
  const GLOBAL uint64_t* sl = gp<uint64_t>(slices);
  const GLOBAL int32_t* off = gp<int32_t>(offsets);
  const GLOBAL uint8_t* s = gp<uint8_t>(src);
  GLOBAL uint8_t* d = gpw<uint8_t>(dst);
  for (int64_t row = (int64_t)blockIdx.x * BLOCK + threadIdx.x; row < count;
       row += (int64_t)gridDim.x * BLOCK) {
    if (!row_valid(valid, row)) continue;
    const uint64_t v = sl[row];
    const uint32_t start = (uint32_t)v, n = (uint32_t)(v >> 32);
    const int64_t o = off[row];
    for (uint32_t i = 0; i < n; ++i) d[o + i] = s[start + i];
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

static int g_num_cus = 0;

// Workgroups of a persistent launch: enough to fill every CU several times.
static int64_t persistent_grid(int64_t tiles) {
  if (g_num_cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      g_num_cus = n;
    else
      g_num_cus = 256;
  }
  const int64_t cap = (int64_t)g_num_cus * 8;
  return tiles < cap ? tiles : cap;
}

// (NREG, R, B) variants compiled: register file size x rows per lane x
// workgroup size. 64-lane workgroups serve programs whose spill slots need
// more LDS than a 256-lane workgroup can have.
#define BSQL_VARIANTS(X) \
  X(8, 2, 256) X(8, 4, 256) X(16, 1, 256) X(16, 2, 256) X(16, 1, 64) X(16, 2, 64) X(8, 4, 64)

// LDS a workgroup may have for spill slots (the static arrays of the
// selection kernels take < 1 KB of the 64 KB).
static int64_t max_spill_lds() { return 60 * 1024; }

// Picks the variant: the smallest register file that holds the program, and
// the rows per lane requested (0 = the default for that register file), for
// workgroup size `block`; false if no variant holds it (or its spill slots).
static bool pick(int nreg, int rpt_req, int block, int n_spill, int* NR, int* RR) {
  struct V { int n, r; };
  const V vs[] = {{8, 4}, {16, 2}, {8, 2}, {16, 1}};
  for (const V& v : vs) {
    if (v.n < nreg) continue;
    if (rpt_req > 0 && v.r != rpt_req) continue;
    bool compiled = false;
#define X(N, RPT, BL) if (N == v.n && RPT == v.r && BL == block) compiled = true;
    BSQL_VARIANTS(X)
#undef X
    if (!compiled) continue;
    if (spill_lds_bytes(n_spill, block, v.r) > max_spill_lds()) continue;
    *NR = v.n;
    *RR = v.r;
    return true;
  }
  if (rpt_req > 0) return pick(nreg, 0, block, n_spill, NR, RR);
  return false;
}

bool variant_fits(int nreg, int rpt, int block, int n_spill) {
  int NR, RR;
  if (!pick(nreg, rpt, block, n_spill, &NR, &RR)) return false;
  return rpt == 0 || RR == rpt;
}

// Rows per lane of the variant a launch with this shape uses (0: none).
int variant_rows(int nreg, int rpt, int block, int n_spill) {
  int NR, RR;
  return pick(nreg, rpt, block, n_spill, &NR, &RR) ? RR : 0;
}

template <int NREG, int R, int B>
static hipError_t launch_execute_t(const DevPlan* d_plan, int64_t num_rows, int n_spill, hipStream_t stream) {
  const int64_t tile = (int64_t)B * R;
  const int64_t grid = persistent_grid((num_rows + tile - 1) / tile);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((vm_rows<NREG, R, false, B>), dim3((unsigned)grid), dim3(B),
                     (unsigned)spill_lds_bytes(n_spill, B, R), stream, d_plan, (const uint32_t*)nullptr,
                     (const int64_t*)nullptr, num_rows);
  return hipGetLastError();
}

// chunks per wave of the selection tiles: 4096-row tiles for the VM,
// 8192-row tiles for range tests
template <int R> struct VmChunks { static constexpr int S = 16 / R; };
// range selection: 4 rows per chunk; 2, 4 or 8 chunks per wave (8, 16 or 32
// rows per lane: sel_rpt 8, 16, 32), 4 by default
constexpr int RANGE_R = 4;
static int range_chunks_of(int sel_rpt) { return sel_rpt == 8 ? 2 : (sel_rpt == 32 ? 8 : 4); }

template <int NREG, int R, int B>
static hipError_t launch_select_vm_t(const DevPlan* d_plan, const SelArgs& sa, int64_t grid, int n_spill,
                                     hipStream_t stream) {
  hipLaunchKernelGGL((select_vm<NREG, R, VmChunks<R>::S, B>), dim3((unsigned)grid), dim3(B),
                     (unsigned)spill_lds_bytes(n_spill, B, R), stream, d_plan, sa);
  return hipGetLastError();
}

template <int R, int S>
static hipError_t launch_slice_ids(const SelArgs& sa, uint32_t* idx, int64_t* d_count, int64_t grid,
                                   hipStream_t stream) {
  hipLaunchKernelGGL((slice_ids<R, S>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, sa, idx, d_count);
  return hipGetLastError();
}

template <int NREG, int R, int B>
static hipError_t launch_gather_t(const DevPlan* d_plan, int64_t num_rows, const uint32_t* idx,
                                  const int64_t* d_count, int64_t capacity, int n_spill, hipStream_t stream) {
  const int64_t tile = (int64_t)B * R;
  const int64_t grid = persistent_grid((num_rows + tile - 1) / tile);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((vm_rows<NREG, R, true, B>), dim3((unsigned)grid), dim3(B),
                     (unsigned)spill_lds_bytes(n_spill, B, R), stream, d_plan, idx, d_count, capacity);
  return hipGetLastError();
}

// `d_plan` is a device copy of `plan` (made by the caller on `stream`).
hipError_t launch_execute(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt, int block,
                          hipStream_t stream) {
  int NR, RR;
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) return launch_execute_t<N, RPT, BL>(d_plan, plan.num_rows, plan.n_spill, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Rows per selection slice and ballots per slice (R * S) of the kernel a
// plan selects with.
static bool slice_shape(const DevPlan& plan, int nreg, int rpt, int block, int64_t* rows, int* ballots) {
  if (plan.n_ranges > 0) {
    const int chunks = range_chunks_of(rpt);
    *rows = (int64_t)64 * RANGE_R * chunks;
    *ballots = RANGE_R * chunks;
    return true;
  }
  int NR, RR;
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return false;
  *rows = 64 * 16;  // R * VmChunks<R>::S == 16 rows per lane
  *ballots = 16;
  return true;
}

// Workgroups of `kernel` that are resident on the device at once (its
// occupancy per CU x CUs), or 8 per CU when the runtime cannot say.
static int64_t resident_workgroups(const void* kernel, int block, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int64_t> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(std::make_tuple(dev, kernel, block, lds));
  if (it != cache.end()) return it->second;
  (void)persistent_grid(1);  // learns the CU count
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess || per_cu <= 0)
    per_cu = 8;
  const int64_t r = (int64_t)per_cu * g_num_cus;
  cache.emplace(std::make_tuple(dev, kernel, block, lds), r);
  return r;
}

// Workgroups of the selection launch (and of its slice_ids): no more than are
// resident at once, so that no second round of workgroups trails the first
// (each takes a contiguous block of slices, ceil(slices / grid) of them).
static int64_t selection_grid(const DevPlan& plan, int nreg, int sel_rpt, int block, int64_t n_slices) {
  const void* k = nullptr;
  int wg = BLOCK;
  size_t lds = 0;
  if (plan.n_ranges > 0) {
    switch (range_chunks_of(sel_rpt)) {
      case 2: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 2>); break;
      case 8: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 8>); break;
      default: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 4>);
    }
  } else {
    int NR, RR;
    if (pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR)) {
      wg = block;
      lds = (size_t)spill_lds_bytes(plan.n_spill, block, RR);
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) k = reinterpret_cast<const void*>(&select_vm<N, RPT, VmChunks<RPT>::S, BL>);
      BSQL_VARIANTS(X)
#undef X
    }
  }
  const int waves = wg / 64;
  int64_t grid = persistent_grid((n_slices + waves - 1) / waves);
  if (k) {
    const int64_t r = resident_workgroups(k, wg, lds);
    if (grid > r) grid = r;
  }
  return grid;
}

int64_t select_scratch_words(const DevPlan& plan, int nreg, int rpt, int block) {
  int64_t rows;
  int ballots;
  if (!slice_shape(plan, nreg, rpt, block, &rows, &ballots)) return -1;
  const int64_t n = (plan.num_rows + rows - 1) / rows;
  const int64_t grid = selection_grid(plan, nreg, rpt, block, n);
  return n * ballots + (n + 2 * grid + 1) / 2;
}

// The output phase of a filter: ops [split, n_ops) over the selected row ids
// (written by the selection phase, count in *d_count).
static hipError_t launch_compact_gather(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt, int block,
                                        const uint32_t* idx, const int64_t* d_count, int64_t capacity,
                                        hipStream_t stream) {
  if (plan.split >= plan.n_ops) return hipSuccess;
  int NR, RR;
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL)                                                                                         \
  if (NR == N && RR == RPT && block == BL)                                                                    \
    return launch_gather_t<N, RPT, BL>(d_plan, plan.num_rows < capacity ? plan.num_rows : capacity, idx,      \
                                       d_count, capacity, plan.n_spill, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Filter (+ project): selection kernel -> ids of the selected rows in `idx`
// (num_rows entries) and their count in *d_count, then the output ops over
// those rows. `sel` holds select_scratch_words(...) words. `phases`: bit 0 =
// selection (+ scan + ids), bit 1 = output phase (a caller that sizes its
// outputs from the count runs them as two calls on one stream, with idx and
// *d_count kept in between). Nothing is cleared first: every slice writes
// its ballots and count, and the scan writes *d_count.
// Whether a filter (+ project) plan runs as the fused one-pass kernel: a
// range-test predicate, output ops that fit the 8-register file without
// spill slots, at most FUSED_MAX_OUT outputs.
bool fused_eligible(const DevPlan& plan, int nreg) {
  return plan.n_ranges > 0 && plan.split < plan.n_ops && plan.n_out <= FUSED_MAX_OUT && nreg <= 8 &&
         plan.n_spill == 0;
}

// Words of scratch the fused kernel uses: one status word per tile, the
// ticket counters.
int64_t fused_scratch_words(int64_t num_rows) { return (num_rows + FUSED_TILE - 1) / FUSED_TILE + 4; }

#define BSQL_FUSED_NREG 8
#define BSQL_FUSED_R 2
#define BSQL_FUSED_SH 4

static hipError_t launch_fused(const DevPlan& plan, const DevPlan* d_plan, uint64_t* words, int64_t* d_count,
                               int64_t capacity, hipStream_t stream) {
  const int64_t tiles = (plan.num_rows + FUSED_TILE - 1) / FUSED_TILE;
  const int64_t n_words = tiles + (FUSED_COUNTERS + 1) / 2;
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, stream, words, n_words, 0ULL);
  if (tiles == 0) {
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, stream, (uint64_t*)d_count, (int64_t)1, 0ULL);
    return hipGetLastError();
  }
  FusedArgs fa{words, (uint32_t*)(words + tiles), d_count, capacity, tiles};
  const size_t lds = (size_t)fused_lds_bytes(plan.n_out);
  const void* k = reinterpret_cast<const void*>(
      &select_project_fused<BSQL_FUSED_NREG, BSQL_FUSED_R, BSQL_FUSED_SH>);
  int64_t grid = resident_workgroups(k, 64, lds);
  if (grid > tiles) grid = tiles;
  // every ticket counter has workgroups
  grid = (grid + FUSED_COUNTERS - 1) / FUSED_COUNTERS * FUSED_COUNTERS;
  hipLaunchKernelGGL((select_project_fused<BSQL_FUSED_NREG, BSQL_FUSED_R, BSQL_FUSED_SH>), dim3((unsigned)grid),
                     dim3(64), (unsigned)lds, stream, d_plan, fa);
  return hipGetLastError();
}

hipError_t launch_compact(const DevPlan& plan, const DevPlan* d_plan, int nreg, int sel_rpt, int rpt, int block,
                          uint64_t* sel, uint32_t* idx, int64_t* d_count, int64_t capacity,
                          hipStream_t stream, int phases, bool fused) {
  if (fused && phases == 3) return launch_fused(plan, d_plan, sel, d_count, capacity, stream);
  if (!(phases & 1))
    return launch_compact_gather(plan, d_plan, nreg, rpt, block, idx, d_count, capacity, stream);
  int64_t rows;
  int ballots;
  if (!slice_shape(plan, nreg, sel_rpt, block, &rows, &ballots)) return hipErrorInvalidValue;
  const int64_t n = (plan.num_rows + rows - 1) / rows;
  const int64_t grid = selection_grid(plan, nreg, sel_rpt, block, n);
  uint32_t* counts = (uint32_t*)(sel + n * ballots);
  SelArgs sa{sel, counts, counts + n, counts + n + grid, n, grid ? (n + grid - 1) / grid : 0};
  hipError_t e = hipSuccess;
  const bool ids = plan.split < plan.n_ops;
  if (n > 0) {
    e = hipErrorInvalidValue;
    if (plan.n_ranges > 0) {
      switch (range_chunks_of(sel_rpt)) {
        case 2:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 2>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
          break;
        case 8:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 8>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
          break;
        default:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 4>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
      }
      e = hipGetLastError();
    } else {
      int NR, RR;
      if (!pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) e = launch_select_vm_t<N, RPT, BL>(d_plan, sa, grid, plan.n_spill, stream);
      BSQL_VARIANTS(X)
#undef X
    }
    if (e != hipSuccess) return e;
  }
  if (ids && n > 0) {
    // the ids kernel also sums the block totals and writes the row total
    if (plan.n_ranges > 0) {
      const int chunks = range_chunks_of(sel_rpt);
      e = chunks == 2 ? launch_slice_ids<RANGE_R, 2>(sa, idx, d_count, grid, stream)
          : chunks == 8 ? launch_slice_ids<RANGE_R, 8>(sa, idx, d_count, grid, stream)
                        : launch_slice_ids<RANGE_R, 4>(sa, idx, d_count, grid, stream);
    } else {
      int NR, RR;
      if (!pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
      e = RR == 1 ? launch_slice_ids<1, VmChunks<1>::S>(sa, idx, d_count, grid, stream)
          : RR == 2 ? launch_slice_ids<2, VmChunks<2>::S>(sa, idx, d_count, grid, stream)
                    : launch_slice_ids<4, VmChunks<4>::S>(sa, idx, d_count, grid, stream);
    }
  } else {
    hipLaunchKernelGGL(block_scan, dim3(1), dim3(SCAN_THREADS), 0, stream, (const uint32_t*)sa.block_sums,
                       sa.block_base, n > 0 ? grid : (int64_t)0, d_count);
    e = hipGetLastError();
  }
  if (e != hipSuccess || !(phases & 2)) return e;
  return launch_compact_gather(plan, d_plan, nreg, rpt, block, idx, d_count, capacity, stream);
}

// Batched launch: `plans` (device) holds n_batches plans, `tile_start`
// (device) n_batches + 1 prefix counts of tiles of block*rpt rows. No
// allocation and no copy: the launch can be captured in a hipGraph.
int batch_tile_rows(int nreg, int rpt, int block) {
  int NR, RR;
  if (!pick(nreg, rpt, block, 0, &NR, &RR)) return -1;
  return block * RR;
}

hipError_t launch_batches(const DevPlan* plans, const int64_t* tile_start, int n_batches, int64_t n_items,
                          int nreg, int rpt, int block, int n_spill, hipStream_t stream) {
  int NR, RR;
  if (!pick(nreg, rpt, block, n_spill, &NR, &RR)) return hipErrorInvalidValue;
  const int64_t grid = persistent_grid(n_items);
  if (grid == 0) return hipSuccess;
#define X(N, RPT, BL)                                                                                  \
  if (NR == N && RR == RPT && block == BL) {                                                           \
    hipLaunchKernelGGL((vm_batches<N, RPT, BL>), dim3((unsigned)grid), dim3(BL),                       \
                       (unsigned)spill_lds_bytes(n_spill, BL, RPT), stream, plans, tile_start, n_batches); \
    return hipGetLastError();                                                                          \
  }
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Assembles string output columns from the slices the VM stored: offsets
// then chars. `status` holds string_tiles(capacity) + 1 words; `d_count`
// (device) or `count` (host) gives the rows.
int64_t string_tiles(int64_t capacity) { return capacity > 0 ? (capacity + STR_TILE - 1) / STR_TILE : 1; }

static void launch_offsets(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                           const int64_t* d_count, int64_t count, int64_t capacity, uint64_t* status,
                           hipStream_t stream) {
  const int64_t tiles = string_tiles(capacity);
  CompactArgs ca{status, (uint32_t*)(status + tiles), nullptr, capacity, tiles};
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((tiles + 1 + 255) / 256)), dim3(256), 0, stream, status,
                     tiles + 1, 0ULL);
  hipLaunchKernelGGL(str_offsets_kernel, dim3((unsigned)tiles), dim3(BLOCK), 0, stream, slices, valid, offsets,
                     d_count, count, ca);
}

// Offsets (count + 1 of them) of a string column from row lengths stored as
// length << 32 in `slices` (null rows: 0). `status`: string_tiles(count) + 1 words.
hipError_t launch_string_offsets(const uint64_t* slices, const uint32_t* valid, int32_t* offsets, int64_t count,
                                 uint64_t* status, hipStream_t stream) {
  launch_offsets(slices, valid, offsets, nullptr, count, count, status, stream);
  return hipGetLastError();
}

hipError_t launch_string_output(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                                const uint8_t* src_chars, uint8_t* dst_chars, const int64_t* d_count,
                                int64_t count, int64_t capacity, uint64_t* status, hipStream_t stream) {
  launch_offsets(slices, valid, offsets, d_count, count, capacity, status, stream);
  const int64_t blocks = persistent_grid((capacity + BLOCK - 1) / BLOCK);
  if (blocks > 0)
    hipLaunchKernelGGL(str_copy_kernel, dim3((unsigned)blocks), dim3(BLOCK), 0, stream, slices, valid, offsets,
                       src_chars, dst_chars, d_count, count, capacity);
  return hipGetLastError();
}

}  // namespace bsql
