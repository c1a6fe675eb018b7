// vm_device.h -- the device code of the row-parallel expression interpreter
// and its filter kernels: column I/O, the VM core (op semantics), the
// register-file step, range tests, selection, compaction and the kernel
// templates. interp_kernels.hip instantiates the kernels for the
// interpreter (ops, column types and ranges read from the launch plan);
// kernels generated for one program (kernel_gen.cpp) include this same text
// through hipRTC and instantiate the templates with a source of constant op
// fields and column types, so both run the same op semantics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// YEAR, MONTH or DAY of a day number (days since 1970-01-01), Hinnant's
// civil_from_days; inline, so that EXTRACT over a DATE column costs no call
__device__ __forceinline__ int64_t extract_ymd(int op, int64_t days) {
  const int64_t z = days + 719468;
  int64_t era;
  unsigned doe;
  if (z == (int64_t)(int32_t)z) {  // every int32 day number: 32-bit divide by the constant
    const int32_t z32 = (int32_t)z;
    const int32_t e32 = (z32 >= 0 ? z32 : z32 - 146096) / 146097;
    era = e32;
    doe = (unsigned)(z32 - e32 * 146097);
  } else {
    era = (z >= 0 ? z : z - 146096) / 146097;
    doe = (unsigned)(z - era * 146097);
  }
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

// EXTRACT (interpreter_ops.cuh:69-99) with Hinnant's civil_from_days
static __device__ int64_t extract(int op, int t, int64_t ticks) {
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
  return extract_ymd(op, days);
}

// Integer division and modulo with the interpreter's edge cases; operands
// that both fit in 32 bits (the usual case) take the 32-bit divide, whose
// truncating quotient and remainder equal the 64-bit ones there.
__device__ __forceinline__ bool fits_i32(int64_t v) { return v == (int64_t)(int32_t)v; }
__device__ __forceinline__ int64_t idiv(int64_t a, int64_t b) {
  if (b == 0) return 0;
  if (b == -1) return (int64_t)(0ULL - (uint64_t)a);
  if (fits_i32(a) && fits_i32(b)) return (int64_t)((int32_t)a / (int32_t)b);
  return a / b;
}
__device__ __forceinline__ int64_t imod(int64_t a, int64_t b) {
  if (b == 0 || b == -1) return 0;
  if (fits_i32(a) && fits_i32(b)) return (int64_t)((int32_t)a % (int32_t)b);
  return a % b;
}

// Philox4x32-10: uniform double in (0, 1] per (seed, row)
static __device__ double philox_uniform(uint64_t seed, uint64_t row) {
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
static __device__ int str_cmp(const GLOBAL uint8_t* a, int32_t na, const GLOBAL uint8_t* b, int32_t nb) {
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
template <int R, bool NT = true>
__device__ __forceinline__ void load_rows(const DevInCol& c, int64_t r0, int n, uint64_t (&v)[R]) {
  const int t = c.type;
  const bool vec = (n == R) && c.aligned16;
#define BSQL_LOAD_CASE(TYPE_ID, CT, EXPR)                                          \
  case TYPE_ID: {                                                                  \
    const GLOBAL CT* p = gp<CT>(c.data) + r0;                                      \
    CT tmp[R];                                                                     \
    if (vec && (sizeof(CT) * R) % 16 == 0) {                                       \
      _Pragma("unroll") for (int k = 0; k < (int)(sizeof(CT) * R / 16); ++k)       \
        ((v4u32*)tmp)[k] = NT ? __builtin_nontemporal_load((const GLOBAL v4u32*)p + k) \
                              : ((const GLOBAL v4u32*)p)[k];                         \
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
          ((v4u32*)tmp)[k] = NT ? __builtin_nontemporal_load((const GLOBAL v4u32*)p + k) : ((const GLOBAL v4u32*)p)[k];
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
static __device__ __noinline__ uint64_t compute_heavy(uint8_t code, uint8_t aux, uint8_t aux2, int64_t imm,
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
// (four bytes at a time from the aligned 32-bit words that cover s[0, n),
// the bytes outside it masked off as in utf8_chars)
static __device__ int32_t utf8_offset(const GLOBAL uint8_t* s, int32_t n, int64_t chars) {
  if (n <= 0 || chars < 0) return n;
  const uintptr_t a = (uintptr_t)s;
  const GLOBAL uint32_t* w = (const GLOBAL uint32_t*)(a & ~(uintptr_t)3);
  const int32_t lead = (int32_t)(a & 3);
  const int32_t end = lead + n;
  int64_t c = 0;
  for (int32_t off = 0; off < end; off += 4) {
    const uint32_t x = w[off >> 2];
    uint32_t m = ~(x & ~(x << 1)) & 0x80808080u;  // bit 7 of every lead byte
    if (off == 0) m &= 0xFFFFFFFFu << (8 * lead);
    if (end - off < 4) m &= (1u << (8 * (end - off))) - 1u;
    const int cnt = __popc(m);
    if (c + cnt > chars) {
      // the (chars - c)-th lead byte of this word
      for (int64_t r = chars - c; r > 0; --r) m &= m - 1u;
      return off + (__ffs(m) - 1) / 8 - lead;
    }
    c += cnt;
  }
  return n;
}
// Characters of the UTF-8 bytes s[0, n): the bytes that are not continuation
// bytes (10xxxxxx), counted four at a time from the aligned 32-bit words that
// cover them, the bytes of those words outside [s, s + n) masked off (an
// aligned word lies in the page of the byte it was read for). The words are
// read four per step, their loads issued together (only words that cover a
// byte of s are read).
static __device__ int64_t utf8_chars(const GLOBAL uint8_t* s, int32_t n) {
  if (n <= 0) return 0;
  const uintptr_t a = (uintptr_t)s;
  const GLOBAL uint32_t* w = (const GLOBAL uint32_t*)(a & ~(uintptr_t)3);
  const int32_t lead = (int32_t)(a & 3);
  const int32_t end = lead + n;  // bytes from w[0]
  int32_t cont = 0;
  for (int32_t off = 0; off < end; off += 16) {
    uint32_t x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = off + 4 * j < end ? w[(off >> 2) + j] : 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int32_t o = off + 4 * j;
      uint32_t m = x[j] & ~(x[j] << 1) & 0x80808080u;  // bit 7 of every continuation byte
      if (o == 0) m &= 0xFFFFFFFFu << (8 * lead);
      if (end - o < 4) m &= end - o > 0 ? (1u << (8 * (end - o))) - 1u : 0u;
      cont += __popc(m);
    }
  }
  return (int64_t)n - cont;
}

// utf8_chars of a lane's R strings (n[k] <= 0: none, 0 characters) at once:
// every string's next four words are read in one step, the loads of all R
// strings issued together, so the lane takes as many steps as its longest
// string needs rather than their sum.
template <int R>
__device__ __forceinline__ void utf8_chars_rows(const GLOBAL uint8_t* const (&s)[R], const int32_t (&n)[R],
                                                int64_t (&chars)[R]) {
  const GLOBAL uint32_t* w[R];
  int32_t lead[R], end[R], cont[R];
  int32_t most = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uintptr_t a = (uintptr_t)s[k];
    w[k] = (const GLOBAL uint32_t*)(a & ~(uintptr_t)3);
    lead[k] = (int32_t)(a & 3);
    end[k] = n[k] > 0 ? lead[k] + n[k] : 0;  // bytes from w[k][0]; 0: nothing read
    cont[k] = 0;
    most = end[k] > most ? end[k] : most;
  }
  for (int32_t off = 0; off < most; off += 16) {
    uint32_t x[R][4];
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) x[k][j] = off + 4 * j < end[k] ? w[k][(off >> 2) + j] : 0u;
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int32_t o = off + 4 * j;
        uint32_t m = x[k][j] & ~(x[k][j] << 1) & 0x80808080u;  // bit 7 of every continuation byte
        if (o == 0) m &= 0xFFFFFFFFu << (8 * lead[k]);
        if (end[k] - o < 4) m &= end[k] - o > 0 ? (1u << (8 * (end[k] - o))) - 1u : 0u;
        cont[k] += __popc(m);
      }
  }
#pragma unroll
  for (int k = 0; k < R; ++k) chars[k] = n[k] > 0 ? (int64_t)n[k] - cont[k] : 0;
}

// The UTF-8 character (its encoded bytes packed little-endian) at s[i].
__device__ __forceinline__ uint32_t utf8_char_at(const GLOBAL uint8_t* s, int32_t i, int32_t len) {
  uint32_t c = 0;
  for (int32_t k = 0; k < len; ++k) c |= (uint32_t)s[i + k] << (8 * k);
  return c;
}

// Whether character c is one to strip: one of the characters of `set`, or
// for an empty set a byte <= ' ' (cudf::strings::strip's rule).
static __device__ bool strip_char(uint32_t c, const GLOBAL uint8_t* set, int32_t nset) {
  if (nset == 0) return c <= (uint32_t)' ';
  for (int32_t i = 0; i < nset;) {
    const int32_t l = utf8_step(set, i, nset);  // >= 1
    if (utf8_char_at(set, i, l) == c) return true;
    i += l;
  }
  return false;
}

// [b0, b1) of s[0, n) after stripping (side bit 0: leading, bit 1: trailing).
static __device__ void strip_range(const GLOBAL uint8_t* s, int32_t n, const GLOBAL uint8_t* set, int32_t nset, int side,
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
static __device__ int64_t str_to_i64(const GLOBAL uint8_t* s, int32_t n) {
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
static __device__ double str_to_f64(const GLOBAL uint8_t* s, int32_t n) {
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
static __device__ int64_t days_from_ymd(int64_t y, int64_t m, int64_t d) {
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
static __device__ int64_t str_to_ticks(const GLOBAL uint8_t* s, int32_t n, const int32_t* fmt, int type) {
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
static __device__ bool regex_char_in_class(const GLOBAL int32_t* prog, int32_t at, uint32_t cp) {
  const int32_t nr = prog[at];
  bool in = false;
  for (int32_t k = 0; k < nr && !in; ++k)
    in = cp >= (uint32_t)prog[at + 2 + 2 * k] && cp <= (uint32_t)prog[at + 3 + 2 * k];
  return prog[at + 1] ? !in : in;
}

static __device__ bool regex_contains(const GLOBAL uint8_t* s, int32_t n, const GLOBAL int32_t* prog) {
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
static __device__ bool like_from(const GLOBAL uint8_t* s, int32_t n, int32_t st, const GLOBAL int32_t* u, int32_t nu,
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
static __device__ bool like_match(const GLOBAL uint8_t* s, int32_t n, const GLOBAL int32_t* pat) {
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
#define AF ((op.flags & F_A_CVT) ? (double)(int64_t)a : as_f(a))
#define BF ((op.flags & F_B_CVT) ? (double)(int64_t)b : as_f(b))
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
    case D_MAGIC_IF_NULL: ROWS(va ? a : (FA ? bits(MAGIC_F64) : (uint64_t)MAGIC_I64), va)
    case D_COALESCE: {  // FIRST_NON_MAGIC(MAGIC_IF_NULL(a), b)
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t va = (VA >> k) & 1u, vb = (VB >> k) & 1u;
        const uint64_t m = va ? A[k] : (FA ? bits(MAGIC_F64) : (uint64_t)MAGIC_I64);
        const bool magic = FA ? as_f(m) == MAGIC_F64 : (int64_t)m == MAGIC_I64;
        D[k] = magic ? B[k] : m;
        VD |= (magic ? vb : va) << k;
      }
      return;
    }
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
    case D_EXTRACT:
      if (op.aux2 == BSQL_TIMESTAMP_DAYS &&
          (op.aux == BSQL_OP_YEAR || op.aux == BSQL_OP_MONTH || op.aux == BSQL_OP_DAY)) {
        ROWS(extract_ymd(op.aux, (int64_t)a), va)
      }
      ROWS(compute_heavy(op.code, op.aux, op.aux2, op.imm, AF, BF, a, rs.row[k]), va)
    case D_SIN: case D_COS: case D_ASIN: case D_ACOS: case D_TAN: case D_COTAN: case D_ATAN:
    case D_LN: case D_LOG:
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
          // = and <> of strings of different lengths need no bytes
          D[k] = kind <= 1 && na != nb ? (uint64_t)(kind == 1)
                                       : (uint64_t)cmp_result(kind, str_cmp(pa, na, pb, nb));
          VD |= 1u << k;
        }
      }
      return;
    }
    case D_CHAR_LENGTH: {
      const GLOBAL uint8_t* p[R];
      int32_t n[R];
      int64_t c[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        p[k] = nullptr;
        n[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          str_at(plan, op.aux, A[k], &p[k], &n[k]);
          VD |= 1u << k;
        }
      }
      utf8_chars_rows<R>(p, n, c);
#pragma unroll
      for (int k = 0; k < R; ++k) D[k] = (uint64_t)c[k];
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
    case D_LIKE_PREFIX: {
      const int kb = (int)((uint64_t)op.imm >> 56);
      const uint64_t want = (uint64_t)op.imm & 0x00FFFFFFFFFFFFFFull;
      const bool whole = op.aux2 != 0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        D[k] = 0;
        if (((rs.exist >> k) & 1u) && ((VA >> k) & 1u)) {
          const GLOBAL uint8_t* p;
          int32_t n;
          str_at(plan, op.aux, A[k], &p, &n);
          bool m = whole ? n == kb : n >= kb;
          if (m) {
            uint64_t got = 0;
#pragma unroll
            for (int j = 0; j < 7; ++j)
              if (j < kb) got |= (uint64_t)p[j] << (8 * j);
            m = got == want;
          }
          D[k] = m ? 1 : 0;
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
static __device__ int64_t look_back(const CompactArgs& ca, int64_t tile, int64_t agg, int lane) {
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

// Exclusive prefix of tile `tile` > 0 whose aggregate `agg` is already
// published (wave 0 only): the look-back of look_back() without its first
// store; publishes the tile's inclusive prefix.
__device__ __forceinline__ int64_t look_back_published(uint64_t* status, int64_t tile, int64_t agg, int lane) {
  int64_t excl = 0;
  int64_t look = tile - 1;
  while (true) {
    const int64_t t = look - lane;
    const uint64_t s = t >= 0 ? ld_status(&status[t]) : ST_PREFIX;
    const uint64_t flag = s & ~ST_VALUE;
    const uint64_t not_ready = __ballot(flag == 0);
    const uint64_t is_p = __ballot(flag == ST_PREFIX);
    const uint64_t upto = is_p ? ((is_p & (~is_p + 1)) << 1) - 1 : ~0ULL;  // lanes [0, first prefix]
    if (not_ready & upto) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += wave_sum(((upto >> lane) & 1) ? (int64_t)(s & ST_VALUE) : 0);
    if (is_p) break;
    look -= 64;
  }
  if (lane == 0) st_status(&status[tile], ST_PREFIX | (uint64_t)(excl + agg));
  return excl;
}

// Exclusive prefix of the tile (wave 0 of its workgroup, or the wave that
// owns the query; the tile's aggregate is published): *s_base, and the row
// count once the last tile knows it.
__device__ __forceinline__ void tile_position(const FusedArgs& fa, int64_t tile, int total, int lane,
                                              int64_t* s_base) {
  const int64_t b = tile == 0 ? 0 : look_back_published(fa.status, tile, total, lane);
  if (lane == 0) {
    *s_base = b;
    if (tile == fa.num_tiles - 1) *gpw<int64_t>(fa.d_count) = b + total;
  }
}

// look_back with a 64*V-tile window per round trip: each lane reads V
// consecutive status words (their loads in flight together). For launches
// with many tiles in flight, so that one round trip usually reaches a
// published prefix.
template <int V>
__device__ int64_t look_back_wide(const CompactArgs& ca, int64_t tile, int64_t agg, int lane) {
  if (tile == 0) {
    if (lane == 0) st_status(&ca.status[0], ST_PREFIX | (uint64_t)agg);
    return 0;
  }
  if (lane == 0) st_status(&ca.status[tile], ST_AGG | (uint64_t)agg);
  int64_t excl = 0;
  int64_t look = tile - 1;
  while (true) {
    uint64_t w[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int64_t t = look - (lane * V + j);
      w[j] = t >= 0 ? ld_status(&ca.status[t]) : ST_PREFIX;
    }
    // this lane's entries, in look-back order, up to and including its first prefix
    bool has_p = false, not_ready = false;
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      if (has_p) continue;
      const uint64_t flag = w[j] & ~ST_VALUE;
      if (flag == 0) not_ready = true;
      sum += (int64_t)(w[j] & ST_VALUE);
      if (flag == ST_PREFIX) has_p = true;
    }
    const uint64_t p = __ballot(has_p);
    const uint64_t upto = p ? ((p & (~p + 1)) << 1) - 1 : ~0ULL;  // lanes [0, first lane with a prefix]
    if (__ballot(not_ready) & upto) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += wave_sum(((upto >> lane) & 1) ? sum : 0);
    if (p) break;
    look -= 64 * V;
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

// Where the device templates take a program's columns, outputs, ranges and
// ops from. PlanSrc reads all of them from the launch plan (the
// interpreter). A kernel generated for one program supplies the same
// structs with its op fields, column types and validity presence as
// constants (literals, column addresses and range bounds still come from the
// plan), so the dispatch on them folds away and the same op semantics run
// straight-line.
struct PlanSrc {
  // the op loop is the kernels' own (a closure over the indexed register
  // file would keep it out of VGPRs)
  static constexpr bool kPlanOps = true;
  // input column whose range-tested values the output ops reuse (-1: none)
  static constexpr int kKeepCol = -1;
  // range tests with all columns' loads in flight together (generated
  // programs, whose range kinds are constants)
  static constexpr bool kEager = false;
  static constexpr int kRanges = 0;
  // rows of a one-pass tile's outputs staged in LDS (0: none)
  static constexpr int kStageRows = 0;
  template <int r> struct RangeAt {};
  __device__ static __forceinline__ DevInCol in(const DevPlan& p, int j) { return p.in[j]; }
  __device__ static __forceinline__ DevOutCol out(const DevPlan& p, int o) { return p.out[o]; }
  __device__ static __forceinline__ DevRange range(const DevPlan& p, int r) { return p.ranges[r]; }
  __device__ static __forceinline__ int n_ranges(const DevPlan& p) { return p.n_ranges; }
  __device__ static __forceinline__ int split(const DevPlan& p) { return p.split; }
  __device__ static __forceinline__ int n_ops(const DevPlan& p) { return p.n_ops; }
  // whether loads of input column j are non-temporal (dflt: the caller's choice)
  __device__ static __forceinline__ bool nt(int j, bool dflt) { return dflt; }
  // f(op) for the ops [q0, q1) in order, until f returns true; the next op's
  // scalar loads are issued before f runs, so their latency hides behind it
  template <class F>
  __device__ static __forceinline__ void ops(const DevPlan& plan, int q0, int q1, F&& f) {
    if (q0 >= q1) return;
    DevOp next = load_op(plan, q0);
    for (int q = q0; q < q1; ++q) {
      const DevOp op = next;
      next = load_op(plan, q + 1 < q1 ? q + 1 : q);
      if (f(op)) return;
    }
  }
};

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
    load_rows<R, NT>(c, rs.row0, rs.nrows, v);
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

// The register file in LDS, for plans read at run time: row k of register
// i of lane t is slot (i*R + k)*64 + t of the wave's file (p points at the
// lane's slot 0), so a dynamic register number costs one address add, the
// R rows are R loads or stores at constant offsets, and the 64 lanes of one
// access touch 64 consecutive slots, which no two lanes share a bank in
// (rows of a lane side by side put lanes 8*R bytes apart: 2-way conflicts
// for R = 4). Each lane touches only its own slots: no barrier. The wave's
// file is NREG*64*R slots.
template <int NREG, int R>
struct LRegs {
  LDS uint64_t* p;
  __device__ __forceinline__ uint64_t at(int k, int i) const { return p[(i * R + k) * 64]; }
  __device__ __forceinline__ void get(int i, uint64_t (&v)[R]) const {
    const LDS uint64_t* q = p + i * (64 * R);
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = q[k * 64];
  }
  __device__ __forceinline__ void put(int i, const uint64_t (&v)[R]) {
    LDS uint64_t* q = p + i * (64 * R);
#pragma unroll
    for (int k = 0; k < R; ++k) q[k * 64] = v[k];
  }
};
template <class T> struct IsLRegs { static constexpr bool value = false; };
template <int NREG, int R> struct IsLRegs<LRegs<NREG, R>> { static constexpr bool value = true; };

// Executes one non-store op. Returns true if it was a KILL that left the
// wave without a selected row.
// Spill slot q of the calling wave: 64 lanes x R rows of data (row k of lane
// t at k*64 + t), then one validity byte per lane (R bits).
template <int R>
__device__ __forceinline__ LDS uint64_t* spill_slot(LDS uint64_t* sp, int64_t q) {
  return sp + q * (64 * R + 8);
}

// One op with the register file in LDS. Operands are read only by the ops
// that use them, and ops without a result write nothing. Column loads are
// not ops here: the tile loop runs a LOADN and its COMMITs as one step
// (lds_loads).
template <int NREG, int R, class P, class RF>
__device__ __forceinline__ bool vm_step_lds(const DevPlan& plan, const DevOp& op, RF& f, uint64_t& vb,
                                            const Rows<R>& rs, uint32_t& alive, LDS uint64_t* sp) {
  constexpr uint32_t RMASK = (1u << R) - 1u;
  const uint8_t code = op.code;
  uint64_t A[R], B[R], D[R];
  uint32_t VA = 0, VB = 0, VD = 0;
  const int dst = op.dst == 0xff ? NREG - 1 : (op.dst & (NREG - 1));
  // operands: both LDS reads issued before either is waited for
  auto operands = [&]() __attribute__((always_inline)) {
    const int ia = op.a & (NREG - 1), ib = op.b & (NREG - 1);
    f.get(ia, A);
    f.get(ib, B);
    VA = (uint32_t)(vb >> (ia * R)) & RMASK;
    VB = (uint32_t)(vb >> (ib * R)) & RMASK;
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
  };
  if (code >= D_ADD_II && code < D_SPILL) {  // the computing ops: one range test, then their switch
    operands();
    compute_rows<R>(plan, op, A, B, VA, VB, D, VD, rs);
  } else if (code == D_KILL || code == D_SEL) {
    operands();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const bool t = code == D_KILL ? truthy(A[k], FA) : (A[k] != 0);
      if (!(((VA >> k) & 1u) && t)) alive &= ~(1u << k);
    }
    return code == D_KILL && __ballot(alive != 0) == 0;
  } else if (code == D_LOADS) {
    const DevInCol c = P::in(plan, op.a);
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
  } else if (code == D_FILL) {
    const int lane = threadIdx.x & 63;
    const LDS uint64_t* slot = spill_slot<R>(sp, op.imm);
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = slot[k * 64 + lane];
    VD = ((const LDS uint8_t*)(slot + 64 * R))[lane];
  } else if (code == D_SPILL) {
    operands();
    const int lane = threadIdx.x & 63;
    LDS uint64_t* slot = spill_slot<R>(sp, op.imm);
#pragma unroll
    for (int k = 0; k < R; ++k) slot[k * 64 + lane] = A[k];
    ((LDS uint8_t*)(slot + 64 * R))[lane] = (uint8_t)VA;
    return false;
  } else {
    return false;  // LOADN and COMMIT are run by lds_loads
  }
  f.put(dst, D);
  vb = (vb & ~((uint64_t)RMASK << (dst * R))) | ((uint64_t)(VD & RMASK) << (dst * R));
  return false;
}

// A LOADN and the COMMITs that follow it (ops q+1 .. q+n) with the register
// file in LDS: the n columns' loads are issued together, then each lands in
// its COMMIT's register.
template <int NREG, int R, bool NT, class P, class RF>
__device__ __forceinline__ void lds_loads(const DevPlan& plan, const DevOp& op, int q, RF& f, uint64_t& vb,
                                          const Rows<R>& rs) {
  constexpr uint32_t RMASK = (1u << R) - 1u;
  const int n = op.aux;
  uint64_t v[4][R];
  uint32_t m[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j >= n) break;
    const int col = (int)((op.imm >> (8 * j)) & 0xff);
    if (P::nt(col, NT)) col_load<R, true>(P::in(plan, col), rs, v[j], m[j]);
    else col_load<R, false>(P::in(plan, col), rs, v[j], m[j]);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j >= n) break;
    const DevOp c = load_op(plan, q + 1 + j);
    const int dst = c.dst & (NREG - 1);
    f.put(dst, v[j]);
    vb = (vb & ~((uint64_t)RMASK << (dst * R))) | ((uint64_t)(m[j] & RMASK) << (dst * R));
  }
}

template <int NREG, int R, bool NT = true, class P = PlanSrc>
__device__ __forceinline__ bool vm_step(const DevPlan& plan, const DevOp& op, LRegs<NREG, R>& f, uint64_t& vb,
                                        LoadTemps<R>&, const Rows<R>& rs, uint32_t& alive, LDS uint64_t* sp) {
  return vm_step_lds<NREG, R, P>(plan, op, f, vb, rs, alive, sp);
}

template <int NREG, int R, bool NT = true, class P = PlanSrc>
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
    const int j0 = (int)(op.imm & 0xff), j1 = (int)((op.imm >> 8) & 0xff), j2 = (int)((op.imm >> 16) & 0xff),
              j3 = (int)((op.imm >> 24) & 0xff);
    if (P::nt(j0, NT)) col_load<R, true>(P::in(plan, j0), rs, T.t0, T.v0);
    else col_load<R, false>(P::in(plan, j0), rs, T.t0, T.v0);
    if (n > 1) {
      if (P::nt(j1, NT)) col_load<R, true>(P::in(plan, j1), rs, T.t1, T.v1);
      else col_load<R, false>(P::in(plan, j1), rs, T.t1, T.v1);
    }
    if (n > 2) {
      if (P::nt(j2, NT)) col_load<R, true>(P::in(plan, j2), rs, T.t2, T.v2);
      else col_load<R, false>(P::in(plan, j2), rs, T.t2, T.v2);
    }
    if (n > 3) {
      if (P::nt(j3, NT)) col_load<R, true>(P::in(plan, j3), rs, T.t3, T.v3);
      else col_load<R, false>(P::in(plan, j3), rs, T.t3, T.v3);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = A[k];
    dst = NREG - 1;
  } else if (code == D_COMMIT) {
    const int j = op.aux;
#pragma unroll
    for (int k = 0; k < R; ++k) D[k] = j == 0 ? T.t0[k] : (j == 1 ? T.t1[k] : (j == 2 ? T.t2[k] : T.t3[k]));
    VD = j == 0 ? T.v0 : (j == 1 ? T.v1 : (j == 2 ? T.v2 : T.v3));
  } else if (code == D_LOADS) {
    const DevInCol c = P::in(plan, op.a);
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
template <int NREG, int R, bool GATHER, class P, class RF>
__device__ __forceinline__ void tile_op(const DevPlan& plan, const DevOp& op, RF& f, uint64_t& vb, LoadTemps<R>& T, const Rows<R>& rs, uint32_t& alive,
                                        LDS uint64_t* sp, int n, int lane) {
  if (op.code != D_STORE) {
    vm_step<NREG, R, true, P>(plan, op, f, vb, T, rs, alive, sp);
    return;
  }
  const int ia = op.a & (NREG - 1);
  uint64_t A[R];
  if constexpr (IsLRegs<RF>::value) {
    f.get(ia, A);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) A[k] = f[k][ia];
  }
  const uint32_t VA = (uint32_t)(vb >> (ia * R));
  if (GATHER) store_out_strided<R>(P::out(plan, op.aux), rs.pos0, rs.exist, A, VA & rs.exist, lane);
  else store_out<R>(P::out(plan, op.aux), rs.pos0, n, A, VA & rs.exist, lane);
}

// One tile: the register file in VGPRs (a fresh array per tile, every
// register written before it is read) or, LF, the wave's LDS file at lfp.
template <int NREG, int R, bool GATHER, int B, class P = PlanSrc, bool LF = false>
__device__ __forceinline__ void vm_tile(const DevPlan& plan, LDS uint64_t* lfp, int64_t tile, int64_t count,
                                        const uint32_t* __restrict__ idx, int lane, LDS uint64_t* sp) {
  constexpr bool kLds = LF;
  constexpr int TILE = B * R;
  typedef typename RegVecT<NREG>::type RegVec;
  RegVec fv[R];
#pragma unroll
  for (int k = 0; k < R; ++k) fv[k] = (RegVec)(0);
  LRegs<NREG, R> fl{lfp};
  auto& f = [&]() -> auto& {
    if constexpr (LF) return fl;
    else return fv;
  }();
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
  if constexpr (P::kPlanOps) {
    // the next op's scalar loads are issued before this op runs, so their
    // latency hides behind it
    const int n_ops = plan.n_ops;
    DevOp next = load_op(plan, plan.split < n_ops ? plan.split : 0);
    for (int q = plan.split; q < n_ops; ++q) {
      const DevOp op = next;
      if constexpr (kLds) {
        if (op.code == D_LOADN) {
          lds_loads<NREG, R, true, P>(plan, op, q, f, vb, rs);
          q += op.aux;
          next = load_op(plan, q + 1 < n_ops ? q + 1 : q);
          continue;
        }
      }
      next = load_op(plan, q + 1 < n_ops ? q + 1 : q);
      tile_op<NREG, R, GATHER, P>(plan, op, f, vb, T, rs, alive, sp, n, lane);
    }
  } else {
    P::ops(plan, P::split(plan), P::n_ops(plan), [&](const DevOp& op) __attribute__((always_inline)) -> bool {
      tile_op<NREG, R, GATHER, P>(plan, op, f, vb, T, rs, alive, sp, n, lane);
      return false;
    });
  }
}

// The calling wave's spill area in the workgroup's dynamic LDS.
template <int R>
__device__ __forceinline__ LDS uint64_t* wave_spill(LDS uint64_t* base, int n_spill) {
  return base + (int64_t)(threadIdx.x >> 6) * n_spill * (64 * R + 8);
}

extern __shared__ uint64_t s_spill_area[];

// The LDS register file of the calling lane (LF kernels): after the
// workgroup's spill slots, NREG*64*R slots per wave, cleared once (every
// register is written before it is read within a tile). nullptr otherwise.
template <int NREG, int R, int B, bool LF>
__device__ __forceinline__ LDS uint64_t* lds_file(int n_spill) {
  if constexpr (!LF) {
    return nullptr;
  } else {
    const int lane = threadIdx.x & 63;
    LDS uint64_t* p = (LDS uint64_t*)s_spill_area + (int64_t)n_spill * (B / 64) * (64 * R + 8) +
                      (int64_t)(threadIdx.x >> 6) * (NREG * 64 * R) + lane;
#pragma unroll
    for (int i = 0; i < NREG * R; ++i) p[i * 64] = 0;
    return p;
  }
}

// perform_interpreter_operation (rows are the positions [0, num_rows)) or
// the output phase of a filter (GATHER: rows are the selected row ids
// idx[0, count), outputs at the positions in that list). Persistent:
// workgroups stride over tiles of B*R positions.
template <int NREG, int R, bool GATHER, int B, class P = PlanSrc, bool LF = false>
__global__ __launch_bounds__(B) void vm_rows(const DevPlan* __restrict__ pplan,
                                             const uint32_t* __restrict__ idx,
                                             const int64_t* __restrict__ d_count, int64_t capacity) {
  constexpr int TILE = B * R;
  const DevPlan& plan = *pplan;  // device memory: ops are scalar-cache loads
  const int lane = threadIdx.x & 63;
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plan.n_spill);
  LDS uint64_t* lfp = lds_file<NREG, R, B, LF>(plan.n_spill);
  int64_t count = plan.num_rows;
  if (GATHER) {
    count = *d_count;
    if (count > capacity) count = capacity;
  }
  for (int64_t tile = blockIdx.x; tile * TILE < count; tile += gridDim.x)
    vm_tile<NREG, R, GATHER, B, P, LF>(plan, lfp, tile, count, idx, lane, sp);
}

// Batched interpreter: every batch has its own plan (program and columns);
// the work items are all batches' tiles, batch b owning items
// [tile_start[b], tile_start[b+1]). Workgroups stride over the items, so
// uneven batch sizes spread over the whole chip in one persistent launch.
template <int NREG, int R, int B, bool LF = false>
__global__ __launch_bounds__(B) void vm_batches(const DevPlan* __restrict__ plans,
                                                const int64_t* __restrict__ tile_start, int n_batches) {
  const int lane = threadIdx.x & 63;
  const int64_t n_items = tile_start[n_batches];
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plans[0].n_spill);
  LDS uint64_t* lfp = lds_file<NREG, R, B, LF>(plans[0].n_spill);
  int lo = 0;  // last batch whose first item <= item (tile_start[0] = 0)
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    // items only grow, so the batch is lo or a later one: one load tells
    // (tile_start[n_batches] = n_items > item), a binary search over the
    // later batches otherwise
    if (tile_start[lo + 1] <= item) {
      int hi = n_batches - 1;
      ++lo;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= item) lo = mid;
        else hi = mid - 1;
      }
    }
    const DevPlan& plan = plans[lo];
    vm_tile<NREG, R, false, B, PlanSrc, LF>(plan, lfp, item - tile_start[lo], plan.num_rows, nullptr, lane, sp);
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

#ifndef BSQL_RTC
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
#endif

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
  if constexpr (__is_same(K, K_B8)) return x != 0;
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

// The slot bits of a raw element, as load_rows converts it.
template <class K>
__device__ __forceinline__ uint64_t kind_slot(typename K::T x) {
  if constexpr (__is_same(K, K_F64)) return bits(x);
  else if constexpr (__is_same(K, K_F32)) return bits((double)x);
  else if constexpr (__is_same(K, K_B8)) return (uint64_t)(x != 0);
  else if constexpr (__is_same(K, K_U8) || __is_same(K, K_U16) || __is_same(K, K_U32)) return (uint64_t)x;
  else return (uint64_t)(int64_t)x;
}

// One range test over this lane's S chunks: returns alive & valid & in range
// (bit k*R + j = row j of chunk k). A full tile issues every chunk's loads
// before the first test; the last, partial tile goes chunk by chunk with
// guarded loads. KEEP: the loaded elements are also returned as slot bits in
// kept[k][j] (defined for the rows still alive on return: their lanes
// loaded them).
template <class K, int R, int S, bool FULL, bool NT, bool KEEP>
__device__ __forceinline__ uint32_t range_chunks_k(const DevInCol& c, const DevRange& rg, int64_t row0,
                                                   int64_t num_rows, uint32_t alive, uint64_t (&kept)[S][R]) {
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
    for (int k = 0; k < S; ++k) {
      pass |= range_bits<K, R>(rg, v[k]) << (k * R);
      if constexpr (KEEP) {
#pragma unroll
        for (int j = 0; j < R; ++j) kept[k][j] = kind_slot<K>(v[k][j]);
      }
    }
    return alive & vm & pass;
  } else {
    uint32_t vm = 0;
    auto chunk = [&](int k) __attribute__((always_inline)) {
      const int64_t r0 = row0 + (int64_t)k * 64 * R;
      const int64_t left = num_rows - r0;
      const int n = ((alive >> (k * R)) & RMASK) == 0 ? 0 : (left <= 0 ? 0 : (left >= R ? R : (int)left));
      CT v[R];
      load_raw<CT, R, NT>(c, r0, n, v);
      vm |= load_valid<R>(c, r0, n) << (k * R);
      pass |= range_bits<K, R>(rg, v) << (k * R);
      if constexpr (KEEP) {
#pragma unroll
        for (int j = 0; j < R; ++j) kept[k][j] = kind_slot<K>(v[j]);
      }
    };
    if constexpr (KEEP) {
#pragma unroll
      for (int k = 0; k < S; ++k) chunk(k);  // kept[k] by constant index
    } else {
#pragma unroll 1
      for (int k = 0; k < S; ++k) chunk(k);
    }
    return alive & vm & pass;
  }
}

template <class K, int R, int S, bool FULL, bool NT = true>
__device__ __forceinline__ uint32_t range_chunks(const DevInCol& c, const DevRange& rg, int64_t row0,
                                                 int64_t num_rows, uint32_t alive) {
  uint64_t none[S][R];
  return range_chunks_k<K, R, S, FULL, NT, false>(c, rg, row0, num_rows, alive, none);
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

// range_test that also returns the elements' slot bits (range_chunks_k, KEEP).
template <int R, int S, bool FULL, bool NT>
__device__ __forceinline__ uint32_t range_test_keep(const DevInCol& c, const DevRange& rg, int64_t row0,
                                                    int64_t num_rows, uint32_t alive, uint64_t (&kept)[S][R]) {
  switch (c.type) {
    case BSQL_INT8: return range_chunks_k<K_I8, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_INT16: return range_chunks_k<K_I16, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_INT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS:
      return range_chunks_k<K_I32, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_UINT8: return range_chunks_k<K_U8, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_UINT16: return range_chunks_k<K_U16, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_UINT32: return range_chunks_k<K_U32, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_BOOL8: return range_chunks_k<K_B8, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_FLOAT32: return range_chunks_k<K_F32, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    case BSQL_FLOAT64: return range_chunks_k<K_F64, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
    default: return range_chunks_k<K_I64, R, S, FULL, NT, true>(c, rg, row0, num_rows, alive, kept);
  }
}

// Row selection by range tests (plan.ranges): one range at a time over the
// wave's whole slice, so each column is loaded with every chunk in flight
// and its type is dispatched once per slice; no register file. Waves stride
// over slices independently.
template <int R, int S, class P = PlanSrc>
__global__ __launch_bounds__(BLOCK) void select_ranges(const DevPlan* __restrict__ pplan, SelArgs sa) {
  typedef SelectTile<R, S> T;
  const DevPlan& plan = *pplan;
  const int lane = threadIdx.x & 63;
  const int64_t num_rows = plan.num_rows;
  const int nr = P::n_ranges(plan);
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
      const DevRange rg = P::range(plan, r);
      if (rg.kind & RK_EMPTY) {
        alive = 0;
        break;
      }
      if (full)
        alive = range_test<R, S, true>(P::in(plan, rg.col), rg, row0, num_rows, alive);
      else
        alive = range_test<R, S, false>(P::in(plan, rg.col), rg, row0, num_rows, alive);
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
__host__ __device__ constexpr int64_t fused_lds_bytes(int n_out) { return (int64_t)n_out * FUSED_VWORDS * 4; }

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

// One pass for a filter of range tests with fixed-width outputs, for kernels
// generated for one program (P supplies constant op fields and column
// types). Workgroup tile blockIdx.x holds W wave slices of S chunks of 64
// lanes x R rows (lane t owns rows t*R .. t*R+R-1 of each chunk).
// Workgroups are dispatched in index order, so every earlier tile is
// resident or done when a tile looks back, and the smallest unfinished tile
// never waits.
//
// Each wave tests the ranges over its slice (each column's loads for the
// whole slice in flight) and the workgroup publishes its row count at once.
// If the tile selected at most P::kStageRows rows, the lanes that own a
// selected row then load their R rows of the output columns (whole
// vectors), run the output ops straight-line and put each result in LDS at
// its position within the tile (8-byte slots, a null bit per row). Only then
// does wave 0 look back for the tile's output position: by that time the
// earlier tiles have mostly published theirs, so one round trip usually
// settles it, and the output loads were in flight meanwhile instead of
// after it. The workgroup writes its outputs as one contiguous run per
// column (consecutive lanes, consecutive rows, non-temporal) and its
// validity as whole words; the two words at the run's ends, shared with
// the neighbouring tiles, are merged with atomics that change only the
// tile's own bits (one OR when none of its rows is null, else an AND/OR
// pair). A tile that selected more rows than LDS holds takes its position
// first and writes each selected row from the registers (per-wave validity
// words in LDS). Only the tile status words need clearing before the
// launch. P::kKeepCol >= 0: that column's values, loaded by its range test,
// are kept in registers and used by the output ops instead of a second load.
template <int R>
__device__ __forceinline__ void temp_set(LoadTemps<R>& T, int j, const uint64_t (&v)[R], uint32_t vm) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (j == 0) T.t0[i] = v[i];
    else if (j == 1) T.t1[i] = v[i];
    else if (j == 2) T.t2[i] = v[i];
    else T.t3[i] = v[i];
  }
  if (j == 0) T.v0 = vm;
  else if (j == 1) T.v1 = vm;
  else if (j == 2) T.v2 = vm;
  else T.v3 = vm;
}

// All range tests [r, P::kRanges) of a generated program over one slice,
// with every column's loads issued before the first test: range r's element
// kind, column and whether its values are kept are constants of the
// program (P::RangeAt<r>). Returns alive & valid & in range for all of them.
template <class P, int R, int S, bool FULL, int r>
__device__ __forceinline__ uint32_t eager_ranges(const DevPlan& plan, int64_t row0, int64_t num_rows,
                                                 uint32_t alive, uint64_t (&kept)[S][R]) {
  if constexpr (r >= P::kRanges) {
    return alive;
  } else {
    typedef typename P::template RangeAt<r> RA;
    typedef typename RA::K K;
    typedef typename K::T CT;
    const DevInCol c = P::in(plan, RA::col);
    CT v[S][R];
    uint32_t vm = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int64_t r0 = row0 + (int64_t)k * 64 * R;
      int n = R;
      if constexpr (!FULL) {
        const int64_t left = num_rows - r0;
        n = left <= 0 ? 0 : (left >= R ? R : (int)left);
      }
      if (P::nt(RA::col, true))
        load_raw<CT, R, true>(c, r0, n, v[k]);
      else
        load_raw<CT, R, false>(c, r0, n, v[k]);
      vm |= load_valid<R>(c, r0, n) << (k * R);
    }
    // the later ranges' loads go out before this range's test
    const uint32_t rest = eager_ranges<P, R, S, FULL, r + 1>(plan, row0, num_rows, alive, kept);
    const DevRange rg = P::range(plan, r);
    uint32_t pass = 0;
    if (!(rg.kind & RK_EMPTY)) {
#pragma unroll
      for (int k = 0; k < S; ++k) pass |= range_bits<K, R>(rg, v[k]) << (k * R);
    }
    if constexpr (RA::keep) {
#pragma unroll
      for (int k = 0; k < S; ++k)
#pragma unroll
        for (int j = 0; j < R; ++j) kept[k][j] = kind_slot<K>(v[k][j]);
    }
    return rest & vm & pass;
  }
}

// The output ops over a wave's selected rows, chunk by chunk: the lanes
// that own a selected row load their R rows (whole vectors) and run the ops
// on them. The rows selected before a lane's first row of chunk k give its
// position: pos_start + (rows in chunks < k) + (rows of lower lanes in k).
// CAP > 0: results go to LDS slot o*CAP + position, a null result sets its
// bit in nulls[o*fp_null_words(CAP) ..] and bit o of *nullmask. CAP == 0:
// results go to base + position (below capacity), their validity bits to
// vbits[o*vw ..] at (base & 31) + position.
template <class P, int NREG, int R, int S, int CAP>
__device__ __forceinline__ void out_rows(const DevPlan& plan, uint32_t alive, int64_t row0, int64_t num_rows,
                                         int pos_start, bool have_kept, const uint64_t (&kept)[S][R],
                                         LDS uint64_t* slots, LDS uint32_t* nulls, LDS uint32_t* nullmask,
                                         int64_t base, int64_t capacity, LDS uint32_t* vbits, int vw) {
  typedef typename RegVecT<NREG>::type RegVec;
  constexpr int CHUNK = 64 * R;
  constexpr uint32_t RMASK = (1u << R) - 1u;
  constexpr int KEEP = P::kKeepCol;
  const int shift = (int)(base & 31);
  int run = pos_start;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const uint32_t a = (alive >> (k * R)) & RMASK;
    int lane_excl, chunk_total;
    chunk_counts<R>(a, lane_excl, chunk_total);
    if (chunk_total == 0) continue;  // wave-uniform
    Rows<R> rs;
    rs.row0 = row0 + (int64_t)k * CHUNK;
    rs.pos0 = rs.row0;
    rs.pstride = 1;
    rs.contig = true;
    const int64_t left = num_rows - rs.row0;
    const int exist = left <= 0 ? 0 : (left >= R ? R : (int)left);
    rs.nrows = a ? exist : 0;  // lanes without a selected row load nothing
    rs.exist = (1u << rs.nrows) - 1u;
#pragma unroll
    for (int j = 0; j < R; ++j) rs.row[j] = rs.row0 + j;
    RegVec f[R];
#pragma unroll
    for (int j = 0; j < R; ++j) f[j] = (RegVec)(0);
    uint64_t vb = 0;
    LoadTemps<R> T;
    uint32_t al = rs.exist;
    const int p0 = run + lane_excl;  // position of the lane's first selected row
    P::ops(plan, P::split(plan), P::n_ops(plan), [&](const DevOp& op) __attribute__((always_inline)) -> bool {
      if (KEEP >= 0 && op.code == D_LOADN) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          if (jj >= op.aux) break;
          const int col = (int)((op.imm >> (8 * jj)) & 0xff);
          uint64_t v[R];
          uint32_t vm = 0;
          if (col == KEEP && have_kept) {
            // selected rows passed the column's range test: valid
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = kept[k][i];
            vm = rs.exist;
          } else if (P::nt(col, true)) {
            col_load<R, true>(P::in(plan, col), rs, v, vm);
          } else {
            col_load<R, false>(P::in(plan, col), rs, v, vm);
          }
          temp_set<R>(T, jj, v, vm);
        }
        return false;
      }
      if (op.code != D_STORE) {
        vm_step<NREG, R, true, P>(plan, op, f, vb, T, rs, al, nullptr);
        return false;
      }
      const int ia = op.a & (NREG - 1);
      const uint32_t VA = (uint32_t)(vb >> (ia * R)) & a;
      const DevOutCol oc = P::out(plan, op.aux);
      int q = p0;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (!((a >> j) & 1u)) continue;
        if constexpr (CAP > 0) {
          slots[op.aux * CAP + q] = f[j][ia];
          if (oc.valid && !((VA >> j) & 1u)) {
            lds_set_bit(nulls + op.aux * fp_null_words(CAP), q);
            __hip_atomic_fetch_or(nullmask, 1u << op.aux, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else if (base + q < capacity) {
          // cached stores: the lanes' pieces of a line meet in L2
          store_elem<false>(oc, base + q, f[j][ia]);
          if (oc.valid && ((VA >> j) & 1u)) lds_set_bit(vbits + op.aux * vw, shift + q);
        }
        ++q;
      }
      return false;
    });
    run += chunk_total;
  }
}

// A one-pass tile's staged rows (LDS slots s_slots + o * CAP, position i of
// the tile) written to positions tbase + i below the capacity, one
// contiguous run per output (consecutive lanes, consecutive rows,
// non-temporal), then their validity as whole words: the two words at the
// run's ends, shared with the neighbouring tiles, merged with atomics that
// change only the tile's bits (one OR when no row of the output is null,
// bit o of nullmask; else an AND/OR pair). s_nulls + o * fp_null_words(CAP):
// the null bits by tile position.
template <class P, int W, int CAP>
__device__ __forceinline__ void fp_flush_staged(const DevPlan& plan, const FusedArgs& fa, int64_t tbase, int total,
                                                const LDS uint64_t* s_slots, const LDS uint32_t* s_nulls,
                                                uint32_t nullmask) {
  constexpr int NULLW = fp_null_words(CAP);
  const int n_out = plan.n_out;
  const int64_t room = fa.capacity - tbase;
  const int nw = room <= 0 ? 0 : (room < total ? (int)room : total);
  for (int o = 0; o < n_out; ++o) {
    const DevOutCol oc = P::out(plan, o);
    const LDS uint64_t* sl = s_slots + o * CAP;
    for (int i = threadIdx.x; i < nw; i += 64 * W) store_elem<true>(oc, tbase + i, sl[i]);
  }
  if (nw == 0) return;
  const int64_t end = tbase + nw;
  const int64_t w0 = tbase >> 5, w1 = (end - 1) >> 5;
  for (int o = 0; o < n_out; ++o) {
    uint32_t* words = P::out(plan, o).valid;
    if (!words) continue;
    const LDS uint32_t* nb = s_nulls + o * NULLW;
    const bool any_null = (nullmask >> o) & 1u;
    for (int64_t g = w0 + threadIdx.x; g <= w1; g += 64 * W) {
      uint32_t own = 0xffffffffu;
      if (g == w0) own &= 0xffffffffu << (tbase & 31);
      if (g == w1 && (end & 31)) own &= 0xffffffffu >> (32 - (end & 31));
      if (!any_null) {
        if (own == 0xffffffffu) gpw<uint32_t>(words)[g] = own;
        else __hip_atomic_fetch_or(gpw<uint32_t>(words) + g, own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        continue;
      }
      // null bits of tile positions [l0, l0 + 32)
      const int64_t l0 = g * 32 - tbase;
      uint32_t x;
      if (l0 < 0) {
        x = nb[0] << (int)(-l0);
      } else {
        const int wi = (int)(l0 >> 5), sh = (int)(l0 & 31);
        const uint64_t two = (uint64_t)nb[wi] | ((uint64_t)nb[wi + 1] << 32);
        x = (uint32_t)(two >> sh);
      }
      merge_word(words, g, own, ~x & own);
    }
  }
}

// A wave's selected rows (alive, its slice from row0) written from the
// registers at base, base + 1, .. (below the capacity), then their validity:
// whole words inside the wave's range, the boundary words merged. vbits: the
// wave's LDS validity words (VW per output), cleared by the caller before a
// barrier.
template <class P, int NREG, int R, int S>
__device__ __forceinline__ void fp_write_direct(const DevPlan& plan, const FusedArgs& fa, uint32_t alive,
                                                int64_t row0, int64_t base, int cnt, bool have_kept,
                                                const uint64_t (&kept)[S][R], LDS uint32_t* vbits) {
  constexpr int VW = S * 64 * R / 32 + 2;
  const int lane = threadIdx.x & 63;
  const int n_out = plan.n_out;
  if (cnt > 0)
    out_rows<P, NREG, R, S, 0>(plan, alive, row0, plan.num_rows, 0, have_kept, kept, nullptr, nullptr, nullptr, base,
                               fa.capacity, vbits, VW);
  __syncthreads();
  int64_t end = base + cnt;
  if (end > fa.capacity) end = fa.capacity;
  if (cnt == 0 || end <= base) return;
  const int64_t w0 = base >> 5, w1 = (end - 1) >> 5;
  for (int o = 0; o < n_out; ++o) {
    uint32_t* words = P::out(plan, o).valid;
    if (!words) continue;
    for (int64_t g = w0 + lane; g <= w1; g += 64) {
      uint32_t own = 0xffffffffu;
      if (g == w0) own &= 0xffffffffu << (base & 31);
      if (g == w1 && (end & 31)) own &= 0xffffffffu >> (32 - (end & 31));
      merge_word(words, g, own, vbits[o * VW + (int)(g - w0)] & own);
    }
  }
}

// One tile (blockIdx.x) of the one-pass filter + project of plan / fa.
template <class P, int NREG, int R, int S, int W>
__device__ __forceinline__ void fp_tile(const DevPlan& plan, const FusedArgs& fa) {
  typedef typename RegVecT<NREG>::type RegVec;
  constexpr int CHUNK = 64 * R;
  constexpr int SLICE = S * CHUNK;
  constexpr int VW = SLICE / 32 + 2;  // validity words a slice's outputs can touch
  constexpr int CAP = P::kStageRows;
  constexpr int NULLW = fp_null_words(CAP);
  static_assert(R * S <= 32, "selection bits per lane");
  static_assert(CAP % 64 == 0 && CAP <= W * SLICE, "staged rows: whole waves, at most a tile");
  __shared__ int s_cnt[W];
  __shared__ int64_t s_base;
  __shared__ uint32_t s_nullmask;  // bit o: a staged row of output o is null
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t num_rows = plan.num_rows;
  const int64_t tile = blockIdx.x;
  const int n_out = plan.n_out;
  // dynamic LDS (fp_lds_bytes): staged slots, their null bits, then each
  // wave's validity words for tiles written from the registers
  LDS uint64_t* s_slots = (LDS uint64_t*)s_spill_area;
  LDS uint32_t* s_nulls = (LDS uint32_t*)(s_slots + (int64_t)n_out * CAP);
  LDS uint32_t* vbits = s_nulls + (int64_t)n_out * NULLW + (int64_t)wave * n_out * VW;
  // ---- range tests over the wave's slice
  const int64_t slice_row0 = (tile * W + wave) * SLICE;
  const int64_t row0 = slice_row0 + lane * R;
  uint32_t alive = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int64_t left = num_rows - (row0 + (int64_t)k * CHUNK);
    const int n = left <= 0 ? 0 : (left >= R ? R : (int)left);
    alive |= ((1u << n) - 1u) << (k * R);
  }
  const bool full = slice_row0 + SLICE <= num_rows;
  const int nr = P::n_ranges(plan);
  constexpr int KEEP = P::kKeepCol;
  uint64_t kept[S][R];
  bool have_kept = false;
  if constexpr (P::kEager) {
    alive = full ? eager_ranges<P, R, S, true, 0>(plan, row0, num_rows, alive, kept)
                 : eager_ranges<P, R, S, false, 0>(plan, row0, num_rows, alive, kept);
    have_kept = KEEP >= 0;
  }
#pragma unroll
  for (int r = 0; r < (P::kEager ? 0 : nr); ++r) {
    if (__ballot(alive != 0) == 0) break;
    const DevRange rg = P::range(plan, r);
    if (rg.kind & RK_EMPTY) {
      alive = 0;
      break;
    }
    const DevInCol c = P::in(plan, rg.col);
    if (KEEP >= 0 && rg.col == KEEP && !have_kept) {
      // rows alive after this test (hence all rows selected in the end)
      // had their values loaded by it
      have_kept = true;
      alive = full ? range_test_keep<R, S, true, true>(c, rg, row0, num_rows, alive, kept)
                   : range_test_keep<R, S, false, true>(c, rg, row0, num_rows, alive, kept);
    } else if (P::nt(rg.col, true)) {
      alive = full ? range_test<R, S, true, true>(c, rg, row0, num_rows, alive)
                   : range_test<R, S, false, true>(c, rg, row0, num_rows, alive);
    } else {
      alive = full ? range_test<R, S, true, false>(c, rg, row0, num_rows, alive)
                   : range_test<R, S, false, false>(c, rg, row0, num_rows, alive);
    }
  }
  const int cnt = (int)wave_sum(__popc(alive));
  for (int i = lane; i < n_out * VW; i += 64) vbits[i] = 0u;
  if constexpr (CAP > 0) {
    for (int i = threadIdx.x; i < n_out * NULLW; i += 64 * W) s_nulls[i] = 0u;
    if (threadIdx.x == 0) s_nullmask = 0u;
  }
  if (lane == 0) s_cnt[wave] = cnt;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    before += w < wave ? s_cnt[w] : 0;
    total += s_cnt[w];
  }
  // the tile's count goes out now; its position is resolved after staging
  if (threadIdx.x == 0) st_status(&fa.status[tile], (tile == 0 ? ST_PREFIX : ST_AGG) | (uint64_t)total);
  const bool stage = CAP > 0 && total <= CAP;  // workgroup-uniform
  if (stage && cnt > 0)
    out_rows<P, NREG, R, S, CAP>(plan, alive, row0, num_rows, before, have_kept, kept, s_slots, s_nulls,
                                 (LDS uint32_t*)&s_nullmask, 0, fa.capacity, nullptr, 0);
  __syncthreads();  // staged rows are in LDS
  if (wave == 0) tile_position(fa, tile, total, lane, &s_base);
  __syncthreads();
  if (total == 0) return;  // workgroup-uniform
  if (stage) {
    fp_flush_staged<P, W, CAP>(plan, fa, s_base, total, s_slots, s_nulls, s_nullmask);
    return;
  }
  // ---- more rows than LDS holds: each wave writes its selected rows at
  // the tile position plus the rows selected before them
  fp_write_direct<P, NREG, R, S>(plan, fa, alive, row0, s_base + before, cnt, have_kept, kept, vbits);
}

template <class P, int NREG, int R, int S, int W>
__global__ __launch_bounds__(64 * W) void filter_project_rows(const DevPlan* __restrict__ pplan, FusedArgs fa) {
  fp_tile<P, NREG, R, S, W>(*pplan, fa);
}

// The outputs of one tile of plan / fa whose position tbase is already
// known (the batched launch finds every query's position before writing
// any): the rows `alive` selects (the lane's bits, wave slices as in
// fp_tile), `before` of them in the tile's lower waves, `cnt` in this wave,
// `total` in the tile. Staged in LDS and written as runs when total <=
// P::kStageRows, else written from the registers; then their validity.
template <class P, int NREG, int R, int S, int W>
__device__ __forceinline__ void fp_write_at(const DevPlan& plan, const FusedArgs& fa, uint32_t alive, int64_t tbase,
                                            int before, int cnt, int total) {
  constexpr int CHUNK = 64 * R;
  constexpr int SLICE = S * CHUNK;
  constexpr int VW = SLICE / 32 + 2;
  constexpr int CAP = P::kStageRows;
  constexpr int NULLW = fp_null_words(CAP);
  static_assert(CAP % 64 == 0 && CAP <= W * SLICE, "staged rows: whole waves, at most a tile");
  __shared__ uint32_t s_nullmask;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n_out = plan.n_out;
  const int64_t tile = blockIdx.x;
  LDS uint64_t* s_slots = (LDS uint64_t*)s_spill_area;
  LDS uint32_t* s_nulls = (LDS uint32_t*)(s_slots + (int64_t)n_out * CAP);
  LDS uint32_t* vbits = s_nulls + (int64_t)n_out * NULLW + (int64_t)wave * n_out * VW;
  const int64_t row0 = (tile * W + wave) * SLICE + lane * R;
  for (int i = lane; i < n_out * VW; i += 64) vbits[i] = 0u;
  if constexpr (CAP > 0) {
    for (int i = threadIdx.x; i < n_out * NULLW; i += 64 * W) s_nulls[i] = 0u;
    if (threadIdx.x == 0) s_nullmask = 0u;
  }
  __syncthreads();
  if (total == 0) return;  // workgroup-uniform
  uint64_t none[S][R];
  if (CAP > 0 && total <= CAP) {
    if (cnt > 0)
      out_rows<P, NREG, R, S, CAP>(plan, alive, row0, plan.num_rows, before, false, none, s_slots, s_nulls,
                                   (LDS uint32_t*)&s_nullmask, 0, fa.capacity, nullptr, 0);
    __syncthreads();
    fp_flush_staged<P, W, CAP>(plan, fa, tbase, total, s_slots, s_nulls, s_nullmask);
    return;
  }
  fp_write_direct<P, NREG, R, S>(plan, fa, alive, row0, tbase + before, cnt, false, none, vbits);
}

// Several filter + project queries over the same table in one launch, for
// kernels generated for the set (B: its selection; Ps: one source per
// query). Workgroup tile blockIdx.x holds W wave slices (as fp_tile). B::select
// tests the queries' ranges column by column: each table column that some
// range reads is loaded once, for the lanes where a query testing it still
// has a row, and every such range is applied to it. The workgroup then
// publishes every query's count, waves look back for the queries' output
// positions in parallel (wave w for queries w, w + W, ..), and each query's
// outputs are written (fp_write_at). Each query has its own plan, tile status
// words, count and outputs; a tile waits only on smaller tiles of the same
// query, which are resident or done (dispatch in index order).
template <int NREG, int R, int S, int W, class B, class... Ps>
__global__ __launch_bounds__(64 * W) void filter_project_batch(MultiArgs ma) {
  constexpr int NQ = sizeof...(Ps);
  constexpr int CHUNK = 64 * R;
  constexpr int SLICE = S * CHUNK;
  static_assert(R * S <= 32, "selection bits per lane");
  __shared__ int s_cnt[NQ][W];
  __shared__ int64_t s_base[NQ];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t tile = blockIdx.x;
  const int64_t num_rows = ma.plan[0]->num_rows;
  const int64_t slice_row0 = (tile * W + wave) * SLICE;
  const int64_t row0 = slice_row0 + lane * R;
  uint32_t exist = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int64_t left = num_rows - (row0 + (int64_t)k * CHUNK);
    const int n = left <= 0 ? 0 : (left >= R ? R : (int)left);
    exist |= ((1u << n) - 1u) << (k * R);
  }
  uint32_t alive[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) alive[q] = exist;
  if (slice_row0 + SLICE <= num_rows)
    B::template select<R, S, true>(ma, row0, num_rows, alive);
  else
    B::template select<R, S, false>(ma, row0, num_rows, alive);
  int cnt[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    cnt[q] = __builtin_amdgcn_readfirstlane((int)wave_sum(__popc(alive[q])));
    if (lane == 0) s_cnt[q][wave] = cnt[q];
  }
  __syncthreads();
  int before[NQ], total[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    int b = 0, t = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int v = s_cnt[q][w];
      b += w < wave ? v : 0;
      t += v;
    }
    before[q] = b;
    total[q] = t;
    if (threadIdx.x == q)
      st_status(&ma.fa[q].status[tile], (tile == 0 ? ST_PREFIX : ST_AGG) | (uint64_t)t);
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
    if (wave == q % W) tile_position(ma.fa[q], tile, total[q], lane, &s_base[q]);
  __syncthreads();
  int q = 0;
  ((fp_write_at<Ps, NREG, R, S, W>(*ma.plan[q], ma.fa[q], alive[q], s_base[q], before[q], cnt[q], total[q]),
    __syncthreads(), ++q),
   ...);
}

// Row selection by ops [0, split) (KILL per conjunct with a wave-uniform
// early exit, or one SEL), chunk by chunk over the wave's slice.
// LF: the register file in LDS (after the spill slots), a LOADN and its
// COMMITs run as one step.
template <int NREG, int R, int S, int B, class P = PlanSrc, bool LF = false>
__global__ __launch_bounds__(B) void select_vm(const DevPlan* __restrict__ pplan, SelArgs sa) {
  typedef typename RegVecT<NREG>::type RegVec;
  typedef SelectTile<R, S> T;
  constexpr int NW = B / 64;
  const DevPlan& plan = *pplan;
  const int lane = threadIdx.x & 63;
  LDS uint64_t* sp = wave_spill<R>((LDS uint64_t*)s_spill_area, plan.n_spill);
  LDS uint64_t* lfp = lds_file<NREG, R, B, LF>(plan.n_spill);
  const int64_t s0 = (int64_t)blockIdx.x * sa.per;
  const int64_t s1 = s0 + sa.per < sa.n_slices ? s0 + sa.per : sa.n_slices;
  uint32_t wave_total = 0;
  const int split = P::split(plan);
  for (int64_t slice = s0 + (threadIdx.x >> 6); slice < s1; slice += NW) {
    RegVec fv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) fv[k] = (RegVec)(0);
    LRegs<NREG, R> fl{lfp};
    auto& f = [&]() -> auto& {
      if constexpr (LF) return fl;
      else return fv;
    }();
    uint64_t vb = 0;
    LoadTemps<R> L;
    uint32_t alive_all = 0;
#pragma unroll 1
    for (int k = 0; k < S; ++k) {
      Rows<R> rs;
      chunk_rows<R>(rs, slice * T::WAVE_ROWS + k * T::CHUNK + lane * R, plan.num_rows);
      uint32_t alive = rs.exist;
      if (__ballot(alive != 0) != 0) {
        if constexpr (P::kPlanOps) {
          DevOp next = load_op(plan, 0);
          for (int q = 0; q < split; ++q) {
            const DevOp op = next;
            if constexpr (LF) {
              if (op.code == D_LOADN) {
                lds_loads<NREG, R, true, P>(plan, op, q, f, vb, rs);
                q += op.aux;
                next = load_op(plan, q + 1 < split ? q + 1 : q);
                continue;
              }
            }
            next = load_op(plan, q + 1 < split ? q + 1 : q);
            if (vm_step<NREG, R, true, P>(plan, op, f, vb, L, rs, alive, sp)) break;
          }
        } else {
          P::ops(plan, 0, split, [&](const DevOp& op) __attribute__((always_inline)) -> bool {
            return vm_step<NREG, R, true, P>(plan, op, f, vb, L, rs, alive, sp);
          });
        }
      }
      alive_all |= alive << (k * R);
    }
    wave_total += emit_slice<R, S>(sa, slice, alive_all);
  }
  emit_block<B>(sa, wave_total);
}

}  // namespace bsql_dev
