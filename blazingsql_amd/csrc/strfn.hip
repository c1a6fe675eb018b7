// strfn.hip -- kernels that build string columns before the interpreter runs
// (the columns libcudf builds in the reference: LogicalProject.cpp:100-742).
//
// Each column is built in three stream-ordered launches: a length pass (one
// thread per row computes the row's byte length and validity), the offsets
// scan shared with the interpreter's string outputs, and a write pass that
// recomputes each row into its place. A row is short and independent, so one
// lane per row with a grid-stride over 64-row groups keeps a wave's validity
// bits in one ballot.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "casemap.h"
#include "strfn.h"
#include "../../include/bsql/interp.h"

namespace bsql_dev {

using bsql::StrFnArgs;
using bsql::StrSrc;

constexpr int SF_BLOCK = 256;

__constant__ uint64_t CASE_UPPER[BSQL_CASE_UPPER_N] = {BSQL_CASE_UPPER_VALUES};
__constant__ uint64_t CASE_LOWER[BSQL_CASE_LOWER_N] = {BSQL_CASE_LOWER_VALUES};
__constant__ uint32_t ALPHA_RANGES[2 * BSQL_ALPHA_RANGES_N] = {BSQL_ALPHA_RANGES_VALUES};

// Bytes out: counts them (length pass) or stores them (write pass).
struct Out {
  uint8_t* p;
  int32_t n;
  __device__ void put(uint8_t c) {
    if (p) p[n] = c;
    ++n;
  }
  __device__ void put(const uint8_t* s, int32_t len) {
    if (p)
      for (int32_t i = 0; i < len; ++i) p[n + i] = s[i];
    n += len;
  }
  __device__ void put_cp(uint32_t cp) {
    if (cp < 0x80) {
      put((uint8_t)cp);
    } else if (cp < 0x800) {
      put((uint8_t)(0xC0 | (cp >> 6)));
      put((uint8_t)(0x80 | (cp & 0x3F)));
    } else {
      put((uint8_t)(0xE0 | (cp >> 12)));
      put((uint8_t)(0x80 | ((cp >> 6) & 0x3F)));
      put((uint8_t)(0x80 | (cp & 0x3F)));
    }
  }
};

__device__ __forceinline__ bool bit_at(const uint32_t* v, int32_t vbit, int64_t row) {
  if (!v) return true;
  const int64_t b = row + vbit;
  return (v[b >> 5] >> (b & 31)) & 1u;
}

// Bytes of the character that starts with lead byte c (1 for a stray byte).
__device__ __forceinline__ int32_t char_bytes(uint8_t c) {
  return c < 0xC0 ? 1 : (c < 0xE0 ? 2 : (c < 0xF0 ? 3 : 4));
}

// Code point of the n-byte character at s (n from char_bytes).
__device__ __forceinline__ uint32_t code_point(const uint8_t* s, int32_t n) {
  if (n == 1) return s[0];
  if (n == 2) return ((uint32_t)(s[0] & 0x1F) << 6) | (s[1] & 0x3F);
  if (n == 3) return ((uint32_t)(s[0] & 0x0F) << 12) | ((uint32_t)(s[1] & 0x3F) << 6) | (s[2] & 0x3F);
  return ((uint32_t)(s[0] & 0x07) << 18) | ((uint32_t)(s[1] & 0x3F) << 12) | ((uint32_t)(s[2] & 0x3F) << 6) |
         (s[3] & 0x3F);
}

// Case-map entry of cp in a table sorted by code point (0 when absent).
__device__ uint64_t case_entry(const uint64_t* table, int n, uint32_t cp) {
  if (cp > 0xFFFF) return 0;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const uint32_t k = (uint32_t)(table[mid] >> 48);
    if (k == cp) return table[mid];
    if (k < cp) lo = mid + 1; else hi = mid;
  }
  return 0;
}

__device__ bool is_alpha(uint32_t cp) {
  if (cp > 0xFFFF) return false;
  int lo = 0, hi = BSQL_ALPHA_RANGES_N;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cp < ALPHA_RANGES[2 * mid]) hi = mid;
    else if (cp >= ALPHA_RANGES[2 * mid + 1]) lo = mid + 1;
    else return true;
  }
  return false;
}

__device__ void put_mapped(Out& o, uint64_t e) {
  o.put_cp((uint32_t)(e >> 32) & 0xFFFF);
  if ((e >> 16) & 0xFFFF) o.put_cp((uint32_t)(e >> 16) & 0xFFFF);
  if (e & 0xFFFF) o.put_cp((uint32_t)e & 0xFFFF);
}

__device__ void put_uint(Out& o, uint64_t v) {
  uint8_t buf[20];
  int n = 0;
  do {
    buf[n++] = (uint8_t)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) o.put(buf[--n]);
}

__device__ void put_padded(Out& o, int64_t v, int width) {
  if (v < 0) {
    o.put('-');
    v = -v;
  }
  uint8_t buf[20];
  int n = 0;
  do {
    buf[n++] = (uint8_t)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n < width) buf[n++] = '0';
  while (n) o.put(buf[--n]);
}

// libcudf 21.08 from_floats (ftos_converter): NaN, Inf, -Inf; otherwise the
// value is brought into [1e-4, 1e9] by exact power-of-two-indexed steps of
// 10^k, split into an integer part and 10 significant digits in all (the
// leftover fraction rounds the last digit half up), trailing zeros dropped
// but one decimal kept, then e+XX / e-XX when the value was rescaled.
__device__ void put_float(Out& o, double value) {
  const double upper10[9] = {10, 100, 10000, 1e8, 1e16, 1e32, 1e64, 1e128, 1e256};
  const double lower10[9] = {.1, .01, .0001, 1e-8, 1e-16, 1e-32, 1e-64, 1e-128, 1e-256};
  const double blower10[9] = {1.0, .1, .001, 1e-7, 1e-15, 1e-31, 1e-63, 1e-127, 1e-255};
  if (value != value) {
    o.put((const uint8_t*)"NaN", 3);
    return;
  }
  if (__builtin_signbit(value)) {
    o.put('-');
    value = -value;
  }
  if (value == __builtin_inf()) {
    o.put((const uint8_t*)"Inf", 3);
    return;
  }
  int exp10 = 0;
  if (value > 1e9) {
    int fx = 256;
    for (int idx = 8; idx >= 0; --idx) {
      if (value >= upper10[idx]) {
        value = __dmul_rn(value, lower10[idx]);
        exp10 += fx;
      }
      fx >>= 1;
    }
  } else if (value > 0.0 && value < 0.0001) {
    int fx = 256;
    for (int idx = 8; idx >= 0; --idx) {
      if (value < blower10[idx]) {
        value = __dmul_rn(value, upper10[idx]);
        exp10 -= fx;
      }
      fx >>= 1;
    }
  }
  int places = 9;
  uint32_t max_digits = 1000000000u;
  uint32_t integer = (uint32_t)value;
  for (uint32_t i = integer; i >= 10; i /= 10) {
    --places;
    max_digits /= 10;
  }
  double rem = __dmul_rn(__dsub_rn(value, (double)integer), (double)max_digits);
  uint32_t decimal = (uint32_t)rem;
  rem = __dsub_rn(rem, (double)decimal);
  decimal += (uint32_t)__dmul_rn(2.0, rem);
  if (decimal >= max_digits) {
    decimal = 0;
    ++integer;
    if (exp10 && integer >= 10) {
      ++exp10;
      integer = 1;
    }
  }
  while (decimal % 10 == 0 && places > 0) {
    decimal /= 10;
    --places;
  }
  put_uint(o, integer);
  o.put('.');
  if (places) {
    uint8_t buf[10];
    for (int k = places - 1; k >= 0; --k) {
      buf[k] = (uint8_t)('0' + decimal % 10);
      decimal /= 10;
    }
    o.put(buf, places);
  } else {
    o.put('0');
  }
  if (exp10) {
    o.put('e');
    o.put(exp10 < 0 ? '-' : '+');
    const int e = exp10 < 0 ? -exp10 : exp10;
    if (e < 10) o.put('0');
    put_uint(o, (uint64_t)e);
  }
}

// days since 1970-01-01 -> proleptic Gregorian y/m/d (Hinnant's civil_from_days)
__device__ void civil(int64_t z, int64_t* y, int* m, int* d) {
  z += 719468;
  const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  const int64_t doe = z - era * 146097;
  const int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const int64_t mp = (5 * doy + 2) / 153;
  *d = (int)(doy - (153 * mp + 2) / 5 + 1);
  *m = (int)(mp < 10 ? mp + 3 : mp - 9);
  *y = yoe + era * 400 + (*m <= 2 ? 1 : 0);
}

__device__ int64_t floor_div(int64_t a, int64_t b) {
  const int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// from_timestamps with "%Y-%m-%d" (days) or "%Y-%m-%d %H:%M:%S" (others)
__device__ void put_timestamp(Out& o, int type, int64_t ticks) {
  int64_t per_sec = 1;
  switch (type) {
    case BSQL_TIMESTAMP_MILLISECONDS: per_sec = 1000; break;
    case BSQL_TIMESTAMP_MICROSECONDS: per_sec = 1000000; break;
    case BSQL_TIMESTAMP_NANOSECONDS: per_sec = 1000000000; break;
    default: break;
  }
  int64_t days, sod = 0;
  if (type == BSQL_TIMESTAMP_DAYS) {
    days = ticks;
  } else {
    const int64_t secs = floor_div(ticks, per_sec);
    days = floor_div(secs, 86400);
    sod = secs - days * 86400;
  }
  int64_t y;
  int m, d;
  civil(days, &y, &m, &d);
  put_padded(o, y, 4);
  o.put('-');
  put_padded(o, m, 2);
  o.put('-');
  put_padded(o, d, 2);
  if (type == BSQL_TIMESTAMP_DAYS) return;
  o.put(' ');
  put_padded(o, sod / 3600, 2);
  o.put(':');
  put_padded(o, (sod / 60) % 60, 2);
  o.put(':');
  put_padded(o, sod % 60, 2);
}

// A string argument of row `row`: bytes and validity.
__device__ __forceinline__ bool str_arg(const StrSrc& a, int64_t row, const uint8_t** s, int32_t* n) {
  if (a.kind == bsql::SRC_LITERAL) {
    *s = a.lit;
    *n = a.lit_len;
    return true;
  }
  if (a.kind == bsql::SRC_NULL) return false;
  const int32_t b = a.offsets[row];
  *s = a.chars + b;
  *n = a.offsets[row + 1] - b;
  return bit_at(a.valid, a.vbit, row);
}

__device__ bool fixed_valid(const StrSrc& a, int64_t row) { return bit_at(a.valid, a.vbit, row); }

template <class T>
__device__ __forceinline__ T fixed_at(const StrSrc& a, int64_t row) {
  return ((const T*)a.data)[row];
}

// ---- regular expressions (the program of regex_compile, in regex.cpp)

constexpr int RX_STACK = 256;       // backtracking frames per lane
constexpr int RX_STEPS = 1 << 20;   // matcher steps per row (over all start positions)
constexpr int32_t RX_RANGE = 1 << 30;  // frame tag: a greedy one-character loop's positions
constexpr int32_t RX_LOOP = 1 << 29;   // frame tag: the body of a lazy loop, entered on backtracking
constexpr int32_t RX_LOOP_SPLIT = 128;  // edge condition bit: the split of an unbounded loop
constexpr int RX_MAX_STATES = 128;

__device__ bool rx_in_class(const int32_t* prog, int32_t at, uint32_t cp) {
  const int32_t nr = prog[at];
  bool in = false;
  for (int32_t k = 0; k < nr && !in; ++k) in = cp >= (uint32_t)prog[at + 2 + 2 * k] && cp <= (uint32_t)prog[at + 3 + 2 * k];
  return prog[at + 1] ? !in : in;
}

// Whether the one-character state k accepts the character at s[pos] (pos < n);
// its byte length in *l.
__device__ __forceinline__ bool rx_accepts(const int32_t* prog, uint32_t k, const uint8_t* s, int32_t pos,
                                           int32_t n, int32_t* l) {
  *l = char_bytes(s[pos]) <= n - pos ? char_bytes(s[pos]) : n - pos;
  const uint32_t cp = code_point(s + pos, *l);
  const uint32_t kk = k & 7u;
  return kk == 0 ? cp == (k >> 3) : (kk == 1 ? cp != '\n' : rx_in_class(prog, (int32_t)(k >> 3), cp));
}

// Leftmost-first (backtracking, first alternative / greedy or lazy order as
// compiled) match of the regex starting exactly at byte `at` of s[0, n); on
// success caps[0..1] is the match and caps[2g..2g+1] group g (-1: unset).
// A greedy loop over one character state runs as far as it can and leaves
// one frame for all its shorter lengths. An iteration of an unbounded loop
// that consumed nothing is not repeated: the match goes on after the loop
// with that iteration's captures (Python's rule). nonempty: a match that
// consumes nothing does not count (the search goes on to the next choice).
// Returns 1 (match), 0 (no match) or -1 when the match needs more than
// RX_STACK pending choices or the row's remaining *steps.
__device__ int rx_match_at(const uint8_t* s, int32_t n, int32_t at, const int32_t* prog, int32_t* caps,
                           int32_t* steps, bool nonempty) {
  const int32_t G = prog[3];
  const int32_t* kind = prog + 4;
  const int32_t* edges = prog + prog[2];
  int32_t stk_a[RX_STACK], stk_b[RX_STACK], stk_c[RX_STACK];
  int32_t iter_start[RX_MAX_STATES];  // an unbounded loop's current iteration: its start (by split state)
  int sp = 0;
  for (int g = 0; g < 2 * (G + 1); ++g) caps[g] = -1;
  caps[0] = at;
  int32_t pc = 0, pos = at;
  // enters the body of the loop whose split is `split` at pos (its
  // iteration start kept, restored on backtracking)
  auto enter_body = [&](int32_t split) -> bool {
    if (sp >= RX_STACK) return false;
    stk_a[sp] = -(64 + split + 1);
    stk_b[sp] = iter_start[split];
    ++sp;
    iter_start[split] = pos;
    pc = split + 1;
    return true;
  };
  while (true) {
    if (--*steps < 0) return -1;
    bool ok = true;
    const uint32_t k = (uint32_t)kind[pc];
    const uint32_t kk = k & 7u;
    if (kk <= 2) {
      int32_t l = 0;
      ok = pos < n && rx_accepts(prog, k, s, pos, n, &l);
      if (ok) {
        pos += l;
        ++pc;
      }
    } else if (kk == 3) {
      if (nonempty && pos == at) {
        ok = false;
      } else {
        caps[1] = pos;
        return 1;
      }
    } else {
      const uint32_t e0 = (uint32_t)edges[2 * pc], e1 = (uint32_t)edges[2 * pc + 1];
      const uint32_t cond = e1 & 3u;
      const bool loop = (e1 & (uint32_t)RX_LOOP_SPLIT) != 0;
      const int slot = (int)(e1 >> 8) - 1;
      const uint32_t x = e0 & 0xFFFFu, y = e0 >> 16;
      if ((cond == 1 && pos != 0) || (cond == 2 && pos != n)) {
        ok = false;
      } else if (x < (uint32_t)pc && y == 0xFFFFu && slot < 0 &&
                 ((uint32_t)edges[2 * x + 1] & (uint32_t)RX_LOOP_SPLIT) && iter_start[x] == pos) {
        // the jump back of a loop whose iteration consumed nothing: leave it
        ++pc;
      } else {
        if (slot >= 0) {
          if (sp >= RX_STACK) return -1;
          stk_a[sp] = -(slot + 1);
          stk_b[sp] = caps[slot];
          ++sp;
          caps[slot] = pos;
        }
        // greedy loop over one character state: split -> (body at +1, out),
        // body -> jump back to the split
        const bool run = slot < 0 && cond == 0 && y != 0xFFFFu && x == (uint32_t)pc + 1 &&
                         ((uint32_t)kind[pc + 1] & 7u) <= 2 && ((uint32_t)kind[pc + 2] & 7u) == 4 &&
                         ((uint32_t)edges[2 * (pc + 2)] & 0xFFFFu) == (uint32_t)pc &&
                         ((uint32_t)edges[2 * (pc + 2)] >> 16) == 0xFFFFu && edges[2 * (pc + 2) + 1] == 0;
        if (run) {
          const uint32_t body = (uint32_t)kind[pc + 1];
          const int32_t from = pos;
          int32_t l = 0;
          while (pos < n && rx_accepts(prog, body, s, pos, n, &l)) pos += l;
          if (pos > from) {  // the shorter lengths, tried longest first on backtracking
            if (sp >= RX_STACK) return -1;
            stk_a[sp] = RX_RANGE | (int32_t)y;
            stk_b[sp] = pos;
            stk_c[sp] = from;
            ++sp;
          }
          pc = (int32_t)y;
        } else {
          if (y != 0xFFFFu) {
            if (sp >= RX_STACK) return -1;
            // the second choice of a lazy loop enters its body
            stk_a[sp] = (loop && y == (uint32_t)pc + 1) ? (RX_LOOP | pc) : (int32_t)y;
            stk_b[sp] = pos;
            ++sp;
          }
          if (loop && x == (uint32_t)pc + 1) {
            if (!enter_body(pc)) return -1;
          } else {
            pc = (int32_t)x;
          }
        }
      }
    }
    if (!ok) {
      while (true) {
        if (sp == 0) return 0;
        const int32_t a = stk_a[sp - 1];
        if (a < 0) {  // restore a capture slot or a loop's iteration start
          if (a <= -65) iter_start[-a - 65] = stk_b[sp - 1];
          else caps[-a - 1] = stk_b[sp - 1];
          --sp;
          continue;
        }
        if (a & RX_RANGE) {  // one character shorter, the frame kept while more remain
          int32_t p = stk_b[sp - 1] - 1;
          while (p > stk_c[sp - 1] && (s[p] & 0xC0) == 0x80) --p;
          if (p <= stk_c[sp - 1]) --sp; else stk_b[sp - 1] = p;
          pc = a & ~RX_RANGE;
          pos = p;
          break;
        }
        --sp;
        pos = stk_b[sp];
        if (a & RX_LOOP) {
          if (!enter_body(a & ~RX_LOOP)) return -1;
        } else {
          pc = a;
        }
        break;
      }
    }
  }
}

// cudf::strings::replace_with_backrefs over s[0, n): matches left to right,
// each replaced by `repl` with \1..\9 the groups (an unset group is empty).
// Empty matches are replaced too; after an empty match at p the next match
// may start at p only if it is not empty (the rule of Python's re.sub,
// whose results the oracle holds). False when a match needs more than the
// matcher's frames or the row's steps: the row's result is null.
__device__ bool rx_replace(Out& o, const uint8_t* s, int32_t n, const int32_t* prog, const uint8_t* repl,
                           int32_t rn) {
  const int32_t G = prog[3];
  int32_t caps[20];
  int32_t steps = RX_STEPS;
  int32_t i = 0;              // copied up to here
  int32_t from = 0;           // the next search starts here
  bool must_advance = false;  // the last match was empty and ended at `from`
  while (from <= n) {
    int32_t at = from;
    bool found = false;
    while (at <= n) {
      const int r = rx_match_at(s, n, at, prog, caps, &steps, must_advance && at == from);
      if (r < 0) return false;
      if (r > 0) {
        found = true;
        break;
      }
      if (at == n) break;
      at += char_bytes(s[at]) <= n - at ? char_bytes(s[at]) : n - at;
    }
    if (!found) break;
    o.put(s + i, at - i);
    for (int32_t k = 0; k < rn; ++k) {
      const uint8_t c = repl[k];
      if (c == '\\' && k + 1 < rn && repl[k + 1] >= '1' && repl[k + 1] <= '9') {
        const int g = repl[k + 1] - '0';
        if (g <= G && caps[2 * g] >= 0 && caps[2 * g + 1] >= caps[2 * g])
          o.put(s + caps[2 * g], caps[2 * g + 1] - caps[2 * g]);
        ++k;
      } else {
        o.put(c);
      }
    }
    const int32_t e = caps[1];
    i = e;
    from = e;
    must_advance = e == at;
  }
  if (i < n) o.put(s + i, n - i);
  return true;
}

// Row `row` of the result into o; false for a null row.
__device__ bool build_row(const StrFnArgs& f, int64_t row, Out& o) {
  const uint8_t* s;
  int32_t n;
  switch (f.op) {
    case bsql::SF_FILL:
    case bsql::SF_CONCAT: {
      const uint8_t* t;
      int32_t m;
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      if (f.op == bsql::SF_FILL) {
        o.put(s, n);
        return true;
      }
      if (!str_arg(f.a[1], row, &t, &m)) return false;
      o.put(s, n);
      o.put(t, m);
      return true;
    }
    case bsql::SF_UPPER:
    case bsql::SF_LOWER: {
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      const uint64_t* table = f.op == bsql::SF_UPPER ? CASE_UPPER : CASE_LOWER;
      const int tn = f.op == bsql::SF_UPPER ? BSQL_CASE_UPPER_N : BSQL_CASE_LOWER_N;
      for (int32_t i = 0; i < n;) {
        int32_t l = char_bytes(s[i]);
        if (i + l > n) l = n - i;
        const uint32_t cp = code_point(s + i, l);
        const uint64_t e = case_entry(table, tn, cp);
        if (e) put_mapped(o, e); else o.put(s + i, l);
        i += l;
      }
      return true;
    }
    case bsql::SF_INITCAP: {
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      bool capitalize = true;
      for (int32_t i = 0; i < n;) {
        int32_t l = char_bytes(s[i]);
        if (i + l > n) l = n - i;
        const uint32_t cp = code_point(s + i, l);
        const bool alpha = is_alpha(cp);
        uint64_t e = 0;
        if (alpha)
          e = capitalize ? case_entry(CASE_UPPER, BSQL_CASE_UPPER_N, cp)
                         : case_entry(CASE_LOWER, BSQL_CASE_LOWER_N, cp);
        if (e) put_mapped(o, e); else o.put(s + i, l);
        capitalize = !alpha;
        i += l;
      }
      return true;
    }
    case bsql::SF_REVERSE: {
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      int32_t end = n;
      while (end > 0) {
        int32_t st = end - 1;
        while (st > 0 && (s[st] & 0xC0) == 0x80) --st;
        o.put(s + st, end - st);
        end = st;
      }
      return true;
    }
    case bsql::SF_REPLACE: {
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      const uint8_t* t = f.a[1].lit;
      const int32_t tn = f.a[1].lit_len;
      const uint8_t* r = f.a[2].lit;
      const int32_t rn = f.a[2].lit_len;
      for (int32_t i = 0; i < n;) {
        bool hit = tn > 0 && i + tn <= n;
        for (int32_t k = 0; hit && k < tn; ++k) hit = s[i + k] == t[k];
        if (hit) {
          o.put(r, rn);
          i += tn;
        } else {
          o.put(s[i]);
          ++i;
        }
      }
      return true;
    }
    case bsql::SF_REGEXP_REPLACE: {
      if (!str_arg(f.a[0], row, &s, &n)) return false;
      // the first `start` characters are kept as they are (slice_strings)
      int32_t b = 0;
      for (int32_t c = 0; c < f.start && b < n; ++c) b += char_bytes(s[b]) <= n - b ? char_bytes(s[b]) : n - b;
      o.put(s, b);
      return rx_replace(o, s + b, n - b, (const int32_t*)f.a[1].lit, f.a[2].lit, f.a[2].lit_len);
    }
    case bsql::SF_CAST: {
      const StrSrc& a = f.a[0];
      if (a.kind == bsql::SRC_STRING) {
        if (!str_arg(a, row, &s, &n)) return false;
        o.put(s, n);
        return true;
      }
      if (!fixed_valid(a, row)) return false;
      switch (a.type) {
        case BSQL_BOOL8:
          if (fixed_at<uint8_t>(a, row)) o.put((const uint8_t*)"true", 4);
          else o.put((const uint8_t*)"false", 5);
          return true;
        case BSQL_INT8: put_padded(o, fixed_at<int8_t>(a, row), 1); return true;
        case BSQL_INT16: put_padded(o, fixed_at<int16_t>(a, row), 1); return true;
        case BSQL_INT32: put_padded(o, fixed_at<int32_t>(a, row), 1); return true;
        case BSQL_INT64: {
          const int64_t v = fixed_at<int64_t>(a, row);
          if (v < 0) {
            o.put('-');
            put_uint(o, 0ULL - (uint64_t)v);
          } else {
            put_uint(o, (uint64_t)v);
          }
          return true;
        }
        case BSQL_UINT8: put_uint(o, fixed_at<uint8_t>(a, row)); return true;
        case BSQL_UINT16: put_uint(o, fixed_at<uint16_t>(a, row)); return true;
        case BSQL_UINT32: put_uint(o, fixed_at<uint32_t>(a, row)); return true;
        case BSQL_UINT64: put_uint(o, fixed_at<uint64_t>(a, row)); return true;
        case BSQL_FLOAT32: put_float(o, (double)fixed_at<float>(a, row)); return true;
        case BSQL_FLOAT64: put_float(o, fixed_at<double>(a, row)); return true;
        case BSQL_TIMESTAMP_DAYS: put_timestamp(o, a.type, fixed_at<int32_t>(a, row)); return true;
        default: put_timestamp(o, a.type, fixed_at<int64_t>(a, row)); return true;
      }
    }
    case bsql::SF_CASE: {
      const StrSrc& c = f.a[0];
      const bool pick = fixed_valid(c, row) && fixed_at<uint8_t>(c, row) != 0;
      if (!str_arg(f.a[pick ? 1 : 2], row, &s, &n)) return false;
      o.put(s, n);
      return true;
    }
    default:
      return false;
  }
}

// Length pass: lens[row] = length << 32, validity words by ballot.
__global__ __launch_bounds__(SF_BLOCK) void strfn_lengths(StrFnArgs f) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * SF_BLOCK;
  for (int64_t base = (int64_t)blockIdx.x * SF_BLOCK + (threadIdx.x & ~63); base < f.num_rows; base += stride) {
    const int64_t row = base + lane;
    bool ok = false;
    if (row < f.num_rows) {
      Out o{nullptr, 0};
      ok = build_row(f, row, o);
      f.lens[row] = ok ? ((uint64_t)(uint32_t)o.n << 32) : 0ULL;
    }
    const uint64_t bits = __ballot(ok);
    if (lane == 0) f.out_valid[base >> 5] = (uint32_t)bits;
    if (lane == 32) f.out_valid[(base >> 5) + 1] = (uint32_t)(bits >> 32);
  }
}

// Write pass: each valid row into out_chars at its offset.
__global__ __launch_bounds__(SF_BLOCK) void strfn_write(StrFnArgs f) {
  const int64_t stride = (int64_t)gridDim.x * SF_BLOCK;
  for (int64_t row = (int64_t)blockIdx.x * SF_BLOCK + threadIdx.x; row < f.num_rows; row += stride) {
    if (!((f.out_valid[row >> 5] >> (row & 31)) & 1u)) continue;
    const int32_t at = f.out_offsets[row];
    if ((int64_t)f.out_offsets[row + 1] > f.chars_cap) continue;  // never past the buffer
    Out o{f.out_chars + at, 0};
    build_row(f, row, o);
  }
}

}  // namespace bsql_dev

namespace bsql {

hipError_t launch_string_offsets(const uint64_t* slices, const uint32_t* valid, int32_t* offsets, int64_t count,
                                 uint64_t* status, hipStream_t stream);

static int64_t strfn_grid(int64_t rows) {
  const int64_t groups = (rows + bsql_dev::SF_BLOCK - 1) / bsql_dev::SF_BLOCK;
  return groups < 2048 ? (groups > 0 ? groups : 1) : 2048;
}

// Builds one string column: lengths + validity, offsets, chars. `lens` holds
// num_rows words and `status` the offsets scan's tile words.
hipError_t launch_strfn(StrFnArgs f, uint64_t* status, hipStream_t stream) {
  if (f.num_rows == 0)  // the offsets scan still writes offsets[0] = 0
    return launch_string_offsets(f.lens, f.out_valid, const_cast<int32_t*>(f.out_offsets), 0, status, stream);
  const unsigned grid = (unsigned)strfn_grid(f.num_rows);
  hipLaunchKernelGGL(bsql_dev::strfn_lengths, dim3(grid), dim3(bsql_dev::SF_BLOCK), 0, stream, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_string_offsets(f.lens, f.out_valid, const_cast<int32_t*>(f.out_offsets), f.num_rows, status, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bsql_dev::strfn_write, dim3(grid), dim3(bsql_dev::SF_BLOCK), 0, stream, f);
  return hipGetLastError();
}

}  // namespace bsql
