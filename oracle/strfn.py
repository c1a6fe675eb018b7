"""TEST INFRASTRUCTURE: host restatement of the string columns the reference
builds with libcudf before its interpreter runs
(engine/src/execution_kernels/LogicalProject.cpp:100-742), and of the step
order that appends them to the table (function_evaluator_transformer,
:670-731). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
use it.

Every function works row by row on Python strings (None = null) and follows
libcudf 21.08 (not in the reference tree; restated from its published
behaviour):
  CONCAT   cudf::strings::concatenate, no separator, a null row stays null
           (LogicalProject.cpp:308-353)
  UPPER / LOWER  cudf::strings::to_upper / to_lower: each character of the
           Basic Multilingual Plane by its full case mapping, every other
           character unchanged (:530-550)
  INITCAP  cudf::strings::title: a letter after a non-letter to upper case,
           a letter after a letter to lower case (:552-564)
  REVERSE  slice_strings with step -1: characters in reverse (:581-595)
  REPLACE  cudf::strings::replace: every non-overlapping occurrence, left
           to right (:132-147)
  CAST(x AS VARCHAR)  from_booleans / from_integers / from_floats /
           from_timestamps("%Y-%m-%d" for days, else "%Y-%m-%d %H:%M:%S")
           (:355-377, cast_to_str_functor :62-91)
  CASE     cudf::copy_if_else: condition valid and true -> first value
           (:612-659)
  REGEXP_REPLACE  cudf::strings::replace_with_backrefs after the first
           position - 1 characters (:148-182), as regexp_replace below
The case mappings come from Python's unicodedata; that the reference's
libcudf character tables agree with them beyond ASCII and Latin letters is
not shown by any vector the reference holds (parity unpinned there).
"""
import datetime
import math

STRING, BOOL8, FLOAT32, FLOAT64 = 23, 11, 9, 10
TS_DAYS, TS_S, TS_MS, TS_US, TS_NS = 12, 13, 14, 15, 16

CONCAT, UPPER, LOWER, INITCAP, REVERSE, REPLACE, CAST, CASE, FILL, REGEXP_REPLACE = range(1, 11)


def _map_bmp(s, fn):
    out = []
    for ch in s:
        m = fn(ch)
        out.append(m if ord(ch) <= 0xFFFF and len(m) <= 3 and all(ord(x) <= 0xFFFF for x in m) else ch)
    return "".join(out)


def upper(s):
    return _map_bmp(s, str.upper)


def lower(s):
    return _map_bmp(s, str.lower)


def initcap(s):
    out = []
    capitalize = True
    for ch in s:
        alpha = ord(ch) <= 0xFFFF and ch.isalpha()
        if alpha:
            out.append(_map_bmp(ch, str.upper) if capitalize else _map_bmp(ch, str.lower))
        else:
            out.append(ch)
        capitalize = not alpha
    return "".join(out)


def regexp_replace(s, pattern, repl, start):
    """cudf::strings::replace_with_backrefs on the characters from `start`
    (the first `start` kept, LogicalProject.cpp:148-182): every match, left
    to right (empty ones too; after an empty match at p the next match may
    start at p only if it is not empty, Python 3.7's re.sub rule), by
    `repl` with \\1..\\9 the groups (an unset or missing group is empty).
    Python's re with ASCII classes; '$' matches only at the end."""
    import re
    rx = re.compile(pattern.replace("$", r"\Z"), re.ASCII)

    def expand(m):
        out, i = [], 0
        while i < len(repl):
            if repl[i] == "\\" and i + 1 < len(repl) and repl[i + 1] in "123456789":
                g = int(repl[i + 1])
                out.append((m.group(g) or "") if g <= rx.groups else "")
                i += 2
            else:
                out.append(repl[i])
                i += 1
        return "".join(out)
    return s[:start] + rx.sub(expand, s[start:])


def float_text(v):
    """libcudf 21.08 ftos_converter: NaN / Inf / -Inf; otherwise rescale into
    [1e-4, 1e9] by the power-of-two-indexed table of 10^k (recording the
    exponent), take the integer part and 10 significant digits in all, round
    the last one half up from the leftover fraction, drop trailing zeros
    (keeping one decimal) and append e+XX / e-XX when rescaled."""
    upper10 = [10.0, 100.0, 10000.0, 1e8, 1e16, 1e32, 1e64, 1e128, 1e256]
    lower10 = [.1, .01, .0001, 1e-8, 1e-16, 1e-32, 1e-64, 1e-128, 1e-256]
    blower10 = [1.0, .1, .001, 1e-7, 1e-15, 1e-31, 1e-63, 1e-127, 1e-255]
    if math.isnan(v):
        return "NaN"
    sign = ""
    if math.copysign(1.0, v) < 0:
        sign = "-"
        v = -v
    if math.isinf(v):
        return sign + "Inf"
    exp10 = 0
    if v > 1e9:
        fx = 256
        for idx in range(8, -1, -1):
            if v >= upper10[idx]:
                v *= lower10[idx]
                exp10 += fx
            fx >>= 1
    elif 0.0 < v < 0.0001:
        fx = 256
        for idx in range(8, -1, -1):
            if v < blower10[idx]:
                v *= upper10[idx]
                exp10 -= fx
            fx >>= 1
    places, max_digits = 9, 1000000000
    integer = int(v)
    i = integer
    while i >= 10:
        places -= 1
        max_digits //= 10
        i //= 10
    rem = (v - float(integer)) * float(max_digits)
    decimal = int(rem)
    rem -= float(decimal)
    decimal += int(2.0 * rem)
    if decimal >= max_digits:
        decimal = 0
        integer += 1
        if exp10 and integer >= 10:
            exp10 += 1
            integer = 1
    while decimal % 10 == 0 and places > 0:
        decimal //= 10
        places -= 1
    text = sign + str(integer) + "." + (str(decimal).zfill(places) if places else "0")
    if exp10:
        text += "e" + ("-" if exp10 < 0 else "+") + str(abs(exp10)).zfill(2)
    return text


def timestamp_text(type_id, ticks):
    per_sec = {TS_S: 1, TS_MS: 1000, TS_US: 10**6, TS_NS: 10**9}
    if type_id == TS_DAYS:
        d = datetime.date(1970, 1, 1) + datetime.timedelta(days=int(ticks))
        return f"{d.year:04d}-{d.month:02d}-{d.day:02d}"
    secs = int(ticks) // per_sec[type_id]
    t = datetime.datetime(1970, 1, 1) + datetime.timedelta(seconds=secs)
    return f"{t.year:04d}-{t.month:02d}-{t.day:02d} {t.hour:02d}:{t.minute:02d}:{t.second:02d}"


def cast_text(type_id, v):
    if type_id == BOOL8:
        return "true" if v else "false"
    if type_id in (FLOAT32, FLOAT64):
        return float_text(float(v))
    if TS_DAYS <= type_id <= TS_NS:
        return timestamp_text(type_id, v)
    return str(int(v))


def column_values(col):
    """Row values of a HostColumn, None where null (strings decoded)."""
    n = col.size
    valid = col.valid if col.valid is not None else [True] * n
    if col.type_id == STRING:
        chars = bytes(col.chars)
        offs = col.offsets
        return [chars[offs[i]:offs[i + 1]].decode() if valid[i] else None for i in range(n)]
    return [col.values[i] if valid[i] else None for i in range(n)]


def build(op, args, n, start=0):
    """One string step over `args` (lists of row values, or ("literal", s) /
    ("null",)); returns a list of strings / None."""
    def arg(k, i):
        a = args[k]
        if isinstance(a, tuple):
            return a[1] if a[0] == "literal" else None
        return a[i]
    out = []
    for i in range(n):
        if op == FILL:
            out.append(arg(0, i))
        elif op == CONCAT:
            x, y = arg(0, i), arg(1, i)
            out.append(None if x is None or y is None else x + y)
        elif op in (UPPER, LOWER, INITCAP, REVERSE):
            x = arg(0, i)
            fn = {UPPER: upper, LOWER: lower, INITCAP: initcap, REVERSE: lambda s: s[::-1]}[op]
            out.append(None if x is None else fn(x))
        elif op == REPLACE:
            x = arg(0, i)
            out.append(None if x is None else x.replace(args[1][1], args[2][1]))
        elif op == CAST:
            t, x = args[0]
            out.append(None if x[i] is None else (x[i] if t == STRING else cast_text(t, x[i])))
        elif op == REGEXP_REPLACE:
            x = arg(0, i)
            out.append(None if x is None else regexp_replace(x, args[1][1], args[2][1], start))
        elif op == CASE:
            c = arg(0, i)
            out.append(arg(1, i) if c is not None and c else arg(2, i))
        else:
            raise ValueError(op)
    return out
