"""Pins the oracle's date/time casts and cross-unit comparisons
(engine/src/Interpreter/interpreter_ops.cuh:623-719, 836-909) against what
fixes them independently of the restatement:

* CAST_DATE of an integer or float is static_cast<int64_t> taken as days;
  of a timestamp it is chrono::time_point_cast<days>, which truncates
  toward zero (duration_cast): known answers at whole days and one tick on
  either side, positive and negative, and Python's datetime for
  non-negative ticks;
* CAST_TIMESTAMP, _SECONDS, _MILLISECONDS and _MICROSECONDS all produce a
  nanosecond count (the output column keeps its coarser unit): checked
  against datetime/timedelta arithmetic in exact integers;
* timestamps of any two units compare as nanoseconds: checked against
  Python datetime comparisons of the same instants;
* durations compare their raw tick counts (the ns conversion the reference
  computes is unused): DURATION_SECONDS 1 < DURATION_MILLISECONDS 500.
CPU only."""
import datetime

import numpy as np
import pytest

UNARY = -1
TS_D, TS_S, TS_MS, TS_US, TS_NS = 12, 13, 14, 15, 16
DUR_D, DUR_S, DUR_MS, DUR_US, DUR_NS = 17, 18, 19, 20, 21
INT8, INT16, INT32, INT64, UINT32, FLOAT32, FLOAT64, BOOL8 = 1, 2, 3, 4, 7, 9, 10, 11
CAST_DATE, CAST_TS_S, CAST_TS_MS, CAST_TS_US, CAST_TS = 33, 34, 35, 36, 37
EQ, NE, LT, GT, LE, GE = 51, 52, 53, 54, 55, 56
TICKS_PER_S = {TS_S: 1, TS_MS: 10**3, TS_US: 10**6, TS_NS: 10**9, DUR_S: 1, DUR_MS: 10**3, DUR_US: 10**6,
               DUR_NS: 10**9}
EPOCH = datetime.datetime(1970, 1, 1)


def _unary(op):
    return {"left": [0], "right": [UNARY], "out": [1], "ops": [op], "final_out": [1]}


def _binary(op):
    return {"left": [0], "right": [1], "out": [2], "ops": [op], "final_out": [2]}


def _run(prog, cols, out_type):
    import oracle
    (v, valid), = oracle.execute(prog, cols, [out_type], cols[0].size)
    assert valid.all()
    return [int(x) for x in v]


def _trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


@pytest.mark.parametrize("t", [TS_S, TS_MS, TS_US, TS_NS])
def test_cast_date_from_timestamps_truncates_toward_zero(oracle_lib, t):
    from oracle import HostColumn
    day = 86400 * TICKS_PER_S[t]
    ticks = [0, 1, -1, day - 1, day, day + 1, -day + 1, -day, -day - 1, 37 * day + 5, -37 * day - 5]
    got = _run(_unary(CAST_DATE), [HostColumn(t, np.asarray(ticks, dtype=np.int64))], TS_D)
    known = [0, 0, 0, 0, 1, 1, 0, -1, -1, 37, -37]
    assert got == known
    # non-negative instants: the calendar date Python computes
    rng = np.random.default_rng(t)
    secs = rng.integers(0, 4 * 10**9, 200)
    ticks = [int(s) * TICKS_PER_S[t] + int(rng.integers(0, TICKS_PER_S[t])) for s in secs]
    got = _run(_unary(CAST_DATE), [HostColumn(t, np.asarray(ticks, dtype=np.int64))], TS_D)
    want = [((EPOCH + datetime.timedelta(seconds=_trunc_div(x, TICKS_PER_S[t]))).date() - EPOCH.date()).days
            for x in ticks]
    assert got == want


@pytest.mark.parametrize("t", [INT8, INT16, INT32, INT64, UINT32, TS_D])
def test_cast_date_from_integers_is_the_value(oracle_lib, t):
    from oracle import HostColumn
    vals = {INT8: [-128, -1, 0, 1, 127], INT16: [-32768, -5, 0, 9, 32767], INT32: [-700000, -1, 0, 1, 2**31 - 1],
            INT64: [-10**6, -1, 0, 1, 10**6], UINT32: [0, 1, 5, 2**31, 2**32 - 1], TS_D: [-719468, -1, 0, 1, 2932896]}[t]
    dt = {INT8: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64, UINT32: np.uint32, TS_D: np.int32}[t]
    got = _run(_unary(CAST_DATE), [HostColumn(t, np.asarray(vals, dtype=dt))], TS_D)
    # the int64 value narrowed to the TIMESTAMP_DAYS column (int32) on write
    assert got == [int(np.int64(v).astype(np.int32)) for v in vals]


@pytest.mark.parametrize("t", [FLOAT32, FLOAT64])
def test_cast_date_from_floats_truncates(oracle_lib, t):
    from oracle import HostColumn
    vals = [0.0, 0.9, -0.9, 1.5, -1.5, 12345.75, -12345.75]
    got = _run(_unary(CAST_DATE), [HostColumn(t, np.asarray(vals))], TS_D)
    assert got == [0, 0, 0, 1, -1, 12345, -12345]


@pytest.mark.parametrize("op,out", [(CAST_TS, TS_NS), (CAST_TS_S, TS_S), (CAST_TS_MS, TS_MS), (CAST_TS_US, TS_US)])
@pytest.mark.parametrize("t", [TS_D, TS_S, TS_MS, TS_US, TS_NS])
def test_cast_timestamp_counts_nanoseconds(oracle_lib, op, out, t):
    """Every CAST_TIMESTAMP variant yields the instant in nanoseconds, whatever
    unit its output column has (interpreter_ops.cuh:873-909)."""
    from oracle import HostColumn
    rng = np.random.default_rng(op * 100 + t)
    if t == TS_D:
        ticks = [0, 1, -1, 18000, -25567] + [int(x) for x in rng.integers(-100000, 100000, 50)]
        want = [d * 86400 * 10**9 for d in ticks]
        # cross-check the arithmetic with datetime for a few in range
        for d in ticks[:5]:
            inst = EPOCH + datetime.timedelta(days=d)
            assert (inst - EPOCH) // datetime.timedelta(microseconds=1) * 1000 == d * 86400 * 10**9
    else:
        ticks = [0, 1, -1] + [int(x) for x in rng.integers(-10**9, 10**9, 50)]
        want = [x * (10**9 // TICKS_PER_S[t]) for x in ticks]
    dt = np.int32 if t == TS_D else np.int64
    got = _run(_unary(op), [HostColumn(t, np.asarray(ticks, dtype=dt))], out)
    assert got == want


def test_cast_timestamp_from_numbers_is_nanoseconds(oracle_lib):
    from oracle import HostColumn
    got = _run(_unary(CAST_TS_S), [HostColumn(INT64, np.asarray([5, -7, 10**12], dtype=np.int64))], TS_S)
    assert got == [5, -7, 10**12]
    got = _run(_unary(CAST_TS), [HostColumn(FLOAT64, np.asarray([2.9, -2.9, 1e9]))], TS_NS)
    assert got == [2, -2, 10**9]


def _instants(rng, n):
    secs = rng.integers(-2 * 10**9, 2 * 10**9, n)
    return [int(s) for s in secs]


@pytest.mark.parametrize("lt,rt", [(TS_S, TS_MS), (TS_MS, TS_US), (TS_D, TS_S), (TS_US, TS_NS), (TS_D, TS_NS),
                                   (TS_NS, TS_S)])
def test_timestamp_comparisons_across_units(oracle_lib, lt, rt):
    """Timestamps compare as instants (both sides in nanoseconds): the same
    answers Python's datetime gives."""
    from oracle import HostColumn
    rng = np.random.default_rng(lt * 31 + rt)
    n = 300

    def ticks(t, secs):
        return secs // 86400 if t == TS_D else secs * TICKS_PER_S[t]

    a = _instants(rng, n)
    b = [x + int(rng.integers(-2, 3)) * (86400 if TS_D in (lt, rt) else 1) if rng.random() < 0.5 else y
         for x, y in zip(a, _instants(rng, n))]
    la = [ticks(lt, x) for x in a]
    rb = [ticks(rt, x) for x in b]
    dt = lambda t: np.int32 if t == TS_D else np.int64
    cols = [HostColumn(lt, np.asarray(la, dtype=dt(lt))), HostColumn(rt, np.asarray(rb, dtype=dt(rt)))]

    def inst(t, v):
        return EPOCH + (datetime.timedelta(days=v) if t == TS_D else datetime.timedelta(seconds=v // TICKS_PER_S[t]))

    ia = [inst(lt, v) for v in la]
    ib = [inst(rt, v) for v in rb]
    import operator
    for op, f in ((EQ, operator.eq), (NE, operator.ne), (LT, operator.lt), (GT, operator.gt), (LE, operator.le),
                  (GE, operator.ge)):
        got = _run(_binary(op), cols, BOOL8)
        assert got == [int(f(x, y)) for x, y in zip(ia, ib)], op


def test_duration_comparisons_use_raw_ticks(oracle_lib):
    """Durations compare their tick counts as stored, not the time they span
    (interpreter_ops.cuh:630-634: the ns values are computed but unused)."""
    from oracle import HostColumn
    a = HostColumn(DUR_S, np.asarray([1, 2, 500, 0], dtype=np.int64))
    b = HostColumn(DUR_MS, np.asarray([500, 2, 1, 0], dtype=np.int64))
    assert _run(_binary(LT), [a, b], BOOL8) == [1, 0, 0, 0]
    assert _run(_binary(EQ), [a, b], BOOL8) == [0, 1, 0, 1]
    assert _run(_binary(GT), [a, b], BOOL8) == [0, 0, 1, 0]
    # a timestamp against a plain integer compares the raw values too
    c = HostColumn(TS_S, np.asarray([10, 20], dtype=np.int64))
    d = HostColumn(INT64, np.asarray([10, 19], dtype=np.int64))
    assert _run(_binary(EQ), [c, d], BOOL8) == [1, 0]
