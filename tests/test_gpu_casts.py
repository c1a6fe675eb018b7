"""GPU parity for the date/time casts and cross-unit comparisons pinned in
test_oracle_casts.py: CAST_DATE and CAST_TIMESTAMP* from integers, floats
and timestamps of every unit (negative ticks included), timestamp
comparisons across units and duration comparisons on raw ticks, through
reference-encoded programs (bsql_program_from_arrays) and compiled text,
bit-exact against the oracle, with nulls."""
import numpy as np
import pytest

from tests.helpers import assert_same

pytestmark = pytest.mark.gpu

UNARY = -1
TS = [12, 13, 14, 15, 16]
DUR = [18, 19, 20, 21]
NUMS = [1, 2, 3, 4, 7, 9, 10]
CASTS = {33: 12, 34: 13, 35: 14, 36: 15, 37: 16}  # op -> output type
CMPS = [51, 52, 53, 54, 55, 56]
NP = {1: np.int8, 2: np.int16, 3: np.int32, 4: np.int64, 7: np.uint32, 9: np.float32, 10: np.float64, 12: np.int32,
      13: np.int64, 14: np.int64, 15: np.int64, 16: np.int64, 18: np.int64, 19: np.int64, 20: np.int64,
      21: np.int64, 11: np.uint8}


def _values(rng, t, n):
    if t in (9, 10):
        return rng.uniform(-1e6, 1e6, n).astype(NP[t])
    if t in (1, 2, 7):
        info = np.iinfo(NP[t])
        return rng.integers(info.min, info.max, n, dtype=np.int64).astype(NP[t])
    if t == 12:
        return rng.integers(-100000, 100000, n).astype(np.int32)
    scale = {13: 1, 14: 10**3, 15: 10**6, 16: 10**9, 3: 1, 4: 1}.get(t, 1)
    v = rng.integers(-3 * 10**9, 3 * 10**9, n).astype(np.int64) * scale
    v += rng.integers(-scale + 1, scale, n) if scale > 1 else 0
    return v.astype(NP[t])


def _run_both(bsql, prog, cols, out_type, n, what):
    import oracle
    from blazingsql_amd.interop import DeviceColumn, Program, perform_interpreter_operation
    p = Program.from_arrays(prog["left"], prog["right"], prog["out"], prog["ops"], prog["final_out"])
    dev = [DeviceColumn.from_numpy(t, v, valid) for t, v, valid in cols]
    host = [oracle.HostColumn(t, v, valid) for t, v, valid in cols]
    out = DeviceColumn.empty(out_type, n)
    perform_interpreter_operation(p, dev, [out], n)
    (ov, oval), = oracle.execute(prog, host, [out_type], n)
    gv, gval = out.to_numpy(n)
    assert_same(gv, gval, ov, oval, what=what)


@pytest.mark.parametrize("src", NUMS + TS)
def test_casts_vs_oracle(bsql, src):
    rng = np.random.default_rng(src)
    n = 6_001
    cols = [(src, _values(rng, src, n), rng.random(n) >= 0.1)]
    for op, out in CASTS.items():
        prog = {"left": [0], "right": [UNARY], "out": [1], "ops": [op], "final_out": [1]}
        _run_both(bsql, prog, cols, out, n, what=f"cast op {op} from type {src}")


@pytest.mark.parametrize("lt,rt", [(12, 13), (13, 14), (14, 15), (15, 16), (16, 13), (12, 16), (14, 14),
                                   (18, 19), (19, 21), (13, 4)])
def test_cross_unit_comparisons_vs_oracle(bsql, lt, rt):
    rng = np.random.default_rng(lt * 40 + rt)
    n = 6_001
    a = _values(rng, lt, n)
    b = _values(rng, rt, n)
    # equal instants in a third of the rows
    tick = {12: 86400 * 10**9, 13: 10**9, 14: 10**6, 15: 10**3, 16: 1}
    if lt in tick and rt in tick:
        same = rng.random(n) < 0.33
        ns = a.astype(np.int64) * tick[lt]
        b = np.where(same, ns // tick[rt], b).astype(NP[rt])
    cols = [(lt, a, rng.random(n) >= 0.05), (rt, b, rng.random(n) >= 0.05)]
    for op in CMPS:
        prog = {"left": [0], "right": [1], "out": [2], "ops": [op], "final_out": [2]}
        _run_both(bsql, prog, cols, 11, n, what=f"compare {op} of types {lt}, {rt}")


def test_casts_through_compiled_text(bsql):
    """The same casts written as Calcite text, with the literal forms the
    front end types (CAST(... ):DATE, timestamp literals in comparisons)."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    rng = np.random.default_rng(3)
    n = 4_001
    types = [14, 12, 4]
    table = [(t, _values(rng, t, n), rng.random(n) >= 0.1) for t in types]
    exprs = ["CAST($0):DATE", "CAST($1):TIMESTAMP", "CAST($2):DATE", ">($0, 1998-09-02 00:00:00)",
             "<=($1, 1995-01-01)", "=(CAST($0):DATE, $1)", "CAST($0):TIMESTAMP_SECONDS"]
    p = bsql.Program.compile(exprs, types)
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    dev = [DeviceColumn.from_numpy(t, v, val) for t, v, val in table]
    host = [oracle.HostColumn(t, v, val) for t, v, val in table]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    perform_interpreter_operation(p, dev, outs, n)
    ora = oracle.evaluate(p, host, n)
    for k in range(len(ot)):
        gv, gval = outs[k].to_numpy(n)
        assert_same(gv, gval, ora[k][0], ora[k][1], what=exprs[k])
