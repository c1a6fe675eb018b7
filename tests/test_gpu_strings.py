"""GPU parity for strings: LIKE, SUBSTRING/LEFT/RIGHT (as slices in the
interpreter), CHAR_LENGTH and comparisons, string projection outputs, and
string columns through filter compaction -- against the reference's golden
vectors and the oracle, bit-exact."""
import numpy as np
import pytest

from tests.helpers import FLOAT64, INT32, INT64, assert_same

pytestmark = pytest.mark.gpu

STRING, BOOL8 = 23, 11


def _run(bsql, exprs, cols, types, n):
    """Compiles and runs on the GPU; returns [(values, valid)] per output."""
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    p = bsql.Program.compile(exprs, types)
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    chars = p.output_chars(cols, n)
    outs = [DeviceColumn.empty(t, n, chars=chars) for t in ot]
    perform_interpreter_operation(p, cols, outs, n)
    return p, ot, [o.to_numpy(n) for o in outs]


def _oracle(bsql, p, host_cols, ot, n):
    import oracle
    return oracle.evaluate(p, host_cols, n)


def _same(g, o, what):
    (gv, gval), (ov, oval) = g, o
    np.testing.assert_array_equal(gval, oval, err_msg=what)
    if isinstance(gv, list):
        assert [a for a, ok in zip(gv, gval) if ok] == [b for b, ok in zip(ov, oval) if ok], what
    else:
        assert_same(gv, gval, ov, oval, what=what)


def test_golden_like(bsql, gpu, golden):
    from blazingsql_amd.interop import DeviceColumn
    g = golden["string_like"]
    col = DeviceColumn.from_strings(g["col1"])
    _, ot, res = _run(bsql, ["LIKE($0, '_')", "LIKE($0, '%o')"], [col], [STRING], len(g["col1"]))
    assert ot == [BOOL8, BOOL8]
    np.testing.assert_array_equal(res[0][0].astype(int), g["expected_col1"])
    np.testing.assert_array_equal(res[1][0].astype(int), g["expected_col2"])


def test_golden_substring(bsql, gpu, golden):
    from blazingsql_amd.interop import DeviceColumn
    g = golden["string_substring"]
    col = DeviceColumn.from_strings(g["col1"])
    _, ot, res = _run(bsql, ["SUBSTRING($0, 3, -1)"], [col], [STRING], len(g["col1"]))
    assert ot == [STRING] and res[0][0] == g["expected"] and res[0][1].all()


def test_golden_substring_computed_length(bsql, gpu, golden):
    from blazingsql_amd.interop import DeviceColumn
    g = golden["string_substring_null"]
    valid = np.asarray(g["valid"], dtype=bool)
    col = DeviceColumn.from_strings(g["col1"], valid)
    _, _, res = _run(bsql, ["SUBSTRING($0, 2, -(CHAR_LENGTH($0), 2))"], [col], [STRING], len(g["col1"]))
    vals, v = res[0]
    np.testing.assert_array_equal(v, np.asarray(g["expected_valid"], dtype=bool))
    assert [a for a, ok in zip(vals, v) if ok] == [e for e, ok in zip(g["expected"], g["expected_valid"]) if ok]


def _string_table(rng, n, alphabet, lo=0, hi=12, null_rate=0.1):
    strs = ["".join(rng.choice(alphabet, int(rng.integers(lo, hi + 1)))) for _ in range(n)]
    valid = rng.random(n) >= null_rate
    return strs, valid


@pytest.mark.parametrize("n", [1, 1000, 70_001])
def test_string_functions_vs_oracle(bsql, gpu, n):
    """Nested slices, LIKE on slices, CHAR_LENGTH, comparisons with literals and
    between columns, string outputs, multi-byte characters, nulls."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n)
    alphabet = list("abcab") + ["é", "ß", "%", "_"]
    s0, v0 = _string_table(rng, n, alphabet)
    s1, v1 = _string_table(rng, n, list("abc"), 0, 4, 0.05)
    k = rng.integers(-3, 8, n).astype(np.int32)
    kv = rng.random(n) >= 0.1
    types = [STRING, STRING, INT32]
    dev = [DeviceColumn.from_strings(s0, v0), DeviceColumn.from_strings(s1, v1),
           DeviceColumn.from_numpy(INT32, k, kv)]
    host = [oracle.HostColumn(STRING, strings=s0, valid=v0), oracle.HostColumn(STRING, strings=s1, valid=v1),
            oracle.HostColumn(INT32, k, kv)]
    exprs = ["LIKE($0, 'a%')", "LIKE($0, '%b_')", "LIKE(SUBSTRING($0, 2), '%é%')", "LIKE($0, '__')",
             "LIKE($0, '%%')", "CHAR_LENGTH($0)", "CHAR_LENGTH(SUBSTRING($0, $2))",
             "SUBSTRING($0, 2, 3)", "SUBSTRING($0, $2, 2)", "SUBSTRING($0, 2, $2)", "LEFT($0, 3)",
             "RIGHT($0, 2)", "LEFT(RIGHT($0, 4), 2)", "=(SUBSTRING($0, 1, 2), 'ab')", ">($0, $1)",
             "<=(SUBSTRING($0, 2), $1)", "$1", "SUBSTRING($1, 2)"]
    for e in exprs:
        p, ot, res = _run(bsql, [e], dev, types, n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


@pytest.mark.parametrize("ranges", [True, False])
def test_filter_with_string_columns_vs_oracle(bsql, gpu, ranges):
    """process_filter over a table with string columns: the strings are
    compacted with the rest (apply_boolean_mask), nulls kept."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    from tests.test_gpu_parity import _full_prog
    n = 150_003
    rng = np.random.default_rng(5)
    s0, v0 = _string_table(rng, n, list("abcdé"), 0, 20, 0.1)
    x = rng.integers(-1000, 1000, n).astype(np.int64)
    y = rng.uniform(-1, 1, n)
    types = [STRING, INT64, FLOAT64]
    dev = [DeviceColumn.from_strings(s0, v0), DeviceColumn.from_numpy(INT64, x), DeviceColumn.from_numpy(FLOAT64, y)]
    host = [oracle.HostColumn(STRING, strings=s0, valid=v0), oracle.HostColumn(INT64, x), oracle.HostColumn(FLOAT64, y)]
    for pred_text in ["AND(>($1, 0), <($2, 0.5:DECIMAL(2, 1)))", "LIKE($0, 'a%')",
                      "AND(>(CHAR_LENGTH($0), 5), <($1, 500))"]:
        pred = bsql.Program.compile(pred_text, types)
        pred.set_launch(ranges=ranges)
        outs = [DeviceColumn.empty(STRING, n, chars=dev[0].chars_size), DeviceColumn.empty(INT64, n, nullable=False),
                DeviceColumn.empty(FLOAT64, n, nullable=False)]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_columns(pred, dev, outs, cnt, n)
        m = int(cnt.item())
        on, ores = oracle.filter_table(_full_prog(bsql, pred_text, types), host, n)
        assert m == on, pred_text
        for k in range(3):
            _same(outs[k].to_numpy(m), ores[k], f"{pred_text} col {k}")


@pytest.mark.parametrize("lds", [False, True])
@pytest.mark.parametrize("n", [5000, 200_003])
def test_strings_math_queries_vs_oracle(bsql, gpu, n, lds):
    """The strings+math workload's four filter+project queries (LIKE prefix,
    CHAR_LENGTH, SUBSTRING equality, their AND; SIN/LN/POWER/FLOOR, a wrapping
    int64 product, int64 division by a column with zeros) against the
    oracle: bit-exact except the transcendentals (within 1 ulp of libm)."""
    import oracle
    import torch
    import bench
    from blazingsql_amd.interop import DeviceColumn, filter_project
    from tests.test_gpu_parity import _full_prog
    rng = np.random.default_rng(n)
    lens = rng.integers(4, 33, n)
    offs = np.zeros(n + 1, dtype=np.int32)
    offs[1:] = np.cumsum(lens)
    # a small alphabet so that every predicate selects rows at these sizes
    chars = rng.choice(np.frombuffer(b"abcd", dtype=np.uint8), int(offs[-1]))
    x = rng.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64, endpoint=True)
    y = (1.0 - rng.random(n)) * 1000.0
    z = (rng.integers(1, 1000, n) * (rng.random(n) >= 0.1)).astype(np.int64)
    types = bench.STR_TYPES
    dev = [DeviceColumn.from_numpy(STRING, None, chars=chars, offsets=offs), DeviceColumn.from_numpy(INT64, x),
           DeviceColumn.from_numpy(FLOAT64, y), DeviceColumn.from_numpy(INT64, z)]
    host = [oracle.HostColumn.strings_from_arrays(offs, chars), oracle.HostColumn(INT64, x),
            oracle.HostColumn(FLOAT64, y), oracle.HostColumn(INT64, z)]
    proj = bsql.Program.compile(bench.STR_PROJ, types)
    proj.set_launch(lds_regs=lds)  # the output phase's register file in LDS or in VGPRs
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    assert ot == [FLOAT64, FLOAT64, FLOAT64, FLOAT64, INT64, INT64]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for pred_text in bench.STR_PREDS:
        pred = bsql.Program.compile(pred_text, types)
        outs = [DeviceColumn.empty(t, n) for t in ot]
        filter_project(pred, proj, dev, outs, cnt, n)
        m = int(cnt.item())
        om, ores = oracle.filter_project(_full_prog(bsql, pred_text, types), _full_prog(bsql, bench.STR_PROJ, types),
                                         host, ot, n)
        assert m == om and m > 0, pred_text
        for k in range(len(ot)):
            gv, gval = outs[k].to_numpy(m)
            ov, oval = ores[k]
            np.testing.assert_array_equal(gval, oval, err_msg=f"{pred_text} out {k}")
            if k < 3:
                a, b = gv[gval], ov[oval]
                assert np.all(np.abs(a - b) <= np.spacing(np.abs(b))), f"{pred_text} out {k}"
            else:
                assert_same(gv, gval, ov, oval, what=f"{pred_text} out {k}")


def test_golden_cast_from_string(bsql, gpu, golden):
    """process_project.cpp:415-446: CAST(varchar) to INTEGER exactly, to DOUBLE
    within cudf's 4-ulp equivalence."""
    from blazingsql_amd.interop import DeviceColumn
    g = golden["cast_from_string"]
    n = len(g["col1"])
    cols = [DeviceColumn.from_strings(g["col1"]), DeviceColumn.from_strings(g["col2"])]
    _, ot, res = _run(bsql, ["CAST($0):INTEGER", "CAST($1):DOUBLE"], cols, [STRING, STRING], n)
    assert ot == [INT32, FLOAT64]
    np.testing.assert_array_equal(res[0][0], g["expected_col1"])
    exp = np.asarray(g["expected_col2"])
    assert np.abs(res[1][0].view(np.int64) - exp.view(np.int64)).max() <= 4


def _numeric_strings(rng, n):
    pieces = ["", "-", "+", "12", "7", "0", ".", "5", "e", "E-", "e+3", "x", " ", "9" * 19, "NaN", "Inf",
              "-Inf", "3.25", "1e308", "e400", "-0"]
    strs = ["".join(rng.choice(pieces, int(rng.integers(0, 4)))) for _ in range(n)]
    return strs, rng.random(n) >= 0.1


@pytest.mark.parametrize("n", [1, 1000, 70_001])
def test_trim_and_string_casts_vs_oracle(bsql, gpu, n):
    """TRIM (three sides; a space, a multi-byte set, the empty set) as string
    outputs, nested under CHAR_LENGTH, LIKE and comparisons; CAST of strings
    to every integer width, FLOAT and DOUBLE, also in arithmetic; nulls."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n + 11)
    s0, v0 = _string_table(rng, n, list("ab ") + ["é", "ß", "\t"], 0, 12, 0.1)
    s1, v1 = _numeric_strings(rng, n)
    types = [STRING, STRING]
    dev = [DeviceColumn.from_strings(s0, v0), DeviceColumn.from_strings(s1, v1)]
    host = [oracle.HostColumn(STRING, strings=s0, valid=v0), oracle.HostColumn(STRING, strings=s1, valid=v1)]
    exprs = ["TRIM(FLAG(BOTH), ' ', $0)", "TRIM(FLAG(LEADING), 'aé', $0)", "TRIM(FLAG(TRAILING), 'ßb ', $0)",
             "TRIM(FLAG(BOTH), '', $0)", "CHAR_LENGTH(TRIM(FLAG(BOTH), ' ', $0))",
             "LIKE(TRIM(FLAG(LEADING), ' ', $0), 'a%')", "=(TRIM(FLAG(BOTH), ' ', $0), 'ab')",
             "SUBSTRING(TRIM(FLAG(BOTH), ' ', $0), 2)", "TRIM(FLAG(BOTH), 'a', SUBSTRING($0, 2))",
             "CAST($1):TINYINT", "CAST($1):SMALLINT", "CAST($1):INTEGER", "CAST($1):BIGINT", "CAST($1):FLOAT",
             "CAST($1):DOUBLE", "+(CAST($1):BIGINT, 1)", "*(CAST($1):DOUBLE, 2)",
             "CAST(TRIM(FLAG(BOTH), ' ', $1)):INTEGER"]
    for e in exprs:
        p, ot, res = _run(bsql, [e], dev, types, n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


@pytest.mark.parametrize("ranges", [True, False])
def test_filter_on_trim_and_casts_vs_oracle(bsql, gpu, ranges):
    """process_filter with predicates over TRIM and string CASTs; the string
    columns compacted with the rest."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    from tests.test_gpu_parity import _full_prog
    n = 120_007
    rng = np.random.default_rng(12)
    s0, v0 = _string_table(rng, n, list("ab "), 0, 6, 0.1)
    s1, v1 = _numeric_strings(rng, n)
    types = [STRING, STRING]
    dev = [DeviceColumn.from_strings(s0, v0), DeviceColumn.from_strings(s1, v1)]
    host = [oracle.HostColumn(STRING, strings=s0, valid=v0), oracle.HostColumn(STRING, strings=s1, valid=v1)]
    for pred_text in ["=(TRIM(FLAG(BOTH), ' ', $0), 'ab')", ">(CAST($1):INTEGER, 5)",
                      "AND(<(CAST($1):DOUBLE, 100), >(CHAR_LENGTH(TRIM(FLAG(TRAILING), ' ', $0)), 2))"]:
        pred = bsql.Program.compile(pred_text, types)
        pred.set_launch(ranges=ranges)
        outs = [DeviceColumn.empty(STRING, n, chars=dev[0].chars_size),
                DeviceColumn.empty(STRING, n, chars=dev[1].chars_size)]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_columns(pred, dev, outs, cnt, n)
        m = int(cnt.item())
        on, ores = oracle.filter_table(_full_prog(bsql, pred_text, types), host, n)
        assert m == on and m > 0, pred_text
        for k in range(2):
            _same(outs[k].to_numpy(m), ores[k], f"{pred_text} col {k}")


def test_golden_string_building(bsql, gpu, golden):
    """process_project.cpp:368-486 on the GPU: || with a literal and between
    columns, CAST to VARCHAR of BOOL8 / INT32 / FLOAT64 / TIMESTAMP_SECONDS,
    string CASE and a comparison of it."""
    from blazingsql_amd.interop import DeviceColumn
    g = golden["string_concat"]
    cols = [DeviceColumn.from_strings(g["col1"]), DeviceColumn.from_strings(g["col2"])]
    _, ot, res = _run(bsql, ["||($0,'XD=D')", "||($0,$1)"], cols, [STRING, STRING], 10)
    assert ot == [STRING, STRING]
    assert res[0][0] == g["expected_col1"] and res[1][0] == g["expected_col2"] and res[0][1].all()
    g = golden["cast_to_string"]
    types = [BOOL8, INT32, FLOAT64, 13]
    cols = [DeviceColumn.from_numpy(BOOL8, np.asarray(g["col1"], dtype=np.uint8)),
            DeviceColumn.from_numpy(INT32, np.asarray(g["col2"], dtype=np.int32)),
            DeviceColumn.from_numpy(FLOAT64, np.asarray(g["col3"])),
            DeviceColumn.from_numpy(13, np.asarray(g["col4"], dtype=np.int64))]
    _, ot, res = _run(bsql, [f"CAST(${k}):VARCHAR" for k in range(4)], cols, types, 10)
    for k in range(4):
        assert res[k][0] == g[f"expected_col{k + 1}"], k
    g = golden["string_case"]
    cols = [DeviceColumn.from_numpy(INT32, np.asarray(g["col1"], dtype=np.int32)),
            DeviceColumn.from_strings(g["col2"])]
    _, ot, res = _run(bsql, ["CASE(=(MOD($0, 2), 0), $1, 'LOL')", "=(CASE(=(MOD($0, 2), 0), $1, 'LOL'), 'LOL')"],
                      cols, [INT32, STRING], 9)
    assert ot == [STRING, BOOL8] and res[0][0] == g["expected"]
    np.testing.assert_array_equal(res[1][0].astype(bool), g["nested_expected"])


def _mixed_table(rng, n):
    from blazingsql_amd.interop import DeviceColumn
    import oracle
    alphabet = list("abcXYZ ,.'-3") + ["é", "ß", "Σ", "ǅ", "ΐ", "中", "ŉ", "ı"]
    s0, v0 = _string_table(rng, n, alphabet, 0, 10, 0.1)
    s1, v1 = _string_table(rng, n, list("ab"), 0, 4, 0.1)
    k = rng.integers(-50, 50, n).astype(np.int32)
    kv = rng.random(n) >= 0.1
    x = rng.normal(0, 1, n) * 10.0 ** rng.integers(-12, 12, n)
    xv = rng.random(n) >= 0.1
    b = (rng.random(n) < 0.5).astype(np.uint8)
    ts = rng.integers(-2 * 10**12, 4 * 10**12, n).astype(np.int64)
    types = [STRING, STRING, INT32, FLOAT64, BOOL8, 14]
    dev = [DeviceColumn.from_strings(s0, v0), DeviceColumn.from_strings(s1, v1),
           DeviceColumn.from_numpy(INT32, k, kv), DeviceColumn.from_numpy(FLOAT64, x, xv),
           DeviceColumn.from_numpy(BOOL8, b), DeviceColumn.from_numpy(14, ts)]
    host = [oracle.HostColumn(STRING, strings=s0, valid=v0), oracle.HostColumn(STRING, strings=s1, valid=v1),
            oracle.HostColumn(INT32, k, kv), oracle.HostColumn(FLOAT64, x, xv), oracle.HostColumn(BOOL8, b),
            oracle.HostColumn(14, ts)]
    return types, dev, host


STRING_BUILDERS = [
    "UPPER($0)", "LOWER($0)", "INITCAP($0)", "REVERSE($0)", "REPLACE($0, 'ab', 'xYz')", "REPLACE($0, 'é', '')",
    "||($0, $1)", "||('pre-', $1)", "||($0, '')", "CAST($2):VARCHAR", "CAST($3):VARCHAR", "CAST($4):VARCHAR",
    "CAST($5):VARCHAR", "CAST(+($2, 1000000)):VARCHAR", "CAST(*($3, 2.5:DECIMAL(2, 1))):VARCHAR",
    "CASE(>($2, 0), $0, $1)", "CASE(>($2, 0), 'pos', <($2, -25), 'low', null)", "CASE(IS NULL($0), $1, 'x')",
    "UPPER(SUBSTRING($0, 2, 3))", "||(UPPER($0), CAST($2):VARCHAR)", "LIKE(UPPER($0), 'A%')",
    "CHAR_LENGTH(||($0, $1))", "=(LOWER($0), $1)", "SUBSTRING(INITCAP($0), 2)", "'literal'",
    "REVERSE(REPLACE(LOWER($0), 'a', 'aa'))",
]


@pytest.mark.parametrize("n", [1, 1000, 70_001])
def test_string_building_vs_oracle(bsql, gpu, n):
    """Every string-building function (and their nesting with interpreted ops
    and numeric expressions) on random tables with nulls, multi-byte and
    case-changing characters: GPU columns equal the oracle's."""
    rng = np.random.default_rng(n + 21)
    types, dev, host = _mixed_table(rng, n)
    for e in STRING_BUILDERS:
        p, ot, res = _run(bsql, [e], dev, types, n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


def test_filter_and_project_with_string_steps_vs_oracle(bsql, gpu):
    """Predicates whose columns are built first (process_filter with the
    gather of every table column), and fused filter+project where the
    projection's steps follow the predicate's."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns, filter_project
    from tests.test_gpu_parity import _full_prog
    n = 90_001
    rng = np.random.default_rng(31)
    types, dev, host = _mixed_table(rng, n)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for pred_text in ["LIKE(UPPER($0), 'A%')", "AND(=(CASE(>($2, 0), $1, 'b'), 'a'), >($3, 0))",
                      "<(CHAR_LENGTH(||($0, $1)), 4)"]:
        pred = bsql.Program.compile(pred_text, types)
        outs = [DeviceColumn.empty(t, n, chars=c.chars_size if t == STRING else 0) for t, c in zip(types, dev)]
        filter_columns(pred, dev, outs, cnt, n)
        m = int(cnt.item())
        host_ext = oracle.run_steps(pred, host, n)
        on, ores = oracle.filter_table(_full_prog(bsql, pred_text, types), host_ext, n)
        assert m == on and m > 0, pred_text
        for k in range(len(types)):
            _same(outs[k].to_numpy(m), ores[k], f"{pred_text} col {k}")
    proj_exprs = ["||(UPPER($0), '!')", "CAST($3):VARCHAR", "+($2, 1)", "CASE(>($2, 10), $0, 'small')"]
    proj = bsql.Program.compile(proj_exprs, types)
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    chars = proj.output_chars(dev, n)
    for pred_text in ["LIKE(UPPER($0), 'A%')", ">($2, 0)"]:
        pred = bsql.Program.compile(pred_text, types)
        outs = [DeviceColumn.empty(t, n, chars=chars) for t in ot]
        filter_project(pred, proj, dev, outs, cnt, n)
        m = int(cnt.item())
        # oracle: the filter over the table, then the projection (with its
        # steps) over the filtered rows
        host_ext = oracle.run_steps(pred, host, n)
        on, filtered = oracle.filter_table(_full_prog(bsql, pred_text, types), host_ext, n)
        assert m == on and m > 0, pred_text
        fcols = []
        for t, (vals, valid) in zip(types, filtered[:len(types)]):
            fcols.append(oracle.HostColumn(t, strings=vals, valid=valid) if t == STRING
                         else oracle.HostColumn(t, vals, valid))
        ores = _oracle(bsql, proj, fcols, ot, m)
        for k in range(len(ot)):
            _same(outs[k].to_numpy(m), ores[k], f"{pred_text} -> {proj_exprs[k]}")


def test_golden_cast_string_to_timestamp(bsql, gpu, golden):
    """process_project.cpp:415-446: CAST(varchar AS TIMESTAMP), nanoseconds."""
    from blazingsql_amd.interop import DeviceColumn
    g = golden["cast_from_string"]
    _, ot, res = _run(bsql, ["CAST($0):TIMESTAMP"], [DeviceColumn.from_strings(g["col3"])], [STRING], 10)
    assert ot == [16]
    np.testing.assert_array_equal(res[0][0], g["expected_col3"])


@pytest.mark.parametrize("n", [1, 5000, 70_001])
def test_string_to_timestamp_vs_oracle(bsql, gpu, n):
    """CAST(varchar AS DATE / TIMESTAMP_*) and TO_DATE / TO_TIMESTAMP with
    other formats, on well-formed, short and malformed strings with nulls,
    also compared and filtered on."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n + 5)
    secs = rng.integers(-9 * 10**9, 9 * 10**9, n)
    import datetime
    epoch = datetime.datetime(1970, 1, 1)
    good = [(epoch + datetime.timedelta(seconds=int(s))).strftime("%Y-%m-%d %H:%M:%S") for s in secs]
    junk = ["".join(rng.choice(list("0123456789-: x"), int(rng.integers(0, 20)))) for _ in range(n)]
    strs = [g if rng.random() < 0.7 else j for g, j in zip(good, junk)]
    valid = rng.random(n) >= 0.1
    dev = [DeviceColumn.from_strings(strs, valid)]
    host = [oracle.HostColumn(STRING, strings=strs, valid=valid)]
    for e in ["CAST($0):DATE", "CAST($0):TIMESTAMP_SECONDS", "CAST($0):TIMESTAMP_MILLISECONDS",
              "CAST($0):TIMESTAMP_MICROSECONDS", "CAST($0):TIMESTAMP", "TO_DATE($0, '%Y-%m-%d')",
              "TO_TIMESTAMP($0, '%Y-%m-%d %H:%M:%S')", "TO_DATE(SUBSTRING($0, 9), '%d')",
              ">(CAST($0):DATE, 1998-09-02)", "EXTRACT(FLAG(YEAR), CAST($0):TIMESTAMP)"]:
        p, ot, res = _run(bsql, [e], dev, [STRING], n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


@pytest.mark.parametrize("n", [1, 2000, 70_001])
def test_like_with_regex_syntax_vs_oracle(bsql, gpu, n):
    """LIKE patterns with characters that are regex syntax after the
    reference's translation ('.', '+', '*', '?', '|', '[..]', '(..)', '{n}',
    '^', '$', escapes): the device's bitset matcher against the oracle's
    ECMAScript regex, with multi-byte characters, newlines and nulls."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n + 3)
    alphabet = list("abc.x-_% ") + ["é", "\n", "1", "ß"]
    strs, valid = _string_table(rng, n, alphabet, 0, 14, 0.1)
    dev = [DeviceColumn.from_strings(strs, valid)]
    host = [oracle.HostColumn(STRING, strings=strs, valid=valid)]
    pats = ["%.", "a.b%", "%a+b%", "(ab)%", "%[ab]c%", "%b{2}%", "a|b%", "%c?d", "%(a|é)_", "[^a]%",
            "%x*x%", "%\\\\.%", "%[.-]_", "^a%", "%$", "%(a|b|c){3}%", "%[a-cé]+$"]
    for pat in pats:
        e = f"LIKE($0, '{pat}')"
        p, ot, res = _run(bsql, [e], dev, [STRING], n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


@pytest.mark.parametrize("n", [1, 3000, 40_001])
def test_regexp_replace_vs_oracle(bsql, gpu, n):
    """REGEXP_REPLACE with groups and backreferences, greedy and lazy loops,
    alternation, anchors, empty matches and a start position, on short and
    long (up to 2000 characters) strings with multi-byte characters and
    nulls; also nested in other functions and in a comparison."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n + 17)
    alphabet = list("abc.x- 12") + ["é"]
    strs, valid = _string_table(rng, n, alphabet, 0, 16, 0.1)
    for i in rng.choice(n, min(n, 20), replace=False):
        strs[i] = "".join(rng.choice(alphabet, int(rng.integers(200, 2000))))
    dev = [DeviceColumn.from_strings(strs, valid)]
    host = [oracle.HostColumn(STRING, strings=strs, valid=valid)]
    exprs = ["REGEXP_REPLACE($0, 'a', 'X')", "REGEXP_REPLACE($0, '(a)(b)', '\\2\\1')",
             "REGEXP_REPLACE($0, '[0-9]+', '#')", "REGEXP_REPLACE($0, 'x*', '-')",
             "REGEXP_REPLACE($0, 'a.*b', '_')", "REGEXP_REPLACE($0, 'a.*?b', '_')",
             "REGEXP_REPLACE($0, '(é|x)+', '\\1\\1')", "REGEXP_REPLACE($0, '^a|b$', '!')",
             "REGEXP_REPLACE($0, '(a)|(b)', '<\\1|\\2>')", "REGEXP_REPLACE($0, '\\d', 'D', 3)",
             "REGEXP_REPLACE($0, '(.)\\.', '\\1!')", "UPPER(REGEXP_REPLACE($0, 'c', 'ç'))",
             "=(REGEXP_REPLACE($0, '[^a]', ''), 'aa')", "CHAR_LENGTH(REGEXP_REPLACE($0, '(a+)', '\\1\\1'))"]
    for e in exprs:
        p, ot, res = _run(bsql, [e], dev, [STRING], n)
        ores = _oracle(bsql, p, host, ot, n)
        _same(res[0], ores[0], e)


def test_regexp_replace_loops_that_match_empty(bsql, gpu):
    """Unbounded loops whose body can match nothing ('(a*)*', '(x?)*',
    '(a|)*'), lazy and greedy, with captures: no iteration that consumed
    nothing is repeated, and the captures of that last empty iteration are
    kept (Python's rule), against the oracle (Python's re)."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    strs = ["baab", "axxb", "", "aaaa", "b", "xax", "ababcab", "aaxaab"]
    valid = np.ones(len(strs), dtype=bool)
    dev = [DeviceColumn.from_strings(strs, valid)]
    host = [oracle.HostColumn(STRING, strings=strs, valid=valid)]
    for pat, rep in [("(a*)*", "[\\1]"), ("(x?)*", "[\\1]"), ("(ab)*", "[\\1]"), ("(a|)*", "<\\1>"),
                     ("(a*?)*", "[\\1]"), ("(a*)*?b", "[\\1]"), ("((a)|(x))*", "\\1\\2\\3|")]:
        e = f"REGEXP_REPLACE($0, '{pat}', '{rep}')"
        p, ot, res = _run(bsql, [e], dev, [STRING], len(strs))
        ores = _oracle(bsql, p, host, ot, len(strs))
        _same(res[0], ores[0], e)
        assert res[0][1].all(), e


def test_regexp_replace_deep_matches(bsql, gpu):
    """A loop with an alternation inside leaves backtracking frames per
    iteration: '(a|b)*c' over runs of 40 characters matches as the oracle
    does; over a 300-character run the match needs more frames than a lane
    holds, and that row's result is null instead of a wrong string."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    strs = ["ab" * 20 + "c", "x" + "ba" * 20 + "cab", "ab" * 150 + "c", "cc", "abab"]
    valid = np.ones(len(strs), dtype=bool)
    dev = [DeviceColumn.from_strings(strs, valid)]
    host = [oracle.HostColumn(STRING, strings=strs, valid=valid)]
    e = "REGEXP_REPLACE($0, '(a|b)*c', '[\\1]')"
    p, ot, res = _run(bsql, [e], dev, [STRING], len(strs))
    ores = _oracle(bsql, p, host, ot, len(strs))
    vals, ok = res[0]
    assert list(ok) == [True, True, False, True, True]
    for i in (0, 1, 3, 4):
        assert vals[i] == ores[0][0][i], (i, vals[i], ores[0][0][i])


@pytest.mark.parametrize("n", [1, 3000, 70001])
def test_like_prefix_and_char_length_fast_paths_vs_oracle(bsql, gpu, n):
    """LIKE with a literal prefix ('ab%') or a whole literal ('ab') of up to 7
    bytes is compared in place, and CHAR_LENGTH counts lead bytes four at a
    time from aligned words: against the oracle on strings at every byte
    alignment, with multi-byte UTF-8, empty strings, nulls, and patterns
    just inside and outside the fast path (8 bytes, '_', '%' first)."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(n)
    alphabet = ["a", "b", "é", "ab", "€", "x", ""]
    words = ["".join(rng.choice(alphabet, int(rng.integers(0, 9)))) for _ in range(n)]
    valid = rng.random(n) > 0.1
    col = DeviceColumn.from_strings(words, valid)
    exprs = ["LIKE($0, 'ab%')", "LIKE($0, 'ab')", "LIKE($0, 'a')", "LIKE($0, 'é%')", "LIKE($0, 'abababa%')",
             "LIKE($0, 'abababab%')", "LIKE($0, 'a_%')", "LIKE($0, '%ab')", "LIKE($0, '')",
             "CHAR_LENGTH($0)", "CHAR_LENGTH(SUBSTRING($0, 2, 5))"]
    p = bsql.Program.compile(exprs, [STRING])
    ot = [p.output_type([STRING], k) for k in range(p.num_outputs)]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    from blazingsql_amd.interop import perform_interpreter_operation
    perform_interpreter_operation(p, [col], outs, n)
    host = [oracle.HostColumn(STRING, valid=valid, strings=words)]
    ores = oracle.evaluate(p, host, n)
    for k, e in enumerate(exprs):
        gv, gval = outs[k].to_numpy(n)
        np.testing.assert_array_equal(gval, ores[k][1], err_msg=e)
        np.testing.assert_array_equal(np.asarray(gv)[gval].astype(np.int64),
                                      np.asarray(ores[k][0])[gval].astype(np.int64), err_msg=e)
