"""Pins the CPU oracle (oracle/interp_oracle.cpp) before it is trusted.

1. Against the golden vectors of the reference's own unit tests
   (tests/golden/reference_vectors.json, transcribed from engine/tests/new-tests).
2. Against what the mathematics fixes: IEEE-754 binary64 arithmetic (numpy),
   two's-complement int64 wrap-around (numpy int64), the proleptic Gregorian
   calendar (Python datetime), and the truth tables of the reference's null
   handling (interpreter_ops.cuh:559-598, 721-727, 744-755).
"""
import datetime

import numpy as np
import pytest

from blazingsql_amd._lib import OP

pytestmark = []


def prog(left, right, out, ops, final_out, left_scalars=None, right_scalars=None):
    n = len(ops)
    return dict(left=left, right=right, out=out, ops=[OP[o] if isinstance(o, str) else o for o in ops],
                final_out=final_out, left_scalars=left_scalars or [None] * n,
                right_scalars=right_scalars or [None] * n)


def test_numeric_program(oracle_lib, golden):
    g = golden["numeric_program"]
    from oracle import HostColumn
    for name, t in g["types"].items():
        rs = [None] * 4
        rs[g["right_scalar_op"]] = (t, g["right_scalar_value"])
        p = prog(g["left"], g["right"], g["out"], g["ops"], g["final_out"], right_scalars=rs)
        cast = {"bool": lambda v: np.asarray(v) != 0}.get(name, lambda v: np.asarray(v))
        cols = [HostColumn(t, cast(g[c])) for c in ("col1", "col2", "col3")]
        (o1, v1), (o2, v2) = oracle_lib.execute(p, cols, [t, t])
        dt = o1.dtype
        exp1 = np.asarray(g["expected_col1"]).astype(np.int64)
        exp2 = np.asarray(g["expected_col2"]).astype(np.int64)
        if name == "bool":
            exp1, exp2 = exp1 != 0, exp2 != 0
        np.testing.assert_array_equal(o1, exp1.astype(dt), err_msg=name)
        np.testing.assert_array_equal(o2, exp2.astype(dt), err_msg=name)
        assert v1.all() and v2.all()


@pytest.mark.parametrize("unit", ["D", "s", "ms", "us", "ns"])
def test_day_of_week(oracle_lib, golden, unit):
    g = golden["day_of_week"]
    from oracle import HostColumn
    t = g["types"][unit]
    p = prog([0], [-1], [1], ["DAYOFWEEK"], [1])
    (o, v), = oracle_lib.execute(p, [HostColumn(t, g["inputs"][unit])], [3])
    exp = g["expected_days"] if unit == "D" else g["expected_other"]
    np.testing.assert_array_equal(o, exp)


@pytest.mark.parametrize("unit", ["D", "s", "ms", "us", "ns"])
def test_timestamp_components(oracle_lib, golden, unit):
    g = golden["timestamp_components"]
    from oracle import HostColumn
    t = g["types"][unit]
    p = prog([0] * 7, [-1] * 7, [1, 2, 3, 4, 5, 6, 7], g["ops"], [1, 2, 3, 4, 5, 6, 7])
    outs = oracle_lib.execute(p, [HostColumn(t, g["inputs"][unit])], [3] * 7)
    exp = g["expected_days"] if unit == "D" else g["expected_other"]
    for (o, v), e in zip(outs, exp):
        np.testing.assert_array_equal(o, e)


@pytest.mark.parametrize("nulls", [False, True])
def test_filter_mod(oracle_lib, golden, nulls):
    g = golden["filter_mod"]
    from oracle import HostColumn
    enc = g["reference_encoding"]
    for name, t in g["types"].items():
        p = prog(enc["left"], enc["right"], enc["out"], enc["ops"], enc["final_out"],
                 right_scalars=[tuple(s) if s else None for s in enc["right_scalars"]])
        conv = (lambda v: np.asarray(v) != 0) if name == "bool" else np.asarray
        valid = g["null_validity"] if nulls else None
        cols = [HostColumn(t, conv(g["col1"]), valid), HostColumn(t, conv(g["col2"]), valid)]
        n, res = oracle_lib.filter_table(p, cols)
        if name == "bool" and not nulls:
            # bool(r) + bool(2r) is 0 or 2: every row passes
            assert n == 40
            continue
        if name == "bool":
            assert n == 20
            continue
        assert n == 20, name
        np.testing.assert_array_equal(res[0][0], conv(g["expected_col1"]).astype(res[0][0].dtype))
        np.testing.assert_array_equal(res[1][0], conv(g["expected_col2"]).astype(res[1][0].dtype))
        assert res[0][1].all() and res[1][1].all()


def test_filter_strings(oracle_lib, golden):
    g = golden["filter_strings"]
    from oracle import HostColumn
    p = prog([0], [-2], [2], ["EQUAL"], [2], right_scalars=[(23, "bar")])
    n, res = oracle_lib.filter_table(p, [HostColumn(23, strings=g["col1"]), HostColumn(3, g["col2"])])
    assert n == 1 and res[0][0] == g["eq_expected"]["col1"] and list(res[1][0]) == g["eq_expected"]["col2"]
    valid = [bool(x) for x in g["null_validity"]]
    p = prog([0], [-2], [2], ["GREATER"], [2], right_scalars=[(23, "d")])
    n, res = oracle_lib.filter_table(p, [HostColumn(23, strings=g["col1"], valid=valid),
                                         HostColumn(3, g["col2"], valid)])
    assert res[0][0] == g["gt_expected"]["col1"] and list(res[1][0]) == g["gt_expected"]["col2"]
    assert res[0][1].all() and res[1][1].all()


def test_project_compare(oracle_lib, golden):
    g = golden["project_compare"]
    from oracle import HostColumn
    # >($0, 3) and <($2, 3) over (col1, col3): slots 0, 1 inputs; outputs 2, 3
    p = prog([0, 1], [-2, -2], [2, 3], ["GREATER", "LESS"], [2, 3], right_scalars=[(3, 3), (3, 3)])
    for t in (1, 2, 3, 4, 9, 10):
        (o1, _), (o2, _) = oracle_lib.execute(p, [HostColumn(t, g["col1"]), HostColumn(t, g["col3"])], [11, 11])
        np.testing.assert_array_equal(o1, g["expected_col1"])
        np.testing.assert_array_equal(o2, g["expected_col2"])
    (o1, _), (o2, _) = oracle_lib.execute(
        p, [HostColumn(11, np.asarray(g["col1"]) != 0), HostColumn(11, np.asarray(g["col3"]) != 0)], [11, 11])
    np.testing.assert_array_equal(o1, g["expected_bool_col1"])
    np.testing.assert_array_equal(o2, g["expected_bool_col2"])


def test_char_length(oracle_lib, golden):
    g = golden["char_length"]
    from oracle import HostColumn
    p = prog([0, 2], [-1, -2], [2, 1], ["CHAR_LENGTH", "SUB"], [1], right_scalars=[None, (3, 1)])
    (o, v), = oracle_lib.execute(p, [HostColumn(23, strings=g["col1"], valid=[bool(x) for x in g["valid"]])], [3])
    np.testing.assert_array_equal(v, np.asarray(g["expected_valid"], dtype=bool))
    np.testing.assert_array_equal(o[v], np.asarray(g["expected"])[v])


@pytest.mark.parametrize("t", [9, 10])
def test_round(oracle_lib, golden, t):
    g = golden["round"]
    from oracle import HostColumn
    # ROUND($0) gets the default INT8 0 argument (expression_tree.hpp:228-236)
    p = prog([0, 0, 0], [-2, -2, -2], [1, 2, 3], ["ROUND"] * 3, [1, 2, 3],
             right_scalars=[(1, 0), (3, 2), (3, 5)])
    dt = np.float32 if t == 9 else np.float64
    outs = oracle_lib.execute(p, [HostColumn(t, np.asarray(g["col1"], dtype=dt))], [t] * 3)
    for (o, v), e in zip(outs, g["expected"]):
        np.testing.assert_array_equal(o, np.asarray(e, dtype=dt))


def test_apply_filter_nulls(oracle_lib, golden):
    g = golden["apply_filter"]
    from oracle import HostColumn
    b = lambda k: [bool(x) for x in g[k]]
    for t in (1, 2, 3, 4, 9, 10):
        cols = [HostColumn(t, g["col1"], b("col1_valid")), HostColumn(23, strings=g["col2"], valid=b("col2_valid")),
                HostColumn(t, g["col3"], b("col3_valid")), HostColumn(11, np.asarray(g["mask"]) != 0, b("mask_valid"))]
        # the predicate is the BOOL8 mask column itself: no ops, output slot 3
        p = prog([], [], [], [], [3])
        n, res = oracle_lib.filter_table(p, cols)
        assert n == 3
        np.testing.assert_array_equal(res[0][0], g["expected_col1"])
        np.testing.assert_array_equal(res[0][1], b("expected_col1_valid"))
        assert res[1][0] == g["expected_col2"]
        np.testing.assert_array_equal(res[1][1], b("expected_col2_valid"))
        np.testing.assert_array_equal(res[2][1], b("expected_col3_valid"))
        np.testing.assert_array_equal(res[2][0][res[2][1]], np.asarray(g["expected_col3"])[b("expected_col3_valid")])


# ---------------------------------------------------------------- mathematics

def test_ieee_binary64_arithmetic(oracle_lib):
    """+ - * / of two FLOAT64 columns round once (numpy is IEEE binary64)."""
    from oracle import HostColumn
    rng = np.random.default_rng(1)
    a = rng.standard_normal(5000) * 10.0 ** rng.integers(-30, 30, 5000)
    b = rng.standard_normal(5000) * 10.0 ** rng.integers(-30, 30, 5000)
    p = prog([0, 0, 0, 0], [1, 1, 1, 1], [2, 3, 4, 5], ["ADD", "SUB", "MUL", "DIV"], [2, 3, 4, 5])
    outs = oracle_lib.execute(p, [HostColumn(10, a), HostColumn(10, b)], [10] * 4)
    with np.errstate(all="ignore"):
        for (o, v), e in zip(outs, (a + b, a - b, a * b, a / b)):
            assert np.array_equal(o.view(np.uint64), e.view(np.uint64))


def test_int64_wraparound_and_division(oracle_lib):
    from oracle import HostColumn
    rng = np.random.default_rng(2)
    a = rng.integers(-2**63, 2**63 - 1, 4000, dtype=np.int64)
    b = rng.integers(-2**63, 2**63 - 1, 4000, dtype=np.int64)
    b[::7] = 0
    b[3::11] = -1
    p = prog([0, 0, 0, 0, 0], [1, 1, 1, 1, 1], [2, 3, 4, 5, 6], ["ADD", "SUB", "MUL", "DIV", "MOD"],
             [2, 3, 4, 5, 6])
    (add, _), (sub, _), (mul, _), (div, vdiv), (mod, _) = oracle_lib.execute(
        p, [HostColumn(4, a), HostColumn(4, b)], [4] * 5)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(add, a + b)
        np.testing.assert_array_equal(sub, a - b)
        np.testing.assert_array_equal(mul, a * b)
    ok = (b != 0) & (b != -1)
    # C++ integer division truncates toward zero; % takes the dividend's sign
    xs, ys = a[ok].tolist(), b[ok].tolist()
    exp_div = np.array([(abs(x) // abs(y)) * (1 if (x >= 0) == (y > 0) else -1) for x, y in zip(xs, ys)],
                       dtype=np.int64)
    np.testing.assert_array_equal(div[ok], exp_div)
    np.testing.assert_array_equal(vdiv, b != 0)   # x / 0 is null (interpreter_ops.cuh:721-724)
    exp_mod = np.array([x - q * y for x, y, q in zip(xs, ys, exp_div.tolist())], dtype=np.int64)
    np.testing.assert_array_equal(mod[ok], exp_mod)


def test_calendar_against_datetime(oracle_lib):
    from oracle import HostColumn
    rng = np.random.default_rng(3)
    days = rng.integers(-200000, 200000, 3000)
    p = prog([0, 0, 0, 0], [-1, -1, -1, -1], [1, 2, 3, 4], ["YEAR", "MONTH", "DAY", "DAYOFWEEK"], [1, 2, 3, 4])
    (y, _), (m, _), (d, _), (w, _) = oracle_lib.execute(p, [HostColumn(12, days)], [3] * 4)
    epoch = datetime.date(1970, 1, 1)
    for i, dd in enumerate(days.tolist()):
        dt = epoch + datetime.timedelta(days=dd)
        assert (y[i], m[i], d[i], w[i]) == (dt.year, dt.month, dt.day, dt.isoweekday())


def test_null_truth_tables(oracle_lib):
    """AND/OR/IS_NOT_DISTINCT_FROM/IS_NOT_TRUE/IS_NOT_FALSE over {false, true, null}^2,
    as the reference's branches define them (interpreter_ops.cuh:575-601, 744-755)."""
    from oracle import HostColumn
    # all nine (left, right) pairs over {false, true, null}
    lvals = [0, 0, 0, 1, 1, 1, 0, 0, 0]
    lv = [1, 1, 1, 1, 1, 1, 0, 0, 0]
    rvals = [0, 1, 0, 0, 1, 0, 0, 1, 0]
    rv = [1, 1, 0, 1, 1, 0, 1, 1, 0]
    a = HostColumn(11, np.asarray(lvals) != 0, np.asarray(lv, dtype=bool))
    b = HostColumn(11, np.asarray(rvals) != 0, np.asarray(rv, dtype=bool))
    p = prog([0, 0, 0, 0, 0], [1, 1, 1, -1, -1], [2, 3, 4, 5, 6],
             ["LOGICAL_AND", "LOGICAL_OR", "IS_NOT_DISTINCT_FROM", "IS_NOT_TRUE", "IS_NOT_FALSE"],
             [2, 3, 4, 5, 6])
    (and_, vand), (or_, vor), (indf, vindf), (nt, vnt), (nf, vnf) = oracle_lib.execute(p, [a, b], [11] * 5)
    # AND: null if either side is null, else l && r
    np.testing.assert_array_equal(vand, [1, 1, 0, 1, 1, 0, 0, 0, 0])
    np.testing.assert_array_equal(and_[vand], [0, 0, 0, 1])
    # OR: both valid -> l||r; left only -> valid iff left is true (data = left);
    # right only -> valid (data = right), so OR(null, false) is a valid false
    np.testing.assert_array_equal(vor, [1, 1, 0, 1, 1, 1, 1, 1, 0])
    np.testing.assert_array_equal(or_[vor], [0, 1, 1, 1, 1, 0, 1])
    # IS NOT DISTINCT FROM: always valid; both null -> true; one null -> false
    assert vindf.all()
    np.testing.assert_array_equal(indf, [1, 0, 0, 0, 1, 0, 0, 0, 1])
    np.testing.assert_array_equal(nt, [1, 1, 1, 0, 0, 0, 1, 1, 1])
    np.testing.assert_array_equal(nf, [0, 0, 0, 1, 1, 1, 1, 1, 1])
    assert vnt.all() and vnf.all()


def test_case_magic_encoding(oracle_lib):
    """CASE(c, t, e) = FIRST_NON_MAGIC(MAGIC_IF_NOT(c, t), e) (expression_tree.hpp:200-220)."""
    from oracle import HostColumn
    c = HostColumn(11, np.array([1, 0, 1, 0]) != 0, np.array([1, 1, 0, 0], dtype=bool))
    t = HostColumn(4, np.array([10, 20, 30, 40]), np.array([1, 1, 1, 0], dtype=bool))
    e = HostColumn(4, np.array([-1, -2, -3, -4]))
    p = prog([0, 4], [1, 2], [4, 3], ["MAGIC_IF_NOT", "FIRST_NON_MAGIC"], [3])
    (o, v), = oracle_lib.execute(p, [c, t, e], [4])
    np.testing.assert_array_equal(o, [10, -2, -3, -4])
    assert v.all()


def test_type_propagation(oracle_lib):
    """Operand types follow interpreter_cpp.cu:502-550 and get_output_type."""
    p = prog([0, 3, 3], [1, -2, 2], [3, 3, 4], ["ADD", "MUL", "LESS"], [4], right_scalars=[None, (10, 2.0), None])
    lt, rt, ot = oracle_lib.op_types(p, [3, 4, 10])
    assert (lt, rt, ot) == ([3, 4, 10], [4, 10, 10], [4, 10, 11])


def _compiled(bsql, exprs, types):
    p = bsql.Program.compile(exprs, types)
    return p, [p.output_type(types, k) for k in range(p.num_outputs)]


def test_string_like(oracle_lib, golden, bsql):
    """LIKE through the reference's regex translation and search."""
    g = golden["string_like"]
    from oracle import HostColumn
    p, ot = _compiled(bsql, ["LIKE($0, '_')", "LIKE($0, '%o')"], [23])
    res = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=g["col1"])], ot)
    np.testing.assert_array_equal(res[0][0].astype(int), g["expected_col1"])
    np.testing.assert_array_equal(res[1][0].astype(int), g["expected_col2"])


def test_string_substring(oracle_lib, golden, bsql):
    g = golden["string_substring"]
    from oracle import HostColumn
    p, ot = _compiled(bsql, ["SUBSTRING($0, 3, -1)"], [23])
    (vals, valid), = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=g["col1"])], ot)
    assert ot == [23] and vals == g["expected"] and valid.all()


def test_string_substring_computed_length_null(oracle_lib, golden, bsql):
    g = golden["string_substring_null"]
    from oracle import HostColumn
    p, ot = _compiled(bsql, ["SUBSTRING($0, 2, -(CHAR_LENGTH($0), 2))"], [23])
    col = HostColumn(23, strings=g["col1"], valid=[bool(x) for x in g["valid"]])
    (vals, valid), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
    np.testing.assert_array_equal(valid, np.asarray(g["expected_valid"], dtype=bool))
    assert [v for v, ok in zip(vals, valid) if ok] == [e for e, ok in zip(g["expected"], g["expected_valid"]) if ok]


def reference_like_regex(pat):
    """like_expression_to_regex_str (LogicalProject.cpp:31-47) written with
    Python's re: the same two left-to-right replacements and anchors."""
    import re
    if not pat:
        return pat
    rx = re.sub(r"([^\\]?|\\{2})%", lambda m: m.group(1) + "(?:.*?)", pat)
    rx = re.sub(r"([^\\]?|\\{2})_", lambda m: m.group(1) + "(?:.)", rx)
    # '$' is written \Z: Python's '$' also matches before a final newline,
    # the ECMAScript '$' of the oracle's regex engine only at the end
    return ("^" if pat[0] != "%" else "") + rx + ("\\Z" if pat[-1] != "%" else "")


def test_like_and_slices_closed_forms(oracle_lib, bsql):
    """LIKE against Python's regex engine running the reference's pattern
    translation (doubled wildcards keep their quirk: '__' is '^_(?:.)$',
    '%%' finds a '%'; '.', '+', '|', classes and groups in a pattern are
    regex syntax there), on random strings with multi-byte characters and
    newlines; SUBSTRING/LEFT/RIGHT against Python slicing."""
    import re
    from oracle import HostColumn
    rng = np.random.default_rng(9)
    alphabet = list("abcd_%") + ["é", "ß", "\n"]
    strs = ["".join(rng.choice(alphabet, int(rng.integers(0, 9)))) for _ in range(600)]
    col = HostColumn(23, strings=strs)
    for pat in ["a%", "%b", "%ab%", "_a%", "a_b", "%", "%é%", "__", "a%b%c", "ß_%", "%%", "%%a", "_%_",
                "a__", "%_%", "", "%a%b", "é",
                # regex syntax keeps its meaning in the reference's translation
                "%.", "a.b%", "%a+b%", "(ab)%", "%[ab]c%", "%b{2}%", "a|b%", "%c?d", "%(a|é)_", "[^a]%"]:
        p, ot = _compiled(bsql, [f"LIKE($0, '{pat}')"], [23])
        (vals, _), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
        rx = re.compile(reference_like_regex(pat), re.ASCII)
        expect = [1 if rx.search(s) else 0 for s in strs]
        np.testing.assert_array_equal(vals.astype(int), expect, err_msg=repr(pat))
    for expr, fn in [("SUBSTRING($0, 2)", lambda s: s[1:]), ("SUBSTRING($0, 0)", lambda s: s),
                     ("SUBSTRING($0, 3, 2)", lambda s: s[2:4]), ("LEFT($0, 3)", lambda s: s[:3]),
                     ("RIGHT($0, 2)", lambda s: s[-2:] if len(s) >= 2 else s), ("RIGHT($0, 0)", lambda s: "")]:
        p, ot = _compiled(bsql, [expr], [23])
        (vals, _), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
        assert vals == [fn(s) for s in strs], expr


def test_mlp_reference_gradients_match_finite_differences():
    """The host MLP reference (the checker of the GPU gradients) against
    central differences of its own loss: a dropped term, a wrong sign or a
    transposed weight in the backward pass fails this."""
    from oracle import mlp_ref
    rng = np.random.default_rng(0)
    n_params = 24 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2
    p = rng.normal(0, 0.2, n_params)
    x = rng.normal(0, 1, (16, 24))
    t = np.stack([rng.random(16), rng.normal(0, 1, 16)], axis=1)
    g = mlp_ref.gradients(p, x, t, cost_weight=0.5)
    h = 1e-6
    for i in rng.choice(n_params, 60, replace=False):
        e = np.zeros(n_params)
        e[i] = h
        fd = (mlp_ref.loss(p + e, x, t, 0.5) - mlp_ref.loss(p - e, x, t, 0.5)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_cast_from_string_golden(oracle_lib, golden, bsql):
    """CAST(varchar) to INTEGER and DOUBLE (process_project.cpp:415-446): the
    integers exactly, the doubles within 4 ulp as cudf's equivalence check."""
    g = golden["cast_from_string"]
    from oracle import HostColumn
    p, ot = _compiled(bsql, ["CAST($0):INTEGER", "CAST($1):DOUBLE"], [23, 23])
    assert ot == [3, 10]
    (a, va), (b, vb) = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=g["col1"]),
                                                               HostColumn(23, strings=g["col2"])], ot)
    np.testing.assert_array_equal(a, g["expected_col1"])
    exp = np.asarray(g["expected_col2"])
    ulps = np.abs(b.view(np.int64) - exp.view(np.int64))
    assert va.all() and vb.all() and ulps.max() <= 4, ulps


def _int_closed_form(s, bits):
    """An optional sign and the longest run of decimal digits after it, the
    value wrapped to two's complement of `bits`."""
    import re
    m = re.match(r"([+-]?)([0-9]*)", s)
    v = int(m.group(2) or "0") * (-1 if m.group(1) == "-" else 1)
    v &= (1 << bits) - 1
    return v - (1 << bits) if v >> (bits - 1) else v


def test_cast_string_to_integers_closed_form(oracle_lib, bsql):
    """cudf::strings::to_integers: sign + digit prefix, wrapping, for every
    signed width; empty and non-numeric strings give 0; nulls stay null."""
    from oracle import HostColumn
    rng = np.random.default_rng(4)
    strs = ["", "+", "-", "0", "-0", "+17", "12abc", "abc", " 5", "3.99", "-1e3", "9" * 25,
            "-9223372036854775808", "9223372036854775808", "18446744073709551617", "-2147483649", "255", "-129"]
    strs += ["".join(rng.choice(list("0123456789+-.a "), int(rng.integers(0, 22)))) for _ in range(400)]
    valid = [i % 7 != 3 for i in range(len(strs))]
    col = HostColumn(23, strings=strs, valid=valid)
    for cast, t, bits in [("TINYINT", 1, 8), ("SMALLINT", 2, 16), ("INTEGER", 3, 32), ("BIGINT", 4, 64)]:
        p, ot = _compiled(bsql, [f"CAST($0):{cast}"], [23])
        assert ot == [t]
        (v, ok), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
        np.testing.assert_array_equal(ok, valid)
        got = [int(x) for x, k in zip(v, ok) if k]
        assert got == [_int_closed_form(s, bits) for s, k in zip(strs, valid) if k], cast


def test_cast_string_to_double_closed_forms(oracle_lib, bsql):
    """cudf::strings::to_floats. Closed forms: an integer string under 2^52
    times 10^e for 0 <= e <= 22 is one rounding of exact operands, so it is
    Python's correctly rounded float(); digits past 52 bits are truncated,
    not rounded; any decimal string lands within 1 ulp of float(); the special
    spellings are exact-match only; exponents past the double range give
    +-inf or +0 (and a zero mantissa with a large exponent infinity)."""
    import math
    from oracle import HostColumn
    rng = np.random.default_rng(5)
    exact = []
    for _ in range(300):
        d = int(rng.integers(0, 2 ** 52))
        e = int(rng.integers(0, 23))
        txt = f"{'-' if rng.random() < 0.5 else ''}{d}e{e}"
        exact.append((txt, float(txt)))
        exact.append((f"{d}", float(d)))
    fixed = [("NaN", math.nan), ("Inf", math.inf), ("-Inf", -math.inf), ("+Inf", 0.0), ("nan", 0.0),
             ("Infinity", 0.0), ("inf", 0.0), ("", 0.0), ("abc", 0.0), ("-abc", -0.0), ("1e309", math.inf),
             ("-1e309", -math.inf), ("1e-400", 0.0), ("-1e-400", 0.0), ("0e500", math.inf),
             ("1.5.5", 1.55), ("12345678901234569999", 1.234567890123456e19), ("4503599627370495", 4503599627370495.0),
             ("45035996273704951", 4503599627370495e1), ("1e", 1.0), ("2E+3", 2000.0), ("7e-0", 7.0),
             (".5", 0.5), ("-.25", -0.25), ("3.", 3.0), ("1e308", 1e308), ("12x3", 12.0)]
    approx = [f"{rng.integers(0, 10 ** 9)}.{rng.integers(0, 10 ** 6):06d}e{int(rng.integers(-300, 290))}"
              for _ in range(300)]
    strs = [s for s, _ in exact + fixed] + approx
    col = HostColumn(23, strings=strs)
    p, ot = _compiled(bsql, ["CAST($0):DOUBLE", "CAST($0):FLOAT"], [23])
    assert ot == [10, 9]
    (v, ok), (f, fok) = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
    assert ok.all() and fok.all()
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(f, v.astype(np.float32))
    n1 = len(exact) + len(fixed)
    want = np.array([x for _, x in exact + fixed])
    np.testing.assert_array_equal(v[:n1].view(np.int64)[~np.isnan(want)], want.view(np.int64)[~np.isnan(want)])
    assert np.isnan(v[:n1][np.isnan(want)]).all()
    ref = np.array([float(s) for s in approx])
    ulps = np.abs(v[n1:].view(np.int64) - ref.view(np.int64))
    assert ulps.max() <= 1, (ulps.max(), approx[int(ulps.argmax())])


def test_trim_closed_forms(oracle_lib, bsql):
    """TRIM(FLAG(BOTH|LEADING|TRAILING), chars, $0) as cudf::strings::strip
    (LogicalProject.cpp:566-579): Python's strip/lstrip/rstrip over the set of
    characters (multi-byte ones included); an empty set strips bytes <= ' '."""
    from oracle import HostColumn
    rng = np.random.default_rng(6)
    alphabet = list("ab ") + ["é", "ß", "\t", "\n", "\x01"]
    strs = ["".join(rng.choice(alphabet, int(rng.integers(0, 9)))) for _ in range(800)] + ["", " ", "é", "ab"]
    valid = [i % 11 != 5 for i in range(len(strs))]
    col = HostColumn(23, strings=strs, valid=valid)
    ws = "".join(chr(c) for c in range(33))
    for side, fn in [("BOTH", str.strip), ("LEADING", str.lstrip), ("TRAILING", str.rstrip)]:
        for chars in [" ", "ab", "é ", "ßé", ""]:
            expr = f"TRIM(FLAG({side}), '{chars}', $0)"
            p, ot = _compiled(bsql, [expr], [23])
            assert ot == [23]
            (v, ok), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
            np.testing.assert_array_equal(ok, valid)
            got = [x for x, k in zip(v, ok) if k]
            assert got == [fn(s, chars or ws) for s, k in zip(strs, valid) if k], expr


def _steps_and_execute(oracle_lib, p, cols, types, n):
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    return ot, oracle_lib.evaluate(p, cols, n)


def test_string_building_golden(oracle_lib, golden, bsql):
    """||, CAST to VARCHAR and string CASE through the steps the compiler
    extracts, against process_project.cpp:368-486."""
    from oracle import HostColumn
    g = golden["string_concat"]
    p = bsql.Program.compile(["||($0,'XD=D')", "||($0,$1)"], [23, 23])
    ot, res = _steps_and_execute(oracle_lib, p, [HostColumn(23, strings=g["col1"]), HostColumn(23, strings=g["col2"])],
                                 [23, 23], 10)
    assert ot == [23, 23] and res[0][0] == g["expected_col1"] and res[1][0] == g["expected_col2"]
    g = golden["cast_to_string"]
    types = [11, 3, 10, 13]
    cols = [HostColumn(11, np.asarray(g["col1"], dtype=np.uint8)), HostColumn(3, g["col2"]), HostColumn(10, g["col3"]),
            HostColumn(13, g["col4"])]
    p = bsql.Program.compile([f"CAST(${k}):VARCHAR" for k in range(4)], types)
    ot, res = _steps_and_execute(oracle_lib, p, cols, types, 10)
    assert ot == [23] * 4
    for k in range(4):
        assert res[k][0] == g[f"expected_col{k + 1}"] and res[k][1].all(), k
    g = golden["string_case"]
    cols = [HostColumn(3, g["col1"]), HostColumn(23, strings=g["col2"])]
    p = bsql.Program.compile(["CASE(=(MOD($0, 2), 0), $1, 'LOL')", "=(CASE(=(MOD($0, 2), 0), $1, 'LOL'), 'LOL')"],
                             [3, 23])
    ot, res = _steps_and_execute(oracle_lib, p, cols, [3, 23], 9)
    assert ot == [23, 11] and res[0][0] == g["expected"]
    np.testing.assert_array_equal(res[1][0].astype(bool), g["nested_expected"])


def test_float_text_closed_form():
    """from_floats as restated, against Python's correctly rounded formatting:
    values with at most 8 significant digits print, inside [1e-4, 1e9], in
    fixed notation with ten digits counting the integer part's, and outside
    as d.ddddddddd e+XX, trailing zeros dropped and one decimal kept; a
    dropped rounding step or a wrong rescale exponent fails this."""
    from oracle.strfn import float_text
    rng = np.random.default_rng(8)
    for _ in range(3000):
        digits = int(rng.integers(1, 9))
        m = int(rng.integers(10 ** (digits - 1), 10 ** digits))
        e = int(rng.integers(-40, 40))
        v = float(f"{'-' if rng.random() < 0.3 else ''}{m}e{e}")
        a = abs(v)
        if 1e-4 <= a <= 1e9:
            # ten digits counting the integer part's: 9 - (integer digits - 1) decimals
            places = 10 - len(str(int(a)))
            if -e > places:
                continue  # the value itself needs rounding: halves are the algorithm's own
            want = format(v, f".{places}f").rstrip("0")
            if want.endswith("."):
                want += "0"
        else:
            mant, ex = format(v, ".9e").split("e")
            mant = mant.rstrip("0")
            mant = mant + "0" if mant.endswith(".") else mant
            want = f"{mant}e{'-' if int(ex) < 0 else '+'}{abs(int(ex)):02d}"
            if mant.lstrip("-") == "1.0":
                # rescaling by inexact powers can land on 10.0 instead of 1.0
                # (the algorithm's own rounding): the same value, one exponent lower
                q = int(ex) - 1
                alt = f"{mant[:-3]}10.0e{'-' if q < 0 else '+'}{abs(q):02d}"
                if float_text(v) == alt:
                    continue
        assert float_text(v) == want, (v, float_text(v), want)
    assert [float_text(x) for x in (0.0, -0.0, float("nan"), float("inf"), float("-inf"))] == \
        ["0.0", "-0.0", "NaN", "Inf", "-Inf"]


def test_case_functions_closed_forms():
    """UPPER / LOWER / INITCAP / REVERSE / REPLACE as restated, against
    Python's whole-string methods on ASCII and Latin-1 letters (where both
    definitions agree): a mapping applied to the wrong characters, a missed
    word boundary or an overlapping replacement fails this."""
    from oracle import strfn
    rng = np.random.default_rng(9)
    alphabet = list("abcXYZ 01,.'-_") + ["é", "Ö", "ñ", "ÿ"]
    strs = ["".join(rng.choice(alphabet, int(rng.integers(0, 16)))) for _ in range(2000)]
    assert strfn.build(strfn.UPPER, [strs], len(strs)) == [s.upper() for s in strs]
    assert strfn.build(strfn.LOWER, [strs], len(strs)) == [s.lower() for s in strs]
    assert strfn.build(strfn.INITCAP, [strs], len(strs)) == [s.title() for s in strs]
    assert strfn.build(strfn.REVERSE, [strs], len(strs)) == ["".join(reversed(s)) for s in strs]
    rep = strfn.build(strfn.REPLACE, [strs, ("literal", "aa"), ("literal", "b")], len(strs))
    import re
    assert rep == [re.sub("aa", "b", s) for s in strs]
    # characters beyond the Basic Multilingual Plane are left unchanged
    assert strfn.upper("a\U00010428") == "A\U00010428"


def test_string_step_errors(bsql):
    """The reference's own rejections (LogicalProject.cpp:132-377)."""
    for expr, types, msg in [("REPLACE($0, '', 'x')", [23], "target"),
                             ("||('a', 'b')", [23], "between literals"),
                             ("UPPER('abc')", [23], "literals"),
                             ("CASE(true, 'a', 'b')", [23], "condition")]:
        with pytest.raises(Exception, match=msg):
            bsql.Program.compile([expr], types)


def test_cast_string_to_timestamp_golden(oracle_lib, golden, bsql):
    """CAST(varchar AS TIMESTAMP) of process_project.cpp:415-446 (nanoseconds)."""
    g = golden["cast_from_string"]
    from oracle import HostColumn
    p, ot = _compiled(bsql, ["CAST($0):TIMESTAMP"], [23])
    assert ot == [16]
    (v, ok), = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=g["col3"])], ot)
    np.testing.assert_array_equal(v, g["expected_col3"])
    assert ok.all()


def test_timestamp_parse_closed_forms(oracle_lib, bsql):
    """to_timestamps as restated, against Python's datetime: well-formed
    strings in the default formats and in TO_DATE / TO_TIMESTAMP formats with
    other field orders and separators, every tick type; a short string keeps
    the fields it does not reach at 1970-01-01 00:00:00."""
    import datetime
    from oracle import HostColumn
    rng = np.random.default_rng(12)
    epoch = datetime.datetime(1970, 1, 1)
    # within +-9.2e9 s, so that nanosecond ticks stay inside int64
    when = [epoch + datetime.timedelta(seconds=int(s)) for s in rng.integers(-9 * 10**9, 9 * 10**9, 400)]
    when = [w for w in when if 1 <= w.year <= 9999]
    secs = [int((w - epoch).total_seconds()) for w in when]
    days = [(w.date() - epoch.date()).days for w in when]
    full = [w.strftime("%Y-%m-%d %H:%M:%S") for w in when]
    col = HostColumn(23, strings=full)
    for cast, t, mult in [("TIMESTAMP_SECONDS", 13, 1), ("TIMESTAMP_MILLISECONDS", 14, 1000),
                          ("TIMESTAMP_MICROSECONDS", 15, 10**6), ("TIMESTAMP", 16, 10**9)]:
        p, ot = _compiled(bsql, [f"CAST($0):{cast}"], [23])
        assert ot == [t]
        (v, _), = oracle_lib.execute(p.to_oracle_dict(), [col], ot)
        assert [int(x) for x in v] == [s * mult for s in secs], cast
    p, ot = _compiled(bsql, ["CAST($0):DATE"], [23])
    (v, _), = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=[w.strftime("%Y-%m-%d") for w in when])], ot)
    assert ot == [12] and [int(x) for x in v] == days
    for fmt in ["%d/%m/%Y", "%Y%m%d", "%m-%d-%Y %H.%M.%S", "%%%Y_%d_%m"]:
        text = [w.strftime(fmt) for w in when]
        for fn, t in [("TO_DATE", 12), ("TO_TIMESTAMP", 16)]:
            p, ot = _compiled(bsql, [f"{fn}($0, '{fmt}')"], [23])
            assert ot == [t]
            (v, _), = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=text)], ot)
            parsed = [datetime.datetime.strptime(x, fmt) for x in text]
            want = [(q.date() - epoch.date()).days if t == 12 else int((q - epoch).total_seconds()) * 10**9
                    for q in parsed]
            assert [int(x) for x in v] == want, (fn, fmt)
    p, ot = _compiled(bsql, ["CAST($0):TIMESTAMP_SECONDS"], [23])
    (v, _), = oracle_lib.execute(p.to_oracle_dict(), [HostColumn(23, strings=["2001-02", "", "1999"])], ot)
    assert [int(x) for x in v] == [int((datetime.datetime(2001, 2, 1) - epoch).total_seconds()), 0,
                                   int((datetime.datetime(1999, 1, 1) - epoch).total_seconds())]
    with pytest.raises(Exception, match="not supported"):
        bsql.Program.compile(["TO_DATE($0, '%Y-%j')"], [23])


def test_regexp_replace_rules(bsql):
    """REGEXP_REPLACE(column, pattern, replacement[, position]) as the
    reference takes it (LogicalProject.cpp:148-182): occurrence and
    match_type are rejected, a position must be above zero; the pattern must
    be one the device matcher takes."""
    p = bsql.Program.compile(["REGEXP_REPLACE($0, '(a)(b)', '\\2\\1', 3)"], [23])
    st = p.steps()
    assert len(st) == 1 and st[0]["op"] == 10 and st[0]["start"] == 2
    assert st[0]["args"][1:] == [("literal", "(a)(b)"), ("literal", "\\2\\1")]
    for expr, msg in [("REGEXP_REPLACE($0, 'a', 'b', 1, 1)", "occurrence"),
                      ("REGEXP_REPLACE($0, 'a', 'b', 0)", "greater than zero"),
                      ("REGEXP_REPLACE($0, '(a', 'b')", "parenthesis"),
                      ("REGEXP_REPLACE($0, '\\1', 'b')", "not supported"),
                      ("REGEXP_REPLACE('abc', 'a', 'b')", "literals")]:
        with pytest.raises(Exception, match=msg):
            bsql.Program.compile([expr], [23])
