"""Host front-end and C ABI checks (no GPU): the Calcite text compiler must emit
the reference's program encoding, accept the reference's golden plans, and
reject what the reference rejects."""
import os
import re

import pytest

from blazingsql_amd._lib import OP, BsqlError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

INT32, INT64, FLOAT64, BOOL8, TS_D, TS_S, STRING = 3, 4, 10, 11, 12, 13, 23


def test_abi_exports_every_declared_symbol(bsql):
    from blazingsql_amd import _lib
    declared = set()
    for h in ("interp.h", "table.h"):
        header = open(os.path.join(ROOT, "include", "bsql", h)).read()
        declared |= set(re.findall(r"\b(bsql_[a-z_]+)\s*\(", header))
    lib = _lib.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == declared


def test_encoding_matches_reference_mod_filter(bsql, golden):
    enc = golden["filter_mod"]["reference_encoding"]
    p = bsql.Program.compile(golden["filter_mod"]["expression"], [INT32, INT32])
    a = p.arrays()
    assert a["left"] == enc["left"] and a["right"] == enc["right"]
    assert a["out"] == enc["out"] and a["final_out"] == enc["final_out"]
    assert a["ops"] == [OP[o] for o in enc["ops"]]
    assert p.scalar(1, 1) == (INT32, 2) and p.scalar(2, 1) == (INT32, 0)


def test_encoding_matches_reference_numeric_program(bsql, golden):
    """interops_tests.cpp:66-79: '+ * + $0 $1 $2 $1 | + $1 2' encodes as the hand-built arrays."""
    g = golden["numeric_program"]
    p = bsql.Program.compile(["+(*(+($0, $1), $2), $1)", "+($1, 2)"], [INT64] * 3)
    a = p.arrays()
    assert a["left"] == g["left"] and a["right"] == g["right"]
    assert a["out"] == g["out"] and a["final_out"] == g["final_out"]
    assert a["ops"] == [OP[o] for o in g["ops"]]
    assert p.num_inputs == 3


def test_used_columns_remap(bsql):
    """Only used columns become interpreter inputs, in table order (LogicalProject.cpp:859-881)."""
    p = bsql.Program.compile(["*($3, $1)", ">($3, 2)"], [INT32] * 5)
    assert p.used_columns() == [1, 3]
    a = p.arrays()
    # $3 -> slot 1, $1 -> slot 0; outputs at 2 and 3; temporaries from 4
    assert a["left"] == [1, 1] and a["right"] == [0, -2] and a["out"] == [2, 3]


def test_and_chain_is_left_deep_in_plan_order(bsql):
    p = bsql.Program.compile("AND(>($0, 1), <($1, 2), =($2, 3), <>($0, 4))", [INT32] * 3)
    ops = p.arrays()["ops"]
    assert ops == [OP[o] for o in ["GREATER", "LESS", "LOGICAL_AND", "EQUAL", "LOGICAL_AND",
                                   "NOT_EQUAL", "LOGICAL_AND"]]
    assert p.num_conjuncts == 4


def test_case_rewrites_to_magic_ops(bsql):
    p = bsql.Program.compile("CASE(=($0, 0), null:DOUBLE, $1)", [INT32, FLOAT64])
    a = p.arrays()
    assert a["ops"] == [OP["EQUAL"], OP["MAGIC_IF_NOT"], OP["FIRST_NON_MAGIC"]]
    assert a["right"][1] == -3  # null THEN value is SCALAR_NULL
    assert p.output_type([INT32, FLOAT64]) == FLOAT64


def test_literal_typing(bsql):
    """expression_tree.cpp:281-322: decimals are FLOAT32 iff float-exact, ints INT32 unless > INT_MAX."""
    p = bsql.Program.compile(["+($0, 0.5)", "+($0, 0.1)", "+($0, 3)", "+($0, 3000000000)",
                              "-(*($1, 2.0:DECIMAL(2, 1)), 1.0:DECIMAL(2, 1))"], [INT64, FLOAT64])
    assert p.scalar(0, 1) == (9, 0.5)
    assert p.scalar(1, 1) == (10, 0.1)
    assert p.scalar(2, 1) == (INT32, 3)
    assert p.scalar(3, 1) == (INT64, 3000000000)
    assert p.scalar(4, 1) == (FLOAT64, 2.0) and p.scalar(5, 1) == (FLOAT64, 1.0)


def test_timestamp_literal(bsql):
    p = bsql.Program.compile("<=($0, 1998-09-02 00:00:00)", [TS_D])
    t, v = p.scalar(0, 1)
    assert t == TS_S and v == 904694400  # 1998-09-02T00:00:00Z in seconds
    p = bsql.Program.compile("<($0, 1995-03-15)", [TS_D])
    assert p.scalar(0, 1) == (TS_D, 9204)


def test_unary_minus_and_extract(bsql):
    p = bsql.Program.compile(["-($0)", "EXTRACT(FLAG(YEAR), $1)"], [INT64, TS_D])
    a = p.arrays()
    assert a["ops"] == [OP["SUB"], OP["YEAR"]] and a["left"][0] == -2


def test_literal_literal_is_rejected(bsql, golden):
    with pytest.raises(BsqlError) as e:
        bsql.Program.compile(golden["errors"]["literal_literal"], [INT32])
    assert e.value.code == 2 and "literals" in str(e.value)


def test_from_arrays_validation(bsql):
    with pytest.raises(BsqlError):
        bsql.Program.from_arrays([0], [64], [5], ["ADD"], [5])  # slot index above 63
    with pytest.raises(BsqlError):
        bsql.Program.from_arrays([-2], [-1], [1], ["ABS"], [1])  # unary op on a literal


def _plan_expressions():
    text = open(os.path.join(ROOT, "tests", "golden", "calcite_plans.txt")).read()
    exprs = []
    for m in re.finditer(r"filters=\[\[(.*?)\]\]", text):
        exprs.append(m.group(1))
    for m in re.finditer(r"LogicalProject\((.*)\)\s*$", text, re.M):
        body = m.group(1)
        for item in re.finditer(r"[\w$]+=\[(.*?)\](?=, [\w$]+=\[|$)", body):
            exprs.append(item.group(1))
    for m in re.finditer(r"LogicalFilter\(condition=\[(.*?)\]\)", text):
        exprs.append(m.group(1))
    return exprs


STRING_FUNCS = ("LIKE", "SUBSTRING", "||", "CONCAT", "TRIM", "UPPER", "LOWER", "VARCHAR", "CHAR(")


def test_golden_plans_compile(bsql):
    """Every filter/project expression of the reference's golden TPC-H plans parses and
    compiles, except the string functions evaluated before the interpreter."""
    exprs = _plan_expressions()
    assert len(exprs) > 40
    compiled = 0
    for e in exprs:
        if any(f in e for f in STRING_FUNCS) or e in ("true",):
            continue
        n = max([int(x) for x in re.findall(r"\$(\d+)", e)] + [0]) + 1
        if not re.search(r"\$\d+", e):
            continue
        types = [INT64] * n
        # columns compared with timestamps are dates
        for k in re.findall(r"\$(\d+), \d{4}-\d{2}-\d{2}", e):
            types[int(k)] = TS_D
        for k in re.findall(r"(?:AND|OR)\(\$(\d+)", e):
            types[int(k)] = BOOL8
        try:
            bsql.Program.compile(e, types)
            compiled += 1
        except BsqlError as err:
            if "string" in str(err) or "literal expression" in str(err):
                continue
            raise AssertionError(f"{e}: {err}")
    assert compiled > 30


def test_benchmark_plans_compile(bsql):
    # 1M-row CPU configuration
    pred = bsql.Program.compile("AND(>($2, 50), <($1, 0.3:DECIMAL(2, 1)))", [INT64, FLOAT64, INT32])
    proj = bsql.Program.compile(["+($0, *($2, 3))", "-(*($1, 2.0:DECIMAL(2, 1)), 1.0:DECIMAL(2, 1))"],
                                [INT64, FLOAT64, INT32])
    assert proj.output_type([INT64, FLOAT64, INT32], 0) == INT64
    assert proj.output_type([INT64, FLOAT64, INT32], 1) == FLOAT64
    assert pred.num_conjuncts == 2
    # 100M-row scan: qty, price, disc, tax, shipdays
    pred = bsql.Program.compile(
        "AND(<=($4, 10471), >=($2, 0.05:DECIMAL(3, 2)), <=($2, 0.07:DECIMAL(3, 2)), <($0, 24))",
        [INT32, FLOAT64, FLOAT64, FLOAT64, INT32])
    assert pred.num_conjuncts == 4
    proj = bsql.Program.compile(["*(*($1, -(1, $2)), +(1, $3))", "*($1, $2)"],
                                [INT32, FLOAT64, FLOAT64, FLOAT64, INT32])
    assert proj.num_outputs == 2


def _scan_cost(order, sel, cost, rows_per_line):
    """Expected cost of a conjunct order under the range-test selection's
    model, summed directly over the order (no subsets)."""
    total, p = 0.0, 1.0
    for j in order:
        total += cost[j] * ((1 - (1 - p) ** rows_per_line[j]) + 0.2 * (1 - (1 - p) ** 1024))
        p *= sel[j]
    return total


def test_order_conjuncts_is_least_cost_order(bsql):
    """bsql_order_conjuncts returns the cheapest order of every permutation
    (brute force over 5 conjuncts, random selectivities and costs); a lone
    very selective term goes first; it rejects bad arguments."""
    import itertools

    import numpy as np

    from blazingsql_amd import learned
    types = [FLOAT64, INT32, FLOAT64, INT32, FLOAT64]
    terms = [">($0, 0.5:DOUBLE)", "<($1, 7)", "<($2, 0.1:DOUBLE)", ">($3, 100)", "<($4, 3.0:DOUBLE)"]
    text = "AND(" + ", ".join(terms) + ")"
    rpl = [16, 32, 16, 32, 16]
    rng = np.random.default_rng(11)
    for _ in range(20):
        sel = rng.uniform(0.001, 1.0, 5).astype(np.float32)
        cost = rng.uniform(50, 120, 5).astype(np.float32)
        pred = bsql.Program.compile(text, types)
        got = learned.order_for_scan(pred, types, sel, cost)
        assert sorted(got) == list(range(5))
        best = min(_scan_cost(o, sel, cost, rpl) for o in itertools.permutations(range(5)))
        assert _scan_cost(got, sel, cost, rpl) <= best * (1 + 1e-9)
    pred = bsql.Program.compile(text, types)
    order = learned.order_for_scan(pred, types, [0.9, 0.9, 0.001, 0.9, 0.9], [100.0] * 5)
    assert order[0] == 2
    with pytest.raises(BsqlError):
        learned.order_for_scan(pred, types, [0.5] * 4, [1.0] * 4)


def test_order_conjuncts_rank_rule_fallback(bsql):
    """Conjuncts that are not all column-versus-literal range tests fall back
    to ascending cost / (1 - selectivity)."""
    from blazingsql_amd import learned
    types = [INT32, INT32, INT32]
    pred = bsql.Program.compile("AND(>(+($0, $1), 3), <($2, 5), =(MOD($0, 2), 0))", types)
    order = learned.order_for_scan(pred, types, [0.5, 0.9, 0.1], [10.0, 10.0, 10.0])
    assert order == [2, 0, 1]
