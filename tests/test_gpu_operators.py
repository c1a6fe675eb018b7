"""GPU tests of the operator surface: the processor functions (process_filter,
process_project, evaluate_expressions, applyBooleanFilter) on tables, with the
reference's own test vectors (filter_tests.cpp, process_project.cpp,
apply_filter.cpp) and the oracle on random tables; the projection bypass and
scan helpers; RETRY on an allocation failure; and the C++ kernel layer
(Filter / Projection / BindableTableScan do_process) through its test binary."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import (BOOL8, FLOAT64, INT32, INT64, TS_D, assert_same, device_columns, host_columns,
                           make_table)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _table(cols, names=None):
    from blazingsql_amd.processor import Table
    return Table.from_columns(cols, names)


def _strings(values, valid=None):
    from blazingsql_amd.interop import DeviceColumn
    return DeviceColumn.from_strings(values, None if valid is None else np.asarray(valid, dtype=bool))


def test_cpp_operator_kernels(bsql, gpu):
    """Filter -> Projection -> BindableTableScan do_process, RETRY and FAIL,
    in a C++ program over the C ABI (tests/cpp/test_kernels.cpp)."""
    binary = os.path.join(ROOT, "tests", "cpp", "test_kernels")
    if not os.path.exists(binary):
        from tests.cpp import build_tests
        build_tests.build()
    print("running", binary, flush=True)
    r = subprocess.run([binary], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


@pytest.mark.parametrize("nulls", [False, True])
def test_process_filter_golden_mod(bsql, gpu, golden, nulls):
    """filter_tests.cpp:75-171: LogicalFilter(condition=[=(MOD(+($0, $1), 2), 0)])."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter
    g = golden["filter_mod"]
    for name, t in g["types"].items():
        if name == "bool":
            continue
        valid = g["null_validity"] if nulls else None
        tab = _table([DeviceColumn.from_numpy(t, np.asarray(g["col1"]), valid),
                      DeviceColumn.from_numpy(t, np.asarray(g["col2"]), valid)], ["a", "b"])
        out = process_filter(tab, g["query"])
        assert out.num_rows == 20 and out.names == ["a", "b"], name
        (o1, v1), (o2, v2) = out.to_numpy()
        np.testing.assert_array_equal(o1, np.asarray(g["expected_col1"]).astype(o1.dtype))
        np.testing.assert_array_equal(o2, np.asarray(g["expected_col2"]).astype(o2.dtype))
        assert v1.all() and v2.all()
        assert (out.column(0).validity is not None) == nulls


def test_process_filter_empty(bsql, gpu, golden):
    """filter_tests.cpp:23-73: an empty input gives an empty table of the same
    schema; a predicate no row passes gives zero rows."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter
    g = golden["filter_mod"]
    empty = _table([DeviceColumn.from_numpy(INT32, np.zeros(0, np.int32)),
                    DeviceColumn.from_numpy(INT64, np.zeros(0, np.int64))], ["a", "b"])
    out = process_filter(empty, g["empty_query"])
    assert out.num_rows == 0 and out.num_columns == 2
    assert [out.column(k).type_id for k in range(2)] == [INT32, INT64]
    tab = _table([DeviceColumn.from_numpy(INT32, np.asarray(g["col1"])),
                  DeviceColumn.from_numpy(INT32, np.asarray(g["col2"]))])
    assert process_filter(tab, g["empty_query"]).num_rows == 0


def test_process_filter_golden_strings(bsql, gpu, golden):
    """filter_tests.cpp:178-214: string equality / inequality filters, with nulls."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter
    g = golden["filter_strings"]
    col2 = np.asarray(g["col2"])
    # NoNulls (:178-195)
    tab = _table([_strings(g["col1"]), DeviceColumn.from_numpy(INT32, col2)])
    eq = process_filter(tab, g["eq_query"]).to_numpy()
    assert eq[0][0] == g["eq_expected"]["col1"] and list(eq[1][0]) == g["eq_expected"]["col2"]
    gt = process_filter(tab, g["gt_query"]).to_numpy()
    assert gt[0][0] == [s for s in g["col1"] if s > "d"]
    # IneqWithNulls (:197-214): null strings drop out of the comparison
    tab = _table([_strings(g["col1"], g["null_validity"]), DeviceColumn.from_numpy(INT32, col2)])
    gt = process_filter(tab, g["gt_query"]).to_numpy()
    assert gt[0][0] == g["gt_expected"]["col1"] and list(gt[1][0]) == g["gt_expected"]["col2"]
    assert process_filter(tab, g["eq_query"]).num_rows == 0  # 'bar' is a null row here


def test_apply_boolean_filter_golden(bsql, gpu, golden):
    """apply_filter.cpp:44-116: a null or false mask row drops the row; each
    column keeps its nulls (strings too)."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import apply_boolean_filter
    g = golden["apply_filter"]
    tab = _table([DeviceColumn.from_numpy(INT32, np.asarray(g["col1"]), np.asarray(g["col1_valid"], bool)),
                  _strings(g["col2"], g["col2_valid"]),
                  DeviceColumn.from_numpy(INT32, np.asarray(g["col3"]), np.asarray(g["col3_valid"], bool))])
    mask = DeviceColumn.from_numpy(BOOL8, np.asarray(g["mask"], np.uint8), np.asarray(g["mask_valid"], bool))
    out = apply_boolean_filter(tab, mask)
    (c1, v1), (c2, v2), (c3, v3) = out.to_numpy()
    np.testing.assert_array_equal(v1, np.asarray(g["expected_col1_valid"], bool))
    np.testing.assert_array_equal(c1[v1], np.asarray(g["expected_col1"])[v1])
    np.testing.assert_array_equal(v2, np.asarray(g["expected_col2_valid"], bool))
    assert [s for s, ok in zip(c2, v2) if ok] == [s for s, ok in zip(g["expected_col2"], g["expected_col2_valid"]) if ok]
    np.testing.assert_array_equal(v3, np.asarray(g["expected_col3_valid"], bool))
    np.testing.assert_array_equal(c3[v3], np.asarray(g["expected_col3"])[v3])


def test_process_project_golden(bsql, gpu, golden):
    """process_project.cpp:18-199: column picks, a literal, comparisons; every
    output nullable and all valid; names from the query part."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_project
    g = golden["project_compare"]
    cols = lambda: [DeviceColumn.from_numpy(INT32, np.asarray(g["col1"])), DeviceColumn.from_numpy(INT32, np.asarray(g["col1"])),
                    DeviceColumn.from_numpy(INT32, np.asarray(g["col3"]))]
    out = process_project(_table(cols(), ["A", "B", "C"]), "LogicalProject(A=[$0], C=[$2])")
    assert out.names == ["A", "C"]
    (a, va), (c, vc) = out.to_numpy()
    assert list(a) == g["col1"] and list(c) == g["col3"] and va.all() and vc.all()
    out = process_project(_table(cols(), ["A", "B", "C"]), "LogicalProject(EXPR$0=[1])")
    (one, v1), = out.to_numpy()
    assert out.column(0).type_id == INT32 and list(one) == [1] * 7 and v1.all()
    out = process_project(_table(cols(), ["A", "B", "C"]), g["query"])
    (e1, w1), (e2, w2) = out.to_numpy()
    assert out.names == ["EXPR$0", "EXPR$1"] and out.column(0).type_id == BOOL8
    assert list(e1) == g["expected_col1"] and list(e2) == g["expected_col2"] and w1.all() and w2.all()


def test_process_project_string_building_golden(bsql, gpu, golden):
    """process_project with the reference's string-building queries
    (process_project.cpp:368-486): the owned output tables are sized by the
    library, including chars longer than any input column's."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter, process_project
    g = golden["string_concat"]
    out = process_project(_table([_strings(g["col1"]), _strings(g["col2"])], ["a", "b"]), g["query"])
    (c1, v1), (c2, v2) = out.to_numpy()
    assert out.names == ["EXPR$0", "EXPR$1"] and c1 == g["expected_col1"] and c2 == g["expected_col2"]
    assert v1.all() and v2.all()
    g = golden["cast_to_string"]
    cols = [DeviceColumn.from_numpy(BOOL8, np.asarray(g["col1"], dtype=np.uint8)),
            DeviceColumn.from_numpy(INT32, np.asarray(g["col2"], dtype=np.int32)),
            DeviceColumn.from_numpy(FLOAT64, np.asarray(g["col3"])),
            DeviceColumn.from_numpy(13, np.asarray(g["col4"], dtype=np.int64))]
    out = process_project(_table(cols, ["a", "b", "c", "d"]), g["query"])
    res = out.to_numpy()
    for k in range(4):
        assert res[k][0] == g[f"expected_col{k + 1}"], k
    g = golden["string_case"]
    tab = lambda: _table([DeviceColumn.from_numpy(INT32, np.asarray(g["col1"], dtype=np.int32)),
                          _strings(g["col2"])], ["a", "b"])
    (c, v), = process_project(tab(), g["query"]).to_numpy()
    assert c == g["expected"]
    # a filter on a built column keeps the table's own columns
    f = process_filter(tab(), "LogicalFilter(condition=[=(CASE(=(MOD($0, 2), 0), $1, 'LOL'), 'LOL')])")
    (a, _), (b, _) = f.to_numpy()
    assert list(a) == [1, 3, 5, 7] and b == ["quick", "fox", "over", "lazy"]


def test_process_project_bypass(bsql, gpu):
    """bypassingProject: $0..$n-1 in order moves the table through (renamed)."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_project, project_bypass
    assert project_bypass("LogicalProject(a=[$0], b=[$1])", ["a", "b"]) == (True, False)
    assert project_bypass("LogicalProject(x=[$0], y=[$1])", ["a", "b"]) == (True, True)
    assert project_bypass("LogicalProject(b=[$1], a=[$0])", ["a", "b"]) == (False, False)
    assert project_bypass("LogicalProject(a=[$0])", ["a", "b"]) == (False, False)
    col = DeviceColumn.from_numpy(INT64, np.arange(10, dtype=np.int64))
    tab = _table([col, DeviceColumn.from_numpy(INT64, np.arange(10, dtype=np.int64) * 2)], ["a", "b"])
    h = tab.handle
    out = process_project(tab, "LogicalProject(x=[$0], y=[$1])")
    assert out.handle == h and out.names == ["x", "y"]
    assert out.column(0) is col     # the same buffers, nothing copied


@pytest.mark.parametrize("seed", [0, 1])
def test_processors_vs_oracle(bsql, gpu, seed):
    """Random nullable tables: process_filter then process_project against the
    oracle's filter then project, bit-exact; evaluate_expressions likewise."""
    import oracle
    from blazingsql_amd.processor import evaluate_expressions, process_filter, process_project
    from tests.test_gpu_parity import _full_prog
    rng = np.random.default_rng(seed)
    types = [INT32, INT64, FLOAT64, INT32]
    n = 100_003
    table = make_table(rng, types, n, [0.0, 0.1, 0.05, 0.2])
    dev, host = device_columns(table), host_columns(table)
    cond = "OR(AND(>($0, -200), <($2, 100.5:DECIMAL(4, 1))), IS NULL($3))"
    # a nested unary minus is written -(0, x): the reference rewrites only a
    # top-level one (fill_minus_op_with_zero, expression_utils.cpp:1155-1164)
    exprs = ["+($1, *($0, 3))", "CASE(>($2, 0), $2, -(0, $2))", "$3", ">($1, $0)"]
    names = ["p", "q", "r", "s"]
    query = "LogicalProject(" + ", ".join(f"{a}=[{e}]" for a, e in zip(names, exprs)) + ")"
    f = process_filter(_table(dev, ["c0", "c1", "c2", "c3"]), f"LogicalFilter(condition=[{cond}])")
    on, fcols = oracle.filter_table(_full_prog(bsql, cond, types), host, n)
    assert f.num_rows == on and f.names == ["c0", "c1", "c2", "c3"]
    for k, (g, o) in enumerate(zip(f.to_numpy(), fcols)):
        assert_same(g[0], g[1], o[0], o[1], what=f"filtered column {k}")
    p = process_project(f, query)
    assert p.names == names
    ot = [p.column(k).type_id for k in range(4)]
    om, ores = oracle.filter_project(_full_prog(bsql, cond, types), _full_prog(bsql, exprs, types), host, ot, n)
    assert om == p.num_rows
    for k, (g, o) in enumerate(zip(p.to_numpy(), ores)):
        assert_same(g[0], g[1], o[0], o[1], what=f"projected column {k}")
    e = evaluate_expressions(_table(dev), exprs)
    oe = oracle.execute(_full_prog(bsql, exprs, types), host, ot, n)
    for k, (g, o) in enumerate(zip(e.to_numpy(), oe)):
        assert_same(g[0], g[1], o[0], o[1], what=f"expression {k}")


def test_scan_projections_and_aliases(bsql, gpu):
    """BindableTableScan: get_projections picks the loaded columns, a pushed-down
    filter indexes them, fix_column_aliases names the result."""
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import fix_column_aliases, process_filter, scan_projections
    q = "BindableTableScan(table=[[main, t]], filters=[[>($1, 5)]], projects=[[0, 2]], aliases=[[a, c]])"
    assert scan_projections(q, 3) == [0, 2]
    assert scan_projections("BindableTableScan(table=[[main, t]], projects=[[]], aliases=[[$f0]])", 3) == [0]
    assert scan_projections("BindableTableScan(table=[[main, t]])", 3) == [0, 1, 2]
    full = [np.arange(10, dtype=np.int64), np.arange(10, dtype=np.int64) * 10, np.arange(10, dtype=np.int64) % 7]
    loaded = [DeviceColumn.from_numpy(INT64, full[k]) for k in scan_projections(q, 3)]
    out = fix_column_aliases(process_filter(_table(loaded, ["x", "y"]), q), q)
    assert out.names == ["a", "c"]
    (a, _), (c, _) = out.to_numpy()
    keep = full[2] > 5
    np.testing.assert_array_equal(a, full[0][keep])
    np.testing.assert_array_equal(c, full[2][keep])


def test_retry_on_allocation_failure(bsql, gpu):
    """A failing allocation surfaces as OutOfMemory (the operators' RETRY) and
    leaves the input usable; the same call succeeds once memory is there."""
    from blazingsql_amd import _lib as L
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter, process_project, use_torch_allocator
    n = 50_000
    x = np.arange(n, dtype=np.int64)
    tab = _table([DeviceColumn.from_numpy(INT64, x)], ["x"])
    use_torch_allocator(limit_bytes=1024)
    try:
        with pytest.raises(L.OutOfMemory):
            process_filter(tab, "LogicalFilter(condition=[>($0, 10)])")
        with pytest.raises(L.OutOfMemory):
            process_project(tab, "LogicalProject(y=[*($0, 2)])")
        assert tab.num_rows == n  # not consumed
    finally:
        use_torch_allocator()
    out = process_filter(tab, "LogicalFilter(condition=[>($0, 10)])")
    assert out.num_rows == n - 11
    y = process_project(tab, "LogicalProject(y=[*($0, 2)])")
    np.testing.assert_array_equal(y.to_numpy()[0][0], x * 2)


def test_invalid_query_parts(bsql, gpu):
    """FAIL cases: a non-boolean condition, a missing column, no condition."""
    from blazingsql_amd import _lib as L
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_filter
    tab = _table([DeviceColumn.from_numpy(INT64, np.arange(5, dtype=np.int64))])
    with pytest.raises(L.BsqlError, match="boolean mask"):
        process_filter(tab, "LogicalFilter(condition=[+($0, 1)])")
    with pytest.raises(L.BsqlError):
        process_filter(tab, "LogicalFilter(condition=[>($3, 1)])")
    with pytest.raises(L.BsqlError):
        process_filter(tab, "LogicalFilter()")


def _random_batches(seed, n_batches, max_rows, types):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_batches):
        rows = 0 if b == 3 else int(np.exp(rng.uniform(np.log(1), np.log(max_rows))))
        out.append(make_table(rng, types, rows, null_rates=list(rng.uniform(0, 0.3, len(types)))))
    return out


@pytest.mark.parametrize("query", [
    "LogicalProject(x=[+(*($0, 3), $1)], y=[CASE(IS NOT NULL($2), $2, *($3, 2.5:DOUBLE))], z=[>($1, $0)], "
    "w=[$2], v=[EXTRACT(FLAG(YEAR), $4)])",
    "LogicalProject(a=[$0], b=[$1], c=[$2], d=[$3], e=[$4])",          # identity: moved through
    "LogicalProject(p=[$1], q=[$0])",                                 # column copies
])
def test_project_batches_vs_one_batch_at_a_time(bsql, gpu, query):
    """bsql_process_project_batches (the waiting batches of a cache in one
    launch) gives every batch the result bsql_process_project gives it alone."""
    from blazingsql_amd.processor import Table, process_project, process_project_batches
    types = [INT32, INT64, FLOAT64, FLOAT64, TS_D]
    names = ["a", "b", "c", "d", "e"]
    hosts = _random_batches(11, 20, 300_000, types)
    one = [process_project(Table.from_columns(device_columns(h), names), query) for h in hosts]
    many = process_project_batches([Table.from_columns(device_columns(h), names) for h in hosts], query)
    assert len(many) == len(one)
    for i, (a, b) in enumerate(zip(many, one)):
        assert a.num_rows == b.num_rows and a.num_columns == b.num_columns
        for k in range(a.num_columns):
            assert a.names[k] == b.names[k]
            ga, va = a.column(k).to_numpy(a.num_rows)
            gb, vb = b.column(k).to_numpy(b.num_rows)
            assert_same(ga, va, gb, vb, what=f"{query[:30]} batch {i} column {k}")


def test_project_batches_strings_and_retry(bsql, gpu):
    """A projection that builds strings runs batch by batch behind the same
    call (same results); a failing allocation consumes no input."""
    from blazingsql_amd import _lib as L
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import process_project, process_project_batches, use_torch_allocator
    words = [["ab", "xyz", "", "q"], ["hello", "w"], ["a" * 40, "b"]]
    query = "LogicalProject(u=[UPPER($0)], n=[+($1, 1)])"

    def batches():
        return [_table([_strings(w), DeviceColumn.from_numpy(INT64, np.arange(len(w), dtype=np.int64))], ["s", "i"])
                for w in words]

    one = [process_project(t, query) for t in batches()]
    ins = batches()
    use_torch_allocator(limit_bytes=64)
    try:
        with pytest.raises(L.OutOfMemory):
            process_project_batches(ins, query)
        assert all(t.num_rows == len(w) for t, w in zip(ins, words))  # not consumed
    finally:
        use_torch_allocator()
    many = process_project_batches(ins, query)
    for a, b in zip(many, one):
        for k in range(2):
            ga, va = a.column(k).to_numpy(a.num_rows)
            gb, vb = b.column(k).to_numpy(b.num_rows)
            assert list(ga) == list(gb) and list(va) == list(vb)


def test_string_steps_use_installed_allocator(bsql, gpu):
    """The columns a projection builds before the interpreter runs (here
    UPPER of a string column, under LIKE) come from the installed allocator:
    a budget that holds the outputs but not the step column fails the call
    with OutOfMemory, and a passing call leaves only its outputs live."""
    from blazingsql_amd import _lib as L
    from blazingsql_amd.processor import process_project, use_torch_allocator
    n = 4000
    words = [("ab" if i % 3 else "xy") + "c" * (i % 20) for i in range(n)]
    query = "LogicalProject(m=[LIKE(UPPER($0), 'AB%')])"
    out_bytes = n + ((n + 31) // 32 * 4 + 63) // 64 * 64    # BOOL8 data + validity
    tab = _table([_strings(words)], ["s"])
    alloc = use_torch_allocator(limit_bytes=out_bytes + 1024)
    try:
        with pytest.raises(L.OutOfMemory):
            process_project(tab, query)
        assert tab.num_rows == n
    finally:
        alloc = use_torch_allocator()
    out = process_project(tab, query)
    got, valid = out.column(0).to_numpy(n)
    assert list(np.asarray(got, dtype=bool)) == [w.upper().startswith("AB") for w in words] and all(valid)
    import torch
    torch.cuda.synchronize()
    assert alloc.live_bytes == out_bytes    # the step column was released
