"""GPU parity: the HIP interpreter / filter / fused filter+project through the
C ABI against the CPU oracle (bit-exact), the reference's golden vectors, and
size-independent properties at large sizes."""
import numpy as np
import pytest

from tests.helpers import (BOOL8, FLOAT32, FLOAT64, INT8, INT16, INT32, INT64, TS_D, ExprGen, assert_same, device_columns,
                           eval_program_both, host_columns, make_table)

pytestmark = pytest.mark.gpu


def _oracle_prog(p):
    return p.to_oracle_dict()


def test_golden_numeric_program(bsql, gpu, golden):
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    g = golden["numeric_program"]
    for name, t in g["types"].items():
        rs = [None] * 4
        rs[g["right_scalar_op"]] = (t, g["right_scalar_value"])
        p = bsql.Program.from_arrays(g["left"], g["right"], g["out"], g["ops"], g["final_out"], right_scalars=rs)
        conv = (lambda v: np.asarray(v) != 0) if name == "bool" else np.asarray
        ins = [DeviceColumn.from_numpy(t, conv(g[c])) for c in ("col1", "col2", "col3")]
        outs = [DeviceColumn.empty(t, 10), DeviceColumn.empty(t, 10)]
        perform_interpreter_operation(p, ins, outs, 10)
        (o1, v1), (o2, v2) = outs[0].to_numpy(), outs[1].to_numpy()
        e1 = np.asarray(g["expected_col1"]).astype(np.int64)
        e2 = np.asarray(g["expected_col2"]).astype(np.int64)
        if name == "bool":
            e1, e2 = e1 != 0, e2 != 0
        np.testing.assert_array_equal(o1, e1.astype(o1.dtype), err_msg=name)
        np.testing.assert_array_equal(o2, e2.astype(o2.dtype), err_msg=name)
        assert v1.all() and v2.all()


@pytest.mark.parametrize("unit", ["D", "s", "ms", "us", "ns"])
def test_golden_timestamps(bsql, gpu, golden, unit):
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    g = golden["timestamp_components"]
    t = g["types"][unit]
    p = bsql.Program.from_arrays([0] * 7, [-1] * 7, [1, 2, 3, 4, 5, 6, 7], g["ops"], [1, 2, 3, 4, 5, 6, 7])
    ins = [DeviceColumn.from_numpy(t, g["inputs"][unit])]
    outs = [DeviceColumn.empty(INT32, 10) for _ in range(7)]
    perform_interpreter_operation(p, ins, outs, 10)
    exp = g["expected_days"] if unit == "D" else g["expected_other"]
    for o, e in zip(outs, exp):
        np.testing.assert_array_equal(o.to_numpy()[0], e)
    g = golden["day_of_week"]
    p = bsql.Program.compile(g["expression"], [t])
    ins = [DeviceColumn.from_numpy(t, g["inputs"][unit])]
    outs = [DeviceColumn.empty(2, 20)]
    perform_interpreter_operation(p, ins, outs, 20)
    np.testing.assert_array_equal(outs[0].to_numpy()[0], g["expected_days"] if unit == "D" else g["expected_other"])


@pytest.mark.parametrize("nulls", [False, True])
def test_golden_filter_mod(bsql, gpu, golden, nulls):
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    g = golden["filter_mod"]
    for name, t in g["types"].items():
        conv = (lambda v: np.asarray(v) != 0) if name == "bool" else np.asarray
        valid = g["null_validity"] if nulls else None
        ins = [DeviceColumn.from_numpy(t, conv(g["col1"]), valid), DeviceColumn.from_numpy(t, conv(g["col2"]), valid)]
        outs = [DeviceColumn.empty(t, 40, nullable=nulls) for _ in range(2)]
        p = bsql.Program.compile(g["expression"], [t, t])
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_columns(p, ins, outs, cnt, 40)
        n = int(cnt.item())
        if name == "bool":
            assert n == (20 if nulls else 40)
            continue
        assert n == 20, name
        (o1, v1), (o2, v2) = outs[0].to_numpy(n), outs[1].to_numpy(n)
        np.testing.assert_array_equal(o1, np.asarray(g["expected_col1"]).astype(o1.dtype))
        np.testing.assert_array_equal(o2, np.asarray(g["expected_col2"]).astype(o2.dtype))
        assert v1.all() and v2.all()


def test_golden_round_and_compare(bsql, gpu, golden):
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    g = golden["round"]
    for t, dt in ((9, np.float32), (10, np.float64)):
        p = bsql.Program.compile(["ROUND($0)", "ROUND($0, 2)", "ROUND($0, 5)"], [t])
        ins = [DeviceColumn.from_numpy(t, np.asarray(g["col1"], dtype=dt))]
        outs = [DeviceColumn.empty(t, 7) for _ in range(3)]
        perform_interpreter_operation(p, ins, outs, 7)
        for o, e in zip(outs, g["expected"]):
            np.testing.assert_array_equal(o.to_numpy()[0], np.asarray(e, dtype=dt))
    g = golden["project_compare"]
    p = bsql.Program.compile([">($0, 3)", "<($2, 3)"], [INT32, INT32, INT32])
    ins = [DeviceColumn.from_numpy(INT32, g["col1"]), DeviceColumn.from_numpy(INT32, g["col1"]),
           DeviceColumn.from_numpy(INT32, g["col3"])]
    outs = [DeviceColumn.empty(BOOL8, 7), DeviceColumn.empty(BOOL8, 7)]
    perform_interpreter_operation(p, ins, outs, 7)
    np.testing.assert_array_equal(outs[0].to_numpy()[0], g["expected_col1"])
    np.testing.assert_array_equal(outs[1].to_numpy()[0], g["expected_col2"])


def test_golden_strings(bsql, gpu, golden):
    """String comparison and CHAR_LENGTH inside the VM (filter_tests.cpp:178-214,
    process_project.cpp:350-366)."""
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    g = golden["filter_strings"]
    strs = [s.encode() for s in g["col1"]]
    offs = np.concatenate([[0], np.cumsum([len(s) for s in strs])]).astype(np.int32)
    chars = np.frombuffer(b"".join(strs) + b"\0", dtype=np.uint8)
    valid = np.asarray(g["null_validity"], dtype=bool)
    col = DeviceColumn.from_numpy(23, None, valid, chars=chars, offsets=offs)
    p = bsql.Program.compile(["=($0, 'bar')", ">($0, 'd')"], [23])
    outs = [DeviceColumn.empty(BOOL8, 10), DeviceColumn.empty(BOOL8, 10)]
    perform_interpreter_operation(p, [col], outs, 10)
    eq, veq = outs[0].to_numpy()
    gt, vgt = outs[1].to_numpy()
    np.testing.assert_array_equal(veq, valid)
    sel = [s for s, v, ok in zip(g["col1"], gt, vgt) if ok and v]
    assert sel == g["gt_expected"]["col1"]
    assert [s for s, v, ok in zip(g["col1"], eq, veq) if ok and v] == []  # 'bar' is null here
    g = golden["char_length"]
    strs = [s.encode() for s in g["col1"]]
    offs = np.concatenate([[0], np.cumsum([len(s) for s in strs])]).astype(np.int32)
    chars = np.frombuffer(b"".join(strs) + b"\0", dtype=np.uint8)
    col = DeviceColumn.from_numpy(23, None, np.asarray(g["valid"], dtype=bool), chars=chars, offsets=offs)
    p = bsql.Program.compile(["-(CHAR_LENGTH($0), 1)"], [23])
    outs = [DeviceColumn.empty(INT32, 4)]
    perform_interpreter_operation(p, [col], outs, 4)
    o, v = outs[0].to_numpy()
    np.testing.assert_array_equal(v, np.asarray(g["expected_valid"], dtype=bool))
    np.testing.assert_array_equal(o[v], np.asarray(g["expected"])[v])


TYPES6 = [INT32, INT64, FLOAT64, FLOAT64, TS_D, INT32]


@pytest.mark.parametrize("lds", [False, True])
@pytest.mark.parametrize("seed", range(6))
def test_random_expressions_vs_oracle(bsql, gpu, seed, lds):
    """Random programs against the oracle, with the interpreter's register
    file in VGPRs or (lds) in LDS."""
    rng = np.random.default_rng(100 + seed)
    types = TYPES6 + [BOOL8]
    n = [1, 31, 257, 1025, 4099, 70001][seed]
    table = make_table(rng, types, n, null_rates=list(rng.uniform(0, 0.3, len(types))))
    gen = ExprGen(rng, types)
    exprs = []
    while len(exprs) < 12:
        e = gen.boolean(3) if rng.random() < 0.4 else gen.numeric(4)
        if e and "$" in e:
            exprs.append(e)
    for e in exprs:
        try:
            res = eval_program_both(bsql, [e], table, n, launch={"lds_regs": lds})
        except Exception as err:  # programs the interpreter rejects must be rejected on compile
            assert "bsql error 2" in str(err), (e, err)
            continue
        for (gv, gvalid), (ov, ovalid), t in res:
            assert_same(gv, gvalid, ov, ovalid, what=e)


def _c1_table(n, seed=7):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1000000, n, dtype=np.int64)
    b = rng.random(n)
    c = rng.integers(0, 100, n, dtype=np.int32)
    cvalid = rng.random(n) >= 0.05
    return [(INT64, a, None), (FLOAT64, b, None), (INT32, c, cvalid)]


C1_PRED = "AND(>($2, 50), <($1, 0.3:DECIMAL(2, 1)))"
C1_PROJ = ["+($0, *($2, 3))", "-(*($1, 2.0:DECIMAL(2, 1)), 1.0:DECIMAL(2, 1))"]


@pytest.mark.parametrize("lds", [False, True])
@pytest.mark.parametrize("n", [0, 1, 1000, 1 << 20])
def test_filter_project_c1_vs_oracle(bsql, gpu, n, lds):
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project
    table = _c1_table(n)
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(C1_PRED, types)
    proj = bsql.Program.compile(C1_PROJ, types)
    proj.set_launch(lds_regs=lds)  # the output phase's register file in LDS or in VGPRs
    outs = [DeviceColumn.empty(INT64, n), DeviceColumn.empty(FLOAT64, n)]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    filter_project(pred, proj, device_columns(table), outs, cnt, n)
    m = int(cnt.item())
    # the oracle runs the two reference operators: filter, then project
    on, ores = oracle.filter_project(_full_prog(bsql, C1_PRED, types), _full_prog(bsql, C1_PROJ, types),
                                     host_columns(table), [INT64, FLOAT64], n)
    assert m == on
    for k in range(2):
        gv, gval = outs[k].to_numpy(m)
        assert_same(gv, gval, ores[k][0], ores[k][1], what=f"C1 output {k}")
    if n:
        # independent restatement of the query in numpy
        a, b, c = table[0][1], table[1][1], table[2][1]
        sel = table[2][2] & (c > 50) & (b < 0.3)
        np.testing.assert_array_equal(outs[0].to_numpy(m)[0], a[sel] + c[sel].astype(np.int64) * 3)
        np.testing.assert_array_equal(outs[1].to_numpy(m)[0], b[sel] * 2.0 - 1.0)


def _full_prog(bsql, exprs, types):
    """A program over the full table for the oracle (identity column map)."""
    import oracle
    return oracle.full_table_program(bsql.Program.compile(exprs, types), len(types))


def c2_table(n, seed=11):
    rng = np.random.default_rng(seed)
    qty = rng.integers(1, 51, n, dtype=np.int32)
    price = rng.uniform(900.0, 105000.0, n)
    disc = rng.integers(0, 11, n) / 100.0
    tax = rng.integers(0, 9, n) / 100.0
    ship = rng.integers(8036, 10562, n, dtype=np.int32)
    return [(INT32, qty, None), (FLOAT64, price, None), (FLOAT64, disc, None), (FLOAT64, tax, None),
            (INT32, ship, None)]


C2_PRED = "AND(<=($4, 10471), >=($2, 0.05:DECIMAL(3, 2)), <=($2, 0.07:DECIMAL(3, 2)), <($0, 24))"
C2_PROJ = ["*(*($1, -(1, $2)), +(1, $3))", "*($1, $2)"]


@pytest.mark.parametrize("mode", ["ranges", "ops", "ops_lazy"])
@pytest.mark.parametrize("order", [None, [3, 1, 2, 0], [1, 0, 3, 2]])
def test_filter_project_c2_orders_vs_oracle(bsql, gpu, order, mode):
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project
    n = 3_000_017
    table = c2_table(n)
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(C2_PRED, types)
    proj = bsql.Program.compile(C2_PROJ, types)
    if order is not None:
        pred.set_conjunct_order(order)
    pred.set_launch(ranges=mode == "ranges", lazy_loads=mode == "ops_lazy")
    outs = [DeviceColumn.empty(FLOAT64, n), DeviceColumn.empty(FLOAT64, n)]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    filter_project(pred, proj, device_columns(table), outs, cnt, n)
    m = int(cnt.item())
    on, ores = oracle.filter_project(_full_prog(bsql, C2_PRED, types), _full_prog(bsql, C2_PROJ, types),
                                     host_columns(table), [FLOAT64, FLOAT64], n)
    assert m == on and 0.10 < m / n < 0.14
    for k in range(2):
        gv, gval = outs[k].to_numpy(m)
        assert_same(gv, gval, ores[k][0], ores[k][1], what=f"C2 output {k}")


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 2047, 2048, 2049, 4096 + 3,
                               100003, 2048 * 64 * 3 + 5])
@pytest.mark.parametrize("sel", ["none", "all", "half", "sparse"])
def test_filter_gather_edges_vs_oracle(bsql, gpu, n, sel):
    """Sizes around slice, chunk, workgroup-block and validity-word boundaries."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    rng = np.random.default_rng(n * 7 + len(sel))
    table = make_table(rng, [INT32, FLOAT64, INT64, INT32], n, null_rates=[0.1, 0.0, 0.3, 0.0])
    thr = {"none": ">($3, 100000)", "all": ">($3, -100000)", "half": ">($3, 0)", "sparse": ">($3, 990)"}[sel]
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(thr, types)
    outs = [DeviceColumn.empty(t, n, nullable=v is not None) for t, _, v in table]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    filter_columns(pred, device_columns(table), outs, cnt, n)
    m = int(cnt.item())
    on, ores = oracle.filter_table(_full_prog(bsql, thr, types), host_columns(table), n)
    assert m == on
    for k in range(4):
        gv, gval = outs[k].to_numpy(m)
        assert_same(gv, gval, ores[k][0], ores[k][1], what=f"col {k}")


def test_transcendentals_within_one_ulp(bsql, gpu):
    """FP64 math functions: within 1 ulp of the host libm."""
    rng = np.random.default_rng(5)
    n = 20000
    y = rng.uniform(1e-3, 1e3, n)
    table = [(FLOAT64, y, None), (FLOAT64, rng.uniform(-1, 1, n), None)]
    exprs = ["SIN($0)", "COS($0)", "TAN($0)", "LN($0)", "LOG10($0)", "POWER($0, 2.5:DECIMAL(2, 1))",
             "ATAN($0)", "ASIN($1)", "ACOS($1)", "FLOOR($0)", "CEIL($0)"]
    res = eval_program_both(bsql, exprs, table, n)
    for e, ((gv, gval), (ov, oval), t) in zip(exprs, res):
        np.testing.assert_array_equal(gval, oval)
        gi, oi = gv.view(np.int64), ov.view(np.int64)
        ulp = np.abs(gi - oi)
        assert ulp.max() <= 1, (e, int(ulp.max()))
        if e.startswith(("FLOOR", "CEIL")):
            assert ulp.max() == 0


def test_empty_and_single_row_execute(bsql, gpu):
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    p = bsql.Program.compile(["+($0, 1)"], [INT32])
    ins = [DeviceColumn.from_numpy(INT32, np.zeros(0, dtype=np.int32))]
    outs = [DeviceColumn.empty(INT32, 0)]
    perform_interpreter_operation(p, ins, outs, 0)
    ins = [DeviceColumn.from_numpy(INT32, np.array([41], dtype=np.int32), np.array([True]))]
    outs = [DeviceColumn.empty(INT32, 1)]
    perform_interpreter_operation(p, ins, outs, 1)
    v, ok = outs[0].to_numpy()
    assert v[0] == 42 and ok[0]


def _range_predicate(rng, types, table):
    """A random AND chain of column-vs-literal comparisons (the shape that
    runs as range tests), sometimes with a literal on the left, a null
    literal, or a mixed int/float domain. Literals are drawn near the column's
    values so that the chains select a fair share of the rows."""
    ops = ["<", ">", "<=", ">=", "<>", "="]
    conj = []
    for _ in range(int(rng.integers(1, 5))):
        c = int(rng.integers(0, len(types)))
        t = types[c]
        v = table[c][1][int(rng.integers(0, len(table[c][1])))]
        r = rng.random()
        if t in (FLOAT64, FLOAT32):
            lit = f"{float(v):.2f}:DECIMAL(6, 2)" if r < 0.7 else str(int(v))
        elif t == TS_D:
            d = np.datetime64(int(v), "D").astype(str)
            lit = d if r < 0.6 and len(d) == 10 else str(int(v))
        else:
            lit = str(int(v)) if r < 0.8 else f"{int(v) + 0.5:.1f}:DECIMAL(8, 1)"
        if rng.random() < 0.04:
            lit = "null:INTEGER"
        op = ops[int(rng.choice(6, p=[0.22, 0.22, 0.22, 0.22, 0.08, 0.04]))]
        conj.append(f"{op}({lit}, ${c})" if rng.random() < 0.3 else f"{op}(${c}, {lit})")
    return conj[0] if len(conj) == 1 else "AND(" + ", ".join(conj) + ")"


@pytest.mark.parametrize("seed", range(8))
def test_range_predicates_vs_oracle(bsql, gpu, seed):
    """Column-vs-literal AND chains: range tests and the general ops select the
    same rows as the oracle, with and without nulls, in both domains."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    rng = np.random.default_rng(1000 + seed)
    n = 200_003
    types = [INT32, FLOAT64, INT64, FLOAT32, TS_D, INT8]
    table = make_table(rng, types, n, null_rates=[0.1, 0.05, 0.0, 0.0, 0.02, 0.0])
    dcols = device_columns(table)
    for _ in range(4):
        text = _range_predicate(rng, types, table)
        on, ores = oracle.filter_table(_full_prog(bsql, text, types), host_columns(table), n)
        for ranges, rpt in ((True, 0), (True, 8), (True, 32), (False, 0)):
            pred = bsql.Program.compile(text, types)
            pred.set_launch(rows_per_thread=rpt, ranges=ranges)
            outs = [DeviceColumn.empty(t, n, nullable=v is not None) for t, _, v in table]
            cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
            filter_columns(pred, dcols, outs, cnt, n)
            m = int(cnt.item())
            assert m == on, (text, ranges, m, on)
            for k in range(len(types)):
                gv, gval = outs[k].to_numpy(m)
                assert_same(gv, gval, ores[k][0], ores[k][1], what=f"{text} ranges={ranges} col {k}")


def test_literal_roots(bsql, gpu):
    """Literal expressions fill their column (make_column_from_scalar,
    LogicalProject.cpp:822-832): every row the value, or every row null; a
    literal predicate keeps all rows or none."""
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns, perform_interpreter_operation
    n = 5000
    x = np.arange(n, dtype=np.int64)
    p = bsql.Program.compile(["5", "null:DOUBLE", "2.5:DOUBLE", "+($0, 1)", "true", "-7:BIGINT"], [INT64])
    ot = [p.output_type([INT64], k) for k in range(6)]
    assert ot == [INT32, FLOAT64, FLOAT64, INT64, BOOL8, INT64]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    perform_interpreter_operation(p, [DeviceColumn.from_numpy(INT64, x)], outs, n)
    vals = [o.to_numpy(n) for o in outs]
    assert (vals[0][0] == 5).all() and vals[0][1].all()
    assert not vals[1][1].any()
    assert (vals[2][0] == 2.5).all() and vals[2][1].all()
    np.testing.assert_array_equal(vals[3][0], x + 1)
    assert (vals[4][0] == 1).all() and (vals[5][0] == -7).all()
    for text, expect in (("true", n), ("false", 0)):
        pred = bsql.Program.compile(text, [INT64])
        out = [DeviceColumn.empty(INT64, n, nullable=False)]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_columns(pred, [DeviceColumn.from_numpy(INT64, x)], out, cnt, n)
        assert int(cnt.item()) == expect
        if expect:
            np.testing.assert_array_equal(out[0].to_numpy(n)[0], x)


def test_alternating_programs_on_one_stream(bsql, gpu):
    """The launch plan copied to the stream's scratch is reused when the same
    program runs over the same columns again; alternating programs, input
    columns and row counts on one stream must each get their own plan."""
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns, perform_interpreter_operation
    rng = np.random.default_rng(77)
    n = 50_003
    a = rng.integers(-100, 100, n).astype(np.int64)
    b = rng.integers(-100, 100, n).astype(np.int64)
    da, db = DeviceColumn.from_numpy(INT64, a), DeviceColumn.from_numpy(INT64, b)
    p1 = bsql.Program.compile(["+($0, 1)"], [INT64])
    p2 = bsql.Program.compile(["*($0, 3)"], [INT64])
    pred = bsql.Program.compile(">($0, 0)", [INT64])
    out = DeviceColumn.empty(INT64, n)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for prog, col, host, rows, fn in [(p1, da, a, n, lambda x: x + 1), (p2, da, a, n, lambda x: x * 3),
                                      (p1, da, a, n, lambda x: x + 1), (p1, db, b, n, lambda x: x + 1),
                                      (p1, da, a, n // 2, lambda x: x + 1), (p1, da, a, n, lambda x: x + 1)]:
        perform_interpreter_operation(prog, [col], [out], rows)
        v, _ = out.to_numpy(rows)
        np.testing.assert_array_equal(v, fn(host[:rows]))
        fo = DeviceColumn.empty(INT64, n, nullable=False)
        filter_columns(pred, [col], [fo], cnt, rows)
        m = int(cnt.item())
        np.testing.assert_array_equal(fo.to_numpy(m)[0], host[:rows][host[:rows] > 0])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_selective_first_range_vs_oracle(bsql, gpu, seed):
    """A very selective first range (0.05-0.5 %) followed by 2-7 ranges over
    every column kind, with nulls (most slices end with no live row, the
    rest with a few): range tests and the general ops select what the oracle
    selects, also with a NOT-EQUAL among the later ranges."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_columns
    rng = np.random.default_rng(2000 + seed)
    n = 300_007
    types = [INT32, FLOAT64, INT64, FLOAT32, TS_D, INT8, INT16, BOOL8]
    table = make_table(rng, types, n, null_rates=[0.0, 0.05, 0.1, 0.0, 0.02, 0.1, 0.0, 0.05])
    dcols = device_columns(table)
    first = table[0][1]
    for trial in range(5):
        # first: a narrow band of column 0 (int32 in [-1000, 1000))
        lo = int(rng.integers(-1000, 990))
        width = int(rng.integers(0, 9))
        terms = [f">=($0, {lo})", f"<=($0, {lo + width})"]
        others = rng.permutation(np.arange(1, len(types)))[:int(rng.integers(2, 7))]
        for c in others:
            t, v, _ = table[c]
            if t == BOOL8:
                terms.append(f"=(${c}, true)")
            elif t in (FLOAT64, FLOAT32):
                q = float(np.quantile(v, rng.uniform(0.2, 0.8)))
                terms.append(f"<(${c}, {q:.6f}:DOUBLE)")
            elif t == TS_D:
                terms.append(f">=(${c}, 1990-01-01)")
            else:
                q = int(np.quantile(v.astype(np.int64), rng.uniform(0.2, 0.8)))
                terms.append(f"{'<>' if trial == 3 else '>='}(${c}, {q})")
        text = "AND(" + ", ".join(terms) + ")"
        on, ores = oracle.filter_table(_full_prog(bsql, text, types), host_columns(table), n)
        assert 0 < on < n // 50 or width == 0, (text, on)
        for ranges in (True, False):
            pred = bsql.Program.compile(text, types)
            pred.set_launch(ranges=ranges)
            outs = [DeviceColumn.empty(t, n, nullable=v is not None) for t, _, v in table]
            cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
            filter_columns(pred, dcols, outs, cnt, n)
            m = int(cnt.item())
            assert m == on, (text, ranges, m, on)
            for k in range(len(types)):
                gv, gval = outs[k].to_numpy(m)
                assert_same(gv, gval, ores[k][0], ores[k][1], what=f"{text} ranges={ranges} col {k}")


FUSED_SIZES = [0, 1, 31, 32, 33, 255, 257, 2047, 2048, 2049, 8191, 8192, 8193, 3 * 8192 + 1000, 250_003,
               1_000_000]


@pytest.mark.parametrize("n", FUSED_SIZES)
def test_fused_filter_project_vs_two_pass_and_oracle(bsql, gpu, n):
    """The one-pass kernel (range predicate, outputs at look-back positions,
    validity merged at wave boundaries) against the selection / row ids /
    output-pass path and the oracle, at sizes around validity-word, chunk,
    slice and tile boundaries, with nulls in predicate and output columns;
    plus the capacity-limited form."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project, filter_project_sized
    rng = np.random.default_rng(n + 3)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, max(n, 1), null_rates=[0.1, 0.0, 0.3, 0.05])
    table = [(t, v[:n], None if val is None else val[:n]) for t, v, val in table]
    dev, host = device_columns(table), host_columns(table)
    pred_text = "AND(>($0, -500), <($3, 800), >=($1, -900.5:DECIMAL(4, 1)))"
    exprs = ["+($2, *($0, 3))", "*($1, 2.5:DECIMAL(2, 1))", "$3", "CASE(IS NOT NULL($2), $2, 0)", "RAND()"]
    om = None
    results = []
    for fused in (True, False):
        pred = bsql.Program.compile(pred_text, types)
        pred.set_launch(fused=fused)
        proj = bsql.Program.compile(exprs, types)
        ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
        outs = [DeviceColumn.empty(t, max(n, 1)) for t in ot]
        cnt = torch.full((1,), -7, dtype=torch.int64, device="cuda")
        print(f"filter_project n={n} fused={fused}", flush=True)
        filter_project(pred, proj, dev, outs, cnt, n)
        torch.cuda.synchronize()
        m = int(cnt.item())
        if om is None:
            om, ores = oracle.filter_project(oracle.full_table_program(pred, 4), oracle.full_table_program(proj, 4),
                                             host, ot, n)
        assert m == om, (fused, m, om)
        got = [o.to_numpy(m) for o in outs]
        for k in range(5):
            assert_same(got[k][0], got[k][1], ores[k][0], ores[k][1], what=f"fused={fused} output {k}")
        results.append(got)
    # RAND numbers the output rows: both paths agree (and with the oracle, above)
    assert_same(results[0][4][0], results[0][4][1], results[1][4][0], results[1][4][1], what="RAND")
    if n >= 1000:
        # pre-sized outputs: the first `cap` rows written, the full count reported
        cap = om // 2
        pred = bsql.Program.compile(pred_text, types)
        pred.set_launch(fused=True)
        proj = bsql.Program.compile(exprs[:4], types)
        outs = [DeviceColumn.empty(t, max(cap, 1)) for t in ot[:4]]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_project_sized(pred, proj, dev, outs, cnt, cap, n)
        torch.cuda.synchronize()
        assert int(cnt.item()) == om
        for k in range(4):
            gv, gval = outs[k].to_numpy(cap)
            assert_same(gv, gval, ores[k][0][:cap], ores[k][1][:cap], what=f"sized output {k}")
