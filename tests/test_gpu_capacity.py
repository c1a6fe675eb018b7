"""GPU parity for every program shape the reference runs: wide sums (loads
issued as registers allow), deep programs whose live values spill to LDS
slots (both spill layouts, 64- and 256-lane workgroups), a 9-deep COALESCE
chain, random reference encodings using 40-63 slots, long expression lists
(bare-column copies, literal fills, 64-slot splits into parts, more than one
output launch) through execute, filter+project and process_project, and RAND.
Everything is checked bit-exact against the oracle on the same inputs."""
import os

import numpy as np
import pytest

from tests.helpers import FLOAT64, INT32, INT64, assert_same, device_columns, host_columns, make_table

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _left_deep(cols):
    e = f"${cols[0]}"
    for k in cols[1:]:
        e = f"+({e}, ${k})"
    return e


def _balanced(cols):
    if len(cols) == 1:
        return f"${cols[0]}"
    h = len(cols) // 2
    return f"+({_balanced(cols[:h])}, {_balanced(cols[h:])})"


def _chain(d, col=0):
    """+(*($c, d), +(*($c, d-1), ...)): post-order keeps d values live."""
    return f"${col}" if d == 0 else f"+(*(${col}, {d}), {_chain(d - 1, col)})"


def _coalesce(n):
    e = f"${n - 1}"
    for k in range(n - 2, -1, -1):
        e = f"CASE(IS NOT NULL(${k}), ${k}, {e})"
    return e


def _execute_vs_oracle(bsql, exprs, table, n, launch=None, what=""):
    import oracle
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    types = [t for t, _, _ in table]
    p = bsql.Program.compile(exprs, types)
    if launch:
        p.set_launch(**launch)
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    print(f"execute {what}: {len(exprs)} expressions, {p.num_ops} ops, launch {launch}", flush=True)
    perform_interpreter_operation(p, device_columns(table), outs, n)
    ora = oracle.evaluate(p, host_columns(table), n)
    for k in range(len(ot)):
        gv, gval = outs[k].to_numpy(n)
        assert_same(gv, gval, ora[k][0], ora[k][1], what=f"{what} output {k}")
    return p


LAUNCHES = [None, {"block_size": 64}, {"spill_wide": True}, {"rows_per_thread": 1}]


@pytest.mark.parametrize("shape,n_cols", [("left", 16), ("left", 32), ("left", 60),
                                          ("balanced", 16), ("balanced", 32), ("balanced", 56)])
def test_wide_float_sums(bsql, shape, n_cols):
    rng = np.random.default_rng(n_cols)
    n = 9_001
    table = make_table(rng, [FLOAT64] * n_cols, n, [0.02] * n_cols)
    cols = list(range(n_cols))
    e = _left_deep(cols) if shape == "left" else _balanced(cols)
    for launch in LAUNCHES[:2]:
        p = _execute_vs_oracle(bsql, [e], table, n, launch, what=f"{shape} sum of {n_cols}")
        assert p.plan_info([FLOAT64] * n_cols)["n_spill"] == 0


@pytest.mark.parametrize("depth", [20, 40, 61])
def test_deep_programs_spill(bsql, depth):
    """Live values beyond the register file go to LDS slots and come back."""
    rng = np.random.default_rng(depth)
    n = 7_777
    table = make_table(rng, [FLOAT64, INT64], n, [0.1, 0.1])
    p = bsql.Program.compile(_chain(depth), [FLOAT64, INT64])
    assert p.plan_info([FLOAT64, INT64])["n_spill"] > 0
    for launch in LAUNCHES:
        _execute_vs_oracle(bsql, [_chain(depth), _chain(depth // 2, 1)], table, n, launch,
                           what=f"chain {depth}")


def test_coalesce_chain_9(bsql):
    rng = np.random.default_rng(9)
    n = 10_000
    table = make_table(rng, [FLOAT64] * 9, n, [0.3] * 9)
    _execute_vs_oracle(bsql, [_coalesce(9)], table, n, what="COALESCE of 9")
    table = make_table(rng, [INT32] * 9, n, [0.5] * 9)
    _execute_vs_oracle(bsql, [_coalesce(9), "+($0, $8)"], table, n, what="INT32 COALESCE of 9")


def _random_deep(rng, d):
    """A random expression whose post-order keeps about d values live: a chain
    of random binary ops over computed left operands."""
    ops = ["+", "-", "*"]
    cols = ["$0", "$1", "$2"]

    def leaf():
        c = cols[int(rng.integers(0, 3))]
        return f"{ops[int(rng.integers(0, 3))]}({c}, {int(rng.integers(1, 9))})"

    e = leaf()
    for _ in range(d):
        op = ops[int(rng.integers(0, 3))]
        if rng.random() < 0.2:
            e = f"CASE(>({leaf()}, 0), {e}, {leaf()})"
        e = f"{op}({leaf()}, {e})"
    return e


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_programs_using_40_to_63_slots(bsql, seed):
    """Reference encodings (from the compiler's first-free slot allocation)
    that use 40-63 slots, run through bsql_program_from_arrays as the
    reference's perform_interpreter_operation takes them."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn, Program, perform_interpreter_operation
    rng = np.random.default_rng(100 + seed)
    n = 5_003
    types = [INT64, FLOAT64, INT32]
    table = make_table(rng, types, n, [0.05, 0.05, 0.05])
    done = 0
    for attempt in range(40):
        e = _random_deep(rng, int(rng.integers(30, 58)))
        try:
            p = bsql.Program.compile(e, types)
        except Exception:  # noqa: BLE001 -- over 63 slots: the reference rejects it too
            continue
        a = p.arrays()
        top = max(a["out"] + a["left"] + a["right"])
        if not 40 <= top <= 63:
            continue
        d = p.to_oracle_dict()
        q = Program.from_arrays(a["left"], a["right"], a["out"], a["ops"], a["final_out"],
                                d["left_scalars"], d["right_scalars"])
        used = p.used_columns()
        ins = [table[c] for c in used]
        ot = [q.output_type([t for t, _, _ in ins], 0)]
        for launch in (None, {"block_size": 64}, {"spill_wide": True}):
            if launch:
                q.set_launch(**launch)
            out = DeviceColumn.empty(ot[0], n)
            print(f"from_arrays program with top slot {top}, launch {launch}", flush=True)
            perform_interpreter_operation(q, device_columns(ins), [out], n)
            (ov, oval), = oracle.execute(d, host_columns(ins), ot, n)
            gv, gval = out.to_numpy(n)
            assert_same(gv, gval, ov, oval, what=f"top slot {top}")
        done += 1
        if done == 3:
            break
    assert done >= 1


def test_bare_columns_literals_and_reorders(bsql):
    rng = np.random.default_rng(5)
    n = 4_099
    table = make_table(rng, [INT32, FLOAT64, INT64] * 7, n, [0.1, 0.0, 0.2] * 7)
    # a 20-column reorder: copies only, no interpreter slots
    _execute_vs_oracle(bsql, [f"${(k * 7) % 21}" for k in range(20)], table, n, what="reorder")
    _execute_vs_oracle(bsql, ["$0", "+($1, 2)", "5", "null:DOUBLE", "*($2, $1)", "$20", "2.5:DOUBLE"], table, n,
                       what="mixed roots")


def test_long_lists_split_and_grouped(bsql):
    rng = np.random.default_rng(6)
    n = 3_001
    table = make_table(rng, [FLOAT64, INT64] * 25, n, [0.05, 0.1] * 25)
    exprs = [f"*(${k % 50}, {k})" for k in range(100)]
    p = _execute_vs_oracle(bsql, exprs, table, n, what="100 expressions")
    assert len(p.parts()) >= 2
    # 100 copies and 30 interpreted outputs: one part, three output launches
    exprs = [f"${k % 50}" for k in range(100)] + [f"+($0, {k})" for k in range(30)]
    p = _execute_vs_oracle(bsql, exprs, table, n, what="130 outputs")
    assert p.parts() == []


def test_filter_project_with_split_projection(bsql):
    """One selection, then an output launch per part and output group."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn, filter_project
    import torch
    rng = np.random.default_rng(7)
    n = 20_011
    types = [FLOAT64, INT64] * 25
    table = make_table(rng, types, n, [0.05, 0.1] * 25)
    pred_text = "AND(>($0, -300.5:DECIMAL(4, 1)), <($3, 500000))"
    exprs = [f"*(${k % 50}, {k})" for k in range(90)] + [f"${k}" for k in range(40)] + ["7"]
    pred = bsql.Program.compile(pred_text, types)
    proj = bsql.Program.compile(exprs, types)
    assert len(proj.parts()) >= 2
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    print("filter_project with a split projection of 131 outputs", flush=True)
    filter_project(pred, proj, device_columns(table), outs, cnt, n)
    torch.cuda.synchronize()
    m = int(cnt.item())
    on, fcols = oracle.filter_table(oracle.full_table_program(pred, len(types)), host_columns(table), n)
    assert m == on
    fhost = [oracle.HostColumn(t, v, val) for t, (v, val) in zip(types, fcols)]
    ora = oracle.evaluate(proj, fhost, on)
    for k in range(len(ot)):
        gv, gval = outs[k].to_numpy(m)
        assert_same(gv, gval, ora[k][0], ora[k][1], what=f"output {k}")


def _split_items(body):
    items, depth, cur, quoted = [], 0, "", False
    for c in body:
        if c == "'":
            quoted = not quoted
        if not quoted:
            if c in "([":
                depth += 1
            elif c in ")]":
                depth -= 1
            elif c == "," and depth == 0:
                items.append(cur.strip())
                cur = ""
                continue
        cur += c
    if cur.strip():
        items.append(cur.strip())
    return items


@pytest.mark.parametrize("lineno", [269, 358, 393])
def test_golden_long_projections_through_process_project(bsql, lineno):
    """The TPC-H projections of 17 and 37 items (tests/golden/calcite_plans.txt)
    through process_project on a typed table, against the oracle."""
    import re
    import oracle
    from blazingsql_amd.processor import Table, process_project
    line = open(os.path.join(HERE, "golden", "calcite_plans.txt")).read().splitlines()[lineno - 1]
    body = line[line.index("LogicalProject(") + len("LogicalProject("):line.rindex(")")]
    items = _split_items(body)
    exprs = [it[it.index("=[") + 2:-1] for it in items]
    n_cols = max(int(x) for x in re.findall(r"\$(\d+)", body)) + 1
    rng = np.random.default_rng(lineno)
    n = 2_049
    types = [FLOAT64, INT64, INT32] * (n_cols // 3 + 1)
    types = types[:n_cols]
    table = make_table(rng, types, n, [0.1] * n_cols)
    query = "LogicalProject(" + body + ")"
    print(f"process_project of plan line {lineno}: {len(exprs)} items over {n_cols} columns", flush=True)
    out = process_project(Table.from_columns(device_columns(table)), query)
    assert out.num_rows == n and out.num_columns == len(exprs)
    p = bsql.Program.compile(exprs, types)
    ora = oracle.evaluate(p, host_columns(table), n)
    for k, (g, o) in enumerate(zip(out.to_numpy(), ora)):
        assert_same(g[0], g[1], o[0], o[1], what=f"item {k} ({exprs[k]})")


def test_rand_range_and_parity(bsql):
    """RAND() (BLZ_RND): every row valid, values in (0, 1] (the reference's
    only check is a [0, 1] range, interops_tests.cpp:123-125), distinct
    across rows, and equal to the oracle's stream (Philox keyed by the
    program's seed and the row)."""
    rng = np.random.default_rng(8)
    n = 100_000
    table = make_table(rng, [FLOAT64], n)
    import oracle
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    p = bsql.Program.compile(["RAND()", "+(RAND(), $0)"], [FLOAT64])
    outs = [DeviceColumn.empty(FLOAT64, n) for _ in range(2)]
    perform_interpreter_operation(p, device_columns(table), outs, n)
    v, valid = outs[0].to_numpy(n)
    assert valid.all()
    assert (v > 0).all() and (v <= 1).all()
    assert len(np.unique(v)) > n - 5
    assert abs(v.mean() - 0.5) < 0.01
    ora = oracle.evaluate(p, host_columns(table), n)
    for k in range(2):
        gv, gval = outs[k].to_numpy(n)
        assert_same(gv, gval, ora[k][0], ora[k][1], what=f"RAND output {k}")
