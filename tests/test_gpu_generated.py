"""GPU parity of kernels generated for one program (BSQL_LAUNCH_GENERATED):
the same calls with and without the flag give bit-identical outputs, and
both match the oracle. Covers execute over random programs (nulls, casts,
CASE, COALESCE, EXTRACT), the one-pass filter + project at sizes around the
2048-row tile and validity-word boundaries, capacity-limited outputs,
timestamp ranges compared across units, string outputs, calls that fall
back to the interpreter, and the kernel cache (a literal change compiles
nothing new)."""
import numpy as np
import pytest

from tests.helpers import (BOOL8, FLOAT64, INT32, INT64, TS_D, ExprGen, assert_same, device_columns,
                           host_columns, make_table)

pytestmark = pytest.mark.gpu

TYPES6 = [INT32, INT64, FLOAT64, FLOAT64, TS_D, INT32, BOOL8]


def _stats():
    from blazingsql_amd.interop import generated_kernel_stats
    return generated_kernel_stats()


def _execute(bsql, exprs, table, n, generated):
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    types = [t for t, _, _ in table]
    p = bsql.Program.compile(exprs, types)
    p.set_launch(generated=generated)
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    outs = [DeviceColumn.empty(t, max(n, 1)) for t in ot]
    perform_interpreter_operation(p, device_columns(table), outs, n)
    return p, [o.to_numpy(n) for o in outs]


@pytest.mark.parametrize("seed", range(4))
def test_generated_execute_random_programs(bsql, gpu, seed):
    import oracle
    rng = np.random.default_rng(900 + seed)
    n = [4099, 70001, 1, 262147][seed]
    table = make_table(rng, TYPES6, n, null_rates=list(rng.uniform(0, 0.3, len(TYPES6))))
    gen = ExprGen(rng, TYPES6)
    exprs = []
    while len(exprs) < 6:
        e = gen.boolean(3) if rng.random() < 0.4 else gen.numeric(4)
        if not e or "$" not in e:
            continue
        try:
            bsql.Program.compile([e], TYPES6)
        except Exception:
            continue
        exprs.append(e)
    k0, l0, _ = _stats()
    print(f"seed {seed} n={n}: {len(exprs)} expressions, generated", flush=True)
    p, got = _execute(bsql, exprs, table, n, True)
    k1, l1, _ = _stats()
    assert l1 > l0, "the generated kernel did not run"
    _, ref = _execute(bsql, exprs, table, n, False)
    ora = oracle.evaluate(p, host_columns(table), n)
    for k, e in enumerate(exprs):
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"generated vs interpreter: {e}")
        assert_same(got[k][0], got[k][1], ora[k][0], ora[k][1], what=f"generated vs oracle: {e}")


SIZES = [0, 1, 31, 255, 2047, 2048, 2049, 8193, 3 * 2048 + 1000, 250_003, 1_000_000]
PRED = "AND(>($0, -500), <($3, 800), >=($1, -900.5:DECIMAL(4, 1)))"
PROJ = ["+($2, *($0, 3))", "*($1, 2.5:DECIMAL(2, 1))", "$3", "CASE(IS NOT NULL($2), $2, 0)",
        "/($1, $2)"]


def _filter_project(bsql, table, n, generated, pred_text=PRED, proj_exprs=PROJ, cap=None):
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project, filter_project_sized
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(pred_text, types)
    pred.set_launch(generated=generated)
    proj = bsql.Program.compile(proj_exprs, types)
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    rows = max(n if cap is None else cap, 1)
    outs = [DeviceColumn.empty(t, rows) for t in ot]
    cnt = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    if cap is None:
        filter_project(pred, proj, device_columns(table), outs, cnt, n)
    else:
        filter_project_sized(pred, proj, device_columns(table), outs, cnt, cap, n)
    torch.cuda.synchronize()
    m = int(cnt.item())
    k = m if cap is None else min(m, cap)
    return pred, proj, ot, m, [o.to_numpy(k) for o in outs]


@pytest.mark.parametrize("n", SIZES)
def test_generated_filter_project_vs_interpreter_and_oracle(bsql, gpu, n):
    import oracle
    rng = np.random.default_rng(n + 11)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, max(n, 1), null_rates=[0.1, 0.0, 0.3, 0.05])
    table = [(t, v[:n], None if val is None else val[:n]) for t, v, val in table]
    print(f"filter_project n={n} generated", flush=True)
    _, l0, _ = _stats()
    pred, proj, ot, m, got = _filter_project(bsql, table, n, True)
    _, l1, _ = _stats()
    if n:
        assert l1 > l0, "the generated kernel did not run"
    _, _, _, m2, ref = _filter_project(bsql, table, n, False)
    om, ores = oracle.filter_project(oracle.full_table_program(pred, 4), oracle.full_table_program(proj, 4),
                                     host_columns(table), ot, n)
    assert m == m2 == om, (m, m2, om)
    for k in range(len(PROJ)):
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"generated vs two-pass output {k}")
        assert_same(got[k][0], got[k][1], ores[k][0], ores[k][1], what=f"generated vs oracle output {k}")


@pytest.mark.parametrize("n", [5000, 100_003])
def test_generated_filter_project_capacity(bsql, gpu, n):
    """Pre-sized outputs: the first `cap` selected rows written, the full count reported."""
    rng = np.random.default_rng(5)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    _, _, _, m_all, full = _filter_project(bsql, table, n, False)
    cap = m_all // 2 + 7
    _, _, _, m, got = _filter_project(bsql, table, n, True, cap=cap)
    assert m == m_all
    for k in range(len(PROJ)):
        assert_same(got[k][0], got[k][1], full[k][0][:cap], full[k][1][:cap], what=f"sized output {k}")


# selective predicates: the one-pass kernel stages each tile's rows in LDS
# (with five outputs, up to 768 of a 4096-row tile) and writes them as one
# run; at about 18 % selected, tiles on either side of that limit mix the
# staged and the direct write in one call
SELECTIVE = {"3%": "AND(>($0, 900), <($3, 800), >=($1, -900.5:DECIMAL(4, 1)))",
             "18%": "AND(>($0, 500), <($3, 800), >=($1, -900.5:DECIMAL(4, 1)))"}


@pytest.mark.parametrize("sel", sorted(SELECTIVE))
@pytest.mark.parametrize("n", [2049, 8193, 1_000_003])
def test_generated_filter_project_staged_rows(bsql, gpu, sel, n):
    """Outputs with nulls (a 30 %-null column, division by zero) written from
    LDS-staged tiles: generated against the two-pass interpreter and the oracle."""
    import oracle
    rng = np.random.default_rng(n + 3)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    pred_text = SELECTIVE[sel]
    print(f"staged filter_project sel={sel} n={n}", flush=True)
    pred, proj, ot, m, got = _filter_project(bsql, table, n, True, pred_text)
    _, _, _, m2, ref = _filter_project(bsql, table, n, False, pred_text)
    om, ores = oracle.filter_project(oracle.full_table_program(pred, 4), oracle.full_table_program(proj, 4),
                                     host_columns(table), ot, n)
    assert m == m2 == om, (m, m2, om)
    assert m > 0
    for k in range(len(PROJ)):
        if k in (0, 4):  # +($2, ...) and /($1, $2) over the 30 %-null $2
            assert not got[k][1].all(), f"output {k} has nulls in the test data"
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"generated vs two-pass output {k}")
        assert_same(got[k][0], got[k][1], ores[k][0], ores[k][1], what=f"generated vs oracle output {k}")


@pytest.mark.parametrize("sel", sorted(SELECTIVE))
def test_generated_filter_project_staged_capacity(bsql, gpu, sel):
    """Pre-sized outputs cut inside a staged tile: the first `cap` rows written,
    the full count reported."""
    n = 300_007
    rng = np.random.default_rng(17)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    _, _, _, m_all, full = _filter_project(bsql, table, n, False, SELECTIVE[sel])
    for cap in (m_all // 3 + 5, m_all - 1):
        _, _, _, m, got = _filter_project(bsql, table, n, True, SELECTIVE[sel], cap=cap)
        assert m == m_all
        for k in range(len(PROJ)):
            assert_same(got[k][0], got[k][1], full[k][0][:cap], full[k][1][:cap], what=f"sized output {k}")


def test_generated_filter_project_scan_shape(bsql, gpu):
    """The 100M-row scan's predicate and projection (range tests on shipdays,
    disc and qty; price * (1 - disc) * (1 + tax) and price * disc) on 3M
    rows: generated against the two-pass interpreter and the oracle."""
    import oracle
    rng = np.random.default_rng(2024)
    n = 3_000_011
    qty = rng.integers(1, 51, n).astype(np.int32)
    price = rng.uniform(900, 105000, n)
    disc = rng.integers(0, 11, n) / 100.0
    tax = rng.integers(0, 9, n) / 100.0
    ship = rng.integers(8036, 10562, n).astype(np.int32)
    table = [(INT32, qty, None), (FLOAT64, price, None), (FLOAT64, disc, None), (FLOAT64, tax, None),
             (INT32, ship, None)]
    pred_text = "AND(<=($4, 10471), >=($2, 0.05:DECIMAL(3, 2)), <=($2, 0.07:DECIMAL(3, 2)), <($0, 24))"
    proj_exprs = ["*(*($1, -(1, $2)), +(1, $3))", "*($1, $2)"]
    pred, proj, ot, m, got = _filter_project(bsql, table, n, True, pred_text, proj_exprs)
    _, _, _, m2, ref = _filter_project(bsql, table, n, False, pred_text, proj_exprs)
    sel = (ship <= 10471) & (disc >= 0.05) & (disc <= 0.07) & (qty < 24)
    assert m == m2 == int(sel.sum())
    np.testing.assert_array_equal(got[0][0], price[sel] * (1 - disc[sel]) * (1 + tax[sel]))
    np.testing.assert_array_equal(got[1][0], price[sel] * disc[sel])
    for k in range(2):
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"scan output {k}")
    om, ores = oracle.filter_project(oracle.full_table_program(pred, 5), oracle.full_table_program(proj, 5),
                                     host_columns([(t, v[:200_000], None) for t, v, _ in table]), ot, 200_000)
    head = int(sel[:200_000].sum())
    assert om == head
    for k in range(2):
        assert_same(got[k][0][:head], got[k][1][:head], ores[k][0], ores[k][1], what=f"scan oracle {k}")


def test_generated_timestamp_ranges_across_units(bsql, gpu):
    """Range tests whose column is compared in nanoseconds (unit factor in
    the generated kernel) against the interpreter and the oracle."""
    import oracle
    rng = np.random.default_rng(8)
    n = 50_001
    ts_ms = rng.integers(0, 2 * 10**12, n).astype(np.int64)  # TIMESTAMP_MILLISECONDS
    x = rng.uniform(-10, 10, n)
    valid = rng.random(n) >= 0.1
    table = [(14, ts_ms, valid), (FLOAT64, x, None)]
    pred_text = "AND(>=($0, 1990-01-01 00:00:00), <($0, 2020-06-01 12:00:00), >($1, -5))"
    proj_exprs = ["*($1, 2)", "CAST($0):DATE"]
    pred, proj, ot, m, got = _filter_project(bsql, table, n, True, pred_text, proj_exprs)
    _, _, _, m2, ref = _filter_project(bsql, table, n, False, pred_text, proj_exprs)
    om, ores = oracle.filter_project(oracle.full_table_program(pred, 2), oracle.full_table_program(proj, 2),
                                     host_columns(table), ot, n)
    assert m == m2 == om and m > 0
    for k in range(2):
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"ts output {k}")
        assert_same(got[k][0], got[k][1], ores[k][0], ores[k][1], what=f"ts oracle {k}")


def test_generated_string_outputs_and_fallbacks(bsql, gpu):
    """A string-building projection runs through the generated execute
    kernel; a LIKE predicate (not range tests) and a RAND projection run on
    the interpreter under the same flag, with the same results."""
    import oracle
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project, perform_interpreter_operation
    rng = np.random.default_rng(3)
    n = 20_000
    words = ["abc", "abxyz", "hello", "", "ab", "zzzzzzzz", "Brand#45"]
    strs = [words[int(i)] for i in rng.integers(0, len(words), n)]
    sval = rng.random(n) >= 0.1
    x = rng.integers(-100, 100, n).astype(np.int64)
    types = [23, INT64]
    dev = [DeviceColumn.from_strings(strs, sval), DeviceColumn.from_numpy(INT64, x)]
    host = [oracle.HostColumn(23, strings=strs, valid=sval), oracle.HostColumn(INT64, x)]
    exprs = ["SUBSTRING($0, 2, 3)", "CHAR_LENGTH($0)", "+($1, 7)"]
    res = []
    for generated in (True, False):
        p = bsql.Program.compile(exprs, types)
        p.set_launch(generated=generated)
        ot = [p.output_type(types, k) for k in range(p.num_outputs)]
        chars = p.output_chars(dev, n)
        outs = [DeviceColumn.empty(t, n, chars=chars) for t in ot]
        perform_interpreter_operation(p, dev, outs, n)
        res.append([o.to_numpy(n) for o in outs])
    ora = oracle.evaluate(p, host, n)
    for k in range(3):
        gv, gval = res[0][k]
        rv, rval = res[1][k]
        assert np.array_equal(np.asarray(gval, bool), np.asarray(rval, bool))
        assert np.array_equal(np.asarray(gval, bool), np.asarray(ora[k][1], bool))
        keep = np.asarray(gval, bool)
        assert [a for a, ok in zip(gv, keep) if ok] == [b for b, ok in zip(rv, keep) if ok]
        assert [a for a, ok in zip(gv, keep) if ok] == [b for b, ok in zip(ora[k][0], keep) if ok]
    # fallbacks: LIKE predicate, RAND output
    for pred_text, proj_exprs in (("LIKE($0, 'ab%')", ["+($1, 1)"]), (">($1, 0)", ["RAND()", "$1"])):
        got = []
        for generated in (True, False):
            pred = bsql.Program.compile(pred_text, types)
            pred.set_launch(generated=generated)
            proj = bsql.Program.compile(proj_exprs, types)
            ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
            outs = [DeviceColumn.empty(t, n) for t in ot]
            cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
            filter_project(pred, proj, dev, outs, cnt, n)
            torch.cuda.synchronize()
            m = int(cnt.item())
            got.append((m, [o.to_numpy(m) for o in outs]))
        assert got[0][0] == got[1][0]
        for k in range(len(proj_exprs)):
            assert_same(got[0][1][k][0], got[0][1][k][1], got[1][1][k][0], got[1][1][k][1], what=pred_text)


def test_generated_kernel_cache_reuses_shapes(bsql, gpu):
    """Programs that differ only in literals share a generated kernel."""
    rng = np.random.default_rng(4)
    n = 10_000
    table = make_table(rng, [INT32, FLOAT64], n, null_rates=[0.1, 0.2])
    _execute(bsql, ["+(*($0, 3), 5)", "*($1, 1.5:DECIMAL(2, 1))"], table, n, True)
    k0, _, _ = _stats()
    p, got = _execute(bsql, ["+(*($0, 7), -2)", "*($1, 4.25:DECIMAL(3, 2))"], table, n, True)
    k1, _, _ = _stats()
    assert k1 == k0, "a literal change compiled a new kernel"
    _, ref = _execute(bsql, ["+(*($0, 7), -2)", "*($1, 4.25:DECIMAL(3, 2))"], table, n, False)
    for k in range(2):
        assert_same(got[k][0], got[k][1], ref[k][0], ref[k][1], what=f"reused kernel output {k}")


BATCH_QUERIES = [
    ("AND(>($0, 900), <($3, 800), >=($1, -900.5:DECIMAL(4, 1)))", PROJ),
    ("<($1, -400.0:DECIMAL(5, 1))", ["*($2, 3)", "+($1, $0)"]),
    ("AND(>=($3, -200), <=($3, 200))", ["/($1, $2)"]),
    (">($2, 0)", ["$0", "CASE(IS NOT NULL($3), $3, -1)"]),
]


def _batch_run(bsql, table, n, queries, caps=None):
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project_batch
    types = [t for t, _, _ in table]
    preds = [bsql.Program.compile(p, types) for p, _ in queries]
    projs = [bsql.Program.compile(j, types) for _, j in queries]
    ots = [[pj.output_type(types, k) for k in range(pj.num_outputs)] for pj in projs]
    outs = [[DeviceColumn.empty(t, max(n, 1)) for t in ot] for ot in ots]
    cnt = torch.full((len(queries),), -7, dtype=torch.int64, device="cuda")
    filter_project_batch(preds, projs, device_columns(table), outs, cnt, n, capacities=caps)
    torch.cuda.synchronize()
    counts = [int(c) for c in cnt.cpu()]
    got = []
    for q in range(len(queries)):
        k = counts[q] if caps is None else min(counts[q], caps[q])
        got.append([o.to_numpy(k) for o in outs[q]])
    return preds, projs, ots, counts, got


@pytest.mark.parametrize("n", [31, 8193, 1_000_003])
def test_filter_project_batch_vs_calls_and_oracle(bsql, gpu, n):
    """Four queries over one table in one generated launch: each query's
    count and outputs equal its own bsql_filter_project call's and the
    oracle's."""
    import oracle
    rng = np.random.default_rng(n + 41)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    print(f"batch n={n}", flush=True)
    _, l0, _ = _stats()
    preds, projs, ots, counts, got = _batch_run(bsql, table, n, BATCH_QUERIES)
    _, l1, _ = _stats()
    assert l1 > l0, "the generated kernel did not run"
    for q, (pt, jt) in enumerate(BATCH_QUERIES):
        _, _, _, m, ref = _filter_project(bsql, table, n, False, pt, jt)
        om, ores = oracle.filter_project(oracle.full_table_program(preds[q], 4),
                                         oracle.full_table_program(projs[q], 4), host_columns(table), ots[q], n)
        assert counts[q] == m == om, (q, counts[q], m, om)
        for k in range(len(jt)):
            assert_same(got[q][k][0], got[q][k][1], ref[k][0], ref[k][1], what=f"query {q} output {k} vs call")
            assert_same(got[q][k][0], got[q][k][1], ores[k][0], ores[k][1], what=f"query {q} output {k} vs oracle")


def test_filter_project_batch_capacities(bsql, gpu):
    """Per-query capacities: the first capacity rows written, full counts."""
    n = 300_007
    rng = np.random.default_rng(77)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    _, _, _, full_counts, full = _batch_run(bsql, table, n, BATCH_QUERIES)
    caps = [c // 2 + 3 for c in full_counts]
    _, _, _, counts, got = _batch_run(bsql, table, n, BATCH_QUERIES, caps=caps)
    assert counts == full_counts
    for q in range(len(BATCH_QUERIES)):
        for k in range(len(got[q])):
            assert_same(got[q][k][0], got[q][k][1], full[q][k][0][:caps[q]], full[q][k][1][:caps[q]],
                        what=f"sized query {q} output {k}")


def test_filter_project_batch_falls_back_per_query(bsql, gpu):
    """A query whose predicate is not range tests: the batch runs each
    query on its own, with the same results."""
    import oracle
    n = 50_001
    rng = np.random.default_rng(5)
    types = [INT32, FLOAT64, INT64, INT32]
    table = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    queries = BATCH_QUERIES[:2] + [(">(*($0, 2), $3)", ["+($0, 1)"])]
    preds, projs, ots, counts, got = _batch_run(bsql, table, n, queries)
    for q in range(len(queries)):
        om, ores = oracle.filter_project(oracle.full_table_program(preds[q], 4),
                                         oracle.full_table_program(projs[q], 4), host_columns(table), ots[q], n)
        assert counts[q] == om
        for k in range(len(got[q])):
            assert_same(got[q][k][0], got[q][k][1], ores[k][0], ores[k][1], what=f"fallback query {q} output {k}")
