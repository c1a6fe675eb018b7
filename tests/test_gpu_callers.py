"""The interpreter's other callers on the GPU, against the oracle:
aggregation inputs (GroupBy.cpp:226-249, 329-352), a join's post-filter
(BatchJoinProcessing.cpp:481: process_filter over the joined table) and the
skip-data pass over row-group statistics (SkipDataProcessor.cpp:242-346),
with its invariant that no row group holding a row that passes the filter is
skipped."""
import numpy as np
import pytest

from tests.helpers import FLOAT64, INT32, INT64, assert_same, make_table

pytestmark = pytest.mark.gpu


def _table(cols, names=None):
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.processor import Table
    dev = [DeviceColumn.from_numpy(t, v, valid) for t, v, valid in cols]
    return Table.from_columns(dev, names)


def test_aggregation_inputs_vs_oracle(bsql, gpu):
    import oracle
    from blazingsql_amd.processor import AGG_COUNT_ALL, AGG_COUNT_VALID, AGG_OTHER, aggregation_input
    rng = np.random.default_rng(3)
    n = 100_003
    types = [INT32, FLOAT64, INT64, INT32]
    cols = make_table(rng, types, n, null_rates=[0.1, 0.0, 0.3, 0.05])
    t = _table(cols)
    # one column: that column itself, borrowed
    one = aggregation_input(t, "$2", AGG_OTHER)
    v, valid = one.to_numpy()[0]
    assert np.array_equal(v[valid], cols[2][1][cols[2][2]]) and np.array_equal(valid, cols[2][2])
    assert one.column(0).data.data_ptr() == t.column(2).data.data_ptr()
    # COUNT(*): the first column
    star = aggregation_input(t, "", AGG_COUNT_ALL)
    assert star.num_rows == n
    # COUNT($0, $2, $3): valid where all three are (the sum of INTEGER casts)
    cnt = aggregation_input(t, "$0, $2, $3", AGG_COUNT_VALID)
    _, cvalid = cnt.to_numpy()[0]
    both = cols[0][2] & cols[2][2] & cols[3][2]
    assert np.array_equal(cvalid, both)
    assert int(cvalid.sum()) == int(both.sum())
    # an expression: evaluated, as the oracle evaluates it
    for expr in ["+($0, *($3, 2))", "CASE(IS NOT NULL($2), $2, 0)"]:
        ev = aggregation_input(t, expr, AGG_OTHER)
        prog = bsql.Program.compile([expr], types)
        ores = oracle.evaluate(prog, [oracle.HostColumn(tt, vv, va) for tt, vv, va in cols], n)
        g = ev.to_numpy()[0]
        assert_same(g[0], g[1], ores[0][0], ores[0][1], what=expr)


def test_join_post_filter_vs_oracle(bsql, gpu):
    """An inner join's output (left key and value, right key and value),
    then the join's remaining condition over it: process_filter, as
    JoinPartitionKernel applies it."""
    import oracle
    from blazingsql_amd.processor import process_filter
    rng = np.random.default_rng(11)
    nl, nr = 20_000, 15_000
    lk = rng.integers(0, 5000, nl).astype(np.int64)
    la = rng.uniform(-100, 100, nl)
    rk = rng.integers(0, 5000, nr).astype(np.int64)
    rb = rng.integers(-50, 50, nr).astype(np.int32)
    rvalid = rng.random(nr) > 0.1
    order = np.argsort(rk, kind="stable")
    lo = np.searchsorted(rk[order], lk, "left")
    hi = np.searchsorted(rk[order], lk, "right")
    li = np.repeat(np.arange(nl), hi - lo)
    ri = np.concatenate([order[a:b] for a, b in zip(lo, hi)]) if len(li) else np.zeros(0, np.int64)
    cols = [(INT64, lk[li], None), (FLOAT64, la[li], None), (INT64, rk[ri], None), (INT32, rb[ri], rvalid[ri])]
    m = len(li)
    assert m > 10_000
    t = _table(cols, ["l_key", "l_a", "r_key", "r_b"])
    for cond in ["AND(>($1, $3), <>($3, 5))", "OR(<($1, -50.5:DECIMAL(3, 1)), =(MOD($0, 7), $3))"]:
        qp = f"LogicalFilter(condition=[{cond}])"
        got = process_filter(t, qp)
        pred = bsql.Program.compile(cond, [c[0] for c in cols])
        on, ores = oracle.filter_table(oracle.full_table_program(pred, 4),
                                       [oracle.HostColumn(tt, vv, va) for tt, vv, va in cols], m)
        assert got.num_rows == on
        g = got.to_numpy()
        for k in range(4):
            assert_same(g[k][0], g[k][1], ores[k][0], ores[k][1], what=f"{cond} column {k}")


def _row_group_stats(cols, names, group):
    """Statistics of consecutive row groups: min_<i>_<name> / max_<i>_<name>
    per column (over the valid rows), then file and row-group ids."""
    n = len(cols[0][1])
    starts = np.arange(0, n, group)
    stats, snames = [], []
    for i, ((t, v, valid), nm) in enumerate(zip(cols, names)):
        lo, hi = [], []
        for s in starts:
            seg = v[s:s + group]
            if valid is not None:
                seg = seg[valid[s:s + group]]
            lo.append(seg.min() if len(seg) else 0)
            hi.append(seg.max() if len(seg) else 0)
        stats += [(t, np.asarray(lo, dtype=v.dtype), None), (t, np.asarray(hi, dtype=v.dtype), None)]
        snames += [f"min_{i}_{nm}", f"max_{i}_{nm}"]
    g = len(starts)
    stats += [(INT32, np.zeros(g, np.int32), None), (INT32, np.arange(g, dtype=np.int32), None)]
    snames += ["file_handle_index", "row_group_index"]
    return stats, snames, starts


@pytest.mark.parametrize("seed", range(3))
def test_skip_data_vs_oracle_and_never_skips_a_match(bsql, gpu, seed):
    import oracle
    from blazingsql_amd.processor import process_skipdata_for_table, skip_data_filter
    rng = np.random.default_rng(seed)
    n, group = 400_000, 1000
    # sorted-ish columns (as files written in key order) make row groups skippable
    a = np.sort(rng.integers(0, 1_000_000, n)).astype(np.int64)
    b = rng.uniform(0, 100, n)
    c = (np.arange(n) // 7 % 5000).astype(np.int32)
    d = rng.integers(0, 10, n).astype(np.int32)
    cols = [(INT64, a, None), (FLOAT64, b, None), (INT32, c, None), (INT32, d, None)]
    names = ["a", "b", "c", "d"]
    stats, snames, starts = _row_group_stats(cols, names, group)
    meta = _table(stats, snames)
    # the scan reads columns a, b, c (projects=[[0, 1, 2]]); d has no statistics
    meta_wo_d = _table(stats[:6] + stats[8:], snames[:6] + snames[8:])
    lo = int(rng.integers(0, 900_000))
    filters = [f"AND(>=($0, {lo}), <($0, {lo + 50_000}))",
               f"OR(=($0, {int(a[n // 2])}), >($2, 4990))",
               f"AND(<($2, 100), >(+($0, $2), {lo}))",
               "AND(<($1, 50.0:DOUBLE), >($0, 990000))"]
    for f in filters:
        scan = f"BindableTableScan(table=[[main, t]], filters=[[{f}]], projects=[[0, 1, 2]], aliases=[[a, b, c]])"
        for m, has in ((meta, [1, 1, 1]), (meta_wo_d, [1, 1, 1])):
            ids, read_all = process_skipdata_for_table(m, names, scan)
            expr = skip_data_filter(f, has)
            assert not read_all and expr
            kept = set(ids.to_numpy()[1][0].tolist())
            # the oracle: the rewritten filter over the same statistics
            proj = []
            for i in (0, 1, 2):
                proj += stats[2 * i:2 * i + 2]
            prog = bsql.Program.compile([expr], [s[0] for s in proj])
            ores = oracle.evaluate(prog, [oracle.HostColumn(t, v, va) for t, v, va in proj], len(starts))
            okeep = set(np.nonzero(ores[0][0].astype(bool) & ores[0][1])[0].tolist())
            assert kept == okeep, f
            # no row group holding a passing row is skipped
            pred = bsql.Program.compile(f, [INT64, FLOAT64, INT32])
            host = [oracle.HostColumn(t, v, va) for t, v, va in cols[:3]]
            ev = oracle.evaluate(pred, host, n)
            hit = ev[0][0].astype(bool) & ev[0][1]
            need = set(np.unique(np.nonzero(hit)[0] // group).tolist())
            assert need <= kept, (f, sorted(need - kept)[:5])
            assert len(kept) < len(starts) or "OR" in f
    # nothing of the filter survives (a column without statistics under OR)
    scan = "BindableTableScan(table=[[main, t]], filters=[[OR(>($0, 5), <($3, 2))]], projects=[[0, 1, 2, 3]], aliases=[[a, b, c, d]])"
    ids, read_all = process_skipdata_for_table(meta_wo_d, names, scan)
    assert read_all and ids is None
