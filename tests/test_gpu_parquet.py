"""Table scan of Parquet files (blazingsql_amd.parquet_scan): the footer's
row-group statistics feed the skip-data pass, the kept row groups are
decoded, copied and filtered by the scan's pushed-down filter, and the
batches carry the scan's aliases. Checked against the oracle's evaluation of
the same filter over the whole file, with the invariant that no row group
holding a passing row is skipped."""
import numpy as np
import pytest

from tests.helpers import FLOAT64, INT32, INT64, TS_D

pytestmark = pytest.mark.gpu


def _write(path, n, group, seed):
    import pyarrow as pa
    import pyarrow.parquet as pq
    rng = np.random.default_rng(seed)
    a = np.sort(rng.integers(0, 1_000_000, n)).astype(np.int64)          # sorted: row groups skippable
    b = rng.uniform(-100, 100, n)
    b_valid = rng.random(n) > 0.1
    c = rng.integers(-1000, 1000, n).astype(np.int32)
    d = (np.arange(n) // 50 - 20000).astype(np.int32)                     # days, increasing
    s = [("ab" if i % 3 else "xy") + str(i % 97) for i in range(n)]
    t = pa.table({"a": pa.array(a), "b": pa.array(b, mask=~b_valid), "c": pa.array(c),
                  "d": pa.array(d, type=pa.int32()).cast(pa.date32()), "s": pa.array(s)})
    pq.write_table(t, path, row_group_size=group)
    return {"a": (INT64, a, None), "b": (FLOAT64, b, b_valid), "c": (INT32, c, None), "d": (TS_D, d, None)}, s


@pytest.mark.parametrize("filt", [
    "AND(>=($0, {lo}), <($0, {hi}))",
    "AND(<($0, {hi}), >($1, 0.0:DOUBLE))",
    "OR(<($0, {lo}), >($3, 1000))",
    "AND(>($3, -19000), <($3, -18800), IS NOT NULL($1))",
])
def test_parquet_scan_vs_oracle(bsql, gpu, tmp_path, filt):
    import oracle
    from blazingsql_amd.parquet_scan import ParquetScan
    n, group = 200_000, 10_000
    path = str(tmp_path / "t.parquet")
    cols, strings = _write(path, n, group, 5)
    a = cols["a"][1]
    lo, hi = int(a[n // 3]), int(a[n // 3 + 25_000])
    f = filt.format(lo=lo, hi=hi)
    # projects a, b, c, d, s ($0..$4 of the scan), named by the aliases
    scan = (f"BindableTableScan(table=[[main, t]], filters=[[{f}]], projects=[[0, 1, 2, 3, 4]], "
            f"aliases=[[x_a, x_b, x_c, x_d, x_s]])")
    ps = ParquetScan([path], scan)
    out = list(ps.batches())
    assert len(out) == len(ps.read) <= n // group
    # the oracle: the filter over the whole file's columns
    types = [cols[k][0] for k in "abcd"]
    pred = bsql.Program.compile(f, types)
    host = [oracle.HostColumn(*cols[k]) for k in "abcd"]
    ev = oracle.evaluate(pred, host, n)
    keep = ev[0][0].astype(bool) & ev[0][1]
    need = set((np.nonzero(keep)[0] // group).tolist())
    read = {rg for _, rg in ps.read}
    assert need <= read, sorted(need - read)[:5]          # never skips a row group with a match
    if "OR" not in f:
        assert len(read) < n // group                      # the statistics skipped something
    # rows: the passing rows of the row groups read, in file order
    got = [t.to_numpy() for t in out]
    assert all(t.names == ["x_a", "x_b", "x_c", "x_d", "x_s"] for t in out)
    rows = np.nonzero(keep)[0]
    for j, k in enumerate("abcd"):
        vals = np.concatenate([g[j][0] for g in got]) if got else np.zeros(0)
        valid = np.concatenate([g[j][1] for g in got]) if got else np.zeros(0, bool)
        t, v, va = cols[k]
        exp_valid = np.ones(len(rows), bool) if va is None else va[rows]
        assert np.array_equal(valid, exp_valid), k
        assert np.array_equal(vals[valid], v[rows][exp_valid]), k
    svals = [x for g in got for x in g[4][0]]
    assert svals == [strings[i] for i in rows]


def test_parquet_scan_without_filter_reads_every_row_group(bsql, gpu, tmp_path):
    from blazingsql_amd.parquet_scan import ParquetScan
    n, group = 30_000, 4_000
    path = str(tmp_path / "u.parquet")
    cols, strings = _write(path, n, group, 9)
    ps = ParquetScan([path], "BindableTableScan(table=[[main, u]], projects=[[2, 0]], aliases=[[cc, aa]])")
    out = list(ps.batches())
    assert len(ps.read) == (n + group - 1) // group
    c = np.concatenate([t.to_numpy()[0][0] for t in out])
    a = np.concatenate([t.to_numpy()[1][0] for t in out])
    assert np.array_equal(c, cols["c"][1]) and np.array_equal(a, cols["a"][1])
    assert out[0].names == ["cc", "aa"]
