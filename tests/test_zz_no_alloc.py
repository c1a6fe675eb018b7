"""bsql_filter_project_noalloc (the entry point for callers that cannot
allocate): its outputs and counts equal bsql_filter_project's on two calls
in a row, the second over a smaller input; sizes beyond its fixed scratch
and programs it cannot serve fail with an error; and a captured call
replays correctly, with eager work between the replays."""
import numpy as np
import pytest

from tests.helpers import FLOAT64, INT32, INT64, assert_same, device_columns, make_table

pytestmark = pytest.mark.gpu

C2_PRED = "AND(<=($4, 10471), >=($2, 0.05:DECIMAL(3, 2)), <=($2, 0.07:DECIMAL(3, 2)), <($0, 24))"
C2_PROJ = ["*(*($1, -(1, $2)), +(1, $3))", "*($1, $2)"]
C1_PRED = "AND(>($2, 50), <($1, 0.3:DECIMAL(2, 1)))"
C1_PROJ = ["+($0, *($2, 3))", "-(*($1, 2.0:DECIMAL(2, 1)), 1.0:DECIMAL(2, 1))", "$2", "CASE(IS NOT NULL($2), $2, 7)"]


def _c2(n, seed):
    rng = np.random.default_rng(seed)
    return [(INT32, rng.integers(1, 51, n, dtype=np.int32), None), (FLOAT64, rng.uniform(900.0, 105000.0, n), None),
            (FLOAT64, rng.integers(0, 11, n) / 100.0, None), (FLOAT64, rng.integers(0, 9, n) / 100.0, None),
            (INT32, rng.integers(8036, 10562, n, dtype=np.int32), None)]


def _c1(n, seed):
    rng = np.random.default_rng(seed)
    return [(INT64, rng.integers(0, 10**6, n, dtype=np.int64), None), (FLOAT64, rng.random(n), None),
            (INT32, rng.integers(0, 100, n, dtype=np.int32), rng.random(n) >= 0.05)]


def _both(bsql, pred_text, proj_text, table, n):
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project, filter_project_noalloc
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(pred_text, types)
    proj = bsql.Program.compile(proj_text, types)
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    dev = device_columns(table)
    res = []
    for fn in (filter_project, filter_project_noalloc):
        outs = [DeviceColumn.empty(t, n) for t in ot]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        fn(pred, proj, dev, outs, cnt, n)
        torch.cuda.synchronize()
        m = int(cnt.item())
        res.append((m, [o.to_numpy(m) for o in outs]))
    (m0, r0), (m1, r1) = res
    assert m0 == m1
    for k, ((a, av), (b, bv)) in enumerate(zip(r0, r1)):
        assert_same(b, bv, a, av, what=f"output {k}")
    return m0


@pytest.mark.parametrize("query", ["c2", "c1"])
def test_noalloc_equals_filter_project_twice(bsql, query):
    make, pred, proj = (_c2, C2_PRED, C2_PROJ) if query == "c2" else (_c1, C1_PRED, C1_PROJ)
    n1, n2 = 3_000_017, 777_777
    m1 = _both(bsql, pred, proj, make(n1, 1), n1)
    m2 = _both(bsql, pred, proj, make(n2, 2), n2)
    assert m1 > m2 > 0


def test_noalloc_limits(bsql):
    import torch
    from blazingsql_amd._lib import BsqlError
    from blazingsql_amd.interop import DeviceColumn, filter_project_noalloc
    n = 40_000_000
    col = DeviceColumn.from_numpy(INT32, np.zeros(n, dtype=np.int32))
    # selection, row ids and output pass: 4 bytes of row id per row exceed
    # the fixed scratch
    pred = bsql.Program.compile(">($0, 5)", [INT32])
    proj = bsql.Program.compile(["+($0, 1)"], [INT32])
    out = DeviceColumn.empty(INT32, n)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(BsqlError, match="scratch"):
        filter_project_noalloc(pred, proj, [col], [out], cnt, n)
    # the fused pass of a range predicate keeps one word per 2048 rows
    pred.set_launch(fused=True)
    filter_project_noalloc(pred, proj, [col], [out], cnt, n)
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
    strs = DeviceColumn.from_strings(["ab", "cd"])
    pred = bsql.Program.compile("LIKE($0, 'a%')", [23])
    proj = bsql.Program.compile(["CHAR_LENGTH($0)"], [23])
    with pytest.raises(BsqlError, match="string literals"):
        filter_project_noalloc(pred, proj, [strs], [DeviceColumn.empty(INT32, 2)], cnt, 2)


def test_noalloc_replays_in_a_graph(bsql):
    """A captured call: every replay (with eager calls and synchronizations
    between them) equals the eager result."""
    import torch
    from blazingsql_amd.interop import DeviceColumn, filter_project_noalloc
    n = 1_000_003
    table = _c1(n, 3)
    types = [t for t, _, _ in table]
    pred = bsql.Program.compile(C1_PRED, types)
    proj = bsql.Program.compile(C1_PROJ, types)
    ot = [proj.output_type(types, k) for k in range(proj.num_outputs)]
    dev = device_columns(table)
    eager = [DeviceColumn.empty(t, n) for t in ot]
    ecnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    filter_project_noalloc(pred, proj, dev, eager, ecnt, n)
    torch.cuda.synchronize()
    m = int(ecnt.item())
    want = [o.to_numpy(m) for o in eager]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            filter_project_noalloc(pred, proj, dev, outs, cnt, n)
    torch.cuda.synchronize()
    other = _c2(500_000, 9)
    for replay in range(3):
        for o in outs:
            o.data.zero_()
            if o.validity is not None:
                o.validity.zero_()
        cnt.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(cnt.item()) == m, replay
        for k, ((a, av), o) in enumerate(zip(want, outs)):
            b, bv = o.to_numpy(m)
            assert_same(b, bv, a, av, what=f"replay {replay} output {k}")
        # eager work on the same scratch between replays
        _both(bsql, C2_PRED, C2_PROJ, other, 500_000)
