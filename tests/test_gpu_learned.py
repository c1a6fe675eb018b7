"""GPU checks of the learned conjunct ordering: the MLP's forward and
backward kernels against the host reference (oracle/mlp_ref.py), the column
sketch kernel against its host restatement, training on records from the
exact interpreter, and the ordering: bit-exact results in every order and a
measured gain over the static plan order on a correlated table."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mlp(bsql, seed=3):
    from blazingsql_amd.learned import Mlp
    return Mlp(seed)


def test_mlp_forward_and_gradients_vs_reference(bsql, gpu):
    import torch
    from oracle import mlp_ref
    rng = np.random.default_rng(1)
    m = _mlp(bsql)
    p = rng.normal(0, 0.15, m.n_params).astype(np.float32)
    m.set_params(p)
    assert np.array_equal(m.params(), p)
    x = rng.normal(0, 1, (64, 24)).astype(np.float32)
    t = np.stack([rng.random(64), rng.normal(0, 1, 64)], axis=1).astype(np.float32)
    y = m.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    _, _, y_ref = mlp_ref.forward(p, x)
    np.testing.assert_allclose(y, y_ref, rtol=1e-4, atol=1e-5)
    g, loss = m.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), cost_weight=0.5)
    g_ref = mlp_ref.gradients(p, x, t, cost_weight=0.5)
    np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=1e-3, atol=2e-6)
    assert abs(float(loss.mean()) - mlp_ref.loss(p, x, t, 0.5)) < 1e-4


def test_adam_step_moves_against_gradient(bsql, gpu):
    import torch
    rng = np.random.default_rng(2)
    m = _mlp(bsql)
    p0 = m.params()
    x = torch.from_numpy(rng.normal(0, 1, (32, 24)).astype(np.float32)).cuda()
    t = torch.from_numpy(np.stack([rng.random(32), rng.normal(0, 1, 32)], 1).astype(np.float32)).cuda()
    g, _ = m.gradients(x, t)
    m.train_step(x, t, lr=1e-3)
    step = m.params() - p0
    gh = g.cpu().numpy().astype(np.float64)
    # Adam's first step (bias-corrected moments m = g, v = g^2) is
    # -lr * g / (|g| + eps), eps = 1e-8; float32 parameters round the result
    expect = -1e-3 * gh / (np.abs(gh) + 1e-8)
    np.testing.assert_allclose(step, expect, rtol=1e-4, atol=6e-8)


def test_sketch_vs_reference(bsql, gpu):
    from blazingsql_amd.interop import DeviceColumn
    from blazingsql_amd.learned import sketch
    from oracle import mlp_ref
    rng = np.random.default_rng(4)
    n = 100_003
    f = rng.normal(5, 3, n)
    i = rng.integers(-1000, 1000, n).astype(np.int32)
    v = rng.random(n) > 0.2
    cols = [DeviceColumn.from_numpy(10, f), DeviceColumn.from_numpy(3, i, v),
            DeviceColumn.from_numpy(10, np.full(n, 2.5))]
    sk = sketch(cols, n)
    for k, (vals, valid) in enumerate([(f, None), (i, v), (np.full(n, 2.5), None)]):
        ref = mlp_ref.sketch(vals, valid, n)
        np.testing.assert_allclose(sk[k], ref.astype(np.float32), rtol=1e-6, atol=1e-7, err_msg=f"column {k}")


def test_training_predicts_selectivity(bsql, gpu):
    from blazingsql_amd import learned
    X, Y = learned.collect_records(2600, rows=1 << 18, seed=5)
    mlp, hist = learned.train(X[:2200], Y[:2200], epochs=40)
    assert hist[-1] < hist[0] * 0.5
    sel, cost = mlp.predict(X[2200:])
    mae = float(np.mean(np.abs(sel - Y[2200:, 0])))
    assert mae < 0.05, mae


def test_learned_order_exact_and_faster(bsql, gpu):
    """A predicate written with its least selective terms first: the learned
    order (selective, cheap terms first) selects exactly the same rows and
    writes the same outputs, in less time."""
    import torch
    from blazingsql_amd import learned
    from blazingsql_amd.interop import DeviceColumn, filter_project
    X, Y = learned.collect_records(2000, rows=1 << 18, seed=6)
    mlp, _ = learned.train(X, Y, epochs=30)
    n = 1 << 24
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    cols = learned.copula_table(n, 0.5, gen)
    rng = np.random.default_rng(3)
    sels = [0.9, 0.8, 0.7, 0.5, 0.3, 0.002]
    terms = [learned.range_term(rng, c, s) for c, s in zip([1, 5, 2, 6, 3, 7], sels)]
    text = "AND(" + ", ".join(terms) + ")"
    types = learned.COPULA_TYPES
    proj_text = ["+($0, $4)", "*($2, 2:DOUBLE)"]

    def run(order_learned):
        pred = bsql.Program.compile(text, types)
        proj = bsql.Program.compile(proj_text, types)
        if order_learned:
            learned.learned_order(mlp, pred, types, cols, n)
        outs = [DeviceColumn.empty(10, n), DeviceColumn.empty(10, n)]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        for _ in range(3):
            filter_project(pred, proj, cols, outs, cnt, n)
        torch.cuda.synchronize()
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(10):
            filter_project(pred, proj, cols, outs, cnt, n)
        z.record()
        torch.cuda.synchronize()
        m = int(cnt.item())
        return a.elapsed_time(z) / 10, m, [o.to_numpy(m) for o in outs]

    t_static, m_s, out_s = run(False)
    t_learned, m_l, out_l = run(True)
    assert m_s == m_l
    for (a, va), (b, vb) in zip(out_s, out_l):
        np.testing.assert_array_equal(va, vb)
        np.testing.assert_array_equal(a.view(np.int64), b.view(np.int64))
    assert t_learned < 0.8 * t_static, (t_learned, t_static)


def test_filter_project_sized_overflow(bsql, gpu):
    """Outputs pre-sized below the count: the first `capacity` rows are
    written and the full count is reported, so the caller can retry."""
    import ctypes
    import torch
    from blazingsql_amd import _lib as L
    from blazingsql_amd.interop import DeviceColumn, columns_array, filter_project
    rng = np.random.default_rng(8)
    n = 200_000
    x = rng.integers(0, 1000, n).astype(np.int64)
    cols = [DeviceColumn.from_numpy(4, x)]
    pred = bsql.Program.compile(">($0, 500)", [4])
    proj = bsql.Program.compile(["*($0, 3)"], [4])
    full = [DeviceColumn.empty(4, n)]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    filter_project(pred, proj, cols, full, cnt, n)
    m = int(cnt.item())
    cap = m // 3
    small = [DeviceColumn.empty(4, cap)]
    cnt2 = torch.zeros(1, dtype=torch.int64, device="cuda")
    L.check(L.lib().bsql_filter_project_sized(pred.handle, proj.handle, columns_array(cols), 1,
                                              columns_array(small), 1, n, cap,
                                              ctypes.c_void_p(cnt2.data_ptr()), None))
    torch.cuda.synchronize()
    assert int(cnt2.item()) == m
    np.testing.assert_array_equal(small[0].to_numpy(cap)[0], full[0].to_numpy(cap)[0])
    # capacities around the count (the output launch is sized by the
    # capacity): every row below min(count, capacity) is written
    from blazingsql_amd.interop import filter_project_sized
    pred2 = bsql.Program.compile(">($0, 995)", [4])
    filter_project(pred2, proj, cols, full, cnt, n)
    m2 = int(cnt.item())
    for cap2 in (0, 1, m2 - 1, m2, m2 + 1, 4 * m2 + 3):
        out2 = [DeviceColumn.empty(4, max(cap2, 1))]
        cnt3 = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        filter_project_sized(pred2, proj, cols, out2, cnt3, cap2, n)
        assert int(cnt3.item()) == m2, cap2
        w = min(m2, cap2)
        np.testing.assert_array_equal(out2[0].to_numpy(w)[0], full[0].to_numpy(w)[0])


@pytest.fixture(scope="module")
def trained_mlp(bsql, gpu):
    from blazingsql_amd import learned
    X, Y = learned.collect_records(1500, rows=1 << 17, seed=21)
    mlp, _ = learned.train(X, Y, epochs=20)
    return mlp


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9])
def test_copula_learned_filter_project_vs_oracle(bsql, gpu, trained_mlp, rho):
    """The learned-ordering workload's shape at 2M rows: Gaussian-copula
    tables (8 columns, pairwise rho), predicates of 4-8 range terms whose
    joint selectivity spans 0.001-0.9, each run in the learned form
    (predicted order, outputs written up to the predicted row bound, a
    second call on overflow; half of them through the generated one-pass
    kernel) and checked bit for bit against the oracle running the static
    plan order."""
    import torch
    import oracle
    from blazingsql_amd import learned
    from blazingsql_amd.interop import DeviceColumn
    n = 2_000_003
    gen = torch.Generator(device="cuda")
    gen.manual_seed(31 + int(rho * 10))
    cols = learned.copula_table(n, rho, gen)
    host = [oracle.HostColumn(c.type_id, c.to_numpy(n)[0]) for c in cols]
    types = learned.COPULA_TYPES
    proj_text = ["+($0, $4)", "*($2, 2:DOUBLE)"]
    rng = np.random.default_rng(int(rho * 10) + 5)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    reordered, selected = 0, []
    for i in range(6):
        text, terms, target = learned.copula_predicate(rng, int(rng.integers(4, 9)))
        print(f"rho={rho} predicate {i}: {len(terms)} terms, target selectivity {target:.4f}", flush=True)
        pred = bsql.Program.compile(text, types)
        pred.set_launch(generated=bool(i % 2))
        proj = bsql.Program.compile(proj_text, types)
        outs = [DeviceColumn.empty(10, n), DeviceColumn.empty(10, n)]
        m, _, order = learned.filter_project_learned(trained_mlp, pred, proj, types, cols, outs, cnt, n)
        static = bsql.Program.compile(text, types)
        om, ores = oracle.filter_project(oracle.full_table_program(static, 8), oracle.full_table_program(proj, 8),
                                         host, [10, 10], n)
        assert m == om, (text, m, om)
        for k in range(2):
            gv, gval = outs[k].to_numpy(m)
            assert np.array_equal(gval, ores[k][1]), f"validity of output {k}"
            assert np.array_equal(gv.view(np.uint64), ores[k][0].view(np.uint64)), f"output {k} of {text}"
        selected.append(m / n)
        reordered += int(list(order) != list(range(len(terms))))
    assert max(selected) > 0.01, selected  # the workload selects rows, not only nothing
    assert reordered > 0


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9])
def test_copula_learned_batch_vs_oracle(bsql, gpu, trained_mlp, rho):
    """The learned and batched form at 2M rows: a copula table's 8 queries
    (4-8 range terms each, joint selectivity 0.001-0.9) in ONE
    bsql_filter_project_batch pass with the learned conjunct orders and the
    predicted row bounds (a query over its bound runs again on its own);
    every query's count, outputs and validity bit for bit against the oracle
    running it in the static order."""
    import torch
    import oracle
    from blazingsql_amd import learned
    from blazingsql_amd.interop import DeviceColumn
    n = 2_000_003
    gen = torch.Generator(device="cuda")
    gen.manual_seed(71 + int(rho * 10))
    cols = learned.copula_table(n, rho, gen)
    host = [oracle.HostColumn(c.type_id, c.to_numpy(n)[0]) for c in cols]
    types = learned.COPULA_TYPES
    proj_text = ["+($0, $4)", "*($2, 2:DOUBLE)"]
    rng = np.random.default_rng(int(rho * 10) + 41)
    texts = [learned.copula_predicate(rng, int(rng.integers(4, 9)))[0] for _ in range(8)]
    print(f"rho={rho}: batch of {len(texts)} queries", flush=True)
    preds = [bsql.Program.compile(t, types) for t in texts]
    projs = [bsql.Program.compile(proj_text, types) for _ in texts]
    outs = [[DeviceColumn.empty(10, n), DeviceColumn.empty(10, n)] for _ in texts]
    bcnt = torch.zeros(len(texts), dtype=torch.int64, device="cuda")
    counts, _ = learned.filter_project_batch_learned(trained_mlp, preds, projs, types, cols, outs, bcnt, n)
    proj = bsql.Program.compile(proj_text, types)
    for q, text in enumerate(texts):
        static = bsql.Program.compile(text, types)
        om, ores = oracle.filter_project(oracle.full_table_program(static, 8), oracle.full_table_program(proj, 8),
                                         host, [10, 10], n)
        assert counts[q] == om, (text, counts[q], om)
        for k in range(2):
            gv, gval = outs[q][k].to_numpy(om)
            assert np.array_equal(gval, ores[k][1]), f"validity of output {k} of {text}"
            assert np.array_equal(gv.view(np.uint64), ores[k][0].view(np.uint64)), f"output {k} of {text}"
    assert max(counts) > 0.01 * n, counts


def test_copula_learned_batch_full_size_matches_calls(bsql, gpu, trained_mlp):
    """The learned and batched form at its full size (64M rows, rho 0.9):
    the one-pass batch selects and writes exactly what each query's own
    bsql_filter_project call in the static order does (batching must not
    change any row, at a size where the oracle is too slow to run)."""
    import torch
    from blazingsql_amd import learned
    from blazingsql_amd.interop import DeviceColumn, filter_project
    n = 1 << 26
    gen = torch.Generator(device="cuda")
    gen.manual_seed(97)
    cols = learned.copula_table(n, 0.9, gen)
    types = learned.COPULA_TYPES
    proj_text = ["+($0, $4)", "*($2, 2:DOUBLE)"]
    rng = np.random.default_rng(97)
    texts = [learned.copula_predicate(rng, int(rng.integers(4, 9)))[0] for _ in range(8)]
    preds = [bsql.Program.compile(t, types) for t in texts]
    projs = [bsql.Program.compile(proj_text, types) for _ in texts]
    outs = [[DeviceColumn.empty(10, n), DeviceColumn.empty(10, n)] for _ in texts]
    bcnt = torch.zeros(len(texts), dtype=torch.int64, device="cuda")
    counts, _ = learned.filter_project_batch_learned(trained_mlp, preds, projs, types, cols, outs, bcnt, n)
    ref = [DeviceColumn.empty(10, n), DeviceColumn.empty(10, n)]
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    proj = bsql.Program.compile(proj_text, types)
    for q, text in enumerate(texts):
        filter_project(bsql.Program.compile(text, types), proj, cols, ref, cnt, n)
        m = int(cnt.item())
        assert counts[q] == m, (text, counts[q], m)
        for k in range(2):
            assert torch.equal(outs[q][k].data[:m].view(torch.int64), ref[k].data[:m].view(torch.int64)), (q, k)
    assert max(counts) > 0.01 * n, counts
