"""Learned conjunct ordering: sketches, the FP32 MLP (trained on the GPU by
the library's own forward/backward/Adam kernels), training data recorded
from the exact interpreter on synthetic batches, and the ordering it drives.

The static plan order is the fallback and the reference result: an order only permutes
the conjuncts of an AND chain, so the selected rows never depend on the
model; a wrong prediction costs time, not correctness.
"""
import ctypes
import time

import numpy as np

from . import _lib as L
from .interop import DeviceColumn, Program, columns_array, _stream_ptr

FEATURE_DIM, SKETCH_DIM = L.FEATURE_DIM, L.SKETCH_DIM


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Mlp:
    """BSQL_FEATURE_DIM -> 64 -> 64 -> 2 ReLU network; output 0 is the
    selectivity logit, output 1 the log cost."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        L.check(L.lib().bsql_mlp_create(seed, ctypes.byref(h)))
        self.handle = h.value
        self.n_params = L.lib().bsql_mlp_num_params(self.handle)

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        L.check(L.lib().bsql_mlp_get_params(self.handle, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        L.check(L.lib().bsql_mlp_set_params(self.handle, p.ctypes.data))

    def forward(self, x, stream=None):
        """x: float32 CUDA tensor (batch, FEATURE_DIM) -> (batch, 2)."""
        import torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], 2), dtype=torch.float32, device=x.device)
        L.check(L.lib().bsql_mlp_forward(self.handle, _ptr(x), x.shape[0], _ptr(y), _stream_ptr(stream)))
        return y

    def gradients(self, x, target, cost_weight=1.0, stream=None):
        import torch
        x, target = x.contiguous(), target.contiguous()
        g = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        loss = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        L.check(L.lib().bsql_mlp_gradients(self.handle, _ptr(x), _ptr(target), x.shape[0], cost_weight,
                                           _ptr(g), _ptr(loss), _stream_ptr(stream)))
        return g, loss

    def train_step(self, x, target, lr, cost_weight=1.0, stream=None):
        import torch
        loss = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        L.check(L.lib().bsql_mlp_train_step(self.handle, _ptr(x), _ptr(target), x.shape[0], lr, cost_weight,
                                            _ptr(loss), _stream_ptr(stream)))
        return loss

    def predict(self, features):
        """(selectivity, cost) arrays for host features (n, FEATURE_DIM)."""
        import torch
        x = torch.from_numpy(np.ascontiguousarray(features, dtype=np.float32)).cuda()
        y = self.forward(x).cpu().numpy()
        return 1.0 / (1.0 + np.exp(-y[:, 0])), np.exp(y[:, 1])

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(L, "lib", None)   # module globals are gone at interpreter exit
        exiting = getattr(L, "exiting", None)
        if h and lib is not None and exiting is not None and not exiting():
            lib().bsql_mlp_free(h)
            self.handle = None


def sketch(cols, num_rows, stream=None):
    """Host (n_cols, SKETCH_DIM) float32 sketch of device columns."""
    import torch
    out = torch.empty(max(len(cols), 1) * SKETCH_DIM, dtype=torch.float32, device="cuda")
    L.check(L.lib().bsql_sketch(columns_array(cols), len(cols), num_rows, _ptr(out), _stream_ptr(stream)))
    return out[:len(cols) * SKETCH_DIM].cpu().numpy().reshape(len(cols), SKETCH_DIM)


def conjunct_features(pred, types, sketch_host):
    t = (ctypes.c_int32 * len(types))(*types)
    sk = np.ascontiguousarray(sketch_host, dtype=np.float32)
    n = L.lib().bsql_conjunct_features(pred.handle, t, len(types), sk.ctypes.data, None, 0)
    if n < 0:
        L.check(-n)
    f = np.zeros((max(n, 1), FEATURE_DIM), dtype=np.float32)
    n2 = L.lib().bsql_conjunct_features(pred.handle, t, len(types), sk.ctypes.data, f.ctypes.data, n)
    if n2 != n:
        L.check(L.ERR_INVALID)
    return f[:n]


def order_by(pred, selectivity, cost):
    sel = np.ascontiguousarray(selectivity, dtype=np.float32)
    c = np.ascontiguousarray(cost, dtype=np.float32)
    L.check(L.lib().bsql_order_by_prediction(pred.handle, sel.ctypes.data, c.ctypes.data, len(sel)))


def order_for_scan(pred, types, selectivity, cost):
    """Sets the order of least expected cost for the range-test selection
    (bsql_order_conjuncts) and returns it."""
    sel = np.ascontiguousarray(selectivity, dtype=np.float32)
    c = np.ascontiguousarray(cost, dtype=np.float32)
    t = np.ascontiguousarray(types, dtype=np.int32)
    out = np.zeros(len(sel), dtype=np.int32)
    L.check(L.lib().bsql_order_conjuncts(pred.handle, t.ctypes.data, len(t), sel.ctypes.data, c.ctypes.data,
                                         len(sel), out.ctypes.data))
    return [int(x) for x in out]


def row_bound(selectivity, num_rows, slack=1.25, floor=4096):
    """Output rows to pre-size a conjunction's compaction for: it passes at
    most the rows of its most selective term, so the smallest predicted term
    selectivity (with slack) bounds the count whatever the correlation."""
    s = float(np.min(np.asarray(selectivity, dtype=np.float64))) if len(selectivity) else 1.0
    return int(min(num_rows, np.ceil(s * slack * num_rows) + floor))


def learned_order(mlp, pred, types, cols, num_rows):
    """Sketches the columns, predicts each conjunct and sets the order.
    Returns (selectivity, cost, order) as predicted."""
    f = conjunct_features(pred, types, sketch(cols, num_rows))
    sel, cost = mlp.predict(f)
    order = order_for_scan(pred, types, sel, cost)
    return sel, cost, order


def filter_project_learned(mlp, pred, proj, types, cols, outs, d_count, num_rows, stream=None, order=True):
    """One filter + project in the learned form: the conjuncts in the order of
    least predicted cost (order=False keeps the order already set), the
    outputs written in one pass up to the predicted row bound
    (bsql_filter_project_sized: positions past it are not written and the
    full count is reported), and only when the count exceeds the bound a
    second call over the whole output. `outs` must hold num_rows rows for
    that second call. Returns (rows selected, whether the bound overflowed,
    the order set or None); the count is read back, so the call synchronizes
    the stream."""
    from .interop import filter_project, filter_project_sized
    chosen = None
    if order:
        sel, _, chosen = learned_order(mlp, pred, types, cols, num_rows)
    else:
        f = conjunct_features(pred, types, sketch(cols, num_rows, stream))
        sel, _ = mlp.predict(f)
    bound = row_bound(sel, num_rows)
    filter_project_sized(pred, proj, cols, outs, d_count, bound, num_rows, stream=stream)
    m = int(d_count.item())
    if m <= bound:
        return m, False, chosen
    filter_project(pred, proj, cols, outs, d_count, num_rows, stream=stream)
    return int(d_count.item()), True, chosen


def plan_batch_learned(mlp, preds, types, cols, num_rows, stream=None):
    """Orders every predicate's conjuncts by least predicted cost and returns
    each one's output row bound: one sketch of the table and one forward
    pass of the MLP over all the predicates' conjuncts."""
    sk = sketch(cols, num_rows, stream)
    feats = [conjunct_features(p, types, sk) for p in preds]
    sel, cost = mlp.predict(np.concatenate(feats, axis=0))
    bounds, at = [], 0
    for p, f in zip(preds, feats):
        s, c = sel[at:at + len(f)], cost[at:at + len(f)]
        at += len(f)
        order_for_scan(p, types, s, c)
        bounds.append(row_bound(s, num_rows))
    return bounds


def filter_project_batch_learned(mlp, preds, projs, types, cols, out_lists, d_counts, num_rows, stream=None,
                                 bounds=None):
    """The learned and batched form of several queries over one table: each
    predicate's conjuncts in the order of least predicted cost, each query's
    outputs written up to its predicted row bound, all the queries in one
    pass (bsql_filter_project_batch); a query whose count exceeds its bound
    runs again on its own over the whole output (`out_lists[q]` must hold
    num_rows rows for that). bounds: from plan_batch_learned (planned here
    when None). Returns (counts, which queries overflowed); the counts are
    read back, so the call synchronizes the stream."""
    from .interop import filter_project, filter_project_batch
    if bounds is None:
        bounds = plan_batch_learned(mlp, preds, types, cols, num_rows, stream)
    filter_project_batch(preds, projs, cols, out_lists, d_counts, num_rows, capacities=bounds, stream=stream)
    counts = [int(c) for c in d_counts.cpu()]
    over = [c > b for c, b in zip(counts, bounds)]
    for q, o in enumerate(over):
        if o:
            filter_project(preds[q], projs[q], cols, out_lists[q], d_counts[q:q + 1], num_rows, stream=stream)
    return counts, over


# ------------------------------------------------------------ synthetic data

COPULA_TYPES = [10, 10, 10, 10, 3, 3, 3, 3]   # 4 x float64, 4 x int32


def copula_table(n, rho, gen, device="cuda"):
    """8 columns through a Gaussian copula with pairwise correlation rho:
    z = sqrt(rho) g0 + sqrt(1 - rho) g_k, u = Phi(z); float64 1e6 u, int32
    floor(1e6 u)."""
    import torch
    g0 = torch.randn(n, generator=gen, device=device, dtype=torch.float64)
    cols = []
    for k, t in enumerate(COPULA_TYPES):
        z = np.sqrt(rho) * g0 + np.sqrt(1.0 - rho) * torch.randn(n, generator=gen, device=device,
                                                                 dtype=torch.float64)
        u = torch.special.ndtr(z)
        data = u * 1e6 if t == 10 else torch.floor(u * 1e6).to(torch.int32)
        cols.append(DeviceColumn(t, data.contiguous()))
    return cols


def range_term(rng, col, sel):
    """`>=($k, lo)` or `<($k, hi)` selecting about `sel` of a uniform column."""
    t = COPULA_TYPES[col]
    if rng.random() < 0.5:
        v = 1e6 * (1.0 - sel)
        lit = f"{v:.2f}:DECIMAL(12, 2)" if t == 10 else str(int(round(v)))
        return f">=(${col}, {lit})"
    v = 1e6 * sel
    lit = f"{v:.2f}:DECIMAL(12, 2)" if t == 10 else str(int(round(v)))
    return f"<(${col}, {lit})"


def random_predicate(rng, n_terms):
    """(text, terms): an AND of n_terms range terms on distinct columns with
    selectivities log-uniform in [0.001, 0.9]."""
    cols = rng.permutation(len(COPULA_TYPES))[:n_terms]
    sels = np.exp(rng.uniform(np.log(0.001), np.log(0.9), n_terms))
    terms = [range_term(rng, int(c), float(s)) for c, s in zip(cols, sels)]
    return ("AND(" + ", ".join(terms) + ")" if n_terms > 1 else terms[0]), terms


def copula_predicate(rng, n_terms, sel_range=(0.001, 0.9)):
    """(text, terms, target): an AND of n_terms range terms on distinct
    columns whose joint selectivity on independent columns is `target`,
    log-uniform in sel_range. The terms all bound the same tail (all
    `>=` high quantiles or all `<` low ones) and split the target by
    Dirichlet weights (term k selects target^w_k); on correlated columns the
    joint selectivity rises towards the smallest term's."""
    cols = rng.permutation(len(COPULA_TYPES))[:n_terms]
    target = float(np.exp(rng.uniform(np.log(sel_range[0]), np.log(sel_range[1]))))
    w = rng.dirichlet(np.ones(n_terms))
    upper = rng.random() < 0.5
    terms = []
    for c, wk in zip(cols, w):
        s = min(max(target ** float(wk), 1e-4), 0.9999)
        t = COPULA_TYPES[int(c)]
        v = 1e6 * (1.0 - s) if upper else 1e6 * s
        lit = f"{v:.2f}:DECIMAL(12, 2)" if t == 10 else str(int(round(v)))
        terms.append(f">=(${int(c)}, {lit})" if upper else f"<(${int(c)}, {lit})")
    return ("AND(" + ", ".join(terms) + ")" if n_terms > 1 else terms[0]), terms, target


def measure_terms(terms, cols, n, reps=3):
    """Exact selectivity and time (microseconds) of each term alone."""
    import torch
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    ins = columns_array(cols)
    sels, times = [], []
    a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for text in terms:
        p = Program.compile(text, COPULA_TYPES)
        L.check(L.lib().bsql_filter_count(p.handle, ins, len(cols), n, _ptr(cnt), _stream_ptr(None)))
        a.record()
        for _ in range(reps):
            L.check(L.lib().bsql_filter_count(p.handle, ins, len(cols), n, _ptr(cnt), _stream_ptr(None)))
        z.record()
        torch.cuda.synchronize()
        sels.append(int(cnt.item()) / n)
        times.append(a.elapsed_time(z) * 1e3 / reps)
    return np.asarray(sels), np.asarray(times)


def collect_records(n_records, rows=1 << 20, seed=0, terms_per_batch=8):
    """Training records from the exact interpreter: features of single range
    terms on sketched synthetic batches, and the term's observed selectivity
    and log cost (microseconds per million rows)."""
    import torch
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    X, Y = [], []
    while len(X) < n_records:
        rho = float(rng.choice([0.0, 0.5, 0.9]))
        cols = copula_table(rows, rho, gen)
        sk = sketch(cols, rows)
        terms = []
        for c in rng.permutation(len(COPULA_TYPES))[:terms_per_batch]:
            sel = float(np.exp(rng.uniform(np.log(0.0005), np.log(0.95))))
            terms.append(range_term(rng, int(c), sel))
        sels, times = measure_terms(terms, cols, rows)
        for text, s, t in zip(terms, sels, times):
            p = Program.compile(text, COPULA_TYPES)
            X.append(conjunct_features(p, COPULA_TYPES, sk)[0])
            Y.append((s, np.log(max(t, 1e-3) * (1 << 20) / rows)))
        del cols
    return np.asarray(X[:n_records], dtype=np.float32), np.asarray(Y[:n_records], dtype=np.float32)


def train(X, Y, epochs=40, batch=256, lr=2e-3, seed=1, cost_weight=0.5, log=None):
    """Trains a fresh Mlp on the GPU; returns it with the per-epoch mean loss."""
    import torch
    mlp = Mlp(seed)
    Xd = torch.from_numpy(X).cuda()
    Yd = torch.from_numpy(Y).cuda()
    rng = np.random.default_rng(seed)
    n = len(X)
    history = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).cuda()
        total = 0.0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            loss = mlp.train_step(Xd[idx].contiguous(), Yd[idx].contiguous(), lr, cost_weight)
            total += float(loss.sum().item())
        history.append(total / n)
        if log:
            log(f"epoch {ep} loss {history[-1]:.4f}")
    return mlp, history
