"""Shared helpers of the parity tests: seeded tables, random Calcite
expressions, and the bit-exact comparison used against the oracle."""
import numpy as np

INT8, INT16, INT32, INT64 = 1, 2, 3, 4
FLOAT32, FLOAT64, BOOL8, TS_D, STRING = 9, 10, 11, 12, 23
NP = {INT8: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64, FLOAT32: np.float32,
      FLOAT64: np.float64, BOOL8: np.uint8, TS_D: np.int32}


def random_column(rng, t, n, null_rate=0.0):
    if t == FLOAT64:
        v = rng.uniform(-1e3, 1e3, n)
        v[rng.random(n) < 0.05] = 0.0
    elif t == FLOAT32:
        v = rng.uniform(-1e3, 1e3, n).astype(np.float32)
    elif t == INT64:
        v = rng.integers(-10**6, 10**6, n, dtype=np.int64)
        wide = rng.random(n) < 0.02  # full-range values exercise int64 wrap-around
        v[wide] = rng.integers(-2**63, 2**63 - 1, int(wide.sum()), dtype=np.int64)
    elif t == INT32:
        v = rng.integers(-1000, 1000, n, dtype=np.int32)
    elif t == INT16:
        v = rng.integers(-300, 300, n, dtype=np.int16)
    elif t == INT8:
        v = rng.integers(-100, 100, n, dtype=np.int8)
    elif t == BOOL8:
        v = (rng.random(n) < 0.5).astype(np.uint8)
    elif t == TS_D:
        v = rng.integers(-20000, 40000, n, dtype=np.int32)
    else:
        raise ValueError(t)
    valid = None
    if null_rate > 0:
        valid = rng.random(n) >= null_rate
    return np.asarray(v, dtype=NP[t]), valid


def make_table(rng, types, n, null_rates=None):
    cols = []
    for i, t in enumerate(types):
        nr = null_rates[i] if null_rates else 0.0
        cols.append((t,) + random_column(rng, t, n, nr))
    return cols


def host_columns(table):
    from oracle import HostColumn
    return [HostColumn(t, v, valid) for t, v, valid in table]


def device_columns(table):
    from blazingsql_amd.interop import DeviceColumn
    return [DeviceColumn.from_numpy(t, v, valid) for t, v, valid in table]


def assert_same(gpu_vals, gpu_valid, ora_vals, ora_valid, what=""):
    """Bit-exact: validity bits equal, and data equal wherever valid."""
    gpu_valid = np.asarray(gpu_valid, dtype=bool)
    ora_valid = np.asarray(ora_valid, dtype=bool)
    assert gpu_valid.shape == ora_valid.shape, what
    bad = np.nonzero(gpu_valid != ora_valid)[0]
    assert bad.size == 0, f"{what}: validity differs at rows {bad[:10]}"
    g = np.asarray(gpu_vals)[gpu_valid]
    o = np.asarray(ora_vals)[ora_valid]
    width = g.dtype.itemsize
    ui = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    gb, ob = g.view(ui), o.astype(g.dtype).view(ui)
    neq = gb != ob
    if g.dtype.kind == "f":
        # IEEE-754 leaves NaN payloads to the implementation: any NaN equals any NaN
        neq &= ~(np.isnan(g) & np.isnan(o.astype(g.dtype)))
    elif width == 8:
        # a CASE whose branches mix BIGINT and DOUBLE hands the double's bits to
        # a BIGINT column unconverted (the interpreter's registers are untyped
        # 64-bit slots); a NaN made by an invalid operation (fmod(x, 0), inf - inf)
        # is negative on x86 and positive on the GPU, so NaN bits match any NaN bits
        neq &= ~(np.isnan(gb.view(np.float64)) & np.isnan(ob.view(np.float64)))
    bad = np.nonzero(neq)[0]
    assert bad.size == 0, f"{what}: data differs at {bad.size} valid rows, e.g. {g[bad[:5]]} vs {o[bad[:5]]}"


# ----------------------------------------------------------- expressions

class ExprGen:
    """Random Calcite expressions over a typed table, restricted to what the
    interpreter evaluates (no string functions, no literal-literal ops)."""

    NUM_OPS = ["+", "-", "*", "/", "MOD"]
    CMP_OPS = ["=", "<>", "<", ">", "<=", ">="]

    def __init__(self, rng, types):
        self.rng = rng
        self.types = types

    def col(self, kinds):
        idx = [i for i, t in enumerate(self.types) if t in kinds]
        return f"${int(self.rng.choice(idx))}" if idx else None

    def literal(self, kind):
        r = self.rng
        if kind == "int":
            return str(int(r.integers(-50, 50)))
        if kind == "float":
            return f"{r.uniform(-100, 100):.3f}:DECIMAL(6, 3)"
        if kind == "bool":
            return "true" if r.random() < 0.5 else "false"
        return "0"

    def numeric(self, depth):
        r = self.rng
        choice = r.random()
        if depth <= 0 or choice < 0.3:
            c = self.col({INT8, INT16, INT32, INT64, FLOAT32, FLOAT64})
            return c
        if choice < 0.6:
            op = self.NUM_OPS[int(r.integers(0, len(self.NUM_OPS)))]
            a = self.numeric(depth - 1)
            b = self.numeric(depth - 1) if r.random() < 0.6 else self.literal("int" if r.random() < 0.6 else "float")
            if r.random() < 0.3:
                a, b = b, a
            if a.split(":")[0].lstrip("-").replace(".", "").isdigit() and b.split(":")[0].lstrip("-").replace(".", "").isdigit():
                b = self.col({INT32, INT64, FLOAT64})
            return f"{op}({a}, {b})"
        if choice < 0.7:
            cast = ["INTEGER", "BIGINT", "DOUBLE", "FLOAT"][int(r.integers(0, 4))]
            return f"CAST({self.numeric(depth - 1)}):{cast}"
        if choice < 0.8:
            d = self.col({TS_D})
            if d:
                return f"EXTRACT(FLAG({['YEAR', 'MONTH', 'DAY'][int(r.integers(0, 3))]}), {d})"
            return self.numeric(depth - 1)
        if choice < 0.87:
            f = ["ABS", "FLOOR", "CEIL"][int(r.integers(0, 3))]
            return f"{f}({self.numeric(depth - 1)})"
        if choice < 0.94:
            x = self.numeric(depth - 1)
            y = self.numeric(depth - 1)
            return f"CASE(IS NOT NULL({x}), {x}, {y})"  # COALESCE as Calcite writes it
        return f"CASE({self.boolean(depth - 1)}, {self.numeric(depth - 1)}, {self.numeric(depth - 1)})"

    def boolean(self, depth):
        r = self.rng
        choice = r.random()
        if depth <= 0 or choice < 0.45:
            a = self.numeric(depth - 1)
            b = self.numeric(depth - 1) if r.random() < 0.4 else self.literal("int")
            return f"{self.CMP_OPS[int(r.integers(0, 6))]}({a}, {b})"
        if choice < 0.7:
            op = "AND" if r.random() < 0.5 else "OR"
            k = int(r.integers(2, 4))
            return f"{op}(" + ", ".join(self.boolean(depth - 1) for _ in range(k)) + ")"
        if choice < 0.8:
            return f"NOT({self.boolean(depth - 1)})"
        if choice < 0.9:
            return f"{['IS NULL', 'IS NOT NULL'][int(r.integers(0, 2))]}({self.numeric(depth - 1)})"
        c = self.col({BOOL8})
        return c if c else self.boolean(depth - 1)


def eval_program_both(bsql, prog_text_list, table, n=None, launch=None):
    """Compiles expressions, runs them on the GPU (with `launch`, keyword
    arguments of Program.set_launch) and in the oracle; returns pairs."""
    import oracle
    from blazingsql_amd.interop import DeviceColumn, perform_interpreter_operation
    types = [t for t, _, _ in table]
    p = bsql.Program.compile(prog_text_list, types)
    if launch:
        p.set_launch(**launch)
    out_types = [p.output_type(types, k) for k in range(p.num_outputs)]
    n = len(table[0][1]) if n is None else n
    dev_in = device_columns(table)
    dev_out = [DeviceColumn.empty(t, n) for t in out_types]
    perform_interpreter_operation(p, dev_in, dev_out, n)
    ora = oracle.evaluate(p, host_columns(table), n)
    res = []
    for k in range(len(out_types)):
        gv, gvalid = dev_out[k].to_numpy(n)
        res.append(((gv, gvalid), ora[k], out_types[k]))
    return res
