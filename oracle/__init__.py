"""TEST INFRASTRUCTURE: ctypes wrapper of the CPU oracle (interp_oracle.cpp).

The oracle restates BlazingSQL's interpreter (engine/src/Interpreter/interpreter_ops.cuh)
and filter path (engine/src/execution_kernels/LogicalFilter.cpp) on the host. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker or the CPU baseline; the product (blazingsql_amd) never loads it.

Parity status of each oracle function is recorded in the header comment of
interp_oracle.cpp and pinned by tests/test_oracle_golden.py.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsql_oracle.so")

_lib = None

STRING = 23
FLOAT32, FLOAT64 = 9, 10
SCALAR_INDEX = -2

_NP = {1: np.int8, 2: np.int16, 3: np.int32, 4: np.int64, 5: np.uint8, 6: np.uint16,
       7: np.uint32, 8: np.uint64, 9: np.float32, 10: np.float64, 11: np.uint8, 12: np.int32,
       13: np.int64, 14: np.int64, 15: np.int64, 16: np.int64, 17: np.int32, 18: np.int64,
       19: np.int64, 20: np.int64, 21: np.int64}


class _Column(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("chars_size", ctypes.c_int32),
                ("size", ctypes.c_int64), ("offset", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("validity", ctypes.c_void_p),
                ("offsets", ctypes.c_void_p)]


class _SV(ctypes.Union):
    _fields_ = [("i64", ctypes.c_int64), ("f64", ctypes.c_double)]


class _Scalar(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("str_len", ctypes.c_int32), ("v", _SV),
                ("str", ctypes.c_char_p)]


class _Program(ctypes.Structure):
    _fields_ = [("left", ctypes.c_void_p), ("right", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("ops", ctypes.c_void_p), ("n_ops", ctypes.c_int32), ("n_final", ctypes.c_int32),
                ("final_out", ctypes.c_void_p), ("left_scalars", ctypes.c_void_p),
                ("right_scalars", ctypes.c_void_p)]


def build():
    from . import build_oracle
    return build_oracle.build()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        P, I, I64, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        l.oracle_execute.argtypes = [P, P, I, P, I, I64, I, U64]
        l.oracle_filter.argtypes = [P, P, I, P, I64, P, I, U64]
        l.oracle_filter_project.argtypes = [P, P, P, I, P, I, I64, P, I, U64]
        l.oracle_op_types.argtypes = [P, P, I, P, P, P]
        for f in (l.oracle_execute, l.oracle_filter, l.oracle_filter_project, l.oracle_op_types):
            f.restype = I
        _lib = l
    return _lib


class HostColumn:
    """Host column for the oracle: numpy values, bool validity (or None), strings."""

    def __init__(self, type_id, values=None, valid=None, strings=None):
        self.type_id = type_id
        if type_id == STRING:
            strs = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
            self.offsets = np.zeros(len(strs) + 1, dtype=np.int32)
            self.offsets[1:] = np.cumsum([len(s) for s in strs]) if strs else []
            self.chars = np.frombuffer(b"".join(strs) + b"\0", dtype=np.uint8).copy()
            self.values = None
            self.size = len(strs)
        else:
            self.values = np.ascontiguousarray(values, dtype=_NP[type_id])
            self.size = len(self.values)
            self.offsets = self.chars = None
        self.valid = None if valid is None else np.asarray(valid, dtype=bool)
        self._words = None if self.valid is None else _pack(self.valid)

    @staticmethod
    def strings_from_arrays(offsets, chars, valid=None):
        """A STRING column straight from int32 offsets (n+1) and uint8 chars."""
        col = HostColumn(STRING, strings=[], valid=valid)
        col.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        col.chars = np.concatenate([np.asarray(chars, dtype=np.uint8), np.zeros(1, dtype=np.uint8)])
        col.size = len(col.offsets) - 1
        return col

    def c(self):
        c = _Column()
        c.type_id = self.type_id
        c.size = self.size
        c.offset = 0
        if self.type_id == STRING:
            c.data = self.chars.ctypes.data
            c.offsets = self.offsets.ctypes.data
            c.chars_size = int(self.offsets[-1]) if self.size else 0
        else:
            c.data = self.values.ctypes.data if self.size else None
        c.validity = self._words.ctypes.data if self._words is not None else None
        return c


def _pack(valid):
    n = len(valid)
    words = (n + 31) // 32
    padded = np.zeros(max(words, 1) * 32, dtype=np.uint8)
    padded[:n] = valid
    return np.packbits(padded.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint32).copy()


def _unpack(words, n):
    if n == 0:
        return np.zeros(0, dtype=bool)
    return np.unpackbits(np.asarray(words, dtype=np.uint32).view(np.uint8), bitorder="little")[:n].astype(bool)


class _ProgHolder:
    def __init__(self, prog):
        n = len(prog["ops"])
        self.left = np.asarray(prog["left"], dtype=np.int16)
        self.right = np.asarray(prog["right"], dtype=np.int16)
        self.out = np.asarray(prog["out"], dtype=np.int16)
        self.ops = np.asarray(prog["ops"], dtype=np.int32)
        self.fo = np.asarray(prog["final_out"], dtype=np.int16)
        self.ls = (_Scalar * max(n, 1))()
        self.rs = (_Scalar * max(n, 1))()
        self._keep = []
        for arr, key in ((self.ls, "left_scalars"), (self.rs, "right_scalars")):
            for i, sc in enumerate(prog.get(key) or [None] * n):
                if sc is None:
                    continue
                t, v = sc[0], sc[1]
                arr[i].type_id = t
                if t == STRING:
                    b = v.encode() if isinstance(v, str) else bytes(v)
                    self._keep.append(b)
                    arr[i].str = b
                    arr[i].str_len = len(b)
                    if len(sc) > 2:
                        arr[i].v.i64 = int(sc[2])  # TRIM's side
                elif t in (FLOAT32, FLOAT64):
                    arr[i].v.f64 = float(v)
                else:
                    arr[i].v.i64 = int(v)
        self.c = _Program(self.left.ctypes.data, self.right.ctypes.data, self.out.ctypes.data,
                          self.ops.ctypes.data, n, len(self.fo), self.fo.ctypes.data,
                          ctypes.addressof(self.ls), ctypes.addressof(self.rs))


def _cols(cols):
    arr = (_Column * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i] = c.c()
    return arr


def _alloc_out(type_id, n, chars_cap=0):
    """Output buffers; a STRING output gets offsets and chars_cap chars."""
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    c = _Column()
    c.type_id = type_id
    c.size = n
    c.validity = words.ctypes.data
    if type_id == STRING:
        chars = np.zeros(max(chars_cap, 1), dtype=np.uint8)
        offs = np.zeros(n + 1, dtype=np.int32)
        c.data = chars.ctypes.data
        c.offsets = offs.ctypes.data
        c.chars_size = chars_cap
        return (chars, offs), words, c
    vals = np.zeros(max(n, 1), dtype=_NP[type_id])
    c.data = vals.ctypes.data
    return vals, words, c


def _chars_bound(in_cols):
    """A string output is a slice or selection of one string input."""
    return max([int(c.offsets[-1]) for c in in_cols if c.type_id == STRING and c.size] or [0])


def _read_out(vals, words, n):
    valid = _unpack(words, n)
    if isinstance(vals, tuple):
        chars, offs = vals
        return [bytes(chars[offs[i]:offs[i + 1]]).decode() for i in range(n)], valid
    return vals[:n], valid


def execute(prog, in_cols, out_types, num_rows=None, threads=0, seed=0x5eed5eed12345678):
    """perform_interpreter_operation on the host: list of (values, valid) per output."""
    num_rows = in_cols[0].size if num_rows is None else num_rows
    ph = _ProgHolder(prog)
    ins = _cols(in_cols)
    outs = (_Column * max(len(out_types), 1))()
    bufs = []
    for k, t in enumerate(out_types):
        vals, words, c = _alloc_out(t, num_rows, _chars_bound(in_cols))
        outs[k] = c
        bufs.append((vals, words))
    rc = lib().oracle_execute(ctypes.byref(ph.c), ins, len(in_cols), outs, len(out_types), num_rows,
                              threads, seed)
    if rc != 0:
        raise RuntimeError(f"oracle_execute failed ({rc})")
    return [_read_out(v, w, num_rows) for v, w in bufs]


def op_types(prog, in_types):
    ph = _ProgHolder(prog)
    n = len(prog["ops"])
    lt, rt, ot = (np.zeros(max(n, 1), dtype=np.int32) for _ in range(3))
    types = np.asarray(in_types, dtype=np.int32)
    rc = lib().oracle_op_types(ctypes.byref(ph.c), types.ctypes.data, len(types), lt.ctypes.data,
                               rt.ctypes.data, ot.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_op_types failed")
    return lt[:n].tolist(), rt[:n].tolist(), ot[:n].tolist()


def filter_project(pred, proj, in_cols, out_types, num_rows=None, threads=0, seed=0x5eed5eed12345678):
    """Filter then project (two reference kernels) on the host."""
    num_rows = in_cols[0].size if num_rows is None else num_rows
    pp, pj = _ProgHolder(pred), _ProgHolder(proj)
    ins = _cols(in_cols)
    outs = (_Column * max(len(out_types), 1))()
    bufs = []
    for k, t in enumerate(out_types):
        vals, words, c = _alloc_out(t, num_rows, _chars_bound(in_cols))
        outs[k] = c
        bufs.append((vals, words))
    count = ctypes.c_int64()
    rc = lib().oracle_filter_project(ctypes.byref(pp.c), ctypes.byref(pj.c), ins, len(in_cols), outs,
                                     len(out_types), num_rows, ctypes.byref(count), threads, seed)
    if rc != 0:
        raise RuntimeError(f"oracle_filter_project failed ({rc})")
    n = count.value
    return n, [_read_out(v, w, n) for v, w in bufs]


def filter_table(pred, in_cols, num_rows=None, threads=0, seed=0x5eed5eed12345678):
    """process_filter on the host: (count, [(values or strings, valid)])."""
    num_rows = in_cols[0].size if num_rows is None else num_rows
    ph = _ProgHolder(pred)
    ins = _cols(in_cols)
    outs = (_Column * max(len(in_cols), 1))()
    bufs = []
    for k, col in enumerate(in_cols):
        c = _Column()
        c.type_id = col.type_id
        c.size = num_rows
        words = np.zeros(max((num_rows + 31) // 32, 1), dtype=np.uint32)
        if col.type_id == STRING:
            chars = np.zeros(max(len(col.chars), 1), dtype=np.uint8)
            offs = np.zeros(num_rows + 1, dtype=np.int32)
            c.data = chars.ctypes.data
            c.offsets = offs.ctypes.data
            bufs.append((chars, offs, words))
        else:
            vals = np.zeros(max(num_rows, 1), dtype=_NP[col.type_id])
            c.data = vals.ctypes.data
            bufs.append((vals, None, words))
        c.validity = words.ctypes.data if col.valid is not None else None
        outs[k] = c
    count = ctypes.c_int64()
    rc = lib().oracle_filter(ctypes.byref(ph.c), ins, len(in_cols), outs, num_rows,
                             ctypes.byref(count), threads, seed)
    if rc != 0:
        raise RuntimeError(f"oracle_filter failed ({rc})")
    n = count.value
    res = []
    for col, (a, offs, words) in zip(in_cols, bufs):
        valid = _unpack(words, n) if col.valid is not None else np.ones(n, dtype=bool)
        if col.type_id == STRING:
            strs = [bytes(a[offs[i]:offs[i + 1]]).decode() for i in range(n)]
            res.append((strs, valid))
        else:
            res.append((a[:n], valid))
    return n, res


def run_steps(prog, cols, num_rows):
    """The table's host columns followed by the columns a compiled program's
    steps build (function_evaluator_transformer, LogicalProject.cpp:670-731):
    expressions through oracle_execute, string functions through strfn.
    `prog` is a compiled program exposing steps() (blazingsql_amd Program)."""
    from . import strfn
    cols = list(cols)
    for st in prog.steps():
        if st["kind"] == "eval":
            sub = st["sub"]
            (vals, valid), = execute(sub.to_oracle_dict(), [cols[c] for c in sub.used_columns()],
                                     [st["out_type"]], num_rows)
            if st["out_type"] == STRING:
                cols.append(HostColumn(STRING, strings=vals, valid=valid))
            else:
                cols.append(HostColumn(st["out_type"], vals, valid))
            continue
        args = [a if isinstance(a, tuple) else strfn.column_values(cols[a]) for a in st["args"]]
        if st["op"] == strfn.CAST:
            args = [(cols[st["args"][0]].type_id, args[0])]
        res = strfn.build(st["op"], args, num_rows, st.get("start", 0))
        cols.append(HostColumn(STRING, strings=[r if r is not None else "" for r in res],
                               valid=[r is not None for r in res]))
    return cols


def _column_values(col, n):
    """(values or strings, validity) of the first n rows of a host column."""
    valid = np.ones(n, dtype=bool) if col.valid is None else np.asarray(col.valid[:n], dtype=bool)
    if col.type_id == STRING:
        strs = [bytes(col.chars[col.offsets[i]:col.offsets[i + 1]]).decode() for i in range(n)]
        return strs, valid
    return np.asarray(col.values[:n]), valid


def evaluate(prog, cols, num_rows=None, threads=0, seed=0x5eed5eed12345678):
    """evaluate_expressions on the host (LogicalProject.cpp:796-941) for a
    compiled program (blazingsql_amd Program) over the table's host columns:
    the columns its steps build, then each output as the reference produces
    it -- interpreted outputs by perform_interpreter_operation over the used
    columns, a bare column root as a copy of that column, a literal root
    filled (a null literal: all rows null) -- part by part for a list split
    at 64 slots. Returns [(values or strings, validity)] per output."""
    num_rows = cols[0].size if num_rows is None else num_rows
    parts = prog.parts()
    if parts:
        res = [None] * prog.num_outputs
        for part, outs in parts:
            for k, v in zip(outs, evaluate(part, cols, num_rows, threads, seed)):
                res[k] = v
        return res
    types = [c.type_id for c in cols]
    ext = run_steps(prog, cols, num_rows) if prog.steps() else list(cols)
    roots = prog.roots()
    out_types = [prog.output_type(types, k) for k in range(len(roots))]
    vm = [k for k, r in enumerate(roots) if r[0] == "vm"]
    vm_res = {}
    if vm:
        d = prog.to_oracle_dict()
        got = execute(d, [ext[c] for c in prog.used_columns()], [out_types[k] for k in vm], num_rows,
                      threads, seed)
        for k, v in zip(vm, got):
            vm_res[k] = v
    res = []
    for k, r in enumerate(roots):
        if r[0] == "vm":
            res.append(vm_res[k])
        elif r[0] == "column":
            res.append(_column_values(ext[r[1]], num_rows))
        elif r[0] == "literal":
            t, v = r[1]
            res.append((np.full(num_rows, v, dtype=_NP[t]), np.ones(num_rows, dtype=bool)))
        else:
            res.append((np.zeros(num_rows, dtype=_NP.get(r[1], np.int64)), np.zeros(num_rows, dtype=bool)))
    return res


def full_table_program(prog, n_table):
    """The oracle program of a compiled program over its whole table (table
    columns, then the columns its steps build), for oracle.filter_project and
    oracle.filter_table: the interpreted encoding with its input slots mapped
    back to table columns and its other slots shifted past them, and every
    output listed -- an interpreted one at its shifted slot, a bare column
    root at the column's own slot (the reference copies the column).
    Literal roots are not expressible this way (use evaluate)."""
    d = prog.to_oracle_dict()
    used = prog.used_columns()
    n_in, n_tab = len(used), n_table + len(prog.steps())

    def remap(s):
        if s < 0:
            return s
        return used[s] if s < n_in else s - n_in + n_tab

    for key in ("left", "right", "out", "final_out"):
        d[key] = [remap(s) for s in d[key]]
    final = []
    for r in d.pop("roots"):
        if r[0] == "vm":
            final.append(d["final_out"][r[1]])
        elif r[0] == "column":
            final.append(r[1])
        else:
            raise ValueError("a literal root has no slot in an oracle program")
    d["final_out"] = final
    return d
