"""Host-side mirror of the interpreter's program and column types over the C ABI.

Program            <-> the postfix plan of add_expression_to_interpreter_plan
                       (engine/src/Interpreter/interpreter_cpp.h:50-59)
perform_interpreter_operation <-> interops::perform_interpreter_operation
                       (engine/src/Interpreter/interpreter_cpp.cu:440-594)
DeviceColumn       <-> a cudf column view: data, 32-bit validity words, offsets
"""
import ctypes

import numpy as np

from . import _lib as L

_NP_DTYPE = {
    L.INT8: np.int8, L.INT16: np.int16, L.INT32: np.int32, L.INT64: np.int64,
    L.UINT8: np.uint8, L.UINT16: np.uint16, L.UINT32: np.uint32, L.UINT64: np.uint64,
    L.FLOAT32: np.float32, L.FLOAT64: np.float64, L.BOOL8: np.uint8,
    L.TIMESTAMP_DAYS: np.int32, L.TIMESTAMP_SECONDS: np.int64,
    L.TIMESTAMP_MILLISECONDS: np.int64, L.TIMESTAMP_MICROSECONDS: np.int64,
    L.TIMESTAMP_NANOSECONDS: np.int64, L.DURATION_DAYS: np.int32,
    L.DURATION_SECONDS: np.int64, L.DURATION_MILLISECONDS: np.int64,
    L.DURATION_MICROSECONDS: np.int64, L.DURATION_NANOSECONDS: np.int64,
}


def np_dtype(type_id):
    return _NP_DTYPE[type_id]


def is_float(type_id):
    return type_id in (L.FLOAT32, L.FLOAT64)


def pack_validity(valid_bools):
    """bool array -> LSB-first 32-bit validity words (numpy uint32)."""
    v = np.asarray(valid_bools, dtype=bool)
    n = v.shape[0]
    words = (n + 31) // 32
    padded = np.zeros(words * 32, dtype=np.uint8)
    padded[:n] = v
    bits = np.packbits(padded.reshape(-1, 8), axis=1, bitorder="little").reshape(-1)
    return bits.view(np.uint32) if words else np.zeros(0, dtype=np.uint32)


def unpack_validity(words, n):
    w = np.asarray(words, dtype=np.uint32)
    if n == 0:
        return np.zeros(0, dtype=bool)
    b = np.unpackbits(w.view(np.uint8), bitorder="little")
    return b[:n].astype(bool)


class Program:
    """An interpreter program (reference encoding), owned by libbsql."""

    def __init__(self, handle, owner=None):
        self._h = ctypes.c_void_p(handle)
        # a sub-program of a step belongs to its owner's handle
        self._owner = owner

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        exiting = getattr(L, "exiting", None)  # module globals are gone late in interpreter exit
        if exiting is None or exiting():
            return
        if h is not None and h.value and getattr(self, "_owner", None) is None:
            try:
                L.lib().bsql_program_free(h)
            except Exception:
                pass
            self._h = None

    @classmethod
    def from_arrays(cls, left, right, out, ops, final_out, left_scalars=None, right_scalars=None):
        """ops are operator_type values (ints) or names from _lib.OPS; scalars are
        per-op (type_id, value) pairs or None."""
        n = len(ops)
        ops_i = [L.OP[o] if isinstance(o, str) else int(o) for o in ops]
        a16 = lambda v: (ctypes.c_int16 * max(len(v), 1))(*v)
        sarr = lambda sc: _scalar_array(sc, n) if sc is not None else None
        ls, rs = sarr(left_scalars), sarr(right_scalars)
        h = ctypes.c_void_p()
        L.check(L.lib().bsql_program_from_arrays(
            a16(left), a16(right), a16(out), (ctypes.c_int32 * max(n, 1))(*ops_i), n,
            a16(final_out), len(final_out), ls, rs, ctypes.byref(h)))
        p = cls(h.value)
        p._keep = (ls, rs)
        return p

    @classmethod
    def compile(cls, exprs, in_types):
        """Calcite expression text(s) -> program (evaluate_expressions front-end)."""
        if isinstance(exprs, str):
            exprs = [exprs]
        arr = (ctypes.c_char_p * len(exprs))(*[e.encode() for e in exprs])
        types = (ctypes.c_int32 * max(len(in_types), 1))(*in_types)
        h = ctypes.c_void_p()
        L.check(L.lib().bsql_program_compile(arr, len(exprs), types, len(in_types), ctypes.byref(h)))
        return cls(h.value)

    @property
    def num_ops(self):
        return L.lib().bsql_program_num_ops(self._h)

    @property
    def num_outputs(self):
        return L.lib().bsql_program_num_outputs(self._h)

    @property
    def num_inputs(self):
        return L.lib().bsql_program_num_inputs(self._h)

    @property
    def num_interpreted(self):
        """Outputs the interpreter produces (the encoded program's final_out)."""
        return L.lib().bsql_program_num_interpreted(self._h)

    def roots(self):
        """How each output is produced: ("vm", j) for interpreted output j,
        ("column", c) for a copy of column c, ("literal", (type, value)) or
        ("null", type) for a filled literal."""
        out = []
        for k in range(self.num_outputs):
            kind, idx = ctypes.c_int32(), ctypes.c_int32()
            s = L.Scalar()
            L.check(L.lib().bsql_program_root(self._h, k, ctypes.byref(kind), ctypes.byref(idx), ctypes.byref(s)))
            if kind.value == L.ROOT_INTERPRETED:
                out.append(("vm", idx.value))
            elif kind.value == L.ROOT_COLUMN:
                out.append(("column", idx.value))
            elif s.str_len == -1:
                out.append(("null", s.type_id))
            else:
                out.append(("literal", (s.type_id, s.v.f64 if is_float(s.type_id) else s.v.i64)))
        return out

    def parts(self):
        """[(part Program, [output indices])] of a list split at 64 slots
        (empty for a program in one piece)."""
        res = []
        for j in range(L.lib().bsql_program_num_parts(self._h)):
            h = ctypes.c_void_p()
            buf = (ctypes.c_int32 * 4096)()
            n = L.lib().bsql_program_part(self._h, j, ctypes.byref(h), buf, 4096)
            if n < 0:
                L.check(-n)
            res.append((Program(h.value, owner=self), list(buf)[:n]))
        return res

    def plan_info(self, in_types):
        """Shape of the device plan for bsql_interp_execute over these column
        types (host-side lowering; no device needed)."""
        info = L.PlanInfo()
        t = (ctypes.c_int32 * max(len(in_types), 1))(*in_types)
        L.check(L.lib().bsql_program_plan_info(self._h, t, len(in_types), ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in L.PlanInfo._fields_}

    def arrays(self):
        n, m = self.num_ops, self.num_interpreted
        left = (ctypes.c_int16 * max(n, 1))()
        right = (ctypes.c_int16 * max(n, 1))()
        out = (ctypes.c_int16 * max(n, 1))()
        ops = (ctypes.c_int32 * max(n, 1))()
        fo = (ctypes.c_int16 * max(m, 1))()
        L.check(L.lib().bsql_program_arrays(self._h, left, right, out, ops, fo))
        return dict(left=list(left)[:n], right=list(right)[:n], out=list(out)[:n],
                    ops=list(ops)[:n], final_out=list(fo)[:m])

    def scalar(self, op, side):
        s = L.Scalar()
        L.check(L.lib().bsql_program_scalar(self._h, op, side, ctypes.byref(s)))
        if s.type_id == L.EMPTY:
            return None
        if s.type_id == L.STRING:
            text = ctypes.string_at(s.str, s.str_len).decode()
            # TRIM's literal also carries the side (1 leading, 2 trailing, 3 both)
            return (s.type_id, text, s.v.i64) if s.v.i64 else (s.type_id, text)
        return (s.type_id, s.v.f64 if is_float(s.type_id) else s.v.i64)

    def output_type(self, in_types, k=0):
        t = ctypes.c_int32()
        types = (ctypes.c_int32 * max(len(in_types), 1))(*in_types)
        L.check(L.lib().bsql_program_output_type(self._h, types, len(in_types), k, ctypes.byref(t)))
        return t.value

    def used_columns(self):
        buf = (ctypes.c_int32 * 64)()
        n = L.lib().bsql_program_used_columns(self._h, buf, 64)
        return list(buf)[:n]

    @property
    def table_columns(self):
        """Width of the table the program was compiled for."""
        return L.lib().bsql_program_table_columns(self._h)

    def steps(self):
        """Columns built before the interpreter runs, in order (step k is column
        table_columns + k): dicts with kind ("eval" with a one-output "sub"
        program, or "strfn" with "op"), "out_type" and "args" (a column index,
        or ("literal", text) / ("null",))."""
        out = []
        for k in range(L.lib().bsql_program_num_steps(self._h)):
            info = L.StepInfo()
            L.check(L.lib().bsql_program_step(self._h, k, ctypes.byref(info)))
            args = []
            for i in range(info.n_args):
                c = info.arg_col[i]
                if c >= 0:
                    args.append(c)
                elif c == -1:
                    args.append(("literal", ctypes.string_at(info.arg_str[i], info.arg_len[i]).decode()))
                else:
                    args.append(("null",))
            d = dict(kind="eval" if info.kind == 0 else "strfn", op=info.op, out_type=info.out_type, args=args,
                     start=info.start)
            if info.kind == 0:
                d["sub"] = Program(info.sub, owner=self)
            out.append(d)
        return out

    def output_chars(self, cols, num_rows):
        """Chars a string output over these device columns can need at most."""
        arr = (L.Column * max(len(cols), 1))(*[c.to_c() for c in cols])
        v = ctypes.c_int64()
        L.check(L.lib().bsql_program_output_chars(self._h, arr, len(cols), num_rows, ctypes.byref(v)))
        return v.value

    @property
    def num_conjuncts(self):
        return L.lib().bsql_program_num_conjuncts(self._h)

    def set_conjunct_order(self, order):
        if order is None:
            L.check(L.lib().bsql_program_set_conjunct_order(self._h, None, 0))
        else:
            arr = (ctypes.c_int32 * len(order))(*order)
            L.check(L.lib().bsql_program_set_conjunct_order(self._h, arr, len(order)))

    def set_launch(self, rows_per_thread=0, block_size=0, ranges=True, lazy_loads=False, spill_wide=False,
                   fused=False, generated=False, lds_regs=True):
        """Launch shape for this program (block_size 0, 64 or 256); for a
        predicate also whether its column-vs-literal conjuncts run as range
        tests and whether conjunct columns are loaded lazily; spill_wide keeps
        at most 7 live values in registers and the rest in LDS slots; fused runs
        a range-test filter with at most 8 fixed-width outputs as one pass;
        generated runs the program through a kernel generated for it (op
        fields and column types compiled in, hipRTC, cached per process);
        lds_regs keeps the interpreter's register file in LDS (execute and a
        filter's output phase; the default, faster on every workload measured),
        lds_regs=False in VGPRs."""
        flags = ((0 if ranges else L.LAUNCH_NO_RANGES) | (L.LAUNCH_LAZY_LOADS if lazy_loads else 0) |
                 (L.LAUNCH_SPILL_WIDE if spill_wide else 0) | (L.LAUNCH_FUSED if fused else 0) |
                 (L.LAUNCH_GENERATED if generated else 0) | (L.LAUNCH_LDS_REGS if lds_regs else 0))
        cfg = L.LaunchConfig(block_size, rows_per_thread, flags)
        L.check(L.lib().bsql_program_set_launch(self._h, ctypes.byref(cfg)))

    def to_oracle_dict(self):
        """Arrays + scalars (the interpreted outputs) in the form the CPU
        oracle consumes; "roots" says how every output is produced."""
        a = self.arrays()
        a["roots"] = self.roots()
        n = len(a["ops"])
        a["left_scalars"] = [self.scalar(i, 0) if a["left"][i] == L.SCALAR_INDEX else None for i in range(n)]
        a["right_scalars"] = [self.scalar(i, 1) if a["right"][i] == L.SCALAR_INDEX else None for i in range(n)]
        return a


def rewrite(expr, stage, types=()):
    """One text stage of the expression front end (bsql_expression_rewrite):
    0 fill_minus_op_with_zero, 1 n-ary CONCAT, 2 ms -> ns units, 3 Reinterpret
    timestamps, 4 interval units, 5 preprocess_expression, 6
    replace_calcite_regex, 7 + expand_if_logical_op, 8 parse + custom-op
    rewrites rebuilt as text."""
    t = (ctypes.c_int32 * max(len(types), 1))(*types)
    n = ctypes.c_size_t()
    L.check(L.lib().bsql_expression_rewrite(expr.encode(), stage, t, len(types), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    L.check(L.lib().bsql_expression_rewrite(expr.encode(), stage, t, len(types), buf, n.value + 1,
                                            ctypes.byref(n)))
    return buf.value.decode()


def set_current_timestamp(ts):
    """The timestamp CURRENT_DATE / CURRENT_TIMESTAMP stand for (None: the clock)."""
    L.check(L.lib().bsql_set_current_timestamp(ts.encode() if ts else None))


def _scalar_array(scalars, n):
    arr = (L.Scalar * max(n, 1))()
    keep = []
    for i, sc in enumerate(scalars):
        if sc is None:
            continue
        t, v = sc[0], sc[1]
        arr[i].type_id = t
        if t == L.STRING:
            b = v.encode() if isinstance(v, str) else bytes(v)
            keep.append(b)
            arr[i].str = b
            arr[i].str_len = len(b)
            if len(sc) > 2:
                arr[i].v.i64 = int(sc[2])
        elif is_float(t):
            arr[i].v.f64 = float(v)
        else:
            arr[i].v.i64 = int(v)
    arr._keep = keep
    return arr


class DeviceColumn:
    """A column in device memory (torch tensors as storage)."""

    def __init__(self, type_id, data, validity=None, offsets=None, size=None, offset=0):
        self.type_id = type_id
        self.data = data          # torch tensor (fixed width) or uint8 chars
        self.validity = validity  # torch int32 tensor of words, or None
        self.offsets = offsets    # torch int32 tensor (strings)
        if size is None:
            size = (offsets.numel() - 1 - offset) if type_id == L.STRING else data.numel() - offset
        self.size = size
        self.offset = offset

    def to_c(self):
        c = L.Column()
        c.type_id = self.type_id
        c.size = self.size
        c.offset = self.offset
        c.data = self.data.data_ptr() if self.data is not None and self.data.numel() else None
        c.validity = self.validity.data_ptr() if self.validity is not None else None
        c.offsets = self.offsets.data_ptr() if self.offsets is not None else None
        if self.type_id == L.STRING:
            c.chars_size = self.chars_size
        return c

    @property
    def chars_size(self):
        """Bytes in the chars buffer (a string column's capacity)."""
        cap = getattr(self, "_chars_cap", None)
        if cap is not None:
            return cap
        if self.offsets is not None and self.offsets.numel():
            return int(self.offsets[-1].item())
        return 0

    @staticmethod
    def from_numpy(type_id, values, valid=None, device="cuda", chars=None, offsets=None):
        import torch
        if type_id == L.STRING:
            data = torch.from_numpy(np.ascontiguousarray(chars, dtype=np.uint8).copy()).to(device)
            offs = torch.from_numpy(np.ascontiguousarray(offsets, dtype=np.int32).copy()).to(device)
            n = len(offsets) - 1
            vt = None
            if valid is not None:
                vt = torch.from_numpy(pack_validity(valid).view(np.int32).copy()).to(device)
            return DeviceColumn(type_id, data, vt, offs, size=n)
        arr = np.ascontiguousarray(values, dtype=np_dtype(type_id))
        data = torch.from_numpy(arr.copy()).to(device)
        vt = None
        if valid is not None:
            vt = torch.from_numpy(pack_validity(valid).view(np.int32).copy()).to(device)
        return DeviceColumn(type_id, data, vt, size=len(arr))

    @staticmethod
    def from_strings(strings, valid=None, device="cuda"):
        """A STRING column from Python str/bytes values."""
        bs = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
        offs = np.zeros(len(bs) + 1, dtype=np.int32)
        offs[1:] = np.cumsum([len(b) for b in bs]) if bs else []
        chars = np.frombuffer(b"".join(bs), dtype=np.uint8)
        return DeviceColumn.from_numpy(L.STRING, None, valid, device, chars=chars, offsets=offs)

    @staticmethod
    def empty(type_id, n, nullable=True, device="cuda", chars=0):
        """Uninitialized output column of n rows (STRING: n+1 offsets and
        `chars` bytes of chars)."""
        import torch
        if type_id == L.STRING:
            data = torch.empty(max(chars, 1), dtype=torch.uint8, device=device)
            offs = torch.zeros(n + 1, dtype=torch.int32, device=device)
            vt = torch.zeros(max((n + 31) // 32, 1), dtype=torch.int32, device=device) if nullable else None
            col = DeviceColumn(type_id, data, vt, offs, size=n)
            col._chars_cap = chars
            return col
        dt = {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}
        width = np.dtype(np_dtype(type_id)).itemsize
        data = torch.empty(max(n, 1), dtype=dt[width], device=device)
        vt = torch.zeros(max((n + 31) // 32, 1), dtype=torch.int32, device=device) if nullable else None
        return DeviceColumn(type_id, data, vt, size=n)

    def to_numpy(self, n=None):
        """(values, valid) of the first n rows; STRING values are a list of str."""
        n = self.size if n is None else n
        if self.type_id == L.STRING:
            offs = self.offsets[self.offset:self.offset + n + 1].cpu().numpy().astype(np.int64)
            chars = self.data.cpu().numpy()
            vals = [bytes(chars[offs[i]:offs[i + 1]]).decode(errors="replace") for i in range(n)]
            if self.validity is None:
                valid = np.ones(n, dtype=bool)
            else:
                valid = unpack_validity(self.validity.cpu().numpy().view(np.uint32), self.offset + n)[self.offset:]
            return vals, valid
        raw = self.data[self.offset:self.offset + n].cpu().numpy() if n else np.zeros(0)
        vals = raw.view(np_dtype(self.type_id)) if n else np.zeros(0, dtype=np_dtype(self.type_id))
        if self.validity is None:
            valid = np.ones(n, dtype=bool)
        else:
            words = self.validity.cpu().numpy().view(np.uint32)
            valid = unpack_validity(words, self.offset + n)[self.offset:]
        return vals, valid


def generated_kernel_stats():
    """(kernels generated so far, launches that ran one, compile seconds)."""
    k, n, s = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
    L.check(L.lib().bsql_generated_kernel_stats(ctypes.byref(k), ctypes.byref(n), ctypes.byref(s)))
    return k.value, n.value, s.value


def check_generated_kernel(pred, proj, in_cols, out_cols, num_rows, arch="gfx950"):
    """Generates and compiles, without a device, the kernel filter_project
    (pred given) or perform_interpreter_operation (pred None) would run with
    generated=True over columns like these (L.Column descriptors or device
    columns; only types, validity presence and alignment are read); returns
    the code object bytes."""
    def arr(cols):
        a = (L.Column * max(len(cols), 1))()
        for k, c in enumerate(cols):
            a[k] = c if isinstance(c, L.Column) else c.to_c()
        return a
    nbytes = ctypes.c_int64()
    L.check(L.lib().bsql_generated_kernel_check(pred.handle if pred is not None else None, proj.handle,
                                                arr(in_cols), len(in_cols), arr(out_cols), len(out_cols),
                                                num_rows, arch.encode(), ctypes.byref(nbytes)))
    return nbytes.value


def columns_array(cols):
    arr = (L.Column * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i] = c.to_c()
    return arr


def _stream_ptr(stream):
    """hipStream_t of `stream`: None = torch's current stream, a
    torch.cuda.Stream, or a raw handle (int)."""
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(getattr(stream, "cuda_stream", stream))


def perform_interpreter_operation(prog, in_cols, out_cols, num_rows=None, stream=None):
    """Evaluates every output of `prog` over `in_cols` into `out_cols`."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    L.check(L.lib().bsql_interp_execute(prog.handle, ins, len(in_cols), outs, len(out_cols),
                                        num_rows, _stream_ptr(stream)))


def filter_columns(pred, in_cols, out_cols, d_count, num_rows=None, stream=None):
    """Stable compaction of every input column by the predicate (count on device)."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    L.check(L.lib().bsql_filter(pred.handle, ins, len(in_cols), outs, num_rows,
                                ctypes.c_void_p(d_count.data_ptr()), _stream_ptr(stream)))


def filter_project(pred, proj, in_cols, out_cols, d_count, num_rows=None, stream=None):
    """Fused WHERE + SELECT: projection outputs of the selected rows, in order."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    L.check(L.lib().bsql_filter_project(pred.handle, proj.handle, ins, len(in_cols), outs,
                                        len(out_cols), num_rows,
                                        ctypes.c_void_p(d_count.data_ptr()), _stream_ptr(stream)))


def filter_project_noalloc(pred, proj, in_cols, out_cols, d_count, num_rows=None, stream=None):
    """filter_project through the allocation-free entry point (fixed device
    scratch, graph-capturable)."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    L.check(L.lib().bsql_filter_project_noalloc(pred.handle, proj.handle, ins, len(in_cols), outs,
                                                len(out_cols), num_rows,
                                                ctypes.c_void_p(d_count.data_ptr()), _stream_ptr(stream)))


def filter_project_sized(pred, proj, in_cols, out_cols, d_count, capacity, num_rows=None, stream=None):
    """filter_project into outputs that hold `capacity` rows: the first
    min(count, capacity) selected rows are written, *d_count gets the full
    count (count > capacity: run again with larger outputs)."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    L.check(L.lib().bsql_filter_project_sized(pred.handle, proj.handle, ins, len(in_cols), outs, len(out_cols),
                                              num_rows, capacity, ctypes.c_void_p(d_count.data_ptr()),
                                              _stream_ptr(stream)))


def _batch_arrays(preds, projs, out_lists):
    nq = len(preds)
    if len(projs) != nq or len(out_lists) != nq:
        raise ValueError("one projection and one output list per predicate")
    pa = (ctypes.c_void_p * max(nq, 1))(*[p.handle for p in preds])
    ja = (ctypes.c_void_p * max(nq, 1))(*[p.handle for p in projs])
    outs = [columns_array(o) for o in out_lists]
    oa = (ctypes.c_void_p * max(nq, 1))(*[ctypes.cast(o, ctypes.c_void_p) for o in outs])
    na = (ctypes.c_int32 * max(nq, 1))(*[len(o) for o in out_lists])
    return pa, ja, outs, oa, na


def filter_project_batch(preds, projs, in_cols, out_lists, d_counts, num_rows=None, capacities=None,
                         stream=None):
    """Queries (preds[q], projs[q]) over the same columns in one pass
    (bsql_filter_project_batch): query q writes out_lists[q] up to
    capacities[q] rows (default: num_rows) and its full count to d_counts[q]
    (an int64 CUDA tensor with one entry per query)."""
    if num_rows is None:
        num_rows = in_cols[0].size if in_cols else 0
    nq = len(preds)
    caps = [num_rows] * nq if capacities is None else [int(c) for c in capacities]
    pa, ja, outs, oa, na = _batch_arrays(preds, projs, out_lists)
    ca = (ctypes.c_int64 * max(nq, 1))(*caps)
    ins = columns_array(in_cols)
    L.check(L.lib().bsql_filter_project_batch(pa, ja, nq, ins, len(in_cols), oa, na, num_rows, ca,
                                              ctypes.c_void_p(d_counts.data_ptr()), _stream_ptr(stream)))


def check_generated_batch_kernel(preds, projs, in_cols, out_lists, num_rows, arch="gfx950"):
    """Compile-only check of filter_project_batch's generated kernel over
    columns like these (L.Column descriptors or device columns): the code
    size in bytes. No GPU needed."""
    def arr(cols):
        return (L.Column * max(len(cols), 1))(*[c if isinstance(c, L.Column) else c.to_c() for c in cols])
    nq = len(preds)
    pa = (ctypes.c_void_p * nq)(*[p.handle for p in preds])
    ja = (ctypes.c_void_p * nq)(*[p.handle for p in projs])
    outs = [arr(o) for o in out_lists]
    oa = (ctypes.c_void_p * nq)(*[ctypes.cast(o, ctypes.c_void_p) for o in outs])
    na = (ctypes.c_int32 * nq)(*[len(o) for o in out_lists])
    nbytes = ctypes.c_int64()
    L.check(L.lib().bsql_generated_batch_kernel_check(pa, ja, nq, arr(in_cols), len(in_cols), oa, na, num_rows,
                                                      arch.encode(), ctypes.byref(nbytes)))
    return nbytes.value


class Batch:
    """Many (program, columns, rows) batches run as one persistent launch
    (bsql_batch_*). The launch allocates and copies nothing, so it can be
    captured in a HIP graph."""

    def __init__(self, items):
        """items: list of (Program, in_cols, out_cols, num_rows)."""
        n = len(items)
        self._keep = []
        descs = (L.BatchDesc * max(n, 1))()
        for i, (prog, ins, outs, rows) in enumerate(items):
            ia, oa = columns_array(ins), columns_array(outs)
            self._keep += [ia, oa, prog, ins, outs]
            descs[i].prog = prog.handle
            descs[i].inp = ctypes.cast(ia, ctypes.c_void_p)
            descs[i].n_in = len(ins)
            descs[i].out = ctypes.cast(oa, ctypes.c_void_p)
            descs[i].n_out = len(outs)
            descs[i].num_rows = rows
        h = ctypes.c_void_p()
        L.check(L.lib().bsql_batch_prepare(descs, n, ctypes.byref(h)))
        self.handle = h.value

    def launch(self, stream=None):
        L.check(L.lib().bsql_batch_launch(self.handle, _stream_ptr(stream)))

    def set_launch(self, rows_per_thread=0, block_size=0, lds_regs=False, spill_wide=False):
        """Launch shape: rows per lane, workgroup size (0, 64 or 256),
        whether the interpreter's register file sits in LDS, and whether every
        program keeps at most 7 live values in registers (the rest in LDS spill
        slots) so that the 8-register file runs them all."""
        flags = (L.LAUNCH_LDS_REGS if lds_regs else 0) | (L.LAUNCH_SPILL_WIDE if spill_wide else 0)
        cfg = L.LaunchConfig(block_size, rows_per_thread, flags)
        L.check(L.lib().bsql_batch_set_launch(self.handle, ctypes.byref(cfg)))

    def autotune(self, stream=None):
        """Times every launch shape; keeps the fastest. Returns its rows per
        lane; `chosen` holds (block_size, rows_per_thread, lds_regs, spill_wide)."""
        cfg = L.LaunchConfig()
        L.check(L.lib().bsql_batch_autotune(self.handle, _stream_ptr(stream), ctypes.byref(cfg)))
        self.chosen = (cfg.block_size, cfg.rows_per_thread, bool(cfg.flags & L.LAUNCH_LDS_REGS),
                       bool(cfg.flags & L.LAUNCH_SPILL_WIDE))
        return cfg.rows_per_thread

    @property
    def total_rows(self):
        return L.lib().bsql_batch_total_rows(self.handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        exiting = getattr(L, "exiting", None)  # module globals are gone late in interpreter exit
        if h and exiting is not None and not exiting():
            L.lib().bsql_batch_free(h)
            self.handle = None


def autotune_program(prog, in_cols, out_cols, num_rows, stream=None):
    """Times the launch shapes for this program on these columns; keeps the
    fastest on the program and returns its rows per lane."""
    ins, outs = columns_array(in_cols), columns_array(out_cols)
    cfg = L.LaunchConfig()
    L.check(L.lib().bsql_program_autotune(prog.handle, ins, len(in_cols), outs, len(out_cols), num_rows,
                                          _stream_ptr(stream), ctypes.byref(cfg)))
    return cfg.rows_per_thread
