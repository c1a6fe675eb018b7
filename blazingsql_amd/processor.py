"""Tables and the processor functions of the filter/project operators, as the
reference's ral::processor namespace exposes them (LogicalFilter.h:14-20,
LogicalProject.h:24-30), over the C ABI of include/bsql/table.h.

    process_filter(table, "LogicalFilter(condition=[...])")      -> Table
    process_project(table, "LogicalProject(a=[...], b=[...])")   -> Table (consumes table)
    evaluate_expressions(table, ["+($0, 1)", ...])               -> Table
    apply_boolean_filter(table, mask_column)                     -> Table

Device buffers of the tables the library creates come from torch's caching
allocator (installed with `use_torch_allocator`, which the first call here
does), so that every column can be read back as torch tensors; an optional
byte budget makes an allocation fail the way RMM's limiting adaptor does, and
the call then raises OutOfMemory (the operators' RETRY).
"""
import ctypes
import threading

import numpy as np

from . import _lib as L
from .interop import DeviceColumn, np_dtype, _stream_ptr

_ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                             ctypes.c_void_p)
_FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


class _TorchAllocator:
    """bsql_alloc_fn / bsql_free_fn over torch.empty on the caller's stream."""

    def __init__(self, limit_bytes=None):
        self.limit = limit_bytes
        self.live = {}          # device pointer -> uint8 tensor
        self.live_bytes = 0
        self.lock = threading.Lock()
        self._alloc = _ALLOC_FN(self._do_alloc)
        self._free = _FREE_FN(self._do_free)

    def _do_alloc(self, ctx, out, nbytes, stream):
        import torch
        try:
            with self.lock:
                if self.limit is not None and self.live_bytes + nbytes > self.limit:
                    return 1
            s = torch.cuda.ExternalStream(stream) if stream else torch.cuda.default_stream()
            with torch.cuda.stream(s):
                t = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        except torch.cuda.OutOfMemoryError:
            return 1
        except Exception:  # noqa: BLE001 -- reported to the library as a device error
            return 2
        with self.lock:
            self.live[t.data_ptr()] = t
            self.live_bytes += int(nbytes)
        out[0] = t.data_ptr()
        return 0

    def _do_free(self, ctx, ptr, nbytes, stream):
        with self.lock:
            t = self.live.pop(ptr, None)
            if t is not None:
                self.live_bytes -= int(nbytes)

    def tensor(self, ptr):
        with self.lock:
            return self.live.get(ptr)


_allocators = []   # every allocator installed (tables keep the one they came from)


def use_torch_allocator(limit_bytes=None):
    """Installs a torch-backed allocator (optionally with a byte budget over
    the buffers it has live) for every table the library creates from now on."""
    a = _TorchAllocator(limit_bytes)
    L.check(L.lib().bsql_set_allocator(ctypes.cast(a._alloc, ctypes.c_void_p),
                                       ctypes.cast(a._free, ctypes.c_void_p), None))
    _allocators.append(a)
    return a


def _ensure_allocator():
    if not _allocators:
        use_torch_allocator()


def _tensor_of(ptr):
    for a in reversed(_allocators):
        t = a.tensor(ptr)
        if t is not None:
            return t
    return None


class Table:
    """A bsql_table handle (BlazingTable): columns + names."""

    def __init__(self, handle, keep=()):
        self._h = handle
        self._keep = list(keep)   # borrowed columns' storage stays alive with the table

    @classmethod
    def from_columns(cls, cols, names=None, num_rows=None):
        """A table borrowing DeviceColumns (nothing copied)."""
        num_rows = (cols[0].size if cols else 0) if num_rows is None else num_rows
        arr = (L.Column * max(len(cols), 1))()
        for i, c in enumerate(cols):
            arr[i] = c.to_c()
        nm = None
        if names is not None:
            nm = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        h = ctypes.c_void_p()
        L.check(L.lib().bsql_table_from_views(arr, nm, len(cols), num_rows, ctypes.byref(h)))
        return cls(h.value, keep=cols)

    @property
    def handle(self):
        if not self._h:
            raise ValueError("table was consumed or freed")
        return self._h

    @property
    def num_rows(self):
        return L.lib().bsql_table_num_rows(self.handle)

    @property
    def num_columns(self):
        return L.lib().bsql_table_num_columns(self.handle)

    @property
    def names(self):
        return [L.lib().bsql_table_name(self.handle, k).decode() for k in range(self.num_columns)]

    def column(self, k):
        """Column k as a DeviceColumn over the table's own buffers."""
        c = L.Column()
        L.check(L.lib().bsql_table_column(self.handle, k, ctypes.byref(c)))
        for kept in self._keep:
            if hasattr(kept, "to_c") and kept.to_c().data == c.data and kept.type_id == c.type_id:
                return kept

        def buf(ptr, dtype):
            if not ptr:
                return None
            t = _tensor_of(ptr)
            if t is None:
                raise ValueError("column buffer not allocated through the torch allocator")
            return t.view(dtype)

        import torch
        if c.type_id == L.STRING:
            data = buf(c.data, torch.uint8)
            if data is None:
                data = torch.zeros(1, dtype=torch.uint8, device="cuda")
            offs = buf(c.offsets, torch.int32)
            if offs is None:
                offs = torch.zeros(1, dtype=torch.int32, device="cuda")
            col = DeviceColumn(c.type_id, data, buf(c.validity, torch.int32), offs, size=c.size, offset=c.offset)
            col._chars_cap = c.chars_size
            return col
        width = np.dtype(np_dtype(c.type_id)).itemsize
        dt = {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[width]
        data = buf(c.data, dt)
        if data is None:
            data = torch.zeros(0, dtype=dt, device="cuda")
        return DeviceColumn(c.type_id, data, buf(c.validity, torch.int32), size=c.size, offset=c.offset)

    def to_numpy(self):
        """[(values, valid)] per column (STRING values as str lists)."""
        n = self.num_rows
        return [self.column(k).to_numpy(n) for k in range(self.num_columns)]

    def _release(self):
        h, self._h = self._h, None
        return h

    def free(self):
        if self._h:
            L.lib().bsql_table_free(self._release())

    def __del__(self):
        exiting = getattr(L, "exiting", None)  # module globals are gone late in interpreter exit
        if exiting is None or exiting():
            return
        try:
            self.free()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


def _out_table(rc, h, keep=()):
    L.check(rc)
    return Table(h.value, keep)


def process_filter(table, query_part, stream=None):
    """ral::processor::process_filter (LogicalFilter.cpp:31-52)."""
    _ensure_allocator()
    h = ctypes.c_void_p()
    rc = L.lib().bsql_process_filter(table.handle, query_part.encode(), ctypes.byref(h), _stream_ptr(stream))
    return _out_table(rc, h)


def process_project(table, query_part, stream=None):
    """ral::processor::process_project (LogicalProject.cpp:943-969). Consumes
    `table` on success (an identity projection returns it renamed)."""
    _ensure_allocator()
    h = ctypes.c_void_p()
    rc = L.lib().bsql_process_project(table.handle, query_part.encode(), ctypes.byref(h), _stream_ptr(stream))
    L.check(rc)
    keep = table._keep
    table._release()
    return Table(h.value, keep)


def process_project_batches(tables, query_part, stream=None):
    """The projection over a cache's waiting batches of one schema in one
    persistent launch (bsql_process_project_batches; the batches a
    ConcatenatingCacheMachine collects, CacheMachine.cpp:866-900). Returns one
    table per batch, in order; consumes every input on success, none on
    OutOfMemory."""
    _ensure_allocator()
    n = len(tables)
    ins = (ctypes.c_void_p * max(n, 1))(*[t.handle for t in tables])
    outs = (ctypes.c_void_p * max(n, 1))()
    L.check(L.lib().bsql_process_project_batches(ins, n, query_part.encode(), outs, _stream_ptr(stream)))
    result = []
    for t, h in zip(tables, outs):
        keep = t._keep
        t._release()
        result.append(Table(h, keep))
    return result


def evaluate_expressions(table, exprs, stream=None):
    """ral::processor::evaluate_expressions (LogicalProject.cpp:796-941)."""
    _ensure_allocator()
    arr = (ctypes.c_char_p * max(len(exprs), 1))(*[e.encode() for e in exprs])
    h = ctypes.c_void_p()
    rc = L.lib().bsql_evaluate_expressions(table.handle, arr, len(exprs), ctypes.byref(h), _stream_ptr(stream))
    return _out_table(rc, h)


def apply_boolean_filter(table, mask, stream=None):
    """ral::processor::applyBooleanFilter (LogicalFilter.cpp:22-29)."""
    _ensure_allocator()
    c = mask.to_c()
    h = ctypes.c_void_p()
    rc = L.lib().bsql_apply_boolean_filter(table.handle, ctypes.byref(c), ctypes.byref(h), _stream_ptr(stream))
    return _out_table(rc, h)


AGG_OTHER, AGG_COUNT_VALID, AGG_COUNT_ALL = 0, 1, 2


def aggregation_input(table, expr, kind=AGG_OTHER, stream=None):
    """The input column of one aggregation (GroupBy.cpp:226-249, 329-352):
    `expr` is the text between the aggregation's parentheses; COUNT over
    several columns (kind=AGG_COUNT_VALID) sums their INTEGER casts, one
    column is borrowed, anything else is evaluated. A one-column Table."""
    _ensure_allocator()
    h = ctypes.c_void_p()
    rc = L.lib().bsql_aggregation_input(table.handle, expr.encode(), kind, ctypes.byref(h), _stream_ptr(stream))
    return _out_table(rc, h, keep=tuple(table._keep) + (table,))


def skip_data_filter(filter_text, has_stats):
    """The filter rewritten over the row-group statistics (column i: min at
    2i, max at 2i + 1); "" when no row group can be skipped."""
    st = (ctypes.c_uint8 * max(len(has_stats), 1))(*[1 if x else 0 for x in has_stats])
    n = ctypes.c_size_t()
    L.check(L.lib().bsql_skip_data_filter(filter_text.encode(), st, len(has_stats), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    L.check(L.lib().bsql_skip_data_filter(filter_text.encode(), st, len(has_stats), buf, n.value + 1,
                                          ctypes.byref(n)))
    return buf.value.decode()


def process_skipdata_for_table(metadata, names, table_scan, stream=None):
    """ral::skip_data::process_skipdata_for_table (SkipDataProcessor.cpp:
    242-346): (table of the file and row-group ids to read, False), or
    (None, True) when no row group can be skipped."""
    _ensure_allocator()
    nm = (ctypes.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    h = ctypes.c_void_p()
    read_all = ctypes.c_int()
    rc = L.lib().bsql_process_skipdata_for_table(metadata.handle, nm, len(names), table_scan.encode(),
                                                 ctypes.byref(h), ctypes.byref(read_all), _stream_ptr(stream))
    L.check(rc)
    if read_all.value:
        return None, True
    return Table(h.value, (metadata,)), False


def project_bypass(query_part, names):
    """bypassingProject (expression_utils.cpp:1115-1150): (bypass, with_aliases)."""
    nm = (ctypes.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    b, w = ctypes.c_int(), ctypes.c_int()
    L.check(L.lib().bsql_project_bypass(query_part.encode(), nm, len(names), ctypes.byref(b), ctypes.byref(w)))
    return bool(b.value), bool(w.value)


def scan_projections(query_part, num_columns):
    """get_projections_wrapper (BatchProcessing.cpp:30-46)."""
    n = L.lib().bsql_scan_projections(query_part.encode(), num_columns, None, 0)
    if n < 0:
        L.check(-n)
    out = (ctypes.c_int32 * max(n, 1))()
    L.lib().bsql_scan_projections(query_part.encode(), num_columns, out, n)
    return list(out[:n])


def fix_column_aliases(table, query_part):
    """fix_column_aliases (expression_utils.cpp:512-535), in place."""
    L.check(L.lib().bsql_table_fix_column_aliases(table.handle, query_part.encode()))
    return table
