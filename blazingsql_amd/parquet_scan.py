"""Table scan of Parquet files with the pushed-down filter.

The reference's BindableTableScan reads a file's row groups through
ral::io::parquet_parser::parse_batch (engine/src/io/data_parser/
ParquetParser.cpp:31-63: cudf::io::read_parquet of the projected columns and
the row groups to read), first asking process_skipdata_for_table
(engine/src/skip_data/SkipDataProcessor.cpp:242-346) which row groups the
row-group min/max statistics cannot rule out, and filters each batch with
process_filter + fix_column_aliases (BindableTableScan::do_process,
engine/src/execution_kernels/BatchProcessing.cpp:261-284).

Here the statistics come from the file footer (pyarrow), the skip-data pass
(the filter rewritten over the statistics and evaluated by the interpreter)
runs on the GPU, the kept row groups' projected columns are decoded by
pyarrow on the host and staged in pinned memory (a few row groups ahead,
on worker threads), and copied to the device on a copy stream while the
previous row group is filtered on the compute stream,
and each batch is filtered by the interpreter (the scan's filters=[[...]]
over the projected columns) and renamed by the scan's aliases. Decoding
stays on the host: the predicate columns cross PCIe once, in their decoded
form, and never round-trip through HBM as anything but the batch itself.
"""
import threading

import numpy as np

from . import _lib as L
from .interop import DeviceColumn
from .processor import Table, fix_column_aliases, process_filter, process_skipdata_for_table, scan_projections


def _bsql_type(arrow_type):
    import pyarrow as pa
    t = arrow_type
    table = [(pa.types.is_int8, L.INT8), (pa.types.is_int16, L.INT16), (pa.types.is_int32, L.INT32),
             (pa.types.is_int64, L.INT64), (pa.types.is_float32, L.FLOAT32), (pa.types.is_float64, L.FLOAT64),
             (pa.types.is_boolean, L.BOOL8), (pa.types.is_date32, 12), (pa.types.is_string, L.STRING)]
    for test, tid in table:
        if test(t):
            return tid
    raise ValueError(f"column type {t} is not read by this scan")


def _stat_value(v, tid):
    """A footer statistic as the column's value (DATE: days since the epoch)."""
    import datetime
    if tid == 12:
        return (v - datetime.date(1970, 1, 1)).days
    return v


class ParquetScan:
    """BindableTableScan over Parquet files: `table_scan` is the plan text
    (table=[[...]], filters=[[...]], projects=[[...]], aliases=[[...]])."""

    def __init__(self, paths, table_scan, decode_threads=8):
        import pyarrow.parquet as pq
        self.paths = list(paths)
        self.files = [pq.ParquetFile(p) for p in self.paths]
        schema = self.files[0].schema_arrow
        self.names = list(schema.names)
        self.types = [_bsql_type(f.type) for f in schema]
        for f in self.files[1:]:
            if list(f.schema_arrow.names) != self.names:
                raise ValueError("the files of one scan must have the same columns")
        self.table_scan = table_scan
        self.projects = scan_projections(table_scan, len(self.names))
        self.filterable = "filters" in table_scan
        self.read = []   # (file, row group) pairs read by the last batches() call
        self.decode_threads = max(1, int(decode_threads))  # row groups decoded ahead, each on a worker thread
        self._local = threading.local()

    def metadata(self):
        """The row-group statistics table process_skipdata_for_table reads:
        min_<i>_<name> / max_<i>_<name> for every fixed-width column whose
        statistics every row group has, then file_handle_index and
        row_group_index (the reference's parquet metadata layout)."""
        cols, names = [], []
        groups = [(fi, rg) for fi, f in enumerate(self.files) for rg in range(f.metadata.num_row_groups)]
        for i, (nm, tid) in enumerate(zip(self.names, self.types)):
            if tid == L.STRING:
                continue
            lo, hi, ok = [], [], True
            for fi, rg in groups:
                st = self.files[fi].metadata.row_group(rg).column(i).statistics
                if st is None or not st.has_min_max:
                    ok = False
                    break
                lo.append(_stat_value(st.min, tid))
                hi.append(_stat_value(st.max, tid))
            if not ok:
                continue
            from .interop import np_dtype
            cols += [DeviceColumn.from_numpy(tid, np.asarray(lo, dtype=np_dtype(tid))),
                     DeviceColumn.from_numpy(tid, np.asarray(hi, dtype=np_dtype(tid)))]
            names += [f"min_{i}_{nm}", f"max_{i}_{nm}"]
        cols += [DeviceColumn.from_numpy(L.INT32, np.asarray([g[0] for g in groups], dtype=np.int32)),
                 DeviceColumn.from_numpy(L.INT32, np.asarray([g[1] for g in groups], dtype=np.int32))]
        names += ["file_handle_index", "row_group_index"]
        return Table.from_columns(cols, names)

    def row_groups(self, stream=None):
        """The (file, row group) pairs the scan must read, in file order."""
        everything = [(fi, rg) for fi, f in enumerate(self.files) for rg in range(f.metadata.num_row_groups)]
        if not self.filterable:
            return everything
        ids, read_all = process_skipdata_for_table(self.metadata(), self.names, self.table_scan, stream)
        if read_all:
            return everything
        fcol, rcol = ids.to_numpy()
        return sorted(zip(fcol[0].tolist(), rcol[0].tolist()))

    def _decode(self, fi, rg):
        """The projected columns of one row group, decoded by pyarrow and
        staged in pinned host memory, on a worker thread (pyarrow's decode
        and numpy's copies release the GIL, so row groups decode and stage
        side by side). A pyarrow ParquetFile is not safe to read from several
        threads at once, so each worker thread reads through its own reader
        of the file. Returns ([(type, rows, values, validity, offsets)], rows)
        with pinned CPU tensors (validity and offsets None when absent)."""
        import pyarrow.parquet as pq
        readers = getattr(self._local, "readers", None)
        if readers is None:
            readers = self._local.readers = {}
        f = readers.get(fi)
        if f is None:
            f = readers[fi] = pq.ParquetFile(self.paths[fi])
        t = f.read_row_group(rg, columns=[self.names[p] for p in self.projects])
        staged = [self._stage(t.column(self.names[p]).combine_chunks(), self.types[p]) for p in self.projects]
        return staged, t.num_rows

    @staticmethod
    def _pinned(np_array):
        import torch
        a = np.ascontiguousarray(np_array)
        host = torch.empty(a.shape, dtype=torch.from_numpy(np.zeros(0, a.dtype)).dtype, pin_memory=True)
        host.numpy()[...] = a
        return host

    def _stage(self, arr, tid):
        """One arrow array in pinned memory: validity words from the arrow
        bitmap (the same bit order; null rows' values are zeros), values, and
        for strings the offsets rebased to the first row and its chars."""
        import pyarrow as pa
        n = len(arr)
        valid = None
        if arr.null_count:
            from .interop import pack_validity
            mask = np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=bool)
            valid = self._pinned(pack_validity(mask).view(np.int32))
        if tid == L.STRING:
            offs = np.frombuffer(arr.buffers()[1], dtype=np.int32)[arr.offset:arr.offset + n + 1].astype(np.int64)
            data = arr.buffers()[2]
            chars = np.frombuffer(data, dtype=np.uint8) if data is not None else np.zeros(0, np.uint8)
            lo, hi = (int(offs[0]), int(offs[-1])) if n else (0, 0)
            body = chars[lo:hi] if hi > lo else np.zeros(1, np.uint8)
            return tid, n, self._pinned(body), valid, self._pinned((offs - lo).astype(np.int32))
        if tid == 12:
            arr = arr.cast(pa.int32())
        if tid == L.BOOL8:
            values = np.asarray(arr.fill_null(False).to_numpy(zero_copy_only=False), dtype=np.uint8)
        else:
            filled = arr.fill_null(0) if arr.null_count else arr
            values = np.asarray(filled.to_numpy(zero_copy_only=False))
        return tid, n, self._pinned(values), valid, None

    @staticmethod
    def _upload(staged, copy_stream, compute_stream):
        """A staged row group's columns copied to the device on
        `copy_stream`; the device tensors are recorded on `compute_stream`,
        which reads them."""
        import torch

        def dev(host):
            if host is None:
                return None
            d = host.to("cuda", non_blocking=True)
            d.record_stream(compute_stream)
            return d
        out = []
        with torch.cuda.stream(copy_stream):
            for tid, n, values, valid, offsets in staged:
                if tid == L.STRING:
                    out.append(DeviceColumn(tid, dev(values), dev(valid), dev(offsets), size=n))
                else:
                    out.append(DeviceColumn(tid, dev(values), dev(valid), size=n))
        return out

    def batches(self, stream=None):
        """The scan's output batches: for every row group the skip-data pass
        keeps, its projected columns filtered by the scan's filter (if any)
        and named by its aliases. Row group k+1 is decoded and copied while
        row group k is filtered."""
        import torch
        from concurrent.futures import ThreadPoolExecutor
        compute = stream if stream is not None else torch.cuda.current_stream()
        copy = torch.cuda.Stream()
        groups = self.row_groups(compute)
        self.read = list(groups)
        names = [self.names[p] for p in self.projects]
        pool = ThreadPoolExecutor(max_workers=self.decode_threads)
        ahead = [pool.submit(self._decode, fi, rg) for fi, rg in groups[:self.decode_threads]]
        pending = None
        for k in range(len(groups) + 1):
            nxt = None
            if k < len(groups):
                staged, rows = ahead.pop(0).result()
                if k + self.decode_threads < len(groups):
                    ahead.append(pool.submit(self._decode, *groups[k + self.decode_threads]))
                cols = self._upload(staged, copy, compute)
                ev = torch.cuda.Event()
                ev.record(copy)
                nxt = (cols, rows, ev)
            if pending is not None:
                cols, rows, ev = pending
                compute.wait_event(ev)
                table = Table.from_columns(cols, names, num_rows=rows)
                if self.filterable:
                    out = process_filter(table, self.table_scan, compute)
                else:
                    out = table
                fix_column_aliases(out, self.table_scan)
                yield out
            pending = nxt
        pool.shutdown(wait=True)
