"""MI355X-native BlazingSQL expression interpreter and filter path.

The product is libbsql.so (C ABI: include/bsql/interp.h) with hand-written
gfx950 HIP kernels; this package is its Python binding, used by the tests and
the benchmark, plus the operator surface of the reference (processor.py).
"""
from . import _lib  # noqa: F401
from .interop import (DeviceColumn, Program, filter_columns, filter_project,  # noqa: F401
                      pack_validity, perform_interpreter_operation, unpack_validity)
