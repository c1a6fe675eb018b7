"""ctypes binding of libbsql.so (include/bsql/interp.h).

The library is built in-tree by blazingsql_amd/build.py. Loading fails loudly
if it is missing: there is no fallback implementation.
"""
import atexit
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsql.so")

# type ids (cudf::type_id numbering)
EMPTY, INT8, INT16, INT32, INT64 = 0, 1, 2, 3, 4
UINT8, UINT16, UINT32, UINT64 = 5, 6, 7, 8
FLOAT32, FLOAT64, BOOL8 = 9, 10, 11
TIMESTAMP_DAYS, TIMESTAMP_SECONDS, TIMESTAMP_MILLISECONDS = 12, 13, 14
TIMESTAMP_MICROSECONDS, TIMESTAMP_NANOSECONDS = 15, 16
DURATION_DAYS, DURATION_SECONDS, DURATION_MILLISECONDS = 17, 18, 19
DURATION_MICROSECONDS, DURATION_NANOSECONDS = 20, 21
STRING = 23

UNARY_INDEX, SCALAR_INDEX, SCALAR_NULL_INDEX, NULLARY_INDEX = -1, -2, -3, -4

OK, ERR_OOM, ERR_INVALID, ERR_DEVICE = 0, 1, 2, 3

# operator_type numbering (engine/src/parser/expression_utils.hpp:11-93)
OPS = ["INVALID_OP", "RAND", "NOT", "ABS", "FLOOR", "CEIL", "SIN", "COS", "ASIN", "ACOS",
       "TAN", "COTAN", "ATAN", "LN", "LOG", "YEAR", "MONTH", "DAY", "DAYOFWEEK", "HOUR",
       "MINUTE", "SECOND", "IS_NULL", "IS_NOT_NULL", "IS_TRUE", "IS_NOT_TRUE", "IS_NOT_FALSE",
       "CAST_TINYINT", "CAST_SMALLINT", "CAST_INTEGER", "CAST_BIGINT", "CAST_FLOAT",
       "CAST_DOUBLE", "CAST_DATE", "CAST_TIMESTAMP_SECONDS", "CAST_TIMESTAMP_MILLISECONDS",
       "CAST_TIMESTAMP_MICROSECONDS", "CAST_TIMESTAMP", "CAST_VARCHAR", "CHAR_LENGTH",
       "STR_LOWER", "STR_UPPER", "STR_INITCAP", "STR_REVERSE", "ADD", "SUB", "MUL", "DIV",
       "MOD", "POW", "ROUND", "EQUAL", "NOT_EQUAL", "LESS", "GREATER", "LESS_EQUAL",
       "GREATER_EQUAL", "BITWISE_AND", "BITWISE_OR", "BITWISE_XOR", "LOGICAL_AND",
       "LOGICAL_OR", "FIRST_NON_MAGIC", "MAGIC_IF_NOT", "STR_LIKE", "STR_SUBSTRING",
       "STR_REGEXP_REPLACE", "STR_CONCAT", "STR_REPLACE", "STR_LEFT", "STR_RIGHT", "TO_DATE",
       "TO_TIMESTAMP", "STR_TRIM", "IS_NOT_DISTINCT_FROM"]
OP = {name: i for i, name in enumerate(OPS)}


class Column(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("chars_size", ctypes.c_int32),
                ("size", ctypes.c_int64), ("offset", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("validity", ctypes.c_void_p),
                ("offsets", ctypes.c_void_p)]


class _ScalarValue(ctypes.Union):
    _fields_ = [("i64", ctypes.c_int64), ("f64", ctypes.c_double)]


class Scalar(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("str_len", ctypes.c_int32),
                ("v", _ScalarValue), ("str", ctypes.c_char_p)]


class StepInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("op", ctypes.c_int32), ("out_type", ctypes.c_int32),
                ("n_args", ctypes.c_int32), ("arg_col", ctypes.c_int32 * 3),
                ("arg_str", ctypes.c_void_p * 3), ("arg_len", ctypes.c_int32 * 3),
                ("sub", ctypes.c_void_p), ("start", ctypes.c_int32)]


STRFN = {"CONCAT": 1, "UPPER": 2, "LOWER": 3, "INITCAP": 4, "REVERSE": 5, "REPLACE": 6, "CAST": 7,
         "CASE": 8, "FILL": 9, "REGEXP_REPLACE": 10}


class PlanInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("n_ops", "n_regs", "n_spill", "n_spill_ops", "n_fill_ops",
                                              "n_inputs", "n_load_issues", "max_loads_in_flight")]


class LaunchConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("rows_per_thread", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


LAUNCH_NO_RANGES, LAUNCH_LAZY_LOADS, LAUNCH_SPILL_WIDE, LAUNCH_FUSED, LAUNCH_GENERATED = 1, 2, 4, 8, 16
LAUNCH_LDS_REGS = 32
ROOT_INTERPRETED, ROOT_COLUMN, ROOT_LITERAL = 0, 1, 2
SKETCH_DIM, FEATURE_DIM, MLP_HIDDEN = 19, 24, 64


class BatchDesc(ctypes.Structure):
    _fields_ = [("prog", ctypes.c_void_p), ("inp", ctypes.c_void_p), ("n_in", ctypes.c_int),
                ("out", ctypes.c_void_p), ("n_out", ctypes.c_int), ("num_rows", ctypes.c_int64)]


# exported symbols and their signatures, as declared in include/bsql/interp.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
SIGNATURES = {
    "bsql_program_from_arrays": (_I, [_P, _P, _P, _P, _I, _P, _I, _P, _P, ctypes.POINTER(_P)]),
    "bsql_program_compile": (_I, [ctypes.POINTER(ctypes.c_char_p), _I, _P, _I, ctypes.POINTER(_P)]),
    "bsql_program_free": (None, [_P]),
    "bsql_program_num_ops": (_I, [_P]),
    "bsql_program_num_outputs": (_I, [_P]),
    "bsql_program_num_inputs": (_I, [_P]),
    "bsql_program_num_interpreted": (_I, [_P]),
    "bsql_program_root": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(Scalar)]),
    "bsql_program_num_parts": (_I, [_P]),
    "bsql_program_part": (_I, [_P, _I, ctypes.POINTER(_P), _P, _I]),
    "bsql_expression_rewrite": (_I, [ctypes.c_char_p, _I, _P, _I, _P, ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_size_t)]),
    "bsql_set_current_timestamp": (_I, [ctypes.c_char_p]),
    "bsql_program_plan_info": (_I, [_P, _P, _I, ctypes.POINTER(PlanInfo)]),
    "bsql_program_arrays": (_I, [_P, _P, _P, _P, _P, _P]),
    "bsql_program_scalar": (_I, [_P, _I, _I, ctypes.POINTER(Scalar)]),
    "bsql_program_output_type": (_I, [_P, _P, _I, _I, ctypes.POINTER(ctypes.c_int32)]),
    "bsql_program_used_columns": (_I, [_P, _P, _I]),
    "bsql_program_num_steps": (_I, [_P]),
    "bsql_program_table_columns": (_I, [_P]),
    "bsql_program_step": (_I, [_P, _I, ctypes.POINTER(StepInfo)]),
    "bsql_program_output_chars": (_I, [_P, _P, _I, _I64, ctypes.POINTER(_I64)]),
    "bsql_interp_execute": (_I, [_P, _P, _I, _P, _I, _I64, _P]),
    "bsql_filter": (_I, [_P, _P, _I, _P, _I64, _P, _P]),
    "bsql_filter_project": (_I, [_P, _P, _P, _I, _P, _I, _I64, _P, _P]),
    "bsql_filter_project_noalloc": (_I, [_P, _P, _P, _I, _P, _I, _I64, _P, _P]),
    "bsql_program_num_conjuncts": (_I, [_P]),
    "bsql_program_set_conjunct_order": (_I, [_P, _P, _I]),
    "bsql_program_set_launch": (_I, [_P, ctypes.POINTER(LaunchConfig)]),
    "bsql_generated_kernel_stats": (_I, [ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_double)]),
    "bsql_generated_kernel_check": (_I, [_P, _P, _P, _I, _P, _I, _I64, ctypes.c_char_p, ctypes.POINTER(_I64)]),
    "bsql_batch_prepare": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "bsql_batch_launch": (_I, [_P, _P]),
    "bsql_batch_set_launch": (_I, [_P, ctypes.POINTER(LaunchConfig)]),
    "bsql_batch_autotune": (_I, [_P, _P, ctypes.POINTER(LaunchConfig)]),
    "bsql_batch_total_rows": (_I64, [_P]),
    "bsql_batch_free": (None, [_P]),
    "bsql_program_autotune": (_I, [_P, _P, _I, _P, _I, _I64, _P, ctypes.POINTER(LaunchConfig)]),
    "bsql_filter_count": (_I, [_P, _P, _I, _I64, _P, _P]),
    "bsql_filter_project_sized": (_I, [_P, _P, _P, _I, _P, _I, _I64, _I64, _P, _P]),
    "bsql_filter_project_batch": (_I, [_P, _P, _I, _P, _I, _P, _P, _I64, _P, _P, _P]),
    "bsql_generated_batch_kernel_check": (_I, [_P, _P, _I, _P, _I, _P, _P, _I64, ctypes.c_char_p,
                                               ctypes.POINTER(_I64)]),
    "bsql_sketch": (_I, [_P, _I, _I64, _P, _P]),
    "bsql_mlp_create": (_I, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "bsql_mlp_num_params": (_I, [_P]),
    "bsql_mlp_get_params": (_I, [_P, _P]),
    "bsql_mlp_set_params": (_I, [_P, _P]),
    "bsql_mlp_forward": (_I, [_P, _P, _I, _P, _P]),
    "bsql_mlp_gradients": (_I, [_P, _P, _P, _I, ctypes.c_float, _P, _P, _P]),
    "bsql_mlp_train_step": (_I, [_P, _P, _P, _I, ctypes.c_float, ctypes.c_float, _P, _P]),
    "bsql_mlp_free": (None, [_P]),
    "bsql_conjunct_features": (_I, [_P, _P, _I, _P, _P, _I]),
    "bsql_order_by_prediction": (_I, [_P, _P, _P, _I]),
    "bsql_order_conjuncts": (_I, [_P, _P, _I, _P, _P, _I, _P]),
    "bsql_set_allocator": (_I, [_P, _P, _P]),
    "bsql_table_create": (_I, [_P, _P, _P, _P, _I, _I64, _P, ctypes.POINTER(_P)]),
    "bsql_table_from_views": (_I, [_P, _P, _I, _I64, ctypes.POINTER(_P)]),
    "bsql_table_free": (None, [_P]),
    "bsql_table_num_rows": (_I64, [_P]),
    "bsql_table_num_columns": (_I, [_P]),
    "bsql_table_column": (_I, [_P, _I, ctypes.POINTER(Column)]),
    "bsql_table_name": (ctypes.c_char_p, [_P, _I]),
    "bsql_table_set_names": (_I, [_P, _P, _I]),
    "bsql_table_record_stream": (_I, [_P, _P]),
    "bsql_process_filter": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_P), _P]),
    "bsql_process_project": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_P), _P]),
    "bsql_process_project_batches": (_I, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_char_p,
                                          ctypes.POINTER(_P), _P]),
    "bsql_table_size_in_bytes": (ctypes.c_int64, [_P]),
    "bsql_evaluate_expressions": (_I, [_P, _P, _I, ctypes.POINTER(_P), _P]),
    "bsql_apply_boolean_filter": (_I, [_P, ctypes.POINTER(Column), ctypes.POINTER(_P), _P]),
    "bsql_aggregation_input": (_I, [_P, ctypes.c_char_p, _I, ctypes.POINTER(_P), _P]),
    "bsql_skip_data_filter": (_I, [ctypes.c_char_p, _P, _I, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "bsql_process_skipdata_for_table": (_I, [_P, _P, _I, ctypes.c_char_p, ctypes.POINTER(_P),
                                             ctypes.POINTER(_I), _P]),
    "bsql_project_bypass": (_I, [ctypes.c_char_p, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "bsql_table_set_project_aliases": (_I, [_P, ctypes.c_char_p]),
    "bsql_table_fix_column_aliases": (_I, [_P, ctypes.c_char_p]),
    "bsql_scan_projections": (_I, [ctypes.c_char_p, _I, _P, _I]),
    "bsql_last_error": (ctypes.c_char_p, []),
    "bsql_version": (_I, []),
}

_lib = None

# Set when the interpreter starts to exit: finalizers then leave native
# objects to the process teardown instead of freeing device memory, since
# the HIP runtime (or a profiler wrapping it) may already be shutting down.
_exiting = False


def _mark_exiting():
    global _exiting
    _exiting = True


atexit.register(_mark_exiting)


def exiting():
    return _exiting


def lib():
    """Loads libbsql.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libbsql.so is not built: run python blazingsql_amd/build.py "
                "(or __graft_entry__.build())")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7.
        # Loading torch first makes libbsql.so bind to that same runtime (same
        # SONAME); loading ours first would bring in a second copy from
        # /opt/rocm, and the two runtimes do not share the device.
        import torch  # noqa: F401
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class BsqlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bsql error {code}: {msg}")
        self.code = code


class OutOfMemory(BsqlError):
    pass


def check(code):
    if code != OK:
        msg = lib().bsql_last_error().decode(errors="replace")
        if code == ERR_OOM:
            raise OutOfMemory(code, msg)
        raise BsqlError(code, msg)
    return code
