"""Kernels generated for one program compile for gfx950 without a GPU
(bsql_generated_kernel_check: the lowering and source of a real call, hipRTC
compile, nothing loaded or launched), and calls that cannot run as a
generated kernel say so. CPU only."""
import pytest

from blazingsql_amd import _lib as L

INT32, FLOAT64 = 3, 10
SCAN_TYPES = [INT32, FLOAT64, FLOAT64, FLOAT64, INT32]
SCAN_PRED = "AND(<=($4, 10471), >=($2, 0.05:DECIMAL(3, 2)), <=($2, 0.07:DECIMAL(3, 2)), <($0, 24))"
SCAN_PROJ = ["*(*($1, -(1, $2)), +(1, $3))", "*($1, $2)"]


def _col(t, n, valid=False, addr=0x10000):
    c = L.Column()
    c.type_id = t
    c.size = n
    c.data = addr
    c.validity = addr + 0x100000 if valid else None
    return c


def test_scan_filter_project_kernel_compiles(bsql):
    from blazingsql_amd.interop import check_generated_kernel
    n = 100_000_000
    pred = bsql.Program.compile(SCAN_PRED, SCAN_TYPES)
    proj = bsql.Program.compile(SCAN_PROJ, SCAN_TYPES)
    nbytes = check_generated_kernel(pred, proj, [_col(t, n) for t in SCAN_TYPES],
                                    [_col(FLOAT64, n, True), _col(FLOAT64, n, True)], n)
    assert nbytes > 1000


def test_execute_kernel_compiles_with_nulls(bsql):
    from blazingsql_amd.interop import check_generated_kernel
    types = [INT32, FLOAT64, 12]
    prog = bsql.Program.compile(["CASE(IS NOT NULL($0), +($0, 1), -1)", "FLOOR(*($1, 2.5:DECIMAL(2, 1)))",
                                 "EXTRACT(FLAG(YEAR), $2)"], types)
    ot = [prog.output_type(types, k) for k in range(3)]
    n = 1000
    nbytes = check_generated_kernel(None, prog, [_col(types[0], n, True), _col(FLOAT64, n), _col(12, n, True)],
                                    [_col(t, n, True) for t in ot], n)
    assert nbytes > 1000


def test_calls_that_stay_on_the_interpreter(bsql):
    from blazingsql_amd._lib import BsqlError
    from blazingsql_amd.interop import check_generated_kernel
    n = 4096
    types = [INT32, FLOAT64]
    cols = [_col(INT32, n), _col(FLOAT64, n)]
    # a RAND output numbers rows by output position: not the one-pass kernel
    pred = bsql.Program.compile(">($0, 3)", types)
    proj = bsql.Program.compile(["RAND()"], types)
    with pytest.raises(BsqlError, match="generated kernel"):
        check_generated_kernel(pred, proj, cols, [_col(FLOAT64, n, True)], n)
    # a predicate that is not made of range tests
    pred = bsql.Program.compile(">(*($0, 2), $1)", types)
    proj = bsql.Program.compile(["+($0, 1)"], types)
    with pytest.raises(BsqlError, match="generated kernel"):
        check_generated_kernel(pred, proj, cols, [_col(4, n, True)], n)
    with pytest.raises(BsqlError):
        check_generated_kernel(pred, proj, cols, [_col(4, n, True)], 0)


def test_batch_kernel_compiles(bsql):
    """Three queries over one table in one generated kernel (the scan's
    predicate and projection, a range on another column with a
    one-output projection, and a nullable table column)."""
    from blazingsql_amd.interop import check_generated_batch_kernel
    n = 1_000_000
    preds = [bsql.Program.compile(SCAN_PRED, SCAN_TYPES), bsql.Program.compile("<($1, 5000.0:DECIMAL(5, 1))", SCAN_TYPES),
             bsql.Program.compile("AND(>($0, 10), <($4, 9000))", SCAN_TYPES)]
    projs = [bsql.Program.compile(SCAN_PROJ, SCAN_TYPES), bsql.Program.compile(["+($0, $4)"], SCAN_TYPES),
             bsql.Program.compile(["*($3, 2:DOUBLE)", "$0"], SCAN_TYPES)]
    ins = [_col(t, n, valid=(i == 0)) for i, t in enumerate(SCAN_TYPES)]
    outs = [[_col(FLOAT64, n, True), _col(FLOAT64, n, True)], [_col(4, n, True)],
            [_col(FLOAT64, n, True), _col(INT32, n, True)]]
    nbytes = check_generated_batch_kernel(preds, projs, ins, outs, n)
    assert nbytes > 1000
