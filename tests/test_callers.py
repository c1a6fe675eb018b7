"""The interpreter's other callers, host side (CPU): the skip-data rewrite of
a table scan's filter over row-group statistics (SkipDataProcessor.cpp:
23-246) and the multi-column COUNT input (expression_utils.cpp:1467-1492).
The expected strings follow the reference's rules by hand: comparisons over
the ranges (= : min <= other max and max >= other min; < <= : min vs other
max; > >= : max vs other min), + and - as (min op min, max op max), column
i's minimum at 2i and maximum at 2i + 1, unsupported operators and columns
without statistics dropped (an AND keeps its other side, anything else
gives up)."""
import pytest

from blazingsql_amd.interop import rewrite
from blazingsql_amd.processor import skip_data_filter

CASES = [
    ("=($0, 5)", [1], "AND(<=($0, 5), >=($1, 5))"),
    ("<($0, 7)", [1], "<($0, 7)"),
    ("<=(7, $0)", [1], "<=(7, $1)"),
    (">($0, 7)", [1], ">($1, 7)"),
    (">=(7, $0)", [1], ">=(7, $0)"),
    ("AND(<($0, 15000), =($1, 5))", [1, 1], "AND(<($0, 15000), AND(<=($2, 5), >=($3, 5)))"),
    ("AND(>($0, 3), <($2, 4))", [1, 1, 0], ">($1, 3)"),
    ("OR(>($0, 3), <($2, 4))", [1, 1, 0], ""),
    ("OR(>($0, 3), <($1, 4))", [1, 1], "OR(>($1, 3), <($2, 4))"),
    (">(+($0, $1), 10)", [1, 1], ">(+($1, $3), 10)"),
    ("<(-($0, 2), $1)", [1, 1], "<(-($0, 2), $3)"),
    ("=(+($0, 1), $1)", [1, 1], "AND(<=(+($0, 1), $3), >=(+($1, 1), $2))"),
    ("AND(IS_NOT_NULL($0), >($1, 2))", [1, 1], ">($3, 2)"),
    ("NOT(>($0, 1))", [1], ""),
    ("OR(AND(<($0, 15000), =($1, 5)), =($0, *($1, $1)), >=($1, 10), <=($2, 500))", [1, 1, 1], ""),
    ("AND(<($0, 15000), =($1, 5), >=($1, 10), <=($2, 500))", [1, 1, 1],
     "AND(AND(AND(<($0, 15000), AND(<=($2, 5), >=($3, 5))), >=($3, 10)), <=($4, 500))"),
    ("AND(>=($3, 1994-01-01 00:00:00), <($3, 1995-01-01 00:00:00), >=($2, 0.05:DECIMAL(4, 2)))", [1, 1, 1, 1],
     "AND(AND(>=($7, 1994-01-01 00:00:00), <($6, 1995-01-01 00:00:00)), >=($5, 0.05))"),
]


@pytest.mark.parametrize("filt,stats,expect", CASES)
def test_skip_data_rewrite(bsql, filt, stats, expect):
    assert skip_data_filter(filt, stats) == expect


def test_multi_column_count_input(bsql):
    assert rewrite("$1, $2", 9) == "+(CAST($1):INTEGER, CAST($2):INTEGER)"
    assert rewrite("$1, $2, $4", 9) == "+(+(CAST($1):INTEGER, CAST($2):INTEGER), CAST($4):INTEGER)"
    assert rewrite("$3", 9) == "$3"
    assert rewrite("", 9) == ""
