"""Writes reference_vectors.json: the input/expected-output vectors of
BlazingSQL's own unit tests for the interpreter/filter path, transcribed as
data (paths relative to the BlazingSQL 21.08 tree). Nothing here imports or
runs the reference; the timestamp inputs are recomputed with the formula the
tests use (start + (range / rows) * i, C++ integer division).

Run: python tests/golden/make_reference_vectors.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


UNITS = {"D": 86400000, "s": 1000, "ms": 1, "us": None, "ns": None}


def ms_to_unit(ms, unit):
    # cuda::std::chrono::time_point_cast from milliseconds (truncation toward zero)
    if unit == "D":
        return trunc_div(ms, 86400000)
    if unit == "s":
        return trunc_div(ms, 1000)
    if unit == "ms":
        return ms
    if unit == "us":
        return ms * 1000
    return ms * 1000000


def series(start_ms, stop_ms, rows, unit):
    start, stop = ms_to_unit(start_ms, unit), ms_to_unit(stop_ms, unit)
    step = trunc_div(stop - start, rows)
    return [start + step * i for i in range(rows)]


TS_TYPES = {"D": 12, "s": 13, "ms": 14, "us": 15, "ns": 16}

vectors = {}

# engine/tests/new-tests/interops_tests.cpp:199-278 -- hand-encoded program
# "+ * + $0 $1 $2 $1 | + $1 2" over col1=2r, col2=10r, col3=20r (10 rows)
vectors["numeric_program"] = {
    "source": "engine/tests/new-tests/interops_tests.cpp:199-278",
    "left": [0, 5, 5, 1], "right": [1, 2, 1, -2], "out": [5, 5, 3, 4], "final_out": [3, 4],
    "ops": ["ADD", "MUL", "ADD", "ADD"], "right_scalar_op": 3, "right_scalar_value": 2,
    "types": {"int8": 1, "int16": 2, "int32": 3, "int64": 4, "bool": 11, "float": 9, "double": 10},
    "rows": 10,
    "col1": [2 * r for r in range(10)], "col2": [10 * r for r in range(10)],
    "col3": [20 * r for r in range(10)],
    "expected_col1": [0, 250, 980, 2190, 3880, 6050, 8700, 11830, 15440, 19530],
    "expected_col2": [2, 12, 22, 32, 42, 52, 62, 72, 82, 92],
}

# interops_tests.cpp:294-455 -- BLZ_DAYOFWEEK over 20 timestamps per unit
vectors["day_of_week"] = {
    "source": "engine/tests/new-tests/interops_tests.cpp:294-455",
    "inputs": {u: series(-2203891200000, 2214086400000, 20, u) for u in TS_TYPES},
    "types": TS_TYPES,
    "expected_days": [4, 5, 6, 7, 1, 2, 3, 4, 5, 6, 7, 1, 2, 3, 4, 5, 6, 7, 1, 2],
    "expected_other": [4, 5, 7, 2, 3, 5, 7, 1, 3, 5, 7, 1, 3, 5, 6, 1, 3, 4, 6, 1],
    "expression": "BL_DOW($0)", "project": "LogicalProject(EXPR$0=[EXTRACT(FLAG(DOW), $0)])",
}

# interops_tests.cpp:457-557 -- YEAR MONTH DAY HOUR MINUTE SECOND DOW
vectors["timestamp_components"] = {
    "source": "engine/tests/new-tests/interops_tests.cpp:457-557",
    "inputs": {u: series(-2500000000000, 2500000000000, 10, u) for u in TS_TYPES},
    "types": TS_TYPES,
    "ops": ["YEAR", "MONTH", "DAY", "HOUR", "MINUTE", "SECOND", "DAYOFWEEK"],
    "expected_days": [
        [1890, 1906, 1922, 1938, 1954, 1970, 1985, 2001, 2017, 2033],
        [10, 8, 6, 4, 2, 1, 11, 9, 7, 5],
        [12, 17, 21, 25, 27, 1, 5, 9, 14, 18],
        [0] * 10, [0] * 10, [0] * 10,
        [7, 5, 3, 1, 6, 4, 2, 7, 5, 3]],
    "expected_other": [
        [1890, 1906, 1922, 1938, 1954, 1970, 1985, 2001, 2017, 2033],
        [10, 8, 6, 4, 2, 1, 11, 9, 7, 5],
        [11, 16, 20, 24, 26, 1, 5, 9, 14, 18],
        [19, 20, 21, 22, 23, 0, 0, 1, 2, 3],
        [33, 26, 20, 13, 6, 0, 53, 46, 40, 33],
        [20, 40, 0, 20, 40, 0, 20, 40, 0, 20],
        [6, 4, 2, 7, 5, 4, 2, 7, 5, 3]],
}

# engine/tests/new-tests/filter_tests.cpp:75-171 -- MOD filter, 40 rows, with/without nulls
vectors["filter_mod"] = {
    "source": "engine/tests/new-tests/filter_tests.cpp:75-171",
    "query": "LogicalFilter(condition=[=(MOD(+($0, $1), 2), 0)])",
    "expression": "=(MOD(+($0, $1), 2), 0)",
    "reference_encoding": {"left": [0, 3, 3], "right": [1, -2, -2], "out": [3, 3, 2],
                           "final_out": [2], "ops": ["ADD", "MOD", "EQUAL"],
                           "right_scalars": [None, [3, 2], [3, 0]]},
    "types": {"int8": 1, "int16": 2, "int32": 3, "int64": 4, "uint8": 5, "uint16": 6,
              "uint32": 7, "uint64": 8, "float": 9, "double": 10, "bool": 11},
    "rows": 40, "col1": list(range(40)), "col2": [2 * r for r in range(40)],
    "null_validity": [r % 2 == 0 for r in range(40)],
    "expected_col1": [2 * r for r in range(20)], "expected_col2": [4 * r for r in range(20)],
    "empty_query": "LogicalFilter(condition=[<(+($0, $1), 0)])",
}

# filter_tests.cpp:178-214 -- string predicates
vectors["filter_strings"] = {
    "source": "engine/tests/new-tests/filter_tests.cpp:178-214",
    "col1": ["foo", "d", "e", "a", "hello", "k", "d", "l", "bar", ""],
    "col2": [10, 8, 6, 4, 2, 1, 11, 9, 7, 5],
    "eq_query": "LogicalFilter(condition=[=($0, 'bar')])",
    "eq_expected": {"col1": ["bar"], "col2": [7]},
    "null_validity": [1, 1, 0, 1, 1, 0, 1, 1, 0, 1],
    "gt_query": "LogicalFilter(condition=[>($0, 'd')])",
    "gt_expected": {"col1": ["foo", "hello", "l"], "col2": [10, 2, 9]},
}

# engine/tests/new-tests/process_project.cpp:129-199 -- comparisons
vectors["project_compare"] = {
    "source": "engine/tests/new-tests/process_project.cpp:129-199",
    "col1": [4, 5, 3, 5, 8, 5, 6], "col3": [10, 40, 70, 5, 2, 10, 11],
    "query": "LogicalProject(EXPR$0=[>($0, 3)], EXPR$1=[<($2, 3)])",
    "expected_col1": [1, 1, 0, 1, 1, 1, 1], "expected_col2": [0, 0, 0, 0, 1, 0, 0],
    "expected_bool_col1": [0] * 7, "expected_bool_col2": [1] * 7,
}

# process_project.cpp:350-366 -- CHAR_LENGTH with a null
vectors["char_length"] = {
    "source": "engine/tests/new-tests/process_project.cpp:350-366",
    "col1": ["abc", "12345", "", "xyzw"], "valid": [1, 1, 0, 1],
    "query": "LogicalProject(EXPR$0=[-(CHAR_LENGTH($0), 1)])",
    "expected": [2, 4, -1, 3], "expected_valid": [1, 1, 0, 1],
}

# process_project.cpp:286-303 -- LIKE with '_' and a leading '%'
vectors["string_like"] = {
    "source": "engine/tests/new-tests/process_project.cpp:286-303",
    "col1": ["foo", "d", "e", "a", "hello", "k", "d", "l", "bar", ""],
    "query": "LogicalProject(EXPR$0=[LIKE($0, '_')], EXPR$1=[LIKE($0, '%o')])",
    "expected_col1": [0, 1, 1, 1, 0, 1, 1, 1, 0, 0],
    "expected_col2": [1, 0, 0, 0, 1, 0, 0, 0, 0, 0],
}

# process_project.cpp:305-321 -- SUBSTRING with a negative literal length (to the end)
vectors["string_substring"] = {
    "source": "engine/tests/new-tests/process_project.cpp:305-321",
    "col1": ["The", "quick", "brown", "fox", "jumps", "over", "the", "lazy", "dog"],
    "query": "LogicalProject(EXPR$0=[SUBSTRING($0, 3, -1)])",
    "expected": ["e", "ick", "own", "x", "mps", "er", "e", "zy", "g"],
}

# process_project.cpp:323-342 -- SUBSTRING with a computed length and a null row
vectors["string_substring_null"] = {
    "source": "engine/tests/new-tests/process_project.cpp:323-342",
    "col1": ["ABC", "123", "", "XYZ"], "valid": [1, 1, 0, 1],
    "query": "LogicalProject(EXPR$0=[SUBSTRING($0, 2, -(CHAR_LENGTH($0), 2))])",
    "expected": ["B", "2", "", "Y"], "expected_valid": [1, 1, 0, 1],
}

# process_project.cpp:368-387 -- || with a literal and between columns
vectors["string_concat"] = {
    "source": "engine/tests/new-tests/process_project.cpp:368-387",
    "col1": ["foo", "d", "e", "a", "hello", "k", "d", "l", "", ""],
    "col2": ["The", "quick", "brown", "fox", "jumps", "over", "the", "lazy", "dog", ""],
    "query": "LogicalProject(EXPR$0=[||($0,'XD=D')], EXPR$1=[||($0,$1)])",
    "expected_col1": ["fooXD=D", "dXD=D", "eXD=D", "aXD=D", "helloXD=D", "kXD=D", "dXD=D", "lXD=D", "XD=D", "XD=D"],
    "expected_col2": ["fooThe", "dquick", "ebrown", "afox", "hellojumps", "kover", "dthe", "llazy", "dog", ""],
}

# process_project.cpp:389-412 -- CAST to VARCHAR of BOOL8, INT32, FLOAT64, TIMESTAMP_SECONDS
vectors["cast_to_string"] = {
    "source": "engine/tests/new-tests/process_project.cpp:389-412",
    "col1": [True, False, True, True, True, False, False, True, False, False],
    "col2": [1, 5, 10, 15, 100, 500, 1000, 5000, 10000, 999999],
    "col3": [1.0, 5.5, 10.00003, 15.45, 100.656, 500.756756, 0.45435436, 0.0000324, 0.1, 999999.001],
    "col4": [0, 10, 2600, 89260, 579500, 6834000, 86796400, 135768000, 715360000, 1230720000],
    "query": "LogicalProject(EXPR$0=[CAST($0):VARCHAR], EXPR$1=[CAST($1):VARCHAR], EXPR$2=[CAST($2):VARCHAR], "
             "EXPR$3=[CAST($3):VARCHAR])",
    "expected_col1": ["true", "false", "true", "true", "true", "false", "false", "true", "false", "false"],
    "expected_col2": ["1", "5", "10", "15", "100", "500", "1000", "5000", "10000", "999999"],
    "expected_col3": ["1.0", "5.5", "10.00003", "15.45", "100.656", "500.756756", "0.45435436", "3.24e-05", "0.1",
                      "999999.001"],
    "expected_col4": ["1970-01-01 00:00:00", "1970-01-01 00:00:10", "1970-01-01 00:43:20", "1970-01-02 00:47:40",
                      "1970-01-07 16:58:20", "1970-03-21 02:20:00", "1972-10-01 14:06:40", "1974-04-21 09:20:00",
                      "1992-09-01 15:06:40", "2008-12-31 10:40:00"],
}

# process_project.cpp:447-486 -- CASE with string values, and compared to a literal
vectors["string_case"] = {
    "source": "engine/tests/new-tests/process_project.cpp:447-486",
    "col1": [0, 1, 2, 3, 4, 5, 6, 7, 8],
    "col2": ["The", "quick", "brown", "fox", "jumps", "over", "the", "lazy", "dog"],
    "query": "LogicalProject(EXPR$0=[CASE(=(MOD($0, 2), 0), $1, 'LOL')])",
    "expected": ["The", "LOL", "brown", "LOL", "jumps", "LOL", "the", "LOL", "dog"],
    "nested_query": "LogicalProject(EXPR$0=[=(CASE(=(MOD($0, 2), 0), $1, 'LOL'), 'LOL')])",
    "nested_expected": [False, True, False, True, False, True, False, True, False],
}

# process_project.cpp:415-446 -- CAST from strings to INTEGER, DOUBLE and TIMESTAMP
vectors["cast_from_string"] = {
    "source": "engine/tests/new-tests/process_project.cpp:415-446",
    "col1": ["1", "5", "10", "15", "100", "500", "1000", "5000", "10000", "999999"],
    "col2": ["1.0", "5.5", "10.00003", "15.45", "100.656", "500.756756", "0.45435436", "3.24e-05", "0.1",
             "999999.001"],
    "col3": ["1970-01-01 00:00:00", "1970-01-01 00:00:10", "1970-01-01 00:43:20", "1970-01-02 00:47:40",
             "1970-01-07 16:58:20", "1970-03-21 02:20:00", "1972-10-01 14:06:40", "1974-04-21 09:20:00",
             "1992-09-01 15:06:40", "2008-12-31 10:40:00"],
    "query": "LogicalProject(EXPR$0=[CAST($0):INTEGER], EXPR$1=[CAST($1):DOUBLE], EXPR$2=[CAST($2):TIMESTAMP])",
    "expected_col3": [0, 10000000000, 2600000000000, 89260000000000, 579500000000000, 6834000000000000,
                      86796400000000000, 135768000000000000, 715360000000000000, 1230720000000000000],
    "expected_col1": [1, 5, 10, 15, 100, 500, 1000, 5000, 10000, 999999],
    "expected_col2": [1.0, 5.5, 10.00003, 15.45, 100.656, 500.756756, 0.45435436, 0.0000324, 0.1, 999999.001],
    "compare_col2": "expect_tables_equivalent: floating point within cudf's ULP tolerance",
}

# process_project.cpp:489-513 -- ROUND with 0, 2 and 5 digits (float and double)
vectors["round"] = {
    "source": "engine/tests/new-tests/process_project.cpp:489-513",
    "col1": [4.0, 5.21, 87317.3, 0.1232387, 0.0000007, 342.9348, 698.3243],
    "query": "LogicalProject(EXPR$0=[ROUND($0)], EXPR$1=[ROUND($0, 2)], EXPR$2=[ROUND($0, 5)])",
    "expected": [[4.0, 5.0, 87317.0, 0.0, 0.0, 343.0, 698.0],
                 [4.00, 5.21, 87317.30, 0.12, 0.00, 342.93, 698.32],
                 [4.00000, 5.21000, 87317.30000, 0.12324, 0.00000, 342.93480, 698.32430]],
}

# engine/tests/new-tests/apply_filter.cpp:44-84 -- compaction with nulls (mask null drops)
vectors["apply_filter"] = {
    "source": "engine/tests/new-tests/apply_filter.cpp:44-116",
    "col1": [5, 4, 3, 5, 8, 5, 6], "col1_valid": [1, 1, 0, 1, 1, 1, 1],
    "col2": ["d", "e", "a", "d", "k", "d", "l"], "col2_valid": [1, 0, 1, 1, 1, 1, 1],
    "col3": [10, 40, 70, 5, 2, 10, 11], "col3_valid": [0, 1, 1, 1, 1, 1, 0],
    "mask": [1, 1, 1, 0, 1, 0, 0], "mask_valid": [1, 1, 0, 1, 1, 1, 0],
    "expected_col1": [5, 4, 8], "expected_col1_valid": [1, 1, 1],
    "expected_col2": ["d", "e", "k"], "expected_col2_valid": [1, 0, 1],
    "expected_col3": [10, 40, 2], "expected_col3_valid": [0, 1, 1],
}

# engine/tests/cache_machine/exception_handling_test.cpp:192-208 -- literal-literal op fails
vectors["errors"] = {
    "source": "engine/tests/cache_machine/exception_handling_test.cpp:174-208",
    "literal_literal": "POWER(-1, 0.5:DECIMAL(2, 1))",
}

with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
    json.dump(vectors, f, indent=1)
print("wrote", os.path.join(HERE, "reference_vectors.json"))
