"""Writes parser_vectors.json: the (input, expected) text pairs of BlazingSQL's
own unit tests for the expression front end, as data:

  engine/tests/parser/expression_tree_test.cpp:12-76
      parse_tree::build + transform_to_custom_op + rebuildExpression
  engine/tests/parser/expression_utils_test.cpp:243-368
      fill_minus_op_with_zero, convert_nary_to_binary_concat
  engine/tests/parser/expression_utils_test.cpp:403-413, 507-642
      convert_ms_to_ns_units, reinterpret_timestamp (with the test's schema)

The test sources are read as text (string literals, schemas and which
function each test calls); nothing of the reference is imported or run.

Run: python tests/golden/make_parser_vectors.py [reference root]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

TYPE_IDS = {"INT8": 1, "INT16": 2, "INT32": 3, "INT64": 4, "FLOAT32": 9, "FLOAT64": 10, "BOOL8": 11,
            "TIMESTAMP_DAYS": 12, "TIMESTAMP_SECONDS": 13, "TIMESTAMP_MILLISECONDS": 14,
            "TIMESTAMP_MICROSECONDS": 15, "TIMESTAMP_NANOSECONDS": 16, "STRING": 23}

STAGES = {"fill_minus_op_with_zero": 0, "convert_nary_to_binary_concat": 1, "convert_ms_to_ns_units": 2,
          "reinterpret_timestamp": 3}

C_STRING = r'"(?:[^"\\]|\\.)*"'


def c_strings(text):
    """The value of adjacent C string literals ("a" "b" -> "ab")."""
    return "".join(bytes(m[1:-1], "utf-8").decode("unicode_escape") for m in re.findall(C_STRING, text))


def tests_of(src, fixture):
    """(name, body, first line) of each TEST_F(fixture, name) in a source text."""
    heads = list(re.finditer(r"TEST_F\(\s*" + fixture + r",\s*(\w+)\s*\)", src))
    for i, h in enumerate(heads):
        end = heads[i + 1].start() if i + 1 < len(heads) else len(src)
        yield h.group(1), src[h.end():end], src.count("\n", 0, h.start()) + 1


def assigned(body, var):
    m = re.search(r"std::string\s+" + var + r"\s*=\s*((?:\s*" + C_STRING + r")+)\s*;", body)
    return c_strings(m.group(1)) if m else None


def main(ref):
    out = {"custom_op": [], "text": []}
    path = os.path.join(ref, "engine/tests/parser/expression_tree_test.cpp")
    src = open(path).read()
    for name, body, line in tests_of(src, "ExpressionTreeTest"):
        m = re.search(r"tree\.build\(((?:\s*" + C_STRING + r")+)\)", body)
        exp = assigned(body, "expected")
        if m and exp is not None:
            out["custom_op"].append({"source": f"engine/tests/parser/expression_tree_test.cpp:{line}",
                                     "name": name, "input": c_strings(m.group(1)), "expected": exp})
    path = os.path.join(ref, "engine/tests/parser/expression_utils_test.cpp")
    src = open(path).read()
    for name, body, line in tests_of(src, "ExpressionUtilsTest"):
        fn = next((f for f in STAGES if re.search(r"\b" + f + r"\s*\(", body)), None)
        if fn is None:
            continue
        inp = assigned(body, "expression")
        if inp is None:
            continue
        schema = [TYPE_IDS[t] for t in re.findall(r"cudf::type_id::(\w+)", body)]
        item = {"source": f"engine/tests/parser/expression_utils_test.cpp:{line}", "name": name,
                "stage": STAGES[fn], "input": inp, "schema": schema}
        if "FAIL()" in body:
            item["error"] = True
        else:
            exp = assigned(body, "expected_str")
            if exp is None:
                exp = assigned(body, "expected_expression")
            if exp is None:
                m = re.search(r"EXPECT_EQ\(\s*\w+\s*,\s*((?:\s*" + C_STRING + r")+)\s*\)", body)
                exp = c_strings(m.group(1)) if m else None
            if exp is None and re.search(r"EXPECT_EQ\(\s*\w+\s*,\s*expression\s*\)", body):
                exp = inp  # the function leaves the text unchanged
            if exp is None:
                exp = inp if "EXPECT" not in body else None  # a call whose result the test ignores
            if exp is None:
                continue
            item["expected"] = exp
        out["text"].append(item)
    dst = os.path.join(HERE, "parser_vectors.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"{dst}: {len(out['custom_op'])} custom-op cases, {len(out['text'])} text cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
