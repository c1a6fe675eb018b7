"""Every filter and projection of the reference's golden TPC-H plans
(tests/golden/calcite_plans.txt, the plan strings of the algebra module's
tpch-reference.bin / selected-queries-reference.bin) runs on the GPU over a
small typed table and matches the oracle bit-exact.

The plans do not carry column types: each $k is typed from how the
expressions use it (a string when it meets a string literal or a string
function, a date when it meets a date literal or EXTRACT, a boolean when it
is an operand of AND/OR/NOT, a double otherwise), and the table is drawn
from values that make the comparisons in the plan select some rows."""
import os
import re

import numpy as np
import pytest

from tests.helpers import assert_same

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INT32, INT64, FLOAT64, BOOL8, TS_D, STRING = 3, 4, 10, 11, 12, 23
STRING_FNS = ("LIKE", "SUBSTRING", "CHAR_LENGTH", "CONCAT", "||", "UPPER", "LOWER", "TRIM", "REPLACE")


def _split_items(body):
    items, depth, cur, quoted = [], 0, "", False
    for c in body:
        if c == "'":
            quoted = not quoted
        if not quoted:
            if c in "([":
                depth += 1
            elif c in ")]":
                depth -= 1
            elif c == "," and depth == 0:
                items.append(cur.strip())
                cur = ""
                continue
        cur += c
    if cur.strip():
        items.append(cur.strip())
    return items


def _plan_parts():
    """(plan, kind, [expressions]) for each filter and projection of each plan."""
    text = open(os.path.join(HERE, "golden", "calcite_plans.txt")).read()
    out = []
    plan = -1
    for line in text.splitlines():
        if line.startswith("#### plan"):
            plan = int(line.split()[-1])
            continue
        s = line.strip()
        for key in ("filters=[[", "condition=["):
            at = s.find(key)
            if at >= 0 and "LogicalJoin" not in s:
                depth, i = 0, at + len(key)
                start = i
                while i < len(s):
                    if s[i] in "([":
                        depth += 1
                    elif s[i] in ")]":
                        if depth == 0:
                            break
                        depth -= 1
                    i += 1
                out.append((plan, "filter", [s[start:i]]))
        if s.startswith("LogicalProject("):
            body = s[len("LogicalProject("):s.rindex(")")]
            exprs = [it[it.index("=[") + 2:-1] for it in _split_items(body)]
            if any(not re.fullmatch(r"\$\d+", e) for e in exprs):
                out.append((plan, "project", exprs))
    return out


def _infer_types(exprs):
    text = " ".join(exprs)
    n = max([int(x) for x in re.findall(r"\$(\d+)", text)] + [0]) + 1
    types = [FLOAT64] * n
    for k in range(n):
        v = rf"\${k}(?!\d)"
        if re.search(rf"(?:{'|'.join(re.escape(f) for f in STRING_FNS)})\({v}", text) or \
                re.search(rf"{v}(?:\):\w+)?, '", text) or re.search(rf"'[^']*', {v}", text) or \
                re.search(rf"[=<>]\({v}, '", text):
            types[k] = STRING
        elif re.search(rf"{v}, \d{{4}}-\d{{2}}-\d{{2}}", text) or re.search(rf"EXTRACT\(FLAG\(\w+\), {v}\)", text) \
                or re.search(rf"\d{{4}}-\d{{2}}-\d{{2}}[ \d:]*, {v}", text):
            types[k] = TS_D
        elif re.search(rf"(?:AND|OR|NOT)\({v}[,)]", text) or re.search(rf", {v}\)", text) and \
                re.search(rf"(?:AND|OR)\([^()]*, {v}\)", text):
            types[k] = BOOL8
    return types


def _literals(text):
    return [m for m in re.findall(r"'((?:[^']|'')*)'", text)]


def _table(rng, types, n, text):
    from blazingsql_amd.interop import DeviceColumn
    from oracle import HostColumn
    lits = _literals(text) or ["x"]
    dev, host = [], []
    for t in types:
        valid = rng.random(n) >= 0.1
        if t == STRING:
            # the plan's literals (and a few variations) so that =, LIKE and
            # SUBSTRING select some rows
            pool = lits + [l.replace("%", "") + "ab" for l in lits] + ["Brand#45", "MEDIUM POLISHED x", "zz"]
            pool = [p.replace("%", "Q").replace("_", "q") for p in pool]
            strs = [pool[int(i)] for i in rng.integers(0, len(pool), n)]
            dev.append(DeviceColumn.from_strings(strs, valid))
            host.append(HostColumn(STRING, strings=strs, valid=valid))
            continue
        if t == TS_D:
            v = rng.integers(8000, 10600, n).astype(np.int32)  # 1991..1999
        elif t == BOOL8:
            v = (rng.random(n) < 0.5).astype(np.uint8)
        else:
            v = np.round(rng.uniform(-5, 60, n), 2)
        dev.append(DeviceColumn.from_numpy(t, v, valid))
        host.append(HostColumn(t, v, valid))
    return dev, host


PARTS = _plan_parts()


@pytest.mark.parametrize("idx", range(len(PARTS)))
def test_golden_plan_part_vs_oracle(bsql, idx):
    import oracle
    import torch
    from blazingsql_amd._lib import BsqlError
    from blazingsql_amd.interop import DeviceColumn, filter_columns, perform_interpreter_operation
    plan, kind, exprs = PARTS[idx]
    types = _infer_types(exprs)
    try:
        p = bsql.Program.compile(exprs, types)
    except BsqlError as e:
        # the few expressions of the plans that need the engine around the
        # interpreter (a subquery's $cor variables, aggregates' COUNT)
        pytest.skip(f"plan {plan} {kind}: {e}")
    rng = np.random.default_rng(idx)
    n = 3_000
    dev, host = _table(rng, types, n, " ".join(exprs))
    print(f"plan {plan} {kind}: {len(exprs)} expressions over {len(types)} columns", flush=True)
    if kind == "filter":
        outs = [DeviceColumn.empty(t, n, nullable=True, chars=c.chars_size if t == STRING else 0)
                for t, c in zip(types, dev)]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        filter_columns(p, dev, outs, cnt, n)
        torch.cuda.synchronize()
        m = int(cnt.item())
        on, ocols = oracle.filter_table(oracle.full_table_program(p, len(types)), host, n)
        assert m == on, (plan, exprs)
        for k, (o, (ov, oval)) in enumerate(zip(outs, ocols)):
            gv, gval = o.to_numpy(m)
            if types[k] == STRING:
                assert [g for g, ok in zip(gv, oval) if ok] == [x for x, ok in zip(ov, oval) if ok]
                assert np.array_equal(np.asarray(gval, bool), np.asarray(oval, bool))
            else:
                assert_same(gv, gval, ov, oval, what=f"plan {plan} filtered column {k}")
        return
    ot = [p.output_type(types, k) for k in range(p.num_outputs)]
    chars = p.output_chars(dev, n)
    outs = [DeviceColumn.empty(t, n, chars=chars) for t in ot]
    perform_interpreter_operation(p, dev, outs, n)
    ora = oracle.evaluate(p, host, n)
    for k in range(len(ot)):
        gv, gval = outs[k].to_numpy(n)
        ov, oval = ora[k]
        if ot[k] == STRING:
            assert np.array_equal(np.asarray(gval, bool), np.asarray(oval, bool))
            assert [g for g, ok in zip(gv, oval) if ok] == [x for x, ok in zip(ov, oval) if ok]
        else:
            assert_same(gv, gval, ov, oval, what=f"plan {plan} expression {k}: {exprs[k]}")


def test_golden_plans_cover_the_filters_and_projections():
    kinds = [k for _, k, _ in PARTS]
    assert kinds.count("filter") >= 20 and kinds.count("project") >= 20
