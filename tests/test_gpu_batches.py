"""GPU parity for the batched path: uneven batches, each with its own random
program (arithmetic, comparisons, CASE, COALESCE, CAST, EXTRACT) and its own
columns with nulls, run as one persistent launch -- against the oracle per
batch, eagerly and replayed from a captured HIP graph; the autotuner's choice
gives the same results."""
import numpy as np
import pytest

from blazingsql_amd._lib import BsqlError
from tests.helpers import (FLOAT64, INT32, INT64, TS_D, ExprGen, assert_same, device_columns,
                           host_columns, make_table)

pytestmark = pytest.mark.gpu

TYPES = [INT32, INT64, FLOAT64, FLOAT64, TS_D, INT32]


def _make_batches(bsql, seed, n_batches, max_rows, wide=0):
    """n_batches batches of up to max_rows rows; batches 1..wide carry
    programs that need more than 8 registers."""
    from blazingsql_amd.interop import DeviceColumn
    rng = np.random.default_rng(seed)
    batches = []
    for b in range(n_batches):
        rows = int(np.exp(rng.uniform(np.log(1), np.log(max_rows))))
        if b == 0:
            rows = 0  # an empty batch
        rates = list(rng.uniform(0, 0.3, len(TYPES)))
        table = make_table(rng, TYPES, rows, null_rates=rates)
        gen = ExprGen(rng, TYPES)
        while True:
            if 1 <= b <= wide:
                exprs = [gen.numeric(5) for _ in range(3)]
            else:
                exprs = [gen.numeric(3) if rng.random() < 0.6 else gen.boolean(3)
                         for _ in range(int(rng.integers(1, 4)))]
            try:
                prog = bsql.Program.compile(exprs, TYPES)
                if 1 <= b <= wide and prog.plan_info(TYPES)["n_regs"] <= 8:
                    continue
                break
            except BsqlError:
                # texts the reference's front end rejects too (an OR of three
                # terms next to FLOOR is not expanded, CalciteExpressionParsing.cpp:341-344)
                continue
        ot = [prog.output_type(TYPES, k) for k in range(prog.num_outputs)]
        outs = [DeviceColumn.empty(t, rows) for t in ot]
        batches.append((prog, table, device_columns(table), outs, rows, ot))
    return batches


def _check(batches, what):
    import oracle
    for i, (prog, table, _, outs, rows, ot) in enumerate(batches):
        ores = oracle.evaluate(prog, host_columns(table), rows)
        for k in range(len(ot)):
            gv, gval = outs[k].to_numpy(rows)
            assert_same(gv, gval, ores[k][0], ores[k][1], what=f"{what} batch {i} output {k}")


def _clear(batches):
    for _, _, _, outs, _, _ in batches:
        for o in outs:
            o.data.fill_(0x55)
            if o.validity is not None:
                o.validity.zero_()


@pytest.mark.parametrize("rpt", [0, 1, 2, 4])
def test_batches_vs_oracle(bsql, gpu, rpt):
    import torch
    from blazingsql_amd.interop import Batch
    batches = _make_batches(bsql, 100 + rpt, 40, 200_000)
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    if rpt:
        batch.set_launch(rpt)
    assert batch.total_rows == sum(b[4] for b in batches)
    batch.launch()
    torch.cuda.synchronize()
    _check(batches, f"rpt={rpt}")


@pytest.mark.parametrize("rpt,block", [(0, 0), (1, 256), (2, 256), (1, 64), (2, 64), (4, 64)])
def test_batches_lds_registers_vs_oracle(bsql, gpu, rpt, block):
    """The register file in LDS: every launch shape against the oracle, then
    a second launch over cleared outputs (the file carries stale values from
    the last tile, which every program overwrites before reading)."""
    import torch
    from blazingsql_amd.interop import Batch
    batches = _make_batches(bsql, 200 + rpt + block, 40, 200_000)
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    batch.set_launch(rpt, block_size=block, lds_regs=True)
    for rep in range(2):
        _clear(batches)
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        _check(batches, f"lds registers rpt={rpt} block={block} launch {rep}")


@pytest.mark.parametrize("lds", [False, True])
def test_batches_in_hip_graph(bsql, gpu, lds):
    """The launch captured in a graph, replayed three times with eager work
    and a synchronization between replays; every replay matches the oracle
    (register file in VGPRs, or in LDS with the 8-register form)."""
    import torch
    from blazingsql_amd.interop import Batch
    batches = _make_batches(bsql, 7, 24, 100_000, wide=2 if lds else 0)
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    if lds:
        batch.set_launch(4, block_size=64, lds_regs=True, spill_wide=True)
    batch.launch()  # warm-up outside capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            batch.launch(stream=s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        _clear(batches)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check(batches, f"replay {rep}")


def test_batch_autotune(bsql, gpu):
    import torch
    from blazingsql_amd.interop import Batch
    batches = _make_batches(bsql, 3, 16, 300_000)
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    r = batch.autotune()
    assert r in (1, 2, 4)
    assert batch.chosen[0] in (64, 256) and batch.chosen[1] == r
    _clear(batches)
    batch.launch()
    torch.cuda.synchronize()
    _check(batches, f"autotuned {batch.chosen}")


def test_program_autotune(bsql, gpu):
    import torch
    from blazingsql_amd.interop import DeviceColumn, autotune_program, perform_interpreter_operation
    import oracle
    rng = np.random.default_rng(1)
    n = 300_000
    table = make_table(rng, TYPES, n, null_rates=[0.1] * len(TYPES))
    prog = bsql.Program.compile(["+(*($1, 3), $0)", "CASE(>($2, 0), $3, $2)"], TYPES)
    ot = [prog.output_type(TYPES, k) for k in range(2)]
    outs = [DeviceColumn.empty(t, n) for t in ot]
    r = autotune_program(prog, device_columns(table), outs, n)
    assert r in (1, 2, 4)
    perform_interpreter_operation(prog, device_columns(table), outs, n)
    torch.cuda.synchronize()
    ores = oracle.evaluate(prog, host_columns(table), n)
    for k in range(2):
        gv, gval = outs[k].to_numpy(n)
        assert_same(gv, gval, ores[k][0], ores[k][1], what=f"output {k}")


@pytest.mark.parametrize("lds", [False, True])
def test_batches_spill_wide_vs_oracle(bsql, gpu, lds):
    """Every program lowered for the 8-register file (values beyond 7 in LDS
    spill slots), with the register file in VGPRs or in LDS, against the
    oracle; the batch mixes programs that need up to 16 registers."""
    import torch
    from blazingsql_amd.interop import Batch
    batches = _make_batches(bsql, 300 + int(lds), 40, 200_000, wide=3)
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    for rpt, block in [(4, 256), (2, 256), (4, 64), (2, 64)]:
        batch.set_launch(rpt, block_size=block, lds_regs=lds, spill_wide=True)
        _clear(batches)
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        _check(batches, f"spill_wide lds={lds} rpt={rpt} block={block}")


def test_c3_shape_vs_oracle(bsql, gpu):
    """The batched workload's shape: programs of 8-24 ops (CASE, COALESCE,
    CAST, EXTRACT, arithmetic, comparisons), batches of 4K-4M rows with null
    rates up to 30 %, autotuned launch, every batch against the oracle."""
    import torch
    from blazingsql_amd.interop import Batch, DeviceColumn
    rng = np.random.default_rng(2024)
    batches = []
    sizes = [4096, 4194304] + [int(np.exp(rng.uniform(np.log(4096), np.log(4194304)))) for _ in range(10)]
    for b, rows in enumerate(sizes):
        table = make_table(rng, TYPES, rows, null_rates=list(rng.uniform(0, 0.3, len(TYPES))))
        gen = ExprGen(rng, TYPES)
        while True:
            exprs = [gen.numeric(4) if rng.random() < 0.7 else gen.boolean(3)
                     for _ in range(int(rng.integers(1, 4)))]
            try:
                prog = bsql.Program.compile(exprs, TYPES)
            except BsqlError:
                continue
            if 8 <= prog.num_ops <= 24:
                break
        ot = [prog.output_type(TYPES, k) for k in range(prog.num_outputs)]
        outs = [DeviceColumn.empty(t, rows) for t in ot]
        batches.append((prog, table, device_columns(table), outs, rows, ot))
    batch = Batch([(p, dev, outs, rows) for p, _, dev, outs, rows, _ in batches])
    batch.autotune()
    _clear(batches)
    torch.cuda.synchronize()
    batch.launch()
    torch.cuda.synchronize()
    _check(batches, f"C3 shape, autotuned {batch.chosen}")
