"""plp_engine_sync_fetch: the result copy overlapped with the decode.

Each decoder publishes a completion word after its row and packing, and the
host copies its results as soon as the word is set.  Every case must give
exactly what plp_engine_sync followed by plp_engine_fetch gives, and the
digests tests/golden/fullsize_digests.json pins (the C oracle and the
reference's g++ build agree on them): the overlapped path itself (one
outstanding run, graph replays, both verdict kinds, shared prefixes, rows
without placements, the block engine), and the cases that copy after the
sync instead (two runs enqueued, decoders rerun at a larger capacity).
"""
import json

import numpy as np
import pytest

from conftest import golden_path

pytestmark = pytest.mark.gpu

EPS, THETA = 0.01, 0.9
with open(golden_path("fullsize_digests.json")) as _f:
    PINNED = json.load(_f)["workloads"]


@pytest.fixture(scope="module")
def plp():
    import plp_amd
    assert plp_amd.device_count() >= 1, "no HIP device"
    return plp_amd


@pytest.fixture(scope="module")
def ctx(plp):
    c = plp.Ctx(0)
    yield c
    c.close()


def _load(plp, ctx, name, n=None, blocks=False, **cfg):
    from plp_amd import synth
    texts, dec_inst, seeds, passes = synth.workload(name)
    if n is not None:
        dec_inst, seeds = dec_inst[:n], seeds[:n]
    eng = plp.Engine(ctx, **cfg)
    eng.load([plp.Instance.from_text(t) for t in texts], dec_inst, seeds,
             plp.params(eps=EPS, theta=THETA), blocks=plp.block_params() if blocks else None)
    return eng, passes


def _out(plp, eng):
    return (plp.RowResult * eng.n_dec)(), np.full((eng.n_dec, eng.max_items(), 7), -7, np.int32)


def _per(plp, rows, pl):
    return plp.results_digest(rows, pl)[1]


@pytest.mark.parametrize("verdicts", ["pair_cache", "point_state"])
def test_overlapped_copy_matches_pins_over_graph_replays(plp, ctx, verdicts):
    """The 256-instance batch (4096 decoders), five graph replays with eager
    work between them; the output arrays start filled with junk."""
    eng, passes = _load(plp, ctx, "mixed_256x16", use_graph=True, verdicts=verdicts)
    pin = PINNED["mixed_256x16"]["per_decoder"]
    for rep in range(5):
        rows, pl = _out(plp, eng)
        eng.reset()
        eng.run(passes)
        eng.sync_fetch(out=(rows, pl))
        assert all(r.err == 0 for r in rows)
        assert _per(plp, rows, pl) == pin, rep
        if rep == 2:
            _, _ = eng.fetch()   # a plain fetch between replays changes nothing
    eng.close()


def test_same_as_sync_then_fetch(plp, ctx):
    """Field for field, rows and the packing rows every decoder returns."""
    eng, passes = _load(plp, ctx, "twenty_1024x4", use_graph=True)
    eng.reset()
    eng.run(passes)
    eng.sync()
    rows_a, pl_a = eng.fetch()
    eng.reset()
    eng.run(passes)
    rows_b, pl_b = eng.sync_fetch()
    for a, b in zip(rows_a, rows_b):
        assert bytes(a) == bytes(b)
    for k, r in enumerate(rows_a):
        assert (pl_a[k, :r.n_boxes] == pl_b[k, :r.n_boxes]).all(), k
    assert _per(plp, rows_b, pl_b) == PINNED["twenty_1024x4"]["per_decoder"]
    eng.close()


def test_rows_without_placements(plp, ctx):
    eng, passes = _load(plp, ctx, "hundred_2048", n=512, use_graph=True)
    eng.reset()
    eng.run(passes)
    rows, pl = eng.sync_fetch(placements=False)
    assert pl is None
    eng.reset()
    eng.run(passes)
    eng.sync()
    want, _ = eng.fetch(placements=False)
    assert [bytes(r) for r in rows] == [bytes(r) for r in want]
    eng.close()


def test_shared_prefixes(plp, ctx):
    """Sharing on: the recording launch writes no completion words; the
    decoders' launch does."""
    eng, passes = _load(plp, ctx, "mixed_256x64", use_graph=True, share_prefix="on")
    for _ in range(2):
        rows, pl = _out(plp, eng)
        eng.reset()
        eng.run(passes)
        eng.sync_fetch(out=(rows, pl))
        assert _per(plp, rows, pl) == PINNED["mixed_256x64"]["per_decoder"]
    eng.close()


def test_two_runs_enqueued_then_run_after_sync(plp, ctx):
    """run(1) twice before the call (copied after the sync), then a fresh
    load decoded by sync(), run, and the overlapped call: the completion
    words of the synchronised run were cleared."""
    from plp_amd import synth
    texts, dec_inst, seeds, passes = synth.workload("twenty_1024x4")
    pin = PINNED["twenty_1024x4"]["per_decoder"]
    eng, _ = _load(plp, ctx, "twenty_1024x4")
    eng.reset()
    eng.run(2)
    eng.run(passes - 2)
    rows, pl = eng.sync_fetch()
    assert _per(plp, rows, pl) == pin
    eng.reset()
    eng.run(1)
    eng.sync()
    eng.run(passes - 1)
    rows, pl = eng.sync_fetch()
    assert _per(plp, rows, pl) == pin
    eng.close()


def test_decoders_rerun_at_a_larger_capacity(plp, ctx):
    """Room for 8 boxes and 16 points: decoders stop, are rerun inside the
    sync, and every result is copied again."""
    eng, passes = _load(plp, ctx, "mixed_256x16", n=512, box_cap=8, ep_cap=16, use_graph=True)
    eng.reset()
    eng.run(passes)
    rows, pl = eng.sync_fetch()
    assert all(r.err == 0 for r in rows)
    assert _per(plp, rows, pl) == PINNED["mixed_256x16"]["per_decoder"][:512]
    eng.close()


def test_block_engine(plp, ctx):
    eng, passes = _load(plp, ctx, "blocks_twenty_1024x4", blocks=True, use_graph=True)
    for _ in range(3):
        rows, pl = _out(plp, eng)
        eng.reset()
        eng.run(passes)
        eng.sync_fetch(out=(rows, pl))
        assert _per(plp, rows, pl) == PINNED["blocks_twenty_1024x4"]["per_decoder"]
    eng.close()
