"""The device engines at the benchmark's full sizes, checked through properties
that hold at any size, plus bit-exact oracle spot checks.

Every returned packing (best accepted, or best of all on fallback) must:
  * lie in the container and never overlap (C2 and pairwise C1);
  * give every box above the floor at least half its footprint in support
    (C3, the integer form 2 * support >= w * d; for the block engine C3 holds
    per block, so this check is left to the oracle comparison there);
  * use only orientations its type allows and at most qty copies per type;
  * report the volume and box count of its rows;
  * carry the stability score the FSR kernel computes for its rows, bit for
    bit, with fallback set exactly when that score is below theta.
A few decoders of each workload are also decoded by the oracle and must match
field for field.  Workloads: the 20-type instance with 1024 decoders x 4
passes, 256 uneven instances x 16 seeds in one launch, the 100-type instance
(600 items, about 100 placed) with 2048 decoders, and the block-building engine on the
20-type instance with 1024 decoders x 4 passes.
"""
import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

EPS, THETA = 0.01, 0.9
FIELDS = ("n_boxes", "volume_cm3", "stability_score", "fallback", "n_passes", "seed",
          "placements")


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


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _orientations(l, lf, w, wf, h, hf):
    """Allowed (w, d, h) triples of a type (plp_common.hpp:269-280)."""
    out = []
    for keep, cand in ((lf, ((w, h, l), (h, w, l))), (wf, ((l, h, w), (h, l, w))),
                       (hf, ((l, w, h), (w, l, h)))):
        if keep:
            for c in cand:
                if c not in out:
                    out.append(c)
    return out


def _check_packing(inst, row, pl, theta, per_box_support=True):
    L, W, H, types = inst
    n = row.n_boxes
    p = np.asarray(pl[:n], np.int64)
    assert n > 0
    t, x, y, z, w, d, h = (p[:, i] for i in range(7))
    assert ((x >= 0) & (y >= 0) & (z >= 0)).all()
    assert ((x + w <= L) & (y + d <= W) & (z + h <= H)).all()
    assert int((w * d * h).sum()) == row.volume
    allowed = [set(_orientations(*ty[:6])) for ty in types]
    for ti, wi, di, hi in zip(t, w, d, h):
        assert (int(wi), int(di), int(hi)) in allowed[int(ti)]
    counts = np.bincount(t, minlength=len(types))
    assert all(counts[i] <= types[i][6] for i in range(len(types)))
    x2, y2, z2 = x + w, y + d, z + h
    ox = np.minimum(x2[:, None], x2[None, :]) - np.maximum(x[:, None], x[None, :])
    oy = np.minimum(y2[:, None], y2[None, :]) - np.maximum(y[:, None], y[None, :])
    oz = np.minimum(z2[:, None], z2[None, :]) - np.maximum(z[:, None], z[None, :])
    ov = (ox > 0) & (oy > 0) & (oz > 0)
    np.fill_diagonal(ov, False)
    assert not ov.any()
    if per_box_support:
        sup = ((z2[None, :] == z[:, None]) * np.maximum(ox, 0) * np.maximum(oy, 0)).sum(axis=1)
        assert ((z == 0) | (2 * sup >= w * d)).all()
    assert row.fallback == (row.score < theta)


def _check_all(plp, ctx, insts, dec_inst, rows, pl, theta, per_box_support=True):
    packs = [pl[k][:rows[k].n_boxes] for k in range(len(rows))]
    _, scores = ctx.fsr_eval(packs)
    for k, r in enumerate(rows):
        assert r.err == 0, k
        assert scores[k] == r.score, k
        _check_packing(insts[dec_inst[k]], r, pl[k], theta, per_box_support)


def _run(plp, ctx, texts, dec_inst, seeds, passes, blocks=None):
    eng = plp.Engine(ctx, use_graph=True)
    eng.load([plp.Instance.from_text(tx) for tx in texts], dec_inst, seeds,
             plp.params(eps=EPS, theta=THETA), blocks=blocks)
    eng.run(passes)
    eng.sync()
    rows, pl = eng.fetch()
    eng.close()
    return rows, pl


def test_twenty_types_1024_decoders_4_passes(plp, ctx, orc):
    from plp_amd import synth
    text = synth.multi_instance()
    inst = orc.parse(text)[0]
    K, P = 1024, 4
    rows, pl = _run(plp, ctx, [text], [0] * K, list(range(1, K + 1)), P)
    _check_all(plp, ctx, [inst], [0] * K, rows, pl, THETA)
    for k in (0, 511, 1023):
        want = orc.run(inst, seed=k + 1, passes=P, eps=EPS, theta=THETA)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k


def test_256_instances_16_seeds(plp, ctx, orc):
    from plp_amd import synth
    texts = synth.mixed_batch(256)
    insts = [orc.parse(tx)[0] for tx in texts]
    dec_inst = [i for i in range(256) for _ in range(16)]
    seeds = [1 + s for _ in range(256) for s in range(16)]
    rows, pl = _run(plp, ctx, texts, dec_inst, seeds, 1)
    _check_all(plp, ctx, insts, dec_inst, rows, pl, THETA)
    for k in (0, 17, 2000, 4095):
        want = orc.run(insts[dec_inst[k]], seed=seeds[k], passes=1, eps=EPS, theta=THETA)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k


def test_hundred_types_2048_decoders(plp, ctx, orc):
    from plp_amd import synth
    text = synth.large_instance()
    inst = orc.parse(text)[0]
    K = 2048
    rows, pl = _run(plp, ctx, [text], [0] * K, list(range(1, K + 1)), 1)
    assert max(r.n_boxes for r in rows) > 64
    _check_all(plp, ctx, [inst], [0] * K, rows, pl, THETA)
    for k in (0, 2047):
        want = orc.run(inst, seed=k + 1, passes=1, eps=EPS, theta=THETA)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k


def test_block_engine_1024_decoders_4_passes(plp, ctx, orc):
    from plp_amd import synth
    text = synth.multi_instance()
    inst = orc.parse(text)[0]
    K, P = 1024, 4
    rows, pl = _run(plp, ctx, [text], [0] * K, list(range(1, K + 1)), P,
                    blocks=plp.block_params())
    # C3 holds for each block's footprint (blockproof/plp_block.cpp tests the
    # block shape), not for every item of a block's bottom layer
    _check_all(plp, ctx, [inst], [0] * K, rows, pl, THETA, per_box_support=False)
    for k in (0, 1023):
        want = orc.block_run(inst, seed=k + 1, passes=P, eps=EPS, theta=THETA)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k
