"""Shared prefixes (plp_engine_config.share_prefix = 2) against the oracle,
the reference's fixtures and the pinned full-size digests.

With sharing on, a run first decodes each instance's trajectory (feas[0] at
every placing step) once; every decoder then replays its own RNG along it and
resumes from the list and boxes of the placing step where its first eps pick
with an index other than 0 leaves it, or takes the trajectory's packing when
none does.  Results must be bit-identical to decoding every decoder from the
origin, for both verdict kinds, every wave count, multi-pass runs split across
calls, capacity stops on the shared part (extreme points and boxes, then the
engine's exact reruns), degenerate instances and graph replays.
"""
import json

import pytest

from conftest import golden_path
from edge_cases import EDGE_TEXTS
from oracle_lib import Oracle, load_text

pytestmark = pytest.mark.gpu

VERDICTS = ("pair_cache", "point_state")
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


def prm(plp, a):
    return plp.params(eps=a["eps"], theta=0.9 if a["theta"] is None else a["theta"], mu=a["mu"],
                      use_theta=a["theta"] is not None, first_fit=a["first_fit"], order=a["order"])


def fields(res):
    return {k: res[k] for k in FIELDS}


def _decode(plp, ctx, insts, dec_inst, seeds, params, passes, **cfg):
    eng = plp.Engine(ctx, share_prefix="on", **cfg)
    eng.load(insts, dec_inst, seeds, params)
    eng.run(passes)
    eng.sync()
    rows, pl = eng.fetch()
    eng.close()
    return rows, pl


@pytest.mark.parametrize("verdicts", VERDICTS)
def test_every_reference_case_shared(plp, ctx, orc, ref_runs, verdicts):
    """Each reference run (the reference's g++ build: eps 0 / 0.01 / 0.3 /
    0.5 / 1.0, 1-6 passes, theta on and off, first fit, input order, 64-bit
    seeds) with 7 more seeds of its instance in the batch, so that the
    trajectory is shared: the run's decoder against the reference's output,
    the others against the oracle."""
    by_params = {}
    for c in ref_runs:
        a = c["args"]
        key = (a["passes"], a["eps"], a["theta"], a["mu"], a["order"], a["first_fit"])
        by_params.setdefault(key, []).append(c)
    for key, cases in by_params.items():
        insts = [plp.Instance.from_text(load_text(c["instance"])) for c in cases]
        dec_inst, seeds = [], []
        for i, c in enumerate(cases):
            s0 = c["args"]["seed"]
            dec_inst += [i] * 8
            seeds += [s0] + [(s0 + 7919 * k) % 2**64 for k in range(1, 8)]
        rows, pl = _decode(plp, ctx, insts, dec_inst, seeds, prm(plp, cases[0]["args"]), key[0],
                           verdicts=verdicts)
        for i, c in enumerate(cases):
            a = c["args"]
            assert rows[8 * i].err == 0, c["case"]
            assert plp.to_json_dict(rows[8 * i], pl[8 * i]) == fields(c["result"]), c["case"]
            o_inst = orc.parse(load_text(c["instance"]))[0]
            for k in range(1, 8):
                want = orc.run(o_inst, seed=seeds[8 * i + k], passes=a["passes"], eps=a["eps"],
                               theta=a["theta"], order=a["order"], first_fit=a["first_fit"],
                               mu=a["mu"])
                assert plp.to_json_dict(rows[8 * i + k], pl[8 * i + k]) == fields(want), \
                    (c["case"], k)


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("wpd", [1, 2, 4, 8])
def test_many_seeds_vs_oracle(plp, ctx, orc, wpd, verdicts):
    """The 20-type instance, 256 seeds x 3 passes at eps 0.01 (most decoders
    leave the trajectory late or never), every decoder against the oracle."""
    text = load_text("synth/multi.txt")
    o_inst = orc.parse(text)[0]
    K = 256
    rows, pl = _decode(plp, ctx, [plp.Instance.from_text(text)], [0] * K,
                       list(range(1, K + 1)), plp.params(eps=0.01, theta=0.9), 3,
                       waves_per_decoder=wpd, verdicts=verdicts)
    for k in range(K):
        want = orc.run(o_inst, seed=k + 1, passes=3, eps=0.01, theta=0.9)
        assert plp.to_json_dict(rows[k], pl[k]) == fields(want), k


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("eps", [0.3, 1.0])
def test_uneven_batch_early_divergence_vs_oracle(plp, ctx, orc, eps, verdicts):
    """Uneven instances (3..100 types) x 6 seeds, eps 0.3 and 1.0: decoders
    leave their trajectories within the first steps, or (eps 1.0, a random
    pick at every placing step) at the first pick with an index other than 0."""
    from plp_amd import synth
    texts = synth.mixed_batch(12, seed0=424_242)
    insts = [plp.Instance.from_text(t) for t in texts]
    dec_inst = [i for i in range(len(texts)) for _ in range(6)]
    seeds = [1000 + 7 * i + s for i in range(len(texts)) for s in range(6)]
    rows, pl = _decode(plp, ctx, insts, dec_inst, seeds, plp.params(eps=eps, theta=0.9), 2,
                       verdicts=verdicts)
    for d, (i, s) in enumerate(zip(dec_inst, seeds)):
        want = orc.run(orc.parse(texts[i])[0], seed=s, passes=2, eps=eps, theta=0.9)
        assert plp.to_json_dict(rows[d], pl[d]) == fields(want), d


@pytest.mark.parametrize("verdicts", VERDICTS)
def test_no_rng_every_decoder_is_the_trajectory(plp, ctx, orc, verdicts):
    """eps 0 and first fit draw nothing: every decoder's passes are its
    instance's trajectory."""
    text = load_text("synth/multi.txt")
    o_inst = orc.parse(text)[0]
    for params, kw in ((plp.params(eps=0.0, use_theta=False), dict(eps=0.0, theta=None)),
                       (plp.params(eps=0.3, theta=0.9, first_fit=True),
                        dict(eps=0.3, theta=0.9, first_fit=True))):
        rows, pl = _decode(plp, ctx, [plp.Instance.from_text(text)], [0] * 16, list(range(16)),
                           params, 2, verdicts=verdicts)
        for k in range(16):
            want = orc.run(o_inst, seed=k, passes=2, **kw)
            assert plp.to_json_dict(rows[k], pl[k]) == fields(want), k


@pytest.mark.parametrize("verdicts", VERDICTS)
def test_passes_in_pieces_and_graph_replays(plp, ctx, orc, verdicts):
    """RNG and D4' state carried across runs (run(1) x 3 == run(3)) and
    graph replays with eager work between them, all shared."""
    text = load_text("synth/multi.txt")
    o_inst = orc.parse(text)[0]
    inst = plp.Instance.from_text(text)
    seeds = list(range(40, 104))
    want = [fields(orc.run(o_inst, seed=s, passes=3, eps=0.01, theta=0.9)) for s in seeds]
    for use_graph in (False, True):
        eng = plp.Engine(ctx, share_prefix="on", use_graph=use_graph, verdicts=verdicts)
        eng.load([inst], [0] * len(seeds), seeds, plp.params(eps=0.01, theta=0.9))
        for rep in range(2):
            eng.reset()
            for _ in range(3):
                eng.run(1)
                eng.sync()
            rows, pl = eng.fetch()
            assert [plp.to_json_dict(r, p) for r, p in zip(rows, pl)] == want, (use_graph, rep)
        eng.close()


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("box_cap,ep_cap", [(8, 0), (16, 16), (48, 64)])
def test_capacity_stops_on_the_shared_part(plp, ctx, box_cap, ep_cap, verdicts):
    """Capacities below what the trajectories reach: a decoder resuming past
    its box capacity, or whose shared part held a longer list than its
    capacity, stops as its own decode would and is rerun exactly; every
    result equals the pinned digests (first 512 decoders of the uneven
    256-instance batch)."""
    from plp_amd import synth
    texts, dec_inst, seeds, passes = synth.workload("mixed_256x16")
    with open(golden_path("fullsize_digests.json")) as f:
        pin = json.load(f)["workloads"]["mixed_256x16"]["per_decoder"][:512]
    insts = [plp.Instance.from_text(t) for t in texts[:32]]
    eng = plp.Engine(ctx, box_cap=box_cap, ep_cap=ep_cap, use_graph=True, verdicts=verdicts,
                     share_prefix="on")
    eng.load(insts, dec_inst[:512], seeds[:512], plp.params(eps=0.01, theta=0.9))
    for _ in range(2):
        eng.reset()
        eng.run(passes)
        eng.sync()
        rows, pl = eng.fetch()
        assert all(r.err == 0 for r in rows)
        assert plp.results_digest(rows, pl)[1] == pin
    eng.close()


@pytest.mark.parametrize("eps", [0.01, 1.0])
def test_edge_instances_shared(plp, ctx, orc, eps):
    """Degenerate instances (nothing fits, one item, types without an allowed
    orientation, an exact tiling) with seeds 0, 1 and 2^64 - 1."""
    names = list(EDGE_TEXTS)
    seeds = [0, 1, 2**64 - 1]
    dec_inst = [i for i in range(len(names)) for _ in seeds]
    dec_seed = [s for _ in names for s in seeds]
    for verdicts in VERDICTS:
        rows, pl = _decode(plp, ctx, [plp.Instance.from_text(EDGE_TEXTS[n]) for n in names],
                           dec_inst, dec_seed, plp.params(eps=eps, theta=0.9), 2,
                           verdicts=verdicts)
        for d, (i, s) in enumerate(zip(dec_inst, dec_seed)):
            want = orc.run(orc.parse(EDGE_TEXTS[names[i]])[0], seed=s, passes=2, eps=eps,
                           theta=0.9)
            assert plp.to_json_dict(rows[d], pl[d]) == fields(want), (names[i], s, verdicts)


@pytest.mark.parametrize("name", ["twenty_1024x4", "mixed_256x16", "hundred_2048",
                                  "hundred_fit_2048", "mixed_256x64"])
def test_full_size_workloads_shared_match_pins(plp, ctx, name):
    """Every decoder of the benchmark workloads, shared, for both verdict
    kinds, against the digests the C oracle and the reference's g++ build
    agree on."""
    from plp_amd import synth
    with open(golden_path("fullsize_digests.json")) as f:
        pin = json.load(f)["workloads"][name]
    texts, dec_inst, seeds, passes = synth.workload(name)
    insts = [plp.Instance.from_text(t) for t in texts]
    for verdicts in VERDICTS:
        rows, pl = _decode(plp, ctx, insts, dec_inst, seeds, plp.params(eps=0.01, theta=0.9),
                           passes, verdicts=verdicts, use_graph=True)
        assert all(r.err == 0 for r in rows)
        digest, per = plp.results_digest(rows, pl)
        bad = [k for k in range(len(rows)) if per[k] != pin["per_decoder"][k]]
        assert not bad, f"{verdicts}: {len(bad)} decoders differ, first {bad[:8]}"
        assert digest == pin["digest"]


@pytest.mark.parametrize("trace_waves,trace_verdicts", [(4, "pair_cache"), (8, "point_state"),
                                                        (8, "pair_cache")])
def test_recording_launch_shapes(plp, ctx, trace_waves, trace_verdicts):
    """The recording launch's own waves and verdict kind (trace_waves,
    trace_verdicts) against 2-wave point-state decoders: the first 512
    decoders of the uneven 256-instance batch equal their pinned digests."""
    from plp_amd import synth
    texts, dec_inst, seeds, passes = synth.workload("mixed_256x16")
    with open(golden_path("fullsize_digests.json")) as f:
        pin = json.load(f)["workloads"]["mixed_256x16"]["per_decoder"][:512]
    rows, pl = _decode(plp, ctx, [plp.Instance.from_text(t) for t in texts[:32]],
                       dec_inst[:512], seeds[:512], plp.params(eps=0.01, theta=0.9), passes,
                       waves_per_decoder=2, verdicts="point_state", use_graph=True,
                       trace_waves=trace_waves, trace_verdicts=trace_verdicts)
    assert all(r.err == 0 for r in rows)
    assert plp.results_digest(rows, pl)[1] == pin
