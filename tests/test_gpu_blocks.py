"""The gfx950 block-building decoder (block_kernel.hip) against the reference
block engine (blockproof/plp_block.cpp): every fixture run of
tests/golden/block_runs.json, and many decoders against the C oracle."""
import json
from collections import defaultdict

import pytest

from conftest import golden_path
from edge_cases import EDGE_TEXTS
from oracle_lib import Oracle, load_text

pytestmark = pytest.mark.gpu

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


def _params(plp, a):
    return (plp.params(eps=a["eps"], theta=0.9 if a["theta"] is None else a["theta"],
                       use_theta=a["theta"] is not None),
            plp.block_params(a["amax"], a["bmax"], a["kmax"], a["topk"], a["pick"]))


# one wave per decoder (the autotuned shape: short lists test several plan
# entries in one pass) and four (one test at a time)
WAVES = [1, 4]


@pytest.mark.parametrize("wpd", WAVES)
def test_every_reference_block_run(plp, ctx, wpd):
    """One engine per parameter set, one decoder per fixture case."""
    with open(golden_path("block_runs.json")) as f:
        cases = json.load(f)
    groups = defaultdict(list)
    for c in cases:
        a = c["args"]
        groups[(a["passes"], a["eps"], a["theta"], a["amax"], a["bmax"], a["kmax"], a["topk"],
                a["pick"])].append(c)
    for key, cs in groups.items():
        prm, bp = _params(plp, cs[0]["args"])
        insts = [plp.Instance.from_text(load_text(c["instance"])) for c in cs]
        eng = plp.Engine(ctx, waves_per_decoder=wpd)
        eng.load(insts, list(range(len(cs))), [c["args"]["seed"] for c in cs], prm, blocks=bp)
        eng.run(key[0])
        eng.sync()
        rows, pl = eng.fetch()
        for i, c in enumerate(cs):
            assert rows[i].err == 0, c["case"]
            got = plp.to_json_dict(rows[i], pl[i])
            assert {k: got[k] for k in FIELDS} == {k: c["result"][k] for k in FIELDS}, c["case"]
        eng.close()


@pytest.mark.parametrize("wpd", [1, 2, 4])
def test_many_block_decoders_vs_oracle(plp, ctx, orc, wpd):
    """128 decoders of the 20-type instance, eps 0.3, 2 passes, first pick."""
    text = load_text("synth/multi.txt")
    o_inst = orc.parse(text)[0]
    K = 128
    eng = plp.Engine(ctx, waves_per_decoder=wpd)
    eng.load([plp.Instance.from_text(text)], [0] * K, list(range(1, K + 1)),
             plp.params(eps=0.3, theta=0.9), blocks=plp.block_params())
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    for k in range(K):
        want = orc.block_run(o_inst, seed=k + 1, passes=2, eps=0.3, theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert {x: got[x] for x in FIELDS} == {x: want[x] for x in FIELDS}, k


@pytest.mark.parametrize("wpd", WAVES)
@pytest.mark.parametrize("pick", ["support", "contact"])
def test_contact_pick_rules_vs_oracle(plp, ctx, orc, pick, wpd):
    text = load_text("data/instances/BR2_050.txt")
    o_inst = orc.parse(text)[0]
    K = 48
    eng = plp.Engine(ctx, waves_per_decoder=wpd)
    eng.load([plp.Instance.from_text(text)], [0] * K, list(range(100, 100 + K)),
             plp.params(eps=0.1, theta=0.9), blocks=plp.block_params(pick=pick, topk=4))
    eng.run(1)
    eng.sync()
    rows, pl = eng.fetch()
    for k in range(K):
        want = orc.block_run(o_inst, seed=100 + k, passes=1, eps=0.1, theta=0.9, pick=pick, topk=4)
        got = plp.to_json_dict(rows[k], pl[k])
        assert {x: got[x] for x in FIELDS} == {x: want[x] for x in FIELDS}, k


@pytest.mark.parametrize("wpd", WAVES)
def test_mixed_block_batch_vs_oracle(plp, ctx, orc, wpd):
    """Uneven instances (3..100 types) x 3 seeds in one launch."""
    from plp_amd import synth
    texts = synth.mixed_batch(12, seed0=515_000)
    insts = [plp.Instance.from_text(t) for t in texts]
    di = [i for i in range(len(texts)) for _ in range(3)]
    ds = [7 + s for _ in range(len(texts)) for s in range(3)]
    eng = plp.Engine(ctx, waves_per_decoder=wpd)
    eng.load(insts, di, ds, plp.params(eps=0.2, theta=0.9), blocks=plp.block_params(kmax=3))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    for k, (i, s) in enumerate(zip(di, ds)):
        want = orc.block_run(orc.parse(texts[i])[0], seed=s, passes=2, eps=0.2, theta=0.9, kmax=3)
        got = plp.to_json_dict(rows[k], pl[k])
        assert {x: got[x] for x in FIELDS} == {x: want[x] for x in FIELDS}, k


@pytest.mark.parametrize("wpd", WAVES)
@pytest.mark.parametrize("eps", [0.0, 1.0])
def test_block_edge_instances_vs_oracle(plp, ctx, orc, eps, wpd):
    """The decoder's degenerate instances (nothing fits, one item, types
    without an allowed orientation, an exact tiling) through the block engine,
    seeds 0, 1 and 2^64 - 1; eps 1.0 shuffles every plan and picks every shape
    at random."""
    names = list(EDGE_TEXTS)
    seeds = [0, 1, 2**64 - 1]
    di = [i for i in range(len(names)) for _ in seeds]
    ds = [s for _ in names for s in seeds]
    eng = plp.Engine(ctx, waves_per_decoder=wpd)
    eng.load([plp.Instance.from_text(EDGE_TEXTS[n]) for n in names], di, ds,
             plp.params(eps=eps, theta=0.9), blocks=plp.block_params())
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    for k, (i, s) in enumerate(zip(di, ds)):
        assert rows[k].err == 0, (names[i], s)
        want = orc.block_run(orc.parse(EDGE_TEXTS[names[i]])[0], seed=s, passes=2, eps=eps,
                             theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert {x: got[x] for x in FIELDS} == {x: want[x] for x in FIELDS}, (names[i], s)


def test_block_capacity_escalation(plp, ctx, orc):
    text = load_text("data/instances/BR1_050.txt")
    eng = plp.Engine(ctx, ep_cap=16)
    eng.load([plp.Instance.from_text(text)], [0] * 4, [1, 2, 3, 4], plp.params(eps=0.3, theta=0.9),
             blocks=plp.block_params())
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    assert eng.describe()["ep_cap"] >= max(r.ep_need for r in rows) > 16
    # waves per decoder left automatic: one for the block engine
    assert eng.describe()["waves_per_decoder"] == 1
    for k in range(4):
        want = orc.block_run(orc.parse(text)[0], seed=k + 1, passes=2, eps=0.3, theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert {x: got[x] for x in FIELDS} == {x: want[x] for x in FIELDS}, k


@pytest.mark.parametrize("box_cap,wpd", [(8, 4), (24, 4), (64, 4), (24, 1)])
def test_block_item_capacity_rerun(plp, ctx, box_cap, wpd):
    """Room for fewer items than the placed blocks hold: decoders stop when a
    block would pass it and are rerun with room for every item; every result
    equals the pinned digests of the 20-type instance x 1024 x 4 passes, run
    after run, and the autotuner's fitted capacity gives them too."""
    import json
    from conftest import golden_path
    from plp_amd import synth
    with open(golden_path("fullsize_digests.json")) as f:
        pin = json.load(f)["workloads"]["blocks_twenty_1024x4"]["per_decoder"]
    texts, dec_inst, seeds, passes = synth.workload("blocks_twenty_1024x4")
    eng = plp.Engine(ctx, box_cap=box_cap, use_graph=True, waves_per_decoder=wpd)
    eng.load([plp.Instance.from_text(t) for t in texts], dec_inst, seeds,
             plp.params(eps=0.01, theta=0.9), blocks=plp.block_params())
    for _ in range(2):
        eng.reset()
        eng.run(passes)
        rows, pl = eng.sync_fetch()
        assert all(r.err == 0 for r in rows)
        assert plp.results_digest(rows, pl)[1] == pin
    tuned, _ = eng.autotune(passes)
    assert 0 < tuned["box_cap"] <= eng.max_items()
    eng.reset()
    eng.run(passes)
    rows, pl = eng.sync_fetch()
    assert plp.results_digest(rows, pl)[1] == pin
    eng.close()


BLOCK_CLI = __import__("os").path.join(__import__("conftest").REPO, "plp-gpu_amd", "build",
                                       "plp_block")


def test_block_cli_matches_reference_cli_runs():
    """`plp_block ... --single-pass --json` prints the reference CLI's JSON
    (time_s aside) for every fixture case the reference CLI ran."""
    import subprocess
    with open(golden_path("block_runs.json")) as f:
        cases = [c for c in json.load(f) if c["reference_cli"] is not None][:12]
    for c in cases:
        a = c["args"]
        args = [BLOCK_CLI, golden_path(c["instance"]), "--single-pass", "--json", "--seed",
                str(a["seed"]), "--eps", repr(a["eps"]), "--pick", a["pick"], "--amax",
                str(a["amax"]), "--bmax", str(a["bmax"]), "--kmax", str(a["kmax"]), "--topk",
                str(a["topk"])] + (["--no-theta"] if a["theta"] is None else
                                   ["--theta", repr(a["theta"])])
        p = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        got = json.loads(p.stdout)
        want = dict(c["reference_cli"])
        got.pop("time_s")
        want.pop("time_s")
        assert got == want, c["case"]
        assert p.stderr.startswith("fill_pct=")
