"""HIP path vs the oracle and the reference's fixtures, through the C ABI.

Bit-exact throughout: placements, volumes, n_boxes, fallback and every FP64
value (stability scores, per-box FSR, random()) are compared as exact values.
"""
import json
import random
import struct

import numpy as np
import pytest

from conftest import golden_path
from edge_cases import EDGE_TEXTS
from oracle_lib import Oracle, load_text

pytestmark = pytest.mark.gpu


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


def hexd(x):
    return format(struct.unpack("<Q", struct.pack("<d", x))[0], "016x")


def inst_from(plp, rel):
    return plp.Instance.from_text(load_text(rel))


def prm(plp, a):
    return plp.params(eps=a["eps"], theta=0.9 if a["theta"] is None else a["theta"], mu=a["mu"],
                      use_theta=a["theta"] is not None, first_fit=a["first_fit"], order=a["order"])


def row_fields(plp, row, pl):
    d = plp.to_json_dict(row, pl)
    return d


FIELDS = ("n_boxes", "volume_cm3", "stability_score", "fallback", "n_passes", "seed",
          "placements")


def ref_fields(res):
    return {k: res[k] for k in FIELDS}


# ------------------------------------------------------------- RNG --------

def test_device_rng_selftest_protocol(ctx, rng_vectors):
    for blk in rng_vectors["selftest"]:
        ops = [0] * 16 + [1] * 16
        args = [0] * 16 + [13] * 16
        f, u = ctx.rng_script([blk["seed"]], ops, args)
        assert [hexd(v) for v in f[0, :16]] == blk["random_hex"]
        assert list(map(int, u[0, 16:])) == blk["getrandbits13"]


def test_device_rng_scripts(ctx, rng_vectors):
    for sc in rng_vectors["scripts"]:
        f, u = ctx.rng_script([sc["seed"]], sc["ops"], sc["args"])
        for i, (op, want) in enumerate(zip(sc["ops"], sc["vals"])):
            got = hexd(f[0, i]) if op == 0 else int(u[0, i])
            assert got == want, (sc["seed"], i, op)


def test_device_rng_many_seeds_vs_oracle(ctx, orc):
    seeds = list(range(0, 300)) + [2**32 + 7, 2**45 + 1]
    ops = [0] * 400 + [2] * 300 + [1] * 50
    rr = random.Random(1)
    args = [0] * 400 + [rr.randint(1, 5000) for _ in range(300)] + [rr.randint(1, 32) for _ in range(50)]
    f, u = ctx.rng_script(seeds, ops, args)
    for si, s in enumerate(seeds):
        r = orc.rng(s)
        for i, (op, arg) in enumerate(zip(ops, args)):
            if op == 0:
                assert f[si, i] == orc.L.orc_rng_random(r)
            elif op == 1:
                assert u[si, i] == orc.L.orc_rng_getrandbits(r, arg)
            else:
                assert u[si, i] == orc.L.orc_rng_randrange(r, arg)


# ------------------------------------------------------- feasibility ------

def _random_state(rnd, L, W, H, n_box, n_cand):
    boxes = []
    tries = 0
    while len(boxes) < n_box and tries < n_box * 50:
        tries += 1
        w, d, h = rnd.randint(5, 60), rnd.randint(5, 60), rnd.randint(5, 60)
        x, y, z = rnd.randint(0, L - w), rnd.randint(0, W - d), rnd.randint(0, H - h)
        c = (x, y, z, x + w, y + d, z + h)
        if all(not (c[0] < b[3] and b[0] < c[3] and c[1] < b[4] and b[1] < c[4] and c[2] < b[5]
                    and b[2] < c[5]) for b in boxes):
            boxes.append(c)
    pts = set()
    for b in boxes:
        for cx in (b[0], b[3]):
            for cy in (b[1], b[4]):
                for cz in (b[2], b[5]):
                    pts.add((cx, cy, cz))
    pts = sorted(pts, key=lambda p: (p[2], p[1], p[0]))[:n_cand] or [(0, 0, 0)]
    return pts, boxes


@pytest.mark.parametrize("n_box,n_cand", [(0, 1), (1, 8), (30, 200), (200, 1500), (1500, 3000)])
def test_feas_check_all_vs_oracle(ctx, orc, n_box, n_cand):
    rnd = random.Random(n_box * 7 + n_cand)
    L, W, H = 587, 233, 220
    pts, boxes = _random_state(rnd, L, W, H, n_box, n_cand)
    for ors in ([(10, 20, 30)], [(40, 30, 20), (30, 40, 20), (40, 20, 30), (20, 40, 30),
                                 (30, 20, 40), (20, 30, 40)], [(1, 1, 1), (587, 233, 220)]):
        got = ctx.feas_check_all(pts, ors, boxes, L, W, H)
        want = orc.check_all(pts, ors, boxes, L, W, H)
        assert np.array_equal(got, want)


def test_feas_check_batch_vs_rows(ctx, orc):
    rnd = random.Random(99)
    L, W, H = 587, 233, 220
    rows = [_random_state(rnd, L, W, H, nb, nc) for nb, nc in ((0, 1), (5, 40), (60, 400), (3, 0))]
    ors = [(55, 45, 35), (45, 55, 35)]
    outs = ctx.feas_check_batch([r[0] if r[0] else np.zeros((0, 3)) for r in rows], ors,
                                [r[1] for r in rows], L, W, H)
    for (pts, boxes), got in zip(rows, outs):
        assert np.array_equal(got, orc.check_all(pts, ors, boxes, L, W, H) if pts else
                              np.zeros(0, np.uint8))


# ------------------------------------------------------------- FSR --------

def test_fsr_eval_bits_vs_reference(ctx, ref_runs):
    packs, want, mus = [], [], []
    for c in ref_runs:
        packs.append(np.asarray(c["result"]["placements"], np.int32).reshape(-1, 7))
        want.append((c["result"]["fsr_hex"], c["result"]["stability_score"], c["args"]["mu"]))
    for mu in sorted({w[2] for w in want}):
        idx = [i for i, w in enumerate(want) if w[2] == mu]
        fsr, score = ctx.fsr_eval([packs[i] for i in idx], mu=mu)
        for j, i in enumerate(idx):
            assert [hexd(v) for v in fsr[j]] == want[i][0], ref_runs[i]["case"]
            assert score[j] == want[i][1], ref_runs[i]["case"]


def test_fsr_eval_hand_kats(ctx):
    tower = [[0, 0, 0, 10 * i, 10, 10, 10] for i in range(10)]
    packs = [[[0, 0, 0, 0, 10, 10, 10]],
             [[0, 0, 0, 0, 14, 10, 10], [0, 14, 0, 0, 14, 10, 10], [0, 5, 0, 10, 14, 10, 10]],
             tower, np.zeros((0, 7))]
    fsr, score = ctx.fsr_eval(packs)
    assert fsr[0][0] == pytest.approx(2 / 3)
    assert list(fsr[1]) == pytest.approx([23 / 21, 19 / 21, 2 / 3])
    assert score[2] == 0.9 and score[3] == 0.0


# ------------------------------------------------ device-resident engine --

# the item decoder's two ways to its verdicts (plp_engine_config.verdicts)
VERDICTS = ("pair_cache", "point_state")


@pytest.mark.parametrize("verdicts", VERDICTS)
def test_engine_every_reference_case(plp, ctx, ref_runs, verdicts):
    """Each reference run as its own decoder in one multi-instance batch."""
    by_params = {}
    for c in ref_runs:
        a = c["args"]
        key = (a["passes"], a["eps"], a["theta"], a["mu"], a["order"], a["first_fit"])
        by_params.setdefault(key, []).append(c)
    for key, cases in by_params.items():
        insts = [inst_from(plp, c["instance"]) for c in cases]
        eng = plp.Engine(ctx, count_work=True, verdicts=verdicts)
        eng.load(insts, list(range(len(cases))), [c["args"]["seed"] for c in cases],
                 prm(plp, cases[0]["args"]))
        eng.run(key[0])
        eng.sync()
        rows, pl = eng.fetch()
        for i, c in enumerate(cases):
            assert rows[i].err == 0, c["case"]
            assert plp.to_json_dict(rows[i], pl[i]) == ref_fields(c["result"]), c["case"]
            assert rows[i].work_pair_box == sum(c["result"]["pass_pair_box"]), c["case"]
            # the engine also counts the list left after the last placement
            assert rows[i].max_ep >= c["result"]["max_cand"], c["case"]
        eng.close()


def test_engine_passes_in_pieces_equal_one_call(plp, ctx, ref_runs):
    """RNG and D4' state persist on the device across runs (run(1) x 4 ==
    run(4)), as the reference's time-budgeted loop needs."""
    c = next(c for c in ref_runs if c["case"] == "BR1_p4_theta")
    inst = inst_from(plp, c["instance"])
    eng = plp.Engine(ctx)
    eng.load([inst], [0], [c["args"]["seed"]], prm(plp, c["args"]))
    for _ in range(4):
        eng.run(1)
        eng.sync()
    rows, pl = eng.fetch()
    assert plp.to_json_dict(rows[0], pl[0]) == ref_fields(c["result"])


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("count_work", [False, True])
def test_engine_capacity_escalation(plp, ctx, ref_runs, count_work, verdicts):
    """A tiny EP capacity overflows (the live list, and with count_work the
    dead list); the engine reruns those decoders from their seeds with a
    larger capacity and the results are unchanged."""
    cases = [c for c in ref_runs if c["case"] in ("BR1_eps03_s1", "BR15_eps0", "plain_c1")]
    for c in cases:
        eng = plp.Engine(ctx, ep_cap=16, count_work=count_work, verdicts=verdicts)
        eng.load([inst_from(plp, c["instance"])], [0], [c["args"]["seed"]], prm(plp, c["args"]))
        eng.run(c["args"]["passes"])
        eng.sync()
        rows, pl = eng.fetch()
        assert rows[0].err == 0
        assert rows[0].ep_need > 16
        assert eng.describe()["ep_cap"] >= rows[0].ep_need
        assert plp.to_json_dict(rows[0], pl[0]) == ref_fields(c["result"]), c["case"]
        if count_work:
            assert rows[0].work_pair_box == sum(c["result"]["pass_pair_box"]), c["case"]
            assert rows[0].max_ep >= c["result"]["max_cand"], c["case"]


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("box_cap,ep_cap", [(8, 0), (16, 64), (48, 0)])
def test_engine_box_capacity_escalation(plp, ctx, box_cap, ep_cap, verdicts):
    """A box capacity below what decoders place stops them; the engine reruns
    them with room for every item (and, with a small extreme-point capacity
    too, more points), then settles on the most boxes any decoder placed.
    Every result equals the pinned digests of the uneven 256-instance batch
    (first 512 decoders) and of a graph replay after the settling."""
    import json as _json
    from plp_amd import synth
    texts, dec_inst, seeds, passes = synth.workload("mixed_256x16")
    with open(golden_path("fullsize_digests.json")) as f:
        pin = _json.load(f)["workloads"]["mixed_256x16"]["per_decoder"][:512]
    insts = [plp.Instance.from_text(t) for t in texts[:32]]
    eng = plp.Engine(ctx, box_cap=box_cap, ep_cap=ep_cap, use_graph=True, verdicts=verdicts)
    eng.load(insts, dec_inst[:512], seeds[:512], plp.params(eps=0.01, theta=0.9))
    for _ in range(2):
        eng.reset()
        eng.run(passes)
        eng.sync()
        rows, pl = eng.fetch()
        assert all(r.err == 0 for r in rows)
        assert plp.results_digest(rows, pl)[1] == pin
    eng.close()


def test_engine_graph_replay_matches_eager(plp, ctx, ref_runs):
    cases = [c for c in ref_runs if c["case"].startswith("multi_c2_s")]
    insts = [inst_from(plp, cases[0]["instance"])]
    seeds = [c["args"]["seed"] for c in cases]
    outs = []
    for use_graph in (False, True):
        eng = plp.Engine(ctx, use_graph=use_graph)
        eng.load(insts, [0] * len(seeds), seeds, prm(plp, cases[0]["args"]))
        for rep in range(3):   # replays with eager work and syncs between them
            eng.reset()
            eng.run(2)
            eng.sync()
            rows, pl = eng.fetch()
            got = [plp.to_json_dict(r, p) for r, p in zip(rows, pl)]
            assert got == [ref_fields(c["result"]) for c in cases], (use_graph, rep)
            outs.append(got)
        eng.close()


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("wpd", [1, 2, 4])
def test_engine_many_decoders_vs_oracle(plp, ctx, orc, wpd, verdicts):
    """K = 256 decoders of the 20-type instance (seeds 1..256, eps 0.01,
    theta 0.9, 2 passes) against the oracle, decoder by decoder, with 1, 2
    and 4 wavefronts per decoder."""
    text = load_text("synth/multi.txt")
    inst = plp.Instance.from_text(text)
    o_inst = orc.parse(text)[0]
    K = 256
    count = wpd == 2     # one width also checks the reference's work count
    eng = plp.Engine(ctx, waves_per_decoder=wpd, count_work=count, verdicts=verdicts)
    eng.load([inst], [0] * K, list(range(1, K + 1)), plp.params(eps=0.01, theta=0.9))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    for k in range(K):
        want = orc.run(o_inst, seed=k + 1, passes=2, eps=0.01, theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert got == {x: want[x] for x in got}, k
        if count:
            assert rows[k].work_pair_box == want["pair_box"], k
            assert 0 < rows[k].work_live_pair_box <= rows[k].work_pair_box, k
        else:
            assert rows[k].work_pair_box == 0 and rows[k].work_live_pair_box == 0, k


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("wpd", [1, 4])
def test_engine_mixed_batch_vs_oracle(plp, ctx, orc, wpd, verdicts):
    """Uneven instances (3..100 types) x 4 seeds in one persistent launch."""
    from plp_amd import synth
    texts = synth.mixed_batch(24, seed0=777_000)
    insts = [plp.Instance.from_text(t) for t in texts]
    dec_inst = [i for i in range(len(texts)) for _ in range(4)]
    seeds = [100 + 3 * i + s for i in range(len(texts)) for s in range(4)]
    eng = plp.Engine(ctx, waves_per_decoder=wpd, verdicts=verdicts)
    eng.load(insts, dec_inst, seeds, plp.params(eps=0.3, theta=0.9))
    eng.run(1)
    eng.sync()
    rows, pl = eng.fetch()
    for d, (i, s) in enumerate(zip(dec_inst, seeds)):
        want = orc.run(orc.parse(texts[i])[0], seed=s, passes=1, eps=0.3, theta=0.9)
        assert plp.to_json_dict(rows[d], pl[d]) == {k: want[k] for k in
                                                     ("n_boxes", "volume_cm3", "stability_score",
                                                      "fallback", "n_passes", "seed",
                                                      "placements")}, d


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("wpd", [2, 4])
def test_engine_many_types_vs_oracle(plp, ctx, orc, wpd, verdicts):
    """130 types of 2-3 copies (131 boxes in a packing, so FSR runs its LDS
    form, beyond 128 boxes), 32 seeds, eps 0.3, 2 passes, against the oracle."""
    from plp_amd import synth
    text = synth.uniform_instance(2040, 130, 20, 90, 1.1)
    inst = plp.Instance.from_text(text)
    o_inst = orc.parse(text)[0]
    K = 32
    eng = plp.Engine(ctx, waves_per_decoder=wpd, verdicts=verdicts)
    eng.load([inst], [0] * K, list(range(1, K + 1)), plp.params(eps=0.3, theta=0.9))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    assert max(r.n_boxes for r in rows) > 128
    for k in range(K):
        want = orc.run(o_inst, seed=k + 1, passes=2, eps=0.3, theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert got == {x: want[x] for x in got}, k


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("wpd", [1, 4])
def test_engine_large_footprints_vs_oracle(plp, ctx, orc, wpd, verdicts):
    """Footprints up to 620 x 590 cm (w * d > 131068, so the pair-state cache
    holds 32-bit support sums instead of clamped 16-bit ones) in a
    1300 x 1000 x 900 container, 48 seeds, eps 0.3, 2 passes."""
    import random
    rr = random.Random(2051)
    types = []
    for i in range(10):
        dims = [rr.randint(180, 620) for _ in range(3)]
        types.append(f"{i + 1} {dims[0]} 1 {dims[1]} {rr.randint(0, 1)} {dims[2]} 1 {rr.randint(1, 4)}")
    text = "1\n1 2051\n1300 1000 900\n10\n" + "\n".join(types) + "\n"
    o_inst = orc.parse(text)[0]
    assert max(a * b for t in o_inst[3] for a, b in ((t[0], t[2]), (t[0], t[4]), (t[2], t[4]))) > 131068
    K = 48
    eng = plp.Engine(ctx, waves_per_decoder=wpd, verdicts=verdicts)
    eng.load([plp.Instance.from_text(text)], [0] * K, list(range(1, K + 1)),
             plp.params(eps=0.3, theta=0.9))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    assert min(r.n_boxes for r in rows) > 1
    for k in range(K):
        want = orc.run(o_inst, seed=k + 1, passes=2, eps=0.3, theta=0.9)
        got = plp.to_json_dict(rows[k], pl[k])
        assert got == {x: want[x] for x in got}, k


@pytest.mark.parametrize("eps", [0.01, 1.0])
def test_engine_edge_instances_vs_oracle(plp, ctx, orc, eps):
    """Degenerate instances (nothing fits, one item, types without an allowed
    orientation, exact tilings) with seeds 0, 1 and 2^64 - 1, in one
    multi-instance launch; eps 1.0 draws a random pick at every placement."""
    names = list(EDGE_TEXTS)
    seeds = [0, 1, 2**64 - 1]
    dec_inst = [i for i in range(len(names)) for _ in seeds]
    dec_seed = [s for _ in names for s in seeds]
    eng = plp.Engine(ctx)
    eng.load([plp.Instance.from_text(EDGE_TEXTS[n]) for n in names], dec_inst, dec_seed,
             plp.params(eps=eps, theta=0.9))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    for d, (i, s) in enumerate(zip(dec_inst, dec_seed)):
        assert rows[d].err == 0, (names[i], s)
        want = orc.run(orc.parse(EDGE_TEXTS[names[i]])[0], seed=s, passes=2, eps=eps, theta=0.9)
        assert plp.to_json_dict(rows[d], pl[d]) == ref_fields(want), (names[i], s)


def test_engine_autotune_keeps_results(plp, ctx, ref_runs):
    cases = [c for c in ref_runs if c["case"].startswith("multi_c2_s")]
    inst = inst_from(plp, cases[0]["instance"])
    seeds = [c["args"]["seed"] for c in cases]
    eng = plp.Engine(ctx)
    eng.load([inst], [0] * len(seeds), seeds, prm(plp, cases[0]["args"]))
    cfg, ms = eng.autotune(1)
    assert cfg["ep_cap"] > 0 and ms > 0
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    assert [plp.to_json_dict(r, p) for r, p in zip(rows, pl)] == \
        [ref_fields(c["result"]) for c in cases]


# ------------------------------------- containers past 32-bit packed keys ---
# Sides up to 2^21 - 1, the reference device form's range
# (gpu/plp_gpu.cu:1263-1268), through 64-bit keys and support sums.

def _big_cases():
    with open(golden_path("big_cases.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


@pytest.mark.parametrize("verdicts", VERDICTS)
@pytest.mark.parametrize("name", ["cube2000", "br1_mm", "hc40_mm", "side_max"])
def test_engine_big_containers_match_reference_and_oracle(plp, ctx, orc, name, verdicts):
    """The reference CLI's eps 0 and eps 0.3 single runs (its g++ host
    build), and 16 seeds x 2 passes against the oracle, at 1 and 4 waves."""
    case = _big_cases()[name]
    inst = plp.Instance.from_text(case["text"])
    o_inst = orc.parse(case["text"])[0]
    for wpd in (1, 4):
        eng = plp.Engine(ctx, waves_per_decoder=wpd, verdicts=verdicts)
        for key, prm, seed in (("eps0", plp.params(eps=0.0, use_theta=False), 0),
                               ("eps03_seed5", plp.params(eps=0.3, theta=0.9), 5)):
            eng.load([inst], [0], [seed], prm)
            eng.run(1)
            eng.sync()
            rows, pl = eng.fetch()
            assert plp.to_json_dict(rows[0], pl[0]) == ref_fields(case["runs"][key]), (key, wpd)
        seeds = list(range(100, 116))
        eng.load([inst], [0] * 16, seeds, plp.params(eps=0.3, theta=0.9))
        eng.run(2)
        eng.sync()
        rows, pl = eng.fetch()
        eng.close()
        for k, sd in enumerate(seeds):
            want = orc.run(o_inst, seed=sd, passes=2, eps=0.3, theta=0.9)
            assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, (k, wpd)


def test_engine_millimetre_instance_is_the_centimetre_one_scaled(plp, ctx):
    """Every decision of the decoder is invariant under scaling all lengths
    by 10 (C2 and C1 compare lengths, C3 compares a ratio, the (z, y, x)
    order is kept, the RNG sees the same feasible counts): the BR1 packings
    in mm are the cm packings times 10, seed for seed."""
    cm = plp.Instance.from_text(load_text("data/instances/BR1_050.txt"))
    mm = plp.Instance.from_text(_big_cases()["br1_mm"]["text"])
    seeds = list(range(1, 65))
    out = []
    for inst in (cm, mm):
        eng = plp.Engine(ctx)
        eng.load([inst], [0] * 64, seeds, plp.params(eps=0.3, use_theta=False))
        eng.run(1)
        eng.sync()
        out.append(eng.fetch())
        eng.close()
    (rc, pc), (rm, pm) = out
    for k in range(64):
        n = rc[k].n_boxes
        assert rm[k].n_boxes == n and rm[k].volume == 1000 * rc[k].volume
        want = np.asarray(pc[k][:n]).copy()
        want[:, 1:] *= 10
        assert (np.asarray(pm[k][:n]) == want).all(), k


@pytest.mark.parametrize("verdicts", VERDICTS)
def test_engine_huge_footprints_64bit_sums(plp, ctx, orc, verdicts):
    """Footprints past 2^31 (support sums past 2^32): the oracle keeps
    Python's exact integers (plp/solver.py), as this build does; the
    reference's g++ build would overflow its int product w * d there."""
    types = [(60000, 1, 50000, 1, 300, 1, 2), (40000, 1, 45000, 1, 200, 1, 3),
             (20000, 1, 70000, 1, 250, 1, 2), (90000, 0, 30000, 1, 150, 1, 2)]
    inst = plp.Instance(100000, 100000, 1000, types)
    o_inst = (100000, 100000, 1000, types)
    seeds = list(range(1, 33))
    eng = plp.Engine(ctx, verdicts=verdicts)
    eng.load([inst], [0] * 32, seeds, plp.params(eps=0.3, theta=0.5))
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    eng.close()
    assert max(r.n_boxes for r in rows) >= 5
    for k, sd in enumerate(seeds):
        want = orc.run(o_inst, seed=sd, passes=2, eps=0.3, theta=0.5)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k
    row, rpl = ctx.plain_run(inst, 7, 2, plp.params(eps=0.3, theta=0.5))
    want = orc.run(o_inst, seed=7, passes=2, eps=0.3, theta=0.5)
    assert plp.to_json_dict(row, rpl) == {f: want[f] for f in FIELDS}


def test_feas_check_big_areas_vs_oracle(plp, ctx, orc):
    rng = np.random.default_rng(11)
    L, W, H = 100000, 100000, 1000
    boxes = []
    for _ in range(40):
        x, y = (int(v) for v in rng.integers(0, 80000, 2))
        w, d = (int(v) for v in rng.integers(5000, 20000, 2))
        z = int(rng.choice([0, 200, 400]))
        boxes.append((x, y, z, x + w, y + d, z + 200))
    cand = [(int(rng.integers(0, 90000)), int(rng.integers(0, 90000)), int(rng.choice([0, 200, 400])))
            for _ in range(300)]
    ors = [(60000, 50000, 100), (20000, 20000, 100), (90000, 40000, 50)]
    got = ctx.feas_check_all(cand, ors, boxes, L, W, H)
    want = orc.check_all(cand, ors, boxes, L, W, H)
    assert (got == want).all() and got.any()


def test_engine_rejects_side_past_21_bits(plp, ctx):
    eng = plp.Engine(ctx)
    big = plp.Instance.from_text(_big_cases()["side_over"]["text"])
    with pytest.raises(plp.PlpError) as ei:
        eng.load([big], [0], [0], plp.params())
    assert ei.value.code == plp.PLP_E_RANGE
    eng.close()


def test_block_engine_big_container_vs_oracle(plp, ctx, orc):
    case = _big_cases()["cube2000"]
    inst = plp.Instance.from_text(case["text"])
    o_inst = orc.parse(case["text"])[0]
    seeds = list(range(1, 17))
    eng = plp.Engine(ctx)
    eng.load([inst], [0] * 16, seeds, plp.params(eps=0.3, theta=0.9), blocks=plp.block_params())
    eng.run(2)
    eng.sync()
    rows, pl = eng.fetch()
    eng.close()
    for k, sd in enumerate(seeds):
        want = orc.block_run(o_inst, seed=sd, passes=2, eps=0.3, theta=0.9)
        assert plp.to_json_dict(rows[k], pl[k]) == {f: want[f] for f in FIELDS}, k


# ----------------------------------------------------------- plain form ---

def test_plain_form_reference_cases(plp, ctx, ref_runs):
    names = {"BR1_eps0", "BR2_eps0", "BR3_eps0", "SYN1_eps0", "BR1_eps03_s7", "BR2_p4_theta",
             "tower10_eps0", "tower9_eps0", "no_orient_eps0", "empty_qty_eps0", "too_big_eps0"}
    for c in ref_runs:
        if c["case"] not in names:
            continue
        a = c["args"]
        row, pl = ctx.plain_run(inst_from(plp, c["instance"]), a["seed"], a["passes"], prm(plp, a))
        assert plp.to_json_dict(row, pl) == ref_fields(c["result"]), c["case"]


def test_fsr_eval_sizes_around_register_limits(plp, ctx, orc):
    """FSR of packings of about 110-150 boxes from the 130-type instance (decoder
    output, varied support structures) and of towers of 128, 129, 200, 256
    and 257 boxes: the register forms (two and four boxes per lane) and the
    LDS form, bit for bit against the oracle."""
    from plp_amd import synth
    text = synth.uniform_instance(2040, 130, 20, 90, 1.1)
    eng = plp.Engine(ctx)
    eng.load([plp.Instance.from_text(text)], [0] * 8, list(range(1, 9)),
             plp.params(eps=0.5, theta=0.9))
    eng.run(1)
    eng.sync()
    rows, pl = eng.fetch()
    packs = [pl[k][:rows[k].n_boxes] for k in range(8)]
    for n in (128, 129, 200, 256, 257):
        packs.append(np.array([[0, 0, 0, 10 * i, 10, 10, 10] for i in range(n)], np.int32))
    assert max(len(p) for p in packs[:8]) > 128
    for mu in (0.2, 0.5):
        fsr, score = ctx.fsr_eval(packs, mu=mu)
        for j, p in enumerate(packs):
            placed7 = [[r[1], r[2], r[3], r[4], r[5], r[6], r[0]] for r in p]
            assert list(fsr[j][:len(p)]) == list(orc.compute_fsr(placed7, mu)), (j, len(p))
            assert score[j] == orc.stability_score(placed7, mu), (j, len(p))


def test_engine_fsr_beyond_reference_cap(plp, ctx, orc):
    """300 stacked boxes: FSR over more boxes than the reference device
    kernel's cap of 256 (plp_gpu.cu:1121-1146), bit-exact against the oracle
    and the closed form (mu / 0.3) * (n - i)."""
    n = 300
    text = f"1\n1 0\n10 10 {10 * n}\n1\n1 10 1 10 1 10 1 {n}\n"
    inst = plp.Instance.from_text(text)
    eng = plp.Engine(ctx)
    eng.load([inst], [0], [0], plp.params(eps=0.0, use_theta=False, mu=0.3))
    eng.run(1)
    eng.sync()
    rows, pl = eng.fetch()
    want = orc.run(orc.parse(text)[0], seed=0, eps=0.0, theta=None, mu=0.3)
    got = plp.to_json_dict(rows[0], pl[0])
    assert got == {k: want[k] for k in got}
    assert rows[0].n_boxes == n
    placed7 = [[r[1], r[2], r[3], r[4], r[5], r[6], r[0]] for r in pl[0][:n]]
    for mu in (0.3, 0.2):
        fsr, score = ctx.fsr_eval([pl[0][:n]], mu=mu)
        assert list(fsr[0]) == list(orc.compute_fsr(placed7, mu))            # bit-exact
        assert list(fsr[0]) == pytest.approx([mu / 0.3 * (n - i) for i in range(n)], rel=1e-12)
        assert score[0] == orc.stability_score(placed7, mu)
    assert rows[0].score == orc.stability_score(placed7, 0.3)
