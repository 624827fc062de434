"""Pins the C oracle (oracle/plp_oracle.c) before it is trusted as the checker.

Every oracle function is checked against something it cannot merely restate:
the reference's checked-in parity JSONs, the reference g++ build's outputs
(tests/golden/ref_runs.json), CPython's own random/repr, the reference's
hand-computed known answers (tests/test_solver.py, tests/test_geometry.py,
tests/test_parser.py of the reference), closed forms and brute force.
"""
import itertools
import json
import math
import random
import struct

import numpy as np
import pytest

from conftest import golden_path
from oracle_lib import Oracle, load_text


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def inst_of(orc, rel):
    return orc.parse(load_text(rel))[0]


def strip(d):
    return {k: v for k, v in d.items() if k not in ("time_s",)}


def as_ref_fields(o):
    return {k: o[k] for k in ("n_boxes", "volume_cm3", "stability_score", "fallback", "n_passes",
                              "seed", "placements")}


def hexd(x):
    return format(struct.unpack("<Q", struct.pack("<d", x))[0], "016x")


# ------------------------------------------------ golden parity (R9, R11, R13)

@pytest.mark.parametrize("name", ["BR1", "BR2", "BR3", "SYN1"])
@pytest.mark.parametrize("kind", ["ref_cpu", "ref_host", "ref_gpu"])
def test_oracle_matches_checked_in_parity(orc, name, kind):
    rel = f"data/synthetic/{name}.txt" if name == "SYN1" else f"data/instances/{name}_050.txt"
    got = orc.run(inst_of(orc, rel), seed=0, passes=1, eps=0.0, theta=None)
    with open(golden_path("parity", kind, name + ".json")) as f:
        want = strip(json.load(f))
    assert as_ref_fields(got) == want


def test_oracle_matches_reference_build_every_case(orc, ref_runs):
    """95 runs of the reference g++ build: eps 0/0.3/0.5/1.0, multi-pass, theta
    on/off, mu 0.5, first-fit, input order, 64-bit seeds, all 19 BR instances,
    synthetic families and the reference's hand-sized cases."""
    for c in ref_runs:
        a = c["args"]
        got = orc.run(inst_of(orc, c["instance"]), seed=a["seed"], passes=a["passes"],
                      eps=a["eps"], theta=a["theta"], order=a["order"],
                      first_fit=a["first_fit"], mu=a["mu"])
        want = strip(c["result"])
        for k in ("n_boxes", "volume_cm3", "stability_score", "fallback", "n_passes", "seed",
                  "placements"):
            assert got[k] == want[k], (c["case"], k)
        assert got["pair_box"] == sum(want["pass_pair_box"]), c["case"]
        assert got["max_cand"] == want["max_cand"], c["case"]


def test_oracle_per_box_fsr_bits(orc, ref_runs):
    """compute_fsr bit patterns of the selected packing of every reference run
    (the reference CLI never prints these; the harness does)."""
    n_checked = 0
    for c in ref_runs:
        res = c["result"]
        rows = [[p[1], p[2], p[3], p[4], p[5], p[6], p[0]] for p in res["placements"]]
        got = [hexd(v) for v in orc.compute_fsr(rows, c["args"]["mu"])]
        assert got == res["fsr_hex"], c["case"]
        n_checked += len(got)
    assert n_checked > 5000


# ------------------------------------------------------------ RNG (R10)

def test_rng_selftest_protocol(orc, rng_vectors):
    for blk in rng_vectors["selftest"]:
        r = orc.rng(blk["seed"])
        assert [hexd(orc.L.orc_rng_random(r)) for _ in range(16)] == blk["random_hex"]
        assert [orc.L.orc_rng_getrandbits(r, 13) for _ in range(16)] == blk["getrandbits13"]


def test_rng_scripts_match_cpython(orc, rng_vectors):
    for sc in rng_vectors["scripts"]:
        r = orc.rng(sc["seed"])
        for op, arg, want in zip(sc["ops"], sc["args"], sc["vals"]):
            if op == 0:
                assert hexd(orc.L.orc_rng_random(r)) == want
            elif op == 1:
                assert orc.L.orc_rng_getrandbits(r, arg) == want
            else:
                assert orc.L.orc_rng_randrange(r, arg) == want


def test_mt19937_textbook_known_answer(orc):
    """Core twist/temper is the textbook generator: after init_genrand(5489)
    the 10000th output is 4123659995 (C++11 [rand.predef]).  The oracle seeds
    through init_by_array, so rebuild the default seeding by hand."""
    r = orc.rng(0)
    mt = [0] * 624
    mt[0] = 5489
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    for i in range(624):
        r.mt[i] = mt[i]
    r.mti = 624
    v = 0
    for _ in range(10000):
        v = orc.L.orc_rng_next(r)
    assert v == 4123659995


# ------------------------------------------------------------ repr (R13)

def test_py_repr_matches_cpython(orc, repr_vectors):
    bad = []
    for hx, want in repr_vectors:
        x = struct.unpack("<d", struct.pack("<Q", int(hx, 16)))[0]
        got = orc.py_repr(x)
        if got != want:
            bad.append((hx, want, got))
    assert not bad, bad[:5]


# ----------------------------------------------------- parser (R1, R2, R3)

def test_parse_real_br1(orc):
    L, W, H, items = inst_of(orc, "data/instances/BR1_050.txt")
    assert (L, W, H) == (587, 233, 220)
    assert items == [(60, 1, 40, 1, 32, 1, 64), (98, 1, 75, 1, 55, 1, 40), (59, 1, 39, 1, 60, 1, 64)] or \
        items[0] == (60, 1, 40, 1, 32, 1, 64)


def test_all_manifest_instances_match_manifest(orc):
    """tests/test_parser.py:42-55 of the reference (its len == 15 assertion is
    stale: the manifest has 19 entries)."""
    with open(golden_path("data", "manifest.json")) as f:
        man = json.load(f)["instances"]
    assert len(man) == 19
    for e in man:
        L, W, H, items = inst_of(orc, "data/" + e["file"])
        assert [L, W, H] == e["container_cm"]
        assert len(items) == e["n_types"]
        assert sum(t[6] for t in items) == e["total_qty"]
        assert L * W * H == e["container_volume_cm3"]
        assert sum(t[0] * t[2] * t[4] * t[6] for t in items) == e["item_volume_cm3"]


def _one(body):
    return "\n".join(["1", "1 1001", *body])


TWO = "\n".join(["2", "1 1001", "587 233 220", "2", "1 60 1 40 1 32 1 64", "2 98 1 75 1 55 1 40",
                 "2 2002", "1200 233 259", "1", "1 100 0 50 1 40 1 30"])


def test_parse_multiple_and_field_mapping(orc):
    a, b = orc.parse(TWO)
    assert a[:3] == (587, 233, 220) and len(a[3]) == 2
    assert b[:3] == (1200, 233, 259) and b[3][0] == (100, 0, 50, 1, 40, 1, 30)
    (inst,) = orc.parse(_one(["587 233 220", "1", "1 10 0 20 1 30 0 5"]))
    assert inst[3][0] == (10, 0, 20, 1, 30, 0, 5)


@pytest.mark.parametrize("bad", [
    _one(["587 233 220", "1", "1 10 1 20 1 30 1"]),
    _one(["587 233 220", "1", "1 10 2 20 1 30 1 5"]),
    _one(["587 0 220", "1", "1 10 1 20 1 30 1 5"]),
    _one(["587 233 220", "1", "1 0 1 20 1 30 1 5"]),
    _one(["587 233 220", "1", "1 10 1 20 1 30 1 -5"]),
    _one(["587 233 220", "1", "1 10.5 1 20 1 30 1 5"]),
    _one(["587 233 220", "2", "1 10 1 20 1 30 1 5", "3 5 1 6 1 7 1 2"]),
    "\n".join(["2", "1 1001", "587 233 220", "1", "1 10 1 20 1 30 1 5"]),
    TWO + "\n999",
    "0",
    "",
    "   \n  \n",
])
def test_parse_rejects_malformed(orc, bad):
    with pytest.raises(ValueError):
        orc.parse(bad)


def test_orientation_order_and_dedup(orc):
    """tests/test_solver.py:58-68 and tests/test_parser.py:101-118."""
    assert orc.orientations(10, 1, 10, 1, 10, 1) == [(10, 10, 10)]
    assert orc.orientations(2, 1, 3, 1, 4, 1) == [(3, 4, 2), (4, 3, 2), (2, 4, 3), (4, 2, 3),
                                                  (2, 3, 4), (3, 2, 4)]
    assert orc.orientations(2, 1, 3, 0, 4, 0) == [(3, 4, 2), (4, 3, 2)]
    assert orc.orientations(2, 0, 3, 0, 4, 0) == []
    # every orientation is a permutation of (l, w, h); brute force over flags
    for l, w, h in [(2, 3, 4), (5, 5, 7), (9, 9, 9), (1, 8, 8)]:
        for lf, wf, hf in itertools.product((0, 1), repeat=3):
            got = orc.orientations(l, lf, w, wf, h, hf)
            assert len(got) == len(set(got))
            vert = {h for (_, _, h) in got}
            allowed = {d for d, f in ((l, lf), (w, wf), (h, hf)) if f}
            assert vert == allowed
            for o in got:
                assert sorted(o) == sorted((l, w, h))


def test_item_sequence(orc):
    inst = (100, 100, 100, [(10, 1, 10, 1, 10, 1, 2), (50, 1, 50, 1, 50, 1, 1),
                            (20, 1, 50, 1, 125, 1, 2)])
    # volumes: 1000, 125000, 125000 -> ties broken by type then copy
    assert orc.item_sequence(inst, "volume") == [1, 2, 2, 0, 0]
    assert orc.item_sequence(inst, "input") == [0, 0, 1, 2, 2]


# ---------------------------------------------------- feasibility (R6)

def _brute_feasible(x, y, z, w, d, h, boxes, L, W, H):
    """Unit-cell definition: C2 box inside the container; C1 no cell of the
    new box is a cell of an old box; C3 at least half of the bottom face's
    cells rest on a cell whose box top is at z (or z == 0)."""
    if x + w > L or y + d > W or z + h > H:
        return 0
    cells = {(i, j, k) for i in range(x, x + w) for j in range(y, y + d) for k in range(z, z + h)}
    for (bx, by, bz, bx2, by2, bz2) in boxes:
        for (i, j, k) in cells:
            if bx <= i < bx2 and by <= j < by2 and bz <= k < bz2:
                return 0
    if z > 0:
        sup = 0
        for i in range(x, x + w):
            for j in range(y, y + d):
                if any(bx <= i < bx2 and by <= j < by2 and bz2 == z
                       for (bx, by, bz, bx2, by2, bz2) in boxes):
                    sup += 1
        if 2 * sup < w * d:
            return 0
    return 1


def test_feasibility_brute_force_tiny(orc):
    rnd = random.Random(7)
    L, W, H = 6, 5, 4
    for trial in range(60):
        boxes = []
        for _ in range(rnd.randint(0, 5)):   # non-overlapping random boxes
            for _ in range(20):
                w, d, h = rnd.randint(1, 3), rnd.randint(1, 3), rnd.randint(1, 2)
                x, y, z = rnd.randint(0, L - w), rnd.randint(0, W - d), rnd.randint(0, H - h)
                cand = (x, y, z, x + w, y + d, z + h)
                if all(not (cand[0] < b[3] and b[0] < cand[3] and cand[1] < b[4] and b[1] < cand[4]
                            and cand[2] < b[5] and b[2] < cand[5]) for b in boxes):
                    boxes.append(cand)
                    break
        pts = [(x, y, z) for x in range(L + 1) for y in range(W + 1) for z in range(H + 1)]
        ors = [(1, 1, 1), (2, 1, 1), (2, 3, 1), (3, 2, 2)]
        bits = orc.check_all(pts, ors, boxes, L, W, H).reshape(len(pts), len(ors))
        for pi, (x, y, z) in enumerate(pts):
            for oi, (w, d, h) in enumerate(ors):
                assert bits[pi, oi] == _brute_feasible(x, y, z, w, d, h, boxes, L, W, H), \
                    (trial, (x, y, z), (w, d, h), boxes)


def test_c3_double_test_is_exactly_2s_lt_wd(orc):
    """(double)s/(double)(w*d) < 0.5  <=>  2s < w*d for s <= w*d < 2^42."""
    rnd = random.Random(3)
    for _ in range(20000):
        wd = rnd.randint(1, 1 << 40)
        s = rnd.choice([wd // 2, wd // 2 + 1, (wd + 1) // 2, (wd - 1) // 2, rnd.randint(0, wd)])
        assert (s / wd < 0.5) == (2 * s < wd)


def test_geometry_kats(orc):
    """tests/test_geometry.py:15-85 via the feasibility predicate."""
    L = W = H = 100
    lower = [(0, 0, 0, 10, 10, 10)]
    f = lambda x, y, z, w, d, h, b: orc.L.orc_feasible_one(  # noqa: E731
        x, y, z, w, d, h, np.asarray(b, np.int32).ctypes.data if b else None, len(b), L, W, H)
    assert f(0, 0, 10, 10, 10, 10, lower) == 1          # full support
    assert f(0, 0, 10, 20, 10, 10, lower) == 1          # exactly half -> feasible
    assert f(7, 0, 10, 10, 10, 10, lower) == 0          # 0.3 support
    assert f(10, 0, 0, 10, 10, 10, lower) == 1          # touching face is no overlap
    assert f(5, 0, 0, 10, 10, 10, lower) == 0           # interior overlap
    assert f(0, 0, 10, 10, 10, 10, [(0, 0, 0, 10, 10, 5)]) == 0   # other level: no support
    two = [(0, 0, 0, 10, 10, 10), (10, 0, 0, 20, 10, 10)]
    assert f(5, 0, 10, 14, 10, 10, two) == 1
    assert f(0, 5, 10, 10, 10, 10, two) == 1            # 0.5 on a
    assert f(0, 0, 0, 100, 50, 10, []) == 1 and f(0, 0, 0, 101, 50, 10, []) == 0


# ------------------------------------------------------------ FSR (R11)

def test_fsr_hand_kats(orc):
    """tests/test_solver.py:107-145 (their approx values; rows x,y,z,w,d,h,type)."""
    f = orc.compute_fsr([[0, 0, 0, 10, 10, 10, 0]])
    assert f[0] == pytest.approx(2 / 3)
    f = orc.compute_fsr([[0, 0, 0, 10, 10, 10, 0], [0, 0, 10, 10, 10, 10, 0]])
    assert list(f) == pytest.approx([4 / 3, 2 / 3])
    f = orc.compute_fsr([[0, 0, 0, 14, 10, 10, 0], [14, 0, 0, 14, 10, 10, 0],
                         [5, 0, 10, 14, 10, 10, 0]])
    assert list(f) == pytest.approx([23 / 21, 19 / 21, 2 / 3])
    tower = [[0, 0, 10 * i, 10, 10, 10, 0] for i in range(3)]
    assert list(orc.compute_fsr(tower)) == pytest.approx([2.0, 4 / 3, 2 / 3])
    assert orc.stability_score(tower) == pytest.approx(2 / 3)
    assert orc.stability_score([]) == 0.0
    t10 = [[0, 0, 10 * i, 10, 10, 10, 0] for i in range(10)]
    assert orc.stability_score(t10) == 0.9


def test_fsr_closed_form_towers(orc):
    """A tower of n equal cubes carries (n - i) masses on box i, so
    FSR_i = (mu / 0.3) * (n - i) exactly when the sum is exact."""
    for n in range(1, 30):
        tower = [[0, 0, 10 * i, 10, 10, 10, 0] for i in range(n)]
        f = orc.compute_fsr(tower, mu=0.3)
        assert list(f) == pytest.approx([float(n - i) for i in range(n)], rel=1e-12)


def test_fsr_conserves_mass(orc):
    """Every box's own mass reaches the floor: the floor boxes' loads sum to
    the total mass (up to rounding)."""
    c = json.load(open(golden_path("ref_runs.json")))
    for case in c[:20]:
        rows = [[p[1], p[2], p[3], p[4], p[5], p[6], p[0]] for p in case["result"]["placements"]]
        if not rows:
            continue
        f = orc.compute_fsr(rows, mu=0.3)   # mu/0.3 = 1 -> fsr = load/m
        m = [50.0 * r[3] * r[4] * r[5] / 1e6 for r in rows]
        floor_load = sum(fi * mi for fi, mi, r in zip(f, m, rows) if r[2] == 0)
        assert floor_load == pytest.approx(sum(m), rel=1e-9)


# --------------------------------------------------- decode invariants (R9)

def test_decode_invariants_all_instances(orc):
    """tests/test_solver.py:288-300: C2, pairwise C1 and final-state C3."""
    with open(golden_path("data", "manifest.json")) as f:
        man = json.load(f)["instances"]
    for e in man:
        inst = inst_of(orc, "data/" + e["file"])
        L, W, H, _ = inst
        pl = orc.decode_pass(inst)
        assert len(pl) > 0
        for i, (x, y, z, w, d, h, t) in enumerate(pl):
            assert 0 <= x and x + w <= L and 0 <= y and y + d <= W and 0 <= z and z + h <= H
            for (x2, y2, z2, w2, d2, h2, _) in pl[i + 1:]:
                assert not (x < x2 + w2 and x2 < x + w and y < y2 + d2 and y2 < y + d
                            and z < z2 + h2 and z2 < z + h)
            if z > 0:
                s = 0
                for (x2, y2, z2, w2, d2, h2, _) in pl:
                    if z2 + h2 == z:
                        ox = min(x + w, x2 + w2) - max(x, x2)
                        oy = min(y + d, y2 + d2) - max(y, y2)
                        if ox > 0 and oy > 0:
                            s += ox * oy
                assert 2 * s >= w * d
        assert int((pl[:, 3] * pl[:, 4] * pl[:, 5]).sum()) <= e["item_volume_cm3"]


def test_first_fit_equals_eps_zero(orc):
    inst = inst_of(orc, "synth/two_types.txt")
    a = orc.run(inst, seed=5, eps=0.3, first_fit=True, theta=None)
    b = orc.run(inst, seed=5, eps=0.0, theta=None)
    assert a["placements"] == b["placements"]
