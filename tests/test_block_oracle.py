"""The block-building engine's C oracle (oracle/plp_block_oracle.c) pinned
against the reference: every fixed-pass run of the reference engine
(blockproof/plp_block.cpp through oracle/block_driver) in
tests/golden/block_runs.json, whose single-pass runs are themselves pinned to
the reference's own CLI binary, plus geometric invariants of the packings."""
import json
import os

import pytest

from conftest import golden_path
from oracle_lib import Oracle, load_text

FIELDS = ("n_boxes", "volume_cm3", "stability_score", "fallback", "n_passes", "seed",
          "placements")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def block_runs():
    with open(golden_path("block_runs.json")) as f:
        return json.load(f)


def _run(orc, c):
    a = c["args"]
    inst = orc.parse(load_text(c["instance"]))[0]
    return orc.block_run(inst, seed=a["seed"], passes=a["passes"], eps=a["eps"], theta=a["theta"],
                         amax=a["amax"], bmax=a["bmax"], kmax=a["kmax"], topk=a["topk"],
                         pick=a["pick"])


def test_fixture_covers_every_pick_rule_and_parameter(block_runs):
    args = [c["args"] for c in block_runs]
    assert {a["pick"] for a in args} == {"first", "support", "contact"}
    assert any(a["eps"] == 0.0 for a in args) and any(a["theta"] is None for a in args)
    assert any((a["amax"], a["bmax"], a["kmax"]) != (3, 3, 4) for a in args)
    assert any(a["passes"] > 1 for a in args) and len(block_runs) >= 60


def test_driver_single_pass_equals_reference_cli(block_runs):
    n = 0
    for c in block_runs:
        if c["reference_cli"] is None:
            continue
        want = dict(c["reference_cli"])
        got = {k: c["result"][k] for k in want if k != "time_s"}
        want.pop("time_s")
        assert got == want, c["case"]
        n += 1
    assert n >= 35


def test_oracle_matches_every_reference_block_run(orc, block_runs):
    for c in block_runs:
        got = _run(orc, c)
        assert {k: got[k] for k in FIELDS} == {k: c["result"][k] for k in FIELDS}, c["case"]


def test_catalogue_matches_enumeration(orc):
    """Shape count of the catalogue (plp_block.cpp:193-218) against a direct
    enumeration, for a few instances and bounds."""
    import ctypes as C
    from oracle_lib import Inst, Item
    for rel in ("data/instances/BR1_050.txt", "synth/multi.txt", "synth/large.txt"):
        L, W, H, types = orc.parse(load_text(rel))[0]
        for amax, bmax, kmax in ((3, 3, 4), (1, 1, 1), (4, 2, 6)):
            want = 0
            for (l, lf, w, wf, h, hf, q) in types:
                for (ow, od, oh) in orc.orientations(l, lf, w, wf, h, hf):
                    for a in range(1, amax + 1):
                        for b in range(1, bmax + 1):
                            for k in range(1, kmax + 1):
                                if a * ow <= L and b * od <= W and k * oh <= H and a * b * k <= q:
                                    want += 1
            inst = orc.make_inst(L, W, H, types)

            class Shape(C.Structure):
                _fields_ = [("ti", C.c_int), ("oi", C.c_int), ("a", C.c_int), ("b", C.c_int),
                            ("k", C.c_int), ("cnt", C.c_longlong), ("bw", C.c_int),
                            ("bd", C.c_int), ("bh", C.c_int)]
            cap = len(types) * 6 * amax * bmax * kmax + 1
            cat = (Shape * cap)()
            by_type = (C.c_int * cap)()
            off = (C.c_int * (len(types) + 1))()
            orc.L.orc_block_catalogue.restype = C.c_int
            n = orc.L.orc_block_catalogue(C.byref(inst), amax, bmax, kmax, cat, cap, by_type, off)
            assert n == want, (rel, amax, bmax, kmax)
            for t in range(len(types)):   # shape_precedence within each type
                keys = [(-cat[i].cnt, -cat[i].bh, -cat[i].a, -cat[i].b, cat[i].oi)
                        for i in by_type[off[t]:off[t + 1]]]
                assert keys == sorted(keys)


def test_packings_are_geometrically_valid(orc, block_runs):
    """Decomposed items lie in the container and never overlap (C1, C2)."""
    for c in block_runs:
        L, W, H, _ = orc.parse(load_text(c["instance"]))[0]
        boxes = [(p[1], p[2], p[3], p[1] + p[4], p[2] + p[5], p[3] + p[6])
                 for p in c["result"]["placements"]]
        for b in boxes:
            assert 0 <= b[0] < b[3] <= L and 0 <= b[1] < b[4] <= W and 0 <= b[2] < b[5] <= H
        for i in range(len(boxes)):
            for j in range(i):
                a, b = boxes[i], boxes[j]
                assert not (a[0] < b[3] and b[0] < a[3] and a[1] < b[4] and b[1] < a[4]
                            and a[2] < b[5] and b[2] < a[5]), (c["case"], i, j)
        assert sum((b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]) for b in boxes) == \
            c["result"]["volume_cm3"]
