"""plp_feas_check_all_noalloc, the allocation-free feasibility step, against
plp_feas_check_all (the entry point it varies) and the oracle: bit for bit,
on two calls in a row, the second with a smaller input (the shared scratch
keeps the first call's contents), plus its size limit."""
import random

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _scene(rnd, L, W, H, n_box, n_cand):
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
    pts = sorted({(cx, cy, cz) for b in boxes for cx in (b[0], b[3]) for cy in (b[1], b[4])
                  for cz in (b[2], b[5])}, key=lambda p: (p[2], p[1], p[0]))[:n_cand]
    return pts or [(0, 0, 0)], boxes


def test_noalloc_matches_check_all_twice_then_smaller():
    import plp_amd
    ctx = plp_amd.Ctx(0)
    orc = Oracle()
    rnd = random.Random(2026)
    L, W, H = 587, 233, 220
    ors = [(40, 30, 20), (30, 40, 20), (40, 20, 30), (20, 40, 30), (30, 20, 40), (20, 30, 40)]
    for n_box, n_cand in ((700, 2500), (40, 90)):
        pts, boxes = _scene(rnd, L, W, H, n_box, n_cand)
        want = ctx.feas_check_all(pts, ors, boxes, L, W, H)
        got = ctx.feas_check_all_noalloc(pts, ors, boxes, L, W, H)
        assert np.array_equal(got, want), (n_box, n_cand)
        assert np.array_equal(got, orc.check_all(pts, ors, boxes, L, W, H))
        assert got.any() and not got.all()
    # 64-bit support sums through the same kernel
    big = [(0, 0, 0, 60000, 50000, 300), (60000, 0, 0, 100000, 50000, 300)]
    cand = [(0, 0, 300), (30000, 0, 300), (50000, 10000, 300)]
    big_ors = [(70000, 40000, 100), (20000, 20000, 100)]
    got = ctx.feas_check_all_noalloc(cand, big_ors, big, 100000, 100000, 1000)
    assert np.array_equal(got, orc.check_all(cand, big_ors, big, 100000, 100000, 1000))
    ctx.close()


def test_noalloc_rejects_more_than_its_scratch():
    import plp_amd
    ctx = plp_amd.Ctx(0)
    pts = np.zeros((3_000_000, 3), np.int32)   # 3e6 x 6 pairs = 18 MB of output alone
    with pytest.raises(plp_amd.PlpError) as ei:
        ctx.feas_check_all_noalloc(pts, [(1, 1, 1)] * 6, [], 587, 233, 220)
    assert ei.value.code == plp_amd.PLP_E_ARG and "scratch" in str(ei.value)
    # the entry point still works after the refusal
    got = ctx.feas_check_all_noalloc([(0, 0, 0)], [(10, 10, 10)], [], 587, 233, 220)
    assert got.tolist() == [1]
    ctx.close()
