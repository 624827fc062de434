"""Continuous-batching server: admission by slots and memory, overlap on the
device, and each request's result identical alone and in the mix."""
import time

import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plp():
    import plp_amd
    assert plp_amd.device_count() >= 1, "no HIP device"
    return plp_amd


def _requests(plp):
    from plp_amd import synth
    reqs = []
    for i, (n_dec, passes, eps) in enumerate([(300, 1, 0.01), (64, 2, 0.3), (1000, 1, 0.01),
                                              (17, 3, 0.5), (512, 1, 0.0), (128, 2, 0.01)]):
        text = synth.mixed_instance(9000 + i)
        reqs.append((text, plp.Instance.from_text(text), n_dec, 100 * i + 1, passes,
                     plp.params(eps=eps, theta=0.9)))
    return reqs


def _alone(plp, ctx, inst, n_dec, seed0, passes, prm):
    eng = plp.Engine(ctx)
    eng.load([inst], [0] * n_dec, [seed0 + k for k in range(n_dec)], prm)
    eng.run(passes)
    eng.sync()
    rows, pl = eng.fetch()
    eng.close()
    return [plp.to_json_dict(r, p) for r, p in zip(rows, pl)]


def _wait(srv, rid, placements=True, timeout=60.0):
    t0 = time.time()
    while True:
        res = srv.poll(rid, placements=placements)
        if res is not None:
            return res
        assert time.time() - t0 < timeout, "request did not finish"
        time.sleep(0.001)


def test_results_same_alone_and_in_the_mix(plp):
    reqs = _requests(plp)
    ctx = plp.Ctx(0)
    alone = [_alone(plp, ctx, inst, n, s0, p, prm) for (_, inst, n, s0, p, prm) in reqs]
    ctx.close()
    srv = plp.Server(0, max_inflight=len(reqs))
    ids = [srv.submit(inst, n, s0, p, prm) for (_, inst, n, s0, p, prm) in reqs]
    assert all(i is not None for i in ids)
    assert srv.inflight() == len(reqs)
    for rid, want, (_, inst, n, s0, p, prm) in zip(ids, alone, reqs):
        rows, pl = _wait(srv, rid)
        got = [plp.to_json_dict(r, q) for r, q in zip(rows, pl)]
        assert got == want
    assert srv.inflight() == 0
    srv.close()


def test_mixed_results_match_oracle_sample(plp):
    reqs = _requests(plp)[:3]
    orc = Oracle()
    srv = plp.Server(0, max_inflight=3)
    ids = [srv.submit(inst, n, s0, p, prm) for (_, inst, n, s0, p, prm) in reqs]
    for rid, (text, inst, n, s0, p, prm) in zip(ids, reqs):
        rows, pl = _wait(srv, rid)
        o_inst = orc.parse(text)[0]
        for k in (0, n // 2, n - 1):
            want = orc.run(o_inst, seed=s0 + k, passes=p, eps=prm.eps, theta=prm.theta)
            got = plp.to_json_dict(rows[k], pl[k])
            assert got == {x: want[x] for x in got}
    srv.close()


def test_queued_requests_share_one_launch(plp):
    """One slot: a long request runs alone; the ten submitted while it runs
    wait and then go to the device together, as one multi-instance launch;
    every result equals its request decoded alone."""
    from plp_amd import synth
    prm = plp.params(eps=0.01, theta=0.9)
    big = plp.Instance.from_text(synth.multi_instance())
    small = [(plp.Instance.from_text(synth.mixed_instance(9100 + i)), 16 + 7 * i, 50 * i + 3)
             for i in range(10)]
    srv = plp.Server(0, max_inflight=1)
    first = srv.submit(big, 4096, 1, 4, prm)
    assert srv.poll(first) is None   # the poll launches it (it runs for milliseconds)
    ids = [srv.submit(inst, n, s0, 4, prm) for (inst, n, s0) in small]
    assert first is not None and all(i is not None for i in ids)
    assert srv.inflight() == 11
    _wait(srv, first, placements=False)
    got = [_wait(srv, rid) for rid in ids]
    assert srv.stats() == {"launches": 2, "requests": 11,
                           "decoders": 4096 + sum(n for (_, n, _) in small)}
    srv.close()
    ctx = plp.Ctx(0)
    for (rows, pl), (inst, n, s0) in zip(got, small):
        assert [plp.to_json_dict(r, q) for r, q in zip(rows, pl)] == \
            _alone(plp, ctx, inst, n, s0, 4, prm)
    ctx.close()


def test_admission_by_memory(plp):
    reqs = _requests(plp)
    (_, inst, n, s0, p, prm) = reqs[2]
    tiny = plp.Server(0, max_inflight=4, mem_budget=1 << 20)
    assert tiny.submit(inst, 4096, 1, 1, prm) is None        # over the memory budget
    small = tiny.submit(inst, 4, 1, 1, prm)
    assert small is not None
    _wait(tiny, small, placements=False)
    again = tiny.submit(inst, 200, 1, 1, prm)                  # returned requests free the budget
    assert again is not None
    _wait(tiny, again, placements=False)
    tiny.close()


def test_poisson_burst_each_response_equals_alone(plp):
    """A replayed burst of Poisson arrivals (1500 requests/s, 256..4096
    decoders log-uniform, one request of 4096) on 4 slots: requests overlap
    on the device and queue behind busy slots, and every response equals the
    same request decoded alone."""
    import math
    import random
    from plp_amd import synth
    rng = random.Random(77)
    prm = plp.params(eps=0.01, theta=0.9)
    arrivals, t = [], 0.0
    for i in range(40):
        t += rng.expovariate(1500.0)
        n = 4096 if i == 5 else int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        arrivals.append((t, plp.Instance.from_text(synth.mixed_instance(9500 + i)), n,
                         1 + 10000 * i))
    srv = plp.Server(0, max_inflight=4)
    pending, results = {}, {}
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(arrivals) or pending:
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and arrivals[nxt][0] <= now:
            (_, inst, n, s0) = arrivals[nxt]
            rid = srv.submit(inst, n, s0, 1, prm)
            assert rid is not None
            pending[rid] = nxt
            nxt += 1
        for rid in list(pending):
            res = srv.poll(rid, placements=True)
            if res is not None:
                results[pending.pop(rid)] = res
        assert time.perf_counter() - t0 < 60.0
    st = srv.stats()
    srv.close()
    assert st["requests"] == 40 and st["launches"] <= 40
    ctx = plp.Ctx(0)
    for i, (_, inst, n, s0) in enumerate(arrivals):
        rows, pl = results[i]
        assert [plp.to_json_dict(r, q) for r, q in zip(rows, pl)] == \
            _alone(plp, ctx, inst, n, s0, 1, prm), i
    ctx.close()


def test_bad_request_is_rejected_and_server_keeps_serving(plp):
    reqs = _requests(plp)
    srv = plp.Server(0, max_inflight=2)
    big = plp.Instance(1 << 21, 2000, 2000, [(10, 1, 10, 1, 10, 1, 1)])   # a side past 21 bits
    with pytest.raises(plp.PlpError) as ei:
        srv.submit(big, 8, 0, 1, plp.params())
    assert ei.value.code == plp.PLP_E_RANGE
    (_, inst, n, s0, p, prm) = reqs[1]
    rid = srv.submit(inst, n, s0, p, prm)
    assert rid is not None
    _wait(srv, rid, placements=False)
    srv.close()
