#!/usr/bin/env python3
"""Decoded-packings/s benchmark of the MI355X pallet-loading decoder.

Headline workload (one GPU): 256 synthetic instances (3..100 item types
each, dims U[25,115], ~1.2x the 587x233x220 container) x 16 seeds = 4096
device-resident decoders in one persistent launch captured in a hipGraph,
launch shape autotuned (waves per decoder x extreme-point capacity x blocks
per CU), eps-greedy eps = 0.01, theta = 0.9, one pass per decoder.  Instance
tables and decoder state are resident in HBM before the timed region.  One
step = reset + run + synchronise + plp_engine_fetch of every decoder's row
and placements into host arrays (the decoders write them into pinned host
memory while they run).  After timing, the digest of the last step's
outputs must equal the one tests/golden/fullsize_digests.json pins (the C
oracle and the reference's g++ build agree on every decoder there).

Printed: one JSON line with the contract fields plus
  roofline      decode_kernel against the VALU issue roof: the VALU lane-ops
                it executes per launch (SQ_INSTS_VALU x 64 from the committed
                counter profile) / kernel time; beside it the live pair x box
                tests it resolves (counted on the device) x 10 lane-ops per
                test, and the reference's brute-force test count over the
                same decoders (work_avoided times the live tests)
  cpu_baseline  the reference's g++ build (oracle/_ref/ref_driver) on the
                same 4096 decoders over the host cores of this job
  extras        the 20-type instance x 1024 decoders x 4 passes, the 100-type
                instance (2048 decoders) in 587x233x220 and in 880x349x330, the
                256 instances x 64 seeds (16384 decoders, shared prefixes),
                the plain one-step-per-launch GPU form, the block-building
                engine (GPU and reference CPU), JSON emission, Poisson serving
usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "plp-gpu_amd"))

HEADLINE = "mixed_256x16"
EPS, THETA = 0.01, 0.9
# One pair x box test of the reference predicate (plp_common.hpp:297-321):
# C1 6 compares + ANDs; C3 1 compare + 2 min + 2 max + 2 sub + 1 mul + 1 add.
OPS_PER_TEST = 10
# MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz VALU issue = 39.3 T lane-ops/s
# (= FP64 vector 78.6 TFLOP/s counting an FMA as two flops); HBM3E 8 TB/s.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
PINNED = os.path.join(REPO, "tests", "golden", "fullsize_digests.json")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cores() -> dict:
    """CPUs this job may run on: the affinity mask, the cgroup quota if any,
    and nproc as the OS reports it."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    usable = aff if quota is None else max(1, min(aff, int(quota)))
    return {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "usable": usable}


def cpu_baseline(texts, seeds_per_inst: int, threads: int, label: str) -> dict | None:
    """The reference g++ build (oracle/_ref/ref_driver bench: each instance
    decoded by seeds 1..seeds_per_inst, one pass) on `threads` threads; where
    it was not built, the C oracle (tests/oracle_lib.py) on a thread pool."""
    ref = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    n_dec = len(texts) * seeds_per_inst
    try:
        if os.path.exists(ref):
            tmp = tempfile.mkdtemp(prefix="plp_bench_", dir=REPO)
            paths = []
            try:
                for i, t in enumerate(texts):
                    paths.append(os.path.join(tmp, f"i{i:04d}.txt"))
                    with open(paths[-1], "w") as f:
                        f.write(t)
                out = json.loads(subprocess.run(
                    [ref, "bench", *paths, "--K", str(seeds_per_inst), "--passes", "1", "--seed",
                     "1", "--eps", str(EPS), "--theta", str(THETA), "--threads", str(threads)],
                    check=True, capture_output=True, text=True, timeout=600).stdout)
            finally:
                for pth in paths:
                    os.remove(pth)
                os.rmdir(tmp)
            kind, secs = "reference", out["seconds"]
        else:
            from concurrent.futures import ThreadPoolExecutor
            sys.path.insert(0, os.path.join(REPO, "tests"))
            from oracle_lib import Oracle
            orc = Oracle()
            insts = [orc.parse(t)[0] for t in texts]
            jobs = [(i, s) for i in range(len(texts)) for s in range(1, seeds_per_inst + 1)]
            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(lambda j: orc.run(insts[j[0]], seed=j[1], passes=1, eps=EPS,
                                              theta=THETA), jobs))
            kind, secs = "port", time.perf_counter() - t0
    except Exception as e:  # noqa: BLE001 -- a missing CPU baseline is reported, not fatal
        log(f"cpu_baseline unavailable: {e}")
        return None
    return {"value": round(n_dec / secs, 2), "unit": "packings/s", "cores": threads,
            "kind": kind,
            "sample": f"{label}: {len(texts)} instances x seeds 1..{seeds_per_inst} "
                      f"({n_dec} decoders, one pass, eps {EPS}, theta {THETA}); {secs:.2f} s "
                      f"wall on {threads} threads"}


def block_cpu_baseline(text: str, threads: int, n: int = 1024, passes: int = 1200) -> dict | None:
    """The reference block engine (blockproof/plp_block.cpp through
    oracle/_ref/block_driver): n decoders x `passes` passes on host threads."""
    drv = os.path.join(REPO, "oracle", "_ref", "block_driver")
    if not os.path.exists(drv):
        return None
    inst_path = os.path.join(REPO, f"_bench_block_{os.getpid()}.txt")
    with open(inst_path, "w") as f:
        f.write(text)
    try:
        out = json.loads(subprocess.run(
            [drv, "bench", inst_path, "--K", str(n), "--passes", str(passes), "--seed", "1",
             "--eps", str(EPS), "--theta", str(THETA), "--threads", str(threads)],
            check=True, capture_output=True, text=True, timeout=600).stdout)
    except Exception as e:  # noqa: BLE001 -- reported, not fatal
        log(f"block cpu baseline unavailable: {e}")
        return None
    finally:
        os.remove(inst_path)
    return {"value": round(out["packings_per_s"], 1), "unit": "packings/s", "cores": threads,
            "kind": "reference",
            "sample": f"{n} decoders x {passes} passes, reference block options; "
                      f"{out['seconds']:.2f} s wall on {threads} threads"}


def serve(plp_amd, synth, seconds: float, rate: float, passes: int, max_inflight: int,
          seed: int = 2031) -> dict:
    """Poisson arrivals at `rate` req/s for `seconds`; each request is one
    instance (types U{3..100}, dims U[25,115]) with log-uniform 256..4096
    decoders.  Latency = completion time - arrival time."""
    import math
    import random
    rng = random.Random(seed)
    t, arrivals = 0.0, []
    while True:
        t += rng.expovariate(rate)
        if t >= seconds:
            break
        n_dec = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        arrivals.append((t, rng.randrange(1 << 30), n_dec))
    log(f"serve: {len(arrivals)} requests over {seconds:.0f} s at {rate:.0f}/s")
    insts = [plp_amd.Instance.from_text(synth.mixed_instance(s)) for (_, s, _) in arrivals]
    prm = plp_amd.params(eps=EPS, theta=THETA)
    srv = plp_amd.Server(0, max_inflight=max_inflight)
    # warm every slot's buffers with the largest request shape (pass counts
    # differ, so that no two of these share a launch and each takes a slot)
    big = max(range(len(arrivals)), key=lambda i: arrivals[i][2] * insts[i].n_items)
    warm = []
    for k in range(max_inflight):
        rid = srv.submit(insts[big], arrivals[big][2], 1, passes + 1 + k, prm)
        if rid is not None:
            warm.append(rid)
    while warm:
        warm = [r for r in warm if srv.poll(r) is None]
    st0 = srv.stats()
    pending, inflight, lat, done_dec = [], {}, [], 0
    nxt = 0
    t0 = time.perf_counter()
    while nxt < len(arrivals) or pending or inflight:
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and arrivals[nxt][0] <= now:
            pending.append(nxt)
            nxt += 1
        while pending:
            i = pending[0]
            rid = srv.submit(insts[i], arrivals[i][2], 1 + i, passes, prm)
            if rid is None:
                break
            inflight[rid] = i
            pending.pop(0)
        for rid in list(inflight):
            res = srv.poll(rid)
            if res is not None:
                i = inflight.pop(rid)
                lat.append(time.perf_counter() - t0 - arrivals[i][0])
                done_dec += arrivals[i][2]
        if not inflight and not pending and nxt < len(arrivals):
            time.sleep(max(0.0, min(0.002, arrivals[nxt][0] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat.sort()
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))]  # noqa: E731
    return {"requests": len(lat), "rate_req_per_s": rate, "seconds": round(wall, 2),
            "passes_per_decoder": passes, "max_inflight": max_inflight,
            "packings_per_s": round(done_dec * passes / wall, 1),
            "p50_ms": round(1e3 * q(0.50), 2), "p99_ms": round(1e3 * q(0.99), 2),
            "max_ms": round(1e3 * lat[-1], 2),
            "launches": st["launches"] - st0["launches"],
            "requests_per_launch": round(len(lat) / max(1, st["launches"] - st0["launches"]), 3)}


def load_pinned() -> dict:
    with open(PINNED) as f:
        return json.load(f)["workloads"]


def timed_workload(plp_amd, synth, ctx, name: str, steps: int, warmup: int, shape: str = "") -> dict:
    """Decoded packings/s of a named workload: autotuned (or fixed) launch
    shape, graph-captured launch; each timed step resets, runs, synchronises
    and fetches every decoder's row and placements into preallocated host
    arrays.  Returns the figures, the engine and the last step's outputs."""
    texts, dec_inst, seeds, passes = synth.workload(name)
    insts = [plp_amd.Instance.from_text(t) for t in texts]
    prm = plp_amd.params(eps=EPS, theta=THETA)
    blocks = plp_amd.block_params() if synth.is_block_workload(name) else None
    if shape:
        wpd, cap, bpc, boxes, verd, share, twpd, tverd = \
            ([int(v) for v in shape.split(",")] + [0] * 5)[:8]
        eng = plp_amd.Engine(ctx, ep_cap=cap, blocks_per_cu=bpc, use_graph=True,
                             waves_per_decoder=wpd, box_cap=boxes, verdicts=verd,
                             share_prefix=share, trace_waves=twpd, trace_verdicts=tverd)
        eng.load(insts, dec_inst, seeds, prm, blocks=blocks)
        # automatic verdicts without the autotuner: point states (2)
        tuned = {"waves_per_decoder": wpd, "ep_cap": cap, "blocks_per_cu": bpc, "box_cap": boxes,
                 "verdicts": verd if (verd or blocks is not None) else 2, "share_prefix": share,
                 "trace_waves": twpd, "trace_verdicts": tverd, "fixed": True}
    else:
        eng = plp_amd.Engine(ctx, use_graph=True)
        eng.load(insts, dec_inst, seeds, prm, blocks=blocks)
        log(f"{name}: tuning launch shape")
        tuned, tune_ms = eng.autotune(passes)
        log(f"{name}: autotune {tuned} ({tune_ms:.3f} ms per launch)")
    n = len(seeds)
    rows = (plp_amd.RowResult * n)()
    pl = np.zeros((n, eng.max_items(), 7), np.int32)
    # each step: reset, run, then sync + fetch in one call whose copy of
    # every decoder's row and packing overlaps the decode
    for _ in range(warmup):
        eng.reset()
        eng.run(passes)
        eng.sync_fetch(out=(rows, pl))
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.reset()
        eng.run(passes)
        eng.sync_fetch(out=(rows, pl))
        kernel_ms.append(eng.last_ms())
    elapsed = time.perf_counter() - t0
    assert all(r.err == 0 for r in rows)
    digest, _ = plp_amd.results_digest(rows, pl)
    return {"name": name, "texts": texts, "insts": insts, "dec_inst": dec_inst, "seeds": seeds,
            "passes": passes, "eng": eng, "tuned": tuned, "steps": steps,
            "packings_per_s": n * passes * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "kernel_ms": sum(kernel_ms) / len(kernel_ms), "digest": digest}


def parity_note(res: dict, pinned: dict) -> dict:
    pin = pinned.get(res["name"])
    ok = pin is not None and pin["digest"] == res["digest"]
    if not ok:
        raise SystemExit(f"{res['name']}: timed outputs differ from the pinned digest "
                         f"({res['digest']} vs {pin and pin['digest']})")
    return {"digest": res["digest"], "matches_pinned": True, "decoders": len(res["seeds"])}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--workload", default=HEADLINE,
                    help="headline workload (plp_amd.synth.workload names)")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape WPD,EP_CAP,BLOCKS_PER_CU[,BOX_CAP[,VERDICTS[,SHARE"
                         "[,TRACE_WAVES,TRACE_VERDICTS]]]] instead of autotuning (the "
                         "plp_engine_config fields of those names)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson serving run in extras (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-inflight", type=int, default=8)
    ap.add_argument("--serve-only", action="store_true",
                    help="only the Poisson serving run; prints its JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        x = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "launches": 3}))
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: decoders never exchange data, so a multi-GPU run is "
              "N independent copies of this job, each GPU decoding its own share of the seeds; "
              "run with --gpus 1", file=sys.stderr)
        return 2
    import plp_amd
    from plp_amd import synth

    if a.serve_only:
        print(json.dumps(serve(plp_amd, synth, a.serve_seconds, a.serve_rate, 1,
                               a.serve_inflight)), flush=True)
        return 0
    pinned = load_pinned()
    ctx = plp_amd.Ctx(0)
    prm = plp_amd.params(eps=EPS, theta=THETA)
    head = timed_workload(plp_amd, synth, ctx, a.workload, a.steps, a.warmup, a.shape)
    parity = parity_note(head, pinned)
    eng = head["eng"]
    shape = eng.describe()
    n_dec, passes = len(head["seeds"]), head["passes"]
    ms_launch = head["kernel_ms"]
    kernel_name = "block_kernel" if synth.is_block_workload(a.workload) else "decode_kernel"
    live = ref_tests = achieved = None
    if kernel_name == "decode_kernel":
        # The pair x box tests of the same decoders, from one untimed run of an
        # engine that also keeps the dead extreme points (same results): the
        # live tests this kernel evaluates, and the reference's brute-force count.
        ec = plp_amd.Engine(ctx, waves_per_decoder=shape["waves_per_decoder"], count_work=True)
        try:
            ec.load(head["insts"], head["dec_inst"], head["seeds"], prm)
            ec.run(passes)
            ec.sync()
            crows, cpl = ec.fetch()
            assert plp_amd.results_digest(crows, cpl)[0] == head["digest"]
            ref_tests = sum(r.work_pair_box for r in crows)
            live = sum(r.work_live_pair_box for r in crows)
            achieved = live * OPS_PER_TEST / (ms_launch * 1e-3) / 1e12
        except plp_amd.PlpError as e:
            # the dead points the count keeps in LDS can outgrow a block (the
            # 880x349x330 workload's reference lists pass 3000 points)
            log(f"work count unavailable: {e}")
        ec.close()
    # Per-launch hardware counters of this workload and launch shape, from the
    # committed rocprofv3 summary (FETCH_SIZE / WRITE_SIZE scaled by the copy
    # calibration; SQ_INSTS_VALU wave instructions x 64 lanes).
    prof = {}
    pfile = os.path.join(REPO, "profiles", f"{kernel_name}_{a.workload}_counters.json")
    if os.path.exists(pfile):
        with open(pfile) as f:
            prof = json.load(f)
    # (only when the profile ran the item decoder with the same verdict kind)
    if prof.get("verdicts", 0) != head["tuned"].get("verdicts", 0):
        log(f"{os.path.basename(pfile)} was taken with verdicts {prof.get('verdicts')}, "
            f"this run uses {head['tuned'].get('verdicts')}: counters not used")
        prof = {}
    traffic = prof.get("hbm_bytes_per_launch")
    executed = None
    if prof.get("valu_insts_per_launch"):
        executed = prof["valu_insts_per_launch"] * 64 / (ms_launch * 1e-3) / 1e12
    texts = head["texts"]
    line = {
        "metric": "decoded packings/s at bit-parity vs plain GPU & g++ host; p50/p99 latency",
        "value": round(head["packings_per_s"], 1),
        "unit": "packings/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+i32",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {len(texts)} instance(s), {n_dec} decoders "
                               f"(plp_amd.synth.workload), eps {EPS}, theta {THETA}, "
                               f"{passes} pass(es) per decoder per step, one persistent "
                               f"hipGraph launch; step = reset + run + sync + fetch of every "
                               f"row and placement (plp_engine_sync_fetch: the copy overlaps "
                               f"the decode)",
                   "decoders": n_dec, "passes_per_step": passes, "parallelism": "none (1 GPU)",
                   "launch": shape, "autotuned": head["tuned"]},
        "parity": parity,
        # achieved: the VALU lane-ops the kernel executes per launch (the
        # committed counter profile of this kernel and verdict kind) over the
        # kernel time; without a matching profile, the live pair x box tests
        # it resolves x 10 lane-ops.  The point-state decoder resolves those
        # tests without executing each of them, so the executed figure is the
        # one bounded by the issue roof.
        "roofline": {"bound": "valu",
                     "achieved": None if (executed or achieved) is None
                     else round(executed if executed is not None else achieved, 3),
                     "achieved_from": "SQ_INSTS_VALU x 64" if executed is not None
                     else "live pair x box tests x 10",
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tlane-op/s",
                     "frac": None if (executed or achieved) is None
                     else round((executed if executed is not None else achieved) / VALU_PEAK_TOPS, 4),
                     "live_test_rate_tops": None if achieved is None else round(achieved, 3),
                     "valu_lane_utilisation": prof.get("valu_lane_utilisation"),
                     "traffic": traffic,
                     "kernel": kernel_name, "kernel_ms": round(ms_launch, 4),
                     "live_pair_box_tests_per_launch": live, "ops_per_test": OPS_PER_TEST,
                     "reference_pair_box_tests_per_launch": ref_tests,
                     "work_avoided": None if live is None else round(ref_tests / max(live, 1), 2),
                     "executed_valu_tops": None if executed is None else round(executed, 3),
                     "executed_valu_frac": None if executed is None
                     else round(executed / VALU_PEAK_TOPS, 4),
                     "hbm_frac": None if traffic is None
                     else round(traffic / (ms_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
                     "counters_from": prof.get("source")},
    }
    extras = {}
    cores = host_cores()
    if not a.no_cpu and kernel_name == "decode_kernel":
        threads = min(16, cores["usable"])
        log(f"cpu baseline on {threads} threads ({cores})")
        ins_per, seeds_per = (len(texts), len(head["seeds"]) // len(texts))
        line["cpu_baseline"] = cpu_baseline(texts, seeds_per, threads, a.workload)
        if line["cpu_baseline"] is not None:
            line["cpu_baseline"]["host"] = cores
        single = cpu_baseline(texts[:max(1, ins_per // 16)], seeds_per, 1, a.workload)
        if single is not None:
            extras["cpu_single_core"] = single
    eng.close()
    if not a.no_extras:
        for name in ("twenty_1024x4", "hundred_2048", "hundred_fit_2048", "mixed_256x64"):
            if name == a.workload:
                continue
            r = timed_workload(plp_amd, synth, ctx, name, 5 if name != "hundred_fit_2048" else 3,
                               1)
            extras[name] = {"packings_per_s": round(r["packings_per_s"], 1),
                            "ms_per_step": round(r["ms_per_step"], 3),
                            "kernel_ms": round(r["kernel_ms"], 3),
                            "launch": r["eng"].describe(), "autotuned": r["tuned"],
                            "parity": parity_note(r, pinned)}
            r["eng"].close()
        # Plain one-step-per-launch form: the host decode loop with one
        # feasibility launch per item step (the reference's CudaBackend).
        text = synth.multi_instance()
        inst = plp_amd.Instance.from_text(text)
        n_plain = 8
        t0 = time.perf_counter()
        for s_ in range(n_plain):
            ctx.plain_run(inst, 1 + s_, 1, prm)
        extras["plain_gpu_packings_per_s"] = round(n_plain / (time.perf_counter() - t0), 2)
        # Block-building engine (blockproof/plp_block.cpp) on the 20-type
        # instance, reference block options, 4 passes per decoder.
        r = timed_workload(plp_amd, synth, ctx, "blocks_twenty_1024x4", 5, 1)
        extras["block_engine_packings_per_s"] = round(r["packings_per_s"], 1)
        extras["block_engine_kernel_ms"] = round(r["kernel_ms"], 3)
        extras["block_engine_launch"] = r["eng"].describe()
        extras["block_engine_autotuned"] = r["tuned"]
        extras["block_engine_parity"] = parity_note(r, pinned)
        r["eng"].close()
        if not a.no_cpu:
            extras["block_engine_cpu_baseline"] = block_cpu_baseline(text, min(16, cores["usable"]))
        # JSON wire format: one 4096-decoder request through `plp_gpu --serve`;
        # format_ms is the host time to write its 4096 reference-format rows.
        cli = os.path.join(REPO, "plp-gpu_amd", "build", "plp_gpu")
        req = json.dumps({"id": 1, "instance_text": text, "decoders": 4096, "seed": 1,
                          "passes": 1, "eps": EPS, "theta": THETA})
        try:
            resp = json.loads(subprocess.run([cli, "--serve"], input=req + "\n", check=True,
                                             capture_output=True, text=True,
                                             timeout=300).stdout)
            extras["json_rows_per_s"] = round(4096 / (resp["format_ms"] * 1e-3), 1)
            extras["json_4096_rows_bytes"] = sum(len(json.dumps(r)) for r in resp["rows"])
        except Exception as e:  # noqa: BLE001 -- reported, not fatal
            log(f"serve wire-format probe failed: {e}")
        if a.serve_seconds > 0:
            extras["serving"] = serve(plp_amd, synth, a.serve_seconds, a.serve_rate, 1,
                                      a.serve_inflight)
    v = line["value"]
    if extras.get("plain_gpu_packings_per_s"):
        extras["x_plain_gpu"] = round(v / extras["plain_gpu_packings_per_s"], 1)
    if line.get("cpu_baseline"):
        extras["x_cpu_all_cores"] = round(v / line["cpu_baseline"]["value"], 1)
    if extras.get("cpu_single_core"):
        extras["x_cpu_single_core"] = round(v / extras["cpu_single_core"]["value"], 1)
    line["extras"] = extras
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
