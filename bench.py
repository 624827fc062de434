#!/usr/bin/env python3
"""Decoded-packings/s benchmark of the MI355X pallet-loading decoder.

Workload (one GPU): the 20-type instance (587x233x220, dims U[25,115], stand
flags p=0.7, ~1.2x container volume) decoded by K = 1024 device-resident
decoders seeded 1..1024 with eps-greedy eps = 0.01, theta = 0.9; one step is
P = 4 passes of every decoder (4096 packings) with D4' selection on the
device.  Instance tables and decoder state are resident in HBM before the
timed region; every step resets the decoders, runs, synchronises and checks
the per-decoder error flags.

Printed: one JSON line with the contract fields plus
  roofline      dominant kernel (decode_kernel) against the VALU issue roof
  cpu_baseline  the reference's g++ build (or the C oracle) on host cores
  extras        plain one-step-per-launch GPU form, the 256-instance batch,
                the 100-type / 2048-decoder instance, the block-building engine
                (GPU and reference CPU), JSON emission, Poisson serving
usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "plp-gpu_amd"))

K_DEC = 1024
PASSES = 4
EPS, THETA = 0.01, 0.9
# Algorithmic unit: one pair x box test of the reference predicate
# (plp_common.hpp:297-321) over every (candidate, orientation, placed box) the
# reference's check_all receives (plp_common.hpp:418), at 10 lane-ops per test
# (C1 6 compares + ANDs; C3 1 compare + 2 min + 2 max + 2 sub + 1 mul + 1 add).
OPS_PER_TEST = 10
# MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz VALU issue = 39.3 T lane-ops/s
# (= FP64 vector 78.6 TFLOP/s counting an FMA as two flops); HBM3E 8 TB/s.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(text: str, threads: int, n: int, passes: int) -> dict | None:
    """The reference g++ build (oracle/_ref/ref_driver) or, when it was not
    built, the C oracle: n decoders x `passes` passes of the same workload
    (1024 x 12 on 16 threads and 64 x 4 on one are each about 5-10 s)."""
    inst_path = os.path.join(REPO, f"_bench_multi_{os.getpid()}.txt")
    with open(inst_path, "w") as f:
        f.write(text)
    ref = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    port = os.path.join(REPO, "oracle", "build", "plp_oracle")
    try:
        if os.path.exists(ref):
            cmd = [ref, "bench", inst_path, "--K", str(n), "--passes", str(passes), "--seed", "1",
                   "--eps", str(EPS), "--theta", str(THETA), "--threads", str(threads)]
            kind = "reference"
        else:
            if not os.path.exists(port):
                subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True,
                               capture_output=True)
            cmd = [port, inst_path, "--K", str(n), "--passes", str(passes), "--seed", "1", "--eps",
                   str(EPS), "--theta", str(THETA), "--threads", str(threads), "--bench"]
            kind = "port"
        out = json.loads(subprocess.run(cmd, check=True, capture_output=True, text=True,
                                        timeout=600).stdout)
    except Exception as e:  # noqa: BLE001 -- a missing CPU baseline is reported, not fatal
        log(f"cpu_baseline unavailable: {e}")
        return None
    finally:
        os.remove(inst_path)
    return {"value": round(out["packings_per_s"], 3), "unit": "packings/s", "cores": threads,
            "kind": kind,
            "sample": f"{n} decoders (seeds 1..{n}) x {passes} passes of the same 20-type "
                      f"instance, eps {EPS}, theta {THETA}; {out['seconds']:.2f} s wall on "
                      f"{threads} threads"}


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
    # warm every slot's buffers with the largest request shape
    big = max(range(len(arrivals)), key=lambda i: arrivals[i][2] * insts[i].n_items)
    warm = []
    for _ in range(max_inflight):
        rid = srv.submit(insts[big], arrivals[big][2], 1, passes, prm)
        if rid is not None:
            warm.append(rid)
    while warm:
        warm = [r for r in warm if srv.poll(r) is None]
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
    srv.close()
    lat.sort()
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))]  # noqa: E731
    return {"requests": len(lat), "rate_req_per_s": rate, "seconds": round(wall, 2),
            "passes_per_decoder": passes, "max_inflight": max_inflight,
            "packings_per_s": round(done_dec * passes / wall, 1),
            "p50_ms": round(1e3 * q(0.50), 2), "p99_ms": round(1e3 * q(0.99), 2),
            "max_ms": round(1e3 * lat[-1], 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape WPD,EP_CAP,BLOCKS_PER_CU instead of autotuning")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of the Poisson serving run in extras (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-inflight", type=int, default=8)
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

    text = synth.multi_instance()
    inst = plp_amd.Instance.from_text(text)
    ctx = plp_amd.Ctx(0)
    prm = plp_amd.params(eps=EPS, theta=THETA)
    if a.shape:
        wpd, cap, bpc = (int(v) for v in a.shape.split(","))
        eng = plp_amd.Engine(ctx, ep_cap=cap, blocks_per_cu=bpc, use_graph=True,
                             waves_per_decoder=wpd)
        eng.load([inst], [0] * K_DEC, list(range(1, K_DEC + 1)), prm)
        tuned = {"waves_per_decoder": wpd, "ep_cap": cap, "blocks_per_cu": bpc, "fixed": True}
    else:
        eng = plp_amd.Engine(ctx, use_graph=True)
        eng.load([inst], [0] * K_DEC, list(range(1, K_DEC + 1)), prm)
        log("tuning launch shape")
        tuned, tune_ms = eng.autotune(PASSES)
        log(f"autotune: {tuned} ({tune_ms:.2f} ms/pass)")
    for _ in range(a.warmup):
        eng.reset()
        eng.run(PASSES)
        eng.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.reset()
        eng.run(PASSES)
        eng.sync()
        kernel_ms.append(eng.last_ms())
    elapsed = time.perf_counter() - t0
    rows, _ = eng.fetch(placements=False)
    assert all(r.err == 0 for r in rows)
    shape = eng.describe()
    # The reference's pair x box tests of the same step, from one untimed run
    # of an engine that also keeps the dead extreme points (same results).
    ec = plp_amd.Engine(ctx, waves_per_decoder=shape["waves_per_decoder"], count_work=True)
    ec.load([inst], [0] * K_DEC, list(range(1, K_DEC + 1)), prm)
    ec.run(PASSES)
    ec.sync()
    crows, _ = ec.fetch(placements=False)
    ec.close()
    assert [(r.n_boxes, r.volume, r.score) for r in crows] == \
        [(r.n_boxes, r.volume, r.score) for r in rows]
    work = sum(r.work_pair_box for r in crows)
    live = sum(r.work_live_pair_box for r in crows)    # the tests this algorithm needs
    packings = K_DEC * PASSES * a.steps
    value = packings / elapsed
    ms_launch = sum(kernel_ms) / len(kernel_ms)
    achieved = work * OPS_PER_TEST / (ms_launch * 1e-3) / 1e12
    # Per-launch hardware counters of the same launch shape, from the
    # committed rocprofv3 summary (FETCH_SIZE / WRITE_SIZE scaled by the copy
    # calibration; SQ_INSTS_VALU wave instructions x 64 lanes).
    prof = {}
    pfile = os.path.join(REPO, "profiles", "decode_kernel_counters.json")
    if os.path.exists(pfile):
        with open(pfile) as f:
            prof = json.load(f)
    traffic = prof.get("hbm_bytes_per_launch")
    executed = None
    if prof.get("valu_insts_per_launch"):
        executed = prof["valu_insts_per_launch"] * 64 / (ms_launch * 1e-3) / 1e12
    line = {
        "metric": "decoded packings/s at bit-parity vs plain GPU & g++ host; p50/p99 latency",
        "value": round(value, 1),
        "unit": "packings/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+i32",
        "data": "synthetic",
        "config": {"workload": "20-type instance 587x233x220, K=1024 decoders seeds 1..1024, "
                               "eps 0.01, theta 0.9, 4 passes per step",
                   "decoders": K_DEC, "passes_per_step": PASSES, "parallelism": "none (1 GPU)",
                   "launch": shape, "autotuned": tuned},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                     "unit": "Tlane-op/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": traffic,
                     "kernel": "decode_kernel", "kernel_ms": round(ms_launch, 3),
                     "reference_pair_box_tests_per_launch": work, "ops_per_test": OPS_PER_TEST,
                     "live_pair_box_tests_per_launch": live,
                     "executed_valu_tops": None if executed is None else round(executed, 3),
                     "executed_valu_frac": None if executed is None
                     else round(executed / VALU_PEAK_TOPS, 4),
                     "hbm_frac": None if traffic is None
                     else round(traffic / (ms_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
                     "counters_from": prof.get("source")},
    }
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(text, threads, n=K_DEC, passes=12)
        single = cpu_baseline(text, 1, n=64, passes=4)
        if single is not None:
            line.setdefault("extras", {})["cpu_single_core"] = single
    if not a.no_extras:
        extras = line.get("extras", {})
        # Plain one-step-per-launch form, one decoder per launch sequence.
        n_plain = 8
        t0 = time.perf_counter()
        for s in range(n_plain):
            ctx.plain_run(inst, 1 + s, 1, prm)
        extras["plain_gpu_packings_per_s"] = round(n_plain / (time.perf_counter() - t0), 2)
        # 256 uneven instances x 16 seeds in one persistent launch.
        texts = synth.mixed_batch(256)
        insts = [plp_amd.Instance.from_text(t) for t in texts]
        e2 = plp_amd.Engine(ctx, use_graph=True)
        e2.load(insts, [i for i in range(256) for _ in range(16)],
                [1 + s for _ in range(256) for s in range(16)], prm)
        extras["batch_256x16_autotuned"], _ = e2.autotune(1)
        e2.run(1)
        e2.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            e2.reset()
            e2.run(1)
            e2.sync()
        extras["batch_256x16_packings_per_s"] = round(3 * 4096 / (time.perf_counter() - t0), 1)
        extras["batch_256x16_kernel_ms"] = round(e2.last_ms(), 3)
        e2.close()
        # 100 types, ~600 boxes, 2048 decoders, one pass.
        e3 = plp_amd.Engine(ctx, use_graph=True)
        e3.load([plp_amd.Instance.from_text(synth.large_instance())], [0] * 2048,
                list(range(1, 2049)), prm)
        extras["large_100types_2048_autotuned"], _ = e3.autotune(1)
        e3.run(1)
        e3.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            e3.reset()
            e3.run(1)
            e3.sync()
        extras["large_100types_2048_packings_per_s"] = round(3 * 2048 / (time.perf_counter() - t0), 1)
        extras["large_100types_2048_kernel_ms"] = round(e3.last_ms(), 3)
        e3.close()
        # Block-building engine (blockproof/plp_block.cpp) on the same instance
        # and seeds, reference block options, 4 passes per decoder.
        eb = plp_amd.Engine(ctx, use_graph=True)
        eb.load([inst], [0] * K_DEC, list(range(1, K_DEC + 1)), prm,
                blocks=plp_amd.block_params())
        eb.run(PASSES)
        eb.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            eb.reset()
            eb.run(PASSES)
            eb.sync()
        extras["block_engine_packings_per_s"] = round(3 * K_DEC * PASSES /
                                                      (time.perf_counter() - t0), 1)
        extras["block_engine_kernel_ms"] = round(eb.last_ms(), 3)
        eb.close()
        if not a.no_cpu:
            extras["block_engine_cpu_baseline"] = block_cpu_baseline(text, min(16, os.cpu_count()
                                                                              or 1))
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
        line["extras"] = extras
    print(json.dumps(line), flush=True)
    eng.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
