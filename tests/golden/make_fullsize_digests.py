#!/usr/bin/env python3
"""Write tests/golden/fullsize_digests.json: the result digest of every decoder
of the full-size benchmark workloads (run in the build container).

For each workload of plp_amd.synth.workload() every decoder is decoded twice
on the host, by the C oracle (oracle/build/liboracle.so) and by the
reference's own g++ build (oracle/_ref/ref_driver, compiled from
gpu/plp_common.hpp by `make -C oracle ref`; for the block-building workload
the block oracle and oracle/_ref/block_driver over blockproof/plp_block.cpp),
and the two must agree decoder for decoder.  The fixture keeps each decoder's digest (plp_amd.decoder_digest
of the reference JSON fields and rows, first 16 hex digits) and the digest of
the whole workload, which the GPU tests and bench.py compare the engine's
results with.

usage: python3 tests/golden/make_fullsize_digests.py [--threads 8] [--only NAME]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "plp-gpu_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

from plp_amd import decoder_digest, synth  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

REF_DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")
BLOCK_DRIVER = os.path.join(REPO, "oracle", "_ref", "block_driver")
OUT = os.path.join(HERE, "fullsize_digests.json")


def oracle_digests(orc, texts, dec_inst, seeds, passes, threads, blocks=False):
    insts = [orc.parse(t)[0] for t in texts]

    def one(k):
        run = orc.block_run if blocks else orc.run
        r = run(insts[dec_inst[k]], seed=seeds[k], passes=passes, eps=synth.EPS, theta=synth.THETA)
        return decoder_digest(r["n_boxes"], r["volume_cm3"], r["stability_score"], r["fallback"],
                              r["n_passes"], r["seed"], r["placements"])

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(len(seeds))))


def reference_digests(texts, dec_inst, seeds, passes, threads, blocks=False):
    """ref_driver rows over runs of consecutive seeds of one instance."""
    jobs = []   # (instance, first seed, count, first decoder)
    k = 0
    while k < len(seeds):
        j = k
        while (j + 1 < len(seeds) and dec_inst[j + 1] == dec_inst[k]
               and seeds[j + 1] == seeds[j] + 1 and j + 1 - k < 64):
            j += 1
        jobs.append((dec_inst[k], seeds[k], j - k + 1, k))
        k = j + 1
    tmp = tempfile.mkdtemp()
    paths = []
    for i, t in enumerate(texts):
        paths.append(os.path.join(tmp, f"i{i}.txt"))
        with open(paths[-1], "w") as f:
            f.write(t)
    out = [None] * len(seeds)

    def run(job):
        inst, s0, n, k0 = job
        p = subprocess.run([BLOCK_DRIVER if blocks else REF_DRIVER, "rows", paths[inst], "--seed",
                            str(s0), "--K", str(n), "--passes", str(passes), "--eps",
                            str(synth.EPS), "--theta", str(synth.THETA)],
                           check=True, capture_output=True, text=True)
        for i, line in enumerate(p.stdout.splitlines()):
            r = json.loads(line)
            out[k0 + i] = decoder_digest(r["n_boxes"], r["volume_cm3"], r["stability_score"],
                                         r["fallback"], r["n_passes"], r["seed"],
                                         [[q[k] for k in ("item_type", "x", "y", "z", "w", "d",
                                                          "h")] if isinstance(q, dict) else q
                                          for q in r["placements"]])

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, jobs))
    for pth in paths:
        os.remove(pth)
    os.rmdir(tmp)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    orc = Oracle()
    fixture = {"workloads": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            fixture = json.load(f)
    for name in synth.WORKLOADS:
        if a.only and name != a.only:
            continue
        texts, dec_inst, seeds, passes = synth.workload(name)
        blocks = synth.is_block_workload(name)
        t0 = time.time()
        mine = oracle_digests(orc, texts, dec_inst, seeds, passes, a.threads, blocks)
        t1 = time.time()
        ref = reference_digests(texts, dec_inst, seeds, passes, a.threads, blocks)
        t2 = time.time()
        bad = [k for k in range(len(seeds)) if mine[k] != ref[k]]
        print(f"{name}: {len(seeds)} decoders, oracle {t1 - t0:.1f} s, reference {t2 - t1:.1f} s, "
              f"{len(bad)} differ", flush=True)
        assert not bad, bad[:10]
        fixture["workloads"][name] = {
            "decoders": len(seeds),
            "passes": passes,
            "digest": hashlib.sha256(b"".join(mine)).hexdigest(),
            "per_decoder": [d.hex()[:16] for d in mine],
            "source": ("C block oracle and reference g++ block engine (block_driver rows)"
                       if blocks else "C oracle and reference g++ build (ref_driver rows)")
                      + ", identical per decoder",
        }
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
