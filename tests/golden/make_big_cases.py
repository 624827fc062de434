#!/usr/bin/env python3
"""Write tests/golden/big_cases.json: instances whose containers need more
than the 31 bits of a 32-bit packed extreme-point key, with the reference
CLI's output on each (run in the build container).

The reference's host build (gpu/plp_gpu.cpp, compiled from its sources into
oracle/_ref/plp_gpu_ref by `make -C oracle ref`) decodes any int container;
its --device form takes sides up to 2^21 - 1 (gpu/plp_gpu.cu:1263-1268).
Cases:
  cube2000     2000 x 2000 x 2000, 12 types with sides U[200, 700]
  br1_mm       BR1 in millimetres: every dimension of BR1_050 times 10
  hc40_mm      BR1's boxes in mm in a 12032 x 2352 x 2698 mm container
  side_max     2097151 x 1000 x 1000 (the largest side the device keys hold)
  side_over    2097152 x 1000 x 1000 (one past it: the reference's --device
               exits 2, its host run decodes)
For each: `--single-pass --eps 0 --no-theta --json` and
`--single-pass --eps 0.3 --seed 5 --json` (theta 0.9).  Footprints stay below
2^31, where the reference's int product w * d is exact.

usage: python3 tests/golden/make_big_cases.py
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_CLI = os.path.join(REPO, "oracle", "_ref", "plp_gpu_ref")
RUNS = {"eps0": ["--single-pass", "--eps", "0", "--no-theta", "--json"],
        "eps03_seed5": ["--single-pass", "--eps", "0.3", "--seed", "5", "--json"]}


def text_of(L, W, H, types, pid=1) -> str:
    lines = ["1", f"1 {pid}", f"{L} {W} {H}", str(len(types))]
    lines += [f"{i} {l} {lf} {w} {wf} {h} {hf} {q}" for i, (l, lf, w, wf, h, hf, q)
              in enumerate(types, 1)]
    return "\n".join(lines) + "\n"


def random_types(seed, n, lo, hi, L, W, H, ratio):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        l, w, h = (rng.randint(lo, hi) for _ in range(3))
        q = max(1, round(ratio * L * W * H / (n * l * w * h)))
        out.append((l, 1, w, 1, h, 1, q))
    return out


def br1_types():
    with open(os.path.join(HERE, "data", "instances", "BR1_050.txt")) as f:
        toks = [ln.split() for ln in f if ln.strip()]
    n = int(toks[3][0])
    return [int(t) for t in toks[2]], [tuple(int(v) for v in toks[4 + i][1:8]) for i in range(n)]


def cases():
    (L, W, H), br1 = br1_types()
    mm = [(l * 10, lf, w * 10, wf, h * 10, hf, q) for (l, lf, w, wf, h, hf, q) in br1]
    yield "cube2000", text_of(2000, 2000, 2000, random_types(2032, 12, 200, 700, 2000, 2000, 2000,
                                                             1.2))
    yield "br1_mm", text_of(L * 10, W * 10, H * 10, mm)
    yield "hc40_mm", text_of(12032, 2352, 2698, mm)
    long_types = [(random.Random(2033 + i).randint(200, 600), 1, 500, 1, 400, 1, 6)
                  for i in range(20)]
    yield "side_max", text_of(2097151, 1000, 1000, long_types)
    yield "side_over", text_of(2097152, 1000, 1000, long_types)


def main() -> None:
    out = []
    for name, text in cases():
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(text)
            path = f.name
        runs = {}
        for key, args in RUNS.items():
            p = subprocess.run([REF_CLI, path, *args], check=True, capture_output=True, text=True,
                               timeout=600)
            runs[key] = json.loads(p.stdout)
            runs[key]["time_s"] = 0.0
            print(name, key, runs[key]["n_boxes"], runs[key]["stability_score"])
        os.remove(path)
        out.append({"name": name, "text": text, "runs": runs})
    with open(os.path.join(HERE, "big_cases.json"), "w") as f:
        json.dump({"source": "reference CLI gpu/plp_gpu.cpp (g++ build)", "args": RUNS,
                   "cases": out}, f)
        f.write("\n")


if __name__ == "__main__":
    main()
