#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ (run in the build container).

Sources, none of which travel as code:
  * data files the reference's own tests use: data/instances/*.txt,
    data/synthetic/SYN1.txt, data/manifest.json, and the checked-in parity
    outputs parity/ref_{cpu,host,gpu}/*.json (copied verbatim as data);
  * the reference's g++ host build and fixed-pass harness compiled from its
    sources by oracle/Makefile into oracle/_ref/ (run here, outputs saved);
  * CPython's own stdlib `random` and `repr` (the MT19937 and float-format
    contract, gpu/make_selftest_ref.py protocol and beyond);
  * the project's synthetic generator (plp-gpu_amd/plp_amd/synth.py).

usage: python3 tests/golden/make_golden.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import json
import os
import random
import shutil
import struct
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))

def _load_synth():
    import importlib.util
    path = os.path.join(REPO, "plp-gpu_amd", "plp_amd", "synth.py")
    spec = importlib.util.spec_from_file_location("plp_synth", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


synth = _load_synth()   # the pure-Python generator only; loads no library

REF_DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")
REF_CLI = os.path.join(REPO, "oracle", "_ref", "plp_gpu_ref")


def hexd(x: float) -> str:
    return format(struct.unpack("<Q", struct.pack("<d", x))[0], "016x")


def copy_reference_data(ref: str) -> None:
    dst = os.path.join(HERE, "data")
    for sub in ("instances", "synthetic"):
        os.makedirs(os.path.join(dst, sub), exist_ok=True)
        for f in sorted(os.listdir(os.path.join(ref, "data", sub))):
            shutil.copyfile(os.path.join(ref, "data", sub, f), os.path.join(dst, sub, f))
    shutil.copyfile(os.path.join(ref, "data", "manifest.json"), os.path.join(dst, "manifest.json"))
    for kind in ("ref_cpu", "ref_host", "ref_gpu"):
        d = os.path.join(HERE, "parity", kind)
        os.makedirs(d, exist_ok=True)
        for f in sorted(os.listdir(os.path.join(ref, "parity", kind))):
            shutil.copyfile(os.path.join(ref, "parity", kind, f), os.path.join(d, f))


def small_instance(L, W, H, types) -> str:
    lines = ["1", "1 0", f"{L} {W} {H}", str(len(types))]
    for i, t in enumerate(types, 1):
        lines.append(f"{i} " + " ".join(str(v) for v in t))
    return "\n".join(lines) + "\n"


def write_synthetic() -> dict:
    d = os.path.join(HERE, "synth")
    os.makedirs(d, exist_ok=True)
    files = {
        "plain": synth.plain_instance(),
        "multi": synth.multi_instance(),
        "large": synth.large_instance(),
    }
    for i, text in enumerate(synth.mixed_batch(16)):
        files[f"mixed_{i:03d}"] = text
    # Hand-sized cases of tests/test_solver.py:182-278, as instance files:
    # item(l, w, h, qty) has every stand flag set.
    def it(l, w, h, q):
        return (l, 1, w, 1, h, 1, q)
    files["tower10"] = small_instance(10, 10, 100, [it(10, 10, 10, 10)])
    files["tower9"] = small_instance(10, 10, 90, [it(10, 10, 10, 9)])
    files["side_by_side"] = small_instance(100, 100, 100, [it(60, 50, 50, 2)])
    files["stack3"] = small_instance(10, 10, 30, [it(10, 10, 10, 3)])
    files["skip"] = small_instance(20, 20, 20, [it(15, 15, 15, 1), it(25, 10, 10, 1)])
    files["vol_vs_input"] = small_instance(100, 100, 100, [it(10, 10, 10, 1), it(50, 50, 50, 1)])
    files["two_types"] = small_instance(100, 100, 100, [it(60, 50, 50, 2), it(30, 30, 30, 4)])
    files["multi_pass"] = small_instance(60, 60, 60, [it(15, 15, 15, 16), it(12, 12, 12, 12)])
    files["no_orient"] = small_instance(50, 50, 50, [(10, 0, 20, 0, 30, 0, 3), it(10, 10, 10, 2)])
    files["empty_qty"] = small_instance(50, 50, 50, [it(10, 10, 10, 0)])
    files["too_big"] = small_instance(20, 20, 20, [it(30, 30, 30, 2)])
    for name, text in files.items():
        with open(os.path.join(d, name + ".txt"), "w") as f:
            f.write(text)
    return {k: os.path.join("synth", k + ".txt") for k in files}


def ref_run(inst_rel: str, seed: int, passes: int, eps: float, theta, order="volume",
            first_fit=False, mu=None) -> dict:
    args = [REF_DRIVER, "run", os.path.join(HERE, inst_rel), "--seed", str(seed), "--passes",
            str(passes), "--eps", repr(eps), "--order", order, "--fsr", "--work"]
    args += ["--no-theta"] if theta is None else ["--theta", repr(theta)]
    if first_fit:
        args.append("--first-fit")
    if mu is not None:
        args += ["--mu", repr(mu)]
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def write_ref_runs(synth_files: dict) -> None:
    cases = []

    def add(name, inst_rel, seed=0, passes=1, eps=0.0, theta=None, order="volume",
            first_fit=False, mu=None):
        res = ref_run(inst_rel, seed, passes, eps, theta, order, first_fit, mu)
        cases.append({"case": name, "instance": inst_rel,
                      "args": {"seed": seed, "passes": passes, "eps": eps, "theta": theta,
                               "order": order, "first_fit": first_fit,
                               "mu": 0.2 if mu is None else mu},
                      "result": res})
        print(name, res["n_boxes"], res["volume_cm3"], res["stability_score"], flush=True)

    br = {f"BR{i}": f"data/instances/BR{i}_050.txt" for i in range(19)}
    br["SYN1"] = "data/synthetic/SYN1.txt"
    for name, rel in br.items():
        add(f"{name}_eps0", rel)
    for name in ("BR1", "BR2", "BR3", "SYN1"):
        for seed in (0, 1, 7):
            add(f"{name}_eps03_s{seed}", br[name], seed=seed, eps=0.3, theta=0.9)
    for name in ("BR1", "BR2", "BR3"):
        add(f"{name}_p4_theta", br[name], seed=3, passes=4, eps=0.3, theta=0.9)
        add(f"{name}_p3_notheta", br[name], seed=5, passes=3, eps=0.3, theta=None)
    add("BR2_mu05_p3", br["BR2"], seed=2, passes=3, eps=0.3, theta=0.5, mu=0.5)
    add("BR4_firstfit_p2", br["BR4"], seed=9, passes=2, eps=0.3, theta=0.9, first_fit=True)
    add("BR5_input_eps0", br["BR5"], order="input")
    add("BR6_input_eps03_p2", br["BR6"], seed=11, passes=2, eps=0.3, theta=0.9, order="input")
    add("BR7_eps1_p2", br["BR7"], seed=4, passes=2, eps=1.0, theta=0.9)
    add("SYN1_bigseed", br["SYN1"], seed=2**32 + 12345, passes=2, eps=0.5, theta=0.9)
    for name in ("tower10", "tower9", "side_by_side", "stack3", "skip", "vol_vs_input",
                 "two_types", "no_orient", "empty_qty", "too_big"):
        add(f"{name}_eps0", synth_files[name], theta=0.9)
    add("vol_vs_input_input", synth_files["vol_vs_input"], order="input")
    add("two_types_eps03_s42", synth_files["two_types"], seed=42, eps=0.3, theta=0.9)
    add("multi_pass_p6", synth_files["multi_pass"], seed=7, passes=6, eps=0.3, theta=0.9)
    add("plain_c1", synth_files["plain"], seed=1, theta=None)
    for seed in range(1, 9):
        add(f"multi_c2_s{seed}", synth_files["multi"], seed=seed, passes=2, eps=0.01, theta=0.9)
    for i in range(16):
        for seed in (1, 2):
            add(f"mixed_{i:03d}_s{seed}", synth_files[f"mixed_{i:03d}"], seed=seed, eps=0.01,
                theta=0.9)
    add("large_c4_s1", synth_files["large"], seed=1, eps=0.01, theta=0.9)
    with open(os.path.join(HERE, "ref_runs.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
        f.write("\n")


def write_rng_vectors() -> None:
    seeds = [0, 1, 7, 123456789, 2**31 - 1, 2**32 - 1, 2**32, 2**32 + 5, 2**40 + 3, 2**63 + 11]
    out = {"selftest": [], "scripts": []}
    for s in (0, 1, 7, 123456789):   # gpu/make_selftest_ref.py:15-35 protocol
        r = random.Random(s)
        out["selftest"].append({"seed": s,
                                "random_hex": [hexd(r.random()) for _ in range(16)],
                                "getrandbits13": [r.getrandbits(13) for _ in range(16)]})
    # Mixed scripts: random() / getrandbits(k) / randrange(n), long enough to
    # cross several 624-word regenerations.
    for s in seeds:
        rr = random.Random(s ^ 0xA5A5)
        ops, args, vals = [], [], []
        r = random.Random(s)
        for _ in range(1500):
            op = rr.choice((0, 0, 1, 2, 2))
            if op == 0:
                ops.append(0); args.append(0); vals.append(hexd(r.random()))
            elif op == 1:
                k = rr.randint(1, 32)
                ops.append(1); args.append(k); vals.append(r.getrandbits(k))
            else:
                n = rr.choice((1, 2, 3, 5, 63, 64, 65, 100, 1000, 4096, 12345, 2**31 - 1))
                ops.append(2); args.append(n); vals.append(r.randrange(n))
        out["scripts"].append({"seed": s, "ops": ops, "args": args, "vals": vals})
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


def write_repr_vectors() -> None:
    vals = [0.0, -0.0, 1.0, 0.5, 0.1, 2.0 / 3.0, 0.5785123966942148, 0.6132075471698113,
            0.6595744680851063, 1e-05, 1e-04, 1e15, 1e16, 0.001, 123456789.5, 1.0 / 121.0,
            1.0 / 106.0, 1e22, 1e-7, 5e-324, 1.7976931348623157e308, 2.0 ** -1074, 9007199254740993.0,
            0.3, 1.0 / 3.0, 123.456, 1e300, 2.5e-5, 0.0001, 0.00001]
    for n in range(1, 701):   # every stability score k/n an n <= 700 packing can have
        for k in sorted({0, 1, n // 3, n // 2, (2 * n) // 3, n - 1, n}):
            vals.append(k / n)
    r = random.Random(12345)
    for _ in range(3000):
        vals.append(struct.unpack("<d", struct.pack("<Q", r.getrandbits(64)))[0])
    for _ in range(1000):
        vals.append(r.random() * 10 ** r.randint(-20, 20))
    out = []
    for v in vals:
        if v != v or v in (float("inf"), float("-inf")):
            continue
        out.append([hexd(v), repr(v)])
    with open(os.path.join(HERE, "repr_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


BLOCK_DRIVER = os.path.join(REPO, "oracle", "_ref", "block_driver")
BLOCK_CLI = os.path.join(REPO, "oracle", "_ref", "plp_block_ref")


def block_args(seed, passes, eps, theta, pick, amax, bmax, kmax, topk):
    a = ["--seed", str(seed), "--eps", repr(eps), "--pick", pick, "--amax", str(amax),
         "--bmax", str(bmax), "--kmax", str(kmax), "--topk", str(topk)]
    return a + (["--no-theta"] if theta is None else ["--theta", repr(theta)])


def write_block_runs(synth_files: dict) -> None:
    """Fixed-pass runs of the reference block engine (blockproof/plp_block.cpp)
    through oracle/block_driver, plus single-pass runs of the reference's own
    CLI that pin the driver."""
    cases = []

    def add(name, inst_rel, seed=0, passes=1, eps=0.3, theta=0.9, pick="first", amax=3, bmax=3,
            kmax=4, topk=8):
        args = block_args(seed, passes, eps, theta, pick, amax, bmax, kmax, topk)
        out = subprocess.run([BLOCK_DRIVER, "run", os.path.join(HERE, inst_rel), "--passes",
                              str(passes)] + args, check=True, capture_output=True, text=True)
        res = json.loads(out.stdout)
        cli = None
        if passes == 1 and seed < 2**31:
            c = subprocess.run([BLOCK_CLI, os.path.join(HERE, inst_rel), "--single-pass", "--json"]
                               + args, check=True, capture_output=True, text=True)
            cli = json.loads(c.stdout)
        cases.append({"case": name, "instance": inst_rel,
                      "args": {"seed": seed, "passes": passes, "eps": eps, "theta": theta,
                               "pick": pick, "amax": amax, "bmax": bmax, "kmax": kmax,
                               "topk": topk},
                      "result": res, "reference_cli": cli})
        print(name, res["n_boxes"], res["volume_cm3"], res["stability_score"], flush=True)

    br = {f"BR{i}": f"data/instances/BR{i}_050.txt" for i in range(19)}
    br["SYN1"] = "data/synthetic/SYN1.txt"
    for name, rel in br.items():
        add(f"blk_{name}_s0", rel)
    for name in ("BR1", "BR2", "BR3"):
        for pick in ("first", "support", "contact"):
            for seed in (1, 2):
                add(f"blk_{name}_{pick}_s{seed}_p3", br[name], seed=seed, passes=3, pick=pick)
        add(f"blk_{name}_eps0", br[name], eps=0.0)
        add(f"blk_{name}_eps0_support", br[name], eps=0.0, pick="support")
        add(f"blk_{name}_notheta_p2", br[name], seed=4, passes=2, theta=None)
    add("blk_BR4_single_items", br["BR4"], seed=3, amax=1, bmax=1, kmax=1)
    add("blk_BR5_222_p2", br["BR5"], seed=5, passes=2, amax=2, bmax=2, kmax=2)
    add("blk_BR6_446", br["BR6"], seed=6, amax=4, bmax=4, kmax=6, pick="contact")
    add("blk_BR7_topk1", br["BR7"], seed=7, topk=1)
    add("blk_BR8_topk3_eps1", br["BR8"], seed=8, eps=1.0, topk=3)
    add("blk_BR9_theta05_p3", br["BR9"], seed=9, passes=3, theta=0.5)
    add("blk_SYN1_bigseed", br["SYN1"], seed=2**32 + 99, passes=2)
    for name in ("tower10", "side_by_side", "stack3", "skip", "two_types", "no_orient",
                 "empty_qty", "too_big"):
        add(f"blk_{name}", synth_files[name])
    for seed in range(1, 5):
        add(f"blk_multi_s{seed}", synth_files["multi"], seed=seed, passes=2, eps=0.01)
    for i in range(6):
        add(f"blk_mixed_{i:03d}", synth_files[f"mixed_{i:03d}"], seed=i + 1, passes=2, eps=0.01,
            pick=("first", "support", "contact")[i % 3])
    add("blk_large_s1", synth_files["large"], seed=1, eps=0.01)
    with open(os.path.join(HERE, "block_runs.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
        f.write("\n")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--block-only", action="store_true",
                    help="only rewrite block_runs.json (the other fixtures stay as they are)")
    a = ap.parse_args()
    if not (os.path.exists(REF_DRIVER) and os.path.exists(REF_CLI)):
        sys.exit("build the reference harness first: make -C oracle ref")
    if a.block_only:
        synth_files = {os.path.splitext(n)[0]: os.path.join("synth", n)
                       for n in os.listdir(os.path.join(HERE, "synth"))}
        write_block_runs(synth_files)
        return
    copy_reference_data(a.ref)
    synth_files = write_synthetic()
    write_rng_vectors()
    write_repr_vectors()
    write_ref_runs(synth_files)
    write_block_runs(synth_files)


if __name__ == "__main__":
    main()
