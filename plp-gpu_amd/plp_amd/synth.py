"""Deterministic synthetic pallet-loading instances (thpack "original" text).

Workload families, in a 587 x 233 x 220 cm container unless stated:

  plain       3 types, dims uniform int [30, 120], every orientation allowed,
              qty so the total box volume is ~1.1x the container
  multi       20 types, dims uniform int [25, 115], each stand flag set with
              p = 0.7 (at least one), qty for ~1.2x the container
  mixed       n types (n uniform int [3, 100] per instance), dims uniform int
              [25, 115], every orientation allowed, ~1.2x the container
  large       100 types, dims log-uniform [10, 150], 6 copies each (~600 boxes)
  large_fit   the same 600 boxes in a container scaled 1.5x per side
              (880 x 349 x 330), so that every box can be placed and the
              extreme-point lists grow into the thousands

Quantity rule: qty_t = max(1, round(r * L*W*H / (n * vol_t))), giving every
type an equal share of r times the container volume.  The generator only uses
Python's random.Random with a fixed seed, so a (family, seed) pair always
produces the same text.
"""
from __future__ import annotations

import math
import random

CONTAINER = (587, 233, 220)


def _qty(ratio: float, n: int, vol: int) -> int:
    L, W, H = CONTAINER
    return max(1, int(round(ratio * L * W * H / (n * vol))))


def _text(types, seed: int, container=CONTAINER) -> str:
    L, W, H = container
    lines = ["1", f"1 {seed}", f"{L} {W} {H}", str(len(types))]
    for i, (l, lf, w, wf, h, hf, q) in enumerate(types, 1):
        lines.append(f"{i} {l} {lf} {w} {wf} {h} {hf} {q}")
    return "\n".join(lines) + "\n"


def uniform_instance(seed: int, n_types: int, lo: int, hi: int, ratio: float,
                     flag_p: float | None = None) -> str:
    rng = random.Random(seed)
    types = []
    for _ in range(n_types):
        l, w, h = (rng.randint(lo, hi) for _ in range(3))
        if flag_p is None:
            flags = (1, 1, 1)
        else:
            while True:
                flags = tuple(1 if rng.random() < flag_p else 0 for _ in range(3))
                if any(flags):
                    break
        types.append((l, flags[0], w, flags[1], h, flags[2], _qty(ratio, n_types, l * w * h)))
    return _text(types, seed)


def plain_instance(seed: int = 2027) -> str:
    return uniform_instance(seed, 3, 30, 120, 1.1)


def multi_instance(seed: int = 2028) -> str:
    return uniform_instance(seed, 20, 25, 115, 1.2, flag_p=0.7)


def mixed_instance(seed: int) -> str:
    n = random.Random(seed ^ 0x5EED).randint(3, 100)
    return uniform_instance(seed, n, 25, 115, 1.2)


def mixed_batch(n_instances: int = 256, seed0: int = 2029_000) -> list[str]:
    return [mixed_instance(seed0 + i) for i in range(n_instances)]


def large_instance(seed: int = 2030, n_types: int = 100, copies: int = 6) -> str:
    rng = random.Random(seed)
    types = []
    for _ in range(n_types):
        l, w, h = (int(round(math.exp(rng.uniform(math.log(10), math.log(150))))) for _ in range(3))
        types.append((l, 1, w, 1, h, 1, copies))
    return _text(types, seed)


def large_fit_instance(seed: int = 2030, scale: float = 1.5) -> str:
    """large_instance's 100 types x 6 copies in a container scaled by `scale`
    per side (880 x 349 x 330 at 1.5: the boxes fill 0.87 of it)."""
    _, _, _, types = parse_types(large_instance(seed))
    L, W, H = (int(v * scale) for v in CONTAINER)
    return _text(types, seed, (L, W, H))


# The benchmark workloads at their full sizes: instance texts, the instance
# of every decoder, its seed, passes per run, and eps / theta.
EPS, THETA = 0.01, 0.9


def workload(name: str):
    """(texts, dec_inst, seeds, passes) of a named full-size workload:
      twenty_1024x4     the 20-type instance, 1024 decoders seeds 1..1024, 4 passes
      mixed_256x16      256 instances of 3..100 types x 16 seeds (4096 decoders), 1 pass
      mixed_256x64      the same 256 instances x 64 seeds (16384 decoders), 1 pass: a queue
                        of eight rounds of the persistent grid
      hundred_2048      the 100-type instance (587x233x220), 2048 decoders, 1 pass
      hundred_fit_2048  the 100-type instance in 880x349x330, 2048 decoders, 1 pass
      blocks_twenty_1024x4  twenty_1024x4 decoded by the block-building engine
                        (reference block options, plp_amd.block_params())
    All with eps 0.01 and theta 0.9."""
    if name.startswith("blocks_"):
        return workload(name[len("blocks_"):])
    if name == "twenty_1024x4":
        return [multi_instance()], [0] * 1024, list(range(1, 1025)), 4
    if name in ("mixed_256x16", "mixed_256x64"):
        k = 16 if name == "mixed_256x16" else 64
        return (mixed_batch(256), [i for i in range(256) for _ in range(k)],
                [1 + s for _ in range(256) for s in range(k)], 1)
    if name == "hundred_2048":
        return [large_instance()], [0] * 2048, list(range(1, 2049)), 1
    if name == "hundred_fit_2048":
        return [large_fit_instance()], [0] * 2048, list(range(1, 2049)), 1
    raise KeyError(name)


WORKLOADS = ("twenty_1024x4", "mixed_256x16", "hundred_2048", "hundred_fit_2048",
             "blocks_twenty_1024x4", "mixed_256x64")


def is_block_workload(name: str) -> bool:
    return name.startswith("blocks_")


def parse_types(text: str):
    """(L, W, H, [(l, lf, w, wf, h, hf, qty), ...]) of a one-instance text."""
    toks = [ln.split() for ln in text.splitlines() if ln.strip()]
    L, W, H = (int(t) for t in toks[2])
    n = int(toks[3][0])
    types = [tuple(int(t) for t in toks[4 + i][1:8]) for i in range(n)]
    return L, W, H, types
