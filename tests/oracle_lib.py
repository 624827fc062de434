"""ctypes access to the C oracle (oracle/build/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it and pin it
against the reference's fixtures.  Nothing in the product imports this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")
CLI = os.path.join(ORACLE_DIR, "build", "plp_oracle")


def ensure_built():
    if not (os.path.exists(LIB) and os.path.exists(CLI)):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)


class Item(C.Structure):
    _fields_ = [("l", C.c_int), ("w", C.c_int), ("h", C.c_int), ("lf", C.c_int),
                ("wf", C.c_int), ("hf", C.c_int), ("qty", C.c_int)]


class Inst(C.Structure):
    _fields_ = [("L", C.c_int), ("W", C.c_int), ("H", C.c_int), ("n_types", C.c_int),
                ("items", C.POINTER(Item))]


class Rng(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


class Work(C.Structure):
    _fields_ = [("pair_box", C.c_uint64), ("pairs", C.c_uint64), ("steps", C.c_int),
                ("max_cand", C.c_int)]


class Result(C.Structure):
    _fields_ = [("n_boxes", C.c_int), ("volume", C.c_longlong), ("score", C.c_double),
                ("fallback", C.c_int), ("n_passes", C.c_int), ("seed", C.c_uint64),
                ("placements", C.POINTER(C.c_int)), ("work", Work)]


class Oracle:
    def __init__(self, path: str = LIB):
        ensure_built()
        L = C.CDLL(path)
        self.L = L
        L.orc_parse.restype = C.c_int
        L.orc_parse.argtypes = [C.c_char_p, C.POINTER(C.POINTER(Inst)), C.c_char_p, C.c_int]
        L.orc_free_instances.argtypes = [C.POINTER(Inst), C.c_int]
        L.orc_orientations.restype = C.c_int
        L.orc_orientations.argtypes = [C.POINTER(Item), C.c_void_p]
        L.orc_item_sequence.restype = C.c_int
        L.orc_item_sequence.argtypes = [C.POINTER(Inst), C.c_int, C.c_void_p, C.c_int]
        L.orc_rng_seed.argtypes = [C.POINTER(Rng), C.c_uint64]
        L.orc_rng_next.restype = C.c_uint32
        L.orc_rng_next.argtypes = [C.POINTER(Rng)]
        L.orc_rng_random.restype = C.c_double
        L.orc_rng_random.argtypes = [C.POINTER(Rng)]
        L.orc_rng_getrandbits.restype = C.c_uint32
        L.orc_rng_getrandbits.argtypes = [C.POINTER(Rng), C.c_int]
        L.orc_rng_randrange.restype = C.c_int
        L.orc_rng_randrange.argtypes = [C.POINTER(Rng), C.c_int]
        L.orc_feasible_one.restype = C.c_int
        L.orc_feasible_one.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_check_all.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_decode_pass.restype = C.c_int
        L.orc_decode_pass.argtypes = [C.POINTER(Inst), C.c_int, C.c_int, C.c_double,
                                      C.POINTER(Rng), C.c_void_p, C.c_int, C.POINTER(Work)]
        L.orc_compute_fsr.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p]
        L.orc_stability_score.restype = C.c_double
        L.orc_stability_score.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(Inst), C.c_uint64, C.c_int, C.c_double, C.c_int, C.c_double,
                              C.c_int, C.c_int, C.c_double, C.POINTER(Result)]
        L.orc_py_repr.restype = C.c_int
        L.orc_py_repr.argtypes = [C.c_double, C.c_char_p, C.c_int]
        L.orc_block_run.restype = C.c_int
        L.orc_block_run.argtypes = [C.POINTER(Inst), C.c_uint64, C.c_int, C.c_double, C.c_int,
                                    C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.POINTER(Result)]

    # ---- parser -------------------------------------------------------
    def parse(self, text: str):
        """List of (L, W, H, [(l, lf, w, wf, h, hf, qty), ...]); raises ValueError."""
        p = C.POINTER(Inst)()
        err = C.create_string_buffer(256)
        n = self.L.orc_parse(text.encode(), C.byref(p), err, 256)
        if n < 0:
            raise ValueError(err.value.decode())
        out = []
        for i in range(n):
            I = p[i]
            items = [(I.items[t].l, I.items[t].lf, I.items[t].w, I.items[t].wf, I.items[t].h,
                      I.items[t].hf, I.items[t].qty) for t in range(I.n_types)]
            out.append((I.L, I.W, I.H, items))
        self.L.orc_free_instances(p, n)
        return out

    @staticmethod
    def make_inst(L, W, H, types):
        arr = (Item * max(1, len(types)))()
        for i, (l, lf, w, wf, h, hf, q) in enumerate(types):
            arr[i] = Item(l, w, h, lf, wf, hf, q)
        inst = Inst(L, W, H, len(types), arr)
        inst._keep = arr
        return inst

    def orientations(self, l, lf, w, wf, h, hf):
        out = (C.c_int * 18)()
        n = self.L.orc_orientations(C.byref(Item(l, w, h, lf, wf, hf, 1)), out)
        return [tuple(out[3 * i:3 * i + 3]) for i in range(n)]

    def item_sequence(self, inst_tuple, order="volume"):
        L, W, H, types = inst_tuple
        inst = self.make_inst(L, W, H, types)
        cap = max(1, sum(t[6] for t in types))
        buf = (C.c_int * cap)()
        n = self.L.orc_item_sequence(C.byref(inst), 0 if order == "volume" else 1, buf, cap)
        return list(buf[:n])

    # ---- rng ----------------------------------------------------------
    def rng(self, seed: int) -> Rng:
        r = Rng()
        self.L.orc_rng_seed(C.byref(r), seed)
        return r

    # ---- feasibility --------------------------------------------------
    def check_all(self, cand, ors, aabb, L, W, H) -> np.ndarray:
        cand = np.ascontiguousarray(np.asarray(cand, np.int32).reshape(-1, 3))
        ors = np.ascontiguousarray(np.asarray(ors, np.int32).reshape(-1, 3))
        aabb = np.ascontiguousarray(np.asarray(aabb, np.int32).reshape(-1, 6))
        bits = np.zeros(len(cand) * len(ors), np.uint8)
        self.L.orc_check_all(cand.ctypes.data, len(cand), ors.ctypes.data, len(ors),
                             aabb.ctypes.data if len(aabb) else None, len(aabb), L, W, H,
                             bits.ctypes.data)
        return bits

    # ---- FSR ----------------------------------------------------------
    def compute_fsr(self, placed7, mu=0.2) -> np.ndarray:
        """placed7: rows (x, y, z, w, d, h, type)."""
        a = np.ascontiguousarray(np.asarray(placed7, np.int32).reshape(-1, 7))
        out = np.zeros(len(a), np.float64)
        if len(a):
            self.L.orc_compute_fsr(a.ctypes.data, len(a), mu, out.ctypes.data)
        return out

    def stability_score(self, placed7, mu=0.2) -> float:
        a = np.ascontiguousarray(np.asarray(placed7, np.int32).reshape(-1, 7))
        return self.L.orc_stability_score(a.ctypes.data if len(a) else None, len(a), mu)

    # ---- run ----------------------------------------------------------
    def run(self, inst_tuple, seed=0, passes=1, eps=0.0, theta=None, order="volume",
            first_fit=False, mu=0.2):
        L, W, H, types = inst_tuple
        inst = self.make_inst(L, W, H, types)
        cap = max(1, sum(t[6] for t in types))
        pl = (C.c_int * (7 * cap))()
        res = Result()
        res.placements = C.cast(pl, C.POINTER(C.c_int))
        rc = self.L.orc_run(C.byref(inst), seed, passes, eps, 0 if theta is None else 1,
                            0.0 if theta is None else theta, 0 if order == "volume" else 1,
                            int(first_fit), mu, C.byref(res))
        assert rc == 0
        return {
            "n_boxes": res.n_boxes,
            "volume_cm3": res.volume,
            "stability_score": res.score,
            "fallback": bool(res.fallback),
            "n_passes": res.n_passes,
            "seed": res.seed,
            "placements": [list(pl[7 * i:7 * i + 7]) for i in range(res.n_boxes)],
            "pair_box": res.work.pair_box,
            "max_cand": res.work.max_cand,
        }

    PICK = {"first": 0, "support": 1, "contact": 2}

    def block_run(self, inst_tuple, seed=0, passes=1, eps=0.3, theta=0.9, amax=3, bmax=3, kmax=4,
                  topk=8, pick="first", mu=0.2):
        """Fixed-pass block-building engine (blockproof/plp_block.cpp)."""
        L, W, H, types = inst_tuple
        inst = self.make_inst(L, W, H, types)
        cap = max(1, sum(t[6] for t in types))
        pl = (C.c_int * (7 * cap))()
        res = Result()
        res.placements = C.cast(pl, C.POINTER(C.c_int))
        rc = self.L.orc_block_run(C.byref(inst), seed, passes, eps, 0 if theta is None else 1,
                                  0.0 if theta is None else theta, amax, bmax, kmax, topk,
                                  self.PICK[pick], mu, C.byref(res))
        assert rc == 0
        return {
            "n_boxes": res.n_boxes,
            "volume_cm3": res.volume,
            "stability_score": res.score,
            "fallback": bool(res.fallback),
            "n_passes": res.n_passes,
            "seed": res.seed,
            "placements": [list(pl[7 * i:7 * i + 7]) for i in range(res.n_boxes)],
        }

    def decode_pass(self, inst_tuple, order="volume", first_fit=False, eps=0.0, rng=None):
        L, W, H, types = inst_tuple
        inst = self.make_inst(L, W, H, types)
        cap = max(1, sum(t[6] for t in types))
        pl = np.zeros((cap, 7), np.int32)
        r = rng if rng is not None else self.rng(0)
        n = self.L.orc_decode_pass(C.byref(inst), 0 if order == "volume" else 1, int(first_fit),
                                   eps, C.byref(r), pl.ctypes.data, cap, None)
        return pl[:n]

    def py_repr(self, x: float) -> str:
        buf = C.create_string_buffer(64)
        self.L.orc_py_repr(x, buf, 64)
        return buf.value.decode()


def load_text(rel: str) -> str:
    with open(os.path.join(REPO, "tests", "golden", rel)) as f:
        return f.read()
