"""Python binding over the C ABI of libplp_hip.so (include/plp_hip.h).

A thin ctypes layer for scripting, tests and benchmarks; every computation
happens in the HIP library.  Importing fails loudly when the built library is
missing -- there is no fallback path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PLP_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "build",
                                                          "libplp_hip.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "build", "plp_gpu")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libplp_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")

_lib = C.CDLL(LIB_PATH)

PLP_OK, PLP_E_ARG, PLP_E_HIP, PLP_E_NOMEM, PLP_E_EPCAP, PLP_E_RANGE, PLP_E_FORMAT, PLP_E_BUSY = range(8)

EXPORTS = [
    "plp_status_string", "plp_last_error", "plp_device_count", "plp_parse_instance",
    "plp_ctx_create", "plp_ctx_destroy", "plp_feas_check_all", "plp_feas_check_all_noalloc",
    "plp_feas_check_batch",
    "plp_fsr_eval", "plp_rng_script", "plp_plain_run", "plp_engine_create",
    "plp_engine_destroy", "plp_engine_load", "plp_engine_load_blocks", "plp_engine_reset", "plp_engine_run", "plp_engine_sync",
    "plp_engine_max_items", "plp_engine_fetch", "plp_engine_sync_fetch", "plp_engine_results",
    "plp_engine_last_ms",
    "plp_engine_describe",
    "plp_engine_autotune", "plp_server_create", "plp_server_destroy", "plp_server_submit",
    "plp_server_poll", "plp_server_inflight", "plp_server_query", "plp_server_stats",
]


class PlpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class InstanceDesc(C.Structure):
    _fields_ = [("L", C.c_int32), ("W", C.c_int32), ("H", C.c_int32),
                ("n_types", C.c_int32), ("types", C.POINTER(C.c_int32))]


class Params(C.Structure):
    _fields_ = [("eps", C.c_double), ("theta", C.c_double), ("mu", C.c_double),
                ("use_theta", C.c_int32), ("first_fit", C.c_int32), ("order", C.c_int32),
                ("reserved", C.c_int32)]


class EngineConfig(C.Structure):
    _fields_ = [("waves_per_decoder", C.c_int32), ("ep_cap", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("use_graph", C.c_int32), ("count_work", C.c_int32),
                ("box_cap", C.c_int32), ("verdicts", C.c_int32), ("share_prefix", C.c_int32),
                ("trace_waves", C.c_int32), ("trace_verdicts", C.c_int32)]


VERDICTS = {"auto": 0, "pair_cache": 1, "point_state": 2}
SHARE = {"auto": 0, "off": 1, "on": 2}


class BlockParams(C.Structure):
    _fields_ = [("amax", C.c_int32), ("bmax", C.c_int32), ("kmax", C.c_int32),
                ("topk", C.c_int32), ("pick", C.c_int32), ("reserved", C.c_int32 * 3)]


PICK = {"first": 0, "support": 1, "contact": 2}


def block_params(amax=3, bmax=3, kmax=4, topk=8, pick="first") -> BlockParams:
    """Block-engine options (reference defaults, blockproof/plp_block.cpp:297-300)."""
    return BlockParams(amax, bmax, kmax, topk, PICK[pick] if isinstance(pick, str) else pick)


class RowResult(C.Structure):
    _fields_ = [("n_boxes", C.c_int32), ("fallback", C.c_int32), ("n_passes", C.c_int32),
                ("err", C.c_int32), ("volume", C.c_int64), ("score", C.c_double),
                ("seed", C.c_uint64), ("work_pair_box", C.c_uint64),
                ("work_live_pair_box", C.c_uint64), ("max_ep", C.c_int32), ("ep_need", C.c_int32)]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_sig = {
    "plp_status_string": (C.c_char_p, [C.c_int]),
    "plp_last_error": (C.c_char_p, []),
    "plp_device_count": (C.c_int, [_I32P]),
    "plp_parse_instance": (C.c_int, [C.c_char_p, _I32P, _I32P, _I32P, _I32P, _I32P, C.c_int32,
                                     C.c_char_p, C.c_int32]),
    "plp_ctx_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "plp_ctx_destroy": (None, [_P]),
    "plp_feas_check_all": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P, C.c_int32,
                                     C.c_int32, C.c_int32, C.c_int32, _P]),
    "plp_feas_check_all_noalloc": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_int32, _P]),
    "plp_feas_check_batch": (C.c_int, [_P, C.c_int32, _P, _P, _P, C.c_int32, _P, _P,
                                       C.c_int32, C.c_int32, C.c_int32, _P]),
    "plp_fsr_eval": (C.c_int, [_P, _P, _P, C.c_int32, C.c_double, _P, _P]),
    "plp_rng_script": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_int32, _P, _P]),
    "plp_plain_run": (C.c_int, [_P, C.POINTER(InstanceDesc), C.c_uint64, C.c_int32,
                                C.POINTER(Params), C.POINTER(RowResult), _P]),
    "plp_engine_create": (C.c_int, [_P, C.POINTER(EngineConfig), C.POINTER(_P)]),
    "plp_engine_destroy": (None, [_P]),
    "plp_engine_load": (C.c_int, [_P, C.POINTER(InstanceDesc), C.c_int32, _P, _P, C.c_int32,
                                  C.POINTER(Params)]),
    "plp_engine_load_blocks": (C.c_int, [_P, C.POINTER(InstanceDesc), C.c_int32, _P, _P,
                                         C.c_int32, C.POINTER(Params), C.POINTER(BlockParams)]),
    "plp_engine_reset": (C.c_int, [_P]),
    "plp_engine_run": (C.c_int, [_P, C.c_int32]),
    "plp_engine_sync": (C.c_int, [_P]),
    "plp_engine_max_items": (C.c_int, [_P, _I32P]),
    "plp_engine_fetch": (C.c_int, [_P, C.POINTER(RowResult), _P]),
    "plp_engine_sync_fetch": (C.c_int, [_P, C.POINTER(RowResult), _P]),
    "plp_engine_results": (C.c_int, [_P, C.POINTER(C.POINTER(RowResult)), C.POINTER(_I32P),
                                     _I32P]),
    "plp_engine_last_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "plp_engine_describe": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P]),
    "plp_engine_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(EngineConfig),
                                      C.POINTER(C.c_float)]),
    "plp_server_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, C.POINTER(_P)]),
    "plp_server_destroy": (None, [_P]),
    "plp_server_submit": (C.c_int, [_P, C.POINTER(InstanceDesc), C.c_int32, C.c_uint64,
                                    C.c_int32, C.POINTER(Params), C.POINTER(C.c_int64)]),
    "plp_server_poll": (C.c_int, [_P, C.c_int64, _I32P, C.POINTER(RowResult), _P, C.c_int32,
                                  _I32P]),
    "plp_server_inflight": (C.c_int, [_P, _I32P]),
    "plp_server_query": (C.c_int, [_P, C.c_int64, _I32P]),
    "plp_server_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _check(rc: int) -> int:
    if rc not in (PLP_OK,):
        raise PlpError(rc, _lib.plp_last_error().decode(errors="replace"))
    return rc


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def device_count() -> int:
    n = C.c_int32(0)
    _check(_lib.plp_device_count(C.byref(n)))
    return n.value


def parse_instance(text: str):
    """(L, W, H, types[n, 7]) of a one-instance thpack text."""
    L, W, H, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    err = C.create_string_buffer(256)
    cap = 4096
    types = np.zeros((cap, 7), np.int32)
    rc = _lib.plp_parse_instance(text.encode(), C.byref(L), C.byref(W), C.byref(H), C.byref(n),
                                 types.ctypes.data_as(_I32P), cap, err, 256)
    _check(rc)
    return L.value, W.value, H.value, types[: n.value].copy()


class Instance:
    """An instance as the C ABI sees it: container + item-type table."""

    def __init__(self, L: int, W: int, H: int, types):
        self.L, self.W, self.H = int(L), int(W), int(H)
        self.types = np.ascontiguousarray(np.asarray(types, np.int32).reshape(-1, 7))

    @classmethod
    def from_text(cls, text: str) -> "Instance":
        return cls(*parse_instance(text))

    @classmethod
    def from_file(cls, path: str) -> "Instance":
        with open(path) as f:
            return cls.from_text(f.read())

    @property
    def n_items(self) -> int:
        return int(self.types[:, 6].sum())

    def desc(self) -> InstanceDesc:
        return InstanceDesc(self.L, self.W, self.H, len(self.types),
                            self.types.ctypes.data_as(_I32P))


def params(eps=0.3, theta=0.9, mu=0.2, use_theta=True, first_fit=False, order="volume") -> Params:
    return Params(float(eps), float(theta), float(mu), int(bool(use_theta)), int(bool(first_fit)),
                  0 if order == "volume" else 1, 0)


# Contexts and servers still open when the interpreter exits are closed
# before the HIP runtime's own exit handlers run.
_OPEN = weakref.WeakSet()


@atexit.register
def _close_all():
    for obj in list(_OPEN):
        obj.close()


class Ctx:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()
        _check(_lib.plp_ctx_create(device, C.byref(self.h)))
        _OPEN.add(self)

    def close(self):
        # Engines borrow the context's stream: destroy them first.
        for child in list(self._children):
            child.close()
        if self.h:
            _lib.plp_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def feas_check_all(self, cand, ors, aabb, L, W, H) -> np.ndarray:
        cand = np.ascontiguousarray(np.asarray(cand, np.int32).reshape(-1, 3))
        ors = np.ascontiguousarray(np.asarray(ors, np.int32).reshape(-1, 3))
        aabb = np.ascontiguousarray(np.asarray(aabb, np.int32).reshape(-1, 6))
        bits = np.zeros(len(cand) * len(ors), np.uint8)
        _check(_lib.plp_feas_check_all(self.h, _ptr(cand), len(cand), _ptr(ors), len(ors),
                                       _ptr(aabb), len(aabb), L, W, H, _ptr(bits)))
        return bits

    def feas_check_all_noalloc(self, cand, ors, aabb, L, W, H, out=None) -> np.ndarray:
        """plp_feas_check_all_noalloc; out (uint8, n_cand * n_ors) may be given."""
        cand = np.ascontiguousarray(np.asarray(cand, np.int32).reshape(-1, 3))
        ors = np.ascontiguousarray(np.asarray(ors, np.int32).reshape(-1, 3))
        aabb = np.ascontiguousarray(np.asarray(aabb, np.int32).reshape(-1, 6))
        bits = np.zeros(len(cand) * len(ors), np.uint8) if out is None else out
        _check(_lib.plp_feas_check_all_noalloc(self.h, _ptr(cand), len(cand), _ptr(ors), len(ors),
                                               _ptr(aabb), len(aabb), L, W, H, _ptr(bits)))
        return bits

    def feas_check_batch(self, cands, ors, aabbs, L, W, H):
        K = len(cands)
        cands = [np.ascontiguousarray(np.asarray(c, np.int32).reshape(-1, 3)) for c in cands]
        aabbs = [np.ascontiguousarray(np.asarray(b, np.int32).reshape(-1, 6)) for b in aabbs]
        ors = np.ascontiguousarray(np.asarray(ors, np.int32).reshape(-1, 3))
        outs = [np.zeros(len(c) * len(ors), np.uint8) for c in cands]
        cp = (C.c_void_p * K)(*[_ptr(c) for c in cands])
        bp = (C.c_void_p * K)(*[_ptr(b) for b in aabbs])
        op = (C.c_void_p * K)(*[_ptr(o) for o in outs])
        nc = np.array([len(c) for c in cands], np.int32)
        nb = np.array([len(b) for b in aabbs], np.int32)
        _check(_lib.plp_feas_check_batch(self.h, K, C.cast(cp, C.c_void_p), _ptr(nc), _ptr(ors),
                                         len(ors), C.cast(bp, C.c_void_p), _ptr(nb), L, W, H,
                                         C.cast(op, C.c_void_p)))
        return outs

    def fsr_eval(self, packings, mu=0.2):
        """packings: list of (n, 7) placement-row arrays -> (per-box FSR list, scores)."""
        rows = [np.asarray(p, np.int32).reshape(-1, 7) for p in packings]
        off = np.zeros(len(rows) + 1, np.int32)
        off[1:] = np.cumsum([len(r) for r in rows])
        flat = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, 7), np.int32))
        fsr = np.zeros(int(off[-1]), np.float64)
        score = np.zeros(len(rows), np.float64)
        _check(_lib.plp_fsr_eval(self.h, _ptr(flat), _ptr(off), len(rows), float(mu), _ptr(fsr),
                                 _ptr(score)))
        return [fsr[off[i]:off[i + 1]] for i in range(len(rows))], score

    def rng_script(self, seeds, ops, args):
        seeds = np.ascontiguousarray(np.asarray(seeds, np.uint64))
        ops = np.ascontiguousarray(np.asarray(ops, np.int32))
        args = np.ascontiguousarray(np.asarray(args, np.int32))
        f = np.zeros((len(seeds), len(ops)), np.float64)
        u = np.zeros((len(seeds), len(ops)), np.uint32)
        _check(_lib.plp_rng_script(self.h, _ptr(seeds), len(seeds), _ptr(ops), _ptr(args),
                                   len(ops), _ptr(f), _ptr(u)))
        return f, u

    def plain_run(self, inst: Instance, seed: int, n_passes: int, prm: Params):
        row = RowResult()
        pl = np.zeros((max(inst.n_items, 1), 7), np.int32)
        _check(_lib.plp_plain_run(self.h, C.byref(inst.desc()), seed, n_passes, C.byref(prm),
                                  C.byref(row), _ptr(pl)))
        return row, pl[: row.n_boxes].copy()


class Engine:
    """Device-resident decoders (seam 3)."""

    def __init__(self, ctx: Ctx, ep_cap=0, blocks_per_cu=0, use_graph=False, waves_per_decoder=0,
                 count_work=False, box_cap=0, verdicts=0, share_prefix=0, trace_waves=0,
                 trace_verdicts=0):
        """verdicts: 0 / "auto", 1 / "pair_cache" or 2 / "point_state"
        (plp_engine_config.verdicts); share_prefix: 0 / "auto", 1 / "off" or
        2 / "on" (plp_engine_config.share_prefix); trace_waves (0, 4, 8) and
        trace_verdicts (as verdicts, 0 = the decoders'): the recording
        launch's shape when prefixes are shared."""
        self.ctx = ctx
        self.h = C.c_void_p()
        if isinstance(verdicts, str):
            verdicts = VERDICTS[verdicts]
        if isinstance(share_prefix, str):
            share_prefix = SHARE[share_prefix]
        if isinstance(trace_verdicts, str):
            trace_verdicts = VERDICTS[trace_verdicts]
        cfg = EngineConfig(int(waves_per_decoder), int(ep_cap), int(blocks_per_cu),
                           int(bool(use_graph)), int(bool(count_work)), int(box_cap),
                           int(verdicts), int(share_prefix), int(trace_waves),
                           int(trace_verdicts))
        _check(_lib.plp_engine_create(ctx.h, C.byref(cfg), C.byref(self.h)))
        ctx._children.add(self)
        self.n_dec = 0

    def close(self):
        if self.h:
            _lib.plp_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def load(self, instances, dec_inst, dec_seed, prm: Params, blocks: BlockParams = None):
        """Item decoders (plp_engine_load), or block-building decoders when
        blocks is given (plp_engine_load_blocks)."""
        insts = list(instances)
        self._keep = insts
        arr = (InstanceDesc * len(insts))(*[i.desc() for i in insts])
        di = np.ascontiguousarray(np.asarray(dec_inst, np.int32))
        ds = np.ascontiguousarray(np.asarray(dec_seed, np.uint64))
        self._arrs = (arr, di, ds)
        self.n_dec = len(di)
        if blocks is None:
            _check(_lib.plp_engine_load(self.h, arr, len(insts), _ptr(di), _ptr(ds), len(di),
                                        C.byref(prm)))
        else:
            _check(_lib.plp_engine_load_blocks(self.h, arr, len(insts), _ptr(di), _ptr(ds),
                                               len(di), C.byref(prm), C.byref(blocks)))

    def reset(self):
        _check(_lib.plp_engine_reset(self.h))

    def run(self, n_passes: int):
        _check(_lib.plp_engine_run(self.h, n_passes))

    def sync(self):
        _check(_lib.plp_engine_sync(self.h))

    def last_ms(self) -> float:
        v = C.c_float()
        _check(_lib.plp_engine_last_ms(self.h, C.byref(v)))
        return v.value

    def describe(self):
        a, b, c, d = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(_lib.plp_engine_describe(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"waves_per_decoder": a.value, "ep_cap": b.value, "lds_bytes": c.value,
                "grid": d.value}

    def autotune(self, n_passes: int):
        cfg = EngineConfig()
        ms = C.c_float()
        _check(_lib.plp_engine_autotune(self.h, n_passes, C.byref(cfg), C.byref(ms)))
        return {"waves_per_decoder": cfg.waves_per_decoder, "ep_cap": cfg.ep_cap,
                "blocks_per_cu": cfg.blocks_per_cu, "box_cap": cfg.box_cap,
                "verdicts": cfg.verdicts, "share_prefix": cfg.share_prefix,
                "trace_waves": cfg.trace_waves, "trace_verdicts": cfg.trace_verdicts}, ms.value

    def max_items(self) -> int:
        mi = C.c_int32()
        _check(_lib.plp_engine_max_items(self.h, C.byref(mi)))
        return mi.value

    def fetch(self, placements=True, out=None):
        """(rows, placements): plp_engine_fetch into new arrays, or into
        out = (RowResult array of n_dec, int32 array (n_dec, max_items, 7))."""
        if out is None:
            rows = (RowResult * self.n_dec)()
            pl = np.zeros((self.n_dec, self.max_items(), 7), np.int32) if placements else None
        else:
            rows, pl = out
        _check(_lib.plp_engine_fetch(self.h, rows, _ptr(pl) if placements else None))
        return (list(rows) if out is None else rows), pl

    def sync_fetch(self, placements=True, out=None):
        """sync() and fetch() in one call (plp_engine_sync_fetch): each
        decoder's results are copied as soon as it finishes, overlapping the
        decode; same arguments and results as fetch()."""
        if out is None:
            rows = (RowResult * self.n_dec)()
            pl = np.zeros((self.n_dec, self.max_items(), 7), np.int32) if placements else None
        else:
            rows, pl = out
        _check(_lib.plp_engine_sync_fetch(self.h, rows, _ptr(pl) if placements else None))
        return (list(rows) if out is None else rows), pl

    def results(self):
        """Zero-copy view of the results in the engine's pinned host memory
        (plp_engine_results): (record array of n_dec RowResult, int32 array
        (n_dec, max_items, 7)); valid until the next load, reset, run or close."""
        rp = C.POINTER(RowResult)()
        pp = _I32P()
        stride = C.c_int32()
        _check(_lib.plp_engine_results(self.h, C.byref(rp), C.byref(pp), C.byref(stride)))
        rows = C.cast(rp, C.POINTER(RowResult * self.n_dec)).contents
        pl = np.ctypeslib.as_array(pp, shape=(self.n_dec, stride.value // 7, 7))
        return rows, pl


class Server:
    """Continuous-batching request loop."""

    def __init__(self, device=0, max_inflight=4, mem_budget=0):
        self.h = C.c_void_p()
        _check(_lib.plp_server_create(device, max_inflight, mem_budget, C.byref(self.h)))
        self._req = {}
        _OPEN.add(self)

    def close(self):
        if self.h:
            _lib.plp_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def submit(self, inst: Instance, n_dec: int, seed0: int, n_passes: int, prm: Params):
        """Returns a request id, or None when the server is busy."""
        rid = C.c_int64()
        d = inst.desc()
        rc = _lib.plp_server_submit(self.h, C.byref(d), n_dec, seed0, n_passes, C.byref(prm),
                                    C.byref(rid))
        if rc == PLP_E_BUSY:
            return None
        _check(rc)
        self._req[rid.value] = (n_dec, inst.n_items)
        return rid.value

    def poll(self, rid: int, placements=False):
        """None while the request runs; then (rows, placements or None), rows
        being a ctypes array of n_dec RowResult."""
        n_dec, n_items = self._req[rid]
        done = C.c_int32()
        _check(_lib.plp_server_query(self.h, rid, C.byref(done)))
        if not done.value:
            return None
        rows = (RowResult * n_dec)()
        mi = C.c_int32()
        cap = n_dec * max(n_items, 1)
        pl = np.zeros((n_dec, max(n_items, 1), 7), np.int32) if placements else None
        rc = _lib.plp_server_poll(self.h, rid, C.byref(done), rows, _ptr(pl) if placements else None,
                                  cap, C.byref(mi))
        if rc not in (PLP_OK, PLP_E_EPCAP):
            _check(rc)
        if not done.value:
            return None
        del self._req[rid]
        return rows, pl

    def inflight(self) -> int:
        n = C.c_int32()
        _check(_lib.plp_server_inflight(self.h, C.byref(n)))
        return n.value

    def stats(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib.plp_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"launches": a.value, "requests": b.value, "decoders": c.value}


def decoder_digest(n_boxes, volume, score, fallback, n_passes, seed, placements) -> bytes:
    """SHA-256 of one decoder's result: the reference JSON fields (score as its
    IEEE-754 bits) and the selected packing's rows (int32 little-endian)."""
    import hashlib
    import struct
    head = struct.pack("<iqdiiQ", int(n_boxes), int(volume), float(score), int(bool(fallback)),
                       int(n_passes), int(seed))
    rows = np.ascontiguousarray(np.asarray(placements, "<i4").reshape(-1, 7)[: int(n_boxes)])
    return hashlib.sha256(head + rows.tobytes()).digest()


def results_digest(rows, placements) -> tuple[str, list[str]]:
    """(digest of all decoders, per-decoder digests as 16 hex digits) of
    engine rows and their placements."""
    import hashlib
    per = [decoder_digest(r.n_boxes, r.volume, r.score, r.fallback, r.n_passes, r.seed,
                          placements[k]) for k, r in enumerate(rows)]
    return hashlib.sha256(b"".join(per)).hexdigest(), [d.hex()[:16] for d in per]


def to_json_dict(row: RowResult, placements) -> dict:
    """The reference JSON fields (solver.py:298-307) except time_s."""
    return {
        "n_boxes": row.n_boxes,
        "volume_cm3": row.volume,
        "stability_score": row.score,
        "fallback": bool(row.fallback),
        "n_passes": row.n_passes,
        "seed": row.seed,
        "placements": [list(map(int, p)) for p in np.asarray(placements)[: row.n_boxes]],
    }
