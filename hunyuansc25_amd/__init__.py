"""hunyuansc25_amd -- Python binding of the MI355X multilevel k-way partitioner.

The product is the C ABI in ``libhgp.so`` (declared in ``include/hunyuangraph.h``);
this module only loads it with ctypes and converts numpy arrays.  There is no
Python or CPU fallback: if the HIP library is missing every call raises.

``PartitionGraph`` mirrors the reference entry point ``hunyuangraph_PartitionGraph``
(hunyuangraph_partitiongraph.h:284): int32 CSR in, ``part`` out.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgp.so")

HGP_OK = 0
ERRORS = {
    -1: "invalid input",
    -2: "device error",
    -3: "out of device memory",
    -4: "unsupported",
    -5: "no HIP device",
    -6: "file error",
}

MAX_LEVELS = 64


class HgpOptions(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("ub_num", C.c_int32),
        ("ub_den", C.c_int32),
        ("match_rounds", C.c_int32),
        ("trials", C.c_int32),
        ("fm_passes", C.c_int32),
        ("patience", C.c_int32),
        ("validate", C.c_int32),
        ("device", C.c_int32),
        ("level_stats", C.c_int32),
        ("fm_kind", C.c_int32),
        ("refiner", C.c_int32),
    ]


class HgpTuning(C.Structure):
    _fields_ = [
        ("tiny_max_vertices", C.c_int32),
        ("few_rows", C.c_int32),
        ("fill_waves", C.c_int32),
        ("seg_grid", C.c_int32),
        ("batch_lanes", C.c_int32),
        ("hub_track", C.c_int32),
        ("graph_max_vertices", C.c_int32),
        ("batch_fuse_max_vertices", C.c_int32),
    ]


class HgpStats(C.Structure):
    _fields_ = [
        ("cut", C.c_int64),
        ("max_pwgt", C.c_int64),
        ("tvwgt", C.c_int64),
        ("levels", C.c_int32),
        ("coarsest_n", C.c_int32),
        ("coarsest_m", C.c_int64),
        ("init_cut", C.c_int64),
        ("refine_iters", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_coarsen", C.c_double),
        ("ms_init", C.c_double),
        ("ms_uncoarsen", C.c_double),
        ("alg_bytes", C.c_double),
        ("level_n", C.c_int64 * MAX_LEVELS),
        ("level_m", C.c_int64 * MAX_LEVELS),
        ("level_iters", C.c_int32 * MAX_LEVELS),
        ("arena_bytes", C.c_int64),
        ("arena_peak", C.c_int64),
        ("level_adjwgt", C.c_int64 * MAX_LEVELS),
        ("level_ms", C.c_double * MAX_LEVELS),
    ]

    def to_dict(self) -> dict:
        L = self.levels
        return {
            "cut": int(self.cut),
            "max_pwgt": int(self.max_pwgt),
            "tvwgt": int(self.tvwgt),
            "levels": int(L),
            "coarsest_n": int(self.coarsest_n),
            "coarsest_m": int(self.coarsest_m),
            "init_cut": int(self.init_cut),
            "refine_iters": int(self.refine_iters),
            "ms_total": float(self.ms_total),
            "ms_coarsen": float(self.ms_coarsen),
            "ms_init": float(self.ms_init),
            "ms_uncoarsen": float(self.ms_uncoarsen),
            "alg_bytes": float(self.alg_bytes),
            "level_n": [int(self.level_n[i]) for i in range(L + 1)],
            "level_m": [int(self.level_m[i]) for i in range(L + 1)],
            "level_iters": [int(self.level_iters[i]) for i in range(L + 1)],
            "arena_bytes": int(self.arena_bytes),
            "arena_peak": int(self.arena_peak),
            "level_adjwgt": [int(self.level_adjwgt[i]) for i in range(L + 1)],
            "level_ms": [float(self.level_ms[i]) for i in range(L + 1)],
        }


class HgpGraph(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int64),
        ("xadj", C.POINTER(C.c_int64)),
        ("adjncy", C.POINTER(C.c_int32)),
        ("adjwgt", C.POINTER(C.c_int32)),
        ("vwgt", C.POINTER(C.c_int32)),
    ]


# Every symbol declared in include/hunyuangraph.h, with its ctypes signature.
_P = C.c_void_p
_I32, _I64, _U64, _F64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
SIGNATURES = {
    "hunyuangraph_PartitionGraph": (None, [_P] * 9),
    "hgp_last_error": (C.c_char_p, []),
    "hgp_version": (C.c_char_p, []),
    "hgp_device_count": (C.c_int, []),
    "hgp_default_options": (None, [_P]),
    "hgp_partition": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P, _P]),
    "hgp_partition_device": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P, _P]),
    "hgp_check_graph_device": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _I32]),
    "hgp_evaluate_device": (C.c_int, [_I32, _P, _P, _P, _P, _P, _I32, _I32, _P]),
    "hgp_evaluate_device_noalloc": (C.c_int, [_I32, _P, _P, _P, _P, _P, _I32, _I32, _P]),
    "hgp_partition_cpu": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _U64, _P, _P]),
    "hgp_get_tuning": (C.c_int, [_P]),
    "hgp_set_tuning": (C.c_int, [_P]),
    "hgp_jet_trace_enable": (C.c_int, [C.c_int]),
    "hgp_jet_trace": (C.c_int64, [_P, C.c_int64]),
    "hgp_jet_trace_dump": (C.c_int, [C.c_int64, C.c_int32]),
    "hgp_jet_trace_dumped": (C.c_int64, [_P, _P, C.c_int64]),
    "hgp_partition_batch": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P]),
    "hgp_server_create": (C.c_int, [_P, _I32, _I64, _P]),
    "hgp_server_submit": (C.c_int64, [_P, _I32, _P, _P, _P, _P, _I32, _U64, _P]),
    "hgp_server_wait": (C.c_int, [_P, _I64, _P, _P]),
    "hgp_server_info": (C.c_int, [_P, _P, _P, _P]),
    "hgp_server_register": (C.c_int64, [_P, _I32, _P, _P, _P, _P]),
    "hgp_server_submit_graph": (C.c_int64, [_P, C.c_int64, _I32, _U64, _P]),
    "hgp_server_unregister": (C.c_int, [_P, C.c_int64]),
    "hgp_server_kept": (C.c_int, [_P, _P, _P, _P, _P]),
    "hgp_server_destroy": (C.c_int, [_P]),
    "hgp_ktimer_enable": (C.c_int, [C.c_int]),
    "hgp_ktimer_report": (C.c_int64, [_P, _I64]),
    "hgp_scan_i32": (C.c_int, [_P, _P, _I64]),
    "hgp_scan_i64": (C.c_int, [_P, _P, _I64]),
    "hgp_part_weights": (C.c_int, [_I32, _P, _P, _I32, _P]),
    "hgp_edge_cut": (C.c_int, [_I32, _P, _P, _P, _P, _P]),
    "hgp_match": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _U64, _I32, _I32, _P, _P]),
    "hgp_contract": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hgp_initpart": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P]),
    "hgp_refine": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _I64, _I32, _P, _P, _P, _P]),
    "hgp_refine_2way_lp": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P]),
    "hgp_graph_free": (None, [_P]),
    "hgp_gen_grid2d": (C.c_int, [_I32, _I32, _P]),
    "hgp_gen_rgg3d": (C.c_int, [_I32, _F64, _U64, _I32, _I32, _I32, _I32, _P]),
    "hgp_gen_rmat": (C.c_int, [_I32, _I64, _F64, _F64, _F64, _U64, _P]),
    "hgp_gen_banded": (C.c_int, [_I32, _I32, _F64, _U64, _P]),
    "hgp_read_graph": (C.c_int, [C.c_char_p, _P]),
    "hgp_write_partition": (C.c_int, [C.c_char_p, _P, _I32, _I32]),
    "hgp_read_mtx": (C.c_int, [C.c_char_p, _P]),
    "hgp_write_graph": (C.c_int, [C.c_char_p, _P, _I32]),
    "hgp_coarsen": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P, _P]),
}

_lib_handle = None


def lib() -> C.CDLL:
    """Loads libhgp.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
        _apply_tuning_file()
    return _lib_handle


def partition_cpu(xadj, adjncy, k, vwgt=None, adjwgt=None, seed=1):
    """The reference's CPU recursive bisection on the host; returns (part, cut)."""
    xadj, adjncy = _i64(xadj), _i32(adjncy)
    n = len(xadj) - 1
    vw = None if vwgt is None else _i32(vwgt)
    aw = None if adjwgt is None else _i32(adjwgt)
    part = np.zeros(max(n, 1), np.int32)
    cut = np.zeros(1, np.int64)
    _check(lib().hgp_partition_cpu(n, _ptr(xadj), _ptr(adjncy), _ptr(vw), _ptr(aw), int(k), int(seed), _ptr(part),
                                   _ptr(cut)))
    return part[:n], int(cut[0])


TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def get_tuning() -> dict:
    t = HgpTuning()
    _check(lib().hgp_get_tuning(C.byref(t)))
    return {name: getattr(t, name) for name, _ in HgpTuning._fields_}


def set_tuning(**kw) -> None:
    """Sets scheduling knobs (see hgp_tuning); results never depend on them."""
    t = HgpTuning()
    _check(lib().hgp_get_tuning(C.byref(t)))
    for name, v in kw.items():
        if name not in dict(HgpTuning._fields_):
            raise KeyError(name)
        setattr(t, name, int(v))
    _check(lib().hgp_set_tuning(C.byref(t)))


def _apply_tuning_file() -> None:
    # values measured by autotune.py on this kind of device, if present
    if os.environ.get("HGP_NO_TUNING_FILE") or not os.path.exists(TUNING_PATH):
        return
    import json
    with open(TUNING_PATH) as f:
        vals = json.load(f).get("tuning", {})
    set_tuning(**{k: v for k, v in vals.items() if k in dict(HgpTuning._fields_)})


class HgpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hgp error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != HGP_OK:
        raise HgpError(rc, lib().hgp_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def options(**kw) -> HgpOptions:
    o = HgpOptions()
    lib().hgp_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def device_count() -> int:
    return int(lib().hgp_device_count())


# ------------------------------------------------------------------ drop-in mirror

def PartitionGraph(nvtxs, xadj, adjncy, vwgt, adjwgt, nparts, tpwgts=None, ubvec=None) -> np.ndarray:
    """Mirror of hunyuangraph_PartitionGraph (hunyuangraph_partitiongraph.h:284):
    int32 CSR in, part[] out.  vwgt/adjwgt may be None (unit weights); ubvec is
    the balance bound (default 1.03); tpwgts are uniform as in the reference."""
    xadj = _i32(xadj)
    adjncy = _i32(adjncy)
    vw = None if vwgt is None else _i32(vwgt)
    aw = None if adjwgt is None else _i32(adjwgt)
    n = C.c_int(int(nvtxs))
    k = C.c_int(int(nparts))
    ub = None if ubvec is None else (C.c_float * 1)(float(ubvec))
    part = np.zeros(int(nvtxs), dtype=np.int32)
    L = lib()
    L.hunyuangraph_PartitionGraph(C.byref(n), _ptr(xadj), _ptr(adjncy), _ptr(vw), _ptr(aw), C.byref(k), None,
                                  None if ub is None else C.cast(ub, C.c_void_p), _ptr(part))
    err = L.hgp_last_error().decode()
    if err:
        raise HgpError(-1, err)
    return part


# ------------------------------------------------------------------ versioned API

def partition(xadj, adjncy, k, vwgt=None, adjwgt=None, **opts):
    """Full pipeline on host arrays; returns (part, stats dict)."""
    xadj = _i64(xadj)
    adjncy = _i32(adjncy)
    n = len(xadj) - 1
    vw = None if vwgt is None else _i32(vwgt)
    aw = None if adjwgt is None else _i32(adjwgt)
    o = options(**opts)
    st = HgpStats()
    part = np.zeros(max(n, 0), dtype=np.int32)
    _check(lib().hgp_partition(n, _ptr(xadj), _ptr(adjncy), _ptr(vw), _ptr(aw), int(k), C.byref(o), _ptr(part),
                               C.byref(st)))
    return part, st.to_dict()


def coarsen(xadj, adjncy, k, vwgt=None, adjwgt=None, **opts):
    """Coarsening phase of partition() alone; returns ((xadj, adjncy, vwgt,
    adjwgt) of the coarsest graph, stats dict)."""
    xadj = _i64(xadj)
    adjncy = _i32(adjncy)
    n = len(xadj) - 1
    vw = None if vwgt is None else _i32(vwgt)
    aw = None if adjwgt is None else _i32(adjwgt)
    o = options(**opts)
    st = HgpStats()
    g = HgpGraph()
    _check(lib().hgp_coarsen(n, _ptr(xadj), _ptr(adjncy), _ptr(vw), _ptr(aw), int(k), C.byref(o), C.byref(g),
                             C.byref(st)))
    return _take_graph(g), st.to_dict()


def partition_batch(graphs, ks, lanes=0, **opts):
    """Partitions many graphs in one call.  graphs: list of (xadj, adjncy,
    vwgt, adjwgt) with vwgt/adjwgt possibly None; ks: parts per graph.
    Returns (list of part arrays, list of stats dicts)."""
    cnt = len(graphs)
    if len(ks) != cnt:
        raise ValueError("one k per graph")
    keep = []
    ns = np.zeros(max(cnt, 1), np.int32)
    arrs = {name: (C.c_void_p * max(cnt, 1))() for name in ("x", "a", "vw", "aw", "p")}
    parts = []
    for i, (x, a, vw, aw) in enumerate(graphs):
        x, a = _i64(x), _i32(a)
        vw = None if vw is None else _i32(vw)
        aw = None if aw is None else _i32(aw)
        p = np.zeros(max(len(x) - 1, 0), np.int32)
        keep += [x, a, vw, aw]
        parts.append(p)
        ns[i] = len(x) - 1
        arrs["x"][i], arrs["a"][i] = x.ctypes.data, a.ctypes.data
        arrs["vw"][i] = None if vw is None else vw.ctypes.data
        arrs["aw"][i] = None if aw is None else aw.ctypes.data
        arrs["p"][i] = p.ctypes.data
    kk = np.ascontiguousarray(ks, np.int32)
    o = options(**opts)
    status = np.zeros(max(cnt, 1), np.int32)
    st = (HgpStats * max(cnt, 1))()
    rc = lib().hgp_partition_batch(cnt, _ptr(ns), arrs["x"], arrs["a"], arrs["vw"], arrs["aw"], _ptr(kk),
                                   C.byref(o), int(lanes), arrs["p"], _ptr(status), st)
    _check(rc)
    return parts, [st[i].to_dict() for i in range(cnt)]


class Server:
    """The serving loop: submit() queues a job and returns a ticket, wait()
    returns (part, stats, latency_ms).  Arrays are kept alive until waited."""

    def __init__(self, slots=4, budget_bytes=200 << 30, **opts):
        self._h = C.c_void_p()
        o = options(**opts)
        _check(lib().hgp_server_create(C.byref(o), int(slots), int(budget_bytes), C.byref(self._h)))
        self._live = {}
        self._graph_n = {}
        import threading
        self._mu = threading.Lock()
        self._waiting = set()

    def submit(self, xadj, adjncy, k, vwgt=None, adjwgt=None, seed=1) -> int:
        x, a = _i64(xadj), _i32(adjncy)
        vw = None if vwgt is None else _i32(vwgt)
        aw = None if adjwgt is None else _i32(adjwgt)
        part = np.zeros(max(len(x) - 1, 0), np.int32)
        t = lib().hgp_server_submit(self._h, len(x) - 1, _ptr(x), _ptr(a), _ptr(vw), _ptr(aw), int(k), int(seed),
                                    _ptr(part))
        if t < 0:
            _check(int(t))
        self._live[t] = (x, a, vw, aw, part)
        return t

    def register(self, xadj, adjncy, vwgt=None, adjwgt=None) -> int:
        """Uploads a graph once; returns its id for submit_graph()."""
        x, a = _i64(xadj), _i32(adjncy)
        vw = None if vwgt is None else _i32(vwgt)
        aw = None if adjwgt is None else _i32(adjwgt)
        g = lib().hgp_server_register(self._h, len(x) - 1, _ptr(x), _ptr(a), _ptr(vw), _ptr(aw))
        if g < 0:
            _check(int(g))
        self._graph_n[g] = len(x) - 1
        return g

    def submit_graph(self, graph, k, seed=1) -> int:
        """Queues a partition of a registered graph; returns a ticket."""
        part = np.zeros(max(self._graph_n.get(graph, 1), 1), np.int32)
        t = lib().hgp_server_submit_graph(self._h, int(graph), int(k), int(seed), _ptr(part))
        if t < 0:
            _check(int(t))
        self._live[t] = (part,)
        return t

    def unregister(self, graph) -> None:
        _check(lib().hgp_server_unregister(self._h, int(graph)))
        self._graph_n.pop(graph, None)

    def kept(self) -> dict:
        h, m, kn, kb = (C.c_int64() for _ in range(4))
        _check(lib().hgp_server_kept(self._h, C.byref(h), C.byref(m), C.byref(kn), C.byref(kb)))
        return {"hits": h.value, "misses": m.value, "kept": kn.value, "kept_bytes": kb.value}

    def wait(self, ticket):
        """A ticket is waited for once: a second wait raises HgpError (-1)
        and leaves the first waiter's job alone."""
        with self._mu:
            if ticket in self._waiting:
                raise HgpError(-1, "ticket is already being waited for")
            self._waiting.add(ticket)
        try:
            st = HgpStats()
            lat = C.c_double()
            rc = lib().hgp_server_wait(self._h, int(ticket), C.byref(st), C.byref(lat))
            arrs = self._live.pop(ticket, None) if rc == HGP_OK or ticket in self._live else None
            _check(rc)
            return arrs[-1], st.to_dict(), lat.value
        finally:
            with self._mu:
                self._waiting.discard(ticket)

    def info(self) -> dict:
        r, p, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().hgp_server_info(self._h, C.byref(r), C.byref(p), C.byref(c)))
        return {"reserved": r.value, "peak_reserved": p.value, "completed": c.value}

    def close(self) -> None:
        if self._h:
            _check(lib().hgp_server_destroy(self._h))
            self._h = C.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def partition_device(n, d_xadj, d_adjncy, d_vwgt, d_adjwgt, k, d_part, **opts) -> dict:
    """Full pipeline on device pointers (ints, e.g. torch tensor data_ptr())."""
    o = options(**opts)
    st = HgpStats()
    _check(lib().hgp_partition_device(int(n), C.c_void_p(d_xadj), C.c_void_p(d_adjncy), C.c_void_p(d_vwgt),
                                      C.c_void_p(d_adjwgt), int(k), C.byref(o), C.c_void_p(d_part),
                                      C.byref(st)))
    return st.to_dict()


def check_graph_device(n, d_xadj, d_adjncy, d_vwgt=0, d_adjwgt=0, symmetric=False, device=0) -> None:
    """Checks a device CSR (device pointers as ints; 0 = unit weights); raises
    HgpError naming the first fault."""
    _check(lib().hgp_check_graph_device(int(n), C.c_void_p(d_xadj), C.c_void_p(d_adjncy),
                                        C.c_void_p(d_vwgt or None), C.c_void_p(d_adjwgt or None),
                                        1 if symmetric else 0, int(device)))


def evaluate_device(n, d_xadj, d_adjncy, d_adjwgt, d_vwgt, d_part, k, d_out, device=0) -> None:
    """Writes {cut, heaviest part, bad part ids, bad neighbour ids} (4 int64)
    to the device array d_out."""
    _check(lib().hgp_evaluate_device(int(n), C.c_void_p(d_xadj), C.c_void_p(d_adjncy),
                                     C.c_void_p(d_adjwgt or None), C.c_void_p(d_vwgt or None),
                                     C.c_void_p(d_part), int(k), int(device), C.c_void_p(d_out)))


def evaluate_device_noalloc(n, d_xadj, d_adjncy, d_adjwgt, d_vwgt, d_part, k, d_out, device=0) -> None:
    """evaluate_device with the library's fixed per-device scratch: allocates
    no device memory (k <= 4096)."""
    _check(lib().hgp_evaluate_device_noalloc(int(n), C.c_void_p(d_xadj), C.c_void_p(d_adjncy),
                                             C.c_void_p(d_adjwgt or None), C.c_void_p(d_vwgt or None),
                                             C.c_void_p(d_part), int(k), int(device), C.c_void_p(d_out)))


# ------------------------------------------------------------------ Jet trace

def jet_trace_enable(on: bool = True) -> None:
    _check(lib().hgp_jet_trace_enable(1 if on else 0))


def jet_trace() -> np.ndarray:
    """Records of the Jet loops since enabling: rows of (level vertices,
    iterations done, mode of the coming iteration, cut so far, loop ended)."""
    total = int(lib().hgp_jet_trace(None, 0))
    buf = np.zeros(max(total, 1), dtype=np.int64)
    lib().hgp_jet_trace(buf.ctypes.data_as(C.c_void_p), total)
    return buf[:total].reshape(-1, 5)


def jet_trace_dump(n, it) -> None:
    """While tracing, keep where/dest of iteration `it` of the level with n vertices."""
    _check(lib().hgp_jet_trace_dump(int(n), int(it)))


def jet_trace_dumped():
    """(where before the moves, dest) of the kept iteration, or None."""
    n = int(lib().hgp_jet_trace_dumped(None, None, 0))
    if n == 0:
        return None
    w = np.zeros(n, dtype=np.int32)
    d = np.zeros(n, dtype=np.int32)
    lib().hgp_jet_trace_dumped(w.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p), n)
    return w, d


# ------------------------------------------------------------------ kernel timer

def ktimer_enable(on: bool = True) -> None:
    _check(lib().hgp_ktimer_enable(1 if on else 0))


def ktimer_report() -> dict:
    """{kernel: (launches, total_ms, algorithmic_bytes)} of the recorded launches."""
    import json
    need = lib().hgp_ktimer_report(None, 0)
    if need < 0:
        _check(int(need))
    buf = C.create_string_buffer(int(need) + 16)
    lib().hgp_ktimer_report(buf, len(buf))
    return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}


# ------------------------------------------------------------------ stages

def scan_i32(a) -> np.ndarray:
    a = _i32(a)
    out = np.empty_like(a)
    _check(lib().hgp_scan_i32(_ptr(a), _ptr(out), a.size))
    return out


def scan_i64(a) -> np.ndarray:
    a = _i64(a)
    out = np.empty_like(a)
    _check(lib().hgp_scan_i64(_ptr(a), _ptr(out), a.size))
    return out


def part_weights(vwgt, where, k) -> np.ndarray:
    vwgt, where = _i32(vwgt), _i32(where)
    out = np.zeros(int(k), dtype=np.int64)
    _check(lib().hgp_part_weights(len(vwgt), _ptr(vwgt), _ptr(where), int(k), _ptr(out)))
    return out


def edge_cut(xadj, adjncy, adjwgt, where) -> int:
    xadj, adjncy, adjwgt, where = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(where)
    out = np.zeros(1, dtype=np.int64)
    _check(lib().hgp_edge_cut(len(xadj) - 1, _ptr(xadj), _ptr(adjncy), _ptr(adjwgt), _ptr(where), _ptr(out)))
    return int(out[0])


def match(xadj, adjncy, vwgt, adjwgt, maxvwgt, seed, level, rounds):
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    m = np.zeros(n, dtype=np.int32)
    cnt = np.zeros(1, dtype=np.int64)
    _check(lib().hgp_match(n, _ptr(xadj), _ptr(adjncy), _ptr(vwgt), _ptr(adjwgt), int(maxvwgt), int(seed),
                           int(level), int(rounds), _ptr(m), _ptr(cnt)))
    return m, int(cnt[0])


def contract(xadj, adjncy, adjwgt, vwgt, match_arr):
    """cmap + canonical contraction of a fixed matching; returns
    (match_normalised, cmap, cxadj, cadjncy, cadjwgt, cvwgt)."""
    xadj, adjncy, adjwgt, vwgt = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(vwgt)
    mt = _i32(match_arr).copy()
    n = len(xadj) - 1
    m = int(xadj[-1])
    cmap = np.zeros(n, dtype=np.int32)
    cxadj = np.zeros(n + 1, dtype=np.int64)
    cadj = np.zeros(max(m, 1), dtype=np.int32)
    cw = np.zeros(max(m, 1), dtype=np.int32)
    cvw = np.zeros(max(n, 1), dtype=np.int32)
    cn = np.zeros(1, dtype=np.int32)
    cm = np.zeros(1, dtype=np.int64)
    _check(lib().hgp_contract(n, _ptr(xadj), _ptr(adjncy), _ptr(adjwgt), _ptr(vwgt), _ptr(mt), _ptr(cmap), _ptr(cn),
                              _ptr(cxadj), _ptr(cadj), _ptr(cw), _ptr(cvw), _ptr(cm)))
    c, e = int(cn[0]), int(cm[0])
    return mt, cmap, cxadj[: c + 1].copy(), cadj[:e].copy(), cw[:e].copy(), cvw[:c].copy()


def initpart(xadj, adjncy, vwgt, adjwgt, k, **opts) -> np.ndarray:
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    o = options(**opts)
    where = np.zeros(n, dtype=np.int32)
    _check(lib().hgp_initpart(n, _ptr(xadj), _ptr(adjncy), _ptr(vwgt), _ptr(adjwgt), int(k), C.byref(o),
                              _ptr(where)))
    return where


def refine(xadj, adjncy, vwgt, adjwgt, k, tvwgt, is_finest, where, cut, **opts):
    """One level of Jet refinement (refiner=1: the older label propagation);
    returns (where, cut, iters)."""
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    o = options(**opts)
    w = _i32(where).copy()
    c = np.array([cut], dtype=np.int64)
    it = np.zeros(1, dtype=np.int32)
    _check(lib().hgp_refine(n, _ptr(xadj), _ptr(adjncy), _ptr(vwgt), _ptr(adjwgt), int(k), int(tvwgt),
                            int(is_finest), C.byref(o), _ptr(w), _ptr(c), _ptr(it)))
    return w, int(c[0]), int(it[0])


def refine_2way_lp(xadj, adjncy, vwgt, adjwgt, where, **opts):
    """The 2-way label propagation pass (jetLP) on a 0/1 partition; returns
    (where, transfer ran)."""
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    o = options(**opts)
    w = _i32(where).copy()
    ran = np.zeros(1, dtype=np.int32)
    _check(lib().hgp_refine_2way_lp(n, _ptr(xadj), _ptr(adjncy), _ptr(vwgt), _ptr(adjwgt), C.byref(o), _ptr(w),
                                    _ptr(ran)))
    return w, bool(ran[0])


# ------------------------------------------------------------------ generators

def _take_graph(g: HgpGraph):
    n, m = int(g.n), int(g.m)
    xadj = np.ctypeslib.as_array(g.xadj, shape=(n + 1,)).copy()
    adjncy = np.ctypeslib.as_array(g.adjncy, shape=(max(m, 1),))[:m].copy()
    adjwgt = np.ctypeslib.as_array(g.adjwgt, shape=(max(m, 1),))[:m].copy()
    vwgt = np.ctypeslib.as_array(g.vwgt, shape=(max(n, 1),))[:n].copy()
    lib().hgp_graph_free(C.byref(g))
    return xadj, adjncy, vwgt, adjwgt


def gen_grid2d(rows, cols):
    g = HgpGraph()
    _check(lib().hgp_gen_grid2d(int(rows), int(cols), C.byref(g)))
    return _take_graph(g)


def gen_rgg3d(n, radius, seed=1, vw=(0, 0), ew=(0, 0)):
    g = HgpGraph()
    _check(lib().hgp_gen_rgg3d(int(n), float(radius), int(seed), int(vw[0]), int(vw[1]), int(ew[0]), int(ew[1]),
                               C.byref(g)))
    return _take_graph(g)


def gen_rmat(scale, nedges, a=0.57, b=0.19, c=0.19, seed=1):
    g = HgpGraph()
    _check(lib().hgp_gen_rmat(int(scale), int(nedges), float(a), float(b), float(c), int(seed), C.byref(g)))
    return _take_graph(g)


def gen_banded(n, bw=32, frac=0.1, seed=1):
    g = HgpGraph()
    _check(lib().hgp_gen_banded(int(n), int(bw), float(frac), int(seed), C.byref(g)))
    return _take_graph(g)


# ------------------------------------------------------------------ files

def read_graph(path):
    """METIS .graph file -> (xadj, adjncy, vwgt, adjwgt), as hunyuangraph_readgraph
    reads it (hunyuangraph_io.h:25-186); raises HgpError on a malformed file."""
    g = HgpGraph()
    _check(lib().hgp_read_graph(os.fsencode(path), C.byref(g)))
    return _take_graph(g)


def read_mtx(path):
    """Matrix Market matrix -> graph of pattern(A + A^T) without the diagonal
    (matrix_to_graph/matrix_to_graph.c:117-281)."""
    g = HgpGraph()
    _check(lib().hgp_read_mtx(os.fsencode(path), C.byref(g)))
    return _take_graph(g)


def write_graph(path, xadj, adjncy, vwgt=None, adjwgt=None):
    """Writes a METIS .graph file; weights are written when given."""
    xadj, adjncy = _i64(xadj), _i32(adjncy)
    vw = None if vwgt is None else _i32(vwgt)
    ew = None if adjwgt is None else _i32(adjwgt)
    g = HgpGraph()
    g.n, g.m = len(xadj) - 1, int(xadj[-1])
    g.xadj = xadj.ctypes.data_as(C.POINTER(C.c_int64))
    g.adjncy = adjncy.ctypes.data_as(C.POINTER(C.c_int32))
    if ew is not None:
        g.adjwgt = ew.ctypes.data_as(C.POINTER(C.c_int32))
    if vw is not None:
        g.vwgt = vw.ctypes.data_as(C.POINTER(C.c_int32))
    weights = (10 if vw is not None else 0) + (1 if ew is not None else 0)
    _check(lib().hgp_write_graph(os.fsencode(path), C.byref(g), weights))


def write_partition(fname, part, nparts):
    """Writes part one value per line to "<fname>.part.<nparts>" (hunyuangraph_io.h:189-205);
    returns that path."""
    p = _i32(part)
    _check(lib().hgp_write_partition(os.fsencode(fname), _ptr(p), len(p), int(nparts)))
    return f"{os.fsdecode(fname)}.part.{int(nparts)}"


def rgg_radius(n: int, degree: float = 12.0) -> float:
    """Radius of a 3-D RGG with the given expected degree before boundary loss."""
    return (3.0 * degree / (4.0 * np.pi * n)) ** (1.0 / 3.0)
