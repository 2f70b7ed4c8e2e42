"""CPU oracle of the partitioning path -- TEST INFRASTRUCTURE ONLY.

ctypes loader for ``oracle/libhgp_oracle.so`` (C sources in this folder, see
``hgp_oracle.h``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg use it, as the checker or as the timed host
path; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgp_oracle.so")
MAX_LEVELS = 64


class OracleOpts(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("ub_num", C.c_int32),
        ("ub_den", C.c_int32),
        ("match_rounds", C.c_int32),
        ("trials", C.c_int32),
        ("fm_passes", C.c_int32),
        ("patience", C.c_int32),
        ("fm_kind", C.c_int32),
        ("refiner", C.c_int32),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("levels", C.c_int32),
        ("coarsest_n", C.c_int32),
        ("coarsest_m", C.c_int64),
        ("init_cut", C.c_int64),
        ("cut", C.c_int64),
        ("max_pwgt", C.c_int64),
        ("refine_iters", C.c_int64),
        ("level_n", C.c_int64 * MAX_LEVELS),
        ("level_m", C.c_int64 * MAX_LEVELS),
        ("level_iters", C.c_int32 * MAX_LEVELS),
        ("level_adjwgt", C.c_int64 * MAX_LEVELS),
    ]

    def to_dict(self):
        L = self.levels
        return {
            "levels": int(L),
            "coarsest_n": int(self.coarsest_n),
            "coarsest_m": int(self.coarsest_m),
            "init_cut": int(self.init_cut),
            "cut": int(self.cut),
            "max_pwgt": int(self.max_pwgt),
            "refine_iters": int(self.refine_iters),
            "level_n": [int(self.level_n[i]) for i in range(L + 1)],
            "level_m": [int(self.level_m[i]) for i in range(L + 1)],
            "level_iters": [int(self.level_iters[i]) for i in range(L + 1)],
            "level_adjwgt": [int(self.level_adjwgt[i]) for i in range(L + 1)],
        }


_P = C.c_void_p
_I32, _I64, _U64, _U32 = C.c_int32, C.c_int64, C.c_uint64, C.c_uint32
_SIG = {
    "oracle_scan_i32": (None, [_P, _I64]),
    "oracle_scan_i64": (None, [_P, _I64]),
    "oracle_part_weights": (None, [_I32, _P, _P, _I32, _P]),
    "oracle_edge_cut": (_I64, [_I32, _P, _P, _P, _P]),
    "oracle_balanced": (C.c_int, [_I32, _P, _P, _I32, _I64, _I64]),
    "oracle_cmap": (_I32, [_I32, _P, _P]),
    "oracle_contract": (_I64, [_I32, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _P]),
    "oracle_default_opts": (None, [_P]),
    "oracle_dga_partition": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P, _P]),
    "oracle_dga_coarsen": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "oracle_coarsen_threshold": (_I32, [_I32]),
    "oracle_maxvwgt": (_I32, [_I64, _I32]),
    "oracle_fb_apply": (None, [_I32, _I32, _P, _P, _P, _P, _I32, _P]),
    "oracle_jet_rw_step": (_I64, [_I32, _P, _P, _P, _P, _I32, _P, _P, _I64, _I64, _P]),
    "oracle_jet_rs_step": (_I64, [_I32, _P, _P, _P, _P, _I32, _P, _P, _I64, _I64, _P]),
    "oracle_cookie_cutter": (None, [_I32, _I32, _P, _I64, _P, _P, _P]),
    "oracle_jet_accept": (C.c_int, [_I64, _I64, _I64, _P, _P, _P]),
    "oracle_dga_match": (_I64, [_I32, _P, _P, _P, _P, _I32, _U64, _I32, _I32, _P]),
    "oracle_edge_hash": (_U32, [_U64, _I32, _I32, _I32]),
    "oracle_dga_initpart": (None, [_I32, _P, _P, _P, _P, _I32, _P, _P]),
    "oracle_jet_trace": (None, [_P, _I64]),
    "oracle_jet_trace_len": (_I64, []),
    "oracle_jet_dump": (None, [_I64, _I32, _P, _P]),
    "oracle_jet_dumped": (_I32, []),
    "oracle_heap_run": (_I32, [_I32, _I32, _P, _P, _P, _I32, _P, _P, _P, _P]),
    "oracle_dga_trial": (_I64, [_I32, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I32, _I32, _U64, _U32, _I32, _I32, _P]),
    "oracle_dga_refine": (_I32, [_I32, _P, _P, _P, _P, _I32, _I64, _I32, _P, _P, _P]),
    "oracle_jet_lp_step": (_I64, [_I32, _P, _P, _P, _I32, _P, _I32, _I32, _P, _P]),
    "oracle_gain_bucket": (_I32, [_I64, _I64]),
    "oracle_cpu_rb_partition": (_I64, [_I32, _P, _P, _P, _P, _I32, _U64, _P]),
    "oracle_lp_threshold": (_I32, [_I32, _I32]),
    "oracle_lp_part_bounds": (None, [_I32, _I64, _P, _P]),
    "oracle_lp_refine": (None, [_I32, _P, _P, _P, _P, _I32, _I64, _P, _P, _P]),
    "oracle_lp2_refine": (_I32, [_I32, _P, _P, _P, _P, _P]),
}

_lib = None


def build() -> None:
    """Compiles the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def opts(**kw) -> OracleOpts:
    o = OracleOpts()
    lib().oracle_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def scan_i32(a):
    a = _i32(a).copy()
    lib().oracle_scan_i32(_p(a), a.size)
    return a


def scan_i64(a):
    a = _i64(a).copy()
    lib().oracle_scan_i64(_p(a), a.size)
    return a


def part_weights(vwgt, where, k):
    vwgt, where = _i32(vwgt), _i32(where)
    out = np.zeros(int(k), dtype=np.int64)
    lib().oracle_part_weights(len(vwgt), _p(vwgt), _p(where), int(k), _p(out))
    return out


def edge_cut(xadj, adjncy, adjwgt, where) -> int:
    xadj, adjncy, adjwgt, where = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(where)
    return int(lib().oracle_edge_cut(len(xadj) - 1, _p(xadj), _p(adjncy), _p(adjwgt), _p(where)))


def balanced(vwgt, where, k, ub_num=103, ub_den=100) -> bool:
    vwgt, where = _i32(vwgt), _i32(where)
    return bool(lib().oracle_balanced(len(vwgt), _p(vwgt), _p(where), int(k), int(ub_num), int(ub_den)))


def cmap(match):
    m = _i32(match).copy()
    cm = np.zeros(len(m), dtype=np.int32)
    cn = lib().oracle_cmap(len(m), _p(m), _p(cm))
    return m, cm, int(cn)


def contract(xadj, adjncy, adjwgt, vwgt, match):
    """(match_normalised, cmap, cxadj, cadjncy, cadjwgt, cvwgt)"""
    xadj, adjncy, adjwgt, vwgt = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(vwgt)
    m, cm, cn = cmap(match)
    n = len(xadj) - 1
    mm = int(xadj[-1])
    cx = np.zeros(cn + 1, dtype=np.int64)
    ca = np.zeros(max(mm, 1), dtype=np.int32)
    cw = np.zeros(max(mm, 1), dtype=np.int32)
    cv = np.zeros(max(cn, 1), dtype=np.int32)
    e = lib().oracle_contract(n, _p(xadj), _p(adjncy), _p(adjwgt), _p(vwgt), _p(m), _p(cm), cn, _p(cx), _p(ca),
                              _p(cw), _p(cv))
    return m, cm, cx, ca[:e].copy(), cw[:e].copy(), cv[:cn].copy()


def coarsen_threshold(k):
    return int(lib().oracle_coarsen_threshold(int(k)))


def maxvwgt(tvwgt, thr):
    return int(lib().oracle_maxvwgt(int(tvwgt), int(thr)))


def edge_hash(seed, level, lo, hi):
    return int(lib().oracle_edge_hash(int(seed), int(level), int(lo), int(hi)))


def dga_match(xadj, adjncy, vwgt, adjwgt, maxvwgt_, seed, level, rounds):
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    m = np.zeros(n, dtype=np.int32)
    cnt = lib().oracle_dga_match(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(maxvwgt_), int(seed), int(level),
                                 int(rounds), _p(m))
    return m, int(cnt)


def dga_partition(xadj, adjncy, k, vwgt=None, adjwgt=None, **kw):
    xadj, adjncy = _i64(xadj), _i32(adjncy)
    n = len(xadj) - 1
    vwgt = np.ones(n, np.int32) if vwgt is None else _i32(vwgt)
    adjwgt = np.ones(len(adjncy), np.int32) if adjwgt is None else _i32(adjwgt)
    o = opts(**kw)
    st = OracleStats()
    part = np.zeros(n, dtype=np.int32)
    rc = lib().oracle_dga_partition(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), C.byref(o), _p(part),
                                    C.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_dga_partition failed: {rc}")
    return part, st.to_dict()


def dga_coarsen(xadj, adjncy, k, vwgt=None, adjwgt=None, **kw):
    """Coarsening phase alone: ((cxadj, cadjncy, cvwgt, cadjwgt), stats)."""
    xadj, adjncy = _i64(xadj), _i32(adjncy)
    n = len(xadj) - 1
    m = len(adjncy)
    vwgt = np.ones(n, np.int32) if vwgt is None else _i32(vwgt)
    adjwgt = np.ones(m, np.int32) if adjwgt is None else _i32(adjwgt)
    o = opts(**kw)
    st = OracleStats()
    cn, cm = C.c_int32(0), C.c_int64(0)
    cx = np.zeros(n + 1, np.int64)
    ca = np.zeros(max(m, 1), np.int32)
    cw = np.zeros(max(m, 1), np.int32)
    cv = np.zeros(n, np.int32)
    rc = lib().oracle_dga_coarsen(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), C.byref(o), C.byref(cn),
                                  C.byref(cm), _p(cx), _p(ca), _p(cw), _p(cv), C.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_dga_coarsen failed: {rc}")
    c, e = cn.value, cm.value
    return (cx[:c + 1].copy(), ca[:e].copy(), cv[:c].copy(), cw[:e].copy()), st.to_dict()


def dga_initpart(xadj, adjncy, vwgt, adjwgt, k, **kw):
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    o = opts(**kw)
    where = np.zeros(n, dtype=np.int32)
    lib().oracle_dga_initpart(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), C.byref(o), _p(where))
    return where


def dga_trial(xadj, adjncy, vwgt, adjwgt, tvwgt, onemin, onemax, ideal0, ntrials, trial, seed, node_id, fm_passes,
              fm_kind=0):
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    w = np.zeros(n, dtype=np.uint8)
    cut = lib().oracle_dga_trial(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(tvwgt), int(onemin), int(onemax),
                                 int(ideal0), int(ntrials), int(trial), int(seed), int(node_id), int(fm_passes),
                                 int(fm_kind), _p(w))
    return w, int(cut)


_trace_buf = None


def jet_trace_start(cap=1 << 22):
    """Starts recording the Jet loops of the following oracle calls."""
    global _trace_buf
    _trace_buf = np.zeros(cap, dtype=np.int64)
    lib().oracle_jet_trace(_p(_trace_buf), cap)


def jet_trace_stop():
    """Stops recording; returns rows of (level vertices, iterations done,
    mode of the coming iteration, cut so far, loop ended)."""
    global _trace_buf
    n = int(lib().oracle_jet_trace_len())
    out = _trace_buf[:min(n, len(_trace_buf))].reshape(-1, 5).copy()
    lib().oracle_jet_trace(None, 0)
    _trace_buf = None
    return out


def heap_run(nv, ops, ties=False):
    """Applies ops = [(op, v, key), ...] (op 0 insert, 1 top, 2 delete,
    3 update) to one device priority queue; returns (tops, vals, keys, loc):
    the vertex each top returned (None for other ops), the final slots and
    the final position index."""
    m = len(ops)
    op = np.array([o[0] for o in ops], dtype=np.int32)
    vv = np.array([o[1] for o in ops], dtype=np.int32)
    kk = np.array([o[2] for o in ops], dtype=np.int64)
    tops = np.full(max(m, 1), -2, dtype=np.int32)
    fv = np.zeros(max(nv, 1), dtype=np.int32)
    fk = np.zeros(max(nv, 1), dtype=np.int64)
    fl = np.zeros(max(nv, 1), dtype=np.int32)
    cnt = lib().oracle_heap_run(int(nv), int(m), _p(op), _p(vv), _p(kk), 1 if ties else 0, _p(tops), _p(fv), _p(fk),
                                _p(fl))
    return ([int(tops[i]) if op[i] == 1 else None for i in range(m)], fv[:cnt].tolist(), fk[:cnt].tolist(),
            fl[:nv].tolist())


def dga_refine(xadj, adjncy, vwgt, adjwgt, k, tvwgt, is_finest, where, cut, **kw):
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    o = opts(**kw)
    w = _i32(where).copy()
    c = np.array([cut], dtype=np.int64)
    it = lib().oracle_dga_refine(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), int(tvwgt), int(is_finest),
                                 C.byref(o), _p(w), _p(c))
    return w, int(c[0]), int(it)


def jet_lp_step(xadj, adjncy, adjwgt, k, where, filter_num, filter_den, lock):
    xadj, adjncy, adjwgt, where = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(where)
    n = len(xadj) - 1
    lk = np.ascontiguousarray(lock, dtype=np.uint8).copy()
    dest = np.zeros(n, dtype=np.int32)
    moves = lib().oracle_jet_lp_step(n, _p(xadj), _p(adjncy), _p(adjwgt), int(k), _p(where), int(filter_num),
                                     int(filter_den), _p(lk), _p(dest))
    return dest, lk, int(moves)


def gain_bucket(gain, vwgt):
    return int(lib().oracle_gain_bucket(int(gain), int(vwgt)))


def cpu_rb_partition(xadj, adjncy, k, vwgt=None, adjwgt=None, seed=1):
    xadj, adjncy = _i64(xadj), _i32(adjncy)
    n = len(xadj) - 1
    vwgt = np.ones(n, np.int32) if vwgt is None else _i32(vwgt)
    adjwgt = np.ones(len(adjncy), np.int32) if adjwgt is None else _i32(adjwgt)
    part = np.zeros(n, dtype=np.int32)
    cut = lib().oracle_cpu_rb_partition(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), int(seed), _p(part))
    return part, int(cut)


def _rebalance(fn, xadj, adjncy, adjwgt, vwgt, k, where, pwgts, maxwgt, opt):
    xadj, adjncy, adjwgt, vwgt, where = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(vwgt), _i32(where)
    pwgts = _i64(pwgts)
    dest = np.zeros(len(where), np.int32)
    m = fn(len(where), _p(xadj), _p(adjncy), _p(adjwgt), _p(vwgt), int(k), _p(where), _p(pwgts), int(maxwgt),
           int(opt), _p(dest))
    return dest, int(m)


def jet_rw_step(xadj, adjncy, adjwgt, vwgt, k, where, pwgts, maxwgt, opt):
    """One weak-rebalance step: (dest, moves)."""
    return _rebalance(lib().oracle_jet_rw_step, xadj, adjncy, adjwgt, vwgt, k, where, pwgts, maxwgt, opt)


def jet_rs_step(xadj, adjncy, adjwgt, vwgt, k, where, pwgts, maxwgt, opt):
    """One strong-rebalance step: (dest, moves)."""
    return _rebalance(lib().oracle_jet_rs_step, xadj, adjncy, adjwgt, vwgt, k, where, pwgts, maxwgt, opt)


def cookie_cutter(bs, maxwgt, pwgts):
    """Slices (es, ee) of the evicted list with weight prefix bs[0..ne]."""
    bs, pwgts = _i64(bs), _i64(pwgts)
    k = len(pwgts)
    es = np.zeros(k, np.int32)
    ee = np.zeros(k, np.int32)
    lib().oracle_cookie_cutter(int(k), int(len(bs) - 1), _p(bs), int(maxwgt), _p(pwgts), _p(es), _p(ee))
    return es, ee


def jet_accept(cur_cut, curmax, maxwgt, best_cut, best_max, count):
    """Acceptance rule: (accepted, best_cut, best_max, count) after one iteration."""
    bc, bm, cn = C.c_int64(best_cut), C.c_int64(best_max), C.c_int32(count)
    r = lib().oracle_jet_accept(int(cur_cut), int(curmax), int(maxwgt), C.byref(bc), C.byref(bm), C.byref(cn))
    return bool(r), int(bc.value), int(bm.value), int(cn.value)


FALLBACKS = {"leaf": 0, "isolate": 1, "twin": 2, "relative": 3, "forced": 4}


def fb_apply(kind, xadj, adjncy, adjwgt, vwgt, maxvwgt, match):
    """One matching fallback pass on a copy of match; returns the new match."""
    xadj, adjncy, adjwgt, vwgt = _i64(xadj), _i32(adjncy), _i32(adjwgt), _i32(vwgt)
    m = np.array(match, dtype=np.int32)
    lib().oracle_fb_apply(FALLBACKS[kind], len(xadj) - 1, _p(xadj), _p(adjncy), _p(adjwgt), _p(vwgt),
                          int(maxvwgt), _p(m))
    return m


def lp_threshold(n, k):
    return int(lib().oracle_lp_threshold(int(n), int(k)))


def lp_part_bounds(k, tvwgt):
    """(fmin, fmax) as the float32 values the older LP compares against."""
    lo, hi = C.c_float(0), C.c_float(0)
    lib().oracle_lp_part_bounds(int(k), int(tvwgt), C.byref(lo), C.byref(hi))
    return np.float32(lo.value), np.float32(hi.value)


def lp_refine(xadj, adjncy, vwgt, adjwgt, k, tvwgt, where):
    """Older k-way label propagation of one level: (where, cut, (rounds with
    moves, rebalancing rounds))."""
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    w = _i32(where).copy()
    c = np.zeros(1, dtype=np.int64)
    it = np.zeros(2, dtype=np.int32)
    lib().oracle_lp_refine(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), int(k), int(tvwgt), _p(w), _p(c), _p(it))
    return w, int(c[0]), (int(it[0]), int(it[1]))


def lp2_refine(xadj, adjncy, vwgt, adjwgt, where):
    """The 2-way label propagation: (where, transfer ran)."""
    xadj, adjncy, vwgt, adjwgt = _i64(xadj), _i32(adjncy), _i32(vwgt), _i32(adjwgt)
    n = len(xadj) - 1
    w = _i32(where).copy()
    ran = lib().oracle_lp2_refine(n, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), _p(w))
    return w, bool(ran)
