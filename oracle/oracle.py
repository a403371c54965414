"""ctypes view of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker for the HIP engine.  The product package never imports this file.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

STAT_KEYS = ("status", "c_loop_iters", "outer_iters", "r_loop_calls", "s_loop_calls",
             "saddle_loop_calls", "c_loops_calls", "edited", "n_max", "n_min",
             "n_saddle1", "n_saddle2")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.oracle_lut.argtypes = [vp, i32]
        L.oracle_link_type.argtypes = [i32, i32]
        L.oracle_classify.argtypes = [vp, i32, i32, i32, vp, vp]
        L.oracle_labels.argtypes = [vp, i32, i32, i32, vp, vp]
        L.oracle_max_label.argtypes = [vp, i32, i32, i32, i32]
        L.oracle_min_label.argtypes = [vp, i32, i32, i32, i32]
        L.oracle_correct.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, vp, vp, i32]
        L.oracle_correct_rule.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, vp, vp, i32, i32]
        L.oracle_correct_rules.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, vp, vp, i32, i32,
                                           i32]
        L.oracle_bitmap3.argtypes = [vp, i64, vp]
        L.oracle_bitmap3.restype = i64
        L.oracle_cost.argtypes = [vp, vp, vp, vp, i64, dbl, vp, vp, vp]
        L.oracle_set_adjacency.argtypes = [i32]
        L.oracle_set_persistence.argtypes = [dbl]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(f):
    D, H, W = f.shape
    return W, H, D


def lut():
    out = np.zeros(24800, np.int32)
    n = lib().oracle_lut(_p(out), out.size)
    return out[:n]


def link_type(key, config):
    return lib().oracle_link_type(key, config)


def classify(field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    t = np.empty(f.shape, np.uint8)
    m = np.empty(f.shape, np.uint16)
    lib().oracle_classify(_p(f), *_dims(f), _p(t), _p(m))
    return t, m


def labels(field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    ds = np.empty(f.shape, np.int32)
    asc = np.empty(f.shape, np.int32)
    lib().oracle_labels(_p(f), *_dims(f), _p(ds), _p(asc))
    return ds, asc


def max_label(field, v):
    f = np.ascontiguousarray(field, dtype=np.float64)
    return lib().oracle_max_label(_p(f), *_dims(f), int(v))


def min_label(field, v):
    f = np.ascontiguousarray(field, dtype=np.float64)
    return lib().oracle_min_label(_p(f), *_dims(f), int(v))


def correct(orig, decp, xi, max_c_iters=0, max_outer=0, threads=0, fixed_step_q=0, f32=False):
    """Returns (fixed, counts, stats dict).  Inputs are not modified.
    fixed_step_q > 0 selects the fixed-step rule set (step -xi/q, threshold q);
    f32 the float32 fork's rule set."""
    o = np.ascontiguousarray(orig, dtype=np.float64)
    d = np.array(decp, dtype=np.float64, order="C", copy=True)
    counts = np.zeros(o.shape, np.uint8)
    st = np.zeros(12, np.int64)
    rc = lib().oracle_correct_rules(_p(o), _p(d), *_dims(o), float(xi), int(max_c_iters),
                                    int(max_outer), _p(counts), _p(st), int(threads),
                                    int(fixed_step_q), int(bool(f32)))
    if rc == -22:
        raise ValueError("invalid input (shape, bound, rule options, or values the float32 "
                         "rule set cannot hold)")
    return d, counts, dict(zip(STAT_KEYS, (int(x) for x in st)))


def bitmap3(vals):
    v = np.ascontiguousarray(vals, dtype=np.uint8).ravel()
    out = np.zeros((3 * v.size + 7) // 8 + 1, np.uint8)
    n = lib().oracle_bitmap3(_p(v), v.size, _p(out))
    if n < 0:
        raise ValueError("value outside 0..6")
    return out[:n]


def cost(orig, before, after, counts, xi):
    """cost() restated: ({changed, snapped, edit_ratio, psnr, fixed_psnr}, bitmap, edits)."""
    o = np.ascontiguousarray(orig, dtype=np.float64).ravel()
    b = np.ascontiguousarray(before, dtype=np.float64).ravel()
    a = np.ascontiguousarray(after, dtype=np.float64).ravel()
    c = np.ascontiguousarray(counts, dtype=np.uint8).ravel()
    out = np.zeros(5, np.float64)
    bm = np.zeros((3 * o.size + 7) // 8 + 1, np.uint8)
    ed = np.zeros(o.size, np.float64)
    rc = lib().oracle_cost(_p(o), _p(b), _p(a), _p(c), o.size, float(xi), _p(out), _p(bm), _p(ed))
    if rc:
        raise ValueError("edit count outside 0..6")
    st = {"changed": int(out[0]), "snapped": int(out[1]), "edit_ratio": float(out[2]),
          "psnr": float(out[3]), "fixed_psnr": float(out[4])}
    return st, bm[:(3 * o.size + 7) // 8], ed[:st["snapped"]]


def split_tree(field):
    """compute_MergeT restated for one field: dict with per-maximum (rank order)
    'max', 'pair_saddle', 'parent' vertex ids, 'branch' vertex lists, 'rounds'."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    L = lib()
    if not hasattr(L.oracle_split_tree, "_set"):
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.oracle_split_tree.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, vp, i32, i32, vp]
        L.oracle_split_tree._set = True
    cap = f.size + 1
    mx, ps, pa = (np.zeros(cap, np.int32) for _ in range(3))
    off = np.zeros(cap + 1, np.int32)
    bv = np.zeros(3 * cap, np.int32)
    rounds = ctypes.c_int(0)
    n = L.oracle_split_tree(_p(f), *_dims(f), _p(mx), _p(ps), _p(pa), _p(off), _p(bv), cap,
                            bv.size, ctypes.byref(rounds))
    if n < 0:
        raise ValueError("buffer too small")
    return {"max": mx[:n].copy(), "pair_saddle": ps[:n].copy(), "parent": pa[:n].copy(),
            "branch": [bv[off[i]:off[i + 1]].tolist() for i in range(n)],
            "rounds": rounds.value}


PHASES = ("presnap", "c_loops", "r_loops", "s_loops", "s_loops_join", "saddle_loops",
          "saddle_loops_join")
MUTATIONS = ("none", "wn_any_label", "wn_all", "wn_desc_order", "fixpath_true",
             "fixpath_false", "wi_last_wins", "sloops_swap", "saddle_no_guard",
             "saddle_join_lower")


def correct_trace(orig, decp, xi, mutation="none"):
    """The correction with its edit trace: (fixed, counts, stats, edits) where
    edits lists (phase name, vertex, value before, value after) for every
    value the pre-snap or an apply_delta wrote, in order.  mutation plants
    one of MUTATIONS (a plausible mistake in one rule) for known-answer tests."""
    L = lib()
    if not hasattr(L.oracle_correct_trace, "_set"):
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_correct_trace.argtypes = [vp, vp, i32, i32, i32, dbl, i32, vp, vp, vp, vp, vp,
                                           vp, i32]
        L.oracle_correct_trace._set = True
    o = np.ascontiguousarray(orig, dtype=np.float64)
    d = np.array(decp, dtype=np.float64, order="C", copy=True)
    counts = np.zeros(o.shape, np.uint8)
    st = np.zeros(12, np.int64)
    cap = 100000
    ph = np.zeros(cap, np.int32)
    vx = np.zeros(cap, np.int32)
    bf = np.zeros(cap, np.float64)
    af = np.zeros(cap, np.float64)
    n = L.oracle_correct_trace(_p(o), _p(d), *_dims(o), float(xi), MUTATIONS.index(mutation),
                               _p(counts), _p(st), _p(ph), _p(vx), _p(bf), _p(af), cap)
    if n < 0:
        raise ValueError(f"oracle_correct_trace failed ({n})")
    edits = [(PHASES[ph[i]], int(vx[i]), float(bf[i]), float(af[i])) for i in range(n)]
    return d, counts, dict(zip(STAT_KEYS, (int(x) for x in st))), edits


RULES = ("wrong_neighbors", "fixpath", "wrong_index_max", "wrong_index_min", "saddles",
         "saddles_join")


def rule_probe(orig, decp, xi, rule, mutation="none"):
    """One rule of the outer loop on the state (orig, decp), nothing applied.

    Returns a dict: wrong_neighbors / fixpath: 'split' and 'join' flagged
    vertex lists; wrong_index_max / _min: 'pairs_0' (direction 0: the first
    ([0]) comparison for maxima, the last for minima) and 'pairs_1' as
    (true, false) tuples; saddles / saddles_join: 'pairs'.  Every rule but
    wrong_neighbors also returns 'proposals', {vertex: delta} of the delta
    buffer before apply_delta."""
    L = lib()
    if not hasattr(L.oracle_rule_probe, "_set"):
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_rule_probe.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, vp, i32, vp, vp,
                                        i32, vp]
        L.oracle_rule_probe._set = True
    o = np.ascontiguousarray(orig, dtype=np.float64)
    d = np.array(decp, dtype=np.float64, order="C", copy=True)
    cap = 4 * o.size + 16
    iout = np.zeros(cap, np.int32)
    pv = np.zeros(o.size, np.int32)
    pd = np.zeros(o.size, np.float64)
    npr = ctypes.c_int(0)
    r = RULES.index(rule)
    rc = L.oracle_rule_probe(_p(o), _p(d), *_dims(o), float(xi), r, MUTATIONS.index(mutation),
                             _p(iout), cap, _p(pv), _p(pd), o.size, ctypes.byref(npr))
    if rc:
        raise ValueError(f"oracle_rule_probe failed ({rc})")
    out = {}
    if r in (0, 1):
        n = int(iout[0])
        out["split"] = iout[1:1 + n].tolist()
        m = int(iout[1 + n])
        out["join"] = iout[2 + n:2 + n + m].tolist()
    else:
        pos = 0
        names = ("pairs_0", "pairs_1") if r in (2, 3) else ("pairs",)
        for nm in names:
            n = int(iout[pos])
            flat = iout[pos + 1:pos + 1 + 2 * n].tolist()
            out[nm] = [(flat[2 * i], flat[2 * i + 1]) for i in range(n)]
            pos += 1 + 2 * n
    if r != 0:
        out["proposals"] = {int(pv[i]): float(pd[i]) for i in range(npr.value)}
    return out


class plane_adjacency:
    """Context manager: the oracle uses the CPU path's plane adjacency (+-x,
    +-y; maxNeighbors = 6 over src/ExtreM.cpp:16-21 on a depth-1 field)
    inside the block, the 14-neighbour stencil outside it."""

    def __enter__(self):
        lib().oracle_set_adjacency(1)
        return self

    def __exit__(self, *a):
        lib().oracle_set_adjacency(0)


class persistence:
    """Context manager: inside the block the oracle's 2-saddle triplets keep
    only the maxima at least thr away from the saddle (the GPU fork's
    findAscPaths filter)."""

    def __init__(self, thr):
        self.thr = float(thr)

    def __enter__(self):
        lib().oracle_set_persistence(self.thr)
        return self

    def __exit__(self, *a):
        lib().oracle_set_persistence(0.0)
