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
        L.oracle_bitmap3.argtypes = [vp, i64, vp]
        L.oracle_bitmap3.restype = i64
        L.oracle_cost.argtypes = [vp, vp, vp, vp, i64, dbl, vp, vp, vp]
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


def correct(orig, decp, xi, max_c_iters=0, max_outer=0, threads=0, fixed_step_q=0):
    """Returns (fixed, counts, stats dict).  Inputs are not modified.
    fixed_step_q > 0 selects the fixed-step rule set (step -xi/q, threshold q)."""
    o = np.ascontiguousarray(orig, dtype=np.float64)
    d = np.array(decp, dtype=np.float64, order="C", copy=True)
    counts = np.zeros(o.shape, np.uint8)
    st = np.zeros(12, np.int64)
    lib().oracle_correct_rule(_p(o), _p(d), *_dims(o), float(xi), int(max_c_iters),
                              int(max_outer), _p(counts), _p(st), int(threads),
                              int(fixed_step_q))
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
