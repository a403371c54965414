"""MI355X ExtreM correction engine: Python view of the C ABI (include/extremz.h).

The engine is libextremz.so (HIP kernels for gfx950 behind a C ABI).  This
module only loads it and converts numpy arrays; there is no CPU fallback: a
missing library or device raises.

Functions mirror the reference's correction path:
  correct()        the main.cpp:65-446 correction call (decp -> fixed)
  classify()       classifyAllVertices (src/ExtreM.cpp:1088-1097)
  labels()         ComputeDescending/AscendingManifold (src/ExtreM.cpp:144-183, 502-541)
  encode()         cost() edit streams (src/ExtreM.cpp:272-389)
  cost()           cost()'s zstd container and statistics (src/ExtreM.cpp:252-389)
  Pool             batched and served corrections over several HIP streams
                   (no reference counterpart: the reference corrects one
                   field per process)
  getdata()        the raw float64 reader + bound of getdata (src/ExtreM.cpp:95-138)
  saveArrayToBin() the raw float64 writer (src/ExtreM.cpp:2918-2928)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libextremz.so")

STAT_KEYS = ("status", "c_loop_iters", "outer_iters", "r_loop_calls", "s_loop_calls",
             "saddle_loop_calls", "c_loops_calls", "edited", "n_max", "n_min", "n_saddle1",
             "n_saddle2", "label_passes", "inc_passes", "host_syncs", "frontier_items")
KERNEL_FAMILIES = ("classify", "check_full", "check_incr", "label_tile", "incremental",
                   "label_chase", "saddle_pass", "slot_gather", "propose", "apply", "reserve",
                   "scan_propose", "saddle_orig", "csr_fill", "_14", "_15")
N_FAMILIES = 16
TIME_ALL = (1 << len(KERNEL_FAMILIES)) - 1

ERRORS = {-3: "iteration cap reached", -5: "HIP runtime error", -11: "not finished",
          -12: "device allocation failed", -19: "no gfx950 device", -22: "invalid argument",
          -38: "libzstd.so.1 not available"}
EAGAIN = -11


class ExtremzError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"extremz error {code} ({ERRORS.get(code, '?')}): {msg}")


class Opts(ctypes.Structure):
    _fields_ = [("max_c_iters", ctypes.c_int), ("max_outer", ctypes.c_int),
                ("timing", ctypes.c_int), ("fixed_step_q", ctypes.c_int),
                ("adjacency", ctypes.c_int), ("rule_set", ctypes.c_int),
                ("persistence", ctypes.c_double)]


ADJ_STENCIL14 = 0  # the 14-neighbour stencil
ADJ_PLANE4 = 1     # the CPU path's plane adjacency (+-x, +-y; depth-1 fields)
RULES_CPU = 0      # the CPU path's midpoint / snap rules (or the fixed step)
RULES_F32 = 1      # the float32 fork's rule set


class Stats(ctypes.Structure):
    _fields_ = [("counts", ctypes.c_int64 * 16), ("kernel_ms", ctypes.c_double * N_FAMILIES),
                ("kernel_launches", ctypes.c_int64 * N_FAMILIES), ("wall_ms", ctypes.c_double)]

    def as_dict(self):
        d = {k: int(self.counts[i]) for i, k in enumerate(STAT_KEYS)}
        d["kernel_ms"] = {k: float(self.kernel_ms[i]) for i, k in enumerate(KERNEL_FAMILIES)
                          if not k.startswith("_")}
        d["kernel_launches"] = {k: int(self.kernel_launches[i])
                                for i, k in enumerate(KERNEL_FAMILIES) if not k.startswith("_")}
        d["wall_ms"] = float(self.wall_ms)
        return d


class Job(ctypes.Structure):
    """extremz_job: one correction handed to a Pool."""
    _fields_ = [("orig", ctypes.c_void_p), ("decp", ctypes.c_void_p), ("counts", ctypes.c_void_p),
                ("W", ctypes.c_int), ("H", ctypes.c_int), ("D", ctypes.c_int),
                ("orig_dtype", ctypes.c_int), ("on_host", ctypes.c_int), ("status", ctypes.c_int),
                ("xi", ctypes.c_double), ("lane", ctypes.c_int), ("reserved", ctypes.c_int),
                ("wait_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("stats", Stats)]


class CostStats(ctypes.Structure):
    _fields_ = [("changed", ctypes.c_int64), ("snapped", ctypes.c_int64),
                ("edit_ratio", ctypes.c_double), ("psnr", ctypes.c_double),
                ("fixed_psnr", ctypes.c_double), ("index_bytes", ctypes.c_int64),
                ("index_zst_bytes", ctypes.c_int64), ("edits_bytes", ctypes.c_int64),
                ("edits_zst_bytes", ctypes.c_int64), ("ocr", ctypes.c_double),
                ("cr", ctypes.c_double), ("obr", ctypes.c_double), ("br", ctypes.c_double)]


class TreeCmp(ctypes.Structure):
    _fields_ = [("n_max_a", ctypes.c_int64), ("n_max_b", ctypes.c_int64),
                ("rounds_a", ctypes.c_int64), ("rounds_b", ctypes.c_int64),
                ("pair_mismatch", ctypes.c_int64), ("branch_mismatch", ctypes.c_int64)]


class PoolInfo(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_int64), ("footprint_bytes", ctypes.c_int64),
                ("peak_footprint_bytes", ctypes.c_int64), ("submitted", ctypes.c_int64),
                ("completed", ctypes.c_int64), ("running", ctypes.c_int64),
                ("queued", ctypes.c_int64), ("lanes", ctypes.c_int), ("reserved", ctypes.c_int)]


_lib = None


def lib():
    """Loads libextremz.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ExtremzError(-19, f"{LIB_PATH} is missing; run extrem_z_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.extremz_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
        L.extremz_ctx_destroy.argtypes = [vp]
        L.extremz_correct.argtypes = [vp, i32, vp, i32, i32, i32, dbl, vp, vp, vp]
        L.extremz_correct_device.argtypes = [vp, vp, vp, i32, i32, i32, dbl, vp, vp, vp]
        L.extremz_classify.argtypes = [vp, i32, i32, i32, vp, vp]
        L.extremz_labels.argtypes = [vp, i32, i32, i32, vp, vp]
        L.extremz_label_pass.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp]
        L.extremz_lut.argtypes = [vp, i32]
        L.extremz_encode.argtypes = [vp, vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp]
        L.extremz_last_error.restype = ctypes.c_char_p
        L.extremz_cost.argtypes = [vp, vp, vp, vp, i64, dbl, i64, i64, i32, vp, i64, vp, i64, vp]
        L.extremz_zstd_bound.argtypes = [i64]
        L.extremz_zstd_bound.restype = i64
        L.extremz_zstd_compress.argtypes = [vp, i64, vp, i64, i32]
        L.extremz_zstd_compress.restype = i64
        L.extremz_zstd_decompress.argtypes = [vp, i64, vp, i64]
        L.extremz_zstd_decompress.restype = i64
        L.extremz_split_tree.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, vp, i32, i32, vp]
        L.extremz_compare_split_trees.argtypes = [vp, vp, i32, i32, i32, vp]
        L.extremz_pool_create.argtypes = [i32, i32, i64, vp, ctypes.POINTER(vp)]
        L.extremz_pool_destroy.argtypes = [vp]
        L.extremz_pool_info_get.argtypes = [vp, vp]
        L.extremz_batch_correct.argtypes = [vp, vp, i32]
        L.extremz_serve_submit.argtypes = [vp, vp, ctypes.POINTER(i64)]
        L.extremz_serve_wait.argtypes = [vp, i64, i32]
        L.extremz_serve_wait_any.argtypes = [vp, i32, ctypes.POINTER(i64)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise ExtremzError(rc, lib().extremz_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(a):
    if a.ndim != 3:
        raise ValueError("fields are 3-D arrays shaped (D, H, W)")
    D, H, W = a.shape
    return W, H, D


def lut():
    """The classification table in LUT.bin layout (24 800 int32)."""
    out = np.zeros(24800, np.int32)
    n = lib().extremz_lut(_p(out), out.size)
    return out[:n]


def classify(field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    t = np.empty(f.shape, np.uint8)
    m = np.empty(f.shape, np.uint16)
    _check(lib().extremz_classify(_p(f), *_dims(f), _p(t), _p(m)))
    return t, m


def labels(field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    ds = np.empty(f.shape, np.int32)
    asc = np.empty(f.shape, np.int32)
    _check(lib().extremz_labels(_p(f), *_dims(f), _p(ds), _p(asc)))
    return ds, asc


def label_pass(field):
    """Types, lower masks and both root labels from one tile pass (the
    original field's pass of the correction).  Returns (types, masks, ds, as)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    t = np.empty(f.shape, np.uint8)
    m = np.empty(f.shape, np.uint16)
    ds = np.empty(f.shape, np.int32)
    asc = np.empty(f.shape, np.int32)
    _check(lib().extremz_label_pass(_p(f), *_dims(f), _p(ds), _p(asc), _p(t), _p(m)))
    return t, m, ds, asc


def _timing_mask(timing):
    """True = every kernel family, False = none, or a bit mask over KERNEL_FAMILIES
    (or an iterable of family names)."""
    if timing is True:
        return TIME_ALL
    if not timing:
        return 0
    if isinstance(timing, int):
        return timing & TIME_ALL
    return sum(1 << KERNEL_FAMILIES.index(n) for n in timing)


def correct(orig, decp, xi, max_c_iters=0, max_outer=0, timing=False, fixed_step_q=0,
            adjacency=ADJ_STENCIL14, rule_set=RULES_CPU, persistence=0.0):
    """Corrects decp against orig within xi.  Returns (fixed, counts, stats).

    orig may be float32 or float64 (promoted exactly); inputs are not modified.
    timing selects the kernel families timed with HIP events (see _timing_mask).
    fixed_step_q > 0 selects the fixed-step rule set (step -xi/q, threshold q).
    adjacency=ADJ_PLANE4 selects the CPU path's plane adjacency (depth 1);
    rule_set=RULES_F32 the float32 fork's rule set; persistence > 0 the GPU
    fork's persistence filter of the 2-saddle triplets (absolute threshold).
    """
    o = np.ascontiguousarray(orig)
    if o.dtype == np.float32:
        dt = 0
    else:
        o = np.ascontiguousarray(o, dtype=np.float64)
        dt = 1
    d = np.array(decp, dtype=np.float64, order="C", copy=True)
    if d.shape != o.shape:
        raise ValueError("orig and decp shapes differ")
    counts = np.zeros(o.shape, np.uint8)
    st = Stats()
    opts = Opts(int(max_c_iters), int(max_outer), _timing_mask(timing), int(fixed_step_q),
                int(adjacency), int(rule_set), float(persistence))
    rc = lib().extremz_correct(_p(o), dt, _p(d), *_dims(o), float(xi), ctypes.byref(opts),
                               ctypes.byref(st), _p(counts))
    _check(rc)
    return d, counts, st.as_dict()


class Context:
    """A device context: one HIP stream plus work buffers (extremz_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().extremz_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h

    def correct_device(self, d_orig, d_decp, shape, xi, d_counts=0, max_c_iters=0, max_outer=0,
                       timing=False, fixed_step_q=0):
        """Device pointers (ints): d_orig float64, d_decp float64 edited in place."""
        W, H, D = shape
        st = Stats()
        opts = Opts(int(max_c_iters), int(max_outer), _timing_mask(timing), int(fixed_step_q))
        rc = lib().extremz_correct_device(self.handle, ctypes.c_void_p(d_orig),
                                          ctypes.c_void_p(d_decp), W, H, D, float(xi),
                                          ctypes.byref(opts), ctypes.byref(st),
                                          ctypes.c_void_p(d_counts or None))
        _check(rc)
        return st.as_dict()

    def close(self):
        if self.handle:
            lib().extremz_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Pool:
    """Lanes (HIP streams + work buffers) on one device with HBM-budget admission
    (extremz_pool).  batch() corrects many fields at once; submit() / wait()
    serve jobs as they arrive.  Every job's result equals the single-field
    correct() result byte for byte."""

    def __init__(self, device=0, lanes=8, hbm_budget=0, max_c_iters=0, max_outer=0, timing=False,
                 fixed_step_q=0):
        h = ctypes.c_void_p()
        opts = Opts(int(max_c_iters), int(max_outer), _timing_mask(timing), int(fixed_step_q))
        _check(lib().extremz_pool_create(int(device), int(lanes), int(hbm_budget),
                                         ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        self._held = {}  # ticket -> (Job, arrays kept alive while the job runs)

    @staticmethod
    def host_job(orig, decp, xi):
        """A Job over host arrays: decp (float64) is corrected in place, counts
        is filled.  Returns (job, keep) where keep holds the arrays."""
        o = np.ascontiguousarray(orig)
        if o.dtype != np.float32:
            o = np.ascontiguousarray(o, dtype=np.float64)
        if decp.dtype != np.float64 or not decp.flags.c_contiguous:
            raise ValueError("decp must be a C-contiguous float64 array (corrected in place)")
        if decp.shape != o.shape:
            raise ValueError("orig and decp shapes differ")
        counts = np.zeros(o.shape, np.uint8)
        W, H, D = _dims(o)
        j = Job(orig=o.ctypes.data, decp=decp.ctypes.data, counts=counts.ctypes.data, W=W, H=H,
                D=D, orig_dtype=0 if o.dtype == np.float32 else 1, on_host=1, xi=float(xi))
        return j, (o, decp, counts)

    @staticmethod
    def device_job(d_orig, d_decp, shape, xi, d_counts=0, orig_dtype=1):
        """A Job over device pointers (ints); orig_dtype 0 = float32, 1 = float64."""
        W, H, D = shape
        return Job(orig=d_orig, decp=d_decp, counts=d_counts or None, W=W, H=H, D=D,
                   orig_dtype=int(orig_dtype), on_host=0, xi=float(xi))

    def batch(self, jobs):
        """Runs a ctypes array / list of Job; returns the first failing status (0 if none)."""
        arr = (Job * len(jobs))(*jobs) if not isinstance(jobs, ctypes.Array) else jobs
        rc = lib().extremz_batch_correct(self.handle, arr, len(arr))
        return rc, arr

    def submit(self, job, keep=None):
        t = ctypes.c_int64(0)
        _check(lib().extremz_serve_submit(self.handle, ctypes.byref(job), ctypes.byref(t)))
        self._held[t.value] = (job, keep)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """Returns the Job once finished (status in job.status), or None on time-out."""
        rc = lib().extremz_serve_wait(self.handle, int(ticket), int(timeout_ms))
        if rc == EAGAIN:
            return None
        held = self._held.pop(ticket, None)
        if held is None:  # another thread received this ticket's status
            raise ExtremzError(rc, lib().extremz_last_error().decode(errors="replace"))
        return held[0]

    def wait_any(self, timeout_ms=-1):
        """Returns (ticket, Job) of the next finished job, or None on time-out."""
        t = ctypes.c_int64(0)
        rc = lib().extremz_serve_wait_any(self.handle, int(timeout_ms), ctypes.byref(t))
        if rc == EAGAIN:
            return None
        job, _ = self._held.pop(t.value)
        return t.value, job

    def info(self):
        i = PoolInfo()
        _check(lib().extremz_pool_info_get(self.handle, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in PoolInfo._fields_ if k != "reserved"}

    def close(self):
        if self.handle:
            lib().extremz_pool_destroy(self.handle)
            self.handle = None
            self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def job_error(job):
    """ExtremzError for a finished Job with a non-zero status, else None."""
    if job.status == 0:
        return None
    return ExtremzError(job.status, lib().extremz_last_error().decode(errors="replace"))


def encode(orig, decp_before, decp_after, counts, xi):
    """cost() streams: (index_bitmap uint8, edits f64, mantissas u64, exponents i32)."""
    o = np.ascontiguousarray(orig, dtype=np.float64).ravel()
    b = np.ascontiguousarray(decp_before, dtype=np.float64).ravel()
    a = np.ascontiguousarray(decp_after, dtype=np.float64).ravel()
    c = np.ascontiguousarray(counts, dtype=np.uint8).ravel()
    n = o.size
    bm = np.zeros((3 * n + 7) // 8, np.uint8)
    edits = np.zeros(n, np.float64)
    mant = np.zeros(n, np.uint64)
    expo = np.zeros(n, np.int32)
    ne = ctypes.c_int64(0)
    _check(lib().extremz_encode(_p(o), _p(b), _p(a), _p(c), n, float(xi), _p(bm), _p(edits),
                                _p(mant), _p(expo), ctypes.byref(ne)))
    k = ne.value
    return bm, edits[:k].copy(), mant[:k].copy(), expo[:k].copy()


def split_tree(field):
    """compute_MergeT of one field: dict with per-maximum (ascending (value, id))
    'max', 'pair_saddle', 'parent' vertex ids, 'branch' vertex lists, 'rounds'."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    cap = f.size + 1
    mx, ps, pa = (np.zeros(cap, np.int32) for _ in range(3))
    off = np.zeros(cap + 1, np.int32)
    bv = np.zeros(3 * cap, np.int32)
    rounds = ctypes.c_int(0)
    n = lib().extremz_split_tree(_p(f), *_dims(f), _p(mx), _p(ps), _p(pa), _p(off), _p(bv), cap,
                                 bv.size, ctypes.byref(rounds))
    if n < 0:
        _check(n if n > -(f.size + 2) * 4 else -22)
    return {"max": mx[:n].copy(), "pair_saddle": ps[:n].copy(), "parent": pa[:n].copy(),
            "branch": [bv[off[i]:off[i + 1]].tolist() for i in range(n)],
            "rounds": rounds.value}


def compare_split_trees(a, b):
    """Branch / pair comparison of main.cpp:480-504 between two fields."""
    fa = np.ascontiguousarray(a, dtype=np.float64)
    fb = np.ascontiguousarray(b, dtype=np.float64)
    if fa.shape != fb.shape:
        raise ValueError("shapes differ")
    out = TreeCmp()
    _check(lib().extremz_compare_split_trees(_p(fa), _p(fb), *_dims(fa), ctypes.byref(out)))
    return {k: getattr(out, k) for k, _ in TreeCmp._fields_}


def zstd_compress(data, level=3):
    """One zstd frame of the bytes of data (host only)."""
    src = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else \
        np.ascontiguousarray(data).view(np.uint8).ravel()
    cap = lib().extremz_zstd_bound(src.size)
    if cap < 0:
        raise ExtremzError(int(cap), "zstd")
    dst = np.empty(max(1, cap), np.uint8)
    n = lib().extremz_zstd_compress(_p(src), src.size, _p(dst), dst.size, int(level))
    if n < 0:
        raise ExtremzError(int(n), "zstd compress")
    return dst[:n].tobytes()


def zstd_decompress(frame):
    """Content of one zstd frame (with its content size, as zstd -f writes)."""
    src = np.frombuffer(bytes(frame), np.uint8)
    need = lib().extremz_zstd_decompress(_p(src), src.size, None, 0)
    if need < 0:
        raise ExtremzError(int(need), "zstd frame")
    dst = np.empty(max(1, need), np.uint8)
    n = lib().extremz_zstd_decompress(_p(src), src.size, _p(dst), dst.size)
    if n < 0:
        raise ExtremzError(int(n), "zstd decompress")
    return dst[:n].tobytes()


def cost(orig, decp_before, decp_after, counts, xi, original_bytes=0, compressed_bytes=0,
         level=3):
    """cost(): (stats dict, index bitmap zstd frame, snapped-edits zstd frame)."""
    o = np.ascontiguousarray(orig, dtype=np.float64).ravel()
    b = np.ascontiguousarray(decp_before, dtype=np.float64).ravel()
    a = np.ascontiguousarray(decp_after, dtype=np.float64).ravel()
    c = np.ascontiguousarray(counts, dtype=np.uint8).ravel()
    n = o.size
    icap = lib().extremz_zstd_bound((3 * n + 7) // 8)
    ecap = lib().extremz_zstd_bound(8 * n)
    if icap < 0 or ecap < 0:
        raise ExtremzError(-38, "libzstd.so.1")
    iz = np.empty(icap, np.uint8)
    ez_ = np.empty(ecap, np.uint8)
    st = CostStats()
    _check(lib().extremz_cost(_p(o), _p(b), _p(a), _p(c), n, float(xi), int(original_bytes),
                              int(compressed_bytes), int(level), _p(iz), iz.size, _p(ez_),
                              ez_.size, ctypes.byref(st)))
    d = {k: getattr(st, k) for k, _ in CostStats._fields_}
    return d, iz[:st.index_zst_bytes].tobytes(), ez_[:st.edits_zst_bytes].tobytes()


def getdata(filename, shape, er):
    """Raw float64 field + bound (max-min)*er, getdata of src/ExtreM.cpp:95-138.

    shape is (W, H, D); a size mismatch raises (the reference prints and
    returns, src/ExtreM.cpp:108-111).
    """
    W, H, D = shape
    a = np.fromfile(filename, dtype="<f8")
    if a.size != W * H * D:
        raise ValueError("File size does not match expected data size.")
    a = a.reshape(D, H, W)
    return a, float((a.max() - a.min()) * er)


def saveArrayToBin(arr, filename):
    """Raw little-endian float64, no header (src/ExtreM.cpp:2918-2928)."""
    np.ascontiguousarray(arr, dtype="<f8").tofile(filename)
