"""Benchmark: corrected voxels/s of the ExtreM correction on one MI355X.

Workload (default "512"): one float32 field 512x512x512, a Gaussian random
field with power spectrum k^-3 (seed 3) normalised to [0, 1], xi = 1e-3,
decp the uniform quantiser 2*xi*rint(orig/(2*xi)).  A step is one full
correction call (main.cpp:65-446 semantics) on device-resident inputs; each
step gets its own pristine copy of decp, prepared before the timed region.
"256" is the 256^3 sum of 32 Gaussians (er 1e-4); "64" the 64^3 CPU-sized
case.  After the timed steps, the last step's corrected field and edit
counts are hashed and compared with the oracle's digest of the same
workload (tests/golden/*_digest.json) when one exists: "verified".

The roofline object is for the dominant single kernel: among the families
that time one kernel per event pair (every family except the sequences
incremental, slot_gather, propose and reserve, whose kernels rocprof lists
one by one, each below the dominant one), the family with the largest
summed time in the warm-up call (every family timed there).  The timed
steps record HIP events around that kernel's launches only, on the stream
it runs on.  achieved = algorithmic bytes per launch / average launch time,
with the per-unit bytes of ALG_BYTES.  traffic is the measured HBM bytes per
launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a
plain copy) from profiles/<kernel>_<workload>_traffic.md when present.
cpu_baseline is the CPU oracle on the 64^3 CPU-sized case, with the oracle's
time on the full workload (from the digest) beside it.

"batch" is the mixed-size batch: 64 float32 fields, shapes log-uniform
per axis from 32^3 to 384x384x256, Gaussian sums, er from {1e-2, 1e-3, 1e-4};
a step is one batch over an extremz_pool, timed beside the plain path (the
same fields one after another on one context).  "serve" is a Poisson stream
of jobs (cubic edges log-uniform 16..512, Gaussian sums and k^-3 random
fields, er from the same set) at offered loads below saturation; it reports
p50 / p99 job latency and the sustained voxels/s per load.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 512|256|64|batch|serve]
    python bench.py --calibrate-copy      # plain 1 GiB copies, for the counter scale
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "corrected voxels/s + % HBM roofline (byte-exact vs CPU); p50/p99 job latency"

# Algorithmic bytes per unit of the single-kernel families (unit, bytes), the
# compulsory bytes of each pass of the reference algorithm:
#   label_tile   per voxel: read decp f64 (8) + write both pointer labels (4 + 4)
#                + write the lower-star mask (2)
#   saddle_pass, saddle_orig  per saddle side: one frontier item, 15 x (8 B value
#                + 4 B label)
#   scan_propose, check_full, classify  per voxel: read f64 (8) + type (1) + mask (2)
#   label_chase  the walk of the pointers left by the tiles: no algorithmic bytes
#                (extra pointer-jumping rounds are excluded from the model)
# apply (30 B per touched voxel), check_incr and csr_fill are list-sized: their
# unit counts stay on the device, so they carry no byte model here.
ALG_BYTES = {
    "label_tile": ("voxel", 18),
    "saddle_pass": ("saddle side", 180),
    "saddle_orig": ("saddle side", 180),
    "scan_propose": ("voxel", 11),
    "check_full": ("voxel", 11),
    "classify": ("voxel", 11),
    "label_chase": ("voxel", 0),
}
LIST_SIZED = ("apply", "check_incr", "csr_fill")
SEQUENCES = ("incremental", "slot_gather", "propose", "reserve")
DIGESTS = {"512": "grf_512_digest.json", "256": "gauss32_256_digest.json"}


def workload(name):
    from extrem_z_amd import synth
    if name == "256":
        o, d, xi = synth.make_case("gauss", (256, 256, 256), seed=2, er=1e-4, n_gauss=32)
        desc = "gauss32-256^3-er1e-4"
    elif name == "512":
        o, d, xi = synth.make_case("grf", (512, 512, 512), seed=3, xi=1e-3)
        desc = "grf-k^-3-512^3-xi1e-3"
    elif name == "64":
        o, d, xi = synth.make_case("gauss", (64, 64, 64), seed=1, er=1e-3, n_gauss=8)
        desc = "gauss8-64^3-er1e-3"
    else:
        raise SystemExit(f"unknown workload {name}")
    return o, d, xi, desc


def units(fam, st, n):
    """Units one launch of family fam processes (host-known for ALG_BYTES)."""
    if ALG_BYTES[fam][0] == "saddle side":
        return st["n_saddle1"] + st["n_saddle2"]
    return n


def traffic_from_profiles(fam, workload_name):
    """HBM bytes per launch from profiles/<fam>_<workload>_traffic.md, if present."""
    path = os.path.join(ROOT, "profiles", f"{fam}_{workload_name}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(threads, workload_name):
    """The oracle on the 64^3 CPU-sized case (sum of 8 Gaussians + 1e-3 noise,
    seed 1, er 1e-3), plus the oracle's time on the full workload from the
    digest fixture (measured in the build container when it was written)."""
    from extrem_z_amd import synth
    from oracle import oracle
    o, d, xi = synth.config(1)
    t = time.perf_counter()
    f, c, st = oracle.correct(o, d, xi, threads=threads)
    dt = time.perf_counter() - t
    out = {"value": o.size / dt, "unit": "voxels/s", "cores": threads, "kind": "port",
           "sample": f"oracle (C++ restatement, OpenMP {threads} threads) on the 64^3 CPU case "
                     f"(8 Gaussians + 1e-3 noise, seed 1, er 1e-3): {dt:.2f} s, "
                     f"{st['outer_iters']} outer iterations"}
    dig = DIGESTS.get(workload_name)
    path = os.path.join(ROOT, "tests", "golden", dig) if dig else None
    if path and os.path.exists(path):
        ref = json.load(open(path))
        n = int(np.prod(ref["shape"]))
        out["full_workload"] = {"seconds": ref["oracle_seconds"], "threads": ref["oracle_threads"],
                                "voxels_per_s": n / ref["oracle_seconds"],
                                "source": f"tests/golden/{dig} (8-CPU build container)"}
    return out


def verify_digest(workload_name, fixed_bytes, counts_bytes, st):
    """Compares one corrected field with the oracle's digest of the workload."""
    import hashlib
    dig = DIGESTS.get(workload_name)
    path = os.path.join(ROOT, "tests", "golden", dig) if dig else None
    if not path or not os.path.exists(path):
        return None
    ref = json.load(open(path))
    same_f = hashlib.sha256(fixed_bytes).hexdigest() == ref["sha256_fixed"]
    same_c = hashlib.sha256(counts_bytes).hexdigest() == ref["sha256_counts"]
    same_s = all(st[k] == v for k, v in ref["stats"].items())
    return {"equal": bool(same_f and same_c and same_s), "fixed_sha256": same_f,
            "counts_sha256": same_c, "counters": same_s, "digest": f"tests/golden/{dig}"}


def calibrate_copy():
    from extrem_z_amd.device import DeviceBuffer, synchronize
    n = 1 << 30
    src = DeviceBuffer.from_numpy(np.ones(n, np.uint8))
    dst = DeviceBuffer(n)
    synchronize()
    t = time.perf_counter()
    for _ in range(5):
        dst.copy_from(src)
    synchronize()
    dt = (time.perf_counter() - t) / 5
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 5,
                      "GB_per_s": 2 * n / dt / 1e9}), flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(np.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def upload_fields(fields, copies):
    """Device copies of orig (float64) and `copies` working copies of decp."""
    from extrem_z_amd.device import DeviceBuffer, synchronize
    out = []
    for f in fields:
        orig, decp, xi = f[0], f[1], f[2]
        pristine = DeviceBuffer.from_numpy(decp.ravel())
        out.append({"shape": orig.shape[::-1], "xi": xi,
                    "orig": DeviceBuffer.from_numpy(orig.ravel()), "pristine": pristine,
                    "work": [pristine.empty_like() for _ in range(copies)], "N": orig.size})
    synchronize()
    return out


def reset(fields, k):
    from extrem_z_amd.device import synchronize
    for f in fields:
        f["work"][k].copy_from(f["pristine"])
    synchronize()


def fam_roofline(stats_list, fam):
    """Achieved GB/s of kernel family fam over a set of job stats."""
    unit, per_unit = ALG_BYTES[fam]
    launches = sum(s["kernel_launches"][fam] for s in stats_list)
    ms = sum(s["kernel_ms"][fam] for s in stats_list)
    alg = 0
    for s in stats_list:
        u = s["n_saddle1"] + s["n_saddle2"] if unit == "saddle side" else s["_N"]
        alg += per_unit * u * s["kernel_launches"][fam]
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": fam,
            "launches": launches, "avg_launch_ms": round(ms / max(1, launches), 4),
            "alg_bytes_per_launch": round(alg / max(1, launches)),
            "alg_bytes_unit": f"{per_unit} B/{unit}"}


def job_stats(job, N):
    import extrem_z_amd as ez
    d = job.stats.as_dict()
    d["_N"] = N
    return d


def run_batch(args):
    """Config 4: 64 mixed fields, pool vs the plain one-context path."""
    import extrem_z_amd as ez
    from extrem_z_amd import synth
    from extrem_z_amd.device import DeviceBuffer, synchronize

    t = time.perf_counter()
    fields = synth.mixed_batch(64)
    log(f"generated 64 fields in {time.perf_counter() - t:.1f} s")
    total = sum(f[0].size for f in fields)
    F = upload_fields(fields, 2)
    del fields
    counts = [DeviceBuffer.zeros(f["N"], np.uint8) for f in F]
    # plain path: one context, fields one after another
    ctx = ez.Context(0)
    big = max(range(len(F)), key=lambda i: F[i]["N"])
    reset(F, 0)
    ctx.correct_device(F[big]["orig"].ptr, F[big]["work"][0].ptr, F[big]["shape"],
                       F[big]["xi"], counts[big].ptr)
    plain_t = []
    for step in range(args.steps):
        reset(F, 0)
        t0 = time.perf_counter()
        for i, f in enumerate(F):
            ctx.correct_device(f["orig"].ptr, f["work"][0].ptr, f["shape"], f["xi"],
                               counts[i].ptr)
        synchronize()
        plain_t.append(time.perf_counter() - t0)
        log(f"plain pass {step}: {plain_t[-1]:.3f} s")
    ctx.close()
    # batched: a pool of lanes
    pool = ez.Pool(0, lanes=args.lanes)

    def jobs(timing):
        pool_opts = []
        for i, f in enumerate(F):
            pool_opts.append(pool.device_job(f["orig"].ptr, f["work"][1].ptr,
                                             f["shape"], f["xi"], counts[i].ptr))
        return pool_opts

    for w in range(max(1, args.warmup)):
        reset(F, 1)
        rc, arr = pool.batch(jobs(False))
        assert rc == 0, rc
        log(f"batch warm-up {w}: done")
    batch_t, lat, stats = [], [], []
    for step in range(args.steps):
        reset(F, 1)
        t0 = time.perf_counter()
        rc, arr = pool.batch(jobs(False))
        synchronize()
        batch_t.append(time.perf_counter() - t0)
        assert rc == 0, rc
        lat += [j.wait_ms + j.run_ms for j in arr]
        log(f"batch pass {step}: {batch_t[-1]:.3f} s")
    same = all(np.array_equal(f["work"][0].numpy(), f["work"][1].numpy()) for f in F)
    info = pool.info()
    pool.close()
    # kernel timing of one more batch pass (every family, events per lane)
    pool = ez.Pool(0, lanes=args.lanes, timing=True)
    reset(F, 1)
    rc, arr = pool.batch(jobs(True))
    stats = [job_stats(j, F[i]["N"]) for i, j in enumerate(arr)]
    pool.close()
    fam_ms = {k: sum(s["kernel_ms"][k] for s in stats) for k in ALG_BYTES}
    dom = max(fam_ms, key=fam_ms.get)
    plain = total / min(plain_t)
    value = total / min(batch_t)
    return {
        "metric": METRIC, "value": value, "unit": "voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": min(batch_t) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batch64-mixed-32^3..384x384x256-er{1e-2,1e-3,1e-4}",
                   "fields": len(F), "voxels": total, "lanes": args.lanes,
                   "plain_voxels_per_s": plain, "speedup_vs_plain": value / plain,
                   "plain_ms": min(plain_t) * 1e3, "identical_to_plain": same,
                   "job_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                   "peak_footprint_GiB": info["peak_footprint_bytes"] / 2 ** 30,
                   "kernel_ms_timed_pass": {k: round(v, 2) for k, v in fam_ms.items()}},
        "roofline": fam_roofline(stats, dom),
    }


def run_serve(args):
    """Config 5: Poisson arrivals at offered loads below saturation."""
    import threading
    import extrem_z_amd as ez
    from extrem_z_amd import synth
    from extrem_z_amd.device import synchronize

    t = time.perf_counter()
    S = synth.serve_set(args.serve_fields)
    log(f"generated {len(S)} jobs in {time.perf_counter() - t:.1f} s: edges "
        f"{[s[0].shape[0] for s in S]}")
    F = upload_fields(S, 0)
    del S
    M = args.serve_jobs
    order = [i % len(F) for i in range(M)]
    work = [F[i]["pristine"].empty_like() for i in order]
    rng = np.random.default_rng(5)
    pool = ez.Pool(0, lanes=args.lanes, hbm_budget=int(args.budget_gib * 2 ** 30))

    def make_jobs():
        for k, i in enumerate(order):
            work[k].copy_from(F[i]["pristine"])
        synchronize()
        return [pool.device_job(F[i]["orig"].ptr, work[k].ptr, F[i]["shape"],
                                F[i]["xi"]) for k, i in enumerate(order)]

    def run(rate):
        """Submits every job at Poisson times (rate jobs/s; inf = all at once)."""
        jobs = make_jobs()
        gaps = rng.exponential(1.0 / rate, size=M) if np.isfinite(rate) else np.zeros(M)
        arrive, done = {}, {}
        t0 = time.perf_counter()

        def feeder():
            due = t0
            for k, j in enumerate(jobs):
                due += gaps[k]
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tk = pool.submit(j)
                arrive[tk] = time.perf_counter()

        th = threading.Thread(target=feeder)
        th.start()
        while len(done) < M:
            r = pool.wait_any(timeout_ms=100000)
            assert r is not None, "no job finished within 100 s"
            tk, j = r
            assert j.status == 0, ez.job_error(j)
            done[tk] = time.perf_counter()
        th.join()
        lat = [(done[k] - arrive[k]) * 1e3 for k in done]
        span = max(done.values()) - t0
        vox = sum(F[i]["N"] for i in order)
        return {"offered_jobs_per_s": rate if np.isfinite(rate) else None,
                "voxels_per_s": vox / span, "jobs_per_s": M / span,
                "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "span_s": span}

    run(float("inf"))  # warm-up: lanes grow their buffers
    sat = run(float("inf"))
    log(f"saturation: {sat['jobs_per_s']:.2f} jobs/s, {sat['voxels_per_s']:.3e} voxels/s")
    loads = []
    for frac in args.loads:
        r = run(frac * sat["jobs_per_s"])
        r["load_fraction"] = frac
        loads.append(r)
        log(f"load {frac:.2f}: p50 {r['p50_ms']:.1f} ms, p99 {r['p99_ms']:.1f} ms, "
            f"{r['voxels_per_s']:.3e} voxels/s")
    info = pool.info()
    pool.close()
    # plain path over the same stream, one job after another
    ctx = ez.Context(0)
    jobs = make_jobs()
    t0 = time.perf_counter()
    for k, i in enumerate(order):
        ctx.correct_device(F[i]["orig"].ptr, work[k].ptr, F[i]["shape"], F[i]["xi"])
    synchronize()
    plain = sum(F[i]["N"] for i in order) / (time.perf_counter() - t0)
    ctx.close()
    top = max(loads, key=lambda r: r["load_fraction"]) if loads else sat
    return {
        "metric": METRIC, "value": top["voxels_per_s"], "unit": "voxels/s", "n_gpus": 1,
        "steps": M, "warmup": 1, "ms_per_step": top["span_s"] * 1e3 / M, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serve-poisson-{M}jobs-edges16..512-gauss+grf-er{{1e-2,1e-3,1e-4}}",
                   "distinct_fields": len(F), "lanes": args.lanes,
                   "budget_GiB": args.budget_gib, "saturation": sat, "loads": loads,
                   "latency_ms": {"p50": top["p50_ms"], "p99": top["p99_ms"],
                                  "load_fraction": top["load_fraction"] if loads else None},
                   "plain_voxels_per_s": plain,
                   "peak_footprint_GiB": info["peak_footprint_bytes"] / 2 ** 30},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--budget-gib", type=float, default=200.0)
    ap.add_argument("--serve-fields", type=int, default=16)
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--loads", type=float, nargs="*", default=[0.3, 0.6, 0.9])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: fields and jobs shard whole per GPU and this build "
              "runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    if args.workload in ("batch", "serve"):
        out = run_batch(args) if args.workload == "batch" else run_serve(args)
        print(json.dumps(out), flush=True)
        return

    import extrem_z_amd as ez
    from extrem_z_amd.device import DeviceBuffer, synchronize

    orig, decp, xi, desc = workload(args.workload)
    D, H, W = orig.shape
    N = orig.size
    log(f"workload {desc}, N={N}, steps={args.steps}, warmup={args.warmup}")
    d_orig = DeviceBuffer.from_numpy(orig.ravel())
    copies = [DeviceBuffer.from_numpy(decp.ravel()) for _ in range(args.steps + max(1, args.warmup))]
    d_counts = DeviceBuffer.zeros(N, np.uint8)
    synchronize()
    ctx = ez.Context(0)
    warm = None
    for i in range(max(1, args.warmup)):
        warm = ctx.correct_device(d_orig.ptr, copies[i].ptr, (W, H, D), xi,
                                  d_counts.ptr, timing=True)
        log(f"warmup {i} done, {warm['wall_ms']:.1f} ms")
    # dominant single kernel of the warm-up call: every family that times one
    # kernel per event pair (the sequences are several kernels each)
    singles = [k for k in warm["kernel_ms"] if k not in SEQUENCES]
    dom = max(singles, key=lambda k: warm["kernel_ms"][k])
    seq_top = max(SEQUENCES, key=lambda k: warm["kernel_ms"][k])
    mask = 1 << ez.KERNEL_FAMILIES.index(dom)
    synchronize()
    stats = []
    base = max(1, args.warmup)
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.correct_device(d_orig.ptr, copies[base + i].ptr, (W, H, D), xi,
                                d_counts.ptr, timing=mask)
        stats.append(st)
    synchronize()
    elapsed = time.perf_counter() - t0
    for st in stats:
        assert st["status"] == 0, st
    ms_per_step = elapsed * 1e3 / args.steps
    value = N * args.steps / elapsed
    # the last timed step's result against the oracle's digest (outside the
    # timed region)
    verified = verify_digest(args.workload, copies[base + args.steps - 1].numpy().tobytes(),
                             d_counts.numpy().tobytes(), stats[-1])
    log(f"verified against the oracle's digest: {verified}")

    launches = sum(s["kernel_launches"][dom] for s in stats)
    avg_ms = sum(s["kernel_ms"][dom] for s in stats) / max(1, launches)
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "traffic": traffic_from_profiles(dom, args.workload),
                "kernel": dom, "launches": launches, "avg_launch_ms": round(avg_ms, 4)}
    if dom in ALG_BYTES:
        unit, per_unit = ALG_BYTES[dom]
        n_units = units(dom, stats[-1], N)
        bytes_launch = per_unit * n_units
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        roofline.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "alg_bytes_per_launch": bytes_launch, "units_per_launch": n_units,
                         "alg_bytes_unit": f"{per_unit} B/{unit}"})
    else:  # a list-sized kernel: its unit counts stay on the device
        roofline.update({"achieved": None, "frac": None, "alg_bytes_per_launch": None})
    roofline = {k: roofline[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic",
                                         "kernel", "launches", "avg_launch_ms",
                                         "alg_bytes_per_launch", "units_per_launch",
                                         "alg_bytes_unit") if k in roofline}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "verified": verified,
        "config": {"workload": desc, "shape": [W, H, D], "xi": xi,
                   "outer_iters": stats[-1]["outer_iters"],
                   "c_loop_iters": stats[-1]["c_loop_iters"],
                   "label_passes": stats[-1]["label_passes"],
                   "inc_passes": stats[-1]["inc_passes"],
                   "edited": stats[-1]["edited"],
                   "host_syncs": stats[-1]["host_syncs"],
                   "call_latency_ms": {"p50": pct([s["wall_ms"] for s in stats], 50),
                                       "p99": pct([s["wall_ms"] for s in stats], 99)},
                   "kernel_ms_warmup_call": {k: round(v, 3) for k, v in warm["kernel_ms"].items()},
                   "launches_per_call": warm["kernel_launches"],
                   "largest_sequence_family": seq_top},
        "roofline": roofline,
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on the 64^3 CPU case)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_threads, args.workload)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
