"""The three large workloads on the GPU, at their full sizes.

- 512^3 Gaussian random field (power spectrum k^-3, normalised to [0, 1],
  xi = 1e-3): status, the bound, the original's types and saddle stars kept,
  every termination counter of main.cpp:277-279 recomputed independently on
  the CPU (tests/topo.py) and zero, the oracle's digest of the same field
  when present, and incremental CP passes equal to full passes.
- The mixed batch: 64 fields, shapes log-uniform per axis from 32^3 to
  384x384x256, Gaussian sums, er in {1e-2, 1e-3, 1e-4}: the pool's results
  equal the plain single-field path byte for byte, and the oracle's on the
  fields up to 64^3.
- The serving stream: jobs of edges 16..512 (Gaussian sums and k^-3 random
  fields) arriving as a Poisson stream into a pool with a 200 GiB budget:
  each result equals the same job corrected alone.
"""
import hashlib
import json
import os
import threading
import time

import numpy as np
import pytest

from tests.topo import violations_fast

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COUNTERS = ("status", "c_loop_iters", "outer_iters", "r_loop_calls", "s_loop_calls",
            "saddle_loop_calls", "c_loops_calls", "edited", "n_max", "n_min", "n_saddle1",
            "n_saddle2")


def ez():
    import extrem_z_amd as m
    return m


def within_bound(orig, fixed, xi):
    tol = xi + np.spacing(np.abs(orig) + xi)
    return bool(np.all(np.abs(fixed - orig) <= tol))


def say(msg):
    print(f"[workloads] {msg}", flush=True)


@pytest.fixture(scope="module")
def grf512():
    from extrem_z_amd import synth
    t = time.perf_counter()
    case = synth.config(3)
    say(f"512^3 GRF generated in {time.perf_counter() - t:.1f} s")
    return case


@pytest.mark.timeout(900)
def test_grf_512_full_size(grf512):
    orig, decp, xi = grf512
    say("512^3 GRF: correction")
    fixed, counts, st = ez().correct(orig.astype(np.float32), decp, xi)
    say(f"512^3 GRF: status {st['status']}, {st['outer_iters']} outer iterations, "
        f"{st['edited']} edited, {st['inc_passes']} incremental CP passes")
    assert st["status"] == 0
    assert within_bound(orig, fixed, xi)
    dig = os.path.join(GOLDEN, "grf_512_digest.json")
    if os.path.exists(dig):
        ref = json.load(open(dig))
        assert ref["xi"] == xi
        assert {k: st[k] for k in COUNTERS} == ref["stats"]
        assert hashlib.sha256(counts.tobytes()).hexdigest() == ref["sha256_counts"]
        assert hashlib.sha256(fixed.tobytes()).hexdigest() == ref["sha256_fixed"]
        say("512^3 GRF: equal to the oracle's digest")
    t0, m0 = ez().classify(orig)
    t1, m1 = ez().classify(fixed)
    assert np.array_equal(t0, t1)
    sad = (t0 >= 1) & (t0 <= 3)
    assert np.array_equal(m0[sad], m1[sad])
    del t0, m0, t1, m1, sad
    say("512^3 GRF: termination counters on the CPU")
    v = violations_fast(orig, fixed)
    assert all(x == 0 for x in v.values()), v
    say(f"512^3 GRF: counters {v}")


@pytest.mark.timeout(900)
def test_grf_512_incremental_equals_full(grf512, monkeypatch):
    """Every incremental CP pass is checked against the full pass inside the
    engine (EXTREMZ_VERIFY_INC), and the result equals the run with full
    passes only."""
    orig, decp, xi = grf512
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "1")
    say("512^3 GRF: correction with every incremental pass verified")
    f1, c1, s1 = ez().correct(orig, decp, xi)
    assert s1["status"] == 0 and s1["inc_passes"] > 0
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "0")
    monkeypatch.setenv("EXTREMZ_FULL_PASSES", "1")
    say("512^3 GRF: correction with full CP passes only")
    f2, c2, s2 = ez().correct(orig, decp, xi)
    assert s2["status"] == 0 and s2["inc_passes"] == 0
    assert f1.tobytes() == f2.tobytes() and np.array_equal(c1, c2)
    assert {k: s1[k] for k in COUNTERS} == {k: s2[k] for k in COUNTERS}


@pytest.mark.timeout(900)
def test_mixed_batch_64_fields():
    from extrem_z_amd import synth
    from oracle import oracle
    t = time.perf_counter()
    fields = synth.mixed_batch(64)
    say(f"mixed batch: 64 fields, {sum(f[0].size for f in fields)} voxels, generated in "
        f"{time.perf_counter() - t:.1f} s")
    alone = []
    for i, (orig, decp, xi) in enumerate(fields):
        fixed, counts, st = ez().correct(orig.astype(np.float32), decp, xi)
        assert st["status"] == 0, i
        assert within_bound(orig, fixed, xi), i
        alone.append((fixed, counts))
    say("mixed batch: plain path done")
    with ez().Pool(0, lanes=8) as pool:
        jobs, keep = [], []
        for orig, decp, xi in fields:
            j, k = pool.host_job(orig.astype(np.float32), decp.copy(), xi)
            jobs.append(j)
            keep.append(k)
        rc, arr = pool.batch(jobs)
        assert rc == 0
        for i in range(len(fields)):
            assert arr[i].status == 0, i
            assert keep[i][1].tobytes() == alone[i][0].tobytes(), i
            assert np.array_equal(keep[i][2], alone[i][1]), i
    say("mixed batch: pool results equal the plain path")
    small = [i for i, f in enumerate(fields) if f[0].size <= 64 ** 3]
    assert small
    for i in small:
        orig, decp, xi = fields[i]
        f0, c0, _ = oracle.correct(orig, decp, xi)
        assert f0.tobytes() == alone[i][0].tobytes(), i
        assert np.array_equal(c0, alone[i][1]), i
    say(f"mixed batch: {len(small)} fields up to 64^3 equal the oracle")


@pytest.mark.timeout(900)
def test_serving_stream_in_mix_equals_alone():
    from extrem_z_amd import synth
    from extrem_z_amd.device import DeviceBuffer, synchronize
    t = time.perf_counter()
    S = synth.serve_set(16)
    say(f"serving stream: {len(S)} jobs, edges {[s[0].shape[0] for s in S]}, generated in "
        f"{time.perf_counter() - t:.1f} s")
    dev = []
    for orig, decp, xi, kind in S:
        dev.append((DeviceBuffer.from_numpy(orig.ravel()), DeviceBuffer.from_numpy(decp.ravel()),
                    orig.shape[::-1], xi))
    synchronize()
    alone = []
    with ez().Context(0) as ctx:
        for o, d, shape, xi in dev:
            w = d.empty_like()
            w.copy_from(d)
            c = DeviceBuffer.zeros(int(np.prod(shape)), np.uint8)
            st = ctx.correct_device(o.ptr, w.ptr, shape, xi, c.ptr)
            assert st["status"] == 0
            alone.append((w.numpy(), c.numpy()))
            del w, c
    say("serving stream: every job corrected alone")
    for (orig, _, xi, _), (fixed, _) in zip(S, alone):
        assert within_bound(orig.ravel(), fixed, xi)
    order = [k % len(S) for k in range(2 * len(S))]
    work = []
    for i in order:
        w = dev[i][1].empty_like()
        w.copy_from(dev[i][1])
        work.append((w, DeviceBuffer.zeros(int(np.prod(dev[i][2])), np.uint8)))
    synchronize()
    rng = np.random.default_rng(5)
    gaps = rng.exponential(0.02, size=len(order))
    with ez().Pool(0, lanes=8, hbm_budget=200 << 30) as pool:
        arrive, done, tickets = {}, {}, {}

        def feeder():
            for k, i in enumerate(order):
                time.sleep(gaps[k])
                o, _, shape, xi = dev[i]
                job = pool.device_job(o.ptr, work[k][0].ptr, shape, xi, work[k][1].ptr)
                tk = pool.submit(job)
                arrive[tk] = time.perf_counter()
                tickets[tk] = k

        th = threading.Thread(target=feeder)
        th.start()
        while len(done) < len(order):
            r = pool.wait_any(timeout_ms=120000)
            assert r is not None, "no job finished within two minutes"
            tk, job = r
            assert job.status == 0, ez().job_error(job)
            done[tk] = time.perf_counter()
        th.join()
        peak = pool.info()["peak_footprint_bytes"]
    lat = sorted((done[t] - arrive[t]) * 1e3 for t in done)
    say(f"serving stream: {len(order)} jobs, latency p50 {lat[len(lat) // 2]:.1f} ms, max "
        f"{lat[-1]:.1f} ms, peak footprint {peak / 2 ** 30:.1f} GiB")
    assert peak <= 200 << 30
    for k, i in enumerate(order):
        assert work[k][0].numpy().tobytes() == alone[i][0].tobytes(), (k, i)
        assert np.array_equal(work[k][1].numpy(), alone[i][1]), (k, i)
