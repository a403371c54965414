"""Batched and served corrections (extremz_pool): every job's result equals the
single-field correction byte for byte, alone or in the mix, under any lane
count and memory budget; the budget bounds the pool's device footprint."""
import ctypes
import threading
import time

import numpy as np
import pytest

from oracle import oracle
from tests.conftest import make_field

pytestmark = pytest.mark.gpu

MIX = [
    ((16, 16, 16), 31, "gauss", 1e-3),
    ((40, 24, 12), 32, "noise", 1e-2),
    ((33, 17, 5), 33, "noise", 5e-2),
    ((48, 48, 32), 34, "gauss", 1e-4),
    ((20, 36, 28), 35, "gauss", 1e-2),
    ((7, 1, 9), 36, "noise", 1e-1),
    ((64, 40, 24), 37, "gauss", 1e-3),
    ((24, 20, 16), 38, "noise", 2e-2),
    ((56, 12, 44), 39, "gauss", 1e-3),
    ((30, 30, 30), 40, "noise", 1e-2),
]


def ez():
    import extrem_z_amd as m
    return m


@pytest.fixture(scope="module")
def cases():
    out = []
    for shape, seed, kind, er in MIX:
        orig, decp, xi = make_field(shape, seed, kind, er)
        fixed, counts, st = ez().correct(orig.astype(np.float32), decp, xi)
        assert st["status"] == 0
        out.append((orig, decp, xi, fixed, counts))
    return out


def test_batch_matches_single_field_and_oracle(cases):
    with ez().Pool(0, lanes=4) as pool:
        jobs, keep = [], []
        for orig, decp, xi, _, _ in cases:
            j, k = pool.host_job(orig.astype(np.float32), decp.copy(), xi)
            jobs.append(j)
            keep.append(k)
        rc, arr = pool.batch(jobs)
        assert rc == 0
        for i, (orig, decp, xi, fixed, counts) in enumerate(cases):
            assert arr[i].status == 0
            assert keep[i][1].tobytes() == fixed.tobytes(), i
            assert np.array_equal(keep[i][2], counts), i
            assert 0 <= arr[i].lane < 4
        info = pool.info()
        assert info["completed"] == len(cases) and info["running"] == 0 and info["queued"] == 0
    for i in (0, 2, 5, 7):  # the oracle on the smaller fields
        orig, decp, xi, fixed, counts = cases[i]
        f0, c0, _ = oracle.correct(orig, decp, xi)
        assert f0.tobytes() == fixed.tobytes()


def test_serve_in_mix_equals_alone(cases):
    """Jobs arrive one by one while earlier ones run; device-resident inputs
    (float32 orig promoted on the lane) and host inputs mixed."""
    from extrem_z_amd.device import DeviceBuffer
    with ez().Pool(0, lanes=3) as pool:
        tickets, bufs = {}, {}
        for rep in range(2):
            for i, (orig, decp, xi, _, _) in enumerate(cases):
                if (i + rep) % 2:
                    o = DeviceBuffer.from_numpy(orig.astype(np.float32).ravel())
                    d = DeviceBuffer.from_numpy(decp.ravel())
                    c = DeviceBuffer.zeros(orig.size, np.uint8)
                    job = pool.device_job(o.ptr, d.ptr, orig.shape[::-1], xi, c.ptr,
                                          orig_dtype=0)
                    t = pool.submit(job, keep=(o, d, c))
                    bufs[t] = ("dev", i, d, c)
                else:
                    job, keep = pool.host_job(orig, decp.copy(), xi)
                    t = pool.submit(job, keep=keep)
                    bufs[t] = ("host", i, keep[1], keep[2])
                tickets[t] = job
                time.sleep(0.002)
        seen = 0
        while seen < len(tickets):
            r = pool.wait_any(timeout_ms=60000)
            assert r is not None, "no job finished within a minute"
            t, job = r
            assert job.status == 0, ez().job_error(job)
            kind, i, d, c = bufs.pop(t)
            fixed, counts = cases[i][3], cases[i][4]
            if kind == "dev":
                d, c = d.numpy(), c.numpy()
            assert d.tobytes() == fixed.ravel().tobytes(), (kind, i)
            assert np.array_equal(c.ravel(), counts.ravel()), (kind, i)
            seen += 1
        assert pool.wait_any(timeout_ms=0) is None


def test_budget_bounds_the_footprint(cases):
    """A budget that holds one large job at a time: jobs queue, the footprint
    never exceeds the budget, and the results are unchanged."""
    big = max(c[0].size for c in cases)
    # working set of a host job: 232 B/voxel fixed + 92 B data-dependent + 17 B staging
    budget = 341 * big + (1 << 20) + (4 << 20)
    with ez().Pool(0, lanes=4, hbm_budget=budget) as pool:
        jobs, keep = [], []
        for orig, decp, xi, _, _ in cases:
            j, k = pool.host_job(orig, decp.copy(), xi)
            jobs.append(j)
            keep.append(k)
        rc, arr = pool.batch(jobs)
        assert rc == 0
        for i, (_, _, _, fixed, counts) in enumerate(cases):
            assert keep[i][1].tobytes() == fixed.tobytes(), i
        info = pool.info()
        assert info["peak_footprint_bytes"] <= budget, info
        # a job larger than the whole budget is refused, the pool stays usable
        orig, decp, xi = make_field((96, 96, 96), 41, "noise", 1e-2)
        j, k = pool.host_job(orig, decp.copy(), xi)
        t = pool.submit(j, keep=k)
        assert pool.wait(t).status == -12
        j, k = pool.host_job(cases[0][0], cases[0][1].copy(), cases[0][2])
        t = pool.submit(j, keep=k)
        assert pool.wait(t).status == 0 and k[1].tobytes() == cases[0][3].tobytes()


def test_invalid_job_rejected_and_concurrent_waiters(cases):
    orig, decp, xi, fixed, _ = cases[1]
    with ez().Pool(0, lanes=2) as pool:
        bad, keep = pool.host_job(orig, decp.copy(), xi)
        bad.xi = 0.0
        with pytest.raises(ez().ExtremzError) as e:
            pool.submit(bad, keep=keep)
        assert e.value.code == -22
        # two host threads waiting on their own tickets
        results = {}

        def client(k):
            j, kp = pool.host_job(orig, decp.copy(), xi)
            t = pool.submit(j, keep=kp)
            results[k] = (pool.wait(t).status, kp[1].tobytes())

        ths = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(120)
        assert all(results[k] == (0, fixed.tobytes()) for k in range(4))


def test_failing_job_does_not_disturb_the_batch(cases):
    """A job whose input is not finite fails with -22 inside its lane; the
    other jobs of the batch, and later jobs on the same lanes, are unaffected."""
    bad_o, bad_d, bad_xi = make_field((20, 20, 20), 42, "noise", 1e-2)
    bad_d = bad_d.copy()
    bad_d[3, 4, 5] = np.nan
    with ez().Pool(0, lanes=2) as pool:
        jobs, keep = [], []
        for k, (orig, decp, xi, _, _) in enumerate(cases[:4]):
            if k == 2:
                j, kp = pool.host_job(bad_o, bad_d.copy(), bad_xi)
            else:
                j, kp = pool.host_job(orig, decp.copy(), xi)
            jobs.append(j)
            keep.append(kp)
        rc, arr = pool.batch(jobs)
        assert rc == -22
        assert [arr[k].status for k in range(4)] == [0, 0, -22, 0]
        for k in (0, 1, 3):
            assert keep[k][1].tobytes() == cases[k][3].tobytes()
        j, kp = pool.host_job(cases[4][0], cases[4][1].copy(), cases[4][2])
        t = pool.submit(j, keep=kp)
        assert pool.wait(t).status == 0 and kp[1].tobytes() == cases[4][3].tobytes()


def test_two_waiters_on_one_ticket(cases):
    """Two threads wait for the same ticket: one receives the status, the
    other a clean error (the job record is not read after it is freed)."""
    orig, decp, xi, fixed, _ = cases[6]
    with ez().Pool(0, lanes=1) as pool:
        for rep in range(3):
            j, kp = pool.host_job(orig, decp.copy(), xi)
            t = pool.submit(j, keep=kp)
            got = []

            def waiter():
                rc = ez().lib().extremz_serve_wait(pool.handle, t, -1)
                got.append(rc)

            ths = [threading.Thread(target=waiter) for _ in range(2)]
            for th in ths:
                th.start()
            for th in ths:
                th.join(120)
            assert sorted(got) == [-22, 0], got
            assert kp[1].tobytes() == fixed.tobytes()
            pool._held.pop(t, None)
        # a waiter on one ticket beside a wait_any reaper
        j, kp = pool.host_job(orig, decp.copy(), xi)
        t = pool.submit(j, keep=kp)
        res = {}

        def one():
            res["one"] = ez().lib().extremz_serve_wait(pool.handle, t, -1)

        th = threading.Thread(target=one)
        th.start()
        tk = ctypes.c_int64(0)
        res["any"] = ez().lib().extremz_serve_wait_any(pool.handle, 20000, ctypes.byref(tk))
        th.join(120)
        # exactly one receives the status; the other finds the ticket gone
        # (-22) or, for wait_any, no unclaimed finished job (-11 at its time-out)
        assert sorted([res["one"], res["any"]]) in ([-22, 0], [-11, 0]), res
        if res["any"] == 0:
            assert tk.value == t
        pool._held.pop(t, None)
