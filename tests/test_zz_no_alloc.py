"""extremz_encode_noalloc (the encoding with no device allocation) against
extremz_encode on the same inputs, through the C ABI: bitmap, edit count,
edits, mantissas and exponents equal, on consecutive calls that share the
static scratch array (the second with a smaller input)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ez():
    import extrem_z_amd as m
    return m


def synthetic(n, seed):
    """Random before/after/counts with many snapped (count 6) changed voxels."""
    rng = np.random.default_rng(seed)
    orig = rng.random(n)
    before = orig + rng.uniform(-0.01, 0.01, n)
    after = before.copy()
    changed = rng.random(n) < 0.5
    after[changed] -= rng.uniform(1e-4, 1e-2, changed.sum())
    counts = np.where(changed, rng.integers(1, 7, n), 0).astype(np.uint8)
    return orig, before, after, counts, 0.01


def corrected(shape, seed):
    from tests.conftest import make_field
    orig, decp, xi = make_field(shape, seed, "noise", 5e-2)
    fixed, counts, st = ez().correct(orig, decp, xi)
    assert st["status"] == 0
    return orig.ravel(), decp.ravel(), fixed.ravel(), counts.ravel(), xi


def run_noalloc(ctx, orig, before, after, counts, xi):
    from extrem_z_amd.device import DeviceBuffer, synchronize
    n = orig.size
    L = ez().lib()
    vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
    L.extremz_encode_noalloc.argtypes = [vp, vp, vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp]
    d = [DeviceBuffer.from_numpy(a) for a in (orig, before, after, counts)]
    nb = (3 * n + 7) // 8
    # outputs start as garbage: every entry read back must be written
    bm = DeviceBuffer.from_numpy(np.full(nb, 0xA5, np.uint8))
    ed = DeviceBuffer.from_numpy(np.full(n, np.nan))
    mt = DeviceBuffer.from_numpy(np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64))
    ex = DeviceBuffer.from_numpy(np.full(n, -7777, np.int32))
    ne = DeviceBuffer.from_numpy(np.array([-1], np.int64))
    rc = L.extremz_encode_noalloc(ctx.handle, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, n, xi,
                                  bm.ptr, ed.ptr, mt.ptr, ex.ptr, ne.ptr)
    assert rc == 0, L.extremz_last_error()
    synchronize()
    k = int(ne.numpy()[0])
    return bm.numpy(), ed.numpy()[:k], mt.numpy()[:k], ex.numpy()[:k]


def check(ctx, case):
    want = ez().encode(*case)
    got = run_noalloc(ctx, *case)
    assert got[0].tobytes() == want[0].tobytes()
    assert got[1].size == want[1].size
    assert got[1].tobytes() == want[1].tobytes()
    assert np.array_equal(got[2], want[2]) and np.array_equal(got[3], want[3])
    return want[1].size


def test_noalloc_encode_equals_encode_on_two_calls():
    with ez().Context(0) as ctx:
        big = synthetic(700_003, 1)          # 684 blocks, the last one partial
        small = corrected((40, 30, 20), 2)   # a real correction's counts
        assert check(ctx, big) > 50_000
        assert check(ctx, small) > 0
        # and back: a small call after a large one, then the large one again
        assert check(ctx, synthetic(1500, 3)) > 0
        check(ctx, big)


def test_noalloc_encode_rejects_sizes_beyond_its_scratch():
    from extrem_z_amd.device import DeviceBuffer
    L = ez().lib()
    vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
    L.extremz_encode_noalloc.argtypes = [vp, vp, vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp]
    b = DeviceBuffer(64)
    with ez().Context(0) as ctx:
        rc = L.extremz_encode_noalloc(ctx.handle, b.ptr, b.ptr, b.ptr, b.ptr, (1 << 30) + 1, 0.1,
                                      b.ptr, None, None, None, b.ptr)
        assert rc == -22
        assert b"EXTREMZ_ENCODE_NOALLOC_MAX_N" in L.extremz_last_error()
