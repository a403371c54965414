"""Pins of the CPU oracle's labelling and correction loop."""
import numpy as np
import pytest

from oracle import oracle
from tests.conftest import make_field
from tests.topo import violations


@pytest.mark.parametrize("shape,seed", [((7, 6, 5), 0), ((9, 4, 6), 1), ((12, 12, 1), 2)])
def test_labels_equal_path_walks(shape, seed):
    W, H, D = shape
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 5, size=(D, H, W)).astype(np.float64)
    ds, asc = oracle.labels(f)
    for v in range(W * H * D):
        assert ds.ravel()[v] == oracle.max_label(f, v)
        assert asc.ravel()[v] == oracle.min_label(f, v)
    # roots are fixed points: no larger (smaller) neighbour
    t, _ = oracle.classify(f)
    assert np.all(t.ravel()[np.unique(ds)] == 4)
    assert np.all(t.ravel()[np.unique(asc)] == 0)


def test_known_answer_false_max_and_min():
    # orig is a monotone ramp; decp swaps vertices 1 and 2.  Hand-derived:
    # pre-snap v0 = 0 - 0.1, v4 = 0.4 - 0.1; v1 is a false max (lowered to the
    # midpoint of orig-xi = 0.0 and 0.18), v2's false min lowers the same v1;
    # after one c_loops iteration the field is monotone again.
    orig = np.array([0.0, 0.1, 0.2, 0.3, 0.4]).reshape(1, 1, 5)
    decp = np.array([0.0, 0.18, 0.12, 0.3, 0.4]).reshape(1, 1, 5)
    fixed, counts, st = oracle.correct(orig, decp, 0.1)
    assert st["status"] == 0 and st["c_loop_iters"] == 1 and st["outer_iters"] == 0
    assert fixed.ravel().tolist() == [-0.1, 0.09, 0.12, 0.3, 0.30000000000000004]
    assert counts.ravel().tolist() == [0, 1, 0, 0, 0]


def test_known_answer_apply_cap():
    # v1 must fall below v2 (orig order) but every midpoint step leaves it
    # above; after five edits apply_delta snaps it to orig - xi with count 6.
    orig = np.array([0.0, 0.5, 0.5000001, 1.0]).reshape(1, 1, 4)
    xi = 0.25
    decp = np.array([0.0, 0.74, 0.26, 1.0]).reshape(1, 1, 4)
    fixed, counts, st = oracle.correct(orig, decp, xi)
    assert st["status"] == 0
    f = fixed.ravel()
    assert f[0] == 0.0 - 0.25 and f[3] == 1.0 - 0.25
    # midpoints toward orig - xi = 0.25: 0.74 -> 0.495 -> 0.3725 -> 0.31125
    # -> 0.280625 -> 0.2653125 (count 5, still above v2 = 0.26); the sixth
    # iteration finds count == 5 and snaps v1 to 0.25 with count 6.
    assert st["c_loop_iters"] == 6
    assert f[1] == 0.25 and counts.ravel()[1] == 6
    assert f[2] == 0.26 and counts.ravel()[2] == 0


def test_known_answer_fixed_step_rule():
    # The apply-cap field under the fixed-step rule with q = 4 (edit_type 0,
    # src/ExtreM_cuda.cu:2905, 3065-3076): every proposal is -xi/4; v1 takes
    # three steps of -0.0625 (count 1..3), still above v2, and the fourth
    # iteration finds count == q - 1 and snaps it to orig - xi with count q.
    orig = np.array([0.0, 0.5, 0.5000001, 1.0]).reshape(1, 1, 4)
    xi = 0.25
    decp = np.array([0.0, 0.74, 0.26, 1.0]).reshape(1, 1, 4)
    fixed, counts, st = oracle.correct(orig, decp, xi, fixed_step_q=4)
    assert st["status"] == 0 and st["c_loop_iters"] == 4
    f = fixed.ravel()
    assert f[1] == 0.5 - 0.25 and counts.ravel().tolist() == [0, 4, 0, 0]
    assert f[2] == 0.26 and f[0] == -0.25 and f[3] == 0.75
    # q = 8: seven steps cover only 7/8 xi, so the same field snaps after
    # eight iterations; a false max that one step clears is fixed in one
    fixed8, counts8, st8 = oracle.correct(orig, decp, xi, fixed_step_q=8)
    assert fixed8.ravel()[1] == 0.25 and counts8.ravel()[1] == 8 and st8["c_loop_iters"] == 8
    near = np.array([0.0, 0.31, 0.28, 1.0]).reshape(1, 1, 4)
    fixed1, counts1, st1 = oracle.correct(orig, near, xi, fixed_step_q=8)
    assert fixed1.ravel()[1] == 0.31 + (-xi / 8) and counts1.ravel()[1] == 1
    assert st1["c_loop_iters"] == 1 and fixed1.ravel()[2] == 0.28


@pytest.mark.parametrize("shape,seed,kind,er", [
    ((16, 16, 16), 3, "gauss", 1e-3),
    ((20, 14, 10), 4, "noise", 2e-2),
    ((12, 18, 9), 5, "gauss", 1e-2),
])
def test_oracle_terminates_with_topology_preserved(shape, seed, kind, er):
    orig, decp, xi = make_field(shape, seed, kind, er)
    fixed, counts, st = oracle.correct(orig, decp, xi)
    assert st["status"] == 0
    tol = xi + np.spacing(np.abs(orig) + xi)
    assert np.all(np.abs(fixed - orig) <= tol)
    assert all(v == 0 for v in violations(orig, fixed).values()), violations(orig, fixed)
    # unedited vertices keep decp except pre-snapped ones (edit count 0)
    assert np.all(counts <= 6)
    changed = fixed != decp
    assert np.all(changed[counts > 0] | (fixed[counts > 0] == decp[counts > 0]))


def test_oracle_deterministic_across_threads():
    orig, decp, xi = make_field((18, 16, 14), 7, "gauss", 1e-3)
    a = oracle.correct(orig, decp, xi, threads=1)
    b = oracle.correct(orig, decp, xi, threads=8)
    assert a[0].tobytes() == b[0].tobytes() and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_bitmap3_known_answer():
    # values 1..6,0,5 -> bits LSB-first: 001 010 011 100 101 110 000 101
    v = np.array([1, 2, 3, 4, 5, 6, 0, 5], np.uint8)
    bits = "".join(format(int(x), "03b")[::-1] for x in v)
    want = [int(bits[i:i + 8][::-1], 2) for i in range(0, 24, 8)]
    assert oracle.bitmap3(v).tolist() == want == [0xD1, 0x58, 0xA3]
    with pytest.raises(ValueError):
        oracle.bitmap3(np.array([7], np.uint8))


@pytest.mark.parametrize("shape,seed,kind,er", [
    ((16, 16, 16), 3, "gauss", 1e-3), ((20, 14, 10), 4, "noise", 2e-2),
    ((48, 40, 1), 10, "gauss", 1e-3), ((7, 1, 9), 9, "noise", 1e-1)])
def test_vectorised_violations_equal_the_loops(shape, seed, kind, er):
    """The vectorised termination counters (used at 512^3) equal the loop
    version on corrected, uncorrected and half-corrected fields."""
    from tests.topo import violations_fast
    orig, decp, xi = make_field(shape, seed, kind, er)
    fixed, _, _ = oracle.correct(orig, decp, xi)
    for g in (decp, fixed, 0.5 * (decp + orig)):
        assert violations(orig, g) == violations_fast(orig, g)


# The float32 fork's rule set (src/ExtreM_cuda_f.cu): fix proposals
# -bound / (1 << 2), applyDeltaBuffer's |delta| > 1e-6, count < 6 and strict
# |orig - (decp + delta)| < bound, else decp := orig - bound with count 7,
# all in float.  Values below are exact in float32.
F32_ORIG = np.array([0.0, 0.5, 0.515625, 1.0]).reshape(1, 1, 4)


def test_known_answer_float32_cap():
    # Pre-snap: v0 -> -0.25, v3 -> 0.75.  v1 (47/64) is a false maximum above
    # v2 (17/64), which is a false minimum whose smallest original neighbour
    # is v1: both propose -0.0625 to v1.  Six steps (count 1..6):
    # 0.671875, 0.609375, 0.546875, 0.484375, 0.421875, 0.359375, each
    # within the bound and still above v2; the seventh iteration finds
    # count 6 and snaps v1 to 0.5 - 0.25 with count 7.
    decp = np.array([0.0, 0.734375, 0.265625, 1.0]).reshape(1, 1, 4)
    fixed, counts, st = oracle.correct(F32_ORIG, decp, 0.25, f32=True)
    assert st["status"] == 0 and st["c_loop_iters"] == 7 and st["outer_iters"] == 0
    assert fixed.ravel().tolist() == [-0.25, 0.25, 0.265625, 0.75]
    assert counts.ravel().tolist() == [0, 7, 0, 0]
    # the fixed-step rule with q = 4 caps at three steps (count 4)
    _, c4, s4 = oracle.correct(F32_ORIG, decp, 0.25, fixed_step_q=4)
    assert c4.ravel().tolist() == [0, 4, 0, 0] and s4["c_loop_iters"] == 4


def test_known_answer_float32_strict_bound():
    # v1 = 0.3125 steps to 0.25 = orig - xi: |0.5 - 0.25| = 0.25 is not below
    # the bound, so the apply snaps (count 7) instead of stepping (count 1).
    decp = np.array([0.0, 0.3125, 0.265625, 1.0]).reshape(1, 1, 4)
    fixed, counts, st = oracle.correct(F32_ORIG, decp, 0.25, f32=True)
    assert fixed.ravel().tolist() == [-0.25, 0.25, 0.265625, 0.75]
    assert counts.ravel().tolist() == [0, 7, 0, 0] and st["c_loop_iters"] == 1


def test_known_answer_float32_rounding():
    # Pre-snap of the global maximum 1 + 2^-23 with bound 3 * 2^-25: in float
    # (1 + 2^-23) - 3 * 2^-25 = 1 + 2^-25 rounds to 1.0; in double it is exact.
    orig = np.array([0.0, 1.0 + 2.0 ** -23]).reshape(1, 1, 2)
    xi = 3 * 2.0 ** -25
    fixed, _, st = oracle.correct(orig, orig, xi, f32=True)
    assert st["status"] == 0
    assert fixed.ravel().tolist() == [-xi, 1.0]
    f64, _, _ = oracle.correct(orig, orig, xi)
    assert f64.ravel().tolist() == [-xi, 1.0 + 2.0 ** -25]


def test_float32_rules_refuse_inexact_inputs():
    with pytest.raises(ValueError):
        oracle.correct(F32_ORIG, F32_ORIG + 0.1, 0.25, f32=True)
