"""GPU parity: the HIP engine against the CPU oracle on the same inputs.

Integer outputs (types, masks, labels, counts) and the corrected float64
field must be bit-identical; iteration counters must match.
"""
import numpy as np
import pytest

from oracle import oracle
from tests.conftest import make_field
from tests.topo import violations

pytestmark = pytest.mark.gpu


def ez():
    import extrem_z_amd as m
    return m


CASES = [
    ((16, 16, 16), 3, "gauss", 1e-3),
    ((20, 14, 10), 4, "noise", 2e-2),
    ((12, 18, 9), 5, "gauss", 1e-2),
    ((33, 17, 5), 6, "noise", 5e-2),
    ((40, 40, 40), 8, "gauss", 1e-3),
    ((7, 1, 9), 9, "noise", 1e-1),
    ((48, 40, 1), 10, "gauss", 1e-3),  # depth 1: the reference's planar triangle mesh
    ((37, 29, 1), 11, "noise", 2e-2),
]


@pytest.mark.parametrize("shape,seed,kind,er", CASES)
def test_classify_parity(shape, seed, kind, er):
    orig, decp, xi = make_field(shape, seed, kind, er)
    for f in (orig, decp):
        t0, m0 = oracle.classify(f)
        t1, m1 = ez().classify(f)
        assert np.array_equal(t0, t1)
        assert np.array_equal(m0, m1)


@pytest.mark.parametrize("shape,seed,kind,er", CASES)
def test_labels_parity(shape, seed, kind, er):
    orig, decp, xi = make_field(shape, seed, kind, er)
    for f in (orig, decp):
        a0, b0 = oracle.labels(f)
        a1, b1 = ez().labels(f)
        assert np.array_equal(a0, a1)
        assert np.array_equal(b0, b1)


COUNTERS = ("status", "c_loop_iters", "outer_iters", "r_loop_calls", "s_loop_calls",
            "saddle_loop_calls", "c_loops_calls", "edited", "n_max", "n_min", "n_saddle1",
            "n_saddle2")


@pytest.mark.parametrize("shape,seed,kind,er", CASES)
def test_correct_parity(shape, seed, kind, er):
    orig, decp, xi = make_field(shape, seed, kind, er)
    f0, c0, s0 = oracle.correct(orig, decp, xi)
    f1, c1, s1 = ez().correct(orig.astype(np.float32), decp, xi)
    assert s0["status"] == 0
    assert f0.tobytes() == f1.tobytes()
    assert np.array_equal(c0, c1)
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}


def test_known_answer_on_gpu():
    orig = np.array([0.0, 0.1, 0.2, 0.3, 0.4]).reshape(1, 1, 5)
    decp = np.array([0.0, 0.18, 0.12, 0.3, 0.4]).reshape(1, 1, 5)
    fixed, counts, st = ez().correct(orig, decp, 0.1)
    assert fixed.ravel().tolist() == [-0.1, 0.09, 0.12, 0.3, 0.30000000000000004]
    assert counts.ravel().tolist() == [0, 1, 0, 0, 0]


def test_config1_parity_64cube():
    from extrem_z_amd import synth
    orig, decp, xi = synth.config(1)
    f0, c0, s0 = oracle.correct(orig, decp, xi)
    f1, c1, s1 = ez().correct(orig.astype(np.float32), decp, xi)
    assert f0.tobytes() == f1.tobytes()
    assert np.array_equal(c0, c1)
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}


def test_repeat_calls_are_identical():
    orig, decp, xi = make_field((24, 20, 16), 11, "gauss", 1e-3)
    a = ez().correct(orig, decp, xi)
    big_o, big_d, big_xi = make_field((30, 30, 30), 12, "noise", 1e-2)
    ez().correct(big_o, big_d, big_xi)          # a larger field in between
    b = ez().correct(orig, decp, xi)
    assert a[0].tobytes() == b[0].tobytes() and np.array_equal(a[1], b[1])


def test_encode_parity():
    orig, decp, xi = make_field((24, 20, 16), 13, "noise", 2e-2)
    fixed, counts, _ = oracle.correct(orig, decp, xi)
    bm, edits, mant, expo = ez().encode(orig, decp, fixed, counts, xi)
    vals = np.where(decp.ravel() != fixed.ravel(), counts.ravel(), 0).astype(np.uint8)
    assert np.array_equal(bm, oracle.bitmap3(vals))
    sel = (decp.ravel() != fixed.ravel()) & (counts.ravel() == 6)
    want = -(decp.ravel()[sel] - (orig.ravel()[sel] - xi))
    assert edits.tobytes() == want.tobytes()
    bits = want.view(np.uint64)
    assert np.array_equal(mant, bits & np.uint64((1 << 52) - 1))
    assert np.array_equal(expo, ((bits >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int32) - 1023)


def test_invariants_at_128cube():
    orig, decp, xi = make_field((128, 128, 128), 21, "gauss", 1e-3, n_gauss=16)
    fixed, counts, st = ez().correct(orig.astype(np.float32), decp, xi)
    assert st["status"] == 0
    tol = xi + np.spacing(np.abs(orig) + xi)
    assert np.all(np.abs(fixed - orig) <= tol)
    t0, m0 = ez().classify(orig)
    t1, m1 = ez().classify(fixed)
    assert np.array_equal(t0, t1)
    sad = (t0 >= 1) & (t0 <= 3)
    assert np.array_equal(m0[sad], m1[sad])


@pytest.mark.parametrize("where,val", [("decp", np.nan), ("orig", np.nan), ("orig", np.inf),
                                       ("orig", -np.inf)])
def test_nonfinite_input_rejected(where, val):
    o = np.zeros((4, 4, 4))
    d = o.copy()
    (d if where == "decp" else o)[1, 1, 1] = val
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(o, d, 0.1)
    assert e.value.code == -22


def test_full_size_256_digest():
    """The 256^3 benchmark workload, corrected on the GPU, against the digest
    of the oracle's result (tests/golden/make_full_digest.py): the same fixed
    field and edit counts byte for byte after 72 outer iterations."""
    import hashlib
    import json
    import os

    from extrem_z_amd import synth
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "gauss32_256_digest.json")))
    r = ref["recipe"]
    orig, decp, xi = synth.make_case(r["kind"], tuple(ref["shape"]), seed=r["seed"], er=r["er"],
                                     n_gauss=r["n_gauss"])
    assert xi == ref["xi"]
    fixed, counts, stats = ez().correct(orig, decp, xi)
    for k, v in ref["stats"].items():
        assert stats[k] == v, k
    assert hashlib.sha256(counts.tobytes()).hexdigest() == ref["sha256_counts"]
    assert hashlib.sha256(fixed.tobytes()).hexdigest() == ref["sha256_fixed"]


@pytest.mark.parametrize("q", [3, 4, 6])
@pytest.mark.parametrize("shape,seed,kind,er", CASES[:4])
def test_fixed_step_rule_parity(shape, seed, kind, er, q):
    """The fixed-step rule set (edit_type 0 of the GPU fork): GPU against the
    oracle, byte for byte, with the counters."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    f0, c0, s0 = oracle.correct(orig, decp, xi, fixed_step_q=q)
    f1, c1, s1 = ez().correct(orig, decp, xi, fixed_step_q=q)
    assert s0["status"] == 0
    assert f0.tobytes() == f1.tobytes()
    assert np.array_equal(c0, c1) and int(c1.max()) <= q
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}
    tol = xi + np.spacing(np.abs(orig) + xi)
    assert np.all(np.abs(f1 - orig) <= tol)


def test_fixed_step_rule_rejects_bad_q():
    orig, decp, xi = make_field((8, 8, 8), 1, "noise", 1e-2)
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(orig, decp, xi, fixed_step_q=255)
    assert e.value.code == -22


INC_CASES = CASES + [((64, 64, 64), 12, "gauss", 1e-3), ((48, 40, 36), 13, "noise", 1e-2)]


@pytest.mark.parametrize("shape,seed,kind,er", INC_CASES)
def test_incremental_passes_equal_full_passes(shape, seed, kind, er, monkeypatch):
    """With EXTREMZ_VERIFY_INC every CP pass of decp that runs incrementally is
    followed inside the engine by the full pass on the same decp, and any
    difference in labels, masks, pointers, triplet ends, gather keys, slots,
    global extrema or the wrong-neighbour candidate count fails the call.  The result equals the run
    with full passes only, and the oracle."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "1")
    f1, c1, s1 = ez().correct(orig, decp, xi)
    assert s1["status"] == 0
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "0")
    monkeypatch.setenv("EXTREMZ_FULL_PASSES", "1")
    f2, c2, s2 = ez().correct(orig, decp, xi)
    assert s2["inc_passes"] == 0
    assert f1.tobytes() == f2.tobytes() and np.array_equal(c1, c2)
    assert {k: s1[k] for k in COUNTERS} == {k: s2[k] for k in COUNTERS}
    if s2["label_passes"] > 2:
        assert s1["inc_passes"] > 0
    if shape[0] * shape[1] * shape[2] <= 40 ** 3:
        f0, c0, _ = oracle.correct(orig, decp, xi)
        assert f0.tobytes() == f1.tobytes()


@pytest.mark.parametrize("shape,seed,kind", [
    ((130, 9, 11), 51, "noise"),   # crosses the 64-wide x tile edges
    ((65, 5, 9), 52, "gauss"),     # one voxel past a tile in x, y and z
    ((64, 4, 4), 53, "noise"),     # exactly one tile
    ((48, 40, 1), 54, "gauss"),
    ((3, 70, 6), 55, "noise"),
    ((2, 9, 13), 56, "noise"),     # W < 3: the tile pass's general path
])
def test_label_pass_classification_equals_k_classify(shape, seed, kind):
    """The original field's types and lower masks come out of the label tile
    pass (its fast interior path and LUT lookup); extremz_classify runs
    k_classify.  Both, and the labels, against the oracle."""
    orig, decp, xi = make_field(shape, seed, kind, 1e-2)
    for f in (orig, decp):
        t, m, ds, asc = ez().label_pass(f)
        tc, mc = ez().classify(f)
        t0, m0 = oracle.classify(f)
        assert np.array_equal(t, tc) and np.array_equal(m, mc)
        assert np.array_equal(t, t0) and np.array_equal(m, m0)
        a0, b0 = oracle.labels(f)
        assert np.array_equal(ds, a0) and np.array_equal(asc, b0)


PLANE_CASES = [
    ((48, 40, 1), 71, "gauss", 1e-3),
    ((37, 29, 1), 72, "noise", 2e-2),
    ((64, 64, 1), 73, "noise", 1e-2),
    ((130, 9, 1), 74, "gauss", 1e-3),
    ((5, 3, 1), 75, "noise", 5e-2),
    ((256, 200, 1), 76, "gauss", 1e-4),
]


@pytest.mark.parametrize("shape,seed,kind,er", PLANE_CASES)
def test_plane_adjacency_parity(shape, seed, kind, er):
    """The CPU path's plane adjacency (maxNeighbors = 6 on a depth-1 field:
    +-x, +-y, classifyVertex1's triangle links and type rule) against the
    oracle in the same mode, byte for byte with the counters."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    with oracle.plane_adjacency():
        f0, c0, s0 = oracle.correct(orig, decp, xi)
    f1, c1, s1 = ez().correct(orig, decp, xi, adjacency=ez().ADJ_PLANE4)
    assert s0["status"] == 0
    assert f0.tobytes() == f1.tobytes()
    assert np.array_equal(c0, c1)
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}
    # and it is a different mode from the stencil's
    f2, _, s2 = ez().correct(orig, decp, xi)
    assert (s2["n_max"], s2["n_min"], s2["n_saddle1"]) != (s1["n_max"], s1["n_min"],
                                                         s1["n_saddle1"])


def test_plane_adjacency_needs_depth_one():
    orig, decp, xi = make_field((8, 8, 2), 77, "noise", 1e-2)
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(orig, decp, xi, adjacency=ez().ADJ_PLANE4)
    assert e.value.code == -22


def test_plane_adjacency_incremental_equals_full(monkeypatch):
    orig, decp, xi = make_field((256, 200, 1), 76, "gauss", 1e-4)
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "1")
    f1, c1, s1 = ez().correct(orig, decp, xi, adjacency=ez().ADJ_PLANE4)
    assert s1["status"] == 0
    monkeypatch.setenv("EXTREMZ_VERIFY_INC", "0")
    monkeypatch.setenv("EXTREMZ_FULL_PASSES", "1")
    f2, c2, s2 = ez().correct(orig, decp, xi, adjacency=ez().ADJ_PLANE4)
    assert f1.tobytes() == f2.tobytes() and np.array_equal(c1, c2)


def f32_case(shape, seed, kind, er):
    """orig, decp and xi that float32 holds exactly."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    return (orig.astype(np.float32).astype(np.float64), decp.astype(np.float32).astype(np.float64),
            float(np.float32(xi)))


def test_float32_rules_known_answers_on_gpu():
    o = np.array([0.0, 0.5, 0.515625, 1.0]).reshape(1, 1, 4)
    for d, want_f, want_c, iters in (([0.0, 0.734375, 0.265625, 1.0], [-0.25, 0.25, 0.265625, 0.75],
                                      [0, 7, 0, 0], 7),
                                     ([0.0, 0.3125, 0.265625, 1.0], [-0.25, 0.25, 0.265625, 0.75],
                                      [0, 7, 0, 0], 1)):
        f, c, st = ez().correct(o, np.array(d).reshape(1, 1, 4), 0.25, rule_set=ez().RULES_F32)
        assert f.ravel().tolist() == want_f and c.ravel().tolist() == want_c
        assert st["c_loop_iters"] == iters
    orig = np.array([0.0, 1.0 + 2.0 ** -23]).reshape(1, 1, 2)
    f, _, _ = ez().correct(orig, orig, 3 * 2.0 ** -25, rule_set=ez().RULES_F32)
    assert f.ravel().tolist() == [-3 * 2.0 ** -25, 1.0]


@pytest.mark.parametrize("shape,seed,kind,er", CASES[:6] + [((64, 64, 64), 81, "gauss", 1e-3)])
def test_float32_rules_parity(shape, seed, kind, er):
    orig, decp, xi = f32_case(shape, seed, kind, er)
    f0, c0, s0 = oracle.correct(orig, decp, xi, f32=True)
    f1, c1, s1 = ez().correct(orig.astype(np.float32), decp, xi, rule_set=ez().RULES_F32)
    assert s0["status"] == 0
    assert f0.tobytes() == f1.tobytes()
    assert np.array_equal(c0, c1) and int(c1.max()) <= 7
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}
    assert np.all(f1.astype(np.float32).astype(np.float64) == f1)


def test_float32_rules_reject_inexact_input():
    orig, decp, xi = make_field((8, 8, 8), 82, "noise", 1e-2)
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(orig, decp + 1e-12, xi, rule_set=ez().RULES_F32)
    assert e.value.code == -22
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(orig, decp, xi, rule_set=ez().RULES_F32, fixed_step_q=4)
    assert e.value.code == -22


@pytest.mark.parametrize("shape,seed,kind,er,sim", [
    ((16, 16, 16), 3, "gauss", 1e-3, 2e-3), ((48, 40, 1), 10, "gauss", 1e-3, 5e-3),
    ((48, 40, 1), 102, "gauss", 1e-3, 1e-2)])
def test_persistence_filter_parity(shape, seed, kind, er, sim):
    """The persistence filter (thr = (max - min) * sim, src/ExtreM_cuda.cu:231)
    against the oracle, byte for byte with the counters; on the last field the
    filter changes the result (no outer iteration instead of one)."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    thr = float(orig.max() - orig.min()) * sim
    with oracle.persistence(thr):
        f0, c0, s0 = oracle.correct(orig, decp, xi)
    f1, c1, s1 = ez().correct(orig, decp, xi, persistence=thr)
    assert s0["status"] == 0
    assert f0.tobytes() == f1.tobytes() and np.array_equal(c0, c1)
    assert {k: s0[k] for k in COUNTERS} == {k: s1[k] for k in COUNTERS}


def test_persistence_filter_cap():
    """A filter that leaves the two fields' triplets apart for good: the loop
    ends at its cap on both sides (EXTREMZ_ENOTCONV)."""
    orig, decp, xi = make_field((20, 14, 10), 4, "noise", 2e-2)
    thr = float(orig.max() - orig.min()) * 0.05
    with oracle.persistence(thr):
        _, _, s0 = oracle.correct(orig, decp, xi, max_outer=20)
    assert s0["status"] == -3
    with pytest.raises(ez().ExtremzError) as e:
        ez().correct(orig, decp, xi, persistence=thr, max_outer=20)
    assert e.value.code == -3
