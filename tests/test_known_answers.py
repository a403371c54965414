"""Hand-derived known answers for the outer-loop rules of the correction.

Every case lives on one 4 x 3 x 1 plane (vertex v = x + 4 y), with values in
units of 1/64 and xi = 6/64, so every midpoint ((orig -/+ xi) + decp) / 2 -
decp is exact.  The neighbours of a vertex, in the direction order of the
stencil (+x, -x, +y, -y, (-1,+1), (+1,-1); src/ExtreM_cuda.cu:155-162 with
z = 0), are the table NB below (checked against the stencil in
test_neighbour_table).  Each expected value was derived by hand by walking
the reference's code on the case:
  get_wrong_split/join_neighbors  src/ExtreM.cpp:2405-2450
  fixpath / r_loops               src/ExtreM.cpp:2452-2593
  get_wrong_index_max/min         src/ExtreM.cpp:2595-2670
  get_wrong_index_saddles(_join)  src/ExtreM.cpp:2672-2715
  fix_wrong_index_*, s_loops*, saddle_loops*  src/ExtreM.cpp:2717-2916
  apply_delta                     src/ExtreM.cpp:1486-1503
  the call sequence               main.cpp:65-446
and the derivation is summarised beside each case.  Each case also names
the plausible mistakes (planted in the oracle only, oracle.MUTATIONS) that it
tells apart from the rule: with the mistake the result differs.

Full runs (cases A, B, C, E) are checked on the oracle (its edit trace) and,
in the gpu-marked tests, on the HIP engine (the corrected field, edit counts
and loop counters).  Single-rule states (S1-S8) go through the oracle's rule
probe (one rule on a given state, nothing applied).
"""
import numpy as np
import pytest

from oracle import oracle

U = 1.0 / 64
XI = 6 * U
SHAPE = (1, 3, 4)  # (D, H, W)

# neighbours of each vertex of the 4 x 3 plane in direction order
NB = {0: [1, 4], 1: [2, 0, 5, 4], 2: [3, 1, 6, 5], 3: [2, 7, 6], 4: [5, 8, 0, 1],
      5: [6, 4, 9, 1, 8, 2], 6: [7, 5, 10, 2, 9, 3], 7: [6, 11, 3, 10], 8: [9, 4, 5],
      9: [10, 8, 5, 6], 10: [11, 9, 6, 7], 11: [10, 7]}


def field(vals):
    return (np.array(vals, np.float64) * U).reshape(SHAPE)


def test_neighbour_table():
    W, H = 4, 3
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1), (-1, 1), (1, -1)]
    for v, want in NB.items():
        x, y = v % W, v // W
        got = [(x + dx) + W * (y + dy) for dx, dy in dirs
               if 0 <= x + dx < W and 0 <= y + dy < H]
        assert got == want, v


# ---------------------------------------------------------------- full runs
# Each: orig, decp (x 1/64), the expected edit trace (phase, vertex, before,
# after; x 1/64), the loop counters, and the mistakes the case detects.
CASES = {
    # A: a wrong split neighbour fixed by fixpath direction 0.
    # orig types: maxima 2 (global), 9, 11; minima 4, 10 (global); 6 is a
    # 2-D saddle (type 3: two upper, two lower link components), 7 a
    # boundary 2-saddle (type 2).  Pre-snap (main.cpp:74-75, 122-131): 10 and
    # 2 to orig - xi, then saddles2 in descending value (6, 7), saddles1 (6).
    # decp then has the original types and saddle stars: c_loops does nothing.
    # Original labels of saddle 7's upper neighbours (6, 11, 3): 2, 11, 2.
    # In decp the steepest neighbour of 6 is 9 (40 > 38), so saddle 7's first
    # upper neighbour 6 reaches 9 != 2: it is flagged (nfc = 1) and
    # get_wrong_split_neighbors returns.  Saddle 6's join side matches
    # (min labels 10, 4, 10, 10), so main.cpp:255's block runs s_loops_join
    # and saddle_loops_join, which find nothing.  Outer iteration 1:
    # fixpath(6, 0) walks from 6: original steepest 2, decp steepest 9:
    # true 2, false 9; the false index is lowered to the midpoint
    # ((40 - 6) + 40) / 2 = 37 (count 1).  Three rounds then find all
    # counters zero.  Lowering the true index instead (2, the global
    # maximum, already at orig - xi) never clears the mismatch.
    "A": dict(
        orig=[16, 36, 44, 24, 4, 12, 28, 8, 32, 40, 0, 20],
        decp=[13, 36, 50, 21, 4, 6, 31, 2, 26, 40, 0, 17],
        trace=[("presnap", 10, 0, -6), ("presnap", 2, 50, 38), ("presnap", 6, 31, 22),
               ("presnap", 7, 2, 2), ("presnap", 6, 22, 22), ("r_loops", 9, 40, 37)],
        fixed=[13, 36, 38, 21, 4, 6, 22, 2, 26, 37, -6, 17],
        counts={9: 1},
        stats=dict(c_loop_iters=0, outer_iters=1, r_loop_calls=1, s_loop_calls=5,
                   saddle_loop_calls=5, c_loops_calls=4, edited=1, n_max=3, n_min=2,
                   n_saddle1=1, n_saddle2=2),
        mistakes=("fixpath_true",)),
    # B: a wrong join neighbour fixed by fixpath direction 1.
    # orig: maxima 0, 6 (global); minima 1 (global), 3, 9, 11; 1-saddles
    # 2, 7, 8, 10; 2-saddle 4.  Pre-snap: 1, 6, then 4, then 2, 8, 7, 10
    # (saddles1 in ascending value).  Original min labels of saddle 10's
    # lower neighbours (11, 9, 7): 11, 9, 11.  In decp, 7 (26) descends to
    # 3 (5 < 10): saddle 10's third lower neighbour reaches 3 != 11, flagged
    # (nfc1 = 1).  The split side matches, so main.cpp:230's block runs
    # s_loops / saddle_loops (nothing to fix).  Outer iteration 1:
    # fixpath(7, 1): original steepest descent 11, decp 3: true 11, false 3;
    # direction 1 lowers the true index: ((4 - 6) + 10) / 2 = 4 (count 1),
    # below 3's 5.  Lowering the false index 3 instead keeps 3 below 11.
    "B": dict(
        orig=[28, 0, 20, 8, 16, 36, 44, 32, 24, 12, 40, 4],
        decp=[22, 3, 26, 5, 19, 30, 47, 26, 24, 6, 43, 10],
        trace=[("presnap", 1, 3, -6), ("presnap", 6, 47, 38), ("presnap", 4, 19, 10),
               ("presnap", 2, 26, 14), ("presnap", 8, 24, 18), ("presnap", 7, 26, 26),
               ("presnap", 10, 43, 34), ("r_loops", 11, 10, 4)],
        fixed=[22, -6, 14, 5, 10, 30, 38, 26, 18, 6, 34, 4],
        counts={11: 1},
        stats=dict(c_loop_iters=0, outer_iters=1, r_loop_calls=1, s_loop_calls=5,
                   saddle_loop_calls=5, c_loops_calls=4, edited=1, n_max=2, n_min=4,
                   n_saddle1=4, n_saddle2=1),
        mistakes=("fixpath_false",)),
    # C: the positional comparison of the side labels (label_count).
    # orig: maxima 3, 5 (global), 11; minima 2 (global), 7, 8; saddles 6, 10
    # (type 3), 4 (type 1).  Split: saddle 10's upper neighbours (11, 6)
    # have original labels 11, 5; in decp 6 ascends to 3 (40 > 38): flagged.
    # Join: saddle 6's lower neighbours (7, 10, 2, 9) have original labels
    # 7, 8, 2, 8; in decp 10 descends to 7 (10 < 14), so the second label is
    # 7 != 8: flagged, although 7 is one of the saddle's original labels (a
    # set comparison would miss it).  Both counters are non-zero: neither
    # first-pass block runs.  Outer iteration 1: fixpath(6, 0) lowers the
    # false index 3 to ((40 - 6) + 40) / 2 = 37; fixpath(10, 1) lowers the
    # true index 9 to ((8 - 6) + 14) / 2 = 8.
    "C": dict(
        orig=[20, 12, 0, 40, 24, 44, 36, 16, 4, 8, 28, 32],
        decp=[14, 12, -6, 40, 27, 44, 39, 10, 4, 14, 34, 29],
        trace=[("presnap", 2, -6, -6), ("presnap", 5, 44, 38), ("presnap", 6, 39, 30),
               ("presnap", 10, 34, 22), ("presnap", 4, 27, 18), ("presnap", 10, 22, 22),
               ("presnap", 6, 30, 30), ("r_loops", 3, 40, 37), ("r_loops", 9, 14, 8)],
        fixed=[14, 12, -6, 37, 18, 38, 30, 10, 4, 8, 22, 29],
        counts={3: 1, 9: 1},
        stats=dict(c_loop_iters=0, outer_iters=1, r_loop_calls=1, s_loop_calls=4,
                   saddle_loop_calls=4, c_loops_calls=4, edited=2, n_max=3, n_min=3,
                   n_saddle1=3, n_saddle2=2),
        mistakes=("wn_any_label",)),
    # E: get_wrong_index_max and the two fix directions of s_loops.
    # orig: minima 0, 2 (global); maxima 1, 7 (global); 1-saddle 4; 2-saddle
    # 9.  Pre-snap lowers 7 (the global maximum) to 38, below maximum 1 (40).
    # Saddle 9 reaches maxima {1, 7}: original first 1, last 7; decp first 7,
    # last 1.  No wrong neighbour, so main.cpp:230's block runs s_loops:
    # pairs (1, 7) (direction 0: lower the true index 1) and (7, 1)
    # (direction 1: lower the false index 1): 1 goes to
    # ((40 - 6) + 40) / 2 = 37, below 7.  wrong_max_counter is 1, so
    # saddle_loops is skipped.  Swapping the directions lowers 7 (already
    # at orig - xi: the snap leaves it there) and never converges.
    "E": dict(
        orig=[16, 40, 0, 36, 32, 4, 20, 44, 28, 8, 12, 24],
        decp=[16, 40, 6, 33, 29, -2, 14, 47, 25, 2, 12, 18],
        trace=[("presnap", 2, 6, -6), ("presnap", 7, 47, 38), ("presnap", 9, 2, 2),
               ("presnap", 4, 29, 26), ("s_loops", 1, 40, 37)],
        fixed=[16, 37, -6, 33, 26, -2, 14, 38, 25, 2, 12, 18],
        counts={1: 1},
        stats=dict(c_loop_iters=0, outer_iters=1, r_loop_calls=1, s_loop_calls=6,
                   saddle_loop_calls=5, c_loops_calls=4, edited=1, n_max=2, n_min=2,
                   n_saddle1=1, n_saddle2=1),
        mistakes=("sloops_swap",)),
}


def expected(case):
    c = CASES[case]
    counts = np.zeros(12, np.uint8)
    for v, k in c["counts"].items():
        counts[v] = k
    return field(c["fixed"]), counts.reshape(SHAPE)


@pytest.mark.parametrize("case", sorted(CASES))
def test_full_run_trace(case):
    c = CASES[case]
    fixed, counts, st, trace = oracle.correct_trace(field(c["orig"]), field(c["decp"]), XI)
    assert st["status"] == 0
    assert [(p, v, b / U, a / U) for p, v, b, a in trace] == \
        [(p, v, float(b), float(a)) for p, v, b, a in c["trace"]]
    want_f, want_c = expected(case)
    assert fixed.tobytes() == want_f.tobytes()
    assert np.array_equal(counts, want_c)
    assert {k: st[k] for k in c["stats"]} == c["stats"]


@pytest.mark.parametrize("case", sorted(CASES))
def test_full_run_detects_its_mistakes(case):
    c = CASES[case]
    ref = oracle.correct_trace(field(c["orig"]), field(c["decp"]), XI)
    for m in c["mistakes"]:
        got = oracle.correct_trace(field(c["orig"]), field(c["decp"]), XI, m)
        assert got[2]["status"] != 0 or got[0].tobytes() != ref[0].tobytes(), m


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_full_run_on_gpu(case):
    import extrem_z_amd as ez
    c = CASES[case]
    fixed, counts, st = ez.correct(field(c["orig"]), field(c["decp"]), XI)
    assert st["status"] == 0
    want_f, want_c = expected(case)
    assert fixed.tobytes() == want_f.tobytes()
    assert np.array_equal(counts, want_c)
    assert {k: st[k] for k in c["stats"]} == c["stats"]


# ------------------------------------------------------- single-rule states
# Each: orig, decp (x 1/64), the rule, the expected probe output (proposal
# deltas x 1/64), and per planted mistake the output it gives instead.
STATES = {
    # S1: saddles in ascending id order; each flags its first mismatch.
    # 2-saddles 2, 5 (type 3), 10 (type 2).  Original labels: saddle 2's
    # upper neighbours (3, 5) -> 3, 4; saddle 5's (4, 9) -> 4, 9; saddle
    # 10's (11, 9, 6) -> 11, 9, 3.  In decp 9 (39) outranks 4 (38): 5
    # ascends to 9 (saddle 2's second label, flagged first) and 6 ascends to
    # 9 (saddle 10's third label, flagged second).  In descending saddle
    # order the list is [6, 5].
    "S1": dict(
        orig=[4, 8, 28, 40, 44, 32, 16, 0, 20, 36, 12, 24],
        decp=[1, 5, 25, 37, 38, 29, 16, 0, 23, 39, 9, 21],
        rule="wrong_neighbors", out={"split": [5, 6], "join": []},
        mistakes={"wn_desc_order": {"split": [6, 5], "join": []}}),
    # S2: a saddle flags only its first mismatch and returns.  1-saddle 7's
    # lower neighbours (11, 3, 10) have original labels 2, 2, 2; in decp
    # 11 and 10 descend (through 10, 9, 5) to 4 (-2 < 3): 11 is flagged and
    # the saddle returns; without the return 10 is flagged too.
    "S2": dict(
        orig=[16, 24, 0, 28, 4, 8, 44, 40, 32, 12, 20, 36],
        decp=[10, 24, 3, 22, -2, 5, 38, 37, 29, 15, 20, 36],
        rule="wrong_neighbors", out={"split": [], "join": [11]},
        mistakes={"wn_all": {"split": [], "join": [11, 10]}}),
    # S3: first saddle wins per maximum.  2-saddles 1, 2, 7.  Triplets
    # (original -> decp, ascending value): saddle 1 {9, 0} -> {9, 0}; saddle
    # 2 {9, 0, 3} -> {3, 9, 0}: first 9 -> 3, last 3 -> 0: pairs (9, 3) and
    # (3, 0); saddle 7 {9, 3} -> {3, 9}: both ends differ but 9 and 3 are
    # flagged already.  Direction 0 lowers the true index 9 to
    # ((36 - 6) + 42) / 2 = 36, direction 1 the false index 0 to
    # ((40 - 6) + 46) / 2 = 40.  If the last saddle won, (3, 9) would replace
    # (3, 0) and 0 would keep its value.
    "S3": dict(
        orig=[40, 24, 12, 44, 4, 32, 8, 20, 16, 36, 28, 0],
        decp=[46, 30, 15, 38, 1, 35, 2, 14, 19, 42, 28, 6],
        rule="wrong_index_max",
        out={"pairs_0": [(9, 3)], "pairs_1": [(3, 0)], "proposals": {0: -6, 9: -6}},
        mistakes={"wi_last_wins": {"pairs_0": [(9, 3)], "pairs_1": [(3, 9)],
                                   "proposals": {9: -6}}}),
    # S4: first saddle wins per minimum.  1-saddles 3, 4.  Triplets
    # (descending value): saddle 3 {1, 7} -> {7, 1}: first 1 -> 7 (pair for
    # direction 1), last 7 -> 1 (direction 0); saddle 4 {1, 8} -> {8, 1}:
    # first 1 -> 8 (1 already flagged), last 8 -> 1.  Direction 0 lowers the
    # true indices 7 and 8: ((4 - 6) + 10) / 2 = 4 and ((0 - 6) + 6) / 2 = 0;
    # direction 1 the false index 7 (the same value).
    "S4": dict(
        orig=[36, 8, 12, 20, 16, 32, 40, 4, 0, 44, 28, 24],
        decp=[36, 2, 6, 23, 13, 35, 40, 10, 6, 50, 31, 27],
        rule="wrong_index_min",
        out={"pairs_0": [(7, 1), (8, 1)], "pairs_1": [(1, 7)],
             "proposals": {7: -6, 8: -6}},
        mistakes={"wi_last_wins": {"pairs_0": [(7, 1), (8, 1)], "pairs_1": [(1, 8)],
                                   "proposals": {7: -6, 8: -6}}}),
    # S5: the two directions of s_loops_join.  Saddle 6 reaches minima
    # {2, 9, 7} (descending original value) and {9, 7, 2} in decp: first
    # 2 -> 9, last 7 -> 2.  Direction 0 lowers the true index 7 to
    # ((0 - 6) + 6) / 2 = 0; direction 1 the false index 9 to
    # ((4 - 6) + 10) / 2 = 4.  Swapped, both would move 2 (by 0).
    "S5": dict(
        orig=[32, 44, 8, 40, 20, 28, 16, 0, 12, 4, 36, 24],
        decp=[38, 41, 2, 40, 23, 34, 22, 6, 18, 10, 42, 24],
        rule="wrong_index_min",
        out={"pairs_0": [(7, 2)], "pairs_1": [(2, 9)], "proposals": {7: -6, 9: -6}},
        mistakes={"sloops_swap": {"pairs_0": [(7, 2)], "pairs_1": [(2, 9)],
                                  "proposals": {2: 0}}}),
    # S6: fix_wrong_index_saddle direction 0 returns when the original
    # saddle is higher.  The original's largest 2-saddle of maximum 0 is 5;
    # decp has a second saddle 1 (24), which is decp's largest for 0 (5
    # reaches 3 and 8 only).  Pair (5, 1); decp[5] = 26 > decp[1] = 24: no
    # proposal.  Without the guard, 1 would go to ((24 - 6) + 24) / 2 = 21.
    "S6": dict(
        orig=[28, 24, 0, 40, 16, 20, 36, 32, 44, 4, 8, 12],
        decp=[34, 24, -3, 43, 22, 26, 30, 26, 41, 1, 5, 15],
        rule="saddles", out={"pairs": [(5, 1)], "proposals": {}},
        mistakes={"saddle_no_guard": {"pairs": [(5, 1)], "proposals": {1: -3}}}),
    # S7: direction 1 raises the true saddle.  The original's smallest
    # 1-saddle of minimum 0 is 5; in decp 5 is regular and saddle 1 (30)
    # reaches 0: pair (5, 1), decp[5] = 28 < 30, so 5 is raised to
    # ((28 + 6) + 28) / 2 = 31.
    "S7": dict(
        orig=[8, 24, 32, 40, 44, 28, 12, 36, 16, 0, 4, 20],
        decp=[11, 30, 38, 46, 44, 28, 18, 30, 19, 3, 4, 14],
        rule="saddles_join", out={"pairs": [(5, 1)], "proposals": {5: 3}},
        mistakes={"saddle_join_lower": {"pairs": [(5, 1)], "proposals": {5: -3}}}),
    # S8: direction 1's guard.  1-saddles 1, 6; the original's smallest for
    # minimum 2 is 6 (20 < 24); in decp 1 (21) is below 6 (23) and reaches 2:
    # pair (6, 1) with the true saddle higher, so no proposal.  This is the
    # only way a join pair arises once every side label matches, which is
    # why a full run never raises a saddle (test below).
    "S8": dict(
        orig=[4, 24, 8, 32, 36, 12, 20, 28, 40, 44, 0, 16],
        decp=[7, 21, 14, 32, 33, 15, 23, 25, 34, 38, -6, 22],
        rule="saddles_join", out={"pairs": [(6, 1)], "proposals": {}},
        mistakes={"saddle_no_guard": {"pairs": [(6, 1)], "proposals": {6: 1.5}}}),
}


def probe(state, mutation="none"):
    s = STATES[state]
    out = oracle.rule_probe(field(s["orig"]), field(s["decp"]), XI, s["rule"], mutation)
    if "proposals" in out:
        out["proposals"] = {v: d / U for v, d in out["proposals"].items()}
    return out


@pytest.mark.parametrize("state", sorted(STATES))
def test_single_rule(state):
    s = STATES[state]
    assert probe(state) == s["out"]
    for m, want in s["mistakes"].items():
        assert want != s["out"]
        assert probe(state, m) == want, m


def test_fixpath_proposals_of_case_c():
    """fixpath alone on case C's state after the pre-snap: the split
    neighbour 6 lowers 3 by 3/64, the join neighbour 10 lowers 9 by 6/64."""
    pre = list(CASES["C"]["decp"])
    for v, val in ((2, -6), (5, 38), (6, 30), (10, 22), (4, 18)):
        pre[v] = val
    out = oracle.rule_probe(field(CASES["C"]["orig"]), field(pre), XI, "fixpath")
    assert out["split"] == [6] and out["join"] == [10]
    assert {v: d / U for v, d in out["proposals"].items()} == {3: -3, 9: -6}


def test_saddle_fix_guard_decides_nothing_in_full_runs():
    """Once every side label matches (the precondition of saddle_loops in
    main.cpp:230-274, 313-396), an original largest (smallest) saddle still
    reaches its extremum in decp, so decp's largest (smallest) one comes
    after (before) it in (value, id) order: direction 0 always lowers the
    false saddle and direction 1 always returns.  On random small fields the
    guard-free and the lowering variants give the same edit traces, and
    saddle_loops_join never edits."""
    rng = np.random.default_rng(11)
    for k in range(150):
        W, H, D = [(4, 3, 1), (5, 4, 1), (3, 3, 2)][k % 3]
        n = W * H * D
        o = rng.permutation(n).reshape(D, H, W) * (4 * U)
        d = o + rng.choice([-1.0, -0.5, 0.0, 0.5, 1.0], size=o.shape) * XI
        ref = oracle.correct_trace(o, d, XI)
        if ref[2]["status"]:
            continue
        assert all(p != "saddle_loops_join" for p, _, _, _ in ref[3])
        for m in ("saddle_no_guard", "saddle_join_lower"):
            got = oracle.correct_trace(o, d, XI, m)
            assert got[3] == ref[3], (k, m)


@pytest.mark.gpu
@pytest.mark.parametrize("state", sorted(STATES))
def test_rule_states_corrected_on_gpu(state):
    """The single-rule states as whole corrections: the HIP engine against the
    oracle, bit for bit."""
    import extrem_z_amd as ez
    s = STATES[state]
    f0, c0, s0 = oracle.correct(field(s["orig"]), field(s["decp"]), XI)
    f1, c1, s1 = ez.correct(field(s["orig"]), field(s["decp"]), XI)
    assert s0["status"] == s1["status"] == 0
    assert f0.tobytes() == f1.tobytes() and np.array_equal(c0, c1)
    assert {k: s0[k] for k in oracle.STAT_KEYS} == {k: s1[k] for k in oracle.STAT_KEYS}


# ------------------------------------------------- persistence filter (thr)
# The GPU fork's findAscPaths keeps a maximum in a 2-saddle's triplet only
# when |f(max) - f(saddle)| >= thr (src/ExtreM_cuda.cu:1877-1881), each field
# with its own values.  Case E after the pre-snap: saddle 9 (orig 8, decp 2)
# reaches maxima 1 (orig 40, decp 40) and 7 (orig 44, decp 38).
E_PRESNAP = [16, 40, -6, 33, 26, -2, 14, 38, 25, 2, 12, 18]


def test_persistence_filter_single_rule():
    # thr = 34/64: the original triplet drops 1 (40 - 8 = 32 < 34) and keeps 7
    # (36): first = last = 7.  decp keeps both (38, 36): ascending 7, 1.  The
    # first ends agree; the last ends give the pair (7, 1) only, whose
    # direction-1 fix lowers the false index 1 by 3/64.
    o, d = field(CASES["E"]["orig"]), field(E_PRESNAP)
    plain = oracle.rule_probe(o, d, XI, "wrong_index_max")
    assert plain["pairs_0"] == [(1, 7)] and plain["pairs_1"] == [(7, 1)]
    with oracle.persistence(34 * U):
        out = oracle.rule_probe(o, d, XI, "wrong_index_max")
    assert out["pairs_0"] == [] and out["pairs_1"] == [(7, 1)]
    assert {v: x / U for v, x in out["proposals"].items()} == {1: -3}


def test_persistence_filter_full_run():
    # thr = 40/64 empties both triplets (36 and 32 / 38 and 36 are below it):
    # no wrong index, nothing to fix, the pre-snapped field stands.
    o, d = field(CASES["E"]["orig"]), field(CASES["E"]["decp"])
    with oracle.persistence(40 * U):
        fixed, counts, st = oracle.correct(o, d, XI)
    assert st["status"] == 0 and st["outer_iters"] == 0
    assert fixed.tobytes() == field(E_PRESNAP).tobytes() and not counts.any()
    # thr = 34/64: the first round lowers 1 to 37 (the single-rule state
    # above); then decp keeps {1 (37 - 2 = 35), 7 (36)}, first 1, while the
    # original keeps {7}: the pair (7, 1) lowers 7, which sits at orig - xi
    # (the midpoint step is 0, so it snaps there again), and the mismatch
    # stays: the loop ends at its cap with 1 at 37.
    with oracle.persistence(34 * U):
        fixed, counts, st = oracle.correct(o, d, XI, max_outer=20)
    assert st["status"] == -3
    assert fixed.ravel()[1] == 37 * U and fixed.ravel()[7] == 38 * U


@pytest.mark.gpu
def test_persistence_filter_on_gpu():
    import extrem_z_amd as ez
    o, d = field(CASES["E"]["orig"]), field(CASES["E"]["decp"])
    fixed, counts, st = ez.correct(o, d, XI, persistence=40 * U)
    assert st["status"] == 0 and st["outer_iters"] == 0
    assert fixed.tobytes() == field(E_PRESNAP).tobytes() and not counts.any()
    with pytest.raises(ez.ExtremzError) as e:
        ez.correct(o, d, XI, persistence=34 * U, max_outer=20)
    assert e.value.code == -3
