"""Split-tree check (compute_MergeT, src/ExtreM.cpp:1863-1918, 2132-2249,
2930-2952; comparison main.cpp:454-504).  The oracle's restatement is pinned
by an independent superlevel-set sweep with union-find (the elder rule:
where components meet, every maximum but the highest dies at that vertex)
and by a closed-form two-bump field; the GPU path is compared with the oracle."""
import numpy as np
import pytest

from oracle import oracle
from tests.conftest import make_field
from tests.topo import neighbours


def sweep_pairs(field):
    """Persistence pair (dying vertex) of every maximum by a descending sweep."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    flat = f.ravel()
    n = flat.size
    order = sorted(range(n), key=lambda v: (flat[v], v), reverse=True)
    parent = list(range(n))
    top = [-1] * n              # highest vertex (the component's maximum)
    seen = [False] * n

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    pair = {}
    for v in order:
        roots = {find(u) for u in neighbours(f.shape, v) if seen[u]}
        seen[v] = True
        if not roots:
            top[v] = v
            continue
        rs = sorted(roots, key=lambda r: (flat[top[r]], top[r]), reverse=True)
        for r in rs[1:]:
            pair[top[r]] = v
        keep = rs[0]
        for r in rs[1:]:
            parent[r] = keep
        parent[v] = keep
    gmax = order[0]
    pair[gmax] = int(min(range(n), key=lambda v: (flat[v], v)))
    return pair


@pytest.mark.parametrize("shape,seed,kind", [((9, 8, 7), 61, "noise"), ((12, 10, 6), 62, "noise"),
                                             ((16, 14, 5), 63, "gauss"), ((7, 9, 1), 64, "noise"),
                                             ((20, 12, 3), 65, "noise")])
def test_oracle_pairs_match_union_find_sweep(shape, seed, kind):
    f, _, _ = make_field(shape, seed, kind, 1e-3, n_gauss=6)
    t = oracle.split_tree(f)
    want = sweep_pairs(f)
    got = {int(m): int(s) for m, s in zip(t["max"], t["pair_saddle"])}
    assert got == want


def test_two_bumps_closed_form():
    """Two Gaussian bumps on a line of the grid: the lower maximum pairs with
    the saddle between them (the lowest point of the ridge), and its branch
    hangs off the higher maximum's branch at that saddle."""
    W, H, D = 24, 9, 5
    x = np.arange(W)[None, None, :]
    y = np.arange(H)[None, :, None]
    z = np.arange(D)[:, None, None]
    r2a = (x - 6) ** 2 + (y - 4) ** 2 + (z - 2) ** 2
    r2b = (x - 17) ** 2 + (y - 4) ** 2 + (z - 2) ** 2
    f = (1.0 * np.exp(-r2a / 8.0) + 0.7 * np.exp(-r2b / 8.0)).astype(np.float64)
    t = oracle.split_tree(f)
    flat = f.ravel()
    hi = int(np.argmax(flat))
    lo = 17 + W * (4 + H * 2)
    assert list(t["max"]) == [lo, hi]
    ridge = [xx + W * (4 + H * 2) for xx in range(6, 18)]
    sad = min(ridge, key=lambda v: flat[v])
    assert t["pair_saddle"][0] == sad
    assert t["parent"].tolist() == [hi, hi]
    assert t["branch"][0] == [lo, sad]
    assert t["branch"][1][:1] == [hi] and sad in t["branch"][1]


def ez():
    import extrem_z_amd as m
    return m


def oracle_compare(a, b):
    """main.cpp:480-504 over the oracle's trees: ranks whose pair / branch differ."""
    ta, tb = oracle.split_tree(a), oracle.split_tree(b)
    n = min(len(ta["max"]), len(tb["max"]))
    extra = max(len(ta["max"]), len(tb["max"])) - n
    pm = extra + sum(int(ta["max"][i] != tb["max"][i] or ta["pair_saddle"][i] != tb["pair_saddle"][i])
                     for i in range(n))
    bm = extra + sum(int(ta["branch"][i] != tb["branch"][i]) for i in range(n))
    return pm, bm


@pytest.mark.gpu
@pytest.mark.parametrize("shape,seed,kind", [((9, 8, 7), 61, "noise"), ((16, 14, 5), 63, "gauss"),
                                             ((7, 9, 1), 64, "noise"), ((24, 20, 16), 66, "noise"),
                                             ((40, 40, 40), 67, "gauss")])
def test_gpu_split_tree_equals_oracle(shape, seed, kind):
    f, _, _ = make_field(shape, seed, kind, 1e-3, n_gauss=6)
    want = oracle.split_tree(f)
    got = ez().split_tree(f)
    for k in ("max", "pair_saddle", "parent"):
        assert np.array_equal(got[k], want[k]), k
    assert got["branch"] == want["branch"]
    assert got["rounds"] == want["rounds"]


@pytest.mark.gpu
@pytest.mark.parametrize("shape,seed,kind,er", [((24, 20, 16), 71, "noise", 2e-2),
                                                ((40, 40, 40), 72, "gauss", 1e-3)])
def test_gpu_tree_comparison_after_correction(shape, seed, kind, er):
    """The check main.cpp runs after the correction, orig against the
    corrected field (and against the uncorrected decp), as the oracle counts it."""
    orig, decp, xi = make_field(shape, seed, kind, er)
    fixed, _, _ = ez().correct(orig, decp, xi)
    for other in (fixed, decp):
        got = ez().compare_split_trees(orig, other)
        pm, bm = oracle_compare(orig, other)
        assert (got["pair_mismatch"], got["branch_mismatch"]) == (pm, bm)
