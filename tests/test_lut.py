"""Phase (a) pins: the LUT against the reference's LUT.bin, and vertex
classification against a brute-force link-component count and closed forms."""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle
from tests.topo import DIRS

LUT_SHA256 = "10af892bf6e660f3a52841a1d59330344386f3bd14db39c10742e60e6077a912"


def test_lut_fixture_is_the_reference_table(golden):
    import hashlib
    data = open(os.path.join(golden, "LUT.bin"), "rb").read()
    assert len(data) == 99200
    assert hashlib.sha256(data).hexdigest() == LUT_SHA256


def test_oracle_lut_matches_reference(golden):
    ref = np.fromfile(os.path.join(golden, "LUT.bin"), dtype="<i4")
    assert np.array_equal(oracle.lut(), ref)
    # value histogram of LUT.bin
    assert np.bincount(ref).tolist() == [27, 6859, 6859, 200, 27, 10828]


def test_product_lut_matches_reference(golden):
    import extrem_z_amd as ez
    ref = np.fromfile(os.path.join(golden, "LUT.bin"), dtype="<i4")
    assert np.array_equal(ez.lut(), ref)


def brute_type(vals, centre_idx, centre, nbr_ids, present):
    """Union-find over the link of one vertex, straight from the definition:
    link vertices are the existing stencil neighbours, link edges join
    neighbours a, b whose offset difference is itself a stencil direction."""
    lower = []
    upper = []
    for d in range(14):
        if not present[d]:
            continue
        v, i = vals[d], nbr_ids[d]
        is_lower = v < centre or (v == centre and i < centre_idx)
        (lower if is_lower else upper).append(d)
    def comps(S):
        parent = {d: d for d in S}
        def f(a):
            while parent[a] != a:
                a = parent[a]
            return a
        for a, b in itertools.combinations(S, 2):
            diff = tuple(DIRS[b][k] - DIRS[a][k] for k in range(3))
            if diff in DIRS:
                parent[f(a)] = f(b)
        return len({f(d) for d in S})
    lc, uc = comps(lower), comps(upper)
    if not upper:
        return 4
    if not lower:
        return 0
    if lc > 1 and uc > 1:
        return 3
    if lc > 1:
        return 1
    if uc > 1:
        return 2
    return 5


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("shape", [(5, 4, 3), (3, 3, 3), (6, 2, 4), (7, 1, 3)])
def test_classification_matches_bruteforce(shape, seed):
    W, H, D = shape
    rng = np.random.default_rng(seed)
    # few distinct values -> many ties broken by index
    f = rng.integers(0, 4, size=(D, H, W)).astype(np.float64)
    types, masks = oracle.classify(f)
    flat = f.ravel()
    for z, y, x in itertools.product(range(D), range(H), range(W)):
        v = x + W * (y + H * z)
        present, vals, ids = [], [], []
        for dx, dy, dz in DIRS:
            nx, ny, nz = x + dx, y + dy, z + dz
            ok = 0 <= nx < W and 0 <= ny < H and 0 <= nz < D
            present.append(ok)
            i = nx + W * (ny + H * nz) if ok else -1
            ids.append(i)
            vals.append(flat[i] if ok else 0.0)
        assert types.ravel()[v] == brute_type(vals, v, flat[v], ids, present)
        want = sum(1 << d for d in range(14) if present[d] and
                   (vals[d] < flat[v] or (vals[d] == flat[v] and ids[d] < v)))
        assert masks.ravel()[v] == want


def grid(shape):
    W, H, D = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


def test_ramp_has_one_min_one_max():
    x, y, z = grid((6, 5, 4))
    t, _ = oracle.classify(x + 2 * y + 4 * z)
    assert np.bincount(t.ravel(), minlength=6).tolist() == [1, 0, 0, 0, 1, 6 * 5 * 4 - 2]
    assert t[0, 0, 0] == 0 and t[-1, -1, -1] == 4


def test_gaussian_bump_has_one_max():
    x, y, z = grid((9, 9, 9))
    f = np.exp(-((x - 4) ** 2 + (y - 4) ** 2 + (z - 4) ** 2) / 8.0)
    t, _ = oracle.classify(f)
    assert (t == 4).sum() == 1 and t[4, 4, 4] == 4


def test_saddles_closed_form():
    x, y, z = grid((5, 5, 5))
    f = (x - 2) ** 2 + (y - 2) ** 2 - 2.5 * (z - 2) ** 2
    t, _ = oracle.classify(f)
    assert t[2, 2, 2] == 1          # two lower cones, one upper ring
    t, _ = oracle.classify(-f)
    assert t[2, 2, 2] == 2


def reference_triangles(W, H):
    """The planar triangles of a W x H x 1 field as the reference enumerates
    them (src/ExtreM.cpp:597-627): cell (x, y) holds (x,y),(x+1,y),(x,y+1)
    and (x+1,y+1),(x,y+1),(x+1,y)."""
    tris = []
    for t in range(2 * (W - 1) * (H - 1)):
        b = t // 2
        x, y = b % (W - 1), b // (W - 1)
        if t % 2 == 0:
            tris.append((x + y * W, x + 1 + y * W, x + (y + 1) * W))
        else:
            tris.append((x + 1 + (y + 1) * W, x + (y + 1) * W, x + 1 + y * W))
    return tris


def mesh_type(flat, v, tris):
    """Critical type of v from its link in the triangle mesh alone: link
    vertices are the other corners of the triangles around v, link edges the
    sides opposite v. No stencil directions are used."""
    link_edges, link = [], set()
    for t in tris:
        if v in t:
            a, b = (u for u in t if u != v)
            link_edges.append((a, b))
            link.update((a, b))
    below = {u for u in link if (flat[u], u) < (flat[v], v)}
    above = link - below

    def comps(S):
        parent = {u: u for u in S}

        def f(a):
            while parent[a] != a:
                a = parent[a]
            return a
        for a, b in link_edges:
            if a in S and b in S:
                parent[f(a)] = f(b)
        return len({f(u) for u in S})
    lc, uc = comps(below), comps(above)
    if not above:
        return 4
    if not below:
        return 0
    return {(True, True): 3, (True, False): 1, (False, True): 2}.get((lc > 1, uc > 1), 5)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("W,H", [(6, 5), (9, 3), (2, 7), (11, 11)])
def test_depth1_is_the_reference_triangle_mesh(W, H, seed):
    """With the 14-neighbour stencil, a field of depth 1 is the planar
    triangle mesh of src/ExtreM.cpp:597-627 with its full links: the in-plane
    stencil directions (+-x, +-y, +-(-1,1,0)) are exactly the mesh's edges, so
    each vertex's type equals the one its link in that mesh gives (interior
    saddles have as many lower as upper link components, so they carry both
    saddle bits).  The reference's CPU run itself (maxNeighbors = 6) is the
    plane adjacency, below."""
    rng = np.random.default_rng(100 + seed)
    f = rng.integers(0, 5, size=(1, H, W)).astype(np.float64)
    types, _ = oracle.classify(f)
    flat = f.ravel()
    tris = reference_triangles(W, H)
    got = [mesh_type(flat, v, tris) for v in range(W * H)]
    assert types.ravel().tolist() == got
    # interior vertices: saddle means both bits (lower and upper cycles alternate)
    t2 = types.reshape(H, W)[1:-1, 1:-1]
    assert not np.isin(t2, (1, 2)).any()


def classify_vertex1(flat, v, W, H, tris):
    """classifyVertex1 (src/ExtreM.cpp:703-891) with maxNeighbors = 6 on a
    depth-1 field, restated: the stars are the adjacency neighbours of
    computeAdjacency (src/ExtreM.cpp:543-575: the first six directions of
    src/ExtreM.cpp:16-21, so +x, -x, +y, -y in the plane); two star
    neighbours of one side are joined when they are the other two corners of
    a triangle around v (vertex_cells, src/ExtreM.cpp:1585-1595); the type
    rule of src/ExtreM.cpp:880-886."""
    x, y = v % W, v // W
    adj = [(x + dx) + W * (y + dy) for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1))
           if 0 <= x + dx < W and 0 <= y + dy < H]
    key = (flat[v], v)
    lower = [u for u in adj if (flat[u], u) < key]
    upper = [u for u in adj if (flat[u], u) > key]
    parent = {u: u for u in adj}

    def find(a):
        while parent[a] != a:
            a = parent[a]
        return a
    for t in tris:
        if v not in t:
            continue
        a, b = (u for u in t if u != v)
        for side in (lower, upper):
            if a in side and b in side:
                parent[find(a)] = find(b)
    lc = len({find(u) for u in lower})
    uc = len({find(u) for u in upper})
    if lc == 0 and uc == 1:
        return 0
    if lc == 1 and uc == 0:
        return 4
    if lc == 1 and uc == 1:
        return 5
    if lc > 1 and uc > 1:
        return 3
    if lc > 1:
        return 1
    if uc > 1:
        return 2
    return 5


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("W,H", [(6, 5), (9, 3), (2, 7), (11, 11), (7, 1)])
def test_plane_adjacency_is_classifyVertex1(W, H, seed):
    """With the CPU path's plane adjacency, the oracle's LUT classification
    and lower stars equal classifyVertex1's, vertex by vertex, on tie-heavy
    fields; its labels are the steepest paths over the four neighbours."""
    rng = np.random.default_rng(300 + seed)
    f = rng.integers(0, 5, size=(1, H, W)).astype(np.float64)
    flat = f.ravel()
    tris = reference_triangles(W, H) if W > 1 and H > 1 else []
    with oracle.plane_adjacency():
        types, masks = oracle.classify(f)
        ds, asc = oracle.labels(f)
        assert all(ds.ravel()[v] == oracle.max_label(f, v) for v in range(W * H))
    assert types.ravel().tolist() == [classify_vertex1(flat, v, W, H, tris)
                                      for v in range(W * H)]
    # lower stars use the four directions only (bits 0-3)
    assert not (masks.astype(np.int64) & ~0xF).any()
    # steepest ascent over the four neighbours, by (value, id)
    for v in range(W * H):
        x, y = v % W, v // W
        cur = v
        while True:
            cx, cy = cur % W, cur // W
            nb = [(cx + dx) + W * (cy + dy) for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1))
                  if 0 <= cx + dx < W and 0 <= cy + dy < H]
            best = max([cur] + nb, key=lambda u: (flat[u], u))
            if best == cur:
                break
            cur = best
        assert ds.ravel()[v] == cur
    # the stencil mode is back outside the block: the diagonals are present
    t14, _ = oracle.classify(f)
    assert t14.shape == types.shape


def test_plane_adjacency_correction_terminates_with_invariants():
    from tests.conftest import make_field
    for shape, seed, kind, er in [((40, 30, 1), 61, "gauss", 1e-3), ((25, 19, 1), 62, "noise", 2e-2)]:
        orig, decp, xi = make_field(shape, seed, kind, er)
        with oracle.plane_adjacency():
            fixed, counts, st = oracle.correct(orig, decp, xi)
            t0, m0 = oracle.classify(orig)
            t1, m1 = oracle.classify(fixed)
        assert st["status"] == 0
        assert np.all(np.abs(fixed - orig) <= xi + np.spacing(np.abs(orig) + xi))
        assert np.array_equal(t0, t1)
        sad = (t0 >= 1) & (t0 <= 3)
        assert np.array_equal(m0[sad], m1[sad])
