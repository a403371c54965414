"""Independent numpy checks of the correction's termination conditions.

Given orig and a corrected field, recomputes from scratch (with the oracle's
classification and labels, themselves pinned by test_lut / test_oracle) what
main.cpp:277-279 requires to be zero: false critical points, saddle stars,
wrong saddle neighbours, wrong first/last extremum per saddle and wrong
largest/smallest saddle per extremum.
"""
import numpy as np

from oracle import oracle

DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (-1, 1, 0),
        (1, -1, 0), (0, 1, 1), (0, -1, -1), (-1, 0, 1), (1, 0, -1), (-1, 1, 1), (1, -1, -1)]


def neighbours(shape, v):
    D, H, W = shape
    z, r = divmod(v, W * H)
    y, x = divmod(r, W)
    out = []
    for dx, dy, dz in DIRS:
        nx, ny, nz = x + dx, y + dy, z + dz
        if 0 <= nx < W and 0 <= ny < H and 0 <= nz < D:
            out.append(nx + W * (ny + H * nz))
    return out


def key(f, v):
    return (f[v], v)


def saddle_roots(f, lab, s, upper, shape):
    out = []
    for n in neighbours(shape, s):
        if (key(f, n) > key(f, s)) == upper:
            out.append(int(lab[n]))
    return out


def violations(orig, fixed):
    """Counts of every condition main.cpp:277-279 tests (all zero when done)."""
    shape = orig.shape
    o, d = orig.ravel(), fixed.ravel()
    ot, om = oracle.classify(orig)
    dt, dm = oracle.classify(fixed)
    ot, om, dt, dm = ot.ravel(), om.ravel(), dt.ravel(), dm.ravel()
    res = {"types": int((ot != dt).sum()),
           "saddle_stars": int(((ot >= 1) & (ot <= 3) & (om != dm)).sum())}
    ods, oas = (a.ravel() for a in oracle.labels(orig))
    dds, das = (a.ravel() for a in oracle.labels(fixed))
    s2 = np.nonzero((ot == 2) | (ot == 3))[0]
    s1 = np.nonzero((ot == 1) | (ot == 3))[0]
    wn = wf = 0
    for S, up, ol, dl in ((s2, True, ods, dds), (s1, False, oas, das)):
        for s in S:
            a = saddle_roots(o, ol, s, up, shape)
            b = saddle_roots(d, dl, s, up, shape)
            wn += a != b
            ka = sorted(set(a), key=lambda r: key(o, r))
            kb = sorted(set(b), key=lambda r: key(d, r))
            if not kb or ka[0] != kb[0] or ka[-1] != kb[-1]:
                wf += 1
    res["wrong_neighbours"] = wn
    res["wrong_first_last"] = wf
    # largest 2-saddle per maximum / smallest 1-saddle per minimum
    ws = 0
    for S, up, ol, dl, ext in ((s2, True, ods, dds, 4), (s1, False, oas, das, 0)):
        E = np.nonzero(ot == ext)[0]
        def best(f, lab):
            ex = sorted(E, key=lambda v: key(f, v))
            glob = ex[-1] if up else ex[0]
            m = {}
            for s in S:
                for r in set(saddle_roots(f, lab, s, up, shape)):
                    if r == glob or ot[r] != ext:
                        continue
                    cur = m.get(r)
                    if cur is None or (key(f, s) > key(f, cur) if up else key(f, s) < key(f, cur)):
                        m[r] = s
            return m
        mo, md = best(o, ol), best(d, dl)
        ws += sum(mo.get(e) is not None and md.get(e) is not None and mo.get(e) != md.get(e)
                  for e in E)
    res["wrong_extremum_saddle"] = ws
    return res


def _shift_index(shape, idx, d):
    """Neighbour ids of the vertices idx in direction d (-1 outside the grid)."""
    D, H, W = shape
    dx, dy, dz = DIRS[d]
    z, r = np.divmod(idx, W * H)
    y, x = np.divmod(r, W)
    nx, ny, nz = x + dx, y + dy, z + dz
    ok = (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H) & (nz >= 0) & (nz < D)
    return np.where(ok, nx + W * (ny + H * nz), -1)


def _side_roots(f, lab, S, up, shape):
    """Per saddle of S and per direction: the root of the neighbour when it is
    upper (up) / lower, else -1.  Shape (14, len(S))."""
    out = np.full((len(DIRS), S.size), -1, np.int64)
    fs = f[S]
    for d in range(len(DIRS)):
        nb = _shift_index(shape, S, d)
        ok = nb >= 0
        n = np.where(ok, nb, 0)
        fn = f[n]
        side = (fn > fs) | ((fn == fs) & (n > S)) if up else (fn < fs) | ((fn == fs) & (n < S))
        side &= ok
        out[d] = np.where(side, lab[n], -1)
    return out


def _ends(f, R):
    """First / last root by (value, id) per column of R (-1 entries skipped)."""
    valid = R >= 0
    Rz = np.where(valid, R, 0)
    fv = f[Rz]
    lo_v = np.full(R.shape[1], np.inf)
    lo_i = np.full(R.shape[1], -1, np.int64)
    hi_v = np.full(R.shape[1], -np.inf)
    hi_i = np.full(R.shape[1], -1, np.int64)
    for d in range(R.shape[0]):
        v, i, ok = fv[d], R[d], valid[d]
        lt = ok & ((v < lo_v) | ((v == lo_v) & (i < lo_i)) | (lo_i < 0))
        lo_v = np.where(lt, v, lo_v)
        lo_i = np.where(lt, i, lo_i)
        gt = ok & ((v > hi_v) | ((v == hi_v) & (i > hi_i)) | (hi_i < 0))
        hi_v = np.where(gt, v, hi_v)
        hi_i = np.where(gt, i, hi_i)
    return lo_i, hi_i


def _best_saddle(f, S, R, up, ot, ext, E):
    """Largest (up) / smallest 2- / 1-saddle per extremum of E over the
    (saddle, root) pairs of R: an array over E of saddle ids (-1: none)."""
    if not E.size or not S.size:
        return np.full(E.size, -1, np.int64)
    ke = np.lexsort((E, f[E]))
    glob = E[ke[-1]] if up else E[ke[0]]
    order = np.lexsort((S, f[S]))          # ascending (value, id)
    rank = np.empty(S.size, np.int64)
    rank[order] = np.arange(S.size)
    best = np.full(f.size, -1 if up else S.size, np.int64)
    for d in range(R.shape[0]):
        r = R[d]
        ok = (r >= 0)
        ok[ok] &= (r[ok] != glob) & (ot[r[ok]] == ext)
        if up:
            np.maximum.at(best, r[ok], rank[ok])
        else:
            np.minimum.at(best, r[ok], rank[ok])
    b = best[E]
    has = (b >= 0) & (b < S.size)
    return np.where(has, S[order[np.clip(b, 0, max(0, S.size - 1))]], -1)


def violations_fast(orig, fixed):
    """violations() vectorised over the saddles (same counts, for large fields)."""
    shape = orig.shape
    o, d = orig.ravel(), fixed.ravel()
    ot, om = oracle.classify(orig)
    dt, dm = oracle.classify(fixed)
    ot, om, dt, dm = ot.ravel(), om.ravel(), dt.ravel(), dm.ravel()
    res = {"types": int((ot != dt).sum()),
           "saddle_stars": int(((ot >= 1) & (ot <= 3) & (om != dm)).sum())}
    del om, dm, dt
    ods, oas = (a.ravel() for a in oracle.labels(orig))
    dds, das = (a.ravel() for a in oracle.labels(fixed))
    s2 = np.nonzero((ot == 2) | (ot == 3))[0].astype(np.int64)
    s1 = np.nonzero((ot == 1) | (ot == 3))[0].astype(np.int64)
    wn = wf = ws = 0
    for S, up, ol, dl, ext in ((s2, True, ods, dds, 4), (s1, False, oas, das, 0)):
        Ra = _side_roots(o, ol, S, up, shape)
        Rb = _side_roots(d, dl, S, up, shape)
        # the lists are compared position by position over the sides' roots in
        # direction order (label_count of get_wrong_split_neighbors)
        pa = np.take_along_axis(Ra, np.argsort(Ra < 0, axis=0, kind="stable"), axis=0)
        pb = np.take_along_axis(Rb, np.argsort(Rb < 0, axis=0, kind="stable"), axis=0)
        wn += int(np.any(pa != pb, axis=0).sum())
        del pa, pb
        fa, la = _ends(o, Ra)
        fb, lb = _ends(d, Rb)
        wf += int(((fb < 0) | (fa != fb) | (la != lb)).sum())
        E = np.nonzero(ot == ext)[0].astype(np.int64)
        mo = _best_saddle(o, S, Ra, up, ot, ext, E)
        md = _best_saddle(d, S, Rb, up, ot, ext, E)
        ws += int(((mo >= 0) & (md >= 0) & (mo != md)).sum())
    res["wrong_neighbours"] = wn
    res["wrong_first_last"] = wf
    res["wrong_extremum_saddle"] = ws
    return res
