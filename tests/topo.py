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
