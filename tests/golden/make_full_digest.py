"""Writes a full-size digest fixture: the oracle's corrected field and edit
counts for one large workload, as sha256 digests plus the iteration counters.

Workloads:
  256  sum of 32 Gaussians + 1e-3 * range noise, 256^3, seed 2, er = 1e-4
       (gauss32_256_digest.json; about 50 minutes on 6 threads)
  512  Gaussian random field with power spectrum k^-3, 512^3, seed 3,
       normalised to [0, 1], xi = 1e-3 (grf_512_digest.json)

The oracle takes tens of minutes to hours for these fields, so the results
are committed and the GPU tests compare digests.

    python tests/golden/make_full_digest.py [256|512] [threads]
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from extrem_z_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402

WORKLOADS = {
    "256": dict(shape=(256, 256, 256), recipe=dict(kind="gauss", seed=2, er=1e-4, n_gauss=32),
                out="gauss32_256_digest.json"),
    "512": dict(shape=(512, 512, 512), recipe=dict(kind="grf", seed=3, xi=1e-3),
                out="grf_512_digest.json"),
}


def make(name):
    w = WORKLOADS[name]
    r = dict(w["recipe"])
    kind = r.pop("kind")
    return synth.make_case(kind, w["shape"], **r)


def main():
    name = sys.argv[1] if len(sys.argv) > 1 else "256"
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 6
    w = WORKLOADS[name]
    o, d, xi = make(name)
    t = time.time()
    fixed, counts, stats = oracle.correct(o, d, xi, threads=threads)
    out = {
        "shape": list(w["shape"]), "recipe": w["recipe"], "xi": xi,
        "sha256_fixed": hashlib.sha256(fixed.tobytes()).hexdigest(),
        "sha256_counts": hashlib.sha256(counts.tobytes()).hexdigest(),
        "stats": stats, "oracle_seconds": round(time.time() - t, 1), "oracle_threads": threads,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), w["out"])
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
