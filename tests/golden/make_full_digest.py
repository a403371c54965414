"""Writes the full-size digest fixture: the oracle's corrected field and edit
counts for the 256^3 benchmark workload (sum of 32 Gaussians + 1e-3 * range
noise, seed 2, er = 1e-4), as sha256 digests plus the iteration counters.

The oracle needs about 50 minutes on 6 threads for this field, so the result
is committed (gauss32_256_digest.json) and the GPU test compares digests.

    python tests/golden/make_full_digest.py [threads]
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

SHAPE = (256, 256, 256)
RECIPE = dict(kind="gauss", seed=2, er=1e-4, n_gauss=32)


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    o, d, xi = synth.make_case(RECIPE["kind"], SHAPE, seed=RECIPE["seed"], er=RECIPE["er"],
                               n_gauss=RECIPE["n_gauss"])
    t = time.time()
    fixed, counts, stats = oracle.correct(o, d, xi, threads=threads)
    out = {
        "shape": list(SHAPE), "recipe": RECIPE, "xi": xi,
        "sha256_fixed": hashlib.sha256(fixed.tobytes()).hexdigest(),
        "sha256_counts": hashlib.sha256(counts.tobytes()).hexdigest(),
        "stats": stats, "oracle_seconds": round(time.time() - t, 1), "oracle_threads": threads,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss32_256_digest.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
