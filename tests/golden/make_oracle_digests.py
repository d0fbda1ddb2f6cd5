"""Digests of the oracle's trees on the benchmarked workloads too large for
the oracle to rerun inside a test.

The oracle's restatement of the reference ring search (mode 1, R = 1) runs
here once per workload; the committed file keeps, per workload, the SHA-256
of its canonical edge array (int32 pairs u < v sorted by (u, v), little
endian), the SHA-256 of the edges' float32 d2 bits in the same order, the
edge count and the FP64 weight.  The points come from the library's own
generator (emst_synth), which is identical on every host, so a GPU test can
rebuild the same cloud and compare digests.

Usage: python tests/golden/make_oracle_digests.py [threads]
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import __graft_entry__ as ge  # noqa: E402
import oracle  # noqa: E402

WORKLOADS = [
    # name, kind, n, dim, seed
    ("C4_powerlaw_2d_1e7", 2, 10_000_000, 2, 4),
]


def digest(pts):
    r = oracle.emst(pts, mode=1, R=1)
    e = oracle.canon(r["u"], r["v"]).astype("<i4")
    d2 = oracle.d2_pairs(r["rescaled"], e[:, 0], e[:, 1])
    return {"edges": int(len(e)), "edges_sha256": hashlib.sha256(e.tobytes()).hexdigest(),
            "d2_sha256": hashlib.sha256(d2.astype("<f4").tobytes()).hexdigest(),
            "weight": r["weight"], "rounds": int(r["rounds"])}


def main():
    pkg = ge.load_package()
    if len(sys.argv) > 1:
        oracle.threads(int(sys.argv[1]))
    path = os.path.join(HERE, "oracle_digests.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for name, kind, n, dim, seed in WORKLOADS:
        t = time.time()
        d = digest(pkg.synth(kind, n, dim, seed))
        d.update({"kind": kind, "n": n, "dim": dim, "seed": seed, "oracle": "mode 1 (ring search), R = 1"})
        out[name] = d
        print(name, d, "%.0f s" % (time.time() - t), flush=True)
        json.dump(out, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
