"""Two contexts on two host threads at once (a batch of 2D clouds and a batch
of 3D clouds, each on its own stream), repeated: every cloud finishes its
tree in every repetition, with the same per-cloud weight, edge count and
last round as the first repetition.  Inside the per-round component search,
waves read component minima that other waves are lowering; a read that
steered a wave's control flow once split a wave and hooked a point to
another cloud (one extra edge in about 1 of 130 repetitions of the
4096-cloud batch)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPEATS = 40


def test_two_threads_two_batches(pkg):
    sizes, dims, kinds, make = pkg.batch_workload(1024, 7)
    groups = []
    for d in (2, 3):
        idx = np.nonzero(dims == d)[0]
        groups.append({"dim": d, "sizes": sizes[idx], "clouds": [make(int(i)) for i in idx]})
    print("clouds", len(sizes), "points", int(sizes.sum()), flush=True)
    # the per-round component search in both halves, with the per-query pass
    opts = [{"bucket_size": 64, "seed_scale": 2.0, "round_search": 1},
            {"bucket_size": 64, "seed_scale": 1.0, "search_split": 3, "round_search": 1}]
    ctxs = [pkg.Calculateur(dim=g["dim"], **o) for g, o in zip(groups, opts)]
    first = None
    try:
        for rep in range(REPEATS):
            errs = []

            def work(c, g):
                try:
                    c.initialize_clouds(g["clouds"])
                    c.run()
                except Exception as ex:     # reported below, on the main thread
                    errs.append(ex)

            ts = [threading.Thread(target=work, args=(c, g)) for c, g in zip(ctxs, groups)]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
            assert not errs, errs
            res = [c.cloud_results() for c in ctxs]
            for (w, e, last), g in zip(res, groups):
                bad = np.nonzero(e != np.maximum(g["sizes"] - 1, 0))[0]
                assert len(bad) == 0, (rep, g["dim"], [(int(i), int(g["sizes"][i]), int(e[i])) for i in bad[:5]])
            if first is None:
                first = res
            else:
                for (w, e, last), (w0, e0, l0) in zip(res, first):
                    assert np.array_equal(w, w0) and np.array_equal(e, e0) and np.array_equal(last, l0), rep
            if rep % 10 == 0:
                print("repetition", rep, "ok", flush=True)
    finally:
        for c in ctxs:
            c.close()
