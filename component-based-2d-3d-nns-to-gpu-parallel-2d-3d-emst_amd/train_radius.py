"""Train the seed-radius MLP from the plain method's own searches.

--search cone (default): the round-1 cone search's model, per point the log
of the radius within which all its cones close.  --search comp: the per-round
component search's model, per searching point and round the log of its
component's shortest outgoing edge (pairs logged by the library during plain
runs, options.log_samples).  Both relative to the point's leaf spacing.

Two training sets (--clouds), 2^20 (features, log true radius) pairs per
dimension, logged from the plain method:
  uniform  the uniform clouds of the reference-path configs only (2D: 10^4
           points seed 0 and 10^6 points seed 1; 3D: 10^6 points seed 2);
  mixed    those plus 10^6-point clouds of the clustered and power-law
           generators at seeds the benchmark does not use (11-16).
Training runs on the GPU with the library's forward/backward kernels
(emst_mlp_train).  The seed offset (added to the predicted log radius) is then
picked by timing the cone search on held-out clouds (uniform; for mixed also
one clustered and one power-law cloud, again at unused seeds).

    python train_radius.py --out <dir> [--clouds mixed] [--search comp]
        (needs a GPU; writes radius_2d.npy, radius_3d.npy and train_radius.json,
        or radius_comp_2d.npy, radius_comp_3d.npy and train_radius_comp.json)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    import importlib.util
    if "emst_amd" in sys.modules:
        return sys.modules["emst_amd"]
    spec = importlib.util.spec_from_file_location("emst_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["emst_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


TRAIN_CLOUDS = {
    "uniform": {2: [(0, 10_000, 0), (0, 1_000_000, 1)], 3: [(0, 1_000_000, 2)]},
    "mixed": {2: [(0, 10_000, 0), (0, 1_000_000, 1), (2, 1_000_000, 11), (1, 1_000_000, 12)],
              3: [(0, 1_000_000, 2), (1, 1_000_000, 13), (2, 1_000_000, 14)]},
}
VALID_CLOUDS = {
    "uniform": {2: [(0, 2_000_000, 101)], 3: [(0, 2_000_000, 102)]},
    "mixed": {2: [(0, 2_000_000, 101), (2, 2_000_000, 15)], 3: [(0, 2_000_000, 102), (1, 2_000_000, 16)]},
}
OFFSETS = {"cone": (-0.25, 0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5),
           "comp": (-0.5, -0.25, 0.0, 0.25, 0.5, 0.75, 1.0)}


def samples(pkg, dim, clouds, search="cone"):
    Xs, Ts = [], []
    for kind, n, seed in TRAIN_CLOUDS[clouds][dim]:
        pts = pkg.synth(kind, n, dim, seed)
        if search == "comp":
            with pkg.Calculateur(dim=dim, round_search=1, log_samples=1) as c:
                c.initialize(pts)
                c.run()
                X, T = c.radius_samples()
        else:
            with pkg.Calculateur(dim=dim) as c:
                c.initialize(pts)
                X, T = c.radius_samples()
        keep = T > -19.0
        Xs.append(X[keep])
        Ts.append(T[keep])
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    rng = np.random.default_rng(dim)
    idx = rng.choice(len(T), size=min(len(T), 1 << 20), replace=False)
    return X[idx], T[idx]


def search_ms(pkg, dim, pts, model, reps=3, search="cone"):
    """Cone search: the search's device time.  Component search: the whole
    run's (its searches are spread over the rounds)."""
    with pkg.Calculateur(dim=dim, round_search=1 if search == "comp" else 0) as c:
        c.initialize(pts)
        if model is not None:
            c.set_radius_model(model)
        best = None
        for _ in range(reps):
            c.initialize(pts)
            c.run()
            t = c.times()[3 if search == "comp" else 1]
            best = t if best is None else min(best, t)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models"))
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--clouds", choices=sorted(TRAIN_CLOUDS), default="uniform")
    ap.add_argument("--search", choices=("cone", "comp"), default="cone")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    pkg = load_pkg()
    report = {"clouds": args.clouds, "search": args.search}
    tag = "comp_" if args.search == "comp" else ""
    for dim in (2, 3):
        t0 = time.time()
        X, T = samples(pkg, dim, args.clouds, args.search)
        print(f"{dim}D: {len(T)} pairs, target mean {T.mean():.3f} var {T.var():.4f}", flush=True)
        model, losses = pkg.train_radius_model(X, T, epochs=args.epochs, batch=1024, lr=2e-3, seed=dim)
        print(f"{dim}D: loss per epoch {np.round(losses, 4).tolist()}", flush=True)
        vclouds = [pkg.synth(kind, n, dim, seed) for kind, n, seed in VALID_CLOUDS[args.clouds][dim]]
        plain = sum(search_ms(pkg, dim, v, None, search=args.search) for v in vclouds)
        timings = {}
        for off in OFFSETS[args.search]:
            m = model.copy()
            m[-1] = off
            timings[off] = sum(search_ms(pkg, dim, v, m, search=args.search) for v in vclouds)
            print(f"{dim}D: offset {off:+.2f}: {args.search} {timings[off]:.2f} ms (plain {plain:.2f})", flush=True)
        best = min(timings, key=timings.get)
        model[-1] = best
        np.save(os.path.join(args.out, f"radius_{tag}{dim}d.npy"), model)
        # features and targets are measured on leaf cells of the library's
        # default size (16 points in 2D, 64 in 3D)
        report[f"{dim}d"] = {"leaf_points": 16 if dim == 2 else 64, "pairs": int(len(T)), "target_var": float(T.var()),
                             "loss_per_epoch": [float(x) for x in losses], "offset": best,
                             "valid_search_ms": {"plain": plain, **{str(k): v for k, v in timings.items()}},
                             "seconds": time.time() - t0}
    with open(os.path.join(args.out, f"train_radius{'_comp' if tag else ''}.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report))


if __name__ == "__main__":
    main()
