"""EMST benchmark: Mpoints/s of the whole GPU path on one MI355X.

Workload (default): 10^7 points in 3D, mixture of 64 isotropic Gaussians
(centres uniform in [0,1)^3, sigma log-uniform 0.002-0.02) plus 10% uniform
background, float32, clipped to the unit cube, generator seed 3.  One step =
one whole EMST from device-resident points to the device-resident edge list
and its FP64 weight: bbox, rescale, Morton sort, hierarchy, the search (the
round-1 cone search, or the per-round component search), Boruvka rounds,
weight.  Every form (cone search plain and learned-seed, component search) is
timed on the same points and must build the same tree; `value` is the
headline form's: the fastest measured, unless --round-search picks a search.

Prints one JSON line (rank 0).  --gpus above 1 is refused: the engine is one
GPU per process and this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def profiled_traffic(args, kernel="k_cone_search"):
    """HBM bytes per launch of `kernel` (for k_comp_search: all its dispatches
    of one EMST) from the committed rocprofv3 summary of this workload
    (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy), or None when no
    summary covers it."""
    tags = {(1, 10_000_000, 3): "gauss64_1e7_3d", (0, 1_000_000, 3): "uniform_1e6_3d"}
    tag = tags.get((args.kind, args.n, args.dim))
    if tag is None or args.cone_dim not in (0, args.dim):
        return None
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--kind", type=int, default=1, help="0 uniform, 1 Gaussian mixture, 2 power-law")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--bucket", type=int, default=0)
    ap.add_argument("--cone-dim", type=int, default=0,
                    help="2 with --dim 3: the 2.5D mode (wedges over the xy plane, 3D distances)")
    ap.add_argument("--plain", action="store_true", help="plain seed only (skip the learned form)")
    ap.add_argument("--round-search", type=int, default=-1, choices=(-1, 0, 1),
                    help="headline form: the fastest of the cone search's forms (0), of the per-round "
                         "component search's (1), or of all (-1)")
    ap.add_argument("--cpu-sample", type=int, default=200_000,
                    help="points of the same distribution timed on the CPU oracle (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (scales rocprof byte counters)")
    ap.add_argument("--batch", type=int, default=0, metavar="CLOUDS",
                    help="batched workload instead: CLOUDS clouds, sizes log-uniform 1e2-1e5, half 2D half "
                         "3D, each uniform or 8 Gaussians (generator seed --batch-seed)")
    ap.add_argument("--batch-seed", type=int, default=5)
    ap.add_argument("--no-autotune", action="store_true", help="batch: keep the default leaf size and seed")
    ap.add_argument("--tsplib", action="store_true",
                    help="time every TSPLIB instance of the reference's 2D tests (tests/golden/tsplib)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the EMST engine runs one context per GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__ as ge

    pkg = ge.load_package()
    pkg.lib()
    if args.batch:
        run_batch(args, pkg)
        return
    if args.tsplib:
        run_tsplib(args, pkg)
        return
    names = {0: "uniform", 1: "gauss64+10%bg", 2: "powerlaw16"}
    pts = pkg.synth(args.kind, args.n, args.dim, args.seed)
    dev = torch.from_numpy(pts).to("cuda:0")
    torch.cuda.synchronize()

    cdim = args.cone_dim or args.dim
    K = 2 * cdim * (1 << (cdim - 1))

    def load_model(tag):
        path = os.path.join(ge.PKG_DIR, "models", f"radius_{tag}{args.dim}d.npy")
        if args.plain or cdim != args.dim or not os.path.exists(path):
            return None, None
        meta = os.path.join(ge.PKG_DIR, "models", f"train_radius{'_comp' if tag else ''}.json")
        return np.load(path), (json.load(open(meta)).get("clouds") if os.path.exists(meta) else None)

    cone_model, model_clouds = load_model("")
    comp_model, _ = load_model("comp_")

    def timed(form_model, round_search):
        ctx = pkg.Calculateur(dim=args.dim, bucket_size=args.bucket, cone_dim=args.cone_dim,
                              round_search=round_search)
        if form_model is not None:
            ctx.set_radius_model(form_model)
        search_ms, total_ms = [], []

        def step():
            ctx.initialize_device(dev.data_ptr(), args.n, args.dim)
            ctx.run()
            return ctx.evaluate()

        for _ in range(args.warmup):
            w, m = step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            w, m = step()
            t = ctx.times()
            search_ms.append(t[1])
            total_ms.append(t[3])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        u, v, _ = ctx.getLinks()
        st = pkg.search_stats(ctx)
        samples = ctx.radius_samples() if (form_model is not None and not round_search) else None
        ctx.close()
        assert m == args.n - 1, f"tree has {m} edges"
        med = lambda xs: sorted(xs)[len(xs) // 2]
        return {"dt": dt, "w": w, "m": m, "search_ms": med(search_ms), "total_ms": med(total_ms),
                "edges_digest": hash((u.tobytes(), v.tobytes())), "stats": st, "samples": samples,
                "mpts": args.n * args.steps / dt / 1e6}

    # every form on the same points; each must build the same tree
    forms = {"cone_plain": timed(None, 0)}
    if cone_model is not None:
        forms["cone_learned"] = timed(cone_model, 0)
    if cdim == args.dim:
        forms["comp_plain"] = timed(None, 1)
        if comp_model is not None:
            forms["comp_learned"] = timed(comp_model, 1)
    ref = forms["cone_plain"]
    for name, f in forms.items():
        assert f["edges_digest"] == ref["edges_digest"] and f["w"] == ref["w"], f"{name} built a different tree"
    # the headline form: the fastest measured (--round-search: the fastest of
    # that search's forms)
    if args.round_search < 0:
        head_name = max(forms, key=lambda k: forms[k]["mpts"])
    else:
        fam = "comp" if args.round_search == 1 else "cone"
        head_name = max((k for k in forms if k.startswith(fam)), key=lambda k: forms[k]["mpts"],
                        default="cone_plain")
    fam = "comp" if head_name.startswith("comp") else "cone"
    rs = 1 if fam == "comp" else 0
    head = forms[head_name]
    plain, learned = forms["cone_plain"], forms.get("cone_learned")
    seed_report = None
    if learned is not None:
        # hit rate: points whose seed already covers every non-empty cone's
        # nearest point (pass 0 closes them all); savings: traversal work
        X, T = learned["samples"]
        has = T > -19.0
        y = pkg.mlp_eval(cone_model, X) + cone_model[-1]
        ps, ls = plain["stats"], learned["stats"]
        seed_report = {
            "hit_rate_learned": round(float(np.mean((y >= T)[has])), 4),
            "hit_rate_plain": round(float(np.mean((np.log(3.0) >= T)[has])), 4),
            "nodes_tested_per_packet": [round(ps["nodes"] / ps["packets"], 1), round(ls["nodes"] / ls["packets"], 1)],
            "pairs_per_point": [round(ps["pairs"] / args.n, 1), round(ls["pairs"] / args.n, 1)],
            "node_test_savings": round(1 - ls["nodes"] / ps["nodes"], 4),
            "pair_savings": round(1 - ls["pairs"] / ps["pairs"], 4),
        }
    ms_per_step = head["dt"] * 1e3 / args.steps
    value = head["mpts"]
    w, m = head["w"], head["m"]

    if fam == "cone":
        # dominant kernel: the round-1 cone search; algorithmic bytes per
        # point = 8*D (own coordinates) + 4 + 8*K (one (id, d2) per cone out);
        # time from HIP events on the context stream
        alg_bytes = args.n * (8 * args.dim + 4 + 8 * K)
        k_ms = head["search_ms"]
        kernel = "k_cone_search"
    else:
        # dominant kernel: the component search's dispatches of one EMST (two
        # per round); algorithmic bytes per query and dispatch = 44: its point
        # (16), label (4), previous key (8) and list entry (4) in, its result
        # (corr 4, key 8) out.  Time from HIP events around each dispatch, in
        # one traced (host-issued) run on the context stream
        with pkg.Calculateur(dim=args.dim, bucket_size=args.bucket, round_search=1, trace=True) as c:
            if head_name == "comp_learned":
                c.set_radius_model(comp_model)
            c.initialize_device(dev.data_ptr(), args.n, args.dim)
            c.run()
            k_ms = c.times()[4]
            queries = pkg.search_stats(c)["comp_queries"]
        alg_bytes = 44 * queries
        kernel = "k_comp_search"
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": profiled_traffic(args, kernel),
                "kernel": kernel, "kernel_ms": round(k_ms, 3)}

    cpu = cpu_baseline(args, pkg, cdim) if args.cpu_sample > 0 else None

    desc = {"cone_plain": "cone search, plain seed", "cone_learned": "cone search, learned seed",
            "comp_plain": "per-round component search, plain seed",
            "comp_learned": "per-round component search, learned seed"}
    mp = lambda k: round(forms[k]["mpts"], 3) if k in forms else None
    out = {
        "metric": "EMST Mpoints/s at 10^7 points (fastest exact form; learned-radius vs plain in config)",
        "value": round(value, 3), "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": f"EMST {args.n} pts {args.dim}D{' (2.5D: xy wedges)' if cdim < args.dim else ''} "
                               f"{names.get(args.kind)} seed {args.seed}",
                   "n": args.n, "dim": args.dim, "form": desc[head_name],
                   "edges": int(m), "weight": w,
                   "gpu_total_ms_median": round(head["total_ms"], 3),
                   "mpts_by_form": {k: mp(k) for k in desc},
                   "same_tree_all_forms": True,
                   # learned vs plain: the cone search's seed model (the
                   # component search ships none: its model measured slower)
                   "plain_mpts": mp("cone_plain"),
                   "plain_search_ms": round(plain["search_ms"], 3),
                   "learned_mpts": mp("cone_learned"),
                   "learned_search_ms": round(learned["search_ms"], 3) if learned else None,
                   "learned_vs_plain": round(plain["dt"] / learned["dt"], 3) if learned else None,
                   "radius_model_training_clouds": model_clouds,
                   "learned_seed": seed_report,
                   "cone_dim": cdim, "cycles_cut": int(forms["cone_plain"]["stats"]["cycles_cut"]),
                   "round_search": rs},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(args, pkg, cdim):
    """The oracle (the project's C restatement of the reference path) timed on
    this host's cores, built here with -O3 -march=native.  value: a sample of
    the benchmarked workload (--cpu-sample points of the same distribution and
    seed; per-point searches in parallel, rounds sequential).  Also the
    reference path's own CPU configuration: 10^4 uniform 2D points, seed 0,
    with 1 thread and with all cores, median of 5 runs after 1 warm-up, over
    the span of the reference's cuda_duree (after the cell build to the
    finished tree, SolutionEMSTRW.cu:85-97,133-165)."""
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    med = lambda xs: sorted(xs)[len(xs) // 2]
    with tempfile.TemporaryDirectory() as tmp:
        try:
            native = oracle.build_native(tmp)
            build = "gcc -O3 -march=native -ffp-contract=off -fopenmp"
        except Exception as e:          # no compiler on this host: the prebuilt -O2 library
            native = oracle.lib()
            build = f"prebuilt gcc -O2 (native build failed: {type(e).__name__})"
        prev = oracle.select(native)
    cores = oracle.threads(0)
    try:
        c1 = pkg.synth(0, 10_000, 2, 0)
        c1_res = {}
        for th in (1, cores):
            oracle.threads(th)
            spans, walls = [], []
            for rep in range(6):
                t1 = time.perf_counter()
                r = oracle.emst(c1, mode=1, R=1)
                wall = time.perf_counter() - t1
                assert len(r["u"]) == len(c1) - 1
                if rep:
                    spans.append(r["ms_after_cells"])
                    walls.append(wall * 1e3)
            c1_res[th] = {"threads": th, "ms_median": round(med(spans), 3),
                          "mpts": round(len(c1) / med(spans) / 1e3, 4),
                          "ms_whole_call_median": round(med(walls), 3)}
        oracle.threads(cores)
        sample = pkg.synth(args.kind, args.cpu_sample, args.dim, args.seed)
        t1 = time.perf_counter()
        r = oracle.emst(sample, mode=1, R=1, cdim=cdim)
        ct = time.perf_counter() - t1
        assert len(r["u"]) == args.cpu_sample - 1
        n1 = max(args.cpu_sample // 5, 2)
        oracle.threads(1)
        t1 = time.perf_counter()
        r1 = oracle.emst(sample[:n1], mode=1, R=1, cdim=cdim)
        c1t = time.perf_counter() - t1
        assert len(r1["u"]) == n1 - 1
    finally:
        oracle.threads(cores)
        oracle.select(prev)
    return {"value": round(args.cpu_sample / ct / 1e6, 5), "unit": "Mpoints/s", "cores": cores, "kind": "port",
            "sample": f"{args.cpu_sample} points of the benchmarked distribution and seed, the oracle's "
                      f"restatement of the reference ring search (R=1), {cores} threads (per-point searches in "
                      f"parallel, rounds sequential), {ct:.1f} s whole call; one thread on the first {n1} points: "
                      f"{n1 / c1t / 1e6:.5f} Mpoints/s ({c1t:.1f} s)",
            "build": build,
            "reference_cpu_config": {"workload": "10^4 uniform 2D points, seed 0, R=1; median of 5 after 1 warm-up; "
                                                 "span after the cell build to the finished tree",
                                     "one_thread": c1_res[1], "all_cores": c1_res[cores]}}


def run_tsplib(args, pkg):
    """The TSPLIB sweep: each instance read with the library's TSPLIB reader,
    its points resident on the device, then W untimed and K timed EMSTs;
    per instance the median device time (emst_get_times total).  value = the
    sum of the medians over all instances."""
    import gzip
    import tempfile

    import numpy as np
    import torch

    tdir = os.path.join(ROOT, "tests", "golden", "tsplib")
    rows = []
    with tempfile.TemporaryDirectory() as tmp:
        for name in sorted(os.listdir(tdir)):
            path = os.path.join(tmp, name[:-3])
            with open(path, "wb") as f:
                f.write(gzip.open(os.path.join(tdir, name)).read())
            pts = pkg.read_points(path, 2)
            n = len(pts)
            dev = torch.from_numpy(np.ascontiguousarray(pts)).to("cuda:0")
            torch.cuda.synchronize()
            with pkg.Calculateur(dim=2) as c:
                ms = []
                for k in range(args.warmup + args.steps):
                    c.initialize_device(dev.data_ptr(), n, 2)
                    c.run()
                    if k >= args.warmup:
                        ms.append(c.times()[3])
                w, m = c.evaluate()
            assert m == n - 1, f"{name}: {m} edges for {n} cities"
            med = sorted(ms)[len(ms) // 2]
            rows.append({"instance": name[:-7], "n": n, "ms": round(med, 3), "mpts": round(n / med / 1e3, 3)})
    total = sum(r["ms"] for r in rows)
    big = max(rows, key=lambda r: r["n"])
    out = {
        "metric": f"TSPLIB sweep: EMST device time summed over {len(rows)} instances", "value": round(total, 3),
        "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "TSPLIB instances of the reference's 2D tests (tests/golden/tsplib)",
        "config": {"workload": "TSPLIB sweep, 2D", "instances": len(rows),
                   "cities": int(sum(r["n"] for r in rows)), "largest": big, "per_instance": rows},
    }
    print(json.dumps(out))


def run_batch(args, pkg):
    """Batched EMST: the 2D clouds and the 3D clouds each run as one batch
    context (emst_set_clouds), the two contexts concurrently from two host
    threads on their own streams.  value = all points / wall time of a step;
    per-cloud latency = build + cone search + the rounds up to the cloud's
    completion round, from the contexts' HIP events in a traced step."""
    import threading

    import numpy as np
    import torch

    B = args.batch
    sizes, dims, kinds, make = pkg.batch_workload(B, args.batch_seed)
    groups = []
    for d in (2, 3):
        idx = np.nonzero(dims == d)[0]
        clouds = [make(i) for i in idx]
        packed = torch.from_numpy(np.concatenate(clouds)).to("cuda:0")
        groups.append({"dim": d, "idx": idx, "sizes": sizes[idx], "dev": packed})
    torch.cuda.synchronize()
    total = int(sizes.sum())

    tuned = [{} for _ in groups]

    def make(trace):
        return [pkg.Calculateur(dim=g["dim"], **dict(t, trace=trace)) for g, t in zip(groups, tuned)]

    def step(ctxs):
        errs = []

        def work(ctx, g):
            try:
                ctx.initialize_clouds_device(g["dev"].data_ptr(), g["sizes"])
                ctx.run()
            except Exception as e:      # surfaced after join
                errs.append(e)

        ts = [threading.Thread(target=work, args=(c, g)) for c, g in zip(ctxs, groups)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        if errs:
            raise errs[0]

    ctxs = make(False)
    if not args.no_autotune:
        # untimed: leaf size and seed scale per dimension from the autotuner
        for i, (c, g) in enumerate(zip(ctxs, groups)):
            c.initialize_clouds_device(g["dev"].data_ptr(), g["sizes"])
            c.autotune()
            tuned[i] = {k: v for k, v in c.options().items()
                        if k in ("bucket_size", "seed_scale", "search_split", "round_search")}
    for _ in range(args.warmup):
        step(ctxs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ctxs)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    edges = 0
    for c, g in zip(ctxs, groups):
        w, e, r = c.cloud_results()
        assert np.array_equal(e, np.maximum(g["sizes"] - 1, 0)), "a cloud did not finish its tree"
        edges += int(e.sum())
        c.close()
    # latency from a traced step
    tctx = make(True)
    step(tctx)
    lat = np.zeros(B)
    for c, g in zip(tctx, groups):
        t = c.times()
        rows = c.trace()
        per_round = [r["time_find_next_closest"] + r["time_find_min_pair"] + r["time_connect_union"] +
                     r["time_flatten"] for r in rows]
        cum = np.concatenate([[0.0], np.cumsum(per_round)])
        _, _, last = c.cloud_results()
        lat[g["idx"]] = t[0] + t[1] + cum[np.minimum(last, len(per_round))]
        c.close()
    out = {
        "metric": f"batched EMST Mpoints/s ({B} clouds)", "value": round(total * args.steps / dt / 1e6, 3),
        "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (project generator, seeded)",
        "config": {"workload": f"{B} clouds, sizes log-uniform 1e2-1e5, half 2D half 3D, uniform or 8 Gaussians",
                   "clouds": B, "points": total, "edges": edges,
                   "clouds_2d": int((dims == 2).sum()), "points_2d": int(sizes[dims == 2].sum()),
                   "autotuned": {f"{g['dim']}d": t for g, t in zip(groups, tuned)},
                   "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                   "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                   "latency_ms_max": round(float(lat.max()), 3)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
