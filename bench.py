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


def profiled(kernel, tag, field):
    """A figure of the committed rocprofv3 summary of `kernel` on the workload
    `tag` (profiles/<kernel>_<tag>.md, a line "<field>: value"), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(field + ":"):
            return float(line.split(":")[1])
    return None


WORKLOAD_TAGS = {(1, 10_000_000, 3): "gauss64_1e7_3d", (2, 10_000_000, 2): "powerlaw16_1e7_2d",
                 (0, 1_000_000, 3): "uniform_1e6_3d"}
NAMES = {0: "uniform", 1: "gauss64+10%bg", 2: "powerlaw16"}
DESC = {"cone_plain": "cone search, plain seed", "cone_learned": "cone search, learned seed",
        "comp_plain": "per-round component search, plain seed",
        "comp_learned": "per-round component search, learned seed"}


def load_model(pkg_dir, tag, dim):
    """(model, training metadata) of models/radius_<tag><dim>d.npy, or (None, None)."""
    import numpy as np
    path = os.path.join(pkg_dir, "models", f"radius_{tag}{dim}d.npy")
    if not os.path.exists(path):
        return None, None
    meta = os.path.join(pkg_dir, "models", f"train_radius{'_comp' if tag else ''}.json")
    return np.load(path), (json.load(open(meta)) if os.path.exists(meta) else None)


def measure_cloud(pkg, ge, kind, n, dim, seed, steps, warmup, bucket=0, cone_dim=0, plain_only=False,
                  families=("cone", "comp")):
    """Every form on the same resident points: W untimed and K timed steps
    each (one step = set_points from device memory + run + weight), all
    forms asserted to build the same tree.  Returns the forms and the models."""
    import numpy as np
    import torch

    pts = pkg.synth(kind, n, dim, seed)
    dev = torch.from_numpy(pts).to("cuda:0")
    torch.cuda.synchronize()
    cdim = cone_dim or dim
    models = {}
    for fam, tag in (("cone", ""), ("comp", "comp_")):
        m, meta = (None, None) if plain_only or cdim != dim else load_model(ge.PKG_DIR, tag, dim)
        models[fam] = (m, meta)

    def timed(model, round_search):
        ctx = pkg.Calculateur(dim=dim, bucket_size=bucket, cone_dim=cone_dim, round_search=round_search)
        if model is not None:
            ctx.set_radius_model(model)
        search_ms, total_ms = [], []

        def step():
            ctx.initialize_device(dev.data_ptr(), n, dim)
            ctx.run()
            return ctx.evaluate()

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            w, m = step()
            t = ctx.times()
            search_ms.append(t[1])
            total_ms.append(t[3])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        u, v, _ = ctx.getLinks()
        samples = ctx.radius_samples() if (model is not None and not round_search) else None
        ctx.close()
        assert m == n - 1, f"tree has {m} edges"
        med = lambda xs: sorted(xs)[len(xs) // 2]
        return {"dt": dt, "w": w, "m": m, "search_ms": med(search_ms), "total_ms": med(total_ms),
                "edges_digest": hash((u.tobytes(), v.tobytes())), "samples": samples,
                "mpts": n * steps / dt / 1e6}

    forms = {}
    if "cone" in families:
        forms["cone_plain"] = timed(None, 0)
        if models["cone"][0] is not None:
            forms["cone_learned"] = timed(models["cone"][0], 0)
    if "comp" in families and cdim == dim:
        forms["comp_plain"] = timed(None, 1)
        if models["comp"][0] is not None:
            forms["comp_learned"] = timed(models["comp"][0], 1)
    ref = next(iter(forms.values()))
    for name, f in forms.items():
        assert f["edges_digest"] == ref["edges_digest"] and f["w"] == ref["w"], f"{name} built a different tree"
    return {"forms": forms, "models": models, "dev": dev, "pts": pts, "cdim": cdim}


def traced_work(pkg, dev, n, dim, model, round_search, bucket=0):
    """One traced run: per-kernel device times (HIP events on the context's
    stream) and the work counters."""
    with pkg.Calculateur(dim=dim, bucket_size=bucket, round_search=round_search, trace=True) as c:
        if model is not None:
            c.set_radius_model(model)
        c.initialize_device(dev.data_ptr(), n, dim)
        c.run()
        return c.times(), pkg.search_stats(c)


def learned_report(forms, fam):
    plain, learned = forms.get(f"{fam}_plain"), forms.get(f"{fam}_learned")
    if plain is None or learned is None:
        return None
    return {"plain_mpts": round(plain["mpts"], 3), "learned_mpts": round(learned["mpts"], 3),
            "learned_vs_plain": round(plain["dt"] / learned["dt"], 3)}


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
    ap.add_argument("--plain", action="store_true", help="plain seed only (skip the learned forms)")
    ap.add_argument("--forms", default="cone,comp", help="search families timed: cone, comp or both")
    ap.add_argument("--round-search", type=int, default=-1, choices=(-1, 0, 1),
                    help="headline form: the fastest of the cone search's forms (0), of the per-round "
                         "component search's (1), or of all (-1)")
    ap.add_argument("--cpu-sample", type=int, default=200_000,
                    help="points of the same distribution timed on the CPU oracle (0 = skip)")
    ap.add_argument("--extra", type=int, default=1,
                    help="1: also time the 10^7-point 2D power-law cloud (seed 4) and the 4096-cloud batch "
                         "(seed 5) into config")
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

    # torch (device buffers for the resident inputs) is loaded before the
    # library: both link the HIP runtime under one soname, and the copy the
    # process loads first then serves both
    import torch  # noqa: F401

    import __graft_entry__ as ge

    pkg = ge.load_package()
    pkg.lib()
    if args.batch:
        print(json.dumps(run_batch(args, pkg, args.batch, args.batch_seed, args.steps, args.warmup)))
        return
    if args.tsplib:
        run_tsplib(args, pkg)
        return

    fams = tuple(f for f in args.forms.split(",") if f)
    res = measure_cloud(pkg, ge, args.kind, args.n, args.dim, args.seed, args.steps, args.warmup,
                        bucket=args.bucket, cone_dim=args.cone_dim, plain_only=args.plain, families=fams)
    forms, models, cdim = res["forms"], res["models"], res["cdim"]
    K = 2 * cdim * (1 << (cdim - 1))
    # the headline form: the fastest measured (--round-search: the fastest of
    # that search's forms)
    if args.round_search < 0:
        head_name = max(forms, key=lambda k: forms[k]["mpts"])
    else:
        fam = "comp" if args.round_search == 1 else "cone"
        head_name = max((k for k in forms if k.startswith(fam)), key=lambda k: forms[k]["mpts"])
    fam = "comp" if head_name.startswith("comp") else "cone"
    head = forms[head_name]
    tag = WORKLOAD_TAGS.get((args.kind, args.n, args.dim)) if args.cone_dim in (0, args.dim) else None

    if fam == "cone":
        # the round-1 cone search; algorithmic bytes per point = 8*D (own
        # coordinates) + 4 + 8*K (one (id, d2) per cone out); time from HIP
        # events on the context stream
        kern = {"k_cone_search": {"alg_bytes": args.n * (8 * args.dim + 4 + 8 * K), "ms": head["search_ms"]}}
        work = None
    else:
        # the NNS kernels of the component search, from one traced run (HIP
        # events around each dispatch on the context stream):
        #   k_knn_first, once per EMST: 16 + 12 * KL + 13 algorithmic bytes
        #   per point (its point 16 in; KL list entries of 12 bytes, target 12
        #   and cursor 1 out), KL = 10 in 3D and 6 in 2D (the library's list
        #   length): 149 and 101;
        #   k_comp_search, two dispatches per round: 44 bytes per query and
        #   dispatch (point 16, label 4, previous key 8, list entry 4 in;
        #   result 12 out)
        t, st = traced_work(pkg, res["dev"], args.n, args.dim,
                            models["comp"][0] if head_name == "comp_learned" else None, 1, args.bucket)
        KL = 10 if args.dim == 3 else 6
        kern = {"k_knn_first": {"alg_bytes": (16 + 12 * KL + 13) * args.n, "ms": t[5]},
                "k_comp_search": {"alg_bytes": 44 * st["comp_queries"], "ms": t[4]}}
        work = {k: st[k] for k in ("comp_queries", "comp_nodes", "comp_leaves", "comp_pairs", "knn_nodes",
                                   "knn_leaves", "knn_pairs", "comp_walks", "knn_walks")}
    for name, kd in kern.items():
        kd["achieved_gbs"] = round(kd["alg_bytes"] / (kd["ms"] * 1e-3) / 1e9, 1) if kd["ms"] > 0 else 0.0
        kd["ms"] = round(kd["ms"], 3)
        tr = [profiled(name, tag, f) for f in ("traffic_bytes_per_launch", "valu_busy")] if tag else [None, None]
        kd["traffic"] = int(tr[0]) if tr[0] is not None else None
        kd["valu_busy"] = tr[1]
        # from the committed SQ counters of the same launch: VALU busy most of
        # the time means instruction issue; idle mostly, waits on dependent loads
        kd["limiter"] = None if tr[1] is None else ("VALU issue" if tr[1] >= 0.5 else "load latency")
    tot_bytes = sum(kd["alg_bytes"] for kd in kern.values())
    tot_ms = sum(kd["ms"] for kd in kern.values())
    achieved = tot_bytes / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else 0.0
    traffics = [kd["traffic"] for kd in kern.values()]
    timed_busy = [(kd["valu_busy"], kd["ms"]) for kd in kern.values() if kd["valu_busy"] is not None]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4),
                "traffic": sum(traffics) if all(x is not None for x in traffics) else None,
                "kernel": "+".join(kern), "kernel_ms": round(tot_ms, 3), "kernels": kern,
                # the HBM roofline is the one the survey prices these kernels
                # on; the counters put their limits elsewhere (per kernel)
                "limiter": "; ".join(f"{k}: {kd['limiter']}" for k, kd in kern.items() if kd.get("limiter"))
                           or None,
                "valu_busy": round(sum(b * m for b, m in timed_busy) / sum(m for _, m in timed_busy), 3)
                             if timed_busy and sum(m for _, m in timed_busy) > 0 else None,
                "work": work}

    extra = {}
    if args.extra and args.n == 10_000_000 and args.kind == 1 and args.dim == 3 and not args.cone_dim:
        # the 10^7-point 2D power-law cloud, every form
        r4 = measure_cloud(pkg, ge, 2, 10_000_000, 2, 4, args.steps, args.warmup, plain_only=args.plain)
        f4 = r4["forms"]
        h4 = max(f4, key=lambda k: f4[k]["mpts"])
        extra["powerlaw_2d_1e7"] = {
            "workload": "EMST 10000000 pts 2D powerlaw16 seed 4", "headline_form": DESC[h4],
            "mpts": round(f4[h4]["mpts"], 3), "ms_per_step": round(f4[h4]["dt"] * 1e3 / args.steps, 3),
            "mpts_by_form": {k: round(v["mpts"], 3) for k, v in f4.items()},
            "learned_vs_plain": {fam_: learned_report(f4, fam_) for fam_ in ("comp", "cone")},
            "same_tree_all_forms": True}
        del r4
        # the 4096-cloud batch (a failed check is reported here, and the line
        # keeps its other figures)
        try:
            extra["batch_4096"] = run_batch(args, pkg, 4096, 5, args.steps, args.warmup)
        except RuntimeError as ex:
            print(f"batch_4096: {ex}", file=sys.stderr)
            extra["batch_4096"] = {"error": str(ex)}

    cpu = cpu_baseline(args, pkg, cdim) if args.cpu_sample > 0 else None
    cone_meta = models["cone"][1] or {}
    comp_meta = models["comp"][1] or {}
    out = {
        "metric": "EMST Mpoints/s at 10^7 points (fastest exact form; learned-radius vs plain in config)",
        "value": round(head["mpts"], 3), "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(head["dt"] * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": f"EMST {args.n} pts {args.dim}D{' (2.5D: xy wedges)' if cdim < args.dim else ''} "
                               f"{NAMES.get(args.kind)} seed {args.seed}",
                   "n": args.n, "dim": args.dim, "form": DESC[head_name],
                   "edges": int(head["m"]), "weight": head["w"],
                   "gpu_total_ms_median": round(head["total_ms"], 3),
                   "mpts_by_form": {k: (round(forms[k]["mpts"], 3) if k in forms else None) for k in DESC},
                   "same_tree_all_forms": True,
                   "learned_vs_plain": {fam_: learned_report(forms, fam_) for fam_ in ("comp", "cone")},
                   "radius_model_training": {"cone": cone_meta.get("clouds"), "comp": comp_meta.get("clouds")},
                   "cone_dim": cdim, "round_search": 1 if fam == "comp" else 0,
                   **extra},
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


def run_batch(args, pkg, B, seed, steps, warmup):
    """Batched EMST: the 2D clouds and the 3D clouds each run as one batch
    context (emst_set_clouds), the two contexts concurrently from two host
    threads on their own streams, in two forms:
      autotuned   leaf size, seed and search form per dimension from
                  emst_autotune, the build issued kernel by kernel and the
                  rounds replayed 4 at a time from a captured graph;
      one_launch  the per-round component search with options.one_launch
                  (and the autotuned leaf size and seed where the autotuner
                  chose that search): each context's whole run is one graph
                  launch.
    Throughput = all points / wall time of a step; value = the faster form's.
    Per-cloud latency = time from the run's start to the end of the cloud's
    completion round: for one_launch from the events inside the launch
    (emst_get_round_ends), for autotuned from a traced step.  Returns the
    result as a dict (the JSON line of --batch)."""
    import threading

    import numpy as np
    import torch

    sizes, dims, kinds, make = pkg.batch_workload(B, seed)
    groups = []
    for d in (2, 3):
        idx = np.nonzero(dims == d)[0]
        clouds = [make(i) for i in idx]
        packed = torch.from_numpy(np.concatenate(clouds)).to("cuda:0")
        groups.append({"dim": d, "idx": idx, "sizes": sizes[idx], "dev": packed})
    torch.cuda.synchronize()
    total = int(sizes.sum())

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

    def timed(opts):
        ctxs = [pkg.Calculateur(dim=g["dim"], **o) for g, o in zip(groups, opts)]
        for _ in range(warmup):
            step(ctxs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(ctxs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        edges = 0
        lat = np.zeros(B)
        for c, g, o in zip(ctxs, groups, opts):
            w, e, last = c.cloud_results()
            bad = np.nonzero(e != np.maximum(g["sizes"] - 1, 0))[0]
            if len(bad):
                raise RuntimeError(
                    f"{len(bad)} of {len(e)} {g['dim']}D clouds did not finish their trees (options {o}): " +
                    ", ".join(f"cloud {int(g['idx'][i])} n={int(g['sizes'][i])} edges={int(e[i])} "
                              f"last round={int(last[i])}" for i in bad[:5]))
            edges += int(e.sum())
            ends = c.round_ends()
            if len(ends):
                t = c.times()
                lat[g["idx"]] = np.where(last > 0, ends[np.maximum(last, 1) - 1], t[0])
            c.close()
        return dt, edges, lat

    # autotuned (untimed tuning)
    tuned = [{} for _ in groups]
    if not args.no_autotune:
        for i, g in enumerate(groups):
            with pkg.Calculateur(dim=g["dim"]) as c:
                c.initialize_clouds_device(g["dev"].data_ptr(), g["sizes"])
                c.autotune()
                tuned[i] = {k: v for k, v in c.options().items()
                            if k in ("bucket_size", "seed_scale", "search_split", "round_search")}
    dt_a, edges, _ = timed(tuned)
    # its latency from a traced step
    tctx = [pkg.Calculateur(dim=g["dim"], **dict(t, trace=True)) for g, t in zip(groups, tuned)]
    step(tctx)
    lat_a = np.zeros(B)
    for c, g in zip(tctx, groups):
        t = c.times()
        rows = c.trace()
        per_round = [r["time_find_next_closest"] + r["time_find_min_pair"] + r["time_connect_union"] +
                     r["time_flatten"] for r in rows]
        cum = np.concatenate([[0.0], np.cumsum(per_round)])
        _, _, last = c.cloud_results()
        lat_a[g["idx"]] = t[0] + t[1] + cum[np.minimum(last, len(per_round))]
        c.close()
    # one launch per context
    dt_o, edges_o, lat_o = timed([dict(t, round_search=1, one_launch=1) if t.get("round_search", 1) == 1
                                  else {"round_search": 1, "one_launch": 1} for t in tuned])
    assert edges_o == edges
    forms = {"autotuned": (dt_a, lat_a), "one_launch": (dt_o, lat_o)}
    best = min(forms, key=lambda k: forms[k][0])
    dt, lat = forms[best]
    rep = lambda d, l: {"mpts": round(total * steps / d / 1e6, 3), "ms_per_step": round(d * 1e3 / steps, 3),
                        "latency_ms_p50": round(float(np.percentile(l, 50)), 3),
                        "latency_ms_p99": round(float(np.percentile(l, 99)), 3),
                        "latency_ms_max": round(float(l.max()), 3)}
    return {
        "metric": f"batched EMST Mpoints/s ({B} clouds)", "value": round(total * steps / dt / 1e6, 3),
        "unit": "Mpoints/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(dt * 1e3 / steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (project generator, seeded)",
        "config": {"workload": f"{B} clouds, sizes log-uniform 1e2-1e5, half 2D half 3D, uniform or 8 Gaussians, "
                               f"seed {seed}",
                   "clouds": B, "points": total, "edges": edges,
                   "clouds_2d": int((dims == 2).sum()), "points_2d": int(sizes[dims == 2].sum()),
                   "form": best, "autotuned": {f"{g['dim']}d": t for g, t in zip(groups, tuned)},
                   "by_form": {k: rep(*v) for k, v in forms.items()},
                   "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                   "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                   "latency_ms_max": round(float(lat.max()), 3)},
    }


if __name__ == "__main__":
    main()
