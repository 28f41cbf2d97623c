"""Benchmark: k-NN graph build throughput and recall on one MI355X.

A step is one calculate(): normalize, RP forest, initial leaf k-NN and the
propagation rounds over the configuration's synthetic points, which are
uploaded to HBM before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config C2]
"""
import argparse
import importlib.util
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "approximate-cuda-knn_amd")

METRIC = "kNN-graph build s + recall@k vs FP64 exact; plain vs chosen form, 1 GPU"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(datasets, config, cfg, sample_n):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    x = datasets.generate(config, n=sample_n)
    m = oracle_lib.METRICS[cfg["metric"]]
    t0 = time.perf_counter()
    oracle_lib.run(x, cfg["k"], cfg["trees"], cfg["leaf"], cfg["props"], m)
    dt = time.perf_counter() - t0
    return {
        "value": sample_n / dt,
        "unit": "points/s",
        "cores": int(oracle_lib.lib().orc_num_threads()),
        "kind": "port",
        "sample": (f"{config} distribution at N={sample_n} (same D, k, trees, leaf, propagations, metric): whole "
                   f"pipeline in {dt:.1f} s on the C oracle (OpenMP over points)"),
    }


# The two forms the metric compares.  On C1-C4 both give the same graph bit
# for bit (checked on every run below); the chosen form only computes less.
FORMS = {
    "plain": {"sketch": 0},
    "chosen": {"sketch": 1, "leaf_mfma": 1},
}

# C5: the plain form is the reference's forest with exact FP32 everywhere;
# the chosen form is the learned one (router bins as the first levels of
# routed median-split trees, FP16 sketch and matrix-core screens), with its
# own trees and rounds.  The graphs differ, so each form's recall is measured.
# Plain: the cheapest reference-forest shape measured to reach recall@20 >=
# 0.95 on this data (T=6, L=256, P=2 gives 0.967; T=4 gives 0.940).  Chosen:
# the same trees and rounds, each tree routed by its own router.
C5_FORMS = {
    "plain": {"opts": {"sketch": 0, "leaf_mfma": 0, "split": 0}, "trees": 6, "leaf": 256, "props": 2},
    "chosen": {"opts": {"sketch": 1, "leaf_mfma": 1, "split": 1, "router_trees": 6, "pca_screen": 1}, "trees": 6,
               "leaf": 256, "props": 2, "routers": 6},
}
ROUTER = {"n": 200_000, "sample_seed": 105, "bin_bits": 8, "hidden": 256, "epochs": 10, "batch": 1024, "lr": 2e-3,
          "sweeps": 20, "trees": 8, "leaf": 256, "props": 3}


def train_routers(aknn, datasets, config, cfg, count):
    """The routers of the learned form, trained on a graph the method builds
    for a separate sample of the same distribution (outside the timed region).
    Router s starts its labels from tree s's leaf order, so each routed tree
    gets a different learned top."""
    rc = ROUTER
    xt = datasets.generate(config, n=rc["n"], sample_seed=rc["sample_seed"])
    t0 = time.perf_counter()
    gt = aknn.Graph()
    gt.initialize(xt, cfg["k"], rc["trees"], rc["leaf"], rc["props"], aknn.METRICS[cfg["metric"]], split=1)
    gt.calculate()
    routers, stats = [], []
    for seed in range(1, count + 1):
        r = aknn.Router(xt.shape[1], rc["hidden"], rc["bin_bits"], seed=seed)
        labels, cut = r.labels(gt, sweeps=rc["sweeps"])
        st = r.fit(gt, labels, epochs=rc["epochs"], batch=rc["batch"], lr=rc["lr"])
        st.update({"cut_before": cut[0], "cut_after": cut[1]})
        routers.append(r)
        stats.append(st)
    gt.close()
    summary = {"routers": count, "train_points": rc["n"], "sample_seed": rc["sample_seed"],
               "bins": 1 << rc["bin_bits"], "hidden": rc["hidden"], "epochs": rc["epochs"],
               "wall_s": time.perf_counter() - t0,
               "accuracy": [round(s["accuracy"], 4) for s in stats],
               "cut_before": [round(s["cut_before"], 4) for s in stats],
               "cut_after": [round(s["cut_after"], 4) for s in stats],
               "train_ms": [round(s["train_ms"], 1) for s in stats]}
    print(f"[bench] routers trained: {summary}", file=sys.stderr, flush=True)
    return routers, summary


def roofline(stats, config, form):
    """Roofline of the step's dominant kernel from its live launch statistics."""
    dom = max(("leaf_knn", "propagate"), key=lambda nm: stats[nm]["total_ms"])
    s = stats[dom]
    avg_s = s["total_ms"] / max(s["launches"], 1) / 1e3
    per_launch = s["units"] / max(s["launches"], 1)
    if dom == "leaf_knn":
        achieved = per_launch / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "leaf_knn_kernel",
                "engine": "FP32 VALU (exact chains, no MFMA)"}
    else:
        achieved = per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": "propagate_kernel"}
    # HBM bytes per launch of this kernel on this workload and form, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against
    # copy kernels)
    prof = os.path.join(ROOT, "profiles", f"traffic_{roof['kernel']}_{config}_{form}.md")
    if os.path.exists(prof):
        with open(prof) as f:
            m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        if m:
            roof["traffic"] = float(m.group(1))
    return roof


def run_form(g, name, opts, args, n, sync):
    for k, v in opts.items():
        g.set_option(k, v)
    for w in range(args.warmup):
        g.calculate()
        print(f"[bench] {name} warmup {w}: {g.stage_times()['total']:.1f} ms", file=sys.stderr, flush=True)
    sync()
    t0 = time.perf_counter()
    stage = []
    for s in range(args.steps):
        g.calculate()
        stage.append(g.stage_times())
    sync()
    wall = (time.perf_counter() - t0) / args.steps
    for s, st in enumerate(stage):
        print(f"[bench] {name} step {s}: " + " ".join(f"{k}={v:.1f}" for k, v in st.items()), file=sys.stderr,
              flush=True)
    stats = {nm: g.kernel_stats(nm) for nm in ("leaf_knn", "propagate", "forest")}
    ex, rs = g.candidate_counts()
    stats["propagate"]["candidates_examined"] = ex
    stats["propagate"]["candidates_rescored"] = rs
    if opts.get("sketch"):
        ex, rs = g.candidate_counts("forest")
        stats["forest"]["side_tests"] = ex
        stats["forest"]["side_tests_exact"] = rs
    if opts.get("leaf_mfma"):
        stats["leaf_mfma"] = g.kernel_stats("leaf_mfma")
        scr, res = g.candidate_counts("leaf_knn")
        stats["leaf_mfma"]["pairs_screened"] = scr
        stats["leaf_mfma"]["pairs_rescored"] = res
    return {
        "value": n / wall,
        "ms_per_step": wall * 1e3,
        "build_s": float(np.mean([st["total"] for st in stage])) / 1e3,
        "stages_ms": {k: float(np.mean([st[k] for st in stage])) for k in stage[0]},
        "kernels": stats,
        "options": opts,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--form", choices=("both", "plain", "chosen"), default="both")
    ap.add_argument("--n", type=int, default=None, help="override the number of points")
    ap.add_argument("--queries", type=int, default=10000)
    ap.add_argument("--cpu-sample", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU sharding is not built: one GPU per box)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("LOCAL_RANK", "0")) != 0:
        return

    aknn = _load("aknn_binding", os.path.join(PKG_DIR, "aknn.py"))
    datasets = _load("aknn_datasets", os.path.join(PKG_DIR, "datasets.py"))
    cfg = dict(datasets.CONFIGS[args.config])
    n = args.n or cfg["n"]
    metric = aknn.METRICS[cfg["metric"]]
    print(f"[bench] generating {args.config} n={n} d={cfg['d']}", file=sys.stderr, flush=True)
    x = datasets.generate(args.config, n=n)

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        torch = None
        sync = lambda: None  # noqa: E731

    names = ["plain", "chosen"] if args.form == "both" else [args.form]
    results, graphs, recalls, shapes = {}, {}, {}, {}
    rng = np.random.default_rng(1234)
    q = np.sort(rng.choice(n, size=min(args.queries, n), replace=False))
    routers, router_stats = [], None
    if args.config == "C5":
        if "chosen" in names:
            routers, router_stats = train_routers(aknn, datasets, args.config, cfg, C5_FORMS["chosen"]["routers"])
        g = None
        for name in names:
            spec = C5_FORMS[name]
            shape = {kk: spec.get(kk, cfg[kk]) for kk in ("trees", "leaf", "props")}
            g = aknn.Graph()
            g.initialize(x, cfg["k"], shape["trees"], shape["leaf"], shape["props"], metric, **spec["opts"])
            for r in routers if spec.get("routers") else []:
                g.add_router(r)
            results[name] = run_form(g, name, spec["opts"], args, n, sync)
            if spec.get("routers"):
                results[name]["router_assign_ms"] = [round(r.stats()["assign_ms"], 3) for r in routers]
            shapes[name] = shape
            if not args.no_recall:
                t1 = time.perf_counter()
                recalls[name] = g.recall(q)
                print(f"[bench] {name} recall {recalls[name]:.4f} ({time.perf_counter() - t1:.2f} s)",
                      file=sys.stderr, flush=True)
            g.set_router(None)
            g.close()
    else:
        g = aknn.Graph()
        g.initialize(x, cfg["k"], cfg["trees"], cfg["leaf"], cfg["props"], metric)
        for name in names:
            results[name] = run_form(g, name, FORMS[name], args, n, sync)
            graphs[name] = g.graph()
            shapes[name] = {kk: cfg[kk] for kk in ("trees", "leaf", "props")}
    head = names[-1]
    r = results[head]
    cfg_name = args.config if args.n is None else f"{args.config}_n{n}"
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.config, "n": n, "d": cfg["d"], "k": cfg["k"], "trees": shapes[head]["trees"],
                   "leaf": shapes[head]["leaf"], "propagations": shapes[head]["props"], "distance": cfg["metric"],
                   "form": head},
        "build_s": r["build_s"],
        "stages_ms": r["stages_ms"],
        "kernels": r["kernels"],
        "roofline": roofline(r["kernels"], cfg_name, head),
    }
    if len(names) > 1:
        other = results["plain"]
        other["roofline"] = roofline(other["kernels"], cfg_name, "plain")
        other["config"] = shapes["plain"]
        line["plain"] = other
        line["chosen_vs_plain"] = r["value"] / other["value"]
        if graphs:
            ga, gb = graphs["plain"], graphs["chosen"]
            line["forms_identical"] = bool((ga[0] == gb[0]).all() and
                                           (ga[1].view(np.uint32) == gb[1].view(np.uint32)).all())
        print(f"[bench] chosen/plain {line['chosen_vs_plain']:.3f}, identical graphs: {line.get('forms_identical')}",
              file=sys.stderr, flush=True)
    if router_stats is not None:
        line["router"] = dict(router_stats)
        if "router_assign_ms" in r:
            line["router"]["assign_ms"] = r["router_assign_ms"]
    if recalls:
        line["recall_at_k"] = recalls[head]
        if "plain" in recalls:
            line["plain"]["recall_at_k"] = recalls["plain"]
        line["recall_queries"] = int(len(q))
        line["recall_reference"] = "FP64 exact k-NN (aknn_exact_knn), self excluded, ties at the k-th distance count"
    elif not args.no_recall:
        # FP64 exact k-NN of seeded queries on the GPU (the library's own
        # kernel); the graph is the last form's, identical for both forms
        t1 = time.perf_counter()
        line["recall_at_k"] = g.recall(q)
        line["recall_queries"] = int(len(q))
        line["recall_reference"] = "FP64 exact k-NN (aknn_exact_knn), self excluded, ties at the k-th distance count"
        print(f"[bench] recall {line['recall_at_k']:.4f} ({time.perf_counter() - t1:.2f} s)", file=sys.stderr,
              flush=True)
        g.close()
    if not args.no_cpu:
        print(f"[bench] CPU baseline on {args.cpu_sample} points", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(datasets, args.config, cfg, min(args.cpu_sample, n))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
