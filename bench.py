"""Benchmark: k-NN graph build throughput and recall on one MI355X.

A step is one calculate(): normalize, forest, initial leaf k-NN and the
propagation rounds over the configuration's synthetic points, which are
uploaded to HBM before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config C5]

C5 (the default, 4M x 128, k=20, EUCLIDEAN) compares three forms at the
shape the plain form needs for recall@20 >= 0.95:
  plain    the reference's sign-split forest, exact FP32 everywhere;
  median   median-split trees on seeded directions, FP16 / principal-subspace
           screens with exact FP32 re-scoring, the mean summed in FP64 over
           point blocks instead of one sequential FP32 chain (no router);
  learned  the same with each tree's first levels routed by its own trained
           router (the chosen form: `value` is its throughput).
Routers are trained before the timed region on a graph the method builds for
a separate 200k-point sample; their inference is inside the bracket.
C1-C4 compare plain with the screened form, which stores the same graph.
"""
import argparse
import importlib.util
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "approximate-cuda-knn_amd")

METRIC = "kNN-graph build s + recall@k vs FP64 exact; plain vs chosen form, 1 GPU"
# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3        # FP32 (vector and matrix, dense)
PEAK_FP16_MFMA_TFLOPS = 2500.0  # FP16 / BF16 matrix cores, dense


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(datasets, config, shape, metric_name, k, sample_n):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    x = datasets.generate(config, n=sample_n)
    m = oracle_lib.METRICS[metric_name]
    t0 = time.perf_counter()
    oracle_lib.run(x, k, shape["trees"], shape["leaf"], shape["props"], m)
    dt = time.perf_counter() - t0
    return {
        "value": sample_n / dt,
        "unit": "points/s",
        "cores": int(oracle_lib.lib().orc_num_threads()),
        "kind": "port",
        "sample": (f"{config} distribution at N={sample_n}, the plain form's shape (k={k}, trees={shape['trees']}, "
                   f"leaf={shape['leaf']}, propagations={shape['props']}, {metric_name}): whole pipeline in "
                   f"{dt:.1f} s on the C oracle (OpenMP over points)"),
    }


# C1-C4: both forms store the same graph bit for bit (checked on every run);
# the chosen form only computes less.
FORMS = {
    "plain": {"sketch": 0, "leaf_mfma": 0},
    "chosen": {"sketch": 1, "leaf_mfma": 1},
}

# C5: the plain form's shape is the cheapest sign-forest shape measured to
# reach recall@20 >= 0.95 on this data (T=6, L=256, P=2 gives 0.967; T=4
# gives 0.940); the median and learned forms keep that shape.
C5_SHAPE = {"trees": 6, "leaf": 256, "props": 2}
C5_FORMS = {
    "plain": {"opts": {"sketch": 0, "leaf_mfma": 0, "split": 0, "pca_screen": 0, "router_trees": 0}},
    "median": {"opts": {"sketch": 1, "leaf_mfma": 1, "split": 1, "pca_screen": 1, "router_trees": 0, "mean_fp64": 1}},
    "learned": {"opts": {"sketch": 1, "leaf_mfma": 1, "split": 1, "pca_screen": 1, "router_trees": 6, "mean_fp64": 1},
                "routers": 6},
}
ROUTER = {"n": 200_000, "sample_seed": 105, "bin_bits": 8, "hidden": 256, "epochs": 10, "batch": 1024, "lr": 2e-3,
          "sweeps": 20, "trees": 8, "leaf": 256, "props": 3}


def train_routers(aknn, datasets, config, cfg, count):
    """The learned form's routers, trained on a graph the method builds for a
    separate sample of the same distribution (outside the timed region).
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
               "wall_s": time.perf_counter() - t0, "outside_timed_region": True,
               "accuracy": [round(s["accuracy"], 4) for s in stats],
               "cut_before": [round(s["cut_before"], 4) for s in stats],
               "cut_after": [round(s["cut_after"], 4) for s in stats],
               "train_ms": [round(s["train_ms"], 1) for s in stats]}
    print(f"[bench] routers trained: {summary}", file=sys.stderr, flush=True)
    return routers, summary


def _entry(ms, units, bound, peak, unit, scale, **extra):
    achieved = units / (ms / 1e3) / scale if ms > 0 else 0.0
    e = {"ms": ms, "units": units, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
         "frac": achieved / peak if peak else None}
    e.update(extra)
    return e


def stage_rooflines(res, d, n, router_info, options):
    """Per stage, the algorithmic units of its kernels over their measured time
    against the engine that bounds them: forest and propagation bytes against
    HBM, exact leaf chains against FP32, screened leaf launches (GEMM-form
    flops of the FP16 products) and router inference (BF16) against the
    matrix cores."""
    k = res["kernels"]
    st = res["stages_ms"]
    out = {"normalize": {"ms": st["normalize"], "bound": "hbm" if options.get("mean_fp64") else "latency",
                         "note": ("mean as FP64 sums over 2048-point blocks" if options.get("mean_fp64") else
                                  "sequential per-dimension FP32 mean (one dependent add per row)")}}
    router_ms = 0.0
    if router_info:
        router_ms = sum(router_info["assign_ms"])
        dp = (d + 15) // 16 * 16
        h, bins = router_info["hidden"], router_info["bins"]
        flops = len(router_info["assign_ms"]) * n * 2.0 * (dp * h + h * h + h * bins)
        out["router_inference"] = _entry(router_ms, flops, "mfma", PEAK_FP16_MFMA_TFLOPS, "TFLOP/s", 1e12,
                                         kernel="router_fused_kernel", engine="BF16 MFMA")
    f = k["forest"]
    out["forest"] = _entry(st["forest"] - router_ms, f["units"], "hbm", PEAK_HBM_GBS, "GB/s", 1e9,
                           levels=f["launches"])
    le = k["leaf_exact"]
    if le["launches"]:
        out["leaf_exact"] = _entry(le["total_ms"], le["units"], "fp32", PEAK_FP32_TFLOPS, "TFLOP/s", 1e12,
                                   kernel="leaf_knn_kernel", launches=le["launches"], engine="FP32 VALU")
    lm = k["leaf_mfma"]
    if lm["launches"]:
        out["leaf_mfma"] = _entry(lm["total_ms"], lm["units"], "mfma", PEAK_FP16_MFMA_TFLOPS, "TFLOP/s", 1e12,
                                  kernel="leaf_mfma_kernel", launches=lm["launches"],
                                  engine="FP16 MFMA screen (GEMM-form flops) + exact FP32 re-score",
                                  pairs_screened=lm.get("pairs_screened"), pairs_rescored=lm.get("pairs_rescored"))
    p = k["propagate"]
    if p["launches"]:
        out["propagation"] = _entry(p["total_ms"], p["units"], "hbm", PEAK_HBM_GBS, "GB/s", 1e9,
                                    kernel="propagate_kernel", launches=p["launches"],
                                    candidates_examined=p.get("candidates_examined"),
                                    candidates_rescored=p.get("candidates_rescored"))
    return out


def roofline(stages, config, form):
    """The top-level roofline: the single kernel with the largest summed launch
    time, its algorithmic units per launch over its mean launch duration (HIP
    events on the launch stream)."""
    # a stage without algorithmic units (the COSINE parity initial pass, whose
    # chunk loop is bound by L2 latency, not by a peak) is not a roofline
    cands = [s for s in ("leaf_exact", "leaf_mfma", "propagation") if s in stages and stages[s]["units"] > 0]
    dom = max(cands, key=lambda s: stages[s]["ms"])
    s = stages[dom]
    bound = "hbm" if dom == "propagation" else "mfma"
    roof = {"bound": bound, "achieved": s["achieved"], "peak": s["peak"], "unit": s["unit"], "frac": s["frac"],
            "traffic": None, "kernel": s["kernel"], "launches": s["launches"],
            "avg_launch_ms": s["ms"] / s["launches"], "units_per_launch": s["units"] / s["launches"]}
    if dom == "leaf_exact":
        roof["engine"] = "FP32 VALU exact chains; peak = MI355X FP32 dense"
    # HBM bytes per launch of this kernel on this workload and form, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against
    # a copy kernel)
    prof = os.path.join(ROOT, "profiles", f"traffic_{roof['kernel']}_{config}_{form}.md")
    if os.path.exists(prof):
        with open(prof) as fh:
            m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", fh.read(), re.M)
        if m:
            roof["traffic"] = float(m.group(1))
    return roof


def run_form(g, name, args, n, sync):
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
    stats = {nm: g.kernel_stats(nm) for nm in ("leaf_knn", "leaf_exact", "leaf_mfma", "propagate", "forest")}
    ex, rs = g.candidate_counts()
    stats["propagate"]["candidates_examined"] = ex
    stats["propagate"]["candidates_rescored"] = rs
    ex, rs = g.candidate_counts("forest")
    stats["forest"]["side_tests"] = ex
    stats["forest"]["side_tests_exact"] = rs
    scr, res = g.candidate_counts("leaf_knn")
    stats["leaf_mfma"]["pairs_screened"] = scr
    stats["leaf_mfma"]["pairs_rescored"] = res
    return {
        "leaf_launch": g.leaf_form(),
        "value": n / wall,
        "ms_per_step": wall * 1e3,
        "build_s": float(np.mean([st["total"] for st in stage])) / 1e3,
        "stages_ms": {k: float(np.mean([st[k] for st in stage])) for k in stage[0]},
        "kernels": stats,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C5")
    ap.add_argument("--form", default="all", help="C5: all, plain, median or learned; others: all, plain or chosen")
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
    c5 = args.config == "C5"
    all_forms = list(C5_FORMS) if c5 else list(FORMS)
    names = all_forms if args.form == "all" else [args.form]
    if any(nm not in all_forms for nm in names):
        ap.error(f"--form must be all or one of {all_forms} for {args.config}")
    head = "learned" if c5 and "learned" in names else ("chosen" if "chosen" in names else names[-1])
    print(f"[bench] generating {args.config} n={n} d={cfg['d']}", file=sys.stderr, flush=True)
    x = datasets.generate(args.config, n=n)

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        sync = lambda: None  # noqa: E731

    results, graphs, recalls, shapes = {}, {}, {}, {}
    rng = np.random.default_rng(1234)
    q = np.sort(rng.choice(n, size=min(args.queries, n), replace=False))
    router_stats = None
    if c5:
        routers = []
        if "learned" in names:
            routers, router_stats = train_routers(aknn, datasets, args.config, cfg, C5_FORMS["learned"]["routers"])
        for name in names:
            spec = C5_FORMS[name]
            shapes[name] = dict(C5_SHAPE)
            g = aknn.Graph()
            g.initialize(x, cfg["k"], C5_SHAPE["trees"], C5_SHAPE["leaf"], C5_SHAPE["props"], metric,
                         **spec["opts"])
            for r in routers if spec.get("routers") else []:
                g.add_router(r)
            results[name] = run_form(g, name, args, n, sync)
            results[name]["options"] = spec["opts"]
            if spec.get("routers"):
                router_stats["assign_ms"] = [round(r.stats()["assign_ms"], 3) for r in routers]
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
            for kk, v in FORMS[name].items():
                g.set_option(kk, v)
            results[name] = run_form(g, name, args, n, sync)
            results[name]["options"] = FORMS[name]
            graphs[name] = g.graph()
            shapes[name] = {kk: cfg[kk] for kk in ("trees", "leaf", "props")}
        if not args.no_recall:
            # the graph is the last form's, the same for both forms
            t1 = time.perf_counter()
            rec = g.recall(q)
            for name in names:
                recalls[name] = rec
            print(f"[bench] recall {rec:.4f} ({time.perf_counter() - t1:.2f} s)", file=sys.stderr, flush=True)
        g.close()
    cfg_name = args.config if args.n is None else f"{args.config}_n{n}"
    for name in names:
        ri = router_stats if (c5 and C5_FORMS[name].get("routers")) else None
        results[name]["stages"] = stage_rooflines(results[name], cfg["d"], n, ri, results[name]["options"])
        results[name]["roofline"] = roofline(results[name]["stages"], cfg_name, name)
        results[name]["config"] = shapes[name]
        if name in recalls:
            results[name]["recall_at_k"] = recalls[name]
    r = results[head]
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
        "stages": r["stages"],
        "roofline": r["roofline"],
        "leaf_launch": r["leaf_launch"],
    }
    if head in recalls:
        line["recall_at_k"] = recalls[head]
        line["recall_queries"] = int(len(q))
        line["recall_reference"] = "FP64 exact k-NN (aknn_exact_knn), self excluded, ties at the k-th distance count"
    line["forms"] = {nm: {kk: results[nm][kk] for kk in ("value", "ms_per_step", "build_s", "stages_ms", "stages",
                                                           "roofline", "config", "options", "kernels", "leaf_launch")
                          if kk in results[nm]} for nm in names}
    for nm in names:
        if nm in recalls:
            line["forms"][nm]["recall_at_k"] = recalls[nm]
    if "plain" in results and head != "plain":
        line["chosen_vs_plain"] = r["value"] / results["plain"]["value"]
        print(f"[bench] {head}/plain {line['chosen_vs_plain']:.3f}", file=sys.stderr, flush=True)
    if c5 and "median" in results and head == "learned":
        line["learned_vs_median"] = r["value"] / results["median"]["value"]
    if graphs and len(graphs) > 1:
        ga, gb = graphs["plain"], graphs["chosen"]
        line["forms_identical"] = bool((ga[0] == gb[0]).all() and (ga[1].view(np.uint32) == gb[1].view(np.uint32)).all())
        print(f"[bench] identical graphs: {line['forms_identical']}", file=sys.stderr, flush=True)
    if router_stats is not None:
        line["router"] = router_stats
    if not args.no_cpu:
        print(f"[bench] CPU baseline on {args.cpu_sample} points", file=sys.stderr, flush=True)
        shape = shapes.get("plain", shapes[head])
        line["cpu_baseline"] = cpu_baseline(datasets, args.config, shape, cfg["metric"], cfg["k"],
                                            min(args.cpu_sample, n))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
