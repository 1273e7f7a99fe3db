"""Benchmark: boosting iterations per second of the MI355X learner on the
10.5M x 28 binary workload (255 leaves, max_bin 255, quantized int8 gradients
with 4 bins, stochastic rounding, leaf renewal), next to the plain f64 GPU
learner and the CPU oracle timed on this host.

One step = one boosting iteration with every input resident in HBM: binary
log-loss gradients on the device, one tree grown by the learner, the score
update. By default the timed steps are the workload's whole job, its 500
iterations from the first tree. Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bankai-lightgbm_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ROUND_PLANNER = 0
HIST_TILING = None  # (min_rows, max_blocks) forced, or None for the autotuner's choice
HIST_OPTS = {}      # bin width, partition-built histograms: forced values, else the learner's defaults
TREE_LOG = None     # a list collecting every tree's row counts (--tree-log), for PMC summaries


def run_gpu(ds, y, params, steps, warmup, lr, profile, every_trees=1):
    """profile: False, True (every kernel kind) or a list of kind names to time,
    on the launches of every `every_trees`-th tree."""
    import numpy as np
    import torch

    import bankai_gpu as bg
    import boosting as bst

    obj = bst.BinaryObjective(y)
    init = obj.boost_from_score(0)
    learner = bg.Learner(ds, params, False)
    learner.set_round_planner(ROUND_PLANNER)
    if HIST_TILING:
        learner.set_hist_tiling(*HIST_TILING)
    if "bin_width" in HIST_OPTS:
        learner.set_hist_bin_width(HIST_OPTS["bin_width"])
    if "partition_histograms" in HIST_OPTS:
        learner.set_partition_histograms(HIST_OPTS["partition_histograms"])
    n = ds.num_data
    dev = torch.device("cuda:0")
    score = torch.full((n,), init, dtype=torch.float64, device=dev)
    label = torch.as_tensor(y.astype(np.float32), device=dev)
    g = torch.empty(n, dtype=torch.float32, device=dev)
    h = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()  # the learner's stream reads these tensors
    trees = []

    def gradients():
        learner.binary_gradients_device(score.data_ptr(), label.data_ptr(), obj.sigmoid, g.data_ptr(), h.data_ptr(),
                                        label_weights=obj.label_weights)

    def step(record):
        # one boosting iteration: the tree, the score update (Tree::apply_shrinkage's
        # values) and the gradients of the next iteration are queued before the
        # host-side bookkeeping of the tree, which then overlaps the device work
        t = learner.train_device(g.data_ptr(), h.data_ptr())
        if TREE_LOG is not None:
            TREE_LOG.append(learner.last_stats())
        learner.add_prediction_to_score_device(bst._round0(t["leaf_value"] * lr), score.data_ptr())
        gradients()
        if record:
            tree = bst.Tree(t)
            tree.apply_shrinkage(lr)
            trees.append(tree)

    gradients()
    for i in range(warmup):
        step(True)
    torch.cuda.synchronize()
    if profile:
        learner.profile(profile, every_trees)
    stats = {"hist_rows": 0, "part_rows": 0, "splits": 0, "rounds": 0, "speculative_waste": 0,
             "committed_hist_rows": 0, "committed_part_rows": 0, "host_ms": 0.0, "wait_ms": 0.0}
    phases = {"replay": 0.0, "plan": 0.0, "launch": 0.0, "readback": 0.0}
    tiling = None
    t0 = time.perf_counter()
    for i in range(steps):
        step(True)
        s = learner.last_stats()
        for k in stats:
            stats[k] += s[k]
        for k in phases:
            phases[k] += s["host_phases_ms"][k]
        tiling = s["hist_tiling"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kstats = learner.kernel_stats() if profile else None
    out_trees = []
    for k, tr in enumerate(trees):
        d = dict(tr.d)
        if k == 0 and abs(init) > bst.K_EPSILON:
            tr2 = bst.Tree(d)
            tr2.add_bias(init)
            d = tr2.d
        out_trees.append(d)
    learner.close()
    stats["hist_tiling"] = tiling
    stats["host_phases_ms"] = phases
    return elapsed, stats, kstats, out_trees


def run_multiclass(rows, steps, warmup, lr, runs, seed=0, num_class=10):
    """Softmax boosting on the 50-feature multiclass workload (40 normal and 10
    categorical columns over 32 categories, max_bin 63): gradients on the device,
    num_class trees per iteration, each tree's score update on the device.
    `runs` holds (params, concurrent) pairs: concurrent runs grow the class trees
    of an iteration at the same time, one learner per class on its own stream and
    host thread (set_tree_slot keeps the trees those of one learner). Returns
    iterations per second for every run."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import torch

    import bankai_gpu as bg
    import boosting as bst
    import synth

    x, y = synth.multiclass_50(rows, seed=seed, num_class=num_class)
    ds = bg.Dataset(x, bg.DatasetParams(max_bin=63, num_threads=min(os.cpu_count() or 1, 16)),
                    categorical=synth.MULTICLASS_CATEGORICAL)
    obj = bst.MulticlassObjective(y, num_class)
    n = ds.num_data
    dev = torch.device("cuda:0")
    out = []
    pool = ThreadPoolExecutor(num_class)
    for params, concurrent in runs:
        learners = [bg.Learner(ds, params, False) for _ in range(num_class if concurrent else 1)]
        if concurrent:
            for k, lk in enumerate(learners):
                lk.set_tree_slot(k, num_class)
        init = np.concatenate([np.full(n, obj.boost_from_score(k)) for k in range(num_class)])
        score = torch.as_tensor(init, device=dev)
        label = torch.as_tensor(y.astype(np.float32), device=dev)
        g = torch.empty(num_class * n, dtype=torch.float32, device=dev)
        h = torch.empty(num_class * n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()

        def tree(k, learner):
            t = learner.train_device(g.data_ptr() + 4 * k * n, h.data_ptr() + 4 * k * n)
            learner.add_prediction_to_score_device(bst._round0(t["leaf_value"] * lr), score.data_ptr() + 8 * k * n)

        def iteration():
            learners[0].softmax_gradients_device(score.data_ptr(), label.data_ptr(), num_class, g.data_ptr(),
                                                 h.data_ptr())
            if concurrent:
                torch.cuda.synchronize()  # the gradients, for every learner's stream
                list(pool.map(lambda k: tree(k, learners[k]), range(num_class)))
                torch.cuda.synchronize()  # every score update, for the next gradients
            else:
                for k in range(num_class):
                    tree(k, learners[0])

        for _ in range(warmup):
            iteration()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            iteration()
        torch.cuda.synchronize()
        out.append(steps / (time.perf_counter() - t0))
        for lk in learners:
            lk.close()
    pool.shutdown()
    return out


def run_ranking(rows, steps, warmup, lr, runs, seed=0):
    """LambdaRank boosting on the 136-feature ranking workload (68 normal and 68
    uniform columns with 20% exact zeros, lognormal query sizes clipped to
    [5, 1000], relevance 0-4): gradients on the device (one block per query), one
    tree and its score update per iteration. Returns iterations per second for
    every parameter set in `runs`."""
    import numpy as np
    import torch

    import bankai_gpu as bg
    import boosting as bst
    import synth

    x, y, q = synth.ranking_136(rows, seed=seed)
    ds = bg.Dataset(x, bg.DatasetParams(num_threads=min(os.cpu_count() or 1, 16)))
    del x
    obj = bst.LambdarankObjective(y, q)
    n = ds.num_data
    dev = torch.device("cuda:0")
    tables, keep = bg.rank_tables_on_device(obj, dev)
    out = []
    for params in runs:
        learner = bg.Learner(ds, params, False)
        score = torch.zeros(n, dtype=torch.float64, device=dev)
        label = torch.as_tensor(y.astype(np.float32), device=dev)
        g = torch.empty(n, dtype=torch.float32, device=dev)
        h = torch.empty(n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()

        def iteration():
            learner.lambdarank_gradients_device(score.data_ptr(), label.data_ptr(), tables, g.data_ptr(), h.data_ptr())
            t = learner.train_device(g.data_ptr(), h.data_ptr())
            learner.add_prediction_to_score_device(bst._round0(t["leaf_value"] * lr), score.data_ptr())

        for _ in range(warmup):
            iteration()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            iteration()
        torch.cuda.synchronize()
        out.append(steps / (time.perf_counter() - t0))
        learner.close()
    return out, len(q)


def device_copy_gbs(mib=1024, reps=10):
    """Bandwidth of a plain streaming kernel (read mib MiB, write mib MiB: an
    elementwise negate), the practical HBM ceiling; its PMC counts also calibrate
    FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.neg(src, out=dst)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * mib * 1024 * 1024 / (ms * 1e-3) / 1e9


# algorithmic bytes per row of each kernel kind (the reference's index-based model):
# root histogram F + P, child histogram F + P + 4 (row index), partition 9
def algo_bytes_per_row(kind, groups, payload):
    return {"root_hist": groups + payload, "child_hist": groups + payload + 4, "partition": 9.0}.get(kind)


def traffic_per_row(kernel):
    """Measured HBM bytes per processed row of `kernel` (fetch + write) from the
    committed rocprofv3 PMC summary: FETCH_SIZE and WRITE_SIZE summed over the
    kernel's dispatches of a counted run, scaled by the same counters on a copy of
    known size, over the rows those dispatches processed (per-tree row counts of
    the same run). None when the summary does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_row")


def bs_steps(args):
    return max(args.breakdown_steps, 1)


def roofline_entry(kind, ks, rows, groups, payload, algo=None):
    """Roofline of one kernel kind over a timed run: algorithmic bytes (rows x the
    per-row model, or `algo` when given) over the summed HIP-event durations of its
    launches, against the 8 TB/s HBM peak; traffic = measured bytes per row x the
    same rows, per launch (the same launch mix as the algorithmic bytes)."""
    launches = max(ks["launches"], 1)
    ms = ks["ms"]
    if algo is None:
        algo = rows * algo_bytes_per_row(kind, groups, payload)
    per_row = algo / rows if rows else 0.0
    achieved = algo / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    tpr = traffic_per_row(kind)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": (tpr * rows / launches) if tpr else None,
            "kernel": kind, "launches": ks["launches"], "avg_launch_us": 1000.0 * ms / launches,
            "rows": rows, "algo_bytes_per_row": per_row, "algo_bytes_per_launch": algo / launches,
            "traffic_bytes_per_row": tpr}


def host_cpus():
    """Threads the CPU baseline uses: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when the host sets it (the GPU box gives a process its share of
    cores that way); the CPU model string from /proc/cpuinfo."""
    visible = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(visible, cap) if cap > 0 else visible
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, visible, model


def cpu_baseline(ds, y, params, threads, budget_s=25.0, max_iters=5):
    """Times the oracle (C++ restatement of the reference CPU learner) on this host
    with `threads` OpenMP threads: whole boosting iterations (host gradients, one
    tree, the score update) until budget_s or max_iters."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    import boosting as bst
    import oracle_binding as ob

    params.num_threads = threads
    os.environ["OMP_NUM_THREADS"] = str(threads)
    obj = bst.BinaryObjective(y)
    score = np.full(ds.num_data, obj.boost_from_score(0))
    learner = ob.OracleLearner(ds, params, False)
    done = 0
    t0 = time.perf_counter()
    while True:
        g, h = obj.gradients(score)
        t = learner.train(g, h)
        learner.add_prediction_to_score(t["leaf_value"] * 0.1, score)
        done += 1
        el = time.perf_counter() - t0
        if el > budget_s or done >= max_iters:
            break
    learner.close()
    return done / el, done, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the workload's job: 500 boosting iterations from the first tree (the trees
    # deepen as boosting goes on, so a sample of early iterations runs faster)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--rows", type=int, default=10_500_000)
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="iterations of the plain FP64 learners (0: the measured run's iterations and warmup)")
    ap.add_argument("--sequential-steps", type=int, default=3)
    ap.add_argument("--device-tree-steps", type=int, default=10,
                    help="iterations timed with the trees planned on the device (0 skips it)")
    ap.add_argument("--breakdown-steps", type=int, default=0,
                    help="iterations of the per-kernel breakdown run (0: the measured run's iterations, so both "
                         "runs grow the same trees)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--holdout", type=int, default=1_050_000)
    ap.add_argument("--auc-iters", type=int, default=500,
                    help="boosting iterations of both forms before the holdout AUC gate (0 skips it)")
    ap.add_argument("--multiclass-rows", type=int, default=2_000_000,
                    help="rows of the 10-class workload measured beside the headline (0 skips it)")
    ap.add_argument("--multiclass-steps", type=int, default=3)
    ap.add_argument("--ranking-rows", type=int, default=5_000_000,
                    help="documents of the LambdaRank workload measured beside the headline (0 skips it)")
    ap.add_argument("--ranking-steps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the device copy measured beside the learner")
    ap.add_argument("--tree-log", default="",
                    help="write every tree's row counts (histogram, partition) to this JSON file")
    ap.add_argument("--hist-tiling", default="",
                    help="force the histogram tiling as min_rows,max_blocks (default: the autotuner's choice)")
    ap.add_argument("--hist-bin-width", type=int, choices=(0, 4, 8), default=0,
                    help="bytes per quantized histogram bin (0: the learner's choice)")
    ap.add_argument("--partition-histograms", choices=("on", "off", "default"), default="default",
                    help="4-byte bins: build one child's histogram inside the partition")
    ap.add_argument("--round-planner", choices=("tree", "rounds", "host"), default="host",
                    help="who plans batched growth: the device for whole trees, the device for the rounds, the host")
    args = ap.parse_args()
    global ROUND_PLANNER, HIST_TILING, TREE_LOG
    TREE_LOG = [] if args.tree_log else None
    ROUND_PLANNER = {"tree": 1, "rounds": 2, "host": 0}[args.round_planner]
    HIST_TILING = tuple(int(v) for v in args.hist_tiling.split(",")) if args.hist_tiling else None
    if args.hist_bin_width:
        HIST_OPTS["bin_width"] = args.hist_bin_width
    if args.partition_histograms != "default":
        HIST_OPTS["partition_histograms"] = args.partition_histograms == "on"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the learner has no multi-GPU path (rows would shard across GPUs "
              "with one histogram all-reduce per build; not built).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import bankai_gpu as bg
    import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    seed = 0
    t0 = time.perf_counter()
    intercept = synth.solve_intercept(seed)
    x, y = synth.binary_28(args.rows, seed=seed, intercept=intercept)
    ds = bg.Dataset(x, bg.DatasetParams(num_threads=min(os.cpu_count() or 1, 16)))
    log(f"data {args.rows} x 28, positive rate {y.mean():.4f}, binned in {time.perf_counter() - t0:.1f}s")
    lr = 0.1
    qparams = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                               stochastic_rounding=1, quant_train_renew_leaf=1, seed=seed, num_threads=16,
                               form=bg.FORM_QUANTIZED, batched=1)
    # per-kernel breakdown: every launch of every kind timed with HIP events on the
    # learner's stream, over the measured run's iterations (the same trees: the
    # runs are deterministic); it names the dominant kernel and gives every
    # roofline below. The events cost host time, so this run is not the measured one.
    if args.breakdown_steps <= 0:
        args.breakdown_steps = args.steps
    _, _, kall, _ = run_gpu(ds, y, qparams, args.breakdown_steps, args.warmup, lr, profile=True)
    dom = max(kall, key=lambda k: kall[k]["ms"])
    # the measured run: no events, the headline's whole-job throughput
    elapsed, stats, _, qtrees = run_gpu(ds, y, qparams, args.steps, args.warmup, lr, profile=False)
    its = args.steps / elapsed
    log(f"quantized: {its:.2f} it/s ({1000 * elapsed / args.steps:.1f} ms/iter), dominant kernel {dom}")

    result = {
        "metric": "boost iters/sec @10.5M x 28, 255 leaves (quantized int8 GPU learner)",
        "value": its,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"binary logloss {args.rows} x 28 dense N(0,1) f32, 53% positive, num_leaves 255, "
                               "max_bin 255, min_data_in_leaf 20, lr 0.1, use_quantized_grad, 4 bins, stochastic "
                               "rounding, quant_train_renew_leaf", "rows": args.rows, "features": 28,
                   "num_leaves": 255, "max_bin": 255, "form": "quantized", "growth": "batched"},
    }
    # roofline of the dominant kernel: algorithmic bytes per launch over the
    # average launch duration (HIP events on the learner's stream, breakdown run)
    ks = kall[dom]
    groups, payload = ds.num_features, 2
    stride = (groups + 1) // 2 * 2 + payload  # row record: group bins, the payload, padded to 16 bytes
    stride = (stride + 15) // 16 * 16
    per_row = algo_bytes_per_row(dom, groups, payload)
    if dom == "partition":
        # rows from the bytes it moves (2 records per row); its algorithmic bytes
        # also hold the child histograms a histogram-building partition adds
        result["roofline"] = roofline_entry(dom, ks, ks["moved_bytes"] / (2 * stride), groups, payload,
                                            algo=ks["algo_bytes"])
    elif per_row:
        result["roofline"] = roofline_entry(dom, ks, ks["algo_bytes"] / per_row, groups, payload)
    else:
        launches = max(ks["launches"], 1)
        achieved = ks["algo_bytes"] / (ks["ms"] * 1e-3) / 1e9 if ks["ms"] > 0 else 0.0
        result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                              "launches": ks["launches"], "avg_launch_us": 1000.0 * ks["ms"] / launches}
    result["roofline"]["timed_trees"] = ("every tree of the breakdown run: the measured run's iterations and trees, "
                                         "timed in a run of their own")
    result["roofline"]["moved_bytes_per_launch"] = ks["moved_bytes"] / max(ks["launches"], 1)
    result["roofline"]["moved_GBs"] = ks["moved_bytes"] / (ks["ms"] * 1e-3) / 1e9 if ks["ms"] > 0 else 0.0
    # the histogram kernel's own rooflines (root and child builds) from the
    # breakdown run, where every launch of every kind is timed
    result["roofline_hist"] = {
        "root": roofline_entry("root_hist", kall["root_hist"], kall["root_hist"]["algo_bytes"] / (groups + payload),
                               groups, payload)}
    if kall.get("child_hist", {}).get("launches", 0) > 0:
        result["roofline_hist"]["child"] = roofline_entry("child_hist", kall["child_hist"],
                                                          kall["child_hist"]["algo_bytes"] / (groups + payload + 4),
                                                          groups, payload)
    else:  # built inside the partition: its bytes and time are the partition's
        part = kall["partition"]
        part_rows = part["moved_bytes"] / (2 * stride)
        result["roofline_hist"]["child"] = {
            "built_in": "partition", "hist_algo_bytes_per_iter": (part["algo_bytes"] - 9.0 * part_rows) / bs_steps(
                args),
            "partition": roofline_entry("partition", part, part_rows, groups, payload, algo=part["algo_bytes"])}
    bs = args.breakdown_steps
    result["kernels"] = {k: {"ms_per_iter": v["ms"] / bs, "launches_per_iter": v["launches"] / bs,
                             "algo_GBs": v["algo_bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0,
                             "moved_GBs": v["moved_bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0}
                         for k, v in kall.items()}
    total_ms = sum(v["ms"] for v in kall.values()) / bs
    result["kernel_ms_share"] = total_ms / (1000.0 * elapsed / args.steps)
    result["rows_per_iter"] = {k: v / args.steps for k, v in stats.items() if k not in ("hist_tiling", "host_phases_ms")}
    result["host_phases_ms_per_iter"] = {k: v / args.steps for k, v in stats["host_phases_ms"].items()}
    result["hist_tiling_autotuned"] = stats["hist_tiling"]
    # whole-iteration algorithmic bytes of the reference's work (committed splits only):
    # hist = sum R_b (F + P + I) with I = 0 for the root; part = sum Q_s * 9;
    # discretizer = N (8 + 16 + 2); score = N (4 + 16)
    f_bytes, p_bytes, n = ds.num_features, 2, args.rows
    hist_b = stats["committed_hist_rows"] * (f_bytes + p_bytes + 4) - args.steps * n * 4
    part_b = stats["committed_part_rows"] * 9
    iter_bytes = (hist_b + part_b) / args.steps + n * (26 + 20)
    result["iteration_algo_bytes"] = iter_bytes
    result["iteration_algo_GBs"] = iter_bytes / (elapsed / args.steps) / 1e9
    result["device_copy_GBs"] = device_copy_gbs(args.copy_mib)

    if args.holdout > 0:
        import boosting as bst

        hx, hy = synth.binary_28(args.holdout, seed=seed, intercept=intercept, stream=3)
        hxt = torch.as_tensor(hx, device="cuda:0")

        def holdout_auc(trees):
            # the host predictor and AUC pinned by the reference's fixtures
            # (tests/test_predict_metric.py), the tree walk run on the device
            return bst.auc(hy, bst.predict_raw(trees, hxt).cpu().numpy())

        result["holdout_auc"] = holdout_auc(qtrees)
    if args.device_tree_steps > 0:
        # the same learner with its trees planned and finished on the device (one
        # host wait per tree); the host-planned rounds above measure faster
        saved, ROUND_PLANNER = ROUND_PLANNER, 1
        del_, _, _, _ = run_gpu(ds, y, qparams, args.device_tree_steps, 2, lr, profile=False)
        ROUND_PLANNER = saved
        result["device_planned_tree_iters_per_sec"] = args.device_tree_steps / del_
        log(f"quantized, trees planned on the device: {args.device_tree_steps / del_:.2f} it/s")
    if args.sequential_steps > 0:
        sparams = dataclasses.replace(qparams, batched=0)
        sel, _, _, _ = run_gpu(ds, y, sparams, args.sequential_steps, 1, lr, profile=False)
        result["sequential_quantized_iters_per_sec"] = args.sequential_steps / sel
        log(f"sequential quantized: {args.sequential_steps / sel:.2f} it/s")
    if not args.no_plain:
        # the plain drop-in learner: FP64 histograms, one split per step as the
        # reference grows a tree; and the same form with batched growth
        # over the measured run's iterations (the same trees up to the forms' tolerance)
        psteps, pwarm = (args.plain_steps, 1) if args.plain_steps > 0 else (args.steps, args.warmup)
        pparams = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, form=bg.FORM_PLAIN_F64, batched=0)
        pel, _, _, _ = run_gpu(ds, y, pparams, psteps, pwarm, lr, profile=False)
        p_its = psteps / pel
        log(f"plain f64 (drop-in, sequential growth): {p_its:.2f} it/s")
        result["plain_f64_iters_per_sec"] = p_its
        result["speedup_vs_plain_f64"] = its / p_its
        bparams = dataclasses.replace(pparams, batched=1)
        bel, _, _, btrees = run_gpu(ds, y, bparams, psteps, pwarm, lr, profile=False)
        result["plain_f64_batched_iters_per_sec"] = psteps / bel
        result["speedup_vs_plain_f64_batched"] = its / (psteps / bel)
        result["plain_iterations"] = {"steps": psteps, "warmup": pwarm}
        log(f"plain f64 batched: {psteps / bel:.2f} it/s")
        # the accuracy gate: both forms trained for the workload's full iteration
        # count (the FP64 form with batched growth, which commits the drop-in
        # learner's trees), holdout AUC of each
        if args.holdout > 0 and args.auc_iters > 0:
            # the measured runs when they are the whole job, else runs of their own
            if args.steps == args.auc_iters and args.warmup == 0:
                el_q, qa = elapsed, qtrees
            else:
                el_q, _, _, qa = run_gpu(ds, y, qparams, args.auc_iters, 0, lr, profile=False)
            if psteps == args.auc_iters and pwarm == 0:
                el_p, pa = bel, btrees
            else:
                el_p, _, _, pa = run_gpu(ds, y, bparams, args.auc_iters, 0, lr, profile=False)
            aq, ap = holdout_auc(qa), holdout_auc(pa)
            result["holdout_auc_gate"] = {"iterations": args.auc_iters, "quantized_batched": aq,
                                          "plain_f64_batched": ap, "gap": abs(aq - ap),
                                          "holdout_rows": args.holdout}
            # the workload's whole run (the headline itself at the default 500 steps
            # from the first tree): its trees deepen as boosting goes on
            result["full_run"] = {"iterations": args.auc_iters,
                                  "quantized_batched_iters_per_sec": args.auc_iters / el_q,
                                  "plain_f64_batched_iters_per_sec": args.auc_iters / el_p,
                                  "speedup": el_p / el_q}
            log(f"whole {args.auc_iters}-iteration runs: quantized {args.auc_iters / el_q:.2f}, f64 batched "
                f"{args.auc_iters / el_p:.2f} iter/s")
            log(f"holdout AUC after {args.auc_iters} iterations: quantized {aq:.5f}, f64 {ap:.5f}")
    if args.multiclass_rows > 0:
        mq = bg.LearnerParams(num_leaves=127, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                              stochastic_rounding=1, quant_train_renew_leaf=1, seed=seed, num_threads=16,
                              form=bg.FORM_QUANTIZED, batched=1)
        mf = bg.LearnerParams(num_leaves=127, min_data_in_leaf=20, form=bg.FORM_PLAIN_F64, batched=1)
        mfs = dataclasses.replace(mf, batched=0)
        qc_its, q_its, f_its, fs_its = run_multiclass(args.multiclass_rows, args.multiclass_steps, 1, lr,
                                                      [(mq, True), (mq, False), (mf, False), (mfs, False)])
        result["other_workloads"] = {
            f"multiclass_{args.multiclass_rows}x50_10_classes": {
                "unit": "iter/s (10 trees each)", "num_leaves": 127, "max_bin": 63,
                "quantized_batched_concurrent_classes": qc_its, "quantized_batched": q_its,
                "plain_f64_batched": f_its, "plain_f64_drop_in": fs_its,
                "data": "synthetic: 40 N(0,1) + 10 categorical columns over 32 categories, "
                        "label = argmax of 10 linear scores + Gumbel noise; softmax gradients on the device"}}
        log(f"multiclass {args.multiclass_rows} x 50: quantized, classes concurrent {qc_its:.2f}, quantized "
            f"{q_its:.2f}, f64 batched {f_its:.2f}, f64 drop-in {fs_its:.2f} iter/s")
    if args.ranking_rows > 0:
        rq = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                              stochastic_rounding=1, quant_train_renew_leaf=1, seed=seed, num_threads=16,
                              form=bg.FORM_QUANTIZED, batched=1)
        rf = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, form=bg.FORM_PLAIN_F64, batched=1)
        (rq_its, rf_its), nq = run_ranking(args.ranking_rows, args.ranking_steps, 1, lr, [rq, rf])
        result.setdefault("other_workloads", {})[f"lambdarank_{args.ranking_rows}x136"] = {
            "unit": "iter/s", "num_leaves": 255, "max_bin": 255, "queries": nq,
            "quantized_batched": rq_its, "plain_f64_batched": rf_its,
            "data": "synthetic: 68 N(0,1) + 68 U(0,1) columns with 20% zeros, lognormal(4,1) query sizes in "
                    "[5, 1000], relevance 0-4; LambdaRank NDCG gradients on the device"}
        log(f"lambdarank {args.ranking_rows} x 136 ({nq} queries): quantized {rq_its:.2f}, f64 batched "
            f"{rf_its:.2f} iter/s")
    if not args.no_cpu_baseline:
        cparams = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                                   stochastic_rounding=1, quant_train_renew_leaf=1, seed=seed, form=bg.FORM_QUANTIZED)
        threads, visible, model = host_cpus()
        v, done, el = cpu_baseline(ds, y, cparams, threads)
        v1, done1, el1 = cpu_baseline(ds, y, cparams, 1, budget_s=20.0, max_iters=1)
        result["cpu_baseline"] = {"value": v, "unit": "iter/s", "cores": threads, "kind": "port",
                                  "sample": f"the first {done} boosting iterations of the same workload (quantized, "
                                            f"full {args.rows} rows) in {el:.1f}s on {threads} threads",
                                  "one_thread": {"value": v1, "unit": "iter/s",
                                                 "sample": f"{done1} iteration(s) in {el1:.1f}s"},
                                  "cpu_model": model, "host_cpus_visible": visible}
        log(f"cpu baseline: {v:.3f} iter/s on {threads} threads ({model}), {v1:.3f} iter/s on 1 thread")
    if args.tree_log:
        with open(args.tree_log, "w") as f:
            json.dump({"rows": args.rows, "trees": TREE_LOG}, f)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
