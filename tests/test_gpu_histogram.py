"""Root histograms of the HIP learner (bk_learner_root_histogram) against the oracle.

  - quantized: exact integer sums, bit for bit, of the oracle's discretized
    gradients (int_histogram.rs:34-107 on the same packed values);
  - plain FP64: within the rounding of two summation orders of the same f32
    inputs, |gpu - seq| <= 2 * n_b * 2^-53 * sum|x| per bin;
  - FP32 compensated: within 2^-30 * sum|x| of the exact per-bin sum (math.fsum),
    a bound that an uncompensated f32 accumulation fails by orders of magnitude;
  - counts: exact.
The regression test at 1,000,000 x 100 (60 normal, 30 uniform, 10 lognormal
columns with 5% NaN) prints the per-bin error of both float forms against the
reference's sequential FP64 histogram.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

import bankai_gpu as bg
import oracle_binding as ob
import synth

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[case] {msg}", flush=True)
    sys.stdout.flush()


def exact_bin_sums(ds, x):
    """Correctly rounded per-bin sums of x (math.fsum), concatenated over groups."""
    out = []
    for col, nb in zip(ds.group_columns(), ds.group_num_total_bin()):
        order = np.argsort(col, kind="stable")
        bounds = np.searchsorted(col[order], np.arange(int(nb) + 1))
        xs = x[order].astype(np.float64)
        out.append([math.fsum(xs[bounds[b]:bounds[b + 1]]) for b in range(int(nb))])
    return np.concatenate([np.asarray(o) for o in out])


def abs_bin_sums(ds, x):
    out = []
    ax = np.abs(x.astype(np.float64))
    for col, nb in zip(ds.group_columns(), ds.group_num_total_bin()):
        out.append(np.bincount(col, weights=ax, minlength=int(nb)))
    return np.concatenate(out)


def bin_counts(ds):
    return np.concatenate([np.bincount(c, minlength=int(nb)) for c, nb in
                           zip(ds.group_columns(), ds.group_num_total_bin())])


@pytest.fixture(scope="module")
def binary_20k():
    x, _ = synth.binary_28(20_000, seed=3)
    return bg.Dataset(x)


@pytest.mark.parametrize("bins,stochastic,const_hess", [(4, 1, 0), (4, 0, 0), (16, 1, 1), (127, 0, 0)])
def test_root_histogram_quantized_exact(binary_20k, bins, stochastic, const_hess):
    announce(f"root int histogram bins={bins} stochastic={stochastic} const_hess={const_hess}")
    ds = binary_20k
    rng = np.random.default_rng(bins)
    g = rng.standard_normal(ds.num_data).astype(np.float32)
    h = np.ones(ds.num_data, np.float32) if const_hess else rng.uniform(0.01, 0.25, ds.num_data).astype(np.float32)
    p = bg.LearnerParams(num_leaves=7, use_quantized_grad=1, num_grad_quant_bins=bins, stochastic_rounding=stochastic,
                         seed=4, num_threads=3, form=bg.FORM_QUANTIZED)
    gpu = bg.Learner(ds, p, const_hess)
    ih, cnt = gpu.root_histogram(g, h)
    # the oracle's first tree discretizes with the same noise offset
    o = ob.OracleLearner(ds, p, const_hess)
    o.train(g, h)
    packed, _, _ = o.quant_state()
    qg = (packed >> 8).astype(np.int64)
    qh = (packed & 0xFF).astype(np.int64)
    want = []
    for col, nb in zip(ds.group_columns(), ds.group_num_total_bin()):
        s = np.zeros((int(nb), 2), np.int64)
        np.add.at(s[:, 0], col, qg)
        np.add.at(s[:, 1], col, qh)
        want.append(s)
    want = np.concatenate(want)
    assert np.array_equal(ih, want)
    assert np.array_equal(cnt, bin_counts(ds))


def oracle_int_root_histogram(ds, p, g, h, const_hess):
    """Exact per-bin sums of the oracle's discretized gradients for its first tree."""
    o = ob.OracleLearner(ds, p, const_hess)
    o.train(g, h)
    packed, _, _ = o.quant_state()
    qg = (packed >> 8).astype(np.int64)
    qh = (packed & 0xFF).astype(np.int64)
    want = []
    for col, nb in zip(ds.group_columns(), ds.group_num_total_bin()):
        s = np.zeros((int(nb), 2), np.int64)
        np.add.at(s[:, 0], col, qg)
        np.add.at(s[:, 1], col, qh)
        want.append(s)
    return np.concatenate(want)


# (min_rows, max_blocks): the autotuner's extremes, and row blocks asked to be far
# larger than one LDS accumulation of 127-bin gradients holds (6144 rows), which
# the planner must split so that no packed partial field overflows
TILINGS = [(2048, 2048), (8192, 512), (65536, 4), (1 << 20, 1)]


# (groups per tile, LDS copies per tile): every layout the autotuner may pick;
# 1024 groups per tile is the whole record
LAYOUTS = [(16, 1), (32, 1), (1024, 1), (16, 2), (32, 2)]


@pytest.fixture(scope="module")
def binary_200k():
    x, _ = synth.binary_28(200_000, seed=8)
    return bg.Dataset(x)


@pytest.mark.parametrize("bins", [4, 127])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tiling", TILINGS)
def test_root_histogram_any_tiling_exact(binary_200k, bins, tiling, layout):
    announce(f"root int histogram 200k x 28 bins={bins} tiling={tiling} layout={layout}")
    ds = binary_200k
    rng = np.random.default_rng(11)
    g = rng.standard_normal(ds.num_data).astype(np.float32)
    h = rng.uniform(0.01, 0.25, ds.num_data).astype(np.float32)
    # every hessian quantizes to the top bins: the hessian fields fill fastest
    h[::2] = 0.25
    p = bg.LearnerParams(num_leaves=7, use_quantized_grad=1, num_grad_quant_bins=bins, stochastic_rounding=0,
                         seed=4, num_threads=1, form=bg.FORM_QUANTIZED)
    gpu = bg.Learner(ds, p, False)
    gpu.set_hist_tiling(*tiling)
    gpu.set_hist_layout(*layout)
    ih, cnt = gpu.root_histogram(g, h)
    assert np.array_equal(ih, oracle_int_root_histogram(ds, p, g, h, False))
    assert np.array_equal(cnt, bin_counts(ds))
    t = gpu.last_stats()["hist_tiling"]
    assert (t["min_rows"], t["max_blocks"], t["tile_groups"], t["sub_histograms"]) == tiling + layout


@pytest.mark.parametrize("layout", LAYOUTS)
def test_root_histogram_layouts_wide_records(layout):
    """100 features with NaN: tiles are also cut where a tile's bins would exceed
    the LDS budget (not only every groups_per_tile groups)."""
    announce(f"root int histogram 100k x 100 layout={layout}")
    x, _ = synth.regression_100(100_000, seed=3)
    ds = bg.Dataset(x)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(ds.num_data).astype(np.float32)
    h = rng.uniform(0.01, 0.25, ds.num_data).astype(np.float32)
    p = bg.LearnerParams(num_leaves=7, use_quantized_grad=1, num_grad_quant_bins=4, stochastic_rounding=0,
                         seed=4, num_threads=1, form=bg.FORM_QUANTIZED)
    gpu = bg.Learner(ds, p, False)
    gpu.set_hist_layout(*layout)
    ih, cnt = gpu.root_histogram(g, h)
    assert np.array_equal(ih, oracle_int_root_histogram(ds, p, g, h, False))
    assert np.array_equal(cnt, bin_counts(ds))
    assert gpu.last_stats()["hist_tiling"]["ntiles"] >= 2


def test_hist_layout_refused_out_of_range(binary_200k):
    announce("histogram layout arguments out of range")
    gpu = bg.Learner(binary_200k, bg.LearnerParams(num_leaves=7, use_quantized_grad=1, form=bg.FORM_QUANTIZED), False)
    for bad in ((8, 1), (24, 1), (16, 3), (16, 0), (0, 1)):
        with pytest.raises(bg.BankaiError):
            gpu.set_hist_layout(*bad)
    gpu.set_hist_layout(0, 0)


@pytest.mark.parametrize("const_hess", [0, 1])
def test_root_histogram_f64(binary_20k, const_hess):
    announce(f"root f64 histogram const_hess={const_hess}")
    ds = binary_20k
    rng = np.random.default_rng(5)
    g = (rng.standard_normal(ds.num_data) * np.exp(rng.uniform(-8, 8, ds.num_data))).astype(np.float32)
    h = np.full(ds.num_data, 0.75, np.float32) if const_hess else rng.uniform(0, 1, ds.num_data).astype(np.float32)
    gpu = bg.Learner(ds, bg.LearnerParams(num_leaves=7, form=bg.FORM_PLAIN_F64), const_hess)
    gh, cnt = gpu.root_histogram(g, h)
    want = ob.root_hist_f64(ds, g, h, const_hess)
    n_b = bin_counts(ds)
    assert np.array_equal(cnt, n_b)
    tol_g = 2.0 * n_b * 2.0 ** -53 * abs_bin_sums(ds, g)
    assert np.all(np.abs(gh[:, 0] - want[:, 0]) <= tol_g)
    if const_hess:
        assert np.array_equal(gh[:, 1], want[:, 1])  # counts times h[0]: exact
    else:
        tol_h = 2.0 * n_b * 2.0 ** -53 * abs_bin_sums(ds, h)
        assert np.all(np.abs(gh[:, 1] - want[:, 1]) <= tol_h)


def test_root_histogram_f32_compensated_bound():
    announce("root f32c histogram 200k x 100 against exact per-bin sums")
    x, _ = synth.regression_100(200_000, seed=1)
    ds = bg.Dataset(x)
    rng = np.random.default_rng(6)
    g = (rng.standard_normal(ds.num_data) * np.exp(rng.uniform(-6, 6, ds.num_data))).astype(np.float32)
    h = rng.uniform(0, 1, ds.num_data).astype(np.float32)
    gpu = bg.Learner(ds, bg.LearnerParams(num_leaves=7, form=bg.FORM_F32_COMPENSATED), False)
    gh, cnt = gpu.root_histogram(g, h)
    assert np.array_equal(cnt, bin_counts(ds))
    for j, v in ((0, g), (1, h)):
        exact = exact_bin_sums(ds, v)
        scale = abs_bin_sums(ds, v)
        err = np.abs(gh[:, j] - exact)
        assert np.all(err <= 2.0 ** -30 * scale), (j, float(np.max(err / np.maximum(scale, 1e-300))))


def per_bin_report(ds, gh, want, x):
    scale = abs_bin_sums(ds, x)
    err = np.abs(gh - want)
    rel = err / np.maximum(np.abs(want), 1e-300)
    return {"max_abs": float(err.max()), "max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
            "max_err_over_abs_sum": float((err / np.maximum(scale, 1e-300)).max()),
            "bins_differing": int(np.count_nonzero(err))}


def test_f32_compensated_error_report_1m_x_100():
    """Per-bin |form - FP64| at 1,000,000 x 100 with NaN missing values, for the
    first L2 iteration (constant hessian) and for variable hessians."""
    announce("per-bin error report 1M x 100")
    x, y = synth.regression_100(1_000_000, seed=0)
    ds = bg.Dataset(x)
    n = ds.num_data
    total_bins = int(np.sum(ds.group_num_total_bin()))
    g_l2 = (np.float64(np.mean(y)) - y).astype(np.float32)  # boost_from_average, then score - label
    rng = np.random.default_rng(7)
    cases = {"l2_first_iteration": (g_l2, np.ones(n, np.float32), True),
             "variable_hessian": (rng.standard_normal(n).astype(np.float32),
                                  rng.uniform(0.0, 0.25, n).astype(np.float32), False)}
    report = {"rows": n, "features": ds.num_features, "total_bins": total_bins}
    for name, (g, h, ch) in cases.items():
        want = ob.root_hist_f64(ds, g, h, ch)
        report[name] = {}
        for form, tag in ((bg.FORM_F32_COMPENSATED, "f32_compensated"), (bg.FORM_PLAIN_F64, "plain_f64")):
            gpu = bg.Learner(ds, bg.LearnerParams(num_leaves=7, form=form), ch)
            gh, cnt = gpu.root_histogram(g, h)
            assert np.array_equal(cnt, bin_counts(ds))
            r = {"g": per_bin_report(ds, gh[:, 0], want[:, 0], g)}
            if not ch:
                r["h"] = per_bin_report(ds, gh[:, 1], want[:, 1], h)
            else:
                assert np.array_equal(gh[:, 1], want[:, 1])
            report[name][tag] = r
            gpu.close()
            assert r["g"]["max_err_over_abs_sum"] <= 2.0 ** -30
    print("[report] " + json.dumps(report), flush=True)
    out = os.environ.get("BK_REPORT_DIR")
    if out:
        with open(os.path.join(out, "f32c_error_1Mx100.json"), "w") as f:
            json.dump(report, f, indent=1)
