"""Pins of the oracle's quantized learner (use_quantized_grad), each against
something other than the oracle's own formulas:

  - the stochastic-rounding noise layout, against the reference's Threading::BlockInfo
    known answers (crates/lgbm-core/src/treelearner/multi_val.rs:636-640 for the
    function at gradient_discretizer/mod.rs:19-29) and the pinned MT19937 engine;
  - the root sums (TL/serial.rs:741-761: sequential f64 sums of int * scale in row
    order), against sums recomputed here from the discretized gradients in Python's
    own IEEE double arithmetic, in row order; a reversed or exact order is shown to
    give other bits on the same data, so the check tells the orders apart;
  - the leaf renewal (TL/serial.rs:634-655: per leaf, sequential f64 sums of the
    original f32 gradients over the leaf's rows in partition order, then
    calculate_splitted_leaf_output<L1, MAX_OUTPUT, !SMOOTH>, feature_histogram.rs:58-100),
    against sums over the oracle's leaf rows recomputed here;
  - the integer orchestration as a whole (integer histograms, fix_histogram,
    sibling subtraction, integer scans, the children's integer sums): with
    gradients that are exact multiples of the gradient scale, nearest rounding and
    a constant hessian, every integer histogram times its scale equals the f64
    histogram exactly, so the quantized learner must grow the tree of the f64
    learner, which the upstream golden models pin (test_golden_train.py). Split
    structure must be identical (the two forms add the reference's kEpsilon in
    different places, so gains agree to ~1e-15 relative, far inside the 1e-9
    tie tolerance); leaf values agree to 1e-12 relative.
"""
import os

import numpy as np
import pytest

import bankai_gpu as bg
import oracle_binding as ob
import synth
from test_golden_train import GOLDEN, build_case, dataset_params, learner_params

K_ZERO_THRESHOLD = float(np.float32(1e-35))


def test_block_info_reference_vectors():
    # crates/lgbm-core/src/treelearner/multi_val.rs:636-640
    assert ob.block_info(1, 10_000, 32) == (1, 10_000)
    assert ob.block_info(4, 10_000, 32) == (4, 2528)
    assert ob.block_info(8, 100, 32) == (4, 32)
    assert ob.block_info(4, 0, 32) == (0, 0)


def test_noise_blocks_start_at_reference_block_size():
    """init_stochastic (gradient_discretizer/mod.rs:87-106) with 4 threads over
    10,000 rows cuts 4 blocks of 2528 rows (the reference vector above: the
    512-row minimum does not bind), block b drawing from MT19937(seed + b) for
    gradients and MT19937(seed + b + 4) for hessians, from their first draw."""
    n, seed, nt, size = 10_000, 11, 4, 2528
    gn, hn = ob.noise(n, seed, nt)
    for b in range(4):
        eg, eh = ob.Mt19937(seed + b), ob.Mt19937(seed + b + nt)
        lo, hi = b * size, min((b + 1) * size, n)
        want_g = [eg.uniform_real() for _ in range(hi - lo)]
        want_h = [eh.uniform_real() for _ in range(hi - lo)]
        assert np.array_equal(gn[lo:hi], want_g), b
        assert np.array_equal(hn[lo:hi], want_h), b
    # one thread: a single stream over every row
    g1, _ = ob.noise(n, seed, 1)
    e = ob.Mt19937(seed)
    assert np.array_equal(g1, [e.uniform_real() for _ in range(n)])


def seq_sum(values):
    """Left-to-right IEEE double sum (Python floats: no fused multiply-add)."""
    s = 0.0
    for v in values:
        s += v
    return s


def quant_params(**kw):
    p = dict(num_leaves=15, min_data_in_leaf=5, use_quantized_grad=1, num_grad_quant_bins=4, stochastic_rounding=1,
             seed=3, num_threads=2, form=bg.FORM_QUANTIZED)
    p.update(kw)
    return bg.LearnerParams(**p)


# bins 6: scales max|g| / 3 and max h / 6 carry full 53-bit significands, so the
# f64 partial sums round and their order shows; bins 4: scales with short
# significands (sums exact in any order)
@pytest.mark.parametrize("bins,stochastic,const_hess,bagged", [(6, 1, 0, 0), (6, 0, 0, 0), (6, 1, 1, 0), (6, 1, 0, 1),
                                                               (6, 0, 1, 1), (4, 1, 0, 0)])
def test_root_sums_sequential_in_row_order(bins, stochastic, const_hess, bagged):
    x, _ = synth.binary_28(6000, seed=5)
    ds = bg.Dataset(x)
    rng = np.random.default_rng(7)
    # gradients over several decades: the f64 partial sums of int * scale round
    g = (rng.standard_normal(ds.num_data) * np.exp(rng.uniform(-3, 3, ds.num_data))).astype(np.float32)
    h = np.ones(ds.num_data, np.float32) if const_hess else rng.uniform(0.01, 0.3, ds.num_data).astype(np.float32)
    rows = np.sort(rng.choice(ds.num_data, 4000, replace=False)).astype(np.int32) if bagged else None
    o = ob.OracleLearner(ds, quant_params(stochastic_rounding=stochastic, num_grad_quant_bins=bins), const_hess)
    o.train(g, h, rows)
    packed, gs, hs = o.quant_state()
    qg = (packed >> 8).astype(np.int64)
    qh = (packed & 0xFF).astype(np.int64)
    order = rows if bagged else np.arange(ds.num_data)
    want_g = seq_sum(float(qg[i]) * gs for i in order)
    want_h = seq_sum(float(qh[i]) * hs for i in order)
    sg, sh, isum = o.root_sums()
    assert sg == want_g and sh == want_h
    ig, ih = int(qg[order].sum()), int(qh[order].sum())
    assert isum == (ig << 32) + ih
    if bins == 6:
        # the check tells summation orders apart on these data
        rev_g = seq_sum(float(qg[i]) * gs for i in order[::-1])
        rev_h = seq_sum(float(qh[i]) * hs for i in order[::-1])
        exact_g = float(np.sum(qg[order])) * gs
        assert rev_g != want_g or exact_g != want_g or (not const_hess and rev_h != want_h)


def leaf_output_renew(sg, sh, l1, l2, mds):
    """calculate_splitted_leaf_output<true, true, false> (feature_histogram.rs:78-100)."""
    sign = lambda v: float((v > 0) - (v < 0))
    reg = sign(sg) * max(abs(sg) - l1, 0.0)
    ret = -reg / (sh + l2)
    if mds > 0.0 and abs(ret) > mds:
        ret = sign(ret) * mds
    return 0.0 if -K_ZERO_THRESHOLD <= ret <= K_ZERO_THRESHOLD else ret


@pytest.mark.parametrize("l1,l2,mds,bagged", [(0.0, 0.0, 0.0, 0), (0.3, 1.5, 0.0, 0), (0.0, 0.5, 0.05, 0),
                                              (0.1, 0.0, 0.0, 1)])
def test_renewal_sequential_over_leaf_rows(l1, l2, mds, bagged):
    x, _ = synth.binary_28(5000, seed=9)
    ds = bg.Dataset(x)
    rng = np.random.default_rng(13)
    # 2^40 of dynamic range: a leaf's f64 sum of f32 values rounds, so its order shows
    g = (rng.standard_normal(ds.num_data) * np.exp(rng.uniform(-14, 14, ds.num_data))).astype(np.float32)
    h = rng.uniform(0.01, 0.3, ds.num_data).astype(np.float32)
    rows = np.sort(rng.choice(ds.num_data, 3500, replace=False)).astype(np.int32) if bagged else None
    p = quant_params(num_leaves=31, quant_train_renew_leaf=1, lambda_l1=l1, lambda_l2=l2, max_delta_step=mds)
    o = ob.OracleLearner(ds, p, False)
    t = o.train(g, h, rows)
    assert t["num_leaves"] > 8
    order_matters = False
    for leaf in range(t["num_leaves"]):
        lr = o.leaf_rows(leaf)
        assert np.all(np.diff(lr) > 0), "partition order is ascending row order"
        assert len(lr) == t["leaf_count"][leaf]
        sg = seq_sum(float(v) for v in g[lr])
        sh = seq_sum(float(v) for v in h[lr])
        assert t["leaf_value"][leaf] == leaf_output_renew(sg, sh, l1, l2, mds), leaf
        order_matters |= seq_sum(float(v) for v in g[lr][::-1]) != sg
    assert order_matters, "the data do not tell summation orders apart"


# golden datasets whose features cover dense, NaN, zero-as-missing and categorical columns
CROSS_CASES = ["binary", "l2_nan", "l2_zero_as_missing", "cat_default", "cat_nan", "l2_regularized"]


def exact_multiple_gradients(x, seed):
    """Integer k in [-2, 2] (both ends present) times 1/8: with 4 quantization bins the
    gradient scale is max|g| / 2 = 1/8 exactly and nearest rounding returns k."""
    rng = np.random.default_rng(seed)
    z = np.nan_to_num(x[:, 0]) + 0.7 * np.nan_to_num(x[:, 1 % x.shape[1]]) + rng.normal(0, 0.5, x.shape[0])
    k = np.clip(np.rint(1.5 * np.tanh(z)), -2, 2)
    k[0], k[1] = -2, 2
    return (k * 0.125).astype(np.float32)


@pytest.mark.parametrize("case", [c for c in CROSS_CASES if os.path.isdir(os.path.join(GOLDEN, c))])
def test_quantized_learner_grows_the_f64_tree_on_exact_gradients(case):
    spec, p, x, y, w, model = build_case(case)
    ds = bg.Dataset(x, dataset_params(p), categorical=spec["categorical"])
    base = learner_params(p)
    base.num_leaves = max(base.num_leaves, 15)
    g = exact_multiple_gradients(x, len(case))
    h = np.ones(ds.num_data, np.float32)
    f64 = ob.OracleLearner(ds, base, True).train(g, h)
    qp = bg.LearnerParams(**{**base.__dict__, "use_quantized_grad": 1, "num_grad_quant_bins": 4,
                             "stochastic_rounding": 0, "quant_train_renew_leaf": 0,
                             "form": bg.FORM_QUANTIZED})
    qo = ob.OracleLearner(ds, qp, True)
    q = qo.train(g, h)
    packed, gs, hs = qo.quant_state()
    assert gs == 0.125 and hs == 1.0
    assert np.array_equal((packed >> 8).astype(np.int64), np.rint(g / 0.125).astype(np.int64))
    assert f64["num_leaves"] == q["num_leaves"] > 3
    for k in ("split_feature_inner", "threshold_in_bin", "decision_type", "left_child", "right_child", "leaf_count",
              "internal_count", "cat_threshold", "cat_boundaries"):
        if k in f64:
            assert np.array_equal(f64[k], q[k]), k
    for k in ("leaf_value", "leaf_weight", "internal_value", "internal_weight"):
        assert np.allclose(f64[k], q[k], rtol=1e-12, atol=1e-300), k
    assert np.allclose(f64["split_gain"], q["split_gain"], rtol=1e-6)
