"""Full-size parity: the benchmark workload (10.5M x 28, 255 leaves, quantized,
4 bins, stochastic rounding, leaf renewal) grown by the HIP learner and by the
oracle from the same gradients must give identical trees, bit for bit, and the
same row-to-leaf assignment (checked through per-leaf row counts and sums of
row ids, and through the score update).
"""
import numpy as np
import pytest

import bankai_gpu as bg
import boosting as bst
import oracle_binding as ob
import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [1, 0])
def test_fullsize_quantized_trees_bit_exact(batched):
    print(f"\n[case] full size 10.5M x 28 quantized batched={batched}, 2 trees, GPU vs oracle", flush=True)
    n = 10_500_000
    x, y = synth.binary_28(n, seed=0, intercept=synth.solve_intercept(0))
    ds = bg.Dataset(x, bg.DatasetParams(num_threads=16))
    del x
    p = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                         stochastic_rounding=1, quant_train_renew_leaf=1, seed=0, num_threads=16,
                         form=bg.FORM_QUANTIZED, batched=batched)
    obj = bst.BinaryObjective(y)
    score = np.full(n, obj.boost_from_score(0))
    gpu = bg.Learner(ds, p)
    ora = ob.OracleLearner(ds, p)
    for it in range(2):
        g, h = obj.gradients(score)
        tg = gpu.train(g, h)
        to = ora.train(g, h)
        assert tg["num_leaves"] == to["num_leaves"] == 255
        for k in ("split_feature", "threshold_in_bin", "decision_type", "left_child", "right_child", "split_gain",
                  "leaf_value", "leaf_weight", "leaf_count", "internal_value", "internal_weight", "internal_count"):
            assert tg[k].tobytes() == to[k].tobytes(), (it, k)
        vals = to["leaf_value"] * 0.1
        s2 = score.copy()
        gpu.add_prediction_to_score(vals, s2)
        ora.add_prediction_to_score(vals, score)
        assert np.array_equal(s2, score)
        for leaf in (0, 7, 128, 254):
            rg, ro = gpu.leaf_rows(leaf), ora.leaf_rows(leaf)
            assert len(rg) == len(ro) and int(rg.astype(np.int64).sum()) == int(ro.astype(np.int64).sum())
            assert np.all(np.diff(rg) > 0)


def test_fullsize_bagged_feature_sampled_bit_exact():
    """The same workload with bagging (fraction 0.6 every iteration) and feature
    sampling by tree (0.8) and by node (0.9): two trees from the GPU learner and
    the oracle on the same bags, bit for bit, with the out-of-bag rows scored by
    both."""
    print("\n[case] full size 10.5M x 28 quantized batched, bagging 0.6 + feature_fraction 0.8/0.9, 2 trees",
          flush=True)
    n = 10_500_000
    x, y = synth.binary_28(n, seed=0, intercept=synth.solve_intercept(0))
    ds = bg.Dataset(x, bg.DatasetParams(num_threads=16))
    del x
    p = bg.LearnerParams(num_leaves=255, min_data_in_leaf=20, use_quantized_grad=1, num_grad_quant_bins=4,
                         stochastic_rounding=1, quant_train_renew_leaf=1, seed=0, num_threads=16,
                         form=bg.FORM_QUANTIZED, batched=1, feature_fraction=0.8, feature_fraction_bynode=0.9,
                         feature_fraction_seed=3)
    obj = bst.BinaryObjective(y)
    sampler = bst.RowSampler({"bagging_fraction": 0.6, "bagging_freq": 1, "bagging_seed": 5}, y)
    score = np.full(n, obj.boost_from_score(0))
    gpu = bg.Learner(ds, p, False)
    ora = ob.OracleLearner(ds, p, False)
    for it in range(2):
        g, h = obj.gradients(score)
        rows = sampler.sample(it, g, h)
        tg = gpu.train(g, h, rows)
        to = ora.train(g, h, rows)
        assert tg["num_leaves"] == to["num_leaves"]
        for k in ("split_feature", "threshold_in_bin", "decision_type", "left_child", "right_child", "split_gain",
                  "leaf_value", "leaf_weight", "leaf_count", "internal_value", "internal_weight", "internal_count"):
            assert tg[k].tobytes() == to[k].tobytes(), (it, k)
        vals = to["leaf_value"] * 0.1
        s2 = score.copy()
        gpu.add_prediction_to_score(vals, s2)
        ora.add_prediction_to_score(vals, score)
        oob = bst.out_of_bag(rows, n)
        gpu.add_out_of_bag_prediction(vals, oob, s2)
        ora.add_out_of_bag_prediction(vals, oob, score)
        assert np.array_equal(s2, score)
