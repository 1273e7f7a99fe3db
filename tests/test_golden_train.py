"""End-to-end pin of the oracle's f64 learner against upstream LightGBM 4.7.0 models.

Each case under tests/golden/train/ is a copy of the reference's fixture
tests/fixtures/train/<case>/ (x.npy, y.npy, params.json, model.txt, plus w.npy for
weights and group.npy for query sizes; the bagging_* and goss* cases train each tree
on a row sample drawn by boosting.RowSampler; produced by
scripts/gen_train_fixtures.py with lightgbm==4.7.0, 1 thread, force_col_wise, seed 7).
The dataset is binned by this repository's host binning (bankai_gpu.Dataset), the
trees grown by the CPU oracle through the boosting loop of boosting.py, and every
tree compared with the upstream model text: structure exactly (split features,
thresholds, decision types, children, counts) and values to 1e-9 relative (the
fixtures were written on darwin/arm64, whose fused multiply-adds move the last bits;
crates/lgbm-core/tests/support/mod.rs:146-250 allows 64 ULP).
"""
import json
import os

import numpy as np
import pytest

import bankai_gpu as bg
import boosting as bst
import oracle_binding as ob

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "train")
CASES = sorted(os.listdir(GOLDEN)) if os.path.isdir(GOLDEN) else []


def build_case(case):
    d = os.path.join(GOLDEN, case)
    spec = json.load(open(os.path.join(d, "params.json")))
    p = spec["params"]
    x = np.load(os.path.join(d, "x.npy"))
    y = np.load(os.path.join(d, "y.npy"))
    w = np.load(os.path.join(d, "w.npy")) if os.path.exists(os.path.join(d, "w.npy")) else None
    model = open(os.path.join(d, "model.txt")).read()
    if os.path.exists(os.path.join(d, "group.npy")):
        spec["group"] = np.load(os.path.join(d, "group.npy"))
    return spec, p, x, y, w, model


def dataset_params(p):
    dp = bg.DatasetParams()
    dp.data_random_seed = bg.derive_data_random_seed(p["seed"]) if "seed" in p else 1
    dp.min_data_in_leaf = p.get("min_data_in_leaf", 20)
    dp.max_bin = p.get("max_bin", 255)
    dp.zero_as_missing = int(p.get("zero_as_missing", False))
    dp.num_threads = 1
    return dp


def learner_params(p):
    lp = bg.LearnerParams()
    for k in ("num_leaves", "max_depth", "min_data_in_leaf", "lambda_l1", "lambda_l2", "max_delta_step", "path_smooth",
              "min_gain_to_split", "min_sum_hessian_in_leaf", "cat_smooth", "cat_l2", "max_cat_threshold",
              "max_cat_to_onehot", "min_data_per_group"):
        if k in p:
            setattr(lp, k, p[k])
    for k in ("feature_fraction", "feature_fraction_bynode"):
        if k in p:
            setattr(lp, k, p[k])
    lp.feature_fraction_seed = bst.resolve_seed(p, "feature_fraction_seed")
    mono = [int(v) for v in str(p.get("monotone_constraints", "")).replace(" ", "").split(",") if v]
    if any(mono):
        raise NotImplementedError("monotone constraints are not supported by the learner")
    if p.get("interaction_constraints"):
        # Config's "[0,1,2],[3,4,5]" form (config/mod.rs:537-552)
        text = p["interaction_constraints"]
        lp.interaction_constraints = tuple(tuple(int(v) for v in grp.split(",") if v.strip())
                                           for grp in text.replace("]", "").split("[") if grp.strip(", "))
    lp.num_threads = 1
    return lp


def make_objective(p, y, w, group=None):
    obj = p["objective"]
    if obj == "lambdarank":
        return bst.LambdarankObjective(y, group, w, sigmoid=p.get("sigmoid", 1.0), norm=p.get("lambdarank_norm", True),
                                       truncation_level=p.get("lambdarank_truncation_level", 30),
                                       label_gain=p.get("label_gain"))
    if obj == "binary":
        return bst.BinaryObjective(y, w, sigmoid=p.get("sigmoid", 1.0), scale_pos_weight=p.get("scale_pos_weight", 1.0),
                                   is_unbalance=p.get("is_unbalance", False))
    if obj == "regression":
        return bst.RegressionL2Objective(y, w)
    if obj == "multiclass":
        return bst.MulticlassObjective(y, p["num_class"], w)
    raise ValueError(obj)


def train_case(case, learner_factory):
    spec, p, x, y, w, model = build_case(case)
    ds = bg.Dataset(x, dataset_params(p), categorical=spec["categorical"])
    objective = make_objective(p, y, w, spec.get("group"))
    lp = learner_params(p)
    sampler = bst.RowSampler(p, y, objective.num_class)
    # train.rs:533: a sampler that rescales hessians (GOSS) turns the constant hessian off
    learners = [learner_factory(ds, lp, objective.const_hess and not sampler.changes_hessians)]
    gbdt = bst.GBDT(learners, objective, ds.num_data, learning_rate=p.get("learning_rate", 0.1),
                    boost_from_average=p.get("boost_from_average", True), sampler=sampler)
    gbdt.train(spec["num_boost_round"])
    return ds, gbdt, model, objective


def compare_models(ds, gbdt, model, objective):
    header = bst.parse_model_header(model)
    assert header["feature_infos"].split(" ") == ds.feature_infos()
    want = bst.parse_model_trees(model)
    got = bst.parse_model_trees(bst.model_to_string(gbdt.models, objective.num_class, gbdt.ntpi, objective.name(),
                                                   [f"Column_{i}" for i in range(ds.view.num_total_features)],
                                                   ds.feature_infos()))
    assert len(got) == len(want)
    exact = ("num_leaves", "num_cat", "split_feature", "decision_type", "left_child", "right_child", "leaf_count",
             "internal_count", "cat_boundaries", "cat_threshold", "threshold")
    approx = ("leaf_value", "leaf_weight", "internal_weight", "internal_value", "split_gain")
    for i, (a, b) in enumerate(zip(got, want)):
        for k in exact:
            if k in b:
                if k == "threshold":
                    va = np.array(a[k].split(), float) if a[k] else np.zeros(0)
                    vb = np.array(b[k].split(), float) if b[k] else np.zeros(0)
                    assert np.array_equal(va, vb), (i, k)
                else:
                    assert a.get(k) == b[k], (i, k, a.get(k), b[k])
        for k in approx:
            va = np.array(a[k].split(), float) if a[k] else np.zeros(0)
            vb = np.array(b[k].split(), float) if b[k] else np.zeros(0)
            # %g with 6 digits for gains/internal values
            rtol = 1e-9 if k in ("leaf_value", "leaf_weight") else 2e-5
            assert np.allclose(va, vb, rtol=rtol, atol=1e-12), (i, k, va, vb)


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_upstream_model(case):
    ds, gbdt, model, objective = train_case(case, lambda ds, lp, ch: ob.OracleLearner(ds, lp, ch))
    compare_models(ds, gbdt, model, objective)
