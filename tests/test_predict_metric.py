"""The host predictor and AUC (the bench's holdout gate) against the reference's
fixtures, copied as data under tests/golden/:

  - predict/<case>/: upstream models (model.txt), test rows (test.npy, label in
    column 0 as crates/lgbm-core/tests/support/mod.rs:78-86 reads it) and raw
    scores (pred_raw.npy), checked by crates/lgbm-core/tests/golden_predict.rs:38-42.
    The cases cover numerical splits, NaN-as-missing (`missing`), zero-as-missing,
    categorical bitsets and five classes. Off the fixtures' platform the
    reference allows 64 ulps or 1e-15 (support/mod.rs:247-257); this predictor
    adds leaf values in tree order from 0.0, as the reference's does, and is
    required to match bit for bit;
  - metrics/binary{,_weighted}/: score, label (and weight) with the expected AUC of
    metric/binary.rs:73-135 (expected.json).
"""
import json
import os

import numpy as np
import pytest

import boosting as bst

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PRED_CASES = sorted(os.listdir(os.path.join(GOLDEN, "predict")))


@pytest.mark.parametrize("case", PRED_CASES)
def test_predict_raw_matches_upstream(case):
    d = os.path.join(GOLDEN, "predict", case)
    k, trees = bst.parse_model(open(os.path.join(d, "model.txt")).read())
    test = np.load(os.path.join(d, "test.npy"))
    want = np.load(os.path.join(d, "pred_raw.npy"))
    got = bst.predict_raw(trees, test[:, 1:], k)
    assert got.shape == want.shape
    assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_predict_raw_routes_every_missing_rule():
    """The `missing` case holds NaN feature values; the prediction differs when
    NaN is read as 0 for every split, so the NaN rule is exercised."""
    d = os.path.join(GOLDEN, "predict", "missing")
    k, trees = bst.parse_model(open(os.path.join(d, "model.txt")).read())
    test = np.load(os.path.join(d, "test.npy"))[:, 1:]
    assert np.isnan(test).any()
    want = np.load(os.path.join(d, "pred_raw.npy"))
    assert not np.array_equal(bst.predict_raw(trees, np.nan_to_num(test), k), want)


@pytest.mark.parametrize("case", ["binary", "binary_weighted"])
def test_auc_matches_upstream(case):
    d = os.path.join(GOLDEN, "metrics", case)
    score = np.load(os.path.join(d, "score.npy"))
    label = np.load(os.path.join(d, "label.npy"))
    w = np.load(os.path.join(d, "weight.npy")) if os.path.exists(os.path.join(d, "weight.npy")) else None
    want = dict(json.load(open(os.path.join(d, "expected.json"))))["auc"]
    assert bst.auc(label, score, w) == want


def test_auc_ties_and_degenerate_classes():
    # brute force: P(score_pos > score_neg) + 0.5 P(tie) over all pairs
    rng = np.random.default_rng(3)
    score = rng.integers(0, 7, 300).astype(np.float64)
    label = (rng.random(300) < 0.4).astype(np.float64)
    p, q = score[label > 0], score[label <= 0]
    want = ((p[:, None] > q[None, :]).sum() + 0.5 * (p[:, None] == q[None, :]).sum()) / (len(p) * len(q))
    assert abs(bst.auc(label, score) - want) < 1e-12
    assert abs(bst.auc(label, score, np.ones(300, np.float32)) - want) < 1e-12
    assert bst.auc(np.ones(10), np.arange(10.0)) == 1.0
    assert bst.auc(np.zeros(10), np.arange(10.0)) == 1.0
