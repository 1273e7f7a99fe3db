"""Boosting loop around a tree learner, and the model text writer.

`GBDT` restates the reference's `Trainer::train_one_iter_with` for the plain
GBDT case (crates/lgbm-core/src/boosting/train.rs:723-839): boost-from-average
init scores (615-642), objective gradients, one tree per class, shrinkage,
score update and `add_bias` on the first iteration. The learner is any object
with the `train(g, h)` / `add_prediction_to_score(values, score)` methods: the
GPU `bankai_gpu.Learner` in production.

Objectives are the closed-form host ones of crates/lgbm-core/src/objective/
(binary.rs:71-124, regression.rs:127-142,240-330 for L2, multiclass.rs:68-105,
rank.rs:183-361 for lambdarank).

`model_to_string` writes the model text of crates/lgbm-core/src/boosting/model_text.rs:88-138
and crates/lgbm-core/src/tree.rs:520-601 (tree blocks; no feature importances or
parameter section).
"""
import numpy as np

K_EPSILON = float(np.float32(1e-15))
K_ZERO_THRESHOLD = float(np.float32(1e-35))


def _round0(v):
    v = np.asarray(v, dtype=np.float64)
    return np.where((v >= -K_ZERO_THRESHOLD) & (v <= K_ZERO_THRESHOLD), 0.0, v)


class Tree:
    """Tree::apply_shrinkage / add_bias / to_model_string over a trained tree dict."""

    def __init__(self, d):
        self.d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        self.shrinkage = 1.0

    @property
    def num_leaves(self):
        return self.d["num_leaves"]

    @property
    def leaf_value(self):
        return self.d["leaf_value"]

    def apply_shrinkage(self, rate):
        """tree.rs:437-462"""
        self.d["leaf_value"] = _round0(self.d["leaf_value"] * rate)
        self.d["internal_value"] = _round0(self.d["internal_value"] * rate)
        self.shrinkage *= rate

    def add_bias(self, val):
        """tree.rs:465-478"""
        self.d["leaf_value"] = _round0(self.d["leaf_value"] + val)
        self.d["internal_value"] = _round0(self.d["internal_value"] + val)
        self.shrinkage = 1.0

    @staticmethod
    def constant(value, count):
        d = {"num_leaves": 1, "num_cat": 0, "leaf_value": np.array([value]), "leaf_weight": np.array([0.0]),
             "leaf_count": np.array([count], np.int32), "internal_value": np.zeros(0), "internal_weight": np.zeros(0),
             "internal_count": np.zeros(0, np.int32), "split_feature": np.zeros(0, np.int32),
             "split_gain": np.zeros(0, np.float32), "threshold": np.zeros(0), "decision_type": np.zeros(0, np.int8),
             "left_child": np.zeros(0, np.int32), "right_child": np.zeros(0, np.int32),
             "cat_boundaries": np.zeros(1, np.int32), "cat_threshold": np.zeros(0, np.uint32)}
        return Tree(d)

    def to_model_string(self):
        d = self.d
        n = d["num_leaves"]
        ni = n - 1
        g6 = lambda v: "%g" % v  # noqa: E731
        g17 = lambda v: "%.17g" % v  # noqa: E731
        join = lambda a, f=str: " ".join(f(x) for x in a)  # noqa: E731
        lines = [f"num_leaves={n}", f"num_cat={d['num_cat']}",
                 "split_feature=" + join(d["split_feature"][:ni], lambda x: str(int(x))),
                 "split_gain=" + join(d["split_gain"][:ni], lambda x: g6(float(x))),
                 "threshold=" + join(d["threshold"][:ni], lambda x: g17(float(x))),
                 "decision_type=" + join(d["decision_type"][:ni], lambda x: str(int(x))),
                 "left_child=" + join(d["left_child"][:ni], lambda x: str(int(x))),
                 "right_child=" + join(d["right_child"][:ni], lambda x: str(int(x))),
                 "leaf_value=" + join(d["leaf_value"][:n], lambda x: g17(float(x))),
                 "leaf_weight=" + join(d["leaf_weight"][:n], lambda x: g17(float(x))),
                 "leaf_count=" + join(d["leaf_count"][:n], lambda x: str(int(x))),
                 "internal_value=" + join(d["internal_value"][:ni], lambda x: g6(float(x))),
                 "internal_weight=" + join(d["internal_weight"][:ni], lambda x: g6(float(x))),
                 "internal_count=" + join(d["internal_count"][:ni], lambda x: str(int(x)))]
        if d["num_cat"] > 0:
            lines.append("cat_boundaries=" + join(d["cat_boundaries"][: d["num_cat"] + 1], lambda x: str(int(x))))
            lines.append("cat_threshold=" + join(d["cat_threshold"], lambda x: str(int(x))))
        lines.append("is_linear=0")
        lines.append("shrinkage=" + g6(self.shrinkage))
        return "\n".join(lines) + "\n\n"


# ---------------------------------------------------------------------------
# Objectives
# ---------------------------------------------------------------------------
class BinaryObjective:
    """objective/binary.rs"""

    def __init__(self, label, weights=None, sigmoid=1.0, scale_pos_weight=1.0, is_unbalance=False):
        self.sigmoid = float(sigmoid)
        self.is_pos = np.asarray(label) > 0
        self.weights = None if weights is None else np.asarray(weights, np.float32)
        cnt_pos = int(self.is_pos.sum())
        cnt_neg = len(self.is_pos) - cnt_pos
        self.need_train = cnt_pos != 0 and cnt_neg != 0
        lw = [1.0, 1.0]
        if is_unbalance and cnt_pos > 0 and cnt_neg > 0:
            if cnt_pos > cnt_neg:
                lw[0] = cnt_pos / cnt_neg
            else:
                lw[1] = cnt_neg / cnt_pos
        lw[1] *= scale_pos_weight
        self.label_weights = np.array(lw)
        self.num_class = 1
        self.const_hess = False

    def name(self):
        return "binary sigmoid:%g" % self.sigmoid

    def gradients(self, score):
        s = self.sigmoid
        lab = np.where(self.is_pos, 1.0, -1.0)
        lw = self.label_weights[self.is_pos.astype(np.int64)]
        response = -lab * s / (1.0 + np.exp(lab * s * score))
        ar = np.abs(response)
        if self.weights is None:
            return (response * lw).astype(np.float32), (ar * (s - ar) * lw).astype(np.float32)
        w = self.weights.astype(np.float64)
        return (response * lw * w).astype(np.float32), (ar * (s - ar) * lw * w).astype(np.float32)

    def boost_from_score(self, k=0):
        if self.weights is None:
            suml = float(np.sum(self.is_pos.astype(np.float64)))
            sumw = float(len(self.is_pos))
        else:
            suml = 0.0
            sumw = 0.0
            for p, w in zip(self.is_pos, self.weights):
                suml += float(np.float32(float(p)) * w)
                sumw += float(w)
        pavg = max(min(suml / sumw, 1.0 - K_EPSILON), K_EPSILON)
        return float(np.log(pavg / (1.0 - pavg)) / self.sigmoid)


class RegressionL2Objective:
    """objective/regression.rs (L2)"""

    def __init__(self, label, weights=None):
        self.label = np.asarray(label, np.float32)
        self.weights = None if weights is None else np.asarray(weights, np.float32)
        self.num_class = 1
        self.const_hess = self.weights is None

    def name(self):
        return "regression"

    def gradients(self, score):
        if self.weights is None:
            return (score - self.label.astype(np.float64)).astype(np.float32), np.ones(len(self.label), np.float32)
        g = (score - self.label.astype(np.float64)).astype(np.float32) * self.weights
        return g.astype(np.float32), self.weights.copy()

    def boost_from_score(self, k=0):
        if self.weights is None:
            s = 0.0
            for v in self.label:
                s += float(v)
            return s / len(self.label)
        suml = sumw = 0.0
        for v, w in zip(self.label, self.weights):
            suml = float(v) * float(w) + suml
            sumw += float(w)
        return suml / sumw


class MulticlassObjective:
    """objective/multiclass.rs (softmax)"""

    def __init__(self, label, num_class, weights=None):
        self.label = np.asarray(label).astype(np.int64)
        self.num_class = int(num_class)
        self.weights = None if weights is None else np.asarray(weights, np.float32)
        self.factor = self.num_class / (self.num_class - 1.0)
        n = len(self.label)
        if self.weights is None:
            cnt = np.bincount(self.label, minlength=self.num_class).astype(np.float64)
            self.init_probs = cnt / n
        else:
            probs = np.zeros(self.num_class)
            for lab, w in zip(self.label, self.weights):
                probs[lab] += float(w)
            self.init_probs = probs / float(np.sum(self.weights.astype(np.float64)))
        self.const_hess = False

    def name(self):
        return "multiclass num_class:%d" % self.num_class

    def gradients(self, score):
        n = len(self.label)
        s = score.reshape(self.num_class, n).T
        m = s.max(axis=1, keepdims=True)
        e = np.exp(s - m)
        p = e / e.sum(axis=1, keepdims=True)
        onehot = np.zeros_like(p)
        onehot[np.arange(n), self.label] = 1.0
        g = p - onehot
        h = self.factor * p * (1.0 - p)
        if self.weights is not None:
            w = self.weights.astype(np.float64)[:, None]
            g = g * w
            h = h * w
        return g.T.reshape(-1).astype(np.float32), h.T.reshape(-1).astype(np.float32)

    def boost_from_score(self, k):
        return float(np.log(max(K_EPSILON, self.init_probs[k])))

    def class_need_train(self, k):
        p = abs(self.init_probs[k])
        return not (p <= K_EPSILON or p >= 1.0 - K_EPSILON)


class LambdarankObjective:
    """LambdarankNdcg (objective/rank.rs:183-361) without position bias.

    Queries are consecutive row ranges (`group_sizes`). The per-pair arithmetic
    follows GetGradientsForOneQuery (rank.rs:263-340) in f64 with f32 lambda and
    hessian accumulators; the pairs of one leading document are evaluated as
    vectors, and that document's own accumulators are summed in pair order
    (np.add.accumulate), so every f32 addition happens in the reference's order.
    exp and log2 are the platform libm's (Python's math module), as in the
    reference (rank.rs:215-223 table, metric/rank.rs:42-47 discounts)."""

    SIGMOID_BINS = 1024 * 1024
    MAX_POSITION = 10_000

    def __init__(self, label, group_sizes, weights=None, sigmoid=1.0, norm=True, truncation_level=30,
                 label_gain=None):
        import math
        if sigmoid <= 0.0:
            raise ValueError(f"Sigmoid param {sigmoid} should be greater than zero")
        self.label = np.asarray(label, np.float32)
        self.weights = None if weights is None else np.asarray(weights, np.float32)
        sizes = np.asarray(group_sizes, np.int64)
        self.bounds = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        if self.bounds[-1] != len(self.label):
            raise ValueError("query sizes do not add up to the number of rows")
        if np.any(sizes > self.MAX_POSITION):
            raise ValueError(f"a query exceeds the upper limit of {self.MAX_POSITION} rows")
        # metric/rank.rs:32-40 default gains 2^i - 1
        self.gain = (np.asarray(label_gain, np.float64) if label_gain
                     else np.array([0.0] + [float((1 << i) - 1) for i in range(1, 31)]))
        lab_i = self.label.astype(np.int64)
        if np.any(np.abs(self.label - lab_i.astype(np.float32)) > K_EPSILON) or np.any(self.label < 0):
            raise ValueError("ranking labels must be non-negative integers")
        if np.any(lab_i >= len(self.gain)):
            raise ValueError("a label is not less than the number of label gains")
        self.discount = np.array([1.0 / math.log2(2.0 + i) for i in range(self.MAX_POSITION)])
        self.sigmoid = float(sigmoid)
        self.norm = bool(norm)
        self.trunc = int(truncation_level)
        # cal_max_dcg over the top min(trunc, cnt) labels (metric/rank.rs:92-121), x86 madd
        k_max = max(self.trunc, 0)
        inv = np.zeros(len(sizes))
        for q in range(len(sizes)):
            lab = np.sort(lab_i[self.bounds[q]:self.bounds[q + 1]])[::-1]
            cur = 0.0
            for j in range(min(k_max, len(lab))):
                cur = self.discount[j] * self.gain[lab[j]] + cur
            inv[q] = 1.0 / cur if cur > 0.0 else cur
        self.inverse_max_dcg = inv
        self.min_in = -50.0 / self.sigmoid / 2.0
        self.max_in = -self.min_in
        self.factor = self.SIGMOID_BINS / (self.max_in - self.min_in)
        self.table = np.array([1.0 / (1.0 + math.exp((i / self.factor + self.min_in) * self.sigmoid))
                               for i in range(self.SIGMOID_BINS)])
        self.num_class = 1
        self.const_hess = False

    def name(self):
        return "lambdarank"

    def boost_from_score(self, k=0):
        return 0.0

    def _sigmoid_of(self, d):
        idx = ((np.clip(d, self.min_in, self.max_in) - self.min_in) * self.factor).astype(np.int64)
        idx = np.minimum(idx, self.SIGMOID_BINS - 1)
        v = self.table[idx]
        v = np.where(d <= self.min_in, self.table[0], v)
        return np.where(d >= self.max_in, self.table[-1], v)

    def _one_query(self, q, label, score, lam, hes):
        cnt = len(label)
        if cnt == 0:
            return
        inv_max_dcg = self.inverse_max_dcg[q]
        order = np.argsort(-score, kind="stable")
        ls = label[order]
        ss = score[order]
        best = ss[0]
        worst_idx = cnt - 1
        if worst_idx > 0 and ss[worst_idx] == -np.inf:
            worst_idx -= 1
        worst = ss[worst_idx]
        sum_lambdas = 0.0
        norm_pairs = self.norm and best != worst
        eps_norm = float(np.float32(0.01))
        neg_sig = -self.sigmoid
        sig2 = self.sigmoid * self.sigmoid
        for i in range(min(cnt - 1, max(self.trunc, 0))):
            if ss[i] == -np.inf:
                continue
            j = np.arange(i + 1, cnt)
            j = j[(ss[j] != -np.inf) & (ls[j] != ls[i])]
            if len(j) == 0:
                continue
            i_high = ls[i] > ls[j]
            high_rank = np.where(i_high, i, j)
            low_rank = np.where(i_high, j, i)
            high = order[high_rank]
            low = order[low_rank]
            delta_score = score[high] - score[low]
            dcg_gap = self.gain[label[high].astype(np.int64)] - self.gain[label[low].astype(np.int64)]
            paired_discount = np.abs(self.discount[high_rank] - self.discount[low_rank])
            dpn = dcg_gap * paired_discount * inv_max_dcg
            if norm_pairs:
                dpn = dpn / (eps_norm + np.abs(delta_score))
            p_lambda = self._sigmoid_of(delta_score)
            p_hessian = p_lambda * (1.0 - p_lambda)
            p_lambda = p_lambda * (neg_sig * dpn)
            p_hessian = p_hessian * (sig2 * dpn)
            pl32 = p_lambda.astype(np.float32)
            ph32 = p_hessian.astype(np.float32)
            # the other document of each pair is touched once for this i
            other = order[j]
            lam[other] = lam[other] + np.where(i_high, -pl32, pl32)
            hes[other] = hes[other] + ph32
            # the leading document accumulates in pair order
            me = order[i]
            lam[me] = np.add.accumulate(np.concatenate([[lam[me]], np.where(i_high, pl32, -pl32)]))[-1]
            hes[me] = np.add.accumulate(np.concatenate([[hes[me]], ph32]))[-1]
            sum_lambdas = np.add.accumulate(np.concatenate([[sum_lambdas], -(2.0 * p_lambda)]))[-1]
        if self.norm and sum_lambdas > 0.0:
            import math
            nf = math.log2(1.0 + sum_lambdas) / sum_lambdas
            lam[:] = (lam.astype(np.float64) * nf).astype(np.float32)
            hes[:] = (hes.astype(np.float64) * nf).astype(np.float32)

    def gradients(self, score):
        n = len(self.label)
        g = np.zeros(n, np.float32)
        h = np.zeros(n, np.float32)
        for q in range(len(self.bounds) - 1):
            lo, hi = int(self.bounds[q]), int(self.bounds[q + 1])
            lam = np.zeros(hi - lo, np.float32)
            hes = np.zeros(hi - lo, np.float32)
            self._one_query(q, self.label[lo:hi], np.asarray(score[lo:hi], np.float64), lam, hes)
            if self.weights is not None:
                lam = lam * self.weights[lo:hi]
                hes = hes * self.weights[lo:hi]
            g[lo:hi] = lam
            h[lo:hi] = hes
        return g, h


# ---------------------------------------------------------------------------
# The boosting loop
# ---------------------------------------------------------------------------
RAND_BLOCK = 1024  # rows per random generator in the row samplers (sample.rs:15-16)


class Random:
    """The LCG of every LightGBM sampler (utils/random.rs:6-80): x = x*214013 + 2531011
    mod 2^32; next_float takes 15 bits, (x >> 16) & 0x7FFF, over 32768 in f32."""

    def __init__(self, seed):
        self.x = int(seed) & 0xFFFFFFFF

    def next_float(self):
        self.x = (self.x * 214013 + 2531011) & 0xFFFFFFFF
        return float(np.float32((self.x >> 16) & 0x7FFF) / np.float32(32768.0))


_SEED_DEFAULTS = {"data_random_seed": 1, "bagging_seed": 3, "drop_seed": 4, "feature_fraction_seed": 2,
                  "objective_seed": 5, "extra_seed": 6}


def resolve_seed(params, name):
    """A sampler seed as Config::set resolves it (config/mod.rs:396-406): an explicit value
    wins; otherwise, when `seed` is given, the seeds are drawn from Random(seed) with
    next_short(0, 32767) in the order data_random_seed, bagging_seed, drop_seed,
    feature_fraction_seed, objective_seed, extra_seed; otherwise the default."""
    if name in params:
        return int(params[name])
    if "seed" not in params:
        return _SEED_DEFAULTS[name]
    x = int(params["seed"]) & 0xFFFFFFFF
    for key in _SEED_DEFAULTS:
        x = (x * 214013 + 2531011) & 0xFFFFFFFF
        if key == name:
            return ((x >> 16) & 0x7FFF) % 32767
    raise KeyError(name)


class BlockDraws:
    """Generators of one row sampler, one per block of 1024 rows seeded seed + block
    (sample.rs:56-60 block_rands), advanced together: draw() gives one next_float per
    row in row order, row i from block i // 1024's generator, and the generators keep
    their state from one call to the next."""

    def __init__(self, seed, n):
        nb = -(-n // RAND_BLOCK)
        self.n = n
        self.x = (np.int64(seed) + np.arange(nb, dtype=np.int64)).astype(np.uint64) & np.uint64(0xFFFFFFFF)
        self.count = np.minimum(RAND_BLOCK, n - np.arange(nb) * RAND_BLOCK)

    def draw(self):
        out = np.zeros((len(self.x), RAND_BLOCK), np.float32)
        for j in range(RAND_BLOCK):
            live = self.count > j
            if not live.any():
                break
            x = (self.x[live] * np.uint64(214013) + np.uint64(2531011)) & np.uint64(0xFFFFFFFF)
            self.x[live] = x
            out[live, j] = ((x >> np.uint64(16)) & np.uint64(0x7FFF)).astype(np.float32) / np.float32(32768.0)
        return out.reshape(-1)[: self.n]


class RowSampler:
    """SampleStrategy (sample.rs:47-242): all rows, bagging (plain or balanced by the
    binary label) or GOSS. sample() returns the ascending rows a tree is trained on, or
    None for all rows, and may rescale the gradients in place (GOSS)."""

    def __init__(self, params, label, num_tree_per_iteration=1):
        p = params
        self.n = len(label)
        self.kind = "full"
        strategy = p.get("data_sample_strategy", "bagging")
        seed = resolve_seed(p, "bagging_seed")
        freq = p.get("bagging_freq", 0)
        fraction = p.get("bagging_fraction", 1.0)
        if strategy == "goss":
            if freq > 0 and fraction != 1.0:
                raise ValueError("Cannot use bagging in GOSS")
            self.top_rate = p.get("top_rate", 0.2)
            self.other_rate = p.get("other_rate", 0.1)
            if self.top_rate + self.other_rate > 1.0 or self.top_rate <= 0.0 or self.other_rate <= 0.0:
                raise ValueError("top_rate + other_rate must be <= 1, both positive")
            self.kind = "goss"
            self.warmup = int(1.0 / p.get("learning_rate", 0.1))
            self.rands = [Random(seed + b) for b in range(-(-self.n // RAND_BLOCK))]
            self.ntpi = num_tree_per_iteration
            return
        if strategy != "bagging":
            raise ValueError(f"data_sample_strategy={strategy}")
        if p.get("bagging_by_query", False):
            raise ValueError("bagging_by_query is not supported")
        positive = np.asarray(label) > 0
        num_pos = int(positive.sum()) if p.get("objective") == "binary" else 0
        pos_f, neg_f = p.get("pos_bagging_fraction", 1.0), p.get("neg_bagging_fraction", 1.0)
        balanced = (pos_f < 1.0 or neg_f < 1.0) and num_pos > 0
        if not ((fraction < 1.0 or balanced) and freq > 0):
            return
        self.kind = "bagging"
        self.freq = freq
        self.threshold = np.where(positive, pos_f, neg_f) if balanced else np.full(self.n, fraction)
        self.draws = BlockDraws(seed, self.n)
        self.bag = np.zeros(0, np.int32)
        self.need_rebag = True

    @property
    def changes_hessians(self):
        return self.kind == "goss"

    def _bagging_draw(self):
        # BaggingHelper / BalancedBaggingHelper: one draw per row, kept when below the
        # row's fraction (sample.rs:203-234), compared in f64
        r = self.draws.draw().astype(np.float64)
        return np.flatnonzero(r < self.threshold).astype(np.int32)

    def _goss_draw(self, g, h):
        # GOSSStrategy::Helper over all rows (sample.rs:244-296)
        n, ntpi = self.n, self.ntpi
        if n == 0:
            return np.zeros(0, np.int32)
        gk = g.reshape(ntpi, n)
        hk = h.reshape(ntpi, n)
        tmp = np.zeros(n, np.float32)
        for k in range(ntpi):
            tmp = (tmp + np.abs(gk[k] * hk[k])).astype(np.float32)
        top_k = max(int(n * self.top_rate), 1)
        other_k = int(n * self.other_rate)
        threshold = np.partition(tmp, n - top_k)[n - top_k]  # the top_k-th largest
        multiply = np.float32(n - top_k) / np.float32(other_k)
        bag = []
        big = 0
        for i in range(n):
            if tmp[i] >= threshold:
                bag.append(i)
                big += 1
            else:
                rest_need = other_k - (len(bag) - big)
                rest_all = (n - i) - (top_k - big)
                prob = rest_need / rest_all
                if self.rands[i // RAND_BLOCK].next_float() < prob:
                    bag.append(i)
                    for k in range(ntpi):
                        gk[k, i] = np.float32(gk[k, i] * multiply)
                        hk[k, i] = np.float32(hk[k, i] * multiply)
        return np.asarray(bag, np.int32)

    def sample(self, it, g, h):
        if self.kind == "full":
            return None
        if self.kind == "goss":
            if it < self.warmup:
                return None
            bag = self._goss_draw(g, h)
            return bag if len(bag) < self.n else None
        if (len(self.bag) < self.n and it % self.freq == 0) or self.need_rebag:
            self.need_rebag = False
            self.bag = self._bagging_draw()
        return self.bag if len(self.bag) < self.n else None


def out_of_bag(rows, n):
    """Rows not in the bag, ascending (train.rs:1282-1291)."""
    if rows is None:
        return None
    mask = np.ones(n, bool)
    mask[rows] = False
    return np.flatnonzero(mask).astype(np.int32)


class GBDT:
    """Trainer::train_one_iter_with (train.rs:723-839) for plain GBDT."""

    def __init__(self, learners, objective, num_data, learning_rate=0.1, boost_from_average=True, sampler=None):
        # one learner per class tree of an iteration may share state: the reference
        # uses a single learner; `learners` is a list of length 1 or num_class
        self.learners = learners
        self.sampler = sampler
        self.obj = objective
        self.n = num_data
        self.lr = learning_rate
        self.boost_from_average = boost_from_average
        self.ntpi = objective.num_class
        self.score = np.zeros(self.n * self.ntpi)
        self.models = []
        self.iter = 0

    def _learner(self, k):
        return self.learners[k if len(self.learners) > 1 else 0]

    def _automatic_init(self, k):
        if not self.models and self.boost_from_average:
            init = self.obj.boost_from_score(k)
            if abs(init) > K_EPSILON:
                return init
        return 0.0

    def train_one_iter(self):
        n = self.n
        inits = []
        for k in range(self.ntpi):
            init = self._automatic_init(k)
            if init != 0.0:
                self.score[k * n:(k + 1) * n] += init
            inits.append(init)
        g, h = self.obj.gradients(self.score)
        # the gradients come first and the row sample after (train.rs:746-769)
        rows = self.sampler.sample(self.iter, g, h) if self.sampler is not None else None
        oob = out_of_bag(rows, n)
        should_continue = False
        for k in range(self.ntpi):
            off = k * n
            need = getattr(self.obj, "class_need_train", lambda _k: True)(k)
            if need:
                d = (self._learner(k).train(g[off:off + n], h[off:off + n]) if rows is None else
                     self._learner(k).train(g[off:off + n], h[off:off + n], rows))
                tree = Tree(d)
            else:
                tree = Tree.constant(0.0, n)
            if tree.num_leaves > 1:
                should_continue = True
                tree.apply_shrinkage(self.lr)
                view = self.score[off:off + n]
                self._learner(k).add_prediction_to_score(tree.leaf_value, view)
                if oob is not None:
                    self._learner(k).add_out_of_bag_prediction(tree.leaf_value, oob, view)
                self.score[off:off + n] = view
                if abs(inits[k]) > K_EPSILON:
                    tree.add_bias(inits[k])
            elif len(self.models) < self.ntpi:
                if not self.boost_from_average:
                    inits[k] = self.obj.boost_from_score(k)
                    self.score[off:off + n] += inits[k]
                tree = Tree.constant(inits[k], n)
            else:
                tree = Tree.constant(0.0, n)
            self.models.append(tree)
        self.iter += 1
        if not should_continue:
            if len(self.models) > self.ntpi:
                del self.models[-self.ntpi:]
            return True
        return False

    def train(self, num_iterations):
        for _ in range(num_iterations):
            if self.train_one_iter():
                break
        return self


def model_to_string(models, num_class, ntpi, objective_name, feature_names, feature_infos):
    """Gbdt::save_model_to_string (model_text.rs:88-138), trees only."""
    head = ["tree", "version=v4", f"num_class={num_class}", f"num_tree_per_iteration={ntpi}", "label_index=0",
            f"max_feature_idx={len(feature_names) - 1}", f"objective={objective_name}",
            "feature_names=" + " ".join(feature_names), "feature_infos=" + " ".join(feature_infos)]
    blocks = [f"Tree={i}\n{t.to_model_string()}" for i, t in enumerate(models)]
    head.append("tree_sizes=" + " ".join(str(len(b.encode())) for b in blocks))
    return "\n".join(head) + "\n\n" + "".join(blocks) + "end of trees\n"


def parse_model_trees(text):
    """Tree blocks of a model text file as dicts of strings (for comparisons)."""
    trees = []
    cur = None
    for line in text.splitlines():
        if line.startswith("Tree="):
            cur = {}
            trees.append(cur)
        elif line.startswith("end of trees"):
            break
        elif cur is not None and "=" in line:
            k, v = line.split("=", 1)
            cur[k] = v
    return trees


def parse_model_header(text):
    out = {}
    for line in text.splitlines():
        if line.startswith("Tree="):
            break
        if "=" in line:
            k, v = line.split("=", 1)
            out[k] = v
    return out


# ---------------------------------------------------------------------------
# Prediction from model text and the AUC metric
# ---------------------------------------------------------------------------
def parse_model(text):
    """Gbdt::from_model_string's tree part (model_text.rs:88-167): the header's
    num_tree_per_iteration and every tree as arrays (split_feature, threshold,
    decision_type, children, leaf_value and the categorical bitsets)."""
    header = parse_model_header(text)
    trees = []
    for b in parse_model_trees(text):
        n = int(b["num_leaves"])
        arr = lambda k, dt: np.array(b[k].split(), dt) if b.get(k) else np.zeros(0, dt)  # noqa: E731
        t = {"num_leaves": n, "num_cat": int(b.get("num_cat", 0)), "leaf_value": arr("leaf_value", np.float64)}
        if n > 1:
            t.update(split_feature=arr("split_feature", np.int64), threshold=arr("threshold", np.float64),
                     decision_type=arr("decision_type", np.int64), left_child=arr("left_child", np.int64),
                     right_child=arr("right_child", np.int64))
        t["cat_boundaries"] = arr("cat_boundaries", np.int64) if t["num_cat"] > 0 else np.zeros(1, np.int64)
        t["cat_threshold"] = arr("cat_threshold", np.int64) if t["num_cat"] > 0 else np.zeros(0, np.int64)
        if b.get("is_linear", "0") != "0":
            raise NotImplementedError("linear trees are not predicted here")
        trees.append(t)
    return int(header.get("num_tree_per_iteration", 1)), trees


def _leaf_of_rows(t, x, torch):
    """Tree::get_leaf (tree.rs:718-783) for every row of x (a float64 tensor,
    rows x features): numerical_decision with the missing-value rules and
    categorical_decision with the bitsets, level by level over the active rows."""
    dev = x.device
    n = x.shape[0]
    leaf = torch.zeros(n, dtype=torch.int64, device=dev)
    if t["num_leaves"] <= 1:
        return leaf
    asT = lambda a, dt: torch.as_tensor(np.asarray(a), dtype=dt, device=dev)  # noqa: E731
    sf, th = asT(t["split_feature"], torch.int64), asT(t["threshold"], torch.float64)
    dt = asT(t["decision_type"], torch.int64)
    lc, rc = asT(t["left_child"], torch.int64), asT(t["right_child"], torch.int64)
    cb, ct = asT(t["cat_boundaries"], torch.int64), asT(t["cat_threshold"], torch.int64)
    rows = torch.arange(n, device=dev)
    node = torch.zeros(n, dtype=torch.int64, device=dev)
    while rows.numel() > 0:
        nd = node
        fval = x[rows, sf[nd]]
        d = dt[nd]
        miss = (d >> 2) & 3
        default_left = (d & 2) != 0
        nan = torch.isnan(fval)
        # numerical: NaN reads as 0 unless NaN is the missing value
        fnum = torch.where(nan & (miss != 2), torch.zeros_like(fval), fval)
        is_zero = (fnum >= -K_ZERO_THRESHOLD) & (fnum <= K_ZERO_THRESHOLD)
        missing = ((miss == 1) & is_zero) | ((miss == 2) & nan)
        go_left = torch.where(missing, default_left, fnum <= th[nd])
        cat = (d & 1) != 0
        if bool(cat.any()):
            # C++ static_cast<int>: truncation toward zero; NaN and negatives go right
            iv = torch.where(nan, torch.full_like(fval, -1.0), fval).clamp(-1.0, 2.0 ** 31 - 1).trunc().to(torch.int64)
            ci = th[nd].to(torch.int64).clamp(min=0, max=max(cb.numel() - 2, 0))
            lo, hi = cb[ci], cb[ci + 1]
            word = lo + torch.div(iv.clamp(min=0), 32, rounding_mode="floor")
            inside = (iv >= 0) & (word < hi)
            bits = ct[word.clamp(max=max(ct.numel() - 1, 0))] if ct.numel() > 0 else torch.zeros_like(word)
            in_set = inside & (((bits >> (iv.clamp(min=0) % 32)) & 1) == 1)
            go_left = torch.where(cat, in_set, go_left)
        child = torch.where(go_left, lc[nd], rc[nd])
        done = child < 0
        leaf[rows[done]] = ~child[done]
        keep = ~done
        rows = rows[keep]
        node = child[keep]
    return leaf


def predict_raw(trees, x, num_tree_per_iteration=1):
    """Raw scores (the predictor's RawScore kind, predictor.rs:162-196): for each
    row and class k, the sum over iterations of tree (i * K + k)'s leaf value, in
    tree order from 0.0. x: a numpy array (computed with torch on the CPU) or a
    torch tensor (computed on its device). trees: dicts as parse_model returns
    them or as the learner's train() returns them. Returns float64, shape
    (rows,) for one class, else (rows, K)."""
    import torch

    as_numpy = isinstance(x, np.ndarray)
    xt = torch.as_tensor(np.ascontiguousarray(x, np.float64)) if as_numpy else x.to(torch.float64)
    K = int(num_tree_per_iteration)
    out = torch.zeros((xt.shape[0], K), dtype=torch.float64, device=xt.device)
    for i, t in enumerate(trees):
        vals = torch.as_tensor(np.asarray(t["leaf_value"][: max(t["num_leaves"], 1)], np.float64), device=xt.device)
        out[:, i % K] += vals[_leaf_of_rows(t, xt, torch)]
    out = out[:, 0] if K == 1 else out
    return out.cpu().numpy() if as_numpy else out


def auc(label, score, weights=None):
    """AUCMetric (metric/binary.rs:73-135): rows by descending score (ties keep row
    order), positives and negatives accumulated per distinct score,
    accum = madd(cur_neg, madd(cur_pos, 0.5, sum_pos), accum) (unfused, as
    utils/common.rs:207-216 on x86_64); 1.0 when one class is absent.

    Per distinct score the sums are taken in row order as the reference's loop
    does; the running sums over the groups are sequential (np.add.accumulate), so
    the result has the reference's rounding. Unweighted group sums are counts,
    exact in any order; weighted ones run the reference's loop literally."""
    label = np.asarray(label)
    score = np.asarray(score, np.float64)
    n = len(label)
    order = np.argsort(-score, kind="stable")
    s = score[order]
    pos = label[order] > 0
    if weights is None:
        sum_w = float(n)
        starts = np.flatnonzero(np.concatenate([[True], s[1:] != s[:-1]]))
        cnt_pos = np.add.reduceat(pos.astype(np.float64), starts)
        cnt_all = np.diff(np.concatenate([starts, [n]])).astype(np.float64)
        cnt_neg = cnt_all - cnt_pos
        sum_pos_before = np.concatenate([[0.0], np.add.accumulate(cnt_pos)[:-1]])
        inner = cnt_pos * 0.5 + sum_pos_before
        accum = float(np.add.accumulate(cnt_neg * inner)[-1]) if n else 0.0
        sum_pos = float(np.add.accumulate(cnt_pos)[-1]) if n else 0.0
    else:
        w = np.asarray(weights, np.float32)[order].astype(np.float64)
        sum_w = 0.0
        for v in np.asarray(weights, np.float32).astype(np.float64):
            sum_w += float(v)
        cur_pos = sum_pos = accum = cur_neg = 0.0
        thr = s[0] if n else 0.0
        for i in range(n):
            if s[i] != thr:
                thr = s[i]
                accum = cur_neg * (cur_pos * 0.5 + sum_pos) + accum
                sum_pos += cur_pos
                cur_neg = cur_pos = 0.0
            if pos[i]:
                cur_pos += w[i]
            else:
                cur_neg += w[i]
        accum = cur_neg * (cur_pos * 0.5 + sum_pos) + accum
        sum_pos += cur_pos
    if sum_pos > 0.0 and sum_pos != sum_w:
        return accum / (sum_pos * (sum_w - sum_pos))
    return 1.0
