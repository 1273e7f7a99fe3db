"""Seeded synthetic workloads of the benchmark configurations.

binary_28: n rows x 28 dense f32 features drawn i.i.d. N(0,1); w ~ N(0, 0.5);
    z = X.w + 0.5 x0 x1 - 0.3 x2^2 (+ an intercept); y ~ Bernoulli(sigmoid(z)).
    The 10k-row case uses no intercept; the 10.5M-row case adds the intercept that
    puts the positive rate at 53%, solved by bisection on a 1M-row pilot drawn
    from its own seed stream.
regression_100: 60 N(0,1), 30 U(0,1), 10 lognormal(0,1) columns with 5% NaN;
    y = sum of 10 random depth-3 step functions + N(0, 0.1).
multiclass_50: 40 N(0,1) columns and 10 categorical columns uniform over {0..31};
    y = argmax over 10 classes of X_dense.W_k + Gumbel noise.
ranking_136: 68 N(0,1) and 68 U(0,1) columns with 20% exact zeros; query sizes
    floor(lognormal(4, 1)) clipped to [5, 1000]; relevance 0-4 from a noisy linear
    score cut at its 50/75/90/97% quantiles (skewed toward 0).
"""
import numpy as np

NUM_FEATURES_BINARY = 28


def _weights(seed):
    return np.random.default_rng([seed, 1]).normal(0.0, 0.5, size=NUM_FEATURES_BINARY)


def _z(x, w):
    return x.astype(np.float64) @ w + 0.5 * x[:, 0].astype(np.float64) * x[:, 1] - 0.3 * x[:, 2].astype(np.float64) ** 2


def _features(n, seed, stream, chunk=1 << 20):
    """n x 28 N(0,1) float32, generated in chunks from a per-stream generator."""
    rng = np.random.default_rng([seed, stream])
    out = np.empty((n, NUM_FEATURES_BINARY), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = rng.standard_normal(size=(e - s, NUM_FEATURES_BINARY), dtype=np.float32)
    return out


def solve_intercept(seed, target=0.53, pilot=1_000_000):
    w = _weights(seed)
    z = _z(_features(pilot, seed, 9), w)
    lo, hi = -5.0, 5.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if np.mean(1.0 / (1.0 + np.exp(-(z + mid)))) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def binary_28(n, seed=0, intercept=0.0, stream=2):
    """(X float32 (n, 28), y float64 0/1)."""
    w = _weights(seed)
    x = _features(n, seed, stream)
    z = _z(x, w) + intercept
    p = 1.0 / (1.0 + np.exp(-z))
    u = np.random.default_rng([seed, stream, 7]).uniform(size=n)
    return x, (u < p).astype(np.float64)


def regression_100(n, seed=0):
    rng = np.random.default_rng([seed, 3])
    x = np.empty((n, 100), np.float32)
    x[:, :60] = rng.standard_normal(size=(n, 60), dtype=np.float32)
    x[:, 60:90] = rng.uniform(size=(n, 30)).astype(np.float32)
    ln = rng.lognormal(0.0, 1.0, size=(n, 10)).astype(np.float32)
    ln[rng.uniform(size=(n, 10)) < 0.05] = np.nan
    x[:, 90:] = ln
    y = np.zeros(n)
    trng = np.random.default_rng([seed, 4])
    for _ in range(10):
        # a random depth-3 step function: 7 internal nodes, 8 leaves
        feats = trng.integers(0, 90, size=7)
        ths = [float(np.quantile(x[: min(n, 100000), f], trng.uniform(0.2, 0.8))) for f in feats]
        vals = trng.normal(size=8)
        node = np.zeros(n, np.int64)
        for depth in range(3):
            f = feats[node]
            t = np.asarray(ths)[node]
            go_right = x[np.arange(n), f] > t
            node = 2 * node + 1 + go_right
        y += vals[node - 7]
    y += trng.normal(0.0, 0.1, size=n)
    return x, y


MULTICLASS_CATEGORICAL = tuple(range(40, 50))


def multiclass_50(n, seed=0, num_class=10):
    rng = np.random.default_rng([seed, 5])
    x = np.empty((n, 50), np.float32)
    x[:, :40] = rng.standard_normal(size=(n, 40), dtype=np.float32)
    x[:, 40:] = rng.integers(0, 32, size=(n, 10)).astype(np.float32)
    w = np.random.default_rng([seed, 6]).normal(size=(40, num_class))
    scores = x[:, :40].astype(np.float64) @ w + rng.gumbel(size=(n, num_class))
    return x, np.argmax(scores, axis=1).astype(np.float64)


def ranking_136(n, seed=0):
    """Returns x, labels, query sizes (the last query is cut to make n rows)."""
    rng = np.random.default_rng([seed, 8])
    sizes = []
    total = 0
    while total < n:
        q = int(np.clip(np.floor(rng.lognormal(4.0, 1.0)), 5, 1000))
        q = min(q, n - total)
        sizes.append(q)
        total += q
    x = np.empty((n, 136), np.float32)
    x[:, :68] = rng.standard_normal(size=(n, 68), dtype=np.float32)
    u = rng.uniform(size=(n, 68)).astype(np.float32)
    u[rng.uniform(size=(n, 68)) < 0.2] = 0.0
    x[:, 68:] = u
    w = np.random.default_rng([seed, 9]).normal(size=136) / np.sqrt(136.0)
    s = x.astype(np.float64) @ w + rng.normal(0.0, 0.5, size=n)
    cuts = np.quantile(s, [0.5, 0.75, 0.9, 0.97])
    return x, np.searchsorted(cuts, s).astype(np.float64), np.asarray(sizes, np.int32)
