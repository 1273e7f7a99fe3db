"""The leaf-renewal certificate used by the GPU learner.

The reference sums a leaf's f32 gradients in f64 one addition at a time in row
order (serial.rs:634-655). The GPU sums them in parallel and claims equality
when n * max|x| < 2^(52+E), E being the exponent of the lowest significand bit
over the leaf's nonzero values. These tests check that claim against exact
rational sums and against sequential sums in other orders, and that the
certificate really does reject cases where the order matters.
"""
import math
from fractions import Fraction

import numpy as np


def lsb_exponent(x):
    """Exponent of the lowest significand bit of a float32 (None for zero)."""
    b = int(np.float32(x).view(np.uint32)) & 0x7FFFFFFF
    if b == 0:
        return None
    be = b >> 23
    assert be != 0xFF
    return -149 if be == 0 else be - 150


def certified(vals):
    es = [e for e in (lsb_exponent(v) for v in vals) if e is not None]
    if not es:
        return True
    e = min(es)
    return len(vals) * float(np.max(np.abs(vals.astype(np.float64)))) < math.ldexp(1.0, 52 + e)


def seq_sum(vals):
    acc = 0.0
    for v in vals:
        acc += float(v)
    return acc


def test_lsb_exponent_known_values():
    assert lsb_exponent(np.float32(1.0)) == -23
    assert lsb_exponent(np.float32(0.5)) == -24
    assert lsb_exponent(np.float32(3.0)) == -22
    assert lsb_exponent(np.float32(0.0)) is None
    assert lsb_exponent(np.float32(-0.0)) is None
    assert lsb_exponent(np.float32(1e-45)) == -149  # smallest denormal
    assert lsb_exponent(np.float32(2.0 ** -126)) == -149  # smallest normal: lsb is 2^-149


def test_certified_sums_are_exact_in_any_order():
    rng = np.random.default_rng(0)
    hits = 0
    for trial in range(300):
        n = int(rng.integers(1, 3000))
        lo, hi = sorted(rng.uniform(-8, 2, 2))
        vals = (rng.normal(size=n) * 10.0 ** rng.uniform(lo, hi, n)).astype(np.float32)
        if not certified(vals):
            continue
        hits += 1
        exact = sum((Fraction(float(v)) for v in vals), Fraction(0))
        s = seq_sum(vals)
        assert Fraction(s) == exact, trial
        assert seq_sum(vals[::-1]) == s
        assert seq_sum(rng.permutation(vals)) == s
        # pairwise (tree) order, like a parallel reduction
        level = [float(v) for v in vals]
        while len(level) > 1:
            level = [level[i] + level[i + 1] if i + 1 < len(level) else level[i] for i in range(0, len(level), 2)]
        assert level[0] == s
    assert hits > 50


def test_certificate_rejects_order_dependent_sums():
    # one large value next to many tiny ones: the sequential sum rounds
    vals = np.array([1.0e4] + [1.0e-7] * 1000, np.float32)
    assert not certified(vals)
    assert seq_sum(vals) != seq_sum(vals[::-1])
    # gradients of a fresh binary model (|g| ~ 0.5, h ~ 0.25) certify for very large leaves
    g = np.where(np.arange(10_000_000) % 2 == 0, 0.47, -0.53).astype(np.float32)
    assert certified(g[:1000])
    e = min(lsb_exponent(v) for v in (np.float32(0.47), np.float32(-0.53)))
    assert 10_500_000 * 0.53 < math.ldexp(1.0, 52 + e)
