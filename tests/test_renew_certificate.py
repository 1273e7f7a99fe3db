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


# The piecewise chain (k_renew_chain): the leaf's values in row order are cut into
# pieces; a piece is added as one exact sum when, from the chain's running sum s,
# |s| + sum|x| < 2^(53 + E) with E the lowest set bit over s and the piece's values,
# and value by value otherwise. Its result must be the sequential sum's, bit for bit.

def lowest_bit(x):
    """Exponent of the lowest set bit of a finite nonzero double (None for zero)."""
    if x == 0.0:
        return None
    f = Fraction(float(x))
    num, den = abs(f.numerator), f.denominator
    e = -(den.bit_length() - 1)  # den is a power of two
    while num % 2 == 0:
        num //= 2
        e += 1
    return e


def piecewise_sum(vals, rows, rng):
    """k_renew_chain's rule; the piece sums are taken in a random order, as atomics do."""
    s, fast, slow = 0.0, 0, 0
    for p0 in range(0, len(vals), rows):
        piece = [float(v) for v in vals[p0:p0 + rows] if v != 0.0]
        if not piece:
            continue
        order = rng.permutation(len(piece))
        t = 0.0
        a = 0.0
        for i in order:
            t += piece[i]
            a += abs(piece[i])
        es = [lowest_bit(v) for v in piece] + ([lowest_bit(s)] if s != 0.0 else [])
        e = min(es)
        if abs(s) + a * (1.0 + 2.0 ** -30) < math.ldexp(1.0, 53 + e):
            # every partial sum is exact, so the piece's sum is too, in any order
            assert Fraction(t) == sum((Fraction(v) for v in piece), Fraction(0))
            s += t
            fast += 1
        else:
            for v in piece:
                s += v
            slow += 1
    return s, fast, slow


def test_piecewise_chain_equals_the_sequential_sum():
    rng = np.random.default_rng(7)
    fast = slow = 0
    for trial in range(400):
        n = int(rng.integers(1, 2500))
        rows = int(rng.choice([1, 7, 64, 256, 4096]))
        lo, hi = sorted(rng.uniform(-12, 3, 2))
        vals = rng.normal(size=n) * 10.0 ** rng.uniform(lo, hi, n)
        if trial % 4 == 1:  # one sign: the running sum grows past the small values' bits
            vals = np.abs(vals)
        if trial % 4 == 2:  # subnormal floats among normal ones
            vals[rng.integers(0, n, max(1, n // 10))] = rng.choice([1e-45, -3e-45, 2.0 ** -130])
        if trial % 4 == 3:  # a large value then many tiny ones, and cancellation
            vals[0] = 1.0e6 * (1 if trial % 8 == 3 else -1)
            vals[n // 2] = -vals[0]
        vals = vals.astype(np.float32)
        got, f, sl = piecewise_sum(vals, rows, rng)
        want = seq_sum(vals)
        assert got == want and math.copysign(1.0, got) == math.copysign(1.0, want), trial
        fast += f
        slow += sl
    # both branches ran often
    assert fast > 1000 and slow > 100


def test_piecewise_chain_on_late_boosting_gradients():
    """One leaf of a confident model: its rows' scores are close (a leaf groups
    similar rows), so most gradients are small and a few near -1; the whole-leaf
    certificate fails, pieces of a few hundred rows pass."""
    rng = np.random.default_rng(3)
    n = 60_000
    p = 1.0 / (1.0 + np.exp(-rng.normal(-9.0, 1.5, n)))
    y = rng.random(n) < p
    g = (p - y).astype(np.float32)
    h = (p * (1.0 - p)).astype(np.float32)
    for vals in (g, h):
        assert not certified(vals)
        got, fast, slow = piecewise_sum(vals, 512, rng)
        assert got == seq_sum(vals)
        assert fast > 10 * slow


# The binade summary (k_renew_binade, k_renew_chain): while s and every partial sum
# stay strictly inside s's binade [2^eb, 2^(eb+1)), fl(s + x) = (k + q + round(f)) u
# with s = k u, u = 2^(eb-52), x/u = q + f, and a tie f = 1/2 goes to the even one;
# the piece's units depend on s only through the parity of k.

def parity_summary(vals, eb):
    """(units added when k is even, units added when k is odd), composed in order."""
    inv_u = math.ldexp(1.0, 52 - eb)
    s0, s1, o0, o1 = 0, 0, 0, 1
    for x in vals:
        y = float(x) * inv_u
        q = math.floor(y)
        f = y - q
        if f == 0.5:
            m0, m1, p0, p1 = q + (q & 1), q + ((q & 1) ^ 1), 0, 0
        else:
            a = q + (1 if f > 0.5 else 0)
            m0, m1, p0, p1 = a, a, a & 1, (a & 1) ^ 1
        s0, s1, o0, o1 = s0 + (m1 if o0 else m0), s1 + (m1 if o1 else m0), (p1 if o0 else p0), (p1 if o1 else p0)
    return s0, s1


def in_binade(s, a, eb, rows):
    slack = a * (1.0 + 2.0 ** -30) + rows * math.ldexp(1.0, eb - 52)
    return abs(s) - slack > math.ldexp(1.0, eb) and abs(s) + slack < math.ldexp(1.0, eb + 1)


def test_binade_summary_equals_the_sequential_sum():
    rng = np.random.default_rng(11)
    checked = ties_seen = 0
    for trial in range(600):
        eb = int(rng.integers(-30, 40))
        u = math.ldexp(1.0, eb - 52)
        s = float(rng.uniform(1.0, 2.0)) * math.ldexp(1.0, eb) * (1 if trial % 3 else -1)
        s = round(s / u) * u  # a double of that binade
        n = int(rng.integers(1, 300))
        mag = 10.0 ** rng.uniform(-18, -6, n) * math.ldexp(1.0, eb)
        vals = rng.normal(size=n) * mag
        if trial % 2 == 0:
            vals = np.abs(vals)
        vals = vals.astype(np.float32)
        # ties: values that are an odd multiple of u/2, exactly as float32
        k = rng.integers(0, n, n // 4 + 1)
        odd = (2 * rng.integers(-(2 ** 20), 2 ** 20, len(k)) + 1).astype(np.float64)
        vals[k] = (odd * (u / 2)).astype(np.float32)
        vals[k] = np.where(vals[k].astype(np.float64) == odd * (u / 2), vals[k], np.float32(0.0))
        a = float(np.sum(np.abs(vals.astype(np.float64))))
        if not in_binade(s, a, eb, 4096):
            continue
        checked += 1
        ties_seen += int(np.count_nonzero(np.abs(vals.astype(np.float64) / u % 1.0 - 0.5) == 0.0))
        want = s
        for v in vals:
            want += float(v)
        s0, s1 = parity_summary(vals, eb)
        kk = int(s / u)
        got = (kk + (s1 if kk & 1 else s0)) * u
        assert got == want, trial
    assert checked > 300 and ties_seen > 1000
