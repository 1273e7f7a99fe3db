"""Pins the oracle's primitives against the reference's own known-answer vectors
and against independent brute force.

Sources of the expected values:
  - MT19937: ISO C++ [rand.predef] 10000th draw, and the libc++ distribution vectors
    of crates/lgbm-core/tests/mt19937.rs:14-75;
  - discretizer: crates/lgbm-core/tests/quantized_gradients.rs:100-158 and the 16
    upstream stumps of tests/fixtures/quantized_gradients.json (copied under tests/golden/);
  - histogram: the ordered per-bin oracle of crates/lgbm-core/tests/simd_histogram.rs:8-54;
  - integer split search: the independent brute force of
    crates/lgbm-core/tests/quantized_split_search.rs:50-223, restated below.
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle_binding as ob

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EPS = float(np.float32(1e-15))


def test_mt19937_standard_vector():
    e = ob.Mt19937()
    v = 0
    for _ in range(10000):
        v = e.next()
    assert v == 4123659995


def test_mt19937_default_seed():
    a, b = ob.Mt19937(), ob.Mt19937(5489)
    assert [a.next() for _ in range(1000)] == [b.next() for _ in range(1000)]


def test_uniform_real_matches_libcxx_bits():
    expected = [0x3fe97d47b66bfc3c, 0x3fc77aca8779b102, 0x3fe8f33a88f76c7f, 0x3fe3196582932836,
                0x3fdc88861e49fc36]
    e = ob.Mt19937(42)
    got = [struct.unpack("<Q", struct.pack("<d", e.uniform_real()))[0] for _ in expected]
    assert got == expected


def test_uniform_int_matches_libcxx():
    e = ob.Mt19937(7)
    assert [e.uniform_int(1000) for _ in range(8)] == [175, 196, 537, 502, 579, 211, 919, 615]
    e = ob.Mt19937(123)
    assert [e.uniform_int(255) for _ in range(4)] == [254, 109, 126, 66]
    e = ob.Mt19937(123)
    assert [e.uniform_int(300) for _ in range(4)] == [98, 230, 17, 83]


def test_nearest_discretization_known_answers():
    r = ob.discretize([-2.0, -0.5, 0.0, 0.5, 2.0], [1.0] * 5, 4)
    assert list(r[0]) == [-508, -252, 4, 260, 516]
    assert (r[1], r[2]) == (1.0, 0.25)
    r = ob.discretize([-4.0, -1.0, 0.0, 1.0, 4.0], [2.0] * 5, 4)
    assert list(r[0]) == [-508, -252, 4, 260, 516]
    assert (r[1], r[2]) == (2.0, 0.5)


def test_discretization_refusals():
    assert ob.discretize([np.nan, 1.0], [1.0, 1.0], 4) is None
    assert ob.discretize([1.0, 1.0], [-1.0, 1.0], 4) is None
    assert ob.discretize([0.0, 0.0], [1.0, 1.0], 4) is None
    assert ob.discretize([-2.0, 2.0], [0.0, 0.0], 4) is None


def test_constant_hessian_discretization():
    r = ob.discretize([-2.0, 2.0], [0.0, 0.0], 4, const_hess=True)
    assert list(r[0]) == [-511, 513]
    assert (r[1], r[2]) == (1.0, 0.0)


def _close(a, b):
    assert abs(a - b) <= 1e-12 * max(abs(b), 1.0), (a, b)


def test_quantized_stumps_match_upstream():
    f = json.load(open(os.path.join(GOLDEN, "quantized_gradients.json")))
    assert f["upstream"] == "4.7.0"
    assert len(f["cases"]) == 16
    for c in f["cases"]:
        g = np.array(c["gradients"], np.float32)
        h = np.array(c["hessians"], np.float32)
        packed, gs, hs = ob.discretize(g, h, c["bins"], const_hess=c["constant_hessian"])
        left = np.zeros(len(g), bool)
        left[c["left_rows"]] = True
        qg = (packed >> 8).astype(np.int64)
        qh = (packed & 0xff).astype(np.int64)
        lg, lh = qg[left].sum() * gs, qh[left].sum() * hs
        rg, rh = qg[~left].sum() * gs, qh[~left].sum() * hs
        _close(lh, c["left_weight"])
        _close(rh, c["right_weight"])
        _close(-lg / lh, c["left_output"])
        _close(-rg / rh, c["right_output"])
        gain = lg * lg / lh + rg * rg / rh - (lg + rg) ** 2 / (lh + rh)
        _close(float(np.float32(gain)), c["split_gain"])


def test_stochastic_noise_layout_and_bounds():
    # block layout: BlockInfo(num_threads, N, 512) with 32-aligned blocks, block b
    # seeded seed+b (gradients) and seed+b+num_threads (hessians)
    n, seed, nt = 5000, 11, 3
    gn, hn = ob.noise(n, seed, nt)
    nb = min(nt, (n + 511) // 512)
    size = ((n + nb - 1) // nb + 31) // 32 * 32
    for b in range(nb):
        eg, eh = ob.Mt19937(seed + b), ob.Mt19937(seed + b + nt)
        s, e = b * size, min((b + 1) * size, n)
        assert all(gn[i] == eg.uniform_real() for i in range(s, min(s + 5, e)))
        assert all(hn[i] == eh.uniform_real() for i in range(s, min(s + 5, e)))
    rng = np.random.default_rng(0)
    g = rng.normal(size=n).astype(np.float32)
    h = rng.uniform(0.01, 0.25, size=n).astype(np.float32)
    for bins in (2, 4, 16, 127):
        packed, gs, hs = ob.discretize(g, h, bins, stochastic=True, seed=seed, num_threads=nt)
        qg = (packed >> 8).astype(np.int64)
        qh = (packed & 0xff).astype(np.int64)
        assert np.all(np.abs(qg) <= bins // 2 + 1) and np.all(qh >= 0) and np.all(qh <= bins + 1)
        # unbiased rounding: E[q * scale] ~ value
        assert abs((qg * gs).mean() - g.mean()) < 4 * gs / np.sqrt(n) + 1e-3


def test_stochastic_rotation_offset():
    # one uniform_int(0, N) draw per tree rotates the noise (mod.rs:181-186)
    n, seed = 777, 5
    rng = np.random.default_rng(1)
    g = rng.normal(size=n).astype(np.float32)
    h = rng.uniform(0.1, 1.0, size=n).astype(np.float32)
    gn, hn = ob.noise(n, seed, 1)
    e = ob.Mt19937(seed)
    e.uniform_int(n)
    start = e.uniform_int(n)
    packed, gs, hs = ob.discretize(g, h, 4, stochastic=True, seed=seed, num_threads=1, draws=1)
    pos = (np.arange(n) + start) % n
    gd = g.astype(np.float64)
    sn = np.where(gd >= 0, gn[pos], -gn[pos])
    qg = np.clip(np.trunc(gd * (1.0 / gs) + sn), -128, 127).astype(np.int64)
    qh = np.clip(np.trunc(h.astype(np.float64) * (1.0 / hs) + hn[pos]), -128, 127).astype(np.int64)
    assert np.array_equal((packed >> 8).astype(np.int64), qg)
    assert np.array_equal((packed & 0xff).astype(np.int64), qh & 0xff)


@pytest.mark.parametrize("num_bin", [16, 255])
@pytest.mark.parametrize("n", [0, 1, 3, 17, 257])
def test_ordered_histogram_bitwise(num_bin, n):
    bins = np.array([(r // 7) % 3 for r in range(n)], np.uint8)
    rows = np.arange(n - 1, -1, -2, dtype=np.int32)
    pattern = np.array([1e20, 1.0, -1e20, -0.0, np.finfo(np.float32).tiny, -3.25, 0.5], np.float32)
    for subset in (False, True):
        idx = rows if subset else np.arange(n)
        g = pattern[idx % 7]
        h = ((idx % 11) * 0.125).astype(np.float32)
        for const in (False, True):
            expect = np.zeros(2 * num_bin)
            for i, r in enumerate(idx):
                b = bins[r] * 2
                expect[b] += float(g[i])
                expect[b + 1] += 1.0 if const else float(h[i])
            got = ob.hist_f64(bins, rows if subset else None, g, h, const, num_bin)
            assert got.tobytes() == expect.tobytes()


def _int_reference(hist, l2, min_data, gs, hs, num_data, min_hess=0.0):
    """Independent brute force of quantized_split_search.rs:58-136 (L2 only, no missing)."""
    tg, th = int(hist[:, 0].sum()), int(hist[:, 1].sum())
    cnt_factor = num_data / th
    out = lambda g, h: -g / (h + l2)  # noqa: E731
    lg_ = lambda g, h: -(2.0 * g * out(g, h) + (h + l2) * out(g, h) * out(g, h))  # noqa: E731
    mgs = lg_(tg * gs, th * hs)
    best = None
    for s in range(1, len(hist)):
        rg, rh = int(hist[s:, 0].sum()), int(hist[s:, 1].sum())
        lgi, lhi = tg - rg, th - rh
        rc = int(rh * cnt_factor + 0.5)
        lc = num_data - rc
        srh, slh = rh * hs, lhi * hs
        if rc < min_data or srh < min_hess or lc < min_data or slh < min_hess:
            continue
        gain = lg_(lgi * gs, slh + EPS) + lg_(rg * gs, srh + EPS)
        if gain <= mgs:
            continue
        if best is None or gain > best[0]:
            best = (gain, s - 1, int(lhi * cnt_factor + 0.5), int(rh * cnt_factor + 0.5), out(lgi * gs, slh),
                    out(rg * gs, srh))
    return best, mgs


@pytest.mark.parametrize("case", [
    ([(3, 10), (-5, 20), (8, 15), (-2, 25), (6, 30)], 1.0, 1.0, 0.0, 1),
    ([(3, 10), (-5, 20), (8, 15), (-2, 25), (6, 30)], 0.5, 1.0, 0.1, 5),
    ([(12, 7), (-9, 31), (4, 12), (-15, 44), (20, 19), (-1, 8)], 0.25, 0.125, 0.0, 1),
    ([(12, 7), (-9, 31), (4, 12), (-15, 44), (20, 19), (-1, 8)], 2.0, 0.5, 1.0, 10),
    ([(3, 4), (-5, 5), (8, 6)], 1.0, 1.0, 0.0, 100),
])
def test_integer_search_matches_brute_force(case):
    rows, gs, hs, l2, min_data = case
    hist = np.array(rows, np.int64)
    tg, th = int(hist[:, 0].sum()), int(hist[:, 1].sum())
    num_data = th
    # offset 0, as the reference test's FeatureMeta (a nonzero most frequent bin)
    r = ob.search(hist, len(hist), most_freq_bin=1, l2=l2, min_hess=0.0, min_data=min_data, int_g=tg, int_h=th, gs=gs, hs=hs,
                  num_data=num_data)
    best, mgs = _int_reference(hist, l2, min_data, gs, hs, num_data)
    assert r["splittable"] == (best is not None)
    if best is None:
        return
    assert r["threshold"] == best[1]
    assert (r["left_count"], r["right_count"]) == (best[2], best[3])
    assert abs(r["gain"] - (best[0] - mgs)) <= 1e-12 * max(abs(best[0] - mgs), 1.0)
    assert abs(r["left_output"] - best[4]) <= 1e-12 and abs(r["right_output"] - best[5]) <= 1e-12


def _f64_brute(hist, l2, min_data, sum_g, sum_h, num_data, min_hess):
    """All thresholds of a numerical feature without missing values, reverse-scan sums."""
    sh = sum_h + 2 * EPS
    cnt_factor = num_data / sh
    mgs = sum_g * sum_g / (sh + l2)
    best = None
    nb = len(hist)
    # the reverse scan visits t = nb-1 .. 1 and keeps the first strictly greater gain
    srg, srh, rc = 0.0, EPS, 0
    for t in range(nb - 1, 0, -1):
        srg += hist[t, 0]
        srh += hist[t, 1]
        rc += int(hist[t, 1] * cnt_factor + 0.5)
        if rc < min_data or srh < min_hess:
            continue
        lc = num_data - rc
        if lc < min_data:
            break
        slh = sh - srh
        if slh < min_hess:
            break
        slg = sum_g - srg
        gain = slg * slg / (slh + l2) + srg * srg / (srh + l2)
        if gain <= mgs:
            continue
        if best is None or gain > best[0]:
            best = (gain, t - 1, lc)
    return best, mgs


def test_f64_search_matches_reverse_scan_brute_force():
    rng = np.random.default_rng(3)
    for trial in range(40):
        nb = int(rng.integers(3, 40))
        cnt = rng.integers(0, 50, size=nb)
        g = rng.normal(size=nb) * cnt
        h = cnt * rng.uniform(0.5, 1.5, size=nb)
        hist = np.stack([g, h], axis=1)
        sum_g, sum_h, n = float(g.sum()), float(h.sum()), int(cnt.sum())
        l2 = float(rng.choice([0.0, 1.0]))
        r = ob.search(hist, nb, most_freq_bin=nb // 2, l2=l2, min_data=5, min_hess=1e-3, sum_g=sum_g, sum_h=sum_h,
                      num_data=n)
        best, mgs = _f64_brute(hist, l2, 5, sum_g, sum_h, n, 1e-3)
        assert r["splittable"] == (best is not None)
        if best is not None:
            assert r["threshold"] == best[1]
            assert r["left_count"] == best[2]
            assert r["gain"] == best[0] - mgs
