// bk_oracle.cpp — CPU restatement of the reference tree learner's hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker (or as the timed CPU baseline), never as the product path.
//
// It restates, in plain C++, the Rust of the reference learner at
//   crates/lgbm-core/src/treelearner/{serial,histogram,feature_histogram,simd,
//   split_info,data_partition}.rs and treelearner/gradient_discretizer/{mod,mt19937}.rs
// with the same arithmetic order. It must be compiled with -ffp-contract=off:
// the reference's `madd` is an unfused a*b+c on x86_64 (utils/common.rs:207-216).
//
// Pinning (see tests/test_oracle_*.py): MT19937 and libc++ distribution vectors
// (tests/mt19937.rs), discretizer known answers and the 16 upstream stumps of
// tests/fixtures/quantized_gradients.json, the ordered-histogram oracle with the
// cancellation patterns of tests/simd_histogram.rs, the brute-force integer split
// reference of tests/quantized_split_search.rs, brute force over all thresholds,
// and end-to-end the upstream golden models tests/fixtures/train/*/model.txt.
//
// Scope: dense single-value groups, numerical and categorical features, bagged
// row subsets, by-tree and by-node feature sampling; no monotone or interaction
// constraints, no forced splits, no CEGB, no linear trees, no extra trees. Those
// inputs are refused.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/bankai_gpu.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

constexpr double K_MIN_SCORE = -std::numeric_limits<double>::infinity();
// kEpsilon = 1e-15f widened to double (meta.rs:7, feature_histogram.rs:14).
const double EPS = static_cast<double>(1e-15f);
// kZeroThreshold = 1e-35f widened (meta.rs:4).
const double K_ZERO_THRESHOLD = static_cast<double>(1e-35f);

// Rust `f64 as i32`: saturating, NaN -> 0.
inline int32_t sat_i32(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return (int32_t)-2147483648LL;
  return (int32_t)x;
}
// Rust `f64 as i8`: saturating, NaN -> 0.
inline int8_t sat_i8(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 127.0) return 127;
  if (x <= -128.0) return -128;
  return (int8_t)x;
}

// ---------------------------------------------------------------------------
// std::mt19937 and libc++'s distributions (gradient_discretizer/mt19937.rs:48-128)
// ---------------------------------------------------------------------------
// The LCG of LightGBM's samplers (utils/random.rs:6-80).
struct LcgRandom {
  uint32_t x = 0;
  explicit LcgRandom(int32_t seed = 0) : x((uint32_t)seed) {}
  int32_t rand_int16() {
    x = x * 214013u + 2531011u;
    return (int32_t)((x >> 16) & 0x7FFFu);
  }
  int32_t rand_int32() {
    x = x * 214013u + 2531011u;
    return (int32_t)(x & 0x7FFFFFFFu);
  }
  int32_t next_int(int32_t lo, int32_t hi) { return rand_int32() % (hi - lo) + lo; }
  float next_float() { return (float)rand_int16() / 32768.0f; }
  // Random::sample: k sorted distinct indices of 0..n-1 (random.rs:52-79)
  std::vector<int> sample(int n, int k) {
    std::vector<int> ret;
    if (k > n || k <= 0) return ret;
    if (k == n) {
      for (int i = 0; i < n; ++i) ret.push_back(i);
      return ret;
    }
    if (k > 1 && (double)k > (double)n / std::log2((double)k)) {
      for (int i = 0; i < n; ++i) {
        const double prob = (double)(k - (int)ret.size()) / (double)(n - i);
        if ((double)next_float() < prob) ret.push_back(i);
      }
    } else {
      std::vector<int> set;
      auto insert = [&](int v) {
        auto it = std::lower_bound(set.begin(), set.end(), v);
        if (it != set.end() && *it == v) return false;
        set.insert(it, v);
        return true;
      };
      for (int r = n - k; r < n; ++r)
        if (!insert(next_int(0, r + 1))) insert(r);
      ret = set;
    }
    return ret;
  }
};

// ColSampler (treelearner/col_sampler.rs:14-218): the by-tree sample is drawn at
// set-up and at every tree, the by-node samples at every search, smaller leaf
// first; interaction constraints restrict a leaf's features to those allowed by
// the real features split on along its branch.
struct ColSampler {
  double frac_bynode = 1.0;
  LcgRandom rnd;
  std::vector<int> valid_inner;  // inner index of each feature, in real-index order
  int nf = 0;
  bool need_reset_bytree = false;
  int used_cnt_bytree = 0;
  std::vector<char> bytree;
  std::vector<int> used_idx;  // indices into valid_inner of the by-tree sample
  std::vector<std::vector<int>> constraints;  // sorted real feature ids per constraint
  std::vector<int> real_of;                   // real feature id of each inner feature

  // allowed_features (col_sampler.rs:178-201): the branch's own features, plus every
  // constraint that holds all of them (every constraint for an empty branch)
  std::vector<int> allowed(const std::vector<int>& branch) const {
    std::vector<int> out(branch.begin(), branch.end());
    for (const auto& c : constraints) {
      bool all = true;
      for (int f : branch)
        if (!std::binary_search(c.begin(), c.end(), f)) {
          all = false;
          break;
        }
      if (all) out.insert(out.end(), c.begin(), c.end());
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }

  static int get_cnt(int total, double fraction) {
    const int mn = std::min(1, total);
    return std::max(mn, (int)((double)total * fraction + 0.5));
  }

  void init(const std::vector<int>& valid, int num_features, double ff, double ffn, int32_t seed) {
    valid_inner = valid;
    nf = num_features;
    frac_bynode = ffn;
    rnd = LcgRandom(seed);
    need_reset_bytree = ff < 1.0;
    used_cnt_bytree = need_reset_bytree ? get_cnt((int)valid.size(), ff) : (int)valid.size();
    bytree.assign((size_t)nf, 1);
    reset_by_tree();
  }

  void reset_by_tree() {
    if (!need_reset_bytree) return;
    std::fill(bytree.begin(), bytree.end(), (char)0);
    used_idx = rnd.sample((int)valid_inner.size(), used_cnt_bytree);
    for (int i : used_idx) bytree[(size_t)valid_inner[(size_t)i]] = 1;
  }

  // get_by_node (col_sampler.rs:111-176)
  std::vector<char> by_node(const std::vector<int>& branch) {
    std::vector<char> ret((size_t)nf, 0);
    const bool constrained = !constraints.empty();
    const std::vector<int> ok = constrained ? allowed(branch) : std::vector<int>();
    auto is_ok = [&](int inner) { return std::binary_search(ok.begin(), ok.end(), real_of[(size_t)inner]); };
    if (frac_bynode >= 1.0) {
      for (int f = 0; f < nf; ++f) ret[(size_t)f] = !constrained || is_ok(f);
      return ret;
    }
    // candidates: the tree's sample (indices into valid_inner) or every feature,
    // filtered by the constraints; the count is taken before the filter
    std::vector<int> cand;
    int cnt;
    if (need_reset_bytree) {
      cnt = get_cnt((int)used_idx.size(), frac_bynode);
      for (int i : used_idx)
        if (!constrained || is_ok(valid_inner[(size_t)i])) cand.push_back(i);
    } else {
      cnt = get_cnt((int)valid_inner.size(), frac_bynode);
      for (int i = 0; i < (int)valid_inner.size(); ++i)
        if (!constrained || is_ok(valid_inner[(size_t)i])) cand.push_back(i);
    }
    if (constrained) cnt = std::min(cnt, (int)cand.size());
    for (int p : rnd.sample((int)cand.size(), cnt)) ret[(size_t)valid_inner[(size_t)cand[(size_t)p]]] = 1;
    return ret;
  }
};

struct Mt19937 {
  uint32_t st[624];
  int idx;
  explicit Mt19937(uint32_t seed = 5489u) { reseed(seed); }
  void reseed(uint32_t seed) {
    st[0] = seed;
    for (int i = 1; i < 624; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    idx = 624;
  }
  void twist() {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (st[i] & 0x80000000u) | (st[(i + 1) % 624] & 0x7fffffffu);
      uint32_t nx = st[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) nx ^= 0x9908b0dfu;
      st[i] = nx;
    }
    idx = 0;
  }
  uint32_t next_u32() {
    if (idx >= 624) twist();
    uint32_t y = st[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // libc++ generate_canonical<double,53>: (g1 + g2*2^32) / 2^64 (mt19937.rs:95-101).
  double uniform_real_01() {
    double g1 = (double)next_u32();
    double g2 = (double)next_u32();
    const double rp = 4294967296.0;
    double sp = g1 + g2 * rp;
    return sp / (rp * rp);
  }
  // libc++ uniform_int_distribution<int>(0, n) by masked rejection (mt19937.rs:109-128).
  uint32_t uniform_int_inclusive(uint32_t n) {
    uint32_t rp = n + 1u;
    if (rp == 1u) return 0;
    if (rp == 0u) return next_u32();
    uint32_t w = 32u - (uint32_t)__builtin_clz(rp) - 1u;
    if (rp & (0xffffffffu >> (32u - w))) w += 1u;
    uint32_t mask = w >= 32u ? 0xffffffffu : ((1u << w) - 1u);
    for (;;) {
      uint32_t u = next_u32() & mask;
      if (u < rp) return u;
    }
  }
};

// ---------------------------------------------------------------------------
// GradientDiscretizer (gradient_discretizer/mod.rs:35-216)
// ---------------------------------------------------------------------------
// Threading::BlockInfo (mod.rs:19-29).
inline void block_info(size_t num_threads, size_t cnt, size_t min_cnt, size_t* n_block, size_t* size) {
  size_t t = num_threads < 1 ? 1 : num_threads;
  size_t m = min_cnt < 1 ? 1 : min_cnt;
  size_t nb = std::min(t, (cnt + m - 1) / m);
  if (nb > 1) {
    size_t s = (cnt + nb - 1) / nb;
    *n_block = nb;
    *size = (s + 31) / 32 * 32;
  } else {
    *n_block = nb;
    *size = cnt;
  }
}

struct Discretizer {
  std::vector<int16_t> packed;
  int bins = 4;
  bool constant_hessian = false;
  double gradient_scale = 0.0, hessian_scale = 0.0;
  bool stochastic = false;
  std::vector<double> grad_rand, hess_rand;
  Mt19937 use_start_eng{0};

  Discretizer(size_t n, int b, bool ch) : packed(n, 0), bins(b), constant_hessian(ch) {}

  // init_stochastic (mod.rs:87-106).
  void init_stochastic(uint32_t seed, size_t num_threads) {
    size_t n = packed.size();
    use_start_eng.reseed(seed);
    grad_rand.assign(n, 0.0);
    hess_rand.assign(n, 0.0);
    size_t threads = num_threads < 1 ? 1 : num_threads;
    size_t nb, bs;
    block_info(threads, n, 512, &nb, &bs);
    for (size_t b = 0; b < nb; ++b) {
      size_t start = b * bs, end = std::min(start + bs, n);
      Mt19937 ge(seed + (uint32_t)b);
      Mt19937 he(seed + (uint32_t)b + (uint32_t)threads);
      for (size_t i = start; i < end; ++i) {
        grad_rand[i] = ge.uniform_real_01();
        hess_rand[i] = he.uniform_real_01();
      }
    }
    stochastic = true;
  }

  // update_scales (mod.rs:134-165). Returns false (and leaves state) on refusal.
  bool update_scales(const float* g, const float* h, double* inv_g, double* inv_h) {
    size_t n = packed.size();
    for (size_t i = 0; i < n; ++i) {
      if (!std::isfinite(g[i]) || !std::isfinite(h[i]) || h[i] < 0.0f) return false;
    }
    float max_g = 0.0f, max_h = 0.0f;
    for (size_t i = 0; i < n; ++i) max_g = std::max(max_g, std::fabs(g[i]));
    for (size_t i = 0; i < n; ++i) max_h = std::max(max_h, h[i]);
    if (max_g == 0.0f || (!constant_hessian && max_h == 0.0f)) return false;
    gradient_scale = (double)max_g / (double)(bins / 2);
    hessian_scale = constant_hessian ? (double)max_h : (double)max_h / (double)bins;
    *inv_g = 1.0 / gradient_scale;
    *inv_h = 1.0 / hessian_scale;
    return true;
  }

  // discretize_nearest (mod.rs:116-130).
  bool nearest(const float* g, const float* h) {
    double ig, ih;
    if (!update_scales(g, h, &ig, &ih)) return false;
    for (size_t i = 0; i < packed.size(); ++i) {
      double gd = (double)g[i];
      double off = gd >= 0.0 ? 0.5 : -0.5;
      int8_t qg = sat_i8(gd * ig + off);
      int8_t qh = constant_hessian ? (int8_t)1 : sat_i8((double)h[i] * ih + 0.5);
      packed[i] = (int16_t)(((int16_t)qg << 8) | (int16_t)(uint8_t)qh);
    }
    return true;
  }

  // discretize_stochastic (mod.rs:172-201).
  bool stochastic_round(const float* g, const float* h) {
    if (!stochastic) return false;
    double ig, ih;
    if (!update_scales(g, h, &ig, &ih)) return false;
    size_t n = packed.size();
    size_t start = use_start_eng.uniform_int_inclusive((uint32_t)n);
    for (size_t i = 0; i < n; ++i) {
      size_t pos = (i + start) % n;
      double gd = (double)g[i];
      double noise = grad_rand[pos];
      double sn = gd >= 0.0 ? noise : -noise;
      int8_t qg = sat_i8(gd * ig + sn);
      int8_t qh = constant_hessian ? (int8_t)1 : sat_i8((double)h[i] * ih + hess_rand[pos]);
      packed[i] = (int16_t)(((int16_t)qg << 8) | (int16_t)(uint8_t)qh);
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
// Gain and output formulas (feature_histogram.rs:58-176, simd.rs:13-37)
// ---------------------------------------------------------------------------
struct SplitParams {
  double l1, l2, max_delta_step, path_smooth;
  int min_data_in_leaf;
  double min_sum_hessian_in_leaf, min_gain_to_split;
};
struct CatParams {
  double cat_smooth, cat_l2;
  int max_cat_threshold, max_cat_to_onehot, min_data_per_group;
};

inline double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
inline int32_t round_int(double x) { return sat_i32(x + 0.5); }
inline double threshold_l1(double s, double l1) {
  double reg = std::max(std::fabs(s) - l1, 0.0);
  return sgn(s) * reg;
}
inline double madd(double a, double b, double c) { return a * b + c; }

template <bool L1, bool MAXO, bool SMOOTH>
inline double leaf_output(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent) {
  double ret = L1 ? -threshold_l1(g, l1) / (h + l2) : -g / (h + l2);
  if (MAXO && mds > 0.0 && std::fabs(ret) > mds) ret = sgn(ret) * mds;
  if (SMOOTH) {
    double nd = (double)n;
    ret = ret * (nd / smooth) / (nd / smooth + 1.0) + parent / (nd / smooth + 1.0);
  }
  return ret;
}
inline double leaf_output_rt(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent,
                             bool L1, bool M, bool S) {
  double ret = L1 ? -threshold_l1(g, l1) / (h + l2) : -g / (h + l2);
  if (M && mds > 0.0 && std::fabs(ret) > mds) ret = sgn(ret) * mds;
  if (S) {
    double nd = (double)n;
    ret = ret * (nd / smooth) / (nd / smooth + 1.0) + parent / (nd / smooth + 1.0);
  }
  return ret;
}
template <bool L1>
inline double gain_given_output(double g, double h, double l1, double l2, double out) {
  double gg = L1 ? threshold_l1(g, l1) : g;
  return -madd(2.0 * gg, out, (h + l2) * out * out);
}
template <bool L1, bool MAXO, bool SMOOTH>
inline double leaf_gain(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent) {
  if (!MAXO && !SMOOTH) {
    if (L1) {
      double sg = threshold_l1(g, l1);
      return (sg * sg) / (h + l2);
    }
    return (g * g) / (h + l2);
  }
  double out = leaf_output<L1, MAXO, SMOOTH>(g, h, l1, l2, mds, smooth, n, parent);
  return gain_given_output<L1>(g, h, l1, l2, out);
}
inline double leaf_gain_rt(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent,
                           bool L1, bool M, bool S) {
  if (!M && !S) {
    if (L1) {
      double sg = threshold_l1(g, l1);
      return (sg * sg) / (h + l2);
    }
    return (g * g) / (h + l2);
  }
  double out = leaf_output_rt(g, h, l1, l2, mds, smooth, n, parent, L1, M, S);
  return L1 ? gain_given_output<true>(g, h, l1, l2, out) : gain_given_output<false>(g, h, l1, l2, out);
}
inline double split_gain_plain(double gl, double hl, double gr, double hr, double l2) {
  return gl * gl / (hl + l2) + gr * gr / (hr + l2);
}
// leaf_output_root (feature_histogram.rs:104-117): clamp keeps NaN.
inline double leaf_output_root(const SplitParams& p, double g, double h, int n) {
  double r = leaf_output<true, true, false>(g, h, p.l1, p.l2, p.max_delta_step, p.path_smooth, n, 0.0);
  if (r < -std::numeric_limits<double>::max()) return -std::numeric_limits<double>::max();
  if (r > std::numeric_limits<double>::max()) return std::numeric_limits<double>::max();
  return r;
}

// ---------------------------------------------------------------------------
// SplitInfo (split_info.rs:9-96)
// ---------------------------------------------------------------------------
struct SplitInfo {
  int32_t feature = -1;
  uint32_t threshold = 0;
  int32_t left_count = 0, right_count = 0, num_cat_threshold = 0;
  double left_output = 0, right_output = 0, gain = K_MIN_SCORE;
  double left_sum_gradient = 0, left_sum_hessian = 0, right_sum_gradient = 0, right_sum_hessian = 0;
  int64_t left_int_sum = 0, right_int_sum = 0;
  std::vector<uint32_t> cat_threshold;
  bool default_left = true;
  void reset() {
    feature = -1;
    gain = K_MIN_SCORE;
  }
  bool better_than(const SplitInfo& o) const {
    double a = gain, b = o.gain;
    if (a != b) return a > b;
    int32_t fa = feature == -1 ? INT32_MAX : feature;
    int32_t fb = o.feature == -1 ? INT32_MAX : o.feature;
    return fa < fb;
  }
};
size_t argmax_split(const std::vector<SplitInfo>& v) {
  size_t best = 0;
  for (size_t i = 1; i < v.size(); ++i)
    if (v[i].better_than(v[best])) best = i;
  return best;
}

struct FeatureMeta {
  int num_bin, missing_type, offset, default_bin, bin_type;
  double penalty;
};

// ---------------------------------------------------------------------------
// Numerical threshold scan, f64 (feature_histogram.rs:219-663), MC off, RAND off
// ---------------------------------------------------------------------------
template <bool L1, bool MAXO, bool SMOOTH, bool REVERSE, bool SKIP_DEFAULT, bool NA_AS_MISSING>
bool sequential_f64(const double* hist, const FeatureMeta& m, const SplitParams& p, double sum_g, double sum_h,
                    int num_data, double min_gain_shift, double parent, SplitInfo& out) {
  const int offset = m.offset;
  bool splittable = false;
  double best_slg = NAN, best_slh = NAN, best_gain = K_MIN_SCORE;
  int best_lc = 0;
  uint32_t best_th = (uint32_t)m.num_bin;
  const double cnt_factor = (double)num_data / sum_h;
  auto G = [&](int t) { return hist[(size_t)t << 1]; };
  auto H = [&](int t) { return hist[((size_t)t << 1) + 1]; };
  auto gains = [&](double slg, double slh, double srg, double srh, int lc, int rc) {
    if (!L1 && !MAXO && !SMOOTH) return split_gain_plain(slg, slh, srg, srh, p.l2);
    return leaf_gain<L1, MAXO, SMOOTH>(slg, slh, p.l1, p.l2, p.max_delta_step, p.path_smooth, lc, parent) +
           leaf_gain<L1, MAXO, SMOOTH>(srg, srh, p.l1, p.l2, p.max_delta_step, p.path_smooth, rc, parent);
  };
  if (REVERSE) {
    double srg = 0.0, srh = EPS;
    int rc = 0;
    int t = m.num_bin - 1 - offset - (NA_AS_MISSING ? 1 : 0);
    const int t_end = 1 - offset;
    for (; t >= t_end; --t) {
      if (SKIP_DEFAULT && (t + offset) == m.default_bin) continue;
      double g = G(t), h = H(t);
      int cnt = round_int(h * cnt_factor);
      srg += g;
      srh += h;
      rc += cnt;
      if (rc < p.min_data_in_leaf || srh < p.min_sum_hessian_in_leaf) continue;
      int lc = num_data - rc;
      if (lc < p.min_data_in_leaf) break;
      double slh = sum_h - srh;
      if (slh < p.min_sum_hessian_in_leaf) break;
      double slg = sum_g - srg;
      double cur = gains(slg, slh, srg, srh, lc, rc);
      if (cur <= min_gain_shift) continue;
      splittable = true;
      if (cur > best_gain) {
        best_lc = lc;
        best_slg = slg;
        best_slh = slh;
        best_th = (uint32_t)(t - 1 + offset);
        best_gain = cur;
      }
    }
  } else {
    double slg = 0.0, slh = EPS;
    int lc = 0;
    int t = 0;
    const int t_end = m.num_bin - 2 - offset;
    if (NA_AS_MISSING && offset == 1) {
      slg = sum_g;
      slh = sum_h - EPS;
      lc = num_data;
      for (int i = 0; i < m.num_bin - offset; ++i) {
        double g = G(i), h = H(i);
        int cnt = round_int(h * cnt_factor);
        slg -= g;
        slh -= h;
        lc -= cnt;
      }
      t = -1;
    }
    for (; t <= t_end; ++t) {
      if (SKIP_DEFAULT && (t + offset) == m.default_bin) continue;
      if (t >= 0) {
        slg += G(t);
        slh += H(t);
        lc += round_int(H(t) * cnt_factor);
      }
      if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) continue;
      int rc = num_data - lc;
      if (rc < p.min_data_in_leaf) break;
      double srh = sum_h - slh;
      if (srh < p.min_sum_hessian_in_leaf) break;
      double srg = sum_g - slg;
      double cur = gains(slg, slh, srg, srh, lc, rc);
      if (cur <= min_gain_shift) continue;
      splittable = true;
      if (cur > best_gain) {
        best_lc = lc;
        best_slg = slg;
        best_slh = slh;
        best_th = (uint32_t)(t + offset);
        best_gain = cur;
      }
    }
  }
  if (splittable && best_gain > out.gain + min_gain_shift) {
    out.threshold = best_th;
    out.left_output = leaf_output<L1, MAXO, SMOOTH>(best_slg, best_slh, p.l1, p.l2, p.max_delta_step, p.path_smooth,
                                                    best_lc, parent);
    out.left_count = best_lc;
    out.left_sum_gradient = best_slg;
    out.left_sum_hessian = best_slh - EPS;
    out.right_output = leaf_output<L1, MAXO, SMOOTH>(sum_g - best_slg, sum_h - best_slh, p.l1, p.l2, p.max_delta_step,
                                                     p.path_smooth, num_data - best_lc, parent);
    out.right_count = num_data - best_lc;
    out.right_sum_gradient = sum_g - best_slg;
    out.right_sum_hessian = sum_h - best_slh - EPS;
    out.gain = best_gain - min_gain_shift;
    out.default_left = REVERSE;
  }
  return splittable;
}

template <bool L1, bool MAXO, bool SMOOTH>
bool numerical_f64(const double* hist, const FeatureMeta& m, const SplitParams& p, double sum_g, double sum_h,
                   int num_data, double parent, SplitInfo& out) {
  double gain_shift = leaf_gain<L1, MAXO, SMOOTH>(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, p.path_smooth, num_data,
                                                  parent);
  double mgs = gain_shift + p.min_gain_to_split;
  bool s = false;
  if (m.num_bin > 2 && m.missing_type != BK_MISSING_NONE) {
    if (m.missing_type == BK_MISSING_ZERO) {
      s |= sequential_f64<L1, MAXO, SMOOTH, true, true, false>(hist, m, p, sum_g, sum_h, num_data, mgs, parent, out);
      s |= sequential_f64<L1, MAXO, SMOOTH, false, true, false>(hist, m, p, sum_g, sum_h, num_data, mgs, parent, out);
    } else {
      s |= sequential_f64<L1, MAXO, SMOOTH, true, false, true>(hist, m, p, sum_g, sum_h, num_data, mgs, parent, out);
      s |= sequential_f64<L1, MAXO, SMOOTH, false, false, true>(hist, m, p, sum_g, sum_h, num_data, mgs, parent, out);
    }
  } else {
    s |= sequential_f64<L1, MAXO, SMOOTH, true, false, false>(hist, m, p, sum_g, sum_h, num_data, mgs, parent, out);
    if (m.missing_type == BK_MISSING_NAN) out.default_left = false;
  }
  return s;
}

bool find_best_threshold_numerical(const double* hist, const FeatureMeta& m, const SplitParams& p, double sum_g,
                                   double sum_h, int num_data, double parent, SplitInfo& out) {
  out.default_left = true;
  out.gain = K_MIN_SCORE;
  double sh = sum_h + 2.0 * EPS;
  bool L1 = p.l1 > 0.0, M = p.max_delta_step > 0.0, S = p.path_smooth > EPS;
  bool s;
#define BKO_GO(a, b, c) s = numerical_f64<a, b, c>(hist, m, p, sum_g, sh, num_data, parent, out)
  if (!L1 && !M && !S) BKO_GO(false, false, false);
  else if (!L1 && !M && S) BKO_GO(false, false, true);
  else if (!L1 && M && !S) BKO_GO(false, true, false);
  else if (!L1 && M && S) BKO_GO(false, true, true);
  else if (L1 && !M && !S) BKO_GO(true, false, false);
  else if (L1 && !M && S) BKO_GO(true, false, true);
  else if (L1 && M && !S) BKO_GO(true, true, false);
  else BKO_GO(true, true, true);
#undef BKO_GO
  out.gain *= m.penalty;
  return s;
}

// ---------------------------------------------------------------------------
// Numerical threshold scan, integer (feature_histogram.rs:673-1142)
// ---------------------------------------------------------------------------
template <bool L1, bool MAXO, bool SMOOTH, bool REVERSE, bool SKIP_DEFAULT, bool NA_AS_MISSING>
bool sequential_int(const int64_t* hist, const FeatureMeta& m, const SplitParams& p, int64_t tg, int64_t th, double gs,
                    double hs, int num_data, double mgs, double parent, SplitInfo& out) {
  const int offset = m.offset;
  bool splittable = false;
  int64_t best_lg = 0, best_lh = 0;
  double best_gain = K_MIN_SCORE;
  uint32_t best_th = (uint32_t)m.num_bin;
  const double cnt_factor = (double)num_data / (double)th;
  auto G = [&](int t) { return hist[(size_t)t * 2]; };
  auto H = [&](int t) { return hist[(size_t)t * 2 + 1]; };
  auto gains = [&](double slg, double slh, double srg, double srh, int lc, int rc) {
    slh += EPS;
    srh += EPS;
    if (!L1 && !MAXO && !SMOOTH) return split_gain_plain(slg, slh, srg, srh, p.l2);
    return leaf_gain<L1, MAXO, SMOOTH>(slg, slh, p.l1, p.l2, p.max_delta_step, p.path_smooth, lc, parent) +
           leaf_gain<L1, MAXO, SMOOTH>(srg, srh, p.l1, p.l2, p.max_delta_step, p.path_smooth, rc, parent);
  };
  if (REVERSE) {
    int64_t rg = 0, rh = 0;
    int t = m.num_bin - 1 - offset - (NA_AS_MISSING ? 1 : 0);
    const int t_end = 1 - offset;
    for (; t >= t_end; --t) {
      if (SKIP_DEFAULT && (t + offset) == m.default_bin) continue;
      rg += G(t);
      rh += H(t);
      int rc = round_int((double)rh * cnt_factor);
      double srh = (double)rh * hs;
      if (rc < p.min_data_in_leaf || srh < p.min_sum_hessian_in_leaf) continue;
      int lc = num_data - rc;
      if (lc < p.min_data_in_leaf) break;
      int64_t lh = th - rh;
      double slh = (double)lh * hs;
      if (slh < p.min_sum_hessian_in_leaf) break;
      int64_t lg = tg - rg;
      double slg = (double)lg * gs;
      double srg = (double)rg * gs;
      double cur = gains(slg, slh, srg, srh, lc, rc);
      if (cur <= mgs) continue;
      splittable = true;
      if (cur > best_gain) {
        best_lg = lg;
        best_lh = lh;
        best_th = (uint32_t)(t - 1 + offset);
        best_gain = cur;
      }
    }
  } else {
    int64_t lg = 0, lh = 0;
    int t = 0;
    const int t_end = m.num_bin - 2 - offset;
    if (NA_AS_MISSING && offset == 1) {
      lg = tg;
      lh = th;
      for (int i = 0; i < m.num_bin - offset; ++i) {
        lg -= G(i);
        lh -= H(i);
      }
      t = -1;
    }
    for (; t <= t_end; ++t) {
      if (SKIP_DEFAULT && (t + offset) == m.default_bin) continue;
      if (t >= 0) {
        lg += G(t);
        lh += H(t);
      }
      int lc = round_int((double)lh * cnt_factor);
      double slh = (double)lh * hs;
      if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) continue;
      int rc = num_data - lc;
      if (rc < p.min_data_in_leaf) break;
      int64_t rh = th - lh;
      double srh = (double)rh * hs;
      if (srh < p.min_sum_hessian_in_leaf) break;
      int64_t rg = tg - lg;
      double slg = (double)lg * gs;
      double srg = (double)rg * gs;
      double cur = gains(slg, slh, srg, srh, lc, rc);
      if (cur <= mgs) continue;
      splittable = true;
      if (cur > best_gain) {
        best_lg = lg;
        best_lh = lh;
        best_th = (uint32_t)(t + offset);
        best_gain = cur;
      }
    }
  }
  if (splittable && best_gain > out.gain + mgs) {
    int64_t rg = tg - best_lg, rh = th - best_lh;
    double slg = (double)best_lg * gs, slh = (double)best_lh * hs;
    double srg = (double)rg * gs, srh = (double)rh * hs;
    int lc = round_int((double)best_lh * cnt_factor);
    int rc = round_int((double)rh * cnt_factor);
    out.threshold = best_th;
    out.left_output = leaf_output<L1, MAXO, SMOOTH>(slg, slh, p.l1, p.l2, p.max_delta_step, p.path_smooth, lc, parent);
    out.left_count = lc;
    out.left_sum_gradient = slg;
    out.left_sum_hessian = slh;
    out.right_output = leaf_output<L1, MAXO, SMOOTH>(srg, srh, p.l1, p.l2, p.max_delta_step, p.path_smooth, rc, parent);
    out.right_count = rc;
    out.right_sum_gradient = srg;
    out.right_sum_hessian = srh;
    out.left_int_sum = (int64_t)(((uint64_t)best_lg << 32) | ((uint64_t)best_lh & 0xffffffffull));
    out.right_int_sum = (int64_t)(((uint64_t)rg << 32) | ((uint64_t)rh & 0xffffffffull));
    out.gain = best_gain - mgs;
    out.default_left = REVERSE;
  }
  return splittable;
}

template <bool L1, bool MAXO, bool SMOOTH>
bool numerical_int(const int64_t* hist, const FeatureMeta& m, const SplitParams& p, int64_t tg, int64_t th, double gs,
                   double hs, double sum_g, double sum_h, int num_data, double parent, SplitInfo& out) {
  double gain_shift = leaf_gain<L1, MAXO, SMOOTH>(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, p.path_smooth, num_data,
                                                  parent);
  double mgs = gain_shift + p.min_gain_to_split;
  bool s = false;
#define BKO_SEQ(r, k, na) s |= sequential_int<L1, MAXO, SMOOTH, r, k, na>(hist, m, p, tg, th, gs, hs, num_data, mgs, parent, out)
  if (m.num_bin > 2 && m.missing_type != BK_MISSING_NONE) {
    if (m.missing_type == BK_MISSING_ZERO) {
      BKO_SEQ(true, true, false);
      BKO_SEQ(false, true, false);
    } else {
      BKO_SEQ(true, false, true);
      BKO_SEQ(false, false, true);
    }
  } else {
    BKO_SEQ(true, false, false);
    if (m.missing_type == BK_MISSING_NAN) out.default_left = false;
  }
#undef BKO_SEQ
  return s;
}

bool find_best_threshold_numerical_int(const int64_t* hist, const FeatureMeta& m, const SplitParams& p, int64_t tg,
                                       int64_t th, double gs, double hs, int num_data, double parent, SplitInfo& out) {
  out.default_left = true;
  out.gain = K_MIN_SCORE;
  double sum_g = (double)tg * gs, sum_h = (double)th * hs;
  bool L1 = p.l1 > 0.0, M = p.max_delta_step > 0.0, S = p.path_smooth > EPS;
  bool s;
#define BKO_GO(a, b, c) s = numerical_int<a, b, c>(hist, m, p, tg, th, gs, hs, sum_g, sum_h, num_data, parent, out)
  if (!L1 && !M && !S) BKO_GO(false, false, false);
  else if (!L1 && !M && S) BKO_GO(false, false, true);
  else if (!L1 && M && !S) BKO_GO(false, true, false);
  else if (!L1 && M && S) BKO_GO(false, true, true);
  else if (L1 && !M && !S) BKO_GO(true, false, false);
  else if (L1 && !M && S) BKO_GO(true, false, true);
  else if (L1 && M && !S) BKO_GO(true, true, false);
  else BKO_GO(true, true, true);
#undef BKO_GO
  out.gain *= m.penalty;
  return s;
}

// ---------------------------------------------------------------------------
// Categorical scans (feature_histogram.rs:1328-1597 f64, 1606-1886 int), MC off, RAND off
// ---------------------------------------------------------------------------
inline double split_gains_cat(double slg, double slh, double srg, double srh, const SplitParams& p, double l2, int lc,
                              int rc, double parent, bool L1, bool M, bool S) {
  return leaf_gain_rt(slg, slh, p.l1, l2, p.max_delta_step, p.path_smooth, lc, parent, L1, M, S) +
         leaf_gain_rt(srg, srh, p.l1, l2, p.max_delta_step, p.path_smooth, rc, parent, L1, M, S);
}

bool find_best_threshold_categorical(const double* hist, const FeatureMeta& m, const SplitParams& p,
                                     const CatParams& c, double sum_g, double sum_h_in, int num_data, double parent,
                                     SplitInfo& out) {
  double sum_h = sum_h_in + 2.0 * EPS;
  out.default_left = false;
  out.gain = K_MIN_SCORE;
  bool L1 = p.l1 > 0.0, M = p.max_delta_step > 0.0, S = p.path_smooth > EPS;
  double gain_shift;
  if (S) {
    gain_shift = L1 ? gain_given_output<true>(sum_g, sum_h, p.l1, p.l2, parent)
                    : gain_given_output<false>(sum_g, sum_h, p.l1, p.l2, parent);
  } else {
    gain_shift = leaf_gain_rt(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, 0.0, num_data, 0.0, L1, M, false);
  }
  double mgs = gain_shift + p.min_gain_to_split;
  const int offset = m.offset;
  const int bin_start = 1 - offset, bin_end = m.num_bin - offset;
  auto G = [&](int t) { return hist[(size_t)t << 1]; };
  auto H = [&](int t) { return hist[((size_t)t << 1) + 1]; };
  const double cnt_factor = (double)num_data / sum_h;
  const bool onehot = m.num_bin <= c.max_cat_to_onehot;
  double l2 = p.l2;
  bool splittable = false;
  double best_gain = K_MIN_SCORE;
  int best_threshold = -1, best_dir = 1, best_lc = 0;
  double best_slg = 0.0, best_slh = 0.0;
  int used_bin = -1;
  std::vector<int> sorted_idx;
  if (onehot) {
    for (int t = bin_start; t < bin_end; ++t) {
      double g = G(t), h = H(t);
      int cnt = round_int(h * cnt_factor);
      if (cnt < p.min_data_in_leaf || h < p.min_sum_hessian_in_leaf) continue;
      int oc = num_data - cnt;
      if (oc < p.min_data_in_leaf) continue;
      double soh = sum_h - h - EPS;
      if (soh < p.min_sum_hessian_in_leaf) continue;
      double sog = sum_g - g;
      double cur = split_gains_cat(sog, soh, g, h + EPS, p, l2, oc, cnt, parent, L1, M, S);
      if (cur <= mgs) continue;
      splittable = true;
      if (cur > best_gain) {
        best_threshold = t;
        best_slg = g;
        best_slh = h + EPS;
        best_lc = cnt;
        best_gain = cur;
      }
    }
  } else {
    for (int i = bin_start; i < bin_end; ++i)
      if ((double)round_int(H(i) * cnt_factor) >= c.cat_smooth) sorted_idx.push_back(i);
    used_bin = (int)sorted_idx.size();
    l2 += c.cat_l2;
    auto ctr = [&](int i) { return G(i) / (H(i) + c.cat_smooth); };
    std::stable_sort(sorted_idx.begin(), sorted_idx.end(), [&](int a, int b) { return ctr(a) < ctr(b); });
    const int dirs[2] = {1, -1};
    const int starts[2] = {0, used_bin - 1};
    int max_num_cat = std::min(c.max_cat_threshold, (used_bin + 1) / 2);
    for (int d = 0; d < 2; ++d) {
      int pos = starts[d];
      int cnt_cur_group = 0;
      double slg = 0.0, slh = EPS;
      int lc = 0;
      for (int i = 0; i < used_bin && i < max_num_cat; ++i) {
        int t = sorted_idx[(size_t)pos];
        pos += dirs[d];
        double g = G(t), h = H(t);
        int cnt = round_int(h * cnt_factor);
        slg += g;
        slh += h;
        lc += cnt;
        cnt_cur_group += cnt;
        if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) continue;
        int rc = num_data - lc;
        if (rc < p.min_data_in_leaf || rc < c.min_data_per_group) break;
        double srh = sum_h - slh;
        if (srh < p.min_sum_hessian_in_leaf) break;
        if (cnt_cur_group < c.min_data_per_group) continue;
        cnt_cur_group = 0;
        double srg = sum_g - slg;
        double cur = split_gains_cat(slg, slh, srg, srh, p, l2, lc, rc, parent, L1, M, S);
        if (cur > mgs) {
          splittable = true;
          if (cur > best_gain) {
            best_lc = lc;
            best_slg = slg;
            best_slh = slh;
            best_threshold = i;
            best_gain = cur;
            best_dir = dirs[d];
          }
        }
      }
    }
  }
  if (splittable) {
    out.left_output = leaf_output_rt(best_slg, best_slh, p.l1, l2, p.max_delta_step, p.path_smooth, best_lc, parent,
                                     L1, M, S);
    out.left_count = best_lc;
    out.left_sum_gradient = best_slg;
    out.left_sum_hessian = best_slh - EPS;
    out.right_output = leaf_output_rt(sum_g - best_slg, sum_h - best_slh, p.l1, l2, p.max_delta_step, p.path_smooth,
                                      num_data - best_lc, parent, L1, M, S);
    out.right_count = num_data - best_lc;
    out.right_sum_gradient = sum_g - best_slg;
    out.right_sum_hessian = sum_h - best_slh - EPS;
    out.gain = best_gain - mgs;
    out.cat_threshold.clear();
    if (onehot) {
      out.num_cat_threshold = 1;
      out.cat_threshold.push_back((uint32_t)(best_threshold + offset));
    } else {
      out.num_cat_threshold = best_threshold + 1;
      for (int i = 0; i <= best_threshold; ++i) {
        int pos = best_dir == 1 ? i : used_bin - 1 - i;
        out.cat_threshold.push_back((uint32_t)(sorted_idx[(size_t)pos] + offset));
      }
    }
  }
  out.gain *= m.penalty;
  return splittable;
}

bool find_best_threshold_categorical_int(const int64_t* hist, const FeatureMeta& m, const SplitParams& p,
                                         const CatParams& c, int64_t tg, int64_t th, double gs, double hs,
                                         int num_data, double parent, SplitInfo& out) {
  double sum_g = (double)tg * gs, sum_h = (double)th * hs;
  out.default_left = false;
  out.gain = K_MIN_SCORE;
  bool L1 = p.l1 > 0.0, M = p.max_delta_step > 0.0, S = p.path_smooth > EPS;
  double gain_shift;
  if (S) {
    gain_shift = L1 ? gain_given_output<true>(sum_g, sum_h, p.l1, p.l2, parent)
                    : gain_given_output<false>(sum_g, sum_h, p.l1, p.l2, parent);
  } else {
    gain_shift = leaf_gain_rt(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, 0.0, num_data, 0.0, L1, M, false);
  }
  double mgs = gain_shift + p.min_gain_to_split;
  const int offset = m.offset;
  const int bin_start = 1 - offset, bin_end = m.num_bin - offset;
  auto IG = [&](int t) { return hist[(size_t)t * 2]; };
  auto IH = [&](int t) { return hist[(size_t)t * 2 + 1]; };
  const double cnt_factor = (double)num_data / (double)th;
  const bool onehot = m.num_bin <= c.max_cat_to_onehot;
  double l2 = p.l2;
  bool splittable = false;
  double best_gain = K_MIN_SCORE;
  int best_threshold = -1, best_dir = 1;
  int64_t best_lg = 0, best_lh = 0;
  int used_bin = -1;
  std::vector<int> sorted_idx;
  if (onehot) {
    for (int t = bin_start; t < bin_end; ++t) {
      int64_t ig = IG(t), ih = IH(t);
      double h = (double)ih * hs;
      int cnt = round_int((double)ih * cnt_factor);
      if (cnt < p.min_data_in_leaf || h < p.min_sum_hessian_in_leaf) continue;
      int oc = num_data - cnt;
      if (oc < p.min_data_in_leaf) continue;
      int64_t ioh = th - ih;
      double soh = (double)ioh * hs;
      if (soh < p.min_sum_hessian_in_leaf) continue;
      double g = (double)ig * gs;
      double sog = (double)(tg - ig) * gs;
      double cur = split_gains_cat(sog, soh, g, h, p, l2, oc, cnt, parent, L1, M, S);
      if (cur <= mgs) continue;
      splittable = true;
      if (cur > best_gain) {
        best_threshold = t;
        best_lg = ig;
        best_lh = ih;
        best_gain = cur;
      }
    }
  } else {
    for (int i = bin_start; i < bin_end; ++i)
      if ((double)round_int((double)IH(i) * cnt_factor) >= c.cat_smooth) sorted_idx.push_back(i);
    used_bin = (int)sorted_idx.size();
    l2 += c.cat_l2;
    auto ctr = [&](int i) { return ((double)IG(i) * gs) / ((double)IH(i) * hs + c.cat_smooth); };
    std::stable_sort(sorted_idx.begin(), sorted_idx.end(), [&](int a, int b) { return ctr(a) < ctr(b); });
    const int dirs[2] = {1, -1};
    const int starts[2] = {0, used_bin - 1};
    int max_num_cat = std::min(c.max_cat_threshold, (used_bin + 1) / 2);
    for (int d = 0; d < 2; ++d) {
      int pos = starts[d];
      int cnt_cur_group = 0;
      int64_t lg = 0, lh = 0;
      int lc = 0;
      for (int i = 0; i < used_bin && i < max_num_cat; ++i) {
        int t = sorted_idx[(size_t)pos];
        pos += dirs[d];
        int64_t ig = IG(t), ih = IH(t);
        int cnt = round_int((double)ih * cnt_factor);
        lg += ig;
        lh += ih;
        lc += cnt;
        cnt_cur_group += cnt;
        double slh = (double)lh * hs;
        if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) continue;
        int rc = num_data - lc;
        if (rc < p.min_data_in_leaf || rc < c.min_data_per_group) break;
        int64_t rh = th - lh;
        double srh = (double)rh * hs;
        if (srh < p.min_sum_hessian_in_leaf) break;
        if (cnt_cur_group < c.min_data_per_group) continue;
        cnt_cur_group = 0;
        double slg = (double)lg * gs;
        double srg = (double)(tg - lg) * gs;
        double cur = split_gains_cat(slg, slh, srg, srh, p, l2, lc, rc, parent, L1, M, S);
        if (cur > mgs) {
          splittable = true;
          if (cur > best_gain) {
            best_lg = lg;
            best_lh = lh;
            best_threshold = i;
            best_gain = cur;
            best_dir = dirs[d];
          }
        }
      }
    }
  }
  if (splittable) {
    double slg = (double)best_lg * gs, slh = (double)best_lh * hs;
    int64_t rg = tg - best_lg, rh = th - best_lh;
    double srg = (double)rg * gs, srh = (double)rh * hs;
    int lc = round_int((double)best_lh * cnt_factor);
    int rc = round_int((double)rh * cnt_factor);
    out.left_output = leaf_output_rt(slg, slh, p.l1, l2, p.max_delta_step, p.path_smooth, lc, parent, L1, M, S);
    out.left_count = lc;
    out.left_sum_gradient = slg;
    out.left_sum_hessian = slh;
    out.right_output = leaf_output_rt(srg, srh, p.l1, l2, p.max_delta_step, p.path_smooth, rc, parent, L1, M, S);
    out.right_count = rc;
    out.right_sum_gradient = srg;
    out.right_sum_hessian = srh;
    out.left_int_sum = (int64_t)(((uint64_t)best_lg << 32) | ((uint64_t)best_lh & 0xffffffffull));
    out.right_int_sum = (int64_t)(((uint64_t)rg << 32) | ((uint64_t)rh & 0xffffffffull));
    out.gain = best_gain - mgs;
    out.cat_threshold.clear();
    if (onehot) {
      out.num_cat_threshold = 1;
      out.cat_threshold.push_back((uint32_t)(best_threshold + offset));
    } else {
      out.num_cat_threshold = best_threshold + 1;
      for (int i = 0; i <= best_threshold; ++i) {
        int pos = best_dir == 1 ? i : used_bin - 1 - i;
        out.cat_threshold.push_back((uint32_t)(sorted_idx[(size_t)pos] + offset));
      }
    }
  }
  out.gain *= m.penalty;
  return splittable;
}

// ---------------------------------------------------------------------------
// Bitsets (utils/common.rs:221-240)
// ---------------------------------------------------------------------------
std::vector<uint32_t> construct_bitset(const std::vector<uint32_t>& vals) {
  std::vector<uint32_t> r;
  for (uint32_t v : vals) {
    size_t i1 = v / 32;
    if (r.size() < i1 + 1) r.resize(i1 + 1, 0);
    r[i1] |= 1u << (v % 32);
  }
  return r;
}
inline bool find_in_bitset(const uint32_t* bits, size_t n, uint32_t pos) {
  size_t i1 = pos / 32;
  if (i1 >= n) return false;
  return (bits[i1] >> (pos % 32)) & 1u;
}

// ---------------------------------------------------------------------------
// Dataset view held by the learner
// ---------------------------------------------------------------------------
struct Group {
  const uint8_t* bins;
  int num_total_bin, feature_start, feature_cnt;
  std::vector<int> bin_offsets;
};
struct Feature {
  int group, sub;
  FeatureMeta meta;
  int most_freq_bin, real_index;
  std::vector<double> upper;
  std::vector<int32_t> bin2cat;
};

// ---------------------------------------------------------------------------
// SplitRule (data_partition.rs:233-410), dense columns
// ---------------------------------------------------------------------------
struct SplitRule {
  bool use_min_bin = false;
  uint32_t minb = 0, maxb = 0, th = 0, t_zero_bin = 0;
  bool miss_is_zero = false, miss_is_na = false, mfb_is_zero = false, mfb_is_na = false;
  bool default_left_route = false, missing_default_left = false, multi_bin = false, max_bin_left = false;
  bool is_cat = false;
  std::vector<uint32_t> cat_bits;
  uint32_t cat_offset = 0;
  const uint8_t* col = nullptr;

  bool route_left(uint32_t bin) const {
    if (is_cat) {
      if ((use_min_bin && (bin < minb || bin > maxb)) || (!use_min_bin && bin == 0)) return default_left_route;
      return find_in_bitset(cat_bits.data(), cat_bits.size(), bin - minb + cat_offset);
    }
    if (multi_bin) {
      if ((miss_is_zero && !mfb_is_zero && bin == t_zero_bin) || (miss_is_na && !mfb_is_na && bin == maxb))
        return missing_default_left;
      if ((use_min_bin && (bin < minb || bin > maxb)) || (!use_min_bin && bin == 0)) {
        if ((miss_is_na && mfb_is_na) || (miss_is_zero && mfb_is_zero)) return missing_default_left;
        return default_left_route;
      }
      return !(bin > th);
    }
    if (miss_is_zero && !mfb_is_zero && bin == t_zero_bin) return missing_default_left;
    if (bin != maxb) {
      if ((miss_is_na && mfb_is_na) || (miss_is_zero && mfb_is_zero)) return missing_default_left;
      return default_left_route;
    }
    if (miss_is_na && !mfb_is_na) return missing_default_left;
    return max_bin_left;
  }
};

void group_view(const Group& g, int sub, uint32_t* minb, uint32_t* maxb, bool* use_min_bin) {
  if (g.feature_cnt == 1) {
    *minb = 1;
    *maxb = (uint32_t)g.bin_offsets[1] - 1;
    *use_min_bin = false;
  } else {
    *minb = (uint32_t)g.bin_offsets[(size_t)sub];
    *maxb = (uint32_t)g.bin_offsets[(size_t)sub + 1] - 1;
    *use_min_bin = true;
  }
}

SplitRule rule_numerical(const Group& g, const Feature& f, uint32_t threshold, bool default_left) {
  SplitRule r;
  uint32_t minb, maxb;
  bool umb;
  group_view(g, f.sub, &minb, &maxb, &umb);
  uint32_t default_bin = (uint32_t)f.meta.default_bin, mfb = (uint32_t)f.most_freq_bin;
  r.col = g.bins;
  r.use_min_bin = umb;
  r.minb = minb;
  r.maxb = maxb;
  switch (f.meta.missing_type) {
    case BK_MISSING_ZERO:
      r.miss_is_zero = true;
      r.mfb_is_zero = default_bin == mfb;
      break;
    case BK_MISSING_NAN: {
      r.miss_is_na = true;
      r.mfb_is_na = umb ? (maxb == mfb + minb && mfb > 0) : (maxb == mfb + 1 && mfb > 0);
      break;
    }
    default:
      break;
  }
  uint32_t th = threshold + minb;
  uint32_t tz = minb + default_bin;
  if (mfb == 0) {
    th -= 1;
    tz -= 1;
  }
  r.th = th;
  r.t_zero_bin = tz;
  r.default_left_route = mfb <= threshold;
  r.missing_default_left = (r.miss_is_zero || r.miss_is_na) && default_left;
  r.max_bin_left = maxb <= th;
  r.multi_bin = minb < maxb;
  return r;
}

SplitRule rule_categorical(const Group& g, const Feature& f, const std::vector<uint32_t>& bin_bits) {
  SplitRule r;
  uint32_t minb, maxb;
  bool umb;
  group_view(g, f.sub, &minb, &maxb, &umb);
  uint32_t mfb = (uint32_t)f.most_freq_bin;
  r.col = g.bins;
  r.is_cat = true;
  r.use_min_bin = umb;
  r.minb = minb;
  r.maxb = maxb;
  r.cat_offset = mfb == 0 ? 1u : 0u;
  r.default_left_route = mfb > 0 && find_in_bitset(bin_bits.data(), bin_bits.size(), mfb);
  r.cat_bits = bin_bits;
  return r;
}

// ---------------------------------------------------------------------------
// HistogramPool (serial.rs:80-162)
// ---------------------------------------------------------------------------
struct Pool {
  std::vector<std::vector<double>> slots;
  std::vector<std::vector<int64_t>> int_slots;  // (g, h) pairs
  std::vector<std::vector<char>> splittable;
  std::vector<int> leaf_to_slot, slot_to_leaf;
  size_t next_free = 0;
  std::vector<uint64_t> last_used;
  uint64_t clock = 0;

  void init(size_t num_leaves, size_t total_bins, size_t nf, size_t cache, bool quant) {
    if (quant) int_slots.assign(cache, std::vector<int64_t>(total_bins * 2, 0));
    else slots.assign(cache, std::vector<double>(total_bins * 2, 0.0));
    splittable.assign(cache, std::vector<char>(nf, 1));
    leaf_to_slot.assign(num_leaves, -1);
    slot_to_leaf.assign(cache, -1);
    last_used.assign(cache < num_leaves ? cache : 0, 0);
  }
  size_t size() const { return splittable.size(); }
  void reset_map() {
    std::fill(leaf_to_slot.begin(), leaf_to_slot.end(), -1);
    std::fill(slot_to_leaf.begin(), slot_to_leaf.end(), -1);
    next_free = 0;
    std::fill(last_used.begin(), last_used.end(), 0);
    clock = 0;
  }
  void touch(size_t s) {
    if (!last_used.empty()) last_used[s] = ++clock;
  }
  std::pair<size_t, bool> get(size_t leaf) {
    if (leaf_to_slot[leaf] >= 0) {
      size_t s = (size_t)leaf_to_slot[leaf];
      touch(s);
      return {s, true};
    }
    size_t s;
    if (next_free < size()) {
      s = next_free++;
    } else {
      size_t oldest = 0;
      for (size_t i = 1; i < last_used.size(); ++i)
        if (last_used[i] < last_used[oldest]) oldest = i;
      if (slot_to_leaf[oldest] >= 0) leaf_to_slot[(size_t)slot_to_leaf[oldest]] = -1;
      s = oldest;
    }
    touch(s);
    leaf_to_slot[leaf] = (int)s;
    slot_to_leaf[s] = (int)leaf;
    return {s, false};
  }
  void move_leaf(size_t src, size_t dst) {
    int s = leaf_to_slot[src];
    if (s < 0) return;
    leaf_to_slot[src] = -1;
    if (leaf_to_slot[dst] >= 0) slot_to_leaf[(size_t)leaf_to_slot[dst]] = -1;
    leaf_to_slot[dst] = s;
    slot_to_leaf[(size_t)s] = (int)dst;
    touch((size_t)s);
  }
};

// ---------------------------------------------------------------------------
// Tree (tree.rs:259-400): training fields only
// ---------------------------------------------------------------------------
struct Tree {
  int num_leaves = 1, num_cat = 0;
  std::vector<int32_t> left_child, right_child, split_feature_inner, split_feature, leaf_parent, leaf_depth, leaf_count,
      internal_count, cat_boundaries;
  std::vector<uint32_t> threshold_in_bin, cat_threshold;
  std::vector<double> threshold, leaf_value, leaf_weight, internal_value, internal_weight;
  std::vector<float> split_gain;
  std::vector<int8_t> decision_type;
  explicit Tree(int max_leaves) {
    size_t ni = max_leaves > 1 ? (size_t)max_leaves - 1 : 0, nl = (size_t)max_leaves;
    left_child.assign(ni, 0);
    right_child.assign(ni, 0);
    split_feature_inner.assign(ni, 0);
    split_feature.assign(ni, 0);
    split_gain.assign(ni, 0.0f);
    threshold.assign(ni, 0.0);
    threshold_in_bin.assign(ni, 0);
    decision_type.assign(ni, 0);
    internal_value.assign(ni, 0.0);
    internal_weight.assign(ni, 0.0);
    internal_count.assign(ni, 0);
    leaf_value.assign(nl, 0.0);
    leaf_weight.assign(nl, 0.0);
    leaf_count.assign(nl, 0);
    leaf_parent.assign(nl, 0);
    if (nl > 0) leaf_parent[0] = -1;
    leaf_depth.assign(nl, 0);
    cat_boundaries.assign(1, 0);
  }
  static double round_to_zero(double v) { return (v >= -K_ZERO_THRESHOLD && v <= K_ZERO_THRESHOLD) ? 0.0 : v; }
  void set_leaf_output(int leaf, double v) { leaf_value[(size_t)leaf] = round_to_zero(v); }
  int split_common(int leaf, int feature, int real, double lv, double rv, int lc, int rc, double lw, double rw,
                   float gain) {
    int new_node = num_leaves - 1, nl = num_leaves;
    int parent = leaf_parent[(size_t)leaf];
    if (parent >= 0) {
      if (left_child[(size_t)parent] == ~leaf) left_child[(size_t)parent] = new_node;
      else right_child[(size_t)parent] = new_node;
    }
    split_feature_inner[(size_t)new_node] = feature;
    split_feature[(size_t)new_node] = real;
    split_gain[(size_t)new_node] = gain;
    left_child[(size_t)new_node] = ~leaf;
    right_child[(size_t)new_node] = ~nl;
    leaf_parent[(size_t)leaf] = new_node;
    leaf_parent[(size_t)nl] = new_node;
    internal_weight[(size_t)new_node] = lw + rw;
    internal_value[(size_t)new_node] = leaf_value[(size_t)leaf];
    internal_count[(size_t)new_node] = lc + rc;
    leaf_value[(size_t)leaf] = std::isnan(lv) ? 0.0 : lv;
    leaf_weight[(size_t)leaf] = lw;
    leaf_count[(size_t)leaf] = lc;
    leaf_value[(size_t)nl] = std::isnan(rv) ? 0.0 : rv;
    leaf_weight[(size_t)nl] = rw;
    leaf_count[(size_t)nl] = rc;
    leaf_depth[(size_t)nl] = leaf_depth[(size_t)leaf] + 1;
    leaf_depth[(size_t)leaf] += 1;
    return new_node;
  }
};

struct LeafSplits {
  int leaf_index = -1, num_data = 0;
  double sum_g = 0, sum_h = 0, weight = 0;
  int64_t int_sum = 0;
};

// ---------------------------------------------------------------------------
// The learner (serial.rs:183-2433)
// ---------------------------------------------------------------------------
struct Learner {
  size_t num_data = 0;
  int nf = 0, num_leaves_cfg = 0, max_depth = 0, num_threads = 1;
  std::vector<Group> groups;
  std::vector<Feature> feats;
  std::vector<size_t> group_start, feat_off, feat_len;
  size_t total_bins = 0;
  SplitParams sp{};
  CatParams cp{};
  bool quant = false, stochastic = false, renew = false, const_hess = false;
  double min_gain_to_split = 0.0;
  Pool pool;
  std::vector<SplitInfo> best;
  LeafSplits smaller, larger;
  size_t smaller_slot = 0;
  long larger_slot = -1, parent_slot = -1;
  // DataPartition (data_partition.rs:66-175)
  std::vector<int32_t> indices, lbuf, rbuf;
  std::vector<size_t> leaf_begin, leaf_cnt;
  std::vector<float> og, oh;
  Discretizer* disc = nullptr;
  double grad_scale = 0.0, hess_scale = 0.0;
  ColSampler cols;
  std::vector<std::vector<int>> branch;  // real features split on along each leaf's path
  // per-tree stats
  int64_t stat_hist_rows = 0, stat_part_rows = 0;
  // The last tree's splits in the order they were made (serial.rs:2003 last_splits):
  // the split leaf, its routing and the new right leaf, replayed on out-of-bag rows.
  struct SplitRecord {
    int leaf;
    SplitRule rule;
    int right;
  };
  std::vector<SplitRecord> splits;

  ~Learner() { delete disc; }

  const int32_t* idx_of(size_t leaf) const { return indices.data() + leaf_begin[leaf]; }

  // construct_histograms (serial.rs:891-1065) for dense single-value groups.
  void construct_hist(const std::vector<char>& used, size_t leaf, size_t slot, const float* g, const float* h) {
    size_t cnt = leaf_cnt[leaf];
    const int32_t* ix = cnt < num_data ? idx_of(leaf) : nullptr;
    stat_hist_rows += (int64_t)cnt;
    const float* pg;
    const float* ph;
    if (ix) {
      og.resize(cnt);
      for (size_t i = 0; i < cnt; ++i) og[i] = g[ix[i]];
      if (!const_hess) {
        oh.resize(cnt);
        for (size_t i = 0; i < cnt; ++i) oh[i] = h[ix[i]];
      }
      pg = og.data();
      ph = const_hess ? h : oh.data();
    } else {
      pg = g;
      ph = h;
    }
    double* hist = pool.slots[slot].data();
    std::vector<size_t> jobs;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      bool u = false;
      for (int f = groups[gi].feature_start; f < groups[gi].feature_start + groups[gi].feature_cnt; ++f) u |= used[(size_t)f] != 0;
      if (u) jobs.push_back(gi);
    }
    const double h0 = (double)h[0];
#pragma omp parallel for schedule(dynamic, 1) if (num_threads > 1 && jobs.size() > 1 && cnt * jobs.size() >= 4096) num_threads(num_threads)
    for (long j = 0; j < (long)jobs.size(); ++j) {
      size_t gi = jobs[(size_t)j];
      const Group& gr = groups[gi];
      double* out = hist + group_start[gi] * 2;
      std::fill(out, out + (size_t)gr.num_total_bin * 2, 0.0);
      const uint8_t* col = gr.bins;
      if (ix) {
        for (size_t i = 0; i < cnt; ++i) {
          size_t ti = (size_t)col[ix[i]] << 1;
          out[ti] += (double)pg[i];
          out[ti + 1] += const_hess ? 1.0 : (double)ph[i];
        }
      } else {
        for (size_t i = 0; i < cnt; ++i) {
          size_t ti = (size_t)col[i] << 1;
          out[ti] += (double)pg[i];
          out[ti + 1] += const_hess ? 1.0 : (double)ph[i];
        }
      }
      if (const_hess)
        for (size_t i = 0; i < (size_t)gr.num_total_bin * 2; i += 2) out[i + 1] *= h0;
    }
  }

  // construct_int_histograms (serial.rs:1074-1164): single-threaded, as written.
  void construct_int_hist(const std::vector<char>& used, size_t leaf, size_t slot) {
    size_t cnt = leaf_cnt[leaf];
    const int32_t* ix = cnt < num_data ? idx_of(leaf) : nullptr;
    stat_hist_rows += (int64_t)cnt;
    int64_t* hist = pool.int_slots[slot].data();
    const int16_t* pk = disc->packed.data();
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      bool u = false;
      for (int f = groups[gi].feature_start; f < groups[gi].feature_start + groups[gi].feature_cnt; ++f) u |= used[(size_t)f] != 0;
      if (!u) continue;
      const Group& gr = groups[gi];
      int64_t* out = hist + group_start[gi] * 2;
      std::fill(out, out + (size_t)gr.num_total_bin * 2, (int64_t)0);
      const uint8_t* col = gr.bins;
      for (size_t i = 0; i < cnt; ++i) {
        size_t row = ix ? (size_t)ix[i] : i;
        int16_t p = pk[row];
        int64_t gg = (int64_t)(int8_t)(p >> 8);
        int64_t hh = (int64_t)(uint8_t)(p & 0xff);
        size_t b = (size_t)col[row] * 2;
        out[b] += gg;
        out[b + 1] += hh;
      }
    }
  }

  // fix_slice (serial.rs:2205-2217) and fix_slice_int (2221-2238)
  void fix_f64(size_t slot, int f, double sg, double sh) {
    int mfb = feats[(size_t)f].most_freq_bin;
    if (mfb <= 0) return;
    int nb = feats[(size_t)f].meta.num_bin;
    double* d = pool.slots[slot].data() + feat_off[(size_t)f] * 2;
    d[mfb << 1] = sg;
    d[(mfb << 1) + 1] = sh;
    for (int i = 0; i < nb; ++i) {
      if (i != mfb) {
        d[mfb << 1] -= d[i << 1];
        d[(mfb << 1) + 1] -= d[(i << 1) + 1];
      }
    }
  }
  void fix_int(size_t slot, int f, int64_t tg, int64_t th) {
    int mfb = feats[(size_t)f].most_freq_bin;
    if (mfb <= 0) return;
    int nb = feats[(size_t)f].meta.num_bin;
    int64_t* d = pool.int_slots[slot].data() + feat_off[(size_t)f] * 2;
    d[(size_t)mfb * 2] = tg;
    d[(size_t)mfb * 2 + 1] = th;
    for (int i = 0; i < nb; ++i) {
      if (i != mfb) {
        d[(size_t)mfb * 2] -= d[(size_t)i * 2];
        d[(size_t)mfb * 2 + 1] -= d[(size_t)i * 2 + 1];
      }
    }
  }

  // search_feature / search_feature_int (serial.rs:2335-2433), no monotone penalty.
  std::pair<bool, SplitInfo> search(size_t slot, int f, const LeafSplits& ls, double parent) {
    SplitInfo s;
    const Feature& ft = feats[(size_t)f];
    bool ok;
    if (quant) {
      const int64_t* h = pool.int_slots[slot].data() + feat_off[(size_t)f] * 2;
      int64_t tg = ls.int_sum >> 32, th = ls.int_sum & 0xffffffffLL;
      if (ft.meta.bin_type == BK_BIN_CATEGORICAL)
        ok = find_best_threshold_categorical_int(h, ft.meta, sp, cp, tg, th, grad_scale, hess_scale, ls.num_data,
                                                 parent, s);
      else
        ok = find_best_threshold_numerical_int(h, ft.meta, sp, tg, th, grad_scale, hess_scale, ls.num_data, parent, s);
    } else {
      const double* h = pool.slots[slot].data() + feat_off[(size_t)f] * 2;
      if (ft.meta.bin_type == BK_BIN_CATEGORICAL)
        ok = find_best_threshold_categorical(h, ft.meta, sp, cp, ls.sum_g, ls.sum_h, ls.num_data, parent, s);
      else
        ok = find_best_threshold_numerical(h, ft.meta, sp, ls.sum_g, ls.sum_h, ls.num_data, parent, s);
    }
    s.feature = ft.real_index;
    return {ok, s};
  }

  double parent_output(const Tree& t, const LeafSplits& ls) const {
    if (t.num_leaves == 1) return leaf_output_root(sp, ls.sum_g, ls.sum_h, ls.num_data);
    return ls.weight;
  }

  // before_find_best_split (serial.rs:779-825)
  bool before_find_best_split(const Tree& t, int left_leaf, int right_leaf) {
    if (max_depth > 0 && t.leaf_depth[(size_t)left_leaf] >= max_depth) {
      best[(size_t)left_leaf].gain = K_MIN_SCORE;
      if (right_leaf >= 0) best[(size_t)right_leaf].gain = K_MIN_SCORE;
      return false;
    }
    int nl = left_leaf >= 0 ? (int)leaf_cnt[(size_t)left_leaf] : 0;
    int nr = right_leaf >= 0 ? (int)leaf_cnt[(size_t)right_leaf] : 0;
    if (nr < sp.min_data_in_leaf * 2 && nl < sp.min_data_in_leaf * 2) {
      best[(size_t)left_leaf].gain = K_MIN_SCORE;
      if (right_leaf >= 0) best[(size_t)right_leaf].gain = K_MIN_SCORE;
      return false;
    }
    parent_slot = -1;
    if (right_leaf < 0) {
      smaller_slot = pool.get((size_t)left_leaf).first;
      larger_slot = -1;
    } else if (nl < nr) {
      auto lg = pool.get((size_t)left_leaf);
      if (lg.second) parent_slot = (long)lg.first;
      larger_slot = (long)lg.first;
      pool.move_leaf((size_t)left_leaf, (size_t)right_leaf);
      smaller_slot = pool.get((size_t)left_leaf).first;
    } else {
      auto lg = pool.get((size_t)left_leaf);
      if (lg.second) parent_slot = (long)lg.first;
      larger_slot = (long)lg.first;
      smaller_slot = pool.get((size_t)right_leaf).first;
    }
    return true;
  }

  // find_best_splits (serial.rs:830-887) + find_best_splits_from_histograms (1322-1566)
  // + find_best_splits_int (1212-1300).
  void find_best_splits(const Tree& t, const float* g, const float* h) {
    std::vector<char> used((size_t)nf, 0);
    for (int f = 0; f < nf; ++f) {
      if (!cols.bytree[(size_t)f]) continue;  // outside the tree's sample: not touched
      if (parent_slot >= 0 && !pool.splittable[(size_t)parent_slot][(size_t)f]) {
        pool.splittable[smaller_slot][(size_t)f] = 0;
        continue;
      }
      used[(size_t)f] = 1;
    }
    bool use_subtract = parent_slot >= 0;
    if (quant) construct_int_hist(used, (size_t)smaller.leaf_index, smaller_slot);
    else construct_hist(used, (size_t)smaller.leaf_index, smaller_slot, g, h);
    if (larger_slot >= 0 && !use_subtract && larger.leaf_index >= 0) {
      if (quant) construct_int_hist(used, (size_t)larger.leaf_index, (size_t)larger_slot);
      else construct_hist(used, (size_t)larger.leaf_index, (size_t)larger_slot, g, h);
    }
    // by-node samples, smaller leaf first (serial.rs:1329-1337)
    const std::vector<char> smask = cols.by_node(branch[(size_t)smaller.leaf_index]);
    const std::vector<char> lmask =
        larger.leaf_index >= 0 ? cols.by_node(branch[(size_t)larger.leaf_index]) : std::vector<char>((size_t)nf, 0);
    double sp_out = parent_output(t, smaller);
    double lp_out = larger.leaf_index >= 0 ? parent_output(t, larger) : 0.0;
    long ls = larger.leaf_index >= 0 ? larger_slot : -1;
    // phase 1: fix and subtract
    for (int f = 0; f < nf; ++f) {
      if (!used[(size_t)f]) continue;
      if (quant) {
        fix_int(smaller_slot, f, smaller.int_sum >> 32, smaller.int_sum & 0xffffffffLL);
        if (ls >= 0) {
          if (use_subtract) {
            int64_t* s = pool.int_slots[smaller_slot].data();
            int64_t* l = pool.int_slots[(size_t)ls].data();
            for (size_t i = feat_off[(size_t)f]; i < feat_off[(size_t)f] + feat_len[(size_t)f]; ++i) {
              l[i * 2] -= s[i * 2];
              l[i * 2 + 1] -= s[i * 2 + 1];
            }
          } else {
            fix_int((size_t)ls, f, larger.int_sum >> 32, larger.int_sum & 0xffffffffLL);
          }
        }
      } else {
        fix_f64(smaller_slot, f, smaller.sum_g, smaller.sum_h);
        if (ls >= 0) {
          if (use_subtract) {
            double* s = pool.slots[smaller_slot].data();
            double* l = pool.slots[(size_t)ls].data();
            for (size_t i = feat_off[(size_t)f] * 2; i < (feat_off[(size_t)f] + feat_len[(size_t)f]) * 2; ++i) l[i] -= s[i];
          } else {
            fix_f64((size_t)ls, f, larger.sum_g, larger.sum_h);
          }
        }
      }
    }
    // phase 2: search, parallel over features as the reference does above 1024 bins
    std::vector<std::pair<bool, SplitInfo>> rs((size_t)nf), rl((size_t)nf);
#pragma omp parallel for schedule(dynamic, 1) if (num_threads > 1) num_threads(num_threads)
    for (int f = 0; f < nf; ++f) {
      if (!used[(size_t)f]) continue;
      rs[(size_t)f] = search(smaller_slot, f, smaller, sp_out);
      if (ls >= 0) rl[(size_t)f] = search((size_t)ls, f, larger, lp_out);
    }
    SplitInfo sb, lb;
    for (int f = 0; f < nf; ++f) {
      if (!used[(size_t)f]) continue;
      pool.splittable[smaller_slot][(size_t)f] = rs[(size_t)f].first;
      if (smask[(size_t)f] && rs[(size_t)f].second.better_than(sb)) sb = rs[(size_t)f].second;
      if (ls >= 0) {
        pool.splittable[(size_t)ls][(size_t)f] = rl[(size_t)f].first;
        if (lmask[(size_t)f] && rl[(size_t)f].second.better_than(lb)) lb = rl[(size_t)f].second;
      }
    }
    best[(size_t)smaller.leaf_index] = sb;
    if (larger.leaf_index >= 0) best[(size_t)larger.leaf_index] = lb;
  }

  // DataPartition::split (data_partition.rs:124-175), sequential stable routing.
  void partition_split(size_t leaf, const SplitRule& rule, size_t right_leaf) {
    size_t begin = leaf_begin[leaf], cnt = leaf_cnt[leaf];
    stat_part_rows += (int64_t)cnt;
    lbuf.clear();
    rbuf.clear();
    const int32_t* ix = indices.data() + begin;
    for (size_t i = 0; i < cnt; ++i) {
      int32_t r = ix[i];
      if (rule.route_left(rule.col[r])) lbuf.push_back(r);
      else rbuf.push_back(r);
    }
    std::copy(lbuf.begin(), lbuf.end(), indices.begin() + (long)begin);
    std::copy(rbuf.begin(), rbuf.end(), indices.begin() + (long)(begin + lbuf.size()));
    leaf_cnt[leaf] = lbuf.size();
    leaf_begin[right_leaf] = begin + lbuf.size();
    leaf_cnt[right_leaf] = cnt - lbuf.size();
  }

  // split (serial.rs:1685-1876), numerical and categorical, no monotone/CEGB.
  std::pair<int, int> split(Tree& t, int best_leaf) {
    SplitInfo info = best[(size_t)best_leaf];
    int inner = -1;
    for (int f = 0; f < nf; ++f)
      if (feats[(size_t)f].real_index == info.feature) inner = f;
    const Feature& ft = feats[(size_t)inner];
    const Group& gr = groups[(size_t)ft.group];
    int next_leaf = t.num_leaves;
    float gain = (float)(info.gain + min_gain_to_split);
    int right;
    if (ft.meta.bin_type == BK_BIN_CATEGORICAL) {
      std::vector<uint32_t> bins(info.cat_threshold.begin(), info.cat_threshold.begin() + info.num_cat_threshold);
      std::vector<uint32_t> real;
      for (uint32_t b : bins) real.push_back((uint32_t)ft.bin2cat[b]);
      std::vector<uint32_t> bitset_real = construct_bitset(real);
      std::vector<uint32_t> bitset_bin = construct_bitset(bins);
      SplitRule rule = rule_categorical(gr, ft, bitset_bin);
      partition_split((size_t)best_leaf, rule, (size_t)next_leaf);
      splits.push_back(SplitRecord{best_leaf, rule, next_leaf});
      info.left_count = (int)leaf_cnt[(size_t)best_leaf];
      info.right_count = (int)leaf_cnt[(size_t)next_leaf];
      int node = t.split_common(best_leaf, inner, info.feature, info.left_output, info.right_output, info.left_count,
                                info.right_count, info.left_sum_hessian, info.right_sum_hessian, gain);
      t.decision_type[(size_t)node] = (int8_t)(1 | (ft.meta.missing_type << 2));
      t.threshold_in_bin[(size_t)node] = (uint32_t)t.num_cat;
      t.threshold[(size_t)node] = (double)t.num_cat;
      t.num_cat += 1;
      t.cat_boundaries.push_back(t.cat_boundaries.back() + (int)bitset_real.size());
      t.cat_threshold.insert(t.cat_threshold.end(), bitset_real.begin(), bitset_real.end());
      t.num_leaves += 1;
      right = t.num_leaves - 1;
    } else {
      double thd = ft.upper[info.threshold];
      SplitRule rule = rule_numerical(gr, ft, info.threshold, info.default_left);
      partition_split((size_t)best_leaf, rule, (size_t)next_leaf);
      splits.push_back(SplitRecord{best_leaf, rule, next_leaf});
      info.left_count = (int)leaf_cnt[(size_t)best_leaf];
      info.right_count = (int)leaf_cnt[(size_t)next_leaf];
      int node = t.split_common(best_leaf, inner, info.feature, info.left_output, info.right_output, info.left_count,
                                info.right_count, info.left_sum_hessian, info.right_sum_hessian, gain);
      int8_t dt = 0;
      if (info.default_left) dt |= 2;
      dt = (int8_t)((dt & 3) | (ft.meta.missing_type << 2));
      t.decision_type[(size_t)node] = dt;
      t.threshold_in_bin[(size_t)node] = info.threshold;
      t.threshold[(size_t)node] = thd;
      t.num_leaves += 1;
      right = t.num_leaves - 1;
    }
    // the branch features of both children: the parent's path and this split's
    // real feature (serial.rs:1719-1722)
    branch[(size_t)right] = branch[(size_t)best_leaf];
    branch[(size_t)right].push_back(info.feature);
    branch[(size_t)best_leaf].push_back(info.feature);
    LeafSplits l{best_leaf, info.left_count, info.left_sum_gradient, info.left_sum_hessian, info.left_output,
                 info.left_int_sum};
    LeafSplits r{right, info.right_count, info.right_sum_gradient, info.right_sum_hessian, info.right_output,
                 info.right_int_sum};
    if (info.left_count < info.right_count) {
      smaller = l;
      larger = r;
    } else {
      smaller = r;
      larger = l;
    }
    best[(size_t)best_leaf] = info;
    return {best_leaf, right};
  }

  // before_train (serial.rs:658-776). rows: the bagged subset, ascending, or null for
  // all rows (DataPartition::init_with / init, data_partition.rs:91-107). The
  // discretizer always sees every row; the root sums run over the bag in order.
  void before_train(const float* g, const float* h, const int32_t* rows, size_t nrows) {
    pool.reset_map();
    leaf_begin.assign((size_t)num_leaves_cfg, 0);
    leaf_cnt.assign((size_t)num_leaves_cfg, 0);
    if (rows) {
      leaf_cnt[0] = nrows;
      std::copy(rows, rows + nrows, indices.begin());
    } else {
      leaf_cnt[0] = num_data;
      for (size_t i = 0; i < num_data; ++i) indices[i] = (int32_t)i;
    }
    for (auto& s : best) s.reset();
    splits.clear();
    branch.assign((size_t)num_leaves_cfg, std::vector<int>());
    cols.reset_by_tree();
    stat_hist_rows = stat_part_rows = 0;
    const size_t cnt = leaf_cnt[0];
    double sg = 0.0, sh = 0.0;
    int64_t isum = 0;
    if (quant) {
      bool ok = stochastic ? disc->stochastic_round(g, h) : disc->nearest(g, h);
      if (ok) {
        grad_scale = disc->gradient_scale;
        hess_scale = disc->hessian_scale;
      }
      const int16_t* pk = disc->packed.data();
      for (size_t j = 0; j < cnt; ++j) {
        size_t i = rows ? (size_t)rows[j] : j;
        int16_t p = pk[i];
        int8_t gg = (int8_t)(p >> 8);
        uint8_t hh = (uint8_t)(p & 0xff);
        sg += (double)gg * grad_scale;
        sh += (double)hh * hess_scale;
        isum += (int64_t)(((uint64_t)(int64_t)gg << 32) | (uint64_t)hh);
      }
    } else {
      for (size_t j = 0; j < cnt; ++j) {
        size_t i = rows ? (size_t)rows[j] : j;
        sg += (double)g[i];
        sh += (double)h[i];
      }
    }
    double carried = smaller.weight;
    smaller = LeafSplits{0, (int)cnt, sg, sh, carried, isum};
    larger = LeafSplits{};
    root_sums = smaller;
  }
  LeafSplits root_sums;  // the root's sums of the last tree (serial.rs:741-761)

  // grow_tree (serial.rs:578-628) + renew_int_grad_tree_output (634-655)
  Tree train(const float* g, const float* h, const int32_t* rows, size_t nrows) {
    before_train(g, h, rows, nrows);
    Tree t(num_leaves_cfg);
    t.set_leaf_output(0, leaf_output_root(sp, smaller.sum_g, smaller.sum_h, smaller.num_data));
    int left_leaf = 0, right_leaf = -1;
    for (int step = 0; step < num_leaves_cfg - 1; ++step) {
      if (before_find_best_split(t, left_leaf, right_leaf)) find_best_splits(t, g, h);
      size_t bl = argmax_split(best);
      if (best[bl].gain <= 0.0) break;
      auto lr = split(t, (int)bl);
      left_leaf = lr.first;
      right_leaf = lr.second;
    }
    if (quant && renew) {
      for (int leaf = 0; leaf < t.num_leaves; ++leaf) {
        const int32_t* ix = idx_of((size_t)leaf);
        size_t c = leaf_cnt[(size_t)leaf];
        double sgr = 0.0, shs = 0.0;
        for (size_t i = 0; i < c; ++i) {
          sgr += (double)g[ix[i]];
          shs += (double)h[ix[i]];
        }
        double o = leaf_output<true, true, false>(sgr, shs, sp.l1, sp.l2, sp.max_delta_step, sp.path_smooth, (int)c,
                                                  0.0);
        t.set_leaf_output(leaf, o);
      }
    }
    return t;
  }
};

void export_tree(const Tree& t, bk_tree_out* o) {
  o->num_leaves = t.num_leaves;
  o->num_cat = t.num_cat;
  size_t ni = (size_t)(t.num_leaves - 1), nl = (size_t)t.num_leaves;
  for (size_t i = 0; i < ni; ++i) {
    o->split_feature_inner[i] = t.split_feature_inner[i];
    o->split_feature[i] = t.split_feature[i];
    o->threshold_in_bin[i] = t.threshold_in_bin[i];
    o->threshold[i] = t.threshold[i];
    o->decision_type[i] = t.decision_type[i];
    o->left_child[i] = t.left_child[i];
    o->right_child[i] = t.right_child[i];
    o->split_gain[i] = t.split_gain[i];
    o->internal_value[i] = t.internal_value[i];
    o->internal_weight[i] = t.internal_weight[i];
    o->internal_count[i] = t.internal_count[i];
  }
  for (size_t i = 0; i < nl; ++i) {
    o->leaf_value[i] = t.leaf_value[i];
    o->leaf_weight[i] = t.leaf_weight[i];
    o->leaf_count[i] = t.leaf_count[i];
    o->leaf_parent[i] = t.leaf_parent[i];
    o->leaf_depth[i] = t.leaf_depth[i];
  }
  for (size_t i = 0; i < t.cat_boundaries.size() && i < nl; ++i) o->cat_boundaries[i] = t.cat_boundaries[i];
  if ((int)t.cat_threshold.size() > o->cat_threshold_cap) throw std::runtime_error("cat_threshold capacity too small");
  for (size_t i = 0; i < t.cat_threshold.size(); ++i) o->cat_threshold[i] = t.cat_threshold[i];
  o->cat_threshold_len = (int)t.cat_threshold.size();
}

}  // namespace

// ===========================================================================
// C API of the oracle (test infrastructure)
// ===========================================================================
extern "C" {

struct bko_learner {
  Learner L;
};

const char* bko_last_error(void) { return g_err.c_str(); }

int bko_learner_create(const bk_dataset_view* ds, const bk_learner_params* p, int is_constant_hessian,
                       bko_learner** out) {
  try {
    if (!ds || !p || !out) throw std::invalid_argument("null argument");
    if (p->extra_trees) throw std::invalid_argument("extra_trees is not supported");
    auto* h = new bko_learner();
    Learner& L = h->L;
    L.num_data = (size_t)ds->num_data;
    L.nf = ds->num_features;
    L.num_leaves_cfg = p->num_leaves;
    L.max_depth = p->max_depth;
    L.num_threads = p->num_threads > 0 ? p->num_threads : 1;
    L.sp = SplitParams{p->lambda_l1, p->lambda_l2, p->max_delta_step, p->path_smooth, p->min_data_in_leaf,
                       p->min_sum_hessian_in_leaf, p->min_gain_to_split};
    L.cp = CatParams{p->cat_smooth, p->cat_l2, p->max_cat_threshold, p->max_cat_to_onehot, p->min_data_per_group};
    L.min_gain_to_split = p->min_gain_to_split;
    L.const_hess = is_constant_hessian != 0;
    L.groups.resize((size_t)ds->num_groups);
    L.group_start.resize((size_t)ds->num_groups);
    size_t cur = 0;
    for (int g = 0; g < ds->num_groups; ++g) {
      Group& G = L.groups[(size_t)g];
      G.bins = ds->group_bins[g];
      G.num_total_bin = ds->group_num_total_bin[g];
      G.feature_start = ds->group_feature_start[g];
      G.feature_cnt = ds->group_feature_cnt[g];
      const int32_t* bo = ds->group_bin_offsets + G.feature_start + g;
      G.bin_offsets.assign(bo, bo + G.feature_cnt + 1);
      if (G.num_total_bin > 256) throw std::invalid_argument("group with more than 256 bins");
      L.group_start[(size_t)g] = cur;
      for (int64_t i = 0; i < ds->num_data; ++i)
        if ((int)G.bins[i] >= G.num_total_bin) throw std::invalid_argument("bin outside its group range");
      cur += (size_t)G.num_total_bin;
    }
    L.total_bins = cur;
    L.feats.resize((size_t)L.nf);
    L.feat_off.resize((size_t)L.nf);
    L.feat_len.resize((size_t)L.nf);
    for (int g = 0; g < ds->num_groups; ++g) {
      const Group& G = L.groups[(size_t)g];
      for (int s = 0; s < G.feature_cnt; ++s) {
        int f = G.feature_start + s;
        Feature& F = L.feats[(size_t)f];
        F.group = g;
        F.sub = s;
        F.most_freq_bin = ds->feature_most_freq_bin[f];
        F.real_index = ds->feature_real_index[f];
        F.meta.num_bin = ds->feature_num_bin[f];
        F.meta.missing_type = ds->feature_missing_type[f];
        F.meta.default_bin = ds->feature_default_bin[f];
        F.meta.bin_type = ds->feature_bin_type[f];
        F.meta.offset = F.most_freq_bin == 0 ? 1 : 0;
        F.meta.penalty = ds->feature_penalty ? ds->feature_penalty[f] : 1.0;
        if (F.meta.bin_type == BK_BIN_CATEGORICAL) {
          F.bin2cat.assign(ds->feature_bin_2_categorical[f], ds->feature_bin_2_categorical[f] + F.meta.num_bin);
        } else {
          F.upper.assign(ds->feature_bin_upper_bound[f], ds->feature_bin_upper_bound[f] + F.meta.num_bin);
        }
        L.feat_off[(size_t)f] = L.group_start[(size_t)g] + (size_t)G.bin_offsets[(size_t)s];
        L.feat_len[(size_t)f] = (size_t)(F.meta.num_bin - F.meta.offset);
      }
    }
    size_t nl = (size_t)p->num_leaves;
    size_t cache = nl;
    if (p->histogram_pool_size > 0.0) {
      size_t bytes = 0;
      for (auto& F : L.feats) bytes += (size_t)F.meta.num_bin * 16;
      size_t c = (size_t)(p->histogram_pool_size * 1024.0 * 1024.0 / (double)bytes);
      cache = std::min(std::max(c, (size_t)2), nl);
    }
    L.quant = p->use_quantized_grad && p->num_grad_quant_bins >= 2 && p->num_grad_quant_bins <= 127;
    L.pool.init(nl, std::max(L.total_bins, (size_t)1), (size_t)L.nf, cache, L.quant);
    L.best.assign(nl, SplitInfo());
    L.indices.resize(L.num_data);
    if (!(p->feature_fraction > 0.0 && p->feature_fraction <= 1.0) ||
        !(p->feature_fraction_bynode > 0.0 && p->feature_fraction_bynode <= 1.0))
      throw std::invalid_argument("feature_fraction and feature_fraction_bynode must be in (0, 1]");
    {
      std::vector<int> valid((size_t)L.nf);
      for (int f = 0; f < L.nf; ++f) valid[(size_t)f] = f;
      std::stable_sort(valid.begin(), valid.end(),
                       [&](int a, int b) { return L.feats[(size_t)a].real_index < L.feats[(size_t)b].real_index; });
      L.cols.init(valid, L.nf, p->feature_fraction, p->feature_fraction_bynode, p->feature_fraction_seed);
      L.cols.real_of.resize((size_t)L.nf);
      for (int f = 0; f < L.nf; ++f) L.cols.real_of[(size_t)f] = L.feats[(size_t)f].real_index;
      for (int i = 0; i < p->num_interaction_constraints; ++i) {
        std::vector<int> c(p->interaction_constraint_features + p->interaction_constraint_offsets[i],
                           p->interaction_constraint_features + p->interaction_constraint_offsets[i + 1]);
        std::sort(c.begin(), c.end());
        L.cols.constraints.push_back(c);
      }
    }
    if (L.quant) {
      L.disc = new Discretizer(L.num_data, p->num_grad_quant_bins, L.const_hess);
      L.stochastic = p->stochastic_rounding != 0;
      L.renew = p->quant_train_renew_leaf != 0;
      if (L.stochastic) L.disc->init_stochastic((uint32_t)p->seed, (size_t)L.num_threads);
    }
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int bko_learner_train(bko_learner* h, const float* g, const float* hs, const int32_t* rows, int64_t nrows,
                      bk_tree_out* out) {
  try {
    if (rows) {
      if (nrows < 0 || (size_t)nrows > h->L.num_data) throw std::invalid_argument("bag larger than the dataset");
      for (int64_t i = 0; i < nrows; ++i)
        if (rows[i] < 0 || (size_t)rows[i] >= h->L.num_data || (i > 0 && rows[i] <= rows[i - 1]))
          throw std::invalid_argument("bagged rows must be ascending and inside the dataset");
    }
    Tree t = h->L.train(g, hs, rows, rows ? (size_t)nrows : 0);
    export_tree(t, out);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int bko_learner_add_prediction_to_score(bko_learner* h, const bk_tree_out* t, double* score) {
  Learner& L = h->L;
  if (t->num_leaves <= 1) return 0;
  for (int leaf = 0; leaf < t->num_leaves; ++leaf) {
    const int32_t* ix = L.idx_of((size_t)leaf);
    for (size_t i = 0; i < L.leaf_cnt[(size_t)leaf]; ++i) score[ix[i]] += t->leaf_value[leaf];
  }
  return 0;
}

// add_out_of_bag_prediction (serial.rs:1995) through add_prediction_for_rows
// (serial.rs:2009): the last tree's splits replayed, in order, on a fresh partition of
// `rows` with the training routing; each row then gets its leaf's value.
int bko_learner_add_out_of_bag_prediction(bko_learner* h, const bk_tree_out* t, const int32_t* rows, int64_t n,
                                          double* score) {
  Learner& L = h->L;
  if (t->num_leaves <= 1 || n <= 0) return 0;
  std::vector<int32_t> ix(rows, rows + n), lb, rb;
  std::vector<size_t> begin((size_t)t->num_leaves, 0), cnt((size_t)t->num_leaves, 0);
  cnt[0] = (size_t)n;
  for (const auto& s : L.splits) {
    size_t b = begin[(size_t)s.leaf], c = cnt[(size_t)s.leaf];
    lb.clear();
    rb.clear();
    for (size_t i = b; i < b + c; ++i) (s.rule.route_left(s.rule.col[ix[i]]) ? lb : rb).push_back(ix[i]);
    std::copy(lb.begin(), lb.end(), ix.begin() + (long)b);
    std::copy(rb.begin(), rb.end(), ix.begin() + (long)(b + lb.size()));
    cnt[(size_t)s.leaf] = lb.size();
    begin[(size_t)s.right] = b + lb.size();
    cnt[(size_t)s.right] = rb.size();
  }
  for (int leaf = 0; leaf < t->num_leaves; ++leaf)
    for (size_t i = begin[(size_t)leaf]; i < begin[(size_t)leaf] + cnt[(size_t)leaf]; ++i)
      score[ix[i]] += t->leaf_value[leaf];
  return 0;
}

int bko_learner_leaf_rows(bko_learner* h, int leaf, int32_t* rows, int64_t* n) {
  Learner& L = h->L;
  const int32_t* ix = L.idx_of((size_t)leaf);
  *n = (int64_t)L.leaf_cnt[(size_t)leaf];
  std::copy(ix, ix + L.leaf_cnt[(size_t)leaf], rows);
  return 0;
}

int bko_learner_stats(bko_learner* h, int64_t* hist_rows, int64_t* part_rows) {
  *hist_rows = h->L.stat_hist_rows;
  *part_rows = h->L.stat_part_rows;
  return 0;
}

// Quantized state of the last tree: packed int8 pairs and scales (for GPU cross-checks).
int bko_learner_quant_state(bko_learner* h, int16_t* packed, double* gs, double* hs) {
  Learner& L = h->L;
  if (!L.disc) return -1;
  if (packed) std::copy(L.disc->packed.begin(), L.disc->packed.end(), packed);
  *gs = L.grad_scale;
  *hs = L.hess_scale;
  return 0;
}

// The last tree's root sums as before_train computed them (serial.rs:741-761):
// sequential f64 sums in row order, and the packed integer sum (quantized).
int bko_learner_root_sums(bko_learner* h, double* sum_g, double* sum_h, int64_t* int_sum) {
  *sum_g = h->L.root_sums.sum_g;
  *sum_h = h->L.root_sums.sum_h;
  *int_sum = h->L.root_sums.int_sum;
  return 0;
}

// Threading::BlockInfo (gradient_discretizer/mod.rs:19-29), as the noise layout uses it.
void bko_block_info(int64_t num_threads, int64_t cnt, int64_t min_cnt, int64_t* n_block, int64_t* size) {
  size_t nb = 0, bs = 0;
  block_info((size_t)num_threads, (size_t)cnt, (size_t)min_cnt, &nb, &bs);
  *n_block = (int64_t)nb;
  *size = (int64_t)bs;
}

// histogram_cache_capacity (serial.rs:250), reset_histogram_pool_config (256-264)
// and continue_after_dataset_reset (266-272) for learners without extra trees.
int64_t bko_learner_histogram_cache_capacity(bko_learner* h) { return (int64_t)h->L.pool.size(); }
void bko_learner_reset_histogram_pool_config(bko_learner* h) { h->L.cols.reset_by_tree(); }
void bko_learner_continue_after_dataset_reset(bko_learner* h, const bko_learner* prev) {
  h->L.cols.rnd = prev->L.cols.rnd;
  h->L.cols.reset_by_tree();
}

void bko_learner_free(bko_learner* h) { delete h; }

// ---- primitive entry points, for the known-answer tests -------------------
void* bko_mt_new(uint32_t seed) { return new Mt19937(seed); }
uint32_t bko_mt_next(void* e) { return ((Mt19937*)e)->next_u32(); }
double bko_mt_uniform_real(void* e) { return ((Mt19937*)e)->uniform_real_01(); }
uint32_t bko_mt_uniform_int(void* e, uint32_t n) { return ((Mt19937*)e)->uniform_int_inclusive(n); }
void bko_mt_free(void* e) { delete (Mt19937*)e; }

// Discretize once: mode 0 nearest, 1 stochastic (seeded with seed / num_threads,
// `draws` earlier trees consumed first). Returns 0 on success, -1 on refusal.
int bko_discretize(const float* g, const float* h, int64_t n, int bins, int const_hess, int mode, uint32_t seed,
                   int num_threads, int draws, int16_t* packed, double* gs, double* hs) {
  Discretizer d((size_t)n, bins, const_hess != 0);
  bool ok;
  if (mode == 1) {
    d.init_stochastic(seed, (size_t)num_threads);
    for (int i = 0; i < draws; ++i) d.use_start_eng.uniform_int_inclusive((uint32_t)n);
    ok = d.stochastic_round(g, h);
  } else {
    ok = d.nearest(g, h);
  }
  if (!ok) return -1;
  std::copy(d.packed.begin(), d.packed.end(), packed);
  *gs = d.gradient_scale;
  *hs = d.hessian_scale;
  return 0;
}

// Noise arrays exactly as init_stochastic lays them out.
void bko_noise(int64_t n, uint32_t seed, int num_threads, double* grad_noise, double* hess_noise) {
  Discretizer d((size_t)n, 4, false);
  d.init_stochastic(seed, (size_t)num_threads);
  std::copy(d.grad_rand.begin(), d.grad_rand.end(), grad_noise);
  std::copy(d.hess_rand.begin(), d.hess_rand.end(), hess_noise);
}

// Ordered f64 histogram of one column (histogram.rs:117-245): out has 2*num_bin doubles.
void bko_hist_f64(const uint8_t* bins, const int32_t* rows, int64_t n, const float* og, const float* oh,
                  int const_hess, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    int64_t r = rows ? rows[i] : i;
    size_t ti = (size_t)bins[r] << 1;
    out[ti] += (double)og[i];
    out[ti + 1] += const_hess ? 1.0 : (double)oh[i];
  }
}

// One numerical/categorical search on a histogram slice. hist: (g,h) pairs of f64
// (mode 0) or int64 (mode 1). Writes [threshold, default_left, left_count,
// right_count, splittable] to iout and [gain, left_output, right_output,
// left_sum_gradient, left_sum_hessian, right_sum_gradient, right_sum_hessian] to dout.
int bko_search(int mode, const void* hist, int num_bin, int missing_type, int default_bin, int most_freq_bin,
               int bin_type, const double* dparams, const int32_t* iparams, double sum_g, double sum_h, int64_t int_g,
               int64_t int_h, double gs, double hs, int num_data, double parent, int32_t* iout, double* dout,
               uint32_t* cat_out, int32_t* num_cat) {
  FeatureMeta m{num_bin, missing_type, most_freq_bin == 0 ? 1 : 0, default_bin, bin_type, 1.0};
  SplitParams p{dparams[0], dparams[1], dparams[2], dparams[3], iparams[0], dparams[4], dparams[5]};
  CatParams c{dparams[6], dparams[7], iparams[1], iparams[2], iparams[3]};
  SplitInfo s;
  bool ok;
  if (mode == 0) {
    const double* h = (const double*)hist;
    ok = bin_type == BK_BIN_CATEGORICAL ? find_best_threshold_categorical(h, m, p, c, sum_g, sum_h, num_data, parent, s)
                                        : find_best_threshold_numerical(h, m, p, sum_g, sum_h, num_data, parent, s);
  } else {
    const int64_t* h = (const int64_t*)hist;
    ok = bin_type == BK_BIN_CATEGORICAL
             ? find_best_threshold_categorical_int(h, m, p, c, int_g, int_h, gs, hs, num_data, parent, s)
             : find_best_threshold_numerical_int(h, m, p, int_g, int_h, gs, hs, num_data, parent, s);
  }
  iout[0] = (int32_t)s.threshold;
  iout[1] = s.default_left;
  iout[2] = s.left_count;
  iout[3] = s.right_count;
  iout[4] = ok;
  dout[0] = s.gain;
  dout[1] = s.left_output;
  dout[2] = s.right_output;
  dout[3] = s.left_sum_gradient;
  dout[4] = s.left_sum_hessian;
  dout[5] = s.right_sum_gradient;
  dout[6] = s.right_sum_hessian;
  *num_cat = s.num_cat_threshold;
  for (size_t i = 0; i < s.cat_threshold.size(); ++i) cat_out[i] = s.cat_threshold[i];
  return 0;
}

}  // extern "C"
