// learner.cpp — host side of the MI355X histogram tree learner and its C ABI.
//
// The control flow mirrors SerialTreeLearner (crates/lgbm-core/src/treelearner/serial.rs):
// histogram pool and LRU (80-162), before_train (658-776), before_find_best_split
// (779-825), find_best_splits (830-887, 1212-1300, 1322-1566), split (1685-1876) and
// renew_int_grad_tree_output (634-655). Row data never leaves the GPU: the host
// keeps O(leaves) bookkeeping (slots, segments, the tree) and reads back one
// small array of per-feature split candidates per step.
//
// Differences of mechanism (same results):
//   - rows are physically partitioned in HBM, so every leaf is a contiguous segment;
//   - histograms carry an exact row count per bin, so both children's sizes are
//     known before their rows move and the smaller child's histogram is built by
//     the partition pass itself;
//   - the quantized root sums (serial.rs:711-740) are taken from exact integer sums
//     when every sequential partial sum is representable in f64 (then the
//     sequential f64 result equals the exact one), else summed on the host in
//     row order.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bankai_gpu.h"
#include "bk_common.h"
#include "kernels.h"

#define BK_HIPC(x)                                                                                          \
  do {                                                                                                      \
    hipError_t e_ = (x);                                                                                    \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " in " #x); \
  } while (0)

namespace bk {

static_assert(sizeof(RuleDev) <= sizeof(((bk_split_record*)nullptr)->rule), "split record too small for a rule");

std::string& last_error() {
  static thread_local std::string e;
  return e;
}

namespace {

const double kEps = (double)1e-15f;
const double kZeroThreshold = (double)1e-35f;
const double kMinScore = -std::numeric_limits<double>::infinity();

// std::mt19937 with libc++'s uniform_int_distribution (gradient_discretizer/mt19937.rs:48-128);
// the learner draws one rotation offset per tree on the host.
struct Mt {
  uint32_t st[624];
  int i = 624;
  explicit Mt(uint32_t seed = 5489u) { seed_with(seed); }
  void seed_with(uint32_t seed) {
    st[0] = seed;
    for (int k = 1; k < 624; ++k) st[k] = 1812433253u * (st[k - 1] ^ (st[k - 1] >> 30)) + (uint32_t)k;
    i = 624;
  }
  uint32_t next() {
    if (i >= 624) {
      for (int k = 0; k < 624; ++k) {
        uint32_t y = (st[k] & 0x80000000u) | (st[(k + 1) % 624] & 0x7fffffffu);
        uint32_t v = st[(k + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        st[k] = v;
      }
      i = 0;
    }
    uint32_t y = st[i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  double real01() {
    double a = (double)next(), b = (double)next();
    const double rp = 4294967296.0;
    return (a + b * rp) / (rp * rp);
  }
  uint32_t uint_incl(uint32_t n) {
    uint32_t rp = n + 1u;
    if (rp == 1u) return 0;
    if (rp == 0u) return next();
    uint32_t w = 32u - (uint32_t)__builtin_clz(rp) - 1u;
    if (rp & (0xffffffffu >> (32u - w))) w += 1u;
    uint32_t mask = w >= 32u ? 0xffffffffu : ((1u << w) - 1u);
    for (;;) {
      uint32_t u = next() & mask;
      if (u < rp) return u;
    }
  }
};

inline double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
// CalculateSplittedLeafOutput<L1=true, MAX=true, SMOOTH=false> with parent 0
// (feature_histogram.rs:76-101), as used for the root and for renewed leaves.
double leaf_output_plain(double g, double h, double l1, double l2, double mds) {
  double reg = std::max(std::fabs(g) - l1, 0.0);
  double ret = -(sgn(g) * reg) / (h + l2);
  if (mds > 0.0 && std::fabs(ret) > mds) ret = sgn(ret) * mds;
  return ret;
}
double leaf_output_root(const bk_learner_params& p, double g, double h) {
  double r = leaf_output_plain(g, h, p.lambda_l1, p.lambda_l2, p.max_delta_step);
  if (r < -std::numeric_limits<double>::max()) return -std::numeric_limits<double>::max();
  if (r > std::numeric_limits<double>::max()) return std::numeric_limits<double>::max();
  return r;
}

// HistogramPool bookkeeping (serial.rs:80-162); the data lives on the device.
struct PoolMap {
  std::vector<std::vector<char>> splittable;
  std::vector<int> leaf_to_slot, slot_to_leaf;
  size_t next_free = 0;
  std::vector<uint64_t> last_used;
  uint64_t clock = 0;
  void init(size_t nl, size_t cache, size_t nf) {
    splittable.assign(cache, std::vector<char>(nf, 1));
    leaf_to_slot.assign(nl, -1);
    slot_to_leaf.assign(cache, -1);
    last_used.assign(cache < nl ? cache : 0, 0);
  }
  size_t size() const { return splittable.size(); }
  void reset() {
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
      for (size_t k = 1; k < last_used.size(); ++k)
        if (last_used[k] < last_used[oldest]) oldest = k;
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

// SplitInfo on the host (split_info.rs:9-85) with the exact child counts.
struct HostSplit {
  SplitOut o{};
  int inner = -1;
  HostSplit() {
    o.gain = kMinScore;
    o.feature = -1;
    o.default_left = 1;
  }
  bool better_than(const HostSplit& b) const {
    double x = o.gain, y = b.o.gain;
    if (x != y) return x > y;
    int fa = o.feature == -1 ? INT32_MAX : o.feature;
    int fb = b.o.feature == -1 ? INT32_MAX : b.o.feature;
    return fa < fb;
  }
  void reset() {
    o.feature = -1;
    o.gain = kMinScore;
  }
};

// Argmax over the leaves' best splits with the reference's rule (split_info.rs:88-96):
// the first maximum under better_than. A tournament tree gives it in O(log L)
// per changed leaf instead of a scan per split. NaN gains (never better, never
// worse) break the ordering a tree needs, so with any NaN present the plain
// scan decides.
struct ArgMax {
  const std::vector<HostSplit>* b = nullptr;
  int n = 0, size = 1, nan_count = 0;
  std::vector<int> t;
  std::vector<char> is_nan;
  bool wins(int i, int j) const {
    if (j < 0) return i >= 0;
    if (i < 0) return false;
    const HostSplit& x = (*b)[(size_t)i];
    const HostSplit& y = (*b)[(size_t)j];
    if (x.better_than(y)) return true;
    if (y.better_than(x)) return false;
    return i < j;
  }
  void init(const std::vector<HostSplit>* best) {
    b = best;
    n = (int)best->size();
    size = 1;
    while (size < n) size <<= 1;
    t.assign((size_t)2 * size, -1);
    is_nan.assign((size_t)n, 0);
    nan_count = 0;
    for (int i = 0; i < n; ++i) {
      t[(size_t)(size + i)] = i;
      note_nan(i);
    }
    for (int k = size - 1; k >= 1; --k) t[(size_t)k] = wins(t[(size_t)2 * k], t[(size_t)2 * k + 1]) ? t[(size_t)2 * k] : t[(size_t)2 * k + 1];
  }
  void note_nan(int i) {
    const double g = (*b)[(size_t)i].o.gain;
    const char now = g != g ? 1 : 0;
    nan_count += now - is_nan[(size_t)i];
    is_nan[(size_t)i] = now;
  }
  void update(int i) {
    note_nan(i);
    for (int k = (size + i) >> 1; k >= 1; k >>= 1)
      t[(size_t)k] = wins(t[(size_t)2 * k], t[(size_t)2 * k + 1]) ? t[(size_t)2 * k] : t[(size_t)2 * k + 1];
  }
  int top() const {
    if (nan_count == 0) return t[1];
    int bl = 0;
    for (int i = 1; i < n; ++i)
      if ((*b)[(size_t)i].better_than((*b)[(size_t)bl])) bl = i;
    return bl;
  }
};

struct LeafSplits {
  int leaf = -1, num_data = 0;
  double sum_g = 0, sum_h = 0, weight = 0;
  long long int_sum = 0;
};

// Training fields of Tree (tree.rs:259-445).
struct Tree {
  int num_leaves = 1, num_cat = 0;
  std::vector<int32_t> left, right, sf_inner, sf, leaf_parent, leaf_depth, leaf_count, internal_count, cat_bound;
  std::vector<uint32_t> th_bin, cat_th;
  std::vector<double> th, leaf_value, leaf_weight, internal_value, internal_weight;
  std::vector<float> gain;
  std::vector<int8_t> dtype;
  explicit Tree(int ml) {
    size_t ni = ml > 1 ? (size_t)ml - 1 : 0, nl = (size_t)ml;
    left.assign(ni, 0);
    right.assign(ni, 0);
    sf_inner.assign(ni, 0);
    sf.assign(ni, 0);
    gain.assign(ni, 0.0f);
    th.assign(ni, 0.0);
    th_bin.assign(ni, 0);
    dtype.assign(ni, 0);
    internal_value.assign(ni, 0.0);
    internal_weight.assign(ni, 0.0);
    internal_count.assign(ni, 0);
    leaf_value.assign(nl, 0.0);
    leaf_weight.assign(nl, 0.0);
    leaf_count.assign(nl, 0);
    leaf_parent.assign(nl, 0);
    if (nl) leaf_parent[0] = -1;
    leaf_depth.assign(nl, 0);
    cat_bound.assign(1, 0);
  }
  static double round0(double v) { return (v >= -kZeroThreshold && v <= kZeroThreshold) ? 0.0 : v; }
  void set_leaf_output(int l, double v) { leaf_value[(size_t)l] = round0(v); }
  int split_common(int leaf, int feature, int real, double lv, double rv, int lc, int rc, double lw, double rw,
                   float g) {
    int node = num_leaves - 1, nl = num_leaves;
    int parent = leaf_parent[(size_t)leaf];
    if (parent >= 0) {
      if (left[(size_t)parent] == ~leaf) left[(size_t)parent] = node;
      else right[(size_t)parent] = node;
    }
    sf_inner[(size_t)node] = feature;
    sf[(size_t)node] = real;
    gain[(size_t)node] = g;
    left[(size_t)node] = ~leaf;
    right[(size_t)node] = ~nl;
    leaf_parent[(size_t)leaf] = node;
    leaf_parent[(size_t)nl] = node;
    internal_weight[(size_t)node] = lw + rw;
    internal_value[(size_t)node] = leaf_value[(size_t)leaf];
    internal_count[(size_t)node] = lc + rc;
    leaf_value[(size_t)leaf] = std::isnan(lv) ? 0.0 : lv;
    leaf_weight[(size_t)leaf] = lw;
    leaf_count[(size_t)leaf] = lc;
    leaf_value[(size_t)nl] = std::isnan(rv) ? 0.0 : rv;
    leaf_weight[(size_t)nl] = rw;
    leaf_count[(size_t)nl] = rc;
    leaf_depth[(size_t)nl] = leaf_depth[(size_t)leaf] + 1;
    leaf_depth[(size_t)leaf] += 1;
    return node;
  }
};

std::vector<uint32_t> bitset_of(const std::vector<uint32_t>& vals) {
  std::vector<uint32_t> r;
  for (uint32_t v : vals) {
    size_t i1 = v / 32;
    if (r.size() < i1 + 1) r.resize(i1 + 1, 0);
    r[i1] |= 1u << (v % 32);
  }
  return r;
}

template <typename T>
T* dmalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  BK_HIPC(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}

// Per-launch timing on the learner's stream: every kernel launched inside a timed
// region gets its own pair of events stamped by the launch itself (no marker
// packets between kernels); the pairs are resolved at the end of a tree.
struct Prof {
  static constexpr size_t kResolveAt = 4096;  // events pending before they are read back
  bool on = false;
  int every = 1;        // time the launches of every `every`-th tree
  int64_t trees = 0;    // train calls since profiling was set
  bool tree_on = true;  // the current tree is timed
  unsigned mask = 0;  // bit k: time launches of kind k
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Rec {
    int kind;
    size_t first, count;  // event pairs first .. first + 2 * count - 1
    double algo, moved;
  };
  std::vector<Rec> pending;
  static constexpr int kKinds = 9;
  double ms[kKinds] = {0};
  int64_t launches[kKinds] = {0};
  double algo[kKinds] = {0}, moved[kKinds] = {0};
  hipEvent_t ev(size_t i) {
    while (pool.size() <= i) {
      hipEvent_t e;
      BK_HIPC(hipEventCreateWithFlags(&e, hipEventDefault));
      pool.push_back(e);
    }
    return pool[i];
  }
  static bool source(void* ctx, void** start, void** stop) {
    Prof* p = (Prof*)ctx;
    const size_t i = p->used;
    p->used += 2;
    *start = (void*)p->ev(i);
    *stop = (void*)p->ev(i + 1);
    p->pending.back().count += 1;
    return true;
  }
  void arm(int kind, double a, double m) {
    pending.push_back(Rec{kind, used, 0, a, m});
    set_event_source(&Prof::source, this);
  }
  void disarm() { set_event_source(nullptr, nullptr); }
  void resolve() {
    for (const Rec& r : pending) {
      if (r.count == 0) continue;
      for (size_t j = 0; j < r.count; ++j) {
        float t = 0.0f;
        BK_HIPC(hipEventElapsedTime(&t, pool[r.first + 2 * j], pool[r.first + 2 * j + 1]));
        ms[r.kind] += t;
      }
      launches[r.kind] += 1;
      algo[r.kind] += r.algo;
      moved[r.kind] += r.moved;
    }
    pending.clear();
    used = 0;
  }
  void reset() {
    for (int k = 0; k < kKinds; ++k) ms[k] = 0, launches[k] = 0, algo[k] = 0, moved[k] = 0;
  }
  // bytes of launches whose work the device decided (known after the tree)
  void add_bytes(int kind, double a, double m) {
    algo[kind] += a;
    moved[kind] += m;
  }
  ~Prof() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

// LightGBM's sampling LCG (utils/random.rs:6-80): x = x * 214013 + 2531011.
struct Lcg {
  uint32_t x = 0;
  uint32_t step() { return x = x * 214013u + 2531011u; }
  float next_float() { return (float)((step() >> 16) & 0x7FFFu) / 32768.0f; }
  int next_int(int lo, int hi) { return (int)(step() & 0x7FFFFFFFu) % (hi - lo) + lo; }
  // k sorted distinct values of 0..n-1 (Random::sample, random.rs:52-79): a
  // selection scan when k is large against n / log2 k, else Floyd's algorithm
  std::vector<int> sample(int n, int k) {
    std::vector<int> out;
    if (k > n || k <= 0) return out;
    out.reserve((size_t)k);
    if (k == n) {
      for (int i = 0; i < n; ++i) out.push_back(i);
    } else if (k > 1 && (double)k > (double)n / std::log2((double)k)) {
      for (int i = 0; i < n; ++i)
        if ((double)next_float() < (double)(k - (int)out.size()) / (double)(n - i)) out.push_back(i);
    } else {
      std::vector<char> taken((size_t)n, 0);
      for (int r = n - k; r < n; ++r) {
        const int v = next_int(0, r + 1);
        taken[(size_t)(taken[(size_t)v] ? r : v)] = 1;
      }
      for (int i = 0; i < n; ++i)
        if (taken[(size_t)i]) out.push_back(i);
    }
    return out;
  }
};

// Feature sampling per tree and per node (ColSampler, treelearner/col_sampler.rs:
// 14-218). One LCG stream: a by-tree draw at set-up and at every tree, by-node draws
// at every search (smaller leaf, then larger). Interaction constraints restrict a
// leaf's features to those allowed by the real features split on along its branch.
struct FeatureSampler {
  Lcg rnd;
  std::vector<int> valid;  // inner features in real-index order
  std::vector<int> in_tree;  // indices into `valid` of the tree's sample
  std::vector<char> bytree;  // per inner feature
  double bynode = 1.0;
  bool per_tree = false;
  int tree_cnt = 0;
  std::vector<std::vector<int>> constraints;  // sorted real feature ids per constraint
  std::vector<int> real_of;                   // real feature id of each inner feature

  static int count(int total, double fraction) {
    return std::max(std::min(1, total), (int)((double)total * fraction + 0.5));
  }
  void init(std::vector<int> order, double ff, double ffn, int32_t seed) {
    valid = std::move(order);
    bytree.assign(valid.size(), 1);
    rnd.x = (uint32_t)seed;
    bynode = ffn;
    per_tree = ff < 1.0;
    tree_cnt = per_tree ? count((int)valid.size(), ff) : (int)valid.size();
    next_tree();
  }
  // the sample of tree `slot` of every `period` trees drawn from the stream (a
  // learner that grows one class tree of each iteration takes the draws of all)
  void next_tree(int slot = 0, int period = 1) {
    if (!per_tree) return;
    for (int j = 0; j < period; ++j) {
      std::vector<int> pick = rnd.sample((int)valid.size(), tree_cnt);
      if (j == slot) in_tree = std::move(pick);
    }
    std::fill(bytree.begin(), bytree.end(), (char)0);
    for (int i : in_tree) bytree[(size_t)valid[(size_t)i]] = 1;
  }
  // by-node masks differ between leaves (sampling or constraints); sampling draws
  bool by_node_active() const { return bynode < 1.0 || !constraints.empty(); }
  bool by_node_draws() const { return bynode < 1.0; }
  // the branch's own features and every constraint holding all of them (every
  // constraint for the root), as allowed_features does (col_sampler.rs:178-201)
  std::vector<int> allowed(const std::vector<int>& branch) const {
    std::vector<int> out(branch);
    for (const auto& c : constraints) {
      bool all = true;
      for (int f : branch) all = all && std::binary_search(c.begin(), c.end(), f);
      if (all) out.insert(out.end(), c.begin(), c.end());
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }
  // get_by_node (col_sampler.rs:111-176) for a leaf whose path split on `branch`
  std::vector<char> by_node(const std::vector<int>& branch) {
    const bool constrained = !constraints.empty();
    const std::vector<int> ok = constrained ? allowed(branch) : std::vector<int>();
    auto admits = [&](int inner) { return !constrained || std::binary_search(ok.begin(), ok.end(), real_of[(size_t)inner]); };
    std::vector<char> m(valid.size(), 0);
    if (bynode >= 1.0) {
      for (size_t f = 0; f < m.size(); ++f) m[f] = admits((int)f) ? 1 : 0;
      return m;
    }
    // candidates (indices into `valid`): the tree's sample or every feature, those
    // the constraints admit; the count comes from the unfiltered set
    std::vector<int> cand;
    int cnt;
    if (per_tree) {
      cnt = count((int)in_tree.size(), bynode);
      for (int i : in_tree)
        if (admits(valid[(size_t)i])) cand.push_back(i);
    } else {
      cnt = count((int)valid.size(), bynode);
      for (int i = 0; i < (int)valid.size(); ++i)
        if (admits(valid[(size_t)i])) cand.push_back(i);
    }
    if (constrained) cnt = std::min(cnt, (int)cand.size());
    for (int p : rnd.sample((int)cand.size(), cnt)) m[(size_t)valid[(size_t)cand[(size_t)p]]] = 1;
    return m;
  }
};

// Bits of a search item's per-feature word: searched (its histogram is scanned and
// its splittable flag kept), and whether it may be the smaller / larger leaf's best
// (the by-node samples).
constexpr int kUseSearch = 1, kUseSmall = 2, kUseLarge = 4, kUseAll = 7;
// Rounds partitioning at most this many rows use the short partition tiles: in
// long tiles such a round is at most 128 blocks for 256 CUs (measured over the
// 500-iteration job at 10.5M rows: 5% less time than long tiles throughout,
// against 3% with the bound at 2M rows and 4.5% at 256K).
constexpr long long kSmallTileRoundRows = 1 << 19;

// A leaf of the growing tree in the batched form: its rows, sums, best split,
// and (after a speculative expansion) its two evaluated children.
struct Node {
  int depth = 0;
  Segment seg{0, 0, 0};
  LeafSplits ls;
  HostSplit best;
  int slot = -1;  // pool slot of its histogram
  int child[2] = {-1, -1};
  bool expanded = false;
  bool kids_searched = false;  // its children's splits were searched
};

struct Learner {
  bk_learner_params P{};
  int N = 0, npad = 0, G = 0, F = 0, TB = 0;
  int form = FORM_F64;
  bool quant = false, const_hess = false, stochastic = false, renew = false;
  bool batched = false;
  int bins = 4;
  std::vector<FeatDev> feats;
  std::vector<int> group_bin_start;
  std::vector<std::vector<double>> upper;
  std::vector<std::vector<int32_t>> bin2cat;
  // group tiles for LDS: small ones for the partition, large ones for segment histograms
  std::vector<int> tile_first;
  std::vector<int> group_bins;  // bins of every group
  int max_tile_bins = 0;
  int tile_width = 16, hist_nsub = 1;  // histogram layout: groups per tile, LDS copies per tile
  int tuned_width = 16, tuned_nsub = 1;
  int stride = 0, pay_off = 0;  // row record layout
  int rows_per_flush = 0;
  size_t cache = 0;       // histogram slots of the reference pool
  size_t num_slots = 0;   // device slots (batched form: one per evaluated leaf)
  int max_items = 0, max_partials = 0, max_part_tiles = 0;
  int part_tile = 0;                   // rows per partition tile
  int part_tile_small = 0;             // rows per short partition tile
  unsigned epoch = 0;  // partition launches so far: the parity picks the cursor bank

  // device memory
  hipStream_t stream = nullptr;
  DevState ds{};
  uint8_t* d_rec[3] = {nullptr, nullptr, nullptr};
  void* d_gh0 = nullptr;
  float *d_grad = nullptr, *d_hess = nullptr;
  double *d_gnoise = nullptr, *d_hnoise = nullptr;
  FeatDev* d_feats = nullptr;
  int* d_group_bin_start = nullptr;
  double* d_pool = nullptr;
  int* d_pool_cnt = nullptr;
  unsigned long long* d_part_q = nullptr;
  double* d_part_g = nullptr;
  double* d_part_h = nullptr;
  unsigned* d_part_c = nullptr;
  unsigned* d_cursors = nullptr;  // partition cursors: 2 banks x max_items x (left, right)
  int* d_err = nullptr;
  int* d_tile_first = nullptr;
  // staging arenas: every round's items go up in one copy, results come back in one
  uint8_t* d_arena = nullptr;
  uint8_t* h_arena = nullptr;
  uint8_t* d_back = nullptr;  // device address of h_back (not a device allocation)
  uint8_t* h_back = nullptr;
  size_t arena_bytes = 0, arena_cur = 0, arena_sent = 0;
  size_t off_fold = 0;
  // host-side staging of a round (packed into the arena when uploaded)
  std::vector<PartItem> v_pitems;
  std::vector<HistItem> v_hitems;
  std::vector<int> v_hblock;
  std::vector<ScanItem> v_sitems;
  std::vector<ReduceItem> v_ritems;
  std::vector<int> v_used;
  int* d_used = nullptr;        // per item x feature
  int* d_group_used = nullptr;  // per group
  SplitOut* d_out = nullptr;
  FoldOut* d_fold = nullptr;
  uint8_t* d_flags = nullptr;
  HistItem* d_hitems = nullptr;
  int* d_hblock = nullptr;
  PartItem* d_pitems = nullptr;
  ScanItem* d_sitems = nullptr;
  ReduceItem* d_ritems = nullptr;
  unsigned* d_max4 = nullptr;
  long long* d_sums3 = nullptr;
  double* d_sum_partials = nullptr;
  Segment* d_segs = nullptr;
  double* d_renew = nullptr;  // device address of the renewal sums inside h_tail
  uint8_t* h_tail = nullptr;  // pinned: [error flag, 16 bytes][renewal sums, 2 per leaf]
  uint8_t* h_misc = nullptr;  // pinned landing area of the per-tree set-up's small reads
  unsigned* h_stamp = nullptr;  // pinned word the stamp kernel writes (sync_spin)
  unsigned stamp_seq = 0;
  int* d_order = nullptr;
  double* d_leaf_vals = nullptr;
  int32_t* d_leaf_of = nullptr;
  int32_t* d_replay = nullptr;      // leaf of every row under a replayed tree, allocated at first use
  int32_t* d_bag = nullptr;         // a host-given bag's rows, allocated at the first bagged tree
  uint32_t* d_bag_bits = nullptr;   // bit r: row r is in the current tree's bag
  bool last_bagged = false;         // the last tree was trained on a bag
  bool sums_bagged = false;         // the kept quantized root sums are over a bag
  NodeDev* d_nodes = nullptr;     // internal nodes of the last tree (row-to-leaf walk)
  NodeRule* d_node_rules = nullptr;
  double* d_renew_part_sums = nullptr;
  int* d_renew_part_meta = nullptr;
  int* d_renew_flags = nullptr;
  RenewPieces renew_pieces;  // the piecewise chain's per-piece sums (renewal leaves whose certificate fails)
  NodeRule* h_nodes = nullptr;
  std::vector<NodeRule> tree_nodes;  // rule and group of every committed split, by internal node
  std::vector<int> last_left, last_right;  // children of the last tree's internal nodes
  bool leaf_of_valid = false;
  int renew_nparts = 0;
  // pinned host staging
  FoldOut* h_fold = nullptr;
  uint8_t* h_flags = nullptr;
  HistItem* h_hitems = nullptr;
  int* h_hblock = nullptr;
  std::vector<Segment> hist_seg;
  std::vector<int> hist_first, hist_count;
  PartItem* h_pitems = nullptr;
  ScanItem* h_sitems = nullptr;
  ReduceItem* h_ritems = nullptr;
  int* h_used = nullptr;
  int* h_group_used = nullptr;
  Segment* h_segs = nullptr;
  int* h_order = nullptr;
  double* h_vals = nullptr;

  // per-tree state (serial.rs:183-245)
  PoolMap pool;
  std::vector<HostSplit> best;
  ArgMax amax;
  LeafSplits smaller, larger;
  size_t smaller_slot = 0;
  long larger_slot = -1, parent_slot = -1;
  std::vector<Segment> segs;  // segment of every leaf of the last tree
  Mt use_start_eng{0};
  int tree_slot = 0, tree_period = 1;  // position in a sequence of trees shared by several learners
  double grad_scale = 0.0, hess_scale = 0.0;
  long long last_int_g = 0, last_int_h = 0;
  double last_root_g = 0.0, last_root_h = 0.0;
  bk_train_stats stats{};
  std::vector<int> used_feat;
  int last_num_leaves = 1;
  double h0 = 1.0;
  Prof prof;
  // The staging arrays are copied asynchronously: the host rewrites them only
  // after the last copy out of them has completed.
  hipEvent_t staged_ev = nullptr;
  bool staged_pending = false;
  // batched form
  std::vector<Node> nodes;
  std::vector<int> leaf_node;
  std::vector<char> node_flags;  // splittable flags of node i at [i * F, (i + 1) * F)
  FeatureSampler cols;
  std::vector<std::vector<int>> branch;  // real features split on along each leaf's path
  // by-node sampling in the batched form: every searched node keeps its per-feature
  // candidates and searched bits; its best is folded on the host when the reference
  // would draw its sample (when its parent's split is committed)
  // device round planner: batched growth whose rounds queue without a host wait
  bool plan_on_device = false;         // bk_learner_set_round_planner
  PlanState* d_pstate = nullptr;
  PlanNode* d_pnodes = nullptr;
  uint8_t* d_pnflags = nullptr;        // splittable flags of every node
  FoldOut* d_pfold = nullptr;          // the round's folded splits (device, read by the planner)
  uint8_t* d_pfflags = nullptr;
  PartItem* d_ppart = nullptr;
  HistItem* d_phist = nullptr;
  ScanItem* d_pscan = nullptr;
  ReduceItem* d_pred = nullptr;
  int* d_pused = nullptr;
  int* d_bytree = nullptr;
  uint8_t* h_plan = nullptr;           // pinned: the state, then the nodes
  PlanState plan_init{};
  std::vector<int> bytree_int;
  std::vector<uint8_t> ones_f;
  int plan_max_nodes = 0, hist_grid_blocks = 0, last_plan_rounds = 0;
  // quantized histograms with 32-bit bins (host-planned batched rounds): the
  // histograms carry no row counts, the partition counts the children and
  // k_children derives the round's histogram, reduce and search items
  bool hist32_ok = false;              // this learner may use them
  int hist_width = 0;                  // bk_learner_set_hist_bin_width: 0 picks, 4 or 8 forces
  bool part_hist_ok = false;           // the partition can build the built child's histogram
  bool part_hist_on = false;           // bk_learner_set_partition_histograms (measured slower: off)
  int rows_per_flush64 = 0, rows_per_flush32 = 0;
  std::vector<ChildSpec> v_cspec;
  HistItem* d_chist = nullptr;
  ScanItem* d_cscan = nullptr;
  ReduceItem* d_cred = nullptr;
  RoundHdr* d_chdr = nullptr;
  int* h_kcount = nullptr;             // pinned: the children's counts, then a status word
  unsigned* d_tickets = nullptr;       // [0]: the partition's blocks (derivation), [1]: the fold's waves (stamp)
  bool derive_in_partition = true;     // the partition's last block derives the round's items
  // device-resident trees: one host wait per tree
  bool resident_on = false;            // bk_learner_set_round_planner(…, 1) turns it on
  DiscState* d_disc = nullptr;
  int* d_node_count = nullptr;
  int* h_start = nullptr;              // pinned: the noise rotation drawn for the tree
  int* h_bytree = nullptr;             // pinned: the tree's feature sample
  DiscState* h_disc = nullptr;         // pinned copy of the discretizer state at a tree's end
  // trees whose root sums and renewal need no host fallback: integer root sums
  // are exact in any order when every partial sum fits the f64 significand,
  // which holds by construction for power-of-two bins and N * (bins + 1) < 2^29
  // (the scales carry at most the 24 significand bits of an f32 maximum)
  bool resident_ok(bool bagged) const {
    if (!resident_on || !planner_ok() || bagged || !d_disc) return false;
    if (quant && ((bins & (bins - 1)) != 0 || (double)N * (bins + 1) >= (double)(1 << 29))) return false;
    return true;
  }
  bool planner_ok() const {
    return plan_on_device && batched && d_pstate && !cols.by_node_active() && P.num_leaves <= kPlanMaxLeaves;
  }
  SplitOut* h_out = nullptr;           // pinned copy of a round's per-feature results
  std::vector<SplitOut> node_cands;    // node i: [i * F, (i + 1) * F)
  std::vector<uint8_t> node_searched;  // node i: [i * F, (i + 1) * F)
  char* flags_of(int node) { return node_flags.data() + (size_t)node * F; }

  static constexpr size_t kMiscBytes = 16 + sizeof(double) * 2 * (2048 + 2);
  static constexpr double kHoldUs = 30.0;  // the hold ahead of a timed round
  unsigned long long hold_ticks = 3000;    // kHoldUs in ticks of the device's wall clock
  int payload_bytes() const { return form == FORM_QUANT ? 2 : 8; }
  template <typename Fn>
  void timed(int kind, double algo, double moved, Fn&& launch) {
    if (!((prof.mask >> kind) & 1u) || !prof.tree_on) {
      launch();
      return;
    }
    prof.arm(kind, algo, moved);
    try {
      launch();
    } catch (...) {
      prof.disarm();
      throw;
    }
    prof.disarm();
  }

  // host time spent blocked on the device during the current train call
  double wait_ms = 0.0;
  void sync() {
    const auto w0 = std::chrono::steady_clock::now();
    BK_HIPC(hipStreamSynchronize(stream));
    wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    staged_pending = false;  // every staged copy has completed
  }

  // The same wait for work whose results land in pinned memory through kernels
  // (no pageable copies before it): a one-lane kernel stamps a sequence number
  // into pinned memory after that work, and the host spins on it. A blocked wait
  // on the stream sleeps on an interrupt and wakes tens of microseconds after the
  // work ends; the spin sees the stamp within a PCIe write. After kSpinMs without
  // the stamp the blocked wait takes over, which also reports any fault.
  static constexpr double kSpinMs = 200.0;
  void sync_spin() {
    const unsigned seq = ++stamp_seq;
    timed(8, 0.0, 0.0, [&] { launch_stamp(dev_alias(h_stamp), seq, stream); });
    wait_stamp(seq);
  }
  // The wait of sync_spin for a stamp that the stream's last kernel writes itself
  // (a stamped fold).
  void wait_stamp(unsigned seq) {
    const auto w0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (unsigned i = 1;; ++i) {
      if (__atomic_load_n(h_stamp, __ATOMIC_ACQUIRE) == seq) {
        seen = true;
        break;
      }
      if ((i & 4095u) == 0u &&
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count() > kSpinMs)
        break;
      __builtin_ia32_pause();
    }
    if (!seen) BK_HIPC(hipStreamSynchronize(stream));
    wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    staged_pending = false;
  }

  ~Learner() { release(); }

  void release() {
    auto f = [](void* p) {
      if (p) (void)hipFree(p);
    };
    auto fh = [this](void* p) {
      if (p) {
        alias_of.erase(p);
        (void)hipHostFree(p);
      }
    };
    for (int b = 0; b < 3; ++b) {
      f(d_rec[b]);
    }
    for (void* p : {(void*)d_grad, (void*)d_hess, (void*)d_gnoise, (void*)d_hnoise, (void*)d_feats,
                    (void*)d_group_bin_start, (void*)d_pool, (void*)d_pool_cnt, (void*)d_part_q, (void*)d_part_g,
                    (void*)d_part_h, (void*)d_part_c, (void*)d_cursors, (void*)d_err, (void*)d_tile_first,
                    (void*)d_gh0, (void*)d_arena, (void*)d_group_used, (void*)d_out,
                    (void*)d_max4, (void*)d_sums3,
                    (void*)d_sum_partials, (void*)d_segs, (void*)d_order, (void*)d_leaf_vals,
                    (void*)d_leaf_of, (void*)d_nodes, (void*)d_node_rules, (void*)d_renew_part_sums,
                    (void*)d_renew_part_meta, (void*)d_renew_flags, (void*)renew_pieces.pieces, (void*)renew_pieces.binade,
                    (void*)renew_pieces.work, (void*)renew_pieces.parts, (void*)renew_pieces.work_count, (void*)d_bag,
                    (void*)d_bag_bits, (void*)d_replay,
                    (void*)d_pstate, (void*)d_pnodes, (void*)d_pnflags, (void*)d_pfold, (void*)d_pfflags,
                    (void*)d_ppart, (void*)d_phist, (void*)d_pscan, (void*)d_pred, (void*)d_pused, (void*)d_bytree,
                    (void*)d_disc, (void*)d_node_count, (void*)d_chist, (void*)d_cscan, (void*)d_cred,
                    (void*)d_chdr, (void*)d_tickets})
      f(p);
    for (void* p : {(void*)h_arena, (void*)h_back, (void*)h_tail, (void*)h_misc, (void*)h_stamp,
                    (void*)h_group_used, (void*)h_segs, (void*)h_order, (void*)h_vals, (void*)h_nodes,
                    (void*)h_out, (void*)h_plan, (void*)h_start, (void*)h_bytree, (void*)h_disc, (void*)h_kcount})
      fh(p);
    if (staged_ev) (void)hipEventDestroy(staged_ev);
    staged_ev = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  // Device address of each pinned host buffer this learner owns, looked up once at
  // allocation; the map is the learner's own, so a learner may move between threads.
  std::unordered_map<void*, void*> alias_of;
  template <typename T>
  void pinned(T** p, size_t n) {
    BK_HIPC(hipHostMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocDefault));
    void* d = nullptr;
    BK_HIPC(hipHostGetDevicePointer(&d, (void*)*p, 0));
    alias_of[(void*)*p] = d;
  }
  template <typename T>
  T* dev_alias(T* h) {
    auto it = alias_of.find((void*)h);
    if (it == alias_of.end()) throw std::logic_error("dev_alias of a buffer this learner did not pin");
    return (T*)it->second;
  }

  void create(const bk_dataset_view* v, const bk_learner_params* p, int is_const_hess) {
    P = *p;
    if (P.extra_trees) throw std::invalid_argument("extra_trees is not supported by the GPU learner");
    if (!(P.feature_fraction > 0.0 && P.feature_fraction <= 1.0) ||
        !(P.feature_fraction_bynode > 0.0 && P.feature_fraction_bynode <= 1.0))
      throw std::invalid_argument("feature_fraction and feature_fraction_bynode must be in (0, 1]");
    if (v->num_data <= 0) throw std::invalid_argument("empty dataset");
    if (v->num_data > (1LL << 30)) throw std::invalid_argument("more than 2^30 rows");
    if (P.num_leaves < 2) throw std::invalid_argument("num_leaves must be at least 2");
    N = (int)v->num_data;
    npad = (N + 255) / 256 * 256;
    G = v->num_groups;
    F = v->num_features;
    const_hess = is_const_hess != 0;
    quant = P.use_quantized_grad && P.num_grad_quant_bins >= 2 && P.num_grad_quant_bins <= 127;
    if (P.form == BK_FORM_QUANTIZED && !quant)
      throw std::invalid_argument("form=quantized needs use_quantized_grad and 2 <= num_grad_quant_bins <= 127");
    form = quant ? FORM_QUANT : (P.form == BK_FORM_F32_COMPENSATED ? FORM_F32C : FORM_F64);
    bins = P.num_grad_quant_bins;
    stochastic = quant && P.stochastic_rounding;
    renew = quant && P.quant_train_renew_leaf;
    // groups and features
    group_bin_start.assign((size_t)G + 1, 0);
    for (int g = 0; g < G; ++g) {
      int nb = v->group_num_total_bin[g];
      if (nb > kMaxBins || nb < 1) throw std::invalid_argument("group bins must be in 1..256 (one byte per row)");
      group_bin_start[(size_t)g + 1] = group_bin_start[(size_t)g] + nb;
    }
    TB = group_bin_start[(size_t)G];
    feats.assign((size_t)F, FeatDev{});
    upper.assign((size_t)F, {});
    bin2cat.assign((size_t)F, {});
    for (int g = 0; g < G; ++g) {
      const int fs = v->group_feature_start[g], fc = v->group_feature_cnt[g];
      const int32_t* bo = v->group_bin_offsets + fs + g;
      for (int s = 0; s < fc; ++s) {
        int f = fs + s;
        FeatDev& m = feats[(size_t)f];
        m.num_bin = v->feature_num_bin[f];
        m.missing_type = v->feature_missing_type[f];
        m.default_bin = v->feature_default_bin[f];
        m.mfb = v->feature_most_freq_bin[f];
        m.offset = m.mfb == 0 ? 1 : 0;
        m.bin_type = v->feature_bin_type[f];
        m.group = g;
        m.bin_base = bo[s];
        m.hist_off = group_bin_start[(size_t)g] + bo[s];
        m.hist_len = m.num_bin - m.offset;
        if (fc == 1) {
          m.minb = 1;
          m.maxb = bo[1] - 1;
          m.use_min_bin = 0;
        } else {
          m.minb = bo[s];
          m.maxb = bo[s + 1] - 1;
          m.use_min_bin = 1;
        }
        m.real = v->feature_real_index[f];
        m.penalty = v->feature_penalty ? v->feature_penalty[f] : 1.0;
        if (m.num_bin > kMaxBins) throw std::invalid_argument("feature with more than 256 bins");
        if (m.bin_type == BK_BIN_CATEGORICAL) {
          if (!v->feature_bin_2_categorical || !v->feature_bin_2_categorical[f])
            throw std::invalid_argument("categorical feature without bin_2_categorical");
          bin2cat[(size_t)f].assign(v->feature_bin_2_categorical[f], v->feature_bin_2_categorical[f] + m.num_bin);
        } else {
          if (!v->feature_bin_upper_bound || !v->feature_bin_upper_bound[f])
            throw std::invalid_argument("numerical feature without bin_upper_bound");
          upper[(size_t)f].assign(v->feature_bin_upper_bound[f], v->feature_bin_upper_bound[f] + m.num_bin);
        }
      }
    }
    // row records: G bins, the payload at pay_off (aligned to its size), padded to 16 bytes
    if (G > kMaxGroups) throw std::invalid_argument("more than 1024 feature groups");
    const int pay_bytes = form == FORM_QUANT ? 2 : 8;
    pay_off = (G + pay_bytes - 1) / pay_bytes * pay_bytes;
    stride = (pay_off + pay_bytes + 15) / 16 * 16;
    group_bins.assign(v->group_num_total_bin, v->group_num_total_bin + G);
    build_tiles(16, 1);
    // rows one LDS accumulation may hold: the packed quantized fields stay exact
    if (form == FORM_QUANT) {
      int maxq = bins + 1;
      rows_per_flush = std::max(kTileRows, ((1 << 20) - 1) / maxq / kTileRows * kTileRows);
      // 32-bit bins: n rows stay exact while n * max h < 2^16 and n * max |g| < 2^15;
      // a discretized row has h <= bins + 1 and |g| <= bins / 2 + 1 (the rounding
      // of the scaled maximum may add one), so n * (bins + 2) < 2^16 bounds both
      rows_per_flush32 = ((1 << 16) - 1) / (bins + 2) / kTileRows * kTileRows;
    } else {
      rows_per_flush = 1 << 30;
    }
    rows_per_flush64 = rows_per_flush;
    cache = (size_t)P.num_leaves;
    if (P.histogram_pool_size > 0.0) {
      size_t bytes = 0;
      for (auto& m : feats) bytes += (size_t)m.num_bin * 16;
      size_t c = (size_t)(P.histogram_pool_size * 1024.0 * 1024.0 / (double)bytes);
      cache = std::min(std::max(c, (size_t)2), (size_t)P.num_leaves);
    }
    // the batched form keeps every evaluated leaf's histogram, so it needs an unbounded pool
    batched = P.batched != 0 && cache == (size_t)P.num_leaves;
    num_slots = batched ? (size_t)P.num_leaves * 2 + 2 : cache;
    max_items = P.num_leaves + 1;
    hist32_ok = form == FORM_QUANT && batched && rows_per_flush32 >= 2 * kTileRows && max_items <= kChildThreads;
    if (hist32_ok) rows_per_flush = std::min(rows_per_flush, rows_per_flush32);  // sizes the partials for both
    part_tile = partition_tile_rows(stride);
    part_tile_small = partition_tile_rows_small(stride);
    max_part_tiles = N / part_tile + max_items + 8;
    // a build has at most one row block per 2048 rows (the tuner's smallest row
    // count), and at least one per rows_per_flush rows: a partial slot never holds
    // more rows than one LDS accumulation (see plan_hist)
    max_partials = std::max({std::min(kHistMaxBlocks, N / 2048 + 1), N / rows_per_flush + 2,
                             (int)std::min<long long>(kHistSpreadBlocks, N / kHistSpreadRows + 1)}) +
                   2 * max_items + 8;
    // a histogram-building partition writes one partial per tile
    if (hist32_ok) max_partials = std::max(max_partials, max_part_tiles);
    // device allocation
    BK_HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // marks the end of staging transfers, which only read pinned host memory
    BK_HIPC(hipEventCreateWithFlags(&staged_ev, hipEventDisableTiming | hipEventDisableSystemFence));
    for (int b = 0; b < 3; ++b) {
      d_rec[b] = dmalloc<uint8_t>((size_t)npad * stride);
    }
    d_gh0 = form == FORM_QUANT ? (void*)dmalloc<int16_t>((size_t)npad) : (void*)dmalloc<float>((size_t)npad * 2);
    BK_HIPC(hipMemsetAsync(d_gh0, 0, (size_t)npad * pay_bytes, stream));
    // the bins arrive column-major; they are transposed into records on the device
    uint8_t* d_colstage = dmalloc<uint8_t>((size_t)G * npad);
    std::unique_ptr<void, void (*)(void*)> colstage_guard(d_colstage, [](void* q) { (void)hipFree(q); });
    for (int g = 0; g < G; ++g)
      BK_HIPC(hipMemcpyAsync(d_colstage + (size_t)g * npad, v->group_bins[g], (size_t)N, hipMemcpyHostToDevice, stream));
    d_grad = dmalloc<float>((size_t)npad);
    d_hess = dmalloc<float>((size_t)npad);
    d_feats = dmalloc<FeatDev>((size_t)F);
    BK_HIPC(hipMemcpyAsync(d_feats, feats.data(), sizeof(FeatDev) * (size_t)F, hipMemcpyHostToDevice, stream));
    d_group_bin_start = dmalloc<int>((size_t)G + 1);
    BK_HIPC(hipMemcpyAsync(d_group_bin_start, group_bin_start.data(), sizeof(int) * (size_t)(G + 1),
                           hipMemcpyHostToDevice, stream));
    d_pool = dmalloc<double>(num_slots * (size_t)TB * 2);
    d_pool_cnt = dmalloc<int>(num_slots * (size_t)TB);
    if (form == FORM_QUANT) {
      d_part_q = dmalloc<unsigned long long>((size_t)max_partials * TB);
    } else {
      d_part_g = dmalloc<double>((size_t)max_partials * TB);
      d_part_h = dmalloc<double>((size_t)max_partials * TB);
      d_part_c = dmalloc<unsigned>((size_t)max_partials * TB);
    }
    d_cursors = dmalloc<unsigned>((size_t)4 * max_items);
    BK_HIPC(hipMemsetAsync(d_cursors, 0, sizeof(unsigned) * (size_t)4 * max_items, stream));
    d_err = dmalloc<int>(1);
    BK_HIPC(hipMemsetAsync(d_err, 0, sizeof(int), stream));
    d_tile_first = dmalloc<int>((size_t)G + 1);  // the most tiles: one per group
    BK_HIPC(hipMemcpyAsync(d_tile_first, tile_first.data(), sizeof(int) * tile_first.size(), hipMemcpyHostToDevice,
                           stream));
    d_group_used = dmalloc<int>((size_t)G);
    d_out = dmalloc<SplitOut>((size_t)F * 2 * max_items);
    {
      // a round's items are packed into one pinned arena and sent in at most two
      // transfers; results come back in one
      v_pitems.resize((size_t)max_items);
      v_hitems.resize((size_t)max_items);
      v_hblock.resize((size_t)kHistMaxBlocks);
      v_sitems.resize((size_t)max_items);
      v_ritems.resize((size_t)2 * max_items);
      v_used.resize((size_t)F * max_items);
      h_pitems = v_pitems.data();
      h_hitems = v_hitems.data();
      h_hblock = v_hblock.data();
      h_sitems = v_sitems.data();
      h_ritems = v_ritems.data();
      h_used = v_used.data();
      auto up = [](size_t x) { return (x + 255) / 256 * 256; };
      // a round's items, or the per-tree renewal upload (segments, order, nodes)
      const size_t round_bytes = up(sizeof(PartItem) * (size_t)max_items) + up(sizeof(HistItem) * (size_t)max_items) +
                                 up(sizeof(int) * (size_t)kHistMaxBlocks) + up(sizeof(ScanItem) * (size_t)max_items) +
                                 up(sizeof(ReduceItem) * 2 * (size_t)max_items) +
                                 up(sizeof(int) * (size_t)F * max_items) + up(sizeof(ChildSpec) * (size_t)max_items);
      const size_t tree_bytes = up(sizeof(Segment) * (size_t)P.num_leaves) + up(sizeof(int) * (size_t)P.num_leaves) +
                                up(sizeof(NodeRule) * (size_t)P.num_leaves);
      const size_t plan_bytes = up(sizeof(PlanState)) + up(sizeof(PlanNode)) + 2 * up(sizeof(int) * (size_t)F);
      arena_bytes = std::max(round_bytes + plan_bytes, tree_bytes) + 256;
      d_arena = dmalloc<uint8_t>(arena_bytes);
      pinned(&h_arena, arena_bytes);
      off_fold = up((size_t)2 * max_items * F);
      const size_t back = off_fold + sizeof(FoldOut) * 2 * (size_t)max_items;
      // the fold writes its results straight into pinned host memory (no copy back)
      pinned(&h_back, back);
      BK_HIPC(hipHostGetDevicePointer((void**)&d_back, h_back, 0));
      d_flags = d_back;
      h_flags = h_back;
      d_fold = (FoldOut*)(d_back + off_fold);
      h_fold = (FoldOut*)(h_back + off_fold);
    }
    if (hist32_ok) {
      v_cspec.resize((size_t)max_items);
      d_chist = dmalloc<HistItem>((size_t)max_items);
      d_cscan = dmalloc<ScanItem>((size_t)max_items);
      d_cred = dmalloc<ReduceItem>((size_t)max_items);
      d_chdr = dmalloc<RoundHdr>(1);
      pinned(&h_kcount, (size_t)2 * max_items + 1);
    }
    d_tickets = dmalloc<unsigned>(2);
    BK_HIPC(hipMemsetAsync(d_tickets, 0, sizeof(unsigned) * 2, stream));
    {
      int dev = 0, khz = 0;
      BK_HIPC(hipGetDevice(&dev));
      BK_HIPC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      hold_ticks = (unsigned long long)((khz > 0 ? khz : 100000) * (kHoldUs / 1000.0));
    }
    d_max4 = dmalloc<unsigned>(4);
    BK_HIPC(hipMemsetAsync(d_max4, 0, sizeof(unsigned) * 4, stream));
    d_sums3 = dmalloc<long long>(4);
    d_sum_partials = dmalloc<double>(2048 + 2);
    d_segs = dmalloc<Segment>((size_t)P.num_leaves);
    // the renewal kernels write the leaf sums straight into pinned host memory
    pinned(&h_tail, 16 + sizeof(double) * 2 * (size_t)P.num_leaves);
    pinned(&h_misc, kMiscBytes);
    pinned(&h_stamp, 16);
    *h_stamp = 0u;
    BK_HIPC(hipHostGetDevicePointer((void**)&d_renew, h_tail + 16, 0));
    d_order = dmalloc<int>((size_t)P.num_leaves);
    d_leaf_vals = dmalloc<double>((size_t)P.num_leaves);
    pinned(&h_group_used, (size_t)G);
    pinned(&h_segs, (size_t)P.num_leaves);
    pinned(&h_order, (size_t)P.num_leaves);
    pinned(&h_vals, (size_t)P.num_leaves);
    d_leaf_of = dmalloc<int32_t>((size_t)npad);
    BK_HIPC(hipMemsetAsync(d_leaf_of, 0, sizeof(int32_t) * (size_t)npad, stream));
    d_nodes = dmalloc<NodeDev>((size_t)P.num_leaves);
    d_node_rules = dmalloc<NodeRule>((size_t)P.num_leaves);
    pinned(&h_nodes, (size_t)P.num_leaves);
    tree_nodes.assign((size_t)P.num_leaves, NodeRule{});
    if (renew) {
      d_renew_flags = dmalloc<int>((size_t)P.num_leaves);
      if (renew_pieces_lds(P.num_leaves) <= (size_t)kLdsBytes) {
        renew_pieces.piece_rows = renew_piece_rows(N, P.num_leaves, (size_t)256 << 20);
        renew_pieces.num_pieces = (int)(((long long)N + renew_pieces.piece_rows - 1) / renew_pieces.piece_rows);
        const size_t entries = (size_t)std::max(renew_pieces.num_pieces, 1) * (size_t)P.num_leaves;
        renew_pieces.pieces = dmalloc<RenewPiece>(entries);
        // binade summaries per piece, leaf and component; the list holds up to 65536 of
        // the pieces the guess picks, with their sub-range summaries (pieces beyond it
        // are added value by value)
        renew_pieces.binade = dmalloc<RenewBinade>(entries * 2);
        renew_pieces.work_cap = (int)std::min<size_t>(entries * 2, (size_t)65536);
        renew_pieces.work = dmalloc<RenewWork>((size_t)renew_pieces.work_cap);
        renew_pieces.parts = dmalloc<RenewPart>((size_t)renew_pieces.work_cap * kRenewParts);
        renew_pieces.work_count = dmalloc<int>(1);
      }
      if (renew_par_lds(P.num_leaves) <= (size_t)kLdsBytes) {
        renew_nparts = renew_parts(N);
        d_renew_part_sums = dmalloc<double>((size_t)renew_nparts * P.num_leaves * 2);
        d_renew_part_meta = dmalloc<int>((size_t)renew_nparts * P.num_leaves * 4);
      }
    }
    if (stochastic) {
      // init_stochastic (gradient_discretizer/mod.rs:87-106): per-block MT19937 noise
      std::vector<double> gn((size_t)N), hn((size_t)N);
      size_t threads = P.num_threads > 0 ? (size_t)P.num_threads : 1;
      size_t nb = std::min(threads, ((size_t)N + 511) / 512);
      size_t bs = (size_t)N;
      if (nb > 1) {
        size_t s = ((size_t)N + nb - 1) / nb;
        bs = (s + 31) / 32 * 32;
      }
      for (size_t b = 0; b < nb; ++b) {
        size_t start = b * bs, end = std::min(start + bs, (size_t)N);
        Mt ge((uint32_t)P.seed + (uint32_t)b), he((uint32_t)P.seed + (uint32_t)b + (uint32_t)threads);
        for (size_t i = start; i < end; ++i) {
          gn[i] = ge.real01();
          hn[i] = he.real01();
        }
      }
      d_gnoise = dmalloc<double>((size_t)N);
      d_hnoise = dmalloc<double>((size_t)N);
      BK_HIPC(hipMemcpyAsync(d_gnoise, gn.data(), sizeof(double) * (size_t)N, hipMemcpyHostToDevice, stream));
      BK_HIPC(hipMemcpyAsync(d_hnoise, hn.data(), sizeof(double) * (size_t)N, hipMemcpyHostToDevice, stream));
      sync();
      use_start_eng.seed_with((uint32_t)P.seed);
    }
    sync();
    // kernel view
    ds.num_data = N;
    ds.npad = npad;
    ds.num_groups = G;
    ds.num_features = F;
    ds.total_bins = TB;
    ds.form = form;
    ds.stride = stride;
    ds.pay_off = pay_off;
    for (int b = 0; b < 3; ++b) {
      ds.rec[b] = d_rec[b];
    }
    ds.gh0 = d_gh0;
    ds.feats = d_feats;
    ds.group_bin_start = d_group_bin_start;
    ds.pool = d_pool;
    ds.pool_cnt = d_pool_cnt;
    ds.part_q = d_part_q;
    ds.part_g = d_part_g;
    ds.part_h = d_part_h;
    ds.part_c = d_part_c;
    ds.err_flag = d_err;
    launch_to_records(ds, d_colstage, (size_t)npad, stream);
    sync();
    pool.init((size_t)P.num_leaves, cache, (size_t)F);
    best.assign((size_t)P.num_leaves, HostSplit());
    segs.assign((size_t)P.num_leaves, Segment{0, 0, 0});
    used_feat.assign((size_t)F, 1);
    {
      std::vector<int> order((size_t)F);
      for (int f = 0; f < F; ++f) order[(size_t)f] = f;
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return v->feature_real_index[a] < v->feature_real_index[b]; });
      cols.init(std::move(order), P.feature_fraction, P.feature_fraction_bynode, P.feature_fraction_seed);
      cols.real_of.assign(v->feature_real_index, v->feature_real_index + F);
      if (P.num_interaction_constraints < 0) throw std::invalid_argument("negative interaction constraint count");
      if (P.num_interaction_constraints > 0 &&
          (!P.interaction_constraint_offsets || !P.interaction_constraint_features))
        throw std::invalid_argument("interaction constraints without their arrays");
      for (int i = 0; i < P.num_interaction_constraints; ++i) {
        const int a = P.interaction_constraint_offsets[i], b = P.interaction_constraint_offsets[i + 1];
        if (a < 0 || b < a) throw std::invalid_argument("interaction constraint offsets must ascend from 0");
        std::vector<int> c(P.interaction_constraint_features + a, P.interaction_constraint_features + b);
        std::sort(c.begin(), c.end());
        cols.constraints.push_back(std::move(c));
      }
      // the arrays are the caller's: they are not kept
      P.interaction_constraint_offsets = nullptr;
      P.interaction_constraint_features = nullptr;
      if (cols.by_node_active()) pinned(&h_out, (size_t)F * 2 * max_items);
    }
    if (batched && P.num_leaves <= kPlanMaxLeaves && !cols.by_node_active()) {
      plan_max_nodes = P.num_leaves * 4 + 4;
      d_pstate = dmalloc<PlanState>(1);
      d_pnodes = dmalloc<PlanNode>((size_t)plan_max_nodes);
      d_pnflags = dmalloc<uint8_t>((size_t)plan_max_nodes * F);
      d_pfold = dmalloc<FoldOut>((size_t)2 * max_items);
      d_pfflags = dmalloc<uint8_t>((size_t)2 * max_items * F);
      d_ppart = dmalloc<PartItem>((size_t)max_items);
      d_phist = dmalloc<HistItem>((size_t)max_items);
      d_pscan = dmalloc<ScanItem>((size_t)max_items);
      d_pred = dmalloc<ReduceItem>((size_t)max_items);
      d_pused = dmalloc<int>((size_t)F * max_items);
      d_bytree = dmalloc<int>((size_t)F);
      pinned(&h_plan, sizeof(PlanState) + sizeof(PlanNode) * (size_t)plan_max_nodes);
      d_disc = dmalloc<DiscState>(1);
      BK_HIPC(hipMemsetAsync(d_disc, 0, sizeof(DiscState), stream));
      d_node_count = dmalloc<int>(1);
      pinned(&h_start, 1);
      pinned(&h_bytree, (size_t)F);
      pinned(&h_disc, 1);
      bytree_int.assign((size_t)F, 1);
      ones_f.assign((size_t)F, 1);
      // the most row blocks a planned round's histogram may cut (its rows are at most N)
      long long nb = std::min<long long>(kHistMaxBlocks, std::max<long long>(1, N / 2048));
      nb = std::max<long long>(nb, ((long long)N + rows_per_flush - 1) / rows_per_flush);
      hist_grid_blocks = (int)nb;
    }
    for (int g = 0; g < G; ++g) h_group_used[g] = 1;
    BK_HIPC(hipMemcpyAsync(d_group_used, h_group_used, sizeof(int) * (size_t)G, hipMemcpyHostToDevice, stream));
    sync();
    if (hist32_ok) {
      DevState s32 = ds;
      s32.hist32 = 1;
      part_hist_ok = partition_hist_fits(s32) && part_tile <= rows_per_flush32 && max_partials >= max_part_tiles;
    }
    use_hist32(hist32_wanted());  // the tuner times the bins the trees will use
    autotune_hist();
  }

  // LDS histogram tiles of whole groups: a new tile every `width` groups (a
  // multiple of the 16 groups of a record chunk; a histogram block reads the
  // tile's chunks of every row), or when nsub copies of the tile would exceed the
  // LDS histogram budget. Narrow tiles mean more blocks and more resident waves
  // per CU for the same rows; a tile of the whole record reads every row once.
  void build_tiles(int width, int nsub) {
    const int per_bin = form == FORM_QUANT ? 8 : 20;  // the widest bin a tree may use
    const int tile_cap = kHistLdsBytes / (per_bin * nsub);
    tile_first.assign(1, 0);
    max_tile_bins = 0;
    int cur = 0;
    for (int g = 0; g < G; ++g) {
      const int nb = group_bins[(size_t)g];
      if (nb > tile_cap) throw std::invalid_argument("a feature group has more bins than the LDS histogram holds");
      if (cur > 0 && (cur + nb > tile_cap || g % width == 0)) {
        tile_first.push_back(g);
        max_tile_bins = std::max(max_tile_bins, cur);
        cur = 0;
      }
      cur += nb;
    }
    tile_first.push_back(G);
    max_tile_bins = std::max(max_tile_bins, cur);
    tile_width = width;
    hist_nsub = nsub;
    if (d_tile_first) {
      BK_HIPC(hipMemcpyAsync(d_tile_first, tile_first.data(), sizeof(int) * tile_first.size(), hipMemcpyHostToDevice,
                             stream));
      sync();
    }
  }
  void hist_launch(const DevState& s, const HistItem* items, int ni, int total, int nblocks, int rpb,
                   const int* block_slot, const RoundHdr* hdr) {
    launch_hist(s, items, ni, total, nblocks, rpb, block_slot, d_group_used, max_tile_bins, d_tile_first,
                (int)tile_first.size() - 1, rows_per_flush, stream, hdr, hist_nsub);
  }

  // 32-bit bins for the next tree: quantized, batched with host-planned rounds,
  // and not turned off by bk_learner_set_hist_bin_width
  bool hist32_wanted() const { return hist32_ok && hist_width != 8 && !planner_ok(); }
  void use_hist32(bool on) {
    ds.hist32 = on ? 1 : 0;
    rows_per_flush = on ? rows_per_flush32 : rows_per_flush64;
  }

  // Autotuner of the histogram tiling, three knobs: the layout (groups per
  // tile: 16, 32 or the whole record; 1 or 2 LDS copies per tile), then the row
  // tiling (rows per workgroup at least hist_min_rows, at most hist_max_blocks
  // row blocks) with the chosen layout. It times the histogram kernel on this
  // dataset's records for every candidate, over all rows (a root build) and over
  // a fifth of them (a child-sized build), and keeps the candidate with the least
  // time for one root and three child builds. The payload is whatever buffer 0
  // holds: the kernel's work does not depend on its values. The choice is kept
  // per (rows, groups, bins, form and bin width) for the process. Results do not
  // depend on it: the quantized histograms are exact integer sums.
  void autotune_hist() {
    static std::mutex mu;
    static std::map<std::array<int, 4>, std::array<int, 4>> picked;
    const std::array<int, 4> key{N, G, TB, form * 2 + ds.hist32};
    std::lock_guard<std::mutex> lock(mu);
    auto it = picked.find(key);
    if (it != picked.end()) {
      apply_tuned(it->second);
      return;
    }
    hipEvent_t e0, e1;
    BK_HIPC(hipEventCreate(&e0));
    BK_HIPC(hipEventCreate(&e1));
    // one root build and three child-sized builds (a fifth of the rows)
    auto time_now = [&]() {
      double t = 0.0;
      for (int part = 0; part < 2; ++part) {
        const int rows = part == 0 ? N : std::max(N / 5, 1);
        hist_seg.assign(1, Segment{0, 0, rows});
        const HistPlan hp = plan_hist(1);
        begin_uploads();
        d_hitems = put(h_hitems, (size_t)hp.ni);
        d_hblock = put(h_hblock, (size_t)hp.nblocks);
        send(true);
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms up, the second is timed
          BK_HIPC(hipEventRecord(e0, stream));
          hist_launch(ds, d_hitems, hp.ni, (int)hp.total, hp.nblocks, hp.rpb, d_hblock, nullptr);
          BK_HIPC(hipEventRecord(e1, stream));
          sync();
          BK_HIPC(hipEventElapsedTime(&ms, e0, e1));
        }
        t += (part == 0 ? 1.0 : 3.0) * ms;
      }
      return t;
    };
    // (1) the layout (groups per tile, LDS copies) at a middle row tiling
    hist_min_rows = 4096;
    hist_max_blocks = 1024;
    std::array<int, 4> best_c{hist_min_rows, hist_max_blocks, 16, 1};
    double best_t = 1e300;
    std::vector<std::vector<int>> seen;
    for (const auto& wn : {std::array<int, 2>{16, 1}, std::array<int, 2>{32, 1}, std::array<int, 2>{kMaxGroups, 1},
                           std::array<int, 2>{16, 2}, std::array<int, 2>{32, 2}}) {
      build_tiles(wn[0], wn[1]);
      std::vector<int> sig = tile_first;  // layouts that cut the same tiles are timed once
      sig.push_back(wn[1]);
      if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
      seen.push_back(sig);
      const double t = time_now();
      if (t < best_t) {
        best_t = t;
        best_c[2] = wn[0];
        best_c[3] = wn[1];
      }
    }
    build_tiles(best_c[2], best_c[3]);
    // (2) the row tiling (rows per workgroup at least, row blocks at most) with it
    best_t = 1e300;
    for (int min_rows : {2048, 4096, 8192})
      for (int max_blocks : {512, 1024, 2048}) {
        hist_min_rows = min_rows;
        hist_max_blocks = max_blocks;
        const double t = time_now();
        if (t < best_t) {
          best_t = t;
          best_c[0] = min_rows;
          best_c[1] = max_blocks;
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    apply_tuned(best_c);
    picked[key] = best_c;
  }
  void apply_tuned(const std::array<int, 4>& c) {
    hist_min_rows = tuned_min_rows = c[0];
    hist_max_blocks = tuned_max_blocks = c[1];
    tuned_width = c[2];
    tuned_nsub = c[3];
    build_tiles(c[2], c[3]);
  }

  ScanParams scan_params() const {
    ScanParams sp{};
    sp.l1 = P.lambda_l1;
    sp.l2 = P.lambda_l2;
    sp.max_delta_step = P.max_delta_step;
    sp.path_smooth = P.path_smooth;
    sp.min_sum_hessian = P.min_sum_hessian_in_leaf;
    sp.min_gain_to_split = P.min_gain_to_split;
    sp.cat_smooth = P.cat_smooth;
    sp.cat_l2 = P.cat_l2;
    sp.min_data_in_leaf = P.min_data_in_leaf;
    sp.max_cat_threshold = P.max_cat_threshold;
    sp.max_cat_to_onehot = P.max_cat_to_onehot;
    sp.min_data_per_group = P.min_data_per_group;
    sp.grad_scale = grad_scale;
    sp.hess_scale = hess_scale;
    sp.h0 = h0;
    sp.const_hess = const_hess ? 1 : 0;
    return sp;
  }

  // group_used on the device: every group (all) or those with a feature in the
  // tree's sample
  void set_groups_used(bool all) {
    stage_wait();
    for (int g = 0; g < G; ++g) h_group_used[g] = all ? 1 : 0;
    if (!all)
      for (int f = 0; f < F; ++f)
        if (cols.bytree[(size_t)f]) h_group_used[feats[(size_t)f].group] = 1;
    timed(8, 0.0, 0.0,
          [&] { launch_copy(d_group_used, dev_alias(h_group_used), sizeof(int) * (size_t)G, stream); });
    staged();
  }

  void ensure_bag_buffers() {
    if (d_bag_bits) return;
    d_bag = dmalloc<int32_t>((size_t)N + 1);
    d_bag_bits = dmalloc<uint32_t>((size_t)N / 32 + 1);
  }

  // The root segment of a bagged tree: the bag's records, ascending, at the start
  // of buffer 1 (DataPartition::init_with, data_partition.rs:100-107); the rows
  // are marked for the row-to-leaf map, and the quantized pairs summed into
  // d_sums3 (zeroed here).
  void gather_bag(const int32_t* d_rows, int nrows) {
    BK_HIPC(hipMemsetAsync(d_bag_bits, 0, sizeof(uint32_t) * ((size_t)N / 32 + 1), stream));
    BK_HIPC(hipMemsetAsync(d_sums3, 0, sizeof(long long) * 4, stream));
    const double db = (double)nrows * (4.0 + 2.0 * stride);
    timed(4, db, db, [&] { launch_gather_bag(ds, d_rows, nrows, 1, d_bag_bits, d_sums3, stream); });
  }

  // The discretizer (gradient_discretizer/mod.rs:134-201) into buffer 0's payload:
  // scales from every row, one rotation draw of the noise per tree, the packed
  // int8 pairs, and their exact sums in d_sums3. Returns false when the
  // reference refuses to discretize (non-finite or all-zero gradients), which
  // keeps the previous discretization (serial.rs:692-703).
  bool discretize(const float* g, const float* h) {
    BK_HIPC(hipMemsetAsync(d_max4, 0, sizeof(unsigned) * 4, stream));
    timed(4, (double)N * 8.0, (double)N * 8.0, [&] { launch_max_abs(g, h, N, d_max4, d_sums3, stream); });
    unsigned mx[4];
    timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_misc), d_max4, sizeof mx, stream); });
    // the maxima start from zero for the next tree on every path
    BK_HIPC(hipMemsetAsync(d_max4, 0, sizeof(unsigned) * 4, stream));
    sync_spin();
    std::memcpy(mx, h_misc, sizeof mx);
    float max_g, max_h;
    std::memcpy(&max_g, &mx[0], 4);
    std::memcpy(&max_h, &mx[1], 4);
    const bool ok = mx[2] == 0u && max_g != 0.0f && (const_hess || max_h != 0.0f);
    if (!ok) return false;
    grad_scale = (double)max_g / (double)(bins / 2);
    hess_scale = const_hess ? (double)max_h : (double)max_h / (double)bins;
    double inv_g = 1.0 / grad_scale, inv_h = 1.0 / hess_scale;
    // the rotation of the noise: one draw per tree; a learner that trains tree
    // tree_slot of every tree_period trees of a shared sequence draws for all of
    // them and keeps its own (serial.rs:692-703 draws only when discretizing)
    int start = 0;
    if (stochastic)
      for (int j = 0; j < tree_period; ++j) {
        const int v = (int)use_start_eng.uint_incl((uint32_t)N);
        if (j == tree_slot) start = v;
      }
    const double db = (double)N * (8.0 + (stochastic ? 16.0 : 0.0) + 2.0);
    timed(4, db, db, [&] {
      launch_discretize(g, h, d_gnoise, d_hnoise, N, start, inv_g, inv_h, const_hess ? 1 : 0, stochastic ? 1 : 0,
                        (int16_t*)d_gh0, d_sums3, stream);
    });
    return true;
  }

  // Discretization and root sums (serial.rs:658-776, gradient_discretizer/mod.rs:134-201).
  // d_rows: the tree's bag on the device (ascending rows), or null for every row.
  // The discretizer's scales always come from every row; the root sums run over
  // the bag.
  LeafSplits before_train(const float* g, const float* h, const int32_t* d_rows = nullptr, int nrows = 0) {
    pool.reset();
    for (auto& s : best) s.reset();
    amax.init(&best);
    std::fill(segs.begin(), segs.end(), Segment{0, 0, 0});
    branch.assign((size_t)P.num_leaves, std::vector<int>());
    const bool bag = d_rows != nullptr;
    const int count = bag ? nrows : N;
    segs[0] = bag ? Segment{1, 0, nrows} : Segment{0, 0, N};
    // the tree's feature sample (col_sampler.rs:87-101): histograms only for the
    // groups of its features
    if (cols.per_tree) {
      cols.next_tree(tree_slot, tree_period);
      set_groups_used(false);
    }
    stats = bk_train_stats{};
    stats.hist_bin_bytes = form == FORM_QUANT ? (ds.hist32 ? 4 : 8) : 0;
    double sg = 0.0, sh = 0.0;
    long long isum = 0;
    if (quant) {
      const bool ok = discretize(g, h);
      // the bag's sums come from the packed pairs in place, fresh or kept from the
      // previous tree when this discretization was refused; kept pairs of a bagged
      // tree are summed again over every row when this tree has no bag
      if (bag) gather_bag(d_rows, nrows);
      const bool resum_all = !ok && !bag && sums_bagged;
      sums_bagged = (ok || bag) ? bag : sums_bagged;
      if (ok || bag || resum_all) {
        long long sums[4] = {0, 0, 0, 0};
        if (!resum_all) {
          timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_misc), d_sums3, sizeof sums, stream); });
          sync_spin();
          std::memcpy(sums, h_misc, sizeof sums);
        }
        if (resum_all) sums_bagged = false;
        // exact when every sequential partial sum fits the f64 significand
        auto exact = [](long long sum_abs, double scale) {
          if (scale == 0.0) return true;
          int e;
          double m = std::frexp(scale, &e);
          uint64_t mant = (uint64_t)std::ldexp(m, 53);
          while (mant && !(mant & 1ull)) mant >>= 1;
          unsigned __int128 prod = (unsigned __int128)(uint64_t)sum_abs * mant;
          return prod < ((unsigned __int128)1 << 53);
        };
        if (!resum_all && exact(sums[2], grad_scale) && exact(sums[1], hess_scale)) {
          last_int_g = sums[0];
          last_int_h = sums[1];
          last_root_g = (double)sums[0] * grad_scale;
          last_root_h = (double)sums[1] * hess_scale;
        } else {
          // sequential f64 sums in row order over the bag or every row
          std::vector<int16_t> pk((size_t)N);
          std::vector<int32_t> rows((size_t)(bag ? nrows : 0));
          BK_HIPC(hipMemcpyAsync(pk.data(), d_gh0, sizeof(int16_t) * (size_t)N, hipMemcpyDeviceToHost, stream));
          if (bag && nrows > 0)
            BK_HIPC(hipMemcpyAsync(rows.data(), d_rows, sizeof(int32_t) * (size_t)nrows, hipMemcpyDeviceToHost,
                                   stream));
          sync();
          double a = 0.0, b = 0.0;
          long long ig = 0, ih = 0;
          for (int j = 0; j < count; ++j) {
            const int i = bag ? rows[(size_t)j] : j;
            if (i < 0 || i >= N) continue;  // flagged by the gather, reported at the end of the tree
            const int8_t gq = (int8_t)(pk[(size_t)i] >> 8);
            const uint8_t hq = (uint8_t)(pk[(size_t)i] & 0xff);
            a += (double)gq * grad_scale;
            b += (double)hq * hess_scale;
            ig += gq;
            ih += hq;
          }
          last_int_g = ig;
          last_int_h = ih;
          last_root_g = a;
          last_root_h = b;
        }
      }
      // on refusal the previous discretization and its sums are kept (serial.rs:692-703)
      sg = last_root_g;
      sh = last_root_h;
      if (d_disc) {
        // a later device-resident tree keeps this state when its discretization is refused
        stage_wait();
        DiscState dsn{};
        dsn.gs = grad_scale;
        dsn.hs = hess_scale;
        dsn.inv_g = grad_scale != 0.0 ? 1.0 / grad_scale : 0.0;
        dsn.inv_h = hess_scale != 0.0 ? 1.0 / hess_scale : 0.0;
        dsn.root_g = last_root_g;
        dsn.root_h = last_root_h;
        dsn.int_g = last_int_g;
        dsn.int_h = last_int_h;
        *h_disc = dsn;
        timed(8, 0.0, 0.0, [&] { launch_copy(d_disc, dev_alias(h_disc), sizeof(DiscState), stream); });
        staged();
      }
      isum = (long long)(((unsigned long long)last_int_g << 32) + (unsigned long long)last_int_h);
    } else {
      launch_pack_f32(g, h, N, const_hess ? 1 : 0, (float*)d_gh0, stream);
      if (bag) gather_bag(d_rows, nrows);
      int nb = launch_sum_gh(g, h, d_rows, count, N, d_sum_partials, stream);
      if ((size_t)nb * 2 * sizeof(double) + 16 > kMiscBytes) throw std::runtime_error("set-up landing area too small");
      timed(8, 0.0, 0.0, [&] {
        launch_copy(dev_alias(h_misc) + 16, d_sum_partials, sizeof(double) * 2 * (size_t)nb, stream);
      });
      timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_misc), h, sizeof(float), stream); });
      sync_spin();
      float hh0 = 0.0f;
      std::memcpy(&hh0, h_misc, sizeof(float));
      std::vector<double> part((size_t)nb * 2);
      std::memcpy(part.data(), h_misc + 16, sizeof(double) * part.size());
      for (int b = 0; b < nb; ++b) {
        sg += part[(size_t)2 * b];
        sh += part[(size_t)2 * b + 1];
      }
      h0 = (double)hh0;
    }
    double carried = smaller.weight;
    LeafSplits root{0, count, sg, sh, carried, isum};
    smaller = root;
    larger = LeafSplits{};
    return root;
  }

  double parent_output(const Tree& t, const LeafSplits& ls) const {
    if (t.num_leaves == 1) return leaf_output_root(P, ls.sum_g, ls.sum_h);
    return ls.weight;
  }

  LeafArgs leaf_args(const Tree& t, const LeafSplits& ls, int slot) const {
    return leaf_args_with(ls, slot, parent_output(t, ls));
  }

  static LeafArgs leaf_args_with(const LeafSplits& ls, int slot, double parent_out) {
    LeafArgs a{};
    a.leaf = ls.leaf;
    a.slot = slot;
    a.num_data = ls.num_data;
    a.sum_g = ls.sum_g;
    a.sum_h = ls.sum_h;
    a.parent_output = parent_out;
    a.int_sum = ls.int_sum;
    return a;
  }

  // ---- launch helpers over staged item arrays -------------------------------
  void staged() {
    BK_HIPC(hipEventRecord(staged_ev, stream));
    staged_pending = true;
  }

  void stage_wait() {
    if (!staged_pending) return;
    const auto w0 = std::chrono::steady_clock::now();
    BK_HIPC(hipEventSynchronize(staged_ev));
    wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    staged_pending = false;
  }

  // Reserves the row blocks of one item of a launch: returns its first block.
  static int add_tasks(int count, int chunk, int& ntasks) {
    const int first = ntasks;
    ntasks += count > 0 ? (count + chunk - 1) / chunk : 0;
    return first;
  }

  // Algorithmic bytes per row (the reference's index-based model): a root build
  // reads F bins and the payload P; a leaf build also reads a 4-byte row index;
  // a partition reads one bin and reads and writes a 4-byte row index. Moved
  // bytes are what these kernels touch: whole row records.
  //
  // Histogram of the staged items (blocks 0 .. ntasks-1); block j writes
  // ---- staging: a round's items are written to the host arena, marked, and
  // copied to the device in one transfer before the round's launches ----------
  struct HistPlan {
    int ni = 0, nblocks = 0, rpb = 0;
    long long total = 0;
  };

  // starts a round's uploads: the previous round's transfers have completed.
  // In a timed tree a hold is queued first: the round's upload and its first
  // timed launch then wait in the queue behind it, so the launch's start event
  // is not stamped before the host has submitted the kernel (the GPU is idle at
  // a round's start, and submitting a timed launch takes the host ~10 us).
  void begin_uploads() {
    stage_wait();
    arena_cur = arena_sent = 0;
    if (prof.mask != 0 && prof.tree_on) launch_hold(hold_ticks, stream);
  }

  // packs n elements into the arena; returns where they will be on the device
  template <typename T>
  T* put(const T* src, size_t n) {
    const size_t off = (arena_cur + 255) / 256 * 256;
    if (off + sizeof(T) * n > arena_bytes) throw std::runtime_error("staging arena too small");
    std::memcpy(h_arena + off, src, sizeof(T) * n);
    arena_cur = off + sizeof(T) * n;
    return (T*)(d_arena + off);
  }

  // sends what was packed since the last send; a caller that synchronizes the
  // stream before the next upload needs no event after the copy (an event is a
  // marker packet that stalls the queue for several microseconds)
  void send(bool sync_follows = false) {
    if (arena_cur > arena_sent) {
      const size_t from = arena_sent & ~(size_t)15;  // 16-byte aligned: the bytes before are sent already
      timed(8, 0.0, 0.0, [&] { launch_copy(d_arena + from, dev_alias(h_arena) + from, arena_cur - from, stream); });
      if (!sync_follows) staged();
    }
    arena_sent = arena_cur;
  }

  // packs a round's partition items (npart) and search items (nscan, with the
  // histogram plan and nr reduce items) and sends them in one transfer
  void upload_round(int npart, const HistPlan* hp, int nscan, int nr, bool sync_follows = false) {
    begin_uploads();
    if (npart > 0) d_pitems = put(h_pitems, (size_t)npart);
    if (nscan > 0) {
      d_hitems = put(h_hitems, (size_t)hp->ni);
      d_hblock = put(h_hblock, (size_t)hp->nblocks);
      d_sitems = put(h_sitems, (size_t)nscan);
      d_ritems = put(h_ritems, (size_t)nr);
      d_used = put(h_used, (size_t)F * nscan);
    }
    send(sync_follows);
  }

  // Histograms of the segments hist_seg[0..n): their rows are concatenated and
  // cut into equal row ranges, one per block (at least hist_min_rows rows, at most
  // hist_max_blocks blocks); a block writes one partial per segment it touches.
  // Segment i's partials are hist_first[i] .. hist_first[i] + hist_count[i] - 1.
  static constexpr int kHistMaxBlocks = 2048;  // largest row-block count the tuner may pick
  int hist_max_blocks = 1024, hist_min_rows = 4096;  // picked by autotune_hist
  int tuned_max_blocks = 1024, tuned_min_rows = 4096;  // the tuner's choice (kept under an override)
  HistPlan plan_hist(int n) {
    HistPlan hp;
    hist_first.assign((size_t)n, 0);
    hist_count.assign((size_t)n, 0);
    std::vector<int> item_of;  // launch item -> segment
    for (int i = 0; i < n; ++i) {
      if (hist_seg[(size_t)i].count <= 0) continue;
      h_hitems[hp.ni] = HistItem{hist_seg[(size_t)i], (int)hp.total, 0};
      item_of.push_back(i);
      hp.total += hist_seg[(size_t)i].count;
      ++hp.ni;
    }
    if (hp.ni == 0) return hp;
    // a block's partial slot holds the sum of all its rows, packed as the LDS bins
    // are: more rows than one LDS accumulation may hold could carry one packed field
    // into the next, so no block takes more than rows_per_flush rows
    const int nblocks = (int)hist_row_blocks(hp.total, hist_min_rows, hist_max_blocks, rows_per_flush);
    hp.rpb = (int)((hp.total + nblocks - 1) / nblocks);
    hp.nblocks = (int)((hp.total + hp.rpb - 1) / hp.rpb);
    // slots: block b gets one per item it touches
    std::vector<int> first_item((size_t)hp.nblocks);
    int slot = 0, it = 0;
    for (int b = 0; b < hp.nblocks; ++b) {
      const long long s0 = (long long)b * hp.rpb, e0 = std::min<long long>(s0 + hp.rpb, hp.total);
      while (h_hitems[it].row0 + h_hitems[it].seg.count <= s0) ++it;
      int last = it;
      while (h_hitems[last].row0 + h_hitems[last].seg.count < e0) ++last;
      first_item[(size_t)b] = it;
      h_hblock[b] = slot;
      slot += last - it + 1;
    }
    if (slot > max_partials) throw std::runtime_error("partial buffer too small");
    for (int k = 0; k < hp.ni; ++k) {
      const int b0 = h_hitems[k].row0 / hp.rpb, b1 = (h_hitems[k].row0 + h_hitems[k].seg.count - 1) / hp.rpb;
      hist_first[(size_t)item_of[(size_t)k]] = h_hblock[b0] + (k - first_item[(size_t)b0]);
      hist_count[(size_t)item_of[(size_t)k]] = b1 - b0 + 1;
    }
    return hp;
  }

  // Algorithmic bytes per row (the reference's index-based model): a root build
  // reads F bins and the payload P; a leaf build also reads a 4-byte row index;
  // a partition reads one bin and reads and writes a 4-byte row index. Moved
  // bytes are what these kernels touch: whole row records.
  void run_hist(const HistPlan& hp, double rows, bool root) {
    if (hp.ni == 0) return;
    const double hb = rows * (G + payload_bytes() + (root ? 0 : 4));
    timed(root ? 0 : 1, hb, rows * (stride + (root ? payload_bytes() : 0)), [&] {
      hist_launch(ds, d_hitems, hp.ni, (int)hp.total, hp.nblocks, hp.rpb, d_hblock, nullptr);
    });
    stats.hist_rows += (int64_t)rows;
  }

  // Partition of the uploaded items (tiles 0 .. ntasks-1).
  void run_partition(int nitems, int ntasks, double part_rows) {
    if (ntasks <= 0) return;
    if (ntasks > max_part_tiles) throw std::runtime_error("too many partition tiles");
    ++epoch;
    timed(2, part_rows * 9.0, part_rows * 2.0 * stride, [&] {
      launch_partition(ds, d_pitems, nitems, ntasks, d_cursors, max_items, epoch, stream);
    });
    stats.part_rows += (int64_t)part_rows;
  }

  // Reduce items of the staged search items; returns their number.
  int plan_scan(int nitems) {
    int nr = 0;
    for (int k = 0; k < nitems; ++k) {
      const ScanItem& si = h_sitems[k];
      h_ritems[nr++] = ReduceItem{si.small.slot, si.small.part_first, si.small.nblk, 0};
      if (si.has_large && !si.use_subtract)
        h_ritems[nr++] = ReduceItem{si.large.slot, si.large.part_first, si.large.nblk, 0};
    }
    return nr;
  }

  // Sums the partials into the slots, searches, folds, and reads the best
  // splits and splittable flags back (k_fold writes them to pinned host memory).
  void run_scan(int nitems, int nr) {
    if (nitems <= 0) return;
    ScanParams sp = scan_params();
    // the batched form with by-node sampling folds on the host, later: the
    // candidates are copied back after the fold, and a stamp follows the copy
    const bool copy_out = batched && cols.by_node_active();
    const unsigned seq = copy_out ? 0u : ++stamp_seq;
    timed(3, 0.0, 0.0, [&] {
      launch_reduce(ds, d_ritems, nr, stream);
      launch_scan(ds, d_sitems, nitems, sp, d_used, d_out, d_fold, d_flags, stream, nullptr,
                  copy_out ? nullptr : dev_alias(h_stamp), seq, d_tickets + 1);
    });
    if (copy_out) {
      timed(8, 0.0, 0.0, [&] {
        launch_copy(dev_alias(h_out), d_out, sizeof(SplitOut) * (size_t)F * 2 * (size_t)nitems, stream);
      });
      sync_spin();
    } else {
      wait_stamp(seq);
    }
    stats.rounds += 1;
  }

  void set_best(int leaf, const HostSplit& v) {
    best[(size_t)leaf] = v;
    amax.update(leaf);
  }
  void set_gain_min(int left_leaf, int right_leaf) {
    best[(size_t)left_leaf].o.gain = kMinScore;
    amax.update(left_leaf);
    if (right_leaf >= 0) {
      best[(size_t)right_leaf].o.gain = kMinScore;
      amax.update(right_leaf);
    }
  }

  // The folded best split of (item, side) k = 2 * item + side and the
  // splittable flags of its features (serial.rs:1500-1566; folded on the device).
  // searched: the features whose flags the search set (null: all); the flags of
  // the others are left as they are, as the reference leaves them
  HostSplit fold(int k, char* splittable, const int* searched = nullptr) const {
    HostSplit b;
    if (h_fold[k].inner >= 0) {
      b.o = h_fold[k].o;
      b.inner = h_fold[k].inner;
    }
    for (int f = 0; f < F; ++f)
      if (!searched || searched[f]) splittable[f] = (char)h_flags[(size_t)k * F + f];
    return b;
  }

  // The fold of one side's per-feature candidates over the features that were
  // searched and are in the by-node sample: the highest gain, ties to the lower
  // real feature index, NaN never (split_info.rs:9-96 better_than), in feature order.
  HostSplit fold_host(const SplitOut* out, const uint8_t* searched, const std::vector<char>& sample) const {
    HostSplit b;
    double bg = kMinScore;
    int bk = INT_MAX, bf = -1;
    for (int f = 0; f < F; ++f) {
      if (!searched[f] || !sample[(size_t)f]) continue;
      const double gain = out[f].gain;
      const int key = out[f].feature == -1 ? INT_MAX : out[f].feature;
      if (!(gain != gain) && (gain > bg || (gain == bg && key < bk))) {
        bg = gain;
        bk = key;
        bf = f;
      }
    }
    if (bf >= 0) {
      b.o = out[bf];
      b.inner = bf;
    }
    return b;
  }

  // The rule and the tree update of one committed split (split, serial.rs:1685-1876).
  // Returns the right leaf index.
  int commit_split(Tree& t, int best_leaf, const HostSplit& info) {
    const int inner = info.inner;
    const FeatDev& m = feats[(size_t)inner];
    const float gain = (float)(info.o.gain + P.min_gain_to_split);
    const int lc = info.o.exact_left, rc = info.o.exact_right;
    if (m.bin_type == BK_BIN_CATEGORICAL) {
      std::vector<uint32_t> real;
      for (uint32_t b = 0; b < 256; ++b)
        if ((info.o.cat_bits[b / 32] >> (b % 32)) & 1u) real.push_back((uint32_t)bin2cat[(size_t)inner][b]);
      std::vector<uint32_t> bits_real = bitset_of(real);
      int node = t.split_common(best_leaf, inner, info.o.feature, info.o.left_output, info.o.right_output, lc, rc,
                                info.o.left_sum_hessian, info.o.right_sum_hessian, gain);
      t.dtype[(size_t)node] = (int8_t)(1 | (m.missing_type << 2));
      t.th_bin[(size_t)node] = (uint32_t)t.num_cat;
      t.th[(size_t)node] = (double)t.num_cat;
      t.num_cat += 1;
      t.cat_bound.push_back(t.cat_bound.back() + (int)bits_real.size());
      t.cat_th.insert(t.cat_th.end(), bits_real.begin(), bits_real.end());
    } else {
      int node = t.split_common(best_leaf, inner, info.o.feature, info.o.left_output, info.o.right_output, lc, rc,
                                info.o.left_sum_hessian, info.o.right_sum_hessian, gain);
      int8_t dt = 0;
      if (info.o.default_left) dt |= 2;
      dt = (int8_t)((dt & 3) | (m.missing_type << 2));
      t.dtype[(size_t)node] = dt;
      t.th_bin[(size_t)node] = (uint32_t)info.o.threshold;
      t.th[(size_t)node] = upper[(size_t)inner][(size_t)info.o.threshold];
    }
    // both children's branch features: the parent's path and this split's real
    // feature (serial.rs:1719-1722)
    branch[(size_t)t.num_leaves] = branch[(size_t)best_leaf];
    branch[(size_t)t.num_leaves].push_back(info.o.feature);
    branch[(size_t)best_leaf].push_back(info.o.feature);
    NodeRule& nd = tree_nodes[(size_t)(t.num_leaves - 1)];
    nd.rule = rule_of(info);
    nd.group = m.group;
    t.num_leaves += 1;
    stats.splits += 1;
    stats.committed_part_rows += (int64_t)lc + rc;
    if (children_searchable(t.leaf_depth[(size_t)best_leaf], lc, rc)) stats.committed_hist_rows += std::min(lc, rc);
    return t.num_leaves - 1;
  }

  RuleDev rule_of(const HostSplit& info) const {
    const FeatDev& m = feats[(size_t)info.inner];
    if (m.bin_type == BK_BIN_CATEGORICAL) return make_rule_categorical(m, info.o.cat_bits);
    return make_rule_numerical(m, (uint32_t)info.o.threshold, info.o.default_left != 0);
  }

  // before_find_best_split's data rules (serial.rs:780-799) for two fresh children
  bool children_searchable(int child_depth, int nl, int nr) const {
    if (P.max_depth > 0 && child_depth >= P.max_depth) return false;
    if (nr < P.min_data_in_leaf * 2 && nl < P.min_data_in_leaf * 2) return false;
    return true;
  }

  // ===========================================================================
  // Sequential form: one split per step, in the reference's order
  // ===========================================================================
  // before_find_best_split (serial.rs:779-825)
  bool before_find_best_split(const Tree& t, int left_leaf, int right_leaf) {
    if (P.max_depth > 0 && t.leaf_depth[(size_t)left_leaf] >= P.max_depth) {
      set_gain_min(left_leaf, right_leaf);
      return false;
    }
    int nl = left_leaf >= 0 ? segs[(size_t)left_leaf].count : 0;
    int nr = right_leaf >= 0 ? segs[(size_t)right_leaf].count : 0;
    if (nr < P.min_data_in_leaf * 2 && nl < P.min_data_in_leaf * 2) {
      set_gain_min(left_leaf, right_leaf);
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

  // is_feature_used from the parent's splittable flags (serial.rs:838-852)
  // with the tree's feature sample and the by-node samples, drawn here for the
  // smaller leaf and then the larger one (serial.rs:1329-1337)
  void compute_used() {
    stage_wait();
    std::vector<char> sm, lg;
    if (cols.by_node_active()) {
      sm = cols.by_node(branch[(size_t)smaller.leaf]);
      if (larger.leaf >= 0) lg = cols.by_node(branch[(size_t)larger.leaf]);
    }
    for (int f = 0; f < F; ++f) {
      used_feat[(size_t)f] = cols.bytree[(size_t)f];
      if (used_feat[(size_t)f] && parent_slot >= 0 && !pool.splittable[(size_t)parent_slot][(size_t)f]) {
        pool.splittable[smaller_slot][(size_t)f] = 0;
        used_feat[(size_t)f] = 0;
      }
      int w = 0;
      if (used_feat[(size_t)f]) {
        w = kUseSearch;
        if (sm.empty() || sm[(size_t)f]) w |= kUseSmall;
        if (lg.empty() || lg[(size_t)f]) w |= kUseLarge;
      }
      h_used[f] = w;
    }
  }

  void fold_sequential(bool has_large) {
    HostSplit sb = fold(0, pool.splittable[smaller_slot].data(), used_feat.data());
    set_best(smaller.leaf, sb);
    if (larger.leaf >= 0)
      set_best(larger.leaf,
               has_large ? fold(1, pool.splittable[(size_t)larger_slot].data(), used_feat.data()) : HostSplit());
  }

  Tree train_sequential(const float* g, const float* h, const int32_t* d_rows, int nrows) {
    const LeafSplits root = before_train(g, h, d_rows, nrows);
    Tree t(P.num_leaves);
    t.set_leaf_output(0, leaf_output_root(P, smaller.sum_g, smaller.sum_h));
    if (root.num_data > 0 && before_find_best_split(t, 0, -1)) {
      compute_used();
      hist_seg.assign(1, segs[0]);
      const HistPlan hp = plan_hist(1);
      ScanItem& si = h_sitems[0];
      si = ScanItem{};
      si.small = leaf_args(t, smaller, (int)smaller_slot);
      si.small.part_first = hist_first[0];
      si.small.nblk = hist_count[0];
      const int nr = plan_scan(1);
      upload_round(0, &hp, 1, nr, true);
      run_hist(hp, root.num_data, true);
      stats.committed_hist_rows += root.num_data;
      run_scan(1, nr);
      fold_sequential(false);
    }
    for (int step = 0; step < P.num_leaves - 1; ++step) {
      const size_t bl = (size_t)amax.top();
      if (best[bl].o.gain <= 0.0) break;
      const HostSplit info = best[bl];
      const int best_leaf = (int)bl;
      const int right = commit_split(t, best_leaf, info);
      const int lc = info.o.exact_left, rc = info.o.exact_right;
      LeafSplits l{best_leaf, lc, info.o.left_sum_gradient, info.o.left_sum_hessian, info.o.left_output,
                   info.o.left_int_sum};
      LeafSplits r{right, rc, info.o.right_sum_gradient, info.o.right_sum_hessian, info.o.right_output,
                   info.o.right_int_sum};
      smaller = lc < rc ? l : r;
      larger = lc < rc ? r : l;
      const Segment ps = segs[(size_t)best_leaf];
      const int dst = ps.buf == 1 ? 2 : 1;
      segs[(size_t)best_leaf] = Segment{dst, ps.begin, lc};
      segs[(size_t)right] = Segment{dst, ps.begin + lc, rc};
      // the next step's before_find_best_split runs now: its pool moves happen
      // before any other pool access, as in the reference
      const bool search = step + 1 < P.num_leaves - 1 && before_find_best_split(t, best_leaf, right);
      if (search) compute_used();
      stage_wait();
      int nt = 0;
      PartItem& pi = h_pitems[0];
      pi = PartItem{};
      pi.seg = ps;
      pi.dst_buf = dst;
      pi.split_group = feats[(size_t)info.inner].group;
      pi.left_total = lc;
      pi.chunk = part_tile;
      pi.rule = rule_of(info);
      pi.first_task = add_tasks(ps.count, pi.chunk, nt);
      pi.ntasks = nt;
      if (!search) {
        upload_round(1, nullptr, 0, 0);
        run_partition(1, nt, ps.count);
      } else {
        // after the partition the smaller child's rows are contiguous: build its
        // histogram, and the larger child's too when the parent's was evicted
        const bool use_subtract = parent_slot >= 0;
        ScanItem& si = h_sitems[0];
        si = ScanItem{};
        si.small = leaf_args(t, smaller, (int)smaller_slot);
        si.has_large = larger.leaf >= 0 && larger_slot >= 0;
        si.use_subtract = use_subtract;
        const Segment sseg = segs[(size_t)smaller.leaf];
        hist_seg.assign(1, sseg);
        double hrows = sseg.count;
        const bool rebuild = si.has_large && !use_subtract;
        if (si.has_large) si.large = leaf_args(t, larger, (int)larger_slot);
        if (rebuild) {
          const Segment lseg = segs[(size_t)larger.leaf];
          hist_seg.push_back(lseg);
          stats.committed_hist_rows += lseg.count;
          hrows += lseg.count;
        }
        const HistPlan hp = plan_hist((int)hist_seg.size());
        si.small.part_first = hist_first[0];
        si.small.nblk = hist_count[0];
        if (rebuild) {
          si.large.part_first = hist_first[1];
          si.large.nblk = hist_count[1];
        }
        const int nr = plan_scan(1);
        upload_round(1, &hp, 1, nr, true);
        run_partition(1, nt, ps.count);
        run_hist(hp, hrows, false);
        run_scan(1, nr);
        fold_sequential(si.has_large != 0);
      }
    }
    return t;
  }

  // ===========================================================================
  // Batched form: every round expands the top frontier leaves in one launch of
  // each kernel; commits replay the reference's best-first order exactly.
  // ===========================================================================
  // The most row blocks plan_hist cuts for any build of at most `rows` rows (0:
  // none): the block target nb grows with the rows, and a build of r rows takes
  // ceil(r / ceil(r / nb(r))) <= nb(r) blocks.
  int hist_blocks_for(long long rows) const {
    return (int)hist_row_blocks(rows, hist_min_rows, hist_max_blocks, rows_per_flush);
  }

  // One round of batched growth with 32-bit bins: the partition counts the
  // children (right rows fill each segment from its end), k_children turns the
  // counts into the round's histogram, reduce and search items, and the host
  // reads the counts back with the folded splits. The round's histogram grid is
  // sized for hist_bound rows (every smaller child holds at most half of its
  // parent's rows); the device plans at most that.
  template <typename Keep>
  void run_derived_round(int ni, int nt, double part_rows, long long hist_bound, const std::vector<int>& order,
                         bool host_fold, bool part_hist, bool small_tiles, Keep&& keep) {
    const auto l0 = std::chrono::steady_clock::now();
    begin_uploads();
    d_pitems = put(h_pitems, (size_t)ni);
    const ChildSpec* u_specs = put(v_cspec.data(), (size_t)ni);
    d_used = put(h_used, (size_t)F * ni);
    send(true);
    ChildParams cp{};
    cp.max_depth = P.max_depth;
    cp.min_data_in_leaf = P.min_data_in_leaf;
    cp.hist_min_rows = hist_min_rows;
    cp.hist_max_blocks = hist_max_blocks;
    cp.rows_per_flush = rows_per_flush;
    cp.max_partials = max_partials;
    cp.F = F;
    cp.part_hist = part_hist ? 1 : 0;
    const int hb = part_hist ? 0 : hist_blocks_for(hist_bound);
    if (hb > max_partials) throw std::runtime_error("partial buffer too small");
    const ScanParams sp = scan_params();
    // the derivation: in the partition's last block, or a launch of its own
    DeriveArgs da{u_specs, d_chist, d_cscan, d_cred, d_chdr, dev_alias(h_kcount), d_tickets, cp,
                  derive_in_partition && ni <= kPartDeriveMax ? 1 : 0, 0};
    if (nt > max_part_tiles) throw std::runtime_error("too many partition tiles");
    ++epoch;
    timed(2, part_rows * 9.0, part_rows * 2.0 * stride, [&] {
      if (part_hist)
        launch_partition_hist(ds, d_pitems, ni, nt, d_cursors, max_items, epoch, d_group_used, stream, &da);
      else
        launch_partition(ds, d_pitems, ni, nt, d_cursors, max_items, epoch, stream, nullptr, &da, small_tiles);
    });
    stats.part_rows += (int64_t)part_rows;
    if (!da.on)
      timed(7, 0.0, 0.0, [&] {
        launch_children(d_pitems, ni, d_cursors, max_items, epoch, u_specs, cp, d_chist, d_cscan, d_cred, d_chdr,
                        dev_alias(h_kcount), stream);
      });
    if (hb > 0)
      timed(1, 0.0, 0.0, [&] {
        hist_launch(ds, d_chist, 0, 0, hb, 0, nullptr, d_chdr);
      });
    const unsigned seq = host_fold ? 0u : ++stamp_seq;
    timed(3, 0.0, 0.0, [&] {
      // ni splits: at most ni search items, one grid row each
      launch_reduce(ds, d_cred, ni, stream, d_chdr, ni);
      launch_scan(ds, d_cscan, ni, sp, d_used, d_out, d_fold, d_flags, stream, d_chdr,
                  host_fold ? nullptr : dev_alias(h_stamp), seq, d_tickets + 1, ni);
    });
    stats.launch_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - l0).count();
    if (host_fold) {
      timed(8, 0.0, 0.0, [&] {
        launch_copy(dev_alias(h_out), d_out, sizeof(SplitOut) * (size_t)F * 2 * (size_t)ni, stream);
      });
      sync_spin();
    } else {
      wait_stamp(seq);
    }
    const auto b0 = std::chrono::steady_clock::now();
    if (h_kcount[2 * ni] != 0) throw std::runtime_error("partition count mismatch on the device");
    long long hist_rows = 0;
    int k = 0;
    for (int j = 0; j < ni; ++j) {
      Node& X = nodes[(size_t)leaf_node[(size_t)order[(size_t)j]]];
      const int lc = h_kcount[2 * j], rc = h_kcount[2 * j + 1];
      if (lc < 0 || rc < 0 || (long long)lc + rc != X.seg.count)
        throw std::runtime_error("partition count mismatch on the device");
      const int side = h_pitems[j].hist_side;
      if (part_hist && side >= 0) hist_rows += side == 0 ? lc : rc;  // built in the partition
      const int dst = X.seg.buf == 1 ? 2 : 1;
      const int xl = X.child[0], xr = X.child[1];
      Node& cl = nodes[(size_t)xl];
      Node& cr = nodes[(size_t)xr];
      cl.seg = Segment{dst, X.seg.begin, lc};
      cr.seg = Segment{dst, X.seg.begin + lc, rc};
      cl.ls.num_data = lc;
      cr.ls.num_data = rc;
      const bool search = children_searchable(X.depth + 1, lc, rc);
      X.kids_searched = search;
      if (!search) continue;
      const bool left_small = part_hist ? side == 0 : lc < rc;  // the built side
      const int xs = left_small ? xl : xr, xg = left_small ? xr : xl;
      nodes[(size_t)xs].slot = v_cspec[(size_t)j].new_slot;
      nodes[(size_t)xg].slot = X.slot;
      if (!part_hist) hist_rows += std::min(lc, rc);
      nodes[(size_t)xs].best = fold(2 * k, flags_of(xs));
      nodes[(size_t)xg].best = fold(2 * k + 1, flags_of(xg));
      if (host_fold) {
        keep(xs, k, 0, j);
        keep(xg, k, 1, j);
        nodes[(size_t)xs].best = nodes[(size_t)xg].best = HostSplit();  // folded when X is committed
      }
      ++k;
    }
    stats.hist_rows += hist_rows;
    if (k > 0) stats.rounds += 1;  // rounds count search launches with work
    stats.readback_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b0).count();
    if (prof.tree_on) {
      // a histogram built in the partition adds its bytes to the partition's
      const double hr = (double)hist_rows;
      const int kind = part_hist ? 2 : 1;
      if ((prof.mask >> kind) & 1u) prof.add_bytes(kind, hr * (G + payload_bytes() + 4), part_hist ? 0.0 : hr * stride);
    }
  }

  Tree train_batched(const float* g, const float* h, const int32_t* d_rows, int nrows) {
    LeafSplits root = before_train(g, h, d_rows, nrows);
    Tree t(P.num_leaves);
    t.set_leaf_output(0, leaf_output_root(P, root.sum_g, root.sum_h));
    nodes.clear();
    const size_t max_nodes = (size_t)P.num_leaves * 4 + 4;
    nodes.reserve(max_nodes);
    if (node_flags.size() < max_nodes * F) node_flags.assign(max_nodes * F, 1);
    Node r0;
    r0.seg = segs[0];
    r0.ls = root;
    r0.slot = 0;
    std::fill(node_flags.begin(), node_flags.begin() + F, (char)1);
    nodes.push_back(r0);
    const bool host_fold = cols.by_node_active();
    if (host_fold && node_cands.size() < max_nodes * F) {
      node_cands.resize(max_nodes * F);
      node_searched.resize(max_nodes * F);
    }
    // a searched node's candidates (search item k2, feature row urow of the
    // round's used table), kept for its fold at its parent's commit
    auto keep = [&](int node, int k2, int side, int urow) {
      std::memcpy(&node_cands[(size_t)node * F], h_out + ((size_t)2 * k2 + side) * F, sizeof(SplitOut) * (size_t)F);
      for (int f = 0; f < F; ++f) node_searched[(size_t)node * F + f] = (uint8_t)(h_used[urow * F + f] & kUseSearch);
    };
    // 32-bit bins: the children's counts come from the partition (k_children),
    // which may also build the histograms (part_hist)
    const bool derive = ds.hist32 != 0;
    const bool part_hist = derive && part_hist_ok && part_hist_on;
    leaf_node.assign((size_t)P.num_leaves, -1);
    leaf_node[0] = 0;
    int next_slot = 1;
    // the root search (serial.rs:591-604 with the root's before_find_best_split)
    if (root.num_data > 0 && root.num_data >= P.min_data_in_leaf * 2) {
      for (int f = 0; f < F; ++f) h_used[f] = cols.bytree[(size_t)f] ? kUseAll : 0;
      hist_seg.assign(1, nodes[0].seg);
      const HistPlan hp = plan_hist(1);
      ScanItem& si = h_sitems[0];
      si = ScanItem{};
      si.small = leaf_args(t, root, 0);
      si.small.part_first = hist_first[0];
      si.small.nblk = hist_count[0];
      si.used_off = 0;
      si.out_off = 0;
      const int nr = plan_scan(1);
      upload_round(0, &hp, 1, nr, true);
      run_hist(hp, root.num_data, true);
      stats.committed_hist_rows += root.num_data;
      run_scan(1, nr);
      nodes[0].best = fold(0, flags_of(0));
      if (host_fold) {
        keep(0, 0, 0, 0);
        nodes[0].best = fold_host(&node_cands[0], &node_searched[0], cols.by_node(branch[0]));
      }
    } else {
      nodes[0].best.o.gain = kMinScore;
    }
    set_best(0, nodes[0].best);
    std::vector<int> order;
    bool done = false;
    using clk = std::chrono::steady_clock;
    auto since_ms = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    while (!done) {
      const auto r0 = clk::now();
      // replay commits in the reference's order while the needed children are known
      for (;;) {
        if (t.num_leaves >= P.num_leaves) {
          done = true;
          break;
        }
        const size_t bl = (size_t)amax.top();
        if (best[bl].o.gain <= 0.0) {
          done = true;
          break;
        }
        const int x = leaf_node[bl];
        if (!nodes[(size_t)x].expanded) break;
        const int cl = nodes[(size_t)x].child[0], cr = nodes[(size_t)x].child[1];
        HostSplit info = best[bl];
        // the children's exact counts (the search's, or the partition's with 32-bit bins)
        info.o.exact_left = nodes[(size_t)cl].ls.num_data;
        info.o.exact_right = nodes[(size_t)cr].ls.num_data;
        const int right = commit_split(t, (int)bl, info);
        nodes[(size_t)cl].ls.leaf = (int)bl;
        nodes[(size_t)cr].ls.leaf = right;
        // the reference searches these children, and draws their by-node samples
        // (smaller child first), only when another split may follow
        if (host_fold && nodes[(size_t)x].kids_searched && t.num_leaves < P.num_leaves) {
          const bool left_small = info.o.exact_left < info.o.exact_right;
          const int sm = left_small ? cl : cr, lg = left_small ? cr : cl;
          const std::vector<char> ms = cols.by_node(branch[(size_t)(left_small ? (int)bl : right)]);
          const std::vector<char> ml = cols.by_node(branch[(size_t)(left_small ? right : (int)bl)]);
          nodes[(size_t)sm].best = fold_host(&node_cands[(size_t)sm * F], &node_searched[(size_t)sm * F], ms);
          nodes[(size_t)lg].best = fold_host(&node_cands[(size_t)lg * F], &node_searched[(size_t)lg * F], ml);
        }
        leaf_node[bl] = cl;
        leaf_node[(size_t)right] = cr;
        set_best((int)bl, nodes[(size_t)cl].best);
        set_best(right, nodes[(size_t)cr].best);
      }
      if (done) break;
      stats.replay_ms += since_ms(r0);
      const auto p0 = clk::now();
      // expand the best unexpanded frontier leaves, at most the remaining budget
      stage_wait();
      order.clear();
      for (int l = 0; l < t.num_leaves; ++l) {
        const Node& nd = nodes[(size_t)leaf_node[(size_t)l]];
        if (!nd.expanded && nd.best.o.gain > 0.0) order.push_back(l);
      }
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return best[(size_t)a].better_than(best[(size_t)b]); });
      const size_t budget = (size_t)(P.num_leaves - t.num_leaves);
      if (order.size() > budget) order.resize(budget);
      int ni = 0, nt = 0, ns = 0;
      double part_rows = 0.0, hist_rows = 0.0;
      long long hist_bound = 0;  // 32-bit bins: the most rows the round's histograms take
      hist_seg.clear();
      // a derived round of few rows partitions in short tiles, so that its
      // blocks still cover the CUs
      long long round_rows = 0;
      for (int l : order) round_rows += nodes[(size_t)leaf_node[(size_t)l]].seg.count;
      const bool small_tiles = derive && !part_hist && round_rows <= kSmallTileRoundRows &&
                               round_rows / part_tile_small + (long long)order.size() + 1 <= max_part_tiles;
      const int tile = small_tiles ? part_tile_small : part_tile;
      for (int l : order) {
        const int x = leaf_node[(size_t)l];
        Node& X = nodes[(size_t)x];
        const HostSplit& info = X.best;
        const int dst = X.seg.buf == 1 ? 2 : 1;
        if (derive) {
          Node cl, cr;
          cl.depth = cr.depth = X.depth + 1;
          cl.ls = LeafSplits{-1, 0, info.o.left_sum_gradient, info.o.left_sum_hessian, info.o.left_output,
                             info.o.left_int_sum};
          cr.ls = LeafSplits{-1, 0, info.o.right_sum_gradient, info.o.right_sum_hessian, info.o.right_output,
                             info.o.right_int_sum};
          cl.best.o.gain = cr.best.o.gain = kMinScore;
          PartItem& pi = h_pitems[ni];
          pi = PartItem{};
          pi.seg = X.seg;
          pi.dst_buf = dst;
          pi.split_group = feats[(size_t)info.inner].group;
          pi.left_total = -1;  // right rows fill the segment from its end
          // the child whose histogram the partition builds: the smaller by the
          // search's estimated counts (any side gives the same integer histograms)
          pi.hist_side = -1;
          if (part_hist && !(P.max_depth > 0 && X.depth + 1 >= P.max_depth))
            pi.hist_side = info.o.left_count < info.o.right_count ? 0 : 1;
          pi.chunk = tile;
          pi.rule = rule_of(info);
          pi.first_task = add_tasks(X.seg.count, pi.chunk, nt);
          pi.ntasks = nt - pi.first_task;
          part_rows += X.seg.count;
          ChildSpec& cs = v_cspec[(size_t)ni];
          cs = ChildSpec{};
          cs.kid[0] = leaf_args_with(cl.ls, -1, cl.ls.weight);
          cs.kid[1] = leaf_args_with(cr.ls, -1, cr.ls.weight);
          cs.depth = X.depth + 1;
          // the smaller child's slot, taken whether or not the pair is searched
          if ((size_t)next_slot >= num_slots) throw std::runtime_error("histogram slots exhausted");
          cs.new_slot = next_slot++;
          cs.parent_slot = X.slot;
          const char* xf = flags_of(x);
          for (int f = 0; f < F; ++f) h_used[ni * F + f] = xf[f] && cols.bytree[(size_t)f] ? kUseAll : 0;
          hist_bound += X.seg.count / 2;
          if (nodes.size() + 2 > max_nodes) throw std::runtime_error("node table too small");
          X.child[0] = (int)nodes.size();
          X.child[1] = (int)nodes.size() + 1;
          X.expanded = true;
          X.kids_searched = false;  // known from the counts
          nodes.push_back(cl);
          nodes.push_back(cr);
          ++ni;
          continue;
        }
        const int lc = info.o.exact_left, rc = info.o.exact_right;
        Node cl, cr;
        cl.depth = cr.depth = X.depth + 1;
        cl.seg = Segment{dst, X.seg.begin, lc};
        cr.seg = Segment{dst, X.seg.begin + lc, rc};
        cl.ls = LeafSplits{-1, lc, info.o.left_sum_gradient, info.o.left_sum_hessian, info.o.left_output,
                           info.o.left_int_sum};
        cr.ls = LeafSplits{-1, rc, info.o.right_sum_gradient, info.o.right_sum_hessian, info.o.right_output,
                           info.o.right_int_sum};
        cl.best.o.gain = cr.best.o.gain = kMinScore;
        const bool search = children_searchable(X.depth + 1, lc, rc);
        const bool left_small = lc < rc;
        PartItem& pi = h_pitems[ni];
        pi = PartItem{};
        pi.seg = X.seg;
        pi.dst_buf = dst;
        pi.split_group = feats[(size_t)info.inner].group;
        pi.left_total = lc;
        pi.chunk = part_tile;
        pi.rule = rule_of(info);
        pi.first_task = add_tasks(X.seg.count, pi.chunk, nt);
        pi.ntasks = nt - pi.first_task;
        part_rows += X.seg.count;
        if (search) {
          hist_rows += std::min(lc, rc);
          // the smaller child gets a new slot; the larger one reuses the parent's
          Node& sm = left_small ? cl : cr;
          Node& lg = left_small ? cr : cl;
          if ((size_t)next_slot >= num_slots) throw std::runtime_error("histogram slots exhausted");
          sm.slot = next_slot++;
          lg.slot = X.slot;
          ScanItem& si = h_sitems[ns];
          si = ScanItem{};
          // a child's parent output is its own output from the split (never the root's formula)
          si.small = leaf_args_with(sm.ls, sm.slot, sm.ls.weight);
          hist_seg.push_back(sm.seg);
          si.large = leaf_args_with(lg.ls, lg.slot, lg.ls.weight);
          si.has_large = 1;
          si.use_subtract = 1;
          si.used_off = ns * F;
          si.out_off = ns * 2 * F;
          const char* xf = flags_of(x);
          for (int f = 0; f < F; ++f) h_used[ns * F + f] = xf[f] && cols.bytree[(size_t)f] ? kUseAll : 0;
          ++ns;
        }
        if (nodes.size() + 2 > max_nodes) throw std::runtime_error("node table too small");
        X.child[0] = (int)nodes.size();
        X.child[1] = (int)nodes.size() + 1;
        X.expanded = true;
        X.kids_searched = search;
        nodes.push_back(cl);
        nodes.push_back(cr);
        ++ni;
      }
      if (ni == 0) break;
      stats.plan_ms += since_ms(p0);
      if (derive) {
        run_derived_round(ni, nt, part_rows, hist_bound, order, host_fold, part_hist, small_tiles, keep);
        continue;
      }
      HistPlan hp;
      int nr = 0;
      if (ns > 0) {
        // histograms of the smaller children, contiguous after the partition
        hp = plan_hist(ns);
        for (int k2 = 0; k2 < ns; ++k2) {
          h_sitems[k2].small.part_first = hist_first[(size_t)k2];
          h_sitems[k2].small.nblk = hist_count[(size_t)k2];
        }
        nr = plan_scan(ns);
      }
      upload_round(ni, &hp, ns, nr, ns > 0);
      run_partition(ni, nt, part_rows);
      if (ns > 0) {
        run_hist(hp, hist_rows, false);
        run_scan(ns, nr);
      }
      // fold the children's candidates
      int k = 0;
      for (int l : order) {
        Node& X = nodes[(size_t)leaf_node[(size_t)l]];
        const HostSplit& info = X.best;
        const int lc = info.o.exact_left, rc = info.o.exact_right;
        if (!children_searchable(X.depth + 1, lc, rc)) continue;
        const bool left_small = lc < rc;
        Node& sm = nodes[(size_t)X.child[left_small ? 0 : 1]];
        Node& lg = nodes[(size_t)X.child[left_small ? 1 : 0]];
        sm.best = fold(2 * k, flags_of(X.child[left_small ? 0 : 1]));
        lg.best = fold(2 * k + 1, flags_of(X.child[left_small ? 1 : 0]));
        if (host_fold) {
          keep(X.child[left_small ? 0 : 1], k, 0, k);
          keep(X.child[left_small ? 1 : 0], k, 1, k);
          sm.best = lg.best = HostSplit();  // folded when X is committed
        }
        ++k;
      }
    }
    for (int l = 0; l < t.num_leaves; ++l) segs[(size_t)l] = nodes[(size_t)leaf_node[(size_t)l]].seg;
    stats.speculative_waste = 0;
    for (const Node& nd : nodes)
      if (nd.expanded) ++stats.speculative_waste;
    stats.speculative_waste -= (t.num_leaves - 1);
    return t;
  }

  PlanParams plan_params() const {
    PlanParams pp{};
    pp.num_leaves = P.num_leaves;
    pp.F = F;
    pp.max_depth = P.max_depth;
    pp.min_data_in_leaf = P.min_data_in_leaf;
    pp.hist_min_rows = hist_min_rows;
    pp.hist_max_blocks = hist_max_blocks;
    pp.rows_per_flush = rows_per_flush;
    pp.part_tile = part_tile;
    pp.max_nodes = plan_max_nodes;
    pp.num_slots = (int)num_slots;
    pp.max_partials = max_partials;
    return pp;
  }
  // the most row blocks a planned histogram cuts: every row in one build
  int plan_hist_grid() const { return (int)hist_row_blocks(N, hist_min_rows, hist_max_blocks, rows_per_flush); }
  void launch_plan_round() {
    const PlanParams pp = plan_params();
    timed(7, 0.0, 0.0, [&] {
      launch_plan(d_pstate, d_pnodes, d_pnflags, d_pfold, d_pfflags, d_feats, d_bytree, pp, d_ppart, d_phist, d_pscan,
                  d_pred, d_pused, stream);
    });
  }
  // `count` planned rounds: the kernels of the round the last plan set up, then
  // the next plan. Grids hold the most work a round can: every row partitioned,
  // every row in a histogram, 2^(r-1) items in round r (the tree's depth at
  // round r is at most r - 1), at most max_items.
  void launch_rounds(int& r, int count, const ScanParams& sp, int hb) {
    const RoundHdr* hdr = &d_pstate->hdr;
    for (int k = 0; k < count; ++k, ++r) {
      const int items = std::min(max_items, 1 << std::min(r - 1, 20));
      ++epoch;
      timed(2, 0.0, 0.0, [&] {
        launch_partition(ds, d_ppart, max_items, max_part_tiles, d_cursors, max_items, epoch, stream, hdr);
      });
      timed(1, 0.0, 0.0, [&] {
        hist_launch(ds, d_phist, 0, 0, hb, 0, nullptr, hdr);
      });
      timed(3, 0.0, 0.0, [&] {
        launch_reduce(ds, d_pred, items, stream, hdr);
        launch_scan(ds, d_pscan, items, sp, d_pused, d_out, d_pfold, d_pfflags, stream, hdr);
      });
      launch_plan_round();
    }
  }

  // The tree from the planner's state and node table read back into h_plan: the
  // reference's commits replayed in order (commit_split), the leaves' segments,
  // and the round statistics.
  void rebuild_from_plan(Tree& t, bool search_root) {
    const PlanState* hs = (const PlanState*)h_plan;
    if (hs->pad) throw std::runtime_error("device planner overflow (nodes, slots or partials)");
    last_plan_rounds = hs->rounds;
    const PlanNode* hn = (const PlanNode*)(h_plan + sizeof(PlanState));
    if (hs->ncommit < 0 || hs->ncommit > P.num_leaves - 1) throw std::runtime_error("device planner: bad commit count");
    for (int c = 0; c < hs->ncommit; ++c) {
      const int bl = hs->commit[2 * c], x = hs->commit[2 * c + 1];
      HostSplit info;
      info.o = hn[x].best.o;
      info.inner = hn[x].best.inner;
      if (bl < 0 || bl >= t.num_leaves || x < 0 || x >= hs->num_nodes || info.inner < 0 || info.inner >= F)
        throw std::runtime_error("device planner: bad commit");
      commit_split(t, bl, info);
    }
    if (t.num_leaves != hs->num_leaves) throw std::runtime_error("device planner: leaf count disagrees");
    if (hs->num_nodes < 1 || hs->num_nodes > plan_max_nodes) throw std::runtime_error("device planner: bad node count");
    for (int l = 0; l < t.num_leaves; ++l) {
      const int x = hs->leaf_node[l];
      if (x < 0 || x >= hs->num_nodes) throw std::runtime_error("device planner: bad leaf node");
      segs[(size_t)l] = hn[x].seg;
    }
    int expanded = 0;
    for (int i = 0; i < hs->num_nodes; ++i) expanded += hn[i].expanded ? 1 : 0;
    stats.speculative_waste = expanded - (t.num_leaves - 1);
    stats.hist_rows += hs->hist_rows;
    stats.part_rows += hs->part_rows;
    stats.rounds += hs->search_rounds + (search_root ? 1 : 0);
    if (prof.tree_on) {
      const double hr = (double)hs->hist_rows, pr = (double)hs->part_rows;
      if ((prof.mask >> 1) & 1u) prof.add_bytes(1, hr * (G + payload_bytes() + 4), hr * stride);
      if ((prof.mask >> 2) & 1u) prof.add_bytes(2, pr * 9.0, pr * 2.0 * stride);
    }
  }

  // ===========================================================================
  // Batched form with the rounds planned on the device: the host launches the
  // root's search, then rounds of [plan, partition, histogram, reduce, search]
  // without waiting between them; k_plan replays the commits and picks each
  // round's work. The host waits once the expected number of rounds is queued,
  // reads the planner's state (complete or not) and the node table, and rebuilds
  // the tree by replaying the same commits (commit_split), so the result is the
  // host-planned batched form's tree, bit for bit.
  // ===========================================================================
  Tree train_planned(const float* g, const float* h, const int32_t* d_rows, int nrows) {
    LeafSplits root = before_train(g, h, d_rows, nrows);
    Tree t(P.num_leaves);
    t.set_leaf_output(0, leaf_output_root(P, root.sum_g, root.sum_h));
    const bool search_root = root.num_data > 0 && root.num_data >= P.min_data_in_leaf * 2;
    const ScanParams sp = scan_params();
    // the root's search, as train_batched runs it
    HistPlan hp;
    int nr = 0;
    if (search_root) {
      for (int f = 0; f < F; ++f) h_used[f] = cols.bytree[(size_t)f] ? kUseAll : 0;
      hist_seg.assign(1, segs[0]);
      hp = plan_hist(1);
      ScanItem& si = h_sitems[0];
      si = ScanItem{};
      si.small = leaf_args(t, root, 0);
      si.small.part_first = hist_first[0];
      si.small.nblk = hist_count[0];
      nr = plan_scan(1);
    }
    // one upload: the root's items and the planner's initial state
    begin_uploads();
    if (search_root) {
      d_hitems = put(h_hitems, (size_t)hp.ni);
      d_hblock = put(h_hblock, (size_t)hp.nblocks);
      d_sitems = put(h_sitems, 1);
      d_ritems = put(h_ritems, (size_t)nr);
      d_used = put(h_used, (size_t)F);
    }
    // the planner's initial state: leaf 0 is node 0, searched by the root item
    plan_init = PlanState{};
    plan_init.num_leaves = 1;
    plan_init.num_nodes = 1;
    plan_init.next_slot = 1;
    plan_init.nabsorb = search_root ? 1 : 0;
    plan_init.absorb[0] = 0;
    plan_init.absorb[1] = -1;
    plan_init.leaf_node[0] = 0;
    PlanNode n0{};
    n0.seg = segs[0];
    n0.child[0] = n0.child[1] = -1;
    n0.num_data = root.num_data;
    n0.sum_g = root.sum_g;
    n0.sum_h = root.sum_h;
    n0.weight = root.weight;
    n0.int_sum = root.int_sum;
    n0.best.o.gain = kMinScore;
    n0.best.o.feature = -1;
    n0.best.o.default_left = 1;
    n0.best.inner = -1;
    for (int f = 0; f < F; ++f) bytree_int[(size_t)f] = cols.bytree[(size_t)f] ? 1 : 0;
    const PlanState* u_state = put(&plan_init, 1);
    const PlanNode* u_node = put(&n0, 1);
    const int* u_bytree = put(bytree_int.data(), (size_t)F);
    const uint8_t* u_ones = put(ones_f.data(), (size_t)F);
    send(true);
    launch_plan_init(u_state, u_node, u_bytree, u_ones, d_pstate, d_pnodes, d_bytree, d_pnflags, F, stream);
    if (search_root) {
      run_hist(hp, root.num_data, true);
      stats.committed_hist_rows += root.num_data;
      timed(3, 0.0, 0.0, [&] {
        launch_reduce(ds, d_ritems, nr, stream);
        launch_scan(ds, d_sitems, 1, sp, d_used, d_out, d_pfold, d_pfflags, stream);
      });
    }
    const int hb = plan_hist_grid();
    launch_plan_round();
    // as many rounds as the last tree planned (trees of one model have similar
    // depths): the host waits once; a deeper tree takes more rounds in batches
    int r = 1, want = last_plan_rounds > 0 ? last_plan_rounds : 8;
    const size_t back = sizeof(PlanState) + sizeof(PlanNode) * (size_t)plan_max_nodes;
    const PlanState* hs = (const PlanState*)h_plan;
    for (;;) {
      launch_rounds(r, want, sp, hb);
      // the state and the node table are two allocations: two copies
      timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_plan), d_pstate, sizeof(PlanState), stream); });
      timed(8, 0.0, 0.0, [&] {
        launch_copy(dev_alias(h_plan) + sizeof(PlanState), d_pnodes, back - sizeof(PlanState), stream);
      });
      sync_spin();
      if (hs->done) break;
      // every planned round expands at least one node, and a tree expands fewer
      // than 2 * num_leaves: more rounds mean a planner that never finishes
      if (r > 2 * P.num_leaves + 8) throw std::runtime_error("device planner did not finish the tree");
      want = 4;
    }
    rebuild_from_plan(t, search_root);
    return t;
  }

  // ===========================================================================
  // Device-resident tree: the planned batched form with the discretizer's
  // scales, the root's sums and search, the final walk rules, the row-to-leaf
  // walk and the leaf renewal all queued on the device; the host waits once, at
  // the tree's end (more rounds than expected: once more per batch of 4), reads
  // the planner's tables and rebuilds the tree with the same commit replay.
  // ===========================================================================
  Tree train_resident(const float* g, const float* h) {
    stats = bk_train_stats{};
    stats.hist_bin_bytes = form == FORM_QUANT ? (ds.hist32 ? 4 : 8) : 0;
    branch.assign((size_t)P.num_leaves, std::vector<int>());
    if (cols.per_tree) {
      cols.next_tree(tree_slot, tree_period);
      set_groups_used(false);
    }
    stage_wait();
    for (int f = 0; f < F; ++f) h_bytree[f] = cols.bytree[(size_t)f] ? 1 : 0;
    // the noise rotation of this tree, drawn on a copy of the engine: the engine
    // advances only when the device accepts the discretization (serial.rs:692-703)
    Mt eng = use_start_eng;
    int start = 0;
    if (quant && stochastic)
      for (int j = 0; j < tree_period; ++j) {
        const int v = (int)eng.uint_incl((uint32_t)N);
        if (j == tree_slot) start = v;
      }
    *h_start = start;
    int nparts = 0;
    if (quant) {
      const double db = (double)N * (8.0 + (stochastic ? 16.0 : 0.0) + 2.0);
      timed(4, db, db, [&] {
        launch_max_abs(g, h, N, d_max4, d_sums3, stream);
        launch_disc_prep(d_max4, dev_alias(h_start), bins, const_hess ? 1 : 0, d_disc, stream);
        launch_discretize(g, h, d_gnoise, d_hnoise, N, 0, 0.0, 0.0, const_hess ? 1 : 0, stochastic ? 1 : 0,
                          (int16_t*)d_gh0, d_sums3, stream, d_disc);
      });
    } else {
      launch_pack_f32(g, h, N, const_hess ? 1 : 0, (float*)d_gh0, stream);
      nparts = launch_sum_gh(g, h, nullptr, N, N, d_sum_partials, stream);
    }
    RootArgs ra{};
    ra.num_data = N;
    ra.min_data_in_leaf = P.min_data_in_leaf;
    ra.quantized = quant ? 1 : 0;
    ra.const_hess = const_hess ? 1 : 0;
    ra.nparts = nparts;
    ra.bins = bins;
    ra.l1 = P.lambda_l1;
    ra.l2 = P.lambda_l2;
    ra.max_delta_step = P.max_delta_step;
    const PlanParams pp = plan_params();
    launch_root_prep(d_disc, d_sums3, d_sum_partials, h, ra, pp, d_pstate, d_pnodes, d_pnflags, dev_alias(h_bytree),
                     d_bytree, d_phist, d_pscan, d_pred, d_pused, d_err, stream);
    ScanParams sp = scan_params();
    sp.dyn = (const double*)d_disc;
    const bool search_root = N >= P.min_data_in_leaf * 2;
    const int hb = plan_hist_grid();
    const RoundHdr* hdr = &d_pstate->hdr;
    // the root's round: its histogram, reduce and search from the root set-up
    timed(0, (double)N * (G + payload_bytes()), (double)N * (stride + payload_bytes()), [&] {
      hist_launch(ds, d_phist, 0, 0, hb, 0, nullptr, hdr);
    });
    if (search_root) {
      stats.hist_rows += N;
      stats.committed_hist_rows += N;
    }
    timed(3, 0.0, 0.0, [&] {
      launch_reduce(ds, d_pred, 1, stream, hdr);
      launch_scan(ds, d_pscan, 1, sp, d_pused, d_out, d_pfold, d_pfflags, stream, hdr);
    });
    launch_plan_round();
    int r = 1, want = last_plan_rounds > 0 ? last_plan_rounds : 8;
    const size_t back = sizeof(PlanState) + sizeof(PlanNode) * (size_t)plan_max_nodes;
    const PlanState* hs = (const PlanState*)h_plan;
    for (;;) {
      launch_rounds(r, want, sp, hb);
      // the tree's walk rules, segments and renewal order, the row-to-leaf walk,
      // the renewal sums, and every table the host reads, before one wait
      launch_finalize(d_pstate, d_pnodes, d_feats, d_node_rules, d_node_count, d_segs, d_order, P.num_leaves, stream);
      timed(6, (double)N * (stride + 4.0), (double)N * (stride + 4.0), [&] {
        launch_leaf_of_tree(ds, d_node_rules, d_nodes, P.num_leaves - 1, nullptr, d_leaf_of, stream, d_node_count);
      });
      if (renew) {
        const int* flags = nullptr;
        if (d_renew_part_sums) {
          timed(5, (double)N * 12.0, (double)N * 12.0, [&] {
            launch_renew_par(N, d_leaf_of, g, h, P.num_leaves, d_renew_part_sums, d_renew_part_meta, stream);
            launch_renew_fix(d_segs, P.num_leaves, renew_nparts, d_renew_part_sums, d_renew_part_meta, d_renew,
                             d_renew_flags, stream);
          });
          flags = d_renew_flags;
        }
        timed(5, 0.0, 0.0,
              [&] {
                launch_renew_sums(N, d_leaf_of, d_order, P.num_leaves, g, h, flags, d_renew, renew_pieces, stream);
              });
      }
      timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_plan), d_pstate, sizeof(PlanState), stream); });
      timed(8, 0.0, 0.0, [&] {
        launch_copy(dev_alias(h_plan) + sizeof(PlanState), d_pnodes, back - sizeof(PlanState), stream);
      });
      timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_disc), d_disc, sizeof(DiscState), stream); });
      timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_tail), d_err, sizeof(int), stream); });
      sync_spin();
      if (hs->done) break;
      if (r > 2 * P.num_leaves + 8) throw std::runtime_error("device planner did not finish the tree");
      want = 4;
    }
    int err = 0;
    std::memcpy(&err, h_tail, sizeof(int));
    if (err) {
      BK_HIPC(hipMemsetAsync(d_err, 0, sizeof(int), stream));
      sync();
      leaf_of_valid = true;
      last_num_leaves = 0;
      if (err & 16) throw std::runtime_error("root sums are not exact in every order (device-resident tree)");
      throw std::runtime_error("partition count mismatch on the device");
    }
    // the discretizer's state as the host path keeps it
    if (quant && h_disc->ok) {
      use_start_eng = eng;
      grad_scale = h_disc->gs;
      hess_scale = h_disc->hs;
      last_int_g = h_disc->int_g;
      last_int_h = h_disc->int_h;
      last_root_g = h_disc->root_g;
      last_root_h = h_disc->root_h;
      sums_bagged = false;
    }
    if (!quant) h0 = h_disc->h0;
    const PlanNode* hn = (const PlanNode*)(h_plan + sizeof(PlanState));
    Tree t(P.num_leaves);
    t.set_leaf_output(0, leaf_output_root(P, hn[0].sum_g, hn[0].sum_h));
    smaller = LeafSplits{0, N, hn[0].sum_g, hn[0].sum_h, 0.0, hn[0].int_sum};
    larger = LeafSplits{};
    rebuild_from_plan(t, search_root);
    if (renew) {
      const double* sums = (const double*)(h_tail + 16);
      for (int leaf = 0; leaf < t.num_leaves; ++leaf)
        t.set_leaf_output(leaf, leaf_output_plain(sums[(size_t)2 * leaf], sums[(size_t)2 * leaf + 1], P.lambda_l1,
                                                  P.lambda_l2, P.max_delta_step));
    }
    return t;
  }

  // Leaf of every row of the last tree: every row walks the tree over its record
  // in original order (the partition moves no row ids).
  // internal nodes of the last tree as the walk takes them
  void fill_nodes(NodeRule* dst) const {
    for (int k = 0; k + 1 < last_num_leaves; ++k) {
      dst[k] = tree_nodes[(size_t)k];
      dst[k].left = last_left[(size_t)k];
      dst[k].right = last_right[(size_t)k];
    }
  }

  // `uploaded`: the nodes already on the device (sent with other per-tree data)
  void ensure_leaf_of(const NodeRule* uploaded = nullptr) {
    if (leaf_of_valid) return;
    const int ni = last_num_leaves - 1;
    const NodeRule* rules = uploaded;
    if (!rules && ni > 0) {
      stage_wait();
      fill_nodes(h_nodes);
      timed(8, 0.0, 0.0,
            [&] { launch_copy(d_node_rules, dev_alias(h_nodes), sizeof(NodeRule) * (size_t)ni, stream); });
      staged();
      rules = d_node_rules;
    }
    timed(6, (double)N * (stride + 4.0), (double)N * (stride + 4.0),
          [&] { launch_leaf_of_tree(ds, rules, d_nodes, ni, last_bagged ? d_bag_bits : nullptr, d_leaf_of, stream); });
    leaf_of_valid = true;
  }

  // d_rows: the tree's bag on the device (ascending rows inside the dataset), or null
  Tree train(const float* g, const float* h, const int32_t* d_rows = nullptr, int nrows = 0) {
    const auto t0 = std::chrono::steady_clock::now();
    wait_ms = 0.0;
    leaf_of_valid = false;
    prof.tree_on = prof.trees++ % prof.every == 0;
    use_hist32(hist32_wanted());
    if (resident_ok(d_rows != nullptr)) {
      Tree t = train_resident(g, h);
      last_bagged = false;
      last_num_leaves = t.num_leaves;
      last_left.assign(t.left.begin(), t.left.begin() + std::max(t.num_leaves - 1, 0));
      last_right.assign(t.right.begin(), t.right.begin() + std::max(t.num_leaves - 1, 0));
      leaf_of_valid = true;  // walked on the device
      if (prof.on && prof.used > Prof::kResolveAt) prof.resolve();
      const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      stats.wait_ms = wait_ms;
      stats.host_ms = total - wait_ms;
      return t;
    }
    Tree t = !batched ? train_sequential(g, h, d_rows, nrows)
                      : (planner_ok() ? train_planned(g, h, d_rows, nrows) : train_batched(g, h, d_rows, nrows));
    last_bagged = d_rows != nullptr;
    last_num_leaves = t.num_leaves;
    last_left.assign(t.left.begin(), t.left.begin() + std::max(t.num_leaves - 1, 0));
    last_right.assign(t.right.begin(), t.right.begin() + std::max(t.num_leaves - 1, 0));
    if (renew) {
      // one upload (synchronized at the end of the tree): the leaf segments, the
      // chain order (largest leaves first: the longest chains start on the first
      // blocks dispatched) and the tree's nodes for the row-to-leaf walk
      const int nl = t.num_leaves;
      for (int i = 0; i < nl; ++i) h_order[i] = i;
      std::stable_sort(h_order, h_order + nl,
                       [&](int a, int b) { return segs[(size_t)a].count > segs[(size_t)b].count; });
      fill_nodes(h_nodes);
      begin_uploads();
      const Segment* u_segs = put(segs.data(), (size_t)nl);
      const int* u_order = put(h_order, (size_t)nl);
      const NodeRule* u_nodes = put(h_nodes, (size_t)std::max(nl - 1, 1));
      send(true);
      const int* flags = nullptr;
      if (d_renew_part_sums && nl > 1) {
        // certified parallel sums; the chain runs only for leaves the certificate rejects.
        // The walk adds the partials as it finds each row's leaf (one pass over the rows).
        bool fused = false;
        if (!leaf_of_valid) {
          timed(6, (double)N * (stride + 12.0), (double)N * (stride + 12.0), [&] {
            fused = launch_leaf_of_tree_renew(ds, u_nodes, d_nodes, nl - 1, last_bagged ? d_bag_bits : nullptr,
                                              d_leaf_of, g, h, nl, d_renew_part_sums, d_renew_part_meta, stream);
          });
          leaf_of_valid = fused;
        }
        if (!fused) {
          ensure_leaf_of(u_nodes);
          timed(5, (double)N * 12.0, (double)N * 12.0, [&] {
            launch_renew_par(N, d_leaf_of, g, h, nl, d_renew_part_sums, d_renew_part_meta, stream);
          });
        }
        timed(5, 0.0, 0.0, [&] {
          launch_renew_fix(u_segs, nl, renew_nparts, d_renew_part_sums, d_renew_part_meta, d_renew, d_renew_flags,
                           stream);
        });
        flags = d_renew_flags;
      }
      timed(5, 0.0, 0.0,
            [&] { launch_renew_sums(N, d_leaf_of, u_order, nl, g, h, flags, d_renew, renew_pieces, stream); });
    }
    // one synchronization for the error flag and the renewal sums
    timed(8, 0.0, 0.0, [&] { launch_copy(dev_alias(h_tail), d_err, sizeof(int), stream); });
    sync_spin();
    int err = 0;
    std::memcpy(&err, h_tail, sizeof(int));
    if (err) {
      // the flag is cleared so the learner stays usable after a refused call
      BK_HIPC(hipMemsetAsync(d_err, 0, sizeof(int), stream));
      sync();
      leaf_of_valid = true;  // no tree: nothing to map
      last_num_leaves = 0;
    }
    if (err & 8) throw std::invalid_argument("bagged rows must be ascending and inside the dataset");
    if (err) throw std::runtime_error("partition count mismatch on the device");
    if (renew) {
      const double* sums = (const double*)(h_tail + 16);
      for (int leaf = 0; leaf < t.num_leaves; ++leaf)
        t.set_leaf_output(leaf, leaf_output_plain(sums[(size_t)2 * leaf], sums[(size_t)2 * leaf + 1], P.lambda_l1,
                                                  P.lambda_l2, P.max_delta_step));
    }
    // events are read back when many are pending (the stream is idle here) or
    // when the statistics are asked for, not per tree
    if (prof.on && prof.used > Prof::kResolveAt) prof.resolve();
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats.wait_ms = wait_ms;
    stats.host_ms = total - wait_ms;
    return t;
  }

  void export_tree(const Tree& t, bk_tree_out* o) const {
    o->num_leaves = t.num_leaves;
    o->num_cat = t.num_cat;
    size_t ni = (size_t)(t.num_leaves - 1), nl = (size_t)t.num_leaves;
    for (size_t i = 0; i < ni; ++i) {
      o->split_feature_inner[i] = t.sf_inner[i];
      o->split_feature[i] = t.sf[i];
      o->threshold_in_bin[i] = t.th_bin[i];
      o->threshold[i] = t.th[i];
      o->decision_type[i] = t.dtype[i];
      o->left_child[i] = t.left[i];
      o->right_child[i] = t.right[i];
      o->split_gain[i] = t.gain[i];
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
    for (size_t i = 0; i < t.cat_bound.size() && i < nl; ++i) o->cat_boundaries[i] = t.cat_bound[i];
    if ((int)t.cat_th.size() > o->cat_threshold_cap) throw std::invalid_argument("cat_threshold capacity too small");
    for (size_t i = 0; i < t.cat_th.size(); ++i) o->cat_threshold[i] = t.cat_th[i];
    o->cat_threshold_len = (int)t.cat_th.size();
  }

  void check_tree_matches(const bk_tree_out* t) const {
    if (t->num_leaves != last_num_leaves)
      throw std::invalid_argument("the tree does not match the learner's last tree (num_leaves)");
  }

  // score[row] += leaf value over the last tree's bag (every row when it had none),
  // or over the rows outside its bag (out_of_bag)
  void add_score_device(const bk_tree_out* t, double* d_score, bool out_of_bag = false) {
    check_tree_matches(t);
    if (t->num_leaves <= 1 || (out_of_bag && !last_bagged)) return;
    ensure_leaf_of();
    stage_wait();
    for (int i = 0; i < t->num_leaves; ++i) h_vals[i] = t->leaf_value[i];
    timed(8, 0.0, 0.0, [&] {
      launch_copy(d_leaf_vals, dev_alias(h_vals), sizeof(double) * (size_t)t->num_leaves, stream);
    });
    staged();
    // no synchronization: the next use of h_vals waits for this copy (stage_wait),
    // and the score is read by later work on the same stream
    timed(6, (double)N * 20.0, (double)N * 20.0, [&] {
      launch_add_score_rows(N, d_leaf_of, d_leaf_vals, t->num_leaves, out_of_bag ? 1 : 0, d_score, stream);
    });
  }

  void add_score_host(const bk_tree_out* t, double* score) {
    check_tree_matches(t);
    if (t->num_leaves <= 1) return;
    ensure_leaf_of();
    std::vector<int32_t> lo((size_t)N);
    BK_HIPC(hipMemcpyAsync(lo.data(), d_leaf_of, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, stream));
    sync();
    staged_pending = false;
    for (int i = 0; i < N; ++i)
      if (lo[(size_t)i] >= 0) score[i] += t->leaf_value[lo[(size_t)i]];
  }

  // add_out_of_bag_prediction (serial.rs:1995-2001): the last tree replayed on
  // `rows` with the training routing, which is the leaf the walk found for them
  void add_oob_host(const bk_tree_out* t, const int32_t* rows, int64_t n, double* score) {
    check_tree_matches(t);
    if (t->num_leaves <= 1 || n <= 0) return;
    for (int64_t j = 0; j < n; ++j)
      if (rows[j] < 0 || rows[j] >= N) throw std::invalid_argument("row outside the dataset");
    ensure_leaf_of();
    std::vector<int32_t> lo((size_t)N);
    BK_HIPC(hipMemcpyAsync(lo.data(), d_leaf_of, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, stream));
    sync();
    staged_pending = false;
    for (int64_t j = 0; j < n; ++j) {
      const int v = lo[(size_t)rows[j]];
      score[rows[j]] += t->leaf_value[v < 0 ? ~v : v];
    }
  }

  // last_splits (serial.rs:2003): the routing of every split of the last tree, by
  // internal node (the order the splits were made)
  int last_splits(bk_split_record* out, int cap) const {
    const int ni = std::max(last_num_leaves - 1, 0);
    if (out && cap < ni) throw std::invalid_argument("split record capacity too small");
    for (int k = 0; out && k < ni; ++k) {
      bk_split_record r{};
      r.group = tree_nodes[(size_t)k].group;
      std::memcpy(r.rule, &tree_nodes[(size_t)k].rule, sizeof(RuleDev));
      out[k] = r;
    }
    return ni;
  }

  // add_prediction_for_rows (serial.rs:2009-2040): any earlier tree of this
  // learner, replayed on `rows` through its split records with the training
  // routing; score[row] += that tree's leaf value. The walk covers every row of
  // the dataset on the device, into a map of its own.
  void add_prediction_for_rows(const bk_tree_out* t, const bk_split_record* splits, int nsplits, const int32_t* rows,
                               int64_t n, double* score) {
    const int nl = t->num_leaves;
    if (nl <= 1 || n <= 0) return;
    if (nsplits != nl - 1) throw std::invalid_argument("one split record per internal node is needed");
    if (nl > P.num_leaves) throw std::invalid_argument("tree larger than the learner's num_leaves");
    for (int64_t j = 0; j < n; ++j)
      if (rows[j] < 0 || rows[j] >= N) throw std::invalid_argument("row outside the dataset");
    const int ni = nl - 1;
    auto child_ok = [&](int c) { return c >= 0 ? c < ni : ~c < nl; };
    std::vector<NodeRule> rules((size_t)ni);
    for (int k = 0; k < ni; ++k) {
      if (splits[k].group < 0 || splits[k].group >= G) throw std::invalid_argument("split record outside the dataset");
      if (!child_ok(t->left_child[k]) || !child_ok(t->right_child[k]))
        throw std::invalid_argument("tree children out of range");
      NodeRule& r = rules[(size_t)k];
      std::memcpy(&r.rule, splits[k].rule, sizeof(RuleDev));
      r.group = splits[k].group;
      r.left = t->left_child[k];
      r.right = t->right_child[k];
      r.pad = 0;
    }
    if (!d_replay) d_replay = dmalloc<int32_t>((size_t)N);
    leaf_of_valid = false;  // the node buffers are shared with the last tree's map
    stage_wait();
    std::memcpy(h_nodes, rules.data(), sizeof(NodeRule) * (size_t)ni);
    timed(8, 0.0, 0.0,
          [&] { launch_copy(d_node_rules, dev_alias(h_nodes), sizeof(NodeRule) * (size_t)ni, stream); });
    launch_leaf_of_tree(ds, d_node_rules, d_nodes, ni, nullptr, d_replay, stream);
    std::vector<int32_t> lo((size_t)N);
    BK_HIPC(hipMemcpyAsync(lo.data(), d_replay, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, stream));
    sync();
    staged_pending = false;
    for (int64_t j = 0; j < n; ++j) score[rows[j]] += t->leaf_value[lo[(size_t)rows[j]]];
  }

  // Root histogram over all rows in the learner's form: what construct_histograms
  // (serial.rs:891-1065) or construct_int_histograms (serial.rs:1074-1164) leave
  // in the root's slot before fix_histogram, with a constant hessian scaled by
  // hessians[0] (serial.rs:982-988). It runs the per-tree set-up (discretization
  // included), so a stochastic learner's noise offset advances as in train.
  void root_histogram(const float* g, const float* h, double* out_gh, int64_t* out_int, int32_t* out_cnt) {
    if (quant ? out_gh != nullptr : out_int != nullptr)
      throw std::invalid_argument(quant ? "the quantized learner returns integer histograms"
                                        : "float learners return f64 histograms");
    use_hist32(false);  // the histogram with its row counts
    before_train(g, h);
    if (cols.per_tree) set_groups_used(true);  // every group, whatever the tree's sample
    hist_seg.assign(1, segs[0]);
    const HistPlan hp = plan_hist(1);
    h_ritems[0] = ReduceItem{0, hist_first[0], hist_count[0], 0};
    begin_uploads();
    d_hitems = put(h_hitems, (size_t)hp.ni);
    d_hblock = put(h_hblock, (size_t)hp.nblocks);
    d_ritems = put(h_ritems, 1);
    send(true);
    run_hist(hp, N, true);
    launch_reduce(ds, d_ritems, 1, stream);
    std::vector<double> gh((size_t)TB * 2);
    std::vector<int32_t> cnt((size_t)TB);
    BK_HIPC(hipMemcpyAsync(gh.data(), ds.pool, sizeof(double) * gh.size(), hipMemcpyDeviceToHost, stream));
    BK_HIPC(hipMemcpyAsync(cnt.data(), ds.pool_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost, stream));
    sync();
    staged_pending = false;
    if (quant) {
      if (out_int) std::memcpy(out_int, gh.data(), sizeof(int64_t) * gh.size());
    } else if (out_gh) {
      for (size_t i = 0; i < (size_t)TB; ++i) {
        out_gh[2 * i] = gh[2 * i];
        out_gh[2 * i + 1] = const_hess ? gh[2 * i + 1] * h0 : gh[2 * i + 1];
      }
    }
    if (out_cnt) std::memcpy(out_cnt, cnt.data(), sizeof(int32_t) * cnt.size());
  }

  // root_histogram on device arrays, allocating nothing: the discretization (or
  // the float payload) goes to buffer 0 as in train, the partial histograms to
  // the file-scope scratch array, and the reduction straight into the caller's
  // d_hist (2 x total_bins f64, or int64 when quantized) and d_count (total_bins
  // int32). No host containers either: the plan is one item over all rows. Waits
  // for the stream before returning. Calls that overlap on one device race on the
  // scratch.
  void root_histogram_noalloc(const float* g, const float* h, void* d_hist, int32_t* d_count) {
    if (!g || !h || !d_hist || !d_count) throw std::invalid_argument("null argument");
    use_hist32(false);  // the histogram with its row counts
    // partials of this tiling: at least one block per rows_per_flush rows (one
    // LDS accumulation per slot), at most what the scratch holds
    const size_t per_partial = (size_t)TB * (form == FORM_QUANT ? 8 : 20);
    const long long cap = (long long)(kScratchBytes / per_partial);
    const long long need = ((long long)N + rows_per_flush - 1) / rows_per_flush;
    if (need > cap || need > kHistMaxBlocks)
      throw std::invalid_argument("root_histogram_noalloc: " + std::to_string(need) + " partial histograms of " +
                                  std::to_string(per_partial) + " bytes exceed the " +
                                  std::to_string(kScratchBytes >> 20) + " MiB scratch");
    long long nb = std::min<long long>(hist_max_blocks, std::max<long long>(1, N / hist_min_rows));
    nb = std::max(std::min(nb, std::min<long long>(cap, kHistMaxBlocks)), need);
    const int rpb = (int)(((long long)N + nb - 1) / nb);
    const int nblocks = (int)(((long long)N + rpb - 1) / rpb);
    if (quant) {
      discretize(g, h);
    } else {
      launch_pack_f32(g, h, N, const_hess ? 1 : 0, (float*)d_gh0, stream);
    }
    set_groups_used(true);
    begin_uploads();
    h_hitems[0] = HistItem{Segment{0, 0, N}, 0, 0};
    for (int b = 0; b < nblocks; ++b) h_hblock[b] = b;
    h_ritems[0] = ReduceItem{0, 0, nblocks, 0};
    d_hitems = put(h_hitems, 1);
    d_hblock = put(h_hblock, (size_t)nblocks);
    d_ritems = put(h_ritems, 1);
    send(true);
    DevState s = ds;
    uint8_t* sc = (uint8_t*)scratch_device_ptr();
    const size_t plane = (size_t)nblocks * TB;
    if (quant) {
      s.part_q = (unsigned long long*)sc;
    } else {
      s.part_g = (double*)sc;
      s.part_h = s.part_g + plane;
      s.part_c = (unsigned*)(s.part_h + plane);
    }
    s.pool = (double*)d_hist;  // slot 0 of this view is the caller's output
    s.pool_cnt = d_count;
    hist_launch(s, d_hitems, 1, N, nblocks, rpb, d_hblock, nullptr);
    launch_reduce(s, d_ritems, 1, stream);
    if (!quant && const_hess) launch_scale_const_hess((double*)d_hist, TB, h, stream);
    sync();
    if (cols.per_tree) set_groups_used(false);
  }

  // Rows of a leaf in ascending order (the order of the reference's partition),
  // read off the row-to-leaf map.
  void leaf_rows(int leaf, int32_t* out, int64_t* n) {
    if (leaf < 0 || leaf >= last_num_leaves) throw std::invalid_argument("leaf out of range");
    ensure_leaf_of();
    std::vector<int32_t> lo((size_t)N);
    BK_HIPC(hipMemcpyAsync(lo.data(), d_leaf_of, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, stream));
    sync();
    int64_t k = 0;
    for (int r = 0; r < N; ++r)
      if (lo[(size_t)r] == leaf) out[k++] = r;
    if (k != segs[(size_t)leaf].count) throw std::runtime_error("leaf rows disagree with the leaf count");
    *n = k;
  }
};

}  // namespace bk

struct bk_learner {
  bk::Learner L;
};

extern "C" {

const char* bk_last_error(void) { return bk::last_error().c_str(); }

int bk_learner_create(const bk_dataset_view* ds, const bk_learner_params* params, int is_constant_hessian,
                      bk_learner** out) {
  return bk::guard([&] {
    if (!ds || !params || !out) throw std::invalid_argument("null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device available");
    auto h = std::make_unique<bk_learner>();
    h->L.create(ds, params, is_constant_hessian);
    *out = h.release();
  });
}

int bk_learner_train(bk_learner* l, const float* g, const float* h, const int32_t* rows, int64_t num_rows,
                     bk_tree_out* tree) {
  return bk::guard([&] {
    if (!l || !g || !h || !tree) throw std::invalid_argument("null argument");
    bk::Learner& L = l->L;
    if (rows) {
      if (num_rows < 0 || num_rows > L.N) throw std::invalid_argument("bag larger than the dataset");
      for (int64_t i = 0; i < num_rows; ++i)
        if (rows[i] < 0 || rows[i] >= L.N || (i > 0 && rows[i] <= rows[i - 1]))
          throw std::invalid_argument("bagged rows must be ascending and inside the dataset");
      L.ensure_bag_buffers();
      if (num_rows > 0)
        BK_HIPC(hipMemcpyAsync(L.d_bag, rows, sizeof(int32_t) * (size_t)num_rows, hipMemcpyHostToDevice, L.stream));
    }
    BK_HIPC(hipMemcpyAsync(L.d_grad, g, sizeof(float) * (size_t)L.N, hipMemcpyHostToDevice, L.stream));
    BK_HIPC(hipMemcpyAsync(L.d_hess, h, sizeof(float) * (size_t)L.N, hipMemcpyHostToDevice, L.stream));
    bk::Tree t = L.train(L.d_grad, L.d_hess, rows ? L.d_bag : nullptr, rows ? (int)num_rows : 0);
    L.export_tree(t, tree);
  });
}

int bk_learner_train_device(bk_learner* l, const float* dg, const float* dh, bk_tree_out* tree) {
  return bk::guard([&] {
    if (!l || !dg || !dh || !tree) throw std::invalid_argument("null argument");
    bk::Tree t = l->L.train(dg, dh);
    l->L.export_tree(t, tree);
  });
}

int bk_learner_add_prediction_to_score(bk_learner* l, const bk_tree_out* tree, double* score) {
  return bk::guard([&] { l->L.add_score_host(tree, score); });
}

int bk_learner_add_prediction_to_score_device(bk_learner* l, const bk_tree_out* tree, double* d_score) {
  return bk::guard([&] { l->L.add_score_device(tree, d_score); });
}

int bk_learner_train_device_bagged(bk_learner* l, const float* dg, const float* dh, const int32_t* d_rows,
                                   int64_t num_rows, bk_tree_out* tree) {
  return bk::guard([&] {
    if (!l || !dg || !dh || !d_rows || !tree) throw std::invalid_argument("null argument");
    bk::Learner& L = l->L;
    if (num_rows < 0 || num_rows > L.N) throw std::invalid_argument("bag larger than the dataset");
    L.ensure_bag_buffers();
    bk::Tree t = L.train(dg, dh, d_rows, (int)num_rows);
    L.export_tree(t, tree);
  });
}

int bk_learner_add_out_of_bag_prediction(bk_learner* l, const bk_tree_out* tree, const int32_t* rows,
                                         int64_t num_rows, double* score) {
  return bk::guard([&] {
    if (!l || !tree || (num_rows > 0 && (!rows || !score))) throw std::invalid_argument("null argument");
    l->L.add_oob_host(tree, rows, num_rows, score);
  });
}

int bk_learner_histogram_cache_capacity(bk_learner* l, int64_t* out) {
  return bk::guard([&] {
    if (!l || !out) throw std::invalid_argument("null argument");
    *out = (int64_t)l->L.cache;
  });
}

int bk_learner_reset_histogram_pool_config(bk_learner* l) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    l->L.cols.next_tree();
  });
}

int bk_learner_continue_after_dataset_reset(bk_learner* l, const bk_learner* previous) {
  return bk::guard([&] {
    if (!l || !previous) throw std::invalid_argument("null argument");
    l->L.cols.rnd = previous->L.cols.rnd;
    l->L.cols.next_tree();
  });
}

int bk_learner_last_splits(bk_learner* l, bk_split_record* out, int32_t capacity, int32_t* num_out) {
  return bk::guard([&] {
    if (!l || !num_out) throw std::invalid_argument("null argument");
    *num_out = l->L.last_splits(out, capacity);
  });
}

int bk_learner_add_prediction_for_rows(bk_learner* l, const bk_tree_out* tree, const bk_split_record* splits,
                                       int32_t num_splits, const int32_t* rows, int64_t num_rows, double* score) {
  return bk::guard([&] {
    if (!l || !tree || (num_rows > 0 && (!rows || !score)) || (num_splits > 0 && !splits))
      throw std::invalid_argument("null argument");
    l->L.add_prediction_for_rows(tree, splits, num_splits, rows, num_rows, score);
  });
}

int bk_learner_add_out_of_bag_prediction_device(bk_learner* l, const bk_tree_out* tree, double* d_score) {
  return bk::guard([&] {
    if (!l || !tree || !d_score) throw std::invalid_argument("null argument");
    l->L.add_score_device(tree, d_score, true);
  });
}

int bk_learner_leaf_rows(bk_learner* l, int32_t leaf, int32_t* rows_out, int64_t* num_out) {
  return bk::guard([&] {
    if (!l || !rows_out || !num_out) throw std::invalid_argument("null argument");
    l->L.leaf_rows(leaf, rows_out, num_out);
  });
}

int bk_learner_root_histogram(bk_learner* l, const float* g, const float* h, double* hist_out,
                              int64_t* int_hist_out, int32_t* count_out) {
  return bk::guard([&] {
    if (!l || !g || !h) throw std::invalid_argument("null argument");
    bk::Learner& L = l->L;
    BK_HIPC(hipMemcpyAsync(L.d_grad, g, sizeof(float) * (size_t)L.N, hipMemcpyHostToDevice, L.stream));
    BK_HIPC(hipMemcpyAsync(L.d_hess, h, sizeof(float) * (size_t)L.N, hipMemcpyHostToDevice, L.stream));
    L.root_histogram(L.d_grad, L.d_hess, hist_out, int_hist_out, count_out);
  });
}

int bk_learner_root_histogram_noalloc(bk_learner* l, const float* d_gradients, const float* d_hessians,
                                      void* d_hist_out, int32_t* d_count_out) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    l->L.root_histogram_noalloc(d_gradients, d_hessians, d_hist_out, d_count_out);
  });
}

int bk_learner_reset_constant_hessian(bk_learner* l, int is_constant_hessian) {
  return bk::guard([&] {
    if (l->L.quant && (is_constant_hessian != 0) != l->L.const_hess)
      throw std::invalid_argument("the quantized learner fixes the hessian mode at creation");
    l->L.const_hess = is_constant_hessian != 0;
  });
}

int bk_learner_binary_gradients_device(bk_learner* l, const double* d_score, const float* d_label,
                                       const float* d_weights, double sigmoid, const double label_weights[2],
                                       float* d_g, float* d_h) {
  return bk::guard([&] {
    if (!l || !d_score || !d_label || !label_weights || !d_g || !d_h) throw std::invalid_argument("null argument");
    if (!(sigmoid > 0.0)) throw std::invalid_argument("sigmoid must be greater than zero");
    bk::Learner& L = l->L;
    const double b = (double)L.N * (d_weights ? 24.0 : 20.0);
    L.timed(6, b, b, [&] {
      bk::launch_binary_gradients(d_score, d_label, d_weights, L.N, sigmoid, label_weights[0], label_weights[1], d_g,
                                  d_h, L.stream);
    });
  });
}

int bk_learner_softmax_gradients_device(bk_learner* l, const double* d_score, const float* d_label, int num_class,
                                        float* d_g, float* d_h) {
  return bk::guard([&] {
    if (num_class < 2) throw std::invalid_argument("softmax needs at least two classes");
    bk::Learner& L = l->L;
    L.timed(6, (double)L.N * num_class * 16.0 + 4.0 * L.N, (double)L.N * num_class * 16.0 + 4.0 * L.N,
            [&] { bk::launch_softmax_gradients(d_score, d_label, L.N, num_class, d_g, d_h, L.stream); });
  });
}

int bk_learner_lambdarank_gradients_device(bk_learner* l, const double* d_score, const float* d_label,
                                           const bk_rank_tables* t, float* d_g, float* d_h) {
  return bk::guard([&] {
    if (!t || !t->query_bounds || !t->inverse_max_dcg || !t->label_gain || !t->discount || !t->sigmoid_table)
      throw std::invalid_argument("null argument");
    if (t->max_query_rows > bk::kRankMaxDocs) throw std::invalid_argument("a query exceeds 2048 rows");
    if (t->table_bins < 1) throw std::invalid_argument("empty sigmoid table");
    bk::RankDev r{};
    r.bounds = t->query_bounds;
    r.inv_max_dcg = t->inverse_max_dcg;
    r.label_gain = t->label_gain;
    r.discount = t->discount;
    r.sigmoid_table = t->sigmoid_table;
    r.sigmoid = t->sigmoid;
    r.min_in = t->min_input;
    r.max_in = t->max_input;
    r.table_factor = t->table_factor;
    r.num_queries = t->num_queries;
    r.num_gain = t->num_gain;
    r.table_bins = t->table_bins;
    r.truncation = t->truncation_level;
    r.norm = t->norm;
    bk::Learner& L = l->L;
    L.timed(6, 0.0, 0.0, [&] { bk::launch_lambdarank_gradients(d_score, d_label, r, d_g, d_h, L.stream); });
  });
}

int bk_learner_set_tree_slot(bk_learner* l, int slot, int trees_per_iteration) {
  return bk::guard([&] {
    if (trees_per_iteration < 1 || slot < 0 || slot >= trees_per_iteration)
      throw std::invalid_argument("slot must lie in [0, trees_per_iteration)");
    if (trees_per_iteration > 1 && l->L.cols.by_node_draws())
      throw std::invalid_argument(
          "by-node feature sampling draws as many samples as the other trees' searches: the trees of one "
          "iteration need one learner");
    l->L.tree_slot = slot;
    l->L.tree_period = trees_per_iteration;
  });
}

int bk_learner_set_round_planner(bk_learner* l, int on_device) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    l->L.plan_on_device = on_device != 0;
    l->L.resident_on = on_device == 1;
  });
}

int bk_learner_set_hist_tiling(bk_learner* l, int32_t min_rows, int32_t max_blocks) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    bk::Learner& L = l->L;
    if (min_rows == 0 && max_blocks == 0) {
      L.hist_min_rows = L.tuned_min_rows;
      L.hist_max_blocks = L.tuned_max_blocks;
      return;
    }
    if (min_rows < 2048 || max_blocks < 1 || max_blocks > bk::Learner::kHistMaxBlocks)
      throw std::invalid_argument("histogram tiling: min_rows >= 2048 and 1 <= max_blocks <= 2048");
    L.hist_min_rows = min_rows;
    L.hist_max_blocks = max_blocks;
  });
}

int bk_learner_set_hist_layout(bk_learner* l, int32_t groups_per_tile, int32_t sub_histograms) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    bk::Learner& L = l->L;
    if (groups_per_tile == 0 && sub_histograms == 0) {
      L.build_tiles(L.tuned_width, L.tuned_nsub);
      return;
    }
    if (groups_per_tile < 16 || groups_per_tile % 16 != 0 || sub_histograms < 1 ||
        sub_histograms > bk::kMaxSubHist)
      throw std::invalid_argument("histogram layout: groups_per_tile a positive multiple of 16, 1 or 2 sub-histograms");
    L.build_tiles(groups_per_tile, sub_histograms);
  });
}

int bk_learner_set_partition_histograms(bk_learner* l, int32_t on) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    l->L.part_hist_on = on != 0;
  });
}

int bk_learner_set_hist_bin_width(bk_learner* l, int32_t bytes) {
  return bk::guard([&] {
    if (!l) throw std::invalid_argument("null argument");
    if (bytes != 0 && bytes != 4 && bytes != 8) throw std::invalid_argument("histogram bin width: 0, 4 or 8 bytes");
    if (bytes == 4 && !l->L.hist32_ok)
      throw std::invalid_argument("32-bit histogram bins need the quantized batched form, num_grad_quant_bins <= 13 "
                                  "and num_leaves <= 1023");
    l->L.hist_width = bytes;
  });
}

int bk_learner_last_stats(bk_learner* l, bk_train_stats* out) {
  return bk::guard([&] {
    if (!l || !out) throw std::invalid_argument("null argument");
    *out = l->L.stats;
    out->hist_min_rows = l->L.hist_min_rows;
    out->hist_max_blocks = l->L.hist_max_blocks;
    out->hist_tile_groups = l->L.tile_width;
    out->hist_ntiles = (int32_t)l->L.tile_first.size() - 1;
    out->hist_sub_histograms = l->L.hist_nsub;
    out->partition_histograms = l->L.ds.hist32 && l->L.part_hist_ok && l->L.part_hist_on;
  });
}

int bk_learner_profile(bk_learner* l, int kinds) {
  return bk::guard([&] {
    l->L.prof.mask = (unsigned)kinds;
    l->L.prof.on = kinds != 0;
    l->L.prof.reset();
    l->L.prof.trees = 0;
    l->L.prof.tree_on = true;
    // create the events now: creating them while launching would add host time
    if (kinds != 0) l->L.prof.ev(bk::Prof::kResolveAt + 512);
  });
}

int bk_learner_profile_trees(bk_learner* l, int every) {
  return bk::guard([&] {
    if (every < 1) throw std::invalid_argument("every must be at least 1");
    l->L.prof.every = every;
    l->L.prof.trees = 0;
  });
}

int bk_learner_kernel_stats(bk_learner* l, int kind, bk_kernel_stats* out) {
  return bk::guard([&] {
    if (kind < 0 || kind >= bk::Prof::kKinds) throw std::invalid_argument("kernel kind out of range");
    if (!l->L.prof.pending.empty()) {
      l->L.sync();
      l->L.prof.resolve();
    }
    out->ms = l->L.prof.ms[kind];
    out->launches = l->L.prof.launches[kind];
    out->algo_bytes = l->L.prof.algo[kind];
    out->moved_bytes = l->L.prof.moved[kind];
  });
}

int bk_learner_free(bk_learner* l) {
  delete l;
  return 0;
}

}  // extern "C"
