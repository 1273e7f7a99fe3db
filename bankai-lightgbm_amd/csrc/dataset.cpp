// dataset.cpp — host-side dataset construction: row sampling, bin finding,
// feature-group layout and the column-major u8 bin matrix the GPU learner
// keeps resident in HBM.
//
// Behaviour follows the reference's in-memory dense path:
//   sampling            crates/lgbm-core/src/io/sample.rs:15-98 (CreateSampleIndices)
//   Random LCG          crates/lgbm-core/src/utils/random.rs:6-80
//   bin finding         crates/lgbm-core/src/io/bin.rs:50-553 (GreedyFindBin,
//                       FindBinWithZeroAsOneBin, NeedFilter, BinMapper::find_bin)
//   value_to_bin        crates/lgbm-core/src/io/bin.rs:558-589
//   layout / groups     crates/lgbm-core/src/io/dataset.rs:19-76,365-447
//   feature bundling    crates/lgbm-core/src/io/dataset.rs:95-363 (FixSampleIndices,
//                       FindGroups, FastFeatureBundling)
//   fill_dense          crates/lgbm-core/src/io/dataset.rs:633-689
// Every stored group is one u8 column. A group of the reference's layout that does
// not fit (a bundle of more than 256 bins, kept in u16 bins by the reference, or a
// multi-value group) is stored exploded: one column per sub-feature, holding the
// values the reference's group holds, in the reference's inner feature order.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bankai_dataset.h"
#include "bk_common.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kZeroThreshold = (double)1e-35f;
const double kSparseThreshold = 0.7;

// Random LCG (utils/random.rs:6-80).
struct Lcg {
  uint32_t x;
  explicit Lcg(int32_t seed) : x((uint32_t)seed) {}
  int32_t rand16() {
    x = x * 214013u + 2531011u;
    return (int32_t)((x >> 16) & 0x7fff);
  }
  int32_t rand32() {
    x = x * 214013u + 2531011u;
    return (int32_t)(x & 0x7fffffff);
  }
  int32_t next_short(int32_t lo, int32_t hi) { return rand16() % (hi - lo) + lo; }
  int32_t next_int(int32_t lo, int32_t hi) { return rand32() % (hi - lo) + lo; }
  float next_float() { return (float)rand16() / 32768.0f; }
  std::vector<int32_t> sample(int32_t n, int32_t k) {
    std::vector<int32_t> ret;
    if (k > n || k <= 0) return ret;
    if (k == n) {
      for (int32_t i = 0; i < n; ++i) ret.push_back(i);
      return ret;
    }
    if (k > 1 && (double)k > (double)n / std::log2((double)k)) {
      for (int32_t i = 0; i < n; ++i) {
        double prob = (double)((size_t)k - ret.size()) / (double)(n - i);
        if ((double)next_float() < prob) ret.push_back(i);
      }
    } else {
      std::set<int32_t> s;
      for (int32_t r = n - k; r < n; ++r) {
        int32_t v = next_int(0, r + 1);
        if (!s.insert(v).second) s.insert(r);
      }
      ret.assign(s.begin(), s.end());
    }
    return ret;
  }
};

inline double next_up(double a) { return std::nextafter(a, std::numeric_limits<double>::infinity()); }
inline bool check_equal_ordered(double a, double b) { return b <= next_up(a); }

// GreedyFindBin (bin.rs:51-131).
std::vector<double> greedy_find_bin(const double* dv, const int32_t* counts, size_t n, int max_bin, size_t total_cnt,
                                    int min_data_in_bin) {
  std::vector<double> ub;
  if ((int)n <= max_bin) {
    int cur = 0;
    for (size_t i = 0; i + 1 < n; ++i) {
      cur += counts[i];
      if (cur >= min_data_in_bin) {
        double val = next_up((dv[i] + dv[i + 1]) / 2.0);
        if (ub.empty() || !check_equal_ordered(ub.back(), val)) {
          ub.push_back(val);
          cur = 0;
        }
      }
    }
    ub.push_back(std::numeric_limits<double>::infinity());
  } else {
    if (min_data_in_bin > 0) {
      max_bin = std::min(max_bin, (int)(total_cnt / (size_t)min_data_in_bin));
      max_bin = std::max(max_bin, 1);
    }
    double mean_bin_size = (double)total_cnt / (double)max_bin;
    int rest_bin_cnt = max_bin;
    int rest_sample_cnt = (int)total_cnt;
    std::vector<char> big(n, 0);
    for (size_t i = 0; i < n; ++i) {
      if ((double)counts[i] >= mean_bin_size) {
        big[i] = 1;
        rest_bin_cnt -= 1;
        rest_sample_cnt -= counts[i];
      }
    }
    mean_bin_size = (double)rest_sample_cnt / (double)rest_bin_cnt;
    size_t mb = (size_t)max_bin;
    std::vector<double> upper(mb, std::numeric_limits<double>::infinity()), lower(mb, std::numeric_limits<double>::infinity());
    size_t bin_cnt = 0;
    lower[0] = dv[0];
    int cur = 0;
    for (size_t i = 0; i + 1 < n; ++i) {
      if (!big[i]) rest_sample_cnt -= counts[i];
      cur += counts[i];
      if (big[i] || (double)cur >= mean_bin_size || (big[i + 1] && (double)cur >= std::max(1.0, mean_bin_size * 0.5))) {
        upper[bin_cnt] = dv[i];
        bin_cnt += 1;
        lower[bin_cnt] = dv[i + 1];
        if ((int)bin_cnt >= max_bin - 1) break;
        cur = 0;
        if (!big[i]) {
          rest_bin_cnt -= 1;
          mean_bin_size = (double)rest_sample_cnt / (double)rest_bin_cnt;
        }
      }
    }
    bin_cnt += 1;
    for (size_t i = 0; i + 1 < bin_cnt; ++i) {
      double val = next_up((upper[i] + lower[i + 1]) / 2.0);
      if (ub.empty() || !check_equal_ordered(ub.back(), val)) ub.push_back(val);
    }
    ub.push_back(std::numeric_limits<double>::infinity());
  }
  return ub;
}

// FindBinWithZeroAsOneBin (bin.rs:221-281), no forced bounds.
std::vector<double> find_bin_zero_as_one(const std::vector<double>& dv, const std::vector<int32_t>& counts, int max_bin,
                                         size_t total_sample_cnt, int min_data_in_bin) {
  size_t n = dv.size();
  std::vector<double> ub;
  int left_cnt_data = 0, cnt_zero = 0, right_cnt_data = 0;
  for (size_t i = 0; i < n; ++i) {
    if (dv[i] <= -kZeroThreshold) left_cnt_data += counts[i];
    else if (dv[i] > kZeroThreshold) right_cnt_data += counts[i];
    else cnt_zero += counts[i];
  }
  size_t left_cnt = n;
  for (size_t i = 0; i < n; ++i)
    if (dv[i] > -kZeroThreshold) {
      left_cnt = i;
      break;
    }
  if (left_cnt > 0 && max_bin > 1) {
    double denom = (double)total_sample_cnt - (double)cnt_zero;
    int left_max_bin = (int)((double)left_cnt_data / denom * (double)(max_bin - 1));
    left_max_bin = std::max(left_max_bin, 1);
    ub = greedy_find_bin(dv.data(), counts.data(), left_cnt, left_max_bin, (size_t)left_cnt_data, min_data_in_bin);
    if (!ub.empty()) ub.back() = -kZeroThreshold;
  }
  long right_start = -1;
  for (size_t i = left_cnt; i < n; ++i)
    if (dv[i] > kZeroThreshold) {
      right_start = (long)i;
      break;
    }
  int right_max_bin = max_bin - 1 - (int)ub.size();
  if (right_start >= 0 && right_max_bin > 0) {
    std::vector<double> r = greedy_find_bin(dv.data() + right_start, counts.data() + right_start, n - (size_t)right_start,
                                            right_max_bin, (size_t)right_cnt_data, min_data_in_bin);
    ub.push_back(kZeroThreshold);
    ub.insert(ub.end(), r.begin(), r.end());
  } else {
    ub.push_back(std::numeric_limits<double>::infinity());
  }
  return ub;
}

struct BinMapper {
  int num_bin = 1;
  int missing_type = BK_MISSING_NONE;
  bool is_trivial = true;
  double sparse_rate = 1.0;
  int bin_type = BK_BIN_NUMERICAL;
  double min_val = 0.0, max_val = 0.0;
  uint32_t default_bin = 0, most_freq_bin = 0;
  std::vector<double> upper;
  std::vector<int32_t> bin2cat;
  std::map<int32_t, uint32_t> cat2bin;

  // value_to_bin (bin.rs:558-589)
  uint32_t value_to_bin(double v) const {
    if (std::isnan(v)) {
      if (bin_type == BK_BIN_CATEGORICAL) return 0;
      if (missing_type == BK_MISSING_NAN) return (uint32_t)(num_bin - 1);
      v = 0.0;
    }
    if (bin_type == BK_BIN_NUMERICAL) {
      int l = 0, r = num_bin - 1;
      if (missing_type == BK_MISSING_NAN) r -= 1;
      while (l < r) {
        int mid = (r + l - 1) / 2;
        if (v <= upper[(size_t)mid]) r = mid;
        else l = mid + 1;
      }
      return (uint32_t)l;
    }
    int32_t iv = bk::sat_i32(v);
    if (iv < 0) return 0;
    auto it = cat2bin.find(iv);
    return it == cat2bin.end() ? 0u : it->second;
  }
};

// NeedFilter (bin.rs:284-305)
bool need_filter(const std::vector<int32_t>& cib, int total, int filter_cnt, int bin_type) {
  if (bin_type == BK_BIN_NUMERICAL) {
    int sum_left = 0;
    for (size_t i = 0; i + 1 < cib.size(); ++i) {
      sum_left += cib[i];
      if (sum_left >= filter_cnt && total - sum_left >= filter_cnt) return false;
    }
  } else if (cib.size() <= 2) {
    for (size_t i = 0; i + 1 < cib.size(); ++i)
      if (cib[i] >= filter_cnt && total - cib[i] >= filter_cnt) return false;
  } else {
    return false;
  }
  return true;
}

struct FindBinParams {
  size_t total_sample_cnt;
  int max_bin, min_data_in_bin, min_split_data;
  bool pre_filter;
  int bin_type;
  bool use_missing, zero_as_missing;
};

// BinMapper::find_bin (bin.rs:342-553), `values` are the sampled non-zero (or NaN) values.
BinMapper find_bin(std::vector<double>& values, const FindBinParams& p) {
  BinMapper m;
  size_t num_sample_values = values.size();
  size_t non_na = 0;
  for (size_t i = 0; i < num_sample_values; ++i)
    if (!std::isnan(values[i])) values[non_na++] = values[i];
  size_t na_cnt = 0;
  if (!p.use_missing) m.missing_type = BK_MISSING_NONE;
  else if (p.zero_as_missing) m.missing_type = BK_MISSING_ZERO;
  else if (non_na == num_sample_values) m.missing_type = BK_MISSING_NONE;
  else {
    na_cnt = num_sample_values - non_na;
    m.missing_type = BK_MISSING_NAN;
  }
  num_sample_values = non_na;
  m.bin_type = p.bin_type;
  m.default_bin = 0;
  int zero_cnt = (int)(p.total_sample_cnt - num_sample_values - na_cnt);
  std::stable_sort(values.begin(), values.begin() + (long)num_sample_values);
  std::vector<double> dv;
  std::vector<int32_t> counts;
  if (num_sample_values == 0 || (values[0] > 0.0 && zero_cnt > 0)) {
    dv.push_back(0.0);
    counts.push_back(zero_cnt);
  }
  if (num_sample_values > 0) {
    dv.push_back(values[0]);
    counts.push_back(1);
  }
  for (size_t i = 1; i < num_sample_values; ++i) {
    if (!check_equal_ordered(values[i - 1], values[i])) {
      if (values[i - 1] < 0.0 && values[i] > 0.0) {
        dv.push_back(0.0);
        counts.push_back(zero_cnt);
      }
      dv.push_back(values[i]);
      counts.push_back(1);
    } else {
      dv.back() = values[i];
      counts.back() += 1;
    }
  }
  if (num_sample_values > 0 && values[num_sample_values - 1] < 0.0 && zero_cnt > 0) {
    dv.push_back(0.0);
    counts.push_back(zero_cnt);
  }
  m.min_val = dv.front();
  m.max_val = dv.back();
  std::vector<int32_t> cib;
  if (p.bin_type == BK_BIN_NUMERICAL) {
    if (m.missing_type == BK_MISSING_ZERO) {
      m.upper = find_bin_zero_as_one(dv, counts, p.max_bin, p.total_sample_cnt, p.min_data_in_bin);
      if (m.upper.size() == 2) m.missing_type = BK_MISSING_NONE;
    } else if (m.missing_type == BK_MISSING_NONE) {
      m.upper = find_bin_zero_as_one(dv, counts, p.max_bin, p.total_sample_cnt, p.min_data_in_bin);
    } else {
      m.upper = find_bin_zero_as_one(dv, counts, p.max_bin - 1, p.total_sample_cnt - na_cnt, p.min_data_in_bin);
      m.upper.push_back(2.0);  // the stored NaN-bin bound is the enum value 2 (bin.rs:434-438)
    }
    m.num_bin = (int)m.upper.size();
    cib.assign((size_t)m.num_bin, 0);
    size_t ib = 0;
    for (size_t i = 0; i < dv.size(); ++i) {
      while (dv[i] > m.upper[ib] && (int)ib < m.num_bin - 1) ++ib;
      cib[ib] += counts[i];
    }
    if (m.missing_type == BK_MISSING_NAN) cib.back() = (int32_t)na_cnt;
  } else {
    std::vector<int32_t> dvi, ci;
    for (size_t i = 0; i < dv.size(); ++i) {
      int32_t val = bk::sat_i32(dv[i]);
      if (val < 0) {
        na_cnt += (size_t)counts[i];
      } else if (dvi.empty() || dvi.back() != val) {
        dvi.push_back(val);
        ci.push_back(counts[i]);
      } else {
        ci.back() += counts[i];
      }
    }
    long rest_cnt = (long)p.total_sample_cnt - (long)na_cnt;
    if (rest_cnt > 0) {
      std::vector<std::pair<int32_t, int32_t>> pairs;
      for (size_t i = 0; i < ci.size(); ++i) pairs.push_back({ci[i], dvi[i]});
      std::stable_sort(pairs.begin(), pairs.end(), [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) { return a.first > b.first; });
      for (size_t i = 0; i < pairs.size(); ++i) {
        ci[i] = pairs[i].first;
        dvi[i] = pairs[i].second;
      }
      float prod = (float)(p.total_sample_cnt - na_cnt) * 0.99f;
      int cut_cnt = bk::sat_i32((double)prod + 0.5);
      size_t cur = 0;
      m.cat2bin.clear();
      m.bin2cat.clear();
      int used_cnt = 0;
      int distinct_cnt = (int)dvi.size();
      if (na_cnt > 0) distinct_cnt += 1;
      int max_bin = std::min(distinct_cnt, p.max_bin);
      m.bin2cat.push_back(-1);
      m.cat2bin[-1] = 0;
      cib.push_back(0);
      m.num_bin = 1;
      while (cur < dvi.size() && (used_cnt < cut_cnt || m.num_bin < max_bin)) {
        if (ci[cur] < p.min_data_in_bin && cur > 1) break;
        m.bin2cat.push_back(dvi[cur]);
        m.cat2bin[dvi[cur]] = (uint32_t)m.num_bin;
        used_cnt += ci[cur];
        cib.push_back(ci[cur]);
        m.num_bin += 1;
        cur += 1;
      }
      m.missing_type = (cur == dvi.size() && na_cnt == 0) ? BK_MISSING_NONE : BK_MISSING_NAN;
      cib[0] = (int32_t)((long)p.total_sample_cnt - (long)used_cnt);
    }
  }
  m.is_trivial = m.num_bin <= 1;
  if (!m.is_trivial && p.pre_filter && need_filter(cib, (int)p.total_sample_cnt, p.min_split_data, p.bin_type))
    m.is_trivial = true;
  if (!m.is_trivial) {
    m.default_bin = m.value_to_bin(0.0);
    size_t am = 0;
    for (size_t i = 1; i < cib.size(); ++i)
      if (cib[i] > cib[am]) am = i;
    m.most_freq_bin = (uint32_t)am;
    double max_sparse_rate = (double)cib[m.most_freq_bin] / (double)p.total_sample_cnt;
    if (m.most_freq_bin != m.default_bin && max_sparse_rate < kSparseThreshold) m.most_freq_bin = m.default_bin;
    m.sparse_rate = (double)cib[m.most_freq_bin] / (double)p.total_sample_cnt;
  } else {
    m.sparse_rate = 1.0;
  }
  return m;
}

std::string fmt_g17(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

}  // namespace

struct bk_dataset {
  int64_t num_data = 0;
  int32_t num_total_features = 0;
  std::vector<BinMapper> mappers;          // per real column
  std::vector<char> used;                  // per real column
  // layout: the real columns of each group in sub-feature order (bundled and
  // shuffled as upstream does); inner features are numbered group by group
  std::vector<std::vector<int32_t>> group_feats;
  // the groups of the reference's layout and how each is stored: as one group, or
  // exploded into one stored group per sub-feature (a multi-value group, or a
  // bundle of more than 256 bins)
  std::vector<std::vector<int32_t>> plan_groups;
  std::vector<char> plan_multi, plan_exploded;
  std::vector<std::vector<uint8_t>> cols;  // per stored group
  std::vector<const uint8_t*> col_ptrs;
  std::vector<int32_t> g_num_total_bin, g_feature_start, g_feature_cnt, g_bin_offsets;
  std::vector<int32_t> f_num_bin, f_missing, f_default, f_mfb, f_type, f_real;
  std::vector<double> f_penalty;
  std::vector<const double*> f_upper;
  std::vector<const int32_t*> f_bin2cat;
  std::vector<std::string> infos;
};

namespace {

// Group columns (fill_dense, dataset.rs:633-689): a row's value of each
// sub-feature, pushed in real-column order, that is not the most frequent bin is
// stored at its sub-feature's offset; within a single-value bundle a later column
// overwrites, so a row holds one sub-feature's value and every other sub-feature
// reads as its most frequent bin. An exploded group keeps exactly that: a
// multi-value group stores every sub-feature's value in its own column; a bundle
// stores, in the column of the last sub-feature the row pushed, that value, and
// bin 0 (the most frequent bin) in the others. Each feature's histogram and each
// split's routing are then those of the reference's group.
void fill_columns(bk_dataset* ds, const double* rowmajor_f64, const float* rowmajor_f32, int64_t nrow, int32_t ncol,
                  int nthreads) {
  const size_t ng = ds->group_feats.size(), np = ds->plan_groups.size();
  ds->cols.assign(ng, std::vector<uint8_t>((size_t)nrow, 0));
  std::vector<size_t> first(np + 1, 0);  // first stored group of each plan group
  for (size_t q = 0; q < np; ++q) first[q + 1] = first[q] + (ds->plan_exploded[q] ? ds->plan_groups[q].size() : 1);
  if (first[np] != ng) throw std::logic_error("stored groups disagree with the plan");
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (long q = 0; q < (long)np; ++q) {
    const std::vector<int32_t>& feats = ds->plan_groups[(size_t)q];
    const bool exploded = ds->plan_exploded[(size_t)q] != 0, multi = ds->plan_multi[(size_t)q] != 0;
    std::vector<size_t> order(feats.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return feats[a] < feats[b]; });
    // stored column and bin offset of each sub-feature
    std::vector<uint8_t*> col(feats.size());
    std::vector<int32_t> off(feats.size());
    for (size_t j = 0; j < feats.size(); ++j) {
      const size_t g = first[(size_t)q] + (exploded ? j : 0);
      col[j] = ds->cols[g].data();
      off[j] = ds->g_bin_offsets[(size_t)ds->g_feature_start[g] + g + (exploded ? 0 : j)];
    }
    for (int64_t r = 0; r < nrow; ++r) {
      for (size_t sub : order) {
        const int32_t real = feats[sub];
        const BinMapper& m = ds->mappers[(size_t)real];
        double v = rowmajor_f64 ? rowmajor_f64[r * ncol + real] : (double)rowmajor_f32[r * ncol + real];
        uint32_t bin = m.value_to_bin(v);
        if (bin == m.most_freq_bin) continue;
        if (m.most_freq_bin == 0) bin -= 1;
        if (exploded && !multi)
          for (size_t j = 0; j < feats.size(); ++j) col[j][r] = 0;  // an earlier push is overwritten
        col[sub][r] = (uint8_t)(bin + (uint32_t)off[sub]);
      }
    }
  }
  ds->col_ptrs.resize(ng);
  for (size_t g = 0; g < ng; ++g) ds->col_ptrs[g] = ds->cols[g].data();
}

// Groups and inner features (Dataset::Construct, dataset.rs:378-447 with
// FeatureGroupLayout::new, dataset.rs:35-76, single-value groups): bin 0 of a group
// is shared by its sub-features' most frequent bins; each sub-feature's other bins
// follow at its offset.
void build_layout(bk_dataset* ds, std::vector<std::vector<int32_t>> groups) {
  ds->group_feats = std::move(groups);
  const size_t ng = ds->group_feats.size();
  size_t nf = 0;
  for (const auto& g : ds->group_feats) nf += g.size();
  ds->g_num_total_bin.resize(ng);
  ds->g_feature_start.resize(ng);
  ds->g_feature_cnt.resize(ng);
  ds->g_bin_offsets.clear();
  ds->f_num_bin.resize(nf);
  ds->f_missing.resize(nf);
  ds->f_default.resize(nf);
  ds->f_mfb.resize(nf);
  ds->f_type.resize(nf);
  ds->f_real.resize(nf);
  ds->f_penalty.assign(nf, 1.0);
  ds->f_upper.assign(nf, nullptr);
  ds->f_bin2cat.assign(nf, nullptr);
  size_t f = 0;
  for (size_t g = 0; g < ng; ++g) {
    ds->g_feature_start[g] = (int32_t)f;
    ds->g_feature_cnt[g] = (int32_t)ds->group_feats[g].size();
    int32_t total = 1;
    ds->g_bin_offsets.push_back(total);
    for (int32_t real : ds->group_feats[g]) {
      const BinMapper& m = ds->mappers[(size_t)real];
      total += m.num_bin - (m.most_freq_bin == 0 ? 1 : 0);
      ds->g_bin_offsets.push_back(total);
      ds->f_num_bin[f] = m.num_bin;
      ds->f_missing[f] = m.missing_type;
      ds->f_default[f] = (int32_t)m.default_bin;
      ds->f_mfb[f] = (int32_t)m.most_freq_bin;
      ds->f_type[f] = m.bin_type;
      ds->f_real[f] = real;
      if (m.bin_type == BK_BIN_NUMERICAL) ds->f_upper[f] = m.upper.data();
      else ds->f_bin2cat[f] = m.bin2cat.data();
      ++f;
    }
    ds->g_num_total_bin[g] = total;
    if (total > 256)
      throw std::invalid_argument("a feature of more than 256 bins cannot be stored one byte per row");
  }
  ds->infos.assign((size_t)ds->num_total_features, "none");
  for (int32_t i = 0; i < ds->num_total_features; ++i) {
    if (!ds->used[(size_t)i]) continue;
    const BinMapper& m = ds->mappers[(size_t)i];
    if (m.bin_type == BK_BIN_CATEGORICAL) {
      std::string s;
      for (size_t k = 0; k < m.bin2cat.size(); ++k) {
        if (k) s += ":";
        s += std::to_string(m.bin2cat[k]);
      }
      ds->infos[(size_t)i] = s;
    } else {
      ds->infos[(size_t)i] = "[" + fmt_g17(m.min_val) + ":" + fmt_g17(m.max_val) + "]";
    }
  }
}

// The sample a bundling decision reads: per column, the sample positions of its
// non-zero values and the values (SampleSet, io/sample.rs).
struct SampleView {
  const std::vector<int32_t>* rows;
  const std::vector<std::vector<double>>* values;
  const std::vector<std::vector<int32_t>>* indices;
  int32_t total;     // sampled rows
  int32_t num_data;  // rows of the dataset (seeds the bundling LCG)
};

// FixSampleIndices (dataset.rs:116-143): when the most frequent bin is not the
// zero bin, the positions whose bin differs from it (zeros included); empty
// otherwise (the non-zero positions then stand).
std::vector<int32_t> fix_sample_indices(const BinMapper& m, int32_t total, const std::vector<int32_t>& ix,
                                        const std::vector<double>& vals) {
  std::vector<int32_t> ret;
  if (m.default_bin == m.most_freq_bin) return ret;
  int32_t i = 0;
  size_t j = 0;
  while (i < total) {
    if (j < ix.size() && ix[j] < i) {
      ++j;
    } else if (j < ix.size() && ix[j] == i) {
      if (m.value_to_bin(vals[j]) != m.most_freq_bin) ret.push_back(i);
      ++i;
    } else {
      ret.push_back(i);
      ++i;
    }
  }
  return ret;
}

// Conflicting positions of `ix` already marked, or -1 once above max_cnt.
int32_t conflict_count(const std::vector<char>& mark, const std::vector<int32_t>& ix, int32_t max_cnt) {
  int32_t ret = 0;
  for (int32_t i : ix) {
    if (mark[(size_t)i]) ++ret;
    if (ret > max_cnt) return -1;
  }
  return ret;
}

// FindGroups (dataset.rs:146-279): greedy exclusive bundling in `order`. A feature
// joins the first of (the newest group, then up to 99 others drawn at random) whose
// marked sample positions it conflicts with at most the group's remaining budget
// (total / 10000 overall) and at most half its own non-zero count; otherwise it
// opens a group. With gpu_limit a group stays within 256 bins. With sparse
// groups enabled, groups used on under 40% of the sample are regrouped into one
// last group, multi-valued when its features conflict (always, with gpu_limit).
void find_groups(const std::vector<BinMapper>& mappers, const std::vector<int32_t>& order,
                 const std::vector<const std::vector<int32_t>*>& ix, const std::vector<int32_t>& num_per_col,
                 int32_t total, int32_t num_data, bool gpu_limit, bool sparse,
                 std::vector<std::vector<int32_t>>& groups_out, std::vector<char>& multi_out) {
  const int32_t kMaxSearchGroup = 100, kMaxBinPerGroup = 256;
  const int32_t max_conflict = total / 10000;
  Lcg rand(num_data);
  std::vector<std::vector<int32_t>> groups;
  std::vector<std::vector<char>> marks;
  std::vector<int32_t> used_rows, total_cnt, num_bins;
  for (int32_t f : order) {
    const BinMapper& m = mappers[(size_t)f];
    const int32_t nz = num_per_col[(size_t)f];
    const int32_t mfb_zero = m.most_freq_bin == 0 ? 1 : 0;
    std::vector<int32_t> avail;
    for (size_t g = 0; g < groups.size(); ++g) {
      const int32_t cur_bins = num_bins[g] + m.num_bin - mfb_zero;
      if (total_cnt[g] + nz <= total + max_conflict && (!gpu_limit || cur_bins <= kMaxBinPerGroup))
        avail.push_back((int32_t)g);
    }
    std::vector<int32_t> search;
    if (!avail.empty()) {
      const int32_t last = (int32_t)avail.size() - 1;
      search.push_back(avail.back());
      for (int32_t i : rand.sample(last, std::min(last, kMaxSearchGroup - 1))) search.push_back(avail[(size_t)i]);
    }
    int32_t best = -1, best_cnt = -1;
    for (int32_t g : search) {
      const int32_t rest = max_conflict - total_cnt[(size_t)g] + used_rows[(size_t)g];
      const int32_t c = conflict_count(marks[(size_t)g], *ix[(size_t)f], rest);
      if (c >= 0 && c <= rest && c <= nz / 2) {
        best = g;
        best_cnt = c;
        break;
      }
    }
    if (best >= 0) {
      groups[(size_t)best].push_back(f);
      total_cnt[(size_t)best] += nz;
      used_rows[(size_t)best] += nz - best_cnt;
      for (int32_t i : *ix[(size_t)f]) marks[(size_t)best][(size_t)i] = 1;
      // the reference adds here with the default bin, elsewhere with the most frequent one
      num_bins[(size_t)best] += m.num_bin - (m.default_bin == 0 ? 1 : 0);
    } else {
      groups.push_back({f});
      marks.emplace_back((size_t)total, (char)0);
      for (int32_t i : *ix[(size_t)f]) marks.back()[(size_t)i] = 1;
      total_cnt.push_back(nz);
      used_rows.push_back(nz);
      num_bins.push_back(1 + m.num_bin - mfb_zero);
    }
  }
  groups_out.clear();
  multi_out.clear();
  if (!sparse) {
    groups_out = std::move(groups);
    multi_out.assign(groups_out.size(), 0);
    return;
  }
  std::vector<int32_t> second;
  for (size_t g = 0; g < groups.size(); ++g) {
    if ((double)used_rows[g] / (double)total >= 0.4) {
      groups_out.push_back(groups[g]);
      multi_out.push_back(0);
    } else {
      second.insert(second.end(), groups[g].begin(), groups[g].end());
    }
  }
  if (second.empty()) return;
  std::vector<char> mark((size_t)total, 0);
  bool is_multi = gpu_limit;
  int32_t conflicts = 0;
  std::vector<int32_t> group;
  for (int32_t f : second) {
    group.push_back(f);
    if (is_multi) continue;
    const int32_t c = conflict_count(mark, *ix[(size_t)f], max_conflict - conflicts);
    conflicts += c;
    if (c < 0 || conflicts > max_conflict) {
      is_multi = true;
      continue;
    }
    for (int32_t i : *ix[(size_t)f]) mark[(size_t)i] = 1;
  }
  groups_out.push_back(group);
  multi_out.push_back(is_multi ? 1 : 0);
}

// FastFeatureBundling (dataset.rs:282-363): bundles in column order and in
// decreasing non-zero count, keeps the fewer groups (the first on a tie), then
// shuffles the groups with the LCG seeded by num_data.
void fast_feature_bundling(const std::vector<BinMapper>& mappers, const SampleView& sv,
                           const std::vector<int32_t>& used, bool gpu_limit, bool sparse,
                           std::vector<std::vector<int32_t>>& groups, std::vector<char>& multi) {
  const size_t ncol = sv.indices->size();
  std::vector<size_t> nz(used.size());
  for (size_t i = 0; i < used.size(); ++i) nz[i] = (*sv.indices)[(size_t)used[i]].size();
  std::vector<size_t> by_cnt(used.size());
  for (size_t i = 0; i < by_cnt.size(); ++i) by_cnt[i] = i;
  std::stable_sort(by_cnt.begin(), by_cnt.end(), [&](size_t a, size_t b) { return nz[a] > nz[b]; });
  std::vector<int32_t> order_by_cnt;
  for (size_t i : by_cnt) order_by_cnt.push_back(used[i]);
  std::vector<std::vector<int32_t>> fixed(ncol);
  std::vector<int32_t> num_per_col(ncol, 0);
  std::vector<const std::vector<int32_t>*> ix(ncol);
  for (size_t f = 0; f < ncol; ++f) ix[f] = &(*sv.indices)[f];
  for (int32_t f : used) {
    fixed[(size_t)f] = fix_sample_indices(mappers[(size_t)f], sv.total, (*sv.indices)[(size_t)f], (*sv.values)[(size_t)f]);
    if (fixed[(size_t)f].empty()) {
      num_per_col[(size_t)f] = (int32_t)(*sv.indices)[(size_t)f].size();
    } else {
      num_per_col[(size_t)f] = (int32_t)fixed[(size_t)f].size();
      ix[(size_t)f] = &fixed[(size_t)f];
    }
  }
  std::vector<std::vector<int32_t>> g2;
  std::vector<char> m2;
  find_groups(mappers, used, ix, num_per_col, sv.total, sv.num_data, gpu_limit, sparse, groups, multi);
  find_groups(mappers, order_by_cnt, ix, num_per_col, sv.total, sv.num_data, gpu_limit, sparse, g2, m2);
  if (groups.size() > g2.size()) {
    groups = std::move(g2);
    multi = std::move(m2);
  }
  Lcg rand(sv.num_data);
  const int32_t ng = (int32_t)groups.size();
  for (int32_t i = 0; i < ng - 1; ++i) {
    const int32_t j = rand.next_short(i + 1, ng);
    std::swap(groups[(size_t)i], groups[(size_t)j]);
    std::swap(multi[(size_t)i], multi[(size_t)j]);
  }
}

// Sampling, bin finding and bundling: the mappers of every column and the groups
// (real columns per group, sub-feature order) with their multi-value flags.
struct Plan {
  std::vector<BinMapper> mappers;
  std::vector<char> used;
  std::vector<std::vector<int32_t>> groups;
  std::vector<char> multi;
};

template <typename T>
Plan plan_dataset(const T* data, int64_t nrow, int32_t ncol, const bk_dataset_params* p, const int32_t* cat,
                  int32_t ncat, int nthreads) {
  if (nrow <= 0 || ncol <= 0) throw std::invalid_argument("empty matrix");
  if (nrow > 2147483647LL) throw std::invalid_argument("too many rows");
  Plan pl;
  // sample rows (sample.rs:15-18) and gather non-zero values (sample.rs:79-98)
  int32_t sample_cnt = std::min((int32_t)nrow, p->bin_construct_sample_cnt);
  std::vector<int32_t> rows = Lcg(p->data_random_seed).sample((int32_t)nrow, sample_cnt);
  std::vector<std::vector<double>> values((size_t)ncol);
  std::vector<std::vector<int32_t>> sidx((size_t)ncol);
  for (size_t i = 0; i < rows.size(); ++i) {
    const T* row = data + (int64_t)rows[i] * ncol;
    for (int32_t k = 0; k < ncol; ++k) {
      double v = (double)row[k];
      if (std::fabs(v) > kZeroThreshold || std::isnan(v)) {
        values[(size_t)k].push_back(v);
        sidx[(size_t)k].push_back((int32_t)i);
      }
    }
  }
  std::set<int32_t> cats(cat, cat + ncat);
  int filter_cnt = (int)(((double)p->min_data_in_leaf * (double)rows.size()) / (double)nrow);
  pl.mappers.resize((size_t)ncol);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int32_t k = 0; k < ncol; ++k) {
    FindBinParams fp{rows.size(), p->max_bin, p->min_data_in_bin, filter_cnt, p->feature_pre_filter != 0,
                     cats.count(k) ? BK_BIN_CATEGORICAL : BK_BIN_NUMERICAL, p->use_missing != 0,
                     p->zero_as_missing != 0};
    pl.mappers[(size_t)k] = find_bin(values[(size_t)k], fp);
  }
  pl.used.assign((size_t)ncol, 0);
  std::vector<int32_t> used_features;
  for (int32_t k = 0; k < ncol; ++k) {
    pl.used[(size_t)k] = !pl.mappers[(size_t)k].is_trivial;
    if (pl.used[(size_t)k]) used_features.push_back(k);
  }
  if (p->enable_bundle && !used_features.empty()) {
    SampleView sv{&rows, &values, &sidx, (int32_t)rows.size(), (int32_t)nrow};
    fast_feature_bundling(pl.mappers, sv, used_features, p->gpu_bundle_limit != 0, p->is_enable_sparse != 0,
                          pl.groups, pl.multi);
  } else {
    for (int32_t k : used_features) pl.groups.push_back({k});
    pl.multi.assign(pl.groups.size(), 0);
  }
  return pl;
}

// Bins of a group as FeatureGroupLayout::new counts them (dataset.rs:35-76),
// multi-value groups included: bin 0 is shared unless the group is a dense
// multi-value one (mean sparse rate under 0.25), which keeps no shared bin, or
// only one when it is group 0 and its first feature's most frequent bin is not 0.
int64_t group_num_total_bin(const Plan& pl, size_t g) {
  const std::vector<int32_t>& feats = pl.groups[g];
  double sparse = 0.0;
  for (int32_t f : feats) sparse += pl.mappers[(size_t)f].sparse_rate;
  sparse /= (double)feats.size();
  int64_t offset = 1;
  bool dense_multi = false;
  if (sparse < (double)0.25f && pl.multi[g]) {
    offset = 0;
    dense_multi = true;
  }
  int64_t total = offset;
  if (g == 0 && !feats.empty() && dense_multi && pl.mappers[(size_t)feats[0]].most_freq_bin > 0) total = 1;
  for (int32_t f : feats) {
    const BinMapper& m = pl.mappers[(size_t)f];
    total += m.num_bin - (m.most_freq_bin == 0 ? offset : 0);
  }
  return total;
}

template <typename T>
bk_dataset* construct(const T* data, int64_t nrow, int32_t ncol, const bk_dataset_params* p, const int32_t* cat,
                      int32_t ncat) {
  // bins are stored one byte per row and feature: u16 bins are not built
  if (p->max_bin < 2 || p->max_bin > 255)
    throw std::invalid_argument("max_bin must be in [2, 255]: bins are stored one byte per row");
  int nthreads = p->num_threads > 0 ? p->num_threads : 1;
#ifdef _OPENMP
  if (p->num_threads <= 0) nthreads = omp_get_max_threads();
#endif
  Plan pl = plan_dataset(data, nrow, ncol, p, cat, ncat, nthreads);
  auto ds = std::make_unique<bk_dataset>();
  ds->num_data = nrow;
  ds->num_total_features = ncol;
  // a group that cannot be one byte per row (multi-valued, or more than 256 bins),
  // or any bundle when asked, is stored exploded, one column per sub-feature
  std::vector<std::vector<int32_t>> stored;
  ds->plan_exploded.assign(pl.groups.size(), 0);
  for (size_t g = 0; g < pl.groups.size(); ++g) {
    const bool explode = pl.groups[g].size() > 1 &&
                         (pl.multi[g] || group_num_total_bin(pl, g) > 256 || p->unbundle_groups != 0);
    ds->plan_exploded[g] = explode ? 1 : 0;
    if (explode)
      for (int32_t f : pl.groups[g]) stored.push_back({f});
    else
      stored.push_back(pl.groups[g]);
  }
  ds->plan_groups = pl.groups;
  ds->plan_multi = pl.multi;
  ds->mappers = std::move(pl.mappers);
  ds->used = std::move(pl.used);
  build_layout(ds.get(), std::move(stored));
  if (std::is_same<T, double>::value) fill_columns(ds.get(), (const double*)data, nullptr, nrow, ncol, nthreads);
  else fill_columns(ds.get(), nullptr, (const float*)data, nrow, ncol, nthreads);
  return ds.release();
}

}  // namespace

extern "C" {

int bk_dataset_from_dense(const double* data, int64_t nrow, int32_t ncol, const bk_dataset_params* p,
                          const int32_t* categorical, int32_t num_categorical, bk_dataset** out) {
  return bk::guard([&] { *out = construct<double>(data, nrow, ncol, p, categorical, num_categorical); });
}

int bk_dataset_from_dense_f32(const float* data, int64_t nrow, int32_t ncol, const bk_dataset_params* p,
                              const int32_t* categorical, int32_t num_categorical, bk_dataset** out) {
  return bk::guard([&] { *out = construct<float>(data, nrow, ncol, p, categorical, num_categorical); });
}

int bk_dataset_from_dense_f32_reference(const float* data, int64_t nrow, int32_t ncol, const bk_dataset* ref,
                                        bk_dataset** out) {
  return bk::guard([&] {
    if (ncol != ref->num_total_features) throw std::invalid_argument("column count differs from the reference");
    auto* ds = new bk_dataset(*ref);
    ds->num_data = nrow;
    fill_columns(ds, nullptr, data, nrow, ncol, 1);
    // re-point metadata into the copy
    for (size_t f = 0; f < ds->f_real.size(); ++f) {
      const BinMapper& m = ds->mappers[(size_t)ds->f_real[f]];
      ds->f_upper[f] = m.bin_type == BK_BIN_NUMERICAL ? m.upper.data() : nullptr;
      ds->f_bin2cat[f] = m.bin_type == BK_BIN_CATEGORICAL ? m.bin2cat.data() : nullptr;
    }
    *out = ds;
  });
}

int bk_dataset_get_view(const bk_dataset* ds, bk_dataset_view* v) {
  return bk::guard([&] {
    std::memset(v, 0, sizeof *v);
    v->num_data = ds->num_data;
    v->num_groups = (int32_t)ds->group_feats.size();
    v->num_features = (int32_t)ds->f_real.size();
    v->num_total_features = ds->num_total_features;
    v->group_bins = ds->col_ptrs.data();
    v->group_num_total_bin = ds->g_num_total_bin.data();
    v->group_feature_start = ds->g_feature_start.data();
    v->group_feature_cnt = ds->g_feature_cnt.data();
    v->group_bin_offsets = ds->g_bin_offsets.data();
    v->feature_num_bin = ds->f_num_bin.data();
    v->feature_missing_type = ds->f_missing.data();
    v->feature_default_bin = ds->f_default.data();
    v->feature_most_freq_bin = ds->f_mfb.data();
    v->feature_bin_type = ds->f_type.data();
    v->feature_real_index = ds->f_real.data();
    v->feature_penalty = ds->f_penalty.data();
    v->feature_bin_upper_bound = ds->f_upper.data();
    v->feature_bin_2_categorical = ds->f_bin2cat.data();
  });
}

int bk_dataset_feature_info(const bk_dataset* ds, int32_t col, char* buf, int32_t cap) {
  return bk::guard([&] {
    if (col < 0 || col >= ds->num_total_features) throw std::invalid_argument("column out of range");
    const std::string& s = ds->infos[(size_t)col];
    if ((int32_t)s.size() + 1 > cap) throw std::invalid_argument("buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int bk_dataset_bundle_layout(const double* data, int64_t nrow, int32_t ncol, const bk_dataset_params* p,
                             const int32_t* categorical, int32_t num_categorical, int32_t* feature_group,
                             int32_t* feature_sub, int64_t* group_num_total_bin_out, int8_t* group_multi_val,
                             int32_t* num_groups) {
  return bk::guard([&] {
    int nthreads = p->num_threads > 0 ? p->num_threads : 1;
    Plan pl = plan_dataset(data, nrow, ncol, p, categorical, num_categorical, nthreads);
    for (int32_t k = 0; k < ncol; ++k) feature_group[k] = feature_sub[k] = -1;
    for (size_t g = 0; g < pl.groups.size(); ++g) {
      for (size_t j = 0; j < pl.groups[g].size(); ++j) {
        feature_group[pl.groups[g][j]] = (int32_t)g;
        feature_sub[pl.groups[g][j]] = (int32_t)j;
      }
      group_num_total_bin_out[g] = group_num_total_bin(pl, g);
      group_multi_val[g] = (int8_t)pl.multi[g];
    }
    *num_groups = (int32_t)pl.groups.size();
  });
}

int bk_dataset_free(bk_dataset* ds) {
  delete ds;
  return 0;
}

}  // extern "C"
