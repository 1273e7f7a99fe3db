/*
 * bankai_dataset.h — host-side dataset construction (sampling, binning, group
 * layout) that produces the bk_dataset_view the learner consumes.
 *
 * Replaces `Dataset::from_dense` (crates/lgbm-core/src/io/dataset.rs:904-915):
 * `SampleSet::from_dense` + `find_bins` (crates/lgbm-core/src/io/sample.rs:31-196),
 * `BinMapper::find_bin` (crates/lgbm-core/src/io/bin.rs:342-553),
 * `DatasetLayout::construct` (dataset.rs:378-447) with `FastFeatureBundling`
 * (dataset.rs:282-363) and `fill_dense` (dataset.rs:633-689). Groups are stored one
 * byte per row; a multi-value group or a bundle of more than 256 bins is stored
 * exploded, one column per sub-feature, holding exactly the values the reference's
 * group holds (for a bundle, the last column a row pushes), with the reference's
 * inner feature order.
 * Same error convention as bankai_gpu.h (0 / -1 and bk_last_error()).
 */
#ifndef BANKAI_DATASET_H
#define BANKAI_DATASET_H

#include <stdint.h>
#include "bankai_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The Config fields dataset construction reads (crates/lgbm-core/src/config/auto.rs:92-99,244-251). */
typedef struct bk_dataset_params {
  int32_t max_bin;                  /* default 255; in [2, 255]: one byte per bin, u16 bins are not built */
  int32_t min_data_in_bin;          /* default 3       */
  int32_t bin_construct_sample_cnt; /* default 200000  */
  int32_t data_random_seed;         /* default 1       */
  int32_t use_missing;              /* default 1       */
  int32_t zero_as_missing;          /* default 0       */
  int32_t feature_pre_filter;       /* default 1       */
  int32_t min_data_in_leaf;         /* default 20 (pre-filter count) */
  int32_t num_threads;              /* 0 = all cores   */
  int32_t enable_bundle;            /* default 1: exclusive feature bundling (EFB) */
  int32_t is_enable_sparse;         /* default 1: regroup sparse groups (FindGroups) */
  int32_t gpu_bundle_limit;         /* default 0; 1 = the bundling of device_type "gpu": groups of
                                       at most 256 bins (dataset.rs:159-181,259) */
  int32_t unbundle_groups;          /* default 0; 1 = store every bundle exploded (one column per
                                       sub-feature); groups that cannot be one byte per row always are */
} bk_dataset_params;

typedef struct bk_dataset bk_dataset;

/* Row-major dense matrix of f64 (nrow x ncol). categorical: real column ids. */
int bk_dataset_from_dense(const double* data, int64_t nrow, int32_t ncol, const bk_dataset_params* params,
                          const int32_t* categorical, int32_t num_categorical, bk_dataset** out);
/* The same from f32 values (each widened exactly to f64 first). */
int bk_dataset_from_dense_f32(const float* data, int64_t nrow, int32_t ncol, const bk_dataset_params* params,
                              const int32_t* categorical, int32_t num_categorical, bk_dataset** out);
/* Bins new rows with the bin mappers and layout of `reference`
 * (Dataset::from_dense_reference, crates/lgbm-core/src/io/dataset.rs:981). */
int bk_dataset_from_dense_f32_reference(const float* data, int64_t nrow, int32_t ncol,
                                        const bk_dataset* reference, bk_dataset** out);
/* Fills a view whose pointers stay valid while the dataset lives. */
int bk_dataset_get_view(const bk_dataset* ds, bk_dataset_view* view);
/* Model-text `feature_infos` entry of real column `col` ("none" when unused). */
int bk_dataset_feature_info(const bk_dataset* ds, int32_t col, char* buf, int32_t cap);
/* The layout Dataset::Construct would give these rows (io/dataset.rs:378-447 with
 * FastFeatureBundling, dataset.rs:282-363), without building the bins: for every
 * column its group and sub-feature index (-1: unused), and per group (capacity ncol)
 * its number of bins as FeatureGroupLayout::new counts them and whether it is
 * multi-valued. Groups that do not fit one byte per row (more than 256 bins,
 * multi-valued) are reported as the reference plans them; bk_dataset_from_dense
 * stores them exploded. */
int bk_dataset_bundle_layout(const double* data, int64_t nrow, int32_t ncol, const bk_dataset_params* params,
                             const int32_t* categorical, int32_t num_categorical, int32_t* feature_group,
                             int32_t* feature_sub, int64_t* group_num_total_bin, int8_t* group_multi_val,
                             int32_t* num_groups);

int bk_dataset_free(bk_dataset* ds);

#ifdef __cplusplus
}
#endif

#endif /* BANKAI_DATASET_H */
