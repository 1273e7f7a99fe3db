/*
 * bankai_gpu.h — C ABI of the MI355X histogram tree learner.
 *
 * The library replaces the seat that `SerialTreeLearner` holds inside the
 * reference's `Trainer` (crates/lgbm-core/src/boosting/train.rs:483). Each
 * entry point below names the reference method it stands in for.
 *
 * Conventions (mirroring crates/lgbm-capi/src/lib.rs:23-26,63-90,181):
 *   - every function returns 0 on success and -1 on failure;
 *   - on failure a thread-local message is kept, read with bk_last_error();
 *   - no C++ exception crosses this boundary;
 *   - host arrays passed in are borrowed for the duration of the call only;
 *     device memory is owned by the handle;
 *   - one handle is driven by one thread at a time.
 *
 * Plain pointers and sizes only: no torch or HIP types appear here.
 */
#ifndef BANKAI_GPU_H
#define BANKAI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Missing-value handling of a feature (crates/lgbm-core/src/meta.rs:10-16). */
enum { BK_MISSING_NONE = 0, BK_MISSING_ZERO = 1, BK_MISSING_NAN = 2 };
/* Bin type of a feature (crates/lgbm-core/src/io/bin.rs BinType). */
enum { BK_BIN_NUMERICAL = 0, BK_BIN_CATEGORICAL = 1 };

/* Learner forms. */
enum {
  BK_FORM_PLAIN_F64 = 0,      /* f64 histograms (the plain drop-in learner)   */
  BK_FORM_F32_COMPENSATED = 1,/* f32 histograms with two-sum compensation     */
  BK_FORM_QUANTIZED = 2       /* int8 gradients, exact integer histograms     */
};

/*
 * Binned dataset, column-major by feature group. Mirrors the layout fields of
 * `DatasetLayout` / `FeatureGroupLayout` (crates/lgbm-core/src/io/dataset.rs:21-93)
 * and the per-feature `BinMapper` metadata (crates/lgbm-core/src/io/bin.rs:21-36)
 * the learner reads. Only dense single-value groups with at most 256 group bins
 * (one byte per row) are accepted.
 */
typedef struct bk_dataset_view {
  int64_t num_data;
  int32_t num_groups;
  int32_t num_features;                 /* inner (used) features              */
  int32_t num_total_features;           /* real columns of the raw matrix      */
  const uint8_t* const* group_bins;     /* [num_groups] -> num_data bytes       */
  const int32_t* group_num_total_bin;   /* [num_groups]                         */
  const int32_t* group_feature_start;   /* [num_groups]                         */
  const int32_t* group_feature_cnt;     /* [num_groups]                         */
  /* bin_offsets of every group, concatenated: group g owns the
   * group_feature_cnt[g] + 1 entries starting at group_feature_start[g] + g. */
  const int32_t* group_bin_offsets;
  /* per inner feature */
  const int32_t* feature_num_bin;
  const int32_t* feature_missing_type;
  const int32_t* feature_default_bin;
  const int32_t* feature_most_freq_bin;
  const int32_t* feature_bin_type;
  const int32_t* feature_real_index;
  const double* feature_penalty;        /* NULL = all 1.0                       */
  /* Numerical: num_bin upper bounds (BinMapper::bin_upper_bound).
   * Categorical: NULL (bin_2_categorical is used instead). */
  const double* const* feature_bin_upper_bound;
  const int32_t* const* feature_bin_2_categorical;
} bk_dataset_view;

/*
 * Learner parameters: the `Config` fields the tree learner reads
 * (SplitParams / CatParams at crates/lgbm-core/src/treelearner/feature_histogram.rs:32-56,1146-1164
 * and the quantized options at crates/lgbm-core/src/config/auto.rs:85-88).
 */
typedef struct bk_learner_params {
  int32_t num_leaves;
  int32_t max_depth;
  int32_t min_data_in_leaf;
  double min_sum_hessian_in_leaf;
  double lambda_l1;
  double lambda_l2;
  double max_delta_step;
  double path_smooth;
  double min_gain_to_split;
  double cat_smooth;
  double cat_l2;
  int32_t max_cat_threshold;
  int32_t max_cat_to_onehot;
  int32_t min_data_per_group;
  int32_t extra_trees;            /* must be 0 (refused otherwise)              */
  double histogram_pool_size;     /* MB; <= 0 keeps one slot per leaf           */
  int32_t use_quantized_grad;
  int32_t num_grad_quant_bins;
  int32_t stochastic_rounding;
  int32_t quant_train_renew_leaf;
  int32_t seed;                   /* Config::seed, seeds the rounding noise      */
  int32_t num_threads;            /* fixes the rounding-noise block layout       */
  int32_t form;                   /* BK_FORM_*                                   */
  int32_t batched;                /* 1 = speculative depth-batched growth        */
  double feature_fraction;        /* ColSampler (treelearner/col_sampler.rs): by-tree  */
  double feature_fraction_bynode; /* and by-node fractions, each in (0, 1]             */
  int32_t feature_fraction_seed;  /* seeds the sampler's LCG (Config::set derives it)  */
  /* interaction_constraints (Config::interaction_constraints_vector): constraint i is the
   * real feature ids interaction_constraint_features[interaction_constraint_offsets[i] ..
   * interaction_constraint_offsets[i + 1]); none when num_interaction_constraints is 0.
   * The arrays are read at creation only. */
  int32_t num_interaction_constraints;
  const int32_t* interaction_constraint_offsets;
  const int32_t* interaction_constraint_features;
} bk_learner_params;

/*
 * One trained tree: the training-time fields of `Tree` (crates/lgbm-core/src/tree.rs:18-49).
 * The caller allocates every array with capacity num_leaves_max (leaf arrays) or
 * num_leaves_max - 1 (node arrays); cat_threshold with cat_threshold_cap words.
 */
typedef struct bk_tree_out {
  int32_t num_leaves;
  int32_t num_cat;
  int32_t* split_feature_inner;
  int32_t* split_feature;
  uint32_t* threshold_in_bin;
  double* threshold;
  int8_t* decision_type;
  int32_t* left_child;
  int32_t* right_child;
  float* split_gain;
  double* internal_value;
  double* internal_weight;
  int32_t* internal_count;
  double* leaf_value;
  double* leaf_weight;
  int32_t* leaf_count;
  int32_t* leaf_parent;
  int32_t* leaf_depth;
  int32_t* cat_boundaries;        /* num_leaves_max entries                      */
  uint32_t* cat_threshold;
  int32_t cat_threshold_cap;
  int32_t cat_threshold_len;
} bk_tree_out;

typedef struct bk_learner bk_learner;

/* Thread-local message of the last failure (crates/lgbm-capi/src/lib.rs:181 LGBM_GetLastError). */
const char* bk_last_error(void);

/* SerialTreeLearner::new (crates/lgbm-core/src/treelearner/serial.rs:284).
 * Uploads the bin columns once and transposes them into row records that stay
 * resident in HBM for the learner's lifetime. */
int bk_learner_create(const bk_dataset_view* ds, const bk_learner_params* params,
                      int is_constant_hessian, bk_learner** out);

/* SerialTreeLearner::train (crates/lgbm-core/src/treelearner/serial.rs:569) with host
 * gradient/hessian arrays of num_data entries. rows: the bagged subset (num_rows
 * distinct rows, ascending, each in [0, num_data)) the tree is grown on, or NULL for
 * every row (DataPartition::init_with / init, data_partition.rs:91-107). The
 * quantized scales always come from every row, as in the reference. */
int bk_learner_train(bk_learner* learner, const float* gradients, const float* hessians,
                     const int32_t* rows, int64_t num_rows, bk_tree_out* tree);

/* Same as bk_learner_train with gradients and hessians already in device memory
 * (num_data floats each). Keeps the boosting loop on the GPU. */
int bk_learner_train_device(bk_learner* learner, const float* d_gradients,
                            const float* d_hessians, bk_tree_out* tree);

/* bk_learner_train on device arrays with a bag in device memory (d_rows: num_rows
 * ascending rows). Rows out of order or outside the dataset fail the call (-1)
 * when the tree ends. */
int bk_learner_train_device_bagged(bk_learner* learner, const float* d_gradients,
                                   const float* d_hessians, const int32_t* d_rows,
                                   int64_t num_rows, bk_tree_out* tree);

/* SerialTreeLearner::add_prediction_to_score (crates/lgbm-core/src/treelearner/serial.rs:1974):
 * score[row] += leaf_value[leaf(row)] for the last trained tree, host scores, over
 * the rows the tree was trained on (its bag, or every row). */
int bk_learner_add_prediction_to_score(bk_learner* learner, const bk_tree_out* tree,
                                       double* score);

/* The same on a device score array of num_data doubles. */
int bk_learner_add_prediction_to_score_device(bk_learner* learner, const bk_tree_out* tree,
                                              double* d_score);

/* SerialTreeLearner::add_out_of_bag_prediction (serial.rs:1995): the last tree
 * replayed on `rows` (any rows of the dataset, usually those outside the bag,
 * train.rs:1282-1291) with the training routing; score[row] += its leaf's value. */
int bk_learner_add_out_of_bag_prediction(bk_learner* learner, const bk_tree_out* tree,
                                         const int32_t* rows, int64_t num_rows, double* score);

/* The same on a device score array, over every row outside the last tree's bag
 * (nothing when it had no bag). */
int bk_learner_add_out_of_bag_prediction_device(bk_learner* learner, const bk_tree_out* tree,
                                                double* d_score);

/* SerialTreeLearner::histogram_cache_capacity (serial.rs:250): the histogram slots
 * of the reference's pool (num_leaves, or what histogram_pool_size allows). */
int bk_learner_histogram_cache_capacity(bk_learner* learner, int64_t* capacity);

/* SerialTreeLearner::reset_histogram_pool_config (serial.rs:256), the budget-only
 * reset: the pool is kept and the feature sampler draws a by-tree sample from its
 * continuing stream. */
int bk_learner_reset_histogram_pool_config(bk_learner* learner);

/* SerialTreeLearner::continue_after_dataset_reset (serial.rs:266): the feature
 * sampler continues `previous`'s stream and draws a by-tree sample from it. */
int bk_learner_continue_after_dataset_reset(bk_learner* learner, const bk_learner* previous);

/* How one split of a tree routes rows (SplitRecord, crates/lgbm-core/src/treelearner/
 * split_info.rs and serial.rs:2003): the group whose bin is read and the learner's
 * routing rule, stored as it was applied. Opaque to the caller, who keeps the
 * records of each tree to replay it later. */
typedef struct bk_split_record {
  int32_t group;
  uint8_t rule[124];
} bk_split_record;

/* SerialTreeLearner::last_splits (serial.rs:2003): the records of the last tree's
 * splits, one per internal node in node order, into out (capacity records; out may
 * be NULL to ask for the count). *num_out receives num_leaves - 1. */
int bk_learner_last_splits(bk_learner* learner, bk_split_record* out, int32_t capacity,
                           int32_t* num_out);

/* SerialTreeLearner::add_prediction_for_rows (serial.rs:2009): an earlier tree of
 * this learner (its children and leaf values from `tree`, its routing from the
 * records last_splits gave for it) replayed on `rows` of the dataset;
 * score[row] += the leaf value (DART and rollback, train.rs:249,961,1071). */
int bk_learner_add_prediction_for_rows(bk_learner* learner, const bk_tree_out* tree,
                                       const bk_split_record* splits, int32_t num_splits,
                                       const int32_t* rows, int64_t num_rows, double* score);

/* DataPartition::indices_on_leaf (crates/lgbm-core/src/treelearner/data_partition.rs:109-112),
 * read through the public `partition` field by Trainer::renew_tree_output
 * (crates/lgbm-core/src/boosting/train.rs:1198-1211). Copies the rows of `leaf`,
 * ascending, into rows_out (capacity num_data); *num_out receives the count. */
int bk_learner_leaf_rows(bk_learner* learner, int32_t leaf, int32_t* rows_out,
                         int64_t* num_out);

/* Root histogram of host gradients over all rows, as
 * SerialTreeLearner::construct_histograms (crates/lgbm-core/src/treelearner/serial.rs:891-1065)
 * leaves it for the root before fix_histogram: per group bin b (bins of all groups
 * concatenated, total_bins = sum of group_num_total_bin),
 *   float forms:  hist_out[2b] = sum of g, hist_out[2b+1] = sum of h (a constant
 *                 hessian counts rows and is scaled by hessians[0], serial.rs:982-988);
 *   quantized:    int_hist_out[2b], [2b+1] = exact integer sums of the discretized
 *                 gradients (construct_int_histograms, serial.rs:1074-1164);
 * count_out[b] = rows in bin b. Each output may be NULL; hist_out must be NULL for the
 * quantized form and int_hist_out for the float forms. It runs the per-tree set-up
 * of train (discretization included), so a stochastic learner's noise offset
 * advances as in a train call. Used to report the per-bin error of the FP32
 * compensated form against an FP64 histogram. */
int bk_learner_root_histogram(bk_learner* learner, const float* gradients, const float* hessians,
                              double* hist_out, int64_t* int_hist_out, int32_t* count_out);

/* bk_learner_root_histogram without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. Inputs are
 * device arrays of num_data gradients and hessians; outputs are device arrays the
 * caller allocated: d_hist_out holds 2 x total_bins values (double for the float
 * forms, int64_t for the quantized form, laid out as bk_learner_root_histogram's
 * outputs) and d_count_out total_bins int32 counts. The partial histograms live in
 * a file-scope 16 MiB device array shared by every call on a device; a call whose
 * partials cannot fit (total_bins x 8 bytes quantized, x 20 bytes float, times
 * ceil(num_data / rows of one LDS accumulation)) fails with an error that says so.
 * Calls that overlap on one device race on that array. The call runs on the
 * learner's stream and returns when the outputs are written; the caller's inputs
 * must be ready on the device before the call. */
int bk_learner_root_histogram_noalloc(bk_learner* learner, const float* d_gradients, const float* d_hessians,
                                      void* d_hist_out, int32_t* d_count_out);

/* SerialTreeLearner::reset_constant_hessian (crates/lgbm-core/src/treelearner/serial.rs:279). */
int bk_learner_reset_constant_hessian(bk_learner* learner, int is_constant_hessian);

/* Binary log-loss gradients and hessians (crates/lgbm-core/src/objective/binary.rs:71-90)
 * from a device score array, written to device arrays, on the learner's stream.
 * label_weights = {negative, positive} as BinaryLogloss::new derives them from
 * is_unbalance and scale_pos_weight (binary.rs:44-52; {1, 1} without them);
 * d_weights: per-row sample weights, or NULL. The device exp is not the host
 * libm's: results agree with the host objective within 2 f32 ulps. A device-side
 * boosting loop uses it; it has no FFI counterpart in the reference, whose
 * objective runs on the host. */
int bk_learner_binary_gradients_device(bk_learner* learner, const double* d_score, const float* d_label,
                                       const float* d_weights, double sigmoid, const double label_weights[2],
                                       float* d_gradients, float* d_hessians);

/* Softmax gradients and hessians (crates/lgbm-core/src/objective/multiclass.rs:68-90,
 * unweighted) for num_class classes from a device score array of num_class * num_data
 * doubles (class k at [k * num_data, (k + 1) * num_data)), written the same way, on the
 * learner's stream. A device-side multiclass loop uses it; no FFI counterpart. */
int bk_learner_softmax_gradients_device(bk_learner* learner, const double* d_score, const float* d_label,
                                        int num_class, float* d_gradients, float* d_hessians);

/* LambdaRank tables (crates/lgbm-core/src/objective/rank.rs:183-240), all in device
 * memory: query row offsets, 1 / max DCG per query, label gains, position discounts
 * and the sigmoid table with its input range. */
typedef struct bk_rank_tables {
  const int32_t* query_bounds;     /* num_queries + 1 */
  const double* inverse_max_dcg;   /* num_queries */
  const double* label_gain;        /* num_gain */
  const double* discount;          /* per rank, at least the largest query's rows */
  const double* sigmoid_table;     /* table_bins */
  double sigmoid, min_input, max_input, table_factor;
  int32_t num_queries, num_gain, table_bins, truncation_level, norm, max_query_rows;
} bk_rank_tables;

/* LambdaRank NDCG gradients and hessians (rank.rs:262-340, unweighted, no position
 * bias) of every query from a device score array, on the learner's stream. The pair
 * terms are the reference's; their float sums run in another order. Queries may hold
 * at most 2048 rows. A device-side ranking loop uses it; no FFI counterpart. */
int bk_learner_lambdarank_gradients_device(bk_learner* learner, const double* d_score, const float* d_label,
                                           const bk_rank_tables* tables, float* d_gradients, float* d_hessians);

/* Trains tree `slot` of every `trees_per_iteration` trees of a sequence shared by
 * several learners, e.g. one learner per class of a multiclass iteration, each
 * on its own stream and host thread, so the class trees grow concurrently. The
 * learner consumes the stochastic-rounding rotation draws of every tree of the
 * sequence (GradientDiscretizer, crates/lgbm-core/src/treelearner/gradient_discretizer/mod.rs:172-201)
 * and the by-tree feature samples of every tree (ColSampler, col_sampler.rs:87-101),
 * and keeps its own, so the trees equal those of one learner training the
 * whole sequence in order, provided no tree of the sequence has its
 * discretization refused (all-zero or non-finite gradients), which skips a draw.
 * Refused (-1) with by-node feature sampling, whose draws depend on the other
 * trees' searches. No reference counterpart: the reference trains the class trees
 * one by one. */
int bk_learner_set_tree_slot(bk_learner* learner, int slot, int trees_per_iteration);

/* Histogram tiling override: rows per workgroup at least min_rows, row blocks at
 * most max_blocks (0, 0: back to what the autotuner picked at creation). Results
 * do not depend on the tiling (integer histograms are exact; the float forms stay
 * within their tolerance); tests use it to check that, and to force row blocks
 * larger than one LDS accumulation. No reference counterpart (tuning only). */
int bk_learner_set_hist_tiling(bk_learner* learner, int32_t min_rows, int32_t max_blocks);

/* Histogram layout override: a new LDS tile every groups_per_tile groups (a
 * positive multiple of 16, the groups of one 16-byte record chunk; a tile is
 * also cut where its bins would exceed the LDS budget), and sub_histograms LDS
 * copies of each tile (1 or 2; wave w adds to copy w % sub_histograms). (0, 0):
 * back to the autotuner's choice. The autotuner picks the row tiling and this
 * layout per dataset shape; results do not depend on either. No reference
 * counterpart (tuning only). */
int bk_learner_set_hist_layout(bk_learner* learner, int32_t groups_per_tile, int32_t sub_histograms);

/* Bytes per histogram bin of the quantized learner's batched growth with
 * host-planned rounds: 4 (16-bit gradient and hessian sums, no row count: the
 * children's exact counts then come from the partition, whose right rows fill
 * each segment from its end), 8 (21-bit sums and a row count, exact counts
 * from the search), or 0 (the default: 4 where it applies). 4 applies to
 * num_grad_quant_bins <= 13 (a row block then holds at least 4096 rows within
 * the 16-bit fields) and num_leaves <= 1023; asking for 4 elsewhere is refused
 * (-1). The sequential form and the device planner use 8. Both widths grow the
 * same tree, bit for bit. The reference picks 8-, 16- or 32-bit histogram
 * fields per leaf from its row count
 * (crates/lgbm-core/src/treelearner/gradient_discretizer/mod.rs:223-232); this
 * choice is per learner and by bins, since a GPU block's rows are bounded
 * instead. */
int bk_learner_set_hist_bin_width(bk_learner* learner, int32_t bytes);

/* With 4-byte bins, on = 1 has the partition also build the histogram of one
 * child of every split: each partition tile adds the rows it routes to that side
 * into an LDS histogram of the whole record and writes it as one partial, so no
 * separate histogram launch reads the child's rows again. The side is the
 * smaller child by the search's estimated counts (the integer histograms, and
 * the trees, do not depend on it). Applies where the record's bins fit in LDS
 * beside the partition's tables (about 23,000 bins). on = 0, the default,
 * builds the smaller child in its own launch: on MI355X at 10.5M x 28 that is
 * faster (the atomics do not overlap the partition's streaming, and a partial
 * per 4096-row tile costs more to reduce). No reference counterpart. */
int bk_learner_set_partition_histograms(bk_learner* learner, int32_t on);

/* Who plans the rounds of batched growth: 0 (the default, measured fastest on
 * MI355X: the host sees a round's end within microseconds and plans it in about
 * ten) the host, waiting for every round's results; 1 the device, for the whole
 * tree: its planner kernel replays the commits and picks each round's work,
 * and the discretizer's scales, the root's sums, the row-to-leaf walk and the
 * leaf renewal also run on the device, so the host waits once per tree; 2 the
 * device plans the rounds, the host the tree's start and end. All grow the same
 * tree. The device
 * planner takes up to 256 leaves and no by-node feature sampling or interaction
 * constraints; a whole-tree device run also needs no bag and, when quantized,
 * power-of-two gradient bins with num_data * (bins + 1) < 2^29 (the root sums are
 * then exact in any order). Otherwise the next mode down runs. No reference
 * counterpart (the reference grows one split per step,
 * crates/lgbm-core/src/treelearner/serial.rs:606-621). */
int bk_learner_set_round_planner(bk_learner* learner, int on_device);

/* Per-tree counters of the last train call. hist_rows / part_rows count every
 * row the kernels processed (the batched form includes speculative expansions
 * that were never committed); committed_hist_rows / committed_part_rows count
 * the reference algorithm's own work: sum of R_b over histogram builds (root N
 * plus each searched smaller child) and sum of Q_s over committed splits.
 * rounds = search launches; speculative_waste = expansions not committed. */
typedef struct bk_train_stats {
  int64_t hist_rows;
  int64_t part_rows;
  int32_t splits;
  int32_t rounds;
  int32_t speculative_waste;
  int32_t hist_bin_bytes;   /* bytes per quantized histogram bin of the last tree (0: float forms) */
  int64_t committed_hist_rows;
  int64_t committed_part_rows;
  double host_ms;  /* wall time of the train call not spent blocked on the device */
  double wait_ms;  /* wall time blocked on the device (synchronizations)           */
  int32_t hist_min_rows;    /* histogram tiling in use (the autotuner's unless forced): */
  int32_t hist_max_blocks;  /* rows per workgroup at least / row blocks at most        */
  int32_t hist_tile_groups; /* groups per LDS tile (a new tile every that many groups) */
  int32_t hist_ntiles;      /* group tiles of a histogram build                       */
  int32_t hist_sub_histograms; /* LDS copies of a tile's histogram                    */
  int32_t partition_histograms; /* 1: the partition built the children's histograms  */
  /* host time of batched growth with 4-byte bins, by phase (within host_ms): */
  double replay_ms;    /* the reference's commits replayed as children become known */
  double plan_ms;      /* the frontier ranked and the round's items written            */
  double launch_ms;    /* the round's upload and launches                              */
  double readback_ms;  /* the children's counts and folded splits read back            */
} bk_train_stats;
int bk_learner_last_stats(bk_learner* learner, bk_train_stats* out);

/* Per-kernel timing with HIP events recorded on the learner's stream around
 * every launch (off by default). Kinds: 0 root histogram, 1 child histograms,
 * 2 partition, 3 split search, 4 discretizer (with the gradient maxima), 5 leaf
 * renewal, 6 score update, tree walk and gradients, 7 the device round planner,
 * 8 transfers (the learner's small copies and stamps on its stream). algo_bytes follows the per-row byte model of the histogram
 * (rows x (features + payload), plus a 4-byte row index for a child) and of the
 * partition (rows x 9: one bin, row id read, row id write), the reference's
 * index-based model; moved_bytes counts what the kernel actually reads and
 * writes (whole row records). No reference counterpart
 * (measurement only). */
typedef struct bk_kernel_stats {
  double ms;
  int64_t launches;
  double algo_bytes;
  double moved_bytes;
} bk_kernel_stats;
/* kinds: bit k turns timing of kind k on (0 = off); the counters are reset. */
int bk_learner_profile(bk_learner* learner, int kinds);
/* Times only the launches of every `every`-th tree from now on (1, the default:
 * every tree). Each timed launch costs about 25 microseconds of host time, which
 * lowers the throughput of a timed run; sampling whole trees keeps the average
 * over every kind of launch of a tree. */
int bk_learner_profile_trees(bk_learner* learner, int every);
int bk_learner_kernel_stats(bk_learner* learner, int kind, bk_kernel_stats* out);

/* Frees the handle and its device memory. NULL is accepted. */
int bk_learner_free(bk_learner* learner);

#ifdef __cplusplus
}
#endif

#endif /* BANKAI_GPU_H */
