// kernels.h — data shared between the host learner (learner.cpp, built by g++)
// and the gfx950 kernels (kernels.hip, built by hipcc): per-feature metadata, the
// split routing rule, the per-feature search result, and the launchers.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define BK_HD __host__ __device__
#else
#define BK_HD
#endif

namespace bk {

enum Form { FORM_F64 = 0, FORM_F32C = 1, FORM_QUANT = 2 };

constexpr int kMaxBins = 256;         // group bins per column (u8 storage)
constexpr int kRowThreads = 256;      // threads of the row-streaming kernels (histogram, partition)
constexpr int kRowTile = 1024;        // rows per tile of those kernels: 4 per thread
constexpr int kTileRows = 2048;       // granularity of the row blocks (tasks) the host cuts
constexpr int kMaxRowBlocks = 512;    // row blocks per histogram build (partials)
constexpr int kLdsBytes = 120 * 1024; // dynamic LDS budget per workgroup on gfx950 (160 KB per CU)
constexpr int kHistLdsBytes = 72 * 1024;  // LDS histogram budget: two row blocks fit on one CU
constexpr int kMaxGroups = 1024;      // groups per row record
constexpr int kScanThreads = 256;    // threads of the split-search blocks
constexpr int kMaxSubHist = 2;       // LDS copies of a histogram tile (waves alternate between them)

// Per inner feature, as the kernels read it.
struct FeatDev {
  int num_bin, missing_type, default_bin, mfb, offset, bin_type;
  int group;       // group index (= column)
  int hist_off;    // first bin of the feature slice in a leaf histogram
  int hist_len;    // num_bin - offset
  int bin_base;    // bin_offsets[sub] inside the group
  int minb, maxb, use_min_bin;  // group_view (data_partition.rs:25-64)
  int real;
  double penalty;
};

// SplitRule (data_partition.rs:233-410), dense u8 columns.
struct RuleDev {
  int is_cat, use_min_bin, multi_bin;
  uint32_t minb, maxb, th, t_zero_bin, cat_offset;
  int miss_is_zero, miss_is_na, mfb_is_zero, mfb_is_na;
  int default_left_route, missing_default_left, max_bin_left;
  uint32_t cat_bits[8];
};

// Word selection with constant indices only: a runtime index into the bitset
// would move the whole struct into per-lane scratch memory on the GPU.
BK_HD inline uint32_t bits_word(const uint32_t* bits, uint32_t i) {
  uint32_t w = 0u;
  for (uint32_t k = 0; k < 8u; ++k)
    if (i == k) w = bits[k];
  return w;
}
BK_HD inline void bits_set(uint32_t* bits, uint32_t pos) {
  for (uint32_t k = 0; k < 8u; ++k)
    if ((pos >> 5) == k) bits[k] |= 1u << (pos & 31u);
}
BK_HD inline bool rule_bit(const RuleDev& r, uint32_t pos) {
  uint32_t i1 = pos / 32u;
  if (i1 >= 8u) return false;
  return (bits_word(r.cat_bits, i1) >> (pos % 32u)) & 1u;
}

// route_bin (data_partition.rs:366-410) for one stored group value.
BK_HD inline bool route_left(const RuleDev& r, uint32_t bin) {
  if (r.is_cat) {
    if ((r.use_min_bin && (bin < r.minb || bin > r.maxb)) || (!r.use_min_bin && bin == 0u)) return r.default_left_route;
    return rule_bit(r, bin - r.minb + r.cat_offset);
  }
  if (r.multi_bin) {
    if ((r.miss_is_zero && !r.mfb_is_zero && bin == r.t_zero_bin) || (r.miss_is_na && !r.mfb_is_na && bin == r.maxb))
      return r.missing_default_left;
    if ((r.use_min_bin && (bin < r.minb || bin > r.maxb)) || (!r.use_min_bin && bin == 0u)) {
      if ((r.miss_is_na && r.mfb_is_na) || (r.miss_is_zero && r.mfb_is_zero)) return r.missing_default_left;
      return r.default_left_route;
    }
    return !(bin > r.th);
  }
  if (r.miss_is_zero && !r.mfb_is_zero && bin == r.t_zero_bin) return r.missing_default_left;
  if (bin != r.maxb) {
    if ((r.miss_is_na && r.mfb_is_na) || (r.miss_is_zero && r.mfb_is_zero)) return r.missing_default_left;
    return r.default_left_route;
  }
  if (r.miss_is_na && !r.mfb_is_na) return r.missing_default_left;
  return r.max_bin_left;
}

// SplitRule::numerical (data_partition.rs:299-363).
BK_HD inline RuleDev make_rule_numerical(const FeatDev& f, uint32_t threshold, bool default_left) {
  RuleDev r{};
  r.is_cat = 0;
  r.use_min_bin = f.use_min_bin;
  r.minb = (uint32_t)f.minb;
  r.maxb = (uint32_t)f.maxb;
  uint32_t default_bin = (uint32_t)f.default_bin, mfb = (uint32_t)f.mfb;
  if (f.missing_type == 1) {
    r.miss_is_zero = 1;
    r.mfb_is_zero = default_bin == mfb;
  } else if (f.missing_type == 2) {
    r.miss_is_na = 1;
    r.mfb_is_na = f.use_min_bin ? (r.maxb == mfb + r.minb && mfb > 0u) : (r.maxb == mfb + 1u && mfb > 0u);
  }
  uint32_t th = threshold + r.minb;
  uint32_t tz = r.minb + default_bin;
  if (mfb == 0u) {
    th -= 1u;
    tz -= 1u;
  }
  r.th = th;
  r.t_zero_bin = tz;
  r.default_left_route = mfb <= threshold;
  r.missing_default_left = (r.miss_is_zero || r.miss_is_na) && default_left;
  r.max_bin_left = r.maxb <= th;
  r.multi_bin = r.minb < r.maxb;
  return r;
}

// SplitRule::categorical (data_partition.rs:266-296); bits = bitset of left bins.
BK_HD inline RuleDev make_rule_categorical(const FeatDev& f, const uint32_t* bits) {
  RuleDev r{};
  r.is_cat = 1;
  r.use_min_bin = f.use_min_bin;
  r.minb = (uint32_t)f.minb;
  r.maxb = (uint32_t)f.maxb;
  r.cat_offset = f.mfb == 0 ? 1u : 0u;
  for (int i = 0; i < 8; ++i) r.cat_bits[i] = bits[i];
  r.default_left_route = f.mfb > 0 && rule_bit(r, (uint32_t)f.mfb);
  return r;
}

// Stored group value of real bin b of feature f (fill_dense, dataset.rs:676-684);
// the most frequent bin is not stored and reads as 0.
BK_HD inline uint32_t stored_value(const FeatDev& f, int b) {
  if (b == f.mfb) return 0u;
  return (uint32_t)(b + f.bin_base - (f.mfb == 0 ? 1 : 0));
}

// One feature's best split for one leaf (SplitInfo, split_info.rs:9-31) plus the
// exact row counts the split routes left and right (from the count channel).
struct SplitOut {
  double gain, left_output, right_output;
  double left_sum_gradient, left_sum_hessian, right_sum_gradient, right_sum_hessian;
  long long left_int_sum, right_int_sum;
  int threshold, left_count, right_count, exact_left, exact_right;
  int default_left, splittable, num_cat, feature, pad;
  uint32_t cat_bits[8];
};

// Scalars of one leaf handed to the search.
struct LeafArgs {
  int leaf, slot, num_data;     // num_data = exact rows of the leaf
  double sum_g, sum_h, parent_output;
  long long int_sum;
  int part_first;               // first partial histogram (task index) of its build
  int nblk;                     // partial histograms to reduce
};

struct ScanParams {
  double l1, l2, max_delta_step, path_smooth, min_sum_hessian, min_gain_to_split;
  double cat_smooth, cat_l2;
  int min_data_in_leaf, max_cat_threshold, max_cat_to_onehot, min_data_per_group;
  double grad_scale, hess_scale, h0;  // quantized scales; h0 = hessians[0] for a constant hessian
  int const_hess;
  const double* dyn;  // when set: grad_scale, hess_scale, h0 are read from the device (a DiscState)
};

// Device buffers of the learner that kernels read.
// Rows are stored as records of `stride` bytes (a multiple of 16): the bin of
// group g at byte g, and in buffers 1 and 2 the row payload at byte pay_off.
// Buffer 0 holds the rows in their original order; its payload lives in gh0
// (rewritten every tree). Buffers 1 and 2 are the partition ping-pong buffers;
// records carry no row id (the row-to-leaf map comes from a tree walk over
// buffer 0).
struct DevState {
  int num_data, npad, num_groups, num_features, total_bins, form;
  int stride, pay_off;
  // quantized histograms with 32-bit bins (16-bit gradient and hessian sums, no
  // row count): partials are u32 in part_q, pool slots carry no counts, and the
  // search leaves exact_left/right to the partition (-1)
  int hist32, pad_h32;
  uint8_t* rec[3];
  void* gh0;                     // payload of buffer 0: int16 packed (quantized) or float2 (g, h)
  const FeatDev* feats;
  const int* group_bin_start;    // first histogram bin of each group
  double* pool;                  // slots x total_bins x 2 (double g,h or int64 g,h)
  int* pool_cnt;                 // slots x total_bins
  unsigned long long* part_q;    // quantized partial histograms: one per task, total_bins each
  double* part_g;                // float forms: gradient, hessian and count planes
  double* part_h;
  unsigned* part_c;
  int* err_flag;
};

// Segment of rows [begin, begin+count) in buffer buf.
struct Segment {
  int buf, begin, count;
};

// A histogram build over a whole segment. The rows of the items of a launch are
// concatenated: item i owns rows row0 .. row0 + seg.count - 1 (row0 ascending,
// from 0, no empty items).
struct HistItem {
  Segment seg;
  int row0, pad;
};

// One leaf split: its rows move from seg (buffer seg.buf) to dst_buf, in tiles of
// `chunk` rows (tasks first_task .. first_task + ntasks - 1). Left rows fill the
// segment from its start; right rows follow them when left_total (the left
// child's exact count) is known, and fill it from its end when left_total < 0,
// the cursors then giving the children's counts. hist_side (0 left, 1 right, -1
// none): the child whose histogram a histogram-building partition accumulates,
// one partial per tile (task).
struct PartItem {
  Segment seg;
  int dst_buf, split_group, left_total, hist_side;
  int first_task, ntasks, chunk, pad1;
  RuleDev rule;
};

// Partial histograms part_first .. part_first + nblk - 1 to be summed into the
// pool slot `slot` (raw: before fix_histogram).
struct ReduceItem {
  int slot, part_first, nblk, pad;
};

// The folded best split of one leaf: the candidate and its inner feature (-1: none).
struct FoldOut {
  SplitOut o;
  int inner, pad[3];
};

// One search launch item: the smaller child (partials) and the larger one
// (subtraction from the parent slot, or its own partials).
struct ScanItem {
  LeafArgs small, large;
  int has_large, use_subtract, used_off, out_off;
};

// ---------------------------------------------------------------------------
// Device round planner (batched growth without a host round trip per round)
// ---------------------------------------------------------------------------
// Work of one planned round, written by k_plan and read by the round's kernels,
// which are launched with grids sized for the most work a round can hold: blocks
// beyond the planned work exit at once.
struct RoundHdr {
  int npart, ntasks;                              // partition items and their tiles
  int nhist, hist_total, hist_nblocks, hist_rpb;  // histogram items, rows, row blocks, rows per block
  int nscan;                                      // search items (one reduce item each)
  int pad;
};

constexpr int kPlanMaxLeaves = 256;  // num_leaves the device planner takes (more: the host plans)

// A node of the batched growth: a leaf that was or may be expanded. Its children
// are nodes child[0] (left) and child[1] (right) once it is expanded.
struct PlanNode {
  Segment seg;
  int depth, slot, expanded, kids_searched;
  int child[2];
  int num_data, pad;
  double sum_g, sum_h, weight;  // the leaf's sums and output (LeafSplits)
  long long int_sum;
  FoldOut best;                 // its best split (inner < 0: none)
};

// Tree state of the planner, in device memory.
struct PlanState {
  int num_leaves, num_nodes, next_slot, done;
  int ncommit, nabsorb, rounds, pad;  // rounds: with work; pad: set on a table overflow
  int search_rounds, pad2;            // rounds with a search
  long long hist_rows, part_rows;     // rows the round kernels processed
  RoundHdr hdr;
  int absorb[2 * kPlanMaxLeaves];     // search item k of the last round: nodes of its smaller and larger side
  int leaf_node[kPlanMaxLeaves];      // node of every leaf
  int commit[2 * kPlanMaxLeaves];     // committed split c: (leaf, node), in the reference's order
};

// A split whose children's counts come from the partition (left_total < 0): the
// children as search arguments (sums, outputs, integer sums; num_data and slot are
// filled in from the counts), the children's depth, and the slots the pair
// takes when searched: new_slot for the smaller child, parent_slot (the split
// leaf's) for the larger.
struct ChildSpec {
  LeafArgs kid[2];
  int depth, new_slot, parent_slot, pad;
};
// Rules and tiling of the items derived from the partition. part_hist: the
// partition built the histogram of each split's hist_side child (one partial per
// tile), which is then the search item's built side, and no histogram items are
// planned.
struct ChildParams {
  int max_depth, min_data_in_leaf, hist_min_rows, hist_max_blocks;
  int rows_per_flush, max_partials, F, part_hist;
};
// Row blocks of a histogram build of `rows` rows: about rows / min_rows, at most
// max_blocks, but a small build spreads over up to kHistSpreadBlocks blocks of at
// least kHistSpreadRows rows so that it occupies the chip, and no block takes
// more than rows_per_flush rows. Grows with rows (a grid sized for a bound holds
// every build within it).
constexpr long long kHistSpreadBlocks = 256, kHistSpreadRows = 512;
BK_HD inline long long hist_row_blocks(long long rows, int min_rows, int max_blocks, int rows_per_flush) {
  if (rows <= 0) return 0;
  long long nb = rows / min_rows;
  if (nb < 1) nb = 1;
  if (nb > max_blocks) nb = max_blocks;
  long long spread = (rows + kHistSpreadRows - 1) / kHistSpreadRows;
  if (spread > kHistSpreadBlocks) spread = kHistSpreadBlocks;
  if (nb < spread) nb = spread;
  const long long flush = (rows + rows_per_flush - 1) / rows_per_flush;
  return nb < flush ? flush : nb;
}
constexpr int kChildThreads = 1024;  // splits one derivation takes (one thread each)
constexpr int kPartDeriveMax = 512;  // splits the partition's last block derives (one thread each)
// The derivation run by the partition's last block (on != 0; at most 512 splits),
// with what launch_children takes: specs, params, the item tables, hdr, the
// host-mapped counts, and a zeroed ticket word the partition's blocks count on
// (the last one resets it).
struct DeriveArgs {
  const ChildSpec* specs;
  HistItem* hitems;
  ScanItem* sitems;
  ReduceItem* ritems;
  RoundHdr* hdr;
  int* counts;
  unsigned* ticket;
  ChildParams cp;
  int on, pad;
};

struct PlanParams {
  int num_leaves, F, max_depth, min_data_in_leaf;
  int hist_min_rows, hist_max_blocks, rows_per_flush, part_tile;
  int max_nodes, num_slots, max_partials, pad1;
};

// ---------------------------------------------------------------------------
// Device-resident trees: the discretizer's scales, the root sums and the tree's
// final tables are computed on the device, so a tree runs without a host wait
// until its end.
// ---------------------------------------------------------------------------
// The discretizer's state, kept across trees (a refused discretization keeps the
// previous one, serial.rs:692-703). The search reads gs, hs, h0 (ScanParams::dyn).
struct DiscState {
  double gs, hs, h0;
  double inv_g, inv_h;
  double root_g, root_h;      // root sums of the last accepted discretization
  long long int_g, int_h;     // and their exact integer sums
  int ok, start, pad0, pad1;  // ok: this tree's discretization was accepted
};
// Scalars of the root set-up.
struct RootArgs {
  int num_data, min_data_in_leaf, quantized, const_hess;
  int nparts, bins, pad0, pad1;  // nparts: f64 partial sums of g and h (float forms)
  double l1, l2, max_delta_step;
};
// Scales from max |g|, max h (zeroed for the next tree) and the noise rotation
// drawn by the host (*start_word).
void launch_disc_prep(unsigned* d_max4, const int* start_word, int bins, int const_hess, DiscState* d, void* stream);
// The root's sums (exact integer sums times the scales when quantized; the
// partial sums in order otherwise), then the planner's initial state, the root
// node, the tree's feature sample and the root's histogram, reduce and search
// items with the round header.
void launch_root_prep(DiscState* d, const long long* sums3, const double* partials, const float* hessians,
                      const RootArgs& ra, const PlanParams& pp, PlanState* st, PlanNode* nodes, uint8_t* node_flags,
                      const int* bytree_src, int* bytree, HistItem* hitems, ScanItem* sitems, ReduceItem* ritems,
                      int* used, int* err, void* stream);
// Kernel timing. While a source is set (per host thread), every launch below takes
// a fresh (start, stop) pair of timing events (hipEvent_t as void*) from it and is
// launched with hipExtLaunchKernel, which stamps the kernel's own begin and end
// into the pair: no marker packets are queued between kernels. The source
// returns false to launch without events.
typedef bool (*EventPairSource)(void* ctx, void** start, void** stop);
void set_event_source(EventPairSource fn, void* ctx);

// Launchers (kernels.hip). All run on `stream` (a hipStream_t passed as void*).
// Records of buffer 0 from column-major bins (group g at cols + g * col_stride).
void launch_to_records(const DevState& s, const uint8_t* cols, size_t col_stride, void* stream);
size_t hist_lds_bytes(int form, int hist32, int max_tile_bins);
// Block b covers concatenated rows [b * rows_per_block, (b + 1) * rows_per_block)
// and writes one partial histogram per item it touches, from slot d_block_slot[b] on.
// With hdr (a planned round): nitems, total_rows, nblocks and rows_per_block are
// read from it, the arguments give the grid (nblocks: the most row blocks), and
// the partial of (block b, item i) is slot b + i (d_block_slot unused).
// Group tiles: tile t covers groups d_tile_first[t] .. d_tile_first[t + 1] - 1
// (max_tile_bins bins at most); nsub LDS copies of a tile's histogram, wave w
// adding to copy w % nsub.
void launch_hist(const DevState& s, const HistItem* d_items, int nitems, int total_rows, int nblocks,
                 int rows_per_block, const int* d_block_slot, const int* d_group_used, int max_tile_bins,
                 const int* d_tile_first, int ntiles, int rows_per_flush, void* stream,
                 const RoundHdr* hdr = nullptr, int nsub = 1);
// Rows per partition tile for a record stride (8192 16-byte chunks per tile),
// and per short tile (2048 chunks), which rounds of few rows take to spread over
// more blocks.
int partition_tile_rows(int stride);
int partition_tile_rows_small(int stride);
// Single-pass partition of every item (row order kept within a tile, not across
// tiles): tasks are the items' tiles in order (chunk == partition_tile_rows, or
// partition_tile_rows_small with small_tiles).
// d_cursors: 2 banks x max_items x (left, right) u32, zeroed before the first
// launch; a launch uses bank (epoch & 1) and zeroes the other, so consecutive
// launches on one stream alternate epoch parity.
// With hdr: nitems and ntasks are read from it; the arguments give the grid.
// derive (may be null): the last block to finish derives the round's items as
// launch_children would (left_total < 0 items, no hdr, nitems <= 512).
void launch_partition(const DevState& s, const PartItem* d_items, int nitems, int ntasks, unsigned* d_cursors,
                      int max_items, unsigned epoch, void* stream, const RoundHdr* hdr = nullptr,
                      const DeriveArgs* derive = nullptr, bool small_tiles = false);
// Whether the partition can build histograms (quantized 32-bit bins): the
// record's chunks divide the block's threads evenly and every bin of the record
// fits in LDS beside the partition's own tables.
bool partition_hist_fits(const DevState& s);
// The partition with histograms: for every item with hist_side >= 0, tile t also
// adds the rows it routes to that side into an LDS histogram of the whole record
// (groups with d_group_used[g] == 0 skipped) and writes it to partial slot t of
// part_q (u32 words, total_bins per slot). Tiles are at most 4096 rows, within a
// 32-bit bin's exact range. Quantized learners with hist32 only.
void launch_partition_hist(const DevState& s, const PartItem* d_items, int nitems, int ntasks, unsigned* d_cursors,
                           int max_items, unsigned epoch, const int* d_group_used, void* stream,
                           const DeriveArgs* derive = nullptr);
// After a partition whose items have left_total < 0: per split j, the children's
// exact counts from bank (epoch & 1) of the cursors (to counts[2j], counts[2j + 1],
// host-mapped memory), and the round's items as the host plans them for known
// counts: a search item for every split whose children are searchable
// (before_find_best_split's max_depth and min_data_in_leaf rules), a histogram
// item for its smaller child and its reduce item, the histogram tiling, and hdr
// (nhist, hist_total, hist_nblocks, hist_rpb, nscan) for the launches that follow.
// counts[2 * nitems] is a status word: 0, or bit 1 (counts that do not add up to
// the segment) / bit 2 (a searchable split whose histogram the partition did not
// build) / bit 5 (a tiling beyond max_partials), and then no work is planned.
// nitems <= kChildThreads.
void launch_children(const PartItem* d_pitems, int nitems, const unsigned* d_cursors, int max_items, unsigned epoch,
                     const ChildSpec* d_specs, const ChildParams& cp, HistItem* hitems, ScanItem* sitems,
                     ReduceItem* ritems, RoundHdr* hdr, int* counts, void* stream);
// Sums the partial histograms of every item into its pool slot; the search then
// reads the raw histogram of the smaller leaf (and of a rebuilt larger leaf)
// from its slot.
// With hdr the grid covers min(nitems, hdr_grid_cap) items at a time (the
// device decides how many of the nitems there are).
void launch_reduce(const DevState& s, const ReduceItem* d_items, int nitems, void* stream,
                   const RoundHdr* hdr = nullptr, int hdr_grid_cap = 64);
// Split search of every (feature, item), then the fold of the best candidate of
// every (item, side) into d_best[2 * item + side] and of every feature's
// splittable flag into d_flags[(2 * item + side) * F + f].
// stamp (may be null): the fold's last wave to finish writes seq to *stamp (a
// host-mapped word), after every fold result is visible to the host; ticket: a
// zeroed word the fold's waves count on (reset by the last).
void launch_scan(const DevState& s, const ScanItem* d_items, int nitems, const ScanParams& sp, const int* d_used,
                 SplitOut* d_out, FoldOut* d_best, uint8_t* d_flags, void* stream, const RoundHdr* hdr = nullptr,
                 unsigned* stamp = nullptr, unsigned seq = 0, unsigned* ticket = nullptr, int hdr_grid_cap = 64);
// One round of planning: absorbs the last round's folded splits into their
// nodes, replays the reference's best-first commits while the best leaf is
// expanded, picks the next frontier (the best unexpanded leaves with positive
// gain, at most the remaining leaf budget) and writes the round's items and
// st->hdr. A complete tree sets st->done and plans no work.
// The planner's initial state, root node, tree feature sample and root flags,
// copied from an uploaded staging area in one launch.
void launch_plan_init(const PlanState* s_state, const PlanNode* s_node, const int* s_bytree, const uint8_t* s_flags,
                      PlanState* st, PlanNode* node0, int* bytree, uint8_t* flags0, int F, void* stream);
void launch_plan(PlanState* st, PlanNode* nodes, uint8_t* node_flags, const FoldOut* fold, const uint8_t* fold_flags,
                 const FeatDev* feats, const int* bytree, const PlanParams& pp, PartItem* pitems, HistItem* hitems,
                 ScanItem* sitems, ReduceItem* ritems, int* used, void* stream);
// max |g|, max h and a bad-value flag into d_out4 (zeroed before); d_zero4 (may
// be null): four int64 zeroed on the way, for the discretizer's sums.
void launch_max_abs(const float* g, const float* h, int n, unsigned* d_out4, long long* d_zero4, void* stream);
// d_state (may be null): start and inverse scales from it; nothing is written when
// its discretization was refused.
void launch_discretize(const float* g, const float* h, const double* gnoise, const double* hnoise, int n, int start,
                       double inv_g, double inv_h, int const_hess, int stochastic, int16_t* packed,
                       long long* d_sums3, void* stream, const DiscState* d_state = nullptr);
void launch_pack_f32(const float* g, const float* h, int n, int const_hess, float* gh, void* stream);
// One internal node of the last tree for the row-to-leaf traversal: the route
// of every bin value under the rule its split applied in the partition (bit b:
// bin b goes left, filled with route_left), the group whose bin it reads, and
// the children (>= 0: an internal node, < 0: ~leaf).
struct NodeDev {
  uint32_t left_bins[8];
  int group, left, right, pad;
};
// A committed split as the host hands it over: its rule, group and children.
struct NodeRule {
  RuleDev rule;
  int group, left, right, pad;
};
// Leaf of every row: the route bitmaps of the nodes are filled from their rules
// (one lane per bin value), then each row walks the tree from node 0 over its
// record in buffer 0 (original row order). num_nodes == 0: every row is in leaf 0.
// bag_bits (bit r: row r was in the tree's bag; null: every row was): a row
// outside the bag gets ~leaf.
// d_num_nodes (may be null): the node count on the device, num_nodes then the most.
void launch_leaf_of_tree(const DevState& s, const NodeRule* d_rules, NodeDev* d_nodes, int num_nodes,
                         const uint32_t* bag_bits, int32_t* leaf_of, void* stream, const int* d_num_nodes = nullptr);
// The committed tree's walk rules (rule, group, children: a later commit of the
// same leaf index, or ~leaf), the leaf segments and the leaves by row count
// (descending, for the renewal chains) for num_leaves leaf slots; *d_count =
// committed splits.
void launch_finalize(const PlanState* st, const PlanNode* nodes, const FeatDev* feats, NodeRule* rules, int* d_count,
                     Segment* segs, int* order, int num_leaves, void* stream);
// Leaf renewal sums: per-block partials over leaf_of, the certificate per leaf,
// then the sequential chain for flagged leaves (d_flags == nullptr runs the
// chain for every leaf).
int renew_parts(int n);
// The walk (staged records) with the leaf renewal's per-block partials
// (launch_renew_par's part_sums and part_meta, renew_parts(num_data) blocks) in
// the same pass over the rows. Returns false, launching nothing, when the
// records, nodes and accumulators do not fit the staged layout.
bool launch_leaf_of_tree_renew(const DevState& s, const NodeRule* d_rules, NodeDev* d_nodes, int num_nodes,
                               const uint32_t* bag_bits, int32_t* leaf_of, const float* g, const float* h,
                               int num_leaves, double* part_sums, int* part_meta, void* stream);
size_t renew_par_lds(int num_leaves);
void launch_renew_par(int n, const int32_t* leaf_of, const float* g, const float* h, int num_leaves,
                      double* part_sums, int* part_meta, void* stream);
void launch_renew_fix(const Segment* d_segs, int num_leaves, int nparts, const double* part_sums, const int* part_meta,
                      double* d_sums, int* d_flags, void* stream);
// One piece: the values of one leaf among rows [p * piece_rows, (p + 1) * piece_rows)
// of one component: their sum (exact where the piece's own certificate holds), the
// sum of their magnitudes and the lowest set bit's exponent over them (INT_MAX: no
// nonzero value, INT_MIN: a NaN or infinity).
struct RenewPiece {
  double tg, th, ag, ah;
  int eg, eh;
};
// Piece storage for the piecewise chain: num_pieces x num_leaves entries (piece-major).
// A piece added from a running sum s inside one binade [2^eb, 2^(eb+1)) of |s| (the
// sum stays there throughout the piece): every addition rounds to a multiple of
// u = 2^(eb-52), to the nearest one whatever s is, except a tie, which goes to the
// even multiple. The piece adds u times s0 when s/u is even, u times s1 when odd.
struct RenewBinade {
  long long s0, s1;
  int eb;  // INT_MIN: no summary
  int work;  // the piece's entry in the list (its sub-range summaries), -1: none
};
// The same tests for one sixteenth of a listed piece's rows: the exact sum, the
// sum of magnitudes and the lowest set bit of the leaf's values there, and the
// binade summary for the binade the guessed running sum has at its start.
constexpr int kRenewParts = 16;
struct RenewPart {
  double t, a;
  long long s0, s1;
  int e, eb;
};
// Piece storage for the piecewise chain: num_pieces x num_leaves entries (piece-major),
// the binade summaries per (piece, leaf, gradients/hessians), and the list of pieces
// whose summaries are computed (piece, leaf, component, binade, guessed running sum
// at the piece's start; a counter first), kRenewParts sub-range summaries per entry.
struct RenewWork {
  int piece, leaf, comp, eb;
  double s;
};
struct RenewPieces {
  RenewPiece* pieces = nullptr;
  RenewBinade* binade = nullptr;
  RenewWork* work = nullptr;
  RenewPart* parts = nullptr;
  int* work_count = nullptr;
  int piece_rows = 0, num_pieces = 0, work_cap = 0;
};
constexpr int kRenewPieceRows = 4096;  // rows per piece at the least
// Rows per piece for n rows and num_leaves leaves: kRenewPieceRows, doubled until the
// pieces take at most max_bytes.
int renew_piece_rows(long long n, int num_leaves, size_t max_bytes);
size_t renew_pieces_lds(int num_leaves);
// The sequential f64 sums of the flagged leaves (d_flags bit 0: gradients, bit 1:
// hessians; null: both of every leaf) in ascending row order, into d_sums[2 * leaf + c].
// With pieces (pieces.pieces != null) the chain runs piece by piece: a piece whose
// every partial sum, started from the chain's running sum, is exactly representable
// is added as one exact sum, the others value by value. Without, one block per leaf
// adds every value.
void launch_renew_sums(int n, const int32_t* leaf_of, const int* d_order, int num_leaves, const float* g,
                       const float* h, const int* d_flags, double* d_sums, const RenewPieces& pieces, void* stream);
// score[row] += d_leaf_values[leaf] over the bag's rows (out_of_bag = 0) or the rows
// outside the bag (out_of_bag = 1).
void launch_add_score_rows(int n, const int32_t* leaf_of, const double* d_leaf_values, int num_leaves, int out_of_bag,
                           double* score, void* stream);
// Per-block f64 sums of g and h over rows[0..n) (null rows: rows 0 .. n-1);
// returns the number of blocks (partials at d_partials[2b], [2b+1]).
int launch_sum_gh(const float* g, const float* h, const int32_t* rows, int n, int num_data, double* d_partials,
                  void* stream);
// The bag's root segment: row bag[j] of buffer 0 (with its payload) becomes row j
// of dst_buf; bag rows are marked in bag_bits (zeroed before); quantized sums of
// the bag's packed pairs are added to sums[0..2] (int g, int h, |int g|). Rows
// not ascending or outside the dataset set err bit 8.
void launch_gather_bag(const DevState& s, const int32_t* bag, int nbag, int dst_buf, uint32_t* bag_bits,
                       long long* sums, void* stream);
// Copies n bytes with a kernel on `stream`. Either side may be pinned host memory
// (its device address): small transfers then stay on the compute queue instead
// of a DMA engine, whose completion the queue would wait for.
void launch_copy(void* dst, const void* src, size_t n, void* stream);
// Writes seq into *word (device address of pinned host memory) after the work
// enqueued on `stream` before it.
void launch_stamp(unsigned* word, unsigned seq, void* stream);
// Keeps `stream` busy for `ticks` of the device's wall clock (timing runs only):
// a timed launch submitted while it runs is stamped by its events when it is
// already queued, not when the host starts submitting it.
void launch_hold(unsigned long long ticks, void* stream);
void launch_softmax_gradients(const double* score, const float* label, int n, int num_class, float* g, float* h,
                              void* stream);
// LambdaRank tables and switches (objective/rank.rs:183-240).
struct RankDev {
  const int* bounds;          // num_queries + 1 row offsets
  const double* inv_max_dcg;  // per query
  const double* label_gain;   // per label value
  const double* discount;     // per rank position
  const double* sigmoid_table;
  double sigmoid, min_in, max_in, table_factor;
  int num_queries, num_gain, table_bins, truncation, norm, pad;
};
constexpr int kRankMaxDocs = 2048;  // documents of one query held in LDS by the gradient kernel
void launch_lambdarank_gradients(const double* score, const float* label, const RankDev& r, float* g, float* h,
                                 void* stream);
// Binary log-loss (objective/binary.rs:71-90): label weights lw_neg / lw_pos,
// optional sample weights (null: none).
// Scratch memory of the allocation-free entry point: a file-scope device array
// of kScratchBytes in kernels.hip. scratch_device_ptr() returns its address on
// the current device (looked up once per device with hipGetSymbolAddress); every
// caller on one device shares it.
constexpr size_t kScratchBytes = (size_t)16 << 20;
void* scratch_device_ptr();
// hist[2 * b + 1] *= (double)hessians[0] for b < total_bins (a constant hessian
// counts rows; construct_histograms scales them, serial.rs:982-988).
void launch_scale_const_hess(double* hist, int total_bins, const float* hessians, void* stream);
void launch_binary_gradients(const double* score, const float* label, const float* weights, int n, double sigmoid,
                             double lw_neg, double lw_pos, float* g, float* h, void* stream);

}  // namespace bk
