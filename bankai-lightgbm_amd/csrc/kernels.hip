// kernels.hip — gfx950 kernels of the histogram tree learner.
//
// Data layout in HBM (one learner):
//   rec[0]         row records in original order: group g's bin at byte g of a
//                  `stride`-byte record (read only after creation); the payload
//                  of these rows lives in gh0 (int16 packed int8 pair when
//                  quantized, float2 (gradient, hessian) otherwise)
//   rec[1..2]      ping-pong copies in partition order, payload inside the record:
//                  every leaf's rows are a contiguous segment of one buffer, so
//                  histogram and partition passes stream whole records (no
//                  per-row gathers)
//   leaf_of        leaf of every row of the last tree (tree walk over rec[0])
//   pool           one histogram per leaf slot: (g, h) per bin, plus an exact
//                  per-bin row count (pool_cnt) used to size partitions in advance
//
// All f64 arithmetic that must match the reference bit for bit is written
// without contraction (the build passes -ffp-contract=off), matching the
// reference's unfused madd on x86_64 (crates/lgbm-core/src/utils/common.rs:207-216).

#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_set>

#include <algorithm>
#include <climits>
#include <mutex>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include <hip/hip_ext.h>

#include "kernels.h"

#define BK_HIP(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " in " #x); \
  } while (0)

namespace bk {

namespace {
thread_local EventPairSource t_event_source = nullptr;
thread_local void* t_event_ctx = nullptr;

template <typename F, typename... Args>
void launch(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, Args... args) {
  void *start = nullptr, *stop = nullptr;
  if (t_event_source && t_event_source(t_event_ctx, &start, &stop))
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, st, (hipEvent_t)start, (hipEvent_t)stop, 0u,
                          args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}
}  // namespace

void set_event_source(EventPairSource fn, void* ctx) {
  t_event_source = fn;
  t_event_ctx = ctx;
}

namespace {

constexpr double kEps = (double)1e-15f;
constexpr double kMinScore = -__builtin_huge_val();

// ---------------------------------------------------------------------------
// Packed quantized LDS accumulator: one 64-bit add per (row, group).
//   bits 43..63: gradient sum (signed), 22..42: hessian sum, 0..21: row count.
// Exact while a flush covers fewer than 2^20 / max|q| rows (host enforces).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long qpack(int16_t p) {
  long long g = (long long)(int8_t)(p >> 8);
  unsigned long long h = (unsigned long long)(uint8_t)(p & 0xff);
  return ((unsigned long long)g << 43) + (h << 22) + 1ull;
}
__device__ __forceinline__ void qunpack(unsigned long long s, long long* g, long long* h, long long* c) {
  *c = (long long)(s & ((1ull << 22) - 1ull));
  *h = (long long)((s >> 22) & ((1ull << 21) - 1ull));
  *g = ((long long)s) >> 43;
}

// 32-bit form: bits 16..31 gradient sum (signed), 0..15 hessian sum, no count.
// A sum of n rows is exact while n * max h < 2^16 and n * max |g| < 2^15: the
// hessian field never carries and the two's-complement word is G * 2^16 + H.
__device__ __forceinline__ unsigned qpack32(int16_t p) {
  const int g = (int)(int8_t)(p >> 8);
  const unsigned h = (unsigned)(uint8_t)(p & 0xff);
  return ((unsigned)g << 16) + h;
}
__device__ __forceinline__ void qunpack32(unsigned w, long long* g, long long* h) {
  *h = (long long)(w & 0xffffu);
  *g = (long long)(((int)w) >> 16);
}

__device__ __forceinline__ int sat_i32_dev(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return (int)(-2147483647 - 1);
  return (int)x;
}
__device__ __forceinline__ int8_t sat_i8_dev(double x) {
  if (x != x) return 0;
  if (x >= 127.0) return 127;
  if (x <= -128.0) return -128;
  return (int8_t)(int)x;
}

// Block-wide exclusive scan of two ints over NT threads.
template <int NT>
__device__ __forceinline__ void block_scan2(int a, int b, int* pa, int* pb, int* ta, int* tb, int* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int xa = a, xb = b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int ya = __shfl_up(xa, o, 64), yb = __shfl_up(xb, o, 64);
    if (lane >= o) {
      xa += ya;
      xb += yb;
    }
  }
  if (lane == 63) {
    scratch[wave] = xa;
    scratch[16 + wave] = xb;
  }
  __syncthreads();
  int oa = 0, ob = 0, sa = 0, sb = 0;
  constexpr int nw = NT / 64;
#pragma unroll
  for (int w = 0; w < nw; ++w) {
    if (w < wave) {
      oa += scratch[w];
      ob += scratch[16 + w];
    }
    sa += scratch[w];
    sb += scratch[16 + w];
  }
  *pa = oa + xa - a;
  *pb = ob + xb - b;
  *ta = sa;
  *tb = sb;
  __syncthreads();
}

// Block-wide sum over NT threads (result in every thread).
template <int NT>
__device__ __forceinline__ int block_sum(int v, int* scratch) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += scratch[w];
  __syncthreads();
  return t;
}

// Block-wide minimum over NT threads (result in every thread).
template <int NT>
__device__ __forceinline__ int block_min(int v, int* scratch) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_down(v, o, 64));
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = INT_MAX;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t = min(t, scratch[w]);
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------------------
// Histogram accumulators in LDS for one group tile.
// ---------------------------------------------------------------------------
template <int FORM>
struct LdsHist;

template <>
struct LdsHist<FORM_QUANT> {
  unsigned long long* q;
  static constexpr int kBytesPerBin = 8;
  __device__ void bind(unsigned char* smem, int nb) { q = (unsigned long long*)smem; }
  __device__ void zero(int nb) {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) q[i] = 0ull;
  }
  __device__ __forceinline__ void add(int b, unsigned long long v) { atomicAdd(&q[b], v); }
  __device__ void flush(const DevState& s, int blk, int bin_lo, int nb, bool first) {
    unsigned long long* dst = s.part_q + (size_t)blk * s.total_bins + bin_lo;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) dst[i] = first ? q[i] : dst[i] + q[i];
  }
};

// Kind of k_hist / k_reduce for the quantized payload with 32-bit bins.
constexpr int kHistQ32 = 3;

template <>
struct LdsHist<kHistQ32> {
  unsigned* q;
  static constexpr int kBytesPerBin = 4;
  __device__ void bind(unsigned char* smem, int nb) { q = (unsigned*)smem; }
  __device__ void zero(int nb) {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) q[i] = 0u;
  }
  __device__ __forceinline__ void add(int b, unsigned v) { atomicAdd(&q[b], v); }
  // partials are u32 words in part_q's memory, total_bins per slot
  __device__ void flush(const DevState& s, int blk, int bin_lo, int nb, bool first) {
    unsigned* dst = (unsigned*)s.part_q + (size_t)blk * s.total_bins + bin_lo;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) dst[i] = first ? q[i] : dst[i] + q[i];
  }
};

template <>
struct LdsHist<FORM_F64> {
  double* g;
  double* h;
  unsigned* c;
  static constexpr int kBytesPerBin = 20;
  __device__ void bind(unsigned char* smem, int nb) {
    g = (double*)smem;
    h = g + nb;
    c = (unsigned*)(h + nb);
  }
  __device__ void zero(int nb) {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      g[i] = 0.0;
      h[i] = 0.0;
      c[i] = 0u;
    }
  }
  __device__ __forceinline__ void add(int b, float gg, float hh) {
    atomicAdd(&g[b], (double)gg);
    atomicAdd(&h[b], (double)hh);
    atomicAdd(&c[b], 1u);
  }
  __device__ void flush(const DevState& s, int blk, int bin_lo, int nb, bool first) {
    size_t o = (size_t)blk * s.total_bins + bin_lo;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      s.part_g[o + i] = first ? g[i] : s.part_g[o + i] + g[i];
      s.part_h[o + i] = first ? h[i] : s.part_h[o + i] + h[i];
      s.part_c[o + i] = first ? c[i] : s.part_c[o + i] + c[i];
    }
  }
};

// f32 sums with a two-sum error term per bin: each atomic returns the old value,
// so the rounding error of that exact addition is recovered and accumulated.
template <>
struct LdsHist<FORM_F32C> {
  float* g;
  float* h;
  float* eg;
  float* eh;
  unsigned* c;
  static constexpr int kBytesPerBin = 20;
  __device__ void bind(unsigned char* smem, int nb) {
    g = (float*)smem;
    h = g + nb;
    eg = h + nb;
    eh = eg + nb;
    c = (unsigned*)(eh + nb);
  }
  __device__ void zero(int nb) {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      g[i] = 0.0f;
      h[i] = 0.0f;
      eg[i] = 0.0f;
      eh[i] = 0.0f;
      c[i] = 0u;
    }
  }
  __device__ __forceinline__ static float two_sum_err(float a, float b) {
    float s = a + b;
    float bb = s - a;
    return (a - (s - bb)) + (b - bb);
  }
  __device__ __forceinline__ void add(int b, float gg, float hh) {
    float og = atomicAdd(&g[b], gg);
    float oh = atomicAdd(&h[b], hh);
    float e1 = two_sum_err(og, gg), e2 = two_sum_err(oh, hh);
    if (e1 != 0.0f) atomicAdd(&eg[b], e1);
    if (e2 != 0.0f) atomicAdd(&eh[b], e2);
    atomicAdd(&c[b], 1u);
  }
  __device__ void flush(const DevState& s, int blk, int bin_lo, int nb, bool first) {
    size_t o = (size_t)blk * s.total_bins + bin_lo;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      double vg = (double)g[i] + (double)eg[i];
      double vh = (double)h[i] + (double)eh[i];
      s.part_g[o + i] = first ? vg : s.part_g[o + i] + vg;
      s.part_h[o + i] = first ? vh : s.part_h[o + i] + vh;
      s.part_c[o + i] = first ? c[i] : s.part_c[o + i] + c[i];
    }
  }
};

// ---------------------------------------------------------------------------
// Row payload: packed int16 (quantized) or float2 (g, h). Buffer 0 keeps it in
// gh0; buffers 1 and 2 carry it inside the record at pay_off.
// ---------------------------------------------------------------------------
template <int FORM>
struct Pay {
  using T = float2;
};
template <>
struct Pay<FORM_QUANT> {
  using T = int16_t;
};
template <>
struct Pay<kHistQ32> {
  using T = int16_t;
};

// Buffer pointers by buffer number. Selecting with branches keeps the kernel
// argument in constant memory; indexing its arrays with a runtime value would
// copy the whole argument into per-lane scratch.
__device__ __forceinline__ uint8_t* rec_of(const DevState& s, int b) {
  return b == 0 ? s.rec[0] : (b == 1 ? s.rec[1] : s.rec[2]);
}

template <int FORM>
__device__ __forceinline__ typename Pay<FORM>::T load_payload(const DevState& s, int buf, long long pos) {
  using T = typename Pay<FORM>::T;
  if (buf == 0) return ((const T*)s.gh0)[pos];
  return *(const T*)(rec_of(s, buf) + (size_t)pos * s.stride + s.pay_off);
}

// Streaming 16-byte accesses: the records a partition moves are read once and
// written once per launch, so they bypass the caches' normal allocation.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 load_stream(const uint8_t* p) {
  const u32x4 x = __builtin_nontemporal_load((const u32x4*)p);
  return make_uint4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void store_stream(uint8_t* p, uint4 v) {
  const u32x4 x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, (u32x4*)p);
}

__device__ __forceinline__ uint32_t lane_word(const uint4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ void set_word(uint4& v, int i, uint32_t w) {
  if (i == 0) v.x = w;
  else if (i == 1) v.y = w;
  else if (i == 2) v.z = w;
  else v.w = w;
}

// writes the payload into a 16-byte record chunk; `off` is its byte offset in the chunk
__device__ __forceinline__ void patch_payload(uint4& v, int off, int16_t p) {
  const int i = off >> 2, sh = (off & 3) * 8;
  set_word(v, i, (lane_word(v, i) & ~(0xffffu << sh)) | ((uint32_t)(uint16_t)p << sh));
}
__device__ __forceinline__ void patch_payload(uint4& v, int off, float2 p) {
  const int i = off >> 2;
  set_word(v, i, __float_as_uint(p.x));
  set_word(v, i + 1, __float_as_uint(p.y));
}

__device__ __forceinline__ void hist_add(LdsHist<FORM_QUANT>& L, int b, int16_t p) { L.add(b, qpack(p)); }
__device__ __forceinline__ void hist_add(LdsHist<kHistQ32>& L, int b, int16_t p) { L.add(b, qpack32(p)); }
template <int FORM>
__device__ __forceinline__ void hist_add(LdsHist<FORM>& L, int b, float2 p) {
  L.add(b, p.x, p.y);
}

template <int FORM>
__device__ __forceinline__ int hist_bytes_of(int nb) {
  return (nb * LdsHist<FORM>::kBytesPerBin + 15) & ~15;
}

// group tile [g0, g1) -> per-group first bin relative to the tile, -1 if skipped
__device__ __forceinline__ void init_gbs(const DevState& s, int* gbs, int g0, int g1, int bin_lo,
                                         const int* __restrict__ group_used) {
  for (int g = threadIdx.x; g < s.num_groups; g += blockDim.x)
    gbs[g] = (g >= g0 && g < g1 && group_used[g]) ? s.group_bin_start[g] - bin_lo : -1;
}

// Item owning task t of a launch: items are listed with ascending first_task,
// item i owning tasks first_task .. first_task + ntasks - 1.
template <typename Item>
__device__ __forceinline__ int find_item(const Item* __restrict__ items, int nitems, int t) {
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].first_task <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Adds one 16-byte chunk of a row with the chunk's group bin offsets in
// registers (off[j] < 0: group outside the tile, unused, or padding).
// The quantized payload of the row whose chunk v this lane holds, when the row's
// nc chunks sit in consecutive lanes of one wave and one of them (pay_lane in the
// row's lanes) holds the payload at byte pay_in: each lane takes the payload
// word out of its own chunk and reads the payload lane's. Every lane of the wave
// calls it.
template <int FORM>
__device__ __forceinline__ typename Pay<FORM>::T row_payload(const uint4& v, int pay_in, int nc, int pay_lane) {
  const int lane = (int)threadIdx.x & 63;
  const int src = lane - lane % nc + pay_lane;
  static_assert(FORM == FORM_QUANT || FORM == kHistQ32, "2-byte payloads");
  const int w = (int)(int16_t)((lane_word(v, pay_in >> 2) >> (8 * (pay_in & 3))) & 0xffffu);
  return (int16_t)__shfl(w, src, 64);
}

template <int FORM>
__device__ __forceinline__ void hist_chunk_regs(LdsHist<FORM>& L, const uint4& v, const int (&off)[16],
                                                typename Pay<FORM>::T p) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (off[j] >= 0) hist_add(L, off[j] + (int)((lane_word(v, j >> 2) >> (8 * (j & 3))) & 0xffu), p);
}

// ---------------------------------------------------------------------------
// Histogram of whole segments: the root, the smaller child of every split (its
// rows are contiguous in the partition buffer), or a leaf whose sibling cannot
// be subtracted. The rows of all items of a launch are concatenated (item i
// owns rows row0 .. row0 + count - 1) and block b takes rows
// [b * rows_per_block, (b + 1) * rows_per_block), so a block covers one item
// or the end of one and the start of the next. It writes one partial
// histogram per item it touches, to slots block_slot[b], block_slot[b] + 1, ...;
// an item's partials are therefore consecutive slots. Grid: (blocks, group
// tiles). Thread i always reads chunk i % nc of the tile's chunk range, so its
// 16 bin offsets stay in registers; consecutive threads read consecutive
// 16-byte chunks.
// ---------------------------------------------------------------------------
constexpr int kHistThreads = 512;
constexpr int kXcds = 8;         // accelerator complexes of an MI355X, each with its own L2
constexpr int kHistRows = 2048;  // rows per payload staging tile

template <int FORM>
__global__ void __launch_bounds__(kHistThreads) k_hist(DevState s, const HistItem* __restrict__ items, int nitems,
                                                       int total_rows, int rows_per_block,
                                                       const int* __restrict__ block_slot,
                                                       const int* __restrict__ group_used,
                                                       const int* __restrict__ tile_first, int rows_per_flush,
                                                       int nblocks, int ntiles, int nsub,
                                                       const RoundHdr* __restrict__ hdr) {
  extern __shared__ __align__(16) unsigned char smem[];
  using PT = typename Pay<FORM>::T;
  if (hdr) {  // a planned round: the work comes from the planner, the grid covers the most it can hold
    nitems = hdr->nhist;
    total_rows = hdr->hist_total;
    rows_per_block = hdr->hist_rpb;
    nblocks = hdr->hist_nblocks;
  }
  // Workgroups go to the XCDs round robin (workgroup i to XCD i % 8). The
  // group tiles of one row block are given workgroups i, i + 8, ... so they
  // run on one XCD one after the other and the later tiles read the rows from
  // that XCD's L2 instead of HBM.
  const int xcd = (int)blockIdx.x % kXcds, k = (int)blockIdx.x / kXcds;
  const int rb = (k / ntiles) * kXcds + xcd, tile = k % ntiles;
  if (rb >= nblocks) return;  // block-uniform
  if (total_rows <= 0) return;  // block-uniform
  const long long start = (long long)rb * rows_per_block;
  const long long end = min(start + (long long)rows_per_block, (long long)total_rows);
  if (start >= end) return;  // block-uniform
  const int g0 = tile_first[tile], g1 = tile_first[tile + 1];
  const int bin_lo = s.group_bin_start[g0], nb = s.group_bin_start[g1] - bin_lo;
  // nsub copies of the tile's histogram; wave w adds to copy w % nsub, and a
  // flush adds the copies up (fewer same-bin atomics per copy)
  const int hb = hist_bytes_of<FORM>(nb);
  LdsHist<FORM> L;
  L.bind(smem + (size_t)hb * (((int)threadIdx.x >> 6) % nsub), nb);
  int* gbs = (int*)(smem + (size_t)hb * nsub);
  PT* pay = (PT*)(gbs + ((s.num_groups + 3) & ~3));
  init_gbs(s, gbs, g0, g1, bin_lo, group_used);
  auto zero_all = [&]() {
    for (int c = 0; c < nsub; ++c) {
      LdsHist<FORM> Z;
      Z.bind(smem + (size_t)hb * c, nb);
      Z.zero(nb);
    }
  };
  // thread i writes partial bin i of every copy in turn, so the adds are ordered
  auto flush_all = [&](int slot, bool first) {
    for (int c = 0; c < nsub; ++c) {
      LdsHist<FORM> Z;
      Z.bind(smem + (size_t)hb * c, nb);
      Z.flush(s, slot, bin_lo, nb, first && c == 0);
    }
  };
  zero_all();
  const int c0 = g0 >> 4, nc = ((g1 - 1) >> 4) + 1 - c0;
  const int pay_chunk = s.pay_off >> 4, pay_in = s.pay_off & 15;
  const int tc = kHistThreads - kHistThreads % nc;  // threads of the chunk loop
  const int rstep = tc / nc;
  const int my_c = c0 + (int)threadIdx.x % nc;
  __syncthreads();
  int off[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int g = 16 * my_c + j;
    off[j] = g < s.num_groups ? gbs[g] : -1;
  }
  // last item starting at or before `start`
  int i = 0;
  {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].row0 <= start) lo = mid;
      else hi = mid - 1;
    }
    i = lo;
  }
  // partial of (block, item): from the host's slot table, or block + item in a
  // planned round (distinct, and consecutive over an item's blocks)
  int slot = block_slot ? block_slot[rb] : rb + i;
  for (long long x = start; x < end; ++i, ++slot) {
    const HistItem it = items[i];
    const Segment seg = it.seg;
    const long long iend = min(end, (long long)it.row0 + seg.count);
    const long long cb = (long long)seg.begin + (x - it.row0);
    const long long ce = (long long)seg.begin + (iend - it.row0);
    const uint8_t* src = rec_of(s, seg.buf) + 16 * my_c;
    // a partition buffer's rows carry their payload: with the payload's chunk
    // in the tile and a row's chunks in one wave, the lanes pass it across (no
    // second read of the records); otherwise it is staged in LDS per piece
    constexpr bool kQuantPay = FORM == FORM_QUANT || FORM == kHistQ32;  // 2-byte payloads
    const bool regs_pay =
        kQuantPay && seg.buf != 0 && pay_chunk >= c0 && pay_chunk < c0 + nc && 64 % nc == 0;  // block-uniform
    bool first = true;
    long long since = 0;
    for (long long t0 = cb; t0 < ce; t0 += kHistRows) {
      const int rows = (int)min((long long)kHistRows, ce - t0);
      if (!regs_pay) {
        for (int r = threadIdx.x; r < rows; r += kHistThreads) pay[r] = load_payload<FORM>(s, seg.buf, t0 + r);
        __syncthreads();
      }
      if ((int)threadIdx.x < tc) {
        const uint8_t* base = src + (size_t)t0 * s.stride;
        for (int r0 = (int)threadIdx.x / nc; r0 < rows; r0 += 4 * rstep) {
          uint4 v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = r0 + u * rstep;
            v[u] = r < rows ? *(const uint4*)(base + (size_t)r * s.stride) : make_uint4(0u, 0u, 0u, 0u);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = r0 + u * rstep;
            if constexpr (kQuantPay) {
              const PT p = regs_pay ? row_payload<FORM>(v[u], pay_in, nc, pay_chunk - c0) : (r < rows ? pay[r] : PT{});
              if (r < rows) hist_chunk_regs<FORM>(L, v[u], off, p);
            } else {
              if (r < rows) hist_chunk_regs<FORM>(L, v[u], off, pay[r]);
            }
          }
        }
      }
      since += rows;
      __syncthreads();
      if (since >= rows_per_flush) {
        flush_all(slot, first);
        __syncthreads();
        zero_all();
        __syncthreads();
        first = false;
        since = 0;
      }
    }
    if (since > 0 || first) {
      flush_all(slot, first);
      __syncthreads();
      zero_all();
      __syncthreads();
    }
    x = iend;
  }
}

// ---------------------------------------------------------------------------
// Single-pass partition (data_partition.rs:124-231): a leaf's rows move to the
// other buffer, left rows first, then right rows. Every item is cut into tiles
// of part_tile_rows(stride) rows, one block per tile. Per tile: the records are
// loaded as 16-byte chunks into registers (consecutive lanes, consecutive
// chunks), the lanes holding the split group's byte route their rows, a block
// scan ranks the rows, and one atomic add per side on the item's cursors claims
// the tile's two output runs; the chunks are then written straight to them.
//
// The reference's partition is stable; this one keeps row order within a tile,
// not across tiles. Nothing computed from the segments depends on that order:
// quantized histograms are exact integer sums, the float forms already add in
// atomic order, and the renewal sums, score updates and leaf rows read the row
// of every leaf from the tree walk over the records in their original order.
// Without an order to keep, no tile waits on another (no look-back).
//
// Cursors: two banks of (left, right) u32 per item; a launch uses bank
// (epoch & 1) and zeroes the other for the next launch. With the left child's
// count known (left_total >= 0) right rows start after the left ones; a tile
// whose claimed run would leave its side of the segment (the routed counts
// disagree with the search's exact counts) writes nothing and sets err bit 2.
// Without it (left_total < 0) right rows fill the segment from its end and the
// final cursors are the children's exact counts.
// ---------------------------------------------------------------------------
constexpr int kPartThreads = 512;
constexpr int kPartChunks = 16;  // 16-byte chunks per thread per tile
constexpr int kPartChunksSmall = 4;  // the same for small rounds: shorter tiles, more blocks per CU
constexpr int kPartSearchItems = 1024;  // items searched in LDS (more: bisection in global memory)
static_assert(kPartDeriveMax == kPartThreads, "the partition's last block derives one split per thread");

__device__ __forceinline__ long long wave_sum64(long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// LDS of the partition: [destinations R x u16][route flags R bytes]; with
// histograms also [payloads R x int16][histogram total_bins x u32]
__host__ __device__ inline size_t partition_lds(int R, int stride) {
  (void)stride;
  return ((size_t)R * 3 + 15) & ~(size_t)15;
}
// The plain partition's LDS: its tables, then the tile's payloads.
__host__ __device__ inline size_t partition_pay_lds(int R, int stride, int pay_bytes) {
  return partition_lds(R, stride) + (((size_t)R * pay_bytes + 15) & ~(size_t)15);
}
__host__ __device__ inline size_t partition_hist_lds(int R, int stride, int total_bins) {
  return partition_lds(R, stride) + (((size_t)R * 2 + 15) & ~(size_t)15) + (size_t)total_bins * 4;
}

template <int NT>
__device__ void derive_children(const PartItem* __restrict__ pitems, int ni, const unsigned* cur,
                                const ChildSpec* __restrict__ specs, const ChildParams& cp,
                                HistItem* __restrict__ hitems, ScanItem* __restrict__ sitems,
                                ReduceItem* __restrict__ ritems, RoundHdr* __restrict__ hdr, int* __restrict__ counts);

// End of a partition block: with a derivation, the block counts itself done
// (thread 0's cursor atomics returned before, so they are complete) and the last
// block derives the round's items from the final cursors.
__device__ __forceinline__ void partition_done(const PartItem* __restrict__ items, int nitems,
                                               const unsigned* cursors, int max_items, unsigned epoch,
                                               const DeriveArgs& da) {
  if (!da.on) return;
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(da.ticket, 1u) == gridDim.x - 1u ? 1 : 0;
  __syncthreads();
  if (!last) return;  // block-uniform
  if (threadIdx.x == 0) *da.ticket = 0u;  // every block has counted: the next launch starts at zero
  derive_children<kPartThreads>(items, nitems, cursors + (size_t)(epoch & 1u) * max_items * 2, da.specs, da.cp,
                                da.hitems, da.sitems, da.ritems, da.hdr, da.counts);
}

template <int FORM, bool HIST, int CH = kPartChunks>
__global__ void __launch_bounds__(kPartThreads) k_partition(DevState s, const PartItem* __restrict__ items,
                                                            int nitems, unsigned* __restrict__ cursors,
                                                            int max_items, unsigned epoch,
                                                            const RoundHdr* __restrict__ hdr,
                                                            const int* __restrict__ group_used, DeriveArgs da) {
  __shared__ int sh_lbase, sh_rbase;
  __shared__ int scratch[32];
  extern __shared__ __align__(16) unsigned char smem[];
  using PT = typename Pay<FORM>::T;
  const int stride = s.stride, nc = stride >> 4;
  const int R = kPartThreads * CH / nc;  // rows per tile
  const int rpt = (R + kPartThreads - 1) / kPartThreads;
  uint16_t* dstp = (uint16_t*)smem;  // R <= 8192
  uint8_t* lflag = (uint8_t*)(dstp + R);
  // the tile's payloads: from gh0 for a tile of buffer 0 (loaded beside the
  // records, patched into them on the way out), and the histogram side's
  PT* payt = (PT*)(smem + partition_lds(R, stride));
  unsigned* hist = (unsigned*)((unsigned char*)payt + (((size_t)R * 2 + 15) & ~(size_t)15));  // HIST: PT is 2 bytes
  const int t = blockIdx.x;
  // the cursors of the next launch start at zero
  unsigned* other = cursors + (size_t)((epoch + 1u) & 1u) * max_items * 2;
  for (int i = t * kPartThreads + (int)threadIdx.x; i < 2 * max_items; i += (int)gridDim.x * kPartThreads) other[i] = 0u;
  if (hdr) {  // a planned round: the grid covers the most tiles a round can hold
    nitems = hdr->npart;
    if (t >= hdr->ntasks) return;  // block-uniform
  }
  // the items' first tiles go to LDS: the item search then costs one global load
  // latency instead of one per bisection step
  __shared__ int sh_first[kPartSearchItems];
  const bool lds_search = nitems <= kPartSearchItems;
  if (lds_search) {
    for (int i = threadIdx.x; i < nitems; i += kPartThreads) sh_first[i] = items[i].first_task;
    __syncthreads();
  }
  int item;
  if (lds_search) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sh_first[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    item = lo;
  } else {
    item = find_item(items, nitems, t);
  }
  const PartItem& it = items[item];
  const int blk = t - it.first_task;
  const Segment seg = it.seg;
  const RuleDev& rule = it.rule;  // by reference: a local copy of its bitset would live in scratch
  const int left_total = it.left_total;
  const long long t0 = (long long)seg.begin + (long long)blk * it.chunk;
  const int rows = (int)min((long long)R, (long long)seg.begin + seg.count - t0);
  const uint8_t* src = rec_of(s, seg.buf) + (size_t)t0 * stride;
  uint8_t* dst = rec_of(s, it.dst_buf);
  const int total = rows * nc;
  const int cs = it.split_group >> 4, bs = it.split_group & 15;
  const int pay_chunk = s.pay_off >> 4, pay_in = s.pay_off & 15;
  // histogram of one side (block-uniform): the tile's rows of that side, over the
  // whole record; a thread's chunk is threadIdx.x % nc for every k (nc divides
  // the block), so its 16 group bin offsets stay in registers
  const int side = HIST ? it.hist_side : -1;
  if (HIST && side >= 0)
    for (int i = threadIdx.x; i < s.total_bins; i += kPartThreads) hist[i] = 0u;
  uint4 v[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int q = (int)threadIdx.x + k * kPartThreads;
    if (q < total) v[k] = load_stream(src + 16 * (size_t)q);
  }
  if (seg.buf == 0)
    for (int r = threadIdx.x; r < rows; r += kPartThreads) payt[r] = ((const PT*)s.gh0)[t0 + r];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int q = (int)threadIdx.x + k * kPartThreads;
    if (q < total) {
      const int r = q / nc, c = q - r * nc;
      if (c == cs) lflag[r] = route_left(rule, (lane_word(v[k], bs >> 2) >> (8 * (bs & 3))) & 0xffu) ? 1 : 0;
      if constexpr (HIST)
        if (side >= 0 && seg.buf != 0 && c == pay_chunk)
          payt[r] = (PT)((lane_word(v[k], pay_in >> 2) >> (8 * (pay_in & 3))) & 0xffffu);
    }
  }
  __syncthreads();
  int vm = 0, lm = 0;
  for (int i = 0; i < rpt; ++i) {
    const int r = (int)threadIdx.x * rpt + i;
    if (r < rows) {
      vm |= 1 << i;
      if (lflag[r]) lm |= 1 << i;
    }
  }
  int pl, pr, tl, tr;
  block_scan2<kPartThreads>(__popc(lm), __popc(vm & ~lm), &pl, &pr, &tl, &tr, scratch);
  if (threadIdx.x == 0) {
    unsigned* cur = cursors + (size_t)(epoch & 1u) * max_items * 2 + 2 * (size_t)item;
    sh_lbase = tl ? (int)atomicAdd(&cur[0], (unsigned)tl) : 0;
    sh_rbase = tr ? (int)atomicAdd(&cur[1], (unsigned)tr) : 0;
  }
  // tile-local destination of every row: left rows at [0, tl), right rows at
  // [tl, tl + tr), both in row order
  {
    int il = pl, ir = tl + pr;
    for (int i = 0; i < rpt; ++i) {
      const int r = (int)threadIdx.x * rpt + i;
      if ((vm >> i) & 1) dstp[r] = (uint16_t)(((lm >> i) & 1) ? il++ : ir++);
    }
  }
  __syncthreads();
  const int lbase = sh_lbase, rbase = sh_rbase;
  // left_total < 0: the right rows fill the segment from its end (each tile's run
  // in row order), and the cursors end at the children's counts
  const bool from_end = left_total < 0;
  const bool bad = from_end ? lbase + tl + rbase + tr > seg.count
                            : (lbase + tl > left_total || rbase + tr > seg.count - left_total);
  if (bad) {  // inconsistent counts: the tile writes nothing (block-uniform)
    if (threadIdx.x == 0) atomicOr(s.err_flag, 2);
    partition_done(items, nitems, cursors, max_items, epoch, da);
    return;
  }
  const long long lpos = (long long)seg.begin + lbase;
  const long long rpos = from_end ? (long long)seg.begin + seg.count - rbase - tr
                                  : (long long)seg.begin + left_total + rbase;
  // every lane writes its chunks straight to their destination rows; the rows of
  // a tile land in two contiguous runs, so the lanes of a wave cover contiguous
  // 16-byte pieces of them
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int q = (int)threadIdx.x + k * kPartThreads;
    if (q < total) {
      const int r = q / nc, c = q - r * nc;
      const int l = (int)dstp[r];
      const long long d = l < tl ? lpos + l : rpos + (l - tl);
      if (seg.buf == 0 && c == pay_chunk) patch_payload(v[k], pay_in, payt[r]);
      store_stream(dst + (size_t)d * stride + 16 * c, v[k]);
    }
  }
  if (!HIST || side < 0) {  // block-uniform
    partition_done(items, nitems, cursors, max_items, epoch, da);
    return;
  }
  // the side's rows into the LDS histogram (the stores above are in flight)
  if constexpr (HIST) {
    const int my_c = (int)threadIdx.x % nc;
    int off[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int g = 16 * my_c + j;
      off[j] = (g < s.num_groups && group_used[g]) ? s.group_bin_start[g] : -1;
    }
    const uint8_t want = side == 0 ? 1 : 0;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int q = (int)threadIdx.x + k * kPartThreads;
      if (q < total) {
        const int r = q / nc;
        if (lflag[r] == want) {
          const unsigned w = qpack32(payt[r]);
#pragma unroll
          for (int j = 0; j < 16; ++j)
            if (off[j] >= 0) atomicAdd(&hist[off[j] + (int)((lane_word(v[k], j >> 2) >> (8 * (j & 3))) & 0xffu)], w);
        }
      }
    }
  }
  __syncthreads();
  unsigned* part = (unsigned*)s.part_q + (size_t)t * s.total_bins;
  for (int i = threadIdx.x; i < s.total_bins; i += kPartThreads) part[i] = hist[i];
  partition_done(items, nitems, cursors, max_items, epoch, da);
}

// Records of buffer 0 from column-major bins: a block stages 64 rows of every
// group in LDS (coalesced column reads), then writes whole records.
constexpr int kRecRows = 64;
__global__ void __launch_bounds__(256) k_to_records(DevState s, const uint8_t* __restrict__ cols, size_t col_stride) {
  extern __shared__ __align__(16) unsigned char smem[];
  const long long row0 = (long long)blockIdx.x * kRecRows;
  const int rows = (int)min((long long)kRecRows, (long long)s.num_data - row0);
  const int stride = s.stride;
  for (int i = threadIdx.x; i < kRecRows * stride / 4; i += blockDim.x) ((uint32_t*)smem)[i] = 0u;
  __syncthreads();
  for (int i = threadIdx.x; i < s.num_groups * kRecRows; i += blockDim.x) {
    const int g = i / kRecRows, r = i - g * kRecRows;
    if (r < rows) smem[r * stride + g] = cols[(size_t)g * col_stride + row0 + r];
  }
  __syncthreads();
  const int nchunk = stride >> 4;
  for (int q = threadIdx.x; q < rows * nchunk; q += blockDim.x)
    *(uint4*)(s.rec[0] + (size_t)row0 * stride + 16 * (size_t)q) = ((const uint4*)smem)[q];
}

// ---------------------------------------------------------------------------
// Split search (feature_histogram.rs:219-1886), one feature per block.
// ---------------------------------------------------------------------------
struct Flags {
  bool l1, mo, sm;
};

__device__ __forceinline__ double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
__device__ __forceinline__ double thr_l1(double s, double l1) {
  double r = fmax(fabs(s) - l1, 0.0);
  return sgn(s) * r;
}
__device__ __forceinline__ int round_int(double x) { return sat_i32_dev(x + 0.5); }

__device__ double leaf_output(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent,
                              Flags f) {
  double ret = f.l1 ? -thr_l1(g, l1) / (h + l2) : -g / (h + l2);
  if (f.mo && mds > 0.0 && fabs(ret) > mds) ret = sgn(ret) * mds;
  if (f.sm) {
    double nd = (double)n;
    ret = ret * (nd / smooth) / (nd / smooth + 1.0) + parent / (nd / smooth + 1.0);
  }
  return ret;
}
__device__ double gain_given_output(double g, double h, double l1, double l2, double out, bool use_l1) {
  double gg = use_l1 ? thr_l1(g, l1) : g;
  double a = 2.0 * gg;
  double prod = a * out;
  double c = (h + l2) * out * out;
  return -(prod + c);
}
__device__ double leaf_gain(double g, double h, double l1, double l2, double mds, double smooth, int n, double parent,
                            Flags f) {
  if (!f.mo && !f.sm) {
    if (f.l1) {
      double sg = thr_l1(g, l1);
      return (sg * sg) / (h + l2);
    }
    return (g * g) / (h + l2);
  }
  double out = leaf_output(g, h, l1, l2, mds, smooth, n, parent, f);
  return gain_given_output(g, h, l1, l2, out, f.l1);
}
__device__ __forceinline__ double split_gain_plain(double gl, double hl, double gr, double hr, double l2) {
  double a = gl * gl / (hl + l2);
  double b = gr * gr / (hr + l2);
  return a + b;
}

// Prefix state of one scan direction at every bin, in LDS: the sums the
// sequential loop of FindBestThresholdSequentially(Int) holds when it evaluates
// bin t (feature_histogram.rs:363-663, 839-1142). Index t + 1 (t may be -1).
// The float forms use the f64 planes, the quantized form the int64 planes:
// they share storage (every access is under the form's branch).
struct ScanPrefix {
  union {
    double g[kMaxBins + 1];
    long long ig[kMaxBins + 1];
  };
  union {
    double h[kMaxBins + 1];
    long long ih[kMaxBins + 1];
  };
  int c[kMaxBins + 1];
  unsigned char active[kMaxBins + 1];
};

// Integer prefix state by a wave scan (exact in any order). Lane l owns bins
// 4l..4l+3; the reverse direction is a suffix scan, the forward one a prefix
// scan starting from the NA-as-missing initial sums at t = -1.
__device__ void build_prefix_int_wave(const long long* IG, const long long* IH, const FeatDev& m, long long tg,
                                      long long th, bool rev, bool skip_default, bool na_missing, ScanPrefix& X) {
  const int lane = threadIdx.x & 63;
  const int offset = m.offset;
  int lo, hi;  // active bins [lo, hi] of this direction
  if (rev) {
    hi = m.num_bin - 1 - offset - (na_missing ? 1 : 0);
    lo = 1 - offset;
  } else {
    lo = 0;
    hi = m.num_bin - 2 - offset;
  }
  long long vg[4], vh[4];
  bool act[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = 4 * lane + k;
    act[k] = t >= lo && t <= hi && !(skip_default && (t + offset) == m.default_bin);
    vg[k] = act[k] ? IG[t] : 0;
    vh[k] = act[k] ? IH[t] : 0;
  }
  if (rev) {
    long long sg[4], sh[4];
    sg[3] = vg[3];
    sh[3] = vh[3];
#pragma unroll
    for (int k = 2; k >= 0; --k) {
      sg[k] = sg[k + 1] + vg[k];
      sh[k] = sh[k + 1] + vh[k];
    }
    // exclusive suffix over the lanes above
    long long eg = sg[0], eh = sh[0];
    for (int o = 1; o < 64; o <<= 1) {
      const long long yg = __shfl_down(eg, o, 64), yh = __shfl_down(eh, o, 64);
      if (lane + o < 64) {
        eg += yg;
        eh += yh;
      }
    }
    eg -= sg[0];
    eh -= sh[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = 4 * lane + k;
      X.ig[t + 1] = sg[k] + eg;
      X.ih[t + 1] = sh[k] + eh;
      X.active[t + 1] = act[k];
    }
    if (lane == 0) X.active[0] = 0;
  } else {
    long long ig0 = 0, ih0 = 0;
    bool start_na = na_missing && offset == 1;
    if (start_na) {
      // tg - sum of every bin of the slice
      long long a = 0, b = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int t = 4 * lane + k;
        if (t < m.num_bin - offset) {
          a += IG[t];
          b += IH[t];
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
      }
      ig0 = tg - a;
      ih0 = th - b;
    }
    long long pg[4], ph[4];
    pg[0] = vg[0];
    ph[0] = vh[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
      pg[k] = pg[k - 1] + vg[k];
      ph[k] = ph[k - 1] + vh[k];
    }
    long long eg = pg[3], eh = ph[3];
    for (int o = 1; o < 64; o <<= 1) {
      const long long yg = __shfl_up(eg, o, 64), yh = __shfl_up(eh, o, 64);
      if (lane >= o) {
        eg += yg;
        eh += yh;
      }
    }
    eg -= pg[3];
    eh -= ph[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = 4 * lane + k;
      X.ig[t + 1] = ig0 + pg[k] + eg;
      X.ih[t + 1] = ih0 + ph[k] + eh;
      X.active[t + 1] = act[k];
    }
    if (lane == 0) {
      X.ig[0] = ig0;
      X.ih[0] = ih0;
      X.active[0] = start_na;
    }
  }
}

// Lane 0 runs the accumulation in the reference's order, so f64 prefix sums
// round exactly as the sequential scan does; integer sums are exact anyway.
template <bool INT>
__device__ void build_prefix(const double* G, const double* H, const long long* IG, const long long* IH,
                             const FeatDev& m, double sum_g, double sum_h, long long tg, long long th, int num_data,
                             double cnt_factor, bool rev, bool skip_default, bool na_missing, ScanPrefix& X) {
  const int offset = m.offset;
  for (int i = 0; i <= kMaxBins; ++i) X.active[i] = 0;
  if (rev) {
    double srg = 0.0, srh = kEps;
    long long rg = 0, rh = 0;
    int rc = 0;
    for (int t = m.num_bin - 1 - offset - (na_missing ? 1 : 0); t >= 1 - offset; --t) {
      if (skip_default && (t + offset) == m.default_bin) continue;
      if (INT) {
        rg += IG[t];
        rh += IH[t];
        X.ig[t + 1] = rg;
        X.ih[t + 1] = rh;
      } else {
        const double g = G[t], h = H[t];
        srg += g;
        srh += h;
        rc += round_int(h * cnt_factor);
        X.g[t + 1] = srg;
        X.h[t + 1] = srh;
        X.c[t + 1] = rc;
      }
      X.active[t + 1] = 1;
    }
  } else {
    double slg = 0.0, slh = kEps;
    long long lg = 0, lh = 0;
    int lc = 0;
    int t = 0;
    if (na_missing && offset == 1) {
      if (INT) {
        lg = tg;
        lh = th;
        for (int i = 0; i < m.num_bin - offset; ++i) {
          lg -= IG[i];
          lh -= IH[i];
        }
      } else {
        slg = sum_g;
        slh = sum_h - kEps;
        lc = num_data;
        for (int i = 0; i < m.num_bin - offset; ++i) {
          const double g = G[i], h = H[i];
          slg -= g;
          slh -= h;
          lc -= round_int(h * cnt_factor);
        }
      }
      t = -1;
    }
    for (; t <= m.num_bin - 2 - offset; ++t) {
      if (skip_default && (t + offset) == m.default_bin) continue;
      if (t >= 0) {
        if (INT) {
          lg += IG[t];
          lh += IH[t];
        } else {
          slg += G[t];
          slh += H[t];
          lc += round_int(H[t] * cnt_factor);
        }
      }
      if (INT) {
        X.ig[t + 1] = lg;
        X.ih[t + 1] = lh;
      } else {
        X.g[t + 1] = slg;
        X.h[t + 1] = slh;
        X.c[t + 1] = lc;
      }
      X.active[t + 1] = 1;
    }
  }
}

struct Best {
  double gain;
  int pos;  // index t + 1
};

// earlier in scan order wins ties: higher t when reversing, lower t otherwise
__device__ __forceinline__ bool better(const Best& a, const Best& b, bool rev) {
  if (a.pos < 0) return false;
  if (b.pos < 0) return true;
  if (a.gain != b.gain) return a.gain > b.gain;
  return rev ? a.pos > b.pos : a.pos < b.pos;
}

// One direction evaluated by the 64 lanes of a wave from the prefix state.
// Returns in `out` what the sequential scan would leave there.
template <bool INT>
__device__ void eval_direction(const FeatDev& m, const ScanParams& p, Flags fl, double sum_g, double sum_h,
                               long long tg, long long th, int num_data, double mgs, double parent,
                               double cnt_factor, bool rev, const ScanPrefix& X, SplitOut& out) {
  const int lane = threadIdx.x & 63;
  const int offset = m.offset;
  // break position: the first bin in scan order where the side constraints fail
  int brk = rev ? -1 : kMaxBins + 2;
  Best best{kMinScore, -1};
  int any = 0;
  double g_pos[5];
  int valid_pos[5];
  for (int j = 0; j < 5; ++j) {
    const int pos = lane + 64 * j;
    valid_pos[j] = 0;
    g_pos[j] = kMinScore;
    if (pos > kMaxBins || !X.active[pos]) continue;
    double slg, slh, srg, srh;
    int lc, rc;
    bool c1, c23;
    if (INT) {
      if (rev) {
        const long long rg = X.ig[pos], rh = X.ih[pos];
        rc = round_int((double)rh * cnt_factor);
        srh = (double)rh * p.hess_scale;
        c1 = rc < p.min_data_in_leaf || srh < p.min_sum_hessian;
        lc = num_data - rc;
        const long long lh = th - rh;
        slh = (double)lh * p.hess_scale;
        c23 = lc < p.min_data_in_leaf || slh < p.min_sum_hessian;
        slg = (double)(tg - rg) * p.grad_scale;
        srg = (double)rg * p.grad_scale;
      } else {
        const long long lg = X.ig[pos], lh = X.ih[pos];
        lc = round_int((double)lh * cnt_factor);
        slh = (double)lh * p.hess_scale;
        c1 = lc < p.min_data_in_leaf || slh < p.min_sum_hessian;
        rc = num_data - lc;
        const long long rh = th - lh;
        srh = (double)rh * p.hess_scale;
        c23 = rc < p.min_data_in_leaf || srh < p.min_sum_hessian;
        slg = (double)lg * p.grad_scale;
        srg = (double)(tg - lg) * p.grad_scale;
      }
    } else {
      if (rev) {
        srg = X.g[pos];
        srh = X.h[pos];
        rc = X.c[pos];
        c1 = rc < p.min_data_in_leaf || srh < p.min_sum_hessian;
        lc = num_data - rc;
        slh = sum_h - srh;
        c23 = lc < p.min_data_in_leaf || slh < p.min_sum_hessian;
        slg = sum_g - srg;
      } else {
        slg = X.g[pos];
        slh = X.h[pos];
        lc = X.c[pos];
        c1 = lc < p.min_data_in_leaf || slh < p.min_sum_hessian;
        rc = num_data - lc;
        srh = sum_h - slh;
        c23 = rc < p.min_data_in_leaf || srh < p.min_sum_hessian;
        srg = sum_g - slg;
      }
    }
    if (c1) continue;
    if (c23) {
      brk = rev ? max(brk, pos) : min(brk, pos);
      continue;
    }
    double gain;
    {
      double a_slh = slh, a_srh = srh;
      if (INT) {
        a_slh += kEps;
        a_srh += kEps;
      }
      if (!fl.l1 && !fl.mo && !fl.sm) {
        gain = split_gain_plain(slg, a_slh, srg, a_srh, p.l2);
      } else {
        double a = leaf_gain(slg, a_slh, p.l1, p.l2, p.max_delta_step, p.path_smooth, lc, parent, fl);
        double b = leaf_gain(srg, a_srh, p.l1, p.l2, p.max_delta_step, p.path_smooth, rc, parent, fl);
        gain = a + b;
      }
    }
    g_pos[j] = gain;
    valid_pos[j] = 1;
  }
  // reduce the break position over the wave
  for (int o = 32; o > 0; o >>= 1) {
    const int other = __shfl_xor(brk, o, 64);
    brk = rev ? max(brk, other) : min(brk, other);
  }
  for (int j = 0; j < 5; ++j) {
    const int pos = lane + 64 * j;
    if (!valid_pos[j]) continue;
    if (rev ? pos <= brk : pos >= brk) continue;
    const double gain = g_pos[j];
    if (gain <= mgs) continue;  // NaN gains are not skipped: they count as splittable
    any = 1;
    Best c{gain, pos};
    if (gain == gain && better(c, best, rev)) best = c;
  }
  for (int o = 32; o > 0; o >>= 1) {
    Best other;
    other.gain = __shfl_xor(best.gain, o, 64);
    other.pos = __shfl_xor(best.pos, o, 64);
    if (better(other, best, rev)) best = other;
    any |= __shfl_xor(any, o, 64);
  }
  if (any) out.splittable = 1;
  if (any && best.pos >= 0 && best.gain > out.gain + mgs) {
    const int pos = best.pos, t = pos - 1;
    out.threshold = rev ? t - 1 + offset : t + offset;
    if (INT) {
      long long blg, blh;
      if (rev) {
        blg = tg - X.ig[pos];
        blh = th - X.ih[pos];
      } else {
        blg = X.ig[pos];
        blh = X.ih[pos];
      }
      const long long rg = tg - blg, rh = th - blh;
      const double slg = (double)blg * p.grad_scale, slh = (double)blh * p.hess_scale;
      const double srg = (double)rg * p.grad_scale, srh = (double)rh * p.hess_scale;
      const int lc = round_int((double)blh * cnt_factor);
      const int rc = round_int((double)rh * cnt_factor);
      out.left_output = leaf_output(slg, slh, p.l1, p.l2, p.max_delta_step, p.path_smooth, lc, parent, fl);
      out.left_count = lc;
      out.left_sum_gradient = slg;
      out.left_sum_hessian = slh;
      out.right_output = leaf_output(srg, srh, p.l1, p.l2, p.max_delta_step, p.path_smooth, rc, parent, fl);
      out.right_count = rc;
      out.right_sum_gradient = srg;
      out.right_sum_hessian = srh;
      out.left_int_sum = (long long)(((unsigned long long)blg << 32) | ((unsigned long long)blh & 0xffffffffull));
      out.right_int_sum = (long long)(((unsigned long long)rg << 32) | ((unsigned long long)rh & 0xffffffffull));
    } else {
      double bslg, bslh;
      int blc;
      if (rev) {
        bslg = sum_g - X.g[pos];
        bslh = sum_h - X.h[pos];
        blc = num_data - X.c[pos];
      } else {
        bslg = X.g[pos];
        bslh = X.h[pos];
        blc = X.c[pos];
      }
      out.left_output = leaf_output(bslg, bslh, p.l1, p.l2, p.max_delta_step, p.path_smooth, blc, parent, fl);
      out.left_count = blc;
      out.left_sum_gradient = bslg;
      out.left_sum_hessian = bslh - kEps;
      out.right_output = leaf_output(sum_g - bslg, sum_h - bslh, p.l1, p.l2, p.max_delta_step, p.path_smooth,
                                     num_data - blc, parent, fl);
      out.right_count = num_data - blc;
      out.right_sum_gradient = sum_g - bslg;
      out.right_sum_hessian = sum_h - bslh - kEps;
    }
    out.gain = best.gain - mgs;
    out.default_left = rev ? 1 : 0;
  }
}

__device__ double leaf_gain_cat(double g, double h, double l1, double l2, double mds, double smooth, int n,
                                double parent, Flags f) {
  return leaf_gain(g, h, l1, l2, mds, smooth, n, parent, f);
}

// Categorical search (feature_histogram.rs:1328-1597, 1606-1886). `sorted` holds
// the eligible bins in stable ctr order (computed by the wave beforehand).
template <bool INT>
__device__ void search_categorical(const double* G, const double* H, const long long* IG, const long long* IH,
                                   const FeatDev& m, const ScanParams& p, double sum_g_in, double sum_h_in,
                                   long long tg, long long th, int num_data, double parent, const int* sorted,
                                   int used_bin, SplitOut& out) {
  double sum_g, sum_h;
  if (INT) {
    sum_g = (double)tg * p.grad_scale;
    sum_h = (double)th * p.hess_scale;
  } else {
    sum_g = sum_g_in;
    sum_h = sum_h_in + 2.0 * kEps;
  }
  out.default_left = 0;
  out.gain = kMinScore;
  Flags fl{p.l1 > 0.0, p.max_delta_step > 0.0, p.path_smooth > kEps};
  double gain_shift;
  if (fl.sm) gain_shift = gain_given_output(sum_g, sum_h, p.l1, p.l2, parent, fl.l1);
  else gain_shift = leaf_gain(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, 0.0, num_data, 0.0, Flags{fl.l1, fl.mo, false});
  const double mgs = gain_shift + p.min_gain_to_split;
  const int offset = m.offset;
  const int bin_start = 1 - offset, bin_end = m.num_bin - offset;
  const double cnt_factor = INT ? (double)num_data / (double)th : (double)num_data / sum_h;
  const bool onehot = m.num_bin <= p.max_cat_to_onehot;
  double l2 = p.l2;
  bool splittable = false;
  double best_gain = kMinScore, best_slg = 0.0, best_slh = 0.0;
  long long best_lg = 0, best_lh = 0;
  int best_threshold = -1, best_dir = 1, best_lc = 0;
  auto gains = [&](double slg, double slh, double srg, double srh, int lc, int rc) -> double {
    double a = leaf_gain_cat(slg, slh, p.l1, l2, p.max_delta_step, p.path_smooth, lc, parent, fl);
    double b = leaf_gain_cat(srg, srh, p.l1, l2, p.max_delta_step, p.path_smooth, rc, parent, fl);
    return a + b;
  };
  if (onehot) {
    for (int t = bin_start; t < bin_end; ++t) {
      if (INT) {
        long long ig = IG[t], ih = IH[t];
        double h = (double)ih * p.hess_scale;
        int cnt = round_int((double)ih * cnt_factor);
        if (cnt < p.min_data_in_leaf || h < p.min_sum_hessian) continue;
        int oc = num_data - cnt;
        if (oc < p.min_data_in_leaf) continue;
        long long ioh = th - ih;
        double soh = (double)ioh * p.hess_scale;
        if (soh < p.min_sum_hessian) continue;
        double g = (double)ig * p.grad_scale;
        double sog = (double)(tg - ig) * p.grad_scale;
        double cur = gains(sog, soh, g, h, oc, cnt);
        if (cur <= mgs) continue;
        splittable = true;
        if (cur > best_gain) {
          best_threshold = t;
          best_lg = ig;
          best_lh = ih;
          best_gain = cur;
        }
      } else {
        double g = G[t], h = H[t];
        int cnt = round_int(h * cnt_factor);
        if (cnt < p.min_data_in_leaf || h < p.min_sum_hessian) continue;
        int oc = num_data - cnt;
        if (oc < p.min_data_in_leaf) continue;
        double soh = sum_h - h - kEps;
        if (soh < p.min_sum_hessian) continue;
        double sog = sum_g - g;
        double cur = gains(sog, soh, g, h + kEps, oc, cnt);
        if (cur <= mgs) continue;
        splittable = true;
        if (cur > best_gain) {
          best_threshold = t;
          best_slg = g;
          best_slh = h + kEps;
          best_lc = cnt;
          best_gain = cur;
        }
      }
    }
  } else {
    l2 += p.cat_l2;
    const int max_num_cat = min(p.max_cat_threshold, (used_bin + 1) / 2);
    for (int di = 0; di < 2; ++di) {
      const int dir = di == 0 ? 1 : -1;
      int pos = di == 0 ? 0 : used_bin - 1;
      int cnt_cur_group = 0;
      double slg = 0.0, slh = kEps;
      long long lg = 0, lh = 0;
      int lc = 0;
      for (int i = 0; i < used_bin && i < max_num_cat; ++i) {
        const int t = sorted[pos];
        pos += dir;
        if (INT) {
          long long ig = IG[t], ih = IH[t];
          int cnt = round_int((double)ih * cnt_factor);
          lg += ig;
          lh += ih;
          lc += cnt;
          cnt_cur_group += cnt;
          double slh_v = (double)lh * p.hess_scale;
          if (lc < p.min_data_in_leaf || slh_v < p.min_sum_hessian) continue;
          int rc = num_data - lc;
          if (rc < p.min_data_in_leaf || rc < p.min_data_per_group) break;
          long long rh = th - lh;
          double srh = (double)rh * p.hess_scale;
          if (srh < p.min_sum_hessian) break;
          if (cnt_cur_group < p.min_data_per_group) continue;
          cnt_cur_group = 0;
          double slg_v = (double)lg * p.grad_scale;
          double srg = (double)(tg - lg) * p.grad_scale;
          double cur = gains(slg_v, slh_v, srg, srh, lc, rc);
          if (cur > mgs) {
            splittable = true;
            if (cur > best_gain) {
              best_lg = lg;
              best_lh = lh;
              best_threshold = i;
              best_gain = cur;
              best_dir = dir;
            }
          }
        } else {
          double g = G[t], h = H[t];
          int cnt = round_int(h * cnt_factor);
          slg += g;
          slh += h;
          lc += cnt;
          cnt_cur_group += cnt;
          if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian) continue;
          int rc = num_data - lc;
          if (rc < p.min_data_in_leaf || rc < p.min_data_per_group) break;
          double srh = sum_h - slh;
          if (srh < p.min_sum_hessian) break;
          if (cnt_cur_group < p.min_data_per_group) continue;
          cnt_cur_group = 0;
          double srg = sum_g - slg;
          double cur = gains(slg, slh, srg, srh, lc, rc);
          if (cur > mgs) {
            splittable = true;
            if (cur > best_gain) {
              best_lc = lc;
              best_slg = slg;
              best_slh = slh;
              best_threshold = i;
              best_gain = cur;
              best_dir = dir;
            }
          }
        }
      }
    }
  }
  out.splittable = splittable ? 1 : 0;
  if (splittable) {
    if (INT) {
      double slg = (double)best_lg * p.grad_scale, slh = (double)best_lh * p.hess_scale;
      long long rg = tg - best_lg, rh = th - best_lh;
      double srg = (double)rg * p.grad_scale, srh = (double)rh * p.hess_scale;
      int lc = round_int((double)best_lh * cnt_factor);
      int rc = round_int((double)rh * cnt_factor);
      out.left_output = leaf_output(slg, slh, p.l1, l2, p.max_delta_step, p.path_smooth, lc, parent, fl);
      out.left_count = lc;
      out.left_sum_gradient = slg;
      out.left_sum_hessian = slh;
      out.right_output = leaf_output(srg, srh, p.l1, l2, p.max_delta_step, p.path_smooth, rc, parent, fl);
      out.right_count = rc;
      out.right_sum_gradient = srg;
      out.right_sum_hessian = srh;
      out.left_int_sum = (long long)(((unsigned long long)best_lg << 32) | ((unsigned long long)best_lh & 0xffffffffull));
      out.right_int_sum = (long long)(((unsigned long long)rg << 32) | ((unsigned long long)rh & 0xffffffffull));
    } else {
      out.left_output = leaf_output(best_slg, best_slh, p.l1, l2, p.max_delta_step, p.path_smooth, best_lc, parent, fl);
      out.left_count = best_lc;
      out.left_sum_gradient = best_slg;
      out.left_sum_hessian = best_slh - kEps;
      out.right_output = leaf_output(sum_g - best_slg, sum_h - best_slh, p.l1, l2, p.max_delta_step, p.path_smooth,
                                     num_data - best_lc, parent, fl);
      out.right_count = num_data - best_lc;
      out.right_sum_gradient = sum_g - best_slg;
      out.right_sum_hessian = sum_h - best_slh - kEps;
    }
    out.gain = best_gain - mgs;
    for (int i = 0; i < 8; ++i) out.cat_bits[i] = 0u;
    if (onehot) {
      out.num_cat = 1;
      bits_set(out.cat_bits, (unsigned)(best_threshold + offset));
    } else {
      out.num_cat = best_threshold + 1;
      for (int i = 0; i <= best_threshold; ++i) {
        int ps = best_dir == 1 ? i : used_bin - 1 - i;
        bits_set(out.cat_bits, (unsigned)(sorted[ps] + offset));
      }
    }
  }
  out.gain *= m.penalty;
}

// One leaf's feature slice in LDS.
struct SliceLds {
  union {
    double g[kMaxBins];
    long long ig[kMaxBins];
  };
  union {
    double h[kMaxBins];
    long long ih[kMaxBins];
  };
  int c[kMaxBins];
  int sorted[kMaxBins];
  int used_bin;
};

// Raw histogram slice of one feature from a pool slot (summed by k_reduce).
// A constant hessian counts rows in the hessian slot; it is scaled by
// hessians[0] here, after the whole sum (histogram.rs:226-232, serial.rs:982-988).
template <int FORM>
__device__ void load_slice_raw(const DevState& s, const FeatDev& m, int slot, const ScanParams& p, SliceLds& S) {
  const double* pool = s.pool + ((size_t)slot * s.total_bins) * 2;
  const long long* pooli = (const long long*)pool;
  const int* cnt = s.pool_cnt + (size_t)slot * s.total_bins;
  const bool counts = !s.hist32;  // 32-bit quantized bins: the slots keep no counts
  for (int i = threadIdx.x; i < m.hist_len; i += blockDim.x) {
    const int bin = m.hist_off + i;
    if (FORM == FORM_QUANT) {
      S.ig[i] = pooli[2 * bin];
      S.ih[i] = pooli[2 * bin + 1];
    } else {
      S.g[i] = pool[2 * bin];
      S.h[i] = p.const_hess ? pool[2 * bin + 1] * p.h0 : pool[2 * bin + 1];
    }
    S.c[i] = counts ? cnt[bin] : 0;
  }
}

template <int FORM>
__device__ void fix_slice(const FeatDev& m, const LeafArgs& a, SliceLds& S, long long* red) {
  const int mfb = m.mfb, nb = m.num_bin;
  if (mfb <= 0) return;  // uniform over the block
  long long sg = 0, sh = 0, sc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    if (i == mfb) continue;
    if (FORM == FORM_QUANT) {
      sg += S.ig[i];
      sh += S.ih[i];
    }
    sc += S.c[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    sg += __shfl_xor(sg, o, 64);
    sh += __shfl_xor(sh, o, 64);
    sc += __shfl_xor(sc, o, 64);
  }
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[3 * wave] = sg;
    red[3 * wave + 1] = sh;
    red[3 * wave + 2] = sc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long tg = 0, th = 0, tc = 0;
    for (int w = 0; w < nw; ++w) {
      tg += red[3 * w];
      th += red[3 * w + 1];
      tc += red[3 * w + 2];
    }
    if (FORM == FORM_QUANT) {
      S.ig[mfb] = (a.int_sum >> 32) - tg;
      S.ih[mfb] = (a.int_sum & 0xffffffffLL) - th;
    } else {
      S.g[mfb] = a.sum_g;
      S.h[mfb] = a.sum_h;
      for (int i = 0; i < nb; ++i)
        if (i != mfb) {
          S.g[mfb] -= S.g[i];
          S.h[mfb] -= S.h[i];
        }
    }
    S.c[mfb] = (int)((long long)a.num_data - tc);
  }
  __syncthreads();
}

template <int FORM>
__device__ void store_slice(const DevState& s, const FeatDev& m, int slot, const SliceLds& S) {
  double* pool = s.pool + ((size_t)slot * s.total_bins) * 2;
  long long* pooli = (long long*)pool;
  int* cnt = s.pool_cnt + (size_t)slot * s.total_bins;
  const bool counts = !s.hist32;
  for (int i = threadIdx.x; i < m.hist_len; i += blockDim.x) {
    const int bin = m.hist_off + i;
    if (FORM == FORM_QUANT) {
      pooli[2 * bin] = S.ig[i];
      pooli[2 * bin + 1] = S.ih[i];
    } else {
      pool[2 * bin] = S.g[i];
      pool[2 * bin + 1] = S.h[i];
    }
    if (counts) cnt[bin] = S.c[i];
  }
}

// Larger leaf = parent (in place in the larger slot) - fixed smaller, over the slice.
template <int FORM>
__device__ void subtract_slice(const DevState& s, const FeatDev& m, int slot, const SliceLds& small, SliceLds& L) {
  double* pool = s.pool + ((size_t)slot * s.total_bins) * 2;
  long long* pooli = (long long*)pool;
  int* cnt = s.pool_cnt + (size_t)slot * s.total_bins;
  const bool counts = !s.hist32;
  for (int i = threadIdx.x; i < m.hist_len; i += blockDim.x) {
    const int bin = m.hist_off + i;
    if (FORM == FORM_QUANT) {
      L.ig[i] = pooli[2 * bin] - small.ig[i];
      L.ih[i] = pooli[2 * bin + 1] - small.ih[i];
    } else {
      L.g[i] = pool[2 * bin] - small.g[i];
      L.h[i] = pool[2 * bin + 1] - small.h[i];
    }
    L.c[i] = counts ? cnt[bin] - small.c[i] : 0;
  }
}

// Bin threadIdx.x of a slot's slice, in registers (the parent's fixed values).
struct ParentBin {
  long long ig, ih;
  double g, h;
  int c;
};

template <int FORM>
__device__ __forceinline__ ParentBin load_parent_bin(const DevState& s, const FeatDev& m, int slot) {
  ParentBin b{};
  const int i = (int)threadIdx.x;
  if (i >= m.hist_len) return b;
  const int bin = m.hist_off + i;
  const double* pool = s.pool + ((size_t)slot * s.total_bins) * 2;
  if (FORM == FORM_QUANT) {
    b.ig = ((const long long*)pool)[2 * bin];
    b.ih = ((const long long*)pool)[2 * bin + 1];
  } else {
    b.g = pool[2 * bin];
    b.h = pool[2 * bin + 1];
  }
  b.c = !s.hist32 ? s.pool_cnt[(size_t)slot * s.total_bins + bin] : 0;
  return b;
}

// subtract_slice with the parent's bin already in registers.
template <int FORM>
__device__ __forceinline__ void subtract_parent_bin(const DevState& s, const FeatDev& m, int slot, const ParentBin& b,
                                                    const SliceLds& small, SliceLds& L) {
  const int i = (int)threadIdx.x;
  if (i >= m.hist_len) return;
  if (FORM == FORM_QUANT) {
    L.ig[i] = b.ig - small.ig[i];
    L.ih[i] = b.ih - small.ih[i];
  } else {
    L.g[i] = b.g - small.g[i];
    L.h[i] = b.h - small.h[i];
  }
  L.c[i] = !s.hist32 ? b.c - small.c[i] : 0;
}

// Stable ctr order of the eligible bins (feature_histogram.rs:1478-1486):
// rank = number of eligible bins with a smaller ctr, or an equal ctr and a lower index.
template <int FORM>
__device__ void cat_sort(const FeatDev& m, const ScanParams& p, const LeafArgs& a, SliceLds& S, int* elig) {
  const int offset = m.offset, bin_start = 1 - offset, bin_end = m.num_bin - offset;
  double cnt_factor;
  if (FORM == FORM_QUANT) cnt_factor = (double)a.num_data / (double)(a.int_sum & 0xffffffffLL);
  else cnt_factor = (double)a.num_data / (a.sum_h + 2.0 * kEps);
  auto ctr = [&](int i) -> double {
    if (FORM == FORM_QUANT)
      return ((double)S.ig[i] * p.grad_scale) / ((double)S.ih[i] * p.hess_scale + p.cat_smooth);
    return S.g[i] / (S.h[i] + p.cat_smooth);
  };
  for (int i = threadIdx.x; i < kMaxBins; i += blockDim.x) {
    int e = 0;
    if (i >= bin_start && i < bin_end) {
      double hv = FORM == FORM_QUANT ? (double)S.ih[i] * cnt_factor : S.h[i] * cnt_factor;
      e = (double)round_int(hv) >= p.cat_smooth ? 1 : 0;
    }
    elig[i] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int i = 0; i < kMaxBins; ++i) n += elig[i];
    S.used_bin = n;
  }
  for (int i = threadIdx.x; i < kMaxBins; i += blockDim.x) {
    if (!elig[i]) continue;
    const double ci = ctr(i);
    int r = 0;
    for (int j = 0; j < kMaxBins; ++j) {
      if (!elig[j] || j == i) continue;
      const double cj = ctr(j);
      if (cj < ci || (!(ci < cj) && j < i)) ++r;
    }
    S.sorted[r] = i;
  }
  __syncthreads();
}

// Exact rows routed left by the chosen split of one leaf (count channel), one wave.
__device__ void exact_counts(const FeatDev& m, const SliceLds& S, int num_data, SplitOut& o) {
  const int lane = threadIdx.x & 63;
  RuleDev r = m.bin_type == 1 ? make_rule_categorical(m, o.cat_bits)
                              : make_rule_numerical(m, (uint32_t)o.threshold, o.default_left != 0);
  long long left = 0, inslice = 0;
  for (int i = lane; i < m.hist_len; i += 64) {
    inslice += S.c[i];
    if (route_left(r, stored_value(m, i + m.offset))) left += S.c[i];
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) {
    left += __shfl_xor(left, o2, 64);
    inslice += __shfl_xor(inslice, o2, 64);
  }
  if (m.offset == 1 && route_left(r, 0u)) left += (long long)num_data - inslice;
  o.exact_left = (int)left;
  o.exact_right = num_data - (int)left;
}

// Numerical search of one leaf by one wave; every wave of the block calls it
// with the same feature (uniform barriers), `enabled` off for an absent leaf.
template <int FORM>
__device__ void search_numerical_wave(const FeatDev& m, const ScanParams& p, const LeafArgs& a, const SliceLds& S,
                                      ScanPrefix& X, bool enabled, SplitOut& out) {
  constexpr bool INT = FORM == FORM_QUANT;
  const int lane = threadIdx.x & 63;
  Flags fl{p.l1 > 0.0, p.max_delta_step > 0.0, p.path_smooth > kEps};
  const long long tg = a.int_sum >> 32, th = a.int_sum & 0xffffffffLL;
  double sum_g, sum_h;
  if (INT) {
    sum_g = (double)tg * p.grad_scale;
    sum_h = (double)th * p.hess_scale;
  } else {
    sum_g = a.sum_g;
    sum_h = a.sum_h + 2.0 * kEps;
  }
  const double cnt_factor = INT ? (double)a.num_data / (double)th : (double)a.num_data / sum_h;
  const double gain_shift =
      leaf_gain(sum_g, sum_h, p.l1, p.l2, p.max_delta_step, p.path_smooth, a.num_data, a.parent_output, fl);
  const double mgs = gain_shift + p.min_gain_to_split;
  out.default_left = 1;
  out.gain = kMinScore;
  bool dirs_rev[2], dirs_skip[2], dirs_na[2];
  int nd;
  if (m.num_bin > 2 && m.missing_type != 0) {
    nd = 2;
    const bool zero = m.missing_type == 1;
    dirs_rev[0] = true;
    dirs_rev[1] = false;
    dirs_skip[0] = dirs_skip[1] = zero;
    dirs_na[0] = dirs_na[1] = !zero;
  } else {
    nd = 1;
    dirs_rev[0] = true;
    dirs_skip[0] = false;
    dirs_na[0] = false;
  }
  for (int d = 0; d < nd; ++d) {
    if (INT) {
      if (enabled) build_prefix_int_wave(S.ig, S.ih, m, tg, th, dirs_rev[d], dirs_skip[d], dirs_na[d], X);
    } else if (enabled && lane == 0) {
      build_prefix<INT>(S.g, S.h, S.ig, S.ih, m, sum_g, sum_h, tg, th, a.num_data, cnt_factor, dirs_rev[d],
                        dirs_skip[d], dirs_na[d], X);
    }
    __syncthreads();
    if (enabled)
      eval_direction<INT>(m, p, fl, sum_g, sum_h, tg, th, a.num_data, mgs, a.parent_output, cnt_factor, dirs_rev[d],
                          X, out);
    __syncthreads();
  }
  if (nd == 1 && m.missing_type == 2) out.default_left = 0;
  out.gain *= m.penalty;
}

// One wave folds the per-feature candidates of (item k, side) in feature order
// with better_than (serial.rs:1500-1566, split_info.rs:67-85): the highest gain
// wins, a tie goes to the lower real feature index (-1 counting as largest), a
// NaN gain never wins. Also writes every feature's splittable flag (0 for
// features not searched).
__device__ void fold_side(const ScanItem& it, int k, int side, const SplitOut* out_all,
                          const int* __restrict__ used_all, int F, FoldOut* __restrict__ best,
                          uint8_t* __restrict__ flags) {
  const int lane = threadIdx.x & 63;
  uint8_t* fl = flags + (size_t)(2 * k + side) * F;
  FoldOut* dst = best + 2 * k + side;
  const bool active = side == 0 || it.has_large;
  const int* used = used_all + it.used_off;
  const SplitOut* out = out_all + it.out_off + (side ? F : 0);
  double bg = kMinScore;
  int bk = INT_MAX, bf = -1;
  // used[f]: bit 0 searched (its splittable flag is kept), bit 1 / bit 2 it may be
  // the smaller / larger leaf's best (the by-node feature samples)
  for (int f = lane; f < F; f += 64) {
    int sp = 0;
    if (active && (used[f] & 1)) {
      const double gain = out[f].gain;
      const int feat = out[f].feature;
      sp = out[f].splittable ? 1 : 0;
      const int key = feat == -1 ? INT_MAX : feat;
      const bool eligible = (used[f] >> (1 + side)) & 1;
      if (eligible && !(gain != gain) && (gain > bg || (gain == bg && key < bk))) {
        bg = gain;
        bk = key;
        bf = f;
      }
    }
    fl[f] = (uint8_t)sp;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double og = __shfl_xor(bg, o, 64);
    const int ok = __shfl_xor(bk, o, 64), of = __shfl_xor(bf, o, 64);
    if (of >= 0 && (bf < 0 || og > bg || (og == bg && ok < bk))) {
      bg = og;
      bk = ok;
      bf = of;
    }
  }
  constexpr int kWords = (int)(sizeof(SplitOut) / 4);
  if (bf >= 0) {
    const uint32_t* src = (const uint32_t*)(out + bf);
    uint32_t* d = (uint32_t*)&dst->o;
    for (int i = lane; i < kWords; i += 64) d[i] = src[i];
  } else if (lane == 0) {
    dst->o = SplitOut{};
    dst->o.gain = kMinScore;
    dst->o.feature = -1;
    dst->o.default_left = 1;
  }
  if (lane == 0) dst->inner = bf;
}

// The search of feature blockIdx.x for search item y (one block; returns are
// block-uniform, except for the waves that search no leaf, which leave before
// the last wave-level steps).
template <int FORM>
__device__ __forceinline__ void scan_item(const DevState& s, const ScanItem* __restrict__ items, int y,
                                          const ScanParams& p0, const int* __restrict__ used_all, SplitOut* out_all) {
  ScanParams p = p0;
  if (p0.dyn) {  // scales computed on the device for this tree
    p.grad_scale = p0.dyn[0];
    p.hess_scale = p0.dyn[1];
    p.h0 = p0.dyn[2];
  }
  __shared__ SliceLds S[2];
  __shared__ ScanPrefix X[2];
  __shared__ int elig[2][kMaxBins];
  const int f = blockIdx.x;
  const ScanItem& item = items[y];
  const int* used = used_all + item.used_off;
  SplitOut* out = out_all + item.out_off;
  const LeafArgs small = item.small, large = item.large;
  const int has_large = item.has_large, use_subtract = item.use_subtract;
  if (!used[f]) return;
  const FeatDev m = s.feats[f];
  // smaller leaf: raw histogram from its slot, fix, store
  __shared__ long long red[3 * (kScanThreads / 64)];
  // the larger leaf's slot holds its parent's fixed slice: with one bin per
  // thread it is loaded beside the smaller leaf's (one global round trip, not two)
  const bool pre = has_large && use_subtract && m.hist_len <= (int)blockDim.x;  // block-uniform
  ParentBin pb{};
  if (pre) pb = load_parent_bin<FORM>(s, m, large.slot);
  load_slice_raw<FORM>(s, m, small.slot, p, S[0]);
  __syncthreads();
  fix_slice<FORM>(m, small, S[0], red);
  store_slice<FORM>(s, m, small.slot, S[0]);
  if (has_large) {
    if (use_subtract) {
      if (pre) subtract_parent_bin<FORM>(s, m, large.slot, pb, S[0], S[1]);
      else subtract_slice<FORM>(s, m, large.slot, S[0], S[1]);
      __syncthreads();
    } else {
      load_slice_raw<FORM>(s, m, large.slot, p, S[1]);
      __syncthreads();
      fix_slice<FORM>(m, large, S[1], red);
    }
    store_slice<FORM>(s, m, large.slot, S[1]);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // waves 0 and 1 search the smaller and the larger leaf; the others only help
  // with the reductions above and take part in the barriers below
  const bool enabled = wave == 0 || (wave == 1 && has_large);
  const int wl = wave < 2 ? wave : 0;
  const LeafArgs a = wl == 0 ? small : large;  // a copy: a reference would put both in scratch
  SplitOut o = SplitOut{};
  o.gain = kMinScore;
  o.default_left = 1;
  o.feature = m.real;
  const long long tg = a.int_sum >> 32, th = a.int_sum & 0xffffffffLL;
  if (m.bin_type == 1) {
    // the ctr sort uses the whole block, one leaf at a time; the scan runs on
    // lane 0 and leaves its result in LDS (a result in registers handed to the
    // call would live in per-lane scratch)
    __shared__ SplitOut cat_out[2];
    cat_sort<FORM>(m, p, small, S[0], elig[0]);
    if (has_large) cat_sort<FORM>(m, p, large, S[1], elig[1]);
    if (threadIdx.x < 2) cat_out[threadIdx.x] = o;
    __syncthreads();
    if (enabled && lane == 0)
      search_categorical<FORM == FORM_QUANT>(S[wl].g, S[wl].h, S[wl].ig, S[wl].ih, m, p, a.sum_g, a.sum_h, tg, th,
                                             a.num_data, a.parent_output, S[wl].sorted, S[wl].used_bin,
                                             cat_out[wl]);
    __syncthreads();
    o = cat_out[wl];
  } else {
    search_numerical_wave<FORM>(m, p, a, S[wl], X[wl], enabled, o);
  }
  if (!enabled) return;
  o.feature = m.real;
  if (o.gain > kMinScore) {
    if (s.hist32) o.exact_left = o.exact_right = -1;  // the partition counts them
    else exact_counts(m, S[wl], a.num_data, o);
  }
  if (lane == 0) out[(wl == 0 ? 0 : s.num_features) + f] = o;
}

// Items blockIdx.y, blockIdx.y + gridDim.y, ...: a planned round's grid is capped
// below the most items a round may hold.
template <int FORM>
__global__ void __launch_bounds__(kScanThreads) k_scan(DevState s, const ScanItem* __restrict__ items, ScanParams p,
                                                       const int* __restrict__ used_all, SplitOut* out_all,
                                                       const RoundHdr* __restrict__ hdr) {
  const int nitems = hdr ? hdr->nscan : (int)gridDim.y;
  for (int y = blockIdx.y; y < nitems; y += gridDim.y) {
    scan_item<FORM>(s, items, y, p, used_all, out_all);
    __syncthreads();  // the LDS slices are reused by the next item
  }
}

// One wave per (item, side): the fold runs after the search launch, whose
// candidates it reads (a kernel boundary orders them across the XCDs' L2s).
// With a stamp, every wave makes its results visible to the host and counts
// itself done; the last one writes seq to the host-mapped stamp word.
__global__ void __launch_bounds__(64) k_fold(const ScanItem* __restrict__ items, const SplitOut* __restrict__ out_all,
                                             const int* __restrict__ used_all, int F, FoldOut* __restrict__ best,
                                             uint8_t* __restrict__ flags, const RoundHdr* __restrict__ hdr,
                                             unsigned* stamp, unsigned seq, unsigned* ticket) {
  const int k = blockIdx.x >> 1, side = blockIdx.x & 1;
  if (!(hdr && k >= hdr->nscan)) fold_side(items[k], k, side, out_all, used_all, F, best, flags);
  if (!stamp) return;
  __threadfence_system();  // this wave's results reach host memory before it counts itself done
  if ((threadIdx.x & 63) == 0 && atomicAdd(ticket, 1u) == gridDim.x - 1u) {
    *ticket = 0u;  // every wave has counted: the next launch starts at zero
    __threadfence_system();
    __hip_atomic_store(stamp, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------------
// Device round planner: the host loop of batched growth, one block per round
// (see launch_plan). Leaves' best splits are ranked by better_than
// (split_info.rs:67-85) with the leaf index last, which is argmax_split's first
// maximum (split_info.rs:88-96) and the stable order of the frontier sort.
// ---------------------------------------------------------------------------
constexpr int kPlanThreads = 256;
static_assert(kPlanMaxLeaves <= kPlanThreads, "one thread per frontier item");
static_assert(kPlanMaxLeaves % 64 == 0, "the commit replay holds kPlanMaxLeaves / 64 leaves per lane");

__device__ __forceinline__ bool plan_better(double ga, int ka, int la, double gb, int kb, int lb) {
  if (ga != gb) return ga > gb;
  if (ka != kb) return ka < kb;
  return la < lb;
}

__device__ __forceinline__ int plan_key(const SplitOut& o) { return o.feature == -1 ? INT_MAX : o.feature; }

__global__ void __launch_bounds__(kPlanThreads) k_plan(PlanState* __restrict__ st, PlanNode* __restrict__ nodes,
                                                       uint8_t* __restrict__ node_flags,
                                                       const FoldOut* __restrict__ fold,
                                                       const uint8_t* __restrict__ fold_flags,
                                                       const FeatDev* __restrict__ feats,
                                                       const int* __restrict__ bytree, PlanParams pp,
                                                       PartItem* __restrict__ pitems, HistItem* __restrict__ hitems,
                                                       ScanItem* __restrict__ sitems, ReduceItem* __restrict__ ritems,
                                                       int* __restrict__ used) {
  __shared__ double lgain[kPlanMaxLeaves], cgain[2][kPlanMaxLeaves];
  __shared__ int lkey[kPlanMaxLeaves], lnode[kPlanMaxLeaves], order[kPlanMaxLeaves], cand[kPlanMaxLeaves];
  __shared__ int ckey[2][kPlanMaxLeaves], cnode[2][kPlanMaxLeaves];
  __shared__ unsigned char lexp[kPlanMaxLeaves];
  __shared__ int scan_node[kPlanMaxLeaves], commit_leaf[kPlanMaxLeaves], commit_node[kPlanMaxLeaves];
  __shared__ int sh_nl, sh_done, sh_first, sh_ncommit;
  __shared__ double sbg[kPlanThreads];
  __shared__ int sbk[kPlanThreads], sbi[kPlanThreads];
  __shared__ int scratch[32];
  const int tid = threadIdx.x, F = pp.F;
  if (st->done) {  // block-uniform: the tree is complete, no work
    if (tid == 0) {
      st->hdr = RoundHdr{};
      st->nabsorb = 0;
    }
    return;
  }
  // (1) the last round's folded splits go to the nodes that were searched: one
  // thread per (item, side), its loads independent of each other
  const int na = st->nabsorb;
  constexpr int kWords = (int)(sizeof(FoldOut) / 4);
  if (tid < 2 * na) {
    const int nd = st->absorb[tid];
    if (nd >= 0) {
      uint32_t w[kWords];
#pragma unroll
      for (int i = 0; i < kWords; ++i) w[i] = ((const uint32_t*)(fold + tid))[i];
#pragma unroll
      for (int i = 0; i < kWords; ++i) ((uint32_t*)&nodes[nd].best)[i] = w[i];
      for (int f = 0; f < F; ++f) node_flags[(size_t)nd * F + f] = fold_flags[(size_t)tid * F + f];
    }
  }
  __syncthreads();
  // (2) every leaf's best split, and for an expanded leaf its children's: a leaf
  // committed in this round becomes its two children, which were searched but
  // not expanded, so the commits below read LDS only
  int nl = st->num_leaves;
  for (int l = tid; l < nl; l += kPlanThreads) {
    const int x = st->leaf_node[l];
    lnode[l] = x;
    lgain[l] = nodes[x].best.o.gain;
    lkey[l] = plan_key(nodes[x].best.o);
    const int e = nodes[x].expanded;
    lexp[l] = (unsigned char)e;
    if (e) {
      for (int side = 0; side < 2; ++side) {
        const int c = nodes[x].child[side];
        cnode[side][l] = c;
        cgain[side][l] = nodes[c].best.o.gain;
        ckey[side][l] = plan_key(nodes[c].best.o);
      }
    }
  }
  __syncthreads();
  // The commits in the reference's order while the best leaf has been expanded.
  // A committed leaf's children are not expanded, so they are never committed
  // in this round: the committed leaves are a prefix of the current leaves in
  // better_than order, and the walk stops at the first position where a child
  // made so far, or the next leaf itself if unexpanded, is the best leaf. The
  // leaves are ranked once (one thread each); lane 0 walks. With a NaN gain
  // (never better nor worse) the order is not total and the reference's plain
  // scan is replayed commit by commit instead.
  bool nan = false;
  for (int l = tid; l < nl; l += kPlanThreads)
    nan |= lgain[l] != lgain[l] || (lexp[l] && (cgain[0][l] != cgain[0][l] || cgain[1][l] != cgain[1][l]));
  const int any_nan = __syncthreads_or(nan ? 1 : 0);
  if (!any_nan) {
    // rank every leaf (one thread each)
    if (tid < nl) {
      const double g = lgain[tid];
      const int k = lkey[tid];
      int r = 0;
#pragma unroll 8
      for (int m = 0; m < nl; ++m) r += (m != tid && plan_better(lgain[m], lkey[m], m, g, k, tid)) ? 1 : 0;
      order[r] = tid;
    }
    __syncthreads();
    // position i holds leaf order[i] and, were it committed as commit i, its two
    // children: the left one takes its index, the right one index nl + i; cb[i]
    // is the best child over positions 0..i (an inclusive scan of better_than)
    const int nl0 = nl;
    double cbg = 0.0;
    int cbk = INT_MAX, cbi = -1;
    if (tid < nl0) {
      const int l = order[tid];
      if (lexp[l]) {
        cbg = cgain[0][l];
        cbk = ckey[0][l];
        cbi = l;
        if (plan_better(cgain[1][l], ckey[1][l], nl0 + tid, cbg, cbk, cbi)) {
          cbg = cgain[1][l];
          cbk = ckey[1][l];
          cbi = nl0 + tid;
        }
      }
    }
    for (int o = 1; o < kPlanThreads; o <<= 1) {
      sbg[tid] = cbg;
      sbk[tid] = cbk;
      sbi[tid] = cbi;
      __syncthreads();
      if (tid >= o && sbi[tid - o] >= 0 &&
          (cbi < 0 || plan_better(sbg[tid - o], sbk[tid - o], sbi[tid - o], cbg, cbk, cbi))) {
        cbg = sbg[tid - o];
        cbk = sbk[tid - o];
        cbi = sbi[tid - o];
      }
      __syncthreads();
    }
    sbg[tid] = cbg;
    sbk[tid] = cbk;
    sbi[tid] = cbi;
    __syncthreads();
    // the walk stops at the first position i where the leaf budget is full, the
    // best leaf (a child of positions < i, or order[i]) has gain <= 0, or it is
    // not expanded (a child never is); positions before it are committed
    int stop = kPlanThreads, why = 0;  // why: 1 done, 2 an unexpanded best leaf
    if (tid <= nl0) {
      const int i = tid;
      const bool have_leaf = i < nl0;
      const int l = have_leaf ? order[i] : 0;
      const bool child_top = i > 0 && sbi[i - 1] >= 0 &&
                             (!have_leaf || plan_better(sbg[i - 1], sbk[i - 1], sbi[i - 1], lgain[l], lkey[l], l));
      if (nl0 + i >= pp.num_leaves || (child_top ? sbg[i - 1] : lgain[l]) <= 0.0) {
        stop = i;
        why = 1;
      } else if (child_top || !lexp[l]) {
        stop = i;
        why = 2;
      }
    }
    // first stopping position (position nl0 always stops: only children remain)
    const int first = block_min<kPlanThreads>(stop * 4 + why, scratch);
    const int ncm = first >> 2;
    if (tid < ncm) {
      const int l = order[tid];
      commit_leaf[tid] = l;
      commit_node[tid] = lnode[l];
    }
    __syncthreads();
    if (tid < ncm) {
      const int l = order[tid], right = nl0 + tid;
      const int c0 = cnode[0][l], c1 = cnode[1][l];
      const double g0 = cgain[0][l], g1 = cgain[1][l];
      const int k0 = ckey[0][l], k1 = ckey[1][l];
      lnode[l] = c0;
      lgain[l] = g0;
      lkey[l] = k0;
      lexp[l] = 0;
      lnode[right] = c1;
      lgain[right] = g1;
      lkey[right] = k1;
      lexp[right] = 0;
    }
    if (tid == 0) {
      sh_done = (first & 3) == 1 ? 1 : 0;
      sh_nl = nl0 + ncm;
      sh_first = st->ncommit;
      sh_ncommit = st->ncommit + ncm;
    }
  } else if (tid < 64) {
    // one wave, commit by commit: lane 0 scans for the first maximum
    const int lane = tid;
    int done = 0, ncommit = st->ncommit, first_commit = ncommit;
    for (;;) {
      if (nl >= pp.num_leaves) {
        done = 1;
        break;
      }
      int b = 0;
      if (lane == 0)
        for (int l = 1; l < nl; ++l) {
          const double x = lgain[l], y = lgain[b];
          if (x != y ? x > y : lkey[l] < lkey[b]) b = l;
        }
      const int bl = __shfl(b, 0, 64);
      if (lgain[bl] <= 0.0) {
        done = 1;
        break;
      }
      if (!lexp[bl]) break;
      const int right = nl;
      if (lane == 0) {
        commit_leaf[ncommit - first_commit] = bl;
        commit_node[ncommit - first_commit] = lnode[bl];
        lnode[right] = cnode[1][bl];
        lgain[right] = cgain[1][bl];
        lkey[right] = ckey[1][bl];
        lexp[right] = 0;
        lnode[bl] = cnode[0][bl];
        lgain[bl] = cgain[0][bl];
        lkey[bl] = ckey[0][bl];
        lexp[bl] = 0;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      ++ncommit;
      ++nl;
    }
    if (lane == 0) {
      sh_done = done;
      sh_nl = nl;
      sh_first = first_commit;
      sh_ncommit = ncommit;
    }
  }
  __syncthreads();
  nl = sh_nl;
  // the commits and the leaves' nodes to the state
  for (int c = sh_first + tid; c < sh_ncommit; c += kPlanThreads) {
    st->commit[2 * c] = commit_leaf[c - sh_first];
    st->commit[2 * c + 1] = commit_node[c - sh_first];
  }
  for (int l = tid; l < nl; l += kPlanThreads) st->leaf_node[l] = lnode[l];
  if (tid == 0) {
    st->ncommit = sh_ncommit;
    st->num_leaves = nl;
  }
  // (3) the frontier: unexpanded leaves with positive gain, best first
  int is_cand = 0;
  if (tid < nl && !sh_done) is_cand = !lexp[tid] && lgain[tid] > 0.0;
  int cpos, unused, ncand, unused2;
  block_scan2<kPlanThreads>(is_cand, 0, &cpos, &unused, &ncand, &unused2, scratch);
  if (is_cand) cand[cpos] = tid;
  __syncthreads();
  if (tid < ncand) {
    const int l = cand[tid];
    const double g = lgain[l];
    const int k = lkey[l];
    int r = 0;
#pragma unroll 8
    for (int i = 0; i < ncand; ++i) {
      const int m = cand[i];
      r += (m != l && plan_better(lgain[m], lkey[m], m, g, k, l)) ? 1 : 0;
    }
    order[r] = l;
  }
  __syncthreads();
  const int ni = sh_done ? 0 : min(ncand, pp.num_leaves - nl);
  // (4) the round's items, one thread per expanded leaf: partition item, two
  // child nodes, and when the children are searchable (before_find_best_split's
  // max_depth and min_data_in_leaf rules, serial.rs:779-799) a search item for
  // the pair, a histogram item for the smaller child (its rows are contiguous
  // after the partition) and its reduce item
  const int j = tid;
  int x = -1, lc = 0, rc = 0, depth = 0, tiles = 0, srch = 0, hrows = 0, cnt = 0;
  Segment seg{0, 0, 0};
  if (j < ni) {
    x = lnode[order[j]];
    seg = nodes[x].seg;
    lc = nodes[x].best.o.exact_left;
    rc = nodes[x].best.o.exact_right;
    depth = nodes[x].depth + 1;
    cnt = seg.count;
    tiles = cnt > 0 ? (cnt + pp.part_tile - 1) / pp.part_tile : 0;
    srch = !(pp.max_depth > 0 && depth >= pp.max_depth) &&
           !(rc < pp.min_data_in_leaf * 2 && lc < pp.min_data_in_leaf * 2);
    hrows = srch ? min(lc, rc) : 0;
  }
  int first_task, sidx, ntasks, nscan, row0, hidx, htotal, nhist;
  block_scan2<kPlanThreads>(tiles, srch, &first_task, &sidx, &ntasks, &nscan, scratch);
  block_scan2<kPlanThreads>(hrows, hrows > 0 ? 1 : 0, &row0, &hidx, &htotal, &nhist, scratch);
  const int part_rows = block_sum<kPlanThreads>(cnt, scratch);
  // histogram tiling of the round: as the host plans it
  int nblocks = 0, rpb = 0;
  if (htotal > 0) {
    const long long nb = hist_row_blocks(htotal, pp.hist_min_rows, pp.hist_max_blocks, pp.rows_per_flush);
    rpb = (int)(((long long)htotal + nb - 1) / nb);
    nblocks = (htotal + rpb - 1) / rpb;
  }
  const int num_nodes = st->num_nodes, next_slot = st->next_slot;
  // tables too small for this round (they are sized so that this cannot happen):
  // nothing is written, the tree stops and the host reports it
  if (num_nodes + 2 * ni > pp.max_nodes || next_slot + nscan > pp.num_slots || nblocks + nhist > pp.max_partials) {
    if (tid == 0) {
      st->done = 1;
      st->pad = 1;
      st->nabsorb = 0;
      st->hdr = RoundHdr{};
    }
    return;
  }
  if (j < ni) {
    const SplitOut o = nodes[x].best.o;
    const int inner = nodes[x].best.inner, xslot = nodes[x].slot;
    const int dst = seg.buf == 1 ? 2 : 1;
    const int base = num_nodes + 2 * j;
    const bool left_small = lc < rc;
    const int sm_slot = srch ? next_slot + sidx : -1;
    const int lg_slot = srch ? xslot : -1;
    // the children (side 0 left, 1 right) as LeafSplits of the split (serial.rs:1685-1876)
    LeafArgs kid[2];
    Segment kseg[2] = {Segment{dst, seg.begin, lc}, Segment{dst, seg.begin + lc, rc}};
    kid[0] = LeafArgs{-1, left_small ? sm_slot : lg_slot, lc, o.left_sum_gradient, o.left_sum_hessian, o.left_output,
                      o.left_int_sum, 0, 0};
    kid[1] = LeafArgs{-1, left_small ? lg_slot : sm_slot, rc, o.right_sum_gradient, o.right_sum_hessian,
                      o.right_output, o.right_int_sum, 0, 0};
    for (int side = 0; side < 2; ++side) {
      PlanNode& c = nodes[base + side];
      c.seg = kseg[side];
      c.depth = depth;
      c.slot = kid[side].slot;
      c.expanded = 0;
      c.kids_searched = 0;
      c.child[0] = c.child[1] = -1;
      c.num_data = kid[side].num_data;
      c.pad = 0;
      c.sum_g = kid[side].sum_g;
      c.sum_h = kid[side].sum_h;
      c.weight = kid[side].parent_output;
      c.int_sum = kid[side].int_sum;
      c.best.o = SplitOut{};
      c.best.o.gain = kMinScore;
      c.best.o.feature = -1;
      c.best.o.default_left = 1;
      c.best.inner = -1;
    }
    nodes[x].child[0] = base;
    nodes[x].child[1] = base + 1;
    nodes[x].expanded = 1;
    nodes[x].kids_searched = srch;
    const FeatDev m = feats[inner];
    PartItem& pi = pitems[j];
    pi.seg = seg;
    pi.dst_buf = dst;
    pi.split_group = m.group;
    pi.left_total = lc;
    pi.hist_side = -1;
    pi.first_task = first_task;
    pi.ntasks = tiles;
    pi.chunk = pp.part_tile;
    pi.pad1 = 0;
    pi.rule = m.bin_type == 1 ? make_rule_categorical(m, o.cat_bits)
                              : make_rule_numerical(m, (uint32_t)o.threshold, o.default_left != 0);
    if (srch) {
      const int sm_side = left_small ? 0 : 1;
      ScanItem si{};
      si.small = kid[sm_side];
      si.large = kid[1 - sm_side];
      if (hrows > 0) {
        const int b0 = row0 / rpb, b1 = (row0 + hrows - 1) / rpb;
        si.small.part_first = b0 + hidx;
        si.small.nblk = b1 - b0 + 1;
        hitems[hidx] = HistItem{kseg[sm_side], row0, 0};
      }
      si.has_large = 1;
      si.use_subtract = 1;
      si.used_off = sidx * F;
      si.out_off = sidx * 2 * F;
      sitems[sidx] = si;
      ritems[sidx] = ReduceItem{sm_slot, si.small.part_first, si.small.nblk, 0};
      st->absorb[2 * sidx] = base + sm_side;
      st->absorb[2 * sidx + 1] = base + 1 - sm_side;
      scan_node[sidx] = x;
    }
  }
  __syncthreads();
  // every search item's features: the expanded node's splittable flags (its
  // children inherit them, serial.rs:838-852) within the tree's feature sample
  for (int i = tid; i < nscan * F; i += kPlanThreads) {
    const int k = i / F, f = i - k * F;
    used[i] = node_flags[(size_t)scan_node[k] * F + f] && bytree[f] ? 7 : 0;
  }
  if (tid == 0) {
    st->done = (sh_done || ni == 0) ? 1 : 0;
    st->num_nodes = num_nodes + 2 * ni;
    st->next_slot = next_slot + nscan;
    st->nabsorb = nscan;
    st->rounds += ni > 0 ? 1 : 0;
    st->search_rounds += nscan > 0 ? 1 : 0;
    st->hist_rows += htotal;
    st->part_rows += part_rows;
    st->hdr = RoundHdr{ni, ntasks, nhist, htotal, nblocks, rpb, nscan, 0};
  }
}

// ---------------------------------------------------------------------------
// A round's items from the partition's counts (quantized 32-bit bins, whose
// histograms count no rows): one thread per split. The children's counts are
// the final cursors of the partition that just ran; the items follow the host's
// plan for known counts: the search item of a searchable pair (smaller child
// first, its histogram built, the larger one by subtraction in the parent's
// slot), the smaller child's histogram item (its rows are contiguous after the
// partition), its reduce item, and the histogram tiling.
// ---------------------------------------------------------------------------
// The derivation of one round over NT threads (ni <= NT), from the partition's
// cursors cur (read with coherent loads: the partition's last block runs it
// while other blocks' cursor atomics have just completed).
template <int NT>
__device__ void derive_children(const PartItem* __restrict__ pitems, int ni, const unsigned* cur,
                                const ChildSpec* __restrict__ specs, const ChildParams& cp,
                                HistItem* __restrict__ hitems, ScanItem* __restrict__ sitems,
                                ReduceItem* __restrict__ ritems, RoundHdr* __restrict__ hdr, int* __restrict__ counts) {
  __shared__ int scratch[32];
  __shared__ int status;
  const int j = threadIdx.x;
  if (j == 0) status = 0;
  __syncthreads();
  int lc = 0, rc = 0, srch = 0, hrows = 0;
  Segment seg{0, 0, 0};
  int dst = 0;
  if (j < ni) {
    seg = pitems[j].seg;
    dst = pitems[j].dst_buf;
    lc = (int)__hip_atomic_load(cur + 2 * j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    rc = (int)__hip_atomic_load(cur + 2 * j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    counts[2 * j] = lc;
    counts[2 * j + 1] = rc;
    if (lc < 0 || rc < 0 || (long long)lc + rc != seg.count) {
      atomicOr(&status, 2);
    } else {
      const int depth = specs[j].depth;
      srch = !(cp.max_depth > 0 && depth >= cp.max_depth) &&
             !(rc < cp.min_data_in_leaf * 2 && lc < cp.min_data_in_leaf * 2);
      if (srch && cp.part_hist && pitems[j].hist_side < 0) atomicOr(&status, 4);  // searched but not built
      hrows = srch && !cp.part_hist ? min(lc, rc) : 0;
    }
  }
  int sidx, row0, nscan, htotal, hidx, unused0, nhist, unused1;
  block_scan2<NT>(srch, hrows, &sidx, &row0, &nscan, &htotal, scratch);
  block_scan2<NT>(hrows > 0 ? 1 : 0, 0, &hidx, &unused0, &nhist, &unused1, scratch);
  int nblocks = 0, rpb = 0;
  if (htotal > 0) {  // as the host cuts a build of htotal rows
    const long long nb = hist_row_blocks(htotal, cp.hist_min_rows, cp.hist_max_blocks, cp.rows_per_flush);
    rpb = (int)(((long long)htotal + nb - 1) / nb);
    nblocks = (htotal + rpb - 1) / rpb;
  }
  if (nblocks + nhist > cp.max_partials) {  // the host sizes the partials so that this cannot happen
    if (j == 0) status |= 32;
    nscan = nhist = htotal = nblocks = 0;
    srch = 0;
  }
  __syncthreads();
  if (status != 0) {  // no work: the host reads the status and stops
    if (j == 0) {
      counts[2 * ni] = status;
      *hdr = RoundHdr{};
    }
    return;  // block-uniform
  }
  if (j < ni && srch) {
    const ChildSpec& cs = specs[j];
    // the built side: the smaller child, or the side the partition built (the
    // children are picked by selects: a local array indexed at run time would
    // live in per-lane scratch memory)
    const int sm = cp.part_hist ? pitems[j].hist_side : (lc < rc ? 0 : 1);
    ScanItem si{};
    si.small = cs.kid[sm];
    si.large = cs.kid[1 - sm];
    si.small.num_data = sm ? rc : lc;
    si.large.num_data = sm ? lc : rc;
    si.small.slot = cs.new_slot;
    si.large.slot = cs.parent_slot;
    si.small.part_first = si.large.part_first = 0;
    si.small.nblk = si.large.nblk = 0;
    const Segment sm_seg = sm ? Segment{dst, seg.begin + lc, rc} : Segment{dst, seg.begin, lc};
    if (cp.part_hist) {  // one partial per partition tile of the split
      si.small.part_first = pitems[j].first_task;
      si.small.nblk = pitems[j].ntasks;
    } else if (hrows > 0) {
      const int b0 = row0 / rpb, b1 = (row0 + hrows - 1) / rpb;
      si.small.part_first = b0 + hidx;  // partial of (block b, item i) is slot b + i
      si.small.nblk = b1 - b0 + 1;
      hitems[hidx] = HistItem{sm_seg, row0, 0};
    }
    si.has_large = 1;
    si.use_subtract = 1;
    si.used_off = j * cp.F;
    si.out_off = sidx * 2 * cp.F;
    sitems[sidx] = si;
    ritems[sidx] = ReduceItem{si.small.slot, si.small.part_first, si.small.nblk, 0};
  }
  if (j == 0) {
    counts[2 * ni] = 0;
    *hdr = RoundHdr{ni, 0, nhist, htotal, nblocks, rpb, nscan, 0};
  }
}

__global__ void __launch_bounds__(kChildThreads) k_children(const PartItem* __restrict__ pitems, int ni,
                                                            const unsigned* __restrict__ cur,
                                                            const ChildSpec* __restrict__ specs, ChildParams cp,
                                                            HistItem* __restrict__ hitems,
                                                            ScanItem* __restrict__ sitems,
                                                            ReduceItem* __restrict__ ritems,
                                                            RoundHdr* __restrict__ hdr, int* __restrict__ counts) {
  derive_children<kChildThreads>(pitems, ni, cur, specs, cp, hitems, sitems, ritems, hdr, counts);
}

// ---------------------------------------------------------------------------
// Gradient discretization (gradient_discretizer/mod.rs:116-201)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_max_abs(const float* __restrict__ g, const float* __restrict__ h, int n, unsigned* out,
                                                 long long* zero4) {
  if (zero4 && blockIdx.x == 0 && threadIdx.x < 4) zero4[threadIdx.x] = 0;  // the discretizer's sums, read later
  unsigned mg = 0u, mh = 0u;
  int bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gv = g[i], hv = h[i];
    if (!isfinite(gv) || !isfinite(hv) || hv < 0.0f) bad = 1;
    mg = max(mg, __float_as_uint(fabsf(gv)));
    mh = max(mh, __float_as_uint(fabsf(hv)));
  }
  for (int o = 32; o > 0; o >>= 1) {
    mg = max(mg, (unsigned)__shfl_down((int)mg, o, 64));
    mh = max(mh, (unsigned)__shfl_down((int)mh, o, 64));
    bad |= __shfl_down(bad, o, 64);
  }
  // one atomic per block: same-address atomics from every wave serialize
  __shared__ unsigned red[3][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = mg;
    red[1][w] = mh;
    red[2][w] = (unsigned)bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      mg = max(mg, red[0][i]);
      mh = max(mh, red[1][i]);
      bad |= (int)red[2][i];
    }
    atomicMax(&out[0], mg);
    atomicMax(&out[1], mh);
    if (bad) atomicOr(&out[2], 1u);
  }
}

__global__ void __launch_bounds__(256) k_discretize(const float* __restrict__ g, const float* __restrict__ h, const double* __restrict__ gn,
                             const double* __restrict__ hn, int n, int start, double inv_g, double inv_h,
                             int const_hess, int stochastic, int16_t* __restrict__ packed, long long* sums,
                             const DiscState* __restrict__ dst) {
  if (dst) {  // a device-resident tree: the set-up kernel decided
    if (!dst->ok) return;  // refused: the previous discretization is kept
    start = dst->start;
    inv_g = dst->inv_g;
    inv_h = dst->inv_h;
  }
  long long sg = 0, sh = 0, sa = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double gd = (double)g[i];
    int pos = i + start;
    if (pos >= n) pos -= n;
    double noise, hnoise;
    if (stochastic) {
      noise = gn[pos];
      hnoise = hn[pos];
    } else {
      noise = 0.5;
      hnoise = 0.5;
    }
    const double sn = gd >= 0.0 ? noise : -noise;
    const double pg = gd * inv_g;
    const int8_t qg = sat_i8_dev(pg + sn);
    int8_t qh = 1;
    if (!const_hess) {
      const double ph = (double)h[i] * inv_h;
      qh = sat_i8_dev(ph + hnoise);
    }
    packed[i] = (int16_t)(uint16_t)(((unsigned)(uint8_t)qg << 8) | (unsigned)(uint8_t)qh);
    sg += qg;
    sh += (uint8_t)qh;
    sa += qg < 0 ? -(int)qg : (int)qg;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sg += __shfl_down(sg, o, 64);
    sh += __shfl_down(sh, o, 64);
    sa += __shfl_down(sa, o, 64);
  }
  __shared__ long long red[3][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = sg;
    red[1][w] = sh;
    red[2][w] = sa;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      sg += red[0][i];
      sh += red[1][i];
      sa += red[2][i];
    }
    atomicAdd((unsigned long long*)&sums[0], (unsigned long long)sg);
    atomicAdd((unsigned long long*)&sums[1], (unsigned long long)sh);
    atomicAdd((unsigned long long*)&sums[2], (unsigned long long)sa);
  }
}

// With a constant hessian the hessian slot counts rows (scaled by hessians[0] after
// the reduction), as histogram.rs:226-232 and serial.rs:982-988 do.
__global__ void k_pack_f32(const float* __restrict__ g, const float* __restrict__ h, int n, int const_hess,
                           float2* __restrict__ gh) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    gh[i] = make_float2(g[i], const_hess ? 1.0f : h[i]);
}

// Root segment of a bagged tree (DataPartition::init_with, data_partition.rs:100-107):
// row bag[j] of buffer 0, with its payload patched in from gh0, becomes row j of
// buffer dst_buf, so the bag sits ascending in one segment as the root leaf. Lanes
// take (row, 16-byte chunk) pairs: the writes are coalesced, the reads whole
// records. The bag's rows are also marked in bag_bits (zeroed by the caller) and,
// for the quantized form, their packed pairs summed exactly: sums[0] = sum of the
// int8 gradients, sums[1] = of the uint8 hessians, sums[2] = of |gradient|. A
// row outside [0, num_data) or not above its predecessor sets err bit 8 and is
// written as a zero record, so every later kernel stays inside its bins.
template <int FORM>
__global__ void __launch_bounds__(256) k_gather_bag(DevState s, const int32_t* __restrict__ bag, int nbag,
                                                    int dst_buf, uint32_t* __restrict__ bag_bits,
                                                    long long* __restrict__ sums) {
  using PT = typename Pay<FORM>::T;
  const int stride = s.stride, nc = stride >> 4;
  const int pay_chunk = s.pay_off >> 4, pay_in = s.pay_off & 15;
  uint8_t* dst = rec_of(s, dst_buf);
  long long sg = 0, sh = 0, sa = 0;
  const long long total = (long long)nbag * nc;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(q / nc), c = (int)(q - (long long)j * nc);
    const int row = bag[j];
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (row >= 0 && row < s.num_data && (j == 0 || bag[j - 1] < row)) {
      v = *(const uint4*)(s.rec[0] + (size_t)row * stride + 16 * c);
      if (c == pay_chunk) {
        const PT p = ((const PT*)s.gh0)[row];
        patch_payload(v, pay_in, p);
        if constexpr (FORM == FORM_QUANT) {
          const long long gq = (long long)(int8_t)(p >> 8);
          sg += gq;
          sh += (long long)(uint8_t)(p & 0xff);
          sa += gq < 0 ? -gq : gq;
        }
      }
      if (c == 0) atomicOr(&bag_bits[row >> 5], 1u << (row & 31));
    } else if (c == 0) {
      atomicOr(s.err_flag, 8);
    }
    *(uint4*)(dst + (size_t)j * stride + 16 * c) = v;
  }
  if constexpr (FORM == FORM_QUANT) {
    sg = wave_sum64(sg);
    sh = wave_sum64(sh);
    sa = wave_sum64(sa);
    if ((threadIdx.x & 63) == 0 && (sg | sh | sa) != 0) {
      atomicAdd((unsigned long long*)&sums[0], (unsigned long long)sg);
      atomicAdd((unsigned long long*)&sums[1], (unsigned long long)sh);
      atomicAdd((unsigned long long*)&sums[2], (unsigned long long)sa);
    }
  }
}

// ---------------------------------------------------------------------------
// renew_int_grad_tree_output (serial.rs:634-655) sums each leaf's f32 gradients
// and hessians in f64, one addition at a time in ascending row order.
//
// Exactness certificate: let E be the exponent of the lowest significand bit
// over a leaf's nonzero values, so every value and every partial sum is a
// multiple of 2^E. If n * max|x| < 2^(52+E), every partial sum, in ANY order,
// is below 2^(53+E) in magnitude and hence exactly representable: no addition
// rounds, and the parallel sum equals the sequential one bit for bit. Leaves
// without the certificate (very small values next to large sums, NaN, Inf) are
// summed by the sequential chain below.
// ---------------------------------------------------------------------------

// exponent of the lowest significand bit of a float: INT_MAX for zero, INT_MIN for NaN/Inf
__device__ __forceinline__ int lsb_exponent(float x) {
  const uint32_t b = __float_as_uint(x) & 0x7fffffffu;
  if (b == 0u) return INT_MAX;
  const int be = (int)(b >> 23);
  if (be == 0xff) return INT_MIN;
  return be == 0 ? -149 : be - 150;
}

// Leaf of every row: the row's record in buffer 0 (original order) walks the
// tree from the root, each internal node sending the row left when the node's
// route bitmap has the bit of the row's bin in its group, until it reaches a
// leaf. The nodes sit in LDS; records of up to 64 bytes are held in registers.
// A path is at most num_nodes long, which bounds every walk.
__device__ __forceinline__ uint32_t record_byte(const uint4& c0, const uint4& c1, const uint4& c2, const uint4& c3,
                                                int g) {
  const int k = g >> 4, w = (g >> 2) & 3;
  const uint32_t w0 = lane_word(c0, w), w1 = lane_word(c1, w), w2 = lane_word(c2, w), w3 = lane_word(c3, w);
  const uint32_t x = k == 0 ? w0 : (k == 1 ? w1 : (k == 2 ? w2 : w3));
  return (x >> (8 * (g & 3))) & 0xffu;
}

// Route bitmap of node blockIdx.x: lane b evaluates the node's rule on bin b.
__global__ void __launch_bounds__(256) k_node_bitmaps(const NodeRule* __restrict__ rules, NodeDev* __restrict__ nodes,
                                                      const int* __restrict__ d_count) {
  if (d_count && (int)blockIdx.x >= *d_count) return;  // block-uniform
  const NodeRule& r = rules[blockIdx.x];
  const uint32_t b = threadIdx.x;
  const unsigned long long m = __ballot(route_left(r.rule, b));
  NodeDev& d = nodes[blockIdx.x];
  if ((threadIdx.x & 63) == 0) {
    d.left_bins[2 * (threadIdx.x >> 6)] = (uint32_t)m;
    d.left_bins[2 * (threadIdx.x >> 6) + 1] = (uint32_t)(m >> 32);
  }
  if (threadIdx.x == 0) {
    d.group = r.group;
    d.left = r.left;
    d.right = r.right;
    d.pad = 0;
  }
}

// Whether a row was in the tree's bag (no bitmap: every row was).
__device__ __forceinline__ bool in_bag(const uint32_t* bag_bits, long long r) {
  return !bag_bits || ((bag_bits[r >> 5] >> (r & 31)) & 1u);
}

// Rows of a grid stride walk the tree; two rows per step of each lane, so two
// independent chains of node loads are in flight. A row outside the tree's bag
// gets ~leaf: the leaf the tree sends it to, marked as not trained on.
// What a walk does with each row's leaf besides writing it: nothing, or the
// leaf renewal's per-block partial sums (RenewAcc).
struct NoEmit {
  __device__ __forceinline__ void operator()(long long, int) const {}
};

template <typename NodePtr, typename Emit = NoEmit>
__device__ __forceinline__ void walk_rows(const DevState& s, NodePtr nd, int num_nodes,
                                          const uint32_t* __restrict__ bag_bits, int32_t* __restrict__ leaf_of,
                                          const Emit& emit = Emit()) {
  const int stride = s.stride, nc = stride >> 4;
  const long long n = s.num_data, step = 2ll * gridDim.x * blockDim.x;
  for (long long r0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; r0 < n; r0 += step) {
    const long long r1 = r0 + (long long)gridDim.x * blockDim.x;
    const bool has1 = r1 < n;
    const uint8_t* rec0 = s.rec[0] + (size_t)r0 * stride;
    const uint8_t* rec1 = s.rec[0] + (size_t)(has1 ? r1 : r0) * stride;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    uint4 a0 = z, a1 = z, a2 = z, a3 = z, b0 = z, b1 = z, b2 = z, b3 = z;
    if (nc <= 4) {
      a0 = *(const uint4*)rec0;
      b0 = *(const uint4*)rec1;
      if (nc > 1) {
        a1 = *(const uint4*)(rec0 + 16);
        b1 = *(const uint4*)(rec1 + 16);
      }
      if (nc > 2) {
        a2 = *(const uint4*)(rec0 + 32);
        b2 = *(const uint4*)(rec1 + 32);
      }
      if (nc > 3) {
        a3 = *(const uint4*)(rec0 + 48);
        b3 = *(const uint4*)(rec1 + 48);
      }
    }
    int n0 = 0, n1 = 0;  // current node, or ~leaf once the walk ended
    for (int k = 0; k < num_nodes && (n0 >= 0 || n1 >= 0); ++k) {
      if (n0 >= 0) {
        const int g = nd[n0].group;
        const uint32_t bin = nc <= 4 ? record_byte(a0, a1, a2, a3, g) : (uint32_t)rec0[g];
        n0 = ((nd[n0].left_bins[bin >> 5] >> (bin & 31u)) & 1u) ? nd[n0].left : nd[n0].right;
      }
      if (n1 >= 0) {
        const int g = nd[n1].group;
        const uint32_t bin = nc <= 4 ? record_byte(b0, b1, b2, b3, g) : (uint32_t)rec1[g];
        n1 = ((nd[n1].left_bins[bin >> 5] >> (bin & 31u)) & 1u) ? nd[n1].left : nd[n1].right;
      }
    }
    const int l0 = n0 < 0 ? ~n0 : 0, l1 = n1 < 0 ? ~n1 : 0;
    const int v0 = in_bag(bag_bits, r0) ? l0 : ~l0, v1 = in_bag(bag_bits, r1) ? l1 : ~l1;
    leaf_of[r0] = v0;
    emit(r0, v0);
    if (has1) {
      leaf_of[r1] = v1;
      emit(r1, v1);
    }
  }
}

// The same walk with each lane's two records staged in LDS (records of up to 64
// bytes): a level then reads the group's byte with one LDS load instead of a
// select chain over the record's registers. A lane's records start every
// pstride = stride + 4 bytes, an odd number of words, so the 64 lanes of a wave
// read from 64 different banks. Each lane reads only what it wrote itself.
template <typename Emit = NoEmit>
__device__ __forceinline__ void walk_rows_staged(const DevState& s, const NodeDev* nd, int num_nodes,
                                                 unsigned char* recs, const uint32_t* __restrict__ bag_bits,
                                                 int32_t* __restrict__ leaf_of, const Emit& emit = Emit()) {
  const int stride = s.stride, nw = stride >> 2, pstride = stride + 4;
  unsigned char* my0 = recs + (size_t)threadIdx.x * pstride;
  unsigned char* my1 = recs + (size_t)(blockDim.x + threadIdx.x) * pstride;
  const long long n = s.num_data, step = 2ll * gridDim.x * blockDim.x;
  for (long long r0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; r0 < n; r0 += step) {
    const long long r1 = r0 + (long long)gridDim.x * blockDim.x;
    const bool has1 = r1 < n;
    const uint32_t* src0 = (const uint32_t*)(s.rec[0] + (size_t)r0 * stride);
    const uint32_t* src1 = (const uint32_t*)(s.rec[0] + (size_t)(has1 ? r1 : r0) * stride);
    for (int w = 0; w < nw; ++w) {
      ((uint32_t*)my0)[w] = src0[w];
      ((uint32_t*)my1)[w] = src1[w];
    }
    int n0 = 0, n1 = 0;  // current node, or ~leaf once the walk ended
    for (int k = 0; k < num_nodes && (n0 >= 0 || n1 >= 0); ++k) {
      if (n0 >= 0) {
        const uint32_t bin = my0[nd[n0].group];
        n0 = ((nd[n0].left_bins[bin >> 5] >> (bin & 31u)) & 1u) ? nd[n0].left : nd[n0].right;
      }
      if (n1 >= 0) {
        const uint32_t bin = my1[nd[n1].group];
        n1 = ((nd[n1].left_bins[bin >> 5] >> (bin & 31u)) & 1u) ? nd[n1].left : nd[n1].right;
      }
    }
    const int l0 = n0 < 0 ? ~n0 : 0, l1 = n1 < 0 ? ~n1 : 0;
    const int v0 = in_bag(bag_bits, r0) ? l0 : ~l0, v1 = in_bag(bag_bits, r1) ? l1 : ~l1;
    leaf_of[r0] = v0;
    emit(r0, v0);
    if (has1) {
      leaf_of[r1] = v1;
      emit(r1, v1);
    }
  }
}

// LDS bytes of the staged walk's records: two per lane of a 256-lane block.
__host__ __device__ inline size_t walk_stage_bytes(int stride) { return (size_t)2 * 256 * (stride + 4); }

__global__ void __launch_bounds__(256) k_leaf_of_tree(DevState s, const NodeDev* __restrict__ nodes, int num_nodes,
                                                      int in_lds, const uint32_t* __restrict__ bag_bits,
                                                      int32_t* __restrict__ leaf_of, const int* __restrict__ d_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (d_count) num_nodes = min(*d_count, num_nodes);  // the LDS holds the most
  if (num_nodes == 0) {
    for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < s.num_data;
         row += (long long)gridDim.x * blockDim.x)
      leaf_of[row] = in_bag(bag_bits, row) ? 0 : ~0;
    return;
  }
  if (in_lds) {
    const int words = num_nodes * (int)(sizeof(NodeDev) / 4);
    for (int i = threadIdx.x; i < words; i += blockDim.x) ((uint32_t*)smem)[i] = ((const uint32_t*)nodes)[i];
    __syncthreads();
    if (in_lds == 2)  // records staged after the nodes
      walk_rows_staged(s, (const NodeDev*)smem, num_nodes, smem + (size_t)words * 4, bag_bits, leaf_of);
    else
      walk_rows(s, (const NodeDev*)smem, num_nodes, bag_bits, leaf_of);  // LDS loads
  } else {
    walk_rows(s, nodes, num_nodes, bag_bits, leaf_of);
  }
}

// The leaf renewal's per-block partials of one row (k_renew_par's accumulation):
// LDS per leaf: sum g, sum h (f64), max|g|, max|h| (float bits), min lsb
// exponents. A row outside the bag (a negative leaf) adds nothing. The sums are
// used only where the certificate holds: then every partial sum is exact and the
// order of the atomics does not matter.
__device__ __forceinline__ void renew_add(double* acc, int* meta, int l, float x, float y) {
  atomicAdd(&acc[2 * l], (double)x);
  atomicAdd(&acc[2 * l + 1], (double)y);
  // NaN and Inf compare above every finite magnitude as unsigned bits and
  // also force INT_MIN exponents, so the certificate fails for them
  const int ax = (int)(__float_as_uint(x) & 0x7fffffffu), ay = (int)(__float_as_uint(y) & 0x7fffffffu);
  const int ex = lsb_exponent(x), ey = lsb_exponent(y);
  if (ax > meta[4 * l]) atomicMax(&meta[4 * l], ax);
  if (ay > meta[4 * l + 1]) atomicMax(&meta[4 * l + 1], ay);
  if (ex < meta[4 * l + 2]) atomicMin(&meta[4 * l + 2], ex);
  if (ey < meta[4 * l + 3]) atomicMin(&meta[4 * l + 3], ey);
}

__device__ __forceinline__ void renew_zero(double* acc, int* meta, int num_leaves) {
  for (int i = threadIdx.x; i < num_leaves; i += blockDim.x) {
    acc[2 * i] = 0.0;
    acc[2 * i + 1] = 0.0;
    meta[4 * i] = 0;
    meta[4 * i + 1] = 0;
    meta[4 * i + 2] = INT_MAX;
    meta[4 * i + 3] = INT_MAX;
  }
}

__device__ __forceinline__ void renew_store(const double* acc, const int* meta, int num_leaves,
                                            double* __restrict__ part_sums, int* __restrict__ part_meta) {
  double* ps = part_sums + (size_t)blockIdx.x * num_leaves * 2;
  int* pm = part_meta + (size_t)blockIdx.x * num_leaves * 4;
  for (int i = threadIdx.x; i < 2 * num_leaves; i += blockDim.x) ps[i] = acc[i];
  for (int i = threadIdx.x; i < 4 * num_leaves; i += blockDim.x) pm[i] = meta[i];
}

struct RenewAcc {
  double* acc;
  int* meta;
  const float* g;
  const float* h;
  int num_leaves;
  __device__ __forceinline__ void operator()(long long r, int l) const {
    if ((unsigned)l < (unsigned)num_leaves) renew_add(acc, meta, l, g[r], h[r]);
  }
};

// The row-to-leaf walk with the leaf renewal's partials in the same pass (the
// rows' leaves are at hand, so leaf_of is not read back): the nodes and the
// staged records in LDS as k_leaf_of_tree's staged form, then the renewal's
// per-leaf accumulators; one partial per block (the grid is renew_parts(n)).
__global__ void __launch_bounds__(256) k_leaf_of_tree_renew(DevState s, const NodeDev* __restrict__ nodes,
                                                            int num_nodes, const uint32_t* __restrict__ bag_bits,
                                                            int32_t* __restrict__ leaf_of,
                                                            const float* __restrict__ g,
                                                            const float* __restrict__ h, int num_leaves,
                                                            double* __restrict__ part_sums,
                                                            int* __restrict__ part_meta) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int words = num_nodes * (int)(sizeof(NodeDev) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) ((uint32_t*)smem)[i] = ((const uint32_t*)nodes)[i];
  unsigned char* recs = smem + (size_t)words * 4;
  double* acc = (double*)(recs + ((walk_stage_bytes(s.stride) + 15) & ~(size_t)15));
  int* meta = (int*)(acc + 2 * (size_t)num_leaves);
  renew_zero(acc, meta, num_leaves);
  __syncthreads();
  walk_rows_staged(s, (const NodeDev*)smem, num_nodes, recs, bag_bits, leaf_of,
                   RenewAcc{acc, meta, g, h, num_leaves});
  __syncthreads();
  renew_store(acc, meta, num_leaves, part_sums, part_meta);
}

// Per-block partial sums per leaf, in row order over a grid stride. LDS holds
// per leaf: sum g, sum h (f64), max|g|, max|h| (float bits), min lsb exponents.
__global__ void __launch_bounds__(256) k_renew_par(int n, const int32_t* __restrict__ leaf_of,
                                                   const float* __restrict__ g, const float* __restrict__ h,
                                                   int num_leaves, double* __restrict__ part_sums,
                                                   int* __restrict__ part_meta) {
  extern __shared__ unsigned char smem[];
  double* acc = (double*)smem;                     // [leaf][2]
  int* meta = (int*)(acc + 2 * (size_t)num_leaves);  // [leaf][4]: maxabs g, maxabs h, lsb g, lsb h
  renew_zero(acc, meta, num_leaves);
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int l = leaf_of[i];
    if ((unsigned)l >= (unsigned)num_leaves) continue;  // a row outside the bag (~leaf)
    renew_add(acc, meta, l, g[i], h[i]);
  }
  __syncthreads();
  renew_store(acc, meta, num_leaves, part_sums, part_meta);
}

// One block per leaf: folds the block partials, checks the certificate, writes
// the certified sums and flags the others for the sequential chain.
__global__ void __launch_bounds__(256) k_renew_fix(const Segment* __restrict__ segs, int num_leaves, int nparts,
                                                   const double* __restrict__ part_sums,
                                                   const int* __restrict__ part_meta, double* __restrict__ out,
                                                   int* __restrict__ flags) {
  __shared__ double rs[2][256];
  __shared__ int rm[4][256];
  const int leaf = blockIdx.x;
  double a = 0.0, b = 0.0;
  int mg = 0, mh = 0, eg = INT_MAX, eh = INT_MAX;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    const size_t o = (size_t)k * num_leaves + leaf;
    a += part_sums[2 * o];
    b += part_sums[2 * o + 1];
    mg = max(mg, part_meta[4 * o]);
    mh = max(mh, part_meta[4 * o + 1]);
    eg = min(eg, part_meta[4 * o + 2]);
    eh = min(eh, part_meta[4 * o + 3]);
  }
  rs[0][threadIdx.x] = a;
  rs[1][threadIdx.x] = b;
  rm[0][threadIdx.x] = mg;
  rm[1][threadIdx.x] = mh;
  rm[2][threadIdx.x] = eg;
  rm[3][threadIdx.x] = eh;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int t = threadIdx.x, u = threadIdx.x + w;
      rs[0][t] += rs[0][u];
      rs[1][t] += rs[1][u];
      rm[0][t] = max(rm[0][t], rm[0][u]);
      rm[1][t] = max(rm[1][t], rm[1][u]);
      rm[2][t] = min(rm[2][t], rm[2][u]);
      rm[3][t] = min(rm[3][t], rm[3][u]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)segs[leaf].count;
    auto certified = [&](int maxbits, int e) {
      if (e == INT_MAX) return true;  // only zeros
      if (e == INT_MIN || maxbits >= 0x7f800000) return false;
      const double bound = n * (double)__uint_as_float((uint32_t)maxbits);
      return bound < ldexp(1.0, 52 + e);
    };
    const bool okg = certified(rm[0][0], rm[2][0]), okh = certified(rm[1][0], rm[3][0]);
    out[2 * leaf] = rs[0][0];
    out[2 * leaf + 1] = rs[1][0];
    flags[leaf] = (okg ? 0 : 1) | (okh ? 0 : 2);
  }
}

// Sequential f64 sums of one leaf's gradients and hessians in ascending row
// order (serial.rs:634-655), for the leaves whose certificate failed: the block
// walks leaf_of in passes of kRenewRows rows; every wave compacts the rows of
// the leaf among its 1024 rows in row order (ballot ranks) into LDS, and two
// lanes then add the values one by one (gradients and hessians side by side).
// The chain is bound by the f64 add latency; the scan of leaf_of is the price of
// keeping no row ids.
constexpr int kRenewRows = 4096;
__global__ void __launch_bounds__(256) k_renew_sums(int n, const int32_t* __restrict__ leaf_of,
                                                    const int* __restrict__ order, const float* __restrict__ g,
                                                    const float* __restrict__ h, const int* __restrict__ flags,
                                                    double* out) {
  __shared__ float vg[kRenewRows], vh[kRenewRows];
  __shared__ int wcnt[4];
  const int leaf = order[blockIdx.x];
  const int need = flags ? flags[leaf] : 3;
  if (need == 0) return;  // block-uniform
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  double ag = 0.0, ah = 0.0;
  for (int base = 0; base < n; base += kRenewRows) {
    unsigned long long m[kRenewRows / 256];
    float gv[kRenewRows / 256], hv[kRenewRows / 256];
    int wc = 0;
#pragma unroll
    for (int j = 0; j < kRenewRows / 256; ++j) {
      const int r = base + wave * (kRenewRows / 4) + 64 * j + lane;
      const bool mine = r < n && leaf_of[r] == leaf;
      m[j] = __ballot(mine);
      wc += __popcll(m[j]);
      gv[j] = mine ? g[r] : 0.0f;
      hv[j] = mine ? h[r] : 0.0f;
    }
    if (lane == 0) wcnt[wave] = wc;
    __syncthreads();
    int pos = 0, total = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wave) pos += wcnt[w];
      total += wcnt[w];
    }
#pragma unroll
    for (int j = 0; j < kRenewRows / 256; ++j) {
      if ((m[j] >> lane) & 1ull) {
        const int at = pos + __popcll(m[j] & below);
        vg[at] = gv[j];
        vh[at] = hv[j];
      }
      pos += __popcll(m[j]);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < total; ++i) ag += (double)vg[i];
    if (threadIdx.x == 64)
      for (int i = 0; i < total; ++i) ah += (double)vh[i];
    __syncthreads();
  }
  if (threadIdx.x == 0 && (need & 1)) out[2 * leaf] = ag;
  if (threadIdx.x == 64 && (need & 2)) out[2 * leaf + 1] = ah;
}

// Piecewise sequential sums. A leaf's sequence of values in ascending row order
// is cut into pieces (its rows within consecutive row ranges). Started from the
// chain's running sum s, a piece adds without any rounding when every partial
// sum s + x_1 + ... + x_i is exactly representable: all of them are multiples of
// 2^E, E = min(lowest set bit of s, lowest set bit of the piece's values), and
// none exceeds |s| + sum|x| in magnitude, so |s| + sum|x| < 2^(53 + E) suffices
// (an integer times 2^E below 2^53 in magnitude is a double, E >= -1074 here).
// Then the chain's result after the piece is s + T exactly, T the piece's exact
// sum (exact in any order under the same bound, so the atomics may add it), and
// the piece costs one addition instead of one per value. Pieces that fail the
// bound are added value by value, as the reference's loop does.

// Exponent of the lowest set bit: INT_MAX for zero, INT_MIN for NaN or infinity.
__device__ __forceinline__ int lowest_bit_f32(float x) {
  const uint32_t b = __float_as_uint(x) & 0x7fffffffu;
  if (b == 0u) return INT_MAX;
  const int be = (int)(b >> 23);
  if (be == 0xff) return INT_MIN;
  const uint32_t m = b & 0x7fffffu;
  return be == 0 ? -149 + __builtin_ctz(m) : (be - 150) + __builtin_ctz(m | 0x800000u);
}

__device__ __forceinline__ int lowest_bit_f64(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
  if (b == 0ull) return INT_MAX;
  const int be = (int)(b >> 52);
  if (be == 0x7ff) return INT_MIN;
  const unsigned long long m = b & 0xfffffffffffffull;
  return be == 0 ? -1074 + __builtin_ctzll(m) : (be - 1075) + __builtin_ctzll(m | 0x10000000000000ull);
}

__device__ __forceinline__ void piece_add(double* t, double* a, int* e, float x) {
  if (x == 0.0f) return;  // adds nothing to either sum (a running sum that starts at +0 stays +0)
  atomicAdd(t, (double)x);
  atomicAdd(a, fabs((double)x));
  const int ex = lowest_bit_f32(x);
  if (ex < *e) atomicMin(e, ex);
}

// One block per piece: for every flagged leaf, its values among the piece's rows.
__global__ void __launch_bounds__(256) k_renew_pieces(int n, int piece_rows, const int32_t* __restrict__ leaf_of,
                                                      const float* __restrict__ g, const float* __restrict__ h,
                                                      int num_leaves, const int* __restrict__ flags,
                                                      RenewPiece* __restrict__ pieces, int* work_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (blockIdx.x == 0 && threadIdx.x == 0 && work_count) *work_count = 0;  // k_renew_guess appends after this launch
  RenewPiece* acc = (RenewPiece*)smem;
  int* need = (int*)(acc + num_leaves);
  int any = 0;
  for (int l = threadIdx.x; l < num_leaves; l += blockDim.x) {
    const int f = flags ? flags[l] : 3;
    need[l] = f;
    any |= f;
    acc[l] = RenewPiece{0.0, 0.0, 0.0, 0.0, INT_MAX, INT_MAX};
  }
  if (!__syncthreads_or(any)) return;  // no leaf needs the chain: block-uniform
  const long long r0 = (long long)blockIdx.x * piece_rows;
  const long long r1 = min(r0 + (long long)piece_rows, (long long)n);
  for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const int l = leaf_of[r];
    if ((unsigned)l >= (unsigned)num_leaves) continue;  // a row outside the bag (~leaf)
    const int f = need[l];
    if (f & 1) piece_add(&acc[l].tg, &acc[l].ag, &acc[l].eg, g[r]);
    if (f & 2) piece_add(&acc[l].th, &acc[l].ah, &acc[l].eh, h[r]);
  }
  __syncthreads();
  RenewPiece* out = pieces + (size_t)blockIdx.x * num_leaves;
  for (int l = threadIdx.x; l < num_leaves; l += blockDim.x)
    if (need[l]) out[l] = acc[l];
}

__device__ __forceinline__ float lane_f32(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

__device__ __forceinline__ double lane_f64(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ long long lane_i64(long long v, int k) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), k);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), k);
  return ((long long)hi << 32) | (long long)(unsigned)lo;
}

// Value-by-value addition of one leaf's values among rows [r0, r1), in row order;
// called by a whole wave, every lane holding the same s. Up to 4096 rows are
// loaded at once (lane i holds rows 4i .. 4i + 3 of every 256: 16-byte loads where
// the rows are aligned and whole), then their matches are added in row order.
__device__ double chain_rows(double s, long long r0, long long r1, int leaf, const int32_t* __restrict__ leaf_of,
                             const float* __restrict__ x) {
  const int lane = (int)threadIdx.x & 63;
  constexpr int kSteps = 16;  // 256-row steps per batch
  const bool aligned = (((uintptr_t)x | (uintptr_t)leaf_of) & 15u) == 0u && (r0 & 3) == 0;
  for (long long b = r0; b < r1; b += 256 * kSteps) {
    int4 lv[kSteps];
    float4 xv[kSteps];
#pragma unroll
    for (int j = 0; j < kSteps; ++j) {
      const long long c = b + 256 * j, r = c + 4 * lane;
      if (aligned && c + 256 <= r1) {
        lv[j] = *(const int4*)(leaf_of + r);
        xv[j] = *(const float4*)(x + r);
      } else {
        // rows past r1 read as leaf -1, which no leaf index equals
        lv[j] = make_int4(r < r1 ? leaf_of[r] : -1, r + 1 < r1 ? leaf_of[r + 1] : -1,
                          r + 2 < r1 ? leaf_of[r + 2] : -1, r + 3 < r1 ? leaf_of[r + 3] : -1);
        xv[j] = make_float4(r < r1 ? x[r] : 0.0f, r + 1 < r1 ? x[r + 1] : 0.0f, r + 2 < r1 ? x[r + 2] : 0.0f,
                            r + 3 < r1 ? x[r + 3] : 0.0f);
      }
    }
#pragma unroll
    for (int j = 0; j < kSteps; ++j) {
      const unsigned long long m0 = __ballot(lv[j].x == leaf), m1 = __ballot(lv[j].y == leaf);
      const unsigned long long m2 = __ballot(lv[j].z == leaf), m3 = __ballot(lv[j].w == leaf);
      unsigned long long mm = m0 | m1 | m2 | m3;
      while (mm) {  // wave-uniform: lanes in order, each lane's four rows in order
        const int k = __builtin_ctzll(mm);
        mm &= mm - 1ull;
        if ((m0 >> k) & 1ull) s += (double)lane_f32(xv[j].x, k);
        if ((m1 >> k) & 1ull) s += (double)lane_f32(xv[j].y, k);
        if ((m2 >> k) & 1ull) s += (double)lane_f32(xv[j].z, k);
        if ((m3 >> k) & 1ull) s += (double)lane_f32(xv[j].w, k);
      }
    }
  }
  return s;
}

// The exact-piece test of the chain: from s, every partial sum of the piece is
// exactly representable (sum|x| from at most 2^22 atomic adds is below the true
// value by less than a factor 2^-31, so a * (1 + 2^-30) bounds it from above; a
// rounded sum below a power of two means the exact sum is below it too).
__device__ __forceinline__ bool piece_exact(double s, double a, int e, double margin) {
  const int em = min(lowest_bit_f64(s), e);
  return em != INT_MIN && fabs(s) + a * (1.0 + 0x1p-30) * margin < ldexp(1.0, 53 + em);
}

// Binade exponent of a normal double (|s| in [2^eb, 2^(eb+1))), INT_MIN otherwise.
__device__ __forceinline__ int binade_of(double s) {
  const int be = (int)(((unsigned long long)__double_as_longlong(s) >> 52) & 0x7ffull);
  return (be == 0 || be == 0x7ff) ? INT_MIN : be - 1023;
}

// The binade test of the chain: s and every partial sum of the piece, rounded or
// not, stay strictly inside s's binade. The drift of the rounded sums from the
// exact ones is at most half a unit per addition, bounded here by piece_rows units.
__device__ __forceinline__ bool piece_in_binade(double s, double a, int eb, int piece_rows) {
  if (eb == INT_MIN || eb < -1022 || eb > 900 || binade_of(s) != eb) return false;
  const double slack = a * (1.0 + 0x1p-30) + (double)piece_rows * ldexp(1.0, eb - 52);
  return fabs(s) - slack > ldexp(1.0, eb) && fabs(s) + slack < ldexp(1.0, eb + 1);
}

// One block of two waves per leaf (gradients, hessians; each only when flagged):
// the chain's running sum guessed piece by piece from the pieces' sums, and the
// pieces that the exact test may reject listed for k_renew_binade with the binade
// of the guess. The guess only picks pieces and binades: the chain checks both.
__global__ void __launch_bounds__(128) k_renew_guess(int num_pieces, int num_leaves, const int* __restrict__ flags,
                                                     const RenewPiece* __restrict__ pieces,
                                                     RenewBinade* __restrict__ binade, RenewWork* __restrict__ work,
                                                     int* work_count, int work_cap) {
  const int leaf = blockIdx.x, comp = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int need = flags ? flags[leaf] : 3;
  if (!((need >> comp) & 1)) return;  // wave-uniform
  double s = 0.0;
  for (int p0 = 0; p0 < num_pieces; p0 += 64) {
    double t = 0.0, a = 0.0;
    int e = INT_MAX;
    if (p0 + lane < num_pieces) {
      const RenewPiece& pc = pieces[(size_t)(p0 + lane) * num_leaves + leaf];
      t = comp ? pc.th : pc.tg;
      a = comp ? pc.ah : pc.ag;
      e = comp ? pc.eh : pc.eg;
    }
    // the running sum at each piece's start: the batch's start plus an exclusive
    // prefix of the pieces' sums (any order: it is a guess)
    const bool valid = e != INT_MAX && e != INT_MIN;
    const double tv = valid ? t : 0.0;
    double incl = tv;
    for (int o = 1; o < 64; o <<= 1) {
      const double up = __shfl_up(incl, o, 64);
      if (lane >= o) incl += up;
    }
    const double mine_s = s + (incl - tv);
    // a margin of 4 on the test: the guess drifts from the chain's sum by roundings
    const int mine_mark = valid && !piece_exact(mine_s, a, e, 4.0) ? 1 : 0;
    const int mine_eb = mine_mark ? binade_of(mine_s) : INT_MIN;
    s += __shfl(incl, 63, 64);
    if (p0 + lane < num_pieces) {
      int eb = INT_MIN, wi = -1;
      if (mine_mark) {
        const int w = atomicAdd(work_count, 1);
        if (w < work_cap) {  // else no room: the chain adds this piece value by value
          work[w] = RenewWork{p0 + lane, leaf, comp, mine_eb, mine_s};
          eb = mine_eb;
          wi = w;
        }
      }
      RenewBinade& bn = binade[((size_t)(p0 + lane) * num_leaves + leaf) * 2 + comp];
      bn.eb = eb;
      bn.work = wi;
    }
  }
}

// The 2-state summary of a run of additions inside one binade: the units added
// and the parity of s/u after the run, for s/u even (0) and odd (1) before it.
struct ParityMap {
  long long s0, s1;
  int o0, o1;
};

__device__ __forceinline__ ParityMap parity_then(const ParityMap& f, const ParityMap& g) {
  ParityMap r;
  r.s0 = f.s0 + (f.o0 ? g.s1 : g.s0);
  r.s1 = f.s1 + (f.o1 ? g.s1 : g.s0);
  r.o0 = f.o0 ? g.o1 : g.o0;
  r.o1 = f.o1 ? g.o1 : g.o0;
  return r;
}

// One addition of x (a value of the leaf) with s/u = k: s + x rounds to
// (k + q + round(f)) u, x/u = q + f; a tie (f = 1/2) goes to the even one of k + q
// and k + q + 1. bad is set when x/u is not a moderate number (not this binade's).
__device__ __forceinline__ ParityMap parity_of(bool mine, float x, double inv_u, bool& bad) {
  ParityMap m{0, 0, 0, 1};
  if (!mine) return m;
  const double y = (double)x * inv_u;  // exact: a power-of-two scaling within range
  if (!(fabs(y) < 0x1p62)) {
    bad = true;
    return m;
  }
  const double q = floor(y), f = y - q;
  const long long qi = (long long)q;
  if (f == 0.5) {
    m.s0 = qi + (qi & 1);        // k even: k + q even if q even, else k + q + 1
    m.s1 = qi + ((qi & 1) ^ 1);  // k odd
    m.o0 = m.o1 = 0;             // either way the result is even
  } else {
    const long long a = qi + (f > 0.5 ? 1 : 0);
    m.s0 = m.s1 = a;
    m.o0 = (int)(a & 1);
    m.o1 = (int)((a & 1) ^ 1);
  }
  return m;
}

__device__ __forceinline__ ParityMap parity_shfl_down(const ParityMap& m, int o) {
  ParityMap r;
  r.s0 = __shfl_down(m.s0, o, 64);
  r.s1 = __shfl_down(m.s1, o, 64);
  r.o0 = __shfl_down(m.o0, o, 64);
  r.o1 = __shfl_down(m.o1, o, 64);
  return r;
}

// The lanes' maps in lane order, composed (lane 0 holds the result).
__device__ __forceinline__ ParityMap parity_reduce(ParityMap m) {
  const int lane = (int)threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const ParityMap r = parity_shfl_down(m, o);
    if ((lane & (2 * o - 1)) == 0) m = parity_then(m, r);
  }
  return m;
}

__device__ __forceinline__ ParityMap parity_lane0(const ParityMap& m) {
  ParityMap r;
  const long long a = m.s0, b = m.s1;
  r.s0 = ((long long)__builtin_amdgcn_readlane((int)(a >> 32), 0) << 32) |
         (long long)(unsigned)__builtin_amdgcn_readlane((int)(a & 0xffffffffll), 0);
  r.s1 = ((long long)__builtin_amdgcn_readlane((int)(b >> 32), 0) << 32) |
         (long long)(unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), 0);
  r.o0 = __builtin_amdgcn_readlane(m.o0, 0);
  r.o1 = __builtin_amdgcn_readlane(m.o1, 0);
  return r;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_min_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// One wave per listed piece: the summary of the leaf's values in the piece's
// rows, in row order, for the listed binade (when it has one), and the same for
// each of its kRenewParts sub-ranges with the binade the guessed running sum has
// at the sub-range's start, with each sub-range's sum, sum of magnitudes and
// lowest set bit. A sub-range's sum is exact wherever the chain uses it (the
// exact test bounds every partial sum of any order).
__global__ void __launch_bounds__(256) k_renew_binade(int n, int piece_rows, int num_leaves,
                                                      const int32_t* __restrict__ leaf_of,
                                                      const float* __restrict__ g, const float* __restrict__ h,
                                                      RenewBinade* __restrict__ binade,
                                                      const RenewWork* __restrict__ work, RenewPart* __restrict__ parts,
                                                      const int* work_count, int work_cap) {
  const int lane = (int)threadIdx.x & 63;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  const int total = min(*work_count, work_cap);
  const int part_rows = piece_rows / kRenewParts;  // a multiple of 256
  for (int w = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); w < total; w += waves) {
    const RenewWork it = work[w];
    const int piece = it.piece, leaf = it.leaf, comp = it.comp, eb = it.eb;
    const float* x = comp ? h : g;
    const bool whole = eb != INT_MIN;
    const double inv_u = whole ? ldexp(1.0, 52 - eb) : 0.0;
    const long long r0 = (long long)piece * piece_rows, r1 = min(r0 + (long long)piece_rows, (long long)n);
    const bool aligned = (((uintptr_t)x | (uintptr_t)leaf_of) & 15u) == 0u;
    ParityMap acc{0, 0, 0, 1};
    bool bad = false;
    double s_run = it.s;  // the guessed running sum, for the sub-ranges' binades
    for (int k = 0; k < kRenewParts; ++k) {
      const long long c0 = r0 + (long long)k * part_rows, c1 = min(c0 + (long long)part_rows, r1);
      const int ebk = binade_of(s_run);
      const bool has_k = ebk != INT_MIN && ebk >= -1022 && ebk <= 900;
      const double inv_uk = has_k ? ldexp(1.0, 52 - ebk) : 0.0;
      ParityMap pk{0, 0, 0, 1};
      bool bad_k = !has_k;
      double t = 0.0, a = 0.0;
      int e = INT_MAX;
      for (long long b = c0; b < c1; b += 256) {
        int lv[4];
        float xv[4];
        const long long r = b + 4 * lane;
        if (aligned && b + 256 <= c1) {  // lane i: rows 4i .. 4i + 3 of the 256, in order
          const int4 l4 = *(const int4*)(leaf_of + r);
          const float4 x4 = *(const float4*)(x + r);
          lv[0] = l4.x, lv[1] = l4.y, lv[2] = l4.z, lv[3] = l4.w;
          xv[0] = x4.x, xv[1] = x4.y, xv[2] = x4.z, xv[3] = x4.w;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            lv[q] = r + q < c1 ? leaf_of[r + q] : -1;
            xv[q] = r + q < c1 ? x[r + q] : 0.0f;
          }
        }
        ParityMap mw{0, 0, 0, 1}, mk{0, 0, 0, 1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool mine = lv[q] == leaf;
          const float v = mine ? xv[q] : 0.0f;
          if (whole) mw = parity_then(mw, parity_of(mine, v, inv_u, bad));
          if (has_k) mk = parity_then(mk, parity_of(mine, v, inv_uk, bad_k));
          if (mine && v != 0.0f) {
            t += (double)v;
            a += fabs((double)v);
            e = min(e, lowest_bit_f32(v));
          }
        }
        if (whole) acc = parity_then(acc, parity_lane0(parity_reduce(mw)));
        if (has_k) pk = parity_then(pk, parity_lane0(parity_reduce(mk)));
      }
      t = wave_sum_f64(t);
      a = wave_sum_f64(a);
      e = wave_min_i32(e);
      bad_k = __ballot(bad_k) != 0ull;
      if (lane == 0) parts[(size_t)w * kRenewParts + k] = RenewPart{t, a, pk.s0, pk.s1, e, bad_k ? INT_MIN : ebk};
      s_run += t;
    }
    const bool any_bad = __ballot(bad) != 0ull;
    if (lane == 0) {
      RenewBinade& out = binade[((size_t)piece * num_leaves + leaf) * 2 + comp];
      out.s0 = acc.s0;
      out.s1 = acc.s1;
      out.eb = whole && !any_bad ? eb : INT_MIN;
    }
  }
}

// s = k u (u = 2^(eb-52), k an integer below 2^53 in magnitude) plus a run inside
// s's binade that adds s0 units when k is even, s1 when odd; the result stays in
// the binade, so it is exact.
__device__ __forceinline__ double binade_add(double s, int eb, long long s0, long long s1) {
  const long long k = (long long)(s * ldexp(1.0, 52 - eb));
  return (double)(k + ((k & 1) ? s1 : s0)) * ldexp(1.0, eb - 52);
}

// One block of two waves per leaf: wave 0 chains the gradients, wave 1 the
// hessians (each only when flagged), over the pieces in row order: an exact piece
// adds its sum, a piece inside the running sum's binade with a summary for that
// binade adds its units, any other piece adds its values one by one.
__global__ void __launch_bounds__(128) k_renew_chain(int n, int piece_rows, int num_pieces,
                                                     const int32_t* __restrict__ leaf_of,
                                                     const float* __restrict__ g, const float* __restrict__ h,
                                                     int num_leaves, const int* __restrict__ flags,
                                                     const RenewPiece* __restrict__ pieces,
                                                     const RenewBinade* __restrict__ binade,
                                                     const RenewPart* __restrict__ parts, double* out) {
  const int leaf = blockIdx.x, comp = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int need = flags ? flags[leaf] : 3;
  if (!((need >> comp) & 1)) return;  // wave-uniform
  const float* x = comp ? h : g;
  double s = 0.0;
  for (int p0 = 0; p0 < num_pieces; p0 += 64) {
    double t = 0.0, a = 0.0;
    int e = INT_MAX, beb = INT_MIN, bw = -1;
    long long bs0 = 0, bs1 = 0;
    if (p0 + lane < num_pieces) {
      const size_t at = (size_t)(p0 + lane) * num_leaves + leaf;
      const RenewPiece& pc = pieces[at];
      t = comp ? pc.th : pc.tg;
      a = comp ? pc.ah : pc.ag;
      e = comp ? pc.eh : pc.eg;
      if (binade) {
        const RenewBinade& bn = binade[at * 2 + comp];
        beb = bn.eb;
        bw = bn.work;
        bs0 = bn.s0;
        bs1 = bn.s1;
      }
    }
    // the whole batch at once: every partial sum of the batch's values, started
    // from s, is exact when |s| + sum|x| < 2^(53 + E) over all of them (E the lowest
    // set bit of s and of every value); the pieces' sums then add exactly in any order
    {
      const int eb_all = wave_min_i32(e);
      const bool any_bad = __ballot(e == INT_MIN) != 0ull;
      if (!any_bad) {
        if (eb_all == INT_MAX) continue;  // no nonzero value of the leaf in the batch
        const double a_all = wave_sum_f64(e == INT_MAX ? 0.0 : a);
        if (piece_exact(s, a_all, eb_all, 1.0)) {
          s += wave_sum_f64(e == INT_MAX ? 0.0 : t);
          continue;
        }
        // or the whole batch inside s's binade, every piece summarized for it: the
        // pieces' 2-state maps compose in row order into the batch's
        const int ebs = binade_of(s);
        const bool summarized = e == INT_MAX || beb == ebs;
        if (ebs != INT_MIN && __ballot(!summarized) == 0ull &&
            piece_in_binade(s, a_all, ebs, piece_rows * min(64, num_pieces - p0))) {
          const ParityMap m = e == INT_MAX ? ParityMap{0, 0, 0, 1}
                                           : ParityMap{bs0, bs1, (int)(bs0 & 1), (int)((bs1 + 1) & 1)};
          const ParityMap all = parity_lane0(parity_reduce(m));
          s = binade_add(s, ebs, all.s0, all.s1);
          continue;
        }
      }
    }
    const int cnt = min(64, num_pieces - p0);
    for (int j = 0; j < cnt; ++j) {
      const int ej = __builtin_amdgcn_readlane(e, j);
      if (ej == INT_MAX) continue;  // no nonzero value of the leaf in this piece
      const double aj = lane_f64(a, j);
      if (piece_exact(s, aj, ej, 1.0)) {
        s += lane_f64(t, j);
        continue;
      }
      const int ebj = __builtin_amdgcn_readlane(beb, j);
      if (ej != INT_MIN && piece_in_binade(s, aj, ebj, piece_rows)) {
        s = binade_add(s, ebj, lane_i64(bs0, j), lane_i64(bs1, j));
        continue;
      }
      const long long r0 = (long long)(p0 + j) * piece_rows, r1 = min(r0 + (long long)piece_rows, (long long)n);
      const int wj = __builtin_amdgcn_readlane(bw, j);
      if (wj < 0 || ej == INT_MIN) {
        s = chain_rows(s, r0, r1, leaf, leaf_of, x);
        continue;
      }
      // the same tests on the piece's sub-ranges (lane k < kRenewParts loads the k-th)
      const int part_rows = piece_rows / kRenewParts;
      RenewPart pt{0.0, 0.0, 0, 0, INT_MAX, INT_MIN};
      if (lane < kRenewParts) pt = parts[(size_t)wj * kRenewParts + lane];
      for (int k = 0; k < kRenewParts; ++k) {
        const int ek = __builtin_amdgcn_readlane(pt.e, k);
        if (ek == INT_MAX) continue;
        const double ak = lane_f64(pt.a, k);
        if (piece_exact(s, ak, ek, 1.0)) {
          s += lane_f64(pt.t, k);
          continue;
        }
        const int ebk = __builtin_amdgcn_readlane(pt.eb, k);
        if (ek != INT_MIN && piece_in_binade(s, ak, ebk, part_rows)) {
          s = binade_add(s, ebk, lane_i64(pt.s0, k), lane_i64(pt.s1, k));
          continue;
        }
        const long long c0 = r0 + (long long)k * part_rows;
        s = chain_rows(s, c0, min(c0 + (long long)part_rows, r1), leaf, leaf_of, x);
      }
    }
  }
  if (lane == 0) out[2 * leaf + comp] = s;
}

// score[row] += vals[leaf of row] over the rows of the tree's bag (out_of_bag = 0)
// or over the rows outside it (out_of_bag = 1, whose leaf_of holds ~leaf).
__global__ void k_add_score_rows(int n, const int32_t* __restrict__ leaf_of, const double* __restrict__ vals,
                                 int num_leaves, int out_of_bag, double* __restrict__ score) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = leaf_of[i];
    const int l = out_of_bag ? ~v : v;
    if (l >= 0 && l < num_leaves) score[i] += vals[l];
  }
}

// Softmax gradients (objective/multiclass.rs:68-90, unweighted) for a device-side
// boosting loop: score and the outputs hold class k's rows at [k * n, (k + 1) * n).
// The exponentials are the device's, so the values can differ in the last bit
// from the host libm's; parity tests compute gradients on the host.
__global__ void k_softmax_gradients(const double* __restrict__ score, const float* __restrict__ label, int n,
                                    int num_class, double factor, float* __restrict__ g, float* __restrict__ h) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double wmax = score[i];
    for (int k = 1; k < num_class; ++k) wmax = fmax(score[(size_t)k * n + i], wmax);
    double wsum = 0.0;
    for (int k = 0; k < num_class; ++k) wsum += exp(score[(size_t)k * n + i] - wmax);
    const int lab = (int)label[i];
    for (int k = 0; k < num_class; ++k) {
      const double p = exp(score[(size_t)k * n + i] - wmax) / wsum;
      const double gi = lab == k ? p - 1.0 : p;
      g[(size_t)k * n + i] = (float)gi;
      h[(size_t)k * n + i] = (float)(factor * p * (1.0 - p));
    }
  }
}

// LambdaRank (NDCG) gradients (objective/rank.rs:262-340), one block per query,
// unweighted, no position bias. The query's documents are ranked by score,
// descending, ties in document order (a bitonic sort of (score, index) in LDS),
// as the reference's stable sort does. For every rank i below the truncation
// level, the lanes take the pairs (i, j > i) with different labels; each pair's
// lambda and hessian are the reference's (same sigmoid table, gains, discounts
// and normalisation) and are added to both documents with LDS float atomics. The
// float sums therefore run in another order than the reference's pair loop;
// parity tests compute gradients on the host.
__device__ __forceinline__ bool rank_before(double sa, int ia, double sb, int ib) {
  // padding (index < 0) sorts last; otherwise score descending, then index
  if (ia < 0 || ib < 0) return ib < 0 && ia >= 0;
  return sa > sb || (sa == sb && ia < ib);
}

__global__ void __launch_bounds__(256) k_lambdarank(const double* __restrict__ score, const float* __restrict__ label,
                                                    RankDev r, float* __restrict__ g, float* __restrict__ h) {
  __shared__ double ss[kRankMaxDocs];
  __shared__ int si[kRankMaxDocs];
  __shared__ float lam[kRankMaxDocs], hes[kRankMaxDocs];
  __shared__ double red[4];
  const int q = blockIdx.x;
  const int lo = r.bounds[q], cnt = r.bounds[q + 1] - lo;
  if (cnt <= 0 || cnt > kRankMaxDocs) return;  // block-uniform; larger queries are refused by the host
  int P = 1;
  while (P < cnt) P <<= 1;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    ss[i] = i < cnt ? score[lo + i] : 0.0;
    si[i] = i < cnt ? i : -1;
    if (i < cnt) {
      lam[i] = 0.0f;
      hes[i] = 0.0f;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int l = i ^ jj;
        if (l > i) {
          const bool up = (i & k) == 0;
          const bool swap = up ? rank_before(ss[l], si[l], ss[i], si[i]) : rank_before(ss[i], si[i], ss[l], si[l]);
          if (swap) {
            const double ts = ss[i];
            ss[i] = ss[l];
            ss[l] = ts;
            const int ti = si[i];
            si[i] = si[l];
            si[l] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  const double inv_max_dcg = r.inv_max_dcg[q];
  const double best = ss[0];
  int worst_idx = cnt - 1;
  if (worst_idx > 0 && ss[worst_idx] == -INFINITY) worst_idx -= 1;
  const double worst = ss[worst_idx];
  const bool norm_pairs = r.norm && best != worst;
  const double eps_norm = (double)0.01f;
  double sum_lambdas = 0.0;
  const int top = min(cnt - 1, max(r.truncation, 0));
  for (int i = 0; i < top; ++i) {
    if (ss[i] == -INFINITY) continue;  // block-uniform
    const float li = label[lo + si[i]];
    for (int j = i + 1 + (int)threadIdx.x; j < cnt; j += blockDim.x) {
      const float lj = label[lo + si[j]];
      if (ss[j] == -INFINITY || li == lj) continue;
      const bool i_high = li > lj;
      const int high_rank = i_high ? i : j, low_rank = i_high ? j : i;
      const int high = si[high_rank], low = si[low_rank];
      const double delta_score = ss[high_rank] - ss[low_rank];
      const double dcg_gap = r.label_gain[(int)label[lo + high]] - r.label_gain[(int)label[lo + low]];
      const double paired_discount = fabs(r.discount[high_rank] - r.discount[low_rank]);
      double dpn = dcg_gap * paired_discount * inv_max_dcg;
      if (norm_pairs) dpn /= eps_norm + fabs(delta_score);
      double pl;
      if (delta_score <= r.min_in) pl = r.sigmoid_table[0];
      else if (delta_score >= r.max_in) pl = r.sigmoid_table[r.table_bins - 1];
      else pl = r.sigmoid_table[min((int)((delta_score - r.min_in) * r.table_factor), r.table_bins - 1)];
      double ph = pl * (1.0 - pl);
      pl = pl * (-r.sigmoid * dpn);
      ph = ph * (r.sigmoid * r.sigmoid * dpn);
      atomicAdd(&lam[low], -(float)pl);
      atomicAdd(&hes[low], (float)ph);
      atomicAdd(&lam[high], (float)pl);
      atomicAdd(&hes[high], (float)ph);
      sum_lambdas -= 2.0 * pl;
    }
  }
  for (int o = 32; o > 0; o >>= 1) sum_lambdas += __shfl_xor(sum_lambdas, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum_lambdas;
  __syncthreads();
  const double total = red[0] + red[1] + red[2] + red[3];
  const bool scale = r.norm && total > 0.0;
  const double nf = scale ? log2(1.0 + total) / total : 1.0;
  for (int d = threadIdx.x; d < cnt; d += blockDim.x) {
    g[lo + d] = scale ? (float)((double)lam[d] * nf) : lam[d];
    h[lo + d] = scale ? (float)((double)hes[d] * nf) : hes[d];
  }
}

// Binary log-loss gradients (objective/binary.rs:71-90) for a device-side boosting
// loop: the same products in the same order (response * label_weight [* w]).
__global__ void k_binary_gradients(const double* __restrict__ score, const float* __restrict__ label,
                                   const float* __restrict__ weights, int n, double sigmoid, double lw_neg,
                                   double lw_pos, float* __restrict__ g, float* __restrict__ h) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const bool pos = label[i] > 0.0f;
    const double lab = pos ? 1.0 : -1.0;
    const double lw = pos ? lw_pos : lw_neg;
    const double response = -lab * sigmoid / (1.0 + exp(lab * sigmoid * score[i]));
    const double ar = fabs(response);
    if (weights) {
      const double w = (double)weights[i];
      g[i] = (float)(response * lw * w);
      h[i] = (float)(ar * (sigmoid - ar) * lw * w);
    } else {
      g[i] = (float)(response * lw);
      h[i] = (float)(ar * (sigmoid - ar) * lw);
    }
  }
}


// Per-block f64 sums of gradients and hessians (root totals of the float forms).
// rows: the rows to sum (a bag, each inside [0, num_data)), or null for rows 0 .. n-1.
__global__ void k_sum_gh(const float* __restrict__ g, const float* __restrict__ h, const int32_t* __restrict__ rows,
                         int n, int num_data, double* out) {
  __shared__ double rg[4], rh[4];
  double sg = 0.0, sh = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = rows ? rows[i] : i;
    if (r < 0 || r >= num_data) continue;  // flagged by the bag's gather
    sg += (double)g[r];
    sh += (double)h[r];
  }
  for (int o = 32; o > 0; o >>= 1) {
    sg += __shfl_down(sg, o, 64);
    sh += __shfl_down(sh, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    rg[threadIdx.x >> 6] = sg;
    rh[threadIdx.x >> 6] = sh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += rg[w];
      b += rh[w];
    }
    out[2 * blockIdx.x] = a;
    out[2 * blockIdx.x + 1] = b;
  }
}


// Sum of the partial histograms of every item into its pool slot. A block
// covers 32 bins of one item with 8 lanes per bin, each lane summing every 8th
// partial (32 consecutive bins of one partial are one coalesced read); the 8
// sums are then added in lane order, so float forms reduce in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kReduceMaxItemsY = 1024;  // grid rows of a reduce whose item count the device decides

template <int FORM>
__global__ void __launch_bounds__(256) k_reduce(DevState s, const ReduceItem* __restrict__ items,
                                                const RoundHdr* __restrict__ hdr) {
  __shared__ long long ra[8][32], rb[8][32];
  __shared__ double da[8][32], db[8][32];
  __shared__ long long rc[8][32];
  // items blockIdx.y, blockIdx.y + gridDim.y, ... (a planned round's grid is
  // capped below the most items a round may hold)
  const int nitems = hdr ? hdr->nscan : (int)gridDim.y;
  for (int y = blockIdx.y; y < nitems; y += gridDim.y) {
    const ReduceItem it = items[y];
    const int lb = threadIdx.x & 31, sp = threadIdx.x >> 5;
    const int bin = blockIdx.x * 32 + lb;
    const size_t stride = (size_t)s.total_bins;
    long long g = 0, h = 0, c = 0;
    double fg = 0.0, fh = 0.0;
    if (bin < s.total_bins) {
      const size_t base = (size_t)it.part_first * stride + bin;
      if (FORM == FORM_QUANT) {
        int j = sp;
        for (; j + 24 < it.nblk; j += 32) {
          unsigned long long w[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) w[u] = s.part_q[base + (size_t)(j + 8 * u) * stride];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            long long gg, hh, cc;
            qunpack(w[u], &gg, &hh, &cc);
            g += gg;
            h += hh;
            c += cc;
          }
        }
        for (; j < it.nblk; j += 8) {
          long long gg, hh, cc;
          qunpack(s.part_q[base + (size_t)j * stride], &gg, &hh, &cc);
          g += gg;
          h += hh;
          c += cc;
        }
      } else if (FORM == kHistQ32) {
        // four partials in flight per step: the sum is latency-bound, not bandwidth-bound
        const unsigned* pq = (const unsigned*)s.part_q;
        int j = sp;
        for (; j + 24 < it.nblk; j += 32) {
          unsigned w[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) w[u] = pq[base + (size_t)(j + 8 * u) * stride];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            long long gg, hh;
            qunpack32(w[u], &gg, &hh);
            g += gg;
            h += hh;
          }
        }
        for (; j < it.nblk; j += 8) {
          long long gg, hh;
          qunpack32(pq[base + (size_t)j * stride], &gg, &hh);
          g += gg;
          h += hh;
        }
      } else {
        for (int j = sp; j < it.nblk; j += 8) {
          fg += s.part_g[base + (size_t)j * stride];
          fh += s.part_h[base + (size_t)j * stride];
          c += s.part_c[base + (size_t)j * stride];
        }
      }
    }
    ra[sp][lb] = g;
    rb[sp][lb] = h;
    rc[sp][lb] = c;
    da[sp][lb] = fg;
    db[sp][lb] = fh;
    __syncthreads();
    if (sp == 0 && bin < s.total_bins) {
      double* pool = s.pool + ((size_t)it.slot * s.total_bins) * 2;
      int* cnt = s.pool_cnt + (size_t)it.slot * s.total_bins;
      long long tc = 0;
      if (FORM == FORM_QUANT || FORM == kHistQ32) {
        long long tg = 0, th = 0;
        for (int k = 0; k < 8; ++k) {
          tg += ra[k][lb];
          th += rb[k][lb];
          tc += rc[k][lb];
        }
        ((long long*)pool)[2 * bin] = tg;
        ((long long*)pool)[2 * bin + 1] = th;
      } else {
        double tg = 0.0, th = 0.0;
        for (int k = 0; k < 8; ++k) {
          tg += da[k][lb];
          th += db[k][lb];
          tc += rc[k][lb];
        }
        pool[2 * bin] = tg;
        pool[2 * bin + 1] = th;
      }
      if (FORM != kHistQ32) cnt[bin] = (int)tc;  // the 32-bit form keeps no counts
    }
    __syncthreads();  // the LDS planes are reused by the next item
  }
}

inline int grid_for(long long n, int threads, int cap = 4096) {
  long long b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}


// Raises the dynamic LDS limit of a kernel to kLdsBytes, once per kernel (kernels
// of one signature, such as the forms of a template, are told apart by address);
// learners on several host threads may call it at once.
template <typename K>
void allow_lds(K kernel) {
  static std::mutex mu;
  static std::unordered_set<const void*> done;
  std::lock_guard<std::mutex> lock(mu);
  if (done.count((const void*)kernel)) return;
  BK_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  done.insert((const void*)kernel);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
size_t hist_lds_bytes(int form, int hist32, int max_tile_bins) {
  const size_t per_bin = form == FORM_QUANT ? (hist32 ? 4 : 8) : 20;
  return ((size_t)max_tile_bins * per_bin + 15) & ~(size_t)15;
}

void launch_hist(const DevState& s, const HistItem* d_items, int nitems, int total_rows, int nblocks,
                 int rows_per_block, const int* d_block_slot, const int* d_group_used, int max_tile_bins,
                 const int* d_tile_first, int ntiles, int rows_per_flush, void* stream, const RoundHdr* hdr,
                 int nsub) {
  if (nblocks <= 0 || (!hdr && total_rows <= 0)) return;
  if (nsub < 1 || nsub > kMaxSubHist) throw std::runtime_error("sub-histogram count out of range");
  const unsigned groups8 = (unsigned)((nblocks + kXcds - 1) / kXcds);
  dim3 grid(groups8 * kXcds * (unsigned)ntiles);
  hipStream_t st = (hipStream_t)stream;
  const size_t pay = s.form == FORM_QUANT ? 2 : 8;
  const size_t lds = hist_lds_bytes(s.form, s.hist32, max_tile_bins) * nsub + 4 * (size_t)((s.num_groups + 3) & ~3) +
                     pay * kHistRows;
  if (lds > (size_t)kLdsBytes) throw std::runtime_error("histogram tile exceeds the LDS budget");
  switch (s.form) {
    case FORM_QUANT:
      if (s.hist32) {
        allow_lds(k_hist<kHistQ32>);
        launch(k_hist<kHistQ32>, grid, dim3(kHistThreads), lds, st, s, d_items, nitems, total_rows, rows_per_block,
               d_block_slot, d_group_used, d_tile_first, rows_per_flush, nblocks, ntiles, nsub, hdr);
        break;
      }
      allow_lds(k_hist<FORM_QUANT>);
      launch(k_hist<FORM_QUANT>, grid, dim3(kHistThreads), lds, st, s, d_items, nitems, total_rows,
                         rows_per_block, d_block_slot, d_group_used, d_tile_first, rows_per_flush, nblocks,
                         ntiles, nsub, hdr);
      break;
    case FORM_F64:
      allow_lds(k_hist<FORM_F64>);
      launch(k_hist<FORM_F64>, grid, dim3(kHistThreads), lds, st, s, d_items, nitems, total_rows,
                         rows_per_block, d_block_slot, d_group_used, d_tile_first, rows_per_flush, nblocks,
                         ntiles, nsub, hdr);
      break;
    default:
      allow_lds(k_hist<FORM_F32C>);
      launch(k_hist<FORM_F32C>, grid, dim3(kHistThreads), lds, st, s, d_items, nitems, total_rows,
                         rows_per_block, d_block_slot, d_group_used, d_tile_first, rows_per_flush, nblocks,
                         ntiles, nsub, hdr);
      break;
  }
  BK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Device-resident trees
// ---------------------------------------------------------------------------
// update_scales (gradient_discretizer/mod.rs:134-165) from the maxima, one lane.
__global__ void k_disc_prep(unsigned* __restrict__ max4, const int* __restrict__ start_word, int bins, int const_hess,
                            DiscState* __restrict__ d) {
  if (threadIdx.x != 0) return;
  const unsigned m0 = max4[0], m1 = max4[1], m2 = max4[2];
  max4[0] = max4[1] = max4[2] = max4[3] = 0u;  // the next tree's maxima start from zero
  const float max_g = __uint_as_float(m0), max_h = __uint_as_float(m1);
  const int ok = m2 == 0u && max_g != 0.0f && (const_hess || max_h != 0.0f);
  d->ok = ok;
  if (!ok) return;
  d->gs = (double)max_g / (double)(bins / 2);
  d->hs = const_hess ? (double)max_h : (double)max_h / (double)bins;
  d->inv_g = 1.0 / d->gs;
  d->inv_h = 1.0 / d->hs;
  d->start = *start_word;
}

// Σ|x| * odd part of the scale's significand < 2^53: every sequential partial sum
// of x_i * scale is exact, so the reference's ordered sum equals the exact one.
__device__ bool sum_exact(long long sum_abs, double scale) {
  if (scale == 0.0) return true;
  int e;
  const double m = frexp(scale, &e);
  unsigned long long mant = (unsigned long long)ldexp(m, 53);
  while (mant && !(mant & 1ull)) mant >>= 1;
  return (unsigned __int128)(unsigned long long)sum_abs * mant < ((unsigned __int128)1 << 53);
}

__device__ double leaf_output_root_dev(double g, double h, double l1, double l2, double mds) {
  const double reg = fmax(fabs(g) - l1, 0.0);
  double ret = -(sgn(g) * reg) / (h + l2);
  if (mds > 0.0 && fabs(ret) > mds) ret = sgn(ret) * mds;
  if (ret < -DBL_MAX) return -DBL_MAX;  // NaN passes through, as on the host
  if (ret > DBL_MAX) return DBL_MAX;
  return ret;
}

__global__ void __launch_bounds__(256) k_root_prep(DiscState* __restrict__ d, const long long* __restrict__ sums3,
                                                   const double* __restrict__ partials,
                                                   const float* __restrict__ hessians, RootArgs ra, PlanParams pp,
                                                   PlanState* __restrict__ st, PlanNode* __restrict__ nodes,
                                                   uint8_t* __restrict__ node_flags,
                                                   const int* __restrict__ bytree_src, int* __restrict__ bytree,
                                                   HistItem* __restrict__ hitems, ScanItem* __restrict__ sitems,
                                                   ReduceItem* __restrict__ ritems, int* __restrict__ used,
                                                   int* __restrict__ err) {
  const int tid = threadIdx.x, F = pp.F, N = ra.num_data;
  for (int f = tid; f < F; f += blockDim.x) {
    const int b = bytree_src[f];
    bytree[f] = b;
    used[f] = b ? 7 : 0;
    node_flags[f] = 1;
  }
  for (int i = tid; i < (int)(sizeof(PlanState) / 4); i += blockDim.x) ((uint32_t*)st)[i] = 0u;
  __syncthreads();
  if (tid != 0) return;
  double sg = 0.0, sh = 0.0;
  long long isum = 0;
  if (ra.quantized) {
    if (d->ok) {
      // before_train's root sums (serial.rs:741-761)
      if (!sum_exact(sums3[2], d->gs) || !sum_exact(sums3[1], d->hs)) atomicOr(err, 16);
      d->int_g = sums3[0];
      d->int_h = sums3[1];
      d->root_g = (double)sums3[0] * d->gs;
      d->root_h = (double)sums3[1] * d->hs;
    }
    sg = d->root_g;
    sh = d->root_h;
    isum = (long long)(((unsigned long long)d->int_g << 32) + (unsigned long long)d->int_h);
  } else {
    for (int b = 0; b < ra.nparts; ++b) {
      sg += partials[2 * b];
      sh += partials[2 * b + 1];
    }
    d->h0 = (double)hessians[0];
  }
  const bool search = N > 0 && N >= ra.min_data_in_leaf * 2;
  PlanNode& n0 = nodes[0];
  n0.seg = Segment{0, 0, N};
  n0.depth = 0;
  n0.slot = 0;
  n0.expanded = 0;
  n0.kids_searched = 0;
  n0.child[0] = n0.child[1] = -1;
  n0.num_data = N;
  n0.pad = 0;
  n0.sum_g = sg;
  n0.sum_h = sh;
  n0.weight = 0.0;
  n0.int_sum = isum;
  n0.best.o = SplitOut{};
  n0.best.o.gain = kMinScore;
  n0.best.o.feature = -1;
  n0.best.o.default_left = 1;
  n0.best.inner = -1;
  st->num_leaves = 1;
  st->num_nodes = 1;
  st->next_slot = 1;
  st->nabsorb = search ? 1 : 0;
  st->absorb[0] = 0;
  st->absorb[1] = -1;
  st->leaf_node[0] = 0;
  // the root's histogram over every row, tiled as the host plans it
  long long nb = min((long long)pp.hist_max_blocks, max(1ll, (long long)N / pp.hist_min_rows));
  nb = max(nb, ((long long)N + pp.rows_per_flush - 1) / pp.rows_per_flush);
  const int rpb = (int)(((long long)N + nb - 1) / nb);
  const int nblocks = N > 0 ? (N + rpb - 1) / rpb : 0;
  hitems[0] = HistItem{Segment{0, 0, N}, 0, 0};
  ritems[0] = ReduceItem{0, 0, nblocks, 0};
  ScanItem si{};
  si.small = LeafArgs{0, 0, N, sg, sh, leaf_output_root_dev(sg, sh, ra.l1, ra.l2, ra.max_delta_step), isum, 0,
                      nblocks};
  sitems[0] = si;
  st->hdr = RoundHdr{0, 0, search ? 1 : 0, N, nblocks, rpb, search ? 1 : 0, 0};
}

__global__ void __launch_bounds__(256) k_finalize(const PlanState* __restrict__ st, const PlanNode* __restrict__ nodes,
                                                  const FeatDev* __restrict__ feats, NodeRule* __restrict__ rules,
                                                  int* __restrict__ d_count, Segment* __restrict__ segs,
                                                  int* __restrict__ order, int num_leaves) {
  __shared__ int cnt[kPlanMaxLeaves];
  const int tid = threadIdx.x, nc = st->ncommit, nl = st->num_leaves;
  if (tid == 0) *d_count = nc;
  for (int c = tid; c < nc; c += blockDim.x) {
    const int leaf = st->commit[2 * c], x = st->commit[2 * c + 1];
    // the node's children: the first later commit that splits the same leaf index
    // (the left child keeps it; the right child is leaf c + 1), else the leaf
    int left = ~leaf, right = ~(c + 1);
    for (int c2 = nc - 1; c2 > c; --c2) {
      const int l2 = st->commit[2 * c2];
      if (l2 == leaf) left = c2;
      if (l2 == c + 1) right = c2;
    }
    const SplitOut& o = nodes[x].best.o;
    const FeatDev m = feats[nodes[x].best.inner];
    NodeRule r{};
    r.rule = m.bin_type == 1 ? make_rule_categorical(m, o.cat_bits)
                             : make_rule_numerical(m, (uint32_t)o.threshold, o.default_left != 0);
    r.group = m.group;
    r.left = left;
    r.right = right;
    rules[c] = r;
  }
  for (int l = tid; l < num_leaves; l += blockDim.x) {
    const Segment sg = l < nl ? nodes[st->leaf_node[l]].seg : Segment{0, 0, 0};
    segs[l] = sg;
    cnt[l] = sg.count;
  }
  __syncthreads();
  // leaves by row count, descending, ties by index
  for (int l = tid; l < num_leaves; l += blockDim.x) {
    int r = 0;
    for (int m = 0; m < num_leaves; ++m) r += (cnt[m] > cnt[l] || (cnt[m] == cnt[l] && m < l)) ? 1 : 0;
    order[r] = l;
  }
}

void launch_disc_prep(unsigned* d_max4, const int* start_word, int bins, int const_hess, DiscState* d, void* stream) {
  launch(k_disc_prep, dim3(1), dim3(64), 0, (hipStream_t)stream, d_max4, start_word, bins, const_hess, d);
  BK_HIP(hipGetLastError());
}

void launch_root_prep(DiscState* d, const long long* sums3, const double* partials, const float* hessians,
                      const RootArgs& ra, const PlanParams& pp, PlanState* st, PlanNode* nodes, uint8_t* node_flags,
                      const int* bytree_src, int* bytree, HistItem* hitems, ScanItem* sitems, ReduceItem* ritems,
                      int* used, int* err, void* stream) {
  launch(k_root_prep, dim3(1), dim3(256), 0, (hipStream_t)stream, d, sums3, partials, hessians, ra, pp, st, nodes,
         node_flags, bytree_src, bytree, hitems, sitems, ritems, used, err);
  BK_HIP(hipGetLastError());
}

void launch_finalize(const PlanState* st, const PlanNode* nodes, const FeatDev* feats, NodeRule* rules, int* d_count,
                     Segment* segs, int* order, int num_leaves, void* stream) {
  if (num_leaves > kPlanMaxLeaves) throw std::invalid_argument("the device planner takes at most 256 leaves");
  launch(k_finalize, dim3(1), dim3(256), 0, (hipStream_t)stream, st, nodes, feats, rules, d_count, segs, order,
         num_leaves);
  BK_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_plan_init(const PlanState* __restrict__ s_state,
                                                  const PlanNode* __restrict__ s_node,
                                                  const int* __restrict__ s_bytree, const uint8_t* __restrict__ s_flags,
                                                  PlanState* st, PlanNode* node0, int* bytree, uint8_t* flags0, int F) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  for (int i = tid; i < (int)(sizeof(PlanState) / 4); i += nt) ((uint32_t*)st)[i] = ((const uint32_t*)s_state)[i];
  for (int i = tid; i < (int)(sizeof(PlanNode) / 4); i += nt) ((uint32_t*)node0)[i] = ((const uint32_t*)s_node)[i];
  for (int i = tid; i < F; i += nt) {
    bytree[i] = s_bytree[i];
    flags0[i] = s_flags[i];
  }
}

void launch_plan_init(const PlanState* s_state, const PlanNode* s_node, const int* s_bytree, const uint8_t* s_flags,
                      PlanState* st, PlanNode* node0, int* bytree, uint8_t* flags0, int F, void* stream) {
  static_assert(sizeof(PlanState) % 4 == 0 && sizeof(PlanNode) % 4 == 0, "word copies");
  launch(k_plan_init, dim3(4), dim3(256), 0, (hipStream_t)stream, s_state, s_node, s_bytree, s_flags, st, node0,
         bytree, flags0, F);
  BK_HIP(hipGetLastError());
}

void launch_plan(PlanState* st, PlanNode* nodes, uint8_t* node_flags, const FoldOut* fold, const uint8_t* fold_flags,
                 const FeatDev* feats, const int* bytree, const PlanParams& pp, PartItem* pitems, HistItem* hitems,
                 ScanItem* sitems, ReduceItem* ritems, int* used, void* stream) {
  if (pp.num_leaves > kPlanMaxLeaves) throw std::invalid_argument("the device planner takes at most 256 leaves");
  launch(k_plan, dim3(1), dim3(kPlanThreads), 0, (hipStream_t)stream, st, nodes, node_flags, fold, fold_flags, feats,
         bytree, pp, pitems, hitems, sitems, ritems, used);
  BK_HIP(hipGetLastError());
}

void launch_children(const PartItem* d_pitems, int nitems, const unsigned* d_cursors, int max_items, unsigned epoch,
                     const ChildSpec* d_specs, const ChildParams& cp, HistItem* hitems, ScanItem* sitems,
                     ReduceItem* ritems, RoundHdr* hdr, int* counts, void* stream) {
  if (nitems <= 0) return;
  if (nitems > kChildThreads || nitems > max_items) throw std::runtime_error("too many splits in one round");
  const unsigned* cur = d_cursors + (size_t)(epoch & 1u) * max_items * 2;
  launch(k_children, dim3(1), dim3(kChildThreads), 0, (hipStream_t)stream, d_pitems, nitems, cur, d_specs, cp, hitems,
         sitems, ritems, hdr, counts);
  BK_HIP(hipGetLastError());
}

int partition_tile_rows(int stride) { return kPartThreads * kPartChunks / (stride >> 4); }
int partition_tile_rows_small(int stride) { return kPartThreads * kPartChunksSmall / (stride >> 4); }

void launch_partition(const DevState& s, const PartItem* d_items, int nitems, int ntasks, unsigned* d_cursors,
                      int max_items, unsigned epoch, void* stream, const RoundHdr* hdr, const DeriveArgs* derive,
                      bool small_tiles) {
  if (ntasks <= 0) return;
  if (nitems > max_items) throw std::runtime_error("more partition items than cursors");
  DeriveArgs da{};
  if (derive && derive->on) {
    if (hdr || nitems > kPartDeriveMax) throw std::runtime_error("the partition derives at most 512 splits, unplanned");
    da = *derive;
  }
  const int R = small_tiles ? partition_tile_rows_small(s.stride) : partition_tile_rows(s.stride);
  if (R < 1) throw std::runtime_error("row record too large for the partition tile");
  const size_t lds = partition_pay_lds(R, s.stride, s.form == FORM_QUANT ? 2 : 8);
  if (lds > (size_t)kLdsBytes) throw std::runtime_error("partition tile exceeds the LDS budget");
  hipStream_t st = (hipStream_t)stream;
  const int* no_groups = nullptr;
  const dim3 grid((unsigned)ntasks), block(kPartThreads);
  if (s.form == FORM_QUANT && small_tiles) {
    allow_lds(k_partition<FORM_QUANT, false, kPartChunksSmall>);
    launch(k_partition<FORM_QUANT, false, kPartChunksSmall>, grid, block, lds, st, s, d_items, nitems, d_cursors,
           max_items, epoch, hdr, no_groups, da);
  } else if (s.form == FORM_QUANT) {
    allow_lds(k_partition<FORM_QUANT, false>);
    launch(k_partition<FORM_QUANT, false>, grid, block, lds, st, s, d_items, nitems, d_cursors, max_items, epoch, hdr,
           no_groups, da);
  } else if (small_tiles) {
    allow_lds(k_partition<FORM_F64, false, kPartChunksSmall>);
    launch(k_partition<FORM_F64, false, kPartChunksSmall>, grid, block, lds, st, s, d_items, nitems, d_cursors,
           max_items, epoch, hdr, no_groups, da);
  } else {
    allow_lds(k_partition<FORM_F64, false>);
    launch(k_partition<FORM_F64, false>, grid, block, lds, st, s, d_items, nitems, d_cursors, max_items, epoch, hdr,
           no_groups, da);
  }
  BK_HIP(hipGetLastError());
}

bool partition_hist_fits(const DevState& s) {
  const int nc = s.stride >> 4;
  if (s.form != FORM_QUANT || !s.hist32 || nc < 1 || kPartThreads % nc != 0) return false;
  const int R = partition_tile_rows(s.stride);
  return partition_hist_lds(R, s.stride, s.total_bins) <= (size_t)kLdsBytes;
}

void launch_partition_hist(const DevState& s, const PartItem* d_items, int nitems, int ntasks, unsigned* d_cursors,
                           int max_items, unsigned epoch, const int* d_group_used, void* stream,
                           const DeriveArgs* derive) {
  if (ntasks <= 0) return;
  if (nitems > max_items) throw std::runtime_error("more partition items than cursors");
  DeriveArgs da{};
  if (derive && derive->on) {
    if (nitems > kPartDeriveMax) throw std::runtime_error("the partition derives at most 512 splits");
    da = *derive;
  }
  if (!partition_hist_fits(s)) throw std::runtime_error("the partition cannot build these histograms");
  const int R = partition_tile_rows(s.stride);
  const size_t lds = partition_hist_lds(R, s.stride, s.total_bins);
  allow_lds(k_partition<FORM_QUANT, true>);
  launch(k_partition<FORM_QUANT, true>, dim3((unsigned)ntasks), dim3(kPartThreads), lds, (hipStream_t)stream, s,
         d_items, nitems, d_cursors, max_items, epoch, (const RoundHdr*)nullptr, d_group_used, da);
  BK_HIP(hipGetLastError());
}

void launch_to_records(const DevState& s, const uint8_t* cols, size_t col_stride, void* stream) {
  const size_t lds = (size_t)kRecRows * s.stride;
  if (lds > (size_t)kLdsBytes) throw std::runtime_error("row record too large");
  allow_lds(k_to_records);
  const unsigned blocks = (unsigned)((s.num_data + kRecRows - 1) / kRecRows);
  launch(k_to_records, dim3(blocks), dim3(256), lds, (hipStream_t)stream, s, cols, col_stride);
  BK_HIP(hipGetLastError());
}

void launch_reduce(const DevState& s, const ReduceItem* d_items, int nitems, void* stream, const RoundHdr* hdr,
                   int hdr_grid_cap) {
  if (nitems <= 0) return;
  // a planned round's kernel loops over its items: a capped grid holds them
  dim3 grid((unsigned)((s.total_bins + 31) / 32),
            (unsigned)(hdr ? std::min(nitems, std::min(hdr_grid_cap, kReduceMaxItemsY)) : nitems));
  hipStream_t st = (hipStream_t)stream;
  switch (s.form) {
    case FORM_QUANT:
      if (s.hist32) launch(k_reduce<kHistQ32>, grid, dim3(256), 0, st, s, d_items, hdr);
      else launch(k_reduce<FORM_QUANT>, grid, dim3(256), 0, st, s, d_items, hdr);
      break;
    case FORM_F64:
      launch(k_reduce<FORM_F64>, grid, dim3(256), 0, st, s, d_items, hdr);
      break;
    default:
      launch(k_reduce<FORM_F32C>, grid, dim3(256), 0, st, s, d_items, hdr);
      break;
  }
  BK_HIP(hipGetLastError());
}


void launch_scan(const DevState& s, const ScanItem* d_items, int nitems, const ScanParams& sp, const int* d_used,
                 SplitOut* d_out, FoldOut* d_best, uint8_t* d_flags, void* stream, const RoundHdr* hdr,
                 unsigned* stamp, unsigned seq, unsigned* ticket, int hdr_grid_cap) {
  if (stamp && !ticket) throw std::runtime_error("a stamped fold needs its ticket word");
  if (nitems <= 0) return;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid((unsigned)s.num_features, (unsigned)(hdr ? std::min(nitems, hdr_grid_cap) : nitems));
  switch (s.form) {
    case FORM_QUANT:
      launch(k_scan<FORM_QUANT>, grid, dim3(kScanThreads), 0, st, s, d_items, sp, d_used, d_out, hdr);
      break;
    case FORM_F64:
      launch(k_scan<FORM_F64>, grid, dim3(kScanThreads), 0, st, s, d_items, sp, d_used, d_out, hdr);
      break;
    default:
      launch(k_scan<FORM_F32C>, grid, dim3(kScanThreads), 0, st, s, d_items, sp, d_used, d_out, hdr);
      break;
  }
  BK_HIP(hipGetLastError());
  launch(k_fold, dim3((unsigned)(2 * nitems)), dim3(64), 0, st, d_items, d_out, d_used, s.num_features,
                     d_best, d_flags, hdr, stamp, seq, ticket);
  BK_HIP(hipGetLastError());
}

void launch_max_abs(const float* g, const float* h, int n, unsigned* d_out4, long long* d_zero4, void* stream) {
  launch(k_max_abs, dim3(grid_for(n, 256 * 8, 1024)), dim3(256), 0, (hipStream_t)stream, g, h, n, d_out4, d_zero4);
  BK_HIP(hipGetLastError());
}

void launch_discretize(const float* g, const float* h, const double* gnoise, const double* hnoise, int n, int start,
                       double inv_g, double inv_h, int const_hess, int stochastic, int16_t* packed,
                       long long* d_sums3, void* stream, const DiscState* d_state) {
  launch(k_discretize, dim3(grid_for(n, 256 * 8, 1024)), dim3(256), 0, (hipStream_t)stream, g, h, gnoise,
                     hnoise, n, start, inv_g, inv_h, const_hess, stochastic, packed, d_sums3, d_state);
  BK_HIP(hipGetLastError());
}

void launch_pack_f32(const float* g, const float* h, int n, int const_hess, float* gh, void* stream) {
  launch(k_pack_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, (hipStream_t)stream, g, h, n,
                     const_hess, (float2*)gh);
  BK_HIP(hipGetLastError());
}

int renew_piece_rows(long long n, int num_leaves, size_t max_bytes) {
  long long rows = kRenewPieceRows;
  while (rows < n && (size_t)((n + rows - 1) / rows) * (size_t)num_leaves * sizeof(RenewPiece) > max_bytes) rows *= 2;
  return (int)std::min<long long>(rows, 1ll << 22);  // the magnitude bound of k_renew_chain holds to 2^22 rows
}

size_t renew_pieces_lds(int num_leaves) { return (size_t)num_leaves * (sizeof(RenewPiece) + sizeof(int)); }

void launch_renew_sums(int n, const int32_t* leaf_of, const int* d_order, int num_leaves, const float* g,
                       const float* h, const int* d_flags, double* d_sums, const RenewPieces& pieces, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = renew_pieces_lds(num_leaves);
  if (pieces.pieces && lds <= (size_t)kLdsBytes && pieces.piece_rows > 0 &&
      (long long)pieces.num_pieces * pieces.piece_rows >= n) {
    const bool summaries = pieces.binade && pieces.work && pieces.parts && pieces.work_count && pieces.work_cap > 0 &&
                           pieces.piece_rows % (256 * kRenewParts) == 0;
    allow_lds(k_renew_pieces);
    launch(k_renew_pieces, dim3((unsigned)std::max(pieces.num_pieces, 1)), dim3(256), lds, st, n, pieces.piece_rows,
           leaf_of, g, h, num_leaves, d_flags, pieces.pieces, summaries ? pieces.work_count : (int*)nullptr);
    BK_HIP(hipGetLastError());
    if (summaries) {
      launch(k_renew_guess, dim3((unsigned)num_leaves), dim3(128), 0, st, pieces.num_pieces, num_leaves, d_flags,
             (const RenewPiece*)pieces.pieces, pieces.binade, pieces.work, pieces.work_count, pieces.work_cap);
      BK_HIP(hipGetLastError());
      // the listed pieces, one wave each (a block's waves exit once the list is done)
      launch(k_renew_binade, dim3(1024), dim3(256), 0, st, n, pieces.piece_rows, num_leaves, leaf_of, g, h,
             pieces.binade, (const RenewWork*)pieces.work, pieces.parts, (const int*)pieces.work_count,
             pieces.work_cap);
      BK_HIP(hipGetLastError());
    }
    launch(k_renew_chain, dim3((unsigned)num_leaves), dim3(128), 0, st, n, pieces.piece_rows, pieces.num_pieces,
           leaf_of, g, h, num_leaves, d_flags, (const RenewPiece*)pieces.pieces,
           summaries ? (const RenewBinade*)pieces.binade : (const RenewBinade*)nullptr,
           (const RenewPart*)pieces.parts, d_sums);
    BK_HIP(hipGetLastError());
    return;
  }
  launch(k_renew_sums, dim3((unsigned)num_leaves), dim3(256), 0, st, n, leaf_of, d_order, g, h, d_flags, d_sums);
  BK_HIP(hipGetLastError());
}

void launch_softmax_gradients(const double* score, const float* label, int n, int num_class, float* g, float* h,
                              void* stream) {
  const double factor = (double)num_class / ((double)num_class - 1.0);
  launch(k_softmax_gradients, dim3(grid_for(n, 256, 4096)), dim3(256), 0, (hipStream_t)stream, score, label, n,
         num_class, factor, g, h);
  BK_HIP(hipGetLastError());
}

void launch_lambdarank_gradients(const double* score, const float* label, const RankDev& r, float* g, float* h,
                                 void* stream) {
  if (r.num_queries <= 0) return;
  launch(k_lambdarank, dim3((unsigned)r.num_queries), dim3(256), 0, (hipStream_t)stream, score, label, r, g, h);
  BK_HIP(hipGetLastError());
}

void launch_binary_gradients(const double* score, const float* label, const float* weights, int n, double sigmoid,
                             double lw_neg, double lw_pos, float* g, float* h, void* stream) {
  launch(k_binary_gradients, dim3(grid_for(n, 256, 4096)), dim3(256), 0, (hipStream_t)stream, score, label, weights,
         n, sigmoid, lw_neg, lw_pos, g, h);
  BK_HIP(hipGetLastError());
}

// The allocation-free entry point's scratch (partial histograms): one array per
// device, held until the process ends. Host code never names it directly: its
// name is a host-side stand-in, so every use goes through scratch_device_ptr().
__device__ __align__(256) static unsigned char scratch[kScratchBytes];

void* scratch_device_ptr() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* ptr[kMaxDevices] = {};
  int dev = -1;
  BK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    BK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
    ptr[dev] = p;
  }
  return ptr[dev];
}

__global__ void k_scale_const_hess(double* __restrict__ hist, int total_bins, const float* __restrict__ hessians) {
  const double h0 = (double)hessians[0];
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < total_bins; b += gridDim.x * blockDim.x)
    hist[2 * b + 1] = hist[2 * b + 1] * h0;
}

void launch_scale_const_hess(double* hist, int total_bins, const float* hessians, void* stream) {
  launch(k_scale_const_hess, dim3(grid_for(total_bins, 256, 1024)), dim3(256), 0, (hipStream_t)stream, hist,
         total_bins, hessians);
  BK_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_copy(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, size_t n,
                                              int vec) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t done = 0;
  if (vec) {
    const size_t n16 = n / 16;
    for (size_t i = i0; i < n16; i += step) ((uint4*)dst)[i] = ((const uint4*)src)[i];
    done = n16 * 16;
  }
  for (size_t i = done + i0; i < n; i += step) dst[i] = src[i];
}

// Stamps `seq` into a word of pinned host memory: launched after a stream's work,
// it lets the host spin on the word instead of sleeping in a blocked wait.
__global__ void k_stamp(unsigned* word, unsigned seq) {
  if (threadIdx.x == 0) __hip_atomic_store(word, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A bounded wait of `ticks` of the device's wall clock (one lane), queued ahead of
// timed launches: launches submitted meanwhile wait in the queue behind it, so
// their timing events are stamped when the kernel is already there.
__global__ void k_hold(unsigned long long ticks) {
  if (threadIdx.x != 0) return;
  const unsigned long long t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

void launch_hold(unsigned long long ticks, void* stream) {
  launch(k_hold, dim3(1), dim3(64), 0, (hipStream_t)stream, ticks);
  BK_HIP(hipGetLastError());
}

void launch_stamp(unsigned* word, unsigned seq, void* stream) {
  launch(k_stamp, dim3(1), dim3(64), 0, (hipStream_t)stream, word, seq);
  BK_HIP(hipGetLastError());
}

void launch_copy(void* dst, const void* src, size_t n, void* stream) {
  if (n == 0) return;
  const int vec = ((((uintptr_t)dst) | ((uintptr_t)src)) & 15u) == 0 ? 1 : 0;
  const size_t units = vec ? (n + 15) / 16 : n;
  const unsigned grid = (unsigned)std::min<size_t>(64, (units + 255) / 256);
  launch(k_copy, dim3(grid), dim3(256), 0, (hipStream_t)stream, (uint8_t*)dst, (const uint8_t*)src, n, vec);
  BK_HIP(hipGetLastError());
}

int launch_sum_gh(const float* g, const float* h, const int32_t* rows, int n, int num_data, double* d_partials,
                  void* stream) {
  const int nb = grid_for(n, 256, 1024);
  launch(k_sum_gh, dim3(nb), dim3(256), 0, (hipStream_t)stream, g, h, rows, n, num_data, d_partials);
  BK_HIP(hipGetLastError());
  return nb;
}

void launch_leaf_of_tree(const DevState& s, const NodeRule* d_rules, NodeDev* d_nodes, int num_nodes,
                         const uint32_t* bag_bits, int32_t* leaf_of, void* stream, const int* d_num_nodes) {
  if (num_nodes > 0) {
    launch(k_node_bitmaps, dim3((unsigned)num_nodes), dim3(256), 0, (hipStream_t)stream, d_rules, d_nodes,
           d_num_nodes);
    BK_HIP(hipGetLastError());
  }
  const size_t node_bytes = (size_t)num_nodes * sizeof(NodeDev);
  const bool staged = s.stride <= 64 && node_bytes + walk_stage_bytes(s.stride) <= (size_t)kLdsBytes;
  const int in_lds = staged ? 2 : (node_bytes <= (size_t)kLdsBytes ? 1 : 0);
  const size_t lds = staged ? node_bytes + walk_stage_bytes(s.stride) : node_bytes;
  allow_lds(k_leaf_of_tree);
  launch(k_leaf_of_tree, dim3(grid_for(s.num_data, 256, 4096)), dim3(256), in_lds ? lds : 0, (hipStream_t)stream, s,
         d_nodes, num_nodes, in_lds, bag_bits, leaf_of, d_num_nodes);
  BK_HIP(hipGetLastError());
}

int renew_parts(int n) { return grid_for(n, 256 * 16, 1024); }

bool launch_leaf_of_tree_renew(const DevState& s, const NodeRule* d_rules, NodeDev* d_nodes, int num_nodes,
                               const uint32_t* bag_bits, int32_t* leaf_of, const float* g, const float* h,
                               int num_leaves, double* part_sums, int* part_meta, void* stream) {
  if (num_nodes <= 0 || s.stride > 64) return false;
  const size_t lds = (size_t)num_nodes * sizeof(NodeDev) + ((walk_stage_bytes(s.stride) + 15) & ~(size_t)15) +
                     renew_par_lds(num_leaves);
  if (lds > (size_t)kLdsBytes) return false;
  launch(k_node_bitmaps, dim3((unsigned)num_nodes), dim3(256), 0, (hipStream_t)stream, d_rules, d_nodes,
         (const int*)nullptr);
  BK_HIP(hipGetLastError());
  allow_lds(k_leaf_of_tree_renew);
  launch(k_leaf_of_tree_renew, dim3(renew_parts(s.num_data)), dim3(256), lds, (hipStream_t)stream, s, d_nodes,
         num_nodes, bag_bits, leaf_of, g, h, num_leaves, part_sums, part_meta);
  BK_HIP(hipGetLastError());
  return true;
}

size_t renew_par_lds(int num_leaves) { return (size_t)num_leaves * (2 * sizeof(double) + 4 * sizeof(int)); }

void launch_renew_par(int n, const int32_t* leaf_of, const float* g, const float* h, int num_leaves,
                      double* part_sums, int* part_meta, void* stream) {
  const size_t lds = renew_par_lds(num_leaves);
  if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_renew_par,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  launch(k_renew_par, dim3(renew_parts(n)), dim3(256), lds, (hipStream_t)stream, n, leaf_of, g, h,
                     num_leaves, part_sums, part_meta);
  BK_HIP(hipGetLastError());
}

void launch_renew_fix(const Segment* d_segs, int num_leaves, int nparts, const double* part_sums, const int* part_meta,
                      double* d_sums, int* d_flags, void* stream) {
  launch(k_renew_fix, dim3((unsigned)num_leaves), dim3(256), 0, (hipStream_t)stream, d_segs, num_leaves,
                     nparts, part_sums, part_meta, d_sums, d_flags);
  BK_HIP(hipGetLastError());
}

void launch_add_score_rows(int n, const int32_t* leaf_of, const double* d_leaf_values, int num_leaves, int out_of_bag,
                           double* score, void* stream) {
  launch(k_add_score_rows, dim3(grid_for(n, 256, 8192)), dim3(256), 0, (hipStream_t)stream, n, leaf_of,
                     d_leaf_values, num_leaves, out_of_bag, score);
  BK_HIP(hipGetLastError());
}

void launch_gather_bag(const DevState& s, const int32_t* bag, int nbag, int dst_buf, uint32_t* bag_bits,
                       long long* sums, void* stream) {
  if (nbag <= 0) return;
  const int nc = s.stride >> 4;
  const dim3 grid(grid_for((long long)nbag * nc, 256, 8192));
  if (s.form == FORM_QUANT)
    launch(k_gather_bag<FORM_QUANT>, grid, dim3(256), 0, (hipStream_t)stream, s, bag, nbag, dst_buf, bag_bits, sums);
  else
    launch(k_gather_bag<FORM_F64>, grid, dim3(256), 0, (hipStream_t)stream, s, bag, nbag, dst_buf, bag_bits, sums);
  BK_HIP(hipGetLastError());
}


}  // namespace bk
