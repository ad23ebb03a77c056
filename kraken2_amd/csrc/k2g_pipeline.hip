// k2g_pipeline: the per-batch launch chain of the classify core captured
// once as a HIP graph and replayed for every batch that fits its capacity.
//
// The chain is the one k2g_classify runs -- element bounds and exact
// hitlist layout, longest-first work units, the persistent classify kernel,
// the dense-table redo, and (for host input) the upload and the writeback of
// results into pinned host memory -- with every step that the direct path
// decides on the host moved onto the device, so one graph launch runs a whole
// batch without a host round trip:
//   * the batch's arrays and sizes come from a device descriptor (PipeDesc)
//     written by the graph's first kernel from a pinned host copy;
//   * work units come from length buckets (a counting sort: per-fragment
//     element counts binned by their top bits, buckets laid out longest
//     first, each bucket cut into units of about unit_elements elements)
//     instead of a radix sort and a selection;
//   * scans, fills and copies are this file's own kernels (a graph replays
//     them with no memset nodes);
//   * the dense-table redo reads its fragment count from device memory;
//   * with reserved (not exact) hitlist space, fragments that outgrow their
//     reservation take exact space from an overflow pool by an atomic bump
//     and are classified again in the same graph;
//   * long nucleotide fragments are split over waves (the direct path's
//     segments and stitch) when the device finds that the batch's longest
//     fragment outweighs a wave's share, so a small batch does not wait on
//     its longest read.
// Reference: the per-block loop of ProcessFiles (src/classify.cc:479-708)
// that the chain replaces, one ClassifySequence per read
// (src/classify.cc:773-948).
#include "k2g.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "k2g_internal.h"

namespace k2g {
__global__ void k_classify_pipe_dna(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                    unsigned long long *);
__global__ void k_classify_pipe_dna_x(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                      unsigned long long *);
__global__ void k_classify_pipe_prot(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                     unsigned long long *);
__global__ void k_classify_dense_pipe_dna(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                          const uint32_t *, const uint32_t *, DenseScratch);
__global__ void k_classify_dense_pipe_prot(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                           const uint32_t *, const uint32_t *, DenseScratch);
__global__ void k_seg_count_pipe(const PipeDesc *, uint32_t, uint32_t, const uint64_t *, const unsigned long long *,
                                 uint64_t, uint64_t, uint32_t *, uint32_t *, uint32_t *);
__global__ void k_seg_plan_pipe(const PipeDesc *, uint32_t, uint32_t, uint32_t, const uint32_t *, const uint64_t *,
                                uint32_t *, SegBatch);
__global__ void k_classify_segments_pipe_dna(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                             SegBatch);
__global__ void k_classify_segments_pipe_dna_x(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut,
                                               SegBatch);
__global__ void k_stitch_pipe_dna(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, SegBatch);
__global__ void k_stitch_pipe_dna_x(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, SegBatch);

namespace {

constexpr uint32_t kMaxShift = 2;                       // up to 4 length buckets per doubling
constexpr uint32_t kMaxBuckets = 1 + (64u << kMaxShift);  // bucket 0: fragments without elements
constexpr uint32_t kTile = 2048;                        // scan tile: 256 threads x 8 elements
constexpr uint32_t kTopThreads = 1024;

// Length bucket of an element count: 0 for none; else 1 + the position of
// its top bit times 2^shift + the next `shift` bits (so a bucket spans a
// factor of at most 2^(1/2^shift) in length).
__device__ __host__ inline uint32_t length_bucket(uint64_t e, uint32_t shift) {
  if (e == 0) return 0;
  const uint32_t msb = 63u - (uint32_t) __builtin_clzll(e);
  const uint32_t mask = (1u << shift) - 1u;
  const uint32_t sub = msb >= shift ? (uint32_t) (e >> (msb - shift)) & mask : (uint32_t) (e << (shift - msb)) & mask;
  return 1u + (msb << shift) + sub;
}

// Smallest element count of bucket b (b >= 1).
__device__ __host__ inline uint64_t bucket_low(uint32_t b, uint32_t shift) {
  const uint32_t msb = (b - 1) >> shift, sub = (b - 1) & ((1u << shift) - 1u);
  const uint64_t lead = (1ull << shift) | sub;
  return msb >= shift ? lead << (msb - shift) : lead >> (shift - msb);
}

// Every device array of a pipeline, passed to its small kernels by value.
struct PipeBufs {
  PipeDesc *desc;              // device copy of the batch descriptor
  const PipeDesc *host_desc;   // pinned host copy (device-mapped address)
  // host input: pinned staging (device-mapped addresses) and device copies
  const char *stage_seq;
  const uint64_t *stage_off;
  const uint8_t *stage_paired;
  char *seq;
  uint64_t *off;
  uint8_t *paired;
  // layout
  uint64_t *run_cap, *run_off, *tiles;
  uint16_t *bucket_of;
  uint32_t *bucket_count, *bucket_fill, *bucket_start, *unit_prefix, *per_unit, *seg_start, *seg_bucket,
      *n_seg;
  uint32_t *order, *unit_start, *n_units;
  unsigned long long *cursor;
  uint32_t *redo, *redo_count;  // [0] dense redo, [1] out of space, [2] out of space only
  uint32_t *space_list, *space_only;  // null with exact hitlist space
  PipeDesc *desc2;                    // descriptor of the out-of-space pass
  unsigned long long *pool_used;
  uint64_t max_space, pool;           // reserved space of a full batch; overflow pool (elements)
  uint32_t *dense_counts, *dense_lists, *dense_lens;
  unsigned int *flags;  // [0] bit 0: a mate of 2^31 or more bases, bit 1: batch over capacity,
                        //     bit 2: overflow pool exhausted; [1] OR of the status words
  unsigned long long *kmer_counts;
  uint32_t node_count;
  // long nucleotide fragments split over waves (null when the pipeline
  // does not split): per fragment its elements and segment count, the split
  // list, segment starts; per segment its layout and outputs; the scratch
  // of the segments' runs and records
  uint64_t *elements;
  unsigned long long *estats;  // [0] largest, [1] total elements, [2] elements of splittable fragments
  uint32_t *sp_cnt, *sp_list, *sp_count;
  uint64_t *sp_wide, *sp_start, *sp_tiles;
  PipeDesc *desc_seg;  // n = the batch's segments (the scan of their run spaces)
  uint32_t *sp_frag, *sp_mate, *sp_a, *sp_b, *sp_runs_n, *sp_nrec, *sp_groups, *sp_status, *sp_runs, *sp_recs;
  uint64_t *sp_cap, *sp_off;
  // results
  uint32_t *call, *groups, *total, *n_runs, *status, *n_recs, *runs, *recs;
  uint64_t *comp_off, *comp_rec_off;
  // host output (device-mapped addresses of pinned buffers)
  uint32_t *h_call, *h_groups, *h_total, *h_runs, *h_recs, *h_words;
  uint64_t *h_run_off, *h_rec_off, *h_kmer_counts;
};

// The descriptor: the host's pinned copy into device memory (one thread);
// with host input its arrays are the pipeline's device copies.
__global__ void k_pipe_desc(PipeBufs b, bool host_io) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  PipeDesc d = *b.host_desc;
  if (host_io) {
    d.seq = b.seq;
    d.off = b.off;
    d.paired = d.paired ? b.paired : nullptr;
  }
  *b.desc = d;
}

// Host input: the batch's bases, offsets and pair flags from pinned staging
// (read over the bus) into device memory.
__global__ __launch_bounds__(256) void k_pipe_upload(PipeBufs b) {
  const uint64_t n = b.desc->n, bases = b.desc->bases;
  const bool flags = b.desc->paired != nullptr;
  const uint64_t t = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x, step = (uint64_t) gridDim.x * blockDim.x;
  const uint64_t words = bases / 16;
  const uint4 *src = reinterpret_cast<const uint4 *>(b.stage_seq);
  uint4 *dst = reinterpret_cast<uint4 *>(b.seq);
  for (uint64_t i = t; i < words; i += step) dst[i] = src[i];
  for (uint64_t i = words * 16 + t; i < bases; i += step) b.seq[i] = b.stage_seq[i];
  for (uint64_t i = t; i < 2 * n + 1; i += step) b.off[i] = b.stage_off[i];
  if (flags)
    for (uint64_t i = t; i < n; i += step) b.paired[i] = b.stage_paired[i];
}

// Counters, cursors and flags of one batch back to zero (a fill, not a
// memset node).
__global__ __launch_bounds__(256) void k_pipe_reset(PipeBufs b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, step = gridDim.x * blockDim.x;
  for (uint32_t i = t; i < kMaxBuckets; i += step) b.bucket_count[i] = b.bucket_fill[i] = 0;
  for (uint32_t i = t; i < (uint32_t) kCursorWords; i += step) b.cursor[i] = 0;
  if (t < 4) b.redo_count[t] = 0;
  if (t < 2) b.flags[t] = 0;
  if (t == 0 && b.pool_used) *b.pool_used = 0;
  if (b.estats && t < 3) b.estats[t] = 0;
  if (b.sp_count && t < 2) b.sp_count[t] = 0;
  if (b.kmer_counts)
    for (uint32_t i = t; i < b.node_count; i += step) b.kmer_counts[i] = 0;
}

// Exact element count (the hitlist space) and length bucket per fragment,
// bucket histogram (LDS, then one atomic per bucket and block).
__global__ __launch_bounds__(256) void k_pipe_bounds(PipeBufs b, uint32_t k, uint32_t dna, uint32_t shift,
                                                      uint32_t pct) {
  __shared__ uint32_t hist[kMaxBuckets];
  for (uint32_t i = threadIdx.x; i < kMaxBuckets; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const PipeDesc &d = *b.desc;
  const uint64_t n = d.n;
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t e = 0;
  if (i < n) {
    const bool pr = d.paired ? d.paired[i] != 0 : d.paired_all != 0;
    const uint64_t l1 = d.off[2 * i + 1] - d.off[2 * i], l2 = pr ? d.off[2 * i + 2] - d.off[2 * i + 1] : 0;
    if ((l1 | l2) >> 31) atomicOr(&b.flags[0], 1u);
    e = fragment_elements(l1, l2, pr, k, dna != 0);
    b.run_cap[i] = run_space(e, pct);
    const uint32_t bk = length_bucket(e, shift);
    b.bucket_of[i] = (uint16_t) bk;
    atomicAdd(&hist[bk], 1u);
    if (b.elements) b.elements[i] = e;
  } else if (i == n) {
    b.run_cap[n] = 0;
  }
  if (b.estats) {  // largest, total and splittable elements: a wave's reduction, one atomic each
    uint64_t mx = e, sum = e, split = e > 2ull * kSegElements ? e : 0;
    for (int dlt = 1; dlt < kWave; dlt <<= 1) {
      const uint64_t o = (uint64_t) __shfl_xor((unsigned long long) mx, dlt);
      mx = o > mx ? o : mx;
      sum += (uint64_t) __shfl_xor((unsigned long long) sum, dlt);
      split += (uint64_t) __shfl_xor((unsigned long long) split, dlt);
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && sum) {
      atomicMax(&b.estats[0], (unsigned long long) mx);
      atomicAdd(&b.estats[1], (unsigned long long) sum);
      if (split) atomicAdd(&b.estats[2], (unsigned long long) split);
    }
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < kMaxBuckets; j += blockDim.x)
    if (hist[j]) atomicAdd(&b.bucket_count[j], hist[j]);
}

// Exclusive scan of v[0 .. n+1) (n from the descriptor) into out, in three
// kernels: tile sums, a scan of the tile sums in one block, the tiles.
__device__ __forceinline__ uint64_t block_exclusive(uint64_t x, uint64_t *wsum, uint64_t *total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint64_t inc = x;
#pragma unroll
  for (int dlt = 1; dlt < kWave; dlt <<= 1) {
    const uint64_t y = (uint64_t) __shfl_up((unsigned long long) inc, dlt);
    if (lane >= dlt) inc += y;
  }
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int j = 0; j < (int) (blockDim.x / kWave); j++) {
    if (j < w) before += wsum[j];
    all += wsum[j];
  }
  if (total) *total = all;
  __syncthreads();
  return before + inc - x;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const PipeDesc *d, const uint64_t *v, uint64_t *tiles) {
  __shared__ uint64_t wsum[256 / kWave];
  const uint64_t m = d->n + 1;
  const uint64_t base = (uint64_t) blockIdx.x * kTile;
  uint64_t s = 0;
  if (base < m)
    for (uint32_t j = 0; j < kTile / 256; j++) {
      const uint64_t i = base + j * 256 + threadIdx.x;
      if (i < m) s += v[i];
    }
  uint64_t tot = 0;
  (void) block_exclusive(s, wsum, &tot);
  if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kTopThreads) void k_scan_top(const PipeDesc *d, uint64_t *tiles) {
  __shared__ uint64_t wsum[kTopThreads / kWave];
  const uint64_t nt = (d->n + 1 + kTile - 1) / kTile;
  const uint64_t per = (nt + kTopThreads - 1) / kTopThreads;
  const uint64_t lo = threadIdx.x * per, hi = lo + per < nt ? lo + per : nt;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; i++) s += tiles[i];
  uint64_t run = block_exclusive(s, wsum, nullptr);
  for (uint64_t i = lo; i < hi; i++) {
    const uint64_t x = tiles[i];
    tiles[i] = run;
    run += x;
  }
}

__global__ __launch_bounds__(256) void k_scan_apply(const PipeDesc *d, const uint64_t *v, const uint64_t *tiles,
                                                     uint64_t *out) {
  __shared__ uint64_t wsum[256 / kWave];
  const uint64_t m = d->n + 1;
  const uint64_t base = (uint64_t) blockIdx.x * kTile;
  if (base >= m) return;
  uint64_t run = tiles[blockIdx.x];
  for (uint32_t j = 0; j < kTile / 256; j++) {
    const uint64_t i = base + j * 256 + threadIdx.x;
    const uint64_t x = i < m ? v[i] : 0;
    uint64_t tot = 0;
    const uint64_t ex = block_exclusive(x, wsum, &tot);
    if (i < m) out[i] = run + ex;
    run += tot;
  }
}

// Buckets laid out longest first, each cut into units of per_unit fragments
// (about unit_elements elements); segments list the non-empty buckets by
// position for the unit pass.
__global__ void k_pipe_buckets(PipeBufs b, uint32_t unit_elements, uint32_t shift, uint64_t capacity) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (b.run_off[b.desc->n] > capacity) {  // more hitlist space than allocated: run nothing
    b.flags[0] |= 2u;
    *b.n_seg = 0;
    *b.n_units = 0;
    return;
  }
  uint32_t pos = 0, units = 0, seg = 0;
  for (int bk = (int) kMaxBuckets - 1; bk >= 0; bk--) {
    const uint32_t c = b.bucket_count[bk];
    const uint64_t lo = bk ? bucket_low((uint32_t) bk, shift) : 0;
    const uint64_t want = lo ? unit_elements / lo : unit_elements;
    const uint32_t per = (uint32_t) (want > 1 ? want : 1);
    b.bucket_start[bk] = pos;
    b.per_unit[bk] = per;
    b.unit_prefix[bk] = units;
    if (c) {
      b.seg_start[seg] = pos;
      b.seg_bucket[seg] = (uint32_t) bk;
      seg++;
    }
    pos += c;
    units += (c + per - 1) / per;
  }
  *b.n_seg = seg;
  *b.n_units = units;
}

// Each fragment to its place in its bucket: an LDS rank within the block, one
// atomic per bucket and block to reserve the block's places.
__global__ __launch_bounds__(256) void k_pipe_scatter(PipeBufs b) {
  __shared__ uint32_t cnt[kMaxBuckets], base[kMaxBuckets];
  if (b.flags[0] & 2u) return;  // over capacity (block-uniform)
  for (uint32_t i = threadIdx.x; i < kMaxBuckets; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  const uint64_t n = b.desc->n;
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bk = 0, rank = 0;
  if (i < n) {
    bk = b.bucket_of[i];
    rank = atomicAdd(&cnt[bk], 1u);
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < kMaxBuckets; j += blockDim.x)
    base[j] = cnt[j] ? atomicAdd(&b.bucket_fill[j], cnt[j]) : 0;
  __syncthreads();
  if (i < n) b.order[b.bucket_start[bk] + base[bk] + rank] = (uint32_t) i;
}

// Unit starts: position p starts a unit when it is a multiple of its bucket's
// per_unit past the bucket's first position.
__global__ __launch_bounds__(256) void k_pipe_units(PipeBufs b) {
  __shared__ uint32_t s_start[kMaxBuckets], s_bucket[kMaxBuckets];
  const uint32_t ns = *b.n_seg;
  for (uint32_t j = threadIdx.x; j < ns; j += blockDim.x) {
    s_start[j] = b.seg_start[j];
    s_bucket[j] = b.seg_bucket[j];
  }
  __syncthreads();
  const uint64_t n = b.desc->n;
  const uint64_t p = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) b.unit_start[*b.n_units] = (uint32_t) n;
  if (p >= n || ns == 0) return;
  uint32_t lo = 0, hi = ns;  // last segment starting at or before p
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (s_start[mid] <= p) lo = mid; else hi = mid;
  }
  const uint32_t bk = s_bucket[lo];
  const uint32_t r = (uint32_t) p - s_start[lo], per = b.per_unit[bk];
  if (r % per == 0) b.unit_start[b.unit_prefix[bk] + r / per] = (uint32_t) p;
}

// Fragments whose hit table overflowed (status bit 0) listed for the dense
// redo; the OR of all status words kept for the host.
__global__ __launch_bounds__(256) void k_pipe_collect(PipeBufs b) {
  if (b.flags[0] & 2u) return;  // over capacity: no status was written (block-uniform)
  const uint64_t n = b.desc->n;
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t st = i < n ? b.status[i] : 0u;
  if (st & 1u) b.redo[atomicAdd(&b.redo_count[0], 1u)] = (uint32_t) i;
  if (b.space_list && (st & 2u)) b.space_list[atomicAdd(&b.redo_count[1], 1u)] = (uint32_t) i;
  if (b.space_list && st == 2u) b.space_only[atomicAdd(&b.redo_count[2], 1u)] = (uint32_t) i;
  uint32_t acc = st;
  for (int dlt = 1; dlt < kWave; dlt <<= 1) acc |= (uint32_t) __shfl_xor((unsigned int) acc, dlt);
  if ((threadIdx.x & (kWave - 1)) == 0 && acc) atomicOr(&b.flags[1], acc);
}

// Out-of-space fragments get exact space from the overflow pool (after the
// batch's reserved space); a fragment the pool cannot hold keeps its
// reservation (the classify kernels never write past run_cap) and flags the
// batch.
__global__ __launch_bounds__(256) void k_pipe_space(PipeBufs b, uint32_t k, uint32_t dna) {
  const uint32_t m = b.redo_count[1];
  const PipeDesc &d = *b.desc;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
    const uint32_t i = b.space_list[j];
    const bool pr = d.paired ? d.paired[i] != 0 : d.paired_all != 0;
    const uint64_t l1 = d.off[2 * i + 1] - d.off[2 * i], l2 = pr ? d.off[2 * i + 2] - d.off[2 * i + 1] : 0;
    const uint64_t e = fragment_elements(l1, l2, pr, k, dna != 0);
    const uint64_t pos = atomicAdd(b.pool_used, (unsigned long long) e);
    if (pos + e > b.pool) {
      atomicOr(&b.flags[0], 4u);
      continue;
    }
    b.run_off[i] = b.max_space + pos;
    b.run_cap[i] = e;
  }
}

// The out-of-space pass's descriptor (the batch's, with the out-of-space-only
// count) and its cleared cursors.
__global__ void k_pipe_desc2(PipeBufs b) {
  if (blockIdx.x != 0) return;
  for (uint32_t i = threadIdx.x; i < (uint32_t) kCursorWords; i += blockDim.x) b.cursor[i] = 0;
  if (threadIdx.x == 0) {
    PipeDesc d = *b.desc;
    d.n = (b.flags[0] & 4u) ? 0 : b.redo_count[2];
    *b.desc2 = d;
  }
}

// The segments' count as the descriptor of the scan of their run spaces
// (whose last entry is zeroed for it).
__global__ void k_pipe_seg_desc(PipeBufs b) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t segs = b.sp_count[0];
  b.sp_cap[segs] = 0;
  PipeDesc d = *b.desc;
  d.n = segs;
  *b.desc_seg = d;
}

// counts (u32, n from the descriptor) widened into v[0 .. n] (v[n] = 0).
__global__ __launch_bounds__(256) void k_pipe_widen(const PipeDesc *d, const uint32_t *counts, uint64_t *v) {
  const uint64_t n = d->n;
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = counts[i];
  if (i == n) v[n] = 0;
}

// Host output: per-fragment scalars and dense offsets into pinned memory.
__global__ __launch_bounds__(256) void k_pipe_out(PipeBufs b, bool hll) {
  if ((uint64_t) blockIdx.x * blockDim.x + threadIdx.x == 0) {
    b.h_words[0] = b.flags[0];
    b.h_words[1] = b.flags[1];
    b.h_words[2] = b.redo_count[0];
    b.h_words[3] = b.redo_count[1];
    b.h_words[4] = b.sp_count ? b.sp_count[1] : 0u;  // fragments split over waves
    b.h_words[5] = b.sp_count ? b.sp_count[0] : 0u;  // and their segments
  }
  if (b.flags[0] & 2u) return;  // over capacity
  const uint64_t n = b.desc->n;
  const uint64_t t = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x, step = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = t; i <= n; i += step) {
    if (i < n) {
      b.h_call[i] = b.call[i];
      b.h_groups[i] = b.groups[i];
      b.h_total[i] = b.total[i];
    }
    b.h_run_off[i] = b.comp_off[i];
    if (hll) b.h_rec_off[i] = b.comp_rec_off[i];
  }
  if (b.kmer_counts)
    for (uint64_t i = t; i < b.node_count; i += step) b.h_kmer_counts[i] = b.kmer_counts[i];
}

// Device results: the flags, the status OR and the redo counts into pinned
// host words (the writeback of host results does it itself).
__global__ void k_pipe_flags(PipeBufs b) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    b.h_words[0] = b.flags[0];
    b.h_words[1] = b.flags[1];
    b.h_words[2] = b.redo_count[0];
    b.h_words[3] = b.redo_count[1];
    b.h_words[4] = b.sp_count ? b.sp_count[1] : 0u;  // fragments split over waves
    b.h_words[5] = b.sp_count ? b.sp_count[0] : 0u;  // and their segments
  }
}

// Each fragment's runs (or records) from its exact space to its dense place
// in pinned host memory.
__global__ __launch_bounds__(64) void k_pipe_compact(const PipeDesc *d, const unsigned int *flags,
                                                     const uint64_t *src_off, const uint32_t *counts,
                                                     const uint64_t *dst_off, const uint32_t *src, uint32_t *dst) {
  if (flags[0] & 2u) return;  // a batch over capacity ran nothing
  const uint64_t n = d->n;
  for (uint64_t f = blockIdx.x; f < n; f += gridDim.x) {
    const uint64_t s = src_off[f], o = dst_off[f];
    const uint32_t c = counts[f];
    for (uint32_t j = threadIdx.x; j < 2 * c; j += 64) dst[2 * o + j] = src[2 * s + j];
  }
}

// Strided sample of a device-resident fragment set (for tuning): lengths,
// then (after a scan on the host side of the library) the bases.
__global__ void k_sample_lengths(const uint64_t *off, uint64_t m, uint64_t stride, uint64_t *len) {
  const uint64_t j = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) {
    const uint64_t i = j * stride;
    len[2 * j] = off[2 * i + 1] - off[2 * i];
    len[2 * j + 1] = off[2 * i + 2] - off[2 * i + 1];
  } else if (j == m) {
    len[2 * m] = 0;
  }
}

__global__ __launch_bounds__(64) void k_sample_copy(const char *seq, const uint64_t *off, const uint8_t *paired,
                                                    uint64_t m, uint64_t stride, const uint64_t *new_off,
                                                    char *out, uint8_t *out_paired) {
  for (uint64_t j = blockIdx.x; j < m; j += gridDim.x) {
    const uint64_t i = j * stride, a = off[2 * i], len = off[2 * i + 2] - a, o = new_off[2 * j];
    for (uint64_t q = threadIdx.x; q < len; q += 64) out[o + q] = seq[a + q];
    if (threadIdx.x == 0 && paired) out_paired[j] = paired[i];
  }
}

}  // namespace

// Pipeline kernels of each classify variant.
struct PipeKernels {
  void (*classify)(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, unsigned long long *);
  void (*dense)(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, const uint32_t *, const uint32_t *,
                DenseScratch);
  // the long-fragment split (null: the variant does not split)
  void (*segments)(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, SegBatch);
  void (*stitch)(DbView, ClassifyParams, const PipeDesc *, FragBatch, FragOut, SegBatch);
};
static const PipeKernels kPipeVariants[3] = {
    {k_classify_pipe_dna, k_classify_dense_pipe_dna, k_classify_segments_pipe_dna, k_stitch_pipe_dna},
    {k_classify_pipe_dna_x, k_classify_dense_pipe_dna, k_classify_segments_pipe_dna_x, k_stitch_pipe_dna_x},
    {k_classify_pipe_prot, k_classify_dense_pipe_prot, nullptr, nullptr},
};

}  // namespace k2g

using namespace k2g;

struct k2g_pipeline {
  k2g_db *db = nullptr;
  k2g_classify_params params{};
  k2g_pipeline_params pp{};
  bool host_io = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t max_elements = 0, max_space = 0, pool = 0;
  uint64_t split_scratch = 0, split_segs = 0;  // 0: no split
  uint32_t pct = 100;
  uint32_t dense_slots = 0;
  int grid = 0;  // persistent workgroups
  std::vector<void *> dev;   // device allocations
  std::vector<void *> host;  // pinned host allocations
  PipeBufs b{};
  PipeDesc *h_desc = nullptr;  // pinned
  char *h_seq = nullptr;
  uint64_t *h_off = nullptr;
  uint8_t *h_paired = nullptr;
  uint32_t *h_call = nullptr, *h_groups = nullptr, *h_total = nullptr, *h_runs = nullptr, *h_recs = nullptr,
           *h_words = nullptr;
  uint64_t *h_run_off = nullptr, *h_rec_off = nullptr, *h_kmer = nullptr;
  uint64_t last_n = 0, last_dense = 0, last_space = 0, last_split = 0, last_segments = 0;
  float last_ms = 0;
};

namespace {

template <class T>
hipError_t dalloc(k2g_pipeline *p, T **ptr, uint64_t n) {
  *ptr = nullptr;
  hipError_t e = hipMalloc((void **) ptr, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) p->dev.push_back(*ptr);
  return e;
}

template <class T>
hipError_t halloc(k2g_pipeline *p, T **ptr, uint64_t n) {
  *ptr = nullptr;
  hipError_t e = hipHostMalloc((void **) ptr, std::max<uint64_t>(n, 1) * sizeof(T), hipHostMallocMapped);
  if (e == hipSuccess) p->host.push_back(*ptr);
  return e;
}

template <class T>
T *mapped(T *host_ptr) {
  T *d = nullptr;
  return hipHostGetDevicePointer((void **) &d, host_ptr, 0) == hipSuccess ? d : nullptr;
}

void release(k2g_pipeline *p) {
  if (p->exec) (void) hipGraphExecDestroy(p->exec);
  if (p->graph) (void) hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  for (void *x : p->dev) (void) hipFree(x);
  for (void *x : p->host) (void) hipHostFree(x);
  p->dev.clear();
  p->host.clear();
}

int variant_of(const k2g_db *db, const k2g_classify_params &p) {
  const bool extras = p.want_hll || p.want_kmer_counts || p.want_probe_stats || p.quick;
  return db->view.dna ? (extras ? 1 : 0) : 2;
}

// Records the chain on the pipeline's stream and instantiates it.
int capture(k2g_pipeline *p) {
  const DbView &v = p->db->view;
  const PipeBufs &b = p->b;
  hipStream_t s = p->stream;
  const uint64_t nmax = p->pp.max_fragments;
  const unsigned g256 = (unsigned) ((nmax + 1 + 255) / 256);
  const unsigned tiles = (unsigned) ((nmax + 1 + kTile - 1) / kTile);
  const int vi = variant_of(p->db, p->params);
  const PipeKernels &K = kPipeVariants[vi];
  const uint32_t shift = (uint32_t) p->pp.bucket_shift;
  const uint32_t T = p->pp.unit_elements > 0 ? (uint32_t) p->pp.unit_elements : 4096u;
  ClassifyParams cp{};
  cp.confidence = p->params.confidence;
  cp.min_hit_groups = p->params.min_hit_groups;
  cp.quick = p->params.quick ? 1u : 0u;
  cp.tab_cap = p->params.table_slots ? (uint32_t) p->params.table_slots : 256u;
  cp.batch_chunks = p->pp.probe_depth > 0 ? (uint32_t) p->pp.probe_depth : 8u;
  FragBatch in{nullptr, nullptr, nullptr, 0, 0, b.order, b.unit_start, b.n_units};
  FragOut out{b.call, b.groups, b.total, b.n_runs, b.status, b.n_recs, b.run_off, b.run_cap, b.runs,
              p->params.want_hll ? b.recs : nullptr, p->params.want_kmer_counts ? b.kmer_counts : nullptr,
              nullptr, nullptr};
  const DenseScratch ds{b.dense_counts, b.dense_lists, b.dense_lens, p->dense_slots};
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_pipe_desc, dim3(1), dim3(64), 0, s, b, p->host_io);
  if (p->host_io) hipLaunchKernelGGL(k_pipe_upload, dim3(2048), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_pipe_reset, dim3(64), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_pipe_bounds, dim3(g256), dim3(256), 0, s, b, v.k, v.dna, shift, p->pct);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap, b.tiles);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, (const PipeDesc *) b.desc, b.tiles);
  hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap, b.tiles,
                     b.run_off);
  hipLaunchKernelGGL(k_pipe_buckets, dim3(1), dim3(64), 0, s, b, T, shift, p->max_space);
  hipLaunchKernelGGL(k_pipe_scatter, dim3(g256), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_pipe_units, dim3(g256), dim3(256), 0, s, b);
  // long fragments split over waves (when the device finds it pays): their
  // segment counts and starts, each mate's cut points and the segments' run
  // spaces, then the segments; the classify kernel skips those fragments and
  // the stitch joins them
  const bool split = b.sp_cnt != nullptr;
  const SegBatch sg{b.sp_frag,   b.sp_mate,  b.sp_a,      b.sp_b,      b.sp_off,    b.sp_cap,
                    b.sp_runs_n, b.sp_nrec,  b.sp_groups, b.sp_status, b.sp_runs,
                    p->params.want_hll ? b.sp_recs : nullptr,
                    b.sp_count,  b.sp_list,  b.sp_cnt,    b.sp_start};
  if (split) {
    const unsigned seg_tiles = (unsigned) ((p->split_segs + 1 + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_seg_count_pipe, dim3(g256), dim3(256), 0, s, (const PipeDesc *) b.desc, v.k,
                       (uint32_t) kSegElements, (const uint64_t *) b.elements, (const unsigned long long *) b.estats,
                       (uint64_t) p->grid, p->split_scratch, b.sp_cnt, b.sp_list, b.sp_count);
    hipLaunchKernelGGL(k_pipe_widen, dim3(g256), dim3(256), 0, s, (const PipeDesc *) b.desc,
                       (const uint32_t *) b.sp_cnt, b.sp_wide);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.sp_wide, b.tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, (const PipeDesc *) b.desc, b.tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.sp_wide, b.tiles,
                       b.sp_start);
    hipLaunchKernelGGL(k_seg_plan_pipe, dim3((unsigned) std::min<uint64_t>(nmax, 4096)), dim3(64), 0, s,
                       (const PipeDesc *) b.desc, v.k, (uint32_t) kSegElements, 100u, (const uint32_t *) b.sp_list,
                       (const uint64_t *) b.sp_start, b.sp_count, sg);
    hipLaunchKernelGGL(k_pipe_seg_desc, dim3(1), dim3(64), 0, s, b);
    hipLaunchKernelGGL(k_scan_tiles, dim3(seg_tiles), dim3(256), 0, s, (const PipeDesc *) b.desc_seg, b.sp_cap,
                       b.sp_tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, (const PipeDesc *) b.desc_seg, b.sp_tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(seg_tiles), dim3(256), 0, s, (const PipeDesc *) b.desc_seg, b.sp_cap,
                       b.sp_tiles, b.sp_off);
    hipLaunchKernelGGL(K.segments, dim3((unsigned) p->grid), dim3(64), 0, s, v, cp, (const PipeDesc *) b.desc, in,
                       out, sg);
    in.split = b.sp_cnt;
  }
  hipLaunchKernelGGL(K.classify, dim3((unsigned) p->grid), dim3(64), 0, s, v, cp, (const PipeDesc *) b.desc, in,
                     out, b.cursor);
  if (split)
    hipLaunchKernelGGL(K.stitch, dim3((unsigned) std::min<uint64_t>(nmax, 16384)), dim3(64), 0, s, v, cp,
                       (const PipeDesc *) b.desc, in, out, sg);
  hipLaunchKernelGGL(k_pipe_collect, dim3(g256), dim3(256), 0, s, b);
  FragOut out2 = out;
  out2.kmer_counts = nullptr;  // the first pass counted every fragment
  if (b.space_list) {
    // out of reserved space: exact space from the pool, then the fragments
    // that only lacked space classified again (in list order)
    hipLaunchKernelGGL(k_pipe_space, dim3(256), dim3(256), 0, s, b, v.k, v.dna);
    hipLaunchKernelGGL(k_pipe_desc2, dim3(1), dim3(64), 0, s, b);
    FragBatch in2{nullptr, nullptr, nullptr, 0, 0, b.space_only, nullptr, nullptr};
    hipLaunchKernelGGL(K.classify, dim3((unsigned) p->grid), dim3(64), 0, s, v, cp, (const PipeDesc *) b.desc2,
                       in2, out2, b.cursor);
  }
  hipLaunchKernelGGL(K.dense, dim3(p->dense_slots), dim3(64), 0, s, v, cp, (const PipeDesc *) b.desc, in, out2,
                     (const uint32_t *) b.redo, (const uint32_t *) b.redo_count, ds);
  if (p->host_io) {
    // dense offsets of the runs (and records), scalars and offsets, then the
    // runs themselves, all into pinned host memory
    hipLaunchKernelGGL(k_pipe_widen, dim3(g256), dim3(256), 0, s, (const PipeDesc *) b.desc,
                       (const uint32_t *) b.n_runs, b.run_cap);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap, b.tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, (const PipeDesc *) b.desc, b.tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap, b.tiles,
                       b.comp_off);
    if (p->params.want_hll) {
      hipLaunchKernelGGL(k_pipe_widen, dim3(g256), dim3(256), 0, s, (const PipeDesc *) b.desc,
                         (const uint32_t *) b.n_recs, b.run_cap);
      hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap,
                         b.tiles);
      hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, (const PipeDesc *) b.desc, b.tiles);
      hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(256), 0, s, (const PipeDesc *) b.desc, b.run_cap, b.tiles,
                         b.comp_rec_off);
    }
    hipLaunchKernelGGL(k_pipe_out, dim3(256), dim3(256), 0, s, b, p->params.want_hll != 0);
    const unsigned gc = (unsigned) std::min<uint64_t>(nmax, 65536);
    hipLaunchKernelGGL(k_pipe_compact, dim3(std::max(1u, gc)), dim3(64), 0, s, (const PipeDesc *) b.desc,
                       (const unsigned int *) b.flags, (const uint64_t *) b.run_off, (const uint32_t *) b.n_runs, (const uint64_t *) b.comp_off,
                       (const uint32_t *) b.runs, b.h_runs);
    if (p->params.want_hll)
      hipLaunchKernelGGL(k_pipe_compact, dim3(std::max(1u, gc)), dim3(64), 0, s, (const PipeDesc *) b.desc,
                         (const unsigned int *) b.flags, (const uint64_t *) b.run_off, (const uint32_t *) b.n_recs, (const uint64_t *) b.comp_rec_off,
                         (const uint32_t *) b.recs, b.h_recs);
  } else {
    hipLaunchKernelGGL(k_pipe_flags, dim3(1), dim3(64), 0, s, b);
  }
  hipGraph_t g = nullptr;
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(s, &g);
  if (le != hipSuccess) {
    if (g) (void) hipGraphDestroy(g);
    return fail(K2G_ESOFTWARE, std::string("pipeline capture: ") + hipGetErrorString(le));
  }
  HIP_TRY(ce);
  p->graph = g;
  HIP_TRY(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
  return K2G_OK;
}

}  // namespace

namespace k2g {
__global__ void k_digest(const uint32_t *call, const uint32_t *groups, const uint32_t *total,
                         const uint32_t *n_runs, const uint64_t *run_off, const uint32_t *runs, uint64_t n,
                         uint64_t *out);

}  // namespace k2g

namespace {

int check_pipeline_params(const k2g_classify_params *p, const k2g_pipeline_params *pp) {
  if (!(p->confidence >= 0 && p->confidence <= 1))
    return fail(K2G_EUSAGE, "confidence threshold must be in [0, 1]");
  const int ts = p->table_slots;
  if (ts != 0 && (ts < 2 || ts > 256 || (ts & (ts - 1)) != 0))
    return fail(K2G_EUSAGE, "table_slots must be a power of two in [2, 256]");
  if (p->run_cap_pct < 0 || p->run_cap_pct > 100) return fail(K2G_EUSAGE, "run_cap_pct must be in [0, 100]");
  if (p->want_probe_stats) return fail(K2G_EUSAGE, "probe statistics need k2g_classify_device");
  if (p->want_hll < 0 || p->want_hll > 1)
    return fail(K2G_EUSAGE, "a pipeline returns per-hit HLL records (want_hll 0 or 1)");
  if (pp->max_fragments == 0 || pp->max_fragments >= 0xFFFFFFFFull)
    return fail(K2G_EUSAGE, "max_fragments must be in [1, 2^32 - 2]");
  if (pp->bucket_shift < 0 || pp->bucket_shift > (int) kMaxShift)
    return fail(K2G_EUSAGE, "bucket_shift must be in [0, 2]");
  if (pp->probe_depth < 0 || pp->probe_depth > 8) return fail(K2G_EUSAGE, "probe_depth must be in [0, 8]");
  if (pp->unit_elements < 0 || pp->waves_per_cu < 0) return fail(K2G_EUSAGE, "negative tuning parameter");
  return K2G_OK;
}

// Grid of the classify kernel: waves_per_cu resident waves per CU (at most
// what its occupancy allows), at least one per cursor.
int pipe_grid(k2g_pipeline *pl, int *grid) {
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kPipeVariants[variant_of(pl->db, pl->params)].classify,
                                                       64, 0));
  per_cu = std::max(1, per_cu);
  const int w = pl->pp.waves_per_cu > 0 ? std::min(pl->pp.waves_per_cu, per_cu) : per_cu;
  *grid = std::max(8, w * std::max(1, pl->db->num_cus));
  return K2G_OK;
}

// (Re)records the graph with the current tuning parameters.
int recapture(k2g_pipeline *pl) {
  if (pl->exec) (void) hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void) hipGraphDestroy(pl->graph);
  pl->exec = nullptr;
  pl->graph = nullptr;
  int rc = pipe_grid(pl, &pl->grid);
  if (rc) return rc;
  return capture(pl);
}

}  // namespace

extern "C" int k2g_pipeline_create(k2g_db *db, const k2g_classify_params *p, const k2g_pipeline_params *pp,
                                   k2g_pipeline **out) {
  if (!db || !p || !pp || !out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  int rc = check_pipeline_params(p, pp);
  if (rc) return rc;
  auto *pl = new k2g_pipeline();
  pl->db = db;
  pl->params = *p;
  pl->pp = *pp;
  if (pl->pp.unit_elements == 0) pl->pp.unit_elements = 4096;
  if (pl->pp.probe_depth == 0) pl->pp.probe_depth = 8;
  pl->host_io = pp->host_io != 0;
  pl->pct = pl->host_io ? 100u : (p->run_cap_pct ? (uint32_t) p->run_cap_pct : 25u);
  const DbView &v = db->view;
  const uint64_t nmax = pp->max_fragments, bmax = pp->max_bases;
  // element bound of a batch: a nucleotide fragment has at most one element
  // per base plus its mate marker; a translated one at most 2L + 11 per mate
  pl->max_elements = v.dna ? bmax + nmax : 2 * bmax + 23 * nmax;
  pl->max_space = pl->pct >= 100 ? pl->max_elements : pl->max_elements * pl->pct / 100 + 64 * nmax;
  // overflow pool of reserved-space pipelines: an eighth of the reserved
  // space, at least 2^20 elements, never more than a batch can need
  pl->pool = pl->pct >= 100 ? 0 : std::min(pl->max_elements, std::max<uint64_t>(pl->max_space / 8, 1u << 20));
  pl->dense_slots = dense_slots_for(nmax, v.node_count);
  const bool hll = p->want_hll != 0, kc = p->want_kmer_counts != 0;
  PipeBufs &b = pl->b;
  auto fin = [&](int code) {
    k2g_pipeline_destroy(pl);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&pl->ev0);
  if (e == hipSuccess) e = hipEventCreate(&pl->ev1);
#define K2G_D(ptr, n) \
  if (e == hipSuccess) e = dalloc(pl, &(ptr), (n))
#define K2G_H(ptr, n) \
  if (e == hipSuccess) e = halloc(pl, &(ptr), (n))
  K2G_D(b.desc, 1);
  K2G_D(b.run_cap, nmax + 1);
  K2G_D(b.run_off, nmax + 1);
  K2G_D(b.tiles, (nmax + 1) / kTile + 2);
  K2G_D(b.bucket_of, nmax);
  K2G_D(b.bucket_count, kMaxBuckets);
  K2G_D(b.bucket_fill, kMaxBuckets);
  K2G_D(b.bucket_start, kMaxBuckets);
  K2G_D(b.unit_prefix, kMaxBuckets);
  K2G_D(b.per_unit, kMaxBuckets);
  K2G_D(b.seg_start, kMaxBuckets);
  K2G_D(b.seg_bucket, kMaxBuckets);
  K2G_D(b.n_seg, 1);
  K2G_D(b.order, nmax);
  K2G_D(b.unit_start, nmax + 1);
  K2G_D(b.n_units, 1);
  K2G_D(b.cursor, kCursorWords);
  K2G_D(b.redo, nmax);
  K2G_D(b.redo_count, 4);
  if (pl->pool) {
    K2G_D(b.space_list, nmax);
    K2G_D(b.space_only, nmax);
    K2G_D(b.desc2, 1);
    K2G_D(b.pool_used, 1);
  }
  K2G_D(b.dense_counts, (uint64_t) pl->dense_slots * v.node_count);
  K2G_D(b.dense_lists, (uint64_t) pl->dense_slots * v.node_count);
  K2G_D(b.dense_lens, pl->dense_slots);
  K2G_D(b.flags, 2);
  if (kc) K2G_D(b.kmer_counts, v.node_count);
  if (kPipeVariants[variant_of(db, *p)].segments && !p->quick) {
    // the split's scratch holds the runs of the split fragments (at most 2^28
    // elements; a batch whose split fragments need more is not split)
    pl->split_scratch = std::max<uint64_t>(std::min<uint64_t>(pl->max_elements, 1ull << 28), 1);
    pl->split_segs = std::min<uint64_t>(2 * pl->split_scratch / kSegElements + 2,
                                        pl->split_scratch / kSegElements + 2 * nmax + 2);
    const uint64_t sn = pl->split_segs;
    K2G_D(b.elements, nmax);
    K2G_D(b.estats, 3);
    K2G_D(b.sp_cnt, nmax + 1);
    K2G_D(b.sp_list, nmax);
    K2G_D(b.sp_count, 2);
    K2G_D(b.sp_wide, nmax + 1);
    K2G_D(b.sp_start, nmax + 1);
    K2G_D(b.sp_tiles, (sn + 1) / kTile + 2);
    K2G_D(b.desc_seg, 1);
    K2G_D(b.sp_frag, sn);
    K2G_D(b.sp_mate, sn);
    K2G_D(b.sp_a, sn);
    K2G_D(b.sp_b, sn);
    K2G_D(b.sp_runs_n, sn);
    K2G_D(b.sp_nrec, sn);
    K2G_D(b.sp_groups, sn);
    K2G_D(b.sp_status, sn);
    K2G_D(b.sp_cap, sn + 1);
    K2G_D(b.sp_off, sn + 1);
    K2G_D(b.sp_runs, 2 * pl->split_scratch);
    if (hll) K2G_D(b.sp_recs, 2 * pl->split_scratch);
  }
  K2G_D(b.call, nmax);
  K2G_D(b.groups, nmax);
  K2G_D(b.total, nmax);
  K2G_D(b.n_runs, nmax);
  K2G_D(b.status, nmax);
  K2G_D(b.n_recs, nmax);
  K2G_D(b.runs, 2 * (pl->max_space + pl->pool));
  if (hll) K2G_D(b.recs, 2 * (pl->max_space + pl->pool));
  K2G_H(pl->h_desc, 1);
  K2G_H(pl->h_words, 6);
  if (pl->host_io) {
    K2G_D(b.seq, bmax + 16);
    K2G_D(b.off, 2 * nmax + 1);
    K2G_D(b.paired, nmax);
    K2G_D(b.comp_off, nmax + 1);
    if (hll) K2G_D(b.comp_rec_off, nmax + 1);
    K2G_H(pl->h_seq, bmax + 16);
    K2G_H(pl->h_off, 2 * nmax + 1);
    K2G_H(pl->h_paired, nmax);
    K2G_H(pl->h_call, nmax);
    K2G_H(pl->h_groups, nmax);
    K2G_H(pl->h_total, nmax);
    K2G_H(pl->h_run_off, nmax + 1);
    K2G_H(pl->h_runs, 2 * pl->max_space);
    if (hll) K2G_H(pl->h_rec_off, nmax + 1);
    if (hll) K2G_H(pl->h_recs, 2 * pl->max_space);
    if (kc) K2G_H(pl->h_kmer, v.node_count);
  }
#undef K2G_D
#undef K2G_H
  if (e != hipSuccess)
    return fin(fail(K2G_EOS, std::string("pipeline allocation: ") + hipGetErrorString(e)));
  b.node_count = v.node_count;
  b.max_space = pl->max_space;
  b.pool = pl->pool;
  b.host_desc = mapped(pl->h_desc);
  b.h_words = mapped(pl->h_words);
  if (pl->host_io) {
    b.stage_seq = mapped(pl->h_seq);
    b.stage_off = mapped(pl->h_off);
    b.stage_paired = mapped(pl->h_paired);
    b.h_call = mapped(pl->h_call);
    b.h_groups = mapped(pl->h_groups);
    b.h_total = mapped(pl->h_total);
    b.h_run_off = mapped(pl->h_run_off);
    b.h_runs = mapped(pl->h_runs);
    if (hll) b.h_rec_off = mapped(pl->h_rec_off);
    if (hll) b.h_recs = mapped(pl->h_recs);
    if (kc) b.h_kmer_counts = mapped(pl->h_kmer);
    if (!b.stage_seq || !b.stage_off || !b.h_call || !b.h_runs)
      return fin(fail(K2G_ESOFTWARE, "pipeline: pinned buffers are not mapped"));
  }
  if (!b.host_desc || !b.h_words) return fin(fail(K2G_ESOFTWARE, "pipeline: pinned buffers are not mapped"));
  rc = recapture(pl);
  if (rc) return fin(rc);
  *out = pl;
  return K2G_OK;
}

extern "C" int k2g_pipeline_destroy(k2g_pipeline *pl) {
  if (!pl) return K2G_OK;
  if (pl->stream) (void) hipStreamSynchronize(pl->stream);
  release(pl);
  if (pl->ev0) (void) hipEventDestroy(pl->ev0);
  if (pl->ev1) (void) hipEventDestroy(pl->ev1);
  if (pl->stream) (void) hipStreamDestroy(pl->stream);
  delete pl;
  return K2G_OK;
}

namespace {
// One graph launch and its wait; then the device's flags.
int launch(k2g_pipeline *pl) {
  HIP_TRY(hipEventRecord(pl->ev0, pl->stream));
  HIP_TRY(hipGraphLaunch(pl->exec, pl->stream));
  HIP_TRY(hipEventRecord(pl->ev1, pl->stream));
  HIP_TRY(hipStreamSynchronize(pl->stream));
  HIP_TRY(hipEventElapsedTime(&pl->last_ms, pl->ev0, pl->ev1));
  if (pl->h_words[0] & 1u) return fail(K2G_EUSAGE, "a read has 2^31 or more bases");
  if (pl->h_words[0] & 2u)
    return fail(K2G_EUSAGE, "the batch's hitlist space exceeds the pipeline's capacity (max_bases)");
  if (pl->h_words[0] & 4u)
    return fail(K2G_EUSAGE, "fragments outgrew their reserved hitlist space by more than the overflow pool (" +
                                std::to_string(pl->pool) + " elements) holds; raise run_cap_pct (100 is exact)");
  pl->last_dense = pl->h_words[2];
  pl->last_space = pl->h_words[3];
  pl->last_split = pl->h_words[4];
  pl->last_segments = pl->h_words[5];
  return K2G_OK;
}
}  // namespace

extern "C" int k2g_pipeline_run(k2g_pipeline *pl, const k2g_fragments *in, k2g_results *out) {
  if (!pl || !in || !out) return fail(K2G_EUSAGE, "null argument");
  if (!pl->host_io) return fail(K2G_EUSAGE, "a device pipeline runs with k2g_pipeline_run_device");
  const uint64_t n = in->n;
  if (n > pl->pp.max_fragments) return fail(K2G_EUSAGE, "batch has more fragments than the pipeline's capacity");
  if (n && (!in->seq || !in->off)) return fail(K2G_EUSAGE, "null fragment arrays");
  const uint64_t bases = n ? in->off[2 * n] : 0;
  if (bases > pl->pp.max_bases) return fail(K2G_EUSAGE, "batch has more bases than the pipeline's capacity");
  for (uint64_t i = 0; i < 2 * n; i++)
    if (in->off[i + 1] < in->off[i]) return fail(K2G_EUSAGE, "fragment offsets must not decrease");
  for (uint64_t i = 0; i < n; i++) {
    const bool pr = in->paired ? in->paired[i] != 0 : pl->params.paired != 0;
    if (((in->off[2 * i + 1] - in->off[2 * i]) | (pr ? in->off[2 * i + 2] - in->off[2 * i + 1] : 0)) >> 31)
      return fail(K2G_EUSAGE, "a read has 2^31 or more bases");
  }
  // into the pinned staging (a caller that filled it in place skips the copies)
  if (bases && in->seq != pl->h_seq) std::memcpy(pl->h_seq, in->seq, bases);
  if (in->off != pl->h_off) std::memcpy(pl->h_off, in->off, (2 * n + 1) * 8);
  if (in->paired && in->paired != pl->h_paired) std::memcpy(pl->h_paired, in->paired, n);
  PipeDesc d{};
  d.seq = nullptr;
  d.off = nullptr;
  d.paired = in->paired ? reinterpret_cast<const uint8_t *>(pl->h_paired) : nullptr;  // marks per-fragment flags
  d.n = n;
  d.bases = bases;
  d.paired_all = pl->params.paired != 0;
  *pl->h_desc = d;
  int rc = launch(pl);
  if (rc) return rc;
  if (pl->h_words[1] & 2u) return fail(K2G_ESOFTWARE, "internal error: hitlist exceeded its bound");
  pl->last_n = n;
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->node_count = pl->db->view.node_count;
  out->call = pl->h_call;
  out->hit_groups = pl->h_groups;
  out->total_kmers = pl->h_total;
  out->run_off = pl->h_run_off;
  out->runs = pl->h_runs;
  // an empty batch has no HLL records or k-mer counts (as k2g_classify)
  if (n && pl->params.want_hll) {
    out->hll_off = pl->h_rec_off;
    out->hll = pl->h_recs;
  }
  if (n && pl->params.want_kmer_counts) out->kmer_counts = pl->h_kmer;
  out->redone = pl->h_words[2];
  out->kernel_ms = pl->last_ms;
  return K2G_OK;
}

extern "C" int k2g_pipeline_staging(k2g_pipeline *pl, char **seq, uint64_t **off, uint8_t **paired) {
  if (!pl) return fail(K2G_EUSAGE, "null pipeline");
  if (!pl->host_io) return fail(K2G_EUSAGE, "a device pipeline has no host staging");
  if (seq) *seq = pl->h_seq;
  if (off) *off = pl->h_off;
  if (paired) *paired = pl->h_paired;
  return K2G_OK;
}

extern "C" int k2g_pipeline_run_device(k2g_pipeline *pl, const char *d_seq, const uint64_t *d_off,
                                       const uint8_t *d_paired, uint64_t n, double *ms) {
  if (!pl) return fail(K2G_EUSAGE, "null pipeline");
  if (pl->host_io) return fail(K2G_EUSAGE, "a host pipeline runs with k2g_pipeline_run");
  if (n > pl->pp.max_fragments) return fail(K2G_EUSAGE, "batch has more fragments than the pipeline's capacity");
  if (n && (!d_seq || !d_off)) return fail(K2G_EUSAGE, "null fragment arrays");
  PipeDesc d{};
  d.seq = d_seq;
  d.off = d_off;
  d.paired = d_paired;
  d.n = n;
  d.bases = 0;
  d.paired_all = pl->params.paired != 0;
  *pl->h_desc = d;
  pl->last_n = 0;
  int rc = launch(pl);
  if (rc) return rc;
  pl->last_n = n;
  if (ms) *ms = pl->last_ms;
  return K2G_OK;
}

extern "C" int k2g_pipeline_digests(k2g_pipeline *pl, uint64_t *out, uint64_t n) {
  if (!pl || (!out && n)) return fail(K2G_EUSAGE, "null argument");
  if (n != pl->last_n) return fail(K2G_EUSAGE, "digest count differs from the last batch's fragments");
  if (!n) return K2G_OK;
  uint64_t *d = nullptr;
  HIP_TRY(hipMalloc((void **) &d, n * 8));
  const PipeBufs &b = pl->b;
  hipLaunchKernelGGL(k_digest, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, pl->stream, b.call, b.groups,
                     b.total, b.n_runs, b.run_off, b.runs, n, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, pl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
  (void) hipFree(d);
  if (e != hipSuccess) return fail(K2G_ESOFTWARE, std::string("pipeline digests: ") + hipGetErrorString(e));
  return K2G_OK;
}

extern "C" int k2g_pipeline_get_params(const k2g_pipeline *pl, k2g_pipeline_params *pp, int32_t *grid) {
  if (!pl || !pp) return fail(K2G_EUSAGE, "null argument");
  *pp = pl->pp;
  if (grid) *grid = pl->grid;
  return K2G_OK;
}

extern "C" int k2g_pipeline_redo_counts(const k2g_pipeline *pl, uint64_t *space, uint64_t *dense) {
  if (!pl) return fail(K2G_EUSAGE, "null pipeline");
  if (space) *space = pl->last_space;
  if (dense) *dense = pl->last_dense;
  return K2G_OK;
}

extern "C" int k2g_pipeline_split_counts(const k2g_pipeline *pl, uint64_t *fragments, uint64_t *segments) {
  if (!pl) return fail(K2G_EUSAGE, "null pipeline");
  if (fragments) *fragments = pl->last_split;
  if (segments) *segments = pl->last_segments;
  return K2G_OK;
}

// ---- autotuner
extern "C" int k2g_pipeline_tune(k2g_db *db, const k2g_classify_params *p, const char *d_seq, const uint64_t *d_off,
                                 const uint8_t *d_paired, uint64_t n, uint64_t stride, k2g_pipeline_params *best,
                                 k2g_tune_report *report) {
  if (!db || !p || !best || !report) return fail(K2G_EUSAGE, "null argument");
  if (n == 0 || !d_seq || !d_off) return fail(K2G_EUSAGE, "no fragments to tune on");
  if (stride == 0) stride = 1;
  std::memset(report, 0, sizeof(*report));
  const uint64_t m = (n + stride - 1) / stride;
  // the sample: every stride-th fragment, copied contiguously
  uint64_t *d_len = nullptr, *d_soff = nullptr;
  char *d_sseq = nullptr;
  uint8_t *d_spaired = nullptr;
  k2g_pipeline *pl = nullptr;
  int rc = K2G_OK;
  std::vector<uint64_t> len(2 * m + 1), soff(2 * m + 1);
  auto cleanup = [&] {
    if (pl) k2g_pipeline_destroy(pl);
    for (void *x : {(void *) d_len, (void *) d_soff, (void *) d_sseq, (void *) d_spaired})
      if (x) (void) hipFree(x);
  };
  hipError_t e = hipMalloc((void **) &d_len, (2 * m + 1) * 8);
  if (e == hipSuccess) e = hipMalloc((void **) &d_soff, (2 * m + 1) * 8);
  if (e == hipSuccess && d_paired) e = hipMalloc((void **) &d_spaired, m);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sample_lengths, dim3((unsigned) ((m + 1 + 255) / 256)), dim3(256), 0, 0, d_off, m, stride,
                       d_len);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(len.data(), d_len, (2 * m + 1) * 8, hipMemcpyDeviceToHost);
  uint64_t bases = 0;
  for (uint64_t i = 0; i < 2 * m; i++) {
    soff[i] = bases;
    bases += len[i];
  }
  soff[2 * m] = bases;
  if (e == hipSuccess) e = hipMemcpy(d_soff, soff.data(), (2 * m + 1) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void **) &d_sseq, bases + 16);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sample_copy, dim3((unsigned) std::min<uint64_t>(m, 65536)), dim3(64), 0, 0, d_seq, d_off,
                       d_paired, m, stride, d_soff, d_sseq, d_spaired);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    cleanup();
    return fail(K2G_ESOFTWARE, std::string("tuning sample: ") + hipGetErrorString(e));
  }
  k2g_pipeline_params pp = *best;
  pp.max_fragments = m;
  pp.max_bases = bases;
  pp.host_io = 0;
  k2g_classify_params exact = *p;
  exact.run_cap_pct = 100;  // no fragment can outgrow its hitlist space
  rc = k2g_pipeline_create(db, &exact, &pp, &pl);
  if (rc) {
    cleanup();
    return rc;
  }
  std::vector<uint64_t> ref(m), dig(m);
  // one configuration: two warm runs, the best of three timed; results must
  // equal the first configuration's
  auto eval = [&](double *ms) -> int {
    int r = recapture(pl);
    if (r) return r;
    double t = 1e30, x = 0;
    for (int i = 0; i < 5; i++) {
      r = k2g_pipeline_run_device(pl, d_sseq, d_soff, d_spaired, m, &x);
      if (r) return r;
      if (i >= 2) t = std::min(t, x);
    }
    r = k2g_pipeline_digests(pl, dig.data(), m);
    if (r) return r;
    if (report->evaluated == 0) ref = dig;
    else if (dig != ref) report->mismatched++;
    report->evaluated++;
    *ms = t;
    return K2G_OK;
  };
  double cur = 0;
  rc = eval(&cur);
  report->default_ms = cur;
  const int32_t units[] = {1024, 2048, 4096, 8192, 16384, 32768};
  const int32_t shifts[] = {0, 1, 2};
  const int32_t depths[] = {2, 4, 6, 8};
  int per_cu = 0;
  if (rc == K2G_OK && hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, kPipeVariants[variant_of(db, *p)].classify, 64, 0) != hipSuccess)
    rc = fail(K2G_ESOFTWARE, "occupancy query failed");
  const int32_t waves[] = {std::max(1, per_cu / 2), std::max(1, per_cu * 3 / 4), 0};
  // coordinate descent, two sweeps over the four parameters
  for (int sweep = 0; sweep < 2 && rc == K2G_OK; sweep++) {
    for (int which = 0; which < 4 && rc == K2G_OK; which++) {
      const int32_t *vals = which == 0 ? units : which == 1 ? shifts : which == 2 ? depths : waves;
      const int count = which == 0 ? 6 : which == 1 ? 3 : which == 2 ? 4 : 3;
      int32_t *field = which == 0 ? &pl->pp.unit_elements
                                  : which == 1 ? &pl->pp.bucket_shift
                                               : which == 2 ? &pl->pp.probe_depth : &pl->pp.waves_per_cu;
      int32_t keep = *field;
      for (int i = 0; i < count && rc == K2G_OK; i++) {
        if (vals[i] == keep) continue;
        *field = vals[i];
        double ms = 0;
        rc = eval(&ms);
        if (rc == K2G_OK && ms < cur) {
          cur = ms;
          keep = vals[i];
        }
      }
      *field = keep;
    }
  }
  if (rc == K2G_OK) {
    report->best_ms = cur;
    report->sample_fragments = m;
    report->sample_bases = bases;
    best->unit_elements = pl->pp.unit_elements;
    best->bucket_shift = pl->pp.bucket_shift;
    best->probe_depth = pl->pp.probe_depth;
    best->waves_per_cu = pl->pp.waves_per_cu;
  }
  cleanup();
  return rc;
}
