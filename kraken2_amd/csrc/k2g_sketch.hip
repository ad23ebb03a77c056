// Per-(input block, taxon) distinct sets of a classify call's HLL records,
// reduced on the device (want_hll == 2).
//
// The reference keeps one HyperLogLogPlusMinus per taxon and input block
// (src/classify.cc:554, 640-645), then merges the block sketches into the
// run's totals (src/hyperloglogplus.cc:572-651).  A block sketch ends in a
// state fixed by the set of sparse-encoded hashes it saw and by whether it
// switched to registers: it switches when its sparse list already holds 1024
// distinct values and another insert arrives (src/hyperloglogplus.cc:
// 471-509), i.e. iff distinct(e_1 .. e_{n-1}) >= 1024 over its records e_1 ..
// e_n in insertion order.  So a block's sketch of a taxon is the distinct set
// of its records plus that flag, and the records themselves need not leave
// the GPU:
//   1. each record becomes a 64-bit key block << (taxon bits + 32) |
//      taxon << 32 | encoded hash, in record order, paired with its record
//      index;
//   2. one radix sort by key groups the (block, taxon) sketches, hashes
//      ascending;
//   3. a run-length pass gives the groups and their record counts, a scan
//      their starts;
//   4. one wave per group counts its distinct hashes d: d >= 1025 switches,
//      d <= 1023 does not, and at d == 1024 the switch happened iff the
//      group's last record in insertion order (the largest record index)
//      repeats an earlier hash;
//   5. a scan of d places each group's distinct hashes, and one wave per
//      group writes them, ascending, with the group's block, taxon, flag and
//      record count, straight into pinned host memory.
// The host folds the groups exactly as it folds per-hit records.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "k2g.h"
#include "k2g_internal.h"

namespace k2g {

void input_blocks(const k2g_fragments *in, bool paired_all, std::vector<uint64_t> &starts) {
  starts.clear();
  const uint64_t n = in->n;
  auto paired = [&](uint64_t i) { return in->paired ? in->paired[i] != 0 : paired_all; };
  for (uint64_t i = 0; i < n;) {
    starts.push_back(i);
    const uint64_t start = i;
    if (paired(i)) {
      while (i < n && paired(i) && i - start < 10000) i++;
    } else {
      uint64_t bases = 0;
      while (i < n && !paired(i) && bases < 3 * 1024 * 1024) {
        bases += in->off[2 * i + 1] - in->off[2 * i];
        i++;
      }
    }
  }
  starts.push_back(n);
}

namespace {

constexpr int kW = 64;

inline uint32_t bit_width(uint64_t x) { return x ? 64u - (uint32_t) __builtin_clzll(x) : 0u; }

struct SketchBits {
  uint32_t tbits, bbits;
};

SketchBits sketch_bits(uint64_t node_count, uint64_t n_blocks) {
  return {std::max(1u, bit_width(node_count ? node_count - 1 : 0)),
          std::max(1u, bit_width(n_blocks ? n_blocks - 1 : 0))};
}

struct HighWord {
  __host__ __device__ uint32_t operator()(uint64_t k) const { return (uint32_t) (k >> 32); }
};
using GroupIter = hipcub::TransformInputIterator<uint32_t, HighWord, const uint64_t *>;

// Step 1: keys and record indices, one wave per fragment.
__global__ __launch_bounds__(kW) void k_sketch_keys(const uint64_t *run_off, const uint32_t *n_recs,
                                                    const uint64_t *rec_off, const uint32_t *recs,
                                                    const uint64_t *block_start, uint32_t n_blocks,
                                                    uint32_t tbits, uint64_t n, uint64_t *keys, uint32_t *idx) {
  for (uint64_t f = blockIdx.x; f < n; f += gridDim.x) {
    uint32_t lo = 0, hi = n_blocks;  // block_start[0] = 0 <= f < block_start[n_blocks] = n
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (block_start[mid] <= f) lo = mid; else hi = mid;
    }
    const uint64_t top = (uint64_t) lo << (tbits + 32);
    const uint64_t s = run_off[f], o = rec_off[f];
    const uint32_t c = n_recs[f];
    for (uint32_t j = threadIdx.x; j < c; j += kW) {
      keys[o + j] = top | ((uint64_t) recs[2 * (s + j)] << 32) | recs[2 * (s + j) + 1];
      idx[o + j] = (uint32_t) (o + j);
    }
  }
}

// counts (u32, *count of them) widened into v; v[*count] = 0.
__global__ __launch_bounds__(256) void k_sketch_widen(const uint32_t *counts, const uint32_t *count, uint64_t cap,
                                                      uint64_t *v) {
  const uint32_t g_n = *count;
  for (uint64_t g = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; g <= g_n && g <= cap;
       g += (uint64_t) gridDim.x * blockDim.x)
    v[g] = g < g_n ? counts[g] : 0;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

// Step 4: distinct hashes and the register switch of each group.
__global__ __launch_bounds__(kW) void k_sketch_groups(const uint64_t *keys, const uint32_t *idx,
                                                      const uint32_t *n_of, const uint64_t *start,
                                                      const uint32_t *count, uint64_t *distinct,
                                                      uint32_t *dense) {
  const uint32_t g_n = *count;
  const int lane = threadIdx.x;
  for (uint32_t g = blockIdx.x; g < g_n; g += gridDim.x) {
    const uint64_t st = start[g], n = n_of[g];
    uint64_t d = 0;
    for (uint64_t base = 0; base < n; base += kW) {
      const uint64_t j = base + lane;
      const bool nw = j < n && (j == 0 || keys[st + j] != keys[st + j - 1]);
      d += (uint64_t) __popcll(__ballot(nw));
    }
    uint32_t sw = d >= 1025 ? 1u : 0u;
    if (d == 1024) {
      // the group's last record in insertion order, and how often its hash occurs
      uint32_t best = 0;
      uint64_t best_key = 0;
      bool any = false;
      for (uint64_t j = lane; j < n; j += kW) {
        const uint32_t r = idx[st + j];
        if (!any || r > best) {
          best = r;
          best_key = keys[st + j];
          any = true;
        }
      }
      for (int dl = 1; dl < kW; dl <<= 1) {
        const uint32_t ob = (uint32_t) __shfl_xor((int) best, dl);
        const uint64_t ok = (uint64_t) __shfl_xor((unsigned long long) best_key, dl);
        const bool oany = __shfl_xor((int) any, dl) != 0;
        if (oany && (!any || ob > best)) {
          best = ob;
          best_key = ok;
          any = true;
        }
      }
      uint64_t occ = 0;
      for (uint64_t base = 0; base < n; base += kW) {
        const uint64_t j = base + lane;
        occ += (uint64_t) __popcll(__ballot(j < n && keys[st + j] == best_key));
      }
      sw = occ > 1 ? 1u : 0u;
    }
    if (lane == 0) {
      distinct[g] = d;
      dense[g] = sw;
    }
  }
  if (blockIdx.x == 0 && lane == 0) distinct[g_n] = 0;
}

struct SketchHost {
  uint32_t *block, *taxon, *dense, *vals;
  uint64_t *n, *off, *words;  // words: groups, distinct values
};

// Step 5: each group's distinct hashes, ascending, and its header into
// pinned host memory.
__global__ __launch_bounds__(kW) void k_sketch_out(const uint64_t *keys, const uint32_t *gkey, const uint32_t *n_of,
                                                   const uint64_t *start, const uint64_t *off,
                                                   const uint32_t *dense, const uint32_t *count, uint32_t tbits,
                                                   SketchHost h) {
  const uint32_t g_n = *count;
  const int lane = threadIdx.x;
  const uint32_t tmask = (uint32_t) ((1ull << tbits) - 1);
  for (uint32_t g = blockIdx.x; g < g_n; g += gridDim.x) {
    const uint64_t st = start[g], n = n_of[g], o = off[g];
    uint64_t rank = 0;
    for (uint64_t base = 0; base < n; base += kW) {
      const uint64_t j = base + lane;
      const bool nw = j < n && (j == 0 || keys[st + j] != keys[st + j - 1]);
      const uint64_t m = __ballot(nw);
      if (nw) h.vals[o + rank + (uint64_t) __popcll(m & lanes_below(lane))] = (uint32_t) keys[st + j];
      rank += (uint64_t) __popcll(m);
    }
    if (lane == 0) {
      h.block[g] = gkey[g] >> tbits;
      h.taxon[g] = gkey[g] & tmask;
      h.dense[g] = dense[g];
      h.n[g] = n;
      h.off[g] = o;
    }
  }
  if (blockIdx.x == 0 && lane == 0) {
    h.off[g_n] = off[g_n];
    h.words[0] = g_n;
    h.words[1] = off[g_n];
  }
}

template <class T>
T *device_view(T *host) {
  T *d = nullptr;
  return hipHostGetDevicePointer((void **) &d, host, 0) == hipSuccess ? d : nullptr;
}

// Groups a call can have at most: one per record, and one per (block, taxon).
uint64_t max_groups(uint64_t records, uint64_t n_blocks, uint64_t node_count) {
  return std::max<uint64_t>(1, std::min<uint64_t>(records, n_blocks * node_count));
}

// The hipcub scratch of the reduction over `records` records and `groups`
// possible groups.
size_t sketch_temp_bytes(uint64_t records, uint64_t groups, uint32_t key_bits) {
  size_t a = 0, b = 0, c = 0;
  uint64_t *k = nullptr;
  uint32_t *u = nullptr;
  hipcub::DoubleBuffer<uint64_t> kb(k, k);
  hipcub::DoubleBuffer<uint32_t> vb(u, u);
  (void) hipcub::DeviceRadixSort::SortPairs(nullptr, a, kb, vb, (int) records, 0, (int) key_bits, (hipStream_t) 0);
  (void) hipcub::DeviceRunLengthEncode::Encode(nullptr, b, GroupIter(k, HighWord()), u, u, u, (int) records,
                                               (hipStream_t) 0);
  (void) hipcub::DeviceScan::ExclusiveSum(nullptr, c, k, k, (int) (groups + 1), (hipStream_t) 0);
  return std::max(a, std::max(b, c));
}

}  // namespace

int sketch_check(const k2g_db *db, uint64_t n_blocks) {
  const SketchBits sb = sketch_bits(db->view.node_count, n_blocks);
  if (sb.tbits + sb.bbits > 32)
    return fail(K2G_EUSAGE, "per-block HLL sets need block and taxon ids within 32 bits; this call has " +
                                std::to_string(n_blocks) + " input blocks over " +
                                std::to_string(db->view.node_count) + " taxa (split it, or use want_hll 1)");
  return K2G_OK;
}

uint64_t sketch_plan_bytes(const k2g_db *db, uint64_t elements, uint64_t n_blocks) {
  const SketchBits sb = sketch_bits(db->view.node_count, n_blocks);
  const uint64_t groups = max_groups(elements, n_blocks, db->view.node_count);
  uint64_t b = region_bytes(n_blocks + 1, 8);                    // block starts
  b += region_bytes(elements, 8) + 2 * region_bytes(elements, 4);  // keys, record indices (two buffers)
  b += 3 * region_bytes(groups, 4) + region_bytes(4, 4);         // group keys, counts, flags; totals
  b += 3 * region_bytes(groups + 1, 8);                          // widened counts, starts, offsets
  b += region_bytes(sketch_temp_bytes(elements, groups, 32 + sb.tbits + sb.bbits), 1);
  return b;
}

int sketch_upload_blocks(k2g_batch *b, const std::vector<uint64_t> &starts) {
  const uint64_t nb = starts.size() - 1;
  int rc = sketch_check(b->db, nb);
  if (rc) return rc;
  HIP_TRY(b->sk_starts.reserve(nb + 1));
  HIP_TRY(hipMemcpyAsync(b->sk_starts.p, starts.data(), (nb + 1) * 8, hipMemcpyHostToDevice, b->stream));
  b->sk_blocks = nb;
  return K2G_OK;
}

int sketch_fetch(k2g_batch *b, HostResults &h, uint64_t records) {
  const uint64_t n = b->n, nb = b->sk_blocks;
  const DbView &v = b->db->view;
  hipStream_t s = b->stream;
  HIP_TRY(h.sk_words.reserve(2));
  h.sk_words.p[0] = h.sk_words.p[1] = 0;
  HIP_TRY(h.sk_off.reserve(1));
  h.sk_off.p[0] = 0;
  if (records == 0) return K2G_OK;
  if (records >= (1ull << 31))
    return fail(K2G_EUSAGE, "per-block HLL sets take fewer than 2^31 records per call (split the batch)");
  const SketchBits sb = sketch_bits(v.node_count, nb);
  const uint32_t key_bits = 32 + sb.tbits + sb.bbits;
  const uint64_t gmax = max_groups(records, nb, v.node_count);
  HIP_TRY(b->sk_keys.reserve(records));
  HIP_TRY(b->sk_idx.reserve(records));
  HIP_TRY(b->sk_idx_alt.reserve(records));
  HIP_TRY(b->sk_gkey.reserve(gmax));
  HIP_TRY(b->sk_n.reserve(gmax));
  HIP_TRY(b->sk_dense.reserve(gmax));
  HIP_TRY(b->sk_count.reserve(4));
  HIP_TRY(b->sk_wide.reserve(gmax + 1));
  HIP_TRY(b->sk_start.reserve(gmax + 1));
  HIP_TRY(b->sk_off.reserve(gmax + 1));
  HIP_TRY(b->temp.reserve(sketch_temp_bytes(records, gmax, key_bits)));
  HIP_TRY(h.sk_block.reserve(gmax));
  HIP_TRY(h.sk_taxon.reserve(gmax));
  HIP_TRY(h.sk_dense.reserve(gmax));
  HIP_TRY(h.sk_n.reserve(gmax));
  HIP_TRY(h.sk_off.reserve(gmax + 1));
  HIP_TRY(h.sk_vals.reserve(records));
  SketchHost out{device_view(h.sk_block.p), device_view(h.sk_taxon.p), device_view(h.sk_dense.p),
                 device_view(h.sk_vals.p),  device_view(h.sk_n.p),     device_view(h.sk_off.p),
                 device_view(h.sk_words.p)};
  if (!out.block || !out.taxon || !out.dense || !out.vals || !out.n || !out.off || !out.words)
    return fail(K2G_ESOFTWARE, "HLL sets: pinned buffers are not mapped");

  // 1. keys in record order (the records' space is the sort's second key buffer)
  const unsigned gf = (unsigned) std::min<uint64_t>(n, 65536);
  hipLaunchKernelGGL(k_sketch_keys, dim3(gf), dim3(kW), 0, s, b->run_off.p, b->n_recs.p, b->comp_rec_off.p,
                     b->recs.p, b->sk_starts.p, (uint32_t) nb, sb.tbits, n, b->sk_keys.p, b->sk_idx.p);
  HIP_TRY(hipGetLastError());
  // 2. sort by (block, taxon, hash); ties keep record order
  uint64_t *alt = reinterpret_cast<uint64_t *>(b->recs.p);  // 2 * space words >= 2 * records
  hipcub::DoubleBuffer<uint64_t> kb(b->sk_keys.p, alt);
  hipcub::DoubleBuffer<uint32_t> vb(b->sk_idx.p, b->sk_idx_alt.p);
  size_t tb = b->temp.cap;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(b->temp.p, tb, kb, vb, (int) records, 0, (int) key_bits, s));
  const uint64_t *keys = kb.Current();
  const uint32_t *idx = vb.Current();
  // 3. groups, their record counts and starts
  tb = b->temp.cap;
  HIP_TRY(hipcub::DeviceRunLengthEncode::Encode(b->temp.p, tb, GroupIter(keys, HighWord()), b->sk_gkey.p,
                                                b->sk_n.p, b->sk_count.p, (int) records, s));
  const unsigned gw = (unsigned) std::min<uint64_t>((gmax + 1 + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sketch_widen, dim3(gw), dim3(256), 0, s, b->sk_n.p, b->sk_count.p, gmax, b->sk_wide.p);
  HIP_TRY(hipGetLastError());
  tb = b->temp.cap;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tb, b->sk_wide.p, b->sk_start.p, (int) (gmax + 1), s));
  // 4. distinct hashes and register switches
  const unsigned gg = (unsigned) std::min<uint64_t>(gmax, 65536);
  hipLaunchKernelGGL(k_sketch_groups, dim3(gg), dim3(kW), 0, s, keys, idx, b->sk_n.p, b->sk_start.p,
                     b->sk_count.p, b->sk_wide.p, b->sk_dense.p);
  HIP_TRY(hipGetLastError());
  // 5. value offsets, then the sets and headers into pinned host memory
  tb = b->temp.cap;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tb, b->sk_wide.p, b->sk_off.p, (int) (gmax + 1), s));
  hipLaunchKernelGGL(k_sketch_out, dim3(gg), dim3(kW), 0, s, keys, b->sk_gkey.p, b->sk_n.p, b->sk_start.p,
                     b->sk_off.p, b->sk_dense.p, b->sk_count.p, sb.tbits, out);
  HIP_TRY(hipGetLastError());
  return K2G_OK;
}

}  // namespace k2g
