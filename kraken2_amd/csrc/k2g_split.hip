// Long nucleotide fragments split over waves (persistent mode).
//
// One wave scans a fragment from its first base to its last, so a batch
// whose longest read outweighs a wave's even share of the batch waits on that
// read: a 50 kb read is about 4 ms of one wave while the rest of the GPU idles.
// Such batches (small service requests with long reads among them) split
// every fragment of more than 2 * kSegElements k-mers into segments of about
// kSegElements k-mers, classified by as many waves (SegBatch):
//   1. k_seg_count: the segments of each long fragment (per mate), listed;
//   2. a scan places each fragment's segments;
//   3. k_seg_plan: each mate's cut points, evenly spaced and moved forward to
//      the first position whose preceding k bases hold no ambiguous base (the
//      k-mer ending there is then a valid previous minimizer for the seam);
//      each segment's run space in the scratch;
//   4. a scan places the segments' run spaces;
//   5. the segment kernel, the persistent kernel (which skips split
//      fragments), then the stitch kernel.
// Results equal the unsplit scan's: a segment re-scans from k bases before its
// first k-mer, so its minimizers, ambiguity, LOOKUP/REPEAT decisions
// (src/classify.cc:797-838) and the taxon a leading repeat takes are the
// ones the whole scan has there.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "k2g.h"
#include "k2g_internal.h"

namespace k2g {

__global__ void k_classify_segments_dna(DbView, ClassifyParams, FragBatch, FragOut, SegBatch);
__global__ void k_classify_segments_dna_x(DbView, ClassifyParams, FragBatch, FragOut, SegBatch);
__global__ void k_stitch_dna(DbView, ClassifyParams, FragBatch, FragOut, SegBatch);
__global__ void k_stitch_dna_x(DbView, ClassifyParams, FragBatch, FragOut, SegBatch);
__global__ void k_widen(const uint32_t *in, uint64_t *out, uint64_t n);
__global__ void k_seg_count(const uint64_t *off, const uint8_t *paired, uint32_t paired_all, uint64_t n, uint32_t k,
                            uint32_t seg, const uint64_t *elements, uint32_t *cnt, uint32_t *list, uint32_t *count);
__global__ void k_seg_plan(const char *seq, const uint64_t *off, const uint8_t *paired, uint32_t paired_all,
                           uint32_t k, uint32_t seg, uint32_t pct, const uint32_t *list, const uint64_t *seg_start,
                           uint32_t *count, uint64_t seg_total_at, SegBatch sg);

bool split_wanted(const k2g_classify_params *p, bool dna, bool persistent, uint64_t max_e, uint64_t sum_e,
                  uint64_t waves) {
  return dna && persistent && !p->quick && !p->want_probe_stats && max_e > 2ull * kSegElements &&
         max_e * std::max<uint64_t>(waves, 1) > sum_e;
}

// Segments of the split fragments (each mate at most ceil(k-mers / kSegElements)).
uint64_t split_max_segments(uint64_t split_sum, uint64_t split_cnt) {
  return split_sum / kSegElements + 2 * split_cnt + 2;
}

uint64_t split_plan_bytes(uint64_t n, uint64_t split_sum, uint64_t split_cnt, bool hll) {
  const uint64_t segs = split_max_segments(split_sum, split_cnt), scratch = split_sum;
  uint64_t b = region_bytes(n + 1, 4) + region_bytes(n, 4) + region_bytes(4, 4);  // counts, list, totals
  b += 2 * region_bytes(n + 1, 8);                                                // widened counts, starts
  b += 4 * region_bytes(segs, 4) + 2 * region_bytes(segs + 1, 8);                 // frag, mate, a, b; cap, off
  b += 4 * region_bytes(segs, 4);                                                 // outputs
  b += region_bytes(2 * std::max<uint64_t>(scratch, 1), 4) * (hll ? 2 : 1);      // scratch runs (+ records)
  size_t t = 0;
  uint64_t *u = nullptr;
  (void) hipcub::DeviceScan::ExclusiveSum(nullptr, t, u, u, (int) (std::max(n, segs) + 1), (hipStream_t) 0);
  b += region_bytes(t, 1);
  return b;
}

int split_launch(k2g_batch *b, const k2g_classify_params *p, const ClassifyParams &cp, FragBatch &in,
                 const FragOut &out, uint64_t split_sum, uint64_t split_cnt, int phase) {
  const DbView &v = b->db->view;
  hipStream_t s = b->stream;
  const uint64_t n = in.n;
  const bool x = p->want_hll || p->want_kmer_counts;
  if (phase == 1) {
    b->sp_active = false;
    // segments get exact space (a segment never runs out of it), so the stitch
    // sees every run of a fragment and its hit table decides the dense redo
    const uint64_t segs = split_max_segments(split_sum, split_cnt), scratch = split_sum;
    if (split_cnt == 0 || segs >= (1ull << 31) || scratch >= (1ull << 62)) return K2G_OK;  // nothing to split
    HIP_TRY(b->sp_cnt.reserve(n + 1));
    HIP_TRY(b->sp_list.reserve(n));
    HIP_TRY(b->sp_count.reserve(4));
    HIP_TRY(b->sp_wide.reserve(n + 1));
    HIP_TRY(b->sp_start.reserve(n + 1));
    HIP_TRY(b->sp_frag.reserve(segs));
    HIP_TRY(b->sp_mate.reserve(segs));
    HIP_TRY(b->sp_a.reserve(segs));
    HIP_TRY(b->sp_b.reserve(segs));
    HIP_TRY(b->sp_cap.reserve(segs + 1));
    HIP_TRY(b->sp_off.reserve(segs + 1));
    HIP_TRY(b->sp_runs_n.reserve(segs));
    HIP_TRY(b->sp_nrec.reserve(segs));
    HIP_TRY(b->sp_groups.reserve(segs));
    HIP_TRY(b->sp_status.reserve(segs));
    HIP_TRY(b->sp_runs.reserve(2 * std::max<uint64_t>(scratch, 1)));
    if (p->want_hll) HIP_TRY(b->sp_recs.reserve(2 * std::max<uint64_t>(scratch, 1)));
    size_t tb = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, b->sp_wide.p, b->sp_start.p, (int) (std::max(n, segs) + 1),
                                             s));
    HIP_TRY(b->temp.reserve(tb));
    HIP_TRY(hipMemsetAsync(b->sp_count.p, 0, 16, s));
    HIP_TRY(hipMemsetAsync(b->sp_cap.p, 0, (segs + 1) * 8, s));
    // 1-2. segment counts and their starts
    const unsigned g1 = (unsigned) ((n + 1 + 255) / 256);
    hipLaunchKernelGGL(k_seg_count, dim3(g1), dim3(256), 0, s, in.off, in.paired, in.paired_all, n, v.k,
                       (uint32_t) kSegElements, (const uint64_t *) b->bounds.p, b->sp_cnt.p, b->sp_list.p,
                       b->sp_count.p);
    hipLaunchKernelGGL(k_widen, dim3(g1), dim3(256), 0, s, (const uint32_t *) b->sp_cnt.p, b->sp_wide.p, n);
    HIP_TRY(hipGetLastError());
    tb = b->temp.cap;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tb, b->sp_wide.p, b->sp_start.p, (int) (n + 1), s));
    // 3-4. cut points and run spaces, then the spaces' offsets
    SegBatch sg{b->sp_frag.p,    b->sp_mate.p,   b->sp_a.p,      b->sp_b.p,      b->sp_off.p,
                b->sp_cap.p,     b->sp_runs_n.p, b->sp_nrec.p,   b->sp_groups.p, b->sp_status.p,
                b->sp_runs.p,    p->want_hll ? b->sp_recs.p : nullptr,
                b->sp_count.p,   b->sp_list.p,   b->sp_cnt.p,    b->sp_start.p};
    const unsigned gp = (unsigned) std::min<uint64_t>(split_cnt, 4096);
    hipLaunchKernelGGL(k_seg_plan, dim3(gp), dim3(64), 0, s, in.seq, in.off, in.paired, in.paired_all, v.k,
                       (uint32_t) kSegElements, 100u, (const uint32_t *) b->sp_list.p,
                       (const uint64_t *) b->sp_start.p, b->sp_count.p, n, sg);
    HIP_TRY(hipGetLastError());
    tb = b->temp.cap;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tb, b->sp_cap.p, b->sp_off.p, (int) (segs + 1), s));
    // 5a. the segments (the persistent kernel that follows skips their fragments)
    const unsigned gs = (unsigned) std::min<uint64_t>(segs, (uint64_t) std::max(b->persistent_blocks[x ? 1 : 0], 8));
    hipLaunchKernelGGL(x ? k_classify_segments_dna_x : k_classify_segments_dna, dim3(gs), dim3(64), 0, s, v, cp, in,
                       out, sg);
    HIP_TRY(hipGetLastError());
    b->sp_segs = segs;
    b->sp_active = true;
    in.split = b->sp_cnt.p;
    return K2G_OK;
  }
  // 5b. phase 2: the split fragments joined and resolved
  if (!b->sp_active) return K2G_OK;
  SegBatch sg{b->sp_frag.p,    b->sp_mate.p,   b->sp_a.p,      b->sp_b.p,      b->sp_off.p,
              b->sp_cap.p,     b->sp_runs_n.p, b->sp_nrec.p,   b->sp_groups.p, b->sp_status.p,
              b->sp_runs.p,    p->want_hll ? b->sp_recs.p : nullptr,
              b->sp_count.p,   b->sp_list.p,   b->sp_cnt.p,    b->sp_start.p};
  const unsigned gt = (unsigned) std::min<uint64_t>(std::max<uint64_t>(b->sp_segs / 2, 1), 16384);
  hipLaunchKernelGGL(x ? k_stitch_dna_x : k_stitch_dna, dim3(gt), dim3(64), 0, s, v, cp, in, out, sg);
  HIP_TRY(hipGetLastError());
  b->split_last = true;
  return K2G_OK;
}

}  // namespace k2g
