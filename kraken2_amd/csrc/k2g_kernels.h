// Kernel-facing batch descriptors shared by the HIP kernels and the host
// driver of the classify core.
#pragma once
#include <stdint.h>

namespace k2g {

struct ClassifyParams {
  double confidence;       // -T
  int32_t min_hit_groups;  // -g
  uint32_t quick;          // -q
  uint32_t tab_cap;        // LDS hit-table slots per fragment (power of 2, <= 256)
};

// Fragment i: mate 1 = seq[off[2i], off[2i+1]), mate 2 = seq[off[2i+1], off[2i+2]).
struct FragBatch {
  const char *seq;
  const uint64_t *off;
  const uint8_t *paired;  // per-fragment flag, or null: use paired_all
  uint32_t paired_all;
  uint64_t n;
  const uint32_t *order;  // processing order for the persistent kernel (null = 0..n-1)
};

struct FragOut {
  uint32_t *call;          // internal taxid
  uint32_t *groups;        // minimizer hit groups
  uint32_t *total;         // total_kmers denominator (src/classify.cc:890-896)
  uint32_t *n_runs;        // hitlist runs written
  uint32_t *status;        // 0 ok, 1 hit-table overflow (redo with the dense table), 2 run bound
  uint32_t *n_recs;        // HLL records written
  const uint64_t *run_off; // n+1 bounds: runs/recs of fragment i live at [run_off[i], run_off[i+1])
  uint32_t *runs;          // (code, count) pairs
  uint32_t *recs;          // (taxon, sparse-encoded HLL hash) pairs, or null
  unsigned long long *kmer_counts;  // per internal taxid, or null
};

// Per-slot scratch of the overflow path: a dense count array per taxon.
struct DenseScratch {
  uint32_t *counts;  // slots * node_count
  uint32_t *lists;   // slots * node_count
  uint32_t *lens;    // slots
  uint32_t slots;
};

}  // namespace k2g
