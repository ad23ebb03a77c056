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
  uint32_t batch_chunks;   // probe depth: chunks whose lookups are queued before one probe
                           // round, in [1, 8] (8 unless tuned)
};

// One batch of a graph-replayed pipeline (k2g_pipeline): the graph's kernels
// read the batch's arrays and sizes from this device-resident descriptor, so
// one instantiated graph serves every batch that fits its capacity.
struct PipeDesc {
  const char *seq;
  const uint64_t *off;
  const uint8_t *paired;  // per-fragment flags, or null: paired_all
  uint64_t n, bases;
  uint32_t paired_all, pad;
};

// Fragment i: mate 1 = seq[off[2i], off[2i+1]), mate 2 = seq[off[2i+1], off[2i+2]).
struct FragBatch {
  const char *seq;
  const uint64_t *off;
  const uint8_t *paired;  // per-fragment flag, or null: use paired_all
  uint32_t paired_all;
  uint64_t n;
  const uint32_t *order;  // processing order for the persistent kernel (null = 0..n-1)
  // Work units of the persistent kernel: unit u is order[unit_start[u] ..
  // unit_start[u+1]), *n_units units in all (device memory).  Null: one
  // fragment per unit.
  const uint32_t *unit_start;
  const uint32_t *n_units;
  // Per fragment, nonzero when the fragment is classified by segments
  // (SegBatch) instead: the persistent kernel skips it.  Null: none is.
  const uint32_t *split;
};

// Long nucleotide fragments split over waves.  Segment s classifies the
// k-mers ending at positions [a, b) of one mate of fragment frag: it scans
// that mate from k bases before a (the k-mer ending at a - 1, free of
// ambiguous bases by the choice of a, supplies the previous minimizer and
// the taxon a repeat of it takes) and writes its runs and HLL records to its
// own space in the scratch.  A stitch pass then joins a fragment's segments
// in order (runs that meet at a seam with one code merge; the mate marker
// sits between the mates), rebuilds the hit table from the joined runs and
// resolves the call.  The first segment of a mate starts at a = k - 1; a
// mate shorter than k has one empty segment.
struct SegBatch {
  uint32_t *frag, *mate, *a, *b;              // per segment
  uint64_t *off, *cap;                        // per segment: its run / record space in the scratch
  uint32_t *n_runs, *nrec, *groups, *status;  // per segment: what it produced
  uint32_t *runs, *recs;                      // the scratch
  const uint32_t *count;                      // [0] segments, [1] split fragments
  const uint32_t *split_list;                 // the split fragments
  const uint32_t *seg_cnt;                    // per fragment: its segments (0: not split)
  const uint64_t *seg_start;                  // per fragment: its first segment
};

struct FragOut {
  uint32_t *call;          // internal taxid
  uint32_t *groups;        // minimizer hit groups
  uint32_t *total;         // total_kmers denominator (src/classify.cc:890-896)
  uint32_t *n_runs;        // hitlist runs written
  uint32_t *status;        // bit 0: hit-table overflow (redo with the dense table);
                           // bit 1: more runs or records than run_cap (redo with the exact bound)
  uint32_t *n_recs;        // HLL records written
  const uint64_t *run_off; // runs/recs of fragment i start at run_off[i] ...
  const uint64_t *run_cap; // ... and hold at most run_cap[i] entries
  uint32_t *runs;          // (code, count) pairs
  uint32_t *recs;          // (taxon, sparse-encoded HLL hash) pairs, or null
  unsigned long long *kmer_counts;  // per internal taxid, or null
  uint32_t *lookups;       // per fragment: table probes issued, or null
  unsigned long long *lines;  // per fragment: distinct 128-byte lines the probes touched
};

// Exact element count of a fragment: one per k-mer end of each scanned
// stream plus border markers (src/classify.cc:797-838); a fragment's
// hitlist runs and HLL records never outnumber it.  L2 is ignored unless
// paired.  Translated databases scan six frames of each mate
// (src/aa_translate.cc:41-100) with five frame markers between them.
__host__ __device__ inline uint64_t fragment_elements(uint64_t L1, uint64_t L2, bool paired,
                                                      uint32_t k, bool dna) {
  uint64_t e = 0;
  for (int m = 0; m < (paired ? 2 : 1); m++) {
    const uint64_t L = m ? L2 : L1;
    if (dna) {
      e += L + 1 > k ? L + 1 - k : 0;
    } else {
      for (int fr = 0; fr < 6; fr++) {
        uint64_t fl;
        if (L < 3) {
          fl = 1;
        } else {
          const uint64_t first = (fr % 3) == 0 ? 3 : ((fr % 3) == 1 ? 4 : 2);
          fl = (L - 1 >= first) ? (L - 1 - first) / 3 + 1 : 0;
        }
        e += fl + 1 > k ? fl + 1 - k : 0;
      }
      e += 5;
    }
  }
  if (paired) e += 1;
  return e;
}

// Hitlist space reserved for a fragment of e elements: min(e, e * pct / 100 + 64).
__host__ __device__ inline uint64_t run_space(uint64_t e, uint32_t pct) {
  const uint64_t c = e * pct / 100 + 64;
  return c < e ? c : e;
}

// The persistent kernel's cursor block: kCursors cursors 128 bytes apart.
constexpr int kCursors = 8;
constexpr int kCursorStride = 16;
constexpr int kCursorWords = kCursors * kCursorStride;

// Per-slot scratch of the overflow path: a dense count array per taxon.
struct DenseScratch {
  uint32_t *counts;  // slots * node_count
  uint32_t *lists;   // slots * node_count
  uint32_t *lens;    // slots
  uint32_t slots;
};

}  // namespace k2g

namespace k2g {

// ---- build side (src/build_db.h:186-326, src/compact_hash.h:416-462)
struct BuildView {
  uint32_t *cells;
  uint64_t capacity;
  uint64_t barrett;
  uint32_t key_bits, value_bits;
  const uint32_t *parent;
  uint32_t k, l, dna, revcom_version;
  uint64_t spaced_mask, toggle_mask, min_clear;
  unsigned long long *size;  // cells filled
  unsigned int *error;       // 1: capacity exceeded
  // capacity estimation mode (cells == null): minimizers whose fmix64 has
  // its low 10 bits below sample_ranges are appended to sample[]
  // (src/estimate_capacity.cc:185-203)
  uint64_t *sample;
  unsigned long long *n_sample;
  uint64_t sample_cap;
  uint32_t sample_ranges;
  // verification mode (cells == null, sample == null): every (minimizer,
  // taxid) pair the build would insert whose fmix64 has its top part_bits
  // bits equal to part is appended to emit_key / emit_tax (at most
  // emit_cap; *n_emit counts them all)
  uint64_t *emit_key;
  uint32_t *emit_tax;
  unsigned long long *n_emit;
  uint64_t emit_cap;
  uint32_t part_bits, part;
};

// Counters of the table verification (k_verify_keys, k_verify_cells).
enum VerifySlot {
  kVOcc,          // (minimizer, taxid) pairs inserted
  kVKeys,         // distinct minimizers
  kVCells,        // distinct cells their lookups end at
  kVMissing,      // minimizers whose lookup finds no cell
  kVNotCovering,  // minimizers whose value is not an ancestor-or-self of each of their taxa
  kVSharedCells,  // cells reached by more than one minimizer (same compacted key)
  kVSharedKeys,   // minimizers in such cells
  kVDiffer,       // minimizers whose value is not the LCA of their own taxa
  kVUnexplained,  // cells whose value is not the LCA of the taxa of every minimizer ending there
  kVSlots
};

// One scanner window of the deterministic build: sequence bytes
// [start, start + len) scanned with a fresh scanner.  first_e is the first
// window-relative k-mer end the reference's scanner reports: k-1 at the
// start of a sequence, k-2 at any later subblock start (src/mmscanner.cc:
// 114-189 returns once str_pos_ >= k, which a start > 0 already satisfies).
struct SubBlock {
  uint64_t start;
  uint32_t len;
  uint32_t taxid;
  uint32_t first_e;
  uint32_t pad;
};

}  // namespace k2g
