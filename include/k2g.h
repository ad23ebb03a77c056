/*
 * k2g.h -- C ABI of the MI355X Kraken 2 classify core (libk2g.so).
 *
 * Plain C types only; no HIP or torch types cross this boundary.  Every
 * function returns 0 on success or a sysexits-style status (64 usage, 65 data
 * error, 66 no input, 70 internal error, 71 OS error, 74 I/O error); the
 * message of the last failure on the calling thread is in k2g_last_error().
 * Nothing here calls exit().
 *
 * Threading: a k2g_db is read-only after k2g_db_open and may be shared by
 * threads (the reference's KeyValueStore is const and shared the same way,
 * src/classify.cc:500-587).  A k2g_batch owns one HIP stream and its buffers
 * and must be used by one thread at a time.
 */
#ifndef K2G_H_
#define K2G_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define K2G_OK 0
#define K2G_EUSAGE 64
#define K2G_EDATA 65
#define K2G_ENOINPUT 66
#define K2G_ESOFTWARE 70
#define K2G_EOS 71
#define K2G_EIO 74

/* Hitlist run codes above any internal taxid. */
#define K2G_CODE_MATE 0xFFFFFFFFu  /* "|:|" mate-pair border   (src/classify.cc:33) */
#define K2G_CODE_FRAME 0xFFFFFFFEu /* "-:-" reading-frame border (src/classify.cc:34) */
#define K2G_CODE_AMBIG 0xFFFFFFFDu /* "A:n" ambiguous span     (src/classify.cc:35) */

typedef struct k2g_db k2g_db;
typedef struct k2g_batch k2g_batch;

typedef struct {
  uint64_t capacity, size, key_bits, value_bits; /* hash.k2d header */
  uint64_t node_count;                           /* taxo.k2d */
  uint64_t k, l, spaced_seed_mask, toggle_mask, minimum_acceptable_hash_value;
  int32_t dna_db, revcom_version; /* opts.k2d */
  uint64_t device_bytes;          /* HBM held by the resident DB */
} k2g_db_info;

/* Replaces load_index (src/classify.cc:269-302): reads hash.k2d, taxo.k2d
 * and opts.k2d, uploads the cell array and the parent array to HBM and keeps
 * them resident.  Rejects 40-bit cells (src/kv_store.h:35-57), taxonomies
 * whose ids are not parent-before-child, and cells holding a taxid that is
 * not a node. */
int k2g_db_open(const char *hash_k2d, const char *taxo_k2d, const char *opts_k2d, k2g_db **out);
int k2g_db_close(k2g_db *db);
int k2g_db_get_info(const k2g_db *db, k2g_db_info *info);
/* External (NCBI) taxid of each internal id, node_count entries. */
int k2g_db_external_ids(const k2g_db *db, uint64_t *out, uint64_t n);
const char *k2g_last_error(void);

/* KeyValueStore::GetBatch (src/kv_store.h:26-33; src/compact_hash.h:354-392):
 * out[i] = value of keys[i] or 0.  Host arrays; synchronous. */
int k2g_get_batch(k2g_db *db, const uint64_t *keys, uint32_t *out, size_t n);
/* Same on device arrays, enqueued on a HIP stream (NULL = default stream). */
int k2g_get_batch_device(k2g_db *db, const uint64_t *d_keys, uint32_t *d_out, size_t n,
                         void *hip_stream);

/* ---- classify: the ProcessFiles -> ClassifySequence seam
 * (src/classify.cc:479-708 -> 773-948) */
typedef struct {
  double confidence;        /* -T, in [0, 1] */
  int32_t min_hit_groups;   /* -g */
  int32_t quick;            /* -q */
  int32_t paired;           /* -P / -S: every fragment is a pair (unless per-fragment flags) */
  int32_t mode;             /* 0: plain, one launch per batch; 1: persistent, longest first;
                             * 2: persistent when fragment lengths are uneven (the longest
                             * above four times the mean), else plain */
  int32_t want_kmer_counts; /* -R: per-taxon LOOKUP-hit counts */
  int32_t want_hll;         /* -K: 1 = per-hit (taxon, encoded HLL hash) records (hll_off, hll);
                             * 2 = the records reduced on the device to one distinct set per
                             * (input block, taxon) with the block sketch's register switch
                             * decided (sketch_*); blocks are the reference's input blocks over
                             * the call's fragments (src/classify.cc:523-543): runs of pairs
                             * 10 000 at a time, runs of single reads until 3 MiB of bases.
                             * 2 needs host fragments (k2g_classify, the server), fewer than
                             * 2^31 records per call, and block and taxon ids that fit 32 bits
                             * together */
  int32_t table_slots;      /* per-fragment LDS hit table (power of 2 in [2,256]; 0 = 256) */
  int32_t want_probe_stats; /* per-fragment probes and distinct 128-byte table lines touched */
  int32_t run_cap_pct;      /* hitlist space reserved per fragment, in percent of its element
                             * count (plus 64); fragments that need more are classified again
                             * into exact space.  0 = 25; 100 = always exact (never redone). */
  int32_t unit_elements;    /* persistent mode: fragments (longest first) are handed to wavefronts
                             * in units of about this many k-mer elements (fewer when a batch is
                             * too small to give every resident wave four units); 0 = 4096 */
} k2g_classify_params;

/* Fragment i: mate 1 = seq[off[2i], off[2i+1]), mate 2 = seq[off[2i+1],
 * off[2i+2]).  A single read's mate 2 range is not read: it may be empty or
 * hold other bytes (the text between records of a FASTA file read in place,
 * which the classify front-end passes as is).  paired may be NULL (then
 * params.paired applies to all) or give a 0/1 flag per fragment. */
typedef struct {
  const char *seq;
  const uint64_t *off; /* 2n+1 entries */
  const uint8_t *paired;
  uint64_t n;
} k2g_fragments;

/* Results of the last k2g_classify on a batch, in host memory owned by the
 * batch (valid until the next call on it).  Runs of fragment i are the
 * (code, count) pairs runs[2*run_off[i] .. 2*run_off[i+1]); codes are
 * internal taxids (0 = no hit) or K2G_CODE_*.  HLL records likewise at
 * hll_off.  kmer_counts has node_count entries. */
typedef struct {
  uint64_t n;
  const uint32_t *call;        /* internal taxid, 0 = unclassified */
  const uint32_t *hit_groups;  /* minimizer_hit_groups */
  const uint32_t *total_kmers; /* confidence denominator */
  const uint64_t *run_off;     /* n+1 */
  const uint32_t *runs;
  const uint64_t *hll_off;     /* n+1, or NULL */
  const uint32_t *hll;         /* (taxon, encoded) pairs, or NULL */
  const uint64_t *kmer_counts; /* or NULL */
  const uint32_t *lookups;     /* per fragment, with want_probe_stats, else NULL */
  const uint64_t *probe_lines; /* per fragment, with want_probe_stats, else NULL */
  uint64_t node_count;
  uint64_t redone;             /* fragments that took the dense-table path */
  double kernel_ms;            /* device time of the classify kernels */
  /* want_hll == 2: one group per (input block, taxon) with HLL records, ordered
   * by block, then taxon.  Group g's distinct encoded hashes are
   * sketch_vals[sketch_off[g] .. sketch_off[g+1]), ascending; sketch_dense[g]
   * says whether the block's sketch of the taxon switched to registers
   * (src/hyperloglogplus.cc:471-509: the sparse list already held 1024
   * distinct values when another insert arrived). */
  uint64_t n_sketches;
  const uint32_t *sketch_block;
  const uint32_t *sketch_taxon;
  const uint32_t *sketch_dense;
  const uint64_t *sketch_n;    /* records (LOOKUP hits) the group observed */
  const uint64_t *sketch_off;  /* n_sketches + 1 */
  const uint32_t *sketch_vals;
} k2g_results;

int k2g_batch_create(k2g_db *db, k2g_batch **out);
int k2g_batch_destroy(k2g_batch *b);

/* Host fragments -> host results (uploads, classifies, downloads; sync). */
int k2g_classify(k2g_batch *b, const k2g_classify_params *p, const k2g_fragments *in,
                 k2g_results *out);

/* Device-resident fragments (all pointers in HBM): classifies on the batch's
 * stream and leaves results in the batch's device buffers; call
 * k2g_batch_fetch to synchronise and copy them to host memory.  order may be
 * NULL; otherwise the persistent kernel visits fragments in that order. */
int k2g_classify_device(k2g_batch *b, const k2g_classify_params *p, const char *d_seq,
                        const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                        const uint32_t *d_order);
int k2g_batch_sync(k2g_batch *b, double *kernel_ms);
int k2g_batch_fetch(k2g_batch *b, k2g_results *out);
/* HIP stream of the batch (a hipStream_t), for callers that enqueue their
 * own work around k2g_classify_device. */
void *k2g_batch_stream(k2g_batch *b);
/* Per-fragment scalars of the last classify call, without hitlists: each
 * non-NULL array receives n entries (n must equal the call's fragment
 * count).  lookups / probe_lines need want_probe_stats on that call.  For
 * callers that need calls and counters only (a report, a digest pass). */
int k2g_batch_fetch_counts(k2g_batch *b, uint32_t *call, uint32_t *hit_groups,
                           uint32_t *total_kmers, uint32_t *n_runs, uint32_t *lookups,
                           uint64_t *probe_lines, uint64_t n);
/* Fragments of the last classify call that were classified a second time:
 * into exact hitlist space (run_cap_pct overflow) and with the dense
 * per-taxon table (more distinct hit taxa than table_slots). */
int k2g_batch_redo_counts(k2g_batch *b, uint64_t *run_space, uint64_t *dense_table);
/* Long fragments of the last call that were split over waves, and their
 * segments (both 0 when the call did not split).  Persistent mode splits the
 * fragments of more than 4096 k-mers into segments of about 2048 when the
 * batch's longest fragment outweighs an even share of the batch per
 * resident wave (nucleotide databases; not with quick mode or probe
 * statistics).  No reference counterpart: results are the same either way. */
int k2g_batch_split_counts(k2g_batch *b, uint64_t *fragments, uint64_t *segments);
/* One 64-bit digest per fragment of the last classify call (call, hit
 * groups, total k-mers and every hitlist run, in order), copied to
 * out[0..n).  Comparing digests checks two runs over the same fragments for
 * equality without moving hitlists to the host.  No reference counterpart:
 * a verification aid for large batches. */
int k2g_batch_digests(k2g_batch *b, uint64_t *out, uint64_t n);

/* ---- classify without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same launch sequence and results as k2g_classify_device with
 * run_cap_pct = 100, but every scratch buffer (element bounds, status words,
 * the longest-first sort and its work units, redo lists, cursors, hipCUB
 * temporary storage, the dense hit tables of the overflow path) is carved
 * from a 128 MiB statically allocated device array per GPU, held until the
 * process ends; at most 524288 fragments per call (larger calls fail with
 * K2G_EUSAGE).  Results go to the caller's device arrays: fragment i's runs
 * are (code, count) pairs runs[2*run_off[i] ..), n_runs[i] of them, inside
 * space of its exact element count (run_off[n] = the batch's element
 * total, which must not exceed run_capacity; K2G_EUSAGE otherwise); HLL
 * records likewise at the same offsets.  The function returns when the
 * last kernels are enqueued on hip_stream (NULL = the null stream); the
 * results are ready when that stream is.  Every call on a GPU uses that
 * GPU's one array: calls that overlap on one GPU (on other host threads or
 * other streams before the previous call's stream finished) race. */
typedef struct {
  uint32_t *call, *hit_groups, *total_kmers, *n_runs; /* n entries each */
  uint64_t *run_off;                                  /* n + 1 */
  uint32_t *runs;                                     /* 2 * run_capacity words */
  uint64_t run_capacity;
  uint32_t *hll;         /* want_hll: 2 * run_capacity words, (taxon, encoded hash) pairs */
  uint32_t *n_hll;       /* want_hll: n */
  uint64_t *kmer_counts; /* want_kmer_counts: node_count */
} k2g_device_results;

int k2g_classify_device_noalloc(k2g_db *db, const k2g_classify_params *p, const char *d_seq,
                                const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                                const k2g_device_results *out, void *hip_stream);

/* ---- report: ReportKrakenStyle (src/reports.cc:241-259) of one classify
 * call's results, written to path.  The fragments (their offsets and paired
 * flags; seq is not read) are cut into the reference's input blocks
 * (src/classify.cc:523-543: single reads until 3 MiB of bases, 10 000 pairs;
 * a change between single and paired fragments starts a new block, as a new
 * input does).  Each classified fragment counts one read for its call.  With
 * kmer_data (-K) the per-taxon k-mer counts and distinct-minimizer
 * estimates are folded block by block from the results' HLL records, the
 * way the reference merges its per-block sketches (src/classify.cc:554,
 * 640-645; src/hyperloglogplus.cc:471-509); the results must then carry
 * them: per-hit records (want_hll 1) or the device's per-(block, taxon)
 * sets (want_hll 2, same blocks).  For a served request this is the report
 * the reference's daemon writes for the same reads run as one classify
 * call. */
typedef struct {
  int32_t report_zeros; /* -z */
  int32_t kmer_data;    /* -K */
  int32_t paired;       /* every fragment is a pair (when in->paired is NULL) */
} k2g_report_params;

int k2g_report_write(k2g_db *db, const k2g_fragments *in, const k2g_results *r,
                     const k2g_report_params *p, const char *path);

/* ---- graph-replayed batch pipeline.  No reference counterpart: the
 * reference classifies block by block on CPU threads (ProcessFiles,
 * src/classify.cc:479-708).  The per-batch launch chain of k2g_classify --
 * upload (host input), element bounds and hitlist layout, longest-first work
 * units, the persistent classify kernel, the dense-table redo and the
 * writeback of results into pinned host memory (host output) -- recorded
 * once as a HIP graph for a batch capacity and replayed for every batch
 * that fits, with no host round trip between its steps.  Work units come
 * from length buckets: fragments are binned by their element counts
 * (2^bucket_shift bins per doubling), the bins laid out longest first and
 * each cut into units of about unit_elements elements.  Results equal
 * k2g_classify's on the same fragments. */
typedef struct k2g_pipeline k2g_pipeline;

typedef struct {
  uint64_t max_fragments; /* batch capacity */
  uint64_t max_bases;
  int32_t host_io;        /* 1: batches from host memory, results into pinned host memory
                           * (k2g_pipeline_run, exact hitlist space); 0: device-resident
                           * batches, results left on the device (k2g_pipeline_run_device;
                           * run_cap_pct of the classify parameters reserved, 0 = 25 as in
                           * k2g_classify_device; fragments that outgrow their reservation
                           * take exact space from an overflow pool of an eighth of the
                           * reserved space, at least 2^20 elements, within the same replay) */
  int32_t unit_elements;  /* elements per work unit (reads per wavefront claim); 0 = 4096 */
  int32_t bucket_shift;   /* 2^bucket_shift length buckets per doubling, in [0, 2] */
  int32_t probe_depth;    /* chunks of 64 k-mers whose lookups one probe round takes, [1, 8]; 0 = 8 */
  int32_t waves_per_cu;   /* resident classify waves per compute unit; 0 = as many as fit */
} k2g_pipeline_params;

int k2g_pipeline_create(k2g_db *db, const k2g_classify_params *p, const k2g_pipeline_params *pp,
                        k2g_pipeline **out);
/* Host fragments -> host results (valid until the next run).  The bases,
 * offsets and flags are copied into the pipeline's pinned staging first
 * (nothing is copied when they already lie there). */
int k2g_pipeline_run(k2g_pipeline *pl, const k2g_fragments *in, k2g_results *out);
/* The host pipeline's pinned staging (max_bases bytes, 2 * max_fragments + 1
 * offsets, max_fragments pair flags): a reader that writes a batch straight
 * into it and passes these pointers to k2g_pipeline_run skips the copy.
 * K2G_EUSAGE for a device pipeline. */
int k2g_pipeline_staging(k2g_pipeline *pl, char **seq, uint64_t **off, uint8_t **paired);
/* Device-resident fragments; results stay in the pipeline's device buffers.
 * ms (may be NULL) receives the graph's device time. */
int k2g_pipeline_run_device(k2g_pipeline *pl, const char *d_seq, const uint64_t *d_off,
                            const uint8_t *d_paired, uint64_t n, double *ms);
/* Per-fragment result digests of the last run (as k2g_batch_digests). */
int k2g_pipeline_digests(k2g_pipeline *pl, uint64_t *out, uint64_t n);
/* The pipeline's parameters and its classify grid (workgroups). */
int k2g_pipeline_get_params(const k2g_pipeline *pl, k2g_pipeline_params *pp, int32_t *grid);
/* Fragments of the last batch classified again: out of reserved hitlist
 * space (into the overflow pool), and with the dense hit table. */
int k2g_pipeline_redo_counts(const k2g_pipeline *pl, uint64_t *space, uint64_t *dense);
/* Fragments of the last batch split over waves, and their segments (long
 * nucleotide fragments, when the batch's longest one outweighs a wave's
 * share).  No counterpart in the reference, which scans each read on one
 * thread; for tests and measurement. */
int k2g_pipeline_split_counts(const k2g_pipeline *pl, uint64_t *fragments, uint64_t *segments);
int k2g_pipeline_destroy(k2g_pipeline *pl);

/* Autotuner: chooses unit_elements, bucket_shift, probe_depth and
 * waves_per_cu for a workload by timing graph replays on a sample of it
 * (every stride-th fragment of the device-resident set), by coordinate
 * descent from *best (two sweeps).  Every configuration must give the same
 * per-fragment results; report->mismatched counts any that did not (0 on a
 * correct build).  *best receives the fastest configuration.  This
 * generalizes the reference's one tuning knob, the prefetch distance of
 * GetBatch (K2_PREFETCH_DIST, src/compact_hash.h:381-384). */
typedef struct {
  uint64_t sample_fragments, sample_bases;
  uint32_t evaluated;  /* configurations timed */
  uint32_t mismatched; /* configurations whose results differed from the first */
  double default_ms;   /* the starting configuration's graph time on the sample */
  double best_ms;      /* the chosen one's */
} k2g_tune_report;

int k2g_pipeline_tune(k2g_db *db, const k2g_classify_params *p, const char *d_seq,
                      const uint64_t *d_off, const uint8_t *d_paired, uint64_t n, uint64_t stride,
                      k2g_pipeline_params *best, k2g_tune_report *report);

/* Pinned host memory for request fragments (and anything else a caller
 * streams to the GPU): uploads from it run at full interconnect speed and
 * asynchronously.  No reference counterpart. */
int k2g_host_alloc(size_t bytes, void **out);
int k2g_host_free(void *p);

/* 64-bit checksum of a result set (every array in order, n and node_count
 * included), computed on the host: two result sets of the same fragments are
 * equal iff their checksums are (up to hash collisions).  No reference
 * counterpart: a verification aid. */
int k2g_results_checksum(const k2g_results *r, uint64_t *out);

/* Random-gather ceiling of the resident table: n_probes independent 16-byte
 * loads at uniformly random 128-byte-aligned offsets of the table's cells
 * (the access a hash probe makes), timed on the device.  Returns the rate in
 * distinct lines per second.  No reference counterpart: the measured
 * ceiling that probe throughput is compared with. */
int k2g_db_gather_ceiling(k2g_db *db, uint64_t n_probes, uint64_t seed, double *lines_per_s,
                          double *kernel_ms);

/* ---- serving: the resident-index request loop.  The reference keeps its
 * index loaded in ClassifyDaemon (src/classify.cc:373-430) and runs whole
 * classify() calls one after another.  Here requests are admitted in arrival
 * order whenever a slot is free and the device memory a request needs fits in
 * what is left of a fixed budget (one allocation made by k2g_server_create,
 * carved per request); admitted requests run concurrently, each on its own
 * slot's HIP stream, while the table stays resident.  A request's results
 * equal those of k2g_classify on the same fragments alone. */
typedef struct k2g_server k2g_server;

typedef struct {
  uint64_t hbm_budget; /* device bytes for request state: reads, hitlists, HLL records, scratch */
  int32_t slots;       /* requests on the GPU at once, in [1, 64] */
} k2g_server_params;

int k2g_server_create(k2g_db *db, const k2g_server_params *p, k2g_server **out);
/* Queues a request (FIFO).  The fragment arrays must stay valid until
 * k2g_server_wait returns for it.  Hitlist space is always exact
 * (params->run_cap_pct is ignored).  K2G_EUSAGE if the request alone needs
 * more than the budget. */
int k2g_server_submit(k2g_server *s, const k2g_classify_params *p, const k2g_fragments *in,
                      uint64_t *ticket);
/* Blocks until the request is done.  out (may be NULL) points into memory
 * that stays valid until k2g_server_release; times_ms (may be NULL)
 * receives {queued, on the GPU, total} milliseconds. */
int k2g_server_wait(k2g_server *s, uint64_t ticket, k2g_results *out, double *times_ms);
int k2g_server_release(k2g_server *s, uint64_t ticket);
/* Highest device memory held by admitted requests, most requests on the GPU
 * at once, and the budget. */
int k2g_server_stats(k2g_server *s, uint64_t *peak_bytes, int32_t *max_concurrent,
                     uint64_t *budget);
/* Finishes queued requests, then frees the server. */
int k2g_server_destroy(k2g_server *s);

/* ---- build side: the table construction of build_db
 * (src/build_db.h:186-326; CompareAndSet src/compact_hash.h:416-462)
 * as lock-free atomicCAS with an in-kernel LCA merge.  Sequences are scanned
 * in the deterministic build's windows (blocks of block_size bases split
 * into subblocks of subblock_size, each overlapping the next by k-1), so the
 * set of (minimizer, taxon) pairs inserted equals the reference's; cell
 * placement may differ from a CPU build (it is order dependent there too). */
typedef struct k2g_builder k2g_builder;

typedef struct {
  uint64_t k, l;
  uint64_t spaced_seed_mask;  /* already expanded (src/utilities.cc:14-27) */
  uint64_t toggle_mask;
  uint64_t min_clear_hash_value;
  uint64_t block_size, subblock_size; /* build_db -B / -b */
  int32_t dna_db;
  int32_t revcom_version; /* 1 for new databases */
} k2g_build_params;

/* capacity cells; value_bits = 0 picks the smallest width that holds every
 * node id (src/build_db.cc:66-75); key_bits = 32 - value_bits. */
int k2g_build_create(const char *taxo_k2d, uint64_t capacity, uint64_t value_bits,
                     const k2g_build_params *p, k2g_builder **out);
/* Sequence i = seq[off[i], off[i+1]) belongs to internal taxon taxid[i]. */
int k2g_build_insert(k2g_builder *b, const char *seq, const uint64_t *off, const uint32_t *taxid,
                     uint64_t n);
/* Same with seq in HBM (off and taxid stay in host memory). */
int k2g_build_insert_device(k2g_builder *b, const char *d_seq, const uint64_t *off,
                            const uint32_t *taxid, uint64_t n);
int k2g_build_get_info(k2g_builder *b, uint64_t *capacity, uint64_t *size, uint64_t *key_bits,
                       uint64_t *value_bits);

/* Verification of a built table against the sequences it was built from.
 * No reference counterpart: a check that src/build_db.cc never makes.  The
 * build windows are scanned again on the device, and every (minimizer, taxid)
 * pair the build inserts (src/build_db.h:186-326) is collected, one hash
 * partition at a time.  Each distinct minimizer is then looked up in the
 * builder's table.  A minimizer whose probe meets a cell holding another
 * minimizer's identical compacted key shares that cell
 * (src/compact_hash.h:432-447), so the cell holds the LCA over both.  The
 * check therefore groups minimizers by the cell their lookup ends at and
 * requires every cell to hold exactly the LCA of the taxa of the minimizers
 * ending there.  Arguments are those of k2g_build_insert_device for the
 * same sequences. */
typedef struct {
  uint64_t occurrences;  /* (minimizer, taxid) pairs inserted */
  uint64_t keys;         /* distinct minimizers */
  uint64_t cells;        /* distinct cells their lookups end at */
  uint64_t missing;      /* minimizers whose lookup finds nothing: must be 0 */
  uint64_t not_covering; /* minimizers whose value is not an ancestor-or-self of each of their
                          * taxa: must be 0 */
  uint64_t shared_cells; /* cells reached by more than one minimizer (compacted-key collisions) */
  uint64_t shared_keys;  /* minimizers in those cells */
  uint64_t differ;       /* minimizers whose value is not the LCA of their own taxa */
  uint64_t unexplained;  /* cells whose value is not the LCA of the taxa of every minimizer
                          * ending there: must be 0, which puts every differing minimizer in
                          * a shared cell */
  uint64_t orphan_cells; /* filled cells no minimizer reaches: must be 0 */
  uint32_t partitions;   /* hash partitions the pairs were collected in */
  uint32_t pad;
} k2g_build_check;

int k2g_build_verify(k2g_builder *b, const char *d_seq, const uint64_t *off, const uint32_t *taxid,
                     uint64_t n, k2g_build_check *out);
int k2g_build_download(k2g_builder *b, uint32_t *cells, uint64_t capacity);
/* hash.k2d in the reference's WriteTable layout (src/compact_hash.h:343-352). */
int k2g_build_write(k2g_builder *b, const char *hash_k2d);
int k2g_build_destroy(k2g_builder *b);

/* estimate_capacity (src/estimate_capacity.cc:66-100, 185-203): exact count
 * of the distinct minimizers whose MurmurHash3 falls in the first n_ranges
 * of 1024 hash ranges, +1, scaled by 1024 / n_ranges.  Whole sequences are
 * scanned (block and subblock sizes are ignored). */
int k2g_estimate_capacity(const k2g_build_params *p, const char *seq, const uint64_t *off,
                          uint64_t n, uint32_t n_ranges, uint64_t *estimate);
int k2g_estimate_capacity_device(const k2g_build_params *p, const char *d_seq,
                                 const uint64_t *off, uint64_t n, uint32_t n_ranges,
                                 uint64_t *estimate);

/* ---- multi-database merge: the per-read core of the reference's merge tool
 * (merge_hit_lists src/merge.cc:494-554, resolve_tree :369-421, the run
 * folding of write_hit_list :328-367), which joins two kraken outputs of
 * the same reads classified against two databases that share a taxonomy.
 * Hit lists hold the values the reference parses into taxid_and_count
 * (src/merge.cc:34-37, 423-469): external taxids and counts as 32-bit
 * signed integers, with INT32_MAX for "A", INT32_MAX-1 for "|" and
 * INT32_MAX-2 for "-" (src/merge.cc:176-178).  Taxids are mapped to nodes as
 * the reference's multi-threaded path maps them (src/merge.cc:594). */
typedef struct k2g_merger k2g_merger;

#define K2G_MERGE_AMBIG 0x7FFFFFFF
#define K2G_MERGE_MATE 0x7FFFFFFE
#define K2G_MERGE_FRAME 0x7FFFFFFD

typedef struct {
  const int32_t *items; /* (taxid, count) pairs */
  const uint64_t *off;  /* n + 1 pair offsets, off[0] = 0: read i is pairs [off[i], off[i+1]) */
  uint64_t n;
} k2g_hit_lists;

typedef struct {
  float confidence; /* merge -i */
  int32_t text;     /* 1: results as printed text (text, text_off); 0: as pairs (token_off, tokens, joined) */
} k2g_merge_params;

typedef struct {
  uint64_t n;
  /* per read: resolve_tree's value, printed as the read's taxid when either
   * input calls the read (src/merge.cc:701-711); the external id of the
   * call, or the node id where the confidence climb returns one (:413) */
  const int32_t *call;
  /* n + 1 offsets: read i prints pairs [token_off[i], token_off[i+1]) of
   * tokens, as int_to_string prints each value (:251-326); a read with no
   * pairs prints no line (:727) */
  const uint64_t *token_off;
  const int32_t *tokens; /* (taxid, count) pairs */
  const uint8_t *joined; /* per pair: 1 = no space follows the pair */
  /* text mode: read i's hit list as write_hit_list prints it (no newline) is
   * text[text_off[i] .. text_off[i+1]); the pair arrays are then NULL */
  const char *text;
  const uint64_t *text_off;
  double kernel_ms;
} k2g_merge_results;

/* The taxonomy (taxo.k2d, the merged taxonomy of both databases) resident
 * on the current device. */
int k2g_merge_open(const char *taxo_k2d, k2g_merger **out);
/* Merges read i of a with read i of b for every i (a->n == b->n).  Results
 * stay valid until the next call on m or k2g_merge_close. */
int k2g_merge(k2g_merger *m, const k2g_merge_params *p, const k2g_hit_lists *a,
              const k2g_hit_lists *b, k2g_merge_results *out);
int k2g_merge_close(k2g_merger *m);

/* The same merge with the hit lists still as text, tokenized on the GPU as
 * parse_hit_list does (src/merge.cc:423-469).  Read i's hit list is the
 * fifth tab-separated field of its kraken line, text[field[i] .. field[i] +
 * len[i]), with the line's newline when it has one; off[i+1] - off[i] must
 * be that field's ':' count (the reference's slot count, :431-435; checked
 * on the device).  off may be NULL: the counts are then taken on the
 * device. */
typedef struct {
  const char *text;
  uint64_t text_bytes;
  const uint64_t *field; /* n */
  const uint32_t *len;   /* n */
  const uint64_t *off;   /* n + 1, off[0] = 0; or NULL (counted on the device) */
  uint64_t n;
} k2g_hit_list_text;

int k2g_merge_text(k2g_merger *m, const k2g_merge_params *p, const k2g_hit_list_text *a,
                   const k2g_hit_list_text *b, k2g_merge_results *out);

#ifdef __cplusplus
}
#endif
#endif /* K2G_H_ */
