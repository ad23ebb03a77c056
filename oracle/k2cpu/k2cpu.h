// k2cpu -- CPU ORACLE (test infrastructure only).
//
// A restatement, in this repository's own code, of the reference Kraken 2
// classify / build path.  It exists to check the GPU path and to time a CPU
// baseline; nothing in the shipped library links or calls it.  Each piece
// names the reference file:line whose behaviour it follows.
//
// Parity pinning (how each function is checked; see tests/):
//   Scanner              golden per-k-mer vectors from the compiled reference
//                        scanner (tests/golden/scan_*), plus a brute-force
//                        window-minimum definition on fuzzed strings.
//   fmix64               closed form (smhasher constants) + golden Get vectors.
//   HashTable::Get       every key of golden tables + table invariants.
//   ResolveTree / line   golden kraken outputs of the compiled reference.
//   Hll                  golden (n_observed, estimate) vectors of the compiled
//                        reference HLL.
//   Reports              golden kreports (with -K) of the compiled reference.
//   build (deterministic)byte-identical hash.k2d against the compiled
//                        reference build on the golden inputs.
//   estimate_capacity    golden estimates of the compiled reference.
//   translate            golden six-frame vectors of the compiled reference.
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace k2cpu {

// ---- src/kraken2_data.h:15-25 (raw 64-byte on-disk layout of opts.k2d)
struct IndexOptions {
  size_t k;
  size_t l;
  uint64_t spaced_seed_mask;
  uint64_t toggle_mask;
  bool dna_db;
  uint64_t minimum_acceptable_hash_value;
  int revcom_version;
  int db_version;
  int db_type;
};
static_assert(sizeof(IndexOptions) == 64, "opts.k2d layout");

bool read_index_options(const std::string &path, IndexOptions &out, std::string &err);
bool write_index_options(const std::string &path, const IndexOptions &o);

constexpr uint64_t kDefaultToggleMask = 0xe37e28c4271b5a2dULL;  // src/mmscanner.h:15
constexpr uint64_t kTaxidMax = ~0ULL;                             // src/kraken2_data.h:28
constexpr uint64_t kMateBorder = kTaxidMax;                       // src/classify.cc:33
constexpr uint64_t kFrameBorder = kTaxidMax - 1;                  // src/classify.cc:34
constexpr uint64_t kAmbigSpan = kTaxidMax - 2;                    // src/classify.cc:35

// ---- src/kv_store.h:59-67
inline uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// ---- src/utilities.cc:14-27: each template bit becomes `factor` bits.
uint64_t expand_spaced_seed_mask(uint64_t mask, int factor);
uint64_t parse_bitstring(const char *s);  // strtol(s, nullptr, 2) semantics

// ---- Taxonomy: src/taxonomy.h:18-26, src/taxonomy.cc:195-272, 275-287
struct TaxNode {
  uint64_t parent_id, first_child, child_count, name_offset, rank_offset,
      external_id, godparent_id;
};

class Taxonomy {
 public:
  bool load(const std::string &path, std::string &err);
  bool write(const std::string &path) const;
  // a is an ancestor of b or equal to it; false if either is 0.
  bool is_ancestor(uint64_t a, uint64_t b) const;
  uint64_t lca(uint64_t a, uint64_t b) const;
  size_t node_count() const { return nodes.size(); }
  const char *name(uint64_t id) const { return names.data() + nodes[id].name_offset; }
  const char *rank(uint64_t id) const { return ranks.data() + nodes[id].rank_offset; }
  uint64_t internal_id(uint64_t external) const;  // 0 if unknown
  void build_external_map();

  std::vector<TaxNode> nodes;
  std::vector<char> names, ranks;
  std::unordered_map<uint64_t, uint64_t> ext_to_int;
};

// NCBI nodes.dmp/names.dmp -> pruned BFS taxonomy (src/taxonomy.cc:19-185)
class NcbiTaxonomy {
 public:
  bool parse(const std::string &nodes_path, const std::string &names_path, std::string &err);
  void mark(uint64_t taxid);
  Taxonomy convert() const;

 private:
  std::map<uint64_t, uint64_t> parent_;
  std::map<uint64_t, std::string> name_, rank_;
  std::map<uint64_t, std::set<uint64_t>> children_;
  std::set<uint64_t> marked_{1};
  std::set<std::string> known_ranks_;
};

// ---- CompactHashTable<CompactHashCell> with LINEAR_PROBING
//      src/compact_hash.h:22-48, 278-462, 493-500
class HashTable {
 public:
  HashTable() = default;
  HashTable(uint64_t capacity, uint64_t key_bits, uint64_t value_bits);
  bool load(const std::string &path, std::string &err);
  bool write(const std::string &path) const;
  uint32_t get(uint64_t key) const;
  void get_batch(const uint64_t *keys, uint32_t *out, size_t n) const;
  bool find_index(uint64_t key, uint64_t *idx) const;
  bool compare_and_set(uint64_t key, uint32_t new_value, uint32_t *old_value);
  // number of distinct 128-byte lines of the cell array a Get(key) touches
  // (home cell through the terminating cell), counted from the table start.
  uint32_t probe_lines(uint64_t key) const;

  uint64_t capacity = 0, size = 0, key_bits = 0, value_bits = 0;
  std::vector<uint32_t> cells;
};

// ---- MinimizerScanner: src/mmscanner.cc:20-220, src/mmscanner.h:26-72
class Scanner {
 public:
  Scanner(ssize_t k, ssize_t l, uint64_t spaced_seed_mask, bool dna,
          uint64_t toggle_mask, int revcom_version = 1);
  void load(const std::string &seq, size_t start = 0, size_t finish = SIZE_MAX);
  const uint64_t *next();
  bool ambiguous() const { return window_len_ < k_ - l_ || recent_ambig_ != 0; }
  ssize_t k() const { return k_; }
  ssize_t l() const { return l_; }

 private:
  uint64_t canonical(uint64_t lmer) const;
  struct Entry {
    uint64_t cand;
    ssize_t pos;
  };
  const std::string *seq_ = nullptr;
  ssize_t k_, l_;
  size_t pos_ = 0, start_ = 0, finish_ = 0;
  uint64_t spaced_, toggle_, lmer_ = 0, lmer_mask_, current_ = 0, recent_ambig_ = 0;
  bool dna_;
  int revcom_version_;
  ssize_t filled_ = 0;      // characters in the current l-mer
  ssize_t window_len_ = 0;  // l-mers seen since the last restart
  // monotone deque stored in a ring
  std::vector<Entry> ring_;
  size_t head_ = 0, count_ = 0;
  uint8_t code_[256];
};

// ---- src/aa_translate.cc:41-100
void translate_all_frames(const std::string &dna, std::vector<std::string> &frames);

// ---- HyperLogLogPlusMinus<uint64_t>, p=12 sparse p'=25, Ertl estimator
//      src/hyperloglogplus.h:46-106, src/hyperloglogplus.cc:117-190, 459-651, 708-739, 816-824
class Hll {
 public:
  void insert(uint64_t item);
  void merge(const Hll &o);
  uint64_t estimate() const;
  uint64_t n_observed = 0;
  bool sparse = true;
  std::set<uint32_t> sparse_list;
  std::vector<uint8_t> regs;  // 4096 when dense

  static uint64_t hash(uint64_t item);                 // fmix64(item + 1)
  static uint32_t encode_sparse(uint64_t hash_value);  // 25-bit index + rank
  static uint8_t encoded_rank(uint32_t enc);
  static uint8_t dense_rank(uint64_t hash_value);
  void to_dense();
};

struct ReadCounter {  // src/readcounts.h:19-88
  uint64_t n_reads = 0, n_kmers = 0;
  Hll kmers;
  void add_kmer(uint64_t m) { n_kmers++; kmers.insert(m); }
  void add(const ReadCounter &o) {
    n_reads += o.n_reads;
    n_kmers += o.n_kmers;
    kmers.merge(o.kmers);
  }
};
using TaxonCounters = std::map<uint64_t, ReadCounter>;

// ---- FASTA/FASTQ records (kseq semantics, src/kseq.h:172-230 via
//      src/seqreader.h:65-194); header stops at the first whitespace.
struct Sequence {
  bool fastq = false;
  std::string header, comment, seq, quals;
  std::string to_string() const;  // src/seqreader.cc:20-49
};

class SeqReader {
 public:
  bool open(const char *path);  // nullptr = stdin
  ~SeqReader();
  bool read(Sequence &s);
  // src/seqreader.h:101-119: records until the sum of lengths >= bytes
  size_t load_block(size_t bytes, std::vector<Sequence> &out);
  // src/seqreader.h:146-163: up to n records
  size_t load_batch(size_t n, std::vector<Sequence> &out);

 private:
  int getc_();
  int fd_ = -1;
  bool own_ = false;
  std::vector<char> buf_;
  size_t beg_ = 0, end_ = 0;
  bool eof_ = false;
  int pending_ = 0;  // header char consumed by the previous record
};

// ---- classify (src/classify.cc)
struct ClassifyOptions {
  std::string index_filename, taxonomy_filename, options_filename, report_filename,
      classified_output_filename, unclassified_output_filename, kraken_output_filename;
  bool mpa_style_report = false, report_kmer_data = false, quick_mode = false,
       report_zero_counts = false, use_translated_search = false,
       print_scientific_name = false, paired_end_processing = false,
       single_file_pairs = false, use_memory_mapping = false;
  double confidence_threshold = 0;
  int num_threads = 1, minimum_quality_score = 0, minimum_hit_groups = 0;
  std::vector<std::string> filenames;
};

struct ClassifyStats {
  uint64_t total_sequences = 0, total_bases = 0, total_classified = 0;
};

struct Index {
  IndexOptions opts{};
  Taxonomy tax;
  HashTable table;
};

bool load_index(const ClassifyOptions &o, Index &idx, std::string &err);

// Per-read result of the classify core, before text formatting.
struct ReadResult {
  uint64_t call = 0;
  int64_t hit_groups = 0;
  uint64_t total_kmers = 0;
  std::vector<uint64_t> taxa;  // one entry per k-mer plus border markers
  uint32_t lookups = 0;        // distinct-minimizer lookups issued
  uint64_t probe_lines = 0;    // 128-byte lines the lookups touched
};

// One fragment (single read or pair) through scan -> GetBatch -> replay ->
// ResolveTree (src/classify.cc:773-906).  counters may be null (no -R).
void classify_fragment(const Index &idx, const ClassifyOptions &o, const std::string &seq1,
                       const std::string *seq2, ReadResult &res, TaxonCounters *counters,
                       bool count_lines = false);
uint64_t resolve_tree(std::unordered_map<uint64_t, uint64_t> &hit_counts, const Taxonomy &tax,
                      size_t total_minimizers, double confidence);
void append_hitlist(std::string &out, const std::vector<uint64_t> &taxa, const Taxonomy &tax);
std::string trim_pair_info(const std::string &id);
// kraken output line for one fragment (src/classify.cc:908-945)
void format_line(std::string &out, const Index &idx, const ClassifyOptions &o, const Sequence &s1,
                 const Sequence *s2, const ReadResult &r);

// Whole-run driver (src/classify.cc:304-352, 479-708): returns exit status.
int run_classify(const ClassifyOptions &o, const Index &idx, ClassifyStats &stats,
                 TaxonCounters &counters, double *seconds);
void report_stats(double seconds, const ClassifyStats &s);  // src/classify.cc:447-477

// ---- reports (src/reports.cc:32-45, 130-259)
void report_kraken_style(const std::string &path, bool report_zeros, bool report_kmer_data,
                         const Taxonomy &tax, const TaxonCounters &call_counters,
                         uint64_t total_seqs, uint64_t total_unclassified);

// ---- build (src/build_db.cc, src/build_db.h)
struct BuildOptions {
  std::string hashtable_filename, id_map_filename, ncbi_taxonomy_directory, options_filename,
      taxonomy_filename;
  size_t block_size = 10 * 1024 * 1024, subblock_size = 1024;
  size_t requested_bits_for_taxid = 0;
  int num_threads = 1;
  bool input_is_protein = false, deterministic = true;
  ssize_t k = 0, l = 0;
  size_t capacity = 0, maximum_capacity = 0;
  uint64_t spaced_seed_mask = 0, toggle_mask = kDefaultToggleMask, min_clear_hash_value = 0;
};
int run_build(BuildOptions &o);
std::vector<std::string> extract_ncbi_sequence_ids(const std::string &header);

// Minimizers the deterministic build inserts for one sequence block, in the
// order it inserts them (src/build_db.h:216-320).
void block_minimizers(const BuildOptions &o, const std::string &seq, size_t block_start,
                      std::vector<uint64_t> &ordered);

// ---- estimate_capacity (src/estimate_capacity.cc:66-100, 185-203)
struct EstimateOptions {
  size_t k = 0, l = 0, n = 4;
  bool input_is_protein = false;
  int threads = 1;
  uint64_t spaced_seed_mask = 0, toggle_mask = kDefaultToggleMask;
};
uint64_t estimate_capacity(const EstimateOptions &o, const char *path);

}  // namespace k2cpu
