// Host-side pieces of the classify front-end around libk2g: sequence input,
// taxonomy names/ranks, distinct-minimizer estimates and the kreport.
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace k2host {

// One FASTA/FASTQ record with kseq field semantics (src/seqreader.h:29-61).
struct Record {
  bool fastq = false;
  std::string header, comment, seq, quals;
  void append_text(std::string &out, const char *header_suffix) const;  // src/seqreader.cc:20-49
};

// Buffered reader (src/kseq.h record rules: '>'/'@' headers, multi-line
// FASTA, '+' separated FASTQ, one trailing '\r' dropped per line).
class Reader {
 public:
  bool open(const char *path);  // nullptr = stdin
  ~Reader();
  bool next(Record &r);
  // src/seqreader.h:101-119 and 146-163
  size_t load_block(size_t bases, std::vector<Record> &out);
  size_t load_batch(size_t records, std::vector<Record> &out);
  // The next records of about `bytes` of input, appended to out in input
  // order: the same records next() returns.  Plain FASTA (every record at a
  // line starting with '>', no line starting with '@' or '+') is cut at
  // record starts and parsed on `threads` threads; anything else, and all
  // input after it, is read by next().  Returns the records appended.
  // pool (may be null): used records whose buffers the parse reuses.
  size_t next_records(size_t bytes, int threads, std::vector<Record> &out,
                      std::vector<Record> *pool = nullptr);
  // Parses bytes held in memory (as if they were the whole input); they must
  // stay valid while the reader is used.
  void open_memory(const char *p, size_t n);

 private:
  int get();
  void line_into(std::string &dst, bool append, int *last);
  const char *base() const { return ext_ ? ext_ : buf_.data(); }
  const char *ext_ = nullptr;  // open_memory: the bytes, read in place
  int fd_ = -1;
  bool own_ = false, eof_ = false;
  std::vector<char> buf_;
  size_t pos_ = 0, end_ = 0;
  int pending_ = 0;
  bool serial_ = false;  // next_records: no more parallel chunks
};

// taxo.k2d for names, ranks, parents and external ids (src/taxonomy.cc:195-238).
struct Taxonomy {
  struct Node {
    uint64_t parent_id, first_child, child_count, name_offset, rank_offset, external_id,
        godparent_id;
  };
  std::vector<Node> nodes;
  std::vector<char> names, ranks;
  bool load(const std::string &path, std::string &err);
  const char *name(uint64_t id) const { return names.data() + nodes[id].name_offset; }
  const char *rank(uint64_t id) const { return ranks.data() + nodes[id].rank_offset; }
};

// NCBI taxonomy as read from nodes.dmp / names.dmp, with the set of taxa to
// keep (the root and ancestors of mapped taxa), written as taxo.k2d
// (src/taxonomy.cc:19-185, 275-287).
struct NcbiTaxonomy {
  std::map<uint64_t, uint64_t> parent;
  std::map<uint64_t, std::string> rank, name;
  std::map<uint64_t, std::vector<uint64_t>> children;
  std::set<uint64_t> marked;
  bool load(const std::string &nodes_path, const std::string &names_path, std::string &err);
  void mark(uint64_t taxid);  // the taxon and its ancestors
  bool write_kraken(const std::string &path, std::string &err) const;
};

// Sequence ids of a FASTA header (src/build_db.cc:113-148).
std::vector<std::string> sequence_ids(const std::string &header);
// src/utilities.cc:14-27
uint64_t expand_spaced_seed_mask(uint64_t mask, int bits_per_char);

// Distinct-minimizer state of one taxon or clade.  The reference's
// HyperLogLogPlusMinus (p=12, sparse p'=25) ends in a state fixed by the set
// of sparse-encoded hashes it saw and whether any per-block sketch switched
// to registers (src/hyperloglogplus.cc:471-509, 572-651); both are kept here.
struct Distinct {
  uint64_t n_observed = 0;
  bool dense = false;
  std::vector<uint32_t> encoded;  // sorted, unique
  void merge(const Distinct &o);
  uint64_t estimate() const;      // Ertl estimator, src/hyperloglogplus.cc:708-739
};

struct TaxonCounter {
  uint64_t n_reads = 0, n_kmers = 0;
  Distinct distinct;
};
using TaxonCounters = std::map<uint64_t, TaxonCounter>;

// Per-block HLL bookkeeping: records (taxon, encoded) in insertion order for
// one input block (src/classify.cc:554, 640-645).
void fold_block_records(const std::vector<std::pair<uint32_t, uint32_t>> &recs,
                        TaxonCounters &total);

// src/reports.cc:241-259
bool report_kraken_style(const std::string &path, bool report_zeros, bool kmer_data,
                         const Taxonomy &tax, const TaxonCounters &calls, uint64_t total_seqs,
                         uint64_t total_unclassified);

}  // namespace k2host
