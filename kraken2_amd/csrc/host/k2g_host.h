// Host-side pieces of the classify front-end around libk2g: sequence input,
// taxonomy names/ranks, distinct-minimizer estimates and the kreport.
#pragma once
#include <cstddef>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace k2host {

// One FASTA/FASTQ record with kseq field semantics (src/seqreader.h:29-61).
struct Record {
  bool fastq = false;
  std::string header, comment, seq, quals;
  void append_text(std::string &out, const char *header_suffix) const;  // src/seqreader.cc:20-49
};

// A record parsed into flat arrays: its id, comment and (FASTQ, unless the
// reader drops them) qualities one after another in a text buffer, its
// bases in a sequence sink.
struct FlatRec {
  uint64_t text_at = 0;  // id, then comment, then seq_len qualities if fastq
  uint64_t seq_at = 0;   // bases in the sink
  uint64_t seq_len = 0;
  uint32_t id_len = 0, cmt_len = 0;
  bool fastq = false;
};

// Where a flat parse writes bases: p[0, n) written, p[n, cap) free.
// grow(sink, need) makes cap >= need and keeps p[0, n) (p may move).
struct SeqSink {
  char *p = nullptr;
  size_t n = 0, cap = 0;
  std::function<void(SeqSink &, size_t)> grow;
  char *room(size_t k) {
    if (n + k > cap) grow(*this, n + k);
    return p + n;
  }
};

// Threads that wait (blocked, not spinning) for work: run(n, fn) calls
// fn(0..n-1) on the workers and the calling thread, items handed out one at
// a time, and returns when all are done.  One caller at a time.
class WorkPool {
 public:
  explicit WorkPool(int threads);
  ~WorkPool();
  void run(size_t n, const std::function<void(size_t)> &fn);
  int size() const { return (int) workers_.size() + 1; }

 private:
  void work();
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)> *fn_ = nullptr;
  size_t n_ = 0, acked_ = 0;
  std::atomic<size_t> next_{0};
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Seconds a reader spent filling its buffer and in the two passes of the
// flat parse.
struct ReadTiming {
  double fill = 0, count = 0, write = 0;
};

// A record whose bases span several lines, for next_inplace to join: its
// index and the end of its text in the buffer.
struct InplaceJoin {
  size_t rec;
  uint64_t end;
};

// A byte buffer that is not zero-filled when it grows: only the first
// `keep` bytes are copied over.
class ByteBuf {
 public:
  char *data() { return p_.get(); }
  const char *data() const { return p_.get(); }
  size_t size() const { return n_; }
  void resize(size_t n, size_t keep) {
    if (n > cap_) {
      std::unique_ptr<char[]> q(new char[n]);
      if (keep) memcpy(q.get(), p_.get(), std::min(keep, n_));
      p_ = std::move(q);
      cap_ = n;
    }
    n_ = n;
  }
  void clear() { n_ = 0; }

 private:
  std::unique_ptr<char[]> p_;
  size_t n_ = 0, cap_ = 0;
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
  // next_records into flat arrays: the records appended to recs, their ids,
  // comments and qualities to text, their bases to seq back to back (a
  // plain-FASTA chunk is counted, then written, by parts side by side).
  size_t next_flat(size_t bytes, int threads, std::vector<FlatRec> &recs, std::vector<char> &text,
                   SeqSink &seq);
  // next_flat for plain FASTA without a counting pass (FASTQ as next_flat
  // reads it): the records are
  // indexed where they lie in the chunk (a record's later sequence lines
  // moved up against its first), then the chunk is copied to raw as it
  // stands: record r's bases are raw.p[seq_at, seq_at + seq_len), with the
  // header text between records left in between.  Other input goes to
  // next_flat's serial path (bases back to back).
  size_t next_inplace(size_t bytes, int threads, std::vector<FlatRec> &recs, std::vector<char> &text,
                      SeqSink &raw);
  // Runs the chunk reads and parse parts on this pool (default: threads
  // started per chunk).
  void set_pool(WorkPool *pool) { pool_ = pool; }
  // next_flat and next_inplace: keep FASTQ qualities in the text (default);
  // without them a record's text is its id and comment only
  void keep_qualities(bool keep) { keep_quals_ = keep; }
  // Room for chunks of up to `bytes` from the start (no regrowth later).
  void reserve_chunk(size_t bytes) { buf_.resize(std::max(buf_.size(), bytes + (1 << 20)), end_); }
  ReadTiming timing;
  // Parses bytes held in memory (as if they were the whole input); they must
  // stay valid while the reader is used.
  void open_memory(const char *p, size_t n);

 private:
  int get();
  void line_into(std::string &dst, bool append, int *last);
  // Reads at least `bytes` unread bytes to the front of the buffer: 0 at
  // the end of input, 1 with cut = the end of the last whole record (plain
  // FASTA from the start), 2 when only next() can read on (serial_ set).
  int fill_chunk(size_t bytes, int threads, size_t &cut);
  // Up to `bytes` more input at dst (several preads side by side for a
  // regular file); sets eof_ when the input ends.
  size_t read_into(char *dst, size_t bytes, int threads);
  // next_inplace's per-part records, text and multi-line records
  std::vector<std::vector<FlatRec>> inplace_recs_;
  std::vector<std::vector<char>> inplace_text_;
  std::vector<std::vector<InplaceJoin>> inplace_joins_;
  const char *base() const { return ext_ ? ext_ : buf_.data(); }
  const char *ext_ = nullptr;  // open_memory: the bytes, read in place
  int fd_ = -1;
  bool own_ = false, eof_ = false;
  ByteBuf buf_;
  size_t pos_ = 0, end_ = 0;
  int pending_ = 0;
  bool serial_ = false;  // next_records: no more parallel chunks
  WorkPool *pool_ = nullptr;
  bool keep_quals_ = true;
  bool flat_fastq(size_t cut, int threads, std::vector<FlatRec> &recs, std::vector<char> &text, SeqSink &seq,
                  size_t before);
  void run(int t, const std::function<void(size_t)> &fn);
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
