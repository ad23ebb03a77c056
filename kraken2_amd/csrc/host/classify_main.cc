// k2g_classify: the classify front-end over libk2g.
//
// Same command line as the reference's classify binary (getopt string of
// src/classify.cc:1076), so scripts/kraken2 and scripts/k2 can exec it
// unchanged.  Input is cut into the reference's blocks (3 MiB of bases for
// single reads, 10 000 pairs from two files, 20 000 interleaved records;
// src/classify.cc:523-543); several blocks go to the GPU as one batch, and
// the output, statistics and per-taxon counters are produced block by block
// in input order exactly as the reference's ordered output does.
#include <getopt.h>
#include <malloc.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "k2g.h"
#include "k2g_host.h"

using k2host::Record;

namespace {

struct Options {
  std::string index_filename, taxonomy_filename, options_filename, report_filename,
      classified_output_filename, unclassified_output_filename, kraken_output_filename;
  bool mpa_style_report = false, report_kmer_data = false, quick_mode = false,
       report_zero_counts = false, print_scientific_name = false, paired_end_processing = false,
       single_file_pairs = false;
  double confidence_threshold = 0;
  int num_threads = 1, minimum_quality_score = 0, minimum_hit_groups = 0;
  int gpu_mode = 2;                 // K2G_MODE: 0 plain, 1 persistent, 2 chosen per batch
  size_t batch_bases = 256u << 20;  // bases per GPU batch (several blocks)
  std::vector<std::string> filenames;
};

struct Stats {
  uint64_t total_sequences = 0, total_bases = 0, total_classified = 0;
};

void usage(int code) {
  std::cerr << "Usage: classify [options] <fasta/fastq file(s)>\n\n"
               "Options: (*mandatory)\n"
               "* -H filename      Kraken 2 index filename\n"
               "* -t filename      Kraken 2 taxonomy filename\n"
               "* -o filename      Kraken 2 options filename\n"
               "  -q               Quick mode\n"
               "  -M               Use memory mapping to access hash & taxonomy\n"
               "  -T NUM           Confidence score threshold (def. 0)\n"
               "  -p NUM           Number of host threads (def. 1)\n"
               "  -Q NUM           Minimum quality score (FASTQ only, def. 0)\n"
               "  -P               Process pairs of reads\n"
               "  -S               Process pairs with mates in same file\n"
               "  -R filename      Print report to filename\n"
               "  -m               In comb. w/ -R, use mpa-style report\n"
               "  -z               In comb. w/ -R, report taxa w/ 0 count\n"
               "  -n               Print scientific name instead of taxid in Kraken output\n"
               "  -g NUM           Minimum number of hit groups needed for call\n"
               "  -C filename      Filename/format to have classified sequences\n"
               "  -U filename      Filename/format to have unclassified sequences\n"
               "  -O filename      Output file for normal Kraken output\n"
               "  -K               In comb. w/ -R, provide minimizer information in report\n";
  exit(code);
}

void die(int code, const std::string &msg) {
  std::cerr << "classify: " << msg << "\n";
  exit(code);
}

void parse(int argc, char **argv, Options &o) {
  int opt;
  while ((opt = getopt(argc, argv, "h?H:t:o:T:p:R:C:U:O:Q:g:nmzqPSMKD")) != -1) {
    switch (opt) {
      case 'h': case '?': usage(0); break;
      case 'H': o.index_filename = optarg; break;
      case 't': o.taxonomy_filename = optarg; break;
      case 'T':
        o.confidence_threshold = std::stod(optarg);
        if (o.confidence_threshold < 0 || o.confidence_threshold > 1)
          die(64, "confidence threshold must be in [0, 1]");
        break;
      case 'o': o.options_filename = optarg; break;
      case 'q': o.quick_mode = true; break;
      case 'p':
        o.num_threads = atoi(optarg);
        if (o.num_threads < 1) die(64, "number of threads can't be less than 1");
        break;
      case 'g': o.minimum_hit_groups = atoi(optarg); break;
      case 'P': o.paired_end_processing = true; break;
      case 'S': o.paired_end_processing = o.single_file_pairs = true; break;
      case 'm': o.mpa_style_report = true; break;
      case 'K': o.report_kmer_data = true; break;
      case 'R': o.report_filename = optarg; break;
      case 'z': o.report_zero_counts = true; break;
      case 'C': o.classified_output_filename = optarg; break;
      case 'U': o.unclassified_output_filename = optarg; break;
      case 'O': o.kraken_output_filename = optarg; break;
      case 'n': o.print_scientific_name = true; break;
      case 'Q': o.minimum_quality_score = atoi(optarg); break;
      case 'M': break;  // the table is resident in HBM either way
      case 'D': die(64, "daemon mode is served by the k2g request loop, not this front-end");
    }
  }
  if (o.index_filename.empty() || o.taxonomy_filename.empty() || o.options_filename.empty()) {
    std::cerr << "classify: mandatory filename missing\n";
    usage(64);
  }
  if (o.mpa_style_report && o.report_filename.empty()) {
    std::cerr << "classify: -m requires -R be used\n";
    usage(64);
  }
  if (o.mpa_style_report) die(64, "mpa-style reports are not provided by this front-end");
  for (int i = optind; i < argc; i++) o.filenames.push_back(argv[i]);
  if (const char *m = getenv("K2G_MODE")) {
    const int v = atoi(m);
    o.gpu_mode = v >= 0 && v <= 2 ? v : 2;
  }
  if (const char *b = getenv("K2G_BATCH_BASES")) o.batch_bases = strtoull(b, nullptr, 10);
}

struct Outputs {
  FILE *kraken = stdout, *c1 = nullptr, *c2 = nullptr, *u1 = nullptr, *u2 = nullptr;
};

FILE *open_out(const std::string &path) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    std::cerr << "\rUnable to open file: " << path << ", reason: " << strerror(errno) << "\n";
    exit(EXIT_FAILURE);
  }
  return f;
}

// src/classify.cc:1008-1059 (paired names expand '#' to _1/_2)
void open_outputs(const Options &o, Outputs &out) {
  auto split = [&](const std::string &fmt, FILE *&a, FILE *&b) {
    std::vector<std::string> parts;
    size_t p = 0;
    for (;;) {
      size_t q = fmt.find('#', p);
      if (q == std::string::npos || parts.size() == 2) {
        parts.push_back(fmt.substr(p));
        break;
      }
      parts.push_back(fmt.substr(p, q - p));
      p = q + 1;
    }
    if (parts.size() < 2) die(65, "Paired filename format missing # character: " + fmt);
    if (parts.size() > 2) die(65, "Paired filename format has >1 # character: " + fmt);
    a = open_out(parts[0] + "_1" + parts[1]);
    b = open_out(parts[0] + "_2" + parts[1]);
  };
  if (!o.classified_output_filename.empty()) {
    if (o.paired_end_processing)
      split(o.classified_output_filename, out.c1, out.c2);
    else
      out.c1 = open_out(o.classified_output_filename);
  }
  if (!o.unclassified_output_filename.empty()) {
    if (o.paired_end_processing)
      split(o.unclassified_output_filename, out.u1, out.u2);
    else
      out.u1 = open_out(o.unclassified_output_filename);
  }
  if (!o.kraken_output_filename.empty())
    out.kraken = o.kraken_output_filename == "-" ? nullptr : open_out(o.kraken_output_filename);
}

void mask_low_quality(Record &r, int q) {  // src/classify.cc:1061-1071
  if (!r.fastq) return;
  if (r.seq.size() != r.quals.size()) {
    std::cerr << "classify: " << r.header << ": Sequence length (" << r.seq.size()
              << ") != Quality string length (" << r.quals.size() << ")\n";
    exit(65);
  }
  for (size_t i = 0; i < r.seq.size(); i++)
    if ((r.quals[i] - '!') < q) r.seq[i] = 'x';
}

// Runs fn(i) for every i in [0, n) on up to `threads` threads, items handed
// out one at a time.  The workers are plain threads that end with the loop:
// OpenMP's workers spin between parallel regions and take the cores that the
// reader thread's parse needs at the same time.
template <class F>
void parallel_for(size_t n, int threads, F fn) {
  const size_t t = std::min<size_t>(n, (size_t) std::max(1, threads));
  if (t <= 1) {
    for (size_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < n;) fn(i);
  };
  std::vector<std::thread> ts;
  for (size_t k = 1; k < t; k++) ts.emplace_back(work);
  work();
  for (auto &th : ts) th.join();
}

// Output text written through a raw pointer into space reserved ahead: no
// per-character bounds checks, and no zero fill once the buffer has grown.
struct Text {
  std::vector<char> v;
  size_t n = 0;
  char *reserve(size_t k) {
    if (n + k > v.size()) v.resize(std::max(2 * v.size(), n + k));
    return v.data() + n;
  }
  void commit(const char *end) { n = end - v.data(); }
};

char *put_uint(char *p, uint64_t v) {
  char buf[20];
  int n = 0;
  do {
    buf[n++] = (char) ('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) *p++ = buf[--n];
  return p;
}

char *put_str(char *p, const char *s, size_t n) {
  memcpy(p, s, n);
  return p + n;
}

struct Block {
  std::vector<Record> a, b;
  size_t frags = 0;   // fragments in this block
  size_t first = 0;   // index of its first fragment in the GPU batch
};

// At most kPieceFrags fragments of one block: the unit of the parallel
// packing and formatting, with its own output text and statistics (written
// out in input order).
constexpr size_t kPieceFrags = 512;
struct Piece {
  size_t block = 0, lo = 0, hi = 0;
  Text kraken;
  std::string c1, c2, u1, u2;
  Stats stats;
};

// One GPU batch in flight: its blocks, its packed bases (pinned, kept across
// batches), its library batch and the results of its last call.  Two slots
// alternate, so one batch's GPU call overlaps the formatting of the batch
// before it and the packing of the batch after it.
struct Slot {
  std::vector<Block> blocks;
  std::vector<Piece> pieces;  // the first npieces are this batch's
  size_t npieces = 0;
  k2g_batch *batch = nullptr;
  char *seq = nullptr;
  size_t seq_cap = 0;
  std::vector<uint64_t> off;
  size_t nfrag = 0;
  k2g_results res{};
  int rc = 0;
  double gpu_s = 0;
};

struct Ctx {
  Options o;
  k2host::Taxonomy tax;
  k2g_db *db = nullptr;
  Slot slots[2];
  bool translated = false;
  Stats stats;
  k2host::TaxonCounters counters;
  std::vector<uint64_t> kmer_totals;
  // K2G_TIMING=1: seconds waiting for parsed input, packing, in GPU calls
  // (k2g_classify, on the GPU thread), formatting and writing, and counters
  bool timing = false;
  double t_wait = 0, t_pack = 0, t_gpu = 0, t_format = 0, t_count = 0;
  double t_alloc = 0, t_offsets = 0;  // parts of t_pack
  // reader thread: parsing chunks, all of its work, and waiting for queue room
  double t_rparse = 0, t_rbusy = 0, t_rfull = 0;
  double t_gpu_wait = 0;  // main thread waiting for a GPU call to finish
};

double now_s() {
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec / 1e6;
}

// One kraken output line (src/classify.cc:908-945, 950-991).
void format_line(const Ctx &c, const Record &s1, const Record *s2, uint32_t call,
                 const uint32_t *runs, uint64_t nruns, Text &out) {
  const auto &tax = c.tax;
  const std::string &id = s1.header;
  size_t id_len = id.size();
  // a pair's id without its /1 or /2 (src/classify.cc:764-771)
  if (s2 && id_len > 2 && id[id_len - 2] == '/' && (id[id_len - 1] == '1' || id[id_len - 1] == '2'))
    id_len -= 2;
  const uint64_t ext = tax.nodes[call].external_id;
  const char *name = nullptr;
  size_t name_len = 0;
  if (c.o.print_scientific_name) {
    name = call ? tax.name(call) : "unclassified";
    name_len = strlen(name);
  }
  // every field's longest text: 20 digits per number, 32 bytes per run
  char *p = out.reserve(id_len + name_len + 128 + 32 * nruns);
  p = put_str(p, call ? "C\t" : "U\t", 2);
  p = put_str(p, id.data(), id_len);
  *p++ = '\t';
  if (name) {
    p = put_str(p, name, name_len);
    p = put_str(p, " (taxid ", 8);
    p = put_uint(p, ext);
    *p++ = ')';
  } else {
    p = put_uint(p, ext);
  }
  *p++ = '\t';
  p = put_uint(p, s1.seq.size());
  if (s2) {
    *p++ = '|';
    p = put_uint(p, s2->seq.size());
  }
  *p++ = '\t';
  if (c.o.quick_mode) {
    p = put_uint(p, ext);
    p = put_str(p, ":Q", 2);
  } else if (nruns == 0) {
    p = put_str(p, "0:0", 3);
  } else {
    for (uint64_t i = 0; i < nruns; i++) {
      const uint32_t code = runs[2 * i], cnt = runs[2 * i + 1];
      const bool last = i + 1 == nruns;
      if (code == K2G_CODE_MATE) {
        p = put_str(p, "|:| ", last ? 3 : 4);
      } else if (code == K2G_CODE_FRAME) {
        p = put_str(p, "-:- ", last ? 3 : 4);
      } else if (code == K2G_CODE_AMBIG) {
        p = put_str(p, "A:", 2);
        p = put_uint(p, cnt);
        *p++ = ' ';
      } else {
        p = put_uint(p, tax.nodes[code].external_id);
        *p++ = ':';
        p = put_uint(p, cnt);
        if (!last) *p++ = ' ';
      }
    }
  }
  *p++ = '\n';
  out.commit(p);
}

// The record pair of fragment i of a block.
void fragment_records(const Options &o, Block &b, size_t i, Record *&r1, Record *&r2) {
  r2 = nullptr;
  if (!o.paired_end_processing) {
    r1 = &b.a[i];
  } else if (o.single_file_pairs) {
    r1 = &b.a[2 * i];
    r2 = &b.a[2 * i + 1];
  } else {
    r1 = &b.a[i];
    r2 = &b.b[i];
  }
}

// Offsets of the slot's fragments, then each block's mates copied into the
// slot's pinned buffer (kept across batches: no page faults after the first
// batch, and the upload reads pinned memory).
void pack_batch(Ctx &c, Slot &sl) {
  const Options &o = c.o;
  const bool paired = o.paired_end_processing;
  double t = now_s();
  size_t nfrag = 0, bases = 0;
  sl.npieces = 0;
  for (size_t bi = 0; bi < sl.blocks.size(); bi++) {
    Block &b = sl.blocks[bi];
    b.first = nfrag;
    if (!paired)
      b.frags = b.a.size();
    else if (o.single_file_pairs)
      b.frags = b.a.size() / 2;
    else
      b.frags = std::min(b.a.size(), b.b.size());
    nfrag += b.frags;
    for (auto &r : b.a) bases += r.seq.size();
    for (auto &r : b.b) bases += r.seq.size();
    for (size_t lo = 0; lo < b.frags; lo += kPieceFrags) {
      if (sl.npieces == sl.pieces.size()) sl.pieces.emplace_back();
      Piece &pc = sl.pieces[sl.npieces++];
      pc.block = bi;
      pc.lo = lo;
      pc.hi = std::min(lo + kPieceFrags, b.frags);
    }
  }
  double ta = now_s();
  if (bases + 1 > sl.seq_cap) {
    if (sl.seq) k2g_host_free(sl.seq);
    sl.seq = nullptr;
    sl.seq_cap = std::max<size_t>(bases + 1, sl.seq_cap + sl.seq_cap / 2);
    const int rc = k2g_host_alloc(sl.seq_cap, (void **) &sl.seq);
    if (rc) die(rc, k2g_last_error());
  }
  double tb = now_s();
  c.t_alloc += tb - ta;
  std::vector<uint64_t> &off = sl.off;
  off.resize(2 * nfrag + 1);
  uint64_t pos = 0;
  for (auto &b : sl.blocks) {
    for (size_t i = 0; i < b.frags; i++) {
      const size_t f = b.first + i;
      Record *r1, *r2;
      fragment_records(o, b, i, r1, r2);
      off[2 * f] = pos;
      pos += r1->seq.size();
      off[2 * f + 1] = pos;
      if (r2) pos += r2->seq.size();
    }
  }
  off[2 * nfrag] = pos;
  ta = now_s();
  c.t_offsets += ta - tb;
  parallel_for(sl.npieces, o.num_threads, [&](size_t pi) {
    const Piece &pc = sl.pieces[pi];
    Block &b = sl.blocks[pc.block];
    for (size_t i = pc.lo; i < pc.hi; i++) {
      const size_t f = b.first + i;
      Record *r1, *r2;
      fragment_records(o, b, i, r1, r2);
      if (o.minimum_quality_score > 0) {
        mask_low_quality(*r1, o.minimum_quality_score);
        if (r2) mask_low_quality(*r2, o.minimum_quality_score);
      }
      std::memcpy(sl.seq + off[2 * f], r1->seq.data(), r1->seq.size());
      if (r2) std::memcpy(sl.seq + off[2 * f + 1], r2->seq.data(), r2->seq.size());
    }
  });
  sl.nfrag = nfrag;
  c.t_pack += now_s() - t;
}

// The slot's GPU call (k2g_classify: upload, classify, results into host
// memory); runs on the GPU thread, so its errors are kept for the main thread.
void gpu_batch(const Ctx &c, Slot &sl) {
  const Options &o = c.o;
  k2g_classify_params p{};
  p.confidence = o.confidence_threshold;
  p.min_hit_groups = o.minimum_hit_groups;
  p.quick = o.quick_mode;
  p.paired = o.paired_end_processing;
  p.mode = o.gpu_mode;
  p.want_kmer_counts = !o.report_filename.empty();
  p.want_hll = !o.report_filename.empty() && o.report_kmer_data;
  k2g_fragments fr{sl.seq, sl.off.data(), nullptr, sl.nfrag};
  const double t = now_s();
  sl.rc = k2g_classify(sl.batch, &p, &fr, &sl.res);
  sl.gpu_s = now_s() - t;
}

// Output lines of the slot's blocks (formatted in parallel, written in
// order), statistics and counters.
void finish_batch(Ctx &c, Slot &sl, Outputs &out, const std::string &gpu_error) {
  const Options &o = c.o;
  if (sl.rc) die(sl.rc, gpu_error);
  c.t_gpu += sl.gpu_s;
  const k2g_results &res = sl.res;
  double t = now_s();
  parallel_for(sl.npieces, o.num_threads, [&](size_t pi) {
    Piece &pc = sl.pieces[pi];
    Block &b = sl.blocks[pc.block];
    pc.kraken.n = 0;
    pc.c1.clear();
    pc.c2.clear();
    pc.u1.clear();
    pc.u2.clear();
    pc.stats = Stats();
    for (size_t i = pc.lo; i < pc.hi; i++) {
      const size_t g = b.first + i;
      Record *r1, *r2;
      fragment_records(o, b, i, r1, r2);
      const uint32_t call = res.call[g];
      pc.stats.total_sequences++;
      if (call) pc.stats.total_classified++;
      pc.stats.total_bases += r1->seq.size() + (r2 ? r2->seq.size() : 0);
      if (out.kraken)
        format_line(c, *r1, r2, call, res.runs + 2 * res.run_off[g], res.run_off[g + 1] - res.run_off[g],
                    pc.kraken);
      if (call) {
        if (out.c1) {
          char suffix[48];
          snprintf(suffix, sizeof(suffix), " kraken:taxid|%llu",
                   (unsigned long long) c.tax.nodes[call].external_id);
          r1->append_text(pc.c1, suffix);
          if (r2) r2->append_text(pc.c2, suffix);
        }
      } else if (out.u1) {
        r1->append_text(pc.u1, nullptr);
        if (r2) r2->append_text(pc.u2, nullptr);
      }
    }
  });
  for (size_t pi = 0; pi < sl.npieces; pi++) {
    const Piece &pc = sl.pieces[pi];
    if (out.kraken) fwrite(pc.kraken.v.data(), 1, pc.kraken.n, out.kraken);
    if (out.c1) fwrite(pc.c1.data(), 1, pc.c1.size(), out.c1);
    if (out.c2) fwrite(pc.c2.data(), 1, pc.c2.size(), out.c2);
    if (out.u1) fwrite(pc.u1.data(), 1, pc.u1.size(), out.u1);
    if (out.u2) fwrite(pc.u2.data(), 1, pc.u2.size(), out.u2);
    c.stats.total_sequences += pc.stats.total_sequences;
    c.stats.total_bases += pc.stats.total_bases;
    c.stats.total_classified += pc.stats.total_classified;
  }
  double t2 = now_s();
  c.t_format += t2 - t;
  if (!o.report_filename.empty()) {
    // read counts per called taxon; k-mer counters per block (HLL state
    // depends on the block a hit falls in, src/classify.cc:554, 640-645)
    for (auto &b : sl.blocks)
      for (size_t i = 0; i < b.frags; i++)
        if (res.call[b.first + i]) c.counters[res.call[b.first + i]].n_reads++;
    if (o.report_kmer_data) {
      std::vector<std::pair<uint32_t, uint32_t>> recs;
      for (auto &b : sl.blocks) {
        recs.clear();
        for (size_t i = 0; i < b.frags; i++) {
          const size_t g = b.first + i;
          for (uint64_t j = res.hll_off[g]; j < res.hll_off[g + 1]; j++)
            recs.emplace_back(res.hll[2 * j], res.hll[2 * j + 1]);
        }
        k2host::fold_block_records(recs, c.counters);
      }
    } else {
      for (uint64_t t = 0; t < res.node_count; t++)
        if (res.kmer_counts[t]) c.counters[t].n_kmers += res.kmer_counts[t];
    }
  }
  c.t_count += now_s() - t2;
}

// Reads the input in the reference's blocks (src/classify.cc:523-543) on a
// reader thread, batches of blocks handed to the caller through a two-deep
// queue, so parsing the next batch overlaps the GPU and the formatting of
// this one.  Batches, blocks and output stay in input order.
int process_files(Ctx &c, const char *f1, const char *f2, Outputs &out) {
  const Options &o = c.o;
  k2host::Reader r1, r2;
  if (!r1.open(f1)) die(66, std::string("unable to open ") + (f1 ? f1 : "stdin"));
  if (f2 && !r2.open(f2)) die(66, std::string("unable to open ") + f2);
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::vector<Block>> ready;
  std::deque<std::vector<Block>> spare;  // consumed batches, their records reused
  bool finished = false;
  std::thread reader([&] {
    bool done = false;
    // Each input file's records, parsed a chunk at a time (side by side on
    // the host threads for plain FASTA), then cut into the reference's
    // blocks; `used` holds records of consumed blocks, whose buffers the
    // next chunk's parse reuses.  Chunks start at 4 MiB, so the first block
    // is ready early, and double up to 64 MiB.
    struct Stream {
      k2host::Reader *r;
      std::vector<Record> parsed, used;
      size_t next = 0;
      size_t chunk = 4u << 20;
    };
    Stream s1{&r1}, s2{&r2};
    auto take = [&](Stream &s, std::vector<Record> &into) -> bool {
      if (s.next == s.parsed.size()) {
        s.parsed.clear();
        s.next = 0;
        const double t0 = now_s();
        const size_t got = s.r->next_records(s.chunk, o.num_threads, s.parsed, &s.used);
        s.chunk = std::min<size_t>(2 * s.chunk, 64u << 20);
        c.t_rparse += now_s() - t0;
        if (!got) return false;
      }
      into.push_back(std::move(s.parsed[s.next++]));
      return true;
    };
    auto take_n = [&](Stream &s, size_t n, std::vector<Record> &into) {
      for (auto &r : into) s.used.push_back(std::move(r));
      into.clear();
      while (into.size() < n && take(s, into)) {
      }
    };
    // The first batches are small, so the GPU starts early; each later one
    // doubles up to batch_bases.
    size_t target = std::min<size_t>(o.batch_bases, 16u << 20);
    while (!done) {
      const double t_batch = now_s();
      std::vector<Block> blocks;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!spare.empty()) {
          blocks.swap(spare.front());
          spare.pop_front();
        }
      }
      size_t bases = 0, used = 0;
      while (bases < target && used < 4096) {
        if (used == blocks.size()) blocks.emplace_back();
        Block &b = blocks[used];
        bool ok;
        if (!o.paired_end_processing) {
          // load_block's rule (src/seqreader.h:101-119): records until
          // 3 MiB of bases
          for (auto &r : b.a) s1.used.push_back(std::move(r));
          b.a.clear();
          size_t got = 0;
          while (got < 3 * 1024 * 1024) {
            if (!take(s1, b.a)) break;
            got += b.a.back().seq.size();
          }
          ok = !b.a.empty();
        } else if (!o.single_file_pairs) {
          // load_batch's rule (src/seqreader.h:146-163) on each file
          take_n(s1, 10000, b.a);
          take_n(s2, 10000, b.b);
          ok = !b.a.empty() && !b.b.empty();
        } else {
          take_n(s1, 20000, b.a);
          ok = !b.a.empty();
        }
        if (!ok) {
          done = true;
          break;
        }
        used++;
        for (auto &r : b.a) bases += r.seq.size();
        for (auto &r : b.b) bases += r.seq.size();
      }
      target = std::min(o.batch_bases, 2 * target);
      blocks.resize(used);
      std::unique_lock<std::mutex> lk(mu);
      const double tw = now_s();
      c.t_rbusy += tw - t_batch;
      cv.wait(lk, [&] { return ready.size() < 2; });
      c.t_rfull += now_s() - tw;
      if (!blocks.empty()) ready.push_back(std::move(blocks));
      if (done) finished = true;
      cv.notify_all();
    }
  });
  // Batch i is packed while batch i-1 is on the GPU; then batch i goes to
  // the GPU thread while batch i-1 is formatted and written, in input order.
  std::thread gpu;
  std::string gpu_error;
  int cur = 0;
  bool prev = false;  // slots[cur ^ 1] holds a batch whose GPU call was started
  auto join_gpu = [&] {
    if (!gpu.joinable()) return;
    const double t = now_s();
    gpu.join();
    c.t_gpu_wait += now_s() - t;
  };
  auto recycle = [&](Slot &sl) {
    std::lock_guard<std::mutex> lk(mu);
    if (spare.size() < 2) spare.push_back(std::move(sl.blocks));
    sl.blocks.clear();
  };
  for (;;) {
    std::vector<Block> blocks;
    {
      const double t = now_s();
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !ready.empty() || finished; });
      c.t_wait += now_s() - t;
      if (ready.empty()) break;
      blocks = std::move(ready.front());
      ready.pop_front();
      cv.notify_all();
    }
    Slot &sl = c.slots[cur];
    sl.blocks = std::move(blocks);
    pack_batch(c, sl);
    join_gpu();
    gpu = std::thread([&c, &sl, &gpu_error] {
      gpu_batch(c, sl);
      if (sl.rc) gpu_error = k2g_last_error();
    });
    if (prev) {
      finish_batch(c, c.slots[cur ^ 1], out, gpu_error);
      recycle(c.slots[cur ^ 1]);
    }
    prev = true;
    cur ^= 1;
  }
  join_gpu();
  if (prev) {
    finish_batch(c, c.slots[cur ^ 1], out, gpu_error);
    recycle(c.slots[cur ^ 1]);
  }
  reader.join();
  return 0;
}

void report_stats(double seconds, const Stats &s) {  // src/classify.cc:447-477
  const uint64_t unc = s.total_sequences - s.total_classified;
  if (isatty(fileno(stderr))) std::cerr << "\r";
  fprintf(stderr, "%llu sequences (%.2f Mbp) processed in %.3fs (%.1f Kseq/m, %.2f Mbp/m).\n",
          (unsigned long long) s.total_sequences, s.total_bases / 1.0e6, seconds,
          s.total_sequences / 1.0e3 / (seconds / 60), s.total_bases / 1.0e6 / (seconds / 60));
  fprintf(stderr, "  %llu sequences classified (%.2f%%)\n", (unsigned long long) s.total_classified,
          s.total_classified * 100.0 / s.total_sequences);
  fprintf(stderr, "  %llu sequences unclassified (%.2f%%)\n", (unsigned long long) unc,
          unc * 100.0 / s.total_sequences);
}

}  // namespace

int main(int argc, char **argv) {
  // records' sequence buffers are allocated by the parser threads and
  // recycled by the main thread: keep freed memory in the heap instead of
  // returning it to the system (and faulting it in again for the next
  // chunk), measured 1.4x faster parsing
  mallopt(M_TRIM_THRESHOLD, 1 << 30);
  mallopt(M_MMAP_THRESHOLD, 1 << 30);
  mallopt(M_TOP_PAD, 256 << 20);
  Ctx c;
  parse(argc, argv, c.o);
  if (const char *t = getenv("K2G_TIMING")) c.timing = atoi(t) != 0;
  std::cerr << "Loading database information...";
  std::string err;
  if (!c.tax.load(c.o.taxonomy_filename, err)) die(65, err);
  int rc = k2g_db_open(c.o.index_filename.c_str(), c.o.taxonomy_filename.c_str(),
                       c.o.options_filename.c_str(), &c.db);
  if (rc) {
    std::cerr << "\n";
    die(rc, k2g_last_error());
  }
  k2g_db_info info;
  k2g_db_get_info(c.db, &info);
  c.translated = !info.dna_db;
  for (Slot &sl : c.slots) {
    rc = k2g_batch_create(c.db, &sl.batch);
    if (rc) die(rc, k2g_last_error());
  }
  std::cerr << " done." << std::endl;

  Outputs out;
  open_outputs(c.o, out);
  struct timeval t1, t2;
  gettimeofday(&t1, nullptr);
  const Options &o = c.o;
  if (o.filenames.empty()) {
    if (o.paired_end_processing && !o.single_file_pairs)
      die(64, "paired end processing used with no files specified");
    process_files(c, nullptr, nullptr, out);
  } else {
    for (size_t i = 0; i < o.filenames.size(); i++) {
      if (o.paired_end_processing && !o.single_file_pairs) {
        if (i + 1 == o.filenames.size()) die(64, "paired end processing used with unpaired file");
        process_files(c, o.filenames[i].c_str(), o.filenames[i + 1].c_str(), out);
        i++;
      } else {
        process_files(c, o.filenames[i].c_str(), nullptr, out);
      }
    }
  }
  for (FILE *f : {out.kraken, out.c1, out.c2, out.u1, out.u2})
    if (f) fflush(f);
  gettimeofday(&t2, nullptr);
  const double secs = (t2.tv_sec - t1.tv_sec) + (t2.tv_usec - t1.tv_usec) / 1e6;
  report_stats(secs, c.stats);
  if (c.timing)
    fprintf(stderr,
            "  timing: input wait %.3fs, packing %.3fs (buffer %.3fs, offsets %.3fs), waiting for the GPU %.3fs "
            "(GPU calls %.3fs, overlapped), format+write %.3fs, counters %.3fs, outside the phases %.3fs; "
            "reader thread: busy %.3fs (chunk reading and parsing %.3fs), waiting for queue room %.3fs\n",
            c.t_wait, c.t_pack, c.t_alloc, c.t_offsets, c.t_gpu_wait, c.t_gpu, c.t_format, c.t_count,
            secs - c.t_wait - c.t_pack - c.t_gpu_wait - c.t_format - c.t_count, c.t_rbusy, c.t_rparse,
            c.t_rfull);
  if (!o.report_filename.empty()) {
    if (!k2host::report_kraken_style(o.report_filename, o.report_zero_counts, o.report_kmer_data,
                                     c.tax, c.counters, c.stats.total_sequences,
                                     c.stats.total_sequences - c.stats.total_classified))
      die(74, "unable to write report " + o.report_filename);
  }
  for (FILE *f : {out.c1, out.c2, out.u1, out.u2})
    if (f) fclose(f);
  if (out.kraken && out.kraken != stdout) fclose(out.kraken);
  for (Slot &sl : c.slots) {
    k2g_batch_destroy(sl.batch);
    if (sl.seq) k2g_host_free(sl.seq);
  }
  k2g_db_close(c.db);
  return 0;
}
