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
#include <memory>
#include <string>
#include <vector>

#include "k2g.h"
#include "k2g_host.h"

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
  // bases per GPU batch (several blocks): small enough that the pinned
  // memory of the batches in flight is touched few times over, large
  // enough that a GPU call's fixed costs stay small
  size_t batch_bases = 64u << 20;
  int gpu_calls = 2;  // K2G_GPU_CALLS: GPU calls in flight at once
  // K2G_PARSE_THREADS / K2G_FORMAT_THREADS: threads of the reader's parse and
  // of the formatting (0: -p); the two run at the same time
  int parse_threads = 0, format_threads = 0;
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
  if (const char *t = getenv("K2G_PARSE_THREADS")) o.parse_threads = std::max(0, atoi(t));
  if (const char *t = getenv("K2G_GPU_CALLS")) o.gpu_calls = std::min(8, std::max(1, atoi(t)));
  if (const char *t = getenv("K2G_FORMAT_THREADS")) o.format_threads = std::max(0, atoi(t));
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

// Bases whose quality is below q become 'x' (src/classify.cc:1061-1071); the
// reader keeps as many qualities as bases.
void mask_low_quality(const k2host::FlatRec &r, const char *quals, char *bases, int q) {
  if (!r.fastq) return;
  for (uint64_t i = 0; i < r.seq_len; i++)
    if ((quals[i] - '!') < q) bases[i] = 'x';
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

// A block of the reference's (src/classify.cc:523-543): the unit of the
// output order and of the per-block HLL state.  Its records are
// recs[rec_first, rec_first + nrec) of the batch, its fragments
// [frag_first, frag_first + frags) (a pair's mates are two records side by
// side; an odd last record of interleaved pairs has no fragment).
struct BlockSpan {
  size_t rec_first = 0, nrec = 0, frag_first = 0, frags = 0;
};

// At most kPieceFrags fragments of one block: the unit of the parallel
// formatting, with its own output text and statistics (written out in
// input order).
constexpr size_t kPieceFrags = 512;
struct Piece {
  size_t block = 0, lo = 0, hi = 0;
  Text kraken;
  std::string c1, c2, u1, u2;
  Stats stats;
};

// The GPU side of a batch: a library batch and the results of its last
// call.  Three slots rotate: two batches' GPU calls overlap each other and
// the formatting of the batch before them.
struct Slot {
  k2g_batch *batch = nullptr;
  k2g_results res{};
  int rc = 0;
  double gpu_s = 0;
  std::string error;
};

// One batch of input: records in fragment order with their bases back to
// back in pinned memory (the GPU input as it stands, mate 1 then mate 2 of
// a pair) and their ids, comments and qualities in `text`, cut into blocks.
struct Batch {
  std::vector<k2host::FlatRec> recs;
  std::vector<char> text;
  k2host::SeqSink seq;
  std::vector<BlockSpan> blocks;
  size_t nfrag = 0, bases = 0;  // fragments and bases of the blocks
  std::vector<uint64_t> off;    // 2 nfrag + 1 fragment offsets into seq
  std::vector<Piece> pieces;    // the first npieces are this batch's
  size_t npieces = 0;
  Slot *slot = nullptr;         // where its GPU call ran
  void clear() {
    recs.clear();
    text.clear();
    seq.n = 0;
    blocks.clear();
    nfrag = bases = 0;
  }
  ~Batch() {
    if (seq.p) k2g_host_free(seq.p);
  }
};

struct Ctx {
  Options o;
  k2host::Taxonomy tax;
  k2g_db *db = nullptr;
  std::vector<Slot> slots;  // gpu_calls + 1
  k2host::WorkPool *pool = nullptr;  // the main thread's workers (formatting)
  bool translated = false;
  Stats stats;
  k2host::TaxonCounters counters;
  std::vector<uint64_t> kmer_totals;
  // K2G_TIMING=1: seconds waiting for parsed input, preparing batches
  // (offsets, masking), waiting for and in GPU calls (k2g_classify, on the
  // GPU thread), formatting and writing, and counters
  bool timing = false;
  double t_wait = 0, t_pack = 0, t_gpu = 0, t_format = 0, t_count = 0;
  // reader thread: parsing chunks, all of its work, and waiting for room
  double t_rparse = 0, t_rbusy = 0, t_rfull = 0;
  k2host::ReadTiming read;  // parts of t_rparse
  double t_ralloc = 0, t_rcopy = 0;  // pinned allocation; copies (carry, pairs)
  double t_gpu_wait = 0;  // main thread waiting for a GPU call to finish
};

double now_s() {
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec / 1e6;
}

// One kraken output line (src/classify.cc:908-945, 950-991).
// id: mate 1's id; len2: mate 2's bases (paired only).
void format_line(const Ctx &c, const char *id, size_t id_len, bool paired, uint64_t len1, uint64_t len2,
                 uint32_t call, const uint32_t *runs, uint64_t nruns, Text &out) {
  const auto &tax = c.tax;
  // a pair's id without its /1 or /2 (src/classify.cc:764-771)
  if (paired && id_len > 2 && id[id_len - 2] == '/' && (id[id_len - 1] == '1' || id[id_len - 1] == '2'))
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
  p = put_str(p, id, id_len);
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
  p = put_uint(p, len1);
  if (paired) {
    *p++ = '|';
    p = put_uint(p, len2);
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

// Record index of mate 1 of fragment i of a block (mate 2 follows it).
size_t mate1(const Options &o, const BlockSpan &b, size_t i) {
  return b.rec_first + (o.paired_end_processing ? 2 * i : i);
}

// One record as FASTA/FASTQ text (src/seqreader.cc:20-49).
void append_record(std::string &out, const Batch &B, const k2host::FlatRec &r, const char *suffix) {
  const char *t = B.text.data() + r.text_at;
  out.push_back(r.fastq ? '@' : '>');
  out.append(t, r.id_len);
  if (suffix) out += suffix;
  if (r.cmt_len) {
    out.push_back(' ');
    out.append(t + r.id_len, r.cmt_len);
  }
  out.push_back('\n');
  out.append(B.seq.p + r.seq_at, r.seq_len);
  if (r.fastq) {
    out += "\n+\n";
    out.append(t + r.id_len + r.cmt_len, r.seq_len);
  }
  out.push_back('\n');
}

// Fragment offsets of the batch (its bases already lie in fragment order),
// pieces for the formatting, and low-quality masking (-Q).
void prepare_batch(Ctx &c, Batch &B) {
  const Options &o = c.o;
  const bool paired = o.paired_end_processing;
  const double t = now_s();
  B.off.resize(2 * B.nfrag + 1);
  B.npieces = 0;
  uint64_t end = 0;
  for (size_t bi = 0; bi < B.blocks.size(); bi++) {
    const BlockSpan &b = B.blocks[bi];
    for (size_t i = 0; i < b.frags; i++) {
      const size_t g = b.frag_first + i;
      const k2host::FlatRec &r1 = B.recs[mate1(o, b, i)];
      B.off[2 * g] = r1.seq_at;
      B.off[2 * g + 1] = r1.seq_at + r1.seq_len;
      end = B.off[2 * g + 1];
      if (paired) end += B.recs[mate1(o, b, i) + 1].seq_len;
    }
    for (size_t lo = 0; lo < b.frags; lo += kPieceFrags) {
      if (B.npieces == B.pieces.size()) B.pieces.emplace_back();
      Piece &pc = B.pieces[B.npieces++];
      pc.block = bi;
      pc.lo = lo;
      pc.hi = std::min(lo + kPieceFrags, b.frags);
    }
  }
  B.off[2 * B.nfrag] = end;
  if (o.minimum_quality_score > 0)
    c.pool->run(B.npieces, [&](size_t pi) {
      const Piece &pc = B.pieces[pi];
      const BlockSpan &b = B.blocks[pc.block];
      for (size_t i = pc.lo; i < pc.hi; i++)
        for (size_t m = 0; m < (paired ? 2u : 1u); m++) {
          const k2host::FlatRec &r = B.recs[mate1(o, b, i) + m];
          mask_low_quality(r, B.text.data() + r.text_at + r.id_len + r.cmt_len, B.seq.p + r.seq_at,
                           o.minimum_quality_score);
        }
    });
  c.t_pack += now_s() - t;
}

// The batch's GPU call (k2g_classify: upload, classify, results into host
// memory); runs on the GPU thread, so its error text is kept in the slot.
void gpu_batch(const Ctx &c, Batch &B, Slot &sl) {
  const Options &o = c.o;
  k2g_classify_params p{};
  p.confidence = o.confidence_threshold;
  p.min_hit_groups = o.minimum_hit_groups;
  p.quick = o.quick_mode;
  p.paired = o.paired_end_processing;
  p.mode = o.gpu_mode;
  p.want_kmer_counts = !o.report_filename.empty();
  p.want_hll = !o.report_filename.empty() && o.report_kmer_data;
  k2g_fragments fr{B.seq.p, B.off.data(), nullptr, B.nfrag};
  const double t = now_s();
  B.slot = &sl;
  sl.rc = k2g_classify(sl.batch, &p, &fr, &sl.res);
  if (sl.rc) sl.error = k2g_last_error();
  sl.gpu_s = now_s() - t;
}

// Output lines of the batch's blocks (formatted in parallel, written in
// order), statistics and counters.
void finish_batch(Ctx &c, Batch &B, Outputs &out) {
  const Options &o = c.o;
  const bool paired = o.paired_end_processing;
  Slot &sl = *B.slot;
  if (sl.rc) die(sl.rc, sl.error);
  c.t_gpu += sl.gpu_s;
  const k2g_results &res = sl.res;
  double t = now_s();
  c.pool->run(B.npieces, [&](size_t pi) {
    Piece &pc = B.pieces[pi];
    const BlockSpan &b = B.blocks[pc.block];
    pc.kraken.n = 0;
    pc.c1.clear();
    pc.c2.clear();
    pc.u1.clear();
    pc.u2.clear();
    pc.stats = Stats();
    for (size_t i = pc.lo; i < pc.hi; i++) {
      const size_t g = b.frag_first + i;
      const k2host::FlatRec &r1 = B.recs[mate1(o, b, i)];
      const k2host::FlatRec *r2 = paired ? &B.recs[mate1(o, b, i) + 1] : nullptr;
      const uint32_t call = res.call[g];
      pc.stats.total_sequences++;
      if (call) pc.stats.total_classified++;
      pc.stats.total_bases += r1.seq_len + (r2 ? r2->seq_len : 0);
      if (out.kraken)
        format_line(c, B.text.data() + r1.text_at, r1.id_len, paired, r1.seq_len, r2 ? r2->seq_len : 0, call,
                    res.runs + 2 * res.run_off[g], res.run_off[g + 1] - res.run_off[g], pc.kraken);
      if (call) {
        if (out.c1) {
          char suffix[48];
          snprintf(suffix, sizeof(suffix), " kraken:taxid|%llu",
                   (unsigned long long) c.tax.nodes[call].external_id);
          append_record(pc.c1, B, r1, suffix);
          if (r2) append_record(pc.c2, B, *r2, suffix);
        }
      } else if (out.u1) {
        append_record(pc.u1, B, r1, nullptr);
        if (r2) append_record(pc.u2, B, *r2, nullptr);
      }
    }
  });
  for (size_t pi = 0; pi < B.npieces; pi++) {
    const Piece &pc = B.pieces[pi];
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
    for (size_t g = 0; g < B.nfrag; g++)
      if (res.call[g]) c.counters[res.call[g]].n_reads++;
    if (o.report_kmer_data) {
      std::vector<std::pair<uint32_t, uint32_t>> recs;
      for (const BlockSpan &b : B.blocks) {
        recs.clear();
        for (size_t g = b.frag_first; g < b.frag_first + b.frags; g++)
          for (uint64_t j = res.hll_off[g]; j < res.hll_off[g + 1]; j++)
            recs.emplace_back(res.hll[2 * j], res.hll[2 * j + 1]);
        k2host::fold_block_records(recs, c.counters);
      }
    } else {
      for (uint64_t t = 0; t < res.node_count; t++)
        if (res.kmer_counts[t]) c.counters[t].n_kmers += res.kmer_counts[t];
    }
  }
  c.t_count += now_s() - t2;
}

// Pinned memory for a batch's bases (grown by copying what it holds).
void grow_pinned(k2host::SeqSink &s, size_t need) {
  const size_t cap = std::max(need, s.cap + s.cap / 2);
  char *p = nullptr;
  const int rc = k2g_host_alloc(cap, (void **) &p);
  if (rc) die(rc, k2g_last_error());
  if (s.n) memcpy(p, s.p, s.n);
  if (s.p) k2g_host_free(s.p);
  s.p = p;
  s.cap = cap;
}

// Records [r0, end) of `from` become the first records of `to` (empty),
// with the bytes of `from` from the first one's bases on.
void carry_records(Batch &from, size_t r0, Batch &to) {
  if (r0 == from.recs.size()) return;
  const uint64_t s0 = from.recs[r0].seq_at, t0 = from.recs[r0].text_at;
  const size_t sn = from.seq.n - s0;
  if (sn) memcpy(to.seq.room(sn), from.seq.p + s0, sn);
  to.seq.n = sn;
  to.text.assign(from.text.begin() + t0, from.text.end());
  for (size_t r = r0; r < from.recs.size(); r++) {
    k2host::FlatRec f = from.recs[r];
    f.seq_at -= s0;
    f.text_at -= t0;
    to.recs.push_back(f);
  }
  from.recs.resize(r0);
  from.text.resize(t0);
  from.seq.n = s0;
}

// Records of one input file parsed a chunk at a time into its own flat
// arrays, handed out in order (two-file pairs, whose mates are copied next
// to each other into the batch).
struct Stage {
  k2host::Reader *r = nullptr;
  std::vector<k2host::FlatRec> recs;
  std::vector<char> text;
  k2host::ByteBuf bases;
  k2host::SeqSink seq;
  size_t next = 0;
  size_t chunk = 4u << 20;
  bool eof = false;
  Stage() {
    seq.grow = [this](k2host::SeqSink &s, size_t need) {
      bases.resize(std::max(need, 2 * bases.size()), s.n);
      s.p = bases.data();
      s.cap = bases.size();
    };
  }
};

// Drops a stage's consumed records, moving the rest (and their text and
// bases) to the front.
void compact_stage(Stage &s) {
  if (s.next == 0) return;
  if (s.next < s.recs.size()) {
    const uint64_t s0 = s.recs[s.next].seq_at, t0 = s.recs[s.next].text_at;
    memmove(s.seq.p, s.seq.p + s0, s.seq.n - s0);
    s.seq.n -= s0;
    s.text.erase(s.text.begin(), s.text.begin() + t0);
    s.recs.erase(s.recs.begin(), s.recs.begin() + s.next);
    for (auto &r : s.recs) {
      r.seq_at -= s0;
      r.text_at -= t0;
    }
  } else {
    s.recs.clear();
    s.text.clear();
    s.seq.n = 0;
  }
  s.next = 0;
}

// Reads the input in the reference's blocks (src/classify.cc:523-543) on a
// reader thread into batches of whole blocks (a two-deep queue), so parsing
// the next batch overlaps the GPU call and the formatting of this one.
// Single reads and interleaved pairs are parsed straight into the batch's
// pinned memory; two-file pairs are parsed per file and their mates copied
// in side by side.  Batches, blocks and output stay in input order.
int process_files(Ctx &c, const char *f1, const char *f2, Outputs &out) {
  const Options &o = c.o;
  const bool paired = o.paired_end_processing, two_files = paired && !o.single_file_pairs;
  k2host::Reader r1, r2;
  if (!r1.open(f1)) die(66, std::string("unable to open ") + (f1 ? f1 : "stdin"));
  if (f2 && !r2.open(f2)) die(66, std::string("unable to open ") + f2);
  const int parse_threads = o.parse_threads ? o.parse_threads : o.num_threads;
  k2host::WorkPool parse_pool(parse_threads);
  r1.set_pool(&parse_pool);
  r2.set_pool(&parse_pool);
  r1.reserve_chunk(64u << 20);
  if (f2) r2.reserve_chunk(64u << 20);
  // FASTQ qualities are kept only for -Q masking and the -C/-U records
  const bool quals = o.minimum_quality_score > 0 || !o.classified_output_filename.empty() ||
                     !o.unclassified_output_filename.empty();
  r1.keep_qualities(quals);
  r2.keep_qualities(quals);
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Batch *> ready, spare;  // filled batches; consumed ones, reused
  std::vector<std::unique_ptr<Batch>> owned;
  const size_t kMaxBatches = c.slots.size() + 3;  // one filling, two ready, the slots'
  bool finished = false;
  const size_t kBlockBases = 3 * 1024 * 1024, kPairBlock = 10000;
  std::thread reader([&] {
    double t_batch = now_s();
    auto get_batch = [&]() -> Batch * {
      std::unique_lock<std::mutex> lk(mu);
      if (spare.empty() && owned.size() < kMaxBatches) {
        owned.emplace_back(new Batch());
        owned.back()->seq.grow = [&c](k2host::SeqSink &sk, size_t need) {
          const double t0 = now_s();
          grow_pinned(sk, need);
          c.t_ralloc += now_s() - t0;
        };
        return owned.back().get();
      }
      const double tw = now_s();
      cv.wait(lk, [&] { return !spare.empty(); });
      c.t_rfull += now_s() - tw;
      Batch *B = spare.front();
      spare.pop_front();
      B->clear();
      return B;
    };
    auto ship = [&](Batch *B, bool last) {
      std::unique_lock<std::mutex> lk(mu);
      const double tw = now_s();
      c.t_rbusy += tw - t_batch;
      cv.wait(lk, [&] { return ready.size() < 2; });
      c.t_rfull += now_s() - tw;
      t_batch = now_s();
      if (B && !B->blocks.empty()) ready.push_back(B);
      else if (B) spare.push_back(B);
      if (last) finished = true;
      cv.notify_all();
    };
    // Single reads keep the file's layout (the header text between them is
    // the single reads' unread mate 2 range); pairs are copied mate next to
    // mate.
    const bool in_place = !paired;
    auto parse_into = [&](k2host::Reader &r, size_t chunk, std::vector<k2host::FlatRec> &recs,
                          std::vector<char> &text, k2host::SeqSink &seq) {
      const double t0 = now_s();
      const size_t got = in_place ? r.next_inplace(chunk, parse_threads, recs, text, seq)
                                  : r.next_flat(chunk, parse_threads, recs, text, seq);
      c.t_rparse += now_s() - t0;
      return got;
    };
    // The first batches are small, so the GPU starts early; each later one
    // doubles up to batch_bases.  Parse chunks grow from 4 MiB to 64 MiB.
    size_t target = std::min<size_t>(o.batch_bases, 16u << 20);
    size_t chunk = 4u << 20;
    const size_t kMaxChunk = 64u << 20;
    // pinned memory costs most on first touch (its pages are mapped then):
    // each batch reserves what it will hold, the chunk sizes keep to that
    auto reserve = [&](Batch *X) { X->seq.room(target + target / 8 + (8u << 20)); };
    Batch *B = get_batch();
    reserve(B);
    auto close_block = [&](size_t rec_first, size_t nrec, size_t frags, uint64_t bases) {
      B->blocks.push_back({rec_first, nrec, B->nfrag, frags});
      B->nfrag += frags;
      B->bases += bases;
    };
    auto full = [&] { return B->bases >= target || B->blocks.size() >= 4096; };
    auto next_batch = [&](size_t carry_from) {
      Batch *N = get_batch();
      target = std::min(o.batch_bases, 2 * target);
      reserve(N);
      const double t0 = now_s();
      carry_records(*B, carry_from, *N);
      c.t_rcopy += now_s() - t0;
      ship(B, false);
      B = N;
    };
    if (!two_files) {
      // single reads: records until 3 MiB of bases (load_block,
      // src/seqreader.h:101-119); interleaved pairs: 20 000 records
      // (load_batch, src/seqreader.h:146-163), one fragment per two
      size_t open = 0, scan = 0;  // first record of the open block; records scanned
      uint64_t open_bases = 0;
      bool eof = false;
      for (;;) {
        bool shipped = false;
        while (scan < B->recs.size()) {
          open_bases += B->recs[scan].seq_len;
          scan++;
          const bool done = paired ? scan - open == 2 * kPairBlock : open_bases >= kBlockBases;
          if (!done) continue;
          close_block(open, scan - open, paired ? (scan - open) / 2 : scan - open, open_bases);
          open = scan;
          open_bases = 0;
          if (full()) {
            next_batch(open);
            open = scan = 0;
            shipped = true;
            break;
          }
        }
        if (shipped) continue;
        if (eof) break;
        // about the bases the batch still needs (and a block's worth), so
        // little is parsed past the batch's last block and carried over
        const size_t need = target > B->seq.n ? target - B->seq.n : 0;
        const size_t want = std::min(chunk, std::max<size_t>(1u << 20, need + need / 16 + kBlockBases));
        if (!parse_into(r1, want, B->recs, B->text, B->seq)) {
          eof = true;
          if (open < B->recs.size()) {
            close_block(open, B->recs.size() - open, paired ? (B->recs.size() - open) / 2 : B->recs.size() - open,
                        open_bases);
            open = scan = B->recs.size();
          }
          break;
        }
        chunk = std::min(2 * chunk, kMaxChunk);
      }
    } else {
      // two files: 10 000 records from each (load_batch on both,
      // src/classify.cc:528-532) make min(n1, n2) pairs; the mates are
      // copied side by side into the batch, pieces of a block in parallel
      Stage s1, s2;
      s1.r = &r1;
      s2.r = &r2;
      // at least k records ahead of s.next, or all that are left
      auto ensure = [&](Stage &s, size_t k) {
        while (!s.eof && s.recs.size() - s.next < k) {
          const double t0 = now_s();
          compact_stage(s);
          c.t_rcopy += now_s() - t0;
          if (!parse_into(*s.r, s.chunk, s.recs, s.text, s.seq))
            s.eof = true;
          else
            s.chunk = std::min(2 * s.chunk, kMaxChunk);
        }
        return std::min(k, s.recs.size() - s.next);
      };
      std::vector<uint64_t> tat, sat;  // text and bases of each pair, then their offsets
      for (;;) {
        const size_t n1 = ensure(s1, kPairBlock);
        const size_t n2 = n1 ? ensure(s2, kPairBlock) : 0;
        if (!n1 || !n2) break;
        const double t0 = now_s();
        // every whole block both files hold now (10 000 pairs each) goes in
        // one parallel copy, up to the block that fills the batch; else the
        // one block that is left, of min(n1, n2) pairs
        const bool whole = n1 == kPairBlock && n2 == kPairBlock;
        size_t n = whole ? std::min(s1.recs.size() - s1.next, s2.recs.size() - s2.next) / kPairBlock * kPairBlock
                         : std::min(n1, n2);
        const size_t first = B->recs.size();
        // a FASTQ record's text holds its qualities only when the readers keep them
        auto text_len = [quals](const k2host::FlatRec &r) {
          return (uint64_t) r.id_len + r.cmt_len + (r.fastq && quals ? r.seq_len : 0);
        };
        tat.resize(n + 1);
        sat.resize(n + 1);
        tat[0] = B->text.size();
        sat[0] = B->seq.n;
        for (size_t k = 0; k < n; k++) {
          const k2host::FlatRec &a = s1.recs[s1.next + k], &b = s2.recs[s2.next + k];
          tat[k + 1] = tat[k] + text_len(a) + text_len(b);
          sat[k + 1] = sat[k] + a.seq_len + b.seq_len;
        }
        size_t blocks = 1;
        if (whole) {
          blocks = 0;
          uint64_t bases = B->bases;
          while (blocks * kPairBlock < n) {
            bases += sat[(blocks + 1) * kPairBlock] - sat[blocks * kPairBlock];
            blocks++;
            if (bases >= target || B->blocks.size() + blocks >= 4096) break;
          }
          n = blocks * kPairBlock;
        }
        B->recs.resize(first + 2 * n);
        B->text.resize(tat[n]);
        B->seq.room(sat[n] - sat[0]);
        const size_t pieces = (n + kPieceFrags - 1) / kPieceFrags;
        parse_pool.run(pieces, [&](size_t pi) {
          for (size_t k = pi * kPieceFrags; k < std::min(n, (pi + 1) * kPieceFrags); k++) {
            uint64_t ta = tat[k], sa = sat[k];
            for (const Stage *st : {&s1, &s2}) {
              const k2host::FlatRec &r = st->recs[st->next + k];
              k2host::FlatRec &f = B->recs[first + 2 * k + (st == &s2)];
              f = r;
              f.text_at = ta;
              f.seq_at = sa;
              const uint64_t tl = text_len(r);
              memcpy(B->text.data() + ta, st->text.data() + r.text_at, tl);
              if (r.seq_len) memcpy(B->seq.p + sa, st->seq.p + r.seq_at, r.seq_len);
              ta += tl;
              sa += r.seq_len;
            }
          }
        });
        B->seq.n = sat[n];
        // records past the shorter file's end are dropped with the last block
        s1.next += whole ? n : n1;
        s2.next += whole ? n : n2;
        c.t_rcopy += now_s() - t0;
        for (size_t bk = 0; bk < blocks; bk++) {
          const size_t lo = whole ? bk * kPairBlock : 0, hi = whole ? lo + kPairBlock : n;
          close_block(first + 2 * lo, 2 * (hi - lo), hi - lo, sat[hi] - sat[lo]);
        }
        if (full()) next_batch(B->recs.size());
      }
    }
    ship(B, true);
  });
  // Batch i goes to the GPU (its own thread and slot) while the batches
  // before it may still be there (gpu_calls in flight, so uploads overlap
  // other batches' kernels); then the oldest one is formatted and written.
  // Batches finish in input order.
  const int kSlots = (int) c.slots.size();
  std::vector<std::thread> gpu(kSlots);
  std::vector<Batch *> inflight(kSlots, nullptr);
  auto finish = [&](int k) {
    if (!inflight[k]) return;
    const double t = now_s();
    gpu[k].join();
    c.t_gpu_wait += now_s() - t;
    finish_batch(c, *inflight[k], out);
    std::lock_guard<std::mutex> lk(mu);
    spare.push_back(inflight[k]);
    inflight[k] = nullptr;
    cv.notify_all();
  };
  size_t i = 0;
  for (;; i++) {
    Batch *B;
    {
      const double t = now_s();
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !ready.empty() || finished; });
      c.t_wait += now_s() - t;
      if (ready.empty()) break;
      B = ready.front();
      ready.pop_front();
      cv.notify_all();
    }
    prepare_batch(c, *B);
    const int k = (int) (i % kSlots);  // batch i - kSlots's slot, finished already
    inflight[k] = B;
    Slot &sl = c.slots[k];
    gpu[k] = std::thread([&c, B, &sl] { gpu_batch(c, *B, sl); });
    finish((int) ((i + 1) % kSlots));  // batch i - kSlots + 1
  }
  for (int j = 1; j < kSlots; j++) finish((int) ((i + j) % kSlots));  // the rest, oldest first
  reader.join();
  for (const k2host::Reader *r : {&r1, &r2}) {
    c.read.fill += r->timing.fill;
    c.read.count += r->timing.count;
    c.read.write += r->timing.write;
  }
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
  // input chunks, parse arrays and output text grow and shrink batch after
  // batch: keep freed memory in the heap instead of returning it to the
  // system (and faulting it in again for the next batch)
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
  c.slots.resize(c.o.gpu_calls + 1);
  for (Slot &sl : c.slots) {
    rc = k2g_batch_create(c.db, &sl.batch);
    if (rc) die(rc, k2g_last_error());
  }
  std::cerr << " done." << std::endl;

  k2host::WorkPool pool(c.o.format_threads ? c.o.format_threads : c.o.num_threads);
  c.pool = &pool;
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
            "  timing: input wait %.3fs, preparing batches %.3fs, waiting for the GPU %.3fs (GPU calls %.3fs, "
            "overlapped), format+write %.3fs, counters %.3fs, outside the phases %.3fs; reader thread: busy %.3fs "
            "(chunk reading and parsing %.3fs: reading %.3fs, counting %.3fs, writing %.3fs; pinned allocation "
            "%.3fs; copies %.3fs), waiting for room %.3fs\n",
            c.t_wait, c.t_pack, c.t_gpu_wait, c.t_gpu, c.t_format, c.t_count,
            secs - c.t_wait - c.t_pack - c.t_gpu_wait - c.t_format - c.t_count, c.t_rbusy, c.t_rparse,
            c.read.fill, c.read.count, c.read.write, c.t_ralloc, c.t_rcopy, c.t_rfull);
  if (!o.report_filename.empty()) {
    if (!k2host::report_kraken_style(o.report_filename, o.report_zero_counts, o.report_kmer_data,
                                     c.tax, c.counters, c.stats.total_sequences,
                                     c.stats.total_sequences - c.stats.total_classified))
      die(74, "unable to write report " + o.report_filename);
  }
  for (FILE *f : {out.c1, out.c2, out.u1, out.u2})
    if (f) fclose(f);
  if (out.kraken && out.kraken != stdout) fclose(out.kraken);
  for (Slot &sl : c.slots) k2g_batch_destroy(sl.batch);
  k2g_db_close(c.db);
  return 0;
}
