// C entry points of the CPU ORACLE for the Python tests (test
// infrastructure only; the shipped library never links this).
#include "k2cpu.h"
#include <cstring>
#include <string>
#include <vector>

using namespace k2cpu;

extern "C" {

// Runs the scanner over seq[start, finish) and writes one (ambiguous,
// minimizer) pair per returned k-mer.  Returns the number of outputs, or -1
// if cap is too small.
long k2cpu_scan_rc(long k, long l, uint64_t mask, uint64_t toggle, int dna, int revcom_version,
                   long start, long finish, const char *seq, long len, uint8_t *amb,
                   uint64_t *mins, long cap) {
  std::string s(seq, (size_t) len);
  Scanner sc(k, l, mask, dna != 0, toggle, revcom_version);
  sc.load(s, (size_t) start, finish < 0 ? SIZE_MAX : (size_t) finish);
  long n = 0;
  const uint64_t *m;
  while ((m = sc.next()) != nullptr) {
    if (n >= cap) return -1;
    amb[n] = sc.ambiguous() ? 1 : 0;
    mins[n] = *m;
    n++;
  }
  return n;
}

long k2cpu_scan(long k, long l, uint64_t mask, uint64_t toggle, int dna, long start, long finish,
                const char *seq, long len, uint8_t *amb, uint64_t *mins, long cap) {
  return k2cpu_scan_rc(k, l, mask, toggle, dna, 1, start, finish, seq, len, amb, mins, cap);
}

// Six frames, NUL-separated, into out (cap bytes).  Returns bytes written.
long k2cpu_translate(const char *seq, long len, char *out, long cap) {
  std::string s(seq, (size_t) len);
  std::vector<std::string> fr(6);
  translate_all_frames(s, fr);
  long pos = 0;
  for (auto &f : fr) {
    if (pos + (long) f.size() + 1 > cap) return -1;
    memcpy(out + pos, f.data(), f.size());
    pos += (long) f.size();
    out[pos++] = 0;
  }
  return pos;
}

// HLL script (same grammar as the reference harness): I/J/M/R/P lines.
// Each P appends (n_observed, estimate) to out.  Returns values written.
long k2cpu_hll_script(const char *script, uint64_t *out, long cap) {
  Hll a, b;
  long n = 0;
  const char *p = script;
  while (*p) {
    const char *e = strchr(p, '\n');
    std::string line = e ? std::string(p, e - p) : std::string(p);
    p = e ? e + 1 : p + strlen(p);
    if (line.empty()) continue;
    char op = line[0];
    if (op == 'I' || op == 'J') {
      const char *q = line.c_str() + 1;
      char *end;
      for (;;) {
        unsigned long long v = strtoull(q, &end, 10);
        if (end == q) break;
        (op == 'I' ? a : b).insert(v);
        q = end;
      }
    } else if (op == 'M') {
      a.merge(b);
      b = Hll();
    } else if (op == 'R') {
      a = Hll();
      b = Hll();
    } else if (op == 'P') {
      if (n + 2 > cap) return -1;
      out[n++] = a.n_observed;
      out[n++] = a.estimate();
    }
  }
  return n;
}

struct K2cpuDb {
  Index idx;
};

void *k2cpu_db_open(const char *hash, const char *taxo, const char *opts) {
  auto *d = new K2cpuDb();
  ClassifyOptions o;
  o.index_filename = hash;
  o.taxonomy_filename = taxo;
  o.options_filename = opts;
  std::string err;
  if (!load_index(o, d->idx, err)) {
    delete d;
    return nullptr;
  }
  return d;
}

void k2cpu_db_close(void *h) { delete (K2cpuDb *) h; }

// The oracle's whole classify driver (input blocks, per-block counters and
// HLL state, kraken output, kreport: src/classify.cc:304-352, 479-708,
// src/reports.cc:241-259) over input files, with the handle's resident
// index instead of loading one: the reference classify command line
// "-p threads [-P] -T confidence -g min_hit_groups [-R report [-K] [-z]]
// -O kraken_out file1 [file2]".  Returns the driver's exit status.
int k2cpu_run_files(void *h, int threads, int paired, double confidence, int min_hit_groups,
                    const char *report, int kmer_data, int zeros, const char *kraken_out,
                    const char *file1, const char *file2) {
  auto *d = (K2cpuDb *) h;
  ClassifyOptions o;
  o.num_threads = threads > 0 ? threads : 1;
  o.paired_end_processing = paired != 0;
  o.confidence_threshold = confidence;
  o.minimum_hit_groups = min_hit_groups;
  o.use_translated_search = !d->idx.opts.dna_db;
  if (report) o.report_filename = report;
  o.report_kmer_data = kmer_data != 0;
  o.report_zero_counts = zeros != 0;
  if (kraken_out) o.kraken_output_filename = kraken_out;
  o.filenames.push_back(file1);
  if (file2) o.filenames.push_back(file2);
  ClassifyStats stats;
  TaxonCounters counters;
  double secs = 0;
  int rc = run_classify(o, d->idx, stats, counters, &secs);
  if (rc) return rc;
  if (report)
    report_kraken_style(o.report_filename, o.report_zero_counts, o.report_kmer_data, d->idx.tax, counters,
                        stats.total_sequences, stats.total_sequences - stats.total_classified);
  return 0;
}

void k2cpu_get_batch(void *h, const uint64_t *keys, uint32_t *out, long n) {
  ((K2cpuDb *) h)->idx.table.get_batch(keys, out, (size_t) n);
}

void k2cpu_probe_lines(void *h, const uint64_t *keys, uint32_t *out, long n) {
  auto &t = ((K2cpuDb *) h)->idx.table;
  for (long i = 0; i < n; i++) out[i] = t.probe_lines(keys[i]);
}

uint64_t k2cpu_db_capacity(void *h) { return ((K2cpuDb *) h)->idx.table.capacity; }

// Classifies n fragments.  Mate sequences are concatenated in seq; mate 1
// of fragment i is seq[off[2i], off[2i+1]), mate 2 seq[off[2i+1], off[2i+2])
// when paired.  Per fragment: call (internal id), hit groups, total k-mers,
// lookups and probe lines; the kraken hitlist field (text after the length
// column) goes NUL-separated into text (cap bytes).  Returns bytes used or
// -1 when text is too small.
long k2cpu_classify(void *h, int paired, int quick, double confidence, int min_hit_groups,
                    const char *seq, const uint64_t *off, long n,
                    uint32_t *call, uint32_t *groups, uint32_t *total, uint32_t *lookups,
                    uint64_t *lines, char *text, long cap) {
  auto *d = (K2cpuDb *) h;
  ClassifyOptions o;
  o.paired_end_processing = paired != 0;
  o.quick_mode = quick != 0;
  o.confidence_threshold = confidence;
  o.minimum_hit_groups = min_hit_groups;
  o.use_translated_search = !d->idx.opts.dna_db;
  ReadResult r;
  long pos = 0;
  std::string s1, s2, line;
  for (long i = 0; i < n; i++) {
    s1.assign(seq + off[2 * i], seq + off[2 * i + 1]);
    if (paired) s2.assign(seq + off[2 * i + 1], seq + off[2 * i + 2]);
    classify_fragment(d->idx, o, s1, paired ? &s2 : nullptr, r, nullptr, lines != nullptr);
    call[i] = (uint32_t) r.call;
    groups[i] = (uint32_t) r.hit_groups;
    total[i] = (uint32_t) r.total_kmers;
    if (lookups) lookups[i] = r.lookups;
    if (lines) lines[i] = r.probe_lines;
    line.clear();
    if (quick)
      line = std::to_string(d->idx.tax.nodes[r.call].external_id) + ":Q";
    else if (r.taxa.empty())
      line = "0:0";
    else
      append_hitlist(line, r.taxa, d->idx.tax);
    if (pos + (long) line.size() + 1 > cap) return -1;
    memcpy(text + pos, line.data(), line.size());
    pos += (long) line.size();
    text[pos++] = 0;
  }
  return pos;
}

}  // extern "C"
