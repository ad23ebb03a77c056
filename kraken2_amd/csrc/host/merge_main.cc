// k2g_merge: the multi-database merge front-end over libk2g.
//
// Same command line as the reference's merge tool (getopt string of
// src/merge.cc:978), so the k2 wrapper's merge_classification_output
// (scripts/k2:3890-3916) can run it unchanged: two kraken outputs of the
// same reads, classified against two databases that share a taxonomy, are
// merged into one kraken output (-o), with the classified read names (-c)
// and a kraken report (-r, -z) as options.  Lines are read in batches; each
// batch is parsed on -T host threads, merged on the GPU (k2g_merge: the
// per-k-mer LCA merge, the new call and the run folding), and printed on
// the host threads in input order.
//
// The output follows the reference's multi-threaded path
// (src/merge.cc:566-793) for every -T: that path maps taxids to taxonomy
// nodes (:594) and ends the -c file with the number of reads (:659-662),
// which the k2 wrapper reads back (scripts/k2:4096-4098).  mpa-style reports
// (-m) are not provided.
#include <getopt.h>
#include <sys/time.h>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <unordered_map>
#include <vector>
#include <omp.h>

#include "k2g.h"
#include "k2g_host.h"

namespace {

constexpr int32_t kAmbig = INT32_MAX, kMate = INT32_MAX - 1, kFrame = INT32_MAX - 2,
                  kValid = INT32_MAX - 3;
constexpr size_t kBatchLines = 1u << 20;

void die(int code, const std::string &msg) {
  std::cerr << "merge: " << msg << "\n";
  exit(code);
}

// src/merge.cc:933-949 (messages and exit codes of usage())
void usage(const char *prog, int err) {
  static const char *errs[] = {"", "Required argument, -o, not specified", "Required argument, -t, not specified",
                               "No input files provided", "Only one input file provided",
                               "Too many input files provided", "Invalid argument"};
  fprintf(stderr, "Error: %s\n", errs[err]);
  fprintf(stderr, "%s -o <output> -t <merged_tax.k2d> [ -r report ] [ -c classified_headers ] input1 input2\n",
          prog);
  exit(err);
}

// src/merge.cc:217-244: the first len characters as decimal digits with no
// checks; more than 20 characters give 0
uint64_t atou(const char *p, size_t len) {
  if (len > 20) return 0;
  uint64_t r = 0;
  for (size_t i = 0; i < len; i++) r = r * 10 + (uint64_t) (int64_t) ((int) (unsigned char) p[i] - '0');
  return r;
}

// src/merge.cc:251-326: an int printed through int_to_string.  Negative
// values print a stale buffer character there; they are outside what
// classify writes, and print '?' here.
void put_int(std::string &o, int32_t v) {
  if (v > kValid || v < 0) {
    o += v == kAmbig ? 'A' : v == kMate ? '|' : v == kFrame ? '-' : '?';
    return;
  }
  char buf[16];
  int n = 0;
  uint32_t x = (uint32_t) v;
  do {
    buf[n++] = (char) ('0' + x % 10);
    x /= 10;
  } while (x);
  while (n) o += buf[--n];
}

// Lines of a file, batch by batch, each line with its '\n' when it has one.
class LineFile {
 public:
  bool open(const char *path) {
    f_ = fopen(path, "rb");
    return f_ != nullptr;
  }
  ~LineFile() {
    if (f_) fclose(f_);
  }
  // Appends up to max_lines lines to buf (cleared first); starts gets each
  // line's offset plus the end offset.
  size_t read(size_t max_lines, std::string &buf, std::vector<size_t> &starts) {
    buf.swap(carry_);
    carry_.clear();
    starts.clear();
    size_t pos = 0, lines = 0;
    for (;;) {
      while (lines < max_lines) {
        const char *nl = (const char *) memchr(buf.data() + pos, '\n', buf.size() - pos);
        if (!nl) break;
        starts.push_back(pos);
        pos = nl - buf.data() + 1;
        lines++;
      }
      if (lines == max_lines) break;
      if (eof_) {
        if (pos < buf.size()) {  // a last line without '\n'
          starts.push_back(pos);
          pos = buf.size();
          lines++;
        }
        break;
      }
      const size_t old = buf.size();
      buf.resize(old + (16u << 20));
      const size_t got = fread(&buf[old], 1, 16u << 20, f_);
      buf.resize(old + got);
      if (got == 0) eof_ = true;
    }
    carry_.assign(buf, pos, std::string::npos);
    buf.resize(pos);
    starts.push_back(pos);
    return lines;
  }

 private:
  FILE *f_ = nullptr;
  bool eof_ = false;
  std::string carry_;
};

struct Fields {  // src/merge.cc:556-564 and the five fields of :573-579
  const char *f[5];
  uint32_t len[5];
};

// Splits on '\t' skipping empty fields; false if fewer than five.
bool get_fields(const char *p, const char *end, Fields &fl) {
  int k = 0;
  while (k < 5) {
    const char *t = (const char *) memchr(p, '\t', end - p);
    const char *e = t ? t : end;
    if (e > p) {
      fl.f[k] = p;
      fl.len[k] = (uint32_t) (e - p);
      k++;
    }
    if (!t) break;
    p = t + 1;
  }
  return k == 5;
}

uint64_t count_items(const char *s, size_t len) {
  uint64_t n = 0;
  for (size_t i = 0; i < len; i++) n += s[i] == ':';
  return n;
}

// src/merge.cc:423-469 into out[0 .. colon count): one slot per ':' (slots
// the tokens do not reach stay {0,0}); the last count loses the field's
// last character (the line's newline).
void parse_hit_list(const char *s, size_t len, int32_t *out, uint64_t slots) {
  std::memset(out, 0, slots * 8);
  const char *p = s, *end = s + len;
  uint64_t i = 0;
  while (p) {
    const char *colon = (const char *) memchr(p, ':', end - p);
    if (!colon) break;
    const char *t = p;
    const size_t len1 = colon - t;
    const char *c = colon + 1;
    const char *sp = (const char *) memchr(c, ' ', end - c);
    size_t len2;
    if (sp) {
      len2 = sp - c;
      p = sp + 1;
    } else {
      len2 = (size_t) (end - c) - 1;
      p = nullptr;
    }
    const char first = len1 ? t[0] : '\0';
    int32_t taxid, count;
    if (first == 'A') {
      taxid = kAmbig;
      count = (int32_t) (uint32_t) atou(c, len2);
    } else if (first == '|') {
      taxid = count = kMate;
    } else if (first == '-') {
      taxid = count = kFrame;
    } else {
      taxid = (int32_t) (uint32_t) atou(t, len1);
      count = (int32_t) (uint32_t) atou(c, len2);
    }
    out[2 * i] = taxid;
    out[2 * i + 1] = count;
    i++;
  }
}

double now_s() {
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec / 1e6;
}

}  // namespace

int main(int argc, char **argv) {
  int threads = 1;
  bool use_names = false, report_zeros = false;
  float conf = 0;
  const char *report = nullptr, *headers = nullptr, *taxo = nullptr, *out_name = nullptr;
  int ch;
  while ((ch = getopt(argc, argv, "b:hi:r:mc:o:t:T:nz")) != -1) {
    switch (ch) {
      case 'b': break;  // lines per reference worker batch: no effect on the output
      case 'c': headers = optarg; break;
      case 'i': conf = strtof(optarg, nullptr); break;
      case 'n': use_names = true; break;
      case 'o': out_name = optarg; break;
      case 'r': report = optarg; break;
      case 'm': die(64, "mpa-style reports are not provided by this front-end");
      case 't': taxo = optarg; break;
      case 'T': threads = (int) strtod(optarg, nullptr); break;
      case 'z': report_zeros = true; break;
      case 'h': usage(argv[0], 0); break;
      default: usage(argv[0], 6);
    }
  }
  const int n_in = argc - optind;
  if (n_in != 2) usage(argv[0], n_in == 0 ? 3 : n_in == 1 ? 4 : 5);
  if (!out_name) usage(argv[0], 1);
  if (!taxo) usage(argv[0], 2);
  if (threads < 1) threads = 1;
  const bool timing = getenv("K2G_TIMING") && atoi(getenv("K2G_TIMING"));

  k2host::Taxonomy tax;
  std::string err;
  if (!tax.load(taxo, err)) die(65, err);
  // GenerateExternalToInternalIDMap (src/taxonomy.cc:289-294)
  std::unordered_map<uint64_t, uint64_t> ext_to_node;
  ext_to_node[0] = 0;
  for (size_t i = 1; i < tax.nodes.size(); i++) ext_to_node[tax.nodes[i].external_id] = i;
  auto node_of = [&](const std::string &taxid) -> uint64_t {  // src/merge.cc:714-716, 737-739
    const uint64_t t = (uint64_t) (int64_t) (int32_t) (uint32_t) atou(taxid.data(), taxid.size());
    auto it = ext_to_node.find(t);
    return it == ext_to_node.end() ? 0 : it->second;
  };
  k2g_merger *mg = nullptr;
  int rc = k2g_merge_open(taxo, &mg);
  if (rc) die(rc, k2g_last_error());

  LineFile in1, in2;
  if (!in1.open(argv[optind])) die(66, std::string("unable to open ") + argv[optind]);
  if (!in2.open(argv[optind + 1])) die(66, std::string("unable to open ") + argv[optind + 1]);
  FILE *out = fopen(out_name, "w");
  if (!out) die(73, std::string("unable to open ") + out_name);
  FILE *hdr = nullptr;
  if (headers && !(hdr = fopen(headers, "w"))) die(73, std::string("unable to open ") + headers);

  std::string buf1, buf2;
  std::vector<size_t> st1, st2;
  std::vector<Fields> fa, fb;
  std::vector<uint64_t> off_a, off_b;
  std::vector<int32_t> items_a, items_b;
  std::vector<std::string> text(threads), names(threads);
  std::vector<k2host::TaxonCounters> part_counters(threads);
  std::vector<uint64_t> part_unclassified(threads);
  uint64_t total_seqs = 0, total_unclassified = 0;
  double t_read = 0, t_parse = 0, t_gpu = 0, t_format = 0, t_kernel = 0;
  const double t_start = now_s();
  omp_set_num_threads(threads);
  for (;;) {
    double t0 = now_s();
    const size_t n = in1.read(kBatchLines, buf1, st1);
    const size_t n2 = in2.read(n ? n : 1, buf2, st2);
    if (n2 != n) die(65, "the two inputs hold different numbers of lines");
    double t1 = now_s();
    t_read += t1 - t0;
    if (n == 0) break;
    fa.resize(n);
    fb.resize(n);
    off_a.assign(n + 1, 0);
    off_b.assign(n + 1, 0);
    bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (size_t i = 0; i < n; i++) {
      if (!get_fields(buf1.data() + st1[i], buf1.data() + st1[i + 1], fa[i]) ||
          !get_fields(buf2.data() + st2[i], buf2.data() + st2[i + 1], fb[i])) {
        bad = true;
        continue;
      }
      off_a[i + 1] = count_items(fa[i].f[4], fa[i].len[4]);
      off_b[i + 1] = count_items(fb[i].f[4], fb[i].len[4]);
    }
    if (bad) die(65, "malformed kraken output line (fewer than five fields)");
    for (size_t i = 0; i < n; i++) {
      off_a[i + 1] += off_a[i];
      off_b[i + 1] += off_b[i];
    }
    items_a.resize(2 * off_a[n]);
    items_b.resize(2 * off_b[n]);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
      parse_hit_list(fa[i].f[4], fa[i].len[4], items_a.data() + 2 * off_a[i], off_a[i + 1] - off_a[i]);
      parse_hit_list(fb[i].f[4], fb[i].len[4], items_b.data() + 2 * off_b[i], off_b[i + 1] - off_b[i]);
    }
    t0 = now_s();
    t_parse += t0 - t1;
    k2g_merge_params p{conf};
    k2g_hit_lists a{items_a.data(), off_a.data(), n}, b{items_b.data(), off_b.data(), n};
    k2g_merge_results res;
    rc = k2g_merge(mg, &p, &a, &b, &res);
    if (rc) die(rc, k2g_last_error());
    t1 = now_s();
    t_gpu += t1 - t0;
    t_kernel += res.kernel_ms / 1e3;
    // print in per-thread contiguous ranges, then write in order
#pragma omp parallel num_threads(threads)
    {
      const int w = omp_get_thread_num(), nw = omp_get_num_threads();
      const size_t lo = n * w / nw, hi = n * (w + 1) / nw;
      std::string &o = text[w], &h = names[w];
      o.clear();
      h.clear();
      std::string taxid;
      uint64_t unc = 0;
      for (size_t i = lo; i < hi; i++) {
        const bool classified = fa[i].f[0][0] == 'C' || fb[i].f[0][0] == 'C';
        taxid.clear();
        if (classified)
          put_int(taxid, res.call[i]);
        else {
          taxid = "0";
          unc++;
        }
        if (use_names) {  // src/merge.cc:713-725
          const char *name = tax.name(node_of(taxid));
          taxid = std::string(name) + " (taxid " + taxid + ")";
        }
        const uint32_t nt = res.n_tokens[i];
        if (nt) {
          o += classified ? "C\t" : "U\t";
          o.append(fa[i].f[1], fa[i].len[1]);
          o += '\t';
          o += taxid;
          o += '\t';
          o.append(fa[i].f[3], fa[i].len[3]);
          o += '\t';
          const uint64_t s = off_a[i] + off_b[i];
          for (uint32_t k = 0; k < nt; k++) {
            put_int(o, res.tokens[2 * (s + k)]);
            o += ':';
            put_int(o, res.tokens[2 * (s + k) + 1]);
            if (!res.joined[s + k]) o += ' ';
          }
          o += '\n';
        }
        if (classified) {
          if (report) part_counters[w][node_of(taxid)].n_reads++;
          if (hdr) {
            h += '>';
            h.append(fa[i].f[1], fa[i].len[1]);
            h += '\n';
          }
        }
      }
      part_unclassified[w] = unc;
    }
    for (int w = 0; w < threads; w++) {
      fwrite(text[w].data(), 1, text[w].size(), out);
      if (hdr) fwrite(names[w].data(), 1, names[w].size(), hdr);
      total_unclassified += part_unclassified[w];
    }
    total_seqs += n;
    t_format += now_s() - t1;
  }
  if (fclose(out) != 0) die(74, std::string("unable to write ") + out_name);
  if (hdr) {
    fprintf(hdr, "%llu\n", (unsigned long long) total_seqs);
    if (fclose(hdr) != 0) die(74, std::string("unable to write ") + headers);
  }
  if (report) {
    k2host::TaxonCounters counters;
    for (auto &pc : part_counters)
      for (auto &kv : pc) counters[kv.first].n_reads += kv.second.n_reads;
    if (!k2host::report_kraken_style(report, report_zeros, false, tax, counters, total_seqs,
                                     total_unclassified))
      die(74, std::string("unable to write report ") + report);
  }
  k2g_merge_close(mg);
  if (timing)
    fprintf(stderr,
            "%llu lines merged in %.3fs (read %.3fs, parse %.3fs, GPU %.3fs of which kernel %.3fs, "
            "format+write %.3fs)\n",
            (unsigned long long) total_seqs, now_s() - t_start, t_read, t_parse, t_gpu, t_kernel, t_format);
  return 0;
}
