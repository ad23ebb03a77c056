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

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>
#include <omp.h>

#include "k2g.h"
#include "k2g_host.h"

namespace {

constexpr int32_t kAmbig = INT32_MAX, kMate = INT32_MAX - 1, kFrame = INT32_MAX - 2,
                  kValid = INT32_MAX - 3;
size_t g_batch_lines = 1u << 18;  // K2G_MERGE_BATCH_LINES

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
  char *e = buf + sizeof(buf), *b = e;
  uint32_t x = (uint32_t) v;
  do {
    *--b = (char) ('0' + x % 10);
    x /= 10;
  } while (x);
  o.append(b, e - b);
}

// A byte buffer that grows without initialising its new bytes; pinned
// buffers (uploaded to the GPU) come from k2g_host_alloc.
struct Bytes {
  char *p = nullptr;
  size_t size = 0, cap = 0;
  bool pinned = false;
  Bytes() = default;
  explicit Bytes(bool pin) : pinned(pin) {}
  Bytes(const Bytes &) = delete;
  Bytes &operator=(const Bytes &) = delete;
  Bytes(Bytes &&o) noexcept : p(o.p), size(o.size), cap(o.cap), pinned(o.pinned) {
    o.p = nullptr;
    o.size = o.cap = 0;
  }
  ~Bytes() { release(p); }
  char *data() { return p; }
  const char *data() const { return p; }
  void release(char *q) {
    if (!q) return;
    if (pinned)
      k2g_host_free(q);
    else
      delete[] q;
  }
  void reserve(size_t n) {
    if (n <= cap) return;
    const size_t c = std::max(n, 2 * cap);
    char *q = nullptr;
    if (pinned) {
      if (k2g_host_alloc(c, (void **) &q)) die(71, k2g_last_error());
    } else {
      q = new char[c];
    }
    if (size) memcpy(q, p, size);
    release(p);
    p = q;
    cap = c;
  }
};

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
  // Up to max_lines lines into buf (replacing its contents); starts gets each
  // line's offset plus the end offset.
  size_t read(size_t max_lines, Bytes &buf, std::vector<size_t> &starts) {
    buf.size = 0;
    buf.reserve(carry_.size + 1);
    if (carry_.size) memcpy(buf.data(), carry_.data(), carry_.size);
    buf.size = carry_.size;
    carry_.size = 0;
    starts.clear();
    size_t pos = 0, lines = 0;
    for (;;) {
      while (lines < max_lines) {
        const char *nl = (const char *) memchr(buf.data() + pos, '\n', buf.size - pos);
        if (!nl) break;
        starts.push_back(pos);
        pos = nl - buf.data() + 1;
        lines++;
      }
      if (lines == max_lines) break;
      if (eof_) {
        if (pos < buf.size) {  // a last line without '\n'
          starts.push_back(pos);
          pos = buf.size;
          lines++;
        }
        break;
      }
      constexpr size_t kChunk = 8u << 20;
      buf.reserve(buf.size + kChunk);
      const size_t got = fread(buf.data() + buf.size, 1, kChunk, f_);
      buf.size += got;
      if (got == 0) eof_ = true;
    }
    carry_.reserve(buf.size - pos + 1);
    carry_.size = buf.size - pos;
    if (carry_.size) memcpy(carry_.data(), buf.data() + pos, carry_.size);
    buf.size = pos;
    starts.push_back(pos);
    return lines;
  }

 private:
  FILE *f_ = nullptr;
  bool eof_ = false;
  Bytes carry_;
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

struct Batch {  // lines of both inputs
  Bytes buf1{true}, buf2{true};
  std::vector<size_t> st1, st2;
  size_t n = 0;
};

// Growable pinned host array (uploads from it run at full link speed).
template <class T>
struct Pinned {
  T *p = nullptr;
  size_t cap = 0;
  void reserve(size_t n) {
    if (n <= cap) return;
    if (p) k2g_host_free(p);
    p = nullptr;
    cap = std::max(n + n / 4, (size_t) 1024);
    if (k2g_host_alloc(cap * sizeof(T), (void **) &p)) die(71, k2g_last_error());
  }
  ~Pinned() {
    if (p) k2g_host_free(p);
  }
};

struct Slot {  // one batch on its way through parse, GPU merge and print
  std::unique_ptr<Batch> batch;
  std::vector<Fields> fa, fb;
  Pinned<uint64_t> field_a, field_b;  // hit-list field offsets
  Pinned<uint32_t> len_a, len_b;                     // hit-list field lengths
  k2g_merge_results res{};
  std::vector<Bytes> text;          // printed lines, one buffer per print thread
  std::vector<std::string> names;   // -c lines, likewise
  bool writing = false;             // queued for the writer thread
};

// Fields of every line pair, on the OpenMP threads (the hit lists are
// counted and tokenized on the GPU).
void parse_batch(Slot &sl, int threads) {
  const Batch &b = *sl.batch;
  const size_t n = b.n;
  sl.fa.resize(n);
  sl.fb.resize(n);
  for (auto *q : {&sl.field_a, &sl.field_b}) q->reserve(n + 1);
  sl.len_a.reserve(n + 1);
  sl.len_b.reserve(n + 1);
  bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad) num_threads(threads)
  for (size_t i = 0; i < n; i++) {
    if (!get_fields(b.buf1.data() + b.st1[i], b.buf1.data() + b.st1[i + 1], sl.fa[i]) ||
        !get_fields(b.buf2.data() + b.st2[i], b.buf2.data() + b.st2[i + 1], sl.fb[i])) {
      bad = true;
      continue;
    }
    sl.field_a.p[i] = sl.fa[i].f[4] - b.buf1.data();
    sl.field_b.p[i] = sl.fb[i].f[4] - b.buf2.data();
    sl.len_a.p[i] = sl.fa[i].len[4];
    sl.len_b.p[i] = sl.fb[i].len[4];
  }
  if (bad) die(65, "malformed kraken output line (fewer than five fields)");
}

// Hands a printed batch's line buffers back to the reader.
void recycle(Slot &sl, std::mutex &mu, std::deque<std::unique_ptr<Batch>> &spare) {
  std::lock_guard<std::mutex> lk(mu);
  if (spare.size() < 2) spare.push_back(std::move(sl.batch));
  sl.batch.reset();
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
  if (const char *v = getenv("K2G_MERGE_BATCH_LINES")) g_batch_lines = std::max(1ull, strtoull(v, nullptr, 10));

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
  // two merge slots: the GPU merges one batch while the host prints the other
  k2g_merger *mg[2] = {nullptr, nullptr};
  for (auto &m : mg) {
    const int rc = k2g_merge_open(taxo, &m);
    if (rc) die(rc, k2g_last_error());
  }
  LineFile in1, in2;
  if (!in1.open(argv[optind])) die(66, std::string("unable to open ") + argv[optind]);
  if (!in2.open(argv[optind + 1])) die(66, std::string("unable to open ") + argv[optind + 1]);
  FILE *out = fopen(out_name, "w");
  if (!out) die(73, std::string("unable to open ") + out_name);
  FILE *hdr = nullptr;
  if (headers && !(hdr = fopen(headers, "w"))) die(73, std::string("unable to open ") + headers);

  // reader thread: both files read side by side, two batches ahead
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::unique_ptr<Batch>> ready, spare;
  bool finished = false, mismatch = false;
  size_t hint1 = 0, hint2 = 0;
  std::thread reader([&] {
    for (;;) {
      std::unique_ptr<Batch> b;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return ready.size() < 2; });
        if (!spare.empty()) {
          b = std::move(spare.front());
          spare.pop_front();
        }
      }
      if (!b) {  // pinned line buffers sized from the batches read so far
        b.reset(new Batch());
        b->buf1.reserve(hint1 + hint1 / 4 + 1);
        b->buf2.reserve(hint2 + hint2 / 4 + 1);
      }
      size_t n2 = 0;
      std::thread second([&] { n2 = in2.read(g_batch_lines, b->buf2, b->st2); });
      b->n = in1.read(g_batch_lines, b->buf1, b->st1);
      second.join();
      hint1 = std::max(hint1, b->buf1.size);
      hint2 = std::max(hint2, b->buf2.size);
      std::lock_guard<std::mutex> lk(mu);
      if (n2 != b->n) {
        mismatch = finished = true;
      } else if (b->n == 0) {
        finished = true;
      } else {
        ready.push_back(std::move(b));
      }
      cv.notify_all();
      if (finished) return;
    }
  });

  Slot slot[2];
  for (auto &sl : slot) {
    sl.text.resize(threads);
    sl.names.resize(threads);
  }
  // writer thread: printed batches go to the files in order while the next
  // batches are parsed, merged and printed
  std::mutex wmu;
  std::condition_variable wcv;
  std::deque<Slot *> wq;
  bool wdone = false;
  std::thread writer([&] {
    for (;;) {
      Slot *sl;
      {
        std::unique_lock<std::mutex> lk(wmu);
        wcv.wait(lk, [&] { return !wq.empty() || wdone; });
        if (wq.empty()) return;
        sl = wq.front();
      }
      for (int w = 0; w < threads; w++) {
        fwrite(sl->text[w].data(), 1, sl->text[w].size, out);
        if (hdr) fwrite(sl->names[w].data(), 1, sl->names[w].size(), hdr);
      }
      std::lock_guard<std::mutex> lk(wmu);
      wq.pop_front();
      sl->writing = false;
      wcv.notify_all();
    }
  });
  std::vector<k2host::TaxonCounters> part_counters(threads);
  std::vector<uint64_t> part_unclassified(threads);
  uint64_t total_seqs = 0, total_unclassified = 0;
  double t_wait = 0, t_parse = 0, t_gpu_wait = 0, t_format = 0, t_kernel = 0;
  const double t_start = now_s();
  omp_set_num_threads(threads);

  // prints one merged batch (src/merge.cc:701-758), in input order
  auto print = [&](Slot &sl) {
    {
      std::unique_lock<std::mutex> lk(wmu);
      wcv.wait(lk, [&] { return !sl.writing; });
    }
    const Batch &bt = *sl.batch;
    const size_t n = bt.n;
    const k2g_merge_results &res = sl.res;
#pragma omp parallel num_threads(threads)
    {
      const int w = omp_get_thread_num(), nw = omp_get_num_threads();
      const size_t lo = n * w / nw, hi = n * (w + 1) / nw;
      Bytes &o = sl.text[w];
      std::string &h = sl.names[w];
      o.size = 0;
      h.clear();
      std::string taxid;
      uint64_t unc = 0;
      for (size_t i = lo; i < hi; i++) {
        const Fields &fa = sl.fa[i], &fb = sl.fb[i];
        const bool classified = fa.f[0][0] == 'C' || fb.f[0][0] == 'C';
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
        const uint64_t t0 = res.text_off[i], t1 = res.text_off[i + 1];
        if (t1 > t0) {
          o.reserve(o.size + 8 + fa.len[1] + taxid.size() + fa.len[3] + (t1 - t0));
          char *q = o.data() + o.size;
          *q++ = classified ? 'C' : 'U';
          *q++ = '\t';
          memcpy(q, fa.f[1], fa.len[1]);
          q += fa.len[1];
          *q++ = '\t';
          memcpy(q, taxid.data(), taxid.size());
          q += taxid.size();
          *q++ = '\t';
          memcpy(q, fa.f[3], fa.len[3]);
          q += fa.len[3];
          *q++ = '\t';
          memcpy(q, res.text + t0, t1 - t0);  // the hit list, printed on the GPU
          q += t1 - t0;
          *q++ = '\n';
          o.size = q - o.data();
        }
        if (classified) {
          if (report) part_counters[w][node_of(taxid)].n_reads++;
          if (hdr) {
            h += '>';
            h.append(fa.f[1], fa.len[1]);
            h += '\n';
          }
        }
      }
      part_unclassified[w] = unc;
    }
    for (int w = 0; w < threads; w++) total_unclassified += part_unclassified[w];
    total_seqs += n;
    std::lock_guard<std::mutex> lk(wmu);
    sl.writing = true;
    wq.push_back(&sl);
    wcv.notify_all();
  };

  Slot *pending = nullptr;  // merged on the GPU, not yet printed
  for (int j = 0;; j ^= 1) {
    double t0 = now_s();
    std::unique_ptr<Batch> b;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !ready.empty() || finished; });
      if (!ready.empty()) {
        b = std::move(ready.front());
        ready.pop_front();
      }
      cv.notify_all();
    }
    double t1 = now_s();
    t_wait += t1 - t0;
    if (!b) break;
    Slot &sl = slot[j];
    sl.batch = std::move(b);
    parse_batch(sl, threads);
    t0 = now_s();
    t_parse += t0 - t1;
    int rc = 0;
    std::string gpu_error;  // the library's last error is per thread
    k2g_merge_params p{conf, 1};
    std::thread gpu([&] {
      const Batch &bt = *sl.batch;
      k2g_hit_list_text a{bt.buf1.data(), bt.buf1.size, sl.field_a.p, sl.len_a.p, nullptr, bt.n};
      k2g_hit_list_text bb{bt.buf2.data(), bt.buf2.size, sl.field_b.p, sl.len_b.p, nullptr, bt.n};
      rc = k2g_merge_text(mg[j], &p, &a, &bb, &sl.res);
      if (rc) gpu_error = k2g_last_error();
    });
    if (pending) {
      print(*pending);
      recycle(*pending, mu, spare);
    }
    t1 = now_s();
    t_format += t1 - t0;
    gpu.join();
    t_gpu_wait += now_s() - t1;
    if (rc) die(rc, gpu_error);
    t_kernel += sl.res.kernel_ms / 1e3;
    pending = &sl;
  }
  reader.join();
  if (mismatch) die(65, "the two inputs hold different numbers of lines");
  if (pending) {
    const double t0 = now_s();
    print(*pending);
    t_format += now_s() - t0;
  }
  {
    std::lock_guard<std::mutex> lk(wmu);
    wdone = true;
    wcv.notify_all();
  }
  writer.join();
  if (ferror(out) || fclose(out) != 0) die(74, std::string("unable to write ") + out_name);
  if (hdr) {
    fprintf(hdr, "%llu\n", (unsigned long long) total_seqs);
    if (ferror(hdr) || fclose(hdr) != 0) die(74, std::string("unable to write ") + headers);
  }
  if (report) {
    k2host::TaxonCounters counters;
    for (auto &pc : part_counters)
      for (auto &kv : pc) counters[kv.first].n_reads += kv.second.n_reads;
    if (!k2host::report_kraken_style(report, report_zeros, false, tax, counters, total_seqs,
                                     total_unclassified))
      die(74, std::string("unable to write report ") + report);
  }
  for (auto &m : mg) k2g_merge_close(m);
  if (timing)
    fprintf(stderr,
            "%llu lines merged in %.3fs (input wait %.3fs, parse %.3fs, print+write %.3fs, GPU wait %.3fs; "
            "kernels %.3fs)\n",
            (unsigned long long) total_seqs, now_s() - t_start, t_wait, t_parse, t_format, t_gpu_wait, t_kernel);
  return 0;
}
