// k2cpu_merge -- CPU ORACLE (test infrastructure only).
//
// A restatement of the reference's multi-database output merge
// (src/merge.cc): two kraken outputs of the same reads, classified against
// two databases that share one taxonomy, are merged k-mer by k-mer (the LCA
// of the two calls at each position), the read is called again from the
// merged hit list, and the merged kraken output, classified-header file and
// kraken report are written.  Same command line as the reference
// (getopt string of src/merge.cc:978).
//
// Semantics follow the reference's multi-threaded path
// (merge_classification_output_parallel, src/merge.cc:566-793) whatever -T
// says: that path builds the external-to-internal taxid map
// (src/merge.cc:594) and ends the classified-header file with the line
// count (:659-662), which the k2 wrapper reads (scripts/k2:4096-4098).  The
// reference's one-thread path (:795-931) builds no map, so every LCA it
// takes maps both taxids to node 0 and every merged taxid prints as 0.
//
// Parity pinning: golden merged outputs, headers and reports of the
// compiled reference merge over reference classify outputs
// (tests/golden/merge, tests/test_oracle_golden.py).
#include <getopt.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <unordered_map>
#include <vector>

#include "k2cpu.h"

using namespace k2cpu;

namespace {

// src/merge.cc:176-179: markers as the merge tool stores them (int)
const int kAmbig = INT32_MAX, kMate = INT32_MAX - 1, kFrame = INT32_MAX - 2,
          kValid = INT32_MAX - 3;

struct TC {
  int taxid, count;
};

// src/merge.cc:217-244: the first len characters as decimal digits, no
// checks; more than 20 characters give 0
uint64_t atou(const char *p, size_t len) {
  if (len > 20) return 0;
  uint64_t r = 0;
  for (size_t i = 0; i < len; i++)
    r = r * 10 + (uint64_t) (int64_t) ((int) (unsigned char) p[i] - '0');
  return r;
}

// src/merge.cc:251-326 (int_to_string of an int widened to uint64)
void put_int(std::string &o, int v) {
  const uint64_t x = (uint64_t) (int64_t) v;
  if (x > (uint64_t) kValid) {
    // negative values print a stale buffer character in the reference:
    // outside the domain of classify outputs
    o += v == kAmbig ? 'A' : v == kMate ? '|' : v == kFrame ? '-' : '?';
    return;
  }
  o += std::to_string(x);
}

// src/merge.cc:556-564: fields split on '\t', empty fields skipped, at most n
int get_fields(const std::string &line, const char *f[], size_t len[], int n) {
  int k = 0;
  size_t p = 0;
  while (k < n) {
    size_t q = line.find('\t', p);
    const size_t e = q == std::string::npos ? line.size() : q;
    if (e > p) {
      f[k] = line.data() + p;
      len[k] = e - p;
      k++;
    }
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return k;
}

// src/merge.cc:423-469.  s[0..len) is the hit-list field (with the line's
// newline when it has one); one slot per ':' and a slot left {0,0} when the
// tokens run out first.
void parse_hit_list(const char *s, size_t len, std::vector<TC> &out) {
  size_t n_items = 0;
  for (size_t i = 0; i < len; i++) n_items += s[i] == ':';
  out.assign(n_items, TC{0, 0});
  const char *p = s, *end = s + len;
  size_t i = 0;
  while (p) {
    const char *colon = (const char *) memchr(p, ':', end - p);
    if (!colon) break;  // the count strsep gets no string
    const char *t = p;
    const size_t len1 = colon - t;
    const char *c = colon + 1;
    const char *sp = (const char *) memchr(c, ' ', end - c);
    size_t len2;
    if (sp) {
      len2 = sp - c;
      p = sp + 1;
    } else {
      len2 = (size_t) (end - c) - 1;  // drops the line's last character
      p = nullptr;
    }
    const char first = len1 ? t[0] : '\0';
    int taxid, count;
    if (first == 'A') {
      taxid = kAmbig;
      count = (int) (uint32_t) atou(c, len2);
    } else if (first == '|') {
      taxid = count = kMate;
    } else if (first == '-') {
      taxid = count = kFrame;
    } else {
      taxid = (int) (uint32_t) atou(t, len1);
      count = (int) (uint32_t) atou(c, len2);
    }
    out[i++] = TC{taxid, count};
  }
}

// src/merge.cc:471-492 (the cache does not change the value)
int lca_external(const Taxonomy &tax, int a, int b) {
  const uint64_t ia = tax.internal_id((uint64_t) (int64_t) a);
  const uint64_t ib = tax.internal_id((uint64_t) (int64_t) b);
  return (int) tax.nodes[tax.lca(ia, ib)].external_id;
}

using HitCounts = std::unordered_map<uint64_t, uint64_t>;

// src/merge.cc:494-554.  The reference reads past the second list when it
// is shorter than the first (undefined); here a missing item is taxid 0 with
// an unbounded count.  An empty first list merges to nothing.
size_t merge_hit_lists(const Taxonomy &tax, std::vector<TC> &h1, std::vector<TC> &h2, HitCounts &hc,
                       std::vector<TC> &merged) {
  size_t i1 = 0, i2 = 0, total = 0;
  if (h1.empty()) return 0;
  for (;;) {
    TC missing{0, INT_MAX};
    TC &a = h1[i1];
    TC &b = i2 < h2.size() ? h2[i2] : missing;
    int final_taxid;
    bool add;
    if (a.taxid > kValid) {
      final_taxid = a.taxid;
      add = false;
    } else {
      final_taxid = lca_external(tax, a.taxid, b.taxid);
      add = true;
    }
    int fc;
    if (a.count < b.count) {
      b.count -= a.count;
      i1++;
      fc = a.count;
      i2 += b.count == 0;
    } else if (b.count < a.count) {
      a.count -= b.count;
      i2++;
      fc = b.count;
      i1 += a.count == 0;
    } else {
      i1++;
      i2++;
      fc = a.count;
    }
    total += (size_t) (int64_t) fc;
    merged.push_back(TC{final_taxid, fc});
    if (add) hc[(uint64_t) (int64_t) final_taxid] += (uint64_t) (int64_t) fc;
    if (i1 >= h1.size()) break;
  }
  return total;
}

// uint32_t r = <float or double>: x86-64 converts through a 64-bit signed
// integer and keeps the low 32 bits
uint32_t to_u32(double r) {
  if (!(r > -9.2e18 && r < 9.2e18)) return 0;
  return (uint32_t) (int64_t) r;
}

// src/merge.cc:369-421.  When the confidence climb finds support it returns
// the INTERNAL id of the clade (:413), which is then printed as a taxid;
// otherwise the external id (:420).
int resolve_tree(const Taxonomy &tax, HitCounts &hc, size_t total, float conf) {
  int max_taxon = 0;
  uint32_t max_score = 0;
  const float prod = conf * (float) total;
  const uint32_t required = to_u32(std::ceil(prod));
  for (auto &kv : hc) {
    const int taxon = (int) tax.internal_id(kv.first);
    uint32_t score = 0;
    for (auto &kv2 : hc)
      if (tax.is_ancestor(tax.internal_id(kv2.first), taxon)) score += (uint32_t) kv2.second;
    if (score > max_score) {
      max_score = score;
      max_taxon = taxon;
    } else if (score == max_score) {
      max_taxon = (int) tax.lca(max_taxon, taxon);
    }
  }
  max_score = (uint32_t) hc[tax.nodes[max_taxon].external_id];
  while (max_taxon && max_score < required) {
    max_score = 0;
    for (auto &kv : hc)
      if (tax.is_ancestor(max_taxon, tax.internal_id(kv.first))) max_score += (uint32_t) kv.second;
    if (max_score >= required) return max_taxon;
    max_taxon = (int) tax.nodes[max_taxon].parent_id;
  }
  return (int) tax.nodes[max_taxon].external_id;
}

// src/merge.cc:328-367: adjacent equal taxa are joined; a {0,0} item prints
// the pending pair with no separator after it
void write_hit_list(const std::vector<TC> &h, std::string &o) {
  if (h.empty()) return;
  int pt = h[0].taxid, pc = h[0].count;
  auto put = [&](int t, int c) {
    put_int(o, t);
    o += ':';
    put_int(o, c);
  };
  for (size_t i = 1; i < h.size(); i++) {
    const int t = h[i].taxid, c = h[i].count;
    if (t == 0 && c == 0) {
      put(pt, pc);
      pt = t;
      pc = c;
    }
    if (pt == t) {
      pc += c;
    } else if (pc == 0 && c > 0) {
      pc = c;
      pt = t;
    } else {
      put(pt, pc);
      o += ' ';
      pt = t;
      pc = c;
    }
  }
  put(pt, pc);
}

bool read_line(std::istream &in, std::string &line) {  // keeps the '\n'
  if (!std::getline(in, line)) return false;
  if (!in.eof()) line += '\n';
  return true;
}

void usage(const char *prog, int code) {
  fprintf(stderr, "%s -o <output> -t <merged_tax.k2d> [ -r report ] [ -c classified_headers ] input1 input2\n",
          prog);
  exit(code);
}

}  // namespace

int main(int argc, char **argv) {
  const char *out_name = nullptr, *taxo = nullptr, *report = nullptr, *headers = nullptr;
  bool use_names = false, report_zeros = false;
  float conf = 0;
  int ch;
  while ((ch = getopt(argc, argv, "b:hi:r:mc:o:t:T:nz")) != -1) {
    switch (ch) {
      case 'b': case 'T': break;  // batch size and threads: no effect on the output
      case 'c': headers = optarg; break;
      case 'i': conf = strtof(optarg, nullptr); break;
      case 'n': use_names = true; break;
      case 'o': out_name = optarg; break;
      case 'r': report = optarg; break;
      case 'm': fprintf(stderr, "mpa-style reports are not provided\n"); return 6;
      case 't': taxo = optarg; break;
      case 'z': report_zeros = true; break;
      case 'h': usage(argv[0], 0); break;
      default: usage(argv[0], 6);
    }
  }
  if (argc - optind != 2) usage(argv[0], argc - optind < 2 ? 3 : 5);
  if (!out_name) usage(argv[0], 1);
  if (!taxo) usage(argv[0], 2);
  Taxonomy tax;
  std::string err;
  if (!tax.load(taxo, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  tax.build_external_map();
  std::ifstream in1(argv[optind]), in2(argv[optind + 1]);
  if (!in1 || !in2) {
    fprintf(stderr, "unable to open input\n");
    return 1;
  }
  std::string l1, l2, out, hdr;
  std::vector<TC> h1, h2, merged;
  HitCounts hc;
  TaxonCounters counters;
  uint64_t total_seqs = 0, total_unclassified = 0;
  while (read_line(in1, l1)) {
    if (!read_line(in2, l2)) {
      fprintf(stderr, "inputs differ in line count\n");
      return 1;
    }
    const char *f1[5], *f2[5];
    size_t n1[5], n2[5];
    if (get_fields(l1, f1, n1, 5) < 5 || get_fields(l2, f2, n2, 5) < 5) {
      fprintf(stderr, "malformed kraken line\n");
      return 1;
    }
    parse_hit_list(f1[4], n1[4], h1);
    parse_hit_list(f2[4], n2[4], h2);
    hc.clear();
    merged.clear();
    const size_t total = merge_hit_lists(tax, h1, h2, hc, merged);
    const int res = resolve_tree(tax, hc, total, conf);
    std::string taxid;
    const bool classified = f1[0][0] == 'C' || f2[0][0] == 'C';
    if (classified) {
      put_int(taxid, res);
    } else {
      taxid = "0";
      total_unclassified++;
    }
    if (use_names) {  // src/merge.cc:713-725
      const uint64_t t = (uint64_t) (int64_t) (int) (uint32_t) atou(taxid.data(), taxid.size());
      const char *name = tax.name(tax.internal_id(t));
      taxid = std::string(name) + " (taxid " + taxid + ")";
    }
    if (!merged.empty()) {
      out += classified ? "C\t" : "U\t";
      out.append(f1[1], n1[1]);
      out += '\t';
      out += taxid;
      out += '\t';
      out.append(f1[3], n1[3]);
      out += '\t';
      write_hit_list(merged, out);
      out += '\n';
    }
    if (report && classified) {  // src/merge.cc:736-740
      const uint64_t t = (uint64_t) (int64_t) (int) (uint32_t) atou(taxid.data(), taxid.size());
      counters[tax.internal_id(t)].n_reads++;
    }
    if (headers && classified) {
      hdr += '>';
      hdr.append(f1[1], n1[1]);
      hdr += '\n';
    }
    total_seqs++;
  }
  if (read_line(in2, l2)) {
    fprintf(stderr, "inputs differ in line count\n");
    return 1;
  }
  FILE *o = fopen(out_name, "w");
  if (!o) return 1;
  fwrite(out.data(), 1, out.size(), o);
  fclose(o);
  if (headers) {
    hdr += std::to_string(total_seqs) + "\n";
    FILE *h = fopen(headers, "w");
    if (!h) return 1;
    fwrite(hdr.data(), 1, hdr.size(), h);
    fclose(h);
  }
  if (report)
    report_kraken_style(report, report_zeros, false, tax, counters, total_seqs, total_unclassified);
  return 0;
}
