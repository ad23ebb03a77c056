// Taxonomy names/ranks, distinct-minimizer estimates and the Kraken-style
// report of the classify front-end.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <unordered_map>
#include <unordered_set>

#include "k2g_host.h"

namespace k2host {

bool Taxonomy::load(const std::string &path, std::string &err) {
  std::ifstream in(path, std::ios::binary);
  char magic[8];
  uint64_t hdr[3];
  in.read(magic, 8);
  if (!in || std::memcmp(magic, "K2TAXDAT", 8) != 0) {
    err = "malformed taxonomy file " + path;
    return false;
  }
  in.read((char *) hdr, 24);
  if (!in) {
    err = "read exhausted taxonomy information in " + path;
    return false;
  }
  nodes.resize(hdr[0]);
  names.resize(hdr[1]);
  ranks.resize(hdr[2]);
  in.read((char *) nodes.data(), hdr[0] * sizeof(Node));
  in.read(names.data(), hdr[1]);
  in.read(ranks.data(), hdr[2]);
  if (!in) {
    err = "read exhausted taxonomy information in " + path;
    return false;
  }
  return true;
}

// ---- HyperLogLog++ state (p = 12, p' = 25)
static constexpr int kP = 12, kPPrime = 25;

// rank carried by a sparse-encoded value (src/hyperloglogplus.cc:135-144)
static inline uint32_t encoded_rank(uint32_t e) {
  if (e & 1u) return (kPPrime - kP) + ((e >> 1) & 0x3fu);
  const uint32_t rest = e << kP;
  return (rest ? (uint32_t) __builtin_clz(rest) : (uint32_t) (32 - kP)) + 1u;
}

static void union_into(std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  if (b.empty()) return;
  if (a.empty()) {
    a = b;
    return;
  }
  std::vector<uint32_t> out;
  out.reserve(a.size() + b.size());
  std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
  a.swap(out);
}

void Distinct::merge(const Distinct &o) {
  if (o.n_observed == 0) return;
  n_observed += o.n_observed;
  dense = dense || o.dense;
  union_into(encoded, o.encoded);
}

// Ertl 2017 (src/hyperloglogplus.cc:300-345, 708-739); same operation order
// so the doubles match bit for bit.
static double ertl_sigma(double x) {
  if (x == 1.0) return std::numeric_limits<double>::infinity();
  double prev, s = x, y = 1.0;
  do {
    prev = s;
    x *= x;
    s += x * y;
    y += y;
  } while (s != prev);
  return s;
}

static double ertl_tau(double x) {
  if (x == 0.0 || x == 1.0) return 0.0;
  double prev, y = 1.0, t = 1 - x;
  do {
    prev = t;
    x = std::sqrt(x);
    y /= 2.0;
    t -= std::pow(1 - x, 2) * y;
  } while (t != prev);
  return t / 3.0;
}

uint64_t Distinct::estimate() const {
  size_t q, m;
  std::vector<int> C;
  if (!dense) {
    q = 64 - kPPrime;
    m = (size_t) 1 << kPPrime;
    C.assign(q + 2, 0);
    for (uint32_t e : encoded) C[encoded_rank(e)]++;
    C[0] = (int) (m - encoded.size());
  } else {
    q = 64 - kP;
    m = (size_t) 1 << kP;
    std::vector<uint8_t> reg(m, 0);
    for (uint32_t e : encoded) {
      const uint32_t idx = e >> (32 - kP);
      const uint8_t r = (uint8_t) encoded_rank(e);
      if (r > reg[idx]) reg[idx] = r;
    }
    C.assign(q + 2, 0);
    for (uint8_t r : reg) C[r]++;
  }
  double den = m * ertl_tau(1.0 - double(C[q + 1]) / double(m));
  for (int k = (int) q; k >= 1; --k) {
    den += C[k];
    den *= 0.5;
  }
  den += m * ertl_sigma(double(C[0]) / double(m));
  const double m_sq_alpha_inf = (m / (2.0 * std::log(2))) * m;
  const double est = m_sq_alpha_inf / den;
  return (n_observed < est) ? n_observed : (uint64_t) std::round(est);
}

void fold_block_records(const std::vector<std::pair<uint32_t, uint32_t>> &recs,
                        TaxonCounters &total) {
  // A block's sketch of a taxon switches to registers iff its sparse list
  // already holds 1024 distinct values when another insert arrives
  // (src/hyperloglogplus.cc:482-484): distinct(e_1..e_{n-1}) >= 1024.
  std::unordered_map<uint32_t, std::vector<uint32_t>> by_taxon;
  for (const auto &r : recs) by_taxon[r.first].push_back(r.second);
  for (auto &kv : by_taxon) {
    std::vector<uint32_t> &v = kv.second;
    const size_t n = v.size();
    const uint32_t last = v.back();
    const size_t occ_last = (size_t) std::count(v.begin(), v.end(), last);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    const size_t before_last = occ_last == 1 ? v.size() - 1 : v.size();
    Distinct d;
    d.n_observed = n;
    d.dense = before_last >= 1024;
    d.encoded = std::move(v);
    TaxonCounter &tc = total[kv.first];
    tc.n_kmers += n;
    tc.distinct.merge(d);
  }
}

// ---- Kraken-style report (src/reports.cc:130-259)
static void report_line(std::string &out, bool kmer_data, uint64_t total_seqs,
                        const TaxonCounter &clade, const TaxonCounter &taxon,
                        const std::string &rank_str, uint64_t taxid, const char *name,
                        int depth) {
  char pct[32];
  snprintf(pct, sizeof(pct), "%6.2f", 100.0 * clade.n_reads / total_seqs);
  out += pct;
  out += "\t" + std::to_string(clade.n_reads) + "\t" + std::to_string(taxon.n_reads) + "\t";
  if (kmer_data)
    out += std::to_string(clade.n_kmers) + "\t" + std::to_string(clade.distinct.estimate()) + "\t";
  out += rank_str + "\t" + std::to_string((uint32_t) taxid) + "\t";
  for (int i = 0; i < depth; i++) out += "  ";
  out += name;
  out += "\n";
}

static void dfs(std::string &out, uint64_t id, bool zeros, bool kmer_data, const Taxonomy &tax,
                const TaxonCounters &clade, const TaxonCounters &calls, uint64_t total_seqs,
                char rank_code, int rank_depth, int depth) {
  static const TaxonCounter empty;
  auto ci = clade.find(id);
  auto ti = calls.find(id);
  const TaxonCounter &cc = ci == clade.end() ? empty : ci->second;
  const TaxonCounter &tc = ti == calls.end() ? empty : ti->second;
  if (!zeros && cc.n_reads == 0) return;
  const Taxonomy::Node &node = tax.nodes[id];
  const std::string rank = tax.rank(id);
  static const struct {
    const char *name;
    char code;
  } major[] = {{"superkingdom", 'D'}, {"domain", 'D'}, {"kingdom", 'K'}, {"phylum", 'P'},
               {"class", 'C'},        {"order", 'O'},  {"family", 'F'},  {"genus", 'G'},
               {"species", 'S'}};
  bool is_major = false;
  for (const auto &m : major)
    if (rank == m.name) {
      rank_code = m.code;
      rank_depth = 0;
      is_major = true;
      break;
    }
  if (!is_major) rank_depth++;
  std::string rank_str(1, rank_code);
  if (rank_depth != 0) rank_str += std::to_string(rank_depth);
  report_line(out, kmer_data, total_seqs, cc, tc, rank_str, node.external_id, tax.name(id), depth);
  if (node.child_count == 0) return;
  std::vector<uint64_t> kids(node.child_count);
  for (uint64_t i = 0; i < node.child_count; i++) kids[i] = node.first_child + i;
  // children by descending clade read count, absent clades last; the same
  // std::sort call as the reference so equal counts keep the same order
  std::sort(kids.begin(), kids.end(), [&](const uint64_t &a, const uint64_t &b) {
    auto ia = clade.find(a);
    if (ia == clade.end()) return false;
    auto ib = clade.find(b);
    if (ib == clade.end()) return true;
    return ia->second.n_reads > ib->second.n_reads;
  });
  for (uint64_t k : kids)
    dfs(out, k, zeros, kmer_data, tax, clade, calls, total_seqs, rank_code, rank_depth, depth + 1);
}

bool report_kraken_style(const std::string &path, bool report_zeros, bool kmer_data,
                         const Taxonomy &tax, const TaxonCounters &calls, uint64_t total_seqs,
                         uint64_t total_unclassified) {
  TaxonCounters clade;
  for (const auto &kv : calls) {
    for (uint64_t t = kv.first; t; t = tax.nodes[t].parent_id) {
      TaxonCounter &c = clade[t];
      c.n_reads += kv.second.n_reads;
      c.n_kmers += kv.second.n_kmers;
      c.distinct.merge(kv.second.distinct);
    }
  }
  std::string out;
  if (total_unclassified != 0 || report_zeros) {
    TaxonCounter u;
    u.n_reads = total_unclassified;
    report_line(out, kmer_data, total_seqs, u, u, "U", 0, "unclassified", 0);
  }
  dfs(out, 1, report_zeros, kmer_data, tax, clade, calls, total_seqs, 'R', -1, 0);
  FILE *f = fopen(path.c_str(), "w");
  if (!f) return false;
  fwrite(out.data(), 1, out.size(), f);
  return fclose(f) == 0;
}

}  // namespace k2host
