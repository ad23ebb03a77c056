// k2cpu -- CPU ORACLE (test infrastructure only; see k2cpu.h for the
// parity-pinning map).  Restates the reference classify/build path in this
// repository's own code.
#include "k2cpu.h"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <fstream>
#include <iostream>
#include <limits>
#include <queue>
#include <sstream>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace k2cpu {

// ======================================================================
// opts.k2d  (src/classify.cc:272-279 reads at most file-size bytes into a
// zeroed struct; src/build_db.cc:94-109 writes the raw struct)
bool read_index_options(const std::string &path, IndexOptions &out, std::string &err) {
  std::memset(&out, 0, sizeof(out));
  struct stat sb;
  if (stat(path.c_str(), &sb) < 0) {
    err = "unable to get filesize of " + path;
    return false;
  }
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    err = "unable to open " + path;
    return false;
  }
  size_t want = std::min<size_t>(sizeof(out), (size_t) sb.st_size);
  size_t got = fread(&out, 1, want, f);
  fclose(f);
  (void) got;
  return true;
}

bool write_index_options(const std::string &path, const IndexOptions &o) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(&o, sizeof(o), 1, f) == 1;
  return fclose(f) == 0 && ok;
}

uint64_t expand_spaced_seed_mask(uint64_t mask, int factor) {
  // src/utilities.cc:14-27: template bit i (counting from the top usable
  // bit) becomes `factor` copies of itself.
  uint64_t out = 0;
  uint64_t ones = (1ULL << factor) - 1;
  for (int i = 64 / factor - 1; i >= 0; i--) {
    out <<= factor;
    if ((mask >> i) & 1) out |= ones;
  }
  return out;
}

uint64_t parse_bitstring(const char *s) { return (uint64_t) strtol(s, nullptr, 2); }

// ======================================================================
// Taxonomy
bool Taxonomy::load(const std::string &path, std::string &err) {
  // src/taxonomy.cc:217-237
  std::ifstream in(path, std::ios::binary);
  char magic[9] = {0};
  in.read(magic, 8);
  if (std::strcmp(magic, "K2TAXDAT") != 0) {
    err = "malformed taxonomy file " + path;
    return false;
  }
  uint64_t n = 0, name_len = 0, rank_len = 0;
  in.read((char *) &n, 8);
  in.read((char *) &name_len, 8);
  in.read((char *) &rank_len, 8);
  if (!in || n > (1ULL << 32)) {
    err = "read exhausted taxonomy information in " + path;
    return false;
  }
  nodes.resize(n);
  names.resize(name_len);
  ranks.resize(rank_len);
  in.read((char *) nodes.data(), n * sizeof(TaxNode));
  in.read(names.data(), name_len);
  in.read(ranks.data(), rank_len);
  if (!in) {
    err = "read exhausted taxonomy information in " + path;
    return false;
  }
  return true;
}

bool Taxonomy::write(const std::string &path) const {
  std::ofstream out(path, std::ios::binary);
  out.write("K2TAXDAT", 8);
  uint64_t n = nodes.size(), nl = names.size(), rl = ranks.size();
  out.write((const char *) &n, 8);
  out.write((const char *) &nl, 8);
  out.write((const char *) &rl, 8);
  out.write((const char *) nodes.data(), n * sizeof(TaxNode));
  out.write(names.data(), nl);
  out.write(ranks.data(), rl);
  return out.good();
}

// src/taxonomy.cc:251-258: ids are BFS-ordered, a parent precedes its child,
// so walk b upwards while it is numerically above a.
bool Taxonomy::is_ancestor(uint64_t a, uint64_t b) const {
  if (a == 0 || b == 0) return false;
  while (b > a) b = nodes[b].parent_id;
  return b == a;
}

// src/taxonomy.cc:262-272
uint64_t Taxonomy::lca(uint64_t a, uint64_t b) const {
  if (a == 0 || b == 0) return a ? a : b;
  while (a != b) {
    if (a > b)
      a = nodes[a].parent_id;
    else
      b = nodes[b].parent_id;
  }
  return a;
}

void Taxonomy::build_external_map() {  // src/taxonomy.cc:289-294
  ext_to_int.clear();
  ext_to_int[0] = 0;
  for (size_t i = 1; i < nodes.size(); i++) ext_to_int[nodes[i].external_id] = i;
}

uint64_t Taxonomy::internal_id(uint64_t external) const {
  auto it = ext_to_int.find(external);
  return it == ext_to_int.end() ? 0 : it->second;
}

// Splits a dmp line (trailing "\t|" already removed) on "\t|\t".
static std::vector<std::string> dmp_fields(const std::string &line, size_t max_fields) {
  std::vector<std::string> f;
  const std::string delim = "\t|\t";
  size_t p = 0;
  while (f.size() < max_fields) {
    size_t q = line.find(delim, p);
    if (q == std::string::npos) {
      f.push_back(line.substr(p));
      break;
    }
    f.push_back(line.substr(p, q - p));
    p = q + delim.size();
  }
  return f;
}

bool NcbiTaxonomy::parse(const std::string &nodes_path, const std::string &names_path,
                         std::string &err) {
  // src/taxonomy.cc:19-113
  std::ifstream nf(nodes_path), mf(names_path);
  if (!nf.good()) {
    err = "error opening " + nodes_path;
    return false;
  }
  if (!mf.good()) {
    err = "error opening " + names_path;
    return false;
  }
  std::string line;
  while (std::getline(nf, line)) {
    if (line.size() >= 2) line.resize(line.size() - 2);
    auto f = dmp_fields(line, 3);
    if (f.size() < 3) continue;
    uint64_t id = std::stoull(f[0]), parent = std::stoull(f[1]);
    if (id == 0) {
      err = "attempt to create taxonomy w/ node ID == 0";
      return false;
    }
    if (id == 1) parent = 0;
    parent_[id] = parent;
    children_[parent].insert(id);
    rank_[id] = f[2];
    known_ranks_.insert(f[2]);
  }
  while (std::getline(mf, line)) {
    if (line.size() >= 2) line.resize(line.size() - 2);
    auto f = dmp_fields(line, 4);
    if (f.size() < 4) continue;
    uint64_t id = std::stoull(f[0]);
    if (id == 0) {
      err = "attempt to create taxonomy w/ node ID == 0";
      return false;
    }
    if (f[3] == "scientific name") name_[id] = f[1];
  }
  return true;
}

void NcbiTaxonomy::mark(uint64_t taxid) {  // src/taxonomy.cc:116-121
  while (!marked_.count(taxid)) {
    marked_.insert(taxid);
    taxid = parent_[taxid];
  }
}

Taxonomy NcbiTaxonomy::convert() const {  // src/taxonomy.cc:123-185
  Taxonomy t;
  t.nodes.assign(marked_.size() + 1, TaxNode{0, 0, 0, 0, 0, 0, 0});
  std::string name_data, rank_data;
  std::map<std::string, uint64_t> rank_off;
  for (const auto &r : known_ranks_) {
    rank_off[r] = rank_data.size();
    rank_data.append(r.c_str(), r.size() + 1);
  }
  std::map<uint64_t, uint64_t> ext_to_int{{0, 0}, {1, 1}};
  std::queue<uint64_t> q;
  q.push(1);
  uint64_t next_id = 0;
  auto lookup = [](const auto &m, const auto &k) {
    auto it = m.find(k);
    return it == m.end() ? decltype(it->second){} : it->second;
  };
  while (!q.empty()) {
    uint64_t ext = q.front();
    q.pop();
    uint64_t id = ++next_id;
    ext_to_int[ext] = id;
    TaxNode n{0, 0, 0, 0, 0, 0, 0};
    n.parent_id = lookup(ext_to_int, lookup(parent_, ext));
    n.external_id = ext;
    auto rk = rank_.find(ext);
    n.rank_offset = lookup(rank_off, rk == rank_.end() ? std::string() : rk->second);
    n.name_offset = name_data.size();
    n.first_child = id + 1 + q.size();
    auto ch = children_.find(ext);
    if (ch != children_.end())
      for (uint64_t c : ch->second)
        if (marked_.count(c)) {
          q.push(c);
          n.child_count++;
        }
    t.nodes[id] = n;
    auto nm = name_.find(ext);
    std::string name = nm == name_.end() ? std::string() : nm->second;
    name_data.append(name.c_str(), name.size() + 1);
  }
  t.names.assign(name_data.begin(), name_data.end());
  t.ranks.assign(rank_data.begin(), rank_data.end());
  return t;
}

// ======================================================================
// Compact hash table
HashTable::HashTable(uint64_t cap, uint64_t kb, uint64_t vb)
    : capacity(cap), size(0), key_bits(kb), value_bits(vb), cells(cap, 0) {}

bool HashTable::load(const std::string &path, std::string &err) {
  // src/compact_hash.h:300-340 (non-mmap path)
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    err = "unable to open " + path;
    return false;
  }
  uint64_t hdr[4];
  if (fread(hdr, 8, 4, f) != 4) {
    fclose(f);
    err = "unexpected end of file reading " + path;
    return false;
  }
  capacity = hdr[0];
  size = hdr[1];
  key_bits = hdr[2];
  value_bits = hdr[3];
  if (key_bits + value_bits != 32) {
    fclose(f);
    err = "Unable to determine width of compact hash cell";
    return false;
  }
  cells.resize(capacity);
  size_t got = fread(cells.data(), 4, capacity, f);
  fclose(f);
  if (got != capacity) {
    err = "unexpected end of file reading " + path;
    return false;
  }
  return true;
}

bool HashTable::write(const std::string &path) const {  // src/compact_hash.h:343-352
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  uint64_t hdr[4] = {capacity, size, key_bits, value_bits};
  bool ok = fwrite(hdr, 8, 4, f) == 4 && fwrite(cells.data(), 4, capacity, f) == capacity;
  return fclose(f) == 0 && ok;
}

// src/compact_hash.h:354-374 with second_hash() == 1 (LINEAR_PROBING)
uint32_t HashTable::get(uint64_t key) const {
  uint64_t hc = fmix64(key);
  uint64_t ck = hc >> (64 - key_bits);
  uint32_t vmask = (1u << value_bits) - 1;
  uint64_t idx = hc % capacity, first = idx;
  for (;;) {
    uint32_t c = cells[idx];
    if ((c & vmask) == 0) break;
    if ((c >> value_bits) == ck) return c & vmask;
    if (++idx == capacity) idx = 0;
    if (idx == first) break;
  }
  return 0;
}

// src/compact_hash.h:379-392: prefetch the home cell 8 keys ahead
void HashTable::get_batch(const uint64_t *keys, uint32_t *out, size_t n) const {
  static const size_t dist = [] {
    const char *e = getenv("K2_PREFETCH_DIST");
    return e ? (size_t) atoi(e) : (size_t) 8;
  }();
  for (size_t i = 0; i < n; i++) {
    if (i + dist < n) __builtin_prefetch(&cells[fmix64(keys[i + dist]) % capacity], 0, 1);
    out[i] = get(keys[i]);
  }
}

uint32_t HashTable::probe_lines(uint64_t key) const {
  uint64_t hc = fmix64(key);
  uint64_t ck = hc >> (64 - key_bits);
  uint32_t vmask = (1u << value_bits) - 1;
  uint64_t idx = hc % capacity, first = idx;
  uint32_t lines = 1;
  uint64_t line = idx / 32;
  for (;;) {
    uint32_t c = cells[idx];
    if ((c & vmask) == 0) break;
    if ((c >> value_bits) == ck) break;
    if (++idx == capacity) idx = 0;
    if (idx == first) break;
    if (idx / 32 != line) {
      line = idx / 32;
      lines++;
    }
  }
  return lines;
}

bool HashTable::find_index(uint64_t key, uint64_t *idx) const {  // src/compact_hash.h:394-414
  uint64_t hc = fmix64(key);
  uint64_t ck = hc >> (64 - key_bits);
  uint32_t vmask = (1u << value_bits) - 1;
  *idx = hc % capacity;
  uint64_t first = *idx;
  for (;;) {
    uint32_t c = cells[*idx];
    if ((c & vmask) == 0) return false;
    if ((c >> value_bits) == ck) return true;
    if (++*idx == capacity) *idx = 0;
    if (*idx == first) break;
  }
  return false;
}

// src/compact_hash.h:416-462 (single writer; the reference's zone locks
// only serialise concurrent writers)
bool HashTable::compare_and_set(uint64_t key, uint32_t new_value, uint32_t *old_value) {
  if (new_value == 0) return false;
  uint64_t hc = fmix64(key);
  uint64_t ck = hc >> (64 - key_bits);
  uint32_t vmask = (1u << value_bits) - 1;
  uint64_t idx = hc % capacity, first = idx;
  for (;;) {
    uint32_t c = cells[idx];
    if ((c & vmask) == 0 || (c >> value_bits) == ck) {
      if (*old_value == (c & vmask)) {
        if ((uint64_t) new_value > vmask) {
          std::cerr << "value too wide for the cell\n";
          exit(70);
        }
        cells[idx] = (uint32_t) (ck << value_bits) | new_value;
        if (*old_value == 0) size++;
        return true;
      }
      *old_value = c & vmask;
      return false;
    }
    if (++idx == capacity) idx = 0;
    if (idx == first) {
      std::cerr << "compact hash table capacity exceeded\n";
      exit(70);
    }
  }
}

// ======================================================================
// Minimizer scanner
Scanner::Scanner(ssize_t k, ssize_t l, uint64_t spaced_seed_mask, bool dna, uint64_t toggle_mask,
                 int revcom_version)
    : k_(k), l_(l), spaced_(spaced_seed_mask), toggle_(toggle_mask), dna_(dna),
      revcom_version_(revcom_version) {
  int bits = dna ? 2 : 4;
  if (l_ > (ssize_t) (63 / bits)) {
    std::cerr << "l exceeds size limits for minimizer scanner\n";
    exit(70);
  }
  lmer_mask_ = (1ULL << (l_ * bits)) - 1;
  toggle_ &= lmer_mask_;
  size_t cap = 1;
  while (cap < (size_t) (k_ - l_ + 2)) cap <<= 1;
  ring_.resize(cap);
  std::memset(code_, 0xff, sizeof(code_));
  auto set = [&](char c, uint8_t v) {
    code_[(uint8_t) c] = v;
    code_[(uint8_t) tolower(c)] = v;
  };
  if (dna) {
    set('A', 0);
    set('C', 1);
    set('G', 2);
    set('T', 3);
  } else {
    // 15-letter reduced amino-acid alphabet (src/mmscanner.cc:41-95); one
    // string per code value 0..15.
    static const char *groups[16] = {"*UO", "A", "NQS", "C", "DE", "F", "G", "H",
                                     "IL",  "K", "P",   "R", "MV", "T", "W", "Y"};
    for (int v = 0; v < 16; v++)
      for (const char *p = groups[v]; *p; p++) set(*p, (uint8_t) v);
  }
}

void Scanner::load(const std::string &seq, size_t start, size_t finish) {
  // src/mmscanner.cc:97-111
  seq_ = &seq;
  start_ = start;
  finish_ = std::min(finish, seq.size());
  pos_ = start_;
  if ((ssize_t) (finish_ - start_) + 1 < l_) pos_ = finish_;
  count_ = 0;
  head_ = 0;
  window_len_ = 0;
  filled_ = 0;
  current_ = ~0ULL;
  recent_ambig_ = 0;
}

uint64_t Scanner::canonical(uint64_t x) const {
  // src/mmscanner.cc:193-220: reverse the 2-bit groups, complement, and
  // align the result to the low 2l bits.
  uint64_t r = __builtin_bswap64(x);
  r = ((r >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((r & 0x0F0F0F0F0F0F0F0FULL) << 4);
  r = ((r >> 2) & 0x3333333333333333ULL) | ((r & 0x3333333333333333ULL) << 2);
  uint64_t mask = (1ULL << (l_ * 2)) - 1;
  uint64_t rc = revcom_version_ == 0 ? (~r) & mask : ((~r) >> (64 - l_ * 2)) & mask;
  return x < rc ? x : rc;
}

const uint64_t *Scanner::next() {
  // src/mmscanner.cc:114-189
  if (pos_ >= finish_) return nullptr;
  const int bits = dna_ ? 2 : 4;
  const uint64_t ambig_code = (1ULL << bits) - 1;
  const size_t rmask = ring_.size() - 1;
  bool changed = false;
  do {
    if (filled_ == l_) filled_--;
    while (filled_ < l_ && pos_ < finish_) {
      filled_++;
      lmer_ <<= bits;
      recent_ambig_ <<= bits;
      uint8_t c = code_[(uint8_t) (*seq_)[pos_++]];
      if (c == 0xff) {
        count_ = 0;
        window_len_ = 0;
        lmer_ = 0;
        filled_ = 0;
        recent_ambig_ |= ambig_code;
      } else {
        lmer_ |= c;
      }
      lmer_ &= lmer_mask_;
      recent_ambig_ &= lmer_mask_;
      if ((pos_ - start_) >= (size_t) k_ && filled_ < l_) return &current_;
    }
    if (filled_ < l_) return nullptr;
    uint64_t cand = dna_ ? canonical(lmer_) : lmer_;
    if (spaced_) cand &= spaced_;
    cand ^= toggle_;
    if (k_ == l_) {
      current_ = cand ^ toggle_;
      return &current_;
    }
    // monotone deque: drop entries from the back that are strictly larger
    while (count_ && ring_[(head_ + count_ - 1) & rmask].cand > cand) count_--;
    if (count_ == 0 && window_len_ >= k_ - l_) changed = true;
    ring_[(head_ + count_) & rmask] = Entry{cand, window_len_};
    count_++;
    if (ring_[head_].pos < window_len_ - k_ + l_) {
      head_ = (head_ + 1) & rmask;
      count_--;
      changed = true;
    }
    if (window_len_ == k_ - l_) changed = true;
    window_len_++;
    if (pos_ >= (size_t) k_) break;
  } while (!changed);
  current_ = ring_[head_].cand ^ toggle_;
  return &current_;
}

// ======================================================================
// six-frame translation (src/aa_translate.cc:37-100).  Codons are indexed
// with A=0,G=1,C=2,T=3 (first base in the high bits).
void translate_all_frames(const std::string &dna, std::vector<std::string> &frames) {
  static const char table[] = "KKNNRRSSTTTTIMIIEEDDGGGGAAAAVVVVQQHHRRRRPPPPLLLL**YY*WCCSSSSLLFF";
  static uint8_t fwd[256], rev[256];
  static bool init = false;
  if (!init) {
    std::memset(fwd, 0xff, sizeof(fwd));
    std::memset(rev, 0xff, sizeof(rev));
    const char *b = "AGCT";
    for (int i = 0; i < 4; i++) {
      fwd[(uint8_t) b[i]] = fwd[(uint8_t) tolower(b[i])] = (uint8_t) i;
      // reverse strand: complement of base i placed in the top codon slot
      rev[(uint8_t) b[i]] = rev[(uint8_t) tolower(b[i])] = (uint8_t) ((3 - i) << 4);
    }
    init = true;
  }
  frames.resize(6);
  size_t max_size = dna.size() / 3 + 1;
  for (auto &f : frames) f.assign(max_size, ' ');
  if (dna.size() < 3) return;
  uint8_t fc = 0, rc = 0;
  int ambig = 0;
  size_t len[6] = {0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < dna.size(); i++) {
    size_t fr = i % 3;
    fc = (uint8_t) ((fc << 2) & 0x3f);
    rc >>= 2;
    if (ambig) ambig--;
    uint8_t f = fwd[(uint8_t) dna[i]], r = rev[(uint8_t) dna[i]];
    if (f == 0xff) {
      ambig = 3;
    } else {
      fc |= f;
      rc |= r;
    }
    if (i >= 2) {
      frames[fr][len[fr]++] = ambig == 0 ? table[fc] : 'X';
      frames[fr + 3][max_size - 1 - len[fr + 3]++] = ambig == 0 ? table[rc] : 'X';
    }
  }
  for (int i = 0; i < 3; i++) frames[i].resize(len[i]);
  for (int i = 3; i < 6; i++) frames[i].erase(0, frames[i].size() - len[i]);
}

// ======================================================================
// HyperLogLog++ (p = 12, sparse p' = 25, 64-bit hashes)
static const int kP = 12, kPPrime = 25;
static const uint32_t kM = 1u << kP;

uint64_t Hll::hash(uint64_t item) { return fmix64(item + 1); }  // hyperloglogplus.cc:816-824

static inline uint8_t clz64_or(uint64_t x, uint8_t none) { return x ? __builtin_clzll(x) : none; }
static inline uint8_t clz32_or(uint32_t x, uint8_t none) { return x ? __builtin_clz(x) : none; }

uint8_t Hll::dense_rank(uint64_t h) {  // getRank(uint64, p): hyperloglogplus.cc:124-131
  return clz64_or(h << kP, 64 - kP) + 1;
}

uint32_t Hll::encode_sparse(uint64_t h) {  // hyperloglogplus.cc:166-190
  uint32_t idx = (uint32_t) ((h >> (64 - kPPrime)) << (32 - kPPrime));
  if ((uint32_t) (idx << kP) == 0) {
    uint8_t extra = clz64_or(h << kPPrime, 64 - kPPrime) + 1;
    return idx | ((uint32_t) extra << 1) | 1;
  }
  return idx;
}

uint8_t Hll::encoded_rank(uint32_t e) {  // hyperloglogplus.cc:135-144
  if (e & 1) return (kPPrime - kP) + ((e >> 1) & 0x3f);
  return clz32_or(e << kP, 32 - kP) + 1;
}

void Hll::to_dense() {  // hyperloglogplus.cc:527-563
  if (!sparse) return;
  sparse = false;
  regs.assign(kM, 0);
  for (uint32_t e : sparse_list) {
    uint32_t idx = e >> (32 - kP);
    uint8_t r = encoded_rank(e);
    if (r > regs[idx]) regs[idx] = r;
  }
  sparse_list.clear();
}

void Hll::insert(uint64_t item) {  // hyperloglogplus.cc:471-509
  n_observed++;
  uint64_t h = hash(item);
  if (sparse && sparse_list.size() + 1 > kM / 4) to_dense();
  if (sparse) {
    sparse_list.insert(encode_sparse(h));
  } else {
    uint32_t idx = (uint32_t) (h >> (64 - kP));
    uint8_t r = dense_rank(h);
    if (r > regs[idx]) regs[idx] = r;
  }
}

void Hll::merge(const Hll &o) {  // hyperloglogplus.cc:613-651
  if (o.n_observed == 0) return;
  if (n_observed == 0) {
    *this = o;
    return;
  }
  n_observed += o.n_observed;
  if (sparse && o.sparse) {
    sparse_list.insert(o.sparse_list.begin(), o.sparse_list.end());
  } else if (o.sparse) {
    for (uint32_t e : o.sparse_list) {
      uint32_t idx = e >> (32 - kP);
      uint8_t r = encoded_rank(e);
      if (r > regs[idx]) regs[idx] = r;
    }
  } else if (sparse) {
    std::set<uint32_t> mine;
    mine.swap(sparse_list);
    sparse = false;
    regs = o.regs;
    for (uint32_t e : mine) {
      uint32_t idx = e >> (32 - kP);
      uint8_t r = encoded_rank(e);
      if (r > regs[idx]) regs[idx] = r;
    }
  } else {
    for (size_t i = 0; i < kM; i++)
      if (o.regs[i] > regs[i]) regs[i] = o.regs[i];
  }
}

// Ertl 2017 corrections (hyperloglogplus.cc:300-345); the loop structure is
// kept so the double results are bit-identical.
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

uint64_t Hll::estimate() const {  // hyperloglogplus.cc:708-739
  size_t q, m;
  std::vector<int> C;
  if (sparse) {
    q = 64 - kPPrime;
    m = (size_t) 1 << kPPrime;
    C.assign(q + 2, 0);
    size_t left = m;
    for (uint32_t e : sparse_list) {
      C[encoded_rank(e)]++;
      left--;
    }
    C[0] = (int) left;
  } else {
    q = 64 - kP;
    m = kM;
    C.assign(q + 2, 0);
    for (uint8_t r : regs) C[r]++;
  }
  double den = m * ertl_tau(1.0 - double(C[q + 1]) / double(m));
  for (int k = (int) q; k >= 1; --k) {
    den += C[k];
    den *= 0.5;
  }
  den += m * ertl_sigma(double(C[0]) / double(m));
  double m_sq_alpha_inf = (m / (2.0 * std::log(2))) * m;
  double est = m_sq_alpha_inf / den;
  return (n_observed < est) ? n_observed : (uint64_t) std::round(est);
}

// ======================================================================
// FASTA/FASTQ reader with kseq record semantics
bool SeqReader::open(const char *path) {
  if (!path) {
    fd_ = 0;
    own_ = false;
  } else {
    fd_ = ::open(path, O_RDONLY);
    own_ = true;
  }
  buf_.resize(1 << 16);
  return fd_ >= 0;
}

SeqReader::~SeqReader() {
  if (own_ && fd_ > 2) ::close(fd_);
}

int SeqReader::getc_() {
  if (beg_ >= end_) {
    if (eof_) return -1;
    ssize_t n = ::read(fd_, buf_.data(), buf_.size());
    if (n <= 0) {
      eof_ = true;
      return -1;
    }
    beg_ = 0;
    end_ = (size_t) n;
  }
  return (unsigned char) buf_[beg_++];
}

bool SeqReader::read(Sequence &s) {
  int c;
  if (pending_ == 0) {
    while ((c = getc_()) >= 0 && c != '>' && c != '@') {
    }
    if (c < 0) return false;
    pending_ = c;
  }
  s.header.clear();
  s.comment.clear();
  s.seq.clear();
  s.quals.clear();
  // name: up to the first whitespace character
  bool any = false;
  while ((c = getc_()) >= 0 && !isspace(c)) {
    s.header.push_back((char) c);
    any = true;
  }
  if (!any && c < 0) return false;
  auto read_line = [&](std::string &dst, bool append) {
    // rest of line into dst; one trailing '\r' dropped if dst has > 1 char
    if (!append) dst.clear();
    int ch;
    while ((ch = getc_()) >= 0 && ch != '\n') dst.push_back((char) ch);
    if (dst.size() > 1 && dst.back() == '\r') dst.pop_back();
    return ch;
  };
  if (c != '\n' && c >= 0) read_line(s.comment, false);
  while ((c = getc_()) >= 0 && c != '>' && c != '+' && c != '@') {
    if (c == '\n') continue;
    s.seq.push_back((char) c);
    read_line(s.seq, true);
  }
  if (c == '>' || c == '@') pending_ = c;
  if (c != '+') {
    s.fastq = false;
    if (c < 0) pending_ = 0;
    return true;
  }
  while ((c = getc_()) >= 0 && c != '\n') {
  }
  if (c < 0) return false;
  int r;
  do {
    r = read_line(s.quals, true);
  } while (r >= 0 && s.quals.size() < s.seq.size());
  pending_ = 0;
  if (s.quals.size() != s.seq.size()) return false;
  s.fastq = true;
  return true;
}

size_t SeqReader::load_block(size_t bytes, std::vector<Sequence> &out) {
  out.clear();
  size_t total = 0;
  while (total < bytes) {
    Sequence s;
    if (!read(s)) break;
    total += s.seq.size();
    out.push_back(std::move(s));
  }
  return out.size();
}

size_t SeqReader::load_batch(size_t n, std::vector<Sequence> &out) {
  out.clear();
  for (size_t i = 0; i < n; i++) {
    Sequence s;
    if (!read(s)) break;
    out.push_back(std::move(s));
  }
  return out.size();
}

std::string Sequence::to_string() const {
  std::string r;
  r.push_back(fastq ? '@' : '>');
  r += header;
  if (!comment.empty()) {
    r.push_back(' ');
    r += comment;
  }
  r.push_back('\n');
  r += seq;
  if (fastq) {
    r += "\n+\n";
    r += quals;
  }
  r.push_back('\n');
  return r;
}

// ======================================================================
// classify
bool load_index(const ClassifyOptions &o, Index &idx, std::string &err) {
  if (!read_index_options(o.options_filename, idx.opts, err)) return false;
  if (!idx.tax.load(o.taxonomy_filename, err)) return false;
  if (!idx.table.load(o.index_filename, err)) return false;
  return true;
}

enum TokKind : uint8_t { T_LOOKUP, T_SKIP, T_REPEAT, T_AMBIG, T_MATE, T_FRAME };
struct Tok {
  uint8_t kind;
  uint32_t key;
};

// src/classify.cc:710-762
uint64_t resolve_tree(std::unordered_map<uint64_t, uint64_t> &hit_counts, const Taxonomy &tax,
                      size_t total_minimizers, double confidence) {
  uint64_t max_taxon = 0;
  uint32_t max_score = 0;
  uint32_t required = (uint32_t) std::ceil(confidence * total_minimizers);
  for (const auto &a : hit_counts) {
    uint32_t score = 0;
    for (const auto &b : hit_counts)
      if (tax.is_ancestor(b.first, a.first)) score += (uint32_t) b.second;
    if (score > max_score) {
      max_score = score;
      max_taxon = a.first;
    } else if (score == max_score) {
      max_taxon = tax.lca(max_taxon, a.first);
    }
  }
  max_score = (uint32_t) hit_counts[max_taxon];
  while (max_taxon && max_score < required) {
    max_score = 0;
    for (const auto &b : hit_counts)
      if (tax.is_ancestor(max_taxon, b.first)) max_score += (uint32_t) b.second;
    if (max_score >= required) return max_taxon;
    max_taxon = tax.nodes[max_taxon].parent_id;
  }
  return max_taxon;
}

void classify_fragment(const Index &idx, const ClassifyOptions &o, const std::string &seq1,
                       const std::string *seq2, ReadResult &res, TaxonCounters *counters,
                       bool count_lines) {
  // scan phase (src/classify.cc:788-838)
  thread_local std::vector<Tok> toks;
  thread_local std::vector<uint64_t> keys;
  thread_local std::vector<uint32_t> vals;
  thread_local std::vector<std::string> frames(6);
  toks.clear();
  keys.clear();
  const IndexOptions &io = idx.opts;
  Scanner sc((ssize_t) io.k, (ssize_t) io.l, io.spaced_seed_mask, io.dna_db, io.toggle_mask,
             io.revcom_version);
  const bool tx = o.use_translated_search;
  const int nframes = tx ? 6 : 1;
  const bool paired = o.paired_end_processing;
  for (int mate = 0; mate < (paired ? 2 : 1); mate++) {
    const std::string &s = mate == 0 ? seq1 : *seq2;
    if (tx) translate_all_frames(s, frames);
    for (int f = 0; f < nframes; f++) {
      sc.load(tx ? frames[f] : s);
      uint64_t last = UINT64_MAX;
      const uint64_t *m;
      while ((m = sc.next()) != nullptr) {
        if (sc.ambiguous()) {
          toks.push_back({T_AMBIG, 0});
        } else if (*m != last) {
          last = *m;
          if (io.minimum_acceptable_hash_value && fmix64(*m) < io.minimum_acceptable_hash_value) {
            toks.push_back({T_SKIP, 0});
          } else {
            toks.push_back({T_LOOKUP, (uint32_t) keys.size()});
            keys.push_back(*m);
          }
        } else {
          toks.push_back({T_REPEAT, 0});
        }
      }
      if (tx && f != 5) toks.push_back({T_FRAME, 0});
    }
    if (paired && mate == 0) toks.push_back({T_MATE, 0});
  }
  // probe phase (src/classify.cc:840-844)
  vals.resize(keys.size());
  if (!keys.empty()) idx.table.get_batch(keys.data(), vals.data(), keys.size());
  res.lookups = (uint32_t) keys.size();
  res.probe_lines = 0;
  if (count_lines)
    for (uint64_t k : keys) res.probe_lines += idx.table.probe_lines(k);
  // replay phase (src/classify.cc:846-888)
  std::unordered_map<uint64_t, uint64_t> hit_counts;
  res.taxa.clear();
  res.call = 0;
  int64_t groups = 0;
  uint64_t last_taxon = 0;
  for (const Tok &t : toks) {
    uint64_t taxon = 0;
    if (t.kind == T_AMBIG) {
      res.taxa.push_back(kAmbigSpan);
      continue;
    } else if (t.kind == T_FRAME) {
      res.taxa.push_back(kFrameBorder);
      continue;
    } else if (t.kind == T_MATE) {
      res.taxa.push_back(kMateBorder);
      continue;
    } else if (t.kind == T_SKIP) {
      taxon = 0;
      last_taxon = 0;
    } else if (t.kind == T_LOOKUP) {
      taxon = vals[t.key];
      last_taxon = taxon;
      if (taxon) {
        groups++;
        if (counters) (*counters)[taxon].add_kmer(keys[t.key]);
      }
    } else {
      taxon = last_taxon;
    }
    if (taxon) {
      if (o.quick_mode && groups >= o.minimum_hit_groups) break;
      hit_counts[taxon]++;
    }
    res.taxa.push_back(taxon);
  }
  // call (src/classify.cc:890-906)
  uint64_t total = res.taxa.size();
  if (paired) total--;
  if (tx) total -= paired ? 4 : 2;
  res.total_kmers = total;
  uint64_t call = resolve_tree(hit_counts, idx.tax, total, o.confidence_threshold);
  if (call && groups < o.minimum_hit_groups) call = 0;
  if (call && counters) (*counters)[call].n_reads++;
  res.call = call;
  res.hit_groups = groups;
}

std::string trim_pair_info(const std::string &id) {  // src/classify.cc:764-771
  size_t n = id.size();
  if (n <= 2) return id;
  if (id[n - 2] == '/' && (id[n - 1] == '1' || id[n - 1] == '2')) return id.substr(0, n - 2);
  return id;
}

void append_hitlist(std::string &out, const std::vector<uint64_t> &taxa, const Taxonomy &tax) {
  // src/classify.cc:950-991
  auto emit = [&](uint64_t code, uint64_t count, bool last) {
    if (code == kMateBorder) {
      out += last ? "|:|" : "|:| ";
    } else if (code == kFrameBorder) {
      out += last ? "-:-" : "-:- ";
    } else if (code == kAmbigSpan) {
      out += "A:" + std::to_string(count) + " ";
    } else {
      out += std::to_string(tax.nodes[code].external_id) + ":" + std::to_string(count);
      if (!last) out += " ";
    }
  };
  uint64_t code = taxa[0], count = 1;
  for (size_t i = 1; i < taxa.size(); i++) {
    if (taxa[i] == code) {
      count++;
    } else {
      emit(code, count, false);
      code = taxa[i];
      count = 1;
    }
  }
  emit(code, count, true);
}

void format_line(std::string &out, const Index &idx, const ClassifyOptions &o, const Sequence &s1,
                 const Sequence *s2, const ReadResult &r) {
  // src/classify.cc:908-945
  out += r.call ? "C\t" : "U\t";
  out += o.paired_end_processing ? trim_pair_info(s1.header) : s1.header;
  out += "\t";
  uint64_t ext = idx.tax.nodes[r.call].external_id;
  if (o.print_scientific_name) {
    out += r.call ? idx.tax.name(r.call) : "unclassified";
    out += " (taxid " + std::to_string(ext) + ")";
  } else {
    out += std::to_string(ext);
  }
  out += "\t";
  out += std::to_string(s1.seq.size());
  if (o.paired_end_processing) out += "|" + std::to_string(s2->seq.size());
  out += "\t";
  if (o.quick_mode)
    out += std::to_string(ext) + ":Q";
  else if (r.taxa.empty())
    out += "0:0";
  else
    append_hitlist(out, r.taxa, idx.tax);
  out += "\n";
}

static void mask_low_quality(Sequence &s, int min_q) {  // src/classify.cc:1061-1071
  if (!s.fastq) return;
  if (s.seq.size() != s.quals.size()) {
    std::cerr << "classify: " << s.header << ": Sequence length (" << s.seq.size()
              << ") != Quality string length (" << s.quals.size() << ")\n";
    exit(65);
  }
  for (size_t i = 0; i < s.seq.size(); i++)
    if ((s.quals[i] - '!') < min_q) s.seq[i] = 'x';
}

static std::vector<std::string> split_hash_name(const std::string &f) {
  std::vector<std::string> parts;
  size_t p = 0;
  for (;;) {
    size_t q = f.find('#', p);
    if (q == std::string::npos || parts.size() == 2) {
      parts.push_back(f.substr(p));
      break;
    }
    parts.push_back(f.substr(p, q - p));
    p = q + 1;
  }
  return parts;
}

struct Outputs {
  FILE *kraken = stdout;
  FILE *c1 = nullptr, *c2 = nullptr, *u1 = nullptr, *u2 = nullptr;
};

static FILE *open_or_die(const std::string &path) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    std::cerr << "\rUnable to open file: " << path << ", reason: " << strerror(errno) << "\n";
    exit(EXIT_FAILURE);
  }
  return f;
}

static void open_outputs(const ClassifyOptions &o, Outputs &out) {  // src/classify.cc:1008-1059
  auto pair_names = [&](const std::string &fmt, FILE *&a, FILE *&b) {
    auto parts = split_hash_name(fmt);
    if (parts.size() < 2) {
      std::cerr << "classify: Paired filename format missing # character: " << fmt << "\n";
      exit(65);
    }
    if (parts.size() > 2) {
      std::cerr << "classify: Paired filename format has >1 # character: " << fmt << "\n";
      exit(65);
    }
    a = open_or_die(parts[0] + "_1" + parts[1]);
    b = open_or_die(parts[0] + "_2" + parts[1]);
  };
  if (!o.classified_output_filename.empty()) {
    if (o.paired_end_processing)
      pair_names(o.classified_output_filename, out.c1, out.c2);
    else
      out.c1 = open_or_die(o.classified_output_filename);
  }
  if (!o.unclassified_output_filename.empty()) {
    if (o.paired_end_processing)
      pair_names(o.unclassified_output_filename, out.u1, out.u2);
    else
      out.u1 = open_or_die(o.unclassified_output_filename);
  }
  if (!o.kraken_output_filename.empty()) {
    if (o.kraken_output_filename == "-")
      out.kraken = nullptr;
    else
      out.kraken = open_or_die(o.kraken_output_filename);
  }
}

struct BlockWork {
  std::vector<Sequence> a, b;
  std::string kraken, c1, c2, u1, u2;
  ClassifyStats stats;
  TaxonCounters counters;
};

static void process_block(const Index &idx, const ClassifyOptions &o, BlockWork &w) {
  const bool paired = o.paired_end_processing;
  const bool report = !o.report_filename.empty();
  ReadResult r;
  size_t n1 = w.a.size();
  size_t i = 0, j = 0;
  for (;;) {
    if (i >= n1) break;
    Sequence *s1 = &w.a[i++];
    Sequence *s2 = nullptr;
    if (paired) {
      if (o.single_file_pairs) {
        if (i >= n1) break;
        s2 = &w.a[i++];
      } else {
        if (j >= w.b.size()) break;
        s2 = &w.b[j++];
      }
    }
    w.stats.total_sequences++;
    if (o.minimum_quality_score > 0) {
      mask_low_quality(*s1, o.minimum_quality_score);
      if (paired) mask_low_quality(*s2, o.minimum_quality_score);
    }
    classify_fragment(idx, o, s1->seq, s2 ? &s2->seq : nullptr, r, report ? &w.counters : nullptr);
    if (r.call) w.stats.total_classified++;
    format_line(w.kraken, idx, o, *s1, s2, r);
    if (r.call) {
      char buf[64];
      snprintf(buf, sizeof(buf), " kraken:taxid|%llu",
               (unsigned long long) idx.tax.nodes[r.call].external_id);
      if (!o.classified_output_filename.empty()) {
        Sequence t = *s1;
        t.header += buf;
        w.c1 += t.to_string();
        if (paired) {
          Sequence t2 = *s2;
          t2.header += buf;
          w.c2 += t2.to_string();
        }
      }
    } else if (!o.unclassified_output_filename.empty()) {
      w.u1 += s1->to_string();
      if (paired) w.u2 += s2->to_string();
    }
    w.stats.total_bases += s1->seq.size();
    if (paired) w.stats.total_bases += s2->seq.size();
  }
}

static int process_files(const char *f1, const char *f2, const Index &idx,
                         const ClassifyOptions &o, ClassifyStats &stats, Outputs &out,
                         TaxonCounters &total) {
  // src/classify.cc:479-708: blocks of 3 MiB (single), 10 000 pairs from
  // two files, or 20 000 interleaved records; output in block order.
  SeqReader r1, r2;
  if (!r1.open(f1)) {
    std::cerr << "classify: unable to open " << (f1 ? f1 : "stdin") << "\n";
    return 66;
  }
  if (f2 && !r2.open(f2)) {
    std::cerr << "classify: unable to open " << f2 << "\n";
    return 66;
  }
  int threads = std::max(1, o.num_threads);
  bool done = false;
  while (!done) {
    std::vector<BlockWork> work(threads);
    int nb = 0;
    for (; nb < threads; nb++) {
      bool ok;
      if (!o.paired_end_processing) {
        ok = r1.load_block(3 * 1024 * 1024, work[nb].a) > 0;
      } else if (!o.single_file_pairs) {
        ok = r1.load_batch(10000, work[nb].a) > 0;
        if (ok) ok = r2.load_batch(10000, work[nb].b) > 0;
      } else {
        ok = r1.load_batch(20000, work[nb].a) > 0;
      }
      if (!ok) {
        done = true;
        break;
      }
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int b = 0; b < nb; b++) process_block(idx, o, work[b]);
    for (int b = 0; b < nb; b++) {
      BlockWork &w = work[b];
      stats.total_sequences += w.stats.total_sequences;
      stats.total_bases += w.stats.total_bases;
      stats.total_classified += w.stats.total_classified;
      if (out.kraken) fwrite(w.kraken.data(), 1, w.kraken.size(), out.kraken);
      if (out.c1) fwrite(w.c1.data(), 1, w.c1.size(), out.c1);
      if (out.c2) fwrite(w.c2.data(), 1, w.c2.size(), out.c2);
      if (out.u1) fwrite(w.u1.data(), 1, w.u1.size(), out.u1);
      if (out.u2) fwrite(w.u2.data(), 1, w.u2.size(), out.u2);
      if (!o.report_filename.empty())
        for (auto &kv : w.counters) total[kv.first].add(kv.second);
    }
  }
  return 0;
}

int run_classify(const ClassifyOptions &o, const Index &idx, ClassifyStats &stats,
                 TaxonCounters &counters, double *seconds) {
  Outputs out;
  open_outputs(o, out);
  struct timeval t1, t2;
  gettimeofday(&t1, nullptr);
  int rc = 0;
  if (o.filenames.empty()) {
    if (o.paired_end_processing && !o.single_file_pairs) {
      std::cerr << "classify: paired end processing used with no files specified\n";
      return 64;
    }
    rc = process_files(nullptr, nullptr, idx, o, stats, out, counters);
  } else {
    for (size_t i = 0; i < o.filenames.size() && rc == 0; i++) {
      if (o.paired_end_processing && !o.single_file_pairs) {
        if (i + 1 == o.filenames.size()) {
          std::cerr << "classify: paired end processing used with unpaired file\n";
          return 64;
        }
        rc = process_files(o.filenames[i].c_str(), o.filenames[i + 1].c_str(), idx, o, stats, out,
                           counters);
        i++;
      } else {
        rc = process_files(o.filenames[i].c_str(), nullptr, idx, o, stats, out, counters);
      }
    }
  }
  for (FILE *f : {out.kraken, out.c1, out.c2, out.u1, out.u2})
    if (f) fflush(f);
  gettimeofday(&t2, nullptr);
  double s = (t2.tv_sec - t1.tv_sec) + (t2.tv_usec - t1.tv_usec) / 1e6;
  if (seconds) *seconds = s;
  for (FILE *f : {out.c1, out.c2, out.u1, out.u2})
    if (f) fclose(f);
  if (out.kraken && out.kraken != stdout) fclose(out.kraken);
  return rc;
}

void report_stats(double seconds, const ClassifyStats &s) {  // src/classify.cc:447-477
  uint64_t unc = s.total_sequences - s.total_classified;
  if (isatty(fileno(stderr))) std::cerr << "\r";
  fprintf(stderr, "%llu sequences (%.2f Mbp) processed in %.3fs (%.1f Kseq/m, %.2f Mbp/m).\n",
          (unsigned long long) s.total_sequences, s.total_bases / 1.0e6, seconds,
          s.total_sequences / 1.0e3 / (seconds / 60), s.total_bases / 1.0e6 / (seconds / 60));
  fprintf(stderr, "  %llu sequences classified (%.2f%%)\n",
          (unsigned long long) s.total_classified, s.total_classified * 100.0 / s.total_sequences);
  fprintf(stderr, "  %llu sequences unclassified (%.2f%%)\n", (unsigned long long) unc,
          unc * 100.0 / s.total_sequences);
}

// ======================================================================
// kreport
static void report_line(std::string &out, bool kmer_data, uint64_t total_seqs,
                        const ReadCounter &clade, const ReadCounter &taxon,
                        const std::string &rank_str, uint64_t taxid, const std::string &name,
                        int depth) {
  // src/reports.cc:130-149
  char pct[32];
  snprintf(pct, sizeof(pct), "%6.2f", 100.0 * clade.n_reads / total_seqs);
  out += pct;
  out += "\t" + std::to_string(clade.n_reads) + "\t" + std::to_string(taxon.n_reads) + "\t";
  if (kmer_data)
    out += std::to_string(clade.n_kmers) + "\t" + std::to_string(clade.kmers.estimate()) + "\t";
  out += rank_str + "\t" + std::to_string((uint32_t) taxid) + "\t";
  for (int i = 0; i < depth; i++) out += "  ";
  out += name + "\n";
}

static void report_dfs(std::string &out, uint64_t taxid, bool zeros, bool kmer_data,
                       const Taxonomy &tax, const TaxonCounters &clade,
                       const TaxonCounters &calls, uint64_t total_seqs, char rank_code,
                       int rank_depth, int depth) {
  // src/reports.cc:152-238
  static const ReadCounter empty;
  auto ci = clade.find(taxid);
  auto ti = calls.find(taxid);
  const ReadCounter &cc = ci == clade.end() ? empty : ci->second;
  const ReadCounter &tc = ti == calls.end() ? empty : ti->second;
  if (!zeros && cc.n_reads == 0) return;
  const TaxNode &node = tax.nodes[taxid];
  std::string rank = tax.rank(taxid);
  static const std::pair<const char *, char> major[] = {
      {"superkingdom", 'D'}, {"domain", 'D'}, {"kingdom", 'K'}, {"phylum", 'P'}, {"class", 'C'},
      {"order", 'O'},        {"family", 'F'}, {"genus", 'G'},   {"species", 'S'}};
  bool is_major = false;
  for (const auto &m : major)
    if (rank == m.first) {
      rank_code = m.second;
      rank_depth = 0;
      is_major = true;
      break;
    }
  if (!is_major) rank_depth++;
  std::string rank_str(1, rank_code);
  if (rank_depth != 0) rank_str += std::to_string(rank_depth);
  report_line(out, kmer_data, total_seqs, cc, tc, rank_str, node.external_id, tax.name(taxid),
              depth);
  if (node.child_count == 0) return;
  std::vector<uint64_t> kids(node.child_count);
  for (uint64_t i = 0; i < node.child_count; i++) kids[i] = node.first_child + i;
  // same comparator and the same std::sort as the reference, so ties land
  // in the same (unstable) order
  std::sort(kids.begin(), kids.end(), [&](const uint64_t &a, const uint64_t &b) {
    auto ia = clade.find(a);
    if (ia == clade.end()) return false;
    auto ib = clade.find(b);
    if (ib == clade.end()) return true;
    return ia->second.n_reads > ib->second.n_reads;
  });
  for (uint64_t k : kids)
    report_dfs(out, k, zeros, kmer_data, tax, clade, calls, total_seqs, rank_code, rank_depth,
               depth + 1);
}

void report_kraken_style(const std::string &path, bool report_zeros, bool report_kmer_data,
                         const Taxonomy &tax, const TaxonCounters &calls, uint64_t total_seqs,
                         uint64_t total_unclassified) {
  // src/reports.cc:32-45, 241-259
  TaxonCounters clade;
  for (const auto &kv : calls) {
    uint64_t t = kv.first;
    while (t) {
      clade[t].add(kv.second);
      t = tax.nodes[t].parent_id;
    }
  }
  std::string out;
  if (total_unclassified != 0 || report_zeros) {
    ReadCounter rc;
    rc.n_reads = total_unclassified;
    report_line(out, report_kmer_data, total_seqs, rc, rc, "U", 0, "unclassified", 0);
  }
  report_dfs(out, 1, report_zeros, report_kmer_data, tax, clade, calls, total_seqs, 'R', -1, 0);
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    std::cerr << "unable to write report " << path << "\n";
    return;
  }
  fwrite(out.data(), 1, out.size(), f);
  fclose(f);
}

// ======================================================================
// build
std::vector<std::string> extract_ncbi_sequence_ids(const std::string &header) {
  // src/build_db.cc:115-146: \x01 separates IDs; whitespace ends an ID
  std::vector<std::string> ids;
  std::string cur;
  bool in_id = true;
  for (char ch : header) {
    if (ch == 0x01) {
      if (!cur.empty()) ids.push_back(cur);
      cur.clear();
      in_id = true;
    } else if (in_id && isspace((unsigned char) ch)) {
      if (!cur.empty()) ids.push_back(cur);
      cur.clear();
      in_id = false;
    } else if (in_id) {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) ids.push_back(cur);
  return ids;
}

void block_minimizers(const BuildOptions &o, const std::string &seq, size_t block_start,
                      std::vector<uint64_t> &ordered) {
  // src/build_db.h:226-275: subblocks of the block, each overlapping the
  // next by k-1, scanned independently; distinct non-ambiguous minimizers
  // grouped by fmix64 % 256, each group in ascending order.
  size_t block_finish = std::min(block_start + o.block_size + o.k - 1, seq.size());
  std::vector<uint64_t> zone_keys[256];
  Scanner sc(o.k, o.l, o.spaced_seed_mask, !o.input_is_protein, o.toggle_mask);
  for (size_t i = block_start; i < block_finish; i += o.subblock_size) {
    size_t sub_finish = std::min(i + o.subblock_size + o.k - 1, block_finish);
    sc.load(seq, i, sub_finish);
    const uint64_t *m;
    while ((m = sc.next()) != nullptr) {
      if (sc.ambiguous()) continue;
      uint64_t hc = fmix64(*m);
      if (o.min_clear_hash_value && hc < o.min_clear_hash_value) continue;
      zone_keys[hc % 256].push_back(*m);
    }
  }
  ordered.clear();
  for (auto &z : zone_keys) {
    std::sort(z.begin(), z.end());
    z.erase(std::unique(z.begin(), z.end()), z.end());
    ordered.insert(ordered.end(), z.begin(), z.end());
  }
}

static void set_minimizer_lca(HashTable &h, uint64_t m, uint64_t taxid, const Taxonomy &tax) {
  // src/build_db.h:186-194
  uint32_t old_value = 0, new_value = (uint32_t) taxid;
  while (!h.compare_and_set(m, new_value, &old_value))
    new_value = (uint32_t) tax.lca(old_value, taxid);
}

int run_build(BuildOptions &o) {
  // src/build_db.cc:40-113
  std::map<std::string, uint64_t> id_map;
  {
    std::ifstream mf(o.id_map_filename);
    if (!mf.good()) {
      std::cerr << "build_db: unable to read from '" << o.id_map_filename << "'\n";
      return 66;
    }
    std::string line;
    while (std::getline(mf, line)) {
      std::istringstream iss(line);
      std::string id;
      uint64_t taxid = 0;
      iss >> id >> taxid;
      if (taxid) id_map[id] = taxid;
    }
  }
  {
    NcbiTaxonomy ncbi;
    std::string err;
    if (!ncbi.parse(o.ncbi_taxonomy_directory + "/nodes.dmp",
                    o.ncbi_taxonomy_directory + "/names.dmp", err)) {
      std::cerr << "build_db: " << err << "\n";
      return 65;
    }
    for (auto &kv : id_map)
      if (kv.second) ncbi.mark(kv.second);
    if (!ncbi.convert().write(o.taxonomy_filename)) {
      std::cerr << "build_db: error writing taxonomy\n";
      return 74;
    }
  }
  std::cerr << "Taxonomy parsed and converted." << std::endl;
  Taxonomy tax;
  std::string err;
  if (!tax.load(o.taxonomy_filename, err)) {
    std::cerr << "build_db: " << err << "\n";
    return 65;
  }
  tax.build_external_map();
  size_t vbits = 1;
  while ((1ULL << vbits) < tax.node_count()) vbits++;
  if (o.requested_bits_for_taxid > 0 && vbits > o.requested_bits_for_taxid) {
    std::cerr << "build_db: more bits required for storing taxid\n";
    return 65;
  }
  if (vbits < o.requested_bits_for_taxid) vbits = o.requested_bits_for_taxid;
  size_t capacity = o.capacity;
  if (o.maximum_capacity) {
    double frac = o.maximum_capacity * 1.0 / o.capacity;
    if (frac > 1) {
      std::cerr << "build_db: maximum capacity larger than requested capacity\n";
      return 65;
    }
    o.min_clear_hash_value = (uint64_t) ((1 - frac) * UINT64_MAX);
    capacity = o.maximum_capacity;
  }
  HashTable table(capacity, 32 - vbits, vbits);
  std::cerr << "CHT created with " << vbits << " bits reserved for taxid." << std::endl;

  SeqReader reader;
  reader.open(nullptr);
  std::vector<Sequence> block;
  std::vector<uint64_t> ordered;
  size_t seq_ct = 0, ch_ct = 0;
  Scanner fast_sc(o.k, o.l, o.spaced_seed_mask, !o.input_is_protein, o.toggle_mask);
  while (reader.load_block(o.block_size, block) > 0) {
    for (auto &s : block) {
      uint64_t taxid = 0;
      for (auto &id : extract_ncbi_sequence_ids(s.header)) {
        auto it = id_map.find(id);
        if (it == id_map.end() || it->second == 0) continue;
        taxid = tax.lca(taxid, tax.internal_id(it->second));
      }
      if (!taxid) continue;
      if (o.input_is_protein && (s.seq.empty() || s.seq.back() != '*')) s.seq.push_back('*');
      if (o.deterministic) {
        for (size_t j = 0; j < s.seq.size(); j += o.block_size) {
          block_minimizers(o, s.seq, j, ordered);
          for (uint64_t m : ordered) set_minimizer_lca(table, m, taxid, tax);
        }
      } else {
        // src/build_db.h:196-214
        fast_sc.load(s.seq);
        const uint64_t *m;
        while ((m = fast_sc.next()) != nullptr) {
          if (fast_sc.ambiguous()) continue;
          if (o.min_clear_hash_value && fmix64(*m) < o.min_clear_hash_value) continue;
          uint32_t existing = 0, nt = (uint32_t) taxid;
          while (!table.compare_and_set(*m, nt, &existing)) nt = (uint32_t) tax.lca(nt, existing);
        }
      }
      seq_ct++;
      ch_ct += s.seq.size();
    }
  }
  std::cerr << "Completed processing of " << seq_ct << " sequences, " << ch_ct << " "
            << (o.input_is_protein ? "aa" : "bp") << std::endl;
  std::cerr << "Writing data to disk... " << std::flush;
  if (!table.write(o.hashtable_filename)) {
    std::cerr << "build_db: unable to write " << o.hashtable_filename << "\n";
    return 74;
  }
  IndexOptions io;
  std::memset(&io, 0, sizeof(io));
  io.k = o.k;
  io.l = o.l;
  io.spaced_seed_mask = o.spaced_seed_mask;
  io.toggle_mask = o.toggle_mask;
  io.dna_db = !o.input_is_protein;
  io.minimum_acceptable_hash_value = o.min_clear_hash_value;
  io.revcom_version = 1;
  io.db_version = 0;
  io.db_type = 0;
  if (!write_index_options(o.options_filename, io)) {
    std::cerr << "build_db: Unable to write options file " << o.options_filename << "\n";
    return 74;
  }
  std::cerr << " complete." << std::endl;
  return 0;
}

uint64_t estimate_capacity(const EstimateOptions &o, const char *path) {
  // src/estimate_capacity.cc:66-100, 185-203
  SeqReader reader;
  reader.open(path);
  std::vector<std::unordered_map<uint64_t, char>> dummy;
  std::vector<std::vector<uint64_t>> sets(o.n);
  Scanner sc(o.k, o.l, o.spaced_seed_mask, !o.input_is_protein, o.toggle_mask);
  Sequence s;
  while (reader.read(s)) {
    if (o.input_is_protein && (s.seq.empty() || s.seq.back() != '*')) s.seq.push_back('*');
    sc.load(s.seq);
    const uint64_t *m;
    while ((m = sc.next()) != nullptr) {
      if (sc.ambiguous()) continue;
      uint64_t hc = fmix64(*m);
      if ((hc & 1023) < o.n) sets[hc & 1023].push_back(*m);
    }
  }
  size_t sum = 0;
  for (auto &v : sets) {
    std::sort(v.begin(), v.end());
    sum += std::unique(v.begin(), v.end()) - v.begin();
  }
  sum++;
  return (uint64_t) (sum * 1024 * 1.0 / o.n);
}

}  // namespace k2cpu
