// Test double of the library entry points the classify front-end calls
// (kraken2_amd/csrc/host/classify_main.cc), so the front-end's host code
// (parsing, blocks and batches, the reader thread and GPU-call threads,
// -Q masking, output formatting, -C/-U records, the report) runs without a
// GPU, built with AddressSanitizer and UndefinedBehaviorSanitizer.  It does
// not classify: each fragment's results are a fixed function of its bases
// that the tests recompute, so a result handed to the wrong fragment, a
// dropped or repeated fragment, or a changed base shows in the output.
//   call: internal id 1 when bit 7 of the CRC-32 of the fragment's bases
//         (mate 1, then mate 2) is set, else 0
//   runs: (0, e1) for mate 1 if it has k-mers, then for a pair the mate
//         border and (0, e2) if mate 2 has k-mers (e = bases - k + 1)
//   hit groups 0, total k-mers e1 + e2, no HLL records, k-mer counts 0.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "k2g.h"

namespace {
thread_local std::string g_error;

int fail(int rc, const std::string &msg) {
  g_error = msg;
  return rc;
}

bool slurp(const char *path, std::vector<char> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

template <class T>
T field(const std::vector<char> &b, size_t at) {
  T v{};
  if (at + sizeof(T) <= b.size()) memcpy(&v, b.data() + at, sizeof(T));
  return v;
}

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320), bit by bit
uint32_t crc32(const char *p, uint64_t n) {
  uint32_t c = 0xFFFFFFFFu;
  for (uint64_t i = 0; i < n; i++) {
    c ^= (uint8_t) p[i];
    for (int b = 0; b < 8; b++) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
  }
  return ~c;
}
}  // namespace

struct k2g_db {
  k2g_db_info info{};
};

struct k2g_batch {
  k2g_db *db = nullptr;
  std::vector<uint32_t> call, groups, total, runs;
  std::vector<uint64_t> run_off, hll_off, kmer_counts;
  std::vector<uint32_t> hll;
};

extern "C" {

const char *k2g_last_error(void) { return g_error.c_str(); }

int k2g_host_alloc(size_t bytes, void **out) {
  *out = malloc(bytes ? bytes : 1);
  return *out ? K2G_OK : fail(K2G_EOS, "host allocation");
}

int k2g_host_free(void *p) {
  free(p);
  return K2G_OK;
}

int k2g_db_open(const char *hash_k2d, const char *taxo_k2d, const char *opts_k2d, k2g_db **out) {
  std::vector<char> h, t, o;
  if (!slurp(hash_k2d, h) || !slurp(taxo_k2d, t) || !slurp(opts_k2d, o) || o.size() < 48)
    return fail(K2G_EIO, "unable to read the database files");
  if (!o[32]) return fail(K2G_EUSAGE, "the test double takes nucleotide databases only");
  auto *db = new k2g_db();
  k2g_db_info &in = db->info;
  in.capacity = field<uint64_t>(h, 0);
  in.size = field<uint64_t>(h, 8);
  in.key_bits = field<uint64_t>(h, 16);
  in.value_bits = field<uint64_t>(h, 24);
  in.node_count = field<uint64_t>(t, 8);
  in.k = field<uint64_t>(o, 0);
  in.l = field<uint64_t>(o, 8);
  in.spaced_seed_mask = field<uint64_t>(o, 16);
  in.toggle_mask = field<uint64_t>(o, 24);
  in.minimum_acceptable_hash_value = field<uint64_t>(o, 40);
  in.dna_db = 1;
  *out = db;
  return K2G_OK;
}

int k2g_db_close(k2g_db *db) {
  delete db;
  return K2G_OK;
}

int k2g_db_get_info(const k2g_db *db, k2g_db_info *info) {
  if (!db || !info) return fail(K2G_EUSAGE, "null argument");
  *info = db->info;
  return K2G_OK;
}

int k2g_batch_create(k2g_db *db, k2g_batch **out) {
  if (!db || !out) return fail(K2G_EUSAGE, "null argument");
  *out = new k2g_batch();
  (*out)->db = db;
  return K2G_OK;
}

int k2g_batch_destroy(k2g_batch *b) {
  delete b;
  return K2G_OK;
}

int k2g_classify(k2g_batch *b, const k2g_classify_params *p, const k2g_fragments *in, k2g_results *r) {
  if (!b || !p || !in || !r) return fail(K2G_EUSAGE, "null argument");
  if (p->want_hll == 2) return fail(K2G_EUSAGE, "the test double writes no HLL sets");
  const uint64_t n = in->n, k = b->db->info.k;
  if (n && (!in->seq || !in->off)) return fail(K2G_EUSAGE, "null fragment arrays");
  b->call.assign(n + 1, 0);
  b->groups.assign(n + 1, 0);
  b->total.assign(n + 1, 0);
  b->run_off.assign(n + 1, 0);
  b->runs.clear();
  for (uint64_t i = 0; i < n; i++) {
    const bool pr = in->paired ? in->paired[i] != 0 : p->paired != 0;
    const uint64_t a = in->off[2 * i], m = in->off[2 * i + 1], z = pr ? in->off[2 * i + 2] : m;
    if (m < a || z < m) return fail(K2G_EUSAGE, "fragment offsets must not decrease");
    b->call[i] = (crc32(in->seq + a, z - a) >> 7) & 1;
    const uint64_t e1 = m - a >= k ? m - a - k + 1 : 0, e2 = pr && z - m >= k ? z - m - k + 1 : 0;
    b->total[i] = (uint32_t) (e1 + e2);
    auto run = [&](uint32_t code, uint32_t count) {
      b->runs.push_back(code);
      b->runs.push_back(count);
    };
    if (e1) run(0, (uint32_t) e1);
    if (pr) run(K2G_CODE_MATE, 1);
    if (e2) run(0, (uint32_t) e2);
    b->run_off[i + 1] = b->runs.size() / 2;
  }
  b->runs.resize(b->runs.size() + 2);  // never empty
  b->hll_off.assign(n + 1, 0);
  b->hll.assign(2, 0);
  b->kmer_counts.assign(b->db->info.node_count + 1, 0);
  memset(r, 0, sizeof(*r));
  r->n = n;
  r->node_count = b->db->info.node_count;
  r->call = b->call.data();
  r->hit_groups = b->groups.data();
  r->total_kmers = b->total.data();
  r->run_off = b->run_off.data();
  r->runs = b->runs.data();
  if (p->want_hll) {
    r->hll_off = b->hll_off.data();
    r->hll = b->hll.data();
  }
  if (p->want_kmer_counts) r->kmer_counts = b->kmer_counts.data();
  return K2G_OK;
}

}  // extern "C"
