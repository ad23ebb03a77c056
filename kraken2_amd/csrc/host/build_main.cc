// k2g_build_db: the reference's build_db command line (getopt string of
// src/build_db.cc:182; defaults src/build_db.cc:40-52) with the table built on
// the GPU through libk2g (k2g_build_*: atomicCAS with an in-kernel LCA merge).
//
// Reads the library FASTA/FASTQ from stdin, writes taxo.k2d from the NCBI
// taxonomy restricted to the mapped taxa, hash.k2d in the reference layout and
// opts.k2d.  Each sequence's taxon is the LCA of the taxa its header ids map
// to (src/build_db.h:100-111); unmapped sequences are skipped.
#include <getopt.h>
#include <sysexits.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "k2g.h"
#include "k2g_host.h"

namespace {

struct Options {
  std::string hash_file, map_file, taxo_file, ncbi_dir, opts_file;
  uint64_t k = 0, l = 0, capacity = 0, max_capacity = 0;
  uint64_t spaced_seed_mask = 0;  // DEFAULT_SPACED_SEED_MASK (src/mmscanner.h:16)
  uint64_t toggle_mask = 0xe37e28c4271b5a2dULL;
  bool protein = false, deterministic = true;
  uint64_t block_size = 10 * 1024 * 1024, subblock_size = 1024;
  uint64_t requested_bits = 0;
  int threads = 1;
};

[[noreturn]] void usage(int code = EX_USAGE) {
  std::cerr << "Usage: k2g_build_db <options>\n"
               "\n"
               "Options (*mandatory):\n"
               "* -H FILENAME   Kraken 2 hash table filename\n"
               "* -m FILENAME   Sequence ID to taxon map filename\n"
               "* -t FILENAME   Kraken 2 taxonomy filename\n"
               "* -n DIR        NCBI taxonomy directory name\n"
               "* -o FILENAME   Kraken 2 options filename\n"
               "* -k INT        Set length of k-mers\n"
               "* -l INT        Set length of minimizers\n"
               "* -c INT        Set capacity of hash table\n"
               "  -M INT        Set maximum capacity of hash table (MiniKraken)\n"
               "  -S BITSTRING  Spaced seed mask\n"
               "  -T BITSTRING  Minimizer toggle mask\n"
               "  -X            Input seqs. are proteins\n"
               "  -p INT        Number of threads (accepted; the GPU builds the table)\n"
               "  -F            Scan whole sequences (no block/subblock windows)\n"
               "  -B INT        Read block size\n"
               "  -b INT        Read subblock size\n"
               "  -r INT        Bit storage requested for taxid\n"
               "  -C INT        CHT cell size (only 32 is supported)" << std::endl;
  std::exit(code);
}

[[noreturn]] void die(int code, const std::string &msg) {
  std::cerr << "k2g_build_db: " << msg << std::endl;
  std::exit(code);
}

long long parse_int(const char *s) { return std::atoll(s); }

void parse(int argc, char **argv, Options &o) {
  int opt;
  while ((opt = getopt(argc, argv, "?hB:b:c:FH:m:n:o:t:k:l:M:p:r:s:S:T:XC:")) != -1) {
    long long v;
    switch (opt) {
      case 'h':
      case '?': usage(0);
      case 'B':
        if ((v = parse_int(optarg)) < 1) die(EX_USAGE, "must have positive block size");
        o.block_size = (uint64_t) v;
        break;
      case 'b':
        if ((v = parse_int(optarg)) < 1) die(EX_USAGE, "must have positive subblock size");
        o.subblock_size = (uint64_t) v;
        break;
      case 'r':
        v = parse_int(optarg);
        if (v < 0) die(EX_USAGE, "can't have negative bit storage");
        if (v > 31) die(EX_USAGE, "can't have more than 31 bits of storage for taxid");
        o.requested_bits = (uint64_t) v;
        break;
      case 'p':
        if ((o.threads = (int) parse_int(optarg)) < 1)
          die(EX_USAGE, "can't have negative number of threads");
        break;
      case 'H': o.hash_file = optarg; break;
      case 'm': o.map_file = optarg; break;
      case 'n': o.ncbi_dir = optarg; break;
      case 'o': o.opts_file = optarg; break;
      case 't': o.taxo_file = optarg; break;
      case 'S': o.spaced_seed_mask = std::strtoull(optarg, nullptr, 2); break;
      case 'T': o.toggle_mask = std::strtoull(optarg, nullptr, 2); break;
      case 'k':
        if ((v = parse_int(optarg)) < 1) die(EX_USAGE, "k must be positive integer");
        o.k = (uint64_t) v;
        break;
      case 'l':
        v = parse_int(optarg);
        if (v < 1) die(EX_USAGE, "l must be positive integer");
        if (v > 31) die(EX_USAGE, "l must be no more than 31");
        o.l = (uint64_t) v;
        break;
      case 'c':
        if ((v = parse_int(optarg)) < 1) die(EX_USAGE, "capacity must be positive integer");
        o.capacity = (uint64_t) v;
        break;
      case 'M':
        if ((v = parse_int(optarg)) < 1) die(EX_USAGE, "max capacity must be positive integer");
        o.max_capacity = (uint64_t) v;
        break;
      case 'F': o.deterministic = false; break;
      case 'X': o.protein = true; break;
      case 'C':
        v = parse_int(optarg);
        if (v != 32 && v != 40) die(EX_USAGE, "CHT cell size should be either 32 or 40 bits");
        if (v == 40) die(EX_USAGE, "40-bit cells are not supported by the GPU build");
        break;
    }
  }
  if (o.spaced_seed_mask != 0)
    o.spaced_seed_mask = k2host::expand_spaced_seed_mask(o.spaced_seed_mask, o.protein ? 4 : 2);
  if (o.hash_file.empty() || o.map_file.empty() || o.ncbi_dir.empty() || o.opts_file.empty() ||
      o.taxo_file.empty()) {
    std::cerr << "missing mandatory filename parameter" << std::endl;
    usage();
  }
  if (o.k == 0 || o.l == 0 || o.capacity == 0) {
    std::cerr << "missing mandatory integer parameter" << std::endl;
    usage();
  }
  if (o.k < o.l) {
    std::cerr << "k cannot be less than l" << std::endl;
    usage();
  }
  if (o.block_size < o.subblock_size) {
    std::cerr << "block size cannot be less than subblock size\n";
    usage();
  }
  if (o.max_capacity > o.capacity) {
    std::cerr << "maximum capacity option shouldn't specify larger capacity than normal"
              << std::endl;
    usage();
  }
}

// src/build_db.cc:152-167
std::map<std::string, uint64_t> read_id_map(const std::string &path) {
  std::ifstream in(path);
  if (!in.good()) die(EX_NOINPUT, "unable to read from '" + path + "'");
  std::map<std::string, uint64_t> m;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::string id;
    uint64_t taxid = 0;
    iss >> id >> taxid;
    if (taxid) m[id] = taxid;
  }
  return m;
}

void check(int rc) {
  if (rc != K2G_OK) die(rc, k2g_last_error());
}

}  // namespace

int main(int argc, char **argv) {
  Options o;
  parse(argc, argv, o);

  auto id_map = read_id_map(o.map_file);
  k2host::NcbiTaxonomy ncbi;
  std::string err;
  if (!ncbi.load(o.ncbi_dir + "/nodes.dmp", o.ncbi_dir + "/names.dmp", err)) die(EX_NOINPUT, err);
  for (auto &kv : id_map) ncbi.mark(kv.second);
  if (!ncbi.write_kraken(o.taxo_file, err)) die(EX_OSERR, err);
  std::cerr << "Taxonomy parsed and converted." << std::endl;

  k2host::Taxonomy tax;
  if (!tax.load(o.taxo_file, err)) die(EX_DATAERR, err);
  std::unordered_map<uint64_t, uint32_t> ext_to_int;
  for (size_t i = 1; i < tax.nodes.size(); i++) ext_to_int[tax.nodes[i].external_id] = (uint32_t) i;
  // src/build_db.cc:66-75
  uint64_t bits_needed = 1;
  while ((1ULL << bits_needed) < tax.nodes.size()) bits_needed++;
  if (o.requested_bits > 0 && bits_needed > o.requested_bits)
    die(EX_DATAERR, "more bits required for storing taxid");
  const uint64_t value_bits = std::max(bits_needed, o.requested_bits);

  uint64_t capacity = o.capacity, min_clear = 0;
  if (o.max_capacity) {  // src/build_db.cc:77-84
    const double frac = o.max_capacity * 1.0 / o.capacity;
    min_clear = (uint64_t) ((1 - frac) * (double) UINT64_MAX);
    capacity = o.max_capacity;
  }

  k2g_build_params bp{};
  bp.k = o.k;
  bp.l = o.l;
  bp.spaced_seed_mask = o.spaced_seed_mask;
  bp.toggle_mask = o.toggle_mask;
  bp.min_clear_hash_value = min_clear;
  // whole-sequence scans (-F) are one window per sequence
  bp.block_size = o.deterministic ? o.block_size : (1ULL << 62);
  bp.subblock_size = o.deterministic ? o.subblock_size : (1ULL << 62);
  bp.dna_db = o.protein ? 0 : 1;
  bp.revcom_version = 1;
  k2g_builder *bld = nullptr;
  check(k2g_build_create(o.taxo_file.c_str(), capacity, value_bits, &bp, &bld));
  std::cerr << "CHT created with " << value_bits << " bits reserved for taxid." << std::endl;

  // library sequences in batches of about 1 Gbp
  k2host::Reader reader;
  if (!reader.open(nullptr)) die(EX_NOINPUT, "unable to read sequences from standard input");
  std::string seq;
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> taxid;
  uint64_t n_seqs = 0, n_bases = 0;
  auto flush = [&] {
    if (taxid.empty()) return;
    check(k2g_build_insert(bld, seq.data(), off.data(), taxid.data(), taxid.size()));
    seq.clear();
    off.assign(1, 0);
    taxid.clear();
  };
  k2host::Record rec;
  while (reader.next(rec)) {
    uint32_t t = 0;  // src/build_db.h:100-111
    for (auto &id : k2host::sequence_ids(rec.header)) {
      auto it = id_map.find(id);
      if (it == id_map.end() || it->second == 0) continue;
      auto jt = ext_to_int.find(it->second);
      const uint32_t b = jt == ext_to_int.end() ? 0 : jt->second;
      // LCA on the BFS-ordered ids (parent < child), src/taxonomy.cc:262-272
      uint32_t a = t;
      if (a == 0 || b == 0) {
        t = a ? a : b;
        continue;
      }
      uint32_t x = a, y = b;
      while (x != y) {
        if (x > y) x = (uint32_t) tax.nodes[x].parent_id;
        else y = (uint32_t) tax.nodes[y].parent_id;
      }
      t = x;
    }
    if (!t) continue;
    if (o.protein && (rec.seq.empty() || rec.seq.back() != '*')) rec.seq.push_back('*');
    seq += rec.seq;
    off.push_back(seq.size());
    taxid.push_back(t);
    n_seqs++;
    n_bases += rec.seq.size();
    if (seq.size() >= (1ULL << 30)) flush();
  }
  flush();
  std::cerr << "Processed " << n_seqs << " sequences (" << n_bases << " "
            << (o.protein ? "aa" : "bp") << ")" << std::endl;
  std::cerr << "Writing data to disk... " << std::flush;
  check(k2g_build_write(bld, o.hash_file.c_str()));
  check(k2g_build_destroy(bld));

  // opts.k2d: the reference's IndexOptions (src/kraken2_data.h:15-25),
  // 64 bytes, padding zeroed
  unsigned char raw[64] = {};
  const uint64_t head[4] = {o.k, o.l, o.spaced_seed_mask, o.toggle_mask};
  std::memcpy(raw, head, 32);
  raw[32] = o.protein ? 0 : 1;
  std::memcpy(raw + 40, &min_clear, 8);
  const int32_t revcom = 1, db_version = 0, db_type = 0;
  std::memcpy(raw + 48, &revcom, 4);
  std::memcpy(raw + 52, &db_version, 4);
  std::memcpy(raw + 56, &db_type, 4);
  std::ofstream of(o.opts_file, std::ios::binary);
  of.write((const char *) raw, sizeof raw);
  if (!of.good()) die(EX_OSERR, "Unable to write options file " + o.opts_file);
  std::cerr << " complete." << std::endl;
  return 0;
}
