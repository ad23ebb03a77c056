// k2synth: seeded synthetic inputs for the classify workloads.
//
//   k2synth db-inputs --out DIR [--seed S] [--levels 4,4,4] [--ranks phylum,genus,species]
//                     [--genome-len N | --genome-lognormal MEDIAN,SIGMA,MIN,MAX]
//                     [--sub 0.05]
//       Writes DIR/taxonomy/{nodes,names}.dmp, DIR/library.fna and
//       DIR/seqid2taxid.map.  The tree is root -> levels[0] children ->
//       levels[1] children each -> ...; the second-to-last level nodes own an
//       i.i.d. "root sequence" with GC ~ U(0.35,0.65) and every leaf is that
//       sequence with --sub random substitutions.
//
//   k2synth reads --lib DIR/library.fna --out R1.fa [--out2 R2.fa] [--seed S]
//                 [--n 100000] [--len 150] [--sub 0.01] [--random-frac 0.1]
//                 [--random-gc 0.5] [--insert-mean 350 --insert-sd 50]
//                 [--long-lognormal MEDIAN,SIGMA,CAP --indel 0.025]
//                 [--fastq]
//       Samples reads uniformly over genomes and positions, either strand.
//
// The generator is deterministic for a given seed and argument list.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) {
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? (uint64_t) (uniform() * n) % n : 0; }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2);
  }
};

const char BASES[4] = {'A', 'C', 'G', 'T'};

char random_base_gc(Rng &r, double gc) {
  double u = r.uniform();
  if (u < gc) return r.uniform() < 0.5 ? 'C' : 'G';
  return r.uniform() < 0.5 ? 'A' : 'T';
}

char substitute(Rng &r, char c) {
  char d;
  do { d = BASES[r.below(4)]; } while (d == c);
  return d;
}

char complement(char c) {
  switch (c) {
    case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; case 'T': return 'A';
    default: return 'N';
  }
}

std::map<std::string, std::string> parse_args(int argc, char **argv, int first) {
  std::map<std::string, std::string> m;
  for (int i = first; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) { std::cerr << "bad argument " << a << "\n"; exit(64); }
    a = a.substr(2);
    if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) m[a] = argv[++i];
    else m[a] = "1";
  }
  return m;
}

std::vector<double> split_nums(const std::string &s) {
  std::vector<double> v;
  size_t p = 0;
  while (p <= s.size()) {
    size_t q = s.find(',', p);
    if (q == std::string::npos) q = s.size();
    if (q > p) v.push_back(atof(s.substr(p, q - p).c_str()));
    p = q + 1;
  }
  return v;
}

std::vector<std::string> split_strs(const std::string &s) {
  std::vector<std::string> v;
  size_t p = 0;
  while (p <= s.size()) {
    size_t q = s.find(',', p);
    if (q == std::string::npos) q = s.size();
    v.push_back(s.substr(p, q - p));
    p = q + 1;
  }
  return v;
}

std::string get(std::map<std::string, std::string> &m, const char *k, const char *def) {
  auto it = m.find(k);
  return it == m.end() ? std::string(def) : it->second;
}

void write_fasta_record(FILE *f, const std::string &name, const std::string &seq, bool fastq) {
  fprintf(f, "%c%s\n", fastq ? '@' : '>', name.c_str());
  fwrite(seq.data(), 1, seq.size(), f);
  if (fastq) {
    fputs("\n+\n", f);
    std::string q(seq.size(), 'I');
    fwrite(q.data(), 1, q.size(), f);
  }
  fputc('\n', f);
}

int db_inputs(std::map<std::string, std::string> &a) {
  std::string out = get(a, "out", "");
  if (out.empty()) { std::cerr << "--out required\n"; return 64; }
  uint64_t seed = strtoull(get(a, "seed", "1").c_str(), nullptr, 10);
  std::vector<double> levels = split_nums(get(a, "levels", "4,4,4"));
  std::vector<std::string> ranks = split_strs(get(a, "ranks", "phylum,genus,species"));
  if (ranks.size() != levels.size() || levels.size() < 2) {
    std::cerr << "--ranks must name one rank per level (at least 2 levels)\n";
    return 64;
  }
  double sub = atof(get(a, "sub", "0.05").c_str());
  std::vector<double> lgn = split_nums(get(a, "genome-lognormal", ""));
  uint64_t glen = strtoull(get(a, "genome-len", "200000").c_str(), nullptr, 10);

  mkdir(out.c_str(), 0755);
  mkdir((out + "/taxonomy").c_str(), 0755);
  FILE *nodes = fopen((out + "/taxonomy/nodes.dmp").c_str(), "w");
  FILE *names = fopen((out + "/taxonomy/names.dmp").c_str(), "w");
  FILE *lib = fopen((out + "/library.fna").c_str(), "w");
  FILE *map = fopen((out + "/seqid2taxid.map").c_str(), "w");
  if (!nodes || !names || !lib || !map) { std::cerr << "cannot write to " << out << "\n"; return 73; }

  fprintf(nodes, "1\t|\t1\t|\tno rank\t|\t\t|\n");
  fprintf(names, "1\t|\troot\t|\t\t|\tscientific name\t|\n");
  Rng rng(seed);
  // level-by-level construction; external ids are sequential from 2
  std::vector<uint64_t> frontier = {1};
  uint64_t next_id = 2;
  std::vector<std::vector<uint64_t>> by_level;
  std::map<uint64_t, uint64_t> parent;
  for (size_t lv = 0; lv < levels.size(); lv++) {
    std::vector<uint64_t> nf;
    for (uint64_t p : frontier) {
      for (int c = 0; c < (int) levels[lv]; c++) {
        uint64_t id = next_id++;
        parent[id] = p;
        fprintf(nodes, "%llu\t|\t%llu\t|\t%s\t|\t\t|\n", (unsigned long long) id,
                (unsigned long long) p, ranks[lv].c_str());
        fprintf(names, "%llu\t|\t%s %llu\t|\t\t|\tscientific name\t|\n", (unsigned long long) id,
                ranks[lv].c_str(), (unsigned long long) id);
        nf.push_back(id);
      }
    }
    by_level.push_back(nf);
    frontier = nf;
  }
  // root sequences at the second-to-last level; leaves are mutated copies
  const auto &owners = by_level[by_level.size() - 2];
  uint64_t genome_no = 0;
  for (uint64_t owner : owners) {
    uint64_t len = glen;
    if (lgn.size() >= 4) {
      double v = std::exp(std::log(lgn[0]) + lgn[1] * rng.normal());
      if (v < lgn[2]) v = lgn[2];
      if (v > lgn[3]) v = lgn[3];
      len = (uint64_t) v;
    }
    double gc = 0.35 + 0.30 * rng.uniform();
    std::string root(len, 'A');
    for (auto &c : root) c = random_base_gc(rng, gc);
    for (uint64_t leaf : by_level.back()) {
      if (parent[leaf] != owner) continue;
      std::string g = root;
      for (auto &c : g)
        if (rng.uniform() < sub) c = substitute(rng, c);
      std::string name = "g" + std::to_string(genome_no++) + "_" + std::to_string(leaf);
      write_fasta_record(lib, name, g, false);
      fprintf(map, "%s\t%llu\n", name.c_str(), (unsigned long long) leaf);
    }
  }
  fclose(nodes); fclose(names); fclose(lib); fclose(map);
  return 0;
}

std::vector<std::string> load_fasta_seqs(const std::string &path) {
  std::vector<std::string> seqs;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    if (line[0] == '>') { seqs.emplace_back(); continue; }
    if (!seqs.empty()) seqs.back() += line;
  }
  return seqs;
}

std::string mutate_read(Rng &r, const std::string &src, double sub, double indel) {
  if (sub <= 0 && indel <= 0) return src;
  std::string out;
  out.reserve(src.size() + 16);
  for (size_t i = 0; i < src.size(); i++) {
    double u = r.uniform();
    if (u < indel) {
      int n = 1 + (int) r.below(3);
      if (r.uniform() < 0.5) {  // deletion: skip n source bases
        i += n - 1;
        continue;
      }
      for (int j = 0; j < n; j++) out.push_back(BASES[r.below(4)]);
      out.push_back(src[i]);
    } else if (u < indel + sub) {
      out.push_back(substitute(r, src[i]));
    } else {
      out.push_back(src[i]);
    }
  }
  return out;
}

int reads(std::map<std::string, std::string> &a) {
  std::string lib = get(a, "lib", ""), out = get(a, "out", ""), out2 = get(a, "out2", "");
  if (lib.empty() || out.empty()) { std::cerr << "--lib and --out required\n"; return 64; }
  uint64_t seed = strtoull(get(a, "seed", "1").c_str(), nullptr, 10);
  uint64_t n = strtoull(get(a, "n", "100000").c_str(), nullptr, 10);
  uint64_t len = strtoull(get(a, "len", "150").c_str(), nullptr, 10);
  double sub = atof(get(a, "sub", "0.01").c_str());
  double indel = atof(get(a, "indel", "0").c_str());
  double rnd = atof(get(a, "random-frac", "0.1").c_str());
  double rgc = atof(get(a, "random-gc", "0.5").c_str());
  double ins_mean = atof(get(a, "insert-mean", "350").c_str());
  double ins_sd = atof(get(a, "insert-sd", "50").c_str());
  std::vector<double> lng = split_nums(get(a, "long-lognormal", ""));
  bool fastq = a.count("fastq") > 0;
  bool paired = !out2.empty();

  std::vector<std::string> genomes = load_fasta_seqs(lib);
  if (genomes.empty()) { std::cerr << "no genomes in " << lib << "\n"; return 65; }
  FILE *f1 = fopen(out.c_str(), "w");
  FILE *f2 = paired ? fopen(out2.c_str(), "w") : nullptr;
  if (!f1 || (paired && !f2)) { std::cerr << "cannot write reads\n"; return 73; }
  Rng rng(seed ^ 0x5eed5eed5eedULL);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t rlen = len;
    if (lng.size() >= 3) {
      double v = std::exp(std::log(lng[0]) + lng[1] * rng.normal());
      if (v > lng[2]) v = lng[2];
      if (v < 1) v = 1;
      rlen = (uint64_t) v;
    }
    std::string name = "r" + std::to_string(i);
    bool off_target = rng.uniform() < rnd;
    if (!paired) {
      std::string s;
      if (off_target) {
        s.resize(rlen);
        for (auto &c : s) c = random_base_gc(rng, rgc);
      } else {
        const std::string &g = genomes[rng.below(genomes.size())];
        uint64_t L = std::min<uint64_t>(rlen, g.size());
        uint64_t pos = rng.below(g.size() - L + 1);
        s = g.substr(pos, L);
        if (rng.uniform() < 0.5) {
          std::string rc(s.rbegin(), s.rend());
          for (auto &c : rc) c = complement(c);
          s = rc;
        }
        s = mutate_read(rng, s, sub, indel);
      }
      write_fasta_record(f1, name, s, fastq);
    } else {
      std::string s1, s2;
      if (off_target) {
        s1.resize(len); s2.resize(len);
        for (auto &c : s1) c = random_base_gc(rng, rgc);
        for (auto &c : s2) c = random_base_gc(rng, rgc);
      } else {
        const std::string &g = genomes[rng.below(genomes.size())];
        double ins = ins_mean + ins_sd * rng.normal();
        uint64_t frag = (uint64_t) std::max<double>((double) len, ins);
        if (frag > g.size()) frag = g.size();
        uint64_t pos = rng.below(g.size() - frag + 1);
        std::string fr = g.substr(pos, frag);
        if (rng.uniform() < 0.5) {
          std::string rc(fr.rbegin(), fr.rend());
          for (auto &c : rc) c = complement(c);
          fr = rc;
        }
        uint64_t L = std::min<uint64_t>(len, fr.size());
        s1 = fr.substr(0, L);
        std::string tail = fr.substr(fr.size() - L);
        s2.assign(tail.rbegin(), tail.rend());
        for (auto &c : s2) c = complement(c);
        s1 = mutate_read(rng, s1, sub, indel);
        s2 = mutate_read(rng, s2, sub, indel);
      }
      write_fasta_record(f1, name + "/1", s1, fastq);
      write_fasta_record(f2, name + "/2", s2, fastq);
    }
  }
  fclose(f1);
  if (f2) fclose(f2);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: k2synth db-inputs|reads --opt value ...\n";
    return 64;
  }
  auto args = parse_args(argc, argv, 2);
  std::string mode = argv[1];
  if (mode == "db-inputs") return db_inputs(args);
  if (mode == "reads") return reads(args);
  std::cerr << "unknown mode " << mode << "\n";
  return 64;
}
