// Seeded synthetic workloads generated directly in HBM (benchmark inputs at
// the real shape: ~10 Gbp of genomes, tens of Gbp of reads).  Every base is
// a pure function of (seed, genome or read id, position), so a workload is
// reproducible and any part of it can be regenerated on the host for
// checking.
//
// Genomes: genome g belongs to group g / per_group (siblings of one genus).
// The group owns an i.i.d. root sequence with GC ~ U(0.35, 0.65); a genome is
// the root with a fraction of its 1-kb segments replaced by genome-specific
// sequence and point substitutions elsewhere.
//
// Reads: kind 0 single, 1 pair (insert ~ N(mean, sd), mate 2 reverse
// complemented), 2 long (substitutions and 1-3 bp indels); a fraction are
// off-target random reads with a fixed GC.  Either strand.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

__host__ __device__ inline uint64_t mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

__host__ __device__ inline uint64_t h3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix(seed ^ mix(a * 0x9E3779B97F4A7C15ULL + b + 0x632BE59BD9B4E019ULL));
}

__host__ __device__ inline double unit(uint64_t x) { return (double) (x >> 11) * (1.0 / 9007199254740992.0); }

__device__ inline char base_gc(uint64_t x, double gc) {
  const double u = unit(x);
  const bool bit = (x & 1) != 0;
  if (u < gc) return bit ? 'C' : 'G';
  return bit ? 'A' : 'T';
}

__device__ inline char other_base(char c, uint64_t x) {
  const char all[4] = {'A', 'C', 'G', 'T'};
  int i = 0;
  for (; i < 4; i++)
    if (all[i] == c) break;
  return all[(i + 1 + (int) (x % 3)) & 3];
}

__device__ inline char comp(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
  }
  return 'N';
}

struct GenomeParams {
  uint64_t seed;
  uint32_t per_group;
  double divergent_frac, snp_rate;
};

__device__ inline char genome_base(const GenomeParams &gp, uint32_t g, uint64_t p) {
  const uint32_t grp = g / gp.per_group;
  const double gc = 0.35 + 0.30 * unit(h3(gp.seed, 0xABCDEF, grp));
  const uint64_t seg = p >> 10;
  if (unit(h3(gp.seed ^ 0x1111, g, seg)) < gp.divergent_frac) return base_gc(h3(gp.seed ^ 0x2222, g, p), gc);
  char b = base_gc(h3(gp.seed ^ 0x3333, grp, p), gc);
  const uint64_t s = h3(gp.seed ^ 0x4444, g, p);
  if (unit(s) < gp.snp_rate) b = other_base(b, s >> 20);
  return b;
}

struct Chunk {
  uint64_t out;  // output offset
  uint64_t pos;  // position in the genome
  uint32_t genome, len;
};

__global__ void k_genomes(GenomeParams gp, const Chunk *chunks, uint64_t n, char *out) {
  const uint64_t c = blockIdx.x;
  if (c >= n) return;
  const Chunk ch = chunks[c];
  for (uint32_t i = threadIdx.x; i < ch.len; i += blockDim.x)
    out[ch.out + i] = genome_base(gp, ch.genome, ch.pos + i);
}

struct ReadParams {
  uint64_t seed;
  double sub_rate, long_sub, long_indel, insert_mean, insert_sd, off_frac, off_gc;
  uint32_t n_genomes;
};

struct Rng {
  uint64_t s;
  __device__ uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    return mix(s);
  }
  __device__ double u() { return unit(next()); }
  __device__ double normal() {
    double a = u(), b = u();
    if (a < 1e-300) a = 1e-300;
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
  }
};

// One thread per read: mate(s) written at their offsets.
__global__ void k_reads(ReadParams rp, const uint8_t *kind, const uint64_t *off, uint64_t n,
                        const char *genomes, const uint64_t *goff, char *seq) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  Rng rng{h3(rp.seed, 0x5EED, r)};
  const uint64_t a0 = off[2 * r], a1 = off[2 * r + 1], a2 = off[2 * r + 2];
  const uint64_t L1 = a1 - a0, L2 = a2 - a1;
  const uint8_t kd = kind[r];
  if (rng.u() < rp.off_frac) {
    for (uint64_t i = a0; i < a2; i++) seq[i] = base_gc(rng.next(), rp.off_gc);
    return;
  }
  const uint32_t g = (uint32_t) (rng.next() % rp.n_genomes);
  const char *G = genomes + goff[g];
  const uint64_t GL = goff[g + 1] - goff[g];
  const bool rev = (rng.next() & 1) != 0;
  if (kd == 0 || kd == 1) {
    uint64_t frag = L1;
    if (kd == 1) {
      double ins = rp.insert_mean + rp.insert_sd * rng.normal();
      frag = (uint64_t) (ins < (double) L1 ? (double) L1 : ins);
      if (frag < L2) frag = L2;
    }
    if (frag > GL) frag = GL;
    const uint64_t pos = rng.next() % (GL - frag + 1);
    // fragment base i on the chosen strand
    auto fb = [&](uint64_t i) -> char { return rev ? comp(G[pos + frag - 1 - i]) : G[pos + i]; };
    for (uint64_t i = 0; i < L1; i++) {
      char b = i < frag ? fb(i) : 'N';
      if (rng.u() < rp.sub_rate) b = other_base(b, rng.next());
      seq[a0 + i] = b;
    }
    for (uint64_t i = 0; i < L2; i++) {  // reverse complement of the fragment end
      char b = i < frag ? comp(fb(frag - 1 - i)) : 'N';
      if (rng.u() < rp.sub_rate) b = other_base(b, rng.next());
      seq[a1 + i] = b;
    }
    return;
  }
  // long read: walk the source with substitutions and 1-3 bp indels
  const uint64_t span = L1 + L1 / 8 + 8;
  const uint64_t use = span < GL ? span : GL;
  uint64_t sp = rng.next() % (GL - use + 1);
  const uint64_t end = sp + use;
  uint64_t o = 0;
  while (o < L1) {
    const double u = rng.u();
    if (u < rp.long_indel) {
      const uint32_t len = 1 + (uint32_t) (rng.next() % 3);
      if (rng.next() & 1) {
        for (uint32_t j = 0; j < len && o < L1; j++) seq[a0 + o++] = "ACGT"[rng.next() & 3];
      } else {
        sp += len;
      }
      continue;
    }
    char b;
    if (sp < end) {
      b = rev ? comp(G[GL - 1 - sp]) : G[sp];
      sp++;
    } else {
      b = "ACGT"[rng.next() & 3];
    }
    if (u < rp.long_indel + rp.long_sub) b = other_base(b, rng.next());
    seq[a0 + o++] = b;
  }
}

// ---- proteomes and back-translated reads (translated search)
//
// Taxon t belongs to group t / per_group.  The group owns n_proteins
// ancestral proteins whose residues are i.i.d. from fixed background
// amino-acid frequencies; taxon t's copy keeps each residue with
// probability `identity` and otherwise draws it again from the background
// (two siblings then agree at about identity^2 of the positions).  Every
// protein is stored followed by '*', as the build appends one.

// Background frequencies (percent x 100) of the 20 amino acids, in the
// order of kAminoAcids; they sum to 10000.
constexpr char kAminoAcids[21] = "ARNDCQEGHILKMFPSTWYV";
struct AaTable {
  uint16_t cum[20];  // cumulative frequencies
  char aa[20];
};

__device__ inline char background_aa(const AaTable &t, uint64_t x) {
  const uint32_t u = (uint32_t) (x % 10000);
  int i = 0;
  while (i < 19 && u >= t.cum[i]) i++;
  return t.aa[i];
}

struct ProteomeParams {
  uint64_t seed;
  uint32_t per_group, n_proteins;
  double identity;
  AaTable bg;
};

struct ProtChunk {
  uint64_t out;   // output offset
  uint32_t taxon, protein, pos, len, prot_len;  // residues [pos, pos+len) of a protein of prot_len
};

__global__ void k_proteomes(ProteomeParams pp, const ProtChunk *chunks, uint64_t n, char *out) {
  const uint64_t c = blockIdx.x;
  if (c >= n) return;
  const ProtChunk ch = chunks[c];
  const uint64_t anc = (uint64_t) (ch.taxon / pp.per_group) * pp.n_proteins + ch.protein;
  const uint64_t own = (uint64_t) ch.taxon * pp.n_proteins + ch.protein;
  for (uint32_t i = threadIdx.x; i < ch.len; i += blockDim.x) {
    const uint32_t q = ch.pos + i;
    char a;
    if (q == ch.prot_len) {
      a = '*';
    } else if (unit(h3(pp.seed ^ 0x5151, own, q)) < pp.identity) {
      a = background_aa(pp.bg, h3(pp.seed ^ 0x6161, anc, q));
    } else {
      a = background_aa(pp.bg, h3(pp.seed ^ 0x7171, own, q));
    }
    out[ch.out + i] = a;
  }
}

// Codons of each amino acid (kAminoAcids order, standard genetic code),
// bases packed two bits each (A=0 C=1 G=2 T=3, first base high).
struct BackTable {
  uint8_t n[20];
  uint8_t codon[20][6];
};

struct ProtReadParams {
  uint64_t seed;
  uint32_t read_len, n_taxa, n_proteins;
  double sub_rate;
  BackTable back;
};

// One thread per read: a random protein of a random taxon, a window of
// read_len/3 residues back-translated with uniformly random synonymous
// codons, 1% substitutions, either strand.  Proteins shorter than the
// window are passed over for another (up to 64 tries, then the read is
// random sequence).
__global__ void k_protein_reads(ProtReadParams rp, const char *prot, const uint64_t *poff, uint64_t n,
                                char *seq) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  Rng rng{h3(rp.seed, 0xBEE5, r)};
  const uint32_t codons = rp.read_len / 3;
  char *o = seq + r * rp.read_len;
  const char *P = nullptr;
  for (int tries = 0; tries < 64 && !P; tries++) {
    const uint64_t s = (rng.next() % rp.n_taxa) * rp.n_proteins + rng.next() % rp.n_proteins;
    const uint64_t L = poff[s + 1] - poff[s] - 1;  // without the '*'
    if (L >= codons) P = prot + poff[s] + rng.next() % (L - codons + 1);
  }
  const bool rev = (rng.next() & 1) != 0;
  for (uint32_t i = 0; i < rp.read_len; i++) o[i] = "ACGT"[rng.next() & 3];
  if (P) {
    for (uint32_t j = 0; j < codons; j++) {
      int a = 0;
      while (a < 19 && kAminoAcids[a] != P[j]) a++;
      const uint8_t c = rp.back.codon[a][rng.next() % rp.back.n[a]];
      for (int b = 0; b < 3; b++) o[3 * j + b] = "ACGT"[(c >> (4 - 2 * b)) & 3];
    }
  }
  for (uint32_t i = 0; i < rp.read_len; i++)
    if (rng.u() < rp.sub_rate) o[i] = other_base(o[i], rng.next());
  if (rev) {
    for (uint32_t i = 0, j = rp.read_len - 1; i < j; i++, j--) {
      const char t = o[i];
      o[i] = comp(o[j]);
      o[j] = comp(t);
    }
    if (rp.read_len & 1) o[rp.read_len / 2] = comp(o[rp.read_len / 2]);
  }
}

}  // namespace

extern "C" {

// Lognormal genome lengths, clipped (host, deterministic).
int k2s_genome_lengths(uint64_t seed, uint32_t n, double median, double sigma, double lo, double hi,
                       uint64_t *lens) {
  std::mt19937_64 eng(seed ^ 0x6E0E);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (uint32_t i = 0; i < n; i++) {
    double v = std::exp(std::log(median) + sigma * nd(eng));
    v = std::min(std::max(v, lo), hi);
    lens[i] = (uint64_t) v;
  }
  return 0;
}

// Genomes into d_out; genome g occupies [off[g], off[g+1]) (host offsets).
int k2s_genomes(uint64_t seed, uint32_t n, uint32_t per_group, double divergent_frac, double snp_rate,
                const uint64_t *off, char *d_out) {
  std::vector<Chunk> chunks;
  for (uint32_t g = 0; g < n; g++) {
    const uint64_t L = off[g + 1] - off[g];
    for (uint64_t p = 0; p < L; p += 4096)
      chunks.push_back(Chunk{off[g] + p, p, g, (uint32_t) std::min<uint64_t>(4096, L - p)});
  }
  if (chunks.empty()) return 0;
  Chunk *d = nullptr;
  if (hipMalloc((void **) &d, chunks.size() * sizeof(Chunk)) != hipSuccess) return 71;
  if (hipMemcpy(d, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice) != hipSuccess) return 71;
  GenomeParams gp{seed, per_group ? per_group : 1, divergent_frac, snp_rate};
  hipLaunchKernelGGL(k_genomes, dim3((unsigned) chunks.size()), dim3(256), 0, 0, gp, d,
                     (uint64_t) chunks.size(), d_out);
  hipError_t e = hipDeviceSynchronize();
  (void) hipFree(d);
  return e == hipSuccess ? 0 : 70;
}

// Read kinds and lengths (host, deterministic): fills off[2n+1] and kind[n].
int k2s_read_layout(uint64_t seed, uint64_t n, double f_single, double f_paired, uint32_t single_len,
                    uint32_t paired_len, double long_median, double long_sigma, uint32_t long_cap,
                    uint64_t *off, uint8_t *kind) {
  std::mt19937_64 eng(seed ^ 0x1A70);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  std::normal_distribution<double> nd(0.0, 1.0);
  uint64_t pos = 0;
  for (uint64_t r = 0; r < n; r++) {
    const double u = ud(eng);
    uint64_t L1, L2 = 0;
    if (u < f_single) {
      kind[r] = 0;
      L1 = single_len;
    } else if (u < f_single + f_paired) {
      kind[r] = 1;
      L1 = L2 = paired_len;
    } else {
      kind[r] = 2;
      double v = std::exp(std::log(long_median) + long_sigma * nd(eng));
      L1 = (uint64_t) std::min<double>(std::max<double>(v, 1.0), (double) long_cap);
    }
    off[2 * r] = pos;
    pos += L1;
    off[2 * r + 1] = pos;
    pos += L2;
  }
  off[2 * n] = pos;
  return 0;
}

// Reads into d_seq (device); kind/off/goff are host arrays.
int k2s_reads(uint64_t seed, uint64_t n, const uint8_t *kind, const uint64_t *off, const char *d_genomes,
              const uint64_t *goff, uint32_t n_genomes, double sub_rate, double long_sub, double long_indel,
              double insert_mean, double insert_sd, double off_frac, double off_gc, char *d_seq) {
  uint8_t *dk = nullptr;
  uint64_t *doff = nullptr, *dg = nullptr;
  if (hipMalloc((void **) &dk, n ? n : 1) != hipSuccess) return 71;
  if (hipMalloc((void **) &doff, (2 * n + 1) * 8) != hipSuccess) return 71;
  if (hipMalloc((void **) &dg, (n_genomes + 1) * 8) != hipSuccess) return 71;
  if (hipMemcpy(dk, kind, n, hipMemcpyHostToDevice) != hipSuccess) return 71;
  if (hipMemcpy(doff, off, (2 * n + 1) * 8, hipMemcpyHostToDevice) != hipSuccess) return 71;
  if (hipMemcpy(dg, goff, (n_genomes + 1) * 8, hipMemcpyHostToDevice) != hipSuccess) return 71;
  ReadParams rp{seed, sub_rate, long_sub, long_indel, insert_mean, insert_sd, off_frac, off_gc, n_genomes};
  if (n)
    hipLaunchKernelGGL(k_reads, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, 0, rp, dk, doff, n,
                       d_genomes, dg, d_seq);
  hipError_t e = hipDeviceSynchronize();
  (void) hipFree(dk);
  (void) hipFree(doff);
  (void) hipFree(dg);
  return e == hipSuccess ? 0 : 70;
}

// Ancestral protein lengths per (group, protein), lognormal, clipped (host,
// deterministic).
int k2s_protein_lengths(uint64_t seed, uint64_t n, double median, double sigma, double lo, double hi,
                        uint32_t *lens) {
  std::mt19937_64 eng(seed ^ 0x9807);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (uint64_t i = 0; i < n; i++) {
    double v = std::exp(std::log(median) + sigma * nd(eng));
    lens[i] = (uint32_t) std::min(std::max(v, lo), hi);
  }
  return 0;
}

// Proteomes into d_out: sequence s = t * n_proteins + p (taxon t, protein
// p) occupies [off[s], off[s+1]) = the protein then '*' (host offsets, whose
// lengths are group lengths + 1).
int k2s_proteomes(uint64_t seed, uint32_t n_taxa, uint32_t per_group, uint32_t n_proteins, double identity,
                  const uint64_t *off, char *d_out) {
  std::vector<ProtChunk> chunks;
  for (uint32_t t = 0; t < n_taxa; t++)
    for (uint32_t p = 0; p < n_proteins; p++) {
      const uint64_t s = (uint64_t) t * n_proteins + p;
      const uint32_t L = (uint32_t) (off[s + 1] - off[s]);  // including '*'
      for (uint32_t q = 0; q < L; q += 4096)
        chunks.push_back(ProtChunk{off[s] + q, t, p, q, std::min<uint32_t>(4096, L - q), L - 1});
    }
  if (chunks.empty()) return 0;
  ProteomeParams pp{seed, per_group ? per_group : 1, n_proteins, identity, {}};
  static const uint16_t freq[20] = {825, 553, 406, 545, 137, 393, 675, 707, 227, 596,
                                    966, 584, 242, 386, 470, 656, 534, 108, 292, 698};
  uint32_t acc = 0;
  for (int i = 0; i < 20; i++) {
    acc += freq[i];
    pp.bg.cum[i] = (uint16_t) acc;
    pp.bg.aa[i] = kAminoAcids[i];
  }
  pp.bg.cum[19] = 10000;
  const size_t step = 1u << 22;  // chunks per launch (grid.x bound)
  ProtChunk *d = nullptr;
  if (hipMalloc((void **) &d, std::min(chunks.size(), step) * sizeof(ProtChunk)) != hipSuccess) return 71;
  hipError_t e = hipSuccess;
  for (size_t c0 = 0; c0 < chunks.size() && e == hipSuccess; c0 += step) {
    const size_t m = std::min(step, chunks.size() - c0);
    e = hipMemcpy(d, chunks.data() + c0, m * sizeof(ProtChunk), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_proteomes, dim3((unsigned) m), dim3(256), 0, 0, pp, d, (uint64_t) m, d_out);
      e = hipDeviceSynchronize();
    }
  }
  (void) hipFree(d);
  return e == hipSuccess ? 0 : 70;
}

// n reads of read_len bases back-translated from the proteomes (d_prot,
// sequence offsets poff on the host as for k2s_proteomes), at d_seq + r *
// read_len.
int k2s_protein_reads(uint64_t seed, uint64_t n, uint32_t read_len, uint32_t n_taxa, uint32_t n_proteins,
                      const char *d_prot, const uint64_t *poff, double sub_rate, char *d_seq) {
  // standard genetic code, codons in TCAG order (first base slowest)
  static const char code[65] = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
  static const int tcag_to_acgt[4] = {3, 1, 0, 2};
  ProtReadParams rp{seed, read_len, n_taxa, n_proteins, sub_rate, {}};
  for (int i = 0; i < 64; i++) {
    const char aa = code[i];
    const char *hit = std::strchr(kAminoAcids, aa);
    if (aa == '*' || !hit) continue;
    const int a = (int) (hit - kAminoAcids);
    const int b0 = tcag_to_acgt[i >> 4], b1 = tcag_to_acgt[(i >> 2) & 3], b2 = tcag_to_acgt[i & 3];
    rp.back.codon[a][rp.back.n[a]++] = (uint8_t) ((b0 << 4) | (b1 << 2) | b2);
  }
  const uint64_t n_seq = (uint64_t) n_taxa * n_proteins;
  uint64_t *dp = nullptr;
  if (hipMalloc((void **) &dp, (n_seq + 1) * 8) != hipSuccess) return 71;
  if (hipMemcpy(dp, poff, (n_seq + 1) * 8, hipMemcpyHostToDevice) != hipSuccess) return 71;
  if (n)
    hipLaunchKernelGGL(k_protein_reads, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, 0, rp, d_prot, dp, n,
                       d_seq);
  hipError_t e = hipDeviceSynchronize();
  (void) hipFree(dp);
  return e == hipSuccess ? 0 : 70;
}

}  // extern "C"
