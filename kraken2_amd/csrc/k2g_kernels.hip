// HIP kernels of the MI355X classify core (gfx950, wave64).
//
// One wavefront owns one fragment (a read or a read pair) and walks it in
// chunks of 64 k-mer end positions.  Per chunk every lane:
//   * loads one base (coalesced) and the wave turns the chunk into 2-bit
//     planes with ballots;
//   * builds the l-mer ending at its position from the planes, takes the
//     canonical form, applies the spaced-seed mask and the toggle
//     (src/mmscanner.cc:150-153);
//   * takes the window minimum over the k-l+1 candidates kept in an LDS ring
//     (src/mmscanner.cc:159-187; after an ambiguous base only the l-mers
//     starting past it count, src/mmscanner.cc:128-146);
//   * decides LOOKUP / REPEAT / AMBIG against the previous non-ambiguous
//     minimizer (src/classify.cc:797-838) and probes the table for LOOKUPs
//     (src/compact_hash.h:354-374);
// then the wave run-length encodes the per-k-mer taxa into hitlist runs,
// accumulates hit counts in an LDS table, and finally resolves the call
// (src/classify.cc:710-762, 890-906).  Nothing here is a dense contraction,
// so no MFMA is involved: the probe is a random 16-byte gather into HBM.
#include "k2g_device.h"
#include "k2g_kernels.h"

namespace k2g {

// Genetic code (A,G,C,T ordering, first base in the high bits) mapped
// straight to the 15-letter reduced protein alphabet codes of the protein
// scanner (src/aa_translate.cc:37; src/mmscanner.cc:41-95).
__constant__ uint8_t c_codon_code[64] = {
    9, 9, 2, 2, 11, 11, 2, 2, 13, 13, 13, 13, 8,  12, 8,  8,  4, 4, 4, 4, 6, 6,
    6, 6, 1, 1, 1,  1,  12, 12, 12, 12, 2,  2,  7,  7,  11, 11, 11, 11, 10, 10,
    10, 10, 8, 8, 8, 8, 0, 0, 15, 15, 0, 14, 3, 3, 2, 2, 2, 2, 8, 8, 5, 5};

// A/C/G/T (either case) -> 0..3, anything else 0xFF, without branches:
// ((c >> 1) ^ (c >> 2)) & 3 maps 'a' 'c' 'g' 't' to 0 1 2 3.
__device__ __forceinline__ uint32_t dna_code(uint8_t ch) {
  const uint32_t c = (uint32_t) ch | 0x20u;
  const bool ok = c == 'a' || c == 'c' || c == 'g' || c == 't';
  return ok ? (((c >> 1) ^ (c >> 2)) & 3u) : 0xFFu;
}

// A=0 G=1 C=2 T=3 (codon index order of the translation table)
__device__ __forceinline__ uint32_t agct_code(uint8_t ch) {
  switch (ch | 0x20) {
    case 'a': return 0;
    case 'g': return 1;
    case 'c': return 2;
    case 't': return 3;
  }
  return 0xFFu;
}

// Length of reading frame fr (0..5) of a DNA mate of length L
// (src/aa_translate.cc:41-100: codons end at i >= 2 with i % 3 == fr % 3;
// a mate shorter than 3 leaves every frame as one blank character).
__device__ __forceinline__ uint64_t frame_len(uint64_t L, int fr) {
  if (L < 3) return 1;
  const uint64_t first = (fr % 3) == 0 ? 3 : ((fr % 3) == 1 ? 4 : 2);
  return (L - 1 >= first) ? (L - 1 - first) / 3 + 1 : 0;
}

// Reduced-alphabet code of position q of reading frame fr, 0xFF = ambiguous.
__device__ __forceinline__ uint32_t frame_code(const char *m, uint64_t L, int fr, uint64_t q) {
  if (L < 3) return 0xFFu;
  const uint64_t first = (fr % 3) == 0 ? 3 : ((fr % 3) == 1 ? 4 : 2);
  const uint64_t nfr = frame_len(L, fr);
  const uint64_t j = fr < 3 ? q : nfr - 1 - q;
  const uint64_t i = first + 3 * j;
  const uint32_t a = agct_code((uint8_t) m[i - 2]);
  const uint32_t b = agct_code((uint8_t) m[i - 1]);
  const uint32_t c = agct_code((uint8_t) m[i]);
  if ((a | b | c) == 0xFFu || a > 3 || b > 3 || c > 3) return 0xFFu;
  const uint32_t idx = fr < 3 ? ((a << 4) | (b << 2) | c) : (((3 - c) << 4) | ((3 - b) << 2) | (3 - a));
  return c_codon_code[idx];
}

// 32-bit sparse encoding of the HLL hash fmix64(m + 1) (p = 12, p' = 25;
// src/hyperloglogplus.cc:166-190, 816-824).
__device__ __forceinline__ uint32_t hll_encode(uint64_t m) {
  const uint64_t h = fmix64(m + 1);
  const uint32_t idx = (uint32_t) ((h >> 39) << 7);
  if ((uint32_t) (idx << 12) == 0) {
    const uint64_t rest = h << 25;
    const uint32_t rank = (rest ? (uint32_t) __clzll((long long) rest) : 39u) + 1u;
    return idx | (rank << 1) | 1u;
  }
  return idx;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  return (uint64_t) __shfl((unsigned long long) v, src);
}
__device__ __forceinline__ uint32_t shfl32(uint32_t v, int src) {
  return (uint32_t) __shfl((unsigned int) v, src);
}

// ----------------------------------------------------------------------
// Per-fragment hit-count tables

struct LdsTable {  // open addressing in LDS; taxon 0 marks an empty slot
  static constexpr bool kSlotsPerLane = true;  // at most 256 slots: 4 per lane
  uint32_t *key, *cnt;
  uint32_t cap;
  __device__ void clear(int lane) {
    for (uint32_t i = lane; i < cap; i += kWave) {
      key[i] = 0;
      cnt[i] = 0;
    }
  }
  __device__ bool add(uint32_t t, uint32_t c) {
    uint32_t h = (t * 2654435761u) & (cap - 1);
    for (uint32_t n = 0; n < cap; n++) {
      const uint32_t old = atomicCAS(&key[h], 0u, t);
      if (old == 0u || old == t) {
        atomicAdd(&cnt[h], c);
        return true;
      }
      h = (h + 1) & (cap - 1);
    }
    return false;
  }
  __device__ uint32_t get(uint32_t t) const {
    uint32_t h = (t * 2654435761u) & (cap - 1);
    for (uint32_t n = 0; n < cap; n++) {
      const uint32_t k = key[h];
      if (k == t) return cnt[h];
      if (k == 0) return 0;
      h = (h + 1) & (cap - 1);
    }
    return 0;
  }
  __device__ uint32_t slots() const { return cap; }
  __device__ uint32_t slot_key(uint32_t i) const { return key[i]; }
  __device__ uint32_t slot_cnt(uint32_t i) const { return cnt[i]; }
  __device__ void release(int) {}
};

// Loads that bypass a possibly stale vector L1 line: the dense table is
// updated with L2 atomics by every lane of the block.
// A segment of a split fragment counts no hits: the stitch pass rebuilds the
// fragment's table from its joined runs.
struct NullTable {
  __device__ __forceinline__ bool add(uint32_t, uint32_t) { return true; }
};

__device__ __forceinline__ uint32_t load_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct DenseTable {  // overflow path: dense counts per taxon in global memory
  static constexpr bool kSlotsPerLane = false;
  uint32_t *counts, *list, *len;
  __device__ void clear(int) {}
  __device__ bool add(uint32_t t, uint32_t c) {
    if (atomicAdd(&counts[t], c) == 0u) store_agent(&list[atomicAdd(len, 1u)], t);
    return true;
  }
  __device__ uint32_t get(uint32_t t) const { return load_agent(&counts[t]); }
  __device__ uint32_t slots() const { return load_agent(len); }
  __device__ uint32_t slot_key(uint32_t i) const { return load_agent(&list[i]); }
  __device__ uint32_t slot_cnt(uint32_t i) const { return load_agent(&counts[load_agent(&list[i])]); }
  __device__ void release(int lane) {
    const uint32_t n = load_agent(len);
    for (uint32_t i = lane; i < n; i += kWave) store_agent(&counts[load_agent(&list[i])], 0u);
    __syncthreads();
    if (lane == 0) store_agent(len, 0u);
    __syncthreads();
  }
};

// ----------------------------------------------------------------------
// Fragment state carried across chunks and streams (wave-uniform).
struct RunState {
  uint32_t run_code = 0, run_count = 0;
  uint32_t n_runs = 0, nrec = 0, groups = 0;
  uint64_t elems = 0;
  bool done = false;
  // per-lane measurement counters (only with out.lookups)
  uint32_t my_lookups = 0;
  uint64_t my_lines = 0;
};

// Writes run idx of the fragment and adds a taxon run's count to the hit
// table (src/classify.cc:846-888).
template <class Table>
__device__ __forceinline__ void close_run(uint32_t *runs_at, uint64_t cap, uint32_t idx,
                                          uint32_t code, uint32_t count, Table &tab, int *bad) {
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  if (idx < cap) {
    u32x2 r;
    r.x = code;
    r.y = count;
    *(u32x2 *) (runs_at + 2 * (uint64_t) idx) = r;
  } else {
    atomicOr(bad, 2);
  }
  if (code != 0 && code < kCodeAmbig && !tab.add(code, count)) atomicOr(bad, 1);
}

// Appends a border marker element (mate "|:|" or frame "-:-").
template <class Table>
__device__ void push_marker(const FragOut &out, uint64_t base, uint64_t cap, RunState &st,
                            uint32_t code, Table &tab, int *bad, int lane) {
  st.elems++;
  if (st.run_count > 0 && st.run_code == code) {  // adjacent markers (empty frames) share a run
    st.run_count++;
    return;
  }
  if (st.run_count > 0) {
    if (lane == 0) close_run(out.runs + 2 * base, cap, st.n_runs, st.run_code, st.run_count, tab, bad);
    st.n_runs++;
  }
  st.run_code = code;
  st.run_count = 1;
}

// A workgroup is one wavefront, whose LDS operations complete in program
// order: ordering the lanes' LDS accesses only needs a compiler barrier.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

constexpr uint64_t kEvenBits = 0x5555555555555555ULL;
constexpr uint64_t kOddBits = 0xAAAAAAAAAAAAAAAAULL;

// Per-wave LDS.  Lookups of up to kBatchChunks chunks are queued (qkey) and
// probed together (qval), so every lane of the wave probes instead of the
// ~1/3 of lanes whose k-mer starts a new minimizer.
constexpr int kQueue = 192;       // queued lookups per probe batch (>= 64 + 64)
constexpr int kRingGuard = 8;     // windows of up to 9 l-mers read without wrapping
constexpr int kBatchChunks = 8;   // chunks (and mate markers) per probe batch
constexpr uint32_t kMateItem = 0xFFFFFFFFu;   // cqb value of a mate-marker item
constexpr uint32_t kFrameItem = 0xFFFFFFFEu;  // cqb value of a frame-marker item

// KEEP_KEYS: the minimizers stay readable after their probes (HLL records
// need them); otherwise each value overwrites its own key's low half, which
// no lane reads again once that key's probe has started, and the wave's LDS
// shrinks by the value array.
template <bool KEEP_KEYS>
struct WaveLds {
  uint64_t ring[kRingGuard + 128];  // window candidates: position p at kRingGuard + (p & 127),
                                    // slots 120..127 mirrored below index kRingGuard
  uint64_t qkey[kQueue];          // minimizers to look up, in k-mer order
  uint32_t qval_[KEEP_KEYS ? kQueue : 1];  // their table values (KEEP_KEYS)
  uint64_t cmask[kBatchChunks][3];  // per queued chunk: non-ambiguous, LOOKUP, k-mer-end masks
  uint32_t cqb[kBatchChunks];     // per queued chunk: queue index of its first lookup
  uint32_t tkey[256];             // LDS hit table
  uint32_t tcnt[256];
  int bad;
  __device__ __forceinline__ uint32_t &qval(uint32_t i) {
    return KEEP_KEYS ? qval_[i] : reinterpret_cast<uint32_t *>(qkey)[2 * i];
  }
};

// CompactHashTable::Get (src/compact_hash.h:354-374, 493-500) as a resumable
// state for a wave where every lane probes: start() hashes the key and finds
// its home cell; each step() loads one 16-byte group of four cells, decides
// all four without branching, and returns true with the value once the probe
// ends.  Keys whose hash is below min_hash are not looked up
// (src/classify.cc:819-823) and read as 0.
// The probe's table constants, copied once per stream into vector registers.
struct ProbeTable {
  const uint32_t *cells;
  uint64_t capacity, barrett, min_hash;
  // vgpr: hold them in vector registers (the lean kernel variants have the
  // registers to spare; the others keep them scalar)
  __device__ __forceinline__ ProbeTable(const DbView &db, bool vgpr)
      : cells(vgpr ? in_vgpr(db.cells) : db.cells),
        capacity(vgpr ? in_vgpr(db.capacity) : db.capacity),
        barrett(vgpr ? in_vgpr(db.barrett) : db.barrett),
        min_hash(vgpr ? in_vgpr(db.min_hash) : db.min_hash) {}
};

struct ProbeState {
  uint64_t idx = 0, left = 0;
  uint32_t ck = 0;
  bool skip = false;  // hash below min_hash (src/classify.cc:819-823): value 0
  __device__ __forceinline__ void start(const DbView &db, const ProbeTable &t, uint64_t key) {
    const uint64_t hc = fmix64(key);
    skip = hc < t.min_hash;
    ck = (uint32_t) (hc >> (64 - db.key_bits));
    if (db.capacity == 1) {
      idx = 0;
    } else {  // hc % capacity with one correction (see mod_capacity)
      const uint64_t r = hc - __umul64hi(hc, t.barrett) * t.capacity;
      idx = r >= t.capacity ? r - t.capacity : r;
    }
    left = t.capacity;
  }
  // start() from a key hashed ahead (prehash_key).
  __device__ __forceinline__ void start_packed(const ProbeTable &t, uint64_t e) {
    idx = (uint32_t) e;
    ck = (uint32_t) (e >> 32) & 0x7FFFFFFFu;
    skip = (e >> 63) != 0;
    left = t.capacity;
  }
  // One 16-byte group: the cells from idx to the group's end, in probe order.
  __device__ __forceinline__ bool step(const DbView &db, const ProbeTable &t, uint32_t &out) {
    return db.lean_probe ? step_lean(db, t, out) : step_bounded(db, t, out);
  }
  // The padding copies the first cells, so cells read past the end of the
  // table are the ones the probe visits after wrapping, and an empty cell
  // somewhere ends every probe: no bounds or remaining-cell count.  A step
  // reads four cells with one 16-byte load (4-byte aligned; the table is
  // padded by at least four cells): those from idx on, or, when idx is among
  // the last three cells of a 128-byte line, the line's last four (so a load
  // never touches a second line the probe may not need).
  __device__ __forceinline__ bool step_lean(const DbView &db, const ProbeTable &t, uint32_t &out) {
    typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
    const uint32_t vb = db.value_bits;
    const uint32_t vmask = (1u << vb) - 1u;
    const uint64_t line_last4 = (idx | 31u) - 3u;
    const uint64_t from = idx < line_last4 ? idx : line_last4;
    const uint32_t off = (uint32_t) (idx - from);  // 0 except near a line's end
    const u32x4_a4 v = *(const u32x4_a4 *) (t.cells + from);
    const uint32_t c4[4] = {v.x, v.y, v.z, v.w};
    uint32_t val = 0;
    bool stop = false;
#pragma unroll
    for (int j = 3; j >= 0; j--) {  // the lowest stopping cell wins
      const uint32_t c = c4[j];
      const bool st = (uint32_t) j >= off && ((c & vmask) == 0u || (c >> vb) == ck);
      val = st ? (c & vmask) : val;  // an empty cell gives 0
      stop |= st;
    }
    if (stop || skip) {
      out = skip ? 0u : val;
      return true;
    }
    idx = from + 4;
    if (idx >= t.capacity) idx -= t.capacity;
    return false;
  }
  // General tables (no empty cell, or fewer than 64 cells): stops at the
  // table's end and after capacity cells.
  __device__ __forceinline__ bool step_bounded(const DbView &db, const ProbeTable &t, uint32_t &out) {
    const uint32_t vb = db.value_bits;
    const uint32_t vmask = (1u << vb) - 1u;
    const uint64_t group = idx & ~3ULL;
    const u32x4 v = *(const u32x4 *) (t.cells + group);
    const uint32_t off = (uint32_t) (idx - group);
    uint64_t avail = 4u - off;
    if (t.capacity - idx < avail) avail = t.capacity - idx;
    if (left < avail) avail = left;
    const uint32_t c4[4] = {v.x, v.y, v.z, v.w};
    uint32_t stop = 0, val = 0;
#pragma unroll
    for (int j = 3; j >= 0; j--) {
      const uint32_t c = c4[j];
      const bool in = (uint32_t) j >= off && (uint64_t) (j - off) < avail;
      const bool empty = (c & vmask) == 0u;
      const bool hit = (c >> vb) == ck;
      const bool st = in && (empty || hit);
      val = st ? (empty ? 0u : (c & vmask)) : val;
      stop |= st ? 1u : 0u;
    }
    left -= avail;
    idx += avail;
    if (idx >= t.capacity) idx = 0;
    if (stop || left == 0 || skip) {
      out = (stop && !skip) ? val : 0u;
      return true;
    }
    return false;
  }
};

// A queued key hashed ahead of its probe (tables of fewer than 2^32 cells):
// home cell in the low 32 bits, compacted key in bits 32-62 (key_bits <= 31),
// bit 63 set when the hash is below min_hash (src/classify.cc:819-823).
__device__ __forceinline__ uint64_t prehash_key(const DbView &db, const ProbeTable &t, uint64_t key) {
  const uint64_t hc = fmix64(key);
  const uint64_t skip = hc < t.min_hash ? 1ULL : 0ULL;
  const uint64_t ck = hc >> (64 - db.key_bits);
  uint64_t idx = 0;
  if (db.capacity != 1) {  // hc % capacity with one correction (see mod_capacity)
    const uint64_t r = hc - __umul64hi(hc, t.barrett) * t.capacity;
    idx = r >= t.capacity ? r - t.capacity : r;
  }
  return idx | (ck << 32) | (skip << 63);
}

// The last probes of a batch, at most kTailProbes left and no key waiting
// (lean tables only: an empty cell ends every probe): probe j of the na
// left takes the G = 64 / 2^ceil(log2 na) lanes j*G .. j*G+G-1, lane i of
// them reading the 4 cells idx + 4i .. idx + 4i + 3 of its probe sequence,
// so a round covers 4G consecutive cells.  The lowest lane of the group with
// an empty or matching cell gives the value (the first such cell in probe
// order); otherwise the group moves on by 4G cells.  A linear probe at load
// 0.7 averages 6 cells but the longest of a batch of a hundred runs to about
// 50, which one lane would walk 4 cells per round.
constexpr uint32_t kTailProbes = 16;
static_assert(kTailProbes <= 16, "probe_tail's groups are at least 4 lanes: at most 16 probes");
__device__ __forceinline__ void probe_tail(const DbView &db, const ProbeTable &t, uint64_t A, const ProbeState &ps,
                                        uint32_t q, uint32_t *qval, uint32_t qstride, int lane) {
  typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
  const uint32_t na = (uint32_t) __popcll(A);
  const uint32_t lg = na <= 1 ? 0u : (na <= 2 ? 1u : (na <= 4 ? 2u : (na <= 8 ? 3u : 4u)));
  const uint32_t G = 64u >> lg;
  const uint32_t j = (uint32_t) lane >> (6 - lg), i = (uint32_t) lane & (G - 1u);
  uint64_t m = A;
  for (uint32_t c = 0; c < j && m; c++) m &= m - 1;  // the j-th probe left
  const bool has = j < na;
  const int src = m ? __ffsll((long long) m) - 1 : 0;
  uint64_t idx = shfl64(ps.idx, src);
  const uint32_t ck = shfl32(ps.ck, src);
  const uint32_t qq = shfl32(q, src);
  const uint64_t gmask = G == 64u ? ~0ULL : (((1ULL << G) - 1ULL) << (j * G));
  const uint32_t vb = db.value_bits;
  const uint32_t vmask = (1u << vb) - 1u;
  bool done = !has;
  uint32_t result = 0;
  while (__ballot(!done)) {
    uint32_t pos = 4u, val = 0;  // first stopping cell of this lane's four (4: none)
    if (!done) {
      uint64_t g = idx + 4u * i;
      while (g >= t.capacity) g -= t.capacity;  // capacity >= 64, 4i < 256
      // cells g .. g+3; past the end the padding holds copies of the first cells
      const u32x4_a4 v = *(const u32x4_a4 *) (t.cells + g);
      const uint32_t c4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int c = 3; c >= 0; c--) {
        const bool st = (c4[c] & vmask) == 0u || (c4[c] >> vb) == ck;
        pos = st ? (uint32_t) c : pos;
        val = st ? (c4[c] & vmask) : val;
      }
    }
    const uint64_t S = __ballot(pos < 4u) & gmask;
    const uint32_t got = shfl32(val, S ? __ffsll((long long) S) - 1 : lane);
    if (!done) {
      if (S) {
        done = true;
        result = got;
      } else {
        idx += 4u * G;
        while (idx >= t.capacity) idx -= t.capacity;
      }
    }
  }
  if (has && i == 0) qval[qq * qstride] = result;
}

// Probes queue entries [0, qn): each lane works one key at a time and takes
// the next queued key as soon as its probe ends (values land at
// qval[q * qstride]).  PACKED: the entries were hashed ahead (prehash_key);
// LEAN: db.lean_probe, and the last probes are split over the idle lanes
// (probe_tail).  on_start(q, ps) runs as a lane starts entry q.
struct NoStartHook {
  __device__ __forceinline__ void operator()(uint32_t, const ProbeState &) const {}
};
template <bool PACKED, bool LEAN, class OnStart = NoStartHook>
__device__ __forceinline__ void probe_loop(const DbView &db, const ProbeTable &ptab, const uint64_t *qkey,
                                           uint32_t *qval, uint32_t qstride, uint32_t qn, int lane,
                                           OnStart on_start = OnStart()) {
  ProbeState ps;
  uint32_t q = (uint32_t) lane;
  bool active = q < qn;
  if (active) {
    if (PACKED)
      ps.start_packed(ptab, qkey[q]);
    else
      ps.start(db, ptab, qkey[q]);
    on_start(q, ps);
  }
  uint32_t next = qn < (uint32_t) kWave ? qn : (uint32_t) kWave;
  for (;;) {
    const uint64_t A = __ballot(active);
    if (!A) break;
    if (LEAN && next >= qn && (uint32_t) __popcll(A) <= kTailProbes) {
      probe_tail(db, ptab, A, ps, q, qval, qstride, lane);
      break;
    }
    if (active) {
      uint32_t val;
      if (LEAN ? ps.step_lean(db, ptab, val) : ps.step_bounded(db, ptab, val)) {
        qval[q * qstride] = val;
        active = false;
      }
    }
    const uint64_t R = __ballot(!active) & (next < qn ? ~0ULL : 0ULL);
    if (R) {
      if (lane_bit(R)) {
        q = next + (uint32_t) __popcll(R & lanemask_lt(lane));
        if (q < qn) {
          active = true;
          if (PACKED)
            ps.start_packed(ptab, qkey[q]);
          else
            ps.start(db, ptab, qkey[q]);
          on_start(q, ps);
        }
      }
      next += (uint32_t) __popcll(R);
      if (next > qn) next = qn;
    }
  }
}

// probe_loop for the table at hand: below 2^32 cells the keys are hashed
// first, densely (each entry turned into its packed home cell and compacted
// key), so a lane that takes a new key only reads it.
__device__ __forceinline__ void probe_queue(const DbView &db, const ProbeTable &ptab, uint64_t *qkey,
                                            uint32_t *qval, uint32_t qstride, uint32_t qn, int lane) {
  if (db.capacity <= 0xFFFFFFFFull) {
    for (uint32_t e = (uint32_t) lane; e < qn; e += kWave) qkey[e] = prehash_key(db, ptab, qkey[e]);
    wave_sync();
    if (db.lean_probe)
      probe_loop<true, true>(db, ptab, qkey, qval, qstride, qn, lane);
    else
      probe_loop<true, false>(db, ptab, qkey, qval, qstride, qn, lane);
  } else {
    if (db.lean_probe)
      probe_loop<false, true>(db, ptab, qkey, qval, qstride, qn, lane);
    else
      probe_loop<false, false>(db, ptab, qkey, qval, qstride, qn, lane);
  }
}

// Scans a fragment's streams in order -- its mates (nucleotide), or the six
// reading frames of each mate (translated), with a marker element between
// streams (src/classify.cc:797-838) -- and appends their k-mer elements to
// the fragment's runs.  Every stream starts a fresh scanner; the streams
// share probe batches, so a short read's frames fill one batch's lanes.  DNA: a nucleotide database
// (else translated search); EXTRAS: per-hit HLL records, per-taxon counts or
// probe statistics may be requested (their code is compiled out otherwise);
// W: the window's l-mer count k-l+1 when fixed at compile time (<= 9), else 0.
// first_emit: k-mers ending before this position of the stream are scanned
// (they set the previous minimizer and its taxon) but are not elements of
// the fragment: no run, hit, record or count (0 for a whole stream).
//
// The stream is processed in probe batches of up to kBatchChunks chunks of
// 64 k-mer end positions:
//   A. each chunk's minimizers and LOOKUP/REPEAT/AMBIG decisions
//      (src/classify.cc:797-838) are computed and its LOOKUP minimizers
//      appended to the wave's queue, with the chunk's masks;
//   B. every lane probes queued keys (64 per round);
//   C. the chunks are replayed in order: each k-mer takes the value of the
//      latest LOOKUP at or before it (src/classify.cc:846-888) and the wave
//      run-length encodes the elements and counts hits.
template <bool DNA, bool EXTRAS, int W, class Table>
__device__ void scan_stream(const DbView &db, const ClassifyParams &P, const FragOut &out,
                            uint64_t base, uint64_t cap, const char *m1, uint64_t L1, const char *m2,
                            uint64_t L2, RunState &st, Table &tab, WaveLds<EXTRAS> &lds, int *bad, int lane,
                            uint32_t first_emit = 0) {
  static_assert(W >= 0 && W <= kRingGuard + 1, "window too wide for the ring guard");
  constexpr bool dna = DNA;
  constexpr uint32_t F = dna ? 1u : 6u;  // streams per mate
  // streams: (mate, frame) = (si / F, si % F); m2 null: one mate
  const uint32_t n_streams = (m2 ? 2u : 1u) * F;
  uint32_t si = 0;
  const char *m = m1;
  uint64_t L = L1;
  int fr = 0;
  // positions within a stream are 32-bit: the library rejects reads of 2^31
  // or more bases before launching
  uint32_t n = (uint32_t) (dna ? L : frame_len(L, fr));
  const uint32_t k = db.k, l = db.l;
  const uint32_t w = k - l + 1;  // l-mers per k-mer window
  const uint64_t lbits = (1ULL << l) - 1ULL;
  int32_t last_amb = -1;
  uint64_t last_min = ~0ULL;
  uint32_t last_taxon = 0;
  uint64_t pp0 = 0, pp1 = 0, pp2 = 0, pp3 = 0;  // previous chunk's planes (translated)
  uint64_t pBEh = 0, pLEh = 0;                   // previous chunk's upper stream words (DNA)
  const uint64_t lmask2 = (1ULL << (2 * l)) - 1ULL;
  constexpr bool kVgprConsts = !EXTRAS;
  const uint64_t tog = !kVgprConsts ? db.toggle_mask : in_vgpr(db.toggle_mask);
  const uint64_t smask = !kVgprConsts ? db.spaced_mask : in_vgpr(db.spaced_mask);
  const ProbeTable ptab(db, kVgprConsts);
  uint32_t *const runs_at = !kVgprConsts ? out.runs + 2 * base : in_vgpr(out.runs + 2 * base);
  const uint64_t cap_v = !kVgprConsts ? cap : in_vgpr(cap);
  uint32_t c = 0;
  bool more = true;
  while (more) {
    // ---- A: scan chunks, queue their lookups
    uint32_t qn = 0, g = 0;
    // wave-uniform, clamped to [1, kBatchChunks] (0 would queue nothing)
    const uint32_t depth = P.batch_chunks - 1u < (uint32_t) kBatchChunks ? P.batch_chunks : (uint32_t) kBatchChunks;
    while (g < depth && qn + kWave <= (uint32_t) kQueue) {
      if (c >= n) {  // this stream is done
        if (si + 1 == n_streams) {
          more = false;
          break;
        }
        // the marker, then the next stream with a fresh scanner state
        si++;
        if (lane == 0) lds.cqb[g] = si % F == 0 ? kMateItem : kFrameItem;
        g++;
        m = si >= F ? m2 : m1;
        L = si >= F ? L2 : L1;
        fr = (int) (si % F);
        n = (uint32_t) (dna ? L : frame_len(L, fr));
        c = 0;
        last_amb = -1;
        last_min = ~0ULL;
        pBEh = pLEh = 0;
        pp0 = pp1 = pp2 = pp3 = 0;
        continue;
      }
      const uint32_t p = c + (uint32_t) lane;
      const bool inb = p < n;
      uint32_t code = 0xFFu;
      if (inb) code = dna ? dna_code((uint8_t) m[p]) : frame_code(m, L, fr, p);
      const bool amb = inb && code == 0xFFu;
      const bool good = inb && !amb;
      const uint64_t P0 = __ballot(good && (code & 1u));
      const uint64_t P1 = __ballot(good && (code & 2u));
      const uint64_t P2 = dna ? 0 : __ballot(good && (code & 4u));
      const uint64_t P3 = dna ? 0 : __ballot(good && (code & 8u));
      const uint64_t A = __ballot(amb);
      // A clean chunk (no ambiguous base in it or in the k-1 positions before
      // it, and past the stream's first k-1 positions) has a valid l-mer and a
      // non-ambiguous k-mer at every position in bounds.
      const bool clean = A == 0 && c + 1 >= k && last_amb < (int32_t) (c + 1) - (int32_t) k;
      bool lvalid, kvalid, nonamb;
      if (clean) {
        lvalid = kvalid = nonamb = inb;
      } else {
        const uint64_t mA = A & lanemask_le(lane);
        const int32_t lastN = mA ? (int32_t) (c + (uint32_t) top_lane(mA)) : last_amb;
        lvalid = inb && p + 1 >= l && lastN < (int32_t) (p + 1) - (int32_t) l;
        kvalid = inb && p + 1 >= k;
        nonamb = kvalid && ((k == l) ? lvalid : (lastN < (int32_t) (p + 2) - (int32_t) k));
      }

      // candidate of the l-mer ending at p
      uint64_t cand = kNoCand;
      if (dna) {
        // The chunk's bases as packed 2-bit streams (wave-uniform): big-endian
        // (first base in the top bits of its 32-base word) for forward l-mers,
        // little-endian complemented bases for reverse complements; each lane
        // then cuts its l-mer out of two adjacent words.
        const uint64_t r0l = replicate_bits((uint32_t) P0), r0h = replicate_bits((uint32_t) (P0 >> 32));
        const uint64_t r1l = replicate_bits((uint32_t) P1), r1h = replicate_bits((uint32_t) (P1 >> 32));
        const uint64_t LEl = ~((r0l & kEvenBits) | (r1l & kOddBits));
        const uint64_t LEh = ~((r0h & kEvenBits) | (r1h & kOddBits));
        const uint64_t BEl = __builtin_bitreverse64((r1l & kEvenBits) | (r0l & kOddBits));
        const uint64_t BEh = __builtin_bitreverse64((r1h & kEvenBits) | (r0h & kOddBits));
        if (lvalid) {
          const bool up = lane >= 32;
          const uint32_t s = 62u - 2u * (uint32_t) (lane & 31);
          const uint64_t flo = up ? BEh : BEl, fhi = up ? BEl : pBEh;
          const uint64_t fwd = ((flo >> s) | ((fhi << 1) << (63u - s))) & lmask2;
          const uint64_t top = up ? LEh : LEl, bot = up ? LEl : pLEh;
          const uint64_t t = (top << s) | ((bot >> 1) >> (63u - s));
          const uint64_t rc = t >> (64u - 2u * l);
          uint64_t canon;
          if (W == 0 && db.revcom_version == 0) {  // src/mmscanner.cc:193-215, old databases
            const uint32_t d = 2u * (32u - l);
            const uint64_t rc0 = ((rc << d) | ((1ULL << d) - 1ULL)) & lmask2;
            canon = fwd < rc0 ? fwd : rc0;
          } else {
            canon = fwd < rc ? fwd : rc;
          }
          if (db.spaced_mask) canon &= smask;
          cand = canon ^ tog;
        }
        pBEh = BEh;
        pLEh = LEh;
      } else if (lvalid) {
        const uint32_t sh = (uint32_t) lane + 65u - l;
        const uint64_t w0 = window128(pp0, P0, sh) & lbits;
        const uint64_t w1 = window128(pp1, P1, sh) & lbits;
        const uint64_t w2 = window128(pp2, P2, sh) & lbits;
        const uint64_t w3 = window128(pp3, P3, sh) & lbits;
        const uint64_t r0 = __builtin_bitreverse64(w0) >> (64 - l);
        const uint64_t r1 = __builtin_bitreverse64(w1) >> (64 - l);
        const uint64_t r2 = __builtin_bitreverse64(w2) >> (64 - l);
        const uint64_t r3 = __builtin_bitreverse64(w3) >> (64 - l);
        cand = (spread4(r3) << 3) | (spread4(r2) << 2) | (spread4(r1) << 1) | spread4(r0);
        if (db.spaced_mask) cand &= smask;
        cand ^= tog;
      }
      const uint32_t slot = (uint32_t) (p & 127);
      lds.ring[kRingGuard + slot] = cand;
      if (slot >= 128u - kRingGuard) lds.ring[slot - (128u - kRingGuard)] = cand;
      wave_sync();

      // k-mer ending at p: minimum over its window of l-mer candidates
      uint64_t mn = ~0ULL;
      if (W != 0) {  // compile-time window: reads at fixed offsets from one address
        const uint64_t *r = &lds.ring[kRingGuard + slot - (W - 1)];
#pragma unroll
        for (int j = 0; j < W; j++) mn = r[j] < mn ? r[j] : mn;
      } else {
        for (uint32_t j = 0; j < w; j++) {
          const uint64_t v = lds.ring[kRingGuard + ((p - j) & 127)];
          mn = v < mn ? v : mn;
        }
      }
      mn ^= tog;

      // token decision against the previous non-ambiguous minimizer (in a
      // clean chunk the lanes in bounds, the next lower lane for each)
      const uint64_t NA = __ballot(nonamb);
      uint64_t prev_min;
      if (clean) {
        prev_min = wave_shr1(mn, last_min);
      } else {
        const uint64_t prevm = NA & lanemask_lt(lane);
        const uint64_t before = shfl64(mn, prevm ? top_lane(prevm) : lane);
        prev_min = prevm ? before : last_min;
      }
      const bool lookup = nonamb && mn != prev_min;
      const uint64_t LK = __ballot(lookup);
      const uint64_t E = __ballot(kvalid && p >= first_emit);
      if (lookup) lds.qkey[qn + __popcll(LK & lanemask_lt(lane))] = mn;
      if (lane == 0) {
        lds.cmask[g][0] = NA;
        lds.cmask[g][1] = LK;
        lds.cmask[g][2] = E;
        lds.cqb[g] = qn;
      }
      qn += (uint32_t) __popcll(LK);
      g++;
      if (A) last_amb = (int32_t) (c + (uint32_t) top_lane(A));
      if (NA) last_min = readlane64(mn, top_lane(NA));
      pp0 = P0;
      pp1 = P1;
      pp2 = P2;
      pp3 = P3;
      c += kWave;
      wave_sync();  // the ring slots of this chunk are read by the next one
    }

    // ---- B: probe the queued minimizers.  Each lane works one key at a time
    // and takes the next queued key as soon as its probe ends, so a step's
    // round trip serves up to 64 probes however unequal their lengths.
    if constexpr (!EXTRAS) {
      probe_queue(db, ptab, lds.qkey, reinterpret_cast<uint32_t *>(lds.qkey), 2, qn, lane);
    } else {  // the keys stay readable (HLL records); probe statistics per lane
      auto stats = [&](uint32_t q, const ProbeState &ps) {
        if (out.lookups && !ps.skip) {
          st.my_lookups++;
          st.my_lines += table_lines(db, lds.qkey[q]);
        }
      };
      if (db.lean_probe)
        probe_loop<false, true>(db, ptab, lds.qkey, &lds.qval(0), 1, qn, lane, stats);
      else
        probe_loop<false, false>(db, ptab, lds.qkey, &lds.qval(0), 1, qn, lane, stats);
    }
    wave_sync();

    // ---- C: replay the chunks in order
    for (uint32_t gi = 0; gi < g && !st.done; gi++) {
      const uint64_t NA = uniform64(lds.cmask[gi][0]);
      const uint64_t LK = uniform64(lds.cmask[gi][1]);
      uint64_t E = uniform64(lds.cmask[gi][2]);
      const uint32_t qb = uniform32(lds.cqb[gi]);
      if (qb == kMateItem || qb == kFrameItem) {  // a new stream: its own latest-LOOKUP value
        push_marker(out, base, cap, st, qb == kMateItem ? kCodeMate : kCodeFrame, tab, bad, lane);
        last_taxon = 0;
        continue;
      }
      const bool nonamb = lane_bit(NA);
      const bool lookup = lane_bit(LK);
      const bool kvalid = lane_bit(E);
      const uint64_t srcm = LK & lanemask_le(lane);
      const uint32_t tq = srcm ? lds.qval(qb + __popcll(srcm) - 1) : last_taxon;
      const uint32_t tax_i = nonamb ? tq : 0u;
      uint64_t HIT = __ballot(lookup && kvalid && tq != 0);

      if (EXTRAS && P.quick) {  // src/classify.cc:880-884: stop at the first qualifying hit
        const uint32_t g_le = st.groups + __popcll(HIT & lanemask_le(lane));
        const uint64_t Q = __ballot(kvalid && tax_i != 0 && (int32_t) g_le >= P.min_hit_groups);
        if (Q) {
          const int cut = __ffsll((long long) Q) - 1;
          E &= lanemask_lt(cut);
          HIT &= lanemask_le(cut);
          st.done = true;
        }
      }

      // per-taxon k-mer counts for LOOKUP hits: one atomic per distinct
      // taxon among the chunk's hits (neighbouring k-mers mostly share one)
      if (EXTRAS && out.kmer_counts && HIT) {
        uint64_t left = HIT;
        while (left) {
          const int f = __ffsll((long long) left) - 1;
          const uint32_t t = readlane32(tq, f);
          const uint64_t same = __ballot(lane_bit(left) && tq == t);
          if (lane == f) atomicAdd(&out.kmer_counts[t], (unsigned long long) __popcll(same));
          left &= ~same;
        }
      }
      // HLL records for LOOKUP hits
      if (EXTRAS && lane_bit(HIT)) {
        if (out.recs) {
          const uint32_t r = st.nrec + __popcll(HIT & lanemask_lt(lane));
          if (r < cap) {
            out.recs[2 * (base + r)] = tq;
            out.recs[2 * (base + r) + 1] = hll_encode(lds.qkey[qb + __popcll(srcm) - 1]);
          } else {
            atomicOr(bad, 2);
          }
        }
      }
      st.nrec += __popcll(HIT);
      st.groups += __popcll(HIT);

      // run-length encode the elements of this chunk
      if (E) {
        const bool isE = lane_bit(E);
        const uint32_t code_i = nonamb ? tax_i : kCodeAmbig;
        // E is a contiguous run of lanes and the code is constant between
        // boundaries, so the run a boundary closes has the lower lane's code
        const uint32_t prev_code = wave_shr1_32(code_i, code_i);
        const int lo = __ffsll((long long) E) - 1;
        const bool boundary =
            isE && (lane == lo ? (st.run_count == 0 || code_i != st.run_code) : code_i != prev_code);
        const uint64_t B = __ballot(boundary);
        const uint64_t pb = B & lanemask_lt(lane);
        const uint32_t drop = (st.run_count == 0) ? 1u : 0u;
        if (boundary) {
          uint32_t cc, cn;
          bool has = true;
          if (pb) {
            cc = prev_code;
            cn = (uint32_t) (lane - top_lane(pb));
          } else {
            cc = st.run_code;
            cn = st.run_count + (uint32_t) (lane - lo);
            has = st.run_count > 0;
          }
          if (has) {
            const uint32_t idx = st.n_runs + __popcll(pb) - drop;
            close_run(runs_at, cap_v, idx, cc, cn, tab, bad);
          }
        }
        st.n_runs += (uint32_t) __popcll(B) - (B ? drop : 0u);
        const int hi = top_lane(E);
        if (B) {
          const int lb = top_lane(B);
          st.run_code = readlane32(code_i, lb);
          st.run_count = (uint32_t) (hi - lb + 1);
        } else {
          st.run_count += (uint32_t) __popcll(E);
        }
        st.elems += (uint64_t) __popcll(E);
      }
      if (LK) last_taxon = uniform32(lds.qval(qb + __popcll(LK) - 1));
    }
    wave_sync();  // the queue is refilled by the next batch
    if (st.done) break;
  }
}

// Wave-wide (score, taxon) reduction: higher score wins, ties take the LCA
// (src/classify.cc:718-740).
__device__ __forceinline__ void reduce_best(const uint32_t *parent, uint32_t &score,
                                            uint32_t &taxon) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t os = (uint32_t) __shfl_xor((unsigned int) score, d);
    const uint32_t ot = (uint32_t) __shfl_xor((unsigned int) taxon, d);
    if (os > score) {
      score = os;
      taxon = ot;
    } else if (os == score) {
      taxon = lca(parent, taxon, ot);
    }
  }
}

__device__ __forceinline__ uint32_t reduce_sum(uint32_t v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += (uint32_t) __shfl_xor((unsigned int) v, d);
  return v;
}

template <bool EXTRAS, class Table>
__device__ void resolve_fragment(const DbView &db, const ClassifyParams &P, const FragOut &out, uint64_t f,
                                 const RunState &st, Table &tab, bool paired, bool tx, int *bad, int lane);

template <bool DNA, bool EXTRAS, class Table>
__device__ void classify_fragment(const DbView &db, const ClassifyParams &P, const FragBatch &in,
                                  const FragOut &out, uint64_t f, Table &tab, WaveLds<EXTRAS> &lds,
                                  int *bad, int lane) {
  const uint64_t o0 = uniform64(in.off[2 * f]), o1 = uniform64(in.off[2 * f + 1]),
                 o2 = uniform64(in.off[2 * f + 2]);
  const bool paired = uniform32(in.paired ? in.paired[f] : in.paired_all) != 0;
  constexpr bool tx = !DNA;
  const uint64_t base = uniform64(out.run_off[f]);
  const uint64_t cap = uniform64(out.run_cap[f]);
  tab.clear(lane);
  if (lane == 0) *bad = 0;
  __syncthreads();

  RunState st;
  // both mates (and, translated, all six frames of each) in one scan, the
  // markers between them
  const char *m2 = paired ? in.seq + o1 : nullptr;
  if (!tx && db.k - db.l == 4 && db.revcom_version != 0)  // the default window (k=35, l=31)
    scan_stream<DNA, EXTRAS, 5>(db, P, out, base, cap, in.seq + o0, o1 - o0, m2, o2 - o1, st, tab, lds, bad,
                                lane);
  else
    scan_stream<DNA, EXTRAS, 0>(db, P, out, base, cap, in.seq + o0, o1 - o0, m2, o2 - o1, st, tab, lds, bad,
                                lane);
  if (st.run_count > 0) {
    if (lane == 0) close_run(out.runs + 2 * base, cap, st.n_runs, st.run_code, st.run_count, tab, bad);
    st.n_runs++;
  }
  __syncthreads();
  resolve_fragment<EXTRAS>(db, P, out, f, st, tab, paired, tx, bad, lane);
}

// ResolveTree (src/classify.cc:710-762) over a fragment's hit table, the
// confidence and minimum-hit-group rules (src/classify.cc:890-905), and the
// fragment's per-fragment outputs.
template <bool EXTRAS, class Table>
__device__ void resolve_fragment(const DbView &db, const ClassifyParams &P, const FragOut &out, uint64_t f,
                                 const RunState &st, Table &tab, bool paired, bool tx, int *bad, int lane) {
  uint64_t total = st.elems;
  if (paired) total--;
  if (tx) total -= paired ? 4 : 2;
  const uint32_t required = (uint32_t) ceil(P.confidence * (double) total);
  uint32_t best = 0, best_t = 0;
  const uint32_t ns = tab.slots();
  // LDS table with at most 64 hit taxa: each lane holds its (up to) four
  // slots with their taxa's preorder spans, and a taxon's score (the counts
  // of its hit ancestors-or-self) is summed over the hit taxa whose span
  // holds it -- no walks up the parent array.  Otherwise (more taxa, or the
  // dense table) every taxon walks its ancestors.
  uint32_t skey[4] = {0, 0, 0, 0}, scnt[4] = {0, 0, 0, 0}, stin[4] = {0, 0, 0, 0};
  bool by_span = false;
  if constexpr (Table::kSlotsPerLane) {
    uint64_t occ[4];
    uint32_t stout[4], hits = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint32_t i = (uint32_t) lane + (uint32_t) (kWave * r);
      const uint32_t t = i < ns ? tab.slot_key(i) : 0u;
      skey[r] = t;
      scnt[r] = t ? tab.slot_cnt(i) : 0u;
      stin[r] = stout[r] = 0;
      if (t) {
        stin[r] = db.span[2 * (uint64_t) t];
        stout[r] = db.span[2 * (uint64_t) t + 1];
      }
      occ[r] = __ballot(t != 0);
      hits += (uint32_t) __popcll(occ[r]);
    }
    if (hits <= (uint32_t) kWave) {
      by_span = true;
      uint32_t sc[4] = {0, 0, 0, 0};
#pragma unroll
      for (int r2 = 0; r2 < 4; r2++) {
        for (uint64_t m = occ[r2]; m; m &= m - 1) {
          const int j = __ffsll((long long) m) - 1;
          const uint32_t at = readlane32(stin[r2], j), ao = readlane32(stout[r2], j);
          const uint32_t ac = readlane32(scnt[r2], j);
#pragma unroll
          for (int r = 0; r < 4; r++) sc[r] += (at <= stin[r] && stin[r] < ao) ? ac : 0u;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (!skey[r]) continue;
        if (sc[r] > best) {
          best = sc[r];
          best_t = skey[r];
        } else if (sc[r] == best) {
          best_t = lca(db.parent, best_t, skey[r]);
        }
      }
    }
  }
  if (!by_span) {
    for (uint32_t i = lane; i < ns; i += kWave) {
      const uint32_t t = tab.slot_key(i);
      if (t == 0) continue;
      uint32_t s = 0;
      for (uint32_t a = t; a != 0; a = db.parent[a]) s += tab.get(a);
      if (s > best) {
        best = s;
        best_t = t;
      } else if (s == best) {
        best_t = lca(db.parent, best_t, t);
      }
    }
  }
  reduce_best(db.parent, best, best_t);
  uint32_t call = uniform32(best_t);
  uint32_t score = uniform32(call ? tab.get(call) : 0);
  while (call && score < required) {
    uint32_t part = 0;
    if (by_span) {
      const uint32_t ct = uniform32(db.span[2 * (uint64_t) call]);
      const uint32_t co = uniform32(db.span[2 * (uint64_t) call + 1]);
#pragma unroll
      for (int r = 0; r < 4; r++) part += (skey[r] && ct <= stin[r] && stin[r] < co) ? scnt[r] : 0u;
    } else {
      for (uint32_t i = lane; i < ns; i += kWave) {
        const uint32_t t = tab.slot_key(i);
        if (t != 0 && is_ancestor(db.parent, call, t)) part += tab.slot_cnt(i);
      }
    }
    score = uniform32(reduce_sum(part));
    if (score >= required) break;
    call = uniform32(db.parent[call]);
  }
  if (call && (int64_t) st.groups < (int64_t) P.min_hit_groups) call = 0;
  __syncthreads();
  if (EXTRAS && out.lookups) {
    uint64_t lk = st.my_lookups, ln = st.my_lines;
    for (int d = 1; d < kWave; d <<= 1) {
      lk += shfl64(lk, lane ^ d);
      ln += shfl64(ln, lane ^ d);
    }
    if (lane == 0) {
      out.lookups[f] = (uint32_t) lk;
      out.lines[f] = ln;
    }
  }
  if (lane == 0) {
    out.call[f] = call;
    out.groups[f] = st.groups;
    out.total[f] = (uint32_t) total;
    out.n_runs[f] = st.n_runs;
    out.n_recs[f] = st.nrec;
    out.status[f] = (uint32_t) *bad;
  }
  tab.release(lane);
}


// Plain path: one wavefront per fragment, one launch per batch.
template <bool DNA, bool EXTRAS>
__device__ __forceinline__ void plain_body(const DbView &db, const ClassifyParams &P,
                                           const FragBatch &in, const FragOut &out) {
  __shared__ WaveLds<EXTRAS> lds;
  const uint64_t f = blockIdx.x;
  if (f >= in.n) return;
  LdsTable tab{lds.tkey, lds.tcnt, P.tab_cap};
  classify_fragment<DNA, EXTRAS>(db, P, in, out, f, tab, lds, &lds.bad, threadIdx.x);
}

// Persistent path: resident wavefronts pull work units (runs of fragments of
// similar length, longest first) from shared cursors, so long and short reads
// share one launch.  The units are dealt to kCursors cursors round robin
// (unit u to cursor u % kCursors, wave w to cursor w % kCursors) so the
// waves' atomics spread over several cache lines; the grid has at least
// min(kCursors, units) waves.

template <bool DNA, bool EXTRAS>
__device__ __forceinline__ void persistent_body(const DbView &db, const ClassifyParams &P,
                                                const FragBatch &in, const FragOut &out,
                                                unsigned long long *cursor) {
  __shared__ WaveLds<EXTRAS> lds;
  LdsTable tab{lds.tkey, lds.tcnt, P.tab_cap};
  const int lane = threadIdx.x;
  const uint64_t n_units = in.unit_start ? uniform32(*in.n_units) : in.n;
  const uint32_t part = blockIdx.x % kCursors;
  unsigned long long *cur = cursor + kCursorStride * part;
  // Every lane adds 1, so each fetch advances the cursor by exactly 64 and
  // any lane's result / 64 is the wave's next unit of its share; no lane-0
  // branch precedes the loop exit and the index is a scalar.  The next fetch
  // is issued before the current unit is classified so its latency hides
  // behind that work.
  uint64_t u = kCursors * (uniform64((uint64_t) atomicAdd(cur, 1ULL)) >> 6) + part;
  while (u < n_units) {
    const unsigned long long next = atomicAdd(cur, 1ULL);
    const uint64_t lo = in.unit_start ? uniform32(in.unit_start[u]) : u;
    const uint64_t hi = in.unit_start ? uniform32(in.unit_start[u + 1]) : u + 1;
    for (uint64_t j = lo; j < hi; j++) {
      const uint64_t f = uniform64(in.order ? (uint64_t) in.order[j] : j);
      if (in.split && uniform32(in.split[f]) != 0) continue;  // classified by segments
      classify_fragment<DNA, EXTRAS>(db, P, in, out, f, tab, lds, &lds.bad, lane);
    }
    u = kCursors * (uniform64((uint64_t) next) >> 6) + part;
  }
}

// Overflow path: fragments whose distinct hit taxa exceeded the LDS table are
// classified again with a dense per-taxon count array (one per block).
// The number of listed fragments is read from device memory (*n_redo), so
// the launch needs no host round trip; a grid of any size covers them.
template <bool DNA>
__device__ __forceinline__ void dense_body(const DbView &db, const ClassifyParams &P,
                                           const FragBatch &in, const FragOut &out,
                                           const uint32_t *redo, const uint32_t *n_redo_p,
                                           const DenseScratch &scratch) {
  __shared__ WaveLds<true> lds;
  const int lane = threadIdx.x;
  const uint32_t slot = blockIdx.x;
  const uint32_t n_redo = uniform32(*n_redo_p);
  if (slot >= n_redo) return;
  DenseTable tab{scratch.counts + (uint64_t) slot * db.node_count,
                 scratch.lists + (uint64_t) slot * db.node_count, scratch.lens + slot};
  // the slot's table starts empty (its memory may hold anything); after each
  // fragment release() clears the entries that fragment used
  for (uint32_t t = lane; t < db.node_count; t += kWave) store_agent(&tab.counts[t], 0u);
  if (lane == 0) store_agent(tab.len, 0u);
  __syncthreads();
  for (uint32_t i = blockIdx.x; i < n_redo; i += gridDim.x)
    classify_fragment<DNA, true>(db, P, in, out, uniform32(redo[i]), tab, lds, &lds.bad, lane);
}

// ---- long fragments split over waves: the segment layout (see SegBatch)
__device__ __host__ inline uint32_t mate_kmers(uint64_t L, uint32_t k) { return L >= k ? (uint32_t) (L - k + 1) : 0u; }

__device__ __host__ inline uint32_t mate_segments(uint32_t e, uint32_t seg) { return e ? (e + seg - 1) / seg : 1u; }

// Step 1: per fragment, its segment count (0: not split); the split
// fragments listed; cnt[n] = 0 for the scan.
__global__ __launch_bounds__(256) void k_seg_count(const uint64_t *off, const uint8_t *paired, uint32_t paired_all,
                                                   uint64_t n, uint32_t k, uint32_t seg, const uint64_t *elements,
                                                   uint32_t *cnt, uint32_t *list, uint32_t *count) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    cnt[n] = 0;
    return;
  }
  uint32_t c = 0;
  if (elements[i] > 2ull * seg) {
    const bool pr = paired ? paired[i] != 0 : paired_all != 0;
    c = mate_segments(mate_kmers(off[2 * i + 1] - off[2 * i], k), seg);
    if (pr) c += mate_segments(mate_kmers(off[2 * i + 2] - off[2 * i + 1], k), seg);
    list[atomicAdd(&count[1], 1u)] = (uint32_t) i;
  }
  cnt[i] = c;
}

// Step 1 inside a graph-replayed pipeline: the batch from its descriptor,
// each fragment's elements from the bounds pass, and the split only when the
// batch's longest fragment outweighs a wave's even share of the batch
// (stats[0] largest, [1] total elements, of `waves`) and the split
// fragments' elements fit the scratch (stats[2] <= scratch).
__global__ __launch_bounds__(256) void k_seg_count_pipe(const PipeDesc *d, uint32_t k, uint32_t seg,
                                                        const uint64_t *elements, const unsigned long long *stats,
                                                        uint64_t waves, uint64_t scratch, uint32_t *cnt,
                                                        uint32_t *list, uint32_t *count) {
  const uint64_t n = d->n;
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    cnt[n] = 0;
    return;
  }
  const bool want = stats[0] > 2ull * seg && stats[0] * waves > stats[1] && stats[2] <= scratch;
  uint32_t c = 0;
  if (want && elements[i] > 2ull * seg) {
    const bool pr = d->paired ? d->paired[i] != 0 : d->paired_all != 0;
    c = mate_segments(mate_kmers(d->off[2 * i + 1] - d->off[2 * i], k), seg);
    if (pr) c += mate_segments(mate_kmers(d->off[2 * i + 2] - d->off[2 * i + 1], k), seg);
    list[atomicAdd(&count[1], 1u)] = (uint32_t) i;
  }
  cnt[i] = c;
}

// Step 3: one wave per split fragment lays out its segments.  Per mate, the
// lanes move the evenly spaced cut points forward past ambiguous bases (the k
// bases before a cut must hold none), then lane 0 walks them from the last to
// the first, keeping each at or below the next, and writes each segment's
// range and reserved run space.
__device__ __forceinline__ void seg_plan_body(const char *seq, const uint64_t *off, const uint8_t *paired,
                                              uint32_t paired_all, uint32_t k, uint32_t seg, uint32_t pct,
                                              const uint32_t *list, const uint64_t *seg_start, uint32_t *count,
                                              uint64_t seg_total_at, const SegBatch &sg) {
  const uint32_t nf = count[1];
  const uint32_t lane = threadIdx.x;
  if (blockIdx.x == 0 && lane == 0) count[0] = (uint32_t) seg_start[seg_total_at];
  for (uint32_t i = blockIdx.x; i < nf; i += gridDim.x) {
    const uint32_t f = list[i];
    const bool pr = paired ? paired[f] != 0 : paired_all != 0;
    uint64_t s = seg_start[f];
    for (uint32_t mate = 0; mate < (pr ? 2u : 1u); mate++) {
      const char *m = seq + off[2 * f + mate];
      const uint32_t L = (uint32_t) (off[2 * f + mate + 1] - off[2 * f + mate]);
      const uint32_t e = mate_kmers(L, k), segs = mate_segments(e, seg);
      for (uint32_t j = lane; j < segs; j += blockDim.x) {
        uint32_t a = e ? (k - 1) + (uint32_t) (((uint64_t) j * e) / segs) : 0u;
        if (j > 0) {  // a >= k: past the last ambiguous base among [a - k, a), until none
          while (a < L) {
            uint32_t amb = 0xFFFFFFFFu;
            for (uint32_t r = a; r > a - k; r--)
              if (dna_code((uint8_t) m[r - 1]) == 0xFFu) {
                amb = r - 1;
                break;
              }
            if (amb == 0xFFFFFFFFu) break;
            a = amb + k + 1;
          }
          if (a > L) a = L;
        }
        sg.a[s + j] = a;
      }
      __syncthreads();
      if (lane == 0) {
        uint32_t next = e ? L : 0u;  // k-mers end at positions [k - 1, L)
        for (uint32_t j = segs; j-- > 0;) {
          uint32_t a = sg.a[s + j];
          if (a > next) a = next;
          sg.frag[s + j] = f;
          sg.mate[s + j] = mate;
          sg.a[s + j] = a;
          sg.b[s + j] = next;
          sg.cap[s + j] = run_space(next > a ? next - a : 0u, pct);
          next = a;
        }
      }
      __syncthreads();
      s += segs;
    }
  }
}

__global__ __launch_bounds__(64) void k_seg_plan(const char *seq, const uint64_t *off, const uint8_t *paired,
                                                 uint32_t paired_all, uint32_t k, uint32_t seg, uint32_t pct,
                                                 const uint32_t *list, const uint64_t *seg_start, uint32_t *count,
                                                 uint64_t seg_total_at, SegBatch sg) {
  seg_plan_body(seq, off, paired, paired_all, k, seg, pct, list, seg_start, count, seg_total_at, sg);
}

// Segments of split fragments (nucleotide): one wave per segment, each a
// fresh scan of its stretch of one mate (see SegBatch).
template <bool EXTRAS>
__device__ __forceinline__ void segment_body(const DbView &db, const ClassifyParams &P, const FragBatch &in,
                                             const FragOut &out, const SegBatch &sg) {
  __shared__ WaveLds<EXTRAS> lds;
  const int lane = threadIdx.x;
  const uint32_t k = db.k;
  const uint32_t ns = uniform32(sg.count[0]);
  FragOut so = out;  // the segment's runs and records go to the scratch
  so.runs = sg.runs;
  so.recs = out.recs ? sg.recs : nullptr;
  so.lookups = nullptr;
  so.lines = nullptr;
  for (uint32_t s = blockIdx.x; s < ns; s += gridDim.x) {
    const uint32_t a = uniform32(sg.a[s]), b = uniform32(sg.b[s]);
    uint32_t nr = 0, nrec = 0, groups = 0, status = 0;
    if (b > a) {
      const uint64_t f = uniform32(sg.frag[s]);
      const uint32_t mate = uniform32(sg.mate[s]);
      const char *m = in.seq + uniform64(in.off[2 * f + mate]);
      const bool first = a + 1 == k;  // the mate's first k-mer: no scan before it
      const uint32_t s0 = first ? 0u : a - k, first_emit = first ? 0u : k;
      const uint64_t base = uniform64(sg.off[s]), cap = uniform64(sg.cap[s]);
      if (lane == 0) lds.bad = 0;
      __syncthreads();
      RunState st;
      NullTable nt;
      if (db.k - db.l == 4 && db.revcom_version != 0)
        scan_stream<true, EXTRAS, 5>(db, P, so, base, cap, m + s0, b - s0, nullptr, 0, st, nt, lds, &lds.bad,
                                     lane, first_emit);
      else
        scan_stream<true, EXTRAS, 0>(db, P, so, base, cap, m + s0, b - s0, nullptr, 0, st, nt, lds, &lds.bad,
                                     lane, first_emit);
      if (st.run_count > 0) {
        if (lane == 0) close_run(so.runs + 2 * base, cap, st.n_runs, st.run_code, st.run_count, nt, &lds.bad);
        st.n_runs++;
      }
      __syncthreads();
      nr = st.n_runs;
      nrec = st.nrec;
      groups = st.groups;
      status = (uint32_t) uniform32((uint32_t) lds.bad);
    }
    if (lane == 0) {
      sg.n_runs[s] = nr;
      sg.nrec[s] = nrec;
      sg.groups[s] = groups;
      sg.status[s] = status;
    }
    __syncthreads();  // the next segment reuses the LDS
  }
}

// Joins each split fragment's segments in order into the fragment's runs
// and records, rebuilds its hit table from the joined runs, and resolves it.
template <bool EXTRAS>
__device__ __forceinline__ void stitch_body(const DbView &db, const ClassifyParams &P, const FragBatch &in,
                                            const FragOut &out, const SegBatch &sg) {
  __shared__ WaveLds<EXTRAS> lds;
  LdsTable tab{lds.tkey, lds.tcnt, P.tab_cap};
  const int lane = threadIdx.x;
  const uint32_t nf = uniform32(sg.count[1]);
  for (uint32_t i = blockIdx.x; i < nf; i += gridDim.x) {
    const uint64_t f = uniform32(sg.split_list[i]);
    const bool paired = uniform32(in.paired ? in.paired[f] : in.paired_all) != 0;
    const uint64_t base = uniform64(out.run_off[f]), cap = uniform64(out.run_cap[f]);
    uint32_t *const runs_at = out.runs + 2 * base;
    tab.clear(lane);
    if (lane == 0) lds.bad = 0;
    __syncthreads();
    RunState st;
    uint32_t status = 0;
    const uint64_t s_lo = uniform64(sg.seg_start[f]), s_hi = s_lo + uniform32(sg.seg_cnt[f]);
    for (uint64_t s = s_lo; s < s_hi; s++) {
      const uint32_t mate = uniform32(sg.mate[s]);
      if (mate == 1 && (s == s_lo || uniform32(sg.mate[s - 1]) == 0))
        push_marker(out, base, cap, st, kCodeMate, tab, &lds.bad, lane);
      const uint32_t a = uniform32(sg.a[s]), b = uniform32(sg.b[s]);
      if (b <= a) continue;
      const uint32_t seg_status = uniform32(sg.status[s]);
      status |= seg_status;
      // a segment that outgrew its space is incomplete: the fragment is
      // classified again whole (status bit 1), so nothing of it is joined
      if (seg_status & 2u) continue;
      st.elems += b - a;
      st.groups += uniform32(sg.groups[s]);
      const uint64_t so = uniform64(sg.off[s]);
      // HLL records, in order
      const uint32_t nrec = uniform32(sg.nrec[s]);
      if (out.recs) {
        for (uint32_t j = (uint32_t) lane; j < nrec; j += kWave) {
          const uint64_t r = st.nrec + j;
          if (r < cap) {
            out.recs[2 * (base + r)] = sg.recs[2 * (so + j)];
            out.recs[2 * (base + r) + 1] = sg.recs[2 * (so + j) + 1];
          } else {
            atomicOr(&lds.bad, 2);
          }
        }
      }
      st.nrec += nrec;
      // runs: the first may continue the open run; the last stays open
      const uint32_t r = uniform32(sg.n_runs[s]);
      if (r == 0) continue;
      const uint32_t *R = sg.runs + 2 * so;
      const uint32_t c0 = uniform32(R[0]);
      uint32_t n0 = uniform32(R[1]);
      if (st.run_count > 0 && st.run_code == c0) {
        n0 += st.run_count;
      } else if (st.run_count > 0) {
        if (lane == 0) close_run(runs_at, cap, st.n_runs, st.run_code, st.run_count, tab, &lds.bad);
        st.n_runs++;
      }
      if (r == 1) {
        st.run_code = c0;
        st.run_count = n0;
        continue;
      }
      if (lane == 0) close_run(runs_at, cap, st.n_runs, c0, n0, tab, &lds.bad);
      for (uint32_t j = 1u + (uint32_t) lane; j + 1 < r; j += kWave)
        close_run(runs_at, cap, st.n_runs + j, R[2 * j], R[2 * j + 1], tab, &lds.bad);
      st.n_runs += r - 1;
      st.run_code = uniform32(R[2 * (r - 1)]);
      st.run_count = uniform32(R[2 * (r - 1) + 1]);
    }
    if (st.run_count > 0) {
      if (lane == 0) close_run(runs_at, cap, st.n_runs, st.run_code, st.run_count, tab, &lds.bad);
      st.n_runs++;
    }
    if (lane == 0 && status) atomicOr(&lds.bad, (int) status);
    __syncthreads();
    resolve_fragment<EXTRAS>(db, P, out, f, st, tab, paired, false, &lds.bad, lane);
    __syncthreads();  // the next fragment reuses the LDS
  }
}

__global__ __launch_bounds__(64) void k_classify_segments_dna(DbView db, ClassifyParams P, FragBatch in,
                                                              FragOut out, SegBatch sg) {
  segment_body<false>(db, P, in, out, sg);
}
__global__ __launch_bounds__(64) void k_classify_segments_dna_x(DbView db, ClassifyParams P, FragBatch in,
                                                                FragOut out, SegBatch sg) {
  segment_body<true>(db, P, in, out, sg);
}
__global__ __launch_bounds__(64) void k_stitch_dna(DbView db, ClassifyParams P, FragBatch in, FragOut out,
                                                   SegBatch sg) {
  stitch_body<false>(db, P, in, out, sg);
}
__global__ __launch_bounds__(64) void k_stitch_dna_x(DbView db, ClassifyParams P, FragBatch in, FragOut out,
                                                     SegBatch sg) {
  stitch_body<true>(db, P, in, out, sg);
}

// Kernel variants: _dna (nucleotide, no per-hit extras: the hot path),
// _dna_x (nucleotide with HLL records / counts / probe statistics / quick
// mode), _prot (translated search, all of them available).
// WAVES: resident waves per SIMD the variant is built for (its LDS allows
// it: 32 one-wave workgroups per CU for _dna, 28 for the others).
// The batch's arrays and sizes of a graph-replayed pipeline, read from its
// descriptor into the kernel's batch view.
__device__ __forceinline__ FragBatch with_desc(FragBatch in, const PipeDesc *d) {
  in.seq = d->seq;
  in.off = d->off;
  in.paired = d->paired;
  in.paired_all = d->paired_all;
  in.n = d->n;
  return in;
}

#define K2G_CLASSIFY_VARIANT(SUFFIX, DNA, EXTRAS, WAVES)                                       \
  __global__ __launch_bounds__(64, WAVES) void k_classify_plain##SUFFIX(                      \
      DbView db, ClassifyParams P, FragBatch in, FragOut out) {                                \
    plain_body<DNA, EXTRAS>(db, P, in, out);                                                   \
  }                                                                                            \
  __global__ __launch_bounds__(64, WAVES) void k_classify_persistent##SUFFIX(                 \
      DbView db, ClassifyParams P, FragBatch in, FragOut out, unsigned long long *cursor) {    \
    persistent_body<DNA, EXTRAS>(db, P, in, out, cursor);                                      \
  }                                                                                            \
  __global__ __launch_bounds__(64, WAVES) void k_classify_pipe##SUFFIX(                       \
      DbView db, ClassifyParams P, const PipeDesc *d, FragBatch in, FragOut out,               \
      unsigned long long *cursor) {                                                            \
    persistent_body<DNA, EXTRAS>(db, P, with_desc(in, d), out, cursor);                        \
  }
K2G_CLASSIFY_VARIANT(_dna, true, false, 8)
K2G_CLASSIFY_VARIANT(_dna_x, true, true, 7)
K2G_CLASSIFY_VARIANT(_prot, false, true, 7)
#undef K2G_CLASSIFY_VARIANT

// The long-fragment split inside a graph-replayed pipeline: the same kernels
// with the batch read from its descriptor.
__global__ __launch_bounds__(64) void k_seg_plan_pipe(const PipeDesc *d, uint32_t k, uint32_t seg, uint32_t pct,
                                                      const uint32_t *list, const uint64_t *seg_start,
                                                      uint32_t *count, SegBatch sg) {
  seg_plan_body(d->seq, d->off, d->paired, d->paired_all, k, seg, pct, list, seg_start, count, d->n, sg);
}
__global__ __launch_bounds__(64) void k_classify_segments_pipe_dna(DbView db, ClassifyParams P, const PipeDesc *d,
                                                                   FragBatch in, FragOut out, SegBatch sg) {
  segment_body<false>(db, P, with_desc(in, d), out, sg);
}
__global__ __launch_bounds__(64) void k_classify_segments_pipe_dna_x(DbView db, ClassifyParams P,
                                                                     const PipeDesc *d, FragBatch in, FragOut out,
                                                                     SegBatch sg) {
  segment_body<true>(db, P, with_desc(in, d), out, sg);
}
__global__ __launch_bounds__(64) void k_stitch_pipe_dna(DbView db, ClassifyParams P, const PipeDesc *d, FragBatch in,
                                                        FragOut out, SegBatch sg) {
  stitch_body<false>(db, P, with_desc(in, d), out, sg);
}
__global__ __launch_bounds__(64) void k_stitch_pipe_dna_x(DbView db, ClassifyParams P, const PipeDesc *d,
                                                          FragBatch in, FragOut out, SegBatch sg) {
  stitch_body<true>(db, P, with_desc(in, d), out, sg);
}

__global__ __launch_bounds__(64) void k_classify_dense_pipe_dna(DbView db, ClassifyParams P, const PipeDesc *d,
                                                                 FragBatch in, FragOut out, const uint32_t *redo,
                                                                 const uint32_t *n_redo, DenseScratch scratch) {
  dense_body<true>(db, P, with_desc(in, d), out, redo, n_redo, scratch);
}
__global__ __launch_bounds__(64) void k_classify_dense_pipe_prot(DbView db, ClassifyParams P, const PipeDesc *d,
                                                                  FragBatch in, FragOut out, const uint32_t *redo,
                                                                  const uint32_t *n_redo, DenseScratch scratch) {
  dense_body<false>(db, P, with_desc(in, d), out, redo, n_redo, scratch);
}
__global__ __launch_bounds__(64) void k_classify_dense_dna(DbView db, ClassifyParams P, FragBatch in,
                                                            FragOut out, const uint32_t *redo,
                                                            const uint32_t *n_redo, DenseScratch scratch) {
  dense_body<true>(db, P, in, out, redo, n_redo, scratch);
}
__global__ __launch_bounds__(64) void k_classify_dense_prot(DbView db, ClassifyParams P,
                                                             FragBatch in, FragOut out,
                                                             const uint32_t *redo, const uint32_t *n_redo,
                                                             DenseScratch scratch) {
  dense_body<false>(db, P, in, out, redo, n_redo, scratch);
}

// KeyValueStore::GetBatch on the device (one lane per key).
__global__ __launch_bounds__(256) void k_get_batch(DbView db, const uint64_t *keys, uint32_t *vals,
                                                    uint64_t n) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vals[i] = table_get(db, keys[i]);
}

// Largest cell value (used to check that every stored taxid is a node).
// out[0]: the largest stored value; out[1]: 1 if some cell is empty.
__global__ __launch_bounds__(256) void k_max_value(DbView db, unsigned int *out) {
  const uint32_t vmask = (1u << db.value_bits) - 1u;
  uint32_t best = 0;
  bool empty = false;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < db.capacity;
       i += (uint64_t) gridDim.x * blockDim.x) {
    const uint32_t v = db.cells[i] & vmask;
    best = v > best ? v : best;
    empty |= v == 0u;
  }
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = (uint32_t) __shfl_xor((unsigned int) best, d);
    best = o > best ? o : best;
  }
  const bool any_empty = __ballot(empty) != 0;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicMax(out, best);
    if (any_empty) atomicOr(out + 1, 1u);
  }
}

// Copies the first cells into the padding after the table (capacity >= 64).
__global__ void k_pad_copy(uint32_t *cells, uint64_t capacity, uint32_t n) {
  const uint32_t t = threadIdx.x;
  if (t < n) cells[capacity + t] = cells[t % capacity];
}

}  // namespace k2g

namespace k2g {

// Reduced amino-acid alphabet of the protein scanner (src/mmscanner.cc:
// 41-95), 0xFF for every other byte.
__device__ __forceinline__ uint32_t prot_code(uint8_t ch) {
  const uint8_t c = (ch >= 'a' && ch <= 'z') ? (uint8_t) (ch - 32) : ch;
  switch (c) {
    case '*': case 'U': case 'O': return 0;
    case 'A': return 1;
    case 'N': case 'Q': case 'S': return 2;
    case 'C': return 3;
    case 'D': case 'E': return 4;
    case 'F': return 5;
    case 'G': return 6;
    case 'H': return 7;
    case 'I': case 'L': return 8;
    case 'K': return 9;
    case 'P': return 10;
    case 'R': return 11;
    case 'M': case 'V': return 12;
    case 'T': return 13;
    case 'W': return 14;
    case 'Y': return 15;
  }
  return 0xFFu;
}

// CompareAndSet + SetMinimizerLCA as one lock-free loop: the first cell of
// the linear probe that is empty or carries the same compacted key receives
// LCA(old value, taxid) (src/build_db.h:186-194, src/compact_hash.h:416-462).
__device__ void table_insert(const BuildView &b, uint64_t key, uint32_t taxid) {
  const uint64_t hc = fmix64(key);
  const uint32_t ck = (uint32_t) (hc >> (64 - b.key_bits));
  const uint32_t vb = b.value_bits;
  const uint32_t vmask = (1u << vb) - 1u;
  uint64_t idx = mod_capacity(hc, b.capacity, b.barrett);
  uint64_t left = b.capacity;
  uint32_t c = load_agent(&b.cells[idx]);
  for (;;) {
    const uint32_t v = c & vmask;
    if (v == 0 || (c >> vb) == ck) {
      const uint32_t nv = v == 0 ? taxid : lca(b.parent, v, taxid);
      if (v != 0 && nv == v) return;
      const uint32_t want = (ck << vb) | nv;
      const uint32_t old = atomicCAS(&b.cells[idx], c, want);
      if (old == c) {
        if (v == 0) atomicAdd(b.size, 1ULL);
        return;
      }
      c = old;  // lost a race: look at the same cell again
      continue;
    }
    if (--left == 0) {
      atomicOr(b.error, 1u);
      return;
    }
    if (++idx == b.capacity) idx = 0;
    c = load_agent(&b.cells[idx]);
  }
}

// Scans one build window and inserts its non-ambiguous minimizers.
__device__ void build_window(const BuildView &b, const char *m, uint64_t n, uint32_t first_e,
                             uint32_t taxid, uint64_t *ring, int lane) {
  const bool dna = b.dna != 0;
  const uint32_t k = b.k, l = b.l;
  const uint64_t lbits = (1ULL << l) - 1ULL;
  int64_t last_amb = -1;
  uint64_t last_min = ~0ULL;
  uint64_t pp0 = 0, pp1 = 0, pp2 = 0, pp3 = 0;
  for (uint64_t c = 0; c < n; c += kWave) {
    const uint64_t p = c + lane;
    const bool inb = p < n;
    uint32_t code = 0xFFu;
    if (inb) code = dna ? dna_code((uint8_t) m[p]) : prot_code((uint8_t) m[p]);
    const bool amb = inb && code == 0xFFu;
    const bool good = inb && !amb;
    const uint64_t P0 = __ballot(good && (code & 1u));
    const uint64_t P1 = __ballot(good && (code & 2u));
    const uint64_t P2 = dna ? 0 : __ballot(good && (code & 4u));
    const uint64_t P3 = dna ? 0 : __ballot(good && (code & 8u));
    const uint64_t A = __ballot(amb);
    const uint64_t mA = A & lanemask_le(lane);
    const int64_t lastN = mA ? (int64_t) (c + top_lane(mA)) : last_amb;
    const bool lvalid = inb && p + 1 >= l && lastN < (int64_t) (p + 1 - l);
    uint64_t cand = kNoCand;
    if (lvalid) {
      const uint32_t s = (uint32_t) lane + 65u - l;
      const uint64_t r0 = __builtin_bitreverse64(window128(pp0, P0, s) & lbits) >> (64 - l);
      const uint64_t r1 = __builtin_bitreverse64(window128(pp1, P1, s) & lbits) >> (64 - l);
      if (dna) {
        cand = canonical_lmer((spread2(r1) << 1) | spread2(r0), l, b.revcom_version);
      } else {
        const uint64_t r2 = __builtin_bitreverse64(window128(pp2, P2, s) & lbits) >> (64 - l);
        const uint64_t r3 = __builtin_bitreverse64(window128(pp3, P3, s) & lbits) >> (64 - l);
        cand = (spread4(r3) << 3) | (spread4(r2) << 2) | (spread4(r1) << 1) | spread4(r0);
      }
      if (b.spaced_mask) cand &= b.spaced_mask;
      cand ^= b.toggle_mask;
    }
    ring[p & 127] = cand;
    __syncthreads();
    const bool kvalid = inb && p >= first_e;
    bool nonamb = false;
    uint64_t mn = ~0ULL;
    if (kvalid) {
      nonamb = (k == l) ? lvalid : (lastN < (int64_t) (p + 2) - (int64_t) k);
      if (nonamb) {
        for (uint32_t j = 0; j <= k - l && j <= p; j++) {
          const uint64_t v = ring[(p - j) & 127];
          mn = v < mn ? v : mn;
        }
        nonamb = mn != kNoCand;
        mn ^= b.toggle_mask;
      }
    }
    __syncthreads();
    const uint64_t NA = __ballot(nonamb);
    const uint64_t prevm = NA & lanemask_lt(lane);
    const uint64_t before = shfl64(mn, prevm ? top_lane(prevm) : lane);
    const uint64_t prev_min = prevm ? before : last_min;
    // consecutive repeats insert the same (minimizer, taxid) pair again,
    // which leaves the cell unchanged; insert each run once
    const bool ins = nonamb && mn != prev_min;
    if (b.emit_key) {
      // verification: the pairs this window inserts, one hash partition at a
      // time, appended with one atomic per wave
      const uint64_t hm = fmix64(mn);
      const bool e = ins && !(b.min_clear && hm < b.min_clear) &&
                     (b.part_bits == 0 || (uint32_t) (hm >> (64 - b.part_bits)) == b.part);
      const uint64_t EM = __ballot(e);
      if (EM) {
        const int fl = __ffsll((long long) EM) - 1;
        unsigned long long at0 = 0;
        if (lane == fl) at0 = atomicAdd(b.n_emit, (unsigned long long) __popcll(EM));
        const uint64_t at = readlane64((uint64_t) at0, fl) + (uint64_t) __popcll(EM & lanemask_lt(lane));
        if (e && at < b.emit_cap) {
          b.emit_key[at] = mn;
          b.emit_tax[at] = taxid;
        }
      }
    } else if (ins) {
      if (b.sample) {
        if ((fmix64(mn) & 1023u) < b.sample_ranges) {
          const unsigned long long at = atomicAdd(b.n_sample, 1ULL);
          if (at < b.sample_cap) b.sample[at] = mn;
        }
      } else if (!(b.min_clear && fmix64(mn) < b.min_clear)) {
        table_insert(b, mn, taxid);
      }
    }
    if (A) last_amb = (int64_t) (c + top_lane(A));
    if (NA) last_min = uniform64(shfl64(mn, top_lane(NA)));
    pp0 = P0;
    pp1 = P1;
    pp2 = P2;
    pp3 = P3;
  }
}

// ---- verification of a built table against the pairs its build inserted
// (sorted by minimizer, then the minimizers sorted by the cell they reach)

// Sum of a per-lane count over the wave, added to *dst by one lane.
__device__ __forceinline__ void wave_add(unsigned long long *dst, uint64_t v, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += shfl64(v, lane ^ d);
  if (lane == 0 && v) atomicAdd(dst, (unsigned long long) v);
}

// One thread per sorted pair; the first pair of each minimizer's run takes
// the LCA of the run's taxa (the value a build inserting only this
// minimizer leaves, src/build_db.h:186-194), looks the minimizer up and
// checks that the value found is an ancestor-or-self of every one of its
// taxa.  Found minimizers are appended as (cell, own LCA) for the per-cell
// pass.
__global__ __launch_bounds__(256) void k_verify_keys(DbView db, const uint64_t *key, const uint32_t *tax,
                                                      uint64_t m, uint64_t *cell_out, uint32_t *lca_out,
                                                      unsigned long long *n_out, unsigned long long *cnt) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool head = i < m && (i == 0 || key[i] != key[i - 1]);
  uint64_t occ = 0, where = db.capacity;
  uint32_t own = 0, val = 0;
  bool cover = true;
  if (head) {
    const uint64_t k = key[i];
    own = tax[i];
    uint64_t j = i + 1;
    for (; j < m && key[j] == k; j++) own = lca(db.parent, own, tax[j]);
    occ = j - i;
    val = table_find(db, k, &where);
    if (val)
      for (uint64_t q = i; q < j; q++) cover = cover && is_ancestor(db.parent, val, tax[q]);
  }
  const bool found = head && val != 0;
  const uint64_t F = __ballot(found);
  if (F) {
    const int fl = __ffsll((long long) F) - 1;
    unsigned long long at0 = 0;
    if (lane == fl) at0 = atomicAdd(n_out, (unsigned long long) __popcll(F));
    const uint64_t at = readlane64((uint64_t) at0, fl) + (uint64_t) __popcll(F & lanemask_lt(lane));
    if (found) {
      cell_out[at] = where;
      lca_out[at] = own;
    }
  }
  wave_add(&cnt[kVOcc], occ, lane);
  const uint64_t H = __ballot(head);
  const uint64_t NC = __ballot(found && !cover);
  const uint64_t D = __ballot(found && val != own);
  if (lane == 0) {
    if (H) atomicAdd(&cnt[kVKeys], (unsigned long long) __popcll(H));
    if (H & ~F) atomicAdd(&cnt[kVMissing], (unsigned long long) __popcll(H & ~F));
    if (NC) atomicAdd(&cnt[kVNotCovering], (unsigned long long) __popcll(NC));
    if (D) atomicAdd(&cnt[kVDiffer], (unsigned long long) __popcll(D));
  }
}

// One thread per (cell, own LCA) pair sorted by cell: the first of each
// cell's run checks that the cell holds the LCA of every minimizer ending
// there (a lone minimizer: its own LCA).
__global__ __launch_bounds__(256) void k_verify_cells(DbView db, const uint64_t *cell, const uint32_t *own,
                                                       uint64_t m, unsigned long long *cnt) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool head = i < m && (i == 0 || cell[i] != cell[i - 1]);
  uint64_t len = 0;
  bool bad = false;
  if (head) {
    const uint64_t c = cell[i];
    uint32_t a = own[i];
    uint64_t j = i + 1;
    for (; j < m && cell[j] == c; j++) a = lca(db.parent, a, own[j]);
    len = j - i;
    const uint32_t vmask = (1u << db.value_bits) - 1u;
    bad = (db.cells[c] & vmask) != a;
  }
  const bool shared = head && len > 1;
  wave_add(&cnt[kVSharedKeys], shared ? len : 0, lane);
  const uint64_t H = __ballot(head), S = __ballot(shared), B = __ballot(bad);
  if (lane == 0) {
    if (H) atomicAdd(&cnt[kVCells], (unsigned long long) __popcll(H));
    if (S) atomicAdd(&cnt[kVSharedCells], (unsigned long long) __popcll(S));
    if (B) atomicAdd(&cnt[kVUnexplained], (unsigned long long) __popcll(B));
  }
}

__global__ __launch_bounds__(64) void k_build(BuildView b, const char *seq, const SubBlock *items,
                                               uint64_t n_items) {
  __shared__ uint64_t ring[128];
  for (uint64_t i = blockIdx.x; i < n_items; i += gridDim.x) {
    const SubBlock it = items[i];
    build_window(b, seq + uniform64(it.start), uniform32(it.len), uniform32(it.first_e),
                 uniform32(it.taxid), ring, threadIdx.x);
    __syncthreads();
  }
}

}  // namespace k2g
