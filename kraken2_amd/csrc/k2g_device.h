// Device-side building blocks of the MI355X classify core: the compact-hash
// probe, taxonomy walks, and the DB view every kernel receives by value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace k2g {

constexpr int kWave = 64;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Special element codes in a hitlist run (never valid internal taxids, which
// are < 2^31 because value_bits <= 31).
constexpr uint32_t kCodeMate = 0xFFFFFFFFu;   // "|:|"
constexpr uint32_t kCodeFrame = 0xFFFFFFFEu;  // "-:-"
constexpr uint32_t kCodeAmbig = 0xFFFFFFFDu;  // "A:n"

constexpr uint64_t kNoCand = ~0ULL;  // ring sentinel: no valid l-mer

// Everything a kernel needs to read the resident DB.
struct DbView {
  // capacity cells and at least 4 more, padded to a multiple of 64; with
  // lean_probe the padding holds copies of the first cells, so a 4-cell read
  // across the end sees the cells a wrapping linear probe visits next
  const uint32_t *cells;
  uint64_t capacity;
  uint64_t barrett;        // floor(2^64 / capacity) (0 when capacity == 1)
  uint32_t key_bits, value_bits;
  const uint32_t *parent;  // internal parent id per node (node 0 -> 0)
  // per node v: span[2v] = its preorder index, span[2v+1] = that plus its
  // subtree size, so a is an ancestor-or-self of b iff b's index lies in a's
  // span (node 0: an empty span)
  const uint32_t *span;
  uint32_t node_count;
  // scanner parameters (opts.k2d)
  uint32_t k, l;
  uint32_t dna;            // 1 = nucleotide DB, 0 = protein (translated search)
  uint32_t revcom_version;
  uint64_t spaced_mask;
  uint64_t toggle_mask;    // already masked to the l-mer width
  uint64_t lmer_mask;
  uint64_t min_hash;       // minimum_acceptable_hash_value
  // 1 when capacity >= 64, the padding copies the first cells and some cell
  // is empty: every probe then ends at an empty or matching cell before it
  // wraps all the way round, so it needs no count of cells left to visit
  uint32_t lean_probe;
};

// MurmurHash3 64-bit finalizer (src/kv_store.h:59-67)
__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Exact x % d with a precomputed m = floor(2^64 / d): q = mulhi(x, m) is
// floor(x / d) or one less, so at most one correction step remains.
__device__ __forceinline__ uint64_t mod_capacity(uint64_t x, uint64_t d, uint64_t m) {
  if (d == 1) return 0;
  const uint64_t q = __umul64hi(x, m);
  const uint64_t r = x - q * d;
  return r >= d ? r - d : r;
}

// CompactHashTable::Get with linear probing (src/compact_hash.h:354-374,
// 493-500).  The probe walks 16-byte aligned groups of four cells so one
// load covers up to four probe steps.
__device__ __forceinline__ uint32_t table_get(const DbView &db, uint64_t key) {
  const uint64_t hc = fmix64(key);
  const uint32_t ck = (uint32_t) (hc >> (64 - db.key_bits));
  const uint32_t vb = db.value_bits;
  const uint32_t vmask = (1u << vb) - 1u;
  uint64_t idx = mod_capacity(hc, db.capacity, db.barrett);
  uint64_t left = db.capacity;  // cells still to examine before a full wrap
  for (;;) {
    const uint64_t base = idx & ~3ULL;
    const u32x4 v = *(const u32x4 *) (db.cells + base);
    const uint32_t c4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if ((uint64_t) j < idx - base) continue;
      const uint32_t c = c4[j];
      if ((c & vmask) == 0) return 0;
      if ((c >> vb) == ck) return c & vmask;
      if (--left == 0) return 0;
      if (base + j + 1 == db.capacity) break;  // wrap to cell 0
    }
    idx = base + 4;
    if (idx >= db.capacity) idx = 0;
  }
}

// table_get that also reports the cell the probe ends at: *where = the
// index of the cell holding key's compacted key, or capacity when the probe
// ends at an empty cell or visits every cell (then the value is 0).
__device__ __forceinline__ uint32_t table_find(const DbView &db, uint64_t key, uint64_t *where) {
  const uint64_t hc = fmix64(key);
  const uint32_t ck = (uint32_t) (hc >> (64 - db.key_bits));
  const uint32_t vb = db.value_bits;
  const uint32_t vmask = (1u << vb) - 1u;
  uint64_t idx = mod_capacity(hc, db.capacity, db.barrett);
  for (uint64_t left = db.capacity; left; left--) {
    const uint32_t c = db.cells[idx];
    if ((c & vmask) == 0) break;
    if ((c >> vb) == ck) {
      *where = idx;
      return c & vmask;
    }
    if (++idx == db.capacity) idx = 0;
  }
  *where = db.capacity;
  return 0;
}

// Distinct 128-byte lines of the cell array that table_get(key) reads, from
// the home cell through the terminating cell (measurement only).
__device__ __forceinline__ uint32_t table_lines(const DbView &db, uint64_t key) {
  const uint64_t hc = fmix64(key);
  const uint32_t ck = (uint32_t) (hc >> (64 - db.key_bits));
  const uint32_t vb = db.value_bits;
  const uint32_t vmask = (1u << vb) - 1u;
  uint64_t idx = mod_capacity(hc, db.capacity, db.barrett);
  uint64_t left = db.capacity, line = idx >> 5;
  uint32_t lines = 1;
  for (;;) {
    const uint32_t c = db.cells[idx];
    if ((c & vmask) == 0 || (c >> vb) == ck) return lines;
    if (--left == 0) return lines;
    if (++idx == db.capacity) idx = 0;
    if ((idx >> 5) != line) {
      line = idx >> 5;
      lines++;
    }
  }
}

// Preorder spans of the taxonomy (see DbView::span) from a parent array in
// which every parent id is below its child's (host code, O(n)).
inline void taxonomy_spans(const uint32_t *parent, size_t n, uint32_t *span) {
  std::vector<uint32_t> size(n, 1), next(n, 0);
  for (size_t i = n; i-- > 1;)
    if (parent[i]) size[parent[i]] += size[i];
  uint32_t cursor = 0;
  if (n) span[0] = span[1] = 0;
  for (size_t i = 1; i < n; i++) {
    uint32_t tin;
    if (parent[i] == 0) {
      tin = cursor;
      cursor += size[i];
    } else {
      tin = next[parent[i]];
      next[parent[i]] += size[i];
    }
    next[i] = tin + 1;
    span[2 * i] = tin;
    span[2 * i + 1] = tin + size[i];
  }
}

// src/taxonomy.cc:251-258 (ids are BFS ordered: parent < child)
__device__ __forceinline__ bool is_ancestor(const uint32_t *parent, uint32_t a, uint32_t b) {
  if (a == 0 || b == 0) return false;
  while (b > a) b = parent[b];
  return b == a;
}

// src/taxonomy.cc:262-272
__device__ __forceinline__ uint32_t lca(const uint32_t *parent, uint32_t a, uint32_t b) {
  if (a == 0 || b == 0) return a ? a : b;
  while (a != b) {
    if (a > b)
      a = parent[a];
    else
      b = parent[b];
  }
  return a;
}

// Wave-uniform values must live in scalar registers: a loop or branch whose
// condition the compiler sees as per-lane (e.g. a shuffled value) is
// structurized for divergent lanes, and around a lane-0-only side effect
// that structure can leave lanes looping without lane 0.  readfirstlane makes
// the uniformity explicit.
__device__ __forceinline__ uint32_t uniform32(uint32_t v) {
  return (uint32_t) __builtin_amdgcn_readfirstlane((int) v);
}
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) v);
  const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) (v >> 32));
  return ((uint64_t) hi << 32) | lo;
}

// A lane's value read by every lane (lane wave-uniform): no LDS round trip.
__device__ __forceinline__ uint32_t readlane32(uint32_t v, int lane) {
  return (uint32_t) __builtin_amdgcn_readlane((int) v, lane);
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) v, lane);
  const uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) (v >> 32), lane);
  return ((uint64_t) hi << 32) | lo;
}

// Lane i receives lane i-1's v and lane 0 receives `first` (a DPP wave
// shift: one move per 32 bits, no LDS).
__device__ __forceinline__ uint32_t wave_shr1_32(uint32_t v, uint32_t first) {
  return (uint32_t) __builtin_amdgcn_update_dpp((int) first, (int) v, 0x138 /* wave_shr:1 */, 0xf,
                                                0xf, false);
}
__device__ __forceinline__ uint64_t wave_shr1(uint64_t v, uint64_t first) {
  const uint32_t lo = (uint32_t) __builtin_amdgcn_update_dpp(
      (int) (uint32_t) first, (int) (uint32_t) v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t) __builtin_amdgcn_update_dpp(
      (int) (uint32_t) (first >> 32), (int) (uint32_t) (v >> 32), 0x138, 0xf, 0xf, false);
  return ((uint64_t) hi << 32) | lo;
}

// A wave-uniform value kept in a vector register: vector instructions read
// it directly, and it takes no scalar register (scalar registers spilled to
// vector lanes cost a readlane at every use).
template <class T>
__device__ __forceinline__ T in_vgpr(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
  return x;
}

// This lane's bit of a wave-uniform mask (the mask itself serves as the
// lane predicate: no vector instructions).
__device__ __forceinline__ bool lane_bit(uint64_t wave_mask) {
  return __builtin_amdgcn_inverse_ballot_w64(wave_mask);
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (1ULL << lane) - 1ULL; }
__device__ __forceinline__ uint64_t lanemask_le(int lane) { return ~0ULL >> (63 - lane); }
__device__ __forceinline__ int top_lane(uint64_t m) { return 63 - __clzll((long long) m); }

// Each bit of a wave-uniform x twice: bit i -> bits 2i and 2i+1 (one scalar
// instruction on the device).
__device__ __forceinline__ uint64_t replicate_bits(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t r;
  asm("s_bitreplicate_b64_b32 %0, %1" : "=s"(r) : "s"(x));
  return r;
#else
  uint64_t r = 0;
  for (int i = 0; i < 32; i++) r |= (uint64_t) ((x >> i) & 1u) * (3ULL << (2 * i));
  return r;
#endif
}

// Spread the low 32 bits so bit t lands on bit 2t.
__device__ __forceinline__ uint64_t spread2(uint64_t x) {
  x &= 0xFFFFFFFFULL;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
  x = (x | (x << 2)) & 0x3333333333333333ULL;
  x = (x | (x << 1)) & 0x5555555555555555ULL;
  return x;
}

// Spread the low 16 bits so bit t lands on bit 4t.
__device__ __forceinline__ uint64_t spread4(uint64_t x) {
  x &= 0xFFFFULL;
  x = (x | (x << 24)) & 0x000000FF000000FFULL;
  x = (x | (x << 12)) & 0x000F000F000F000FULL;
  x = (x | (x << 6)) & 0x0303030303030303ULL;
  x = (x | (x << 3)) & 0x1111111111111111ULL;
  return x;
}

// 2-bit groups of x in reverse order (src/mmscanner.cc:193-215 before the
// complement step).
__device__ __forceinline__ uint64_t reverse_pairs(uint64_t x) {
  uint64_t r = __builtin_bitreverse64(x);
  return ((r >> 1) & 0x5555555555555555ULL) | ((r & 0x5555555555555555ULL) << 1);
}

// Canonical l-mer (src/mmscanner.cc:193-220): min of the l-mer and its
// reverse complement in the reference's two revcom conventions.
__device__ __forceinline__ uint64_t canonical_lmer(uint64_t x, uint32_t l, uint32_t version) {
  const uint64_t mask = (l >= 32) ? ~0ULL : ((1ULL << (2 * l)) - 1ULL);
  const uint64_t r = ~reverse_pairs(x);
  const uint64_t rc = version == 0 ? (r & mask) : ((r >> (64 - 2 * l)) & mask);
  return x < rc ? x : rc;
}

// Bits [s, s+width) of the 128-bit value hi:lo (s in [1, 127]).
__device__ __forceinline__ uint64_t window128(uint64_t lo, uint64_t hi, uint32_t s) {
  if (s >= 64) return hi >> (s - 64);
  return (lo >> s) | (hi << (64 - s));
}

}  // namespace k2g
