// Device code of the multi-database merge (the reference's merge tool,
// src/merge.cc), shared by the library (k2g_merge.hip) and the CPU
// execution harness of the tests.
//
// One thread per read.  Per read: the two run-length hit lists are walked
// together and each stretch gets the LCA of the two taxa
// (merge_hit_lists, src/merge.cc:494-554); the read is called again from
// the merged list (resolve_tree, :369-421); the merged list is folded into
// the runs the reference prints (write_hit_list, :328-367).  A read's
// merged list is at most as long as its two input lists together, so read
// i owns the slots [a.off[i] + b.off[i], a.off[i+1] + b.off[i+1]) of every
// per-item array: no counting pass, no allocation per read.  The call
// compares every merged item against every other (the reference's pairwise
// score over distinct taxa; duplicates give the same scores and LCA ties),
// with ancestry from preorder spans.
#pragma once
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "k2g.h"
#include "k2g_device.h"

namespace k2g {

constexpr int32_t kValid = INT32_MAX - 3;  // larger values are "A", "|" and "-"

struct MergeTax {
  const uint32_t *parent;   // node -> parent (parent < child)
  const uint32_t *span;     // preorder spans: [2v] preorder index, [2v+1] end of subtree
  const int32_t *ext32;     // node -> external id as the merge tool's int holds it
  const uint64_t *ext64;    // node -> external id
  const uint64_t *key;      // sorted external ids of the external-to-node map
  const uint32_t *key_node; // node of each key
  uint32_t n_keys;
  const uint32_t *direct;   // node of every external id below n_direct (0 if none), or null
  uint64_t n_direct;
};

// Taxonomy::GetInternalID (src/taxonomy.h:81-85) of the int taxid widened
// as the reference widens it: 0 when unknown.
__device__ __forceinline__ uint32_t node_of(const MergeTax &t, int32_t v) {
  const uint64_t k = (uint64_t) (int64_t) v;
  if (t.direct) return k < t.n_direct ? t.direct[k] : 0;
  uint32_t lo = 0, hi = t.n_keys;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (t.key[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < t.n_keys && t.key[lo] == k ? t.key_node[lo] : 0;
}

// IsAAncestorOfB (src/taxonomy.cc:251-258): a is b or above it
__device__ __forceinline__ bool above(const MergeTax &t, uint32_t a, uint32_t b) {
  if (!a || !b) return false;
  const uint32_t pb = t.span[2 * b];
  return t.span[2 * a] <= pb && pb < t.span[2 * a + 1];
}

// uint32_t r = <floating value>: through a signed 64-bit integer, low 32
// bits kept (x86-64 code for the reference's conversion)
__device__ __forceinline__ uint32_t to_u32(float r) {
  if (!(r > -9.2e18f && r < 9.2e18f)) return 0;
  return (uint32_t) (int64_t) r;
}

__device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) {
  return (int32_t) ((uint32_t) a + (uint32_t) b);
}

// Characters int_to_string (src/merge.cc:251-326) prints for an int:
// decimal digits, or one marker character ("A", "|", "-"; '?' for negative
// values, which classify never writes and the reference prints from a stale
// buffer).
__device__ __forceinline__ uint32_t printed_len(int32_t v) {
  if (v > kValid || v < 0) return 1;
  uint32_t n = 1;
  for (uint32_t x = (uint32_t) v; x >= 10; x /= 10) n++;
  return n;
}

__device__ __forceinline__ char *print_int(char *o, int32_t v) {
  if (v > kValid || v < 0) {
    *o = v == INT32_MAX ? 'A' : v == INT32_MAX - 1 ? '|' : v == INT32_MAX - 2 ? '-' : '?';
    return o + 1;
  }
  const uint32_t n = printed_len(v);
  uint32_t x = (uint32_t) v;
  for (uint32_t i = n; i-- > 0;) {
    o[i] = (char) ('0' + x % 10);
    x /= 10;
  }
  return o + n;
}

__global__ void __launch_bounds__(256)
k_merge(MergeTax t, const int32_t *__restrict__ a_items, const uint64_t *__restrict__ a_off,
        const int32_t *__restrict__ b_items, const uint64_t *__restrict__ b_off, uint64_t n,
        float conf, int32_t *__restrict__ w_print, int32_t *__restrict__ w_count,
        uint32_t *__restrict__ w_node, uint32_t *__restrict__ w_dcount, int32_t *__restrict__ tokens,
        uint8_t *__restrict__ joined,
        uint64_t *__restrict__ n_tokens, uint64_t *__restrict__ text_len, int32_t *__restrict__ call) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint64_t a0 = a_off[r], a1 = a_off[r + 1], b0 = b_off[r], b1 = b_off[r + 1];
  const uint64_t s = a0 + b0;  // this read's slots
  int32_t *pr = w_print + s, *ct = w_count + s;
  // distinct nodes of the merged taxa and their summed counts: the
  // reference's hit_counts map (:505, :546-548) keyed by node
  uint32_t *dn = w_node + s, *dc = w_dcount + s;
  uint32_t d = 0;

  // ---- merge_hit_lists (src/merge.cc:494-554).  A second list shorter
  // than the first reads past its end in the reference; here a missing item
  // is taxid 0 with an unbounded count.
  uint32_t m = 0;
  uint64_t total = 0;
  if (a1 > a0) {
    uint64_t i1 = a0, i2 = b0;
    int32_t ta = a_items[2 * i1], ca = a_items[2 * i1 + 1];
    uint32_t na = ta > kValid ? 0 : node_of(t, ta);
    int32_t tb = 0, cb = INT32_MAX;
    uint32_t nb = 0;
    if (i2 < b1) {
      tb = b_items[2 * i2];
      cb = b_items[2 * i2 + 1];
      nb = node_of(t, tb);
    }
    for (;;) {
      int32_t p;
      uint32_t node = 0;
      if (ta > kValid) {
        p = ta;  // "A", "|" or "-" of the first list passes through
      } else {
        p = t.ext32[lca(t.parent, na, nb)];  // get_lca, :471-492
        node = node_of(t, p);                // the call keys taxa by that value
      }
      int32_t fc;
      bool adv1, adv2;
      if (ca < cb) {
        cb = (int32_t) ((uint32_t) cb - (uint32_t) ca);
        fc = ca;
        adv1 = true;
        adv2 = cb == 0;
      } else if (cb < ca) {
        ca = (int32_t) ((uint32_t) ca - (uint32_t) cb);
        fc = cb;
        adv2 = true;
        adv1 = ca == 0;
      } else {
        fc = ca;
        adv1 = adv2 = true;
      }
      total += (uint64_t) (int64_t) fc;
      pr[m] = p;
      ct[m] = fc;
      m++;
      if (p <= kValid) {
        uint32_t k = d;
        if (d && dn[d - 1] == node) {
          k = d - 1;
        } else {
          for (k = 0; k < d && dn[k] != node; k++) {
          }
          if (k == d) {
            dn[d] = node;
            dc[d] = 0;
            d++;
          }
        }
        dc[k] += (uint32_t) fc;
      }
      if (adv1) {
        if (++i1 >= a1) break;
        ta = a_items[2 * i1];
        ca = a_items[2 * i1 + 1];
        na = ta > kValid ? 0 : node_of(t, ta);
      }
      if (adv2) ++i2;
      if (adv2 || i2 >= b1) {
        if (i2 < b1) {
          tb = b_items[2 * i2];
          cb = b_items[2 * i2 + 1];
          nb = node_of(t, tb);
        } else {
          tb = 0;
          cb = INT32_MAX;
          nb = 0;
        }
      }
    }
  }

  // ---- resolve_tree (src/merge.cc:369-421) over the distinct nodes;
  // markers carry no taxon.  Scores and sums wrap at 32 bits as the
  // reference's do.
  const uint32_t required = to_u32(ceilf(conf * (float) total));
  uint32_t max_taxon = 0, max_score = 0;
  for (uint32_t j = 0; j < d; j++) {
    const uint32_t taxon = dn[j];
    uint32_t score = 0;
    for (uint32_t k = 0; k < d; k++)
      if (above(t, dn[k], taxon)) score += dc[k];
    if (score > max_score) {
      max_score = score;
      max_taxon = taxon;
    } else if (score == max_score) {
      max_taxon = lca(t.parent, max_taxon, taxon);
    }
  }
  // hit_counts[external id of the call]: items whose taxid value equals it
  const uint64_t call_ext = t.ext64[max_taxon];
  max_score = 0;
  for (uint32_t k = 0; k < m; k++)
    if (pr[k] <= kValid && (uint64_t) (int64_t) pr[k] == call_ext) max_score += (uint32_t) ct[k];
  bool node_id = false;
  while (max_taxon && max_score < required) {
    max_score = 0;
    for (uint32_t k = 0; k < d; k++)
      if (above(t, max_taxon, dn[k])) max_score += dc[k];
    if (max_score >= required) {
      node_id = true;  // the node id is returned, as the reference does (:413)
      break;
    }
    max_taxon = t.parent[max_taxon];
  }
  call[r] = node_id ? (int32_t) max_taxon : t.ext32[max_taxon];

  // ---- write_hit_list (src/merge.cc:328-367): printed pairs
  uint32_t nt = 0;
  uint64_t chars = 0;
  if (m) {
    int32_t *tk = tokens + 2 * s;
    uint8_t *jn = joined + s;
    int32_t pt = pr[0], pc = ct[0];
    for (uint32_t j = 1; j < m; j++) {
      const int32_t tj = pr[j], cj = ct[j];
      if (tj == 0 && cj == 0) {
        tk[2 * nt] = pt;
        tk[2 * nt + 1] = pc;
        jn[nt++] = 1;
        pt = tj;
        pc = cj;
      }
      if (pt == tj) {
        pc = wrap_add(pc, cj);
      } else if (pc == 0 && cj > 0) {
        pc = cj;
        pt = tj;
      } else {
        tk[2 * nt] = pt;
        tk[2 * nt + 1] = pc;
        jn[nt++] = 0;
        pt = tj;
        pc = cj;
      }
    }
    tk[2 * nt] = pt;
    tk[2 * nt + 1] = pc;
    jn[nt++] = 1;
    for (uint32_t k = 0; k < nt; k++)
      chars += printed_len(tk[2 * k]) + 1 + printed_len(tk[2 * k + 1]) + (jn[k] ? 0 : 1);
  }
  n_tokens[r] = nt;
  text_len[r] = chars;
}

// parse_hit_list (src/merge.cc:423-469) of one hit-list field s[0..len)
// (the fifth field of a kraken line, with the line's newline) into
// out[0 .. slots) pairs, slots being the field's ':' count: tokens split at
// the first ':' and the next ' ', numbers read digit by digit with no
// checks (:217-244), the last count one character short (the newline), and
// slots the tokens do not reach left {0,0}.
__device__ __forceinline__ uint64_t field_atou(const char *p, uint64_t len) {
  if (len > 20) return 0;
  uint64_t r = 0;
  for (uint64_t i = 0; i < len; i++) r = r * 10 + (uint64_t) (int64_t) ((int) (unsigned char) p[i] - '0');
  return r;
}

__device__ __forceinline__ void parse_field(const char *s, uint64_t len, int32_t *out, uint64_t slots) {
  const char *p = s, *end = s + len;
  uint64_t i = 0;
  bool more = true;
  while (more && i < slots) {
    const char *colon = p;
    while (colon < end && *colon != ':') colon++;
    if (colon == end) break;
    const char *c = colon + 1, *sp = c;
    while (sp < end && *sp != ' ') sp++;
    const uint64_t len1 = colon - p;
    uint64_t len2;
    if (sp < end) {
      len2 = sp - c;
    } else {
      len2 = (uint64_t) (end - c) - 1;
      more = false;
    }
    const char first = len1 ? p[0] : '\0';
    int32_t taxid, count;
    if (first == 'A') {
      taxid = INT32_MAX;
      count = (int32_t) (uint32_t) field_atou(c, len2);
    } else if (first == '|') {
      taxid = count = INT32_MAX - 1;
    } else if (first == '-') {
      taxid = count = INT32_MAX - 2;
    } else {
      taxid = (int32_t) (uint32_t) field_atou(p, len1);
      count = (int32_t) (uint32_t) field_atou(c, len2);
    }
    out[2 * i] = taxid;
    out[2 * i + 1] = count;
    i++;
    p = sp + 1;
  }
  for (; i < slots; i++) out[2 * i] = out[2 * i + 1] = 0;
}

// ':' count of each read's hit-list field (its item slots), one thread per
// read; counts[n] is left for the caller (the scan's total).
__global__ void __launch_bounds__(256)
k_merge_count(const char *__restrict__ text, const uint64_t *__restrict__ field, const uint32_t *__restrict__ len,
              uint64_t n, uint64_t *__restrict__ counts) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const char *f = text + field[r];
  uint64_t colons = 0;
  for (uint32_t i = 0; i < len[r]; i++) colons += f[i] == ':';
  counts[r] = colons;
}

// One thread per read: field[r], len[r] locate read r's hit-list field in
// text; off (n+1) gives its item slots.  check: the offsets are the caller's,
// so each read's slot count must be its field's ':' count (offsets counted on
// the device are those counts already).
__global__ void __launch_bounds__(256)
k_merge_parse(const char *__restrict__ text, const uint64_t *__restrict__ field, const uint32_t *__restrict__ len,
              const uint64_t *__restrict__ off, uint64_t n, int32_t *__restrict__ items, uint32_t *__restrict__ bad,
              bool check) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const char *f = text + field[r];
  const uint64_t slots = off[r + 1] - off[r];
  if (check) {
    uint64_t colons = 0;
    for (uint32_t i = 0; i < len[r]; i++) colons += f[i] == ':';
    if (colons != slots) {
      atomicOr(bad, 1u);
      return;
    }
  }
  parse_field(f, len[r], items + 2 * off[r], slots);
}

// Sets bad bit 1 when a read's two hit lists hold 2^32 or more items together
// (k_merge counts a read's merged items in 32 bits).
__global__ void __launch_bounds__(256)
k_merge_limits(const uint64_t *__restrict__ a_off, const uint64_t *__restrict__ b_off, uint64_t n,
               uint32_t *__restrict__ bad) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if (a_off[r + 1] - a_off[r] + b_off[r + 1] - b_off[r] > 0xFFFFFFFFull) atomicOr(bad, 2u);
}

// Prints read r's pairs ("taxid:count", a space after each unless joined)
// at text[text_off[r]] (text_off: exclusive sum of text_len).
__global__ void __launch_bounds__(256)
k_merge_print(const uint64_t *__restrict__ a_off, const uint64_t *__restrict__ b_off, uint64_t n,
              const uint64_t *__restrict__ n_tokens, const uint64_t *__restrict__ text_off,
              const int32_t *__restrict__ tokens, const uint8_t *__restrict__ joined, char *__restrict__ text) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint64_t s = a_off[r] + b_off[r], nt = n_tokens[r];
  char *o = text + text_off[r];
  for (uint64_t k = 0; k < nt; k++) {
    o = print_int(o, tokens[2 * (s + k)]);
    *o++ = ':';
    o = print_int(o, tokens[2 * (s + k) + 1]);
    if (!joined[s + k]) *o++ = ' ';
  }
}

// Copies read r's printed pairs from its slots to [tok_off[r], tok_off[r+1])
// of the dense arrays (tok_off: exclusive sum of n_tokens).
__global__ void __launch_bounds__(256)
k_merge_compact(const uint64_t *__restrict__ a_off, const uint64_t *__restrict__ b_off, uint64_t n,
                const uint64_t *__restrict__ tok_off, const int32_t *__restrict__ tokens,
                const uint8_t *__restrict__ joined, int32_t *__restrict__ d_tokens, uint8_t *__restrict__ d_joined) {
  const uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint64_t s = a_off[r] + b_off[r], d = tok_off[r], nt = tok_off[r + 1] - d;
  const uint64_t *src = (const uint64_t *) (tokens + 2 * s);  // a (taxid, count) pair
  uint64_t *dst = (uint64_t *) (d_tokens + 2 * d);
  for (uint64_t k = 0; k < nt; k++) {
    dst[k] = src[k];
    d_joined[d + k] = joined[s + k];
  }
}

// Parent and external ids of taxo.k2d's nodes (TaxonomyNode,
// src/taxonomy.h:18-26); 0 or a K2G error code with err set.
inline int read_taxonomy_nodes(const char *path, std::vector<uint32_t> &parent, std::vector<uint64_t> &ext,
                               std::string &err) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    err = std::string("unable to open ") + path;
    return K2G_EOS;
  }
  char magic[8];
  uint64_t hdr[3];
  if (fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "K2TAXDAT", 8) != 0 || fread(hdr, 8, 3, f) != 3 ||
      hdr[0] == 0 || hdr[0] >= (1ULL << 31) - 4) {
    fclose(f);
    err = std::string("malformed taxonomy file ") + path;
    return K2G_EDATA;
  }
  std::vector<uint64_t> raw(hdr[0] * 7);
  const size_t got = fread(raw.data(), 56, hdr[0], f);
  fclose(f);
  if (got != hdr[0]) {
    err = std::string("read exhausted taxonomy information in ") + path;
    return K2G_EDATA;
  }
  parent.resize(hdr[0]);
  ext.resize(hdr[0]);
  for (uint64_t i = 0; i < hdr[0]; i++) {
    const uint64_t pid = raw[7 * i];
    if (i > 0 && pid >= i) {
      err = "taxonomy node " + std::to_string(i) + " has parent id " + std::to_string(pid) +
            " (ids must increase away from the root)";
      return K2G_EDATA;
    }
    parent[i] = (uint32_t) pid;
    ext[i] = raw[7 * i + 5];
  }
  parent[0] = 0;
  return K2G_OK;
}

// Host tables behind MergeTax.  External ids below 2^26 (NCBI taxids are)
// also get a direct-mapped table, one load per lookup instead of a search.
struct MergeTables {
  std::vector<uint32_t> span, key_node, direct;
  std::vector<int32_t> ext32;
  std::vector<uint64_t> key;
};

constexpr uint64_t kMergeDirectLimit = 1ull << 26;

inline void merge_tables(const std::vector<uint32_t> &parent, const std::vector<uint64_t> &ext, MergeTables &t,
                         uint64_t direct_limit = kMergeDirectLimit) {
  const size_t n = parent.size();
  t.span.resize(2 * n);
  taxonomy_spans(parent.data(), n, t.span.data());
  t.ext32.resize(n);
  for (size_t i = 0; i < n; i++) t.ext32[i] = (int32_t) (uint32_t) ext[i];
  // GenerateExternalToInternalIDMap (src/taxonomy.cc:289-294): 0 -> 0, then
  // every node in id order, a later node overwriting an equal external id
  std::map<uint64_t, uint32_t> m;
  m[0] = 0;
  for (size_t i = 1; i < n; i++) m[ext[i]] = (uint32_t) i;
  t.key.clear();
  t.key_node.clear();
  for (const auto &kv : m) {
    t.key.push_back(kv.first);
    t.key_node.push_back(kv.second);
  }
  t.direct.clear();
  if (t.key.back() < direct_limit) {
    t.direct.assign(t.key.back() + 1, 0);
    for (const auto &kv : m) t.direct[kv.first] = kv.second;
  }
}

}  // namespace k2g
