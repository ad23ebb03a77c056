// Host-side internals of libk2g.so shared by the classify driver
// (k2g_api.hip) and the request server (k2g_server.hip): device and pinned
// buffers, the resident database and the batch state.
#pragma once
#include "k2g.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "host/k2g_host.h"
#include "k2g_device.h"
#include "k2g_kernels.h"

namespace k2g {

// Sets the calling thread's last error message and returns code.
int fail(int code, const std::string &msg);
// Last error message of the calling thread.
const std::string &last_error();

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::k2g::fail(K2G_ESOFTWARE, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

// A device range that a batch carves its buffers from by bumping `used`
// (never freed piecemeal).  Lets a caller bound a batch's device memory to
// a range it admitted in advance.
struct Region {
  char *base = nullptr;
  size_t size = 0;
  size_t used = 0;
};

constexpr size_t kRegionAlign = 256;

inline size_t region_bytes(size_t n, size_t elem) {
  const size_t b = std::max<size_t>(n, 64) * elem;
  return (b + kRegionAlign - 1) / kRegionAlign * kRegionAlign;
}

// Growable device buffer: hipMalloc'd, or carved from a Region when one is
// bound (then growth takes new space in the region and release is a no-op).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;
  Region *region = nullptr;
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    const size_t want = std::max<size_t>(n, 64);
    if (region) {
      const size_t off = (region->used + kRegionAlign - 1) / kRegionAlign * kRegionAlign;
      const size_t bytes = region_bytes(want, sizeof(T));
      if (off + bytes > region->size) return hipErrorOutOfMemory;
      p = reinterpret_cast<T *>(region->base + off);
      cap = want;
      region->used = off + bytes;
      return hipSuccess;
    }
    if (p) (void) hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void **) &p, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p && !region) (void) hipFree(p);
    p = nullptr;
    cap = 0;
  }
  // Drops the current storage (freeing it if owned) and allocates from r
  // (or with hipMalloc when r is null) from now on.
  void bind(Region *r) {
    release();
    region = r;
  }
  // Uses the caller's n elements at q (never freed: r must be non-null, the
  // region the rest of the batch is carved from).
  void adopt(T *q, size_t n, Region *r) {
    region = r;
    p = q;
    cap = q ? n : 0;
  }
};

// First-fit allocator of byte ranges of [0, size) (not thread-safe).
class ByteArena {
 public:
  explicit ByteArena(uint64_t size = 0) {
    if (size) free_[0] = size;
  }
  bool alloc(uint64_t bytes, uint64_t *off) {
    if (bytes == 0) {
      *off = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes;
      const uint64_t at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    if (bytes == 0) return;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<uint64_t, uint64_t> free_;
};

// One pinned, device-mapped host allocation carved into buffers (thread-safe).
// Freeing pinned memory waits for the whole device, so a request loop that
// grows and drops result buffers takes them from here instead.
struct HostArena {
  char *base = nullptr;
  uint64_t size = 0;
  std::mutex mu;
  ByteArena ranges;
  // frees idle memory of the arena's owner back into it (false: none left);
  // called without the arena's lock held
  std::function<bool()> reclaim;
  void *take(uint64_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    for (;;) {
      {
        std::lock_guard<std::mutex> lk(mu);
        uint64_t off = 0;
        if (ranges.alloc(bytes, &off)) return base + off;
      }
      if (!reclaim || !reclaim()) return nullptr;
    }
  }
  void give(void *p, uint64_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    std::lock_guard<std::mutex> lk(mu);
    ranges.release((uint64_t) ((char *) p - base), bytes);
  }
};

// Growable pinned host buffer, also addressed by kernels (zero-copy writes);
// from a HostArena when one is set (hipHostMalloc when it is full).
template <class T>
struct PinnedBuf {
  T *p = nullptr;
  size_t cap = 0;
  HostArena *arena = nullptr;
  bool in_arena = false;
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    release();
    const size_t want = std::max<size_t>(n + n / 2, 1024);
    if (arena && (p = (T *) arena->take(want * sizeof(T))) != nullptr) {
      in_arena = true;
      cap = want;
      return hipSuccess;
    }
    hipError_t e = hipHostMalloc((void **) &p, want * sizeof(T), hipHostMallocMapped);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p && in_arena) arena->give(p, cap * sizeof(T));
    else if (p) (void) hipHostFree(p);
    p = nullptr;
    cap = 0;
    in_arena = false;
  }
};

// Host copies of a batch's last results (hitlists and HLL records are
// written by the compaction kernel straight into the pinned buffers).  A
// batch owns one; the request server hands it to the finished request and
// gives the batch another from its pool, so results are never copied.
struct HostResults {
  // all pinned, so the device copies into them asynchronously
  PinnedBuf<uint32_t> call, groups, total, lookups, runs, recs, words;
  PinnedBuf<uint64_t> run_off, rec_off, kmer_counts, lines;
  // per-(block, taxon) HLL sets (want_hll == 2); sk_words: groups, values
  PinnedBuf<uint32_t> sk_block, sk_taxon, sk_dense, sk_vals;
  PinnedBuf<uint64_t> sk_n, sk_off, sk_words;
  uint64_t served = 0;  // largest element bound of a call these buffers held (request server)
  HostResults() = default;
  explicit HostResults(HostArena *a) {
    for (auto *x : {&call, &groups, &total, &lookups, &runs, &recs, &words, &sk_block, &sk_taxon, &sk_dense,
                    &sk_vals})
      x->arena = a;
    for (auto *x : {&run_off, &rec_off, &kmer_counts, &lines, &sk_n, &sk_off, &sk_words}) x->arena = a;
  }
  // pinned bytes held
  uint64_t bytes() const {
    uint64_t b = 0;
    for (auto *x : {&call, &groups, &total, &lookups, &runs, &recs, &words, &sk_block, &sk_taxon, &sk_dense,
                    &sk_vals})
      b += x->cap * 4;
    for (auto *x : {&run_off, &rec_off, &kmer_counts, &lines, &sk_n, &sk_off, &sk_words}) b += x->cap * 8;
    return b;
  }
  HostResults(const HostResults &) = delete;
  HostResults &operator=(const HostResults &) = delete;
  ~HostResults() {
    for (auto *b : {&call, &groups, &total, &lookups, &runs, &recs, &words, &sk_block, &sk_taxon, &sk_dense,
                    &sk_vals})
      b->release();
    for (auto *b : {&run_off, &rec_off, &kmer_counts, &lines, &sk_n, &sk_off, &sk_words}) b->release();
  }
};

}  // namespace k2g

struct k2g_db {
  k2g::DbView view{};
  uint32_t *d_cells = nullptr;
  uint32_t *d_parent = nullptr;
  uint32_t *d_span = nullptr;
  std::vector<uint64_t> external_ids;
  k2host::Taxonomy tax;  // names and ranks, for reports
  k2g_db_info info{};
  int device = 0;
  int num_cus = 0;
};

struct k2g_batch {
  k2g_db *db = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // inputs (host-path uploads)
  k2g::DevBuf<char> seq;
  k2g::DevBuf<uint64_t> off;
  k2g::DevBuf<uint8_t> paired;
  // per-fragment device outputs
  k2g::DevBuf<uint32_t> call, groups, total, n_runs, status, n_recs, order, redo;
  k2g::DevBuf<uint64_t> run_off, run_cap, bounds, comp_off, comp_rec_off, space_bounds, space_off,
      digest;
  k2g::DevBuf<uint32_t> runs, recs;
  k2g::DevBuf<unsigned long long> kmer_counts, cursor, lines;
  k2g::DevBuf<uint32_t> lookups;
  k2g::DevBuf<uint32_t> redo_count, space, space_only;
  k2g::DevBuf<unsigned char> temp;
  k2g::DevBuf<uint32_t> sort_keys, sort_keys_out, sort_vals;
  k2g::DevBuf<uint32_t> dense_counts, dense_lists, dense_lens;
  k2g::DevBuf<uint64_t> unit_pos, unit_scan;
  k2g::DevBuf<uint8_t> unit_flag;
  k2g::DevBuf<uint32_t> unit_start, n_units;
  k2g::DevBuf<unsigned int> fetch_word;  // OR of the status words at fetch
  // per-(block, taxon) HLL sets (want_hll == 2): the call's input-block
  // starts, sort keys and record indices, groups
  k2g::DevBuf<uint64_t> sk_starts, sk_keys, sk_wide, sk_start, sk_off;
  k2g::DevBuf<uint32_t> sk_idx, sk_idx_alt, sk_gkey, sk_n, sk_dense, sk_count;
  // long fragments split over waves (persistent mode): per-fragment segment
  // counts and starts, the split list, the segments and their scratch
  k2g::DevBuf<uint32_t> sp_cnt, sp_list, sp_count, sp_frag, sp_mate, sp_a, sp_b, sp_runs_n, sp_nrec, sp_groups,
      sp_status, sp_runs, sp_recs;
  k2g::DevBuf<uint64_t> sp_wide, sp_start, sp_cap, sp_off;
  uint64_t sp_segs = 0;       // segment capacity of the last split call
  bool sp_active = false;     // the call being launched splits
  bool split_last = false;    // the last call split fragments
  uint64_t sk_blocks = 0;    // input blocks of the last call with want_hll == 2
  bool sk_pending = false;   // k2g_classify uploaded the next call's block starts
  uint32_t dense_slots = 0;
  // host copies of the last results
  std::unique_ptr<k2g::HostResults> host = std::make_unique<k2g::HostResults>();
  // last launch
  uint64_t n = 0;
  k2g_classify_params params{};
  uint64_t redone = 0, space_redone = 0;
  bool ran_persistent = false;  // which kernel the last call launched
  double kernel_ms = 0;
  int persistent_blocks[3] = {0, 0, 0};  // per kernel variant: resident waves
  // outputs (call .. n_recs, run_off, runs, recs, kmer_counts) are the
  // caller's arrays, adopted with their sizes: never grown or freed
  bool fixed_outputs = false;
  // the next call's batch statistics, when the caller computed them on the
  // host (the server's admission plan): the launch then reads nothing back
  struct HostPlan {
    bool valid = false, too_long = false;
    uint64_t space = 0, max_e = 0, sum_e = 0;
    uint64_t split_sum = 0, split_cnt = 0;  // elements and number of the fragments a split would cut
  } plan;
  // the dense-table redo of the last call was sized on the device: its
  // count is read when results are fetched
  bool redo_on_device = false;

  template <class F>
  void for_each_device_buf(F f) {
    f(seq), f(off), f(paired);
    f(call), f(groups), f(total), f(n_runs), f(status), f(n_recs), f(order), f(redo);
    f(run_off), f(run_cap), f(bounds), f(comp_off), f(comp_rec_off), f(space_bounds),
        f(space_off), f(digest);
    f(runs), f(recs), f(kmer_counts), f(cursor), f(lines), f(lookups);
    f(redo_count), f(space), f(space_only), f(temp), f(sort_keys), f(sort_keys_out), f(sort_vals);
    f(dense_counts), f(dense_lists), f(dense_lens);
    f(unit_pos), f(unit_scan), f(unit_flag), f(unit_start), f(n_units), f(fetch_word);
    f(sk_starts), f(sk_keys), f(sk_wide), f(sk_start), f(sk_off);
    f(sk_idx), f(sk_idx_alt), f(sk_gkey), f(sk_n), f(sk_dense), f(sk_count);
    f(sp_cnt), f(sp_list), f(sp_count), f(sp_frag), f(sp_mate), f(sp_a), f(sp_b), f(sp_runs_n), f(sp_nrec),
        f(sp_groups), f(sp_status), f(sp_runs), f(sp_recs);
    f(sp_wide), f(sp_start), f(sp_cap), f(sp_off);
  }
  // Allocates every device buffer from r from now on (null: hipMalloc).
  void bind_region(k2g::Region *r) {
    for_each_device_buf([r](auto &buf) { buf.bind(r); });
    dense_slots = 0;
    split_last = false;
  }
};

namespace k2g {
// The launch sequence of k2g_classify_device (results left in the batch's
// device buffers, or in adopted caller arrays when b->fixed_outputs).
int classify_launch(k2g_batch *b, const k2g_classify_params *p, const char *d_seq,
                    const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                    const uint32_t *d_order);
// Long nucleotide fragments split over waves: segments of about this many
// k-mers, for fragments of more than twice as many, in batches whose longest
// fragment outweighs an even share of the batch per resident wave.
constexpr uint32_t kSegElements = 2048;
bool split_wanted(const k2g_classify_params *p, bool dna, bool persistent, uint64_t max_e, uint64_t sum_e,
                  uint64_t waves);
// Device bytes a split takes (upper bound, for Region plans), given the
// elements and number of the fragments it cuts.
uint64_t split_plan_bytes(uint64_t n, uint64_t split_sum, uint64_t split_cnt, bool hll);
// phase 1 (before the persistent kernel): plan and launch the segments, and
// set in.split; phase 2 (after it): the stitch.  Enqueued on b->stream.
int split_launch(k2g_batch *b, const k2g_classify_params *p, const ClassifyParams &cp, FragBatch &in,
                 const FragOut &out, uint64_t split_sum, uint64_t split_cnt, int phase);
// The reference's input blocks over a call's fragments (src/classify.cc:
// 523-543): the first fragment of each block, then n.
void input_blocks(const k2g_fragments *in, bool paired_all, std::vector<uint64_t> &starts);
// Per-(block, taxon) HLL sets (want_hll == 2): whether a call's block and
// taxon ids fit the sort key; the device bytes the reduction takes over a
// call of `elements` k-mer elements; the upload of a call's block starts
// (sets b->sk_blocks); the reduction of the last call's records into h
// (enqueued on the batch's stream; the caller waits).
int sketch_check(const k2g_db *db, uint64_t n_blocks);
uint64_t sketch_plan_bytes(const k2g_db *db, uint64_t elements, uint64_t n_blocks);
int sketch_upload_blocks(k2g_batch *b, const std::vector<uint64_t> &starts);
int sketch_fetch(k2g_batch *b, HostResults &h, uint64_t records);
// Resident waves of each persistent kernel variant on the current device,
// into b->persistent_blocks.
int persistent_grid(k2g_batch *b, int num_cus);
// Device bytes a k2g_classify call (host fragments, results fetched) takes
// from a bound Region, as an upper bound; requires run_cap_pct == 100.
uint64_t plan_classify_bytes(const k2g_db *db, const k2g_classify_params *p,
                             const k2g_fragments *in, k2g_batch::HostPlan *stats = nullptr);
// Dense hit tables the overflow path of a call over n fragments takes when
// it is launched without a host round trip (at most 64 MiB of them).
inline uint32_t dense_slots_for(uint64_t n, uint64_t node_count) {
  const uint64_t fit = (64ull << 20) / (8 * (node_count ? node_count : 1));
  return (uint32_t) std::max<uint64_t>(1, std::min<uint64_t>(std::min<uint64_t>(n, 256), fit));
}
}  // namespace k2g
