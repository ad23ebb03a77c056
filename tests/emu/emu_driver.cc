// Host driver of the SIMT emulation: runs the classify kernels of
// kraken2_amd/csrc/k2g_kernels.hip block by block with 64 host threads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../kraken2_amd/csrc/k2g_kernels.hip"

namespace emu {
thread_local Dim3 threadIdx_, blockIdx_;
Dim3 blockDim_, gridDim_;
std::barrier<> *bar = nullptr;
unsigned long long vote_slot[64];
uint64_t shfl_slot[64];
std::atomic<unsigned long long> collectives{0};
}  // namespace emu

template <class F>
static void run_grid(unsigned grid, F body) {
  std::barrier<> b(64);
  emu::bar = &b;
  emu::gridDim_.x = grid;
  emu::blockDim_.x = 64;
  for (unsigned blk = 0; blk < grid; blk++) {
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < 64; t++)
      ts.emplace_back([&, t, blk] {
        emu::threadIdx_.x = t;
        emu::blockIdx_.x = blk;
        body();
      });
    for (auto &t : ts) t.join();
  }
}

// segment size of the split path (the library's kSegElements)
constexpr uint32_t kSegElementsEmu = 2048;

extern "C" {

// Classifies n fragments with the device code.  cells padded to a multiple
// of 64; parent has node_count entries.  mode 0 plain, 1 persistent with the
// given grid.  run_off/run_cap (n) and runs/recs buffers prepared by the caller.
int emu_classify(const uint32_t *cells, uint64_t capacity, uint32_t kb, uint32_t vb, const uint32_t *parent,
                 uint32_t node_count, uint32_t k, uint32_t l, uint32_t dna, uint64_t spaced, uint64_t toggle,
                 uint64_t min_hash, uint32_t revcom_version, double conf, int32_t min_groups, uint32_t quick, uint32_t tab_cap,
                 const char *seq, const uint64_t *off, const uint8_t *paired, uint32_t paired_all, uint64_t n,
                 const uint32_t *order, int mode, unsigned grid, const uint64_t *run_off,
                 const uint64_t *run_cap, const uint32_t *unit_start, const uint32_t *n_units,
                 uint32_t *call,
                 uint32_t *groups, uint32_t *total, uint32_t *n_runs, uint32_t *status, uint32_t *n_recs,
                 uint32_t *runs, uint32_t *recs, unsigned long long *kmer_counts) {
  using namespace k2g;
  DbView v{};
  v.cells = cells;
  v.capacity = capacity;
  v.barrett = capacity > 1 ? (uint64_t) (((unsigned __int128) 1 << 64) / capacity) : 0;
  v.key_bits = kb;
  v.value_bits = vb;
  v.parent = parent;
  std::vector<uint32_t> span(2 * (size_t) node_count);
  taxonomy_spans(parent, node_count, span.data());
  v.span = span.data();
  v.node_count = node_count;
  v.k = k;
  v.l = l;
  v.dna = dna;
  v.revcom_version = revcom_version;
  const uint32_t bits = dna ? 2 : 4;
  v.lmer_mask = (1ULL << (l * bits)) - 1;
  v.spaced_mask = spaced;
  v.toggle_mask = toggle & v.lmer_mask;
  v.min_hash = min_hash;
  // the library's load-time rule: an empty cell and at least 64 cells give
  // the lean probe (the caller pads the cells with copies of the first ones)
  {
    const uint32_t vmask = (1u << vb) - 1u;
    bool empty = false;
    for (uint64_t i = 0; i < capacity && !empty; i++) empty = (cells[i] & vmask) == 0;
    v.lean_probe = (empty && capacity >= 64) ? 1u : 0u;
  }
  ClassifyParams P{conf, min_groups, quick, tab_cap, 8};
  FragBatch in{seq, off, paired, paired_all, n, order, unit_start, n_units};
  FragOut out{call,    groups, total, n_runs, status,      n_recs,  run_off,
              run_cap, runs,   recs,  kmer_counts, nullptr, nullptr};
  // the variant the library would launch for this database and these outputs
  const bool extras = recs || kmer_counts || quick;
  if (mode == 0) {
    run_grid((unsigned) n, [&] {
      if (!dna) k_classify_plain_prot(v, P, in, out);
      else if (extras) k_classify_plain_dna_x(v, P, in, out);
      else k_classify_plain_dna(v, P, in, out);
    });
  } else {
    // the library launches at least min(8, units) waves (one per cursor)
    static unsigned long long cursor[kCursorWords];
    for (auto &x : cursor) x = 0;
    if (grid < 8 && grid < n) grid = (unsigned) (n < 8 ? n : 8);
    run_grid(grid, [&] {
      if (!dna) k_classify_persistent_prot(v, P, in, out, cursor);
      else if (extras) k_classify_persistent_dna_x(v, P, in, out, cursor);
      else k_classify_persistent_dna(v, P, in, out, cursor);
    });
  }
  return 0;
}

// The split path of the library's persistent mode: the long fragments'
// segment layout (k_seg_count, a host scan, k_seg_plan, a host scan of the
// run spaces), the segments, the persistent kernel skipping split fragments,
// then the stitch.  elements: each fragment's exact element count.  Same
// arguments as emu_classify otherwise (seg_pct: the segments' reserved run
// space in percent; 100 = exact).  Returns the number of segments.
int emu_classify_split(const uint32_t *cells, uint64_t capacity, uint32_t kb, uint32_t vb, const uint32_t *parent,
                       uint32_t node_count, uint32_t k, uint32_t l, uint64_t spaced, uint64_t toggle,
                       uint64_t min_hash, uint32_t revcom_version, double conf, int32_t min_groups, uint32_t tab_cap,
                       const char *seq, const uint64_t *off, const uint8_t *paired, uint64_t n,
                       const uint32_t *order, unsigned grid, const uint64_t *run_off, const uint64_t *run_cap,
                       const uint64_t *elements, uint32_t seg_pct, uint32_t *call, uint32_t *groups,
                       uint32_t *total, uint32_t *n_runs, uint32_t *status, uint32_t *n_recs, uint32_t *runs,
                       uint32_t *recs, unsigned long long *kmer_counts) {
  using namespace k2g;
  DbView v{};
  v.cells = cells;
  v.capacity = capacity;
  v.barrett = capacity > 1 ? (uint64_t) (((unsigned __int128) 1 << 64) / capacity) : 0;
  v.key_bits = kb;
  v.value_bits = vb;
  v.parent = parent;
  std::vector<uint32_t> span(2 * (size_t) node_count);
  taxonomy_spans(parent, node_count, span.data());
  v.span = span.data();
  v.node_count = node_count;
  v.k = k;
  v.l = l;
  v.dna = 1;
  v.revcom_version = revcom_version;
  v.lmer_mask = (1ULL << (l * 2)) - 1;
  v.spaced_mask = spaced;
  v.toggle_mask = toggle & v.lmer_mask;
  v.min_hash = min_hash;
  {
    const uint32_t vmask = (1u << vb) - 1u;
    bool empty = false;
    for (uint64_t i = 0; i < capacity && !empty; i++) empty = (cells[i] & vmask) == 0;
    v.lean_probe = (empty && capacity >= 64) ? 1u : 0u;
  }
  ClassifyParams P{conf, min_groups, 0u, tab_cap, 8};
  FragBatch in{seq, off, paired, 0u, n, order, nullptr, nullptr};
  FragOut out{call,    groups, total, n_runs, status,      n_recs,  run_off,
              run_cap, runs,   recs,  kmer_counts, nullptr, nullptr};
  const bool extras = recs || kmer_counts;
  // 1-2. segment counts, split list, starts
  std::vector<uint32_t> cnt(n + 1), list(n + 1), count(4, 0);
  run_grid((unsigned) ((n + 1 + 63) / 64), [&] {
    k_seg_count(off, paired, 0u, n, k, kSegElementsEmu, elements, cnt.data(), list.data(), count.data());
  });
  std::vector<uint64_t> start(n + 1, 0);
  for (uint64_t i = 0; i < n; i++) start[i + 1] = start[i] + cnt[i];
  const uint64_t segs = start[n];
  std::vector<uint32_t> frag(segs + 1), mate(segs + 1), a(segs + 1), b(segs + 1), nr(segs + 1), nrec(segs + 1),
      grp(segs + 1), st(segs + 1);
  std::vector<uint64_t> cap(segs + 1, 0), soff(segs + 1, 0);
  SegBatch sg{frag.data(), mate.data(), a.data(), b.data(), soff.data(), cap.data(), nr.data(), nrec.data(),
              grp.data(),  st.data(),   nullptr,  nullptr,  count.data(), list.data(), cnt.data(), start.data()};
  // 3-4. cut points and run spaces
  run_grid((unsigned) std::max<uint32_t>(1, count[1]), [&] {
    k_seg_plan(seq, off, paired, 0u, k, kSegElementsEmu, seg_pct, list.data(), start.data(), count.data(), n, sg);
  });
  for (uint64_t s = 0; s < segs; s++) soff[s + 1] = soff[s] + cap[s];
  std::vector<uint32_t> sruns(2 * std::max<uint64_t>(soff[segs], 1)), srecs(2 * std::max<uint64_t>(soff[segs], 1));
  sg.runs = sruns.data();
  sg.recs = recs ? srecs.data() : nullptr;
  // 5. segments, persistent (skipping split fragments), stitch
  run_grid(3, [&] {
    if (extras) k_classify_segments_dna_x(v, P, in, out, sg);
    else k_classify_segments_dna(v, P, in, out, sg);
  });
  in.split = cnt.data();
  static unsigned long long cursor[kCursorWords];
  for (auto &x : cursor) x = 0;
  if (grid < 8 && grid < n) grid = (unsigned) (n < 8 ? n : 8);
  run_grid(grid, [&] {
    if (extras) k_classify_persistent_dna_x(v, P, in, out, cursor);
    else k_classify_persistent_dna(v, P, in, out, cursor);
  });
  run_grid(2, [&] {
    if (extras) k_stitch_dna_x(v, P, in, out, sg);
    else k_stitch_dna(v, P, in, out, sg);
  });
  return (int) segs;
}

unsigned long long emu_collectives() { return emu::collectives.load(); }

}  // extern "C"
