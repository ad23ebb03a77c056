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

unsigned long long emu_collectives() { return emu::collectives.load(); }

}  // extern "C"
