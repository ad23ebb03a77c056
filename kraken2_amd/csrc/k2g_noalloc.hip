// k2g_classify_device_noalloc: the classify launch sequence of
// k2g_classify_device for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Every scratch buffer of the sequence -- element bounds, run
// offsets' source, status words, the longest-first sort and its work units,
// redo lists, cursors, hipCUB temporary storage, the dense hit tables of the
// overflow path -- is carved from one statically allocated device array per
// GPU; the results go to device arrays the caller passes.
#include "k2g.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "k2g_internal.h"

namespace {

constexpr size_t kScratchBytes = 128ull << 20;      // per GPU, held until the process ends
constexpr uint64_t kMaxFragments = 1ull << 19;      // per call
constexpr int kMaxDevices = 64;

// The scratch array.  In host code its name is a stand-in, not a device
// address: only the pointer hipGetSymbolAddress returns on a device is ever
// passed to kernels or HIP calls.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
  std::mutex lock;        // host-side state of one device's calls
  bool ready = false;
  void *ptr = nullptr;    // this device's copy of `scratch`
  k2g_batch *batch = nullptr;
  k2g::Region region;
};
DeviceScratch g_dev[kMaxDevices];

// Once per device, with that device current: the address of its scratch
// array and the batch state whose buffers are carved from it (kernel
// occupancy for the persistent grid).  No device memory is allocated.
int setup(int dev, DeviceScratch &d) {
  if (d.ready) return K2G_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return k2g::fail(K2G_ESOFTWARE, std::string("scratch address on device ") + std::to_string(dev) + ": " +
                                        hipGetErrorString(e));
  auto *b = new k2g_batch();
  for (int vi = 0; vi < 3; vi++) b->persistent_blocks[vi] = 0;
  d.ptr = ptr;
  d.batch = b;
  d.region.base = (char *) ptr;
  d.region.size = kScratchBytes;
  d.ready = true;
  return K2G_OK;
}

}  // namespace

extern "C" int k2g_classify_device_noalloc(k2g_db *db, const k2g_classify_params *p, const char *d_seq,
                                           const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                                           const k2g_device_results *out, void *hip_stream) {
  using namespace k2g;
  if (!db || !p || !out) return fail(K2G_EUSAGE, "null argument");
  if (n > kMaxFragments)
    return fail(K2G_EUSAGE, "k2g_classify_device_noalloc takes at most " + std::to_string(kMaxFragments) +
                                " fragments per call (" + std::to_string(n) + " given)");
  if (p->want_probe_stats) return fail(K2G_EUSAGE, "probe statistics need k2g_classify_device");
  if (n && (!out->call || !out->hit_groups || !out->total_kmers || !out->n_runs || !out->run_off))
    return fail(K2G_EUSAGE, "null result array");
  if (p->want_hll && (!out->hll || !out->n_hll)) return fail(K2G_EUSAGE, "want_hll needs hll and n_hll arrays");
  if (p->want_kmer_counts && !out->kmer_counts) return fail(K2G_EUSAGE, "want_kmer_counts needs kmer_counts");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(K2G_EUSAGE, "device index out of range");
  if (dev != db->device) return fail(K2G_EUSAGE, "the database is resident on another device");
  DeviceScratch &d = g_dev[dev];
  std::lock_guard<std::mutex> guard(d.lock);
  int rc = setup(dev, d);
  if (rc) return rc;
  k2g_batch *b = d.batch;
  if (!b->persistent_blocks[0]) {
    rc = persistent_grid(b, db->num_cus);
    if (rc) return rc;
  }
  // every call carves the whole scratch anew: nothing carries over
  d.region.used = 0;
  b->bind_region(&d.region);
  b->db = db;
  b->stream = (hipStream_t) hip_stream;
  b->ev0 = b->ev1 = nullptr;
  b->fixed_outputs = true;
  const uint64_t m = n ? n : 1;
  b->call.adopt(out->call, m, &d.region);
  b->groups.adopt(out->hit_groups, m, &d.region);
  b->total.adopt(out->total_kmers, m, &d.region);
  b->n_runs.adopt(out->n_runs, m, &d.region);
  b->n_recs.adopt(p->want_hll ? out->n_hll : nullptr, p->want_hll ? m : 0, &d.region);
  b->run_off.adopt(out->run_off, m + 1, &d.region);
  b->runs.adopt(out->runs, out->runs ? 2 * out->run_capacity : 0, &d.region);
  b->recs.adopt(p->want_hll ? out->hll : nullptr, p->want_hll ? 2 * out->run_capacity : 0, &d.region);
  b->kmer_counts.adopt(reinterpret_cast<unsigned long long *>(p->want_kmer_counts ? out->kmer_counts : nullptr),
                       p->want_kmer_counts ? db->view.node_count : 0, &d.region);
  // exact hitlist space: no fragment is classified a second time for room
  k2g_classify_params q = *p;
  q.run_cap_pct = 100;
  rc = classify_launch(b, &q, d_seq, d_off, d_paired, n, nullptr);
  b->stream = nullptr;
  return rc;
}
