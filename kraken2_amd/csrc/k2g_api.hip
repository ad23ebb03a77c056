// Host side of libk2g.so: resident-DB management, batch buffers and the
// launch sequence of the classify core.  Implements include/k2g.h.
#include "k2g.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "k2g_device.h"
#include "k2g_internal.h"
#include "k2g_kernels.h"

namespace k2g {
#define K2G_DECLARE_VARIANT(SUFFIX)                                                      \
  __global__ void k_classify_plain##SUFFIX(DbView, ClassifyParams, FragBatch, FragOut);   \
  __global__ void k_classify_persistent##SUFFIX(DbView, ClassifyParams, FragBatch, FragOut, \
                                                unsigned long long *);
K2G_DECLARE_VARIANT(_dna)
K2G_DECLARE_VARIANT(_dna_x)
K2G_DECLARE_VARIANT(_prot)
#undef K2G_DECLARE_VARIANT
__global__ void k_classify_dense_dna(DbView, ClassifyParams, FragBatch, FragOut, const uint32_t *,
                                     const uint32_t *, DenseScratch);
__global__ void k_classify_dense_prot(DbView, ClassifyParams, FragBatch, FragOut, const uint32_t *,
                                      const uint32_t *, DenseScratch);

// The kernel variants a batch launches: nucleotide without per-hit extras or
// quick mode (the hot path), nucleotide with them, translated search.
struct ClassifyKernels {
  void (*plain)(DbView, ClassifyParams, FragBatch, FragOut);
  void (*persistent)(DbView, ClassifyParams, FragBatch, FragOut, unsigned long long *);
  void (*dense)(DbView, ClassifyParams, FragBatch, FragOut, const uint32_t *, const uint32_t *,
                DenseScratch);
};
inline int variant_index(bool dna, bool extras) { return dna ? (extras ? 1 : 0) : 2; }
const ClassifyKernels kVariants[3] = {
    {k_classify_plain_dna, k_classify_persistent_dna, k_classify_dense_dna},
    {k_classify_plain_dna_x, k_classify_persistent_dna_x, k_classify_dense_dna},
    {k_classify_plain_prot, k_classify_persistent_prot, k_classify_dense_prot},
};
__global__ void k_get_batch(DbView, const uint64_t *, uint32_t *, uint64_t);
__global__ void k_max_value(DbView, unsigned int *);
__global__ void k_pad_copy(uint32_t *, uint64_t, uint32_t);
__global__ void k_build(BuildView, const char *, const SubBlock *, uint64_t);
__global__ void k_verify_keys(DbView, const uint64_t *, const uint32_t *, uint64_t, uint64_t *,
                              uint32_t *, unsigned long long *, unsigned long long *);
__global__ void k_verify_cells(DbView, const uint64_t *, const uint32_t *, uint64_t,
                               unsigned long long *);
}  // namespace k2g

using namespace k2g;

namespace k2g {

namespace {
thread_local std::string g_error;
}  // namespace

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

const std::string &last_error() { return g_error; }

// Grows buf to n elements keeping its first `keep`.
template <class T>
hipError_t grow_keep(DevBuf<T> &buf, size_t n, size_t keep, hipStream_t s) {
  if (n <= buf.cap) return hipSuccess;
  if (buf.region) {  // fresh space in the region, then copy
    DevBuf<T> q;
    q.region = buf.region;
    hipError_t e = q.reserve(n);
    if (e == hipSuccess && keep)
      e = hipMemcpyAsync(q.p, buf.p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    buf.p = q.p;
    buf.cap = q.cap;
    return hipSuccess;
  }
  T *q = nullptr;
  hipError_t e = hipMalloc((void **) &q, n * sizeof(T));
  if (e != hipSuccess) return e;
  if (keep) e = hipMemcpyAsync(q, buf.p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void) hipFree(q);
    return e;
  }
  if (buf.p) (void) hipFree(buf.p);
  buf.p = q;
  buf.cap = n;
  return hipSuccess;
}

}  // namespace k2g

namespace {

bool read_fully(int fd, void *buf, size_t n, off_t off) {
  char *p = (char *) buf;
  while (n) {
    ssize_t r = pread(fd, p, n, off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    p += r;
    n -= (size_t) r;
    off += r;
  }
  return true;
}

}  // namespace

extern "C" const char *k2g_last_error(void) { return k2g::last_error().c_str(); }

// ----------------------------------------------------------------------
// DB open: src/classify.cc:269-302, src/compact_hash.h:278-341,
// src/taxonomy.cc:195-238
extern "C" int k2g_db_open(const char *hash_k2d, const char *taxo_k2d, const char *opts_k2d,
                           k2g_db **out) {
  if (!hash_k2d || !taxo_k2d || !opts_k2d || !out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  // opts.k2d: raw IndexOptions (src/kraken2_data.h:15-25), read up to the
  // file size into a zeroed struct (src/classify.cc:272-279).
  struct RawOpts {
    uint64_t k, l, spaced, toggle;
    uint8_t dna, pad[7];
    uint64_t min_hash;
    int32_t revcom, db_version, db_type, pad2;
  } ro;
  static_assert(sizeof(RawOpts) == 64, "opts.k2d layout");
  std::memset(&ro, 0, sizeof(ro));
  {
    int fd = open(opts_k2d, O_RDONLY);
    if (fd < 0) return fail(K2G_EOS, std::string("unable to open ") + opts_k2d);
    struct stat sb;
    fstat(fd, &sb);
    size_t want = std::min<size_t>(sizeof(ro), (size_t) sb.st_size);
    bool ok = read_fully(fd, &ro, want, 0);
    close(fd);
    if (!ok) return fail(K2G_EIO, std::string("error reading ") + opts_k2d);
  }
  const bool dna = ro.dna != 0;
  const uint32_t bits = dna ? 2 : 4;
  if (ro.l == 0 || ro.l > (uint64_t) (63 / bits))
    return fail(K2G_ESOFTWARE, std::string("l exceeds size limits for minimizer ") +
                                   (dna ? "nucleotide" : "protein") + " scanner");
  if (ro.k < ro.l) return fail(K2G_EDATA, "opts.k2d: k < l");
  if (ro.k - ro.l > 63) return fail(K2G_EDATA, "opts.k2d: k - l > 63 is not supported by the GPU scan");
  if (ro.revcom > 1 || ro.revcom < 0) return fail(K2G_EDATA, "opts.k2d: unknown revcom version");

  // taxo.k2d
  std::vector<uint32_t> parent;
  std::vector<uint64_t> ext;
  {
    FILE *f = fopen(taxo_k2d, "rb");
    if (!f) return fail(K2G_EOS, std::string("unable to open ") + taxo_k2d);
    char magic[8];
    uint64_t hdr[3];
    if (fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "K2TAXDAT", 8) != 0) {
      fclose(f);
      return fail(K2G_EDATA, std::string("malformed taxonomy file ") + taxo_k2d);
    }
    if (fread(hdr, 8, 3, f) != 3 || hdr[0] == 0 || hdr[0] >= (1ULL << 31)) {
      fclose(f);
      return fail(K2G_EDATA, std::string("read exhausted taxonomy information in ") + taxo_k2d);
    }
    const uint64_t nodes = hdr[0];
    std::vector<uint64_t> raw(nodes * 7);
    size_t got = fread(raw.data(), 56, nodes, f);
    fclose(f);
    if (got != nodes)
      return fail(K2G_EDATA, std::string("read exhausted taxonomy information in ") + taxo_k2d);
    parent.resize(nodes);
    ext.resize(nodes);
    for (uint64_t i = 0; i < nodes; i++) {
      const uint64_t pid = raw[i * 7 + 0];
      // IsAAncestorOfB / LowestCommonAncestor rely on parent < child
      // (src/taxonomy.cc:249-272); anything else could walk forever.
      if (i > 0 && pid >= i)
        return fail(K2G_EDATA, "taxonomy node " + std::to_string(i) + " has parent id " +
                                   std::to_string(pid) + " (parents must precede children)");
      parent[i] = (uint32_t) (i == 0 ? 0 : pid);
      ext[i] = raw[i * 7 + 5];
    }
  }

  // hash.k2d header
  int fd = open(hash_k2d, O_RDONLY);
  if (fd < 0) return fail(K2G_EOS, std::string("unable to open ") + hash_k2d);
  uint64_t hh[4];
  if (!read_fully(fd, hh, sizeof(hh), 0)) {
    close(fd);
    return fail(K2G_EDATA, std::string("unexpected end of file reading ") + hash_k2d);
  }
  const uint64_t capacity = hh[0], kb = hh[2], vb = hh[3];
  if (kb + vb == 40) {
    close(fd);
    return fail(K2G_EDATA, "40-bit compact hash cells are not supported");
  }
  if (kb + vb != 32 || kb == 0 || vb == 0 || vb > 31 || capacity == 0) {
    close(fd);
    return fail(K2G_EDATA, "Unable to determine width of compact hash cell");
  }
  struct stat sb;
  fstat(fd, &sb);
  if ((uint64_t) sb.st_size < 32 + capacity * 4) {
    close(fd);
    return fail(K2G_EDATA, std::string("unexpected end of file reading ") + hash_k2d);
  }

  auto *db = new k2g_db();
  hipError_t e = hipGetDevice(&db->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&db->num_cus, hipDeviceAttributeMultiprocessorCount, db->device);
  if (e != hipSuccess) {
    close(fd);
    delete db;
    return fail(K2G_ESOFTWARE, std::string("no HIP device: ") + hipGetErrorString(e));
  }
  // at least 4 cells of padding: a probe's 16-byte read from its last cell
  const uint64_t padded = (capacity + 4 + 63) / 64 * 64;
  e = hipMalloc((void **) &db->d_cells, padded * 4);
  if (e == hipSuccess) e = hipMemset(db->d_cells + capacity, 0, (padded - capacity) * 4);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_parent, parent.size() * 4);
  if (e == hipSuccess)
    e = hipMemcpy(db->d_parent, parent.data(), parent.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    std::vector<uint32_t> span(2 * parent.size());
    taxonomy_spans(parent.data(), parent.size(), span.data());
    e = hipMalloc((void **) &db->d_span, span.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(db->d_span, span.data(), span.size() * 4, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    close(fd);
    k2g_db_close(db);
    return fail(K2G_EOS, std::string("unable to allocate device memory for the table: ") +
                             hipGetErrorString(e));
  }
  // stream the cell array through a pinned staging buffer
  {
    const size_t chunk = 64ull << 20;
    char *stage[2] = {nullptr, nullptr};
    hipStream_t s;
    e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipHostMalloc((void **) &stage[0], chunk, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **) &stage[1], chunk, hipHostMallocDefault);
    hipEvent_t done[2] = {nullptr, nullptr};
    if (e == hipSuccess) e = hipEventCreate(&done[0]);
    if (e == hipSuccess) e = hipEventCreate(&done[1]);
    // record both once so every later wait is on an event that was recorded
    if (e == hipSuccess) e = hipEventRecord(done[0], s);
    if (e == hipSuccess) e = hipEventRecord(done[1], s);
    const uint64_t bytes = capacity * 4;
    bool ok = e == hipSuccess;
    int which = 0;
    for (uint64_t pos = 0; ok && pos < bytes; pos += chunk, which ^= 1) {
      const size_t n = (size_t) std::min<uint64_t>(chunk, bytes - pos);
      if (hipEventSynchronize(done[which]) != hipSuccess) {
        ok = false;
        break;
      }
      if (!read_fully(fd, stage[which], n, (off_t) (32 + pos))) {
        ok = false;
        break;
      }
      if (hipMemcpyAsync((char *) db->d_cells + pos, stage[which], n, hipMemcpyHostToDevice, s) !=
              hipSuccess ||
          hipEventRecord(done[which], s) != hipSuccess)
        ok = false;
    }
    (void) hipStreamSynchronize(s);
    if (done[0]) (void) hipEventDestroy(done[0]);
    if (done[1]) (void) hipEventDestroy(done[1]);
    if (stage[0]) (void) hipHostFree(stage[0]);
    if (stage[1]) (void) hipHostFree(stage[1]);
    (void) hipStreamDestroy(s);
    close(fd);
    if (!ok) {
      k2g_db_close(db);
      return fail(K2G_EIO, std::string("error reading ") + hash_k2d);
    }
  }

  DbView &v = db->view;
  v.cells = db->d_cells;
  v.capacity = capacity;
  v.barrett = capacity > 1 ? (uint64_t) (((unsigned __int128) 1 << 64) / capacity) : 0;
  v.key_bits = (uint32_t) kb;
  v.value_bits = (uint32_t) vb;
  v.parent = db->d_parent;
  v.span = db->d_span;
  v.node_count = (uint32_t) parent.size();
  v.k = (uint32_t) ro.k;
  v.l = (uint32_t) ro.l;
  v.dna = dna ? 1 : 0;
  v.revcom_version = (uint32_t) ro.revcom;
  v.lmer_mask = (1ULL << (ro.l * bits)) - 1;
  v.spaced_mask = ro.spaced;
  v.toggle_mask = ro.toggle & v.lmer_mask;
  v.min_hash = ro.min_hash;

  // every stored value must be a node of the taxonomy; an empty cell and at
  // least 64 cells allow the lean probe, whose padding copies the first cells
  {
    unsigned int *d_max = nullptr, h_max[2] = {0, 0};
    e = hipMalloc((void **) &d_max, sizeof(h_max));
    if (e == hipSuccess) e = hipMemset(d_max, 0, sizeof(h_max));
    if (e == hipSuccess) {
      const unsigned grid = (unsigned) std::min<uint64_t>((capacity + 255) / 256, 65536);
      hipLaunchKernelGGL(k_max_value, dim3(grid), dim3(256), 0, 0, v, d_max);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h_max, d_max, sizeof(h_max), hipMemcpyDeviceToHost);
    if (d_max) (void) hipFree(d_max);
    if (e == hipSuccess && h_max[1] && capacity >= 64) {
      if (padded > capacity) {
        hipLaunchKernelGGL(k_pad_copy, dim3(1), dim3(128), 0, 0, db->d_cells, capacity,
                           (uint32_t) (padded - capacity));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
      }
      v.lean_probe = 1;
    }
    if (e != hipSuccess) {
      k2g_db_close(db);
      return fail(K2G_ESOFTWARE, std::string("table check failed: ") + hipGetErrorString(e));
    }
    if (h_max[0] >= v.node_count) {
      k2g_db_close(db);
      return fail(K2G_EDATA, "hash table holds taxid " + std::to_string(h_max[0]) +
                                 " but the taxonomy has only " + std::to_string(v.node_count) +
                                 " nodes");
    }
  }

  db->external_ids = std::move(ext);
  {
    std::string err;
    if (!db->tax.load(taxo_k2d, err)) {
      k2g_db_close(db);
      return fail(K2G_EDATA, err);
    }
  }
  k2g_db_info &inf = db->info;
  inf.capacity = capacity;
  inf.size = hh[1];
  inf.key_bits = kb;
  inf.value_bits = vb;
  inf.node_count = parent.size();
  inf.k = ro.k;
  inf.l = ro.l;
  inf.spaced_seed_mask = ro.spaced;
  inf.toggle_mask = ro.toggle;
  inf.minimum_acceptable_hash_value = ro.min_hash;
  inf.dna_db = dna ? 1 : 0;
  inf.revcom_version = ro.revcom;
  inf.device_bytes = padded * 4 + parent.size() * 12;
  *out = db;
  return K2G_OK;
}

extern "C" int k2g_db_close(k2g_db *db) {
  if (!db) return K2G_OK;
  if (db->d_cells) (void) hipFree(db->d_cells);
  if (db->d_parent) (void) hipFree(db->d_parent);
  if (db->d_span) (void) hipFree(db->d_span);
  delete db;
  return K2G_OK;
}

extern "C" int k2g_db_get_info(const k2g_db *db, k2g_db_info *info) {
  if (!db || !info) return fail(K2G_EUSAGE, "null argument");
  *info = db->info;
  return K2G_OK;
}

extern "C" int k2g_db_external_ids(const k2g_db *db, uint64_t *out, uint64_t n) {
  if (!db || !out) return fail(K2G_EUSAGE, "null argument");
  if (n < db->external_ids.size()) return fail(K2G_EUSAGE, "output too small for node_count ids");
  std::memcpy(out, db->external_ids.data(), db->external_ids.size() * 8);
  return K2G_OK;
}

// ----------------------------------------------------------------------
extern "C" int k2g_get_batch_device(k2g_db *db, const uint64_t *d_keys, uint32_t *d_out, size_t n,
                                    void *hip_stream) {
  if (!db) return fail(K2G_EUSAGE, "null db");
  if (n == 0) return K2G_OK;
  if (!d_keys || !d_out) return fail(K2G_EUSAGE, "null array");
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_get_batch, dim3((unsigned) blocks), dim3(256), 0,
                     (hipStream_t) hip_stream, db->view, d_keys, d_out, (uint64_t) n);
  HIP_TRY(hipGetLastError());
  return K2G_OK;
}

extern "C" int k2g_get_batch(k2g_db *db, const uint64_t *keys, uint32_t *out, size_t n) {
  if (!db) return fail(K2G_EUSAGE, "null db");
  if (n == 0) return K2G_OK;
  if (!keys || !out) return fail(K2G_EUSAGE, "null array");
  uint64_t *dk = nullptr;
  uint32_t *dv = nullptr;
  HIP_TRY(hipMalloc((void **) &dk, n * 8));
  hipError_t e = hipMalloc((void **) &dv, n * 4);
  if (e == hipSuccess) e = hipMemcpy(dk, keys, n * 8, hipMemcpyHostToDevice);
  int rc = K2G_OK;
  if (e == hipSuccess) {
    rc = k2g_get_batch_device(db, dk, dv, n, nullptr);
    if (rc == K2G_OK) e = hipMemcpy(out, dv, n * 4, hipMemcpyDeviceToHost);
  }
  (void) hipFree(dk);
  if (dv) (void) hipFree(dv);
  if (e != hipSuccess) return fail(K2G_ESOFTWARE, std::string("get_batch: ") + hipGetErrorString(e));
  return rc;
}

// ----------------------------------------------------------------------
// batch
namespace k2g {

// Element count e and reserved hitlist space of every fragment; a fragment
// that needs more than its reserved space is classified again into e.
// Element bound, reserved run space and sort key per fragment, and the batch
// statistics at stats (zeroed by the caller): [0] set when a mate has 2^31
// or more bases (the kernels count positions within a read in 32 bits);
// u64 at [2]: the largest element count, at [4]: their sum (want_lengths;
// launched with 256 threads per block).
__global__ void k_bounds(const uint64_t *off, const uint8_t *paired, uint32_t paired_all,
                         uint64_t n, uint32_t k, uint32_t dna, uint32_t pct, uint64_t *bounds,
                         uint64_t *caps, uint32_t *keys, uint32_t *stats, bool want_lengths) {
  // stats: [0] a read of 2^31 or more bases; [2..3] largest, [4..5] total
  // element count; [6..7] elements and [8..9] number of the fragments with
  // more than 2 * kSegElements (split candidates)
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t e = 0;
  if (i < n) {
    const bool pr = paired ? paired[i] != 0 : paired_all != 0;
    const uint64_t l1 = off[2 * i + 1] - off[2 * i], l2 = pr ? off[2 * i + 2] - off[2 * i + 1] : 0;
    if ((l1 | l2) >> 31) atomicOr(stats, 1u);
    e = fragment_elements(l1, l2, pr, k, dna != 0);
    bounds[i] = e;
    caps[i] = run_space(e, pct);
    if (keys) keys[i] = (uint32_t) (e < 0xFFFFFFFFull ? e : 0xFFFFFFFFull);
  } else if (i == n) {
    bounds[n] = 0;
    caps[n] = 0;
  }
  if (!want_lengths) return;
  // one atomic each per block for the batch's statistics
  __shared__ uint64_t bmax[256 / kWave], bsum[256 / kWave], bls[256 / kWave], blc[256 / kWave];
  const bool lng = e > 2ull * kSegElements;
  uint64_t mx = e, sum = e, ls = lng ? e : 0, lc = lng ? 1 : 0;
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t om = (uint64_t) __shfl_xor((unsigned long long) mx, d);
    sum += (uint64_t) __shfl_xor((unsigned long long) sum, d);
    ls += (uint64_t) __shfl_xor((unsigned long long) ls, d);
    lc += (uint64_t) __shfl_xor((unsigned long long) lc, d);
    mx = om > mx ? om : mx;
  }
  const unsigned wv = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    bmax[wv] = mx;
    bsum[wv] = sum;
    bls[wv] = ls;
    blc[wv] = lc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned j = 1; j < blockDim.x / kWave; j++) {
      mx = bmax[j] > mx ? bmax[j] : mx;
      sum += bsum[j];
      ls += bls[j];
      lc += blc[j];
    }
    if (lc) {
      atomicAdd(reinterpret_cast<unsigned long long *>(stats + 6), (unsigned long long) ls);
      atomicAdd(reinterpret_cast<unsigned long long *>(stats + 8), (unsigned long long) lc);
    }
    if (sum) {
      atomicMax(reinterpret_cast<unsigned long long *>(stats + 2), (unsigned long long) mx);
      atomicAdd(reinterpret_cast<unsigned long long *>(stats + 4), (unsigned long long) sum);
    }
  }
}

// Work units over the longest-first order: position j starts a unit when its
// element prefix enters a new window of T elements.  T is the requested unit
// size, lowered for small batches so that every resident wave gets about four
// units (prefix[n] is the batch's total).
__global__ void k_unit_flags(const uint64_t *prefix, uint64_t n, uint64_t T_max, uint64_t waves,
                             uint8_t *flag) {
  const uint64_t j = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t T = prefix[n] / (4 * waves);
  T = T < 1 ? 1 : (T > T_max ? T_max : T);
  flag[j] = (j == 0 || prefix[j] / T != prefix[j - 1] / T) ? 1 : 0;
}

__global__ void k_unit_end(uint32_t *unit_start, const uint32_t *n_units, uint64_t n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) unit_start[*n_units] = (uint32_t) n;
}

__global__ void k_iota(uint32_t *v, uint64_t n) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t) i;
}

// Fragments to classify again: count[0] dense-table list (status bit 0),
// count[1] all fragments that ran out of hitlist space (bit 1), count[2]
// those of them whose hit table did not overflow (status == 2).
__global__ void k_collect_redo(const uint32_t *status, uint64_t n, uint32_t *dense,
                               uint32_t *space, uint32_t *space_only, uint32_t *count) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t st = status[i];
  if (st & 1u) dense[atomicAdd(&count[0], 1u)] = (uint32_t) i;
  if (st & 2u) space[atomicAdd(&count[1], 1u)] = (uint32_t) i;
  if (st == 2u) space_only[atomicAdd(&count[2], 1u)] = (uint32_t) i;
}

__global__ void k_gather_bounds(const uint32_t *list, uint32_t m, const uint64_t *bounds,
                                uint64_t *out) {
  const uint64_t j = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = bounds[list[j]];
  if (j == m) out[m] = 0;
}

// Moves the listed fragments' runs to exact-size space after `base`.
__global__ void k_repoint(const uint32_t *list, uint32_t m, const uint64_t *tail_off,
                          uint64_t base, const uint64_t *bounds, uint64_t *run_off,
                          uint64_t *run_cap) {
  const uint64_t j = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t f = list[j];
  run_off[f] = base + tail_off[j];
  run_cap[f] = bounds[f];
}

// One digest per fragment over its call, counters and runs, in order.
__global__ void k_digest(const uint32_t *call, const uint32_t *groups, const uint32_t *total,
                         const uint32_t *n_runs, const uint64_t *run_off, const uint32_t *runs,
                         uint64_t n, uint64_t *out) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t h = 0x6A09E667F3BCC909ull;
  auto mix = [&h](uint64_t w) { h = fmix64(h ^ (w + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2))); };
  mix(call[i]);
  mix(((uint64_t) groups[i] << 32) | total[i]);
  const uint32_t m = n_runs[i];
  mix(m);
  const uint32_t *r = runs + 2 * run_off[i];
  for (uint32_t j = 0; j < m; j++) mix(((uint64_t) r[2 * j] << 32) | r[2 * j + 1]);
  out[i] = h;
}

// Random-gather ceiling: each lane makes `per_lane` independent 16-byte loads
// at uniformly random 128-byte lines of the cell array.
__global__ __launch_bounds__(256) void k_gather_ceiling(const uint32_t *cells, uint64_t n_lines,
                                                        uint32_t per_lane, uint64_t seed,
                                                        uint32_t *sink) {
  const uint64_t t = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t acc = 0;
  for (uint32_t j = 0; j < per_lane; j += 8) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint64_t h = fmix64(seed + (t * per_lane + j + u) * 0x9E3779B97F4A7C15ull);
      const uint64_t line = __umul64hi(h, n_lines);
      v[u] = *reinterpret_cast<const uint4 *>(cells + line * 32);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  if (acc == 0x5EEDF00Du) sink[t & 63] = acc;
}

// Copies each fragment's runs (2 words each) into a dense array.
__global__ __launch_bounds__(64) void k_compact(const uint64_t *src_off, const uint32_t *counts,
                                                const uint64_t *dst_off, const uint32_t *src,
                                                uint32_t *dst, uint64_t n) {
  for (uint64_t f = blockIdx.x; f < n; f += gridDim.x) {
    const uint64_t s = src_off[f], d = dst_off[f];
    const uint32_t c = counts[f];
    for (uint32_t j = threadIdx.x; j < 2 * c; j += 64) dst[2 * d + j] = src[2 * s + j];
  }
}

// v[0 .. n) = x (a fill that, unlike a memset node, is safe to record in a
// graph and replay).
__global__ void k_fill_u32(uint32_t *v, uint32_t n, uint32_t x) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = x;
}

// *out |= every status word (one atomic per wave).
__global__ __launch_bounds__(256) void k_status_or(const uint32_t *status, uint64_t n, unsigned int *out) {
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
    acc |= status[i];
  for (int d = 1; d < kWave; d <<= 1) acc |= (uint32_t) __shfl_xor((unsigned int) acc, d);
  if ((threadIdx.x & (kWave - 1)) == 0 && acc) atomicOr(out, acc);
}

__global__ void k_widen(const uint32_t *in, uint64_t *out, uint64_t n) {
  const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
  if (i == n) out[n] = 0;
}

}  // namespace k2g

extern "C" int k2g_batch_create(k2g_db *db, k2g_batch **out) {
  if (!db || !out) return fail(K2G_EUSAGE, "null argument");
  auto *b = new k2g_batch();
  b->db = db;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    k2g_batch_destroy(b);
    return fail(K2G_ESOFTWARE, std::string("batch setup: ") + hipGetErrorString(e));
  }
  const int rc = k2g::persistent_grid(b, db->num_cus);
  if (rc) {
    k2g_batch_destroy(b);
    return rc;
  }
  *out = b;
  return K2G_OK;
}

// Resident one-wave workgroups of each persistent kernel variant on the
// current device (at least one per cursor).
int k2g::persistent_grid(k2g_batch *b, int num_cus) {
  for (int vi = 0; vi < 3; vi++) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kVariants[vi].persistent, 64, 0));
    b->persistent_blocks[vi] = std::max(8, std::max(1, per_cu) * std::max(1, num_cus));
  }
  return K2G_OK;
}

extern "C" int k2g_batch_destroy(k2g_batch *b) {
  if (!b) return K2G_OK;
  if (b->stream) (void) hipStreamSynchronize(b->stream);
  b->for_each_device_buf([](auto &buf) { buf.release(); });
  b->host.reset();
  if (b->ev0) (void) hipEventDestroy(b->ev0);
  if (b->ev1) (void) hipEventDestroy(b->ev1);
  if (b->stream) (void) hipStreamDestroy(b->stream);
  delete b;
  return K2G_OK;
}

extern "C" void *k2g_batch_stream(k2g_batch *b) { return b ? (void *) b->stream : nullptr; }

namespace k2g {
// Largest hipcub scratch a classify call over n fragments asks for (scans of
// n + 1 words; the longest-first sort of n keys and its unit selection).
static size_t temp_bytes_for(uint64_t n, bool sort) {
  size_t a = 0, c = 0;
  uint64_t *u = nullptr;
  (void) hipcub::DeviceScan::ExclusiveSum(nullptr, a, u, u, n + 1, (hipStream_t) 0);
  size_t d = 0;
  if (sort) {
    uint32_t *k = nullptr;
    uint8_t *fl = nullptr;
    (void) hipcub::DeviceRadixSort::SortPairsDescending(nullptr, c, k, k, k, k, (int) n, 0, 32,
                                                        (hipStream_t) 0);
    (void) hipcub::DeviceSelect::Flagged(nullptr, d, hipcub::CountingInputIterator<uint32_t>(0), fl,
                                         k, k, (int) n, (hipStream_t) 0);
  }
  return std::max(a, std::max(c, d));
}
}  // namespace k2g

static int check_params(const k2g_classify_params *p) {
  if (!p) return fail(K2G_EUSAGE, "null params");
  if (!(p->confidence >= 0 && p->confidence <= 1))
    return fail(K2G_EUSAGE, "confidence threshold must be in [0, 1]");
  if (p->mode < 0 || p->mode > 2)
    return fail(K2G_EUSAGE, "mode must be 0 (plain), 1 (persistent) or 2 (chosen per batch)");
  const int s = p->table_slots;
  if (s != 0 && (s < 2 || s > 256 || (s & (s - 1)) != 0))
    return fail(K2G_EUSAGE, "table_slots must be a power of two in [2, 256]");
  if (p->run_cap_pct < 0 || p->run_cap_pct > 100)
    return fail(K2G_EUSAGE, "run_cap_pct must be in [0, 100]");
  if (p->want_hll < 0 || p->want_hll > 2) return fail(K2G_EUSAGE, "want_hll must be 0, 1 or 2");
  return K2G_OK;
}

extern "C" int k2g_classify_device(k2g_batch *b, const k2g_classify_params *p, const char *d_seq,
                                   const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                                   const uint32_t *d_order) {
  return k2g::classify_launch(b, p, d_seq, d_off, d_paired, n, d_order);
}

int k2g::classify_launch(k2g_batch *b, const k2g_classify_params *p, const char *d_seq,
                         const uint64_t *d_off, const uint8_t *d_paired, uint64_t n,
                         const uint32_t *d_order) {
  if (!b) return fail(K2G_EUSAGE, "null batch");
  const bool sketch_blocks = b->sk_pending;  // this call's input blocks were uploaded
  b->sk_pending = false;
  int rc = check_params(p);
  if (rc) return rc;
  if (p->want_hll == 2 && n && !sketch_blocks)
    return fail(K2G_EUSAGE, "want_hll 2 needs host fragments (k2g_classify or the server): the input blocks "
                            "come from their offsets");
  k2g_db *db = b->db;
  const DbView &v = db->view;
  b->n = n;
  b->params = *p;
  b->redone = 0;
  b->space_redone = 0;
  b->kernel_ms = 0;
  b->split_last = false;
  if (n == 0) return K2G_OK;
  if (n >= (1ULL << 32)) return fail(K2G_EUSAGE, "batch too large (fragments must fit 32 bits)");
  if (!d_seq || !d_off) return fail(K2G_EUSAGE, "null fragment arrays");
  hipStream_t s = b->stream;

  HIP_TRY(b->call.reserve(n));
  HIP_TRY(b->groups.reserve(n));
  HIP_TRY(b->total.reserve(n));
  HIP_TRY(b->n_runs.reserve(n));
  HIP_TRY(b->status.reserve(n));
  HIP_TRY(b->n_recs.reserve(n));
  HIP_TRY(b->run_off.reserve(n + 1));
  HIP_TRY(b->run_cap.reserve(n + 1));
  HIP_TRY(b->bounds.reserve(n + 1));
  HIP_TRY(b->sort_keys.reserve(n));
  HIP_TRY(b->redo_count.reserve(12));  // three redo counts, then k_bounds's statistics
  HIP_TRY(b->cursor.reserve(kCursorWords));
  // in a bounded region, the largest scratch any step needs is taken once
  const bool may_sort = p->mode != 0 && !d_order;
  if (b->temp.region) HIP_TRY(b->temp.reserve(temp_bytes_for(n, may_sort)));

  // element bounds -> reserved run space -> run offsets
  const uint32_t pct = p->run_cap_pct ? (uint32_t) p->run_cap_pct : 25u;
  const unsigned g1 = (unsigned) ((n + 1 + 255) / 256);
  uint32_t *const stats = b->redo_count.p + 2;  // 8-byte aligned (see k_bounds)
  HIP_TRY(hipMemsetAsync(stats, 0, 40, s));
  hipLaunchKernelGGL(k_bounds, dim3(g1), dim3(256), 0, s, d_off, d_paired,
                     (uint32_t) (p->paired != 0), n, v.k, v.dna, pct, b->bounds.p, b->run_cap.p,
                     may_sort ? b->sort_keys.p : nullptr, stats, p->mode != 0);
  HIP_TRY(hipGetLastError());
  size_t tbytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, b->run_cap.p, b->run_off.p, n + 1, s));
  HIP_TRY(b->temp.reserve(tbytes));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tbytes, b->run_cap.p, b->run_off.p, n + 1, s));
  uint64_t total_space = 0, max_e = 0, sum_e = 0, split_sum = 0, split_cnt = 0;
  if (b->plan.valid) {  // the caller's host-side statistics: nothing to read back
    b->plan.valid = false;
    if (b->plan.too_long) return fail(K2G_EUSAGE, "a read has 2^31 or more bases");
    total_space = b->plan.space;
    max_e = b->plan.max_e;
    sum_e = b->plan.sum_e;
    split_sum = b->plan.split_sum;
    split_cnt = b->plan.split_cnt;
  } else {
    uint32_t st[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(&total_space, b->run_off.p + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(st, stats, 40, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (st[0]) return fail(K2G_EUSAGE, "a read has 2^31 or more bases");
    max_e = st[2] | ((uint64_t) st[3] << 32);
    sum_e = st[4] | ((uint64_t) st[5] << 32);
    split_sum = st[6] | ((uint64_t) st[7] << 32);
    split_cnt = st[8] | ((uint64_t) st[9] << 32);
  }
  // mode 2: the persistent kernel pays for its sort and work units only when
  // lengths are uneven (the longest fragment above four times the mean);
  // batches of similar reads run one plain launch
  const bool persistent = p->mode == 1 || (p->mode == 2 && (d_order || max_e * n > 4 * sum_e));
  b->ran_persistent = persistent;
  if (b->fixed_outputs && 2 * total_space > b->runs.cap)
    return fail(K2G_EUSAGE, "run capacity " + std::to_string(b->runs.cap / 2) +
                                " is below the batch's element count " + std::to_string(total_space));
  HIP_TRY(b->runs.reserve(2 * std::max<uint64_t>(total_space, 1)));
  if (p->want_hll) HIP_TRY(b->recs.reserve(2 * std::max<uint64_t>(total_space, 1)));
  if (p->want_kmer_counts) {
    HIP_TRY(b->kmer_counts.reserve(v.node_count));
    HIP_TRY(hipMemsetAsync(b->kmer_counts.p, 0, v.node_count * 8, s));
  }

  // persistent order: longest fragments first
  const uint32_t *order = d_order;
  const uint32_t *units = nullptr;
  if (persistent && !order) {
    HIP_TRY(b->order.reserve(n));
    HIP_TRY(b->sort_vals.reserve(n));
    HIP_TRY(b->sort_keys_out.reserve(n));
    const unsigned gi = (unsigned) ((n + 255) / 256);
    hipLaunchKernelGGL(k_iota, dim3(gi), dim3(256), 0, s, b->sort_vals.p, n);
    size_t sbytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sbytes, b->sort_keys.p,
                                                         b->sort_keys_out.p, b->sort_vals.p,
                                                         b->order.p, (int) n, 0, 32, s));
    HIP_TRY(b->temp.reserve(sbytes));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(b->temp.p, sbytes, b->sort_keys.p,
                                                         b->sort_keys_out.p, b->sort_vals.p,
                                                         b->order.p, (int) n, 0, 32, s));
    order = b->order.p;
    // work units: windows of about T elements of the sorted order
    const uint64_t T = p->unit_elements > 0 ? (uint64_t) p->unit_elements : 4096u;
    HIP_TRY(b->unit_pos.reserve(n + 1));
    HIP_TRY(b->unit_scan.reserve(n + 1));
    HIP_TRY(b->unit_flag.reserve(n));
    HIP_TRY(b->unit_start.reserve(n + 1));
    HIP_TRY(b->n_units.reserve(1));
    hipLaunchKernelGGL(k_widen, dim3((unsigned) ((n + 1 + 255) / 256)), dim3(256), 0, s,
                       b->sort_keys_out.p, b->unit_pos.p, n);
    size_t ub = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, ub, b->unit_pos.p, b->unit_scan.p, n + 1, s));
    HIP_TRY(b->temp.reserve(ub));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, ub, b->unit_pos.p, b->unit_scan.p, n + 1, s));
    const int vi_units = variant_index(v.dna != 0, p->want_hll || p->want_kmer_counts ||
                                                         p->want_probe_stats || p->quick);
    hipLaunchKernelGGL(k_unit_flags, dim3(gi), dim3(256), 0, s, b->unit_scan.p, n, T,
                       (uint64_t) b->persistent_blocks[vi_units], b->unit_flag.p);
    size_t fb = 0;
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, fb, hipcub::CountingInputIterator<uint32_t>(0),
                                          b->unit_flag.p, b->unit_start.p, b->n_units.p, (int) n, s));
    HIP_TRY(b->temp.reserve(fb));
    HIP_TRY(hipcub::DeviceSelect::Flagged(b->temp.p, fb, hipcub::CountingInputIterator<uint32_t>(0),
                                          b->unit_flag.p, b->unit_start.p, b->n_units.p, (int) n, s));
    hipLaunchKernelGGL(k_unit_end, dim3(1), dim3(64), 0, s, b->unit_start.p, b->n_units.p, n);
    HIP_TRY(hipGetLastError());
    units = b->unit_start.p;
  }

  ClassifyParams cp{};
  cp.confidence = p->confidence;
  cp.min_hit_groups = p->min_hit_groups;
  cp.quick = p->quick ? 1u : 0u;
  cp.tab_cap = p->table_slots ? (uint32_t) p->table_slots : 256u;
  cp.batch_chunks = 8;
  FragBatch in{d_seq, d_off,  d_paired, (uint32_t) (p->paired != 0), n, order, units,
               units ? b->n_units.p : nullptr};
  if (p->want_probe_stats) {
    HIP_TRY(b->lookups.reserve(n));
    HIP_TRY(b->lines.reserve(n));
  }
  FragOut out{b->call.p,    b->groups.p,  b->total.p,   b->n_runs.p,
              b->status.p,  b->n_recs.p,  b->run_off.p, b->run_cap.p,
              b->runs.p,    p->want_hll ? b->recs.p : nullptr,
              p->want_kmer_counts ? b->kmer_counts.p : nullptr,
              p->want_probe_stats ? b->lookups.p : nullptr,
              p->want_probe_stats ? b->lines.p : nullptr};

  const int vi = variant_index(v.dna != 0, p->want_hll || p->want_kmer_counts ||
                                               p->want_probe_stats || p->quick);
  const ClassifyKernels &K = kVariants[vi];
  if (b->ev0) HIP_TRY(hipEventRecord(b->ev0, s));
  // a batch whose longest fragment outweighs a wave's share splits its long
  // fragments over waves (segments before, stitch after the persistent kernel)
  const bool split = persistent && !d_order && !b->fixed_outputs &&
                     split_wanted(p, v.dna != 0, persistent, max_e, sum_e, (uint64_t) b->persistent_blocks[vi]);
  if (split) {
    rc = split_launch(b, p, cp, in, out, split_sum, split_cnt, 1);
    if (rc) return rc;
  }
  if (!persistent) {
    hipLaunchKernelGGL(K.plain, dim3((unsigned) n), dim3(64), 0, s, v, cp, in, out);
  } else {
    HIP_TRY(hipMemsetAsync(b->cursor.p, 0, kCursorWords * 8, s));
    const unsigned grid = (unsigned) std::min<uint64_t>(b->persistent_blocks[vi], n);
    hipLaunchKernelGGL(K.persistent, dim3(grid), dim3(64), 0, s, v, cp, in, out, b->cursor.p);
  }
  HIP_TRY(hipGetLastError());
  if (split) {
    rc = split_launch(b, p, cp, in, out, split_sum, split_cnt, 2);
    if (rc) return rc;
  }
  if (b->ev1) HIP_TRY(hipEventRecord(b->ev1, s));

  // fragments to classify again: out of hitlist space, or hit table overflowed
  HIP_TRY(b->redo.reserve(n));
  HIP_TRY(b->space.reserve(n));
  HIP_TRY(b->space_only.reserve(n));
  HIP_TRY(hipMemsetAsync(b->redo_count.p, 0, 12, s));
  hipLaunchKernelGGL(k_collect_redo, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s,
                     b->status.p, n, b->redo.p, b->space.p, b->space_only.p, b->redo_count.p);
  HIP_TRY(hipGetLastError());
  // with exact hitlist space no fragment can need more room, and the dense
  // redo takes its fragment count from the device: no host round trip
  const bool on_device = pct >= 100;
  uint32_t counts[3] = {0, 0, 0};
  if (!on_device) {
    HIP_TRY(hipMemcpyAsync(counts, b->redo_count.p, 12, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  b->redo_on_device = on_device;
  const uint32_t n_redo = counts[0], n_space = counts[1], n_space_only = counts[2];
  FragOut out2 = out;
  out2.kmer_counts = nullptr;  // the first pass counted every fragment
  out2.lookups = nullptr;
  out2.lines = nullptr;
  if (n_space) {
    // exact space after the reserved space, then classify those fragments again
    HIP_TRY(b->space_bounds.reserve(n_space + 1));
    HIP_TRY(b->space_off.reserve(n_space + 1));
    const unsigned gs = (unsigned) ((n_space + 1 + 255) / 256);
    hipLaunchKernelGGL(k_gather_bounds, dim3(gs), dim3(256), 0, s, b->space.p, n_space,
                       b->bounds.p, b->space_bounds.p);
    HIP_TRY(hipGetLastError());
    size_t sb = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, b->space_bounds.p, b->space_off.p,
                                             n_space + 1, s));
    HIP_TRY(b->temp.reserve(sb));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, sb, b->space_bounds.p, b->space_off.p,
                                             n_space + 1, s));
    uint64_t tail = 0;
    HIP_TRY(hipMemcpyAsync(&tail, b->space_off.p + n_space, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(grow_keep(b->runs, 2 * (total_space + tail), 2 * total_space, s));
    if (p->want_hll) HIP_TRY(grow_keep(b->recs, 2 * (total_space + tail), 2 * total_space, s));
    hipLaunchKernelGGL(k_repoint, dim3(gs), dim3(256), 0, s, b->space.p, n_space, b->space_off.p,
                       total_space, b->bounds.p, b->run_off.p, b->run_cap.p);
    HIP_TRY(hipGetLastError());
    out2.runs = b->runs.p;
    out2.recs = p->want_hll ? b->recs.p : nullptr;
    if (n_space_only) {
      FragBatch in2 = in;
      in2.n = n_space_only;
      in2.order = b->space_only.p;
      in2.unit_start = nullptr;
      in2.n_units = nullptr;
      in2.split = nullptr;  // classified whole
      HIP_TRY(hipMemsetAsync(b->cursor.p, 0, kCursorWords * 8, s));
      const unsigned grid = (unsigned) std::min<uint64_t>(b->persistent_blocks[vi], n_space_only);
      hipLaunchKernelGGL(K.persistent, dim3(grid), dim3(64), 0, s, v, cp, in2, out2,
                         b->cursor.p);
      HIP_TRY(hipGetLastError());
    }
    b->space_redone = n_space;
  }
  if (n_redo || on_device) {
    // hit table overflowed: classify again with a dense per-taxon table
    uint32_t slots = on_device ? dense_slots_for(n, v.node_count) : std::min<uint32_t>(n_redo, 256);
    if (b->fixed_outputs && b->seq.region) {
      // a fixed region: as many slots as its remaining room holds
      const Region &r = *b->seq.region;
      const uint64_t per_slot = 2 * region_bytes(v.node_count, 4) + kRegionAlign;
      const uint64_t room = r.size > r.used + region_bytes(slots, 4) ? r.size - r.used - region_bytes(slots, 4)
                                                                      : 0;
      slots = (uint32_t) std::min<uint64_t>(slots, room / per_slot);
      if (slots == 0)
        return fail(K2G_EUSAGE, "no room left in the fixed scratch for one dense hit table of " +
                                    std::to_string(v.node_count) + " taxa");
    }
    if (b->dense_slots < slots) {  // each slot's table is cleared by the kernel that uses it
      HIP_TRY(b->dense_counts.reserve((uint64_t) slots * v.node_count));
      HIP_TRY(b->dense_lists.reserve((uint64_t) slots * v.node_count));
      HIP_TRY(b->dense_lens.reserve(slots));
      b->dense_slots = slots;
    }
    DenseScratch ds{b->dense_counts.p, b->dense_lists.p, b->dense_lens.p, slots};
    hipLaunchKernelGGL(K.dense, dim3(slots), dim3(64), 0, s, v, cp, in, out2, b->redo.p,
                       b->redo_count.p, ds);
    HIP_TRY(hipGetLastError());
    b->redone = n_redo;
  }
  return K2G_OK;
}

namespace k2g {
// The dense-redo count of a call whose redo was sized on the device.
static int settle_redo_count(k2g_batch *b) {
  if (!b->redo_on_device || !b->n) return K2G_OK;
  uint32_t c = 0;
  HIP_TRY(hipMemcpyAsync(&c, b->redo_count.p, 4, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->redone = c;
  b->redo_on_device = false;
  return K2G_OK;
}
}  // namespace k2g

extern "C" int k2g_batch_fetch_counts(k2g_batch *b, uint32_t *call, uint32_t *hit_groups,
                                      uint32_t *total_kmers, uint32_t *n_runs, uint32_t *lookups,
                                      uint64_t *probe_lines, uint64_t n) {
  if (!b) return fail(K2G_EUSAGE, "null batch");
  if (n != b->n) return fail(K2G_EUSAGE, "count differs from the last batch's fragments");
  if ((lookups || probe_lines) && !b->params.want_probe_stats)
    return fail(K2G_EUSAGE, "probe statistics were not requested for the last batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (!n) return K2G_OK;
  std::vector<uint32_t> st(n);
  HIP_TRY(hipMemcpy(st.data(), b->status.p, n * 4, hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < n; i++)
    if (st[i] & 2u) return fail(K2G_ESOFTWARE, "internal error: hitlist exceeded its bound");
  if (call) HIP_TRY(hipMemcpy(call, b->call.p, n * 4, hipMemcpyDeviceToHost));
  if (hit_groups) HIP_TRY(hipMemcpy(hit_groups, b->groups.p, n * 4, hipMemcpyDeviceToHost));
  if (total_kmers) HIP_TRY(hipMemcpy(total_kmers, b->total.p, n * 4, hipMemcpyDeviceToHost));
  if (n_runs) HIP_TRY(hipMemcpy(n_runs, b->n_runs.p, n * 4, hipMemcpyDeviceToHost));
  if (lookups) HIP_TRY(hipMemcpy(lookups, b->lookups.p, n * 4, hipMemcpyDeviceToHost));
  if (probe_lines) HIP_TRY(hipMemcpy(probe_lines, b->lines.p, n * 8, hipMemcpyDeviceToHost));
  return K2G_OK;
}

extern "C" int k2g_batch_redo_counts(k2g_batch *b, uint64_t *run_space, uint64_t *dense_table) {
  if (!b) return fail(K2G_EUSAGE, "null batch");
  const int rc = k2g::settle_redo_count(b);
  if (rc) return rc;
  if (run_space) *run_space = b->space_redone;
  if (dense_table) *dense_table = b->redone;
  return K2G_OK;
}

extern "C" int k2g_batch_split_counts(k2g_batch *b, uint64_t *fragments, uint64_t *segments) {
  if (!b) return fail(K2G_EUSAGE, "null batch");
  uint32_t c[2] = {0, 0};
  if (b->split_last && b->sp_count.p) {
    HIP_TRY(hipMemcpyAsync(c, b->sp_count.p, 8, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  }
  if (fragments) *fragments = c[1];
  if (segments) *segments = c[0];
  return K2G_OK;
}

extern "C" int k2g_batch_digests(k2g_batch *b, uint64_t *out, uint64_t n) {
  if (!b || (!out && n)) return fail(K2G_EUSAGE, "null argument");
  if (n != b->n) return fail(K2G_EUSAGE, "digest count differs from the last batch's fragments");
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (!n) return K2G_OK;
  HIP_TRY(b->digest.reserve(n));
  hipLaunchKernelGGL(k_digest, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, b->stream,
                     b->call.p, b->groups.p, b->total.p, b->n_runs.p, b->run_off.p, b->runs.p, n,
                     b->digest.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, b->digest.p, n * 8, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return K2G_OK;
}

extern "C" int k2g_batch_sync(k2g_batch *b, double *kernel_ms) {
  if (!b) return fail(K2G_EUSAGE, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0;
  if (b->n) HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->kernel_ms = ms;
  if (kernel_ms) *kernel_ms = ms;
  return K2G_OK;
}

extern "C" int k2g_batch_fetch(k2g_batch *b, k2g_results *r) {
  if (!b || !r) return fail(K2G_EUSAGE, "null argument");
  const uint64_t n = b->n;
  hipStream_t s = b->stream;
  std::memset(r, 0, sizeof(*r));
  r->n = n;
  r->node_count = b->db->view.node_count;
  HostResults &h = *b->host;
  HIP_TRY(h.call.reserve(std::max<uint64_t>(n, 1)));
  HIP_TRY(h.groups.reserve(std::max<uint64_t>(n, 1)));
  HIP_TRY(h.total.reserve(std::max<uint64_t>(n, 1)));
  HIP_TRY(h.run_off.reserve(n + 1));
  HIP_TRY(h.words.reserve(8));
  h.run_off.p[0] = 0;
  const bool hll = n && b->params.want_hll, stats = n && b->params.want_probe_stats;
  const bool sketches = hll && b->params.want_hll == 2;
  if (hll) HIP_TRY(h.rec_off.reserve(n + 1));
  if (n) {
    // phase 1, no host round trip: OR of every status word, the results'
    // scalars into pinned host memory, and the scans that lay out the
    // hitlists and HLL records densely; then one wait for the totals
    HIP_TRY(b->fetch_word.reserve(1));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, b->fetch_word.p, 1u, 0u);
    hipLaunchKernelGGL(k_status_or, dim3((unsigned) std::min<uint64_t>((n + 255) / 256, 2048)), dim3(256), 0, s,
                       b->status.p, n, b->fetch_word.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h.words.p, b->fetch_word.p, 4, hipMemcpyDeviceToHost, s));
    if (b->redo_on_device) HIP_TRY(hipMemcpyAsync(h.words.p + 1, b->redo_count.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h.call.p, b->call.p, n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h.groups.p, b->groups.p, n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h.total.p, b->total.p, n * 4, hipMemcpyDeviceToHost, s));
    auto scan = [&](const uint32_t *d_counts, DevBuf<uint64_t> &off, DevBuf<uint64_t> &widened,
                    PinnedBuf<uint64_t> &h_off) -> int {
      HIP_TRY(off.reserve(n + 1));
      HIP_TRY(widened.reserve(n + 1));
      hipLaunchKernelGGL(k_widen, dim3((unsigned) ((n + 1 + 255) / 256)), dim3(256), 0, s, d_counts,
                         widened.p, n);
      size_t tb = 0;
      HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, widened.p, off.p, n + 1, s));
      HIP_TRY(b->temp.reserve(tb));
      HIP_TRY(hipcub::DeviceScan::ExclusiveSum(b->temp.p, tb, widened.p, off.p, n + 1, s));
      HIP_TRY(hipMemcpyAsync(h_off.p, off.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
      return K2G_OK;
    };
    int rc = scan(b->n_runs.p, b->comp_off, b->bounds, h.run_off);
    if (rc) return rc;
    if (hll) {
      rc = scan(b->n_recs.p, b->comp_rec_off, b->bounds, h.rec_off);  // stream order: bounds is free again
      if (rc) return rc;
    }
    if (stats) {
      HIP_TRY(h.lookups.reserve(n));
      HIP_TRY(h.lines.reserve(n));
      HIP_TRY(hipMemcpyAsync(h.lookups.p, b->lookups.p, n * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(h.lines.p, b->lines.p, n * 8, hipMemcpyDeviceToHost, s));
    }
    if (b->params.want_kmer_counts) {
      HIP_TRY(h.kmer_counts.reserve(b->db->view.node_count));
      HIP_TRY(hipMemcpyAsync(h.kmer_counts.p, b->kmer_counts.p, b->db->view.node_count * 8,
                             hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (h.words.p[0] & 2u) return fail(K2G_ESOFTWARE, "internal error: hitlist exceeded its bound");
    if (b->redo_on_device) {
      b->redone = h.words.p[1];
      b->redo_on_device = false;
    }
    // phase 2: each fragment's runs (and records) written straight into
    // pinned, mapped host memory at their dense offsets; one more wait
    auto compact = [&](const uint32_t *d_counts, const uint32_t *d_src, const DevBuf<uint64_t> &off,
                       uint64_t tot, PinnedBuf<uint32_t> &dst) -> int {
      HIP_TRY(dst.reserve(2 * std::max<uint64_t>(tot, 1)));
      uint32_t *d_dst = nullptr;
      HIP_TRY(hipHostGetDevicePointer((void **) &d_dst, dst.p, 0));
      const unsigned grid = (unsigned) std::min<uint64_t>(n, 65536);
      hipLaunchKernelGGL(k_compact, dim3(grid), dim3(64), 0, s, b->run_off.p, d_counts, off.p, d_src, d_dst, n);
      HIP_TRY(hipGetLastError());
      return K2G_OK;
    };
    rc = compact(b->n_runs.p, b->runs.p, b->comp_off, h.run_off.p[n], h.runs);
    if (rc) return rc;
    if (hll && !sketches) {
      rc = compact(b->n_recs.p, b->recs.p, b->comp_rec_off, h.rec_off.p[n], h.recs);
      if (rc) return rc;
    }
    if (sketches) {  // the records reduced to per-(block, taxon) sets on the device
      rc = sketch_fetch(b, h, h.rec_off.p[n]);
      if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s));
  }
  float ms = 0;
  if (n && b->ev0) HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->kernel_ms = ms;
  r->kernel_ms = ms;
  r->redone = b->redone;
  r->call = h.call.p;
  r->hit_groups = h.groups.p;
  r->total_kmers = h.total.p;
  r->run_off = h.run_off.p;
  r->runs = h.runs.p;
  if (hll && !sketches) {
    r->hll_off = h.rec_off.p;
    r->hll = h.recs.p;
  }
  if (sketches) {
    r->n_sketches = h.sk_words.p[0];
    r->sketch_block = h.sk_block.p;
    r->sketch_taxon = h.sk_taxon.p;
    r->sketch_dense = h.sk_dense.p;
    r->sketch_n = h.sk_n.p;
    r->sketch_off = h.sk_off.p;
    r->sketch_vals = h.sk_vals.p;
  }
  if (stats) {
    r->lookups = h.lookups.p;
    r->probe_lines = h.lines.p;
  }
  if (n && b->params.want_kmer_counts) r->kmer_counts = h.kmer_counts.p;
  return K2G_OK;
}

extern "C" int k2g_classify(k2g_batch *b, const k2g_classify_params *p, const k2g_fragments *in,
                            k2g_results *out) {
  if (!b || !in || !out) return fail(K2G_EUSAGE, "null argument");
  int rc = check_params(p);
  if (rc) return rc;
  const uint64_t n = in->n;
  if (n && (!in->seq || !in->off)) return fail(K2G_EUSAGE, "null fragment arrays");
  const uint64_t bases = n ? in->off[2 * n] : 0;
  for (uint64_t i = 0; i < 2 * n; i++)
    if (in->off[i + 1] < in->off[i]) return fail(K2G_EUSAGE, "fragment offsets must not decrease");
  hipStream_t s = b->stream;
  if (n) {
    HIP_TRY(b->seq.reserve(bases + 1));
    HIP_TRY(b->off.reserve(2 * n + 1));
    if (bases) HIP_TRY(hipMemcpyAsync(b->seq.p, in->seq, bases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->off.p, in->off, (2 * n + 1) * 8, hipMemcpyHostToDevice, s));
    if (in->paired) {
      HIP_TRY(b->paired.reserve(n));
      HIP_TRY(hipMemcpyAsync(b->paired.p, in->paired, n, hipMemcpyHostToDevice, s));
    }
    if (p->want_hll == 2) {  // the per-(block, taxon) HLL sets need the input blocks
      std::vector<uint64_t> starts;
      k2g::input_blocks(in, p->paired != 0, starts);
      rc = k2g::sketch_upload_blocks(b, starts);
      if (rc) return rc;
      b->sk_pending = true;
    }
  }
  rc = k2g_classify_device(b, p, b->seq.p, b->off.p, in->paired ? b->paired.p : nullptr, n,
                           nullptr);
  if (rc) return rc;
  return k2g_batch_fetch(b, out);
}

// ----------------------------------------------------------------------
// build
struct k2g_builder {
  BuildView view{};
  uint32_t *d_cells = nullptr, *d_parent = nullptr;
  unsigned long long *d_size = nullptr;
  unsigned int *d_error = nullptr;
  uint64_t capacity = 0, node_count = 0;
  k2g_build_params params{};
  hipStream_t stream = nullptr;
  DevBuf<char> seq;
  DevBuf<SubBlock> items;
  int num_cus = 0;
};

static int read_parents(const char *taxo_k2d, std::vector<uint32_t> &parent) {
  FILE *f = fopen(taxo_k2d, "rb");
  if (!f) return fail(K2G_EOS, std::string("unable to open ") + taxo_k2d);
  char magic[8];
  uint64_t hdr[3];
  if (fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "K2TAXDAT", 8) != 0 ||
      fread(hdr, 8, 3, f) != 3 || hdr[0] == 0 || hdr[0] >= (1ULL << 31)) {
    fclose(f);
    return fail(K2G_EDATA, std::string("malformed taxonomy file ") + taxo_k2d);
  }
  std::vector<uint64_t> raw(hdr[0] * 7);
  const size_t got = fread(raw.data(), 56, hdr[0], f);
  fclose(f);
  if (got != hdr[0]) return fail(K2G_EDATA, std::string("read exhausted taxonomy information in ") + taxo_k2d);
  parent.resize(hdr[0]);
  for (uint64_t i = 0; i < hdr[0]; i++) {
    const uint64_t pid = raw[7 * i];
    if (i > 0 && pid >= i)
      return fail(K2G_EDATA, "taxonomy node " + std::to_string(i) + " has parent id " +
                                 std::to_string(pid) + " (parents must precede children)");
    parent[i] = (uint32_t) (i ? pid : 0);
  }
  return K2G_OK;
}

extern "C" int k2g_build_create(const char *taxo_k2d, uint64_t capacity, uint64_t value_bits,
                                const k2g_build_params *p, k2g_builder **out) {
  if (!taxo_k2d || !p || !out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  if (capacity == 0) return fail(K2G_EUSAGE, "capacity must be positive integer");
  const uint32_t bits = p->dna_db ? 2 : 4;
  if (p->l == 0 || p->l > 63 / bits) return fail(K2G_EUSAGE, "l exceeds size limits for the minimizer scanner");
  if (p->k < p->l) return fail(K2G_EUSAGE, "k cannot be less than l");
  if (p->k - p->l > 63) return fail(K2G_EUSAGE, "k - l > 63 is not supported by the GPU scan");
  if (p->subblock_size == 0 || p->block_size < p->subblock_size)
    return fail(K2G_EUSAGE, "block size cannot be less than subblock size");
  std::vector<uint32_t> parent;
  int rc = read_parents(taxo_k2d, parent);
  if (rc) return rc;
  uint64_t vb = 1;
  while ((1ULL << vb) < parent.size()) vb++;
  if (value_bits) {
    if (value_bits < vb) return fail(K2G_EDATA, "more bits required for storing taxid");
    vb = value_bits;
  }
  if (vb > 31) return fail(K2G_EUSAGE, "can't have more than 31 bits of storage for taxid");
  auto *b = new k2g_builder();
  b->capacity = capacity;
  b->node_count = parent.size();
  b->params = *p;
  const uint64_t padded = (capacity + 63) / 64 * 64;
  hipError_t e = hipMalloc((void **) &b->d_cells, padded * 4);
  if (e == hipSuccess) e = hipMemset(b->d_cells, 0, padded * 4);
  if (e == hipSuccess) e = hipMalloc((void **) &b->d_parent, parent.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(b->d_parent, parent.data(), parent.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void **) &b->d_size, 8);
  if (e == hipSuccess) e = hipMemset(b->d_size, 0, 8);
  if (e == hipSuccess) e = hipMalloc((void **) &b->d_error, 4);
  if (e == hipSuccess) e = hipMemset(b->d_error, 0, 4);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  int dev = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    k2g_build_destroy(b);
    return fail(K2G_EOS, std::string("unable to allocate hash table memory: ") + hipGetErrorString(e));
  }
  BuildView &v = b->view;
  v.cells = b->d_cells;
  v.capacity = capacity;
  v.barrett = capacity > 1 ? (uint64_t) (((unsigned __int128) 1 << 64) / capacity) : 0;
  v.key_bits = (uint32_t) (32 - vb);
  v.value_bits = (uint32_t) vb;
  v.parent = b->d_parent;
  v.k = (uint32_t) p->k;
  v.l = (uint32_t) p->l;
  v.dna = p->dna_db ? 1 : 0;
  v.revcom_version = (uint32_t) p->revcom_version;
  const uint64_t lmask = (1ULL << (p->l * bits)) - 1;
  v.spaced_mask = p->spaced_seed_mask;
  v.toggle_mask = p->toggle_mask & lmask;
  v.min_clear = p->min_clear_hash_value;
  v.size = b->d_size;
  v.error = b->d_error;
  *out = b;
  return K2G_OK;
}

extern "C" int k2g_build_destroy(k2g_builder *b) {
  if (!b) return K2G_OK;
  if (b->stream) (void) hipStreamSynchronize(b->stream);
  if (b->d_cells) (void) hipFree(b->d_cells);
  if (b->d_parent) (void) hipFree(b->d_parent);
  if (b->d_size) (void) hipFree(b->d_size);
  if (b->d_error) (void) hipFree(b->d_error);
  b->seq.release();
  b->items.release();
  if (b->stream) (void) hipStreamDestroy(b->stream);
  delete b;
  return K2G_OK;
}

// Windows of the deterministic build (src/build_db.h:224-244): blocks of
// block_size, each extended by k-1, cut into subblocks of subblock_size,
// each extended by k-1 within its block.
static void build_windows(const k2g_build_params &p, const uint64_t *off, const uint32_t *taxid,
                          uint64_t n, uint64_t base, std::vector<SubBlock> &items) {
  const uint64_t k = p.k, B = p.block_size, b = p.subblock_size;
  for (uint64_t s = 0; s < n; s++) {
    const uint64_t L = off[s + 1] - off[s];
    if (taxid[s] == 0) continue;
    for (uint64_t j = 0; j < L; j += B) {
      const uint64_t bfin = std::min(j + B + k - 1, L);
      for (uint64_t i = j; i < bfin; i += b) {
        const uint64_t fin = std::min(i + b + k - 1, bfin);
        SubBlock it;
        it.start = off[s] - base + i;
        it.len = (uint32_t) (fin - i);
        it.taxid = taxid[s];
        it.first_e = (uint32_t) (i == 0 ? k - 1 : k - 2);
        it.pad = 0;
        items.push_back(it);
      }
    }
  }
}

static int build_launch(k2g_builder *b, const char *d_seq, const std::vector<SubBlock> &items) {
  if (items.empty()) return K2G_OK;
  HIP_TRY(b->items.reserve(items.size()));
  HIP_TRY(hipMemcpyAsync(b->items.p, items.data(), items.size() * sizeof(SubBlock),
                         hipMemcpyHostToDevice, b->stream));
  const uint64_t grid = std::min<uint64_t>(items.size(), (uint64_t) std::max(1, b->num_cus) * 32);
  hipLaunchKernelGGL(k_build, dim3((unsigned) grid), dim3(64), 0, b->stream, b->view, d_seq,
                     b->items.p, (uint64_t) items.size());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  unsigned int err = 0;
  HIP_TRY(hipMemcpy(&err, b->d_error, 4, hipMemcpyDeviceToHost));
  if (err) return fail(K2G_ESOFTWARE, "compact hash table capacity exceeded");
  return K2G_OK;
}

static int check_window_lengths(const k2g_build_params &p) {
  if (p.subblock_size + p.k > 0xFFFFFFFFull) return fail(K2G_EUSAGE, "subblock size too large");
  return K2G_OK;
}

extern "C" int k2g_build_insert_device(k2g_builder *b, const char *d_seq, const uint64_t *off,
                                       const uint32_t *taxid, uint64_t n) {
  if (!b || (n && (!d_seq || !off || !taxid))) return fail(K2G_EUSAGE, "null argument");
  int rc = check_window_lengths(b->params);
  if (rc) return rc;
  for (uint64_t i = 0; i < n; i++) {
    if (off[i + 1] < off[i]) return fail(K2G_EUSAGE, "sequence offsets must not decrease");
    if (taxid[i] >= b->node_count) return fail(K2G_EDATA, "taxid is not a node of the taxonomy");
  }
  // windows in chunks so the item list stays modest
  std::vector<SubBlock> items;
  uint64_t s = 0;
  while (s < n) {
    items.clear();
    uint64_t e = s;
    while (e < n && items.size() < (1u << 22)) {
      build_windows(b->params, off + e, taxid + e, 1, 0, items);
      e++;
    }
    rc = build_launch(b, d_seq, items);
    if (rc) return rc;
    s = e;
  }
  return K2G_OK;
}

extern "C" int k2g_build_insert(k2g_builder *b, const char *seq, const uint64_t *off,
                                const uint32_t *taxid, uint64_t n) {
  if (!b || (n && (!seq || !off || !taxid))) return fail(K2G_EUSAGE, "null argument");
  if (n == 0) return K2G_OK;
  const uint64_t base = off[0], bytes = off[n] - off[0];
  HIP_TRY(b->seq.reserve(bytes + 1));
  HIP_TRY(hipMemcpy(b->seq.p, seq + base, bytes, hipMemcpyHostToDevice));
  std::vector<uint64_t> rel(n + 1);
  for (uint64_t i = 0; i <= n; i++) rel[i] = off[i] - base;
  return k2g_build_insert_device(b, b->seq.p, rel.data(), taxid, n);
}

// Table verification: the build's (minimizer, taxid) pairs collected in
// hash partitions (every minimizer sharing a cell with another has the same
// compacted key, so a cell's minimizers all fall in one partition), sorted
// by minimizer, looked up, then sorted by the cell they reach.
extern "C" int k2g_build_verify(k2g_builder *b, const char *d_seq, const uint64_t *off,
                                const uint32_t *taxid, uint64_t n, k2g_build_check *out) {
  if (!b || !out || (n && (!d_seq || !off || !taxid))) return fail(K2G_EUSAGE, "null argument");
  std::memset(out, 0, sizeof(*out));
  int rc = check_window_lengths(b->params);
  if (rc) return rc;
  uint64_t bases = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (off[i + 1] < off[i]) return fail(K2G_EUSAGE, "sequence offsets must not decrease");
    if (taxid[i] >= b->node_count) return fail(K2G_EDATA, "taxid is not a node of the taxonomy");
    if (taxid[i]) bases += off[i + 1] - off[i];
  }
  std::vector<SubBlock> items;
  build_windows(b->params, off, taxid, n, 0, items);
  hipStream_t s = b->stream;
  HIP_TRY(hipStreamSynchronize(s));
  // pairs per window <= its k-mers; minimizer windows of w l-mers make about
  // 2/(w+1) of the k-mers new minimizers: room for 1.5 times that, cut into
  // partitions of at most 2^30 pairs that use at most a third of free HBM
  const uint64_t w = b->params.k - b->params.l + 1;
  const uint64_t expect = bases * 2 / (w + 1) * 3 / 2 + (1u << 20);
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  uint32_t parts = 1;
  auto per_part = [&](uint32_t p) { return expect / p + (1u << 20); };
  while (parts < (1u << 16) &&
         (per_part(parts) > (3ull << 29) || per_part(parts) * 40 > free_b / 3))
    parts *= 2;
  uint32_t part_bits = (uint32_t) __builtin_ctz(parts);
  uint64_t cap = per_part(parts);
  DevBuf<SubBlock> d_items;
  DevBuf<uint64_t> ka, kb;
  DevBuf<uint32_t> ta, tb;
  DevBuf<unsigned long long> cnt;
  DevBuf<unsigned char> temp;
  auto release = [&] {
    d_items.release();
    ka.release();
    kb.release();
    ta.release();
    tb.release();
    cnt.release();
    temp.release();
  };
  auto run = [&]() -> int {
    HIP_TRY(d_items.reserve(std::max<size_t>(items.size(), 1)));
    if (!items.empty())
      HIP_TRY(hipMemcpyAsync(d_items.p, items.data(), items.size() * sizeof(SubBlock),
                             hipMemcpyHostToDevice, s));
    HIP_TRY(ka.reserve(cap));
    HIP_TRY(kb.reserve(cap));
    HIP_TRY(ta.reserve(cap));
    HIP_TRY(tb.reserve(cap));
    HIP_TRY(cnt.reserve(kVSlots + 2));
    HIP_TRY(hipMemsetAsync(cnt.p, 0, (kVSlots + 2) * 8, s));
    size_t tb1 = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, ka.p, kb.p, ta.p, tb.p, (int) cap, 0, 64, s));
    HIP_TRY(temp.reserve(tb1));
    DbView dv{};
    dv.cells = b->d_cells;
    dv.capacity = b->capacity;
    dv.barrett = b->view.barrett;
    dv.key_bits = b->view.key_bits;
    dv.value_bits = b->view.value_bits;
    dv.parent = b->d_parent;
    dv.node_count = (uint32_t) b->node_count;
    const int key_bits_sorted = b->params.dna_db ? (int) (2 * b->params.l) : (int) (4 * b->params.l);
    int cell_bits = 1;
    while (cell_bits < 64 && (1ull << cell_bits) <= b->capacity) cell_bits++;
    unsigned long long *const n_emit = cnt.p + kVSlots, *const n_found = cnt.p + kVSlots + 1;
    const uint64_t grid_b = std::min<uint64_t>(std::max<size_t>(items.size(), 1),
                                               (uint64_t) std::max(1, b->num_cus) * 32);
    for (uint32_t part = 0; part < parts; part++) {
      HIP_TRY(hipMemsetAsync(n_emit, 0, 16, s));
      BuildView v = b->view;
      v.cells = nullptr;
      v.sample = nullptr;
      v.emit_key = ka.p;
      v.emit_tax = ta.p;
      v.n_emit = n_emit;
      v.emit_cap = cap;
      v.part_bits = part_bits;
      v.part = part;
      if (!items.empty()) {
        hipLaunchKernelGGL(k_build, dim3((unsigned) grid_b), dim3(64), 0, s, v, d_seq, d_items.p,
                           (uint64_t) items.size());
        HIP_TRY(hipGetLastError());
      }
      unsigned long long m = 0;
      HIP_TRY(hipMemcpyAsync(&m, n_emit, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (m > cap) return 1;  // partition overflow: more partitions
      if (m == 0) continue;
      size_t tbytes = tb1;
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(temp.p, tbytes, ka.p, kb.p, ta.p, tb.p, (int) m, 0,
                                                 key_bits_sorted, s));
      const unsigned g = (unsigned) ((m + 255) / 256);
      hipLaunchKernelGGL(k_verify_keys, dim3(g), dim3(256), 0, s, dv, kb.p, tb.p, (uint64_t) m, ka.p,
                         ta.p, n_found, cnt.p);
      HIP_TRY(hipGetLastError());
      unsigned long long u = 0;
      HIP_TRY(hipMemcpyAsync(&u, n_found, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (u == 0) continue;
      tbytes = tb1;
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(temp.p, tbytes, ka.p, kb.p, ta.p, tb.p, (int) u, 0,
                                                 cell_bits, s));
      hipLaunchKernelGGL(k_verify_cells, dim3((unsigned) ((u + 255) / 256)), dim3(256), 0, s, dv, kb.p,
                         tb.p, (uint64_t) u, cnt.p);
      HIP_TRY(hipGetLastError());
    }
    unsigned long long c[kVSlots];
    HIP_TRY(hipMemcpyAsync(c, cnt.p, sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out->occurrences = c[kVOcc];
    out->keys = c[kVKeys];
    out->cells = c[kVCells];
    out->missing = c[kVMissing];
    out->not_covering = c[kVNotCovering];
    out->shared_cells = c[kVSharedCells];
    out->shared_keys = c[kVSharedKeys];
    out->differ = c[kVDiffer];
    out->unexplained = c[kVUnexplained];
    out->partitions = parts;
    return 0;
  };
  for (;;) {  // a partition that overflows its room: twice as many partitions
    rc = run();
    if (rc != 1 || parts >= (1u << 16)) break;
    parts *= 2;
    part_bits++;
    cap = per_part(parts);
  }
  release();
  if (rc == 1) return fail(K2G_ESOFTWARE, "verification partition overflow");
  if (rc) return rc;
  uint64_t size = 0;
  rc = k2g_build_get_info(b, nullptr, &size, nullptr, nullptr);
  if (rc) return rc;
  out->orphan_cells = size > out->cells ? size - out->cells : 0;
  return K2G_OK;
}

extern "C" int k2g_build_get_info(k2g_builder *b, uint64_t *capacity, uint64_t *size,
                                  uint64_t *key_bits, uint64_t *value_bits) {
  if (!b) return fail(K2G_EUSAGE, "null builder");
  unsigned long long sz = 0;
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(&sz, b->d_size, 8, hipMemcpyDeviceToHost));
  if (capacity) *capacity = b->capacity;
  if (size) *size = sz;
  if (key_bits) *key_bits = b->view.key_bits;
  if (value_bits) *value_bits = b->view.value_bits;
  return K2G_OK;
}

extern "C" int k2g_build_download(k2g_builder *b, uint32_t *cells, uint64_t capacity) {
  if (!b || !cells) return fail(K2G_EUSAGE, "null argument");
  if (capacity < b->capacity) return fail(K2G_EUSAGE, "output smaller than the table");
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(cells, b->d_cells, b->capacity * 4, hipMemcpyDeviceToHost));
  return K2G_OK;
}

extern "C" int k2g_build_write(k2g_builder *b, const char *hash_k2d) {
  if (!b || !hash_k2d) return fail(K2G_EUSAGE, "null argument");
  uint64_t hdr[4];
  int rc = k2g_build_get_info(b, &hdr[0], &hdr[1], &hdr[2], &hdr[3]);
  if (rc) return rc;
  FILE *f = fopen(hash_k2d, "wb");
  if (!f) return fail(K2G_EOS, std::string("unable to write ") + hash_k2d);
  bool ok = fwrite(hdr, 8, 4, f) == 4;
  const uint64_t chunk = 64ull << 20;
  std::vector<uint32_t> buf;
  for (uint64_t pos = 0; ok && pos < b->capacity; pos += chunk) {
    const uint64_t m = std::min(chunk, b->capacity - pos);
    buf.resize(m);
    if (hipMemcpy(buf.data(), b->d_cells + pos, m * 4, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    if (ok) ok = fwrite(buf.data(), 4, m, f) == m;
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(K2G_EIO, std::string("error writing ") + hash_k2d);
  return K2G_OK;
}

// ----------------------------------------------------------------------
// capacity estimate (src/estimate_capacity.cc:66-100, 185-203): the exact
// number of distinct minimizers among those whose hash falls in the first
// n_ranges of 1024 ranges, scaled by 1024 / n_ranges.
extern "C" int k2g_estimate_capacity_device(const k2g_build_params *p, const char *d_seq,
                                            const uint64_t *off, uint64_t n, uint32_t n_ranges,
                                            uint64_t *estimate) {
  if (!p || !estimate || (n && (!d_seq || !off))) return fail(K2G_EUSAGE, "null argument");
  if (n_ranges < 1 || n_ranges > 1024) return fail(K2G_EUSAGE, "n must be in [1, 1024]");
  const uint32_t bits = p->dna_db ? 2 : 4;
  if (p->l == 0 || p->l > 63 / bits || p->k < p->l || p->k - p->l > 63)
    return fail(K2G_EUSAGE, "unsupported k/l for the GPU scan");
  // whole-sequence scans cut into windows overlapping by k-1; a window that
  // continues a sequence reports from its (k-1)-th base on, like one long scan
  const uint64_t W = 1u << 16;
  std::vector<SubBlock> items;
  uint64_t bases = 0;
  for (uint64_t s = 0; s < n; s++) {
    if (off[s + 1] < off[s]) return fail(K2G_EUSAGE, "sequence offsets must not decrease");
    const uint64_t L = off[s + 1] - off[s];
    bases += L;
    for (uint64_t i = 0; i < L; i += W) {
      SubBlock it;
      it.start = off[s] + i;
      it.len = (uint32_t) std::min<uint64_t>(W + p->k - 1, L - i);
      it.taxid = 0;
      it.first_e = (uint32_t) (i == 0 ? p->k - 1 : (p->k - 1));
      it.pad = 0;
      if (i > 0) {
        // resume: the window starts k-1 bases back so its first report is
        // the k-mer ending at position i of the sequence
        it.start -= p->k - 1;
        it.len = (uint32_t) std::min<uint64_t>(W + p->k - 1, L - i + p->k - 1);
      }
      items.push_back(it);
    }
  }
  const uint64_t cap = bases * n_ranges / 1024 + 1024 * 1024;
  DevBuf<SubBlock> d_items;
  DevBuf<uint64_t> sample, sorted;
  DevBuf<unsigned long long> cnt;
  DevBuf<unsigned char> temp;
  HIP_TRY(d_items.reserve(std::max<size_t>(items.size(), 1)));
  HIP_TRY(sample.reserve(cap));
  HIP_TRY(sorted.reserve(cap));
  HIP_TRY(cnt.reserve(2));
  HIP_TRY(hipMemset(cnt.p, 0, 16));
  if (!items.empty())
    HIP_TRY(hipMemcpy(d_items.p, items.data(), items.size() * sizeof(SubBlock), hipMemcpyHostToDevice));
  BuildView v{};
  v.k = (uint32_t) p->k;
  v.l = (uint32_t) p->l;
  v.dna = p->dna_db ? 1 : 0;
  v.revcom_version = (uint32_t) p->revcom_version;
  v.spaced_mask = p->spaced_seed_mask;
  v.toggle_mask = p->toggle_mask & ((1ULL << (p->l * bits)) - 1);
  v.sample = sample.p;
  v.n_sample = cnt.p;
  v.sample_cap = cap;
  v.sample_ranges = n_ranges;
  int dev = 0, cus = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (!items.empty()) {
    const uint64_t grid = std::min<uint64_t>(items.size(), (uint64_t) std::max(1, cus) * 32);
    hipLaunchKernelGGL(k_build, dim3((unsigned) grid), dim3(64), 0, 0, v, d_seq, d_items.p,
                       (uint64_t) items.size());
    HIP_TRY(hipGetLastError());
  }
  unsigned long long m = 0;
  HIP_TRY(hipMemcpy(&m, cnt.p, 8, hipMemcpyDeviceToHost));
  if (m > cap) return fail(K2G_ESOFTWARE, "capacity sample overflow");
  uint64_t distinct = 0;
  if (m) {
    size_t tb = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, sample.p, sorted.p, (int) m));
    HIP_TRY(temp.reserve(tb));
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(temp.p, tb, sample.p, sorted.p, (int) m));
    size_t tb2 = 0;
    HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, tb2, sorted.p, sample.p, cnt.p + 1, (int) m));
    HIP_TRY(temp.reserve(tb2));
    HIP_TRY(hipcub::DeviceSelect::Unique(temp.p, tb2, sorted.p, sample.p, cnt.p + 1, (int) m));
    unsigned long long u = 0;
    HIP_TRY(hipMemcpy(&u, cnt.p + 1, 8, hipMemcpyDeviceToHost));
    distinct = u;
  }
  d_items.release();
  sample.release();
  sorted.release();
  cnt.release();
  temp.release();
  const uint64_t sum = distinct + 1;
  *estimate = (uint64_t) (sum * 1024 * 1.0 / n_ranges);
  return K2G_OK;
}

extern "C" int k2g_estimate_capacity(const k2g_build_params *p, const char *seq,
                                     const uint64_t *off, uint64_t n, uint32_t n_ranges,
                                     uint64_t *estimate) {
  if (!p || !estimate || (n && (!seq || !off))) return fail(K2G_EUSAGE, "null argument");
  const uint64_t base = n ? off[0] : 0, bytes = n ? off[n] - off[0] : 0;
  DevBuf<char> d;
  HIP_TRY(d.reserve(bytes + 1));
  if (bytes) HIP_TRY(hipMemcpy(d.p, seq + base, bytes, hipMemcpyHostToDevice));
  std::vector<uint64_t> rel(n + 1);
  for (uint64_t i = 0; i <= n; i++) rel[i] = off[i] - base;
  int rc = k2g_estimate_capacity_device(p, d.p, rel.data(), n, n_ranges, estimate);
  d.release();
  return rc;
}

// ----------------------------------------------------------------------
// random-gather ceiling of the resident table
extern "C" int k2g_db_gather_ceiling(k2g_db *db, uint64_t n_probes, uint64_t seed,
                                     double *lines_per_s, double *kernel_ms) {
  if (!db || !lines_per_s) return fail(K2G_EUSAGE, "null argument");
  const uint64_t n_lines = db->view.capacity / 32;  // 128-byte lines of whole cells
  if (n_lines == 0) return fail(K2G_EUSAGE, "table smaller than one 128-byte line");
  constexpr uint32_t per_lane = 32;
  constexpr unsigned block = 256;
  const uint64_t blocks64 = std::max<uint64_t>(1, (n_probes + per_lane * block - 1) / (per_lane * block));
  if (blocks64 > 0x7FFFFFFFull) return fail(K2G_EUSAGE, "too many probes");
  const unsigned grid = (unsigned) blocks64;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint32_t *sink = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMalloc((void **) &sink, 64 * 4);
  float ms = 0;
  for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // first launch warms up
    e = hipEventRecord(e0, s);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_gather_ceiling, dim3(grid), dim3(block), 0, s, db->view.cells, n_lines,
                         per_lane, seed + rep, sink);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  }
  if (sink) (void) hipFree(sink);
  if (e0) (void) hipEventDestroy(e0);
  if (e1) (void) hipEventDestroy(e1);
  if (s) (void) hipStreamDestroy(s);
  if (e != hipSuccess) return fail(K2G_ESOFTWARE, std::string("gather ceiling: ") + hipGetErrorString(e));
  const double probes = (double) grid * block * per_lane;
  *lines_per_s = ms > 0 ? probes / (ms * 1e-3) : 0.0;
  if (kernel_ms) *kernel_ms = ms;
  return K2G_OK;
}

// ----------------------------------------------------------------------
// device memory plan of one k2g_classify call in a bound region (mirrors
// the reservations of k2g_classify, k2g_classify_device and k2g_batch_fetch)
uint64_t k2g::plan_classify_bytes(const k2g_db *db, const k2g_classify_params *p,
                                  const k2g_fragments *in, k2g_batch::HostPlan *stats) {
  const uint64_t n = in->n;
  if (n == 0) return 0;
  const DbView &v = db->view;
  const uint64_t bases = in->off[2 * n];
  const uint32_t pct = p->run_cap_pct ? (uint32_t) p->run_cap_pct : 25u;
  uint64_t space = 0, max_e = 0, sum_e = 0, split_sum = 0, split_cnt = 0;
  bool too_long = false;
  for (uint64_t i = 0; i < n; i++) {
    const bool pr = in->paired ? in->paired[i] != 0 : p->paired != 0;
    const uint64_t l1 = in->off[2 * i + 1] - in->off[2 * i];
    const uint64_t l2 = pr ? in->off[2 * i + 2] - in->off[2 * i + 1] : 0;
    too_long |= ((l1 | l2) >> 31) != 0;
    const uint64_t e = fragment_elements(l1, l2, pr, v.k, v.dna != 0);
    space += run_space(e, pct);
    max_e = std::max(max_e, e);
    sum_e += e;
    if (e > 2ull * kSegElements) {
      split_sum += e;
      split_cnt++;
    }
  }
  if (stats) {
    stats->valid = true;
    stats->too_long = too_long;
    stats->space = space;
    stats->max_e = max_e;
    stats->sum_e = sum_e;
    stats->split_sum = split_sum;
    stats->split_cnt = split_cnt;
  }
  const bool sort = p->mode != 0;
  uint64_t b = 0;
  b += region_bytes(bases + 1, 1) + region_bytes(2 * n + 1, 8);  // seq, off
  if (in->paired) b += region_bytes(n, 1);
  b += 6 * region_bytes(n, 4);                                    // call .. n_recs
  b += 3 * region_bytes(n + 1, 8);                                // run_off, run_cap, bounds
  b += region_bytes(n, 4) + region_bytes(8, 4) + region_bytes(kCursorWords, 8);  // sort keys, counters, cursors
  b += region_bytes(temp_bytes_for(n, sort), 1);
  b += region_bytes(2 * std::max<uint64_t>(space, 1), 4);         // runs
  if (p->want_hll) b += region_bytes(2 * std::max<uint64_t>(space, 1), 4);
  if (p->want_kmer_counts) b += region_bytes(v.node_count, 8);
  if (sort) {
    b += 3 * region_bytes(n, 4);                                  // order, sort values, sorted keys
    b += 2 * region_bytes(n + 1, 8) + region_bytes(n, 1);         // unit prefix, scan, flags
    b += region_bytes(n + 1, 4) + region_bytes(1, 4);             // unit starts, unit count
  }
  if (p->want_probe_stats) b += region_bytes(n, 4) + region_bytes(n, 8);
  if (sort && v.dna && !p->quick && !p->want_probe_stats && max_e > 2ull * kSegElements)
    b += split_plan_bytes(n, split_sum, split_cnt, p->want_hll != 0);  // long fragments split over waves
  b += 3 * region_bytes(n, 4);                                    // redo lists
  const uint64_t slots = pct >= 100 ? dense_slots_for(n, v.node_count)  // dense-table redo, worst case
                                    : std::min<uint64_t>(n, 256);
  b += 2 * region_bytes(slots * v.node_count, 4) + region_bytes(slots, 4);
  if (pct < 100) {                                                // exact-space redo, worst case
    b += 2 * region_bytes(n + 1, 8);
    uint64_t exact = 0;
    for (uint64_t i = 0; i < n; i++) {
      const bool pr = in->paired ? in->paired[i] != 0 : p->paired != 0;
      exact += fragment_elements(in->off[2 * i + 1] - in->off[2 * i],
                                 pr ? in->off[2 * i + 2] - in->off[2 * i + 1] : 0, pr, v.k,
                                 v.dna != 0);
    }
    b += region_bytes(2 * (space + exact), 4) * (p->want_hll ? 2 : 1);
  }
  b += region_bytes(n + 1, 8) * (p->want_hll ? 2 : 1);           // fetch offsets
  b += region_bytes(1, 4);                                        // fetch status word
  if (p->want_hll == 2) {                                         // per-(block, taxon) HLL sets
    std::vector<uint64_t> starts;
    input_blocks(in, p->paired != 0, starts);
    b += sketch_plan_bytes(db, sum_e, starts.size() - 1);
  }
  return b;
}

// ----------------------------------------------------------------------
// kreport of one call's results (src/reports.cc:241-259), counters folded in
// the reference's input blocks (src/classify.cc:523-543, 554, 640-645)
extern "C" int k2g_report_write(k2g_db *db, const k2g_fragments *in, const k2g_results *r,
                                const k2g_report_params *p, const char *path) {
  if (!db || !in || !r || !p || !path) return fail(K2G_EUSAGE, "null argument");
  const uint64_t n = in->n;
  if (r->n != n) return fail(K2G_EUSAGE, "results and fragments differ in count");
  if (n && (!in->off || !r->call)) return fail(K2G_EUSAGE, "null fragment offsets or calls");
  const bool sets = r->sketch_off != nullptr;
  if (p->kmer_data && n && !(r->hll_off && r->hll) && !(sets && r->sketch_vals))
    return fail(K2G_EUSAGE, "a k-mer data report needs results with HLL records or sets (want_hll)");
  if (r->node_count != db->tax.nodes.size()) return fail(K2G_EUSAGE, "results of another database");
  std::vector<uint64_t> starts;
  k2g::input_blocks(in, p->paired != 0, starts);
  const uint64_t n_blocks = starts.size() - 1;
  k2host::TaxonCounters counters;
  uint64_t classified = 0;
  for (uint64_t f = 0; f < n; f++) {
    if (r->call[f] >= r->node_count) return fail(K2G_EUSAGE, "call is not a node of the taxonomy");
    if (r->call[f]) {
      counters[r->call[f]].n_reads++;
      classified++;
    }
  }
  if (p->kmer_data && sets) {
    // the device's per-(block, taxon) sets: folded like the block sketches
    // they stand for (a block's sketch merges into its taxon's total)
    for (uint64_t g = 0; g < r->n_sketches; g++) {
      if (r->sketch_block[g] >= n_blocks || r->sketch_taxon[g] >= r->node_count ||
          r->sketch_off[g + 1] < r->sketch_off[g])
        return fail(K2G_EUSAGE, "HLL sets of another call");
      k2host::Distinct d;
      d.n_observed = r->sketch_n[g];
      d.dense = r->sketch_dense[g] != 0;
      d.encoded.assign(r->sketch_vals + r->sketch_off[g], r->sketch_vals + r->sketch_off[g + 1]);
      k2host::TaxonCounter &tc = counters[r->sketch_taxon[g]];
      tc.n_kmers += d.n_observed;
      tc.distinct.merge(d);
    }
  } else if (p->kmer_data) {
    std::vector<std::pair<uint32_t, uint32_t>> recs;
    for (uint64_t blk = 0; blk < n_blocks; blk++) {
      recs.clear();
      for (uint64_t f = starts[blk]; f < starts[blk + 1]; f++)
        for (uint64_t j = r->hll_off[f]; j < r->hll_off[f + 1]; j++) recs.emplace_back(r->hll[2 * j], r->hll[2 * j + 1]);
      k2host::fold_block_records(recs, counters);
    }
  }
  if (!k2host::report_kraken_style(path, p->report_zeros != 0, p->kmer_data != 0, db->tax, counters, n,
                                   n - classified))
    return fail(K2G_EIO, std::string("unable to write report ") + path);
  return K2G_OK;
}

// ----------------------------------------------------------------------
// pinned host memory and result checksums
extern "C" int k2g_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) return fail(K2G_EOS, std::string("pinned allocation: ") + hipGetErrorString(e));
  return K2G_OK;
}

extern "C" int k2g_host_free(void *p) {
  if (p) HIP_TRY(hipHostFree(p));
  return K2G_OK;
}

namespace {
struct Checksum {
  uint64_t h = 0x6A09E667F3BCC909ull;
  void word(uint64_t w) {
    uint64_t x = h ^ (w + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2));
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    h = x;
  }
  // n elements of p (p may be null when n is 0: an empty array has no data).
  // Four independent multiply-rotate accumulators over the elements, folded
  // into the running hash, so long arrays hash at memory speed.
  template <class T>
  void array(const T *p, uint64_t n) {
    word(n);
    uint64_t a[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull,
                     0x85EBCA77C2B2AE63ull};
    uint64_t i = 0;
    for (; i + 4 <= n; i += 4)
      for (int j = 0; j < 4; j++) {
        const uint64_t x = (a[j] ^ (uint64_t) p[i + j]) * 0x9FB21C651E98DF25ull;
        a[j] = (x << 29) | (x >> 35);
      }
    for (int j = 0; j < 4; j++) word(a[j]);
    for (; i < n; i++) word((uint64_t) p[i]);
  }
};
}  // namespace

extern "C" int k2g_results_checksum(const k2g_results *r, uint64_t *out) {
  if (!r || !out) return fail(K2G_EUSAGE, "null argument");
  Checksum c;
  const uint64_t n = r->n;
  c.word(n);
  c.word(r->node_count);
  c.array(r->call, n);
  c.array(r->hit_groups, n);
  c.array(r->total_kmers, n);
  // which optional outputs are present is part of the result; a present
  // output's arrays are hashed by length and content only
  c.word((r->run_off ? 1u : 0u) | (r->hll_off ? 2u : 0u) | (r->kmer_counts ? 4u : 0u));
  c.array(r->run_off, r->run_off ? n + 1 : 0);
  c.array(r->runs, r->run_off ? 2 * r->run_off[n] : 0);
  c.array(r->hll_off, r->hll_off ? n + 1 : 0);
  c.array(r->hll, r->hll_off ? 2 * r->hll_off[n] : 0);
  c.array(r->kmer_counts, r->kmer_counts ? r->node_count : 0);
  if (r->sketch_off) {  // per-(block, taxon) HLL sets
    const uint64_t g = r->n_sketches;
    c.word(8u);
    c.array(r->sketch_block, g);
    c.array(r->sketch_taxon, g);
    c.array(r->sketch_dense, g);
    c.array(r->sketch_n, g);
    c.array(r->sketch_off, g + 1);
    c.array(r->sketch_vals, r->sketch_off[g]);
  }
  *out = c.h;
  return K2G_OK;
}
