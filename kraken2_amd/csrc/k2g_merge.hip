// Multi-database merge (the reference's merge tool, src/merge.cc): the
// resident merged taxonomy and the batch entry point over k_merge
// (k2g_merge_kernel.h).
#include "k2g_internal.h"
#include "k2g_merge_kernel.h"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstring>

using namespace k2g;

namespace {

template <class T>
hipError_t upload(DevBuf<T> &d, const T *h, size_t n, hipStream_t s) {
  hipError_t e = d.reserve(n);
  if (e == hipSuccess && n) e = hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
  return e;
}

}  // namespace

struct k2g_merger {
  MergeTax view{};
  DevBuf<uint32_t> parent, span, key_node, direct;
  DevBuf<int32_t> ext32;
  DevBuf<uint64_t> ext64, key;
  DevBuf<int32_t> a_items, b_items, w_print, w_count, tokens, d_tokens, call;
  DevBuf<uint64_t> a_off, b_off, n_tokens, tok_off, text_len;
  DevBuf<uint32_t> w_node, w_dcount;
  DevBuf<uint8_t> joined, d_joined, scan_tmp;
  DevBuf<char> text, a_text, b_text;
  DevBuf<uint64_t> a_field, b_field;
  DevBuf<uint32_t> a_len, b_len, bad;
  PinnedBuf<int32_t> h_call, h_tokens;
  PinnedBuf<uint64_t> h_tok_off;
  PinnedBuf<uint8_t> h_joined;
  PinnedBuf<char> h_text;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~k2g_merger() {
    for (auto *b : {&parent, &span, &key_node, &direct, &w_node, &w_dcount}) b->release();
    for (auto *b : {&ext32, &a_items, &b_items, &w_print, &w_count, &tokens, &d_tokens, &call}) b->release();
    for (auto *b : {&ext64, &key, &a_off, &b_off, &n_tokens, &tok_off, &text_len}) b->release();
    for (auto *b : {&joined, &d_joined, &scan_tmp}) b->release();
    for (auto *b : {&text, &a_text, &b_text}) b->release();
    for (auto *b : {&a_field, &b_field}) b->release();
    for (auto *b : {&a_len, &b_len, &bad}) b->release();
    h_text.release();
    h_call.release();
    h_tokens.release();
    h_tok_off.release();
    h_joined.release();
    if (ev0) (void) hipEventDestroy(ev0);
    if (ev1) (void) hipEventDestroy(ev1);
    if (stream) (void) hipStreamDestroy(stream);
  }
};

extern "C" int k2g_merge_open(const char *taxo_k2d, k2g_merger **out) {
  if (!taxo_k2d || !out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  std::vector<uint32_t> parent;
  std::vector<uint64_t> ext;
  std::string err;
  int rc = read_taxonomy_nodes(taxo_k2d, parent, ext, err);
  if (rc) return fail(rc, err);
  const size_t n = parent.size();
  MergeTables tb;
  merge_tables(parent, ext, tb);
  const std::vector<uint32_t> &span = tb.span, &key_node = tb.key_node;
  const std::vector<int32_t> &ext32 = tb.ext32;
  const std::vector<uint64_t> &key = tb.key;
  auto *g = new k2g_merger();
  hipStream_t s;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) g->stream = s;
  if (e == hipSuccess) e = hipEventCreate(&g->ev0);
  if (e == hipSuccess) e = hipEventCreate(&g->ev1);
  if (e == hipSuccess) e = upload(g->parent, parent.data(), n, s);
  if (e == hipSuccess) e = upload(g->span, span.data(), 2 * n, s);
  if (e == hipSuccess) e = upload(g->ext32, ext32.data(), n, s);
  if (e == hipSuccess) e = upload(g->ext64, ext.data(), n, s);
  if (e == hipSuccess) e = upload(g->key, key.data(), key.size(), s);
  if (e == hipSuccess) e = upload(g->key_node, key_node.data(), key_node.size(), s);
  if (e == hipSuccess && !tb.direct.empty()) e = upload(g->direct, tb.direct.data(), tb.direct.size(), s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete g;
    return fail(K2G_ESOFTWARE, std::string("merge taxonomy upload failed: ") + hipGetErrorString(e));
  }
  g->view = MergeTax{g->parent.p, g->span.p, g->ext32.p, g->ext64.p, g->key.p, g->key_node.p,
                     (uint32_t) key.size(), tb.direct.empty() ? nullptr : g->direct.p, tb.direct.size()};
  *out = g;
  return K2G_OK;
}

static int run_merge(k2g_merger *g, const k2g_merge_params *p, uint64_t n, uint64_t slots,
                     k2g_merge_results *out);

static int check_lists(const k2g_hit_lists *h, const char *which) {
  if (!h->off) return fail(K2G_EUSAGE, std::string("null offsets in hit lists ") + which);
  if (h->off[0] != 0) return fail(K2G_EUSAGE, std::string("hit lists ") + which + " must start at offset 0");
  for (uint64_t i = 0; i < h->n; i++)
    if (h->off[i + 1] < h->off[i])
      return fail(K2G_EUSAGE, std::string("offsets of hit lists ") + which + " must not decrease");
  if (h->off[h->n] && !h->items) return fail(K2G_EUSAGE, std::string("null items in hit lists ") + which);
  if (h->off[h->n] > (uint64_t) UINT32_MAX)
    return fail(K2G_EUSAGE, std::string("hit lists ") + which + " hold 2^32 or more items");
  return K2G_OK;
}

extern "C" int k2g_merge(k2g_merger *g, const k2g_merge_params *p, const k2g_hit_lists *a,
                         const k2g_hit_lists *b, k2g_merge_results *out) {
  if (!g || !p || !a || !b || !out) return fail(K2G_EUSAGE, "null argument");
  if (a->n != b->n) return fail(K2G_EUSAGE, "the two inputs hold different numbers of reads");
  int rc = check_lists(a, "a");
  if (!rc) rc = check_lists(b, "b");
  if (rc) return rc;
  for (uint64_t i = 0; i < a->n; i++)
    if (a->off[i + 1] - a->off[i] + b->off[i + 1] - b->off[i] > (uint64_t) UINT32_MAX)
      return fail(K2G_EUSAGE, "a read's hit lists hold 2^32 or more items");
  const uint64_t n = a->n, na = a->off[n], nb = b->off[n];
  hipStream_t s = g->stream;
  HIP_TRY(upload(g->a_items, a->items, 2 * na, s));
  HIP_TRY(upload(g->b_items, b->items, 2 * nb, s));
  HIP_TRY(upload(g->a_off, a->off, n + 1, s));
  HIP_TRY(upload(g->b_off, b->off, n + 1, s));
  return run_merge(g, p, n, na + nb, out);
}

// Merges the n reads whose items and offsets are on the device.
static int run_merge(k2g_merger *g, const k2g_merge_params *p, uint64_t n, uint64_t slots, k2g_merge_results *out) {
  hipStream_t s = g->stream;
  HIP_TRY(g->w_print.reserve(slots));
  HIP_TRY(g->w_count.reserve(slots));
  HIP_TRY(g->w_node.reserve(slots));
  HIP_TRY(g->w_dcount.reserve(slots));
  HIP_TRY(g->tokens.reserve(2 * slots));
  HIP_TRY(g->joined.reserve(slots));
  HIP_TRY(g->n_tokens.reserve(n + 1));
  HIP_TRY(g->text_len.reserve(n + 1));
  HIP_TRY(g->tok_off.reserve(n + 1));
  HIP_TRY(g->call.reserve(n));
  const bool text = p->text != 0;
  uint64_t *counts = text ? g->text_len.p : g->n_tokens.p;  // what the scan lays out
  size_t scan_bytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts, g->tok_off.p, n + 1, s));
  HIP_TRY(g->scan_tmp.reserve(scan_bytes + 1));
  HIP_TRY(hipEventRecord(g->ev0, s));
  const uint32_t blocks = (uint32_t) ((n + 255) / 256);
  HIP_TRY(hipMemsetAsync(counts + n, 0, 8, s));
  if (n) {
    hipLaunchKernelGGL(k_merge, dim3(blocks), dim3(256), 0, s, g->view, g->a_items.p, g->a_off.p, g->b_items.p,
                       g->b_off.p, n, p->confidence, g->w_print.p, g->w_count.p, g->w_node.p, g->w_dcount.p,
                       g->tokens.p, g->joined.p, g->n_tokens.p, g->text_len.p, g->call.p);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(g->scan_tmp.p, scan_bytes, counts, g->tok_off.p, n + 1, s));
  uint64_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, g->tok_off.p + n, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (text) {
    HIP_TRY(g->text.reserve(total + 1));
    if (n) {
      hipLaunchKernelGGL(k_merge_print, dim3(blocks), dim3(256), 0, s, g->a_off.p, g->b_off.p, n, g->n_tokens.p,
                         g->tok_off.p, g->tokens.p, g->joined.p, g->text.p);
      HIP_TRY(hipGetLastError());
    }
  } else {
    HIP_TRY(g->d_tokens.reserve(2 * total));
    HIP_TRY(g->d_joined.reserve(total));
    if (n) {
      hipLaunchKernelGGL(k_merge_compact, dim3(blocks), dim3(256), 0, s, g->a_off.p, g->b_off.p, n, g->tok_off.p,
                         g->tokens.p, g->joined.p, g->d_tokens.p, g->d_joined.p);
      HIP_TRY(hipGetLastError());
    }
  }
  HIP_TRY(hipEventRecord(g->ev1, s));
  HIP_TRY(g->h_call.reserve(n + 1));
  HIP_TRY(g->h_tok_off.reserve(n + 1));
  if (n) HIP_TRY(hipMemcpyAsync(g->h_call.p, g->call.p, n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(g->h_tok_off.p, g->tok_off.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
  if (text) {
    HIP_TRY(g->h_text.reserve(total + 1));
    if (total) HIP_TRY(hipMemcpyAsync(g->h_text.p, g->text.p, total, hipMemcpyDeviceToHost, s));
  } else {
    HIP_TRY(g->h_tokens.reserve(2 * total + 2));
    HIP_TRY(g->h_joined.reserve(total + 1));
    if (total) {
      HIP_TRY(hipMemcpyAsync(g->h_tokens.p, g->d_tokens.p, 2 * total * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(g->h_joined.p, g->d_joined.p, total, hipMemcpyDeviceToHost, s));
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  out->n = n;
  out->call = g->h_call.p;
  out->token_off = text ? nullptr : g->h_tok_off.p;
  out->tokens = text ? nullptr : g->h_tokens.p;
  out->joined = text ? nullptr : g->h_joined.p;
  out->text = text ? g->h_text.p : nullptr;
  out->text_off = text ? g->h_tok_off.p : nullptr;
  out->kernel_ms = ms;
  return K2G_OK;
}

static int check_text(const k2g_hit_list_text *h, const char *which) {
  const std::string w = std::string(" of hit-list text ") + which;
  if (h->n && (!h->field || !h->len)) return fail(K2G_EUSAGE, "null arrays" + w);
  if (h->off && h->off[0] != 0) return fail(K2G_EUSAGE, "item offsets" + w + " must start at 0");
  for (uint64_t i = 0; i < h->n; i++) {
    if (h->off && h->off[i + 1] < h->off[i]) return fail(K2G_EUSAGE, "item offsets" + w + " must not decrease");
    if (h->field[i] > h->text_bytes || h->len[i] > h->text_bytes - h->field[i])
      return fail(K2G_EUSAGE, "a field" + w + " lies outside the text");
  }
  if (h->text_bytes && !h->text) return fail(K2G_EUSAGE, "null text" + w);
  if (h->off && h->off[h->n] > (uint64_t) UINT32_MAX)
    return fail(K2G_EUSAGE, "hit lists" + w + " hold 2^32 or more items");
  return K2G_OK;
}

extern "C" int k2g_merge_text(k2g_merger *g, const k2g_merge_params *p, const k2g_hit_list_text *a,
                              const k2g_hit_list_text *b, k2g_merge_results *out) {
  if (!g || !p || !a || !b || !out) return fail(K2G_EUSAGE, "null argument");
  if (a->n != b->n) return fail(K2G_EUSAGE, "the two inputs hold different numbers of reads");
  int rc = check_text(a, "a");
  if (!rc) rc = check_text(b, "b");
  if (rc) return rc;
  const uint64_t n = a->n;
  if (a->off && b->off)
    for (uint64_t i = 0; i < n; i++)
      if (a->off[i + 1] - a->off[i] + b->off[i + 1] - b->off[i] > (uint64_t) UINT32_MAX)
        return fail(K2G_EUSAGE, "a read's hit lists hold 2^32 or more items");
  hipStream_t s = g->stream;
  HIP_TRY(upload(g->a_text, a->text, a->text_bytes, s));
  HIP_TRY(upload(g->b_text, b->text, b->text_bytes, s));
  HIP_TRY(upload(g->a_field, a->field, n, s));
  HIP_TRY(upload(g->b_field, b->field, n, s));
  HIP_TRY(upload(g->a_len, a->len, n, s));
  HIP_TRY(upload(g->b_len, b->len, n, s));
  // item offsets: the caller's, or the fields' ':' counts scanned on the device
  const uint32_t blocks = (uint32_t) ((n + 255) / 256);
  uint64_t tot[2] = {0, 0};
  for (int side = 0; side < 2; side++) {
    const k2g_hit_list_text *h = side ? b : a;
    DevBuf<uint64_t> &off = side ? g->b_off : g->a_off;
    if (h->off) {
      HIP_TRY(upload(off, h->off, n + 1, s));
      tot[side] = h->off[n];
      continue;
    }
    HIP_TRY(off.reserve(n + 1));
    HIP_TRY(g->n_tokens.reserve(n + 1));  // scratch here: the counts
    HIP_TRY(hipMemsetAsync(g->n_tokens.p + n, 0, 8, s));
    if (n) {
      hipLaunchKernelGGL(k_merge_count, dim3(blocks), dim3(256), 0, s, side ? g->b_text.p : g->a_text.p,
                         side ? g->b_field.p : g->a_field.p, side ? g->b_len.p : g->a_len.p, n, g->n_tokens.p);
      HIP_TRY(hipGetLastError());
    }
    size_t scan_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, g->n_tokens.p, off.p, n + 1, s));
    HIP_TRY(g->scan_tmp.reserve(scan_bytes + 1));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(g->scan_tmp.p, scan_bytes, g->n_tokens.p, off.p, n + 1, s));
    HIP_TRY(hipMemcpyAsync(&tot[side], off.p + n, 8, hipMemcpyDeviceToHost, s));
  }
  if (!a->off || !b->off) HIP_TRY(hipStreamSynchronize(s));  // one wait for the device totals
  if (tot[0] > (uint64_t) UINT32_MAX || tot[1] > (uint64_t) UINT32_MAX)
    return fail(K2G_EUSAGE, std::string("hit lists ") + (tot[0] > (uint64_t) UINT32_MAX ? "a" : "b") +
                                " hold 2^32 or more items");
  const uint64_t na = tot[0], nb = tot[1];
  HIP_TRY(g->a_items.reserve(2 * na + 2));
  HIP_TRY(g->b_items.reserve(2 * nb + 2));
  HIP_TRY(g->bad.reserve(1));
  HIP_TRY(hipMemsetAsync(g->bad.p, 0, 4, s));
  if (n) {
    hipLaunchKernelGGL(k_merge_parse, dim3(blocks), dim3(256), 0, s, g->a_text.p, g->a_field.p, g->a_len.p,
                       g->a_off.p, n, g->a_items.p, g->bad.p, a->off != nullptr);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_merge_parse, dim3(blocks), dim3(256), 0, s, g->b_text.p, g->b_field.p, g->b_len.p,
                       g->b_off.p, n, g->b_items.p, g->bad.p, b->off != nullptr);
    HIP_TRY(hipGetLastError());
    if (!a->off || !b->off) {  // the per-read limit the host checks when both offsets are given
      hipLaunchKernelGGL(k_merge_limits, dim3(blocks), dim3(256), 0, s, (const uint64_t *) g->a_off.p,
                         (const uint64_t *) g->b_off.p, n, g->bad.p);
      HIP_TRY(hipGetLastError());
    }
  }
  uint32_t bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, g->bad.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (bad & 2u) return fail(K2G_EUSAGE, "a read's hit lists hold 2^32 or more items");
  if (bad) return fail(K2G_EUSAGE, "item offsets do not match the ':' count of the hit-list fields");
  return run_merge(g, p, n, na + nb, out);
}

extern "C" int k2g_merge_close(k2g_merger *g) {
  delete g;
  return K2G_OK;
}
