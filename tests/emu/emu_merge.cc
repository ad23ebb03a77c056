// CPU execution of the merge kernel (kraken2_amd/csrc/k2g_merge_kernel.h)
// behind the library's merge entry points, so the k2g_merge front-end
// (kraken2_amd/csrc/host/merge_main.cc) runs end to end without a GPU in
// the tests.  The kernel runs one host thread per read, in read order.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "k2g.h"
#include "k2g_merge_kernel.h"

namespace emu {
thread_local Dim3 threadIdx_, blockIdx_;
Dim3 blockDim_, gridDim_;
std::barrier<> *bar = nullptr;
unsigned long long vote_slot[64];
uint64_t shfl_slot[64];
std::atomic<unsigned long long> collectives{0};
}  // namespace emu

namespace {
thread_local std::string g_error;
}

struct k2g_merger {
  std::vector<uint32_t> parent;
  std::vector<uint64_t> ext;
  k2g::MergeTables tb;
  k2g::MergeTax view{};
  std::vector<int32_t> w_print, w_count, tokens, call, d_tokens;
  std::vector<uint32_t> w_node, w_dcount;
  std::vector<uint64_t> n_tokens, tok_off, text_len;
  std::vector<uint8_t> joined, d_joined;
  std::vector<char> text;
};

extern "C" {

const char *k2g_last_error(void) { return g_error.c_str(); }

int k2g_host_alloc(size_t bytes, void **out) {
  *out = malloc(bytes ? bytes : 1);
  return *out ? K2G_OK : K2G_EOS;
}

int k2g_host_free(void *p) {
  free(p);
  return K2G_OK;
}

int k2g_merge_open(const char *taxo_k2d, k2g_merger **out) {
  auto *g = new k2g_merger();
  int rc = k2g::read_taxonomy_nodes(taxo_k2d, g->parent, g->ext, g_error);
  if (rc) {
    delete g;
    return rc;
  }
  // K2G_EMU_MERGE_SEARCH=1: no direct-mapped table, every lookup searches
  const char *search = getenv("K2G_EMU_MERGE_SEARCH");
  k2g::merge_tables(g->parent, g->ext, g->tb, search && atoi(search) ? 0 : k2g::kMergeDirectLimit);
  g->view = k2g::MergeTax{g->parent.data(),     g->tb.span.data(),     g->tb.ext32.data(),
                          g->ext.data(),        g->tb.key.data(),      g->tb.key_node.data(),
                          (uint32_t) g->tb.key.size(),
                          g->tb.direct.empty() ? nullptr : g->tb.direct.data(), g->tb.direct.size()};
  *out = g;
  return K2G_OK;
}

int k2g_merge(k2g_merger *g, const k2g_merge_params *p, const k2g_hit_lists *a, const k2g_hit_lists *b,
              k2g_merge_results *out);

int k2g_merge_text(k2g_merger *g, const k2g_merge_params *p, const k2g_hit_list_text *a,
                   const k2g_hit_list_text *b, k2g_merge_results *out) {
  if (a->n != b->n || (a->off && a->off[0]) || (b->off && b->off[0])) {
    g_error = "bad hit lists";
    return K2G_EUSAGE;
  }
  const uint64_t n = a->n;
  // offsets from the caller, or the fields' ':' counts (k_merge_count)
  std::vector<uint64_t> oa(n + 1, 0), ob(n + 1, 0);
  for (int side = 0; side < 2; side++) {
    const k2g_hit_list_text *h = side ? b : a;
    std::vector<uint64_t> &o = side ? ob : oa;
    if (h->off) {
      o.assign(h->off, h->off + n + 1);
      continue;
    }
    std::vector<uint64_t> cnt(n + 1, 0);
    for (uint64_t r = 0; r < n; r++) {
      emu::blockIdx_.x = (unsigned) r;
      emu::threadIdx_.x = 0;
      k2g::k_merge_count(h->text, h->field, h->len, n, cnt.data());
    }
    for (uint64_t r = 0; r < n; r++) o[r + 1] = o[r] + cnt[r];
  }
  std::vector<int32_t> ia(2 * oa[n] + 2), ib(2 * ob[n] + 2);
  uint32_t bad = 0;
  for (uint64_t r = 0; r < n; r++) {
    emu::blockIdx_.x = (unsigned) r;
    emu::threadIdx_.x = 0;
    k2g::k_merge_parse(a->text, a->field, a->len, oa.data(), n, ia.data(), &bad, a->off != nullptr);
    k2g::k_merge_parse(b->text, b->field, b->len, ob.data(), n, ib.data(), &bad, b->off != nullptr);
    if (!a->off || !b->off) k2g::k_merge_limits(oa.data(), ob.data(), n, &bad);
  }
  if (bad & 2u) {
    g_error = "a read's hit lists hold 2^32 or more items";
    return K2G_EUSAGE;
  }
  if (bad) {
    g_error = "item offsets do not match the ':' count of the hit-list fields";
    return K2G_EUSAGE;
  }
  k2g_hit_lists la{ia.data(), oa.data(), n}, lb{ib.data(), ob.data(), n};
  return k2g_merge(g, p, &la, &lb, out);
}

int k2g_merge(k2g_merger *g, const k2g_merge_params *p, const k2g_hit_lists *a, const k2g_hit_lists *b,
              k2g_merge_results *out) {
  if (a->n != b->n || a->off[0] || b->off[0]) {
    g_error = "bad hit lists";
    return K2G_EUSAGE;
  }
  emu::blockDim_.x = 1;
  const uint64_t n = a->n, slots = a->off[n] + b->off[n];
  g->w_print.assign(slots, 0);
  g->w_count.assign(slots, 0);
  g->w_node.assign(slots, 0);
  g->w_dcount.assign(slots, 0);
  g->tokens.assign(2 * slots, 0);
  g->joined.assign(slots, 0);
  g->n_tokens.assign(n + 1, 0);
  g->text_len.assign(n + 1, 0);
  g->call.assign(n, 0);
  emu::blockDim_.x = 1;
  emu::gridDim_.x = (unsigned) n;
  for (uint64_t r = 0; r < n; r++) {
    emu::blockIdx_.x = (unsigned) r;
    emu::threadIdx_.x = 0;
    k2g::k_merge(g->view, a->items, a->off, b->items, b->off, n, p->confidence, g->w_print.data(),
                 g->w_count.data(), g->w_node.data(), g->w_dcount.data(), g->tokens.data(), g->joined.data(),
                 g->n_tokens.data(),
                 g->text_len.data(), g->call.data());
  }
  const bool text = p->text != 0;
  const std::vector<uint64_t> &counts = text ? g->text_len : g->n_tokens;
  g->tok_off.assign(n + 1, 0);
  for (uint64_t r = 0; r < n; r++) g->tok_off[r + 1] = g->tok_off[r] + counts[r];
  const uint64_t total = g->tok_off[n];
  if (text) {
    g->text.assign(total + 1, 0);
    for (uint64_t r = 0; r < n; r++) {
      emu::blockIdx_.x = (unsigned) r;
      k2g::k_merge_print(a->off, b->off, n, g->n_tokens.data(), g->tok_off.data(), g->tokens.data(),
                         g->joined.data(), g->text.data());
    }
  } else {
    g->d_tokens.assign(2 * total, 0);
    g->d_joined.assign(total, 0);
    for (uint64_t r = 0; r < n; r++) {
      emu::blockIdx_.x = (unsigned) r;
      k2g::k_merge_compact(a->off, b->off, n, g->tok_off.data(), g->tokens.data(), g->joined.data(),
                           g->d_tokens.data(), g->d_joined.data());
    }
  }
  out->n = n;
  out->call = g->call.data();
  out->token_off = text ? nullptr : g->tok_off.data();
  out->tokens = text ? nullptr : g->d_tokens.data();
  out->joined = text ? nullptr : g->d_joined.data();
  out->text = text ? g->text.data() : nullptr;
  out->text_off = text ? g->tok_off.data() : nullptr;
  out->kernel_ms = 0;
  return K2G_OK;
}

int k2g_merge_close(k2g_merger *g) {
  delete g;
  return K2G_OK;
}

}  // extern "C"
