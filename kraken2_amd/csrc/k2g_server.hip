// Request server of libk2g.so: the resident-index request loop.
//
// The reference keeps its index loaded between requests in ClassifyDaemon
// (src/classify.cc:373-430) and runs each request's whole classify() call
// one after another.  Here requests are admitted in arrival order whenever a
// slot is free and the device memory the request needs fits in what is left
// of a fixed budget; admitted requests run concurrently, each on its own
// slot's HIP stream, so their transfers and kernels overlap on the GPU while
// the table stays resident.  The budget is one device allocation made at
// start-up and carved per request (first fit), so the server never holds
// more device memory than it was given.
#include "k2g.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "k2g_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

using Arena = k2g::ByteArena;

// Host result sets the server keeps for reuse (at most this many, holding at
// most half its result memory).
constexpr size_t kPoolSets = 64;

struct Request {
  uint64_t ticket = 0;
  k2g_classify_params params{};
  k2g_fragments in{};
  uint64_t bytes = 0;
  k2g_batch::HostPlan plan;  // the batch statistics the admission plan computed
  int state = 0;  // 0 queued, 1 on the GPU, 2 done
  int rc = K2G_OK;
  std::string error;
  Clock::time_point t_submit, t_start, t_done;
  // results: the slot batch's host results, handed over without a copy
  std::unique_ptr<k2g::HostResults> res;
  k2g_results out{};
};

}  // namespace

struct k2g_server {
  k2g_db *db = nullptr;
  char *d_base = nullptr;
  uint64_t budget = 0;
  Arena arena{0};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<uint64_t, Request *> requests;
  std::vector<std::thread> workers;
  bool stop = false;
  uint64_t next_ticket = 1;
  uint64_t in_use = 0, peak_in_use = 0;
  int running = 0, max_running = 0;
  int worker_rc = K2G_OK;
  std::string worker_error;
  // host result sets of released requests, reused by the slots (their
  // vectors and pinned buffers keep their capacity)
  std::vector<std::unique_ptr<k2g::HostResults>> pool;
  // pinned host memory the result sets are carved from (k2g_server_destroy
  // returns every set to it before freeing it)
  k2g::HostArena host;
};

namespace {

void worker_main(k2g_server *s) {
  k2g_batch *b = nullptr;
  int rc = hipSetDevice(s->db->device) == hipSuccess ? K2G_OK : K2G_ESOFTWARE;
  if (rc == K2G_OK) rc = k2g_batch_create(s->db, &b);
  if (rc != K2G_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->worker_rc = rc;
    s->worker_error = "server slot setup: " + k2g::last_error();
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    return;
  }
  k2g::Region region;
  for (;;) {
    Request *q = nullptr;
    uint64_t off = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (!s->queue.empty() && s->arena.alloc(s->queue.front()->bytes, &off)) {
          q = s->queue.front();
          s->queue.pop_front();
          break;
        }
        if (s->stop && s->queue.empty()) break;
        s->cv_work.wait(lk);
      }
      if (!q) break;
      q->state = 1;
      q->t_start = Clock::now();
      s->in_use += q->bytes;
      s->peak_in_use = std::max(s->peak_in_use, s->in_use);
      s->running++;
      s->max_running = std::max(s->max_running, s->running);
    }
    {
      // the host result set that fits this request best, the slot's own or a
      // pooled one: the smallest that held a call of at least this request's
      // element bound, else the largest (pinned buffers grow by reallocation,
      // and freeing pinned memory waits for the whole device)
      std::lock_guard<std::mutex> lk(s->mu);
      const uint64_t need = q->plan.space;
      auto better = [need](uint64_t c, uint64_t than) {
        return (c >= need) != (than >= need) ? c >= need : (c >= need ? c < than : c > than);
      };
      size_t pick = s->pool.size();
      uint64_t got = b->host->served;
      for (size_t i = 0; i < s->pool.size(); i++)
        if (better(s->pool[i]->served, got)) {
          pick = i;
          got = s->pool[i]->served;
        }
      if (pick < s->pool.size()) std::swap(b->host, s->pool[pick]);
    }
    region.base = s->d_base + off;
    region.size = q->bytes;
    region.used = 0;
    b->bind_region(&region);
    b->plan = q->plan;  // the launch reads no statistics back
    int r = k2g_classify(b, &q->params, &q->in, &q->out);
    b->plan.valid = false;
    if (r == K2G_OK) {
      b->host->served = std::max(b->host->served, q->plan.space);
      q->res = std::move(b->host);  // q->out points into these buffers
    } else {
      q->rc = r;
      q->error = k2g::last_error();
    }
    b->bind_region(nullptr);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (!b->host) {
        if (!s->pool.empty()) {
          b->host = std::move(s->pool.back());
          s->pool.pop_back();
        } else {
          b->host = std::make_unique<k2g::HostResults>(&s->host);
        }
      }
      s->arena.release(off, q->bytes);
      s->in_use -= q->bytes;
      s->running--;
      q->state = 2;
      q->t_done = Clock::now();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
  k2g_batch_destroy(b);
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

using k2g::fail;

extern "C" int k2g_server_create(k2g_db *db, const k2g_server_params *p, k2g_server **out) {
  if (!db || !p || !out) return fail(K2G_EUSAGE, "null argument");
  *out = nullptr;
  if (p->slots < 1 || p->slots > 64) return fail(K2G_EUSAGE, "slots must be in [1, 64]");
  if (p->hbm_budget < (1u << 20)) return fail(K2G_EUSAGE, "hbm_budget must be at least 1 MiB");
  auto *s = new k2g_server();
  s->db = db;
  s->budget = p->hbm_budget / k2g::kRegionAlign * k2g::kRegionAlign;
  hipError_t e = hipSetDevice(db->device);
  if (e == hipSuccess) e = hipMalloc((void **) &s->d_base, s->budget);
  if (e != hipSuccess) {
    delete s;
    return fail(K2G_EOS, std::string("server budget allocation: ") + hipGetErrorString(e));
  }
  s->arena = Arena(s->budget);
  // result sets: pinned host memory of a quarter of the device budget,
  // within [256 MiB, 16 GiB], halved while the host cannot pin that much
  // (without any, result buffers are pinned one by one)
  for (uint64_t want = std::min<uint64_t>(std::max<uint64_t>(s->budget / 4, 256ull << 20), 16ull << 30);
       want >= (256ull << 20); want /= 2) {
    if (hipHostMalloc((void **) &s->host.base, want, hipHostMallocMapped) == hipSuccess) {
      s->host.size = want;
      break;
    }
    s->host.base = nullptr;
  }
  (void) hipGetLastError();  // a refused size is not an error of the next call
  s->host.ranges = k2g::ByteArena(s->host.size);
  // a slot whose results do not fit takes back the largest pooled set's
  // memory before pinning more (freeing pinned memory waits for the whole
  // device, which stalls every slot)
  s->host.reclaim = [s] {
    std::unique_ptr<k2g::HostResults> victim;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->pool.empty()) return false;
      size_t big = 0;
      for (size_t i = 1; i < s->pool.size(); i++)
        if (s->pool[i]->bytes() > s->pool[big]->bytes()) big = i;
      victim = std::move(s->pool[big]);
      s->pool.erase(s->pool.begin() + big);
    }
    return true;  // victim's buffers go back to the arena here
  };
  for (int i = 0; i < p->slots; i++) s->workers.emplace_back(worker_main, s);
  *out = s;
  return K2G_OK;
}

extern "C" int k2g_server_submit(k2g_server *s, const k2g_classify_params *p,
                                 const k2g_fragments *in, uint64_t *ticket) {
  if (!s || !p || !in || !ticket) return fail(K2G_EUSAGE, "null argument");
  if (in->n && (!in->seq || !in->off)) return fail(K2G_EUSAGE, "null fragment arrays");
  for (uint64_t i = 0; i < 2 * in->n; i++)
    if (in->off[i + 1] < in->off[i]) return fail(K2G_EUSAGE, "fragment offsets must not decrease");
  auto *q = new Request();
  q->params = *p;
  q->params.run_cap_pct = 100;  // exact hitlist space: the plan is an upper bound
  q->in = *in;
  q->bytes = k2g::plan_classify_bytes(s->db, &q->params, in, &q->plan);
  q->bytes = (q->bytes + k2g::kRegionAlign - 1) / k2g::kRegionAlign * k2g::kRegionAlign;
  if (q->bytes > s->budget) {
    const uint64_t need = q->bytes;
    delete q;
    return fail(K2G_EUSAGE, "request needs " + std::to_string(need) +
                                " bytes of device memory; the server budget is " +
                                std::to_string(s->budget));
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->worker_rc != K2G_OK) {
      delete q;
      return fail(s->worker_rc, s->worker_error);
    }
    q->ticket = s->next_ticket++;
    q->t_submit = Clock::now();
    s->requests[q->ticket] = q;
    s->queue.push_back(q);
    *ticket = q->ticket;
  }
  s->cv_work.notify_all();
  return K2G_OK;
}

extern "C" int k2g_server_wait(k2g_server *s, uint64_t ticket, k2g_results *out, double *times_ms) {
  if (!s) return fail(K2G_EUSAGE, "null server");
  Request *q = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return fail(K2G_EUSAGE, "unknown ticket");
    q = it->second;
    s->cv_done.wait(lk, [&] { return q->state == 2 || (s->worker_rc != K2G_OK && q->state == 0); });
    if (q->state != 2) return fail(s->worker_rc, s->worker_error);
  }
  if (times_ms) {
    times_ms[0] = ms_between(q->t_submit, q->t_start);
    times_ms[1] = ms_between(q->t_start, q->t_done);
    times_ms[2] = ms_between(q->t_submit, q->t_done);
  }
  if (q->rc != K2G_OK) return fail(q->rc, q->error);
  if (out) *out = q->out;  // valid until the request is released
  return K2G_OK;
}

extern "C" int k2g_server_release(k2g_server *s, uint64_t ticket) {
  if (!s) return fail(K2G_EUSAGE, "null server");
  std::unique_ptr<k2g::HostResults> drop;  // freed after the lock is released
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return fail(K2G_EUSAGE, "unknown ticket");
    if (it->second->state != 2) return fail(K2G_EUSAGE, "request still in progress");
    // result sets are kept for reuse up to kPoolSets sets and half the result
    // memory; beyond that they are freed
    if (it->second->res) {
      uint64_t held = 0;
      for (const auto &x : s->pool) held += x->bytes();
      if (s->pool.size() < kPoolSets && held + it->second->res->bytes() <= s->host.size / 2)
        s->pool.push_back(std::move(it->second->res));
      else
        drop = std::move(it->second->res);
    }
    delete it->second;
    s->requests.erase(it);
  }
  return K2G_OK;
}

extern "C" int k2g_server_stats(k2g_server *s, uint64_t *peak_bytes, int32_t *max_concurrent,
                                uint64_t *budget) {
  if (!s) return fail(K2G_EUSAGE, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (peak_bytes) *peak_bytes = s->peak_in_use;
  if (max_concurrent) *max_concurrent = s->max_running;
  if (budget) *budget = s->budget;
  return K2G_OK;
}

extern "C" int k2g_server_destroy(k2g_server *s) {
  if (!s) return K2G_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->requests) delete kv.second;
  s->pool.clear();  // every result set is back in the arena before it is freed
  if (s->d_base) (void) hipFree(s->d_base);
  if (s->host.base) (void) hipHostFree(s->host.base);
  delete s;
  return K2G_OK;
}
