// Serving loop: secret-shared query requests (a table sort, or GROUP BY key
// SUM(value)) on one GPU with continuous batching.
//
// Each of `slots` worker threads owns a context (its own streams, scratch and
// device-memory pool).  A request is admitted in arrival order as soon as a
// slot is free and its estimated device footprint fits the HBM budget next to
// the requests already running; admitted requests run concurrently on their
// slots' streams, so small requests fill the CUs a large one leaves idle and
// one request's host-side waits overlap another's kernels.  Every request
// re-keys its slot's context with the request's own keys, so its result
// shares depend only on the request: the same alone and in any mix.  With
// batch_rows set, the small radix sorts of concurrent requests meet in a
// rendezvous and run as one batched sort (segmented launches across the
// requests' contexts), which leaves every request's results unchanged.
//
// The request operations are the library's EncodedTable::sort
// (encoded_table.h:900-998) and EncodedTable::aggregate with its defaults
// (encoded_table.h:1059-1188: sort on valid + key, pad to a power of two,
// distinct, forward aggregation network, unpad, filter on the distinct bits).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"

struct orqg_server {
  orqg_server_config cfg;
  std::vector<orqg_ctx *> ctxs;
  orqg::SortBatcher batcher;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  struct Pending {
    orqg_request req;
    double t_submit;
    size_t need;
  };
  std::deque<Pending> queue;
  std::deque<orqg_completion> done;
  size_t reserved = 0;
  int running = 0;
  bool stop = false;
  orqg_server_stats stats{};
  std::chrono::steady_clock::time_point t0;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace orqg {
namespace {

// Estimated device bytes of one request of n rows (the sharded permutations
// of its radix sorts dominate: ~70 of them x 3 groups x 2 directions x 4 B).
size_t request_bytes(const orqg_request &r) {
  const size_t n = r.n;
  size_t pad = 1;
  while (pad < n) pad <<= 1;
  return n * 1800 + pad * 6 * r.width * 3 + (64u << 20);
}

struct TableGuard {
  orqg_et *t = nullptr;
  ~TableGuard() {
    if (t) orqg_et_free(t);
  }
};

#define ORQG_CALL(expr)                                         \
  do {                                                          \
    int rc_ = (expr);                                           \
    if (rc_) throw ::orqg::Error(rc_, orqg_last_error(c));      \
  } while (0)

// The request as the reference's table calls make it: a table {[K], V} with
// its valid column, then EncodedTable::sort on [K] or EncodedTable::aggregate
// on [K] with V summed (both with their default options and the DEFAULT
// protocol, radix sort on the server's contexts).
void run_request(orqg_ctx *c, const orqg_request &r) {
  ORQG_CALL(orqg_ctx_rekey(c, r.keys));
  const char *names[2] = {"[K]", "V"};
  TableGuard g;
  ORQG_CALL(orqg_et_create(c, r.width, r.n, names, 2, &g.t));
  ORQG_CALL(orqg_vec_write(c, orqg_et_column(g.t, "[K]"), r.key_shares));
  ORQG_CALL(orqg_vec_write(c, orqg_et_column(g.t, "V"), r.val_shares));
  const char *keys[1] = {"[K]"};
  if (r.op == ORQG_REQ_SORT) {
    ORQG_CALL(orqg_et_sort(g.t, keys, nullptr, 1, ORQG_SORT_DEFAULT));
  } else if (r.op == ORQG_REQ_GROUPBY_SUM) {
    const orqg_et_agg sum = {"V", "V", ORQG_AGG_SUM};
    ORQG_CALL(orqg_et_aggregate(g.t, keys, 1, &sum, 1, 0, 1, 1, nullptr));
  } else {
    throw Error(ORQG_E_ARG, "request: unknown op");
  }
  if (r.out_key) ORQG_CALL(orqg_vec_read(c, orqg_et_column(g.t, "[K]"), r.out_key));
  if (r.out_val) ORQG_CALL(orqg_vec_read(c, orqg_et_column(g.t, "V"), r.out_val));
  if (r.out_valid) ORQG_CALL(orqg_vec_read(c, orqg_et_column(g.t, "[##V]"), r.out_valid));
}

void worker(orqg_server *s, int slot) {
  orqg_ctx *c = s->ctxs[slot];
  (void)hipSetDevice(c->device);
  for (;;) {
    orqg_server::Pending p;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        if (s->queue.empty()) return false;
        // the head is admitted when it fits next to the running requests
        // (alone it always runs: its footprint was checked at submission)
        return s->running == 0 || s->reserved + s->queue.front().need <= s->cfg.hbm_budget;
      });
      if (s->stop && s->queue.empty()) return;
      if (s->queue.empty()) continue;
      p = s->queue.front();
      s->queue.pop_front();
      s->reserved += p.need;
      s->running++;
      s->stats.max_running = std::max<uint64_t>(s->stats.max_running, s->running);
      s->stats.peak_reserved = std::max<uint64_t>(s->stats.peak_reserved, s->reserved);
    }
    orqg_completion done{};
    done.id = p.req.id;
    done.t_submit_ms = p.t_submit;
    done.t_start_ms = s->now_ms();
    auto note = [&](const char *m) { std::snprintf(done.msg, sizeof(done.msg), "%s", m); };
    s->batcher.enter();
    try {
      run_request(c, p.req);
      done.status = ORQG_OK;
    } catch (const Error &e) {
      done.status = e.code;
      note(e.what());
    } catch (const std::exception &e) {
      done.status = ORQG_E_DEVICE;
      note(e.what());
    } catch (...) {
      done.status = ORQG_E_DEVICE;
      note("unknown exception");
    }
    s->batcher.leave();
    if (done.status != ORQG_OK) {
      // leave the slot's context clean for the next request: wait for both
      // streams and read (and so clear) a pending swap-resolution check, so
      // that this failure is not reported again by the next request's rekey
      (void)orqg_sync(c);
      (void)hipStreamSynchronize(c->stream2);
      (void)hipGetLastError();
    }
    done.t_done_ms = s->now_ms();
    done.slot = slot;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->reserved -= p.need;
      s->running--;
      s->stats.completed++;
      if (done.status != ORQG_OK) s->stats.failed++;
      s->stats.peak_live_bytes = std::max<uint64_t>(s->stats.peak_live_bytes, live_bytes());
      s->done.push_back(done);
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
}

}  // namespace
}  // namespace orqg

extern "C" {

int orqg_server_create(const orqg_server_config *cfg, orqg_server **out) {
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 64) return ORQG_E_ARG;
  orqg_server *s = new orqg_server;
  s->cfg = *cfg;
  if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = (size_t)200 << 30;
  s->t0 = std::chrono::steady_clock::now();
  s->batcher.max_rows = cfg->batch_rows;
  s->batcher.window_us = std::max(0, cfg->batch_window_us);
  for (int i = 0; i < cfg->slots; i++) {
    orqg_config cc{};
    cc.device = cfg->device;
    cc.batch_size = cfg->batch_size;
    orqg_ctx *c = nullptr;
    int rc = orqg_ctx_create(&cc, &c);
    if (rc) {
      for (orqg_ctx *x : s->ctxs) orqg_ctx_destroy(x);
      delete s;
      return rc;
    }
    if (cfg->batch_rows) c->batcher = &s->batcher;
    // DEFAULT_SORT_PROTO=RADIXSORT (sorting.h:6-9): small radix sorts of
    // concurrent requests can then share batched launches
    c->default_sort = ORQG_RADIXSORT;
    s->ctxs.push_back(c);
  }
  for (int i = 0; i < cfg->slots; i++) s->workers.emplace_back(orqg::worker, s, i);
  *out = s;
  return ORQG_OK;
}

int orqg_server_submit(orqg_server *s, const orqg_request *req) {
  if (!s || !req || !req->key_shares || !req->val_shares || (req->width != 4 && req->width != 8))
    return ORQG_E_ARG;
  orqg_server::Pending p;
  p.req = *req;
  p.t_submit = s->now_ms();
  p.need = orqg::request_bytes(*req);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (p.need > s->cfg.hbm_budget) {
      orqg_completion d{};
      d.id = req->id;
      d.status = ORQG_E_BUDGET;
      d.t_submit_ms = d.t_start_ms = d.t_done_ms = p.t_submit;
      d.slot = -1;
      s->done.push_back(d);
      s->stats.rejected++;
    } else {
      s->queue.push_back(p);
      s->stats.submitted++;
    }
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  return ORQG_OK;
}

int orqg_server_poll(orqg_server *s, orqg_completion *out, int max, int *count, int wait_ms) {
  if (!s || !out || !count || max < 1) return ORQG_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *count = k;
  return ORQG_OK;
}

int orqg_server_pending(orqg_server *s, int *queued, int *running) {
  if (!s) return ORQG_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int)s->queue.size();
  if (running) *running = s->running;
  return ORQG_OK;
}

int orqg_server_get_stats(orqg_server *s, orqg_server_stats *st) {
  if (!s || !st) return ORQG_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  orqg::alloc_counters(&st->mallocs, &st->pool_trims);
  {
    std::lock_guard<std::mutex> lb(s->batcher.mu);
    st->sort_batches = s->batcher.batches;
    st->sorts_batched = s->batcher.batched;
  }
  return ORQG_OK;
}

double orqg_server_now_ms(orqg_server *s) { return s ? s->now_ms() : 0.0; }

int orqg_server_destroy(orqg_server *s) {
  if (!s) return ORQG_E_ARG;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // drain: queued requests still run
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (orqg_ctx *c : s->ctxs) orqg_ctx_destroy(c);
  delete s;
  return ORQG_OK;
}

}  // extern "C"
