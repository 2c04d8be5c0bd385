// extern "C" entry points of liborqg (declared in include/orqg.h) and the
// context / vector / message-log plumbing behind them.
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>
#include <unordered_map>
#include <unordered_set>

#include "internal.h"

using namespace orqg;

size_t orqg_ctx::timer_event() {
  // the pool only grows inside one API call; it is recycled at the next
  // sync / timer read, when no region is open
  if (ev_used == ev_pool.size()) {
    hipEvent_t e;
    ORQG_HIP(hipEventCreate(&e));
    ev_pool.push_back(e);
  }
  return ev_used++;
}

void orqg_ctx::timer_flush() {
  if (pending.empty()) {
    ev_used = 0;
    return;
  }
  ORQG_HIP(hipEventSynchronize(ev_pool[pending.back().b]));
  for (const auto &t : pending) {
    float ms = 0;
    ORQG_HIP(hipEventElapsedTime(&ms, ev_pool[t.a], ev_pool[t.b]));
    auto &tm = timers[t.fam];
    tm.ms += ms;
    tm.launches += t.launches;
  }
  pending.clear();
  ev_used = 0;
}

void *orqg_ctx::scratch_get(const std::string &name, size_t bytes) {
  Scratch &s = scratch[name];
  if (s.bytes < bytes) {
    if (s.p) dev_free(s.p);
    s.p = nullptr;
    s.bytes = 0;
    size_t want = bytes + bytes / 8;
    s.p = dev_alloc(this, want);
    s.bytes = want;
  }
  return s.p;
}

// Byte accounting per (op, from, to), first-seen order: the reference's
// Communicator counts bytes per send (communicator.h:22-33).
void orqg_ctx::log_msg(const char *op, int from, int to, uint64_t bytes) {
  if (bytes == 0) return;
  for (auto &m : log)
    if (m.from == from && m.to == to && std::strcmp(m.op, op) == 0) {
      m.bytes += bytes;
      return;
    }
  orqg_msg m;
  std::memset(&m, 0, sizeof(m));
  std::strncpy(m.op, op, sizeof(m.op) - 1);
  m.from = from;
  m.to = to;
  m.bytes = bytes;
  log.push_back(m);
}

namespace orqg {
// Device memory: a caching allocator.  Blocks come in size classes (four per
// power of two); a freed block goes back to the pool of the context that
// allocated it and is reused by that context's later allocations.  All of a
// context's device work is ordered on its stream, so a block reused by the
// same context is never touched by earlier work still in flight, and no
// hipFree (which waits for the whole device) sits on the steady-state path.
// g_live_bytes counts everything held, pooled blocks included, and is what a
// context's HBM budget is checked against.
//
// A pool holds at most kPoolCap bytes: a served stream of many sizes would
// otherwise keep one set of blocks per size class it has seen.  Over the cap
// the oldest blocks are released.  Released blocks are handed to hipFree only
// after the allocator lock is dropped: hipFree waits for all device work, and
// other contexts' threads must not queue behind that wait for the lock.
struct Block {
  size_t bytes;
  orqg_ctx *owner;
};
static size_t g_live_bytes = 0;
static std::unordered_map<void *, Block> g_blocks;
static std::unordered_set<orqg_ctx *> g_alive;
static std::mutex g_alloc_mu;
static uint64_t g_age = 0, g_mallocs = 0, g_trimmed = 0;
constexpr size_t kPoolCap = (size_t)24 << 30;

size_t live_bytes() {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  return g_live_bytes;
}

void alloc_counters(uint64_t *mallocs, uint64_t *trimmed) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  if (mallocs) *mallocs = g_mallocs;
  if (trimmed) *trimmed = g_trimmed;
}

static size_t size_class(size_t b) {
  if (b <= 256) return 256;
  const int k = 63 - __builtin_clzll((unsigned long long)(b - 1));  // 2^k < b <= 2^(k+1)
  const size_t step = (size_t)1 << (k - 2);
  return (b + step - 1) / step * step;
}

// take the oldest pooled block of c out of the books (lock held); the caller
// hipFrees it after dropping the lock
static void *evict_oldest_locked(orqg_ctx *c) {
  Pool &P = c->pool;
  auto it = P.by_age.begin();
  const size_t cls = it->second.first;
  void *p = it->second.second;
  auto range = P.by_size.equal_range(cls);
  for (auto j = range.first; j != range.second; ++j)
    if (j->second.first == p) {
      P.by_size.erase(j);
      break;
    }
  P.by_age.erase(it);
  P.bytes -= cls;
  g_live_bytes -= cls;
  g_blocks.erase(p);
  g_trimmed++;
  return p;
}

// every pooled block of every context (lock held)
static void take_all_pools_locked(std::vector<void *> &out) {
  for (orqg_ctx *c : g_alive)
    while (!c->pool.by_age.empty()) out.push_back(evict_oldest_locked(c));
}

static void free_blocks(const std::vector<void *> &v) {
  for (void *p : v) (void)hipFree(p);
}

void register_ctx(orqg_ctx *c) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_alive.insert(c);
}

// The context's pooled blocks are released; blocks it still lends out (vectors
// the caller has not freed) are released by hipFree when they are freed.
void unregister_ctx(orqg_ctx *c) {
  std::vector<void *> victims;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    while (!c->pool.by_age.empty()) victims.push_back(evict_oldest_locked(c));
    for (auto &kv : g_blocks)
      if (kv.second.owner == c) kv.second.owner = nullptr;
    g_alive.erase(c);
  }
  free_blocks(victims);
}

void *dev_alloc(orqg_ctx *c, size_t bytes) {
  const size_t cls = size_class(bytes ? bytes : 16);
  std::vector<void *> victims;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    Pool &P = c->pool;
    auto it = P.by_size.find(cls);
    if (it != P.by_size.end()) {
      void *p = it->second.first;
      P.by_age.erase(it->second.second);
      P.by_size.erase(it);
      P.bytes -= cls;
      return p;
    }
    if (c->budget && g_live_bytes + cls > c->budget) take_all_pools_locked(victims);
  }
  free_blocks(victims);
  victims.clear();
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (c->budget && g_live_bytes + cls > c->budget)
      throw Error(ORQG_E_BUDGET, "allocation exceeds the configured HBM budget");
  }
  void *p = nullptr;
  if (hipMalloc(&p, cls) != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(g_alloc_mu);
      take_all_pools_locked(victims);
    }
    free_blocks(victims);
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      throw Error(ORQG_E_OOM, "hipMalloc of " + std::to_string(cls) + " bytes failed");
    }
  }
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_mallocs++;
  g_live_bytes += cls;
  g_blocks[p] = {cls, c};
  return p;
}

void dev_free(void *p) {
  if (!p) return;
  std::vector<void *> victims;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_blocks.find(p);
    if (it == g_blocks.end()) {
      victims.push_back(p);  // not ours (never expected)
    } else {
      orqg_ctx *owner = it->second.owner;
      const size_t cls = it->second.bytes;
      if (owner && g_alive.count(owner)) {
        Pool &P = owner->pool;
        const uint64_t age = ++g_age;
        auto pos = P.by_size.emplace(cls, std::make_pair(p, age));
        (void)pos;
        P.by_age.emplace(age, std::make_pair(cls, p));
        P.bytes += cls;
        while (P.bytes > kPoolCap) victims.push_back(evict_oldest_locked(owner));
      } else {
        g_live_bytes -= cls;
        g_blocks.erase(it);
        victims.push_back(p);
      }
    }
  }
  free_blocks(victims);
}
}  // namespace orqg

#define API_BEGIN(c) \
  try {              \
    (void)(c);
#define API_END(c)                                      \
  return ORQG_OK;                                       \
  }                                                     \
  catch (const orqg::Error &e) {                        \
    if (c) (c)->err = e.what();                         \
    return e.code;                                      \
  }                                                     \
  catch (const std::exception &e) {                     \
    if (c) (c)->err = e.what();                         \
    return ORQG_E_DEVICE;                               \
  }

extern "C" {

int orqg_ctx_create(const orqg_config *cfg, orqg_ctx **out) {
  if (!cfg || !out) return ORQG_E_ARG;
  orqg_ctx *c = new orqg_ctx;
  try {
    c->device = cfg->device;
    c->batch_size = cfg->batch_size == 0 ? -12 : cfg->batch_size;
    if (c->batch_size > 0 && c->batch_size % 256)
      throw Error(ORQG_E_ARG, "ctx_create: a positive batch size must be a multiple of 256");
    c->budget = cfg->hbm_budget;
    std::memcpy(c->keys, cfg->keys, sizeof(c->keys));
    ORQG_HIP(hipSetDevice(c->device));
    noalloc_setup(c->device);
    register_ctx(c);
    ORQG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    ORQG_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    ORQG_HIP(hipEventCreateWithFlags(&c->ev_aux, hipEventDisableTiming));
    hipDeviceProp_t prop;
    ORQG_HIP(hipGetDeviceProperties(&prop, c->device));
    c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    aes_make_tables(&c->tabs);
    // the round keys of every worker thread: thread 0's keys, then those of
    // threads 1.. from thread_keys
    const int T = cfg->threads > 1 ? cfg->threads : 1;
    if (T > 64) throw Error(ORQG_E_ARG, "ctx_create: at most 64 worker threads");
    if (T > 1 && !cfg->thread_keys) throw Error(ORQG_E_ARG, "ctx_create: thread_keys missing");
    c->threads = T;
    c->tnonce.assign(T, std::array<uint64_t, ORQG_NUM_KEYS>{});
    std::vector<uint32_t> rk((size_t)T * ORQG_NUM_KEYS * 60);
    for (int t = 0; t < T; t++)
      for (int k = 0; k < ORQG_NUM_KEYS; k++)
        aes256_expand(c->tabs,
                      t == 0 ? c->keys[k] : cfg->thread_keys + ((size_t)(t - 1) * ORQG_NUM_KEYS + k) * 32,
                      rk.data() + ((size_t)t * ORQG_NUM_KEYS + k) * 60);
    c->d_rk_all = (uint32_t *)dev_alloc(c, rk.size() * 4);
    c->d_rk = c->d_rk_all;
    c->d_t0 = (uint32_t *)dev_alloc(c, sizeof(c->tabs.t0));  // T0; kernels replicate it in LDS
    // copies on the context's own stream, never the legacy null stream: a
    // null-stream operation would be ordered against every other context's
    // streams (and is refused while any of them is capturing a graph)
    ORQG_HIP(hipMemcpyAsync(c->d_rk_all, rk.data(), rk.size() * 4, hipMemcpyHostToDevice, c->stream));
    ORQG_HIP(hipMemcpyAsync(c->d_t0, c->tabs.t0, sizeof(c->tabs.t0), hipMemcpyHostToDevice,
                            c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));
  } catch (const Error &e) {
    std::fprintf(stderr, "orqg_ctx_create: %s\n", e.what());
    unregister_ctx(c);
    delete c;
    return e.code;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "orqg_ctx_create: %s\n", e.what());
    unregister_ctx(c);
    delete c;
    return ORQG_E_DEVICE;
  }
  *out = c;
  return ORQG_OK;
}

int orqg_ctx_destroy(orqg_ctx *c) {
  if (!c) return ORQG_E_ARG;
  (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto &kv : c->scratch) dev_free(kv.second.p);
  for (auto &p : c->queue) free_perm(p);
  dev_free(c->d_rk_all);
  dev_free(c->d_t0);
  if (c->host_pinned) (void)hipHostFree(c->host_pinned);
  unregister_ctx(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_aux) (void)hipEventDestroy(c->ev_aux);
  if (c->stream2) {
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamDestroy(c->stream2);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ORQG_OK;
}

int orqg_ctx_rekey(orqg_ctx *c, const uint8_t keys[ORQG_NUM_KEYS][32]) {
  API_BEGIN(c)
  if (!keys) throw Error(ORQG_E_ARG, "ctx_rekey: NULL keys");
  require_one_thread(c, "ctx_rekey");
  ORQG_HIP(hipStreamSynchronize(c->stream));
  check_swaps(c);
  ORQG_HIP(hipStreamSynchronize(c->stream2));
  std::memcpy(c->keys, keys, sizeof(c->keys));
  uint32_t rk[ORQG_NUM_KEYS * 60];
  for (int k = 0; k < ORQG_NUM_KEYS; k++) aes256_expand(c->tabs, c->keys[k], rk + 60 * k);
  ORQG_HIP(hipMemcpyAsync(c->d_rk_all, rk, sizeof(rk), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < ORQG_NUM_KEYS; k++) c->nonce[k] = 0;
  for (auto &p : c->queue) free_perm(p);
  c->queue.clear();
  c->stored_size = 0;
  c->log.clear();
  c->timer_flush();
  c->timers.clear();
  API_END(c)
}

int orqg_ctx_set_default_sort(orqg_ctx *c, int protocol) {
  API_BEGIN(c)
  if (protocol != ORQG_QUICKSORT && protocol != ORQG_RADIXSORT)
    throw Error(ORQG_E_ARG, "ctx_set_default_sort: QUICKSORT or RADIXSORT");
  c->default_sort = protocol;
  API_END(c)
}

int orqg_ctx_default_sort(orqg_ctx *c) { return c ? c->default_sort : ORQG_E_ARG; }

const char *orqg_last_error(orqg_ctx *c) { return c ? c->err.c_str() : "null context"; }

#ifndef ORQG_SOURCE_HASH
#define ORQG_SOURCE_HASH "unknown"
#endif
const char *orqg_build_info(void) { return ORQG_SOURCE_HASH; }

int orqg_sync(orqg_ctx *c) {
  API_BEGIN(c)
  ORQG_HIP(hipStreamSynchronize(c->stream));
  check_swaps(c);
  c->timer_flush();
  API_END(c)
}

void *orqg_stream(orqg_ctx *c) { return c ? (void *)c->stream : nullptr; }

int orqg_prg_nonce(orqg_ctx *c, int role, uint64_t *nonce) {
  if (!c || role < 0 || role >= ORQG_NUM_KEYS || !nonce) return ORQG_E_ARG;
  *nonce = c->nonce[role];
  return ORQG_OK;
}

int orqg_prg_nonce_thread(orqg_ctx *c, int thread, int role, uint64_t *nonce) {
  if (!c || thread < 0 || thread >= c->threads || role < 0 || role >= ORQG_NUM_KEYS || !nonce)
    return ORQG_E_ARG;
  *nonce = thread == c->cur_thread ? c->nonce[role] : c->tnonce[thread][role];
  return ORQG_OK;
}

int orqg_dev_alloc(orqg_ctx *c, size_t bytes, void **dev_out) {
  API_BEGIN(c)
  if (!dev_out) throw Error(ORQG_E_ARG, "dev_alloc: null output");
  *dev_out = dev_alloc(c, bytes);
  API_END(c)
}

int orqg_dev_free(void *dev) {
  dev_free(dev);
  return ORQG_OK;
}

int orqg_dev_to_host(orqg_ctx *c, void *host, const void *dev, size_t bytes) {
  API_BEGIN(c)
  ORQG_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  check_swaps(c);
  API_END(c)
}

int orqg_vec_create(orqg_ctx *c, size_t n, int width, int encoding, orqg_vec **out) {
  API_BEGIN(c)
  if (!out || (width != 4 && width != 8)) throw Error(ORQG_E_ARG, "vec_create: width must be 4 or 8");
  orqg_vec *v = new orqg_vec;
  v->ctx = c;
  v->n = n;
  v->width = width;
  v->enc = encoding ? ORQG_BSHARED : ORQG_ASHARED;
  try {
    v->d = dev_alloc(c, 3 * n * (size_t)width + 64);
    ORQG_HIP(hipMemsetAsync(v->d, 0, 3 * n * (size_t)width, c->stream));
  } catch (...) {
    delete v;
    throw;
  }
  *out = v;
  API_END(c)
}

// EVector::resize (e_vector.h): keep the first min(n, old) rows; new rows
// hold zero shares.
int orqg_vec_resize(orqg_ctx *c, orqg_vec *v, size_t n) {
  API_BEGIN(c)
  if (!v) throw Error(ORQG_E_ARG, "vec_resize: NULL vector");
  if (n == v->n) return ORQG_OK;
  const size_t w = v->width;
  void *d = dev_alloc(c, 3 * n * w + 64);
  const size_t keep = n < v->n ? n : v->n;
  ORQG_HIP(hipMemsetAsync(d, 0, 3 * n * w, c->stream));
  ORQG_HIP(hipMemcpyAsync(d, v->d, 3 * keep * w, hipMemcpyDeviceToDevice, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  dev_free(v->d);
  v->d = d;
  v->n = n;
  API_END(c)
}
// EncodedTable::tail (encoded_table.h:1407-1418): keep the last n rows.
int orqg_vec_tail(orqg_ctx *c, orqg_vec *v, size_t n) {
  API_BEGIN(c)
  if (!v || n > v->n) throw Error(ORQG_E_ARG, "vec_tail: n exceeds the vector size");
  if (n == v->n) return ORQG_OK;
  const size_t w = v->width;
  void *d = dev_alloc(c, 3 * n * w + 64);
  ORQG_HIP(hipMemcpyAsync(d, (char *)v->d + 3 * (v->n - n) * w, 3 * n * w,
                          hipMemcpyDeviceToDevice, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  dev_free(v->d);
  v->d = d;
  v->n = n;
  API_END(c)
}

// dst rows [dst_start, +count) = src rows [src_start, +count)
// (EncodedTable::copy_column_typed, encoded_table.h:1749-1754)
int orqg_vec_copy_rows(orqg_ctx *c, const orqg_vec *src, size_t src_start, orqg_vec *dst,
                       size_t dst_start, size_t count) {
  API_BEGIN(c)
  if (!src || !dst || src->width != dst->width || src_start + count > src->n ||
      dst_start + count > dst->n)
    throw Error(ORQG_E_ARG, "vec_copy_rows: range outside the vectors");
  const size_t w = src->width;
  if (count)
    ORQG_HIP(hipMemcpyAsync((char *)dst->d + 3 * dst_start * w,
                            (const char *)src->d + 3 * src_start * w, 3 * count * w,
                            hipMemcpyDeviceToDevice, c->stream));
  API_END(c)
}
// dst rows [dst_start, +count) = the three shares of src row `row`
// (a repeated_subset_reference assignment, mapping_access_vector.h:854-866)
int orqg_vec_broadcast_row(orqg_ctx *c, const orqg_vec *src, size_t row, orqg_vec *dst,
                           size_t dst_start, size_t count) {
  API_BEGIN(c)
  if (!src || !dst || src->width != dst->width || row >= src->n || dst_start + count > dst->n)
    throw Error(ORQG_E_ARG, "vec_broadcast_row: range outside the vectors");
  broadcast_row(c, src, row, dst, dst_start, count);
  API_END(c)
}

// The block goes back to its context's pool (or to hipFree when the context
// is gone); the context itself is not touched here.
int orqg_vec_free(orqg_vec *v) {
  if (!v) return ORQG_E_ARG;
  dev_free(v->d);
  delete v;
  return ORQG_OK;
}

size_t orqg_vec_size(const orqg_vec *v) { return v ? v->n : 0; }
int orqg_vec_width(const orqg_vec *v) { return v ? v->width : 0; }
void *orqg_vec_data(orqg_vec *v) { return v ? v->d : nullptr; }

// host [3][n] <-> device [n][3]
int orqg_vec_write(orqg_ctx *c, orqg_vec *v, const void *host) {
  API_BEGIN(c)
  size_t n = v->n, w = v->width;
  std::vector<uint8_t> aos(3 * n * w);
  const uint8_t *h = (const uint8_t *)host;
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) std::memcpy(&aos[(3 * i + k) * w], h + (k * n + i) * w, w);
  ORQG_HIP(hipMemcpyAsync(v->d, aos.data(), aos.size(), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int orqg_vec_read(orqg_ctx *c, const orqg_vec *v, void *host) {
  API_BEGIN(c)
  size_t n = v->n, w = v->width;
  std::vector<uint8_t> aos(3 * n * w);
  ORQG_HIP(hipMemcpyAsync(aos.data(), v->d, aos.size(), hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  check_swaps(c);
  uint8_t *h = (uint8_t *)host;
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) std::memcpy(h + (k * n + i) * w, &aos[(3 * i + k) * w], w);
  API_END(c)
}

int orqg_vec_copy(orqg_ctx *c, const orqg_vec *src, orqg_vec *dst) {
  API_BEGIN(c)
  if (src->n != dst->n || src->width != dst->width) throw Error(ORQG_E_ARG, "vec_copy: shape");
  ORQG_HIP(hipMemcpyAsync(dst->d, src->d, 3 * src->n * (size_t)src->width,
                          hipMemcpyDeviceToDevice, c->stream));
  dst->enc = src->enc;
  API_END(c)
}

int orqg_prg_keystream(orqg_ctx *c, int role, size_t bytes, void *dev_out) {
  API_BEGIN(c)
  prg_keystream(c, role, bytes, dev_out);
  API_END(c)
}

int orqg_secret_share(orqg_ctx *c, const void *host_plain, orqg_vec *v) {
  API_BEGIN(c)
  secret_share(c, host_plain, v);
  API_END(c)
}

int orqg_open(orqg_ctx *c, const orqg_vec *v, void *host_out) {
  API_BEGIN(c)
  open_vec(c, v, host_out);
  check_swaps(c);
  API_END(c)
}

int orqg_public_share(orqg_ctx *c, const void *host_plain, orqg_vec *v) {
  API_BEGIN(c)
  size_t n = v->n, w = v->width;
  std::vector<uint8_t> sh(3 * n * w, 0);
  std::memcpy(sh.data(), host_plain, n * w);  // s_0 = x, s_1 = s_2 = 0
  int rc = orqg_vec_write(c, v, sh.data());
  if (rc) return rc;
  API_END(c)
}

static void check3(const orqg_vec *x, const orqg_vec *y, const orqg_vec *z) {
  if (!x || !z || (y && (y->n != x->n || y->width != x->width)) || z->n != x->n ||
      z->width != x->width)
    throw Error(ORQG_E_ARG, "operand shapes differ");
}

int orqg_xor_b(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  elementwise_local(c, kXor, x, y, z);
  API_END(c)
}
int orqg_add_a(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  elementwise_local(c, kAdd, x, y, z);
  API_END(c)
}
int orqg_sub_a(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  elementwise_local(c, kSub, x, y, z);
  API_END(c)
}
int orqg_not_b(orqg_ctx *c, const orqg_vec *x, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, nullptr, z);
  elementwise_local(c, kNot, x, nullptr, z);
  API_END(c)
}
int orqg_and_b(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  mult(c, false, x, y, z);
  API_END(c)
}
int orqg_mul_a(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  mult(c, true, x, y, z);
  API_END(c)
}
int orqg_b2a_bit(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  API_BEGIN(c)
  check3(x, nullptr, y);
  b2a_bit(c, x, y);
  y->enc = ORQG_ASHARED;
  API_END(c)
}
int orqg_compare(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *eq, orqg_vec *gt) {
  API_BEGIN(c)
  check3(x, y, eq ? eq : gt ? gt : x);
  if (gt) check3(x, y, gt);
  if (x->enc != ORQG_BSHARED || y->enc != ORQG_BSHARED)
    throw Error(ORQG_E_ARG, "compare: operands must be boolean-shared");
  compare(c, x, y, eq, gt, true, 0, 0);
  if (eq) eq->enc = ORQG_BSHARED;
  if (gt) gt->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_cmp(orqg_ctx *c, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  if (!y) throw Error(ORQG_E_ARG, "cmp: missing operand");
  if (x->enc != ORQG_BSHARED || y->enc != ORQG_BSHARED)
    throw Error(ORQG_E_ARG, "cmp: operands must be boolean-shared");
  switch (op) {
    case ORQG_EQ: compare(c, x, y, z, nullptr, false, 0, 0); break;
    case ORQG_NE: compare(c, x, y, z, nullptr, false, 1, 0); break;
    case ORQG_GT: compare(c, x, y, nullptr, z, true, 0, 0); break;
    case ORQG_LT: compare(c, y, x, nullptr, z, true, 0, 0); break;
    case ORQG_LE: compare(c, x, y, nullptr, z, true, 0, 1); break;
    case ORQG_GE: compare(c, y, x, nullptr, z, true, 0, 1); break;
    default: throw Error(ORQG_E_ARG, "cmp: op");
  }
  z->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_b2a(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  API_BEGIN(c)
  check3(x, nullptr, y);
  if (x == y) throw Error(ORQG_E_ARG, "b2a: output must not alias the input");
  if (x->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "b2a: input must be boolean-shared");
  b2a_full(c, x, y);
  y->enc = ORQG_ASHARED;
  API_END(c)
}
static void check_b(const orqg_vec *x) {
  if (x->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "operand must be boolean-shared");
}
int orqg_add_b(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  check_b(x);
  check_b(y);
  adder(c, 0, x, y, z);
  z->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_sub_b(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  check_b(x);
  check_b(y);
  adder(c, 1, x, y, z);
  z->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_ppa(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, int subtract, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  check_b(x);
  check_b(y);
  adder(c, subtract ? 3 : 2, x, y, z);
  z->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_rca_compare(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  API_BEGIN(c)
  check3(x, y, z);
  check_b(x);
  check_b(y);
  adder(c, 4, x, y, z);
  z->enc = ORQG_BSHARED;
  API_END(c)
}
int orqg_a2b(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  API_BEGIN(c)
  check3(x, nullptr, y);
  if (x->enc != ORQG_ASHARED) throw Error(ORQG_E_ARG, "a2b: input must be arithmetic-shared");
  if (x == y) throw Error(ORQG_E_ARG, "a2b: output must not alias the input");
  a2b(c, x, y);
  API_END(c)
}
int orqg_secret_share_from(orqg_ctx *c, const void *host, orqg_vec *v, int owner) {
  API_BEGIN(c)
  if (!v || !host || owner < 0 || owner > 2) throw Error(ORQG_E_ARG, "secret_share_from");
  secret_share(c, host, v, owner);
  API_END(c)
}
int orqg_reshare(orqg_ctx *c, orqg_vec *v, int group) {
  API_BEGIN(c)
  reshare(c, v, group);
  API_END(c)
}

static void check_perm(const orqg_vec *p) {
  if (!p || p->width != 4) throw Error(ORQG_E_ARG, "permutation must be a width-4 vector");
}
static void check_cols(orqg_vec *const *cols, size_t k, size_t n) {
  if (k && !cols) throw Error(ORQG_E_ARG, "column array is NULL");
  for (size_t i = 0; i < k; i++)
    if (!cols[i] || cols[i]->n != n) throw Error(ORQG_E_ARG, "columns differ in size");
}

int orqg_perm_identity(orqg_ctx *c, orqg_vec *perm, int enc) {
  API_BEGIN(c)
  check_perm(perm);
  if (enc != ORQG_ASHARED && enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "encoding");
  perm_identity(c, perm, enc);
  API_END(c)
}
int orqg_perm_invert(orqg_ctx *c, orqg_vec *perm) {
  API_BEGIN(c)
  check_perm(perm);
  perm_invert(c, perm);
  ORQG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}
int orqg_perm_b2a(orqg_ctx *c, orqg_vec *perm) {
  API_BEGIN(c)
  check_perm(perm);
  perm_b2a(c, perm);
  API_END(c)
}
int orqg_perm_negate(orqg_ctx *c, orqg_vec *perm) {
  API_BEGIN(c)
  check_perm(perm);
  perm_negate(c, perm);
  API_END(c)
}
int orqg_perm_reverse(orqg_ctx *c, orqg_vec *perm) {
  API_BEGIN(c)
  check_perm(perm);
  perm_reverse(c, perm);
  API_END(c)
}
int orqg_perm_compose(orqg_ctx *c, orqg_vec *sigma, orqg_vec *rho) {
  API_BEGIN(c)
  check_perm(sigma);
  check_perm(rho);
  perm_compose(c, sigma, rho);
  API_END(c)
}
int orqg_shuffle(orqg_ctx *c, orqg_vec *x) {
  API_BEGIN(c)
  if (!x) throw Error(ORQG_E_ARG, "shuffle: NULL vector");
  if (x->n > 0) {
    ShardedPerm pi = next_perm(c, x->n);
    try {
      apply_sharded(c, x, pi, false);
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      free_perm(pi);
      throw;
    }
    ORQG_HIP(hipStreamSynchronize(c->stream));
    free_perm(pi);
  }
  API_END(c)
}
int orqg_table_shuffle(orqg_ctx *c, orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols,
                       size_t nb) {
  API_BEGIN(c)
  size_t n = na ? a_cols[0]->n : nb ? b_cols[0]->n : 0;
  check_cols(a_cols, na, n);
  check_cols(b_cols, nb, n);
  table_shuffle(c, a_cols, na, b_cols, nb);
  API_END(c)
}
int orqg_table_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, const int *orders,
                    const int *single_bit, orqg_vec *const *a_cols, size_t na,
                    orqg_vec *const *b_cols, size_t nb) {
  API_BEGIN(c)
  if (nkeys == 0 || !keys || !keys[0]) throw Error(ORQG_E_ARG, "table_sort: no sort keys");
  const size_t n = keys[0]->n;
  check_cols(keys, nkeys, n);
  check_cols(a_cols, na, n);
  check_cols(b_cols, nb, n);
  for (size_t j = 0; j < nkeys; j++)
    if (keys[j]->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "table_sort: keys must be B-shared");
  table_sort(c, keys, nkeys, orders, single_bit, a_cols, na, b_cols, nb);
  API_END(c)
}
int orqg_table_sort_protocol(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, const int *orders,
                             const int *single_bit, orqg_vec *const *a_cols, size_t na,
                             orqg_vec *const *b_cols, size_t nb, int protocol) {
  API_BEGIN(c)
  if (nkeys == 0 || !keys || !keys[0]) throw Error(ORQG_E_ARG, "table_sort: no sort keys");
  const size_t n = keys[0]->n;
  check_cols(keys, nkeys, n);
  check_cols(a_cols, na, n);
  check_cols(b_cols, nb, n);
  for (size_t j = 0; j < nkeys; j++)
    if (keys[j]->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "table_sort: keys must be B-shared");
  table_sort(c, keys, nkeys, orders, single_bit, a_cols, na, b_cols, nb, protocol);
  API_END(c)
}
static void check_files(const char *const f[3]) {
  if (!f || !f[0] || !f[1] || !f[2]) throw Error(ORQG_E_ARG, "csv: three party file paths needed");
}
int orqg_vec_read_share_files(orqg_ctx *c, orqg_vec *v, const char *const files[3]) {
  API_BEGIN(c)
  if (!v) throw Error(ORQG_E_ARG, "vec_read_share_files: NULL vector");
  check_files(files);
  vec_read_share_files(c, v, files);
  API_END(c)
}
int orqg_vec_write_share_files(orqg_ctx *c, const orqg_vec *v, const char *const files[3]) {
  API_BEGIN(c)
  if (!v) throw Error(ORQG_E_ARG, "vec_write_share_files: NULL vector");
  check_files(files);
  vec_write_share_files(c, v, files);
  API_END(c)
}
int orqg_table_read_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                              orqg_vec *const *cols, size_t ncols) {
  API_BEGIN(c)
  check_files(files);
  if (ncols && (!names || !cols)) throw Error(ORQG_E_ARG, "table_read_share_csv: NULL columns");
  check_cols(cols, ncols, ncols ? cols[0]->n : 0);
  table_read_share_csv(c, files, names, cols, ncols);
  API_END(c)
}
int orqg_table_write_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                               orqg_vec *const *cols, size_t ncols) {
  API_BEGIN(c)
  check_files(files);
  if (ncols && (!names || !cols)) throw Error(ORQG_E_ARG, "table_write_share_csv: NULL columns");
  check_cols(cols, ncols, ncols ? cols[0]->n : 0);
  table_write_share_csv(c, files, names, cols, ncols);
  API_END(c)
}
int orqg_table_read_csv_data(orqg_ctx *c, const char *path, const char *const *names,
                             orqg_vec *const *cols, size_t ncols, int input_party) {
  API_BEGIN(c)
  if (!path || (ncols && (!names || !cols))) throw Error(ORQG_E_ARG, "table_read_csv_data: NULL");
  check_cols(cols, ncols, ncols ? cols[0]->n : 0);
  table_read_csv_data(c, path, names, cols, ncols, input_party);
  API_END(c)
}
int orqg_quicksort(orqg_ctx *c, orqg_vec *v, int order, orqg_vec *perm_out) {
  API_BEGIN(c)
  if (!v || v->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "quicksort: v must be B-shared");
  quicksort(c, v, order, perm_out);
  API_END(c)
}

int orqg_bitonic_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, const int *orders,
                      orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols, size_t nb) {
  API_BEGIN(c)
  if (nkeys == 0 || !keys || !keys[0]) throw Error(ORQG_E_ARG, "bitonic_sort: no sort keys");
  const size_t n = keys[0]->n;
  check_cols(keys, nkeys, n);
  check_cols(a_cols, na, n);
  check_cols(b_cols, nb, n);
  for (size_t j = 0; j < nkeys; j++)
    if (keys[j]->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "bitonic_sort: keys must be B-shared");
  bitonic_sort(c, keys, nkeys, orders, a_cols, na, b_cols, nb);
  API_END(c)
}

int orqg_aggregate(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys,
                   const orqg_agg_spec *b_specs, size_t nb, const orqg_agg_spec *a_specs,
                   size_t na, int reverse, const orqg_vec *sel) {
  API_BEGIN(c)
  if ((nkeys && !keys) || (nb && !b_specs) || (na && !a_specs))
    throw Error(ORQG_E_ARG, "aggregate: NULL array");
  for (size_t k = 0; k < nkeys; k++)
    if (!keys[k] || keys[k]->enc != ORQG_BSHARED)
      throw Error(ORQG_E_ARG, "aggregate: keys must be B-shared");
  aggregate(c, keys, nkeys, b_specs, nb, a_specs, na, reverse, sel);
  API_END(c)
}
int orqg_distinct(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, orqg_vec *res) {
  API_BEGIN(c)
  if (!res || (nkeys && !keys)) throw Error(ORQG_E_ARG, "distinct: NULL argument");
  distinct(c, keys, nkeys, res);
  API_END(c)
}
int orqg_vec_unary(orqg_ctx *c, int op, const orqg_vec *x, uint64_t arg, orqg_vec *z) {
  API_BEGIN(c)
  vec_unary(c, op, x, arg, z);
  API_END(c)
}
int orqg_prefix_sum(orqg_ctx *c, orqg_vec *v) {
  API_BEGIN(c)
  prefix_sum(c, v);
  API_END(c)
}
int orqg_distinct_noalloc(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, orqg_vec *res) {
  API_BEGIN(c)
  if (!res || (nkeys && !keys)) throw Error(ORQG_E_ARG, "distinct_noalloc: NULL argument");
  distinct_noalloc(c, keys, nkeys, res);
  API_END(c)
}
int orqg_public_fill(orqg_ctx *c, orqg_vec *v, uint64_t value, int enc) {
  API_BEGIN(c)
  if (!v) throw Error(ORQG_E_ARG, "public_fill: NULL vector");
  public_fill(c, v, value);
  v->enc = enc ? ORQG_BSHARED : ORQG_ASHARED;
  API_END(c)
}

int orqg_reserve_perms(orqg_ctx *c, size_t n, size_t count) {
  API_BEGIN(c)
  reserve_perms(c, n, count);
  API_END(c)
}
size_t orqg_perm_queue_size(orqg_ctx *c) { return c ? c->queue.size() : 0; }

int orqg_peek_perm(orqg_ctx *c, size_t index, int group, int32_t *host_out) {
  API_BEGIN(c)
  if (index >= c->queue.size() || group < 0 || group > 2) throw Error(ORQG_E_ARG, "peek_perm");
  ShardedPerm &p = c->queue[index];
  perm_wait(c, p);
  ORQG_HIP(hipMemcpyAsync(host_out, p.fwd[group], p.n * 4, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  check_swaps(c);
  API_END(c)
}

int orqg_apply_sharded_perm(orqg_ctx *c, orqg_vec *x, int inverse) {
  API_BEGIN(c)
  ShardedPerm pi = next_perm(c, x->n);
  try {
    apply_sharded(c, x, pi, inverse != 0);
  } catch (...) {
    free_perm(pi);
    throw;
  }
  free_perm(pi);
  API_END(c)
}

int orqg_apply_elementwise_perm(orqg_ctx *c, orqg_vec *x, const orqg_vec *perm) {
  API_BEGIN(c)
  if (perm->n != x->n || perm->width != 4) throw Error(ORQG_E_ARG, "apply_elementwise: perm");
  apply_elementwise(c, x, perm);
  API_END(c)
}

int orqg_radix_sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out) {
  API_BEGIN(c)
  if (!perm_out || perm_out->n != v->n || perm_out->width != 4)
    throw Error(ORQG_E_ARG, "radix_sort: perm_out must be a width-4 vector of the same size");
  if (bits < 1 || bits > 8 * v->width) throw Error(ORQG_E_ARG, "radix_sort: bits");
  if (v->n >= (1ull << 31)) throw Error(ORQG_E_ARG, "radix_sort: at most 2^31-1 rows");
  radix_sort(c, v, order, bits, perm_out);
  perm_out->enc = ORQG_ASHARED;
  API_END(c)
}

int orqg_radix_sort_batch(orqg_ctx *const *ctxs, orqg_vec *const *vs, const int *orders,
                          const int *bits, orqg_vec *const *perms_out, size_t m) {
  orqg_ctx *c0 = (ctxs && m) ? ctxs[0] : nullptr;
  API_BEGIN(c0)
  if (m == 0) return ORQG_OK;
  if (!ctxs || !vs || !orders || !bits || !perms_out)
    throw Error(ORQG_E_ARG, "radix_sort_batch: NULL array");
  std::vector<orqg_ctx *> cs(ctxs, ctxs + m);
  std::vector<orqg_vec *> v(vs, vs + m), p(perms_out, perms_out + m);
  std::vector<int> o(orders, orders + m), b(bits, bits + m);
  std::set<orqg_ctx *> seen;
  for (size_t q = 0; q < m; q++) {
    const std::string at = "radix_sort_batch: request " + std::to_string(q) + ": ";
    if (!cs[q] || !v[q] || !p[q]) throw Error(ORQG_E_ARG, at + "NULL context or vector");
    if (!seen.insert(cs[q]).second) throw Error(ORQG_E_ARG, at + "a context appears twice");
    if (cs[q]->device != cs[0]->device) throw Error(ORQG_E_ARG, at + "contexts on different devices");
    if (v[q]->ctx != cs[q] || p[q]->ctx != cs[q])
      throw Error(ORQG_E_ARG, at + "vectors must belong to the request's context");
    if (p[q]->n != v[q]->n || p[q]->width != 4)
      throw Error(ORQG_E_ARG, at + "perm_out must be a width-4 vector of the same size");
    if (b[q] < 1 || b[q] > 8 * v[q]->width) throw Error(ORQG_E_ARG, at + "bits");
    if (v[q]->n >= (1ull << 31)) throw Error(ORQG_E_ARG, at + "at most 2^31-1 rows");
  }
  radix_sort_batch(cs, v, o, b, p);
  for (size_t q = 0; q < m; q++) p[q]->enc = ORQG_ASHARED;
  API_END(c0)
}

size_t orqg_msglog_count(orqg_ctx *c) { return c ? c->log.size() : 0; }
int orqg_msglog_get(orqg_ctx *c, size_t i, orqg_msg *out) {
  if (!c || !out || i >= c->log.size()) return ORQG_E_ARG;
  *out = c->log[i];
  return ORQG_OK;
}
int orqg_msglog_clear(orqg_ctx *c) {
  if (!c) return ORQG_E_ARG;
  c->log.clear();
  return ORQG_OK;
}

int orqg_timer_enable(orqg_ctx *c, int enable) {
  if (!c) return ORQG_E_ARG;
  c->timing = enable != 0;
  return ORQG_OK;
}
int orqg_timer_get(orqg_ctx *c, const char *family, double *ms, uint64_t *launches) {
  if (!c || !family) return ORQG_E_ARG;
  try {
    c->timer_flush();
  } catch (const Error &e) {
    c->err = e.what();
    return e.code;
  }
  auto it = c->timers.find(family);
  if (it == c->timers.end()) {
    if (ms) *ms = 0;
    if (launches) *launches = 0;
    return ORQG_OK;
  }
  if (ms) *ms = it->second.ms;
  if (launches) *launches = it->second.launches;
  return ORQG_OK;
}
int orqg_timer_reset(orqg_ctx *c) {
  if (!c) return ORQG_E_ARG;
  try {
    c->timer_flush();
  } catch (const Error &e) {
    c->err = e.what();
    return e.code;
  }
  c->timers.clear();
  return ORQG_OK;
}

}  // extern "C"
