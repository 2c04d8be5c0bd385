// Internal definitions shared by the translation units of liborqg.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/orqg.h"
#include "aes.h"

namespace orqg {

// Largest request the reference's AESPRGAlgorithm hands to one AES-GCM call
// (prg_algorithm.h:14-15): each such chunk consumes one nonce.
constexpr uint64_t kChunkBytes = 1ull << 20;
constexpr uint32_t kBlocksPerChunk = (uint32_t)(kChunkBytes / 16);

// Key roles, see orqg.h.
constexpr int kLocal0 = 0, kPair0 = 3, kGroup0 = 6;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ORQG_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::orqg::Error(ORQG_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// ---- runtime batch schedule -------------------------------------------------
// RunTime::getThreadBatchBoundaries + Task::execute with one worker thread
// (include/backend/common/runtime.h:198-247, task.h:34-44,89-94).
inline uint64_t batch_elems(uint64_t n, long batch_size) {
  if (n == 0) return 1;
  if (batch_size < 0) {
    uint64_t nb = (uint64_t)(-batch_size);
    uint64_t chunks = (n + 255) / 256;
    return (chunks + nb - 1) / nb * 256;
  }
  return (uint64_t)batch_size;
}

// Keystream positions of one runtime call: n elements of `tsz` bytes, batches
// of `bs` elements, and `calls` getNext calls per batch on each PRG involved.
// Inside batch b the j-th call starts at nonce offset b*calls*cfull + j*c(b),
// where c(b) = ceil(|b| * tsz / 1 MiB).
struct CallSched {
  uint64_t n, bs, cfull;
  uint32_t tsz, calls;
};

inline CallSched make_sched(uint64_t n, long batch_size, uint32_t tsz, uint32_t calls) {
  CallSched s;
  s.n = n;
  s.bs = batch_elems(n, batch_size);
  s.tsz = tsz;
  s.calls = calls;
  s.cfull = (s.bs * tsz + kChunkBytes - 1) / kChunkBytes;
  return s;
}

// nonces consumed on each PRG by the whole call
inline uint64_t sched_nonces(const CallSched &s) {
  if (s.n == 0) return 0;
  uint64_t nb = (s.n + s.bs - 1) / s.bs;
  uint64_t last = s.n - (nb - 1) * s.bs;
  return (nb - 1) * s.calls * s.cfull + (uint64_t)s.calls * ((last * s.tsz + kChunkBytes - 1) / kChunkBytes);
}

// Locate element e's bytes for call j: nonce offset, CTR counter and the byte
// offset inside the 16-byte block.
__device__ __host__ inline void sched_locate(const CallSched &s, uint64_t e, uint32_t j,
                                             uint64_t &nonce_off, uint32_t &ctr,
                                             uint32_t &byte_in_block) {
  uint64_t b = e / s.bs;
  uint64_t l = e - b * s.bs;
  uint64_t size_b = s.n - b * s.bs;
  if (size_b > s.bs) size_b = s.bs;
  uint64_t cb = (size_b * s.tsz + kChunkBytes - 1) / kChunkBytes;
  uint64_t byte = l * s.tsz;
  nonce_off = b * s.calls * s.cfull + j * cb + (byte >> 20);
  uint32_t in_chunk = (uint32_t)(byte & (kChunkBytes - 1));
  ctr = 2 + (in_chunk >> 4);
  byte_in_block = in_chunk & 15;
}

// ---- device memory pool --------------------------------------------------------
struct Scratch {
  void *p = nullptr;
  size_t bytes = 0;
};

// A context's free device blocks (caching allocator, guarded by the allocator
// lock): by size class for reuse, and by age for trimming the oldest first.
struct Pool {
  std::multimap<size_t, std::pair<void *, uint64_t>> by_size;  // class -> (block, age)
  std::map<uint64_t, std::pair<size_t, void *>> by_age;        // age -> (class, block)
  size_t bytes = 0;
};

struct ShardedPerm {
  size_t n = 0;
  int32_t *fwd[3] = {nullptr, nullptr, nullptr};  // pi_g      (scatter form)
  int32_t *inv[3] = {nullptr, nullptr, nullptr};  // pi_g^{-1} (gather form)
  hipEvent_t ready = nullptr;  // recorded after its swap resolution (second stream)
  // one device allocation holds every permutation of a reserve call; it goes
  // back to the allocator when the last of them is freed
  std::shared_ptr<void> slab;
};

struct Timer {
  double ms = 0;
  uint64_t launches = 0;
};

struct SortBatcher;

}  // namespace orqg

struct orqg_ctx {
  int device = 0;
  int default_sort = ORQG_QUICKSORT;  // SortingProtocol::DEFAULT (sorting.h:6-9)
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // auxiliary stream (same default priority)
  hipEvent_t ev_aux = nullptr;
  // pinned host staging of the quicksort rounds (hipHostMalloc; grown on
  // demand, freed with the context)
  void *host_pinned = nullptr;
  size_t host_pinned_bytes = 0;
  long batch_size = -12;
  size_t budget = 0;
  uint64_t nonce[ORQG_NUM_KEYS] = {0};
  uint8_t keys[ORQG_NUM_KEYS][32];
  orqg::AesTables tabs;
  uint32_t *d_rk = nullptr;  // ORQG_NUM_KEYS x 60 round-key words of the current worker thread
  // worker threads (the reference's RunTime threads): the round keys of all
  // of them, and the nonces of those not current (`nonce` holds the current
  // thread's; outside a thread scope that is thread 0)
  int threads = 1;
  uint32_t *d_rk_all = nullptr;
  std::vector<std::array<uint64_t, ORQG_NUM_KEYS>> tnonce;
  int cur_thread = 0;
  int scope_depth = 0;
  uint32_t *d_t0 = nullptr;  // 256 words
  std::string err;
  std::vector<orqg_msg> log;
  std::map<std::string, orqg::Scratch> scratch;
  orqg::Pool pool;  // free device blocks of this context
  std::deque<orqg::ShardedPerm> queue;
  // a served context's small radix sorts join those of the server's other
  // slots (null: the context sorts on its own)
  orqg::SortBatcher *batcher = nullptr;
  uint32_t *fy_err = nullptr;       // swap-resolution check flags, read by check_swaps
  bool fy_pending = false;
  size_t stored_size = 0;
  int num_cu = 256;
  bool timing = false;
  std::map<std::string, orqg::Timer> timers;
  // timed regions record event pairs from a pool without waiting; the
  // elapsed times are folded into `timers` at the next sync / timer read
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct PendingTimer {
    std::string fam;
    size_t a, b;
    uint64_t launches;  // kernel launches the region holds
  };
  std::vector<PendingTimer> pending;
  size_t timer_event();
  void timer_flush();

  void *scratch_get(const std::string &name, size_t bytes);
  void log_msg(const char *op, int from, int to, uint64_t bytes);
  const uint32_t *rk(int role) const { return d_rk + 60 * role; }
};

struct orqg_vec {
  orqg_ctx *ctx = nullptr;
  size_t n = 0;
  int width = 4;
  int enc = ORQG_BSHARED;
  void *d = nullptr;  // [n][3] elements of `width` bytes
};

// rows [off, off + m) of v, as a vector of its own
inline orqg_vec vec_rows(const orqg_vec *v, uint64_t off, uint64_t m) {
  orqg_vec s = *v;
  s.n = m;
  s.d = (char *)v->d + off * 3 * (uint64_t)v->width;
  return s;
}

namespace orqg {

// Allocation that counts against the HBM budget (a caching allocator per context).
void *dev_alloc(orqg_ctx *c, size_t bytes);
void dev_free(void *p);
size_t live_bytes();
// allocator counters since load: hipMalloc calls, pooled blocks released by
// trimming (a context's pool over its cap, or every pool after a failed
// hipMalloc / an allocation over the budget)
void alloc_counters(uint64_t *mallocs, uint64_t *trimmed);

// Scoped timing of one kernel family on the context stream: two events
// bracket the region's launches; nothing waits for them here.
struct TimedRegion {
  orqg_ctx *c;
  const char *fam;
  size_t a = 0;
  uint64_t launches;
  TimedRegion(orqg_ctx *c_, const char *f, uint64_t nl = 1) : c(c_), fam(f), launches(nl) {
    if (c->timing) {
      a = c->timer_event();
      ORQG_HIP(hipEventRecord(c->ev_pool[a], c->stream));
    }
  }
  ~TimedRegion() noexcept(false) {
    if (c->timing) {
      size_t b = c->timer_event();
      ORQG_HIP(hipEventRecord(c->ev_pool[b], c->stream));
      c->pending.push_back({fam, a, b, launches});
    }
  }
};

// ---- worker threads ------------------------------------------------------------
// RunTime::getThreadBatchBoundaries (include/backend/common/runtime.h:198-247):
// whole batches of the batch size (256 elements when it is negative) dealt to
// the threads in contiguous runs, the first threads one extra when they do not
// divide evenly, the last thread also the partial batch at the end.  Each
// thread's Task then batches its own range (task.h:34-44,89-94), which is
// make_sched on the range's length.
inline std::vector<std::pair<uint64_t, uint64_t>> thread_bounds(uint64_t n, long batch_size,
                                                                int threads) {
  const uint64_t B = batch_size < 0 ? 256 : (uint64_t)batch_size;
  const uint64_t whole = n / B, per = whole / (uint64_t)threads;
  uint64_t extra = whole % (uint64_t)threads, start = 0;
  std::vector<std::pair<uint64_t, uint64_t>> out;
  for (int t = 0; t < threads; t++) {
    uint64_t m = per * B;
    if (extra > 0) {
      m += B;
      extra--;
    }
    out.push_back({start, start + m});
    start += m;
  }
  out.back().second += n % B;
  return out;
}

// Makes worker thread t current: its nonces in c->nonce, its round keys at c->d_rk.
struct ThreadScope {
  orqg_ctx *c;
  int prev;
  ThreadScope(orqg_ctx *c_, int t) : c(c_), prev(c_->cur_thread) {
    c->scope_depth++;
    if (t == prev) return;
    for (int k = 0; k < ORQG_NUM_KEYS; k++) c->tnonce[prev][k] = c->nonce[k];
    for (int k = 0; k < ORQG_NUM_KEYS; k++) c->nonce[k] = c->tnonce[t][k];
    c->d_rk = c->d_rk_all + (size_t)t * ORQG_NUM_KEYS * 60;
    c->cur_thread = t;
  }
  ~ThreadScope() {
    c->scope_depth--;
    const int t = c->cur_thread;
    if (t == prev) return;
    for (int k = 0; k < ORQG_NUM_KEYS; k++) c->tnonce[t][k] = c->nonce[k];
    for (int k = 0; k < ORQG_NUM_KEYS; k++) c->nonce[k] = c->tnonce[prev][k];
    c->d_rk = c->d_rk_all + (size_t)prev * ORQG_NUM_KEYS * 60;
    c->cur_thread = prev;
  }
};

// f(offset, count) once per worker thread with a non-empty range of a runtime
// call over n elements, with that thread current (one call, f(0, n), for a
// single thread)
template <class F>
void for_threads(orqg_ctx *c, uint64_t n, F &&f) {
  if (n == 0) return;
  if (c->threads <= 1) {
    f((uint64_t)0, n);
    return;
  }
  if (c->scope_depth) throw Error(ORQG_E_PARITY, "internal: nested worker-thread ranges");
  const auto b = thread_bounds(n, c->batch_size, c->threads);
  for (int t = 0; t < c->threads; t++) {
    if (b[t].second == b[t].first) continue;
    ThreadScope ts(c, t);
    f(b[t].first, b[t].second - b[t].first);
  }
}

inline void require_one_thread(const orqg_ctx *c, const char *what) {
  if (c->threads > 1)
    throw Error(ORQG_E_ARG, std::string(what) +
                                ": more than one worker thread is not supported here (it is "
                                "for secret sharing, and_b, multiply_a, b2a_bit, reshare, "
                                "permutations, radix and table sort, comparisons, b2a, the "
                                "adders and a2b)");
}

inline unsigned grid_for(uint64_t items, unsigned block, unsigned cap) {
  uint64_t g = (items + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---- entry points implemented in the .hip files -------------------------------
void prg_keystream(orqg_ctx *c, int role, size_t bytes, void *out);
void secret_share(orqg_ctx *c, const void *host, orqg_vec *v, int owner = 0);
void open_vec(orqg_ctx *c, const orqg_vec *v, void *host_out);
void elementwise_local(orqg_ctx *c, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
void mult(orqg_ctx *c, bool arith, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
void b2a_bit(orqg_ctx *c, const orqg_vec *x, orqg_vec *y);
void reshare(orqg_ctx *c, orqg_vec *v, int group);
void reserve_perms(orqg_ctx *c, size_t n, size_t count);
// reserve_perms of several distinct contexts on one device; the small sizes
// are generated together
void reserve_perms_multi(const std::vector<orqg_ctx *> &cs, const std::vector<size_t> &ns,
                         const std::vector<size_t> &counts);
ShardedPerm next_perm(orqg_ctx *c, size_t n);
void free_perm(ShardedPerm &p);
// the main stream waits (on the device) until p's swap resolution is done
void perm_wait(orqg_ctx *c, ShardedPerm &p);
// swap resolution runs on the second stream after reserve_perms returns; its
// device-side check is read here (waits for that stream), throws if it failed
void check_swaps(orqg_ctx *c);
void apply_sharded(orqg_ctx *c, orqg_vec *x, const ShardedPerm &pi, bool inverse);
void apply_elementwise(orqg_ctx *c, orqg_vec *x, const orqg_vec *perm);
void radix_sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out);
void radix_sort_batch(const std::vector<orqg_ctx *> &cs, const std::vector<orqg_vec *> &vs,
                      const std::vector<int> &orders, const std::vector<int> &bits,
                      const std::vector<orqg_vec *> &perms);

// Rendezvous of concurrent radix sorts (the serving loop's slots): a sort of
// at most max_rows rows waits until every active request is waiting here or
// window_us has passed since the batch's first sort arrived, and the batch
// then runs as one radix_sort_batch (each request's results unchanged).
struct SortBatcher {
  size_t max_rows = 0;
  int window_us = 0;
  std::mutex mu;
  std::condition_variable cv;
  int active = 0;  // requests running on the slots
  struct Job {
    orqg_ctx *c;
    orqg_vec *v, *perm;
    int order, bits;
    bool done = false;
    std::exception_ptr err;
  };
  std::vector<Job *> waiting;
  bool leading = false;
  uint64_t batches = 0, batched = 0;
  void enter();
  void leave();
  void sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out);
};
void compare(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *eq, orqg_vec *gt,
             bool full, int neg_eq, int neg_gt);
void b2a_full(orqg_ctx *c, const orqg_vec *x, orqg_vec *y);
void perm_identity(orqg_ctx *c, orqg_vec *perm, int enc);
void public_fill_raw(orqg_ctx *c, void *d, uint64_t m, uint64_t value, int w);
void public_fill(orqg_ctx *c, orqg_vec *v, uint64_t value);
void broadcast_row(orqg_ctx *c, const orqg_vec *src, uint64_t row, orqg_vec *dst,
                   uint64_t dst_start, uint64_t count);
void aggregate(orqg_ctx *c, orqg_vec *const *keys, size_t nk, const orqg_agg_spec *bs, size_t nbs,
               const orqg_agg_spec *as, size_t nas, int reverse, const orqg_vec *sel);
void distinct(orqg_ctx *c, orqg_vec *const *keys, size_t nk, orqg_vec *res);
// the allocation-free distinct: its one scratch buffer is a fixed static
// device array of kNoAllocScratchBytes (set up per device by noalloc_setup,
// which context creation calls with the device current)
constexpr size_t kNoAllocScratchBytes = (size_t)48 << 20;
void noalloc_setup(int device);
void distinct_noalloc(orqg_ctx *c, orqg_vec *const *keys, size_t nk, orqg_vec *res);
void adder(orqg_ctx *c, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
void a2b(orqg_ctx *c, const orqg_vec *x, orqg_vec *y);
void perm_b2a(orqg_ctx *c, orqg_vec *perm);
void perm_negate(orqg_ctx *c, orqg_vec *perm);
void perm_reverse(orqg_ctx *c, orqg_vec *perm);
void perm_invert(orqg_ctx *c, orqg_vec *perm);
void perm_compose(orqg_ctx *c, orqg_vec *sigma, orqg_vec *rho);
void table_shuffle(orqg_ctx *c, orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols,
                   size_t nb);
void bitonic_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nk, const int *orders,
                  orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols, size_t nb);
void table_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, const int *orders,
                const int *single_bit, orqg_vec *const *a_cols, size_t na,
                orqg_vec *const *b_cols, size_t nb, int protocol = ORQG_RADIXSORT);
void quicksort(orqg_ctx *c, orqg_vec *v, int order, orqg_vec *perm_out);
void vec_read_share_files(orqg_ctx *c, orqg_vec *v, const char *const files[3]);
void vec_write_share_files(orqg_ctx *c, const orqg_vec *v, const char *const files[3]);
void table_read_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                          orqg_vec *const *cols, size_t ncols);
void table_write_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                           orqg_vec *const *cols, size_t ncols);
void table_read_csv_data(orqg_ctx *c, const char *path, const char *const *names,
                         orqg_vec *const *cols, size_t ncols, int input_party);

enum LocalOp { kXor = 0, kAdd = 1, kSub = 2, kNot = 3 };
void vec_unary(orqg_ctx *c, int op, const orqg_vec *x, uint64_t arg, orqg_vec *z);
void prefix_sum(orqg_ctx *c, orqg_vec *v);

}  // namespace orqg
