// CPU restatement of ORQ's 3-party replicated (honest-majority) protocol path:
// AES PRG keystream, zero sharings, boolean/arithmetic primitives, b2a_bit,
// honest-majority sharded permutations, elementwise-permutation application
// and oblivious radix sort.
//
// THIS IS TEST INFRASTRUCTURE (the oracle). Only the tests, the smoke check and
// the CPU-baseline leg of the benchmark load it; the GPU library never does.
//
// Every party is simulated separately: party p holds the replicated pair
// (s_p, s_{p+1}) of every shared vector and its own copies of every PRG object,
// exactly as one reference process would.  Messages are copies between party
// structs and are logged with their byte counts.  Runtime batching follows the
// reference's schedule: a call over n elements is split into contiguous
// thread ranges (T worker threads, each with its own PRG objects), and each
// range into batches (batch_size < 0 => |b| equal batches of 256-multiples),
// because every batch makes its own PRG calls and therefore owns its own
// nonces.
//
// Parity pins (see tests/test_oracle.py):
//   AES keystream ........ AES-GCM spec test case 14 + a second KAT; OpenSSL CTR
//   zero sharings ........ XOR / sum over parties is 0
//   and/mul/b2a_bit/open . plaintext closed forms on seeded inputs
//   gen_perm ............. brute-force enumeration on tiny n; bijection
//   sharded perms ........ apply then inverse = identity, reshare keeps values
//   radix_sort ........... plaintext stable sort (signed order); perm applies
// and, beyond these, against the reference's own code compiled in place
// (ref_driver.cpp, ref_party.cpp; tests/test_ref_oracle.py): keystream,
// gen_perm, sharded permutations, batching, every primitive's shares, radix
// sort, quicksort, the comparison and adder circuits and the table operators
// share for share, with the bytes each party sends.
#include <openssl/evp.h>

#include <algorithm>
#include <array>
#include <cassert>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

typedef unsigned __int128 u128;
typedef __int128 i128;

// The three parties run on three threads where each one only touches its own
// shares and its own PRG objects (the reference runs each party as its own
// process with one worker thread): PARTIES_PARALLEL in front of a
// `for (int p = 0; p < 3; p++)` loop.  Results do not depend on it.
#define ORQ_PRAGMA(x) _Pragma(#x)
#define PARTIES_PARALLEL ORQ_PRAGMA(omp parallel for num_threads(3) schedule(static, 1))

namespace {

// ---------------------------------------------------------------- PRG ------
// AESPRGAlgorithm (include/core/random/prg/prg_algorithm.h:139-231):
// fillBytes splits a request into <=1 MiB chunks (:161-172); each chunk is one
// AES-256-GCM encryption of zeros under IV = LE64(nonce) || 0^4, after which
// the nonce increments (:214-231).  GCM's keystream starts at counter block
// IV || BE32(2), so a chunk equals AES-256-CTR from that block.
constexpr size_t kMaxAesQuery = 1u << 20;

struct Prg {
  uint8_t key[32];
  uint64_t nonce = 0;
  EVP_CIPHER_CTX *ctx = nullptr;

  Prg() {}
  explicit Prg(const uint8_t *k) { std::memcpy(key, k, 32); }
  Prg(const Prg &o) : nonce(o.nonce), ctx(nullptr) { std::memcpy(key, o.key, 32); }
  Prg &operator=(const Prg &o) {
    std::memcpy(key, o.key, 32);
    nonce = o.nonce;
    return *this;
  }
  ~Prg() {
    if (ctx) EVP_CIPHER_CTX_free(ctx);
  }

  // one aesGenerateValues call (prg_algorithm.h:214-231)
  void chunk(uint8_t *dst, size_t bytes) {
    static const uint8_t zeros[4096] = {0};
    uint8_t iv[16] = {0};
    for (int b = 0; b < 8; b++) iv[b] = (uint8_t)(nonce >> (8 * b));
    iv[15] = 2;
    if (!ctx) ctx = EVP_CIPHER_CTX_new();
    if (EVP_EncryptInit_ex(ctx, EVP_aes_256_ctr(), nullptr, key, iv) != 1)
      throw std::runtime_error("EVP init failed");
    size_t done = 0;
    while (done < bytes) {
      int len = (int)std::min(bytes - done, sizeof(zeros));
      int out = 0;
      EVP_EncryptUpdate(ctx, dst + done, &out, zeros, len);
      done += (size_t)len;
    }
    nonce++;
  }

  // fillBytes (prg_algorithm.h:161-172)
  void fill(uint8_t *dst, size_t bytes) {
    size_t idx = 0;
    while (idx < bytes) {
      size_t b = std::min(kMaxAesQuery, bytes - idx);
      chunk(dst + idx, b);
      idx += b;
    }
  }

  // PRGAlgorithm::getNext(Vector<T>) -> getNextBuffered with a 1 MiB buffer
  // (prg_algorithm.h:56-62, 87-104): one fillBytes per buffer-full of elements.
  template <typename T>
  void getNext(T *dst, size_t n) {
    const size_t per = kMaxAesQuery / sizeof(T);
    size_t idx = 0;
    while (idx < n) {
      size_t m = std::min(per, n - idx);
      fill(reinterpret_cast<uint8_t *>(dst + idx), m * sizeof(T));
      idx += m;
    }
  }
};

// ------------------------------------------------------- batch schedule ---
// RunTime::getThreadBatchBoundaries (include/backend/common/runtime.h:198-247):
// whole batches of the call's batch size (256 elements when it is negative)
// are dealt to the T threads in contiguous runs, the first threads taking one
// extra when they do not divide evenly, and the last thread also takes the
// partial batch at the end.
std::vector<std::pair<size_t, size_t>> thread_bounds(size_t n, long batch_size, int threads) {
  const size_t B = batch_size < 0 ? 256 : (size_t)batch_size;
  const size_t whole = n / B, per = whole / (size_t)threads;
  size_t extra = whole % (size_t)threads;
  std::vector<std::pair<size_t, size_t>> out;
  size_t start = 0;
  for (int t = 0; t < threads; t++) {
    size_t m = per * B;
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

// Each thread's Task splits its range into batches of
// compute_equal_batches(...) elements when batch_size < 0, else batch_size
// (include/backend/common/task.h:34-44, 89-94).
struct Batch {
  size_t s, e;
  int t;  // the worker thread whose PRG objects the batch draws from
};

std::vector<Batch> batches_for(size_t n, long batch_size, int threads = 1) {
  std::vector<Batch> out;
  if (n == 0) return out;
  const auto bounds = thread_bounds(n, batch_size, threads);
  for (int t = 0; t < threads; t++) {
    const size_t s0 = bounds[t].first, e0 = bounds[t].second;
    if (e0 == s0) continue;
    size_t bs;
    if (batch_size < 0) {
      size_t nb = (size_t)(-batch_size);
      size_t chunks = (e0 - s0 + 255) / 256;
      bs = (chunks + nb - 1) / nb * 256;
    } else {
      bs = (size_t)batch_size;
    }
    for (size_t i = s0; i < e0; i += bs) out.push_back({i, std::min(i + bs, e0), t});
  }
  return out;
}

// --------------------------------------------------------- message log ---
struct Msg {
  char op[24];
  int from, to;
  uint64_t bytes;
};

// ------------------------------------------------------------- parties ---
// Groups in the order generateGroups(3,2,1) yields them
// (include/core/protocols/protocol.h:309-375): {0,1},{0,2},{1,2}; index 3 is
// the everyone group {0,1,2}.
const int kGroups[4][3] = {{0, 1, -1}, {0, 2, -1}, {1, 2, -1}, {0, 1, 2}};
bool in_group(int g, int p) {
  for (int k = 0; k < 3; k++)
    if (kGroups[g][k] == p) return true;
  return false;
}

struct Party {
  int id;
  Prg local;     // localPRG: secret sharing (replicated_3pc.h:338-392)
  Prg prev;      // commonPRGManager->get(-1): key of party id-1
  Prg next;      // commonPRGManager->get(+1): this party's own pairwise key
  Prg group[4];  // group PRGs, valid only where in_group(g, id)
};

// Key roles (include/backend/common/rand_setup.h:90-146): party p generates
// local_key_p, sends it to p+1 and uses it for relative rank +1 (next); the key
// received from p-1 is used for relative rank -1 (prev).  Group keys come from
// the lowest member (:21-78).  Here they are given explicitly:
//   keys[0..2]  localPRG of party p
//   keys[3..5]  pairwise key K_p (next of p, prev of p+1)
//   keys[6..9]  group keys {0,1},{0,2},{1,2},{0,1,2}
}  // namespace

constexpr int kMaxThreads = 64;

struct oc_ctx {
  // the parties' PRG objects of each worker thread (setup_random_generation
  // per thread, include/backend/common/setup.h:48-57)
  std::vector<std::array<Party, 3>> TP;
  int threads = 1;
  long batch_size = -12;
  int gen_everyone = 0;  // also run gen_perm for the unused {0,1,2} group
  std::vector<Msg> log;
  // PermutationManager queue (permutation_manager.h:102-181): each entry is
  // one honest-majority sharded permutation, per party a map group->perm.
  struct Sharded {
    size_t n;
    std::vector<int32_t> perm[3][4];  // [party][group]
  };
  std::deque<Sharded> queue;
  size_t stored_size = 0;
};

struct oc_vec {
  size_t n;
  int width;  // bytes per element: 4, 8 or 16
  std::vector<uint8_t> sh[3][2];  // party p: (s_p, s_{p+1})
};

namespace {

// Byte accounting per (op, from, to) in first-seen order; the reference's
// Communicator counts bytes per send (communicator.h:22-33).
void log_msg(oc_ctx *c, const char *op, int from, int to, uint64_t bytes) {
  if (bytes == 0) return;
  for (auto &m : c->log)
    if (m.from == from && m.to == to && std::strcmp(m.op, op) == 0) {
      m.bytes += bytes;
      return;
    }
  Msg m;
  std::snprintf(m.op, sizeof(m.op), "%s", op);
  m.from = from;
  m.to = to;
  m.bytes = bytes;
  c->log.push_back(m);
}

template <typename T>
T *S(oc_vec *v, int p, int i) {
  return reinterpret_cast<T *>(v->sh[p][i].data());
}

oc_vec *new_vec(size_t n, int width) {
  oc_vec *v = new oc_vec;
  v->n = n;
  v->width = width;
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < 2; i++) v->sh[p][i].assign(n * width, 0);
  return v;
}

oc_vec *clone_vec(const oc_vec *x) { return new oc_vec(*x); }

// ZeroSharingGenerator::getNextBinary / getNextArithmetic on a Vector
// (include/core/random/correlation/zero_sharing_generator.h:102-147).
template <typename T>
void zero_binary(Party &pt, T *out, size_t n) {
  std::vector<T> a(n), b(n);
  pt.prev.getNext(a.data(), n);
  pt.next.getNext(b.data(), n);
  for (size_t i = 0; i < n; i++) out[i] = a[i] ^ b[i];
}
template <typename T>
void zero_arith(Party &pt, T *out, size_t n) {
  std::vector<T> b(n);
  pt.prev.getNext(out, n);
  pt.next.getNext(b.data(), n);
  for (size_t i = 0; i < n; i++) out[i] -= b[i];
}

// ---------------------------------------------------------- primitives ---
// Replicated_3PC::and_b (include/core/protocols/replicated_3pc.h:166-191) and
// multiply_a (:44-67), evaluated batch by batch as RunTime::eval_protocol_2arg
// does (runtime.h:690-702).  z may alias x or y.
template <typename T, bool kArith>
void mult_batch(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *z, size_t s, size_t e,
                const char *op, int t) {
  size_t nb = e - s;
  std::vector<T> local[3];
  PARTIES_PARALLEL
  for (int p = 0; p < 3; p++) {
    local[p].resize(nb);
    if (kArith)
      zero_arith<T>(c->TP[t][p], local[p].data(), nb);
    else
      zero_binary<T>(c->TP[t][p], local[p].data(), nb);
    const T *x0 = S<T>(x, p, 0) + s, *x1 = S<T>(x, p, 1) + s;
    const T *y0 = S<T>(y, p, 0) + s, *y1 = S<T>(y, p, 1) + s;
    for (size_t i = 0; i < nb; i++) {
      if (kArith)
        local[p][i] += x0[i] * y0[i] + x0[i] * y1[i] + x1[i] * y0[i];
      else
        local[p][i] ^= (x0[i] & y0[i]) ^ (x0[i] & y1[i]) ^ (x1[i] & y0[i]);
    }
  }
  // exchangeShares(local, remote, 2, 1): send to p-1, receive from p+1
  for (int p = 0; p < 3; p++) {
    std::memcpy(S<T>(z, p, 0) + s, local[p].data(), nb * sizeof(T));
    std::memcpy(S<T>(z, p, 1) + s, local[(p + 1) % 3].data(), nb * sizeof(T));
    log_msg(c, op, p, (p + 2) % 3, nb * sizeof(T));
  }
}

template <typename T, bool kArith>
void mult(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *z, const char *op) {
  for (auto b : batches_for(x->n, c->batch_size, c->threads))
    mult_batch<T, kArith>(c, x, y, z, b.s, b.e, op, b.t);
}

// Replicated_3PC::b2a_bit (replicated_3pc.h:220-251), per batch.
template <typename T>
void b2a_bit_batch(oc_ctx *c, oc_vec *x, oc_vec *y, size_t s, size_t e, int t) {
  size_t nb = e - s;
  oc_vec *s0 = new_vec(nb, sizeof(T)), *s1 = new_vec(nb, sizeof(T));
  std::vector<T> send0[3], send1[3];
  PARTIES_PARALLEL
  for (int p = 0; p < 3; p++) {
    T *a0 = S<T>(s0, p, 0), *a1 = S<T>(s0, p, 1), *b0 = S<T>(s1, p, 0);
    const T *x0 = S<T>(x, p, 0) + s, *x1 = S<T>(x, p, 1) + s;
    std::vector<T> lsb0(nb), lsb1(nb);
    for (size_t i = 0; i < nb; i++) {
      lsb0[i] = x0[i] & 1;
      lsb1[i] = x1[i] & 1;
      if (p == 0) lsb1[i] ^= lsb0[i];
    }
    zero_arith<T>(c->TP[t][p], a0, nb);
    zero_arith<T>(c->TP[t][p], b0, nb);
    for (size_t i = 0; i < nb; i++) {
      if (p == 0) a0[i] += lsb1[i];
      if (p == 1) b0[i] += lsb1[i];
    }
    (void)a1;
    send0[p].assign(a0, a0 + nb);
    send1[p].assign(b0, b0 + nb);
  }
  for (int p = 0; p < 3; p++) {
    std::memcpy(S<T>(s0, p, 1), send0[(p + 1) % 3].data(), nb * sizeof(T));
    std::memcpy(S<T>(s1, p, 1), send1[(p + 1) % 3].data(), nb * sizeof(T));
    log_msg(c, "b2a_bit", p, (p + 2) % 3, 2 * nb * sizeof(T));
  }
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a = S<T>(s0, p, k), *b = S<T>(s1, p, k);
      for (size_t i = 0; i < nb; i++) a[i] -= b[i];
    }
  // y = s0 * s0 (multiply_a on the same batch)
  oc_vec *out = new_vec(nb, sizeof(T));
  mult_batch<T, true>(c, s0, s0, out, 0, nb, "b2a_bit", t);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++)
      std::memcpy(S<T>(y, p, k) + s, S<T>(out, p, k), nb * sizeof(T));
  delete s0;
  delete s1;
  delete out;
}

template <typename T>
void b2a_bit(oc_ctx *c, oc_vec *x, oc_vec *y) {
  for (auto b : batches_for(x->n, c->batch_size, c->threads)) b2a_bit_batch<T>(c, x, y, b.s, b.e, b.t);
}

// open_shares_a / open_shares_b (replicated_3pc.h:310-330)
template <typename T, bool kArith>
std::vector<T> open_vec(oc_ctx *c, oc_vec *x) {
  std::vector<T> out(x->n);
  for (auto b : batches_for(x->n, c->batch_size, c->threads)) {
    for (int p = 0; p < 3; p++) log_msg(c, "open", p, (p + 2) % 3, (b.e - b.s) * sizeof(T));
  }
  // every party reconstructs the same value; compute party 0's view and
  // check the others agree.
  for (int p = 0; p < 3; p++) {
    const T *a = S<T>(x, p, 0), *b = S<T>(x, p, 1), *r = S<T>(x, (p + 1) % 3, 1);
    for (size_t i = 0; i < x->n; i++) {
      T v = kArith ? (T)(a[i] + b[i] + r[i]) : (T)(a[i] ^ b[i] ^ r[i]);
      if (p == 0)
        out[i] = v;
      else if (out[i] != v)
        throw std::runtime_error("open: parties disagree (inconsistent replicated shares)");
    }
  }
  return out;
}

// public_share (replicated_3pc.h:456-470): P0 (x,0), P1 (0,0), P2 (0,x).
template <typename T>
void public_share(oc_vec *v, const T *x) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) std::fill(v->sh[p][k].begin(), v->sh[p][k].end(), 0);
  std::memcpy(S<T>(v, 0, 0), x, v->n * sizeof(T));
  std::memcpy(S<T>(v, 2, 1), x, v->n * sizeof(T));
}

// secret_share_b / secret_share_a from party 0 (replicated_3pc.h:401-448),
// batched by eval_protocol_1arg_alloc (runtime.h:584-603).
template <typename T, bool kArith>
void secret_share(oc_ctx *c, const T *data, oc_vec *v, int owner = 0) {
  for (auto b : batches_for(v->n, c->batch_size, c->threads)) {
    size_t nb = b.e - b.s;
    std::vector<T> s1(nb), s2(nb), s3(nb);
    c->TP[b.t][owner].local.getNext(s1.data(), nb);
    c->TP[b.t][owner].local.getNext(s2.data(), nb);
    for (size_t i = 0; i < nb; i++)
      s3[i] = kArith ? (T)(data[b.s + i] - s1[i] - s2[i]) : (T)(data[b.s + i] ^ s1[i] ^ s2[i]);
    // the owner keeps (s1, s2), its successor gets (s2, s3), its predecessor (s3, s1)
    const std::vector<T> *parts[3][2] = {{&s1, &s2}, {&s2, &s3}, {&s3, &s1}};
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 2; k++)
        std::memcpy(S<T>(v, (owner + j) % 3, k) + b.s, parts[j][k]->data(), nb * sizeof(T));
    log_msg(c, "secret_share", owner, (owner + 1) % 3, 2 * nb * sizeof(T));
    log_msg(c, "secret_share", owner, (owner + 2) % 3, 2 * nb * sizeof(T));
  }
}

// Protocol::reshare (include/core/protocols/protocol.h:500-541) for 3PC,
// per batch.  Members draw rand[0], rand[1] from the group PRG, rand[2] is
// dependent (zero_sharing_generator.h:159-183); party p masks its shares
// (p, p+1) with rand[p], rand[p+1].  The outsider c then receives share c
// from c-1 and share c+1 from c+1 (generate_send/recv_assignment :108-209).
template <typename T, bool kArith>
void reshare(oc_ctx *c, oc_vec *v, int g) {
  int outsider = -1;
  for (int p = 0; p < 3; p++)
    if (!in_group(g, p)) outsider = p;
  for (auto b : batches_for(v->n, c->batch_size, c->threads)) {
    size_t nb = b.e - b.s;
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++) {
      if (p == outsider) continue;
      std::vector<T> r[3];
      for (int k = 0; k < 3; k++) r[k].assign(nb, 0);
      for (int k = 0; k < 2; k++) {
        c->TP[b.t][p].group[g].getNext(r[k].data(), nb);
        for (size_t i = 0; i < nb; i++) {
          if (kArith)
            r[2][i] -= r[k][i];
          else
            r[2][i] ^= r[k][i];
        }
      }
      for (int k = 0; k < 2; k++) {
        T *d = S<T>(v, p, k) + b.s;
        const T *m = r[(p + k) % 3].data();
        for (size_t i = 0; i < nb; i++) d[i] = kArith ? (T)(d[i] + m[i]) : (T)(d[i] ^ m[i]);
      }
    }
    int o = outsider, pm = (o + 2) % 3, pp = (o + 1) % 3;
    std::memcpy(S<T>(v, o, 0) + b.s, S<T>(v, pm, 1) + b.s, nb * sizeof(T));
    std::memcpy(S<T>(v, o, 1) + b.s, S<T>(v, pp, 0) + b.s, nb * sizeof(T));
    log_msg(c, "reshare", pm, o, nb * sizeof(T));
    log_msg(c, "reshare", pp, o, nb * sizeof(T));
  }
}

// ------------------------------------------------------- permutations ---
// gen_perm (include/core/random/permutations/hm_sharded_permutation_generator.h:70-108)
// Fisher-Yates driven by a 2^20-entry uint32 buffer refilled eagerly from the
// group PRG, rejection-sampled against the fixed mask 2^bit_width(n)-1.
void gen_perm(std::vector<int32_t> &perm, Prg &g, uint64_t *draws_out) {
  size_t size = perm.size();
  std::iota(perm.begin(), perm.end(), 0);
  const size_t buf = kMaxAesQuery;  // elements (uint32), i.e. 4 MiB
  std::vector<uint32_t> random(buf);
  g.getNext(random.data(), buf);
  int bits = 0;
  for (size_t t = size; t; t >>= 1) bits++;
  uint32_t mask = (uint32_t)((1ull << bits) - 1);
  size_t rand_index = 0;
  uint64_t draws = 0;
  for (size_t i = 0; i < size; i++) {
    uint32_t rnd = (uint32_t)size;
    while ((size_t)rnd > size - i - 1) {
      rnd = random[rand_index] & mask;
      rand_index++;
      draws++;
      if (rand_index == buf) {
        g.getNext(random.data(), buf);
        rand_index = 0;
      }
    }
    std::swap(perm[i], perm[i + rnd]);
  }
  if (draws_out) *draws_out = draws;
}

// PermutationManager::reserve -> RunTime::generate_permutations
// (runtime.h:749-788) -> HMShardedPermutationGenerator::generateBatch
// (hm_sharded_permutation_generator.h:178-194): the new permutations are
// dealt to the threads in contiguous runs (count / T each, the first
// count % T threads one more), queue order thread by thread; thread t makes
// each of its permutations from its own PRG objects, for each group
// containing the party (in group order), with gen_perm.
void reserve(oc_ctx *c, size_t n, size_t count) {
  if (!c->queue.empty() && c->stored_size != n) c->queue.clear();
  c->stored_size = n;
  if (count <= c->queue.size()) return;
  count -= c->queue.size();
  std::vector<oc_ctx::Sharded> fresh(count);
  std::vector<int> owner(count);
  {
    const size_t T = (size_t)c->threads, base = count / T, rem = count % T;
    size_t k = 0;
    for (size_t t = 0; t < T; t++)
      for (size_t j = 0; j < base + (t < rem ? 1 : 0); j++) owner[k++] = (int)t;
  }
  // each party generates its groups' permutations in queue order from its
  // own group PRGs
  PARTIES_PARALLEL
  for (int p = 0; p < 3; p++)
    for (size_t k = 0; k < count; k++) {
      oc_ctx::Sharded &sp = fresh[k];
      sp.n = n;
      for (int g = 0; g < 4; g++) {
        if (!in_group(g, p)) continue;
        if (g == 3 && !c->gen_everyone) continue;
        sp.perm[p][g].resize(n);
        gen_perm(sp.perm[p][g], c->TP[owner[k]][p].group[g], nullptr);
      }
    }
  for (auto &sp : fresh) c->queue.push_back(std::move(sp));
}

oc_ctx::Sharded next_perm(oc_ctx *c, size_t n) {
  if (c->queue.empty() || c->stored_size != n) reserve(c, n, 1);
  oc_ctx::Sharded s = std::move(c->queue.front());
  c->queue.pop_front();
  return s;
}

// local_apply_perm (include/core/operators/shuffle.h:113-133): new[pi[i]] = x[i]
template <typename T>
void scatter(T *x, const int32_t *pi, size_t n) {
  std::vector<T> t(n);
  for (size_t i = 0; i < n; i++) t[pi[i]] = x[i];
  std::memcpy(x, t.data(), n * sizeof(T));
}
// local_apply_inverse_perm (shuffle.h:250-270): new[i] = x[pi[i]]
template <typename T>
void gather(T *x, const int32_t *pi, size_t n) {
  std::vector<T> t(n);
  for (size_t i = 0; i < n; i++) t[i] = x[pi[i]];
  std::memcpy(x, t.data(), n * sizeof(T));
}

// hm_oblivious_apply_sharded_perm / _inverse_ (shuffle.h:349-412)
template <typename T, bool kArith>
void apply_sharded(oc_ctx *c, oc_vec *x, const oc_ctx::Sharded &sp, bool inverse) {
  for (int gi = 0; gi < 3; gi++) {
    int g = inverse ? 2 - gi : gi;
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++) {
      if (!in_group(g, p)) continue;
      for (int k = 0; k < 2; k++) {
        if (inverse)
          gather<T>(S<T>(x, p, k), sp.perm[p][g].data(), x->n);
        else
          scatter<T>(S<T>(x, p, k), sp.perm[p][g].data(), x->n);
      }
    }
    reshare<T, kArith>(c, x, g);
  }
}

// oblivious_apply_elementwise_perm (shuffle.h:687-706) for an int32 A- or
// B-shared permutation `perm` applied to x.
template <typename T, bool kArith>
void apply_elementwise(oc_ctx *c, oc_vec *x, oc_vec *perm, bool perm_arith) {
  oc_vec *cp = clone_vec(perm);
  oc_ctx::Sharded pi = next_perm(c, x->n);  // getNextPair: (pi, clone(pi))
  apply_sharded<T, kArith>(c, x, pi, false);
  std::vector<int32_t> opened;
  if (perm_arith) {
    apply_sharded<uint32_t, true>(c, cp, pi, false);
    auto o = open_vec<uint32_t, true>(c, cp);
    opened.assign(o.begin(), o.end());
  } else {
    apply_sharded<uint32_t, false>(c, cp, pi, false);
    auto o = open_vec<uint32_t, false>(c, cp);
    opened.assign(o.begin(), o.end());
  }
  for (size_t i = 0; i < x->n; i++)
    if (opened[i] < 0 || (size_t)opened[i] >= x->n)
      throw std::runtime_error("opened permutation out of range");
  PARTIES_PARALLEL
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) scatter<T>(S<T>(x, p, k), opened.data(), x->n);
  delete cp;
}

// ElementwisePermutation::b2a (include/core/containers/permutation.h:222-258)
void perm_b2a(oc_ctx *c, oc_vec *perm) {
  oc_vec *cp = clone_vec(perm);
  oc_ctx::Sharded pi = next_perm(c, perm->n);
  apply_sharded<uint32_t, false>(c, cp, pi, false);
  auto o = open_vec<uint32_t, false>(c, cp);
  public_share<uint32_t>(perm, o.data());
  apply_sharded<uint32_t, true>(c, perm, pi, true);
  delete cp;
}

// ElementwisePermutation::invert (permutation.h:205-212)
void perm_invert(oc_ctx *c, oc_vec *perm) {
  oc_vec *inv = new_vec(perm->n, 4);
  std::vector<uint32_t> id(perm->n);
  std::iota(id.begin(), id.end(), 0u);
  public_share<uint32_t>(inv, id.data());
  apply_elementwise<uint32_t, true>(c, inv, perm, true);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) perm->sh[p][k] = inv->sh[p][k];
  delete inv;
}

// multiply by the public constant -1, then add -1 (ElementwisePermutation::negate,
// permutation.h:268-280; the result aliases the permutation's storage).
void perm_negate(oc_ctx *c, oc_vec *perm) {
  oc_vec *m1 = new_vec(perm->n, 4);
  std::vector<uint32_t> neg(perm->n, 0xffffffffu);
  public_share<uint32_t>(m1, neg.data());
  mult<uint32_t, true>(c, perm, m1, perm, "mul_a");
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      uint32_t *a = S<uint32_t>(perm, p, k), *b = S<uint32_t>(m1, p, k);
      for (size_t i = 0; i < perm->n; i++) a[i] += b[i];
    }
  delete m1;
}

void reverse_vec(oc_vec *v) {
  size_t w = v->width;
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      auto &d = v->sh[p][k];
      for (size_t i = 0; i < v->n / 2; i++)
        std::swap_ranges(d.begin() + i * w, d.begin() + (i + 1) * w,
                         d.begin() + (v->n - 1 - i) * w);
    }
}

// ----------------------------------------------------- comparison circuits
typedef unsigned __int128 u128;

// std::popcount of a share word of any width (bit_xor's parity,
// mapping_access_vector.h:430-439)
template <typename T>
int popc(T x) {
  int c = 0;
  for (unsigned s = 0; s < 8 * sizeof(T); s += 64) c += __builtin_popcountll((unsigned long long)(x >> s));
  return c;
}

template <typename T>
void not_vec(oc_vec *v) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a = S<T>(v, p, k);
      for (size_t i = 0; i < v->n; i++) a[i] = ~a[i];
    }
}

// BSharedVector::bit_same (include/core/containers/b_shared_vector.h:164-178):
// e = ~(x ^ y); for l < log2(bits): e &= e >> 2^l (arithmetic shift of each share).
template <typename T>
oc_vec *bit_same(oc_ctx *c, oc_vec *x, oc_vec *y) {
  typedef typename std::make_signed<T>::type ST;
  const size_t n = x->n;
  oc_vec *e = new_vec(n, sizeof(T)), *t = new_vec(n, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(x, p, k), *b = S<T>(y, p, k);
      T *d = S<T>(e, p, k);
      for (size_t i = 0; i < n; i++) d[i] = ~(a[i] ^ b[i]);
    }
  int logb = 0;
  while ((1 << logb) < (int)(8 * sizeof(T))) logb++;  // bit_width(bits - 1)
  for (int l = 0; l < logb; l++) {
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        const T *a = S<T>(e, p, k);
        T *d = S<T>(t, p, k);
        for (size_t i = 0; i < n; i++) d[i] = (T)((ST)a[i] >> (1 << l));
      }
    mult<T, false>(c, e, t, e, "and_b");
  }
  delete t;
  return e;
}

// BSharedVector::_compare (include/core/operators/circuits.h:231-295).
template <typename T>
void compare(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *eq, oc_vec *gt) {
  typedef typename std::make_signed<T>::type ST;
  const size_t n = x->n;
  const int B = 8 * sizeof(T);
  const size_t nc = n / B + (n % B > 0);
  oc_vec *e = bit_same<T>(c, x, y);
  // gt = parity(((e >> 1) ^ e) & x)
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(e, p, k);
      T *d = S<T>(gt, p, k);
      for (size_t i = 0; i < n; i++) d[i] = (T)((ST)a[i] >> 1) ^ a[i];
    }
  mult<T, false>(c, gt, x, gt, "and_b");
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *d = S<T>(gt, p, k);
      for (size_t i = 0; i < n; i++) d[i] = (T)(popc<T>(d[i]) & 1);
    }
  // signed inputs: fix the sign bit on MSB-packed words
  oc_vec *s1 = new_vec(nc, sizeof(T)), *s2 = new_vec(nc, sizeof(T)), *r = new_vec(nc, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *xa = S<T>(x, p, k), *ya = S<T>(y, p, k), *g = S<T>(gt, p, k);
      T *a1 = S<T>(s1, p, k), *a2 = S<T>(s2, p, k), *ar = S<T>(r, p, k);
      for (size_t w = 0; w < nc; w++) {  // pack_from (mapping_access_vector.h:974-991)
        T v1 = 0, v2 = 0, vr = 0;
        for (int q = 0; q < B && w * B + q < n; q++) {
          v1 |= (T)((xa[w * B + q] >> (B - 1)) & 1) << q;
          v2 |= (T)((ya[w * B + q] >> (B - 1)) & 1) << q;
          vr |= (T)(g[w * B + q] & 1) << q;
        }
        a1[w] = v1;
        a2[w] = v2;
        ar[w] = vr;
      }
      // r ^= s2 ; s2 ^= s1
      for (size_t w = 0; w < nc; w++) {
        ar[w] ^= a2[w];
        a2[w] ^= a1[w];
      }
    }
  // s2 |= r  ->  s2 = ~(~s2 & ~r)  (BSharedVector::operator|=, b_shared_vector.h:452-456)
  not_vec<T>(s2);
  not_vec<T>(r);
  mult<T, false>(c, s2, r, s2, "and_b");
  not_vec<T>(s2);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a1 = S<T>(s1, p, k);
      const T *a2 = S<T>(s2, p, k);
      T *g = S<T>(gt, p, k);
      for (size_t w = 0; w < nc; w++) a1[w] ^= a2[w];  // s1 ^= s2
      for (size_t i = 0; i < n; i++)  // unpack_from(s1, 0), then mask(1)
        g[i] = (T)((a1[i / B] >> (i % B)) & 1);
    }
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(e, p, k);
      T *d = S<T>(eq, p, k);
      for (size_t i = 0; i < n; i++) d[i] = a[i] & 1;  // eq_bits.mask(1)
    }
  delete e;
  delete s1;
  delete s2;
  delete r;
}

// BSharedVector::b2a (b_shared_vector.h:296-329): sum over bits of
// b2a_bit(bit i) * 2^i, the scaling being a secure multiply_a by a public
// sharing of 2^i.
template <typename T>
void b2a_full(oc_ctx *c, oc_vec *x, oc_vec *y) {
  const size_t n = x->n;
  const int B = 8 * sizeof(T);
  oc_vec *bit = new_vec(n, sizeof(T)), *ba = new_vec(n, sizeof(T)), *pc = new_vec(n, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) std::fill(y->sh[p][k].begin(), y->sh[p][k].end(), 0);
  for (int i = 0; i < B; i++) {
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        const T *a = S<T>(x, p, k);
        T *d = S<T>(bit, p, k);
        for (size_t j = 0; j < n; j++) d[j] = (a[j] >> i) & 1;  // logical shift, mask(1)
      }
    b2a_bit<T>(c, bit, ba);
    if (i > 0) {
      std::vector<T> cst(n, (T)((T)1 << i));
      public_share<T>(pc, cst.data());
      mult<T, true>(c, ba, pc, ba, "mul_a");
    }
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        T *d = S<T>(y, p, k);
        const T *a = S<T>(ba, p, k);
        for (size_t j = 0; j < n; j++) d[j] += a[j];
      }
  }
  delete bit;
  delete ba;
  delete pc;
}

// BSharedVector comparison operators (b_shared_vector.h:493-580):
// op 0 ==, 1 !=, 2 <, 3 >, 4 <=, 5 >=.  `!` is not_b_1 (replicated_3pc.h:207):
// every share becomes !(s & 1).
template <typename T>
void cmp_op(oc_ctx *c, int op, oc_vec *x, oc_vec *y, oc_vec *z) {
  oc_vec *r;
  if (op == 0 || op == 1) {
    r = bit_same<T>(c, x, y);
  } else {
    oc_vec *eq = new_vec(x->n, sizeof(T));
    r = new_vec(x->n, sizeof(T));
    bool swap = (op == 2 || op == 5);  // x < y and x >= y are built on y > x
    compare<T>(c, swap ? y : x, swap ? x : y, eq, r);
    delete eq;
  }
  bool neg = (op == 1 || op == 4 || op == 5);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(r, p, k);
      T *d = S<T>(z, p, k);
      for (size_t i = 0; i < x->n; i++) d[i] = neg ? (T)((a[i] & 1) == 0) : (T)(a[i] & 1);
    }
  delete r;
}

// ---------------------------------------------------------- radix sort ---
// radix_sort (include/core/operators/radixsort.h:205-229) with pad_input /
// remove_padding (include/core/operators/sorting.h:431-495).  K is the key
// share type (uint32_t or uint64_t), PT the padded type (PadWidth, :409-411).
template <typename K, typename PT, typename SPT>
void radix_sort(oc_ctx *c, oc_vec *v, int bits, bool desc, oc_vec *perm_out) {
  const size_t n = v->n;
  const bool full_width = (int)(sizeof(K) * 8) == bits;
  // pad_input
  oc_vec *pad = new_vec(n, sizeof(PT));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const K *src = S<K>(v, p, k);
      PT *dst = S<PT>(pad, p, k);
      for (size_t i = 0; i < n; i++) {
        typedef typename std::make_signed<K>::type SK;
        dst[i] = (PT)(SPT)(SK)src[i];  // sign-extending copy-cast
        dst[i] <<= 32;
      }
    }
  for (size_t i = 0; i < n; i++) {
    uint32_t idx = desc ? (uint32_t)(-1 - (int64_t)i) : (uint32_t)i;
    S<PT>(pad, 0, 0)[i] |= idx;
    S<PT>(pad, 2, 1)[i] |= idx;
  }
  if (desc) reverse_vec(pad);

  // radix_sort_body (radixsort.h:116-193)
  reserve(c, n, (size_t)bits + 2);  // 1 permutation + (bits+1) pairs, pairs are singles in 3PC
  oc_vec *vprime = new_vec(n, 4), *f1 = new_vec(n, 4), *s = new_vec(2 * n, 4);
  oc_vec *s1 = new_vec(n, 4), *perm = new_vec(n, 4);
  for (int i = 0; i < bits; i++) {
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        const PT *src = S<PT>(pad, p, k);
        uint32_t *dst = S<uint32_t>(vprime, p, k);
        for (size_t j = 0; j < n; j++) {
          PT sh = (PT)((SPT)src[j] >> (32 + i));  // arithmetic shift
          if (full_width && i == bits - 1) sh = ~sh;  // inplace_invert (not_b)
          dst[j] = (uint32_t)sh;                     // copy-cast to int32
        }
      }
    b2a_bit<uint32_t>(c, vprime, f1);
    // s0 = one - f1 ; s1 = f1 ; s[0] -= 1 ; prefix_sum over 2n
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        uint32_t *sp = S<uint32_t>(s, p, k);
        const uint32_t *f = S<uint32_t>(f1, p, k);
        uint32_t one = ((p == 0 && k == 0) || (p == 2 && k == 1)) ? 1u : 0u;
        for (size_t j = 0; j < n; j++) {
          sp[j] = one - f[j];
          sp[n + j] = f[j];
        }
        if (n > 0) sp[0] -= one;
        uint32_t acc = 0;
        for (size_t j = 0; j < 2 * n; j++) {
          acc += sp[j];
          sp[j] = acc;
        }
      }
    // s1 -= s0 ; f1 *= s1 ; s0 += f1
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        uint32_t *sp = S<uint32_t>(s, p, k), *d = S<uint32_t>(s1, p, k);
        for (size_t j = 0; j < n; j++) d[j] = sp[n + j] - sp[j];
      }
    mult<uint32_t, true>(c, f1, s1, f1, "mul_a");
    PARTIES_PARALLEL
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 2; k++) {
        uint32_t *sp = S<uint32_t>(s, p, k), *d = S<uint32_t>(perm, p, k);
        const uint32_t *f = S<uint32_t>(f1, p, k);
        for (size_t j = 0; j < n; j++) d[j] = sp[j] + f[j];
      }
    apply_elementwise<PT, false>(c, pad, perm, true);
  }
  delete vprime;
  delete f1;
  delete s;
  delete s1;
  delete perm;

  if (desc) reverse_vec(pad);
  // remove_padding: low 32 bits -> B-shared perm; (padded >> 32) -> v
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const PT *src = S<PT>(pad, p, k);
      uint32_t *pd = S<uint32_t>(perm_out, p, k);
      K *vd = S<K>(v, p, k);
      for (size_t j = 0; j < n; j++) {
        pd[j] = (uint32_t)src[j];
        vd[j] = (K)(PT)((SPT)src[j] >> 32);
      }
    }
  perm_b2a(c, perm_out);
  if (desc) perm_negate(c, perm_out);
  perm_invert(c, perm_out);
  delete pad;
}

// ---------------------------------------------------------- quicksort ---
// quicksort (include/core/operators/quicksort.h:212-244): pad_input
// (sorting.h:431-462) to PadWidth -- int64 for 32-bit keys, __int128 for
// 64-bit keys --, quicksort_body (quicksort.h:62-198), remove_padding
// (sorting.h:474-495) and invert.  The comparison bits are opened each round
// and the partition is public, as in the reference.

template <typename T>
oc_vec *gather_rows(const oc_vec *v, const std::vector<size_t> &idx);

// swap rows a and b in every party's share copies
template <typename P>
void swap_rows(oc_vec *v, size_t a, size_t b) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) std::swap(S<P>(v, p, k)[a], S<P>(v, p, k)[b]);
}

// partition (quicksort.h:28-50): Lomuto over [start, end] with the element
// at `start` as pivot; cmp[j] == 0 (not above the pivot) moves j left.  The
// pivot lands at i, which is marked, and i + 1 heads the right part.
template <typename P>
void qs_partition(oc_vec *v, const std::vector<uint8_t> &cmp, long start, long end,
                  std::vector<long> &pivots) {
  long i = start;
  for (long j = start + 1; j <= end; j++)
    if (cmp[j] == 0) {
      i++;
      swap_rows<P>(v, (size_t)j, (size_t)i);
    }
  swap_rows<P>(v, (size_t)start, (size_t)i);
  pivots[i] = i;
  if (i + 1 <= end) pivots[i + 1] = i + 1;
}

// quicksort_body (quicksort.h:62-198) on the padded (unique) values
template <typename P>
void quicksort_body(oc_ctx *c, oc_vec *v) {
  const size_t N = v->n;
  {  // v.shuffle(): one sharded permutation (shuffle.h:752-757)
    oc_ctx::Sharded sp = next_perm(c, N);
    apply_sharded<P, false>(c, v, sp, false);
  }
  std::vector<long> pivots(N, -1);
  pivots[0] = 0;
  std::vector<uint8_t> exp_cmp(N, 0);
  for (;;) {
    // non-pivots, and the pivot of each: the last marked position before it
    // (pivot_temp = prefix max of pivots)
    std::vector<size_t> red_idx, piv_idx;
    long run = -1;
    for (size_t j = 0; j < N; j++) {
      run = std::max(run, pivots[j]);
      if (pivots[j] < 0) {
        red_idx.push_back(j);
        piv_idx.push_back((size_t)run);
      }
    }
    const size_t size = red_idx.size();
    oc_vec *red = gather_rows<P>(v, red_idx), *piv = gather_rows<P>(v, piv_idx);
    oc_vec *eq = new_vec(size, sizeof(P)), *gt = new_vec(size, sizeof(P));
    compare<P>(c, red, piv, eq, gt);  // reduced_vec._compare(pivot_vec, temp, comparisons)
    std::vector<P> cmp = open_vec<P, false>(c, gt);
    for (size_t k = 0; k < size; k++) exp_cmp[red_idx[k]] = (uint8_t)(cmp[k] & 1);
    delete red;
    delete piv;
    delete eq;
    delete gt;
    long index = -1;
    for (long p = 0; p < (long)N; p++) {
      if (pivots[p] < 0) continue;
      if (p - index == 1) {  // adjacent pivots
        index++;
        continue;
      }
      qs_partition<P>(v, exp_cmp, index, p - 1, pivots);
      index = p;
    }
    qs_partition<P>(v, exp_cmp, index, (long)N - 1, pivots);
    size_t num_pivots = 0;
    for (long x : pivots) num_pivots += x >= 0;
    if (num_pivots >= N) return;
  }
}

// K: the key's unsigned type (32 or 64 bits); P: its PadWidth, twice as wide.
// Padded share = sign-extending copy-cast of the key share, << 32; the public
// index (i, or -1 - i for DESC) sits in the low 32 bits of share 0.
template <typename K, typename P>
void quicksort_any(oc_ctx *c, oc_vec *v, bool desc, oc_vec *perm_out) {
  typedef typename std::make_signed<K>::type SK;
  typedef typename std::make_signed<P>::type SP;
  const size_t n = v->n;
  reserve(c, n, 3);  // 1 for the shuffle, 1 for remove_padding's b2a, 1 pair for invert
  if (n == 0) return;
  oc_vec *pad = new_vec(n, sizeof(P));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const K *src = S<K>(v, p, k);
      P *dst = S<P>(pad, p, k);
      for (size_t i = 0; i < n; i++) dst[i] = (P)((SP)(SK)src[i]) << 32;
    }
  for (size_t i = 0; i < n; i++) {
    uint32_t idx = desc ? (uint32_t)(-1 - (int64_t)i) : (uint32_t)i;
    S<P>(pad, 0, 0)[i] |= idx;
    S<P>(pad, 2, 1)[i] |= idx;
  }
  if (desc) reverse_vec(pad);
  quicksort_body<P>(c, pad);
  if (desc) reverse_vec(pad);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const P *src = S<P>(pad, p, k);
      uint32_t *pd = S<uint32_t>(perm_out, p, k);
      K *vd = S<K>(v, p, k);
      for (size_t j = 0; j < n; j++) {
        pd[j] = (uint32_t)src[j];
        vd[j] = (K)(P)((SP)src[j] >> 32);
      }
    }
  delete pad;
  perm_b2a(c, perm_out);
  if (desc) perm_negate(c, perm_out);
  perm_invert(c, perm_out);
}

// ElementwisePermutation::reverse (permutation.h:177-198): p * (-1) + (n - 1)
void perm_reverse(oc_ctx *c, oc_vec *perm) {
  const size_t n = perm->n;
  oc_vec *m1 = new_vec(n, 4), *ad = new_vec(n, 4);
  std::vector<uint32_t> neg(n, 0xffffffffu), nm1(n, (uint32_t)(n - 1));
  public_share<uint32_t>(m1, neg.data());
  public_share<uint32_t>(ad, nm1.data());
  mult<uint32_t, true>(c, perm, m1, perm, "mul_a");
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      uint32_t *a = S<uint32_t>(perm, p, k), *b = S<uint32_t>(ad, p, k);
      for (size_t i = 0; i < n; i++) a[i] += b[i];
    }
  delete m1;
  delete ad;
}

// compose_permutations (shuffle.h:716-743): rho <- rho o sigma
void compose(oc_ctx *c, oc_vec *sigma, oc_vec *rho, bool arith) {
  oc_ctx::Sharded pi = next_perm(c, sigma->n);
  std::vector<int32_t> opened;
  if (arith) {
    apply_sharded<uint32_t, true>(c, sigma, pi, false);
    auto o = open_vec<uint32_t, true>(c, sigma);
    opened.assign(o.begin(), o.end());
  } else {
    apply_sharded<uint32_t, false>(c, sigma, pi, false);
    auto o = open_vec<uint32_t, false>(c, sigma);
    opened.assign(o.begin(), o.end());
  }
  for (size_t i = 0; i < rho->n; i++)
    if (opened[i] < 0 || (size_t)opened[i] >= rho->n)
      throw std::runtime_error("compose: opened permutation out of range");
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) gather<uint32_t>(S<uint32_t>(rho, p, k), opened.data(), rho->n);
  if (arith)
    apply_sharded<uint32_t, true>(c, rho, pi, true);
  else
    apply_sharded<uint32_t, false>(c, rho, pi, true);
}

void apply_ew_any(oc_ctx *c, oc_vec *x, bool x_arith, oc_vec *perm, bool perm_arith) {
  if (x->width == 4) {
    if (x_arith) apply_elementwise<uint32_t, true>(c, x, perm, perm_arith);
    else apply_elementwise<uint32_t, false>(c, x, perm, perm_arith);
  } else {
    if (x_arith) apply_elementwise<uint64_t, true>(c, x, perm, perm_arith);
    else apply_elementwise<uint64_t, false>(c, x, perm, perm_arith);
  }
}

// operators::shuffle of a table (shuffle.h:768-786)
void table_shuffle(oc_ctx *c, oc_vec **a_cols, size_t na, oc_vec **b_cols, size_t nb) {
  oc_vec *first = na ? a_cols[0] : nb ? b_cols[0] : nullptr;
  if (!first) return;
  const size_t n = first->n;
  oc_vec *perm = new_vec(n, 4);
  std::vector<uint32_t> id(n);
  std::iota(id.begin(), id.end(), 0u);
  public_share<uint32_t>(perm, id.data());
  oc_ctx::Sharded pi = next_perm(c, n);
  apply_sharded<uint32_t, false>(c, perm, pi, false);
  for (size_t i = 0; i < na; i++) apply_ew_any(c, a_cols[i], true, perm, false);
  for (size_t i = 0; i < nb; i++) apply_ew_any(c, b_cols[i], false, perm, false);
  delete perm;
}

void radix_any(oc_ctx *c, oc_vec *v, int bits, bool desc, oc_vec *perm_out) {
  if (v->width == 4)
    radix_sort<uint32_t, uint64_t, int64_t>(c, v, bits, desc, perm_out);
  else
    radix_sort<uint64_t, u128, i128>(c, v, bits, desc, perm_out);
}

// operators::table_sort (sorting.h:509-612) with protocol 0 RADIXSORT or
// 1 QUICKSORT (32-bit keys); single-bit columns use 1-bit radix sort either way
void table_sort(oc_ctx *c, oc_vec **keys, size_t nkeys, const int *orders, const int *single_bit,
                oc_vec **a_cols, size_t na, oc_vec **b_cols, size_t nb, int protocol) {
  const size_t n = keys[0]->n;
  int ns = 0;
  for (size_t j = 0; j < nkeys; j++) ns += single_bit && single_bit[j] ? 1 : 0;
  const int nk = (int)nkeys - ns, nc = (int)(na + nb), L = 8 * keys[0]->width;
  long perms_required = nk + ns - 1 + nk + ns;
  long pairs_required = 4L * ns + 3L * nk + nc - 1;
  if (protocol == 1)
    perms_required += nk;
  else
    pairs_required += (long)nk * L;
  if (n == 0) return;
  reserve(c, n, (size_t)(perms_required + pairs_required));
  auto sort_sub = [&](int col, oc_vec *perm_out) {
    const bool desc = orders && orders[col] == 1;
    if (single_bit && single_bit[col]) {
      radix_any(c, keys[col], 1, desc, perm_out);
    } else if (protocol == 1) {
      if (keys[col]->width == 4)
        quicksort_any<uint32_t, uint64_t>(c, keys[col], desc, perm_out);
      else
        quicksort_any<uint64_t, u128>(c, keys[col], desc, perm_out);
    } else {
      radix_any(c, keys[col], 8 * keys[col]->width, desc, perm_out);
    }
  };
  int C = (int)nkeys - 1;
  oc_vec *sort_perm = new_vec(n, 4);
  oc_vec *orig = clone_vec(keys[C]);
  sort_sub(C, sort_perm);
  *keys[C] = *orig;
  delete orig;
  for (C--; C >= 0; C--) {
    orig = clone_vec(keys[C]);
    apply_ew_any(c, keys[C], false, sort_perm, true);
    oc_vec *next = new_vec(n, 4);
    sort_sub(C, next);
    *keys[C] = *orig;
    delete orig;
    compose(c, sort_perm, next, true);
    delete sort_perm;
    sort_perm = next;
  }
  for (size_t j = 0; j < nkeys; j++) apply_ew_any(c, keys[j], false, sort_perm, true);
  for (size_t i = 0; i < na; i++) apply_ew_any(c, a_cols[i], true, sort_perm, true);
  for (size_t i = 0; i < nb; i++) apply_ew_any(c, b_cols[i], false, sort_perm, true);
  delete sort_perm;
}

// ---- bitonic sort (operators/sorting.h:80-353, operators/common.h:20-48) ----
template <typename T>
oc_vec *gather_rows(const oc_vec *v, const std::vector<size_t> &idx) {
  oc_vec *o = new_vec(idx.size(), sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(const_cast<oc_vec *>(v), p, k);
      T *d = S<T>(o, p, k);
      for (size_t i = 0; i < idx.size(); i++) d[i] = a[idx[i]];
    }
  return o;
}
template <typename T>
void scatter_rows(oc_vec *v, const oc_vec *src, const std::vector<size_t> &idx) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(const_cast<oc_vec *>(src), p, k);
      T *d = S<T>(v, p, k);
      for (size_t i = 0; i < idx.size(); i++) d[idx[i]] = a[i];
    }
}
// z = x op y share-wise (local)
template <typename T>
oc_vec *local_op(const oc_vec *x, const oc_vec *y, int op) {
  oc_vec *z = new_vec(x->n, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(const_cast<oc_vec *>(x), p, k), *b = S<T>(const_cast<oc_vec *>(y), p, k);
      T *d = S<T>(z, p, k);
      for (size_t i = 0; i < x->n; i++)
        d[i] = op == 0 ? (T)(a[i] ^ b[i]) : op == 1 ? (T)(a[i] + b[i]) : (T)(a[i] - b[i]);
    }
  return z;
}
// multiplex (common.h:20-48): boolean a ^ (extend_lsb(sel) & (b ^ a)),
// arithmetic a + sel * (b - a)
template <typename T>
oc_vec *multiplex_b(oc_ctx *c, const oc_vec *sel, const oc_vec *a, const oc_vec *b) {
  oc_vec *se = new_vec(sel->n, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *s = S<T>(const_cast<oc_vec *>(sel), p, k);
      T *d = S<T>(se, p, k);
      for (size_t i = 0; i < sel->n; i++) d[i] = (T)(0 - (s[i] & 1));  // extend_lsb
    }
  oc_vec *d = local_op<T>(b, a, 0);
  mult<T, false>(c, se, d, d, "and_b");
  oc_vec *r = local_op<T>(a, d, 0);
  delete se;
  delete d;
  return r;
}
template <typename T>
oc_vec *multiplex_a(oc_ctx *c, const oc_vec *sel, const oc_vec *a, const oc_vec *b) {
  oc_vec *d = local_op<T>(b, a, 2);
  mult<T, true>(c, const_cast<oc_vec *>(sel), d, d, "mul_a");
  oc_vec *r = local_op<T>(a, d, 1);
  delete d;
  return r;
}

template <typename T>
void bitonic_sort(oc_ctx *c, oc_vec **keys, size_t nk, const int *orders, oc_vec **a_cols,
                  size_t na, oc_vec **b_cols, size_t nb) {
  const size_t n = keys[0]->n;
  if (n < 2) return;
  if (n & (n - 1)) throw std::runtime_error("bitonic_sort: size must be a power of two");
  int rounds = 0;
  while (((size_t)1 << rounds) < n) rounds++;
  const size_t m = n / 2;
  std::vector<size_t> xi(m), yi(m);
  for (int i = 0; i < rounds; i++)
    for (int j = 0; j <= i; j++) {
      const size_t h = (size_t)1 << (i - j);
      const bool mirror = j == 0;  // box_direction_2 == -1
      for (size_t q = 0; q < m; q++) {
        size_t box = q / h, t = q % h;
        xi[q] = box * 2 * h + t;
        yi[q] = mirror ? box * 2 * h + 2 * h - 1 - t : box * 2 * h + h + t;
      }
      // compare_rows (sorting.h:80-108)
      std::vector<oc_vec *> X(nk), Y(nk);
      for (size_t k = 0; k < nk; k++) {
        X[k] = gather_rows<T>(keys[k], xi);
        Y[k] = gather_rows<T>(keys[k], yi);
      }
      oc_vec *eq = new_vec(m, sizeof(T)), *gt = new_vec(m, sizeof(T));
      {
        bool inv = orders && orders[0] == 1;
        compare<T>(c, inv ? Y[0] : X[0], inv ? X[0] : Y[0], eq, gt);
      }
      for (size_t k = 1; k < nk; k++) {
        bool inv = orders && orders[k] == 1;
        oc_vec *neq = new_vec(m, sizeof(T)), *ngt = new_vec(m, sizeof(T));
        compare<T>(c, inv ? Y[k] : X[k], inv ? X[k] : Y[k], neq, ngt);
        oc_vec *t = new_vec(m, sizeof(T));
        mult<T, false>(c, ngt, eq, t, "and_b");  // new_gt & eq
        oc_vec *g2 = local_op<T>(gt, t, 0);
        delete gt;
        gt = g2;
        mult<T, false>(c, eq, neq, eq, "and_b");  // eq & new_eq
        delete neq;
        delete ngt;
        delete t;
      }
      for (int p = 0; p < 3; p++)
        for (int k = 0; k < 2; k++) {
          T *g = S<T>(gt, p, k);
          for (size_t q = 0; q < m; q++) g[q] &= 1;  // mask(1)
        }
      // swap (sorting.h:146-162): tmp = mux(bits, x, y); y = mux(bits, y, x); x = tmp
      auto swap_b = [&](oc_vec *col) {
        oc_vec *x = gather_rows<T>(col, xi), *y = gather_rows<T>(col, yi);
        oc_vec *tmp = multiplex_b<T>(c, gt, x, y);
        oc_vec *ny = multiplex_b<T>(c, gt, y, x);
        scatter_rows<T>(col, tmp, xi);
        scatter_rows<T>(col, ny, yi);
        delete x;
        delete y;
        delete tmp;
        delete ny;
      };
      for (size_t k = 0; k < nk; k++) swap_b(keys[k]);
      for (size_t k = 0; k < nb; k++) swap_b(b_cols[k]);
      if (na > 0) {
        oc_vec *ga = new_vec(m, sizeof(T));
        b2a_bit<T>(c, gt, ga);
        for (size_t k = 0; k < na; k++) {
          oc_vec *x = gather_rows<T>(a_cols[k], xi), *y = gather_rows<T>(a_cols[k], yi);
          oc_vec *tmp = multiplex_a<T>(c, ga, x, y);
          oc_vec *ny = multiplex_a<T>(c, ga, y, x);
          scatter_rows<T>(a_cols[k], tmp, xi);
          scatter_rows<T>(a_cols[k], ny, yi);
          delete x;
          delete y;
          delete tmp;
          delete ny;
        }
        delete ga;
      }
      for (size_t k = 0; k < nk; k++) {
        delete X[k];
        delete Y[k];
      }
      delete eq;
      delete gt;
    }
}

// ---- aggregation network (operators/aggregation.h:27-346) and distinct ----
enum AggFn { kAggSum = 0, kAggCount = 1, kAggCopy = 2, kAggMax = 3, kAggMin = 4, kAggOr = 5,
             kAggValid = 6 };
struct AggSpec {
  oc_vec *in, *out;
  int fn;
};

template <typename T>
oc_vec *slice_of(const oc_vec *v, size_t start, size_t len) {
  std::vector<size_t> idx(len);
  for (size_t i = 0; i < len; i++) idx[i] = start + i;
  return gather_rows<T>(v, idx);
}
template <typename T>
void write_slice(oc_vec *v, size_t start, const oc_vec *src) {
  std::vector<size_t> idx(src->n);
  for (size_t i = 0; i < src->n; i++) idx[i] = start + i;
  scatter_rows<T>(v, src, idx);
}
template <typename T>
oc_vec *not_copy(const oc_vec *x) {
  oc_vec *r = clone_vec(x);
  not_vec<T>(r);
  return r;
}
// x | y = ~(~x & ~y) (b_shared_vector.h:438-441)
template <typename T>
oc_vec *or_b(oc_ctx *c, const oc_vec *x, const oc_vec *y) {
  oc_vec *nx = not_copy<T>(x), *ny = not_copy<T>(y);
  mult<T, false>(c, nx, ny, nx, "and_b");
  not_vec<T>(nx);
  delete ny;
  return nx;
}
template <typename T>
void mask1(oc_vec *v) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a = S<T>(v, p, k);
      for (size_t i = 0; i < v->n; i++) a[i] &= 1;
    }
}

// the boolean aggregation functions of aggregation.h:27-140 on (g, a, b);
// returns the new accumulator a
template <typename T>
oc_vec *agg_b(oc_ctx *c, int fn, const oc_vec *g, const oc_vec *a, const oc_vec *b) {
  switch (fn) {
    case kAggCopy:
      return multiplex_b<T>(c, g, a, b);
    case kAggOr: {
      oc_vec *ext = new_vec(g->n, sizeof(T));
      for (int p = 0; p < 3; p++)
        for (int k = 0; k < 2; k++) {
          const T *s = S<T>(const_cast<oc_vec *>(g), p, k);
          T *d = S<T>(ext, p, k);
          for (size_t i = 0; i < g->n; i++) d[i] = (T)(0 - (s[i] & 1));
        }
      mult<T, false>(c, ext, const_cast<oc_vec *>(b), ext, "and_b");
      oc_vec *r = or_b<T>(c, a, ext);
      delete ext;
      return r;
    }
    case kAggMax:
    case kAggMin: {
      oc_vec *eq = new_vec(a->n, sizeof(T)), *bg = new_vec(a->n, sizeof(T));
      compare<T>(c, const_cast<oc_vec *>(b), const_cast<oc_vec *>(a), eq, bg);  // b > a
      if (fn == kAggMin)
        for (int p = 0; p < 3; p++)
          for (int k = 0; k < 2; k++) {
            T *d = S<T>(bg, p, k);
            for (size_t i = 0; i < bg->n; i++) d[i] = (T)((d[i] & 1) == 0);  // not_b_1
          }
      oc_vec *inner = multiplex_b<T>(c, bg, a, b);
      oc_vec *r = multiplex_b<T>(c, g, a, inner);
      delete eq;
      delete bg;
      delete inner;
      return r;
    }
    case kAggValid: {
      oc_vec *agg = new_vec(a->n, sizeof(T));
      mult<T, false>(c, const_cast<oc_vec *>(a), const_cast<oc_vec *>(b), agg, "and_b");
      oc_vec *r = multiplex_b<T>(c, g, a, agg);
      delete agg;
      return r;
    }
    default:
      throw std::runtime_error("aggregate: not a boolean aggregation");
  }
}
template <typename T>
oc_vec *agg_a(oc_ctx *c, int fn, const oc_vec *g, const oc_vec *a, const oc_vec *b) {
  switch (fn) {
    case kAggSum:
    case kAggCount: {  // a + g * b
      oc_vec *t = new_vec(a->n, sizeof(T));
      mult<T, true>(c, const_cast<oc_vec *>(g), const_cast<oc_vec *>(b), t, "mul_a");
      oc_vec *r = local_op<T>(a, t, 1);
      delete t;
      return r;
    }
    case kAggCopy:
      return multiplex_a<T>(c, g, a, b);
    default:
      throw std::runtime_error("aggregate: not an arithmetic aggregation");
  }
}

template <typename T>
void aggregate(oc_ctx *c, oc_vec **keys, size_t nk, AggSpec *bs, size_t nb, AggSpec *as,
               size_t na, bool reverse, oc_vec *sel) {
  size_t n = nk ? keys[0]->n : nb ? bs[0].in->n : na ? as[0].in->n : 0;
  if (!nk && !nb && !na) throw std::runtime_error("Empty aggregation!");
  if (n & (n - 1)) throw std::runtime_error("aggregate: size must be a power of two");
  bool a_copy = false, a_noncopy = false, b_copy = false, b_noncopy = false;
  std::vector<T> onesv(n, 1);
  for (size_t i = 0; i < na; i++) {
    if (as[i].fn == kAggCount) {
      if (sel) {
        b2a_bit<T>(c, sel, as[i].out);
      } else {
        public_share<T>(as[i].out, onesv.data());
      }
    } else if (as[i].out != as[i].in) {
      *as[i].out = *as[i].in;
    }
    (as[i].fn == kAggCopy ? a_copy : a_noncopy) = true;
  }
  for (size_t i = 0; i < nb; i++) {
    if (bs[i].out != bs[i].in) *bs[i].out = *bs[i].in;
    (bs[i].fn == kAggCopy ? b_copy : b_noncopy) = true;
  }
  int log_size = 0;
  while (((size_t)1 << (log_size + 1)) <= n) log_size++;
  for (int lv = 1; lv <= log_size; lv++) {
    size_t d = n / ((size_t)1 << lv);
    if (reverse) d = n / ((size_t)1 << (log_size - lv + 1));
    const size_t m = n - d;
    oc_vec *g;
    if (nk == 0) {
      g = new_vec(m, sizeof(T));
      std::vector<T> o(m, 1);
      public_share<T>(g, o.data());
    } else {
      oc_vec *x = slice_of<T>(keys[0], 0, m), *y = slice_of<T>(keys[0], d, m);
      g = bit_same<T>(c, x, y);
      mask1<T>(g);  // operator==
      delete x;
      delete y;
      for (size_t k = 1; k < nk; k++) {
        x = slice_of<T>(keys[k], 0, m);
        y = slice_of<T>(keys[k], d, m);
        oc_vec *e = bit_same<T>(c, x, y);
        mask1<T>(e);
        mult<T, false>(c, g, e, g, "and_b");  // group_bits_b &= ...
        delete e;
        delete x;
        delete y;
      }
    }
    oc_vec *jg = clone_vec(g);
    if (sel && (a_noncopy || b_noncopy)) {
      oc_vec *s1 = slice_of<T>(sel, d, m), *s0 = slice_of<T>(sel, 0, m);
      oc_vec *sx = local_op<T>(s1, s0, 0);
      not_vec<T>(sx);
      mult<T, false>(c, g, sx, g, "and_b");
      delete s1;
      delete s0;
      delete sx;
    }
    mask1<T>(g);
    for (size_t i = 0; i < nb; i++) {
      oc_vec *a = slice_of<T>(bs[i].out, 0, m), *b = slice_of<T>(bs[i].out, d, m);
      const oc_vec *gg = bs[i].fn == kAggCopy ? jg : g;
      if (reverse) {
        oc_vec *r = agg_b<T>(c, bs[i].fn, gg, b, a);
        write_slice<T>(bs[i].out, d, r);
        delete r;
      } else {
        oc_vec *r = agg_b<T>(c, bs[i].fn, gg, a, b);
        write_slice<T>(bs[i].out, 0, r);
        delete r;
      }
      delete a;
      delete b;
    }
    oc_vec *ga = nullptr, *jga = nullptr;
    if (nk > 0 && na > 0) {
      if (a_noncopy) {
        ga = new_vec(m, sizeof(T));
        b2a_bit<T>(c, g, ga);
      }
      if (a_copy) {
        jga = new_vec(m, sizeof(T));
        b2a_bit<T>(c, jg, jga);
      }
    } else if (na > 0) {  // no keys: public ones
      std::vector<T> o(m, 1);
      ga = new_vec(m, sizeof(T));
      public_share<T>(ga, o.data());
      jga = clone_vec(ga);
    }
    for (size_t i = 0; i < na; i++) {
      oc_vec *a = slice_of<T>(as[i].out, 0, m), *b = slice_of<T>(as[i].out, d, m);
      const oc_vec *gg = as[i].fn == kAggCopy ? jga : ga;
      if (reverse) {
        oc_vec *r = agg_a<T>(c, as[i].fn, gg, b, a);
        write_slice<T>(as[i].out, d, r);
        delete r;
      } else {
        oc_vec *r = agg_a<T>(c, as[i].fn, gg, a, b);
        write_slice<T>(as[i].out, 0, r);
        delete r;
      }
      delete a;
      delete b;
    }
    delete g;
    delete jg;
    delete ga;
    delete jga;
  }
}

// operators::distinct (distinct.h:19-44): res[0] = public 1; for each key,
// res[1..n) |= key[0..n-1) != key[1..n)
template <typename T>
void distinct(oc_ctx *c, oc_vec **keys, size_t nk, oc_vec *res) {
  const size_t n = res->n;
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) std::fill(res->sh[p][k].begin(), res->sh[p][k].end(), 0);
  if (n == 0) return;
  oc_vec *one = new_vec(1, sizeof(T));
  T o1 = 1;
  public_share<T>(one, &o1);
  write_slice<T>(res, 0, one);
  delete one;
  if (n == 1) return;
  for (size_t k = 0; k < nk; k++) {
    oc_vec *a = slice_of<T>(keys[k], 0, n - 1), *b = slice_of<T>(keys[k], 1, n - 1);
    oc_vec *e = bit_same<T>(c, a, b);
    for (int p = 0; p < 3; p++)  // (a == b) -> mask(1) -> not_b_1
      for (int q = 0; q < 2; q++) {
        T *d = S<T>(e, p, q);
        for (size_t i = 0; i < e->n; i++) d[i] = (T)(((d[i] & 1) & 1) == 0);
      }
    oc_vec *rest = slice_of<T>(res, 1, n - 1);
    // rest |= ne: invert rest, rest &= ~ne, invert rest (b_shared_vector.h:452-456)
    not_vec<T>(rest);
    oc_vec *nne = not_copy<T>(e);
    mult<T, false>(c, rest, nne, rest, "and_b");
    not_vec<T>(rest);
    write_slice<T>(res, 1, rest);
    delete a;
    delete b;
    delete e;
    delete rest;
    delete nne;
  }
}

// ---- adders (operators/circuits.h:35-216) and a2b (a_shared_vector.h:138-144) ----
// pack_from (mapping_access_vector.h:974-991): word w, bit q = bit `pos` of
// element w*B + q
template <typename T>
oc_vec *pack_bits(const oc_vec *v, int pos) {
  const int B = 8 * sizeof(T);
  const size_t n = v->n, nc = n / B + (n % B > 0);
  oc_vec *o = new_vec(nc, sizeof(T));
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = S<T>(const_cast<oc_vec *>(v), p, k);
      T *d = S<T>(o, p, k);
      for (size_t e = 0; e < n; e++) d[e / B] |= (T)((a[e] >> pos) & 1) << (e % B);
    }
  return o;
}
// unpack_from (mapping_access_vector.h:993-1013): bit `pos` of element e =
// bit e%B of word e/B; other bits kept
template <typename T>
void unpack_bits(oc_vec *v, const oc_vec *w, int pos) {
  const int B = 8 * sizeof(T);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a = S<T>(v, p, k);
      const T *d = S<T>(const_cast<oc_vec *>(w), p, k);
      for (size_t e = 0; e < v->n; e++) {
        T bit = (d[e / B] >> (e % B)) & 1;
        a[e] = (T)((a[e] & ~((T)1 << pos)) | (bit << pos));
      }
    }
}
template <typename T>
void xor_into(oc_vec *x, const oc_vec *y) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *a = S<T>(x, p, k);
      const T *b = S<T>(const_cast<oc_vec *>(y), p, k);
      for (size_t i = 0; i < x->n; i++) a[i] ^= b[i];
    }
}

// ripple_carry_adder (circuits.h:35-77): bit-serial on bit-packed words,
// carry c_i = ((c_{i-1} ^ a_i) & axb_i) ^ a_i, B - 1 and_b calls on n/B words
template <typename T>
void rca(oc_ctx *c, const oc_vec *a, const oc_vec *b, bool carry_in, oc_vec *sum) {
  const int B = 8 * sizeof(T);
  const size_t n = a->n, nc = n / B + (n % B > 0);
  oc_vec *axb = local_op<T>(a, b, 0);
  oc_vec *carry = new_vec(nc, sizeof(T));
  if (carry_in) not_vec<T>(carry);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) std::fill(sum->sh[p][k].begin(), sum->sh[p][k].end(), 0);
  for (int i = 0; i < B; i++) {
    oc_vec *axb_i = pack_bits<T>(axb, i), *a_i = pack_bits<T>(a, i);
    oc_vec *t = local_op<T>(axb_i, carry, 0);
    unpack_bits<T>(sum, t, i);
    delete t;
    if (i < B - 1) {
      xor_into<T>(carry, a_i);
      mult<T, false>(c, carry, axb_i, carry, "and_b");
      xor_into<T>(carry, a_i);
    }
    delete axb_i;
    delete a_i;
  }
  delete axb;
  delete carry;
}

// rca_compare (circuits.h:93-136): a < b as the sign of a + ~b + 1
template <typename T>
void rca_compare(oc_ctx *c, const oc_vec *a, const oc_vec *b, oc_vec *out) {
  const int B = 8 * sizeof(T);
  const size_t n = a->n, nc = n / B + (n % B > 0);
  oc_vec *nb = not_copy<T>(b);
  oc_vec *axb = local_op<T>(a, nb, 0);
  oc_vec *carry = new_vec(nc, sizeof(T));
  not_vec<T>(carry);  // + 1
  oc_vec *axb_i = nullptr;
  for (int i = 0; i < B; i++) {
    delete axb_i;
    axb_i = pack_bits<T>(axb, i);
    if (i == B - 1) break;
    oc_vec *a_i = pack_bits<T>(a, i);
    xor_into<T>(carry, a_i);
    mult<T, false>(c, carry, axb_i, carry, "and_b");
    xor_into<T>(carry, a_i);
    delete a_i;
  }
  xor_into<T>(carry, axb_i);
  oc_vec *sum = new_vec(n, sizeof(T));
  unpack_bits<T>(sum, carry, B - 1);
  typedef typename std::make_signed<T>::type ST;
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {  // ltz of each share
      const T *s = S<T>(sum, p, k);
      T *d = S<T>(out, p, k);
      for (size_t i = 0; i < n; i++) d[i] = (T)((ST)s[i] < 0);
    }
  delete nb;
  delete axb;
  delete carry;
  delete axb_i;
  delete sum;
}

// parallel_prefix_adder (circuits.h:138-216), Kogge-Stone
template <typename T>
void ppa(oc_ctx *c, const oc_vec *x, const oc_vec *y, bool carry_in, oc_vec *out) {
  const int B = 8 * sizeof(T);
  const int LOG = B == 32 ? 5 : 6;
  const size_t n = x->n;
  oc_vec *p = local_op<T>(x, y, 0);
  oc_vec *g = new_vec(n, sizeof(T));
  mult<T, false>(c, const_cast<oc_vec *>(x), const_cast<oc_vec *>(y), g, "and_b");
  if (carry_in) {
    oc_vec *prop = clone_vec(p);
    mask1<T>(prop);
    for (int k = 0; k < 2; k++)  // p ^ public 1: s_0 only
      for (int q = 0; q < 3; q++)
        if ((q == 0 && k == 0) || (q == 2 && k == 1)) {
          T *a = S<T>(p, q, k);
          for (size_t i = 0; i < n; i++) a[i] ^= 1;
        }
    xor_into<T>(g, prop);
    delete prop;
  }
  oc_vec *propagate = clone_vec(p);
  for (int i = 0; i < LOG; i++) {
    oc_vec *gs = new_vec(n, sizeof(T)), *ps = new_vec(n, sizeof(T));
    for (int q = 0; q < 3; q++)
      for (int k = 0; k < 2; k++) {
        const T *ga = S<T>(g, q, k), *pa = S<T>(p, q, k);
        T *gd = S<T>(gs, q, k), *pd = S<T>(ps, q, k);
        for (size_t e = 0; e < n; e++) {
          gd[e] = (T)(ga[e] << (1 << i));
          pd[e] = (T)(pa[e] << (1 << i));
        }
      }
    // prefix_sum((g, p), (gs, ps)) (circuits.h:138-164)
    mult<T, false>(c, ps, p, ps, "and_b");  // p_y &= p_x
    mult<T, false>(c, gs, p, gs, "and_b");  // g_y &= p_x
    oc_vec *gpg = new_vec(n, sizeof(T));
    mult<T, false>(c, g, gs, gpg, "and_b");  // g_x & g_y
    oc_vec *t = local_op<T>(g, gs, 0);
    xor_into<T>(t, gpg);
    delete g;
    delete p;
    g = t;
    p = ps;
    delete gs;
    delete gpg;
  }
  for (int q = 0; q < 3; q++)
    for (int k = 0; k < 2; k++) {
      const T *ga = S<T>(g, q, k), *pr = S<T>(propagate, q, k);
      T *d = S<T>(out, q, k);
      for (size_t e = 0; e < n; e++) d[e] = pr[e] ^ (T)(ga[e] << 1);
    }
  delete g;
  delete p;
  delete propagate;
}

// ASharedVector::a2b (a_shared_vector.h:138-144) with redistribute_shares_b
// (replicated_3pc.h:259-262): party 0 shares s0 + s1, party 2 shares s2, and
// the default adder (ripple carry) adds them.
template <typename T>
void a2b(oc_ctx *c, const oc_vec *x, oc_vec *y) {
  const size_t n = x->n;
  std::vector<T> v0(n), v2(n);
  const T *x00 = S<T>(const_cast<oc_vec *>(x), 0, 0), *x01 = S<T>(const_cast<oc_vec *>(x), 0, 1);
  const T *x20 = S<T>(const_cast<oc_vec *>(x), 2, 0);
  for (size_t i = 0; i < n; i++) {
    v0[i] = (T)(x00[i] + x01[i]);
    v2[i] = x20[i];
  }
  oc_vec *f = new_vec(n, sizeof(T)), *sv = new_vec(n, sizeof(T));
  secret_share<T, false>(c, v0.data(), f, 0);
  secret_share<T, false>(c, v2.data(), sv, 2);
  rca<T>(c, f, sv, false, y);
  delete f;
  delete sv;
}

}  // namespace

// =================================================================== C ABI
// Local (message-free) share operations, each party on its two share copies:
// op 0 xor_b (z = x ^ y, protocol.h:649), 1 add_a (z = x + y, protocol.h:561),
// 2 sub_a (z = x - y, protocol.h:570), 3 not_b (z = ~x, replicated_3pc.h:199),
// 4 neg_a (z = -x, protocol.h:587).  y is not read by ops 3 and 4.
template <typename T>
static void local_op(int op, const oc_vec *x, const oc_vec *y, oc_vec *z) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = (const T *)x->sh[p][k].data();
      const T *b = (const T *)y->sh[p][k].data();
      T *d = (T *)z->sh[p][k].data();
      for (size_t i = 0; i < x->n; i++) {
        switch (op) {
          case 0: d[i] = a[i] ^ b[i]; break;
          case 1: d[i] = (T)(a[i] + b[i]); break;
          case 2: d[i] = (T)(a[i] - b[i]); break;
          case 3: d[i] = (T)~a[i]; break;
          default: d[i] = (T)(T(0) - a[i]); break;
        }
      }
    }
}

// Share-local Vector maps, each party on both of its shares
// (mapping_access_vector.h): 0 bit_left_shift (:413-421), 1
// bit_arithmetic_right_shift (:376-384), 2 bit_logical_right_shift, 3 bit_xor
// (:430-439, popcount parity), 4 extend_lsb (:1541-1547), 5 ltz (:1528-1535),
// 6 mask (:1312-1316).  Shift amounts at or over the width give 0 (sign fill
// for the arithmetic shift), where the reference's expression is undefined.
template <typename T>
static void unary_op(int op, const oc_vec *x, uint64_t arg, oc_vec *z) {
  typedef typename std::make_signed<T>::type S;
  const int B = 8 * sizeof(T);
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      const T *a = (const T *)x->sh[p][k].data();
      T *d = (T *)z->sh[p][k].data();
      for (size_t i = 0; i < x->n; i++) {
        const T v = a[i];
        T r;
        switch (op) {
          case 0: r = arg >= (uint64_t)B ? 0 : (T)(v << arg); break;
          case 1: r = (T)((S)v >> (arg >= (uint64_t)B ? B - 1 : (int)arg)); break;
          case 2: r = arg >= (uint64_t)B ? 0 : (T)(v >> arg); break;
          case 3: r = (T)(__builtin_popcountll((unsigned long long)v) & 1); break;
          case 4: r = (T)(-(v & 1)); break;
          case 5: r = (T)((S)v < 0); break;
          default: r = v & (T)arg; break;
        }
        d[i] = r;
      }
    }
}

// SharedVector::prefix_sum (shared_vector.h:249): std::inclusive_scan of
// each share (mapping_access_vector.h:446)
template <typename T>
static void prefix_sum_op(oc_vec *v) {
  for (int p = 0; p < 3; p++)
    for (int k = 0; k < 2; k++) {
      T *d = (T *)v->sh[p][k].data();
      std::inclusive_scan(d, d + v->n, d);
    }
}

extern "C" {

// `threads` worker threads, thread t with the ten keys at keys + 320 t
oc_ctx *oc_create_threads(const uint8_t *keys /* threads x 10 x 32 bytes */, int threads,
                          long batch_size, int gen_everyone) {
  if (threads < 1 || threads > kMaxThreads) return nullptr;
  oc_ctx *c = new oc_ctx;
  c->batch_size = batch_size;
  c->gen_everyone = gen_everyone;
  c->threads = threads;
  c->TP.resize(threads);
  for (int t = 0; t < threads; t++) {
    const uint8_t *k = keys + 320 * (size_t)t;
    for (int p = 0; p < 3; p++) {
      Party &pt = c->TP[t][p];
      pt.id = p;
      pt.local = Prg(k + 32 * p);
      pt.next = Prg(k + 32 * (3 + p));
      pt.prev = Prg(k + 32 * (3 + (p + 2) % 3));
      for (int g = 0; g < 4; g++) pt.group[g] = Prg(k + 32 * (6 + g));
    }
  }
  return c;
}

oc_ctx *oc_create(const uint8_t *keys /* 10 x 32 bytes */, long batch_size, int gen_everyone) {
  return oc_create_threads(keys, 1, batch_size, gen_everyone);
}

void oc_destroy(oc_ctx *c) { delete c; }

// nonce of one PRG object of worker thread t: role 0 local, 1 prev, 2 next,
// 3+g group g
uint64_t oc_prg_nonce_thread(oc_ctx *c, int t, int party, int role) {
  Party &pt = c->TP[t][party];
  if (role == 0) return pt.local.nonce;
  if (role == 1) return pt.prev.nonce;
  if (role == 2) return pt.next.nonce;
  return pt.group[role - 3].nonce;
}

// nonce of one PRG object of thread 0
uint64_t oc_prg_nonce(oc_ctx *c, int party, int role) {
  Party &pt = c->TP[0][party];
  if (role == 0) return pt.local.nonce;
  if (role == 1) return pt.prev.nonce;
  if (role == 2) return pt.next.nonce;
  return pt.group[role - 3].nonce;
}

// AESPRGAlgorithm keystream for `bytes` bytes starting at `nonce`;
// returns the nonce after the call.
uint64_t oc_keystream(const uint8_t *key, uint64_t nonce, uint8_t *out, size_t bytes) {
  Prg g(key);
  g.nonce = nonce;
  g.fill(out, bytes);
  return g.nonce;
}

// gen_perm from a PRG at `*nonce` (advanced in place); `draws` gets the number
// of uint32 words consumed before the last acceptance, inclusive.
void oc_gen_perm(const uint8_t *key, uint64_t *nonce, int32_t *out, size_t n, uint64_t *draws) {
  Prg g(key);
  g.nonce = *nonce;
  std::vector<int32_t> perm(n);
  gen_perm(perm, g, draws);
  std::memcpy(out, perm.data(), n * sizeof(int32_t));
  *nonce = g.nonce;
}

oc_vec *oc_vec_new(size_t n, int width) { return new_vec(n, width); }
void oc_vec_free(oc_vec *v) { delete v; }
// keep the first min(n, size) rows, new rows zero
void oc_vec_resize(oc_vec *v, size_t n) {
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < 2; i++) v->sh[p][i].resize(n * v->width, 0);
  v->n = n;
}
// keep the last n rows
void oc_vec_tail(oc_vec *v, size_t n) {
  const size_t w = v->width, drop = v->n - n;
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < 2; i++) v->sh[p][i].erase(v->sh[p][i].begin(), v->sh[p][i].begin() + drop * w);
  v->n = n;
}
void oc_vec_copy_rows(const oc_vec *src, size_t s0, oc_vec *dst, size_t d0, size_t count) {
  const size_t w = src->width;
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < 2; i++)
      std::memcpy(dst->sh[p][i].data() + d0 * w, src->sh[p][i].data() + s0 * w, count * w);
}
void oc_vec_broadcast_row(const oc_vec *src, size_t row, oc_vec *dst, size_t d0, size_t count) {
  const size_t w = src->width;
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < 2; i++)
      for (size_t r = 0; r < count; r++)
        std::memcpy(dst->sh[p][i].data() + (d0 + r) * w, src->sh[p][i].data() + row * w, w);
}
size_t oc_vec_size(oc_vec *v) { return v->n; }
void oc_vec_set(oc_vec *v, int party, int idx, const void *src) {
  std::memcpy(v->sh[party][idx].data(), src, v->n * v->width);
}
void oc_vec_get(oc_vec *v, int party, int idx, void *dst) {
  std::memcpy(dst, v->sh[party][idx].data(), v->n * v->width);
}

#define OC_TRY(...)                                  \
  try {                                              \
    __VA_ARGS__;                                     \
    return 0;                                        \
  } catch (const std::exception &e) {                \
    std::fprintf(stderr, "oracle: %s\n", e.what()); \
    return -1;                                       \
  }

int oc_secret_share(oc_ctx *c, const void *data, oc_vec *v, int arith) {
  OC_TRY({
    if (v->width == 4)
      arith ? secret_share<uint32_t, true>(c, (const uint32_t *)data, v)
            : secret_share<uint32_t, false>(c, (const uint32_t *)data, v);
    else if (v->width == 8)
      arith ? secret_share<uint64_t, true>(c, (const uint64_t *)data, v)
            : secret_share<uint64_t, false>(c, (const uint64_t *)data, v);
    else
      throw std::runtime_error("secret_share: width");
  })
}

int oc_open(oc_ctx *c, oc_vec *v, int arith, void *out) {
  OC_TRY({
    if (v->width == 4) {
      auto o = arith ? open_vec<uint32_t, true>(c, v) : open_vec<uint32_t, false>(c, v);
      std::memcpy(out, o.data(), o.size() * 4);
    } else if (v->width == 8) {
      auto o = arith ? open_vec<uint64_t, true>(c, v) : open_vec<uint64_t, false>(c, v);
      std::memcpy(out, o.data(), o.size() * 8);
    } else
      throw std::runtime_error("open: width");
  })
}

int oc_local(oc_ctx *, int op, oc_vec *x, oc_vec *y, oc_vec *z) {
  OC_TRY({
    if (op < 0 || op > 4 || x->n != z->n || x->width != z->width ||
        (op < 3 && (y->n != x->n || y->width != x->width)))
      throw std::runtime_error("local op: argument");
    if (x->width == 4)
      local_op<uint32_t>(op, x, y, z);
    else if (x->width == 8)
      local_op<uint64_t>(op, x, y, z);
    else
      throw std::runtime_error("local op: width");
  })
}

int oc_unary(oc_ctx *, int op, oc_vec *x, uint64_t arg, oc_vec *z) {
  OC_TRY({
    if (op < 0 || op > 6 || x->n != z->n || x->width != z->width)
      throw std::runtime_error("unary: argument");
    if (x->width == 4)
      unary_op<uint32_t>(op, x, arg, z);
    else if (x->width == 8)
      unary_op<uint64_t>(op, x, arg, z);
    else
      throw std::runtime_error("unary: width");
  })
}

int oc_prefix_sum(oc_ctx *, oc_vec *v) {
  OC_TRY({
    if (v->width == 4)
      prefix_sum_op<uint32_t>(v);
    else if (v->width == 8)
      prefix_sum_op<uint64_t>(v);
    else
      throw std::runtime_error("prefix_sum: width");
  })
}

int oc_and_b(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *z) {
  OC_TRY({
    if (x->width == 4)
      mult<uint32_t, false>(c, x, y, z, "and_b");
    else if (x->width == 8)
      mult<uint64_t, false>(c, x, y, z, "and_b");
    else
      throw std::runtime_error("and_b: width");
  })
}

int oc_mul_a(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *z) {
  OC_TRY({
    if (x->width == 4)
      mult<uint32_t, true>(c, x, y, z, "mul_a");
    else if (x->width == 8)
      mult<uint64_t, true>(c, x, y, z, "mul_a");
    else
      throw std::runtime_error("mul_a: width");
  })
}

int oc_b2a_bit(oc_ctx *c, oc_vec *x, oc_vec *y) {
  OC_TRY({
    if (x->width == 4)
      b2a_bit<uint32_t>(c, x, y);
    else if (x->width == 8)
      b2a_bit<uint64_t>(c, x, y);
    else
      throw std::runtime_error("b2a_bit: width");
  })
}

int oc_reshare(oc_ctx *c, oc_vec *v, int group, int arith) {
  OC_TRY({
    if (v->width == 4)
      arith ? reshare<uint32_t, true>(c, v, group) : reshare<uint32_t, false>(c, v, group);
    else if (v->width == 8)
      arith ? reshare<uint64_t, true>(c, v, group) : reshare<uint64_t, false>(c, v, group);
    else
      throw std::runtime_error("reshare: width");
  })
}

int oc_reserve(oc_ctx *c, size_t n, size_t count) { OC_TRY(reserve(c, n, count)) }
size_t oc_queue_size(oc_ctx *c) { return c->queue.size(); }

// copy one queued sharded permutation's local permutation for a group as
// party `party` sees it (index 0 = front of the queue)
int oc_peek_perm(oc_ctx *c, size_t index, int party, int group, int32_t *out) {
  OC_TRY({
    if (index >= c->queue.size()) throw std::runtime_error("peek: index");
    const auto &v = c->queue[index].perm[party][group];
    if (v.empty()) throw std::runtime_error("peek: party not in group");
    std::memcpy(out, v.data(), v.size() * 4);
  })
}

int oc_apply_sharded(oc_ctx *c, oc_vec *x, int arith, int inverse) {
  OC_TRY({
    oc_ctx::Sharded pi = next_perm(c, x->n);
    if (x->width == 4)
      arith ? apply_sharded<uint32_t, true>(c, x, pi, inverse)
            : apply_sharded<uint32_t, false>(c, x, pi, inverse);
    else if (x->width == 8)
      arith ? apply_sharded<uint64_t, true>(c, x, pi, inverse)
            : apply_sharded<uint64_t, false>(c, x, pi, inverse);
    else
      throw std::runtime_error("apply_sharded: width");
  })
}

int oc_apply_elementwise(oc_ctx *c, oc_vec *x, int x_arith, oc_vec *perm, int perm_arith) {
  OC_TRY({
    if (x->width == 4)
      x_arith ? apply_elementwise<uint32_t, true>(c, x, perm, perm_arith)
              : apply_elementwise<uint32_t, false>(c, x, perm, perm_arith);
    else if (x->width == 8)
      x_arith ? apply_elementwise<uint64_t, true>(c, x, perm, perm_arith)
              : apply_elementwise<uint64_t, false>(c, x, perm, perm_arith);
    else
      throw std::runtime_error("apply_elementwise: width");
  })
}

// radix_sort of a B-shared int32/int64 vector (in place); perm_out (width 4)
// receives the A-shared sorting permutation.
int oc_radix_sort(oc_ctx *c, oc_vec *v, int bits, int desc, oc_vec *perm_out) {
  OC_TRY({
    if (perm_out->n != v->n || perm_out->width != 4) throw std::runtime_error("radix: perm_out");
    if (v->width == 4)
      radix_sort<uint32_t, uint64_t, int64_t>(c, v, bits, desc, perm_out);
    else if (v->width == 8)
      radix_sort<uint64_t, u128, i128>(c, v, bits, desc, perm_out);
    else
      throw std::runtime_error("radix_sort: width");
  })
}

// thread ranges of one runtime call over n elements: out[2t], out[2t+1]
void oc_thread_bounds(size_t n, long batch_size, int threads, uint64_t *out) {
  auto b = thread_bounds(n, batch_size, threads);
  for (int t = 0; t < threads; t++) {
    out[2 * t] = b[t].first;
    out[2 * t + 1] = b[t].second;
  }
}

// batch sizes of one runtime call over n elements (returns the batch count)
size_t oc_batches(size_t n, long batch_size, uint64_t *sizes, size_t cap) {
  auto b = batches_for(n, batch_size);
  for (size_t i = 0; i < b.size() && i < cap; i++) sizes[i] = b[i].e - b[i].s;
  return b.size();
}

// x._compare(y, eq, gt): boolean shares of x == y and x > y (signed)
int oc_compare(oc_ctx *c, oc_vec *x, oc_vec *y, oc_vec *eq, oc_vec *gt) {
  OC_TRY({
    if (x->width == 4)
      compare<uint32_t>(c, x, y, eq, gt);
    else if (x->width == 8)
      compare<uint64_t>(c, x, y, eq, gt);
    else
      throw std::runtime_error("compare: width");
  })
}

int oc_cmp(oc_ctx *c, int op, oc_vec *x, oc_vec *y, oc_vec *z) {
  OC_TRY({
    if (op < 0 || op > 5) throw std::runtime_error("cmp: op");
    if (x->width == 4)
      cmp_op<uint32_t>(c, op, x, y, z);
    else
      cmp_op<uint64_t>(c, op, x, y, z);
  })
}

// full-width b2a of a boolean-shared vector
int oc_b2a(oc_ctx *c, oc_vec *x, oc_vec *y) {
  OC_TRY({
    if (x->width == 4)
      b2a_full<uint32_t>(c, x, y);
    else if (x->width == 8)
      b2a_full<uint64_t>(c, x, y);
    else
      throw std::runtime_error("b2a: width");
  })
}

int oc_perm_invert(oc_ctx *c, oc_vec *perm, int arith) {
  OC_TRY({
    if (!arith) throw std::runtime_error("perm_invert: only A-shared permutations");
    perm_invert(c, perm);
  })
}
int oc_perm_b2a(oc_ctx *c, oc_vec *perm) { OC_TRY(perm_b2a(c, perm)) }
int oc_perm_negate(oc_ctx *c, oc_vec *perm) { OC_TRY(perm_negate(c, perm)) }
int oc_perm_reverse(oc_ctx *c, oc_vec *perm) { OC_TRY(perm_reverse(c, perm)) }
int oc_perm_compose(oc_ctx *c, oc_vec *sigma, oc_vec *rho, int arith) {
  OC_TRY(compose(c, sigma, rho, arith != 0))
}
int oc_shuffle(oc_ctx *c, oc_vec *x, int arith) {
  OC_TRY({
    oc_ctx::Sharded pi = next_perm(c, x->n);
    if (x->width == 4) {
      if (arith) apply_sharded<uint32_t, true>(c, x, pi, false);
      else apply_sharded<uint32_t, false>(c, x, pi, false);
    } else {
      if (arith) apply_sharded<uint64_t, true>(c, x, pi, false);
      else apply_sharded<uint64_t, false>(c, x, pi, false);
    }
  })
}
int oc_table_shuffle(oc_ctx *c, oc_vec **a_cols, size_t na, oc_vec **b_cols, size_t nb) {
  OC_TRY(table_shuffle(c, a_cols, na, b_cols, nb))
}
int oc_table_sort(oc_ctx *c, oc_vec **keys, size_t nkeys, const int *orders, const int *single_bit,
                  oc_vec **a_cols, size_t na, oc_vec **b_cols, size_t nb, int protocol) {
  OC_TRY(table_sort(c, keys, nkeys, orders, single_bit, a_cols, na, b_cols, nb, protocol))
}

int oc_quicksort(oc_ctx *c, oc_vec *v, int desc, oc_vec *perm_out) {
  OC_TRY({
    if (v->width == 4)
      quicksort_any<uint32_t, uint64_t>(c, v, desc != 0, perm_out);
    else if (v->width == 8)
      quicksort_any<uint64_t, u128>(c, v, desc != 0, perm_out);
    else
      throw std::runtime_error("quicksort: 32- or 64-bit keys");
  })
}

int oc_bitonic_sort(oc_ctx *c, oc_vec **keys, size_t nkeys, const int *orders, oc_vec **a_cols,
                    size_t na, oc_vec **b_cols, size_t nb) {
  OC_TRY({
    if (keys[0]->width == 4)
      bitonic_sort<uint32_t>(c, keys, nkeys, orders, a_cols, na, b_cols, nb);
    else
      bitonic_sort<uint64_t>(c, keys, nkeys, orders, a_cols, na, b_cols, nb);
  })
}

// specs: arrays of (in, out) vector pointers and function ids
int oc_aggregate(oc_ctx *c, oc_vec **keys, size_t nk, oc_vec **b_in, oc_vec **b_out,
                 const int *b_fn, size_t nb, oc_vec **a_in, oc_vec **a_out, const int *a_fn,
                 size_t na, int reverse, oc_vec *sel) {
  OC_TRY({
    std::vector<AggSpec> bs(nb), as(na);
    for (size_t i = 0; i < nb; i++) bs[i] = {b_in[i], b_out[i], b_fn[i]};
    for (size_t i = 0; i < na; i++) as[i] = {a_in[i], a_out[i], a_fn[i]};
    int w = nk ? keys[0]->width : nb ? b_in[0]->width : a_in[0]->width;
    if (w == 4)
      aggregate<uint32_t>(c, keys, nk, bs.data(), nb, as.data(), na, reverse != 0, sel);
    else
      aggregate<uint64_t>(c, keys, nk, bs.data(), nb, as.data(), na, reverse != 0, sel);
  })
}
int oc_distinct(oc_ctx *c, oc_vec **keys, size_t nk, oc_vec *res) {
  OC_TRY({
    if (res->width == 4)
      distinct<uint32_t>(c, keys, nk, res);
    else
      distinct<uint64_t>(c, keys, nk, res);
  })
}

// op 0: RCA add, 1: RCA subtract (x + ~y + 1), 2: PPA add, 3: PPA subtract,
// 4: rca_compare (x < y), 5: a2b of x
int oc_adder(oc_ctx *c, int op, oc_vec *x, oc_vec *y, oc_vec *z) {
  OC_TRY({
    auto run = [&](auto tag) {
      typedef decltype(tag) T;
      oc_vec *ny = (op == 1 || op == 3) ? not_copy<T>(y) : nullptr;
      switch (op) {
        case 0: rca<T>(c, x, y, false, z); break;
        case 1: rca<T>(c, x, ny, true, z); break;
        case 2: ppa<T>(c, x, y, false, z); break;
        case 3: ppa<T>(c, x, ny, true, z); break;
        case 4: rca_compare<T>(c, x, y, z); break;
        case 5: a2b<T>(c, x, z); break;
        default: throw std::runtime_error("adder: op");
      }
      delete ny;
    };
    if (x->width == 4)
      run(uint32_t{});
    else
      run(uint64_t{});
  })
}
int oc_secret_share_from(oc_ctx *c, const void *data, oc_vec *v, int arith, int owner) {
  OC_TRY({
    if (v->width == 4) {
      if (arith) secret_share<uint32_t, true>(c, (const uint32_t *)data, v, owner);
      else secret_share<uint32_t, false>(c, (const uint32_t *)data, v, owner);
    } else {
      if (arith) secret_share<uint64_t, true>(c, (const uint64_t *)data, v, owner);
      else secret_share<uint64_t, false>(c, (const uint64_t *)data, v, owner);
    }
  })
}

size_t oc_msglog_count(oc_ctx *c) { return c->log.size(); }
void oc_msglog_get(oc_ctx *c, size_t i, char *op, int *from, int *to, uint64_t *bytes) {
  const Msg &m = c->log[i];
  std::memcpy(op, m.op, sizeof(m.op));
  *from = m.from;
  *to = m.to;
  *bytes = m.bytes;
}
void oc_msglog_clear(oc_ctx *c) { c->log.clear(); }

}  // extern "C"
