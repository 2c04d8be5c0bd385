// Test infrastructure only: a driver that runs the reference's OWN code for the
// randomness, the 3-party replicated primitives and the batch schedule, so that
// the oracle (orq_oracle.cpp) can be checked against it share for share.
//
// Nothing here restates an algorithm.  The reference's headers are compiled in
// place (include path /root/reference/include, see oracle/Makefile) and called
// through their public types:
//   - AESPRGAlgorithm / CommonPRG (core/random/prg/prg_algorithm.h:139-231,
//     common_prg.h:23-151): keystream framing and nonces;
//   - gen_perm and HMShardedPermutationGenerator::getNext
//     (core/random/permutations/hm_sharded_permutation_generator.h:70-194);
//   - ZeroSharingGenerator (core/random/correlation/zero_sharing_generator.h);
//   - Replicated_3PC<int32/int64> (core/protocols/replicated_3pc.h) and
//     Protocol::reshare (core/protocols/protocol.h:500-541);
//   - the runtime's per-call batching, through the reference's own Task classes
//     (backend/common/task.h:34-94) with one worker thread.
//
// What this file supplies is only what the reference obtains from its MPI
// launch: the three parties run as three host threads, keys are passed in
// explicitly instead of drawn by randombytes and exchanged, and messages travel
// through an in-process implementation of the reference's Communicator
// interface (core/communication/communicator.h:14-183).  Key roles follow
// backend/common/rand_setup.h:90-146: party p's CommonPRG at relative rank +1 is
// keyed by pair key K_p (its own key, sent to p+1), the one at -1 by K_{p-1};
// group g's CommonPRG is keyed by group key G_g (rand_setup.h:21-77); the local
// PRG by L_p.  The key block is the oracle's: 10 x 32 bytes, L_0..L_2,
// K_0..K_2, G_{01}, G_{02}, G_{12}, G_{012}.
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "debug/orq_debug.h"
#include "core/containers/vector.h"
#include "core/containers/e_vector.h"
#include "core/communication/communicator.h"
#include "core/random/prg/common_prg.h"
#include "core/random/correlation/zero_sharing_generator.h"
#include "core/random/permutations/hm_sharded_permutation_generator.h"
#include "core/random/manager.h"
#include "core/protocols/replicated_3pc.h"
#include "backend/common/task.h"

namespace {

using Bytes = std::vector<uint8_t>;

// One FIFO of byte messages per ordered pair of parties.
struct Mailboxes {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Bytes> box[3][3];  // [from][to]

  void put(int from, int to, Bytes &&b) {
    {
      std::lock_guard<std::mutex> g(mu);
      box[from][to].push_back(std::move(b));
    }
    cv.notify_all();
  }
  Bytes take(int from, int to) {
    std::unique_lock<std::mutex> g(mu);
    cv.wait(g, [&] { return !box[from][to].empty(); });
    Bytes b = std::move(box[from][to].front());
    box[from][to].pop_front();
    return b;
  }
};

// The reference's Communicator interface over Mailboxes.  Vectors are read and
// written element by element through operator[], so batch windows and index
// mappings behave as they do for the reference's own communicators.
class InProcComm : public orq::Communicator {
  Mailboxes *mb;
  static constexpr int kParties = 3;

  int abs_id(orq::PartyID rel) const { return ((kParties + rel + currentId) % kParties + kParties) % kParties; }

  template <typename T>
  void send_vec(const orq::Vector<T> &v, orq::PartyID rel, size_t size) {
    Bytes b(size * sizeof(T));
    for (size_t i = 0; i < size; i++) {
      T e = v[i];
      std::memcpy(b.data() + i * sizeof(T), &e, sizeof(T));
    }
    bytes_sent += size * sizeof(T);
    mb->put(currentId, abs_id(rel), std::move(b));
  }
  template <typename T>
  void recv_vec(orq::Vector<T> &v, orq::PartyID rel, size_t size) {
    Bytes b = mb->take(abs_id(rel), currentId);
    if (b.size() != size * sizeof(T)) throw std::runtime_error("InProcComm: message size mismatch");
    for (size_t i = 0; i < size; i++) {
      T e;
      std::memcpy(&e, b.data() + i * sizeof(T), sizeof(T));
      v[i] = e;
    }
  }
  template <typename T>
  void send_one(T e, orq::PartyID rel) {
    Bytes b(sizeof(T));
    std::memcpy(b.data(), &e, sizeof(T));
    bytes_sent += sizeof(T);
    mb->put(currentId, abs_id(rel), std::move(b));
  }
  template <typename T>
  void recv_one(T &e, orq::PartyID rel) {
    Bytes b = mb->take(abs_id(rel), currentId);
    if (b.size() != sizeof(T)) throw std::runtime_error("InProcComm: message size mismatch");
    std::memcpy(&e, b.data(), sizeof(T));
  }

 public:
  InProcComm(int id, Mailboxes *m) : orq::Communicator(id), mb(m) {}

#define ORQ_REF_COMM_TYPE(T)                                                                     \
  void sendShare(T e, orq::PartyID id) override { send_one(e, id); }                             \
  void receiveShare(T &e, orq::PartyID id) override { recv_one(e, id); }                         \
  void sendShares(const orq::Vector<T> &s, orq::PartyID id, size_t size) override {              \
    send_vec(s, id, size);                                                                       \
  }                                                                                              \
  void receiveShares(orq::Vector<T> &s, orq::PartyID id, size_t size) override {                 \
    recv_vec(s, id, size);                                                                       \
  }                                                                                              \
  void exchangeShares(orq::Vector<T> sent, orq::Vector<T> &recv, orq::PartyID id,                \
                      size_t size) override {                                                    \
    send_vec(sent, id, size);                                                                    \
    recv_vec(recv, id, size);                                                                    \
  }                                                                                              \
  void exchangeShares(orq::Vector<T> sent, orq::Vector<T> &recv, orq::PartyID to,                \
                      orq::PartyID from, size_t size) override {                                 \
    send_vec(sent, to, size);                                                                    \
    recv_vec(recv, from, size);                                                                  \
  }                                                                                              \
  void sendShares(const std::vector<orq::Vector<T>> &s, std::vector<orq::PartyID> ids) override { \
    for (size_t i = 0; i < ids.size(); i++) send_vec(s[i], ids[i], s[i].size());                 \
  }                                                                                              \
  void receiveBroadcast(std::vector<orq::Vector<T>> &s, std::vector<orq::PartyID> ids) override { \
    for (size_t i = 0; i < ids.size(); i++) recv_vec(s[i], ids[i], s[i].size());                 \
  }                                                                                              \
  void exchangeShares(const std::vector<orq::Vector<T>> &s, std::vector<orq::Vector<T>> &r,      \
                      std::vector<orq::PartyID> to, std::vector<orq::PartyID> from) override {   \
    for (size_t i = 0; i < to.size(); i++) send_vec(s[i], to[i], s[i].size());                   \
    for (size_t i = 0; i < from.size(); i++) recv_vec(r[i], from[i], r[i].size());               \
  }

  ORQ_REF_COMM_TYPE(int8_t)
  ORQ_REF_COMM_TYPE(int16_t)
  ORQ_REF_COMM_TYPE(int32_t)
  ORQ_REF_COMM_TYPE(int64_t)
#undef ORQ_REF_COMM_TYPE
  // __int128 has no single-element overloads in the interface
  void sendShares(const orq::Vector<__int128_t> &s, orq::PartyID id, size_t size) override {
    send_vec(s, id, size);
  }
  void receiveShares(orq::Vector<__int128_t> &s, orq::PartyID id, size_t size) override {
    recv_vec(s, id, size);
  }
  void exchangeShares(orq::Vector<__int128_t> sent, orq::Vector<__int128_t> &recv,
                      orq::PartyID id, size_t size) override {
    send_vec(sent, id, size);
    recv_vec(recv, id, size);
  }
  void exchangeShares(orq::Vector<__int128_t> sent, orq::Vector<__int128_t> &recv,
                      orq::PartyID to, orq::PartyID from, size_t size) override {
    send_vec(sent, to, size);
    recv_vec(recv, from, size);
  }
  void sendShares(const std::vector<orq::Vector<__int128_t>> &s,
                  std::vector<orq::PartyID> ids) override {
    for (size_t i = 0; i < ids.size(); i++) send_vec(s[i], ids[i], s[i].size());
  }
  void receiveBroadcast(std::vector<orq::Vector<__int128_t>> &s,
                        std::vector<orq::PartyID> ids) override {
    for (size_t i = 0; i < ids.size(); i++) recv_vec(s[i], ids[i], s[i].size());
  }
  void exchangeShares(const std::vector<orq::Vector<__int128_t>> &s,
                      std::vector<orq::Vector<__int128_t>> &r, std::vector<orq::PartyID> to,
                      std::vector<orq::PartyID> from) override {
    for (size_t i = 0; i < to.size(); i++) send_vec(s[i], to[i], s[i].size());
    for (size_t i = 0; i < from.size(); i++) recv_vec(r[i], from[i], r[i].size());
  }
};

template <typename T>
using EV = orq::EVector<T, 2>;
template <typename T>
using Proto = orq::Replicated_3PC<T, std::vector<T>, orq::Vector<T>, EV<T>>;

std::shared_ptr<orq::random::CommonPRG> make_prg(const uint8_t *key, int rank, uint64_t skip = 0) {
  std::vector<unsigned char> seed(key, key + 32);
  auto alg = std::make_unique<orq::random::AESPRGAlgorithm>(seed);
  for (uint64_t i = 0; i < skip; i++) alg->incrementNonce();
  return std::make_shared<orq::random::CommonPRG>(std::move(alg), rank);
}

const std::vector<std::set<int>> kGroups = {{0, 1}, {0, 2}, {1, 2}, {0, 1, 2}};

struct Party {
  int id;
  std::shared_ptr<orq::random::CommonPRG> local;
  std::shared_ptr<orq::random::CommonPRGManager> mgr;
  std::shared_ptr<orq::random::ZeroSharingGenerator> zsg;
  std::unique_ptr<orq::random::RandomnessManager> rm;
  std::unique_ptr<InProcComm> comm;
  std::unique_ptr<Proto<int32_t>> p32;
  std::unique_ptr<Proto<int64_t>> p64;
};

}  // namespace

struct rf_ctx {
  Mailboxes mb;
  Party P[3];
  long batch_size;
};

namespace {

// Run f(party) on three host threads, as the reference's three processes do.
template <typename F>
int run3(rf_ctx *c, F f) {
  std::string err[3];
  std::thread th[3];
  for (int p = 0; p < 3; p++)
    th[p] = std::thread([&, p] {
      try {
        f(c->P[p]);
      } catch (const std::exception &e) {
        err[p] = e.what();
      } catch (...) {
        err[p] = "unknown exception";
      }
    });
  for (auto &t : th) t.join();
  for (auto &e : err)
    if (!e.empty()) return -1;
  return 0;
}

// Faithful layout of one replicated vector: buf[(party * 2 + idx) * n + i].
template <typename T>
EV<T> load_ev(const void *buf, int party, size_t n) {
  const T *b = static_cast<const T *>(buf);
  std::vector<T> s0(b + (party * 2 + 0) * n, b + (party * 2 + 0) * n + n);
  std::vector<T> s1(b + (party * 2 + 1) * n, b + (party * 2 + 1) * n + n);
  return EV<T>(std::vector<orq::Vector<T>>{orq::Vector<T>(std::move(s0)), orq::Vector<T>(std::move(s1))});
}
template <typename T>
void store_ev(const EV<T> &v, void *buf, int party, size_t n) {
  T *b = static_cast<T *>(buf);
  for (int k = 0; k < 2; k++) {
    const orq::Vector<T> &s = v(k);
    for (size_t i = 0; i < n; i++) b[(party * 2 + k) * n + i] = s[i];
  }
}

template <typename T>
Proto<T> *proto(Party &pt) {
  if constexpr (sizeof(T) == 4)
    return pt.p32.get();
  else
    return pt.p64.get();
}

template <typename T>
int and_mul(rf_ctx *c, size_t n, const void *x, const void *y, void *z, bool arith) {
  return run3(c, [&](Party &pt) {
    EV<T> X = load_ev<T>(x, pt.id, n), Y = load_ev<T>(y, pt.id, n), Z(n);
    orq::service::Task_2_ref<EV<T>, EV<T>> task(
        X, Y, Z, 0, n, c->batch_size, [&](EV<T> &a, EV<T> &b, EV<T> &r) {
          if (arith)
            proto<T>(pt)->multiply_a(a, b, r);
          else
            proto<T>(pt)->and_b(a, b, r);
        });
    task.execute();
    store_ev<T>(Z, z, pt.id, n);
  });
}

template <typename T>
int b2a_bit(rf_ctx *c, size_t n, const void *x, void *y) {
  return run3(c, [&](Party &pt) {
    EV<T> X = load_ev<T>(x, pt.id, n), Y(n);
    orq::service::Task_1_ref<EV<T>, EV<T>> task(
        X, Y, 0, n, c->batch_size, [&](EV<T> &a, EV<T> &r) { proto<T>(pt)->b2a_bit(a, r); });
    task.execute();
    store_ev<T>(Y, y, pt.id, n);
  });
}

template <typename T>
int reshare(rf_ctx *c, size_t n, void *v, int group, bool binary) {
  // each party reads and writes only its own two rows of v
  return run3(c, [&](Party &pt) {
    EV<T> V = load_ev<T>(v, pt.id, n);
    orq::service::Task_1_void<EV<T>> task(V, 0, n, c->batch_size, [&](EV<T> &a) {
      proto<T>(pt)->reshare(a, kGroups[group], binary);
    });
    task.execute();
    store_ev<T>(V, v, pt.id, n);
  });
}

template <typename T>
int secret_share(rf_ctx *c, size_t n, const void *data, int owner, bool binary, void *out) {
  return run3(c, [&](Party &pt) {
    // only the owner's data is read; the others pass a vector of the same size
    std::vector<T> d(n, 0);
    if (pt.id == owner) std::memcpy(d.data(), data, n * sizeof(T));
    orq::Vector<T> D(std::move(d));
    EV<T> R(n);
    orq::service::Task_1_ref<orq::Vector<T>, EV<T>> task(
        D, R, 0, n, c->batch_size, [&](orq::Vector<T> &a, EV<T> &r) {
          r = binary ? proto<T>(pt)->secret_share_b(a, owner) : proto<T>(pt)->secret_share_a(a, owner);
        });
    task.execute();
    store_ev<T>(R, out, pt.id, n);
  });
}

template <typename T>
int open(rf_ctx *c, size_t n, const void *v, bool binary, void *out /* [3][n] */) {
  return run3(c, [&](Party &pt) {
    EV<T> V = load_ev<T>(v, pt.id, n);
    orq::Vector<T> R(n);
    orq::service::Task_1_ref<EV<T>, orq::Vector<T>> task(
        V, R, 0, n, c->batch_size, [&](EV<T> &a, orq::Vector<T> &r) {
          r = binary ? proto<T>(pt)->open_shares_b(a) : proto<T>(pt)->open_shares_a(a);
        });
    task.execute();
    T *o = static_cast<T *>(out) + pt.id * n;
    for (size_t i = 0; i < n; i++) o[i] = R[i];
  });
}

template <typename T>
int zero_sharing(rf_ctx *c, size_t n, bool binary, void *out /* [3][n] */) {
  return run3(c, [&](Party &pt) {
    orq::Vector<T> z(n);
    if (binary)
      pt.zsg->getNextBinary(z);
    else
      pt.zsg->getNextArithmetic(z);
    T *o = static_cast<T *>(out) + pt.id * n;
    for (size_t i = 0; i < n; i++) o[i] = z[i];
  });
}

}  // namespace

extern "C" {

rf_ctx *rf_create(const uint8_t *keys /* 10 x 32 bytes */, long batch_size) {
  if (sodium_init() < 0) return nullptr;
  rf_ctx *c = new rf_ctx;
  c->batch_size = batch_size;
  for (int p = 0; p < 3; p++) {
    Party &pt = c->P[p];
    pt.id = p;
    pt.local = make_prg(keys + 32 * p, p);
    pt.mgr = std::make_shared<orq::random::CommonPRGManager>(3);
    auto next = make_prg(keys + 32 * (3 + p), p);
    auto prev = make_prg(keys + 32 * (3 + (p + 2) % 3), p);
    pt.mgr->add(next, +1);
    pt.mgr->add(prev, -1);
    for (int g = 0; g < 4; g++)
      if (kGroups[g].contains(p)) {
        auto gp = make_prg(keys + 32 * (6 + g), p);
        pt.mgr->add(gp, kGroups[g]);
      }
    pt.zsg = std::make_shared<orq::random::ZeroSharingGenerator>(3, pt.mgr, p);
    pt.rm = std::make_unique<orq::random::RandomnessManager>(pt.local, pt.mgr, pt.zsg);
    pt.comm = std::make_unique<InProcComm>(p, &c->mb);
    pt.p32 = std::make_unique<Proto<int32_t>>(p, pt.comm.get(), pt.rm.get());
    pt.p64 = std::make_unique<Proto<int64_t>>(p, pt.comm.get(), pt.rm.get());
  }
  return c;
}

void rf_destroy(rf_ctx *c) { delete c; }

uint64_t rf_bytes_sent(rf_ctx *c, int party) { return c->P[party].comm->getBytesSent(); }

#define RF_DISPATCH(w, call4, call8) ((w) == 4 ? (call4) : (w) == 8 ? (call8) : -2)

int rf_and_b(rf_ctx *c, int w, size_t n, const void *x, const void *y, void *z) {
  return RF_DISPATCH(w, (and_mul<int32_t>(c, n, x, y, z, false)), (and_mul<int64_t>(c, n, x, y, z, false)));
}
int rf_mul_a(rf_ctx *c, int w, size_t n, const void *x, const void *y, void *z) {
  return RF_DISPATCH(w, (and_mul<int32_t>(c, n, x, y, z, true)), (and_mul<int64_t>(c, n, x, y, z, true)));
}
int rf_b2a_bit(rf_ctx *c, int w, size_t n, const void *x, void *y) {
  return RF_DISPATCH(w, (b2a_bit<int32_t>(c, n, x, y)), (b2a_bit<int64_t>(c, n, x, y)));
}
int rf_reshare(rf_ctx *c, int w, size_t n, void *v, int group, int binary) {
  if (group < 0 || group > 2) return -2;
  return RF_DISPATCH(w, (reshare<int32_t>(c, n, v, group, binary)), (reshare<int64_t>(c, n, v, group, binary)));
}
int rf_secret_share(rf_ctx *c, int w, size_t n, const void *data, int owner, int binary, void *out) {
  if (owner < 0 || owner > 2) return -2;
  return RF_DISPATCH(w, (secret_share<int32_t>(c, n, data, owner, binary, out)),
                     (secret_share<int64_t>(c, n, data, owner, binary, out)));
}
int rf_open(rf_ctx *c, int w, size_t n, const void *v, int binary, void *out) {
  return RF_DISPATCH(w, (open<int32_t>(c, n, v, binary, out)), (open<int64_t>(c, n, v, binary, out)));
}
int rf_zero_sharing(rf_ctx *c, int w, size_t n, int binary, void *out) {
  return RF_DISPATCH(w, (zero_sharing<int32_t>(c, n, binary, out)), (zero_sharing<int64_t>(c, n, binary, out)));
}

// HMShardedPermutationGenerator::getNext(n) of every party, `count` times:
// out[((k * 3 + party) * 4 + group) * n + i], groups {01},{02},{12},{012};
// a group the party is not in is left untouched.
int rf_sharded_perms(rf_ctx *c, size_t n, int count, int32_t *out) {
  return run3(c, [&](Party &pt) {
    orq::random::HMShardedPermutationGenerator gen(pt.id, pt.mgr, kGroups);
    for (int k = 0; k < count; k++) {
      auto sp = std::static_pointer_cast<orq::random::HMShardedPermutation>(gen.getNext(n));
      auto m = sp->getPermMap();
      for (int g = 0; g < 4; g++) {
        auto it = m->find(kGroups[g]);
        if (it == m->end()) continue;
        if (it->second.size() != n) throw std::runtime_error("sharded perm size");
        std::memcpy(out + ((size_t)(k * 3 + pt.id) * 4 + g) * n, it->second.data(), n * sizeof(int32_t));
      }
    }
  });
}

// The keystream a CommonPRG keyed by `key`, `skip` nonces in, returns for
// getNext(Vector<T>) of `count` elements of `width` bytes (1, 4 or 8).
int rf_keystream(const uint8_t *key, uint64_t skip, int width, size_t count, void *out) {
  if (sodium_init() < 0) return -1;
  auto prg = make_prg(key, 0, skip);
  if (width == 1) {
    orq::Vector<uint8_t> v(count);
    prg->getNext(v);
    for (size_t i = 0; i < count; i++) static_cast<uint8_t *>(out)[i] = v[i];
  } else if (width == 4) {
    orq::Vector<uint32_t> v(count);
    prg->getNext(v);
    for (size_t i = 0; i < count; i++) static_cast<uint32_t *>(out)[i] = v[i];
  } else if (width == 8) {
    orq::Vector<uint64_t> v(count);
    prg->getNext(v);
    for (size_t i = 0; i < count; i++) static_cast<uint64_t *>(out)[i] = v[i];
  } else {
    return -2;
  }
  return 0;
}

// `count` successive gen_perm calls of size n on one CommonPRG, `skip` nonces in.
int rf_gen_perm(const uint8_t *key, uint64_t skip, size_t n, int count, int32_t *out) {
  if (sodium_init() < 0) return -1;
  auto prg = make_prg(key, 0, skip);
  std::vector<int> perm(n);
  for (int k = 0; k < count; k++) {
    orq::random::gen_perm(perm, prg);
    std::memcpy(out + (size_t)k * n, perm.data(), n * sizeof(int32_t));
  }
  return 0;
}

// The runtime's batch size for a call over [start, end) (task.h:34-44).
long rf_equal_batches(size_t start, size_t end, long batches) {
  return (long)orq::service::compute_equal_batches(start, end, batches);
}

}  // extern "C"
