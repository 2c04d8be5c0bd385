// Test infrastructure only: one party of a 3-party run of the reference's OWN
// runtime and operators, for checking the oracle's composite operators share
// for share.  Three copies of this program run side by side (oracle/ref.py
// starts them), one per party, as the reference's three processes do.
//
// The reference's headers are compiled in place (oracle/Makefile, target
// `ref`); everything computed goes through them: RunTime and its workers
// (backend/common/runtime.h, worker.h), Replicated_3PC via the reference's
// DefaultProtocolFactory, HMShardedPermutationGenerator behind the
// PermutationManager, and the operators (radix_sort, quicksort, table sort,
// aggregate, distinct, inner join, the comparison and adder circuits) called
// as the reference's tests call them, and radix sort timed as its
// micro-benchmark times it.
//
// What this file supplies is what orq_runtime_init (backend/common/setup.h:8-57)
// gets from its launcher: the party id, a Communicator, and the randomness
// setup.  The communicator is an implementation of the reference's interface
// (core/communication/communicator.h:14-183) over a socket per peer.  The
// randomness is set up as setup_random_generation does
// (backend/common/rand_setup.h:306-381), with the same objects in the same
// roles, but keyed from an explicit 10 x 32-byte key block (the oracle's:
// L_0..L_2, K_0..K_2, G_{01}, G_{02}, G_{12}, G_{012}) instead of keys drawn
// by randombytes and exchanged.
//
// Usage: orq_ref_party <party> <fds> <in> <out>
//   fds: per worker thread, the sockets to party+1 and party+2 ("n0,p0,n1,p1")
//   in:  i64 threads | keys[320 x threads] | i64 batch_size | i64 scenario |
//        i64 params[8] | payload
//   out: the scenario's arrays, then u64 bytes sent (all threads)
#include <sys/socket.h>
#include <unistd.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <fstream>
#include <mutex>
#include <thread>

#include "orq.h"

using namespace orq::service;
using namespace orq::operators;
namespace R3 = orq::service::nocopy_service::replicated_3pc;

namespace {

// ---- Communicator over one stream socket per peer ---------------------------
// Sends are queued to a sender thread per peer, so that a party never blocks
// on a send while its peer blocks on a send to it (exchangeShares).
class SocketComm : public orq::Communicator {
  struct Peer {
    int fd = -1;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::vector<uint8_t>> q;
    bool stop = false;
    std::thread th;
  };
  Peer peers[3];

  int abs_id(orq::PartyID rel) const { return ((3 + rel + currentId) % 3 + 3) % 3; }

  void sender(Peer *p) {
    for (;;) {
      std::vector<uint8_t> b;
      {
        std::unique_lock<std::mutex> g(p->mu);
        p->cv.wait(g, [&] { return p->stop || !p->q.empty(); });
        if (p->q.empty()) return;
        b = std::move(p->q.front());
        p->q.pop_front();
      }
      size_t off = 0;
      while (off < b.size()) {
        ssize_t k = ::send(p->fd, b.data() + off, b.size() - off, 0);
        if (k <= 0) throw std::runtime_error("SocketComm: send failed");
        off += (size_t)k;
      }
    }
  }
  void put(int to, std::vector<uint8_t> &&b) {
    Peer &p = peers[to];
    {
      std::lock_guard<std::mutex> g(p.mu);
      p.q.push_back(std::move(b));
    }
    p.cv.notify_all();
  }
  void get(int from, uint8_t *dst, size_t bytes) {
    size_t off = 0;
    while (off < bytes) {
      ssize_t k = ::recv(peers[from].fd, dst + off, bytes - off, 0);
      if (k <= 0) throw std::runtime_error("SocketComm: recv failed");
      off += (size_t)k;
    }
  }

  template <typename T>
  void send_vec(const orq::Vector<T> &v, orq::PartyID rel, size_t size) {
    std::vector<uint8_t> b(size * sizeof(T));
    for (size_t i = 0; i < size; i++) {
      T e = v[i];
      std::memcpy(b.data() + i * sizeof(T), &e, sizeof(T));
    }
    bytes_sent += size * sizeof(T);
    put(abs_id(rel), std::move(b));
  }
  template <typename T>
  void recv_vec(orq::Vector<T> &v, orq::PartyID rel, size_t size) {
    std::vector<uint8_t> b(size * sizeof(T));
    get(abs_id(rel), b.data(), b.size());
    for (size_t i = 0; i < size; i++) {
      T e;
      std::memcpy(&e, b.data() + i * sizeof(T), sizeof(T));
      v[i] = e;
    }
  }
  template <typename T>
  void send_one(T e, orq::PartyID rel) {
    std::vector<uint8_t> b(sizeof(T));
    std::memcpy(b.data(), &e, sizeof(T));
    bytes_sent += sizeof(T);
    put(abs_id(rel), std::move(b));
  }
  template <typename T>
  void recv_one(T &e, orq::PartyID rel) {
    get(abs_id(rel), reinterpret_cast<uint8_t *>(&e), sizeof(T));
  }

 public:
  SocketComm(int id, int fd_next, int fd_prev) : orq::Communicator(id) {
    peers[(id + 1) % 3].fd = fd_next;
    peers[(id + 2) % 3].fd = fd_prev;
    for (int q = 0; q < 3; q++)
      if (q != id) peers[q].th = std::thread(&SocketComm::sender, this, &peers[q]);
  }
  // every queued message sent: the sender threads drain their queues and end
  void flush() {
    for (int q = 0; q < 3; q++) {
      if (q == currentId || !peers[q].th.joinable()) continue;
      {
        std::lock_guard<std::mutex> g(peers[q].mu);
        peers[q].stop = true;
      }
      peers[q].cv.notify_all();
      peers[q].th.join();
    }
  }
  ~SocketComm() { flush(); }

#define ORQ_REF_COMM_TYPE(T)                                                                      \
  void sendShare(T e, orq::PartyID id) override { send_one(e, id); }                              \
  void receiveShare(T &e, orq::PartyID id) override { recv_one(e, id); }
  ORQ_REF_COMM_TYPE(int8_t)
  ORQ_REF_COMM_TYPE(int16_t)
  ORQ_REF_COMM_TYPE(int32_t)
  ORQ_REF_COMM_TYPE(int64_t)
#undef ORQ_REF_COMM_TYPE
#define ORQ_REF_COMM_VEC(T)                                                                        \
  void sendShares(const orq::Vector<T> &s, orq::PartyID id, size_t size) override {                \
    send_vec(s, id, size);                                                                         \
  }                                                                                                \
  void receiveShares(orq::Vector<T> &s, orq::PartyID id, size_t size) override {                   \
    recv_vec(s, id, size);                                                                         \
  }                                                                                                \
  void exchangeShares(orq::Vector<T> sent, orq::Vector<T> &recv, orq::PartyID id, size_t size)     \
      override {                                                                                   \
    send_vec(sent, id, size);                                                                      \
    recv_vec(recv, id, size);                                                                      \
  }                                                                                                \
  void exchangeShares(orq::Vector<T> sent, orq::Vector<T> &recv, orq::PartyID to,                  \
                      orq::PartyID from, size_t size) override {                                   \
    send_vec(sent, to, size);                                                                      \
    recv_vec(recv, from, size);                                                                    \
  }                                                                                                \
  void sendShares(const std::vector<orq::Vector<T>> &s, std::vector<orq::PartyID> ids) override {  \
    for (size_t i = 0; i < ids.size(); i++) send_vec(s[i], ids[i], s[i].size());                   \
  }                                                                                                \
  void receiveBroadcast(std::vector<orq::Vector<T>> &s, std::vector<orq::PartyID> ids) override {  \
    for (size_t i = 0; i < ids.size(); i++) recv_vec(s[i], ids[i], s[i].size());                   \
  }                                                                                                \
  void exchangeShares(const std::vector<orq::Vector<T>> &s, std::vector<orq::Vector<T>> &r,        \
                      std::vector<orq::PartyID> to, std::vector<orq::PartyID> from) override {     \
    for (size_t i = 0; i < to.size(); i++) send_vec(s[i], to[i], s[i].size());                     \
    for (size_t i = 0; i < from.size(); i++) recv_vec(r[i], from[i], r[i].size());                 \
  }
  ORQ_REF_COMM_VEC(int8_t)
  ORQ_REF_COMM_VEC(int16_t)
  ORQ_REF_COMM_VEC(int32_t)
  ORQ_REF_COMM_VEC(int64_t)
  ORQ_REF_COMM_VEC(__int128_t)
#undef ORQ_REF_COMM_VEC
};

std::shared_ptr<orq::random::CommonPRG> make_prg(const uint8_t *key, int rank) {
  std::vector<unsigned char> seed(key, key + 32);
  return std::make_shared<orq::random::CommonPRG>(std::make_unique<orq::random::AESPRGAlgorithm>(seed), rank);
}

// setup_random_generation (rand_setup.h:306-381) for 3 parties, explicit keys
std::unique_ptr<orq::random::RandomnessManager> make_randomness(int rank, const uint8_t *keys,
                                                           const std::vector<std::set<int>> &groups) {
  auto localPRG = make_prg(keys + 32 * rank, rank);
  auto mgr = std::make_shared<orq::random::CommonPRGManager>(3);
  // create_common_prg_among_group for every group, then the everyone group
  // (already present, and CommonPRGManager::add keeps the first entry)
  const std::vector<std::set<int>> key_groups = {{0, 1}, {0, 2}, {1, 2}, {0, 1, 2}};
  auto group_prg = [&](const std::set<int> &g) {
    if (!g.contains(rank)) return;
    for (int k = 0; k < 4; k++)
      if (key_groups[k] == g) {
        auto prg = make_prg(keys + 32 * (6 + k), rank);
        mgr->add(prg, g);
      }
  };
  for (const auto &g : groups) group_prg(g);
  group_prg({0, 1, 2});
  // setup_3pc_common_prgs (rand_setup.h:90-146): relative rank 1 is the
  // party's own key K_rank, relative rank 2 the key K_{rank-1} it received
  auto next = make_prg(keys + 32 * (3 + rank), rank);
  auto prev = make_prg(keys + 32 * (3 + (rank + 2) % 3), rank);
  mgr->add(next, 1);
  mgr->add(prev, 2);
  std::map<orq::random::RandomnessManager::typed_correlation, orq::random::CorrelationGenerator *> cg;
  auto sharded = new orq::random::HMShardedPermutationGenerator(rank, mgr, groups);
  cg[{typeid(int8_t), orq::random::Correlation::ShardedPermutation}] = sharded;
  cg[{typeid(int16_t), orq::random::Correlation::ShardedPermutation}] = sharded;
  cg[{typeid(int32_t), orq::random::Correlation::ShardedPermutation}] = sharded;
  cg[{typeid(int64_t), orq::random::Correlation::ShardedPermutation}] = sharded;
  cg[{typeid(__int128_t), orq::random::Correlation::ShardedPermutation}] = sharded;
  auto zsg = std::make_shared<orq::random::ZeroSharingGenerator>(3, mgr, rank);
  return std::make_unique<orq::random::RandomnessManager>(localPRG, mgr, zsg, cg);
}

struct Out {
  std::vector<uint8_t> bytes;
  template <typename T>
  void put(const orq::Vector<T> &v) {
    for (size_t i = 0; i < v.size(); i++) {
      T e = v[i];
      const uint8_t *p = reinterpret_cast<const uint8_t *>(&e);
      bytes.insert(bytes.end(), p, p + sizeof(T));
    }
  }
  template <typename T>
  void put_ev(const orq::EVector<T, 2> &v) {
    put(v(0));
    put(v(1));
  }
  void put_u64(uint64_t x) {
    const uint8_t *p = reinterpret_cast<const uint8_t *>(&x);
    bytes.insert(bytes.end(), p, p + 8);
  }
};

template <typename T>
orq::Vector<T> vec_from(const uint8_t *p, size_t n) {
  orq::Vector<T> v(n);
  for (size_t i = 0; i < n; i++) std::memcpy(&v[i], p + i * sizeof(T), sizeof(T));
  return v;
}

// scenario 1: v = secret_share_b(data, 0); perm = radix_sort(v, order, bits)
// (radixsort.h:205-229, as tests/test_sorting.cpp:134-174 calls it).
// params: width, bits, desc.  out: v's two shares, perm's two shares.
template <typename T>
void radix(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {
  auto data = vec_from<T>(payload, n);
  R3::BSharedVector<T> v = R3::secret_share_b(data, 0);
  auto order = prm[2] ? SortOrder::DESC : SortOrder::ASC;
  auto perm = radix_sort(v, order, (size_t)prm[1]);
  out.put_ev(v.vector);
  out.put_ev(perm.sharedVector.vector);
}

// scenario 2: quicksort (quicksort.h:28-246), 32- or 64-bit keys, same
// outputs.  params: width, -, desc
template <typename T>
void quick(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {
  auto data = vec_from<T>(payload, n);
  R3::BSharedVector<T> v = R3::secret_share_b(data, 0);
  auto order = prm[2] ? SortOrder::DESC : SortOrder::ASC;
  auto perm = quicksort(v, order);
  out.put_ev(v.vector);
  out.put_ev(perm.sharedVector.vector);
}

// ---- tables (EncodedTable, encoded_table.h) --------------------------------

template <typename T>
using Table = R3::EncodedTable<T>;
template <typename T>
using TA = R3::ASharedVector<T>;
template <typename T>
using TB = R3::BSharedVector<T>;

// every column in the table's own (std::map) order: name, encoding
// (0 A-shared, 1 B-shared), rows, and this party's two shares
template <typename T>
void put_table(Table<T> &t, Out &out) {
  auto names = t.getColumnNames();
  out.put_u64(names.size());
  for (const auto &nm : names) {
    out.put_u64(nm.size());
    out.bytes.insert(out.bytes.end(), nm.begin(), nm.end());
    const bool b = t[nm].encoding == orq::Encoding::BShared;
    out.put_u64(b ? 1 : 0);
    if (b) {
      auto v = t.asBSharedVector(nm);
      out.put_u64(v.size());
      out.put_ev(v.vector);
    } else {
      auto v = t.asASharedVector(nm);
      out.put_u64(v.size());
      out.put_ev(v.vector);
    }
  }
}

// `ncols` plaintext columns of n elements each, party 0's data
template <typename T>
std::vector<orq::Vector<T>> columns(const uint8_t *&p, size_t ncols, size_t n) {
  std::vector<orq::Vector<T>> c;
  for (size_t k = 0; k < ncols; k++) {
    c.push_back(vec_from<T>(p, n));
    p += n * sizeof(T);
  }
  return c;
}

// scenario 3: {[K1], ([K2],) [P], A} shared from party 0, then
// sort(keys, order, protocol) (encoded_table.h:900-998).
// params: width, nkeys, protocol (SortingProtocol), desc
template <typename T>
void table_sort(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {
  const int nkeys = (int)prm[1];
  std::vector<std::string> schema = {"[K1]"};
  if (nkeys == 2) schema.push_back("[K2]");
  schema.push_back("[P]");
  schema.push_back("A");
  const uint8_t *p = payload;
  auto cols = columns<T>(p, schema.size(), n);
  Table<T> t = R3::secret_share(cols, schema);
  std::vector<std::string> keys(schema.begin(), schema.begin() + nkeys);
  t.sort(keys, prm[3] ? orq::operators::SortOrder::DESC : orq::operators::SortOrder::ASC,
         (orq::SortingProtocol)prm[2]);
  put_table(t, out);
}

// scenario 4: R(key, attr) JOIN L(key, value) ON key copying R's attribute,
// then GROUP BY key SUM(value), COUNT(*), MAX(value) (join.h:31-233,
// encoded_table.h:1059-1188).  params: width, nr, nl
template <typename T>
void join_groupby(const int64_t *prm, const uint8_t *payload, size_t, Out &out) {
  using namespace orq::aggregators;
  const size_t nr = (size_t)prm[1], nl = (size_t)prm[2];
  const uint8_t *p = payload;
  auto rc = columns<T>(p, 2, nr);
  auto lc = columns<T>(p, 2, nl);
  lc.push_back(lc[1]);
  Table<T> r = R3::secret_share(rc, {"[key]", "[attr]"});
  Table<T> l = R3::secret_share(lc, {"[key]", "value", "[vmax]"});
  auto j = r.inner_join(l, {"[key]"}, {{"[attr]", "[attr]", copy<TB<T>>}});
  j.addColumns({"cnt"});
  j.aggregate({"[key]"}, {{"value", "value", sum<TA<T>>},
                          {"value", "cnt", count<TA<T>>},
                          {"[vmax]", "[vmax]", max<TB<T>>}});
  put_table(j, out);
}

// scenario 5: {[K], [V2]} then distinct({[K], [V2]}) (encoded_table.h:1198-1237)
template <typename T>
void table_distinct(const int64_t *, const uint8_t *payload, size_t n, Out &out) {
  const uint8_t *p = payload;
  auto cols = columns<T>(p, 2, n);
  Table<T> t = R3::secret_share(cols, {"[K]", "[V2]"});
  t.distinct({"[K]", "[V2]"});
  put_table(t, out);
}

// scenario 6: {[K], A, C, [B], [M]} then aggregate([K]; sum A, count into C,
// max [B], min [M]), forward or reverse.  params: width, reverse
template <typename T>
void table_aggregate(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {
  using namespace orq::aggregators;
  const uint8_t *p = payload;
  auto cols = columns<T>(p, 5, n);
  Table<T> t = R3::secret_share(cols, {"[K]", "A", "C", "[B]", "[M]"});
  orq::relational::AggregationOptions opt;
  opt.reverse = prm[1] != 0;
  t.aggregate({"[K]"},
              {{"A", "A", sum<TA<T>>},
               {"A", "C", count<TA<T>>},
               {"[B]", "[B]", max<TB<T>>},
               {"[M]", "[M]", min<TB<T>>}},
              opt);
  put_table(t, out);
}

// scenario 7: {[K], A} then shuffle() (encoded_table.h:1003-1018)
template <typename T>
void table_shuffle(const int64_t *, const uint8_t *payload, size_t n, Out &out) {
  const uint8_t *p = payload;
  auto cols = columns<T>(p, 2, n);
  Table<T> t = R3::secret_share(cols, {"[K]", "A"});
  t.shuffle();
  put_table(t, out);
}

// scenario 8: x, y boolean-shared and x arithmetic-shared from party 0, then
// ==, !=, <, >, <=, >= (b_shared_vector.h:493-590, circuits.h:231-295), full
// b2a (:296-329), + and - (the ripple-carry adder, circuits.h:35-77) and a2b
// (a_shared_vector.h:138-144), in that order.  params: width
template <typename T>
void circuits(const int64_t *, const uint8_t *payload, size_t n, Out &out) {
  auto x = vec_from<T>(payload, n);
  auto y = vec_from<T>(payload + n * sizeof(T), n);
  TB<T> bx = R3::secret_share_b(x, 0);
  TB<T> by = R3::secret_share_b(y, 0);
  TA<T> ax = R3::secret_share_a(x, 0);
  out.put_ev((*(bx == by)).vector);
  out.put_ev((*(bx != by)).vector);
  out.put_ev((*(bx < by)).vector);
  out.put_ev((*(bx > by)).vector);
  out.put_ev((*(bx <= by)).vector);
  out.put_ev((*(bx >= by)).vector);
  out.put_ev((*bx.b2a()).vector);
  out.put_ev((*(bx + by)).vector);
  out.put_ev((*(bx - by)).vector);
  out.put_ev((*ax.a2b()).vector);
}

// scenario 9: the CPU baseline.  v = secret_share_b(data, 0) (not timed),
// then radix_sort(v) timed as bench/micro/micro_sorting.cpp:31-38 times it,
// with its permutations reserved first (radixsort.h:121-128) and timed apart,
// so that the online part can be reported too.  params: width, bits.
// out: v's two shares, then f64 seconds for the reserve and for the whole.
template <typename T>
void radix_bench(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {
  auto data = vec_from<T>(payload, n);
  R3::BSharedVector<T> v = R3::secret_share_b(data, 0);
  const size_t bits = (size_t)prm[1];
  const auto t0 = std::chrono::steady_clock::now();
  orq::random::PermutationManager::get()->reserve(n, 1, bits + 1);
  const auto t1 = std::chrono::steady_clock::now();
  auto perm = radix_sort(v, SortOrder::ASC, bits);
  const auto t2 = std::chrono::steady_clock::now();
  (void)perm;
  out.put_ev(v.vector);
  const double s_res = std::chrono::duration<double>(t1 - t0).count();
  const double s_all = std::chrono::duration<double>(t2 - t0).count();
  out.bytes.insert(out.bytes.end(), reinterpret_cast<const uint8_t *>(&s_res),
                   reinterpret_cast<const uint8_t *>(&s_res) + 8);
  out.bytes.insert(out.bytes.end(), reinterpret_cast<const uint8_t *>(&s_all),
                   reinterpret_cast<const uint8_t *>(&s_all) + 8);
}

#define ORQ_REF_BY_WIDTH(fn)                                                             \
  void fn##_any(const int64_t *prm, const uint8_t *payload, size_t n, Out &out) {           \
    if (prm[0] == 4)                                                                        \
      fn<int32_t>(prm, payload, n, out);                                                    \
    else                                                                                    \
      fn<int64_t>(prm, payload, n, out);                                                    \
  }
ORQ_REF_BY_WIDTH(radix)
ORQ_REF_BY_WIDTH(quick)
ORQ_REF_BY_WIDTH(table_sort)
ORQ_REF_BY_WIDTH(join_groupby)
ORQ_REF_BY_WIDTH(table_distinct)
ORQ_REF_BY_WIDTH(table_aggregate)
ORQ_REF_BY_WIDTH(table_shuffle)
ORQ_REF_BY_WIDTH(circuits)
ORQ_REF_BY_WIDTH(radix_bench)
#undef ORQ_REF_BY_WIDTH

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s party fds in out\n", argv[0]);
    return 2;
  }
  const int party = std::atoi(argv[1]);
  // per worker thread t: the sockets to party+1 and party+2, "n0,p0,n1,p1,..."
  std::vector<int> fds;
  for (const char *f = argv[2]; *f;) {
    fds.push_back(std::atoi(f));
    while (*f && *f != ',') f++;
    if (*f == ',') f++;
  }
  std::ifstream in(argv[3], std::ios::binary);
  std::vector<uint8_t> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (buf.size() < 8) return 2;
  int64_t threads;
  std::memcpy(&threads, buf.data(), 8);
  if (threads < 1 || (size_t)threads * 2 != fds.size() ||
      buf.size() < 8 + 320 * (size_t)threads + 8 * 10)
    return 2;
  const uint8_t *keys = buf.data() + 8;  // thread t's key block at keys + 320 t
  int64_t hdr[10];
  std::memcpy(hdr, keys + 320 * threads, sizeof(hdr));
  const int64_t batch_size = hdr[0], scenario = hdr[1];
  const int64_t *prm = hdr + 2;
  const uint8_t *payload = keys + 320 * threads + sizeof(hdr);
  const size_t n = (size_t)prm[7];
  if (sodium_init() < 0) return 3;

  try {
    // orq_runtime_init (setup.h:8-57): every worker thread gets its own
    // communicator, randomness and protocol objects
    runTime = std::make_unique<RunTime>(batch_size, (int)threads, true);
    runTime->setup_workers(party);
    R3::ProtocolFactory protocolFactory(party, 3);
    auto groups = orq::ProtocolBase::generateRandomnessGroups(3, 2, 1);
    for (int t = 0; t < threads; t++) {
      auto &w = runTime->workers[t];
      w.attach_comm(std::make_unique<SocketComm>(party, fds[2 * t], fds[2 * t + 1]));
      w.attach_rand(make_randomness(party, keys + 320 * t, groups));
      w.init_proto<int8_t>(protocolFactory);
      w.init_proto<int16_t>(protocolFactory);
      w.init_proto<int32_t>(protocolFactory);
      w.init_proto<int64_t>(protocolFactory);
      w.init_proto<__int128_t>(protocolFactory);
    }

    Out out;
    switch (scenario) {
      case 1: radix_any(prm, payload, n, out); break;
      case 2: quick_any(prm, payload, n, out); break;
      case 3: table_sort_any(prm, payload, n, out); break;
      case 4: join_groupby_any(prm, payload, n, out); break;
      case 5: table_distinct_any(prm, payload, n, out); break;
      case 6: table_aggregate_any(prm, payload, n, out); break;
      case 7: table_shuffle_any(prm, payload, n, out); break;
      case 8: circuits_any(prm, payload, n, out); break;
      case 9: radix_bench_any(prm, payload, n, out); break;
      default:
        std::fprintf(stderr, "unknown scenario %ld\n", (long)scenario);
        return 2;
    }
    uint64_t sent = 0;
    for (int t = 0; t < threads; t++) {
      auto *comm = runTime->workers[t].getCommunicator();
      sent += comm->getBytesSent();
      // the last messages must reach the peers before this process ends
      static_cast<SocketComm *>(comm)->flush();
    }
    out.put_u64(sent);
    std::ofstream o(argv[4], std::ios::binary);
    o.write(reinterpret_cast<const char *>(out.bytes.data()), (std::streamsize)out.bytes.size());
    o.close();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "party %d: %s\n", party, e.what());
    std::fflush(stderr);
    _exit(1);
  }
  std::fflush(stdout);
  // the worker threads and communicator senders are not torn down one by one
  _exit(0);
}
