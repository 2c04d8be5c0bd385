/*
 * orqg_adapter.hpp -- the C++ adapter an ORQ build includes to run its
 * 3-party operators on liborqg (include/orqg.h).
 *
 * It is generic over ORQ's share container: `EV` is any type with the shape
 * of orq::EVector<T, 2> (include/core/containers/e_vector.h:130-533) as a
 * party holds it -- `size()`, and `operator()(k)` giving the k-th replicated
 * share as an indexable vector of T.  Party p holds (s_p, s_{p+1}), so the
 * three distinct shares are the first share of each party.
 *
 * Errors come back from the C ABI as status codes; the adapter turns them
 * into orqg_adapter::Error carrying the code and orqg_last_error's text.
 */
#ifndef ORQG_ADAPTER_HPP_
#define ORQG_ADAPTER_HPP_

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "orqg.h"

namespace orqg_adapter {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// One context = the three parties' PRGs, permutation queue and message log
// (replicated_3pc::orq_init with explicit keys, include/backend/common/setup.h:7-87).
class Context {
 public:
  explicit Context(const uint8_t keys[ORQG_NUM_KEYS][32], int device = 0, long batch_size = -12,
                   size_t hbm_budget = 0) {
    orqg_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.device = device;
    std::memcpy(cfg.keys, keys, sizeof(cfg.keys));
    cfg.batch_size = batch_size;  // RunTime's default is -12 (runtime.h:35)
    cfg.hbm_budget = hbm_budget;
    const int rc = orqg_ctx_create(&cfg, &ctx_);
    if (rc != ORQG_OK) throw Error(rc, "orqg_ctx_create failed");
  }
  ~Context() {
    if (ctx_) orqg_ctx_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  orqg_ctx *get() const { return ctx_; }
  void check(int rc) const {
    if (rc != ORQG_OK) throw Error(rc, orqg_last_error(ctx_));
  }
  void sync() const { check(orqg_sync(ctx_)); }

 private:
  orqg_ctx *ctx_ = nullptr;
};

// A device-resident shared vector ([n][3] shares), freed with the object.
class Vec {
 public:
  Vec(Context &c, size_t n, int width, int encoding) : c_(&c) {
    c.check(orqg_vec_create(c.get(), n, width, encoding, &v_));
  }
  ~Vec() {
    if (v_) orqg_vec_free(v_);
  }
  Vec(const Vec &) = delete;
  Vec &operator=(const Vec &) = delete;
  Vec(Vec &&o) noexcept : c_(o.c_), v_(o.v_) { o.v_ = nullptr; }

  orqg_vec *get() const { return v_; }
  size_t size() const { return orqg_vec_size(v_); }

  // host [3][n] shares s_0, s_1, s_2
  template <typename T>
  void write(const std::vector<T> &host) {
    if (host.size() != 3 * size() || sizeof(T) != (size_t)orqg_vec_width(v_))
      throw Error(ORQG_E_ARG, "Vec::write: host buffer shape");
    c_->check(orqg_vec_write(c_->get(), v_, host.data()));
  }
  template <typename T>
  std::vector<T> read() const {
    std::vector<T> host(3 * size());
    if (sizeof(T) != (size_t)orqg_vec_width(v_)) throw Error(ORQG_E_ARG, "Vec::read: width");
    c_->check(orqg_vec_read(c_->get(), v_, host.data()));
    return host;
  }

 private:
  Context *c_;
  orqg_vec *v_ = nullptr;
};

// The three parties' EVectors -> one device vector.  parties[p] is party p's
// (s_p, s_{p+1}); s_k is taken from party k's first share.
template <typename T, typename EV>
Vec upload(Context &g, const std::array<EV, 3> &parties, int encoding) {
  const size_t n = parties[0].size();
  Vec v(g, n, sizeof(T), encoding);
  std::vector<T> host(3 * n);
  for (int k = 0; k < 3; k++) {
    const auto &s = parties[k](0);
    for (size_t i = 0; i < n; i++) host[k * n + i] = (T)s[i];
  }
  v.write(host);
  return v;
}

// The device vector -> the three parties' EVectors ((s_p, s_{p+1}) each).
template <typename T, typename EV>
void download(const Vec &v, std::array<EV, 3> &parties) {
  const std::vector<T> host = v.read<T>();
  const size_t n = v.size();
  for (int p = 0; p < 3; p++)
    for (int j = 0; j < 2; j++) {
      auto &s = parties[p](j);
      const size_t k = (size_t)((p + j) % 3);
      for (size_t i = 0; i < n; i++) s[i] = host[k * n + i];
    }
}

// operators::radix_sort (include/core/operators/radixsort.h:205-229) for the
// three parties at once: v sorted in place on its low `bits` bits, perm_out
// (width 4, A-shared) the sorting permutation.
inline void radix_sort(Context &g, Vec &v, bool descending, int bits, Vec &perm_out) {
  g.check(orqg_radix_sort(g.get(), v.get(), descending ? ORQG_DESC : ORQG_ASC, bits,
                          perm_out.get()));
}

// operators::table_sort with SortingProtocol::RADIXSORT (sorting.h:509-612).
inline void table_sort(Context &g, const std::vector<Vec *> &keys, const std::vector<int> &orders,
                       const std::vector<int> &single_bit, const std::vector<Vec *> &a_cols,
                       const std::vector<Vec *> &b_cols) {
  std::vector<orqg_vec *> k, a, b;
  for (Vec *x : keys) k.push_back(x->get());
  for (Vec *x : a_cols) a.push_back(x->get());
  for (Vec *x : b_cols) b.push_back(x->get());
  g.check(orqg_table_sort(g.get(), k.data(), k.size(), orders.data(),
                          single_bit.empty() ? nullptr : single_bit.data(), a.data(), a.size(),
                          b.data(), b.size()));
}

// EncodedTable (include/core/containers/tabular/encoded_table.h) in the
// library: named columns ('[' in a name = boolean-shared), the valid column
// "[##V]", and the relational methods, freed with the object.
class Table {
 public:
  Table(Context &g, int width, size_t rows, const std::vector<std::string> &names) : g_(&g) {
    std::vector<const char *> p;
    for (const std::string &x : names) p.push_back(x.c_str());
    g.check(orqg_et_create(g.get(), width, rows, p.data(), p.size(), &t_));
  }
  ~Table() {
    if (t_) orqg_et_free(t_);
  }
  Table(const Table &) = delete;
  Table &operator=(const Table &) = delete;
  Table(Table &&o) noexcept : g_(o.g_), t_(o.t_) { o.t_ = nullptr; }

  orqg_et *get() const { return t_; }
  size_t rows() const { return orqg_et_rows(t_); }
  // the column (lent: valid until the table changes)
  orqg_vec *column(const std::string &name) const {
    orqg_vec *v = orqg_et_column(t_, name.c_str());
    if (!v) throw Error(ORQG_E_ARG, "Table: no column " + name);
    return v;
  }
  // host [3][n] shares into / out of a column
  template <typename T>
  void write(const std::string &name, const std::vector<T> &host) {
    if (host.size() != 3 * rows()) throw Error(ORQG_E_ARG, "Table::write: host buffer shape");
    g_->check(orqg_vec_write(g_->get(), column(name), host.data()));
  }
  template <typename T>
  std::vector<T> read(const std::string &name) const {
    std::vector<T> host(3 * rows());
    g_->check(orqg_vec_read(g_->get(), column(name), host.data()));
    return host;
  }
  void add_columns(const std::vector<std::string> &names) {
    std::vector<const char *> p = c_names(names);
    g_->check(orqg_et_add_columns(t_, p.data(), p.size()));
  }
  void sort(const std::vector<std::string> &keys, const std::vector<int> &orders = {},
            int protocol = ORQG_SORT_DEFAULT) {
    std::vector<const char *> p = c_names(keys);
    g_->check(orqg_et_sort(t_, p.data(), orders.empty() ? nullptr : orders.data(), p.size(), protocol));
  }
  // aggregate with its defaults (sort on valid + keys, one valid row per group)
  void aggregate(const std::vector<std::string> &keys, const std::vector<orqg_et_agg> &spec,
                 bool reverse = false) {
    std::vector<const char *> p = c_names(keys);
    g_->check(orqg_et_aggregate(t_, p.data(), p.size(), spec.data(), spec.size(), reverse, 1, 1,
                                nullptr));
  }
  // operators::inner_join (operators/join.h:31-233): this (unique keys) JOIN right
  Table inner_join(Table &right, const std::vector<std::string> &keys,
                   const std::vector<orqg_et_agg> &spec, bool trim_invalid = true) {
    std::vector<const char *> p = c_names(keys);
    orqg_et *out = nullptr;
    g_->check(orqg_et_inner_join(t_, right.t_, p.data(), p.size(), spec.data(), spec.size(),
                                 trim_invalid, &out));
    return Table(*g_, out);
  }

 private:
  Table(Context &g, orqg_et *t) : g_(&g), t_(t) {}
  static std::vector<const char *> c_names(const std::vector<std::string> &names) {
    std::vector<const char *> p;
    for (const std::string &x : names) p.push_back(x.c_str());
    return p;
  }
  Context *g_;
  orqg_et *t_ = nullptr;
};

// open_shares_b / _a (replicated_3pc.h:310-330)
template <typename T>
std::vector<T> open(Context &g, const orqg_vec *v) {
  std::vector<T> out(orqg_vec_size(v));
  g.check(orqg_open(g.get(), v, out.data()));
  return out;
}
template <typename T>
std::vector<T> open(Context &g, const Vec &v) {
  return open<T>(g, v.get());
}

}  // namespace orqg_adapter

#endif  // ORQG_ADAPTER_HPP_
