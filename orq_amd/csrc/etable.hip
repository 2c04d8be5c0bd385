// EncodedTable (include/core/containers/tabular/encoded_table.h) in the
// library: a table of named shared columns and the relational methods on this
// path -- sort, filter, project, distinct, aggregate, concatenate and the
// inner join of operators/join.h -- as control flow over the C ABI, so that a
// C or C++ caller reaches the join and the grouped aggregation without a host
// mirror of its own.
//
// As in the reference:
//   * the schema is ordered by column name (std::map, encoded_table.h:147);
//   * a column whose name contains '[' is boolean-shared, any other
//     arithmetic-shared (EncodedTable::isBShared, encoded_table.h:1672);
//   * every table carries the valid column "[##V]", a public sharing of ones
//     (configureValid, encoded_table.h:274-285);
//   * sort keys named "[##V]" or "[##TID]" are sorted on one bit
//     (encoded_table.h:915-935).
// Every step that draws randomness is made in the reference's order, so the
// result shares equal those of the same calls made one by one.
#include <algorithm>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "internal.h"

namespace {

const char *const kValid = "[##V]";
const char *const kJoinId = "[##TID]";
const char *const kUniq = "[##UNIQ]";
const char *const kValidTemp = "[##VALID_TEMP]";

bool is_bshared(const std::string &name) { return name.find('[') != std::string::npos; }

}  // namespace

struct orqg_et {
  orqg_ctx *ctx = nullptr;
  int width = 4;
  size_t rows = 0;
  std::map<std::string, orqg_vec *> cols;  // owned
  std::vector<std::string> names;          // column_name() storage, rebuilt on reads
  ~orqg_et() {
    for (auto &kv : cols) orqg_vec_free(kv.second);
  }
};

namespace orqg {
namespace {

void ck(orqg_ctx *c, int rc) {
  if (rc != ORQG_OK) throw Error(rc, orqg_last_error(c));
}

orqg_vec *new_vec(orqg_ctx *c, size_t n, int width, int enc) {
  orqg_vec *v = nullptr;
  ck(c, orqg_vec_create(c, n, width, enc, &v));
  return v;
}

// a zero column of the encoding its name gives
orqg_vec *zero_col(orqg_et *t, const std::string &name, size_t rows) {
  return new_vec(t->ctx, rows, t->width, is_bshared(name) ? ORQG_BSHARED : ORQG_ASHARED);
}

orqg_vec *ones(orqg_et *t, size_t rows) {
  orqg_vec *v = new_vec(t->ctx, rows, t->width, ORQG_BSHARED);
  ck(t->ctx, orqg_public_fill(t->ctx, v, 1, ORQG_BSHARED));
  return v;
}

void set_col(orqg_et *t, const std::string &name, orqg_vec *v) {
  auto it = t->cols.find(name);
  if (it != t->cols.end()) {
    if (it->second == v) return;
    orqg_vec_free(it->second);
    it->second = v;
  } else {
    t->cols[name] = v;
  }
}

orqg_vec *col(orqg_et *t, const std::string &name) {
  auto it = t->cols.find(name);
  if (it == t->cols.end()) throw Error(ORQG_E_ARG, "table: no column " + name);
  return it->second;
}

void et_add_columns(orqg_et *t, const std::vector<std::string> &names) {
  for (const std::string &c : names) set_col(t, c, zero_col(t, c, t->rows));
}

void et_delete_columns(orqg_et *t, const std::vector<std::string> &names) {
  for (const std::string &c : names) {
    auto it = t->cols.find(c);
    if (it == t->cols.end()) continue;
    orqg_vec_free(it->second);
    t->cols.erase(it);
  }
}

// EncodedTable::project: keep `keep` and the reserved columns (encoded_table.h:38)
void et_project(orqg_et *t, const std::set<std::string> &keep_in) {
  std::set<std::string> keep = keep_in;
  keep.insert(kJoinId);
  keep.insert(kValid);
  keep.insert(kUniq);
  std::vector<std::string> drop;
  for (auto &kv : t->cols)
    if (!keep.count(kv.first)) drop.push_back(kv.first);
  et_delete_columns(t, drop);
}

// filter (encoded_table.h:298-314): valid &= x
void et_filter(orqg_et *t, orqg_vec *x) {
  orqg_vec *v = col(t, kValid);
  orqg_vec *z = new_vec(t->ctx, t->rows, t->width, ORQG_BSHARED);
  ck(t->ctx, orqg_and_b(t->ctx, v, x, z));
  set_col(t, kValid, z);
}

void et_resize(orqg_et *t, size_t n) {
  for (auto &kv : t->cols) ck(t->ctx, orqg_vec_resize(t->ctx, kv.second, n));
  t->rows = n;
}

void et_tail(orqg_et *t, size_t n) {
  for (auto &kv : t->cols) ck(t->ctx, orqg_vec_tail(t->ctx, kv.second, n));
  t->rows = n;
}

void et_pad_power_of_two(orqg_et *t) {
  size_t p = 1;
  while (p < t->rows) p <<= 1;
  et_resize(t, p);
}

// EncodedTable::sort (encoded_table.h:900-998) on every column
void et_sort(orqg_et *t, const std::vector<std::string> &keys, const std::vector<int> &orders,
          int protocol) {
  if (protocol == ORQG_SORT_DEFAULT)
    // SortingProtocol::DEFAULT: the context's DEFAULT_SORT_PROTO (QUICKSORT
    // unless set, sorting.h:6-9)
    protocol = t->ctx->default_sort;
  if (protocol != ORQG_RADIXSORT && protocol != ORQG_QUICKSORT && protocol != ORQG_BITONICSORT)
    throw Error(ORQG_E_ARG, "table sort: unknown protocol");
  std::vector<orqg_vec *> kv, a, b;
  std::vector<int> single;
  for (const std::string &k : keys) {
    if (!is_bshared(k)) throw Error(ORQG_E_ARG, "table sort: key " + k + " is not boolean-shared");
    kv.push_back(col(t, k));
    single.push_back(k == kValid || k == kJoinId ? 1 : 0);
  }
  for (auto &c : t->cols) {
    if (std::find(keys.begin(), keys.end(), c.first) != keys.end()) continue;
    (is_bshared(c.first) ? b : a).push_back(c.second);
  }
  if (protocol == ORQG_BITONICSORT) {
    const size_t original = t->rows;
    et_pad_power_of_two(t);
    ck(t->ctx, orqg_bitonic_sort(t->ctx, kv.data(), kv.size(), orders.data(), a.data(), a.size(),
                                 b.data(), b.size()));
    auto vk = std::find(keys.begin(), keys.end(), std::string(kValid));
    if (vk != keys.end()) {  // padded rows are invalid: drop them again
      if (orders[vk - keys.begin()] == ORQG_ASC) et_tail(t, original);
      else et_resize(t, original);
    }
    return;
  }
  ck(t->ctx, orqg_table_sort_protocol(t->ctx, kv.data(), kv.size(), orders.data(), single.data(),
                                      a.data(), a.size(), b.data(), b.size(), protocol));
}

void et_sort_asc(orqg_et *t, const std::vector<std::string> &keys) {
  et_sort(t, keys, std::vector<int>(keys.size(), ORQG_ASC), ORQG_SORT_DEFAULT);
}

// distinct of `keys` into the column `res` (operators/distinct.h:19-44)
void et_distinct_into(orqg_et *t, const std::vector<std::string> &keys, const std::string &res) {
  std::vector<orqg_vec *> kv;
  for (const std::string &k : keys) kv.push_back(col(t, k));
  ck(t->ctx, orqg_distinct(t->ctx, kv.data(), kv.size(), col(t, res)));
}

// EncodedTable::distinct (encoded_table.h:1199-1237)
void et_distinct(orqg_et *t, const std::vector<std::string> &keys) {
  std::vector<std::string> sk = {kValid};
  sk.insert(sk.end(), keys.begin(), keys.end());
  et_sort_asc(t, sk);
  et_add_columns(t, {kUniq});
  et_distinct_into(t, keys, kUniq);
  et_filter(t, col(t, kUniq));
  et_delete_columns(t, {kUniq});
}

struct Agg {
  std::string data, result;
  int func;
};

// EncodedTable::aggregate (encoded_table.h:1059-1188)
void et_aggregate(orqg_et *t, std::vector<std::string> keys, const std::vector<Agg> &spec, bool reverse,
               bool do_sort, bool mark_valid, const std::string &table_id) {
  if (do_sort) {
    keys.insert(keys.begin(), kValid);
    et_sort_asc(t, keys);
  }
  const size_t original = t->rows;
  et_pad_power_of_two(t);
  std::vector<orqg_vec *> kv;
  for (const std::string &k : keys) kv.push_back(col(t, k));
  std::vector<orqg_agg_spec> bs, as;
  bool has_any = false;
  for (const Agg &s : spec) {
    if (is_bshared(s.data) != is_bshared(s.result))
      throw Error(ORQG_E_ARG, "aggregation " + s.data + " -> " + s.result + ": encodings differ");
    if (s.func < ORQG_AGG_SUM || s.func > ORQG_AGG_VALID)
      throw Error(ORQG_E_ARG, "aggregation: unknown function");
    // copy and valid are not aggregations (AggregationSelector::isAggregation)
    if (s.func != ORQG_AGG_COPY && s.func != ORQG_AGG_VALID) has_any = true;
    orqg_agg_spec a = {col(t, s.data), col(t, s.result), s.func};
    (is_bshared(s.data) ? bs : as).push_back(a);
  }
  const orqg_vec *sel = nullptr;
  if (mark_valid) {
    et_add_columns(t, {kUniq});
    et_distinct_into(t, keys, kUniq);
    if (!table_id.empty()) sel = col(t, table_id);
  }
  ck(t->ctx, orqg_aggregate(t->ctx, kv.data(), kv.size(), bs.data(), bs.size(), as.data(), as.size(),
                            reverse ? 1 : 0, sel));
  if (original < t->rows) et_resize(t, original);
  if (mark_valid) {
    if (has_any) {
      if (reverse) {
        // valid[0 .. n-1) &= uniq[1 .. n): the last row of each group stays
        const size_t n = t->rows;
        if (n > 1) {
          orqg_ctx *c = t->ctx;
          orqg_vec *v = zero_col(t, kValid, n - 1), *u = zero_col(t, kUniq, n - 1);
          orqg_vec *z = new_vec(c, n - 1, t->width, ORQG_BSHARED);
          try {
            ck(c, orqg_vec_copy_rows(c, col(t, kValid), 0, v, 0, n - 1));
            ck(c, orqg_vec_copy_rows(c, col(t, kUniq), 1, u, 0, n - 1));
            ck(c, orqg_and_b(c, v, u, z));
            ck(c, orqg_vec_copy_rows(c, z, 0, col(t, kValid), 0, n - 1));
          } catch (...) {
            for (orqg_vec *x : {v, u, z}) orqg_vec_free(x);
            throw;
          }
          for (orqg_vec *x : {v, u, z}) orqg_vec_free(x);
        }
      } else {
        et_filter(t, col(t, kUniq));
      }
    }
    et_delete_columns(t, {kUniq});
  }
}

// EncodedTable::concatenate (encoded_table.h:1548-1616): this table's rows,
// then the other's; the table id "[##TID]" is 0 above and a secret sharing of
// 1 repeated over the other table's rows
orqg_et *et_concatenate(orqg_et *a, orqg_et *b) {
  if (a->ctx != b->ctx || a->width != b->width)
    throw Error(ORQG_E_ARG, "concatenate: tables on different contexts or widths");
  orqg_et *t = new orqg_et;
  try {
    t->ctx = a->ctx;
    t->width = a->width;
    t->rows = a->rows + b->rows;
    std::set<std::string> names = {kJoinId};
    for (auto &kv : a->cols) names.insert(kv.first);
    for (auto &kv : b->cols) names.insert(kv.first);
    for (const std::string &c : names) set_col(t, c, zero_col(t, c, t->rows));
    set_col(t, kValid, ones(t, t->rows));  // configureValid, overwritten below
    orqg_ctx *c = t->ctx;
    for (auto &kv : a->cols) ck(c, orqg_vec_copy_rows(c, kv.second, 0, col(t, kv.first), 0, a->rows));
    for (auto &kv : b->cols)
      ck(c, orqg_vec_copy_rows(c, kv.second, 0, col(t, kv.first), a->rows, b->rows));
    orqg_vec *one = new_vec(c, 1, t->width, ORQG_BSHARED);
    try {
      const uint64_t plain = 1;
      const uint32_t plain32 = 1;
      ck(c, orqg_secret_share(c, t->width == 8 ? (const void *)&plain : (const void *)&plain32, one));
      ck(c, orqg_vec_broadcast_row(c, one, 0, col(t, kJoinId), a->rows, b->rows));
    } catch (...) {
      orqg_vec_free(one);
      throw;
    }
    orqg_vec_free(one);
  } catch (...) {
    delete t;
    throw;
  }
  return t;
}

// operators::inner_join (operators/join.h:31-233, _join with the inner-join
// valid rule): this (unique keys) JOIN right (foreign keys) ON keys
orqg_et *et_inner_join(orqg_et *left, orqg_et *right, const std::vector<std::string> &keys_in,
                    const std::vector<Agg> &spec_in, bool trim_invalid) {
  orqg_et *t = et_concatenate(left, right);
  try {
    std::set<std::string> out(keys_in.begin(), keys_in.end());
    for (auto &kv : right->cols) out.insert(kv.first);
    for (const Agg &s : spec_in) {
      out.insert(s.data);
      out.insert(s.result);
    }
    et_project(t, out);
    std::vector<std::string> keys = {kValid};
    keys.insert(keys.end(), keys_in.begin(), keys_in.end());
    std::vector<std::string> sk = keys;
    sk.push_back(kJoinId);
    et_sort_asc(t, sk);
    et_add_columns(t, {kUniq, kValidTemp});
    std::vector<Agg> spec = spec_in;
    spec.push_back({kValidTemp, kValidTemp, ORQG_AGG_VALID});
    et_distinct_into(t, keys, kUniq);
    // inner join: a row is valid if its key group is not a singleton
    orqg_ctx *c = t->ctx;
    orqg_vec *nu = new_vec(c, t->rows, t->width, ORQG_BSHARED);
    orqg_vec *vt = nullptr;
    try {
      ck(c, orqg_not_b(c, col(t, kUniq), nu));
      vt = new_vec(c, t->rows, t->width, ORQG_BSHARED);
      ck(c, orqg_and_b(c, col(t, kValid), nu, vt));
    } catch (...) {
      orqg_vec_free(nu);
      if (vt) orqg_vec_free(vt);
      throw;
    }
    orqg_vec_free(nu);
    set_col(t, kValidTemp, vt);
    et_delete_columns(t, {kUniq});
    et_aggregate(t, keys, spec, true, false, true, kJoinId);
    et_filter(t, col(t, kValidTemp));
    et_delete_columns(t, {kValidTemp, kJoinId});
    if (trim_invalid) {
      // drop rows that cannot be valid when the right table is small enough
      // that a sort on the valid bit pays (join.h)
      const size_t L = left->rows, R = right->rows;
      const size_t w = 8 * (size_t)left->width;
      const size_t cc = left->cols.size() - 1;
      if (R < L * (w + cc - 1) / (cc + 2)) {
        et_sort(t, {kValid}, {ORQG_DESC}, ORQG_SORT_DEFAULT);
        et_resize(t, R);
      }
    }
  } catch (...) {
    delete t;
    throw;
  }
  return t;
}

std::vector<std::string> names_of(const char *const *names, size_t n) {
  if (n && !names) throw Error(ORQG_E_ARG, "table: NULL column names");
  std::vector<std::string> out;
  for (size_t i = 0; i < n; i++) {
    if (!names[i]) throw Error(ORQG_E_ARG, "table: NULL column name");
    out.emplace_back(names[i]);
  }
  return out;
}

std::vector<Agg> specs_of(const orqg_et_agg *spec, size_t n) {
  if (n && !spec) throw Error(ORQG_E_ARG, "table: NULL aggregation spec");
  std::vector<Agg> out;
  for (size_t i = 0; i < n; i++) {
    if (!spec[i].data || !spec[i].result) throw Error(ORQG_E_ARG, "aggregation: NULL column name");
    out.push_back({spec[i].data, spec[i].result, spec[i].func});
  }
  return out;
}

}  // namespace
}  // namespace orqg

using orqg::Error;

#define ET_BEGIN(t)                                          \
  orqg_ctx *c_ = (t) ? (t)->ctx : nullptr;                   \
  if (!(t)) return ORQG_E_ARG;                               \
  try {
#define ET_END                                               \
  return ORQG_OK;                                            \
  }                                                          \
  catch (const Error &e) {                                   \
    if (c_) c_->err = e.what();                              \
    return e.code;                                           \
  }                                                          \
  catch (const std::exception &e) {                          \
    if (c_) c_->err = e.what();                              \
    return ORQG_E_DEVICE;                                    \
  }

extern "C" {

int orqg_et_create(orqg_ctx *ctx, int width, size_t rows, const char *const *names, size_t ncols,
                   orqg_et **out) {
  if (!ctx || !out || (width != 4 && width != 8)) return ORQG_E_ARG;
  orqg_et *t = new orqg_et;
  t->ctx = ctx;
  t->width = width;
  t->rows = rows;
  try {
    for (const std::string &c : orqg::names_of(names, ncols))
      orqg::set_col(t, c, orqg::zero_col(t, c, rows));
    orqg::set_col(t, kValid, orqg::ones(t, rows));
  } catch (const Error &e) {
    ctx->err = e.what();
    delete t;
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    delete t;
    return ORQG_E_DEVICE;
  }
  *out = t;
  return ORQG_OK;
}

int orqg_et_free(orqg_et *t) {
  if (!t) return ORQG_E_ARG;
  delete t;
  return ORQG_OK;
}

size_t orqg_et_rows(const orqg_et *t) { return t ? t->rows : 0; }
size_t orqg_et_ncols(const orqg_et *t) { return t ? t->cols.size() : 0; }

const char *orqg_et_column_name(orqg_et *t, size_t i) {
  if (!t || i >= t->cols.size()) return nullptr;
  t->names.clear();
  for (auto &kv : t->cols) t->names.push_back(kv.first);
  return t->names[i].c_str();
}

orqg_vec *orqg_et_column(orqg_et *t, const char *name) {
  if (!t || !name) return nullptr;
  auto it = t->cols.find(name);
  return it == t->cols.end() ? nullptr : it->second;
}

int orqg_et_set_column(orqg_et *t, const char *name, const orqg_vec *v) {
  ET_BEGIN(t)
  if (!name || !v || v->n != t->rows || v->width != t->width)
    throw Error(ORQG_E_ARG, "et_set_column: the column must have the table's rows and width");
  orqg_vec *x = orqg::zero_col(t, name, t->rows);
  const int rc = orqg_vec_copy(t->ctx, v, x);
  if (rc) {
    orqg_vec_free(x);
    throw Error(rc, orqg_last_error(t->ctx));
  }
  x->enc = is_bshared(name) ? ORQG_BSHARED : ORQG_ASHARED;
  orqg::set_col(t, name, x);
  ET_END
}

int orqg_et_add_columns(orqg_et *t, const char *const *names, size_t n) {
  ET_BEGIN(t)
  orqg::et_add_columns(t, orqg::names_of(names, n));
  ET_END
}

int orqg_et_delete_columns(orqg_et *t, const char *const *names, size_t n) {
  ET_BEGIN(t)
  orqg::et_delete_columns(t, orqg::names_of(names, n));
  ET_END
}

int orqg_et_project(orqg_et *t, const char *const *names, size_t n) {
  ET_BEGIN(t)
  const std::vector<std::string> v = orqg::names_of(names, n);
  orqg::et_project(t, std::set<std::string>(v.begin(), v.end()));
  ET_END
}

int orqg_et_filter(orqg_et *t, const char *name) {
  ET_BEGIN(t)
  if (!name) throw Error(ORQG_E_ARG, "et_filter: NULL column name");
  orqg::et_filter(t, orqg::col(t, name));
  ET_END
}

int orqg_et_resize(orqg_et *t, size_t n) {
  ET_BEGIN(t)
  orqg::et_resize(t, n);
  ET_END
}

int orqg_et_tail(orqg_et *t, size_t n) {
  ET_BEGIN(t)
  if (n > t->rows) throw Error(ORQG_E_ARG, "et_tail: more rows than the table has");
  orqg::et_tail(t, n);
  ET_END
}

int orqg_et_shuffle(orqg_et *t) {
  ET_BEGIN(t)
  std::vector<orqg_vec *> a, b;
  for (auto &kv : t->cols) (is_bshared(kv.first) ? b : a).push_back(kv.second);
  orqg::ck(t->ctx, orqg_table_shuffle(t->ctx, a.data(), a.size(), b.data(), b.size()));
  ET_END
}

int orqg_et_sort(orqg_et *t, const char *const *keys, const int *orders, size_t nk, int protocol) {
  ET_BEGIN(t)
  if (nk == 0) throw Error(ORQG_E_ARG, "et_sort: no sort keys");
  std::vector<int> o(nk, ORQG_ASC);
  if (orders)
    for (size_t i = 0; i < nk; i++) o[i] = orders[i] == ORQG_DESC ? ORQG_DESC : ORQG_ASC;
  orqg::et_sort(t, orqg::names_of(keys, nk), o, protocol);
  ET_END
}

int orqg_et_distinct(orqg_et *t, const char *const *keys, size_t nk) {
  ET_BEGIN(t)
  orqg::et_distinct(t, orqg::names_of(keys, nk));
  ET_END
}

int orqg_et_aggregate(orqg_et *t, const char *const *keys, size_t nk, const orqg_et_agg *spec,
                      size_t ns, int reverse, int do_sort, int mark_valid, const char *table_id) {
  ET_BEGIN(t)
  orqg::et_aggregate(t, orqg::names_of(keys, nk), orqg::specs_of(spec, ns), reverse != 0, do_sort != 0,
                  mark_valid != 0, table_id ? table_id : "");
  ET_END
}

int orqg_et_concatenate(orqg_et *a, orqg_et *b, orqg_et **out) {
  ET_BEGIN(a)
  if (!b || !out) throw Error(ORQG_E_ARG, "et_concatenate: NULL table");
  *out = orqg::et_concatenate(a, b);
  ET_END
}

int orqg_et_inner_join(orqg_et *left, orqg_et *right, const char *const *keys, size_t nk,
                       const orqg_et_agg *spec, size_t ns, int trim_invalid, orqg_et **out) {
  ET_BEGIN(left)
  if (!right || !out) throw Error(ORQG_E_ARG, "et_inner_join: NULL table");
  *out = orqg::et_inner_join(left, right, orqg::names_of(keys, nk), orqg::specs_of(spec, ns),
                          trim_invalid != 0);
  ET_END
}

}  // extern "C"
