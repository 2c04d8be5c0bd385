// Compiles include/orqg_adapter.hpp (the adapter INTEGRATION.md gives an ORQ
// build) against include/orqg.h, links liborqg.so and runs it.
//
//   adapter_test cpu   argument and error paths only (no device is touched)
//   adapter_test gpu   three parties' EVector-shaped shares -> radix sort of
//                      2^16 int32 keys -> back to per-party shares; checks the
//                      reconstructed order against std::stable_sort and the
//                      sorting permutation against its ranks; then a join
//                      R JOIN L ON key + GROUP BY key SUM through the
//                      adapter's Table, against the expected rows
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "orqg_adapter.hpp"

namespace {

// The shape of orq::EVector<T, 2> as one party holds it: two replicated shares.
template <typename T>
struct PartyShares {
  std::array<std::vector<T>, 2> s;
  explicit PartyShares(size_t n = 0) : s{std::vector<T>(n), std::vector<T>(n)} {}
  size_t size() const { return s[0].size(); }
  std::vector<T> &operator()(int k) { return s[k]; }
  const std::vector<T> &operator()(int k) const { return s[k]; }
};

uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int fail(const char *what) {
  std::fprintf(stderr, "adapter_test: %s\n", what);
  return 1;
}

int run_cpu() {
  // argument checks happen before any device call
  if (orqg_ctx_create(nullptr, nullptr) != ORQG_E_ARG) return fail("ctx_create(NULL)");
  if (std::strcmp(orqg_last_error(nullptr), "null context") != 0) return fail("last_error(NULL)");
  if (orqg_vec_size(nullptr) != 0 || orqg_vec_width(nullptr) != 0) return fail("vec accessors");
  if (orqg_vec_free(nullptr) != ORQG_E_ARG) return fail("vec_free(NULL)");
  if (orqg_server_create(nullptr, nullptr) != ORQG_E_ARG) return fail("server_create(NULL)");
  orqg_adapter::Error e(ORQG_E_PARITY, "x");
  if (e.code != ORQG_E_PARITY) return fail("Error code");
  std::printf("adapter_test cpu ok\n");
  return 0;
}

int run_gpu() {
  const size_t n = 1 << 16;
  uint8_t keys[ORQG_NUM_KEYS][32];
  uint64_t seed = 0x0A11CE01;
  for (auto &k : keys)
    for (int i = 0; i < 32; i += 8) {
      const uint64_t w = splitmix64(seed);
      std::memcpy(k + i, &w, 8);
    }
  // plaintext keys and a client's XOR sharing of them, laid out as the three
  // parties' (s_p, s_{p+1}) pairs
  std::vector<int32_t> plain(n);
  std::array<std::vector<uint32_t>, 3> s;
  for (auto &v : s) v.resize(n);
  uint64_t r = 7;
  for (size_t i = 0; i < n; i++) {
    plain[i] = (int32_t)(uint32_t)splitmix64(r);
    s[0][i] = (uint32_t)splitmix64(r);
    s[1][i] = (uint32_t)splitmix64(r);
    s[2][i] = (uint32_t)plain[i] ^ s[0][i] ^ s[1][i];
  }
  std::array<PartyShares<uint32_t>, 3> parties{PartyShares<uint32_t>(n), PartyShares<uint32_t>(n),
                                               PartyShares<uint32_t>(n)};
  for (int p = 0; p < 3; p++) {
    parties[p](0) = s[p];
    parties[p](1) = s[(p + 1) % 3];
  }
  try {
    orqg_adapter::Context g(keys);
    orqg_adapter::Vec v = orqg_adapter::upload<uint32_t>(g, parties, ORQG_BSHARED);
    orqg_adapter::Vec perm(g, n, 4, ORQG_ASHARED);
    orqg_adapter::radix_sort(g, v, false, 32, perm);
    std::array<PartyShares<uint32_t>, 3> out{PartyShares<uint32_t>(n), PartyShares<uint32_t>(n),
                                             PartyShares<uint32_t>(n)};
    orqg_adapter::download<uint32_t>(v, out);
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return plain[a] < plain[b]; });
    for (size_t i = 0; i < n; i++) {
      // every party's pair is consistent with its neighbours'
      for (int p = 0; p < 3; p++)
        if (out[p](1)[i] != out[(p + 1) % 3](0)[i]) return fail("replicated shares disagree");
      const int32_t got = (int32_t)(out[0](0)[i] ^ out[1](0)[i] ^ out[2](0)[i]);
      if (got != plain[order[i]]) return fail("sorted keys differ from std::stable_sort");
    }
    const std::vector<uint32_t> pv = orqg_adapter::open<uint32_t>(g, perm);
    for (size_t i = 0; i < n; i++)
      if (pv[order[i]] != i) return fail("sorting permutation differs from the ranks");
    // an error comes back as an exception with the library's message
    bool threw = false;
    try {
      orqg_adapter::Vec bad(g, n, 3, ORQG_BSHARED);
    } catch (const orqg_adapter::Error &err) {
      threw = err.code == ORQG_E_ARG;
    }
    if (!threw) return fail("width 3 not refused");
    // R(key, [attr]) JOIN L(key, value) ON key copying [attr] down, then
    // GROUP BY key SUM(value), through the library's EncodedTable
    {
      const size_t nr = 4, nl = 9;
      const uint32_t rkey[nr] = {1, 2, 3, 5}, attr[nr] = {10, 20, 30, 50};
      const uint32_t lkey[nl] = {2, 1, 2, 7, 5, 2, 1, 3, 9}, val[nl] = {1, 2, 3, 4, 5, 6, 7, 8, 9};
      auto bshare = [](const uint32_t *x, size_t m) {  // s_0 = x, s_1 = s_2 = 0
        std::vector<uint32_t> h(3 * m, 0);
        std::copy(x, x + m, h.begin());
        return h;
      };
      orqg_adapter::Table R(g, 4, nr, {"[key]", "[attr]"});
      R.write("[key]", bshare(rkey, nr));
      R.write("[attr]", bshare(attr, nr));
      orqg_adapter::Table L(g, 4, nl, {"[key]", "value"});
      L.write("[key]", bshare(lkey, nl));
      L.write("value", bshare(val, nl));
      orqg_adapter::Table J = R.inner_join(L, {"[key]"}, {{"[attr]", "[attr]", ORQG_AGG_COPY}});
      J.aggregate({"[key]"}, {{"value", "value", ORQG_AGG_SUM}});
      const std::vector<uint32_t> k = orqg_adapter::open<uint32_t>(g, J.column("[key]"));
      const std::vector<uint32_t> a = orqg_adapter::open<uint32_t>(g, J.column("[attr]"));
      const std::vector<uint32_t> v = orqg_adapter::open<uint32_t>(g, J.column("value"));
      const std::vector<uint32_t> ok = orqg_adapter::open<uint32_t>(g, J.column("[##V]"));
      std::vector<std::array<uint32_t, 3>> got;
      for (size_t i = 0; i < J.rows(); i++)
        if (ok[i]) got.push_back({k[i], a[i], v[i]});
      std::sort(got.begin(), got.end());
      const std::vector<std::array<uint32_t, 3>> want = {
          {1, 10, 9}, {2, 20, 10}, {3, 30, 8}, {5, 50, 5}};
      if (got != want) return fail("join + group-by through the adapter's Table differs");
    }
  } catch (const std::exception &ex) {
    std::fprintf(stderr, "adapter_test: %s\n", ex.what());
    return 1;
  }
  std::printf("adapter_test gpu ok: 65536-row radix sort and a join + group-by through the C++ "
              "adapter\n");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  return mode == "gpu" ? run_gpu() : run_cpu();
}
