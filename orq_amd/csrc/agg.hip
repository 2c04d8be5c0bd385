// The sorting-network aggregation (operators/aggregation.h:27-346) and
// distinct (operators/distinct.h:19-44).
//
// Level l of the aggregation pairs row j with row j + d for j < m = n - d.
// The reference updates the accumulator rows from the values all rows had
// before the level (its right-hand sides are materialised before the
// assignment), so a level here reads the column, writes the m new values to
// a temporary and copies them back over the accumulator rows.  Every runtime
// call of the reference (the equality circuits, the and_b / multiply_a of
// the aggregation functions, b2a_bit of the group bits) draws its keystream
// at its own nonce offset, so shares, nonces and messages match it.
#include <map>
#include <mutex>
#include <string>

#include "internal.h"
#include "kern_common.h"
#include "cmp_core.h"

namespace orqg {
namespace {

enum AggFn { kSum = 0, kCount = 1, kCopy = 2, kMax = 3, kMin = 4, kOr = 5, kValid = 6 };

struct PairArgs {
  uint64_t m;            // pairs (rows j < m paired with j + d)
  CallSched s1, sp, s3;  // and_b over m, the packed and_b of _compare, b2a_bit
  uint64_t N1, Np, N3;
};

template <typename T, int E>
__device__ inline void load_run(const T *__restrict__ col, uint64_t row0, uint64_t m, uint64_t q0,
                                T (&v)[E][3]) {
  // rows past the run read its last row (callers run only with m >= 1) and
  // are zeroed: no branch around a load, so the 3E loads are in flight together
#pragma unroll
  for (int i = 0; i < E; i++) {
    const bool in = q0 + i < m;
    const uint64_t q = in ? q0 + i : m - 1;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      const T x = col[3 * (row0 + q) + p];
      v[i][p] = in ? x : (T)0;
    }
  }
}
template <typename T, int E>
__device__ inline void store_run(T *__restrict__ col, uint64_t m, uint64_t q0, const T (&v)[E][3]) {
#pragma unroll
  for (int i = 0; i < E; i++)
    if (q0 + i < m)
#pragma unroll
      for (int p = 0; p < 3; p++) col[3 * (q0 + i) + p] = v[i][p];
}

template <typename T>
__device__ inline void mul3(const T (&x)[3], const T (&y)[3], const T (&S)[3], T (&z)[3]) {
#pragma unroll
  for (int p = 0; p < 3; p++) {
    int q = (p + 1) % 3, pm = (p + 2) % 3;
    z[p] = (T)(S[pm] - S[p]) + x[p] * y[p] + x[p] * y[q] + x[q] * y[p];
  }
}

// a ^ (extend_lsb(s) & (b ^ a)) with one and_b keystream (operators/common.h:20-31)
template <typename T>
__device__ inline void mux3(const T (&s)[3], const T (&a)[3], const T (&b)[3], const T (&ks)[3],
                            T (&r)[3]) {
  T se[3], d[3], z[3];
#pragma unroll
  for (int p = 0; p < 3; p++) {
    se[p] = (T)0 - (s[p] & 1);
    d[p] = b[p] ^ a[p];
  }
  and3<T>(se, d, ks, z);
#pragma unroll
  for (int p = 0; p < 3; p++) r[p] = a[p] ^ z[p];
}

// x._compare(y) / x == y over pairs: x = xcol[xo + j], y = ycol[yo + j];
// out = eq (CMP false) or gt (CMP true), with not_b_1 when neg.
template <typename T, bool CMP>
__global__ void __launch_bounds__(kAesThreads)
    k_pair_cmp(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
               NonceTriple nb, uint64_t noff, PairArgs a, const T *__restrict__ xcol, uint64_t xo,
               const T *__restrict__ ycol, uint64_t yo, T *__restrict__ out, int neg) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const int lane = threadIdx.x & 63;
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t gw = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); gw < ngrp;
       gw += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = (gw + lane) * E;
    T X[E][3], Y[E][3], EQ[E][3], GT[E][3];
    load_run<T, E>(xcol, xo, a.m, q0, X);
    load_run<T, E>(ycol, yo, a.m, q0, Y);
    compare_group<T, CMP>(A, rk_pair, nb, noff, a.N1, a.s1, a.sp, gw, lane, X, Y, EQ, GT);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int p = 0; p < 3; p++) {
        T v = CMP ? GT[i][p] : EQ[i][p];
        EQ[i][p] = neg ? (T)((v & 1) ^ 1) : v;
      }
    store_run<T, E>(out, a.m, q0, EQ);
  }
}

// g = g & x, one and_b at nonce offset noff; with SEL, x = ~(sel[d + j] ^ sel[j]).
template <typename T, bool SEL>
__global__ void __launch_bounds__(kAesThreads)
    k_pair_and(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
               NonceTriple nb, uint64_t noff, PairArgs a, T *__restrict__ g,
               const T *__restrict__ x, uint64_t d) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t grp = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; grp < ngrp;
       grp += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = grp * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(a.s1, q0, 0, off, ctr, bib);
    T G[E][3], X[E][3];
    load_run<T, E>(g, 0, a.m, q0, G);
    if (SEL) {
      T X2[E][3];
      load_run<T, E>(x, d, a.m, q0, X);
      load_run<T, E>(x, 0, a.m, q0, X2);
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int p = 0; p < 3; p++) X[i][p] = ~(X[i][p] ^ X2[i][p]);
    } else {
      load_run<T, E>(x, 0, a.m, q0, X);
    }
    T S[3][E];
    pair_stream<T, true>(A, rk_pair, nb, noff, off, ctr, S);
#pragma unroll
    for (int i = 0; i < E; i++) {
      T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
      and3<T>(G[i], X[i], ks, z);
#pragma unroll
      for (int p = 0; p < 3; p++) G[i][p] = z[p];
    }
    store_run<T, E>(g, a.m, q0, G);
  }
}

// b2a_bit of the LSBs of `in` (m values) at nonce offset noff
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_pair_b2a(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
               NonceTriple nb, uint64_t noff, PairArgs a, const T *__restrict__ in,
               T *__restrict__ out) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  const NonceTriple nbb = {{nb.v[0] + noff, nb.v[1] + noff, nb.v[2] + noff}};
  for (uint64_t grp = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; grp < ngrp;
       grp += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = grp * E;
    T X[E][3], Y[E][3];
    load_run<T, E>(in, 0, a.m, q0, X);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int p = 0; p < 3; p++) X[i][p] &= 1;
    b2a_bit_group<T>(A, rk_pair, nbb, a.s3, q0, X, Y);
    store_run<T, E>(out, a.m, q0, Y);
  }
}

// One boolean or arithmetic aggregation function on the pairs: acc = col[ao + j],
// in = col[io + j]; the new accumulator values go to out[j].  g is the group
// bit vector the function uses (join group bits for copy), bg = (in > acc)
// for max / min.  Calls start at nonce offset noff.
template <typename T, int FN>
__global__ void __launch_bounds__(kAesThreads)
    k_pair_agg(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
               NonceTriple nb, uint64_t noff, PairArgs a, const T *__restrict__ col, uint64_t ao,
               uint64_t io, const T *__restrict__ g, const T *__restrict__ bg,
               T *__restrict__ out, int arith) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t grp = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; grp < ngrp;
       grp += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = grp * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(a.s1, q0, 0, off, ctr, bib);
    T AC[E][3], IN[E][3], G[E][3], R[E][3];
    load_run<T, E>(col, ao, a.m, q0, AC);
    load_run<T, E>(col, io, a.m, q0, IN);
    load_run<T, E>(g, 0, a.m, q0, G);
    T S[3][E];
    pair_stream<T, true>(A, rk_pair, nb, noff, off, ctr, S);
    if (arith) {
      // sum / count: acc + g * in;  copy: acc + g * (in - acc)
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]}, d[3], z[3];
#pragma unroll
        for (int p = 0; p < 3; p++) d[p] = FN == kCopy ? (T)(IN[i][p] - AC[i][p]) : IN[i][p];
        mul3<T>(G[i], d, ks, z);
#pragma unroll
        for (int p = 0; p < 3; p++) R[i][p] = AC[i][p] + z[p];
      }
    } else if (FN == kCopy) {
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]};
        mux3<T>(G[i], AC[i], IN[i], ks, R[i]);
      }
    } else {
      // first call
      T F[E][3];
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]};
        if (FN == kOr) {  // t = extend_lsb(g) & in
          T se[3];
#pragma unroll
          for (int p = 0; p < 3; p++) se[p] = (T)0 - (G[i][p] & 1);
          and3<T>(se, IN[i], ks, F[i]);
        } else if (FN == kValid) {  // agg = acc & in
          and3<T>(AC[i], IN[i], ks, F[i]);
        } else {  // max / min: inner = mux(bg, acc, in)
          T B2[3];
#pragma unroll
          for (int p = 0; p < 3; p++) {
            const T x = bg[3 * (q0 + i < a.m ? q0 + i : a.m - 1) + p];
            B2[p] = q0 + i < a.m ? x : (T)0;
          }
          mux3<T>(B2, AC[i], IN[i], ks, F[i]);
        }
      }
      // second call
      pair_stream<T, true>(A, rk_pair, nb, noff + a.N1, off, ctr, S);
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]};
        if (FN == kOr) {  // acc | t = ~(~acc & ~t)
          T na[3], nt[3], z[3];
#pragma unroll
          for (int p = 0; p < 3; p++) {
            na[p] = ~AC[i][p];
            nt[p] = ~F[i][p];
          }
          and3<T>(na, nt, ks, z);
#pragma unroll
          for (int p = 0; p < 3; p++) R[i][p] = ~z[p];
        } else {  // valid / max / min: mux(g, acc, f)
          mux3<T>(G[i], AC[i], F[i], ks, R[i]);
        }
      }
    }
    store_run<T, E>(out, a.m, q0, R);
  }
}

// distinct: rest[j] |= ne[j] for rest = res[1 + j], as ~(~rest & ~ne)
// (b_shared_vector.h:452-456), one and_b at nonce offset noff.
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_distinct_or(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
                  NonceTriple nb, uint64_t noff, PairArgs a, T *__restrict__ res,
                  const T *__restrict__ ne) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t grp = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; grp < ngrp;
       grp += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = grp * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(a.s1, q0, 0, off, ctr, bib);
    T R[E][3], X[E][3];
    load_run<T, E>(res, 1, a.m, q0, R);
    load_run<T, E>(ne, 0, a.m, q0, X);
    T S[3][E];
    pair_stream<T, true>(A, rk_pair, nb, noff, off, ctr, S);
#pragma unroll
    for (int i = 0; i < E; i++) {
      T ks[3] = {S[0][i], S[1][i], S[2][i]}, nr[3], nx[3], z[3];
#pragma unroll
      for (int p = 0; p < 3; p++) {
        nr[p] = ~R[i][p];
        nx[p] = ~X[i][p];
      }
      and3<T>(nr, nx, ks, z);
#pragma unroll
      for (int p = 0; p < 3; p++) R[i][p] = ~z[p];
    }
#pragma unroll
    for (int i = 0; i < E; i++)
      if (q0 + i < a.m)
#pragma unroll
        for (int p = 0; p < 3; p++) res[3 * (1 + q0 + i) + p] = R[i][p];
  }
}

// ---- host side -----------------------------------------------------------------------
struct Launcher {
  orqg_ctx *c;
  int w;
  PairArgs a;
  unsigned grid;
  NonceTriple nb;
  uint64_t noff = 0;  // offset inside the current runtime-call sequence

  Launcher(orqg_ctx *c_, int w_, uint64_t m) : c(c_), w(w_) {
    const int B = 8 * w;
    a.m = m;
    a.s1 = make_sched(m, c->batch_size, w, 1);
    a.sp = make_sched(m / B + (m % B > 0), c->batch_size, w, 1);
    a.s3 = make_sched(m, c->batch_size, w, 3);
    a.N1 = sched_nonces(a.s1);
    a.Np = sched_nonces(a.sp);
    a.N3 = sched_nonces(a.s3);
    grid = grid_for((m * w + 15) / 16, kAesThreads, c->num_cu);
    nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
  }
  // advance the pair keys and log the messages of calls issued so far
  void and_calls(uint64_t k) { account("and_b", k * a.N1, k * a.m * w); }
  void mul_calls(uint64_t k) { account("mul_a", k * a.N1, k * a.m * w); }
  void b2a_call() { account("b2a_bit", a.N3, 3 * a.m * w); }
  void packed_call() {
    const uint64_t nw = a.sp.n;
    account("and_b", a.Np, nw * w);
  }
  void account(const char *op, uint64_t nonces, uint64_t bytes) {
    noff += nonces;
    for (int p = 0; p < 3; p++) c->log_msg(op, p, (p + 2) % 3, bytes);
  }
  void finish() {
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += noff;
  }

  template <typename T>
  void cmp(bool full, const void *x, uint64_t xo, const void *y, uint64_t yo, void *out, int neg) {
    if (full)
      hipLaunchKernelGGL((k_pair_cmp<T, true>), dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, noff, a, (const T *)x, xo, (const T *)y, yo,
                         (T *)out, neg);
    else
      hipLaunchKernelGGL((k_pair_cmp<T, false>), dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, noff, a, (const T *)x, xo, (const T *)y, yo,
                         (T *)out, neg);
    ORQG_HIP(hipGetLastError());
    const int L = sizeof(T) == 4 ? 5 : 6;
    and_calls(L);
    if (full) {
      and_calls(1);
      packed_call();
    }
  }
  template <typename T>
  void and_(void *g, const void *x, bool sel, uint64_t d) {
    if (sel)
      hipLaunchKernelGGL((k_pair_and<T, true>), dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, noff, a, (T *)g, (const T *)x, d);
    else
      hipLaunchKernelGGL((k_pair_and<T, false>), dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, noff, a, (T *)g, (const T *)x, d);
    ORQG_HIP(hipGetLastError());
    and_calls(1);
  }
  template <typename T>
  void b2a(const void *in, void *out) {
    hipLaunchKernelGGL(k_pair_b2a<T>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                       c->rk(kPair0), nb, noff, a, (const T *)in, (T *)out);
    ORQG_HIP(hipGetLastError());
    b2a_call();
  }
  template <typename T, int FN>
  void agg_launch(const void *col, uint64_t ao, uint64_t io, const void *g, const void *bg,
                  void *out, int arith) {
    hipLaunchKernelGGL((k_pair_agg<T, FN>), dim3(grid), dim3(kAesThreads), 0, c->stream,
                       c->d_t0, c->rk(kPair0), nb, noff, a, (const T *)col, ao, io, (const T *)g,
                       (const T *)bg, (T *)out, arith);
    ORQG_HIP(hipGetLastError());
  }
  template <typename T>
  void agg(int fn, bool arith, const void *col, uint64_t ao, uint64_t io, const void *g,
           const void *bg, void *out) {
    switch (fn) {
      case kSum:
      case kCount: agg_launch<T, kSum>(col, ao, io, g, bg, out, 1); break;
      case kCopy: agg_launch<T, kCopy>(col, ao, io, g, bg, out, arith ? 1 : 0); break;
      case kOr: agg_launch<T, kOr>(col, ao, io, g, bg, out, 0); break;
      case kValid: agg_launch<T, kValid>(col, ao, io, g, bg, out, 0); break;
      default: agg_launch<T, kMax>(col, ao, io, g, bg, out, 0); break;  // max / min
    }
    if (arith)
      mul_calls(1);
    else
      and_calls(fn == kCopy ? 1 : 2);
  }
  template <typename T>
  void distinct_or(void *res, const void *ne) {
    hipLaunchKernelGGL(k_distinct_or<T>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                       c->rk(kPair0), nb, noff, a, (T *)res, (const T *)ne);
    ORQG_HIP(hipGetLastError());
    and_calls(1);
  }
};

template <typename T>
void aggregate_t(orqg_ctx *c, orqg_vec *const *keys, size_t nk, const orqg_agg_spec *bs,
                 size_t nbs, const orqg_agg_spec *as, size_t nas, bool reverse,
                 const orqg_vec *sel, uint64_t n) {
  constexpr int w = sizeof(T);
  bool a_copy = false, a_noncopy = false, b_copy = false, b_noncopy = false;
  // preprocessing (aggregation.h:206-246)
  for (size_t i = 0; i < nas; i++) {
    if (as[i].func == kCount) {
      if (sel)
        b2a_bit(c, sel, as[i].out);
      else
        public_fill(c, as[i].out, 1);
    } else if (as[i].out != as[i].in) {
      ORQG_HIP(hipMemcpyAsync(as[i].out->d, as[i].in->d, 3 * n * w, hipMemcpyDeviceToDevice,
                              c->stream));
    }
    as[i].out->enc = ORQG_ASHARED;
    (as[i].func == kCopy ? a_copy : a_noncopy) = true;
  }
  for (size_t i = 0; i < nbs; i++) {
    if (bs[i].out != bs[i].in)
      ORQG_HIP(hipMemcpyAsync(bs[i].out->d, bs[i].in->d, 3 * n * w, hipMemcpyDeviceToDevice,
                              c->stream));
    bs[i].out->enc = ORQG_BSHARED;
    (bs[i].func == kCopy ? b_copy : b_noncopy) = true;
  }
  const size_t vb = 3 * n * w + 64;
  T *g = (T *)c->scratch_get("agg_g", vb), *jg = (T *)c->scratch_get("agg_jg", vb);
  T *e = (T *)c->scratch_get("agg_e", vb), *bgb = (T *)c->scratch_get("agg_bg", vb);
  T *ga = (T *)c->scratch_get("agg_ga", vb), *jga = (T *)c->scratch_get("agg_jga", vb);
  T *tmp = (T *)c->scratch_get("agg_tmp", vb);
  int log_size = 0;
  while ((2ull << log_size) <= n) log_size++;
  TimedRegion tr(c, "aggregate");
  for (int lv = 1; lv <= log_size; lv++) {
    uint64_t d = n >> lv;
    if (reverse) d = n >> (log_size - lv + 1);
    const uint64_t m = n - d;
    const size_t mb = 3 * m * w;
    Launcher L(c, w, m);
    // group bits (aggregation.h:269-291)
    if (nk == 0) {
      public_fill_raw(c, g, m, 1, w);
    } else {
      L.cmp<T>(false, keys[0]->d, 0, keys[0]->d, d, g, 0);
      for (size_t k = 1; k < nk; k++) {
        L.cmp<T>(false, keys[k]->d, 0, keys[k]->d, d, e, 0);
        L.and_<T>(g, e, false, 0);
      }
    }
    ORQG_HIP(hipMemcpyAsync(jg, g, mb, hipMemcpyDeviceToDevice, c->stream));
    if (sel && (a_noncopy || b_noncopy)) L.and_<T>(g, sel->d, true, d);
    // group_bits_b.mask(1) happens inside the consumers: every use reads the LSB
    const uint64_t ao = reverse ? d : 0, io = reverse ? 0 : d;
    for (size_t i = 0; i < nbs; i++) {
      const int fn = bs[i].func;
      if (fn == kMax || fn == kMin)
        L.cmp<T>(true, bs[i].out->d, io, bs[i].out->d, ao, bgb, fn == kMin ? 1 : 0);
      L.agg<T>(fn, false, bs[i].out->d, ao, io, fn == kCopy ? (const void *)jg : (const void *)g,
               bgb, tmp);
      ORQG_HIP(hipMemcpyAsync((char *)bs[i].out->d + 3 * ao * w, tmp, mb,
                              hipMemcpyDeviceToDevice, c->stream));
    }
    if (nas > 0) {
      if (nk > 0) {
        if (a_noncopy) L.b2a<T>(g, ga);
        if (a_copy) L.b2a<T>(jg, jga);
      } else {
        public_fill_raw(c, ga, m, 1, w);
        public_fill_raw(c, jga, m, 1, w);
      }
    }
    for (size_t i = 0; i < nas; i++) {
      const int fn = as[i].func;
      L.agg<T>(fn, true, as[i].out->d, ao, io, fn == kCopy ? (const void *)jga : (const void *)ga,
               nullptr, tmp);
      ORQG_HIP(hipMemcpyAsync((char *)as[i].out->d + 3 * ao * w, tmp, mb,
                              hipMemcpyDeviceToDevice, c->stream));
    }
    L.finish();
  }
}

// ne_buf: the comparison results (3 n w bytes); NULL takes the context's
// scratch.  The caller-supplied form zeroes res with a kernel instead of a
// memset node, so that it can sit in a captured graph.
template <typename T>
void distinct_t(orqg_ctx *c, orqg_vec *const *keys, size_t nk, orqg_vec *res, T *ne_buf) {
  constexpr int w = sizeof(T);
  const uint64_t n = res->n;
  if (ne_buf)
    public_fill_raw(c, res->d, n, 0, w);  // res->zero()
  else
    ORQG_HIP(hipMemsetAsync(res->d, 0, 3 * n * w, c->stream));
  res->enc = ORQG_BSHARED;
  if (n == 0) return;
  public_fill_raw(c, res->d, 1, 1, w);  // res[0] = public 1
  if (n == 1) return;
  // every row of ne that distinct_or reads (0 .. n-2) is written by the
  // comparison first, so the buffer needs no clearing
  T *ne = ne_buf ? ne_buf : (T *)c->scratch_get("distinct_ne", 3 * n * w + 64);
  TimedRegion tr(c, "distinct");
  for (size_t k = 0; k < nk; k++) {
    Launcher L(c, w, n - 1);
    L.cmp<T>(false, keys[k]->d, 0, keys[k]->d, 1, ne, 1);  // a != b = !(a == b)
    L.distinct_or<T>(res->d, ne);
    L.finish();
  }
}

}  // namespace

void aggregate(orqg_ctx *c, orqg_vec *const *keys, size_t nk, const orqg_agg_spec *bs, size_t nbs,
               const orqg_agg_spec *as, size_t nas, int reverse, const orqg_vec *sel) {
  require_one_thread(c, "aggregate");
  const orqg_vec *first = nk ? keys[0] : nbs ? bs[0].in : nas ? as[0].in : nullptr;
  if (!first) throw Error(ORQG_E_ARG, "Empty aggregation!");
  const uint64_t n = first->n;
  const int w = first->width;
  if (n & (n - 1)) throw Error(ORQG_E_ARG, "aggregate: size must be a power of two");
  auto chk = [&](const orqg_vec *v) {
    if (!v || v->n != n || v->width != w)
      throw Error(ORQG_E_ARG, "aggregate: columns differ in size or width");
  };
  for (size_t k = 0; k < nk; k++) chk(keys[k]);
  for (size_t i = 0; i < nbs; i++) {
    chk(bs[i].in);
    chk(bs[i].out);
    if (bs[i].func == kSum || bs[i].func == kCount || bs[i].func > kValid)
      throw Error(ORQG_E_ARG, "aggregate: not a boolean aggregation");
  }
  for (size_t i = 0; i < nas; i++) {
    chk(as[i].in);
    chk(as[i].out);
    if (as[i].func != kSum && as[i].func != kCount && as[i].func != kCopy)
      throw Error(ORQG_E_ARG, "aggregate: not an arithmetic aggregation");
  }
  if (sel) chk(sel);
  if (n < 2) return;
  if (w == 4)
    aggregate_t<uint32_t>(c, keys, nk, bs, nbs, as, nas, reverse != 0, sel, n);
  else
    aggregate_t<uint64_t>(c, keys, nk, bs, nbs, as, nas, reverse != 0, sel, n);
}

void distinct(orqg_ctx *c, orqg_vec *const *keys, size_t nk, orqg_vec *res) {
  require_one_thread(c, "distinct");
  for (size_t k = 0; k < nk; k++)
    if (!keys[k] || keys[k]->n != res->n || keys[k]->width != res->width)
      throw Error(ORQG_E_ARG, "distinct: columns differ in size or width");
  if (res->width == 4)
    distinct_t<uint32_t>(c, keys, nk, res, nullptr);
  else
    distinct_t<uint64_t>(c, keys, nk, res, nullptr);
}

}  // namespace orqg

// ---- distinct without allocation ---------------------------------------------
// The comparison buffer of orqg_distinct_noalloc lives in this static device
// array: every GPU the library runs on holds one copy of it until the process
// ends, and every call on that GPU shares it (calls that overlap on one GPU
// race on it).  In host code `scratch` names a host-side stand-in, not device
// memory: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device at context creation.
static __device__ __align__(256) uint8_t scratch[orqg::kNoAllocScratchBytes];

namespace orqg {
namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device ordinal -> its copy of `scratch`
}  // namespace

void noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  void *p = nullptr;
  ORQG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
}

void distinct_noalloc(orqg_ctx *c, orqg_vec *const *keys, size_t nk, orqg_vec *res) {
  require_one_thread(c, "distinct_noalloc");
  for (size_t k = 0; k < nk; k++)
    if (!keys[k] || keys[k]->n != res->n || keys[k]->width != res->width)
      throw Error(ORQG_E_ARG, "distinct: columns differ in size or width");
  const size_t need = 3 * res->n * (size_t)res->width;
  if (need > kNoAllocScratchBytes)
    throw Error(ORQG_E_ARG, "distinct_noalloc: " + std::to_string(res->n) + " rows of width " +
                                std::to_string(res->width) + " need " + std::to_string(need) +
                                " bytes of scratch, over its fixed " +
                                std::to_string(kNoAllocScratchBytes) +
                                " (at most 2^21 rows of width 8, 2^22 of width 4)");
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(c->device);
    if (it != g_scratch_ptr.end()) ptr = it->second;
  }
  if (!ptr) throw Error(ORQG_E_DEVICE, "distinct_noalloc: scratch of this device not set up");
  if (res->width == 4)
    distinct_t<uint32_t>(c, keys, nk, res, (uint32_t *)ptr);
  else
    distinct_t<uint64_t>(c, keys, nk, res, (uint64_t *)ptr);
}

}  // namespace orqg
