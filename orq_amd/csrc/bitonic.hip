// Bitonic sorting network over secret-shared rows (operators/sorting.h:80-353).
// Each compare-exchange step runs over the n/2 row pairs of the step as one
// fused persistent kernel per sort key (its _compare; from the second key on,
// one more for the lexicographic combination) and one for the swaps: the oblivious
// multiplexes of the key and boolean data columns (operators/common.h:20-48)
// and, when arithmetic data columns exist, one for b2a_bit of the swap bits
// and one for the arithmetic multiplexes.  Every
// one of those runtime calls draws its zero-sharing keystream at the nonce
// offset the call would have in the reference's sequence, so the shares,
// nonces and message counts match it.
#include <vector>

#include "internal.h"
#include "kern_common.h"
#include "cmp_core.h"

namespace orqg {
namespace {

constexpr int kMaxCols = 8;

struct StepArgs {
  uint64_t m;       // row pairs in the step (n / 2)
  uint32_t hshift;  // half box size h = 2^hshift
  int mirror;       // first step of a round: y runs backwards through the box
  int nk, nbc, nac;
  uint32_t desc_mask;
  // device table of column pointers: keys, then boolean, then arithmetic data
  void *const *cols;
  CallSched s1, sp, s3;  // and_b over m rows, the packed and_b, b2a_bit
  uint64_t N1, Np, N3;   // nonces each of those calls uses on every pair key
};

// Row indices of pair q (alternating_subset_reference and its reversed form,
// mapping_access_vector.h:788-845, as bitonic_sort builds them).
__device__ inline void pair_rows(const StepArgs &a, uint64_t q, uint64_t &xr, uint64_t &yr) {
  const uint64_t h = 1ull << a.hshift;
  const uint64_t box = q >> a.hshift, t = q & (h - 1);
  xr = box * 2 * h + t;
  yr = a.mirror ? box * 2 * h + 2 * h - 1 - t : box * 2 * h + h + t;
}

template <typename T, int E>
__device__ inline void load_rows(const T *__restrict__ col, const uint64_t (&r)[E],
                                 const bool (&ok)[E], T (&v)[E][3]) {
#pragma unroll
  for (int i = 0; i < E; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) v[i][k] = ok[i] ? col[3 * r[i] + k] : (T)0;
}
template <typename T, int E>
__device__ inline void store_rows(T *__restrict__ col, const uint64_t (&r)[E], const bool (&ok)[E],
                                  const T (&v)[E][3]) {
#pragma unroll
  for (int i = 0; i < E; i++)
    if (ok[i])
#pragma unroll
      for (int k = 0; k < 3; k++) col[3 * r[i] + k] = v[i][k];
}

// multiply_a of the three parties (replicated_3pc.h:44-67)
template <typename T>
__device__ inline void mul3(const T (&x)[3], const T (&y)[3], const T (&S)[3], T (&z)[3]) {
#pragma unroll
  for (int p = 0; p < 3; p++) {
    int q = (p + 1) % 3, pm = (p + 2) % 3;
    z[p] = (T)(S[pm] - S[p]) + x[p] * y[p] + x[p] * y[q] + x[q] * y[p];
  }
}

// compare_rows (sorting.h:80-108), key k of the step: t._compare(o) with
// t = x (ASC) or y (DESC) at nonce offset noff, written to (eqo, gto).
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_bitonic_cmp(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
                  NonceTriple nb, StepArgs a, int k, uint64_t noff, T *__restrict__ eqo,
                  T *__restrict__ gto) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const int lane = threadIdx.x & 63;
  const uint64_t ngrp = (a.m + E - 1) / E;
  const bool desc = (a.desc_mask >> k) & 1;
  const T *col = (const T *)a.cols[k];
  for (uint64_t gw = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); gw < ngrp;
       gw += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = (gw + lane) * E;
    T X[E][3], Y[E][3];
    Grp3<T> EQ, GT;
#pragma unroll
    for (int i = 0; i < E; i++) {
      const bool ok = q0 + i < a.m;
      uint64_t xr, yr;
      pair_rows(a, ok ? q0 + i : 0, xr, yr);
      const uint64_t tr = desc ? yr : xr, orow = desc ? xr : yr;
#pragma unroll
      for (int p = 0; p < 3; p++) {
        X[i][p] = ok ? col[3 * tr + p] : (T)0;
        Y[i][p] = ok ? col[3 * orow + p] : (T)0;
      }
    }
    compare_group<T, true>(A, rk_pair, nb, noff, a.N1, a.s1, a.sp, gw, lane, X, Y, EQ.v, GT.v);
    if (q0 < a.m) {
      EQ.store(eqo, q0, a.m);
      GT.store(gto, q0, a.m);
    }
  }
}

// Lexicographic combination for a later key (sorting.h:98-103):
// gt ^= new_gt & eq, then eq &= new_eq, two and_b calls from nonce offset noff.
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_bitonic_combine(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
                      NonceTriple nb, StepArgs a, uint64_t noff, T *__restrict__ eqb,
                      T *__restrict__ gtb, const T *__restrict__ neq,
                      const T *__restrict__ ngt) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = g * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(a.s1, q0, 0, off, ctr, bib);
    Grp3<T> EQ, GT, NEQ, NGT;
    EQ.load(eqb, q0, a.m);
    GT.load(gtb, q0, a.m);
    NEQ.load(neq, q0, a.m);
    NGT.load(ngt, q0, a.m);
    T S[3][E];
    pair_stream<T, true>(A, rk_pair, nb, noff, off, ctr, S);
#pragma unroll
    for (int i = 0; i < E; i++) {
      T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
      and3<T>(NGT.v[i], EQ.v[i], ks, z);
#pragma unroll
      for (int p = 0; p < 3; p++) GT.v[i][p] ^= z[p];
    }
    pair_stream<T, true>(A, rk_pair, nb, noff + a.N1, off, ctr, S);
#pragma unroll
    for (int i = 0; i < E; i++) {
      T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
      and3<T>(EQ.v[i], NEQ.v[i], ks, z);
#pragma unroll
      for (int p = 0; p < 3; p++) EQ.v[i][p] = z[p];
    }
    EQ.store(eqb, q0, a.m);
    GT.store(gtb, q0, a.m);
  }
}

// swap (sorting.h:146-162) with the bits gt.mask(1).  Boolean columns
// [c0, c1) of the pointer table: x' = mux(bits, x, y), y' = mux(bits, y, x),
// mux(s, a, b) = a ^ (extend_lsb(s) & (b ^ a)); with ARITH, arithmetic
// columns and mux(s, a, b) = a + s * (b - a), s = bits.b2a_bit() read from
// `sel`.  Two runtime calls per column, starting at nonce offset noff0.
template <typename T, bool ARITH>
__global__ void __launch_bounds__(kAesThreads)
    k_bitonic_swap(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
                   NonceTriple nb, uint64_t noff0, StepArgs a, int c0, int c1,
                   const T *__restrict__ sel) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = g * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(a.s1, q0, 0, off, ctr, bib);
    uint64_t noff = noff0;
#pragma unroll 1
    for (int c = c0; c < c1; c++) {
      T *col = (T *)a.cols[c];
      T tmpx[E][3];
#pragma unroll 1
      for (int half = 0; half < 2; half++) {
        T S[3][E];
        pair_stream<T, true>(A, rk_pair, nb, noff, off, ctr, S);
        noff += a.N1;
#pragma unroll
        for (int i = 0; i < E; i++) {
          if (q0 + i >= a.m) continue;
          uint64_t xr, yr;
          pair_rows(a, q0 + i, xr, yr);
          const uint64_t ra = half == 0 ? xr : yr, rb = half == 0 ? yr : xr;
          T va[3], vb[3], d[3], z[3], sv[3], ks[3] = {S[0][i], S[1][i], S[2][i]};
#pragma unroll
          for (int p = 0; p < 3; p++) {
            va[p] = col[3 * ra + p];
            vb[p] = col[3 * rb + p];
            sv[p] = sel[3 * (q0 + i) + p];
          }
          if (!ARITH) {
#pragma unroll
            for (int p = 0; p < 3; p++) {
              d[p] = vb[p] ^ va[p];
              sv[p] = (T)0 - (sv[p] & 1);  // extend_lsb
            }
            and3<T>(sv, d, ks, z);
#pragma unroll
            for (int p = 0; p < 3; p++) z[p] ^= va[p];
          } else {
#pragma unroll
            for (int p = 0; p < 3; p++) d[p] = vb[p] - va[p];
            mul3<T>(sv, d, ks, z);
#pragma unroll
            for (int p = 0; p < 3; p++) z[p] += va[p];
          }
          // both halves read the columns as they were: store after the second
          if (half == 0) {
#pragma unroll
            for (int p = 0; p < 3; p++) tmpx[i][p] = z[p];
          } else {
#pragma unroll
            for (int p = 0; p < 3; p++) {
              col[3 * xr + p] = tmpx[i][p];
              col[3 * yr + p] = z[p];
            }
          }
        }
      }
    }
  }
}

// bits.b2a_bit() of the swap bits (replicated_3pc.h:220-251) at nonce offset noff
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_bitonic_b2a(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
                  NonceTriple nb, uint64_t noff, StepArgs a, const T *__restrict__ gtb,
                  T *__restrict__ out) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (a.m + E - 1) / E;
  const NonceTriple nbb = {{nb.v[0] + noff, nb.v[1] + noff, nb.v[2] + noff}};
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q0 = g * E;
    Grp3<T> X, Y;
    X.load(gtb, q0, a.m);
    T bits[E][3];
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int p = 0; p < 3; p++) bits[i][p] = X.v[i][p] & 1;
    b2a_bit_group<T>(A, rk_pair, nbb, a.s3, q0, bits, Y.v);
    Y.store(out, q0, a.m);
  }
}

}  // namespace

// operators::bitonic_sort (sorting.h:266-353) on B-shared keys with A- and
// B-shared data columns; n must be a power of two.
void bitonic_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nk, const int *orders,
                  orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols, size_t nb) {
  require_one_thread(c, "bitonic_sort");
  if (nk == 0) throw Error(ORQG_E_ARG, "bitonic_sort: no sort keys");
  if (nk > (size_t)kMaxCols || na > (size_t)kMaxCols || nb > (size_t)kMaxCols)
    throw Error(ORQG_E_ARG, "bitonic_sort: at most 8 key, 8 A-shared and 8 B-shared columns");
  const uint64_t n = keys[0]->n;
  const int w = keys[0]->width;
  for (size_t k = 0; k < nk; k++)
    if (keys[k]->width != w) throw Error(ORQG_E_ARG, "bitonic_sort: columns differ in width");
  for (size_t k = 0; k < na; k++)
    if (a_cols[k]->width != w) throw Error(ORQG_E_ARG, "bitonic_sort: columns differ in width");
  for (size_t k = 0; k < nb; k++)
    if (b_cols[k]->width != w) throw Error(ORQG_E_ARG, "bitonic_sort: columns differ in width");
  if (n < 2) return;
  if (n & (n - 1)) throw Error(ORQG_E_ARG, "bitonic_sort: size must be a power of two");
  int rounds = 0;
  while ((1ull << rounds) < n) rounds++;

  StepArgs a{};
  a.m = n / 2;
  a.nk = (int)nk;
  a.nbc = (int)nb;
  a.nac = (int)na;
  std::vector<void *> ptrs;
  for (size_t k = 0; k < nk; k++) {
    ptrs.push_back(keys[k]->d);
    if (orders && orders[k] == ORQG_DESC) a.desc_mask |= 1u << k;
  }
  for (size_t k = 0; k < nb; k++) ptrs.push_back(b_cols[k]->d);
  for (size_t k = 0; k < na; k++) ptrs.push_back(a_cols[k]->d);
  void **dptrs = (void **)c->scratch_get("bitonic_cols", ptrs.size() * sizeof(void *) + 64);
  ORQG_HIP(hipMemcpyAsync(dptrs, ptrs.data(), ptrs.size() * sizeof(void *),
                          hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  a.cols = dptrs;
  const int B = 8 * w, L = w == 4 ? 5 : 6;
  const uint64_t nwords = a.m / B + (a.m % B > 0);
  a.s1 = make_sched(a.m, c->batch_size, w, 1);
  a.sp = make_sched(nwords, c->batch_size, w, 1);
  a.s3 = make_sched(a.m, c->batch_size, w, 3);
  a.N1 = sched_nonces(a.s1);
  a.Np = sched_nonces(a.sp);
  a.N3 = sched_nonces(a.s3);
  const uint64_t used = nk * ((uint64_t)(L + 1) * a.N1 + a.Np) + (nk - 1) * 2 * a.N1 +
                        2 * (nk + nb) * a.N1 + (na ? a.N3 + 2 * na * a.N1 : 0);
  const uint64_t mw = a.m * (uint64_t)w;
  const uint64_t and_bytes = nk * ((uint64_t)(L + 1) * mw + nwords * w) + (nk - 1) * 2 * mw +
                             2 * (nk + nb) * mw;
  // nonces of the comparison: key 0 uses (L+1) and_b + the packed one,
  // every later key two more and_b
  const uint64_t cmp_key = (uint64_t)(L + 1) * a.N1 + a.Np;
  const uint64_t cmp_used = nk * cmp_key + (nk - 1) * 2 * a.N1;
  unsigned grid = grid_for((a.m + (16 / w) - 1) / (16 / w), kAesThreads, c->num_cu);
  void *eqb = c->scratch_get("bitonic_eq", 3 * a.m * (size_t)w + 64);
  void *gtb = c->scratch_get("bitonic_gt", 3 * a.m * (size_t)w + 64);
  void *eq2 = nk > 1 ? c->scratch_get("bitonic_eq2", 3 * a.m * (size_t)w + 64) : nullptr;
  void *gt2 = nk > 1 ? c->scratch_get("bitonic_gt2", 3 * a.m * (size_t)w + 64) : nullptr;
  TimedRegion tr(c, "bitonic");
  for (int i = 0; i < rounds; i++)
    for (int j = 0; j <= i; j++) {
      a.hshift = (uint32_t)(i - j);
      a.mirror = j == 0;
      NonceTriple nbt = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
      for (int k = 0; k < (int)nk; k++) {
        const uint64_t noff = k == 0 ? 0 : cmp_key + (uint64_t)(k - 1) * (cmp_key + 2 * a.N1);
        void *eo = k == 0 ? eqb : eq2, *go = k == 0 ? gtb : gt2;
        if (w == 4)
          hipLaunchKernelGGL(k_bitonic_cmp<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                             c->d_t0, c->rk(kPair0), nbt, a, k, noff, (uint32_t *)eo,
                             (uint32_t *)go);
        else
          hipLaunchKernelGGL(k_bitonic_cmp<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                             c->d_t0, c->rk(kPair0), nbt, a, k, noff, (uint64_t *)eo,
                             (uint64_t *)go);
        if (k > 0) {
          if (w == 4)
            hipLaunchKernelGGL(k_bitonic_combine<uint32_t>, dim3(grid), dim3(kAesThreads), 0,
                               c->stream, c->d_t0, c->rk(kPair0), nbt, a, noff + cmp_key,
                               (uint32_t *)eqb, (uint32_t *)gtb, (const uint32_t *)eq2,
                               (const uint32_t *)gt2);
          else
            hipLaunchKernelGGL(k_bitonic_combine<uint64_t>, dim3(grid), dim3(kAesThreads), 0,
                               c->stream, c->d_t0, c->rk(kPair0), nbt, a, noff + cmp_key,
                               (uint64_t *)eqb, (uint64_t *)gtb, (const uint64_t *)eq2,
                               (const uint64_t *)gt2);
        }
      }
      const int cb = (int)(nk + nb), ca = (int)(nk + nb + na);
      const uint64_t b2a_off = cmp_used + 2 * (uint64_t)cb * a.N1;
      if (w == 4) {
        hipLaunchKernelGGL((k_bitonic_swap<uint32_t, false>), dim3(grid), dim3(kAesThreads), 0,
                           c->stream, c->d_t0, c->rk(kPair0), nbt, cmp_used, a, 0, cb,
                           (const uint32_t *)gtb);
        if (na) {
          hipLaunchKernelGGL(k_bitonic_b2a<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                             c->d_t0, c->rk(kPair0), nbt, b2a_off, a, (const uint32_t *)gtb,
                             (uint32_t *)eqb);
          hipLaunchKernelGGL((k_bitonic_swap<uint32_t, true>), dim3(grid), dim3(kAesThreads), 0,
                             c->stream, c->d_t0, c->rk(kPair0), nbt, b2a_off + a.N3, a, cb, ca,
                             (const uint32_t *)eqb);
        }
      } else {
        hipLaunchKernelGGL((k_bitonic_swap<uint64_t, false>), dim3(grid), dim3(kAesThreads), 0,
                           c->stream, c->d_t0, c->rk(kPair0), nbt, cmp_used, a, 0, cb,
                           (const uint64_t *)gtb);
        if (na) {
          hipLaunchKernelGGL(k_bitonic_b2a<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                             c->d_t0, c->rk(kPair0), nbt, b2a_off, a, (const uint64_t *)gtb,
                             (uint64_t *)eqb);
          hipLaunchKernelGGL((k_bitonic_swap<uint64_t, true>), dim3(grid), dim3(kAesThreads), 0,
                             c->stream, c->d_t0, c->rk(kPair0), nbt, b2a_off + a.N3, a, cb, ca,
                             (const uint64_t *)eqb);
        }
      }
      ORQG_HIP(hipGetLastError());
      for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
      for (int p = 0; p < 3; p++) c->log_msg("and_b", p, (p + 2) % 3, and_bytes);
      if (na) {
        for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * mw);
        for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, 2 * na * mw);
      }
    }
}

}  // namespace orqg
