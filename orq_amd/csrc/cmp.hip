// Boolean comparison circuits and the full boolean-to-arithmetic conversion,
// each fused into one persistent kernel: all three parties' shares of an
// element stay in registers across every and_b / b2a_bit / multiply_a round
// of the circuit, and only the inputs and the final shares touch HBM.  Every
// round still draws its randomness from the keystream positions that the
// round's own runtime call would use, so the shares, nonces and message
// counts are those of the reference's sequence of calls.
#include <cstring>
#include <vector>

#include "internal.h"
#include "kern_common.h"
#include "cmp_core.h"

namespace orqg {

// BSharedVector::bit_same and _compare over a whole vector, one fused kernel
// (compare_group); neg_eq / neg_gt apply not_b_1 (replicated_3pc.h:207) to
// the output.
template <typename T, bool CMP>
__global__ void __launch_bounds__(kAesThreads)
    k_compare(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
              NonceTriple nb, uint64_t stride, CallSched s, CallSched sp,
              const T *__restrict__ x, const T *__restrict__ y, T *__restrict__ eq_out,
              T *__restrict__ gt_out, int neg_eq, int neg_gt,
              const PackedThreads *__restrict__ pt, uint64_t woff) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const int lane = threadIdx.x & 63;
  const uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t gw = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); gw < ngrp;
       gw += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = gw + lane;
    const uint64_t e0 = g * E;
    Grp3<T> X, Y, EQ, GT;
    X.load(x, e0, s.n);
    Y.load(y, e0, s.n);
    compare_group<T, CMP>(A, rk_pair, nb, 0, stride, s, sp, gw, lane, X.v, Y.v, EQ.v, GT.v, pt,
                          woff);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        EQ.v[i][k] ^= (T)neg_eq;
        GT.v[i][k] ^= (T)neg_gt;
      }
    if (g < ngrp) {
      if (eq_out) EQ.store(eq_out, e0, s.n);
      if (CMP && gt_out) GT.store(gt_out, e0, s.n);
    }
  }
}

void compare(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *eq, orqg_vec *gt,
             bool full, int neg_eq, int neg_gt) {
  const uint64_t n = x->n;
  if (n == 0) return;
  const int w = x->width;
  const int B = 8 * w;
  const int L = w == 4 ? 5 : w == 8 ? 6 : 7;
  const uint64_t nwords = n / B + (n % B > 0);
  TimedRegion tr(c, full ? "compare" : "equal");
  // With worker threads, each thread evaluates the element-wise calls over
  // its range of elements with its keys; the packed call (the sign fix) is
  // split by ranges of words, which the kernels look up in a table.
  const PackedThreads *dpt = nullptr;
  std::vector<std::pair<uint64_t, uint64_t>> eb(1, {0, n}), pb(1, {0, nwords});
  if (c->threads > 1) {
    const int T = c->threads;
    eb = thread_bounds(n, c->batch_size, T);
    pb = thread_bounds(nwords, c->batch_size, T);
    for (int t = 0; t < T; t++)
      if (eb[t].first % B) throw Error(ORQG_E_ARG, "compare: thread range not word-aligned");
    PackedThreads h;
    std::memset(&h, 0, sizeof(h));
    h.T = (uint32_t)T;
    for (int t = 0; t < T; t++) {
      ThreadScope ts(c, t);
      const uint64_t stride = sched_nonces(make_sched(eb[t].second - eb[t].first, c->batch_size, w, 1));
      PackedThread &P = h.t[t];
      P.lo = pb[t].first;
      P.hi = pb[t].second;
      for (int k = 0; k < 3; k++) P.nonce[k] = c->nonce[kPair0 + k] + (uint64_t)(L + 1) * stride;
      P.s = make_sched(P.hi - P.lo, c->batch_size, w, 1);
      P.rk = c->rk(kPair0);
    }
    PackedThreads *d = (PackedThreads *)c->scratch_get("cmp_threads", sizeof(PackedThreads));
    ORQG_HIP(hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));  // h is on the host stack
    dpt = d;
  }
  const char *xd = (const char *)x->d, *yd = (const char *)y->d;
  char *eqd = eq ? (char *)eq->d : nullptr, *gtd = gt ? (char *)gt->d : nullptr;
  for_threads(c, n, [&](uint64_t off, uint64_t m) {
    CallSched s = make_sched(m, c->batch_size, w, 1);
    CallSched sp = make_sched(m / B + (m % B > 0), c->batch_size, w, 1);
    const uint64_t stride = sched_nonces(s);
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    unsigned grid = grid_for((m * w + 15) / 16, kAesThreads, c->num_cu);
    const uint64_t boff = 3 * off * (uint64_t)w;
#define ORQG_CMP_LAUNCH(T)                                                                      \
  hipLaunchKernelGGL((full ? k_compare<T, true> : k_compare<T, false>), dim3(grid),            \
                     dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(kPair0), nb, stride, s, sp, \
                     (const T *)(xd + boff), (const T *)(yd + boff),                            \
                     eqd ? (T *)(eqd + boff) : (T *)nullptr, gtd ? (T *)(gtd + boff) : (T *)nullptr, \
                     neg_eq, neg_gt, dpt, off / B)
    if (w == 4)
      ORQG_CMP_LAUNCH(uint32_t);
    else if (w == 8)
      ORQG_CMP_LAUNCH(uint64_t);
    else
      ORQG_CMP_LAUNCH(u128);  // quicksort's padded 64-bit keys
#undef ORQG_CMP_LAUNCH
    ORQG_HIP(hipGetLastError());
  });
  // each thread's pair PRGs: its element-wise calls, then its packed call
  for (int t = 0; t < (int)eb.size(); t++) {
    ThreadScope ts(c, c->threads > 1 ? t : c->cur_thread);
    const uint64_t stride =
        sched_nonces(make_sched(eb[t].second - eb[t].first, c->batch_size, w, 1));
    uint64_t used = (uint64_t)L * stride;
    if (full) used += stride + sched_nonces(make_sched(pb[t].second - pb[t].first, c->batch_size, w, 1));
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  }
  const uint64_t bytes = (uint64_t)(full ? L + 1 : L) * n * w + (full ? nwords * (uint64_t)w : 0);
  for (int p = 0; p < 3; p++) c->log_msg("and_b", p, (p + 2) % 3, bytes);
}

// BSharedVector::b2a (b_shared_vector.h:296-329): for every bit i, b2a_bit of
// (x >> i) & 1 (one runtime call, three getNext calls per batch), then for
// i > 0 a multiply_a by the public sharing (2^i, 0, 0) (one call), summed.
// Call i's nonce base: base + i*n3 + max(i-1, 0)*n1.
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_b2a_full(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
               NonceTriple nb, uint64_t n3, uint64_t n1, CallSched s3, CallSched s1,
               const T *__restrict__ x, T *__restrict__ y) {
  constexpr int E = 16 / sizeof(T);
  constexpr int B = 8 * sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (s3.n + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e0 = g * E;
    Grp3<T> X, Y;
    X.load(x, e0, s3.n);
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s1, e0, 0, off, ctr, bib);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) Y.v[i][k] = 0;
#pragma unroll 1
    for (int b = 0; b < B; b++) {
      const uint64_t base = (uint64_t)b * n3 + (b > 0 ? (uint64_t)(b - 1) * n1 : 0);
      T bits[E][3], ya[E][3];
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) bits[i][k] = (X.v[i][k] >> b) & 1;
      NonceTriple nbb = {{nb.v[0] + base, nb.v[1] + base, nb.v[2] + base}};
      b2a_bit_group<T>(A, rk_pair, nbb, s3, e0, bits, ya);
      if (b > 0) {
        const T cst = (T)1 << b;
        T S[3][E];
        pair_stream<T>(A, rk_pair, nbb, n3, off, ctr, S);
#pragma unroll
        for (int i = 0; i < E; i++) {
          // multiply_a by (c, 0, 0): z_0 = (a_0 + a_1) c, z_1 = 0, z_2 = a_2 c
          ya[i][0] = (T)(S[2][i] - S[0][i]) + (ya[i][0] + ya[i][1]) * cst;
          ya[i][1] = (T)(S[0][i] - S[1][i]);
          ya[i][2] = (T)(S[1][i] - S[2][i]) + ya[i][2] * cst;
        }
      }
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Y.v[i][k] += ya[i][k];
    }
    Y.store(y, e0, s3.n);
  }
}

void b2a_full(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  const uint64_t n = x->n;
  if (n == 0) return;
  const int B = 8 * x->width;
  TimedRegion tr(c, "b2a");
  // every call is over the n elements, so each worker thread runs the whole
  // conversion over its range with its own keys and nonces
  for_threads(c, n, [&](uint64_t off, uint64_t m) {
    CallSched s3 = make_sched(m, c->batch_size, x->width, 3);
    CallSched s1 = make_sched(m, c->batch_size, x->width, 1);
    const uint64_t n3 = sched_nonces(s3), n1 = sched_nonces(s1);
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    unsigned grid = grid_for((m * x->width + 15) / 16, kAesThreads, c->num_cu);
    if (x->width == 4)
      hipLaunchKernelGGL(k_b2a_full<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, n3, n1, s3, s1,
                         (const uint32_t *)x->d + 3 * off, (uint32_t *)y->d + 3 * off);
    else
      hipLaunchKernelGGL(k_b2a_full<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kPair0), nb, n3, n1, s3, s1,
                         (const uint64_t *)x->d + 3 * off, (uint64_t *)y->d + 3 * off);
    ORQG_HIP(hipGetLastError());
    const uint64_t used = (uint64_t)B * n3 + (uint64_t)(B - 1) * n1;
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  });
  const uint64_t nb_bytes = n * (uint64_t)x->width;
  for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * nb_bytes * B);
  for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, nb_bytes * (B - 1));
}

}  // namespace orqg
