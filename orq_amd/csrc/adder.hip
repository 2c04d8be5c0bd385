// Boolean adders (operators/circuits.h:35-216) and the arithmetic-to-boolean
// conversion (ASharedVector::a2b, a_shared_vector.h:138-144).
//
// Ripple carry: the reference runs the carry recurrence on bit-packed words
// (bit q of word w is element w*B + q), one and_b call per bit position.
// Every operation on the packed words is bitwise, so here each lane of a wave
// carries one element and runs the recurrence on its own bits; it needs only
// its bit of each packed call's keystream.  A wave covers whole packed words,
// so all 3*(B-1) AES blocks a wave needs are computed up front, spread over
// its 64 lanes, and each step reads its block with cross-lane shuffles.
#include <algorithm>

#include "internal.h"
#include "kern_common.h"
#include "cmp_core.h"

namespace orqg {
namespace {

// RCA over n elements: mode 0 sum (operator+ / operator-), mode 1 compare
// (rca_compare: the sign of x + ~y + 1, as 0/1 shares).  negy complements y
// on load (subtraction: x + ~y with carry-in 1).
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_rca(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair, NonceTriple nb,
          CallSched sp, uint64_t Np, uint64_t n, const T *__restrict__ x,
          const T *__restrict__ y, int negy, int carry_in, int mode, T *__restrict__ out) {
  constexpr int B = 8 * sizeof(T);
  constexpr int WPB = 16 / sizeof(T);  // packed words per AES block
  constexpr int ITEMS = 3 * (B - 1);   // (call, key) pairs
  constexpr int ROUNDS = (ITEMS + 63) / 64;
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const int lane = threadIdx.x & 63;
  const uint64_t nwave = (n + 63) / 64;
  const uint64_t nwords = sp.n;
  for (uint64_t wv = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; wv < nwave;
       wv += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t e = wv * 64 + lane;
    const bool ok = e < n;
    // the wave's packed words all lie in the AES block of its first word
    const uint64_t w0 = (wv * 64) / B;
    const uint64_t blk_word = w0 & ~(uint64_t)(WPB - 1);
    uint64_t poff = 0;
    uint32_t pctr = 0, pbib;
    if (blk_word < nwords) sched_locate(sp, blk_word, 0, poff, pctr, pbib);
    uint4 kb[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
      const int t = r * 64 + lane;
      kb[r] = make_uint4(0, 0, 0, 0);
      if (t < ITEMS && blk_word < nwords) {
        const int call = t / 3, key = t % 3;
        kb[r] = aes_ctr_block<kCopies>(A, rk_pair + 60 * key,
                                       nb.v[key] + (uint64_t)call * Np + poff, pctr);
      }
    }
    // this lane's bit sits in dword D of each block at bit b
    const uint64_t wrd = e / B;
    const int q = (int)(e % B);
    const int D = (int)(wrd % WPB) * (B / 32) + q / 32, bpos = q % 32;
    T xs[3], ys[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      xs[k] = ok ? x[3 * e + k] : (T)0;
      ys[k] = ok ? y[3 * e + k] : (T)0;
      if (negy) ys[k] = ~ys[k];
    }
    T sum[3] = {0, 0, 0};
    uint32_t cb[3];  // carry bit of each share
#pragma unroll
    for (int k = 0; k < 3; k++) cb[k] = carry_in ? 1u : 0u;
#pragma unroll 1
    for (int i = 0; i < B; i++) {
      uint32_t xb[3], ab[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        xb[k] = (uint32_t)((xs[k] ^ ys[k]) >> i) & 1u;
        ab[k] = (uint32_t)(xs[k] >> i) & 1u;
        sum[k] |= (T)(xb[k] ^ cb[k]) << i;
      }
      if (i == B - 1) {
        if (mode == 1)
#pragma unroll
          for (int k = 0; k < 3; k++) cb[k] ^= xb[k];
        break;
      }
      uint32_t ks[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int t = 3 * i + k, r = t / 64, src = t % 64;
        uint4 v = kb[0];
#pragma unroll
        for (int rr = 1; rr < ROUNDS; rr++)
          if (r == rr) v = kb[rr];
        uint32_t d0 = __shfl(v.x, src), d1 = __shfl(v.y, src), d2 = __shfl(v.z, src),
                 d3 = __shfl(v.w, src);
        const uint32_t dw = D == 0 ? d0 : D == 1 ? d1 : D == 2 ? d2 : d3;
        ks[k] = (dw >> bpos) & 1u;
      }
      // c = ((c ^ a_i) & axb_i) ^ a_i, the and_b on one bit of each share
      uint32_t cc[3], z[3];
#pragma unroll
      for (int k = 0; k < 3; k++) cc[k] = cb[k] ^ ab[k];
      and3<uint32_t>(cc, xb, ks, z);
#pragma unroll
      for (int k = 0; k < 3; k++) cb[k] = z[k] ^ ab[k];
    }
    if (ok) {
#pragma unroll
      for (int k = 0; k < 3; k++) out[3 * e + k] = mode == 1 ? (T)cb[k] : sum[k];
    }
  }
}

// parallel_prefix_adder (circuits.h:138-216), 1 + 3*log2(B) and_b calls
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_ppa(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair, NonceTriple nb,
          CallSched s, uint64_t N1, const T *__restrict__ x, const T *__restrict__ y, int negy,
          int carry_in, T *__restrict__ out) {
  constexpr int E = 16 / sizeof(T);
  constexpr int LOG = sizeof(T) == 4 ? 5 : 6;
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g0 < ngrp;
       g0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e0 = g0 * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s, e0, 0, off, ctr, bib);
    Grp3<T> X, Y, R;
    X.load(x, e0, s.n);
    Y.load(y, e0, s.n);
    T p[E][3], g[E][3], prop[E][3];
    T S[3][E];
    pair_stream<T, true>(A, rk_pair, nb, 0, off, ctr, S);  // g = x & y
#pragma unroll
    for (int i = 0; i < E; i++) {
      T yy[3], ks[3] = {S[0][i], S[1][i], S[2][i]};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        yy[k] = negy ? (T)~Y.v[i][k] : Y.v[i][k];
        p[i][k] = X.v[i][k] ^ yy[k];
      }
      and3<T>(X.v[i], yy, ks, g[i]);
      if (carry_in) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const T lsb = p[i][k] & 1;
          if (k == 0) p[i][k] ^= 1;  // p ^ public 1 changes s_0
          g[i][k] ^= lsb;
        }
      }
#pragma unroll
      for (int k = 0; k < 3; k++) prop[i][k] = p[i][k];
    }
#pragma unroll 1
    for (int l = 0; l < LOG; l++) {
      const int sh = 1 << l;
      T ps[E][3], gs[E][3];
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
          ps[i][k] = (T)(p[i][k] << sh);
          gs[i][k] = (T)(g[i][k] << sh);
        }
      const uint64_t base = (uint64_t)(1 + 3 * l) * N1;
      pair_stream<T, true>(A, rk_pair, nb, base, off, ctr, S);  // p_y &= p_x
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
        and3<T>(ps[i], p[i], ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) ps[i][k] = z[k];
      }
      pair_stream<T, true>(A, rk_pair, nb, base + N1, off, ctr, S);  // g_y &= p_x
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
        and3<T>(gs[i], p[i], ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) gs[i][k] = z[k];
      }
      pair_stream<T, true>(A, rk_pair, nb, base + 2 * N1, off, ctr, S);  // g_x & g_y
#pragma unroll
      for (int i = 0; i < E; i++) {
        T ks[3] = {S[0][i], S[1][i], S[2][i]}, z[3];
        and3<T>(g[i], gs[i], ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          g[i][k] = g[i][k] ^ gs[i][k] ^ z[k];
          p[i][k] = ps[i][k];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) R.v[i][k] = prop[i][k] ^ (T)(g[i][k] << 1);
    R.store(out, e0, s.n);
  }
}

// redistribute_shares_b (replicated_3pc.h:259-262): party 0 secret-shares
// s0 + s1 with its local PRG, party 2 secret-shares s2 with its own
// (get_shares_b, replicated_3pc.h:384-392; owner d keeps (r1, r2), d+1 gets
// (r2, r3), d+2 gets (r3, r1)).
template <typename T>
__global__ void __launch_bounds__(kAesThreads)
    k_redistribute(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk0,
                   const uint32_t *__restrict__ rk2, uint64_t n0, uint64_t n2, CallSched s,
                   const T *__restrict__ x, T *__restrict__ f, T *__restrict__ sv) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g0 < ngrp;
       g0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e0 = g0 * E;
    T r0[2][E], r2[2][E];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      uint64_t off;
      uint32_t ctr, bib;
      sched_locate(s, e0, j, off, ctr, bib);
      unpack_block<T>(aes_ctr_block<kCopies>(A, rk0, n0 + off, ctr), r0[j]);
      unpack_block<T>(aes_ctr_block<kCopies>(A, rk2, n2 + off, ctr), r2[j]);
    }
    Grp3<T> X, F, SV;
    X.load(x, e0, s.n);
#pragma unroll
    for (int i = 0; i < E; i++) {
      const T v0 = X.v[i][0] + X.v[i][1], v2 = X.v[i][2];
      F.v[i][0] = r0[0][i];
      F.v[i][1] = r0[1][i];
      F.v[i][2] = v0 ^ r0[0][i] ^ r0[1][i];
      SV.v[i][2] = r2[0][i];
      SV.v[i][0] = r2[1][i];
      SV.v[i][1] = v2 ^ r2[0][i] ^ r2[1][i];
    }
    F.store(f, e0, s.n);
    SV.store(sv, e0, s.n);
  }
}

}  // namespace

// op: 0 RCA add, 1 RCA subtract, 2 PPA add, 3 PPA subtract, 4 rca_compare
void adder(orqg_ctx *c, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  const uint64_t n = x->n;
  if (n == 0) return;
  const int w = x->width, B = 8 * w;
  const bool sub = op == 1 || op == 3;
  TimedRegion tr(c, "adder");
  if (op == 2 || op == 3) {
    // every call is over the n elements: each worker thread runs the whole
    // adder over its element range with its own keys and nonces
    const int LOG = w == 4 ? 5 : 6;
    const uint64_t calls = 1 + 3 * (uint64_t)LOG;
    for_threads(c, n, [&](uint64_t off, uint64_t m) {
      NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
      CallSched s = make_sched(m, c->batch_size, w, 1);
      const uint64_t N1 = sched_nonces(s);
      unsigned grid = grid_for((m * w + 15) / 16, kAesThreads, c->num_cu);
      if (w == 4)
        hipLaunchKernelGGL(k_ppa<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                           c->rk(kPair0), nb, s, N1, (const uint32_t *)x->d + 3 * off,
                           (const uint32_t *)y->d + 3 * off, sub ? 1 : 0, sub ? 1 : 0,
                           (uint32_t *)z->d + 3 * off);
      else
        hipLaunchKernelGGL(k_ppa<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                           c->rk(kPair0), nb, s, N1, (const uint64_t *)x->d + 3 * off,
                           (const uint64_t *)y->d + 3 * off, sub ? 1 : 0, sub ? 1 : 0,
                           (uint64_t *)z->d + 3 * off);
      ORQG_HIP(hipGetLastError());
      for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += calls * N1;
    });
    for (int p = 0; p < 3; p++) c->log_msg("and_b", p, (p + 2) % 3, calls * n * w);
    return;
  }
  // the ripple-carry adder's calls are over the packed words (B elements
  // each): a worker thread's range of words is a range of elements
  const uint64_t nwords = n / B + (n % B > 0);
  const int negy = (op == 1 || op == 4) ? 1 : 0, cin = negy, mode = op == 4 ? 1 : 0;
  for_threads(c, nwords, [&](uint64_t woff, uint64_t wm) {
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    CallSched sp = make_sched(wm, c->batch_size, w, 1);
    const uint64_t Np = sched_nonces(sp);
    const uint64_t off = woff * B, m = std::min<uint64_t>(n - off, wm * B);
    unsigned grid = grid_for((m + 63) / 64 * 64, kAesThreads, c->num_cu);
    if (w == 4)
      hipLaunchKernelGGL(k_rca<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                         c->rk(kPair0), nb, sp, Np, m, (const uint32_t *)x->d + 3 * off,
                         (const uint32_t *)y->d + 3 * off, negy, cin, mode,
                         (uint32_t *)z->d + 3 * off);
    else
      hipLaunchKernelGGL(k_rca<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                         c->rk(kPair0), nb, sp, Np, m, (const uint64_t *)x->d + 3 * off,
                         (const uint64_t *)y->d + 3 * off, negy, cin, mode,
                         (uint64_t *)z->d + 3 * off);
    ORQG_HIP(hipGetLastError());
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += (uint64_t)(B - 1) * Np;
  });
  for (int p = 0; p < 3; p++) c->log_msg("and_b", p, (p + 2) % 3, (uint64_t)(B - 1) * nwords * w);
}

void a2b(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  const uint64_t n = x->n;
  if (n == 0) return;
  const int w = x->width;
  orqg_vec f = *x, sv = *x;
  f.enc = sv.enc = ORQG_BSHARED;
  f.d = c->scratch_get("a2b_f", 3 * n * w + 64);
  sv.d = c->scratch_get("a2b_s", 3 * n * w + 64);
  for_threads(c, n, [&](uint64_t off, uint64_t m) {
    CallSched s = make_sched(m, c->batch_size, w, 2);
    unsigned grid = grid_for((m * w + 15) / 16, kAesThreads, c->num_cu);
    if (w == 4)
      hipLaunchKernelGGL(k_redistribute<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kLocal0), c->rk(kLocal0 + 2), c->nonce[kLocal0],
                         c->nonce[kLocal0 + 2], s, (const uint32_t *)x->d + 3 * off,
                         (uint32_t *)f.d + 3 * off, (uint32_t *)sv.d + 3 * off);
    else
      hipLaunchKernelGGL(k_redistribute<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream,
                         c->d_t0, c->rk(kLocal0), c->rk(kLocal0 + 2), c->nonce[kLocal0],
                         c->nonce[kLocal0 + 2], s, (const uint64_t *)x->d + 3 * off,
                         (uint64_t *)f.d + 3 * off, (uint64_t *)sv.d + 3 * off);
    ORQG_HIP(hipGetLastError());
    const uint64_t used = sched_nonces(s);
    c->nonce[kLocal0] += used;
    c->nonce[kLocal0 + 2] += used;
  });
  const uint64_t bytes = 2 * n * (uint64_t)w;
  c->log_msg("secret_share", 0, 1, bytes);
  c->log_msg("secret_share", 0, 2, bytes);
  c->log_msg("secret_share", 2, 0, bytes);
  c->log_msg("secret_share", 2, 1, bytes);
  adder(c, 0, &f, &sv, y);
  y->enc = ORQG_BSHARED;
}

}  // namespace orqg
