// Boolean comparison circuits and the full boolean-to-arithmetic conversion,
// each fused into one persistent kernel: all three parties' shares of an
// element stay in registers across every and_b / b2a_bit / multiply_a round
// of the circuit, and only the inputs and the final shares touch HBM.  Every
// round still draws its randomness from the keystream positions that the
// round's own runtime call would use, so the shares, nonces and message
// counts are those of the reference's sequence of calls.
#include "internal.h"
#include "kern_common.h"

namespace orqg {

// bit_width(bits - 1): the number of shift-and-AND levels of bit_same
// (b_shared_vector.h:164-178, LOG_MAX_BITS_NUMBER).
template <typename T>
constexpr int same_levels() {
  return sizeof(T) == 4 ? 5 : 6;
}

// z = x & y for the three parties (replicated_3pc.h:166-191) with the pair
// keystream S[k] of key K_k: z_p = S_{p-1} ^ S_p ^ x_p y_p ^ x_p y_{p+1} ^ x_{p+1} y_p.
template <typename T>
__device__ inline void and3(const T (&x)[3], const T (&y)[3], const T (&S)[3], T (&z)[3]) {
#pragma unroll
  for (int p = 0; p < 3; p++) {
    int q = (p + 1) % 3, pm = (p + 2) % 3;
    z[p] = (S[pm] ^ S[p]) ^ (x[p] & y[p]) ^ (x[p] & y[q]) ^ (x[q] & y[p]);
  }
}

// Keystream of one and_b call for the E elements starting at e0: S[k][i].
template <typename T>
__device__ inline void pair_stream(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                   const NonceTriple &nb, uint64_t nonce_add, uint64_t off,
                                   uint32_t ctr, T (&S)[3][16 / sizeof(T)]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + nonce_add + off, ctr);
    unpack_block<T>(blk, S[k]);
  }
}

// BSharedVector::bit_same (b_shared_vector.h:164-178) and, with CMP,
// BSharedVector::_compare (circuits.h:231-295) on signed elements.
//
// Runtime calls and their nonce bases on the three pair keys (stride = nonces
// one and_b over n elements uses):
//   level l < L        and_b(e, e >> 2^l)                 base + l*stride
//   CMP: L             and_b((e >> 1) ^ e, x)             base + L*stride
//   CMP: L + 1         and_b(~s2, ~r) on n/B packed words base + (L+1)*stride
// The packed words hold one bit per element (pack_from/unpack_from,
// mapping_access_vector.h:974-1013) and every operation on them is bitwise,
// so each lane evaluates its own elements' bits; it only needs its bits of
// the packed call's keystream, which a few lanes of the wave compute and
// the others read with cross-lane shuffles.  The loop is wave-uniform so the
// shuffles always run with the whole wave.
//
// neg_eq / neg_gt apply not_b_1 (replicated_3pc.h:207) to the output.
template <typename T, bool CMP>
__global__ void __launch_bounds__(kAesThreads)
    k_compare(const uint32_t *__restrict__ t0, const uint32_t *__restrict__ rk_pair,
              NonceTriple nb, uint64_t stride, CallSched s, CallSched sp,
              const T *__restrict__ x, const T *__restrict__ y, T *__restrict__ eq_out,
              T *__restrict__ gt_out, int neg_eq, int neg_gt) {
  typedef typename std::make_signed<T>::type ST;
  constexpr int E = 16 / sizeof(T);
  constexpr int B = 8 * sizeof(T);
  constexpr int L = same_levels<T>();
  constexpr int kBlkPerWave = 64 / B;  // AES blocks of packed words per wave
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const int lane = threadIdx.x & 63;
  const uint64_t ngrp = (s.n + E - 1) / E;
  const uint64_t nwords = sp.n;
  for (uint64_t gw = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); gw < ngrp;
       gw += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = gw + lane;
    const bool valid = g < ngrp;
    const uint64_t e0 = g * E;
    Grp3<T> X, Y;
    X.load(x, e0, s.n);
    Y.load(y, e0, s.n);
    // lanes past the end run the same arithmetic on zeros and store nothing
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s, e0, 0, off, ctr, bib);

    T e[E][3];
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) e[i][k] = ~(X.v[i][k] ^ Y.v[i][k]);
#pragma unroll 1
    for (int l = 0; l < L; l++) {
      T S[3][E];
      pair_stream<T>(A, rk_pair, nb, l * stride, off, ctr, S);
#pragma unroll
      for (int i = 0; i < E; i++) {
        T t[3], ks[3] = {S[0][i], S[1][i], S[2][i]};
#pragma unroll
        for (int k = 0; k < 3; k++) t[k] = (T)((ST)e[i][k] >> (1 << l));
        T z[3];
        and3<T>(e[i], t, ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) e[i][k] = z[k];
      }
    }

    Grp3<T> EQ, GT;
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) EQ.v[i][k] = (e[i][k] & 1) ^ (T)neg_eq;

    if (CMP) {
      // gt = parity(((e >> 1) ^ e) & x)
      T S[3][E];
      pair_stream<T>(A, rk_pair, nb, L * stride, off, ctr, S);
      T r[E][3];
#pragma unroll
      for (int i = 0; i < E; i++) {
        T a[3], z[3], ks[3] = {S[0][i], S[1][i], S[2][i]};
#pragma unroll
        for (int k = 0; k < 3; k++) a[k] = (T)((ST)e[i][k] >> 1) ^ e[i][k];
        and3<T>(a, X.v[i], ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) r[i][k] = (T)(__popcll((unsigned long long)z[k]) & 1);
      }

      // keystream of the packed and_b: lane j < 3*kBlkPerWave computes key
      // j%3 of the wave's AES block j/3
      const uint64_t wave_e0 = gw * E;  // multiple of 64*E, hence of B
      uint4 kb = make_uint4(0, 0, 0, 0);
      if (lane < 3 * kBlkPerWave) {
        const int key = lane % 3, blk = lane / 3;
        const uint64_t w = wave_e0 / B + (uint64_t)blk * E;
        if (w < nwords) {
          uint64_t poff;
          uint32_t pctr, pbib;
          sched_locate(sp, w, 0, poff, pctr, pbib);
          kb = aes_ctr_block<kCopies>(A, rk_pair + 60 * key, nb.v[key] + (L + 1) * stride + poff,
                                      pctr);
        }
      }
      // this lane's elements sit in dword D of the wave's packed keystream,
      // at bits b0 .. b0+E-1
      const int rel = lane * E;
      const int D = rel / 32, b0 = rel % 32;
      uint32_t kw[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int src = (D / 4) * 3 + k;
        uint32_t d0 = __shfl(kb.x, src), d1 = __shfl(kb.y, src), d2 = __shfl(kb.z, src),
                 d3 = __shfl(kb.w, src);
        const int dd = D % 4;
        kw[k] = dd == 0 ? d0 : dd == 1 ? d1 : dd == 2 ? d2 : d3;
      }
#pragma unroll
      for (int i = 0; i < E; i++) {
        T s1[3], s2[3], rr[3], a[3], b[3], z[3], ks[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          s1[k] = (X.v[i][k] >> (B - 1)) & 1;
          s2[k] = (Y.v[i][k] >> (B - 1)) & 1;
          rr[k] = r[i][k] ^ s2[k];  // r ^= s2
          s2[k] ^= s1[k];           // s2 ^= s1
          a[k] = s2[k] ^ 1;         // s2 |= r  as  ~(~s2 & ~r)
          b[k] = rr[k] ^ 1;
          ks[k] = (kw[k] >> (b0 + i)) & 1;
        }
        and3<T>(a, b, ks, z);
#pragma unroll
        for (int k = 0; k < 3; k++) GT.v[i][k] = (s1[k] ^ (z[k] ^ 1)) ^ (T)neg_gt;  // s1 ^= s2
      }
    }
    if (valid) {
      if (eq_out) EQ.store(eq_out, e0, s.n);
      if (CMP && gt_out) GT.store(gt_out, e0, s.n);
    }
  }
}

void compare(orqg_ctx *c, const orqg_vec *x, const orqg_vec *y, orqg_vec *eq, orqg_vec *gt,
             bool full, int neg_eq, int neg_gt) {
  const uint64_t n = x->n;
  if (n == 0) return;
  const int B = 8 * x->width;
  const int L = x->width == 4 ? 5 : 6;
  CallSched s = make_sched(n, c->batch_size, x->width, 1);
  const uint64_t nwords = n / B + (n % B > 0);
  CallSched sp = make_sched(nwords, c->batch_size, x->width, 1);
  const uint64_t stride = sched_nonces(s);
  NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
  unsigned grid = grid_for((n * x->width + 15) / 16, kAesThreads, c->num_cu);
  TimedRegion tr(c, full ? "compare" : "equal");
#define ORQG_CMP_LAUNCH(T)                                                                     \
  hipLaunchKernelGGL((full ? k_compare<T, true> : k_compare<T, false>), dim3(grid),           \
                     dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(kPair0), nb, stride, s, sp, \
                     (const T *)x->d, (const T *)y->d, eq ? (T *)eq->d : (T *)nullptr,        \
                     gt ? (T *)gt->d : (T *)nullptr, neg_eq, neg_gt)
  if (x->width == 4)
    ORQG_CMP_LAUNCH(uint32_t);
  else
    ORQG_CMP_LAUNCH(uint64_t);
#undef ORQG_CMP_LAUNCH
  ORQG_HIP(hipGetLastError());
  uint64_t used = (uint64_t)L * stride;
  if (full) used += stride + sched_nonces(sp);
  for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  const uint64_t bytes = (uint64_t)(full ? L + 1 : L) * n * x->width +
                         (full ? nwords * (uint64_t)x->width : 0);
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
  CallSched s3 = make_sched(n, c->batch_size, x->width, 3);
  CallSched s1 = make_sched(n, c->batch_size, x->width, 1);
  const uint64_t n3 = sched_nonces(s3), n1 = sched_nonces(s1);
  NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
  unsigned grid = grid_for((n * x->width + 15) / 16, kAesThreads, c->num_cu);
  TimedRegion tr(c, "b2a");
  if (x->width == 4)
    hipLaunchKernelGGL(k_b2a_full<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                       c->rk(kPair0), nb, n3, n1, s3, s1, (const uint32_t *)x->d,
                       (uint32_t *)y->d);
  else
    hipLaunchKernelGGL(k_b2a_full<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                       c->rk(kPair0), nb, n3, n1, s3, s1, (const uint64_t *)x->d,
                       (uint64_t *)y->d);
  ORQG_HIP(hipGetLastError());
  const uint64_t used = (uint64_t)B * n3 + (uint64_t)(B - 1) * n1;
  for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  const uint64_t nb_bytes = n * (uint64_t)x->width;
  for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * nb_bytes * B);
  for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, nb_bytes * (B - 1));
}

}  // namespace orqg
