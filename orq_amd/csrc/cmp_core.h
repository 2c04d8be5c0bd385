// Device form of the comparison circuits, shared by the vector comparison
// kernels and the bitonic sorting network.
#pragma once
#include "internal.h"
#include "kern_common.h"

namespace orqg {

// bit_width(bits - 1): the number of shift-and-AND levels of bit_same
// (b_shared_vector.h:164-178, LOG_MAX_BITS_NUMBER).
template <typename T>
constexpr int same_levels() {
  return sizeof(T) == 4 ? 5 : sizeof(T) == 8 ? 6 : 7;
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
// SEQ computes the three keys' blocks one after the other, which holds fewer
// registers than interleaving them (for kernels with much live state).
template <typename T, bool SEQ = false>
__device__ inline void pair_stream_impl(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                   const NonceTriple &nb, uint64_t nonce_add, uint64_t off,
                                   uint32_t ctr, T (&S)[3][16 / sizeof(T)]) {
#pragma unroll 1
  for (int k = 0; k < 3; k++) {
    uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + nonce_add + off, ctr);
    unpack_block<T>(blk, S[k]);
  }
}

template <typename T, bool SEQ = false>
__device__ inline void pair_stream(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                   const NonceTriple &nb, uint64_t nonce_add, uint64_t off,
                                   uint32_t ctr, T (&S)[3][16 / sizeof(T)]) {
  if (SEQ) {
    pair_stream_impl<T, true>(A, rk_pair, nb, nonce_add, off, ctr, S);
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + nonce_add + off, ctr);
      unpack_block<T>(blk, S[k]);
    }
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
// With more than one worker thread the packed call is split into the
// threads' ranges of packed words, which need not match their ranges of
// elements: each thread's word range, its pair keys and its nonces at the
// start of its packed call.
constexpr int kMaxWorkerThreads = 64;
struct PackedThread {
  uint64_t lo, hi;    // packed words [lo, hi) of the whole call
  uint64_t nonce[3];  // the thread's pair-key nonces where its packed call starts
  CallSched s;        // schedule of its hi - lo words
  const uint32_t *rk; // its three pair keys (3 x 60 words)
};
struct PackedThreads {
  uint32_t T;
  PackedThread t[kMaxWorkerThreads];
};

// compare_group: the circuit for the E elements of one thread; X, Y hold
// their three shares.  The runtime calls start at nonce offset `noff`; `s`
// schedules the element-wise calls and `sp` the packed one; `gw` is the
// first group index of the calling wave (the loop around it must be
// wave-uniform).  EQ / GT receive the masked boolean shares (GT only with CMP).
// pt (worker threads, else null): the packed call's thread table, the
// elements being those of a range that starts at packed word woff.
template <typename T, bool CMP>
__device__ inline void compare_group(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                     const NonceTriple &nb, uint64_t noff, uint64_t stride,
                                     const CallSched &s, const CallSched &sp, uint64_t gw,
                                     int lane, const T (&X)[16 / sizeof(T)][3],
                                     const T (&Y)[16 / sizeof(T)][3], T (&EQ)[16 / sizeof(T)][3],
                                     T (&GT)[16 / sizeof(T)][3],
                                     const PackedThreads *__restrict__ pt = nullptr,
                                     uint64_t woff = 0) {
  typedef typename SignedOf<T>::type ST;
  constexpr int E = 16 / sizeof(T);
  constexpr int B = 8 * sizeof(T);
  constexpr int L = same_levels<T>();
  // AES blocks of packed words per wave: a block packs 128 elements' bits and
  // a wave holds 64 E of them (for 16-byte words, half a block)
  constexpr int kBlkPerWave = 64 / B > 0 ? 64 / B : 1;
  const uint64_t nwords = sp.n;
  const uint64_t e0 = (gw + lane) * E;
  // lanes past the end run the same arithmetic on zeros; the caller stores nothing
  uint64_t off;
  uint32_t ctr, bib;
  sched_locate(s, e0, 0, off, ctr, bib);
  T e[E][3];
#pragma unroll
  for (int i = 0; i < E; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) e[i][k] = ~(X[i][k] ^ Y[i][k]);
#pragma unroll 1
  for (int l = 0; l < L; l++) {
    T S[3][E];
    pair_stream<T>(A, rk_pair, nb, noff + l * stride, off, ctr, S);
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

#pragma unroll
  for (int i = 0; i < E; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) EQ[i][k] = e[i][k] & 1;

  if (CMP) {
    // gt = parity(((e >> 1) ^ e) & x)
    T S[3][E];
    pair_stream<T>(A, rk_pair, nb, noff + L * stride, off, ctr, S);
    T r[E][3];
#pragma unroll
    for (int i = 0; i < E; i++) {
      T a[3], z[3], ks[3] = {S[0][i], S[1][i], S[2][i]};
#pragma unroll
      for (int k = 0; k < 3; k++) a[k] = (T)((ST)e[i][k] >> 1) ^ e[i][k];
      and3<T>(a, X[i], ks, z);
#pragma unroll
      for (int k = 0; k < 3; k++) r[i][k] = (T)(popc<T>(z[k]) & 1);
    }

    // keystream of the packed and_b: lane j < 3*kBlkPerWave computes key
    // j%3 of the wave's AES block j/3
    const uint64_t wave_e0 = gw * E;  // multiple of 64*E: of 128 unless E = 1
    uint4 kb = make_uint4(0, 0, 0, 0);
    if (lane < 3 * kBlkPerWave) {
      const int key = lane % 3, blk = lane / 3;
      const uint64_t w = wave_e0 / B + (uint64_t)blk * E;
      if (w < nwords) {
        uint64_t poff;
        uint32_t pctr, pbib;
        if (pt) {
          // the thread whose word range holds word woff + w (ranges are
          // whole batches, so a block's words share one thread)
          const uint64_t g = woff + w;
          uint32_t t = 0;
          while (t + 1 < pt->T && g >= pt->t[t].hi) t++;
          const PackedThread &P = pt->t[t];
          sched_locate(P.s, g - P.lo, 0, poff, pctr, pbib);
          kb = aes_ctr_block<kCopies>(A, P.rk + 60 * key, P.nonce[key] + poff, pctr);
        } else {
          sched_locate(sp, w, 0, poff, pctr, pbib);
          kb = aes_ctr_block<kCopies>(A, rk_pair + 60 * key,
                                    nb.v[key] + noff + (L + 1) * stride + poff, pctr);
        }
      }
    }
    // this lane's elements sit in dword D of the wave's packed keystream,
    // at bits b0 .. b0+E-1 (a wave of 16-byte words starts 0 or 64 elements
    // into its block)
    const int rel = (int)(wave_e0 % 128) + lane * E;
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
        s1[k] = (X[i][k] >> (B - 1)) & 1;
        s2[k] = (Y[i][k] >> (B - 1)) & 1;
        rr[k] = r[i][k] ^ s2[k];  // r ^= s2
        s2[k] ^= s1[k];           // s2 ^= s1
        a[k] = s2[k] ^ 1;         // s2 |= r  as  ~(~s2 & ~r)
        b[k] = rr[k] ^ 1;
        ks[k] = (kw[k] >> (b0 + i)) & 1;
      }
      and3<T>(a, b, ks, z);
#pragma unroll
      for (int k = 0; k < 3; k++) GT[i][k] = s1[k] ^ (z[k] ^ 1);  // s1 ^= s2
    }
  }
}

}  // namespace orqg
