// Device helpers shared by the kernels: share-group loads/stores, keystream
// unpacking and the group-PRG resharing masks.
#pragma once
#include <type_traits>

#include "internal.h"

namespace orqg {

// LDS replication factor of the AES tables for ordinary kernels.
constexpr int kCopies = 32;       // AES table replication (conflict-free LDS lookups)
constexpr int kAesThreads = 1024; // threads per AES workgroup (one workgroup per CU)
#define AES_LDS __shared__ __align__(16) uint32_t lds[AesLds<kCopies>::kWords]

struct NonceTriple {
  uint64_t v[3];
};

// 16-byte share words: quicksort's padded values for 64-bit keys
// (PadWidth __int128, sorting.h:409-416)
typedef unsigned __int128 u128;

// the signed type of a share word, for arithmetic shifts
template <typename T>
struct SignedOf {
  typedef typename std::make_signed<T>::type type;
};
template <>
struct SignedOf<u128> {
  typedef __int128 type;
};

// population count of a share word of any width
template <typename T>
__device__ inline int popc(T x) {
  int c = 0;
#pragma unroll
  for (unsigned s = 0; s < 8 * sizeof(T); s += 64) c += __popcll((unsigned long long)(x >> s));
  return c;
}

// Split a 16-byte keystream block into little-endian elements of type T.
template <typename T>
__device__ inline void unpack_block(uint4 b, T (&out)[16 / sizeof(T)]);
template <>
__device__ inline void unpack_block<uint32_t>(uint4 b, uint32_t (&out)[4]) {
  out[0] = b.x;
  out[1] = b.y;
  out[2] = b.z;
  out[3] = b.w;
}
template <>
__device__ inline void unpack_block<uint64_t>(uint4 b, uint64_t (&out)[2]) {
  out[0] = (uint64_t)b.x | ((uint64_t)b.y << 32);
  out[1] = (uint64_t)b.z | ((uint64_t)b.w << 32);
}
template <>
__device__ inline void unpack_block<u128>(uint4 b, u128 (&out)[1]) {
  out[0] = (u128)((uint64_t)b.x | ((uint64_t)b.y << 32)) |
           ((u128)((uint64_t)b.z | ((uint64_t)b.w << 32)) << 64);
}

// E = 16/sizeof(T) consecutive elements of an [n][3] share array.
template <typename T>
struct Grp3 {
  static constexpr int E = 16 / sizeof(T);
  T v[E][3];
  __device__ inline void load(const T *__restrict__ base, uint64_t e0, uint64_t n) {
    if (e0 + E <= n) {
      const uint4 *p = reinterpret_cast<const uint4 *>(base + 3 * e0);
      uint4 q[3] = {p[0], p[1], p[2]};
      const T *t = reinterpret_cast<const T *>(q);
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) v[i][k] = t[3 * i + k];
    } else {
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) v[i][k] = (e0 + i < n) ? base[3 * (e0 + i) + k] : (T)0;
    }
  }
  __device__ inline void store(T *__restrict__ base, uint64_t e0, uint64_t n) const {
    if (e0 + E <= n) {
      uint4 q[3];
      T *t = reinterpret_cast<T *>(q);
#pragma unroll
      for (int i = 0; i < E; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) t[3 * i + k] = v[i][k];
      uint4 *p = reinterpret_cast<uint4 *>(base + 3 * e0);
      p[0] = q[0];
      p[1] = q[1];
      p[2] = q[2];
    } else {
#pragma unroll
      for (int i = 0; i < E; i++)
        if (e0 + i < n)
#pragma unroll
          for (int k = 0; k < 3; k++) base[3 * (e0 + i) + k] = v[i][k];
    }
  }
};

// Resharing masks of one batch-call pair (zero_sharing_generator.h:159-183):
// rand[0] = call 0, rand[1] = call 1, rand[2] = rand[0]^rand[1] (binary) or
// -(rand[0]+rand[1]) (arithmetic).  m[k][i] masks share k of element e0+i.
template <typename T, bool ARITH, int COPIES>
__device__ inline void group_masks(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                   uint64_t nonce0, const CallSched &s, uint64_t e0,
                                   T (&m)[3][16 / sizeof(T)]) {
  constexpr int E = 16 / sizeof(T);
#pragma unroll
  for (int j = 0; j < 2; j++) {
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s, e0, j, off, ctr, bib);
    uint4 blk = aes_ctr_block<COPIES>(A, rk, nonce0 + off, ctr);
    unpack_block<T>(blk, m[j]);
  }
#pragma unroll
  for (int i = 0; i < E; i++) m[2][i] = ARITH ? (T)(0 - m[0][i] - m[1][i]) : (T)(m[0][i] ^ m[1][i]);
}

// Replicated_3PC::b2a_bit (replicated_3pc.h:220-251) per batch: two arithmetic
// zero sharings (calls 0 and 1), P0 adds x0^x1 (LSBs) to the first, P1 adds x2
// to the second, both are exchanged, d = first - second, y = d*d via
// multiply_a (call 2).  Party p's zero share of a call is R_{p-1} - R_p, R_k
// being pairwise key k's keystream, so d_p = D_{p-1} - D_p (+ the bits) with
// D_k = R_k(call 0) - R_k(call 1), and call 2 adds R2_{p-1} - R2_p to y_p.
// For 4-byte elements the blocks are computed key by key, the three calls of
// a key together: one key schedule live at a time, three independent blocks
// in flight.
template <typename T>
__device__ inline void b2a_bit_group(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                     const NonceTriple &nb, const CallSched &s, uint64_t e0,
                                     const T (&bits)[16 / sizeof(T)][3],
                                     T (&y)[16 / sizeof(T)][3]) {
  constexpr int E = 16 / sizeof(T);
  if constexpr (sizeof(T) == 4) {
    uint64_t off[3];
    uint32_t ctr[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      uint32_t bib;
      sched_locate(s, e0, j, off[j], ctr[j], bib);
    }
    T d[E][3], m[E][3];
#pragma unroll
    for (int i = 0; i < E; i++) {
      d[i][0] = (T)(bits[i][0] ^ bits[i][1]);
      d[i][1] = (T)(0 - bits[i][2]);
      d[i][2] = 0;
      m[i][0] = m[i][1] = m[i][2] = 0;
    }
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
      const uint32_t *rk = rk_pair + 60 * k;
      T R0[E], R1[E], R2[E];
      unpack_block<T>(aes_ctr_block<kCopies>(A, rk, nb.v[k] + off[0], ctr[0]), R0);
      unpack_block<T>(aes_ctr_block<kCopies>(A, rk, nb.v[k] + off[1], ctr[1]), R1);
      unpack_block<T>(aes_ctr_block<kCopies>(A, rk, nb.v[k] + off[2], ctr[2]), R2);
#pragma unroll
      for (int i = 0; i < E; i++) {
        const T D = (T)(R0[i] - R1[i]);
#pragma unroll
        for (int p = 0; p < 3; p++) {
          // key k is party k's "next" key (-) and party k+1's "prev" key (+)
          if (p == k) {
            d[i][p] -= D;
            m[i][p] -= R2[i];
          }
          if (p == (k + 1) % 3) {
            d[i][p] += D;
            m[i][p] += R2[i];
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const int q = (p + 1) % 3;
        y[i][p] = (T)(m[i][p] + d[i][p] * d[i][p] + d[i][p] * d[i][q] + d[i][q] * d[i][p]);
      }
  } else {
    // 8-byte elements (two per block): call by call, the three keys' blocks
    // together, which measured faster for them
    T d[E][3];
    // the three calls are consumed one after the other to bound register use
#pragma unroll 1
    for (int j = 0; j < 3; j++) {
      uint64_t off;
      uint32_t ctr, bib;
      sched_locate(s, e0, j, off, ctr, bib);
      T R[3][E];  // [key][elem]
#pragma unroll
      for (int k = 0; k < 3; k++) {
        uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + off, ctr);
        unpack_block<T>(blk, R[k]);
      }
#pragma unroll
      for (int i = 0; i < E; i++) {
        if (j == 0) {
#pragma unroll
          for (int p = 0; p < 3; p++) {
            T a = (T)(R[(p + 2) % 3][i] - R[p][i]);
            if (p == 0) a += bits[i][0] ^ bits[i][1];
            d[i][p] = a;
          }
        } else if (j == 1) {
#pragma unroll
          for (int p = 0; p < 3; p++) {
            T b = (T)(R[(p + 2) % 3][i] - R[p][i]);
            if (p == 1) b += bits[i][2];
            d[i][p] -= b;
          }
        } else {
#pragma unroll
          for (int p = 0; p < 3; p++) {
            int q = (p + 1) % 3, pm = (p + 2) % 3;
            y[i][p] = (T)(R[pm][i] - R[p][i]) + d[i][p] * d[i][p] + d[i][p] * d[i][q] +
                      d[i][q] * d[i][p];
          }
        }
      }
    }
  }
}

// outsider of group g ({0,1},{0,2},{1,2}) is party 2-g; it receives share o
// from o-1 and share o+1 from o+1 (protocol.h:108-209).
inline void log_reshare(orqg_ctx *c, int g, uint64_t bytes) {
  int o = 2 - g;
  c->log_msg("reshare", (o + 2) % 3, o, bytes);
  c->log_msg("reshare", (o + 1) % 3, o, bytes);
}

}  // namespace orqg
