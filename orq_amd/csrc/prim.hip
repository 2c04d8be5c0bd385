// Protocol primitives of Replicated_3PC for all three parties at once:
// keystream, secret sharing, open, local ops, and_b / multiply_a, b2a_bit and
// group resharing.  Every kernel works on groups of E = 16/sizeof(T)
// consecutive elements so that one AES block per PRG call covers the group
// (batches start at multiples of 256 elements and chunks at 1 MiB, so a group
// never straddles either).
#include <cstring>
#include <cmath>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

// ---- raw keystream (AESPRGAlgorithm::fillBytes) --------------------------------
template <int COPIES>
__global__ void __launch_bounds__(kAesThreads) k_keystream(const uint32_t *__restrict__ t0,
                                                   const uint32_t *__restrict__ rk,
                                                   uint64_t nonce0, uint64_t nbytes,
                                                   uint8_t *__restrict__ out) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  aes_load_tables<COPIES>(lds, t0);
  AesCtx<COPIES> A(lds);
  uint64_t nblk = (nbytes + 15) / 16;
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nblk;
       b += (uint64_t)gridDim.x * blockDim.x) {
    uint4 k = aes_ctr_block<COPIES>(A, rk, nonce0 + b / kBlocksPerChunk,
                                    2 + (uint32_t)(b % kBlocksPerChunk));
    if (b * 16 + 16 <= nbytes) {
      reinterpret_cast<uint4 *>(out)[b] = k;
    } else {
      uint32_t w[4] = {k.x, k.y, k.z, k.w};
      for (uint64_t i = b * 16; i < nbytes; i++) out[i] = (uint8_t)(w[(i & 15) / 4] >> (8 * (i & 3)));
    }
  }
}

void prg_keystream(orqg_ctx *c, int role, size_t bytes, void *out) {
  if (role < 0 || role >= ORQG_NUM_KEYS) throw Error(ORQG_E_ARG, "keystream: role");
  if (bytes == 0) return;
  uint64_t nblk = (bytes + 15) / 16;
  TimedRegion tr(c, "keystream");
  hipLaunchKernelGGL(k_keystream<kCopies>, dim3(grid_for(nblk, kAesThreads, c->num_cu)),
                     dim3(kAesThreads), 0,
                     c->stream, c->d_t0, c->rk(role), c->nonce[role], (uint64_t)bytes,
                     (uint8_t *)out);
  ORQG_HIP(hipGetLastError());
  c->nonce[role] += (bytes + kChunkBytes - 1) / kChunkBytes;
}

// ---- secret sharing from a data owner ---------------------------------------------
// get_shares_b/_a (replicated_3pc.h:338-392) and secret_share_b/_a (:401-448):
// share_1, share_2 from the owner's localPRG (two getNext calls per batch),
// share_3 = x ^ s1 ^ s2 (x - s1 - s2).
template <typename T, bool ARITH>
__global__ void __launch_bounds__(kAesThreads) k_secret_share(const uint32_t *__restrict__ t0,
                                                      const uint32_t *__restrict__ rk,
                                                      uint64_t nonce0, CallSched s,
                                                      const T *__restrict__ plain,
                                                      T *__restrict__ out, int owner) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e0 = g * E;
    T r[2][E];
    for (int j = 0; j < 2; j++) {
      uint64_t off;
      uint32_t ctr, bib;
      sched_locate(s, e0, j, off, ctr, bib);
      uint4 k = aes_ctr_block<kCopies>(A, rk, nonce0 + off, ctr);
      unpack_block<T>(k, r[j]);
    }
    for (int i = 0; i < E && e0 + i < s.n; i++) {
      T x = plain[e0 + i];
      T s2 = ARITH ? (T)(x - r[0][i] - r[1][i]) : (T)(x ^ r[0][i] ^ r[1][i]);
      // the owner keeps (r1, r2), its successor gets (r2, r3), its predecessor (r3, r1)
      T *o = out + 3 * (e0 + i);
      o[owner] = r[0][i];
      o[(owner + 1) % 3] = r[1][i];
      o[(owner + 2) % 3] = s2;
    }
  }
}

void secret_share(orqg_ctx *c, const void *host, orqg_vec *v, int owner) {
  size_t bytes = v->n * v->width;
  if (v->n == 0) return;
  void *tmp = c->scratch_get("plain_in", bytes);
  ORQG_HIP(hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, c->stream));
  int role = kLocal0 + owner;  // the data owner's local PRG
  bool arith = v->enc == ORQG_ASHARED;
  // each worker thread shares its range with its own keys
  for_threads(c, v->n, [&](uint64_t off, uint64_t m) {
    CallSched s = make_sched(m, c->batch_size, v->width, 2);
    unsigned grid = grid_for((m + 3) / 4, kAesThreads, c->num_cu);
    const char *in = (const char *)tmp + off * v->width;
    char *out = (char *)v->d + off * 3 * v->width;
    if (v->width == 4) {
      auto k = arith ? k_secret_share<uint32_t, true> : k_secret_share<uint32_t, false>;
      hipLaunchKernelGGL(k, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(role),
                         c->nonce[role], s, (const uint32_t *)in, (uint32_t *)out, owner);
    } else {
      auto k = arith ? k_secret_share<uint64_t, true> : k_secret_share<uint64_t, false>;
      hipLaunchKernelGGL(k, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(role),
                         c->nonce[role], s, (const uint64_t *)in, (uint64_t *)out, owner);
    }
    ORQG_HIP(hipGetLastError());
    c->nonce[role] += sched_nonces(s);
  });
  c->log_msg("secret_share", owner, (owner + 1) % 3, 2 * bytes);
  c->log_msg("secret_share", owner, (owner + 2) % 3, 2 * bytes);
}

// ---- open --------------------------------------------------------------------------
template <typename T, bool ARITH>
__global__ void k_open(uint64_t n, const T *__restrict__ v, T *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const T *s = v + 3 * i;
    out[i] = ARITH ? (T)(s[0] + s[1] + s[2]) : (T)(s[0] ^ s[1] ^ s[2]);
  }
}

void open_vec(orqg_ctx *c, const orqg_vec *v, void *host_out) {
  size_t bytes = v->n * v->width;
  if (v->n == 0) return;
  void *tmp = c->scratch_get("plain_out", bytes);
  unsigned grid = grid_for(v->n, 256, c->num_cu * 16);
  bool arith = v->enc == ORQG_ASHARED;
  if (v->width == 4)
    hipLaunchKernelGGL((arith ? k_open<uint32_t, true> : k_open<uint32_t, false>), dim3(grid),
                       dim3(256), 0, c->stream, (uint64_t)v->n, (const uint32_t *)v->d,
                       (uint32_t *)tmp);
  else if (v->width == 8)
    hipLaunchKernelGGL((arith ? k_open<uint64_t, true> : k_open<uint64_t, false>), dim3(grid),
                       dim3(256), 0, c->stream, (uint64_t)v->n, (const uint64_t *)v->d,
                       (uint64_t *)tmp);
  else  // 16-byte words (quicksort's padded 64-bit keys)
    hipLaunchKernelGGL((arith ? k_open<u128, true> : k_open<u128, false>), dim3(grid), dim3(256),
                       0, c->stream, (uint64_t)v->n, (const u128 *)v->d, (u128 *)tmp);
  ORQG_HIP(hipGetLastError());
  ORQG_HIP(hipMemcpyAsync(host_out, tmp, bytes, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, bytes);
}

// ---- local share-wise ops ------------------------------------------------------
template <typename T, int OP>
__global__ void k_local(uint64_t m, const T *__restrict__ x, const T *__restrict__ y,
                        T *__restrict__ z) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T a = x[i];
    T r;
    if (OP == kXor) r = a ^ y[i];
    else if (OP == kAdd) r = a + y[i];
    else if (OP == kSub) r = a - y[i];
    else r = ~a;  // not_b: every party complements both shares (replicated_3pc.h:199)
    z[i] = r;
  }
}

void elementwise_local(orqg_ctx *c, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  uint64_t m = 3 * (uint64_t)x->n;
  if (m == 0) return;
  unsigned grid = grid_for(m, 256, c->num_cu * 16);
#define ORQG_LOCAL_CASE(T)                                                                    \
  {                                                                                           \
    const T *xp = (const T *)x->d, *yp = y ? (const T *)y->d : xp;                             \
    T *zp = (T *)z->d;                                                                        \
    switch (op) {                                                                             \
      case kXor: hipLaunchKernelGGL((k_local<T, kXor>), dim3(grid), dim3(256), 0, c->stream, m, xp, yp, zp); break; \
      case kAdd: hipLaunchKernelGGL((k_local<T, kAdd>), dim3(grid), dim3(256), 0, c->stream, m, xp, yp, zp); break; \
      case kSub: hipLaunchKernelGGL((k_local<T, kSub>), dim3(grid), dim3(256), 0, c->stream, m, xp, yp, zp); break; \
      default: hipLaunchKernelGGL((k_local<T, kNot>), dim3(grid), dim3(256), 0, c->stream, m, xp, yp, zp); break; \
    }                                                                                         \
  }
  if (x->width == 4)
    ORQG_LOCAL_CASE(uint32_t)
  else
    ORQG_LOCAL_CASE(uint64_t)
#undef ORQG_LOCAL_CASE
  ORQG_HIP(hipGetLastError());
}

// ---- and_b / multiply_a -------------------------------------------------------------
// Party p: z_p = alpha_p (+|^) x_p y_p (+|^) x_p y_{p+1} (+|^) x_{p+1} y_p, with the
// zero sharing alpha_p = PRG_prev - PRG_next (prev ^ next for binary).  Party p's
// prev PRG is key K_{p-1} and its next PRG is K_p, so alpha_p = S_{p-1} op S_p where
// S_k is K_k's keystream for this batch.  z_p is sent to p-1, whose second share
// it becomes: the stored s_p.
template <typename T, bool ARITH>
__global__ void __launch_bounds__(kAesThreads) k_mult(const uint32_t *__restrict__ t0,
                                              const uint32_t *__restrict__ rk_pair,
                                              NonceTriple nb, CallSched s,
                                              const T *__restrict__ x, const T *__restrict__ y,
                                              T *__restrict__ z) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e0 = g * E;
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s, e0, 0, off, ctr, bib);
    T S[3][E];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + off, ctr);
      unpack_block<T>(blk, S[k]);
    }
    Grp3<T> X, Y, Z;
    X.load(x, e0, s.n);
    Y.load(y, e0, s.n);
#pragma unroll
    for (int i = 0; i < E; i++) {
#pragma unroll
      for (int p = 0; p < 3; p++) {
        int q = (p + 1) % 3, pm = (p + 2) % 3;
        T a0 = X.v[i][p], a1 = X.v[i][q], b0 = Y.v[i][p], b1 = Y.v[i][q];
        if (ARITH)
          Z.v[i][p] = (T)(S[pm][i] - S[p][i]) + a0 * b0 + a0 * b1 + a1 * b0;
        else
          Z.v[i][p] = (S[pm][i] ^ S[p][i]) ^ (a0 & b0) ^ (a0 & b1) ^ (a1 & b0);
      }
    }
    Z.store(z, e0, s.n);
  }
}

void mult(orqg_ctx *c, bool arith, const orqg_vec *x, const orqg_vec *y, orqg_vec *z) {
  if (x->n == 0) return;
  TimedRegion tr(c, "mult");
  // each worker thread's range with its own pair keys and batch schedule
  for_threads(c, x->n, [&](uint64_t off, uint64_t m) {
    const orqg_vec xs = vec_rows(x, off, m), ys = vec_rows(y, off, m), zs = vec_rows(z, off, m);
    CallSched s = make_sched(m, c->batch_size, x->width, 1);
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    unsigned grid = grid_for((m * x->width + 15) / 16, kAesThreads, c->num_cu);
    if (x->width == 4)
      hipLaunchKernelGGL((arith ? k_mult<uint32_t, true> : k_mult<uint32_t, false>), dim3(grid),
                         dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(kPair0), nb, s,
                         (const uint32_t *)xs.d, (const uint32_t *)ys.d, (uint32_t *)zs.d);
    else
      hipLaunchKernelGGL((arith ? k_mult<uint64_t, true> : k_mult<uint64_t, false>), dim3(grid),
                         dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(kPair0), nb, s,
                         (const uint64_t *)xs.d, (const uint64_t *)ys.d, (uint64_t *)zs.d);
    ORQG_HIP(hipGetLastError());
    uint64_t used = sched_nonces(s);
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  });
  const char *op = arith ? "mul_a" : "and_b";
  for (int p = 0; p < 3; p++) c->log_msg(op, p, (p + 2) % 3, x->n * (uint64_t)x->width);
}

// ---- b2a_bit (group helper b2a_bit_group lives in kern_common.h) ----
template <typename T>
__global__ void __launch_bounds__(kAesThreads) k_b2a_bit(const uint32_t *__restrict__ t0,
                                                 const uint32_t *__restrict__ rk_pair,
                                                 NonceTriple nb, CallSched s,
                                                 const T *__restrict__ x, T *__restrict__ y) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e0 = g * E;
    Grp3<T> X, Y;
    X.load(x, e0, s.n);
    T bits[E][3];
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) bits[i][k] = X.v[i][k] & 1;
    b2a_bit_group<T>(A, rk_pair, nb, s, e0, bits, Y.v);
    Y.store(y, e0, s.n);
  }
}

void b2a_bit(orqg_ctx *c, const orqg_vec *x, orqg_vec *y) {
  if (x->n == 0) return;
  for_threads(c, x->n, [&](uint64_t off, uint64_t m) {
    const orqg_vec xs = vec_rows(x, off, m), ys = vec_rows(y, off, m);
    CallSched s = make_sched(m, c->batch_size, x->width, 3);
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    unsigned grid = grid_for((m * x->width + 15) / 16, kAesThreads, c->num_cu);
    if (x->width == 4)
      hipLaunchKernelGGL(k_b2a_bit<uint32_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                         c->rk(kPair0), nb, s, (const uint32_t *)xs.d, (uint32_t *)ys.d);
    else
      hipLaunchKernelGGL(k_b2a_bit<uint64_t>, dim3(grid), dim3(kAesThreads), 0, c->stream, c->d_t0,
                         c->rk(kPair0), nb, s, (const uint64_t *)xs.d, (uint64_t *)ys.d);
    ORQG_HIP(hipGetLastError());
    uint64_t used = sched_nonces(s);
    for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
  });
  for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * x->n * (uint64_t)x->width);
}

// ---- reshare ------------------------------------------------------------------------
// Protocol::reshare (protocol.h:500-541): the two members of group g draw
// rand[0], rand[1] from the group PRG (two calls per batch), rand[2] is the
// dependent vector, and share k is masked with rand[k]; the outsider receives
// its two shares.
template <typename T, bool ARITH>
__global__ void __launch_bounds__(kAesThreads) k_reshare(const uint32_t *__restrict__ t0,
                                                 const uint32_t *__restrict__ rk, uint64_t nonce0,
                                                 CallSched s, T *__restrict__ v) {
  constexpr int E = 16 / sizeof(T);
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  uint64_t ngrp = (s.n + E - 1) / E;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngrp;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e0 = g * E;
    T m[3][E];
    group_masks<T, ARITH, kCopies>(A, rk, nonce0, s, e0, m);
    Grp3<T> V;
    V.load(v, e0, s.n);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) V.v[i][k] = ARITH ? (T)(V.v[i][k] + m[k][i]) : (T)(V.v[i][k] ^ m[k][i]);
    V.store(v, e0, s.n);
  }
}

void reshare(orqg_ctx *c, orqg_vec *v, int group) {
  if (group < 0 || group > 2) throw Error(ORQG_E_ARG, "reshare: group");
  if (v->n == 0) return;
  int role = kGroup0 + group;
  bool arith = v->enc == ORQG_ASHARED;
  for_threads(c, v->n, [&](uint64_t off, uint64_t m) {
    const orqg_vec vs = vec_rows(v, off, m);
    CallSched s = make_sched(m, c->batch_size, v->width, 2);
    unsigned grid = grid_for((m * v->width + 15) / 16, kAesThreads, c->num_cu);
    if (v->width == 4)
      hipLaunchKernelGGL((arith ? k_reshare<uint32_t, true> : k_reshare<uint32_t, false>), dim3(grid),
                         dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(role), c->nonce[role], s,
                         (uint32_t *)vs.d);
    else
      hipLaunchKernelGGL((arith ? k_reshare<uint64_t, true> : k_reshare<uint64_t, false>), dim3(grid),
                         dim3(kAesThreads), 0, c->stream, c->d_t0, c->rk(role), c->nonce[role], s,
                         (uint64_t *)vs.d);
    ORQG_HIP(hipGetLastError());
    c->nonce[role] += sched_nonces(s);
  });
  log_reshare(c, group, v->n * (uint64_t)v->width);
}

}  // namespace orqg
