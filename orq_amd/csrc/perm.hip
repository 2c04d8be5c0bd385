// Sharded and elementwise permutation application on the GPU
// (include/core/operators/shuffle.h:113-133, 250-270, 349-412, 687-706).
// The permutations themselves come from genperm.hip.
#include <algorithm>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

// ---- sharded-permutation rounds on API vectors -----------------------------------
// One round for group g: the members apply pi_g to both their shares
// (local_apply_perm: new[pi[i]] = x[i], i.e. new[j] = x[pi^{-1}[j]]; for the
// inverse application new[i] = x[pi[i]]) and then reshare: share k gets the
// group PRG mask rand[k] at its new position.
template <typename T, bool ARITH>
__global__ void __launch_bounds__(kAesThreads) k_round(const uint32_t *__restrict__ t0,
                                               const uint32_t *__restrict__ rk, uint64_t nonce0,
                                               CallSched s, const int32_t *__restrict__ src_idx,
                                               const T *__restrict__ in, T *__restrict__ out) {
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
#pragma unroll
    for (int i = 0; i < E; i++) {
      if (e0 + i < s.n) {
        const T *p = in + 3 * (uint64_t)src_idx[e0 + i];
#pragma unroll
        for (int k = 0; k < 3; k++) V.v[i][k] = ARITH ? (T)(p[k] + m[k][i]) : (T)(p[k] ^ m[k][i]);
      }
    }
    V.store(out, e0, s.n);
  }
}

// The reshare after local_apply_perm is a runtime call over the permuted
// vector: each worker thread masks its range of destination rows with its own
// group keys.
template <typename T>
void launch_round(orqg_ctx *c, bool arith, int g, const int32_t *idx, const T *in, T *out,
                  uint64_t n) {
  int role = kGroup0 + g;
  for_threads(c, n, [&](uint64_t off, uint64_t m) {
    CallSched s = make_sched(m, c->batch_size, sizeof(T), 2);
    unsigned grid = grid_for((m * sizeof(T) + 15) / 16, kAesThreads, c->num_cu);
    hipLaunchKernelGGL((arith ? k_round<T, true> : k_round<T, false>), dim3(grid), dim3(kAesThreads),
                       0, c->stream, c->d_t0, c->rk(role), c->nonce[role], s, idx + off, in,
                       out + 3 * off);
    ORQG_HIP(hipGetLastError());
    c->nonce[role] += sched_nonces(s);
  });
  log_reshare(c, g, n * sizeof(T));
}
template void launch_round<uint32_t>(orqg_ctx *, bool, int, const int32_t *, const uint32_t *,
                                     uint32_t *, uint64_t);
template void launch_round<uint64_t>(orqg_ctx *, bool, int, const int32_t *, const uint64_t *,
                                     uint64_t *, uint64_t);
template void launch_round<u128>(orqg_ctx *, bool, int, const int32_t *, const u128 *, u128 *,
                                 uint64_t);

// hm_oblivious_apply_sharded_perm / _inverse_ (shuffle.h:349-412)
void apply_sharded(orqg_ctx *c, orqg_vec *x, const ShardedPerm &pi, bool inverse) {
  if (x->n == 0) return;
  size_t bytes = 3 * x->n * (size_t)x->width;
  void *tmp = c->scratch_get("round_tmp", bytes);
  bool arith = x->enc == ORQG_ASHARED;
  void *cur = x->d, *nxt = tmp;
  for (int gi = 0; gi < 3; gi++) {
    int g = inverse ? 2 - gi : gi;
    const int32_t *idx = inverse ? pi.fwd[g] : pi.inv[g];
    if (x->width == 4)
      launch_round<uint32_t>(c, arith, g, idx, (const uint32_t *)cur, (uint32_t *)nxt, x->n);
    else if (x->width == 8)
      launch_round<uint64_t>(c, arith, g, idx, (const uint64_t *)cur, (uint64_t *)nxt, x->n);
    else
      launch_round<u128>(c, arith, g, idx, (const u128 *)cur, (u128 *)nxt, x->n);
    std::swap(cur, nxt);
  }
  // three rounds: result is in tmp
  ORQG_HIP(hipMemcpyAsync(x->d, cur, bytes, hipMemcpyDeviceToDevice, c->stream));
}

// open an A- or B-shared int32 permutation and scatter x by it
template <typename T>
__global__ void k_open_scatter(uint64_t n, bool arith, const uint32_t *__restrict__ perm,
                               const T *__restrict__ in, T *__restrict__ out,
                               uint32_t *__restrict__ err) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *p = perm + 3 * i;
    uint32_t sigma = arith ? p[0] + p[1] + p[2] : p[0] ^ p[1] ^ p[2];
    if (sigma >= n) {
      atomicOr(err, 1u);
      continue;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[3 * (uint64_t)sigma + k] = in[3 * i + k];
  }
}

// oblivious_apply_elementwise_perm (shuffle.h:687-706): one sharded permutation
// pi is applied to x and to a copy of perm, pi(perm) is opened, and pi(x) is
// scattered by it.
void apply_elementwise(orqg_ctx *c, orqg_vec *x, const orqg_vec *perm) {
  if (x->n == 0) return;
  ShardedPerm pi = next_perm(c, x->n);
  orqg_vec cp = *perm;
  size_t pbytes = 3 * perm->n * 4;
  cp.d = c->scratch_get("ew_perm_copy", pbytes);
  ORQG_HIP(hipMemcpyAsync(cp.d, perm->d, pbytes, hipMemcpyDeviceToDevice, c->stream));
  apply_sharded(c, x, pi, false);
  apply_sharded(c, &cp, pi, false);
  for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, perm->n * 4ull);
  size_t bytes = 3 * x->n * (size_t)x->width;
  void *tmp = c->scratch_get("ew_tmp", bytes);
  uint32_t *err = (uint32_t *)c->scratch_get("ew_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  unsigned grid = grid_for(x->n, 256, c->num_cu * 16);
  bool arith = perm->enc == ORQG_ASHARED;
  if (x->width == 4)
    hipLaunchKernelGGL(k_open_scatter<uint32_t>, dim3(grid), dim3(256), 0, c->stream,
                       (uint64_t)x->n, arith, (const uint32_t *)cp.d, (const uint32_t *)x->d,
                       (uint32_t *)tmp, err);
  else
    hipLaunchKernelGGL(k_open_scatter<uint64_t>, dim3(grid), dim3(256), 0, c->stream,
                       (uint64_t)x->n, arith, (const uint32_t *)cp.d, (const uint64_t *)x->d,
                       (uint64_t *)tmp, err);
  ORQG_HIP(hipGetLastError());
  ORQG_HIP(hipMemcpyAsync(x->d, tmp, bytes, hipMemcpyDeviceToDevice, c->stream));
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  free_perm(pi);
  if (herr) throw Error(ORQG_E_PARITY, "apply_elementwise: opened permutation out of range");
}

}  // namespace orqg
