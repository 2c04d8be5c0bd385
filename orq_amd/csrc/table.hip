// ElementwisePermutation operations, permutation composition, shuffles and the
// table sort, composed from the sharded-permutation rounds, the elementwise
// application and the radix sort.  Each function follows the reference's
// sequence of protocol calls, so it consumes the same sharded permutations,
// PRG nonces and messages in the same order.
#include <vector>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

// ---- public sharings and opened permutations ------------------------------------
// public_share (replicated_3pc.h:456-470): share s_0 = value, s_1 = s_2 = 0.
__global__ void k_public_iota(uint64_t n, uint32_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    out[3 * i] = (uint32_t)i;
    out[3 * i + 1] = 0;
    out[3 * i + 2] = 0;
  }
}
__global__ void k_public_fill(uint64_t n, uint32_t value, uint32_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    out[3 * i] = value;
    out[3 * i + 1] = 0;
    out[3 * i + 2] = 0;
  }
}
// x += public c: only s_0 changes
__global__ void k_add_public(uint64_t n, uint32_t value, uint32_t *__restrict__ x) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[3 * i] += value;
}
// open an int32 sharing and write it back as a public arithmetic sharing
__global__ void k_open_public(uint64_t n, bool arith_in, const uint32_t *__restrict__ in,
                              uint32_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *s = in + 3 * i;
    uint32_t o = arith_in ? s[0] + s[1] + s[2] : s[0] ^ s[1] ^ s[2];
    out[3 * i] = o;
    out[3 * i + 1] = 0;
    out[3 * i + 2] = 0;
  }
}
// open the permutation `perm` and gather x through it:
// local_apply_inverse_perm (shuffle.h:250-270), out[i] = x[perm[i]]
template <typename T>
__global__ void k_open_gather(uint64_t n, bool arith, const uint32_t *__restrict__ perm,
                              const T *__restrict__ in, T *__restrict__ out,
                              uint32_t *__restrict__ err) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *p = perm + 3 * i;
    uint32_t src = arith ? p[0] + p[1] + p[2] : p[0] ^ p[1] ^ p[2];
    if (src >= n) {
      atomicOr(err, 1u);
      continue;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[3 * i + k] = in[3 * (uint64_t)src + k];
  }
}

template <typename T>
__global__ void k_public_fill_t(uint64_t n, T value, T *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    out[3 * i] = value;
    out[3 * i + 1] = 0;
    out[3 * i + 2] = 0;
  }
}

static unsigned flat_grid(orqg_ctx *c, uint64_t n) { return grid_for(n, 256, c->num_cu * 16); }

// m rows of a public sharing of `value` (replicated_3pc.h:456-470) at d
void public_fill_raw(orqg_ctx *c, void *d, uint64_t m, uint64_t value, int w) {
  if (m == 0) return;
  if (w == 4)
    hipLaunchKernelGGL(k_public_fill_t<uint32_t>, dim3(flat_grid(c, m)), dim3(256), 0, c->stream,
                       m, (uint32_t)value, (uint32_t *)d);
  else
    hipLaunchKernelGGL(k_public_fill_t<uint64_t>, dim3(flat_grid(c, m)), dim3(256), 0, c->stream,
                       m, (uint64_t)value, (uint64_t *)d);
  ORQG_HIP(hipGetLastError());
}
void public_fill(orqg_ctx *c, orqg_vec *v, uint64_t value) {
  public_fill_raw(c, v->d, v->n, value, v->width);
}

template <typename T>
__global__ void k_broadcast_row(uint64_t count, const T *__restrict__ src, T *__restrict__ dst) {
  const T s0 = src[0], s1 = src[1], s2 = src[2];
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    dst[3 * i] = s0;
    dst[3 * i + 1] = s1;
    dst[3 * i + 2] = s2;
  }
}
void broadcast_row(orqg_ctx *c, const orqg_vec *src, uint64_t row, orqg_vec *dst,
                   uint64_t dst_start, uint64_t count) {
  if (count == 0) return;
  if (src->width == 4)
    hipLaunchKernelGGL(k_broadcast_row<uint32_t>, dim3(flat_grid(c, count)), dim3(256), 0,
                       c->stream, count, (const uint32_t *)src->d + 3 * row,
                       (uint32_t *)dst->d + 3 * dst_start);
  else
    hipLaunchKernelGGL(k_broadcast_row<uint64_t>, dim3(flat_grid(c, count)), dim3(256), 0,
                       c->stream, count, (const uint64_t *)src->d + 3 * row,
                       (uint64_t *)dst->d + 3 * dst_start);
  ORQG_HIP(hipGetLastError());
}

// ElementwisePermutation(size, encoding) (permutation.h:62-69): a public
// sharing of the identity.
void perm_identity(orqg_ctx *c, orqg_vec *perm, int enc) {
  if (perm->width != 4) throw Error(ORQG_E_ARG, "permutation: width must be 4");
  perm->enc = enc;
  if (perm->n == 0) return;
  hipLaunchKernelGGL(k_public_iota, dim3(flat_grid(c, perm->n)), dim3(256), 0, c->stream,
                     (uint64_t)perm->n, (uint32_t *)perm->d);
  ORQG_HIP(hipGetLastError());
}

// ElementwisePermutation::b2a for 3PC (permutation.h:222-258): pi on a copy,
// open, public arithmetic sharing, inverse pi.
void perm_b2a(orqg_ctx *c, orqg_vec *perm) {
  const uint64_t n = perm->n;
  if (perm->enc != ORQG_BSHARED) throw Error(ORQG_E_ARG, "perm_b2a: permutation is not B-shared");
  if (n == 0) {
    perm->enc = ORQG_ASHARED;
    return;
  }
  ShardedPerm pi = next_perm(c, n);
  orqg_vec cp = *perm;
  cp.d = c->scratch_get("pb2a_copy", n * 12 + 64);
  ORQG_HIP(hipMemcpyAsync(cp.d, perm->d, n * 12, hipMemcpyDeviceToDevice, c->stream));
  apply_sharded(c, &cp, pi, false);
  for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, n * 4);
  hipLaunchKernelGGL(k_open_public, dim3(flat_grid(c, n)), dim3(256), 0, c->stream, n, false,
                     (const uint32_t *)cp.d, (uint32_t *)perm->d);
  ORQG_HIP(hipGetLastError());
  perm->enc = ORQG_ASHARED;
  apply_sharded(c, perm, pi, true);
  free_perm(pi);
}

// x = x * public(a) + public(b), the secure multiply_a consuming its zero
// sharing as ElementwisePermutation::negate / reverse do
// (permutation.h:177-198,268-280).
static void perm_affine(orqg_ctx *c, orqg_vec *perm, uint32_t a, uint32_t b) {
  const uint64_t n = perm->n;
  if (perm->enc != ORQG_ASHARED) throw Error(ORQG_E_ARG, "permutation is not A-shared");
  if (n == 0) return;
  orqg_vec m = *perm;
  m.d = c->scratch_get("perm_const", n * 12 + 64);
  hipLaunchKernelGGL(k_public_fill, dim3(flat_grid(c, n)), dim3(256), 0, c->stream, n, a,
                     (uint32_t *)m.d);
  mult(c, true, perm, &m, perm);
  hipLaunchKernelGGL(k_add_public, dim3(flat_grid(c, n)), dim3(256), 0, c->stream, n, b,
                     (uint32_t *)perm->d);
  ORQG_HIP(hipGetLastError());
}
void perm_negate(orqg_ctx *c, orqg_vec *perm) { perm_affine(c, perm, 0xffffffffu, 0xffffffffu); }
void perm_reverse(orqg_ctx *c, orqg_vec *perm) {
  perm_affine(c, perm, 0xffffffffu, (uint32_t)(perm->n - 1));
}

// ElementwisePermutation::invert (permutation.h:205-212): apply the
// permutation to a sharing of the identity with the same encoding.
void perm_invert(orqg_ctx *c, orqg_vec *perm) {
  const uint64_t n = perm->n;
  if (n == 0) return;
  orqg_vec inv = *perm;
  inv.d = c->scratch_get("perm_inv", n * 12 + 64);
  perm_identity(c, &inv, perm->enc);
  apply_elementwise(c, &inv, perm);
  ORQG_HIP(hipMemcpyAsync(perm->d, inv.d, n * 12, hipMemcpyDeviceToDevice, c->stream));
}

// compose_permutations (shuffle.h:716-743): rho <- rho o sigma; sigma is left
// holding pi(sigma), as in the reference.
void perm_compose(orqg_ctx *c, orqg_vec *sigma, orqg_vec *rho) {
  const uint64_t n = sigma->n;
  if (rho->n != n || sigma->width != 4 || rho->width != 4)
    throw Error(ORQG_E_ARG, "compose: permutations differ in size");
  if (sigma->enc != rho->enc) throw Error(ORQG_E_ARG, "compose: encodings differ");
  if (n == 0) return;
  ShardedPerm pi = next_perm(c, n);
  apply_sharded(c, sigma, pi, false);
  for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, n * 4);
  uint32_t *tmp = (uint32_t *)c->scratch_get("compose_tmp", n * 12 + 64);
  uint32_t *err = (uint32_t *)c->scratch_get("compose_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  hipLaunchKernelGGL(k_open_gather<uint32_t>, dim3(flat_grid(c, n)), dim3(256), 0, c->stream, n,
                     sigma->enc == ORQG_ASHARED, (const uint32_t *)sigma->d,
                     (const uint32_t *)rho->d, tmp, err);
  ORQG_HIP(hipGetLastError());
  ORQG_HIP(hipMemcpyAsync(rho->d, tmp, n * 12, hipMemcpyDeviceToDevice, c->stream));
  apply_sharded(c, rho, pi, true);
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  free_perm(pi);
  if (herr) throw Error(ORQG_E_PARITY, "compose: opened permutation out of range");
}

// operators::shuffle of a table (shuffle.h:768-786): a B-shared identity
// permutation is shuffled by one sharded permutation, then applied
// elementwise to every arithmetic column and then every binary column.
void table_shuffle(orqg_ctx *c, orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols,
                   size_t nb) {
  orqg_vec *first = na ? a_cols[0] : nb ? b_cols[0] : nullptr;
  if (!first) return;
  const uint64_t n = first->n;
  orqg_vec perm;
  perm.ctx = c;
  perm.n = n;
  perm.width = 4;
  perm.d = dev_alloc(c, n * 12 + 64);
  try {
    perm_identity(c, &perm, ORQG_BSHARED);
    if (n > 0) {
      ShardedPerm pi = next_perm(c, n);
      apply_sharded(c, &perm, pi, false);
      free_perm(pi);
    }
    for (size_t i = 0; i < na; i++) apply_elementwise(c, a_cols[i], &perm);
    for (size_t i = 0; i < nb; i++) apply_elementwise(c, b_cols[i], &perm);
  } catch (...) {
    ORQG_HIP(hipStreamSynchronize(c->stream));
    dev_free(perm.d);
    throw;
  }
  ORQG_HIP(hipStreamSynchronize(c->stream));
  dev_free(perm.d);
}

// operators::table_sort with SortingProtocol::RADIXSORT (sorting.h:509-612).
void table_sort(orqg_ctx *c, orqg_vec *const *keys, size_t nkeys, const int *orders,
                const int *single_bit, orqg_vec *const *a_cols, size_t na,
                orqg_vec *const *b_cols, size_t nb, int protocol) {
  if (nkeys == 0) throw Error(ORQG_E_ARG, "table_sort: no sort keys");
  if (protocol != ORQG_RADIXSORT && protocol != ORQG_QUICKSORT)
    throw Error(ORQG_E_ARG, "table_sort: unknown protocol");
  const uint64_t n = keys[0]->n;
  int ns = 0;
  for (size_t j = 0; j < nkeys; j++) ns += single_bit && single_bit[j] ? 1 : 0;
  const int nk = (int)nkeys - ns;
  const int nc = (int)(na + nb);
  const int L = 8 * keys[0]->width;
  // sorting.h:531-550 for 3PC (pairs are single permutations): quicksort
  // needs one more permutation per multi-bit key (its shuffle), radix sort
  // one more pair per key bit
  long perms_required = nk + ns - 1 + nk + ns;
  long pairs_required = 4L * ns + 3L * nk + nc - 1;
  if (protocol == ORQG_QUICKSORT)
    perms_required += nk;
  else
    pairs_required += (long)nk * L;
  if (n == 0) return;
  reserve_perms(c, n, (size_t)(perms_required + pairs_required));

  std::vector<void *> owned;
  auto make = [&](size_t width) {
    orqg_vec v;
    v.ctx = c;
    v.n = n;
    v.width = (int)width;
    v.d = dev_alloc(c, 3 * n * width + 64);
    owned.push_back(v.d);
    return v;
  };
  try {
    auto sort_sub = [&](int col, orqg_vec *perm_out) {
      const int ord = orders ? orders[col] : ORQG_ASC;
      if (single_bit && single_bit[col])
        radix_sort(c, keys[col], ord, 1, perm_out);
      else if (protocol == ORQG_QUICKSORT)
        quicksort(c, keys[col], ord, perm_out);
      else
        radix_sort(c, keys[col], ord, 8 * keys[col]->width, perm_out);
    };
    int C = (int)nkeys - 1;
    orqg_vec sort_perm = make(4), next = make(4);
    orqg_vec orig = make(keys[C]->width);
    ORQG_HIP(hipMemcpyAsync(orig.d, keys[C]->d, 3 * n * keys[C]->width, hipMemcpyDeviceToDevice,
                            c->stream));
    sort_sub(C, &sort_perm);
    ORQG_HIP(hipMemcpyAsync(keys[C]->d, orig.d, 3 * n * keys[C]->width,
                            hipMemcpyDeviceToDevice, c->stream));
    for (C--; C >= 0; C--) {
      orqg_vec o2 = make(keys[C]->width);
      ORQG_HIP(hipMemcpyAsync(o2.d, keys[C]->d, 3 * n * keys[C]->width, hipMemcpyDeviceToDevice,
                              c->stream));
      apply_elementwise(c, keys[C], &sort_perm);
      sort_sub(C, &next);
      ORQG_HIP(hipMemcpyAsync(keys[C]->d, o2.d, 3 * n * keys[C]->width,
                              hipMemcpyDeviceToDevice, c->stream));
      // sort_permutation = compose(sort_permutation, next_permutation)
      perm_compose(c, &sort_perm, &next);
      std::swap(sort_perm.d, next.d);
      sort_perm.enc = next.enc;
    }
    for (size_t j = 0; j < nkeys; j++) apply_elementwise(c, keys[j], &sort_perm);
    for (size_t i = 0; i < na; i++) apply_elementwise(c, a_cols[i], &sort_perm);
    for (size_t i = 0; i < nb; i++) apply_elementwise(c, b_cols[i], &sort_perm);
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    for (void *p : owned) dev_free(p);
    throw;
  }
  ORQG_HIP(hipStreamSynchronize(c->stream));
  for (void *p : owned) dev_free(p);
}

}  // namespace orqg
