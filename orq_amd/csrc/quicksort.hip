// operators::quicksort (include/core/operators/quicksort.h:212-244) for
// 32-bit keys, the reference's default sort protocol (sorting.h:5-22).
//
// The keys are padded to 64 bits with their index in the low half
// (pad_input, sorting.h:431-462), so every padded value is unique, and
// shuffled with one sharded permutation (shuffle.h:752-757).  Then, round by
// round (quicksort_body, quicksort.h:62-198): every element that is not yet a
// pivot is compared with the pivot of its segment, the last marked position
// before it; the comparison bits are opened; and each segment is
// partitioned around its first element (Lomuto, quicksort.h:28-50), marking
// the pivot's final position and the head of the right part.  The sort ends
// when every position is marked.  Finally the padding is removed, which
// gives the sorting permutation (remove_padding, sorting.h:474-495), and the
// permutation is inverted.
//
// What runs where: the comparisons (the only secret-dependent arithmetic:
// _compare, PRG draws and messages as the reference makes them) and every
// gather of shares run on the device.  The pivot bookkeeping and the
// partition act on opened, public bits, so the host does them: they are a
// sequential walk per segment whose outcome (a row permutation) the device
// then applies to all three shares in one gather.
#include <algorithm>
#include <vector>

#include "internal.h"

namespace orqg {
namespace {

// pad_input for int32 keys: padded share k = share k << 32 (the sign
// extension of the copy-cast is shifted out), and the public index -- i, or
// -1 - i for DESC -- ORed into the low half of s_0 (public_share holds a
// public value in s_0 only).  DESC reverses the rows afterwards.
__global__ void k_qs_pad(uint64_t n, int desc, const uint32_t *__restrict__ v,
                         uint64_t *__restrict__ pad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t idx = desc ? (uint32_t)(-1 - (int64_t)i) : (uint32_t)i;
    const uint64_t r = desc ? n - 1 - i : i;
    pad[3 * r] = ((uint64_t)v[3 * i] << 32) | idx;
    pad[3 * r + 1] = (uint64_t)v[3 * i + 1] << 32;
    pad[3 * r + 2] = (uint64_t)v[3 * i + 2] << 32;
  }
}

// remove_padding after the (DESC) reversal: low halves -> the B-shared
// permutation, high halves (padded >> 32, copy-cast to 32 bits) -> v
__global__ void k_qs_unpad(uint64_t n, int desc, const uint64_t *__restrict__ pad,
                           uint32_t *__restrict__ v, uint32_t *__restrict__ perm) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t src = desc ? n - 1 - r : r;
    for (int k = 0; k < 3; k++) {
      const uint64_t x = pad[3 * src + k];
      perm[3 * r + k] = (uint32_t)x;
      v[3 * r + k] = (uint32_t)(x >> 32);
    }
  }
}

// dst row i = src row idx[i], all three shares
__global__ void k_qs_gather(uint64_t m, const uint32_t *__restrict__ idx,
                            const uint64_t *__restrict__ src, uint64_t *__restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = idx[i];
    dst[3 * i] = src[3 * j];
    dst[3 * i + 1] = src[3 * j + 1];
    dst[3 * i + 2] = src[3 * j + 2];
  }
}

orqg_vec view(orqg_ctx *c, void *d, uint64_t n, int width, int enc) {
  orqg_vec v;
  v.ctx = c;
  v.n = n;
  v.width = width;
  v.enc = enc;
  v.d = d;
  return v;
}

// partition (quicksort.h:28-50) of rows [start, end] around the row at
// start, applied to the row map `src` (src[pos] = row now at pos)
void partition(std::vector<uint32_t> &src, const std::vector<uint8_t> &cmp, int64_t start,
               int64_t end, std::vector<int32_t> &pivots) {
  int64_t i = start;
  for (int64_t j = start + 1; j <= end; j++)
    if (cmp[j] == 0) {
      i++;
      std::swap(src[j], src[i]);
    }
  std::swap(src[start], src[i]);
  pivots[i] = (int32_t)i;
  if (i + 1 <= end) pivots[i + 1] = (int32_t)(i + 1);
}

// quicksort_body on the padded, shuffled vector P (n rows of 3 x uint64);
// returns the buffer that holds the result (P or its twin Q)
uint64_t *quicksort_rounds(orqg_ctx *c, uint64_t *P, uint64_t *Q, uint64_t n) {
  std::vector<int32_t> pivots(n, -1);
  pivots[0] = 0;
  std::vector<uint8_t> cmp(n, 0);  // exp_cmp_plaintext: the last opened bit of each row
  std::vector<uint32_t> idx, src(n);
  std::vector<uint64_t> opened;
  uint32_t *d_idx = (uint32_t *)c->scratch_get("qs_idx", 8 * n + 64);
  uint32_t *d_src = (uint32_t *)c->scratch_get("qs_src", 4 * n + 64);
  uint64_t *red = (uint64_t *)c->scratch_get("qs_red", 24 * n + 64);
  uint64_t *piv = (uint64_t *)c->scratch_get("qs_piv", 24 * n + 64);
  uint64_t *eq = (uint64_t *)c->scratch_get("qs_eq", 24 * n + 64);
  uint64_t *gt = (uint64_t *)c->scratch_get("qs_gt", 24 * n + 64);
  uint64_t num_pivots = 1;
  for (;;) {
    // the non-pivot rows, then the pivot of each (prefix max of pivots)
    const uint64_t size = n - num_pivots;
    idx.resize(2 * size);
    uint64_t k = 0;
    int32_t run = -1;
    for (uint64_t j = 0; j < n; j++) {
      run = std::max(run, pivots[j]);
      if (pivots[j] < 0) {
        idx[k] = (uint32_t)j;
        idx[size + k] = (uint32_t)run;
        k++;
      }
    }
    if (size) {
      ORQG_HIP(hipMemcpyAsync(d_idx, idx.data(), 8 * size, hipMemcpyHostToDevice, c->stream));
      const unsigned g = grid_for(size, 256, c->num_cu * 16);
      hipLaunchKernelGGL(k_qs_gather, dim3(g), dim3(256), 0, c->stream, size, d_idx, P, red);
      hipLaunchKernelGGL(k_qs_gather, dim3(g), dim3(256), 0, c->stream, size, d_idx + size, P,
                         piv);
      ORQG_HIP(hipGetLastError());
      // reduced_vec._compare(pivot_vec, temp, comparisons): gt = reduced > pivot
      orqg_vec vr = view(c, red, size, 8, ORQG_BSHARED), vp = view(c, piv, size, 8, ORQG_BSHARED);
      orqg_vec ve = view(c, eq, size, 8, ORQG_BSHARED), vg = view(c, gt, size, 8, ORQG_BSHARED);
      compare(c, &vr, &vp, &ve, &vg, true, 0, 0);
      opened.resize(size);
      open_vec(c, &vg, opened.data());  // comparisons.open()
      for (uint64_t t = 0; t < size; t++) cmp[idx[t]] = (uint8_t)(opened[t] & 1);
    }
    // partition every segment between marked positions (quicksort.h:149-165)
    for (uint64_t j = 0; j < n; j++) src[j] = (uint32_t)j;
    int64_t index = -1;
    for (int64_t p = 0; p < (int64_t)n; p++) {
      if (pivots[p] < 0) continue;
      if (p - index == 1) {  // adjacent pivots
        index++;
        continue;
      }
      partition(src, cmp, index, p - 1, pivots);
      index = p;
    }
    partition(src, cmp, index, (int64_t)n - 1, pivots);
    // (cmp stays indexed by position, as in the reference: a partition reads
    // cmp[j] before any swap has touched position j, and the next round
    // rewrites every non-pivot position before reading it)
    ORQG_HIP(hipMemcpyAsync(d_src, src.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_qs_gather, dim3(grid_for(n, 256, c->num_cu * 16)), dim3(256), 0,
                       c->stream, n, d_src, P, Q);
    ORQG_HIP(hipGetLastError());
    std::swap(P, Q);
    // the host vectors above are reused next round: the copies out of them
    // must have been consumed
    ORQG_HIP(hipStreamSynchronize(c->stream));
    num_pivots = 0;
    for (int32_t x : pivots) num_pivots += x >= 0;
    if (num_pivots >= n) return P;
  }
}

}  // namespace

void quicksort(orqg_ctx *c, orqg_vec *v, int order, orqg_vec *perm_out) {
  require_one_thread(c, "quicksort");
  if (!v || !perm_out || v->width != 4 || perm_out->width != 4 || perm_out->n != v->n)
    throw Error(ORQG_E_ARG, "quicksort: 32-bit keys and a 32-bit permutation of the same size");
  const uint64_t n = v->n;
  if (n >= (1ull << 31)) throw Error(ORQG_E_ARG, "quicksort: n must be < 2^31");
  const bool desc = order == ORQG_DESC;
  // 1 permutation for the shuffle, 1 for remove_padding's b2a, 1 pair for
  // the inversion; pairs are single permutations in 3PC (quicksort.h:215-224)
  reserve_perms(c, n, 3);
  if (n == 0) return;
  TimedRegion tr(c, "quicksort");
  uint64_t *P = (uint64_t *)c->scratch_get("qs_P", 24 * n + 64);
  uint64_t *Q = (uint64_t *)c->scratch_get("qs_Q", 24 * n + 64);
  const unsigned grid = grid_for(n, 256, c->num_cu * 16);
  hipLaunchKernelGGL(k_qs_pad, dim3(grid), dim3(256), 0, c->stream, n, (int)desc,
                     (const uint32_t *)v->d, P);
  ORQG_HIP(hipGetLastError());
  {  // v.shuffle()
    orqg_vec pv = view(c, P, n, 8, ORQG_BSHARED);
    ShardedPerm pi = next_perm(c, n);
    try {
      apply_sharded(c, &pv, pi, false);
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      free_perm(pi);
      throw;
    }
    ORQG_HIP(hipStreamSynchronize(c->stream));
    free_perm(pi);
  }
  P = quicksort_rounds(c, P, Q, n);
  hipLaunchKernelGGL(k_qs_unpad, dim3(grid), dim3(256), 0, c->stream, n, (int)desc, P,
                     (uint32_t *)v->d, (uint32_t *)perm_out->d);
  ORQG_HIP(hipGetLastError());
  perm_out->enc = ORQG_BSHARED;
  perm_b2a(c, perm_out);
  if (desc) perm_negate(c, perm_out);
  perm_invert(c, perm_out);
  ORQG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace orqg
