// operators::quicksort (include/core/operators/quicksort.h:212-244) for
// 32- and 64-bit keys, the reference's default sort protocol (sorting.h:5-22).
//
// The keys are padded to PadWidth -- 64 bits for 32-bit keys, 128 for 64-bit
// keys (sorting.h:409-416) -- with their index in the low 32 bits
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

#include <thread>
#include <utility>

#include "internal.h"
#include "kern_common.h"

namespace orqg {
namespace {

// pad_input: padded share k = the sign-extending copy-cast of key share k to
// P (bit-sliced sign extension, as the reference's copy-cast of each share),
// shifted left by 32, and the public index -- i, or -1 - i for DESC -- ORed
// into the low 32 bits of s_0 (public_share holds a public value in s_0
// only).  DESC reverses the rows afterwards.
template <typename K, typename P>
__global__ void k_qs_pad(uint64_t n, int desc, const K *__restrict__ v, P *__restrict__ pad) {
  typedef typename SignedOf<K>::type SK;
  typedef typename SignedOf<P>::type SP;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t idx = desc ? (uint32_t)(-1 - (int64_t)i) : (uint32_t)i;
    const uint64_t r = desc ? n - 1 - i : i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const P x = (P)(SP)(SK)v[3 * i + k] << 32;
      pad[3 * r + k] = k == 0 ? (P)(x | idx) : x;
    }
  }
}

// remove_padding after the (DESC) reversal: the low 32 bits -> the B-shared
// permutation, padded >> 32 copy-cast to the key width -> v
template <typename K, typename P>
__global__ void k_qs_unpad(uint64_t n, int desc, const P *__restrict__ pad, K *__restrict__ v,
                           uint32_t *__restrict__ perm) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t src = desc ? n - 1 - r : r;
    for (int k = 0; k < 3; k++) {
      const P x = pad[3 * src + k];
      perm[3 * r + k] = (uint32_t)x;
      v[3 * r + k] = (K)(x >> 32);
    }
  }
}

// dst row i = src row idx[i], all three shares
template <typename P>
__global__ void k_qs_gather(uint64_t m, const uint32_t *__restrict__ idx,
                            const P *__restrict__ src, P *__restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = idx[i];
    dst[3 * i] = src[3 * j];
    dst[3 * i + 1] = src[3 * j + 1];
    dst[3 * i + 2] = src[3 * j + 2];
  }
}

// bit 0 of the opened B-shared rows (s0 ^ s1 ^ s2), one byte per row
template <typename P>
__global__ void k_qs_open_bit(uint64_t m, const P *__restrict__ v, uint8_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint8_t)((v[3 * i] ^ v[3 * i + 1] ^ v[3 * i + 2]) & 1u);
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

constexpr int kHostThreads = 8;  // host threads for the partitions of one round

// partition (quicksort.h:28-50) of rows [start, end] around the row at
// start, applied to the row map `src` (src[pos] = row now at pos)
uint64_t partition(uint32_t *src, const std::vector<uint8_t> &cmp, int64_t start, int64_t end,
                   std::vector<int32_t> &pivots) {
  int64_t i = start;
  for (int64_t j = start + 1; j <= end; j++)
    if (cmp[j] == 0) {
      i++;
      std::swap(src[j], src[i]);
    }
  std::swap(src[start], src[i]);
  // position start is marked already (it is the segment's pivot position);
  // i and i + 1 are new marks unless i is start
  uint64_t marked = i != start ? 1 : 0;
  pivots[i] = (int32_t)i;
  if (i + 1 <= end) {
    pivots[i + 1] = (int32_t)(i + 1);
    marked++;
  }
  return marked;
}

// f(lo, hi) over [0, m) split into up to kHostThreads ranges on host threads
// (one range, on the caller, below 2^16)
template <typename F>
void host_ranges(uint64_t m, F &&f) {
  const int T = m >= (1u << 16) ? kHostThreads : 1;
  if (T == 1) {
    f((uint64_t)0, m);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([&, t] { f(m * t / T, m * (t + 1) / T); });
  for (auto &x : th) x.join();
}

// quicksort_body on the padded, shuffled vector P (n rows of 3 words of type
// W); returns the buffer that holds the result (P or its twin Q)
template <typename W>
W *quicksort_rounds(orqg_ctx *c, W *P, W *Q, uint64_t n) {
  constexpr int kW = (int)sizeof(W);
  std::vector<int32_t> pivots(n, -1);
  pivots[0] = 0;
  std::vector<uint8_t> cmp(n, 0);  // exp_cmp_plaintext: the last opened bit of each row
  // host staging in pinned memory: idx (2n words), src (n words), the opened
  // comparison bits (n bytes)
  const size_t pin_bytes = 13 * (size_t)n + 64;
  if (c->host_pinned_bytes < pin_bytes) {
    if (c->host_pinned) ORQG_HIP(hipHostFree(c->host_pinned));
    c->host_pinned = nullptr;
    c->host_pinned_bytes = 0;
    ORQG_HIP(hipHostMalloc(&c->host_pinned, pin_bytes, hipHostMallocDefault));
    c->host_pinned_bytes = pin_bytes;
  }
  uint32_t *idx = (uint32_t *)c->host_pinned;
  uint32_t *src = idx + 2 * n;
  uint8_t *bits = (uint8_t *)(src + n);
  uint8_t *d_bits = (uint8_t *)c->scratch_get("qs_bits", n + 64);
  uint32_t *d_idx = (uint32_t *)c->scratch_get("qs_idx", 8 * n + 64);
  uint32_t *d_src = (uint32_t *)c->scratch_get("qs_src", 4 * n + 64);
  W *red = (W *)c->scratch_get("qs_red", 3 * kW * n + 64);
  W *piv = (W *)c->scratch_get("qs_piv", 3 * kW * n + 64);
  W *eq = (W *)c->scratch_get("qs_eq", 3 * kW * n + 64);
  W *gt = (W *)c->scratch_get("qs_gt", 3 * kW * n + 64);
  uint64_t num_pivots = 1;
  std::vector<std::pair<int64_t, int64_t>> segs;
  for (;;) {
    // the non-pivot rows, then the pivot of each (prefix max of pivots)
    const uint64_t size = n - num_pivots;
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
      ORQG_HIP(hipMemcpyAsync(d_idx, idx, 4 * size, hipMemcpyHostToDevice, c->stream));
      ORQG_HIP(hipMemcpyAsync(d_idx + size, idx + size, 4 * size, hipMemcpyHostToDevice,
                              c->stream));
      const unsigned g = grid_for(size, 256, c->num_cu * 16);
      hipLaunchKernelGGL(k_qs_gather<W>, dim3(g), dim3(256), 0, c->stream, size, d_idx, P, red);
      hipLaunchKernelGGL(k_qs_gather<W>, dim3(g), dim3(256), 0, c->stream, size, d_idx + size,
                         P, piv);
      ORQG_HIP(hipGetLastError());
      // reduced_vec._compare(pivot_vec, temp, comparisons): gt = reduced > pivot
      orqg_vec vr = view(c, red, size, kW, ORQG_BSHARED), vp = view(c, piv, size, kW, ORQG_BSHARED);
      orqg_vec ve = view(c, eq, size, kW, ORQG_BSHARED), vg = view(c, gt, size, kW, ORQG_BSHARED);
      compare(c, &vr, &vp, &ve, &vg, true, 0, 0);
      // comparisons.open(): only bit 0 of each opened row is used, so only
      // it leaves the device; the message log carries the whole rows
      hipLaunchKernelGGL(k_qs_open_bit<W>, dim3(g), dim3(256), 0, c->stream, size, gt, d_bits);
      ORQG_HIP(hipGetLastError());
      ORQG_HIP(hipMemcpyAsync(bits, d_bits, size, hipMemcpyDeviceToHost, c->stream));
      ORQG_HIP(hipStreamSynchronize(c->stream));
      for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, size * kW);
      host_ranges(size, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t t = lo; t < hi; t++) cmp[idx[t]] = bits[t];  // distinct rows
      });
    }
    // partition every segment between marked positions (quicksort.h:149-165).
    // The segments are found first (a partition only marks positions inside
    // its own segment), then partitioned on host threads: each segment's rows
    // of src, cmp and pivots are its own.
    host_ranges(n, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t j = lo; j < hi; j++) src[j] = (uint32_t)j;
    });
    segs.clear();
    int64_t index = -1;
    for (int64_t p = 0; p < (int64_t)n; p++) {
      if (pivots[p] < 0) continue;
      if (p - index == 1) {  // adjacent pivots
        index++;
        continue;
      }
      segs.emplace_back(index, p - 1);
      index = p;
    }
    segs.emplace_back(index, (int64_t)n - 1);
    const int nthr = n >= (1u << 16) && segs.size() > 1
                         ? (int)std::min<size_t>(segs.size(), kHostThreads)
                         : 1;
    uint64_t marked = 0;  // positions the partitions mark as pivots
    if (nthr == 1) {
      for (const auto &sg : segs) marked += partition(src, cmp, sg.first, sg.second, pivots);
    } else {
      // contiguous runs of segments with about n / nthr rows each
      std::vector<size_t> cut(nthr + 1, segs.size());
      cut[0] = 0;
      for (size_t q = 0, t = 1; q < segs.size() && t < (size_t)nthr; q++)
        if ((uint64_t)(segs[q].second + 1) * nthr >= (uint64_t)t * n) cut[t++] = q + 1;
      std::vector<uint64_t> mk(nthr, 0);
      std::vector<std::thread> th;
      for (int t = 0; t < nthr; t++)
        th.emplace_back([&, t] {
          for (size_t q = cut[t]; q < cut[t + 1]; q++)
            mk[t] += partition(src, cmp, segs[q].first, segs[q].second, pivots);
        });
      for (auto &x : th) x.join();
      for (uint64_t x : mk) marked += x;
    }
    // (cmp stays indexed by position, as in the reference: a partition reads
    // cmp[j] before any swap has touched position j, and the next round
    // rewrites every non-pivot position before reading it)
    ORQG_HIP(hipMemcpyAsync(d_src, src, 4 * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_qs_gather<W>, dim3(grid_for(n, 256, c->num_cu * 16)), dim3(256), 0,
                       c->stream, n, d_src, P, Q);
    ORQG_HIP(hipGetLastError());
    std::swap(P, Q);
    // the host vectors above are reused next round: the copies out of them
    // must have been consumed
    ORQG_HIP(hipStreamSynchronize(c->stream));
    num_pivots += marked;
    if (num_pivots >= n) return P;
  }
}

// pad, shuffle, sort rounds and unpad for keys of type K padded to P
template <typename K, typename P>
void quicksort_padded(orqg_ctx *c, orqg_vec *v, bool desc, orqg_vec *perm_out) {
  const uint64_t n = v->n;
  constexpr int kP = (int)sizeof(P);
  P *Pb = (P *)c->scratch_get("qs_P", 3 * kP * n + 64);
  P *Qb = (P *)c->scratch_get("qs_Q", 3 * kP * n + 64);
  const unsigned grid = grid_for(n, 256, c->num_cu * 16);
  hipLaunchKernelGGL((k_qs_pad<K, P>), dim3(grid), dim3(256), 0, c->stream, n, (int)desc,
                     (const K *)v->d, Pb);
  ORQG_HIP(hipGetLastError());
  {  // v.shuffle()
    orqg_vec pv = view(c, Pb, n, kP, ORQG_BSHARED);
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
  Pb = quicksort_rounds<P>(c, Pb, Qb, n);
  hipLaunchKernelGGL((k_qs_unpad<K, P>), dim3(grid), dim3(256), 0, c->stream, n, (int)desc, Pb,
                     (K *)v->d, (uint32_t *)perm_out->d);
  ORQG_HIP(hipGetLastError());
}

}  // namespace

void quicksort(orqg_ctx *c, orqg_vec *v, int order, orqg_vec *perm_out) {
  require_one_thread(c, "quicksort");
  if (!v || !perm_out || (v->width != 4 && v->width != 8) || perm_out->width != 4 ||
      perm_out->n != v->n)
    throw Error(ORQG_E_ARG,
                "quicksort: 32- or 64-bit keys and a 32-bit permutation of the same size");
  const uint64_t n = v->n;
  if (n >= (1ull << 31)) throw Error(ORQG_E_ARG, "quicksort: n must be < 2^31");
  const bool desc = order == ORQG_DESC;
  // 1 permutation for the shuffle, 1 for remove_padding's b2a, 1 pair for
  // the inversion; pairs are single permutations in 3PC (quicksort.h:215-224)
  reserve_perms(c, n, 3);
  if (n == 0) return;
  TimedRegion tr(c, "quicksort");
  if (v->width == 4)
    quicksort_padded<uint32_t, uint64_t>(c, v, desc, perm_out);
  else
    quicksort_padded<uint64_t, u128>(c, v, desc, perm_out);
  perm_out->enc = ORQG_BSHARED;
  perm_b2a(c, perm_out);
  if (desc) perm_negate(c, perm_out);
  perm_invert(c, perm_out);
  ORQG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace orqg
