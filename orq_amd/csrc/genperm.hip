// gen_perm on the GPU, reproduced exactly: the local permutations of the
// honest-majority sharded permutations
// (include/core/random/permutations/hm_sharded_permutation_generator.h:70-108)
// with the same keystream words, the same rejections and the same swaps.
//
// gen_perm is a sequential Fisher-Yates: for i = 0..n-1 it draws uint32 words
// from the group PRG (refilled 2^20 words = 4 nonces at a time, eagerly), takes
// v = w & (2^bit_width(n) - 1) and accepts the first v <= n-1-i; then it swaps
// positions i and i+v.  With k = n - i open positions, draw d is accepted iff
// v_d < k at that moment.  The work splits into
//   (1) acceptance: which draws are accepted -- r_i = v of the i-th acceptance
//       and D = draws consumed (fixes the next permutation's first nonce);
//   (2) swap resolution: the permutation those swaps produce.
//
// (1) walks the draw stream of each group PRG ({0,1}, {0,2}, {1,2}: three
// independent chains, processed by the same launches) window by window.  In a
// window that starts with k open positions:
//   * k_gp_draws: keystream -> draws; draws with v >= k can never be accepted
//     (k only decreases); the others ("candidates") are compacted per tile;
//   * k_gp_classify: candidate a (rank in the window) is accepted for sure if
//     v < k - a (at most a acceptances precede it); otherwise it is
//     "ambiguous" and is listed, per tile, in rank order;
//   * k_gp_resolve: one workgroup per chain decides the ambiguous candidates
//     exactly with alternating optimistic passes, builds the sorted list of
//     rejected ranks, and advances the chain (next window, or the next
//     permutation at the refill boundary after the last acceptance);
//   * k_gp_commit: accepted candidate a becomes r_{i0 + a - #rejected before a}.
// The window size keeps the expected number of ambiguous candidates within the
// resolver's capacity; an overflowing window is retried at a quarter of its
// size, down to kWinMin draws, which always fits.
//
// (2) uses that position i is final after step i: Fisher-Yates leaves at i
// what position t_i = i + r_i held just before step i, found by walking "last
// earlier step that targeted this position" links.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

namespace {

constexpr int kGpThreads = kAesThreads;                   // 1024
constexpr int kDrawsPerThread = 32;                       // 8 AES blocks per thread
constexpr uint32_t kTile = kGpThreads * kDrawsPerThread;  // 32768 draws
constexpr uint32_t kWinMax = 1u << 26;                    // draws per window
constexpr uint32_t kWinMin = 1024;
constexpr uint32_t kMaxTiles = kWinMax / kTile;           // 2048
constexpr uint32_t kAmbCap = 65536;  // ambiguous candidates per window (level 1)
constexpr uint32_t kL2Cap = 16384;   // undecided after the first exact pass (in LDS)
constexpr uint64_t kRefillWords = 1ull << 20;  // gen_perm buffer: 2^20 uint32
constexpr uint64_t kWordsPerNonce = 1ull << 18;

struct ChainState {
  uint64_t nonce;   // nonce where the current permutation's draws start
  uint64_t d0;      // draw index (inside the current permutation) of the window start
  uint32_t k;       // index of the permutation being generated
  uint32_t i;       // acceptances so far in it
  uint32_t W;       // draws in the current window
  uint32_t done;
  uint32_t err;
  // the window just resolved, for k_gp_commit
  uint32_t c_valid, c_i0, c_k, c_ntiles, c_R;
};

struct ChainBufs {
  ChainState *st;      // [3]
  uint64_t *tc;        // [3][kWinMax]   per tile: (draw offset << 32) | v of its candidates
  uint32_t *tcnt;      // [3][kMaxTiles] candidates per tile
  uint32_t *toff;      // [3][kMaxTiles] rank of the tile's first candidate
  uint64_t *ta;        // [3][kWinMax]   per tile: (rank << 32) | v of its ambiguous candidates
  uint32_t *tacnt;     // [3][kMaxTiles]
  uint32_t *amb_a;     // [3][kAmbCap]   ambiguous candidates in rank order: rank
  int32_t *amb_u;      // [3][kAmbCap]   ... and U = rank - (k - v)
  uint32_t *m_tot;     // [3]            ambiguous candidates in the window
  uint32_t *rej;       // [3][kAmbCap]   ranks of rejected candidates, ascending
  uint32_t *r_out;     // [3][count][n]  accepted offsets r_i
  uint64_t *D;         // [3][count]     draws consumed per permutation
  uint64_t *nonce_at;  // [3][count+1]   nonce at the start of each permutation
};

// Window for k open positions: about W^2 k / (2 maskN^2) candidates are
// ambiguous; 256 maskN / sqrt(k) keeps that near half of kAmbCap.
__host__ __device__ inline uint32_t choose_window(uint32_t k, uint32_t maskN) {
  double w = 256.0 * (double)maskN / sqrt((double)(k ? k : 1));
  if (w >= (double)kWinMax) return kWinMax;
  uint32_t W = w >= kTile ? (uint32_t)(w / kTile) * kTile : (uint32_t)(w / kWinMin) * kWinMin;
  return W < kWinMin ? kWinMin : W;
}

__device__ inline uint32_t ntiles_of(const ChainState &s) {
  return s.done ? 0u : (s.W + kTile - 1) / kTile;
}

// exclusive block scan of one value per thread (blockDim multiple of 64)
__device__ inline uint32_t block_scan(uint32_t x, uint32_t *s_w, uint32_t &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  __syncthreads();
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  uint32_t base = 0;
  total = 0;
  for (int w = 0; w < nw; w++) {
    if (w < wv) base += s_w[w];
    total += s_w[w];
  }
  return base + inc - x;
}

// (chain, tile) of flat task `task` over the three chains' tiles; false if none
__device__ inline bool task_tile(const uint32_t (&nt)[3], uint32_t task, int &chain,
                                 uint32_t &tile) {
  if (task < nt[0]) {
    chain = 0;
    tile = task;
    return true;
  }
  task -= nt[0];
  if (task < nt[1]) {
    chain = 1;
    tile = task;
    return true;
  }
  task -= nt[1];
  if (task < nt[2]) {
    chain = 2;
    tile = task;
    return true;
  }
  return false;
}

// ---- (a) keystream -> candidates, compacted per tile ------------------------------
template <int COPIES>
__global__ void __launch_bounds__(kGpThreads) k_gp_draws(const uint32_t *__restrict__ tab,
                                                         const uint32_t *__restrict__ rk_groups,
                                                         ChainBufs B, uint32_t n, uint32_t mask) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  __shared__ uint32_t s_w[kGpThreads / 64];
  uint32_t nt[3];
  for (int c = 0; c < 3; c++) nt[c] = ntiles_of(B.st[c]);
  const uint32_t total = nt[0] + nt[1] + nt[2];
  if (blockIdx.x >= total) return;
  aes_load_tables<COPIES>(lds, tab);
  AesCtx<COPIES> A(lds);
  for (uint32_t task = blockIdx.x; task < total; task += gridDim.x) {
    int chain;
    uint32_t tile;
    task_tile(nt, task, chain, tile);
    const ChainState &st = B.st[chain];
    const uint32_t kstart = n - st.i, W = st.W;
    const uint32_t off0 = tile * kTile + threadIdx.x * kDrawsPerThread;
    uint32_t v[kDrawsPerThread];
    if (off0 < W) {
      const uint32_t *rk = rk_groups + 60 * chain;
      const uint64_t dbase = st.d0 + off0;
#pragma unroll
      for (int b = 0; b < kDrawsPerThread / 4; b++) {
        const uint64_t d = dbase + 4 * b;
        uint4 blk = aes_ctr_block<COPIES>(A, rk, st.nonce + d / kWordsPerNonce,
                                          2 + (uint32_t)((d % kWordsPerNonce) / 4));
        v[4 * b + 0] = blk.x & mask;
        v[4 * b + 1] = blk.y & mask;
        v[4 * b + 2] = blk.z & mask;
        v[4 * b + 3] = blk.w & mask;
      }
    } else {
#pragma unroll
      for (int j = 0; j < kDrawsPerThread; j++) v[j] = 0xffffffffu;  // never a candidate
    }
    uint32_t nc = 0;
#pragma unroll
    for (int j = 0; j < kDrawsPerThread; j++) nc += v[j] < kstart;
    uint32_t tot;
    uint32_t pos = block_scan(nc, s_w, tot);
    uint64_t *out = B.tc + (size_t)chain * kWinMax + (size_t)tile * kTile;
#pragma unroll
    for (int j = 0; j < kDrawsPerThread; j++)
      if (v[j] < kstart) out[pos++] = ((uint64_t)(off0 + j) << 32) | v[j];
    if (threadIdx.x == 0) B.tcnt[(size_t)chain * kMaxTiles + tile] = tot;
    __syncthreads();
  }
}

// ---- (b) ranks, sure accepts, ambiguous lists -----------------------------------
__global__ void __launch_bounds__(kGpThreads) k_gp_classify(ChainBufs B, uint32_t n) {
  __shared__ uint32_t s_w[kGpThreads / 64];
  uint32_t nt[3];
  for (int c = 0; c < 3; c++) nt[c] = ntiles_of(B.st[c]);
  const uint32_t total = nt[0] + nt[1] + nt[2];
  for (uint32_t task = blockIdx.x; task < total; task += gridDim.x) {
    int chain;
    uint32_t tile;
    task_tile(nt, task, chain, tile);
    const uint32_t kstart = n - B.st[chain].i;
    const uint32_t *tcnt = B.tcnt + (size_t)chain * kMaxTiles;
    // rank of the tile's first candidate: sum over the earlier tiles
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < tile; t += blockDim.x) part += tcnt[t];
    uint32_t off;
    block_scan(part, s_w, off);
    const uint32_t cnt = tcnt[tile];
    const uint64_t *in = B.tc + (size_t)chain * kWinMax + (size_t)tile * kTile;
    uint64_t e[kDrawsPerThread];
    uint32_t na = 0;
#pragma unroll
    for (int q = 0; q < kDrawsPerThread; q++) {
      const uint32_t j = threadIdx.x * kDrawsPerThread + q;
      e[q] = j < cnt ? in[j] : 0;
      const uint32_t v = (uint32_t)e[q];
      if (j < cnt && !(off + j < kstart - v)) na++;
    }
    uint32_t atot;
    uint32_t pos = block_scan(na, s_w, atot);
    uint64_t *out = B.ta + (size_t)chain * kWinMax + (size_t)tile * kTile;
#pragma unroll
    for (int q = 0; q < kDrawsPerThread; q++) {
      const uint32_t j = threadIdx.x * kDrawsPerThread + q;
      const uint32_t v = (uint32_t)e[q];
      if (j < cnt && !(off + j < kstart - v)) out[pos++] = ((uint64_t)(off + j) << 32) | v;
    }
    if (threadIdx.x == 0) {
      B.toff[(size_t)chain * kMaxTiles + tile] = off;
      B.tacnt[(size_t)chain * kMaxTiles + tile] = atot;
    }
    __syncthreads();
  }
}

// ---- (b') ambiguous candidates into one rank-ordered array ------------------------
__global__ void __launch_bounds__(kGpThreads) k_gp_gather(ChainBufs B, uint32_t n) {
  __shared__ uint32_t s_w[kGpThreads / 64];
  uint32_t nt[3];
  for (int c = 0; c < 3; c++) nt[c] = ntiles_of(B.st[c]);
  const uint32_t total = nt[0] + nt[1] + nt[2];
  for (uint32_t task = blockIdx.x; task < total; task += gridDim.x) {
    int chain;
    uint32_t tile;
    task_tile(nt, task, chain, tile);
    const uint32_t kstart = n - B.st[chain].i;
    const uint32_t *tacnt = B.tacnt + (size_t)chain * kMaxTiles;
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < tile; t += blockDim.x) part += tacnt[t];
    uint32_t off;
    block_scan(part, s_w, off);
    const uint32_t cnt = tacnt[tile];
    if (tile == nt[chain] - 1 && threadIdx.x == 0) B.m_tot[chain] = off + cnt;
    if (off >= kAmbCap) continue;
    const uint64_t *in = B.ta + (size_t)chain * kWinMax + (size_t)tile * kTile;
    uint32_t *oa = B.amb_a + (size_t)chain * kAmbCap;
    int32_t *ou = B.amb_u + (size_t)chain * kAmbCap;
    for (uint32_t j = threadIdx.x; j < cnt && off + j < kAmbCap; j += blockDim.x) {
      const uint64_t e = in[j];
      const uint32_t a = (uint32_t)(e >> 32);
      oa[off + j] = a;
      ou[off + j] = (int32_t)(a - (kstart - (uint32_t)e));
    }
  }
}

// ---- (c) exact resolution of the ambiguous candidates, chain advance -------------
// Ambiguous candidate j (list index t_j, rank a_j, T_j = k - v_j <= a_j) is
// accepted iff fewer than T_j acceptances precede it, i.e. iff more than
// U_j = a_j - T_j of the ambiguous candidates before it are rejected.
//   type B list (accept iff rejected-before > U): index t <= U -> rejected;
//     U < 0 -> accepted; otherwise accepted iff
//     accepted-before-among-undecided < t - (sure accepts before) - U;
//   type A list (accept iff accepted-before < T): t < T -> accepted;
//     T <= 0 -> rejected; otherwise accepted iff
//     rejected-before-among-undecided > t - (sure rejects before) - T.
// The first pass (type B over all ambiguous candidates) runs over global
// memory; the undecided rest, in order, is resolved in LDS.  Every pass
// decides at least its first item, so the loop ends.
__global__ void __launch_bounds__(kGpThreads) k_gp_resolve(ChainBufs B, uint32_t n,
                                                           uint32_t count, uint32_t maskN) {
  constexpr int PMAX = kAmbCap / kGpThreads;  // 64 items per thread at most
  __shared__ uint16_t s_t[kL2Cap];   // level-1 index of each undecided item
  __shared__ uint16_t s_idx[kL2Cap]; // current list -> undecided item
  __shared__ int32_t s_thr[kL2Cap];
  __shared__ uint8_t s_dec[kL2Cap];  // 1 accepted, 2 rejected
  __shared__ uint32_t s_w[kGpThreads / 64];
  __shared__ uint32_t s_l2;
  const int chain = blockIdx.x;
  ChainState *st = B.st + chain;
  if (st->done) {
    if (threadIdx.x == 0) st->c_valid = 0;
    return;
  }
  const uint32_t kstart = n - st->i;
  const uint32_t ntiles = (st->W + kTile - 1) / kTile;
  const uint32_t *tcnt = B.tcnt + (size_t)chain * kMaxTiles;
  const uint32_t *amb_a = B.amb_a + (size_t)chain * kAmbCap;
  const int32_t *amb_u = B.amb_u + (size_t)chain * kAmbCap;
  uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
  const uint32_t m = B.m_tot[chain];
  uint32_t C;
  {
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) part += tcnt[t];
    block_scan(part, s_w, C);
  }
  const bool over = m > kAmbCap;
  // level 1 (type B over all m; U >= 0, so no sure accepts): each thread owns
  // the contiguous run [t0, t0 + P); bit q of `rejm` = item t0+q rejected
  const uint32_t P = over ? 0 : (m + kGpThreads - 1) / kGpThreads;
  const uint32_t t0 = threadIdx.x * P;
  uint64_t rejm = 0, undm = 0;
  for (uint32_t q = 0; q < P && t0 + q < m; q++) {
    const int32_t U = amb_u[t0 + q];
    if ((int32_t)(t0 + q) <= U) rejm |= 1ull << q;
    else undm |= 1ull << q;
  }
  uint32_t l2;
  uint32_t upos = block_scan((uint32_t)__builtin_popcountll(undm), s_w, l2);
  const uint32_t ubase = upos;
  if (threadIdx.x == 0) s_l2 = l2;
  __syncthreads();
  if (over || s_l2 > kL2Cap) {
    // too many ambiguous candidates: retry this window at a quarter size
    if (threadIdx.x == 0) {
      uint32_t w = st->W / 4;
      st->W = w >= kTile ? w / kTile * kTile : (w >= kWinMin ? w / kWinMin * kWinMin : kWinMin);
      st->c_valid = 0;
    }
    return;
  }
  // undecided items continue as a type A list with T' = t - U
  for (uint64_t mm = undm; mm; mm &= mm - 1) {
    const uint32_t q = (uint32_t)__builtin_ctzll(mm), t = t0 + q;
    s_t[upos] = (uint16_t)t;
    s_idx[upos] = (uint16_t)upos;
    s_thr[upos] = (int32_t)t - amb_u[t];
    s_dec[upos] = 0;
    upos++;
  }
  __syncthreads();
  constexpr int PER = kL2Cap / kGpThreads;  // 16
  uint32_t len = s_l2;
  int type_b = 0;
  while (len > 0) {
    const uint32_t beg = threadIdx.x * PER;
    uint8_t dec[PER];
    uint32_t n_sa = 0, n_sr = 0, n_un = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const uint32_t t = beg + q;
      dec[q] = 3;
      if (t < len) {
        const int32_t thr = s_thr[t];
        if (type_b) dec[q] = ((int32_t)t <= thr) ? 2 : (thr < 0 ? 1 : 0);
        else dec[q] = ((int32_t)t < thr) ? 1 : (thr <= 0 ? 2 : 0);
        n_sa += dec[q] == 1;
        n_sr += dec[q] == 2;
        n_un += dec[q] == 0;
      }
    }
    uint32_t tot;
    uint32_t sa = block_scan(n_sa, s_w, tot);
    uint32_t sr = block_scan(n_sr, s_w, tot);
    uint32_t un = block_scan(n_un, s_w, tot);
    const uint32_t new_len = tot;
    int32_t nthr[PER];
    uint16_t nidx[PER];
    uint32_t npos[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const uint32_t t = beg + q;
      npos[q] = 0xffffffffu;
      if (dec[q] == 3) continue;
      const uint16_t id = s_idx[t];
      if (dec[q] != 0) {
        s_dec[id] = dec[q];
      } else {
        const int32_t thr = s_thr[t];
        nthr[q] = type_b ? (int32_t)t - (int32_t)sa - thr : (int32_t)t - (int32_t)sr - thr;
        nidx[q] = id;
        npos[q] = un++;
      }
      sa += dec[q] == 1;
      sr += dec[q] == 2;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (npos[q] != 0xffffffffu) {
        s_idx[npos[q]] = nidx[q];
        s_thr[npos[q]] = nthr[q];
      }
    __syncthreads();
    len = new_len;
    type_b = !type_b;
  }
  // this thread's undecided items sit at l2 positions ubase, ubase+1, ...
  {
    uint32_t u = ubase;
    for (uint64_t mm = undm; mm; mm &= mm - 1) {
      if (s_dec[u] == 2) rejm |= 1ull << __builtin_ctzll(mm);
      u++;
    }
  }
  (void)PMAX;
  // rejected ranks in ascending order (the ambiguous list is in rank order)
  uint32_t R;
  uint32_t rpos = block_scan((uint32_t)__builtin_popcountll(rejm), s_w, R);
  for (uint64_t mm = rejm; mm; mm &= mm - 1) rej[rpos++] = amb_a[t0 + __builtin_ctzll(mm)];
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t acc = C - R;
    if (acc > kstart) atomicOr(&st->err, 4u);
    st->c_valid = 1;
    st->c_i0 = st->i;
    st->c_k = st->k;
    st->c_ntiles = ntiles;
    st->c_R = R;
    if (acc >= kstart) {
      // the permutation ends here: rank a* of the kstart-th acceptance is
      // kstart-1 shifted past every rejected rank at or before it
      uint32_t a = kstart - 1;
      for (uint32_t q = 0; q < R; q++) {
        if (rej[q] <= a) a++;
        else break;
      }
      const uint32_t *toff = B.toff + (size_t)chain * kMaxTiles;
      uint32_t lo = 0, hi = ntiles;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (toff[mid] <= a) lo = mid;
        else hi = mid;
      }
      const uint64_t e = B.tc[(size_t)chain * kWinMax + (size_t)lo * kTile + (a - toff[lo])];
      const uint64_t D_end = st->d0 + (e >> 32) + 1;
      const uint32_t k = st->k;
      B.D[(size_t)chain * count + k] = D_end;
      // 4 nonces per 2^20-word refill, one refill up front
      st->nonce += 4 * (1 + D_end / kRefillWords);
      st->k = k + 1;
      st->i = 0;
      st->d0 = 0;
      B.nonce_at[(size_t)chain * (count + 1) + k + 1] = st->nonce;
      if (st->k >= count) st->done = 1;
      st->W = choose_window(n, maskN);
    } else {
      st->i += acc;
      st->d0 += st->W;
      st->W = choose_window(n - st->i, maskN);
    }
  }
}

// ---- (d) r_i for the accepted candidates --------------------------------------------
__global__ void __launch_bounds__(kGpThreads) k_gp_commit(ChainBufs B, uint32_t n, uint32_t count) {
  uint32_t nt[3];
  for (int c = 0; c < 3; c++) nt[c] = B.st[c].c_valid ? B.st[c].c_ntiles : 0;
  const uint32_t total = nt[0] + nt[1] + nt[2];
  for (uint32_t task = blockIdx.x; task < total; task += gridDim.x) {
    int chain;
    uint32_t tile;
    task_tile(nt, task, chain, tile);
    const ChainState &st = B.st[chain];
    const uint32_t R = st.c_R, i0 = st.c_i0, k = st.c_k;
    const uint32_t kstart = n - i0;
    const uint32_t cnt = B.tcnt[(size_t)chain * kMaxTiles + tile];
    const uint32_t off = B.toff[(size_t)chain * kMaxTiles + tile];
    const uint64_t *in = B.tc + (size_t)chain * kWinMax + (size_t)tile * kTile;
    const uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
    uint32_t *r = B.r_out + ((size_t)chain * count + k) * (size_t)n;
    // each thread commits a contiguous run of candidates: one binary search
    // into the ascending rejected ranks, then a merge walk
    const uint32_t per = (cnt + blockDim.x - 1) / blockDim.x;
    const uint32_t j0 = threadIdx.x * per;
    if (j0 >= cnt) continue;
    const uint32_t j1 = j0 + per < cnt ? j0 + per : cnt;
    uint32_t lo = 0, hi = R;  // lower_bound of off + j0 in rej
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (rej[mid] < off + j0) lo = mid + 1;
      else hi = mid;
    }
    for (uint32_t j = j0; j < j1; j++) {
      const uint32_t a = off + j;
      if (lo < R && rej[lo] == a) {
        lo++;
        continue;
      }
      const uint32_t accr = a - lo;
      if (accr >= kstart) continue;  // cannot happen for an exact resolution
      r[i0 + accr] = (uint32_t)in[j];
    }
  }
}

__global__ void k_gp_done(ChainBufs B, uint32_t *flag) {
  if (threadIdx.x == 0) {
    flag[0] = B.st[0].done & B.st[1].done & B.st[2].done;
    flag[1] = B.st[0].err | B.st[1].err | B.st[2].err;
    for (int g = 0; g < 3; g++) flag[2 + g] = B.st[g].k;  // permutations completed
  }
}

// ---- swap resolution ----------------------------------------------------------
// Step j swaps positions j and t_j = j + r_j.  Steps are chained per target
// position (head[t] -> most recently linked step -> next[...]); list order is
// arbitrary, the walks take maxima.
__global__ void k_fy_link(uint32_t n, const uint32_t *__restrict__ r_all, size_t chain_stride,
                          int32_t *__restrict__ head, int32_t *__restrict__ next,
                          uint32_t *__restrict__ err) {
  const uint32_t *r = r_all + blockIdx.y * chain_stride;
  int32_t *hd = head + (size_t)blockIdx.y * n, *nx = next + (size_t)blockIdx.y * n;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t t = j + r[j];
    if (t >= n) {
      atomicOr(err, 8u);
      nx[j] = -1;
      continue;
    }
    nx[j] = atomicExch(&hd[t], (int32_t)j);
  }
}

// largest step index < lim whose target is pos, or -1
__device__ inline int32_t fy_last_before(const int32_t *__restrict__ hd,
                                         const int32_t *__restrict__ nx, uint32_t pos,
                                         uint32_t lim, uint32_t n) {
  int32_t best = -1;
  int32_t q = hd[pos];
  for (uint32_t guard = 0; q >= 0 && guard < n; guard++) {
    if ((uint32_t)q < lim && q > best) best = q;
    q = nx[q];
  }
  return best;
}

// Position j is final after step j; it receives what position t_j held just
// before step j: V(q) for the last earlier step q that targeted t_j, or t_j
// itself.  V(m), the value at position m just before step m, is V(p) for the
// last step p < m that targeted m, or m.
__global__ void k_fy_walk(uint32_t n, const uint32_t *__restrict__ r_all, size_t chain_stride,
                          const int32_t *__restrict__ head, const int32_t *__restrict__ next,
                          int32_t *fwd0, int32_t *fwd1, int32_t *fwd2, int32_t *inv0, int32_t *inv1,
                          int32_t *inv2) {
  const uint32_t *r = r_all + blockIdx.y * chain_stride;
  const int32_t *hd = head + (size_t)blockIdx.y * n, *nx = next + (size_t)blockIdx.y * n;
  int32_t *fwd = blockIdx.y == 0 ? fwd0 : blockIdx.y == 1 ? fwd1 : fwd2;
  int32_t *inv = blockIdx.y == 0 ? inv0 : blockIdx.y == 1 ? inv1 : inv2;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t t = j + r[j];
    if (t >= n) continue;  // flagged by k_fy_link
    int32_t q = fy_last_before(hd, nx, t, j, n);
    uint32_t val = t;
    if (q >= 0) {
      uint32_t m = (uint32_t)q;
      for (uint32_t guard = 0; guard < n; guard++) {
        int32_t p = fy_last_before(hd, nx, m, m, n);
        if (p < 0) break;
        m = (uint32_t)p;
      }
      val = m;
    }
    fwd[j] = (int32_t)val;
    inv[val] = (int32_t)j;
  }
}

}  // namespace

void free_perm(ShardedPerm &p) {
  for (int g = 0; g < 3; g++) {
    dev_free(p.fwd[g]);
    dev_free(p.inv[g]);
    p.fwd[g] = p.inv[g] = nullptr;
  }
}

// PermutationManager::reserve for 3PC (permutation_manager.h:102-181):
// `count` new sharded permutations, each with one local permutation per pair
// group, generated in queue order from each group's PRG
// (HMShardedPermutationGenerator::generateBatch,
// hm_sharded_permutation_generator.h:178-194).  The reference also generates a
// permutation for the {0,1,2} group, which no operator on this path applies and
// whose PRG feeds nothing else here, so it is not generated.
void reserve_perms(orqg_ctx *c, size_t n_, size_t count_) {
  if (!c->queue.empty() && c->stored_size != n_) {
    for (auto &p : c->queue) free_perm(p);
    c->queue.clear();
  }
  c->stored_size = n_;
  if (count_ <= c->queue.size()) return;
  count_ -= c->queue.size();
  if (n_ >= (1ull << 31)) throw Error(ORQG_E_ARG, "reserve: n must be < 2^31");
  const uint32_t n = (uint32_t)n_, count = (uint32_t)count_;
  uint32_t bits = 0;
  for (uint32_t t = n; t; t >>= 1) bits++;
  const uint32_t maskN = 1u << bits;
  const uint32_t mask = maskN - 1;

  std::vector<ShardedPerm> fresh(count);
  for (auto &p : fresh) {
    p.n = n;
    for (int g = 0; g < 3; g++) {
      p.fwd[g] = (int32_t *)dev_alloc(c, (size_t)n * 4 + 16);
      p.inv[g] = (int32_t *)dev_alloc(c, (size_t)n * 4 + 16);
    }
  }
  if (n == 0) {
    // gen_perm still fills its buffer once: 4 nonces per local permutation
    for (int g = 0; g < 3; g++) c->nonce[kGroup0 + g] += 4ull * count;
    for (auto &p : fresh) c->queue.push_back(p);
    return;
  }

  TimedRegion tr(c, "gen_perm");
  ChainBufs B;
  B.st = (ChainState *)c->scratch_get("gp_state", 3 * sizeof(ChainState));
  B.tc = (uint64_t *)c->scratch_get("gp_tc", 3ull * kWinMax * 8);
  B.tcnt = (uint32_t *)c->scratch_get("gp_tcnt", 3ull * kMaxTiles * 4);
  B.toff = (uint32_t *)c->scratch_get("gp_toff", 3ull * kMaxTiles * 4);
  B.ta = (uint64_t *)c->scratch_get("gp_ta", 3ull * kWinMax * 8);
  B.tacnt = (uint32_t *)c->scratch_get("gp_tacnt", 3ull * kMaxTiles * 4);
  B.amb_a = (uint32_t *)c->scratch_get("gp_amb_a", 3ull * kAmbCap * 4);
  B.amb_u = (int32_t *)c->scratch_get("gp_amb_u", 3ull * kAmbCap * 4);
  B.m_tot = (uint32_t *)c->scratch_get("gp_m_tot", 64);
  B.rej = (uint32_t *)c->scratch_get("gp_rej", 3ull * kAmbCap * 4);
  B.r_out = (uint32_t *)c->scratch_get("gp_r", 3ull * count * n * 4);
  B.D = (uint64_t *)c->scratch_get("gp_D", 3ull * count * 8);
  B.nonce_at = (uint64_t *)c->scratch_get("gp_nonce_at", 3ull * (count + 1) * 8);
  uint32_t *flag = (uint32_t *)c->scratch_get("gp_flag", 32);

  ChainState init[3];
  std::memset(init, 0, sizeof(init));
  std::vector<uint64_t> nat(3ull * (count + 1), 0);
  for (int g = 0; g < 3; g++) {
    init[g].nonce = c->nonce[kGroup0 + g];
    init[g].W = choose_window(n, maskN);
    nat[(size_t)g * (count + 1)] = init[g].nonce;
  }
  ORQG_HIP(hipMemcpyAsync(B.st, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.nonce_at, nat.data(), nat.size() * 8, hipMemcpyHostToDevice, c->stream));

  const unsigned grid = std::max(1, c->num_cu);
  uint32_t host_flag[5] = {0, 0, 0, 0, 0};
  const int kBatch = 16;
  // The window kernels take all their state from device memory, so a batch of
  // windows is one static launch sequence: it is captured once as a graph.
  auto enqueue_batch = [&]() {
    for (int w = 0; w < kBatch; w++) {
      hipLaunchKernelGGL(k_gp_draws<kCopies>, dim3(grid), dim3(kGpThreads), 0, c->stream, c->d_t0,
                         c->rk(kGroup0), B, n, mask);
      hipLaunchKernelGGL(k_gp_classify, dim3(grid), dim3(kGpThreads), 0, c->stream, B, n);
      hipLaunchKernelGGL(k_gp_gather, dim3(grid), dim3(kGpThreads), 0, c->stream, B, n);
      hipLaunchKernelGGL(k_gp_resolve, dim3(3), dim3(kGpThreads), 0, c->stream, B, n, count,
                         maskN);
      hipLaunchKernelGGL(k_gp_commit, dim3(grid), dim3(kGpThreads), 0, c->stream, B, n, count);
    }
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  ORQG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_batch();
  ORQG_HIP(hipStreamEndCapture(c->stream, &graph));
  ORQG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  struct GraphGuard {
    hipGraph_t g;
    hipGraphExec_t e;
    ~GraphGuard() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard{graph, gexec};
  // Swap resolution of a finished local permutation runs on a second stream
  // while the windows of the next ones proceed: it is memory-latency bound,
  // the windows are AES (VALU/LDS) bound.
  int32_t *head = (int32_t *)c->scratch_get("fy_head", (size_t)n * 4 + 64);
  int32_t *next = (int32_t *)c->scratch_get("fy_next", (size_t)n * 4 + 64);
  uint32_t *err = (uint32_t *)c->scratch_get("fy_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  ORQG_HIP(hipEventRecord(c->ev_aux, c->stream));
  ORQG_HIP(hipStreamWaitEvent(c->stream2, c->ev_aux, 0));
  const unsigned fgrid = grid_for(n, 256, c->num_cu * 8);
  const size_t chain_stride = (size_t)count * n;
  uint32_t launched[3] = {0, 0, 0};
  auto resolve_swaps = [&](int g, uint32_t upto) {
    for (uint32_t k = launched[g]; k < upto; k++) {
      const uint32_t *r = B.r_out + (size_t)g * chain_stride + (size_t)k * n;
      ORQG_HIP(hipMemsetAsync(head, 0xff, (size_t)n * 4, c->stream2));
      hipLaunchKernelGGL(k_fy_link, dim3(fgrid, 1), dim3(256), 0, c->stream2, n, r, chain_stride,
                         head, next, err);
      hipLaunchKernelGGL(k_fy_walk, dim3(fgrid, 1), dim3(256), 0, c->stream2, n, r, chain_stride,
                         head, next, fresh[k].fwd[g], fresh[k].fwd[g], fresh[k].fwd[g],
                         fresh[k].inv[g], fresh[k].inv[g], fresh[k].inv[g]);
    }
    if (upto > launched[g]) launched[g] = upto;
  };
  // every window either resolves or is retried at a quarter size down to
  // kWinMin draws, which always resolves; the cap only guards logic errors
  const uint64_t max_windows = 64ull * count * (1 + (uint64_t)n / 1024) + (1u << 20);
  for (uint64_t it = 0;; it++) {
    if (it * kBatch > max_windows) throw Error(ORQG_E_PARITY, "gen_perm: window limit exceeded");
    ORQG_HIP(hipGraphLaunch(gexec, c->stream));
    hipLaunchKernelGGL(k_gp_done, dim3(1), dim3(64), 0, c->stream, B, flag);
    ORQG_HIP(hipGetLastError());
    ORQG_HIP(hipMemcpyAsync(host_flag, flag, sizeof(host_flag), hipMemcpyDeviceToHost, c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));
    if (host_flag[1])
      throw Error(ORQG_E_PARITY,
                  "gen_perm: device-side check failed (flags " + std::to_string(host_flag[1]) + ")");
    // r of permutations below host_flag[2+g] is complete (the commits ran
    // before the synchronisation above)
    for (int g = 0; g < 3; g++) resolve_swaps(g, std::min(host_flag[2 + g], count));
    if (host_flag[0]) break;
  }
  ChainState fin[3];
  ORQG_HIP(hipMemcpyAsync(fin, B.st, sizeof(fin), hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  for (int g = 0; g < 3; g++) c->nonce[kGroup0 + g] = fin[g].nonce;
  for (int g = 0; g < 3; g++) resolve_swaps(g, count);
  ORQG_HIP(hipGetLastError());
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream2));
  ORQG_HIP(hipStreamSynchronize(c->stream2));
  if (herr) throw Error(ORQG_E_PARITY, "gen_perm: swap resolution check failed");
  for (auto &p : fresh) c->queue.push_back(p);
}

ShardedPerm next_perm(orqg_ctx *c, size_t n) {
  if (c->queue.empty() || c->stored_size != n) reserve_perms(c, n, 1);
  ShardedPerm p = c->queue.front();
  c->queue.pop_front();
  return p;
}

}  // namespace orqg
