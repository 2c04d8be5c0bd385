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
//   * k_gp_window: keystream -> draws; draws with v >= k can never be accepted
//     (k only decreases); the others ("candidates") get their rank a in the
//     window through a decoupled look-back over the tiles and are written in
//     rank order.  Candidate a is accepted for sure if v < k - a (at most a
//     acceptances precede it); otherwise it is "ambiguous" and is written, in
//     rank order too, to the resolver's list (a second look-back).  The same
//     launch also commits the previously resolved window (tasks taken after
//     the tiles): accepted candidate a becomes r_{i0 + a - #rejected before a};
//   * k_gp_resolve: one workgroup per chain decides the ambiguous candidates
//     exactly (a sure-reject test, then an in-order walk of the rest), builds
//     the sorted list of rejected ranks, and advances the chain (next window,
//     or the next permutation at the refill boundary after the last
//     acceptance).
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
// sparse windows (candidate density <= 1/256) use super-tiles of 4 AES rounds
// per workgroup, their few candidates stashed in LDS between the rounds
constexpr uint32_t kSuperRounds = 4;
constexpr uint32_t kSuperTile = kTile * kSuperRounds;  // 131072 draws
constexpr int kStash = 192;                             // candidates per wave
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
  uint32_t seq;     // window sequence number (tags the look-back status words)
  uint32_t buf;     // candidate buffer of the current window
  uint32_t C_tot;   // candidates in the current window (written by its last tile)
  uint32_t m_tot;   // ambiguous candidates in it
  uint32_t ovf;     // a sparse tile's candidate stash overflowed
  uint32_t force_dense;  // retry the current window with dense tiles
  // the window just resolved, committed by the next k_gp_window launch
  uint32_t c_valid, c_i0, c_k, c_C, c_R, c_buf;
};

struct ChainBufs {
  ChainState *st;      // [3]
  uint32_t *cand_v;    // [2][3][cand_cap] candidates in rank order: v
  uint32_t *cand_d;    // [2][3][cand_cap] ... and the draw offset in the window, written
                       // only when the window can end the permutation (W >= k)
  uint32_t cand_cap;
  uint64_t *stat_c;    // [3][kMaxTiles] look-back status of the candidate counts
  uint64_t *stat_a;    // [3][kMaxTiles] ... of the ambiguous counts
  uint32_t *ctr;       // [2] tile ticket, finished workgroups
  uint32_t *amb_a;     // [3][kAmbCap]   ambiguous candidates in rank order: rank
  int32_t *amb_u;      // [3][kAmbCap]   ... and U = rank - (k - v)
  uint32_t *rej;       // [3][kAmbCap]   ranks of rejected candidates, ascending
  uint32_t *r_out;     // [3][count][n]  accepted offsets r_i
  uint64_t *D;         // [3][count]     draws consumed per permutation
  uint64_t *nonce_at;  // [3][count+1]   nonce at the start of each permutation
};

// Window for k open positions: about W^2 k / (2 maskN^2) candidates are
// ambiguous; 256 maskN / sqrt(k) keeps that near half of kAmbCap.  Near the
// end of a permutation that is far more than the maskN * H(k) draws the k
// remaining positions take on average, so it is capped at three times those.
__host__ __device__ inline uint32_t choose_window(uint32_t k, uint32_t maskN) {
  const double kk = (double)(k ? k : 1);
  double w = 256.0 * (double)maskN / sqrt(kk);
  const double rest = 3.0 * (double)maskN * (log(kk) + 0.58);
  if (rest < w) w = rest;
  if (w >= (double)kWinMax) return kWinMax;
  uint32_t W = w >= kTile ? (uint32_t)(w / kTile) * kTile : (uint32_t)(w / kWinMin) * kWinMin;
  return W < kWinMin ? kWinMin : W;
}

// super-tiles when candidates are rare and the window still gives every
// workgroup at least one (otherwise dense tiles spread it over more CUs)
__device__ inline bool sparse_window(const ChainState &s, uint32_t n, uint32_t maskN) {
  return !s.force_dense && (uint64_t)(n - s.i) * 256 <= maskN &&
         (uint64_t)s.W >= (uint64_t)kSuperTile * gridDim.x;
}

__device__ inline uint32_t ntiles_of(const ChainState &s, uint32_t n, uint32_t maskN) {
  const uint32_t td = sparse_window(s, n, maskN) ? kSuperTile : kTile;
  return s.done ? 0u : (s.W + td - 1) / td;
}

// exclusive block scan of one value per thread (blockDim multiple of 64)
template <typename T>
__device__ inline T block_scan(T x, T *s_w, T &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  __syncthreads();
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  T base = 0;
  total = 0;
  for (int w = 0; w < nw; w++) {
    if (w < wv) base += s_w[w];
    total += s_w[w];
  }
  return base + inc - x;
}

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ inline uint64_t lanes_below() {
  const int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
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

// ---- decoupled look-back over the tiles of one chain's window ----------------------
// Status word of a tile: (seq << 2 | kind) << 32 | value, kind 1 = the tile's own
// count, 2 = inclusive prefix; a word with another seq is not published yet.
// Tiles are taken in increasing order from a ticket, so every tile waited on
// belongs to a running (or finished) workgroup and publishes its count before
// waiting itself.  Run by one whole wave; returns the exclusive prefix.
constexpr uint32_t kSpinMax = 1u << 20;

__device__ inline uint32_t look_back(uint64_t *stat, uint32_t tile, uint32_t seq, uint32_t cnt,
                                     uint32_t *err) {
  const int lane = threadIdx.x & 63;
  const uint64_t tag_own = (uint64_t)((seq << 2) | 1u) << 32;
  const uint64_t tag_inc = (uint64_t)((seq << 2) | 2u) << 32;
  if (tile == 0) {
    if (lane == 0) __hip_atomic_store(stat, tag_inc | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0)
    __hip_atomic_store(stat + tile, tag_own | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t excl = 0;
  for (int32_t base = (int32_t)tile - 1;; base -= 64) {
    const int32_t idx = base - lane;
    uint32_t kind = 2, val = 0;  // before tile 0: an inclusive prefix of 0
    if (idx >= 0) {
      for (uint32_t spin = 0;; spin++) {
        const uint64_t s =
            __hip_atomic_load(stat + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t hi = (uint32_t)(s >> 32);
        if ((hi >> 2) == seq && (hi & 3u)) {
          kind = hi & 3u;
          val = (uint32_t)s;
          break;
        }
        if (spin >= kSpinMax) {  // never expected: give up loudly instead of hanging
          atomicOr(err, 16u);
          kind = 2;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    const uint64_t incm = __ballot(kind == 2);
    const int first = __builtin_ctzll(incm | (1ull << 63));
    excl += wave_sum(lane <= first ? val : 0u);
    if (incm) break;
  }
  if (lane == 0)
    __hip_atomic_store(stat + tile, tag_inc | (excl + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// first index of the ascending a[0..len) with a[i] >= key, by one wave: 64-ary
// search (a few dependent coalesced loads instead of log2(len) scalar ones)
__device__ inline uint32_t wave_lower_bound(const uint32_t *__restrict__ a, uint32_t len,
                                            uint32_t key) {
  const int lane = threadIdx.x & 63;
  uint32_t lo = 0, hi = len;  // answer in [lo, hi]
  while (hi - lo > 64) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t p = lo + step * lane;
    const uint32_t c = __builtin_popcountll(__ballot(p < hi && a[p] < key));
    const uint32_t nlo = c == 0 ? lo : lo + step * (c - 1) + 1;
    hi = lo + step * c < hi ? lo + step * c : hi;
    lo = nlo;
  }
  const uint32_t p = lo + lane;
  return lo + __builtin_popcountll(__ballot(p < hi && a[p] < key));
}

// ---- commit: r_i for the accepted candidates of a resolved window ------------------
// Accepted candidate of rank a becomes r_{i0 + a - #rejected ranks below a}.  A
// wave takes 1024 consecutive ranks: the rejected ones among them (a slice of
// the ascending rejected list, found by a wave search) become a 1024-bit mask
// in LDS, and lane l handles ranks base + 64 j + l, so loads and stores coalesce.
__device__ inline void commit_block(const ChainBufs &B, int chain, uint32_t b, uint32_t n,
                                    uint32_t count, unsigned long long *s_mask) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = lanes_below();
  const ChainState &st = B.st[chain];
  const uint32_t C = st.c_C, R = st.c_R, i0 = st.c_i0, kstart = n - i0;
  if (b * 1024 >= C) return;
  const uint32_t *cand = B.cand_v + ((size_t)st.c_buf * 3 + chain) * B.cand_cap;
  const uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
  uint32_t *r = B.r_out + ((size_t)chain * count + st.c_k) * (size_t)n;
  const uint32_t base = b * 1024;
  uint32_t v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const uint32_t f = base + 64 * j + lane;
    v[j] = f < C ? cand[f] : 0;
  }
  const uint32_t lo0 = wave_lower_bound(rej, R, base);
  if (lane < 16) s_mask[lane] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t p0 = lo0;; p0 += 64) {
    const uint32_t p = p0 + lane;
    const uint32_t e = p < R ? rej[p] - base : 0xffffffffu;
    const bool in = e < 1024;
    if (in) atomicOr(&s_mask[e >> 6], 1ull << (e & 63));
    if (~__ballot(in)) break;  // the slice ends in these 64
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  uint32_t rb = lo0;  // rejected ranks below the current row
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const uint64_t M = s_mask[j];
    const uint32_t f = base + 64 * j + lane;
    if (f < C && !((M >> lane) & 1)) {
      const uint32_t accr = f - rb - (uint32_t)__builtin_popcountll(M & below);
      if (accr < kstart) r[i0 + accr] = v[j];  // always, for an exact resolution
    }
    rb += (uint32_t)__builtin_popcountll(M);
  }
  __builtin_amdgcn_wave_barrier();
}

// ---- one window step: commit the last resolved windows, then draw and classify -----
// Dense tile (32768 draws): wave w covers draws [w * 2048, (w + 1) * 2048); lane
// l computes AES blocks j * 64 + l (j < 8) of that range, i.e. draws
// j * 256 + 4 l + q, so rank order inside a wave is (j, l, q) and candidates
// are compacted with ballots into consecutive addresses.
// Sparse tile (131072 draws): wave w covers [w * 8192, (w + 1) * 8192) in four
// rounds of 2048 laid out as above; its candidates go to an LDS stash in rank
// order, so one ticket and one pair of look-backs serve four AES rounds.
// Candidate a (rank in the window) is accepted for sure if v < k - a (at most a
// acceptances precede it); otherwise it is ambiguous and goes to the resolver.
template <int COPIES>
__device__ inline void draw_round(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                  const ChainState &st, uint32_t W, uint32_t mask, uint32_t base,
                                  uint32_t (&v)[kDrawsPerThread]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const uint32_t d = base + j * 256 + lane * 4;
    if (d < W) {
      const uint64_t dabs = st.d0 + d;
      const uint4 blk = aes_ctr_block<COPIES>(A, rk, st.nonce + dabs / kWordsPerNonce,
                                              2 + (uint32_t)((dabs % kWordsPerNonce) / 4));
      v[4 * j + 0] = blk.x & mask;
      v[4 * j + 1] = blk.y & mask;
      v[4 * j + 2] = blk.z & mask;
      v[4 * j + 3] = blk.w & mask;
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) v[4 * j + q] = 0xffffffffu;  // never a candidate
    }
  }
}

template <int COPIES>
__global__ void __launch_bounds__(kGpThreads) k_gp_window(const uint32_t *__restrict__ tab,
                                                          const uint32_t *__restrict__ rk_groups,
                                                          ChainBufs B, uint32_t n, uint32_t mask,
                                                          uint32_t count, uint32_t draws_on) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  __shared__ uint32_t s_wc[kGpThreads / 64], s_wa[kGpThreads / 64];
  __shared__ uint32_t s_next, s_off, s_aoff;
  __shared__ unsigned long long s_mask[kGpThreads / 64][16];
  __shared__ uint32_t s_sd[kGpThreads / 64][kStash], s_sv[kGpThreads / 64][kStash];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = lanes_below();
  const uint32_t maskN = mask + 1;

  // tasks, in ticket order: the tiles of the current windows, then commit
  // chunks (16 blocks of 1024 ranks, one per wave) of the windows resolved
  // last, which workgroups without a tile (or done with theirs) take while
  // the tiles' look-backs wait
  uint32_t nc[3], nt[3];
  for (int c = 0; c < 3; c++) {
    nc[c] = B.st[c].c_valid ? (B.st[c].c_C + 16383) / 16384 : 0u;
    nt[c] = draws_on ? ntiles_of(B.st[c], n, maskN) : 0u;
  }
  const uint32_t ntile = nt[0] + nt[1] + nt[2];
  const uint32_t total = ntile + nc[0] + nc[1] + nc[2];
  // s_next holds the next task, taken one task ahead: a tile takes its
  // successor's ticket while its first look-back waits anyway
  if (threadIdx.x == 0 && total > 0) s_next = atomicAdd(&B.ctr[0], 1u);
  bool tables = false;
  for (;;) {
    if (total == 0) break;
    __syncthreads();
    const uint32_t task = s_next;
    __syncthreads();  // read by all before it is rewritten
    if (task >= total) break;
    if (task >= ntile) {
      int chain;
      uint32_t chunk;
      task_tile(nc, task - ntile, chain, chunk);
      commit_block(B, chain, chunk * (kGpThreads / 64) + wv, n, count, s_mask[wv]);
      if (threadIdx.x == 0) s_next = atomicAdd(&B.ctr[0], 1u);
      continue;
    }
    if (!tables) {
      aes_load_tables<COPIES>(lds, tab);
      tables = true;
    }
    AesCtx<COPIES> A(lds);
    int chain;
    uint32_t tile;
    task_tile(nt, task, chain, tile);
    ChainState &st = B.st[chain];
    const uint32_t kstart = n - st.i, W = st.W, seq = st.seq;
    const bool sparse = sparse_window(st, n, maskN);
    const uint32_t *rk = rk_groups + 60 * chain;
    uint32_t *cand_v = B.cand_v + ((size_t)st.buf * 3 + chain) * B.cand_cap;
    uint32_t *cand_d = B.cand_d + ((size_t)st.buf * 3 + chain) * B.cand_cap;
    const uint32_t cap = B.cand_cap;
    const bool can_end = W >= kstart;
    uint32_t *amb_a = B.amb_a + (size_t)chain * kAmbCap;
    int32_t *amb_u = B.amb_u + (size_t)chain * kAmbCap;
    uint32_t v[kDrawsPerThread];
    uint32_t wc = 0;  // this wave's candidates
    uint32_t wbase;
    if (sparse) {
      wbase = tile * kSuperTile + (uint32_t)wv * (kSuperTile / (kGpThreads / 64));
      for (uint32_t r = 0; r < kSuperRounds; r++) {
        const uint32_t rbase = wbase + r * 2048;
        draw_round<COPIES>(A, rk, st, W, mask, rbase, v);
#pragma unroll
        for (int j = 0; j < 8; j++) {
          uint64_t b[4];
          uint32_t pre = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            b[q] = __ballot(v[4 * j + q] < kstart);
            pre += __builtin_popcountll(b[q] & below);
          }
          if (!(b[0] | b[1] | b[2] | b[3])) continue;
          uint32_t i = wc + pre;
          const uint32_t d = rbase + j * 256 + lane * 4;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const uint32_t x = v[4 * j + q];
            if (x < kstart) {
              if (i < (uint32_t)kStash) {
                s_sd[wv][i] = d + q;
                s_sv[wv][i] = x;
              }
              i++;
            }
          }
#pragma unroll
          for (int q = 0; q < 4; q++) wc += __builtin_popcountll(b[q]);
        }
      }
      if (wc > (uint32_t)kStash && lane == 0) atomicOr(&st.ovf, 1u);  // window retried dense
    } else {
      wbase = tile * kTile + (uint32_t)wv * 2048;
      draw_round<COPIES>(A, rk, st, W, mask, wbase, v);
#pragma unroll
      for (int e = 0; e < kDrawsPerThread; e++) wc += __builtin_popcountll(__ballot(v[e] < kstart));
    }
    // candidate counts per wave, tile offset by look-back (wave 0 also takes
    // the next ticket meanwhile)
    if (lane == 0) s_wc[wv] = wc;
    __syncthreads();
    if (wv == 0) {
      uint32_t tk = 0;
      if (lane == 0) tk = atomicAdd(&B.ctr[0], 1u);
      uint32_t cnt = wave_sum(lane < kGpThreads / 64 ? s_wc[lane] : 0u);
      uint32_t off = look_back(B.stat_c + (size_t)chain * kMaxTiles, tile, seq, cnt, &st.err);
      if (lane == 0) {
        s_next = tk;
        s_off = off;
        if (tile == nt[chain] - 1) st.C_tot = off + cnt;
      }
    }
    __syncthreads();
    uint32_t run = s_off;
    for (int w = 0; w < wv; w++) run += s_wc[w];
    const uint32_t run0 = run;
    // candidate records in rank order, ambiguous counts
    uint32_t wa = 0;
    if (sparse) {
      const uint32_t ns = wc < (uint32_t)kStash ? wc : (uint32_t)kStash;
      for (uint32_t i0 = 0; i0 < ns; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool in = i < ns;
        const uint32_t d = in ? s_sd[wv][i] : 0u, x = in ? s_sv[wv][i] : 0u;
        const uint32_t a = run + i;
        if (in && a < cap) {
          cand_v[a] = x;
          if (can_end) cand_d[a] = d;
        }
        wa += __builtin_popcountll(__ballot(in && a + x >= kstart));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        uint64_t b[4];
        uint32_t pre = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          b[q] = __ballot(v[4 * j + q] < kstart);
          pre += __builtin_popcountll(b[q] & below);
        }
        if (!(b[0] | b[1] | b[2] | b[3])) continue;  // no candidate in this row
        uint32_t a = run + pre;
        const uint32_t d = wbase + j * 256 + lane * 4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t x = v[4 * j + q];
          const bool is_c = x < kstart;
          if (is_c && a < cap) {
            cand_v[a] = x;
            if (can_end) cand_d[a] = d + q;
          }
          wa += __builtin_popcountll(__ballot(is_c && a + x >= kstart));
          a += is_c;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) run += __builtin_popcountll(b[q]);
      }
    }
    if (lane == 0) s_wa[wv] = wa;
    __syncthreads();
    if (wv == 0) {
      uint32_t cnt = wave_sum(lane < kGpThreads / 64 ? s_wa[lane] : 0u);
      uint32_t off = look_back(B.stat_a + (size_t)chain * kMaxTiles, tile, seq, cnt, &st.err);
      if (lane == 0) {
        s_aoff = off;
        if (tile == nt[chain] - 1) st.m_tot = off + cnt;
      }
    }
    __syncthreads();
    // ambiguous records, in rank order
    uint32_t arun = s_aoff;
    for (int w = 0; w < wv; w++) arun += s_wa[w];
    run = run0;
    if (sparse) {
      const uint32_t ns = wc < (uint32_t)kStash ? wc : (uint32_t)kStash;
      for (uint32_t i0 = 0; i0 < ns && wa; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool in = i < ns;
        const uint32_t x = in ? s_sv[wv][i] : 0u;
        const uint32_t a = run + i;
        const bool amb = in && a + x >= kstart;
        const uint64_t ab = __ballot(amb);
        const uint32_t r = arun + __builtin_popcountll(ab & below);
        if (amb && r < kAmbCap) {
          amb_a[r] = a;
          amb_u[r] = (int32_t)(a + x - kstart);
        }
        arun += __builtin_popcountll(ab);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8 && wa; j++) {
        uint64_t b[4];
        uint32_t pre = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          b[q] = __ballot(v[4 * j + q] < kstart);
          pre += __builtin_popcountll(b[q] & below);
        }
        if (!(b[0] | b[1] | b[2] | b[3])) continue;  // no candidate in this row
        uint32_t a = run + pre;
        bool amb[4];
        uint32_t apre = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t x = v[4 * j + q];
          const bool is_c = x < kstart;
          amb[q] = is_c && a + x >= kstart;
          const uint64_t ab = __ballot(amb[q]);
          apre += __builtin_popcountll(ab & below);
          a += is_c;
        }
        // second walk: write this lane's ambiguous records after the lower lanes'
        a = run + pre;
        uint32_t r = arun + apre;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t x = v[4 * j + q];
          if (amb[q]) {
            if (r < kAmbCap) {
              amb_a[r] = a;
              amb_u[r] = (int32_t)(a + x - kstart);
            }
            r++;
          }
          a += x < kstart;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          run += __builtin_popcountll(b[q]);
          arun += __builtin_popcountll(__ballot(amb[q]));
        }
      }
    }
  }
  // the last workgroup out resets the ticket and retires the committed windows
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t fin = atomicAdd(&B.ctr[1], 1u);
    if (fin == gridDim.x - 1) {
      B.ctr[0] = 0;
      B.ctr[1] = 0;
      for (int c = 0; c < 3; c++) B.st[c].c_valid = 0;
      __threadfence();
    }
  }
}

// ---- exact resolution of the ambiguous candidates, chain advance -------------------
// Ambiguous candidate t (list index t, rank a_t, T_t = k - v_t <= a_t) is
// accepted iff fewer than T_t acceptances precede it, i.e. iff more than
// U_t = a_t - T_t >= 0 of the ambiguous candidates before it are rejected.
// At most t of them are, so t <= U_t means rejected for sure.  The first
// stage applies that test to all ambiguous candidates, streaming global
// memory row by row (item t = row * 1024 + thread, compacted with ballots).
// Every earlier candidate of an undecided one is either rejected for sure or
// undecided, so an undecided item is accepted iff fewer than T' = t - U_t of
// the undecided items before it are accepted: the second stage walks that
// short list in order (in LDS, one wave, 64 items per step).
__global__ void __launch_bounds__(kGpThreads) k_gp_resolve(ChainBufs B, uint32_t n,
                                                           uint32_t count, uint32_t maskN) {
  constexpr uint32_t kRows = kAmbCap / kGpThreads;  // 64
  __shared__ int32_t s_thr[kL2Cap];
  __shared__ uint8_t s_dec[kL2Cap];   // 1 accepted, 2 rejected
  __shared__ uint32_t s_cnt[kRows * (kGpThreads / 64)];
  __shared__ uint32_t s_w[kGpThreads / 64];
  __shared__ uint32_t s_l2;
  const int chain = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = lanes_below();
  ChainState *st = B.st + chain;
  if (st->done) return;
  const uint32_t ovf = st->ovf;
  __syncthreads();
  if (ovf) {
    // a sparse tile's stash overflowed: the same window again, dense
    if (threadIdx.x == 0) {
      st->ovf = 0;
      st->force_dense = 1;
      st->c_valid = 0;
      st->seq++;
    }
    return;
  }
  const uint32_t kstart = n - st->i;
  const uint32_t *amb_a = B.amb_a + (size_t)chain * kAmbCap;
  const int32_t *amb_u = B.amb_u + (size_t)chain * kAmbCap;
  uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
  const uint32_t m = st->m_tot, C = st->C_tot;
  const bool over = m > kAmbCap || C > B.cand_cap;
  const uint32_t rows = over ? 0 : (m + kGpThreads - 1) / kGpThreads;
  // level 1: bit q of undm = item q * 1024 + threadIdx.x undecided, of rejm rejected
  uint64_t undm = 0, rejm = 0;
  for (uint32_t q0 = 0; q0 < rows; q0 += 8) {
    int32_t u[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t t = (q0 + j) * kGpThreads + threadIdx.x;
      u[j] = t < m ? amb_u[t] : 0;
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t q = q0 + j, t = q * kGpThreads + threadIdx.x;
      if (q >= rows) break;
      const bool in = t < m, rj = in && (int32_t)t <= u[j], ud = in && !rj;
      rejm |= (uint64_t)rj << q;
      undm |= (uint64_t)ud << q;
      const uint64_t b = __ballot(ud);
      if (lane == 0) s_cnt[q * (kGpThreads / 64) + wv] = __builtin_popcountll(b);
    }
  }
  // undecided positions: exclusive scan over (row, wave) counts
  __syncthreads();
  const uint32_t nent = rows * (kGpThreads / 64);
  {
    const uint32_t x = threadIdx.x < nent ? s_cnt[threadIdx.x] : 0u;
    uint32_t l2;
    const uint32_t ex = block_scan(x, s_w, l2);
    __syncthreads();
    if (threadIdx.x < nent) s_cnt[threadIdx.x] = ex;
    if (threadIdx.x == 0) s_l2 = l2;
    __syncthreads();
  }
  if (over || s_l2 > kL2Cap) {
    // too many candidates or ambiguous ones: retry this window at a quarter size
    if (threadIdx.x == 0) {
      uint32_t w = st->W / 4;
      st->W = w >= kTile ? w / kTile * kTile : (w >= kWinMin ? w / kWinMin * kWinMin : kWinMin);
      st->c_valid = 0;
      st->seq++;
    }
    return;
  }
  // undecided items continue as a type A list with T' = t - U
  for (uint32_t q0 = 0; q0 < rows; q0 += 8) {
    int32_t u[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t q = q0 + j;
      u[j] = q < rows && ((undm >> q) & 1) ? amb_u[q * kGpThreads + threadIdx.x] : 0;
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t q = q0 + j;
      if (q >= rows) break;
      const bool ud = (undm >> q) & 1;
      const uint64_t b = __ballot(ud);
      if (!ud) continue;
      const uint32_t t = q * kGpThreads + threadIdx.x;
      const uint32_t p = s_cnt[q * (kGpThreads / 64) + wv] + __builtin_popcountll(b & below);
      s_thr[p] = (int32_t)t - u[j];
      s_dec[p] = 0;
    }
  }
  __syncthreads();
  // The undecided items, in order: item i is accepted iff fewer than T'_i of
  // the undecided items before it are accepted.  One wave walks the list in
  // chunks of 64: given the acceptances before the chunk (acc), the chunk's
  // decisions are the fixed point of d_i = [acc + #(accepted before i in the
  // chunk) < T'_i], reached by iterating from the optimistic guess with one
  // ballot per round (item i is exact after round i, so at most 64 rounds; a
  // repeated ballot is the unique fixed point, i.e. the sequential answer).
  if (wv == 0) {
    const uint32_t len = s_l2;
    uint32_t acc = 0;
    for (uint32_t c0 = 0; c0 < len; c0 += 64) {
      const uint32_t i = c0 + lane;
      const bool in = i < len;
      const int32_t T = in ? s_thr[i] : 0;
      uint64_t dm = __ballot(in && (int32_t)acc < T);
      for (int round = 0; round < 65; round++) {
        const uint32_t before = acc + (uint32_t)__builtin_popcountll(dm & below);
        const uint64_t nm = __ballot(in && (int32_t)before < T);
        if (nm == dm) break;
        dm = nm;
      }
      if (in) s_dec[i] = (dm >> lane) & 1 ? 1 : 2;
      acc += (uint32_t)__builtin_popcountll(dm);
    }
  }
  __syncthreads();
  // rejected items: level-1 rejects and the undecided ones the walk rejected
  for (uint32_t q = 0; q < rows; q++) {
    const bool ud = (undm >> q) & 1;
    const uint64_t b = __ballot(ud);
    if (ud) {
      const uint32_t p = s_cnt[q * (kGpThreads / 64) + wv] + __builtin_popcountll(b & below);
      if (s_dec[p] == 2) rejm |= 1ull << q;
    }
  }
  __syncthreads();
  for (uint32_t q = 0; q < rows; q++) {
    const uint64_t b = __ballot((rejm >> q) & 1);
    if (lane == 0) s_cnt[q * (kGpThreads / 64) + wv] = __builtin_popcountll(b);
  }
  __syncthreads();
  uint32_t R;
  {
    const uint32_t x = threadIdx.x < nent ? s_cnt[threadIdx.x] : 0u;
    const uint32_t ex = block_scan(x, s_w, R);
    __syncthreads();
    if (threadIdx.x < nent) s_cnt[threadIdx.x] = ex;
    __syncthreads();
  }
  // rejected ranks in ascending order (the ambiguous list is in rank order);
  // loads go out 8 rows at a time
  for (uint32_t q0 = 0; q0 < rows; q0 += 8) {
    uint32_t a[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t q = q0 + j;
      a[j] = q < rows && ((rejm >> q) & 1) ? amb_a[q * kGpThreads + threadIdx.x] : 0;
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t q = q0 + j;
      if (q >= rows) break;
      const bool rj = (rejm >> q) & 1;
      const uint64_t b = __ballot(rj);
      if (rj) rej[s_cnt[q * (kGpThreads / 64) + wv] + __builtin_popcountll(b & below)] = a[j];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t acc = C - R;
    if (acc > kstart) atomicOr(&st->err, 4u);
    const uint32_t buf = st->buf;
    st->c_valid = 1;
    st->c_i0 = st->i;
    st->c_k = st->k;
    st->c_C = C;
    st->c_R = R;
    st->c_buf = buf;
    st->buf = buf ^ 1u;
    st->seq++;
    st->force_dense = 0;
    if (acc >= kstart) {
      // the permutation ends here: rank a* of the kstart-th acceptance is
      // kstart-1 shifted past every rejected rank at or before it (rej is
      // strictly increasing, so rej[q] - q is non-decreasing: the rejected
      // ranks passed are those with rej[q] - q <= kstart - 1)
      uint32_t qlo = 0, qhi = R;
      while (qlo < qhi) {
        const uint32_t mid = (qlo + qhi) >> 1;
        if (rej[mid] - mid <= kstart - 1) qlo = mid + 1;
        else qhi = mid;
      }
      const uint32_t a = kstart - 1 + qlo;
      const uint32_t e = B.cand_d[((size_t)buf * 3 + chain) * B.cand_cap + a];
      const uint64_t D_end = st->d0 + e + 1;
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

// Each target's list sorted by step index, descending, in place: the lists
// come out of k_fy_link nearly sorted (steps are linked in roughly increasing
// order), so a list is only checked, and the rare unsorted one gets an
// insertion sort.  Then parent(t) = the largest step < t that targeted t:
// the head, or the one after it if the head is step t itself (r_t = 0).
__global__ void k_fy_sort(uint32_t n, int32_t *__restrict__ head, int32_t *__restrict__ next,
                          int32_t *__restrict__ parent, uint32_t *__restrict__ err) {
  int32_t *hd = head + (size_t)blockIdx.y * n, *nx = next + (size_t)blockIdx.y * n;
  int32_t *par = parent + (size_t)blockIdx.y * n;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    int32_t h = hd[t];
    bool sorted = true;
    uint32_t len = 0;
    for (int32_t q = h; q >= 0; len++) {
      const int32_t nq = nx[q];
      if (nq > q || len >= n) {  // (len >= n: a cycle, never for valid links)
        sorted = false;
        break;
      }
      q = nq;
    }
    if (!sorted) {
      int32_t sh = -1;  // sorted list built so far (descending)
      uint32_t guard = 0;
      for (int32_t q = h; q >= 0 && guard < n; guard++) {
        const int32_t nq = nx[q];
        if (sh < 0 || q > sh) {
          nx[q] = sh;
          sh = q;
        } else {
          int32_t p = sh;
          for (int32_t pn = nx[p]; pn > q; pn = nx[p]) p = pn;
          nx[q] = nx[p];
          nx[p] = q;
        }
        q = nq;
      }
      if (guard >= n) atomicOr(err, 32u);
      h = sh;
      hd[t] = h;
    }
    par[t] = h == (int32_t)t ? nx[t] : h;
  }
}

// Position j is final after step j; it receives what position t_j held just
// before step j: V(q) for q = the last earlier step that targeted t_j (the next
// one in t_j's sorted list), or t_j itself.  V(m), the value at position m just
// before step m, is V(parent(m)), or m when no earlier step targeted m.
__global__ void k_fy_walk(uint32_t n, const uint32_t *__restrict__ r_all, size_t chain_stride,
                          const int32_t *__restrict__ next, const int32_t *__restrict__ parent,
                          int32_t *fwd0, int32_t *fwd1, int32_t *fwd2, int32_t *inv0, int32_t *inv1,
                          int32_t *inv2) {
  const uint32_t *r = r_all + blockIdx.y * chain_stride;
  const int32_t *nx = next + (size_t)blockIdx.y * n, *par = parent + (size_t)blockIdx.y * n;
  int32_t *fwd = blockIdx.y == 0 ? fwd0 : blockIdx.y == 1 ? fwd1 : fwd2;
  int32_t *inv = blockIdx.y == 0 ? inv0 : blockIdx.y == 1 ? inv1 : inv2;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t t = j + r[j];
    if (t >= n) continue;  // flagged by k_fy_link
    const int32_t q = nx[j];
    uint32_t val = t;
    if (q >= 0) {
      uint32_t m = (uint32_t)q;
      for (uint32_t guard = 0; guard < n; guard++) {
        const int32_t p = par[m];
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
  // candidates per window: about 256 sqrt(k) <= 256 sqrt(n); 4x headroom (a
  // window over capacity is retried smaller)
  const uint32_t cand_cap = (uint32_t)std::min<double>(
      kWinMax, std::max<double>(1 << 16, 1024.0 * std::ceil(std::sqrt((double)n))));
  B.cand_cap = cand_cap;
  B.cand_v = (uint32_t *)c->scratch_get("gp_cand_v", 6ull * cand_cap * 4);
  B.cand_d = (uint32_t *)c->scratch_get("gp_cand_d", 6ull * cand_cap * 4);
  B.stat_c = (uint64_t *)c->scratch_get("gp_stat_c", 3ull * kMaxTiles * 8);
  B.stat_a = (uint64_t *)c->scratch_get("gp_stat_a", 3ull * kMaxTiles * 8);
  B.ctr = (uint32_t *)c->scratch_get("gp_ctr", 64);
  B.amb_a = (uint32_t *)c->scratch_get("gp_amb_a", 3ull * kAmbCap * 4);
  B.amb_u = (int32_t *)c->scratch_get("gp_amb_u", 3ull * kAmbCap * 4);
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
    init[g].seq = 1;  // the status words are cleared to seq 0 below
    nat[(size_t)g * (count + 1)] = init[g].nonce;
  }
  ORQG_HIP(hipMemsetAsync(B.stat_c, 0, 3ull * kMaxTiles * 8, c->stream));
  ORQG_HIP(hipMemsetAsync(B.stat_a, 0, 3ull * kMaxTiles * 8, c->stream));
  ORQG_HIP(hipMemsetAsync(B.ctr, 0, 64, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.st, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.nonce_at, nat.data(), nat.size() * 8, hipMemcpyHostToDevice, c->stream));

  const unsigned grid = std::max(1, c->num_cu);
  uint32_t host_flag[5] = {0, 0, 0, 0, 0};
  const int kBatch = 16;
  // The window kernels take all their state from device memory, so a batch of
  // windows is one static launch sequence: it is captured once as a graph.
  // Each step commits the previously resolved windows, draws and classifies
  // the current ones; a last commit-only step leaves every resolved window of
  // the batch committed, so the completed-permutation count read after a
  // batch covers committed offsets only.
  auto enqueue_batch = [&]() {
    for (int w = 0; w <= kBatch; w++) {
      hipLaunchKernelGGL(k_gp_window<kCopies>, dim3(grid), dim3(kGpThreads), 0, c->stream,
                         c->d_t0, c->rk(kGroup0), B, n, mask, count, (uint32_t)(w < kBatch));
      if (w < kBatch)
        hipLaunchKernelGGL(k_gp_resolve, dim3(3), dim3(kGpThreads), 0, c->stream, B, n, count,
                           maskN);
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
  int32_t *parent = (int32_t *)c->scratch_get("fy_parent", (size_t)n * 4 + 64);
  uint32_t *err = (uint32_t *)c->scratch_get("fy_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  ORQG_HIP(hipEventRecord(c->ev_aux, c->stream));
  ORQG_HIP(hipStreamWaitEvent(c->stream2, c->ev_aux, 0));
  // while windows run: one 4-wave workgroup per two CUs, enough random loads
  // in flight to finish in the windows' shadow and room beside them for the
  // window and resolve workgroups, which need whole CUs
  const unsigned fgrid_bg = grid_for(n, 256, std::max(1, c->num_cu / 2));
  unsigned fgrid = fgrid_bg;
  const size_t chain_stride = (size_t)count * n;
  uint32_t launched[3] = {0, 0, 0};
  auto resolve_swaps = [&](int g, uint32_t upto) {
    for (uint32_t k = launched[g]; k < upto; k++) {
      const uint32_t *r = B.r_out + (size_t)g * chain_stride + (size_t)k * n;
      ORQG_HIP(hipMemsetAsync(head, 0xff, (size_t)n * 4, c->stream2));
      hipLaunchKernelGGL(k_fy_link, dim3(fgrid, 1), dim3(256), 0, c->stream2, n, r, chain_stride,
                         head, next, err);
      hipLaunchKernelGGL(k_fy_sort, dim3(fgrid, 1), dim3(256), 0, c->stream2, n, head, next,
                         parent, err);
      hipLaunchKernelGGL(k_fy_walk, dim3(fgrid, 1), dim3(256), 0, c->stream2, n, r, chain_stride,
                         next, parent, fresh[k].fwd[g], fresh[k].fwd[g], fresh[k].fwd[g],
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
  // the windows are done: what is left runs alone on the whole GPU
  fgrid = grid_for(n, 256, c->num_cu * 8);
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
