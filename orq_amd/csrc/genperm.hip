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
// independent chains, processed by the same launches) window by window.  One
// launch (k_gp_step) per step s:
//   * its workgroups 3.. draw window s: keystream -> draws; draws with
//     v >= k_hi can never be accepted (k only decreases); the others
//     ("candidates") get their rank a in the window through a decoupled
//     look-back over the tiles and are written in rank order.  Candidate a is
//     accepted for sure if v < k_lo - a (at most a acceptances precede it);
//     otherwise it is "listed" and written, in rank order too, to the
//     resolver's list (a second look-back).  They also commit the window
//     resolved at step s - 1 (tasks taken after the tiles): accepted candidate
//     a becomes r_{i0 + a - #rejected before a};
//   * its workgroups 0..2 (one per chain) resolve window s - 1 exactly (sure
//     acceptances and rejections, then an in-order walk of the rest), build
//     the sorted list of rejected ranks, and advance the chain (next window,
//     or the next permutation at the refill boundary after the last
//     acceptance).
// The window size keeps the expected number of listed candidates within the
// resolver's capacity; an overflowing window is retried at a quarter of its
// size, down to kWinMin draws, which always fits.
//
// (2) uses that position i is final after step i: Fisher-Yates leaves at i
// what position t_i = i + r_i held just before step i, found by walking "last
// earlier step that targeted this position" links.
#include <algorithm>
#include <cmath>
#include <cstdlib>
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
// sparse windows (candidate density <= 3/128) use super-tiles of 2, 4 or 8 AES
// rounds per workgroup, their few candidates stashed in LDS between the rounds
constexpr uint32_t kSuperRounds2 = 8;  // the most rounds per super-tile
constexpr int kStash = 192;                             // candidates per wave
constexpr uint64_t kRefillWords = 1ull << 20;  // gen_perm buffer: 2^20 uint32
constexpr uint64_t kWordsPerNonce = 1ull << 18;

// Windows are drawn and resolved in a two-stage pipeline: step s draws window
// s while window s - 1 is resolved.  Window s is drawn before the exact number
// of open positions at its start is known: window s - 1 accepts C - m + A of
// its C candidates (m listed, A of them accepted), so k lies in
// [k_lo, k_hi] = n - i - C + m - [L, 0] (i: exact acceptances before window
// s - 1, L: a bound on A), and window s classifies with those bounds
// (candidates: v < k_hi; sure acceptances: v < k_lo - a).  A window that may
// end the permutation or needs a retry is not followed by a speculative one:
// its chain skips a step and restarts from the exact state; a bound that
// fails is caught by the resolver, which redraws the window exactly.

// A chain's exact state before window s (ring slot s & 1), written by the
// resolve launch of step s.
struct ResState {
  uint64_t nonce;        // nonce where the current permutation's draws start
  uint64_t d0;           // restart: draw offset (in the permutation) of the next window
  uint32_t k;            // index of the permutation being generated
  uint32_t i;            // acceptances in it before window s
  uint32_t W;            // restart: draws of the next window
  uint32_t force_dense;  // restart: dense tiles (a sparse tile's stash overflowed)
  uint32_t done;
  uint32_t void_next;    // window s was drawn on a state that did not hold: discard it
  // window s - 1 resolved, committed by the window launch of step s + 1
  uint32_t c_valid, c_i0, c_k, c_C, c_R, c_cbuf, c_rbuf;
};

// Window s of a chain (ring slot s & 1), planned by every workgroup of the
// window launch of step s from slot s - 1 (the same answer everywhere).
struct WinDesc {
  uint64_t d0;         // draw offset in the permutation
  uint32_t active;     // drawn in this step
  uint32_t W;          // draws
  uint32_t klo, khi;   // bounds on the open positions at its start
  uint32_t sparse;     // super-tiles
  uint32_t nt;         // tiles
  uint32_t seq;        // look-back tag (step + 1)
  uint32_t can_end;    // W >= klo: the window may complete the permutation
  uint32_t C, m;       // candidates, ambiguous candidates (written by the last tile)
  uint32_t ovf;        // = seq when a sparse tile's stash overflowed
};

struct ChainBufs {
  ResState *rs;        // [2][3]
  WinDesc *wd;         // [2][3]
  uint32_t *ctl;       // [0] step, [2] tile ticket, [3] finished workgroups
  uint32_t *err;       // device-side check flags
  uint32_t l2cap;      // undecided items a resolve takes (kL2Cap; smaller only to test retries)
  uint32_t bound_shift;  // k_lo margin divided by 2^bound_shift (0; > 0 only to test redraws)
  uint32_t *cand_v;    // [3][3][cand_cap] candidates in rank order: v (buffer s % 3)
  uint32_t *cand_d;    // [3][3][cand_cap] ... and the draw offset in the window, written
                       // only when the window can end the permutation
  uint32_t cand_cap;
  uint64_t *stat_c;    // [3][kMaxTiles] look-back status of the candidate counts
  uint64_t *stat_a;    // [3][kMaxTiles] ... of the ambiguous counts
  uint32_t *amb_a;     // [2][3][kAmbCap] ambiguous candidates in rank order: rank
  int32_t *amb_u;      // [2][3][kAmbCap] ... and rank - (k_lo - v)
  uint32_t *rej;       // [2][3][kAmbCap] ranks of rejected candidates, ascending
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

// Plan window s of `chain` from the state of step s - 1.  Speculates (window
// s - 1 still being resolved) only when window s - 1 can neither complete the
// permutation (acceptances <= C < k_lo) nor be retried for capacity; a
// window the resolver discards or retries leaves an exact restart.
__device__ inline WinDesc plan_window(const ChainBufs &B, int chain, uint32_t s, uint32_t n,
                                      uint32_t maskN, uint32_t nwork) {
  // copies, so that every field is loaded at once rather than one load per
  // short-circuited test (slot s - 1 is not written during this step)
  const ResState r = B.rs[((s - 1) & 1) * 3 + chain];
  const WinDesc p = B.wd[((s - 1) & 1) * 3 + chain];
  WinDesc w;
  w.active = 0;
  w.C = w.m = w.ovf = 0;
  w.seq = s + 1;
  w.nt = 0;
  if (r.done) return w;
  bool force_dense = false;
  if (p.active && !r.void_next) {
    if (p.ovf == p.seq || p.m > kAmbCap || p.C > B.cand_cap || p.C >= p.klo)
      return w;  // wait for the exact state
    // Window s - 1 accepts C - m + A of its candidates, A = its listed
    // candidates that are accepted: its sure acceptances listed because of
    // the bounds (about C (khi - klo) / k) and the accepted undecided ones (at
    // most the undecided, about C^3 / (6 k^2)).  k_lo assumes A below a bound
    // with ample margin; the resolver checks it and redraws if it fails.
    const double kp = (double)(n - r.i), Cd = (double)p.C;
    const double mu = Cd * Cd * Cd / (6.0 * kp * kp) + Cd * (double)(p.khi - p.klo) / kp;
    const double Lf = mu + 8.0 * sqrt(mu) + 256.0;
    const uint32_t L = (Lf >= (double)p.m ? p.m : (uint32_t)Lf) >> B.bound_shift;
    // the undecided items a bound L adds to the next resolve (about m L / k)
    // must leave room in its list
    if (mu + (double)p.m * (double)L / kp > (double)(B.l2cap / 2)) return w;
    w.khi = n - (r.i + p.C) + p.m;
    w.klo = w.khi - L;
    w.d0 = p.d0 + p.W;
    w.W = choose_window(w.khi, maskN);
  } else {
    w.klo = w.khi = n - r.i;
    w.d0 = r.d0;
    w.W = r.W;
    force_dense = r.force_dense != 0;
  }
  w.active = 1;
  // super-tiles when candidates are rare and the window still gives every
  // workgroup at least one (otherwise dense tiles spread it over more CUs)
  // R rounds per super-tile: the most of 8, 4, 2 for which a wave expects at
  // most 96 candidates (R * 2048 * k_hi / maskN; its stash of 192 is then ten
  // standard deviations above) and the window still gives every drawing
  // workgroup a super-tile
  w.sparse = 0;
  if (!force_dense)
    for (uint32_t R = kSuperRounds2; R >= 2; R >>= 1)
      if ((uint64_t)w.khi * 2048 * R <= 96ull * maskN &&
          (uint64_t)w.W >= (uint64_t)kTile * R * nwork) {
        w.sparse = R;
        break;
      }
  const uint32_t td = w.sparse ? kTile * w.sparse : kTile;
  // whole rounds of tiles over the drawing workgroups (a chain may draw
  // alone in a step); retried windows below one round keep their size
  const uint64_t round = (uint64_t)td * nwork;
  if (w.W >= round) w.W = (uint32_t)(w.W / round * round);
  w.can_end = w.W >= w.klo;
  w.nt = (w.W + td - 1) / td;
  return w;
}

// When the three windows are dense and their tiles fill r whole rounds over
// the nwork drawing workgroups plus a last round under a quarter full, the
// last round's tiles come off the windows with the most tiles (one at a time,
// the lowest chain first on ties), so that the step runs r rounds.  Any
// window size is exact (the next window starts where this one ends); only
// windows of whole tiles of at least two tiles are shortened.
__device__ inline void trim_last_round(WinDesc (&p)[3], uint32_t nwork) {
  uint32_t T = 0;
  for (int c = 0; c < 3; c++) {
    if (!p[c].active || p[c].nt == 0) continue;
    if (p[c].sparse || p[c].W % kTile != 0) return;
    T += p[c].nt;
  }
  if (T <= nwork) return;
  const uint32_t extra = T % nwork;
  if (extra == 0 || extra * 4 > nwork) return;
  for (uint32_t e = 0; e < extra; e++) {
    int best = -1;
    for (int c = 0; c < 3; c++)
      if (p[c].active && p[c].nt > 1 && (best < 0 || p[c].nt > p[best].nt)) best = c;
    if (best < 0) break;
    p[best].nt--;
    p[best].W -= kTile;
  }
  for (int c = 0; c < 3; c++)
    if (p[c].active) p[c].can_end = p[c].W >= p[c].klo;
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
// A wait is bounded in time (the 100 MHz constant clock), not in polls, so
// that a GPU shared with other streams' kernels cannot trip it early: a wait
// longer than kSpinTicks reports error flag 16 instead of hanging.
constexpr uint64_t kSpinTicks = 400000000ull;  // 4 s

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
      const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
      for (;;) {
        const uint64_t s =
            __hip_atomic_load(stat + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t hi = (uint32_t)(s >> 32);
        if ((hi >> 2) == seq && (hi & 3u)) {
          kind = hi & 3u;
          val = (uint32_t)s;
          break;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > kSpinTicks) {  // never expected
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
__device__ inline void commit_block(const ChainBufs &B, const ResState &st, int chain, uint32_t b,
                                    uint32_t n, uint32_t count, unsigned long long *s_mask) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = lanes_below();
  const uint32_t C = st.c_C, R = st.c_R, i0 = st.c_i0, kstart = n - i0;
  if (b * 1024 >= C) return;
  const uint32_t *cand = B.cand_v + ((size_t)st.c_cbuf * 3 + chain) * B.cand_cap;
  const uint32_t *rej = B.rej + ((size_t)st.c_rbuf * 3 + chain) * kAmbCap;
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
// Sparse tile (R = 4 or 8 rounds, R * 32768 draws): wave w covers
// [w * R * 2048, (w + 1) * R * 2048) in R rounds of 2048 laid out as above; its
// candidates go to an LDS stash in rank order, so one ticket and one pair of
// look-backs serve R AES rounds.
// Candidate a (rank in the window) is accepted for sure if v < k_lo - a (at
// most a acceptances precede it); otherwise it is listed for the resolver.
template <int COPIES>
__device__ inline void draw_round(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                  uint64_t nonce, uint64_t d0, uint32_t W, uint32_t mask,
                                  uint32_t base, uint32_t (&v)[kDrawsPerThread]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const uint32_t d = base + j * 256 + lane * 4;
    if (d < W) {
      const uint64_t dabs = d0 + d;
      const uint4 blk = aes_ctr_block<COPIES>(A, rk, nonce + dabs / kWordsPerNonce,
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

constexpr uint32_t kRows = kAmbCap / kGpThreads;  // 64
constexpr uint32_t kWaves = kGpThreads / 64;
// the resolver loads items (row, thread) of its ambiguous lists without a
// per-item condition (a load behind a branch is waited for on its own): every
// row < kRows indexes inside the list's kAmbCap entries
static_assert(kRows * kGpThreads == kAmbCap, "resolver rows must tile the ambiguous lists");
// LDS of a resolving workgroup (it uses the AES table space of the step kernel)
struct ResolveLds {
  int32_t thr[kL2Cap];
  uint64_t cnt[kRows * kWaves];  // (sure acceptances << 32 | undecided) before each
                                 // (row, wave), then rejected before each
  uint64_t w[kWaves];
  uint8_t dec[kL2Cap];           // 1 accepted, 2 rejected
  uint32_t l2;
};

__device__ void resolve_window(const ChainBufs &B, int chain, uint32_t step, uint32_t n,
                               uint32_t count, uint32_t maskN, ResolveLds &L);
static_assert(sizeof(ResolveLds) <= AesLds<kCopies>::kWords * 4, "resolver LDS exceeds the tables'");

template <int COPIES>
__global__ void __launch_bounds__(kGpThreads) k_gp_step(const uint32_t *__restrict__ tab,
                                                          const uint32_t *__restrict__ rk_groups,
                                                          ChainBufs B, uint32_t n, uint32_t mask,
                                                          uint32_t count, uint32_t draws_on) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  __shared__ uint32_t s_wc[kGpThreads / 64], s_wa[kGpThreads / 64];
  __shared__ uint32_t s_next, s_off, s_aoff;
  __shared__ unsigned long long s_mask[kGpThreads / 64][16];
  __shared__ uint32_t s_sd[kGpThreads / 64][kStash], s_sv[kGpThreads / 64][kStash];
  __shared__ WinDesc s_plan[3];
  __shared__ uint64_t s_nonce[3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = lanes_below();
  const uint32_t maskN = mask + 1;
  const uint32_t step = B.ctl[0];
  const ResState *rs_prev = B.rs + ((step - 1) & 1) * 3;
  // workgroups 0..2 resolve window step - 1 (one chain each) on their own
  // CUs while the others draw window step (draws_on = 0: a commit-only
  // launch, no step)
  const uint32_t nres = draws_on ? 3u : 0u, nwork = gridDim.x - nres;
  if (blockIdx.x < nres) {
    resolve_window(B, blockIdx.x, step, n, count, maskN, *reinterpret_cast<ResolveLds *>(lds));
  } else {
  // the windows resolved at the previous step (committed below): loaded here,
  // unconditionally, so that their latency overlaps the planning
  uint32_t cvalid[3], cC[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cvalid[c] = rs_prev[c].c_valid;
    cC[c] = rs_prev[c].c_C;
  }
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    WinDesc w = plan_window(B, c, step, n, maskN, nwork);
    if (!draws_on) {
      w.active = 0;
      w.nt = 0;
    }
    s_plan[c] = w;
    s_nonce[c] = rs_prev[c].nonce;
  }
  __syncthreads();
  // the three windows' dense tiles as whole rounds over the drawing
  // workgroups, where the last round would be under a quarter full: its
  // tiles are taken off the largest windows (every workgroup computes the
  // same, from the same plans)
  if (threadIdx.x == 0 && draws_on) trim_last_round(s_plan, nwork);
  __syncthreads();
  if (threadIdx.x < 3 && draws_on && blockIdx.x == nres) {
    const int c = threadIdx.x;
    const WinDesc &w = s_plan[c];
    WinDesc &o = B.wd[(step & 1) * 3 + c];
    o.d0 = w.d0;
    o.active = w.active;
    o.W = w.W;
    o.klo = w.klo;
    o.khi = w.khi;
    o.sparse = w.sparse;
    o.nt = w.nt;
    o.seq = w.seq;
    o.can_end = w.can_end;
  }
  // tasks, in ticket order: the tiles of the current windows, then commit
  // chunks (16 blocks of 1024 ranks, one per wave) of the windows resolved
  // at the previous step, which workgroups without a tile (or done with
  // theirs) take while the tiles' look-backs wait
  uint32_t nc[3], nt[3];
  for (int c = 0; c < 3; c++) {
    nc[c] = cvalid[c] ? (cC[c] + 16383) / 16384 : 0u;
    nt[c] = s_plan[c].nt;
  }
  const uint32_t ntile = nt[0] + nt[1] + nt[2];
  const uint32_t total = ntile + nc[0] + nc[1] + nc[2];
  // s_next holds the next task, taken one task ahead: a tile takes its
  // successor's ticket while its first look-back waits anyway
  if (threadIdx.x == 0 && total > 0) s_next = atomicAdd(&B.ctl[2], 1u);
  bool tables = false;
  for (;;) {
    if (total == 0) break;
    __syncthreads();
    const uint32_t task = __builtin_amdgcn_readfirstlane(s_next);  // uniform: scalar registers
    __syncthreads();  // read by all before it is rewritten
    if (task >= total) break;
    if (task >= ntile) {
      int chain;
      uint32_t chunk;
      task_tile(nc, task - ntile, chain, chunk);
      commit_block(B, rs_prev[chain], chain, chunk * (kGpThreads / 64) + wv, n, count,
                   s_mask[wv]);
      if (threadIdx.x == 0) s_next = atomicAdd(&B.ctl[2], 1u);
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
    const WinDesc &pl = s_plan[chain];
    // candidates: v < khi; sure acceptances: a + v < klo
    const uint32_t khi = __builtin_amdgcn_readfirstlane(pl.khi);
    const uint32_t klo = __builtin_amdgcn_readfirstlane(pl.klo);
    const uint32_t W = __builtin_amdgcn_readfirstlane(pl.W);
    const uint32_t seq = __builtin_amdgcn_readfirstlane(pl.seq);
    const uint32_t srounds = __builtin_amdgcn_readfirstlane(pl.sparse);  // 0: dense
    const bool sparse = srounds != 0;
    const uint64_t d0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(pl.d0 >> 32)) << 32) |
                        __builtin_amdgcn_readfirstlane((uint32_t)pl.d0);
    const uint64_t nonce =
        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(s_nonce[chain] >> 32)) << 32) |
        __builtin_amdgcn_readfirstlane((uint32_t)s_nonce[chain]);
    WinDesc &wd = B.wd[(step & 1) * 3 + chain];
    const uint32_t *rk = rk_groups + 60 * chain;
    const size_t cb = (size_t)(step % 3) * 3 + chain;
    uint32_t *cand_v = B.cand_v + cb * B.cand_cap;
    uint32_t *cand_d = B.cand_d + cb * B.cand_cap;
    const uint32_t cap = B.cand_cap;
    const bool can_end = __builtin_amdgcn_readfirstlane(pl.can_end) != 0;
    uint32_t *amb_a = B.amb_a + ((size_t)(step & 1) * 3 + chain) * kAmbCap;
    int32_t *amb_u = B.amb_u + ((size_t)(step & 1) * 3 + chain) * kAmbCap;
    uint32_t v[kDrawsPerThread];
    uint32_t wc = 0;  // this wave's candidates
    uint32_t wbase;
    if (sparse) {
      wbase = tile * srounds * kTile + (uint32_t)wv * srounds * 2048u;
      for (uint32_t r = 0; r < srounds; r++) {
        const uint32_t rbase = wbase + r * 2048;
        draw_round<COPIES>(A, rk, nonce, d0, W, mask, rbase, v);
#pragma unroll
        for (int j = 0; j < 8; j++) {
          uint64_t b[4];
          uint32_t pre = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            b[q] = __ballot(v[4 * j + q] < khi);
            pre += __builtin_popcountll(b[q] & below);
          }
          if (!(b[0] | b[1] | b[2] | b[3])) continue;
          uint32_t i = wc + pre;
          const uint32_t d = rbase + j * 256 + lane * 4;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const uint32_t x = v[4 * j + q];
            if (x < khi) {
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
      if (wc > (uint32_t)kStash && lane == 0) wd.ovf = seq;  // window retried dense
    } else {
      wbase = tile * kTile + (uint32_t)wv * 2048;
      draw_round<COPIES>(A, rk, nonce, d0, W, mask, wbase, v);
#pragma unroll
      for (int e = 0; e < kDrawsPerThread; e++) wc += __builtin_popcountll(__ballot(v[e] < khi));
    }
    // candidate counts per wave, tile offset by look-back (wave 0 also takes
    // the next ticket meanwhile)
    if (lane == 0) s_wc[wv] = wc;
    __syncthreads();
    if (wv == 0) {
      uint32_t tk = 0;
      if (lane == 0) tk = atomicAdd(&B.ctl[2], 1u);
      uint32_t cnt = wave_sum(lane < kGpThreads / 64 ? s_wc[lane] : 0u);
      uint32_t off = look_back(B.stat_c + (size_t)chain * kMaxTiles, tile, seq, cnt, B.err);
      if (lane == 0) {
        s_next = tk;
        s_off = off;
        if (tile == nt[chain] - 1) wd.C = off + cnt;
      }
    }
    __syncthreads();
    uint32_t run = __builtin_amdgcn_readfirstlane(s_off);
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
        wa += __builtin_popcountll(__ballot(in && a + x >= klo));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        uint64_t b[4];
        uint32_t pre = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          b[q] = __ballot(v[4 * j + q] < khi);
          pre += __builtin_popcountll(b[q] & below);
        }
        if (!(b[0] | b[1] | b[2] | b[3])) continue;  // no candidate in this row
        uint32_t a = run + pre;
        const uint32_t d = wbase + j * 256 + lane * 4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t x = v[4 * j + q];
          const bool is_c = x < khi;
          if (is_c && a < cap) {
            cand_v[a] = x;
            if (can_end) cand_d[a] = d + q;
          }
          wa += __builtin_popcountll(__ballot(is_c && a + x >= klo));
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
      uint32_t off = look_back(B.stat_a + (size_t)chain * kMaxTiles, tile, seq, cnt, B.err);
      if (lane == 0) {
        s_aoff = off;
        if (tile == nt[chain] - 1) wd.m = off + cnt;
      }
    }
    __syncthreads();
    // ambiguous records, in rank order
    uint32_t arun = __builtin_amdgcn_readfirstlane(s_aoff);
    for (int w = 0; w < wv; w++) arun += s_wa[w];
    run = run0;
    if (sparse) {
      const uint32_t ns = wc < (uint32_t)kStash ? wc : (uint32_t)kStash;
      for (uint32_t i0 = 0; i0 < ns && wa; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool in = i < ns;
        const uint32_t x = in ? s_sv[wv][i] : 0u;
        const uint32_t a = run + i;
        const bool amb = in && a + x >= klo;
        const uint64_t ab = __ballot(amb);
        const uint32_t r = arun + __builtin_popcountll(ab & below);
        if (amb && r < kAmbCap) {
          amb_a[r] = a;
          amb_u[r] = (int32_t)(a + x - klo);
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
          b[q] = __ballot(v[4 * j + q] < khi);
          pre += __builtin_popcountll(b[q] & below);
        }
        if (!(b[0] | b[1] | b[2] | b[3])) continue;  // no candidate in this row
        uint32_t a = run + pre;
        bool amb[4];
        uint32_t apre = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t x = v[4 * j + q];
          const bool is_c = x < khi;
          amb[q] = is_c && a + x >= klo;
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
              amb_u[r] = (int32_t)(a + x - klo);
            }
            r++;
          }
          a += x < khi;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          run += __builtin_popcountll(b[q]);
          arun += __builtin_popcountll(__ballot(amb[q]));
        }
      }
    }
  }
  }
  // the last workgroup out resets the ticket, retires the committed windows
  // and advances the step
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t fin = atomicAdd(&B.ctl[3], 1u);
    if (fin == gridDim.x - 1) {
      B.ctl[2] = 0;
      B.ctl[3] = 0;
      for (int c = 0; c < 3; c++) B.rs[((step - 1) & 1) * 3 + c].c_valid = 0;
      if (draws_on) B.ctl[0] = step + 1;
      __threadfence();
    }
  }
}

// ---- exact resolution of the ambiguous candidates, chain advance -------------------
// With k the exact open positions at the window start, listed candidate t
// (list index t, rank a_t, T_t = k - v_t) is accepted iff fewer than T_t
// acceptances precede it, i.e. iff more than U_t = a_t - T_t of the
// candidates before it are rejected; only listed ones can be.  U_t < 0 means
// accepted for sure (listed because the window was drawn with k_lo < k), and
// at most t of the listed items before t are rejected, so t <= U_t means
// rejected for sure.  An undecided item t is accepted iff fewer than
// T' = t - SA_t - U_t of the undecided items before it are accepted (SA_t:
// sure acceptances listed before it; T' <= 0 rejects): the walk decides that
// short list in order (in LDS, one wave, 64 items per step).  Item t of the
// list is handled by thread t % 1024 as its row t / 1024; loads go out 16
// rows at a time.
__device__ void resolve_window(const ChainBufs &B, int chain, uint32_t step, uint32_t n,
                               uint32_t count, uint32_t maskN, ResolveLds &L) {
  constexpr uint32_t kHalf = 16;  // rows per batch of loads
  constexpr uint32_t kNb = kRows / kHalf;
  int32_t *s_thr = L.thr;
  uint8_t *s_dec = L.dec;
  uint64_t *s_cnt = L.cnt;
  uint64_t *s_w = L.w;
  uint32_t &s_l2 = L.l2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = lanes_below();
  // copies of slot step - 1 (not written during this step): every field is
  // loaded at once, not one load per short-circuited test
  const ResState r = B.rs[((step - 1) & 1) * 3 + chain];
  const WinDesc p = B.wd[((step - 1) & 1) * 3 + chain];
  ResState &o = B.rs[(step & 1) * 3 + chain];
  const uint32_t kx = n - r.i;  // exact open positions at the window start
  // nothing drawn, or drawn on a state that did not hold: carry the state over
  bool skip = r.done || !p.active || r.void_next;
  if (threadIdx.x == 0) {
    o = r;
    o.c_valid = 0;
    o.void_next = 0;
    if (!skip && (p.ovf == p.seq || p.klo > kx || kx > p.khi)) {
      // a sparse tile's stash overflowed (the same window again, dense), or
      // the bounds it was drawn with do not hold (the same window again,
      // exact); the window drawn after it is discarded
      o.d0 = p.d0;
      o.W = p.W;
      o.force_dense = p.ovf == p.seq;
      o.void_next = 1;
    }
  }
  if (!skip && (p.ovf == p.seq || p.klo > kx || kx > p.khi)) skip = true;
  const int32_t delta = (int32_t)(kx - p.klo);
  const size_t ab = ((size_t)((step - 1) & 1) * 3 + chain) * kAmbCap;
  const uint32_t *amb_a = B.amb_a + ab;
  const int32_t *amb_u = B.amb_u + ab;
  uint32_t *rej = B.rej + ab;
  const uint32_t m = p.m, C = p.C;
  const bool over = m > kAmbCap || C > B.cand_cap;
  const uint32_t rows = skip || over ? 0 : (m + kGpThreads - 1) / kGpThreads;
  const uint32_t nent = rows * kWaves;
  // level 1: bit q of undm = item q * 1024 + threadIdx.x undecided, of rejm
  // rejected, of sam accepted for sure
  uint64_t undm = 0, rejm = 0, sam = 0;
#pragma unroll 1
  for (uint32_t h = 0; h < kNb; h++) {
    if (h * kHalf >= rows) break;
    int32_t u[kHalf];
#pragma unroll
    for (uint32_t j = 0; j < kHalf; j++) {
      const uint32_t q = h * kHalf + j, t = q * kGpThreads + threadIdx.x;
      const int32_t x = amb_u[t];  // unconditional (t < kAmbCap): loads in flight together
      u[j] = q < rows && t < m ? x - delta : 0;
    }
#pragma unroll
    for (uint32_t j = 0; j < kHalf; j++) {
      const uint32_t q = h * kHalf + j, t = q * kGpThreads + threadIdx.x;
      if (q >= rows) break;
      const bool in = t < m, sa = in && u[j] < 0, rj = in && !sa && (int32_t)t <= u[j],
                 ud = in && !sa && !rj;
      rejm |= (uint64_t)rj << q;
      undm |= (uint64_t)ud << q;
      sam |= (uint64_t)sa << q;
      const uint64_t bu = __ballot(ud), bs = __ballot(sa);
      if (lane == 0)
        s_cnt[q * kWaves + wv] =
            ((uint64_t)__builtin_popcountll(bs) << 32) | (uint64_t)__builtin_popcountll(bu);
    }
  }
  // undecided positions and sure acceptances before each (row, wave): one
  // exclusive scan over both counts
  __syncthreads();
  {
    const uint64_t x = threadIdx.x < nent ? s_cnt[threadIdx.x] : 0ull;
    uint64_t tot;
    const uint64_t ex = block_scan<uint64_t>(x, s_w, tot);
    __syncthreads();
    if (threadIdx.x < nent) s_cnt[threadIdx.x] = ex;
    if (threadIdx.x == 0) s_l2 = (uint32_t)tot;
    __syncthreads();
  }
  if (!skip && (over || s_l2 > B.l2cap)) {
    // too many candidates or ambiguous ones: retry this window at a quarter size
    if (threadIdx.x == 0) {
      const uint32_t w = p.W / 4;
      o.d0 = p.d0;
      o.W = w >= kTile ? w / kTile * kTile : (w >= kWinMin ? w / kWinMin * kWinMin : kWinMin);
      o.force_dense = 0;
      o.void_next = 1;
    }
    skip = true;
  }
  if (!skip) {
    // undecided items with their thresholds T' = t - SA_t - U_t
#pragma unroll 1
    for (uint32_t h = 0; h < kNb; h++) {
      if (h * kHalf >= rows) break;
      int32_t u[kHalf];
#pragma unroll
      for (uint32_t j = 0; j < kHalf; j++) {
        const uint32_t q = h * kHalf + j, t = q * kGpThreads + threadIdx.x;
        const int32_t x = amb_u[t];
        u[j] = q < rows && ((undm >> q) & 1) ? x - delta : 0;
      }
#pragma unroll
      for (uint32_t j = 0; j < kHalf; j++) {
        const uint32_t q = h * kHalf + j;
        if (q >= rows) break;
        const bool ud = (undm >> q) & 1;
        const uint64_t bu = __ballot(ud), bs = __ballot((sam >> q) & 1);
        if (ud) {
          const uint32_t t = q * kGpThreads + threadIdx.x;
          const uint64_t e = s_cnt[q * kWaves + wv];
          const uint32_t pos = (uint32_t)e + __builtin_popcountll(bu & below);
          const uint32_t sat = (uint32_t)(e >> 32) + __builtin_popcountll(bs & below);
          s_thr[pos] = (int32_t)(t - sat) - u[j];
          s_dec[pos] = 0;
        }
      }
    }
    // ranks of the first batch's items that may be rejected, in flight
    // during the walk
    uint32_t a[kHalf];
#pragma unroll
    for (uint32_t q = 0; q < kHalf; q++) {
      const uint32_t t = q * kGpThreads + threadIdx.x;
      const uint32_t x = amb_a[t];
      a[q] = ((rejm | undm) >> q) & 1 ? x : 0u;  // (no bit is set at or past rows)
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
    // rejected items: the sure ones and the undecided ones the walk rejected
#pragma unroll
    for (uint32_t q = 0; q < kRows; q++) {
      if (q >= rows) break;
      const bool ud = (undm >> q) & 1;
      const uint64_t bu = __ballot(ud);
      if (ud && s_dec[(uint32_t)s_cnt[q * kWaves + wv] + __builtin_popcountll(bu & below)] == 2)
        rejm |= 1ull << q;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kRows; q++) {
      if (q >= rows) break;
      const uint64_t b = __ballot((rejm >> q) & 1);
      if (lane == 0) s_cnt[q * kWaves + wv] = __builtin_popcountll(b);
    }
    __syncthreads();
    uint64_t R64;
    {
      const uint64_t x = threadIdx.x < nent ? s_cnt[threadIdx.x] : 0ull;
      const uint64_t ex = block_scan<uint64_t>(x, s_w, R64);
      __syncthreads();
      if (threadIdx.x < nent) s_cnt[threadIdx.x] = ex;
      __syncthreads();
    }
    const uint32_t R = (uint32_t)R64;
    // rejected ranks in ascending order (the list is in rank order); the
    // later rows load their ranks now
#pragma unroll 1
    for (uint32_t h = 0; h < kNb; h++) {
      if (h * kHalf >= rows) break;
      if (h > 0) {
#pragma unroll
        for (uint32_t q = 0; q < kHalf; q++) {
          const uint32_t qq = h * kHalf + q, t = qq * kGpThreads + threadIdx.x;
          const uint32_t x = amb_a[t];
          a[q] = (rejm >> qq) & 1 ? x : 0u;
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < kHalf; q++) {
        const uint32_t qq = h * kHalf + q;
        if (qq >= rows) break;
        const bool rj = (rejm >> qq) & 1;
        const uint64_t b = __ballot(rj);
        if (rj) rej[(uint32_t)s_cnt[qq * kWaves + wv] + __builtin_popcountll(b & below)] = a[q];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      const uint32_t acc = C - R;
      if (acc > kx) atomicOr(B.err, 4u);
      o.c_valid = 1;
      o.c_i0 = r.i;
      o.c_k = r.k;
      o.c_C = C;
      o.c_R = R;
      o.c_cbuf = (step - 1) % 3;
      o.c_rbuf = (step - 1) & 1;
      o.force_dense = 0;
      if (acc >= kx) {
        // the permutation ends here: rank a* of the kx-th acceptance is kx-1
        // shifted past every rejected rank at or before it (rej is strictly
        // increasing, so rej[q] - q is non-decreasing: the rejected ranks
        // passed are those with rej[q] - q <= kx - 1)
        uint32_t qlo = 0, qhi = R;
        while (qlo < qhi) {
          const uint32_t mid = (qlo + qhi) >> 1;
          if (rej[mid] - mid <= kx - 1) qlo = mid + 1;
          else qhi = mid;
        }
        const uint32_t a = kx - 1 + qlo;
        const uint32_t e = B.cand_d[((size_t)((step - 1) % 3) * 3 + chain) * B.cand_cap + a];
        const uint64_t D_end = p.d0 + e + 1;
        const uint32_t k = r.k;
        B.D[(size_t)chain * count + k] = D_end;
        // 4 nonces per 2^20-word refill, one refill up front
        o.nonce = r.nonce + 4 * (1 + D_end / kRefillWords);
        o.k = k + 1;
        o.i = 0;
        o.d0 = 0;
        B.nonce_at[(size_t)chain * (count + 1) + k + 1] = o.nonce;
        if (o.k >= count) o.done = 1;
        o.W = choose_window(n, maskN);
        o.void_next = 1;
      } else {
        o.i = r.i + acc;
        o.d0 = p.d0 + p.W;
        o.W = choose_window(n - o.i, maskN);
      }
    }
  }
}

// after a batch: all done, check flags, permutations completed per chain (the
// state the last resolve launch wrote)
__global__ void k_gp_done(ChainBufs B, uint32_t *flag) {
  if (threadIdx.x == 0) {
    const ResState *r = B.rs + ((B.ctl[0] - 1) & 1) * 3;
    flag[0] = r[0].done & r[1].done & r[2].done;
    flag[1] = *B.err;
    for (int g = 0; g < 3; g++) flag[2 + g] = r[g].k;
    for (int g = 0; g < 3; g++) reinterpret_cast<uint64_t *>(flag + 6)[g] = r[g].nonce;
  }
}

// ---- swap resolution ----------------------------------------------------------
// Step j swaps positions j and t_j = j + r_j.  Steps are chained per target
// position (head[t] -> most recently linked step -> next[...]).  A list is
// sorted (descending) iff every link put a larger step in front of a smaller
// one; a link that did not marks its target dirty, and only dirty lists are
// sorted.
__device__ inline void fy_link(uint32_t n, const uint32_t *__restrict__ r, int32_t *__restrict__ hd,
                               int32_t *__restrict__ nx, uint8_t *__restrict__ dirty,
                               uint32_t *__restrict__ err) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t t = j + r[j];
    if (t >= n) {
      atomicOr(err, 8u);
      nx[j] = -1;
      continue;
    }
    const int32_t old = atomicExch(&hd[t], (int32_t)j);
    nx[j] = old;
    if (old > (int32_t)j) dirty[t] = 1;
  }
}

__global__ void k_fy_link(uint32_t n, const uint32_t *__restrict__ r, int32_t *__restrict__ head,
                          int32_t *__restrict__ next, uint8_t *__restrict__ dirty,
                          uint32_t *__restrict__ err) {
  fy_link(n, r, head, next, dirty, err);
}

// Each dirty target's list sorted by step index, descending, in place (the
// flag is cleared for the next permutation): the lists come out of k_fy_link
// nearly sorted, since steps are linked in roughly increasing order.  A
// list's first kSortRegs members are read into registers; an unsorted list
// that short (the usual case: the late targets' lists) is sorted there by a
// sorting network and relinked, a longer one gets an insertion sort in place.
// Then parent(t) = the largest step < t that targeted t: the head, or the one
// after it if the head is step t itself (r_t = 0).
constexpr int kSortRegs = 16;

__device__ inline void cswap_desc(int32_t &a, int32_t &b) {
  const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
  a = hi;
  b = lo;
}

__device__ inline void fy_sort(uint32_t n, int32_t *__restrict__ hd, int32_t *__restrict__ nx,
                               int32_t *__restrict__ par, uint8_t *__restrict__ dirty,
                               uint32_t *__restrict__ err) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    int32_t h = hd[t];
    if (!dirty[t]) {
      par[t] = h == (int32_t)t ? nx[t] : h;
      continue;
    }
    dirty[t] = 0;
    int32_t mem[kSortRegs];
    bool sorted = true;
    uint32_t len = 0;
    int32_t q = h;
#pragma unroll
    for (int k = 0; k < kSortRegs; k++) {
      mem[k] = -1;  // -1 sorts last (descending)
      if (q >= 0) {
        mem[k] = q;
        const int32_t nq = nx[q];
        if (nq > q) sorted = false;
        q = nq;
        len++;
      }
    }
    // the rest of a longer list is only checked
    for (; q >= 0 && sorted; len++) {
      const int32_t nq = nx[q];
      if (nq > q || len >= n) {  // (len >= n: a cycle, never for valid links)
        sorted = false;
        break;
      }
      q = nq;
    }
    if (!sorted && q < 0 && len <= (uint32_t)kSortRegs) {
      // odd-even transposition network, descending; the -1 padding stays last
#pragma unroll
      for (int r = 0; r < kSortRegs; r++) {
#pragma unroll
        for (int k = r & 1; k + 1 < kSortRegs; k += 2) cswap_desc(mem[k], mem[k + 1]);
      }
#pragma unroll
      for (int k = 0; k < kSortRegs; k++)
        if (mem[k] >= 0) nx[mem[k]] = k + 1 < kSortRegs ? mem[k + 1] : -1;
      h = mem[0];
      hd[t] = h;
    } else if (!sorted) {
      int32_t sh = -1;  // sorted list built so far (descending)
      uint32_t guard = 0;
      for (int32_t q2 = h; q2 >= 0 && guard < n; guard++) {
        const int32_t nq = nx[q2];
        if (sh < 0 || q2 > sh) {
          nx[q2] = sh;
          sh = q2;
        } else {
          int32_t p = sh;
          for (int32_t pn = nx[p]; pn > q2; pn = nx[p]) p = pn;
          nx[q2] = nx[p];
          nx[p] = q2;
        }
        q2 = nq;
      }
      if (guard >= n) atomicOr(err, 32u);
      h = sh;
      hd[t] = h;
    }
    par[t] = h == (int32_t)t ? nx[t] : h;
  }
}

__global__ void k_fy_sort(uint32_t n, int32_t *__restrict__ head, int32_t *__restrict__ next,
                          int32_t *__restrict__ parent, uint8_t *__restrict__ dirty,
                          uint32_t *__restrict__ err) {
  fy_sort(n, head, next, parent, dirty, err);
}

// Position j is final after step j; it receives what position t_j held just
// before step j: V(q) for q = the last earlier step that targeted t_j (the next
// one in t_j's sorted list), or t_j itself.  V(m), the value at position m just
// before step m, is V(parent(m)), or m when no earlier step targeted m.
__device__ inline void fy_walk(uint32_t n, const uint32_t *__restrict__ r,
                               const int32_t *__restrict__ nx, const int32_t *__restrict__ par,
                               int32_t *__restrict__ fwd, int32_t *__restrict__ inv) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t t = j + r[j];
    if (t >= n) continue;  // flagged by fy_link
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

__global__ void k_fy_walk(uint32_t n, const uint32_t *__restrict__ r, const int32_t *__restrict__ next,
                          const int32_t *__restrict__ parent, int32_t *fwd, int32_t *inv) {
  fy_walk(n, r, next, parent, fwd, inv);
}

// ---- small permutations ----------------------------------------------------------
// A permutation whose draws fit one refill (fewer than 2^20: the usual case up
// to a few times 10^4 rows) starts exactly 4 nonces after the previous one, so
// every local permutation of a reserve -- of several contexts at once -- is
// drawn by its own workgroup in one launch, from a start nonce known in
// advance (a permutation that needed more is detected afterwards from its
// draw count and the later ones of its chain are drawn again).  The workgroup
// draws kSmallDraws words per pass (one AES block per thread) into LDS; then
// its first wave scans them 64 at a time: with k open positions, the draws
// v < k are candidates, and candidate a (rank among them) is accepted for
// sure if v < k - a; a group of 64 with any other candidate is walked lane by
// lane.  The swap resolution then runs for every local permutation in one
// launch sequence (grid y = permutation).
// draws per pass: one AES block per thread of waves 1.. (wave 0 scans)
constexpr uint32_t kSmallDraws = (kGpThreads - 64) * 4;
constexpr uint32_t kScanGroup = 20;  // 64-draw chunks a scanning lane holds at once
static_assert(kSmallDraws % (64 * kScanGroup) == 0, "scan groups must tile a pass");

struct SmallTask {
  const uint32_t *rk;  // the group PRG's round keys
  uint64_t nonce;      // nonce where the permutation's draws start
  uint32_t *r;         // [n] accepted offsets r_i
  uint64_t *D;         // draws consumed
  uint64_t max_draws;  // guard: a permutation drawing more sets an error flag
  uint32_t n, mask;
};

// Pass p's draws go to buffer p & 1: waves 1.. draw pass p + 1 while wave 0
// scans pass p.  Wave 0 leaves the open count after pass p in s_k[p & 1],
// which every wave reads after the pass's barrier (two slots: wave 0 may
// already be scanning pass p + 1 while a slower wave still reads slot p & 1).
template <int COPIES>
__global__ void __launch_bounds__(kGpThreads) k_gp_small(const uint32_t *__restrict__ tab,
                                                           const SmallTask *__restrict__ tasks,
                                                           uint32_t *__restrict__ err) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  __shared__ __align__(16) uint32_t buf[2][kSmallDraws];
  __shared__ uint32_t s_k[2];
  aes_load_tables<COPIES>(lds, tab);
  AesCtx<COPIES> A(lds);
  const SmallTask t = tasks[blockIdx.x];
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t below = lanes_below();
  auto draw = [&](uint64_t base, uint32_t *dst) {
    const uint32_t q = threadIdx.x - 64;
    const uint64_t d = base + q * 4;
    const uint4 blk = aes_ctr_block<COPIES>(A, t.rk, t.nonce + d / kWordsPerNonce,
                                            2 + (uint32_t)((d % kWordsPerNonce) / 4));
    reinterpret_cast<uint4 *>(dst)[q] =
        make_uint4(blk.x & t.mask, blk.y & t.mask, blk.z & t.mask, blk.w & t.mask);
  };
  uint32_t k = t.n, i = 0;  // open positions, acceptances (wave 0)
  uint64_t D = 0;
  if (threadIdx.x >= 64) draw(0, buf[0]);
  __syncthreads();
  for (uint64_t p = 0;; p++) {
    const uint64_t base = p * kSmallDraws;
    if (threadIdx.x >= 64) {
      draw(base + kSmallDraws, buf[(p + 1) & 1]);
    } else {
      const uint32_t *b = buf[p & 1];
      for (uint32_t g0 = 0; g0 < kSmallDraws / 64 && k > 0; g0 += kScanGroup) {
        uint32_t vv[kScanGroup];
#pragma unroll
        for (uint32_t u = 0; u < kScanGroup; u++) vv[u] = b[(g0 + u) * 64 + lane];
#pragma unroll
        for (uint32_t u = 0; u < kScanGroup; u++) {
          const uint32_t v = vv[u];
          const uint64_t cand = __ballot(v < k);
          if (!cand) continue;
          const uint32_t a = (uint32_t)__popcll(cand & below);
          uint64_t acc = __ballot(v < k && v + a < k);
          if (acc != cand) {
            acc = 0;
            uint32_t kk = k;
            for (uint64_t rest = cand; rest && kk;) {
              const int L = __ffsll((long long)rest) - 1;
              rest &= rest - 1;
              const uint32_t vL = (uint32_t)__builtin_amdgcn_readlane((int)v, L);
              if (vL < kk) {
                acc |= 1ull << L;
                kk--;
              }
            }
          }
          if ((acc >> lane) & 1ull) t.r[i + (uint32_t)__popcll(acc & below)] = v;
          const uint32_t na = (uint32_t)__popcll(acc);
          i += na;
          k -= na;
          if (k == 0) {
            D = base + (g0 + u) * 64 + (uint64_t)(63 - __clzll((long long)acc)) + 1;
            break;
          }
        }
      }
      if (threadIdx.x == 0) s_k[p & 1] = k;
    }
    __syncthreads();
    if (s_k[p & 1] == 0) break;
    if (base + 2 * kSmallDraws >= t.max_draws) {
      if (threadIdx.x == 0) atomicOr(err, 64u);
      return;
    }
  }
  if (threadIdx.x == 0) *t.D = D;
}

struct FyTask {
  const uint32_t *r;
  int32_t *head, *next, *parent;
  uint8_t *dirty;
  int32_t *fwd, *inv;
  uint32_t n;
};

__global__ void k_fyb_link(const FyTask *__restrict__ tasks, uint32_t *__restrict__ err) {
  const FyTask t = tasks[blockIdx.y];
  fy_link(t.n, t.r, t.head, t.next, t.dirty, err);
}
__global__ void k_fyb_sort(const FyTask *__restrict__ tasks, uint32_t *__restrict__ err) {
  const FyTask t = tasks[blockIdx.y];
  fy_sort(t.n, t.head, t.next, t.parent, t.dirty, err);
}
__global__ void k_fyb_walk(const FyTask *__restrict__ tasks) {
  const FyTask t = tasks[blockIdx.y];
  fy_walk(t.n, t.r, t.next, t.parent, t.fwd, t.inv);
}

}  // namespace

void free_perm(ShardedPerm &p) {
  if (p.ready) (void)hipEventDestroy(p.ready);
  p.ready = nullptr;
  for (int g = 0; g < 3; g++) p.fwd[g] = p.inv[g] = nullptr;
  p.slab.reset();
}

namespace {

// RunTime::generate_permutations (runtime.h:749-788): the new permutations
// dealt to the worker threads in contiguous runs, count / T each and one more
// for the first count % T threads.  Returns (first, count) per thread.
std::vector<std::pair<uint32_t, uint32_t>> thread_runs(const orqg_ctx *c, uint32_t count) {
  const uint32_t T = (uint32_t)c->threads, base = count / T, rem = count % T;
  std::vector<std::pair<uint32_t, uint32_t>> out;
  uint32_t k = 0;
  for (uint32_t t = 0; t < T; t++) {
    const uint32_t m = base + (t < rem ? 1 : 0);
    out.push_back({k, m});
    k += m;
  }
  return out;
}

// One context's share of a reserve: the permutations it still needs.
struct Reserve {
  orqg_ctx *c = nullptr;
  uint32_t n = 0, count = 0;
  std::vector<ShardedPerm> fresh;
};

// The queue bookkeeping of PermutationManager::reserve and the device arrays
// of the new permutations; false when nothing is left to generate.
bool reserve_begin(orqg_ctx *c, size_t n_, size_t count_, Reserve &R) {
  check_swaps(c);  // the previous call's swap resolution (its scratch is reused)
  if (!c->queue.empty() && c->stored_size != n_) {
    for (auto &p : c->queue) free_perm(p);
    c->queue.clear();
  }
  c->stored_size = n_;
  if (count_ <= c->queue.size()) return false;
  count_ -= c->queue.size();
  if (n_ >= (1ull << 31)) throw Error(ORQG_E_ARG, "reserve: n must be < 2^31");
  const uint32_t n = (uint32_t)n_, count = (uint32_t)count_;
  R.c = c;
  R.n = n;
  R.count = count;
  // the six arrays of every new permutation in one allocation (a served
  // stream of many sizes then costs one allocator entry per reserve, not six
  // per permutation); a consumer of the last of them waits for that one's
  // swap resolution, which is the last second-stream work on the slab
  R.fresh.assign(count, ShardedPerm());
  const size_t stride = ((size_t)n * 4 + 16 + 255) / 256 * 256;
  void *slab_p = dev_alloc(c, stride * 6 * count);
  std::shared_ptr<void> slab(slab_p, [](void *q) { dev_free(q); });
  for (uint32_t k = 0; k < count; k++) {
    ShardedPerm &p = R.fresh[k];
    p.n = n;
    p.slab = slab;
    for (int g = 0; g < 3; g++) {
      p.fwd[g] = (int32_t *)((char *)slab_p + stride * (6ull * k + g));
      p.inv[g] = (int32_t *)((char *)slab_p + stride * (6ull * k + 3 + g));
    }
  }
  if (n == 0) {
    // gen_perm still fills its buffer once: 4 nonces per local permutation,
    // on the group PRGs of the thread that makes it
    const auto runs = thread_runs(c, count);
    for (int t = 0; t < c->threads; t++) {
      if (!runs[t].second) continue;
      ThreadScope ts(c, t);
      for (int g = 0; g < 3; g++) c->nonce[kGroup0 + g] += 4ull * runs[t].second;
    }
    for (auto &p : R.fresh) c->queue.push_back(p);
    return false;
  }
  return true;
}

void reserve_end(Reserve &R) {
  for (auto &p : R.fresh) R.c->queue.push_back(p);
  R.fresh.clear();
}

// The windowed pipeline over the whole GPU, for one context's reserve.
void gen_windows(orqg_ctx *c, Reserve &R) {
  const uint32_t n = R.n, count = R.count;
  std::vector<ShardedPerm> &fresh = R.fresh;
  uint32_t bits = 0;
  for (uint32_t t = n; t; t >>= 1) bits++;
  const uint32_t maskN = 1u << bits;
  const uint32_t mask = maskN - 1;

  TimedRegion tr(c, "gen_perm");
  ChainBufs B;
  B.rs = (ResState *)c->scratch_get("gp_rs", 6 * sizeof(ResState));
  B.wd = (WinDesc *)c->scratch_get("gp_wd", 6 * sizeof(WinDesc));
  B.ctl = (uint32_t *)c->scratch_get("gp_ctl", 64);
  B.err = (uint32_t *)c->scratch_get("gp_err", 16);
  // candidates per window: about 256 sqrt(k) <= 256 sqrt(n); 4x headroom (a
  // window over capacity is retried smaller)
  const uint32_t cand_cap = (uint32_t)std::min<double>(
      kWinMax, std::max<double>(1 << 16, 1024.0 * std::ceil(std::sqrt((double)n))));
  B.cand_cap = cand_cap;
  // test hooks (environment, read per call): a smaller undecided-list cap
  // forces window retries, a smaller k_lo margin forces bound failures and
  // redraws; the permutations stay exact either way
  B.l2cap = kL2Cap;
  B.bound_shift = 0;
  if (const char *e = std::getenv("ORQG_GP_L2CAP")) B.l2cap = std::max(64, std::min<int>((int)kL2Cap, std::atoi(e)));
  if (const char *e = std::getenv("ORQG_GP_BOUND_SHIFT")) B.bound_shift = (uint32_t)std::max(0, std::min(31, std::atoi(e)));
  B.cand_v = (uint32_t *)c->scratch_get("gp_cand_v", 9ull * cand_cap * 4);
  B.cand_d = (uint32_t *)c->scratch_get("gp_cand_d", 9ull * cand_cap * 4);
  B.stat_c = (uint64_t *)c->scratch_get("gp_stat_c", 3ull * kMaxTiles * 8);
  B.stat_a = (uint64_t *)c->scratch_get("gp_stat_a", 3ull * kMaxTiles * 8);
  B.amb_a = (uint32_t *)c->scratch_get("gp_amb_a", 6ull * kAmbCap * 4);
  B.amb_u = (int32_t *)c->scratch_get("gp_amb_u", 6ull * kAmbCap * 4);
  B.rej = (uint32_t *)c->scratch_get("gp_rej", 6ull * kAmbCap * 4);
  B.r_out = (uint32_t *)c->scratch_get("gp_r", 3ull * count * n * 4);
  B.D = (uint64_t *)c->scratch_get("gp_D", 3ull * count * 8);
  B.nonce_at = (uint64_t *)c->scratch_get("gp_nonce_at", 3ull * (count + 1) * 8);
  uint32_t *flag = (uint32_t *)c->scratch_get("gp_flag", 64);

  // the exact state before window 0 sits in ring slot 1 (step -1), with no
  // window drawn at step -1
  ResState init[6];
  WinDesc wd0[6];
  std::memset(init, 0, sizeof(init));
  std::memset(wd0, 0, sizeof(wd0));
  std::vector<uint64_t> nat(3ull * (count + 1), 0);
  for (int g = 0; g < 3; g++) {
    ResState &r = init[3 + g];
    r.nonce = c->nonce[kGroup0 + g];
    r.W = choose_window(n, maskN);
    nat[(size_t)g * (count + 1)] = r.nonce;
  }
  ORQG_HIP(hipMemsetAsync(B.stat_c, 0, 3ull * kMaxTiles * 8, c->stream));
  ORQG_HIP(hipMemsetAsync(B.stat_a, 0, 3ull * kMaxTiles * 8, c->stream));
  ORQG_HIP(hipMemsetAsync(B.ctl, 0, 64, c->stream));
  ORQG_HIP(hipMemsetAsync(B.err, 0, 16, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.rs, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.wd, wd0, sizeof(wd0), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.nonce_at, nat.data(), nat.size() * 8, hipMemcpyHostToDevice, c->stream));

  // one workgroup per CU: three resolve, the others draw (a workgroup fills a
  // CU's LDS)
  const unsigned grid = std::max(4, c->num_cu);
  uint32_t host_flag[12] = {0};
  // steps per graph batch: large permutations take ~50 steps each, so long
  // batches save host round trips; small ones finish in a few steps, and the
  // steps of a batch left after the last permutation are empty launches
  const int kBatch = n >= (1u << 22) ? 128 : n >= (1u << 20) ? 64 : 16;
  // The kernel takes all its state from device memory, so a batch of steps
  // is one static launch sequence, built once as a graph: each step draws
  // window s (committing the windows resolved at step s - 1 as well) and
  // resolves window s - 1; a last commit-only launch leaves every resolved
  // window of the batch committed, so the completed-permutation count read
  // after a batch covers committed offsets.  The graph is built node by node
  // rather than captured from a stream: a capture in one context's thread
  // breaks (and leaves its stream stuck capturing) when another thread makes
  // a null-stream call meanwhile.
  struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~GraphGuard() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard;
  ORQG_HIP(hipGraphCreate(&guard.g, 0));
  {
    const uint32_t *tab = c->d_t0;
    const uint32_t *rkg = c->rk(kGroup0);
    uint32_t on = 1, off = 0;
    void *args_on[] = {(void *)&tab, (void *)&rkg, (void *)&B, (void *)&n, (void *)&mask,
                       (void *)&count, (void *)&on};
    void *args_off[] = {(void *)&tab, (void *)&rkg, (void *)&B, (void *)&n, (void *)&mask,
                        (void *)&count, (void *)&off};
    hipGraphNode_t prev = nullptr;
    for (int w = 0; w <= kBatch; w++) {
      hipKernelNodeParams kp;
      std::memset(&kp, 0, sizeof(kp));
      kp.func = reinterpret_cast<void *>(&k_gp_step<kCopies>);
      kp.gridDim = dim3(grid);
      kp.blockDim = dim3(kGpThreads);
      kp.sharedMemBytes = 0;
      kp.kernelParams = w < kBatch ? args_on : args_off;
      kp.extra = nullptr;
      hipGraphNode_t node;
      ORQG_HIP(hipGraphAddKernelNode(&node, guard.g, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
      prev = node;
    }
  }
  ORQG_HIP(hipGraphInstantiate(&guard.e, guard.g, nullptr, nullptr, 0));
  hipGraphExec_t gexec = guard.e;
  // Swap resolution of a finished local permutation runs on a second stream
  // while the windows of the next ones proceed: it is memory-latency bound,
  // the windows are AES (VALU/LDS) bound.
  int32_t *head = (int32_t *)c->scratch_get("fy_head", (size_t)n * 4 + 64);
  int32_t *next = (int32_t *)c->scratch_get("fy_next", (size_t)n * 4 + 64);
  int32_t *parent = (int32_t *)c->scratch_get("fy_parent", (size_t)n * 4 + 64);
  uint8_t *dirty = (uint8_t *)c->scratch_get("fy_dirty", (size_t)n + 64);
  uint32_t *err = (uint32_t *)c->scratch_get("fy_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  ORQG_HIP(hipMemsetAsync(dirty, 0, n, c->stream));  // then cleared by k_fy_sort
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
      hipLaunchKernelGGL(k_fy_link, dim3(fgrid), dim3(256), 0, c->stream2, n, r, head, next, dirty,
                         err);
      hipLaunchKernelGGL(k_fy_sort, dim3(fgrid), dim3(256), 0, c->stream2, n, head, next, parent,
                         dirty, err);
      hipLaunchKernelGGL(k_fy_walk, dim3(fgrid), dim3(256), 0, c->stream2, n, r, next, parent,
                         fresh[k].fwd[g], fresh[k].inv[g]);
    }
    if (upto > launched[g]) launched[g] = upto;
  };
  // a permutation is ready once all three of its swap resolutions are queued
  uint32_t marked = 0;
  auto mark_ready = [&]() {
    const uint32_t all = std::min(launched[0], std::min(launched[1], launched[2]));
    for (; marked < all; marked++) {
      ORQG_HIP(hipEventCreateWithFlags(&fresh[marked].ready, hipEventDisableTiming));
      ORQG_HIP(hipEventRecord(fresh[marked].ready, c->stream2));
    }
  };
  // every window either resolves or is retried at a quarter size down to
  // kWinMin draws, which always resolves; the cap only guards logic errors
  const uint64_t max_windows = 64ull * count * (1 + (uint64_t)n / 1024) + (1u << 20);
  for (uint64_t it = 0;; it++) {
    if (it * kBatch > 2 * max_windows) throw Error(ORQG_E_PARITY, "gen_perm: window limit exceeded");
    {
      // family k_gp_step: the batch's kBatch step launches and its
      // commit-only launch, back to back in one graph
      TimedRegion ts(c, "k_gp_step", kBatch + 1);
      ORQG_HIP(hipGraphLaunch(gexec, c->stream));
    }
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
    mark_ready();
    if (host_flag[0]) break;
  }
  for (int g = 0; g < 3; g++) {
    uint64_t nonce;
    std::memcpy(&nonce, host_flag + 6 + 2 * g, 8);
    c->nonce[kGroup0 + g] = nonce;
  }
  // the windows are done: the rest of the swap resolution gets a full grid
  // and runs on while the caller goes on (a consumer of a permutation waits
  // for its event; the check flags are read by check_swaps)
  fgrid = grid_for(n, 256, c->num_cu * 8);
  for (int g = 0; g < 3; g++) resolve_swaps(g, count);
  mark_ready();
  ORQG_HIP(hipGetLastError());
  c->fy_err = err;
  c->fy_pending = true;
}

// Sizes up to this take the small path (ORQG_GP_SMALL_MAX overrides it; 0
// turns the small path off).  At 2^16 rows a permutation draws about 1.5 * 2^20
// words, so its nonce count (8) is still predicted right nearly always.
uint32_t small_max() {
  if (const char *e = std::getenv("ORQG_GP_SMALL_MAX")) return (uint32_t)std::max(0L, std::atol(e));
  // the window pipeline's test hooks act on that path only
  if (std::getenv("ORQG_GP_L2CAP") || std::getenv("ORQG_GP_BOUND_SHIFT")) return 0;
  return 1u << 16;
}

// Nonces a permutation of n rows most likely consumes: 4 per refill of 2^20
// draws, with the expected draw count maskN * H(n) (position i takes
// maskN / (n - i) draws on average).
uint64_t likely_nonces(uint32_t n, uint32_t maskN) {
  double h = 0;
  for (uint32_t k = n; k >= 1; k--) h += 1.0 / k;
  const double draws = (double)maskN * h;
  return 4ull * (1 + (uint64_t)(draws / (double)kRefillWords));
}

// Every local permutation of the reserves in `rs` (contexts of the same
// device, n <= small_max()), drawn in one launch on lead's stream and resolved
// in one launch sequence.  The contexts' streams are ordered before and after.
void gen_small(orqg_ctx *lead, const std::vector<Reserve *> &rs) {
  TimedRegion tr(lead, "gen_perm");
  struct Chain {
    Reserve *R;
    int g;
    size_t first;  // task index of its permutation 0
    uint64_t q;    // predicted nonces per permutation
  };
  std::vector<Chain> chains;
  std::vector<SmallTask> task;
  std::vector<uint32_t *> rbuf(rs.size());
  size_t ntask = 0;
  for (Reserve *R : rs) ntask += 3ull * R->count;
  uint64_t *dD = (uint64_t *)lead->scratch_get("gps_D", ntask * 8 + 64);
  uint32_t *derr = (uint32_t *)lead->scratch_get("gps_err", 16);
  for (size_t q = 0; q < rs.size(); q++) {
    Reserve &R = *rs[q];
    uint32_t bits = 0;
    for (uint32_t t = R.n; t; t >>= 1) bits++;
    const uint32_t maskN = 1u << bits;
    uint64_t per = likely_nonces(R.n, maskN);
    // test hook: a fixed (wrong) prediction forces the redraws below
    if (const char *e = std::getenv("ORQG_GP_SMALL_PREDICT")) per = (uint64_t)std::max(0L, std::atol(e));
    rbuf[q] = (uint32_t *)R.c->scratch_get("gp_r", 3ull * R.count * R.n * 4 + 64);
    for (int g = 0; g < 3; g++) {
      chains.push_back({&R, g, task.size(), per});
      for (uint32_t k = 0; k < R.count; k++) {
        SmallTask t;
        t.rk = R.c->rk(kGroup0 + g);
        t.nonce = R.c->nonce[kGroup0 + g] + per * k;
        t.r = rbuf[q] + ((size_t)g * R.count + k) * R.n;
        t.D = dD + task.size();
        // far beyond any real draw count (expected maskN * H(n))
        t.max_draws = 1024ull * maskN * (bits + 1);
        t.n = R.n;
        t.mask = maskN - 1;
        task.push_back(t);
      }
    }
  }
  // every context's earlier work (its queue, its scratch) precedes the draws
  for (Reserve *R : rs)
    if (R->c != lead) {
      ORQG_HIP(hipEventRecord(R->c->ev_aux, R->c->stream));
      ORQG_HIP(hipStreamWaitEvent(lead->stream, R->c->ev_aux, 0));
    }
  SmallTask *dtask = (SmallTask *)lead->scratch_get("gps_tasks", ntask * sizeof(SmallTask) + 64);
  ORQG_HIP(hipMemsetAsync(derr, 0, 4, lead->stream));
  std::vector<uint64_t> D(ntask, 0);
  std::vector<SmallTask> todo = task;
  for (int round = 0; !todo.empty(); round++) {
    if (round > (int)ntask + 1) throw Error(ORQG_E_PARITY, "gen_perm: start nonces did not settle");
    ORQG_HIP(hipMemcpyAsync(dtask, todo.data(), todo.size() * sizeof(SmallTask),
                            hipMemcpyHostToDevice, lead->stream));
    {
      TimedRegion tk(lead, "k_gp_small");
      hipLaunchKernelGGL(k_gp_small<kCopies>, dim3((unsigned)todo.size()), dim3(kGpThreads), 0,
                         lead->stream, lead->d_t0, dtask, derr);
    }
    ORQG_HIP(hipGetLastError());
    uint32_t herr = 0;
    ORQG_HIP(hipMemcpyAsync(D.data(), dD, ntask * 8, hipMemcpyDeviceToHost, lead->stream));
    ORQG_HIP(hipMemcpyAsync(&herr, derr, 4, hipMemcpyDeviceToHost, lead->stream));
    ORQG_HIP(hipStreamSynchronize(lead->stream));
    if (herr) throw Error(ORQG_E_PARITY, "gen_perm: draw limit exceeded");
    // exact start nonces from the draw counts; a chain whose permutation k
    // started elsewhere is drawn again from k on
    todo.clear();
    for (const Chain &ch : chains) {
      uint64_t e = ch.R->c->nonce[kGroup0 + ch.g];
      for (uint32_t k = 0; k < ch.R->count; k++) {
        SmallTask &t = task[ch.first + k];
        if (t.nonce != e) {
          for (uint32_t k2 = k; k2 < ch.R->count; k2++) {
            task[ch.first + k2].nonce = e + ch.q * (k2 - k);
            todo.push_back(task[ch.first + k2]);
          }
          break;
        }
        e += 4ull * (1 + D[ch.first + k] / kRefillWords);
      }
    }
  }
  for (const Chain &ch : chains) {
    const SmallTask &last = task[ch.first + ch.R->count - 1];
    ch.R->c->nonce[kGroup0 + ch.g] = last.nonce + 4ull * (1 + D[ch.first + ch.R->count - 1] / kRefillWords);
  }
  // swap resolution of all of them
  size_t tot = 0;
  uint32_t nmax = 0;
  for (Reserve *R : rs) {
    tot += 3ull * R->count * R->n;
    nmax = std::max(nmax, R->n);
  }
  int32_t *head = (int32_t *)lead->scratch_get("gps_head", tot * 4 + 64);
  int32_t *next = (int32_t *)lead->scratch_get("gps_next", tot * 4 + 64);
  int32_t *parent = (int32_t *)lead->scratch_get("gps_parent", tot * 4 + 64);
  uint8_t *dirty = (uint8_t *)lead->scratch_get("gps_dirty", tot + 64);
  std::vector<FyTask> fy;
  fy.reserve(ntask);
  size_t off = 0;
  for (const Chain &ch : chains)
    for (uint32_t k = 0; k < ch.R->count; k++) {
      FyTask f;
      f.r = task[ch.first + k].r;
      f.head = head + off;
      f.next = next + off;
      f.parent = parent + off;
      f.dirty = dirty + off;
      f.fwd = ch.R->fresh[k].fwd[ch.g];
      f.inv = ch.R->fresh[k].inv[ch.g];
      f.n = ch.R->n;
      fy.push_back(f);
      off += ch.R->n;
    }
  FyTask *dfy = (FyTask *)lead->scratch_get("gps_fy", fy.size() * sizeof(FyTask) + 64);
  ORQG_HIP(hipMemcpyAsync(dfy, fy.data(), fy.size() * sizeof(FyTask), hipMemcpyHostToDevice,
                          lead->stream));
  ORQG_HIP(hipMemsetAsync(head, 0xff, tot * 4, lead->stream));
  ORQG_HIP(hipMemsetAsync(dirty, 0, tot, lead->stream));
  const unsigned gx = grid_for(nmax, 256, 1u << 16);
  for (size_t t0 = 0; t0 < fy.size(); t0 += 65535) {
    const unsigned ny = (unsigned)std::min<size_t>(65535, fy.size() - t0);
    hipLaunchKernelGGL(k_fyb_link, dim3(gx, ny), dim3(256), 0, lead->stream, dfy + t0, derr);
    hipLaunchKernelGGL(k_fyb_sort, dim3(gx, ny), dim3(256), 0, lead->stream, dfy + t0, derr);
    hipLaunchKernelGGL(k_fyb_walk, dim3(gx, ny), dim3(256), 0, lead->stream, dfy + t0);
  }
  ORQG_HIP(hipGetLastError());
  // (the wait also keeps the host task tables alive until they are copied)
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, derr, 4, hipMemcpyDeviceToHost, lead->stream));
  ORQG_HIP(hipStreamSynchronize(lead->stream));
  if (herr) throw Error(ORQG_E_PARITY, "gen_perm: swap resolution check failed");
}

}  // namespace

// PermutationManager::reserve for 3PC (permutation_manager.h:102-181):
// `count` new sharded permutations, each with one local permutation per pair
// group, generated in queue order from each group's PRG
// (HMShardedPermutationGenerator::generateBatch,
// hm_sharded_permutation_generator.h:178-194).  The reference also generates a
// permutation for the {0,1,2} group, which no operator on this path applies and
// whose PRG feeds nothing else here, so it is not generated.
void reserve_perms(orqg_ctx *c, size_t n, size_t count) {
  Reserve R;
  if (!reserve_begin(c, n, count, R)) return;
  if (c->threads <= 1) {
    if (R.n <= small_max())
      gen_small(c, {&R});
    else
      gen_windows(c, R);
    reserve_end(R);
    return;
  }
  // each worker thread makes its run of the permutations from its own group
  // keys, in queue order
  const auto runs = thread_runs(c, R.count);
  for (int t = 0; t < c->threads; t++) {
    if (!runs[t].second) continue;
    check_swaps(c);  // the previous run's swap resolution (its scratch is reused)
    Reserve Rt;
    Rt.c = c;
    Rt.n = R.n;
    Rt.count = runs[t].second;
    Rt.fresh.assign(R.fresh.begin() + runs[t].first,
                    R.fresh.begin() + runs[t].first + runs[t].second);
    {
      ThreadScope ts(c, t);
      if (Rt.n <= small_max())
        gen_small(c, {&Rt});
      else
        gen_windows(c, Rt);
    }
    std::copy(Rt.fresh.begin(), Rt.fresh.end(), R.fresh.begin() + runs[t].first);
  }
  reserve_end(R);
}

// reserve_perms of several contexts (distinct, one device): the small ones
// share one launch sequence on the first one's stream.
void reserve_perms_multi(const std::vector<orqg_ctx *> &cs, const std::vector<size_t> &ns,
                         const std::vector<size_t> &counts) {
  for (orqg_ctx *c : cs) require_one_thread(c, "batched reserve");
  std::vector<Reserve> R(cs.size());
  std::vector<Reserve *> small;
  for (size_t q = 0; q < cs.size(); q++) {
    if (!reserve_begin(cs[q], ns[q], counts[q], R[q])) continue;
    if (R[q].n <= small_max()) {
      small.push_back(&R[q]);
    } else {
      gen_windows(cs[q], R[q]);
      reserve_end(R[q]);
    }
  }
  if (small.empty()) return;
  gen_small(small[0]->c, small);
  for (Reserve *r : small) {
    // the lead's stream finished the generation (gen_small synchronised it)
    reserve_end(*r);
  }
}

void perm_wait(orqg_ctx *c, ShardedPerm &p) {
  if (p.ready) ORQG_HIP(hipStreamWaitEvent(c->stream, p.ready, 0));
}

void check_swaps(orqg_ctx *c) {
  if (!c->fy_pending) return;
  c->fy_pending = false;
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, c->fy_err, 4, hipMemcpyDeviceToHost, c->stream2));
  ORQG_HIP(hipStreamSynchronize(c->stream2));
  if (herr) throw Error(ORQG_E_PARITY, "gen_perm: swap resolution check failed");
}

ShardedPerm next_perm(orqg_ctx *c, size_t n) {
  if (c->queue.empty() || c->stored_size != n) reserve_perms(c, n, 1);
  ShardedPerm p = c->queue.front();
  c->queue.pop_front();
  perm_wait(c, p);
  return p;
}

}  // namespace orqg
