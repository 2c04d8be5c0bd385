// Honest-majority sharded permutations on the GPU:
//   * gen_perm (include/core/random/permutations/hm_sharded_permutation_generator.h:70-108)
//     reproduced exactly -- same keystream words, same rejections, same swaps;
//   * local application, resharing rounds and the elementwise application of
//     include/core/operators/shuffle.h:113-133, 250-270, 349-412, 687-706.
//
// gen_perm is a sequential Fisher-Yates: for i = 0..n-1 draw uint32 words w
// from the group PRG (refilled 2^20 words = 4 nonces at a time), take
// v = w & (2^bit_width(n)-1) and accept the first v <= n-1-i; swap(i, i+v).
// Writing k = n - i for the number of positions still open, draw d is
// accepted iff v_d < k at that moment.  We split this into
//   (1) acceptance: which draws are accepted (gives r_i = v of the i-th
//       acceptance and D = draws consumed), and
//   (2) swap resolution: the permutation those swaps produce.
//
// (1) runs window by window over the draw stream of each group PRG; the three
// group PRGs ({0,1},{0,2},{1,2}) are three independent chains processed by the
// same launches.  Inside a window that starts with k open positions:
//   * draws with v >= k are rejected whatever happens (k only decreases);
//   * the others ("candidates") get their rank a among candidates (a
//     decoupled look-back scan across tiles).  If every earlier candidate was
//     accepted, k at candidate a is k - a, so v < k - a means "accepted" for
//     sure (at most a candidates were accepted before it).
//   * the rest ("ambiguous", rare) are resolved exactly by one workgroup with
//     alternating optimistic passes, each of which decides most remaining
//     items and hands the rest to the next pass.
// (2) uses that position i is final after step i: the value Fisher-Yates
// leaves at i is whatever position t_i = i + r_i held just before step i,
// which is found by walking "last earlier step that targeted this position"
// pointers over per-position buckets of steps.
#include <algorithm>

#include <cstring>
#include <cmath>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

namespace {

constexpr int kTileThreads = 1024;
constexpr int kDrawsPerThread = 16;                         // 4 AES blocks
constexpr uint32_t kTile = kTileThreads * kDrawsPerThread;  // 16384 draws
constexpr int kGpCopies = kCopies;                          // AES table replication
constexpr uint32_t kWinMax = 1u << 24;                      // draws per window
constexpr uint32_t kWinMin = 1024;  // <= kAmbCap candidates, so a window always resolves
constexpr uint32_t kMaxTiles = kWinMax / kTile;
constexpr uint32_t kAmbCap = 8192;
constexpr int kResolveThreads = 1024;
constexpr uint64_t kRefillWords = 1ull << 20;  // gen_perm buffer: 2^20 uint32
constexpr uint64_t kWordsPerNonce = 1ull << 18;

struct ChainState {
  uint64_t nonce;   // nonce where the current permutation's draws start
  uint64_t d0;      // draw index (inside the current permutation) of the window start
  uint64_t D_end;   // draws consumed by the permutation that just finished
  uint32_t k;       // index of the permutation being generated
  uint32_t i;       // acceptances so far in it
  uint32_t W;       // window size in draws (multiple of kTile)
  uint32_t ticket;  // tile ticket counter
  uint32_t C;       // candidates in the window
  uint32_t m;       // ambiguous items in the window
  uint32_t R;       // rejected ambiguous items
  uint32_t acc;     // accepted in the window
  uint32_t ended;   // the permutation finished in this window
  uint32_t overflow;
  uint32_t done;
  uint32_t err;
  // what k_gp_commit writes for the window just resolved
  uint32_t c_valid, c_i0, c_k, c_C, c_R;
};

struct ChainBufs {
  ChainState *st;      // [3]
  uint64_t *status;    // [3][kMaxTiles]
  uint64_t *cand;      // [3][kWinMax]   (draw offset << 32) | v, by candidate rank
  uint64_t *amb;       // [3][kAmbCap]   (rank << 32) | v, in rank order
  uint32_t *rej;       // [3][kAmbCap]   ranks of rejected ambiguous candidates
  uint32_t *r_out;     // [3][count][n]  accepted offsets r_i
  uint64_t *D;         // [3][count]     draws consumed per permutation
  uint64_t *nonce_at;  // [3][count+1]   nonce at the start of each permutation
};

// Window size for k open positions: the expected number of ambiguous
// candidates is about W^2 k / (2 maskN^2); 90 maskN / sqrt(k) keeps it near
// half of kAmbCap.  An overflowing window is retried at a quarter size.
__host__ __device__ inline uint32_t choose_window(uint32_t k, uint32_t maskN) {
  double w = 90.0 * (double)maskN / sqrt((double)(k ? k : 1));
  if (w >= (double)kWinMax) return kWinMax;
  uint32_t W = w >= kTile ? (uint32_t)(w / kTile) * kTile : (uint32_t)(w / kWinMin) * kWinMin;
  return W < kWinMin ? kWinMin : W;
}

// status word: [63:62] state (see the look-back in k_gp_window), [61:31]
// candidate count, [30:0] ambiguous count.
__device__ inline uint64_t pack_status(uint32_t flag, uint32_t cnt, uint32_t amb) {
  return ((uint64_t)flag << 62) | ((uint64_t)(cnt & 0x7fffffffu) << 31) | (amb & 0x7fffffffu);
}

// Wave-wide look-back for tile `tile`: walks predecessors 64 at a time and
// returns the exclusive prefix of one field.  A predecessor whose state is
// >= incl_state holds an inclusive prefix of that field (the walk stops
// there); states below it but nonzero hold aggregates; `field` 1 is the
// candidate count, 0 the ambiguous count.  For the ambiguous field states 1
// carry no value yet and are waited for.  Waits are bounded: after 2^22
// polls an error flag is raised and the walk gives up.
__device__ inline uint32_t lookback(const uint64_t *status, uint32_t tile, uint32_t incl_state,
                                    int field, int lane, uint32_t *err) {
  uint32_t acc = 0;
  int64_t base = (int64_t)tile - 1;
  uint32_t polls = 0;
  const uint32_t ready_state = field ? 1u : 2u;
  while (base >= 0) {
    int64_t p = base - lane;
    uint32_t flag = incl_state, val = 0;  // before tile 0: inclusive zero
    if (p >= 0) {
      uint64_t s = __hip_atomic_load(&status[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      flag = (uint32_t)(s >> 62);
      val = field ? (uint32_t)((s >> 31) & 0x7fffffffu) : (uint32_t)(s & 0x7fffffffu);
    }
    uint64_t incl = __ballot(flag >= incl_state);
    uint64_t notready = __ballot(flag < ready_state);
    int first = incl ? __builtin_ctzll(incl) : 64;
    uint64_t before = first >= 63 ? ~0ull : ((1ull << (first + 1)) - 1);
    if (notready & before) {
      if (++polls > (1u << 22)) {
        if (lane == 0) atomicOr(err, field ? 1u : 2u);
        return acc;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t x = lane <= first ? val : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    acc += x;
    if (first < 64) break;
    base -= 64;
  }
  return acc;
}

// Draws of one thread of tile `tile`: d = d0 + tile*kTile + threadIdx*kDrawsPerThread + j.
// Draws at window offset >= W are returned as 0xffffffff, which is never a
// candidate (masked values are < 2^31).
template <int COPIES>
__device__ inline void gp_draws(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                uint64_t nonce, uint64_t d0, uint32_t tile, uint32_t mask,
                                uint32_t W, uint32_t (&v)[kDrawsPerThread]) {
  const uint32_t off0 = tile * kTile + threadIdx.x * kDrawsPerThread;
  if (off0 >= W) {
#pragma unroll
    for (int j = 0; j < kDrawsPerThread; j++) v[j] = 0xffffffffu;
    return;
  }
  const uint64_t dbase = d0 + off0;
#pragma unroll
  for (int b = 0; b < kDrawsPerThread / 4; b++) {
    uint64_t d = dbase + 4 * b;
    uint4 blk = aes_ctr_block<COPIES>(A, rk, nonce + d / kWordsPerNonce,
                                      2 + (uint32_t)((d % kWordsPerNonce) / 4));
    v[4 * b + 0] = blk.x & mask;
    v[4 * b + 1] = blk.y & mask;
    v[4 * b + 2] = blk.z & mask;
    v[4 * b + 3] = blk.w & mask;
  }
}

// One window of one chain (blockIdx.y): keystream -> candidates -> ranks ->
// sure accepts / ambiguous list.  Tiles are taken from a ticket counter; the
// keystream of the next tile is computed before the look-back of the current
// one so that the look-back latency overlaps AES work.
template <int COPIES>
__global__ void __launch_bounds__(kTileThreads) k_gp_window(const uint32_t *__restrict__ t0,
                                                            const uint32_t *__restrict__ rk_groups,
                                                            ChainBufs B, uint32_t n,
                                                            uint32_t mask) {
  __shared__ __align__(16) uint32_t lds[AesLds<COPIES>::kWords];
  __shared__ uint32_t s_ticket;
  __shared__ uint32_t s_wc[kTileThreads / 64], s_wa[kTileThreads / 64];
  __shared__ uint32_t s_prefix_c, s_prefix_a;
  const int chain = blockIdx.y;
  ChainState *st = B.st + chain;
  if (st->done) return;
  aes_load_tables<COPIES>(lds, t0);
  AesCtx<COPIES> A(lds);
  const uint32_t *rk = rk_groups + 60 * chain;
  const uint64_t nonce = st->nonce, d0 = st->d0;
  const uint32_t kstart = n - st->i, W = st->W;
  const uint32_t ntiles = (W + kTile - 1) / kTile;
  uint64_t *status = B.status + (size_t)chain * kMaxTiles;
  uint64_t *cand = B.cand + (size_t)chain * kWinMax;
  uint64_t *amb = B.amb + (size_t)chain * kAmbCap;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

  if (threadIdx.x == 0) s_ticket = atomicAdd(&st->ticket, 1u);
  __syncthreads();
  uint32_t tile = s_ticket;
  __syncthreads();
  uint32_t v[kDrawsPerThread];
  if (tile < ntiles) gp_draws<COPIES>(A, rk, nonce, d0, tile, mask, W, v);

  while (tile < ntiles) {
    uint32_t nc = 0;
#pragma unroll
    for (int j = 0; j < kDrawsPerThread; j++) nc += v[j] < kstart;
    uint32_t incl = nc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wc[wv] = incl;
    __syncthreads();
    uint32_t wbase = 0, tile_c = 0;
    for (int w = 0; w < kTileThreads / 64; w++) {
      if (w < wv) wbase += s_wc[w];
      tile_c += s_wc[w];
    }
    const uint32_t excl = wbase + incl - nc;
    // Decoupled look-back with three states per tile:
    //   1: candidate aggregate        2: candidate prefix + ambiguous aggregate
    //   3: candidate prefix + ambiguous prefix (both inclusive)
    // A tile only waits for lower tiles, which hold earlier tickets and are
    // therefore being worked on by running blocks.
    if (threadIdx.x == 0) {
      if (tile > 0)
        __hip_atomic_store(&status[tile], pack_status(1, tile_c, 0), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      s_ticket = atomicAdd(&st->ticket, 1u);
    }
    __syncthreads();
    const uint32_t next_tile = s_ticket;
    // wave 0 runs the look-back while the other waves already compute the
    // next tile's keystream; wave 0 computes its own share afterwards
    if (wv == 0) {
      uint32_t pc = lookback(status, tile, 2, 1, lane, &st->err);
      if (lane == 0) s_prefix_c = pc;
    }
    uint32_t vn[kDrawsPerThread];
    if (next_tile < ntiles) gp_draws<COPIES>(A, rk, nonce, d0, next_tile, mask, W, vn);
    __syncthreads();
    const uint32_t pre_c = s_prefix_c;
    // ambiguous flags need the global rank a = pre_c + excl + (index among own)
    uint32_t na = 0;
    {
      uint32_t a = pre_c + excl;
#pragma unroll
      for (int j = 0; j < kDrawsPerThread; j++) {
        if (v[j] < kstart) {
          if (!(a < kstart - v[j])) na++;
          a++;
        }
      }
    }
    uint32_t inca = na;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t t = __shfl_up(inca, o, 64);
      if (lane >= o) inca += t;
    }
    if (lane == 63) s_wa[wv] = inca;
    __syncthreads();
    uint32_t wabase = 0, tile_a = 0;
    for (int w = 0; w < kTileThreads / 64; w++) {
      if (w < wv) wabase += s_wa[w];
      tile_a += s_wa[w];
    }
    const uint32_t excla = wabase + inca - na;
    if (wv == 0) {
      if (tile > 0 && lane == 0)
        __hip_atomic_store(&status[tile], pack_status(2, pre_c + tile_c, tile_a), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      uint32_t pa = lookback(status, tile, 3, 0, lane, &st->err);
      if (lane == 0) {
        s_prefix_a = pa;
        __hip_atomic_store(&status[tile], pack_status(3, pre_c + tile_c, pa + tile_a),
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (tile == ntiles - 1) {
          st->C = pre_c + tile_c;
          st->m = pa + tile_a;
        }
      }
    }
    __syncthreads();
    const uint32_t pre_a = s_prefix_a;
    {
      uint32_t a = pre_c + excl, q = pre_a + excla;
      const uint32_t doff = tile * kTile + threadIdx.x * kDrawsPerThread;
#pragma unroll
      for (int j = 0; j < kDrawsPerThread; j++) {
        if (v[j] < kstart) {
          cand[a] = ((uint64_t)(doff + j) << 32) | v[j];
          if (!(a < kstart - v[j])) {
            if (q < kAmbCap) amb[q] = ((uint64_t)a << 32) | v[j];
            q++;
          }
          a++;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kDrawsPerThread; j++) v[j] = vn[j];
    tile = next_tile;
    __syncthreads();
  }
}

// Exact resolution of the ambiguous candidates of one chain's window.
// Item j (rank a_j, value v_j, T_j = k - v_j <= a_j) is accepted iff the number
// of acceptances before it is < T_j, i.e. iff the number of rejected
// ambiguous items before it exceeds U_j = a_j - T_j.
//   type B list (accept iff rejected-before > U): item at list index t with
//     t <= U is surely rejected, U < 0 surely accepted, otherwise it is
//     accepted iff accepted-before-among-undecided < t - (sure accepts before) - U;
//   type A list (accept iff accepted-before < T): t < T surely accepted,
//     T <= 0 surely rejected, otherwise accepted iff
//     rejected-before-among-undecided > t - (sure rejects before) - T.
// Each pass decides most items; the undecided keep their order.
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t *s_w, uint32_t &total) {
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

__global__ void __launch_bounds__(kResolveThreads) k_gp_resolve(ChainBufs B, uint32_t n,
                                                                uint32_t count, uint32_t maskN) {
  constexpr int PER = kAmbCap / kResolveThreads;  // items per thread (8)
  __shared__ uint16_t s_idx[kAmbCap];
  __shared__ int32_t s_thr[kAmbCap];
  __shared__ uint8_t s_dec[kAmbCap];  // 0 undecided, 1 accepted, 2 rejected
  __shared__ uint32_t s_w[kResolveThreads / 64];
  __shared__ uint32_t s_cnt;
  const int chain = blockIdx.x;
  ChainState *st = B.st + chain;
  if (st->done) {
    if (threadIdx.x == 0) st->c_valid = 0;
    return;
  }
  const uint32_t m = st->m;
  const uint32_t *dummy = nullptr;
  (void)dummy;
  uint64_t *amb = B.amb + (size_t)chain * kAmbCap;
  uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
  uint64_t *cand = B.cand + (size_t)chain * kWinMax;
  const uint32_t kstart = n - st->i;
  uint64_t *status = B.status + (size_t)chain * kMaxTiles;
  const uint32_t used_tiles = (st->W + kTile - 1) / kTile;
  if (m > kAmbCap) {
    // too many ambiguous candidates: retry the same window at a quarter size
    for (uint32_t t = threadIdx.x; t < used_tiles; t += blockDim.x) status[t] = 0;
    if (threadIdx.x == 0) {
      st->c_valid = 0;
      uint32_t w = st->W / 4;
      st->W = w >= kTile ? w / kTile * kTile : (w >= kWinMin ? w / kWinMin * kWinMin : kWinMin);
      st->ticket = 0;
      st->C = 0;
      st->m = 0;
    }
    return;
  }
  // initial list: all ambiguous items, type B thresholds U = a - (k - v)
  for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) {
    uint64_t e = amb[j];
    uint32_t a = (uint32_t)(e >> 32), v = (uint32_t)e;
    s_idx[j] = (uint16_t)j;
    s_thr[j] = (int32_t)(a - (kstart - v));
    s_dec[j] = 0;
  }
  __syncthreads();
  uint32_t len = m;
  int type_b = 1;
  while (len > 0) {
    // each thread owns a contiguous run of PER list slots
    uint32_t beg = threadIdx.x * PER;
    uint8_t dec[PER];
    uint32_t n_sure_acc = 0, n_sure_rej = 0, n_und = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      uint32_t t = beg + q;
      dec[q] = 3;  // past the end
      if (t < len) {
        int32_t thr = s_thr[t];
        if (type_b) {
          if ((int32_t)t <= thr) dec[q] = 2;
          else if (thr < 0) dec[q] = 1;
          else dec[q] = 0;
        } else {
          if ((int32_t)t < thr) dec[q] = 1;
          else if (thr <= 0) dec[q] = 2;
          else dec[q] = 0;
        }
        n_sure_acc += dec[q] == 1;
        n_sure_rej += dec[q] == 2;
        n_und += dec[q] == 0;
      }
    }
    uint32_t tot;
    uint32_t sa = block_excl_scan(n_sure_acc, s_w, tot);
    uint32_t sr = block_excl_scan(n_sure_rej, s_w, tot);
    uint32_t un = block_excl_scan(n_und, s_w, tot);
    const uint32_t new_len = tot;
    // read phase done for old entries before compaction overwrites
    int32_t nthr[PER];
    uint16_t nidx[PER];
    uint32_t npos[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      uint32_t t = beg + q;
      npos[q] = 0xffffffffu;
      if (dec[q] == 3) continue;
      uint16_t id = s_idx[t];
      if (dec[q] == 1 || dec[q] == 2) {
        s_dec[id] = dec[q];
      } else {
        int32_t thr = s_thr[t];
        int32_t nt = type_b ? (int32_t)t - (int32_t)sa - thr : (int32_t)t - (int32_t)sr - thr;
        nthr[q] = nt;
        nidx[q] = id;
        npos[q] = un;
        un++;
      }
      sa += dec[q] == 1;
      sr += dec[q] == 2;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
      if (npos[q] != 0xffffffffu) {
        s_idx[npos[q]] = nidx[q];
        s_thr[npos[q]] = nthr[q];
      }
    }
    __syncthreads();
    len = new_len;
    type_b = !type_b;
  }
  // rejected ranks, in rank order
  uint32_t beg = threadIdx.x * PER;
  uint32_t nr = 0;
#pragma unroll
  for (int q = 0; q < PER; q++) nr += (beg + q < m) && s_dec[beg + q] == 2;
  uint32_t tot;
  uint32_t pos = block_excl_scan(nr, s_w, tot);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    uint32_t j = beg + q;
    if (j < m && s_dec[j] == 2) rej[pos++] = (uint32_t)(amb[j] >> 32);
  }
  for (uint32_t t = threadIdx.x; t < used_tiles; t += blockDim.x) status[t] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t C = st->C, R = tot, acc = C - tot;
    if (acc > kstart) atomicOr(&st->err, 4u);
    st->c_valid = 1;
    st->c_i0 = st->i;
    st->c_k = st->k;
    st->c_C = C;
    st->c_R = R;
    if (acc >= kstart) {
      // the permutation ends inside this window: find the candidate rank of
      // the kstart-th acceptance, a* = kstart-1 shifted past every rejected
      // rank at or before it
      uint32_t a = kstart - 1;
      for (uint32_t q = 0; q < R; q++) {
        if (rej[q] <= a) a++;
        else break;
      }
      const uint64_t D_end = st->d0 + (cand[a] >> 32) + 1;
      const uint32_t k = st->k;
      B.D[(size_t)chain * count + k] = D_end;
      // gen_perm refills 4 nonces per 2^20 draws, eagerly
      // (hm_sharded_permutation_generator.h:93-102)
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
    st->ticket = 0;
    st->C = 0;
    st->m = 0;
  }
}

// Write r_i for accepted candidates; locate the draw that finished the permutation.
__global__ void __launch_bounds__(256) k_gp_commit(ChainBufs B, uint32_t n, uint32_t count) {
  const int chain = blockIdx.y;
  const ChainState *st = B.st + chain;
  if (!st->c_valid) return;
  const uint32_t C = st->c_C, R = st->c_R, i0 = st->c_i0, k = st->c_k;
  const uint32_t kstart = n - i0;
  const uint64_t *cand = B.cand + (size_t)chain * kWinMax;
  const uint32_t *rej = B.rej + (size_t)chain * kAmbCap;
  uint32_t *r = B.r_out + ((size_t)chain * count + k) * (size_t)n;
  for (uint32_t a = blockIdx.x * blockDim.x + threadIdx.x; a < C; a += gridDim.x * blockDim.x) {
    // lower_bound of a in rej[0..R)
    uint32_t lo = 0, hi = R;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (rej[mid] < a) lo = mid + 1;
      else hi = mid;
    }
    if (lo < R && rej[lo] == a) continue;
    uint32_t accr = a - lo;
    if (accr >= kstart) continue;  // cannot happen for an exact resolution
    r[i0 + accr] = (uint32_t)cand[a];
  }
}

__global__ void k_gp_done(ChainBufs B, uint32_t *flag) {
  if (threadIdx.x == 0) {
    uint32_t all = B.st[0].done & B.st[1].done & B.st[2].done;
    flag[0] = all;
    flag[1] = B.st[0].err | B.st[1].err | B.st[2].err;
  }
}

// ---- swap resolution ----------------------------------------------------------
// Step j of Fisher-Yates swaps positions j and t_j = j + r_j.  Steps are
// chained per target position (head[t] -> most recently linked step ->
// next[...]); list order is arbitrary, the walks below take maxima.
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
__device__ inline int32_t fy_last_before(const int32_t *__restrict__ hd, const int32_t *__restrict__ nx,
                                         uint32_t pos, uint32_t lim, uint32_t n) {
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

// ---- device-wide exclusive scan of uint32 (three phases) -----------------------
constexpr int kScanBlock = 1024;
constexpr int kScanPer = 4;

__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(uint64_t n, const uint32_t *__restrict__ in,
                                                            uint32_t *__restrict__ sums) {
  __shared__ uint32_t s_w[kScanBlock / 64];
  uint64_t base = (uint64_t)blockIdx.x * kScanBlock * kScanPer + threadIdx.x * kScanPer;
  uint32_t x = 0;
  for (int q = 0; q < kScanPer; q++)
    if (base + q < n) x += in[base + q];
  uint32_t tot;
  block_excl_scan(x, s_w, tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_sums(uint32_t nb, uint32_t *__restrict__ sums) {
  __shared__ uint32_t s_w[kScanBlock / 64];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += kScanBlock) {
    uint32_t j = b0 + threadIdx.x;
    uint32_t x = j < nb ? sums[j] : 0;
    uint32_t tot;
    uint32_t e = block_excl_scan(x, s_w, tot);
    uint32_t carry = s_carry;
    if (j < nb) sums[j] = carry + e;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + tot;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(uint64_t n, const uint32_t *__restrict__ in,
                                                           const uint32_t *__restrict__ sums,
                                                           uint32_t *__restrict__ out) {
  __shared__ uint32_t s_w[kScanBlock / 64];
  uint64_t base = (uint64_t)blockIdx.x * kScanBlock * kScanPer + threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  uint32_t x = 0;
  for (int q = 0; q < kScanPer; q++) {
    v[q] = base + q < n ? in[base + q] : 0;
    x += v[q];
  }
  uint32_t tot;
  uint32_t e = block_excl_scan(x, s_w, tot) + sums[blockIdx.x];
  for (int q = 0; q < kScanPer; q++) {
    if (base + q < n) out[base + q] = e;
    e += v[q];
  }
}

}  // namespace

// out[i] = sum_{j<i} in[j] for i in [0, n]  (out has n+1 entries)
void exclusive_scan_u32(orqg_ctx *c, const uint32_t *in, uint32_t *out, uint64_t n_plus_1) {
  uint64_t per_block = (uint64_t)kScanBlock * kScanPer;
  uint32_t nb = (uint32_t)((n_plus_1 + per_block - 1) / per_block);
  uint32_t *sums = (uint32_t *)c->scratch_get("scan_sums", (size_t)nb * 4 + 16);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, c->stream, n_plus_1, in, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, nb, sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, c->stream, n_plus_1, in, sums, out);
  ORQG_HIP(hipGetLastError());
}

void free_perm(ShardedPerm &p) {
  for (int g = 0; g < 3; g++) {
    dev_free(p.fwd[g]);
    dev_free(p.inv[g]);
    p.fwd[g] = p.inv[g] = nullptr;
  }
}

// PermutationManager::reserve for 3PC: `count` new sharded permutations, each
// with one local permutation per pair group.  The {0,1,2} group's permutation
// is generated by the reference but never applied (its PRG feeds nothing else
// on this path), so it is not generated here.
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
  B.status = (uint64_t *)c->scratch_get("gp_status", 3ull * kMaxTiles * 8);
  B.cand = (uint64_t *)c->scratch_get("gp_cand", 3ull * kWinMax * 8);
  B.amb = (uint64_t *)c->scratch_get("gp_amb", 3ull * kAmbCap * 8);
  B.rej = (uint32_t *)c->scratch_get("gp_rej", 3ull * kAmbCap * 4);
  B.r_out = (uint32_t *)c->scratch_get("gp_r", 3ull * count * n * 4);
  B.D = (uint64_t *)c->scratch_get("gp_D", 3ull * count * 8);
  B.nonce_at = (uint64_t *)c->scratch_get("gp_nonce_at", 3ull * (count + 1) * 8);
  uint32_t *flag = (uint32_t *)c->scratch_get("gp_flag", 16);

  ChainState init[3];
  memset(init, 0, sizeof(init));
  std::vector<uint64_t> nat(3ull * (count + 1), 0);
  for (int g = 0; g < 3; g++) {
    init[g].nonce = c->nonce[kGroup0 + g];
    init[g].W = choose_window(n, maskN);
    nat[(size_t)g * (count + 1)] = init[g].nonce;
  }
  ORQG_HIP(hipMemcpyAsync(B.st, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemcpyAsync(B.nonce_at, nat.data(), nat.size() * 8, hipMemcpyHostToDevice, c->stream));
  ORQG_HIP(hipMemsetAsync(B.status, 0, 3ull * kMaxTiles * 8, c->stream));

  const unsigned grid_win = std::max(1, c->num_cu / 3);  // one workgroup per CU over 3 chains
  const unsigned grid_commit = std::max(1, c->num_cu);
  uint32_t host_flag[2] = {0, 0};
  const int kBatch = 16;
  // every window either resolves or retries at a quarter size down to kWinMin,
  // which always resolves; the cap below only guards against logic errors
  const uint64_t max_windows = 64ull * count * (1 + (uint64_t)n / 1024) + (1u << 20);
  for (uint64_t it = 0;; it++) {
    if (it * kBatch > max_windows) throw Error(ORQG_E_PARITY, "gen_perm: window limit exceeded");
    for (int w = 0; w < kBatch; w++) {
      hipLaunchKernelGGL(k_gp_window<kGpCopies>, dim3(grid_win, 3), dim3(kTileThreads), 0,
                         c->stream, c->d_t0, c->rk(kGroup0), B, n, mask);
      hipLaunchKernelGGL(k_gp_resolve, dim3(3), dim3(kResolveThreads), 0, c->stream, B, n, count,
                         maskN);
      hipLaunchKernelGGL(k_gp_commit, dim3(grid_commit, 3), dim3(256), 0, c->stream, B, n, count);
    }
    hipLaunchKernelGGL(k_gp_done, dim3(1), dim3(64), 0, c->stream, B, flag);
    ORQG_HIP(hipGetLastError());
    ORQG_HIP(hipMemcpyAsync(host_flag, flag, 8, hipMemcpyDeviceToHost, c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));
    if (host_flag[1]) throw Error(ORQG_E_PARITY, "gen_perm: device-side check failed (flags " +
                                                     std::to_string(host_flag[1]) + ")");
    if (host_flag[0]) break;
  }
  ChainState fin[3];
  ORQG_HIP(hipMemcpyAsync(fin, B.st, sizeof(fin), hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  for (int g = 0; g < 3; g++) c->nonce[kGroup0 + g] = fin[g].nonce;

  // swap resolution: the three local permutations of one sharded permutation per launch
  int32_t *head = (int32_t *)c->scratch_get("fy_head", 3ull * n * 4 + 64);
  int32_t *next = (int32_t *)c->scratch_get("fy_next", 3ull * n * 4 + 64);
  uint32_t *err = (uint32_t *)c->scratch_get("fy_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  unsigned grid = grid_for(n, 256, c->num_cu * 8);
  const size_t chain_stride = (size_t)count * n;
  for (uint32_t k = 0; k < count; k++) {
    const uint32_t *r = B.r_out + (size_t)k * n;
    ORQG_HIP(hipMemsetAsync(head, 0xff, 3ull * n * 4, c->stream));
    hipLaunchKernelGGL(k_fy_link, dim3(grid, 3), dim3(256), 0, c->stream, n, r, chain_stride, head,
                       next, err);
    hipLaunchKernelGGL(k_fy_walk, dim3(grid, 3), dim3(256), 0, c->stream, n, r, chain_stride, head,
                       next, fresh[k].fwd[0], fresh[k].fwd[1], fresh[k].fwd[2], fresh[k].inv[0],
                       fresh[k].inv[1], fresh[k].inv[2]);
  }
  ORQG_HIP(hipGetLastError());
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  if (herr) throw Error(ORQG_E_PARITY, "gen_perm: swap resolution check failed");
  for (auto &p : fresh) c->queue.push_back(p);
}

ShardedPerm next_perm(orqg_ctx *c, size_t n) {
  if (c->queue.empty() || c->stored_size != n) reserve_perms(c, n, 1);
  ShardedPerm p = c->queue.front();
  c->queue.pop_front();
  return p;
}

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

template <typename T>
void launch_round(orqg_ctx *c, bool arith, int g, const int32_t *idx, const T *in, T *out,
                  uint64_t n) {
  int role = kGroup0 + g;
  CallSched s = make_sched(n, c->batch_size, sizeof(T), 2);
  unsigned grid = grid_for((n * sizeof(T) + 15) / 16, kAesThreads, c->num_cu);
  hipLaunchKernelGGL((arith ? k_round<T, true> : k_round<T, false>), dim3(grid), dim3(kAesThreads), 0,
                     c->stream, c->d_t0, c->rk(role), c->nonce[role], s, idx, in, out);
  ORQG_HIP(hipGetLastError());
  c->nonce[role] += sched_nonces(s);
  log_reshare(c, g, n * sizeof(T));
}
template void launch_round<uint32_t>(orqg_ctx *, bool, int, const int32_t *, const uint32_t *,
                                     uint32_t *, uint64_t);
template void launch_round<uint64_t>(orqg_ctx *, bool, int, const int32_t *, const uint64_t *,
                                     uint64_t *, uint64_t);

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
    else
      launch_round<uint64_t>(c, arith, g, idx, (const uint64_t *)cur, (uint64_t *)nxt, x->n);
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
