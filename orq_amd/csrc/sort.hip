// Oblivious radix sort (include/core/operators/radixsort.h:116-229) for all
// three parties on one GPU.
//
// The padded vector of pad_input (include/core/operators/sorting.h:431-462)
// is kept as L 32-bit words per share, in one 64-byte record per element
// (words k*L .. k*L+L-1 = share k; the bit permutation's three shares ride in
// words 12..14 during a bit iteration):
//   int32 keys -> PadWidth int64  -> L = 2 (index word, key word)
//   int64 keys -> PadWidth int128 -> L = 3 (index word, 2 key words); bits
//   96..127 of the reference's int128 shares never reach an output (the sort
//   reads bits 32..95, remove_padding keeps bits 0..95), so they are not
//   stored, but their keystream is still generated: resharing masks are drawn
//   as whole 16-byte elements exactly as the reference draws them.
//
// One bit iteration (radixsort.h:141-192):
//   k_bit_b2a   bit 32+i of each share -> b2a_bit (3 zero sharings per batch)
//               -> f1, plus per-block sums of f1 for the prefix scan
//   k_scan3     exclusive scan of the block sums (3 shares)
//   k_bit_perm  s = prefix_sum([1-f1 ; f1]) - 1 in closed form, then
//               perm = s0 + f1*(s1-s0) with multiply_a (1 zero sharing),
//               written into the records' perm words
//   k_round x3  sharded permutation rounds: gather whole records through
//               pi_g^{-1} (one aligned 64-byte piece of a line per element)
//               and add the group-g resharing masks to v and perm; the third
//               round opens pi(perm) and scatters pi(v) by it (the record
//               buffers alternate between bits, so it never writes a buffer
//               it gathers from)
#include <algorithm>
#include <numeric>

#include <cstring>
#include <cmath>

#include "internal.h"
#include "kern_common.h"

namespace orqg {

template <typename T>
void launch_round(orqg_ctx *c, bool arith, int g, const int32_t *idx, const T *in, T *out,
                  uint64_t n);

namespace {

constexpr int kBT = kAesThreads;    // threads per block
constexpr int kEPB = kBT * 4;       // elements per scan block
constexpr int kRec = 16;            // words per element record: v (3L words), perm at kRecPerm
constexpr int kRecPerm = 12;

__device__ inline uint32_t bexcl(uint32_t x, uint32_t *s_w, uint32_t &total) {
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

// exclusive block scan of three values per thread at once (two barriers); s_w
// holds 3 words per wave
__device__ inline void bexcl3(const uint32_t (&x)[3], uint32_t *s_w, uint32_t (&ex)[3],
                              uint32_t (&total)[3]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t inc[3] = {x[0], x[1], x[2]};
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint32_t t = __shfl_up(inc[k], o, 64);
      if (lane >= o) inc[k] += t;
    }
  }
  __syncthreads();
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 3; k++) s_w[3 * wv + k] = inc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; k++) {
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < nw; w++) {
      const uint32_t v = s_w[3 * w + k];
      if (w < wv) base += v;
      tot += v;
    }
    ex[k] = base + inc[k] - x[k];
    total[k] = tot;
  }
}

// ---- pad_input ----------------------------------------------------------------
// share k of element i: (sign-extended key << 32) | idx_k, idx_0 = i (ASC) or
// (uint32)(-1-i) (DESC), idx_1 = idx_2 = 0 (public_share, replicated_3pc.h:456-470);
// DESC then reverses the element order.
template <int L, typename K>
__global__ void k_pad(uint64_t n, bool desc, const K *__restrict__ v, uint32_t *__restrict__ P) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t dst = desc ? n - 1 - i : i;
    uint32_t idx = desc ? (uint32_t)(-1 - (int64_t)i) : (uint32_t)i;
    uint32_t w[kRec];
#pragma unroll
    for (int q = 0; q < kRec; q++) w[q] = 0u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      K key = v[3 * i + k];
      w[k * L] = k == 0 ? idx : 0u;
      w[k * L + 1] = (uint32_t)key;
      if (L == 3) w[k * L + 2] = (uint32_t)((uint64_t)key >> 32);
    }
    uint4 *o = reinterpret_cast<uint4 *>(P + dst * kRec);
#pragma unroll
    for (int q = 0; q < kRec / 4; q++)
      o[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
  }
}

// ---- one bit: b2a_bit + block sums -------------------------------------------------
// A superblock's 12 words per thread (4 rows x 3 shares of the key word that
// holds the bit), issued together: a row past the end reads row n-1 (no
// branch around a load, which would serialize them) and is zeroed when used.
template <int L>
__device__ inline void bit_load_raw(const uint32_t *__restrict__ P, uint64_t n, int word,
                                    uint64_t e0, uint32_t (&raw)[4][3]) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint64_t e = e0 + i < n ? e0 + i : n - 1;
#pragma unroll
    for (int k = 0; k < 3; k++) raw[i][k] = P[e * kRec + k * L + word];
  }
}

__device__ inline void bit_extract(const uint32_t (&raw)[4][3], uint64_t e0, uint64_t n, int bit,
                                   uint32_t flip, uint32_t (&bits)[4][3]) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) bits[i][k] = (e0 + i < n) ? (((raw[i][k] >> bit) & 1u) ^ flip) : 0u;
}

// b2a_bit of one superblock's bits: f1 to F, the superblock's share sums to
// bsum3[0..2]
__device__ inline void bit_b2a_block(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                     const NonceTriple &nb, const CallSched &s, uint64_t e0,
                                     const uint32_t (&bits)[4][3], uint32_t *__restrict__ F,
                                     uint32_t *__restrict__ bsum3, uint32_t *s_w) {
  uint32_t f[4][3];
  b2a_bit_group<uint32_t>(A, rk_pair, nb, s, e0, bits, f);
  uint32_t loc[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (e0 + i >= s.n) f[i][k] = 0u;
      loc[k] += f[i][k];
    }
  // the thread's 12 words are contiguous and 16-byte aligned: three vector stores
  if (e0 + 4 <= s.n) {
    uint4 *o = reinterpret_cast<uint4 *>(F + e0 * 3);
    o[0] = make_uint4(f[0][0], f[0][1], f[0][2], f[1][0]);
    o[1] = make_uint4(f[1][1], f[1][2], f[2][0], f[2][1]);
    o[2] = make_uint4(f[2][2], f[3][0], f[3][1], f[3][2]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (e0 + i < s.n)
#pragma unroll
        for (int k = 0; k < 3; k++) F[(e0 + i) * 3 + k] = f[i][k];
  }
  uint32_t ex[3], tot[3];
  bexcl3(loc, s_w, ex, tot);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; k++) bsum3[k] = tot[k];
}

template <int L>
__global__ void __launch_bounds__(kBT) k_bit_b2a(const uint32_t *__restrict__ t0,
                                                 const uint32_t *__restrict__ rk_pair,
                                                 NonceTriple nb, CallSched s,
                                                 const uint32_t *__restrict__ P, int word,
                                                 int bit, uint32_t flip, uint32_t *__restrict__ F,
                                                 uint32_t *__restrict__ bsum) {
  AES_LDS;
  __shared__ uint32_t s_w[3 * (kBT / 64)];
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint32_t nsb = (uint32_t)((s.n + kEPB - 1) / kEPB);
  uint32_t raw[4][3];
  if (blockIdx.x < nsb) bit_load_raw<L>(P, s.n, word, (uint64_t)blockIdx.x * kEPB + threadIdx.x * 4, raw);
  for (uint32_t sb = blockIdx.x; sb < nsb; sb += gridDim.x) {
    const uint64_t e0 = (uint64_t)sb * kEPB + threadIdx.x * 4;
    uint32_t bits[4][3];
    bit_extract(raw, e0, s.n, bit, flip, bits);
    // the next superblock's words are in flight during this one's keystream
    // (the last superblock reloads itself rather than branch)
    const uint32_t nx = sb + gridDim.x < nsb ? sb + gridDim.x : sb;
    bit_load_raw<L>(P, s.n, word, (uint64_t)nx * kEPB + threadIdx.x * 4, raw);
    bit_b2a_block(A, rk_pair, nb, s, e0, bits, F, bsum + sb * 3, s_w);
  }
}

// exclusive scan of [nb][3] block sums in place; totals to tot[3]
__global__ void __launch_bounds__(1024) k_scan3(uint32_t nb, uint32_t *__restrict__ bsum,
                                                uint32_t *__restrict__ tot) {
  __shared__ uint32_t s_w[1024 / 64];
  __shared__ uint32_t s_carry[3];
  if (threadIdx.x < 3) s_carry[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
    uint32_t j = b0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint32_t x = j < nb ? bsum[j * 3 + k] : 0;
      uint32_t t;
      uint32_t e = bexcl(x, s_w, t);
      uint32_t carry = s_carry[k];
      if (j < nb) bsum[j * 3 + k] = carry + e;
      __syncthreads();
      if (threadIdx.x == 0) s_carry[k] = carry + t;
      __syncthreads();
    }
  }
  if (threadIdx.x < 3) tot[threadIdx.x] = s_carry[threadIdx.x];
}

// s0_k(e) = one_k*e - F_k(e), s1_k(e) = one_k*(n-1) - Ftot_k + F_k(e) with F the
// inclusive prefix of f1 and one = (1,0,0); perm = s0 + f1*(s1 - s0).  One
// superblock: boff3 = its exclusive offsets, tot = the totals (3 words each).
__device__ inline void bit_perm_block(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                      const NonceTriple &nb, const CallSched &s, uint64_t e0,
                                      const uint32_t *__restrict__ F,
                                      const uint32_t *__restrict__ boff3,
                                      const uint32_t *__restrict__ tot, uint32_t *__restrict__ R,
                                      uint32_t *s_w) {
  uint32_t f[4][3];
  uint32_t loc[3] = {0, 0, 0};
  if (e0 + 4 <= s.n) {
    const uint4 *src = reinterpret_cast<const uint4 *>(F + e0 * 3);
    const uint4 a = src[0], b = src[1], c = src[2];
    f[0][0] = a.x; f[0][1] = a.y; f[0][2] = a.z; f[1][0] = a.w;
    f[1][1] = b.x; f[1][2] = b.y; f[2][0] = b.z; f[2][1] = b.w;
    f[2][2] = c.x; f[3][0] = c.y; f[3][1] = c.z; f[3][2] = c.w;
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) f[i][k] = (e0 + i < s.n) ? F[(e0 + i) * 3 + k] : 0u;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) loc[k] += f[i][k];
  uint32_t pre[3];
  {
    uint32_t t[3];
    bexcl3(loc, s_w, pre, t);
#pragma unroll
    for (int k = 0; k < 3; k++) pre[k] += boff3[k];
  }
  uint32_t S[3][4];
  {
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(s, e0, 0, off, ctr, bib);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint4 blk = aes_ctr_block<kCopies>(A, rk_pair + 60 * k, nb.v[k] + off, ctr);
      unpack_block<uint32_t>(blk, S[k]);
    }
  }
  const uint32_t n1 = (uint32_t)(s.n - 1);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint32_t s0[3], y[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      pre[k] += f[i][k];  // inclusive
      uint32_t one = k == 0 ? 1u : 0u;
      s0[k] = one * (uint32_t)(e0 + i) - pre[k];
      uint32_t s1 = one * n1 - tot[k] + pre[k];
      y[k] = s1 - s0[k];
    }
    if (e0 + i < s.n) {
      uint32_t pr[3];
#pragma unroll
      for (int p = 0; p < 3; p++) {
        int q = (p + 1) % 3, pm = (p + 2) % 3;
        uint32_t z = (S[pm][i] - S[p][i]) + f[i][p] * y[p] + f[i][p] * y[q] + f[i][q] * y[p];
        pr[p] = s0[p] + z;
      }
      // into the element record's perm words, as one whole 32-byte piece
      // (words 8..15: word 8 is v's last word for L = 3, 9..11 are zero) so
      // that no partial write needs a read-modify-write
      uint32_t *rec = R + (e0 + i) * kRec;
      const uint32_t w8 = rec[8];
      uint4 *o = reinterpret_cast<uint4 *>(rec + 8);
      o[0] = make_uint4(w8, 0u, 0u, 0u);
      o[1] = make_uint4(pr[0], pr[1], pr[2], 0u);
    }
  }
}

template <int L>
__global__ void __launch_bounds__(kBT) k_bit_perm(const uint32_t *__restrict__ t0,
                                                  const uint32_t *__restrict__ rk_pair,
                                                  NonceTriple nb, CallSched s,
                                                  const uint32_t *__restrict__ F,
                                                  const uint32_t *__restrict__ boff,
                                                  const uint32_t *__restrict__ tot,
                                                  uint32_t *__restrict__ R) {
  AES_LDS;
  __shared__ uint32_t s_w[3 * (kBT / 64)];
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint32_t nsb = (uint32_t)((s.n + kEPB - 1) / kEPB);
  for (uint32_t sb = blockIdx.x; sb < nsb; sb += gridDim.x)
    bit_perm_block(A, rk_pair, nb, s, (uint64_t)sb * kEPB + threadIdx.x * 4, F, boff + sb * 3, tot,
                   R, s_w);
}

// ---- sharded round on element records: gather through idx, then add the
// resharing masks of v (binary, tszv-byte masks) and perm (arithmetic, 4-byte
// masks), drawn from group PRG g ----------------------------------------------------
// A wave takes 128 consecutive elements per step, lane l elements base + l
// and base + 64 + l, so a quad of lanes holds the 4 elements one perm mask
// block covers.  Each lane computes its elements' two v mask blocks and one of
// the quad's four perm blocks (call l & 1 for the block of 64 (l >> 1) & 1),
// broadcast in the quad with DPP.  A step's gathers are issued before its AES
// work and consumed after it.  Records are written after a 4 x 4 transpose of
// 16-byte quarters inside the quad, so one store instruction writes 16 whole
// 64-byte records instead of a quarter of 64.
template <int CTRL>
__device__ inline uint32_t quad_dpp(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ inline uint4 quad_dpp4(const uint4 &x) {
  return make_uint4(quad_dpp<CTRL>(x.x), quad_dpp<CTRL>(x.y), quad_dpp<CTRL>(x.z), quad_dpp<CTRL>(x.w));
}
__device__ inline uint32_t word_of(const uint4 &b, uint32_t s) {
  return s == 0 ? b.x : s == 1 ? b.y : s == 2 ? b.z : b.w;
}
// lane s of a quad holds quarters X[0..3] of record s; afterwards it holds
// quarter s of records 0..3 (two butterfly stages, xor 1 then xor 2)
__device__ inline void quad_transpose(uint4 (&X)[4], uint32_t s) {
#pragma unroll
  for (int h = 0; h < 2; h++) {  // quarters 2h, 2h + 1 with lane s ^ 1
    const uint4 p0 = quad_dpp4<0xB1>(X[2 * h]), p1 = quad_dpp4<0xB1>(X[2 * h + 1]);
    if (s & 1u) X[2 * h] = p1;
    else X[2 * h + 1] = p0;
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {  // quarters h, h + 2 with lane s ^ 2
    const uint4 p0 = quad_dpp4<0x4E>(X[h]), p2 = quad_dpp4<0x4E>(X[h + 2]);
    if (s & 2u) X[h] = p2;
    else X[h + 2] = p0;
  }
}

// One wave step of a sharded round: rows [128 w, 128 w + 128) of the n = sv.n
// rows (every wave of the launch takes its own steps).
template <int L, bool SCATTER>
__device__ inline void round2_step(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk,
                                   uint64_t nv, const CallSched &sv, uint64_t np,
                                   const CallSched &sp, const int32_t *__restrict__ idx,
                                   const uint32_t *__restrict__ Rin, uint32_t *__restrict__ Rout,
                                   uint32_t *__restrict__ err, uint64_t w) {
  constexpr int kVq = (3 * L + 3) / 4;  // 16-byte pieces holding v
  const uint64_t n = sv.n;
  const uint32_t lane = threadIdx.x & 63, s = lane & 3;
  const uint64_t nstep = (n + 127) / 128;
  // source row of this lane's element i of step ww, or -1
  auto load_idx = [&](uint64_t ww, int32_t (&ix)[2]) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const uint64_t e = ww * 128 + 64 * i + lane;
      ix[i] = ww < nstep && e < n ? idx[e] : -1;
    }
  };
  auto load_rec = [&](const int32_t (&ix)[2], uint4 (&pv)[2][kVq], uint4 (&pq)[2]) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      if (ix[i] >= 0) {
        const uint4 *rs = reinterpret_cast<const uint4 *>(Rin + (uint64_t)(uint32_t)ix[i] * kRec);
#pragma unroll
        for (int q = 0; q < kVq; q++) pv[i][q] = rs[q];
        pq[i] = rs[kRecPerm / 4];
      }
    }
  };
  const uint64_t base = w * 128;
  // the step's records, in flight during its AES work (loading a step
  // ahead measured slower: the registers it holds cost more than it hides)
  int32_t ci[2];
  uint4 cv[2][kVq], cq[2];
  load_idx(w, ci);
  load_rec(ci, cv, cq);
  // this lane's perm block: call lane & 1 of the group of block (lane >> 1) & 1
  uint4 pb;
  {
    uint64_t off;
    uint32_t ctr, bib;
    sched_locate(sp, base + 64 * ((lane >> 1) & 1) + (lane & ~3u), lane & 1, off, ctr, bib);
    pb = aes_ctr_block<kCopies>(A, rk, np + off, ctr);
  }
  // word s of quad lane c's block, for each c: calls 0 and 1 of blocks 0 and 1
  uint32_t pm[4];
  {
    // every lane takes word s: each word is broadcast from quad lane c (all
    // lanes active), then selected
    const uint4 b0 = quad_dpp4<0x00>(pb), b1 = quad_dpp4<0x55>(pb);
    pm[0] = word_of(b0, s);
    pm[1] = word_of(b1, s);
    const uint4 b2 = quad_dpp4<0xAA>(pb), b3 = quad_dpp4<0xFF>(pb);
    pm[2] = word_of(b2, s);
    pm[3] = word_of(b3, s);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const uint64_t e = base + 64 * i + lane;
    const bool live = e < n;
    const uint32_t m0 = pm[2 * i], m1 = pm[2 * i + 1];
    const uint32_t m2 = 0u - m0 - m1;
    // v masks: element e's bytes of calls 0 and 1
    uint32_t r[2][L];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      uint64_t off;
      uint32_t ctr, bib;
      sched_locate(sv, live ? e : 0, j, off, ctr, bib);
      uint4 blk = aes_ctr_block<kCopies>(A, rk, nv + off, ctr);
      uint32_t w4[4] = {blk.x, blk.y, blk.z, blk.w};
#pragma unroll
      for (int q = 0; q < L; q++) r[j][q] = w4[bib / 4 + q];
    }
    uint32_t wd[kRec];
#pragma unroll
    for (int q = 0; q < kVq; q++) {
      wd[4 * q] = cv[i][q].x;
      wd[4 * q + 1] = cv[i][q].y;
      wd[4 * q + 2] = cv[i][q].z;
      wd[4 * q + 3] = cv[i][q].w;
    }
#pragma unroll
    for (int q = 4 * kVq; q < kRec; q++) wd[q] = 0u;
#pragma unroll
    for (int q = 0; q < L; q++) {
      wd[0 * L + q] ^= r[0][q];
      wd[1 * L + q] ^= r[1][q];
      wd[2 * L + q] ^= r[0][q] ^ r[1][q];
    }
    wd[kRecPerm + 0] = cq[i].x + m0;
    wd[kRecPerm + 1] = cq[i].y + m1;
    wd[kRecPerm + 2] = cq[i].z + m2;
    // destination row: the element itself, or (scatter) its opened perm
    uint32_t dst = (uint32_t)e;
    bool ok = live;
    if (SCATTER) {
      dst = wd[kRecPerm] + wd[kRecPerm + 1] + wd[kRecPerm + 2];
      if (live && dst >= n) {
        atomicOr(err, 1u);
        ok = false;
      }
    }
    uint4 X[4];
#pragma unroll
    for (int q = 0; q < 4; q++) X[q] = make_uint4(wd[4 * q], wd[4 * q + 1], wd[4 * q + 2], wd[4 * q + 3]);
    quad_transpose(X, s);
    // store t: quarter s of the record of quad lane t (whole records per
    // instruction); a dropped row is marked by an out-of-range destination
    const uint32_t dk = ok ? dst : 0xffffffffu;
#define ORQG_STORE(t, CTRL)                                                          \
{                                                                                  \
  const uint32_t dt = quad_dpp<CTRL>(dk);                                          \
  if (dt != 0xffffffffu) reinterpret_cast<uint4 *>(Rout + (uint64_t)dt * kRec)[s] = X[t]; \
}
    ORQG_STORE(0, 0x00)
    ORQG_STORE(1, 0x55)
    ORQG_STORE(2, 0xAA)
    ORQG_STORE(3, 0xFF)
#undef ORQG_STORE
  }
}

template <int L, bool SCATTER>
__global__ void __launch_bounds__(kBT) k_round2(const uint32_t *__restrict__ t0,
                                                const uint32_t *__restrict__ rk, uint64_t nv,
                                                CallSched sv, uint64_t np, CallSched sp,
                                                const int32_t *__restrict__ idx,
                                                const uint32_t *__restrict__ Rin,
                                                uint32_t *__restrict__ Rout,
                                                uint32_t *__restrict__ err) {
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t nstep = (sv.n + 127) / 128;
  const uint64_t wstride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nstep; w += wstride)
    round2_step<L, SCATTER>(A, rk, nv, sv, np, sp, idx, Rin, Rout, err, w);
}

// remove_padding (sorting.h:475-495): perm_b = low 32 bits (B-shared int32),
// v = bits 32.. of each share.  DESC reverses the element order first.
template <int L, typename K>
__global__ void k_unpad(uint64_t n, bool desc, const uint32_t *__restrict__ P, K *__restrict__ v,
                        uint32_t *__restrict__ pb) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t src = desc ? n - 1 - i : i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const uint32_t *p = P + src * kRec + k * L;
      pb[3 * i + k] = p[0];
      uint64_t key = p[1];
      if (L == 3) key |= (uint64_t)p[2] << 32;
      v[3 * i + k] = (K)key;
    }
  }
}

// open a [n][3] uint32 vector (XOR or sum) and write it as the public sharing
// (o, 0, 0) into out (open_shares_* then public_share)
// Returns the record buffer that holds the sorted vector (P or the second one).
template <int L>
uint32_t *sort_body(orqg_ctx *c, uint32_t *P, uint64_t n, int bits, bool full_width) {
  const uint32_t tszv = L == 3 ? 16 : 8;  // PadWidth element size
  uint32_t *F = (uint32_t *)c->scratch_get("rs_F", n * 12 + 64);
  uint32_t *R1 = (uint32_t *)c->scratch_get("rs_R1", n * kRec * 4 + 64);
  const uint32_t nblk = (uint32_t)((n + kEPB - 1) / kEPB);
  uint32_t *bsum = (uint32_t *)c->scratch_get("rs_bsum", (size_t)nblk * 12 + 64);
  uint32_t *tot = (uint32_t *)c->scratch_get("rs_tot", 64);
  uint32_t *err = (uint32_t *)c->scratch_get("rs_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  const unsigned grid_rounds = grid_for((n + 3) / 4, kBT, c->num_cu);
  const unsigned grid_pers = std::min<unsigned>(nblk, (unsigned)c->num_cu);
  uint32_t *cur = P, *other = R1;

  for (int i = 0; i < bits; i++) {
    const int bitpos = 32 + i;
    const uint32_t flip = (full_width && i == bits - 1) ? 1u : 0u;  // inplace_invert on the MSB
    NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
    {
      TimedRegion tr(c, "sort_b2a");
      CallSched s = make_sched(n, c->batch_size, 4, 3);
      hipLaunchKernelGGL(k_bit_b2a<L>, dim3(grid_pers), dim3(kBT), 0, c->stream, c->d_t0, c->rk(kPair0),
                         nb, s, cur, bitpos / 32, bitpos % 32, flip, F, bsum);
      uint64_t used = sched_nonces(s);
      for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
      for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * n * 4);
      hipLaunchKernelGGL(k_scan3, dim3(1), dim3(1024), 0, c->stream, nblk, bsum, tot);
    }
    {
      TimedRegion tr(c, "sort_perm");
      NonceTriple nb2 = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
      CallSched s = make_sched(n, c->batch_size, 4, 1);
      hipLaunchKernelGGL(k_bit_perm<L>, dim3(grid_pers), dim3(kBT), 0, c->stream, c->d_t0,
                         c->rk(kPair0), nb2, s, F, bsum, tot, cur);
      uint64_t used = sched_nonces(s);
      for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
      for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, n * 4);
    }
    // oblivious_apply_elementwise_perm(v, perm): pi from the queue
    ShardedPerm pi = next_perm(c, n);
    {
      TimedRegion tr(c, "sort_round");
      // key G_g is used first by v's round g, then by perm's round g
      // cur -> other -> cur -> other (scattered)
      uint32_t *rin = cur, *rout = other;
      CallSched sv = make_sched(n, c->batch_size, tszv, 2);
      CallSched sp = make_sched(n, c->batch_size, 4, 2);
      const uint64_t uv = sched_nonces(sv), up = sched_nonces(sp);
      for (int g = 0; g < 3; g++) {
        const int role = kGroup0 + g;
        const uint64_t nv = c->nonce[role], np = nv + uv;
        {
          TimedRegion tk(c, g < 2 ? "k_round2" : "k_round2_scatter");
          if (g < 2)
            hipLaunchKernelGGL((k_round2<L, false>), dim3(grid_rounds), dim3(kBT), 0, c->stream,
                               c->d_t0, c->rk(role), nv, sv, np, sp, pi.inv[g], rin, rout, err);
          else
            hipLaunchKernelGGL((k_round2<L, true>), dim3(grid_rounds), dim3(kBT), 0, c->stream,
                               c->d_t0, c->rk(role), nv, sv, np, sp, pi.inv[g], rin, rout, err);
        }
        c->nonce[role] += uv + up;
        std::swap(rin, rout);
      }
      // log in the reference's order: three rounds on v, then three on perm
      for (int g = 0; g < 3; g++) log_reshare(c, g, n * tszv);
      for (int g = 0; g < 3; g++) log_reshare(c, g, n * 4);
      for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, n * 4);
      std::swap(cur, other);
    }
    free_perm(pi);
    ORQG_HIP(hipGetLastError());
  }
  uint32_t herr = 0;
  ORQG_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  ORQG_HIP(hipStreamSynchronize(c->stream));
  if (herr) throw Error(ORQG_E_PARITY, "radix_sort: opened bit permutation out of range");
  check_swaps(c);
  return cur;
}

}  // namespace

void radix_sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out) {
  const uint64_t n = v->n;
  const bool desc = order == ORQG_DESC;
  const bool full_width = bits == 8 * v->width;
  if (n == 0) return;
  // 1 permutation + (bits + 1) pairs; pairs are single permutations in 3PC
  // (radixsort.h:121-128, permutation_manager.h:108-111)
  reserve_perms(c, n, (size_t)bits + 2);
  const int L = v->width == 8 ? 3 : 2;
  uint32_t *P = (uint32_t *)c->scratch_get("rs_P", n * kRec * 4 + 64);
  const unsigned grid_flat = grid_for(n, 256, c->num_cu * 16);
  if (L == 3) {
    hipLaunchKernelGGL((k_pad<3, uint64_t>), dim3(grid_flat), dim3(256), 0, c->stream, n, desc,
                       (const uint64_t *)v->d, P);
    P = sort_body<3>(c, P, n, bits, full_width);
  } else {
    hipLaunchKernelGGL((k_pad<2, uint32_t>), dim3(grid_flat), dim3(256), 0, c->stream, n, desc,
                       (const uint32_t *)v->d, P);
    P = sort_body<2>(c, P, n, bits, full_width);
  }
  // remove_padding
  uint32_t *pb = (uint32_t *)perm_out->d;
  if (L == 3)
    hipLaunchKernelGGL((k_unpad<3, uint64_t>), dim3(grid_flat), dim3(256), 0, c->stream, n, desc, P,
                       (uint64_t *)v->d, pb);
  else
    hipLaunchKernelGGL((k_unpad<2, uint32_t>), dim3(grid_flat), dim3(256), 0, c->stream, n, desc, P,
                       (uint32_t *)v->d, pb);
  ORQG_HIP(hipGetLastError());
  TimedRegion tr(c, "sort_tail");
  // ElementwisePermutation::b2a, negate (DESC) and invert
  // (permutation.h:205-280), as radix_sort does after remove_padding
  // (radixsort.h:222-227, sorting.h:475-495)
  perm_out->enc = ORQG_BSHARED;
  perm_b2a(c, perm_out);
  if (desc) perm_negate(c, perm_out);
  perm_invert(c, perm_out);
  ORQG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace orqg
