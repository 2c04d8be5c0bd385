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
#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <thread>
#include <string>

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

// exclusive scan of [nb][3] block sums in place; totals to tot[3] (one
// workgroup of 1024 threads)
__device__ inline void scan3_body(uint32_t nb, uint32_t *__restrict__ bsum,
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

__global__ void __launch_bounds__(1024) k_scan3(uint32_t nb, uint32_t *__restrict__ bsum,
                                                uint32_t *__restrict__ tot) {
  scan3_body(nb, bsum, tot);
}

// s0_k(e) = one_k*e - F_k(e), s1_k(e) = one_k*(n-1) - Ftot_k + F_k(e) with F the
// inclusive prefix of f1 and one = (1,0,0); perm = s0 + f1*(s1 - s0), with
// multiply_a's zero sharing from the pair PRGs at nb.  Rows e0..e0+3 of one
// superblock: f = their f1 shares, pre = F's exclusive prefix at row e0,
// tot = Ftot.  The rows R and the schedule s cover start at global row row0
// of a vector of ntot rows (one worker thread's range; 0 and s.n for one
// thread).  The result goes into the rows' records as one 32-byte piece
// (words 8..15): word 8 is v's last word for L = 3 (kept; 0 for L = 2),
// words 9..11 take fw (f1's shares, or zeros), 12..14 the permutation.
template <int L>
__device__ inline void bit_perm_rows(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                     const NonceTriple &nb, const CallSched &s, uint64_t e0,
                                     const uint32_t (&f)[4][3], uint32_t (&pre)[3],
                                     const uint32_t (&tot)[3], uint32_t *R, uint64_t row0,
                                     uint64_t ntot, bool keep_f) {
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
  const uint32_t n1 = (uint32_t)(ntot - 1);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint32_t s0[3], y[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      pre[k] += f[i][k];  // inclusive
      uint32_t one = k == 0 ? 1u : 0u;
      s0[k] = one * (uint32_t)(row0 + e0 + i) - pre[k];
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
      // one whole 32-byte piece, so that no partial write needs a
      // read-modify-write
      uint32_t *rec = R + (e0 + i) * kRec;
      const uint32_t w8 = L == 3 ? rec[8] : 0u;
      uint4 *o = reinterpret_cast<uint4 *>(rec + 8);
      o[0] = keep_f ? make_uint4(w8, f[i][0], f[i][1], f[i][2]) : make_uint4(w8, 0u, 0u, 0u);
      o[1] = make_uint4(pr[0], pr[1], pr[2], 0u);
    }
  }
}

// bit_perm_rows from f1 stored in F and the superblock's exclusive offsets
// boff3 (3 words)
template <int L>
__device__ inline void bit_perm_block(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk_pair,
                                      const NonceTriple &nb, const CallSched &s, uint64_t e0,
                                      const uint32_t *__restrict__ F,
                                      const uint32_t *__restrict__ boff3,
                                      const uint32_t *__restrict__ tot, uint32_t *__restrict__ R,
                                      uint32_t *s_w, uint64_t row0 = 0, uint64_t ntot = 0) {
  if (ntot == 0) ntot = s.n;
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
  uint32_t pre[3], t3[3] = {tot[0], tot[1], tot[2]};
  {
    uint32_t t[3];
    bexcl3(loc, s_w, pre, t);
#pragma unroll
    for (int k = 0; k < 3; k++) pre[k] += boff3[k];
  }
  bit_perm_rows<L>(A, rk_pair, nb, s, e0, f, pre, t3, R, row0, ntot, false);
}

// ---- one bit in one pass (one worker thread): b2a_bit, then the permutation
// with F's superblock offset B and total T left out.  Both enter the shares
// linearly: with D = 2B - T, perm_p gains -B_p + f_p*(D_p + D_q) + f_q*D_p
// (q = p+1 mod 3), which the first sharded round adds to each gathered row
// from the row's f1 shares (record words 9..11) and its superblock's (B, D).
template <int L>
__global__ void __launch_bounds__(kBT) k_bit_fused(const uint32_t *__restrict__ t0,
                                                   const uint32_t *__restrict__ rk_pair,
                                                   NonceTriple nb, CallSched s, NonceTriple nb2,
                                                   CallSched s2, uint32_t *P, int word, int bit,
                                                   uint32_t flip, uint32_t *__restrict__ bsum) {
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
    // the next superblock's words are in flight during this one's work; its
    // rows are not the ones this superblock writes
    const uint32_t nx = sb + gridDim.x < nsb ? sb + gridDim.x : sb;
    bit_load_raw<L>(P, s.n, word, (uint64_t)nx * kEPB + threadIdx.x * 4, raw);
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
    uint32_t pre[3], btot[3];
    bexcl3(loc, s_w, pre, btot);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < 3; k++) bsum[sb * 3 + k] = btot[k];
    const uint32_t zero3[3] = {0u, 0u, 0u};
    bit_perm_rows<L>(A, rk_pair, nb2, s2, e0, f, pre, zero3, P, 0, s.n, true);
  }
}

// the exclusive scan of k_bit_fused's superblock sums into C[sb] = (B0, B1,
// B2, 0), one 16-byte entry per superblock; the totals T to tot
__device__ inline void scan3_corr_body(uint32_t nb, uint32_t *__restrict__ bsum,
                                       uint32_t *__restrict__ tot, uint4 *__restrict__ C) {
  scan3_body(nb, bsum, tot);
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x)
    C[j] = make_uint4(bsum[j * 3], bsum[j * 3 + 1], bsum[j * 3 + 2], 0u);
}

__global__ void __launch_bounds__(1024) k_scan3_corr(uint32_t nb, uint32_t *__restrict__ bsum,
                                                     uint32_t *__restrict__ tot,
                                                     uint4 *__restrict__ C) {
  scan3_corr_body(nb, bsum, tot, C);
}

template <int L>
__global__ void __launch_bounds__(kBT) k_bit_perm(const uint32_t *__restrict__ t0,
                                                  const uint32_t *__restrict__ rk_pair,
                                                  NonceTriple nb, CallSched s,
                                                  const uint32_t *__restrict__ F,
                                                  const uint32_t *__restrict__ boff,
                                                  const uint32_t *__restrict__ tot,
                                                  uint32_t *__restrict__ R, uint64_t row0,
                                                  uint64_t ntot) {
  AES_LDS;
  __shared__ uint32_t s_w[3 * (kBT / 64)];
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint32_t nsb = (uint32_t)((s.n + kEPB - 1) / kEPB);
  for (uint32_t sb = blockIdx.x; sb < nsb; sb += gridDim.x)
    bit_perm_block<L>(A, rk_pair, nb, s, (uint64_t)sb * kEPB + threadIdx.x * 4, F, boff + sb * 3, tot,
                   R, s_w, row0, ntot);
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
// rows (every wave of the launch takes its own steps).  Those rows are global
// rows row0.. of a vector of ntot rows (one worker thread's destination
// range; 0 and n for one thread): idx is indexed from row0, a gathered row
// and an opened destination are global.
//
// CORR (the first round after k_bit_fused): each gathered row's permutation
// shares first gain the terms of F's superblock offset B and total T that
// k_bit_fused left out, from the row's f1 shares (record words 9..11),
// C[row / kEPB] = B and T = tot[0..2].
template <int L, bool SCATTER, bool CORR = false>
__device__ inline void round2_step(const AesCtx<kCopies> &A, const uint32_t *__restrict__ rk,
                                   uint64_t nv, const CallSched &sv, uint64_t np,
                                   const CallSched &sp, const int32_t *__restrict__ idx,
                                   const uint32_t *__restrict__ Rin, uint32_t *__restrict__ Rout,
                                   uint32_t *__restrict__ err, uint64_t w, uint64_t row0 = 0,
                                   uint64_t ntot = 0, const uint4 *__restrict__ C = nullptr,
                                   const uint32_t *__restrict__ tot = nullptr) {
  constexpr int kVq = (3 * L + 3) / 4;  // 16-byte pieces holding v
  const uint64_t n = sv.n;
  if (ntot == 0) ntot = n;
  const uint32_t lane = threadIdx.x & 63, s = lane & 3;
  const uint64_t nstep = (n + 127) / 128;
  // source row of this lane's element i of step ww, or -1
  auto load_idx = [&](uint64_t ww, int32_t (&ix)[2]) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const uint64_t e = ww * 128 + 64 * i + lane;
      ix[i] = ww < nstep && e < n ? idx[row0 + e] : -1;
    }
  };
  // CORR: pf = record words 8..11 (for L = 3 that is pv's last piece), pc
  // = the row's B
  constexpr int kF = CORR && L != 3 ? 1 : 0, kC = CORR ? 1 : 0;
  auto load_rec = [&](const int32_t (&ix)[2], uint4 (&pv)[2][kVq], uint4 (&pq)[2],
                      uint4 (&pf)[2][kF + 1], uint4 (&pc)[2][kC + 1]) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      if (ix[i] >= 0) {
        const uint4 *rs = reinterpret_cast<const uint4 *>(Rin + (uint64_t)(uint32_t)ix[i] * kRec);
#pragma unroll
        for (int q = 0; q < kVq; q++) pv[i][q] = rs[q];
        pq[i] = rs[kRecPerm / 4];
        if (kF) pf[i][0] = rs[2];
        if (kC) pc[i][0] = C[(uint32_t)ix[i] / (uint32_t)kEPB];
      }
    }
  };
  const uint64_t base = w * 128;
  // the step's records, in flight during its AES work (loading a step
  // ahead measured slower: the registers it holds cost more than it hides)
  int32_t ci[2];
  uint4 cv[2][kVq], cq[2], cf[2][kF + 1], cc[2][kC + 1];
  load_idx(w, ci);
  load_rec(ci, cv, cq, cf, cc);
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
    uint32_t pq3[3] = {cq[i].x, cq[i].y, cq[i].z};
    if (CORR) {
      const uint4 fw = kF ? cf[i][0] : cv[i][kVq - 1];  // words 8..11
      const uint32_t f[3] = {fw.y, fw.z, fw.w};
      const uint32_t B[3] = {cc[i][0].x, cc[i][0].y, cc[i][0].z};
      const uint32_t D[3] = {2u * B[0] - tot[0], 2u * B[1] - tot[1], 2u * B[2] - tot[2]};
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const int q = (p + 1) % 3;
        pq3[p] += f[p] * (D[p] + D[q]) + f[q] * D[p] - B[p];
      }
    }
    wd[kRecPerm + 0] = pq3[0] + m0;
    wd[kRecPerm + 1] = pq3[1] + m1;
    wd[kRecPerm + 2] = pq3[2] + m2;
    // destination row: the element itself, or (scatter) its opened perm
    uint32_t dst = (uint32_t)(row0 + e);
    bool ok = live;
    if (SCATTER) {
      dst = wd[kRecPerm] + wd[kRecPerm + 1] + wd[kRecPerm + 2];
      if (live && dst >= ntot) {
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

template <int L, bool SCATTER, bool CORR = false>
__global__ void __launch_bounds__(kBT) k_round2(const uint32_t *__restrict__ t0,
                                                const uint32_t *__restrict__ rk, uint64_t nv,
                                                CallSched sv, uint64_t np, CallSched sp,
                                                const int32_t *__restrict__ idx,
                                                const uint32_t *__restrict__ Rin,
                                                uint32_t *__restrict__ Rout,
                                                uint32_t *__restrict__ err, uint64_t row0,
                                                uint64_t ntot, const uint4 *__restrict__ C = nullptr,
                                                const uint32_t *__restrict__ tot = nullptr) {
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const uint64_t nstep = (sv.n + 127) / 128;
  const uint64_t wstride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nstep; w += wstride)
    round2_step<L, SCATTER, CORR>(A, rk, nv, sv, np, sp, idx, Rin, Rout, err, w, row0, ntot, C,
                                  tot);
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

// ---- segmented bit kernels: one bit iteration of many sorts per launch -------------
// Each sort (a "segment": its own context, size, buffers, keys, nonces and
// permutation) is described by a SegBit per bit.  The superblocks (k_bit_b2a,
// k_bit_perm) and wave steps (rounds) of all segments are numbered flat; a
// workgroup (wave, for the rounds) finds its segment by binary search over
// the segments' first flat indices and runs the unsegmented kernel's body on
// it.  Which workgroup takes which item changes nothing in the results, so
// the launch shapes are free to tune.
struct SegBit {
  const uint32_t *rk_pair;    // pair keys (3 x 60 words)
  const uint32_t *rk_grp[3];  // group keys
  uint32_t *cur, *other;      // record buffers: the bit reads cur, its scatter writes other
  uint32_t *F, *bsum, *tot, *err;
  const int32_t *idx[3];      // pi_g^{-1}
  NonceTriple nb, nb2;        // b2a_bit, multiply_a
  CallSched s_b2a, s_perm, sv, sp;
  uint64_t nv[3], np[3];      // round g: v masks, perm masks
  uint64_t st0;               // first flat wave step
  uint32_t sb0, nsb;          // first flat superblock, superblocks
  uint32_t flip;
};

// The descriptors are read through the constant address space (they do not
// change during a launch), so that a wave-uniform index gives scalar loads and
// the fields live in scalar registers, as kernel arguments do.
typedef const __attribute__((address_space(4))) SegBit *SegPtr;

__device__ inline SegPtr seg_ptr(const SegBit *segs) { return (SegPtr)segs; }

// field-wise copies out of the constant address space
__device__ inline CallSched ld_sched(const __attribute__((address_space(4))) CallSched &x) {
  CallSched r;
  r.n = x.n;
  r.bs = x.bs;
  r.cfull = x.cfull;
  r.tsz = x.tsz;
  r.calls = x.calls;
  return r;
}
__device__ inline NonceTriple ld_nonces(const __attribute__((address_space(4))) NonceTriple &x) {
  NonceTriple r;
  r.v[0] = x.v[0];
  r.v[1] = x.v[1];
  r.v[2] = x.v[2];
  return r;
}

__device__ inline uint32_t seg_of_sb(SegPtr segs, uint32_t nseg, uint32_t f) {
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].sb0 <= f) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t seg_of_step(SegPtr segs, uint32_t nseg, uint64_t w) {
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].st0 <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

// k_bit_fused over the superblocks of every request (numbered flat); a
// request's superblock prefix entries C go to its F buffer, which the fused
// bit does not otherwise use (16 bytes per superblock of 4096 rows fit in the
// 12 bytes per row F holds)
template <int L>
__global__ void __launch_bounds__(kBT) k_bit_fused_seg(const uint32_t *__restrict__ t0,
                                                       const SegBit *__restrict__ segs,
                                                       uint32_t nseg, uint32_t total, int word,
                                                       int bit) {
  AES_LDS;
  __shared__ uint32_t s_w[3 * (kBT / 64)];
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const SegPtr sp = seg_ptr(segs);
  for (uint32_t f = blockIdx.x; f < total; f += gridDim.x) {
    const SegPtr S = sp + seg_of_sb(sp, nseg, f);
    const CallSched sch = ld_sched(S->s_b2a), sch2 = ld_sched(S->s_perm);
    const NonceTriple nb = ld_nonces(S->nb), nb2 = ld_nonces(S->nb2);
    const uint32_t sb = f - S->sb0;
    const uint64_t e0 = (uint64_t)sb * kEPB + threadIdx.x * 4;
    uint32_t raw[4][3], bits[4][3];
    bit_load_raw<L>(S->cur, sch.n, word, e0, raw);
    bit_extract(raw, e0, sch.n, bit, S->flip, bits);
    uint32_t fv[4][3];
    b2a_bit_group<uint32_t>(A, S->rk_pair, nb, sch, e0, bits, fv);
    uint32_t loc[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (e0 + i >= sch.n) fv[i][k] = 0u;
        loc[k] += fv[i][k];
      }
    uint32_t pre[3], btot[3];
    bexcl3(loc, s_w, pre, btot);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < 3; k++) S->bsum[sb * 3 + k] = btot[k];
    const uint32_t zero3[3] = {0u, 0u, 0u};
    bit_perm_rows<L>(A, S->rk_pair, nb2, sch2, e0, fv, pre, zero3, S->cur, 0, sch.n, true);
  }
}
__global__ void __launch_bounds__(1024) k_scan3_corr_seg(const SegBit *__restrict__ segs) {
  const SegPtr S = seg_ptr(segs) + blockIdx.x;
  scan3_corr_body(S->nsb, S->bsum, S->tot, reinterpret_cast<uint4 *>(S->F));
}
template <int L, bool SCATTER, bool CORR = false>
__global__ void __launch_bounds__(kBT) k_round2_seg(const uint32_t *__restrict__ t0,
                                                    const SegBit *__restrict__ segs, uint32_t nseg,
                                                    uint64_t total, int g) {
  AES_LDS;
  aes_load_tables<kCopies>(lds, t0);
  AesCtx<kCopies> A(lds);
  const SegPtr segp = seg_ptr(segs);
  const uint64_t wstride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  for (uint64_t w = w0; w < total; w += wstride) {
    // w is the same across the wave (total < 2^32 steps): the search and the
    // segment's fields then use scalar loads
    const uint32_t wu = __builtin_amdgcn_readfirstlane((uint32_t)w);
    const SegPtr S = segp + seg_of_step(segp, nseg, wu);
    const CallSched sv = ld_sched(S->sv), sp = ld_sched(S->sp);
    const uint32_t *rin = g == 1 ? S->other : S->cur;
    uint32_t *rout = g == 1 ? S->cur : S->other;
    round2_step<L, SCATTER, CORR>(A, S->rk_grp[g], S->nv[g], sv, S->np[g], sp, S->idx[g], rin,
                                  rout, S->err, wu - S->st0, 0, 0,
                                  reinterpret_cast<const uint4 *>(S->F), S->tot);
  }
}

// Returns the record buffer that holds the sorted vector (P or the second one).
template <int L>
uint32_t *sort_body(orqg_ctx *c, uint32_t *P, uint64_t n, int bits, bool full_width) {
  const uint32_t tszv = L == 3 ? 16 : 8;  // PadWidth element size
  uint32_t *F = (uint32_t *)c->scratch_get("rs_F", n * 12 + 64);
  uint32_t *R1 = (uint32_t *)c->scratch_get("rs_R1", n * kRec * 4 + 64);
  // superblocks are counted per worker-thread range (one range per thread),
  // their sums laid out range after range
  const uint32_t nblk_max = (uint32_t)((n + kEPB - 1) / kEPB) + (uint32_t)c->threads;
  uint32_t *bsum = (uint32_t *)c->scratch_get("rs_bsum", (size_t)nblk_max * 12 + 64);
  uint32_t *tot = (uint32_t *)c->scratch_get("rs_tot", 64);
  uint32_t *err = (uint32_t *)c->scratch_get("rs_err", 16);
  ORQG_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  uint32_t *cur = P, *other = R1;
  // one worker thread: b2a_bit and the permutation in one pass (k_bit_fused),
  // F's offsets and total added by the first round from (B, D) per superblock
  const bool fused = c->threads <= 1;
  const uint32_t nsb1 = (uint32_t)((n + kEPB - 1) / kEPB);
  uint4 *corr = fused ? (uint4 *)c->scratch_get("rs_corr", (size_t)nsb1 * 16 + 64) : nullptr;

  for (int i = 0; i < bits; i++) {
    const int bitpos = 32 + i;
    const uint32_t flip = (full_width && i == bits - 1) ? 1u : 0u;  // inplace_invert on the MSB
    if (fused) {
      TimedRegion tr(c, "sort_b2a");
      NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
      CallSched s = make_sched(n, c->batch_size, 4, 3);
      const uint64_t used = sched_nonces(s);
      NonceTriple nb2 = {{nb.v[0] + used, nb.v[1] + used, nb.v[2] + used}};
      CallSched s2 = make_sched(n, c->batch_size, 4, 1);
      const uint64_t used2 = sched_nonces(s2);
      hipLaunchKernelGGL(k_bit_fused<L>, dim3(std::min<unsigned>(nsb1, (unsigned)c->num_cu)),
                         dim3(kBT), 0, c->stream, c->d_t0, c->rk(kPair0), nb, s, nb2, s2, cur,
                         bitpos / 32, bitpos % 32, flip, bsum);
      hipLaunchKernelGGL(k_scan3_corr, dim3(1), dim3(1024), 0, c->stream, nsb1, bsum, tot, corr);
      for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used + used2;
      for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * n * 4);
      for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, n * 4);
    } else {
      TimedRegion tr(c, "sort_b2a");
      uint32_t sb0 = 0;
      for_threads(c, n, [&](uint64_t off, uint64_t m) {
        NonceTriple nb = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
        CallSched s = make_sched(m, c->batch_size, 4, 3);
        const uint32_t nsb = (uint32_t)((m + kEPB - 1) / kEPB);
        hipLaunchKernelGGL(k_bit_b2a<L>, dim3(std::min<unsigned>(nsb, (unsigned)c->num_cu)),
                           dim3(kBT), 0, c->stream, c->d_t0, c->rk(kPair0), nb, s, cur + off * kRec,
                           bitpos / 32, bitpos % 32, flip, F + off * 3, bsum + (size_t)sb0 * 3);
        uint64_t used = sched_nonces(s);
        for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
        sb0 += nsb;
      });
      for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * n * 4);
      hipLaunchKernelGGL(k_scan3, dim3(1), dim3(1024), 0, c->stream, sb0, bsum, tot);
    }
    if (!fused) {
      TimedRegion tr(c, "sort_perm");
      uint32_t sb0 = 0;
      for_threads(c, n, [&](uint64_t off, uint64_t m) {
        NonceTriple nb2 = {{c->nonce[kPair0], c->nonce[kPair0 + 1], c->nonce[kPair0 + 2]}};
        CallSched s = make_sched(m, c->batch_size, 4, 1);
        const uint32_t nsb = (uint32_t)((m + kEPB - 1) / kEPB);
        hipLaunchKernelGGL(k_bit_perm<L>, dim3(std::min<unsigned>(nsb, (unsigned)c->num_cu)),
                           dim3(kBT), 0, c->stream, c->d_t0, c->rk(kPair0), nb2, s, F + off * 3,
                           bsum + (size_t)sb0 * 3, tot, cur + off * kRec, off, n);
        uint64_t used = sched_nonces(s);
        for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += used;
        sb0 += nsb;
      });
      for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, n * 4);
    }
    // oblivious_apply_elementwise_perm(v, perm): pi from the queue
    ShardedPerm pi = next_perm(c, n);
    {
      TimedRegion tr(c, "sort_round");
      // key G_g is used first by v's round g, then by perm's round g (each
      // worker thread's own G_g, over its range of destination rows)
      // cur -> other -> cur -> other (scattered)
      uint32_t *rin = cur, *rout = other;
      for (int g = 0; g < 3; g++) {
        const int role = kGroup0 + g;
        TimedRegion tk(c, g < 2 ? "k_round2" : "k_round2_scatter");
        for_threads(c, n, [&](uint64_t off, uint64_t m) {
          CallSched sv = make_sched(m, c->batch_size, tszv, 2);
          CallSched sp = make_sched(m, c->batch_size, 4, 2);
          const uint64_t uv = sched_nonces(sv), up = sched_nonces(sp);
          const uint64_t nv = c->nonce[role], np = nv + uv;
          const unsigned grid = grid_for((m + 3) / 4, kBT, c->num_cu);
          if (g == 0 && fused)
            hipLaunchKernelGGL((k_round2<L, false, true>), dim3(grid), dim3(kBT), 0, c->stream,
                               c->d_t0, c->rk(role), nv, sv, np, sp, pi.inv[g], rin, rout, err,
                               off, n, (const uint4 *)corr, (const uint32_t *)tot);
          else if (g < 2)
            hipLaunchKernelGGL((k_round2<L, false>), dim3(grid), dim3(kBT), 0, c->stream, c->d_t0,
                               c->rk(role), nv, sv, np, sp, pi.inv[g], rin, rout, err, off, n);
          else
            hipLaunchKernelGGL((k_round2<L, true>), dim3(grid), dim3(kBT), 0, c->stream, c->d_t0,
                               c->rk(role), nv, sv, np, sp, pi.inv[g], rin, rout, err, off, n);
          c->nonce[role] += uv + up;
        });
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

namespace {

// remove_padding (into v and perm_out) and the permutation's b2a, negate
// (DESC) and invert, on c's stream; P holds the sorted records
void sort_finish(orqg_ctx *c, orqg_vec *v, uint32_t *P, bool desc, orqg_vec *perm_out) {
  const uint64_t n = v->n;
  const unsigned grid_flat = grid_for(n, 256, c->num_cu * 16);
  uint32_t *pb = (uint32_t *)perm_out->d;
  if (v->width == 8)
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
}

void pad_records(orqg_ctx *c, hipStream_t st, const orqg_vec *v, bool desc, uint32_t *P) {
  const uint64_t n = v->n;
  const unsigned grid_flat = grid_for(n, 256, c->num_cu * 16);
  if (v->width == 8)
    hipLaunchKernelGGL((k_pad<3, uint64_t>), dim3(grid_flat), dim3(256), 0, st, n, desc,
                       (const uint64_t *)v->d, P);
  else
    hipLaunchKernelGGL((k_pad<2, uint32_t>), dim3(grid_flat), dim3(256), 0, st, n, desc,
                       (const uint32_t *)v->d, P);
  ORQG_HIP(hipGetLastError());
}

// ---- launch shapes of the segmented kernels, tuned online ------------------------
// The workgroups of a launch are a free choice (results do not depend on it).
// For each kernel kind, record layout and size class (log2 of the flat items)
// the candidates are tried on live bit iterations -- one launch each, timed
// with events read after the call's final synchronisation -- and the fastest
// is kept for the rest of the process.  ORQG_TUNE=0 keeps the defaults.
enum TuneKind { kTuneB2a = 0, kTuneRound = 2 };

struct Tuner {
  std::mutex mu;
  std::map<uint64_t, unsigned> best;
  std::map<uint64_t, std::vector<unsigned>> cands;
  std::map<uint64_t, std::map<unsigned, std::pair<float, int>>> tried;  // (fastest ms, samples)
};
Tuner &tuner() {
  static Tuner t;
  return t;
}
bool tune_on() {
  const char *e = std::getenv("ORQG_TUNE");
  return !(e && e[0] == '0');
}
uint64_t tune_key(int kind, int L, uint64_t items) {
  uint64_t b = 0;
  while ((1ull << b) < items) b++;
  return ((uint64_t)kind << 48) | ((uint64_t)L << 40) | b;
}
// candidate workgroup counts for `units` workgroup-sized pieces of work
std::vector<unsigned> grid_candidates(uint64_t units, int num_cu) {
  std::vector<unsigned> out;
  for (unsigned g : {(unsigned)num_cu / 4, (unsigned)num_cu / 2, (unsigned)num_cu,
                     2u * (unsigned)num_cu}) {
    const unsigned x = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, units));
    if (std::find(out.begin(), out.end(), x) == out.end()) out.push_back(x);
  }
  return out;
}
constexpr int kTuneSamples = 2;  // timed launches per candidate (the fastest counts)

struct TuneTrial {
  uint64_t key;
  unsigned grid;
  hipEvent_t a, b;
};

// the launch shape for this launch: settled, a candidate still short of
// samples (then timed), or the default (the first candidate >= min(units,
// num_cu), the unsegmented kernels' choice)
unsigned tune_pick(orqg_ctx *c, int kind, int L, uint64_t units, std::vector<TuneTrial> &trials,
                   bool &timed) {
  const std::vector<unsigned> cand = grid_candidates(units, c->num_cu);
  unsigned dflt = cand.back();
  for (unsigned g : cand)
    if (g >= (unsigned)std::min<uint64_t>(units, (uint64_t)c->num_cu)) {
      dflt = g;
      break;
    }
  timed = false;
  if (!tune_on() || cand.size() == 1) return dflt;
  const uint64_t key = tune_key(kind, L, units);
  Tuner &T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  auto it = T.best.find(key);
  if (it != T.best.end()) return std::min<unsigned>(it->second, (unsigned)units);
  // a size class's candidates are those of the first size seen in it
  auto ci = T.cands.find(key);
  if (ci == T.cands.end()) ci = T.cands.emplace(key, cand).first;
  auto &done = T.tried[key];
  for (unsigned g : ci->second) {
    int samples = 0;
    auto d = done.find(g);
    if (d != done.end()) samples = d->second.second;
    for (const TuneTrial &t : trials) samples += t.key == key && t.grid == g;
    if (samples < kTuneSamples) {
      timed = true;
      return std::min<unsigned>(g, (unsigned)units);
    }
  }
  return dflt;
}

// fold the call's timed launches into the tuner (after the final
// synchronisation); a size class whose candidates all have their samples
// settles on the fastest
void tune_settle(std::vector<TuneTrial> &trials) {
  Tuner &T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  for (TuneTrial &t : trials) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      auto &e = T.tried[t.key][t.grid];
      e.first = e.second ? std::min(e.first, ms) : ms;
      e.second++;
    }
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (TuneTrial &t : trials) {
    if (T.best.count(t.key)) continue;
    const auto &cand = T.cands[t.key];
    const auto &done = T.tried[t.key];
    unsigned pick = 0;
    float fastest = 0;
    bool complete = true;
    for (unsigned g : cand) {
      auto d = done.find(g);
      if (d == done.end() || d->second.second < kTuneSamples) {
        complete = false;
        break;
      }
      if (!pick || d->second.first < fastest) {
        pick = g;
        fastest = d->second.first;
      }
    }
    if (complete) T.best[t.key] = pick;
  }
  trials.clear();
}

struct BatchReq {
  orqg_ctx *c;
  orqg_vec *v, *perm;
  bool desc, full_width;
  int bits;
  uint32_t *P, *R1, *cur, *other, *F, *bsum, *tot, *err;
};

// The bit iterations of the requests in `rq` (all with record layout L), on
// lead's stream; returns after every request's tail ran on its own stream.
template <int L>
void batch_body(orqg_ctx *lead, std::vector<BatchReq> &rq) {
  const uint32_t tszv = L == 3 ? 16 : 8;  // PadWidth element size
  int maxbits = 0;
  for (BatchReq &r : rq) maxbits = std::max(maxbits, r.bits);
  // descriptors of every bit, built with radix_sort's own nonce, queue and
  // message-log bookkeeping per request
  std::vector<SegBit> all;
  std::vector<size_t> first((size_t)maxbits + 1);
  std::vector<uint32_t> tot_sb(maxbits);
  std::vector<uint64_t> tot_st(maxbits);
  std::vector<ShardedPerm> used;
  auto release = [&]() {
    for (ShardedPerm &p : used) free_perm(p);
    used.clear();
  };
  try {
    for (int i = 0; i < maxbits; i++) {
      first[i] = all.size();
      uint32_t sb = 0;
      uint64_t st = 0;
      for (BatchReq &r : rq) {
        if (i >= r.bits) continue;
        orqg_ctx *c = r.c;
        const uint64_t n = r.v->n;
        SegBit S;
        std::memset(&S, 0, sizeof(S));
        S.rk_pair = c->rk(kPair0);
        for (int g = 0; g < 3; g++) S.rk_grp[g] = c->rk(kGroup0 + g);
        S.cur = r.cur;
        S.other = r.other;
        S.F = r.F;
        S.bsum = r.bsum;
        S.tot = r.tot;
        S.err = r.err;
        S.flip = (r.full_width && i == r.bits - 1) ? 1u : 0u;  // inplace_invert on the MSB
        for (int k = 0; k < 3; k++) S.nb.v[k] = c->nonce[kPair0 + k];
        S.s_b2a = make_sched(n, c->batch_size, 4, 3);
        uint64_t u = sched_nonces(S.s_b2a);
        for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += u;
        for (int p = 0; p < 3; p++) c->log_msg("b2a_bit", p, (p + 2) % 3, 3 * n * 4);
        for (int k = 0; k < 3; k++) S.nb2.v[k] = c->nonce[kPair0 + k];
        S.s_perm = make_sched(n, c->batch_size, 4, 1);
        u = sched_nonces(S.s_perm);
        for (int k = 0; k < 3; k++) c->nonce[kPair0 + k] += u;
        for (int p = 0; p < 3; p++) c->log_msg("mul_a", p, (p + 2) % 3, n * 4);
        // oblivious_apply_elementwise_perm(v, perm): pi from the queue
        if (c->queue.empty() || c->stored_size != n)
          throw Error(ORQG_E_PARITY, "radix_sort_batch: permutation queue out of step");
        ShardedPerm pi = c->queue.front();
        c->queue.pop_front();
        if (pi.ready) ORQG_HIP(hipStreamWaitEvent(lead->stream, pi.ready, 0));
        used.push_back(pi);
        for (int g = 0; g < 3; g++) S.idx[g] = pi.inv[g];
        S.sv = make_sched(n, c->batch_size, tszv, 2);
        S.sp = make_sched(n, c->batch_size, 4, 2);
        const uint64_t uv = sched_nonces(S.sv), up = sched_nonces(S.sp);
        for (int g = 0; g < 3; g++) {
          S.nv[g] = c->nonce[kGroup0 + g];
          S.np[g] = S.nv[g] + uv;
          c->nonce[kGroup0 + g] += uv + up;
        }
        for (int g = 0; g < 3; g++) log_reshare(c, g, n * tszv);
        for (int g = 0; g < 3; g++) log_reshare(c, g, n * 4);
        for (int p = 0; p < 3; p++) c->log_msg("open", p, (p + 2) % 3, n * 4);
        S.nsb = (uint32_t)((n + kEPB - 1) / kEPB);
        S.sb0 = sb;
        S.st0 = st;
        sb += S.nsb;
        st += (n + 127) / 128;
        all.push_back(S);
        std::swap(r.cur, r.other);
      }
      tot_sb[i] = sb;
      tot_st[i] = st;
    }
    first[maxbits] = all.size();
    if (first[maxbits] == 0) return;
    SegBit *dseg = (SegBit *)lead->scratch_get("rsb_seg", all.size() * sizeof(SegBit) + 64);
    ORQG_HIP(hipMemcpyAsync(dseg, all.data(), all.size() * sizeof(SegBit), hipMemcpyHostToDevice,
                            lead->stream));
    std::vector<TuneTrial> trials;
    auto launch_timed = [&](int kind, uint64_t units, auto &&fn) {
      bool timed = false;
      const unsigned grid = tune_pick(lead, kind, L, units, trials, timed);
      TuneTrial t{tune_key(kind, L, units), grid, nullptr, nullptr};
      if (timed) {
        ORQG_HIP(hipEventCreate(&t.a));
        ORQG_HIP(hipEventCreate(&t.b));
        ORQG_HIP(hipEventRecord(t.a, lead->stream));
      }
      fn(grid);
      if (timed) {
        ORQG_HIP(hipEventRecord(t.b, lead->stream));
        trials.push_back(t);
      }
    };
    for (int i = 0; i < maxbits; i++) {
      const SegBit *d = dseg + first[i];
      const uint32_t nseg = (uint32_t)(first[i + 1] - first[i]);
      const int bitpos = 32 + i;
      const uint32_t tsb = tot_sb[i];
      const uint64_t tst = tot_st[i];
      {
        // b2a_bit and the bit permutation in one pass (k_bit_fused), the
        // prefix's superblock offsets and totals added by the first round
        TimedRegion tr(lead, "sortb_b2a", 2);
        launch_timed(kTuneB2a, tsb, [&](unsigned grid) {
          hipLaunchKernelGGL(k_bit_fused_seg<L>, dim3(grid), dim3(kBT), 0, lead->stream,
                             lead->d_t0, d, nseg, tsb, bitpos / 32, bitpos % 32);
        });
        hipLaunchKernelGGL(k_scan3_corr_seg, dim3(nseg), dim3(1024), 0, lead->stream, d);
      }
      {
        TimedRegion tr(lead, "sortb_round", 3);
        // a workgroup takes 16 wave steps at a time
        launch_timed(kTuneRound, (tst + 15) / 16, [&](unsigned grid) {
          hipLaunchKernelGGL((k_round2_seg<L, false, true>), dim3(grid), dim3(kBT), 0,
                             lead->stream, lead->d_t0, d, nseg, tst, 0);
          hipLaunchKernelGGL((k_round2_seg<L, false>), dim3(grid), dim3(kBT), 0, lead->stream,
                             lead->d_t0, d, nseg, tst, 1);
          hipLaunchKernelGGL((k_round2_seg<L, true>), dim3(grid), dim3(kBT), 0, lead->stream,
                             lead->d_t0, d, nseg, tst, 2);
        });
      }
      ORQG_HIP(hipGetLastError());
    }
    // the requests' opened-permutation checks, then each request's tail on
    // its own stream
    std::vector<uint32_t> herr(rq.size(), 0);
    for (size_t q = 0; q < rq.size(); q++)
      ORQG_HIP(hipMemcpyAsync(&herr[q], rq[q].err, 4, hipMemcpyDeviceToHost, lead->stream));
    ORQG_HIP(hipStreamSynchronize(lead->stream));
    tune_settle(trials);
    release();
    for (size_t q = 0; q < rq.size(); q++)
      if (herr[q])
        throw Error(ORQG_E_PARITY, "radix_sort_batch: request " + std::to_string(q) +
                                       ": opened bit permutation out of range");
  } catch (...) {
    (void)hipStreamSynchronize(lead->stream);
    release();
    throw;
  }
}

}  // namespace

void radix_sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out) {
  const uint64_t n = v->n;
  const bool desc = order == ORQG_DESC;
  const bool full_width = bits == 8 * v->width;
  if (n == 0) return;
  if (c->batcher && n <= c->batcher->max_rows) {
    c->batcher->sort(c, v, order, bits, perm_out);
    return;
  }
  // 1 permutation + (bits + 1) pairs; pairs are single permutations in 3PC
  // (radixsort.h:121-128, permutation_manager.h:108-111)
  reserve_perms(c, n, (size_t)bits + 2);
  uint32_t *P = (uint32_t *)c->scratch_get("rs_P", n * kRec * 4 + 64);
  pad_records(c, c->stream, v, desc, P);
  if (v->width == 8)
    P = sort_body<3>(c, P, n, bits, full_width);
  else
    P = sort_body<2>(c, P, n, bits, full_width);
  sort_finish(c, v, P, desc, perm_out);
  ORQG_HIP(hipStreamSynchronize(c->stream));
}

// radix_sort of m requests, each on its own context, with the bit iterations
// of all of them in shared launches: every request gets the shares, nonces,
// permutation queue and message log that radix_sort on its context gives.
void radix_sort_batch(const std::vector<orqg_ctx *> &cs, const std::vector<orqg_vec *> &vs,
                      const std::vector<int> &orders, const std::vector<int> &bits,
                      const std::vector<orqg_vec *> &perms) {
  std::vector<size_t> live;
  for (size_t q = 0; q < cs.size(); q++)
    if (vs[q]->n > 0) live.push_back(q);
  if (live.empty()) return;
  {
    std::vector<orqg_ctx *> rc;
    std::vector<size_t> rn, rk;
    for (size_t q : live) {
      rc.push_back(cs[q]);
      rn.push_back(vs[q]->n);
      rk.push_back((size_t)bits[q] + 2);
    }
    reserve_perms_multi(rc, rn, rk);
  }
  orqg_ctx *lead = cs[live[0]];
  for (int L : {3, 2}) {
    std::vector<BatchReq> rq;
    for (size_t q : live) {
      if ((vs[q]->width == 8 ? 3 : 2) != L) continue;
      BatchReq r;
      orqg_ctx *c = cs[q];
      const uint64_t n = vs[q]->n;
      r.c = c;
      r.v = vs[q];
      r.perm = perms[q];
      r.desc = orders[q] == ORQG_DESC;
      r.bits = bits[q];
      r.full_width = bits[q] == 8 * vs[q]->width;
      r.P = (uint32_t *)c->scratch_get("rs_P", n * kRec * 4 + 64);
      r.R1 = (uint32_t *)c->scratch_get("rs_R1", n * kRec * 4 + 64);
      r.F = (uint32_t *)c->scratch_get("rs_F", n * 12 + 64);
      r.bsum = (uint32_t *)c->scratch_get("rs_bsum", (n + kEPB - 1) / kEPB * 12 + 64);
      r.tot = (uint32_t *)c->scratch_get("rs_tot", 64);
      r.err = (uint32_t *)c->scratch_get("rs_err", 16);
      r.cur = r.P;
      r.other = r.R1;
      rq.push_back(r);
    }
    if (rq.empty()) continue;
    // every request's earlier work (its reserve, the producers of its
    // vector) precedes the shared launches
    for (BatchReq &r : rq)
      if (r.c != lead) {
        ORQG_HIP(hipEventRecord(r.c->ev_aux, r.c->stream));
        ORQG_HIP(hipStreamWaitEvent(lead->stream, r.c->ev_aux, 0));
      }
    for (BatchReq &r : rq) {
      ORQG_HIP(hipMemsetAsync(r.err, 0, 4, lead->stream));
      pad_records(r.c, lead->stream, r.v, r.desc, r.P);
    }
    if (L == 3)
      batch_body<3>(lead, rq);
    else
      batch_body<2>(lead, rq);
    // (batch_body synchronised lead's stream) each request's tail on its own
    // stream, then its swap-resolution checks, as radix_sort ends.  The
    // tails are many small launches with host waits between them; the
    // requests' contexts are independent, so a few host threads issue them
    // side by side.
    const size_t nthr = std::min<size_t>(rq.size(), 8);
    std::vector<std::exception_ptr> errs(rq.size());
    auto tail = [&](size_t first) {
      for (size_t q = first; q < rq.size(); q += nthr) {
        try {
          ORQG_HIP(hipSetDevice(rq[q].c->device));
          sort_finish(rq[q].c, rq[q].v, rq[q].cur, rq[q].desc, rq[q].perm);
          ORQG_HIP(hipStreamSynchronize(rq[q].c->stream));
          check_swaps(rq[q].c);
        } catch (...) {
          errs[q] = std::current_exception();
        }
      }
    };
    {
      std::vector<std::thread> pool;
      for (size_t k = 1; k < nthr; k++) pool.emplace_back(tail, k);
      tail(0);
      for (std::thread &t : pool) t.join();
    }
    for (std::exception_ptr &e : errs)
      if (e) std::rethrow_exception(e);
  }
}

void SortBatcher::enter() {
  std::lock_guard<std::mutex> lk(mu);
  active++;
}

void SortBatcher::leave() {
  std::lock_guard<std::mutex> lk(mu);
  active--;
  cv.notify_all();  // a gathering leader may now have everyone
}

void SortBatcher::sort(orqg_ctx *c, orqg_vec *v, int order, int bits, orqg_vec *perm_out) {
  Job j;
  j.c = c;
  j.v = v;
  j.perm = perm_out;
  j.order = order;
  j.bits = bits;
  std::unique_lock<std::mutex> lk(mu);
  waiting.push_back(&j);
  cv.notify_all();
  while (!j.done) {
    if (leading) {
      cv.wait(lk, [&] { return j.done || !leading; });
      continue;
    }
    // lead: gather, then run the batch outside the lock (the other jobs'
    // threads wait, so their contexts are not in use meanwhile)
    leading = true;
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(window_us);
    cv.wait_until(lk, until, [&] { return (int)waiting.size() >= active; });
    std::vector<Job *> take;
    take.swap(waiting);
    lk.unlock();
    std::exception_ptr err;
    try {
      std::vector<orqg_ctx *> cs;
      std::vector<orqg_vec *> vs, ps;
      std::vector<int> os, bs;
      for (Job *t : take) {
        cs.push_back(t->c);
        vs.push_back(t->v);
        ps.push_back(t->perm);
        os.push_back(t->order);
        bs.push_back(t->bits);
      }
      radix_sort_batch(cs, vs, os, bs, ps);
    } catch (...) {
      err = std::current_exception();
    }
    lk.lock();
    for (Job *t : take) {
      t->err = err;
      t->done = true;
    }
    batches++;
    batched += take.size();
    leading = false;
    cv.notify_all();
  }
  lk.unlock();
  if (j.err) std::rethrow_exception(j.err);
}

}  // namespace orqg
