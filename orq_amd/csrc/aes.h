// AES-256 encryption of counter blocks for the reference's PRG keystream.
//
// Keystream framing (include/core/random/prg/prg_algorithm.h:214-231): a PRG
// call of B bytes is split into <=1 MiB chunks; chunk c is the AES-256-GCM
// encryption of zeros under IV = LE64(nonce + c) || 0^4, i.e. AES-256-CTR
// starting at counter block IV || BE32(2).
//
// State words are little-endian columns: w_c = s_{0,c} | s_{1,c}<<8 |
// s_{2,c}<<16 | s_{3,c}<<24.  One full round is
//   t_c = T0[b0(w_c)] ^ T1[b1(w_{c+1})] ^ T2[b2(w_{c+2})] ^ T3[b3(w_{c+3})] ^ rk
// with T0[x] = (2S, S, S, 3S) and T1..T3 its byte rotations (FIPS-197 5.1,
// table form).  The tables live in LDS, replicated 32 times so that lane l
// reads copy (l % 32): every lane of a half-wave hits its own bank and table
// lookups are conflict-free.  The 128 KiB that takes leaves room for one
// workgroup per CU, so AES kernels run 1024-thread persistent workgroups.
//
// Entry x of table j, copy c sits at LDS byte
//   (j >> 1) * 65536 + x * 256 + (j & 1) * 128 + c * 4,
// so one v_perm_b32 builds the whole lookup address from the state word (byte
// x lands in address byte 1) and a per-lane base (c * 4 in byte 0, the region
// bit in byte 2), and (j & 1) * 128 rides in the ds_read offset: one VALU op
// per lookup instead of an extract plus an add (measured on MI355X: 1.10 TB/s
// of keystream, against 0.89 TB/s with a shift-and-add address).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orqg {

// ---- host-side tables and key schedule ------------------------------------
struct AesTables {
  uint8_t sbox[256];
  uint32_t t0[256];
};

inline uint8_t aes_xtime(uint8_t x) { return (uint8_t)((x << 1) ^ ((x & 0x80) ? 0x1b : 0)); }

// S-box from the multiplicative inverse in GF(2^8) and the affine map,
// walking the multiplicative group with generator 3.
inline void aes_make_tables(AesTables *t) {
  uint8_t p = 1, q = 1;
  do {
    p = (uint8_t)(p ^ aes_xtime(p));  // p *= 3
    q ^= (uint8_t)(q << 1);           // q /= 3
    q ^= (uint8_t)(q << 2);
    q ^= (uint8_t)(q << 4);
    if (q & 0x80) q ^= 0x09;
    uint8_t x = (uint8_t)(q ^ ((q << 1) | (q >> 7)) ^ ((q << 2) | (q >> 6)) ^
                          ((q << 3) | (q >> 5)) ^ ((q << 4) | (q >> 4)));
    t->sbox[p] = (uint8_t)(x ^ 0x63);
  } while (p != 1);
  t->sbox[0] = 0x63;
  for (int i = 0; i < 256; i++) {
    uint8_t s = t->sbox[i], s2 = aes_xtime(s), s3 = (uint8_t)(s2 ^ s);
    t->t0[i] = (uint32_t)s2 | ((uint32_t)s << 8) | ((uint32_t)s << 16) | ((uint32_t)s3 << 24);
  }
}

// AES-256 key expansion (FIPS-197 5.2) into 60 little-endian words.
inline void aes256_expand(const AesTables &t, const uint8_t key[32], uint32_t rk[60]) {
  for (int i = 0; i < 8; i++)
    rk[i] = (uint32_t)key[4 * i] | ((uint32_t)key[4 * i + 1] << 8) |
            ((uint32_t)key[4 * i + 2] << 16) | ((uint32_t)key[4 * i + 3] << 24);
  uint8_t rcon = 1;
  auto sub = [&](uint32_t w) {
    return (uint32_t)t.sbox[w & 0xff] | ((uint32_t)t.sbox[(w >> 8) & 0xff] << 8) |
           ((uint32_t)t.sbox[(w >> 16) & 0xff] << 16) | ((uint32_t)t.sbox[w >> 24] << 24);
  };
  for (int i = 8; i < 60; i++) {
    uint32_t tmp = rk[i - 1];
    if (i % 8 == 0) {
      tmp = sub((tmp >> 8) | (tmp << 24)) ^ rcon;
      rcon = aes_xtime(rcon);
    } else if (i % 8 == 4) {
      tmp = sub(tmp);
    }
    rk[i] = rk[i - 8] ^ tmp;
  }
}

// ---- device side ------------------------------------------------------------
// LDS table layout (words): entry (table j, index x, copy c) at
// (j >> 1) * 16384 + x * 64 + (j & 1) * 32 + c.  Table j = rotl(T0, 8 j).
template <int COPIES>
struct AesLds {
  static_assert(COPIES == 32, "the v_perm addressing assumes 32 table copies");
  static constexpr int kWords = 4 * 256 * COPIES;
  static constexpr int kBytes = kWords * 4;
};

// Build the replicated tables in LDS (word layout above) from T0 in global
// memory (1 KiB, so a launch's workgroups read 1 KiB each, not the 128 KiB
// image).  Consecutive lanes write consecutive 16-byte chunks; chunk w holds
// copies 4 (w & 7) .. + 3 of entry x = (w >> 4) & 255 of table
// j = 2 (w >> 12) + ((w >> 3) & 1).  lds must be 16-byte aligned.
template <int COPIES>
__device__ inline void aes_load_tables(uint32_t *lds, const uint32_t *__restrict__ g_t0) {
  uint4 *dst = reinterpret_cast<uint4 *>(lds);
  for (int w = threadIdx.x; w < AesLds<COPIES>::kWords / 4; w += blockDim.x) {
    const uint32_t j = ((uint32_t)(w >> 12) << 1) | ((uint32_t)(w >> 3) & 1u);
    const uint32_t t = g_t0[(w >> 4) & 255];
    const uint32_t v = j ? (t << (8 * j)) | (t >> (32 - 8 * j)) : t;
    dst[w] = make_uint4(v, v, v, v);
  }
  __syncthreads();
}

// three-input XOR in one VALU op (gfx950 v_bitop3_b32, truth table 0x96)
__device__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

template <int COPIES>
struct AesCtx {
  const char *t;  // LDS tables
  uint32_t base;  // this lane's copy offset (byte 0) and the region bit (byte 2)
  __device__ explicit AesCtx(const uint32_t *lds)
      : t(reinterpret_cast<const char *>(lds)), base(((threadIdx.x % COPIES) << 2) | 0x10000u) {}
  // table J at byte K of state word s.  v_perm_b32 selectors: 0-3 pick bytes of
  // the second operand (base), 4-7 bytes of the first (s), 12 gives 0x00.
  template <int J, int K>
  __device__ inline uint32_t T(uint32_t s) const {
    constexpr uint32_t sel = 0x0c000000u | ((J >> 1) ? 0x020000u : 0x0c0000u) | ((4u + K) << 8);
    const uint32_t a = __builtin_amdgcn_perm(s, base, sel);
    return *reinterpret_cast<const uint32_t *>(t + a + (J & 1) * 128);
  }
};

// Encrypt one block (4 LE words) under round keys rk[60] (wave-uniform).
template <int COPIES>
__device__ inline void aes256_encrypt(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                      uint32_t &w0, uint32_t &w1, uint32_t &w2, uint32_t &w3) {
  uint32_t s0 = w0 ^ rk[0], s1 = w1 ^ rk[1], s2 = w2 ^ rk[2], s3 = w3 ^ rk[3];
#pragma unroll
  for (int r = 1; r < 14; r++) {
    uint32_t t0 = xor3(xor3(A.template T<0, 0>(s0), A.template T<1, 1>(s1), A.template T<2, 2>(s2)),
                       A.template T<3, 3>(s3), rk[4 * r + 0]);
    uint32_t t1 = xor3(xor3(A.template T<0, 0>(s1), A.template T<1, 1>(s2), A.template T<2, 2>(s3)),
                       A.template T<3, 3>(s0), rk[4 * r + 1]);
    uint32_t t2 = xor3(xor3(A.template T<0, 0>(s2), A.template T<1, 1>(s3), A.template T<2, 2>(s0)),
                       A.template T<3, 3>(s1), rk[4 * r + 2]);
    uint32_t t3 = xor3(xor3(A.template T<0, 0>(s3), A.template T<1, 1>(s0), A.template T<2, 2>(s1)),
                       A.template T<3, 3>(s2), rk[4 * r + 3]);
    s0 = t0;
    s1 = t1;
    s2 = t2;
    s3 = t3;
  }
  // last round: SubBytes + ShiftRows + AddRoundKey; S[x] is byte 0 of T2,
  // byte 1 of T3, byte 2 of T0 and byte 3 of T1, merged with two byte permutes
  auto last = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t k) {
    uint32_t lo = __builtin_amdgcn_perm(A.template T<3, 1>(b), A.template T<2, 0>(a), 0x0c0c0500u);
    uint32_t hi = __builtin_amdgcn_perm(A.template T<1, 3>(d), A.template T<0, 2>(c), 0x07020c0cu);
    return xor3(lo, hi, k);
  };
  const uint32_t o0 = last(s0, s1, s2, s3, rk[56]), o1 = last(s1, s2, s3, s0, rk[57]),
                 o2 = last(s2, s3, s0, s1, rk[58]), o3 = last(s3, s0, s1, s2, rk[59]);
  w0 = o0;
  w1 = o1;
  w2 = o2;
  w3 = o3;
}

// Keystream block `ctr` of nonce `nonce`: encrypt LE64(nonce) || 0^4 || BE32(ctr).
template <int COPIES>
__device__ inline uint4 aes_ctr_block(const AesCtx<COPIES> &A, const uint32_t *__restrict__ rk,
                                      uint64_t nonce, uint32_t ctr) {
  uint32_t w0 = (uint32_t)nonce, w1 = (uint32_t)(nonce >> 32), w2 = 0, w3 = __builtin_bswap32(ctr);
  aes256_encrypt<COPIES>(A, rk, w0, w1, w2, w3);
  return make_uint4(w0, w1, w2, w3);
}

}  // namespace orqg
