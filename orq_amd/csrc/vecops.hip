// Share-local Vector operations: every share of every party is transformed
// on its own, with no message and no randomness, because each of these maps
// is linear over the sharing (bits move or copy within an element for XOR
// shares; sums of shares for additive ones):
//   bit_left_shift / bit_arithmetic_right_shift / bit_logical_right_shift
//     (include/core/containers/mapping_access_vector.h:376-421), as
//     BSharedVector operator<< / >> (b_shared_vector.h:100-127,384-414)
//   bit_xor, the parity of each element (mapping_access_vector.h:430-439;
//     b_shared_vector.h:135-137)
//   extend_lsb (mapping_access_vector.h:1541-1547; b_shared_vector.h:145-147)
//   ltz, the sign bit (mapping_access_vector.h:1528-1535; protocol.h:682)
//   mask, AND with a public constant (mapping_access_vector.h:1312-1316;
//     b_shared_vector.h:462-464)
//   prefix_sum, an inclusive scan of each share (SharedVector::prefix_sum,
//     shared_vector.h:249 -> mapping_access_vector.h:446)
//
// Layout: the [n][3] share array is 3n contiguous elements; the unary maps
// run over them flat.  The scan runs per share column over the rows, in
// three passes over tiles of kScanRows rows: tile sums, an exclusive scan of
// the tile sums in one workgroup, then each tile rescanned with its offset.
#include <type_traits>

#include "internal.h"

namespace orqg {
namespace {

template <typename T, int OP>
__global__ void k_unary(uint64_t m, const T *x, T *z, uint64_t arg) {  // z may be x
  using S = typename std::make_signed<T>::type;
  constexpr int B = 8 * sizeof(T);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const T a = x[i];
    T r;
    if (OP == ORQG_SHL)
      r = arg >= (uint64_t)B ? (T)0 : (T)(a << arg);
    else if (OP == ORQG_ASR)
      r = (T)((S)a >> (arg >= (uint64_t)B ? B - 1 : (int)arg));
    else if (OP == ORQG_LSR)
      r = arg >= (uint64_t)B ? (T)0 : (T)(a >> arg);
    else if (OP == ORQG_PARITY)
      r = (T)(__popcll((unsigned long long)a) & 1);
    else if (OP == ORQG_EXTEND_LSB)
      r = (T)((T)0 - (a & 1));
    else if (OP == ORQG_LTZ)
      r = (T)(a >> (B - 1));
    else
      r = a & (T)arg;  // ORQG_MASK
    z[i] = r;
  }
}

constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;                                // rows per thread
constexpr uint64_t kScanRows = kScanThreads * kScanPer;     // rows per tile

// exclusive scan of one value per thread across the workgroup (kScanThreads)
template <typename T>
__device__ inline T block_excl_scan(T v, T *s_w) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  T inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  T base = 0;
  for (int w = 0; w < wv; w++) base += s_w[w];
  __syncthreads();
  return base + inc - v;
}

// pass 1: per tile and share, the sum of its rows
template <typename T>
__global__ void __launch_bounds__(kScanThreads)
    k_ps_tile_sums(uint64_t n, const T *__restrict__ x, T *__restrict__ tsum) {
  __shared__ T s_w[kScanThreads / 64];
  const uint64_t r0 = blockIdx.x * kScanRows + (uint64_t)threadIdx.x * kScanPer;
  T acc[3] = {0, 0, 0};
  for (int j = 0; j < kScanPer; j++) {
    const uint64_t r = r0 + j;
    if (r < n)
      for (int k = 0; k < 3; k++) acc[k] += x[3 * r + k];
  }
  for (int k = 0; k < 3; k++) {
    const T before = block_excl_scan<T>(acc[k], s_w);
    if (threadIdx.x == kScanThreads - 1) tsum[3 * blockIdx.x + k] = before + acc[k];
  }
}

// pass 2: exclusive scan of the tile sums, one workgroup, in place
template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_ps_scan_tiles(uint64_t ntiles, T *tsum) {
  __shared__ T s_w[kScanThreads / 64];
  __shared__ T s_tot;
  T carry[3] = {0, 0, 0};
  for (uint64_t base = 0; base < ntiles; base += kScanThreads) {
    const uint64_t t = base + threadIdx.x;
    for (int k = 0; k < 3; k++) {
      const T v = t < ntiles ? tsum[3 * t + k] : (T)0;
      const T before = block_excl_scan<T>(v, s_w);
      if (t < ntiles) tsum[3 * t + k] = carry[k] + before;
      // the chunk's total, from the last thread's inclusive value
      if (threadIdx.x == kScanThreads - 1) s_tot = before + v;
      __syncthreads();
      carry[k] += s_tot;
      __syncthreads();
    }
  }
}

// pass 3: each tile rescanned with its offset, written to z (may alias x)
template <typename T>
__global__ void __launch_bounds__(kScanThreads)
    k_ps_apply(uint64_t n, const T *x, T *z, const T *__restrict__ tsum) {
  __shared__ T s_w[kScanThreads / 64];
  const uint64_t r0 = blockIdx.x * kScanRows + (uint64_t)threadIdx.x * kScanPer;
  T v[kScanPer][3];
  T acc[3] = {0, 0, 0};
  for (int j = 0; j < kScanPer; j++) {
    const uint64_t r = r0 + j;
    for (int k = 0; k < 3; k++) {
      v[j][k] = r < n ? x[3 * r + k] : (T)0;
      acc[k] += v[j][k];
    }
  }
  T off[3];
  for (int k = 0; k < 3; k++) off[k] = tsum[3 * blockIdx.x + k] + block_excl_scan<T>(acc[k], s_w);
  for (int j = 0; j < kScanPer; j++) {
    const uint64_t r = r0 + j;
    for (int k = 0; k < 3; k++) {
      off[k] += v[j][k];
      if (r < n) z[3 * r + k] = off[k];
    }
  }
}

template <typename T>
void unary_t(orqg_ctx *c, int op, const orqg_vec *x, uint64_t arg, orqg_vec *z) {
  const uint64_t m = 3 * x->n;
  if (m == 0) return;
  const unsigned grid = grid_for(m, 256, c->num_cu * 16);
  const T *xp = (const T *)x->d;
  T *zp = (T *)z->d;
#define ORQG_UNARY(OPC)                                                                       \
  case OPC:                                                                                   \
    hipLaunchKernelGGL((k_unary<T, OPC>), dim3(grid), dim3(256), 0, c->stream, m, xp, zp, arg); \
    break;
  switch (op) {
    ORQG_UNARY(ORQG_SHL)
    ORQG_UNARY(ORQG_ASR)
    ORQG_UNARY(ORQG_LSR)
    ORQG_UNARY(ORQG_PARITY)
    ORQG_UNARY(ORQG_EXTEND_LSB)
    ORQG_UNARY(ORQG_LTZ)
    ORQG_UNARY(ORQG_MASK)
    default:
      throw Error(ORQG_E_ARG, "vec_unary: unknown op");
  }
#undef ORQG_UNARY
  ORQG_HIP(hipGetLastError());
}

template <typename T>
void prefix_sum_t(orqg_ctx *c, orqg_vec *v) {
  const uint64_t n = v->n;
  if (n == 0) return;
  const uint64_t ntiles = (n + kScanRows - 1) / kScanRows;
  T *tsum = (T *)c->scratch_get("prefix_tiles", 3 * ntiles * sizeof(T) + 64);
  T *d = (T *)v->d;
  TimedRegion tr(c, "prefix_sum");
  hipLaunchKernelGGL(k_ps_tile_sums<T>, dim3((unsigned)ntiles), dim3(kScanThreads), 0, c->stream,
                     n, (const T *)d, tsum);
  hipLaunchKernelGGL(k_ps_scan_tiles<T>, dim3(1), dim3(kScanThreads), 0, c->stream, ntiles, tsum);
  hipLaunchKernelGGL(k_ps_apply<T>, dim3((unsigned)ntiles), dim3(kScanThreads), 0, c->stream, n,
                     (const T *)d, d, (const T *)tsum);
  ORQG_HIP(hipGetLastError());
}

}  // namespace

void vec_unary(orqg_ctx *c, int op, const orqg_vec *x, uint64_t arg, orqg_vec *z) {
  if (!x || !z || x->n != z->n || x->width != z->width)
    throw Error(ORQG_E_ARG, "vec_unary: vectors differ in size or width");
  if (op < ORQG_SHL || op > ORQG_MASK) throw Error(ORQG_E_ARG, "vec_unary: unknown op");
  if (x->width == 4)
    unary_t<uint32_t>(c, op, x, arg, z);
  else
    unary_t<uint64_t>(c, op, x, arg, z);
  z->enc = x->enc;
}

void prefix_sum(orqg_ctx *c, orqg_vec *v) {
  if (!v) throw Error(ORQG_E_ARG, "prefix_sum: NULL vector");
  if (v->n > (uint64_t)UINT32_MAX * kScanRows) throw Error(ORQG_E_ARG, "prefix_sum: too many rows");
  if (v->width == 4)
    prefix_sum_t<uint32_t>(c, v);
  else
    prefix_sum_t<uint64_t>(c, v);
}

}  // namespace orqg
