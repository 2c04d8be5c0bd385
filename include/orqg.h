/*
 * orqg.h -- C ABI of the MI355X execution backend for ORQ's 3-party
 * replicated (honest-majority) oblivious operators.
 *
 * All three parties live on one GPU.  A shared vector stores the three
 * distinct replicated shares s_0, s_1, s_2 once, element-major
 * ([n][3] elements); party p's view is (s_p, s_{p+1}).  Every message the
 * reference would send between parties becomes a device-buffer handoff whose
 * byte count is appended to the context's message log.
 *
 * ORQ itself is a header-only C++ template library with no FFI; the entry
 * points below replace the per-party protocol primitives and operators cited
 * beside each declaration (paths relative to the ORQ source tree).
 *
 * Conventions: every function returns ORQG_OK (0) or a negative ORQG_E_*
 * code; orqg_last_error() gives the message.  Calls are stream-ordered on the
 * context's HIP stream; functions that return host data synchronize.
 * A context is single-owner (one host thread at a time), like a reference
 * worker that owns its Protocol, PRGs and Communicator.
 */
#ifndef ORQG_H_
#define ORQG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORQG_OK = 0,
  ORQG_E_ARG = -1,    /* bad argument, size or width */
  ORQG_E_OOM = -2,    /* device allocation failed */
  ORQG_E_BUDGET = -3, /* request exceeds the configured HBM budget */
  ORQG_E_DEVICE = -4, /* HIP runtime error */
  ORQG_E_PARITY = -5, /* a device-side consistency check failed */
};

enum { ORQG_BSHARED = 1, ORQG_ASHARED = 0 }; /* include/core/containers/encoding.h:8 */
enum { ORQG_ASC = 0, ORQG_DESC = 1 };
/* BSharedVector comparison operators (include/core/containers/b_shared_vector.h:493-580) */
enum { ORQG_EQ = 0, ORQG_NE = 1, ORQG_LT = 2, ORQG_GT = 3, ORQG_LE = 4, ORQG_GE = 5 };

/* PRG roles of the 10 AES-256 keys (include/backend/common/rand_setup.h:21-146):
 *   0..2  localPRG of party p (secret sharing)
 *   3..5  pairwise key K_p: "next" PRG of party p and "prev" PRG of party p+1
 *   6..9  group keys {0,1}, {0,2}, {1,2}, {0,1,2}                           */
#define ORQG_NUM_KEYS 10

typedef struct {
  int device;                        /* HIP device ordinal */
  uint8_t keys[ORQG_NUM_KEYS][32];   /* see roles above (worker thread 0) */
  long batch_size;                   /* runtime batch size; <0 = |b| equal batches (ref default -12);
                                        a positive one must be a multiple of 256 (the kernels
                                        keep a batch's elements in whole 16-byte keystream
                                        blocks and whole packed words) */
  size_t hbm_budget;                 /* bytes; 0 = no limit */
  /* worker threads of the reference's RunTime (0 or 1: one).  With T > 1, a
   * call over n elements is split into the threads' ranges
   * (runtime.h:198-247), each batched on its own and drawing from its own
   * thread's keys, and reserved permutations are dealt to the threads in
   * runs (runtime.h:749-788).  Supported: secret sharing, open, and_b /
   * multiply_a, b2a_bit, reshare, sharded and elementwise permutations and
   * their ElementwisePermutation ops, radix and table sort (RADIXSORT), the
   * comparisons (==, <, ...), full b2a, the adders and a2b; the sorting
   * networks, quicksort, aggregation and distinct return ORQG_E_ARG.        */
  int threads;
  const uint8_t *thread_keys;        /* T > 1: (T-1) x ORQG_NUM_KEYS x 32 bytes, threads 1.. */
} orqg_config;

typedef struct orqg_ctx orqg_ctx;
typedef struct orqg_vec orqg_vec;

/* ---- context --------------------------------------------------------- */
/* replaces replicated_3pc::orq_init / orq_runtime_init
 * (include/backend/common/setup.h:7-87) with deterministic keys           */
int orqg_ctx_create(const orqg_config *cfg, orqg_ctx **out);
int orqg_ctx_destroy(orqg_ctx *ctx);
/* Give the context fresh keys: every PRG nonce back to 0, the permutation
 * queue emptied, the message log and timers cleared (a new session of the
 * three parties, as RandomnessManager setup does in rand_setup.h:21-146).   */
int orqg_ctx_rekey(orqg_ctx *ctx, const uint8_t keys[ORQG_NUM_KEYS][32]);
/* The protocol that SortingProtocol::DEFAULT stands for in this context's
 * table sorts (EncodedTable::sort, aggregate, distinct, inner_join): the
 * reference's compile-time DEFAULT_SORT_PROTO (operators/sorting.h:6-9),
 * QUICKSORT unless set (ORQG_QUICKSORT or ORQG_RADIXSORT; the ORQG_*SORT
 * values are below).                                                        */
int orqg_ctx_set_default_sort(orqg_ctx *ctx, int protocol);
int orqg_ctx_default_sort(orqg_ctx *ctx);
const char *orqg_last_error(orqg_ctx *ctx);
/* The hash of the sources the library was built from (SHA-256 prefix over
 * orq_amd/csrc, this header and the compile flags; orq_amd/build.py).  No
 * counterpart in the reference.                                              */
const char *orqg_build_info(void);
int orqg_sync(orqg_ctx *ctx);
/* the context's HIP stream (hipStream_t), for callers that order their own work */
void *orqg_stream(orqg_ctx *ctx);
/* current nonce of PRG `role` (0..9, see above): AESPRGAlgorithm::nonce
 * (include/core/random/prg/prg_algorithm.h:200-207)                         */
int orqg_prg_nonce(orqg_ctx *ctx, int role, uint64_t *nonce);
/* the nonce of key `role` of worker thread `thread` */
int orqg_prg_nonce_thread(orqg_ctx *ctx, int thread, int role, uint64_t *nonce);

/* ---- raw device buffers (plumbing for FFI callers and tests) ---------- */
int orqg_dev_alloc(orqg_ctx *ctx, size_t bytes, void **dev_out);
int orqg_dev_free(void *dev);
/* stream-ordered copy device -> host, then synchronize */
int orqg_dev_to_host(orqg_ctx *ctx, void *host, const void *dev, size_t bytes);

/* ---- shared vectors (EVector / A-/BSharedVector storage) -------------- */
/* width: 4 (int32) or 8 (int64) bytes per share element
 * (include/core/containers/e_vector.h:130-533)                              */
int orqg_vec_create(orqg_ctx *ctx, size_t n, int width, int encoding, orqg_vec **out);
int orqg_vec_free(orqg_vec *v);
size_t orqg_vec_size(const orqg_vec *v);
int orqg_vec_width(const orqg_vec *v);
/* device pointer to the [n][3] share array */
void *orqg_vec_data(orqg_vec *v);
/* host I/O of the three distinct shares, host layout [3][n] (s_0, s_1, s_2) */
int orqg_vec_write(orqg_ctx *ctx, orqg_vec *v, const void *host_shares);
int orqg_vec_read(orqg_ctx *ctx, const orqg_vec *v, void *host_shares);
int orqg_vec_copy(orqg_ctx *ctx, const orqg_vec *src, orqg_vec *dst);
/* EVector resize as EncodedTable::resize uses it (encoded_table.h:1426-1436):
 * keep the first min(n, size) rows, new rows are zero shares                */
int orqg_vec_resize(orqg_ctx *ctx, orqg_vec *v, size_t n);
/* keep the last n rows (EncodedTable::tail, encoded_table.h:1407-1418)       */
int orqg_vec_tail(orqg_ctx *ctx, orqg_vec *v, size_t n);
/* dst rows [dst_start, +count) = src rows [src_start, +count)
 * (EncodedTable::copy_column_typed, encoded_table.h:1749-1754)              */
int orqg_vec_copy_rows(orqg_ctx *ctx, const orqg_vec *src, size_t src_start, orqg_vec *dst,
                       size_t dst_start, size_t count);
/* dst rows [dst_start, +count) = row `row` of src, all three shares (the
 * repeated_subset_reference assignment of mapping_access_vector.h:854-866)  */
int orqg_vec_broadcast_row(orqg_ctx *ctx, const orqg_vec *src, size_t row, orqg_vec *dst,
                           size_t dst_start, size_t count);

/* ---- share-local Vector operations (no messages, no randomness) ------ */
/* z = op(x) on every share; z may be x.  Each map is linear over the
 * sharing, so it acts on the shared value:
 *   ORQG_SHL / ORQG_ASR / ORQG_LSR   x << arg, arithmetic and logical x >> arg
 *     (bit_left_shift / bit_arithmetic_right_shift / bit_logical_right_shift,
 *     include/core/containers/mapping_access_vector.h:376-421, as
 *     BSharedVector operator<< / >>, b_shared_vector.h:100-127,384-414)
 *   ORQG_PARITY      XOR of all bits (bit_xor, mapping_access_vector.h:430-439)
 *   ORQG_EXTEND_LSB  every bit = the LSB (extend_lsb, :1541-1547)
 *   ORQG_LTZ         the sign bit as 0/1 (ltz, :1528-1535; protocol.h:682)
 *   ORQG_MASK        x & arg (mask, :1312-1316; b_shared_vector.h:462-464)  */
enum { ORQG_SHL = 0, ORQG_ASR = 1, ORQG_LSR = 2, ORQG_PARITY = 3, ORQG_EXTEND_LSB = 4,
       ORQG_LTZ = 5, ORQG_MASK = 6 };
int orqg_vec_unary(orqg_ctx *ctx, int op, const orqg_vec *x, uint64_t arg, orqg_vec *z);
/* SharedVector::prefix_sum (include/core/containers/shared_vector.h:249):
 * in place inclusive scan of every share (std::inclusive_scan,
 * mapping_access_vector.h:446), so of the A-shared value                   */
int orqg_prefix_sum(orqg_ctx *ctx, orqg_vec *v);

/* ---- randomness ------------------------------------------------------ */
/* AESPRGAlgorithm::fillBytes of PRG `role` into device memory, advancing its
 * nonce (prg_algorithm.h:161-231)                                            */
int orqg_prg_keystream(orqg_ctx *ctx, int role, size_t bytes, void *dev_out);

/* ---- protocol primitives (all three parties per call) ---------------- */
/* Replicated_3PC::secret_share_b/_a from party 0 (replicated_3pc.h:401-448) */
int orqg_secret_share(orqg_ctx *ctx, const void *host_plain, orqg_vec *v);
/* open_shares_b/_a (replicated_3pc.h:310-330); host_out gets n elements    */
int orqg_open(orqg_ctx *ctx, const orqg_vec *v, void *host_out);
/* public_share (replicated_3pc.h:456-470)                                    */
int orqg_public_share(orqg_ctx *ctx, const void *host_plain, orqg_vec *v);
/* xor_b / add_a / sub_a (protocol.h:561-649)                                 */
int orqg_xor_b(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
int orqg_add_a(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
int orqg_sub_a(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* not_b (replicated_3pc.h:199)                                               */
int orqg_not_b(orqg_ctx *ctx, const orqg_vec *x, orqg_vec *z);
/* and_b (replicated_3pc.h:166-191)                                           */
int orqg_and_b(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* multiply_a (replicated_3pc.h:44-67)                                        */
int orqg_mul_a(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* b2a_bit (replicated_3pc.h:220-251): y = A-sharing of the LSB of x          */
int orqg_b2a_bit(orqg_ctx *ctx, const orqg_vec *x, orqg_vec *y);
/* BSharedVector::_compare (include/core/operators/circuits.h:231-295):
 * eq = (x == y) and gt = (x > y) as boolean shares of the LSB, elements taken
 * as signed two's complement.  Consumes the pair-PRG randomness of its
 * log2(bits) + 1 and_b calls over n elements and one over n/bits packed
 * words.  eq or gt may be NULL when not wanted.                              */
int orqg_compare(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *eq,
                 orqg_vec *gt);
/* z = (x op y) for op = ORQG_EQ..ORQG_GE (b_shared_vector.h:493-580):
 * == is bit_same & 1 (b_shared_vector.h:164-178), < is y > x, != / <= / >=
 * apply not_b_1 (replicated_3pc.h:207) to ==, > and <.                        */
int orqg_cmp(orqg_ctx *ctx, int op, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* BSharedVector::b2a (b_shared_vector.h:296-329): y = A-sharing of the whole
 * boolean-shared value, one b2a_bit and one multiply_a by 2^i per bit.        */
int orqg_b2a(orqg_ctx *ctx, const orqg_vec *x, orqg_vec *y);
/* ---- boolean adders and a2b (include/core/operators/circuits.h:35-216) -- */
/* BSharedVector::operator+ / operator- with the default adder, the ripple
 * carry adder (b_shared_vector.h:588-621, circuits.h:35-77): z = x + y,
 * z = x + ~y + 1                                                             */
int orqg_add_b(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
int orqg_sub_b(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* parallel_prefix_adder (circuits.h:138-216, Kogge-Stone): z = x + y, or
 * x + ~y + 1 when subtract != 0                                              */
int orqg_ppa(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, int subtract, orqg_vec *z);
/* rca_compare (circuits.h:93-136): z = sign bit of x - y (mod 2^bits), as
 * 0/1 boolean shares                                                         */
int orqg_rca_compare(orqg_ctx *ctx, const orqg_vec *x, const orqg_vec *y, orqg_vec *z);
/* ASharedVector::a2b (a_shared_vector.h:138-144): redistribute_shares_b
 * (replicated_3pc.h:259-262) then the ripple carry adder; y B-shared         */
int orqg_a2b(orqg_ctx *ctx, const orqg_vec *x, orqg_vec *y);
/* secret_share_b / _a with data owner `owner` (replicated_3pc.h:401-448)     */
int orqg_secret_share_from(orqg_ctx *ctx, const void *host_data, orqg_vec *v, int owner);

/* Protocol::reshare for group 0..2 = {0,1},{0,2},{1,2} (protocol.h:500-541) */
int orqg_reshare(orqg_ctx *ctx, orqg_vec *v, int group);

/* ---- permutations ---------------------------------------------------- */
/* PermutationManager::reserve (permutation_manager.h:102-181): generate
 * `count` honest-majority sharded permutations of size n with gen_perm
 * (hm_sharded_permutation_generator.h:70-108) and queue them.  Returns once
 * the draws are done; the last swap resolutions may still run on a second
 * stream.  Every consumer of a permutation waits for it on the device, and
 * a failed device-side check of the swap resolution is reported by the next
 * orqg_sync / host read / reserve (ORQG_E_PARITY).                          */
int orqg_reserve_perms(orqg_ctx *ctx, size_t n, size_t count);
size_t orqg_perm_queue_size(orqg_ctx *ctx);
/* local permutation of queued sharded permutation `index` for group 0..2
 * (host copy, for tests)                                                    */
int orqg_peek_perm(orqg_ctx *ctx, size_t index, int group, int32_t *host_out);
/* oblivious_apply_sharded_perm / _inverse_ (shuffle.h:349-412,598-660) with
 * the next queued permutation                                               */
int orqg_apply_sharded_perm(orqg_ctx *ctx, orqg_vec *x, int inverse);
/* oblivious_apply_elementwise_perm (shuffle.h:687-706); perm: width-4 vector */
int orqg_apply_elementwise_perm(orqg_ctx *ctx, orqg_vec *x, const orqg_vec *perm);


/* ---- ElementwisePermutation, shuffle, table sort ------------------------- */
/* ElementwisePermutation(size, encoding) (include/core/containers/permutation.h:62-69):
 * perm (width 4) = public sharing of the identity, encoding enc              */
int orqg_perm_identity(orqg_ctx *ctx, orqg_vec *perm, int enc);
/* ElementwisePermutation::invert (permutation.h:205-212)                     */
int orqg_perm_invert(orqg_ctx *ctx, orqg_vec *perm);
/* ElementwisePermutation::b2a for 3PC (permutation.h:222-258)                */
int orqg_perm_b2a(orqg_ctx *ctx, orqg_vec *perm);
/* ElementwisePermutation::negate (permutation.h:268-280): p = -p - 1         */
int orqg_perm_negate(orqg_ctx *ctx, orqg_vec *perm);
/* ElementwisePermutation::reverse (permutation.h:177-198): p = n - 1 - p     */
int orqg_perm_reverse(orqg_ctx *ctx, orqg_vec *perm);
/* compose_permutations(sigma, rho) (include/core/operators/shuffle.h:716-743):
 * rho <- the composition; sigma is left holding pi(sigma) as in the reference */
int orqg_perm_compose(orqg_ctx *ctx, orqg_vec *sigma, orqg_vec *rho);
/* operators::shuffle of one vector (shuffle.h:752-757)                       */
int orqg_shuffle(orqg_ctx *ctx, orqg_vec *x);
/* operators::shuffle of a table (shuffle.h:768-786), called by
 * EncodedTable::shuffle (encoded_table.h:1003-1018) with the A-shared and the
 * B-shared columns in schema order                                            */
int orqg_table_shuffle(orqg_ctx *ctx, orqg_vec *const *a_cols, size_t na,
                       orqg_vec *const *b_cols, size_t nb);
/* ---- aggregation network and distinct -------------------------------- */
/* aggregation functions of include/core/operators/aggregation.h:27-140:
 * sum, count, copy (A-shared); copy, max, min, bit_or, valid (B-shared)      */
enum { ORQG_AGG_SUM = 0, ORQG_AGG_COUNT = 1, ORQG_AGG_COPY = 2, ORQG_AGG_MAX = 3,
       ORQG_AGG_MIN = 4, ORQG_AGG_OR = 5, ORQG_AGG_VALID = 6 };
typedef struct {
  orqg_vec *in;  /* input column */
  orqg_vec *out; /* result column (may be `in`) */
  int func;      /* ORQG_AGG_* */
} orqg_agg_spec;
/* aggregators::aggregate (operators/aggregation.h:167-346): the odd-even
 * aggregation network over n = 2^k rows, grouped by equality of all `keys`.
 * reverse != 0 runs Direction::Reverse; sel (NULL: none) is the selection
 * column of joins (table id).  Forward leaves each group's aggregate in its
 * first row, reverse in its last.                                            */
int orqg_aggregate(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys,
                   const orqg_agg_spec *b_specs, size_t nb, const orqg_agg_spec *a_specs,
                   size_t na, int reverse, const orqg_vec *sel);
/* operators::distinct (operators/distinct.h:19-44): res[0] = 1 and
 * res[i] = (row i differs from row i-1 on some key), as boolean shares      */
int orqg_distinct(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys, orqg_vec *res);
/* orqg_distinct without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its one
 * scratch buffer (3 * n * width bytes of comparison results) is a fixed
 * 48 MiB static device array per GPU, held until the process ends, instead
 * of context memory, so n is at most 2^21 rows of width 8 (2^22 of width 4);
 * a larger input returns ORQG_E_ARG.  res is zeroed by a kernel, not a memset.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from different contexts or streams) race.  Results, nonces and message log
 * are those of orqg_distinct.                                               */
int orqg_distinct_noalloc(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys, orqg_vec *res);
/* public_share of one value in every row (replicated_3pc.h:456-470)         */
int orqg_public_fill(orqg_ctx *ctx, orqg_vec *v, uint64_t value, int enc);

/* operators::bitonic_sort (operators/sorting.h:266-353): keys[0] most
 * significant, orders[j] ORQG_ASC/ORQG_DESC; all columns of one width, the
 * row count a power of two; at most 8 columns of each kind.  Not stable.     */
int orqg_bitonic_sort(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys, const int *orders,
                      orqg_vec *const *a_cols, size_t na, orqg_vec *const *b_cols, size_t nb);
/* operators::table_sort with SortingProtocol::RADIXSORT (operators/sorting.h:509-612):
 * keys[0] is the most significant sort column; orders[j] ORQG_ASC/ORQG_DESC;
 * single_bit[j] != 0 sorts column j on its LSB only (NULL: none).  Every
 * key, A-shared and B-shared column ends up in sorted row order.            */
int orqg_table_sort(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys, const int *orders,
                    const int *single_bit, orqg_vec *const *a_cols, size_t na,
                    orqg_vec *const *b_cols, size_t nb);
/* SortingProtocol (sorting.h:15-22); QUICKSORT is the reference's DEFAULT   */
enum { ORQG_RADIXSORT = 0, ORQG_QUICKSORT = 1, ORQG_BITONICSORT = 2, ORQG_SORT_DEFAULT = 3 };
/* operators::table_sort with an explicit protocol: RADIXSORT as above, or
 * QUICKSORT for the multi-bit sort columns (single-bit columns use 1-bit
 * radix sort under either), with the reserve counts of sorting.h:531-550
 * for that protocol.                                                        */
int orqg_table_sort_protocol(orqg_ctx *ctx, orqg_vec *const *keys, size_t nkeys,
                             const int *orders, const int *single_bit, orqg_vec *const *a_cols,
                             size_t na, orqg_vec *const *b_cols, size_t nb, int protocol);
/* operators::quicksort (include/core/operators/quicksort.h:212-244) of a
 * B-shared 32- or 64-bit vector in place (signed order; stable, through the
 * index padding to 64 or 128 bits): shuffle, rounds of pivot comparisons
 * whose bits are opened,
 * public partitions; perm_out (width 4) gets the A-shared sorting
 * permutation.  Consumes 3 permutations (reserved if the queue lacks them). */
int orqg_quicksort(orqg_ctx *ctx, orqg_vec *v, int order, orqg_vec *perm_out);

/* ---- operators ------------------------------------------------------- */
/* radix_sort (include/core/operators/radixsort.h:205-229): sorts the
 * B-shared vector v in place on its low `bits` bits (bits == 8*width sorts
 * signed), returns the A-shared sorting permutation in perm_out (width 4). */
int orqg_radix_sort(orqg_ctx *ctx, orqg_vec *v, int order, int bits, orqg_vec *perm_out);
/* radix_sort of m independent requests at once (a served batch).  Request q
 * sorts vs[q] on its own context ctxs[q] (all distinct, one device) with
 * orders[q] and bits[q], into perms_out[q], and gets exactly what
 * orqg_radix_sort on that context gives: the same shares, nonces,
 * permutation queue and message log.  The permutations of the small requests
 * are generated in shared launches, and the bit iterations
 * (radixsort.h:141-192) of all requests run in shared, segmented launches,
 * whose shapes are tuned online.  The first context carries the error text. */
int orqg_radix_sort_batch(orqg_ctx *const *ctxs, orqg_vec *const *vs, const int *orders,
                          const int *bits, orqg_vec *const *perms_out, size_t m);

/* ---- share files (the reference's CSV formats) ------------------------ */
/* Each call covers the three parties' files: party_files[p] holds party p's
 * replicated pair (s_p, s_{p+1}); a pair whose two holders disagree is
 * refused (ORQG_E_PARITY).  Values are signed decimals of the share width.  */
/* EVector(size, file) (include/core/containers/e_vector.h:316-348): one row
 * "a,b" (or "a b") per element; a first line that is not numeric is skipped;
 * rows past v's size are ignored, missing rows are zero shares.             */
int orqg_vec_read_share_files(orqg_ctx *ctx, orqg_vec *v, const char *const party_files[3]);
/* EVector::output (e_vector.h:249-271)                                      */
int orqg_vec_write_share_files(orqg_ctx *ctx, const orqg_vec *v,
                               const char *const party_files[3]);
/* EncodedTable::inputCSVTableSecretShares (encoded_table.h:491-560): header
 * "[COL]_0,[COL]_1,..."; names[j] names cols[j] (all of one size n).  When the
 * table has "[##V]" and the files lack it, party 0 secret-shares ones into it
 * over the rows read.                                                        */
int orqg_table_read_share_csv(orqg_ctx *ctx, const char *const party_files[3],
                              const char *const *names, orqg_vec *const *cols, size_t ncols);
/* EncodedTable::outputCSVTableSecretShares (encoded_table.h:571-628), columns
 * in the order given (the reference writes its schema, i.e. sorted names)   */
int orqg_table_write_share_csv(orqg_ctx *ctx, const char *const party_files[3],
                               const char *const *names, orqg_vec *const *cols, size_t ncols);
/* EncodedTable::inputCSVTableData (encoded_table.h:391-488): the input
 * party's plaintext CSV (header of column names, all in the table); every
 * column is then secret-shared from input_party in the order given, "[##V]"
 * being 1 on the rows read when the file lacks it.                          */
int orqg_table_read_csv_data(orqg_ctx *ctx, const char *path, const char *const *names,
                             orqg_vec *const *cols, size_t ncols, int input_party);

/* ---- message log (Communicator byte accounting, communicator.h:14-183) - */
typedef struct {
  char op[24];
  int from, to;
  uint64_t bytes;
} orqg_msg;
size_t orqg_msglog_count(orqg_ctx *ctx);
int orqg_msglog_get(orqg_ctx *ctx, size_t i, orqg_msg *out);
int orqg_msglog_clear(orqg_ctx *ctx);

/* ---- timing (HIP events on the context stream) ------------------------ */
/* per-kernel-family device time accumulated since the last reset, in ms;
 * family: "aes_genperm", "sort_round", ...; returns 0 if unknown          */
int orqg_timer_enable(orqg_ctx *ctx, int enable);
int orqg_timer_get(orqg_ctx *ctx, const char *family, double *ms, uint64_t *launches);
int orqg_timer_reset(orqg_ctx *ctx);


/* ---- EncodedTable (include/core/containers/tabular/encoded_table.h) ---------
 * A table of named shared columns of one width, ordered by name (std::map,
 * encoded_table.h:147).  A name containing '[' is boolean-shared, any other
 * arithmetic-shared (isBShared, encoded_table.h:1672); every table has the
 * valid column "[##V]" (configureValid, encoded_table.h:274-285); the sort
 * keys "[##V]" and "[##TID]" are sorted on one bit (encoded_table.h:915-935).
 * Columns are owned by the table; orqg_et_column lends one until the table
 * changes it or is freed.  The protocols of orqg_et_sort are RADIXSORT,
 * QUICKSORT, BITONICSORT and SORT_DEFAULT (the context's default sort
 * protocol, orqg_ctx_set_default_sort: QUICKSORT, as in the reference).     */
typedef struct orqg_et orqg_et;
typedef struct {
  const char *data;    /* input column */
  const char *result;  /* result column (may be `data`) */
  int func;            /* ORQG_AGG_* */
} orqg_et_agg;
/* EncodedTable(name, schema, rows) (encoded_table.h:244-249): zero columns
 * plus the valid column set to ones                                         */
int orqg_et_create(orqg_ctx *ctx, int width, size_t rows, const char *const *names, size_t ncols,
                   orqg_et **out);
int orqg_et_free(orqg_et *t);
size_t orqg_et_rows(const orqg_et *t);
size_t orqg_et_ncols(const orqg_et *t);
/* name of column i in schema order (valid until the next call on t) */
const char *orqg_et_column_name(orqg_et *t, size_t i);
orqg_vec *orqg_et_column(orqg_et *t, const char *name);
/* copy v's shares into column `name` (created if absent; its encoding follows the name) */
int orqg_et_set_column(orqg_et *t, const char *name, const orqg_vec *v);
/* addColumns / deleteColumns / project (encoded_table.h:724-803)           */
int orqg_et_add_columns(orqg_et *t, const char *const *names, size_t n);
int orqg_et_delete_columns(orqg_et *t, const char *const *names, size_t n);
int orqg_et_project(orqg_et *t, const char *const *names, size_t n);
/* filter (encoded_table.h:298-314): [##V] &= column                         */
int orqg_et_filter(orqg_et *t, const char *name);
/* resize / tail (encoded_table.h:1407-1436)                                 */
int orqg_et_resize(orqg_et *t, size_t n);
int orqg_et_tail(orqg_et *t, size_t n);
/* shuffle (encoded_table.h:1003-1018)                                       */
int orqg_et_shuffle(orqg_et *t);
/* sort (encoded_table.h:900-998): every column by the keys, keys[0] most
 * significant; orders NULL = all ascending                                   */
int orqg_et_sort(orqg_et *t, const char *const *keys, const int *orders, size_t nk, int protocol);
/* distinct (encoded_table.h:1199-1237): sort on [##V] + keys, keep the first
 * row of each key group valid                                               */
int orqg_et_distinct(orqg_et *t, const char *const *keys, size_t nk);
/* aggregate (encoded_table.h:1059-1188) with its options: reverse direction,
 * do_sort (on [##V] + keys first), mark_valid (keep one row per group valid),
 * table_id (the join selector column, or NULL)                              */
int orqg_et_aggregate(orqg_et *t, const char *const *keys, size_t nk, const orqg_et_agg *spec,
                      size_t ns, int reverse, int do_sort, int mark_valid, const char *table_id);
/* concatenate (encoded_table.h:1548-1616): a's rows then b's, with the table
 * id column "[##TID]" (0 for a, a sharing of 1 for b)                       */
int orqg_et_concatenate(orqg_et *a, orqg_et *b, orqg_et **out);
/* operators::inner_join (operators/join.h:31-233): left (unique keys) JOIN
 * right ON keys, with aggregations `spec` (e.g. copying left's columns down);
 * trim_invalid drops rows that cannot be valid when that pays (join.h)      */
int orqg_et_inner_join(orqg_et *left, orqg_et *right, const char *const *keys, size_t nk,
                       const orqg_et_agg *spec, size_t ns, int trim_invalid, orqg_et **out);

/* ---- serving loop (no counterpart in the reference: its batched/served form) --
 * Secret-shared query requests on one GPU with continuous batching: `slots`
 * worker contexts run admitted requests concurrently; a request is admitted
 * in arrival order when a slot is free and its estimated device footprint fits
 * hbm_budget beside the running ones.  Each request re-keys its slot with its
 * own keys, so its result shares equal those of the same request run alone.
 * The slot contexts sort with radix sort as their DEFAULT protocol (the
 * reference built with DEFAULT_SORT_PROTO=RADIXSORT, sorting.h:6-9), so that
 * the small sorts of concurrent requests can share batched launches.
 * Host buffers named by a request must stay valid until its completion.      */
enum { ORQG_REQ_SORT = 0,        /* EncodedTable::sort on [K], data V and [##V]  */
       ORQG_REQ_GROUPBY_SUM = 1  /* EncodedTable::aggregate([K], {V, V, sum})     */ };
typedef struct {
  uint64_t id;
  int op;                            /* ORQG_REQ_* */
  int width;                         /* 4 or 8 bytes per element */
  size_t n;                          /* rows */
  const void *key_shares;            /* host [3][n]: s_0, s_1, s_2 of the B-shared key */
  const void *val_shares;            /* host [3][n]: the A-shared value column */
  uint8_t keys[ORQG_NUM_KEYS][32];   /* the request's PRG keys (roles as above) */
  void *out_key, *out_val, *out_valid; /* host [3][n] result shares (NULL: skip) */
} orqg_request;
typedef struct {
  uint64_t id;
  int status;                        /* ORQG_OK or an ORQG_E_* code */
  int slot;                          /* worker that ran it, -1 if rejected */
  double t_submit_ms, t_start_ms, t_done_ms; /* server clock */
  char msg[192];                     /* error text when status != ORQG_OK */
} orqg_completion;
typedef struct {
  int device;
  int slots;                         /* concurrent requests, 1..64 */
  size_t hbm_budget;                 /* admission budget in bytes (0: 200 GiB) */
  long batch_size;                   /* runtime batch size as orqg_config */
  size_t batch_rows;                 /* radix sorts of at most this many rows from
                                        concurrent requests run as one
                                        orqg_radix_sort_batch (0: each alone) */
  int batch_window_us;               /* longest wait for more sorts to join a batch */
} orqg_server_config;
typedef struct {
  uint64_t submitted, completed, rejected, max_running, peak_reserved, peak_live_bytes;
  uint64_t failed;      /* completed with a status other than ORQG_OK */
  uint64_t mallocs;     /* hipMalloc calls of the process's allocator */
  uint64_t pool_trims;  /* pooled blocks it released (pool cap, or memory pressure) */
  uint64_t sort_batches, sorts_batched; /* shared radix-sort batches, the sorts in them */
} orqg_server_stats;
typedef struct orqg_server orqg_server;
int orqg_server_create(const orqg_server_config *cfg, orqg_server **out);
int orqg_server_submit(orqg_server *srv, const orqg_request *req);
/* up to `max` completions; waits up to wait_ms for the first one           */
int orqg_server_poll(orqg_server *srv, orqg_completion *out, int max, int *count, int wait_ms);
int orqg_server_pending(orqg_server *srv, int *queued, int *running);
int orqg_server_get_stats(orqg_server *srv, orqg_server_stats *stats);
double orqg_server_now_ms(orqg_server *srv);
/* runs every queued request to completion, then stops the workers           */
int orqg_server_destroy(orqg_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* ORQG_H_ */
