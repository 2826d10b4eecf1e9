/*
 * provekit_gpu — C ABI of the MI355X (gfx950) implementation of ProveKit's
 * WHIR commitment hot path: Reed-Solomon encoding by NTT over BN254 Fr,
 * Skyscraper / Poseidon2 Merkle hashing, and the R1CS sparse products and
 * sumcheck rounds.
 *
 * Conventions
 *  - Field elements are ark-ff `ark_bn254::Fr` values as laid out in memory:
 *    4 little-endian u64 limbs of x * 2^256 mod p (Montgomery form), fully
 *    reduced.  A `uint64_t*` "Fr array" of n elements holds 4n limbs.
 *  - Skyscraper byte buffers are the reference's `[u8]` views: each 32-byte
 *    word is a little-endian 256-bit integer (any value; reduced mod p).
 *  - Every function returns a pkgpu_status; on failure the calling thread's
 *    message is available from pkgpu_last_error() (the FFI pattern of
 *    tooling/provekit-ffi/src/ffi.rs:29-87).  Nothing unwinds across the ABI.
 *  - "host" entry points take caller-owned host buffers (the reference's
 *    slices) and return when the result is in them.  "_device" entry points
 *    take device pointers (from pkgpu_malloc or any HIP allocation on the
 *    current device) and a hipStream_t passed as void* (NULL: the calling
 *    thread's stream); they return after enqueueing.
 *  - All entry points use the calling thread's current HIP device
 *    (pkgpu_init selects it).  The library is thread-safe.
 */
#ifndef PROVEKIT_GPU_H
#define PROVEKIT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PKGPU_ABI_VERSION 1

typedef enum {
  PKGPU_OK = 0,
  PKGPU_ERR_ARG = 1,      /* null pointer, length mismatch, bad option */
  PKGPU_ERR_NOT_POW2 = 2, /* a size must be a power of two */
  PKGPU_ERR_SIZE = 3,     /* divisibility or range violation (e.g. order above 2^28) */
  PKGPU_ERR_OOM = 4,      /* device allocation failed */
  PKGPU_ERR_DEVICE = 5,   /* HIP runtime error */
  PKGPU_ERR_BUDGET = 6,   /* serving admission refused: HBM budget */
  PKGPU_ERR_STATE = 7,    /* unknown handle or handle in the wrong state */
  PKGPU_ERR_INTERNAL = 8  /* a self-check failed */
} pkgpu_status;

/* ------------------------------------------------------------ library */
/* Select device `device` for the calling thread and initialise it.
 * Replaces provekit_common::register_ntt (provekit/common/src/lib.rs:44-63)
 * as the one-time set-up call; idempotent. */
int pkgpu_init(int device);
int pkgpu_abi_version(void);
/* Message of the calling thread's last failure ("" if none). */
const char* pkgpu_last_error(void);
int pkgpu_device_synchronize(void);

/* Device memory helpers for callers without their own HIP allocator. */
int pkgpu_malloc(void** ptr, size_t bytes);
int pkgpu_free(void* ptr);
int pkgpu_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int pkgpu_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int pkgpu_stream_synchronize(void* stream);

/* ------------------------------------------------------------ ntt crate */
/* ntt::ntt_nr(values, codeword_size, num_groups) (ntt/src/ntt.rs:105-108):
 * in-place NTT, natural -> bit-reversed order, of len / codeword_size
 * interleaved polynomials, starting at stage log2(num_groups).
 * Errors as the reference's asserts (ntt.rs:164-169): len % num_groups,
 * non-power-of-two codeword_size. */
int pkgpu_ntt_nr(uint64_t* values, size_t len, size_t codeword_size, size_t num_groups);
int pkgpu_ntt_nr_device(uint64_t* values, size_t len, size_t codeword_size, size_t num_groups,
                        void* stream);
/* ntt::intt_rn(values) (ntt/src/ntt.rs:326-330): inverse, bit-reversed ->
 * natural order. */
int pkgpu_intt_rn(uint64_t* values, size_t len);
int pkgpu_intt_rn_device(uint64_t* values, size_t len, void* stream);
/* ntt's intt_nr(values) (ntt/src/ntt.rs:333-348, a private fn there that
 * intt_rn and the crate's tests call): inverse, natural-order evaluations ->
 * coefficients in bit-reversed order.  PKGPU_ERR_NOT_POW2 where ntt_nr's
 * assert panics (a length that is not a power of two). */
int pkgpu_intt_nr(uint64_t* values, size_t len);
int pkgpu_intt_nr_device(uint64_t* values, size_t len, void* stream);
/* ntt's reverse_order(values, codeword_size) (ntt/src/ntt.rs:310-323, private
 * there): the first codeword_size entries of the len-element slice in
 * bit-reversed order, in place.  PKGPU_ERR_NOT_POW2 where its assert panics
 * (codeword_size 0 or not a power of two), PKGPU_ERR_SIZE where its swap
 * would index past the slice (codeword_size > len). */
int pkgpu_reverse_order(uint64_t* values, size_t len, size_t codeword_size);
int pkgpu_reverse_order_device(uint64_t* values, size_t len, size_t codeword_size, void* stream);

/* NTT schedule tuning (no counterpart in the reference, whose ntt crate runs
 * one fixed schedule, ntt/src/ntt.rs:150-248).  Every transform above picks
 * its pass shape (tile of 2^tile_log elements in LDS, butterflies per thread,
 * minimum consecutive run 2^min_run in early passes) from a table keyed by
 * (log2 codeword, stages run).  The table is read when the library starts
 * from $PKGPU_NTT_TUNING, else from ntt_tuning.json beside the library;
 * missing lengths take the nearest tuned one, an empty table tiles of 2^11.
 * autotune times every shape for log2 codeword in [log_min, log_max] (1..27)
 * and 0..max_skip skipped stages (blowup 2^max_skip), keeps the fastest and,
 * if save_path is given, writes the table there. */
int pkgpu_ntt_autotune(int log_min, int log_max, int max_skip, const char* save_path,
                       double* seconds);
int pkgpu_ntt_tuning_load(const char* path);
int pkgpu_ntt_tuning_get(int log_codeword, int stages, int* tile_log, int* bpt, int* min_run);

/* ------------------------------------------------------------ RSFr plugin */
/* ReedSolomon<Fr> for RSFr (provekit/common/src/ntt.rs:12-105). */
/* next_order: PKGPU_ERR_SIZE where the reference returns None (> 2^28). */
int pkgpu_rs_next_order(size_t size, size_t* order);
/* generator(codeword_length): the primitive codeword_length-th root of unity. */
int pkgpu_rs_generator(size_t codeword_length, uint64_t out[4]);
/* evaluation_points: out[t] = generator^{bitrev(indices[t])}. */
int pkgpu_rs_evaluation_points(size_t masked_message_length, size_t codeword_length,
                               const size_t* indices, size_t n, uint64_t* out);
/* interleaved_encode(messages, masks, codeword_length): messages[r] points at
 * message_length elements; out receives num_messages * codeword_length. */
int pkgpu_rs_interleaved_encode(const uint64_t* const* messages, size_t num_messages,
                                size_t message_length, const uint64_t* masks, size_t masks_len,
                                size_t codeword_length, uint64_t* out);
/* Device form: messages is one row-major num_messages x message_length
 * array.  coset != 0 evaluates on the coset 5*H (coefficient j scaled by 5^j;
 * 5 = multiplicative generator of Fr) instead of H. */
int pkgpu_rs_interleaved_encode_device(const uint64_t* messages, size_t num_messages,
                                       size_t message_length, const uint64_t* masks,
                                       size_t masks_len, size_t codeword_length, int coset,
                                       uint64_t* out, void* stream);

/* ------------------------------------------------------------ Skyscraper */
/* skyscraper::simple::compress(l, r) (skyscraper/core/src/simple.rs:12-14). */
int pkgpu_skyscraper_compress(const uint64_t l[4], const uint64_t r[4], uint64_t out[4]);
/* skyscraper::simple::compress_many(messages, hashes) (simple.rs:3-9):
 * messages_len == 2 * hashes_len, multiples of 64 and 32 bytes. */
int pkgpu_skyscraper_compress_many(const uint8_t* messages, size_t messages_len,
                                   uint8_t* hashes, size_t hashes_len);
/* SkyscraperHashEngine::hash_many(size, input, output)
 * (provekit/common/src/skyscraper/whir.rs:42-100): count = output digests,
 * input_len must equal size * count, size a positive multiple of 32. */
int pkgpu_skyscraper_hash_many(size_t size, const uint8_t* input, size_t input_len,
                               uint8_t* output, size_t count);
int pkgpu_skyscraper_hash_many_device(size_t size, const void* input, size_t count, void* output,
                                      void* stream);
/* Binary Merkle tree of n (power of two) 32-byte leaf digests with
 * compress(left, right); nodes receives 2n-1 digests level by level, root
 * last (nodes may equal leaves). */
int pkgpu_skyscraper_merkle_device(const void* leaves, size_t n, void* nodes, void* stream);

/* Openings from a device-resident commitment (the codeword and nodes of
 * pkgpu_commit_device, or a Poseidon2 tree with the same node layout): for
 * each of the n host leaf indices (< leaves, a power of two), values (host)
 * receives the leaf's leaf_elems elements and paths (host) the log2(leaves)
 * sibling digests of 32 bytes, leaf level first.  The data a whir query
 * opening reads (the protocol itself lives in the unvendored whir crate);
 * the tree is not copied back. */
int pkgpu_merkle_open_device(const uint64_t* codeword, const void* nodes, size_t leaves,
                             size_t leaf_elems, const uint64_t* leaf_indices, size_t n,
                             uint64_t* values, uint8_t* paths, void* stream);

/* ------------------------------------------------------ proof of work */
/* skyscraper::pow::threshold (skyscraper/core/src/pow.rs:18-26): 4 LE u64.
 * PKGPU_ERR_ARG ("Difficulty must be in the range [0, 80)") outside [0, 80),
 * where the reference panics. */
int pkgpu_skyscraper_pow_threshold(double difficulty, uint64_t out[4]);
/* skyscraper::pow::verify (pow.rs:28-30): *ok = 1 when difficulty is 0 or
 * compress(challenge, [nonce, 0, 0, 0]) < threshold(difficulty).  The
 * challenge is any 256-bit integer (4 LE u64, the reference's
 * transmute of the 32-byte spongefish challenge). */
int pkgpu_skyscraper_pow_verify(const uint64_t challenge[4], double difficulty, uint64_t nonce,
                                int* ok);
/* skyscraper::pow::solve (pow.rs:36-45, generic.rs:42-71): the smallest nonce
 * passing threshold(difficulty + 0.01); 0 when difficulty is 0.  Searches
 * nonces [0, max_nonces) (max_nonces 0: no bound, as the reference) and sets
 * *found = 0 if none passes.  Synchronises `stream` (NULL: the calling
 * thread's stream) once per window of nonces. */
int pkgpu_skyscraper_pow_solve(const uint64_t challenge[4], double difficulty,
                               uint64_t max_nonces, uint64_t* nonce, int* found, void* stream);

/* ------------------------------------------------------------ Poseidon2 */
/* poseidon2::poseidon2_permutation (poseidon2/src/permutation.rs:174-176) on
 * n states of 4 Fr (4n elements, row-major). */
int pkgpu_poseidon2_permutation_many(const uint64_t* states, size_t n, uint64_t* out);
/* poseidon2::poseidon2_hash (poseidon2/src/hash.rs:25-55) on one message. */
int pkgpu_poseidon2_hash(const uint64_t* inputs, size_t len, uint64_t out[4]);
/* poseidon2_hash on `count` messages of `len` elements each (row-major). */
int pkgpu_poseidon2_hash_many(const uint64_t* inputs, size_t len, size_t count, uint64_t* out);
int pkgpu_poseidon2_hash_many_device(const uint64_t* inputs, size_t len, size_t count,
                                     uint64_t* out, void* stream);
/* Merkle tree with poseidon2_hash([left, right]) per node, as
 * compute_merkle_root (playground/passport-input-gen/src/commitment.rs:225-237)
 * walks it: n leaves (power of two); nodes (2n-1 Fr, may be NULL) level by
 * level, root last. */
int pkgpu_poseidon2_merkle(const uint64_t* leaves, size_t n, uint64_t* nodes, uint64_t root[4]);
int pkgpu_poseidon2_merkle_device(const uint64_t* leaves, size_t n, uint64_t* nodes,
                                  void* stream);
/* The device tree's level launches recorded once into a hipGraph bound to
 * these buffers; replay with pkgpu_graph_launch, free with pkgpu_graph_destroy. */
int pkgpu_poseidon2_merkle_graph(const uint64_t* leaves, size_t n, uint64_t* nodes,
                                 void** graph);
int pkgpu_graph_launch(void* graph, void* stream);
int pkgpu_graph_destroy(void* graph);

/* ------------------------------------------------------------ R1CS / sumcheck */
/* HydratedSparseMatrix * &[Fr] (provekit/common/src/sparse_matrix.rs:606-624):
 * CSR with row r spanning [row_starts[r], row_starts[r+1]) (the last row ends
 * at nnz), u32 column indices, values expanded from the interner.  Fails like
 * the reference's assert when z_len != num_cols. */
int pkgpu_sparse_mul(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                     const uint32_t* cols, const uint64_t* vals, const uint64_t* z, size_t z_len,
                     uint64_t* out);
int pkgpu_sparse_mul_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                            size_t nnz, const uint32_t* cols, const uint64_t* vals,
                            const uint64_t* z, uint64_t* out, void* stream);
/* The same product on the reference's own storage: HydratedSparseMatrix is
 * a SparseMatrix whose entries hold InternedFieldElement ids into an
 * Interner (sparse_matrix.rs:56-71,583-624, interner.rs:7-42).  ids: nnz u32
 * indices into interner (interner_len Fr, Montgomery form).  The host form
 * fails with the reference's "Value not in interner." for an id past the
 * table; the device form trusts the ids (one past the table reads as zero,
 * never out of bounds).  Streams 4 bytes per entry instead of 32. */
int pkgpu_sparse_mul_interned(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                              size_t nnz, const uint32_t* cols, const uint32_t* ids,
                              const uint64_t* interner, size_t interner_len, const uint64_t* z,
                              size_t z_len, uint64_t* out);
int pkgpu_sparse_mul_interned_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                     size_t nnz, const uint32_t* cols, const uint32_t* ids,
                                     const uint64_t* interner, size_t interner_len,
                                     const uint64_t* z, uint64_t* out, void* stream);
/* No counterpart in the reference: pkgpu_sparse_mul_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results; its
 * scratch (chunk flags and partial sums) is a fixed device array inside the
 * library, one per GPU, held until the process ends, so it calls no
 * hipMalloc/hipFree.  nnz <= PKGPU_SPMV_NOALLOC_MAX_NNZ (else PKGPU_ERR_SIZE).
 * stream NULL means the device's null stream.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other streams, other threads)
 * race; serialise them. */
#define PKGPU_SPMV_NOALLOC_MAX_NNZ ((size_t)1 << 26)
int pkgpu_sparse_mul_device_noalloc(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                    size_t nnz, const uint32_t* cols, const uint64_t* vals,
                                    const uint64_t* z, uint64_t* out, void* stream);
/* SparseMatrix::transpose (provekit/common/src/sparse_matrix.rs:409-449):
 * the num_cols x num_rows matrix whose entries are the (col, row, value)
 * triples in (col, row) order.  t_row_starts receives num_cols entries (row c
 * ends where row c+1 starts, the last at nnz), t_cols and t_vals nnz each.
 * The host-slice form checks the input like pkgpu_sparse_mul; the device form
 * trusts it (column indices below num_cols) and stays in bounds regardless. */
int pkgpu_sparse_transpose(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                           size_t nnz, const uint32_t* cols, const uint64_t* vals,
                           uint32_t* t_row_starts, uint32_t* t_cols, uint64_t* t_vals);
int pkgpu_sparse_transpose_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                  size_t nnz, const uint32_t* cols, const uint64_t* vals,
                                  uint32_t* t_row_starts, uint32_t* t_cols, uint64_t* t_vals,
                                  void* stream);
/* SparseMatrix::transpose on the reference's own storage
 * (provekit/common/src/sparse_matrix.rs:409-449 moves InternedFieldElement
 * ids, not field elements): the same transpose with the u32 value ids moving
 * into t_ids (nnz).  The ids are not looked up, so no interner is needed. */
int pkgpu_sparse_transpose_interned(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                    size_t nnz, const uint32_t* cols, const uint32_t* ids,
                                    uint32_t* t_row_starts, uint32_t* t_cols, uint32_t* t_ids);
int pkgpu_sparse_transpose_interned_device(size_t num_rows, size_t num_cols,
                                           const uint32_t* row_starts, size_t nnz,
                                           const uint32_t* cols, const uint32_t* ids,
                                           uint32_t* t_row_starts, uint32_t* t_cols,
                                           uint32_t* t_ids, void* stream);
/* One matrix of multiply_transposed_by_eq_alpha
 * (provekit/common/src/utils/sumcheck.rs:203-227): out (t_num_rows Fr) =
 * At * eq(alpha, .), the eq table over the m device-resident elements of
 * alpha truncated to t_num_cols entries (the constraint count).  At is the
 * transposed matrix, e.g. from pkgpu_sparse_transpose_device. */
int pkgpu_multiply_transposed_by_eq_alpha_device(size_t t_num_rows, size_t t_num_cols,
                                                 const uint32_t* t_row_starts, size_t nnz,
                                                 const uint32_t* t_cols, const uint64_t* t_vals,
                                                 const uint64_t* alpha, size_t m, uint64_t* out,
                                                 void* stream);
/* The same product on an interned At (the reference's at.hydrate(interner) *
 * eq_alpha, sumcheck.rs:216-224, without the hydrated copy): t_ids index the
 * interner table of interner_len Fr; an id beyond it reads as zero. */
int pkgpu_multiply_transposed_by_eq_alpha_interned_device(
    size_t t_num_rows, size_t t_num_cols, const uint32_t* t_row_starts, size_t nnz,
    const uint32_t* t_cols, const uint32_t* t_ids, const uint64_t* interner, size_t interner_len,
    const uint64_t* alpha, size_t m, uint64_t* out, void* stream);
/* ---------------------------------------------------- R1CS wire format */
/* Device CSR filled by the decoders below (caller-allocated buffers). */
typedef struct {
  size_t num_rows, num_cols, nnz;
  uint32_t* row_starts; /* device, num_rows entries (row r ends at r+1's start, the last at nnz) */
  uint32_t* cols;       /* device, nnz entries */
  uint64_t* vals;       /* device, nnz Fr (Montgomery form) */
} pkgpu_csr;

/* The postcard encodings (the `postcard` crate, Cargo.toml:124) of Interner
 * (provekit/common/src/interner.rs:6-10 with utils/serde_ark.rs:11-30),
 * SparseMatrix (provekit/common/src/sparse_matrix.rs:149-290: delta-coded
 * columns) and R1CS (provekit/common/src/r1cs.rs:59-66), as
 * postcard::to_allocvec writes them, decoded on the device.  Blobs are host
 * memory.  Each call writes the sizes; it also fills the outputs when they are
 * given (interner decode: out != NULL; matrix decode: any of m's pointers set;
 * R1CS decode: interner != NULL, and then every mats[i] buffer of non-zero
 * size), which must hold the sizes a first call reported.  Matrix values come
 * out hydrated (InternedFieldElement -> Fr).  Malformed input
 * (PKGPU_ERR_ARG): a bad or truncated varint, num_rows different from the
 * row-start count, row starts not starting at 0, decreasing or past the value
 * count, delta count != value count, a column >= num_cols, an id >= the
 * interner size, an interner integer >= p (the reference's deserializer
 * checks the first and last of these; the others would panic later or give an
 * inconsistent matrix there).  Synchronises the stream. */
int pkgpu_interner_decode_device(const uint8_t* blob, size_t len, uint64_t* out, size_t* count,
                                 void* stream);
int pkgpu_sparse_matrix_decode_device(const uint8_t* blob, size_t len, const uint64_t* interner,
                                      size_t interner_len, pkgpu_csr* m, void* stream);
int pkgpu_r1cs_decode_device(const uint8_t* blob, size_t len, size_t* num_public_inputs,
                             uint64_t* interner, size_t* interner_len, pkgpu_csr mats[3],
                             void* stream);

/* calculate_witness_bounds (provekit/common/src/utils/sumcheck.rs:163-179):
 * a = A w, b = B w, c = a o b, each zero-padded to padded_len =
 * num_rows.next_power_of_two() (written to *padded_len; the out arrays must
 * hold that many elements). */
int pkgpu_calculate_witness_bounds(size_t num_rows, size_t num_cols, const uint32_t* a_row_starts,
                                   size_t a_nnz, const uint32_t* a_cols, const uint64_t* a_vals,
                                   const uint32_t* b_row_starts, size_t b_nnz,
                                   const uint32_t* b_cols, const uint64_t* b_vals,
                                   const uint64_t* w, size_t w_len, uint64_t* a_out,
                                   uint64_t* b_out, uint64_t* c_out, size_t* padded_len);
/* calculate_evaluations_over_boolean_hypercube_for_eq(r, num_entries)
 * (sumcheck.rs:108-120): eq(r, x) for x < num_entries, r[0] on the top bit. */
int pkgpu_eq_evaluations(const uint64_t* r, size_t m, size_t num_entries, uint64_t* out);
int pkgpu_eq_evaluations_device(const uint64_t* r, size_t m, size_t num_entries, uint64_t* out,
                                void* stream);
/* sumcheck_fold_map_reduce([a, b, c, eq], fold, map) with the R1CS map of
 * run_zk_sumcheck_prover (provekit/prover/src/whir_r1cs.rs:551-559), on
 * device-resident MLEs of `size` elements.  fold: host pointer to one Fr or
 * NULL (round 0).  With a fold the four MLEs are folded in place and the
 * caller keeps the first size/2 elements (the reference's truncate).
 * out: host, 3 Fr = (h(0), h(-1), h(inf)). */
int pkgpu_sumcheck_r1cs_round_device(uint64_t* a, uint64_t* b, uint64_t* c, uint64_t* eq,
                                     size_t size, const uint64_t* fold, uint64_t* out,
                                     void* stream);
/* The whole sumcheck of run_zk_sumcheck_prover
 * (provekit/prover/src/whir_r1cs.rs:519-621) on the device, with this
 * library's Skyscraper transcript (seeded) in place of spongefish.
 * a, b, c: device, 2^m elements each, consumed (m = 0: one element each, run
 * zero-padded to two like pad_to_pow2_len_min2).  blinding: host m x 4 Fr
 * (reference ZK mode: sum_g absorbed, rho squeezed, claim rho * sum_g, and
 * after the last round blinding_eval = sum_i g_i(alpha_i) absorbed, :613-618;
 * the univariates are the 4m elements the commit wrote at w1[offset..], so
 * this is <expand_powers::<4>(alpha), w1[offset..offset+4m]>; needs m >= 1)
 * or NULL (plain mode: rho = 0, claim = sum eq (a b - c) computed on device,
 * blinding_eval = 0 and not absorbed).
 * Outputs (host): coeffs m x 4 Fr, alphas m Fr, r m Fr (the eq point),
 * final_claim 1 Fr (the last cubic at alpha_{m-1}), rho and blinding_eval 1
 * Fr each (may be NULL).  The verifier's check is final_claim - rho *
 * blinding_eval = eq(r, alpha)(a(alpha) b(alpha) - c(alpha)).  Returns
 * PKGPU_ERR_INTERNAL if a round identity fails. */
int pkgpu_sumcheck_prove_device(uint64_t* a, uint64_t* b, uint64_t* c, size_t m,
                                const uint64_t* blinding, uint64_t seed, uint64_t* coeffs,
                                uint64_t* alphas, uint64_t* r, uint64_t* final_claim,
                                uint64_t* rho, uint64_t* blinding_eval, void* stream);

/* ------------------------------------------------------------ commit */
/* WHIR-style commitment of one polynomial of n coefficients (n, blowup powers
 * of two): RS-encode to n*blowup elements (RSFr order; coset != 0: on 5*H),
 * hash leaves of leaf_elems consecutive codeword elements by left fold
 * (hash_many with size 32*leaf_elems), binary Skyscraper tree to the root.
 * This is the work of WhirR1CSScheme::commit -> whir commit
 * (provekit/prover/src/whir_r1cs.rs:82-140) with the leaf layout fixed here.
 * Host form: codeword (n*blowup Fr) and nodes (2*leaves-1 digests) may be
 * NULL; root receives 32 bytes. */
int pkgpu_commit(const uint64_t* coeffs, size_t n, size_t blowup, int coset, size_t leaf_elems,
                 uint64_t* codeword, uint8_t* nodes, uint8_t root[32]);
int pkgpu_commit_device(const uint64_t* coeffs, size_t n, size_t blowup, int coset,
                        size_t leaf_elems, uint64_t* codeword, void* nodes, void* stream);

/* WhirR1CSScheme::new_from_dimensions (provekit/r1cs-compiler/src/whir_r1cs.rs:55-84):
 * the scheme's witness variables m = max(log2 w1, log2 w2, 13) (plus one when
 * w1's zero padding has no room for 4 m0 coefficients) and sumcheck
 * variables m0 = max(log2 num_constraints, 1), log2 rounded up. */
int pkgpu_whir_r1cs_scheme_size(size_t num_witnesses, size_t num_constraints, size_t w1_size,
                                size_t* m, size_t* m0);
/* WhirR1CSScheme::commit (provekit/prover/src/whir_r1cs.rs:82-140): the
 * witness (witness_size Fr) zero-padded to 2^num_vars (pad_to_power_of_two,
 * then resize; num_vars is whir's num_witness_variables, the scheme's m),
 * the blinding univariates (m0 x 4 Fr, row-major; NULL for the w2
 * commitment) written at offset = witness_size, then committed as
 * pkgpu_commit.  polynomial receives the shaped 2^num_vars elements (the
 * commitment's `polynomial`, whose [offset, offset + 4 m0) slice the
 * sumcheck's blinding_eval reads).  Fails with the reference's messages when
 * the witness exceeds 2^num_vars or the coefficients run past the end.
 * Host form: polynomial, codeword, nodes may be NULL.  Device form: all
 * device pointers; witness may equal polynomial, blinding must not overlap
 * it. */
int pkgpu_whir_r1cs_commit(const uint64_t* witness, size_t witness_size, size_t num_vars,
                           const uint64_t* blinding, size_t m0, size_t blowup, int coset,
                           size_t leaf_elems, uint64_t* polynomial, uint64_t* codeword,
                           uint8_t* nodes, uint8_t root[32]);
int pkgpu_whir_r1cs_commit_device(const uint64_t* witness, size_t witness_size, size_t num_vars,
                                  const uint64_t* blinding, size_t m0, size_t blowup, int coset,
                                  size_t leaf_elems, uint64_t* polynomial, uint64_t* codeword,
                                  void* nodes, void* stream);

/* ------------------------------------------------------------ bn254_multiplier */
/* bn254_multiplier::scalar_mul / scalar_sqr
 * (skyscraper/bn254-multiplier/src/scalar.rs:12-132) element-wise over n
 * elements: out[i] = a[i] b[i] 2^-256 (a[i]^2 2^-256) mod p as the reference
 * computes it (the log-jump reduction, then reduce_ct), below 2^256 - 2p and
 * not canonical, byte for byte the reference's output.  Inputs: 4 LE u64
 * limbs each, at most 2^256 - 2p (the reference's contract). */
int pkgpu_scalar_mul_many(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
int pkgpu_scalar_sqr_many(const uint64_t* a, uint64_t* out, size_t n);
int pkgpu_scalar_mul_many_device(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n,
                                 void* stream);
int pkgpu_scalar_sqr_many_device(const uint64_t* a, uint64_t* out, size_t n, void* stream);

/* ------------------------------------------------------------ synthetic inputs */
/* No counterpart in the reference (its benchmarks draw Fr::rand from
 * rand_chacha on the host): n seeded uniform field elements on the device.
 * out[i] = the ChaCha20 block (RFC 8439, nonce 0) with counter first + i
 * under `key` (32 bytes), read as a 512-bit little-endian integer mod p, used
 * as the Montgomery limbs.  first + n <= 2^32. */
int pkgpu_random_fr_device(const uint8_t key[32], uint64_t first, size_t n, uint64_t* out,
                           void* stream);

/* ------------------------------------------------------------ batched commit */
/* Many polynomials of uneven lengths committed together (no counterpart
 * call in the reference, which commits one witness per whir call): a plan
 * binds device buffers once; each run enqueues the coset scaling, one NTT
 * launch per pass round for all polynomials, and ONE persistent launch that
 * hashes every leaf and builds every tree from a device work queue.
 * Polynomial i: coeffs[i] (n[i] Fr, power of two), blowup[i] (power of two),
 * coset[i] (0: H, else 5*H), codewords[i] (n[i]*blowup[i] Fr),
 * nodes[i] (2*leaves-1 digests, leaves = n[i]*blowup[i]/leaf_elems).
 * tile_log: 0 takes each polynomial's tuned NTT shape; 6..11 forces tiles
 * of 2^tile_log elements with the default shape. */
int pkgpu_batch_plan_create(size_t count, const uint64_t* const* coeffs, const size_t* n,
                            const size_t* blowup, const int* coset, size_t leaf_elems,
                            uint64_t* const* codewords, void* const* nodes, int tile_log,
                            void** plan);
int pkgpu_batch_run(void* plan, void* stream);
/* The run recorded once into a hipGraph (cached in the plan) and launched. */
int pkgpu_batch_graph_launch(void* plan, void* stream);
/* Copy the count roots of the last run to host (32 bytes each) after
 * synchronising `stream`; fails if the hashing queue reported a timeout. */
int pkgpu_batch_roots(void* plan, uint8_t* roots, void* stream);
int pkgpu_batch_plan_destroy(void* plan);

/* ------------------------------------------------------------ serving loop */
/* A request loop for R1CS sumcheck + witness-commit jobs with continuous
 * batching (no counterpart in the reference, which proves one witness per
 * process).  Each job: a, b, c = A z, B z, C z (2^m rows, 2^m columns), the
 * whole sumcheck over eq(r, x)(a b - c) in plain mode with the library
 * transcript seeded by `seed`, and the commit of z (blowup, coset,
 * leaf_elems as pkgpu_commit).  Jobs are admitted in arrival order when a
 * stream slot is free and their HBM footprint fits the budget; admitted jobs
 * overlap on the device.  Inputs may be host or device pointers and must
 * stay valid until the job has finished.  The codeword and tree stay resident
 * (counted against the budget) until pkgpu_server_release. */
typedef struct {
  size_t m;
  const uint32_t* row_starts[3]; /* A, B, C */
  const uint32_t* cols[3];
  const uint64_t* vals[3];
  size_t nnz[3];
  const uint64_t* z;
  uint64_t seed;
  size_t blowup;
  int coset;
  size_t leaf_elems;
} pkgpu_job;

typedef struct {
  uint64_t root[4];        /* Merkle root of the witness commit (32 bytes) */
  uint64_t final_claim[4]; /* last round's cubic at the last challenge */
  double latency_ms;       /* submit -> finished */
  double queue_ms;         /* submit -> admitted */
  double run_ms;           /* admitted -> finished */
  int status;              /* pkgpu_status of the job */
} pkgpu_job_result;

int pkgpu_server_create(int device, size_t hbm_budget_bytes, int slots, void** server);
/* PKGPU_ERR_BUDGET if the job alone exceeds the budget. */
int pkgpu_server_submit(void* server, const pkgpu_job* job, uint64_t* job_id);
/* Blocks until the job finished (timeout_ms <= 0: no limit).  coeffs (m x 4
 * Fr), alphas (m Fr), r (m Fr) may be NULL. */
int pkgpu_server_wait(void* server, uint64_t job_id, double timeout_ms, pkgpu_job_result* result,
                      uint64_t* coeffs, uint64_t* alphas, uint64_t* r);
int pkgpu_server_poll(void* server, uint64_t job_id, int* done);
/* Openings from the resident tree: for each leaf index, the leaf's
 * leaf_elems elements (Fr) and the sibling digest at every level
 * (log2(leaves) x 32 bytes, leaf level first). */
int pkgpu_server_open(void* server, uint64_t job_id, const uint64_t* leaf_indices, size_t n,
                      uint64_t* leaf_values, uint8_t* paths);
int pkgpu_server_release(void* server, uint64_t job_id);
int pkgpu_server_stats(void* server, uint64_t* bytes_in_use, uint64_t* peak_in_flight);
int pkgpu_server_destroy(void* server);

#ifdef __cplusplus
}
#endif
#endif /* PROVEKIT_GPU_H */
