/* ORACLE — test infrastructure only (the CPU restatement that checks the HIP
 * path and times the reference's CPU algorithm).  Nothing in the product
 * links, loads or calls this.
 *
 * Every exported function restates one reference function; the citation is
 * on the definition.  Field elements are `ofr` (oracle/fr.h): 4 LE u64 limbs
 * in Montgomery form, canonical.  Byte buffers for Skyscraper are raw LE
 * 256-bit integers, as in the reference (`[u8]` views of `[u64; 4]`).
 */
#ifndef PK_ORACLE_H
#define PK_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#include "fr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* threads used by the parallel loops (rayon stand-in); 0 = all cores */
void or_set_threads(int n);
int or_get_threads(void);

/* ---- ntt (ntt/src/ntt.rs) ---- */
ofr or_root_of_unity(size_t n);
void or_roots_table(size_t order, ofr* out); /* order/2 entries */
void or_ntt_nr(ofr* values, size_t len, size_t codeword_size, size_t num_groups);
void or_intt_rn(ofr* values, size_t len);
void or_intt_nr(ofr* values, size_t len);
void or_reverse_order(ofr* values, size_t n);
/* RSFr (provekit/common/src/ntt.rs) */
int or_rs_next_order(size_t size, size_t* order);
void or_rs_evaluation_points(size_t codeword_length, const size_t* indices, size_t n, ofr* out);
/* messages: num_messages rows of message_length, row-major; out: num_messages*codeword_length */
int or_rs_interleaved_encode(const ofr* messages, size_t num_messages, size_t message_length,
                             const ofr* masks, size_t masks_len, size_t codeword_length, ofr* out);

/* ---- skyscraper (skyscraper/core, skyscraper/bn254-multiplier) ---- */
uint8_t or_sky_sbox(uint8_t v);
void or_sky_ss(int round, const uint64_t l[4], const uint64_t r[4], uint64_t ol[4], uint64_t orr[4]);
void or_sky_bb(int round, const uint64_t l[4], const uint64_t r[4], uint64_t ol[4], uint64_t orr[4]);
void or_sky_permute(const uint64_t l[4], const uint64_t r[4], uint64_t ol[4], uint64_t orr[4]);
void or_sky_compress_spec(const uint64_t l[4], const uint64_t r[4], uint64_t out[4]);
void or_sky_compress(const uint64_t l[4], const uint64_t r[4], uint64_t out[4]);
void or_scalar_sqr(const uint64_t a[4], uint64_t out[4]);
void or_scalar_mul(const uint64_t a[4], const uint64_t b[4], uint64_t out[4]);
void or_sky_compress_many(const uint8_t* messages, uint8_t* hashes, size_t count);
int or_sky_hash_many(size_t size, const uint8_t* input, size_t count, uint8_t* output);
/* proof of work (skyscraper/core/src/pow.rs) */
void or_f64_to_u256(double f, uint64_t out[4]);
int or_pow_threshold(double difficulty, uint64_t out[4]);
int or_pow_verify(const uint64_t challenge[4], double difficulty, uint64_t nonce);
int or_pow_solve(const uint64_t challenge[4], double difficulty, uint64_t limit, uint64_t* nonce);

/* ---- commit: RS encode + leaf hashing + binary Skyscraper tree ---- */
/* tree: 2*num_leaves-1 digests: nodes[0..num_leaves) leaves, then each level, root last */
void or_merkle_tree(const uint8_t* leaf_digests, size_t num_leaves, uint8_t* nodes);
int or_commit(const ofr* coeffs, size_t n, size_t blowup, int coset, size_t leaf_elems,
              ofr* codeword, uint8_t* nodes);
/* WhirR1CSScheme sizes and the witness shaping of WhirR1CSScheme::commit */
void or_whir_r1cs_params(size_t num_witnesses, size_t num_constraints, size_t w1_size, size_t* m,
                         size_t* m0);
int or_whir_r1cs_shape(const ofr* witness, size_t witness_size, size_t num_vars, const ofr* g,
                       size_t m0, ofr* out);

/* ---- poseidon2 (poseidon2/src) ---- */
void or_p2_permutation(const ofr in[4], ofr out[4]);
ofr or_p2_hash(const ofr* inputs, size_t n);
void or_p2_merkle_levels(const ofr* leaves, size_t n, ofr* nodes); /* 2n-1 nodes, root last */

/* ---- sumcheck / sparse (provekit/common/src/utils/sumcheck.rs, sparse_matrix.rs) ---- */
void or_eq_evals(const ofr* r, size_t m, size_t num_entries, ofr* out);
ofr or_calculate_eq(const ofr* r, const ofr* alpha, size_t n);
void or_sparse_mul(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                   const uint32_t* cols, const ofr* vals, const ofr* z, ofr* out);
int or_sparse_transpose(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                        const uint32_t* cols, const ofr* vals, uint32_t* t_row_starts,
                        uint32_t* t_cols, ofr* t_vals);
/* ---- wire format (postcard R1CS / Interner / SparseMatrix) ---- */
int or_varint_decode(const uint8_t* b, size_t len, int max_bytes, int bits, uint64_t* out,
                     size_t* used);
int or_interner_decode(const uint8_t* blob, size_t len, ofr* out, size_t cap, size_t* count);
int or_sparse_matrix_decode(const uint8_t* blob, size_t len, const ofr* interner,
                            size_t interner_len, size_t cap_rows, size_t cap_nnz, uint32_t* rs,
                            uint32_t* cols, ofr* vals, size_t* num_rows, size_t* num_cols,
                            size_t* nnz);
int or_r1cs_decode(const uint8_t* blob, size_t len, size_t sizes[11], uint32_t* const rs[3],
                   uint32_t* const cols[3], ofr* const vals[3], const size_t caps[6]);
void or_sumcheck_r1cs_round(ofr* a, ofr* b, ofr* c, ofr* eq, size_t size, const ofr* fold,
                            ofr out[3]);
ofr or_eval_cubic(const ofr poly[4], ofr x);
int or_sumcheck_prove(ofr* a, ofr* b, ofr* c, size_t m, const ofr* blinding, uint64_t seed,
                      ofr* coeffs, ofr* alphas, ofr* r_out, ofr* final_claim, ofr* extra);
void or_sumcheck_plant_fault(int kind);
void or_transcript_run(uint64_t seed, const ofr* absorbs, size_t n_abs, ofr* squeezes,
                       size_t n_sq);

ofr or_fr_mul(ofr a, ofr b);
ofr or_fr_add(ofr a, ofr b);
ofr or_fr_sub(ofr a, ofr b);

#ifdef __cplusplus
}
#endif
#endif
