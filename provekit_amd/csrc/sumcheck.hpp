// Device-resident sumcheck state and launchers for the R1CS products.
//
// Transcript (this project's; the reference's spongefish is not vendored):
// C is Skyscraper compression,
//   init(seed):     s = C([seed, 0, 0, 0], [kTrTag0, kTrTag1, 0, 0])   (u64 limbs)
//   absorb1(x):     s = C(s, x)
//   squeeze(k):     out_j = C(s, [j + 1, 0, 0, 0]) for j < k; s = C(s, 0)
//   round(c0..3):   s = C(C(s, c0), C(c2, c3)); the new state is the round
//                   challenge.  c1 is not absorbed: the verifier derives it
//                   from the claim (c1 = S - 2 c0 - c2 - c3), which the
//                   transcript already fixes.  Two dependent compressions.
// Field elements enter as their Montgomery limbs; a challenge is the Fr whose
// Montgomery limbs are out_j.  The compressions of one step that do not depend
// on each other run on separate lanes of one wave.
#pragma once
#include "runtime.hpp"

namespace pk {

constexpr uint64_t kTrTag0 = 0x70726f76656b6974ULL;  // "provekit"
constexpr uint64_t kTrTag1 = 0x73756d636865636bULL;  // "sumcheck"
constexpr int kMaxSumcheckVars = 28;

struct ScState {
  fr_t S;                                  // current claim
  fr_t rho;                                // blinding challenge (0 in plain mode)
  fr_t sum_g;                              // sum of the blinding polynomial over the cube
  fr_t blinding_eval;                      // sum_i g_i(alpha_i), after the last round
  fr_t state;                              // transcript state
  fr_t one, half;                          // Montgomery 1 and 1/2
  fr_t r[kMaxSumcheckVars];                // eq point
  fr_t alpha[kMaxSumcheckVars];            // round challenges
  fr_t coeffs[kMaxSumcheckVars][4];        // round polynomials c0..c3
  fr_t g[kMaxSumcheckVars][4];             // blinding univariates (reference mode)
  int m;                                   // number of variables / rounds
  int mode;                                // 0 plain (rho = 0, honest claim), 1 reference ZK
  int error;                               // set if a round identity fails
  int pad;
};

// CSR matrix on the device (SparseMatrix, provekit/common/src/sparse_matrix.rs:56-71,
// with interned values expanded to field elements).
struct CsrDev {
  uint64_t rows, ncols, nnz;
  const uint32_t* row_starts;  // rows entries; row r ends where row r+1 starts (last: nnz)
  const uint32_t* cols;
  const fr_t* vals;  // hydrated values (nnz), or null when ids are set
  // interned values: ids (nnz) into the interner table (interner_len Fr);
  // only the SpMV reads this form
  const uint32_t* ids = nullptr;
  const fr_t* interner = nullptr;
  uint64_t interner_len = 0;
};

void spmv_dev(const CsrDev& A, const fr_t* z, fr_t* out, hipStream_t s);
// The SpMV launches on caller-provided scratch: flags holds
// 2 * spmv_chunks(nnz) + 1 u32, partials 2 * spmv_chunks(nnz) elements (both
// fully rewritten by every call; the flags are cleared by a kernel).
uint64_t spmv_chunks(uint64_t nnz);
void spmv_launch(const CsrDev& A, const fr_t* z, fr_t* out, uint32_t* flags, fr_t* partials,
                 hipStream_t s);
// The same product with its scratch in a file-scope device array of this
// library (no allocation); nnz <= kNoAllocSpmvMaxNnz.  Calls overlapping on
// one device share the array and race.
constexpr uint64_t kNoAllocSpmvMaxNnz = 1ull << 26;
void noalloc_setup(int device);
void spmv_dev_noalloc(const CsrDev& A, const fr_t* z, fr_t* out, hipStream_t s);
// SparseMatrix::transpose: t_row_starts gets A.ncols entries, t_cols and
// t_vals A.nnz (entries in (column, row) order).
void sparse_transpose_dev(const CsrDev& A, uint32_t* t_row_starts, uint32_t* t_cols,
                          fr_t* t_vals, hipStream_t s);
// The same transpose of an interned matrix (A.ids set): the ids move with
// their entries into t_ids (A.nnz), as the reference's transpose moves them.
void sparse_transpose_interned_dev(const CsrDev& A, uint32_t* t_row_starts, uint32_t* t_cols,
                                   uint32_t* t_ids, hipStream_t s);
// out = At * eq(alpha, .) with the eq table truncated to At.ncols entries
// (one matrix of multiply_transposed_by_eq_alpha); alpha: m device elements.
void transposed_eq_mul_dev(const CsrDev& At, const fr_t* alpha, int m, fr_t* out,
                           hipStream_t s);
void hadamard_dev(const fr_t* a, const fr_t* b, fr_t* c, uint64_t n, uint64_t n_pad, hipStream_t s);
void zero_dev(fr_t* p, uint64_t n, hipStream_t s);
// eq(r, x) for x < num_entries (r: m device elements).
void eq_table_dev(const fr_t* r, int m, uint64_t num_entries, fr_t* out, hipStream_t s);
// One fold-map-reduce round; out3 receives (h(0), h(-1), h(inf)), canonical.
// partials: sc_partials_elems() elements.
void sc_round_dev(fr_t* a, fr_t* b, fr_t* c, fr_t* eq, uint64_t size, const fr_t* fold,
                  fr_t* partials, fr_t* out3, hipStream_t s);
size_t sc_partials_elems();
// Whole sumcheck on the device.  st: device ScState prepared by
// sc_state_init_host and uploaded; a, b, c: 2^m elements (consumed); eq:
// 2^m scratch; work: sc_partials_elems() + 3 elements.
void sc_prove_dev(ScState* st, fr_t* a, fr_t* b, fr_t* c, fr_t* eq, int m, int mode,
                  uint64_t seed, fr_t* work, hipStream_t s);
void sc_state_init_host(ScState& h, int m, int mode, const fr_t* blinding);

}  // namespace pk
