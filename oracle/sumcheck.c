/* ORACLE — test infrastructure only.  CPU restatement of the reference's
 * R1CS sumcheck pieces (provekit/common/src/utils/sumcheck.rs,
 * provekit/common/src/sparse_matrix.rs:606-624,
 * provekit/prover/src/whir_r1cs.rs:427-621) and of the Fiat-Shamir
 * transcript this project defines in place of spongefish (not vendored).
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* eval_eq (sumcheck.rs:125-153): r[0] selects the top half of the table */
static void eval_eq(const ofr* r, size_t m, ofr* out, size_t out_len, ofr scalar,
                    size_t subtree) {
  if (m == 0) {
    out[0] = ofr_add(out[0], scalar);
    return;
  }
  size_t half = subtree / 2;
  size_t left = out_len < half ? out_len : half;
  size_t right = out_len > half ? out_len - half : 0;
  ofr s1 = ofr_mul(scalar, r[0]);
  ofr s0 = ofr_sub(scalar, s1);
  eval_eq(r + 1, m - 1, out, left, s0, half);
  if (right) eval_eq(r + 1, m - 1, out + left, right, s1, half);
}

/* calculate_evaluations_over_boolean_hypercube_for_eq (sumcheck.rs:108-120) */
void or_eq_evals(const ofr* r, size_t m, size_t num_entries, ofr* out) {
  if (num_entries == 0) return;
  memset(out, 0, sizeof(ofr) * num_entries);
  eval_eq(r, m, out, num_entries, OFR_ONE, (size_t)1 << m);
}

/* calculate_eq (sumcheck.rs:182-188) */
ofr or_calculate_eq(const ofr* r, const ofr* alpha, size_t n) {
  ofr acc = OFR_ONE;
  for (size_t i = 0; i < n; i++) {
    ofr t = ofr_add(ofr_mul(r[i], alpha[i]),
                    ofr_mul(ofr_sub(OFR_ONE, r[i]), ofr_sub(OFR_ONE, alpha[i])));
    acc = ofr_mul(acc, t);
  }
  return acc;
}

/* HydratedSparseMatrix * &[Fr] (sparse_matrix.rs:606-624): CSR rows, row r
 * spans [row_starts[r], row_starts[r+1]) (the last row ends at nnz). */
void or_sparse_mul(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                   const uint32_t* cols, const ofr* vals, const ofr* z, ofr* out) {
  (void)num_cols;
  int nt = or_get_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
  for (size_t r = 0; r < num_rows; r++) {
    size_t b = row_starts[r], e = r + 1 < num_rows ? row_starts[r + 1] : nnz;
    ofr acc = {{0, 0, 0, 0}};
    for (size_t i = b; i < e; i++) acc = ofr_add(acc, ofr_mul(vals[i], z[cols[i]]));
    out[r] = acc;
  }
}

/* SparseMatrix::transpose (provekit/common/src/sparse_matrix.rs:409-449):
 * the entries (col, row, value) in (col, row) order; t_row_starts gets
 * num_cols entries.  A counting sort by column over the rows in order gives
 * that order (within a column the rows come out increasing).  Returns -1 if
 * a column index is out of range. */
int or_sparse_transpose(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                        const uint32_t* cols, const ofr* vals, uint32_t* t_row_starts,
                        uint32_t* t_cols, ofr* t_vals) {
  size_t* next = (size_t*)calloc(num_cols + 1, sizeof(size_t));
  if (!next) return -2;
  for (size_t i = 0; i < nnz; i++) {
    if (cols[i] >= num_cols) {
      free(next);
      return -1;
    }
    next[cols[i] + 1]++;
  }
  for (size_t c = 0; c < num_cols; c++) next[c + 1] += next[c];
  for (size_t c = 0; c < num_cols; c++) t_row_starts[c] = (uint32_t)next[c];
  for (size_t r = 0; r < num_rows; r++) {
    size_t b = row_starts[r], e = r + 1 < num_rows ? row_starts[r + 1] : nnz;
    for (size_t i = b; i < e; i++) {
      size_t k = next[cols[i]]++;
      t_cols[k] = (uint32_t)r;
      t_vals[k] = vals[i];
    }
  }
  free(next);
  return 0;
}

/* The R1CS map of run_zk_sumcheck_prover (whir_r1cs.rs:552-559) on one pair */
static void r1cs_map(ofr a0, ofr a1, ofr b0, ofr b1, ofr c0, ofr c1, ofr e0, ofr e1, ofr o[3]) {
  ofr f0 = ofr_mul(e0, ofr_sub(ofr_mul(a0, b0), c0));
  ofr ea = ofr_sub(ofr_add(e0, e0), e1);
  ofr aa = ofr_sub(ofr_add(a0, a0), a1);
  ofr bb = ofr_sub(ofr_add(b0, b0), b1);
  ofr cc = ofr_sub(ofr_add(c0, c0), c1);
  ofr fm1 = ofr_mul(ea, ofr_sub(ofr_mul(aa, bb), cc));
  ofr finf = ofr_mul(ofr_mul(ofr_sub(e1, e0), ofr_sub(a1, a0)), ofr_sub(b1, b0));
  o[0] = f0;
  o[1] = fm1;
  o[2] = finf;
}

/* sumcheck_fold_map_reduce with N = 4 MLEs [a, b, c, eq] and the R1CS map
 * (sumcheck.rs:14-102).  With a fold value the four quarters are folded in
 * place first (p0 += f (p2 - p0), p1 += f (p3 - p1)) and the map runs over
 * (p0, p1); the caller then keeps the first half.  Without, the map runs
 * over the two halves. */
void or_sumcheck_r1cs_round(ofr* a, ofr* b, ofr* c, ofr* eq, size_t size, const ofr* fold,
                            ofr out[3]) {
  ofr* m[4] = {a, b, c, eq};
  int nt = or_get_threads();
  ofr tot[3] = {{{0}}, {{0}}, {{0}}};
  size_t h = fold ? size / 4 : size / 2;
#pragma omp parallel num_threads(nt)
  {
    ofr loc[3] = {{{0}}, {{0}}, {{0}}};
#pragma omp for schedule(static)
    for (size_t i = 0; i < h; i++) {
      ofr lo[4], hi[4];
      for (int j = 0; j < 4; j++) {
        if (fold) {
          ofr* p = m[j];
          p[i] = ofr_add(p[i], ofr_mul(*fold, ofr_sub(p[i + 2 * h], p[i])));
          p[i + h] = ofr_add(p[i + h], ofr_mul(*fold, ofr_sub(p[i + 3 * h], p[i + h])));
        }
        lo[j] = m[j][i];
        hi[j] = m[j][i + h];
      }
      ofr o[3];
      r1cs_map(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], lo[3], hi[3], o);
      for (int k = 0; k < 3; k++) loc[k] = ofr_add(loc[k], o[k]);
    }
#pragma omp critical(or_sc)
    for (int k = 0; k < 3; k++) tot[k] = ofr_add(tot[k], loc[k]);
  }
  for (int k = 0; k < 3; k++) out[k] = tot[k];
}

/* eval_cubic_poly (sumcheck.rs:156-158) */
ofr or_eval_cubic(const ofr p[4], ofr x) {
  return ofr_add(p[0], ofr_mul(x, ofr_add(p[1], ofr_mul(x, ofr_add(p[2], ofr_mul(x, p[3]))))));
}

/* field primitives exported for tests */
ofr or_fr_mul(ofr a, ofr b) { return ofr_mul(a, b); }
ofr or_fr_add(ofr a, ofr b) { return ofr_add(a, b); }
ofr or_fr_sub(ofr a, ofr b) { return ofr_sub(a, b); }
