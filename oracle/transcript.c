/* ORACLE — test infrastructure only.
 *
 * The Fiat-Shamir transcript is spongefish in the reference (git rev fcc277f,
 * not vendored, so its challenge bytes cannot be reproduced).  This project
 * defines its own transcript from the pinned Skyscraper compression C, shaped
 * so that a device can run it with few sequential compressions; this file
 * restates it independently of the product for the parity tests:
 *   init(seed):     s = C([seed, 0, 0, 0], [TAG0, TAG1, 0, 0])
 *   absorb1(x):     s = C(s, x)
 *   squeeze(k):     out_j = C(s, [j + 1, 0, 0, 0]) for j < k, then
 *                   s = C(s, [0, 0, 0, 0]); a challenge is the Fr whose
 *                   Montgomery limbs are out_j (C reduces, so out_j < p)
 *   round(c0..3):   s = C(C(s, c0), C(c2, c3)), and the new state is the
 *                   round's challenge.  c1 is not absorbed: the verifier
 *                   derives it from the claim (c1 = S - 2 c0 - c2 - c3), and
 *                   the claim is fixed by the transcript so far.  Two
 *                   dependent compressions per round.
 * Field elements are absorbed as their 4 Montgomery limbs.
 * The ZK sumcheck driver restates run_zk_sumcheck_prover
 * (provekit/prover/src/whir_r1cs.rs:519-621) with
 * compute_blinding_coefficients_for_round (:427-492) and
 * sum_over_hypercube (:494-501).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define TAG0 0x70726f76656b6974ULL /* "provekit" */
#define TAG1 0x73756d636865636bULL /* "sumcheck" */

typedef struct {
  uint64_t s[4];
} or_transcript;

static void tr_init(or_transcript* t, uint64_t seed) {
  uint64_t l[4] = {seed, 0, 0, 0}, r[4] = {TAG0, TAG1, 0, 0};
  or_sky_compress(l, r, t->s);
}
static void tr_absorb1(or_transcript* t, ofr x) { or_sky_compress(t->s, x.l, t->s); }
static void tr_round(or_transcript* t, const ofr c[4]) {
  uint64_t h1[4], h2[4];
  or_sky_compress(t->s, c[0].l, h1);
  or_sky_compress(c[2].l, c[3].l, h2);
  or_sky_compress(h1, h2, t->s);
}
static void tr_squeeze(or_transcript* t, size_t k, ofr* out) {
  for (size_t j = 0; j < k; j++) {
    uint64_t c[4] = {j + 1, 0, 0, 0};
    or_sky_compress(t->s, c, out[j].l);
  }
  static const uint64_t z[4] = {0, 0, 0, 0};
  or_sky_compress(t->s, z, t->s);
}

/* Planted mistakes for the tests that show the verifier's checks catch them
 * (0 = none): 1 prefix multiplier doubled once too often, 2 suffix
 * multiplier equal to the prefix multiplier (not half of it), 3 rho * g
 * subtracted from c0 instead of added, 4 blinding_eval without the last
 * univariate.  Test infrastructure: set only by tests/test_sumcheck.py. */
static int g_fault;
void or_sumcheck_plant_fault(int kind) { g_fault = kind; }

/* compute_blinding_coefficients_for_round (whir_r1cs.rs:427-492) */
static void blinding_coeffs(const ofr* g, size_t n, size_t compute_for, const ofr* alphas,
                            ofr out[4]) {
  int all_fixed = 0;
  if (compute_for == n) {
    all_fixed = 1;
    compute_for = n - 1;
  }
  ofr zero = {{0}};
  ofr prefix = zero;
  for (size_t i = 0; i < compute_for; i++) prefix = ofr_add(prefix, or_eval_cubic(g + 4 * i, alphas[i]));
  ofr suffix = zero;
  for (size_t i = compute_for + 1; i < n; i++)
    suffix = ofr_add(suffix, ofr_add(or_eval_cubic(g + 4 * i, zero), or_eval_cubic(g + 4 * i, OFR_ONE)));
  ofr pm = OFR_ONE;
  for (size_t i = 0; i < n - 1 - compute_for; i++) pm = ofr_add(pm, pm);
  if (g_fault == 1) pm = ofr_add(pm, pm);
  ofr two = ofr_add(OFR_ONE, OFR_ONE);
  ofr sm = g_fault == 2 ? pm : ofr_mul(pm, ofr_inv(two));
  ofr other = ofr_add(ofr_mul(pm, prefix), ofr_mul(sm, suffix));
  const ofr* cur = g + 4 * compute_for;
  ofr c[4] = {ofr_add(ofr_mul(pm, cur[0]), other), ofr_mul(pm, cur[1]), ofr_mul(pm, cur[2]),
              ofr_mul(pm, cur[3])};
  if (all_fixed) {
    out[0] = or_eval_cubic(c, alphas[compute_for]);
    out[1] = out[2] = out[3] = zero;
    return;
  }
  for (int k = 0; k < 4; k++) out[k] = c[k];
}

/* sum_over_hypercube (whir_r1cs.rs:494-501) */
static ofr sum_over_hypercube(const ofr* g, size_t n) {
  ofr c[4];
  blinding_coeffs(g, n, 0, NULL, c);
  ofr zero = {{0}};
  return ofr_add(or_eval_cubic(c, zero), or_eval_cubic(c, OFR_ONE));
}

/* The blinding evaluation of run_zk_sumcheck_prover (whir_r1cs.rs:613-618):
 * <expand_powers::<4>(alpha), w1[offset .. offset + 4m]>
 * (provekit/common/src/prefix_covector.rs:156-166), where the commit wrote
 * the univariates g (m x 4, row-major) at w1[offset..] (whir_r1cs.rs:118-124),
 * so the dot product is sum_i sum_j alpha_i^j g_i[j]. */
static ofr blinding_eval(const ofr* g, size_t m, const ofr* alphas) {
  ofr acc = {{0}};
  size_t n = g_fault == 4 && m ? m - 1 : m;
  for (size_t i = 0; i < n; i++) {
    ofr pw = OFR_ONE;
    for (int j = 0; j < 4; j++) {
      acc = ofr_add(acc, ofr_mul(pw, g[4 * i + j]));
      pw = ofr_mul(pw, alphas[i]);
    }
  }
  return acc;
}

/* Full sumcheck over eq(r, x) * (a(x) b(x) - c(x)), size = 2^m.
 * blinding != NULL: reference mode (whir_r1cs.rs:519-621): absorb sum_g,
 *   squeeze rho, claimed sum rho * sum_g, blinding g (m x 4) mixed in; after
 *   the last round blinding_eval is computed and absorbed.  Needs m >= 1 (the
 *   reference's compute_for = n - 1 underflows at m = 0).
 * blinding == NULL: plain mode: rho = 0, the claimed sum is computed as
 *   sum_x eq(r,x)(a b - c) and absorbed in place of sum_g; no blinding_eval.
 * a, b, c are consumed (folded in place; 2 elements each when m = 0, the
 *   reference's pad_to_pow2_len_min2).  coeffs: m x 4, alphas: m,
 * r_out: m (the eq point), final_claim: the last cubic at alpha_{m-1}.
 * extra (may be NULL): [rho, sum_g, blinding_eval, transcript state at the end].
 * Returns 0, -1 if the round identity c0+c0+c1+c2+c3 = S fails, -2 for a
 * blinding polynomial with m = 0. */
int or_sumcheck_prove(ofr* a, ofr* b, ofr* c, size_t m, const ofr* blinding, uint64_t seed,
                      ofr* coeffs, ofr* alphas, ofr* r_out, ofr* final_claim, ofr* extra) {
  if (blinding && m == 0) return -2;
  size_t size = (size_t)1 << m;
  if (size < 2) size = 2;
  or_transcript tr;
  tr_init(&tr, seed);
  if (m) tr_squeeze(&tr, m, r_out);
  ofr* eq = (ofr*)calloc(size, sizeof(ofr));
  or_eq_evals(r_out, m, (size_t)1 << m, eq);
  ofr zero = {{0}};
  ofr rho = zero, sg = zero, S;
  if (blinding) {
    sg = sum_over_hypercube(blinding, m);
    tr_absorb1(&tr, sg);
    tr_squeeze(&tr, 1, &rho);
    S = ofr_mul(rho, sg);
  } else {
    S = zero;
    for (size_t i = 0; i < size; i++)
      S = ofr_add(S, ofr_mul(eq[i], ofr_sub(ofr_mul(a[i], b[i]), c[i])));
    tr_absorb1(&tr, S);
  }
  ofr half = ofr_inv(ofr_add(OFR_ONE, OFR_ONE));
  int ok = 0;
  size_t cur = size;
  for (size_t idx = 0; idx < m; idx++) {
    ofr h[3];
    const ofr* fold = idx ? &alphas[idx - 1] : NULL;
    or_sumcheck_r1cs_round(a, b, c, eq, cur, fold, h);
    if (fold) cur /= 2;
    ofr g[4] = {zero, zero, zero, zero};
    if (blinding) blinding_coeffs(blinding, m, idx, alphas, g);
    ofr* co = coeffs + 4 * idx;
    co[0] = g_fault == 3 ? ofr_sub(h[0], ofr_mul(rho, g[0])) : ofr_add(h[0], ofr_mul(rho, g[0]));
    ofr g_m1 = ofr_sub(ofr_add(ofr_sub(g[0], g[1]), g[2]), g[3]);
    ofr at_m1 = ofr_add(h[1], ofr_mul(rho, g_m1));
    ofr three_c0 = ofr_add(ofr_add(co[0], co[0]), co[0]);
    co[2] = ofr_mul(half, ofr_sub(ofr_add(S, at_m1), three_c0));
    co[3] = ofr_add(h[2], ofr_mul(rho, g[3]));
    co[1] = ofr_sub(ofr_sub(ofr_sub(ofr_sub(S, co[0]), co[0]), co[3]), co[2]);
    ofr chk = ofr_add(ofr_add(ofr_add(ofr_add(co[0], co[0]), co[1]), co[2]), co[3]);
    if (!ofr_eq(chk, S)) ok = -1;
    tr_round(&tr, co);
    memcpy(alphas[idx].l, tr.s, sizeof tr.s);
    S = or_eval_cubic(co, alphas[idx]);
  }
  *final_claim = S;
  ofr be = zero;
  if (blinding) {
    be = blinding_eval(blinding, m, alphas);
    tr_absorb1(&tr, be);
  }
  if (extra) {
    extra[0] = rho;
    extra[1] = sg;
    extra[2] = be;
    memcpy(extra[3].l, tr.s, sizeof tr.s);
  }
  free(eq);
  return ok;
}

/* transcript primitives exported for tests: absorb1 each element, then
 * squeeze n_sq challenges at once */
void or_transcript_run(uint64_t seed, const ofr* absorbs, size_t n_abs, ofr* squeezes,
                       size_t n_sq) {
  or_transcript t;
  tr_init(&t, seed);
  for (size_t i = 0; i < n_abs; i++) tr_absorb1(&t, absorbs[i]);
  if (n_sq) tr_squeeze(&t, n_sq, squeezes);
}
