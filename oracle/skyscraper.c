/* ORACLE — test infrastructure only.  CPU restatement of Skyscraper-v2 over
 * BN254 Fr (skyscraper/core) and of the whir HashEngine plugin
 * (provekit/common/src/skyscraper/whir.rs), plus the commit layout this
 * project fixes for one polynomial (leaves of consecutive codeword elements,
 * binary 2-to-1 tree).
 *
 * Two forms of the compression are kept:
 *  - the spec form (skyscraper/core/src/reference.rs:41-98): values mod p,
 *    checked against the reference's Sage known-answer vectors;
 *  - the fast form (skyscraper/core/src/generic.rs:77-102 with
 *    bn254_multiplier::scalar_sqr, scalar.rs:12-70): partial reductions, used
 *    as the timed CPU baseline; tests check it equals the spec form.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef unsigned __int128 u128;

/* k*p, k = 0..5 (skyscraper/core/src/constants.rs:8-15 lists the same
 * multiples; they are just the integers k*p). */
static const uint64_t PM[6][4] = {
    {0, 0, 0, 0},
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    {0x87c3eb27e0000002ULL, 0x5067d090f372e122ULL, 0x70a08b6d0302b0baULL, 0x60c89ce5c2634053ULL},
    {0xcba5e0bbd0000003ULL, 0x789bb8d96d2c51b3ULL, 0x28f0d12384840917ULL, 0x912ceb58a394e07dULL},
    {0x0f87d64fc0000004ULL, 0xa0cfa121e6e5c245ULL, 0xe14116da06056174ULL, 0xc19139cb84c680a6ULL},
    {0x5369cbe3b0000005ULL, 0xc903896a609f32d6ULL, 0x99915c908786b9d1ULL, 0xf1f5883e65f820d0ULL},
};

/* Skyscraper round constants for BN254 Fr, t = 1, as integers < p
 * (skyscraper/core/src/constants.rs:30-49; the Sage reference's
 * Sky_BN254_1.rcons).  rc[0] = rc[17] = 0. */
static const uint64_t RC[18][4] = {
    {0, 0, 0, 0},
    {0x903c4324270bd744ULL, 0x873125f708a7d269ULL, 0x081dd27906c83855ULL, 0x276b1823ea6d7667ULL},
    {0x7ac8edbb4b378d71ULL, 0xe29d79f3d99e2cb7ULL, 0x751417914c1a5a18ULL, 0x0cf02bd758a484a6ULL},
    {0xfa7adc6769e5bc36ULL, 0x1c3f8e297cca387dULL, 0x0eb7730d63481db0ULL, 0x25b0e03f18ede544ULL},
    {0x57847e652f03cfb7ULL, 0x33440b9668873404ULL, 0x955a32e849af80bcULL, 0x002882fcbe14ae70ULL},
    {0x979231396257d4d7ULL, 0x29989c3e1b37d3c1ULL, 0x12ef02b47f1277baULL, 0x039ad8571e2b7a9cULL},
    {0xb5b48465abbb7887ULL, 0xa72a6bc5e6ba2d2bULL, 0x4cd48043712f7b29ULL, 0x1142d5410fc1fc1aULL},
    {0x7ab2c156059075d3ULL, 0x17cb3594047999b2ULL, 0x44f2c93598f289f7ULL, 0x1d78439f69bc0becULL},
    {0x05d7a965138b8edbULL, 0x36ef35a3d55c48b1ULL, 0x8ddfb8a1ac6f1628ULL, 0x258588a508f4ff82ULL},
    {0x1596fb9afccb49e9ULL, 0x9a7367d69a09a95bULL, 0x9bc43f6984e4c157ULL, 0x13087879d2f514feULL},
    {0x295ccd233b4109faULL, 0xe1d72f89ed868012ULL, 0x2e9e1eea4bc88a8eULL, 0x17dadee898c45232ULL},
    {0x9a8590b4aa1f486fULL, 0xb75834b430e9130eULL, 0xb8e90b1034d5de31ULL, 0x295c6d1546e7f4a6ULL},
    {0x850adcb74c6eb892ULL, 0x07699ef305b92fc3ULL, 0x4ef96a2ba1720f2dULL, 0x1288ca0e1d3ed446ULL},
    {0x01960f9349d1b5eeULL, 0x8ccad30769371c69ULL, 0xe5c81e8991c98662ULL, 0x17563b4d1ae023f3ULL},
    {0x6ba01e9476b32917ULL, 0xa1cb0a3add977bc9ULL, 0x86815a945815f030ULL, 0x2869043be91a1eeaULL},
    {0x81776c885511d976ULL, 0x7475d34f47f414e7ULL, 0x5d090056095d96cfULL, 0x14941f0aff59e79aULL},
    {0xbc40b4fd8fc8c034ULL, 0xbb7142c3cce4fd48ULL, 0x318356758a39005aULL, 0x1ce337a190f4379fULL},
    {0, 0, 0, 0},
};

/* ------------------------------------------------------------ 256-bit ints */
static void add4(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) {
  u128 c = 0;
  for (int i = 0; i < 4; i++) {
    c += (u128)a[i] + b[i];
    r[i] = (uint64_t)c;
    c >>= 64;
  }
}
static int sub4(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - b[i] - br;
    r[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  return (int)br;
}
static int lt4(const uint64_t a[4], const uint64_t b[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] < b[i]) return 1;
    if (a[i] > b[i]) return 0;
  }
  return 0;
}
/* any 256-bit -> [0, p) */
static void reduce_full(const uint64_t x[4], uint64_t r[4]) {
  uint64_t t[4];
  memcpy(t, x, 32);
  while (!lt4(t, PM[1])) sub4(t, PM[1], t);
  memcpy(r, t, 32);
}

/* ------------------------------------------------------------ spec form */

uint8_t or_sky_sbox(uint8_t v) {
  /* skyscraper/core/src/reference.rs:96-98 (Table 3 of the paper) */
  uint8_t nv = (uint8_t)~v;
  uint8_t r1 = (uint8_t)((nv << 1) | (nv >> 7));
  uint8_t r2 = (uint8_t)((v << 2) | (v >> 6));
  uint8_t r3 = (uint8_t)((v << 3) | (v >> 5));
  uint8_t x = (uint8_t)(v ^ (r1 & r2 & r3));
  return (uint8_t)((x << 1) | (x >> 7));
}

/* reference.rs:80-94: canonical LE bytes, swap the 16-byte halves, S-box each
 * byte, read back mod p. */
static void bar_spec(const uint64_t x[4], uint64_t r[4]) {
  uint8_t b[32], o[32];
  memcpy(b, x, 32);
  for (int i = 0; i < 32; i++) o[i] = or_sky_sbox(b[(i + 16) & 31]);
  uint64_t t[4];
  memcpy(t, o, 32);
  reduce_full(t, r);
}

/* field ops on canonical integers (values, not Montgomery): the square times
 * sigma^-1 = 2^-256 is exactly one Montgomery product. */
static void fadd(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) {
  ofr x, y;
  memcpy(x.l, a, 32);
  memcpy(y.l, b, 32);
  ofr z = ofr_add(x, y);
  memcpy(r, z.l, 32);
}
static void sq_sigma(const uint64_t a[4], uint64_t r[4]) {
  ofr x;
  memcpy(x.l, a, 32);
  ofr z = ofr_mul(x, x);
  memcpy(r, z.l, 32);
}

/* one Feistel step: r += F(l) + rc[round]; swap */
static void step(int round, int use_bar, uint64_t l[4], uint64_t r[4]) {
  uint64_t f[4], t[4];
  if (use_bar)
    bar_spec(l, f);
  else
    sq_sigma(l, f);
  fadd(f, RC[round], t);
  fadd(r, t, t);
  memcpy(r, l, 32);
  memcpy(l, t, 32);
}

/* reference.rs:63-69 */
void or_sky_ss(int round, const uint64_t l_in[4], const uint64_t r_in[4], uint64_t ol[4],
               uint64_t orr[4]) {
  uint64_t l[4], r[4];
  reduce_full(l_in, l);
  reduce_full(r_in, r);
  step(round, 0, l, r);
  step(round + 1, 0, l, r);
  memcpy(ol, l, 32);
  memcpy(orr, r, 32);
}
/* reference.rs:72-78 */
void or_sky_bb(int round, const uint64_t l_in[4], const uint64_t r_in[4], uint64_t ol[4],
               uint64_t orr[4]) {
  uint64_t l[4], r[4];
  reduce_full(l_in, l);
  reduce_full(r_in, r);
  step(round, 1, l, r);
  step(round + 1, 1, l, r);
  memcpy(ol, l, 32);
  memcpy(orr, r, 32);
}
/* reference.rs:49-60: ss(0) ss(2) ss(4) bb(6) ss(8) bb(10) ss(12) ss(14) ss(16) */
void or_sky_permute(const uint64_t l_in[4], const uint64_t r_in[4], uint64_t ol[4],
                    uint64_t orr[4]) {
  uint64_t l[4], r[4];
  reduce_full(l_in, l);
  reduce_full(r_in, r);
  for (int round = 0; round < 18; round++) {
    int use_bar = (round == 6 || round == 7 || round == 10 || round == 11);
    step(round, use_bar, l, r);
  }
  memcpy(ol, l, 32);
  memcpy(orr, r, 32);
}
/* reference.rs:41-46: permute(l, r).0 + l */
void or_sky_compress_spec(const uint64_t l_in[4], const uint64_t r_in[4], uint64_t out[4]) {
  uint64_t l[4], pl[4], pr[4];
  reduce_full(l_in, l);
  or_sky_permute(l_in, r_in, pl, pr);
  fadd(pl, l, out);
}

/* ------------------------------------------------------------ fast form */
/* 2^-64k mod p for k = 1, 2, 3 (the I1, I2, I3 of
 * skyscraper/bn254-multiplier/src/constants.rs:71-89) */
static const uint64_t I1[4] = {0x2d3e8053e396ee4dULL, 0xca478dbeab3c92cdULL,
                               0xb2d8f06f77f52a93ULL, 0x24d6ba07f7aa8f04ULL};
static const uint64_t I2[4] = {0x18ee753c76f9dc6fULL, 0x54ad7e14a329e70fULL,
                               0x2b16366f4f7684dfULL, 0x133100d71fdf3579ULL};
static const uint64_t I3[4] = {0x9bacb016127cbe4eULL, 0x0b2051fa31944124ULL,
                               0xb064eea46091c76cULL, 0x2b062aaa49f80c7dULL};
static const uint64_t MU0 = 0xc2e1f593efffffffULL; /* -p^-1 mod 2^64 */

/* The part of bn254_multiplier::scalar_sqr / scalar_mul after the
 * schoolbook product (scalar.rs:46-69, :103-131): fold the three low words
 * of t with 2^-192, 2^-128, 2^-64 (the I3, I2, I1 products, 5 words each,
 * summed mod 2^320 as `addv` does), one Montgomery step on the low word, keep
 * words 1..4, then reduce_ct (utils.rs:52-56): subtract 2p if bit 255 is set. */
static void log_jump_reduce(const uint64_t t[8], uint64_t out[4]) {
  uint64_t s[5] = {t[3], t[4], t[5], t[6], t[7]};
  const uint64_t* I[3] = {I3, I2, I1};
  for (int k = 0; k < 3; k++) {
    uint64_t r5[5];
    u128 c = 0;
    for (int j = 0; j < 4; j++) {
      c += (u128)t[k] * I[k][j];
      r5[j] = (uint64_t)c;
      c >>= 64;
    }
    r5[4] = (uint64_t)c;
    c = 0;
    for (int j = 0; j < 5; j++) {
      c += (u128)s[j] + r5[j];
      s[j] = (uint64_t)c;
      c >>= 64;
    }
  }
  uint64_t m = MU0 * s[0];
  u128 c = 0;
  for (int j = 0; j < 4; j++) {
    c += (u128)m * PM[1][j] + s[j];
    s[j] = (uint64_t)c;
    c >>= 64;
  }
  s[4] += (uint64_t)c;
  uint64_t r[4] = {s[1], s[2], s[3], s[4]};
  if (r[3] >> 63) sub4(r, PM[2], r);
  memcpy(out, r, 32);
}

static void schoolbook(const uint64_t a[4], const uint64_t b[4], uint64_t t[8]) {
  memset(t, 0, 64);
  for (int i = 0; i < 4; i++) {
    u128 c = 0;
    for (int j = 0; j < 4; j++) {
      c += (u128)a[i] * b[j] + t[i + j];
      t[i + j] = (uint64_t)c;
      c >>= 64;
    }
    t[i + 4] = (uint64_t)c;
  }
}

/* bn254_multiplier::scalar_sqr (scalar.rs:12-70): the full product a*a (all
 * 16 digit products, as the reference forms them), then log_jump_reduce.
 * Input < 2p; output < 2^256 - 2p, congruent to a^2 2^-256, not canonical. */
void or_scalar_sqr(const uint64_t a[4], uint64_t out[4]) {
  uint64_t t[8];
  schoolbook(a, a, t);
  log_jump_reduce(t, out);
}

/* bn254_multiplier::scalar_mul (scalar.rs:73-132): the same on a*b; inputs
 * up to 2^256 - 2p (the reference's safe_bn254_montgomery_input). */
void or_scalar_mul(const uint64_t a[4], const uint64_t b[4], uint64_t out[4]) {
  uint64_t t[8];
  schoolbook(a, b, t);
  log_jump_reduce(t, out);
}

/* reduce_partial (skyscraper/core/src/reduce.rs:35-41) */
static void reduce_partial(uint64_t x[4]) {
  uint64_t q = x[3] / (PM[1][3] + 1);
  sub4(x, PM[q], x);
}
/* reduce_partial_add_rc (reduce.rs:48-55): x - (q*p - rc) */
static void reduce_partial_add_rc(uint64_t x[4], int rc) {
  uint64_t q = x[3] / (PM[1][3] + 1);
  uint64_t k[4];
  sub4(PM[q], RC[rc], k); /* wraps for q = 0, as MODULUS_N_MINUS_RC[0] does */
  sub4(x, k, x);
}
/* bar (skyscraper/core/src/bar.rs:15-31) on a value < 2p */
static void bar_fast(const uint64_t x[4], uint64_t r[4]) {
  uint64_t t[4];
  if (sub4(x, PM[1], t)) memcpy(t, x, 32); /* reduce_1 */
  uint64_t s[4] = {t[2], t[3], t[0], t[1]};
  uint8_t* b = (uint8_t*)s;
  for (int i = 0; i < 32; i++) b[i] = or_sky_sbox(b[i]);
  reduce_partial(s);
  memcpy(r, s, 32);
}

/* generic::compress (skyscraper/core/src/generic.rs:77-102) */
void or_sky_compress(const uint64_t l_in[4], const uint64_t r_in[4], uint64_t out[4]) {
  uint64_t l[4], r[4], t[4], f[4], nl[4];
  memcpy(l, l_in, 32);
  memcpy(r, r_in, 32);
  reduce_partial(l);
  reduce_partial(r);
  memcpy(t, l, 32);
  for (int round = 0; round < 17; round++) {
    int use_bar = (round == 6 || round == 7 || round == 10 || round == 11);
    if (use_bar)
      bar_fast(l, f);
    else
      or_scalar_sqr(l, f);
    add4(r, f, nl);
    if (round == 0)
      reduce_partial(nl);
    else
      reduce_partial_add_rc(nl, round);
    memcpy(r, l, 32);
    memcpy(l, nl, 32);
  }
  or_scalar_sqr(l, f);
  add4(r, f, nl);
  add4(nl, t, nl);
  reduce_partial(nl); /* reduce = reduce_1(reduce_partial(x)) */
  if (!sub4(nl, PM[1], t)) memcpy(nl, t, 32);
  memcpy(out, nl, 32);
}

/* simple::compress_many (skyscraper/core/src/simple.rs:3-9) */
void or_sky_compress_many(const uint8_t* messages, uint8_t* hashes, size_t count) {
  int nt = or_get_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
  for (size_t i = 0; i < count; i++) {
    uint64_t l[4], r[4], o[4];
    memcpy(l, messages + 64 * i, 32);
    memcpy(r, messages + 64 * i + 32, 32);
    or_sky_compress(l, r, o);
    memcpy(hashes + 32 * i, o, 32);
  }
}

/* SkyscraperHashEngine::hash_many (provekit/common/src/skyscraper/whir.rs:42-100):
 * size 32 copies, size 64 compresses, larger sizes left-fold the 32-byte
 * chunks.  Returns -1 on an unsupported size. */
int or_sky_hash_many(size_t size, const uint8_t* input, size_t count, uint8_t* output) {
  if (size == 0 || size % 32) return -1;
  if (size == 32) {
    memcpy(output, input, 32 * count);
    return 0;
  }
  size_t chunks = size / 32;
  int nt = or_get_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
  for (size_t i = 0; i < count; i++) {
    const uint8_t* m = input + i * size;
    uint64_t acc[4], x[4];
    memcpy(acc, m, 32);
    for (size_t k = 1; k < chunks; k++) {
      memcpy(x, m + 32 * k, 32);
      or_sky_compress(acc, x, acc);
    }
    memcpy(output + 32 * i, acc, 32);
  }
  return 0;
}

/* Binary tree over leaf digests, nodes laid out level by level. */
void or_merkle_tree(const uint8_t* leaves, size_t n, uint8_t* nodes) {
  memcpy(nodes, leaves, 32 * n);
  size_t off = 0;
  int nt = or_get_threads();
  while (n > 1) {
    const uint8_t* lvl = nodes + 32 * off;
    uint8_t* up = nodes + 32 * (off + n);
    size_t h = n / 2;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t i = 0; i < h; i++) {
      uint64_t l[4], r[4], o[4];
      memcpy(l, lvl + 64 * i, 32);
      memcpy(r, lvl + 64 * i + 32, 32);
      or_sky_compress(l, r, o);
      memcpy(up + 32 * i, o, 32);
    }
    off += n;
    n = h;
  }
}

/* One-polynomial commit with the layout this project fixes: zero-pad n
 * coefficients to n*blowup, evaluate by RSFr's ntt_nr (bit-reversed order,
 * num_groups = blowup since the padded coefficients are the replicated
 * coset chunk), optionally on the coset 5*H (coefficients pre-scaled by 5^j),
 * hash leaves of leaf_elems consecutive codeword elements (their in-memory
 * Montgomery bytes) with hash_many, then the binary tree. */
int or_commit(const ofr* coeffs, size_t n, size_t blowup, int coset, size_t leaf_elems,
              ofr* codeword, uint8_t* nodes) {
  size_t N = n * blowup;
  if (N == 0 || (N & (N - 1)) || (n & (n - 1)) || N % leaf_elems) return -1;
  ofr* msg = (ofr*)malloc(sizeof(ofr) * n);
  memcpy(msg, coeffs, sizeof(ofr) * n);
  if (coset) {
    ofr g = ofr_from_u64(5), pw = OFR_ONE;
    for (size_t j = 0; j < n; j++) {
      msg[j] = ofr_mul(msg[j], pw);
      pw = ofr_mul(pw, g);
    }
  }
  int rc = or_rs_interleaved_encode(msg, 1, n, NULL, 0, N, codeword);
  free(msg);
  if (rc) return rc;
  size_t leaves = N / leaf_elems;
  uint8_t* ld = (uint8_t*)malloc(32 * leaves);
  or_sky_hash_many(32 * leaf_elems, (const uint8_t*)codeword, leaves, ld);
  or_merkle_tree(ld, leaves, nodes);
  free(ld);
  return 0;
}

/* ceil(log2(n)), 0 for n <= 1: next_power_of_two (provekit/common/src/utils/mod.rs:71-79) */
static size_t ceil_log2(size_t n) {
  size_t k = 0;
  while (((size_t)1 << k) < n) k++;
  return k;
}

/* WhirR1CSScheme::new_from_dimensions (provekit/r1cs-compiler/src/whir_r1cs.rs:55-84):
 * m = max(log w1, log w2, 13), m0 = max(log constraints, 1), and one more
 * variable when w1's zero padding has no room for the 4 m0 blinding
 * coefficients. */
void or_whir_r1cs_params(size_t num_witnesses, size_t num_constraints, size_t w1_size, size_t* m,
                         size_t* m0) {
  size_t w2 = num_witnesses - w1_size;
  size_t mr = ceil_log2(w1_size);
  if (ceil_log2(w2) > mr) mr = ceil_log2(w2);
  if (mr < 13) mr = 13;
  size_t z = ceil_log2(num_constraints);
  if (z < 1) z = 1;
  if (((size_t)1 << mr) - w1_size < 4 * z) mr++;
  *m = mr;
  *m0 = z;
}

/* The witness shaping of WhirR1CSScheme::commit (provekit/prover/src/whir_r1cs.rs:103-126):
 * pad_to_power_of_two (utils/mod.rs:84-91), zero-extend to 2^num_vars, and
 * (g != NULL, the w1 commitment) write the m0 x 4 blinding coefficients at
 * offset = witness_size.  out: 2^num_vars elements.  Returns -1 if the witness
 * exceeds 2^num_vars ("R1CS witness length exceeds scheme capacity"), -2 if
 * the coefficients run past the end (the reference's index panic). */
int or_whir_r1cs_shape(const ofr* witness, size_t witness_size, size_t num_vars, const ofr* g,
                       size_t m0, ofr* out) {
  size_t target = (size_t)1 << num_vars;
  if (witness_size > target) return -1;
  size_t padded = (size_t)1 << ceil_log2(witness_size);
  if (padded < target) padded = target;
  if (g && witness_size + 4 * m0 > padded) return -2;
  memset(out, 0, sizeof(ofr) * padded);
  memcpy(out, witness, sizeof(ofr) * witness_size);
  if (g)
    for (size_t i = 0; i < m0; i++)
      for (size_t j = 0; j < 4; j++) out[witness_size + 4 * i + j] = g[4 * i + j];
  return 0;
}
