/* ORACLE — test infrastructure only.  CPU restatement of the reference's
 * Poseidon2 permutation and sponge (poseidon2/src/permutation.rs,
 * poseidon2/src/hash.rs) and of the 2-to-1 Merkle step
 * `poseidon2_hash(&[left, right])`
 * (playground/passport-input-gen/src/commitment.rs:225-237).
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "poseidon2_params.h"

static ofr P2_RC[64][4];
static ofr P2_DIAG[4];
static int p2_ready = 0;

/* big-endian hex -> Montgomery form (poseidon2/src/constants.rs:11-27) */
static ofr from_hex(const char* h) {
  ofr x = {{0, 0, 0, 0}};
  size_t n = strlen(h);
  for (size_t i = 0; i < n; i++) {
    char ch = h[n - 1 - i];
    uint64_t v = (ch >= '0' && ch <= '9') ? (uint64_t)(ch - '0')
                 : (ch >= 'a' && ch <= 'f') ? (uint64_t)(ch - 'a' + 10)
                                            : (uint64_t)(ch - 'A' + 10);
    x.l[i / 16] |= v << (4 * (i % 16));
  }
  return ofr_from_int(x);
}

/* POSEIDON2_CONFIG (poseidon2/src/permutation.rs:28-68) */
static void p2_init(void) {
#pragma omp critical(or_p2)
  {
    if (!p2_ready) {
      memset(P2_RC, 0, sizeof P2_RC);
      for (int i = 0; i < 4; i++) P2_DIAG[i] = from_hex(OR_P2_DIAG[i]);
      for (int r = 0; r < 4; r++)
        for (int j = 0; j < 4; j++) P2_RC[r][j] = from_hex(OR_P2_FULL1[r][j]);
      for (int r = 0; r < 56; r++) P2_RC[4 + r][0] = from_hex(OR_P2_PARTIAL[r]);
      for (int r = 0; r < 4; r++)
        for (int j = 0; j < 4; j++) P2_RC[60 + r][j] = from_hex(OR_P2_FULL2[r][j]);
      p2_ready = 1;
    }
  }
}

static ofr sbox(ofr x) {
  ofr s = ofr_mul(x, x);
  return ofr_mul(ofr_mul(s, s), x);
}

/* matrix_multiplication_4x4 (permutation.rs:105-124) */
static void ext_mds(ofr s[4]) {
  ofr t0 = ofr_add(s[0], s[1]);
  ofr t1 = ofr_add(s[2], s[3]);
  ofr t2 = ofr_add(ofr_add(s[1], s[1]), t1);
  ofr t3 = ofr_add(ofr_add(s[3], s[3]), t0);
  ofr t4 = ofr_add(t1, t1);
  t4 = ofr_add(t4, t4);
  t4 = ofr_add(t4, t3);
  ofr t5 = ofr_add(t0, t0);
  t5 = ofr_add(t5, t5);
  t5 = ofr_add(t5, t2);
  ofr t6 = ofr_add(t3, t5);
  ofr t7 = ofr_add(t2, t4);
  s[0] = t6;
  s[1] = t5;
  s[2] = t7;
  s[3] = t4;
}

/* internal_m_multiplication (permutation.rs:126-131) */
static void int_mds(ofr s[4]) {
  ofr sum = ofr_add(ofr_add(s[0], s[1]), ofr_add(s[2], s[3]));
  for (int i = 0; i < 4; i++) s[i] = ofr_add(ofr_mul(s[i], P2_DIAG[i]), sum);
}

/* Poseidon2::permutation (permutation.rs:138-170) */
void or_p2_permutation(const ofr in[4], ofr out[4]) {
  if (!p2_ready) p2_init();
  ofr s[4] = {in[0], in[1], in[2], in[3]};
  ext_mds(s);
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < 4; i++) s[i] = sbox(ofr_add(s[i], P2_RC[r][i]));
    ext_mds(s);
  }
  for (int r = 4; r < 60; r++) {
    s[0] = sbox(ofr_add(s[0], P2_RC[r][0]));
    int_mds(s);
  }
  for (int r = 60; r < 64; r++) {
    for (int i = 0; i < 4; i++) s[i] = sbox(ofr_add(s[i], P2_RC[r][i]));
    ext_mds(s);
  }
  for (int i = 0; i < 4; i++) out[i] = s[i];
}

/* poseidon2_hash (poseidon2/src/hash.rs:25-55): rate 3, IV = len * 2^64 in
 * lane 3, duplex on every full cache, one final permutation. */
ofr or_p2_hash(const ofr* inputs, size_t n) {
  ofr z = {{0, 0, 0, 0}};
  ofr two64 = ofr_from_int((ofr){{0, 1, 0, 0}});
  ofr st[4] = {z, z, z, ofr_mul(ofr_from_u64(n), two64)};
  ofr cache[3] = {z, z, z};
  size_t cs = 0;
  for (size_t i = 0; i < n; i++) {
    if (cs == 3) {
      for (int j = 0; j < 3; j++) st[j] = ofr_add(st[j], cache[j]);
      cache[0] = cache[1] = cache[2] = z;
      cs = 0;
      or_p2_permutation(st, st);
    }
    cache[cs++] = inputs[i];
  }
  for (size_t j = 0; j < cs; j++) st[j] = ofr_add(st[j], cache[j]);
  or_p2_permutation(st, st);
  return st[0];
}

/* Level-by-level 2-to-1 reduction with poseidon2_hash([l, r]) as in
 * compute_merkle_root (commitment.rs:225-237).  nodes: 2n-1, root last. */
void or_p2_merkle_levels(const ofr* leaves, size_t n, ofr* nodes) {
  if (!p2_ready) p2_init();
  memcpy(nodes, leaves, sizeof(ofr) * n);
  size_t off = 0;
  int nt = or_get_threads();
  while (n > 1) {
    size_t h = n / 2;
    const ofr* lvl = nodes + off;
    ofr* up = nodes + off + n;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t i = 0; i < h; i++) up[i] = or_p2_hash(lvl + 2 * i, 2);
    off += n;
    n = h;
  }
}
