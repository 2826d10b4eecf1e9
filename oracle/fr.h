/* ORACLE — test infrastructure only.  This is the CPU restatement used to
 * check the HIP path; the product never links or calls it.
 *
 * BN254 scalar field Fr, as ark-ff 0.5 `Fp256<MontBackend<FrConfig, 4>>`
 * (third-party, pinned in the reference's Cargo.lock:231-234,301-304) keeps
 * it: 4 little-endian u64 limbs holding x*2^256 mod p, always fully reduced
 * (< p).  Multiplication is textbook CIOS Montgomery with 64-bit digits.
 * Semantics pinned by tests against Python big integers (closed form
 * a*b*2^-256 mod p) and by the reference's KATs for the callers.
 */
#ifndef PK_ORACLE_FR_H
#define PK_ORACLE_FR_H
#include <stdint.h>
#include <string.h>

typedef struct { uint64_t l[4]; } ofr;
typedef unsigned __int128 u128;

static const uint64_t OFR_P[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                                  0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t OFR_NP0 = 0xc2e1f593efffffffULL; /* -p^-1 mod 2^64 */
static const ofr OFR_ONE = {{0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL,
                             0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL}}; /* R mod p */
static const ofr OFR_R2 = {{0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL,
                            0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}}; /* R^2 mod p */

static inline int ofr_geq_p(const uint64_t a[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] > OFR_P[i]) return 1;
    if (a[i] < OFR_P[i]) return 0;
  }
  return 1;
}
static inline void ofr_sub_p(uint64_t a[4]) {
  u128 br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - OFR_P[i] - br;
    a[i] = (uint64_t)d;
    br = (d >> 64) ? 1 : 0;
  }
}
static inline ofr ofr_add(ofr a, ofr b) {
  ofr r;
  u128 c = 0;
  for (int i = 0; i < 4; i++) {
    c += (u128)a.l[i] + b.l[i];
    r.l[i] = (uint64_t)c;
    c >>= 64;
  }
  if (ofr_geq_p(r.l)) ofr_sub_p(r.l); /* a + b < 2p < 2^256 */
  return r;
}
static inline ofr ofr_sub(ofr a, ofr b) {
  ofr r;
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a.l[i] - b.l[i] - br;
    r.l[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
      c += (u128)r.l[i] + OFR_P[i];
      r.l[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  return r;
}
static inline ofr ofr_neg(ofr a) {
  ofr z = {{0, 0, 0, 0}};
  return ofr_sub(z, a);
}
/* a*b*2^-256 mod p, canonical. */
static inline ofr ofr_mul(ofr a, ofr b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    u128 c = 0;
    for (int j = 0; j < 4; j++) {
      c += (u128)a.l[j] * b.l[i] + t[j];
      t[j] = (uint64_t)c;
      c >>= 64;
    }
    c += t[4];
    t[4] = (uint64_t)c;
    t[5] = (uint64_t)(c >> 64);
    uint64_t m = t[0] * OFR_NP0;
    c = (u128)m * OFR_P[0] + t[0];
    c >>= 64;
    for (int j = 1; j < 4; j++) {
      c += (u128)m * OFR_P[j] + t[j];
      t[j - 1] = (uint64_t)c;
      c >>= 64;
    }
    c += t[4];
    t[3] = (uint64_t)c;
    t[4] = t[5] + (uint64_t)(c >> 64);
  }
  ofr r = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || ofr_geq_p(r.l)) ofr_sub_p(r.l);
  return r;
}
static inline ofr ofr_sqr(ofr a) { return ofr_mul(a, a); }
static inline int ofr_eq(ofr a, ofr b) { return memcmp(&a, &b, sizeof a) == 0; }
static inline int ofr_is_zero(ofr a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3]) == 0; }
/* integer (any 256-bit value) -> Montgomery form of (x mod p) */
static inline ofr ofr_from_int(ofr x) { return ofr_mul(x, OFR_R2); }
/* Montgomery form -> canonical integer */
static inline ofr ofr_to_int(ofr x) {
  ofr one = {{1, 0, 0, 0}};
  return ofr_mul(x, one);
}
static inline ofr ofr_from_u64(uint64_t v) {
  ofr x = {{v, 0, 0, 0}};
  return ofr_from_int(x);
}
static inline ofr ofr_pow(ofr b, uint64_t e) {
  ofr r = OFR_ONE;
  while (e) {
    if (e & 1) r = ofr_mul(r, b);
    b = ofr_mul(b, b);
    e >>= 1;
  }
  return r;
}
/* inverse by Fermat: a^(p-2) */
static inline ofr ofr_inv(ofr a) {
  static const uint64_t E[4] = {0x43e1f593efffffffULL, 0x2833e84879b97091ULL,
                                0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  ofr r = OFR_ONE;
  for (int i = 3; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      r = ofr_mul(r, r);
      if ((E[i] >> b) & 1) r = ofr_mul(r, a);
    }
  return r;
}
#endif
