// Host-side BN254 Fr arithmetic for set-up work only: roots of unity, the
// coset generator, 1/n, parameter conversion.  Same storage as the device:
// 4 LE u64 limbs of x*2^256 mod p, canonical.  The hot path never runs here.
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>

namespace pk {

struct hfr {
  uint64_t l[4];
};

namespace hf {
typedef unsigned __int128 u128;
static const uint64_t P[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                              0x30644e72e131a029ULL};
static const uint64_t NP0 = 0xc2e1f593efffffffULL;
static const hfr R2 = {{0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL,
                        0x0216d0b17f4e44a5ULL}};
static const hfr ONE = {{0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL,
                         0x0e0a77c19a07df2fULL}};

inline bool geq_p(const uint64_t a[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != P[i]) return a[i] > P[i];
  }
  return true;
}
inline void sub_p(uint64_t a[4]) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - P[i] - br;
    a[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}
inline hfr add(hfr a, const hfr& b) {
  u128 c = 0;
  for (int i = 0; i < 4; i++) {
    c += (u128)a.l[i] + b.l[i];
    a.l[i] = (uint64_t)c;
    c >>= 64;
  }
  if (geq_p(a.l)) sub_p(a.l);
  return a;
}
inline hfr sub(hfr a, const hfr& b) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a.l[i] - b.l[i] - br;
    a.l[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
      c += (u128)a.l[i] + P[i];
      a.l[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  return a;
}
// Montgomery product, canonical output (4x64 CIOS).
inline hfr mul(const hfr& a, const hfr& b) {
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
    uint64_t m = t[0] * NP0;
    c = ((u128)m * P[0] + t[0]) >> 64;
    for (int j = 1; j < 4; j++) {
      c += (u128)m * P[j] + t[j];
      t[j - 1] = (uint64_t)c;
      c >>= 64;
    }
    c += t[4];
    t[3] = (uint64_t)c;
    t[4] = t[5] + (uint64_t)(c >> 64);
  }
  hfr r = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || geq_p(r.l)) sub_p(r.l);
  return r;
}
inline hfr from_u64(uint64_t v) { return mul(hfr{{v, 0, 0, 0}}, R2); }
// any 256-bit integer -> Montgomery form of its residue
inline hfr from_int(const hfr& x) { return mul(x, R2); }
inline hfr pow(hfr b, uint64_t e) {
  hfr r = ONE;
  while (e) {
    if (e & 1) r = mul(r, b);
    b = mul(b, b);
    e >>= 1;
  }
  return r;
}
inline hfr inv(const hfr& a) {
  // a^(p-2)
  static const uint64_t E[4] = {0x43e1f593efffffffULL, 0x2833e84879b97091ULL,
                                0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  hfr r = ONE;
  for (int i = 3; i >= 0; i--)
    for (int b = 63; b >= 0; b--) {
      r = mul(r, r);
      if ((E[i] >> b) & 1) r = mul(r, a);
    }
  return r;
}
// Big-endian hex string (canonical value) -> Montgomery form.
inline hfr from_hex(const char* h) {
  hfr x = {{0, 0, 0, 0}};
  size_t n = strlen(h);
  for (size_t i = 0; i < n && i < 64; i++) {
    char ch = h[n - 1 - i];
    uint64_t v = (ch >= '0' && ch <= '9')   ? (uint64_t)(ch - '0')
                 : (ch >= 'a' && ch <= 'f') ? (uint64_t)(ch - 'a' + 10)
                                            : (uint64_t)(ch - 'A' + 10);
    x.l[i / 16] |= v << (4 * (i % 16));
  }
  return from_int(x);
}
// The 2^28-th root of unity of BN254 Fr that ark-bn254 fixes
// (TWO_ADIC_ROOT_OF_UNITY = 5^((p-1)/2^28)), as an integer:
// 19103219067921713944291392827692070036145651957329286315305642004821462161904.
inline hfr two_adic_root() {
  return from_int(hfr{{0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL,
                       0x2a3c09f0a58a7e85ULL}});
}
// get_root_of_unity(2^k): the 2^28-th root squared 28-k times.
inline hfr root_of_unity_log(unsigned k) {
  hfr g = two_adic_root();
  for (unsigned i = k; i < 28; i++) g = mul(g, g);
  return g;
}
}  // namespace hf
}  // namespace pk
