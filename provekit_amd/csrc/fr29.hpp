// BN254 Fr in 9 x 29-bit limbs, for long chains of products (Poseidon2).
//
// With 29-bit digits a 32x32 -> 64-bit multiply-add never needs a carry-out:
// a column of up to 18 digit products (9 of a*b, 9 of m*p) plus the incoming
// carry stays below 2^64, so a Montgomery product is 162 multiply-adds into
// one 64-bit accumulator and no add-with-carry chain (the 8 x 32 form spends
// one add-with-carry per digit product).  Squares use the symmetric cross
// products once and double the column sum.
//
// Montgomery radix R' = 2^261 (nine 29-bit steps).  Values in this form are
// "internal": I(x) = x * 2^261 mod p, kept lazily (any integer < 2^261 that is
// congruent).  Storage values S(x) = x * 2^256 mod p (ark-ff Montgomery)
// convert with one product each way: I = mont29(S, 2^266 mod p) and
// S = mont29(I, 2^256 mod p).  Since p / 2^261 < 2^-7.4, a product of two
// values below k*p is below (k^2 * 0.006 + 1) * p, so products pull values
// back towards p by themselves; sums grow and are reduced where a chain
// would otherwise let them climb (reduce29).
//
// Limb bounds: mont29 needs one operand's limbs < 2^31 and the other's
// < 2^29 (or both < 2^30); its output has limbs 0..7 < 2^29.  Limb-wise sums
// are brought back with norm29 (carry propagation) before they are squared.
#pragma once
#include "fr.hpp"

namespace pk {

struct f29 {
  uint32_t l[9];
};

constexpr uint32_t kM29 = 0x1fffffffu;

// p in base 2^29
__device__ __forceinline__ uint32_t p29_limb(int i) {
  switch (i) {
    case 0: return 0x10000001u; case 1: return 0x1f0fac9fu; case 2: return 0x0e5c2450u;
    case 3: return 0x07d090f3u; case 4: return 0x1585d283u; case 5: return 0x02db40c0u;
    case 6: return 0x00a6e141u; case 7: return 0x0e5c2634u; default: return 0x0030644eu;
  }
}

// a * b * 2^-261 mod p (lazy).  The quotient digit uses -p^-1 = 2^28 - 1
// (mod 2^29): m = -x (1 + 2^28) mod 2^29.
template <bool SQR>
__device__ __forceinline__ f29 mont29(const f29& a, const f29& b) {
  uint32_t m[9];
  f29 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    if (SQR) {
      uint64_t cr = 0;
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j > i && j < 9) cr += (uint64_t)a.l[i] * a.l[j];
      }
      acc += cr << 1;
      if ((k & 1) == 0 && k / 2 < 9) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
    } else {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j >= 0 && j < 9) acc += (uint64_t)a.l[i] * b.l[j];
      }
    }
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int q = k - j;
      if (j < k && q >= 1 && q < 9) acc += (uint64_t)m[j] * p29_limb(q);
    }
    if (k < 9) {
      const uint32_t x = (uint32_t)acc;
      m[k] = (0u - (x + (x << 28))) & kM29;
      acc += (uint64_t)m[k] * p29_limb(0);  // the low 29 bits become zero
    } else {
      r.l[k - 9] = (uint32_t)acc & kM29;
    }
    acc >>= 29;
  }
  r.l[8] = (uint32_t)acc;
  return r;
}
__device__ __forceinline__ f29 mul29(const f29& a, const f29& b) { return mont29<false>(a, b); }
__device__ __forceinline__ f29 sqr29(const f29& a) { return mont29<true>(a, a); }

__device__ __forceinline__ f29 add29(const f29& a, const f29& b) {
  f29 r;
#pragma unroll
  for (int k = 0; k < 9; k++) r.l[k] = a.l[k] + b.l[k];
  return r;
}

// carries up: limbs 0..7 < 2^29, limb 8 takes the rest
__device__ __forceinline__ void norm29(f29& a) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    a.l[k + 1] += a.l[k] >> 29;
    a.l[k] &= kM29;
  }
}

// a (normalized, value < 2^262, so a.l[8] < 2^30) -> a value below 4p, same
// residue: subtract q p with q a lower estimate of a / p from the top limb,
// in integers: q = floor(a.l[8] floor(2^52 / 3171407) / 2^52) - 1, where
// p >> 232 = 3171406.45..., so q <= floor(a / p) and a - q p < 2.0001 p
// (checked at both ends of every top-limb value's range in a CPU model).
__device__ __forceinline__ void reduce29(f29& a) {
  constexpr uint64_t kRecip = (1ull << 52) / 3171407;  // 0x54a4736b
  uint32_t q = (uint32_t)(((uint64_t)a.l[8] * kRecip) >> 52);
  q = q ? q - 1 : 0;
  int64_t carry = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int64_t t = (int64_t)a.l[k] - (int64_t)((uint64_t)q * p29_limb(k)) + carry;
    if (k < 8) {
      a.l[k] = (uint32_t)t & kM29;
      carry = t >> 29;  // arithmetic: -1, 0 or a small negative multiple
    } else {
      a.l[8] = (uint32_t)t;
    }
  }
}

// 8 x 32 <-> 9 x 29, same integer (from29 needs normalized limbs and a value
// below 2^256)
__device__ __forceinline__ f29 to29(const fr_t& s) {
  f29 r;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int b = 29 * k, w = b >> 5, sh = b & 31;
    uint64_t v = s.l[w] >> sh;
    if (w + 1 < 8 && sh > 3) v |= (uint64_t)s.l[w + 1] << (32 - sh);
    r.l[k] = (uint32_t)v & kM29;
  }
  return r;
}
__device__ __forceinline__ fr_t from29(const f29& a) {
  fr_t s;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int b = 32 * j, k = b / 29, sh = b % 29;
    uint64_t v = (uint64_t)a.l[k] >> sh;
    v |= (uint64_t)a.l[k + 1] << (29 - sh);
    if (k + 2 < 9 && 58 - sh < 32) v |= (uint64_t)a.l[k + 2] << (58 - sh);
    s.l[j] = (uint32_t)v;
  }
  return s;
}

}  // namespace pk
