// BN254 scalar field Fr on gfx950: 256-bit Montgomery arithmetic on the VALU.
//
// Storage layout is the ark-ff one used by the reference (`ark_bn254::Fr`):
// 4 little-endian u64 limbs of x*2^256 mod p, which are the same bytes as
// 8 little-endian u32 limbs.  Kernels keep elements in 8 VGPRs.
//
// Value ranges used by the kernels:
//   canonical  [0, p)
//   lazy       [0, 2p)   (4p < 2^256, so sums of two lazy values never wrap)
//   NTT lazy   [0, 4p)   (between butterfly stages only)
// Montgomery products of two lazy values are lazy without a final subtraction
// because a*b/2^256 + p < 4p^2/2^256 + p < 2p; the same holds for a canonical
// times a value below 4p (4p^2/2^256 + p < 1.76p).
//
// Multiplication is Montgomery with 32-bit digits in product-scanning form:
// every digit product is one v_mad_u64_u32 (32x32 + 64 -> 64) plus one
// add-with-carry.  There is no floating point anywhere.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pk {

struct alignas(16) fr_t {
  uint32_t l[8];
};

// p = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
#define PK_P0 0xf0000001u
#define PK_P1 0x43e1f593u
#define PK_P2 0x79b97091u
#define PK_P3 0x2833e848u
#define PK_P4 0x8181585du
#define PK_P5 0xb85045b6u
#define PK_P6 0xe131a029u
#define PK_P7 0x30644e72u
// -p^{-1} mod 2^32
#define PK_NP0 0xefffffffu

__device__ __forceinline__ uint32_t p_limb(int i) {
  switch (i) {
    case 0: return PK_P0; case 1: return PK_P1; case 2: return PK_P2; case 3: return PK_P3;
    case 4: return PK_P4; case 5: return PK_P5; case 6: return PK_P6; default: return PK_P7;
  }
}
// 2p
__device__ __forceinline__ uint32_t p2_limb(int i) {
  switch (i) {
    case 0: return 0xe0000002u; case 1: return 0x87c3eb27u; case 2: return 0xf372e122u;
    case 3: return 0x5067d090u; case 4: return 0x0302b0bau; case 5: return 0x70a08b6du;
    case 6: return 0xc2634053u; default: return 0x60c89ce5u;
  }
}

// ---------------------------------------------------------------- carries
__device__ __forceinline__ uint32_t addc(uint32_t a, uint32_t b, uint32_t& c) {
  uint32_t co;
  uint32_t r = __builtin_addc(a, b, c, &co);
  c = co;
  return r;
}
__device__ __forceinline__ uint32_t subb(uint32_t a, uint32_t b, uint32_t& br) {
  uint32_t bo;
  uint32_t r = __builtin_subc(a, b, br, &bo);
  br = bo;
  return r;
}

// r = a - m if a >= m else a, with m given limb-wise by F (p or 2p).
template <int WHICH>
__device__ __forceinline__ void cond_sub(fr_t& a) {
  fr_t d;
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = subb(a.l[i], WHICH == 1 ? p_limb(i) : p2_limb(i), br);
  // br == 1 -> a < m -> keep a
#pragma unroll
  for (int i = 0; i < 8; i++) a.l[i] = br ? a.l[i] : d.l[i];
}

// lazy -> canonical
__device__ __forceinline__ void fr_canon(fr_t& a) { cond_sub<1>(a); }
// [0, 4p) -> lazy
__device__ __forceinline__ void fr_reduce2p(fr_t& a) { cond_sub<2>(a); }

// lazy + lazy -> lazy
__device__ __forceinline__ fr_t fr_add(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(a.l[i], b.l[i], c);
  cond_sub<2>(r);
  return r;
}
// lazy - lazy -> lazy  (a - b, adding 2p back on borrow)
__device__ __forceinline__ fr_t fr_sub(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = subb(a.l[i], b.l[i], br);
  uint32_t mask = 0u - br;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(r.l[i], p2_limb(i) & mask, c);
  return r;
}
// a + b without reduction (caller guarantees a + b < 2^256)
__device__ __forceinline__ fr_t fr_add_raw(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(a.l[i], b.l[i], c);
  return r;
}
// a - b + 2p for a, b in [0, 2p): the result lies in (0, 4p), so the wrap of
// the subtraction and the carry of the addition cancel modulo 2^256.
__device__ __forceinline__ fr_t fr_sub_plus_2p(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = subb(a.l[i], b.l[i], br);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(r.l[i], p2_limb(i), c);
  return r;
}
// canonical + canonical -> canonical
__device__ __forceinline__ fr_t fr_add_canon(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(a.l[i], b.l[i], c);
  cond_sub<1>(r);
  return r;
}
// canonical - canonical -> canonical
__device__ __forceinline__ fr_t fr_sub_canon(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = subb(a.l[i], b.l[i], br);
  uint32_t mask = 0u - br;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = addc(r.l[i], p_limb(i) & mask, c);
  return r;
}

// ---------------------------------------------------------------- Montgomery
// One column step of product scanning: (c2:c01) += a*b.  The 64-bit add of the
// product into the aligned accumulator pair carries out into an SGPR lane
// mask, which the following add-with-carry folds into the third word.  Plain C
// does not get the compiler to use the carry-out of v_mad_u64_u32, and 64-bit
// operands must sit in even-aligned register pairs, so a CIOS row written in C
// spends about three moves per digit product; this pair spends none.
// ALU only: no memory operand, nothing left in flight.
__device__ __forceinline__ void macc(uint64_t& c01, uint32_t& c2, uint32_t a, uint32_t b) {
  uint64_t cy;
  asm("v_mad_u64_u32 %0, %2, %3, %4, %0\n\t"
      "v_addc_co_u32_e64 %1, %2, 0, %1, %2"
      : "+v"(c01), "+v"(c2), "=&s"(cy)
      : "v"(a), "v"(b));
}

// a, b lazy -> a*b*2^-256 lazy.
// Montgomery multiplication in finely integrated product scanning (FIPS) form:
// column k of a*b and of m*p are summed in one 96-bit accumulator; for k < 8
// the quotient digit m_k is chosen so that the column's low word cancels, for
// k >= 8 the low word is output digit k-8.
__device__ __forceinline__ fr_t fr_mul(const fr_t& a, const fr_t& b) {
  uint32_t m[8];
  fr_t r;
  uint64_t c01 = 0;
  uint32_t c2 = 0;
#pragma unroll
  for (int k = 0; k < 15; k++) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int j = k - i;
      if (j >= 0 && j < 8) macc(c01, c2, a.l[i], b.l[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int q = k - j;
      if (j < k && q >= 1 && q < 8) macc(c01, c2, m[j], p_limb(q));
    }
    if (k < 8) {
      // m = x * (-p^-1 mod 2^32) with -p^-1 = -(1 + 2^28): two full-rate ops
      // instead of a quarter-rate v_mul_lo_u32
      const uint32_t x = (uint32_t)c01;
      m[k] = 0u - (x + (x << 28));
      macc(c01, c2, m[k], PK_P0);
    } else {
      r.l[k - 8] = (uint32_t)c01;
    }
    c01 = (c01 >> 32) | ((uint64_t)c2 << 32);
    c2 = 0;
  }
  r.l[7] = (uint32_t)c01;
  return r;
}

// a lazy -> a^2*2^-256 lazy.
__device__ __forceinline__ fr_t fr_sqr(const fr_t& a) { return fr_mul(a, a); }

__device__ __forceinline__ bool fr_is_zero(const fr_t& a) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o |= a.l[i];
  return o == 0;
}

// ---------------------------------------------------------------- memory
__device__ __forceinline__ fr_t fr_load(const fr_t* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 x = q[0], y = q[1];
  fr_t r;
  r.l[0] = x.x; r.l[1] = x.y; r.l[2] = x.z; r.l[3] = x.w;
  r.l[4] = y.x; r.l[5] = y.y; r.l[6] = y.z; r.l[7] = y.w;
  return r;
}
__device__ __forceinline__ void fr_store(fr_t* p, const fr_t& v) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  q[1] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}

}  // namespace pk
