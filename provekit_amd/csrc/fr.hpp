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
// lazy + lazy -> lazy, the same 24 instructions as fr_add with the add
// chain (carry in an SGPR pair) and the subtract-2p chain (borrow in vcc)
// interleaved: an independent instruction sits between every carry's write
// and its read, so a lone wave waits less on the carry chains (164 against
// 272 cycles per addition on one wave; the compiler emits the two chains one
// after the other).  The 2p limbs are VGPR operands: a VOP2 carry op cannot take a
// literal next to vcc (one constant-bus read).  For latency-bound code (the
// Skyscraper rounds); ALU only, nothing left in flight.
__device__ __forceinline__ fr_t fr_add_ilv(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t d[8];
  uint64_t c;
  // add k (carry: SGPR pair), subtract-2p k (borrow: vcc), one wait state:
  // every carry and borrow is read two instructions after it is written
  asm(
      "v_add_co_u32_e64 %0, %16, %17, %25\n\t"
      "v_subrev_co_u32_e32 %8, vcc, 0xe0000002, %0\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %1, %16, %18, %26, %16\n\t"
      "v_subb_co_u32_e32 %9, vcc, %1, %33, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %2, %16, %19, %27, %16\n\t"
      "v_subb_co_u32_e32 %10, vcc, %2, %34, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %3, %16, %20, %28, %16\n\t"
      "v_subb_co_u32_e32 %11, vcc, %3, %35, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %4, %16, %21, %29, %16\n\t"
      "v_subb_co_u32_e32 %12, vcc, %4, %36, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %5, %16, %22, %30, %16\n\t"
      "v_subb_co_u32_e32 %13, vcc, %5, %37, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %6, %16, %23, %31, %16\n\t"
      "v_subb_co_u32_e32 %14, vcc, %6, %38, vcc\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %7, %16, %24, %32, %16\n\t"
      "v_subb_co_u32_e32 %15, vcc, %7, %39, vcc\n\t"
      "s_nop 1\n\t"
      "v_cndmask_b32_e32 %0, %8, %0, vcc\n\t"
      "v_cndmask_b32_e32 %1, %9, %1, vcc\n\t"
      "v_cndmask_b32_e32 %2, %10, %2, vcc\n\t"
      "v_cndmask_b32_e32 %3, %11, %3, vcc\n\t"
      "v_cndmask_b32_e32 %4, %12, %4, vcc\n\t"
      "v_cndmask_b32_e32 %5, %13, %5, vcc\n\t"
      "v_cndmask_b32_e32 %6, %14, %6, vcc\n\t"
      "v_cndmask_b32_e32 %7, %15, %7, vcc"
      : "=&v"(r.l[0]), "=&v"(r.l[1]), "=&v"(r.l[2]), "=&v"(r.l[3]), "=&v"(r.l[4]), "=&v"(r.l[5]), "=&v"(r.l[6]), "=&v"(r.l[7]),
        "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3]), "=&v"(d[4]), "=&v"(d[5]), "=&v"(d[6]), "=&v"(d[7]), "=&s"(c)
      : "v"(a.l[0]), "v"(a.l[1]), "v"(a.l[2]), "v"(a.l[3]),
        "v"(a.l[4]), "v"(a.l[5]), "v"(a.l[6]), "v"(a.l[7]),
        "v"(b.l[0]), "v"(b.l[1]), "v"(b.l[2]), "v"(b.l[3]),
        "v"(b.l[4]), "v"(b.l[5]), "v"(b.l[6]), "v"(b.l[7]),
        "v"(p2_limb(1)), "v"(p2_limb(2)), "v"(p2_limb(3)), "v"(p2_limb(4)), "v"(p2_limb(5)), "v"(p2_limb(6)), "v"(p2_limb(7))
      : "vcc");
  (void)c;
  return r;
}
// lazy - lazy -> lazy, the fr_add_ilv scheme for a subtraction: the
// difference's borrow chain (SGPR pair) interleaved with the chain that adds
// 2p back to it (vcc), then the final borrow picks the one that is in range.
// 24 instructions like fr_sub; for latency-bound code.
__device__ __forceinline__ fr_t fr_sub_ilv(const fr_t& a, const fr_t& b) {
  fr_t r;
  uint32_t u[8];
  uint64_t c;
  // subtract k (borrow: SGPR pair), add-2p k (carry: vcc), one wait state,
  // as in fr_add_ilv
  asm(
      "v_sub_co_u32_e64 %0, %16, %17, %25\n\t"
      "v_add_co_u32_e32 %8, vcc, 0xe0000002, %0\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %1, %16, %18, %26, %16\n\t"
      "v_addc_co_u32_e32 %9, vcc, %33, %1, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %2, %16, %19, %27, %16\n\t"
      "v_addc_co_u32_e32 %10, vcc, %34, %2, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %3, %16, %20, %28, %16\n\t"
      "v_addc_co_u32_e32 %11, vcc, %35, %3, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %4, %16, %21, %29, %16\n\t"
      "v_addc_co_u32_e32 %12, vcc, %36, %4, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %5, %16, %22, %30, %16\n\t"
      "v_addc_co_u32_e32 %13, vcc, %37, %5, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %6, %16, %23, %31, %16\n\t"
      "v_addc_co_u32_e32 %14, vcc, %38, %6, vcc\n\t"
      "s_nop 0\n\t"
      "v_subb_co_u32_e64 %7, %16, %24, %32, %16\n\t"
      "v_addc_co_u32_e32 %15, vcc, %39, %7, vcc\n\t"
      "s_nop 0\n\t"
      "v_cndmask_b32_e64 %0, %0, %8, %16\n\t"
      "v_cndmask_b32_e64 %1, %1, %9, %16\n\t"
      "v_cndmask_b32_e64 %2, %2, %10, %16\n\t"
      "v_cndmask_b32_e64 %3, %3, %11, %16\n\t"
      "v_cndmask_b32_e64 %4, %4, %12, %16\n\t"
      "v_cndmask_b32_e64 %5, %5, %13, %16\n\t"
      "v_cndmask_b32_e64 %6, %6, %14, %16\n\t"
      "v_cndmask_b32_e64 %7, %7, %15, %16"
      : "=&v"(r.l[0]), "=&v"(r.l[1]), "=&v"(r.l[2]), "=&v"(r.l[3]), "=&v"(r.l[4]), "=&v"(r.l[5]), "=&v"(r.l[6]), "=&v"(r.l[7]),
        "=&v"(u[0]), "=&v"(u[1]), "=&v"(u[2]), "=&v"(u[3]), "=&v"(u[4]), "=&v"(u[5]), "=&v"(u[6]), "=&v"(u[7]), "=&s"(c)
      : "v"(a.l[0]), "v"(a.l[1]), "v"(a.l[2]), "v"(a.l[3]),
        "v"(a.l[4]), "v"(a.l[5]), "v"(a.l[6]), "v"(a.l[7]),
        "v"(b.l[0]), "v"(b.l[1]), "v"(b.l[2]), "v"(b.l[3]),
        "v"(b.l[4]), "v"(b.l[5]), "v"(b.l[6]), "v"(b.l[7]),
        "v"(p2_limb(1)), "v"(p2_limb(2)), "v"(p2_limb(3)), "v"(p2_limb(4)), "v"(p2_limb(5)), "v"(p2_limb(6)), "v"(p2_limb(7))
      : "vcc");
  (void)c;
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
// Each digit product is a multiply-add pair: v_mad_u64_u32 adds the 32x32
// product into the aligned 64-bit accumulator pair and carries out into an
// SGPR lane mask, which the following v_addc_co_u32 folds into the third
// word.  Plain C does not get the compiler to use that carry-out, and 64-bit
// operands must sit in even-aligned register pairs, so a row written in C
// spends about three moves per digit product; the pair spends none.  ALU
// only: no memory operand, nothing left in flight.
// a, b lazy -> a*b*2^-256 lazy.
// Montgomery multiplication in finely integrated product scanning (FIPS) form:
// column k of a*b and of m*p are summed in one 96-bit accumulator; for k < 8
// the quotient digit m_k is chosen so that the column's low word cancels, for
// k >= 8 the low word is output digit k-8.  The body (fr_mul_asm.inc, written
// by gen_fr_mul.py) holds one asm statement per column, not per multiply-add:
// the compiler follows every asm statement with a wait state before its
// outputs are used, which per multiply-add cost a lone wave ~18 % of its
// product latency (1700 -> 1400 cycles measured; the same rate at full
// occupancy).  The quotient digit x * 0xefffffff mod 2^32 is asm too (the low
// word of a v_mad_u64_u32): written in C the compiler emits v_mul_lo_u32.
// Operand bound: a_7 + b_7 <= 2^32 - 18 on the top limbs, which lazy x lazy
// (0.76 * 2^32) and canonical x any value below 4p (0.95 * 2^32) meet.  In
// columns 7-14 the products that hold a top limb open the column and are
// summed without a carry add (the accumulator cannot pass 2^64 there): 16
// fewer instructions per product.
__device__ __forceinline__ fr_t fr_mul(const fr_t& a, const fr_t& b) {
  uint32_t m[8];
  fr_t r;
  uint64_t c01 = 0;
  uint32_t c2 = 0;
  uint64_t cy[3], mq;
#include "fr_mul_asm.inc"
  (void)cy;
  return r;
}

// a lazy (a < 2p) -> a^2*2^-256 lazy, the value and bytes fr_mul(a, a)
// gives.  Each cross product a_i a_j (i < j) is formed once, against the
// doubled limbs of a: d[j] = (a_j << 1) | (a_{j-1} >> 31) and, next to the
// diagonal, u[j] = a_j << 1 (fr_sqr_asm.inc, written by gen_fr_mul.py, which
// states the identity): 36 digit products for the a*a half instead of 64,
// ~40 fewer instructions per squaring, 14 squarings per Skyscraper
// compression.
__device__ __forceinline__ fr_t fr_sqr(const fr_t& a) {
  uint32_t u[8], d[8];
#pragma unroll
  for (int j = 1; j < 8; j++) {
    u[j] = a.l[j] << 1;
    d[j] = __builtin_amdgcn_alignbit(a.l[j], a.l[j - 1], 31);
  }
  uint32_t m[8];
  fr_t r;
  uint64_t c01 = 0;
  uint32_t c2 = 0;
  uint64_t cy[3], mq;
#include "fr_sqr_asm.inc"
  (void)cy;
  return r;
}

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
