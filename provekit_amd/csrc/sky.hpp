// Skyscraper-v2 2-to-1 compression over BN254 Fr on the VALU.
//
// Semantics (skyscraper/core/src/reference.rs:41-98; the fast form in
// generic.rs:77-102 computes the same function): inputs are two arbitrary
// 256-bit little-endian integers l, r taken mod p; 18 Feistel rounds
//   r += F_i(l) + rc[i]; swap(l, r)
// with F_i(x) = x^2 * 2^-256 (one Montgomery squaring) for rounds 0-5, 8, 9,
// 12-17 and F_i = bar for rounds 6, 7, 10, 11; the output is (l + l_in) mod p,
// canonical.  bar(x): the canonical value's 32 LE bytes rotated by 16, the
// S-box applied to every byte, read back mod p.
//
// Values are kept in [0, 2p) between rounds; bar canonicalises its input and
// the output is canonical, so any representation choice gives the same bytes.
#pragma once
#include "fr.hpp"

namespace pk {

// Round constants rc[0..17] as integers < p, 8 LE u32 limbs each (the
// Skyscraper BN254 t = 1 instance; skyscraper/core/src/constants.rs:30-49).
__constant__ static const uint32_t kSkyRC[18][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {0x270bd744u, 0x903c4324u, 0x08a7d269u, 0x873125f7u, 0x06c83855u, 0x081dd279u, 0xea6d7667u, 0x276b1823u},
    {0x4b378d71u, 0x7ac8edbbu, 0xd99e2cb7u, 0xe29d79f3u, 0x4c1a5a18u, 0x75141791u, 0x58a484a6u, 0x0cf02bd7u},
    {0x69e5bc36u, 0xfa7adc67u, 0x7cca387du, 0x1c3f8e29u, 0x63481db0u, 0x0eb7730du, 0x18ede544u, 0x25b0e03fu},
    {0x2f03cfb7u, 0x57847e65u, 0x68873404u, 0x33440b96u, 0x49af80bcu, 0x955a32e8u, 0xbe14ae70u, 0x002882fcu},
    {0x6257d4d7u, 0x97923139u, 0x1b37d3c1u, 0x29989c3eu, 0x7f1277bau, 0x12ef02b4u, 0x1e2b7a9cu, 0x039ad857u},
    {0xabbb7887u, 0xb5b48465u, 0xe6ba2d2bu, 0xa72a6bc5u, 0x712f7b29u, 0x4cd48043u, 0x0fc1fc1au, 0x1142d541u},
    {0x059075d3u, 0x7ab2c156u, 0x047999b2u, 0x17cb3594u, 0x98f289f7u, 0x44f2c935u, 0x69bc0becu, 0x1d78439fu},
    {0x138b8edbu, 0x05d7a965u, 0xd55c48b1u, 0x36ef35a3u, 0xac6f1628u, 0x8ddfb8a1u, 0x08f4ff82u, 0x258588a5u},
    {0xfccb49e9u, 0x1596fb9au, 0x9a09a95bu, 0x9a7367d6u, 0x84e4c157u, 0x9bc43f69u, 0xd2f514feu, 0x13087879u},
    {0x3b4109fau, 0x295ccd23u, 0xed868012u, 0xe1d72f89u, 0x4bc88a8eu, 0x2e9e1eeau, 0x98c45232u, 0x17dadee8u},
    {0xaa1f486fu, 0x9a8590b4u, 0x30e9130eu, 0xb75834b4u, 0x34d5de31u, 0xb8e90b10u, 0x46e7f4a6u, 0x295c6d15u},
    {0x4c6eb892u, 0x850adcb7u, 0x05b92fc3u, 0x07699ef3u, 0xa1720f2du, 0x4ef96a2bu, 0x1d3ed446u, 0x1288ca0eu},
    {0x49d1b5eeu, 0x01960f93u, 0x69371c69u, 0x8ccad307u, 0x91c98662u, 0xe5c81e89u, 0x1ae023f3u, 0x17563b4du},
    {0x76b32917u, 0x6ba01e94u, 0xdd977bc9u, 0xa1cb0a3au, 0x5815f030u, 0x86815a94u, 0xe91a1eeau, 0x2869043bu},
    {0x5511d976u, 0x81776c88u, 0x47f414e7u, 0x7475d34fu, 0x095d96cfu, 0x5d090056u, 0xff59e79au, 0x14941f0au},
    {0x8fc8c034u, 0xbc40b4fdu, 0xcce4fd48u, 0xbb7142c3u, 0x8a39005au, 0x31835675u, 0x90f4379fu, 0x1ce337a1u},
    {0, 0, 0, 0, 0, 0, 0, 0},
};

// 4p (< 2^256), for reducing arbitrary 256-bit integers.
__device__ __forceinline__ uint32_t p4_limb(int i) {
  switch (i) {
    case 0: return 0xc0000004u; case 1: return 0x0f87d64fu; case 2: return 0xe6e5c245u;
    case 3: return 0xa0cfa121u; case 4: return 0x06056174u; case 5: return 0xe14116dau;
    case 6: return 0x84c680a6u; default: return 0xc19139cbu;
  }
}
// any 256-bit integer (< 2^256 < 5.3p) -> [0, 2p): a - 4p if that is not
// negative (< 1.3p), else a - 2p if not negative, else a.  The two borrow
// chains are independent, so a lone wave fills the wait states of one with
// the other (952 -> 856 cycles per bar on one wave, same instruction count as
// subtracting 4p and then 2p in turn).
__device__ __forceinline__ void reduce_any(fr_t& a) {
  fr_t d2, d4;
  uint32_t b2 = 0, b4 = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    d2.l[i] = subb(a.l[i], p2_limb(i), b2);
    d4.l[i] = subb(a.l[i], p4_limb(i), b4);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t x = b2 ? a.l[i] : d2.l[i];
    a.l[i] = b4 ? x : d4.l[i];
  }
}

// Byte S-box on the four bytes of a word: rotl1(v ^ (rotl1(~v) & rotl2(v) & rotl3(v))).
__device__ __forceinline__ uint32_t sbox4(uint32_t v) {
  const uint32_t nv = ~v;
  const uint32_t r1 = ((nv & 0x7f7f7f7fu) << 1) | ((nv >> 7) & 0x01010101u);
  const uint32_t r2 = ((v & 0x3f3f3f3fu) << 2) | ((v >> 6) & 0x03030303u);
  const uint32_t r3 = ((v & 0x1f1f1f1fu) << 3) | ((v >> 5) & 0x07070707u);
  const uint32_t t = v ^ (r1 & r2 & r3);
  return ((t & 0x7f7f7f7fu) << 1) | ((t >> 7) & 0x01010101u);
}

// bar on a lazy value; lazy result.
__device__ __forceinline__ fr_t sky_bar(fr_t x) {
  fr_canon(x);
  fr_t y;
#pragma unroll
  for (int i = 0; i < 8; i++) y.l[i] = sbox4(x.l[(i + 4) & 7]);
  reduce_any(y);
  return y;
}

__device__ __forceinline__ fr_t sky_rc(int round) {
  fr_t c;
#pragma unroll
  for (int i = 0; i < 8; i++) c.l[i] = kSkyRC[round][i];
  return c;
}

// Double round dr of the reference's nine (ss / bb, reference.rs:49-78): two
// Feistel rounds update r then l in place, so nothing is swapped.
__device__ __forceinline__ void sky_double_round(fr_t& l, fr_t& r, int dr) {
  const bool bb = (dr == 3) | (dr == 5);  // rounds 6, 7 and 10, 11
  const fr_t f0 = bb ? sky_bar(l) : fr_sqr(l);
  r = fr_add_ilv(r, f0);
  if (dr != 0) r = fr_add_ilv(r, sky_rc(2 * dr));  // rc[0] = 0
  const fr_t f1 = bb ? sky_bar(r) : fr_sqr(r);
  l = fr_add_ilv(l, f1);
  if (dr != 8) l = fr_add_ilv(l, sky_rc(2 * dr + 1));  // rc[17] = 0
}

// compress(l, r) for arbitrary 256-bit inputs; canonical output.  The nine
// double rounds are straight-line code (~5.0k instructions, ~34 KB): no
// branches, the round constants as literals, and the scheduler works across
// rounds.  On one wave this took 30.3k cycles against 35.3k for a loop over
// the double rounds (21.7k now, with fr_add_ilv, fr_sqr and the top-limb
// terms), and the full-occupancy rate is ~2 % higher.  Straight-line code starting cold costs
// little (~0.3 us per launch, measured), but several inlined copies in a
// kernel whose waves all stream through them would overflow the 64 KB
// instruction cache: every kernel here calls it from one place.  Two entry
// forms of the same function: by value (the leaf fold's accumulator) and by
// reference; each kernel keeps the form its measured code was built with.
__device__ __forceinline__ fr_t sky_compress_v(fr_t l, fr_t r) {
  reduce_any(l);
  reduce_any(r);
  const fr_t t = l;
#pragma unroll
  for (int dr = 0; dr < 9; dr++) sky_double_round(l, r, dr);
  fr_t o = fr_add_ilv(l, t);
  fr_canon(o);
  return o;
}
__device__ __forceinline__ fr_t sky_compress(const fr_t& l, const fr_t& r) {
  return sky_compress_v(l, r);
}

// ---------------------------------------------------------------- table form
// The reference's fused reduction (reduce_partial_add_rc, reduce.rs:48-55,
// with MODULUS_N_MINUS_RC, constants.rs:51-58): a round's sum x = r + F(l)
// (below 4p, no reduction) becomes x - (k p - rc) with k = the multiple of p
// read off x's top limb, one subtraction chain instead of two reduced
// additions.  k is per lane, so the 6 candidates k p - rc (k = 0..5) of each
// round sit in a 3.4 KB LDS table, filled once per workgroup.
//   k = floor(x_7 / (p_7 + 1)) = (x_7 * 0xa948e8c5) >> 61 exactly for every
//   32-bit x_7 (checked exhaustively on the CPU); one full-rate
//   v_mad_u64_u32, then a shift and a mask give the byte offset 32 k.
//   k p <= x_7 2^224 <= x, and
//   x - k p < p + (k + 1) 2^224, so x - k p + rc < 2p for every round
//   constant (all lie below p - 2^250): the lazy range [0, 2p) is kept.
// Half-round h (0..17) uses row h: rc[h] (rows 0 and 17 have rc = 0, the
// input, S-box and output reductions use row 0).  Per round: an 8-limb add,
// the multiple and its offset (3) and an 8-limb subtraction: 19 VALU
// instructions against 56 for two reduced additions with the constant moved
// into registers.
// The same LDS block holds the byte S-box as a 256-entry table (sb), which
// the table-form bar reads instead of evaluating the bit-sliced form.
constexpr int kSkyTabRows = 18, kSkyTabMults = 6;
struct SkyTab {
  uint4 e[kSkyTabRows * kSkyTabMults * 2];  // entry (h, k): e[2 (6h + k)], e[2 (6h + k) + 1]
  uint8_t sb[256];                          // sb[x] = S-box of byte x
};

// Fill the table; every thread of the workgroup calls it, then a barrier.
__device__ __forceinline__ void sky_tab_fill(SkyTab& tab) {
  for (int t = threadIdx.x; t < 256; t += blockDim.x) tab.sb[t] = (uint8_t)sbox4((uint32_t)t);
  for (int t = threadIdx.x; t < kSkyTabRows * kSkyTabMults; t += blockDim.x) {
    const int h = t / kSkyTabMults, k = t % kSkyTabMults;
    uint32_t v[8];
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {  // k p
      c += (uint64_t)k * p_limb(i);
      v[i] = (uint32_t)c;
      c >>= 32;
    }
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = subb(v[i], kSkyRC[h][i], br);  // - rc, mod 2^256
    tab.e[2 * t] = make_uint4(v[0], v[1], v[2], v[3]);
    tab.e[2 * t + 1] = make_uint4(v[4], v[5], v[6], v[7]);
  }
}

// 32 floor(x7 / (p_7 + 1)), the byte offset of multiple k within a table
// row, exact for all 32-bit x7: ((t >> 61) << 5) = (t_hi >> 24) & 0xe0, a
// shift and a mask after the multiply
__device__ __forceinline__ uint32_t sky_top_off(uint32_t x7) {
  uint64_t t, cy;
  asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(t), "=s"(cy) : "v"(x7), "s"(0xa948e8c5u));
  (void)cy;
  return ((uint32_t)(t >> 32) >> 24) & 0xe0u;
}

// x (any value below 2^256) -> x - (k p - rc[h]) in [0, 2p) (rc = 0: [0, p + 6 2^224))
__device__ __forceinline__ fr_t sky_tab_reduce(const fr_t& x, const SkyTab& tab, int h) {
  const char* row = reinterpret_cast<const char*>(&tab.e[2 * kSkyTabMults * h]) + sky_top_off(x.l[7]);
  const uint4 a = *reinterpret_cast<const uint4*>(row);
  const uint4 b = *reinterpret_cast<const uint4*>(row + 16);
  fr_t r;
  uint32_t br = 0;
  r.l[0] = subb(x.l[0], a.x, br);
  r.l[1] = subb(x.l[1], a.y, br);
  r.l[2] = subb(x.l[2], a.z, br);
  r.l[3] = subb(x.l[3], a.w, br);
  r.l[4] = subb(x.l[4], b.x, br);
  r.l[5] = subb(x.l[5], b.y, br);
  r.l[6] = subb(x.l[6], b.z, br);
  r.l[7] = subb(x.l[7], b.w, br);
  return r;
}

// S-box of the four bytes of a word by table: four byte reads (the byte
// extracts are the addresses) and two byte permutes plus an OR to merge them,
// 7 VALU instructions against 16 for sbox4 (shifts, bit-field inserts and
// 3-input bit operations).  The 256-byte table is 64 dwords, so a wave's
// reads touch at most two dwords per bank.
__device__ __forceinline__ uint32_t sbox4_t(uint32_t v, const uint8_t* sb) {
  const uint32_t b0 = sb[v & 0xffu], b1 = sb[(v >> 8) & 0xffu];
  const uint32_t b2 = sb[(v >> 16) & 0xffu], b3 = sb[v >> 24];
  // v_perm_b32 selector bytes: 0-3 pick bytes of the second operand, 4-7 of
  // the first, 0x0c gives zero
  const uint32_t lo = __builtin_amdgcn_perm(b1, b0, 0x0c0c0400u);  // [b0, b1, 0, 0]
  const uint32_t hi = __builtin_amdgcn_perm(b3, b2, 0x04000c0cu);  // [0, 0, b2, b3]
  return lo | hi;
}

// lazy -> canonical (fr_canon), with p read from the table (row 0, multiple
// 1: p - rc[0] = p) instead of moved into registers limb by limb: a VOP2
// subtract-with-borrow cannot take a literal next to vcc, so fr_canon spends
// seven moves on the constant
__device__ __forceinline__ void sky_canon_t(fr_t& x, const SkyTab& tab) {
  const uint4 a = tab.e[2], b = tab.e[3];
  const uint32_t pl[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  fr_t d;
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = subb(x.l[i], pl[i], br);
#pragma unroll
  for (int i = 0; i < 8; i++) x.l[i] = br ? x.l[i] : d.l[i];
}

// bar with the table: the raw S-box output (any 256-bit value) comes back
// below p + 6 2^224
__device__ __forceinline__ fr_t sky_bar_t(fr_t x, const SkyTab& tab) {
  sky_canon_t(x, tab);
  fr_t y;
#pragma unroll
  for (int i = 0; i < 8; i++) y.l[i] = sbox4_t(x.l[(i + 4) & 7], tab.sb);
  return sky_tab_reduce(y, tab, 0);
}

// compress(l, r) with the table reductions (generic.rs:77-102 order: r +=
// F(l), then the fused reduction with rc[h]); canonical output, the same
// bytes as sky_compress.
__device__ __forceinline__ fr_t sky_compress_t(fr_t l, fr_t r, const SkyTab& tab) {
  l = sky_tab_reduce(l, tab, 0);
  r = sky_tab_reduce(r, tab, 0);
  const fr_t t = l;
#pragma unroll
  for (int dr = 0; dr < 9; dr++) {
    const bool bb = (dr == 3) | (dr == 5);  // rounds 6, 7 and 10, 11
    const fr_t f0 = bb ? sky_bar_t(l, tab) : fr_sqr(l);
    r = sky_tab_reduce(fr_add_raw(r, f0), tab, 2 * dr);
    const fr_t f1 = bb ? sky_bar_t(r, tab) : fr_sqr(r);
    l = sky_tab_reduce(fr_add_raw(l, f1), tab, 2 * dr + 1);
  }
  // l + t < 2p + p + 6 2^224: the row-0 reduction leaves it below
  // p + 6 2^224, one conditional subtraction of p makes it canonical
  fr_t o = sky_tab_reduce(fr_add_raw(l, t), tab, 0);
  sky_canon_t(o, tab);
  return o;
}

// SkyscraperHashEngine::hash_many of one message of `chunks` words
// (provekit/common/src/skyscraper/whir.rs:42-100): acc = m0, then
// acc = compress(acc, m_k).  Words are read four at a time, one 128-byte line
// per lane: neighbouring lanes hash neighbouring messages, so reading one word
// per compression would touch every line four times far apart in time, and
// the lines get evicted in between.  Four compressions take tens of
// microseconds per wave, so the other waves of the SIMD hide the next load.
// The compressions of a group run in a rolled loop over a shifting window,
// so the kernel holds one copy of the compression.
__device__ __forceinline__ fr_t sky_hash_message(const fr_t* m, uint64_t chunks) {
  constexpr int G = 4;
  fr_t acc = {};
  for (uint64_t g = 0; g < chunks; g += G) {
    fr_t w[G] = {};
#pragma unroll
    for (int j = 0; j < G; j++)
      if (g + j < chunks) w[j] = fr_load(m + g + j);
    int n = chunks - g < (uint64_t)G ? (int)(chunks - g) : G;
    if (g == 0) {  // m0 starts the fold
      acc = w[0];
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
      n -= 1;
    }
#pragma unroll 1
    for (int j = 0; j < n; j++) {
      acc = sky_compress_v(acc, w[0]);
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
    }
  }
  return acc;
}

// sky_hash_message with the table-form compression.
__device__ __forceinline__ fr_t sky_hash_message_t(const fr_t* m, uint64_t chunks,
                                                   const SkyTab& tab) {
  constexpr int G = 4;
  fr_t acc = {};
  for (uint64_t g = 0; g < chunks; g += G) {
    fr_t w[G] = {};
#pragma unroll
    for (int j = 0; j < G; j++)
      if (g + j < chunks) w[j] = fr_load(m + g + j);
    int n = chunks - g < (uint64_t)G ? (int)(chunks - g) : G;
    if (g == 0) {  // m0 starts the fold
      acc = w[0];
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
      n -= 1;
    }
#pragma unroll 1
    for (int j = 0; j < n; j++) {
      acc = sky_compress_t(acc, w[0], tab);
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
    }
  }
  return acc;
}

// Wave-level helpers for task loops whose waves work independently (no
// workgroup barrier).  Every lane runs them: a lane-0-only branch inside such
// a loop lets the compiler restructure the loop so that lane 0 leaves it at
// the end of an iteration while the other lanes go on, after which lane 0's
// broadcasts no longer reach them.
//
// Next task id of the wave: one atomic (lane 0 adds 1, the others 0), lane
// 0's result to all lanes.
__device__ __forceinline__ uint32_t wave_take(uint32_t* counter) {
  const uint32_t add = (threadIdx.x & 63) == 0 ? 1u : 0u;
  const uint32_t v = __hip_atomic_fetch_add(counter, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_amdgcn_readfirstlane(v);
}
// Bounded wait for flags [first, first + count) (first < 0: none): each lane
// polls one flag of a group of 64 with a relaxed load, and the wave moves on
// when every flag of the group is set; gives up when the error flag is
// raised or after `spin` polls (then raises it).
__device__ __forceinline__ bool wave_wait_range(const uint32_t* done, int64_t first, uint32_t count,
                                                uint32_t* err, uint64_t spin) {
  if (first < 0) return true;
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t base = 0; base < count; base += 64) {
    const bool mine = base + lane < count;
    const uint32_t* f = done + first + base + (mine ? lane : 0);
    uint64_t it = 0;
    for (;; it++) {
      const uint32_t v = mine ? __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1u;
      if (__all(v != 0)) break;
      if (it + 1 >= spin) {
        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
      if (__builtin_amdgcn_readfirstlane(
              __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        return false;
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return true;
}
__device__ __forceinline__ fr_t fr_shfl(const fr_t& v, int src) {
  fr_t r;
#pragma unroll
  for (int k = 0; k < 8; k++) r.l[k] = (uint32_t)__shfl((int)v.l[k], src, 64);
  return r;
}

}  // namespace pk
