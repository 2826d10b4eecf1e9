// bn254_multiplier::{scalar_mul, scalar_sqr}
// (skyscraper/bn254-multiplier/src/scalar.rs:12-132) over many elements, one
// lane per element, restating the reference's integer steps exactly so that
// the outputs (below 2^256 - 2p, not canonical) are the same bytes:
//   t = a * b (512 bits), then with 64-bit words t0, t1, t2 of t
//   s = t[3..8) + t0 * 2^-192 + t1 * 2^-128 + t2 * 2^-64   (mod 2^320, addv)
//   m = mu0 * s0 (mod 2^64), r = words 1..4 of s + m p     (mod 2^320)
//   r -= 2p if bit 255 of r is set                          (reduce_ct)
// with 2^-64k standing for the constants I3, I2, I1 (2^-64k mod p,
// constants.rs:71-89).  Plain 32-bit schoolbook arithmetic: these are
// utility entry points, not on the commit path (the compression inlines its
// own squaring, fr.hpp).
#include "kernels.hpp"

namespace pk {

namespace {

// 2^-192, 2^-128, 2^-64 mod p and p as 32-bit limbs, little-endian
__constant__ uint32_t kI3[8] = {0x127cbe4e, 0x9bacb016, 0x31944124, 0x0b2051fa,
                                0x6091c76c, 0xb064eea4, 0x49f80c7d, 0x2b062aaa};
__constant__ uint32_t kI2[8] = {0x76f9dc6f, 0x18ee753c, 0xa329e70f, 0x54ad7e14,
                                0x4f7684df, 0x2b16366f, 0x1fdf3579, 0x133100d7};
__constant__ uint32_t kI1[8] = {0xe396ee4d, 0x2d3e8053, 0xab3c92cd, 0xca478dbe,
                                0x77f52a93, 0xb2d8f06f, 0xf7aa8f04, 0x24d6ba07};
__constant__ uint32_t kP[8] = {0xf0000001, 0x43e1f593, 0x79b97091, 0x2833e848,
                               0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72};
__constant__ uint32_t k2P[8] = {0xe0000002, 0x87c3eb27, 0xf372e122, 0x5067d090,
                                0x0302b0ba, 0x70a08b6d, 0xc2634053, 0x60c89ce5};
constexpr uint64_t kMu0 = 0xc2e1f593efffffffULL;  // -p^-1 mod 2^64

// acc (10 limbs, mod 2^320) += x (2 limbs) * y (8 limbs)
__device__ __forceinline__ void mac_2x8(uint32_t (&acc)[10], uint32_t x0, uint32_t x1,
                                        const uint32_t* y) {
  uint32_t prod[10];
#pragma unroll
  for (int k = 0; k < 10; k++) prod[k] = 0;
  const uint32_t x[2] = {x0, x1};
#pragma unroll
  for (int i = 0; i < 2; i++) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      c += (uint64_t)x[i] * y[j] + prod[i + j];
      prod[i + j] = (uint32_t)c;
      c >>= 32;
    }
    prod[i + 8] = (uint32_t)c;
  }
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < 10; k++) {
    c += (uint64_t)acc[k] + prod[k];
    acc[k] = (uint32_t)c;
    c >>= 32;
  }
}

__global__ __launch_bounds__(256) void k_scalar_mul(const fr_t* a, const fr_t* b, fr_t* out,
                                                    uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const fr_t x = fr_load(a + e);
  const fr_t y = b ? fr_load(b + e) : x;
  uint32_t t[16];
#pragma unroll
  for (int k = 0; k < 16; k++) t[k] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      c += (uint64_t)x.l[i] * y.l[j] + t[i + j];
      t[i + j] = (uint32_t)c;
      c >>= 32;
    }
    t[i + 8] = (uint32_t)c;
  }
  uint32_t s[10];
#pragma unroll
  for (int k = 0; k < 10; k++) s[k] = t[6 + k];  // 64-bit words 3..7
  mac_2x8(s, t[0], t[1], kI3);
  mac_2x8(s, t[2], t[3], kI2);
  mac_2x8(s, t[4], t[5], kI1);
  const uint64_t m = kMu0 * ((uint64_t)s[1] << 32 | s[0]);
  mac_2x8(s, (uint32_t)m, (uint32_t)(m >> 32), kP);
  fr_t r;
#pragma unroll
  for (int k = 0; k < 8; k++) r.l[k] = s[2 + k];
  if (r.l[7] >> 31) {
    uint64_t br = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint64_t d = (uint64_t)r.l[k] - k2P[k] - br;
      r.l[k] = (uint32_t)d;
      br = (d >> 32) & 1;
    }
  }
  fr_store(out + e, r);
}

}  // namespace

void scalar_mul_many_dev(const fr_t* a, const fr_t* b, fr_t* out, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_scalar_mul<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, b, out, n);
  PK_HIP(hipGetLastError());
}

}  // namespace pk
