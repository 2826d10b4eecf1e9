// Seeded synthetic field elements generated on the device (no counterpart in
// the reference, whose benchmarks draw Fr::rand from rand_chacha on the host).
// Element i is the ChaCha20 block (RFC 8439, 20 rounds, nonce 0) with block
// counter first + i under a caller-given 256-bit key, read as a 512-bit
// little-endian integer and reduced mod p; the canonical residue v is used
// directly as the Montgomery limbs (the field element v 2^-256), as the host
// generator does.  One lane per element, no rejection, so any sub-range can
// be generated on its own; the reduction's bias is below 2^-250.
#include <string.h>

#include "kernels.hpp"

namespace pk {

namespace {

// 2^512 mod p, canonical (R^2 in Montgomery terms)
__constant__ uint32_t kR2[8] = {0xae216da7, 0x1bb8e645, 0xe35c59e3, 0x53fe3ab1,
                                0x53bb8085, 0x8c49833d, 0x7f4e44a5, 0x0216d0b1};

__device__ __forceinline__ uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

__device__ __forceinline__ void qr(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
  a += b; d ^= a; d = rotl(d, 16);
  c += d; b ^= c; b = rotl(b, 12);
  a += b; d ^= a; d = rotl(d, 8);
  c += d; b ^= c; b = rotl(b, 7);
}

struct Key {
  uint32_t w[8];
};

__global__ __launch_bounds__(256) void k_random_fr(Key key, uint32_t first, uint64_t n,
                                                   fr_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t st[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574,
                     key.w[0], key.w[1], key.w[2], key.w[3],
                     key.w[4], key.w[5], key.w[6], key.w[7],
                     first + (uint32_t)i, 0, 0, 0};
  uint32_t x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = st[k];
#pragma unroll
  for (int r = 0; r < 10; r++) {
    qr(x[0], x[4], x[8], x[12]);
    qr(x[1], x[5], x[9], x[13]);
    qr(x[2], x[6], x[10], x[14]);
    qr(x[3], x[7], x[11], x[15]);
    qr(x[0], x[5], x[10], x[15]);
    qr(x[1], x[6], x[11], x[12]);
    qr(x[2], x[7], x[8], x[13]);
    qr(x[3], x[4], x[9], x[14]);
  }
  fr_t lo, hi, c;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    lo.l[k] = x[k] + st[k];
    hi.l[k] = x[8 + k] + st[8 + k];
    c.l[k] = kR2[k];
  }
  // each half below 2^256 < 6p: five conditional subtractions make it canonical
#pragma unroll
  for (int k = 0; k < 5; k++) {
    fr_canon(lo);
    fr_canon(hi);
  }
  fr_t h = fr_mul(hi, c);  // hi 2^256 mod p, below 2p
  fr_canon(h);
  fr_store(out + i, fr_add_canon(lo, h));
}

}  // namespace

void random_fr_dev(const uint8_t key[32], uint64_t first, uint64_t n, fr_t* out, hipStream_t s) {
  PK_REQUIRE(first + n <= (1ull << 32), PK_ERR_SIZE,
             "random_fr: block counters past 2^32 (RFC 8439's 32-bit counter)");
  if (!n) return;
  Key k;
  memcpy(k.w, key, 32);
  k_random_fr<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(k, (uint32_t)first, n, out);
  PK_HIP(hipGetLastError());
}

}  // namespace pk
