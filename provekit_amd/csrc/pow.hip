// Skyscraper proof-of-work grinding on gfx950 (skyscraper/core/src/pow.rs:18-45
// with the solver of skyscraper/core/src/generic.rs:42-71; spongefish's
// PowStrategy in provekit/common/src/skyscraper/pow.rs:14-36).
//
// solve(ch, d) is the smallest nonce with compress(ch, [nonce, 0, 0, 0]) below
// threshold(d + 0.01).  The search runs over windows of nonces in increasing
// order, one launch per window and one lane per nonce.  A workgroup whose
// first nonce is above the smallest hit so far returns at once; every other
// workgroup hashes all of its nonces, so when a window holds a hit, every
// nonce below the smallest one has been hashed and the answer is exact.  The
// host stops at the first window with a hit.  Windows start small (the usual
// difficulty, 10 bits, needs about a thousand hashes) and grow four-fold.
#include <cmath>
#include <cstring>

#include "kernels.hpp"
#include "sky.hpp"

namespace pk {

namespace {

// a < b as 256-bit integers (skyscraper/core/src/arithmetic.rs:50-59)
__device__ __forceinline__ bool lt256(const fr_t& a, const fr_t& b) {
  bool lt = false, decided = false;
#pragma unroll
  for (int i = 7; i >= 0; i--) {
    lt = decided ? lt : a.l[i] < b.l[i];
    decided = decided || a.l[i] != b.l[i];
  }
  return decided && lt;
}

__global__ __launch_bounds__(256) void k_pow_window(fr_t challenge, fr_t thr, uint64_t base,
                                                    uint64_t count,
                                                    unsigned long long* best) {
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const uint64_t first = base + (uint64_t)blockIdx.x * blockDim.x;
  if (first > __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint64_t off = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (off >= count) return;
  const uint64_t nonce = base + off;
  fr_t r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = 0;
  r.l[0] = (uint32_t)nonce;
  r.l[1] = (uint32_t)(nonce >> 32);
  const fr_t h = sky_compress_t(challenge, r, tab);
  if (lt256(h, thr)) atomicMin(best, (unsigned long long)nonce);
}

}  // namespace

void f64_to_u256(double f, uint64_t out[4]) {
  uint64_t bits;
  memcpy(&bits, &f, 8);
  const bool neg = (bits >> 63) != 0;
  const int eb = (int)((bits >> 52) & 0x7ff);
  const uint64_t frac = bits & ((1ULL << 52) - 1);
  const int e = eb == 0 ? -1022 : eb - 1023;
  const uint64_t sig = eb == 0 ? frac : frac | (1ULL << 52);
  out[0] = out[1] = out[2] = out[3] = 0;
  if (neg) return;
  if (e > 256) {  // also infinity and NaN
    out[0] = out[1] = out[2] = out[3] = ~0ULL;
    return;
  }
  const int sh = e - 52;
  if (sh < 0) {
    out[0] = (uint64_t)std::round(f);  // halves away from zero, as Rust's f64::round
    return;
  }
  const int w = sh / 64, b = sh % 64;
  out[w] = sig << b;
  if (b && w < 3) out[w + 1] = sig >> (64 - b);
}

void pow_threshold(double difficulty, uint64_t out[4]) {
  PK_REQUIRE(difficulty >= 0.0 && difficulty < 80.0, PK_ERR_ARG,
             "Difficulty must be in the range [0, 80)");
  // p's top limb as f64, scaled by 2^192 (pow.rs:23)
  const double modulus = (double)0x30644e72e131a029ULL * std::ldexp(1.0, 192);
  f64_to_u256(std::exp2(-difficulty) * modulus, out);
}

bool sky_pow_solve_dev(const uint64_t challenge[4], double difficulty, uint64_t limit,
                       uint64_t* nonce, hipStream_t s) {
  if (difficulty == 0.0) {
    *nonce = 0;
    return true;
  }
  uint64_t thr64[4];
  pow_threshold(difficulty + 0.01, thr64);  // the prover's bias (pow.rs:10,40)
  fr_t ch, thr;
  memcpy(ch.l, challenge, 32);
  memcpy(thr.l, thr64, 32);
  Scratch best(sizeof(unsigned long long), s);
  PK_HIP(hipMemsetAsync(best.p, 0xff, sizeof(unsigned long long), s));
  unsigned long long host = ~0ULL;
  uint64_t window = 1u << 16;
  for (uint64_t base = 0; base < limit;) {
    const uint64_t count = std::min(window, limit - base);
    k_pow_window<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(
        ch, thr, base, count, best.as<unsigned long long>());
    PK_HIP(hipGetLastError());
    PK_HIP(hipMemcpyAsync(&host, best.p, sizeof host, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    if (host != ~0ULL) {
      *nonce = host;
      return true;
    }
    base += count;
    window = std::min<uint64_t>(window * 4, 1ull << 26);
  }
  return false;
}

}  // namespace pk
