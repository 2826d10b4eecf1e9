/* ORACLE — test infrastructure only.  CPU restatement of the Skyscraper
 * proof-of-work grinding (skyscraper/core/src/pow.rs:18-100 and the solver
 * loop of skyscraper/core/src/generic.rs:42-71), used by
 * provekit/common/src/skyscraper/pow.rs:14-36 as spongefish's PowStrategy.
 *
 *  - threshold(d) = f64_to_u256(2^-d * (p's top limb as f64) * 2^192);
 *  - verify(ch, d, nonce): d == 0, or compress(ch, [nonce, 0, 0, 0]) <
 *    threshold(d) as 256-bit integers;
 *  - solve(ch, d): 0 if d == 0, else the smallest nonce whose compression is
 *    below threshold(d + 0.01).  The reference's threads scan interleaved
 *    chunks and keep the minimum hit; every chunk below the final minimum is
 *    scanned in full, so its answer is the smallest such nonce.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>
#include "oracle.h"

/* f64_parts + f64_to_u256 (pow.rs:47-88): the float's value rounded to the
 * nearest integer (halves away from zero), clamped to [0, 2^256 - 1]; NaN
 * maps to the maximum. */
void or_f64_to_u256(double f, uint64_t out[4]) {
  uint64_t bits;
  memcpy(&bits, &f, 8);
  const int sign = (int)(bits >> 63);
  const int exp_bits = (int)((bits >> 52) & 0x7ff);
  const uint64_t frac = bits & ((1ULL << 52) - 1);
  const int exp = exp_bits == 0 ? -1022 : exp_bits - 1023;
  const uint64_t sig = exp_bits == 0 ? frac : frac + (1ULL << 52);
  memset(out, 0, 32);
  if (sign) return;
  if (exp > 256) {
    out[0] = out[1] = out[2] = out[3] = ~0ULL;
    return;
  }
  const int shift = exp - 52;
  if (shift < 0) {
    out[0] = (uint64_t)round(f);
    return;
  }
  const int limb = shift / 64, s = shift % 64;
  out[limb] = sig << s;
  if (s != 0 && limb < 3) out[limb + 1] = sig >> (64 - s);
}

/* threshold (pow.rs:18-26); -1 outside [0, 80) */
int or_pow_threshold(double difficulty, uint64_t out[4]) {
  if (!(difficulty >= 0.0 && difficulty < 80.0)) return -1;
  const double modulus = (double)0x30644e72e131a029ULL * ldexp(1.0, 192);
  const double prob = exp2(-difficulty);
  or_f64_to_u256(prob * modulus, out);
  return 0;
}

/* arithmetic::less_than (skyscraper/core/src/arithmetic.rs:50-59) */
static int lt256(const uint64_t a[4], const uint64_t b[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] < b[i]) return 1;
    if (a[i] > b[i]) return 0;
  }
  return 0;
}

static int below(const uint64_t ch[4], uint64_t nonce, const uint64_t thr[4]) {
  const uint64_t r[4] = {nonce, 0, 0, 0};
  uint64_t h[4];
  or_sky_compress(ch, r, h);
  return lt256(h, thr);
}

/* verify (pow.rs:28-30); -1 on a bad difficulty */
int or_pow_verify(const uint64_t challenge[4], double difficulty, uint64_t nonce) {
  if (difficulty == 0.0) return 1;
  uint64_t thr[4];
  if (or_pow_threshold(difficulty, thr)) return -1;
  return below(challenge, nonce, thr);
}

/* solve (pow.rs:36-45, generic.rs:42-71): the smallest nonce below `limit`
 * that passes threshold(d + 0.01).  Returns 1 and sets *nonce when found, 0
 * when no nonce below `limit` passes, -1 on a bad difficulty.  Chunks of
 * 4096 nonces are scanned by the OpenMP threads in order, several at a time;
 * the first chunk batch holding a hit gives the minimum. */
int or_pow_solve(const uint64_t challenge[4], double difficulty, uint64_t limit,
                 uint64_t* nonce) {
  if (difficulty == 0.0) {
    *nonce = 0;
    return 1;
  }
  uint64_t thr[4];
  if (or_pow_threshold(difficulty + 0.01, thr)) return -1;
  const uint64_t chunk = 4096;
  const int nt = or_get_threads();
  const uint64_t batch = chunk * (uint64_t)nt;
  for (uint64_t base = 0; base < limit; base += batch) {
    uint64_t best = UINT64_MAX;
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(min : best)
    for (int t = 0; t < nt; t++) {
      const uint64_t lo = base + (uint64_t)t * chunk;
      for (uint64_t n = lo; n < lo + chunk && n < limit; n++)
        if (below(challenge, n, thr)) {
          if (n < best) best = n;
          break;
        }
    }
    if (best != UINT64_MAX) {
      *nonce = best;
      return 1;
    }
  }
  return 0;
}
