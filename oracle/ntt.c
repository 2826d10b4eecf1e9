/* ORACLE — test infrastructure only.  CPU restatement of the reference NTT
 * (ntt/src/ntt.rs) and of the RSFr Reed-Solomon plugin
 * (provekit/common/src/ntt.rs), used to check the HIP path and as the timed
 * CPU baseline.  OpenMP loops stand in for rayon at the same split points.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ntt/src/ntt.rs:18-21 — 32 KiB of Fr */
#define WORKLOAD_SIZE 1024

static unsigned ilog2(size_t n) {
  unsigned k = 0;
  while (((size_t)1 << k) < n) k++;
  return k;
}
/* ntt/src/ntt.rs:255-260 */
static size_t reverse_bits(size_t v, unsigned bits) {
  size_t r = 0;
  for (unsigned i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}

/* The 2^28-th root of unity ark-bn254 uses (TWO_ADIC_ROOT_OF_UNITY =
 * GENERATOR^T with GENERATOR = 5, T = (p-1)/2^28), as an integer:
 * 0x2a3c09f0a58a7e8500e0a7eb8ef62abc402d111e41112ed49bd61b6e725b19f0
 * = 19103219067921713944291392827692070036145651957329286315305642004821462161904.
 * get_root_of_unity(n) = that ^ (2^28 / n). */
ofr or_root_of_unity(size_t n) {
  ofr g = {{0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL,
            0x2a3c09f0a58a7e85ULL}};
  g = ofr_from_int(g);
  unsigned k = ilog2(n);
  for (unsigned i = k; i < 28; i++) g = ofr_mul(g, g);
  return g;
}

/* init_roots_reverse_ordered (ntt/src/ntt.rs:275-306) */
void or_roots_table(size_t order, ofr* out) {
  size_t n = order / 2;
  if (n == 0) return;
  if (n == 1) {
    out[0] = OFR_ONE;
    return;
  }
  ofr root = or_root_of_unity(order);
  unsigned bits = ilog2(n);
  ofr w = OFR_ONE;
  for (size_t i = 0; i < n; i++) {
    out[reverse_bits(i, bits)] = w;
    w = ofr_mul(w, root);
  }
}

static ofr* g_roots = NULL;
static size_t g_roots_order = 0;
/* NTTEngine / extend_roots_table (ntt/src/ntt.rs:24-125): a process-global,
 * prefix-consistent table, grown on demand.  Initial order is
 * workload_size.next_power_of_two(). */
static const ofr* roots_for(size_t codeword_size) {
  size_t want = codeword_size < WORKLOAD_SIZE ? WORKLOAD_SIZE : codeword_size;
  const ofr* cur;
#pragma omp critical(or_roots)
  {
    if (g_roots_order < want) {
      /* the smaller table is kept, not freed: a transform on another thread
       * may still be reading it (the reference's readers hold its lock) */
      ofr* t = (ofr*)malloc(sizeof(ofr) * (want / 2));
      or_roots_table(want, t);
      g_roots = t;
      g_roots_order = want;
    }
    cur = g_roots;
  }
  return cur;
}

static inline void butterfly(ofr* e, ofr* o, ofr w) {
  ofr t = ofr_mul(w, *o);
  ofr a = ofr_add(*e, t);
  ofr b = ofr_sub(*e, t);
  *e = a;
  *o = b;
}

/* dit_nr_cache (ntt/src/ntt.rs:229-248) */
static void dit_nr_cache(const ofr* rr, size_t segment, ofr* input, size_t len, size_t size) {
  size_t elements_in_group = len;
  size_t num_of_groups = 1;
  while (num_of_groups < size) {
    size_t twiddle_base = segment * num_of_groups;
    for (size_t k = 0; k < len / elements_in_group; k++) {
      ofr w = rr[twiddle_base + k];
      ofr* g = input + k * elements_in_group;
      size_t h = elements_in_group / 2;
      for (size_t i = 0; i < h; i++) butterfly(&g[i], &g[i + h], w);
    }
    elements_in_group /= 2;
    num_of_groups *= 2;
  }
}

/* interleaved_ntt_nr (ntt/src/ntt.rs:150-227) */
static void interleaved_ntt_nr(const ofr* rr, ofr* values, size_t len, size_t codeword_size,
                               size_t num_groups) {
  if (codeword_size <= 1) return;
  size_t elements_in_group = len / num_groups;
  int nt = or_get_threads();
  size_t lim = codeword_size < 32 ? codeword_size : 32;
  /* ntt.rs:189-203: sequential over groups, parallel over pairs */
  while (num_groups < lim && elements_in_group > WORKLOAD_SIZE) {
    for (size_t k = 0; k < num_groups; k++) {
      ofr w = rr[k];
      ofr* g = values + k * elements_in_group;
      size_t h = elements_in_group / 2;
#pragma omp parallel for num_threads(nt) schedule(static)
      for (size_t i = 0; i < h; i++) butterfly(&g[i], &g[i + h], w);
    }
    elements_in_group /= 2;
    num_groups *= 2;
  }
  /* ntt.rs:205-219: parallel over groups */
  while (num_groups < codeword_size && elements_in_group > WORKLOAD_SIZE) {
    size_t h = elements_in_group / 2;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < num_groups; k++) {
      ofr w = rr[k];
      ofr* g = values + k * elements_in_group;
      for (size_t i = 0; i < h; i++) butterfly(&g[i], &g[i + h], w);
    }
    elements_in_group /= 2;
    num_groups *= 2;
  }
  /* ntt.rs:221-226 */
  size_t chunks = len / elements_in_group;
  size_t sub = codeword_size / num_groups;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (size_t k = 0; k < chunks; k++)
    dit_nr_cache(rr, k, values + k * elements_in_group, elements_in_group, sub);
}

/* ntt_nr (ntt/src/ntt.rs:105-108) */
void or_ntt_nr(ofr* values, size_t len, size_t codeword_size, size_t num_groups) {
  if (codeword_size <= 1) return;
  const ofr* rr = roots_for(codeword_size);
  interleaved_ntt_nr(rr, values, len, codeword_size, num_groups);
}

/* reverse_order (ntt/src/ntt.rs:310-323): the first n entries (n = the
 * codeword size, a power of two) in bit-reversed order. */
void or_reverse_order(ofr* values, size_t n) {
  if (n <= 1) return;
  unsigned bits = ilog2(n);
  for (size_t i = 0; i < n; i++) {
    size_t r = reverse_bits(i, bits);
    if (i < r) {
      ofr t = values[i];
      values[i] = values[r];
      values[r] = t;
    }
  }
}

/* intt_nr (ntt/src/ntt.rs:333-348): natural-order evaluations ->
 * coefficients in bit-reversed order. */
void or_intt_nr(ofr* values, size_t n) {
  if (n == 0) return;
  for (size_t i = 1, j = n - 1; i < j; i++, j--) {
    ofr t = values[i];
    values[i] = values[j];
    values[j] = t;
  }
  or_ntt_nr(values, n, n, 1);
  ofr factor = ofr_inv(ofr_from_u64(n));
  for (size_t i = 0; i < n; i++) values[i] = ofr_mul(values[i], factor);
}

/* intt_rn (ntt/src/ntt.rs:326-330) */
void or_intt_rn(ofr* values, size_t len) {
  or_reverse_order(values, len);
  or_intt_nr(values, len);
  or_reverse_order(values, len);
}

/* RSFr::next_order (provekit/common/src/ntt.rs:13-20) */
int or_rs_next_order(size_t size, size_t* order) {
  size_t o = 1;
  while (o < size) o <<= 1;
  if (o > ((size_t)1 << 28)) return 0;
  *order = o;
  return 1;
}

/* RSFr::evaluation_points (provekit/common/src/ntt.rs:22-42) */
void or_rs_evaluation_points(size_t codeword_length, const size_t* indices, size_t n, ofr* out) {
  unsigned bits = 0;
  if (codeword_length > 1) {
    size_t x = codeword_length - 1;
    while (x) {
      bits++;
      x >>= 1;
    }
  }
  ofr g = or_root_of_unity(codeword_length);
  for (size_t t = 0; t < n; t++) {
    size_t k = bits == 0 ? indices[t] : reverse_bits(indices[t], bits);
    out[t] = ofr_pow(g, k);
  }
}

/* RSFr::interleaved_encode (provekit/common/src/ntt.rs:51-100) */
int or_rs_interleaved_encode(const ofr* messages, size_t num_messages, size_t message_length,
                             const ofr* masks, size_t masks_len, size_t codeword_length,
                             ofr* out) {
  if (num_messages == 0) return 0;
  size_t total = num_messages * codeword_length;
  /* the reference indexes result[column * M + row] and slices the masks into
   * the M x codeword_length vector: past its end it panics (provekit/common/src/ntt.rs:70-80) */
  if (message_length > codeword_length || message_length * num_messages + masks_len > total) return -1;
  if (codeword_length == 0) return 0; /* empty messages, empty codeword */
  memset(out, 0, sizeof(ofr) * total);
  for (size_t col = 0; col < message_length; col++)
    for (size_t row = 0; row < num_messages; row++)
      out[col * num_messages + row] = messages[row * message_length + col];
  memcpy(out + message_length * num_messages, masks, sizeof(ofr) * masks_len);
  size_t mask_length = masks_len / num_messages;
  size_t coset_size;
  if (!or_rs_next_order(message_length + mask_length, &coset_size)) return -1;
  while (codeword_length % coset_size) {
    if (!or_rs_next_order(coset_size + 1, &coset_size)) return -1;
  }
  /* ntt_nr asserts a power-of-two codeword above 1 (ntt/src/ntt.rs:160-169) */
  if (codeword_length > 1 && (codeword_length & (codeword_length - 1))) return -1;
  size_t num_cosets = codeword_length / coset_size;
  size_t chunk = coset_size * num_messages;
  for (size_t k = 1; k < num_cosets; k++) memmove(out + k * chunk, out, sizeof(ofr) * chunk);
  or_ntt_nr(out, total, codeword_length, num_cosets);
  return 0;
}
