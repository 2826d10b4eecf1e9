// R1CS products and the ZK sumcheck on gfx950.
//
//  * k_spmv: HydratedSparseMatrix * &[Fr] (provekit/common/src/sparse_matrix.rs:606-624),
//    one lane per row (R1CS rows hold a handful of entries), values already
//    expanded from the interner.
//  * eq table (provekit/common/src/utils/sumcheck.rs:108-153): eq(r, x) with
//    r[0] on the top index bit, as eq_hi[x >> k] * eq_lo[x & (2^k - 1)] from two
//    small tables, one product per entry.
//  * k_sc_round: sumcheck_fold_map_reduce with the R1CS map
//    (sumcheck.rs:14-102, provekit/prover/src/whir_r1cs.rs:552-559): optional
//    in-place fold of the four MLEs, the three map outputs, a workgroup sum in
//    LDS, one partial per workgroup.
//  * k_sc_finalize: sums the partials, forms the round polynomial
//    (whir_r1cs.rs:567-609, with compute_blinding_coefficients_for_round
//    :427-492), runs the transcript step and writes the next challenge, in
//    one workgroup, so a whole sumcheck runs without the host.
// The transcript is the project's Skyscraper transcript (init / absorb1 /
// round step / squeeze with parallel compressions, defined in sumcheck.hpp).
#include <algorithm>
#include <map>
#include <mutex>

#include "host_fr.hpp"
#include "kernels.hpp"
#include "sky.hpp"
#include "sumcheck.hpp"

namespace pk {

namespace {

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ fr_t fr_zero() {
  fr_t z;
#pragma unroll
  for (int i = 0; i < 8; i++) z.l[i] = 0;
  return z;
}

// ------------------------------------------------------------------ SpMV
// Device-resident matrices are not validated on the host, so the kernels
// keep every access in bounds whatever the indices: row ranges are clipped
// to [0, nnz] and entries with a column >= ncols are skipped (a malformed
// matrix gives an unspecified result, never an out-of-bounds read).
//
// Rows of at most kShortRow entries (R1CS rows hold a handful) take one lane
// each.  A longer row on one lane would be a straggler (a transposed matrix
// has a row per witness column, and the constant-one column can appear in
// most constraints), so long rows are summed by entry chunks instead: the
// row's lane flags the kChunk-entry chunks it covers, one workgroup per
// flagged chunk sums every long row's entries inside the chunk (rows that
// start and end inside it are finished there; a row crossing the chunk's
// first entry leaves a head partial, one crossing its last entry a tail
// partial), and a last pass adds the partials of each row that spans
// chunks.  The lanes of long rows list the spanning ones, so no host
// round-trip is needed.
constexpr uint64_t kShortRow = 32;
constexpr uint64_t kChunk = 2048;
constexpr int kMaxChunkRows = (int)(kChunk / (kShortRow + 1)) + 2;  // long rows meeting one chunk

__device__ __forceinline__ void row_range(const uint32_t* row_starts, uint64_t rows, uint64_t nnz,
                                          uint64_t r, uint64_t& b, uint64_t& e) {
  e = (r + 1 < rows) ? min((uint64_t)row_starts[r + 1], nnz) : nnz;
  b = min((uint64_t)row_starts[r], e);
}

// Where entry i's value comes from: hydrated (one Fr per entry, 32 bytes
// streamed per entry) or interned, as the reference stores a SparseMatrix
// (sparse_matrix.rs:56-71: an id per entry into the Interner's table,
// interner.rs:7-42, hydrated at use, :583-592): 4 bytes per entry and a
// table that real constraint systems keep small (a few distinct
// coefficients), so it stays in cache.  An id past the table reads as zero
// (device data is not validated; never an out-of-bounds read).
struct ValsHydrated {
  const fr_t* v;
  __device__ __forceinline__ fr_t operator()(uint64_t i) const { return fr_load(v + i); }
};
struct ValsInterned {
  const uint32_t* id;
  const fr_t* table;
  uint64_t n;
  __device__ __forceinline__ fr_t operator()(uint64_t i) const {
    const uint32_t k = id[i];
    return k < n ? fr_load(table + k) : fr_zero();
  }
};

// sum of v[i] z[col[i]] over [b, e), one lane
template <class V>
__device__ __forceinline__ fr_t row_dot(uint64_t b, uint64_t e, uint64_t ncols, const uint32_t* cols,
                                        const V& vals, const fr_t* z) {
  fr_t acc = fr_zero();
  // take the entries four at a time, with all index, value and gather loads
  // issued before the products, so a lane keeps several independent HBM
  // requests in flight instead of one chain (index -> gather -> product)
  constexpr int G = 4;
  for (uint64_t i0 = b; i0 < e; i0 += G) {
    uint32_t c[G];
    fr_t v[G], x[G];
#pragma unroll
    for (int k = 0; k < G; k++) c[k] = (i0 + k < e) ? cols[i0 + k] : 0xffffffffu;
#pragma unroll
    for (int k = 0; k < G; k++) {
      if (c[k] < ncols) {
        v[k] = vals(i0 + k);
        x[k] = fr_load(z + c[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < G; k++)
      if (c[k] < ncols) acc = fr_add(acc, fr_mul(v[k], x[k]));
  }
  return acc;
}

// spans: [0] = count, then the spanning long rows (capacity n_chunks)
template <class V>
__global__ __launch_bounds__(256) void k_spmv(uint64_t rows, uint64_t ncols,
                                              const uint32_t* row_starts, uint64_t nnz,
                                              const uint32_t* cols, V vals,
                                              const fr_t* z, fr_t* out, uint32_t* chunk_flags,
                                              uint32_t* spans, uint64_t n_chunks) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  uint64_t b, e;
  row_range(row_starts, rows, nnz, r, b, e);
  if (e - b > kShortRow) {
    const uint64_t c0 = b / kChunk, c1 = (e - 1) / kChunk;
    for (uint64_t c = c0; c <= c1 && c < n_chunks; c++) chunk_flags[c] = 1u;
    if (c1 > c0) {
      const uint32_t slot = atomicAdd(spans, 1u);
      if (slot < n_chunks) spans[1 + slot] = (uint32_t)r;
    }
    return;
  }
  fr_t acc = row_dot(b, e, ncols, cols, vals, z);
  fr_canon(acc);
  fr_store(out + r, acc);
}

// Workgroup sum (256 threads) of one value per thread; valid on thread 0.
__device__ __forceinline__ fr_t block_sum1(fr_t v) {
  __shared__ fr_t part[4];
  const int t = threadIdx.x, lane = t & 63;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) v = fr_add(v, fr_shfl(v, lane ^ w));
  __syncthreads();  // part may still be read from a previous call
  if (lane == 0) part[t >> 6] = v;
  __syncthreads();
  fr_t s = fr_zero();
  if (t == 0) s = fr_add(fr_add(part[0], part[1]), fr_add(part[2], part[3]));
  return s;
}

template <class V>
__global__ __launch_bounds__(256) void k_spmv_chunks(uint64_t rows, uint64_t ncols,
                                                     const uint32_t* row_starts, uint64_t nnz,
                                                     const uint32_t* cols, V vals,
                                                     const fr_t* z, fr_t* out,
                                                     const uint32_t* chunk_flags, fr_t* head,
                                                     fr_t* tail, uint64_t n_chunks) {
  __shared__ uint64_t r_lo;
  __shared__ uint32_t list[kMaxChunkRows];
  __shared__ int nlist;
  const int t = threadIdx.x;
  // a grid of at most 1024 workgroups strides over the chunks: most chunks
  // of most matrices have no long row, and a workgroup per chunk would cost
  // more to launch than to skip
  for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    if (!chunk_flags[c]) continue;
    const uint64_t c0 = c * kChunk, c1 = min(c0 + kChunk, nnz);
    if (t == 0) {
      // the last row starting at or before c0 (0 if none)
      uint64_t lo = 0, hi = rows;
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (row_starts[mid] <= c0)
          lo = mid;
        else
          hi = mid;
      }
      r_lo = lo;
      nlist = 0;
    }
    __syncthreads();
    for (uint64_t rr = r_lo + t; rr < rows; rr += blockDim.x) {
      uint64_t b, e;
      row_range(row_starts, rows, nnz, rr, b, e);
      if (b >= c1) break;
      if (e > c0 && e - b > kShortRow) {
        const int i = atomicAdd(&nlist, 1);
        if (i < kMaxChunkRows) list[i] = (uint32_t)rr;
      }
    }
    __syncthreads();
    const int n = min(nlist, kMaxChunkRows);
    for (int k = 0; k < n; k++) {
      const uint64_t rr = list[k];
      uint64_t b, e;
      row_range(row_starts, rows, nnz, rr, b, e);
      const uint64_t lo = max(b, c0), hi = min(e, c1);
      fr_t acc = fr_zero();
      for (uint64_t i = lo + t; i < hi; i += blockDim.x) {
        const uint32_t cc = cols[i];
        if (cc < ncols) acc = fr_add(acc, fr_mul(vals(i), fr_load(z + cc)));
      }
      fr_t sum = block_sum1(acc);
      if (t == 0) {
        fr_canon(sum);
        if (b >= c0 && e <= c1)
          fr_store(out + rr, sum);   // the whole row is in this chunk
        else if (b < c0)
          fr_store(head + c, sum);   // continues from earlier chunks
        else
          fr_store(tail + c, sum);   // starts here, continues after
      }
    }
    __syncthreads();  // r_lo, list and nlist are rewritten for the next chunk
  }
}

// Rows spanning chunks: the tail partial of the first chunk plus the head
// partials of the others.
__global__ __launch_bounds__(256) void k_spmv_spans(uint64_t rows, const uint32_t* row_starts,
                                                    uint64_t nnz, const uint32_t* spans,
                                                    uint64_t n_chunks, const fr_t* head,
                                                    const fr_t* tail, fr_t* out) {
  const uint64_t count = min((uint64_t)spans[0], n_chunks);
  const int t = threadIdx.x;
  for (uint64_t j = blockIdx.x; j < count; j += gridDim.x) {
    const uint64_t rr = min((uint64_t)spans[1 + j], rows - 1);
    uint64_t b, e;
    row_range(row_starts, rows, nnz, rr, b, e);
    fr_t acc = fr_zero();
    if (e > b) {
      const uint64_t cf = b / kChunk, cl = min((e - 1) / kChunk, n_chunks - 1);
      for (uint64_t cc = cf + 1 + t; cc <= cl; cc += blockDim.x) acc = fr_add(acc, fr_load(head + cc));
      if (t == 0 && cf < n_chunks) acc = fr_add(acc, fr_load(tail + cf));
    }
    fr_t sum = block_sum1(acc);
    if (t == 0) {
      fr_canon(sum);
      fr_store(out + rr, sum);
    }
  }
}

__global__ __launch_bounds__(256) void k_spmv_clear(uint32_t* flags, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = 0u;
}

// c[i] = a[i] * b[i] for i < n, zero up to n_pad
__global__ __launch_bounds__(256) void k_hadamard(const fr_t* a, const fr_t* b, fr_t* c,
                                                  uint64_t n, uint64_t n_pad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  fr_t v = fr_zero();
  if (i < n) {
    v = fr_mul(fr_load(a + i), fr_load(b + i));
    fr_canon(v);
  }
  fr_store(c + i, v);
}

__global__ void k_zero(fr_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fr_store(p + i, fr_zero());
}

// ------------------------------------------------------------------ eq
// out[x] = prod_j (bit_{k-1-j}(x) ? r[j] : 1 - r[j]) for x < n, k variables.
__global__ __launch_bounds__(256) void k_eq_direct(const fr_t* r, int k, uint64_t n, fr_t* out,
                                                   fr_t one) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  fr_t acc = one;
  for (int j = 0; j < k; j++) {
    const fr_t rj = fr_load(r + j);
    const fr_t f = ((x >> (k - 1 - j)) & 1) ? rj : fr_sub(one, rj);
    acc = fr_mul(acc, f);
  }
  fr_canon(acc);
  fr_store(out + x, acc);
}
__global__ __launch_bounds__(256) void k_eq_combine(const fr_t* hi, const fr_t* lo, int lo_bits,
                                                    uint64_t n, fr_t* out) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  fr_t v = fr_mul(fr_load(hi + (x >> lo_bits)), fr_load(lo + (x & ((1ull << lo_bits) - 1))));
  fr_canon(v);
  fr_store(out + x, v);
}

// ------------------------------------------------------------------ rounds
constexpr int kRoundThreads = 256;
constexpr unsigned kMaxRoundBlocks = 1024;

// Workgroup sum of NV values per lane; lane 0 writes them (canonical) to dst.
template <int NV>
__device__ __forceinline__ void block_sum(const fr_t (&v)[NV], fr_t* dst) {
  __shared__ fr_t sh[NV][kRoundThreads];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NV; q++) sh[q][t] = v[q];
  __syncthreads();
  for (int w = kRoundThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < NV; q++) sh[q][t] = fr_add(sh[q][t], sh[q][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      fr_t x = sh[q][0];
      fr_canon(x);
      fr_store(dst + q, x);
    }
  }
}
__device__ __forceinline__ void block_sum3(fr_t v0, fr_t v1, fr_t v2, fr_t* dst) {
  const fr_t v[3] = {v0, v1, v2};
  block_sum<3>(v, dst);
}

// One sumcheck_fold_map_reduce call.  fold == nullptr: map over the halves of
// size; else fold the quarters in place with *fold first, map over (p0, p1).
// With WITH_CLAIM (round 0 of the plain mode, no fold) each workgroup also
// writes sum eq1 (a1 b1 - c1) over its high halves as a fourth partial: the
// claim is h(0) + h(1) = sum_x eq (a b - c), so no separate pass over the
// four MLEs is needed for it.
__device__ __forceinline__ fr_t r1cs_term(const fr_t& e, const fr_t& a, const fr_t& b, const fr_t& c) {
  return fr_mul(e, fr_sub(fr_mul(a, b), c));
}

template <bool WITH_CLAIM>
__global__ __launch_bounds__(kRoundThreads) void k_sc_round(fr_t* a, fr_t* b, fr_t* c, fr_t* eq,
                                                            uint64_t size, const fr_t* fold,
                                                            fr_t* partials) {
  const uint64_t h = fold ? size / 4 : size / 2;
  fr_t f;
  if (fold) f = fr_load(fold);
  fr_t s0 = fr_zero(), s1 = fr_zero(), s2 = fr_zero(), s3 = fr_zero();
  fr_t* mle[4] = {a, b, c, eq};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < h;
       i += (uint64_t)gridDim.x * blockDim.x) {
    fr_t lo[4], hi[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      fr_t x0 = fr_load(mle[j] + i), x1 = fr_load(mle[j] + i + h);
      if (fold) {
        const fr_t x2 = fr_load(mle[j] + i + 2 * h), x3 = fr_load(mle[j] + i + 3 * h);
        x0 = fr_add(x0, fr_mul(f, fr_sub(x2, x0)));
        x1 = fr_add(x1, fr_mul(f, fr_sub(x3, x1)));
        fr_t w0 = x0, w1 = x1;
        fr_canon(w0);
        fr_canon(w1);
        fr_store(mle[j] + i, w0);
        fr_store(mle[j] + i + h, w1);
      }
      lo[j] = x0;
      hi[j] = x1;
    }
    // [eq0 (a0 b0 - c0), (2eq0 - eq1)((2a0 - a1)(2b0 - b1) - (2c0 - c1)),
    //  (eq1 - eq0)(a1 - a0)(b1 - b0)]
    const fr_t f0 = r1cs_term(lo[3], lo[0], lo[1], lo[2]);
    const fr_t ea = fr_sub(fr_add(lo[3], lo[3]), hi[3]);
    const fr_t aa = fr_sub(fr_add(lo[0], lo[0]), hi[0]);
    const fr_t bb = fr_sub(fr_add(lo[1], lo[1]), hi[1]);
    const fr_t cc = fr_sub(fr_add(lo[2], lo[2]), hi[2]);
    const fr_t fm1 = fr_mul(ea, fr_sub(fr_mul(aa, bb), cc));
    const fr_t finf = fr_mul(fr_mul(fr_sub(hi[3], lo[3]), fr_sub(hi[0], lo[0])), fr_sub(hi[1], lo[1]));
    s0 = fr_add(s0, f0);
    s1 = fr_add(s1, fm1);
    s2 = fr_add(s2, finf);
    if (WITH_CLAIM) s3 = fr_add(s3, r1cs_term(hi[3], hi[0], hi[1], hi[2]));
  }
  if (WITH_CLAIM) {
    const fr_t v[4] = {s0, s1, s2, s3};
    block_sum<4>(v, partials + 4 * blockIdx.x);
  } else {
    block_sum3(s0, s1, s2, partials + 3 * blockIdx.x);
  }
}

// sum_x eq(x) (a(x) b(x) - c(x)) partials: the claimed sum of the plain mode.
__global__ __launch_bounds__(kRoundThreads) void k_sc_claim(const fr_t* a, const fr_t* b,
                                                            const fr_t* c, const fr_t* eq,
                                                            uint64_t size, fr_t* partials) {
  fr_t s0 = fr_zero();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const fr_t v = fr_mul(fr_load(eq + i), fr_sub(fr_mul(fr_load(a + i), fr_load(b + i)), fr_load(c + i)));
    s0 = fr_add(s0, v);
  }
  block_sum3(s0, fr_zero(), fr_zero(), partials + 3 * blockIdx.x);
}

// Sum n partial triples into out[0..3) (one workgroup).
__global__ __launch_bounds__(kRoundThreads) void k_sum_partials(const fr_t* partials, unsigned n,
                                                                fr_t* out) {
  fr_t s0 = fr_zero(), s1 = fr_zero(), s2 = fr_zero();
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    s0 = fr_add(s0, fr_load(partials + 3 * i));
    s1 = fr_add(s1, fr_load(partials + 3 * i + 1));
    s2 = fr_add(s2, fr_load(partials + 3 * i + 2));
  }
  block_sum3(s0, s1, s2, out);
}

// ------------------------------------------------------------------ driver kernels
__device__ __forceinline__ fr_t cubic(const fr_t* p, const fr_t& x) {
  fr_t v = p[3];
  v = fr_add_ilv(p[2], fr_mul(x, v));
  v = fr_add_ilv(p[1], fr_mul(x, v));
  v = fr_add_ilv(p[0], fr_mul(x, v));
  fr_canon(v);
  return v;
}

__device__ __forceinline__ fr_t limb_word(uint32_t w0, uint32_t w2) {
  fr_t z = fr_zero();
  z.l[0] = w0;
  z.l[2] = w2;
  return z;
}

// compute_blinding_coefficients_for_round (whir_r1cs.rs:427-492) for
// compute_for < n (one lane).
__device__ void blinding_coeffs(const ScState* st, int idx, fr_t out[4]) {
  const int n = st->m;
  const fr_t zero = fr_zero();
  fr_t prefix = zero, suffix = zero;
  for (int i = 0; i < idx; i++) prefix = fr_add(prefix, cubic(st->g[i], st->alpha[i]));
  for (int i = idx + 1; i < n; i++) {
    // g_i(0) + g_i(1) = 2 c0 + c1 + c2 + c3
    const fr_t* gi = st->g[i];
    suffix = fr_add(suffix, fr_add(fr_add(fr_add(gi[0], gi[0]), gi[1]), fr_add(gi[2], gi[3])));
  }
  fr_t pm = st->one;
  for (int i = 0; i < n - 1 - idx; i++) pm = fr_add(pm, pm);
  const fr_t sm = fr_mul(pm, st->half);
  const fr_t other = fr_add(fr_mul(pm, prefix), fr_mul(sm, suffix));
  const fr_t* cur = st->g[idx];
  out[0] = fr_add(fr_mul(pm, cur[0]), other);
  out[1] = fr_mul(pm, cur[1]);
  out[2] = fr_mul(pm, cur[2]);
  out[3] = fr_mul(pm, cur[3]);
}

// Transcript init and the m eq-point challenges: one wave, lanes in parallel.
__global__ __launch_bounds__(64) void k_sc_init(ScState* st, uint64_t seed) {
  __shared__ fr_t s;
  const int t = threadIdx.x;
  if (t == 0) {
    fr_t l = fr_zero(), r = fr_zero();
    l.l[0] = (uint32_t)seed;
    l.l[1] = (uint32_t)(seed >> 32);
    r.l[0] = kTrTag0 & 0xffffffffu;
    r.l[1] = (uint32_t)(kTrTag0 >> 32);
    r.l[2] = kTrTag1 & 0xffffffffu;
    r.l[3] = (uint32_t)(kTrTag1 >> 32);
    s = sky_compress(l, r);
  }
  __syncthreads();
  const int m = st->m;
  // squeeze(m): r_j = C(s, [j + 1]) for lanes j < m, lane m: s = C(s, 0),
  // one call for all lanes (operands picked per lane)
  fr_t v;
  if (t <= m) v = sky_compress(s, t < m ? limb_word(t + 1, 0) : fr_zero());
  __syncthreads();
  if (t < m) st->r[t] = v;
  if (t == m) st->state = v;
}

// After the claim reduction: mode P absorbs the computed claim; mode R
// absorbs sum_g, squeezes rho and sets the claim to rho * sum_g.
__global__ __launch_bounds__(64) void k_sc_claim_finalize(ScState* st, const fr_t* claim) {
  __shared__ fr_t s;
  const int t = threadIdx.x;
  if (t == 0) {
    if (st->mode == 0) {
      st->S = claim[0];
      st->rho = fr_zero();
      s = sky_compress(st->state, claim[0]);
    } else {
      // sum_over_hypercube(g) (whir_r1cs.rs:494-501): p(0) + p(1) of round 0's coefficients
      fr_t c[4];
      blinding_coeffs(st, 0, c);
      fr_t sg = fr_add(fr_add(fr_add(c[0], c[0]), c[1]), fr_add(c[2], c[3]));
      fr_canon(sg);
      st->sum_g = sg;
      s = sky_compress(st->state, sg);
    }
  }
  __syncthreads();
  if (st->mode == 1) {
    fr_t v;
    if (t < 2) v = sky_compress(s, t == 0 ? limb_word(1, 0) : fr_zero());
    __syncthreads();
    if (t == 0) {
      st->rho = v;
      fr_t S = fr_mul(v, st->sum_g);
      fr_canon(S);
      st->S = S;
    }
    if (t == 1) st->state = v;
  } else if (t == 0) {
    st->state = s;
  }
}

// After the last round of the reference mode: blinding_eval =
// <expand_powers::<4>(alpha), w1[offset .. offset + 4m]> (whir_r1cs.rs:613-618,
// prefix_covector.rs:156-166).  The commit placed the univariates g at
// w1[offset..], so this is sum_i g_i(alpha_i); it is absorbed like the
// reference's prover_message.  One lane: m cubics and one compression.
__global__ __launch_bounds__(64) void k_sc_blinding_eval(ScState* st) {
  if (threadIdx.x != 0) return;
  const int m = st->m;
  fr_t be = fr_zero();
  for (int i = 0; i < m; i++) be = fr_add(be, cubic(st->g[i], st->alpha[i]));
  fr_canon(be);
  st->blinding_eval = be;
  st->state = sky_compress(st->state, be);
}

// What thread 0 of a finalizing workgroup reads from ScState: loaded once at
// kernel start, next to the partial sums, instead of after them (a lone
// wave would wait out each load), and carried in registers across the rounds
// of k_sc_tail.
struct RoundCtx {
  fr_t state, S, half;
  int mode;
};
__device__ __forceinline__ RoundCtx load_ctx(const ScState* st) {
  RoundCtx c;
  c.state = st->state;
  c.S = st->S;
  c.half = st->half;
  c.mode = st->mode;
  return c;
}

// The end of round idx, run by every thread of the calling workgroup (it
// holds barriers): h0, hm1, hinf are the round's sums and ctx the round's
// state (both read on thread 0).  With `claim_hi` (round 0 of the plain mode)
// thread 0 first takes S = claim = h(0) + h(1) and absorbs it (the order of
// k_sc_claim_finalize).  Then: form the cubic, check c0 + c0 + c1 + c2 + c3 =
// S, the transcript round step (the new state is alpha_idx), S = cubic(alpha).
// Thread 0 leaves the new state and S in ctx and in st.
__device__ __forceinline__ void finalize_round(ScState* st, int idx, RoundCtx& ctx, const fr_t& h0,
                                               const fr_t& hm1, const fr_t& hinf,
                                               const fr_t* claim_hi) {
  __shared__ fr_t co[4];
  __shared__ fr_t hh[2];
  __shared__ fr_t state;
  const int t = threadIdx.x;
  if (t == 0) {
    fr_t stt = ctx.state;
    if (claim_hi) {
      fr_t cl = fr_add_ilv(h0, *claim_hi);
      fr_canon(cl);
      ctx.S = cl;
      st->S = cl;
      st->rho = fr_zero();
      stt = sky_compress(stt, cl);
    }
    state = stt;
    const fr_t S = ctx.S;
    fr_t c0 = h0, at_m1 = hm1, c3 = hinf;
    if (ctx.mode == 1) {
      // the blinding polynomial's part, rho * g (plain mode: rho = 0)
      fr_t g[4];
      blinding_coeffs(st, idx, g);
      const fr_t rho = st->rho;
      const fr_t gm1 = fr_sub_ilv(fr_add_ilv(fr_sub_ilv(g[0], g[1]), g[2]), g[3]);
      c0 = fr_add_ilv(c0, fr_mul(rho, g[0]));
      at_m1 = fr_add_ilv(at_m1, fr_mul(rho, gm1));
      c3 = fr_add_ilv(c3, fr_mul(rho, g[3]));
    }
    const fr_t three_c0 = fr_add_ilv(fr_add_ilv(c0, c0), c0);
    fr_t c2 = fr_mul(ctx.half, fr_sub_ilv(fr_add_ilv(S, at_m1), three_c0));
    fr_t c1 = fr_sub_ilv(fr_sub_ilv(fr_sub_ilv(fr_sub_ilv(S, c0), c0), c3), c2);
    fr_canon(c0);
    fr_canon(c1);
    fr_canon(c2);
    fr_canon(c3);
    co[0] = c0;
    co[1] = c1;
    co[2] = c2;
    co[3] = c3;
    fr_t chk = fr_add_ilv(fr_add_ilv(fr_add_ilv(fr_add_ilv(c0, c0), c1), c2), c3);
    fr_canon(chk);
    fr_t Sc = S;
    fr_canon(Sc);
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 8; q++) ok &= chk.l[q] == Sc.l[q];
    if (!ok) st->error = 1;
    st->coeffs[idx][0] = c0;
    st->coeffs[idx][1] = c1;
    st->coeffs[idx][2] = c2;
    st->coeffs[idx][3] = c3;
  }
  __syncthreads();
  // round step: thread 0 hashes (state, c0), thread 1 (c2, c3), then one
  // more.  The operands are picked per lane before one call, so both lanes
  // run the same instructions at once (two calls on divergent branches would
  // run one after the other).
  if (t < 2) {
    const fr_t x = t == 0 ? state : co[2];
    const fr_t y = t == 0 ? co[0] : co[3];
    hh[t] = sky_compress(x, y);
  }
  __syncthreads();
  if (t == 0) {
    const fr_t alpha = sky_compress(hh[0], hh[1]);
    st->alpha[idx] = alpha;
    st->state = alpha;
    const fr_t cf[4] = {co[0], co[1], co[2], co[3]};
    const fr_t S = cubic(cf, alpha);
    st->S = S;
    ctx.state = alpha;
    ctx.S = S;
  }
}

// Sum of np <= kMaxRoundBlocks partials of NV values (stride NV) by a
// workgroup of kFinThreads: every thread issues its loads (at most four
// partials) together, then wave sums by shuffles and the per-wave sums on
// thread 0, where the result is read.  One wave looping over the partials
// had waited out up to sixteen dependent loads (~30 us at 1024 partials).
constexpr int kFinThreads = 256;
template <int NV>
__device__ __forceinline__ void block_sum_partials(const fr_t* partials, unsigned np,
                                                   fr_t (&out)[NV]) {
  constexpr int K = kMaxRoundBlocks / kFinThreads;
  __shared__ fr_t wsum[NV][kFinThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  fr_t x[K][NV];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const unsigned i = t + k * kFinThreads;
#pragma unroll
    for (int q = 0; q < NV; q++) x[k][q] = i < np ? fr_load(partials + NV * i + q) : fr_zero();
  }
  fr_t acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    acc[q] = x[0][q];
#pragma unroll
    for (int k = 1; k < K; k++) acc[q] = fr_add_ilv(acc[q], x[k][q]);
  }
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] = fr_add_ilv(acc[q], fr_shfl(acc[q], lane ^ w));
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) wsum[q][wave] = acc[q];
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      out[q] = wsum[q][0];
#pragma unroll
      for (int w = 1; w < kFinThreads / 64; w++) out[q] = fr_add_ilv(out[q], wsum[q][w]);
    }
  }
}

// Round idx after k_sc_round: sum the np workgroup partials, then
// finalize_round.  NV = 4: round 0 of the plain mode, whose partials also
// carry the high-half sums of the claim.
template <int NV>
__global__ __launch_bounds__(kFinThreads) void k_sc_finalize(ScState* st, const fr_t* partials,
                                                             unsigned np, int idx) {
  RoundCtx ctx;
  if (threadIdx.x == 0) ctx = load_ctx(st);
  fr_t sums[NV];
  block_sum_partials<NV>(partials, np, sums);
  const fr_t hi = sums[NV - 1];
  finalize_round(st, idx, ctx, sums[0], sums[1], sums[2], NV == 4 ? &hi : nullptr);
}

// The last rounds, from an input of cur <= kTailElems elements down to 2, in
// one workgroup: fold + map (at most one pair per thread), workgroup sums,
// finalize_round, then the next round, with no launch in between.  Below
// ~2^11 elements a round is latency-bound (a short chain of products per lane
// and the transcript step), so one workgroup is as fast as a grid and saves
// two launches per round.
constexpr int kTailThreads = 512;
constexpr uint64_t kTailElems = 4 * (uint64_t)kTailThreads;

__global__ __launch_bounds__(kTailThreads) void k_sc_tail(ScState* st, fr_t* a, fr_t* b, fr_t* c,
                                                          fr_t* eq, uint64_t cur, int idx0) {
  __shared__ fr_t alpha_sh;
  __shared__ fr_t wsum[3][kTailThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  fr_t* mle[4] = {a, b, c, eq};
  // in LDS, not registers: the tail's fold and the unrolled compressions
  // already fill the 256 VGPRs a 512-thread workgroup allows
  __shared__ RoundCtx ctx;
  if (t == 0) {
    ctx = load_ctx(st);
    alpha_sh = ctx.state;  // the state after round idx0 - 1 is its challenge
  }
  __syncthreads();
  const int m = st->m;
  for (int idx = idx0; idx < m; idx++, cur >>= 1) {
    const uint64_t h = cur / 4;
    const fr_t f = alpha_sh;
    fr_t v[3] = {fr_zero(), fr_zero(), fr_zero()};
    if ((uint64_t)t < h) {
      const uint64_t i = t;
      fr_t lo[4], hi[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        fr_t x0 = fr_load(mle[j] + i), x1 = fr_load(mle[j] + i + h);
        const fr_t x2 = fr_load(mle[j] + i + 2 * h), x3 = fr_load(mle[j] + i + 3 * h);
        x0 = fr_add(x0, fr_mul(f, fr_sub(x2, x0)));
        x1 = fr_add(x1, fr_mul(f, fr_sub(x3, x1)));
        fr_t w0 = x0, w1 = x1;
        fr_canon(w0);
        fr_canon(w1);
        fr_store(mle[j] + i, w0);
        fr_store(mle[j] + i + h, w1);
        lo[j] = x0;
        hi[j] = x1;
      }
      v[0] = r1cs_term(lo[3], lo[0], lo[1], lo[2]);
      const fr_t ea = fr_sub(fr_add(lo[3], lo[3]), hi[3]);
      const fr_t aa = fr_sub(fr_add(lo[0], lo[0]), hi[0]);
      const fr_t bb = fr_sub(fr_add(lo[1], lo[1]), hi[1]);
      const fr_t cc = fr_sub(fr_add(lo[2], lo[2]), hi[2]);
      v[1] = fr_mul(ea, fr_sub(fr_mul(aa, bb), cc));
      v[2] = fr_mul(fr_mul(fr_sub(hi[3], lo[3]), fr_sub(hi[0], lo[0])), fr_sub(hi[1], lo[1]));
    }
    // wave sums by shuffles, then the per-wave sums on thread 0
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
#pragma unroll
      for (int q = 0; q < 3; q++) v[q] = fr_add_ilv(v[q], fr_shfl(v[q], lane ^ w));
    }
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 3; q++) wsum[q][wave] = v[q];
    }
    __syncthreads();
    fr_t h3[3] = {fr_zero(), fr_zero(), fr_zero()};
    if (t == 0) {
      for (int w = 0; w < kTailThreads / 64; w++) {
#pragma unroll
        for (int q = 0; q < 3; q++) h3[q] = fr_add_ilv(h3[q], wsum[q][w]);
      }
#pragma unroll
      for (int q = 0; q < 3; q++) fr_canon(h3[q]);
    }
    finalize_round(st, idx, ctx, h3[0], h3[1], h3[2], nullptr);
    if (t == 0) alpha_sh = ctx.state;
    __syncthreads();
  }
}

}  // namespace

// ------------------------------------------------------------------ launchers
uint64_t spmv_chunks(uint64_t nnz) { return nnz ? (nnz + kChunk - 1) / kChunk : 1; }

template <class V>
static void spmv_launch_v(const CsrDev& A, V vals, const fr_t* z, fr_t* out, uint32_t* flags,
                          fr_t* partials, hipStream_t s) {
  const uint64_t n_chunks = spmv_chunks(A.nnz);
  // flags[n_chunks], spans[1 + n_chunks], head/tail partials[n_chunks]
  uint32_t* spans = flags + n_chunks;
  fr_t* head = partials;
  fr_t* tail = head + n_chunks;
  k_spmv_clear<<<nblk(n_chunks + 1, 256), 256, 0, s>>>(flags, n_chunks + 1);  // flags and the count
  k_spmv<V><<<nblk(A.rows, 256), 256, 0, s>>>(A.rows, A.ncols, A.row_starts, A.nnz, A.cols, vals,
                                               z, out, flags, spans, n_chunks);
  k_spmv_chunks<V><<<(unsigned)std::min<uint64_t>(n_chunks, 1024), 256, 0, s>>>(
      A.rows, A.ncols, A.row_starts, A.nnz, A.cols, vals, z, out, flags, head, tail, n_chunks);
  k_spmv_spans<<<64, 256, 0, s>>>(A.rows, A.row_starts, A.nnz, spans, n_chunks, head, tail, out);
  PK_HIP(hipGetLastError());
}

void spmv_launch(const CsrDev& A, const fr_t* z, fr_t* out, uint32_t* flags, fr_t* partials,
                 hipStream_t s) {
  if (!A.rows) return;
  if (A.ids)
    spmv_launch_v(A, ValsInterned{A.ids, A.interner, A.interner_len}, z, out, flags, partials, s);
  else
    spmv_launch_v(A, ValsHydrated{A.vals}, z, out, flags, partials, s);
}

void spmv_dev(const CsrDev& A, const fr_t* z, fr_t* out, hipStream_t s) {
  if (!A.rows) return;
  const uint64_t n_chunks = spmv_chunks(A.nnz);
  Scratch flags(4 * (2 * n_chunks + 1), s), part(2 * sizeof(fr_t) * n_chunks, s);
  spmv_launch(A, z, out, flags.as<uint32_t>(), part.as<fr_t>(), s);
}

void hadamard_dev(const fr_t* a, const fr_t* b, fr_t* c, uint64_t n, uint64_t n_pad, hipStream_t s) {
  if (!n_pad) return;
  k_hadamard<<<nblk(n_pad, 256), 256, 0, s>>>(a, b, c, n, n_pad);
  PK_HIP(hipGetLastError());
}

void zero_dev(fr_t* p, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_zero<<<nblk(n, 256), 256, 0, s>>>(p, n);
  PK_HIP(hipGetLastError());
}

void eq_table_dev(const fr_t* r, int m, uint64_t num_entries, fr_t* out, hipStream_t s) {
  if (!num_entries) return;
  fr_t one;
  memcpy(&one, &hf::ONE, 32);
  if (m <= 12) {
    k_eq_direct<<<nblk(num_entries, 256), 256, 0, s>>>(r, m, num_entries, out, one);
    PK_HIP(hipGetLastError());
    return;
  }
  // eq(r, x) = eq(r_hi, x_hi) eq(r_lo, x_lo); only the half-table entries
  // that x < num_entries reaches are built, so a large m with few entries
  // (the reference's truncation) stays small
  const int lo_bits = m / 2, hi_bits = m - lo_bits;
  const uint64_t lo_n = std::min<uint64_t>(num_entries, 1ull << lo_bits);
  const uint64_t hi_n = ((num_entries - 1) >> lo_bits) + 1;
  Scratch hi(sizeof(fr_t) * hi_n, s), lo(sizeof(fr_t) * lo_n, s);
  k_eq_direct<<<nblk(hi_n, 256), 256, 0, s>>>(r, hi_bits, hi_n, hi.as<fr_t>(), one);
  k_eq_direct<<<nblk(lo_n, 256), 256, 0, s>>>(r + hi_bits, lo_bits, lo_n, lo.as<fr_t>(), one);
  k_eq_combine<<<nblk(num_entries, 256), 256, 0, s>>>(hi.as<fr_t>(), lo.as<fr_t>(), lo_bits,
                                                      num_entries, out);
  PK_HIP(hipGetLastError());
}

// Workgroups of a grid-stride round kernel: one per 256 pairs, at most as
// many as the device holds at once (the kernel's occupancy times the CUs,
// at most kMaxRoundBlocks).  Every workgroup strides over the same share of
// the pairs, so a grid larger than one generation would run its remainder
// as a second, thinly occupied generation (k_sc_round holds three waves per
// SIMD: 768 resident workgroups on 256 CUs, against the 1024 it took before).
template <class K>
static unsigned round_blocks(uint64_t work, K kernel) {
  static std::mutex mtx;
  static std::map<std::pair<int, const void*>, unsigned> cap_of;
  int dev = 0;
  PK_HIP(hipGetDevice(&dev));
  unsigned cap;
  {
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cap_of.find({dev, reinterpret_cast<const void*>(kernel)});
    if (it == cap_of.end()) {
      int per_cu = 0, cus = 0;
      PK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kRoundThreads, 0));
      PK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const uint64_t c = (uint64_t)std::max(per_cu, 1) * (uint64_t)std::max(cus, 1);
      it = cap_of.emplace(std::make_pair(dev, reinterpret_cast<const void*>(kernel)),
                          (unsigned)std::min<uint64_t>(c, kMaxRoundBlocks)).first;
    }
    cap = it->second;
  }
  uint64_t b = (work + kRoundThreads - 1) / kRoundThreads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

void sc_round_dev(fr_t* a, fr_t* b, fr_t* c, fr_t* eq, uint64_t size, const fr_t* fold,
                  fr_t* partials, fr_t* out3, hipStream_t s) {
  const uint64_t h = fold ? size / 4 : size / 2;
  const unsigned nb = round_blocks(h, k_sc_round<false>);
  k_sc_round<false><<<nb, kRoundThreads, 0, s>>>(a, b, c, eq, size, fold, partials);
  k_sum_partials<<<1, kRoundThreads, 0, s>>>(partials, nb, out3);
  PK_HIP(hipGetLastError());
}

size_t sc_partials_elems() { return 4 * (size_t)kMaxRoundBlocks; }

void sc_prove_dev(ScState* st, fr_t* a, fr_t* b, fr_t* c, fr_t* eq, int m, int mode,
                  uint64_t seed, fr_t* work, hipStream_t s) {
  const uint64_t size = m >= 1 ? (1ull << m) : 2;
  fr_t* partials = work;
  fr_t* h3 = work + sc_partials_elems();
  k_sc_init<<<1, 64, 0, s>>>(st, seed);
  PK_HIP(hipGetLastError());
  eq_table_dev(st->r, m, 1ull << m, eq, s);
  if (m == 0) zero_dev(eq + 1, 1, s);
  // The plain mode's claim comes with round 0's partials (k_sc_round<true>);
  // with no rounds (m = 0) it takes its own pass.  The reference mode's
  // claim is rho * sum_g, from the blinding polynomial alone.
  if (mode == 0 && m == 0) {
    const unsigned nb = round_blocks(size, k_sc_claim);
    k_sc_claim<<<nb, kRoundThreads, 0, s>>>(a, b, c, eq, size, partials);
    k_sum_partials<<<1, kRoundThreads, 0, s>>>(partials, nb, h3);
  }
  if (mode != 0 || m == 0) k_sc_claim_finalize<<<1, 64, 0, s>>>(st, h3);
  PK_HIP(hipGetLastError());
  uint64_t cur = size;
  for (int idx = 0; idx < m; idx++) {
    const fr_t* fold = idx ? &st->alpha[idx - 1] : nullptr;
    if (fold && cur <= kTailElems) {
      // the remaining rounds in one workgroup
      k_sc_tail<<<1, kTailThreads, 0, s>>>(st, a, b, c, eq, cur, idx);
      PK_HIP(hipGetLastError());
      break;
    }
    // two launches per round: fold + map + workgroup sums, then the
    // sum of the partials and the transcript step in one workgroup
    const uint64_t pairs = fold ? cur / 4 : cur / 2;
    const unsigned nb = (idx == 0 && mode == 0) ? round_blocks(pairs, k_sc_round<true>)
                                                : round_blocks(pairs, k_sc_round<false>);
    if (idx == 0 && mode == 0) {
      k_sc_round<true><<<nb, kRoundThreads, 0, s>>>(a, b, c, eq, cur, fold, partials);
      k_sc_finalize<4><<<1, kFinThreads, 0, s>>>(st, partials, nb, idx);
    } else {
      k_sc_round<false><<<nb, kRoundThreads, 0, s>>>(a, b, c, eq, cur, fold, partials);
      k_sc_finalize<3><<<1, kFinThreads, 0, s>>>(st, partials, nb, idx);
    }
    if (fold) cur /= 2;
    PK_HIP(hipGetLastError());
  }
  if (mode != 0 && m >= 1) {
    k_sc_blinding_eval<<<1, 64, 0, s>>>(st);
    PK_HIP(hipGetLastError());
  }
}

void sc_state_init_host(ScState& h, int m, int mode, const fr_t* blinding) {
  memset(&h, 0, sizeof h);
  h.m = m;
  h.mode = mode;
  memcpy(&h.one, &hf::ONE, 32);
  hfr half = hf::inv(hf::from_u64(2));
  memcpy(&h.half, &half, 32);
  if (blinding) memcpy(h.g, blinding, sizeof(fr_t) * 4 * m);
}

}  // namespace pk
