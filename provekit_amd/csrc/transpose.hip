// CSR transpose on gfx950 (SparseMatrix::transpose,
// provekit/common/src/sparse_matrix.rs:409-449) and the product with
// eq(alpha, .) that follows the sumcheck (multiply_transposed_by_eq_alpha,
// provekit/common/src/utils/sumcheck.rs:203-227).
//
// The transposed entries are the (col, row, value) triples in (col, row)
// order.  The CSR arrays list entries row by row, so a STABLE sort of the
// entries by column gives that order.  The sort is a least-significant-digit
// radix sort written for this path: the column bits (22 at 2^22 columns) in
// passes of at most 8 bits (256 bins; 22 bits: 8 + 8 + 6), each pass moving
// the column (key) and an 8-byte payload (row, third word), where the third
// word is the entry's index (hydrated values) or its interned value id.  The
// first pass takes the keys and builds the payloads from the CSR arrays; the
// last writes the sorted columns, the rows (t_cols) and the third words (the
// ids, or the permutation that k_gather_vals follows to fetch the 32-byte
// values in output order, one random read per entry and sequential writes,
// while it fills the row starts).
//
// One pass over tiles of 4096 entries (1024 threads, 256 entries a wave):
//   k_radix_hist     each tile's digit counts, tile-major: H[tile][bin];
//   k_radix_scan     one workgroup per bin: H[.][bin] becomes the exclusive
//                    prefix over tiles, T[bin] the bin's total;
//   k_radix_scatter  each wave ranks its entries 64 at a time: the lanes
//                    holding the same digit are found with one ballot per
//                    digit bit, a lane's rank is the number of such lanes
//                    below it plus the wave's running count of that digit (an
//                    LDS counter per wave and digit that the lowest such lane
//                    advances with an atomic add and hands to the others);
//                    the waves' counts are prefix-summed per digit, the tile's
//                    keys are staged in LDS in digit order, and slot i of that
//                    order goes to (sum of T below its digit) + H[tile][digit]
//                    + (i - the digit's first slot in the tile), so that the
//                    lanes of a wave write runs of consecutive records.
//                    Entries with equal digits keep their order, so every pass
//                    is stable and the passes together sort by the whole
//                    column.
// Around the sort:
//   k_row_of         the row of every entry (for the first pass's records):
//                    each wave takes 64 rows and walks their entries 64 at a
//                    time, finding each entry's row by a binary search over
//                    the lanes' row starts (shuffles);
//   k_row_starts     t_row_starts from the sorted columns: slot k opens every
//                    column in (col[k-1], col[k]]; the last slot also closes
//                    the columns after it at nnz.
// Every kernel keeps its accesses in bounds for any input.  The device path
// trusts the CSR invariants (non-decreasing row starts within nnz, columns
// below num_cols); a column beyond num_cols sorts by its low bits.
#include <initializer_list>

#include "kernels.hpp"
#include "sumcheck.hpp"

namespace pk {

namespace {

constexpr int kMaxDigitBits = 8;
constexpr uint32_t kBins = 1u << kMaxDigitBits;
constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kChunks = 4;                        // 64-entry chunks per wave
constexpr int kPerWave = 64 * kChunks;            // 256
constexpr uint64_t kTile = (uint64_t)kWaves * kPerWave;  // 4096
constexpr int kScanThreads = 256;

inline unsigned nblocks(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Exclusive prefix sum over the workgroup (NT threads) of one value per
// thread; *total is the workgroup's sum.  Every thread must call it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t ws[NT / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; i++) {
    off += i < w ? ws[i] : 0u;
    tot += ws[i];
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// ------------------------------------------------------------ row of entry
__global__ __launch_bounds__(256) void k_row_of(const uint32_t* row_starts, uint64_t rows,
                                                uint64_t nnz, uint32_t* row_of) {
  const uint64_t r0 = (uint64_t)blockIdx.x * 256 + (threadIdx.x & ~63u);
  if (r0 >= rows) return;  // the whole wave leaves together
  const int lane = threadIdx.x & 63;
  const int last = (int)(rows - r0 < 64 ? rows - r0 - 1 : 63);  // last lane with a row
  const uint64_t r = r0 + (uint64_t)(lane < last ? lane : last);
  const uint64_t e = r + 1 < rows ? (row_starts[r + 1] < nnz ? row_starts[r + 1] : nnz) : nnz;
  const uint64_t b = row_starts[r] < e ? row_starts[r] : e;
  const uint32_t b32 = (uint32_t)b;
  // the first wave also takes entries before row 0's start (row 0), so every
  // entry gets a row even when row_starts[0] > 0
  const uint64_t lo_e = r0 == 0 ? 0 : (uint32_t)__shfl((int)b32, 0, 64);
  const uint64_t hi_e = (uint32_t)__shfl((int)(uint32_t)e, last, 64);
  for (uint64_t p0 = lo_e; p0 < hi_e; p0 += 64) {
    const uint64_t p = p0 + lane;
    // largest lane l <= last with b_l <= p (row starts are non-decreasing)
    int l = 0;
#pragma unroll
    for (int step = 32; step; step >>= 1) {
      const int c = l + step;
      const uint32_t bc = (uint32_t)__shfl((int)b32, c & 63, 64);
      if (c <= last && bc <= p) l = c;
    }
    if (p < hi_e) row_of[p] = (uint32_t)(r0 + l);
  }
}

// ------------------------------------------------------------ pass ends
// Sources: where a pass reads its entries.
struct SrcCsr {  // the CSR arrays (first pass)
  const uint32_t* cols;
  const uint32_t* row_of;
  const uint32_t* third;  // interned ids, or null: the entry index
  __device__ __forceinline__ uint32_t key(uint64_t i) const { return cols[i]; }
  __device__ __forceinline__ uint2 pay(uint64_t i) const {
    return make_uint2(row_of[i], third ? third[i] : (uint32_t)i);
  }
};
struct SrcArr {  // the previous pass's keys and payloads
  const uint32_t* keys;
  const uint2* pays;
  __device__ __forceinline__ uint32_t key(uint64_t i) const { return keys[i]; }
  __device__ __forceinline__ uint2 pay(uint64_t i) const { return pays[i]; }
};
// Sinks: where a pass writes slot pos.
struct SinkArr {
  uint32_t* keys;
  uint2* pays;
  __device__ __forceinline__ void put(uint64_t pos, uint32_t k, const uint2& v) const {
    keys[pos] = k;
    pays[pos] = v;
  }
};
struct SinkOut {  // last pass: sorted columns, rows, third words
  uint32_t* keys;
  uint32_t* rows;
  uint32_t* third;
  __device__ __forceinline__ void put(uint64_t pos, uint32_t k, const uint2& v) const {
    keys[pos] = k;
    rows[pos] = v.x;
    third[pos] = v.y;
  }
};

// ------------------------------------------------------------ radix passes
template <class Src>
__global__ __launch_bounds__(kThreads) void k_radix_hist(Src src, uint64_t n, int shift, uint32_t bins,
                                                         uint32_t* H) {
  __shared__ uint32_t h[kBins];
  for (uint32_t b = threadIdx.x; b < bins; b += kThreads) h[b] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  const uint64_t end = base + kTile < n ? base + kTile : n;
  for (uint64_t i = base + threadIdx.x; i < end; i += kThreads)
    atomicAdd(&h[(src.key(i) >> shift) & (bins - 1)], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < bins; b += kThreads) H[(uint64_t)blockIdx.x * bins + b] = h[b];
}

// one workgroup per bin: exclusive prefix of H[.][bin] over the tiles, the
// tiles taken kScanPer a thread (loads issued together)
constexpr int kScanPer = 16;
__global__ __launch_bounds__(kScanThreads) void k_radix_scan(uint32_t* H, uint64_t tiles, uint32_t bins,
                                                             uint32_t* T) {
  const uint32_t b = blockIdx.x;
  uint32_t carry = 0;
  for (uint64_t c0 = 0; c0 < tiles; c0 += (uint64_t)kScanThreads * kScanPer) {
    const uint64_t t0 = c0 + (uint64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
      v[k] = t0 + k < tiles ? H[(t0 + k) * bins + b] : 0u;
      sum += v[k];
    }
    uint32_t total;
    uint32_t x = carry + block_excl_scan<kScanThreads>(sum, &total);
#pragma unroll
    for (int k = 0; k < kScanPer; k++)
      if (t0 + k < tiles) {
        H[(t0 + k) * bins + b] = x;
        x += v[k];
      }
    carry += total;
  }
  if (threadIdx.x == 0) T[b] = carry;
}

template <class Src, class Sink>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(Src src, Sink sink, uint64_t n, int shift,
                                                            int digit_bits, const uint32_t* H,
                                                            const uint32_t* T) {
  __shared__ uint32_t cnt[kWaves][kBins];  // per wave and digit: running count, then first slot
  __shared__ uint32_t first[kBins];        // output slot of the tile's digit-order slot 0, per digit
  __shared__ uint32_t stage_key[kTile];
  __shared__ uint16_t stage_idx[kTile];
  const uint32_t bins = 1u << digit_bits;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const uint64_t tile = blockIdx.x;
  for (uint32_t i = t; i < kWaves * kBins; i += kThreads) (&cnt[0][0])[i] = 0;

  const uint64_t base = tile * kTile + (uint64_t)w * kPerWave;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes under this one
  uint32_t key[kChunks], rank[kChunks];
#pragma unroll
  for (int j = 0; j < kChunks; j++) {
    const uint64_t idx = base + (uint64_t)j * 64 + lane;
    key[j] = idx < n ? src.key(idx) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kChunks; j++) {
    const uint64_t idx = base + (uint64_t)j * 64 + lane;
    const bool valid = idx < n;
    const uint32_t d = (key[j] >> shift) & (bins - 1);
    // eight ballots whatever the digit width: bits at and above digit_bits
    // are zero in every lane, so their rounds leave peers unchanged, and the
    // loop unrolls
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kMaxDigitBits; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t r = (uint32_t)__popcll(peers & below);
    uint32_t old = 0;
    if (valid && r == 0) old = atomicAdd(&cnt[w][d], (uint32_t)__popcll(peers));
    const int leader = peers ? __ffsll((unsigned long long)peers) - 1 : 0;
    rank[j] = (uint32_t)__shfl((int)old, leader, 64) + r;
  }
  __syncthreads();
  {
    // thread b < bins: the tile's count of digit b, its in-tile first slot
    // (x), the bin's first output slot, and the waves' first slots
    const bool own = (uint32_t)t < bins;
    uint32_t sum = 0;
#pragma unroll
    for (int v = 0; v < kWaves; v++) sum += own ? cnt[v][t & (kBins - 1)] : 0u;
    uint32_t total;
    const uint32_t x = block_excl_scan<kThreads>(sum, &total);
    const uint32_t below_bins = block_excl_scan<kThreads>(own ? T[t & (kBins - 1)] : 0u, &total);
    if (own) {
      first[t] = below_bins + H[tile * bins + t] - x;
      uint32_t run = x;
#pragma unroll
      for (int v = 0; v < kWaves; v++) {
        const uint32_t c = cnt[v][t];
        cnt[v][t] = run;
        run += c;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kChunks; j++) {
    const uint64_t idx = base + (uint64_t)j * 64 + lane;
    if (idx >= n) continue;
    const uint32_t slot = cnt[w][(key[j] >> shift) & (bins - 1)] + rank[j];
    if (slot < kTile) {
      stage_key[slot] = key[j];
      stage_idx[slot] = (uint16_t)(idx - tile * kTile);
    }
  }
  __syncthreads();
  const uint64_t tn = n - tile * kTile < kTile ? n - tile * kTile : kTile;
  for (uint32_t i = t; i < tn; i += kThreads) {
    const uint32_t k = stage_key[i];
    const uint64_t pos = (uint64_t)first[(k >> shift) & (bins - 1)] + i;
    const uint64_t from = tile * kTile + stage_idx[i];
    if (pos < n && from < n) sink.put(pos, k, src.pay(from));
  }
}

__global__ __launch_bounds__(256) void k_row_starts(const uint32_t* sorted_cols, uint64_t nnz,
                                                    uint64_t ncols, uint32_t* t_row_starts) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nnz) return;
  const uint64_t lo = k == 0 ? 0 : (uint64_t)sorted_cols[k - 1] + 1;
  const uint64_t hi = k == nnz ? ncols : (uint64_t)sorted_cols[k] + 1;
  for (uint64_t c = lo; c < hi && c < ncols; c++) t_row_starts[c] = (uint32_t)k;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// hydrated last step, thread k <= nnz: t_vals[k] = vals[perm[k]] and the
// row starts that slot k opens (as k_row_starts)
__global__ __launch_bounds__(256) void k_gather_vals(const uint32_t* perm, const uint32_t* sorted_cols,
                                                     const fr_t* vals, uint64_t n, uint64_t ncols,
                                                     fr_t* t_vals, uint32_t* t_row_starts) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  if (k < n) {
    const uint32_t e = perm[k];
    fr_store(t_vals + k, fr_load(vals + (e < n ? e : 0)));
  }
  const uint64_t lo = k == 0 ? 0 : (uint64_t)sorted_cols[k - 1] + 1;
  const uint64_t hi = k == n ? ncols : (uint64_t)sorted_cols[k] + 1;
  for (uint64_t c = lo; c < hi && c < ncols; c++) t_row_starts[c] = (uint32_t)k;
}

struct Plan {
  uint64_t n, tiles;
  int bits, passes, digit;
  int shift(int p) const { return p * digit; }
  int width(int p) const { return p + 1 < passes ? digit : bits - p * digit; }
};

template <class Src, class Sink>
void radix_pass(const Plan& pl, int p, Src src, Sink sink, uint32_t* H, uint32_t* T, hipStream_t s) {
  const int shift = pl.shift(p), db = pl.width(p);
  const uint32_t bins = 1u << db;
  k_radix_hist<<<(unsigned)pl.tiles, kThreads, 0, s>>>(src, pl.n, shift, bins, H);
  k_radix_scan<<<bins, kScanThreads, 0, s>>>(H, pl.tiles, bins, T);
  k_radix_scatter<<<(unsigned)pl.tiles, kThreads, 0, s>>>(src, sink, pl.n, shift, db, H, T);
  PK_HIP(hipGetLastError());
}

// Sub-buffers carved from one scratch allocation (256-byte aligned).
struct Carve {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(uint64_t count) {
    off = (off + 255) & ~size_t(255);
    T* r = reinterpret_cast<T*>(base + off);
    off += sizeof(T) * (count ? count : 1);
    return r;
  }
  static size_t size(std::initializer_list<uint64_t> bytes) {
    size_t t = 0;
    for (uint64_t b : bytes) t += ((b ? b : 1) + 255) & ~size_t(255);
    return t;
  }
};

Plan make_plan(const CsrDev& A) {
  Plan pl;
  pl.n = A.nnz;
  pl.tiles = (pl.n + kTile - 1) / kTile;
  pl.bits = 1;
  while (pl.bits < 32 && (1ull << pl.bits) < A.ncols) pl.bits++;
  pl.passes = (pl.bits + kMaxDigitBits - 1) / kMaxDigitBits;
  pl.digit = (pl.bits + pl.passes - 1) / pl.passes;
  return pl;
}

// scratch bytes of sort_by_column: row_of, H, T, and the ping-pong keys and
// payloads
size_t sort_bytes(const Plan& pl) {
  const uint64_t na = pl.passes > 1 ? pl.n : 1, nb = pl.passes > 2 ? pl.n : 1;
  return Carve::size({4 * pl.n, 4 * pl.tiles * kBins, 4 * kBins, 4 * na, 8 * na, 4 * nb, 8 * nb});
}

// Stable sort of the entries by column; the last pass writes through `last`
// (whose keys array receives the sorted columns).  ws: sort_bytes(pl) bytes.
void sort_by_column(const CsrDev& A, const Plan& pl, const uint32_t* third, SinkOut last, Carve ws,
                    hipStream_t s) {
  const uint64_t na = pl.passes > 1 ? pl.n : 1, nb = pl.passes > 2 ? pl.n : 1;
  uint32_t* row_of = ws.take<uint32_t>(pl.n);
  uint32_t* h = ws.take<uint32_t>(pl.tiles * kBins);
  uint32_t* tt = ws.take<uint32_t>(kBins);
  SinkArr cur{ws.take<uint32_t>(na), ws.take<uint2>(na)};
  SinkArr other{ws.take<uint32_t>(nb), ws.take<uint2>(nb)};
  if (A.rows)
    k_row_of<<<nblocks(A.rows, 256), 256, 0, s>>>(A.row_starts, A.rows, pl.n, row_of);
  PK_HIP(hipGetLastError());
  const SrcCsr csr{A.cols, row_of, third};
  if (pl.passes == 1) {
    radix_pass(pl, 0, csr, last, h, tt, s);
    return;
  }
  radix_pass(pl, 0, csr, cur, h, tt, s);
  for (int p = 1; p + 1 < pl.passes; p++) {
    radix_pass(pl, p, SrcArr{cur.keys, cur.pays}, other, h, tt, s);
    std::swap(cur, other);
  }
  radix_pass(pl, pl.passes - 1, SrcArr{cur.keys, cur.pays}, last, h, tt, s);
}

void fill_row_starts(const uint32_t* sorted_keys, uint64_t nnz, uint64_t ncols,
                     uint32_t* t_row_starts, hipStream_t s) {
  k_row_starts<<<nblocks(nnz + 1, 256), 256, 0, s>>>(sorted_keys, nnz, ncols, t_row_starts);
  PK_HIP(hipGetLastError());
}

bool transpose_trivial(const CsrDev& A, uint32_t* t_row_starts, hipStream_t s) {
  PK_REQUIRE(A.nnz < (1ull << 32) && A.rows < (1ull << 32) && A.ncols < (1ull << 32), PK_ERR_SIZE,
             "transpose: sizes must fit u32 indices");
  PK_REQUIRE(A.rows > 0 || A.nnz == 0, PK_ERR_ARG, "transpose: entries without rows");
  if (!A.ncols) return true;
  if (!A.nnz) {
    k_fill_u32<<<nblocks(A.ncols, 256), 256, 0, s>>>(t_row_starts, A.ncols, 0u);
    PK_HIP(hipGetLastError());
    return true;
  }
  return false;
}

}  // namespace

void sparse_transpose_dev(const CsrDev& A, uint32_t* t_row_starts, uint32_t* t_cols,
                          fr_t* t_vals, hipStream_t s) {
  if (transpose_trivial(A, t_row_starts, s)) return;
  const Plan pl = make_plan(A);
  const size_t sb = sort_bytes(pl);
  Scratch ws(sb + Carve::size({4 * A.nnz, 4 * A.nnz}), s);
  Carve c{ws.as<char>(), sb};
  uint32_t* keys = c.take<uint32_t>(A.nnz);
  uint32_t* perm = c.take<uint32_t>(A.nnz);
  sort_by_column(A, pl, nullptr, SinkOut{keys, t_cols, perm}, Carve{ws.as<char>()}, s);
  k_gather_vals<<<nblocks(A.nnz + 1, 256), 256, 0, s>>>(perm, keys, A.vals, A.nnz, A.ncols, t_vals,
                                                          t_row_starts);
  PK_HIP(hipGetLastError());
}

void sparse_transpose_interned_dev(const CsrDev& A, uint32_t* t_row_starts, uint32_t* t_cols,
                                   uint32_t* t_ids, hipStream_t s) {
  if (transpose_trivial(A, t_row_starts, s)) return;
  const Plan pl = make_plan(A);
  const size_t sb = sort_bytes(pl);
  Scratch ws(sb + Carve::size({4 * A.nnz}), s);
  Carve c{ws.as<char>(), sb};
  uint32_t* keys = c.take<uint32_t>(A.nnz);
  sort_by_column(A, pl, A.ids, SinkOut{keys, t_cols, t_ids}, Carve{ws.as<char>()}, s);
  fill_row_starts(keys, A.nnz, A.ncols, t_row_starts, s);
}

void transposed_eq_mul_dev(const CsrDev& At, const fr_t* alpha, int m, fr_t* out,
                           hipStream_t s) {
  if (!At.rows) return;
  // eq(alpha, .) truncated to the constraint count, which is At's column count
  Scratch eq(sizeof(fr_t) * (At.ncols ? At.ncols : 1), s);
  eq_table_dev(alpha, m, At.ncols, eq.as<fr_t>(), s);
  spmv_dev(At, eq.as<fr_t>(), out, s);
}

}  // namespace pk
