// CSR transpose on gfx950 (SparseMatrix::transpose,
// provekit/common/src/sparse_matrix.rs:409-449) and the product with
// eq(alpha, .) that follows the sumcheck (multiply_transposed_by_eq_alpha,
// provekit/common/src/utils/sumcheck.rs:203-227).
//
// The transposed entries are the (col, row, value) triples in (col, row)
// order.  The CSR arrays list entries row by row, so a stable sort by column
// of the pairs (row << 32 | entry index) gives that order and hands each
// output slot its row (the new column index) and the entry whose value it
// takes.  The sort is rocPRIM's stable LSD radix sort over the column bits
// only (22 bits at 2^22 columns: three passes).  Around it:
//   k_pack / k_mark_rows + scan  the (row << 32 | e) pairs: every entry
//                 starts as e, each non-empty row r stamps r into the high
//                 word at its first entry, and an inclusive scan with
//                 (a, b) -> (max(a.hi, b.hi), b.lo) carries the row forward
//                 (associative; one streaming pass instead of a search per
//                 entry);
//   k_gather      t_cols[k] = row, t_vals[k] = vals[e] (one 32-byte gather
//                 per entry);
//   k_row_starts  t_row_starts from the sorted columns: slot k opens every
//                 column in (col[k-1], col[k]]; the last slot also closes the
//                 columns after it at nnz.
// Every kernel keeps its accesses in bounds for any input (row starts
// beyond nnz are ignored, entry indices come from [0, nnz)).  Column indices
// must be below num_cols (the reference's invariant); larger ones sort by
// their low bits and land in some row of the result.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "kernels.hpp"
#include "sumcheck.hpp"

namespace pk {

namespace {

inline unsigned nblocks(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__global__ __launch_bounds__(256) void k_pack(uint64_t* v, uint64_t nnz) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) v[e] = e;
}

// non-empty row r stamps r into the high word of its first entry
__global__ __launch_bounds__(256) void k_mark_rows(const uint32_t* row_starts, uint64_t rows,
                                                   uint64_t nnz, uint64_t* v) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const uint64_t b = row_starts[r];
  const uint64_t e = r + 1 < rows ? row_starts[r + 1] : nnz;
  if (b < e && b < nnz) v[b] = (r << 32) | b;
}

struct CarryRow {
  __device__ __host__ uint64_t operator()(uint64_t a, uint64_t b) const {
    const uint64_t ha = a >> 32, hb = b >> 32;
    return ((ha > hb ? ha : hb) << 32) | (b & 0xffffffffull);
  }
};

__global__ __launch_bounds__(256) void k_gather(const uint64_t* sorted, const fr_t* vals,
                                                uint64_t nnz, uint32_t* t_cols, fr_t* t_vals) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint64_t v = sorted[k];
  uint64_t e = v & 0xffffffffull;
  e = e < nnz ? e : 0;
  t_cols[k] = (uint32_t)(v >> 32);
  fr_store(t_vals + k, fr_load(vals + e));
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

}  // namespace

void sparse_transpose_dev(const CsrDev& A, uint32_t* t_row_starts, uint32_t* t_cols,
                          fr_t* t_vals, hipStream_t s) {
  PK_REQUIRE(A.nnz < (1ull << 32) && A.rows < (1ull << 32) && A.ncols < (1ull << 32), PK_ERR_SIZE,
             "transpose: sizes must fit u32 indices");
  if (!A.ncols) return;
  if (!A.nnz) {
    k_fill_u32<<<nblocks(A.ncols, 256), 256, 0, s>>>(t_row_starts, A.ncols, 0u);
    PK_HIP(hipGetLastError());
    return;
  }
  const uint64_t nnz = A.nnz;
  unsigned end_bit = 1;
  while (end_bit < 32 && (1ull << end_bit) < A.ncols) end_bit++;
  Scratch packed(8 * nnz, s), with_rows(8 * nnz, s), keys(4 * nnz, s), sorted(8 * nnz, s);
  k_pack<<<nblocks(nnz, 256), 256, 0, s>>>(packed.as<uint64_t>(), nnz);
  if (A.rows)
    k_mark_rows<<<nblocks(A.rows, 256), 256, 0, s>>>(A.row_starts, A.rows, nnz,
                                                     packed.as<uint64_t>());
  PK_HIP(hipGetLastError());
  size_t scan_bytes = 0, sort_bytes = 0;
  PK_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, packed.as<uint64_t>(),
                                 with_rows.as<uint64_t>(), (size_t)nnz, CarryRow(), s));
  PK_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, A.cols, keys.as<uint32_t>(),
                                   with_rows.as<uint64_t>(), sorted.as<uint64_t>(), (size_t)nnz,
                                   0u, end_bit, s));
  Scratch temp(std::max<size_t>(std::max(scan_bytes, sort_bytes), 16), s);
  PK_HIP(rocprim::inclusive_scan(temp.p, scan_bytes, packed.as<uint64_t>(),
                                 with_rows.as<uint64_t>(), (size_t)nnz, CarryRow(), s));
  PK_HIP(rocprim::radix_sort_pairs(temp.p, sort_bytes, A.cols, keys.as<uint32_t>(),
                                   with_rows.as<uint64_t>(), sorted.as<uint64_t>(), (size_t)nnz,
                                   0u, end_bit, s));
  k_gather<<<nblocks(nnz, 256), 256, 0, s>>>(sorted.as<uint64_t>(), A.vals, nnz, t_cols, t_vals);
  PK_HIP(hipGetLastError());
  k_row_starts<<<nblocks(nnz + 1, 256), 256, 0, s>>>(keys.as<uint32_t>(), nnz, A.ncols,
                                                     t_row_starts);
  PK_HIP(hipGetLastError());
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
