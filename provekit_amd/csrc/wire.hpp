// R1CS wire format (postcard) to device CSR; see wire.hip.
#pragma once
#include "runtime.hpp"

namespace pk {

struct WireCsrSizes {
  uint64_t num_rows = 0, num_cols = 0, nnz = 0;
  uint64_t first = 0;  // index of the matrix's first varint in its stream
};
struct WireCsrOut {
  uint32_t* row_starts;  // num_rows
  uint32_t* cols;        // nnz
  fr_t* vals;            // nnz
};
struct R1csSizes {
  uint64_t num_public_inputs = 0, interner_len = 0;
  WireCsrSizes mats[3];
};

// Interner blob -> Montgomery values in `out` (device; nullptr: count only).
uint64_t interner_decode_dev(const uint8_t* blob, size_t len, fr_t* out, hipStream_t s);
// SparseMatrix blob -> CSR in `out` (nullptr: sizes only); interner: ni
// device values.
WireCsrSizes sparse_matrix_decode_dev(const uint8_t* blob, size_t len, const fr_t* interner,
                                      uint64_t ni, const WireCsrOut* out, hipStream_t s);
// R1CS blob -> interner values and the CSR of a, b, c (interner_out nullptr:
// sizes only; otherwise outs[3] must hold buffers of those sizes).
R1csSizes r1cs_decode_dev(const uint8_t* blob, size_t len, fr_t* interner_out,
                          const WireCsrOut* outs, hipStream_t s);

}  // namespace pk
