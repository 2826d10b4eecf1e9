// One-polynomial WHIR-style commitment on the device: RS encode (NTT) ->
// leaf hashing -> Skyscraper Merkle tree.  Inputs and outputs stay resident.
// whir_r1cs_commit_dev adds the witness shaping of WhirR1CSScheme::commit
// (provekit/prover/src/whir_r1cs.rs:103-126) in front: zero padding to
// 2^num_vars and the blinding coefficients at offset = witness size.
#include "kernels.hpp"

namespace pk {

static void commit_check(uint64_t n, uint64_t blowup, uint64_t leaf_elems) {
  PK_REQUIRE(is_pow2(n) && is_pow2(blowup), PK_ERR_NOT_POW2,
             "commit: length and blowup must be powers of two");
  const uint64_t N = n * blowup;
  PK_REQUIRE(N <= (1ull << 28), PK_ERR_SIZE, "commit: codeword above 2^28");
  PK_REQUIRE(leaf_elems >= 1 && N % leaf_elems == 0 && is_pow2(leaf_elems), PK_ERR_SIZE,
             "commit: leaf size must be a power of two dividing the codeword");
}

void commit_dev(Device& d, const fr_t* coeffs, uint64_t n, uint64_t blowup, int coset,
                uint64_t leaf_elems, fr_t* codeword, fr_t* nodes, hipStream_t s) {
  commit_check(n, blowup, leaf_elems);
  const uint64_t N = n * blowup;
  rs_encode_dev(d, coeffs, 1, n, nullptr, 0, N, codeword, coset, s);
  const uint64_t leaves = N / leaf_elems;
  sky_hash_many_dev(leaf_elems, codeword, leaves, nodes, s);
  sky_merkle_dev(nodes, leaves, nodes, s);
}

namespace {

// polynomial[i] = witness[i] below witness_size, then the blinding
// coefficients g (row-major m0 x 4), then zeros; in place when
// witness == polynomial (each thread reads its own element first).
__global__ __launch_bounds__(256) void k_shape_witness(const fr_t* w, uint64_t ws, const fr_t* g,
                                                       uint64_t ng, fr_t* out, uint64_t len) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  fr_t v;
  if (i < ws) {
    v = fr_load(w + i);
  } else if (i - ws < ng) {
    v = fr_load(g + (i - ws));
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) v.l[k] = 0;
  }
  fr_store(out + i, v);
}

}  // namespace

void whir_r1cs_shape_dev(const fr_t* witness, uint64_t witness_size, int num_vars, const fr_t* g,
                         uint64_t m0, fr_t* polynomial, hipStream_t s) {
  const uint64_t len = 1ull << num_vars;
  PK_REQUIRE(witness_size <= len, PK_ERR_SIZE, "R1CS witness length exceeds scheme capacity");
  const uint64_t ng = g ? 4 * m0 : 0;
  PK_REQUIRE(witness_size + ng <= len, PK_ERR_SIZE,
             "index out of bounds: the blinding coefficients run past the padded witness");
  k_shape_witness<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(witness, witness_size, g, ng,
                                                                polynomial, len);
  PK_HIP(hipGetLastError());
}

void whir_r1cs_commit_dev(Device& d, const fr_t* witness, uint64_t witness_size, int num_vars,
                          const fr_t* g, uint64_t m0, uint64_t blowup, int coset,
                          uint64_t leaf_elems, fr_t* polynomial, fr_t* codeword, fr_t* nodes,
                          hipStream_t s) {
  PK_REQUIRE(num_vars >= 0 && num_vars <= 28, PK_ERR_SIZE, "commit: more than 2^28 coefficients");
  commit_check(1ull << num_vars, blowup, leaf_elems);  // before the polynomial is written
  whir_r1cs_shape_dev(witness, witness_size, num_vars, g, m0, polynomial, s);
  commit_dev(d, polynomial, 1ull << num_vars, blowup, coset, leaf_elems, codeword, nodes, s);
}

}  // namespace pk
