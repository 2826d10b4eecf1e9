// One-polynomial WHIR-style commitment on the device: RS encode (NTT) ->
// leaf hashing -> Skyscraper Merkle tree.  Inputs and outputs stay resident.
#include "kernels.hpp"

namespace pk {

void commit_dev(Device& d, const fr_t* coeffs, uint64_t n, uint64_t blowup, int coset,
                uint64_t leaf_elems, fr_t* codeword, fr_t* nodes, hipStream_t s) {
  PK_REQUIRE(is_pow2(n) && is_pow2(blowup), PK_ERR_NOT_POW2,
             "commit: length and blowup must be powers of two");
  const uint64_t N = n * blowup;
  PK_REQUIRE(N <= (1ull << 28), PK_ERR_SIZE, "commit: codeword above 2^28");
  PK_REQUIRE(leaf_elems >= 1 && N % leaf_elems == 0 && is_pow2(leaf_elems), PK_ERR_SIZE,
             "commit: leaf size must be a power of two dividing the codeword");
  rs_encode_dev(d, coeffs, 1, n, nullptr, 0, N, codeword, coset, s);
  const uint64_t leaves = N / leaf_elems;
  sky_hash_many_dev(leaf_elems, codeword, leaves, nodes, s);
  sky_merkle_dev(nodes, leaves, nodes, s);
}

}  // namespace pk
