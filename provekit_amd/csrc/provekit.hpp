// C++ host mirror of the reference's interfaces for the WHIR commit path,
// backed by the gfx950 kernels.  Names, argument meaning and failure
// behaviour follow the Rust crates (a precondition violation that panics in
// the reference throws pk::Error here, with the reference's message):
//   ntt::{ntt_nr, intt_rn}                      ntt/src/ntt.rs:105,326
//   RSFr : ReedSolomon<Fr>                      provekit/common/src/ntt.rs:12-105
//   skyscraper::simple::{compress, compress_many} skyscraper/core/src/simple.rs:3-14
//   SkyscraperHashEngine : HashEngine           provekit/common/src/skyscraper/whir.rs:29-101
//   skyscraper::pow, SkyscraperPoW              skyscraper/core/src/pow.rs:18-45,
//                                               provekit/common/src/skyscraper/pow.rs:8-36
// Host slices are copied to HBM, transformed there, and copied back.
#pragma once
#include <array>
#include <optional>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace provekit_gpu {

using Fr = std::array<uint64_t, 4>;  // Montgomery limbs, as ark_bn254::Fr
using Hash = std::array<uint8_t, 32>;

void ntt_nr(Fr* values, size_t len, size_t codeword_size, size_t num_groups);
void intt_rn(Fr* values, size_t len);

struct RSFr {
  std::optional<size_t> next_order(size_t size) const;
  std::vector<Fr> evaluation_points(size_t masked_message_length, size_t codeword_length,
                                    const std::vector<size_t>& indices) const;
  std::vector<Fr> interleaved_encode(const std::vector<const Fr*>& messages,
                                     size_t message_length, const std::vector<Fr>& masks,
                                     size_t codeword_length) const;
  Fr generator(size_t codeword_length) const;
};

namespace skyscraper {
Fr compress(const Fr& l, const Fr& r);
void compress_many(const uint8_t* messages, size_t messages_len, uint8_t* hashes,
                   size_t hashes_len);
// skyscraper::pow (skyscraper/core/src/pow.rs:18-45).  The challenge is any
// 256-bit integer (4 LE u64).
namespace pow {
Fr threshold(double difficulty);
bool verify(const Fr& challenge, double difficulty, uint64_t nonce);
uint64_t solve(const Fr& challenge, double difficulty);
}  // namespace pow
}  // namespace skyscraper

// SkyscraperPoW : spongefish PowStrategy (provekit/common/src/skyscraper/pow.rs:8-36)
struct SkyscraperPoW {
  struct Solution {
    std::array<uint8_t, 32> challenge;
    uint64_t nonce;
  };
  SkyscraperPoW(const std::array<uint8_t, 32>& challenge, double bits);
  bool check(uint64_t nonce) const;
  Solution solution(uint64_t nonce) const { return Solution{challenge_, nonce}; }
  Solution solve() const;

 private:
  std::array<uint8_t, 32> challenge_;
  double bits_;
};

struct SkyscraperHashEngine {
  std::string name() const { return "skyscraper"; }
  bool supports_size(size_t size) const { return size > 0 && size % 32 == 0; }
  // The reference returns skyscraper::WIDTH_LCM = 12 (the SIMD width of its
  // CPU kernels).  One lane per message on 256 CUs wants tens of thousands of
  // messages per call, so the GPU engine asks for large batches.
  size_t preferred_batch_size() const { return 1u << 16; }
  void hash_many(size_t size, const uint8_t* input, size_t input_len, Hash* output,
                 size_t count) const;
};

// Commit one polynomial; see pkgpu_commit.
Hash commit(const Fr* coeffs, size_t n, size_t blowup, bool coset, size_t leaf_elems,
            Fr* codeword_out, Hash* nodes_out);

}  // namespace provekit_gpu
