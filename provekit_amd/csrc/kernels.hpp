// Launchers for the gfx950 kernels.  All pointers are device pointers; every
// launcher enqueues on `s` and returns without synchronising.
#pragma once
#include "runtime.hpp"

namespace pk {

// ---- NTT (ntt.hip) ----
// One NTT pass over a tile grid: stages [s0, s0 + r) of codeword index bits,
// tiles of 2^tl = 2^(r + jb) elements, bpt butterflies per thread per stage
// (see k_ntt_pass).
struct NttPassDesc {
  fr_t* data;
  const fr_t* src;  // first pass only: read element g from src[g % chunk]
  const fr_t* rr;   // root table
  uint64_t chunk;
  uint64_t Q;       // low stride: 2^q * M elements below the window
  int r, jb, canon, tl, bpt;
  int chunk_pow2;   // chunk is a power of two: g % chunk is a mask
};
// Schedule knobs of one transform: tiles of at most 2^tile_log elements,
// bpt butterflies per thread, earlier passes keep runs of >= 2^min_run
// consecutive elements.  tile_log 0 takes the tuned shape for the length.
struct NttShape {
  int tile_log = 0;
  int bpt = 0;       // 0: the tile's default
  int min_run = -1;  // -1: 2
};
// Tile run by workgroup b of a pass with T tiles per row h and H rows.  A
// first pass that reads a replicated coset chunk has one row per replica
// (h = g / chunk), so the H tiles (h, pos) read the same chunk elements.
// Workgroups go to the 8 XCDs round-robin (b mod 8), each with its own L2:
// b = 8 (H u + h) + x runs tile (h, 8 u + x), so the H replicas of a
// position run on one XCD within 8 H workgroups of each other and all but
// the first read the chunk from that L2.  Other passes, and rows that are
// not a multiple of 8 tiles, keep b.  A bijection on [0, H T).
__host__ __device__ inline uint64_t ntt_tile_order(bool replicated, uint64_t T, uint64_t H,
                                                   uint64_t b) {
  if (!replicated || (T & 7) != 0 || H == 0) return b;
  const uint64_t x = b & 7, rest = b >> 3;
  return (rest % H) * T + (rest / H) * 8 + x;
}
bool ntt_shape_supported(int tile_log, int bpt);
int ntt_default_bpt(int tile_log);
// The passes of ntt_nr over `data`.
std::vector<NttPassDesc> ntt_plan(Device& d, fr_t* data, uint64_t len, uint64_t codeword,
                                  uint64_t num_groups, const fr_t* src, uint64_t chunk,
                                  NttShape shape = NttShape());
void ntt_run_plan(const std::vector<NttPassDesc>& plan, uint64_t len, hipStream_t s);
// Run one batched pass round: workgroup b runs tile map[b].y of descs[map[b].x].
void ntt_batch_round_dev(const NttPassDesc* descs, const uint2* map, uint32_t blocks, int tl,
                         int bpt, hipStream_t s);
// Tuned shapes (tune.cpp): keyed by (log2 codeword, stages run).  The table
// is read once from the tuning file (PKGPU_NTT_TUNING, else ntt_tuning.json
// beside the library); lengths without an entry use the nearest tuned one
// with the same number of skipped stages, else tiles of 2^11.
NttShape ntt_tuned_shape(int log_codeword, int stages);
void ntt_tuning_put(int log_codeword, int stages, NttShape shape, double us);
void ntt_tuning_load(const std::string& path);
void ntt_tuning_save(const std::string& path);
// Times every supported shape for log_codeword in [log_min, log_max] and
// skipped stages 0..max_skip on `s`, keeps the fastest; returns seconds spent.
double ntt_autotune(Device& d, int log_min, int log_max, int max_skip, hipStream_t s);
// In-place DIT NTT, natural -> bit-reversed order, over `len` elements of
// `len / codeword` interleaved polynomials, starting at stage log2(num_groups)
// (ntt::ntt_nr).  If `src` is given, the first pass reads element g from
// src[g % chunk] instead of data[g] (replication of the first coset chunk).
void ntt_nr_dev(Device& d, fr_t* data, uint64_t len, uint64_t codeword, uint64_t num_groups,
                hipStream_t s, const fr_t* src = nullptr, uint64_t chunk = 0);
// ntt::intt_rn on device (uses one scratch buffer of len elements).
void intt_rn_dev(Device& d, fr_t* data, uint64_t len, hipStream_t s);
// ntt's intt_nr (natural-order evaluations -> coefficients in bit-reversed
// order) on device, one scratch buffer of len elements.
void intt_nr_dev(Device& d, fr_t* data, uint64_t len, hipStream_t s);
// Bit-reversal permutation of n elements, out-of-place.
void bitrev_permute_dev(const fr_t* in, fr_t* out, uint64_t n, hipStream_t s);
// ntt's reverse_order: the first `codeword` (a power of two) entries in
// bit-reversed order, in place.
void reverse_order_dev(fr_t* data, uint64_t codeword, hipStream_t s);
// RSFr::interleaved_encode into `out` (num_messages * codeword elements).
// msgs: num_messages rows of msg_len (row-major).  coset != 0 evaluates on the
// coset 5*H instead of H (coefficient of index j scaled by 5^j).
void rs_encode_dev(Device& d, const fr_t* msgs, uint64_t num_messages, uint64_t msg_len,
                   const fr_t* masks, uint64_t masks_len, uint64_t codeword, fr_t* out,
                   int coset, hipStream_t s);
// Evaluation points w^{bitrev(i)} for the given indices (RSFr::evaluation_points).
void rs_points_dev(Device& d, uint64_t codeword, const uint64_t* idx, uint64_t n, fr_t* out,
                   hipStream_t s);

// ---- Skyscraper (skyscraper.hip) ----
void sky_compress_many_dev(const fr_t* pairs, fr_t* out, uint64_t n, hipStream_t s);
// SkyscraperHashEngine::hash_many for messages of `chunks` 32-byte words.
void sky_hash_many_dev(uint64_t chunks, const fr_t* in, uint64_t count, fr_t* out, hipStream_t s);
// Binary tree over n (power of two) leaf digests: nodes[0..n) = leaves (may
// alias `leaves`), levels follow, root at nodes[2n-2].
void sky_merkle_dev(const fr_t* leaves, uint64_t n, fr_t* nodes, hipStream_t s);
// Openings from a resident tree (any 2-to-1 tree in this node layout): for
// query q with leaf index idx[q] < leaves (device array), values gets the
// leaf's leaf_elems codeword elements and paths the log2(leaves) sibling
// digests, leaf level first.
void merkle_open_dev(const fr_t* codeword, const fr_t* nodes, uint64_t leaves, uint64_t leaf_elems,
                     const uint64_t* idx, uint64_t n, fr_t* values, fr_t* paths, hipStream_t s);

// ---- proof of work (pow.hip) ----
// f64_to_u256 and threshold of skyscraper/core/src/pow.rs:18-26,61-88
// (threshold throws PK_ERR_ARG outside [0, 80)).
void f64_to_u256(double f, uint64_t out[4]);
void pow_threshold(double difficulty, uint64_t out[4]);
// The smallest nonce below `limit` whose compression with `challenge` (any
// 256-bit integer) is below threshold(difficulty + 0.01); false if none.
// Synchronises `s` once per window of nonces.
bool sky_pow_solve_dev(const uint64_t challenge[4], double difficulty, uint64_t limit,
                       uint64_t* nonce, hipStream_t s);

// ---- Poseidon2 (poseidon2.hip) ----
// n states of 4 elements (row-major) -> permuted states.
void p2_permutation_many_dev(Device& d, const fr_t* states, fr_t* out, uint64_t n, hipStream_t s);
// poseidon2_hash of `count` messages of `len` elements each -> count digests.
void p2_hash_many_dev(Device& d, const fr_t* in, uint64_t len, uint64_t count, fr_t* out,
                      hipStream_t s);
// 2-to-1 tree (hash([l, r])) over n leaves: nodes[0..n) leaves, root at 2n-2.
void p2_merkle_dev(Device& d, const fr_t* leaves, uint64_t n, fr_t* nodes, hipStream_t s);
// The same launches recorded once as an executable graph.
hipGraphExec_t p2_merkle_graph(Device& d, const fr_t* leaves, uint64_t n, fr_t* nodes);

// ---- commit pipeline (commit.hip) ----
// One polynomial of n coefficients: RS encode at `blowup` (optionally on the
// coset 5*H), leaves of `leaf_elems` consecutive codeword elements hashed by
// left fold, binary Skyscraper tree.  codeword: n*blowup; nodes: 2*leaves-1.
void commit_dev(Device& d, const fr_t* coeffs, uint64_t n, uint64_t blowup, int coset,
                uint64_t leaf_elems, fr_t* codeword, fr_t* nodes, hipStream_t s);
// WhirR1CSScheme::commit: polynomial (2^num_vars) = witness (witness_size),
// then g (m0 x 4, or none when g is null), then zeros; committed as above.
// witness may equal polynomial; g must not overlap it.
void whir_r1cs_shape_dev(const fr_t* witness, uint64_t witness_size, int num_vars, const fr_t* g,
                         uint64_t m0, fr_t* polynomial, hipStream_t s);
void whir_r1cs_commit_dev(Device& d, const fr_t* witness, uint64_t witness_size, int num_vars,
                          const fr_t* g, uint64_t m0, uint64_t blowup, int coset,
                          uint64_t leaf_elems, fr_t* polynomial, fr_t* codeword, fr_t* nodes,
                          hipStream_t s);
// Synthetic inputs: element i = ChaCha20 block first + i under `key`, as a
// 512-bit integer mod p (inputs.hip).
void random_fr_dev(const uint8_t key[32], uint64_t first, uint64_t n, fr_t* out, hipStream_t s);
// bn254_multiplier::{scalar_mul, scalar_sqr} over n elements (b null: squares).
void scalar_mul_many_dev(const fr_t* a, const fr_t* b, fr_t* out, uint64_t n, hipStream_t s);

}  // namespace pk
