// Poseidon2 over BN254 Fr (t = 4, x^5, 4 + 56 + 4 rounds) on gfx950.
//
// Same permutation as poseidon2/src/permutation.rs:138-170: external MDS,
// 4 full rounds (add constants, x^5 on every lane, external MDS), 56 partial
// rounds (constant and x^5 on lane 0, internal matrix s_i <- s_i d_i + sum),
// 4 full rounds.  The sponge (poseidon2/src/hash.rs:25-55) has rate 3 and
// IV = len * 2^64 in lane 3; the Merkle step is hash([l, r]), one
// permutation of [l, r, 0, 2 * 2^64] keeping lane 0
// (playground/passport-input-gen/src/commitment.rs:225-237).
//
// One lane owns one permutation: 488 Montgomery products, the state in 32
// VGPRs, round constants in constant memory read with uniform indices (they
// arrive as scalar operands).  Values stay in [0, 2p) and leave canonical.
#include "host_fr.hpp"
#include "kernels.hpp"
#include "poseidon2_params.hpp"

namespace pk {

__constant__ fr_t kP2RC[64][4];
__constant__ fr_t kP2DiagM[4];
__constant__ fr_t kP2IV2;  // 2 * 2^64, the IV of a two-element message

namespace {

__device__ __forceinline__ fr_t sbox5(const fr_t& x) {
  const fr_t x2 = fr_mul(x, x);
  const fr_t x4 = fr_mul(x2, x2);
  return fr_mul(x4, x);
}

// matrix_multiplication_4x4 (permutation.rs:105-124)
__device__ __forceinline__ void ext_mds(fr_t s[4]) {
  const fr_t t0 = fr_add(s[0], s[1]);
  const fr_t t1 = fr_add(s[2], s[3]);
  const fr_t t2 = fr_add(fr_add(s[1], s[1]), t1);
  const fr_t t3 = fr_add(fr_add(s[3], s[3]), t0);
  fr_t t4 = fr_add(t1, t1);
  t4 = fr_add(fr_add(t4, t4), t3);
  fr_t t5 = fr_add(t0, t0);
  t5 = fr_add(fr_add(t5, t5), t2);
  const fr_t t6 = fr_add(t3, t5);
  const fr_t t7 = fr_add(t2, t4);
  s[0] = t6;
  s[1] = t5;
  s[2] = t7;
  s[3] = t4;
}

__device__ __forceinline__ void full_round(fr_t s[4], int r) {
#pragma unroll
  for (int i = 0; i < 4; i++) s[i] = sbox5(fr_add(s[i], kP2RC[r][i]));
  ext_mds(s);
}

__device__ __noinline__ void p2_permute(fr_t s[4]) {
  ext_mds(s);
#pragma unroll 1
  for (int r = 0; r < 4; r++) full_round(s, r);
#pragma unroll 1
  for (int r = 4; r < 60; r++) {
    s[0] = sbox5(fr_add(s[0], kP2RC[r][0]));
    const fr_t sum = fr_add(fr_add(s[0], s[1]), fr_add(s[2], s[3]));
#pragma unroll
    for (int i = 0; i < 4; i++) s[i] = fr_add(fr_mul(s[i], kP2DiagM[i]), sum);
  }
#pragma unroll 1
  for (int r = 60; r < 64; r++) full_round(s, r);
#pragma unroll
  for (int i = 0; i < 4; i++) fr_canon(s[i]);
}

__device__ __forceinline__ fr_t p2_two_to_one(const fr_t& l, const fr_t& r) {
  fr_t s[4];
  s[0] = l;
  s[1] = r;
  for (int k = 0; k < 8; k++) s[2].l[k] = 0;
  s[3] = kP2IV2;
  p2_permute(s);
  return s[0];
}

__global__ __launch_bounds__(256) void k_p2_perm_many(const fr_t* in, fr_t* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t s[4];
#pragma unroll
  for (int k = 0; k < 4; k++) s[k] = fr_load(in + 4 * i + k);
  p2_permute(s);
#pragma unroll
  for (int k = 0; k < 4; k++) fr_store(out + 4 * i + k, s[k]);
}

// poseidon2_hash over `count` messages of `len` elements each.
__global__ __launch_bounds__(256) void k_p2_hash_many(const fr_t* in, uint64_t len, uint64_t count,
                                                      fr_t iv, fr_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const fr_t* m = in + i * len;
  fr_t s[4];
  for (int k = 0; k < 3; k++)
    for (int q = 0; q < 8; q++) s[k].l[q] = 0;
  s[3] = iv;
  uint64_t pos = 0;
  // every full block of 3 is added then permuted; the tail (1..3 elements,
  // or none for an empty message) is added before the final permutation
  while (len - pos > 3) {
    for (int k = 0; k < 3; k++) s[k] = fr_add(s[k], fr_load(m + pos + k));
    p2_permute(s);
    pos += 3;
  }
  for (uint64_t k = 0; pos + k < len; k++) s[k] = fr_add(s[k], fr_load(m + pos + k));
  p2_permute(s);
  fr_store(out + i, s[0]);
}

__global__ __launch_bounds__(256) void k_p2_level(const fr_t* in, fr_t* out, uint64_t n_out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  fr_store(out + i, p2_two_to_one(fr_load(in + 2 * i), fr_load(in + 2 * i + 1)));
}

// last levels (<= 64 nodes) in one wave
__global__ __launch_bounds__(64) void k_p2_tail(fr_t* nodes, uint64_t n_in) {
  __shared__ fr_t lvl[64];
  const int t = threadIdx.x;
  if (t < (int)n_in) lvl[t] = fr_load(nodes + t);
  __syncthreads();
  uint64_t off = n_in;
  for (uint64_t n = n_in; n > 1; n >>= 1) {
    const uint64_t h = n >> 1;
    fr_t v;
    if ((uint64_t)t < h) v = p2_two_to_one(lvl[2 * t], lvl[2 * t + 1]);
    __syncthreads();
    if ((uint64_t)t < h) {
      lvl[t] = v;
      fr_store(nodes + off + t, v);
    }
    __syncthreads();
    off += h;
  }
}

__global__ void k_p2_copy(const fr_t* in, fr_t* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fr_store(out + i, fr_load(in + i));
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

std::mutex g_p2_mtx;

}  // namespace

// Parameters in Montgomery form in this device's constant memory (once).
void poseidon2_consts(Device& d) {
  std::lock_guard<std::mutex> lk(g_p2_mtx);
  if (!d.p2_ready) {
    fr_t rc[64][4];
    memset(rc, 0, sizeof rc);
    auto put = [](fr_t& dst, const hfr& v) { memcpy(&dst, &v, 32); };
    for (int r = 0; r < 4; r++)
      for (int j = 0; j < 4; j++) put(rc[r][j], hf::from_hex(kP2Full1[r][j]));
    for (int r = 0; r < 56; r++) put(rc[4 + r][0], hf::from_hex(kP2Partial[r]));
    for (int r = 0; r < 4; r++)
      for (int j = 0; j < 4; j++) put(rc[60 + r][j], hf::from_hex(kP2Full2[r][j]));
    fr_t diag[4];
    for (int i = 0; i < 4; i++) put(diag[i], hf::from_hex(kP2Diag[i]));
    fr_t iv2;
    put(iv2, hf::mul(hf::from_u64(2), hf::from_int(hfr{{0, 1, 0, 0}})));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2RC), rc, sizeof rc));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2DiagM), diag, sizeof diag));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2IV2), &iv2, sizeof iv2));
    d.p2_ready = true;
  }
}

void p2_permutation_many_dev(Device& d, const fr_t* states, fr_t* out, uint64_t n, hipStream_t s) {
  if (!n) return;
  poseidon2_consts(d);
  k_p2_perm_many<<<nblk(n, 256), 256, 0, s>>>(states, out, n);
  PK_HIP(hipGetLastError());
}

void p2_hash_many_dev(Device& d, const fr_t* in, uint64_t len, uint64_t count, fr_t* out,
                      hipStream_t s) {
  if (!count) return;
  poseidon2_consts(d);
  hfr iv = hf::mul(hf::from_u64(len), hf::from_int(hfr{{0, 1, 0, 0}}));
  fr_t ivd;
  memcpy(&ivd, &iv, 32);
  k_p2_hash_many<<<nblk(count, 256), 256, 0, s>>>(in, len, count, ivd, out);
  PK_HIP(hipGetLastError());
}

void p2_merkle_dev(Device& d, const fr_t* leaves, uint64_t n, fr_t* nodes, hipStream_t s) {
  PK_REQUIRE(is_pow2(n), PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
  poseidon2_consts(d);
  if (leaves != nodes) {
    k_p2_copy<<<nblk(n, 256), 256, 0, s>>>(leaves, nodes, n);
    PK_HIP(hipGetLastError());
  }
  fr_t* lvl = nodes;
  while (n > 64) {
    const unsigned t = n / 2 >= (1u << 16) ? 256 : 64;
    k_p2_level<<<nblk(n / 2, t), t, 0, s>>>(lvl, lvl + n, n / 2);
    PK_HIP(hipGetLastError());
    lvl += n;
    n /= 2;
  }
  if (n > 1) {
    k_p2_tail<<<1, 64, 0, s>>>(lvl, n);
    PK_HIP(hipGetLastError());
  }
}

// The level launches of p2_merkle_dev recorded once into a hipGraph.
hipGraphExec_t p2_merkle_graph(Device& d, const fr_t* leaves, uint64_t n, fr_t* nodes) {
  poseidon2_consts(d);
  hipStream_t cs;
  PK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  PK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  try {
    p2_merkle_dev(d, leaves, n, nodes, cs);
  } catch (...) {
    hipGraph_t junk;
    (void)hipStreamEndCapture(cs, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    (void)hipStreamDestroy(cs);
    throw;
  }
  PK_HIP(hipStreamEndCapture(cs, &g));
  hipGraphExec_t ex;
  PK_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  PK_HIP(hipGraphDestroy(g));
  PK_HIP(hipStreamDestroy(cs));
  return ex;
}

}  // namespace pk
