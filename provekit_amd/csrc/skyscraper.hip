// Skyscraper hashing on gfx950: batch 2-to-1 compression, leaf hashing by
// left fold (SkyscraperHashEngine::hash_many,
// provekit/common/src/skyscraper/whir.rs:42-100) and binary Merkle trees.
//
// Every compression is ~3.9k VALU instructions (14 Montgomery squarings, 4
// byte-S-box layers read from an LDS table, lazy reductions), so these
// kernels are VALU-bound; one lane owns one message, which keeps every lane
// busy with no cross-lane traffic.  The upper tree levels are latency-bound
// (a compression is a long dependent chain on one wave), so they run one wave
// per workgroup and the last six levels run inside a single wave.
#include "kernels.hpp"
#include "sky.hpp"

namespace pk {

namespace {

// Every kernel here runs the table form of the compression (sky_compress_t):
// the workgroup fills its LDS reduction table before any lane exits.
__global__ __launch_bounds__(256) void k_compress_many(const fr_t* pairs, fr_t* out, uint64_t n) {
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fr_t l = fr_load(pairs + 2 * i), r = fr_load(pairs + 2 * i + 1);
  fr_store(out + i, sky_compress_t(l, r, tab));
}

// One lane per message of `chunks` words.  At most 128 VGPRs (four waves
// per SIMD, the whole 2^18-leaf grid resident at once): left to itself the
// compiler spends 167 on the unrolled compression and the word window, which
// fits three.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_hash_fold(const fr_t* in, uint64_t chunks, uint64_t count,
                                                   fr_t* out) {
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  fr_store(out + i, sky_hash_message_t(in + i * chunks, chunks, tab));
}

// The tree's upper levels are a chain of latency-bound launches (one
// compression per level on the critical path).  When another commit's
// kernels share the SIMDs (commits pipelined on two streams), their waves
// would take most issue slots from these few: the tree waves raise their
// issue priority (s_setprio) so the chain runs at close to its lone-wave
// speed while the other work fills the remaining slots.
__global__ __launch_bounds__(256) void k_merkle_level(const fr_t* in, fr_t* out, uint64_t n_out) {
  __builtin_amdgcn_s_setprio(3);
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  fr_store(out + i, sky_compress_t(fr_load(in + 2 * i), fr_load(in + 2 * i + 1), tab));
}

// The last levels, n_in <= 128 nodes down to the root, in one wave: lane t
// hashes nodes 2t and 2t + 1 of the first of them from memory, then every
// level takes its two children from neighbouring lanes by shuffles, so no
// level waits on LDS, a barrier or its own stores.  These levels are
// latency-bound (one compression per level on the critical path), so a
// single wave is as fast as any spread and needs no launches.
__global__ __launch_bounds__(64) void k_merkle_tail(fr_t* nodes, uint64_t n_in) {
  __builtin_amdgcn_s_setprio(3);
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const int t = threadIdx.x;
  bool on = (uint64_t)t < (n_in >> 1);
  fr_t x = {}, y = {};
  if (on) {
    x = fr_load(nodes + 2 * t);
    y = fr_load(nodes + 2 * t + 1);
  }
  uint64_t off = n_in;
  for (uint64_t n = n_in; n > 1; n >>= 1) {
    fr_t v = {};
    if (on) {
      v = sky_compress_t(x, y, tab);
      fr_store(nodes + off + t, v);
    }
    off += n >> 1;
    x = fr_shfl(v, (2 * t) & 63);
    y = fr_shfl(v, (2 * t + 1) & 63);
    on = (uint64_t)t < (n >> 2);
  }
}

__global__ void k_copy(const fr_t* in, fr_t* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fr_store(out + i, fr_load(in + i));
}

// Leaf opening: query q's leaf index i -> the leaf's elements and the
// sibling digest at every level (level-by-level node layout, root last).
__global__ void k_open(const fr_t* codeword, const fr_t* nodes, uint64_t leaves, uint64_t leaf_elems,
                       int depth, const uint64_t* idx, uint64_t n, fr_t* values, fr_t* paths) {
  const uint64_t q = blockIdx.x;
  if (q >= n) return;
  const uint64_t i = idx[q];
  if (i >= leaves) return;  // the host checks; never read outside the tree
  for (uint64_t k = threadIdx.x; k < leaf_elems; k += blockDim.x)
    fr_store(values + q * leaf_elems + k, fr_load(codeword + i * leaf_elems + k));
  for (int l = threadIdx.x; l < depth; l += blockDim.x) {
    const uint64_t off = 2 * leaves - 2 * (leaves >> l);
    fr_store(paths + q * depth + l, fr_load(nodes + off + (((i >> l) ^ 1))));
  }
}

inline unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void sky_compress_many_dev(const fr_t* pairs, fr_t* out, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_compress_many<<<blocks_for(n, 256), 256, 0, s>>>(pairs, out, n);
  PK_HIP(hipGetLastError());
}

void sky_hash_many_dev(uint64_t chunks, const fr_t* in, uint64_t count, fr_t* out, hipStream_t s) {
  if (!count) return;
  if (chunks == 1) {
    k_copy<<<blocks_for(count, 256), 256, 0, s>>>(in, out, count);
  } else {
    k_hash_fold<<<blocks_for(count, 256), 256, 0, s>>>(in, chunks, count, out);
  }
  PK_HIP(hipGetLastError());
}

void sky_merkle_dev(const fr_t* leaves, uint64_t n, fr_t* nodes, hipStream_t s) {
  PK_REQUIRE(is_pow2(n), PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
  if (leaves != nodes) {
    k_copy<<<blocks_for(n, 256), 256, 0, s>>>(leaves, nodes, n);
    PK_HIP(hipGetLastError());
  }
  fr_t* lvl = nodes;
  while (n > 128) {
    // one-wave workgroups: the narrow upper levels spread over CUs instead of
    // sharing SIMDs
    const unsigned t = n / 2 >= (1u << 16) ? 256 : 64;
    k_merkle_level<<<blocks_for(n / 2, t), t, 0, s>>>(lvl, lvl + n, n / 2);
    PK_HIP(hipGetLastError());
    lvl += n;
    n /= 2;
  }
  if (n > 1) {
    k_merkle_tail<<<1, 64, 0, s>>>(lvl, n);
    PK_HIP(hipGetLastError());
  }
}

void merkle_open_dev(const fr_t* codeword, const fr_t* nodes, uint64_t leaves, uint64_t leaf_elems,
                     const uint64_t* idx, uint64_t n, fr_t* values, fr_t* paths, hipStream_t s) {
  if (!n) return;
  PK_REQUIRE(is_pow2(leaves), PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
  PK_REQUIRE(n < (1ull << 31), PK_ERR_SIZE, "merkle open: too many queries");
  k_open<<<(unsigned)n, 64, 0, s>>>(codeword, nodes, leaves, leaf_elems, log2_exact(leaves), idx, n,
                                     values, paths);
  PK_HIP(hipGetLastError());
}

}  // namespace pk
