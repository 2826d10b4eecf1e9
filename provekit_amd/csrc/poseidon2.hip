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
// One lane owns one permutation: 488 Montgomery products, round constants
// in constant memory read with uniform indices (scalar operands).  Inputs are
// storage-form elements below 2^256; outputs leave canonical.
//
// The permutation runs in the 9 x 29-bit form of fr29.hpp: the state is
// converted once on entry and once on exit (one product per lane each way),
// the round constants and the internal diagonal are kept in that form, and
// value growth is bounded as follows (units of p):
//   - full rounds: the S-box input is below 30 p, x^2 < 6 p, x^4 < 1.2 p,
//     x^5 < 1.3 p, so the external MDS (row sums up to 16) stays below 21 p;
//   - partial rounds: the sum of the four lanes is reduced below 2 p, so
//     s_i d_i + sum stays below 4 p.
// Limbs are carried before each S-box and inside and after the MDS (whose
// doublings would otherwise overflow 32-bit limbs).
#include "fr29.hpp"
#include "host_fr.hpp"
#include "kernels.hpp"
#include "poseidon2_params.hpp"

namespace pk {

__constant__ fr_t kP2IV2;  // 2 * 2^64, the IV of a two-element message
__constant__ f29 kP2RC29[64][4];  // round constants, internal form
__constant__ f29 kP2Diag29[4];
__constant__ f29 kP2In29;         // 2^266 mod p: storage -> internal
__constant__ f29 kP2Out29;        // 2^256 mod p: internal -> storage

namespace {

// x^5 in the internal form; x normalized
__device__ __forceinline__ f29 sbox5_29(const f29& x) {
  const f29 x2 = sqr29(x);
  const f29 x4 = sqr29(x2);
  return mul29(x4, x);
}

// matrix_multiplication_4x4 (permutation.rs:105-124) on normalized lanes
__device__ __forceinline__ void ext_mds29(f29 s[4]) {
  f29 t0 = add29(s[0], s[1]);
  f29 t1 = add29(s[2], s[3]);
  f29 t2 = add29(add29(s[1], s[1]), t1);
  f29 t3 = add29(add29(s[3], s[3]), t0);
  norm29(t0);
  norm29(t1);
  norm29(t2);
  norm29(t3);
  f29 t4 = add29(t1, t1);
  t4 = add29(add29(t4, t4), t3);
  f29 t5 = add29(t0, t0);
  t5 = add29(add29(t5, t5), t2);
  s[0] = add29(t3, t5);
  s[1] = t5;
  s[2] = add29(t2, t4);
  s[3] = t4;
  // limbs up to 6 * 2^29 here: carried so that the partial rounds' sum of
  // four lanes and the products' operand bounds hold
#pragma unroll
  for (int i = 0; i < 4; i++) norm29(s[i]);
}

// One full round on the four lanes, written out lane by lane: as a loop over
// the lanes inside the rolled round loop the compiler left the lane index
// dynamic and kept the state in scratch memory (private segment 304 bytes,
// 116 VGPRs; written out: none, 76).
__device__ __forceinline__ f29 rc_sbox29(const f29& v, const f29& rc) {
  f29 x = add29(v, rc);
  norm29(x);
  return sbox5_29(x);
}
__device__ __forceinline__ void full_round29(f29 s[4], int r) {
  const f29 s0 = rc_sbox29(s[0], kP2RC29[r][0]);
  const f29 s1 = rc_sbox29(s[1], kP2RC29[r][1]);
  const f29 s2 = rc_sbox29(s[2], kP2RC29[r][2]);
  const f29 s3 = rc_sbox29(s[3], kP2RC29[r][3]);
  s[0] = s0;
  s[1] = s1;
  s[2] = s2;
  s[3] = s3;
  ext_mds29(s);
}

// Inlined into each kernel, which calls it from one place (~7.3k
// instructions, 44 KB): an out-of-line call passes the state through scratch
// memory.
__device__ __forceinline__ void p2_permute29(fr_t st[4]) {
  f29 s[4];
#pragma unroll
  for (int i = 0; i < 4; i++) s[i] = mul29(to29(st[i]), kP2In29);
  ext_mds29(s);
#pragma unroll 1
  for (int r = 0; r < 4; r++) full_round29(s, r);
#pragma unroll 1
  for (int r = 4; r < 60; r++) {
    f29 x = add29(s[0], kP2RC29[r][0]);
    norm29(x);
    s[0] = sbox5_29(x);
    f29 sum = add29(add29(s[0], s[1]), add29(s[2], s[3]));
    norm29(sum);
    reduce29(sum);
#pragma unroll
    for (int i = 0; i < 4; i++) s[i] = add29(mul29(s[i], kP2Diag29[i]), sum);
  }
#pragma unroll 1
  for (int r = 60; r < 64; r++) full_round29(s, r);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f29 v = mul29(s[i], kP2Out29);  // < 1.2 p: one subtraction makes it canonical
    norm29(v);
    st[i] = from29(v);
    fr_canon(st[i]);
  }
}

__device__ __forceinline__ void p2_permute(fr_t s[4]) { p2_permute29(s); }

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
  // or none for an empty message) is added before the final permutation.
  // One loop, so the kernel holds one copy of the permutation.
  do {
    const uint64_t take = len - pos > 3 ? 3 : len - pos;
#pragma unroll
    for (int k = 0; k < 3; k++)  // static indices keep the state in registers
      if ((uint64_t)k < take) s[k] = fr_add(s[k], fr_load(m + pos + k));
    pos += take;
    p2_permute(s);
  } while (pos < len);
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

// Parameters in this device's constant memory (once): the round constants
// and the internal diagonal in the 29-bit internal form, the IV of a
// two-element message in storage form.
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
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2IV2), &iv2, sizeof iv2));
    // internal (2^261 Montgomery, 29-bit limbs) forms: I(c) = S(c) * 32 mod p
    const hfr s32 = hf::from_u64(32);
    auto limbs29 = [](f29& dst, const hfr& v) {  // v: an integer below 2^256
      for (int k = 0; k < 9; k++) {
        const int b = 29 * k, w = b >> 6, sh = b & 63;
        unsigned __int128 x = v.l[w] >> sh;
        if (w + 1 < 4 && sh > 35) x |= (unsigned __int128)v.l[w + 1] << (64 - sh);
        dst.l[k] = (uint32_t)(x & 0x1fffffffu);
      }
    };
    static f29 rc29[64][4], diag29[4], in29, out29;
    for (int r = 0; r < 64; r++)
      for (int j = 0; j < 4; j++) {
        hfr v;
        memcpy(&v, &rc[r][j], 32);
        limbs29(rc29[r][j], hf::mul(v, s32));
      }
    for (int i = 0; i < 4; i++) {
      hfr v;
      memcpy(&v, &diag[i], 32);
      limbs29(diag29[i], hf::mul(v, s32));
    }
    limbs29(in29, hf::from_u64(1024));  // 1024 * 2^256 = 2^266 (mod p)
    limbs29(out29, hf::ONE);            // 2^256 (mod p)
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2RC29), rc29, sizeof rc29));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2Diag29), diag29, sizeof diag29));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2In29), &in29, sizeof in29));
    PK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kP2Out29), &out29, sizeof out29));
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
